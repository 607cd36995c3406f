// gfx950 design-matrix assembly: neighbour build, cluster enumeration and the
// 2/3/4-body A-row kernels of the ChIMES least-squares path.
//
// Reference behaviour reproduced (paths relative to the ChIMES-lsq tree):
//   neighbour sets      NEIGHBORS::DO_UPDATE_BIG/SMALL  src/ClassDefs.C:137-393
//   2B rows             Cheby::Deriv_2B                 src/Cheby.C:709-883
//   3B rows             Cheby::Deriv_3B                 src/Cheby.C:885-1277
//   4B rows             Cheby::Deriv_4B                 src/Cheby.C:1279-1729
//   row layout          A_MAT::PRINT_FRAME              src/A_Matrix.C:210-567
//
// Design: every cluster is re-ordered into the canonical atom order of its
// cluster type when it is enumerated, so the power tables of a type are the
// same for every cluster of that type.  One workgroup owns (frame, cluster
// type, tile of columns); its lanes walk the clusters of that (frame, type)
// bucket, keep per-pair Chebyshev tables in vector registers (indexed by the
// wave-uniform power of the current allowed power set), and accumulate the
// force / stress / energy rows of the tile in LDS.  The tile is written to A
// once, coalesced along its columns.
#include <hip/hip_runtime.h>
#include <type_traits>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "device.hpp"
#include "setup.hpp"

namespace cg {

thread_local bool t_no_grow = false;
thread_local uint64_t t_grow_events = 0;


// ------------------------------------------------------------ host geometry
// Cell setup of BOX::UPDATE_CELL (src/ClassDefs.C:1940-1971 and callees).
static void host_cell(const double *cell, DevFrame &f) {
  double ax = cell[0], ay = cell[1], az = cell[2], bx = cell[3], by = cell[4], bz = cell[5], cx = cell[6],
         cy = cell[7], cz = cell[8];
  double H[9] = {ax, bx, cx, ay, by, cy, az, bz, cz};
  f.ortho = (std::fabs(ay) + std::fabs(az) + std::fabs(bx) + std::fabs(bz) + std::fabs(cx) + std::fabs(cy) > 1.0e-10) ? 0 : 1;
  double det = H[0] * (H[4] * H[8] - H[5] * H[7]) - H[1] * (H[3] * H[8] - H[5] * H[6]) + H[2] * (H[3] * H[7] - H[4] * H[6]);
  if (det == 0) throw Error(CG_ERR_INPUT, "ERROR: H-matrix determinant of zero computed in INVERT_HMAT");
  double T[9] = {(H[4] * H[8] - H[5] * H[7]), -(H[3] * H[8] - H[5] * H[6]), (H[3] * H[7] - H[4] * H[6]),
                 -(H[1] * H[8] - H[2] * H[7]), (H[0] * H[8] - H[2] * H[6]), -(H[0] * H[7] - H[1] * H[6]),
                 (H[1] * H[5] - H[2] * H[4]), -(H[0] * H[5] - H[2] * H[3]), (H[0] * H[4] - H[1] * H[3])};
  double I[9] = {T[0], T[3], T[6], T[1], T[4], T[7], T[2], T[5], T[8]};
  for (int i = 0; i < 9; ++i) { f.H[i] = H[i]; f.Hi[i] = I[i] / det; }
  auto mag = [](double x, double y, double z) { return std::sqrt(x * x + y * y + z * z); };
  auto ang = [&](double x1, double y1, double z1, double x2, double y2, double z2) {
    return std::acos((x1 * x2 + y1 * y2 + z1 * z2) / (mag(x1, y1, z1) * mag(x2, y2, z2)));
  };
  double la = mag(ax, ay, az), lb = mag(bx, by, bz), lc = mag(cx, cy, cz);
  double al = ang(bx, by, bz, cx, cy, cz), be = ang(cx, cy, cz, ax, ay, az), ga = ang(ax, ay, az, bx, by, bz);
  double xy = lb * std::cos(ga); if (std::fabs(xy) < 1e-12) xy = 0.0;
  double xz = lc * std::cos(be); if (std::fabs(xz) < 1e-12) xz = 0.0;
  double ly = std::sqrt(lb * lb - xy * xy);
  double yz = (lb * lc * std::cos(al) - xy * xz) / ly; if (std::fabs(yz) < 1e-12) yz = 0.0;
  double lz = std::sqrt(lc * lc - xz * xz - yz * yz);
  f.lx = la; f.ly = ly; f.lz = lz;
  {  // BOX::UPDATE_EXTENT
    double xlo = 0, ylo = 0, t = 0;
    if (xy < xlo) xlo = xy;
    if (xz < xlo) xlo = xz;
    if (xy + xz < xlo) xlo = xy + xz;
    if (yz < ylo) ylo = yz;
    double xhi = ax;
    if (xy > t) t = xy;
    if (xz > t) t = xz;
    if (xy + xz > t) t = xy + xz;
    xhi += t;
    double yhi = by;
    if (yz > 0.0) yhi += yz;
    f.min_extent = std::min(xhi - xlo, std::min(yhi - ylo, cz));
  }
  double vol;
  if (f.ortho) vol = ax * by * cz;
  else {
    double v = 1 + 2 * std::cos(al) * std::cos(be) * std::cos(ga) - std::cos(al) * std::cos(al) -
               std::cos(be) * std::cos(be) - std::cos(ga) * std::cos(ga);
    vol = la * lb * lc * std::sqrt(v);
  }
  f.inv_vol = 1.0 / vol;
  // perpendicular widths |det| / |b x c| etc.
  auto cross = [](const double *u, const double *v, double *w) {
    w[0] = u[1] * v[2] - u[2] * v[1]; w[1] = u[2] * v[0] - u[0] * v[2]; w[2] = u[0] * v[1] - u[1] * v[0];
  };
  const double *A = cell, *B = cell + 3, *C = cell + 6;
  double w[3];
  cross(B, C, w); f.width[0] = std::fabs(det) / mag(w[0], w[1], w[2]);
  cross(C, A, w); f.width[1] = std::fabs(det) / mag(w[0], w[1], w[2]);
  cross(A, B, w); f.width[2] = std::fabs(det) / mag(w[0], w[1], w[2]);
}

// ------------------------------------------------------------ device helpers
// Cartesian image position: wrapped coordinate + H * shift.
__device__ __forceinline__ void image_pos(const DevFrame &f, const double *wrap, uint32_t r, double *p) {
  int a = ref_atom(r);
  double sx = ref_sx(r), sy = ref_sy(r), sz = ref_sz(r);
  const double *w = wrap + 3 * (f.atom_off + a);
  p[0] = w[0] + f.H[0] * sx + f.H[1] * sy + f.H[2] * sz;
  p[1] = w[1] + f.H[3] * sx + f.H[4] * sy + f.H[5] * sz;
  p[2] = w[2] + f.H[6] * sx + f.H[7] * sy + f.H[8] * sz;
}

// BOX::WRAP_ATOM (src/ClassDefs.C:2053-2110).
__global__ void k_wrap(const DevFrame *frames, const int *atom_frame, const double *xyz, double *wrap, int natoms) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= natoms) return;
  const DevFrame &f = frames[atom_frame[i]];
  const double *u = xyz + 3 * i;
  double *w = wrap + 3 * i;
  if (f.ortho) {
    double ix = floor(u[0] / f.lx), iy = floor(u[1] / f.ly), iz = floor(u[2] / f.lz);
    w[0] = u[0] - ix * f.lx; w[1] = u[1] - iy * f.ly; w[2] = u[2] - iz * f.lz;
  } else {
    double s[3];
    for (int r = 0; r < 3; ++r) s[r] = f.Hi[3 * r] * u[0] + f.Hi[3 * r + 1] * u[1] + f.Hi[3 * r + 2] * u[2];
    for (int r = 0; r < 3; ++r) s[r] -= floor(s[r]);
    for (int r = 0; r < 3; ++r) w[r] = f.H[3 * r] * s[0] + f.H[3 * r + 1] * s[1] + f.H[3 * r + 2] * s[2];
  }
}

// Cell list: bin of each atom in fractional coordinates (frames that use the
// cell search), as a sort key; other frames sort after every bin.
__global__ void k_bin_keys(const DevFrame *frames, const int *atom_frame, const double *wrap, int natoms,
                           int nbins_total, int *key, int *val) {
  const int gi = blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= natoms) return;
  const DevFrame &f = frames[atom_frame[gi]];
  val[gi] = gi;
  if (!f.cell) { key[gi] = nbins_total; return; }
  const double *w = wrap + 3 * gi;
  int b[3];
  for (int r = 0; r < 3; ++r) {
    const double s = f.Hi[3 * r] * w[0] + f.Hi[3 * r + 1] * w[1] + f.Hi[3 * r + 2] * w[2];
    b[r] = min(f.nbin[r] - 1, max(0, (int)floor(s * f.nbin[r])));
  }
  key[gi] = f.bin_off + (b[2] * f.nbin[1] + b[1]) * f.nbin[0] + b[0];
}

// Wave-wide exclusive prefix sum of one int per lane (and the wave total).
__device__ __forceinline__ int wave_excl_scan(int v, int &total) {
  const int lane = threadIdx.x & 63;
  int x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  total = __shfl(x, 63, 64);
  return x - v;
}

// Neighbour images of atom i within rc: every (j, shift) with shift in
// [-L, L]^3 (the reference's ghost layers, L = the frame's layer count), restricted to shifts that can lie
// within rc along each axis.  Big cells keep j >= i (half list with the
// reference's a1 <= PARENT[a2] rule); small cells keep every image.
// Frames at least three list cutoffs thick along every axis search a cell
// list (27 fractional-coordinate bins around the atom, the entries in bin
// order); the others scan every atom, one wave per atom with lanes over j,
// entries sorted by parent j then shift (sx, sy, sz).  Both give the same set
// of (j, shift) entries.  FILL also stores, per entry, the relative vector
// image(j) - i with its length, the owning atom and the parent's type.
template <bool FILL>
__global__ __launch_bounds__(256) void k_neigh(const DevFrame *frames, const int *atom_frame, const double *wrap,
                                                const int *type, const int *bin_start, const int *bin_atom,
                                                int natoms, double rc, int *count, const int *offset,
                                                uint32_t *list, double4 *nvec, int *nown, int *ntyp, int cap) {
  const int gi = blockIdx.x * 4 + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (gi >= natoms) return;
  const DevFrame &f = frames[atom_frame[gi]];
  const int i = gi - f.atom_off;
  const double *wi = wrap + 3 * gi;
  double fi[3];
  for (int r = 0; r < 3; ++r) fi[r] = f.Hi[3 * r] * wi[0] + f.Hi[3 * r + 1] * wi[1] + f.Hi[3 * r + 2] * wi[2];
  const double rc2 = rc * rc;
  double reach[3];
  for (int r = 0; r < 3; ++r) reach[r] = rc / f.width[r] + 1e-9;
  int run = FILL ? offset[gi] : 0;
  if (f.cell) {
    // 27 neighbouring bins (each >= rc thick along its axis, at least 3 per
    // axis): every image within rc lies there, with shift -1/0/+1 per axis
    // where the stencil crosses the cell face
    int bi[3];
    for (int r = 0; r < 3; ++r) bi[r] = min(f.nbin[r] - 1, max(0, (int)floor(fi[r] * f.nbin[r])));
    for (int st = 0; st < 27; ++st) {
      int bn[3], sh[3];
      const int d[3] = {st % 3 - 1, (st / 3) % 3 - 1, st / 9 - 1};
      for (int r = 0; r < 3; ++r) {
        bn[r] = bi[r] + d[r];
        sh[r] = 0;
        if (bn[r] < 0) { bn[r] += f.nbin[r]; sh[r] = -1; }
        else if (bn[r] >= f.nbin[r]) { bn[r] -= f.nbin[r]; sh[r] = 1; }
      }
      const int bin = f.bin_off + (bn[2] * f.nbin[1] + bn[1]) * f.nbin[0] + bn[0];
      const int c0 = bin_start[bin], c1 = bin_start[bin + 1];
      for (int cb = c0; cb < c1; cb += 64) {
        const int c = cb + lane;
        int n = 0, j = -1;
        double v[3] = {0, 0, 0}, r2 = 0.0;
        if (c < c1) {
          const int gj = bin_atom[c];
          j = gj - f.atom_off;
          if (j >= i && !(j == i && sh[0] == 0 && sh[1] == 0 && sh[2] == 0)) {
            const double *wj = wrap + 3 * gj;
            for (int r = 0; r < 3; ++r)
              v[r] = wj[r] - wi[r] + f.H[3 * r] * sh[0] + f.H[3 * r + 1] * sh[1] + f.H[3 * r + 2] * sh[2];
            r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
            n = r2 < rc2 ? 1 : 0;
          }
        }
        int tot;
        const int pos = run + wave_excl_scan(n, tot);
        if (FILL && n && pos < cap) {
          list[pos] = make_ref(j, sh[0], sh[1], sh[2]);
          nvec[pos] = make_double4(v[0], v[1], v[2], sqrt(r2));
          nown[pos] = gi;
          ntyp[pos] = type[f.atom_off + j];
        }
        run += tot;
      }
    }
    if (!FILL && lane == 0) count[gi] = run;
    return;
  }
  for (int jb = f.small ? 0 : i; jb < f.natoms; jb += 64) {
    const int j = jb + lane;
    int n = 0;
    int lo[3] = {0, 0, 0}, hi[3] = {-1, -1, -1};
    const double *wj = wrap + 3 * (f.atom_off + (j < f.natoms ? j : i));
    if (j < f.natoms) {
      double fj[3], d[3];
      for (int r = 0; r < 3; ++r) fj[r] = f.Hi[3 * r] * wj[0] + f.Hi[3 * r + 1] * wj[1] + f.Hi[3 * r + 2] * wj[2];
      for (int r = 0; r < 3; ++r) d[r] = fj[r] - fi[r];
      for (int r = 0; r < 3; ++r) {
        lo[r] = max(-f.nlayers, (int)ceil(-d[r] - reach[r]));
        hi[r] = min(f.nlayers, (int)floor(-d[r] + reach[r]));
      }
    }
    // pass 1: count this lane's images; pass 2 (FILL): write them in order
    for (int pass = 0; pass < (FILL ? 2 : 1); ++pass) {
      int base = 0;
      if (pass == 1) {
        int tot;
        base = run + wave_excl_scan(n, tot);
        run += tot;
      }
      int m = 0;
      for (int sx = lo[0]; sx <= hi[0]; ++sx)
        for (int sy = lo[1]; sy <= hi[1]; ++sy)
          for (int sz = lo[2]; sz <= hi[2]; ++sz) {
            if (j == i && sx == 0 && sy == 0 && sz == 0) continue;
            double v[3];
            for (int r = 0; r < 3; ++r)
              v[r] = wj[r] - wi[r] + f.H[3 * r] * sx + f.H[3 * r + 1] * sy + f.H[3 * r + 2] * sz;
            const double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
            if (r2 < rc2) {
              if (pass == 1 && base + m < cap) {
                list[base + m] = make_ref(j, sx, sy, sz);
                nvec[base + m] = make_double4(v[0], v[1], v[2], sqrt(r2));
                nown[base + m] = gi;
                ntyp[base + m] = type[f.atom_off + j];
              }
              ++m;
            }
          }
      n = m;
    }
    if (!FILL) {
      int tot;
      wave_excl_scan(n, tot);
      run += tot;
    }
  }
  if (!FILL && lane == 0) count[gi] = run;
}

struct EnumArgs {
  const DevFrame *frames;
  const int *atom_frame;
  const int *type;      // per atom (batch)
  const int *noff;      // neighbour offsets [natoms+1]
  const uint32_t *nlist;
  const double4 *nvec;  // relative vector + length per entry
  const int *nown;      // owning atom per entry
  const int *ntyp;      // parent atom type per entry
  int nnb, natoms, ntypes_atom, nkeys_type;   // nnb: capacity of the entry arrays
  int ccap;             // capacity of the cluster arrays (fill passes write below it)
  const int *ntot;      // device-side number of valid entries (<= nnb)
  double cutoff;
  int *count;           // [nnb]
  const int *coff;      // [nnb+1]
  uint32_t *refs;       // [ncl * nb]
  int *key;             // [ncl] frame * types + type
  // family tables
  const double *smin, *smax;
  // per type tuple (index sum_q t_q * ntypes_atom^q of the tuple in
  // enumeration order): cluster type (-1: none), S_MAXIM of each enumerated
  // pair slot, canonical slot permutation
  const int *cct;
  const double *csmax;
  const int *cperm;
  const int *cslot;     // canonical pair slot of each enumerated pair slot (PAIR_INDICES)
  unsigned long long *slot_min;  // fit statistics: per (cluster type, slot) smallest distance, or null
  const int *int_pair_map;
  int md;               // force-evaluation pair set
};

__device__ __forceinline__ double dist3(const double4 &u, const double4 &v) {
  double dx = v.x - u.x, dy = v.y - u.y, dz = v.z - u.z;
  return sqrt(dx * dx + dy * dy + dz * dz);
}

// Cluster-type lookup (make_id_int + INT_MAP) and PROCEED on every pair with
// its canonical pair's S_MAXIM (src/Cheby.C:1015-1054, 1387-1500), through
// tables built on the host per type tuple (index sum_q t_q * ntypes^q, types
// in enumeration order): one small cached row per candidate.  d[] holds the
// pair distances in enumeration order (0,1) (0,2) (0,3) (1,2) (1,3) (2,3).
template <int NB>
__device__ __forceinline__ int cluster_type_c(const EnumArgs &a, const int *types, const double *d,
                                              const int *&perm, int &cid_out) {
  constexpr int NP = NB * (NB - 1) / 2;
  int cid = 0;
#pragma unroll
  for (int q = NB - 1; q >= 0; --q) cid = cid * a.ntypes_atom + types[q];
  const int ct = a.cct[cid];
  if (ct < 0) return -1;
  const double *row = a.csmax + (size_t)cid * NP;
#pragma unroll
  for (int pf = 0; pf < NP; ++pf)
    if (!(d[pf] < row[pf])) return -1;
  perm = a.cperm + cid * NB;
  cid_out = cid;
  return ct;
}

// Smallest distance of each canonical pair slot of a cluster type, as the
// reference tracks it (MIN_FOUND[PAIR_INDICES[f]] = r_f, src/Cheby.C:1059-1078,
// 1506-1519): bit patterns of positive doubles order like the values, and the
// plain read first keeps the atomics to the few clusters that lower a minimum.
template <int NP>
__device__ __forceinline__ void record_slot_min(const EnumArgs &a, int ct, int cid, const double *d) {
  const int *slot = a.cslot + cid * NP;
#pragma unroll
  for (int pf = 0; pf < NP; ++pf) {
    unsigned long long *m = a.slot_min + ct * NP + slot[pf];
    const unsigned long long b = __builtin_bit_cast(unsigned long long, d[pf]);
    if (b < __hip_atomic_load(m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMin(m, b);
  }
}

// 2-body enumeration (the pair loop of Deriv_2B / Force_all): one thread per
// neighbour entry; LSQ keeps S_MINIM < r < S_MAXIM, force evaluation every
// r < S_MAXIM.  FILL stores the pair with its (frame, pair type) key.
template <bool FILL>
__global__ __launch_bounds__(256) void k_enum2(EnumArgs a) {
  const int k2 = blockIdx.x * blockDim.x + threadIdx.x;
  if (k2 >= a.nnb) return;
  if (k2 >= *a.ntot) {
    if (!FILL) a.count[k2] = 0;
    return;
  }
  const double4 v2 = a.nvec[k2];
  int n = 0;
  if (v2.w < a.cutoff) {
    const int gi = a.nown[k2];
    const int fidx = a.atom_frame[gi];
    const DevFrame &f = a.frames[fidx];
    const uint32_t r2 = a.nlist[k2];
    const int p = a.int_pair_map[a.type[gi] * a.ntypes_atom + a.ntyp[k2]];
    const double r = v2.w;
    if (r < a.smax[p] && (a.md || r > a.smin[p])) {
      const int out = FILL ? a.coff[k2] : 0;
      if (FILL && out < a.ccap) {
        a.refs[2 * out] = make_ref(gi - f.atom_off, 0, 0, 0);
        a.refs[2 * out + 1] = r2;
        a.key[out] = fidx * a.nkeys_type + p;
      }
      n = 1;
    }
  }
  if (!FILL) a.count[k2] = n;
}

// Per-family neighbour sublist: the entries of each atom's list closer than
// the family cutoff, in list order (so still sorted by parent atom).  One
// wave per atom; COUNT writes the per-atom size, otherwise the entries are
// compacted to soff[atom].
template <bool COUNT>
__global__ __launch_bounds__(256) void k_sub(const int *noff, const uint32_t *nlist, const double4 *nvec,
                                             const int *nown, const int *ntyp, int natoms, double cut, int *cnt,
                                             const int *soff, uint32_t *slist, double4 *svec, int *sown, int *styp,
                                             int cap) {
  const int gi = blockIdx.x * 4 + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (gi > natoms) return;
  if (gi == natoms) {
    if (COUNT && lane == 0) cnt[gi] = 0;
    return;
  }
  // entries past the list capacity were not stored (the caller detects the
  // overflow from the device totals and repeats the call with exact sizes)
  const int b = min(noff[gi], cap), e = min(noff[gi + 1], cap);
  int run = COUNT ? 0 : soff[gi];
  for (int k0 = b; k0 < e; k0 += 64) {
    const int k = k0 + lane;
    double4 v = make_double4(0, 0, 0, 0);
    bool keep = false;
    if (k < e) {
      v = nvec[k];
      keep = v.w < cut;
    }
    int tot;
    const int pos = run + wave_excl_scan(keep ? 1 : 0, tot);
    if (!COUNT && keep) {
      slist[pos] = nlist[k];
      svec[pos] = v;
      sown[pos] = nown[k];
      styp[pos] = ntyp[k];
    }
    run += tot;
  }
  if (COUNT && lane == 0) cnt[gi] = run;
}

// 3- and 4-body enumeration, one wave per anchor atom: the atom's (sub)list
// is staged in LDS, lanes take the (k2, k3) entry pairs of the atom in order,
// and each lane walks k4 for quadruplets.  The same rules as k_enum apply
// (parent(a2) <= parent(a3) <= parent(a4) in big cells, all orderings in
// small cells, every pair below the family cutoff, then the type lookup and
// PROCEED), and the clusters of an atom are written in (k2, k3, k4) order.
// COUNT writes the per-atom cluster count.  Lists longer than the LDS stage
// are read from global memory instead.
constexpr int ENUM_STAGE = 160;

struct StageLds {
  const double4 *v; const uint32_t *r; const int *t;
  __device__ __forceinline__ double4 vec(int k) const { return v[k]; }
  __device__ __forceinline__ uint32_t ref(int k) const { return r[k]; }
  __device__ __forceinline__ int typ(int k) const { return t[k]; }
};
struct StageGlobal {
  const double4 *v; const uint32_t *r; const int *t;
  __device__ __forceinline__ double4 vec(int k) const { return v[k]; }
  __device__ __forceinline__ uint32_t ref(int k) const { return r[k]; }
  __device__ __forceinline__ int typ(int k) const { return t[k]; }
};

constexpr int ENUM_ADJ_WORDS = (ENUM_STAGE + 63) / 64;

template <int NB, bool COUNT, class S>
__device__ __forceinline__ int enum_atom(const EnumArgs &a, const S &st, int n, int qa, int qb, bool ordered,
                                         uint32_t r1, int t1, int kbase, int out, const unsigned long long *adj) {
  const int lane = threadIdx.x & 63;
  const double cut = a.cutoff;
  int run = 0;
  for (int q0 = qa; q0 < qb; q0 += 64) {
    const int q = q0 + lane;
    int k2 = 0, k3 = 0;
    bool ok = q < qb;
    double4 v2 = make_double4(0, 0, 0, 0), v3 = v2;
    uint32_t r2 = 0, r3 = 0;
    double r23 = 0.0;
    if (ok) {
      k2 = q / n;
      k3 = q - k2 * n;
      ok = k3 != k2;
    }
    if (ok) {
      r2 = st.ref(k2);
      r3 = st.ref(k3);
      ok = !ordered || ref_atom(r2) <= ref_atom(r3);
    }
    if (ok) {
      v2 = st.vec(k2);
      v3 = st.vec(k3);
      ok = v2.w < cut && v3.w < cut;
    }
    if (ok) {
      r23 = dist3(v2, v3);
      ok = r23 < cut;
    }
    int types[4] = {t1, 0, 0, 0};
    if (ok) {
      types[1] = st.typ(k2);
      types[2] = st.typ(k3);
    }
    // pass 0 counts this lane's clusters, pass 1 writes them (not for COUNT);
    // with an adjacency table pass 0 also records the accepted k4 so that
    // pass 1 revisits only those
    int mine = 0;
    unsigned long long kept[ENUM_ADJ_WORDS];
#pragma unroll
    for (int w = 0; w < ENUM_ADJ_WORDS; ++w) kept[w] = 0ull;
    for (int pass = 0; pass < (COUNT ? 1 : 2); ++pass) {
      int base = 0;
      if (pass == 1) {
        int tot;
        base = out + run + wave_excl_scan(mine, tot);
        run += tot;
      }
      int m = 0;
      if (ok) {
        if (NB == 3) {
          const double d[3] = {v2.w, v3.w, r23};
          const int *perm;
          int cid;
          const int ct = cluster_type_c<3>(a, types, d, perm, cid);
          if (ct >= 0) {
            if (pass == 1 && base + m < a.ccap) {
              const uint32_t rr[3] = {r1, r2, r3};
              for (int x = 0; x < 3; ++x) a.refs[3 * (base + m) + x] = rr[perm[x]];
              a.key[base + m] = kbase + ct;
              if (a.slot_min) record_slot_min<3>(a, ct, cid, d);
            }
            ++m;
          }
        } else {
          const int p3 = ref_atom(r3);
          // candidates k4: with an adjacency table, the entries within the
          // cutoff of both k2 and k3 (set bits, in index order); else all
          const bool replay = adj && pass == 1;
          int word = 0;
          unsigned long long bits = 0;
          if (adj) bits = replay ? kept[0] : (adj[k2 * ENUM_ADJ_WORDS] & adj[k3 * ENUM_ADJ_WORDS]);
          for (int kk = 0;; ++kk) {
            int k4;
            if (adj) {
              while (!bits && ++word < ENUM_ADJ_WORDS)
                bits = replay ? kept[word] : (adj[k2 * ENUM_ADJ_WORDS + word] & adj[k3 * ENUM_ADJ_WORDS + word]);
              if (!bits) break;
              k4 = word * 64 + __builtin_ctzll(bits);
              bits &= bits - 1;
            } else {
              if (kk >= n) break;
              k4 = kk;
            }
            if (k4 == k2 || k4 == k3) continue;
            const uint32_t r4 = st.ref(k4);
            if (ordered && ref_atom(r4) < p3) continue;
            const double4 v4 = st.vec(k4);
            if (!(v4.w < cut)) continue;
            const double r24 = dist3(v2, v4);
            if (!(r24 < cut)) continue;
            const double r34 = dist3(v3, v4);
            if (!(r34 < cut)) continue;
            types[3] = st.typ(k4);
            const double d[6] = {v2.w, v3.w, v4.w, r23, r24, r34};
            const int *perm;
            int cid;
            const int ct = cluster_type_c<4>(a, types, d, perm, cid);
            if (ct < 0) continue;
            if (pass == 0 && adj) {
#pragma unroll
              for (int w = 0; w < ENUM_ADJ_WORDS; ++w)
                if (k4 / 64 == w) kept[w] |= 1ull << (k4 % 64);
            }
            if (pass == 1 && base + m < a.ccap) {
              const uint32_t rr[4] = {r1, r2, r3, r4};
              for (int x = 0; x < 4; ++x) a.refs[4 * (base + m) + x] = rr[perm[x]];
              a.key[base + m] = kbase + ct;
              if (a.slot_min) record_slot_min<6>(a, ct, cid, d);
            }
            ++m;
          }
        }
      }
      mine = m;
    }
    if (COUNT) {
      int tot;
      wave_excl_scan(mine, tot);
      run += tot;
    }
  }
  return run;
}

// One workgroup per anchor atom: the atom's list is staged in LDS (and, for
// 4-body, its adjacency bitmasks built) by the whole workgroup, then wave w
// takes the w-th quarter of the (k2, k3) entry pairs.  Counts and outputs are
// per (atom, wave), so the clusters of an atom keep their (k2, k3, k4) order.
template <int NB, bool COUNT>
__global__ __launch_bounds__(256) void k_enum_wave(EnumArgs a) {
  __shared__ double4 s_vec[ENUM_STAGE];
  __shared__ uint32_t s_ref[ENUM_STAGE];
  __shared__ int s_typ[ENUM_STAGE];
  // 4-body: bit l of row k set when entries k and l are closer than the cutoff
  __shared__ unsigned long long s_adj[NB == 4 ? ENUM_STAGE * ENUM_ADJ_WORDS : 1];
  const int wave = threadIdx.x / 64, lane = threadIdx.x & 63;
  const int gi = blockIdx.x;
  if (gi >= a.natoms) return;  // whole workgroup
  const int b = min(a.noff[gi], a.nnb);
  const int n = min(a.noff[gi + 1], a.nnb) - b;
  const bool staged = n <= ENUM_STAGE;
  if (staged)
    for (int k = threadIdx.x; k < n; k += 256) {
      s_vec[k] = a.nvec[b + k];
      s_ref[k] = a.nlist[b + k];
      s_typ[k] = a.ntyp[b + k];
    }
  __syncthreads();
  if (NB == 4 && staged) {
    // one row per thread (rows k, k+256, ...), written by its owner only
    for (int k = threadIdx.x; k < n; k += 256) {
      const double4 vk = s_vec[k];
      for (int w = 0; w < ENUM_ADJ_WORDS; ++w) {
        unsigned long long m = 0;
        for (int l = w * 64; l < min(n, w * 64 + 64); ++l)
          if (l != k && dist3(vk, s_vec[l]) < a.cutoff) m |= 1ull << (l - w * 64);
        s_adj[k * ENUM_ADJ_WORDS + w] = m;
      }
    }
  }
  __syncthreads();
  const int fidx = a.atom_frame[gi];
  const DevFrame &f = a.frames[fidx];
  const uint32_t r1 = make_ref(gi - f.atom_off, 0, 0, 0);
  const int t1 = a.type[gi];
  const int kbase = fidx * a.nkeys_type;
  const int slot = 4 * gi + wave;
  const int out = COUNT ? 0 : a.coff[slot];
  const int npair = n * n, per = (npair + 3) / 4;
  const int qa = min(npair, wave * per), qb = min(npair, qa + per);
  int tot;
  if (staged)
    tot = enum_atom<NB, COUNT>(a, StageLds{s_vec, s_ref, s_typ}, n, qa, qb, !f.small, r1, t1, kbase, out,
                               NB == 4 ? s_adj : nullptr);
  else
    tot = enum_atom<NB, COUNT>(a, StageGlobal{a.nvec + b, a.nlist + b, a.ntyp + b}, n, qa, qb, !f.small, r1, t1,
                               kbase, out, nullptr);
  if (COUNT && lane == 0) a.count[slot] = tot;
}

// Fit statistics over the neighbour list (padded by the reference's 0.3 A):
// the smallest distance of each pair type among the pairs of its 2-body list
// (MIN_FOUND_DIST, src/Cheby.C:769-770), entries closer than cut only.
__global__ void k_pair_min(const int *type, const int *nown, const int *ntyp, const double4 *nvec, const int *ntot,
                           int nnb, const int *int_pair_map, int ntypes, double cut, unsigned long long *pmin) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnb || k >= *ntot) return;
  const double r = nvec[k].w;
  if (!(r < cut)) return;
  unsigned long long *m = pmin + int_pair_map[type[nown[k]] * ntypes + ntyp[k]];
  const unsigned long long b = __builtin_bit_cast(unsigned long long, r);
  if (b < __hip_atomic_load(m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMin(m, b);
}

// Size of the reference's flat 3- or 4-body interaction list of each frame
// (NEIGHBORS::UPDATE_3B/4B_INTERACTION, src/ClassDefs.C:1513-1625): from
// atom i's list at the family cutoff plus padding (half list, every image in
// small cells), the ordered entry pairs (j, k), j != k, parent(j) <=
// parent(k) unless small, with r_jk < cut (3-body) or r_jk <= cut and a
// third entry l != j, k, parent(k), parent(j) <= parent(l) unless small,
// r_jl < cut and r_kl < cut (4-body).  One wave per atom, lanes over (j, k).
template <int NB>
__global__ __launch_bounds__(256) void k_flat_count(const DevFrame *frames, const int *atom_frame, const int *noff,
                                                    const uint32_t *nlist, const double4 *nvec, int natoms, int nnb,
                                                    double cut, unsigned long long *per_frame) {
  const int gi = blockIdx.x * 4 + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (gi >= natoms) return;
  const int b = min(noff[gi], nnb), n = min(noff[gi + 1], nnb) - b;
  const int fidx = atom_frame[gi];
  const bool ordered = !frames[fidx].small;
  unsigned long long cnt = 0;
  for (int q = lane; q < n * n; q += 64) {
    const int j = q / n, k = q - j * n;
    if (j == k) continue;
    const double4 vj = nvec[b + j], vk = nvec[b + k];
    if (!(vj.w < cut) || !(vk.w < cut)) continue;
    const int pj = ref_atom(nlist[b + j]), pk = ref_atom(nlist[b + k]);
    if (ordered && pj > pk) continue;
    const double rjk = dist3(vj, vk);
    if (NB == 3) {
      if (rjk < cut) ++cnt;
      continue;
    }
    if (rjk > cut) continue;
    for (int l = 0; l < n; ++l) {
      if (l == j || l == k) continue;
      const double4 vl = nvec[b + l];
      if (!(vl.w < cut)) continue;
      const int pl = ref_atom(nlist[b + l]);
      if (ordered && (pk > pl || pj > pl)) continue;
      if (dist3(vj, vl) < cut && dist3(vk, vl) < cut) ++cnt;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
  if (lane == 0 && cnt) atomicAdd(per_frame + fidx, cnt);
}

__global__ void k_iota(int *v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = i;
}

template <int NB>
__global__ void k_gather_refs(const uint32_t *refs, const int *idx, int ncl, uint32_t *out) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncl) return;
  const int s = idx[c];
#pragma unroll
  for (int q = 0; q < NB; ++q) out[NB * c + q] = refs[NB * s + q];
}

// boff[q] = first cluster whose key is >= q (keys sorted), q in [0, nkeys].
__global__ void k_bucket_bounds(const int *key, int ncl, int nkeys, int *boff) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > nkeys) return;
  int lo = 0, hi = ncl;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (key[mid] < q) lo = mid + 1;
    else hi = mid;
  }
  boff[q] = lo;
}

// Keys of the slots in [*tot, cap) of a capacity-sized cluster list get the
// sentinel key (one past the last bucket), so that sorting and bucket bounds
// over the whole capacity see only the valid clusters in the buckets.
__global__ void k_pad_keys(int *key, const int *tot, int cap, int sentinel) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cap && i >= *tot) key[i] = sentinel;
}

// Device-side task list of the force kernel: chunks of at most `chunk`
// clusters of one (frame, type) bucket.  k_task_counts gives the chunks per
// bucket (scanned into toff, toff[nkeys] = number of tasks), k_task_fill one
// task per thread.
__global__ void k_task_counts(const int *bounds, int nkeys, int chunk, int *cnt) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nkeys) cnt[q] = (bounds[q + 1] - bounds[q] + chunk - 1) / chunk;
  else if (q == nkeys) cnt[q] = 0;
}

// ------------------------------------------------------------ assembly kernel

__global__ void k_task_fill(const int *bounds, const int *toff, int nkeys, int ntypes, int chunk, int max_tasks,
                            Task *tasks) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= max_tasks || t >= toff[nkeys]) return;
  int lo = 0, hi = nkeys - 1;  // last bucket with toff[q] <= t
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (toff[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  const int j = t - toff[lo];
  const int c0 = bounds[lo] + j * chunk;
  tasks[t] = Task{lo / ntypes, lo % ntypes, 0, 0, c0, min(bounds[lo + 1], c0 + chunk)};
}

// Totals of a call gathered for one copy back: entries of the neighbour list
// and clusters of each family (-1 where a pointer is null).
struct TotalPtrs { const int *p[4]; };
__global__ void k_totals(TotalPtrs t, double *out) {
  const int q = threadIdx.x;
  if (q < 4) out[q] = t.p[q] ? (double)*t.p[q] : -1.0;
}

struct AsmArgs {
  const Task *tasks;
  // bucket bounds of the family: clusters of (frame f, type t) are
  // [bounds[f*ntypes+t], bounds[f*ntypes+t+1]) of refs (null: no clusters)
  const int *bounds;
  int ntypes;
  const DevFrame *frames;
  const double *wrap;
  const uint32_t *refs;     // bucketed cluster refs [ncl*NB]
  const uint32_t *perm;     // packed powers
  const int *sets8;         // set records: 4-body 8 ints (2 * power of pairs 0..5, 0, 0), 3-body 4 ints
  const int *colptr;        // [family ncols + 1]
  const int *col_rel;       // [ntypes] family-relative first column of a type
  const int *col_abs;       // [ntypes] absolute first column of a type
  const double *smin, *smax, *xavg, *xdiff, *lambda;  // [ntypes*NP]
  const double *recip;                                 // [ntypes*NP*3]: 1/lambda, 1/xdiff, 1/smax
  void *A;
  long long ld;
  int transform, fix_type, fcut_type, stress_mode;
  double deriv_const, smooth_dist, fcut_offset;
  int nb_scale;             // index into perm_scale (2, 3, 4)
  int det;                  // waves add to the LDS tile in a fixed order (bitwise-reproducible A)
  int qform;                // 4-body set loops in quotient form (F, Q = D / F tables)
  // 3/4-body with the hand-scheduled set loops: pair tables per neighbour
  // entry (FP64 A: F[VW] then Q or D [VW], doubles; FP32 A: (F, Q or D)
  // float pairs in the register order) and the entry of each cluster pair
  // (-1: build it here); null: every cluster builds its tables
  const void *ptab;
  const int *pidx;
};

template <class T, int N> using vecN = T __attribute__((ext_vector_type(N)));

// Per-pair table of (F[i], D[i]) for powers i = 0..VW-1, packed into one
// vector of 32-bit elements so that a wave-uniform power index is lowered to
// GPR-index-mode moves (s_set_gpr_idx_on + one v_mov per dword) rather than
// to a chain of v_cndmask selects, which LLVM picks for short vectors.
template <class T, int VW, bool SPLIT = (VW > 8)> struct PairTab;
// FP64, VW <= 8: one vector of 4*VW dwords (F lo, F hi, D lo, D hi per power)
template <int VW> struct PairTab<double, VW, false> {
  vecN<uint32_t, 4 * VW> v;
  __device__ __forceinline__ void zero() { v = 0u; }
  __device__ __forceinline__ void set(int i, double f, double d) {
    const uint64_t fb = __builtin_bit_cast(uint64_t, f), db = __builtin_bit_cast(uint64_t, d);
    v[4 * i] = (uint32_t)fb; v[4 * i + 1] = (uint32_t)(fb >> 32);
    v[4 * i + 2] = (uint32_t)db; v[4 * i + 3] = (uint32_t)(db >> 32);
  }
  __device__ __forceinline__ void get(unsigned i, double &f, double &d) const {
    f = __builtin_bit_cast(double, (uint64_t)v[4 * i] | ((uint64_t)v[4 * i + 1] << 32));
    d = __builtin_bit_cast(double, (uint64_t)v[4 * i + 2] | ((uint64_t)v[4 * i + 3] << 32));
  }
};
// FP64, 8 < VW <= 16: separate F and D vectors of 2*VW dwords (register
// vectors are limited to 32 dwords; longer ones would live in scratch)
template <int VW> struct PairTab<double, VW, true> {
  vecN<uint32_t, 2 * VW> F, D;
  __device__ __forceinline__ void zero() { F = 0u; D = 0u; }
  __device__ __forceinline__ void set(int i, double f, double d) {
    const uint64_t fb = __builtin_bit_cast(uint64_t, f), db = __builtin_bit_cast(uint64_t, d);
    F[2 * i] = (uint32_t)fb; F[2 * i + 1] = (uint32_t)(fb >> 32);
    D[2 * i] = (uint32_t)db; D[2 * i + 1] = (uint32_t)(db >> 32);
  }
  __device__ __forceinline__ void get(unsigned i, double &f, double &d) const {
    f = __builtin_bit_cast(double, (uint64_t)F[2 * i] | ((uint64_t)F[2 * i + 1] << 32));
    d = __builtin_bit_cast(double, (uint64_t)D[2 * i] | ((uint64_t)D[2 * i + 1] << 32));
  }
};
template <int VW, bool SPLIT> struct PairTab<float, VW, SPLIT> {
  static constexpr int N = VW < 5 ? 5 : VW;  // >= 10 elements keeps index mode
  vecN<float, 2 * N> v;                      // F, D per power
  __device__ __forceinline__ void zero() { v = 0.0f; }
  __device__ __forceinline__ void set(int i, float f, float d) { v[2 * i] = f; v[2 * i + 1] = d; }
  __device__ __forceinline__ void get(unsigned i, float &f, float &d) const { f = v[2 * i]; d = v[2 * i + 1]; }
};

// 1/r for r > 0 from the hardware reciprocal estimate and two Newton steps
// y <- y + y (1 - r y) (each squares the relative error, so the estimate's
// ~2^-26 ends below one ulp), in 5 VALU instead of a correctly rounded
// division's ~12.
__device__ __forceinline__ double recip_nr(double r) {
  double y = __builtin_amdgcn_rcp(r);
  y = fma(y, fma(-r, y, 1.0), y);
  return fma(y, fma(-r, y, 1.0), y);
}

// Chebyshev tables of one pair: F[i] = fc*T_{i+BASE}(x), D[i] = d/dr of that,
// with set_polys / set_polys_out_of_range / FCUT::get_fcut semantics.
template <class T, int VW, int BASE, class Tab>
__device__ __forceinline__ void pair_table(const AsmArgs &a, double r, double smin, double smax, double xavg,
                                           const double *recip, Tab &P) {
  // recip: 1/lambda, 1/xdiff, 1/smax of the pair type (host-computed: the
  // table is rebuilt for every column tile, and FP64 divisions cost ~10 VALU)
  const double ilambda = recip[0], ixdiff = recip[1], ismax = recip[2];
  double re = r < smin ? smin : r;
  double x, ex = 0.0, dxdr;
  if (a.transform == CG_MORSE) {
    ex = exp(-re * ilambda);
    x = (ex - xavg) * ixdiff;
    dxdr = (-ex * ilambda) * ixdiff;
  } else if (a.transform == CG_INVRSE_R) {
    x = (1.0 / re - xavg) * ixdiff;
    dxdr = -ixdiff / (re * re);
  } else {
    x = (re - xavg) * ixdiff;
    dxdr = ixdiff;
  }
  dxdr *= a.deriv_const;
  double damp = 1.0, shift = 0.0, dscale = 1.0;
  bool zero_d = false;
  if (r < smin) {
    if (a.fix_type == CG_FIX_ZERO_DERIV) zero_d = true;
    else if (a.fix_type == CG_FIX_CONSTANT_DERIV) shift = r - smin;
    else {
      damp = exp((r - smin) / a.smooth_dist);
      shift = a.smooth_dist * (damp - 1.0);
      dscale = damp;
    }
  }
  double fc, fcd;
  if (a.fcut_type == CG_FCUT_CUBIC) {
    double f0 = 1.0 - r * ismax;
    fc = f0 * f0 * f0;
    fcd = f0 * f0 * (-3.0 * ismax);
  } else {
    double th = smax - a.fcut_offset * smax;
    if (r < th) { fc = 1.0; fcd = 0.0; }
    else if (r > smax) { fc = 0.0; fcd = 0.0; }
    else {
      double f0 = (r - th) / (smax - th) * 3.14159265359 + 3.14159265359 / 2.0;
      fc = 0.5 + 0.5 * sin(f0);
      fcd = 0.5 * cos(f0) * (3.14159265359 / (smax - th));
    }
  }
  // T_n by recurrence; dT_n/dr = n * U_{n-1}(x) * dx/dr with U by recurrence
  constexpr int NT = VW + BASE;
  double tn[NT + 1], un[NT + 1];
  tn[0] = 1.0; tn[1] = x;
  un[0] = 1.0; un[1] = 2.0 * x;
#pragma unroll
  for (int n = 2; n <= NT; ++n) {
    tn[n] = 2.0 * x * tn[n - 1] - tn[n - 2];
    un[n] = 2.0 * x * un[n - 1] - un[n - 2];
  }
#pragma unroll
  for (int n = BASE; n < NT; ++n) {
    double tdn = n == 0 ? 0.0 : n * dxdr * un[n - 1];
    if (zero_d) tdn = 0.0;
    const double tv = tn[n] + shift * tdn;
    const double tdv = tdn * dscale;
    P.set(n - BASE, (T)(fc * tv), (T)(fc * tdv + fcd * tv));
  }
}

// 4-body FP64 tables for the hand-scheduled set loops: F and D (product
// form) or F and Q = D / F (quotient form) of the six pairs in four 32-dword
// register blocks held in fixed registers v[128:255] (F of pairs 0-2, F of
// pairs 3-5, D or Q of pairs 0-2, D or Q of pairs 3-5; pair slot of 10
// dwords, entry of power p at 2p).  A wave-uniform power selects an entry
// through GPR-index mode applied to src0 of the arithmetic itself, so a table
// read costs no instruction of its own.
//
// With Q the set loop needs only the product P = prod_k F_k of a set:
// E += P and G_k += Q_k P (= D_k prod_{j != k} F_j).  Multiplication and
// division keep relative errors, so Q_k P carries a few roundings more than
// the direct product however small F_k is; F entries below 1e-60 in
// magnitude (T_n(x) = 0 at a Chebyshev node) are held at 1e-60, which keeps
// Q finite and changes E by at most 1e-60 times the other factors.  The
// quotients of a pair come from one division (the reciprocal of the product
// of its F entries, unwound entry by entry), which the bound keeps clear of
// underflow for up to five entries.
typedef vecN<uint32_t, 32> v32u;
constexpr double kTabFmin = 1e-60;
struct Tab4 {
  v32u FA, FB, DA, DB;
  __device__ __forceinline__ void zero() { FA = 0u; FB = 0u; DA = 0u; DB = 0u; }
};
struct Tab4Pair {
  Tab4 &t;
  int k;
  bool q;  // quotient form: F held at >= kTabFmin (tab4_quotient then turns D into Q)
  __device__ __forceinline__ void set(int i, double f, double d) {
    const double fs = q && fabs(f) < kTabFmin ? kTabFmin : f;
    const uint64_t fb = __builtin_bit_cast(uint64_t, fs), db = __builtin_bit_cast(uint64_t, d);
    const int o = 10 * (k % 3) + 2 * i;
    if (k < 3) {
      t.FA[o] = (uint32_t)fb; t.FA[o + 1] = (uint32_t)(fb >> 32);
      t.DA[o] = (uint32_t)db; t.DA[o + 1] = (uint32_t)(db >> 32);
    } else {
      t.FB[o] = (uint32_t)fb; t.FB[o + 1] = (uint32_t)(fb >> 32);
      t.DB[o] = (uint32_t)db; t.DB[o + 1] = (uint32_t)(db >> 32);
    }
  }
};
__device__ __forceinline__ double tab_get(const v32u &v, int o) {
  return __builtin_bit_cast(double, ((uint64_t)v[o + 1] << 32) | (uint64_t)v[o]);
}
__device__ __forceinline__ void tab_put(v32u &v, int o, double x) {
  const uint64_t b = __builtin_bit_cast(uint64_t, x);
  v[o] = (uint32_t)b; v[o + 1] = (uint32_t)(b >> 32);
}
// D -> Q = D / F for the VW entries of pair k: c_i = F_0 ... F_i, one
// division 1 / c_{VW-1}, then 1 / F_i = (1 / c_i) c_{i-1} and
// 1 / c_{i-1} = (1 / c_i) F_i from the last entry down.
template <int VW> __device__ __forceinline__ void tab4_quotient(Tab4 &t, int k) {
  v32u &Fv = k < 3 ? t.FA : t.FB;
  v32u &Dv = k < 3 ? t.DA : t.DB;
  const int o = 10 * (k % 3);
  double f[VW], c[VW];
#pragma unroll
  for (int i = 0; i < VW; ++i) f[i] = tab_get(Fv, o + 2 * i);
  c[0] = f[0];
#pragma unroll
  for (int i = 1; i < VW; ++i) c[i] = c[i - 1] * f[i];
  double inv = 1.0 / c[VW - 1];
#pragma unroll
  for (int i = VW - 1; i >= 1; --i) {
    tab_put(Dv, o + 2 * i, tab_get(Dv, o + 2 * i) * (inv * c[i - 1]));
    inv *= f[i];
  }
  tab_put(Dv, o, tab_get(Dv, o) * inv);
}

// Pair tables once per neighbour entry of a 3- or 4-body family list (the
// entries closer than the family cutoff): entry e holds F[0..VW) and then
// D[0..VW) (or Q = D / F for the 4-body quotient form) at tab + 2 * VW * e,
// from the same table code as a cluster's own build (pair_table, and for
// 4-body the Tab4 slot with its F bound and tab4_quotient), with the
// parameters of the entry's 2-body pair type (ptp: smin, smax, xavg,
// 1/lambda, 1/xdiff, 1/smax).  A column tile then loads its clusters' tables
// instead of building them: a frame has far fewer pairs within the family
// cutoff than (cluster, pair) slots.  One thread per entry below the device
// total.
template <int VW> struct MemTab {
  double f[VW], d[VW];
  __device__ __forceinline__ void set(int i, double x, double y) { f[i] = x; d[i] = y; }
};
template <int NB, int VW, int BASE>
__global__ void k_pair_tab(AsmArgs a, const double4 *svec, const int *sown, const int *styp, const int *type,
                           const int *int_pair_map, int ntypes_atom, const double *ptp, const int *ntot, int cap,
                           void *tab_out) {
  double *tab = static_cast<double *>(tab_out);
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= min(*ntot, cap)) return;
  const double *p = ptp + 6 * int_pair_map[type[sown[e]] * ntypes_atom + styp[e]];
  double v[2 * VW];
  if constexpr (NB == 4) {
    Tab4 t;
    t.zero();
    Tab4Pair tp{t, 0, a.qform != 0};
    pair_table<double, VW, BASE>(a, svec[e].w, p[0], p[1], p[2], p + 3, tp);
    if (a.qform) tab4_quotient<VW>(t, 0);
#pragma unroll
    for (int i = 0; i < VW; ++i) {
      v[i] = tab_get(t.FA, 2 * i);
      v[VW + i] = tab_get(t.DA, 2 * i);
    }
  } else {
    MemTab<VW> t;
    pair_table<double, VW, BASE>(a, svec[e].w, p[0], p[1], p[2], p + 3, t);
#pragma unroll
    for (int i = 0; i < VW; ++i) {
      v[i] = t.f[i];
      v[VW + i] = t.d[i];
    }
  }
  double2 *o = reinterpret_cast<double2 *>(tab + (size_t)e * (2 * VW));
#pragma unroll
  for (int i = 0; i < VW; ++i) o[i] = make_double2(v[2 * i], v[2 * i + 1]);
}

// Neighbour entry of each pair of a bucketed 3- or 4-body cluster, pairs in
// the assembly's order (0,1) (0,2) [(0,3)] (1,2) [(1,3) (2,3)] of its refs: the
// entry of the lower-indexed atom's list holding the other atom at their
// relative image shift (lists hold j >= i, every image in small cells).  -1
// where that list lacks it (a distance within rounding of the cutoff, or a
// shift the refs cannot encode): the assembly builds that table itself.
// Capacity slots past the device total carry the sentinel key.
template <int NB>
__global__ void k_pair_idx(const uint32_t *refs, const int *keys, int ncl, int nkeys_type, int nkeys,
                           const DevFrame *frames, const int *soff, const uint32_t *slist, int cap, int *pidx) {
  constexpr int NP = NB * (NB - 1) / 2;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncl) return;
  const int key = keys[c];
  int out[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) out[k] = -1;
  if (key >= 0 && key < nkeys) {
    const int aoff = frames[key / nkeys_type].atom_off;
    uint32_t r[NB];
#pragma unroll
    for (int s = 0; s < NB; ++s) r[s] = refs[(size_t)NB * c + s];
    int k = 0;
#pragma unroll
    for (int x = 0; x < NB; ++x)
#pragma unroll
      for (int y = x + 1; y < NB; ++y, ++k) {
        uint32_t ra = r[x], rb = r[y];
        if (ref_atom(rb) < ref_atom(ra)) { const uint32_t t = ra; ra = rb; rb = t; }
        const int dx = ref_sx(rb) - ref_sx(ra), dy = ref_sy(rb) - ref_sy(ra), dz = ref_sz(rb) - ref_sz(ra);
        const int lo = -kRefShiftBias, hi = 15 - kRefShiftBias;
        if (dx < lo || dx > hi || dy < lo || dy > hi || dz < lo || dz > hi) continue;
        const uint32_t want = make_ref(ref_atom(rb), dx, dy, dz);
        const int g = aoff + ref_atom(ra);
        const int b = min(soff[g], cap), en = min(soff[g + 1], cap);
        for (int i = b; i < en; ++i)
          if (slist[i] == want) {
            out[k] = i;
            break;
          }
      }
  }
#pragma unroll
  for (int k = 0; k < NP; ++k) pidx[(size_t)NP * c + k] = out[k];
}

// One allowed set (indices i0..i5 = 2 * power of each pair in SGPRs, index
// mode on): the product runs through the six pair regions, E and G_5 are
// added in the last one, and G_4 .. G_0 walking the regions back, so a set
// takes 11 index changes and 13 VALU.  F_k sits at v[128 + 10k] (k < 3) /
// v[160 + 10(k-3)], Q_k 64 registers higher.
#define CG_SET4Q(i0, i1, i2, i3, i4, i5)                      \
  "s_set_gpr_idx_idx " #i0 "\n\t"                            \
  "v_mov_b64 %[X], v[128:129]\n\t"                           \
  "s_set_gpr_idx_idx " #i1 "\n\t"                            \
  "v_mul_f64 %[X], v[138:139], %[X]\n\t"                     \
  "s_set_gpr_idx_idx " #i2 "\n\t"                            \
  "v_mul_f64 %[X], v[148:149], %[X]\n\t"                     \
  "s_set_gpr_idx_idx " #i3 "\n\t"                            \
  "v_mul_f64 %[X], v[160:161], %[X]\n\t"                     \
  "s_set_gpr_idx_idx " #i4 "\n\t"                            \
  "v_mul_f64 %[X], v[170:171], %[X]\n\t"                     \
  "s_set_gpr_idx_idx " #i5 "\n\t"                            \
  "v_mul_f64 %[P], v[180:181], %[X]\n\t"                     \
  "v_fma_f64 %[E], 1.0, %[P], %[E]\n\t"                      \
  "v_fmac_f64 %[G5], v[244:245], %[P]\n\t"                   \
  "s_set_gpr_idx_idx " #i4 "\n\t"                            \
  "v_fmac_f64 %[G4], v[234:235], %[P]\n\t"                   \
  "s_set_gpr_idx_idx " #i3 "\n\t"                            \
  "v_fmac_f64 %[G3], v[224:225], %[P]\n\t"                   \
  "s_set_gpr_idx_idx " #i2 "\n\t"                            \
  "v_fmac_f64 %[G2], v[212:213], %[P]\n\t"                   \
  "s_set_gpr_idx_idx " #i1 "\n\t"                            \
  "v_fmac_f64 %[G1], v[202:203], %[P]\n\t"                   \
  "s_set_gpr_idx_idx " #i0 "\n\t"                            \
  "v_fmac_f64 %[G0], v[192:193], %[P]\n\t"

// Two sets a, b interleaved region by region: independent product chains,
// then the sums in an order that puts another instruction between the two
// updates of each accumulator.
#define CG_SET4QX2(a0, a1, a2, a3, a4, a5, b0, b1, b2, b3, b4, b5)                   \
  "s_set_gpr_idx_idx " #a0 "\n\t" "v_mov_b64 %[X], v[128:129]\n\t"                 \
  "s_set_gpr_idx_idx " #b0 "\n\t" "v_mov_b64 %[XB], v[128:129]\n\t"                \
  "s_set_gpr_idx_idx " #a1 "\n\t" "v_mul_f64 %[X], v[138:139], %[X]\n\t"           \
  "s_set_gpr_idx_idx " #b1 "\n\t" "v_mul_f64 %[XB], v[138:139], %[XB]\n\t"         \
  "s_set_gpr_idx_idx " #a2 "\n\t" "v_mul_f64 %[X], v[148:149], %[X]\n\t"           \
  "s_set_gpr_idx_idx " #b2 "\n\t" "v_mul_f64 %[XB], v[148:149], %[XB]\n\t"         \
  "s_set_gpr_idx_idx " #a3 "\n\t" "v_mul_f64 %[X], v[160:161], %[X]\n\t"           \
  "s_set_gpr_idx_idx " #b3 "\n\t" "v_mul_f64 %[XB], v[160:161], %[XB]\n\t"         \
  "s_set_gpr_idx_idx " #a4 "\n\t" "v_mul_f64 %[X], v[170:171], %[X]\n\t"           \
  "s_set_gpr_idx_idx " #b4 "\n\t" "v_mul_f64 %[XB], v[170:171], %[XB]\n\t"         \
  "s_set_gpr_idx_idx " #a5 "\n\t" "v_mul_f64 %[P], v[180:181], %[X]\n\t"           \
  "s_set_gpr_idx_idx " #b5 "\n\t" "v_mul_f64 %[PB], v[180:181], %[XB]\n\t"         \
  "v_fmac_f64 %[G5], v[244:245], %[PB]\n\t"                                 \
  "v_fma_f64 %[E], 1.0, %[P], %[E]\n\t"                                     \
  "s_set_gpr_idx_idx " #a4 "\n\t" "v_fmac_f64 %[G4], v[234:235], %[P]\n\t"  \
  "s_set_gpr_idx_idx " #a5 "\n\t" "v_fmac_f64 %[G5], v[244:245], %[P]\n\t"  \
  "v_fma_f64 %[E], 1.0, %[PB], %[E]\n\t"                                    \
  "s_set_gpr_idx_idx " #b4 "\n\t" "v_fmac_f64 %[G4], v[234:235], %[PB]\n\t" \
  "s_set_gpr_idx_idx " #b3 "\n\t" "v_fmac_f64 %[G3], v[224:225], %[PB]\n\t" \
  "s_set_gpr_idx_idx " #a2 "\n\t" "v_fmac_f64 %[G2], v[212:213], %[P]\n\t"  \
  "s_set_gpr_idx_idx " #a3 "\n\t" "v_fmac_f64 %[G3], v[224:225], %[P]\n\t"  \
  "s_set_gpr_idx_idx " #b2 "\n\t" "v_fmac_f64 %[G2], v[212:213], %[PB]\n\t" \
  "s_set_gpr_idx_idx " #b1 "\n\t" "v_fmac_f64 %[G1], v[202:203], %[PB]\n\t" \
  "s_set_gpr_idx_idx " #a0 "\n\t" "v_fmac_f64 %[G0], v[192:193], %[P]\n\t"  \
  "s_set_gpr_idx_idx " #a1 "\n\t" "v_fmac_f64 %[G1], v[202:203], %[P]\n\t"  \
  "s_set_gpr_idx_idx " #b0 "\n\t" "v_fmac_f64 %[G0], v[192:193], %[PB]\n\t"

// Record stream of the 4-body set loops (FP64 and FP32): cp points at the
// column's entry of colptr (first and end set), records are 32 bytes at
// sets8 = (lo, hi).  Two sets per step, their records loaded one pair ahead
// into alternating SGPR blocks s[40:55] / s[60:75]; an odd last set alone.
// The loop reads at most two records past a column's last, which must exist.
// SET1 / SET2 name the one-set and two-set macros.
#define CG_LOOP4(SET1, SET2, TAG)                                                        \
  "s_mov_b32 s59, m0\n\t"                                                                \
  "s_load_dwordx2 s[40:41], %[cp], 0x0\n\t"                                              \
  "s_waitcnt lgkmcnt(0)\n\t"                                                             \
  "s_sub_i32 s58, s41, s40\n\t"                                                          \
  "s_cmp_lt_i32 s58, 1\n\t"                                                              \
  "s_cbranch_scc1 L_" TAG "_done_%=\n\t"                                                 \
  "s_mov_b32 s41, 0\n\t"                                                                 \
  "s_lshl_b64 s[40:41], s[40:41], 5\n\t"                                                 \
  "s_add_u32 s56, s40, %[lo]\n\t"                                                        \
  "s_addc_u32 s57, s41, %[hi]\n\t"                                                       \
  "s_load_dwordx16 s[40:55], s[56:57], 0x0\n\t"                                          \
  "s_waitcnt lgkmcnt(0)\n\t"                                                             \
  "s_set_gpr_idx_on s40, gpr_idx(SRC0)\n"                                                \
  "L_" TAG "_loop_%=:\n\t"                                                               \
  "s_cmp_lt_i32 s58, 2\n\t"                                                              \
  "s_cbranch_scc1 L_" TAG "_tail_a_%=\n\t"                                               \
  "s_load_dwordx16 s[60:75], s[56:57], 0x40\n\t"                                         \
  SET2(s40, s41, s42, s43, s44, s45, s48, s49, s50, s51, s52, s53)                       \
  "s_add_u32 s56, s56, 0x40\n\t"                                                         \
  "s_addc_u32 s57, s57, 0\n\t"                                                           \
  "s_add_i32 s58, s58, -2\n\t"                                                           \
  "s_waitcnt lgkmcnt(0)\n\t"                                                             \
  "s_cmp_lt_i32 s58, 2\n\t"                                                              \
  "s_cbranch_scc1 L_" TAG "_tail_b_%=\n\t"                                               \
  "s_load_dwordx16 s[40:55], s[56:57], 0x40\n\t"                                         \
  SET2(s60, s61, s62, s63, s64, s65, s68, s69, s70, s71, s72, s73)                       \
  "s_add_u32 s56, s56, 0x40\n\t"                                                         \
  "s_addc_u32 s57, s57, 0\n\t"                                                           \
  "s_add_i32 s58, s58, -2\n\t"                                                           \
  "s_waitcnt lgkmcnt(0)\n\t"                                                             \
  "s_branch L_" TAG "_loop_%=\n"                                                         \
  "L_" TAG "_tail_b_%=:\n\t"                                                             \
  "s_cmp_eq_u32 s58, 0\n\t"                                                              \
  "s_cbranch_scc1 L_" TAG "_off_%=\n\t"                                                  \
  SET1(s60, s61, s62, s63, s64, s65)                                                     \
  "s_branch L_" TAG "_off_%=\n"                                                          \
  "L_" TAG "_tail_a_%=:\n\t"                                                             \
  "s_cmp_eq_u32 s58, 0\n\t"                                                              \
  "s_cbranch_scc1 L_" TAG "_off_%=\n\t"                                                  \
  SET1(s40, s41, s42, s43, s44, s45)                                                     \
  "L_" TAG "_off_%=:\n\t"                                                                \
  "s_set_gpr_idx_off\n"                                                                  \
  "L_" TAG "_done_%=:\n\t"                                                               \
  "s_waitcnt lgkmcnt(0)\n\t"                                                             \
  "s_mov_b32 m0, s59\n\t"

#define CG_LOOP4_CLOBBERS                                                                                      \
  "s40", "s41", "s42", "s43", "s44", "s45", "s46", "s47", "s48", "s49", "s50", "s51", "s52", "s53", "s54", "s55", \
      "s56", "s57", "s58", "s59", "s60", "s61", "s62", "s63", "s64", "s65", "s66", "s67", "s68", "s69", "s70",    \
      "s71", "s72", "s73", "s74", "s75", "scc"

// Sum over the allowed sets of one column, quotient form (FP64 F and Q tables and sums).
__device__ __forceinline__ void set_loop4q(const Tab4 &t, const int *cp, uint32_t lo, uint32_t hi, double &E,
                                          double *G) {
  double X, P, XB, PB;
  double G0 = G[0], G1 = G[1], G2 = G[2], G3 = G[3], G4 = G[4], G5 = G[5];
  asm volatile(CG_LOOP4(CG_SET4Q, CG_SET4QX2, "set4q")
      : [E] "+v"(E), [G0] "+v"(G0), [G1] "+v"(G1), [G2] "+v"(G2), [G3] "+v"(G3), [G4] "+v"(G4), [G5] "+v"(G5),
        [X] "=&v"(X), [P] "=&v"(P), [XB] "=&v"(XB), [PB] "=&v"(PB)
      : [cp] "s"(cp), [lo] "s"(lo), [hi] "s"(hi), "{v[128:159]}"(t.FA), "{v[160:191]}"(t.FB), "{v[192:223]}"(t.DA),
        "{v[224:255]}"(t.DB)
      : CG_LOOP4_CLOBBERS);
  G[0] = G0; G[1] = G1; G[2] = G2; G[3] = G3; G[4] = G4; G[5] = G5;
}
#undef CG_SET4Q
#undef CG_SET4QX2

// Product form (F and D tables), for families with few sets per column
// where the quotients would not repay their set-up.  One allowed set
// (indices i0..i5 = 2 * power of each pair in SGPRs):
// E += prod_k F_k, G_k += D_k prod_{j != k} F_j, as L = F0 F1 F2 and
// R = F3 F4 F5 with their partial derivatives, each pair's index region
// entered once.  F_k sits at v[128 + 10k] (k < 3) / v[160 + 10(k-3)], D_k
// 64 registers higher.
#define CG_SET4(i0, i1, i2, i3, i4, i5)                      \
  "s_set_gpr_idx_on " #i0 ", gpr_idx(SRC0)\n\t"              \
  "v_mov_b64 %[X0], v[128:129]\n\t"                          \
  "v_mov_b64 %[Y0], v[192:193]\n\t"                          \
  "s_set_gpr_idx_idx " #i5 "\n\t"                            \
  "v_mov_b64 %[X5], v[180:181]\n\t"                          \
  "v_mov_b64 %[Y5], v[244:245]\n\t"                          \
  "s_set_gpr_idx_idx " #i1 "\n\t"                            \
  "v_mul_f64 %[M], v[202:203], %[X0]\n\t"                    \
  "v_mul_f64 %[N], v[138:139], %[Y0]\n\t"                    \
  "v_mul_f64 %[L], v[138:139], %[X0]\n\t"                    \
  "s_set_gpr_idx_idx " #i4 "\n\t"                            \
  "v_mul_f64 %[M4], v[234:235], %[X5]\n\t"                   \
  "v_mul_f64 %[N4], v[170:171], %[Y5]\n\t"                   \
  "v_mul_f64 %[R], v[170:171], %[X5]\n\t"                    \
  "s_set_gpr_idx_idx " #i2 "\n\t"                            \
  "v_mul_f64 %[DL2], v[212:213], %[L]\n\t"                   \
  "v_mul_f64 %[L], v[148:149], %[L]\n\t"                     \
  "v_mul_f64 %[M], v[148:149], %[M]\n\t"                     \
  "v_mul_f64 %[N], v[148:149], %[N]\n\t"                     \
  "s_set_gpr_idx_idx " #i3 "\n\t"                            \
  "v_mul_f64 %[DR3], v[224:225], %[R]\n\t"                   \
  "v_mul_f64 %[R], v[160:161], %[R]\n\t"                     \
  "v_mul_f64 %[M4], v[160:161], %[M4]\n\t"                   \
  "v_mul_f64 %[N4], v[160:161], %[N4]\n\t"                   \
  "s_set_gpr_idx_off\n\t"                                    \
  "v_fmac_f64 %[E], %[L], %[R]\n\t"                          \
  "v_fmac_f64 %[G0], %[N], %[R]\n\t"                         \
  "v_fmac_f64 %[G1], %[M], %[R]\n\t"                         \
  "v_fmac_f64 %[G2], %[DL2], %[R]\n\t"                       \
  "v_fmac_f64 %[G3], %[DR3], %[L]\n\t"                       \
  "v_fmac_f64 %[G4], %[M4], %[L]\n\t"                        \
  "v_fmac_f64 %[G5], %[N4], %[L]\n\t"

// Sum over the allowed sets of one column: cp points at the column's entry
// of colptr (first and end set), records are 32 bytes at sets8 = (lo, hi);
// the two records after a column's last are read ahead and must exist.  Set
// records are loaded one ahead into alternating SGPR blocks.
__device__ __forceinline__ void set_loop4(const Tab4 &t, const int *cp, uint32_t lo, uint32_t hi, double &E,
                                          double *G) {
  double X0, Y0, X5, Y5, L, M, N, DL2, R, M4, N4, DR3;
  double G0 = G[0], G1 = G[1], G2 = G[2], G3 = G[3], G4 = G[4], G5 = G[5];
  asm volatile(
      "s_mov_b32 s59, m0\n\t"
      "s_load_dwordx2 s[40:41], %[cp], 0x0\n\t"
      "s_waitcnt lgkmcnt(0)\n\t"
      "s_sub_i32 s58, s41, s40\n\t"
      "s_cmp_lt_i32 s58, 1\n\t"
      "s_cbranch_scc1 L_set4_done_%=\n\t"
      "s_mov_b32 s41, 0\n\t"
      "s_lshl_b64 s[40:41], s[40:41], 5\n\t"
      "s_add_u32 s56, s40, %[lo]\n\t"
      "s_addc_u32 s57, s41, %[hi]\n\t"
      "s_load_dwordx8 s[40:47], s[56:57], 0x0\n\t"
      "s_waitcnt lgkmcnt(0)\n"
      "L_set4_%=:\n\t"
      "s_load_dwordx8 s[48:55], s[56:57], 0x20\n\t"
      CG_SET4(s40, s41, s42, s43, s44, s45)
      "s_add_i32 s58, s58, -1\n\t"
      "s_cmp_eq_u32 s58, 0\n\t"
      "s_cbranch_scc1 L_set4_done_%=\n\t"
      "s_waitcnt lgkmcnt(0)\n\t"
      "s_load_dwordx8 s[40:47], s[56:57], 0x40\n\t"
      CG_SET4(s48, s49, s50, s51, s52, s53)
      "s_add_u32 s56, s56, 0x40\n\t"
      "s_addc_u32 s57, s57, 0\n\t"
      "s_add_i32 s58, s58, -1\n\t"
      "s_cmp_eq_u32 s58, 0\n\t"
      "s_waitcnt lgkmcnt(0)\n\t"
      "s_cbranch_scc0 L_set4_%=\n"
      "L_set4_done_%=:\n\t"
      "s_waitcnt lgkmcnt(0)\n\t"
      "s_mov_b32 m0, s59\n\t"
      : [E] "+v"(E), [G0] "+v"(G0), [G1] "+v"(G1), [G2] "+v"(G2), [G3] "+v"(G3), [G4] "+v"(G4), [G5] "+v"(G5),
        [X0] "=&v"(X0), [Y0] "=&v"(Y0), [X5] "=&v"(X5), [Y5] "=&v"(Y5), [L] "=&v"(L), [M] "=&v"(M), [N] "=&v"(N),
        [DL2] "=&v"(DL2), [R] "=&v"(R), [M4] "=&v"(M4), [N4] "=&v"(N4), [DR3] "=&v"(DR3)
      : [cp] "s"(cp), [lo] "s"(lo), [hi] "s"(hi), "{v[128:159]}"(t.FA), "{v[160:191]}"(t.FB), "{v[192:223]}"(t.DA),
        "{v[224:255]}"(t.DB)
      : "s40", "s41", "s42", "s43", "s44", "s45", "s46", "s47", "s48", "s49", "s50", "s51", "s52", "s53", "s54", "s55",
        "s56", "s57", "s58", "s59", "scc");
  G[0] = G0; G[1] = G1; G[2] = G2; G[3] = G3; G[4] = G4; G[5] = G5;
}
#undef CG_SET4

// 3-body FP64 tables for the hand-scheduled set loop (powers below 9): F
// of pair k at v[128 + 18k], D at v[182 + 18k], entry of power p at 2p, in
// three 32-dword blocks and one 12-dword block held in v[128:235].
struct Tab3 {
  v32u B[3];
  vecN<uint32_t, 12> B3;
  __device__ __forceinline__ void zero() { B[0] = 0u; B[1] = 0u; B[2] = 0u; B3 = 0u; }
};
struct Tab3Pair {
  Tab3 &t;
  int k;
  __device__ __forceinline__ void put(int o, double x) {
    const uint64_t b = __builtin_bit_cast(uint64_t, x);
    if (o >= 96) {
      t.B3[o - 96] = (uint32_t)b;
      t.B3[o - 95] = (uint32_t)(b >> 32);
    } else {
      t.B[o >> 5][o & 31] = (uint32_t)b;
      t.B[o >> 5][(o & 31) + 1] = (uint32_t)(b >> 32);
    }
  }
  __device__ __forceinline__ void set(int i, double f, double d) {
    put(18 * k + 2 * i, f);
    put(54 + 18 * k + 2 * i, d);
  }
};

// One allowed 3-body set: E += F0 F1 F2, G_k += D_k prod_{j != k} F_j.
#define CG_SET3(i0, i1, i2)                                  \
  "s_set_gpr_idx_on " #i0 ", gpr_idx(SRC0)\n\t"              \
  "v_mov_b64 %[X0], v[128:129]\n\t"                          \
  "v_mov_b64 %[Y0], v[182:183]\n\t"                          \
  "s_set_gpr_idx_idx " #i1 "\n\t"                            \
  "v_mul_f64 %[L], v[146:147], %[X0]\n\t"                    \
  "v_mul_f64 %[M], v[200:201], %[X0]\n\t"                    \
  "v_mul_f64 %[N], v[146:147], %[Y0]\n\t"                    \
  "s_set_gpr_idx_idx " #i2 "\n\t"                            \
  "v_fmac_f64 %[E], v[164:165], %[L]\n\t"                    \
  "v_fmac_f64 %[G2], v[218:219], %[L]\n\t"                   \
  "v_fmac_f64 %[G1], v[164:165], %[M]\n\t"                   \
  "v_fmac_f64 %[G0], v[164:165], %[N]\n\t"                   \
  "s_set_gpr_idx_off\n\t"

// As set_loop4 for 3-body records of 16 bytes (2 * power of pairs 0..2, 0).
__device__ __forceinline__ void set_loop3(const Tab3 &t, const int *cp, uint32_t lo, uint32_t hi, double &E,
                                          double *G) {
  double X0, Y0, L, M, N;
  double G0 = G[0], G1 = G[1], G2 = G[2];
  asm volatile(
      "s_mov_b32 s59, m0\n\t"
      "s_load_dwordx2 s[40:41], %[cp], 0x0\n\t"
      "s_waitcnt lgkmcnt(0)\n\t"
      "s_sub_i32 s58, s41, s40\n\t"
      "s_cmp_lt_i32 s58, 1\n\t"
      "s_cbranch_scc1 L_set3_done_%=\n\t"
      "s_mov_b32 s41, 0\n\t"
      "s_lshl_b64 s[40:41], s[40:41], 4\n\t"
      "s_add_u32 s56, s40, %[lo]\n\t"
      "s_addc_u32 s57, s41, %[hi]\n\t"
      "s_load_dwordx4 s[40:43], s[56:57], 0x0\n\t"
      "s_waitcnt lgkmcnt(0)\n"
      "L_set3_%=:\n\t"
      "s_load_dwordx4 s[44:47], s[56:57], 0x10\n\t"
      CG_SET3(s40, s41, s42)
      "s_add_i32 s58, s58, -1\n\t"
      "s_cmp_eq_u32 s58, 0\n\t"
      "s_cbranch_scc1 L_set3_done_%=\n\t"
      "s_waitcnt lgkmcnt(0)\n\t"
      "s_load_dwordx4 s[40:43], s[56:57], 0x20\n\t"
      CG_SET3(s44, s45, s46)
      "s_add_u32 s56, s56, 0x20\n\t"
      "s_addc_u32 s57, s57, 0\n\t"
      "s_add_i32 s58, s58, -1\n\t"
      "s_cmp_eq_u32 s58, 0\n\t"
      "s_waitcnt lgkmcnt(0)\n\t"
      "s_cbranch_scc0 L_set3_%=\n"
      "L_set3_done_%=:\n\t"
      "s_waitcnt lgkmcnt(0)\n\t"
      "s_mov_b32 m0, s59\n\t"
      : [E] "+v"(E), [G0] "+v"(G0), [G1] "+v"(G1), [G2] "+v"(G2), [X0] "=&v"(X0), [Y0] "=&v"(Y0), [L] "=&v"(L),
        [M] "=&v"(M), [N] "=&v"(N)
      : [cp] "s"(cp), [lo] "s"(lo), [hi] "s"(hi), "{v[128:159]}"(t.B[0]), "{v[160:191]}"(t.B[1]),
        "{v[192:223]}"(t.B[2]), "{v[224:235]}"(t.B3)
      : "s40", "s41", "s42", "s43", "s44", "s45", "s46", "s47", "s56", "s57", "s58", "s59", "scc");
  G[0] = G0; G[1] = G1; G[2] = G2;
}
#undef CG_SET3

// FP32 tables of the FP32-A path (4-body powers below 5, 3-body below 9):
// F and D (4-body: Q = D / F) of power p of a pair are one 64-bit entry (F
// low) at offset 2p of the pair's slot, so the FP64 set records (2 * power)
// select entries unchanged.  4-body: pair k at v[64 + 10k]
// (k < 3) / v[96 + 10(k-3)]; 3-body: pair k at v[64 + 18k].  The tables sit
// in v[64:127].
struct Tab4f {
  v32u A, B;
  __device__ __forceinline__ void zero() { A = 0u; B = 0u; }
};
// One FP32 4-body table entry: F and D, or in quotient form F held at >= 1e-30
// in magnitude and Q = D / F.
__device__ __forceinline__ void tab4f_entry(bool q, float f, float d, float &fs, float &dq) {
  fs = q && fabsf(f) < 1e-30f ? 1e-30f : f;
  dq = q ? d * __builtin_amdgcn_rcpf(fs) : d;
}
struct Tab4fPair {
  Tab4f &t;
  int k;
  bool q;  // quotient form: F and Q = D / F (F held at >= 1e-30 in magnitude)
  __device__ __forceinline__ void set(int i, float f, float d) {
    float fs, dq;
    tab4f_entry(q, f, d, fs, dq);
    const int o = 10 * (k % 3) + 2 * i;
    if (k < 3) {
      t.A[o] = __builtin_bit_cast(uint32_t, fs); t.A[o + 1] = __builtin_bit_cast(uint32_t, dq);
    } else {
      t.B[o] = __builtin_bit_cast(uint32_t, fs); t.B[o + 1] = __builtin_bit_cast(uint32_t, dq);
    }
  }
};
struct Tab3f {
  v32u A, B;
  __device__ __forceinline__ void zero() { A = 0u; B = 0u; }
};
struct Tab3fPair {
  Tab3f &t;
  int k;
  __device__ __forceinline__ void put(int o, float x) {
    if (o >= 32) t.B[o - 32] = __builtin_bit_cast(uint32_t, x);
    else t.A[o] = __builtin_bit_cast(uint32_t, x);
  }
  __device__ __forceinline__ void set(int i, float f, float d) {
    put(18 * k + 2 * i, f);
    put(18 * k + 2 * i + 1, d);
  }
};

// FP32 pair tables once per neighbour entry (FP32 A): as k_pair_tab, with
// the FP32 table code (pair_table in FP32 and, for 4-body, tab4f_entry's bound
// and quotient), entry e holding the float pairs (F_i, Q_i or D_i) at
// tab + 2 * VW * e, in the order of the registers they are loaded into.
template <int VW> struct MemTabF {
  bool q;
  float v[2 * VW];
  __device__ __forceinline__ void set(int i, float f, float d) {
    if (q) tab4f_entry(true, f, d, v[2 * i], v[2 * i + 1]);
    else { v[2 * i] = f; v[2 * i + 1] = d; }
  }
};
template <int NB, int VW, int BASE>
__global__ void k_pair_tab_f(AsmArgs a, const double4 *svec, const int *sown, const int *styp, const int *type,
                             const int *int_pair_map, int ntypes_atom, const double *ptp, const int *ntot, int cap,
                             void *tab_out) {
  float *tab = static_cast<float *>(tab_out);
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= min(*ntot, cap)) return;
  const double *p = ptp + 6 * int_pair_map[type[sown[e]] * ntypes_atom + styp[e]];
  MemTabF<VW> t{NB == 4 && a.qform != 0};
  pair_table<float, VW, BASE>(a, svec[e].w, p[0], p[1], p[2], p + 3, t);
  float2 *o = reinterpret_cast<float2 *>(tab + (size_t)e * (2 * VW));
#pragma unroll
  for (int i = 0; i < VW; ++i) o[i] = make_float2(t.v[2 * i], t.v[2 * i + 1]);
}

// One allowed 4-body set in FP32, as the FP64 CG_SET4Q (F_k at the even,
// Q_k = D_k / F_k at the odd register of an entry).
#define CG_SET4QF(i0, i1, i2, i3, i4, i5)                     \
  "s_set_gpr_idx_idx " #i0 "\n\t"                            \
  "v_mov_b32 %[X], v64\n\t"                                  \
  "s_set_gpr_idx_idx " #i1 "\n\t"                            \
  "v_mul_f32 %[X], v74, %[X]\n\t"                            \
  "s_set_gpr_idx_idx " #i2 "\n\t"                            \
  "v_mul_f32 %[X], v84, %[X]\n\t"                            \
  "s_set_gpr_idx_idx " #i3 "\n\t"                            \
  "v_mul_f32 %[X], v96, %[X]\n\t"                            \
  "s_set_gpr_idx_idx " #i4 "\n\t"                            \
  "v_mul_f32 %[X], v106, %[X]\n\t"                           \
  "s_set_gpr_idx_idx " #i5 "\n\t"                            \
  "v_mul_f32 %[P], v116, %[X]\n\t"                           \
  "v_fmac_f32 %[E], 1.0, %[P]\n\t"                           \
  "v_fmac_f32 %[G5], v117, %[P]\n\t"                         \
  "s_set_gpr_idx_idx " #i4 "\n\t"                            \
  "v_fmac_f32 %[G4], v107, %[P]\n\t"                         \
  "s_set_gpr_idx_idx " #i3 "\n\t"                            \
  "v_fmac_f32 %[G3], v97, %[P]\n\t"                          \
  "s_set_gpr_idx_idx " #i2 "\n\t"                            \
  "v_fmac_f32 %[G2], v85, %[P]\n\t"                          \
  "s_set_gpr_idx_idx " #i1 "\n\t"                            \
  "v_fmac_f32 %[G1], v75, %[P]\n\t"                          \
  "s_set_gpr_idx_idx " #i0 "\n\t"                            \
  "v_fmac_f32 %[G0], v65, %[P]\n\t"

#define CG_SET4QFX2(a0, a1, a2, a3, a4, a5, b0, b1, b2, b3, b4, b5)                  \
  "s_set_gpr_idx_idx " #a0 "\n\t" "v_mov_b32 %[X], v64\n\t"                        \
  "s_set_gpr_idx_idx " #b0 "\n\t" "v_mov_b32 %[XB], v64\n\t"                       \
  "s_set_gpr_idx_idx " #a1 "\n\t" "v_mul_f32 %[X], v74, %[X]\n\t"                  \
  "s_set_gpr_idx_idx " #b1 "\n\t" "v_mul_f32 %[XB], v74, %[XB]\n\t"                \
  "s_set_gpr_idx_idx " #a2 "\n\t" "v_mul_f32 %[X], v84, %[X]\n\t"                  \
  "s_set_gpr_idx_idx " #b2 "\n\t" "v_mul_f32 %[XB], v84, %[XB]\n\t"                \
  "s_set_gpr_idx_idx " #a3 "\n\t" "v_mul_f32 %[X], v96, %[X]\n\t"                  \
  "s_set_gpr_idx_idx " #b3 "\n\t" "v_mul_f32 %[XB], v96, %[XB]\n\t"                \
  "s_set_gpr_idx_idx " #a4 "\n\t" "v_mul_f32 %[X], v106, %[X]\n\t"                 \
  "s_set_gpr_idx_idx " #b4 "\n\t" "v_mul_f32 %[XB], v106, %[XB]\n\t"               \
  "s_set_gpr_idx_idx " #a5 "\n\t" "v_mul_f32 %[P], v116, %[X]\n\t"                 \
  "s_set_gpr_idx_idx " #b5 "\n\t" "v_mul_f32 %[PB], v116, %[XB]\n\t"               \
  "v_fmac_f32 %[G5], v117, %[PB]\n\t"                                 \
  "v_fmac_f32 %[E], 1.0, %[P]\n\t"                                    \
  "s_set_gpr_idx_idx " #a4 "\n\t" "v_fmac_f32 %[G4], v107, %[P]\n\t"  \
  "s_set_gpr_idx_idx " #a5 "\n\t" "v_fmac_f32 %[G5], v117, %[P]\n\t"  \
  "v_fmac_f32 %[E], 1.0, %[PB]\n\t"                                   \
  "s_set_gpr_idx_idx " #b4 "\n\t" "v_fmac_f32 %[G4], v107, %[PB]\n\t" \
  "s_set_gpr_idx_idx " #b3 "\n\t" "v_fmac_f32 %[G3], v97, %[PB]\n\t"  \
  "s_set_gpr_idx_idx " #a2 "\n\t" "v_fmac_f32 %[G2], v85, %[P]\n\t"   \
  "s_set_gpr_idx_idx " #a3 "\n\t" "v_fmac_f32 %[G3], v97, %[P]\n\t"   \
  "s_set_gpr_idx_idx " #b2 "\n\t" "v_fmac_f32 %[G2], v85, %[PB]\n\t"  \
  "s_set_gpr_idx_idx " #b1 "\n\t" "v_fmac_f32 %[G1], v75, %[PB]\n\t"  \
  "s_set_gpr_idx_idx " #a0 "\n\t" "v_fmac_f32 %[G0], v65, %[P]\n\t"   \
  "s_set_gpr_idx_idx " #a1 "\n\t" "v_fmac_f32 %[G1], v75, %[P]\n\t"   \
  "s_set_gpr_idx_idx " #b0 "\n\t" "v_fmac_f32 %[G0], v65, %[PB]\n\t"

// set_loop4q with FP32 tables and sums.
__device__ __forceinline__ void set_loop4qf(const Tab4f &t, const int *cp, uint32_t lo, uint32_t hi, float &E,
                                           float *G) {
  float X, P, XB, PB;
  float G0 = G[0], G1 = G[1], G2 = G[2], G3 = G[3], G4 = G[4], G5 = G[5];
  asm volatile(CG_LOOP4(CG_SET4QF, CG_SET4QFX2, "set4qf")
      : [E] "+v"(E), [G0] "+v"(G0), [G1] "+v"(G1), [G2] "+v"(G2), [G3] "+v"(G3), [G4] "+v"(G4), [G5] "+v"(G5),
        [X] "=&v"(X), [P] "=&v"(P), [XB] "=&v"(XB), [PB] "=&v"(PB)
      : [cp] "s"(cp), [lo] "s"(lo), [hi] "s"(hi), "{v[64:95]}"(t.A), "{v[96:127]}"(t.B)
      : CG_LOOP4_CLOBBERS);
  G[0] = G0; G[1] = G1; G[2] = G2; G[3] = G3; G[4] = G4; G[5] = G5;
}
#undef CG_SET4QF
#undef CG_SET4QFX2

// One allowed 4-body set in FP32, the products of CG_SET4: (F0, D0) and
// (F5, D5) are copied to v[60:63], the other pairs are read in place.
#define CG_SET4F(i0, i1, i2, i3, i4, i5)                     \
  "s_set_gpr_idx_on " #i0 ", gpr_idx(SRC0)\n\t"              \
  "v_mov_b64 v[60:61], v[64:65]\n\t"                         \
  "s_set_gpr_idx_idx " #i5 "\n\t"                            \
  "v_mov_b64 v[62:63], v[116:117]\n\t"                       \
  "s_set_gpr_idx_idx " #i1 "\n\t"                            \
  "v_mul_f32 %[M], v75, v60\n\t"                             \
  "v_mul_f32 %[N], v74, v61\n\t"                             \
  "v_mul_f32 %[L], v74, v60\n\t"                             \
  "s_set_gpr_idx_idx " #i4 "\n\t"                            \
  "v_mul_f32 %[M4], v107, v62\n\t"                           \
  "v_mul_f32 %[N4], v106, v63\n\t"                           \
  "v_mul_f32 %[R], v106, v62\n\t"                            \
  "s_set_gpr_idx_idx " #i2 "\n\t"                            \
  "v_mul_f32 %[DL2], v85, %[L]\n\t"                          \
  "v_mul_f32 %[L], v84, %[L]\n\t"                            \
  "v_mul_f32 %[M], v84, %[M]\n\t"                            \
  "v_mul_f32 %[N], v84, %[N]\n\t"                            \
  "s_set_gpr_idx_idx " #i3 "\n\t"                            \
  "v_mul_f32 %[DR3], v97, %[R]\n\t"                          \
  "v_mul_f32 %[R], v96, %[R]\n\t"                            \
  "v_mul_f32 %[M4], v96, %[M4]\n\t"                          \
  "v_mul_f32 %[N4], v96, %[N4]\n\t"                          \
  "s_set_gpr_idx_off\n\t"                                    \
  "v_fmac_f32 %[E], %[L], %[R]\n\t"                          \
  "v_fmac_f32 %[G0], %[N], %[R]\n\t"                         \
  "v_fmac_f32 %[G1], %[M], %[R]\n\t"                         \
  "v_fmac_f32 %[G2], %[DL2], %[R]\n\t"                       \
  "v_fmac_f32 %[G3], %[DR3], %[L]\n\t"                       \
  "v_fmac_f32 %[G4], %[M4], %[L]\n\t"                        \
  "v_fmac_f32 %[G5], %[N4], %[L]\n\t"

// set_loop4 with FP32 tables and sums (same record stream and read-ahead).
__device__ __forceinline__ void set_loop4f(const Tab4f &t, const int *cp, uint32_t lo, uint32_t hi, float &E,
                                           float *G) {
  float L, M, N, DL2, R, M4, N4, DR3;
  float G0 = G[0], G1 = G[1], G2 = G[2], G3 = G[3], G4 = G[4], G5 = G[5];
  asm volatile(
      "s_mov_b32 s59, m0\n\t"
      "s_load_dwordx2 s[40:41], %[cp], 0x0\n\t"
      "s_waitcnt lgkmcnt(0)\n\t"
      "s_sub_i32 s58, s41, s40\n\t"
      "s_cmp_lt_i32 s58, 1\n\t"
      "s_cbranch_scc1 L_set4f_done_%=\n\t"
      "s_mov_b32 s41, 0\n\t"
      "s_lshl_b64 s[40:41], s[40:41], 5\n\t"
      "s_add_u32 s56, s40, %[lo]\n\t"
      "s_addc_u32 s57, s41, %[hi]\n\t"
      "s_load_dwordx8 s[40:47], s[56:57], 0x0\n\t"
      "s_waitcnt lgkmcnt(0)\n"
      "L_set4f_%=:\n\t"
      "s_load_dwordx8 s[48:55], s[56:57], 0x20\n\t"
      CG_SET4F(s40, s41, s42, s43, s44, s45)
      "s_add_i32 s58, s58, -1\n\t"
      "s_cmp_eq_u32 s58, 0\n\t"
      "s_cbranch_scc1 L_set4f_done_%=\n\t"
      "s_waitcnt lgkmcnt(0)\n\t"
      "s_load_dwordx8 s[40:47], s[56:57], 0x40\n\t"
      CG_SET4F(s48, s49, s50, s51, s52, s53)
      "s_add_u32 s56, s56, 0x40\n\t"
      "s_addc_u32 s57, s57, 0\n\t"
      "s_add_i32 s58, s58, -1\n\t"
      "s_cmp_eq_u32 s58, 0\n\t"
      "s_waitcnt lgkmcnt(0)\n\t"
      "s_cbranch_scc0 L_set4f_%=\n"
      "L_set4f_done_%=:\n\t"
      "s_waitcnt lgkmcnt(0)\n\t"
      "s_mov_b32 m0, s59\n\t"
      : [E] "+v"(E), [G0] "+v"(G0), [G1] "+v"(G1), [G2] "+v"(G2), [G3] "+v"(G3), [G4] "+v"(G4), [G5] "+v"(G5),
        [L] "=&v"(L), [M] "=&v"(M), [N] "=&v"(N), [DL2] "=&v"(DL2), [R] "=&v"(R), [M4] "=&v"(M4), [N4] "=&v"(N4),
        [DR3] "=&v"(DR3)
      : [cp] "s"(cp), [lo] "s"(lo), [hi] "s"(hi), "{v[64:95]}"(t.A), "{v[96:127]}"(t.B)
      : "v60", "v61", "v62", "v63", "s40", "s41", "s42", "s43", "s44", "s45", "s46", "s47", "s48", "s49", "s50",
        "s51", "s52", "s53", "s54", "s55", "s56", "s57", "s58", "s59", "scc");
  G[0] = G0; G[1] = G1; G[2] = G2; G[3] = G3; G[4] = G4; G[5] = G5;
}
#undef CG_SET4F
#undef CG_LOOP4
#undef CG_LOOP4_CLOBBERS

// One allowed 3-body set in FP32 (products of CG_SET3).
#define CG_SET3F(i0, i1, i2)                                 \
  "s_set_gpr_idx_on " #i0 ", gpr_idx(SRC0)\n\t"              \
  "v_mov_b64 v[60:61], v[64:65]\n\t"                         \
  "s_set_gpr_idx_idx " #i1 "\n\t"                            \
  "v_mul_f32 %[L], v82, v60\n\t"                             \
  "v_mul_f32 %[M], v83, v60\n\t"                             \
  "v_mul_f32 %[N], v82, v61\n\t"                             \
  "s_set_gpr_idx_idx " #i2 "\n\t"                            \
  "v_fmac_f32 %[E], v100, %[L]\n\t"                          \
  "v_fmac_f32 %[G2], v101, %[L]\n\t"                         \
  "v_fmac_f32 %[G1], v100, %[M]\n\t"                         \
  "v_fmac_f32 %[G0], v100, %[N]\n\t"                         \
  "s_set_gpr_idx_off\n\t"

__device__ __forceinline__ void set_loop3f(const Tab3f &t, const int *cp, uint32_t lo, uint32_t hi, float &E,
                                           float *G) {
  float L, M, N;
  float G0 = G[0], G1 = G[1], G2 = G[2];
  asm volatile(
      "s_mov_b32 s59, m0\n\t"
      "s_load_dwordx2 s[40:41], %[cp], 0x0\n\t"
      "s_waitcnt lgkmcnt(0)\n\t"
      "s_sub_i32 s58, s41, s40\n\t"
      "s_cmp_lt_i32 s58, 1\n\t"
      "s_cbranch_scc1 L_set3f_done_%=\n\t"
      "s_mov_b32 s41, 0\n\t"
      "s_lshl_b64 s[40:41], s[40:41], 4\n\t"
      "s_add_u32 s56, s40, %[lo]\n\t"
      "s_addc_u32 s57, s41, %[hi]\n\t"
      "s_load_dwordx4 s[40:43], s[56:57], 0x0\n\t"
      "s_waitcnt lgkmcnt(0)\n"
      "L_set3f_%=:\n\t"
      "s_load_dwordx4 s[44:47], s[56:57], 0x10\n\t"
      CG_SET3F(s40, s41, s42)
      "s_add_i32 s58, s58, -1\n\t"
      "s_cmp_eq_u32 s58, 0\n\t"
      "s_cbranch_scc1 L_set3f_done_%=\n\t"
      "s_waitcnt lgkmcnt(0)\n\t"
      "s_load_dwordx4 s[40:43], s[56:57], 0x20\n\t"
      CG_SET3F(s44, s45, s46)
      "s_add_u32 s56, s56, 0x20\n\t"
      "s_addc_u32 s57, s57, 0\n\t"
      "s_add_i32 s58, s58, -1\n\t"
      "s_cmp_eq_u32 s58, 0\n\t"
      "s_waitcnt lgkmcnt(0)\n\t"
      "s_cbranch_scc0 L_set3f_%=\n"
      "L_set3f_done_%=:\n\t"
      "s_waitcnt lgkmcnt(0)\n\t"
      "s_mov_b32 m0, s59\n\t"
      : [E] "+v"(E), [G0] "+v"(G0), [G1] "+v"(G1), [G2] "+v"(G2), [L] "=&v"(L), [M] "=&v"(M), [N] "=&v"(N)
      : [cp] "s"(cp), [lo] "s"(lo), [hi] "s"(hi), "{v[64:95]}"(t.A), "{v[96:127]}"(t.B)
      : "v60", "v61", "s40", "s41", "s42", "s43", "s44", "s45", "s46", "s47", "s56", "s57", "s58", "s59", "scc");
  G[0] = G0; G[1] = G1; G[2] = G2;
}
#undef CG_SET3F

template <class T> __device__ __forceinline__ void lds_add(T *p, T v) { unsafeAtomicAdd(p, v); }

// LDS tile geometry of k_assemble (shared with the host task builder).
constexpr int SE_REP = 16;  // replicas of each stress / energy row
__host__ __device__ constexpr int tile_stride(int w) { return w | 1; }
__host__ __device__ constexpr size_t tile_bytes(int natoms, int w) {
  return (3 * (size_t)natoms + 7 * SE_REP) * (size_t)tile_stride(w) * sizeof(double);
}

// Instances whose tables need more than ~150 registers run 256 threads and may
// use more than 256 VGPRs (one wave per SIMD); the others are held to two
// waves per SIMD so that 256-thread workgroups can share a CU.
template <int NB, int VW, class T> constexpr bool asm_heavy() { return 2 * (NB * (NB - 1) / 2) * VW * (int)sizeof(T) / 4 > 150; }

// Workgroups go to the eight XCDs round-robin (workgroup b to XCD b % 8).
// Runs of eight consecutive tasks are kept on one XCD instead (runs still go
// round-robin, which keeps the XCDs' loads even): neighbouring column tiles
// of a bucket write adjacent pieces (W x 8 bytes) of the same rows of A, and
// side by side on one XCD its L2 can merge them into whole lines.
__device__ __forceinline__ int xcd_task(int b, int n) {
  constexpr int G = 8, SPAN = 8 * G;
  const int base = b - b % SPAN;
  if (n - base < SPAN) return b;  // the last, partial span keeps its order
  return base + (b & 7) * G + (b >> 3) % G;
}

// GACC: frames too large for the LDS tile accumulate straight into A (zeroed
// first) with global atomics: 12 force, the stress rows and the energy rows
// per (cluster, column).
template <int NB, int VW, int BASE, class T, int BS, bool GACC = false>
__global__ __launch_bounds__(BS, (asm_heavy<NB, VW, T>() ? 1 : 2)) void k_assemble(AsmArgs a) {
  constexpr int NP = NB * (NB - 1) / 2;
  // 4-body with powers below 5 and 3-body with powers below 9 take the
  // hand-scheduled set loops: FP64 tables and sums for FP64 A, FP32 ones for
  // FP32 A (F32A); other kernels compute tables, products and projections in T
  constexpr bool ASM4 = NB == 4 && VW <= 5;
  constexpr bool ASM3 = NB == 3 && VW <= 9;
  constexpr bool F32A = std::is_same<T, float>::value && (ASM4 || ASM3);
  // the set-loop sums of F32A are widened once per column: positions,
  // projection and tile stay in FP64
  using TP = typename std::conditional<ASM4 || ASM3, double, T>::type;
  // The tile is accumulated in FP64 for both precisions: LDS ds_add_f64 is
  // fast on gfx950 while ds_add_f32 stalls the LDS pipeline, and the FP32
  // path then rounds each A entry once, at the write-out.
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double *acc = reinterpret_cast<double *>(smem_raw);
  const Task t = a.tasks[xcd_task((int)blockIdx.x, (int)gridDim.x)];
  const DevFrame &f = a.frames[t.frame];
  const int nf = 3 * f.natoms;
  const int W = t.ncols;
  const int ty = t.type;
  // cluster range of the task's bucket, read from the device bounds
  int cl0 = 0, cl1 = 0;
  if (a.bounds) {
    cl0 = a.bounds[t.frame * a.ntypes + ty];
    cl1 = a.bounds[t.frame * a.ntypes + ty + 1];
  }
  const int ncl = cl1 - cl0;
  if (ncl <= 0) {
    // a bucket without clusters: its tile of A is zero (block-uniform exit)
    T *Aout = reinterpret_cast<T *>(a.A);
    const int nrow = nf + f.stress_rows + f.ener_rows;
    const int cabs = a.col_abs[ty] + t.col0;
    // element i = r * W + cc of the tile, stepped by BS without a division
    const int dr = BS / W, dc = BS - dr * W;
    int r = threadIdx.x / W, cc = threadIdx.x - r * W;
    while (r < nrow) {
      Aout[(f.row_off + r) * a.ld + cabs + cc] = (T)0;
      r += dr;
      cc += dc;
      if (cc >= W) { cc -= W; ++r; }
    }
    return;
  }
  // tile layout: force component c of atom i in column w at
  // acc[(i * Wp + w) * 3 + c] (components are instruction immediates; the odd
  // atom stride 3 * Wp spreads different atoms over distinct LDS banks);
  // then SE_REP replicas of the 6 stress and 1 energy sums, replica q
  // of column c holding its 7 components at se[(q * Wp + c) * 7 + comp]
  // (component offsets are instruction immediates, and the odd replica
  // stride 7 * Wp keeps the 16 replicas on distinct banks)
  const int Wp = tile_stride(W);
  const int nacc = (nf + 7 * SE_REP) * Wp;
  if constexpr (!GACC) {
    for (int i = threadIdx.x; i < nacc; i += BS) acc[i] = 0.0;
    __syncthreads();
  }
  const int colbase = a.col_rel[ty] + t.col0;
  const TP ps = (TP)f.perm_scale[a.nb_scale];
  const bool full_stress = a.stress_mode == CG_STRESS_ALL;
  const bool any_stress = a.stress_mode != CG_STRESS_NONE;
  const int rep = threadIdx.x & (SE_REP - 1);
  double *se = acc + (size_t)nf * Wp;
  // thread i takes clusters [i*R, i*R + R) of the bucket: neighbouring lanes
  // hold clusters of different anchor atoms, so their force-row atomics do
  // not collide on one LDS address
  const int R = (ncl + BS - 1) / BS;
  for (int rnd = 0; rnd < R; ++rnd) {
    const int j = threadIdx.x * R + rnd;
    // a wave whose first lane is past the bucket's end has no cluster in this
    // round (or any later one): it leaves the loop (there is no barrier in it)
    // (in deterministic mode every wave stays for the barriers of each column)
    if (!a.det && (int)(threadIdx.x & ~63u) * R + rnd >= ncl) break;
    const bool valid = j < ncl;
    // lanes past the bucket's end build the tables of its last cluster (all
    // registers then hold defined values without a clearing pass) and drop
    // their sums at the projection
    const int ci = cl0 + min(j, ncl - 1);
    // 4-body FP64 with powers below 5 takes the hand-scheduled set loop
    PairTab<T, VW> tab[ASM4 || ASM3 ? 1 : NP];
    Tab4 tab4;
    Tab3 tab3;
    Tab4f tab4f;
    Tab3f tab3f;
    // positions relative to the cluster's first atom and, per pair,
    // PERM_SCALE / r: pair vector v = d[y] - d[x] (d[0] = 0), force on x is
    // +g v / r, and the stress of the cluster is sum_s F_s (x) d_s
    TP dpos[NB][3], sc[NP];
    int atom[NB];
    // 3/4-body with entry tables: the neighbour entry of each pair, read
    // before the geometry so that the table loads can start early
    constexpr bool PENT = ASM4 || ASM3;
    int pent[PENT ? NP : 1];
#pragma unroll
    for (int k = 0; k < (PENT ? NP : 1); ++k) pent[k] = -1;
    if constexpr (ASM4) {
      if (a.pidx) {
        const int2 *pp = reinterpret_cast<const int2 *>(a.pidx + (size_t)NP * ci);
#pragma unroll
        for (int k = 0; k < NP / 2; ++k) {
          const int2 w = pp[k];
          pent[2 * k] = w.x;
          pent[2 * k + 1] = w.y;
        }
      }
    } else if constexpr (PENT) {
      if (a.pidx) {
#pragma unroll
        for (int k = 0; k < NP; ++k) pent[k] = a.pidx[(size_t)NP * ci + k];
      }
    }
    {
      double p[NB][3];
      dpos[0][0] = dpos[0][1] = dpos[0][2] = (TP)0;
#pragma unroll
      for (int s = 0; s < NB; ++s) {
        uint32_t r = a.refs[(size_t)NB * ci + s];
        atom[s] = ref_atom(r);
        image_pos(f, a.wrap, r, p[s]);
      }
#pragma unroll
      for (int s = 1; s < NB; ++s)
#pragma unroll
        for (int c = 0; c < 3; ++c) dpos[s][c] = (TP)(p[s][c] - p[0][c]);
      int k = 0;
#pragma unroll
      for (int x = 0; x < NB; ++x)
#pragma unroll
        for (int y = x + 1; y < NB; ++y) {
          double v0 = p[y][0] - p[x][0], v1 = p[y][1] - p[x][1], v2 = p[y][2] - p[x][2];
          double r = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
          sc[k] = (TP)((double)ps * recip_nr(r));
          const int q = ty * NP + k;
          if constexpr (ASM4 && F32A) {
            if (pent[k] >= 0) {
              const float2 *src = static_cast<const float2 *>(a.ptab) + (size_t)pent[k] * VW;
              v32u &Fv = k < 3 ? tab4f.A : tab4f.B;
#pragma unroll
              for (int i = 0; i < VW; ++i) {
                const float2 w = src[i];
                Fv[10 * (k % 3) + 2 * i] = __builtin_bit_cast(uint32_t, w.x);
                Fv[10 * (k % 3) + 2 * i + 1] = __builtin_bit_cast(uint32_t, w.y);
              }
            } else {
              Tab4fPair tp{tab4f, k, a.qform != 0};
              pair_table<float, VW, BASE>(a, r, a.smin[q], a.smax[q], a.xavg[q], a.recip + 3 * q, tp);
            }
          } else if constexpr (ASM3 && F32A) {
            Tab3fPair tp{tab3f, k};
            if (pent[k] >= 0) {
              const float2 *src = static_cast<const float2 *>(a.ptab) + (size_t)pent[k] * VW;
#pragma unroll
              for (int i = 0; i < VW; ++i) {
                const float2 w = src[i];
                tp.put(18 * k + 2 * i, w.x);
                tp.put(18 * k + 2 * i + 1, w.y);
              }
            } else {
              pair_table<float, VW, BASE>(a, r, a.smin[q], a.smax[q], a.xavg[q], a.recip + 3 * q, tp);
            }
          } else if constexpr (ASM4) {
            if (pent[k] >= 0) {
              const double2 *src = static_cast<const double2 *>(a.ptab) + (size_t)pent[k] * VW;
              double v[2 * VW];
#pragma unroll
              for (int i = 0; i < VW; ++i) {
                const double2 w = src[i];
                v[2 * i] = w.x;
                v[2 * i + 1] = w.y;
              }
              v32u &Fv = k < 3 ? tab4.FA : tab4.FB;
              v32u &Dv = k < 3 ? tab4.DA : tab4.DB;
#pragma unroll
              for (int i = 0; i < VW; ++i) {
                tab_put(Fv, 10 * (k % 3) + 2 * i, v[i]);
                tab_put(Dv, 10 * (k % 3) + 2 * i, v[VW + i]);
              }
            } else {
              Tab4Pair tp{tab4, k, a.qform != 0};
              pair_table<double, VW, BASE>(a, r, a.smin[q], a.smax[q], a.xavg[q], a.recip + 3 * q, tp);
              if (a.qform) tab4_quotient<VW>(tab4, k);
            }
          } else if constexpr (ASM3) {
            Tab3Pair tp{tab3, k};
            if (pent[k] >= 0) {
              const double2 *src = static_cast<const double2 *>(a.ptab) + (size_t)pent[k] * VW;
              double v[2 * VW];
#pragma unroll
              for (int i = 0; i < VW; ++i) {
                const double2 w = src[i];
                v[2 * i] = w.x;
                v[2 * i + 1] = w.y;
              }
#pragma unroll
              for (int i = 0; i < VW; ++i) tp.set(i, v[i], v[VW + i]);
            } else {
              pair_table<double, VW, BASE>(a, r, a.smin[q], a.smax[q], a.xavg[q], a.recip + 3 * q, tp);
            }
          } else {
            pair_table<T, VW, BASE>(a, r, a.smin[q], a.smax[q], a.xavg[q], a.recip + 3 * q, tab[k]);
          }
          ++k;
        }
    }
    for (int cc = 0; cc < W; ++cc) {
      const int col = colbase + cc;
      double *se_col = se + (size_t)(rep * Wp + cc) * 7;
      TP e = (TP)0, g[NP];
#pragma unroll
      for (int k = 0; k < NP; ++k) g[k] = (T)0;
      if constexpr (F32A) {
        float ef = 0.0f, gf[NP];
#pragma unroll
        for (int k = 0; k < NP; ++k) gf[k] = 0.0f;
        if constexpr (ASM4) {
          if (a.qform)
            set_loop4qf(tab4f, a.colptr + col, (uint32_t)(uintptr_t)a.sets8, (uint32_t)((uintptr_t)a.sets8 >> 32), ef, gf);
          else
            set_loop4f(tab4f, a.colptr + col, (uint32_t)(uintptr_t)a.sets8, (uint32_t)((uintptr_t)a.sets8 >> 32), ef, gf);
        } else
          set_loop3f(tab3f, a.colptr + col, (uint32_t)(uintptr_t)a.sets8, (uint32_t)((uintptr_t)a.sets8 >> 32), ef, gf);
        e = ef;
#pragma unroll
        for (int k = 0; k < NP; ++k) g[k] = gf[k];
      } else if constexpr (ASM4) {
        if (a.qform)
          set_loop4q(tab4, a.colptr + col, (uint32_t)(uintptr_t)a.sets8, (uint32_t)((uintptr_t)a.sets8 >> 32), e, g);
        else
          set_loop4(tab4, a.colptr + col, (uint32_t)(uintptr_t)a.sets8, (uint32_t)((uintptr_t)a.sets8 >> 32), e, g);
      } else if constexpr (ASM3) {
        set_loop3(tab3, a.colptr + col, (uint32_t)(uintptr_t)a.sets8, (uint32_t)((uintptr_t)a.sets8 >> 32), e, g);
      } else {
      const int pb = a.colptr[col], pe = a.colptr[col + 1];
      uint32_t wnext = pb < pe ? a.perm[pb] : 0u;
      for (int q = pb; q < pe; ++q) {
        const uint32_t w = wnext;
        if (q + 1 < pe) wnext = a.perm[q + 1];
        T fv[NP], dv[NP];
#pragma unroll
        for (int k = 0; k < NP; ++k) tab[k].get((w >> (4 * k)) & 15u, fv[k], dv[k]);
        T pre[NP + 1], suf[NP + 1];
        pre[0] = (T)1;
#pragma unroll
        for (int k = 0; k < NP; ++k) pre[k + 1] = pre[k] * fv[k];
        suf[NP] = (T)1;
#pragma unroll
        for (int k = NP - 1; k >= 0; --k) suf[k] = suf[k + 1] * fv[k];
        e += pre[NP];
#pragma unroll
        for (int k = 0; k < NP; ++k) g[k] += dv[k] * (pre[k] * suf[k + 1]);
      }
      }
      // force on each canonical atom: sum over its pairs of +-g v / r (lanes
      // without a cluster compute it too and drop it)
      TP F[NB][3];
#pragma unroll
      for (int s = 0; s < NB; ++s) F[s][0] = F[s][1] = F[s][2] = (TP)0;
      {
        int k = 0;
#pragma unroll
        for (int x = 0; x < NB; ++x)
#pragma unroll
          for (int y = x + 1; y < NB; ++y) {
            const TP gk = g[k] * sc[k];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              const TP v = x == 0 ? dpos[y][c] : dpos[y][c] - dpos[x][c];
              F[x][c] += gk * v;
              F[y][c] -= gk * v;
            }
            ++k;
          }
      }
      if constexpr (GACC) {
        if (valid) {
          T *Ag = reinterpret_cast<T *>(a.A);
          const long long colg = a.col_abs[ty] + t.col0 + cc;
#pragma unroll
          for (int s = 0; s < NB; ++s)
#pragma unroll
            for (int c = 0; c < 3; ++c)
              unsafeAtomicAdd(&Ag[(f.row_off + 3LL * atom[s] + c) * a.ld + colg], (T)F[s][c]);
          const int nsr = f.stress_rows, ner = f.ener_rows;
          // stress row k holds component k (3 rows) or sel9[k] (9 rows), x 1/V
          for (int k = 0; k < nsr; ++k) {
            const int comp = nsr == 9 ? (k == 0 ? 0 : k == 1 || k == 3 ? 3 : k == 2 || k == 6 ? 4 : k == 4 ? 1
                                                  : k == 5 || k == 7 ? 5 : 2) : k;
            const int ia = comp < 3 ? comp : (comp == 5 ? 1 : 0);
            const int ib = comp < 3 ? comp : (comp == 3 ? 1 : 2);
            TP sv = 0;
#pragma unroll
            for (int s = 1; s < NB; ++s) sv += F[s][ia] * dpos[s][ib];
            unsafeAtomicAdd(&Ag[(f.row_off + nf + k) * a.ld + colg], (T)((double)sv * f.inv_vol));
          }
          for (int k = 0; k < ner; ++k)
            unsafeAtomicAdd(&Ag[(f.row_off + nf + nsr + k) * a.ld + colg], (T)(double)(ps * e));
        }
        continue;
      }
      auto emit = [&]() {
#pragma unroll
        for (int s = 0; s < NB; ++s) {
          double *dst = acc + ((size_t)atom[s] * Wp + cc) * 3;
          lds_add(dst, (double)F[s][0]);
          lds_add(dst + 1, (double)F[s][1]);
          lds_add(dst + 2, (double)F[s][2]);
        }
        if (any_stress) {
          const int nsc = full_stress ? 6 : 3;
#pragma unroll
          for (int comp = 0; comp < 6; ++comp) {
            if (comp >= nsc) break;
            const int ia = comp < 3 ? comp : (comp == 5 ? 1 : 0);
            const int ib = comp < 3 ? comp : (comp == 3 ? 1 : 2);
            // -sum_pairs g r_a r_b / r = sum_atoms F_a (p - p_0)_b
            TP sv = 0;
#pragma unroll
            for (int s = 1; s < NB; ++s) sv += F[s][ia] * dpos[s][ib];
            lds_add(&se_col[comp], (double)sv);
          }
        }
        lds_add(&se_col[6], (double)(ps * e));
      };
      if (a.det) {
        // deterministic order: the waves add their column in turn (the lanes
        // of one LDS atomic are served in a fixed order), so the tile, and
        // A, are the same bits on every run
        const int wv = threadIdx.x >> 6;
        for (int q = 0; q < BS / 64; ++q) {
          if (wv == q && valid) emit();
          __syncthreads();
        }
      } else if (valid) {
        emit();
      }
    }
  }
  if constexpr (GACC) return;
  __syncthreads();
  // write the tile: force rows, stress rows (x 1/V), energy rows (x3)
  T *Aout = reinterpret_cast<T *>(a.A);
  const long long r0 = f.row_off;
  const int cabs = a.col_abs[ty] + t.col0;
  {
    // element i = r * W + cc of the tile, stepped by BS without a division
    const int dr = BS / W, dc = BS - dr * W;
    int r = threadIdx.x / W, cc = threadIdx.x - r * W;
    while (r < nf) {
      const int ia = r / 3;
      Aout[(r0 + r) * a.ld + cabs + cc] = (T)acc[(ia * Wp + cc) * 3 + (r - 3 * ia)];
      r += dr;
      cc += dc;
      if (cc >= W) { cc -= W; ++r; }
    }
  }
  const int nsr = f.stress_rows, ner = f.ener_rows;
  const int sel9[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};
  for (int i = threadIdx.x; i < (nsr + ner) * W; i += BS) {
    const int k = i / W, cc = i - k * W;
    int comp;
    double scale;
    if (k < nsr) { comp = nsr == 9 ? sel9[k] : k; scale = f.inv_vol; }
    else { comp = 6; scale = 1.0; }
    double s = 0;
#pragma unroll
    for (int q = 0; q < SE_REP; ++q) s += se[(q * Wp + cc) * 7 + comp];
    Aout[(r0 + nf + k) * a.ld + cabs + cc] = (T)(s * (double)scale);
  }
}

// Force / energy / virial of clusters with fitted coefficients
// (Cheby::Force_all 2-body part, Force_3B, Force_4B; src/Cheby.C:1731-2675).
//
// The reference sums, for every cluster, coef[column(s)] * prod_k F_k[p_k]
// over the allowed power sets s of the cluster type.  Here the coefficients
// of a type are laid out as a dense tensor C[p_0, ..., p_{NP-1}] (N^NP
// entries, zero where a power combination is not allowed), and energy and
// pair gradients are one nested contraction of C with the per-pair tables:
// the two innermost pairs are unrolled (table entries are plain registers,
// C entries are wave-uniform scalar loads), the outer pairs are runtime
// loops with one uniform table lookup per iteration.  Lanes walk clusters
// of one (frame, type) bucket; pair forces are scattered with atomics, and
// energy and virial are reduced over the workgroup.
struct ForceArgs {
  const Task *tasks;
  const int *ntasks;     // device-side task count (the grid may be larger)
  const DevFrame *frames;
  const double *wrap;
  const uint32_t *refs;
  const double *smin, *smax, *xavg, *xdiff, *lambda;
  const double *recip;  // [ntypes*NP*3]: 1/lambda, 1/xdiff, 1/smax
  const double *coef;    // [ntypes * tsize] coefficient tensors
  int tsize;             // N^NP
  double *F;             // [batch atoms * 3]
  double *EV;            // [nframes * 10]: energy, virial xx xy xz yx yy yz zx zy zz
  int transform, fix_type, fcut_type;
  double deriv_const, smooth_dist, fcut_offset;
  double penalty_scale, penalty_dist;
  int nb_scale;
};

// E = sum_p C[p] F0[p], g0 = sum_p C[p] D0[p]
template <int N>
__device__ __forceinline__ void contract1(const PairTab<double, N> *t, const double *__restrict__ C, double &E,
                                          double *g) {
  double e = 0.0, g0 = 0.0;
#pragma unroll
  for (int p = 0; p < N; ++p) {
    double f, d;
    t[0].get(p, f, d);
    const double c = C[p];
    e += c * f;
    g0 += c * d;
  }
  E = e;
  g[0] = g0;
}

// Inner two levels: S = sum_{a,b} C[a*N+b] Fx[a] Fy[b], Tx = sum C Dx Fy, Ty = sum C Fx Dy
template <int N>
__device__ __forceinline__ void contract_inner2(const PairTab<double, N> &tx, const PairTab<double, N> &ty,
                                                const double *__restrict__ C, double &S, double &Tx, double &Ty) {
  double s = 0.0, sx = 0.0, sy = 0.0;
#pragma unroll
  for (int pa = 0; pa < N; ++pa) {
    double s2 = 0.0, t2 = 0.0;
#pragma unroll
    for (int pb = 0; pb < N; ++pb) {
      double f, d;
      ty.get(pb, f, d);
      const double c = C[pa * N + pb];
      s2 += c * f;
      t2 += c * d;
    }
    double fa, da;
    tx.get(pa, fa, da);
    s += fa * s2;
    sy += fa * t2;
    sx += da * s2;
  }
  S = s;
  Tx = sx;
  Ty = sy;
}

// Inner three levels fully unrolled (small orders): S = sum C Fx Fy Fz over
// N^3 entries plus the three partial gradients.
template <int N>
__device__ __forceinline__ void contract_inner3(const PairTab<double, N> &tx, const PairTab<double, N> &ty,
                                                const PairTab<double, N> &tz, const double *__restrict__ C,
                                                double &S, double &Tx, double &Ty, double &Tz) {
  double s = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
#pragma unroll
  for (int pa = 0; pa < N; ++pa) {
    double S2, T2y, T2z;
    contract_inner2<N>(ty, tz, C + pa * N * N, S2, T2y, T2z);
    double fa, da;
    tx.get(pa, fa, da);
    s += fa * S2;
    sx += da * S2;
    sy += fa * T2y;
    sz += fa * T2z;
  }
  S = s;
  Tx = sx;
  Ty = sy;
  Tz = sz;
}

template <int N>
__device__ __forceinline__ void contract3(const PairTab<double, N> *t, const double *__restrict__ C, double &E,
                                          double *g) {
  if constexpr (N <= 8) {  // whole tensor unrolled: one batch of scalar loads
    double S, T0, T1, T2;
    contract_inner3<N>(t[0], t[1], t[2], C, S, T0, T1, T2);
    E = S;
    g[0] = T0; g[1] = T1; g[2] = T2;
    return;
  } else {
    double e = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
    for (int p0 = 0; p0 < N; ++p0) {
      double f0, d0;
      t[0].get(p0, f0, d0);
      double S, T1, T2;
      contract_inner2<N>(t[1], t[2], C + p0 * N * N, S, T1, T2);
      e += f0 * S;
      g0 += d0 * S;
      g1 += f0 * T1;
      g2 += f0 * T2;
    }
    E = e;
    g[0] = g0; g[1] = g1; g[2] = g2;
  }
}

template <int N>
__device__ __forceinline__ void contract6(const PairTab<double, N> *t, const double *__restrict__ C, double &E,
                                          double *g) {
  double e = 0.0, G[6] = {0, 0, 0, 0, 0, 0};
  if constexpr (N <= 5) {  // pairs 3..5 unrolled (N^3 coefficients per outer iteration)
    for (int p0 = 0; p0 < N; ++p0) {
      double f0, d0;
      t[0].get(p0, f0, d0);
      double S1 = 0.0, T1[5] = {0, 0, 0, 0, 0};
      for (int p1 = 0; p1 < N; ++p1) {
        double f1, d1;
        t[1].get(p1, f1, d1);
        double S2 = 0.0, T2[4] = {0, 0, 0, 0};
        for (int p2 = 0; p2 < N; ++p2) {
          double f2, d2;
          t[2].get(p2, f2, d2);
          double S3, T33, T34, T35;
          contract_inner3<N>(t[3], t[4], t[5], C + ((p0 * N + p1) * N + p2) * N * N * N, S3, T33, T34, T35);
          S2 += f2 * S3;
          T2[0] += d2 * S3;
          T2[1] += f2 * T33;
          T2[2] += f2 * T34;
          T2[3] += f2 * T35;
        }
        S1 += f1 * S2;
        T1[0] += d1 * S2;
#pragma unroll
        for (int q = 0; q < 4; ++q) T1[q + 1] += f1 * T2[q];
      }
      e += f0 * S1;
      G[0] += d0 * S1;
#pragma unroll
      for (int q = 0; q < 5; ++q) G[q + 1] += f0 * T1[q];
    }
    E = e;
#pragma unroll
    for (int q = 0; q < 6; ++q) g[q] = G[q];
    return;
  } else {
    for (int p0 = 0; p0 < N; ++p0) {
      double f0, d0;
      t[0].get(p0, f0, d0);
      double S1 = 0.0, T1[5] = {0, 0, 0, 0, 0};  // gradients of pairs 1..5
      for (int p1 = 0; p1 < N; ++p1) {
        double f1, d1;
        t[1].get(p1, f1, d1);
        double S2 = 0.0, T2[4] = {0, 0, 0, 0};    // pairs 2..5
        for (int p2 = 0; p2 < N; ++p2) {
          double f2, d2;
          t[2].get(p2, f2, d2);
          double S3 = 0.0, T3[3] = {0, 0, 0};     // pairs 3..5
          const double *Cb = C + ((p0 * N + p1) * N + p2) * N * N * N;
          for (int p3 = 0; p3 < N; ++p3) {
            double f3, d3;
            t[3].get(p3, f3, d3);
            double S4, T44, T45;
            contract_inner2<N>(t[4], t[5], Cb + p3 * N * N, S4, T44, T45);
            S3 += f3 * S4;
            T3[0] += d3 * S4;
            T3[1] += f3 * T44;
            T3[2] += f3 * T45;
          }
          S2 += f2 * S3;
          T2[0] += d2 * S3;
  #pragma unroll
          for (int q = 0; q < 3; ++q) T2[q + 1] += f2 * T3[q];
        }
        S1 += f1 * S2;
        T1[0] += d1 * S2;
  #pragma unroll
        for (int q = 0; q < 4; ++q) T1[q + 1] += f1 * T2[q];
      }
      e += f0 * S1;
      G[0] += d0 * S1;
  #pragma unroll
      for (int q = 0; q < 5; ++q) G[q + 1] += f0 * T1[q];
    }
    E = e;
  #pragma unroll
    for (int q = 0; q < 6; ++q) g[q] = G[q];
  }
}

constexpr int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }
constexpr int FORCE_LDS_COEF = 4096;  // coefficient tensors up to 32 KB are staged in LDS

template <int NB, int N, int BASE, int BS>
__global__ __launch_bounds__(BS) void k_force(ForceArgs a) {
  constexpr int NP = NB * (NB - 1) / 2;
  constexpr int TS = ipow(N, NP);
  constexpr bool STAGE = TS <= FORCE_LDS_COEF;
  __shared__ double red[10][BS / 64];
  // The workgroup's clusters share one type: its coefficient tensor is read
  // once into LDS (uniform-address reads are broadcasts) instead of through
  // the scalar cache, which thrashes between the tensors of different types.
  __shared__ double sC[STAGE ? TS : 1];
  if ((int)blockIdx.x >= *a.ntasks) return;  // whole workgroup, before any barrier
  const Task t = a.tasks[blockIdx.x];
  const DevFrame &f = a.frames[t.frame];
  const int ty = t.type;
  const double ps = f.perm_scale[a.nb_scale];
  const double *C = a.coef + (size_t)ty * a.tsize;
  if constexpr (STAGE) {
    for (int i = threadIdx.x; i < TS; i += BS) sC[i] = C[i];
    __syncthreads();
    C = sC;
  }
  AsmArgs aa{};
  aa.transform = a.transform; aa.fix_type = a.fix_type; aa.fcut_type = a.fcut_type;
  aa.deriv_const = a.deriv_const; aa.smooth_dist = a.smooth_dist; aa.fcut_offset = a.fcut_offset;
  double ev[10];
#pragma unroll
  for (int q = 0; q < 10; ++q) ev[q] = 0.0;
  const int ci = t.c0 + threadIdx.x;
  if (ci < t.c1) {
    PairTab<double, N> tab[NP];
    double rh[NP][3], rl[NP];
    int atom[NB];
    double p[NB][3];
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      uint32_t r = a.refs[(size_t)NB * ci + q];
      atom[q] = ref_atom(r);
      image_pos(f, a.wrap, r, p[q]);
    }
    {
      int k = 0;
#pragma unroll
      for (int x = 0; x < NB; ++x)
#pragma unroll
        for (int y = x + 1; y < NB; ++y) {
          double v0 = p[y][0] - p[x][0], v1 = p[y][1] - p[x][1], v2 = p[y][2] - p[x][2];
          double r = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
          const double ir = 1.0 / r;
          rh[k][0] = v0 * ir; rh[k][1] = v1 * ir; rh[k][2] = v2 * ir; rl[k] = r;
          const int q = ty * NP + k;
          pair_table<double, N, BASE>(aa, r, a.smin[q], a.smax[q], a.xavg[q], a.recip + 3 * q, tab[k]);
          ++k;
        }
    }
    double e, g[NP];
    if (NP == 1) contract1<N>(tab, C, e, g);
    else if (NP == 3) contract3<N>(tab, C, e, g);
    else contract6<N>(tab, C, e, g);
    e *= ps;
#pragma unroll
    for (int k = 0; k < NP; ++k) g[k] *= ps;
    if (NB == 2 && a.penalty_scale > 0.0) {
      const double sm = a.smin[ty];
      if (rl[0] - a.penalty_dist < sm) {
        const double rp = sm + a.penalty_dist - rl[0];
        g[0] += -3.0 * rp * rp * a.penalty_scale;
        e += rp * rp * rp * a.penalty_scale;
      }
    }
    // per-atom sums of the pair forces first: 3 atomics per atom of the cluster
    double fa[NB][3];
#pragma unroll
    for (int q = 0; q < NB; ++q) fa[q][0] = fa[q][1] = fa[q][2] = 0.0;
    int k = 0;
#pragma unroll
    for (int x = 0; x < NB; ++x)
#pragma unroll
      for (int y = x + 1; y < NB; ++y) {
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
          const double d = g[k] * rh[k][cc];
          fa[x][cc] += d;
          fa[y][cc] -= d;
        }
#pragma unroll
        for (int u = 0; u < 3; ++u)
#pragma unroll
          for (int v = 0; v < 3; ++v) ev[1 + 3 * u + v] -= g[k] * rl[k] * rh[k][u] * rh[k][v];
        ++k;
      }
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      double *Fq = a.F + 3 * (size_t)(f.atom_off + atom[q]);
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) atomicAdd(Fq + cc, fa[q][cc]);
    }
    ev[0] += e;
  }
  // workgroup reduction of energy + virial
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
#pragma unroll
  for (int q = 0; q < 10; ++q) {
    double v = ev[q];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (lane == 0) red[q][wave] = v;
  }
  __syncthreads();
  if (threadIdx.x < 10) {
    double v = 0.0;
    for (int w = 0; w < BS / 64; ++w) v += red[threadIdx.x][w];
    atomicAdd(&a.EV[t.frame * 10 + threadIdx.x], v);
  }
}

// Energy-offset columns: atom counts on energy rows (A_MAT::add_col_of_ones).
template <class T>
__global__ void k_offsets(const DevFrame *frames, const int *type, int nframes, int ntypes, int col0, T *A,
                          long long ld) {
  int fi = blockIdx.x;
  if (fi >= nframes) return;
  const DevFrame &f = frames[fi];
  __shared__ int cnt[CG_MAX_ATOM_TYPES];
  if (threadIdx.x < CG_MAX_ATOM_TYPES) cnt[threadIdx.x] = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < f.natoms; i += blockDim.x) atomicAdd(&cnt[type[f.atom_off + i]], 1);
  __syncthreads();
  // every row of the frame: zero on force and stress rows, the atom counts
  // on the energy rows
  const int nrows = 3 * f.natoms + f.stress_rows + f.ener_rows;
  const int r_ener = 3 * f.natoms + f.stress_rows;
  for (int k = threadIdx.x; k < nrows * ntypes; k += blockDim.x) {
    const int rr = k / ntypes, t = k - rr * ntypes;
    A[(f.row_off + rr) * ld + col0 + t] = rr >= r_ener ? (T)cnt[t] : (T)0;
  }
}

// ------------------------------------------------------------ host side
static void upload(Ctx &c, void *dst, const void *src, size_t bytes) {
  CG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c.stream));
}

template <class T> static T *dalloc_copy(const std::vector<T> &v) {
  T *p = nullptr;
  size_t n = std::max<size_t>(v.size(), 1);
  CG_HIP(hipMalloc(&p, n * sizeof(T)));
  if (!v.empty()) CG_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return p;
}

static void build_family(Ctx &c, int which) {
  const cg_model &m = c.m;
  DevFamily &F = c.fam[which];
  std::vector<int> colptr, col_rel, col_abs;
  std::vector<uint32_t> perm;
  std::vector<double> smin, smax, xavg, xdiff, lam;
  if (which == 0) {
    F.np = 1; F.nb = 2; F.ntypes = m.npairs; F.ncols = m.n2b_cols; F.max_pow = m.order2b;
    F.cutoff = m.max_cutoff_2b;
    for (int p = 0; p < m.npairs; ++p) {
      col_rel.push_back(p * m.order2b);
      col_abs.push_back(p * m.order2b);
      F.h_col_offset.push_back(p * m.order2b);
      F.h_ncols.push_back(m.order2b);
      F.h_excluded.push_back(0);
      for (int i = 0; i < m.order2b; ++i) { colptr.push_back((int)perm.size()); perm.push_back((uint32_t)i); }
      smin.push_back(m.pair_smin[p]); smax.push_back(m.pair_smax[p]);
      xavg.push_back(m.pair_xavg[p]); xdiff.push_back(m.pair_xdiff[p]); lam.push_back(m.pair_lambda[p]);
    }
    colptr.push_back((int)perm.size());
  } else {
    const cg_cluster_set &cs = which == 1 ? m.c3 : m.c4;
    F.np = cs.npairs; F.nb = cs.nbody; F.ntypes = cs.ntypes; F.ncols = cs.ncols; F.max_pow = cs.order;
    F.cutoff = cs.max_cutoff;
    int fam_base = cs.ntypes ? cs.col_offset[0] : 0;
    for (int t = 0; t < cs.ntypes; ++t) {
      col_rel.push_back(cs.col_offset[t] - fam_base);
      col_abs.push_back(cs.col_offset[t]);
      F.h_col_offset.push_back(cs.col_offset[t]);
      F.h_ncols.push_back(cs.n_unique[t]);
      F.h_excluded.push_back(cs.excluded[t]);
      int prev = -1;
      for (int i = cs.perm_offset[t]; i < cs.perm_offset[t + 1]; ++i) {
        int col = cs.perm_col[i];
        if (col != prev) {
          if (col != prev + 1) throw Error(CG_ERR_INPUT, "allowed powers are not grouped by column");
          colptr.push_back((int)perm.size());
          prev = col;
        }
        uint32_t w = 0;
        for (int k = 0; k < cs.npairs; ++k) {
          int p = cs.perm_powers[(size_t)i * cs.npairs + k];
          if (p < 0 || p > 15) throw Error(CG_ERR_UNSUPPORTED, "power > 15 is not supported");
          w |= uint32_t(p) << (4 * k);
        }
        perm.push_back(w);
      }
      if (prev + 1 != cs.n_unique[t]) throw Error(CG_ERR_INPUT, "column count mismatch");
      for (int k = 0; k < cs.npairs; ++k) {
        smin.push_back(cs.s_min[t * cs.npairs + k]); smax.push_back(cs.s_max[t * cs.npairs + k]);
        xavg.push_back(cs.x_avg[t * cs.npairs + k]); xdiff.push_back(cs.x_diff[t * cs.npairs + k]);
        lam.push_back(m.pair_lambda[cs.pair_type[t * cs.npairs + k]]);
      }
    }
    colptr.push_back((int)perm.size());
    // Type-tuple tables for the enumeration: the reference's INT_MAP,
    // INT_MAP atom and pair maps are indexed by sum_q (t_q+1) 10^q; here the
    // tuple index is dense, sum_q t_q ntypes^q, and the pair map is resolved
    // to the canonical pair's S_MAXIM.
    const int nt = m.ntypes, nb = cs.nbody, np = cs.npairs;
    int ntup = 1;
    for (int q = 0; q < nb; ++q) ntup *= nt;
    std::vector<int> cct(ntup, -1), cperm((size_t)ntup * nb, 0), cslot((size_t)ntup * np, 0);
    std::vector<double> csmax((size_t)ntup * np, -1.0);
    for (int cid = 0; cid < ntup; ++cid) {
      int id = 0, w = 1, r = cid;
      for (int q = 0; q < nb; ++q) {
        id += w * (r % nt + 1);
        r /= nt;
        w *= CG_MAX_ATOM_TYPES;
      }
      if (id >= cs.id_size) throw Error(CG_ERR_INPUT, "cluster id table too small");
      const int ctp = cs.id_map[id];
      cct[cid] = ctp;
      if (ctp < 0) continue;
      for (int q = 0; q < nb; ++q) cperm[(size_t)cid * nb + q] = cs.id_atom_perm[(size_t)id * nb + q];
      for (int pf = 0; pf < np; ++pf) {
        const int cp = cs.id_pair_map[(size_t)id * np + pf];
        if (cp < 0 || cp >= np) throw Error(CG_ERR_INPUT, "bad cluster pair map");
        csmax[(size_t)cid * np + pf] = cs.s_max[ctp * np + cp];
        cslot[(size_t)cid * np + pf] = cp;
      }
    }
    F.cct = dalloc_copy(cct);
    F.cperm = dalloc_copy(cperm);
    F.cslot = dalloc_copy(cslot);
    F.csmax = dalloc_copy(csmax);
  }
  F.colptr = dalloc_copy(colptr);
  F.perm = dalloc_copy(perm);
  if (F.nb == 3 && F.max_pow <= 9) {
    // records of the hand-scheduled 3-body set loop, plus the two it reads ahead
    std::vector<int> s4((perm.size() + 2) * 4, 0);
    for (size_t i = 0; i < perm.size(); ++i)
      for (int k = 0; k < 3; ++k) {
        const int pw = (perm[i] >> (4 * k)) & 15u;
        if (pw >= F.max_pow) throw Error(CG_ERR_INPUT, "3-body power out of range");
        s4[4 * i + k] = 2 * pw;
      }
    F.sets8 = dalloc_copy(s4);
  }
  if (F.nb == 4 && F.max_pow <= 5) {
    // The quotient form saves 12 VALU per set and costs one division and
    // ~4 VALU per table entry for every (cluster, column tile): families
    // averaging 6 or more sets per column take it.
    F.qform = perm.size() >= 6 * (size_t)F.ncols ? 1 : 0;
    // records of the hand-scheduled 4-body set loop, plus the two it reads ahead
    std::vector<int> s8((perm.size() + 2) * 8, 0);
    for (size_t i = 0; i < perm.size(); ++i)
      for (int k = 0; k < 6; ++k) {
        const int pw = (perm[i] >> (4 * k)) & 15u;
        if (pw >= F.max_pow) throw Error(CG_ERR_INPUT, "4-body power out of range");
        s8[8 * i + k] = 2 * pw;
      }
    F.sets8 = dalloc_copy(s8);
  }
  F.smin = dalloc_copy(smin); F.smax = dalloc_copy(smax);
  F.xavg = dalloc_copy(xavg); F.xdiff = dalloc_copy(xdiff); F.lambda = dalloc_copy(lam);
  std::vector<double> recip(3 * smin.size());
  for (size_t q = 0; q < smin.size(); ++q) {
    recip[3 * q] = 1.0 / lam[q];
    recip[3 * q + 1] = 1.0 / xdiff[q];
    recip[3 * q + 2] = 1.0 / smax[q];
  }
  F.recip = dalloc_copy(recip);
  if ((F.nb == 4 && F.max_pow <= 5) || (F.nb == 3 && F.max_pow <= 9)) {
    // Pair tables per neighbour entry need one parameter set per 2-body pair
    // type: every (cluster type, pair slot) of a pair type must carry the
    // same cutoffs (SPECIFIC 3B/4B cutoffs may differ per cluster type, and
    // then the tables stay per cluster).
    const cg_cluster_set &cs = F.nb == 4 ? m.c4 : m.c3;
    std::vector<double> ptp(6 * (size_t)std::max(m.npairs, 1), 0.0);
    std::vector<char> seen(std::max(m.npairs, 1), 0);
    bool same = true;
    for (int q = 0; q < (int)smin.size() && same; ++q) {
      const int pt = cs.pair_type[q];
      if (pt < 0 || pt >= m.npairs) { same = false; break; }
      const double v[6] = {smin[q], smax[q], xavg[q], recip[3 * q], recip[3 * q + 1], recip[3 * q + 2]};
      if (!seen[pt]) {
        seen[pt] = 1;
        for (int i = 0; i < 6; ++i) ptp[6 * pt + i] = v[i];
      } else {
        for (int i = 0; i < 6; ++i) same = same && ptp[6 * pt + i] == v[i];
      }
    }
    // 3-body families with few sets per column leave too little work per
    // (cluster, column tile) to hide the table loads (C/H/N/O batch, 1.7
    // sets per column: 16.8 -> 18.6 ms with them, carbon, 4.5: 5.1 -> 4.8 ms)
    const bool enough = F.nb == 4 || perm.size() >= 3 * (size_t)F.ncols;
    F.pcache = same && enough ? 1 : 0;
    if (same) F.ptp = dalloc_copy(ptp);
  }
  std::vector<int> both(col_rel);
  both.insert(both.end(), col_abs.begin(), col_abs.end());
  F.col_offset = dalloc_copy(both);  // [0,ntypes): relative, [ntypes,2*ntypes): absolute
  F.ncol_type = dalloc_copy(F.h_ncols);
}

// Tile configurations of the assembly: threads per workgroup, LDS budget of a
// workgroup's tile, and the most columns a tile may take.  0: one workgroup
// of 512 per CU with the widest tile; 1: one of 256 (fewer idle lanes on
// small (frame, type) buckets); 2: two of 256 per CU, tiles half as wide;
// 3 and 4: 0 and 2 with at most 16 columns per tile (more, shorter tasks for
// small frames).  cg_autotune picks one per (family, frame-size class).
struct TileCfg { int bs; size_t lds; int wcap; };
static constexpr TileCfg kTileCfg[] = {{512, 150 * 1024, 64}, {256, 150 * 1024, 64}, {256, 75 * 1024, 64},
                                       {512, 150 * 1024, 16}, {256, 75 * 1024, 16}};
constexpr int kNumTileCfg = sizeof(kTileCfg) / sizeof(kTileCfg[0]);
static_assert(kNumTileCfg == CG_TILE_CONFIGS, "tile configurations");

// Frame-size class of the tile configuration: atoms below 128, 512, 1024, above.
static int tile_class(int natoms) { return natoms < 128 ? 0 : natoms < 512 ? 1 : natoms < 1024 ? 2 : 3; }

// Choices of cg_autotune kept for the life of the process, per model
// (species, orders, columns), family and frame-size class; contexts created
// later for the same model start from them.
static std::mutex g_tile_mu;
static std::map<std::string, int> g_tile_cache;
static std::string tile_key(const cg_model &m, int w, int cls) {
  std::string k;
  for (int t = 0; t < m.ntypes; ++t) {
    if (m.type_names) k.append(m.type_names + 8 * t, strnlen(m.type_names + 8 * t, 8));
    k += ',';
  }
  k += std::to_string(m.order2b) + ',' + std::to_string(m.c3.order) + ',' + std::to_string(m.c4.order) + ',' +
       std::to_string(m.ncols) + '|' + std::to_string(w) + '|' + std::to_string(cls);
  return k;
}

// Which kernel instance handles a family.
template <class T>
static void launch_family(Ctx &c, int which, size_t ntasks, const Task *d_tasks, const DevFrame *d_frames,
                          const uint32_t *refs, const int *bounds, size_t lds_bytes, void *A, long long ld,
                          int want_bs, hipStream_t stream, bool gacc = false, const void *ptab = nullptr,
                          const int *pidx = nullptr) {
  const DevFamily &F = c.fam[which];
  AsmArgs a{};
  a.tasks = d_tasks;
  a.bounds = bounds;
  a.ntypes = F.ntypes;
  a.det = c.asm_det ? 1 : 0;
  a.frames = d_frames;
  a.wrap = c.b_wrap.as<double>(0);
  a.refs = refs;
  a.perm = F.perm;
  a.sets8 = F.sets8;
  a.colptr = F.colptr;
  a.col_rel = F.col_offset;
  a.col_abs = F.col_offset + F.ntypes;
  a.smin = F.smin; a.smax = F.smax; a.xavg = F.xavg; a.xdiff = F.xdiff; a.lambda = F.lambda; a.recip = F.recip;
  a.A = A;
  a.ld = ld;
  a.transform = c.m.transform; a.fix_type = c.m.fix_type; a.fcut_type = c.m.fcut_type;
  a.stress_mode = c.m.fit_stress;
  a.deriv_const = c.m.deriv_const; a.smooth_dist = c.m.smooth_dist; a.fcut_offset = c.m.fcut_offset;
  a.nb_scale = F.nb;
  a.qform = F.qform;
  a.ptab = ptab;
  a.pidx = ptab ? pidx : nullptr;
  dim3 grid((unsigned)ntasks), block(512);
  (void)block;
  if (ntasks == 0) return;
  if (which == 0) want_bs = 512;
#define CG_LAUNCH_BS(NBv, VWv, BASEv, BSv)                                                                  \
  do {                                                                                                      \
    auto kern = k_assemble<NBv, VWv, BASEv, T, BSv>;                                                        \
    block = dim3(BSv);                                                                                      \
    CG_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes)); \
    hipLaunchKernelGGL(kern, grid, block, lds_bytes, stream, a);                                            \
    CG_HIP(hipGetLastError());                                                                              \
  } while (0)
  // register-heavy instances run 256 threads whatever the configuration;
  // global accumulation (frames beyond the LDS tile) runs 256 threads
#define CG_LAUNCH(NBv, VWv, BASEv)                                                                          \
  do {                                                                                                      \
    constexpr bool heavy = asm_heavy<NBv, VWv, T>();                                                      \
    if (gacc) {                                                                                             \
      auto kern = k_assemble<NBv, VWv, BASEv, T, 256, true>;                                                \
      hipLaunchKernelGGL(kern, grid, dim3(256), 0, stream, a);                                              \
      CG_HIP(hipGetLastError());                                                                            \
    } else if constexpr (heavy || NBv == 2) {                                                               \
      CG_LAUNCH_BS(NBv, VWv, BASEv, heavy ? 256 : 512);                                                     \
    } else {                                                                                                \
      if (want_bs == 256) CG_LAUNCH_BS(NBv, VWv, BASEv, 256);                                               \
      else CG_LAUNCH_BS(NBv, VWv, BASEv, 512);                                                              \
    }                                                                                                       \
  } while (0)
  if (which == 0) {
    if (F.max_pow <= 16) CG_LAUNCH(2, 16, 1);
    else throw Error(CG_ERR_UNSUPPORTED, "2-body order above 16 is not supported");
  } else if (which == 1) {
    if (F.max_pow <= 4) CG_LAUNCH(3, 4, 0);
    else if (F.max_pow <= 8) CG_LAUNCH(3, 8, 0);
    else if (F.max_pow <= 9) CG_LAUNCH(3, 9, 0);
    else if (F.max_pow <= 12) CG_LAUNCH(3, 12, 0);
    else if (F.max_pow <= 16) CG_LAUNCH(3, 16, 0);
    else throw Error(CG_ERR_UNSUPPORTED, "3-body order above 16 is not supported");
  } else {
    if (F.max_pow <= 4) CG_LAUNCH(4, 4, 0);
    else if (F.max_pow <= 5) CG_LAUNCH(4, 5, 0);
    else if (F.max_pow <= 8) CG_LAUNCH(4, 8, 0);
    else throw Error(CG_ERR_UNSUPPORTED, "4-body order above 8 is not supported");
  }
#undef CG_LAUNCH
#undef CG_LAUNCH_BS
}

static size_t exclusive_scan(Ctx &c, const int *in, int *out, int n) {
  size_t tmp = 0;
  CG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, n + 1, c.stream));
  void *t = c.b_scan.get(tmp);
  if (!t) throw Error(CG_ERR_MEMORY, "scan scratch");
  CG_HIP(hipcub::DeviceScan::ExclusiveSum(t, tmp, in, out, n + 1, c.stream));
  return tmp;
}

static int read_int(Ctx &c, const int *d) {
  int v = 0;
  CG_HIP(hipMemcpyAsync(&v, d, sizeof(int), hipMemcpyDeviceToHost, c.stream));
  CG_HIP(hipStreamSynchronize(c.stream));
  return v;
}

struct FamOut {
  int ncl = 0;                 // clusters (exact mode) or capacity of the list (capacity mode)
  uint32_t *refs = nullptr;
  std::vector<int> boff;       // host bucket bounds (exact mode only)
  int *d_bounds = nullptr;     // device bucket bounds [nkeys + 1]
  const int *keys = nullptr;   // device keys of the bucketed clusters (sentinel nkeys past the total)
  int nkeys = 0;
};

// A 3- or 4-body family's neighbour list (its sublist, or the full list when
// its cutoff is the list cutoff): offsets per atom, refs, vectors, owner
// atoms, target types, the device total and the capacity.
struct FamList {
  const int *off = nullptr;
  const uint32_t *list = nullptr;
  const double4 *vec = nullptr;
  const int *own = nullptr, *typ = nullptr;
  const int *tot = nullptr;
  int cap = 0;
};

struct Batch {
  std::vector<DevFrame> hf;
  const cg_frame *frames = nullptr;  // host frames of the call
  int nframes = 0;
  int natoms = 0;
  long long rows = 0;
  int nbins_total = 0;         // cell-list bins over all frames
  double mx = 0.0;             // list cutoff (largest family cutoff)
  DevFrame *d_frames = nullptr;
  int *d_af = nullptr;         // frame of each atom
  int *d_type = nullptr;
  double *d_xyz = nullptr, *d_wrap = nullptr;
  FamOut fo[3];
  FamList fl[3];               // 3- and 4-body family lists (entries 1, 2)
  TotalPtrs tot{};             // device totals: neighbour entries, clusters per family
};

// Zero n ints / doubles with a kernel (not hipMemsetAsync, which is kept out
// of captured graphs).
__global__ void k_zero_i(int *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}
__global__ void k_zero_d(double *p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}
static void zero_i(Ctx &c, int *p, int n) {
  if (n > 0) hipLaunchKernelGGL(k_zero_i, dim3((n + 255) / 256), dim3(256), 0, c.stream, p, n);
}
static void zero_d(Ctx &c, double *p, size_t n) {
  if (n > 0) hipLaunchKernelGGL(k_zero_d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, p, n);
}

// Upload a batch of frames, wrap atoms, build neighbour lists and enumerate
// + bucket the clusters of every family.  md != 0 selects the force-evaluation
// pair set (every pair below S_MAXIM, Cheby::Force_all) instead of the
// least-squares one (S_MINIM < r < S_MAXIM, Cheby::Deriv_2B).
// Host part of a call: frame table (geometry, row offsets, small-cell and
// cell-list decisions), batch arrays, and their upload.
// The reference pads its neighbour lists by 0.3 A (NEIGHBORS::RCUT_PADDING,
// src/ClassDefs.C:16); the fit statistics count over padded lists.
constexpr double kListPadding = 0.3;

static void prepare_host(Ctx &c, const cg_frame *frames, int nframes, Batch &B, bool sticky) {
  std::vector<DevFrame> &hf = B.hf;
  B.frames = frames;
  const cg_model &m = c.m;
  if (nframes <= 0) throw Error(CG_ERR_ARG, "no frames");
  // ---- host frame table
  hf.assign(nframes, DevFrame{});
  std::vector<int> atom_frame;
  std::vector<double> xyz;
  std::vector<int> type;
  long long rows = 0;
  int natoms = 0;
  int nbins_total = 0;
  double mx = m.max_cutoff_2b;
  if (m.c3.nbody) mx = std::max(mx, m.c3.max_cutoff);
  if (m.c4.nbody) mx = std::max(mx, m.c4.max_cutoff);
  // list cutoff: the largest family cutoff, padded when a fit gathers statistics
  const double list_rc = mx + ((sticky && c.stats.on) ? kListPadding : 0.0);
  for (int i = 0; i < nframes; ++i) {
    const cg_frame &fr = frames[i];
    if (fr.natoms <= 0 || fr.natoms >= (1 << 20)) throw Error(CG_ERR_ARG, "bad atom count");
    DevFrame &d = hf[i];
    host_cell(fr.cell, d);
    d.natoms = fr.natoms;
    d.atom_off = natoms;
    d.row_off = rows;
    d.stress_rows = (m.fit_stress != CG_STRESS_NONE && (m.nstress < 0 || fr.index < m.nstress))
                        ? (m.fit_stress == CG_STRESS_ALL ? 9 : 3) : 0;
    d.ener_rows = (m.fit_energy && (m.nener < 0 || fr.index < m.nener)) ? 3 : 0;
    // Image layers: a fit searches the NLAYERS ghost layers of its input, as
    // the reference does (it only warns when the cutoff exceeds half a
    // layered cell vector, BOX::IS_RCUT_SAFE, src/ClassDefs.C:2166-2190,2538);
    // a force evaluation searches as many layers as the cutoff needs, so that
    // no image within the cutoff is missed in a small cell.
    if (sticky) {
      d.nlayers = m.nlayers;
    } else {
      const double wmin = std::min(d.width[0], std::min(d.width[1], d.width[2]));
      d.nlayers = std::max(1, (int)std::ceil(list_rc / wmin));
      if (d.nlayers > kMaxLayers)
        throw Error(CG_ERR_INPUT, "configuration " + std::to_string(i) + ": cell width " + std::to_string(wmin) +
                                      " A is below cutoff / " + std::to_string(kMaxLayers) +
                                      " (more image layers than are supported)");
    }
    if (d.nlayers < 1 || d.nlayers > kMaxLayers) throw Error(CG_ERR_UNSUPPORTED, "NLAYERS must be 1 to 7");
    // NEIGHBORS small-cell mode is sticky across the frames of a fit; force
    // evaluations treat every configuration on its own
    const int small_here = (m.small_mode || d.min_extent <= mx) ? 1 : 0;
    if (sticky) {
      if (small_here) c.sticky_small = 1;
      d.small = c.sticky_small;
    } else {
      d.small = small_here;
    }
    // cell list when every axis holds at least three bins of thickness >= mx
    // and there are at least 125 bins (the 27-bin stencil then visits under a
    // quarter of the frame; below that the direct scan is as cheap), with the
    // bins per frame kept to O(atoms) for sparse cells
    d.cell = 0;
    d.nbin[0] = d.nbin[1] = d.nbin[2] = 1;
    d.bin_off = 0;
    if (!d.small && mx > 0.0) {
      int nb[3];
      bool ok = true;
      for (int r = 0; r < 3; ++r) {
        nb[r] = (int)std::min(256.0, std::floor(d.width[r] / list_rc));
        ok = ok && nb[r] >= 3;
      }
      if (ok && (long long)nb[0] * nb[1] * nb[2] < 125) ok = false;
      if (ok) {
        while ((long long)nb[0] * nb[1] * nb[2] > 4LL * fr.natoms + 125) {
          int r = nb[0] >= nb[1] && nb[0] >= nb[2] ? 0 : (nb[1] >= nb[2] ? 1 : 2);
          if (nb[r] <= 5) break;
          --nb[r];
        }
        d.cell = 1;
        for (int r = 0; r < 3; ++r) d.nbin[r] = nb[r];
        d.bin_off = nbins_total;
        nbins_total += nb[0] * nb[1] * nb[2];
      }
    }
    d.perm_scale[0] = d.perm_scale[1] = 1.0;
    for (int j = 2; j < 5; ++j) d.perm_scale[j] = d.small ? d.perm_scale[j - 1] / j : 1.0;
    rows += 3LL * fr.natoms + d.stress_rows + d.ener_rows;
    for (int a = 0; a < fr.natoms; ++a) {
      atom_frame.push_back(i);
      for (int k = 0; k < 3; ++k) xyz.push_back(fr.xyz[3 * a + k]);
      if (fr.type[a] < 0 || fr.type[a] >= m.ntypes) throw Error(CG_ERR_ARG, "bad atom type");
      type.push_back(fr.type[a]);
    }
    natoms += fr.natoms;
  }
  DevFrame *d_frames = c.b_frames.as<DevFrame>(nframes);
  int *d_af = c.b_type.as<int>(2 * (size_t)natoms);
  int *d_type = d_af + natoms;
  double *d_xyz = c.b_xyz.as<double>(3 * (size_t)natoms);
  double *d_wrap = c.b_wrap.as<double>(3 * (size_t)natoms);
  if (!d_frames || !d_af || !d_xyz || !d_wrap) throw Error(CG_ERR_MEMORY, "frame buffers");
  upload(c, d_frames, hf.data(), hf.size() * sizeof(DevFrame));
  upload(c, d_af, atom_frame.data(), natoms * sizeof(int));
  upload(c, d_type, type.data(), natoms * sizeof(int));
  upload(c, d_xyz, xyz.data(), xyz.size() * sizeof(double));
  B.nframes = nframes;
  B.natoms = natoms;
  B.rows = rows;
  B.nbins_total = nbins_total;
  B.mx = list_rc;
  B.d_frames = d_frames;
  B.d_af = d_af;
  B.d_type = d_type;
  B.d_xyz = d_xyz;
  B.d_wrap = d_wrap;
}

// Device part: wrap, cell lists, neighbour lists, sublists, cluster
// enumeration and bucketing.  With caps (capacity mode) it reads nothing back
// and can be captured into a graph; without, it reads the list sizes and the
// bucket bounds back to the host (exact mode).
static void prepare_device(Ctx &c, int md, Batch &B, const Caps *caps) {
  const cg_model &m = c.m;
  const int nframes = B.nframes, natoms = B.natoms, nbins_total = B.nbins_total;
  const double mx = B.mx;
  DevFrame *d_frames = B.d_frames;
  int *d_af = B.d_af, *d_type = B.d_type;
  double *d_xyz = B.d_xyz, *d_wrap = B.d_wrap;
  const int TB = 256;
  hipLaunchKernelGGL(k_wrap, dim3((natoms + TB - 1) / TB), dim3(TB), 0, c.stream, d_frames, d_af, d_xyz, d_wrap, natoms);
  // ---- cell lists: atoms sorted (stably) by bin, bin starts
  const int *d_bin_start = nullptr, *d_bin_atom = nullptr;
  if (nbins_total > 0) {
    char *bb = static_cast<char *>(c.b_bins.get((4 * (size_t)natoms + nbins_total + 1) * sizeof(int) + 256));
    if (!bb) throw Error(CG_ERR_MEMORY, "cell lists");
    int *key = reinterpret_cast<int *>(bb), *val = key + natoms, *key2 = val + natoms, *val2 = key2 + natoms;
    int *bstart = val2 + natoms;
    hipLaunchKernelGGL(k_bin_keys, dim3((natoms + TB - 1) / TB), dim3(TB), 0, c.stream, d_frames, d_af, d_wrap,
                       natoms, nbins_total, key, val);
    int bits = 1;
    while ((1LL << bits) <= (long long)nbins_total) ++bits;
    size_t tmp = 0;
    CG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, key, key2, val, val2, natoms, 0, bits, c.stream));
    void *t = c.b_scan.get(tmp);
    if (!t) throw Error(CG_ERR_MEMORY, "sort scratch");
    CG_HIP(hipcub::DeviceRadixSort::SortPairs(t, tmp, key, key2, val, val2, natoms, 0, bits, c.stream));
    hipLaunchKernelGGL(k_bucket_bounds, dim3((nbins_total + 1 + TB - 1) / TB), dim3(TB), 0, c.stream, key2, natoms,
                       nbins_total, bstart);
    CG_HIP(hipGetLastError());
    d_bin_start = bstart;
    d_bin_atom = val2;
  }
  // ---- neighbour lists (one list at the largest family cutoff)
  int *d_ncount = c.b_ncount.as<int>(natoms + 1);
  int *d_noff = c.b_noff.as<int>(natoms + 1);
  if (!d_ncount || !d_noff) throw Error(CG_ERR_MEMORY, "neighbour counts");
  zero_i(c, d_ncount, natoms + 1);
  hipLaunchKernelGGL(k_neigh<false>, dim3((natoms + 3) / 4), dim3(256), 0, c.stream, d_frames, d_af, d_wrap,
                     (const int *)d_type, d_bin_start, d_bin_atom, natoms, mx, d_ncount,
                     (const int *)nullptr, (uint32_t *)nullptr, (double4 *)nullptr, (int *)nullptr, (int *)nullptr, 0);
  exclusive_scan(c, d_ncount, d_noff, natoms);
  // exact mode reads the list size back (host sync); capacity mode sizes the
  // list from the caller's capacity
  const int nnb = caps ? caps->nnb : read_int(c, d_noff + natoms);
  const size_t nn1 = std::max(nnb, 1);
  B.tot = TotalPtrs{{d_noff + natoms, nullptr, nullptr, nullptr}};
  uint32_t *d_nlist = c.b_nlist.as<uint32_t>(nn1);
  double4 *d_nvec = c.b_nvec.as<double4>(nn1);
  int *d_nown = c.b_nown.as<int>(2 * nn1);
  int *d_ntyp = d_nown + nn1;
  if (!d_nlist || !d_nvec || !d_nown) throw Error(CG_ERR_MEMORY, "neighbour list");
  hipLaunchKernelGGL(k_neigh<true>, dim3((natoms + 3) / 4), dim3(256), 0, c.stream, d_frames, d_af, d_wrap,
                     (const int *)d_type, d_bin_start, d_bin_atom, natoms, mx, (int *)nullptr,
                     (const int *)d_noff, d_nlist, d_nvec, d_nown, d_ntyp, (int)nn1);
  // ---- fit statistics over the padded lists: pair minima, flat list sizes
  if (c.stats.on && !md && nnb > 0) {
    LsqStats &S = c.stats;
    hipLaunchKernelGGL(k_pair_min, dim3((nnb + TB - 1) / TB), dim3(TB), 0, c.stream, (const int *)d_type,
                       (const int *)d_nown, (const int *)d_ntyp, (const double4 *)d_nvec, (const int *)(d_noff + natoms),
                       nnb, (const int *)c.d_int_pair_map, m.ntypes, m.max_cutoff_2b + kListPadding, S.d_min);
    if (S.flat_cap < 2 * (size_t)nframes) {
      if (S.d_flat) (void)hipFree(S.d_flat);
      S.d_flat = nullptr;
      S.flat_cap = 0;
      CG_HIP(hipMalloc(&S.d_flat, 2 * (size_t)nframes * sizeof(long long)));
      S.flat_cap = 2 * (size_t)nframes;
    }
    CG_HIP(hipMemsetAsync(S.d_flat, 0, 2 * (size_t)nframes * sizeof(long long), c.stream));
    unsigned long long *flat = reinterpret_cast<unsigned long long *>(S.d_flat);
    if (c.fam[1].ntypes > 0 && c.fam[1].ncols > 0)
      hipLaunchKernelGGL(k_flat_count<3>, dim3((natoms + 3) / 4), dim3(256), 0, c.stream, (const DevFrame *)d_frames,
                         (const int *)d_af, (const int *)d_noff, (const uint32_t *)d_nlist, (const double4 *)d_nvec,
                         natoms, nnb, c.fam[1].cutoff + kListPadding, flat);
    if (c.fam[2].ntypes > 0 && c.fam[2].ncols > 0)
      hipLaunchKernelGGL(k_flat_count<4>, dim3((natoms + 3) / 4), dim3(256), 0, c.stream, (const DevFrame *)d_frames,
                         (const int *)d_af, (const int *)d_noff, (const uint32_t *)d_nlist, (const double4 *)d_nvec,
                         natoms, nnb, c.fam[2].cutoff + kListPadding, flat + nframes);
    CG_HIP(hipGetLastError());
  }
  // ---- per-family sublists (3B / 4B cutoffs below the list cutoff), sized
  // by the full list so that no host read is needed
  const uint32_t *fl_list[3] = {d_nlist, d_nlist, d_nlist};
  const double4 *fl_vec[3] = {d_nvec, d_nvec, d_nvec};
  const int *fl_own[3] = {d_nown, d_nown, d_nown};
  const int *fl_typ[3] = {d_ntyp, d_ntyp, d_ntyp};
  const int *fl_off[3] = {d_noff, d_noff, d_noff};
  const dim3 wg((unsigned)((natoms + 1 + 3) / 4)), wb(256);  // one wave per atom (+1 for the tail slot)
  for (int w = 1; w < 3; ++w) {
    const DevFamily &F = c.fam[w];
    if (F.ntypes == 0 || F.ncols == 0 || nnb == 0 || !(F.cutoff < mx)) continue;
    DBuf &sb = c.b_sub[w - 1];
    // layout: vec [nn1] (double4) | list [nn1] | own [nn1] | typ [nn1] | cnt [natoms+1] | off [natoms+1]
    char *base = static_cast<char *>(
        sb.get(nn1 * (sizeof(double4) + 3 * sizeof(int)) + 2 * ((size_t)natoms + 1) * sizeof(int) + 256));
    if (!base) throw Error(CG_ERR_MEMORY, "neighbour sublist");
    double4 *svec = reinterpret_cast<double4 *>(base);
    uint32_t *slist = reinterpret_cast<uint32_t *>(svec + nn1);
    int *sown = reinterpret_cast<int *>(slist + nn1);
    int *styp = sown + nn1;
    int *scnt = styp + nn1;
    int *soff = scnt + natoms + 1;
    hipLaunchKernelGGL(k_sub<true>, wg, wb, 0, c.stream, d_noff, d_nlist, d_nvec, d_nown, d_ntyp, natoms, F.cutoff,
                       scnt, (const int *)nullptr, (uint32_t *)nullptr, (double4 *)nullptr, (int *)nullptr,
                       (int *)nullptr, (int)nn1);
    exclusive_scan(c, scnt, soff, natoms);
    hipLaunchKernelGGL(k_sub<false>, wg, wb, 0, c.stream, d_noff, d_nlist, d_nvec, d_nown, d_ntyp, natoms, F.cutoff,
                       (int *)nullptr, (const int *)soff, slist, svec, sown, styp, (int)nn1);
    fl_list[w] = slist; fl_vec[w] = svec; fl_own[w] = sown; fl_typ[w] = styp; fl_off[w] = soff;
  }
  for (int w = 0; w < 3; ++w) {
    FamList &L = B.fl[w];
    L = FamList{};
    if (nnb == 0 || w == 0) continue;
    L.off = fl_off[w]; L.list = fl_list[w]; L.vec = fl_vec[w]; L.own = fl_own[w]; L.typ = fl_typ[w];
    L.tot = fl_off[w] + natoms;
    L.cap = nnb;
  }
  // ---- per-family cluster counts: 2-body one thread per neighbour entry,
  // 3/4-body one wave per atom over its LDS-staged list
  FamOut *fo = B.fo;
  const size_t cslot = std::max(nn1, 4 * (size_t)natoms) + 1;
  int *d_ccount = c.b_ccount.as<int>(3 * cslot);
  int *d_coff = c.b_coff.as<int>(3 * cslot);
  if (!d_ccount || !d_coff) throw Error(CG_ERR_MEMORY, "cluster counts");
  EnumArgs ea[3] = {};
  bool active[3] = {false, false, false};
  int nitems[3] = {0, 0, 0};  // count slots: entries (2B) or (atom, wave) pairs (3B/4B)
  const dim3 eg((unsigned)((nnb + TB - 1) / TB)), eb(TB);
  const dim3 ag((unsigned)natoms), ab(256);  // 3/4-body enumeration: one workgroup per atom
  for (int w = 0; w < 3; ++w) {
    const DevFamily &F = c.fam[w];
    if (F.ntypes == 0 || F.ncols == 0 || nnb == 0) continue;
    active[w] = true;
    EnumArgs &E = ea[w];
    E.frames = d_frames; E.atom_frame = d_af; E.type = d_type; E.noff = fl_off[w]; E.nlist = fl_list[w];
    E.nvec = fl_vec[w]; E.nown = fl_own[w]; E.ntyp = fl_typ[w]; E.nnb = nnb; E.ntot = d_noff + natoms;
    E.natoms = natoms; E.ntypes_atom = m.ntypes; E.nkeys_type = F.ntypes;
    E.cutoff = F.cutoff;
    E.count = d_ccount + w * cslot;
    E.coff = d_coff + w * cslot;
    E.cct = F.cct; E.csmax = F.csmax; E.cperm = F.cperm; E.cslot = F.cslot;
    E.slot_min = (c.stats.on && !md && w > 0) ? c.stats.d_min + (w == 1 ? c.stats.off3 : c.stats.off4) : nullptr;
    E.smin = F.smin; E.smax = F.smax; E.int_pair_map = c.d_int_pair_map;
    E.md = md;
    nitems[w] = w == 0 ? nnb : 4 * natoms;
    zero_i(c, E.count + nitems[w], 1);
    if (w == 0) hipLaunchKernelGGL((k_enum2<false>), eg, eb, 0, c.stream, E);
    else if (w == 1) hipLaunchKernelGGL((k_enum_wave<3, true>), ag, ab, 0, c.stream, E);
    else hipLaunchKernelGGL((k_enum_wave<4, true>), ag, ab, 0, c.stream, E);
    CG_HIP(hipGetLastError());
    exclusive_scan(c, E.count, const_cast<int *>(E.coff), nitems[w]);
  }
  int tot[3] = {0, 0, 0};
  for (int w = 0; w < 3; ++w)
    if (active[w]) B.tot.p[1 + w] = ea[w].coff + nitems[w];
  if (caps) {
    for (int w = 0; w < 3; ++w) tot[w] = caps->ncl[w];
  } else {
    for (int w = 0; w < 3; ++w)
      if (active[w])
        CG_HIP(hipMemcpyAsync(&tot[w], ea[w].coff + nitems[w], sizeof(int), hipMemcpyDeviceToHost, c.stream));
    CG_HIP(hipStreamSynchronize(c.stream));                                  // host sync 2 (exact mode)
  }
  // ---- fill, bucket by (frame, type), bucket bounds
  std::vector<int> hbounds[3];
  int *d_bounds[3] = {nullptr, nullptr, nullptr};
  DBuf *fbufs[3] = {&c.b_fam[0], &c.b_fam[1], &c.b_fam[2]};
  for (int w = 0; w < 3; ++w) {
    const DevFamily &F = c.fam[w];
    fo[w].ncl = 0;
    fo[w].refs = nullptr;
    fo[w].keys = nullptr;
    fo[w].boff.assign((size_t)nframes * std::max(F.ntypes, 1) + 1, 0);
    if (!active[w]) continue;
    const int ncl = tot[w];
    const int nb = F.nb;
    const int nkeys = nframes * F.ntypes;
    const bool sort = F.ntypes > 1 && ncl > 1;
    // capacity mode: slots past the device total carry the sentinel key
    const size_t n1 = std::max(ncl, 1);
    // refs [n1*nb] | key [n1] | bounds [nkeys+1] | (sort) key2 [n1] | idx [n1] | idx2 [n1] | refs2 [n1*nb]
    const size_t words = n1 * nb + n1 + (size_t)nkeys + 1 + (sort ? n1 * (3 + nb) : 0);
    uint32_t *blk = fbufs[w]->as<uint32_t>(words);
    if (!blk) throw Error(CG_ERR_MEMORY, "cluster list");
    EnumArgs &E = ea[w];
    E.refs = blk;
    E.key = reinterpret_cast<int *>(blk + n1 * nb);
    d_bounds[w] = E.key + n1;
    fo[w].ncl = ncl;
    fo[w].refs = E.refs;
    fo[w].nkeys = nkeys;
    fo[w].d_bounds = d_bounds[w];
    E.ccap = ncl;
    const int *keys = E.key;
    if (ncl > 0) {
      if (w == 0) hipLaunchKernelGGL((k_enum2<true>), eg, eb, 0, c.stream, E);
      else if (w == 1) hipLaunchKernelGGL((k_enum_wave<3, false>), ag, ab, 0, c.stream, E);
      else hipLaunchKernelGGL((k_enum_wave<4, false>), ag, ab, 0, c.stream, E);
      CG_HIP(hipGetLastError());
      if (caps)
        hipLaunchKernelGGL(k_pad_keys, dim3((ncl + TB - 1) / TB), dim3(TB), 0, c.stream, E.key,
                           (const int *)(ea[w].coff + nitems[w]), ncl, nkeys);
      if (sort) {
        // stable radix sort by (frame, type): clusters of a bucket keep their
        // enumeration order, so the bucketed list is deterministic
        int *key2 = d_bounds[w] + nkeys + 1;
        int *idx = key2 + n1, *idx2 = idx + n1;
        uint32_t *refs2 = reinterpret_cast<uint32_t *>(idx2 + n1);
        hipLaunchKernelGGL(k_iota, dim3((ncl + TB - 1) / TB), dim3(TB), 0, c.stream, idx, ncl);
        int bits = 1;
        while ((1LL << bits) <= (long long)nkeys) ++bits;  // keys 0..nkeys (nkeys: sentinel)
        size_t tmp = 0;
        CG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, E.key, key2, idx, idx2, ncl, 0, bits, c.stream));
        void *t = c.b_scan.get(tmp);
        if (!t) throw Error(CG_ERR_MEMORY, "sort scratch");
        CG_HIP(hipcub::DeviceRadixSort::SortPairs(t, tmp, E.key, key2, idx, idx2, ncl, 0, bits, c.stream));
        if (nb == 2) hipLaunchKernelGGL(k_gather_refs<2>, dim3((ncl + TB - 1) / TB), dim3(TB), 0, c.stream, E.refs, idx2, ncl, refs2);
        else if (nb == 3) hipLaunchKernelGGL(k_gather_refs<3>, dim3((ncl + TB - 1) / TB), dim3(TB), 0, c.stream, E.refs, idx2, ncl, refs2);
        else hipLaunchKernelGGL(k_gather_refs<4>, dim3((ncl + TB - 1) / TB), dim3(TB), 0, c.stream, E.refs, idx2, ncl, refs2);
        fo[w].refs = refs2;
        keys = key2;
      }
    }
    fo[w].keys = keys;
    hipLaunchKernelGGL(k_bucket_bounds, dim3((nkeys + 1 + TB - 1) / TB), dim3(TB), 0, c.stream, keys, ncl, nkeys,
                       d_bounds[w]);
    CG_HIP(hipGetLastError());
    if (!caps) {
      fo[w].boff.resize((size_t)nkeys + 1);
      CG_HIP(hipMemcpyAsync(fo[w].boff.data(), d_bounds[w], ((size_t)nkeys + 1) * sizeof(int),
                            hipMemcpyDeviceToHost, c.stream));
    }
  }
  if (!caps) CG_HIP(hipStreamSynchronize(c.stream));                         // host sync 3 (exact mode)
  if (!c.capturing) CG_HIP(hipEventRecord(c.ev[1], c.stream));
}

// 3/4-body pair tables once per neighbour entry (FP64 or FP32, as the
// assembly), and the entry of each pair of every bucketed cluster, for a
// family whose tables depend on the entry alone (DevFamily::pcache) and whose
// tile kernel takes the hand-scheduled set loops (4-body orders up to 5,
// 3-body up to 9).  Runs over the list and cluster capacities with device
// totals, so it needs no host read.
static void pair_cache(Ctx &c, const Batch &B, int w, bool fp32, const void *&ptab, const int *&pidx) {
  const DevFamily &F = c.fam[w];
  const FamOut &o = B.fo[w];
  const FamList &L = B.fl[w];
  ptab = nullptr;
  pidx = nullptr;
  if (w < 1 || !c.asm_pcache || !F.pcache || !F.ptp || !o.refs || !o.keys || o.ncl <= 0 || !L.off || L.cap <= 0)
    return;
  const int nb = F.nb, np = nb * (nb - 1) / 2, TB = 256;
  // table width of the kernel instance launch_family picks for the order
  const int VW = nb == 4 ? (F.max_pow <= 4 ? 4 : 5) : (F.max_pow <= 4 ? 4 : F.max_pow <= 8 ? 8 : 9);
  void *tab = c.b_ptab[w - 1].get((size_t)L.cap * 2 * VW * (fp32 ? sizeof(float) : sizeof(double)));
  int *idx = c.b_pidx[w - 1].as<int>((size_t)o.ncl * np);
  if (!tab || !idx) throw Error(CG_ERR_MEMORY, "pair tables");
  AsmArgs a{};
  a.transform = c.m.transform; a.fix_type = c.m.fix_type; a.fcut_type = c.m.fcut_type;
  a.deriv_const = c.m.deriv_const; a.smooth_dist = c.m.smooth_dist; a.fcut_offset = c.m.fcut_offset;
  a.qform = F.qform;
  const dim3 tg((unsigned)((L.cap + TB - 1) / TB)), tb(TB);
  const int *pm = c.d_int_pair_map;
  const double *ptp = F.ptp;
#define CG_PTAB(NBv, VWv)                                                                                   \
  do {                                                                                                      \
    if (fp32)                                                                                               \
      hipLaunchKernelGGL((k_pair_tab_f<NBv, VWv, 0>), tg, tb, 0, c.stream, a, L.vec, L.own, L.typ,          \
                         (const int *)B.d_type, pm, c.m.ntypes, ptp, L.tot, L.cap, tab);                    \
    else                                                                                                    \
      hipLaunchKernelGGL((k_pair_tab<NBv, VWv, 0>), tg, tb, 0, c.stream, a, L.vec, L.own, L.typ,            \
                         (const int *)B.d_type, pm, c.m.ntypes, ptp, L.tot, L.cap, tab);                    \
  } while (0)
  if (nb == 4) {
    if (VW == 4) CG_PTAB(4, 4);
    else CG_PTAB(4, 5);
  } else {
    if (VW == 4) CG_PTAB(3, 4);
    else if (VW == 8) CG_PTAB(3, 8);
    else CG_PTAB(3, 9);
  }
#undef CG_PTAB
  const dim3 ig((unsigned)((o.ncl + TB - 1) / TB)), ib(TB);
  if (nb == 4)
    hipLaunchKernelGGL(k_pair_idx<4>, ig, ib, 0, c.stream, o.refs, o.keys, o.ncl, F.ntypes, o.nkeys,
                       (const DevFrame *)B.d_frames, L.off, L.list, L.cap, idx);
  else
    hipLaunchKernelGGL(k_pair_idx<3>, ig, ib, 0, c.stream, o.refs, o.keys, o.ncl, F.ntypes, o.nkeys,
                       (const DevFrame *)B.d_frames, L.off, L.list, L.cap, idx);
  CG_HIP(hipGetLastError());
  ptab = tab;
  pidx = idx;
}

static void prepare(Ctx &c, const cg_frame *frames, int nframes, int md, Batch &B, bool sticky = true,
                    const Caps *caps = nullptr) {
  prepare_host(c, frames, nframes, B, sticky);
  prepare_device(c, md, B, caps);
}

constexpr size_t kMaxGraphs = 16;  // captured shapes kept per context

// List capacities of a capacity-mode call from the largest per-atom rates the
// context has seen (x1.25), rounded up to 8 steps per octave so that the
// capacities (part of a graph's shape) stay put while the counts drift from
// call to call; a call with as many atoms as the last one of its kind keeps
// the last capacities while they leave 5 % headroom over those rates.
static Caps plan_caps(Ctx &c, Ctx::CapPlan &plan, int natoms) {
  auto cap = [&](double rate) {
    const double v = std::min(1.9e9, natoms * rate * 1.25 + 1024.0);
    const double g = std::exp2(std::floor(std::log2(v)) - 3.0);
    return (int)(std::ceil(v / g) * g);
  };
  Caps caps;
  caps.nnb = cap(c.rate_nb);
  for (int w = 0; w < 3; ++w) caps.ncl[w] = cap(c.rate_cl[w]);
  if (natoms == plan.natoms) {
    bool keep = plan.last.nnb >= natoms * c.rate_nb * 1.05;
    for (int w = 0; w < 3; ++w) keep = keep && plan.last.ncl[w] >= natoms * c.rate_cl[w] * 1.05;
    if (keep) caps = plan.last;
  }
  plan.last = caps;
  plan.natoms = natoms;
  return caps;
}

// Every grow-only device buffer of a context.
template <class Fn> static void for_each_dbuf(Ctx &c, Fn fn) {
  for (DBuf *b : {&c.b_frames, &c.b_xyz, &c.b_type, &c.b_wrap, &c.b_ncount, &c.b_noff, &c.b_nlist, &c.b_nvec,
                  &c.b_nown, &c.b_sub[0], &c.b_sub[1], &c.b_bins, &c.b_ccount, &c.b_coff, &c.b_fam[0], &c.b_fam[1],
                  &c.b_fam[2], &c.b_tasks[0], &c.b_tasks[1], &c.b_tasks[2], &c.b_tcount[0], &c.b_tcount[1],
                  &c.b_tcount[2], &c.b_A, &c.b_scan, &c.b_x, &c.b_atvp, &c.b_ptab[0], &c.b_ptab[1], &c.b_pidx[0], &c.b_pidx[1], &c.b_fitkeep, &c.b_force, &c.b_coef[0], &c.b_coef[1],
                  &c.b_coef[2], &c.b_tmp, &c.b_ewald, &c.b_qr, &c.b_fitw, &c.b_lars, &c.b_atot, &c.b_atasks[0], &c.b_atasks[1], &c.b_atasks[2]})
    fn(*b);
}

// Addresses of the context's buffers (and pinned result buffer), hashed: a
// captured graph is replayed only while they are the ones it was captured
// with.
static uint64_t buffers_fingerprint(Ctx &c) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint64_t v) { h = (h ^ v) * 1099511628211ull; };
  for_each_dbuf(c, [&](DBuf &b) { mix((uint64_t)(uintptr_t)b.p); mix(b.cap); });
  mix((uint64_t)(uintptr_t)c.h_pin);
  mix(c.h_pin_cap);
  return h;
}

static void drop_graphs(Ctx &c) {
  for (auto &g : c.graphs)
    if (g.second.exec) (void)hipGraphExecDestroy(g.second.exec);
  c.graphs.clear();
}

// Capture what body enqueues on the context stream into a graph; false
// (nothing enqueued) when the capture fails, e.g. because a buffer would
// have to grow.
template <class Body> static bool capture_graph(Ctx &c, Body body, hipGraphExec_t *exec) {
  hipGraph_t graph = nullptr;
  c.capturing = true;
  t_no_grow = true;
  bool ok = hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
  if (ok) {
    try {
      body();
    } catch (const Error &) {
      ok = false;
    }
    if (hipStreamEndCapture(c.stream, &graph) != hipSuccess) ok = false;
  }
  c.capturing = false;
  t_no_grow = false;
  (void)hipGetLastError();
  if (ok && graph) ok = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0) == hipSuccess;
  else ok = false;
  if (graph) (void)hipGraphDestroy(graph);
  return ok;
}

// Keep a new graph under key (evicting the least recently used one when the
// cache is full).
static void keep_graph(Ctx &c, const std::string &key, hipGraphExec_t exec, uint64_t fp) {
  if (c.graphs.size() >= kMaxGraphs) {
    auto lru = c.graphs.begin();
    for (auto g = c.graphs.begin(); g != c.graphs.end(); ++g)
      if (g->second.used < lru->second.used) lru = g;
    (void)hipGraphExecDestroy(lru->second.exec);
    c.graphs.erase(lru);
  }
  c.graphs[key] = Ctx::Graph{exec, fp, ++c.graph_clock};
}

// The graph kept under key, if its buffers are still the context's (a
// buffer that moved since the capture makes every graph of the context
// stale: they are dropped).
static Ctx::Graph *find_graph(Ctx &c, const std::string &key, uint64_t fp) {
  auto it = c.graphs.find(key);
  if (it == c.graphs.end()) return nullptr;
  if (it->second.buffers != fp) {
    drop_graphs(c);
    return nullptr;
  }
  it->second.used = ++c.graph_clock;
  return &it->second;
}

// Assembly tasks of one family: every (frame, type) bucket in column tiles
// as wide as the frame's LDS tile allows under its size class's tile
// configuration, whatever the bucket's cluster count (the kernel reads the
// bucket's cluster range from the device bounds, and a bucket without
// clusters writes its tile of zeros).  The list depends only on the frame
// sizes and the tile choice, so it is built before any device work.  Tasks
// of adjacent classes with one configuration share a launch (each class is
// launched alone while the autotuner times them); within a launch larger
// frames come first (order: the frames by decreasing atom count).
static void build_tasks(const Ctx &c, int w, const std::vector<DevFrame> &hf, const std::vector<int> &by_size,
                        std::vector<Task> &tasks, std::vector<TaskGroup> &groups) {
  const DevFamily &F = c.fam[w];
  auto cfg_of = [&](int cls) { return w == 0 ? 0 : c.tile_cfg[w][cls]; };
  tasks.clear();
  groups.clear();
  // frames whose force rows do not fit the largest LDS tile (or above the
  // context's limit) accumulate in A itself: their tasks come last, in a
  // launch of their own, 64 columns wide
  auto global_acc = [&](int n) {
    return (c.gacc_above > 0 && n > c.gacc_above) || tile_bytes(n, 1) > kTileCfg[0].lds;
  };
  for (int cls = 0; cls < CG_TILE_CLASSES; ++cls) {
    const TileCfg &cfg = kTileCfg[cfg_of(cls)];
    const size_t first = tasks.size();
    size_t lds = 0;
    for (int fi : by_size) {
      const int n = hf[fi].natoms;
      if (tile_class(n) != cls || global_acc(n)) continue;
      int wmax = cfg.wcap;
      while (wmax > 0 && tile_bytes(n, wmax) > cfg.lds) --wmax;
      if (wmax < 1) wmax = 1;  // (configurations with half tiles: the frame fits one column of the full tile)
      for (int t = 0; t < F.ntypes; ++t) {
        const int nc = F.h_ncols[t];
        if (nc <= 0) continue;
        const int ntile = (nc + wmax - 1) / wmax;
        const int wt = (nc + ntile - 1) / ntile;
        lds = std::max(lds, tile_bytes(n, wt));
        for (int col = 0; col < nc; col += wt) tasks.push_back(Task{fi, t, col, std::min(wt, nc - col), 0, 0});
      }
    }
    if (tasks.size() == first) continue;
    if (!groups.empty() && !c.split_classes && groups.back().cls1 == cls - 1 &&
        cfg_of(groups.back().cls0) == cfg_of(cls)) {
      // merged launch: its tasks keep the order of their classes (classes
      // of larger frames come later, so the heavy tasks of the merged
      // launch are reordered first)
      std::vector<Task> merged(tasks.begin() + first, tasks.end());
      merged.insert(merged.end(), tasks.begin() + groups.back().begin, tasks.begin() + first);
      std::copy(merged.begin(), merged.end(), tasks.begin() + groups.back().begin);
      groups.back().end = tasks.size();
      groups.back().cls1 = cls;
      groups.back().lds = std::max(groups.back().lds, lds);
    } else {
      groups.push_back(TaskGroup{first, tasks.size(), cls, cls, lds});
    }
  }
  const size_t first = tasks.size();
  for (int fi : by_size) {
    if (!global_acc(hf[fi].natoms)) continue;
    for (int t = 0; t < F.ntypes; ++t) {
      const int nc = F.h_ncols[t];
      if (nc <= 0) continue;
      const int ntile = (nc + 63) / 64;
      const int wt = (nc + ntile - 1) / ntile;
      for (int col = 0; col < nc; col += wt) tasks.push_back(Task{fi, t, col, std::min(wt, nc - col), 0, 0});
    }
  }
  if (tasks.size() > first) {
    TaskGroup g{first, tasks.size(), CG_TILE_CLASSES - 1, CG_TILE_CLASSES - 1, 0};
    g.gacc = true;
    groups.push_back(g);
  }
}

// Shape of a capacity-mode assembly: everything its launches depend on
// besides the contents of the uploaded arrays.
static std::string asm_shape_key(const Ctx &c, const Batch &B, const Caps &caps, int prec) {
  std::string k = "A|";
  auto add = [&](long long v) { k += std::to_string(v); k += ','; };
  add(prec); add(c.asm_det); add(c.asm_pcache); add(B.nframes); add(caps.nnb); add(caps.ncl[0]); add(caps.ncl[1]); add(caps.ncl[2]);
  for (int w = 0; w < 3; ++w)
    for (int cls = 0; cls < CG_TILE_CLASSES; ++cls) add(c.tile_cfg[w][cls]);
  for (const DevFrame &d : B.hf) {
    add(d.natoms); add(d.small); add(d.cell); add(d.nbin[0]); add(d.nbin[1]); add(d.nbin[2]); add(d.nlayers);
    add(d.stress_rows); add(d.ener_rows);
  }
  return k;
}

// Design-matrix assembly of a batch of frames into the context's A.
//
// The first assembly on a context reads the neighbour and cluster totals back
// to size its lists exactly.  Later ones (without fit statistics) size them
// from the per-atom rates seen so far, as force calls do, and read nothing
// back before the end: the kernels write nothing past a capacity, and the
// device totals come back with the end of the call (a call that did not fit
// is repeated with exact sizes).  An assembly of the same shape as the
// previous call (frame sizes and cells, capacities, tile choice) is captured
// into a HIP graph and replayed by later calls of that shape, after the
// frame arrays and the task lists have been uploaded to the same buffers.
// Every entry of A is written by a kernel (tiles of zeros for empty
// buckets), so A is cleared first only when the model has FITCOUL charge
// columns.
static void assemble_impl(Ctx &c, const cg_frame *frames, int nframes, int prec, double *b_host, bool exact) {
  const cg_model &m = c.m;
  ++c.n_asm;
  c.asm_replayed = false;
  int natoms = 0;
  for (int i = 0; i < nframes; ++i) natoms += std::max(frames[i].natoms, 0);
  Caps caps;
  const bool capmode = !exact && c.rate_nb > 0.0 && !c.stats.on;
  if (capmode) caps = plan_caps(c, c.cap_asm, natoms);
  Batch B;
  CG_HIP(hipEventRecord(c.ev[0], c.stream));
  prepare_host(c, frames, nframes, B, true);
  std::vector<DevFrame> &hf = B.hf;
  const long long rows = B.rows;
  // ---- b (A_MAT::PRINT_FRAME b.txt order)
  c.h_b.assign(rows, 0.0);
  {
    long long r = 0;
    const int sel9[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};
    for (int i = 0; i < nframes; ++i) {
      const cg_frame &fr = frames[i];
      for (int k = 0; k < 3 * fr.natoms; ++k) c.h_b[r++] = fr.forces ? fr.forces[k] : 0.0;
      for (int k = 0; k < hf[i].stress_rows; ++k)
        c.h_b[r++] = fr.stress[hf[i].stress_rows == 9 ? sel9[k] : k] / 6.9476955;
      for (int k = 0; k < hf[i].ener_rows; ++k) c.h_b[r++] = fr.energy;
    }
    if (b_host) std::memcpy(b_host, c.h_b.data(), rows * sizeof(double));
  }
  // ---- A and the task lists (uploaded before any capture)
  const size_t esz = prec == CG_FP32 ? sizeof(float) : sizeof(double);
  const long long ld = m.ncols;
  void *A = c.b_A.get((size_t)rows * ld * esz);
  if (!A) throw Error(CG_ERR_MEMORY, "A matrix does not fit in device memory");
  // columns written by the assembly kernels: the families' (every bucket)
  // and the energy offsets (k_offsets, every row)
  int covered = m.n_offset_cols;
  for (int w = 0; w < 3; ++w)
    if (c.fam[w].ntypes > 0) covered += c.fam[w].ncols;
  bool clear_A = covered != m.ncols || m.n_charge_cols > 0;
  // task lists: rebuilt and uploaded when the frame sizes or the tile choice
  // differ from the last assembly's
  {
    std::string tkey;
    for (int w = 0; w < 3; ++w)
      for (int cls = 0; cls < CG_TILE_CLASSES; ++cls) tkey += std::to_string(c.tile_cfg[w][cls]) + ',';
    tkey += c.split_classes ? "s|" : "m|";
    tkey += std::to_string(c.gacc_above) + '|';
    for (const DevFrame &d : hf) tkey += std::to_string(d.natoms) + ',';
    bool same = tkey == c.asm_task_key;
    for (int w = 0; w < 3 && same; ++w)
      same = c.asm_tasks[w].empty() || c.b_atasks[w].cap >= c.asm_tasks[w].size() * sizeof(Task);
    if (!same) {
      c.asm_task_key.clear();
      std::vector<int> by_size(nframes);
      for (int i = 0; i < nframes; ++i) by_size[i] = i;
      std::stable_sort(by_size.begin(), by_size.end(), [&](int x, int y) { return hf[x].natoms > hf[y].natoms; });
      for (int w = 0; w < 3; ++w) {
        c.asm_tasks[w].clear();
        c.asm_groups[w].clear();
        const DevFamily &F = c.fam[w];
        if (F.ntypes == 0 || F.ncols == 0) continue;
        build_tasks(c, w, hf, by_size, c.asm_tasks[w], c.asm_groups[w]);
        if (c.asm_tasks[w].empty()) continue;
        Task *d = c.b_atasks[w].as<Task>(c.asm_tasks[w].size());
        if (!d) throw Error(CG_ERR_MEMORY, "task list");
        upload(c, d, c.asm_tasks[w].data(), c.asm_tasks[w].size() * sizeof(Task));
      }
      c.asm_task_key = tkey;
    }
  }
  const std::vector<Task> *atasks = c.asm_tasks;
  const std::vector<TaskGroup> *fam_groups = c.asm_groups;
  // frames accumulated in A itself need it zeroed first
  for (int w = 0; w < 3; ++w)
    for (const TaskGroup &g : fam_groups[w]) clear_A = clear_A || g.gacc;
  Task *d_tasks[3] = {nullptr, nullptr, nullptr};
  for (int w = 0; w < 3; ++w)
    if (!atasks[w].empty()) d_tasks[w] = c.b_atasks[w].as<Task>(atasks[w].size());
  double *d_tot = c.b_atot.as<double>(8);
  if (!d_tot) throw Error(CG_ERR_MEMORY, "assembly totals");
  // ---- device work: lists, clusters, families, offsets, totals
  auto body = [&]() {
    if (clear_A) CG_HIP(hipMemsetAsync(A, 0, (size_t)rows * ld * esz, c.stream));
    prepare_device(c, 0, B, capmode ? &caps : nullptr);
    const void *ptab[3] = {nullptr, nullptr, nullptr};
    const int *pidx[3] = {nullptr, nullptr, nullptr};
    for (int w = 1; w < 3; ++w)
      if (d_tasks[w]) pair_cache(c, B, w, prec == CG_FP32, ptab[w], pidx[w]);
    for (int w = 0; w < 3; ++w) {
      if (!d_tasks[w]) continue;
      const std::vector<TaskGroup> &groups = fam_groups[w];
      if (!c.capturing) CG_HIP(hipEventRecord(c.fev[w][0], c.stream));
      // launches of different configurations run side by side (their tails
      // overlap); the autotuner's timing runs keep them in sequence
      if (groups.size() > CG_TILE_CLASSES + 1) throw Error(CG_ERR_DEVICE, "assembly launch groups");
      const bool fork = groups.size() > 1 && !c.split_classes;
      if (fork) CG_HIP(hipEventRecord(c.fork, c.stream));
      for (size_t gi = 0; gi < groups.size(); ++gi) {
        const TaskGroup &g = groups[gi];
        hipStream_t st = fork && gi > 0 ? c.aux[gi - 1] : c.stream;
        if (st != c.stream) CG_HIP(hipStreamWaitEvent(st, c.fork, 0));
        const size_t lds = g.lds;
        const int bs = kTileCfg[w == 0 ? 0 : c.tile_cfg[w][g.cls0]].bs;
        const bool timed = !c.capturing && !g.gacc;
        if (timed) CG_HIP(hipEventRecord(c.cev[w][g.cls0][0], st));
        if (prec == CG_FP32)
          launch_family<float>(c, w, g.end - g.begin, d_tasks[w] + g.begin, B.d_frames, B.fo[w].refs,
                               B.fo[w].d_bounds, lds, A, ld, bs, st, g.gacc, ptab[w], pidx[w]);
        else
          launch_family<double>(c, w, g.end - g.begin, d_tasks[w] + g.begin, B.d_frames, B.fo[w].refs,
                                B.fo[w].d_bounds, lds, A, ld, bs, st, g.gacc, ptab[w], pidx[w]);
        if (timed) CG_HIP(hipEventRecord(c.cev[w][g.cls0][1], st));
        if (st != c.stream) {
          CG_HIP(hipEventRecord(c.join[gi - 1], st));
          CG_HIP(hipStreamWaitEvent(c.stream, c.join[gi - 1], 0));
        }
      }
      if (!c.capturing) CG_HIP(hipEventRecord(c.fev[w][1], c.stream));
    }
    if (m.n_offset_cols > 0) {
      if (prec == CG_FP32)
        hipLaunchKernelGGL(k_offsets<float>, dim3(nframes), dim3(256), 0, c.stream, B.d_frames, B.d_type, nframes,
                           m.ntypes, m.ncols_short, (float *)A, ld);
      else
        hipLaunchKernelGGL(k_offsets<double>, dim3(nframes), dim3(256), 0, c.stream, B.d_frames, B.d_type, nframes,
                           m.ntypes, m.ncols_short, (double *)A, ld);
    }
    hipLaunchKernelGGL(k_totals, dim3(1), dim3(64), 0, c.stream, B.tot, d_tot);
    CG_HIP(hipMemcpyAsync(c.h_atot, d_tot, 4 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  };
  bool replayed = false;
  // calls with Ewald columns (per-call k tables), a cleared A or statistics
  // stay out of graphs
  if (capmode && c.asm_graphs && !clear_A && !c.split_classes) {
    const std::string key = asm_shape_key(c, B, caps, prec);
    const uint64_t fp = buffers_fingerprint(c);
    if (Ctx::Graph *g = find_graph(c, key, fp)) {
      CG_HIP(hipGraphLaunch(g->exec, c.stream));
      ++c.n_asm_graph;
      replayed = true;
    } else if (key == c.last_asm_key) {
      hipGraphExec_t exec = nullptr;
      if (capture_graph(c, body, &exec)) {
        ++c.n_asm_capture;
        keep_graph(c, key, exec, fp);
        CG_HIP(hipGraphLaunch(exec, c.stream));
        replayed = true;
      }
    }
    c.last_asm_key = key;
  }
  if (!replayed) body();
  if (m.n_charge_cols > 0) ewald_columns(c, frames, hf, B.d_af, B.d_type, B.d_wrap, B.natoms, A, ld, prec);
  CG_HIP(hipEventRecord(c.ev[2], c.stream));
  CG_HIP(hipEventSynchronize(c.ev[2]));
  CG_HIP(hipGetLastError());
  // ---- totals: list sizes against the capacities, rates for later calls
  const double *tot = c.h_atot;
  const double na = std::max(natoms, 1);
  if (capmode) {
    bool fits = tot[0] <= caps.nnb;
    for (int w = 0; w < 3; ++w) fits = fits && tot[1 + w] <= caps.ncl[w];
    if (!fits) {
      c.rate_nb = std::max(c.rate_nb, tot[0] / na);
      for (int w = 0; w < 3; ++w) c.rate_cl[w] = std::max(c.rate_cl[w], tot[1 + w] / na);
      ++c.n_asm_repeat;
      --c.n_asm;
      assemble_impl(c, frames, nframes, prec, b_host, true);
      return;
    }
  }
  c.rate_nb = std::max(c.rate_nb, tot[0] / na);
  for (int w = 0; w < 3; ++w) {
    c.rate_cl[w] = std::max(c.rate_cl[w], tot[1 + w] / na);
    c.ncl[w] = tot[1 + w] > 0 ? (long long)tot[1 + w] : 0;
  }
  if (c.stats.on) {
    LsqStats &S = c.stats;
    FamOut *fo = B.fo;
    for (int w = 0; w < 3; ++w) {
      const DevFamily &F = c.fam[w];
      if (S.count[w].size() != (size_t)F.ntypes) S.count[w].assign(F.ntypes, 0);
      if (fo[w].boff.size() < (size_t)nframes * F.ntypes + 1 || fo[w].ncl == 0) continue;
      for (int fi = 0; fi < nframes; ++fi)
        for (int t = 0; t < F.ntypes; ++t)
          S.count[w][t] += fo[w].boff[(size_t)fi * F.ntypes + t + 1] - fo[w].boff[(size_t)fi * F.ntypes + t];
    }
    S.flat3.assign(nframes, 0);
    S.flat4.assign(nframes, 0);
    if (S.d_flat && B.natoms > 0) {
      std::vector<long long> h(2 * (size_t)nframes);
      CG_HIP(hipMemcpyAsync(h.data(), S.d_flat, h.size() * sizeof(long long), hipMemcpyDeviceToHost, c.stream));
      CG_HIP(hipStreamSynchronize(c.stream));
      std::copy(h.begin(), h.begin() + nframes, S.flat3.begin());
      std::copy(h.begin() + nframes, h.end(), S.flat4.begin());
    }
  }
  // ---- timing: per family and class on eager calls; a replayed call has
  // only its total (the graph's kernels carry no timing events)
  c.asm_replayed = replayed;
  for (int w = 0; w < 3; ++w) {
    c.ms_fam[w] = 0.f;
    for (int cls = 0; cls < CG_TILE_CLASSES; ++cls) c.ms_cls[w][cls] = 0.f;
  }
  if (replayed) {
    c.ms_enum = 0.f;
    CG_HIP(hipEventElapsedTime(&c.ms_asm, c.ev[0], c.ev[2]));
    for (int w = 0; w < 3; ++w) c.ms_fam[w] = -1.f;
  } else {
    CG_HIP(hipEventElapsedTime(&c.ms_enum, c.ev[0], c.ev[1]));
    CG_HIP(hipEventElapsedTime(&c.ms_asm, c.ev[1], c.ev[2]));
    for (int w = 0; w < 3; ++w) {
      if (!d_tasks[w]) continue;
      CG_HIP(hipEventElapsedTime(&c.ms_fam[w], c.fev[w][0], c.fev[w][1]));
      for (const TaskGroup &g : fam_groups[w]) {
        if (g.gacc) continue;
        float ms = 0.f;
        CG_HIP(hipEventElapsedTime(&ms, c.cev[w][g.cls0][0], c.cev[w][g.cls0][1]));
        for (int cls = g.cls0; cls <= g.cls1; ++cls) c.ms_cls[w][cls] = ms;  // shared by a merged launch
      }
    }
  }
  c.ms_total = c.ms_enum + c.ms_asm;
  c.A = A;
  c.A_prec = prec;
  c.A_rows = rows;
  c.A_ld = (int)ld;
}

void assemble(Ctx &c, const cg_frame *frames, int nframes, int prec, double *b_host) {
  assemble_impl(c, frames, nframes, prec, b_host, false);
}


// Contraction order of a family: tensor dimension per pair (2-body: the
// pair's SNUM columns; 3B/4B: largest power + 1).
static int force_order(const Ctx &c, int w) { return w == 0 ? c.m.order2b : c.fam[w].max_pow; }

// Coefficient tensors of every family for the coefficient vector x (host
// build, uploaded when x differs from the last one seen by this context).
static void set_coef(Ctx &c, const double *x) {
  const cg_model &m = c.m;
  if (c.coef_valid && std::equal(x, x + m.ncols, c.h_x_last.begin())) return;
  for (int w = 0; w < 3; ++w) {
    const DevFamily &F = c.fam[w];
    c.coef_size[w] = 0;
    if (F.ntypes == 0 || F.ncols == 0) continue;
    const int N = force_order(c, w);
    size_t ts = 1;
    for (int k = 0; k < F.np; ++k) ts *= (size_t)N;
    std::vector<double> T((size_t)F.ntypes * ts, 0.0);
    if (w == 0) {
      for (int t = 0; t < F.ntypes; ++t)
        for (int q = 0; q < N; ++q) T[(size_t)t * ts + q] = x[F.h_col_offset[t] + q];
    } else {
      const cg_cluster_set &cs = w == 1 ? m.c3 : m.c4;
      for (int t = 0; t < cs.ntypes; ++t) {
        if (cs.excluded[t]) continue;
        for (int sidx = cs.perm_offset[t]; sidx < cs.perm_offset[t + 1]; ++sidx) {
          size_t idx = 0;
          for (int k = 0; k < cs.npairs; ++k) idx = idx * N + (size_t)cs.perm_powers[(size_t)sidx * cs.npairs + k];
          T[(size_t)t * ts + idx] = x[cs.col_offset[t] + cs.perm_col[sidx]];
        }
      }
    }
    double *d = c.b_coef[w].as<double>(T.size());
    if (!d) throw Error(CG_ERR_MEMORY, "coefficient tensors");
    CG_HIP(hipMemcpy(d, T.data(), T.size() * sizeof(double), hipMemcpyHostToDevice));
    c.coef_size[w] = (int)ts;
  }
  if (m.n_charge_cols > 0) ewald_type_charges(m, x, c.q_type);
  c.h_x_last.assign(x, x + m.ncols);
  c.coef_valid = true;
}

constexpr int FORCE_BS = 256;  // threads (= clusters) per force workgroup

template <int NB, int N, int BASE>
static void launch_force_n(Ctx &c, dim3 grid, const ForceArgs &a) {
  hipLaunchKernelGGL((k_force<NB, N, BASE, FORCE_BS>), grid, dim3(FORCE_BS), 0, c.stream, a);
}

template <int NB, int BASE, int... Ns>
static bool dispatch_force(Ctx &c, int N, dim3 grid, const ForceArgs &a, std::integer_sequence<int, Ns...>) {
  bool done = false;
  ((N == Ns + 1 ? (launch_force_n<NB, Ns + 1, BASE>(c, grid, a), done = true) : false), ...);
  return done;
}

// Device part of a force call after prepare_device: per-family task lists
// and force kernels, the call's totals, and the copy of results and totals
// into the pinned host buffer.
// Host destination of a force call's results: F (3 * atoms doubles) and the
// rest (energy / virial per frame and the call totals); F == nullptr means
// the context's pinned buffer receives everything.
struct HostOut {
  double *F = nullptr;
  double *rest = nullptr;
};

static void force_device(Ctx &c, const Batch &B, double *d_F, size_t nout, const HostOut &ho = HostOut{}) {
  const cg_model &m = c.m;
  const int TB = 256;
  double *d_EV = d_F + 3 * (size_t)B.natoms;
  zero_d(c, d_F, nout);
  for (int w = 0; w < 3; ++w) {
    const DevFamily &Fm = c.fam[w];
    const FamOut &fo = B.fo[w];
    if (Fm.ntypes == 0 || Fm.ncols == 0 || fo.ncl == 0 || !fo.d_bounds) continue;
    // tasks: chunks of at most FORCE_BS clusters of one (frame, type) bucket
    const int nkeys = fo.nkeys;
    const int max_tasks = nkeys + (fo.ncl + FORCE_BS - 1) / FORCE_BS;
    int *d_cnt = c.b_tcount[w].as<int>(2 * ((size_t)nkeys + 1));
    Task *d_tasks = c.b_tasks[w].as<Task>(max_tasks);
    if (!d_cnt || !d_tasks) throw Error(CG_ERR_MEMORY, "task list");
    int *d_toff = d_cnt + nkeys + 1;
    hipLaunchKernelGGL(k_task_counts, dim3((nkeys + 1 + TB - 1) / TB), dim3(TB), 0, c.stream, fo.d_bounds, nkeys,
                       FORCE_BS, d_cnt);
    exclusive_scan(c, d_cnt, d_toff, nkeys);
    hipLaunchKernelGGL(k_task_fill, dim3((max_tasks + TB - 1) / TB), dim3(TB), 0, c.stream, fo.d_bounds,
                       (const int *)d_toff, nkeys, Fm.ntypes, FORCE_BS, max_tasks, d_tasks);
    ForceArgs a{};
    a.tasks = d_tasks; a.ntasks = d_toff + nkeys;
    a.frames = B.d_frames; a.wrap = B.d_wrap; a.refs = fo.refs;
    a.smin = Fm.smin; a.smax = Fm.smax; a.xavg = Fm.xavg; a.xdiff = Fm.xdiff; a.lambda = Fm.lambda;
    a.recip = Fm.recip;
    a.coef = c.b_coef[w].as<double>(0); a.tsize = c.coef_size[w];
    a.F = d_F; a.EV = d_EV;
    a.transform = m.transform; a.fix_type = m.fix_type; a.fcut_type = m.fcut_type;
    a.deriv_const = m.deriv_const; a.smooth_dist = m.smooth_dist; a.fcut_offset = m.fcut_offset;
    a.penalty_scale = w == 0 ? c.penalty_scale : 0.0;
    a.penalty_dist = c.penalty_dist;
    a.nb_scale = Fm.nb;
    const dim3 grid((unsigned)max_tasks);
    const int N = force_order(c, w);
    bool ok;
    if (w == 0) ok = dispatch_force<2, 1>(c, N, grid, a, std::make_integer_sequence<int, 16>{});
    else if (w == 1) ok = dispatch_force<3, 0>(c, N, grid, a, std::make_integer_sequence<int, 16>{});
    else ok = dispatch_force<4, 0>(c, N, grid, a, std::make_integer_sequence<int, 8>{});
    if (!ok) throw Error(CG_ERR_UNSUPPORTED, "force evaluation supports 2/3-body orders up to 16 and 4-body up to 8");
    CG_HIP(hipGetLastError());
  }
  // Ewald terms of a model fitted with charges (chimes_md's ZCalc_Ewald)
  if (m.n_charge_cols > 0) ewald_md(c, B.frames, B.hf, B.d_af, B.d_type, B.d_wrap, B.natoms, d_F, d_EV);
  hipLaunchKernelGGL(k_totals, dim3(1), dim3(64), 0, c.stream, B.tot, d_F + nout);
  if (ho.F) {
    const size_t nf = 3 * (size_t)B.natoms;
    CG_HIP(hipMemcpyAsync(ho.F, d_F, nf * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    CG_HIP(hipMemcpyAsync(ho.rest, d_F + nf, (nout + 4 - nf) * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  } else {
    CG_HIP(hipMemcpyAsync(c.h_pin, d_F, (nout + 4) * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  }
}

// Shape of a capacity-mode call: everything its launches depend on besides
// the contents of the uploaded arrays.
static std::string shape_key(const Ctx &c, const Batch &B, const Caps &caps) {
  std::string k;
  auto add = [&](long long v) { k += std::to_string(v); k += ','; };
  // the penalty settings are kernel arguments of the captured launches
  add((long long)__builtin_bit_cast(uint64_t, c.penalty_dist));
  add((long long)__builtin_bit_cast(uint64_t, c.penalty_scale));
  add(B.nframes); add(caps.nnb); add(caps.ncl[0]); add(caps.ncl[1]); add(caps.ncl[2]);
  for (const DevFrame &d : B.hf) {
    add(d.natoms); add(d.small); add(d.cell); add(d.nbin[0]); add(d.nbin[1]); add(d.nbin[2]); add(d.nlayers);
  }
  return k;
}

// Force / energy / virial of a batch of configurations, each standing alone
// (no small-cell stickiness across them).  F: [3 * sum natoms], E: [nframes],
// vir: [9 * nframes].
//
// Once the context has seen a call, the neighbour and cluster lists are
// sized from per-atom rates (the largest seen, plus a margin) instead of
// being read back, the force tasks are built on the device, and the only
// host synchronisation is the copy of the results.  The device totals come
// back with them; a call that did not fit is repeated with exact sizes.
// A call of the same shape as the previous one is captured into a HIP graph
// (kept per shape and replayed for later calls of that shape, after the
// frame arrays have been uploaded), while the buffers it uses stay in place.
static void force_batch_impl(Ctx &c, const double *x_host, const cg_frame *frames, int nframes, double *F,
                             double *E, double *vir, bool exact, double *rest_host = nullptr) {
  const cg_model &m = c.m;
  // rest_host (allocation-free calls): results go straight to F and to this
  // host array of 10 * nframes + 4 doubles, with exact list sizes (no pinned
  // buffer, no capacity estimates, no graphs)
  const bool noalloc = rest_host != nullptr;
  set_coef(c, x_host);
  int natoms = 0;
  for (int i = 0; i < nframes; ++i) natoms += frames[i].natoms;
  Caps caps;
  const bool capmode = !noalloc && !exact && c.rate_nb > 0.0;
  if (capmode) caps = plan_caps(c, c.cap_force, natoms);
  Batch B;
  prepare_host(c, frames, nframes, B, false);
  const size_t nout = 3 * (size_t)natoms + 10 * (size_t)nframes;
  double *d_F = c.b_force.as<double>(nout + 16);
  if (!d_F) throw Error(CG_ERR_MEMORY, "force buffers");
  if (!noalloc && c.h_pin_cap < nout + 4) {
    const size_t old_cap = c.h_pin_cap;
    // the old block is freed with the context (hipHostFree would wait for
    // the whole device, stalling other contexts' work)
    if (c.h_pin) c.h_pin_retired.push_back(c.h_pin);
    c.h_pin = nullptr;
    c.h_pin_cap = 0;
    ++t_grow_events;
    const size_t n = std::max(nout + nout / 4 + 64, std::min(2 * old_cap, 2 * nout + 64));
    CG_HIP(hipHostMalloc((void **)&c.h_pin, n * sizeof(double), hipHostMallocDefault));
    c.h_pin_cap = n;
  }
  bool done = false;
  // calls with Ewald terms upload per-call k tables and stay out of graphs
  if (capmode && m.n_charge_cols == 0) {
    const std::string key = shape_key(c, B, caps);
    const uint64_t fp = buffers_fingerprint(c);
    if (Ctx::Graph *g = find_graph(c, key, fp)) {
      CG_HIP(hipGraphLaunch(g->exec, c.stream));
      ++c.n_graph;
      done = true;
    } else if (key == c.last_key) {
      hipGraphExec_t exec = nullptr;
      auto body = [&]() {
        prepare_device(c, 1, B, &caps);
        force_device(c, B, d_F, nout);
      };
      if (capture_graph(c, body, &exec)) {
        ++c.n_capture;
        keep_graph(c, key, exec, fp);
        CG_HIP(hipGraphLaunch(exec, c.stream));
        done = true;
      }
    }
    c.last_key = key;
  }
  HostOut ho;
  if (noalloc) {
    ho.F = F;
    ho.rest = rest_host;
  }
  if (!done) {
    prepare_device(c, 1, B, capmode ? &caps : nullptr);
    force_device(c, B, d_F, nout, ho);
  }
  CG_HIP(hipStreamSynchronize(c.stream));
  // results: F then energy / virial per frame then the call totals
  const double *res_rest = noalloc ? rest_host : c.h_pin + 3 * (size_t)natoms;
  const double *tot = res_rest + 10 * (size_t)nframes;
  const double na = std::max(natoms, 1);
  bool fits = true;
  if (capmode) {
    fits = tot[0] <= caps.nnb;
    for (int w = 0; w < 3; ++w) fits = fits && tot[1 + w] <= caps.ncl[w];
  }
  c.rate_nb = std::max(c.rate_nb, tot[0] / na);
  for (int w = 0; w < 3; ++w) c.rate_cl[w] = std::max(c.rate_cl[w], tot[1 + w] / na);
  if (!fits) {
    ++c.n_repeat;
    force_batch_impl(c, x_host, frames, nframes, F, E, vir, true);
    return;
  }
  if (F && !noalloc) std::memcpy(F, c.h_pin, 3 * (size_t)natoms * sizeof(double));
  const double *ev = res_rest;
  for (int fi = 0; fi < nframes; ++fi) {
    double eo = 0.0;  // per-type energy offsets: atom count x offset coefficient
    if (m.n_offset_cols > 0) {
      int cnt[CG_MAX_ATOM_TYPES] = {};
      for (int a = 0; a < frames[fi].natoms; ++a) ++cnt[frames[fi].type[a]];
      for (int t = 0; t < m.n_offset_cols && t < m.ntypes; ++t) eo += cnt[t] * x_host[m.ncols_short + t];
    }
    if (E) E[fi] = ev[10 * fi] + eo;
    if (vir) for (int q = 0; q < 9; ++q) vir[9 * fi + q] = ev[10 * fi + 1 + q];
  }
}

void force_batch(Ctx &c, const double *x_host, const cg_frame *frames, int nframes, double *F, double *E,
                 double *vir) {
  ++c.n_force;
  force_batch_impl(c, x_host, frames, nframes, F, E, vir, false);
}

// Scratch of the allocation-free force entry point: one fixed array per
// GPU (it lives as long as the process).  In host code the array's name is
// not a device address; its address on each device comes from
// hipGetSymbolAddress, taken once per device when the first context is
// created there, and only that address is handed to kernels and copies.
constexpr size_t kNoAllocScratchBytes = (size_t)CG_NOALLOC_SCRATCH_MIB << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxDevices = 64;
static void *g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu;
thread_local Arena *t_arena = nullptr;

// Device address of this device's scratch array (the device must be current).
static void scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) throw Error(CG_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void *p = nullptr;
  CG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) throw Error(CG_ERR_DEVICE, "scratch array has no device address");
  g_scratch_ptr[device] = p;
}

static void *scratch_ptr(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return (device >= 0 && device < kMaxDevices) ? g_scratch_ptr[device] : nullptr;
}

// cg_force with every device buffer of the call carved from the scratch
// array (no hipMalloc / hipHostMalloc / hipFree): coefficient tensors,
// frame arrays, neighbour and cluster lists, sort scratch, task lists and
// results.  Results are copied to F and to a stack array.
void force_noalloc(Ctx &c, const double *x, const cg_frame *f, double *F, double *E, double *vir) {
  void *base = scratch_ptr(c.device);
  if (!base) throw Error(CG_ERR_DEVICE, "scratch array not set up on this device");
  // work of an earlier call on this stream (one that failed half way) may
  // still use the scratch; the coefficient upload below is not stream-ordered
  CG_HIP(hipStreamSynchronize(c.stream));
  Arena arena;
  arena.base = static_cast<char *>(base);
  arena.cap = kNoAllocScratchBytes;
  arena.reset();
  double rest[10 + 4];
  // installed for this call only, whichever way it ends; the context's own
  // coefficient tensors are not the ones in the scratch
  struct Install {
    Ctx &c;
    Install(Ctx &cc, Arena *a) : c(cc) { t_arena = a; c.coef_valid = false; }
    ~Install() { t_arena = nullptr; c.coef_valid = false; }
  };
  try {
    Install guard(c, &arena);
    ++c.n_force;
    force_batch_impl(c, x, f, 1, F, E, vir, true, rest);
  } catch (const Error &e) {
    if (arena.overflow)
      throw Error(CG_ERR_MEMORY, "cg_force_noalloc: a configuration of " + std::to_string(f->natoms) +
                                     " atoms needs more than the " + std::to_string(CG_NOALLOC_SCRATCH_MIB) +
                                     " MiB scratch (at least " + std::to_string((arena.wanted + (1 << 20) - 1) >> 20) +
                                     " MiB); use cg_force");
    throw;
  }
}

void ctx_init(Ctx &c, const cg_model *m, int device) {
  c.model = m;
  c.m = *m;
  c.device = device;
  CG_HIP(hipSetDevice(device));
  scratch_setup(device);
  CG_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  for (auto &st : c.aux) CG_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  // buffers grow stream-ordered on the context stream; the device's pool
  // keeps freed blocks for reuse instead of returning them at every sync
  for_each_dbuf(c, [&](DBuf &b) { b.s = c.stream; });
  {
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool) {
      uint64_t keep = UINT64_MAX;
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    (void)hipGetLastError();
  }
  CG_HIP(hipEventCreateWithFlags(&c.fork, hipEventDisableTiming));
  for (auto &e : c.join) CG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  for (auto &e : c.ev) CG_HIP(hipEventCreate(&e));
  for (auto &fe : c.fev) for (auto &e : fe) CG_HIP(hipEventCreate(&e));
  for (auto &fc : c.cev) for (auto &fe : fc) for (auto &e : fe) CG_HIP(hipEventCreate(&e));
  CG_HIP(hipHostMalloc((void **)&c.h_atot, 8 * sizeof(double), hipHostMallocDefault));
  {
    // tile choices made by cg_autotune for this model earlier in the process
    std::lock_guard<std::mutex> lk(g_tile_mu);
    for (int w = 1; w < 3; ++w)
      for (int cls = 0; cls < CG_TILE_CLASSES; ++cls) {
        auto it = g_tile_cache.find(tile_key(*m, w, cls));
        if (it != g_tile_cache.end()) c.tile_cfg[w][cls] = it->second;
      }
  }
  std::vector<int> ipm(m->int_pair_map, m->int_pair_map + m->ntypes * m->ntypes);
  c.d_int_pair_map = dalloc_copy(ipm);
  build_family(c, 0);
  if (m->c3.nbody && m->c3.ncols > 0) build_family(c, 1);
  if (m->c4.nbody && m->c4.ncols > 0) build_family(c, 2);
}

void ctx_free(Ctx &c) {
  if (c.stream) (void)hipStreamSynchronize(c.stream);
  for_each_dbuf(c, [](DBuf &b) { b.release(); });
  if (c.stream) (void)hipStreamSynchronize(c.stream);
  for (auto &F : c.fam) {
    for (void *p : {(void *)F.ncol_type, (void *)F.col_offset, (void *)F.colptr, (void *)F.perm, (void *)F.smin, (void *)F.smax,
                    (void *)F.xavg, (void *)F.xdiff, (void *)F.lambda, (void *)F.recip, (void *)F.cct, (void *)F.csmax,
                    (void *)F.cperm, (void *)F.cslot, (void *)F.sets8, (void *)F.ptp})
      if (p) (void)hipFree(p);
  }
  drop_graphs(c);
  if (c.stats.d_min) (void)hipFree(c.stats.d_min);
  if (c.stats.d_flat) (void)hipFree(c.stats.d_flat);
  c.stats = LsqStats{};
  if (c.h_pin) (void)hipHostFree(c.h_pin);
  for (double *p : c.h_pin_retired) (void)hipHostFree(p);
  c.h_pin_retired.clear();
  c.h_pin = nullptr;
  c.h_pin_cap = 0;
  if (c.h_atot) (void)hipHostFree(c.h_atot);
  c.h_atot = nullptr;
  if (c.d_int_pair_map) (void)hipFree(c.d_int_pair_map);
  for (auto &e : c.ev) if (e) (void)hipEventDestroy(e);
  for (auto &fe : c.fev) for (auto &e : fe) if (e) (void)hipEventDestroy(e);
  for (auto &fc : c.cev) for (auto &fe : fc) for (auto &e : fe) if (e) (void)hipEventDestroy(e);
  for (auto &st : c.aux) if (st) (void)hipStreamDestroy(st);
  if (c.fork) (void)hipEventDestroy(c.fork);
  for (auto &e : c.join) if (e) (void)hipEventDestroy(e);
  if (c.stream) (void)hipStreamDestroy(c.stream);
}

}  // namespace cg

// ------------------------------------------------------------ C ABI

extern "C" {

cg_status cg_create(const cg_model *m, int device, cg_ctx **out) {
  if (!m || !out) { cg::set_error("null argument"); return CG_ERR_ARG; }
  auto *x = new cg_ctx;
  try {
    cg::ctx_init(x->c, m, device);
    *out = x;
    return CG_OK;
  } catch (const cg::Error &e) {
    cg::set_error(e.what());
    cg::ctx_free(x->c);
    delete x;
    return e.code;
  }
}

void cg_destroy(cg_ctx *ctx) {
  if (!ctx) return;
  cg::ctx_free(ctx->c);
  delete ctx;
}

cg_status cg_assemble(cg_ctx *ctx, const cg_frame *frames, int nframes, int prec, const void **A_dev,
                      double *b_host, int64_t *rows_out) {
  if (!ctx || !frames) { cg::set_error("null argument"); return CG_ERR_ARG; }
  if (prec != CG_FP64 && prec != CG_FP32) { cg::set_error("bad precision"); return CG_ERR_ARG; }
  try {
    CG_HIP(hipSetDevice(ctx->c.device));
    cg::assemble(ctx->c, frames, nframes, prec, b_host);
    if (A_dev) *A_dev = ctx->c.A;
    if (rows_out) *rows_out = ctx->c.A_rows;
    return CG_OK;
  } catch (const cg::Error &e) {
    cg::set_error(e.what());
    return e.code;
  }
}

cg_status cg_autotune(cg_ctx *ctx, const cg_frame *frames, int nframes, int prec, int *choice, double *ms_table) {
  if (!ctx || !frames || nframes <= 0) { cg::set_error("cg_autotune: bad argument"); return CG_ERR_ARG; }
  if (prec != CG_FP64 && prec != CG_FP32) { cg::set_error("bad precision"); return CG_ERR_ARG; }
  constexpr int NK = cg::kNumTileCfg, NC = CG_TILE_CLASSES;
  cg::Ctx &c = ctx->c;
  int keep[3][NC];
  std::memcpy(keep, c.tile_cfg, sizeof(keep));
  // every tuning assembly runs eagerly (a graph replay has no per-launch times)
  struct GraphsOff {
    cg::Ctx &c;
    bool was;
    explicit GraphsOff(cg::Ctx &cc) : c(cc), was(cc.asm_graphs) { c.asm_graphs = false; }
    ~GraphsOff() { c.asm_graphs = was; }
  } graphs_off(c);
  try {
    CG_HIP(hipSetDevice(c.device));
    double ms[NK][3][NC] = {};
    bool present[NC] = {};
    for (int i = 0; i < nframes; ++i) present[cg::tile_class(frames[i].natoms)] = true;
    c.split_classes = true;
    for (int k = 0; k < NK; ++k) {
      for (int w = 0; w < 3; ++w)
        for (int cls = 0; cls < NC; ++cls) c.tile_cfg[w][cls] = w == 0 ? 0 : k;
      if (k == 0) cg::assemble(c, frames, nframes, prec, nullptr);  // warm-up (buffers, code)
      cg::assemble(c, frames, nframes, prec, nullptr);
      for (int w = 0; w < 3; ++w)
        for (int cls = 0; cls < NC; ++cls) ms[k][w][cls] = c.ms_cls[w][cls];
    }
    c.split_classes = false;
    // fastest configuration per (family, class) from the timing runs
    int best[3][NC] = {};
    for (int w = 1; w < 3; ++w)
      for (int cls = 0; cls < NC; ++cls)
        for (int k = 1; k < NK; ++k)
          if (ms[k][w][cls] > 0.0 && ms[k][w][cls] < ms[best[w][cls]][w][cls]) best[w][cls] = k;
    // a family keeps its per-class choices only if they beat the default tiles
    // on a whole assembly (one launch per configuration, launches side by side)
    float whole[2][3];
    for (int pass = 0; pass < 2; ++pass) {
      for (int w = 0; w < 3; ++w)
        for (int cls = 0; cls < NC; ++cls) c.tile_cfg[w][cls] = pass == 0 ? 0 : best[w][cls];
      cg::assemble(c, frames, nframes, prec, nullptr);
      for (int w = 0; w < 3; ++w) whole[pass][w] = c.ms_fam[w];
    }
    std::lock_guard<std::mutex> lk(cg::g_tile_mu);
    for (int w = 0; w < 3; ++w) {
      const bool keep_best = w > 0 && whole[1][w] < whole[0][w];
      for (int cls = 0; cls < NC; ++cls) {
        const int k = keep_best ? best[w][cls] : 0;
        c.tile_cfg[w][cls] = k;
        if (w > 0 && present[cls] && ms[0][w][cls] > 0.0) cg::g_tile_cache[cg::tile_key(c.m, w, cls)] = k;
        if (choice) choice[w * NC + cls] = present[cls] ? k : -1;
      }
    }
    if (ms_table)
      for (int k = 0; k < NK; ++k)
        for (int w = 0; w < 3; ++w)
          for (int cls = 0; cls < NC; ++cls) ms_table[(k * 3 + w) * NC + cls] = ms[k][w][cls];
    return CG_OK;
  } catch (const cg::Error &e) {
    c.split_classes = false;
    std::memcpy(c.tile_cfg, keep, sizeof(keep));  // back to the configuration before the call
    cg::set_error(e.what());
    return e.code;
  }
}

cg_status cg_tile_choice(cg_ctx *ctx, int *choice) {
  if (!ctx || !choice) { cg::set_error("cg_tile_choice: bad argument"); return CG_ERR_ARG; }
  for (int w = 0; w < 3; ++w)
    for (int cls = 0; cls < CG_TILE_CLASSES; ++cls) choice[w * CG_TILE_CLASSES + cls] = ctx->c.tile_cfg[w][cls];
  return CG_OK;
}

cg_status cg_tile_set(cg_ctx *ctx, const int *choice) {
  if (!ctx || !choice) { cg::set_error("cg_tile_set: bad argument"); return CG_ERR_ARG; }
  for (int i = 0; i < 3 * CG_TILE_CLASSES; ++i)
    if (choice[i] < 0 || choice[i] >= cg::kNumTileCfg || (i < CG_TILE_CLASSES && choice[i] != 0)) {
      cg::set_error("cg_tile_set: configuration out of range (the 2-body family has only configuration 0)");
      return CG_ERR_ARG;
    }
  for (int w = 0; w < 3; ++w)
    for (int cls = 0; cls < CG_TILE_CLASSES; ++cls) ctx->c.tile_cfg[w][cls] = choice[w * CG_TILE_CLASSES + cls];
  return CG_OK;
}

cg_status cg_copy_A(cg_ctx *ctx, int64_t r0, int64_t nrows, double *host) {
  if (!ctx || !host || !ctx->c.A) { cg::set_error("nothing assembled"); return CG_ERR_ARG; }
  auto &c = ctx->c;
  if (r0 < 0 || r0 + nrows > c.A_rows) { cg::set_error("row range"); return CG_ERR_ARG; }
  try {
    size_t n = (size_t)nrows * c.A_ld;
    if (c.A_prec == CG_FP64) {
      CG_HIP(hipMemcpy(host, (double *)c.A + r0 * c.A_ld, n * sizeof(double), hipMemcpyDeviceToHost));
    } else {
      std::vector<float> tmp(n);
      CG_HIP(hipMemcpy(tmp.data(), (float *)c.A + r0 * c.A_ld, n * sizeof(float), hipMemcpyDeviceToHost));
      for (size_t i = 0; i < n; ++i) host[i] = tmp[i];
    }
    return CG_OK;
  } catch (const cg::Error &e) {
    cg::set_error(e.what());
    return e.code;
  }
}

cg_status cg_append_rows(cg_ctx *ctx, const double *rows, int nrows) {
  if (!ctx || !ctx->c.A) { cg::set_error("nothing assembled"); return CG_ERR_ARG; }
  if (nrows < 0 || (nrows > 0 && !rows)) { cg::set_error("cg_append_rows: bad rows"); return CG_ERR_ARG; }
  if (nrows == 0) return CG_OK;
  auto &c = ctx->c;
  try {
    CG_HIP(hipSetDevice(c.device));
    const size_t esz = c.A_prec == CG_FP32 ? sizeof(float) : sizeof(double);
    const size_t old_bytes = (size_t)c.A_rows * c.A_ld * esz;
    const size_t add = (size_t)nrows * c.A_ld;
    if (old_bytes + add * esz > c.b_A.cap) {
      // grow, keeping the assembled rows
      void *p = nullptr;
      const size_t n = old_bytes + add * esz;
      if (hipMalloc(&p, n + n / 4 + 256) != hipSuccess) throw cg::Error(CG_ERR_MEMORY, "A does not fit in device memory");
      CG_HIP(hipMemcpyAsync(p, c.A, old_bytes, hipMemcpyDeviceToDevice, c.stream));
      CG_HIP(hipStreamSynchronize(c.stream));
      c.b_A.release();
      c.b_A.p = p;
      c.b_A.cap = n + n / 4 + 256;
      c.A = p;
    }
    char *dst = static_cast<char *>(c.A) + old_bytes;
    if (c.A_prec == CG_FP32) {
      std::vector<float> tmp(rows, rows + add);
      CG_HIP(hipMemcpy(dst, tmp.data(), add * sizeof(float), hipMemcpyHostToDevice));
    } else {
      CG_HIP(hipMemcpy(dst, rows, add * sizeof(double), hipMemcpyHostToDevice));
    }
    c.A_rows += nrows;
    return CG_OK;
  } catch (const cg::Error &e) {
    cg::set_error(e.what());
    return e.code;
  }
}

cg_status cg_force_batch(cg_ctx *ctx, const double *x, const cg_frame *frames, int nframes, double *F, double *E,
                         double *virial9) {
  if (!ctx || !x || !frames || nframes <= 0 || !F) { cg::set_error("null argument"); return CG_ERR_ARG; }
  try {
    CG_HIP(hipSetDevice(ctx->c.device));
    cg::force_batch(ctx->c, x, frames, nframes, F, E, virial9);
    return CG_OK;
  } catch (const cg::Error &e) {
    cg::set_error(e.what());
    return e.code;
  }
}

cg_status cg_force(cg_ctx *ctx, const double *x, const cg_frame *f, double *F, double *E, double *virial9) {
  if (!ctx || !x || !f || !F) { cg::set_error("null argument"); return CG_ERR_ARG; }
  try {
    CG_HIP(hipSetDevice(ctx->c.device));
    cg::force_batch(ctx->c, x, f, 1, F, E, virial9);
    return CG_OK;
  } catch (const cg::Error &e) {
    cg::set_error(e.what());
    return e.code;
  }
}

cg_status cg_force_noalloc(cg_ctx *ctx, const double *x, const cg_frame *f, double *F, double *E, double *virial9) {
  if (!ctx || !x || !f || !F) { cg::set_error("null argument"); return CG_ERR_ARG; }
  if (f->natoms <= 0 || !f->xyz || !f->type) { cg::set_error("cg_force_noalloc: bad frame"); return CG_ERR_ARG; }
  try {
    CG_HIP(hipSetDevice(ctx->c.device));
    cg::force_noalloc(ctx->c, x, f, F, E, virial9);
    return CG_OK;
  } catch (const cg::Error &e) {
    cg::set_error(e.what());
    return e.code;
  }
}

cg_status cg_last_detail(cg_ctx *ctx, double *ms_family3, int64_t *clusters3) {
  if (!ctx) return CG_ERR_ARG;
  for (int w = 0; w < 3; ++w) {
    if (ms_family3) ms_family3[w] = ctx->c.ms_fam[w];
    if (clusters3) clusters3[w] = ctx->c.ncl[w];
  }
  return CG_OK;
}

cg_status cg_copy_probe(cg_ctx *ctx, size_t bytes) {
  if (!ctx || bytes == 0) { cg::set_error("bad argument"); return CG_ERR_ARG; }
  try {
    cg::Ctx &c = ctx->c;
    CG_HIP(hipSetDevice(c.device));
    char *buf = static_cast<char *>(c.b_tmp.get(2 * bytes));
    if (!buf) throw cg::Error(CG_ERR_MEMORY, "copy probe buffers");
    CG_HIP(hipMemsetAsync(buf, 1, bytes, c.stream));
    CG_HIP(hipStreamSynchronize(c.stream));
    CG_HIP(hipMemcpyAsync(buf + bytes, buf, bytes, hipMemcpyDeviceToDevice, c.stream));
    CG_HIP(hipStreamSynchronize(c.stream));
    return CG_OK;
  } catch (const cg::Error &e) {
    cg::set_error(e.what());
    return e.code;
  }
}

cg_status cg_force_stats(cg_ctx *ctx, int64_t *calls, int64_t *repeats, int64_t *graph_replays) {
  if (!ctx) return CG_ERR_ARG;
  if (calls) *calls = ctx->c.n_force;
  if (repeats) *repeats = ctx->c.n_repeat;
  if (graph_replays) *graph_replays = ctx->c.n_graph;
  return CG_OK;
}

cg_status cg_set_penalty(cg_ctx *ctx, double dist, double scale) {
  if (!ctx || !(dist >= 0.0) || !(scale >= 0.0)) { cg::set_error("bad penalty"); return CG_ERR_ARG; }
  cg::Ctx &c = ctx->c;
  if (dist != c.penalty_dist || scale != c.penalty_scale) {
    // captured force calls hold the penalty in their kernel arguments
    cg::drop_graphs(c);
    c.last_key.clear();
  }
  c.penalty_dist = dist;
  c.penalty_scale = scale;
  return CG_OK;
}

cg_status cg_stats_enable(cg_ctx *ctx, int on) {
  if (!ctx) { cg::set_error("null argument"); return CG_ERR_ARG; }
  try {
    cg::Ctx &c = ctx->c;
    CG_HIP(hipSetDevice(c.device));
    cg::LsqStats &S = c.stats;
    S.on = on != 0;
    if (!S.on) return CG_OK;
    const int n2 = c.fam[0].ntypes, n3 = c.fam[1].ntypes * 3, n4 = c.fam[2].ntypes * 6;
    S.off3 = n2;
    S.off4 = n2 + n3;
    if (S.n_min < n2 + n3 + n4 || !S.d_min) {
      if (S.d_min) (void)hipFree(S.d_min);
      S.d_min = nullptr;
      CG_HIP(hipMalloc(&S.d_min, std::max(1, n2 + n3 + n4) * sizeof(unsigned long long)));
    }
    S.n_min = n2 + n3 + n4;
    // every minimum starts at 1e10, the reference's "not found" value
    std::vector<unsigned long long> init(std::max(1, S.n_min), __builtin_bit_cast(unsigned long long, 1.0e10));
    CG_HIP(hipMemcpy(S.d_min, init.data(), init.size() * sizeof(unsigned long long), hipMemcpyHostToDevice));
    for (int w = 0; w < 3; ++w) S.count[w].assign(c.fam[w].ntypes, 0);
    S.flat3.clear();
    S.flat4.clear();
    return CG_OK;
  } catch (const cg::Error &e) {
    cg::set_error(e.what());
    return e.code;
  }
}

cg_status cg_stats_get(cg_ctx *ctx, cg_lsq_stats *out) {
  if (!ctx || !out) { cg::set_error("null argument"); return CG_ERR_ARG; }
  cg::Ctx &c = ctx->c;
  cg::LsqStats &S = c.stats;
  if (!S.d_min) { cg::set_error("statistics were not enabled"); return CG_ERR_ARG; }
  try {
    CG_HIP(hipSetDevice(c.device));
    std::vector<unsigned long long> h(std::max(1, S.n_min));
    CG_HIP(hipStreamSynchronize(c.stream));
    CG_HIP(hipMemcpy(h.data(), S.d_min, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    auto val = [&](int i) { return __builtin_bit_cast(double, h[i]); };
    const int n2 = c.fam[0].ntypes, nt3 = c.fam[1].ntypes, nt4 = c.fam[2].ntypes;
    for (int t = 0; t < n2; ++t) {
      if (out->pair_min) out->pair_min[t] = val(t);
      if (out->pair_count) out->pair_count[t] = S.count[0].empty() ? 0 : S.count[0][t];
    }
    for (int t = 0; t < nt3; ++t) {
      for (int k = 0; k < 3; ++k) if (out->trip_min) out->trip_min[3 * t + k] = val(S.off3 + 3 * t + k);
      if (out->trip_count) out->trip_count[t] = S.count[1].empty() ? 0 : S.count[1][t];
    }
    for (int t = 0; t < nt4; ++t) {
      for (int k = 0; k < 6; ++k) if (out->quad_min) out->quad_min[6 * t + k] = val(S.off4 + 6 * t + k);
      if (out->quad_count) out->quad_count[t] = S.count[2].empty() ? 0 : S.count[2][t];
    }
    return CG_OK;
  } catch (const cg::Error &e) {
    cg::set_error(e.what());
    return e.code;
  }
}

cg_status cg_stats_frames(cg_ctx *ctx, int64_t *flat3, int64_t *flat4, int *nframes) {
  if (!ctx) { cg::set_error("null argument"); return CG_ERR_ARG; }
  const cg::LsqStats &S = ctx->c.stats;
  if (nframes) *nframes = (int)S.flat3.size();
  for (size_t i = 0; i < S.flat3.size(); ++i) {
    if (flat3) flat3[i] = S.flat3[i];
    if (flat4) flat4[i] = S.flat4[i];
  }
  return CG_OK;
}

cg_status cg_synchronize(cg_ctx *ctx) {
  if (!ctx) { cg::set_error("null argument"); return CG_ERR_ARG; }
  try {
    CG_HIP(hipSetDevice(ctx->c.device));
    CG_HIP(hipStreamSynchronize(ctx->c.stream));
    return CG_OK;
  } catch (const cg::Error &e) {
    cg::set_error(e.what());
    return e.code;
  }
}

cg_status cg_assembly_graphs(cg_ctx *ctx, int on) {
  if (!ctx) { cg::set_error("cg_assembly_graphs: bad argument"); return CG_ERR_ARG; }
  ctx->c.asm_graphs = on != 0;
  return CG_OK;
}

cg_status cg_assembly_deterministic(cg_ctx *ctx, int on) {
  if (!ctx) { cg::set_error("cg_assembly_deterministic: bad argument"); return CG_ERR_ARG; }
  ctx->c.asm_det = on != 0;
  return CG_OK;
}

cg_status cg_assembly_pair_tables(cg_ctx *ctx, int on, int *active) {
  if (!ctx) { cg::set_error("cg_assembly_pair_tables: bad argument"); return CG_ERR_ARG; }
  ctx->c.asm_pcache = on != 0;
  if (active) {
    *active = 0;
    for (int w = 1; w < 3; ++w)
      if (on && ctx->c.fam[w].pcache && ctx->c.fam[w].ncols > 0) *active |= w;
  }
  return CG_OK;
}

cg_status cg_assembly_global_above(cg_ctx *ctx, int natoms) {
  if (!ctx || natoms < 0) { cg::set_error("cg_assembly_global_above: bad argument"); return CG_ERR_ARG; }
  ctx->c.gacc_above = natoms;
  return CG_OK;
}

cg_status cg_assembly_stats(cg_ctx *ctx, int64_t *calls, int64_t *repeats, int64_t *graph_replays,
                            int64_t *captures) {
  if (!ctx) { cg::set_error("cg_assembly_stats: bad argument"); return CG_ERR_ARG; }
  const cg::Ctx &c = ctx->c;
  if (calls) *calls = c.n_asm;
  if (repeats) *repeats = c.n_asm_repeat;
  if (graph_replays) *graph_replays = c.n_asm_graph;
  if (captures) *captures = c.n_asm_capture;
  return CG_OK;
}

cg_status cg_last_timing(cg_ctx *ctx, double *ms_enum, double *ms_asm, double *ms_total) {
  if (!ctx) return CG_ERR_ARG;
  if (ms_enum) *ms_enum = ctx->c.ms_enum;
  if (ms_asm) *ms_asm = ctx->c.ms_asm;
  if (ms_total) *ms_total = ctx->c.ms_total;
  return CG_OK;
}

}  // extern "C"
