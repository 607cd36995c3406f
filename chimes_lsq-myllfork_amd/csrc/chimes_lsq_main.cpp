// chimes_lsq_gpu: drop-in for `chimes_lsq fm_setup.in` (src/chimes_lsq.C:108-424)
// with the design matrix assembled on the GPU, optionally followed by the fit
// that `chimes_lsq.py` performs (src/chimes_lsq.py:169-672).
//
//   chimes_lsq_gpu fm_setup.in [--fit] [--algorithm svd|ridge|lassolars] [--eps E] [--alpha A] [--weights FILE]
//                  [--prec fp64|fp32] [--gram qr|fp64|bf16x3]
//                  [--refine K] [--batch FRAMES] [--ranks N] [--no-A] [--no-text]
//
// Files written in the working directory: A.txt b.txt b-labeled.txt natoms.txt
// dim.txt params.header ff_groups.map frames.NNNN.log; with --fit also
// params.txt force.txt.
// --ranks N lays the frames out as the reference's N MPI ranks would process
// them (contiguous blocks, or round robin with SKPFRMS), rows in rank order;
// with SPLITFI true each rank's rows go to A.<rank>.txt with dim.<rank>.txt.
// --no-A skips the A text (b, natoms and dim files are still written).
// Standard output is what chimes_lsq prints on rank 0 of an N-rank run (the
// reference test suite's fm_setup.out); this program's own summary (GPU
// times, the fit) goes to standard error.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/chimes_gpu.h"

namespace {
struct OwnedFrame {
  cg_frame f;
  std::vector<double> xyz, forces;
  std::vector<int> type;
};

int die(const char *what, cg_status st, const void *ctx = nullptr) {
  std::fprintf(stderr, "Error (%s): status %d: %s\n", what, st, cg_last_error(ctx));
  return 1;
}
}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "Usage: chimes_lsq_gpu input_file [--fit] [--algorithm svd|ridge|lassolars] [--eps E] [--alpha A] [--weights FILE] "
                         "[--prec fp64|fp32] [--gram qr|fp64|bf16x3] [--refine K] "
                         "[--batch FRAMES] [--ranks N] [--no-A] [--no-text]\n");
    return 1;
  }
  const char *infile = argv[1];
  bool do_fit = false, text = true, write_A = true;
  double eps = 1.0e-5, alpha = 1.0e-4;
  std::string algorithm = "svd", weights;
  int prec = CG_FP64, refine = 3, batch = 0, gram = CG_FIT_QR, nranks = 1;
  for (int i = 2; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "--fit") do_fit = true;
    else if (a == "--no-text") text = false;
    else if (a == "--no-A") write_A = false;
    else if (a == "--ranks" && i + 1 < argc) nranks = std::atoi(argv[++i]);
    else if (a == "--eps" && i + 1 < argc) eps = std::atof(argv[++i]);
    else if (a == "--alpha" && i + 1 < argc) alpha = std::atof(argv[++i]);
    else if (a == "--algorithm" && i + 1 < argc) algorithm = argv[++i];
    else if (a == "--weights" && i + 1 < argc) weights = argv[++i];
    else if (a == "--refine" && i + 1 < argc) refine = std::atoi(argv[++i]);
    else if (a == "--batch" && i + 1 < argc) batch = std::atoi(argv[++i]);
    else if (a == "--prec" && i + 1 < argc) {
      std::string p = argv[++i];
      prec = p == "fp32" ? CG_FP32 : CG_FP64;
    } else if (a == "--gram" && i + 1 < argc) {
      std::string p = argv[++i];
      gram = p == "bf16x3" ? CG_GRAM_BF16X3 : p == "fp64" ? CG_GRAM_FP64 : CG_FIT_QR;
    } else {
      std::fprintf(stderr, "unknown option %s\n", a.c_str());
      return 1;
    }
  }
  auto t0 = std::chrono::steady_clock::now();
  cg_setup *setup = nullptr;
  cg_status st = cg_setup_read(infile, &setup);
  if (st) return die("reading input", st);
  const cg_model *m = cg_setup_model(setup);
  if ((st = cg_report_setup(setup, stdout))) return die("report", st);
  // read every frame
  std::string dir = ".";
  cg_traj *traj = nullptr;
  if ((st = cg_traj_open(setup, dir.c_str(), &traj))) return die("opening trajectory", st);
  const int nframes = cg_setup_nframes(setup);
  std::vector<OwnedFrame> frames(nframes);
  for (int i = 0; i < nframes; ++i) {
    cg_frame f;
    if ((st = cg_traj_next(traj, &f))) return die("reading trajectory", st);
    OwnedFrame &o = frames[i];
    o.xyz.assign(f.xyz, f.xyz + 3 * f.natoms);
    o.forces.assign(f.forces, f.forces + 3 * f.natoms);
    o.type.assign(f.type, f.type + f.natoms);
    o.f = f;
    o.f.xyz = o.xyz.data();
    o.f.forces = o.forces.data();
    o.f.type = o.type.data();
  }
  cg_traj_close(traj);
  // frames in the order the reference's ranks process them
  if (nranks < 1) { std::fprintf(stderr, "--ranks must be >= 1\n"); return 1; }
  std::vector<int> order(nframes), rank_of(nframes);
  int nproc = 0;
  if ((st = cg_frame_order(setup, nranks, order.data(), rank_of.data(), &nproc))) return die("frame order", st);
  cg_ctx *ctx = nullptr;
  if ((st = cg_create(m, 0, &ctx))) return die("creating device context", st);
  if ((st = cg_stats_enable(ctx, 1))) return die("statistics", st, ctx);
  std::vector<int64_t> flat3(nframes, 0), flat4(nframes, 0);
  std::vector<double> frame_ms(nframes, 0.0);
  if (text && (st = cg_write_open_opts(setup, dir.c_str(), write_A ? 0 : CG_WRITE_NO_A)))
    return die("opening A/b files", st);
  std::vector<cg_frame> cf(nproc);
  for (int i = 0; i < nproc; ++i) cf[i] = frames[order[i]].f;
  const int nf_run = nproc;
  if (do_fit || batch <= 0) batch = std::max(nf_run, 1);
  double ms_asm = 0;
  int64_t total_rows = 0;
  std::vector<double> b_all;
  for (int f0 = 0; f0 < nf_run; f0 += batch) {
    int nb = std::min(batch, nf_run - f0);
    int64_t rows = 0;
    for (int i = 0; i < nb; ++i) rows += cg_model_frame_rows(m, cf[f0 + i].natoms, cf[f0 + i].index);
    std::vector<double> b(rows);
    const void *A = nullptr;
    if ((st = cg_assemble(ctx, cf.data() + f0, nb, prec, &A, b.data(), &rows))) return die("assembling", st, ctx);
    double te, ta, tt;
    cg_last_timing(ctx, &te, &ta, &tt);
    ms_asm += tt;
    {
      std::vector<int64_t> f3(nb), f4(nb);
      int got = 0;
      if ((st = cg_stats_frames(ctx, f3.data(), f4.data(), &got))) return die("statistics", st, ctx);
      for (int i = 0; i < nb && i < got; ++i) {
        flat3[order[f0 + i]] = f3[i];
        flat4[order[f0 + i]] = f4[i];
      }
      // a frame's share of the batch's GPU time, by its rows
      for (int i = 0; i < nb; ++i)
        frame_ms[order[f0 + i]] = tt * cg_model_frame_rows(m, cf[f0 + i].natoms, cf[f0 + i].index) / std::max<int64_t>(rows, 1);
    }
    b_all.insert(b_all.end(), b.begin(), b.end());
    if (text) {
      std::vector<double> host(write_A ? (size_t)rows * m->ncols : 0);
      if (write_A && (st = cg_copy_A(ctx, 0, rows, host.data()))) return die("copying A", st, ctx);
      int64_t r = 0;
      for (int i = 0; i < nb; ++i) {
        int nr = (int)cg_model_frame_rows(m, cf[f0 + i].natoms, cf[f0 + i].index);
        if ((st = cg_write_rank(setup, rank_of[f0 + i]))) return die("writing A", st);
        if ((st = cg_write_frame(setup, &cf[f0 + i], write_A ? host.data() + r * m->ncols : nullptr, nr)))
          return die("writing A", st);
        r += nr;
      }
    }
    total_rows += rows;
  }
  if (text && (st = cg_write_close(setup))) return die("closing A/b files", st);
  if ((st = cg_write_header(setup, dir.c_str()))) return die("writing params.header", st);
  {
    // the reference's standard output, rank 0 of an nranks run: every frame
    // read in order, rank 0's frames processed, then the statistics
    std::vector<int> rank_in(nframes, -1);
    for (int i = 0; i < nproc; ++i) rank_in[order[i]] = rank_of[i];
    bool first = true;
    int last0 = -1;
    int64_t force_rows = 0;
    for (int i = 0; i < nframes; ++i) {
      if ((st = cg_report_frame_read(setup, &frames[i].f, stdout))) return die("report", st);
      if (rank_in[i] >= 0) force_rows += 3LL * frames[i].f.natoms;
      if (rank_in[i] != 0) continue;
      if ((st = cg_report_frame_done(setup, &frames[i].f, first ? 1 : 0, nranks, flat3[i], flat4[i], stdout)))
        return die("report", st);
      first = false;
      last0 = i;
    }
    std::vector<int> counts(m->ntypes, 0);
    if (last0 >= 0)
      for (int a = 0; a < frames[last0].f.natoms; ++a) ++counts[frames[last0].f.type[a]];
    const int nt3 = m->c3.nbody ? m->c3.ntypes : 0, nt4 = m->c4.nbody ? m->c4.ntypes : 0;
    std::vector<double> pmin(m->npairs), tmin(3 * (size_t)nt3 + 1), qmin(6 * (size_t)nt4 + 1);
    std::vector<int64_t> pcnt(m->npairs), tcnt(nt3 + 1), qcnt(nt4 + 1);
    cg_lsq_stats sts{pmin.data(), pcnt.data(), tmin.data(), tcnt.data(), qmin.data(), qcnt.data()};
    if ((st = cg_stats_get(ctx, &sts))) return die("statistics", st, ctx);
    if ((st = cg_report_finish(setup, force_rows, counts.data(), &sts, stdout))) return die("report", st);
    std::vector<cg_frame> all(nframes);
    std::vector<double> secs(nframes);
    for (int i = 0; i < nframes; ++i) {
      all[i] = frames[i].f;
      secs[i] = frame_ms[i] * 1e-3;
    }
    if ((st = cg_write_frame_logs(setup, dir.c_str(), nranks, all.data(), secs.data()))) return die("frame logs", st);
  }
  std::fprintf(stderr, "assembly: %lld rows x %d columns, %.3f ms on the GPU (%.3e rows/s)\n",
               (long long)total_rows, m->ncols, ms_asm, total_rows / (ms_asm * 1e-3));
  if (do_fit) {
    // FITCOUL charge-constraint rows follow the last frame's rows
    int ncons = 0;
    if ((st = cg_setup_constraint_rows(setup, nullptr, nullptr, &ncons))) return die("constraint rows", st);
    if (ncons > 0) {
      std::vector<double> cr((size_t)ncons * m->ncols), cb(ncons);
      if ((st = cg_setup_constraint_rows(setup, cr.data(), cb.data(), &ncons))) return die("constraint rows", st);
      if ((st = cg_append_rows(ctx, cr.data(), ncons))) return die("constraint rows", st, ctx);
      b_all.insert(b_all.end(), cb.begin(), cb.end());
      total_rows += ncons;
    }
    std::vector<double> x(m->ncols), y(total_rows);
    // hierarchical exclusions: fit the kept columns; params.txt takes the
    // kept coefficients in order, as chimes_lsq.py reads the reduced A.txt
    std::vector<int> keep(m->ncols);
    int nkeep = 0;
    if ((st = cg_setup_fit_columns(setup, keep.data(), &nkeep))) return die("fit columns", st);
    if ((st = cg_set_fit_columns(ctx, keep.data(), nkeep < m->ncols ? nkeep : 0))) return die("fit columns", st, ctx);
    cg_fit_result res{};
    if (!weights.empty() && algorithm == "svd") {
      // one weight per row of b (src/chimes_lsq.py:75-129)
      std::vector<double> w;
      std::FILE *wf = std::fopen(weights.c_str(), "r");
      if (!wf) { std::fprintf(stderr, "cannot read %s\n", weights.c_str()); return 1; }
      double v;
      while (std::fscanf(wf, "%lf", &v) == 1) w.push_back(v);
      std::fclose(wf);
      if ((int64_t)w.size() != total_rows) { std::printf("Wrong number of lines in WEIGHTS file\n"); return 1; }
      if ((st = cg_set_fit_weights(ctx, w.data(), (int64_t)w.size()))) return die("weights", st, ctx);
      gram = CG_FIT_QR;
    }
    if (!weights.empty() && algorithm == "lassolars") {
      std::fprintf(stderr, "--weights with --algorithm lassolars is not supported\n");
      return 1;
    }
    if (algorithm == "ridge") {
      if ((st = cg_fit_ridge(ctx, b_all.data(), alpha, x.data(), &res))) return die("fitting", st, ctx);
    } else if (algorithm == "lassolars") {
      if ((st = cg_fit_lassolars(ctx, b_all.data(), alpha, 100000, x.data(), &res))) return die("fitting", st, ctx);
    } else if (algorithm == "svd") {
      if ((st = cg_fit(ctx, b_all.data(), eps, refine, gram, x.data(), &res))) return die("fitting", st, ctx);
    } else {
      std::printf("Unrecognized fitting algorithm\n");
      return 1;
    }
    if (!weights.empty()) std::snprintf(res.weights_file, sizeof res.weights_file, "%s", weights.c_str());
    if ((st = cg_matvec(ctx, x.data(), y.data()))) return die("computing A x", st, ctx);
    std::vector<double> xk(nkeep);
    for (int i = 0; i < nkeep; ++i) xk[i] = x[keep[i]];
    if ((st = cg_write_params(setup, dir.c_str(), xk.data(), nkeep, total_rows, &res, y.data(), b_all.data())))
      return die("writing params.txt", st);
    std::fprintf(stderr, "fit: RMS %.6e, %d of %d variables kept, gram %.3f ms, solve %.3f ms, refine %.3f ms\n",
                 res.rms, res.nvars, m->ncols, res.ms_gram, res.ms_solve, res.ms_refine);
  }
  cg_destroy(ctx);
  cg_setup_free(setup);
  std::chrono::duration<double> dt = std::chrono::steady_clock::now() - t0;
  std::fprintf(stderr, "total wall time %.3f s\n", dt.count());
  return 0;
}
