// Force-evaluation server with continuous batching.
//
// The reference evaluates one configuration per MD step through ZCalc
// (src/functions.C:815-891, src/chimes_md.C:755); it has no request queue.
// Here independent configurations are submitted from any thread and served
// by worker threads, each owning a device context (own HIP stream and
// buffers).  A worker takes every queued request that fits the batch limits
// (atom count, and an HBM budget shared by all batches in flight) and
// evaluates them in one cg_force_batch pass; requests that arrive while a
// batch runs are admitted by the next free worker, so batches overlap on the
// device.  Results are kept per ticket until polled.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "device.hpp"
#include "setup.hpp"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  int64_t ticket = 0;
  std::vector<double> xyz;
  std::vector<int> type;
  cg_frame f{};
  size_t est_bytes = 0;
  std::vector<double> F;
  double E = 0.0;
  double vir[9] = {};
  cg_status status = CG_OK;
  std::string err;
  bool done = false;
  Clock::time_point t_submit, t_done;
};

// Device bytes a configuration needs in one batch: neighbour entries within
// the largest cutoff (half list, 44 bytes each plus two sublists), clusters
// per entry for the 3- and 4-body families, and per-atom arrays.
size_t estimate_bytes(const cg_model &m, const cg_frame &f) {
  const double *h = f.cell;
  const double vol = std::fabs(h[0] * (h[4] * h[8] - h[5] * h[7]) - h[1] * (h[3] * h[8] - h[5] * h[6]) +
                               h[2] * (h[3] * h[7] - h[4] * h[6]));
  double rc = m.max_cutoff_2b;
  if (m.c3.nbody) rc = std::max(rc, m.c3.max_cutoff);
  if (m.c4.nbody) rc = std::max(rc, m.c4.max_cutoff);
  const double rho = vol > 0 ? f.natoms / vol : 0.1;
  const double per_atom = 4.18879 * rc * rc * rc * rho;  // images within rc
  const double entries = 0.5 * f.natoms * per_atom + 27.0 * f.natoms;
  const double clusters = entries * std::max(1.0, 0.5 * per_atom);
  return (size_t)(4096 + 160.0 * f.natoms + 3.0 * 44.0 * entries + 96.0 * clusters);
}

struct Server {
  const cg_model *m = nullptr;
  int device = 0;
  std::vector<double> x;
  size_t budget = 0;
  int max_atoms = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> queue;
  std::unordered_map<int64_t, std::shared_ptr<Request>> live;
  int64_t next_ticket = 1;
  size_t inflight = 0;
  bool stop = false;
  bool hold = false;  // workers take no new batch while set (cg_serve_hold)
  int64_t n_batches = 0, n_requests = 0;
  int64_t n_isolated = 0;  // failed batches whose requests were re-run one by one
  // summed over the workers' contexts: calls replayed from a graph, calls
  // captured into a new graph, calls repeated with exact list sizes, and
  // buffer moves (allocations)
  int64_t n_replay = 0, n_capture = 0, n_repeat = 0, n_grow = 0;
  // largest neighbour / cluster counts per atom any worker has seen: every
  // worker sizes its lists from them, so that a configuration one worker has
  // met is not repeated with exact sizes on another
  double rate_nb = 0.0, rate_cl[3] = {0.0, 0.0, 0.0};
  std::vector<std::thread> workers;
  std::string init_error;
  int ready = 0;

  // Forces of requests [i0, i1) of a batch in one cg::force_batch pass into
  // F / E / V (resized); any failure, device or host, becomes a status.
  cg_status evaluate(cg::Ctx &ctx, const std::vector<std::shared_ptr<Request>> &batch, size_t i0, size_t i1,
                     std::vector<double> &F, std::vector<double> &E, std::vector<double> &V, std::string &err) {
    std::vector<cg_frame> frames;
    size_t tot = 0;
    for (size_t i = i0; i < i1; ++i) {
      frames.push_back(batch[i]->f);
      tot += 3 * (size_t)batch[i]->f.natoms;
    }
    try {
      F.assign(tot, 0.0);
      E.assign(frames.size(), 0.0);
      V.assign(9 * frames.size(), 0.0);
      CG_HIP(hipSetDevice(device));
      cg::force_batch(ctx, x.data(), frames.data(), (int)frames.size(), F.data(), E.data(), V.data());
      return CG_OK;
    } catch (const cg::Error &e) {
      err = e.what();
      return e.code;
    } catch (const std::bad_alloc &) {
      err = "host memory exhausted";
      return CG_ERR_MEMORY;
    } catch (const std::exception &e) {
      err = e.what();
      return CG_ERR_DEVICE;
    }
  }

  void worker() {
    cg::Ctx ctx;
    std::string ierr;
    try {
      cg::ctx_init(ctx, m, device);
    } catch (const std::exception &e) {
      ierr = e.what();
    }
    if (!ierr.empty()) {
      try { cg::ctx_free(ctx); } catch (...) {}
      std::lock_guard<std::mutex> lk(mu);
      init_error = ierr;
      ++ready;
      cv_done.notify_all();
      return;
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      ++ready;
      cv_done.notify_all();
    }
    std::vector<std::shared_ptr<Request>> batch;
    std::vector<double> F, E, V;
    for (;;) {
      size_t bytes = 0;
      batch.clear();
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] {
          return stop || (!hold && !queue.empty() && (inflight == 0 || inflight + queue.front()->est_bytes <= budget));
        });
        if (queue.empty()) break;  // stop requested and nothing left
        int atoms = 0;
        while (!queue.empty()) {
          const auto &r = queue.front();
          if (!batch.empty() && (atoms + r->f.natoms > max_atoms || inflight + bytes + r->est_bytes > budget)) break;
          atoms += r->f.natoms;
          bytes += r->est_bytes;
          batch.push_back(r);
          queue.pop_front();
        }
        inflight += bytes;
        ctx.rate_nb = std::max(ctx.rate_nb, rate_nb);
        for (int w = 0; w < 3; ++w) ctx.rate_cl[w] = std::max(ctx.rate_cl[w], rate_cl[w]);
      }
      // one pass over the whole batch; if it fails, each request is
      // evaluated alone so that only the requests that fail by themselves
      // report an error
      const int64_t g0 = ctx.n_graph, c0 = ctx.n_capture, r0 = ctx.n_repeat;
      const uint64_t a0 = cg::t_grow_events;
      std::string err;
      const cg_status st = evaluate(ctx, batch, 0, batch.size(), F, E, V, err);
      if (st == CG_OK) {
        size_t o = 0;
        for (size_t i = 0; i < batch.size(); ++i) {
          Request &r = *batch[i];
          const size_t n3 = 3 * (size_t)r.f.natoms;
          r.F.assign(F.begin() + o, F.begin() + o + n3);
          r.E = E[i];
          std::memcpy(r.vir, &V[9 * i], sizeof(r.vir));
          r.status = CG_OK;
          o += n3;
        }
      } else if (batch.size() == 1) {
        batch[0]->status = st;
        batch[0]->err = err;
      } else {
        for (size_t i = 0; i < batch.size(); ++i) {
          Request &r = *batch[i];
          std::string e1;
          r.status = evaluate(ctx, batch, i, i + 1, F, E, V, e1);
          r.err = e1;
          if (r.status == CG_OK) {
            r.F = F;
            r.E = E[0];
            std::memcpy(r.vir, V.data(), sizeof(r.vir));
          }
        }
      }
      const auto now = Clock::now();
      {
        std::lock_guard<std::mutex> lk(mu);
        for (auto &rp : batch) {
          rp->done = true;
          rp->t_done = now;
        }
        if (st != CG_OK && batch.size() > 1) ++n_isolated;
        n_replay += ctx.n_graph - g0;
        n_capture += ctx.n_capture - c0;
        n_repeat += ctx.n_repeat - r0;
        n_grow += (int64_t)(cg::t_grow_events - a0);
        rate_nb = std::max(rate_nb, ctx.rate_nb);
        for (int w = 0; w < 3; ++w) rate_cl[w] = std::max(rate_cl[w], ctx.rate_cl[w]);
        inflight -= bytes;
        ++n_batches;
        n_requests += (int64_t)batch.size();
      }
      cv_done.notify_all();
      cv_work.notify_all();
    }
    cg::ctx_free(ctx);
  }
};

}  // namespace

struct cg_server {
  Server s;
};

extern "C" {

cg_status cg_serve_create(const cg_model *m, int device, const double *x, int nworkers, size_t hbm_budget_bytes,
                          int max_batch_atoms, cg_server **out) {
  if (!m || !x || !out || nworkers < 1 || nworkers > 8 || max_batch_atoms < 1) {
    cg::set_error("cg_serve_create: bad argument");
    return CG_ERR_ARG;
  }
  *out = nullptr;
  auto *sv = new cg_server;
  Server &s = sv->s;
  s.m = m;
  s.device = device;
  s.x.assign(x, x + m->ncols);
  s.budget = hbm_budget_bytes ? hbm_budget_bytes : (size_t)4 << 30;
  s.max_atoms = max_batch_atoms;
  for (int i = 0; i < nworkers; ++i) s.workers.emplace_back([&s] { s.worker(); });
  {
    std::unique_lock<std::mutex> lk(s.mu);
    s.cv_done.wait(lk, [&] { return s.ready == nworkers; });
  }
  if (!s.init_error.empty()) {
    const std::string e = s.init_error;
    {
      std::lock_guard<std::mutex> lk(s.mu);
      s.stop = true;
    }
    s.cv_work.notify_all();
    for (auto &t : s.workers) t.join();
    delete sv;
    cg::set_error("cg_serve_create: " + e);
    return CG_ERR_DEVICE;
  }
  *out = sv;
  return CG_OK;
}

cg_status cg_serve_submit(cg_server *sv, const cg_frame *f, int64_t *ticket) {
  if (!sv || !f || !ticket || f->natoms <= 0 || !f->xyz || !f->type) {
    cg::set_error("cg_serve_submit: bad argument");
    return CG_ERR_ARG;
  }
  Server &s = sv->s;
  // validate here so that one bad request cannot fail the batch it joins
  for (int i = 0; i < f->natoms; ++i)
    if (f->type[i] < 0 || f->type[i] >= s.m->ntypes) {
      cg::set_error("cg_serve_submit: atom type out of range");
      return CG_ERR_ARG;
    }
  if (f->natoms >= (1 << 20)) {
    cg::set_error("cg_serve_submit: too many atoms");
    return CG_ERR_ARG;
  }
  auto r = std::make_shared<Request>();
  r->xyz.assign(f->xyz, f->xyz + 3 * (size_t)f->natoms);
  r->type.assign(f->type, f->type + f->natoms);
  r->f = *f;
  r->f.xyz = r->xyz.data();
  r->f.type = r->type.data();
  r->f.forces = nullptr;
  r->est_bytes = estimate_bytes(*s.m, r->f);
  r->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s.mu);
    if (s.stop) {
      cg::set_error("cg_serve_submit: server is stopping");
      return CG_ERR_ARG;
    }
    r->ticket = s.next_ticket++;
    s.live[r->ticket] = r;
    s.queue.push_back(r);
    *ticket = r->ticket;
  }
  s.cv_work.notify_one();
  return CG_OK;
}

cg_status cg_serve_poll(cg_server *sv, int64_t ticket, int wait, int *done, double *F, double *E, double *virial9,
                        double *latency_ms) {
  if (!sv || !done) {
    cg::set_error("cg_serve_poll: bad argument");
    return CG_ERR_ARG;
  }
  Server &s = sv->s;
  std::shared_ptr<Request> r;
  {
    std::unique_lock<std::mutex> lk(s.mu);
    auto it = s.live.find(ticket);
    if (it == s.live.end()) {
      cg::set_error("cg_serve_poll: unknown ticket");
      return CG_ERR_ARG;
    }
    r = it->second;
    if (wait) s.cv_done.wait(lk, [&] { return r->done; });
    *done = r->done ? 1 : 0;
    if (!r->done) return CG_OK;
    s.live.erase(it);
  }
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
  if (r->status != CG_OK) {
    cg::set_error(r->err);
    return r->status;
  }
  if (F) std::memcpy(F, r->F.data(), r->F.size() * sizeof(double));
  if (E) *E = r->E;
  if (virial9) std::memcpy(virial9, r->vir, sizeof(r->vir));
  return CG_OK;
}

cg_status cg_serve_hold(cg_server *sv, int hold) {
  if (!sv) {
    cg::set_error("cg_serve_hold: bad argument");
    return CG_ERR_ARG;
  }
  {
    std::lock_guard<std::mutex> lk(sv->s.mu);
    sv->s.hold = hold != 0;
  }
  sv->s.cv_work.notify_all();
  return CG_OK;
}

cg_status cg_serve_stats(cg_server *sv, int64_t *batches, int64_t *requests) {
  if (!sv) return CG_ERR_ARG;
  std::lock_guard<std::mutex> lk(sv->s.mu);
  if (batches) *batches = sv->s.n_batches;
  if (requests) *requests = sv->s.n_requests;
  return CG_OK;
}

cg_status cg_serve_counters(cg_server *sv, int64_t *out, int n) {
  if (!sv || !out || n < 0) {
    cg::set_error("cg_serve_counters: bad argument");
    return CG_ERR_ARG;
  }
  const Server &s = sv->s;
  std::lock_guard<std::mutex> lk(sv->s.mu);
  const int64_t v[7] = {s.n_batches, s.n_requests, s.n_isolated, s.n_replay, s.n_capture, s.n_repeat, s.n_grow};
  for (int i = 0; i < n && i < 7; ++i) out[i] = v[i];
  return CG_OK;
}

void cg_serve_destroy(cg_server *sv) {
  if (!sv) return;
  Server &s = sv->s;
  {
    std::lock_guard<std::mutex> lk(s.mu);
    s.stop = true;
  }
  s.cv_work.notify_all();
  for (auto &t : s.workers) t.join();
  delete sv;
}

}  // extern "C"
