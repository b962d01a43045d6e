// oa_serve.hip -- request loop for independent electronic minimisations on
// one GPU (include/oa_serve.h).
//
// One scheduler thread owns every admitted context.  Each pass of its loop:
//   1. admits queued requests in arrival order, skipping over those that do
//      not fit yet, while the slot count and the HBM budget allow (the charge
//      is the context's exact device footprint, oa_ctx_bytes);
//   2. enqueues the next CP step of every admitted request that has none in
//      flight, on that request's own stream, followed by an event;
//   3. retires finished steps (event complete): reads the step's energies and
//      |F|, and finishes the request at |F| <= f_tol (tol_cp_min,
//      CP_State_GSpacePlane.C:158-191), at max_steps, or on an error (a
//      step that could not be enqueued, or an event that reports anything
//      but completion or "not ready"),
//      releasing its memory so that queued requests can be admitted; a
//      mixed-precision request that reaches fp64_below or stalls continues
//      in an FP64 context.
// A finished request's context stays resident in a pool, keyed by
// everything oa_create consumed (cell, grid, cutoff, states, k-vectors,
// occupations, options): the next request of the same shape takes it over
// with oa_psi_set and its own potential, without device allocation or host
// layout work.  Pooled contexts count against the HBM budget and are
// released, least recently pooled first, when a new shape needs the room.
// Steps of different requests overlap on the device; no request waits for
// another one to finish.
//
// Every HIP call of the loop is made on the scheduler thread (contexts are
// created, loaded and released there too).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <mutex>
#include <thread>
#include <vector>

#include "oa_internal.h"
#include "oa_serve.h"
#include "oa_serve_poll.h"

namespace {

using clk = std::chrono::steady_clock;

double ms_since(clk::time_point a, clk::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

// What oa_create consumed for a request: two contexts with equal keys hold
// the same layout and options.
struct ShapeKey {
  std::vector<double> scalars;  // cell, cutoff, options (integers exactly as doubles)
  std::vector<int> ints;        // grid, states, points
  std::vector<int> kx, ky, kz;
  std::vector<double> occ;
  bool operator==(const ShapeKey& o) const {
    return ints == o.ints && scalars == o.scalars && kx == o.kx && ky == o.ky && kz == o.kz &&
           occ == o.occ;
  }
};

// A request with its own copies of every input.
struct Job {
  oa_request r{};
  std::vector<int> kx, ky, kz;
  std::vector<double> occ, psi0, vks;
  // self-consistent request: the ions' local pseudopotential (copied)
  bool scf = false;
  oa_local_pseudo lp{};
  std::vector<double> lp_pos, lp_vps, lp_vps_g0;
  std::vector<int> lp_species;
  size_t bytes = 0;
  clk::time_point t_submit;
  ShapeKey key;       // of the requested precision (set at submit)
  long long seq = 0;  // submission order
};

ShapeKey shape_key(const Job& j, int precision) {
  ShapeKey k;
  const oa_system& sys = j.r.sys;
  const oa_options& o = j.r.opt;
  k.ints = {sys.grid[0], sys.grid[1], sys.grid[2], sys.nstates, sys.npts, precision,
            o.ns_max_iter, o.use_lambda, o.use_cg, o.want_resid, o.fft_mode, o.use_graph,
            o.ns_mode, o.vks_from_rho, o.autotune, o.gemm_split, o.ns_fp32, o.fft_full};
  k.scalars.assign(sys.hmat, sys.hmat + 9);
  for (double v : {sys.ecut_ry, o.ns_tol, o.dt, o.tau_mass, o.ecut_mass}) k.scalars.push_back(v);
  k.kx = j.kx;
  k.ky = j.ky;
  k.kz = j.kz;
  k.occ = j.occ;
  return k;
}


// the key of the same request's context in another precision
ShapeKey with_precision(ShapeKey k, int precision) {
  k.ints[5] = precision;
  return k;
}

// A finished request's context, resident for the next request of its shape.
struct Pooled {
  oa_ctx* ctx = nullptr;
  size_t bytes = 0;
  ShapeKey key;
};

constexpr size_t kPoolMax = 64;  // pooled contexts at most

using oa_serve::classify_event;
using oa_serve::StepPoll;

struct Active {
  Job job;
  oa_ctx* ctx = nullptr;
  bool from_pool = false;  // ctx was taken over from the pool at admission
  hipEvent_t ev = nullptr;
  bool inflight = false;
  bool issue_failed = false;
  int steps = 0;
  clk::time_point t_admit;
  oa_step_out last{};
  // mixed precision: best |F| so far and when, and the FP64 continuation
  double best_f = 1e300;
  int best_step = 0;
  bool in_fp64 = false;
  int fp64_steps = 0;
};

constexpr int kStallSteps = 50;  // steps without a new best |F| that count as a stall

}  // namespace

struct oa_server {
  int device = 0;
  size_t budget = 0;
  int max_active = 1;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Job> queue;
  std::deque<oa_result> done;
  size_t used = 0, peak_used = 0;
  int n_active = 0, peak_active = 0;
  long long steps_run = 0, submitted = 0, finished = 0;
  bool stop = false;
  std::atomic<long long> cancel_upto{-1};  // withdraw requests with seq <= this
  long long seq = 0;                       // submission counter (under mu)
  std::thread th;
  // guarded by mu; its bytes are part of `used`
  std::list<Pooled> pool;
  long long pool_hits = 0;
};

namespace {

void finish(oa_server* s, Active* a, oa_status st) {
  oa_result res{};
  res.tag = a->job.r.tag;
  res.status = st;
  res.steps = a->steps;
  res.eelec = a->last.eelec;
  res.fnorm = a->last.fnorm;
  if (a->ctx && a->job.r.psi_out && (st == OA_OK || st == OA_ENOCONV))
    (void)oa_psi_get(a->ctx, a->job.r.psi_out);
  const auto now = clk::now();
  res.queue_ms = ms_since(a->job.t_submit, a->t_admit);
  res.run_ms = ms_since(a->t_admit, now);
  res.total_ms = ms_since(a->job.t_submit, now);
  res.hbm_bytes = a->job.bytes;
  res.fp64_steps = a->fp64_steps;
  res.reused = a->from_pool ? 1 : 0;
  if (a->ev) (void)hipEventDestroy(a->ev);
  // a context that ran without a device error goes back to the pool
  const bool keep = a->ctx && (st == OA_OK || st == OA_ENOCONV || st == OA_ECANCELED);
  Pooled p;
  std::vector<oa_ctx*> release;
  if (keep) {
    p.ctx = a->ctx;
    p.bytes = oa_ctx_device_bytes(a->ctx);
    p.key = a->in_fp64 ? with_precision(a->job.key, OA_FP64) : a->job.key;
  } else if (a->ctx) {
    release.push_back(a->ctx);
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->used -= a->job.bytes;
    if (keep) {
      s->used += p.bytes;
      s->pool.push_back(std::move(p));
      while (s->pool.size() > kPoolMax) {
        s->used -= s->pool.front().bytes;
        release.push_back(s->pool.front().ctx);
        s->pool.pop_front();
      }
    }
    s->n_active--;
    s->finished++;
    s->done.push_back(res);
    s->cv_done.notify_all();
  }
  for (oa_ctx* c : release) oa_destroy(c);
}

// Take a pooled context of this shape (caller holds s->mu), or nullptr.
oa_ctx* pool_take(oa_server* s, const ShapeKey& key) {
  for (auto it = s->pool.begin(); it != s->pool.end(); ++it) {
    if (it->key == key) {
      oa_ctx* c = it->ctx;
      s->used -= it->bytes;
      s->pool.erase(it);
      s->pool_hits++;
      return c;
    }
  }
  return nullptr;
}

oa_system job_system(Job& j) {
  oa_system sys = j.r.sys;
  sys.kx = j.kx.empty() ? nullptr : j.kx.data();
  sys.ky = j.ky.empty() ? nullptr : j.ky.data();
  sys.kz = j.kz.empty() ? nullptr : j.kz.data();
  sys.occ = j.occ.empty() ? nullptr : j.occ.data();
  return sys;
}

// V(r) of a new context: the fixed input, or the ions whose local
// pseudopotential enters V_KS[rho] at every step
oa_status load_potential(Active* a) {
  if (a->job.scf) {
    oa_local_pseudo lp = a->job.lp;
    lp.pos = a->job.lp_pos.data();
    lp.species = a->job.lp_species.data();
    lp.vps = a->job.lp_vps.data();
    lp.vps_g0 = a->job.lp_vps_g0.data();
    return oa_local_pseudo_set(a->ctx, &lp);
  }
  return oa_vks_set(a->ctx, a->job.vks.data());
}

// create and load one admitted request's context
oa_status start(oa_server* s, Active* a) {
  oa_system sys = job_system(a->job);
  oa_status st = a->ctx ? OA_OK : oa_create(s->device, &sys, &a->job.r.opt, &a->ctx);
  if (st == OA_OK) st = oa_psi_set(a->ctx, a->job.psi0.data());
  if (st == OA_OK) st = load_potential(a);
  if (st == OA_OK && hipEventCreateWithFlags(&a->ev, hipEventDisableTiming) != hipSuccess)
    st = OA_EHIP;
  return st;
}

// Continue a mixed-precision request in FP64: its coefficients go through
// the host copy of psi0 into an FP64 context of its shape (pooled, else
// new).  The mixed context goes to the pool for the next request of its
// shape; the request's reservation (the larger of the two precisions) covers
// the FP64 one, and pooled contexts of other shapes are released, least
// recently pooled first, while the pool holds more than the budget allows.
oa_status to_fp64(oa_server* s, Active* a) {
  oa_status st = oa_psi_get(a->ctx, a->job.psi0.data());
  if (st != OA_OK) return st;
  std::vector<oa_ctx*> release;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    // a pooled FP64 context of this shape: the request's reservation already
    // covers one (admission charged the larger of the two precisions)
    oa_ctx* fp64 = pool_take(s, with_precision(a->job.key, OA_FP64));
    Pooled p;
    p.ctx = a->ctx;
    p.bytes = oa_ctx_device_bytes(a->ctx);
    p.key = a->job.key;
    s->used += p.bytes;
    s->pool.push_back(std::move(p));
    while ((s->used > s->budget && !s->pool.empty()) || s->pool.size() > kPoolMax) {
      s->used -= s->pool.front().bytes;
      release.push_back(s->pool.front().ctx);
      s->pool.pop_front();
    }
    a->ctx = fp64;
    a->job.r.opt.precision = OA_FP64;
  }
  for (oa_ctx* c : release) oa_destroy(c);
  oa_system sys = job_system(a->job);
  st = a->ctx ? OA_OK : oa_create(s->device, &sys, &a->job.r.opt, &a->ctx);
  if (st == OA_OK) st = oa_psi_set(a->ctx, a->job.psi0.data());
  if (st == OA_OK) st = load_potential(a);
  a->in_fp64 = true;
  return st;
}

void scheduler(oa_server* s) {
  (void)hipSetDevice(s->device);
  std::list<Active*> active;
  std::vector<oa_ctx*> release;  // pooled contexts given up for room
  for (;;) {
    // 1. admission (arrival order, skipping requests that do not fit yet)
    std::vector<Active*> admitted;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      if (active.empty() && s->queue.empty()) {
        s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty(); });
        if (s->stop && s->queue.empty()) break;
      }
      for (auto it = s->queue.begin(); it != s->queue.end() && s->n_active < s->max_active;) {
        // a pooled context of the same shape, else room made by releasing
        // pooled contexts of other shapes, least recently pooled first
        oa_ctx* reuse = pool_take(s, it->key);
        while (!reuse && s->used + it->bytes > s->budget && !s->pool.empty()) {
          s->used -= s->pool.front().bytes;
          release.push_back(s->pool.front().ctx);
          s->pool.pop_front();
        }
        if (reuse || s->used + it->bytes <= s->budget) {
          Active* a = new Active();
          a->job = std::move(*it);
          a->ctx = reuse;
          a->from_pool = reuse != nullptr;
          it = s->queue.erase(it);
          s->used += a->job.bytes;
          s->n_active++;
          s->peak_used = std::max(s->peak_used, s->used);
          s->peak_active = std::max(s->peak_active, s->n_active);
          admitted.push_back(a);
        } else {
          ++it;
        }
      }
    }
    for (oa_ctx* c : release) oa_destroy(c);
    release.clear();
    for (Active* a : admitted) {
      a->t_admit = clk::now();
      const oa_status st = start(s, a);
      if (st == OA_OK) {
        active.push_back(a);
      } else {
        finish(s, a, st);
        delete a;
      }
    }
    bool progressed = !admitted.empty();
    // 2. issue the next step of every idle request
    for (Active* a : active) {
      if (a->inflight || a->job.seq <= s->cancel_upto.load(std::memory_order_relaxed)) continue;
      const oa_status st = oa_step(a->ctx, nullptr);
      if (st != OA_OK || hipEventRecord(a->ev, oa_ctx_stream(a->ctx)) != hipSuccess) {
        a->issue_failed = true;  // retired below
        continue;
      }
      a->inflight = true;
    }
    // 3. retire finished steps
    for (auto it = active.begin(); it != active.end();) {
      Active* a = *it;
      oa_status st = OA_OK;
      bool end = false;
      const bool cancelled = a->job.seq <= s->cancel_upto.load(std::memory_order_relaxed);
      const StepPoll poll = a->issue_failed ? StepPoll::Failed
                            : a->inflight   ? classify_event(hipEventQuery(a->ev))
                                            : StepPoll::Running;
      if (poll == StepPoll::Failed) {  // not enqueued, or the device reported an error
        (void)hipGetLastError();
        a->inflight = false;
        st = OA_EHIP;
        end = true;
      } else if (poll == StepPoll::Done) {
        a->inflight = false;
        a->steps++;
        progressed = true;
        {
          std::lock_guard<std::mutex> lk(s->mu);
          s->steps_run++;
        }
        st = oa_step_result(a->ctx, &a->last);
        if (a->in_fp64) a->fp64_steps++;
        if (st != OA_OK) {
          end = true;
        } else if (a->last.fnorm <= a->job.r.f_tol) {
          end = true;
        } else if (a->steps >= a->job.r.max_steps) {
          st = OA_ENOCONV;
          end = true;
        } else if (!a->in_fp64 && a->job.r.opt.precision == OA_MIXED &&
                   a->job.r.fp64_below > 0.0) {
          if (a->last.fnorm < a->best_f) {
            a->best_f = a->last.fnorm;
            a->best_step = a->steps;
          }
          if (a->last.fnorm <= a->job.r.fp64_below || a->steps - a->best_step >= kStallSteps) {
            st = to_fp64(s, a);
            if (st != OA_OK) end = true;
          }
        }
      }
      if (!end && cancelled && !a->inflight) {
        st = OA_ECANCELED;  // withdrawn: ends once no step is in flight
        end = true;
      }
      if (end) {
        finish(s, a, st);
        delete a;
        it = active.erase(it);
        progressed = true;
      } else {
        ++it;
      }
    }
    if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
  std::list<Pooled> pool;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    pool.swap(s->pool);
    for (const Pooled& p : pool) s->used -= p.bytes;
  }
  for (Pooled& p : pool) oa_destroy(p.ctx);
}

}  // namespace

extern "C" {

oa_status oa_server_create(int device, size_t hbm_budget_bytes, int max_active,
                           oa_server** out) {
  if (!out || max_active < 1 || hbm_budget_bytes == 0) return OA_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return OA_EHIP;
  oa_server* s = new oa_server();
  s->device = device;
  s->budget = hbm_budget_bytes;
  s->max_active = max_active;
  s->th = std::thread(scheduler, s);
  *out = s;
  return OA_OK;
}

oa_status oa_server_submit(oa_server* s, const oa_request* req) {
  if (!s || !req || !req->psi0 || (!req->vks && !req->pseudo) || req->max_steps < 1)
    return OA_EINVAL;
  Job j;
  j.r = *req;
  const oa_system& sys = req->sys;
  if (!req->vks) {
    const oa_local_pseudo& p = *req->pseudo;
    if (p.natoms < 0 || p.nspecies < 1 || p.nspline < 1 || (p.natoms > 0 && (!p.pos || !p.species)) ||
        !p.vps || !p.vps_g0)
      return OA_EINVAL;
    j.scf = true;
    j.r.opt.vks_from_rho = 1;
    j.lp = p;
    j.lp_pos.assign(p.pos, p.pos + 3 * (size_t)p.natoms);
    j.lp_species.assign(p.species, p.species + p.natoms);
    j.lp_vps.assign(p.vps, p.vps + 4 * (size_t)p.nspecies * p.nspline);
    j.lp_vps_g0.assign(p.vps_g0, p.vps_g0 + p.nspecies);
  } else {
    j.r.opt.vks_from_rho = 0;
  }
  j.r.pseudo = nullptr;
  (void)hipSetDevice(s->device);
  j.bytes = oa_ctx_bytes(&sys, &j.r.opt);
  if (req->opt.precision == OA_MIXED && req->fp64_below > 0.0) {
    oa_options o64 = j.r.opt;
    o64.precision = OA_FP64;
    j.bytes = std::max(j.bytes, oa_ctx_bytes(&sys, &o64));
  }
  if (j.bytes == 0 || j.bytes > s->budget) return OA_EINVAL;
  int npts = sys.npts;
  if (npts > 0) {
    if (!sys.kx || !sys.ky || !sys.kz) return OA_EINVAL;
    j.kx.assign(sys.kx, sys.kx + npts);
    j.ky.assign(sys.ky, sys.ky + npts);
    j.kz.assign(sys.kz, sys.kz + npts);
  } else {
    npts = oa_kvecs(sys.ecut_ry, sys.hmat, nullptr, nullptr, nullptr);
  }
  if (sys.occ) j.occ.assign(sys.occ, sys.occ + sys.nstates);
  j.psi0.assign(req->psi0, req->psi0 + (size_t)2 * sys.nstates * npts);
  if (!j.scf) {
    const size_t N = (size_t)sys.grid[0] * sys.grid[1] * sys.grid[2];
    j.vks.assign(req->vks, req->vks + N);
  }
  j.key = shape_key(j, j.r.opt.precision);
  j.t_submit = clk::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    j.seq = s->seq++;
    s->queue.push_back(std::move(j));
    s->submitted++;
  }
  s->cv_work.notify_one();
  return OA_OK;
}

int oa_server_poll(oa_server* s, oa_result* res, int max) {
  if (!s || !res || max < 1) return 0;
  std::lock_guard<std::mutex> lk(s->mu);
  int n = 0;
  while (n < max && !s->done.empty()) {
    res[n++] = s->done.front();
    s->done.pop_front();
  }
  return n;
}

oa_status oa_server_drain(oa_server* s) {
  if (!s) return OA_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->finished == s->submitted; });
  return OA_OK;
}

oa_status oa_server_cancel(oa_server* s) {
  if (!s) return OA_EINVAL;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->cancel_upto.store(s->seq - 1, std::memory_order_relaxed);
    const auto now = clk::now();
    for (const Job& j : s->queue) {
      oa_result r{};
      r.tag = j.r.tag;
      r.status = OA_ECANCELED;
      r.queue_ms = r.total_ms = ms_since(j.t_submit, now);
      s->done.push_back(r);
      s->finished++;
    }
    s->queue.clear();
  }
  s->cv_done.notify_all();
  s->cv_work.notify_all();
  return OA_OK;
}

oa_status oa_server_stats(oa_server* s, double st[8]) {
  if (!s || !st) return OA_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  st[0] = s->n_active;
  st[1] = (double)s->queue.size();
  st[2] = (double)s->finished;
  st[3] = (double)s->used;
  st[4] = s->peak_active;
  st[5] = (double)s->peak_used;
  st[6] = (double)s->steps_run;
  st[7] = (double)s->submitted;
  return OA_OK;
}

oa_status oa_server_destroy(oa_server* s) {
  if (!s) return OA_EINVAL;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  (void)oa_server_drain(s);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->th.joinable()) s->th.join();
  delete s;
  return OA_OK;
}

}  // extern "C"
