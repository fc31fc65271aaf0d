// The serving loop: continuous batching of FMM solves of uneven sizes and methods on
// one device.  No reference counterpart (the reference's nearest pieces are the plan
// and operator reuse of src/device_passes.jl:136-146 and src/operators/cache.jl:5-46
// and build_expansions! on a fixed geometry, src/fmm.jl:264-298).
//
// * Slots: each worker thread owns a context (its own HIP stream and rocBLAS handle),
//   so the solves admitted at one time run concurrently on the device.
// * Memory: each slot keeps the blocks its solves free for the next one (no hipMalloc /
//   hipFree between requests of similar sizes), up to its share of the device memory the
//   budget leaves free; an allocation that fails returns the idle blocks of every
//   context and tries again.
// * Admission: a free slot takes the first queued request whose device-memory
//   estimate fits the budget next to what running solves and resident plans hold
//   (evicting resident plans, least recently used first, to make room); a request
//   that has waited longer than the aging limit blocks the requests behind it from
//   overtaking, so large solves are not starved.
// * Residency: a request with a geometry key keeps its plan (tree, lists, buffers) on
//   the device; a later request with the same cache and key skips the tree build and
//   runs the passes only (build_expansions! on a fixed geometry).
// * Latency: from submission to the completion of the last output copy.
#include "ffmm_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

struct ffmm_server_s;

namespace ffmm {

static double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

template <class T>
static size_t cap_bytes(const DBuf<T> &b) {
  return b.cap * sizeof(T);
}

size_t plan_bytes(const Plan &p) {
  size_t s = 0;
  s += cap_bytes(p.lam_d) + cap_bytes(p.node_leafslot) + cap_bytes(p.hscale) + cap_bytes(p.khat);
  s += cap_bytes(p.cell) + cap_bytes(p.level) + cap_bytes(p.parent) + cap_bytes(p.child_first) + cap_bytes(p.plo) +
       cap_bytes(p.phi) + cap_bytes(p.col) + cap_bytes(p.perm) + cap_bytes(p.point_node);
  s += cap_bytes(p.leaf) + cap_bytes(p.ghost) + cap_bytes(p.xlist) + cap_bytes(p.has_source);
  s += cap_bytes(p.sx) + cap_bytes(p.sy) + cap_bytes(p.sz) + cap_bytes(p.centers) + cap_bytes(p.sx32) +
       cap_bytes(p.sy32) + cap_bytes(p.sz32);
  s += cap_bytes(p.s2m_nodes) + cap_bytes(p.leaf_nodes);
  for (const auto &li : p.lists)
    s += cap_bytes(li.ptr) + cap_bytes(li.src) + cap_bytes(li.far) + cap_bytes(li.csrc) + cap_bytes(li.cfar) +
         cap_bytes(li.ctgt) +
         cap_bytes(li.cstart);
  s += cap_bytes(p.Mp) + cap_bytes(p.Pl) + cap_bytes(p.Lc) + cap_bytes(p.Lc32) + cap_bytes(p.tau) + cap_bytes(p.tau32);
  s += cap_bytes(p.pM) + cap_bytes(p.tup) + cap_bytes(p.Mup) + cap_bytes(p.tM) + cap_bytes(p.G) + cap_bytes(p.tdn) +
       cap_bytes(p.Gc) + cap_bytes(p.Yc) + cap_bytes(p.Gc32) + cap_bytes(p.Yc32) + cap_bytes(p.pM32) + cap_bytes(p.Mp32) + cap_bytes(p.tM32);
  s += cap_bytes(p.spec_S) + cap_bytes(p.spec_U) + cap_bytes(p.spec_S32) + cap_bytes(p.spec_U32);
  s += cap_bytes(p.Zc) + cap_bytes(p.Zc32) + cap_bytes(p.gin) + cap_bytes(p.gout);
  for (const auto &b : p.spec_gptr) s += cap_bytes(b);
  for (const auto &b : p.spec_cand) s += cap_bytes(b);
  for (const auto &b : p.spec_gcol) s += cap_bytes(b);
  for (const auto &b : p.spec_tiles) s += cap_bytes(b);
  for (const auto *v : {&p.spec_ulist, &p.spec_fit, &p.spec_over})
    for (const auto &b : *v) s += cap_bytes(b);
  for (const auto &b : p.spec_uslot) s += cap_bytes(b);
  s += cap_bytes(p.tgt_sorted) + cap_bytes(p.tgt_node) + cap_bytes(p.tgt_leafslot) + cap_bytes(p.tgt_order) +
       cap_bytes(p.tgt_ptr) + cap_bytes(p.tgt_keys);
  return s;
}

// a conservative estimate of a plan's device memory before its tree exists: per node
// (about 12 N / ppl of them) the multipole and partial-local columns and, spectral,
// the single-precision spectra; per leaf (about 5 N / ppl) the local column and, cube
// method, the check potentials and pseudo-inverse intermediates of all leaves at once;
// per point the sorted points, densities, outputs and the sort's buffers
static int64_t estimate_bytes(const OpCache &o, int64_t n, int ppl, int m2l, int64_t nt) {
  const double nodes = 12.0 * (double)n / ppl + 64, leaves = 5.0 * (double)n / ppl + 64;
  double b = nodes * (2.0 * o.Ns * 8) + leaves * (o.Nl * 8.0) + 256.0 * n + 96.0 * nt;
  if (o.method == METHOD_CUBE) b += leaves * (o.Nl + 3.0 * o.Ns) * 8;
  if (m2l == FFMM_M2L_SPECTRAL) b += 0.6 * nodes * 2304 * 16;
  // dense class M2L: the class-major pair table, and the gathered columns and products of
  // one target chunk (at most about 2 GiB)
  if (m2l == FFMM_M2L_DENSE) b += 4.0 * o.nclass * nodes + std::min<double>(2.2e9, 2.0 * o.Ns * 8 * 200.0 * nodes);
  return (int64_t)b + (64ll << 20);
}

struct Request {
  int64_t ticket = 0;
  ffmm_request r{};
  OpCache *cache = nullptr;
  int64_t est = 0;
  double t_submit = 0, t_done = 0;
  bool done = false;
  // out of device memory while other solves ran: queued again, admitted only when it
  // would run alone (then its own footprint is all the device holds)
  bool alone = false;
  int status = FFMM_OK;
  std::string error;
};

struct Resident {
  OpCache *cache = nullptr;
  int64_t geometry = -1;
  std::unique_ptr<Plan> plan;
  size_t bytes = 0;
  bool busy = false;
  uint64_t last_use = 0;
};

struct Worker {
  Context ctx;
  DBuf<double> xyz, dens, out, tgt, tout;
  std::thread th;
};

}  // namespace ffmm

struct ffmm_server_s {
  int device = 0;
  int64_t budget = 0;
  double aging_ms = 2000.0;
  std::vector<std::unique_ptr<ffmm::Worker>> workers;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<ffmm::Request *> queue;
  std::map<int64_t, std::unique_ptr<ffmm::Request>> reqs;
  std::vector<std::unique_ptr<ffmm::Resident>> resident;
  std::vector<std::pair<ffmm::OpCache *, int64_t>> building;   // geometries whose plan a slot is building
  int64_t in_use = 0, peak = 0;
  int64_t next_ticket = 1;
  uint64_t clock = 0;
  int64_t completed = 0, built = 0, reused = 0, evictions = 0;
  int64_t worst_ratio = 0;   // largest (plan bytes after the solve) / (admission estimate), in thousandths
  int64_t retried = 0;       // requests queued again after running out of device memory
  int running = 0;
  bool stop = false;
  // plan data in and out of the resident set: a context of its own (created at the first
  // use), so that no worker's stream or BLAS handle is shared
  std::mutex io_mu;
  std::unique_ptr<ffmm::Context> io;
};

namespace ffmm {

static int ncomp_of(const OpCache &o, int output) {
  if (o.method == METHOD_HARMONIC && o.kernel == FFMM_KERNEL_STOKESLET) return output == FFMM_PRESSURE ? 1 : 3;
  return output == FFMM_POTENTIAL ? 1 : output == FFMM_GRADIENT ? 3 : 4;
}

static Resident *find_resident(ffmm_server_s &S, OpCache *c, int64_t g) {
  if (g < 0) return nullptr;
  for (auto &r : S.resident)
    if (r->cache == c && r->geometry == g) return r.get();
  return nullptr;
}

// under the lock: drop the least recently used idle resident plan; false if none
static bool evict_one(ffmm_server_s &S, std::vector<std::unique_ptr<Resident>> &graveyard) {
  int best = -1;
  for (int i = 0; i < (int)S.resident.size(); i++)
    if (!S.resident[i]->busy && (best < 0 || S.resident[i]->last_use < S.resident[best]->last_use)) best = i;
  if (best < 0) return false;
  S.in_use -= (int64_t)S.resident[best]->bytes;
  graveyard.push_back(std::move(S.resident[best]));
  S.resident.erase(S.resident.begin() + best);
  S.evictions++;
  return true;
}

// the idle blocks the slots' contexts keep for reuse (device memory the server holds)
static int64_t idle_bytes(ffmm_server_s &S) {
  size_t b = 0;
  for (auto &w : S.workers) b += context_idle_bytes(w->ctx);
  return (int64_t)b;
}

// under the lock: the next admissible request (or null), with its resident plan if any.
// The budget covers the estimates of the running solves and the resident plans (their
// trees and lists; the work buffers go back to the slots); to make room the resident
// plans are evicted, least recently used first.  The slots' idle cached blocks are not
// counted: an allocation that fails returns every context's idle blocks and tries again.
static Request *pick(ffmm_server_s &S, Resident *&res, std::vector<std::unique_ptr<Resident>> &graveyard) {
  const double t = now_ms();
  for (auto it = S.queue.begin(); it != S.queue.end(); ++it) {
    Request *q = *it;
    if (q->alone && S.running > 0) return nullptr;   // it goes next, alone
    Resident *r = find_resident(S, q->cache, q->r.geometry);
    if (r) {
      if (r->busy) continue;             // another slot holds this geometry's plan
      // its work buffers are allocated again: the request's estimate is reserved
      r->busy = true;                    // (not evicted to make room for itself)
      while (S.in_use + q->est > S.budget && evict_one(S, graveyard)) {
      }
      if (S.in_use + q->est <= S.budget || S.running == 0) {
        S.in_use += q->est;
        S.peak = std::max(S.peak, S.in_use);
        res = r;
        S.queue.erase(it);
        return q;
      }
      r->busy = false;
      if (t - q->t_submit > S.aging_ms) return nullptr;
      continue;
    }
    if (q->r.geometry >= 0 &&
        std::find(S.building.begin(), S.building.end(), std::make_pair(q->cache, q->r.geometry)) != S.building.end())
      continue;                          // its plan is being built: wait for it to become resident
    while (S.in_use + q->est > S.budget && evict_one(S, graveyard)) {
    }
    if (S.in_use + q->est <= S.budget || S.running == 0) {
      S.in_use += q->est;
      S.peak = std::max(S.peak, S.in_use);
      res = nullptr;
      if (q->r.geometry >= 0) S.building.emplace_back(q->cache, q->r.geometry);
      S.queue.erase(it);
      return q;
    }
    if (t - q->t_submit > S.aging_ms) return nullptr;   // let no one overtake a starving request
  }
  return nullptr;
}

static void run_request(Worker &w, Request &q, Plan *plan, bool fresh) {
  Context &ctx = w.ctx;
  const ffmm_request &r = q.r;
  OpCache &o = *q.cache;
  hipStream_t s = ctx.stream;
  FFMM_HIP(hipSetDevice(ctx.device));
  AllocScope scope(ctx);
  if (fresh) {
    w.xyz.resize(3 * (size_t)r.n);
    FFMM_HIP(hipMemcpyAsync(w.xyz.p, r.points, sizeof(double) * 3 * r.n, hipMemcpyHostToDevice, s));
    plan_init(*plan, ctx, o, w.xyz.p, r.n, r.ppl, r.origin, r.width, r.e0, r.max_levels > 0 ? r.max_levels : 30,
              r.precision, r.m2l);
  }
  // a resident plan runs on the stream and BLAS handle of the slot that holds it now
  plan->ctx = &ctx;
  if (plan->N != r.n) fail(FFMM_ERR_DIMENSION, "the resident geometry has " + std::to_string(plan->N) + " points");
  const int nd = r.nd > 0 ? r.nd : 1;
  w.dens.resize((size_t)nd * r.n);
  FFMM_HIP(hipMemcpyAsync(w.dens.p, r.dens, sizeof(double) * nd * r.n, hipMemcpyHostToDevice, s));
  plan_build(*plan, w.dens.p, nd, 0);
  const int nc = ncomp_of(o, r.output);
  if (r.out_sources) {
    w.out.resize((size_t)nc * r.n);
    plan_eval_sources(*plan, w.out.p, r.output);
    FFMM_HIP(hipMemcpyAsync(r.out_sources, w.out.p, sizeof(double) * nc * r.n, hipMemcpyDeviceToHost, s));
  }
  if (r.targets && r.nt > 0) {
    w.tgt.resize(3 * (size_t)r.nt);
    w.tout.resize((size_t)nc * r.nt);
    FFMM_HIP(hipMemcpyAsync(w.tgt.p, r.targets, sizeof(double) * 3 * r.nt, hipMemcpyHostToDevice, s));
    plan_eval_targets(*plan, w.tgt.p, r.nt, w.tout.p, r.output);
    FFMM_HIP(hipMemcpyAsync(r.out_targets, w.tout.p, sizeof(double) * nc * r.nt, hipMemcpyDeviceToHost, s));
  }
  ctx.sync();
}

static void worker_loop(ffmm_server_s *S, Worker *w) {
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    Request *q = nullptr;
    Resident *res = nullptr;
    std::vector<std::unique_ptr<Resident>> graveyard;
    S->cv_work.wait_for(lk, std::chrono::milliseconds(50), [&] {
      if (S->stop && S->queue.empty()) return true;
      q = pick(*S, res, graveyard);
      return q != nullptr;
    });
    if (!q) {
      if (S->stop && S->queue.empty()) break;
      if (!graveyard.empty()) {
        lk.unlock();
        for (auto &g : graveyard) plan_release(*g->plan);
        graveyard.clear();
        lk.lock();
      }
      continue;
    }
    S->running++;
    lk.unlock();
    for (auto &g : graveyard) plan_release(*g->plan);
    graveyard.clear();
    std::unique_ptr<Plan> own;
    Plan *plan = res ? res->plan.get() : nullptr;
    if (!plan) {
      own = std::make_unique<Plan>();
      plan = own.get();
    }
    int status = FFMM_OK;
    std::string err;
    try {
      run_request(*w, *q, plan, res == nullptr);
    } catch (const Error &e) {
      status = e.code;
      err = e.what();
    } catch (const std::exception &e) {
      status = FFMM_ERR_STATE;
      err = e.what();
    }
    size_t real = own ? plan_bytes(*own) : 0;
    const int64_t ratio = own && q->est > 0 ? (int64_t)(1000.0 * (double)real / (double)q->est) : 0;
    bool keep = own && status == FFMM_OK && q->r.geometry >= 0;
    // a kept plan holds its tree and lists only; its work buffers go back to this slot
    try {
      if (keep) {
        plan_shrink(*own);
        real = plan_bytes(*own);
      } else if (res) {
        plan_shrink(*res->plan);
      }
    } catch (const std::exception &) {
      keep = false;
    }
    if (own && !keep) {
      plan_release(*own);
      own.reset();
    }
    lk.lock();
    S->running--;
    S->worst_ratio = std::max(S->worst_ratio, ratio);
    // out of memory next to other solves: give the reservation back and queue the
    // request again at the front, to run alone
    const bool retry = status == FFMM_ERR_MEMORY && !q->alone && S->running > 0;
    if (retry) {
      S->in_use -= q->est;
      if (res) {
        res->busy = false;
      } else {
        auto bi = std::find(S->building.begin(), S->building.end(), std::make_pair(q->cache, q->r.geometry));
        if (bi != S->building.end()) S->building.erase(bi);
      }
      q->alone = true;
      S->retried++;
      S->queue.push_front(q);
      S->cv_work.notify_all();
      if (own) {
        lk.unlock();
        plan_release(*own);
        own.reset();
        lk.lock();
      }
      continue;
    }
    S->in_use -= q->est;
    if (res) {
      res->busy = false;
      res->last_use = ++S->clock;
      S->reused++;
    } else {
      if (status == FFMM_OK) S->built++;
      auto bi = std::find(S->building.begin(), S->building.end(), std::make_pair(q->cache, q->r.geometry));
      if (bi != S->building.end()) S->building.erase(bi);
      // two slots may have built the same geometry at once: keep the first
      if (keep && find_resident(*S, q->cache, q->r.geometry)) keep = false;
      if (keep) {
        auto rr = std::make_unique<Resident>();
        rr->cache = q->cache;
        rr->geometry = q->r.geometry;
        rr->bytes = real;
        rr->last_use = ++S->clock;
        rr->plan = std::move(own);
        S->in_use += (int64_t)real;
        S->peak = std::max(S->peak, S->in_use);
        S->resident.push_back(std::move(rr));
      }
    }
    q->status = status;
    q->error = err;
    q->t_done = now_ms();
    q->done = true;
    S->completed++;
    S->cv_done.notify_all();
    S->cv_work.notify_all();
    if (own && !keep) {
      lk.unlock();
      plan_release(*own);
      own.reset();
      lk.lock();
    }
  }
}

}  // namespace ffmm

using namespace ffmm;

extern "C" {

int ffmm_server_create(int device, int nslots, int64_t hbm_budget, ffmm_server *out) {
  return guard([&] {
    if (!out) fail(FFMM_ERR_ARGUMENT, "null output handle");
    *out = nullptr;
    if (nslots < 1 || nslots > 16) fail(FFMM_ERR_ARGUMENT, "nslots must lie in 1:16");
    if (hbm_budget <= 0) fail(FFMM_ERR_ARGUMENT, "the memory budget must be positive");
    auto S = std::make_unique<ffmm_server_s>();
    S->device = device;
    S->budget = hbm_budget;
    for (int i = 0; i < nslots; i++) {
      auto w = std::make_unique<Worker>();
      context_init(w->ctx, device, nullptr);
      S->workers.push_back(std::move(w));
    }
    // the slots' idle blocks live in the device memory the budget leaves free: each slot
    // keeps at most its share of it (at least 256 MiB), so that the budget plus the idle
    // blocks fit the device
    size_t fr = 0, tot = 0;
    FFMM_HIP(hipMemGetInfo(&fr, &tot));
    const size_t outside = fr > (size_t)hbm_budget ? fr - (size_t)hbm_budget : 0;
    const size_t share = std::max<size_t>((size_t)256 << 20, outside / (size_t)nslots);
    for (auto &w : S->workers) w->ctx.blocks->limit = share;
    for (auto &w : S->workers) w->th = std::thread(worker_loop, S.get(), w.get());
    *out = S.release();
  });
}

int ffmm_server_submit(ffmm_server S, const ffmm_request *req, int64_t *ticket) {
  return guard([&] {
    if (!S || !req || !ticket) fail(FFMM_ERR_ARGUMENT, "null argument");
    if (!req->cache) fail(FFMM_ERR_ARGUMENT, "null cache");
    OpCache *c = &req->cache->c;
    if (!c->ready) fail(FFMM_ERR_STATE, "the operator cache holds no operators");
    if (c->ctx->device != S->device) fail(FFMM_ERR_ARGUMENT, "the cache lives on another device");
    if (req->n <= 0 || !req->points || !req->dens) fail(FFMM_ERR_ARGUMENT, "a request needs points and densities");
    if (req->nt > 0 && (!req->targets || !req->out_targets)) fail(FFMM_ERR_ARGUMENT, "null targets or target output");
    if (!req->out_sources && req->nt <= 0) fail(FFMM_ERR_ARGUMENT, "a request needs an output");
    if (req->output < 0 || req->output > 3) fail(FFMM_ERR_ARGUMENT, "unknown output");
    if (req->output == FFMM_PRESSURE && !(c->method == METHOD_HARMONIC && c->kernel == FFMM_KERNEL_STOKESLET))
      fail(FFMM_ERR_ARGUMENT, "Pressure is an output of the Stokes kernels");
    auto q = std::make_unique<Request>();
    q->r = *req;
    q->cache = c;
    q->est = estimate_bytes(*c, req->n, std::max(1, req->ppl), req->m2l, std::max<int64_t>(0, req->nt));
    q->t_submit = now_ms();
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) fail(FFMM_ERR_STATE, "the server is shutting down");
    q->ticket = S->next_ticket++;
    *ticket = q->ticket;
    S->queue.push_back(q.get());
    S->reqs[q->ticket] = std::move(q);
    S->cv_work.notify_one();
  });
}

int ffmm_server_wait(ffmm_server S, int64_t ticket, double *latency_ms) {
  int status = FFMM_OK;
  std::string err;
  int rc = guard([&] {
    if (!S) fail(FFMM_ERR_ARGUMENT, "null server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) fail(FFMM_ERR_ARGUMENT, "unknown ticket");
    Request *q = it->second.get();
    S->cv_done.wait(lk, [&] { return q->done; });
    if (latency_ms) *latency_ms = q->t_done - q->t_submit;
    status = q->status;
    err = q->error;
    S->reqs.erase(it);
  });
  if (rc != FFMM_OK) return rc;
  if (status != FFMM_OK) g_last_error = err;
  return status;
}

int ffmm_server_stats(ffmm_server S, int64_t *stats, int n) {
  return guard([&] {
    if (!S || !stats) fail(FFMM_ERR_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    const int64_t v[12] = {S->completed, S->built,     S->reused,     (int64_t)S->resident.size(),
                           S->in_use,    S->peak,      S->evictions,  (int64_t)S->queue.size(),
                           idle_bytes(*S), S->worst_ratio, S->retried, allocation_flushes()};
    for (int i = 0; i < n && i < 12; i++) stats[i] = v[i];
  });
}

static ffmm::Context &server_io(ffmm_server_s &S) {
  if (!S.io) {
    auto c = std::make_unique<ffmm::Context>();
    ffmm::context_init(*c, S.device, nullptr);
    S.io = std::move(c);
  }
  return *S.io;
}

int ffmm_server_export(ffmm_server S, ffmm_opcache cache, int64_t geometry, void *buf, int64_t cap, int64_t *size) {
  return guard([&] {
    if (!S || !cache || !size) fail(FFMM_ERR_ARGUMENT, "null argument");
    if (cap < 0) fail(FFMM_ERR_ARGUMENT, "negative capacity");
    std::lock_guard<std::mutex> io_lk(S->io_mu);
    Resident *r = nullptr;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      r = find_resident(*S, &cache->c, geometry);
      if (!r) fail(FFMM_ERR_ARGUMENT, "no resident plan of this cache and geometry");
      if (r->busy) fail(FFMM_ERR_STATE, "the resident plan is in use");
      r->busy = true;   // no slot takes it meanwhile, nor is it evicted
    }
    struct Release {
      ffmm_server_s *S;
      Resident *r;
      ~Release() {
        {
          std::lock_guard<std::mutex> lk(S->mu);
          r->busy = false;
        }
        S->cv_work.notify_all();
      }
    } rel{S, r};
    Context &io = server_io(*S);
    r->plan->ctx = &io;
    AllocScope scope(io);
    const size_t need = plan_save(*r->plan, nullptr, 0);
    *size = (int64_t)need;
    if (buf) {
      if ((size_t)cap < need) fail(FFMM_ERR_ARGUMENT, "the buffer is smaller than the plan data");
      plan_save(*r->plan, buf, (size_t)cap);
    }
  });
}

int ffmm_server_import(ffmm_server S, ffmm_opcache cache, int64_t geometry, const void *buf, int64_t size) {
  return guard([&] {
    if (!S || !cache || !buf) fail(FFMM_ERR_ARGUMENT, "null argument");
    if (geometry < 0) fail(FFMM_ERR_ARGUMENT, "a resident plan needs a geometry key >= 0");
    if (size < 0) fail(FFMM_ERR_ARGUMENT, "negative size");
    OpCache *c = &cache->c;
    if (c->ctx->device != S->device) fail(FFMM_ERR_ARGUMENT, "the cache lives on another device");
    {
      std::lock_guard<std::mutex> lk(S->mu);
      if (find_resident(*S, c, geometry)) fail(FFMM_ERR_ARGUMENT, "a plan of this cache and geometry is resident");
    }
    std::lock_guard<std::mutex> io_lk(S->io_mu);
    auto rr = std::make_unique<Resident>();
    rr->cache = c;
    rr->geometry = geometry;
    rr->plan = std::make_unique<Plan>();
    Context &io = server_io(*S);
    try {
      plan_restore(*rr->plan, io, *c, buf, (size_t)size);
    } catch (...) {
      plan_release(*rr->plan);
      throw;
    }
    try {
      plan_shrink(*rr->plan);
    } catch (...) {
      plan_release(*rr->plan);
      throw;
    }
    rr->bytes = plan_bytes(*rr->plan);
    std::vector<std::unique_ptr<Resident>> graveyard;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      if (find_resident(*S, c, geometry)) {
        plan_release(*rr->plan);
        fail(FFMM_ERR_ARGUMENT, "a plan of this cache and geometry is resident");
      }
      while (S->in_use + (int64_t)rr->bytes > S->budget && evict_one(*S, graveyard)) {}
      if (S->in_use + (int64_t)rr->bytes > S->budget) {
        plan_release(*rr->plan);
        fail(FFMM_ERR_MEMORY, "the plan does not fit the server's memory budget");
      }
      rr->last_use = ++S->clock;
      S->in_use += (int64_t)rr->bytes;
      S->peak = std::max(S->peak, S->in_use);
      S->resident.push_back(std::move(rr));
    }
    for (auto &g : graveyard) plan_release(*g->plan);
    S->cv_work.notify_all();
  });
}

int ffmm_server_destroy(ffmm_server S) {
  return guard([&] {
    if (!S) return;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->stop = true;
    }
    S->cv_work.notify_all();
    for (auto &w : S->workers)
      if (w->th.joinable()) w->th.join();
    for (auto &r : S->resident) plan_release(*r->plan);
    S->resident.clear();
    if (S->io) context_release(*S->io);
    for (auto &w : S->workers) {
      w->xyz.release();
      w->dens.release();
      w->out.release();
      w->tgt.release();
      w->tout.release();
      context_release(w->ctx);
    }
    delete S;
  });
}

}  // extern "C"
