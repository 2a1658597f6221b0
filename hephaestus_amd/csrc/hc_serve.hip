// Continuous-batching server of eddy-current time-step solves.
//
// A request is one backward-Euler step of an operator that stays resident on
// the device (TimeDomainEquationSystemProblemOperator::ImplicitSolve,
// src/problem_operators/time_domain_equation_system_problem_operator.cpp:30-51,
// followed by the BackwardEulerSolver update): with A = dt K_nu + M_sigma the
// operator of hc_op_setup(op, dt),
//     f = b_src - K_nu u_n,   A k = f  (Jacobi-PCG or FP32 PCG + FP64 refinement),
//     u_{n+1} = u_n + dt k.
// A scheduler thread admits waiting requests as slots whenever their work
// vectors fit in the device-memory budget (an arena carved once at creation),
// advances every admitted slot in rounds of the batched solve engine (one
// launch per ND order for the operator applies of all slots, one launch per
// vector kernel), and retires finished slots at round boundaries, so requests
// that arrive while others run join the running batch.  Each slot's
// arithmetic does not depend on the other slots: a request's result is the
// same alone and in any mix (same tile variant per order).
//
// With HC_SOLVE_MG the solve of a request is the multigrid PCG of hc_mg.hip,
// whose V-cycle does not batch across operators: up to max_active worker
// threads (default 4) each take the next arrived request whose operator is
// not in use by another worker, solve it on their own stream and retire it, so
// the requests of different operators overlap on the device and a request
// waits only for admission.  The result of a request is again independent of
// the others (its operator's hierarchy, built on first use, is the same).
#include "hc_op.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <thread>

namespace
{

using Clock = std::chrono::steady_clock;

double
ms_between(Clock::time_point a, Clock::time_point b)
{
  return std::chrono::duration<double, std::milli>(b - a).count();
}

// First-fit allocator over [0, size) with coalescing free list.
class Arena
{
public:
  void reset(size_t size)
  {
    free_.clear();
    if (size)
      free_[0] = size;
    used_ = peak_ = 0;
  }
  bool alloc(size_t n, size_t & off)
  {
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= n)
      {
        off = it->first;
        const size_t rest = it->second - n;
        free_.erase(it);
        if (rest)
          free_[off + n] = rest;
        used_ += n;
        peak_ = std::max(peak_, used_);
        return true;
      }
    return false;
  }
  void release(size_t off, size_t n)
  {
    used_ -= n;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first)
    {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin())
    {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first)
      {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t used() const { return used_; }
  size_t peak() const { return peak_; }

private:
  std::map<size_t, size_t> free_;
  size_t used_ = 0, peak_ = 0;
};

struct Request
{
  hc_op_t op;
  const double * bsrc;
  double * u;
  double dt;
  uint64_t tag;
  Clock::time_point arrival, admitted;
  size_t off = 0, bytes = 0;
  hc::Workspace ws;
  double *f = nullptr, *k = nullptr;
  hc::PcgState *in = nullptr, *out = nullptr;
  double * part = nullptr;
  bool fp64 = false; // in the engine's FP64 iteration group (hc::EngineInst::fp64)
};

size_t
request_bytes(int64_t N, bool ir)
{
  using hc::align256;
  return align256(hc::Workspace::bytes_for(N, 1, ir)) + 2 * align256(sizeof(double) * N) +
         2 * align256(sizeof(hc::PcgState)) + align256(sizeof(double) * 4 * hc::RED_BLOCKS);
}

// multigrid mode: f and k (the solver's vectors belong to the operator)
size_t
request_bytes_mg(int64_t N)
{
  return 2 * hc::align256(sizeof(double) * N);
}

constexpr int MG_WORKERS_DEFAULT = 4;

} // namespace

struct hc_server_s
{
  hc_server_config cfg{};
  hipStream_t st = nullptr;
  hc::DevBlock mem;
  Arena arena;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request *> pending;
  std::vector<Request *> active;
  std::deque<hc_request_result> results;
  std::thread th;
  // multigrid mode: worker threads, operators in use
  bool mg = false;
  std::vector<std::thread> workers;
  std::set<hc_op_t> busy;
  int in_flight = 0;
  bool stop = false;
  int error = HC_OK;
  std::string error_msg;
  Clock::time_point t0;
  hc::Engine engine;
  const hc::TileCandidate * cands64[hc::PMAX + 1] = {};
  const hc::TileCandidate * cands32[hc::PMAX + 1] = {};
  // counters
  int64_t n_submitted = 0, n_completed = 0, n_rounds = 0;
  int peak_active = 0;

  void loop();
  void mg_worker();
  int mg_request(Request * r, hipStream_t ws);
  int admit(std::vector<Request *> & fresh);
  int start_request(Request * r);
  int set_engine();
};

// Engine tables over the active requests (one slot each, in order).
int
hc_server_s::set_engine()
{
  std::vector<hc::EngineInst> inst;
  for (Request * r : active)
  {
    hc::EngineInst e{};
    e.op = r->op;
    e.nrhs = 1;
    e.b = r->f;
    e.x = r->k;
    e.rtol = cfg.rtol;
    e.maxit = cfg.maxit;
    e.ws = &r->ws;
    e.in = r->in;
    e.out = r->out;
    e.part = r->part;
    e.fp64 = r->fp64;
    inst.push_back(e);
  }
  return engine.set(inst, cfg.mode, 0, true, cands64, cands32, st);
}

int
hc_server_s::start_request(Request * r)
{
  hc_op_t op = r->op;
  const int64_t N = op->N;
  const bool ir = hc::prec_mode(cfg.mode) == HC_SOLVE_FP32_IR;
  char * p = (char *)mem.base + r->off;
  const size_t wsb = hc::align256(hc::Workspace::bytes_for(N, 1, ir));
  r->ws.carve(p, N, 1, ir);
  p += wsb;
  r->f = (double *)p;
  p += hc::align256(sizeof(double) * N);
  r->k = (double *)p;
  p += hc::align256(sizeof(double) * N);
  r->in = (hc::PcgState *)p;
  p += hc::align256(sizeof(hc::PcgState));
  r->out = (hc::PcgState *)p;
  p += hc::align256(sizeof(hc::PcgState));
  r->part = (double *)p;
  // f = b_src - K u_n  (K from the operator's quadrature data; its apply
  // writes u_ess on PEC rows, zeroed next), k = 0
  // the geometry follows the server's tile variant of this order (the
  // operator's own plan may hold another decomposition)
  const hc::TileCandidate * cand = cands64[op->p] ? cands64[op->p] : op->plan64.cand;
  hc::Geom g;
  hc::PlaneSegs segs;
  hc::geom_for(op, cand, g, segs);
  HC_TRY_HIP(hc::apply_scaled<double>(g, segs, op->tb64, op->qd64, op->ef(false), r->u, r->f, r->ws.part64, 0.0,
                                      1.0 / op->last_dt, cand, op->desc_dev, st));
  HC_TRY_HIP(hc::axpby(N, 1.0, r->bsrc, -1.0, r->f, st));
  HC_TRY_HIP(hc::zero_ess(op->g, r->f, st));
  HC_TRY_HIP(hc::axpby(N, 0.0, r->k, 0.0, r->k, st));
  hc::PcgState hin, hout;
  hc::initial_states(hin, hout, cfg.mode, op->fp32_inner_ok());
  HC_TRY_HIP(hipMemcpyAsync(r->in, &hin, sizeof(hin), hipMemcpyHostToDevice, st));
  HC_TRY_HIP(hipMemcpyAsync(r->out, &hout, sizeof(hout), hipMemcpyHostToDevice, st));
  HC_TRY_HIP(hipStreamSynchronize(st)); // the host state copies above are from the stack
  return HC_OK;
}

int
hc_server_s::admit(std::vector<Request *> & fresh)
{
  const bool ir = hc::prec_mode(cfg.mode) == HC_SOLVE_FP32_IR;
  std::unique_lock<std::mutex> lk(mu);
  const auto now = Clock::now();
  while (!pending.empty())
  {
    Request * r = pending.front();
    if (r->arrival > now)
      break;
    if (cfg.max_active > 0 && (int)(active.size() + fresh.size()) >= cfg.max_active)
      break;
    const size_t need = request_bytes(r->op->N, ir);
    size_t off;
    if (!arena.alloc(need, off))
      break; // first in, first admitted: wait for memory to free up
    r->off = off;
    r->bytes = need;
    r->admitted = now;
    pending.pop_front();
    fresh.push_back(r);
  }
  return HC_OK;
}

void
hc_server_s::loop()
{
  (void)hipSetDevice(0);
  auto set_error = [&](int rc) {
    std::lock_guard<std::mutex> lk(mu);
    error = rc;
    error_msg = hc_last_error();
    cv_done.notify_all();
  };
  for (;;)
  {
    {
      std::unique_lock<std::mutex> lk(mu);
      for (;;)
      {
        if (stop && pending.empty() && active.empty())
          return;
        if (!active.empty())
          break;
        if (!pending.empty())
        {
          const auto wake = pending.front()->arrival;
          if (wake <= Clock::now())
            break;
          cv_work.wait_until(lk, wake);
          continue;
        }
        cv_work.wait(lk);
      }
    }
    std::vector<Request *> fresh;
    admit(fresh);
    int rc = HC_OK;
    for (Request * r : fresh)
      if ((rc = start_request(r)))
        break;
    if (rc)
    {
      // the admitted requests stay owned by the server (released with it)
      {
        std::lock_guard<std::mutex> lk(mu);
        active.insert(active.end(), fresh.begin(), fresh.end());
      }
      set_error(rc);
      return;
    }
    bool changed = !fresh.empty();
    {
      std::lock_guard<std::mutex> lk(mu);
      active.insert(active.end(), fresh.begin(), fresh.end());
      peak_active = std::max(peak_active, (int)active.size());
    }
    if (active.empty())
      continue;
    if (changed && (rc = set_engine()))
    {
      set_error(rc);
      return;
    }
    if ((rc = engine.round(cfg.round_iters, st)) || (rc = engine.gather(st)))
    {
      set_error(rc);
      return;
    }
    n_rounds++;
    // retire finished slots
    const hc::PcgState * hs = engine.host_states();
    std::vector<Request *> keep, gone;
    bool grow64 = false; // a request whose slot continues in FP64 joins the FP64 group
    for (size_t i = 0; i < active.size(); i++)
    {
      Request * r = active[i];
      if (hs[i].prec64 && hs[i].phase != hc::PH_DONE && !r->fp64 && r->op->fp32_inner_ok() &&
          hc::prec_mode(cfg.mode) == HC_SOLVE_FP32_IR)
      {
        r->fp64 = true;
        grow64 = true;
      }
      (hs[i].phase == hc::PH_DONE ? gone : keep).push_back(r);
    }
    if (gone.empty())
    {
      if (grow64 && (rc = set_engine()))
      {
        set_error(rc);
        return;
      }
      continue;
    }
    std::vector<hc_request_result> res;
    for (size_t i = 0; i < active.size(); i++)
    {
      if (hs[i].phase != hc::PH_DONE)
        continue;
      Request * r = active[i];
      // u_{n+1} = u_n + dt k
      if (hipSuccess != hc::axpby(r->op->N, r->dt, r->k, 1.0, r->u, st))
      {
        set_error(hc::fail(HC_EHIP, "server: update launch failed"));
        return;
      }
      const hc::SlotStats s = hc::Engine::stats_of(hs[i]);
      hc_request_result o{};
      o.tag = r->tag;
      o.status = s.status ? s.status : (s.converged ? HC_OK : HC_ENOTCONV);
      o.converged = s.converged;
      o.iterations = s.inner;
      o.outer_iterations = s.outer;
      o.rel_res = s.rel_res;
      o.fp64_fallback = (hc::prec_mode(cfg.mode) == HC_SOLVE_FP32_IR) ? s.fp64 : 0;
      res.push_back(o);
    }
    if (hipStreamSynchronize(st) != hipSuccess)
    {
      set_error(hc::fail(HC_EHIP, "server: stream synchronisation failed"));
      return;
    }
    const auto now = Clock::now();
    {
      std::lock_guard<std::mutex> lk(mu);
      size_t g = 0;
      for (Request * r : gone)
      {
        res[g].latency_ms = ms_between(r->arrival, now);
        res[g].queue_ms = ms_between(r->arrival, r->admitted);
        res[g].arrival_ms = ms_between(t0, r->arrival);
        results.push_back(res[g]);
        arena.release(r->off, r->bytes);
        delete r;
        g++;
        n_completed++;
      }
      active.swap(keep);
      cv_done.notify_all();
    }
    if (!active.empty() && (rc = set_engine()))
    {
      set_error(rc);
      return;
    }
  }
}

// One multigrid request on the worker's stream: f = b_src - K u, A k = f by
// multigrid PCG (PCG with the gradient-space correction when the operator's
// mesh has no hierarchy), u += dt k.
int
hc_server_s::mg_request(Request * r, hipStream_t ws)
{
  hc_op_t op = r->op;
  const int64_t N = op->N;
  char * p = (char *)mem.base + r->off;
  r->f = (double *)p;
  r->k = (double *)(p + hc::align256(sizeof(double) * N));
  if (int rc = hc_op_apply_curl(op, r->u, r->f, ws))
    return rc;
  HC_TRY_HIP(hc::axpby(N, 1.0, r->bsrc, -1.0, r->f, ws));
  HC_TRY_HIP(hc::zero_ess(op->g, r->f, ws));
  HC_TRY_HIP(hc::axpby(N, 0.0, r->k, 0.0, r->k, ws));
  int nlev = 0;
  int64_t nc = 0;
  const bool has_mg = hc::mg_levels(op, ws, &nlev, &nc) == HC_OK;
  hc_solve_stats s{};
  int rc = has_mg ? hc::mg_pcg_solve(op, r->f, r->k, cfg.rtol, cfg.maxit, &s, ws, false)
                  : hc_pcg_solve(op, r->f, r->k, 1, cfg.rtol, cfg.maxit, HC_SOLVE_FP64 | HC_SOLVE_AUX, &s, ws);
  if (rc && rc != HC_ENOTCONV && rc != HC_ENAN)
    return rc;
  HC_TRY_HIP(hc::axpby(N, r->dt, r->k, 1.0, r->u, ws));
  HC_TRY_HIP(hipStreamSynchronize(ws));
  hc_request_result o{};
  o.tag = r->tag;
  o.status = s.status ? s.status : (s.converged ? HC_OK : HC_ENOTCONV);
  o.converged = s.converged;
  o.iterations = s.iterations;
  o.outer_iterations = s.outer_iterations;
  o.rel_res = s.rel_res;
  o.fp64_fallback = 0;
  const auto now = Clock::now();
  o.latency_ms = ms_between(r->arrival, now);
  o.queue_ms = ms_between(r->arrival, r->admitted);
  o.arrival_ms = ms_between(t0, r->arrival);
  std::lock_guard<std::mutex> lk(mu);
  results.push_back(o);
  n_completed++;
  cv_done.notify_all();
  return HC_OK;
}

void
hc_server_s::mg_worker()
{
  (void)hipSetDevice(0);
  hipStream_t ws = nullptr;
  if (hipStreamCreate(&ws) != hipSuccess)
  {
    std::lock_guard<std::mutex> lk(mu);
    error = hc::fail(HC_EHIP, "server: hipStreamCreate failed");
    error_msg = hc_last_error();
    cv_done.notify_all();
    return;
  }
  for (;;)
  {
    Request * r = nullptr;
    {
      std::unique_lock<std::mutex> lk(mu);
      for (;;)
      {
        if (error || (stop && pending.empty()))
          break;
        // the first arrived request whose operator is free and whose vectors fit
        const auto now = Clock::now();
        Clock::time_point wake = Clock::time_point::max();
        for (auto it = pending.begin(); it != pending.end(); ++it)
        {
          if ((*it)->arrival > now)
          {
            wake = (*it)->arrival;
            break;
          }
          if (busy.count((*it)->op))
            continue;
          size_t off;
          const size_t need = request_bytes_mg((*it)->op->N);
          if (!arena.alloc(need, off))
            break; // first in, first admitted: wait for memory to free up
          r = *it;
          r->off = off;
          r->bytes = need;
          r->admitted = now;
          pending.erase(it);
          busy.insert(r->op);
          in_flight++;
          peak_active = std::max(peak_active, in_flight);
          break;
        }
        if (r)
          break;
        if (wake != Clock::time_point::max())
          cv_work.wait_until(lk, wake);
        else
          cv_work.wait(lk);
      }
    }
    if (!r)
      break;
    const int rc = mg_request(r, ws);
    {
      std::lock_guard<std::mutex> lk(mu);
      arena.release(r->off, r->bytes);
      busy.erase(r->op);
      in_flight--;
      if (rc && !error)
      {
        error = rc;
        error_msg = hc_last_error();
      }
      delete r;
      cv_done.notify_all();
    }
    cv_work.notify_all(); // an operator or memory became free
    if (rc)
      break;
  }
  (void)hipStreamSynchronize(ws);
  (void)hipStreamDestroy(ws);
  cv_work.notify_all();
}

// Runs a server thread's body; what it throws becomes the server's error
// (the thread ends; polls then report the error).
template <typename F>
static void
guarded(hc_server_t s, F && body)
{
  try
  {
    body();
  }
  catch (const std::exception & ex)
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (!s->error)
    {
      s->error = hc::fail(HC_EHIP, std::string("server thread: ") + ex.what());
      s->error_msg = hc_last_error();
    }
    s->cv_done.notify_all();
  }
}

extern "C" {

int64_t
hc_server_request_bytes(int64_t ndofs, int mode)
{
  if (ndofs <= 0)
    return -1;
  if (mode & HC_SOLVE_MG)
    return (int64_t)request_bytes_mg(ndofs);
  return (int64_t)request_bytes(ndofs, hc::prec_mode(mode) == HC_SOLVE_FP32_IR);
}

int
hc_server_create(const hc_server_config * cfg, hc_server_t * out)
try
{
  if (!cfg || !out)
    return hc::fail(HC_EINVAL, "hc_server_create: null argument");
  if (cfg->hbm_budget_bytes <= 0 || !(cfg->rtol > 0) || cfg->maxit < 1 ||
      (hc::prec_mode(cfg->mode) != HC_SOLVE_FP64 && hc::prec_mode(cfg->mode) != HC_SOLVE_FP32_IR) ||
      (cfg->mode & ~(HC_SOLVE_AUX | HC_SOLVE_MG | 0xF)))
    return hc::fail(HC_EINVAL, "hc_server_create: budget > 0, rtol > 0, maxit >= 1 and a solve mode are required");
  if ((cfg->mode & HC_SOLVE_MG) && hc::prec_mode(cfg->mode) != HC_SOLVE_FP64)
    return hc::fail(HC_EINVAL, "hc_server_create: HC_SOLVE_MG runs in FP64 (mode HC_SOLVE_FP64 | HC_SOLVE_MG)");
  hc_server_t s = new hc_server_s();
  s->cfg = *cfg;
  if (s->cfg.round_iters <= 0)
    s->cfg.round_iters = 16;
  if (hipStreamCreate(&s->st) != hipSuccess)
  {
    delete s;
    return hc::fail(HC_EHIP, "hc_server_create: hipStreamCreate failed");
  }
  int rc = s->mem.alloc((size_t)cfg->hbm_budget_bytes);
  if (rc)
  {
    (void)hipStreamDestroy(s->st);
    delete s;
    return rc;
  }
  s->arena.reset((size_t)cfg->hbm_budget_bytes);
  for (int p = 1; p <= hc::PMAX; p++)
  {
    const hc::TileCandidate * c;
    if (hc::tile_candidates(p, &c) > 0)
      s->cands64[p] = s->cands32[p] = &c[0];
  }
  s->t0 = Clock::now();
  s->mg = (cfg->mode & HC_SOLVE_MG) != 0;
  // (no C++ exception may cross the C ABI: a thread that cannot be started
  // is reported as an error; with some multigrid workers running, they serve)
  try
  {
    if (s->mg)
    {
      const int nw = cfg->max_active > 0 ? cfg->max_active : MG_WORKERS_DEFAULT;
      for (int i = 0; i < nw; i++)
        s->workers.emplace_back([s] { guarded(s, [s] { s->mg_worker(); }); });
    }
    else
      s->th = std::thread([s] { guarded(s, [s] { s->loop(); }); });
  }
  catch (const std::exception & ex)
  {
    if (!s->mg || s->workers.empty())
    {
      hc_server_destroy(s);
      return hc::fail(HC_EHIP, std::string("hc_server_create: cannot start a thread: ") + ex.what());
    }
  }
  *out = s;
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_server_set_tile(hc_server_t s, int order, int prec_bits, int tx, int ty, int tz, int mode)
try
{
  if (!s || order < 1 || order > hc::PMAX || (prec_bits != 64 && prec_bits != 32))
    return hc::fail(HC_EINVAL, "hc_server_set_tile: bad argument");
  const hc::TileCandidate * c;
  const int n = hc::tile_candidates(order, &c);
  for (int i = 0; i < n; i++)
    if (c[i].shape.tx == tx && c[i].shape.ty == ty && c[i].shape.tz == tz && (mode < 0 || c[i].mode == mode))
    {
      std::lock_guard<std::mutex> lk(s->mu);
      if (!s->active.empty() || !s->pending.empty())
        return hc::fail(HC_EBUSY, "hc_server_set_tile: requests are in flight");
      (prec_bits == 64 ? s->cands64 : s->cands32)[order] = &c[i];
      return HC_OK;
    }
  return hc::fail(HC_EINVAL, "hc_server_set_tile: no compiled tile variant with this shape");
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_server_submit(hc_server_t s, hc_op_t op, const double * b_src, double * u, double dt, uint64_t tag,
                 double arrival_ms)
try
{
  if (!s || !op || !b_src || !u)
    return hc::fail(HC_EINVAL, "hc_server_submit: null argument");
  if (!op->setup_done || std::fabs(dt - op->last_dt) > 1e-12 * dt)
    return hc::fail(HC_EINVAL, "hc_server_submit: the operator must be set up with this time step");
  const bool ir = hc::prec_mode(s->cfg.mode) == HC_SOLVE_FP32_IR;
  if (ir && !op->qd32)
    return hc::fail(HC_EINVAL, "hc_server_submit: FP32 refinement needs an operator created with HC_PREC_FP32");
  if (s->mg && !op->qd64)
    return hc::fail(HC_EINVAL, "hc_server_submit: multigrid needs an operator created with HC_PREC_FP64");
  if ((s->mg ? request_bytes_mg(op->N) : request_bytes(op->N, ir)) > (size_t)s->cfg.hbm_budget_bytes)
    return hc::fail(HC_EBUSY, "hc_server_submit: the request's work vectors exceed the whole memory budget");
  Request * r = new Request();
  r->op = op;
  r->bsrc = b_src;
  r->u = u;
  r->dt = dt;
  r->tag = tag;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->error)
    {
      delete r;
      return hc::fail(s->error, "hc_server_submit: the server stopped on an error: " + s->error_msg);
    }
    const auto now = Clock::now();
    r->arrival = arrival_ms < 0 ? now
                                : s->t0 + std::chrono::duration_cast<Clock::duration>(
                                              std::chrono::duration<double, std::milli>(arrival_ms));
    // keep the queue ordered by arrival
    auto it = s->pending.end();
    while (it != s->pending.begin() && (*std::prev(it))->arrival > r->arrival)
      --it;
    s->pending.insert(it, r);
    s->n_submitted++;
  }
  s->cv_work.notify_all();
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_server_poll(hc_server_t s, hc_request_result * out, int max, int wait_ms)
try
{
  if (!s || (max > 0 && !out))
    return hc::fail(HC_EINVAL, "hc_server_poll: bad argument");
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->results.empty() && wait_ms > 0 && !s->error)
    s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms),
                        [&] { return !s->results.empty() || s->error; });
  if (s->error && s->results.empty())
    return hc::fail(s->error, "hc_server_poll: the server stopped on an error: " + s->error_msg);
  int n = 0;
  while (n < max && !s->results.empty())
  {
    out[n++] = s->results.front();
    s->results.pop_front();
  }
  return n;
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_server_info(hc_server_t s, hc_server_stats * st)
try
{
  if (!s || !st)
    return hc::fail(HC_EINVAL, "hc_server_info: null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  st->submitted = s->n_submitted;
  st->completed = s->n_completed;
  st->rounds = s->n_rounds;
  st->pending = (int64_t)s->pending.size();
  st->active = s->mg ? (int64_t)s->in_flight : (int64_t)s->active.size();
  st->peak_active = s->peak_active;
  st->budget_bytes = s->cfg.hbm_budget_bytes;
  st->peak_bytes = (int64_t)s->arena.peak();
  st->elapsed_ms = ms_between(s->t0, Clock::now());
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_server_destroy(hc_server_t s)
try
{
  if (!s)
    return hc::fail(HC_EINVAL, "hc_server_destroy: null handle");
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    if (s->error)
    {
      // the scheduler has returned; drop what is left
      for (Request * r : s->pending)
        delete r;
      for (Request * r : s->active)
        delete r;
      s->pending.clear();
      s->active.clear();
    }
  }
  s->cv_work.notify_all();
  if (s->th.joinable())
    s->th.join();
  for (auto & w : s->workers)
    if (w.joinable())
      w.join();
  (void)hipStreamSynchronize(s->st);
  (void)hipStreamDestroy(s->st);
  delete s;
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

} // extern "C"
