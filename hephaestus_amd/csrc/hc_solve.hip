// Batched solve engine: Jacobi-PCG (the device side of DefaultJacobiPCGSolver,
// src/solvers/hephaestus_solvers.hpp:34-57) and FP32 PCG + FP64 iterative
// refinement over the right-hand sides of one or more operator instances, and
// the C ABI entry points built on it (hc_pcg_solve, hc_batch_*).
//
// Every right-hand side of every instance is one solve slot (hc_pcg.hip); the
// operator applies of one iteration are one launch per ND order over a device
// table of instances (hc_apply.hip), each gated by the slots' done flags.
#include "hc_op.hpp"

#include <chrono>
#include <condition_variable>
#include <mutex>
#include <set>
#include <thread>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>

namespace hc
{

namespace
{

constexpr int MAX_OUTER_FP64 = 5;  // PCG runs (1 + residual-replacement restarts)
constexpr int MAX_OUTER_IR = 400;  // FP64 residual replacements of the FP32 PCG
constexpr int GRAPH_MIN_ITERS = 64; // rounds of this many iterations replay a captured graph
// FP32 PCG: an FP64 residual replacement each time ||r|| fell by 10x
constexpr double RELIABLE_DELTA2 = 1e-2;

} // namespace

void
geom_for(const hc_op_s * op, const TileCandidate * c, Geom & g, PlaneSegs & segs)
{
  g = op->g;
  g.T[0] = c->shape.tx;
  g.T[1] = c->shape.ty;
  g.T[2] = c->shape.tz;
  for (int a = 0; a < 3; a++)
    g.nt[a] = (g.n[a] + g.T[a] - 1) / g.T[a];
  segs = make_plane_segs(g);
}

void
initial_states(PcgState & in, PcgState & out, int mode, bool fp32_ok)
{
  std::memset(&in, 0, sizeof(in));
  std::memset(&out, 0, sizeof(out));
  in.done = 1; // no inner solve running
  out.phase = PH_FIRST;
  out.prec64 = (prec_mode(mode) == HC_SOLVE_FP32_IR && fp32_ok) ? 0 : 1;
}

Engine::~Engine()
{
  drop_graphs();
  if (own_st_)
    (void)hipStreamDestroy(own_st_);
  if (host_)
    (void)hipHostFree(host_);
}

int
Engine::max_outer() const
{
  return mode_ == HC_SOLVE_FP32_IR ? MAX_OUTER_IR : MAX_OUTER_FP64;
}

int
Engine::set(const std::vector<EngineInst> & inst, int mode, int rhs_block, bool force_batched,
            const TileCandidate * const * cands64, const TileCandidate * const * cands32, hipStream_t st)
{
  drop_graphs(); // captured rounds hold the old tables' sizes
  inst_ = inst;
  mode_ = prec_mode(mode);
  aux_ = (mode & HC_SOLVE_AUX) != 0;
  const bool ir = (mode_ == HC_SOLVE_FP32_IR);
  nslots_ = 0;
  int nb_max = 1;
  int owned = 0; // slots whose states the engine owns
  for (auto & e : inst_)
  {
    if (e.nrhs < 1)
      return fail(HC_EINVAL, "solve: nrhs must be >= 1");
    if (!e.op->setup_done)
      return fail(HC_EINVAL, "solve: hc_op_setup has not been called on an operator of the batch");
    if (ir && !e.op->qd32)
      return fail(HC_EINVAL, "solve: FP32 refinement needs operators created with HC_PREC_FP32");
    if (!e.ws || e.ws->nrhs < e.nrhs || (ir && !e.ws->has32))
      return fail(HC_EINVAL, "solve: work vectors missing");
    if (aux_ && (!e.op->ginv64 || (ir && !e.op->ginv32)))
      return fail(HC_EINVAL, "solve: gradient-space data missing (hc_op_setup has not run)");
    e.slot0 = nslots_;
    nslots_ += e.nrhs;
    if (!e.in)
      owned += e.nrhs;
    nb_max = std::max(nb_max, red_blocks(e.op->N));
  }
  if (nslots_ == 0)
    return fail(HC_EINVAL, "solve: empty batch");
  std::vector<std::vector<int>> by_order(PMAX + 1);
  for (int i = 0; i < (int)inst_.size(); i++)
    by_order[inst_[i].op->p].push_back(i);
  size_t ndesc_total = 0;
  for (int p = 1; p <= PMAX; p++)
    ndesc_total += by_order[p].size();
  // device layout: slot table, gathered states, counters, apply tables; the
  // states the engine owns live in own_ (kept when a later set() passes them
  // back, as the compaction of a running batch does)
  const size_t o_slots = 0;
  const size_t o_gath = align256(o_slots + sizeof(SolveSlot) * nslots_);
  const size_t o_git = align256(o_gath + sizeof(PcgState) * nslots_);
  const size_t o_blk = align256(o_git + 2 * sizeof(int));
  const size_t o_desc = align256(o_blk + sizeof(int) * (nslots_ + 1));
  const size_t total = o_desc + sizeof(ApplyDesc) * ndesc_total * K_COUNT;
  const size_t own_in = 0;
  const size_t own_out = align256(own_in + sizeof(PcgState) * owned);
  const size_t own_part = align256(own_out + sizeof(PcgState) * owned);
  const size_t own_total = own_part + sizeof(double) * 4 * RED_BLOCKS * (size_t)owned;
  if (owned > 0 && own_.bytes < own_total)
  {
    int rc = own_.alloc(own_total);
    if (rc)
      return rc;
  }
  // the iteration counter survives a rebuild (slots may be mid-solve)
  int git_host[2] = {0, 0};
  if (git_)
  {
    HC_TRY_HIP(hipMemcpyAsync(git_host, git_, sizeof(git_host), hipMemcpyDeviceToHost, st));
    HC_TRY_HIP(hipStreamSynchronize(st));
  }
  if (mem_.bytes < total)
  {
    int rc = mem_.alloc(total);
    if (rc)
      return rc;
  }
  char * base = (char *)mem_.base;
  SolveSlot * d_slots = (SolveSlot *)(base + o_slots);
  gathered_ = (PcgState *)(base + o_gath);
  git_ = (int *)(base + o_git);
  int * d_blk0 = (int *)(base + o_blk);
  ApplyDesc * d_desc = (ApplyDesc *)(base + o_desc);
  int opos = 0;
  for (auto & e : inst_)
    if (!e.in)
    {
      char * ob = (char *)own_.base;
      e.in = (PcgState *)(ob + own_in) + opos;
      e.out = (PcgState *)(ob + own_out) + opos;
      e.part = (double *)(ob + own_part) + (size_t)opos * 4 * RED_BLOCKS;
      opos += e.nrhs;
    }
  rhs_block_ = rhs_block;
  force_batched_ = force_batched;
  for (int p = 0; p <= PMAX; p++)
  {
    cands64_[p] = cands64 ? cands64[p] : nullptr;
    cands32_[p] = cands32 ? cands32[p] : nullptr;
  }
  // slots
  std::vector<SolveSlot> hs(nslots_);
  for (auto & e : inst_)
  {
    const hc_op_s * op = e.op;
    const int64_t N = op->N;
    const Workspace & w = *e.ws;
    for (int r = 0; r < e.nrhs; r++)
    {
      SolveSlot & s = hs[e.slot0 + r];
      std::memset(&s, 0, sizeof(s));
      const int64_t o = (int64_t)r * N;
      s.n = N;
      s.rtol = e.rtol;
      s.maxit = e.maxit;
      s.delta2 = ir ? RELIABLE_DELTA2 : 0.0;
      s.b = e.b + o;
      s.x = e.x + o;
      s.r64 = w.r64 + o;
      s.p64 = w.p64 + o;
      s.q64 = w.q64 + o;
      s.dinv64 = op->dinv64;
      if (ir)
      {
        s.r32 = w.r32 + o;
        s.x32 = w.x32 + o;
        s.p32 = w.p32 + o;
        s.q32 = w.q32 + o;
        s.dinv32 = op->dinv32;
      }
      s.in = e.in + r;
      s.out = e.out + r;
      s.part = e.part + (size_t)r * 4 * RED_BLOCKS;
      s.g64 = (!ir || !op->fp32_inner_ok() || e.fp64) ? 1 : 0;
      if (aux_)
      {
        s.aux = 1;
        s.ag = op->ag;
        s.ginv64 = op->ginv64;
        s.t64 = w.t64 + (size_t)r * w.NNC;
        if (ir)
        {
          s.ginv32 = op->ginv32;
          s.t32 = w.t32 + (size_t)r * w.NNC;
        }
      }
    }
  }
  // flattened slot grid: the blocks of slot i are [blk0[i], blk0[i + 1])
  std::vector<int> hblk(nslots_ + 1, 0);
  for (int i = 0; i < nslots_; i++)
    hblk[i + 1] = hblk[i] + red_blocks(hs[i].n);
  table_.slots = d_slots;
  table_.nslots = nslots_;
  table_.nb_max = nb_max;
  table_.any_aux = aux_ ? 1 : 0;
  table_.nblk = hblk[nslots_];
  table_.blk0 = d_blk0;
  // new slots start in the mode's precision (FP64 where the operator does not
  // admit FP32 inner solves)
  for (auto & e : inst_)
  {
    if (ir && e.op->fp32_inner_ok())
      any32_ = true;
    else
      any64_ = true;
  }
  // apply groups
  std::vector<ApplyDesc> hd(ndesc_total * K_COUNT);
  size_t dpos = 0;
  for (int k = 0; k < K_COUNT; k++)
  {
    groups_[k].clear();
    if (k == K_PQ32 && !ir)
      continue;
    const bool prec32 = (k == K_PQ32);
    for (int p = 1; p <= PMAX; p++)
    {
      // the FP32 group holds the instances whose slots may run FP32 inner
      // solves (work items of the others would only be gated off)
      std::vector<int> ids;
      for (int i : by_order[p])
      {
        const bool may32 = ir && inst_[i].op->fp32_inner_ok();
        // FP32 group: instances that may run FP32; FP64 iteration group in
        // FP32 + refinement: the others and those promoted after a slot
        // continued in FP64 (the residual group K_XQ holds every instance)
        if (prec32 ? may32 : (k != K_PQ64 || !may32 || inst_[i].fp64))
          ids.push_back(i);
      }
      if (ids.empty())
        continue;
      const hc_op_s * op0 = inst_[ids[0]].op;
      ApplyGroup G{};
      G.order = p;
      G.prec = prec32 ? 32 : 64;
      const TileCandidate * over = prec32 ? (cands32 ? cands32[p] : nullptr) : (cands64 ? cands64[p] : nullptr);
      G.cand = over ? over : (prec32 ? op0->plan32.cand : op0->plan64.cand);
      G.tables = prec32 ? (const void *)&op0->tb32 : (const void *)&op0->tb64;
      G.batched = force_batched || ids.size() > 1 || inst_[ids[0]].nrhs > 1;
      G.descs = d_desc + dpos;
      G.ndesc = (int)ids.size();
      G.grouped = 0;
      int w0 = 0, pw0 = 0;
      bool all_ef = true; // every instance has element factors: one launch reads them all
      for (int i : ids)
      {
        const EngineInst & e = inst_[i];
        const hc_op_s * op = e.op;
        const Workspace & w = *e.ws;
        ApplyDesc d{};
        geom_for(op, G.cand, d.g, d.segs);
        d.prows = (int)d.segs.start[d.segs.nseg];
        d.ntiles = d.g.nt[0] * d.g.nt[1] * d.g.nt[2];
        d.nrhs = e.nrhs;
        d.rb = G.batched ? (rhs_block > 0 ? std::min(rhs_block, e.nrhs) : default_rhs_block(d.ntiles, e.nrhs)) : 1;
        d.qd = prec32 ? (const void *)op->qd32 : (const void *)op->qd64;
        d.ef = op->ef(prec32);
        if (!d.ef.qc)
          all_ef = false;
        if (k == K_PQ32)
        {
          d.x = w.p32;
          d.y = w.q32;
          d.part = w.part32;
          d.done = &e.in->run;
          d.gate = 1;
        }
        else if (k == K_PQ64)
        {
          d.x = w.p64;
          d.y = w.q64;
          d.part = w.part64;
          d.done = &e.in->run;
          d.gate = 2;
        }
        else
        {
          d.x = e.x;
          d.y = w.q64;
          d.part = w.part64;
          d.done = &e.out->phase;
          d.gate = PH_RESID;
        }
        d.done_stride = STATE_STRIDE_INT;
        d.w0 = w0;
        d.pw0 = pw0;
        w0 += d.ntiles * ((d.nrhs + d.rb - 1) / d.rb);
        pw0 += d.prows * d.nrhs;
        if (d.rb > 1)
          G.grouped = 1;
        if (!G.batched)
          G.single = d;
        hd[dpos++] = d;
      }
      if (all_ef)
        G.grouped |= LAUNCH_EF;
      else // mixed: every instance reads its quadrature-point factors
        for (size_t q = dpos - ids.size(); q < dpos; q++)
          hd[q].ef = ElemFactors{};
      if (!G.batched)
        G.single = hd[dpos - 1];
      G.nwork = w0;
      G.npwork = pw0;
      groups_[k].push_back(G);
    }
  }
  HC_TRY_HIP(hipMemcpyAsync(d_slots, hs.data(), sizeof(SolveSlot) * nslots_, hipMemcpyHostToDevice, st));
  HC_TRY_HIP(hipMemcpyAsync(d_blk0, hblk.data(), sizeof(int) * (nslots_ + 1), hipMemcpyHostToDevice, st));
  HC_TRY_HIP(hipMemcpyAsync(d_desc, hd.data(), sizeof(ApplyDesc) * hd.size(), hipMemcpyHostToDevice, st));
  HC_TRY_HIP(hipMemcpyAsync(git_, git_host, sizeof(git_host), hipMemcpyHostToDevice, st));
  HC_TRY_HIP(hipStreamSynchronize(st)); // the host tables above are about to go out of scope
  if (host_cap_ < nslots_)
  {
    if (host_)
      (void)hipHostFree(host_);
    host_ = nullptr;
    host_cap_ = 0;
    if (hipHostMalloc((void **)&host_, sizeof(PcgState) * nslots_, hipHostMallocDefault) != hipSuccess)
      return fail(HC_ENOMEM, "hipHostMalloc of the solver state failed");
    host_cap_ = nslots_;
  }
  return HC_OK;
}

int
Engine::apply(Kind k, hipStream_t st)
{
  for (const ApplyGroup & G : groups_[k])
  {
    hipError_t e;
    if (G.batched)
      e = G.prec == 64 ? apply_multi<double>(G.cand, *(const Tables<double> *)G.tables, G.descs, G.ndesc, G.nwork,
                                             G.npwork, G.grouped, st)
                       : apply_multi<float>(G.cand, *(const Tables<float> *)G.tables, G.descs, G.ndesc, G.nwork,
                                            G.npwork, G.grouped, st);
    else
    {
      e = G.cand->launch(G.single, nullptr, 1, G.nwork, 0, G.tables, G.prec, st);
      if (e == hipSuccess)
        e = G.prec == 64 ? assemble_planes<double>(G.single, nullptr, 1, G.npwork, st)
                         : assemble_planes<float>(G.single, nullptr, 1, G.npwork, st);
    }
    if (e != hipSuccess)
      return fail(HC_EHIP, std::string("operator apply launch: ") + hipGetErrorString(e));
  }
  return HC_OK;
}

int
Engine::reset_counter(hipStream_t st)
{
  HC_TRY_HIP(zero_int(git_, 2, st));
  return HC_OK;
}

namespace
{
std::atomic<int> g_graphs{0};
} // namespace

bool
graphs_enabled()
{
  return g_graphs.load() != 0;
}

void
Engine::drop_graphs()
{
  for (auto & g : graphs_)
    (void)hipGraphExecDestroy(g.second);
  graphs_.clear();
}

int
Engine::round(int iters, hipStream_t st, bool graph)
{
  const bool ir = (mode_ == HC_SOLVE_FP32_IR);
  // the precisions that have slots running, as of the last gathered states
  // (a slot that switched to FP64 in this round starts iterating next round)
  const bool run32 = ir && any32_, run64 = !ir || any64_;
  if (!graph || !st || !graphs_enabled())
    return round_launches(iters, run32, run64, st);
  const int key = 4 * iters + 2 * (run32 ? 1 : 0) + (run64 ? 1 : 0);
  hipGraphExec_t ex = nullptr;
  for (auto & g : graphs_)
    if (g.first == key)
      ex = g.second;
  if (!ex)
  {
    HC_TRY_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int rc = round_launches(iters, run32, run64, st);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (rc || e != hipSuccess)
    {
      if (g)
        (void)hipGraphDestroy(g);
      return rc ? rc : fail(HC_EHIP, std::string("round capture: ") + hipGetErrorString(e));
    }
    const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess)
      return fail(HC_EHIP, std::string("round graph instantiation: ") + hipGetErrorString(ei));
    graphs_.push_back({key, ex});
  }
  HC_TRY_HIP(hipGraphLaunch(ex, st));
  return HC_OK;
}

int
Engine::round_launches(int iters, bool run32, bool run64, hipStream_t st)
{
  const bool ir = (mode_ == HC_SOLVE_FP32_IR);
  int rc;
  HC_TRY_HIP(phase_prep(table_, st));
  if (ir)
    HC_TRY_HIP(ir_update(table_, st));
  if ((rc = apply(K_XQ, st)))
    return rc;
  HC_TRY_HIP(residual64(table_, max_outer(), MAX_OUTER_FP64, st));
  if (ir)
    HC_TRY_HIP(pcg_init_reliable(table_, git_, st));
  HC_TRY_HIP(pcg_init64(table_, git_, st));
  for (int k = 0; k < iters; k++)
  {
    if (run32)
    {
      if ((rc = apply(K_PQ32, st)))
        return rc;
      HC_TRY_HIP(pcg_iterate<float>(table_, git_, st));
    }
    if (run64)
    {
      if ((rc = apply(K_PQ64, st)))
        return rc;
      HC_TRY_HIP(pcg_iterate<double>(table_, git_, st));
    }
  }
  return HC_OK;
}

int
Engine::gather(hipStream_t st)
{
  HC_TRY_HIP(gather_states(table_, gathered_, st));
  HC_TRY_HIP(hipMemcpyAsync(host_, gathered_, sizeof(PcgState) * nslots_, hipMemcpyDeviceToHost, st));
  HC_TRY_HIP(hipStreamSynchronize(st));
  any32_ = any64_ = false;
  for (int i = 0; i < nslots_; i++)
    if (host_[i].phase != PH_DONE)
      (host_[i].prec64 ? any64_ : any32_) = true;
  return HC_OK;
}

bool
Engine::promote64()
{
  if (mode_ != HC_SOLVE_FP32_IR)
    return false;
  bool grow = false;
  for (auto & e : inst_)
  {
    if (e.fp64)
      continue;
    for (int r = 0; r < e.nrhs; r++)
    {
      const PcgState & h = host_[e.slot0 + r];
      if (h.prec64 && h.phase != PH_DONE)
      {
        e.fp64 = true;
        grow = grow || e.op->fp32_inner_ok(); // the others are in the group already
        break;
      }
    }
  }
  return grow;
}

SlotStats
Engine::stats_of(const PcgState & s)
{
  SlotStats o{};
  o.inner = s.inner;
  o.outer = s.iters;
  o.status = s.status;
  o.rel_res = s.bb > 0 ? std::sqrt(s.rr / s.bb) : 0.0;
  o.converged = (s.status == 0 && s.rr <= s.tol2 * s.bb * (1.0 + 1e-6)) ? 1 : 0;
  o.fp64 = s.prec64;
  return o;
}

int
Engine::solve(int mode, hipStream_t st_in, std::vector<SlotStats> & stats, double * ms)
{
  if (prec_mode(mode) != mode_ || ((mode & HC_SOLVE_AUX) != 0) != aux_)
    return fail(HC_EINVAL, "solve: mode differs from the one the batch was set up for");
  // work of a solve issued on the null stream runs on the engine's own
  // (blocking, default-priority) stream, which orders with the null stream and
  // can be captured into graphs
  if (!st_in && !own_st_)
    HC_TRY_HIP(hipStreamCreate(&own_st_));
  hipStream_t st = st_in ? st_in : own_st_;
  int maxit_all = 0;
  for (auto & e : inst_)
    maxit_all = std::max(maxit_all, e.maxit);
  // fresh states for every slot
  std::vector<PcgState> hin(nslots_), hout(nslots_);
  for (int i = 0; i < nslots_; i++)
    initial_states(hin[i], hout[i], mode, true);
  any32_ = any64_ = false;
  for (auto & e : inst_)
    for (int r = 0; r < e.nrhs; r++)
    {
      initial_states(hin[e.slot0 + r], hout[e.slot0 + r], mode, e.op->fp32_inner_ok());
      (hout[e.slot0 + r].prec64 ? any64_ : any32_) = true;
    }
  for (auto & e : inst_)
  {
    HC_TRY_HIP(hipMemcpyAsync(e.in, hin.data() + e.slot0, sizeof(PcgState) * e.nrhs, hipMemcpyHostToDevice, st));
    HC_TRY_HIP(hipMemcpyAsync(e.out, hout.data() + e.slot0, sizeof(PcgState) * e.nrhs, hipMemcpyHostToDevice, st));
  }
  hipEvent_t e0, e1;
  HC_TRY_HIP(hipEventCreate(&e0));
  HC_TRY_HIP(hipEventCreate(&e1));
  HC_TRY_HIP(hipEventRecord(e0, st));
  int rc;
  if ((rc = reset_counter(st)))
    return rc;
  // rounds until every slot is done; the iteration count per round grows so
  // that the host looks at the states rarely once solves are long.  Instances
  // whose slots are all done leave the tables once a quarter of those still
  // running has finished, so a long tail does not launch work for them.
  const std::vector<EngineInst> all = inst_;
  std::vector<int> orig(all.size()); // original index of each current instance
  for (size_t i = 0; i < all.size(); i++)
    orig[i] = (int)i;
  stats.assign(nslots_, SlotStats{});
  int chunk = 8, rounds = 0;
  const int max_rounds = max_outer() * (maxit_all / 8 + 4) + 8;
  for (;;)
  {
    if ((rc = round(chunk, st, chunk >= GRAPH_MIN_ITERS)))
      return rc;
    if ((rc = gather(st)))
      return rc;
    const bool grow64 = promote64();
    std::vector<EngineInst> keep;
    std::vector<int> keep_orig;
    for (size_t j = 0; j < inst_.size(); j++)
    {
      const EngineInst & e = inst_[j];
      bool done = true;
      for (int r = 0; r < e.nrhs && done; r++)
        done = host_[e.slot0 + r].phase == PH_DONE;
      if (done)
        for (int r = 0; r < e.nrhs; r++)
          stats[all[orig[j]].slot0 + r] = stats_of(host_[e.slot0 + r]);
      else
      {
        keep.push_back(e);
        keep_orig.push_back(orig[j]);
      }
    }
    if (keep.empty())
      break;
    if (++rounds > max_rounds)
      return fail(HC_EHIP, "solve: rounds exceeded the bound without every slot finishing");
    chunk = std::min(chunk * 2, 64);
    if ((int)(inst_.size() - keep.size()) * 4 >= (int)inst_.size())
    {
      const TileCandidate * c64[PMAX + 1], *c32[PMAX + 1];
      for (int p = 0; p <= PMAX; p++)
      {
        c64[p] = cands64_[p];
        c32[p] = cands32_[p];
      }
      if ((rc = set(keep, mode_ | (aux_ ? HC_SOLVE_AUX : 0), rhs_block_, force_batched_, c64, c32, st)))
        return rc;
      orig.swap(keep_orig);
    }
    else if (grow64)
    {
      const std::vector<EngineInst> cur = inst_;
      const TileCandidate * c64[PMAX + 1], *c32[PMAX + 1];
      for (int p = 0; p <= PMAX; p++)
      {
        c64[p] = cands64_[p];
        c32[p] = cands32_[p];
      }
      if ((rc = set(cur, mode_ | (aux_ ? HC_SOLVE_AUX : 0), rhs_block_, force_batched_, c64, c32, st)))
        return rc;
    }
  }
  HC_TRY_HIP(hipEventRecord(e1, st));
  HC_TRY_HIP(hipEventSynchronize(e1));
  float t = 0;
  HC_TRY_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms)
    *ms = t;
  // the full instance list again (the caller may solve the same batch again)
  const TileCandidate * c64[PMAX + 1], *c32[PMAX + 1];
  for (int p = 0; p <= PMAX; p++)
  {
    c64[p] = cands64_[p];
    c32[p] = cands32_[p];
  }
  return set(all, mode_ | (aux_ ? HC_SOLVE_AUX : 0), rhs_block_, force_batched_, c64, c32, st);
}

} // namespace hc

// ---------------------------------------------------------------------------
// operator workspace
// ---------------------------------------------------------------------------
int
hc_op_s::ensure_ws(int nrhs, bool need32)
{
  if (ws.nrhs >= nrhs && (ws.has32 || !need32) && ws_mem.base)
    return HC_OK;
  const int k = std::max(nrhs, ws.nrhs);
  const bool n32 = need32 || ws.has32;
  const size_t sz = hc::Workspace::bytes_for(N, k, n32);
  bytes -= (int64_t)ws_mem.bytes;
  ws = hc::Workspace{};
  int rc = ws_mem.alloc(sz);
  if (rc)
    return rc;
  bytes += (int64_t)sz;
  ws.carve(ws_mem.base, N, k, n32);
  return HC_OK;
}

hc_op_s::~hc_op_s()
{
  delete engine;
  hc::mg_free(mg);
  if (av_host)
    (void)hipHostFree(av_host);
  if (dfs_op)
    hc_op_destroy(dfs_op);
  for (void * v : allocs)
    (void)hipFree(v);
}

namespace
{

void
fold_stats(const std::vector<hc::SlotStats> & s, int begin, int count, double ms, int mode, hc_solve_stats * out)
{
  hc_solve_stats r{};
  r.converged = 1;
  for (int i = begin; i < begin + count; i++)
  {
    r.iterations = std::max(r.iterations, s[i].inner);
    r.outer_iterations = std::max(r.outer_iterations, s[i].outer);
    r.rel_res = std::max(r.rel_res, s[i].rel_res);
    if (!s[i].converged)
      r.converged = 0;
    if (s[i].status && !r.status)
      r.status = s[i].status;
    if (hc::prec_mode(mode) == HC_SOLVE_FP32_IR)
      r.fp64_fallbacks += s[i].fp64;
  }
  if (r.status == 0 && !r.converged)
    r.status = HC_ENOTCONV;
  r.solve_ms = ms;
  *out = r;
}

int
status_rc(const hc_solve_stats & s, const char * who)
{
  if (s.status == HC_ENAN)
    return hc::fail(HC_ENAN, std::string(who) + ": NaN/Inf met in a reduction");
  if (!s.converged)
  {
    char buf[200];
    std::snprintf(buf, sizeof(buf), "%s: not converged after %d iterations, true relative residual %.3e", who,
                  s.iterations, s.rel_res);
    return hc::fail(HC_ENOTCONV, buf);
  }
  return HC_OK;
}

} // namespace

// ---------------------------------------------------------------------------
// batches of operator instances
// ---------------------------------------------------------------------------
struct hc_batch_s
{
  struct Entry
  {
    hc_op_t op;
    const double * b;
    double * x;
    int nrhs;
    double rtol;
    int maxit;
  };
  std::vector<Entry> entries;
  std::vector<std::unique_ptr<hc::DevBlock>> ws_mem;
  std::vector<hc::Workspace> ws;
  hc::Engine engine;
  int rhs_block = 0;
  bool dirty = true;
  int mode = -1;
  std::vector<hc::SlotStats> stats;
};

static int
total_rhs(const hc_batch_s * bt)
{
  int n = 0;
  for (const auto & e : bt->entries)
    n += e.nrhs;
  return n;
}

extern "C" {

int
hc_engine_set_graphs(int enable)
try
{
  hc::g_graphs.store(enable ? 1 : 0);
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_pcg_solve(hc_op_t op, const double * b, double * x, int nrhs, double rtol, int maxit, int mode,
             hc_solve_stats * stats, void * stream)
try
{
  if (!op || !b || !x)
    return hc::fail(HC_EINVAL, "hc_pcg_solve: null argument");
  if (!op->setup_done)
    return hc::fail(HC_EINVAL, "hc_pcg_solve: hc_op_setup has not been called");
  if (nrhs < 1)
    return hc::fail(HC_EINVAL, "hc_pcg_solve: nrhs must be >= 1");
  if (!(rtol > 0) || maxit < 1)
    return hc::fail(HC_EINVAL, "hc_pcg_solve: rtol must be > 0 and maxit >= 1");
  if (mode & HC_SOLVE_MG)
  {
    // geometric multigrid V-cycle preconditioner, FP64 (HC_SOLVE_AUX implied:
    // the hybrid smoother carries the gradient-space correction)
    if (hc::prec_mode(mode & ~HC_SOLVE_MG) != HC_SOLVE_FP64 || (mode & ~(HC_SOLVE_MG | HC_SOLVE_AUX)))
      return hc::fail(HC_EINVAL, "hc_pcg_solve: HC_SOLVE_MG runs in FP64 (mode HC_SOLVE_FP64 | HC_SOLVE_MG)");
    if (nrhs > 1)
    {
      // all right-hand sides through one V-cycle per iteration
      bool handled = false;
      const int rc = hc::mg_pcg_solve_multi(op, b, x, nrhs, rtol, maxit, stats, (hipStream_t)stream, handled);
      if (handled || rc)
        return rc;
    }
    hc_solve_stats all{};
    all.converged = 1;
    int rc_all = HC_OK;
    for (int r = 0; r < nrhs; r++)
    {
      hc_solve_stats s{};
      const int rc = hc::mg_pcg_solve(op, b + (int64_t)r * op->N, x + (int64_t)r * op->N, rtol, maxit, &s,
                                      (hipStream_t)stream);
      if (rc && rc != HC_ENOTCONV && rc != HC_ENAN)
        return rc;
      all.iterations = std::max(all.iterations, s.iterations);
      all.outer_iterations = std::max(all.outer_iterations, s.outer_iterations);
      all.rel_res = std::max(all.rel_res, s.rel_res);
      all.solve_ms += s.solve_ms;
      if (!s.converged)
        all.converged = 0;
      if (rc && !rc_all)
      {
        rc_all = rc;
        all.status = rc;
      }
    }
    if (stats)
      *stats = all;
    return rc_all ? hc::fail(rc_all, std::string("hc_pcg_solve: ") + hc_last_error()) : HC_OK;
  }
  const int pm = hc::prec_mode(mode);
  if ((pm != HC_SOLVE_FP64 && pm != HC_SOLVE_FP32_IR) || (mode & ~(HC_SOLVE_AUX | 0xF)))
    return hc::fail(HC_EINVAL, "hc_pcg_solve: unknown mode");
  if (pm == HC_SOLVE_FP32_IR && !op->qd32)
    return hc::fail(HC_EINVAL, "hc_pcg_solve: FP32 refinement needs an operator created with HC_PREC_FP32");
  int rc = op->ensure_ws(nrhs, pm == HC_SOLVE_FP32_IR);
  if (rc)
    return rc;
  if (!op->engine)
    op->engine = new hc::Engine();
  std::vector<hc::EngineInst> inst{{op, nrhs, b, x, rtol, maxit, &op->ws, 0}};
  if ((rc = op->engine->set(inst, mode, op->rhs_block, false, nullptr, nullptr, (hipStream_t)stream)))
    return rc;
  std::vector<hc::SlotStats> s;
  double ms = 0;
  if ((rc = op->engine->solve(mode, (hipStream_t)stream, s, &ms)))
    return rc;
  hc_solve_stats out{};
  fold_stats(s, 0, nrhs, ms, mode, &out);
  if (stats)
    *stats = out;
  return status_rc(out, "hc_pcg_solve");
}
catch (...)
{
  return hc::fail_exception();
}

// HC_SOLVE_MG batches: the multigrid PCG of every entry (its right-hand sides
// through one V-cycle per iteration; the gradient-space PCG for an operator
// whose mesh has no hierarchy), run by MG_BATCH_WORKERS host threads with
// their own streams; a worker takes the next entry whose operator no other
// worker is using, so solves of different operators overlap on the device.
// Each result is that of the same solve run alone.  stats[i].solve_ms: the
// batch's wall time.
constexpr int MG_BATCH_WORKERS = 4;

static int
mg_batch_solve(hc_batch_t bt, hc_solve_stats * stats, hipStream_t caller)
{
  HC_TRY_HIP(hipStreamSynchronize(caller)); // b and x as the caller's stream left them
  const size_t ntask = bt->entries.size(); // one task per entry
  std::vector<hc_solve_stats> ts(ntask);
  std::vector<char> done(ntask, 0);
  std::mutex mu;
  std::condition_variable cv;
  std::set<hc_op_t> busy;
  size_t next_free = 0; // tasks before it are all taken
  int err = HC_OK;
  std::string err_msg;
  const auto t0 = std::chrono::steady_clock::now();
  auto worker = [&]() {
    (void)hipSetDevice(0);
    hipStream_t st = nullptr;
    if (hipStreamCreate(&st) != hipSuccess)
    {
      std::lock_guard<std::mutex> lk(mu);
      if (!err)
      {
        err = hc::fail(HC_EHIP, "hc_batch_solve: hipStreamCreate failed");
        err_msg = hc_last_error();
      }
      cv.notify_all();
      return;
    }
    for (;;)
    {
      int k = -1;
      {
        std::unique_lock<std::mutex> lk(mu);
        for (;;)
        {
          if (err || next_free >= ntask)
            break;
          for (size_t j = next_free; j < ntask; j++)
            if (!done[j] && !busy.count(bt->entries[j].op))
            {
              k = (int)j;
              break;
            }
          if (k >= 0)
            break;
          cv.wait(lk);
        }
        if (k < 0)
          break;
        done[k] = 1; // taken
        busy.insert(bt->entries[k].op);
        while (next_free < ntask && done[next_free])
          next_free++;
      }
      const auto & e = bt->entries[k];
      hc_op_t op = e.op;
      int nlev = 0;
      int64_t nc = 0;
      hc_solve_stats s{};
      int rc = HC_OK;
      if (hc::mg_levels(op, st, &nlev, &nc) != HC_OK)
        rc = hc_pcg_solve(op, e.b, e.x, e.nrhs, e.rtol, e.maxit, HC_SOLVE_FP64 | HC_SOLVE_AUX, &s, st);
      else
      {
        bool handled = false;
        if (e.nrhs > 1)
          rc = hc::mg_pcg_solve_multi(op, e.b, e.x, e.nrhs, e.rtol, e.maxit, &s, st, handled);
        if (!handled && !rc)
        {
          // one right-hand side: one solve
          s = hc_solve_stats{};
          s.converged = 1;
          for (int r = 0; r < e.nrhs && !rc; r++)
          {
            hc_solve_stats s1{};
            const int64_t off = (int64_t)r * op->N;
            rc = hc::mg_pcg_solve(op, e.b + off, e.x + off, e.rtol, e.maxit, &s1, st, false);
            if (rc == HC_ENOTCONV || rc == HC_ENAN)
              rc = HC_OK;
            s.iterations = std::max(s.iterations, s1.iterations);
            s.outer_iterations = std::max(s.outer_iterations, s1.outer_iterations);
            s.rel_res = std::max(s.rel_res, s1.rel_res);
            if (!s1.converged)
              s.converged = 0;
            if (s1.status && !s.status)
              s.status = s1.status;
          }
        }
      }
      if (rc == HC_ENOTCONV || rc == HC_ENAN)
        rc = HC_OK; // reported in the statistics
      if (!rc && hipStreamSynchronize(st) != hipSuccess)
        rc = hc::fail(HC_EHIP, "hc_batch_solve: stream synchronisation failed");
      {
        std::lock_guard<std::mutex> lk(mu);
        ts[k] = s;
        busy.erase(op);
        if (rc && !err)
        {
          err = rc;
          err_msg = hc_last_error();
        }
      }
      cv.notify_all();
      if (rc)
        break;
    }
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    cv.notify_all();
  };
  // no C++ exception may cross the C ABI: a worker body reports what it
  // throws as an error, and when no thread can be started the calling thread
  // runs the worker itself
  auto guarded = [&]() {
    try
    {
      worker();
    }
    catch (const std::exception & ex)
    {
      std::lock_guard<std::mutex> lk(mu);
      if (!err)
      {
        err = hc::fail(HC_EHIP, std::string("hc_batch_solve: ") + ex.what());
        err_msg = hc_last_error();
      }
      cv.notify_all();
    }
  };
  std::vector<std::thread> pool; // no more workers (and streams) than tasks
  try
  {
    for (size_t w = 0; w < std::min<size_t>(MG_BATCH_WORKERS, ntask); w++)
      pool.emplace_back(guarded);
  }
  catch (const std::exception &)
  {
  }
  if (pool.empty() && ntask > 0)
    guarded();
  for (auto & t : pool)
    t.join();
  if (err)
    return hc::fail(err, "hc_batch_solve: " + err_msg);
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  hc_solve_stats all{};
  all.converged = 1;
  for (size_t i = 0; i < ntask; i++)
  {
    hc_solve_stats s = ts[i];
    s.solve_ms = ms;
    if (stats)
      stats[i] = s;
    all.iterations = std::max(all.iterations, s.iterations);
    all.rel_res = std::max(all.rel_res, s.rel_res);
    if (!s.converged)
      all.converged = 0;
    if (s.status && !all.status)
      all.status = s.status;
  }
  all.solve_ms = ms;
  return status_rc(all, "hc_batch_solve");
}

int
hc_batch_create(hc_batch_t * out)
try
{
  if (!out)
    return hc::fail(HC_EINVAL, "hc_batch_create: null argument");
  *out = new hc_batch_s();
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_batch_add(hc_batch_t bt, hc_op_t op, const double * b, double * x, int nrhs, double rtol, int maxit)
try
{
  if (!bt || !op || !b || !x)
    return hc::fail(HC_EINVAL, "hc_batch_add: null argument");
  if (nrhs < 1 || !(rtol > 0) || maxit < 1)
    return hc::fail(HC_EINVAL, "hc_batch_add: nrhs >= 1, rtol > 0 and maxit >= 1 are required");
  if (!op->setup_done)
    return hc::fail(HC_EINVAL, "hc_batch_add: hc_op_setup has not been called");
  bt->entries.push_back({op, b, x, nrhs, rtol, maxit});
  bt->dirty = true;
  return (int)bt->entries.size() - 1;
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_batch_set_rhs_block(hc_batch_t bt, int rb)
try
{
  if (!bt || rb < 0)
    return hc::fail(HC_EINVAL, "hc_batch_set_rhs_block: bad argument");
  bt->rhs_block = rb;
  bt->dirty = true;
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_batch_solve(hc_batch_t bt, int mode, hc_solve_stats * stats, void * stream)
try
{
  if (!bt)
    return hc::fail(HC_EINVAL, "hc_batch_solve: null handle");
  if (bt->entries.empty())
    return hc::fail(HC_EINVAL, "hc_batch_solve: empty batch");
  if (mode & HC_SOLVE_MG)
  {
    if (hc::prec_mode(mode) != HC_SOLVE_FP64 || (mode & ~(HC_SOLVE_MG | HC_SOLVE_AUX)))
      return hc::fail(HC_EINVAL, "hc_batch_solve: HC_SOLVE_MG runs in FP64 (mode HC_SOLVE_FP64 | HC_SOLVE_MG)");
    return mg_batch_solve(bt, stats, (hipStream_t)stream);
  }
  const int pm = hc::prec_mode(mode);
  if ((pm != HC_SOLVE_FP64 && pm != HC_SOLVE_FP32_IR) || (mode & ~(HC_SOLVE_AUX | 0xF)))
    return hc::fail(HC_EINVAL, "hc_batch_solve: unknown mode");
  const bool ir = pm == HC_SOLVE_FP32_IR;
  int rc;
  if (bt->dirty || bt->mode != mode)
  {
    // one workspace per entry
    bt->ws.resize(bt->entries.size());
    bt->ws_mem.resize(bt->entries.size());
    for (size_t i = 0; i < bt->entries.size(); i++)
    {
      const auto & e = bt->entries[i];
      hc::Workspace & w = bt->ws[i];
      if (!bt->ws_mem[i])
        bt->ws_mem[i].reset(new hc::DevBlock());
      if (w.N != e.op->N || w.nrhs < e.nrhs || (ir && !w.has32))
      {
        const size_t sz = hc::Workspace::bytes_for(e.op->N, e.nrhs, ir);
        if ((rc = bt->ws_mem[i]->alloc(sz)))
          return rc;
        w.carve(bt->ws_mem[i]->base, e.op->N, e.nrhs, ir);
      }
    }
    std::vector<hc::EngineInst> inst;
    for (size_t i = 0; i < bt->entries.size(); i++)
    {
      const auto & e = bt->entries[i];
      inst.push_back({e.op, e.nrhs, e.b, e.x, e.rtol, e.maxit, &bt->ws[i], 0});
    }
    if ((rc = bt->engine.set(inst, mode, bt->rhs_block, true, nullptr, nullptr, (hipStream_t)stream)))
      return rc;
    bt->dirty = false;
    bt->mode = mode;
  }
  double ms = 0;
  if ((rc = bt->engine.solve(mode, (hipStream_t)stream, bt->stats, &ms)))
  {
    // a failed solve may leave the engine on a compacted instance list: the
    // next call rebuilds the tables from every entry
    bt->dirty = true;
    return rc;
  }
  if (bt->stats.size() < (size_t)total_rhs(bt))
  {
    bt->dirty = true;
    return hc::fail(HC_EHIP, "hc_batch_solve: engine returned fewer slot statistics than right-hand sides");
  }
  int pos = 0;
  hc_solve_stats all{};
  all.converged = 1;
  for (size_t i = 0; i < bt->entries.size(); i++)
  {
    hc_solve_stats s{};
    fold_stats(bt->stats, pos, bt->entries[i].nrhs, ms, mode, &s);
    pos += bt->entries[i].nrhs;
    if (stats)
      stats[i] = s;
    all.iterations = std::max(all.iterations, s.iterations);
    all.outer_iterations = std::max(all.outer_iterations, s.outer_iterations);
    all.rel_res = std::max(all.rel_res, s.rel_res);
    if (!s.converged)
      all.converged = 0;
    if (s.status && !all.status)
      all.status = s.status;
    all.fp64_fallbacks += s.fp64_fallbacks;
  }
  all.solve_ms = ms;
  return status_rc(all, "hc_batch_solve");
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_batch_destroy(hc_batch_t bt)
try
{
  if (!bt)
    return hc::fail(HC_EINVAL, "hc_batch_destroy: null handle");
  (void)hipDeviceSynchronize();
  delete bt;
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

} // extern "C"
