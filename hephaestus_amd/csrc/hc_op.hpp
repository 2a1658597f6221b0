// Library-internal state of an operator handle (hc_op_t) and of the batched
// solve engine shared by the C ABI entry points.
#pragma once

#include "../../include/hc_hcurl.h"
#include "hc_kernels.hpp"

#include <string>
#include <vector>

namespace hc
{

int fail(int code, const std::string & msg);
// The error code of a C++ exception caught at the C ABI (HC_ENOMEM for an
// allocation failure, else HC_EHIP), with its message as the last error.
int fail_exception();

#define HC_TRY_HIP(expr)                                                                                        \
  do                                                                                                             \
  {                                                                                                              \
    hipError_t _e = (expr);                                                                                      \
    if (_e != hipSuccess)                                                                                        \
      return ::hc::fail(HC_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e));                             \
  } while (0)

// Device memory block owned by one object (one hipMalloc, carved into
// 256-byte aligned pieces).
struct DevBlock
{
  void * base = nullptr;
  size_t bytes = 0;
  DevBlock() = default;
  DevBlock(const DevBlock &) = delete;
  DevBlock & operator=(const DevBlock &) = delete;
  ~DevBlock() { release(); }
  void release()
  {
    if (base)
      (void)hipFree(base);
    base = nullptr;
    bytes = 0;
  }
  int alloc(size_t sz)
  {
    release();
    if (sz == 0)
      sz = 256;
    if (hipMalloc(&base, sz) != hipSuccess)
    {
      base = nullptr;
      return fail(HC_ENOMEM, "hipMalloc of " + std::to_string(sz) + " bytes failed");
    }
    bytes = sz;
    return HC_OK;
  }
};

// Precision part of a solve mode (HC_SOLVE_FP64 / HC_SOLVE_FP32_IR).
inline int
prec_mode(int mode)
{
  return mode & ~(HC_SOLVE_AUX | HC_SOLVE_MG);
}

inline size_t
align256(size_t v)
{
  return (v + 255) & ~(size_t)255;
}

// Work vectors of the solves of nrhs right-hand sides of one operator
// (N DOFs each, rhs r at offset r*N), the tile-face partials of the apply
// (4 N per rhs and precision) and the node vectors of the gradient-space
// correction (nn_cap(N) per rhs and precision, node vector r at r*nn_cap(N)).
struct Workspace
{
  int64_t N = 0;
  int64_t NNC = 0;
  int nrhs = 0;
  bool has32 = false;
  double *r64 = nullptr, *q64 = nullptr, *p64 = nullptr, *part64 = nullptr, *t64 = nullptr;
  float *b32 = nullptr, *x32 = nullptr, *r32 = nullptr, *p32 = nullptr, *q32 = nullptr, *part32 = nullptr,
        *t32 = nullptr;

  // An upper bound of the H1 node count of any structured ND space with N
  // DOFs: with L_a = n_a p >= 1, N = sum_c L_c prod_{a != c} (L_a + 1)
  // >= 3/2 prod_a (L_a + 1).
  static int64_t nn_cap(int64_t N) { return 2 * N / 3 + 1; }
  static size_t bytes_for(int64_t N, int nrhs, bool need32)
  {
    const size_t v64 = align256(sizeof(double) * (size_t)N * nrhs);
    const size_t v32 = align256(sizeof(float) * (size_t)N * nrhs);
    const size_t t64 = align256(sizeof(double) * (size_t)nn_cap(N) * nrhs);
    const size_t t32 = align256(sizeof(float) * (size_t)nn_cap(N) * nrhs);
    size_t b = 3 * v64 + 4 * v64 + t64;
    if (need32)
      b += 5 * v32 + 4 * v32 + t32;
    return b;
  }
  // carve the pieces out of [base, base + bytes_for(...))
  void carve(void * base, int64_t n, int k, bool need32)
  {
    N = n;
    NNC = nn_cap(n);
    nrhs = k;
    has32 = need32;
    char * p = (char *)base;
    const size_t v64 = align256(sizeof(double) * (size_t)N * nrhs);
    const size_t v32 = align256(sizeof(float) * (size_t)N * nrhs);
    r64 = (double *)p;
    p += v64;
    q64 = (double *)p;
    p += v64;
    p64 = (double *)p;
    p += v64;
    part64 = (double *)p;
    p += 4 * v64;
    t64 = (double *)p;
    p += align256(sizeof(double) * (size_t)NNC * nrhs);
    if (need32)
    {
      b32 = (float *)p;
      p += v32;
      x32 = (float *)p;
      p += v32;
      r32 = (float *)p;
      p += v32;
      p32 = (float *)p;
      p += v32;
      q32 = (float *)p;
      p += v32;
      part32 = (float *)p;
      p += 4 * v32;
      t32 = (float *)p;
    }
    else
      b32 = x32 = r32 = p32 = q32 = part32 = t32 = nullptr;
  }
};

// Apply kernel variant of one precision: geometry with the tile decomposition,
// the tile-face assembly segments and the compiled candidate.
struct ApplyPlan
{
  Geom g{};
  PlaneSegs segs{};
  const TileCandidate * cand = nullptr;
};

class Engine;
struct MgHier;
void mg_free(MgHier * h);

} // namespace hc

struct hc_op_s
{
  hc::Geom g{}; // base geometry (tile fields follow plan64)
  hc::ApplyPlan plan64, plan32;
  hc::Tables<double> tb64{};
  hc::Tables<float> tb32{};
  int p = 0;
  int precision = 0;
  int64_t E = 0, N = 0, n_ess = 0;
  bool box = true;
  double h[3] = {0, 0, 0};
  double x0[3] = {0, 0, 0};
  std::vector<double> nu, sig;
  bool coef_dirty = true;
  bool setup_done = false;
  double last_dt = 0.0;
  int64_t bytes = 0;
  // device
  double * d_vert = nullptr;
  double * d_nu = nullptr;
  double * d_sig = nullptr;
  double * qd64 = nullptr;
  float * qd32 = nullptr;
  // element factors of a box mesh (dt nu_e, sigma_e per element) and the axis
  // constants: the apply reads them instead of qd64/qd32 while ef_mode is 1
  double * qc64 = nullptr;
  float * qc32 = nullptr;
  double ck[3] = {0, 0, 0}, mk[3] = {0, 0, 0};
  int ef_mode = 1;
  hc::ElemFactors ef(bool f32) const
  {
    hc::ElemFactors f;
    if (ef_mode && qc64)
    {
      f.qc = f32 ? (const void *)qc32 : (const void *)qc64;
      for (int a = 0; a < 3; a++)
      {
        f.ck[a] = ck[a];
        f.mk[a] = mk[a];
      }
    }
    return f;
  }
  double * diag = nullptr;
  double * dinv64 = nullptr;
  float * dinv32 = nullptr;
  // gradient-space correction: node diagonal inverse (computed at setup)
  hc::AuxGeom ag{};
  double * ginv64 = nullptr;
  float * ginv32 = nullptr;
  // min over elements of sigma h_eff^2 / nu (set at setup); FP32 + FP64
  // refinement runs its inner solves in FP32 only when this is at least
  // FP32_MIN_TERM_RATIO, in FP64 otherwise
  double term_ratio = 0.0;
  double * d_ratio = nullptr; // TERM_RATIO_BLOCKS block minima
  bool fp32_inner_ok() const { return term_ratio >= hc::FP32_MIN_TERM_RATIO; }
  hc::ApplyDesc * desc_dev = nullptr; // staging of the batched apply descriptor
  int rhs_block = 0;                    // right-hand sides per work item (0: automatic)
  std::vector<void *> allocs;
  // work vectors of hc_op_apply (nrhs > 1) and hc_pcg_solve, grown on demand
  hc::DevBlock ws_mem;
  hc::Workspace ws;
  hc::Engine * engine = nullptr;
  // geometric multigrid hierarchy of HC_SOLVE_MG (hc_mg.hip), built on first
  // use and rebuilt when setup_version (counts the re-assemblies) moved on
  hc::MgHier * mg = nullptr;
  // work vectors of hc_op_av_solve and hc_op_div_free_source (hc_av.hip),
  // allocated on first use and kept; the unit-mass factors of the latter
  // depend on the mesh alone and are computed once
  hc::DevBlock av_mem, dfs_mem;
  double * av_host = nullptr; // pinned copy of the solver state
  bool dfs_ready = false;
  // unit-mass operator on the same box mesh (nu = 0, sigma = 1) whose node
  // hierarchy preconditions the Helmholtz projection of hc_op_div_free_source
  hc_op_t dfs_op = nullptr;
  int dfs_mg = 0; // 0 not tried, 1 available, -1 not available
  int setup_version = 0;

  template <typename T>
  int alloc(T *& ptr, int64_t count)
  {
    void * v = nullptr;
    const size_t sz = sizeof(T) * (size_t)(count > 0 ? count : 1);
    if (hipMalloc(&v, sz) != hipSuccess)
      return hc::fail(HC_ENOMEM, "hipMalloc of " + std::to_string(sz) + " bytes failed");
    ptr = (T *)v;
    allocs.push_back(v);
    bytes += (int64_t)sz;
    return HC_OK;
  }
  // workspace for at least nrhs right-hand sides (and FP32 vectors if need32)
  int ensure_ws(int nrhs, bool need32);
  ~hc_op_s();
};

namespace hc
{

// Batched Jacobi-PCG / FP32+FP64 refinement over the right-hand sides of one
// or more operator instances (see hc_solve.hip).
struct EngineInst
{
  hc_op_t op;
  int nrhs;
  const double * b; // nrhs * N, rhs r at r*N
  double * x;
  double rtol;
  int maxit;
  const Workspace * ws;
  int slot0;
  // per-rhs device states and reduction partials (nrhs each); null: owned by
  // the engine (then they do not survive a later set())
  PcgState * in = nullptr;
  PcgState * out = nullptr;
  double * part = nullptr; // nrhs * 4 * RED_BLOCKS
  // FP32 + refinement: the instance is in the FP64 iteration group (always
  // when its operator does not admit FP32 inner solves; otherwise once one of
  // its slots continues in FP64 -- Engine::promote64)
  bool fp64 = false;
};

struct SlotStats
{
  int inner;     // inner PCG iterations (summed over refinement steps / restarts)
  int outer;     // refinement steps or PCG runs
  int status;
  int converged;
  double rel_res; // final FP64 true relative residual
  int fp64;       // finished in FP64
};

// Initial values of the states of a slot that starts a solve in `mode`
// (fp32_ok: the operator admits FP32 inner solves, hc_op_s::fp32_inner_ok).
void initial_states(PcgState & in, PcgState & out, int mode, bool fp32_ok);

class Engine
{
public:
  Engine() = default;
  ~Engine();
  Engine(const Engine &) = delete;
  Engine & operator=(const Engine &) = delete;

  // Replace the instance list and rebuild the device tables (no work of the
  // engine may be in flight).  Per order the
  // tile variant is cands64/cands32[order] when given, else the plan of the
  // first instance of that order.  rhs_block: right-hand sides per work item
  // of the batched apply (0: default).
  // mode: HC_SOLVE_FP64 or HC_SOLVE_FP32_IR, optionally | HC_SOLVE_AUX
  // st: the stream of the solves that follow; the table uploads (and the read
  // of the iteration counter) are ordered on it, not on the null stream, so
  // that engines on other streams are not synchronised with it.
  int set(const std::vector<EngineInst> & inst, int mode, int rhs_block = 0, bool force_batched = false,
          const TileCandidate * const * cands64 = nullptr, const TileCandidate * const * cands32 = nullptr,
          hipStream_t st = nullptr);
  // One round: phase moves, FP64 residuals and refinement updates, inner
  // starts, then `iters` PCG iterations of every slot in an inner solve.
  // graph: replay a captured hipGraph of this round's launches (captured on
  // first use per (iters, running precisions) and table set; st must not be
  // the null stream).
  int round(int iters, hipStream_t st, bool graph = false);
  // Copies the outer states of all slots to host_states() (synchronises).
  int gather(hipStream_t st);
  const PcgState * host_states() const { return host_; }
  // Zero the iteration counter (before the first round of a fresh batch).
  int reset_counter(hipStream_t st);
  // After gather(): instances with a slot that continues in FP64 and that are
  // not yet in the FP64 iteration group join it; true if the tables must be
  // rebuilt (set() with instances()) before the next round.
  bool promote64();
  const std::vector<EngineInst> & instances() const { return inst_; }
  // Batch solve from fresh states to completion; per-slot statistics.
  int solve(int mode, hipStream_t st, std::vector<SlotStats> & stats, double * ms);
  int nslots() const { return nslots_; }
  int max_outer() const;
  static SlotStats stats_of(const PcgState & out);

private:
  int round_launches(int iters, bool run32, bool run64, hipStream_t st);
  void drop_graphs();
  struct ApplyGroup
  {
    int order;
    bool batched;      // descriptor table (BATCH kernel) or by-value single instance
    ApplyDesc single;  // by-value descriptor (batched == false)
    const ApplyDesc * descs; // device
    int ndesc, nwork, npwork;
    int grouped; // bit 0: some instance groups rb > 1 right-hand sides per work item;
                 // LAUNCH_EF: every instance reads element factors
    const TileCandidate * cand;
    const void * tables; // Tables<double> or Tables<float>
    int prec;
  };
  enum Kind
  {
    K_PQ32 = 0, // A p -> q in FP32, slots running FP32 iterations
    K_PQ64,     // A p -> q in FP64, slots running FP64 iterations
    K_XQ,       // A x -> q64 (FP64), slots in PH_RESID
    K_COUNT
  };
  int apply(Kind k, hipStream_t st);

  std::vector<EngineInst> inst_;
  int nslots_ = 0;
  int mode_ = 0;     // precision mode (without HC_SOLVE_AUX)
  bool aux_ = false; // gradient-space correction
  int rhs_block_ = 0;
  bool force_batched_ = false;
  const TileCandidate * cands64_[PMAX + 1] = {};
  const TileCandidate * cands32_[PMAX + 1] = {};
  DevBlock mem_;
  DevBlock own_; // states of instances that came without their own
  SlotTable table_{};
  int * git_ = nullptr; // iteration counters of the FP32 and FP64 groups
  PcgState * gathered_ = nullptr; // device, nslots
  PcgState * host_ = nullptr;     // pinned, nslots
  int host_cap_ = 0;
  bool any32_ = false, any64_ = true; // precisions with running slots
  std::vector<ApplyGroup> groups_[K_COUNT];
  // captured rounds of the current tables: key = 4 iters + 2 run32 + run64
  std::vector<std::pair<int, hipGraphExec_t>> graphs_;
  hipStream_t own_st_ = nullptr; // solves issued on the null stream run here
};

// Resolves the current device's address of the static scratch of
// hc_op_apply_noalloc (hc_noalloc.hip); hc_op_create calls it once per operator.
int noalloc_setup();

// Geometry of op with the tile decomposition of candidate c, and the
// tile-face assembly segments of that decomposition.
void geom_for(const hc_op_s * op, const TileCandidate * c, Geom & g, PlaneSegs & segs);

// PCG with the geometric multigrid preconditioner (hc_mg.hip), FP64, one
// right-hand side; mg_levels reports the hierarchy (built if needed).
// allow_graph = false: never capture the V-cycle (callers on worker threads)
int mg_pcg_solve(hc_op_t op, const double * b, double * x, double rtol, int maxit, hc_solve_stats * stats,
                 hipStream_t st, bool allow_graph = true);
// r = b - A x (FP64, nrhs right-hand sides at stride N) with the apply
// kernels forming the difference as they store (no separate pass over r).
int op_residual(hc_op_t op, const double * b, const double * x, double * r, hipStream_t st, int nrhs = 1);
int mg_levels(hc_op_t op, hipStream_t st, int * nlev, int64_t * ncoarse);
// k right-hand sides at once (handled = false: hierarchy with an inexact
// coarsest level, not solved here)
int mg_pcg_solve_multi(hc_op_t op, const double * B, double * X, int k, double rtol, int maxit,
                       hc_solve_stats * stats, hipStream_t st, bool & handled);
// AV block preconditioner (hc_mg.hip): availability (builds the ND and node
// hierarchies) and application za = V-cycle(ra), zv = node V-cycle(rv).
bool mg_av_ready(hc_op_t op, hipStream_t st);
int mg_av_precond(hc_op_t op, const double * ra, const double * rv, double * za, double * zv, hipStream_t st);
// z = node V-cycle of G^T M G on r (op's node hierarchy, after mg_av_ready)
int mg_node_precond(hc_op_t op, const double * r, double * z, hipStream_t st);

// Graph capture of the engine's steady-state rounds (process-wide switch,
// off by default; hc_engine_set_graphs).
bool graphs_enabled();

} // namespace hc
