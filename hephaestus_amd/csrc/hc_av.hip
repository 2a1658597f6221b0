// Solve of the AV formulation's block system (AVEquationSystem,
// src/formulations/AV/av_formulation.cpp:150-196) on the device.
//
// The block operator (hc_op_av_apply) with the weak-divergence block in the
// form the reference states, (sigma u, grad V')
// (src/kernels/vector_fe_weak_divergence_kernel.hpp:8), is symmetric positive
// semi-definite: its energy is dt (curl a, nu curl a) + (a0 + G v, sigma
// (a0 + G v)), zero exactly on the gauge fields (-G phi, phi).  The reference
// solves it with HypreGMRES + BoomerAMG (problem_builder_base.cpp:206-218);
// here it is preconditioned CG with a block-diagonal preconditioner: the
// multigrid V-cycles of hc_mg.hip on the ND block and on the node block
// G^T M G where the mesh has the hierarchies (with exact blocks the
// preconditioned operator has the eigenvalues 0, 1 and 2 only on the free
// DOFs, since K G = 0 makes A G = M G there), else block Jacobi diag(A_AA)^-1 (1 on PEC DOFs) and
// diag(G^T M G)^-1; it converges on consistent right-hand sides.  Reductions:
// per-block partial sums on the device, summed in a fixed order, so a solve
// is bitwise reproducible.
#include "hc_aux.hpp"
#include "hc_kernels.hpp"
#include "hc_op.hpp"

#include <cmath>
#include <initializer_list>
#include <vector>

#include "hc_pcgdev.hpp"

namespace hc
{

namespace
{

// r = b - q; out[b] = block parts of (r, r)
__global__ void __launch_bounds__(AV_THREADS)
k_av_resid(int64_t n, const double * __restrict__ b, const double * __restrict__ q, double * __restrict__ out)
{
  __shared__ double sh[AV_THREADS];
  double rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * AV_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * AV_THREADS)
  {
    const double ri = b[i] - q[i];
    rr += ri * ri;
  }
  rr = av_block_sum(rr, sh);
  if (threadIdx.x == 0)
    out[blockIdx.x] = rr;
}

__global__ void
k_fill_value(int64_t n, double v, double * __restrict__ y)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = v;
}

// Pieces carved out of one device block, 256-byte aligned.
struct Carver
{
  char * p;
  double * take(int64_t n)
  {
    double * v = (double *)p;
    p += align256(sizeof(double) * (size_t)(n > 0 ? n : 1));
    return v;
  }
};

size_t
bytes_of(std::initializer_list<int64_t> counts)
{
  size_t b = 0;
  for (int64_t n : counts)
    b += align256(sizeof(double) * (size_t)(n > 0 ? n : 1));
  return b;
}

} // namespace

} // namespace hc

using hc::AV_BLOCKS;
using hc::AV_THREADS;

// DivFreeSource::Apply (src/sources/div_free_source.cpp:71-114) for the
// source linear form j = (J, phi_i):
//   g = M^-1 j                 (H(curl) mass with unit coefficient, no
//                               essential DOFs: div_free_source.cpp:76-89)
//   g <- g - G Q,  (G^T M G) Q = G^T M g   (the Helmholtz projection
//                               P(g) = g - grad Q, (grad.P(g), q) = 0 for all q
//                               in H1, div_free_source.cpp:7-17,93-106, no
//                               boundary conditions)
//   lf += M g                  (div_free_source.cpp:108-113)
// Both solves are Jacobi-preconditioned CG to rtol (the reference uses its
// default GMRES, hephaestus_solvers.hpp); the projection's singular system
// (constants) has a consistent right-hand side (G 1 = 0).  src receives the
// projected source field g, Q (optional) the potential.
extern "C" int
hc_op_div_free_source(hc_op_t op, const double * j, double * src, double * lf, double * Q, int project,
                      double rtol, int maxit, hc_solve_stats * stats, void * stream)
try
{
  if (!op || !j || !src || !lf)
    return hc::fail(HC_EINVAL, "hc_op_div_free_source: null argument");
  if (!(rtol > 0.0) || maxit < 1)
    return hc::fail(HC_EINVAL, "hc_op_div_free_source: rtol must be > 0 and maxit >= 1");
  int rc = op->ensure_ws(1, false);
  if (rc)
    return rc;
  hipStream_t st = (hipStream_t)stream;
  const int64_t N = op->N, NN = op->ag.NN, E = op->E;
  const int Qp = op->p + 1;
  const int64_t NQ = (int64_t)12 * E * Qp * Qp * Qp;
  // work vectors, kept with the operator: unit-mass factors and the two
  // diagonals (mesh only, computed once), the ND and node PCG vectors
  if (!op->dfs_mem.base)
  {
    const size_t need = hc::bytes_of({NQ, E, E, N, N, N, N, N, N, NN, NN, NN, NN, NN, NN, NN, hc::PCG_PARTS}) +
                        hc::align256(sizeof(hc::PcgDev));
    if ((rc = op->dfs_mem.alloc(need)))
      return rc;
    op->dfs_ready = false;
  }
  if (!op->av_host && hipHostMalloc((void **)&op->av_host, sizeof(hc::PcgDev)) != hipSuccess)
  {
    op->av_host = nullptr;
    return hc::fail(HC_ENOMEM, "hc_op_div_free_source: pinned host allocation failed");
  }
  hc::Carver cv{(char *)op->dfs_mem.base};
  double *qdu = cv.take(NQ), *c0 = cv.take(E), *c1 = cv.take(E), *dM = cv.take(N);
  double *r = cv.take(N), *z = cv.take(N), *p = cv.take(N), *q = cv.take(N), *t = cv.take(N);
  double *rn = cv.take(NN), *zn = cv.take(NN), *pn = cv.take(NN), *qn = cv.take(NN), *dS = cv.take(NN),
         *Qv = cv.take(NN), *bn = cv.take(NN);
  double * part = cv.take(hc::PCG_PARTS);
  hc::PcgDev * sdev = (hc::PcgDev *)cv.p;
  hc::PcgDev * shost = (hc::PcgDev *)op->av_host;
  // the projection's node Laplacian G^T M G: preconditioned by the node-space
  // V-cycle of a unit-mass operator on the same box mesh (built once, outside the timed region), else Jacobi
  if (project && op->box && op->dfs_mg == 0)
  {
    op->dfs_mg = -1;
    hc_mesh_desc md{};
    md.nx = op->g.n[0];
    md.ny = op->g.n[1];
    md.nz = op->g.n[2];
    md.vert = nullptr;
    for (int a = 0; a < 3; a++)
    {
      md.x0[a] = op->x0[a];
      md.len[a] = op->h[a] * op->g.n[a];
    }
    md.pec_faces = 0;
    std::vector<double> zero(E, 0.0), one(E, 1.0);
    if (!op->dfs_op && hc_op_create(&md, op->p, zero.data(), one.data(), HC_PREC_FP64, &op->dfs_op) != HC_OK)
      op->dfs_op = nullptr;
    if (op->dfs_op && hc_op_setup(op->dfs_op, 1.0, stream) == HC_OK && hc::mg_av_ready(op->dfs_op, st))
      op->dfs_mg = 1;
  }
  const bool node_mg = project && op->dfs_mg == 1;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return hc::fail(HC_EHIP, "hc_op_div_free_source: event creation failed");
  struct Ev
  {
    hipEvent_t a, b;
    ~Ev()
    {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } ev{e0, e1};
  (void)hipEventRecord(e0, st);
  const hc::ApplyPlan & pl = op->plan64;
  hc::Geom g0 = pl.g;
  g0.pec = 0;
  if (!op->dfs_ready)
  {
    // unit-coefficient mass factors on the operator's geometry, their
    // diagonal and the node diagonal of G^T M G
    hipLaunchKernelGGL(hc::k_fill_value, dim3(256), dim3(256), 0, st, E, 0.0, c0);
    hipLaunchKernelGGL(hc::k_fill_value, dim3(256), dim3(256), 0, st, E, 1.0, c1);
    HC_TRY_HIP(hc::pa_setup(op->g, op->tb64, op->d_vert, op->h, c0, c1, 1.0, qdu, st));
    HC_TRY_HIP(hc::diagonal(g0, op->tb64, qdu, dM, st));
    HC_TRY_HIP(hc::invert_diag(dM, dM, nullptr, N, st));
    HC_TRY_HIP(hc::aux_diag_inv(g0, op->ag, op->tb64, qdu, dS, dS, nullptr, st));
    op->dfs_ready = true;
  }
  auto mass = [&](const double * x, double * y) -> int {
    HC_TRY_HIP(hc::apply_scaled<double>(g0, pl.segs, op->tb64, qdu, hc::ElemFactors{}, x, y, op->ws.part64, 1.0, 0.0, pl.cand,
                                        op->desc_dev, st));
    return HC_OK;
  };
  int it1 = 0, it2 = 0, s1 = HC_OK, s2 = HC_OK;
  bool c1ok = false, c2ok = true;
  // g = M^-1 j
  hipLaunchKernelGGL(hc::k_fill_value, dim3(2048), dim3(256), 0, st, N, 0.0, src);
  {
    const hc::PcgBlock blk{N, dM, j, src, r, z, p, q};
    auto ap = [&](const double * x0, const double *, double * y0, double *) -> int { return mass(x0, y0); };
    if ((rc = hc::pcg_dev(1, &blk, ap, rtol, maxit, part, sdev, shost, st, it1, c1ok, s1)))
      return rc;
  }
  hipLaunchKernelGGL(hc::k_fill_value, dim3(2048), dim3(256), 0, st, NN, 0.0, Qv);
  if (project && s1 == HC_OK)
  {
    // b_n = G^T M g; S = G^T M G on the nodes, diag(S)^-1 from the unit mass factors
    if ((rc = mass(src, t)))
      return rc;
    HC_TRY_HIP(hc::grad_transpose(op->ag, t, bn, st));
    auto stiff = [&](const double * x0, const double *, double * y0, double *) -> int {
      HC_TRY_HIP(hc::grad_apply(op->ag, x0, t, st));
      int e = mass(t, q);
      if (e)
        return e;
      HC_TRY_HIP(hc::grad_transpose(op->ag, q, y0, st));
      return HC_OK;
    };
    // the stiffness apply uses t and q: the node PCG's own vectors are rn..qn
    const hc::PcgBlock blk{NN, dS, bn, Qv, rn, zn, pn, qn};
    auto npr = [&](const double * r0, const double *, double * z0, double *) -> int {
      return hc::mg_node_precond(op->dfs_op, r0, z0, st);
    };
    if ((rc = hc::pcg_dev(1, &blk, stiff, rtol, maxit, part, sdev, shost, st, it2, c2ok, s2, node_mg, npr)))
      return rc;
    // g <- g - G Q
    HC_TRY_HIP(hc::grad_apply(op->ag, Qv, t, st));
    HC_TRY_HIP(hc::axpby(N, -1.0, t, 1.0, src, st));
  }
  // lf += M g (left untouched when a solve failed)
  if (s1 == HC_OK && s2 == HC_OK)
  {
    if ((rc = mass(src, t)))
      return rc;
    HC_TRY_HIP(hc::axpby(N, 1.0, t, 1.0, lf, st));
  }
  if (Q)
    HC_TRY_HIP(hipMemcpyAsync(Q, Qv, sizeof(double) * NN, hipMemcpyDeviceToDevice, st));
  (void)hipEventRecord(e1, st);
  HC_TRY_HIP(hipStreamSynchronize(st));
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  const int status = s1 != HC_OK ? s1 : s2;
  if (stats)
  {
    *stats = hc_solve_stats{};
    stats->iterations = it1 + it2;
    stats->outer_iterations = project ? 2 : 1;
    stats->converged = (c1ok && c2ok && status == HC_OK) ? 1 : 0;
    stats->status = status;
    stats->solve_ms = ms;
  }
  if (status != HC_OK)
    return hc::fail(status, status == HC_ENAN ? "hc_op_div_free_source: NaN/Inf in a reduction"
                                              : "hc_op_div_free_source: tolerance not reached in maxit");
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

extern "C" int
hc_op_av_solve(hc_op_t op, const double * ba, const double * bv, double * a, double * v, double rtol, int maxit,
               hc_solve_stats * stats, void * stream)
try
{
  if (!op || !ba || !bv || !a || !v)
    return hc::fail(HC_EINVAL, "hc_op_av_solve: null argument");
  if (!op->setup_done)
    return hc::fail(HC_EINVAL, "hc_op_av_solve: hc_op_setup has not been called");
  if (!(rtol > 0.0) || maxit < 1)
    return hc::fail(HC_EINVAL, "hc_op_av_solve: rtol must be > 0 and maxit >= 1");
  hipStream_t st = (hipStream_t)stream;
  const int64_t N = op->N, NN = op->ag.NN;
  int rc;
  if (!op->av_mem.base &&
      (rc = op->av_mem.alloc(hc::bytes_of({N, N, N, N, NN, NN, NN, NN, NN, hc::PCG_PARTS}) +
                             hc::align256(sizeof(hc::PcgDev)))))
    return rc;
  if (!op->av_host && hipHostMalloc((void **)&op->av_host, sizeof(hc::PcgDev)) != hipSuccess)
  {
    op->av_host = nullptr;
    return hc::fail(HC_ENOMEM, "hc_op_av_solve: pinned host allocation failed");
  }
  hc::Carver cv{(char *)op->av_mem.base};
  double *ra = cv.take(N), *za = cv.take(N), *pa = cv.take(N), *qa = cv.take(N);
  double *rv = cv.take(NN), *zv = cv.take(NN), *pv = cv.take(NN), *qv = cv.take(NN), *vinv = cv.take(NN);
  double * part = cv.take(hc::PCG_PARTS);
  hc::PcgDev * sdev = (hc::PcgDev *)cv.p;
  hc::PcgDev * shost = (hc::PcgDev *)op->av_host;
  // block preconditioner: multigrid on both blocks where the mesh has the
  // hierarchies (built on first use, outside the timed region), else block Jacobi
  const bool mg = hc::mg_av_ready(op, st);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return hc::fail(HC_EHIP, "hc_op_av_solve: event creation failed");
  struct Ev
  {
    hipEvent_t a, b;
    ~Ev()
    {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } ev{e0, e1};
  (void)hipEventRecord(e0, st);
  // node preconditioner: diag(G^T M G)^-1 on every node (V has no essential nodes)
  hc::Geom g0 = op->g;
  g0.pec = 0;
  if (hc::aux_diag_inv(g0, op->ag, op->tb64, op->qd64, vinv, vinv, nullptr, st) != hipSuccess)
    return hc::fail(HC_EHIP, "hc_op_av_solve: node diagonal failed");
  const hc::PcgBlock blk[2] = {{N, op->dinv64, ba, a, ra, za, pa, qa}, {NN, vinv, bv, v, rv, zv, pv, qv}};
  auto ap = [&](const double * x0, const double * x1, double * y0, double * y1) -> int {
    return hc_op_av_apply(op, x0, x1, y0, y1, stream);
  };
  auto pr = [&](const double * r0, const double * r1, double * z0, double * z1) -> int {
    return hc::mg_av_precond(op, r0, r1, z0, z1, st);
  };
  int it = 0, status = HC_OK;
  bool conv = false;
  if ((rc = hc::pcg_dev(2, blk, ap, rtol, maxit, part, sdev, shost, st, it, conv, status, mg, pr)))
    return rc;
  // true residual of the final iterate
  if ((rc = hc_op_av_apply(op, a, v, qa, qv, stream)))
    return rc;
  const dim3 G(AV_BLOCKS), T(AV_THREADS);
  hipLaunchKernelGGL(hc::k_av_resid, G, T, 0, st, N, ba, qa, part);
  hipLaunchKernelGGL(hc::k_av_resid, G, T, 0, st, NN, bv, qv, part + AV_BLOCKS);
  (void)hipEventRecord(e1, st);
  std::vector<double> h(2 * AV_BLOCKS);
  HC_TRY_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * 2 * AV_BLOCKS, hipMemcpyDeviceToHost, st));
  HC_TRY_HIP(hipStreamSynchronize(st));
  double tr = 0.0;
  for (double x : h)
    tr += x;
  const double bb = shost->bb;
  const double rel = bb > 0.0 ? std::sqrt(tr / bb) : 0.0;
  conv = conv && status == HC_OK && rel <= rtol * (1.0 + 1e-6);
  if (status == HC_OK && !conv)
    status = HC_ENOTCONV;
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  if (stats)
  {
    *stats = hc_solve_stats{};
    stats->iterations = it;
    stats->outer_iterations = mg ? 2 : 1; // 2: multigrid block preconditioner, 1: block Jacobi
    stats->converged = conv ? 1 : 0;
    stats->status = status;
    stats->rel_res = rel;
    stats->solve_ms = ms;
  }
  if (status != HC_OK)
    return hc::fail(status, status == HC_ENAN ? "hc_op_av_solve: NaN/Inf in a reduction"
                                                     : "hc_op_av_solve: tolerance not reached in maxit");
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}
