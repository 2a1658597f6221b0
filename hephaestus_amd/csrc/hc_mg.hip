// Geometric multigrid preconditioner for the H(curl) system A = dt K_nu + M_sigma
// on structured box meshes, and the PCG solve that uses it (HC_SOLVE_MG).
//
// The reference preconditions this system with HypreAMS
// (src/formulations/HCurl/hcurl_formulation.cpp:69-79, AMS in singular mode):
// an auxiliary-space method whose node-space (gradient) and vector-node-space
// corrections are multigrid cycles.  This is the geometric counterpart on the
// nested structured meshes: a V-cycle whose levels halve the mesh (while the
// element counts are even and the coefficients constant on 2x2x2 blocks, so
// that each coarse operator is the Galerkin product of the finer one), then
// drop the order to 1 on the same mesh, then (order 1, only while too large
// for the coarsest solve) halve the mesh with averaged coefficients; the
// hybrid smoother of Hiptmair on every level (B = D^-1 + G Dg^-1 G^T: Jacobi
// on the edges plus Jacobi on the node space through the discrete gradient,
// the correction that AMS's gradient part performs) inside a Chebyshev
// iteration; and a dense Cholesky solve on the coarsest level.
//
// Transfers: the ND spaces are nested (ND_p on a mesh inside ND_p on its
// refinement, ND_1 inside ND_p on the same mesh), so prolongation evaluates
// the coarse field at the fine DOF functionals -- per component a tensor
// product of 1-D interpolations (open Gauss-Legendre basis along the
// component's axis, closed Gauss-Lobatto basis along the others) times
// h_f / h_c along the component's axis (covariant scaling of the tangential
// DOF); restriction is its transpose.  Both run as three 1-D passes.  The
// V-cycle (the same polynomial smoother before and after an exact coarse
// solve) is a fixed symmetric positive definite operator, so plain PCG
// applies.
#include "hc_aux.hpp"
#include "hc_basis.hpp"
#include "hc_op.hpp"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

namespace hc
{

namespace
{

constexpr int MG_SMOOTH = 2;            // pre- and post-smoothing steps
// Chebyshev smoothing range lam_hi / lam_lo per order (measured on the
// two-region coil problems: wider ranges pay off at higher order)
constexpr double MG_CHEB_RANGE[4] = {0.0, 6.0, 6.0, 20.0};
constexpr int64_t MG_COARSE_TARGET = 4000; // stop coarsening at this many DOFs
constexpr int64_t MG_COARSE_MAX = 12000;   // largest coarsest level solved by its dense inverse
// A coarsest level above the dense limit is approximated by this many steps
// of the Chebyshev iteration with the hybrid smoother on [lam / RANGE, 1.1
// lam] (measured on the 64-cube mix: 16 steps, range 100 solve its 59
// multigrid instances in 383 ms, against 653 ms with PCG to rtol 0.2 there)
constexpr int MG_COARSE_CHEB_STEPS = 16;
constexpr double MG_COARSE_CHEB_RANGE = 100.0;
constexpr int MG_POWER_ITERS = 15;      // power iterations for the smoother's lambda_max
constexpr int DOT_BLOCKS = 512;
// replay the finest-level V-cycle from a captured graph (off by default: a
// capture must not overlap synchronous HIP calls of other host threads, such
// as the hierarchy builds of the worker threads of the multigrid server and
// batch, which therefore always launch eagerly)
std::atomic<int> g_mg_graphs{0};
constexpr int DOT_THREADS = 256;

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// One 1-D pass of a tensor-product transfer: out = scale * M (along axis) in,
// + out when accumulate.  in: dims d[0..2] (x fastest); out: the same dims but
// d[axis] -> m (the rows of M, CSR rp/ci/w).  Grid: x over the (i0, i1)
// plane of the output, y over i2.
__global__ void __launch_bounds__(256)
k_pass1d(const double * __restrict__ in, double * __restrict__ out, int d0, int d1, int d2, int axis, int m,
         const int * __restrict__ rp, const int * __restrict__ ci, const double * __restrict__ w, double scale,
         int accumulate, int64_t sin, int64_t sout)
{
  in += blockIdx.z * sin; // right-hand side blockIdx.z (strides 0 for one)
  out += blockIdx.z * sout;
  const unsigned o0 = axis == 0 ? m : d0, o1 = axis == 1 ? m : d1;
  const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= o0 * o1)
    return;
  const int i1 = (int)(t / o0);
  const int i0 = (int)(t - (unsigned)i1 * o0);
  const int i2 = (int)blockIdx.y;
  const int row = axis == 0 ? i0 : (axis == 1 ? i1 : i2);
  // input offset of (j, i) along the axis: base + j * stride
  const int64_t stride = axis == 0 ? 1 : (axis == 1 ? (int64_t)d0 : (int64_t)d0 * d1);
  const int64_t base = ((int64_t)(axis == 2 ? 0 : i2) * d1 + (axis == 1 ? 0 : i1)) * d0 + (axis == 0 ? 0 : i0);
  double s = 0.0;
  for (int k = rp[row]; k < rp[row + 1]; k++)
    s += w[k] * in[base + ci[k] * stride];
  s *= scale;
  const int64_t o = ((int64_t)i2 * o1 + i1) * o0 + i0;
  out[o] = accumulate ? out[o] + s : s;
}

// The same pass for the three components of an ND transfer in one launch
// (blockIdx.z = 3 rhs + component): each component has its own dims, 1-D
// matrix, scale and input/output vectors; the per-entry arithmetic is that
// of k_pass1d.
struct Pass3
{
  const double * in[3];
  double * out[3];
  int d[3][3]; // input dims per component
  int m[3];    // rows of the component's 1-D matrix
  const int * rp[3];
  const int * ci[3];
  const double * w[3];
  double scale[3];
  int64_t sin, sout; // right-hand-side strides
  int axis, accumulate;
  int zero_pec; // PEC bits (HC_FACE_*) whose entries of the output are set to zero, or 0
};

__global__ void __launch_bounds__(256)
k_pass3(Pass3 q)
{
  const int comp = (int)(blockIdx.z % 3), rhs = (int)(blockIdx.z / 3);
  const int axis = q.axis, m = q.m[comp];
  const int d0 = q.d[comp][0], d1 = q.d[comp][1], d2 = q.d[comp][2];
  const unsigned o0 = axis == 0 ? m : d0, o1 = axis == 1 ? m : d1, o2 = axis == 2 ? m : d2;
  const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= o0 * o1 || blockIdx.y >= o2)
    return;
  const double * __restrict__ in = q.in[comp] + rhs * q.sin;
  double * __restrict__ out = q.out[comp] + rhs * q.sout;
  const int * __restrict__ rp = q.rp[comp];
  const int * __restrict__ ci = q.ci[comp];
  const double * __restrict__ w = q.w[comp];
  const int i1 = (int)(t / o0);
  const int i0 = (int)(t - (unsigned)i1 * o0);
  const int i2 = (int)blockIdx.y;
  const int row = axis == 0 ? i0 : (axis == 1 ? i1 : i2);
  const int64_t stride = axis == 0 ? 1 : (axis == 1 ? (int64_t)d0 : (int64_t)d0 * d1);
  const int64_t base = ((int64_t)(axis == 2 ? 0 : i2) * d1 + (axis == 1 ? 0 : i1)) * d0 + (axis == 0 ? 0 : i0);
  double s = 0.0;
  for (int k = rp[row]; k < rp[row + 1]; k++)
    s += w[k] * in[base + ci[k] * stride];
  s *= q.scale[comp];
  const int64_t o = ((int64_t)i2 * o1 + i1) * o0 + i0;
  if (q.zero_pec)
  {
    // PEC entries of the output (a restriction's last pass): tangential DOFs
    // on a PEC face of one of the component's closed axes
    const int ii[3] = {i0, i1, i2};
    const unsigned od[3] = {o0, o1, o2};
    bool ess = false;
#pragma unroll
    for (int b = 0; b < 3; b++)
      if (b != comp)
        ess = ess || (ii[b] == 0 && ((q.zero_pec >> (2 * b)) & 1)) ||
              ((unsigned)ii[b] == od[b] - 1 && ((q.zero_pec >> (2 * b + 1)) & 1));
    if (ess)
      s = 0.0;
  }
  out[o] = q.accumulate ? out[o] + s : s;
}

// Block partials of (a, b) and (a, a) (fixed grid and order).
__global__ void __launch_bounds__(DOT_THREADS)
k_dot2(const double * __restrict__ a, const double * __restrict__ b, int64_t n, double * __restrict__ part)
{
  __shared__ double sh[2][DOT_THREADS];
  double s0 = 0.0, s1 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * DOT_THREADS + threadIdx.x; i < n; i += (int64_t)DOT_BLOCKS * DOT_THREADS)
  {
    const double ai = a[i];
    s0 += ai * b[i];
    s1 += ai * ai;
  }
  double v[2] = {s0, s1};
  group_sum<DOT_THREADS, 2>(v, &sh[0][0]);
  if (threadIdx.x == 0)
  {
    part[2 * blockIdx.x] = v[0];
    part[2 * blockIdx.x + 1] = v[1];
  }
}

// x += alpha p; r -= alpha q (one pass)
__global__ void
k_xr_update(int64_t n, double alpha, const double * __restrict__ p, const double * __restrict__ q,
            double * __restrict__ x, double * __restrict__ r)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
  {
    x[i] += alpha * p[i];
    r[i] -= alpha * q[i];
  }
}

// Block partials of (a, b) and (a, a) per right-hand side blockIdx.y (vectors
// at stride n; the same blocks and order as k_dot2 for each of them).
__global__ void __launch_bounds__(DOT_THREADS)
k_dot2_k(const double * __restrict__ a, const double * __restrict__ b, int64_t n, double * __restrict__ part)
{
  __shared__ double sh[2][DOT_THREADS];
  a += blockIdx.y * n;
  b += blockIdx.y * n;
  double s0 = 0.0, s1 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * DOT_THREADS + threadIdx.x; i < n; i += (int64_t)DOT_BLOCKS * DOT_THREADS)
  {
    const double ai = a[i];
    s0 += ai * b[i];
    s1 += ai * ai;
  }
  double v[2] = {s0, s1};
  group_sum<DOT_THREADS, 2>(v, &sh[0][0]);
  if (threadIdx.x == 0)
  {
    double * o = part + 2 * ((int64_t)blockIdx.y * DOT_BLOCKS + blockIdx.x);
    o[0] = v[0];
    o[1] = v[1];
  }
}

// x += c p, r -= c q with c = coef[rhs] (vectors at stride n, rhs blockIdx.y)
__global__ void
k_xr_update_k(int64_t n, const double * __restrict__ coef, const double * __restrict__ p,
              const double * __restrict__ q, double * __restrict__ x, double * __restrict__ r)
{
  const double c = coef[blockIdx.y];
  if (c == 0.0)
    return;
  const int64_t o = blockIdx.y * n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
  {
    x[o + i] += c * p[o + i];
    r[o + i] -= c * q[o + i];
  }
}

// p = z + c p with c = coef[rhs]
__global__ void
k_zpbp_k(int64_t n, const double * __restrict__ coef, const double * __restrict__ z, double * __restrict__ p)
{
  const double c = coef[blockIdx.y];
  const int64_t o = blockIdx.y * n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[o + i] = z[o + i] + c * p[o + i];
}

// Chebyshev step: d = a z + b d (d = a z with flag 1: d not read); x = x + d
// (x = d with flag 2: x not read)
__global__ void
k_cheb_update(int64_t n, double a, double b, const double * __restrict__ z, double * __restrict__ d,
              double * __restrict__ x, int flags)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
  {
    const double di = (flags & 1) ? a * z[i] : a * z[i] + b * d[i];
    d[i] = di;
    x[i] = (flags & 2) ? di : x[i] + di;
  }
}

// Upper triangle of a column-major n x n matrix from its lower triangle.
__global__ void
k_symmetrize_lower(double * __restrict__ A, int64_t n)
{
  const int64_t total = n * n;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (int64_t)gridDim.x * blockDim.x)
  {
    const int64_t i = k % n, j = k / n; // element (i, j)
    if (i < j)
      A[k] = A[j + i * n];
  }
}

// y = A x for a dense symmetric n x n matrix stored full, column-major: row i
// of A is its column i (contiguous).  One workgroup per row, fixed-order sum.
__global__ void __launch_bounds__(256)
k_dense_symv(const double * __restrict__ A, int64_t n, const double * __restrict__ x, double * __restrict__ y)
{
  x += blockIdx.y * n; // right-hand side blockIdx.y
  y += blockIdx.y * n;
  __shared__ double sh[256];
  const int64_t i = blockIdx.x;
  const double * __restrict__ col = A + i * n;
  double s = 0.0;
  for (int64_t j = threadIdx.x; j < n; j += 256)
    s += col[j] * x[j];
  s = group_sum<256>(s, sh);
  if (threadIdx.x == 0)
    y[i] = s;
}

// Identity columns j0 .. j0 + nb - 1 of an n x n matrix (column-major, nb columns).
__global__ void
k_identity_cols(double * __restrict__ e, int64_t n, int j0, int nb)
{
  const int64_t total = n * nb;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (int64_t)gridDim.x * blockDim.x)
  {
    const int64_t i = k % n, j = k / n;
    e[k] = (i == j0 + j) ? 1.0 : 0.0;
  }
}

// A_jj *= 1 + delta (diagonal boost of a dense n x n matrix)
__global__ void
k_boost_diag(double * __restrict__ A, int64_t n, double delta)
{
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x)
    A[j * n + j] *= 1.0 + delta;
}

int
grid_for(int64_t n)
{
  const int64_t b = (n + 255) / 256;
  return (int)std::min<int64_t>(std::max<int64_t>(b, 1), 8192);
}

// ---------------------------------------------------------------------------
// 1-D transfer matrices
// ---------------------------------------------------------------------------

double
lagrange1(int n, const double * nodes, int j, double t)
{
  double v = 1.0;
  for (int m = 0; m < n; m++)
    if (m != j)
      v *= (t - nodes[m]) / (nodes[j] - nodes[m]);
  return v;
}

// Gauss-Legendre / Gauss-Lobatto points on [0,1] (hc_basis.hpp, the points
// of the operator's tables)
void
points01(int n, bool lobatto, double * x)
{
  if (lobatto)
  {
    basis::gauss_lobatto01(n, x);
    return;
  }
  double w[QMAXC + 1] = {};
  basis::gauss_legendre01(n, x, w);
}

struct Csr1
{
  std::vector<int> rp, ci;
  std::vector<double> w;
  int rows = 0, cols = 0;
};

// Prolongation of 1-D DOFs from nc coarse elements of order pc to nf = r nc
// fine elements (r = 1 or 2) of order pf >= pc: open (p Gauss-Legendre points
// per element) or closed (p+1 Gauss-Lobatto, shared at element ends).  Row =
// fine DOF, entries = the coarse basis at its point (the coarse field is in
// the fine space: ND_pc on a mesh is contained in ND_pf on its refinement).
Csr1
prolong_1d(int nc, int pc, int nf, int pf, bool closed)
{
  double gf[PMAX + 2], gc[PMAX + 2];
  points01(closed ? pf + 1 : pf, closed, gf);
  points01(closed ? pc + 1 : pc, closed, gc);
  Csr1 m;
  const int r = nf / nc;
  const int kc = closed ? pc + 1 : pc; // coarse local nodes per element
  m.rows = closed ? nf * pf + 1 : nf * pf;
  m.cols = closed ? nc * pc + 1 : nc * pc;
  m.rp.push_back(0);
  for (int g = 0; g < m.rows; g++)
  {
    const int ef = std::min(g / pf, nf - 1);
    const int jf = g - ef * pf;
    const int ec = ef / r, sub = ef % r;
    const double t = (sub + gf[jf]) / r; // position in the coarse element
    for (int jc = 0; jc < kc; jc++)
    {
      m.ci.push_back(ec * pc + jc);
      m.w.push_back(lagrange1(kc, gc, jc, t));
    }
    m.rp.push_back((int)m.ci.size());
  }
  return m;
}

Csr1
transpose(const Csr1 & a)
{
  Csr1 t;
  t.rows = a.cols;
  t.cols = a.rows;
  std::vector<std::vector<std::pair<int, double>>> rows(t.rows);
  for (int i = 0; i < a.rows; i++)
    for (int k = a.rp[i]; k < a.rp[i + 1]; k++)
      rows[a.ci[k]].push_back({i, a.w[k]}); // ascending fine index: a fixed summation order
  t.rp.push_back(0);
  for (auto & r : rows)
  {
    for (auto & e : r)
    {
      t.ci.push_back(e.first);
      t.w.push_back(e.second);
    }
    t.rp.push_back((int)t.ci.size());
  }
  return t;
}

// device copy of a CSR 1-D matrix (carved out of a DevBlock)
struct Csr1Dev
{
  const int *rp = nullptr, *ci = nullptr;
  const double * w = nullptr;
  int rows = 0, cols = 0;
};

} // namespace

// One level of the hierarchy.
struct MgLevel
{
  hc_op_t op = nullptr;
  bool owned = false;
  int n[3] = {0, 0, 0};
  int64_t N = 0;
  double *x = nullptr, *b = nullptr, *r = nullptr, *z = nullptr, *d = nullptr;
  double omega = 1.0; // Richardson damping 1 / (1.05 lambda_max)
  double lam = 1.0;   // lambda_max(B A) estimate
  double cheb = 0.0;  // Chebyshev range lam_hi / lam_lo of the smoother (<= 1: Richardson)
  // transfers to this level from the next coarser one: [axis][open=0/closed=1]
  Csr1Dev P[3][2], R[3][2];
};

struct NodeHier;
void node_free(NodeHier * n);
struct MgMulti;
void multi_free(MgMulti * m);

struct MgHier
{
  std::vector<MgLevel> lev;
  NodeHier * nodes = nullptr; // node-space hierarchy of the AV preconditioner (built on demand)
  MgMulti * multi = nullptr;  // work vectors of multi-RHS solves (allocated on demand)
  int nodes_state = 0;        // 0 not tried, 1 built, -1 not available
  int version = -1;
  DevBlock mem;        // level vectors, temporaries, transfer matrices
  double * tmp1 = nullptr;
  double * tmp2 = nullptr;
  double * zz = nullptr;   // preconditioned residual of the PCG
  double * part = nullptr; // DOT_BLOCKS * 2
  double * hpart = nullptr; // pinned copy
  double * Ainv = nullptr; // coarsest level, dense inverse (symmetric, stored full)
  bool coarse_iter = false; // coarsest level above MG_COARSE_MAX: Chebyshev steps instead of its inverse
  DevBlock ainv_mem;
  int64_t NL = 0;
  rocblas_handle blas = nullptr;
  int nsmooth = MG_SMOOTH;   // smoothing steps before and after the coarse correction
  // the finest-level V-cycle captured as a graph for fixed input and output
  // vectors (replayed on the solve's stream); rebuilt when a level operator's
  // workspace moved
  hipStream_t cap = nullptr;
  hipGraphExec_t vc = nullptr;
  const double * vc_in = nullptr;
  double * vc_out = nullptr;
  std::vector<void *> vc_ws;
  int vc_seen = 0; // eager V-cycles with the current pointers (capture from the second)
  ~MgHier()
  {
    node_free(nodes);
    multi_free(multi);
    for (auto & l : lev)
      if (l.owned && l.op)
        hc_op_destroy(l.op);
    if (hpart)
      (void)hipHostFree(hpart);
    if (blas)
      rocblas_destroy_handle(blas);
    if (vc)
      (void)hipGraphExecDestroy(vc);
    if (cap)
      (void)hipStreamDestroy(cap);
  }
};

void
mg_free(MgHier * h)
{
  delete h;
}

namespace
{

// dims of component c on a level with element counts n and order p
void
comp_dims(const int n[3], int p, int c, int d[3])
{
  for (int a = 0; a < 3; a++)
    d[a] = (a == c) ? n[a] * p : n[a] * p + 1;
}

hipError_t
dot2(MgHier & h, const double * a, const double * b, int64_t n, double & ab, double & aa, hipStream_t st)
{
  hipLaunchKernelGGL(k_dot2, dim3(DOT_BLOCKS), dim3(DOT_THREADS), 0, st, a, b, n, h.part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return e;
  if ((e = hipMemcpyAsync(h.hpart, h.part, sizeof(double) * 2 * DOT_BLOCKS, hipMemcpyDeviceToHost, st)) != hipSuccess)
    return e;
  if ((e = hipStreamSynchronize(st)) != hipSuccess)
    return e;
  ab = aa = 0.0;
  for (int i = 0; i < DOT_BLOCKS; i++)
  {
    ab += h.hpart[2 * i];
    aa += h.hpart[2 * i + 1];
  }
  return hipSuccess;
}

// y (level l) += P x (level l+1)   (prolong = true), or
// y (level l+1) = R x (level l)    (prolong = false)
hipError_t
transfer(MgHier & h, int l, bool prolong, const double * x, double * y, hipStream_t st, int nrhs = 1,
         double * tmp1 = nullptr, double * tmp2 = nullptr)
{
  const MgLevel & f = h.lev[l];
  const MgLevel & c = h.lev[l + 1];
  // nrhs vectors at strides N_in / N_out; the intermediates of every
  // right-hand side at stride N of the finest level
  if (!tmp1)
  {
    tmp1 = h.tmp1;
    tmp2 = h.tmp2;
  }
  const int64_t sx = prolong ? c.N : f.N, sy = prolong ? f.N : c.N, sT = h.lev[0].N;
  // per component: source, destination, and its intermediates at the
  // component's offset on the finer level (every intermediate of a component
  // is at most the size of its fine-level block, so the blocks do not meet)
  const double * src[3];
  double * dst[3];
  double *t1[3], *t2[3];
  int d[3][3];
  double scale[3];
  int64_t offf = 0, offc = 0;
  for (int comp = 0; comp < 3; comp++)
  {
    int df[3], dc[3];
    comp_dims(f.n, f.op->p, comp, df);
    comp_dims(c.n, c.op->p, comp, dc);
    // covariant DOF (h times the tangential component): h_f / h_c along the component's axis
    scale[comp] = (double)c.n[comp] / f.n[comp];
    src[comp] = x + (prolong ? offc : offf);
    dst[comp] = y + (prolong ? offf : offc);
    t1[comp] = tmp1 + offf;
    t2[comp] = tmp2 + offf;
    std::memcpy(d[comp], prolong ? dc : df, sizeof(d[comp]));
    offf += (int64_t)df[0] * df[1] * df[2];
    offc += (int64_t)dc[0] * dc[1] * dc[2];
  }
  // pass a over every component: x (src -> t1), y (t1 -> t2), z (t2 -> dst)
  for (int a = 0; a < 3; a++)
  {
    Pass3 q{};
    int64_t plane = 0, o2 = 0;
    for (int comp = 0; comp < 3; comp++)
    {
      const int kind = (a == comp) ? 0 : 1;
      const Csr1Dev & m = prolong ? f.P[a][kind] : f.R[a][kind];
      q.in[comp] = (a == 0) ? src[comp] : (a == 1 ? t1[comp] : t2[comp]);
      q.out[comp] = (a == 2) ? dst[comp] : (a == 0 ? t1[comp] : t2[comp]);
      for (int b = 0; b < 3; b++)
        q.d[comp][b] = d[comp][b];
      q.m[comp] = m.rows;
      q.rp[comp] = m.rp;
      q.ci[comp] = m.ci;
      q.w[comp] = m.w;
      q.scale[comp] = a == 2 ? scale[comp] : 1.0;
      plane = std::max(plane, (int64_t)(a == 0 ? m.rows : d[comp][0]) * (a == 1 ? m.rows : d[comp][1]));
      o2 = std::max(o2, (int64_t)(a == 2 ? m.rows : d[comp][2]));
      d[comp][a] = m.rows;
    }
    q.axis = a;
    q.accumulate = (a == 2 && prolong) ? 1 : 0;
    q.zero_pec = (a == 2 && !prolong) ? c.op->g.pec : 0;
    q.sin = a == 0 ? sx : sT;
    q.sout = a == 2 ? sy : sT;
    hipLaunchKernelGGL(k_pass3, dim3((unsigned)((plane + 255) / 256), (unsigned)o2, (unsigned)(3 * nrhs)), dim3(256),
                       0, st, q);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
      return e;
  }
  return hipSuccess; // (a restriction's PEC entries are zeroed by its last pass)
}

// r = b - A x on level l (x == nullptr: r = b)
int
residual(MgLevel & L, const double * b, const double * x, double * r, hipStream_t st)
{
  if (!x)
  {
    HC_TRY_HIP(hipMemcpyAsync(r, b, sizeof(double) * L.N, hipMemcpyDeviceToDevice, st));
    return HC_OK;
  }
  return op_residual(L.op, b, x, r, st);
}

// h.nsmooth steps of the smoother on a level, from x = 0 when zero_start.
// Richardson (cheb_range 0): x += omega B (b - A x).  Chebyshev (cheb_range
// alpha > 1): the Chebyshev iteration for B A on [lam_hi / alpha, lam_hi],
// lam_hi = 1.1 lambda_max (Saad, Iterative Methods for Sparse Linear
// Systems, 2nd ed., Algorithm 12.1, with the residual recomputed each step).
// Either error propagator is a polynomial in B A, self-adjoint in the A
// inner product, so the same smoother before and after the coarse correction
// keeps the V-cycle symmetric.
int
smooth(MgHier & h, MgLevel & L, bool zero_start, hipStream_t st, int steps = -1)
{
  const int nsteps = steps < 0 ? h.nsmooth : steps;
  double theta = 0, delta = 0, sigma1 = 0, rho0 = 0;
  if (L.cheb > 1.0)
  {
    const double hi = 1.1 * L.lam, lo = hi / L.cheb;
    theta = 0.5 * (hi + lo);
    delta = 0.5 * (hi - lo);
    sigma1 = theta / delta;
    rho0 = 1.0 / sigma1;
  }
  for (int s = 0; s < nsteps; s++)
  {
    const bool first = zero_start && s == 0;
    if (int rc = residual(L, L.b, first ? nullptr : L.x, L.r, st))
      return rc;
    if (L.cheb > 1.0)
    {
      double a, b;
      if (s == 0)
      {
        a = 1.0 / theta;
        b = 0.0;
      }
      else
      {
        const double rho1 = 1.0 / (2.0 * sigma1 - rho0);
        a = 2.0 * rho1 / delta;
        b = rho1 * rho0;
        rho0 = rho1;
      }
      // z = B r and the update in one pass after the node stage
      if (int rc = L.op->ensure_ws(1, false))
        return rc;
      HC_TRY_HIP(aux_smooth_step(L.op->ag, L.op->dinv64, L.op->ginv64, L.r, L.op->ws.t64, a, b, L.d, L.x,
                                 (s == 0 ? 1 : 0) | (first ? 2 : 0), st));
    }
    else
    {
      if (int rc = hc_op_precond_apply(L.op, HC_SOLVE_FP64 | HC_SOLVE_AUX, L.r, L.z, st))
        return rc;
      HC_TRY_HIP(axpby(L.N, L.omega, L.z, first ? 0.0 : 1.0, L.x, st));
    }
  }
  return HC_OK;
}

// x_l <- V-cycle(b_l) with x_l starting from zero
int
vcycle(MgHier & h, int l, hipStream_t st)
{
  MgLevel & L = h.lev[l];
  if (l == (int)h.lev.size() - 1 && h.coarse_iter)
    return smooth(h, L, true, st, MG_COARSE_CHEB_STEPS); // a fixed polynomial in B A, applied to b
  if (l == (int)h.lev.size() - 1)
  {
    // the inverse is stored full (symmetric): x = A^-1 b reads it by columns
    hipLaunchKernelGGL(k_dense_symv, dim3((unsigned)L.N), dim3(256), 0, st, (const double *)h.Ainv, L.N,
                       (const double *)L.b, L.x);
    HC_TRY_HIP(hipGetLastError());
    return HC_OK;
  }
  if (int rc = smooth(h, L, true, st)) // pre-smoothing from x = 0
    return rc;
  if (int rc = residual(L, L.b, L.x, L.r, st))
    return rc;
  HC_TRY_HIP(transfer(h, l, false, L.r, h.lev[l + 1].b, st));
  if (int rc = vcycle(h, l + 1, st))
    return rc;
  HC_TRY_HIP(transfer(h, l, true, h.lev[l + 1].x, L.x, st));
  return smooth(h, L, false, st); // post-smoothing
}

// z = V-cycle(r) on the finest level.  With graphs on and an exact coarsest
// solve the launches are captured once per (r, z, level workspaces) -- on the
// second call, after an eager one has made every lazy allocation -- and
// replayed on st; the replay runs the same kernels with the same arguments in
// the same order as the eager cycle.
int
vcycle_top(MgHier & h, double * rin, double * zout, hipStream_t st, bool allow_graph)
{
  MgLevel & L = h.lev[0];
  auto eager = [&](hipStream_t s) -> int {
    double * const b0 = L.b;
    double * const x0 = L.x;
    L.b = rin;
    L.x = zout;
    const int rc = vcycle(h, 0, s);
    L.b = b0;
    L.x = x0;
    return rc;
  };
  if (!allow_graph || !g_mg_graphs.load() || h.coarse_iter)
    return eager(st);
  std::vector<void *> ws;
  for (auto & l : h.lev)
    ws.push_back(l.op->ws_mem.base);
  const bool same = h.vc_in == rin && h.vc_out == zout && ws == h.vc_ws;
  if (!same)
  {
    if (h.vc)
      (void)hipGraphExecDestroy(h.vc);
    h.vc = nullptr;
    h.vc_in = rin;
    h.vc_out = zout;
    h.vc_ws = ws;
    h.vc_seen = 0;
  }
  if (!h.vc)
  {
    if (h.vc_seen++ == 0)
      return eager(st);
    if (!h.cap)
      HC_TRY_HIP(hipStreamCreateWithFlags(&h.cap, hipStreamNonBlocking));
    HC_TRY_HIP(hipStreamBeginCapture(h.cap, hipStreamCaptureModeThreadLocal));
    const int rc = eager(h.cap);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(h.cap, &g);
    if (rc || e != hipSuccess)
    {
      if (g)
        (void)hipGraphDestroy(g);
      return rc ? rc : fail(HC_EHIP, std::string("multigrid: V-cycle capture: ") + hipGetErrorString(e));
    }
    const hipError_t ei = hipGraphInstantiate(&h.vc, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess)
    {
      h.vc = nullptr;
      return fail(HC_EHIP, std::string("multigrid: V-cycle graph instantiation: ") + hipGetErrorString(ei));
    }
  }
  HC_TRY_HIP(hipGraphLaunch(h.vc, st));
  return HC_OK;
}

} // namespace

// Work vectors of a solve of k right-hand sides at once (every vector holds
// the k right-hand sides one after another).
struct MgMulti
{
  int k = 0;
  DevBlock mem;
  std::vector<double *> x, b, r, z, d; // per level, k N_l each
  double *tmp1 = nullptr, *tmp2 = nullptr;
  double *R = nullptr, *P = nullptr, *Q = nullptr, *Z = nullptr;
  double *part = nullptr, *coef = nullptr;
  double *hpart = nullptr, *hcoef = nullptr; // pinned
  ~MgMulti()
  {
    if (hpart)
      (void)hipHostFree(hpart);
    if (hcoef)
      (void)hipHostFree(hcoef);
  }
};

void
multi_free(MgMulti * m)
{
  delete m;
}

namespace
{

int
mm_ensure(MgHier & h, int k)
{
  if (h.multi && h.multi->k >= k)
    return HC_OK;
  multi_free(h.multi);
  h.multi = nullptr;
  auto * mm = new MgMulti();
  std::unique_ptr<MgMulti> guard(mm);
  mm->k = k;
  const int64_t N0 = h.lev[0].N;
  size_t bytes = 0;
  for (auto & L : h.lev)
    bytes += 5 * align256(sizeof(double) * L.N * k);
  bytes += 6 * align256(sizeof(double) * N0 * k) + align256(sizeof(double) * 2 * DOT_BLOCKS * k) +
           align256(sizeof(double) * k);
  if (int rc = mm->mem.alloc(bytes))
    return rc;
  char * q = (char *)mm->mem.base;
  auto take = [&](int64_t n) {
    double * v = (double *)q;
    q += align256(sizeof(double) * n);
    return v;
  };
  for (auto & L : h.lev)
  {
    mm->x.push_back(take(L.N * k));
    mm->b.push_back(take(L.N * k));
    mm->r.push_back(take(L.N * k));
    mm->z.push_back(take(L.N * k));
    mm->d.push_back(take(L.N * k));
  }
  mm->tmp1 = take(N0 * k);
  mm->tmp2 = take(N0 * k);
  mm->R = take(N0 * k);
  mm->P = take(N0 * k);
  mm->Q = take(N0 * k);
  mm->Z = take(N0 * k);
  mm->part = take(2 * DOT_BLOCKS * k);
  mm->coef = take(k);
  HC_TRY_HIP(hipHostMalloc((void **)&mm->hpart, sizeof(double) * 2 * DOT_BLOCKS * k));
  HC_TRY_HIP(hipHostMalloc((void **)&mm->hcoef, sizeof(double) * k));
  for (auto & L : h.lev)
    if (int rc = L.op->ensure_ws(k, false))
      return rc;
  h.multi = guard.release();
  return HC_OK;
}

// per right-hand side (a, b) and (a, a) of k vectors of n entries
hipError_t
dot2_k(MgMulti & mm, const double * a, const double * b, int64_t n, int k, std::vector<double> & ab,
       std::vector<double> & aa, hipStream_t st)
{
  hipLaunchKernelGGL(k_dot2_k, dim3(DOT_BLOCKS, k), dim3(DOT_THREADS), 0, st, a, b, n, mm.part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return e;
  if ((e = hipMemcpyAsync(mm.hpart, mm.part, sizeof(double) * 2 * DOT_BLOCKS * k, hipMemcpyDeviceToHost, st)) !=
      hipSuccess)
    return e;
  if ((e = hipStreamSynchronize(st)) != hipSuccess)
    return e;
  ab.assign(k, 0.0);
  aa.assign(k, 0.0);
  for (int r = 0; r < k; r++)
    for (int i = 0; i < DOT_BLOCKS; i++)
    {
      ab[r] += mm.hpart[2 * ((int64_t)r * DOT_BLOCKS + i)];
      aa[r] += mm.hpart[2 * ((int64_t)r * DOT_BLOCKS + i) + 1];
    }
  return hipSuccess;
}

int
smooth_k(MgHier & h, MgMulti & mm, int l, bool zero_start, int k, hipStream_t st, int steps = -1)
{
  const int nsteps = steps < 0 ? h.nsmooth : steps;
  MgLevel & L = h.lev[l];
  const int64_t n = L.N * (int64_t)k;
  double *x = mm.x[l], *b = mm.b[l], *r = mm.r[l], *z = mm.z[l], *d = mm.d[l];
  double theta = 0, delta = 0, sigma1 = 0, rho0 = 0;
  if (L.cheb > 1.0)
  {
    const double hi = 1.1 * L.lam, lo = hi / L.cheb;
    theta = 0.5 * (hi + lo);
    delta = 0.5 * (hi - lo);
    sigma1 = theta / delta;
    rho0 = 1.0 / sigma1;
  }
  for (int s = 0; s < nsteps; s++)
  {
    const bool first = zero_start && s == 0;
    if (first)
      HC_TRY_HIP(hipMemcpyAsync(r, b, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    else if (int rc = op_residual(L.op, b, x, r, st, k))
      return rc;
    if (L.cheb > 1.0)
    {
      double a, bb;
      if (s == 0)
      {
        a = 1.0 / theta;
        bb = 0.0;
      }
      else
      {
        const double rho1 = 1.0 / (2.0 * sigma1 - rho0);
        a = 2.0 * rho1 / delta;
        bb = rho1 * rho0;
        rho0 = rho1;
      }
      // z = B r and the update in one pass after the node stage
      HC_TRY_HIP(aux_smooth_step(L.op->ag, L.op->dinv64, L.op->ginv64, r, L.op->ws.t64, a, bb, d, x,
                                 (s == 0 ? 1 : 0) | (first ? 2 : 0), st, k, L.N, L.op->ws.NNC));
    }
    else
    {
      HC_TRY_HIP(aux_precond_apply(L.op->ag, L.op->dinv64, L.op->ginv64, r, L.op->ws.t64, z, st, k, L.N,
                                   L.op->ws.NNC));
      HC_TRY_HIP(axpby(n, L.omega, z, first ? 0.0 : 1.0, x, st));
    }
  }
  return HC_OK;
}

int
vcycle_k(MgHier & h, MgMulti & mm, int l, int k, hipStream_t st)
{
  MgLevel & L = h.lev[l];
  if (l == (int)h.lev.size() - 1 && h.coarse_iter)
    return smooth_k(h, mm, l, true, k, st, MG_COARSE_CHEB_STEPS);
  if (l == (int)h.lev.size() - 1)
  {
    hipLaunchKernelGGL(k_dense_symv, dim3((unsigned)L.N, (unsigned)k), dim3(256), 0, st, (const double *)h.Ainv, L.N,
                       (const double *)mm.b[l], mm.x[l]);
    HC_TRY_HIP(hipGetLastError());
    return HC_OK;
  }
  if (int rc = smooth_k(h, mm, l, true, k, st))
    return rc;
  if (int rc = op_residual(L.op, mm.b[l], mm.x[l], mm.r[l], st, k))
    return rc;
  HC_TRY_HIP(transfer(h, l, false, mm.r[l], mm.b[l + 1], st, k, mm.tmp1, mm.tmp2));
  if (int rc = vcycle_k(h, mm, l + 1, k, st))
    return rc;
  HC_TRY_HIP(transfer(h, l, true, mm.x[l + 1], mm.x[l], st, k, mm.tmp1, mm.tmp2));
  return smooth_k(h, mm, l, false, k, st);
}

Csr1Dev
upload(const Csr1 & m, char *& p)
{
  Csr1Dev d;
  d.rows = m.rows;
  d.cols = m.cols;
  const size_t a = align256(sizeof(int) * m.rp.size()), b = align256(sizeof(int) * m.ci.size());
  (void)hipMemcpy(p, m.rp.data(), sizeof(int) * m.rp.size(), hipMemcpyHostToDevice);
  d.rp = (const int *)p;
  p += a;
  (void)hipMemcpy(p, m.ci.data(), sizeof(int) * m.ci.size(), hipMemcpyHostToDevice);
  d.ci = (const int *)p;
  p += b;
  (void)hipMemcpy(p, m.w.data(), sizeof(double) * m.w.size(), hipMemcpyHostToDevice);
  d.w = (const double *)p;
  p += align256(sizeof(double) * m.w.size());
  return d;
}

size_t
csr_bytes(const Csr1 & m)
{
  return align256(sizeof(int) * m.rp.size()) + align256(sizeof(int) * m.ci.size()) +
         align256(sizeof(double) * m.w.size());
}

// lambda_max of B A on a level by power iteration (from a seeded vector)
int
estimate_lambda(MgHier & h, MgLevel & L, hipStream_t st, double & lam)
{
  HC_TRY_HIP(fill_uniform(&L.op->g, 91, L.x, L.N, st));
  lam = 1.0;
  for (int it = 0; it < MG_POWER_ITERS; it++)
  {
    double ab, nn;
    HC_TRY_HIP(dot2(h, L.x, L.x, L.N, ab, nn, st));
    const double s = 1.0 / std::sqrt(std::max(nn, 1e-300));
    HC_TRY_HIP(axpby(L.N, s, L.x, 0.0, L.x, st));
    if (int rc = hc_op_apply(L.op, L.x, L.r, 1, st))
      return rc;
    if (int rc = hc_op_precond_apply(L.op, HC_SOLVE_FP64 | HC_SOLVE_AUX, L.r, L.z, st))
      return rc;
    // Rayleigh quotient in the A inner product: (x, B A x)_A / (x, x)_A
    double xz, zz, xr, xx;
    HC_TRY_HIP(dot2(h, L.r, L.z, L.N, xz, zz, st)); // (A x, B A x)
    HC_TRY_HIP(dot2(h, L.x, L.r, L.N, xr, xx, st)); // (x, A x)
    lam = xz / xr;
    HC_TRY_HIP(hipMemcpyAsync(L.x, L.z, sizeof(double) * L.N, hipMemcpyDeviceToDevice, st));
  }
  return HC_OK;
}

// Hierarchy builds run one at a time in the process: a build is one-time
// setup (level operators, their autotuning, rocSOLVER on the coarsest level,
// power iterations), and the multigrid server and batch workers otherwise
// start several of them together on their first requests.  Solves on built
// hierarchies run concurrently.
std::mutex g_build_mu;

// the hierarchy of op (built on first use; rebuilt when op was set up again)
int
mg_build(hc_op_t op, hipStream_t st)
{
  if (op->mg && op->mg->version == op->setup_version)
    return HC_OK;
  std::lock_guard<std::mutex> lk(g_build_mu);
  mg_free(op->mg);
  op->mg = nullptr;
  if (!op->box)
    return fail(HC_EINVAL, "multigrid: needs a box mesh (nested coarse meshes)");
  MgHier * h = new MgHier();
  std::unique_ptr<MgHier> guard(h);
  MgLevel L0;
  L0.op = op;
  for (int a = 0; a < 3; a++)
    L0.n[a] = op->g.n[a];
  L0.N = op->N;
  h->lev.push_back(L0);
  std::vector<double> nu = op->nu, sig = op->sig;
  // Coarsening, one step per level: halve the mesh while every element count
  // is even and the coefficients are constant on each 2x2x2 block of children
  // (the coarse operator is then the Galerkin product); else drop the order
  // to 1 on the same mesh (ND_1 is contained in ND_p); else, at order 1 and
  // only while the level is too large for the dense coarsest solve, halve the
  // mesh with the coefficients averaged over the children.
  while (h->lev.back().N > MG_COARSE_TARGET)
  {
    const MgLevel & f = h->lev.back();
    const int fp = f.op->p;
    const bool even = f.n[0] % 2 == 0 && f.n[1] % 2 == 0 && f.n[2] % 2 == 0;
    bool aligned = even;
    for (int k = 0; aligned && k < f.n[2]; k++)
      for (int j = 0; aligned && j < f.n[1]; j++)
        for (int i = 0; aligned && i < f.n[0]; i++)
        {
          const int64_t e = ((int64_t)k * f.n[1] + j) * f.n[0] + i;
          const int64_t e0 = ((int64_t)(k & ~1) * f.n[1] + (j & ~1)) * f.n[0] + (i & ~1);
          aligned = nu[e] == nu[e0] && sig[e] == sig[e0];
        }
    int step; // 0: halve (exact), 1: order -> 1, 2: halve (averaged)
    if (aligned)
      step = 0;
    else if (fp > 1)
      step = 1;
    else if (even && f.N > MG_COARSE_MAX)
      step = 2;
    else
      break;
    MgLevel c;
    for (int a = 0; a < 3; a++)
      c.n[a] = step == 1 ? f.n[a] : f.n[a] / 2;
    const int cp = step == 1 ? 1 : fp;
    std::vector<double> nuc, sigc;
    if (step == 1)
    {
      nuc = nu;
      sigc = sig;
    }
    else
    {
      const int64_t Ec = (int64_t)c.n[0] * c.n[1] * c.n[2];
      nuc.assign(Ec, 0.0);
      sigc.assign(Ec, 0.0);
      for (int k = 0; k < c.n[2]; k++)
        for (int j = 0; j < c.n[1]; j++)
          for (int i = 0; i < c.n[0]; i++)
          {
            const int64_t ec = ((int64_t)k * c.n[1] + j) * c.n[0] + i;
            for (int dz = 0; dz < 2; dz++)
              for (int dy = 0; dy < 2; dy++)
                for (int dx = 0; dx < 2; dx++)
                {
                  const int64_t ef = ((int64_t)(2 * k + dz) * f.n[1] + (2 * j + dy)) * f.n[0] + (2 * i + dx);
                  nuc[ec] += 0.125 * nu[ef];
                  sigc[ec] += 0.125 * sig[ef];
                }
          }
    }
    hc_mesh_desc d{};
    d.nx = c.n[0];
    d.ny = c.n[1];
    d.nz = c.n[2];
    d.vert = nullptr;
    for (int a = 0; a < 3; a++)
    {
      d.x0[a] = op->x0[a];
      d.len[a] = op->h[a] * op->g.n[a];
    }
    d.pec_faces = op->g.pec;
    int rc = hc_op_create(&d, cp, nuc.data(), sigc.data(), HC_PREC_FP64, &c.op);
    if (rc)
      return rc;
    c.owned = true;
    h->lev.push_back(c);
    if ((rc = hc_op_setup(h->lev.back().op, op->last_dt, st)))
      return rc;
    // the level's own fastest tile variant (its mesh differs from the finest)
    if ((rc = hc_op_autotune(h->lev.back().op, 3, nullptr, st)))
      return rc;
    h->lev.back().N = h->lev.back().op->N;
    nu.swap(nuc);
    sig.swap(sigc);
  }
  for (auto & L : h->lev)
    L.cheb = MG_CHEB_RANGE[std::min(std::max(L.op->p, 1), 3)];
  const int nl = (int)h->lev.size();
  // a single level (a mesh small enough already) is solved by the dense inverse
  // alone; a coarsest level above the dense limit (odd element counts at order
  // 1) is approximated by MG_COARSE_CHEB_STEPS Chebyshev steps with the hybrid
  // smoother (a fixed symmetric positive definite polynomial, so the V-cycle
  // stays a fixed SPD preconditioner for plain PCG)
  h->NL = h->lev.back().N;
  h->coarse_iter = h->NL > MG_COARSE_MAX;
  if (h->coarse_iter)
    h->lev.back().cheb = MG_COARSE_CHEB_RANGE;
  if (h->coarse_iter && nl < 2)
    return fail(HC_EINVAL, "multigrid: coarsest level has " + std::to_string(h->NL) +
                               " DOFs and the mesh cannot be coarsened (odd element counts at order 1; at most " +
                               std::to_string(MG_COARSE_MAX) + " DOFs for a single level)");
  // transfer matrices
  std::vector<Csr1> mats; // per level pair: [axis][kind] P then R
  size_t bytes = 0;
  for (int l = 0; l + 1 < nl; l++)
    for (int a = 0; a < 3; a++)
      for (int kind = 0; kind < 2; kind++)
      {
        Csr1 P = prolong_1d(h->lev[l + 1].n[a], h->lev[l + 1].op->p, h->lev[l].n[a], h->lev[l].op->p, kind == 1);
        Csr1 R = transpose(P);
        bytes += csr_bytes(P) + csr_bytes(R);
        mats.push_back(std::move(P));
        mats.push_back(std::move(R));
      }
  // vectors: x, b, r, z per level; two temporaries of the finest size; dot partials
  for (auto & L : h->lev)
    bytes += 5 * align256(sizeof(double) * L.N);
  bytes += 3 * align256(sizeof(double) * op->N) + align256(sizeof(double) * 2 * DOT_BLOCKS);
  if (int rc = h->mem.alloc(bytes))
    return rc;
  char * q = (char *)h->mem.base;
  size_t mi = 0;
  for (int l = 0; l + 1 < nl; l++)
    for (int a = 0; a < 3; a++)
      for (int kind = 0; kind < 2; kind++)
      {
        h->lev[l].P[a][kind] = upload(mats[mi++], q);
        h->lev[l].R[a][kind] = upload(mats[mi++], q);
      }
  for (auto & L : h->lev)
  {
    double ** v[5] = {&L.x, &L.b, &L.r, &L.z, &L.d};
    for (double ** pp : v)
    {
      *pp = (double *)q;
      q += align256(sizeof(double) * L.N);
    }
  }
  h->tmp1 = (double *)q;
  q += align256(sizeof(double) * op->N);
  h->tmp2 = (double *)q;
  q += align256(sizeof(double) * op->N);
  h->zz = (double *)q;
  q += align256(sizeof(double) * op->N);
  h->part = (double *)q;
  HC_TRY_HIP(hipHostMalloc((void **)&h->hpart, sizeof(double) * 2 * DOT_BLOCKS));
  // smoother damping per level: 1 / lambda_max(B A)
  for (int l = 0; l + 1 < nl || (l < nl && h->coarse_iter); l++)
  {
    double lam = 1.0;
    if (int rc = estimate_lambda(*h, h->lev[l], st, lam))
      return rc;
    h->lev[l].lam = lam;
    h->lev[l].omega = 1.0 / (1.05 * lam);
  }
  // coarsest level: dense matrix by probing, Cholesky inverse (lower triangle)
  if (!h->coarse_iter)
  {
    MgLevel & C = h->lev.back();
    const int64_t n = C.N;
    if (int rc = h->ainv_mem.alloc(sizeof(double) * n * n + sizeof(int) * 64))
      return rc;
    h->Ainv = (double *)h->ainv_mem.base;
    int * info = (int *)((char *)h->ainv_mem.base + sizeof(double) * n * n);
    const int nb = 256;
    DevBlock eblk;
    if (int rc = eblk.alloc(sizeof(double) * n * nb))
      return rc;
    if (rocblas_create_handle(&h->blas) != rocblas_status_success)
      return fail(HC_EHIP, "multigrid: rocblas_create_handle failed");
    if (rocblas_set_stream(h->blas, st) != rocblas_status_success)
      return fail(HC_EHIP, "multigrid: rocblas_set_stream failed");
    // A coarsest level with a tiny sigma (a gradient-dominated null space
    // near h^3 sigma against a curl part near 1 / h) is SPD but can be so
    // ill-conditioned that the rounding of its probed columns (which depends
    // on the level's tile variant) leaves a Cholesky pivot negative; then
    // the probe is repeated with the diagonal raised by a relative 1e-12,
    // 1e-10, 1e-8 (the coarsest solve is part of a preconditioner).  The
    // first attempt is the plain matrix.
    constexpr double BOOST[4] = {0.0, 1e-12, 1e-10, 1e-8};
    int hinfo[2] = {0, 0};
    for (int attempt = 0; attempt < 4; attempt++)
    {
      for (int j0 = 0; j0 < n; j0 += nb)
      {
        const int k = (int)std::min<int64_t>(nb, n - j0);
        hipLaunchKernelGGL(k_identity_cols, dim3(grid_for(n * k)), dim3(256), 0, st, (double *)eblk.base, n, j0, k);
        HC_TRY_HIP(hipGetLastError());
        if (int rc = hc_op_apply(C.op, (const double *)eblk.base, h->Ainv + (int64_t)j0 * n, k, st))
          return rc;
      }
      if (BOOST[attempt] > 0.0)
      {
        hipLaunchKernelGGL(k_boost_diag, dim3(grid_for(n)), dim3(256), 0, st, h->Ainv, n, BOOST[attempt]);
        HC_TRY_HIP(hipGetLastError());
      }
      if (rocsolver_dpotrf(h->blas, rocblas_fill_lower, (rocblas_int)n, h->Ainv, (rocblas_int)n, info) !=
              rocblas_status_success ||
          rocsolver_dpotri(h->blas, rocblas_fill_lower, (rocblas_int)n, h->Ainv, (rocblas_int)n, info + 1) !=
              rocblas_status_success)
        return fail(HC_EHIP, "multigrid: rocsolver Cholesky of the coarsest level failed");
      HC_TRY_HIP(hipMemcpyAsync(hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost, st));
      HC_TRY_HIP(hipStreamSynchronize(st));
      if (!hinfo[0] && !hinfo[1])
        break;
    }
    if (hinfo[0] || hinfo[1])
      return fail(HC_EHIP, "multigrid: coarsest-level matrix is not positive definite (info " +
                               std::to_string(hinfo[0]) + ", " + std::to_string(hinfo[1]) + ")");
    hipLaunchKernelGGL(k_symmetrize_lower, dim3(grid_for(n * n)), dim3(256), 0, st, h->Ainv, n);
    HC_TRY_HIP(hipGetLastError());
  }
  h->version = op->setup_version;
  op->mg = guard.release();
  return HC_OK;
}

} // namespace

// PCG with the multigrid V-cycle as preconditioner, FP64, one right-hand side:
// the stopping rule of hc_pcg_solve (recursive residual, then the true FP64
// residual, restarted from the current x up to MAX_RESTARTS times).
int
mg_pcg_solve(hc_op_t op, const double * b, double * x, double rtol, int maxit, hc_solve_stats * stats,
             hipStream_t st, bool allow_graph)
{
  constexpr int MAX_RESTARTS = 4;
  if (int rc = mg_build(op, st))
    return rc;
  MgHier & h = *op->mg;
  MgLevel & L = h.lev[0];
  const int64_t N = op->N;
  // PCG vectors: r, p, q in the operator's workspace, z in the hierarchy (the
  // V-cycle works in the level vectors)
  if (int rc = op->ensure_ws(1, false))
    return rc;
  double * r = op->ws.r64;
  double * pv = op->ws.p64;
  double * q = op->ws.q64;
  double * zz = h.zz;
  hipEvent_t e0, e1;
  HC_TRY_HIP(hipEventCreate(&e0));
  HC_TRY_HIP(hipEventCreate(&e1));
  HC_TRY_HIP(hipEventRecord(e0, st));
  double bb, dummy;
  HC_TRY_HIP(dot2(h, b, b, N, dummy, bb, st));
  const double tol2 = rtol * rtol * bb;
  int iters = 0, outer = 0, passes = 0, status = HC_OK; // passes: PCG runs from a true residual
  bool conv = false;
  double rr = 0.0;
  // the V-cycle reads the finest level's b and writes its x: point them at
  // the PCG's residual and preconditioned residual for the call
  auto precond = [&](double * rin, double * zout) -> int { return vcycle_top(h, rin, zout, st, allow_graph); };
  for (outer = 1; outer <= 1 + MAX_RESTARTS && !conv && status == HC_OK; outer++)
  {
    // r = b - A x
    if (int rc = hc_op_apply(op, x, r, 1, st))
      return rc;
    HC_TRY_HIP(axpby(N, 1.0, b, -1.0, r, st));
    HC_TRY_HIP(dot2(h, r, r, N, dummy, rr, st));
    if (rr <= tol2)
    {
      conv = true;
      break;
    }
    passes++;
    if (int rc = precond(r, zz))
      return rc;
    HC_TRY_HIP(hipMemcpyAsync(pv, zz, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
    double gam;
    HC_TRY_HIP(dot2(h, r, zz, N, gam, rr, st));
    while (iters < maxit)
    {
      if (int rc = hc_op_apply(op, pv, q, 1, st))
        return rc;
      double pq, pp;
      HC_TRY_HIP(dot2(h, pv, q, N, pq, pp, st));
      if (!std::isfinite(pq) || !std::isfinite(gam))
      {
        status = HC_ENAN;
        break;
      }
      const double alpha = gam / pq;
      hipLaunchKernelGGL(k_xr_update, dim3(grid_for(N)), dim3(256), 0, st, N, alpha, pv, q, x, r);
      HC_TRY_HIP(hipGetLastError());
      iters++;
      HC_TRY_HIP(dot2(h, r, r, N, dummy, rr, st));
      if (rr <= tol2)
        break;
      if (int rc = precond(r, zz))
        return rc;
      double gnew, rr2;
      HC_TRY_HIP(dot2(h, r, zz, N, gnew, rr2, st));
      const double beta = gnew / gam;
      HC_TRY_HIP(axpby(N, 1.0, zz, beta, pv, st));
      gam = gnew;
    }
    if (status != HC_OK)
      break;
    // true residual of the result
    if (int rc = hc_op_apply(op, x, r, 1, st))
      return rc;
    HC_TRY_HIP(axpby(N, 1.0, b, -1.0, r, st));
    HC_TRY_HIP(dot2(h, r, r, N, dummy, rr, st));
    conv = rr <= tol2;
    if (iters >= maxit)
      break;
  }
  HC_TRY_HIP(hipEventRecord(e1, st));
  HC_TRY_HIP(hipEventSynchronize(e1));
  float ms = 0;
  HC_TRY_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (stats)
  {
    hc_solve_stats s{};
    s.iterations = iters;
    s.outer_iterations = passes;
    s.converged = conv ? 1 : 0;
    s.status = status != HC_OK ? status : (conv ? HC_OK : HC_ENOTCONV);
    s.rel_res = bb > 0 ? std::sqrt(rr / bb) : 0.0;
    s.solve_ms = ms;
    *stats = s;
  }
  if (status != HC_OK)
    return fail(status, "hc_pcg_solve (multigrid): NaN met in a reduction");
  if (!conv)
    return fail(HC_ENOTCONV, "hc_pcg_solve (multigrid): not converged after " + std::to_string(iters) +
                                 " iterations, true relative residual " + std::to_string(bb > 0 ? std::sqrt(rr / bb) : 0.0));
  return HC_OK;
}

// PCG with the multigrid V-cycle on k right-hand sides at once (columns of
// B, X at stride N): every level operation runs as one launch over the k
// vectors (the operator applies as one multi-RHS launch), each right-hand
// side keeps its own PCG scalars and stops on its own (its later updates are
// zero), with the stopping rule of the one-RHS solve.  Each right-hand side's
// arithmetic is that of the same solve alone.  `handled` stays false when
// the hierarchy cannot be built (the returned error; the caller then solves
// the right-hand sides another way).
int
mg_pcg_solve_multi(hc_op_t op, const double * B, double * X, int k, double rtol, int maxit, hc_solve_stats * stats,
                   hipStream_t st, bool & handled)
{
  constexpr int MAX_RESTARTS = 4;
  handled = false;
  if (int rc = mg_build(op, st))
    return rc;
  MgHier & h = *op->mg;
  handled = true;
  if (int rc = mm_ensure(h, k))
    return rc;
  MgMulti & mm = *h.multi;
  const int64_t N = op->N, n = N * (int64_t)k;
  double *R = mm.R, *P = mm.P, *Q = mm.Q, *Z = mm.Z;
  hipEvent_t e0, e1;
  HC_TRY_HIP(hipEventCreate(&e0));
  HC_TRY_HIP(hipEventCreate(&e1));
  HC_TRY_HIP(hipEventRecord(e0, st));
  std::vector<double> bb, dummy, rr, gam, gnew, pq;
  HC_TRY_HIP(dot2_k(mm, B, B, N, k, dummy, bb, st));
  std::vector<double> tol2(k);
  for (int r = 0; r < k; r++)
    tol2[r] = rtol * rtol * bb[r];
  std::vector<int> iters(k, 0), active(k, 0), stat(k, HC_OK);
  std::vector<char> conv(k, 0);
  // the V-cycle's finest level reads R and writes Z
  auto precond = [&]() -> int {
    double * const b0 = mm.b[0];
    double * const x0 = mm.x[0];
    mm.b[0] = R;
    mm.x[0] = Z;
    const int rc = vcycle_k(h, mm, 0, k, st);
    mm.b[0] = b0;
    mm.x[0] = x0;
    return rc;
  };
  auto set_coef = [&](const std::vector<double> & c) -> hipError_t {
    for (int r = 0; r < k; r++)
      mm.hcoef[r] = c[r];
    return hipMemcpyAsync(mm.coef, mm.hcoef, sizeof(double) * k, hipMemcpyHostToDevice, st);
  };
  auto resid = [&]() -> int {
    if (int rc = hc_op_apply(op, X, R, k, st))
      return rc;
    HC_TRY_HIP(axpby(n, 1.0, B, -1.0, R, st));
    HC_TRY_HIP(dot2_k(mm, R, R, N, k, dummy, rr, st));
    return HC_OK;
  };
  const dim3 vg(std::max(1, std::min(grid_for(N), 2048 / std::max(1, k / 4))), k);
  int passes = 0; // PCG runs from a true residual
  for (int outer = 1; outer <= 1 + MAX_RESTARTS; outer++)
  {
    if (int rc = resid())
      return rc;
    bool any = false;
    for (int r = 0; r < k; r++)
    {
      conv[r] = rr[r] <= tol2[r];
      active[r] = !conv[r] && stat[r] == HC_OK && iters[r] < maxit;
      any = any || active[r];
    }
    if (!any)
      break;
    passes++;
    if (int rc = precond())
      return rc;
    HC_TRY_HIP(dot2_k(mm, R, Z, N, k, gam, dummy, st));
    HC_TRY_HIP(hipMemcpyAsync(P, Z, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    for (;;)
    {
      if (int rc = hc_op_apply(op, P, Q, k, st))
        return rc;
      HC_TRY_HIP(dot2_k(mm, P, Q, N, k, pq, dummy, st));
      std::vector<double> alpha(k, 0.0);
      for (int r = 0; r < k; r++)
        if (active[r])
        {
          if (!std::isfinite(pq[r]) || !std::isfinite(gam[r]))
          {
            stat[r] = HC_ENAN;
            active[r] = 0;
          }
          else
            alpha[r] = gam[r] / pq[r];
        }
      HC_TRY_HIP(set_coef(alpha));
      hipLaunchKernelGGL(k_xr_update_k, vg, dim3(256), 0, st, N, (const double *)mm.coef, (const double *)P,
                         (const double *)Q, X, R);
      HC_TRY_HIP(hipGetLastError());
      HC_TRY_HIP(dot2_k(mm, R, R, N, k, dummy, rr, st));
      bool any_active = false;
      for (int r = 0; r < k; r++)
        if (active[r])
        {
          iters[r]++;
          if (rr[r] <= tol2[r] || iters[r] >= maxit)
            active[r] = 0;
          any_active = any_active || active[r];
        }
      if (!any_active)
        break;
      if (int rc = precond())
        return rc;
      HC_TRY_HIP(dot2_k(mm, R, Z, N, k, gnew, dummy, st));
      std::vector<double> beta(k, 0.0);
      for (int r = 0; r < k; r++)
        if (active[r])
          beta[r] = gnew[r] / gam[r];
      HC_TRY_HIP(set_coef(beta));
      hipLaunchKernelGGL(k_zpbp_k, vg, dim3(256), 0, st, N, (const double *)mm.coef, (const double *)Z, P);
      HC_TRY_HIP(hipGetLastError());
      for (int r = 0; r < k; r++)
        if (active[r])
          gam[r] = gnew[r];
    }
  }
  // true residuals of the results
  if (int rc = resid())
    return rc;
  HC_TRY_HIP(hipEventRecord(e1, st));
  HC_TRY_HIP(hipEventSynchronize(e1));
  float ms = 0;
  HC_TRY_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  hc_solve_stats all{};
  all.converged = 1;
  int status = HC_OK;
  for (int r = 0; r < k; r++)
  {
    conv[r] = rr[r] <= tol2[r];
    all.iterations = std::max(all.iterations, iters[r]);
    all.rel_res = std::max(all.rel_res, bb[r] > 0 ? std::sqrt(rr[r] / bb[r]) : 0.0);
    if (!conv[r])
      all.converged = 0;
    if (stat[r] && !status)
      status = stat[r];
  }
  all.outer_iterations = passes;
  all.status = status ? status : (all.converged ? HC_OK : HC_ENOTCONV);
  all.solve_ms = ms;
  if (stats)
    *stats = all;
  if (status)
    return fail(status, "hc_pcg_solve (multigrid): NaN met in a reduction");
  if (!all.converged)
    return fail(HC_ENOTCONV, "hc_pcg_solve (multigrid): not every right-hand side converged, true relative residual " +
                                 std::to_string(all.rel_res));
  return HC_OK;
}

int
mg_levels(hc_op_t op, hipStream_t st, int * nlev, int64_t * ncoarse)
{
  if (int rc = mg_build(op, st))
    return rc;
  *nlev = (int)op->mg->lev.size();
  *ncoarse = op->mg->NL;
  return HC_OK;
}


// ---------------------------------------------------------------------------
// Node-space multigrid for the AV formulation's scalar-potential block
// S = G^T M_sigma G (M_sigma: the operator's mass term without PEC
// elimination).  Its levels are the node spaces of the ND hierarchy's levels
// (Q_p Lagrange on the closed points, nested like the ND spaces), transfers
// are closed 1-D interpolations along every axis (the nodal counterpart of
// the ND transfers: G_f P_nodes = P_ND G_c), the smoother is the Chebyshev
// iteration with Jacobi, and the coarsest level, singular with the constants
// as kernel (V has no essential nodes), is solved by the dense inverse of
// S_c + alpha 1 1^T (alpha = trace(S_c) / n_c^2), which inverts S_c on the
// complement of the constants.
// ---------------------------------------------------------------------------

struct NodeLevel
{
  hc_op_t op = nullptr; // the ND level's operator (borrowed)
  Geom g0{};            // its geometry without PEC
  int64_t NN = 0, N = 0;
  double *x = nullptr, *b = nullptr, *r = nullptr, *z = nullptr, *d = nullptr, *dinv = nullptr;
  double *t1 = nullptr, *t2 = nullptr; // ND-size temporaries of the S apply
  double lam = 1.0;
  Csr1Dev P[3], R[3]; // from the next coarser level, per axis
};

struct NodeHier
{
  std::vector<NodeLevel> lev;
  DevBlock mem, sinv_mem;
  double * Sinv = nullptr;
  int64_t NC = 0;
};

void
node_free(NodeHier * n)
{
  delete n;
}

namespace
{

constexpr int64_t NODE_COARSE_MAX = 4096; // largest coarsest node level (dense inverse)
constexpr double NODE_CHEB_RANGE = 20.0;

__global__ void
k_add_const(double * __restrict__ A, int64_t n, double v)
{
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    A[k] += v;
}

__global__ void
k_trace_part(const double * __restrict__ A, int64_t n, double * __restrict__ out)
{
  // one workgroup, fixed order
  __shared__ double sh[256];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256)
    s += A[i * n + i];
  s = group_sum<256>(s, sh);
  if (threadIdx.x == 0)
    out[0] = s;
}

// y = S x on a node level
int
node_apply(NodeLevel & L, const double * x, double * y, hipStream_t st)
{
  hc_op_t op = L.op;
  HC_TRY_HIP(grad_apply(op->ag, x, L.t1, st));
  HC_TRY_HIP(apply_scaled<double>(L.g0, op->plan64.segs, op->tb64, op->qd64, op->ef(false), L.t1, L.t2, op->ws.part64, 1.0, 0.0,
                                  op->plan64.cand, op->desc_dev, st));
  HC_TRY_HIP(grad_transpose(op->ag, L.t2, y, st));
  return HC_OK;
}

// y (level l) += P x (level l+1) or y (level l+1) = P^T x (level l)
hipError_t
node_transfer(MgHier & h, NodeHier & nh, int l, bool prolong, const double * x, double * y, hipStream_t st)
{
  const NodeLevel & f = nh.lev[l];
  const NodeLevel & c = nh.lev[l + 1];
  int d[3];
  for (int a = 0; a < 3; a++)
    d[a] = prolong ? c.op->ag.M[a] : f.op->ag.M[a];
  for (int a = 0; a < 3; a++)
  {
    const Csr1Dev & m = prolong ? f.P[a] : f.R[a];
    double * out = (a == 2) ? y : (a == 0 ? h.tmp1 : h.tmp2);
    const double * in = (a == 0) ? x : (a == 1 ? h.tmp1 : h.tmp2);
    const int64_t plane = (int64_t)(a == 0 ? m.rows : d[0]) * (a == 1 ? m.rows : d[1]);
    const int o2 = a == 2 ? m.rows : d[2];
    hipLaunchKernelGGL(k_pass1d, dim3((unsigned)((plane + 255) / 256), (unsigned)o2), dim3(256), 0, st, in, out, d[0],
                       d[1], d[2], a, m.rows, m.rp, m.ci, m.w, 1.0, (a == 2 && prolong) ? 1 : 0, (int64_t)0,
                       (int64_t)0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
      return e;
    d[a] = m.rows;
  }
  return hipSuccess;
}

int
node_smooth(MgHier & h, NodeLevel & L, bool zero_start, hipStream_t st)
{
  const double hi = 1.1 * L.lam, lo = hi / NODE_CHEB_RANGE;
  const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma1 = theta / delta;
  double rho0 = 1.0 / sigma1;
  for (int s = 0; s < h.nsmooth; s++)
  {
    const bool first = zero_start && s == 0;
    if (first)
      HC_TRY_HIP(hipMemcpyAsync(L.r, L.b, sizeof(double) * L.NN, hipMemcpyDeviceToDevice, st));
    else
    {
      if (int rc = node_apply(L, L.x, L.r, st))
        return rc;
      HC_TRY_HIP(axpby(L.NN, 1.0, L.b, -1.0, L.r, st));
    }
    HC_TRY_HIP(mul(L.NN, L.dinv, L.r, L.z, st));
    double a, b;
    if (s == 0)
    {
      a = 1.0 / theta;
      b = 0.0;
    }
    else
    {
      const double rho1 = 1.0 / (2.0 * sigma1 - rho0);
      a = 2.0 * rho1 / delta;
      b = rho1 * rho0;
      rho0 = rho1;
    }
    hipLaunchKernelGGL(k_cheb_update, dim3(grid_for(L.NN)), dim3(256), 0, st, L.NN, a, b, L.z, L.d, L.x,
                       (s == 0 ? 1 : 0) | (first ? 2 : 0));
    HC_TRY_HIP(hipGetLastError());
  }
  return HC_OK;
}

int
node_vcycle(MgHier & h, NodeHier & nh, int l, hipStream_t st)
{
  NodeLevel & L = nh.lev[l];
  if (l == (int)nh.lev.size() - 1)
  {
    hipLaunchKernelGGL(k_dense_symv, dim3((unsigned)L.NN), dim3(256), 0, st, (const double *)nh.Sinv, L.NN,
                       (const double *)L.b, L.x);
    HC_TRY_HIP(hipGetLastError());
    return HC_OK;
  }
  if (int rc = node_smooth(h, L, true, st))
    return rc;
  if (int rc = node_apply(L, L.x, L.r, st))
    return rc;
  HC_TRY_HIP(axpby(L.NN, 1.0, L.b, -1.0, L.r, st));
  HC_TRY_HIP(node_transfer(h, nh, l, false, L.r, nh.lev[l + 1].b, st));
  if (int rc = node_vcycle(h, nh, l + 1, st))
    return rc;
  HC_TRY_HIP(node_transfer(h, nh, l, true, nh.lev[l + 1].x, L.x, st));
  return node_smooth(h, L, false, st);
}

// Node hierarchy over the ND levels (the ND hierarchy must be built and have a
// dense coarsest level); HC_OK when built, a status when not available.
int
node_build(MgHier & h, hipStream_t st)
{
  if (h.coarse_iter)
    return fail(HC_EINVAL, "multigrid: no node hierarchy for an inexact coarsest level");
  const int nl = (int)h.lev.size();
  const int64_t NC = h.lev.back().op->ag.NN;
  if (NC > NODE_COARSE_MAX)
    return fail(HC_EINVAL, "multigrid: coarsest node level too large for its dense inverse");
  auto * nh = new NodeHier();
  std::unique_ptr<NodeHier> guard(nh);
  size_t bytes = 0;
  std::vector<Csr1> mats;
  for (int l = 0; l < nl; l++)
  {
    NodeLevel L;
    L.op = h.lev[l].op;
    L.g0 = L.op->plan64.g;
    L.g0.pec = 0;
    L.NN = L.op->ag.NN;
    L.N = L.op->N;
    if (int rc = L.op->ensure_ws(1, false))
      return rc;
    nh->lev.push_back(L);
    bytes += 6 * align256(sizeof(double) * L.NN) + 2 * align256(sizeof(double) * L.N);
  }
  for (int l = 0; l + 1 < nl; l++)
    for (int a = 0; a < 3; a++)
    {
      const MgLevel & f = h.lev[l];
      const MgLevel & c = h.lev[l + 1];
      Csr1 P = prolong_1d(c.n[a], c.op->p, f.n[a], f.op->p, true);
      Csr1 R = transpose(P);
      bytes += csr_bytes(P) + csr_bytes(R);
      mats.push_back(std::move(P));
      mats.push_back(std::move(R));
    }
  if (int rc = nh->mem.alloc(bytes))
    return rc;
  char * q = (char *)nh->mem.base;
  size_t mi = 0;
  for (int l = 0; l + 1 < nl; l++)
    for (int a = 0; a < 3; a++)
    {
      nh->lev[l].P[a] = upload(mats[mi++], q);
      nh->lev[l].R[a] = upload(mats[mi++], q);
    }
  for (auto & L : nh->lev)
  {
    double ** v[6] = {&L.x, &L.b, &L.r, &L.z, &L.d, &L.dinv};
    for (double ** pp : v)
    {
      *pp = (double *)q;
      q += align256(sizeof(double) * L.NN);
    }
    L.t1 = (double *)q;
    q += align256(sizeof(double) * L.N);
    L.t2 = (double *)q;
    q += align256(sizeof(double) * L.N);
    HC_TRY_HIP(aux_diag_inv(L.g0, L.op->ag, L.op->tb64, L.op->qd64, L.dinv, L.dinv, nullptr, st));
  }
  // lambda_max(D^-1 S) per smoothed level by power iteration
  for (int l = 0; l + 1 < nl; l++)
  {
    NodeLevel & L = nh->lev[l];
    HC_TRY_HIP(fill_uniform(nullptr, 93, L.x, L.NN, st));
    double lam = 1.0;
    for (int it = 0; it < MG_POWER_ITERS; it++)
    {
      double ab, nn;
      HC_TRY_HIP(dot2(h, L.x, L.x, L.NN, ab, nn, st));
      HC_TRY_HIP(axpby(L.NN, 1.0 / std::sqrt(std::max(nn, 1e-300)), L.x, 0.0, L.x, st));
      if (int rc = node_apply(L, L.x, L.r, st))
        return rc;
      HC_TRY_HIP(mul(L.NN, L.dinv, L.r, L.z, st));
      double xr, xx, xz, zz;
      HC_TRY_HIP(dot2(h, L.x, L.r, L.NN, xr, xx, st)); // (x, S x)
      HC_TRY_HIP(dot2(h, L.r, L.z, L.NN, xz, zz, st)); // (S x, D^-1 S x)
      lam = xr > 0 ? xz / xr : 1.0;
      HC_TRY_HIP(hipMemcpyAsync(L.x, L.z, sizeof(double) * L.NN, hipMemcpyDeviceToDevice, st));
    }
    L.lam = lam;
  }
  // coarsest: dense S_c by probing, + alpha 1 1^T, Cholesky inverse
  {
    NodeLevel & C = nh->lev.back();
    const int64_t n = C.NN;
    if (int rc = nh->sinv_mem.alloc(sizeof(double) * (n * n + n + 64)))
      return rc;
    nh->Sinv = (double *)nh->sinv_mem.base;
    double * e = nh->Sinv + n * n;
    int * info = (int *)(e + n);
    double * tr = (double *)(info + 16);
    for (int64_t j = 0; j < n; j++)
    {
      hipLaunchKernelGGL(k_identity_cols, dim3(grid_for(n)), dim3(256), 0, st, e, n, (int)j, 1);
      HC_TRY_HIP(hipGetLastError());
      if (int rc = node_apply(C, e, nh->Sinv + j * n, st))
        return rc;
    }
    hipLaunchKernelGGL(k_trace_part, dim3(1), dim3(256), 0, st, (const double *)nh->Sinv, n, tr);
    double htr = 0.0;
    HC_TRY_HIP(hipMemcpyAsync(&htr, tr, sizeof(double), hipMemcpyDeviceToHost, st));
    HC_TRY_HIP(hipStreamSynchronize(st));
    if (!(htr > 0.0) || !std::isfinite(htr))
      return fail(HC_EINVAL, "multigrid: node coarsest level has no positive trace");
    hipLaunchKernelGGL(k_add_const, dim3(grid_for(n * n)), dim3(256), 0, st, nh->Sinv, n * n, htr / ((double)n * n));
    HC_TRY_HIP(hipGetLastError());
    if (rocblas_set_stream(h.blas, st) != rocblas_status_success ||
        rocsolver_dpotrf(h.blas, rocblas_fill_lower, (rocblas_int)n, nh->Sinv, (rocblas_int)n, info) !=
            rocblas_status_success ||
        rocsolver_dpotri(h.blas, rocblas_fill_lower, (rocblas_int)n, nh->Sinv, (rocblas_int)n, info + 1) !=
            rocblas_status_success)
      return fail(HC_EHIP, "multigrid: rocsolver Cholesky of the node coarsest level failed");
    int hinfo[2] = {0, 0};
    HC_TRY_HIP(hipMemcpyAsync(hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost, st));
    HC_TRY_HIP(hipStreamSynchronize(st));
    if (hinfo[0] || hinfo[1])
      return fail(HC_EHIP, "multigrid: node coarsest matrix is not positive definite");
    hipLaunchKernelGGL(k_symmetrize_lower, dim3(grid_for(n * n)), dim3(256), 0, st, nh->Sinv, n);
    HC_TRY_HIP(hipGetLastError());
    nh->NC = n;
  }
  h.nodes = guard.release();
  return HC_OK;
}

} // namespace

// Whether the AV block preconditioner diag(V-cycle of A, node V-cycle of S)
// is available for op (builds both hierarchies on first use).
bool
mg_av_ready(hc_op_t op, hipStream_t st)
{
  if (!op->box || mg_build(op, st) != HC_OK)
    return false;
  MgHier & h = *op->mg;
  if (h.nodes_state == 0)
    h.nodes_state = node_build(h, st) == HC_OK ? 1 : -1;
  return h.nodes_state == 1;
}

int
mg_node_precond(hc_op_t op, const double * r, double * z, hipStream_t st)
{
  MgHier & h = *op->mg;
  NodeLevel & N0 = h.nodes->lev[0];
  double * const nb0 = N0.b;
  double * const nx0 = N0.x;
  N0.b = (double *)r;
  N0.x = z;
  const int rc = node_vcycle(h, *h.nodes, 0, st);
  N0.b = nb0;
  N0.x = nx0;
  return rc;
}

// za = V-cycle of the ND system on ra, zv = node V-cycle of S on rv (both
// fixed linear operators: a symmetric positive definite block preconditioner)
int
mg_av_precond(hc_op_t op, const double * ra, const double * rv, double * za, double * zv, hipStream_t st)
{
  MgHier & h = *op->mg;
  MgLevel & L = h.lev[0];
  double * const b0 = L.b;
  double * const x0 = L.x;
  L.b = (double *)ra;
  L.x = za;
  int rc = vcycle(h, 0, st);
  L.b = b0;
  L.x = x0;
  if (rc)
    return rc;
  NodeLevel & N0 = h.nodes->lev[0];
  double * const nb0 = N0.b;
  double * const nx0 = N0.x;
  N0.b = (double *)rv;
  N0.x = zv;
  rc = node_vcycle(h, *h.nodes, 0, st);
  N0.b = nb0;
  N0.x = nx0;
  return rc;
}

} // namespace hc

extern "C" {

int
hc_mg_set_graphs(int enable)
try
{
  hc::g_mg_graphs.store(enable ? 1 : 0);
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_mg_levels(hc_op_t op, int * nlevels, int64_t * coarse_ndofs, void * stream)
try
{
  if (!op || !nlevels || !coarse_ndofs)
    return hc::fail(HC_EINVAL, "hc_mg_levels: null argument");
  if (!op->setup_done)
    return hc::fail(HC_EINVAL, "hc_mg_levels: hc_op_setup has not been called");
  return hc::mg_levels(op, (hipStream_t)stream, nlevels, coarse_ndofs);
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_mg_level_op(hc_op_t op, int level, hc_op_t * level_op)
try
{
  if (!op || !level_op)
    return hc::fail(HC_EINVAL, "hc_mg_level_op: null argument");
  if (!op->mg || op->mg->version != op->setup_version)
    return hc::fail(HC_EINVAL, "hc_mg_level_op: no current hierarchy (call hc_mg_levels first)");
  if (level < 0 || level >= (int)op->mg->lev.size())
    return hc::fail(HC_EINVAL, "hc_mg_level_op: level out of range");
  *level_op = op->mg->lev[level].op;
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

int
hc_mg_transfer(hc_op_t op, int level, int dir, const double * x, double * y, void * stream)
try
{
  if (!op || !x || !y)
    return hc::fail(HC_EINVAL, "hc_mg_transfer: null argument");
  if (!op->mg || op->mg->version != op->setup_version)
    return hc::fail(HC_EINVAL, "hc_mg_transfer: no current hierarchy (call hc_mg_levels first)");
  if (level < 0 || level + 1 >= (int)op->mg->lev.size())
    return hc::fail(HC_EINVAL, "hc_mg_transfer: level must be below the coarsest");
  if (dir != 0 && dir != 1)
    return hc::fail(HC_EINVAL, "hc_mg_transfer: dir must be 0 (restrict) or 1 (prolong)");
  hipStream_t st = (hipStream_t)stream;
  if (dir == 1)
    HC_TRY_HIP(hipMemsetAsync(y, 0, sizeof(double) * op->mg->lev[level].N, st));
  HC_TRY_HIP(hc::transfer(*op->mg, level, dir == 1, x, y, st));
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

} // extern "C"
