// Geometric multigrid preconditioner for the H(curl) system A = dt K_nu + M_sigma
// on structured box meshes, and the PCG solve that uses it (HC_SOLVE_MG).
//
// The reference preconditions this system with HypreAMS
// (src/formulations/HCurl/hcurl_formulation.cpp:69-79, AMS in singular mode):
// an auxiliary-space method whose node-space (gradient) and vector-node-space
// corrections are multigrid cycles.  This is the geometric counterpart on the
// nested structured meshes: a V-cycle over the meshes n, n/2, n/4, ... with
// the same ND_p space on each, the hybrid smoother of Hiptmair on every level
// (damped Richardson with B = D^-1 + G Dg^-1 G^T: Jacobi on the edges plus
// Jacobi on the node space through the discrete gradient, the correction that
// AMS's gradient part performs), and a dense Cholesky solve on the coarsest
// level.
//
// Levels: level 0 is the caller's operator; level l+1 halves every element
// count of level l (all must be even), with the coefficients averaged over the
// eight children, the same order, PEC faces and dt.  Transfers: the ND spaces
// are nested, so prolongation evaluates the coarse field at the fine DOF
// functionals -- per component a tensor product of 1-D interpolations (open
// Gauss-Legendre basis along the component's axis, closed Gauss-Lobatto basis
// along the others) times h_f / h_c = 1/2 (covariant scaling of the
// tangential DOF); restriction is its transpose.  Both run as three 1-D passes.
// The V-cycle (nu_1 = nu_2 smoothing steps with the same damped symmetric
// smoother, exact coarse solve) is a fixed symmetric positive definite
// operator, so plain PCG applies.
#include "hc_op.hpp"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

namespace hc
{

namespace
{

constexpr int MG_SMOOTH = 2;            // pre- and post-smoothing steps
constexpr int64_t MG_COARSE_TARGET = 4000; // stop coarsening at this many DOFs
constexpr int64_t MG_COARSE_MAX = 12000;   // largest coarsest level (dense inverse)
constexpr int MG_POWER_ITERS = 15;      // power iterations for the smoother's lambda_max
constexpr int DOT_BLOCKS = 512;
constexpr int DOT_THREADS = 256;

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// One 1-D pass of a tensor-product transfer: out = scale * M (along axis) in,
// + out when accumulate.  in: dims d[0..2] (x fastest); out: the same dims but
// d[axis] -> m (the rows of M, CSR rp/ci/w).
__global__ void
k_pass1d(const double * __restrict__ in, double * __restrict__ out, int d0, int d1, int d2, int axis, int m,
         const int * __restrict__ rp, const int * __restrict__ ci, const double * __restrict__ w, double scale,
         int accumulate)
{
  const int o0 = axis == 0 ? m : d0, o1 = axis == 1 ? m : d1, o2 = axis == 2 ? m : d2;
  const int64_t total = (int64_t)o0 * o1 * o2;
  for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < total; o += (int64_t)gridDim.x * blockDim.x)
  {
    const int i0 = (int)(o % o0);
    const int i1 = (int)((o / o0) % o1);
    const int i2 = (int)(o / ((int64_t)o0 * o1));
    const int row = axis == 0 ? i0 : (axis == 1 ? i1 : i2);
    double s = 0.0;
    for (int k = rp[row]; k < rp[row + 1]; k++)
    {
      const int j = ci[k];
      const int j0 = axis == 0 ? j : i0, j1 = axis == 1 ? j : i1, j2 = axis == 2 ? j : i2;
      s += w[k] * in[((int64_t)j2 * d1 + j1) * d0 + j0];
    }
    s *= scale;
    out[o] = accumulate ? out[o] + s : s;
  }
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
  sh[0][threadIdx.x] = s0;
  sh[1][threadIdx.x] = s1;
  __syncthreads();
  for (int s = DOT_THREADS / 2; s > 0; s >>= 1)
  {
    if ((int)threadIdx.x < s)
    {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
  {
    part[2 * blockIdx.x] = sh[0][0];
    part[2 * blockIdx.x + 1] = sh[1][0];
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

struct Csr1
{
  std::vector<int> rp, ci;
  std::vector<double> w;
  int rows = 0, cols = 0;
};

// Prolongation of 1-D DOFs from nc coarse to 2 nc fine elements: open (p
// Gauss-Legendre points per element) or closed (p+1 Gauss-Lobatto, shared at
// element ends).  Row = fine DOF, entries = coarse basis values at its point.
Csr1
prolong_1d(int nc, int p, bool closed, const double * gl, const double * gll)
{
  Csr1 m;
  const int nf = 2 * nc;
  const int k = closed ? p + 1 : p; // local nodes per element
  const double * nodes = closed ? gll : gl;
  m.rows = closed ? nf * p + 1 : nf * p;
  m.cols = closed ? nc * p + 1 : nc * p;
  m.rp.push_back(0);
  for (int gf = 0; gf < m.rows; gf++)
  {
    const int ef = std::min(gf / p, nf - 1);
    const int jf = gf - ef * p;
    const int ec = ef / 2, s = ef % 2;
    const double t = (s + nodes[jf]) / 2.0; // position in the coarse element
    for (int jc = 0; jc < k; jc++)
    {
      const double v = lagrange1(k, nodes, jc, t);
      m.ci.push_back(ec * p + jc);
      m.w.push_back(v);
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

// Gauss-Legendre / Gauss-Lobatto points on [0,1] (Newton on P_n / P'_{n-1})
void
points01(int n, bool lobatto, double * x)
{
  auto leg = [](int k, double z, double & P, double & dP) {
    double p0 = 1.0, p1 = z;
    if (k == 0)
    {
      P = 1.0;
      dP = 0.0;
      return;
    }
    for (int j = 2; j <= k; j++)
    {
      const double p2 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p0) / j;
      p0 = p1;
      p1 = p2;
    }
    P = p1;
    dP = k * (z * p1 - p0) / (z * z - 1.0);
  };
  if (!lobatto)
  {
    for (int i = 0; i < n; i++)
    {
      double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), P, dP;
      for (int it = 0; it < 100; it++)
      {
        leg(n, z, P, dP);
        const double dz = P / dP;
        z -= dz;
        if (std::fabs(dz) < 1e-16)
          break;
      }
      x[n - 1 - i] = 0.5 * (z + 1.0);
    }
    return;
  }
  const int p = n - 1;
  x[0] = 0.0;
  x[p] = 1.0;
  for (int i = 1; i < p; i++)
  {
    double z = -std::cos(M_PI * i / p), P, dP;
    for (int it = 0; it < 100; it++)
    {
      leg(p, z, P, dP);
      const double d2P = (2.0 * z * dP - p * (p + 1.0) * P) / (1.0 - z * z);
      const double dz = dP / d2P;
      z -= dz;
      if (std::fabs(dz) < 1e-16)
        break;
    }
    x[i] = 0.5 * (z + 1.0);
  }
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
  double *x = nullptr, *b = nullptr, *r = nullptr, *z = nullptr;
  double omega = 1.0;
  // transfers to this level from the next coarser one: [axis][open=0/closed=1]
  Csr1Dev P[3][2], R[3][2];
};

struct MgHier
{
  std::vector<MgLevel> lev;
  int version = -1;
  DevBlock mem;        // level vectors, temporaries, transfer matrices
  double * tmp1 = nullptr;
  double * tmp2 = nullptr;
  double * zz = nullptr;   // preconditioned residual of the PCG
  double * part = nullptr; // DOT_BLOCKS * 2
  double * hpart = nullptr; // pinned copy
  double * Ainv = nullptr; // coarsest level, dense inverse (lower triangle)
  DevBlock ainv_mem;
  int64_t NL = 0;
  rocblas_handle blas = nullptr;
  ~MgHier()
  {
    for (auto & l : lev)
      if (l.owned && l.op)
        hc_op_destroy(l.op);
    if (hpart)
      (void)hipHostFree(hpart);
    if (blas)
      rocblas_destroy_handle(blas);
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
transfer(MgHier & h, int l, bool prolong, const double * x, double * y, hipStream_t st)
{
  const MgLevel & f = h.lev[l];
  const MgLevel & c = h.lev[l + 1];
  const int p = f.op->p;
  int64_t offf = 0, offc = 0;
  for (int comp = 0; comp < 3; comp++)
  {
    int df[3], dc[3];
    comp_dims(f.n, p, comp, df);
    comp_dims(c.n, p, comp, dc);
    const double * src = x + (prolong ? offc : offf);
    double * dst = y + (prolong ? offf : offc);
    int d[3];
    std::memcpy(d, prolong ? dc : df, sizeof(d));
    for (int a = 0; a < 3; a++)
    {
      const int kind = (a == comp) ? 0 : 1;
      const Csr1Dev & m = prolong ? f.P[a][kind] : f.R[a][kind];
      double * out = (a == 2) ? dst : (a == 0 ? h.tmp1 : h.tmp2);
      const double * in = (a == 0) ? src : (a == 1 ? h.tmp1 : h.tmp2);
      int64_t total = (int64_t)m.rows;
      for (int q = 0; q < 3; q++)
        if (q != a)
          total *= d[q];
      hipLaunchKernelGGL(k_pass1d, dim3(grid_for(total)), dim3(256), 0, st, in, out, d[0], d[1], d[2], a, m.rows,
                         m.rp, m.ci, m.w, a == 2 ? 0.5 : 1.0, (a == 2 && prolong) ? 1 : 0);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess)
        return e;
      d[a] = m.rows;
    }
    offf += (int64_t)df[0] * df[1] * df[2];
    offc += (int64_t)dc[0] * dc[1] * dc[2];
  }
  if (!prolong)
    return zero_ess(c.op->g, y, st);
  return hipSuccess;
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
  if (int rc = hc_op_apply(L.op, x, r, 1, st))
    return rc;
  HC_TRY_HIP(axpby(L.N, 1.0, b, -1.0, r, st));
  return HC_OK;
}

// x_l <- V-cycle(b_l) with x_l starting from zero
int
vcycle(MgHier & h, int l, hipStream_t st)
{
  MgLevel & L = h.lev[l];
  if (l == (int)h.lev.size() - 1)
  {
    const double one = 1.0, zero = 0.0;
    if (rocblas_set_stream(h.blas, st) != rocblas_status_success ||
        rocblas_dsymv(h.blas, rocblas_fill_lower, (rocblas_int)L.N, &one, h.Ainv, (rocblas_int)L.N, L.b, 1, &zero,
                      L.x, 1) != rocblas_status_success)
      return fail(HC_EHIP, "multigrid: coarse solve (rocblas_dsymv) failed");
    return HC_OK;
  }
  // pre-smoothing from x = 0
  for (int s = 0; s < MG_SMOOTH; s++)
  {
    if (int rc = residual(L, L.b, s == 0 ? nullptr : L.x, L.r, st))
      return rc;
    if (int rc = hc_op_precond_apply(L.op, HC_SOLVE_FP64 | HC_SOLVE_AUX, L.r, L.z, st))
      return rc;
    HC_TRY_HIP(axpby(L.N, L.omega, L.z, s == 0 ? 0.0 : 1.0, L.x, st));
  }
  if (int rc = residual(L, L.b, L.x, L.r, st))
    return rc;
  HC_TRY_HIP(transfer(h, l, false, L.r, h.lev[l + 1].b, st));
  if (int rc = vcycle(h, l + 1, st))
    return rc;
  HC_TRY_HIP(transfer(h, l, true, h.lev[l + 1].x, L.x, st));
  for (int s = 0; s < MG_SMOOTH; s++)
  {
    if (int rc = residual(L, L.b, L.x, L.r, st))
      return rc;
    if (int rc = hc_op_precond_apply(L.op, HC_SOLVE_FP64 | HC_SOLVE_AUX, L.r, L.z, st))
      return rc;
    HC_TRY_HIP(axpby(L.N, L.omega, L.z, 1.0, L.x, st));
  }
  return HC_OK;
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

// the hierarchy of op (built on first use; rebuilt when op was set up again)
int
mg_build(hc_op_t op, hipStream_t st)
{
  if (op->mg && op->mg->version == op->setup_version)
    return HC_OK;
  mg_free(op->mg);
  op->mg = nullptr;
  if (!op->box)
    return fail(HC_EINVAL, "multigrid: needs a box mesh (nested coarse meshes)");
  MgHier * h = new MgHier();
  std::unique_ptr<MgHier> guard(h);
  const int p = op->p;
  MgLevel L0;
  L0.op = op;
  for (int a = 0; a < 3; a++)
    L0.n[a] = op->g.n[a];
  L0.N = op->N;
  h->lev.push_back(L0);
  std::vector<double> nu = op->nu, sig = op->sig;
  while (h->lev.back().N > MG_COARSE_TARGET)
  {
    const MgLevel & f = h->lev.back();
    if ((f.n[0] % 2) || (f.n[1] % 2) || (f.n[2] % 2) || f.n[0] < 2 || f.n[1] < 2 || f.n[2] < 2)
      break;
    MgLevel c;
    for (int a = 0; a < 3; a++)
      c.n[a] = f.n[a] / 2;
    const int64_t Ec = (int64_t)c.n[0] * c.n[1] * c.n[2];
    std::vector<double> nuc(Ec, 0.0), sigc(Ec, 0.0);
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
    int rc = hc_op_create(&d, p, nuc.data(), sigc.data(), HC_PREC_FP64, &c.op);
    if (rc)
      return rc;
    c.owned = true;
    h->lev.push_back(c);
    if ((rc = hc_op_setup(h->lev.back().op, op->last_dt, st)))
      return rc;
    h->lev.back().N = h->lev.back().op->N;
    nu.swap(nuc);
    sig.swap(sigc);
  }
  const int nl = (int)h->lev.size();
  // a single level (a mesh small enough already) is solved by the dense inverse alone
  h->NL = h->lev.back().N;
  if (h->NL > MG_COARSE_MAX)
    return fail(HC_EINVAL, "multigrid: coarsest level has " + std::to_string(h->NL) +
                               " DOFs (element counts must halve, while even, down to a mesh of at most " +
                               std::to_string(MG_COARSE_MAX) + " DOFs)");
  // transfer matrices
  double gl[PMAX + 1], gll[PMAX + 2];
  points01(p, false, gl);
  points01(p + 1, true, gll);
  std::vector<Csr1> mats; // per level pair: [axis][kind] P then R
  size_t bytes = 0;
  for (int l = 0; l + 1 < nl; l++)
    for (int a = 0; a < 3; a++)
      for (int kind = 0; kind < 2; kind++)
      {
        Csr1 P = prolong_1d(h->lev[l + 1].n[a], p, kind == 1, gl, gll);
        Csr1 R = transpose(P);
        bytes += csr_bytes(P) + csr_bytes(R);
        mats.push_back(std::move(P));
        mats.push_back(std::move(R));
      }
  // vectors: x, b, r, z per level; two temporaries of the finest size; dot partials
  for (auto & L : h->lev)
    bytes += 4 * align256(sizeof(double) * L.N);
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
    double ** v[4] = {&L.x, &L.b, &L.r, &L.z};
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
  for (int l = 0; l + 1 < nl; l++)
  {
    double lam = 1.0;
    if (int rc = estimate_lambda(*h, h->lev[l], st, lam))
      return rc;
    h->lev[l].omega = 1.0 / (1.05 * lam);
  }
  // coarsest level: dense matrix by probing, Cholesky inverse (lower triangle)
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
    for (int j0 = 0; j0 < n; j0 += nb)
    {
      const int k = (int)std::min<int64_t>(nb, n - j0);
      hipLaunchKernelGGL(k_identity_cols, dim3(grid_for(n * k)), dim3(256), 0, st, (double *)eblk.base, n, j0, k);
      HC_TRY_HIP(hipGetLastError());
      if (int rc = hc_op_apply(C.op, (const double *)eblk.base, h->Ainv + (int64_t)j0 * n, k, st))
        return rc;
    }
    if (rocblas_create_handle(&h->blas) != rocblas_status_success)
      return fail(HC_EHIP, "multigrid: rocblas_create_handle failed");
    if (rocblas_set_stream(h->blas, st) != rocblas_status_success)
      return fail(HC_EHIP, "multigrid: rocblas_set_stream failed");
    if (rocsolver_dpotrf(h->blas, rocblas_fill_lower, (rocblas_int)n, h->Ainv, (rocblas_int)n, info) !=
            rocblas_status_success ||
        rocsolver_dpotri(h->blas, rocblas_fill_lower, (rocblas_int)n, h->Ainv, (rocblas_int)n, info + 1) !=
            rocblas_status_success)
      return fail(HC_EHIP, "multigrid: rocsolver Cholesky of the coarsest level failed");
    int hinfo[2] = {0, 0};
    HC_TRY_HIP(hipMemcpyAsync(hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost, st));
    HC_TRY_HIP(hipStreamSynchronize(st));
    if (hinfo[0] || hinfo[1])
      return fail(HC_EHIP, "multigrid: coarsest-level matrix is not positive definite (info " +
                               std::to_string(hinfo[0]) + ", " + std::to_string(hinfo[1]) + ")");
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
             hipStream_t st)
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
  int iters = 0, outer = 0, status = HC_OK;
  bool conv = false;
  double rr = 0.0;
  auto precond = [&](const double * rin, double * zout) -> int {
    HC_TRY_HIP(hipMemcpyAsync(L.b, rin, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
    if (int rc = vcycle(h, 0, st))
      return rc;
    HC_TRY_HIP(hipMemcpyAsync(zout, L.x, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
    return HC_OK;
  };
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
      HC_TRY_HIP(axpby(N, 1.0, zz, gnew / gam, pv, st));
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
    s.outer_iterations = std::min(outer, 1 + MAX_RESTARTS);
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

int
mg_levels(hc_op_t op, hipStream_t st, int * nlev, int64_t * ncoarse)
{
  if (int rc = mg_build(op, st))
    return rc;
  *nlev = (int)op->mg->lev.size();
  *ncoarse = op->mg->NL;
  return HC_OK;
}

} // namespace hc

extern "C" {

int
hc_mg_levels(hc_op_t op, int * nlevels, int64_t * coarse_ndofs, void * stream)
{
  if (!op || !nlevels || !coarse_ndofs)
    return hc::fail(HC_EINVAL, "hc_mg_levels: null argument");
  if (!op->setup_done)
    return hc::fail(HC_EINVAL, "hc_mg_levels: hc_op_setup has not been called");
  return hc::mg_levels(op, (hipStream_t)stream, nlevels, coarse_ndofs);
}

int
hc_mg_level_op(hc_op_t op, int level, hc_op_t * level_op)
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

int
hc_mg_transfer(hc_op_t op, int level, int dir, const double * x, double * y, void * stream)
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

} // extern "C"
