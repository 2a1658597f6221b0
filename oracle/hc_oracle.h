/*
 * hc_oracle.h -- CPU FP64 restatement of the H(curl) curl-curl + sigma-mass
 * operator path, used ONLY as test infrastructure (the parity checker) and as
 * the CPU baseline leg of bench.py.  Nothing in the product library links it.
 *
 * What it restates (reference file:line, paths relative to the hephaestus
 * source root):
 *   - the bilinear form a(du,w) = (dt*alpha curl du, curl w) + (beta du, w)
 *     that CurlCurlEquationSystem::AddKernels registers
 *     (src/formulations/HCurl/hcurl_formulation.cpp:130-157; the same block in
 *     src/formulations/AV/av_formulation.cpp:159-173), built from
 *     CurlCurlKernel (src/kernels/curl_curl_kernel.cpp:20-24) and
 *     VectorFEMassKernel (src/kernels/vector_fe_mass_kernel.cpp:20-24);
 *   - the PEC elimination of EquationSystem::FormLinearSystem
 *     (src/equation_systems/equation_system.cpp:132-190) with essential
 *     tangential Dirichlet DOFs from BCMap::ApplyEssentialBCs
 *     (src/boundary_conditions/boundary_conditions.cpp:32-51);
 *   - the operator apply EquationSystem::Mult (equation_system.cpp:200-204);
 *   - DefaultJacobiPCGSolver (src/solvers/hephaestus_solvers.hpp:34-57):
 *     PCG preconditioned by the inverse diagonal.
 * The element arithmetic itself lives in MFEM (CurlCurlIntegrator,
 * VectorFEMassIntegrator on ND_HexahedronElement) and hypre (PCG,
 * DiagScale); neither is available, so this file restates the published
 * mathematics of those objects: open Gauss-Legendre x closed Gauss-Lobatto
 * tensor bases per vector component, covariant Piola map, exact (Q = p+1)
 * Gauss quadrature.
 *
 * Conventions fixed here (the reference does not pin them):
 *   - reference element [0,1]^3, quadrature weights sum to 1;
 *   - structured global numbering per component, all edges/faces oriented
 *     along +axes (no sign flips);
 *   - qdata layout [e][qz][k][qy][qx], k = 0..5 mass (xx,xy,xz,yy,yz,zz),
 *     k = 6..11 curl (same order);
 *   - essential DOFs: y_ess = x_ess, diag_ess = 1, x_ess treated as 0 in the
 *     element products (A_ess,ess = I, eliminated rows/cols).
 */
#ifndef HC_ORACLE_H
#define HC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HCO_PMAX 4

typedef struct
{
  int nx, ny, nz;      /* elements per direction */
  const double * vert; /* (nx+1)(ny+1)(nz+1)*3 vertex coords, x fastest; NULL => box */
  double x0[3];        /* box origin (vert == NULL) */
  double len[3];       /* box edge lengths (vert == NULL) */
  int pec_faces;       /* bit0 x-min, bit1 x-max, bit2 y-min, bit3 y-max, bit4 z-min, bit5 z-max */
} hco_mesh;

/* 1D tables for order p: Q = p+1 Gauss points on (0,1).
 * Bo[q*p + d], Bc[q*(p+1) + d], Gc[q*(p+1) + d], qpts[q], qw[q]. */
int hco_basis(int p, double * Bo, double * Bc, double * Gc, double * qpts, double * qw);
/* Gauss-Legendre (n points) and Gauss-Lobatto (n points) on [0,1]. */
void hco_gauss_legendre(int n, double * x, double * w);
void hco_gauss_lobatto(int n, double * x);

int64_t hco_ndofs(const hco_mesh * m, int p);
void hco_comp_dims(const hco_mesh * m, int p, int c, int64_t dims[3]);
void hco_ess_mask(const hco_mesh * m, int p, uint8_t * ess);
/* Element-local -> global DOF map (nd entries) of element e. */
void hco_elem_dofs(const hco_mesh * m, int p, int64_t e, int64_t * gid);

/* Partial-assembly setup: qdata[E*Q^3*12]. */
void hco_pa_setup(const hco_mesh * m,
                  int p,
                  const double * curl_coef,
                  const double * mass_coef,
                  double * qdata);
/* y = A x with PEC constraint; sum-factorised, element colouring (deterministic). */
void hco_pa_apply(const hco_mesh * m,
                  int p,
                  const double * qdata,
                  const uint8_t * ess,
                  const double * x,
                  double * y);
/* Same operator, no PEC handling and no colouring: y += sum_e R_e^T A_e R_e x
 * over the element range [e0, e1) (used by the element-matrix tests). */
void hco_pa_apply_raw(const hco_mesh * m,
                      int p,
                      const double * qdata,
                      int64_t e0,
                      int64_t e1,
                      const double * x,
                      double * y);
void hco_pa_diagonal(const hco_mesh * m,
                     int p,
                     const double * qdata,
                     const uint8_t * ess,
                     double * diag);

/* Dense element matrix of element e (nd x nd, row-major) by full quadrature of
 * explicitly evaluated basis functions (no sum factorisation). */
void hco_elem_matrix(const hco_mesh * m, int p, const double * qdata, int64_t e, double * Ae);

/* Full assembly into CSR with PEC elimination (what the reference executes:
 * ParBilinearForm::Assemble + FormLinearSystem).  Arrays are malloc'd; free
 * with hco_csr_free. */
typedef struct
{
  int64_t n, nnz;
  int64_t * rowptr;
  int64_t * col;
  double * val;
} hco_csr;
int hco_csr_assemble(const hco_mesh * m, int p, const double * qdata, const uint8_t * ess, hco_csr * A);
void hco_csr_free(hco_csr * A);
void hco_csr_apply(const hco_csr * A, const double * x, double * y);

/* Jacobi-PCG.  Stopping rule ||r_k||_2 <= rtol ||b||_2 on the recursive
 * residual; returns iterations (or -1 if not converged in maxit).  res_hist
 * (optional, length maxit+1) receives ||r_k||/||b||. */
typedef struct
{
  int iterations;
  int converged;
  double rel_res;      /* recursive residual at exit */
  double true_rel_res; /* ||b - A x|| / ||b||, recomputed with the operator */
} hco_pcg_stats;
int hco_pcg_pa(const hco_mesh * m,
               int p,
               const double * qdata,
               const uint8_t * ess,
               const double * diag,
               const double * b,
               double * x,
               double rtol,
               int maxit,
               hco_pcg_stats * st,
               double * res_hist);
int hco_pcg_csr(const hco_csr * A,
                const double * diag,
                const double * b,
                double * x,
                double rtol,
                int maxit,
                hco_pcg_stats * st,
                double * res_hist);

/* Counter-based splitmix64: out[i] = i-th output of splitmix64 seeded with
 * `seed`, mapped to [-1,1) as 2*((u >> 11) * 2^-53) - 1. */
void hco_rand_uniform(uint64_t seed, int64_t n, double * out);
/* RHS rule "uniform per free true DOF": b_i = rand_i for free DOFs, 0 on ess. */
void hco_rhs_uniform(const hco_mesh * m, int p, uint64_t seed, const uint8_t * ess, double * b);
/* Coil-like divergence-free current J = curl psi, psi = sum_k a_k exp(-|x-c_k|^2/(2 s_k^2)),
 * b_i = int J . phi_i with Gauss (p+2)^3 quadrature; ess entries 0. */
void hco_rhs_coil(const hco_mesh * m, int p, uint64_t seed, const uint8_t * ess, double * b);

int hco_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
