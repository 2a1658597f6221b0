/*
 * hc_hcurl.h -- C ABI of the MI355X-native H(curl) curl-curl + sigma-mass
 * operator library (libhcurl.so).
 *
 * Reference interfaces these entry points replace (paths relative to the
 * hephaestus source root; MFEM/hypre calls marked [ext]):
 *
 *   hc_op_create    mfem::ParBilinearForm on the ND space + the
 *                   CurlCurlKernel / VectorFEMassKernel coefficient binding:
 *                   src/kernels/curl_curl_kernel.cpp:11-18 (Init resolves the
 *                   coefficient), src/kernels/vector_fe_mass_kernel.cpp:11-18,
 *                   src/equation_systems/equation_system.cpp:324-347
 *                   (BuildBilinearForms creates the form per test variable).
 *   hc_op_setup     ParBilinearForm::Assemble [ext] as driven by
 *                   EquationSystem::BuildBilinearForms (equation_system.cpp:345)
 *                   and TimeDependentEquationSystem::SetTimeStep
 *                   (equation_system.cpp:413-426: re-assemble when dt changes);
 *                   here a partial-assembly setup of quadrature-point factors.
 *   hc_op_apply     EquationSystem::Mult -> HypreParMatrix::Mult [ext]
 *                   (equation_system.cpp:200-204) on the system formed by
 *                   FormLinearSystem (equation_system.cpp:132-190): y = A x with
 *                   essential (PEC) rows/cols eliminated, A_ess,ess = I.
 *   hc_op_apply_f32 the same operator applied in FP32 (no reference
 *                   counterpart; the inner operator of FP32+FP64 refinement).
 *   hc_op_diagonal  HypreDiagScale's diagonal [ext]
 *                   (src/solvers/hephaestus_solvers.hpp:52).
 *   hc_pcg_solve    DefaultJacobiPCGSolver::Mult = HyprePCG + HypreDiagScale
 *                   [ext] (hephaestus_solvers.hpp:34-57), called from
 *                   TimeDomainEquationSystemProblemOperator::ImplicitSolve
 *                   (src/problem_operators/time_domain_equation_system_problem_operator.cpp:46-48).
 *   hc_op_destroy   ~ParBilinearForm [ext].
 *   hc_last_error   MFEM_ABORT's message (the reference aborts; this library
 *                   returns a negative status and keeps the message).
 *
 * Conventions
 *   - Every entry point returns HC_OK (0) or a negative HC_E* status and never
 *     aborts; hc_last_error() returns the calling thread's last message.
 *   - Vector arguments are DEVICE pointers (hipMalloc'd or torch-allocated);
 *     the mesh vertex and coefficient arrays of hc_op_create are HOST pointers.
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *   - A handle is not thread-safe: one host thread per handle at a time.
 *   - DOF numbering: component c = x,y,z has n_d*p points along its own axis
 *     (open, Gauss-Legendre) and n_d*p+1 along the others (closed,
 *     Gauss-Lobatto); x-component first, each component x-fastest.  All edges
 *     and faces are oriented along +axes (no sign flips).
 *   - Multi-RHS vectors: nrhs vectors of N entries stored one after the other
 *     (right-hand side r at offset r*N; column-major N x nrhs, leading
 *     dimension N).  Each right-hand side's solve is independent of the others
 *     in the batch: same iterations, same result bits.
 */
#ifndef HC_HCURL_H
#define HC_HCURL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HC_OK 0
#define HC_EINVAL (-1)   /* bad argument */
#define HC_ENOMEM (-2)   /* device (or host) allocation failed */
#define HC_EHIP (-3)     /* HIP runtime error */
#define HC_ENOTCONV (-4) /* solver did not reach the tolerance in maxit */
#define HC_ENAN (-5)     /* NaN/Inf met in a reduction */
#define HC_EBUSY (-6)    /* resource budget exceeded (admission refused) */

/* pec_faces bits: faces with essential tangential (PEC) conditions. */
#define HC_FACE_XMIN 0x01
#define HC_FACE_XMAX 0x02
#define HC_FACE_YMIN 0x04
#define HC_FACE_YMAX 0x08
#define HC_FACE_ZMIN 0x10
#define HC_FACE_ZMAX 0x20
#define HC_FACE_ALL 0x3F

/* precision flags of hc_op_create */
#define HC_PREC_FP64 0x1
#define HC_PREC_FP32 0x2

/* solve modes of hc_pcg_solve */
#define HC_SOLVE_FP64 0    /* plain FP64 Jacobi-PCG */
#define HC_SOLVE_FP32_IR 1 /* FP32 Jacobi-PCG inner solves + FP64 refinement; an
                              operator whose hc_op_term_ratio is below 6e-6 runs its
                              inner solves in FP64 (FP32 rounding noise would swamp
                              the sigma term there) and counts in fp64_fallbacks */
/* flag, or-ed into either mode: precondition with Jacobi plus the
 * discrete-gradient auxiliary-space correction B = D^-1 + G Dg^-1 G^T
 * (Dg = diag(G^T A G) on the H1 nodes; the one-level form of the HypreAMS
 * preconditioner that HCurlFormulation::ConstructJacobianPreconditioner sets,
 * src/formulations/HCurl/hcurl_formulation.cpp:69-79) instead of Jacobi alone */
#define HC_SOLVE_AUX 0x10
/* flag for hc_pcg_solve (with HC_SOLVE_FP64; box meshes): precondition with a
 * multigrid V-cycle.  Levels: the mesh halves while its element counts are
 * even and the coefficients are constant on each 2x2x2 block of children;
 * then the order drops to 1 on the same mesh; then, at order 1 and only
 * while a level exceeds 12000 DOFs, the mesh halves with the coefficients
 * averaged (same PEC faces and dt throughout).  On every level two pre- and
 * two post-smoothing steps of Chebyshev iteration with the hybrid smoother
 * D^-1 + G Dg^-1 G^T (edge Jacobi plus node Jacobi through the discrete
 * gradient), the transfers of the nested ND spaces, and a dense Cholesky
 * solve on the coarsest level; a coarsest level still above 12000 DOFs (odd
 * element counts at order 1) is approximated instead by 16 steps of the same
 * Chebyshev iteration on [lambda_max / 100, 1.1 lambda_max] (a fixed
 * polynomial, so the V-cycle stays a fixed SPD preconditioner), and a mesh
 * that has only such a level (order 1, odd counts, above 12000 DOFs) is
 * refused (HC_EINVAL).  With nrhs > 1 the right-hand sides go through one
 * multi-RHS V-cycle per iteration, each with its own PCG
 * scalars and stopping test; the results equal those of one solve each up
 * to round-off of the multi-RHS kernels.  The geometric counterpart, on structured meshes, of the
 * multigrid auxiliary-space preconditioner HypreAMS that
 * HCurlFormulation::ConstructJacobianPreconditioner sets
 * (src/formulations/HCurl/hcurl_formulation.cpp:69-79).  Built on first use
 * and kept with the operator until its next re-assembly. */
#define HC_SOLVE_MG 0x20

typedef struct hc_op_s * hc_op_t;
typedef struct hc_batch_s * hc_batch_t;

typedef struct
{
  int nx, ny, nz;     /* elements per direction */
  const double * vert; /* host; (nx+1)(ny+1)(nz+1)*3 coords, x fastest; NULL = box */
  double x0[3];       /* box origin (vert == NULL) */
  double len[3];      /* box edge lengths (vert == NULL) */
  int pec_faces;      /* HC_FACE_* bits */
} hc_mesh_desc;

typedef struct
{
  int iterations;       /* inner PCG iterations (summed over refinement steps or
                           restarts; max over the right-hand sides) */
  int outer_iterations; /* HC_SOLVE_FP32_IR: refinement steps; HC_SOLVE_FP64:
                           PCG runs (1 + residual-replacement restarts) */
  int converged;        /* 1 if every RHS met rtol */
  int status;           /* HC_OK or HC_ENOTCONV / HC_ENAN */
  double rel_res;       /* max over RHS of the final FP64 true relative residual */
  double solve_ms;      /* device time of the solve (hipEvents) */
  int fp64_fallbacks;   /* HC_SOLVE_FP32_IR: right-hand sides that finished in
                           FP64 (their FP32 iteration stopped making progress) */
  int pad;
} hc_solve_stats;

/* Create an operator on a structured hex mesh.  nu_e / sigma_e: per-element
 * curl-curl and mass coefficients (host arrays, nx*ny*nz each).  precision:
 * HC_PREC_* bits selecting which quadrature-data copies are kept. */
int hc_op_create(const hc_mesh_desc * mesh,
                 int order,
                 const double * nu_e,
                 const double * sigma_e,
                 int precision,
                 hc_op_t * out);
/* Replace the per-element coefficients (host arrays); takes effect at the next
 * hc_op_setup. */
int hc_op_set_coefficients(hc_op_t op, const double * nu_e, const double * sigma_e);
/* Partial-assembly setup for a(u,w) = (dt*nu curl u, curl w) + (sigma u, w):
 * FP64 quadrature-point factors (and the FP32 copy when requested) plus the
 * Jacobi diagonal.  Re-running with an unchanged dt (|ddt| <= 1e-12 dt) and
 * unchanged coefficients is a no-op. */
int hc_op_setup(hc_op_t op, double dt, void * stream);
int64_t hc_op_ndofs(hc_op_t op);
int hc_op_ess_count(hc_op_t op, int64_t * n_ess);
/* Device bytes owned by the handle. */
int64_t hc_op_device_bytes(hc_op_t op);

/* z = B r with the preconditioner of a solve mode: B = D^-1 (Jacobi, the
 * HypreDiagScale of DefaultJacobiPCGSolver, hephaestus_solvers.hpp:52), or,
 * with HC_SOLVE_AUX in mode, B = D^-1 + G Dg^-1 G^T (the gradient-space
 * correction; one-level form of HypreAMS, hcurl_formulation.cpp:69-79).
 * r, z: N doubles on the device.  Needs hc_op_setup. */
int hc_op_precond_apply(hc_op_t op, int mode, const double * r, double * z, void * stream);

/* min over elements of (trace of the sigma-mass factors) / (trace of the
 * curl-curl factors) at a quadrature point, divided by p^2: sigma h^2 / (nu p^2)
 * on an affine mesh, the relative size of the operator on gradient fields.
 * No counterpart in the reference (it selects the precision of FP32 + FP64
 * refinement).  Needs hc_op_setup. */
int hc_op_term_ratio(hc_op_t op, double * ratio);

/* Factor storage read by the apply kernels.  mode 1 (default): on an
 * axis-aligned box mesh (hc_mesh_desc.vert == NULL) the kernels read two
 * factors per element, dt nu_e and sigma_e, and form the quadrature-point
 * factors from them (J = diag(h): mass sigma_e w |det J| / h_a^2, curl
 * dt nu_e w h_a^2 / |det J|, no off-diagonal terms); other meshes read the
 * 12 (p+1)^3 quadrature-point factors per element of the partial assembly.
 * mode 0: the quadrature-point factors on every mesh.  The two agree to
 * round-off (a few units in the last place of each factor).  No counterpart
 * in the reference (MFEM's partial assembly stores the quadrature-point
 * factors).  hc_op_factor_mode: the storage the kernels read (0 or 1) or a
 * negative status. */
int hc_op_set_factor_mode(hc_op_t op, int mode);
int hc_op_factor_mode(hc_op_t op);

/* Process-wide switch (default off): batch and single solves replay their
 * steady-state rounds (64 iterations) from a captured hipGraph, one per table
 * set, instead of launching every kernel from the host.  No counterpart in the
 * reference.  Results are identical either way; on one MI355X the replay was
 * not faster (a round's kernels, not their launches, set its time). */
int hc_engine_set_graphs(int enable);

/* y = A x (PEC eliminated).  x, y: device, nrhs vectors (rhs r at r*N).
 * With nrhs > 1 a work item of the kernel applies the operator to a group of
 * right-hand sides, reading each tile's quadrature data once per group. */
int hc_op_apply(hc_op_t op, const double * x, double * y, int nrhs, void * stream);
int hc_op_apply_f32(hc_op_t op, const float * x, float * y, int nrhs, void * stream);
/* y = A x (FP64, one right-hand side) without any device allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops).  Same kernels and result bits
 * as hc_op_apply with nrhs = 1; the tile-face partials live in a static device
 * array of hc_noalloc_scratch_doubles() doubles per GPU (32 MiB), shared by
 * every call on that GPU: calls that overlap on one GPU race.  HC_EINVAL when
 * the operator's partials (4 slots of its tile-plane DOFs) exceed it.  The
 * scratch address is resolved at hc_op_create (or at the first call on a
 * device), so a call inside a capture only launches kernels. */
int hc_op_apply_noalloc(hc_op_t op, const double * x, double * y, void * stream);
int64_t hc_noalloc_scratch_doubles(void);
/* y = K x with K the curl-curl part alone, (nu curl u, curl w), same PEC
 * convention as hc_op_apply (y_ess = x_ess).  The operator of
 * WeakCurlCurlKernel (src/kernels/weak_curl_curl_kernel.cpp:22-31: the
 * right-hand side term -K u_n of a backward-Euler step), applied from the
 * quadrature data of hc_op_setup (no second assembly). */
int hc_op_apply_curl(hc_op_t op, const double * x, double * y, void * stream);
/* H1 nodes of the operator's mesh and order: (n0 p + 1)(n1 p + 1)(n2 p + 1),
 * x fastest (the scalar-potential space of the AV block operator). */
int64_t hc_op_nnodes(hc_op_t op);
/* Block operator of the AV formulation (AVEquationSystem::AddKernels,
 * src/formulations/AV/av_formulation.cpp:150-196) on the operator's mesh,
 * coefficients and dt.  Unknowns: a = dA/dt on the ND DOFs, v = V on the
 * H1 nodes (hc_op_nnodes).  With G the discrete gradient (H1 -> ND) and a0 = a
 * zeroed on the PEC DOFs:
 *   ya = (dt K_nu + M_sigma) a0 + M_sigma G v   CurlCurl + VectorFEMass +
 *        MixedVectorGradient (sigma grad V, u') (mixed_vector_gradient_kernel.hpp:8)
 *   yv = G^T M_sigma a0 + G^T M_sigma G v       VectorFEWeakDivergence (sigma u,
 *        grad V') (vector_fe_weak_divergence_kernel.hpp:8) + Diffusion
 *        (sigma grad V, grad V') (diffusion_kernel.hpp:8)
 * and ya = a on the PEC DOFs (hc_op_apply's elimination); v has no essential
 * nodes.  Outputs must not alias the inputs.  Device pointers, FP64. */
int hc_op_av_apply(hc_op_t op, const double * a, const double * v, double * ya, double * yv, void * stream);
/* Solve of the AV block system (the reference: HypreGMRES + BoomerAMG on the
 * monolithic matrix, src/problem_builders/problem_builder_base.cpp:206-218,
 * for AVFormulation): the block operator of hc_op_av_apply is symmetric
 * positive semi-definite (zero on the gauge fields (-G phi, phi)), so this is
 * PCG with a block-diagonal preconditioner: on box meshes with the multigrid
 * hierarchy of HC_SOLVE_MG (dense coarsest level), the V-cycle of that
 * hierarchy on the ND block and a node-space V-cycle on G^T M G (stats->
 * outer_iterations = 2); elsewhere block Jacobi (diag(A)^-1 on the ND DOFs, 1
 * on PEC DOFs; diag(G^T M G)^-1 on the nodes; outer_iterations = 1).  With
 * exact blocks the preconditioned operator has the eigenvalues 0 (gauge), 1
 * and 2 on the free DOFs (K G = 0 makes A G = M G).  (a, v) hold the initial
 * guess and
 * receive the solution (its gauge component is not determined).  Stops
 * when ||b - AV x|| <= rtol ||b|| (FP64, both blocks); stats->rel_res is the
 * true residual of the result.  Its work vectors are allocated on the first
 * call and kept with the operator; the PCG scalars stay on the device (the
 * host reads the solver state every few iterations, at most 8). */
int hc_op_av_solve(hc_op_t op, const double * ba, const double * bv, double * a, double * v, double rtol,
                   int maxit, hc_solve_stats * stats, void * stream);
/* DivFreeSource::Apply (src/sources/div_free_source.cpp:71-114) for a source
 * linear form j = (J, phi_i) (e.g. from hc_op_rhs_coil): g = M^-1 j with the
 * unit-coefficient H(curl) mass matrix (no essential DOFs); when project != 0
 * the Helmholtz projection g <- g - G Q, (G^T M G) Q = G^T M g (Q up to a
 * constant, no boundary conditions, as the reference's HelmholtzProjector with
 * an empty BCMap); then lf += M g.  src receives g, Q (optional, H1 nodes) the
 * potential.  Both solves: PCG to rtol (the reference: its default GMRES),
 * Jacobi-preconditioned, the projection's node Laplacian on box meshes by
 * the node-space multigrid V-cycle of a unit-mass operator on the same mesh.  stats: iterations of both solves summed; lf is left unchanged when
 * a solve fails.  Its work vectors and the unit-mass quadrature factors
 * (which depend on the mesh alone) are set up on the first call and kept with
 * the operator; PCG scalars on the device as in hc_op_av_solve. */
int hc_op_div_free_source(hc_op_t op, const double * j, double * src, double * lf, double * Q, int project,
                          double rtol, int maxit, hc_solve_stats * stats, void * stream);
/* Right-hand sides per work item of the multi-RHS apply and solve (0 =
 * automatic: one per work item; a larger group holds the tile's quadrature
 * data in registers across its right-hand sides).  Results do not depend on
 * it. */
int hc_op_set_rhs_block(hc_op_t op, int rhs_block);
/* Element-tile shape of the apply kernel.  Each order has a set of compiled
 * tile variants: tile extent tx x ty x tz elements and a load mode (0:
 * quadrature data loaded where it is used, 1: at kernel entry), plus 4 when each
 * pencil's quadrature points are split over two wave groups.  Results differ
 * between shapes by round-off only and are bitwise reproducible for a fixed
 * shape.  prec_bits: 64 or 32; mode -1 in hc_op_set_tile matches any mode. */
int hc_op_set_tile(hc_op_t op, int prec_bits, int tx, int ty, int tz, int mode);
/* Compiled tile variants of an order: writes (tx, ty, tz, mode) for up to
 * `max` variants into out and returns the number of variants. */
int hc_tile_list(int order, int * out, int max);
int hc_op_get_tile(hc_op_t op, int prec_bits, int * tx, int * ty, int * tz, int * mode);
/* Autotuner: times every compiled tile variant on this operator (after
 * hc_op_setup), `reps` applies each, and keeps the fastest per precision.
 * best_us (optional, 2 entries): the winning FP64 / FP32 apply times.
 * No reference counterpart (the reference applies a CSR matrix). */
int hc_op_autotune(hc_op_t op, int reps, double * best_us, void * stream);
/* Jacobi diagonal (1 on PEC DOFs) into device array d. */
int hc_op_diagonal(hc_op_t op, double * d, void * stream);

/* Jacobi-PCG on A x = b (b, x device; x holds the initial guess).  Stopping
 * rule ||b - A x||_2 <= rtol ||b||_2 per RHS, checked on the recursive residual
 * and confirmed with the FP64 operator.  maxit bounds the inner iterations of
 * each PCG; mode HC_SOLVE_FP64 or HC_SOLVE_FP32_IR. */
int hc_pcg_solve(hc_op_t op,
                 const double * b,
                 double * x,
                 int nrhs,
                 double rtol,
                 int maxit,
                 int mode,
                 hc_solve_stats * stats,
                 void * stream);

/* The multigrid hierarchy of HC_SOLVE_MG (built here if needed), for callers
 * that inspect or test it; no reference counterpart (HypreAMS keeps its
 * hierarchy internal, src/formulations/HCurl/hcurl_formulation.cpp:69-79).
 * hc_mg_levels: number of levels (level 0 = op itself) and the DOFs of the
 * coarsest.  hc_mg_level_op: the operator of a level (borrowed: valid until
 * op is set up again or destroyed; level 0 returns op).  hc_mg_transfer:
 * prolong (dir 1: y on level l = P x, x on level l+1) or restrict (dir 0:
 * y on level l+1 = P^T x with PEC entries zeroed, x on level l). */
int hc_mg_levels(hc_op_t op, int * nlevels, int64_t * coarse_ndofs, void * stream);
/* Process-wide: replay the finest-level V-cycle of HC_SOLVE_MG solves from a
 * captured hipGraph (default off; hierarchies with an exact coarsest solve,
 * solves called by the caller's thread -- the multigrid server and batch
 * workers always launch eagerly).  The replay launches the same kernels with
 * the same arguments in the same order, so results are bitwise those of eager
 * launches.  A capture must not overlap synchronous HIP calls (allocations,
 * blocking copies) of other host threads: enable it only where solves do not
 * run beside such calls.  No reference counterpart. */
int hc_mg_set_graphs(int enable);
int hc_mg_level_op(hc_op_t op, int level, hc_op_t * level_op);
int hc_mg_transfer(hc_op_t op, int level, int dir, const double * x, double * y, void * stream);

/* Batches of solves over several operator instances (differing sizes and
 * orders), advanced together: every iteration applies all instances of one
 * order in one launch and updates all right-hand sides in one launch per
 * vector kernel; converged right-hand sides drop out through device flags.
 * Replaces a loop of ImplicitSolve calls
 * (src/problem_operators/time_domain_equation_system_problem_operator.cpp:30-51)
 * over independent problems; the reference has no batched counterpart.
 * The tile variant of each order is that of the batch's first instance of that
 * order (hc_op_set_tile / hc_op_autotune).  Each instance's results equal
 * those of the same solve run alone with the same tile variant.
 *
 * hc_batch_add: b, x are device arrays of nrhs*N doubles of op (x = initial
 * guess); returns the entry index (>= 0) or a negative status.
 * hc_batch_solve: stats (optional) receives one entry per added instance.
 * Mode HC_SOLVE_FP64 | HC_SOLVE_MG: every entry is solved by the multigrid
 * PCG of hc_pcg_solve (its right-hand sides through one multi-RHS V-cycle per
 * iteration; the gradient-space PCG where a mesh has no hierarchy) by 4 host
 * workers with their own streams, each taking the next entry whose operator
 * no other worker is using (the caller's stream is synchronised first; the
 * call returns when all are done; solve_ms is the batch's wall time). */
int hc_batch_create(hc_batch_t * out);
int hc_batch_add(hc_batch_t batch, hc_op_t op, const double * b, double * x, int nrhs, double rtol, int maxit);
/* Right-hand sides per work item of the batched apply (0 = automatic). */
int hc_batch_set_rhs_block(hc_batch_t batch, int rhs_block);
int hc_batch_solve(hc_batch_t batch, int mode, hc_solve_stats * stats, void * stream);
int hc_batch_destroy(hc_batch_t batch);

/* Continuous-batching server of backward-Euler time-step solves
 * (TimeDomainEquationSystemProblemOperator::ImplicitSolve + the
 * BackwardEulerSolver update, time_domain_equation_system_problem_operator.cpp:30-51):
 * for a request (op, b_src, u) with A = dt K_nu + M_sigma set up by
 * hc_op_setup(op, dt):  f = b_src - K_nu u,  A k = f,  u <- u + dt k.
 * Operators stay resident (owned by the caller, used by the server alone while
 * its requests are in flight).  A scheduler thread admits requests in arrival
 * order as soon as their work vectors fit in the device-memory budget and
 * fewer than max_active are running, advances all admitted requests together
 * (one launch per ND order per operator apply), and retires each as soon as it
 * converges; a request's result does not depend on what else is running.
 * With mode HC_SOLVE_FP64 | HC_SOLVE_MG each request is solved by the
 * multigrid PCG of hc_pcg_solve (the gradient-space PCG for an operator whose
 * mesh has no hierarchy): max_active worker threads (0 = 4), each with its own
 * stream, take the next arrived request whose operator no other worker is
 * using, so requests of different operators run concurrently; round_iters and
 * hc_server_set_tile do not apply (operators keep their own tile variants).
 * The reference serves one problem per process and has no counterpart. */
typedef struct hc_server_s * hc_server_t;
typedef struct
{
  int64_t hbm_budget_bytes; /* device memory for the work vectors of admitted requests */
  int max_active;           /* most requests running at once (0 = no limit) */
  int mode;                 /* HC_SOLVE_FP64 or HC_SOLVE_FP32_IR (| HC_SOLVE_AUX), or
                               HC_SOLVE_FP64 | HC_SOLVE_MG */
  double rtol;              /* true relative residual of each solve */
  int maxit;                /* inner PCG iterations per request */
  int round_iters;          /* PCG iterations between scheduling decisions (0 = 16) */
} hc_server_config;
typedef struct
{
  uint64_t tag;
  int status;           /* HC_OK, HC_ENOTCONV or HC_ENAN */
  int converged;
  int iterations;       /* inner PCG iterations */
  int outer_iterations; /* refinement steps (FP32+IR) or PCG runs (FP64) */
  double rel_res;       /* final FP64 true relative residual of A k = f */
  double latency_ms;    /* arrival -> result */
  double queue_ms;      /* arrival -> admission */
  double arrival_ms;    /* arrival, from the server's creation */
  int fp64_fallback;    /* 1 if an FP32+IR solve finished in FP64 */
  int pad;
} hc_request_result;
typedef struct
{
  int64_t submitted, completed, rounds, pending, active;
  int64_t peak_active;
  int64_t budget_bytes, peak_bytes;
  double elapsed_ms;
} hc_server_stats;
/* Work-vector bytes one request of an operator with N DOFs takes from the
 * budget (for the server's mode; with HC_SOLVE_MG the solver's vectors belong
 * to the operator and only f, k count). */
int64_t hc_server_request_bytes(int64_t ndofs, int mode);
int hc_server_create(const hc_server_config * cfg, hc_server_t * out);
/* Tile variant used for all operators of an order (before any submit). */
int hc_server_set_tile(hc_server_t srv, int order, int prec_bits, int tx, int ty, int tz, int mode);
/* b_src, u: device, N doubles, valid until the request's result is polled.
 * arrival_ms >= 0: the request arrives that many ms after hc_server_create
 * (scheduled arrivals); < 0: now.  HC_EBUSY if it can never fit the budget. */
int hc_server_submit(hc_server_t srv, hc_op_t op, const double * b_src, double * u, double dt, uint64_t tag,
                     double arrival_ms);
/* Completed requests (up to max), waiting up to wait_ms for the first;
 * returns the number written or a negative status. */
int hc_server_poll(hc_server_t srv, hc_request_result * out, int max, int wait_ms);
int hc_server_info(hc_server_t srv, hc_server_stats * stats);
/* Finishes every submitted request, then releases the server. */
int hc_server_destroy(hc_server_t srv);

int hc_op_destroy(hc_op_t op);
const char * hc_last_error(void);

/* Fill device vector v (n entries) with b_i = 2*u01(splitmix64_i(seed)) - 1,
 * zeroed on the operator's PEC DOFs when op != NULL. */
int hc_fill_uniform(hc_op_t op, uint64_t seed, double * v, int64_t n, void * stream);

/* Timing of `reps` single-RHS applies y = A x (prec_bits 64 or 32, x/y of
 * that precision) with HIP events around each kernel on `stream`: out_us[0] =
 * mean duration of the tile kernel, out_us[1] = of the tile-face assembly
 * kernel.  No reference counterpart (measurement entry point). */
int hc_op_profile_apply(hc_op_t op, int prec_bits, const void * x, void * y, int reps, double * out_us,
                        void * stream);
/* Device vector helpers of the host layer.
 * hc_axpby: y = a x + b y (n doubles; b == 0 ignores y's old contents).
 * hc_op_zero_ess: v_i = 0 on the operator's PEC DOFs -- the right-hand side
 * part of FormLinearSystem with zero tangential boundary values
 * (src/equation_systems/equation_system.cpp:132-190, B_ess = x_bc = 0). */
int hc_axpby(int64_t n, double a, const double * x, double b, double * y, void * stream);
int hc_op_zero_ess(hc_op_t op, double * v, void * stream);
/* Coil-like right-hand side b_i = int J . phi_i of a divergence-free current
 * J = curl psi, psi = sum_{k<4} a_k exp(-|x - c_k|^2 / (2 s_k^2)) with
 * a_k in U[-1,1]^3, c_k in U[0.3,0.7]^3, s_k in U[0.05,0.15] drawn in that
 * order from splitmix64(seed); Gauss rule with p+2 points per direction; zero
 * on PEC DOFs.  b: device, N doubles.  Plays the role of the coil source
 * terms of the reference's examples (DivFreeSource, src/sources/div_free_source.cpp). */
int hc_op_rhs_coil(hc_op_t op, uint64_t seed, double * b, void * stream);

#ifdef __cplusplus
}
#endif
#endif
