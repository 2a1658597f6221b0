// Launchers of the operator and solver kernels (host side of hc_kernels.hip and
// hc_pcg.hip).
#pragma once

#include "hc_common.hpp"
#include "hc_aux.hpp"

namespace hc
{

struct TileShape
{
  int tx, ty, tz, nthr;
};

// Segments (component c, plane axis a) of the tile-face assembly: the DOFs of
// component c on the interior tile planes normal to axis a, stored compactly
// as rows (plane k, position along the second in-plane axis) of entries along
// the first in-plane axis.  start[] is the prefix of rows, eoff[c][a] the first
// entry of segment (c, a) in the compact partial arrays.  A DOF on planes of
// two axes belongs to the segment of the lower-numbered axis.
struct PlaneSegs
{
  int nseg;
  int comp[6];
  int axis[6];
  int64_t start[7];
  int64_t eoff[3][3];
};

// One operator instance of an apply launch.  Work items are (tile, rhs group)
// with the group index fastest; a group is rb consecutive right-hand sides (the
// single-instance kernel uses rb = 1).  The items of instance i occupy
// [w0, w0 + ntiles*ceil(nrhs/rb)) of the launch's flattened work range.
// Vectors of rhs r start at r*N.
// Element factors of an axis-aligned box mesh.  With J = diag(h) the
// quadrature-point factors reduce to mass_a = sigma_e w_q |det J| / h_a^2 and
// curl_a = dt nu_e w_q h_a^2 / |det J| (no off-diagonal terms), so a box
// operator needs two numbers per element, qc[2e] = dt nu_e, qc[2e+1] =
// sigma_e (element type of the launch), plus the axis constants
// ck[a] = h_a^2 / |det J| and mk[a] = |det J| / h_a^2 and the 1D weights.
// qc == nullptr: the quadrature-point factors qd are used.
struct ElemFactors
{
  const void * qc = nullptr;
  double ck[3] = {0, 0, 0};
  double mk[3] = {0, 0, 0};
};

struct ApplyDesc
{
  Geom g;
  PlaneSegs segs;
  const void * qd; // quadrature data (element type of the launch)
  ElemFactors ef;  // element factors (box meshes), used instead of qd when set
  const void * resid = nullptr; // residual mode: y = resid - A x (vector at the rhs offset of y)
  const void * x;
  void * y;
  void * part;     // 4 slots per rhs: slot s of compact plane entry j at s*pstride + j
  const int * done;// per-rhs flags at done[r * done_stride], or null: rhs r
  int done_stride; // is applied only while its flag equals `gate`
  int gate = 0;
  int nrhs;
  int rb;          // right-hand sides per work item of the batched kernels
  int w0;          // first tile work item
  int ntiles;
  int pw0;         // first plane-assembly work item
  int prows;       // plane rows per rhs
  double mscale = 1.0; // weight of the mass term (0: curl-curl part only)
  double cscale = 1.0; // weight of the curl-curl term
  int64_t pstride = 0; // distance between the partial slots (0: N); a rhs takes 4 slots
};

// Slot stride of the tile-face partials of an instance.
__host__ __device__ inline int64_t
part_stride(const ApplyDesc & d)
{
  return d.pstride ? d.pstride : d.g.off[3];
}

// Locate the instance of a flattened work item: largest i with key(i) <= w.
template <typename F>
__device__ inline int
find_instance(int w, int ndesc, F key)
{
  int lo = 0, hi = ndesc - 1;
  while (lo < hi)
  {
    const int mid = (lo + hi + 1) >> 1;
    if (key(mid) <= w)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Workgroups are dispatched to the GPU's 8 XCDs round-robin (blockIdx mod 8),
// each XCD with its own L2.  This renumbering gives every XCD a contiguous
// range of work items, so neighbouring tiles (which read the same x halo rows
// and write parts of the same cache lines of y) run on one XCD at about the
// same time and meet in its L2.  A bijection of [0, nb).
// Pointers typed as global memory.  A pointer read from a descriptor table in
// device memory is otherwise generic, and generic (flat) accesses also count
// on the LDS counter: every LDS wait of a kernel would then wait for its
// outstanding loads and stores as well.
template <typename T>
using gptr = __attribute__((address_space(1))) T *;
template <typename T>
__device__ inline gptr<T>
gmem(T * p)
{
  return (gptr<T>)p;
}

// A wave-uniform 64-bit value held in scalar registers.
__device__ inline int64_t
uniform64(int64_t v)
{
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Offsets of the compact-partials segments (component c, plane axis a) of an
// instance, read once before a write-out loop into scalar registers.  Named
// fields, no array: an array selected by a per-lane axis would be placed in
// scratch memory, and inside the loop any per-lane load makes each face DOF
// wait for every store the thread issued before it (loads and stores share
// one memory counter).
struct SegOffsets
{
  int64_t x0, x1, x2, y0, y1, y2, z0, z1, z2;
};
__device__ inline SegOffsets
seg_offsets(const PlaneSegs & segs)
{
  SegOffsets o;
  o.x0 = uniform64(segs.eoff[0][0]);
  o.x1 = uniform64(segs.eoff[0][1]);
  o.x2 = uniform64(segs.eoff[0][2]);
  o.y0 = uniform64(segs.eoff[1][0]);
  o.y1 = uniform64(segs.eoff[1][1]);
  o.y2 = uniform64(segs.eoff[1][2]);
  o.z0 = uniform64(segs.eoff[2][0]);
  o.z1 = uniform64(segs.eoff[2][1]);
  o.z2 = uniform64(segs.eoff[2][2]);
  return o;
}
template <int c>
__device__ inline int64_t
seg_offset(const SegOffsets & o, int a)
{
  const int64_t e0 = c == 0 ? o.x0 : (c == 1 ? o.y0 : o.z0);
  const int64_t e1 = c == 0 ? o.x1 : (c == 1 ? o.y1 : o.z1);
  const int64_t e2 = c == 0 ? o.x2 : (c == 1 ? o.y2 : o.z2);
  return (a == 0) ? e0 : ((a == 1) ? e1 : e2);
}
__device__ inline int64_t
seg_offset(const SegOffsets & o, int c, int a)
{
  return (c == 0) ? seg_offset<0>(o, a) : ((c == 1) ? seg_offset<1>(o, a) : seg_offset<2>(o, a));
}

constexpr int NXCD = 8;
__device__ inline int
xcd_remap(int b, int nb)
{
  const int q = nb / NXCD, r = nb % NXCD;
  const int k = b % NXCD, i = b / NXCD;
  return k * q + (k < r ? k : r) + i;
}

// A compiled tile variant of the apply kernel (shape + load mode).
// descs == nullptr: the single instance d0; otherwise ndesc instances.
// grouped != 0 (with descs): work items are groups of rb right-hand sides.
// grouped bit 0: rhs groups; bit 1 (LAUNCH_EF): every instance of descs carries
// element factors (the single instance d0 is read from d0.ef.qc).
constexpr int LAUNCH_EF = 2;
using TileLaunch = hipError_t (*)(const ApplyDesc & d0, const ApplyDesc * descs, int ndesc, int nwork, int grouped,
                                  const void * tables, int prec_bits, hipStream_t st);
struct TileCandidate
{
  int p;
  TileShape shape;
  int mode; // 0: qdata loaded at the quadrature stage, 1: at entry; +2: per-component
            // transposed-x stage and write-out (lean LDS); +4: split qz range; 8: qp kernel
  TileLaunch launch;
};
// The one-lane-per-quadrature-point tile kernel (hc_apply_qp.hip), mode 8.
template <int P, int TX, int TY, int TZ>
hipError_t launch_qp(const ApplyDesc & d0, const ApplyDesc * descs, int ndesc, int nwork, int grouped,
                     const void * tables, int prec_bits, hipStream_t st);
constexpr int MODE_QP = 8;
// Candidates of order p (contiguous array); returns their count.
int tile_candidates(int p, const TileCandidate ** out);
TileShape default_tile(int p);
PlaneSegs make_plane_segs(const Geom & g);

hipError_t pa_setup(const Geom & g,
                    const Tables<double> & tb,
                    const double * vert,
                    const double h[3],
                    const double * curl_coef,
                    const double * mass_coef,
                    double curl_scale,
                    double * qd,
                    hipStream_t st);
hipError_t cast_f32(const double * a, float * b, int64_t n, hipStream_t st);
// Element factors of a box mesh (ElemFactors::qc): dt nu_e, sigma_e per
// element, FP64 and (qc32 != nullptr) FP32.
hipError_t elem_factors(int64_t E, const double * nu, const double * sig, double dt, double * qc, float * qc32,
                        hipStream_t st);

// Coil-like current J = curl psi, psi = sum_k a_k exp(-|x - c_k|^2 / (2 s_k^2)),
// with its Gauss rule (p + 2 points) and the ND basis at those points.
struct CoilParams
{
  double a[4][3], c[4][3], s[4];
  double qp[QMAXC + 1], qw[QMAXC + 1];
  double Bo[(QMAXC + 1) * PMAX], Bc[(QMAXC + 1) * (PMAX + 1)];
};
// y = a x + b y (b == 0: y = a x)
hipError_t axpby(int64_t n, double a, const double * x, double b, double * y, hipStream_t st);
// y = d * x (entrywise)
hipError_t mul(int64_t n, const double * d, const double * x, double * y, hipStream_t st);
// v_i = 0 on the PEC DOFs of g
hipError_t zero_ess(const Geom & g, double * v, hipStream_t st, int nrhs = 1); // nrhs vectors at stride N
// b_i = int J . phi_i (zero on PEC DOFs)
hipError_t rhs_coil(const Geom & g, const CoilParams & cp, const double * vert, const double h[3],
                    const double x0[3], double * b, hipStream_t st);
hipError_t diagonal(const Geom & g, const Tables<double> & tb, const double * qd, double * d, hipStream_t st);

// y = A x (PEC eliminated) for nrhs right-hand sides (vector r at r*N).
// part: 4*N*nrhs scratch for tile-face partials.  done: optional device flags
// done[r*done_stride]; the kernels skip rhs r when its flag is set.  rb > 1:
// the batched kernel with rb right-hand sides per work item; its descriptor
// is staged in desc_dev (device, one ApplyDesc).
template <typename T>
hipError_t apply_op(const Geom & g,
                    const PlaneSegs & segs,
                    const Tables<T> & tb,
                    const T * qd,
                    const ElemFactors & ef,
                    const T * x,
                    T * y,
                    T * part,
                    const int * done,
                    hipStream_t st,
                    const TileCandidate * cand = nullptr,
                    int nrhs = 1,
                    int done_stride = 0,
                    int rb = 1,
                    ApplyDesc * desc_dev = nullptr,
                    const T * resid_b = nullptr);
// Right-hand sides per work item for a batch of nrhs on ntiles tiles.
int default_rhs_block(int ntiles, int nrhs);
// y = (mscale M + cscale K') x (K' = the curl part of the quadrature data,
// i.e. dt K) with the batched kernel; its descriptor is staged in desc_dev.
// part: 4 N scratch.
template <typename T>
hipError_t apply_scaled(const Geom & g, const PlaneSegs & segs, const Tables<T> & tb, const T * qd,
                        const ElemFactors & ef, const T * x, T * y, T * part, double mscale, double cscale,
                        const TileCandidate * cand, ApplyDesc * desc_dev, hipStream_t st);

// Fill the plane-assembly fields of a descriptor (segments, rows per rhs).
void plane_rows(ApplyDesc & d);

// Apply over a device table of instances (all of one order, one tile variant).
// nwork = sum of ntiles*nrhs, npwork = sum of prows*nrhs.
template <typename T>
hipError_t apply_multi(const TileCandidate * cand, const Tables<T> & tb, const ApplyDesc * descs_dev, int ndesc,
                       int nwork, int npwork, int grouped, hipStream_t st);

// Sum of the tile-face partials (fixed order) into y.
template <typename T>
hipError_t assemble_planes(const ApplyDesc & d0, const ApplyDesc * descs, int ndesc, int npwork, hipStream_t st);

// ---------------------------------------------------------------------------
// Jacobi-PCG over solve slots (hc_pcg.hip)
// ---------------------------------------------------------------------------
constexpr int RED_BLOCKS = 2048; // most blocks per slot of the reduction kernels
constexpr int RED_THREADS = 256;

// Blocks of a slot of length n (a function of n only: the summation order of a
// slot does not depend on the rest of the batch).
__host__ __device__ inline int
red_blocks(int64_t n)
{
  const int64_t b = (n + 4 * RED_THREADS - 1) / (4 * RED_THREADS);
  return b < 1 ? 1 : (b > RED_BLOCKS ? RED_BLOCKS : (int)b);
}

// Device-resident solver scalars of one slot (inner PCG or outer refinement).
struct PcgState
{
  double gam[2]; // (r, z) of the previous / current iteration (by parity)
  double bb;     // (b, b)
  double rr;     // (r, r) of the last iterate
  double tol2;   // tolerance^2 relative to bb
  int done;      // inner: 1 once converged, out of iterations, or NaN met
  int iters;     // inner: iterations of this PCG; outer: refinement steps
  int status;    // 0, or HC_ENAN
  int it0;       // inner: value of the iteration counter at the start
  int maxit;     // inner: iteration cap
  int inner;     // outer: inner iterations summed over the refinement steps
  int phase;     // outer: PH_* below
  int first;     // outer: 1 while the residual of the initial guess is due;
                 // inner (reliable mode): 1 when paused for a replacement
  double rrmax;  // inner: largest (r, r) since the last residual replacement;
                 // outer: (r, r) of the previous FP64 residual
  int counted;   // inner: iterations already added to the outer count
  int prec64;    // outer: 1 when the slot iterates in FP64, 0 in FP32
  int run;       // inner: 0 idle, 1 FP32 iterations running, 2 FP64 running
  int stall;     // outer: FP32 replacements in a row without progress
  int revert;    // outer: 1 while the last FP32 increment is being undone
  int outer64;   // outer: finished FP64 inner solves
};
constexpr int STATE_STRIDE_INT = sizeof(PcgState) / sizeof(int);

// Phases of a solve slot (outer state).  Each round of the engine moves every
// slot through: PH_FIRST/finished inner -> PH_RESID (FP64 residual of x, the
// refinement update first when an FP32 inner solve finished) -> PH_DONE or
// PH_INIT (inner PCG start) -> PH_INNER (iterations until the inner done flag).
enum SlotPhase
{
  PH_FIRST = 1,
  PH_RESID = 2,
  PH_INIT = 3,
  PH_INNER = 4,
  PH_DONE = 5
};

// One right-hand side of one operator instance.  The inner vectors (ib..iq,
// idinv) have the precision of the inner solve (FP64 solve: b, x, r64, p64,
// q64 and the FP64 diagonal inverse; refinement: FP32 work vectors).
// One right-hand side of one operator instance.  FP64 PCG iterates on
// (b, x, r64, p64, q64, dinv64); FP32 PCG (reliable updates) on (r32, x32,
// p32, q32, dinv32) with x and the residual replaced in FP64.
struct SolveSlot
{
  int64_t n;
  double rtol;
  int maxit;
  int pad;
  double delta2;   // FP32: replacement when (r, r) < delta2 * rrmax
  const double * b;
  double * x;
  double * r64;    // FP64 residual
  double * p64;
  double * q64;    // FP64 A p, or A x for the residual
  const double * dinv64;
  float * r32;
  float * x32;     // FP32 iterate since the last replacement
  float * p32;
  float * q32;
  const float * dinv32;
  PcgState * in;
  PcgState * out;
  double * part; // RED_BLOCKS * 4 partial sums
  // gradient-space correction of the preconditioner (aux != 0):
  // z = D^-1 r + G Dg^-1 G^T r (hc_aux.hpp); t = Dg^-1 G^T r at the nodes
  int aux;
  int g64; // the FP64 apply group serves this slot: an FP64 inner start waits for it
  const double * ginv64;
  const float * ginv32;
  double * t64;
  float * t32;
  AuxGeom ag;
};

struct SlotTable
{
  const SolveSlot * slots; // device
  int nslots;
  int nb_max;  // max red_blocks(n) over the slots
  int any_aux; // some slot uses the gradient-space correction
  int nblk;    // blocks of the flattened slot grid: sum of red_blocks(n)
  const int * blk0; // device, nslots + 1: first block of each slot
};

// Phase moves at the start of a round (one thread per slot).
hipError_t phase_prep(const SlotTable & t, hipStream_t st);
// x += (double) x32 for slots in PH_RESID after FP32 iterations (reliable mode).
hipError_t ir_update(const SlotTable & t, hipStream_t st);
// r64 = b - q64 for slots in PH_RESID; outer state update and the move to
// PH_DONE or PH_INIT (or back to PH_RESID after undoing an FP32 increment
// that increased the residual).  A slot stops after max_outer outer steps, or
// after max_outer64 finished FP64 inner solves.
hipError_t residual64(const SlotTable & t, int max_outer, int max_outer64, hipStream_t st);
// The iteration counters git_pair[0] (FP32 iterations) and git_pair[1] (FP64)
// give each slot its iteration index: counter - value at the slot's start.
// FP32 PCG (re)start for FP32 slots in PH_INIT: r32 = (float) r64, x32 = 0,
// p = D^-1 r32 unless paused for a replacement (direction kept).
hipError_t pcg_init_reliable(const SlotTable & t, const int * git_pair, hipStream_t st);
// FP64 PCG start for FP64 slots in PH_INIT: r = b - A x (q64 = A x), p =
// D^-1 r; moves them to PH_INNER.
hipError_t pcg_init64(const SlotTable & t, const int * git_pair, hipStream_t st);
// One PCG iteration after q = A p for the slots running in precision T;
// advances the counter of T.
template <typename T>
hipError_t pcg_iterate(const SlotTable & t, int * git_pair, hipStream_t st);
// Copies the outer states of the slots into dst[0..nslots).
hipError_t gather_states(const SlotTable & t, PcgState * dst, hipStream_t st);
hipError_t zero_int(int * v, int n, hipStream_t st);
hipError_t invert_diag(const double * d, double * dinv64, float * dinv32, int64_t n, hipStream_t st);
hipError_t fill_uniform(const Geom * g, uint64_t seed, double * v, int64_t n, hipStream_t st);

} // namespace hc
