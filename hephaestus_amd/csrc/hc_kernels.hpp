// Launchers of the operator and solver kernels (host side of hc_kernels.hip and
// hc_pcg.hip).
#pragma once

#include "hc_common.hpp"

namespace hc
{

struct TileShape
{
  int tx, ty, tz, nthr;
};

// Segments (component, plane axis) of the tile-face assembly; start[] is the
// prefix of plane rows.
struct PlaneSegs
{
  int nseg;
  int comp[6];
  int axis[6];
  int64_t start[7];
};

// One operator instance of an apply launch.  Work items are (tile, rhs) with
// the rhs index fastest; the items of instance i occupy [w0, w0 + ntiles*nrhs)
// of the launch's flattened work range.  Vectors of rhs r start at r*N.
struct ApplyDesc
{
  Geom g;
  PlaneSegs segs;
  const void * qd; // quadrature data (element type of the launch)
  const void * x;
  void * y;
  void * part;     // 4 N per rhs
  const int * done;// per-rhs flags at done[r * done_stride], or null
  int done_stride;
  int nrhs;
  int w0;          // first tile work item
  int ntiles;
  int pw0;         // first plane-assembly work item
  int prows;       // plane rows per rhs
};

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

// A compiled tile variant of the apply kernel (shape + load mode).
// descs == nullptr: the single instance d0; otherwise ndesc instances.
using TileLaunch = hipError_t (*)(const ApplyDesc & d0, const ApplyDesc * descs, int ndesc, int nwork,
                                  const void * tables, int prec_bits, hipStream_t st);
struct TileCandidate
{
  int p;
  TileShape shape;
  int mode; // 0: qdata loaded at the quadrature stage, 1: at entry, 2: persistent pipelined
  TileLaunch launch;
};
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
hipError_t diagonal(const Geom & g, const Tables<double> & tb, const double * qd, double * d, hipStream_t st);

// y = A x (PEC eliminated).  part: 4*N scratch for tile-face partials.  done:
// optional device flag; when *done != 0 the kernels return at once.
template <typename T>
hipError_t apply_op(const Geom & g,
                    const PlaneSegs & segs,
                    const Tables<T> & tb,
                    const T * qd,
                    const T * x,
                    T * y,
                    T * part,
                    const int * done,
                    hipStream_t st,
                    const TileCandidate * cand = nullptr,
                    int nrhs = 1,
                    int done_stride = 0);

// Fill the plane-assembly fields of a descriptor (segments, rows per rhs).
void plane_rows(ApplyDesc & d);

// Apply over a device table of instances (all of one order, one tile variant).
// nwork = sum of ntiles*nrhs, npwork = sum of prows*nrhs.
template <typename T>
hipError_t apply_multi(const TileCandidate * cand, const Tables<T> & tb, const ApplyDesc * descs_dev, int ndesc,
                       int nwork, int npwork, hipStream_t st);

// Sum of the tile-face partials (fixed order) into y.
template <typename T>
hipError_t assemble_planes(const ApplyDesc & d0, const ApplyDesc * descs, int ndesc, int npwork, hipStream_t st);

// ---------------------------------------------------------------------------
// Jacobi-PCG (hc_pcg.hip)
// ---------------------------------------------------------------------------
constexpr int RED_BLOCKS = 512; // fixed grid of the reduction kernels
constexpr int RED_THREADS = 256;

// Device-resident solver scalars of one right-hand side.
struct PcgState
{
  double gam[2]; // (r, z) of the previous / current iteration (by parity)
  double bb;     // (b, b)
  double rr;     // (r, r) of the last iterate
  double tol2;   // rtol^2
  int done;      // 1 once converged (or NaN met)
  int iters;     // iterations performed
  int status;    // 0, or HC_ENAN
  int pad;
};

struct PcgWork
{
  double * part;    // RED_BLOCKS * 4 partial sums
  PcgState * state; // device
};

template <typename T>
hipError_t pcg_init(const T * b, const T * q, const T * dinv, T * r, T * p, int64_t n, double rtol, PcgWork w,
                    hipStream_t st);
template <typename T>
hipError_t pcg_dot_pq(const T * p, const T * q, int64_t n, PcgWork w, hipStream_t st);
template <typename T>
hipError_t pcg_update_xr(int it, T * x, const T * p, T * r, const T * q, const T * dinv, int64_t n, PcgWork w,
                         hipStream_t st);
template <typename T>
hipError_t pcg_update_p(int it, T * p, const T * r, const T * dinv, int64_t n, PcgWork w, hipStream_t st);

// r = b - Ax (FP64), state->rr = (r,r), state->bb = (b,b) when set_bb.
hipError_t residual64(const double * b, const double * Ax, double * r, int64_t n, PcgWork w, bool set_bb,
                      hipStream_t st);
// r32 = (float)(r64 / sqrt(state->rr)).
hipError_t to_f32_scaled(const double * r64, float * r32, int64_t n, const PcgState * s, hipStream_t st);
// x += sqrt(state->rr) * (double)d32   (the scale of the matching to_f32_scaled)
hipError_t ir_update(double * x, const float * d32, int64_t n, const PcgState * s, hipStream_t st);
hipError_t invert_diag(const double * d, double * dinv64, float * dinv32, int64_t n, hipStream_t st);
hipError_t fill_uniform(const Geom * g, uint64_t seed, double * v, int64_t n, hipStream_t st);

} // namespace hc
