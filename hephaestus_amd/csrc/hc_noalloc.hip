// hc_op_apply_noalloc: y = A x (FP64, one right-hand side) with no device
// allocation, for callers that cannot allocate (code under HIP graph capture,
// tight serving loops).  It runs the same kernels as hc_op_apply (the
// operator's FP64 tile variant, then the tile-face assembly) and differs only
// in where the tile-face partials live: a file-scope device array of fixed
// size instead of the handle's workspace, which hc_op_apply grows on demand.
//
// Every call on a GPU shares that GPU's array (the last call's partials stay
// in it; every partial slot that the assembly reads is written by the tile
// kernel of the same call, so nothing needs clearing).  Calls that overlap on
// one GPU (different streams or host threads) race on it.
#include "hc_op.hpp"

#include <mutex>

namespace
{

// Tile-face partials of one apply: 4 slots of the operator's compact
// tile-plane entries (2^22 doubles = 32 MiB per GPU, held until the process ends).
constexpr int64_t SCRATCH_DOUBLES = int64_t(1) << 22;
__device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
double * g_scratch[MAX_DEVICES] = {};
std::mutex g_mu;

// Device address of this GPU's scratch array (the host-side name `scratch` is
// only a stand-in for it).  Resolved once per device, with that device current.
int
scratch_of_current_device(double ** out)
{
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess)
    return hc::fail(HC_EHIP, std::string("hc_op_apply_noalloc: hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= MAX_DEVICES)
    return hc::fail(HC_EINVAL, "hc_op_apply_noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev])
  {
    void * p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return hc::fail(HC_EHIP, std::string("hc_op_apply_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
    g_scratch[dev] = (double *)p;
  }
  *out = g_scratch[dev];
  return HC_OK;
}

// Compact tile-plane entries of a decomposition (the partial slot length).
int64_t
plane_entries(const hc::Geom & g, const hc::PlaneSegs & s)
{
  int64_t S = 0;
  for (int k = 0; k < s.nseg; k++)
  {
    const int c = s.comp[k], a = s.axis[k];
    const int b0 = (a == 0) ? 1 : 0;
    const int64_t end = s.eoff[c][a] + (s.start[k + 1] - s.start[k]) * g.cd[c][b0];
    S = end > S ? end : S;
  }
  return S;
}

} // namespace

namespace hc
{

// Set-up of the current device's scratch address (hc_op_create calls it, so a
// later hc_op_apply_noalloc on that device makes no runtime call but launches).
int
noalloc_setup()
{
  double * p = nullptr;
  return scratch_of_current_device(&p);
}

} // namespace hc

extern "C" {

int64_t
hc_noalloc_scratch_doubles(void)
{
  return SCRATCH_DOUBLES;
}

int
hc_op_apply_noalloc(hc_op_t op, const double * x, double * y, void * stream)
try
{
  if (!op || !x || !y)
    return hc::fail(HC_EINVAL, "hc_op_apply_noalloc: null argument");
  if (!op->setup_done)
    return hc::fail(HC_EINVAL, "hc_op_apply_noalloc: hc_op_setup has not been called");
  const hc::ApplyPlan & pl = op->plan64;
  const int64_t S = plane_entries(pl.g, pl.segs);
  if (4 * S > SCRATCH_DOUBLES)
    return hc::fail(HC_EINVAL, "hc_op_apply_noalloc: the tile-face partials of this operator need " +
                                   std::to_string(4 * S) + " doubles, more than the " +
                                   std::to_string(SCRATCH_DOUBLES) +
                                   " of the static scratch (use hc_op_apply, or a tile variant with fewer tile planes)");
  double * part = nullptr;
  if (int rc = scratch_of_current_device(&part))
    return rc;
  hipStream_t st = (hipStream_t)stream;
  hc::ApplyDesc d{};
  d.g = pl.g;
  d.segs = pl.segs;
  d.prows = (int)pl.segs.start[pl.segs.nseg];
  d.qd = op->qd64;
  d.ef = op->ef(false);
  d.x = x;
  d.y = y;
  d.part = part;
  d.pstride = S > 0 ? S : 1;
  d.nrhs = 1;
  d.rb = 1;
  d.w0 = 0;
  d.pw0 = 0;
  d.ntiles = pl.g.nt[0] * pl.g.nt[1] * pl.g.nt[2];
  HC_TRY_HIP(pl.cand->launch(d, nullptr, 1, d.ntiles, 0, &op->tb64, 64, st));
  HC_TRY_HIP(hc::assemble_planes<double>(d, nullptr, 1, d.prows, st));
  return HC_OK;
}
catch (...)
{
  return hc::fail_exception();
}

} // extern "C"
