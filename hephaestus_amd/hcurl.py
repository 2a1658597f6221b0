"""ctypes binding of libhcurl.so (the C ABI in include/hc_hcurl.h).

PyTorch is only plumbing here: device buffers are torch tensors whose data
pointers are handed to the C ABI, and the HIP stream is torch's current stream.
The library is loaded after torch so that both share one HIP runtime.  There is
no fallback: if libhcurl.so is missing or no HIP device is present the calls
raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhcurl.so")

HC_OK = 0
HC_EINVAL = -1
HC_ENOMEM = -2
HC_EHIP = -3
HC_ENOTCONV = -4
HC_ENAN = -5
HC_EBUSY = -6

FACE_ALL = 0x3F
PREC_FP64 = 0x1
PREC_FP32 = 0x2
SOLVE_FP64 = 0
SOLVE_FP32_IR = 1
SOLVE_AUX = 0x10  # flag: Jacobi + discrete-gradient auxiliary-space correction
SOLVE_MG = 0x20  # flag: geometric multigrid V-cycle preconditioner (FP64)

# Entry points declared in include/hc_hcurl.h (checked by the CPU test-suite).
EXPORTS = [
    "hc_op_create", "hc_op_set_coefficients", "hc_op_setup", "hc_op_ndofs", "hc_op_ess_count",
    "hc_op_device_bytes", "hc_op_apply", "hc_op_apply_f32", "hc_op_diagonal", "hc_op_precond_apply",
    "hc_op_term_ratio", "hc_engine_set_graphs", "hc_pcg_solve",
    "hc_op_destroy", "hc_last_error", "hc_fill_uniform", "hc_op_set_tile", "hc_op_get_tile",
    "hc_op_autotune", "hc_tile_list", "hc_op_set_rhs_block", "hc_batch_create", "hc_batch_add", "hc_batch_set_rhs_block",
    "hc_batch_solve", "hc_batch_destroy", "hc_op_rhs_coil", "hc_op_profile_apply", "hc_op_apply_curl", "hc_axpby", "hc_op_zero_ess",
    "hc_op_nnodes", "hc_op_av_apply", "hc_op_av_solve", "hc_op_div_free_source",
    "hc_server_request_bytes", "hc_server_create", "hc_server_set_tile", "hc_server_submit", "hc_server_poll",
    "hc_server_info", "hc_server_destroy", "hc_op_apply_noalloc", "hc_noalloc_scratch_doubles",
    "hc_mg_levels", "hc_mg_level_op", "hc_mg_transfer", "hc_mg_set_graphs",
    "hc_op_set_factor_mode", "hc_op_factor_mode",
]


class HcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class MeshDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("vert", C.POINTER(C.c_double)),
                ("x0", C.c_double * 3), ("len", C.c_double * 3),
                ("pec_faces", C.c_int)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_int64), ("max_active", C.c_int), ("mode", C.c_int),
                ("rtol", C.c_double), ("maxit", C.c_int), ("round_iters", C.c_int)]


class RequestResult(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("status", C.c_int), ("converged", C.c_int), ("iterations", C.c_int),
                ("outer_iterations", C.c_int), ("rel_res", C.c_double), ("latency_ms", C.c_double),
                ("queue_ms", C.c_double), ("arrival_ms", C.c_double), ("fp64_fallback", C.c_int),
                ("pad", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("rounds", C.c_int64),
                ("pending", C.c_int64), ("active", C.c_int64), ("peak_active", C.c_int64),
                ("budget_bytes", C.c_int64), ("peak_bytes", C.c_int64), ("elapsed_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("outer_iterations", C.c_int), ("converged", C.c_int),
                ("status", C.c_int), ("rel_res", C.c_double), ("solve_ms", C.c_double),
                ("fp64_fallbacks", C.c_int), ("pad", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def load():
    """Load libhcurl.so (after torch, so that one HIP runtime is shared)."""
    global _lib
    if _lib is not None:
        return _lib
    import torch  # noqa: F401  (HIP runtime first)
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (hipcc) first")
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    lib.hc_op_create.argtypes = [C.POINTER(MeshDesc), C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                 C.c_int, C.POINTER(vp)]
    lib.hc_op_set_coefficients.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.hc_op_setup.argtypes = [vp, C.c_double, vp]
    lib.hc_op_ndofs.argtypes = [vp]
    lib.hc_op_ndofs.restype = C.c_int64
    lib.hc_op_ess_count.argtypes = [vp, C.POINTER(C.c_int64)]
    lib.hc_op_device_bytes.argtypes = [vp]
    lib.hc_op_device_bytes.restype = C.c_int64
    lib.hc_op_apply.argtypes = [vp, vp, vp, C.c_int, vp]
    lib.hc_op_apply_f32.argtypes = [vp, vp, vp, C.c_int, vp]
    lib.hc_op_apply_noalloc.argtypes = [vp, vp, vp, vp]
    lib.hc_noalloc_scratch_doubles.restype = C.c_int64
    lib.hc_op_diagonal.argtypes = [vp, vp, vp]
    lib.hc_op_precond_apply.argtypes = [vp, C.c_int, vp, vp, vp]
    lib.hc_op_term_ratio.argtypes = [vp, C.POINTER(C.c_double)]
    lib.hc_op_set_factor_mode.argtypes = [vp, C.c_int]
    lib.hc_op_factor_mode.argtypes = [vp]
    lib.hc_engine_set_graphs.argtypes = [C.c_int]
    lib.hc_pcg_solve.argtypes = [vp, vp, vp, C.c_int, C.c_double, C.c_int, C.c_int, C.POINTER(SolveStats), vp]
    lib.hc_mg_levels.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int64), vp]
    lib.hc_mg_level_op.argtypes = [vp, C.c_int, C.POINTER(vp)]
    lib.hc_mg_transfer.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp]
    lib.hc_mg_set_graphs.argtypes = [C.c_int]
    lib.hc_op_destroy.argtypes = [vp]
    lib.hc_last_error.restype = C.c_char_p
    lib.hc_fill_uniform.argtypes = [vp, C.c_uint64, vp, C.c_int64, vp]
    lib.hc_op_set_tile.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    lib.hc_op_get_tile.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                   C.POINTER(C.c_int)]
    lib.hc_op_autotune.argtypes = [vp, C.c_int, C.POINTER(C.c_double), vp]
    lib.hc_tile_list.argtypes = [C.c_int, C.POINTER(C.c_int), C.c_int]
    lib.hc_op_set_rhs_block.argtypes = [vp, C.c_int]
    lib.hc_batch_create.argtypes = [C.POINTER(vp)]
    lib.hc_batch_add.argtypes = [vp, vp, vp, vp, C.c_int, C.c_double, C.c_int]
    lib.hc_batch_set_rhs_block.argtypes = [vp, C.c_int]
    lib.hc_batch_solve.argtypes = [vp, C.c_int, C.POINTER(SolveStats), vp]
    lib.hc_batch_destroy.argtypes = [vp]
    lib.hc_op_rhs_coil.argtypes = [vp, C.c_uint64, vp, vp]
    lib.hc_op_apply_curl.argtypes = [vp, vp, vp, vp]
    lib.hc_op_nnodes.argtypes = [vp]
    lib.hc_op_nnodes.restype = C.c_int64
    lib.hc_op_av_apply.argtypes = [vp, vp, vp, vp, vp, vp]
    lib.hc_op_av_solve.argtypes = [vp, vp, vp, vp, vp, C.c_double, C.c_int, C.POINTER(SolveStats), vp]
    lib.hc_op_div_free_source.argtypes = [vp, vp, vp, vp, vp, C.c_int, C.c_double, C.c_int,
                                          C.POINTER(SolveStats), vp]
    lib.hc_axpby.argtypes = [C.c_int64, C.c_double, vp, C.c_double, vp, vp]
    lib.hc_op_zero_ess.argtypes = [vp, vp, vp]
    lib.hc_server_request_bytes.argtypes = [C.c_int64, C.c_int]
    lib.hc_server_request_bytes.restype = C.c_int64
    lib.hc_server_create.argtypes = [C.POINTER(ServerConfig), C.POINTER(vp)]
    lib.hc_server_set_tile.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    lib.hc_server_submit.argtypes = [vp, vp, vp, vp, C.c_double, C.c_uint64, C.c_double]
    lib.hc_server_poll.argtypes = [vp, C.POINTER(RequestResult), C.c_int, C.c_int]
    lib.hc_server_info.argtypes = [vp, C.POINTER(ServerStats)]
    lib.hc_server_destroy.argtypes = [vp]
    lib.hc_op_profile_apply.argtypes = [vp, C.c_int, vp, vp, C.c_int, C.POINTER(C.c_double), vp]
    _lib = lib
    return lib


def _check(rc, what=""):
    if rc != HC_OK:
        msg = _lib.hc_last_error().decode()
        raise HcError(rc, f"{what}: {msg}" if what else msg)


def _stream(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _nrhs(v, n, dtype):
    """Right-hand sides of a vector argument: shape (N,) or (nrhs, N), contiguous."""
    import torch
    if v.dtype != dtype or not v.is_cuda or not v.is_contiguous():
        raise ValueError(f"expected a contiguous {dtype} device tensor")
    if v.dim() == 1 and v.shape[0] == n:
        return 1
    if v.dim() == 2 and v.shape[1] == n:
        return int(v.shape[0])
    raise ValueError(f"expected shape ({n},) or (nrhs, {n}), got {tuple(v.shape)}")


def set_graphs(enable):
    """Process-wide: replay the solve engine's steady-state rounds from captured hipGraphs."""
    _check(load().hc_engine_set_graphs(1 if enable else 0), "hc_engine_set_graphs")


def set_mg_graphs(enable):
    """Process-wide: replay multigrid V-cycles from captured hipGraphs (default off)."""
    _check(load().hc_mg_set_graphs(1 if enable else 0), "hc_mg_set_graphs")


def tile_list(order):
    """Compiled tile variants (tx, ty, tz, prefetch) of the apply kernel for an order."""
    lib = load()
    n = lib.hc_tile_list(int(order), None, 0)
    buf = (C.c_int * (4 * n))()
    lib.hc_tile_list(int(order), buf, n)
    return [tuple(buf[4 * i:4 * i + 4]) for i in range(n)]


class Operator:
    """dt*nu curl-curl + sigma mass on a structured hex mesh (ND order p), PEC
    faces eliminated.  Vectors are torch tensors on the HIP device."""

    def __init__(self, n, order, nu, sigma, precision=PREC_FP64 | PREC_FP32, pec_faces=FACE_ALL,
                 vert=None, x0=(0.0, 0.0, 0.0), length=(1.0, 1.0, 1.0)):
        lib = load()
        if np.isscalar(n):
            n = (int(n),) * 3
        self.n = tuple(int(v) for v in n)
        self.order = int(order)
        E = self.n[0] * self.n[1] * self.n[2]
        self.nelem = E
        nu = np.broadcast_to(np.asarray(nu, dtype=np.float64), (E,)).copy()
        sigma = np.broadcast_to(np.asarray(sigma, dtype=np.float64), (E,)).copy()
        self._vert = None if vert is None else np.ascontiguousarray(vert, dtype=np.float64)
        desc = MeshDesc(self.n[0], self.n[1], self.n[2],
                        self._vert.ctypes.data_as(C.POINTER(C.c_double)) if self._vert is not None else None,
                        (C.c_double * 3)(*x0), (C.c_double * 3)(*length), int(pec_faces))
        h = C.c_void_p()
        _check(lib.hc_op_create(C.byref(desc), self.order,
                                nu.ctypes.data_as(C.POINTER(C.c_double)),
                                sigma.ctypes.data_as(C.POINTER(C.c_double)), int(precision), C.byref(h)),
               "hc_op_create")
        self._h = h
        self.ndofs = int(lib.hc_op_ndofs(h))

    def set_coefficients(self, nu, sigma):
        E = self.nelem
        nu = np.broadcast_to(np.asarray(nu, dtype=np.float64), (E,)).copy()
        sigma = np.broadcast_to(np.asarray(sigma, dtype=np.float64), (E,)).copy()
        _check(_lib.hc_op_set_coefficients(self._h, nu.ctypes.data_as(C.POINTER(C.c_double)),
                                           sigma.ctypes.data_as(C.POINTER(C.c_double))), "set_coefficients")

    def setup(self, dt=1.0, stream=None):
        _check(_lib.hc_op_setup(self._h, float(dt), _stream(stream)), "hc_op_setup")
        return self

    @property
    def ess_count(self):
        v = C.c_int64()
        _check(_lib.hc_op_ess_count(self._h, C.byref(v)))
        return int(v.value)

    @property
    def device_bytes(self):
        return int(_lib.hc_op_device_bytes(self._h))

    def apply(self, x, y=None, stream=None):
        """y = A x; x of shape (N,) or (nrhs, N), float64."""
        import torch
        nrhs = _nrhs(x, self.ndofs, torch.float64)
        if y is None:
            y = torch.empty_like(x)
        _check(_lib.hc_op_apply(self._h, _ptr(x), _ptr(y), nrhs, _stream(stream)), "hc_op_apply")
        return y

    def apply_curl(self, x, y=None, stream=None):
        """y = K x, the curl-curl part alone (nu curl u, curl w); x of shape (N,)."""
        import torch
        if _nrhs(x, self.ndofs, torch.float64) != 1:
            raise ValueError("apply_curl takes one vector")
        if y is None:
            y = torch.empty_like(x)
        _check(_lib.hc_op_apply_curl(self._h, _ptr(x), _ptr(y), _stream(stream)), "hc_op_apply_curl")
        return y

    @property
    def nnodes(self):
        """H1 nodes of the mesh and order (the AV scalar-potential space)."""
        return int(_lib.hc_op_nnodes(self._h))

    def av_apply(self, a, v, ya=None, yv=None, stream=None):
        """AV block operator: (ya, yv) = [[dt K + M, M G], [G^T M, G^T M G]] (a, v),
        PEC DOFs of a eliminated; a of shape (N,), v of shape (nnodes,)."""
        import torch
        if _nrhs(a, self.ndofs, torch.float64) != 1:
            raise ValueError("av_apply takes one vector a")
        nn = self.nnodes
        if v.dtype != torch.float64 or v.numel() != nn or not v.is_contiguous():
            raise ValueError(f"v must be a contiguous float64 tensor of {nn} entries")
        if ya is None:
            ya = torch.empty_like(a)
        if yv is None:
            yv = torch.empty_like(v)
        _check(_lib.hc_op_av_apply(self._h, _ptr(a), _ptr(v), _ptr(ya), _ptr(yv), _stream(stream)),
               "hc_op_av_apply")
        return ya, yv

    def av_solve(self, ba, bv, a, v, rtol=1e-8, maxit=10000, stream=None, raise_on_fail=True):
        """PCG on the AV block system; (a, v) hold the initial guess and receive
        the solution."""
        import torch
        for t, n in ((ba, self.ndofs), (a, self.ndofs), (bv, self.nnodes), (v, self.nnodes)):
            if t.dtype != torch.float64 or t.numel() != n or not t.is_contiguous():
                raise ValueError("av_solve: contiguous float64 vectors of ndofs / nnodes entries")
        st = SolveStats()
        rc = _lib.hc_op_av_solve(self._h, _ptr(ba), _ptr(bv), _ptr(a), _ptr(v), float(rtol), int(maxit),
                                 C.byref(st), _stream(stream))
        if rc != HC_OK and (raise_on_fail or rc not in (HC_ENOTCONV, HC_ENAN)):
            _check(rc, "hc_op_av_solve")
        d = st.as_dict()
        d["rc"] = rc
        return d

    def div_free_source(self, j, lf, project=True, rtol=1e-10, maxit=20000, stream=None):
        """DivFreeSource::Apply: g = M^-1 j (unit mass), optional Helmholtz
        projection g - grad Q, lf += M g.  Returns (g, Q, stats)."""
        import torch
        _nrhs(j, self.ndofs, torch.float64)
        _nrhs(lf, self.ndofs, torch.float64)
        g = torch.empty_like(j)
        Q = torch.empty(self.nnodes, dtype=torch.float64, device=j.device)
        st = SolveStats()
        _check(_lib.hc_op_div_free_source(self._h, _ptr(j), _ptr(g), _ptr(lf), _ptr(Q), int(bool(project)),
                                          float(rtol), int(maxit), C.byref(st), _stream(stream)),
               "hc_op_div_free_source")
        return g, Q, st.as_dict()

    def apply_noalloc(self, x, y=None, stream=None):
        """y = A x (FP64, one right-hand side) through hc_op_apply_noalloc (no
        device allocation; partials in the library's static scratch)."""
        import torch
        _nrhs(x, self.ndofs, torch.float64)
        if y is None:
            y = torch.empty_like(x)
        _check(_lib.hc_op_apply_noalloc(self._h, _ptr(x), _ptr(y), _stream(stream)), "hc_op_apply_noalloc")
        return y

    def apply_f32(self, x, y=None, stream=None):
        import torch
        nrhs = _nrhs(x, self.ndofs, torch.float32)
        if y is None:
            y = torch.empty_like(x)
        _check(_lib.hc_op_apply_f32(self._h, _ptr(x), _ptr(y), nrhs, _stream(stream)), "hc_op_apply_f32")
        return y

    def diagonal(self, stream=None):
        import torch
        d = torch.empty(self.ndofs, dtype=torch.float64, device="cuda")
        _check(_lib.hc_op_diagonal(self._h, _ptr(d), _stream(stream)), "hc_op_diagonal")
        return d

    def set_factor_mode(self, mode):
        """1 (default): element factors on box meshes; 0: quadrature-point factors."""
        _check(_lib.hc_op_set_factor_mode(self._h, int(mode)), "hc_op_set_factor_mode")
        return self

    @property
    def factor_mode(self):
        """The factor storage the apply kernels read (1: element factors, 0: quadrature points)."""
        r = _lib.hc_op_factor_mode(self._h)
        if r < 0:
            _check(r, "hc_op_factor_mode")
        return int(r)

    @property
    def term_ratio(self):
        """min over elements of sigma h^2 / (nu p^2) (from the quadrature factors)."""
        v = C.c_double()
        _check(_lib.hc_op_term_ratio(self._h, C.byref(v)), "hc_op_term_ratio")
        return v.value

    def precond_apply(self, r, mode=SOLVE_FP64 | SOLVE_AUX, z=None, stream=None):
        """z = B r with the preconditioner of a solve mode (Jacobi, or with
        SOLVE_AUX Jacobi + the gradient-space correction)."""
        import torch
        _nrhs(r, self.ndofs, torch.float64)
        if z is None:
            z = torch.empty_like(r)
        _check(_lib.hc_op_precond_apply(self._h, int(mode), _ptr(r), _ptr(z), _stream(stream)), "hc_op_precond_apply")
        return z

    def solve(self, b, x, rtol=1e-8, maxit=10000, mode=SOLVE_FP64, stream=None, raise_on_fail=True):
        """Jacobi-PCG (mode SOLVE_FP64) or FP32 PCG + FP64 refinement on A x = b;
        mode | SOLVE_AUX adds the gradient-space correction to the Jacobi
        preconditioner; SOLVE_FP64 | SOLVE_MG preconditions with the geometric
        multigrid V-cycle.  b, x of shape (N,) or (nrhs, N), x holds the initial guess."""
        import torch
        nrhs = _nrhs(b, self.ndofs, torch.float64)
        if _nrhs(x, self.ndofs, torch.float64) != nrhs:
            raise ValueError("b and x differ in their number of right-hand sides")
        st = SolveStats()
        rc = _lib.hc_pcg_solve(self._h, _ptr(b), _ptr(x), nrhs, float(rtol), int(maxit), int(mode),
                               C.byref(st), _stream(stream))
        if rc != HC_OK and (raise_on_fail or rc not in (HC_ENOTCONV, HC_ENAN)):
            _check(rc, "hc_pcg_solve")
        d = st.as_dict()
        d["rc"] = rc
        return d

    def set_rhs_block(self, rb):
        """Right-hand sides per work item of multi-RHS applies/solves (0 = automatic)."""
        _check(_lib.hc_op_set_rhs_block(self._h, int(rb)), "hc_op_set_rhs_block")

    def set_tile(self, prec_bits, tx, ty, tz, prefetch=-1):
        _check(_lib.hc_op_set_tile(self._h, prec_bits, tx, ty, tz, prefetch), "hc_op_set_tile")

    def get_tile(self, prec_bits=64):
        v = [C.c_int() for _ in range(4)]
        _check(_lib.hc_op_get_tile(self._h, prec_bits, *[C.byref(a) for a in v]), "hc_op_get_tile")
        return tuple(int(a.value) for a in v)

    def autotune(self, reps=5, stream=None):
        best = (C.c_double * 2)()
        _check(_lib.hc_op_autotune(self._h, int(reps), best, _stream(stream)), "hc_op_autotune")
        return {"fp64_us": best[0], "fp32_us": best[1], "tile64": self.get_tile(64), "tile32": self.get_tile(32)}

    def fill_uniform(self, seed, v, stream=None):
        _check(_lib.hc_fill_uniform(self._h, C.c_uint64(seed), _ptr(v), v.numel(), _stream(stream)),
               "hc_fill_uniform")
        return v

    def profile_apply(self, x, y, reps=20, stream=None):
        """Mean per-launch durations (us) of the tile kernel and of the tile-face
        assembly kernel over `reps` applies, timed with HIP events on the stream."""
        import torch
        bits = 64 if x.dtype == torch.float64 else 32
        _nrhs(x, self.ndofs, x.dtype)
        out = (C.c_double * 2)()
        _check(_lib.hc_op_profile_apply(self._h, bits, _ptr(x), _ptr(y), int(reps), out, _stream(stream)),
               "hc_op_profile_apply")
        return {"tile_us": out[0], "planes_us": out[1]}

    def rhs_coil(self, seed, b=None, stream=None):
        """Coil-like divergence-free current right-hand side (seeded)."""
        import torch
        if b is None:
            b = torch.empty(self.ndofs, dtype=torch.float64, device="cuda")
        _nrhs(b, self.ndofs, torch.float64)
        _check(_lib.hc_op_rhs_coil(self._h, C.c_uint64(seed), _ptr(b), _stream(stream)), "hc_op_rhs_coil")
        return b

    def mg_levels(self, stream=None):
        """(levels, DOFs of the coarsest) of the SOLVE_MG hierarchy (built if needed)."""
        nl, nc = C.c_int(), C.c_int64()
        _check(_lib.hc_mg_levels(self._h, C.byref(nl), C.byref(nc), _stream(stream)), "hc_mg_levels")
        return int(nl.value), int(nc.value)

    def mg_level(self, level):
        """The operator of a hierarchy level (borrowed: valid until this
        operator is set up again or closed)."""
        h = C.c_void_p()
        _check(_lib.hc_mg_level_op(self._h, int(level), C.byref(h)), "hc_mg_level_op")
        return _BorrowedOperator(h, self)

    def mg_transfer(self, level, prolong, x, y, stream=None):
        """y = P x (prolong: level+1 -> level) or y = P^T x, PEC entries zeroed
        (restrict: level -> level+1)."""
        _check(_lib.hc_mg_transfer(self._h, int(level), 1 if prolong else 0, _ptr(x), _ptr(y), _stream(stream)),
               "hc_mg_transfer")
        return y

    def close(self):
        if getattr(self, "_h", None) and self._h.value and not getattr(self, "_borrowed", False):
            _lib.hc_op_destroy(self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _BorrowedOperator(Operator):
    """An operator handle owned by another (a multigrid level): never destroyed here."""

    def __init__(self, h, owner):
        self._borrowed = True
        self._h = h
        self._owner = owner
        self.ndofs = int(_lib.hc_op_ndofs(h))


class Batch:
    """Solves over several operator instances advanced together (hc_batch_*).
    Keeps references to the operators and vectors it was given."""

    def __init__(self):
        lib = load()
        h = C.c_void_p()
        _check(lib.hc_batch_create(C.byref(h)), "hc_batch_create")
        self._h = h
        self._keep = []

    def add(self, op, b, x, rtol=1e-8, maxit=10000):
        import torch
        nrhs = _nrhs(b, op.ndofs, torch.float64)
        if _nrhs(x, op.ndofs, torch.float64) != nrhs:
            raise ValueError("b and x differ in their number of right-hand sides")
        rc = _lib.hc_batch_add(self._h, op._h, _ptr(b), _ptr(x), nrhs, float(rtol), int(maxit))
        if rc < 0:
            _check(rc, "hc_batch_add")
        self._keep.append((op, b, x))
        return rc

    def set_rhs_block(self, rb):
        _check(_lib.hc_batch_set_rhs_block(self._h, int(rb)), "hc_batch_set_rhs_block")

    def solve(self, mode=SOLVE_FP32_IR, stream=None, raise_on_fail=True):
        n = len(self._keep)
        st = (SolveStats * max(n, 1))()
        rc = _lib.hc_batch_solve(self._h, int(mode), st, _stream(stream))
        if rc != HC_OK and (raise_on_fail or rc not in (HC_ENOTCONV, HC_ENAN)):
            _check(rc, "hc_batch_solve")
        out = [st[i].as_dict() for i in range(n)]
        return rc, out

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.hc_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching server of backward-Euler step solves (hc_server_*).
    A request (op, b_src, u, dt) computes u <- u + dt k with
    (dt K + M) k = b_src - K u on the operator set up with that dt."""

    def __init__(self, budget_bytes, mode=SOLVE_FP32_IR, rtol=1e-8, maxit=20000, max_active=0, round_iters=16):
        lib = load()
        cfg = ServerConfig(int(budget_bytes), int(max_active), int(mode), float(rtol), int(maxit), int(round_iters))
        h = C.c_void_p()
        _check(lib.hc_server_create(C.byref(cfg), C.byref(h)), "hc_server_create")
        self._h = h
        self.mode = mode
        self._keep = {}

    @staticmethod
    def request_bytes(ndofs, mode=SOLVE_FP32_IR):
        return int(load().hc_server_request_bytes(int(ndofs), int(mode)))

    def set_tile(self, order, prec_bits, tx, ty, tz, mode=-1):
        _check(_lib.hc_server_set_tile(self._h, order, prec_bits, tx, ty, tz, mode), "hc_server_set_tile")

    def submit(self, op, b_src, u, dt, tag, arrival_ms=-1.0):
        import torch
        _nrhs(b_src, op.ndofs, torch.float64)
        _nrhs(u, op.ndofs, torch.float64)
        _check(_lib.hc_server_submit(self._h, op._h, _ptr(b_src), _ptr(u), float(dt), C.c_uint64(tag),
                                     float(arrival_ms)), "hc_server_submit")
        self._keep[int(tag)] = (op, b_src, u)

    def poll(self, max_results=256, wait_ms=100):
        buf = (RequestResult * max_results)()
        n = _lib.hc_server_poll(self._h, buf, int(max_results), int(wait_ms))
        if n < 0:
            _check(n, "hc_server_poll")
        out = [buf[i].as_dict() for i in range(n)]
        for r in out:
            self._keep.pop(int(r["tag"]), None)
        return out

    def info(self):
        st = ServerStats()
        _check(_lib.hc_server_info(self._h, C.byref(st)), "hc_server_info")
        return st.as_dict()

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.hc_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
