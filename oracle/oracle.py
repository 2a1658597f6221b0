"""ctypes front-end of the CPU FP64 oracle (oracle/hc_oracle.c).

TEST INFRASTRUCTURE ONLY.  This module is the parity checker and the CPU
baseline: only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg
of bench.py may import it.  The product package (hephaestus_amd) never does.

It restates, on the CPU, the operator that hephaestus assembles in
src/formulations/HCurl/hcurl_formulation.cpp:130-157 (dt*alpha curl-curl +
beta mass, from CurlCurlKernel / VectorFEMassKernel), the PEC elimination of
src/equation_systems/equation_system.cpp:132-190 and the Jacobi-PCG of
src/solvers/hephaestus_solvers.hpp:34-57.  See hc_oracle.h for the
conventions (basis, numbering, qdata layout).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libhcoracle.so")
_lib = None

PEC_ALL = 0x3F


def build(force: bool = False) -> str:
    """Compile libhcoracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)
    return _LIB_PATH


class _Mesh(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("vert", C.POINTER(C.c_double)),
                ("x0", C.c_double * 3), ("len", C.c_double * 3),
                ("pec_faces", C.c_int)]


class _Csr(C.Structure):
    _fields_ = [("n", C.c_int64), ("nnz", C.c_int64),
                ("rowptr", C.POINTER(C.c_int64)), ("col", C.POINTER(C.c_int64)),
                ("val", C.POINTER(C.c_double))]


class _Stats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int),
                ("rel_res", C.c_double), ("true_rel_res", C.c_double)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _lib.hco_ndofs.restype = C.c_int64
        _lib.hco_num_threads.restype = C.c_int
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _u8(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


class Mesh:
    """Structured hex mesh: box [x0, x0+len] or explicit vertex array."""

    def __init__(self, nx, ny=None, nz=None, vert=None, x0=(0, 0, 0),
                 length=(1, 1, 1), pec_faces=PEC_ALL):
        self.nx = int(nx)
        self.ny = int(ny if ny is not None else nx)
        self.nz = int(nz if nz is not None else nx)
        self.vert = None if vert is None else np.ascontiguousarray(vert, dtype=np.float64)
        if self.vert is not None:
            assert self.vert.shape == ((self.nz + 1) * (self.ny + 1) * (self.nx + 1), 3)
        self.x0 = tuple(float(v) for v in x0)
        self.len = tuple(float(v) for v in length)
        self.pec_faces = int(pec_faces)
        self._c = _Mesh(self.nx, self.ny, self.nz,
                        _dp(self.vert) if self.vert is not None else None,
                        (C.c_double * 3)(*self.x0), (C.c_double * 3)(*self.len),
                        self.pec_faces)

    @property
    def nelem(self):
        return self.nx * self.ny * self.nz

    def box_vertices(self):
        """Vertex array of the box mesh (x fastest)."""
        xs = self.x0[0] + self.len[0] * np.arange(self.nx + 1) / self.nx
        ys = self.x0[1] + self.len[1] * np.arange(self.ny + 1) / self.ny
        zs = self.x0[2] + self.len[2] * np.arange(self.nz + 1) / self.nz
        Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
        return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)

    @property
    def ptr(self):
        return C.byref(self._c)


def basis(p):
    Q = p + 1
    Bo = np.zeros(Q * p)
    Bc = np.zeros(Q * (p + 1))
    Gc = np.zeros(Q * (p + 1))
    qp = np.zeros(Q)
    qw = np.zeros(Q)
    rc = lib().hco_basis(p, _dp(Bo), _dp(Bc), _dp(Gc), _dp(qp), _dp(qw))
    assert rc == 0
    return (Bo.reshape(Q, p), Bc.reshape(Q, p + 1), Gc.reshape(Q, p + 1), qp, qw)


def gauss_legendre(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().hco_gauss_legendre(n, _dp(x), _dp(w))
    return x, w


def gauss_lobatto(n):
    x = np.zeros(n)
    lib().hco_gauss_lobatto(n, _dp(x))
    return x


def ndofs(m: Mesh, p):
    return int(lib().hco_ndofs(m.ptr, p))


def ess_mask(m: Mesh, p):
    e = np.zeros(ndofs(m, p), dtype=np.uint8)
    lib().hco_ess_mask(m.ptr, p, _u8(e))
    return e


def elem_dofs(m: Mesh, p, e):
    nd = 3 * p * (p + 1) ** 2
    g = np.zeros(nd, dtype=np.int64)
    lib().hco_elem_dofs(m.ptr, p, C.c_int64(e), g.ctypes.data_as(C.POINTER(C.c_int64)))
    return g


def elem_coef(m: Mesh, val):
    a = np.asarray(val, dtype=np.float64)
    if a.ndim == 0:
        a = np.full(m.nelem, float(a))
    assert a.shape == (m.nelem,)
    return np.ascontiguousarray(a)


def pa_setup(m: Mesh, p, curl_coef, mass_coef):
    Q = p + 1
    qd = np.zeros(m.nelem * Q ** 3 * 12)
    cc = elem_coef(m, curl_coef)
    mc = elem_coef(m, mass_coef)
    lib().hco_pa_setup(m.ptr, p, _dp(cc), _dp(mc), _dp(qd))
    return qd


def pa_apply(m: Mesh, p, qd, ess, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    lib().hco_pa_apply(m.ptr, p, _dp(qd), _u8(ess), _dp(x), _dp(y))
    return y


def pa_apply_raw(m: Mesh, p, qd, x, e0=0, e1=None):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    e1 = m.nelem if e1 is None else e1
    lib().hco_pa_apply_raw(m.ptr, p, _dp(qd), C.c_int64(e0), C.c_int64(e1), _dp(x), _dp(y))
    return y


def pa_diagonal(m: Mesh, p, qd, ess):
    d = np.zeros(ndofs(m, p))
    lib().hco_pa_diagonal(m.ptr, p, _dp(qd), _u8(ess), _dp(d))
    return d


def elem_matrix(m: Mesh, p, qd, e):
    nd = 3 * p * (p + 1) ** 2
    A = np.zeros((nd, nd))
    lib().hco_elem_matrix(m.ptr, p, _dp(qd), C.c_int64(e), _dp(A))
    return A


class Csr:
    """Oracle-owned CSR (full assembly with PEC elimination)."""

    def __init__(self, m: Mesh, p, qd, ess):
        self._c = _Csr()
        rc = lib().hco_csr_assemble(m.ptr, p, _dp(qd), _u8(ess), C.byref(self._c))
        assert rc == 0
        self.n = int(self._c.n)
        self.nnz = int(self._c.nnz)

    def apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        lib().hco_csr_apply(C.byref(self._c), _dp(x), _dp(y))
        return y

    def to_scipy(self):
        import scipy.sparse as sp
        n, nnz = self.n, self.nnz
        rp = np.ctypeslib.as_array(self._c.rowptr, shape=(n + 1,)).copy()
        col = np.ctypeslib.as_array(self._c.col, shape=(nnz,)).copy()
        val = np.ctypeslib.as_array(self._c.val, shape=(nnz,)).copy()
        return sp.csr_matrix((val, col, rp), shape=(n, n))

    def pcg(self, diag, b, rtol, maxit, x0=None, hist=False):
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64)
        st = _Stats()
        h = np.zeros(maxit + 1) if hist else None
        lib().hco_pcg_csr(C.byref(self._c), _dp(diag), _dp(np.ascontiguousarray(b)), _dp(x),
                          C.c_double(rtol), maxit, C.byref(st), _dp(h) if hist else None)
        return x, _stats(st, h)

    def __del__(self):
        try:
            lib().hco_csr_free(C.byref(self._c))
        except Exception:
            pass


def _stats(st, h):
    d = {"iterations": st.iterations, "converged": bool(st.converged),
         "rel_res": st.rel_res, "true_rel_res": st.true_rel_res}
    if h is not None:
        d["hist"] = h[: st.iterations + 1].copy()
    return d


def pcg_pa(m: Mesh, p, qd, ess, diag, b, rtol, maxit, x0=None, hist=False):
    n = ndofs(m, p)
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    st = _Stats()
    h = np.zeros(maxit + 1) if hist else None
    lib().hco_pcg_pa(m.ptr, p, _dp(qd), _u8(ess), _dp(diag), _dp(np.ascontiguousarray(b)),
                     _dp(x), C.c_double(rtol), maxit, C.byref(st), _dp(h) if hist else None)
    return x, _stats(st, h)


def rand_uniform(seed, n):
    out = np.zeros(n)
    lib().hco_rand_uniform(C.c_uint64(seed), C.c_int64(n), _dp(out))
    return out


def rhs_uniform(m: Mesh, p, seed, ess):
    b = np.zeros(ndofs(m, p))
    lib().hco_rhs_uniform(m.ptr, p, C.c_uint64(seed), _u8(ess), _dp(b))
    return b


def rhs_coil(m: Mesh, p, seed, ess):
    b = np.zeros(ndofs(m, p))
    lib().hco_rhs_coil(m.ptr, p, C.c_uint64(seed), _u8(ess), _dp(b))
    return b


def num_threads():
    return int(lib().hco_num_threads())


def two_region_coef(n, inner, outer, lo=0.25, hi=0.75):
    """Per-element coefficient of the two-region cube: `inner` on elements whose
    centroid lies in [lo, hi]^3, `outer` elsewhere (x fastest ordering)."""
    c = (np.arange(n) + 0.5) / n
    ins = (c >= lo) & (c <= hi)
    Z, Y, X = np.meshgrid(ins, ins, ins, indexing="ij")
    m = (X & Y & Z).ravel()
    return np.where(m, inner, outer).astype(np.float64)


# ---------------------------------------------------------------------------
# AV formulation blocks (src/formulations/AV/av_formulation.cpp:150-196): the
# A rows hold CurlCurl (dt alpha), VectorFEMass (beta) and MixedVectorGradient
# (sigma grad V, u') (src/kernels/mixed_vector_gradient_kernel.hpp:8); the V
# rows VectorFEWeakDivergence (sigma u, grad V')
# (src/kernels/vector_fe_weak_divergence_kernel.hpp:8) and Diffusion
# (sigma grad V, grad V') (src/kernels/diffusion_kernel.hpp:8).  V lives on the
# continuous Q_p space of the closed (Gauss-Lobatto) points, one node per point
# of the (n0 p + 1)(n1 p + 1)(n2 p + 1) grid, x fastest.  Its gradients lie in
# the ND_p space, so every block is a product of the ND mass matrix M_sigma, the
# curl-curl matrix and the discrete gradient G.
# ---------------------------------------------------------------------------
def _lagrange_derivatives(nodes, x):
    """D[i, m] = l_m'(x_i) for the Lagrange basis l_m on `nodes`:
    l_m'(x) = l_m(x) sum_{k != m} 1 / (x - x_k), evaluated by the product rule
    term by term (no division by x - x_k, so x may coincide with a node)."""
    n = len(nodes)
    D = np.zeros((len(x), n))
    for m in range(n):
        den = np.prod([nodes[m] - nodes[k] for k in range(n) if k != m])
        for i, xi in enumerate(x):
            s = 0.0
            for k in range(n):
                if k == m:
                    continue
                s += np.prod([xi - nodes[j] for j in range(n) if j not in (m, k)])
            D[i, m] = s / den
    return D


def node_dims(m: Mesh, p):
    return (m.nx * p + 1, m.ny * p + 1, m.nz * p + 1)


def grad_matrix(m: Mesh, p):
    """Discrete gradient G (H1 nodes -> ND DOFs) as scipy CSR.  The ND DOF of
    component c at open point i of element e along c (closed indices j, k on
    the other axes) is the tangential derivative in reference coordinates,
    sum_m l_m'(xi_i) phi(node e p + m along c, j, k): exact on Q_p fields."""
    import scipy.sparse as sp
    D = _lagrange_derivatives(gauss_lobatto(p + 1), gauss_legendre(p)[0])  # [open i][closed m]
    n = (m.nx, m.ny, m.nz)
    M = node_dims(m, p)
    rows, cols, vals = [], [], []
    off = 0
    for c in range(3):
        dims = [n[a] * p if a == c else n[a] * p + 1 for a in range(3)]
        k = np.arange(dims[0] * dims[1] * dims[2])
        idx = [k % dims[0], (k // dims[0]) % dims[1], k // (dims[0] * dims[1])]
        e, i = idx[c] // p, idx[c] % p
        for mm in range(p + 1):
            node = list(idx)
            node[c] = e * p + mm
            rows.append(off + k)
            cols.append((node[2] * M[1] + node[1]) * M[0] + node[0])
            vals.append(D[i, mm])
        off += k.size
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(off, M[0] * M[1] * M[2]))


def av_apply(m: Mesh, p, nu, sigma, dt, a, v):
    """AV block operator on (a = dA/dt on the ND DOFs, v = V on the nodes):
        ya = (dt K_nu + M_sigma) a0 + M_sigma G v,   ya = a on the PEC DOFs,
        yv = G^T M_sigma a0 + G^T M_sigma G v,
    a0 = a zeroed on the PEC DOFs (the elimination of
    src/equation_systems/equation_system.cpp:132-190); v has no essential nodes."""
    ess = ess_mask(m, p).astype(bool)
    G = grad_matrix(m, p)
    a = np.asarray(a, dtype=np.float64)
    a0 = np.where(ess, 0.0, a)
    s = a0 + G @ v
    Ms = pa_apply_raw(m, p, pa_setup(m, p, 0.0, sigma), s)
    Ka = pa_apply_raw(m, p, pa_setup(m, p, dt * elem_coef(m, nu), 0.0), a0)
    ya = np.where(ess, a, Ka + Ms)
    yv = G.T @ Ms
    return ya, yv
