"""Discretisation check of the HIP solve (needs an MI355X): the method of
manufactured solutions on hex meshes up to 32^3 (24^3 at order 3; solves to
rtol 1e-11, far below the discretisation error), the GPU-path counterpart of
the reference's MMS rate test (test/integration/test_aform_source.cpp:197-208,
window 1.85 < r < 3.0 for order 2).

u = (g(x) s(y) s(z), g(y) s(z) s(x), g(z) s(x) s(y)) with s = sin pi t and
g = cos pi t has zero tangential trace on the unit cube (PEC on every face)
and varies along every axis in every component (so ND_p reaches only its
order-p rate; a field constant along its own component's axis, such as
(sin pi y sin pi z, ...), converges one order faster).  With
curl curl u = grad div u - lap u,
  (curl curl u)_x = s'(x) g'(y) s(z) + s'(x) s(y) g'(z) + 2 pi^2 g(x) s(y) s(z)
(and cyclic), so every component of f = nu curl curl u + sigma u is a sum of
products of one-dimensional functions (checked against finite differences of
u), and the load vector b_i = (f, phi_i) of the tensor-product ND basis is a
sum of Kronecker products of assembled 1-D load vectors (h^2 from the Jacobian
and the covariant Piola factor 1/h); it is pinned against element-by-element
brute-force quadrature of tests/brute.py on a small mesh.  The L2 error of the
solution x is computed without evaluating the field: ||u_h - u||^2 =
x.M x - 2 x.(u, phi) + ||u||^2 with M the unit mass matrix (the library's
operator with nu = 0, sigma = 1) and ||u||^2 = 3/8.
The rate between consecutive meshes must exceed p - 0.15 for p = 1, 2, 3 and,
for p = 2, lie in the reference's window.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NU, SIG = 1.0, 1.0


@pytest.fixture(scope="module")
def env():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import hephaestus_amd as H
    H.load()
    return torch, H


def _gl(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return (x + 1.0) / 2.0, w / 2.0


def _gll(p):
    inner = np.polynomial.legendre.Legendre.basis(p).deriv().roots()
    return np.concatenate([[0.0], np.sort((inner + 1.0) / 2.0), [1.0]])


def _lagrange(nodes, t):
    """L[q, j] = l_j(t_q) for the Lagrange basis of the nodes."""
    L = np.ones((len(t), len(nodes)))
    for j, xj in enumerate(nodes):
        for m, xm in enumerate(nodes):
            if m != j:
                L[:, j] *= (t - xm) / (xj - xm)
    return L


def _load_1d(n, p, g, closed):
    """Assembled 1-D load: sum over elements of int_0^1 g(x(xi)) l_j(xi) dxi."""
    h = 1.0 / n
    nodes = _gll(p) if closed else _gl(p)[0]
    t, w = _gl(12)
    L = _lagrange(nodes, t)
    out = np.zeros(n * p + 1 if closed else n * p)
    for e in range(n):
        out[e * p + np.arange(len(nodes))] += (w * g((e + t) * h)) @ L
    return out


S = lambda t: np.sin(np.pi * t)
DS = lambda t: np.pi * np.cos(np.pi * t)
G = lambda t: np.cos(np.pi * t)
DG = lambda t: -np.pi * np.sin(np.pi * t)


def _terms_f():
    """f_x as (coefficient, a(x), b(y), c(z)) terms; f_y, f_z are its cyclic shifts."""
    return [(NU, DS, DG, S), (NU, DS, S, DG), (2 * np.pi ** 2 * NU + SIG, G, S, S)]


def _u_terms():
    return [(1.0, G, S, S)]


def _load(n, p, terms):
    """(sum of terms, phi_i) for all ND DOFs of the n^3 unit cube (structured
    numbering); a term of component x is (coef, a(x), b(y), c(z)) and is
    shifted cyclically for components y and z."""
    h = 1.0 / n
    out = []
    for c in range(3):
        bc = 0.0
        for (k, a0, a1, a2) in terms:
            # function of the component's own axis first, then the next two axes
            f = [None, None, None]
            f[c], f[(c + 1) % 3], f[(c + 2) % 3] = a0, a1, a2
            L = [_load_1d(n, p, f[ax], ax != c) for ax in range(3)]
            bc = bc + k * np.kron(L[2], np.kron(L[1], L[0]))
        out.append(bc)
    return h ** 2 * np.concatenate(out)


def _mms_rhs(n, p):
    return _load(n, p, _terms_f())


def _u_exact(X):
    x, y, z = X[:, 0], X[:, 1], X[:, 2]
    return np.stack([G(x) * S(y) * S(z), G(y) * S(z) * S(x), G(z) * S(x) * S(y)], axis=1)


def _f_exact(X):
    x, y, z = X[:, 0], X[:, 1], X[:, 2]
    out = []
    for (a, b, c) in ((x, y, z), (y, z, x), (z, x, y)):
        out.append(sum(k * fa(a) * fb(b) * fc(c) for (k, fa, fb, fc) in _terms_f()))
    return np.stack(out, axis=1)


def _ess_mask(n, p):
    """PEC on all faces: a DOF is essential on a face it is tangential to."""
    cl, op = n * p + 1, n * p
    masks = []
    for c in range(3):
        dims = [op if a == c else cl for a in range(3)]
        iz, iy, ix = np.meshgrid(*(np.arange(d) for d in dims[::-1]), indexing="ij")
        idx = [ix, iy, iz]
        m = np.zeros(ix.shape, bool)
        for a in range(3):
            if a != c:
                m |= (idx[a] == 0) | (idx[a] == dims[a] - 1)
        masks.append(m.ravel())
    return np.concatenate(masks)


def test_mms_source_is_curl_curl():
    """f = nu curl curl u + sigma u at random points, curl by central
    differences of u (h = 1e-4), and u vanishes tangentially on the faces."""
    rng = np.random.default_rng(3)
    X = rng.uniform(0.1, 0.9, (20, 3))
    d = 1e-4

    def curl(F, X):
        J = np.zeros((len(X), 3, 3))  # J[:, i, j] = d F_i / d x_j
        for j in range(3):
            e = np.zeros(3)
            e[j] = d
            J[:, :, j] = (F(X + e) - F(X - e)) / (2 * d)
        return np.stack([J[:, 2, 1] - J[:, 1, 2], J[:, 0, 2] - J[:, 2, 0], J[:, 1, 0] - J[:, 0, 1]], axis=1)

    cc = curl(lambda Y: curl(_u_exact, Y), X)
    f = NU * cc + SIG * _u_exact(X)
    assert np.abs(f - _f_exact(X)).max() <= 1e-5 * np.abs(_f_exact(X)).max()
    for ax in range(3):
        for v in (0.0, 1.0):
            Y = rng.uniform(0, 1, (10, 3))
            Y[:, ax] = v
            u = _u_exact(Y)
            tang = [c for c in range(3) if c != ax]
            assert np.abs(u[:, tang]).max() <= 1e-15


def test_mms_load_vector_vs_bruteforce():
    """The Kronecker load vector equals element-by-element quadrature of the
    explicitly built basis (tests/brute.py) on a 3^3 mesh, p = 1..3."""
    import brute as B
    for p in (1, 2, 3):
        n = 3
        h = 1.0 / n
        pts, w = B.quad_rule(p + 8)
        phi, _ = B.ref_basis(B.Basis1D(p), pts)
        N = 3 * (n * p) * (n * p + 1) ** 2
        b = np.zeros(N)
        for ez in range(n):
            for ey in range(n):
                for ex in range(n):
                    X = (pts + np.array([ex, ey, ez])) * h
                    f = _f_exact(X) * h ** 2
                    np.add.at(b, B.elem_gids((n, n, n), p, ex, ey, ez), np.einsum("iqr,qr,q->i", phi, f, w))
        k = _mms_rhs(n, p)
        assert np.linalg.norm(k - b) <= 1e-11 * np.linalg.norm(b), p
        assert np.array_equal(_ess_mask(n, p), B.ess_mask((n, n, n), p).astype(bool)), p


@pytest.mark.parametrize("p,ns", [(1, (8, 16, 32)), (2, (8, 16, 32)), (3, (6, 12, 24))])
def test_mms_rate_gpu(env, p, ns):
    torch, H = env
    errs = []
    for n in ns:
        print(f"case mms p={p} n={n}", flush=True)
        b = _mms_rhs(n, p)
        ess = _ess_mask(n, p)
        bu = _load(n, p, _u_terms())
        b[ess] = 0.0
        op = H.Operator(n, p, NU, SIG, precision=H.PREC_FP64).setup(1.0)
        bd = torch.from_numpy(b).cuda()
        x = torch.zeros_like(bd)
        st = op.solve(bd, x, rtol=1e-11, maxit=50000, mode=H.SOLVE_FP64 | H.SOLVE_AUX)
        assert st["converged"], st
        mass = H.Operator(n, p, 0.0, 1.0, precision=H.PREC_FP64).setup(1.0)
        Mx = mass.apply(x)
        e2 = float(torch.dot(x, Mx)) - 2.0 * float(torch.dot(x, torch.from_numpy(bu).cuda())) + 3.0 / 8.0
        errs.append(np.sqrt(max(e2, 0.0)))
        op.close()
        mass.close()
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    print(f"MMS GPU p={p} n={ns} L2 errors {errs} rates {rates}")
    assert rates[-1] > p - 0.15, (errs, rates)
    assert rates[-1] < p + 0.5, (errs, rates)
    if p == 2:
        assert 1.85 < rates[-1] < 3.0, (errs, rates)
