"""Pins of the oracle's operator, diagonal, CSR assembly and Jacobi-PCG (CPU).

Reference path restated by the oracle: the bilinear form of
CurlCurlEquationSystem::AddKernels (src/formulations/HCurl/hcurl_formulation.cpp:130-157),
PEC elimination of EquationSystem::FormLinearSystem
(src/equation_systems/equation_system.cpp:132-190), EquationSystem::Mult
(equation_system.cpp:200-204) and DefaultJacobiPCGSolver
(src/solvers/hephaestus_solvers.hpp:34-57).  MFEM and hypre are absent, so the
pins are an independent numpy route (tests/brute.py), textbook closed forms,
invariants and scipy.
"""
import numpy as np
import pytest
import scipy.linalg as sl
import scipy.sparse.linalg as spl

import brute as B


def _perturbed_mesh(oracle, n, seed, amp=0.08):
    rng = np.random.default_rng(seed)
    m0 = oracle.Mesh(*n)
    v = m0.box_vertices() + amp * rng.standard_normal(((n[0] + 1) * (n[1] + 1) * (n[2] + 1), 3)) / max(n)
    return oracle.Mesh(*n, vert=v), v


@pytest.mark.parametrize("p", [1, 2, 3])
def test_element_matrices_vs_bruteforce_trilinear(oracle, p):
    n = (2, 3, 2)
    m, vert = _perturbed_mesh(oracle, n, 10 + p)
    rng = np.random.default_rng(p)
    nu = rng.uniform(0.5, 2, m.nelem)
    sg = rng.uniform(0.5, 2, m.nelem)
    qd = oracle.pa_setup(m, p, nu, sg)
    for e in (0, 5, m.nelem - 1):
        Ae = oracle.elem_matrix(m, p, qd, e)
        idx = (e % n[0], (e // n[0]) % n[1], e // (n[0] * n[1]))
        K, M = B.element_matrices(vert, n, p, nu[e], sg[e], idx)
        assert np.abs(Ae - (K + M)).max() <= 1e-13 * np.abs(K + M).max()
        assert np.abs(Ae - Ae.T).max() <= 1e-14 * np.abs(Ae).max()


@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("pec", [0x3F, 0x15])
def test_pa_apply_diag_csr_vs_bruteforce(oracle, p, pec):
    n = (3, 2, 2)
    m, vert = _perturbed_mesh(oracle, n, 20 + p)
    m = oracle.Mesh(*n, vert=vert, pec_faces=pec)
    rng = np.random.default_rng(100 + p)
    nu = rng.uniform(0.5, 2, m.nelem)
    sg = 10.0 ** rng.uniform(-3, 3, m.nelem)
    qd = oracle.pa_setup(m, p, nu, sg)
    A, Kg, Mg, ess = B.assemble(n, p, nu, sg, vert=vert, pec_faces=pec)
    ess_o = oracle.ess_mask(m, p)
    np.testing.assert_array_equal(ess_o.astype(bool), ess)
    x = rng.standard_normal(A.shape[0])
    y = oracle.pa_apply(m, p, qd, ess_o, x)
    yr = A @ x
    assert np.linalg.norm(y - yr) <= 1e-14 * np.linalg.norm(yr) * 10
    # raw (no PEC) operator
    y_raw = oracle.pa_apply_raw(m, p, qd, x)
    assert np.linalg.norm(y_raw - (Kg + Mg) @ x) <= 1e-13 * np.linalg.norm(y_raw)
    d = oracle.pa_diagonal(m, p, qd, ess_o)
    np.testing.assert_allclose(d, A.diagonal(), rtol=1e-13, atol=0)
    Cs = oracle.Csr(m, p, qd, ess_o).to_scipy()
    assert abs(Cs - A).max() <= 1e-13 * abs(A).max()


@pytest.mark.parametrize("p", [1, 2, 3])
def test_symmetry(oracle, p):
    m = oracle.Mesh(3, 3, 2)
    rng = np.random.default_rng(7)
    qd = oracle.pa_setup(m, p, rng.uniform(1, 2, m.nelem), rng.uniform(1, 2, m.nelem))
    ess = oracle.ess_mask(m, p)
    N = oracle.ndofs(m, p)
    x1, x2 = rng.standard_normal(N), rng.standard_normal(N)
    a = x1 @ oracle.pa_apply(m, p, qd, ess, x2)
    b = x2 @ oracle.pa_apply(m, p, qd, ess, x1)
    assert abs(a - b) <= 1e-13 * abs(a)


def _whitney_brick(a, b, c):
    """Closed-form lowest-order (Whitney) brick matrices, DOF = edge circulation.

    Edge order: x-edges (dz,dy), y-edges (dz,dx), z-edges (dy,dx), each with the
    first index slowest.  m = [[1/3,1/6],[1/6,1/3]], s = [[1,-1],[-1,1]]."""
    m = np.array([[1 / 3, 1 / 6], [1 / 6, 1 / 3]])
    s = np.array([[1.0, -1.0], [-1.0, 1.0]])
    sg = np.array([-1.0, 1.0])  # integral of the derivative of the 1D hats
    M = np.zeros((12, 12))
    K = np.zeros((12, 12))
    M[0:4, 0:4] = b * c / a * np.kron(m, m)
    M[4:8, 4:8] = a * c / b * np.kron(m, m)
    M[8:12, 8:12] = a * b / c * np.kron(m, m)
    K[0:4, 0:4] = (1 / a) * (b / c * np.kron(s, m) + c / b * np.kron(m, s))
    K[4:8, 4:8] = (1 / b) * (a / c * np.kron(s, m) + c / a * np.kron(m, s))
    K[8:12, 8:12] = (1 / c) * (a / b * np.kron(s, m) + b / a * np.kron(m, s))
    for i in range(4):
        dz_i, dy_i = divmod(i, 2)
        for j in range(4):
            dz_j, dx_j = divmod(j, 2)
            K[i, 4 + j] = -c / (a * b) * sg[dy_i] * sg[dx_j] * m[dz_i, dz_j]
            dy_j, dx_j2 = divmod(j, 2)
            K[i, 8 + j] = -b / (a * c) * sg[dx_j2] * sg[dz_i] * m[dy_i, dy_j]
    for i in range(4):
        dz_i, dx_i = divmod(i, 2)
        for j in range(4):
            dy_j, dx_j = divmod(j, 2)
            K[4 + i, 8 + j] = -a / (b * c) * m[dx_i, dx_j] * sg[dy_j] * sg[dz_i]
    K = np.triu(K) + np.triu(K, 1).T
    return K, M


def test_whitney_brick_closed_form(oracle):
    a, b, c = 0.7, 1.3, 0.4
    m = oracle.Mesh(1, 1, 1, length=(a, b, c), pec_faces=0)
    qk = oracle.pa_setup(m, 1, 1.0, 0.0)
    qm = oracle.pa_setup(m, 1, 0.0, 1.0)
    Kr, Mr = _whitney_brick(a, b, c)
    np.testing.assert_allclose(oracle.elem_matrix(m, 1, qk, 0), Kr, atol=1e-14)
    np.testing.assert_allclose(oracle.elem_matrix(m, 1, qm, 0), Mr, atol=1e-14)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_gradient_null_space(oracle, p):
    n = (3, 2, 2)
    m = oracle.Mesh(*n)
    qd = oracle.pa_setup(m, p, np.ones(m.nelem), np.zeros(m.nelem))
    G, hd = B.h1_gradient(n, p)
    phi = np.random.default_rng(p).standard_normal(G.shape[1])
    iz, iy, ix = np.meshgrid(*(np.arange(d) for d in hd[::-1]), indexing="ij")
    bd = (ix == 0) | (iy == 0) | (iz == 0) | (ix == hd[0] - 1) | (iy == hd[1] - 1) | (iz == hd[2] - 1)
    phi[bd.ravel()] = 0
    u = G @ phi
    ess = oracle.ess_mask(m, p)
    assert np.all(u[ess.astype(bool)] == 0)  # gradients of H1_0 satisfy PEC
    x = np.random.default_rng(9).standard_normal(len(u))
    scale = np.linalg.norm(oracle.pa_apply_raw(m, p, qd, x)) / np.linalg.norm(x)
    assert np.linalg.norm(oracle.pa_apply_raw(m, p, qd, u)) <= 1e-13 * scale * np.linalg.norm(u)


_FIELDS = {
    1: (lambda x, y, z: (1 + y + z + y * z, 2 - x + 3 * z * x, 0.5 + x * y),
        lambda x, y, z: (-2 * x, 1 + 0 * y, -2 + 2 * z)),
    2: (lambda x, y, z: (x * y * y, y * z * z, z * x * x),
        lambda x, y, z: (-2 * y * z, -2 * z * x, -2 * x * y)),
    3: (lambda x, y, z: (x * x * y ** 3, y * y * z ** 3 + x, z * z * x ** 3 * y),
        lambda x, y, z: (z * z * x ** 3 - 3 * y * y * z * z,
                         -3 * z * z * x * x * y,
                         1 - 3 * x * x * y * y)),
}


@pytest.mark.parametrize("p", [1, 2, 3])
def test_energies_of_polynomial_fields(oracle, p):
    """x^T M x = int sigma |u|^2 and x^T K x = int nu |curl u|^2 for u in ND_p."""
    n, Lb = (2, 2, 3), (1.0, 1.5, 0.7)
    m = oracle.Mesh(*n, length=Lb)
    f, curl = _FIELDS[p]
    u = B.interpolate(n, p, f, Lb)
    qm = oracle.pa_setup(m, p, 0.0, 2.0)
    qk = oracle.pa_setup(m, p, 3.0, 0.0)
    eM = u @ oracle.pa_apply_raw(m, p, qm, u)
    eK = u @ oracle.pa_apply_raw(m, p, qk, u)
    pts, w = B.quad_rule(8)
    X, Y, Z = pts[:, 0] * Lb[0], pts[:, 1] * Lb[1], pts[:, 2] * Lb[2]
    w = w * np.prod(Lb)
    fx, fy, fz = f(X, Y, Z)
    cx, cy, cz = curl(X, Y, Z)
    assert abs(eM - 2 * np.sum(w * (fx ** 2 + fy ** 2 + fz ** 2))) <= 1e-13 * abs(eM)
    exK = 3 * np.sum(w * (cx ** 2 + cy ** 2 + cz ** 2))
    assert abs(eK - exK) <= 1e-12 * abs(exK)


@pytest.mark.parametrize("p,nn", [(2, 4), (3, 3)])
def test_maxwell_cavity_eigenvalues(oracle, p, nn):
    """PEC unit cube: K x = lambda M x has (n p - 1)^3 zero modes (gradients)
    and then 2pi^2 (x3), 3pi^2 (x2), 5pi^2 (x6), 6pi^2 (x6)."""
    m = oracle.Mesh(nn)
    ess = oracle.ess_mask(m, p).astype(bool)
    qk = oracle.pa_setup(m, p, 1.0, 0.0)
    qm = oracle.pa_setup(m, p, 0.0, 1.0)
    fr = np.where(~ess)[0]
    N = len(ess)
    K = np.zeros((len(fr), len(fr)))
    M = np.zeros((len(fr), len(fr)))
    for j, col in enumerate(fr):
        e = np.zeros(N)
        e[col] = 1.0
        K[:, j] = oracle.pa_apply_raw(m, p, qk, e)[fr]
        M[:, j] = oracle.pa_apply_raw(m, p, qm, e)[fr]
    ev = sl.eigh(K, M, eigvals_only=True)
    small = ev <= 1e-8 * ev.max()
    assert small.sum() == (nn * p - 1) ** 3
    nz = ev[~small][:17] / np.pi ** 2
    exact = np.array([2] * 3 + [3] * 2 + [5] * 6 + [6] * 6, dtype=float)
    np.testing.assert_allclose(nz, exact, rtol=1e-2)


def test_pcg_vs_scipy_config1(oracle):
    """Jacobi-PCG on the 16^3, p=1 cube with uniform coefficients: same solution as
    a direct solve, same iterates as the CSR route."""
    m = oracle.Mesh(16)
    p = 1
    qd = oracle.pa_setup(m, p, 1.0, 1.0)
    ess = oracle.ess_mask(m, p)
    d = oracle.pa_diagonal(m, p, qd, ess)
    b = oracle.rhs_uniform(m, p, 1, ess)
    x, st = oracle.pcg_pa(m, p, qd, ess, d, b, 1e-12, 2000, hist=True)
    assert st["converged"] and st["true_rel_res"] < 2e-12
    csr = oracle.Csr(m, p, qd, ess)
    A = csr.to_scipy()
    xd = spl.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - xd) <= 1e-10 * np.linalg.norm(xd)
    x2, st2 = csr.pcg(d, b, 1e-12, 2000, hist=True)
    # Two routes to the same operator differ in round-off only; CG's residual
    # oscillates near the end, so the stopping iteration may move a little.
    assert abs(st2["iterations"] - st["iterations"]) <= 0.03 * st["iterations"]
    # Finite-precision CG amplifies round-off once orthogonality is lost: the two
    # residual histories agree to ~1e-12 for the first ~80 iterations only.
    np.testing.assert_allclose(st["hist"][:60], st2["hist"][:60], rtol=1e-9)
    assert np.linalg.norm(x2 - x) <= 1e-10 * np.linalg.norm(x)
    # scipy's own Jacobi-PCG reaches the same solution
    Mi = spl.LinearOperator(A.shape, matvec=lambda v: v / d)
    xs, info = spl.cg(A, b, rtol=1e-12, maxiter=2000, M=Mi)
    assert info == 0
    assert np.linalg.norm(xs - xd) <= 1e-9 * np.linalg.norm(xd)


def test_coil_rhs_vs_bruteforce(oracle):
    """b_i = int J . phi_i, J = curl psi; compared with numpy quadrature, and the
    discrete divergence G^T b of a divergence-free source is small."""
    n, p = (4, 4, 4), 2
    m = oracle.Mesh(*n)
    ess = oracle.ess_mask(m, p)
    b = oracle.rhs_coil(m, p, 3, ess)
    # numpy route: same psi parameters from the splitmix64 stream
    r = oracle.rand_uniform(3, 28)
    u = (r + 1) / 2
    a = np.zeros((4, 3)); c = np.zeros((4, 3)); s = np.zeros(4)
    k = 0
    for j in range(4):
        a[j] = 2 * u[k:k + 3] - 1; k += 3
        c[j] = 0.3 + 0.4 * u[k:k + 3]; k += 3
        s[j] = 0.05 + 0.1 * u[k]; k += 1

    def J(X):
        out = np.zeros_like(X)
        for j in range(4):
            d = X - c[j]
            g = np.exp(-(d ** 2).sum(1) / (2 * s[j] ** 2))
            gr = -d / s[j] ** 2 * g[:, None]
            out += np.cross(gr, a[j])
        return out

    bb = B.Basis1D(p)
    pts, w = B.quad_rule(p + 2)
    phi, _ = B.ref_basis(bb, pts)
    br = np.zeros(len(b))
    h = 1.0 / n[0]
    for ez in range(n[2]):
        for ey in range(n[1]):
            for ex in range(n[0]):
                X = (pts + np.array([ex, ey, ez])) * h
                f = J(X) * h ** 3 / h  # u = phi_hat / h, dx = h^3 dxi
                g = B.elem_gids(n, p, ex, ey, ez)
                np.add.at(br, g, np.einsum("iqr,qr,q->i", phi, f, w))
    br[ess.astype(bool)] = 0
    assert np.linalg.norm(b - br) <= 1e-13 * np.linalg.norm(br)
    G, hd = B.h1_gradient(n, p)
    iz, iy, ix = np.meshgrid(*(np.arange(d) for d in hd[::-1]), indexing="ij")
    interior = ~((ix == 0) | (iy == 0) | (iz == 0) | (ix == hd[0] - 1) | (iy == hd[1] - 1) | (iz == hd[2] - 1)).ravel()
    div = (G.T @ b)[interior]
    assert np.linalg.norm(div) <= 0.1 * np.linalg.norm(b)


@pytest.mark.slow
@pytest.mark.parametrize("p", [1, 2])
def test_mms_convergence_rate(oracle, p):
    """PEC-compatible MMS u = (sin pi y sin pi z, sin pi z sin pi x, sin pi x sin pi y):
    curl curl u = 2 pi^2 u, so f = (2 pi^2 nu + sigma) u.  The L2 error rate of the
    Jacobi-PCG solution is about p (the reference's MMS rate window idea,
    test/integration/test_aform_source.cpp:197-208)."""
    nu, sig = 1.0, 1.0
    pi = np.pi

    def uex(X):
        x, y, z = X[:, 0], X[:, 1], X[:, 2]
        return np.stack([np.sin(pi * y) * np.sin(pi * z), np.sin(pi * z) * np.sin(pi * x),
                         np.sin(pi * x) * np.sin(pi * y)], axis=1)

    errs = []
    ns = [2, 4, 8]
    bb = B.Basis1D(p)
    pts, w = B.quad_rule(p + 3)
    phi, _ = B.ref_basis(bb, pts)
    for nn in ns:
        n = (nn, nn, nn)
        m = oracle.Mesh(nn)
        h = 1.0 / nn
        ess = oracle.ess_mask(m, p)
        qd = oracle.pa_setup(m, p, nu, sig)
        d = oracle.pa_diagonal(m, p, qd, ess)
        b = np.zeros(oracle.ndofs(m, p))
        for ez in range(nn):
            for ey in range(nn):
                for ex in range(nn):
                    X = (pts + np.array([ex, ey, ez])) * h
                    f = (2 * pi ** 2 * nu + sig) * uex(X) * h ** 2
                    np.add.at(b, B.elem_gids(n, p, ex, ey, ez), np.einsum("iqr,qr,q->i", phi, f, w))
        b[ess.astype(bool)] = 0
        x, st = oracle.pcg_pa(m, p, qd, ess, d, b, 1e-13, 5000)
        assert st["converged"]
        e2 = 0.0
        for ez in range(nn):
            for ey in range(nn):
                for ex in range(nn):
                    X = (pts + np.array([ex, ey, ez])) * h
                    uh = np.einsum("i,iqr->qr", x[B.elem_gids(n, p, ex, ey, ez)], phi) / h
                    e2 += np.sum(w * ((uh - uex(X)) ** 2).sum(1)) * h ** 3
        errs.append(np.sqrt(e2))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    print("MMS p=%d errors %s rates %s" % (p, errs, rates))
    assert rates[-1] > p - 0.15, (errs, rates)
    assert rates[-1] < p + 1.5, (errs, rates)
