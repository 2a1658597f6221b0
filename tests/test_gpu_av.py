"""AV formulation block operator on the GPU (needs an MI355X), through the C ABI
(hc_op_av_apply): against the oracle's av_apply on small box and trilinear
meshes, orders 1..3, PEC on all faces and on a subset (<= 1e-13 relative), and
at 48^3 order 3 (9,082,800 ND DOFs, 3,048,625 nodes) by its gauge null space
(a, v) = (-G phi, phi) without PEC and its symmetry with PEC.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import hephaestus_amd as H
    H.load()
    from oracle import oracle as O
    O.build()
    return torch, H, O


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


CASES = [(1, (5, 4, 3), 0x3F, False), (2, (4, 3, 5), 0x3F, True), (3, (3, 4, 3), 0x3F, False),
         (2, (4, 4, 3), 0x19, False), (3, (3, 3, 4), 0x26, True), (1, (6, 3, 4), 0, True)]


@pytest.mark.parametrize("p,n,pec,tri", CASES)
def test_av_apply_vs_oracle(env, p, n, pec, tri):
    torch, H, O = env
    print(f"case av p={p} n={n} pec={pec:#x} trilinear={tri}", flush=True)
    rng = np.random.default_rng(900 + p + pec)
    m0 = O.Mesh(*n)
    vert = None
    if tri:
        vert = m0.box_vertices() + 0.04 * rng.standard_normal(((n[0] + 1) * (n[1] + 1) * (n[2] + 1), 3)) / max(n)
    m = O.Mesh(*n, vert=vert, pec_faces=pec)
    E = m.nelem
    nu, sig = rng.uniform(0.5, 2.0, E), 10.0 ** rng.uniform(-3, 3, E)
    dt = 0.37
    a = rng.standard_normal(O.ndofs(m, p))
    v = rng.standard_normal(int(np.prod(O.node_dims(m, p))))
    ya_ref, yv_ref = O.av_apply(m, p, nu, sig, dt, a, v)
    op = H.Operator(n, p, nu, sig, pec_faces=pec, vert=vert).setup(dt)
    assert op.nnodes == v.size
    ya, yv = op.av_apply(torch.from_numpy(a).cuda(), torch.from_numpy(v).cuda())
    assert _rel(ya.cpu().numpy(), ya_ref) <= 1e-13
    assert _rel(yv.cpu().numpy(), yv_ref) <= 1e-13


def _gll(p):
    inner = np.polynomial.legendre.Legendre.basis(p).deriv().roots()
    return np.concatenate([[0.0], np.sort((inner + 1.0) / 2.0), [1.0]])


def test_av_apply_full_size_properties(env):
    """48^3 order 3.  Gauge null space without PEC: phi = x + 2y - z at the
    nodes has G phi = (h, 2h, -h) on the x, y, z DOFs (DOF = h times the
    tangential component), so (a, v) = (-G phi, phi) maps to zero.  Symmetry
    with PEC on all faces on PEC-consistent random vectors."""
    torch, H, O = env
    n, p = 48, 3
    print(f"case av full size n={n} p={p}", flush=True)
    sig = O.two_region_coef(n, 1e4, 1e-6)
    h = 1.0 / n
    op = H.Operator(n, p, 1.0, sig, pec_faces=0).setup(1.0)
    N, NN = op.ndofs, op.nnodes
    assert NN == (n * p + 1) ** 3
    e = np.minimum(np.arange(n * p + 1) // p, n - 1)
    xs = (e + _gll(p)[np.arange(n * p + 1) - e * p]) * h
    Z, Y, X = np.meshgrid(xs, xs, xs, indexing="ij")
    phi = torch.from_numpy((X + 2.0 * Y - Z).ravel()).cuda()
    nc = N // 3   # the three components have equal counts on a cube
    a = torch.empty(N, dtype=torch.float64, device="cuda")
    a[:nc] = -h
    a[nc:2 * nc] = -2.0 * h
    a[2 * nc:] = h
    ya, yv = op.av_apply(a, phi)
    r = torch.empty_like(a)
    op.fill_uniform(9, r)
    w = torch.empty_like(phi)
    w.uniform_(-1.0, 1.0, generator=torch.Generator(device="cuda").manual_seed(4))
    za, zv = op.av_apply(r, w)
    assert float(torch.linalg.norm(ya)) <= 1e-12 * float(torch.linalg.norm(za))
    assert float(torch.linalg.norm(yv)) <= 1e-12 * float(torch.linalg.norm(zv))
    del op
    op = H.Operator(n, p, 1.0, sig).setup(1.0)
    a1, a2 = torch.empty(N, dtype=torch.float64, device="cuda"), torch.empty(N, dtype=torch.float64, device="cuda")
    op.fill_uniform(11, a1)
    op.fill_uniform(12, a2)
    v2 = torch.empty_like(phi)
    v2.uniform_(-1.0, 1.0, generator=torch.Generator(device="cuda").manual_seed(5))
    ya1, yv1 = op.av_apply(a1, w)
    ya2, yv2 = op.av_apply(a2, v2)
    lhs = float(torch.dot(a1, ya2) + torch.dot(w, yv2))
    rhs = float(torch.dot(ya1, a2) + torch.dot(yv1, v2))
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + abs(rhs))


@pytest.mark.parametrize("p,n,pec", [(1, (6, 5, 4), 0x3F), (2, (4, 4, 3), 0x3F), (3, (3, 3, 3), 0x3F),
                                     (2, (4, 3, 4), 0x19)])
def test_av_solve_consistent_rhs(env, p, n, pec):
    """b = AV x_true; PCG from zero converges to rtol in the true residual
    (recomputed by the oracle's av_apply), and the gauge-invariant parts of
    the solution equal those of x_true: s = a0 + G v (the field sigma acts on)
    and a on the PEC DOFs."""
    torch, H, O = env
    print(f"case av solve p={p} n={n} pec={pec:#x}", flush=True)
    rng = np.random.default_rng(1200 + p + pec)
    m = O.Mesh(*n, pec_faces=pec)
    E = m.nelem
    nu, sig = rng.uniform(0.5, 2.0, E), 10.0 ** rng.uniform(-1, 1, E)
    dt = 0.5
    a_t = rng.standard_normal(O.ndofs(m, p))
    v_t = rng.standard_normal(int(np.prod(O.node_dims(m, p))))
    ba, bv = O.av_apply(m, p, nu, sig, dt, a_t, v_t)
    op = H.Operator(n, p, nu, sig, pec_faces=pec).setup(dt)
    a = torch.zeros(a_t.size, dtype=torch.float64, device="cuda")
    v = torch.zeros(v_t.size, dtype=torch.float64, device="cuda")
    rtol = 1e-10
    st = op.av_solve(torch.from_numpy(ba).cuda(), torch.from_numpy(bv).cuda(), a, v, rtol=rtol, maxit=20000)
    assert st["converged"] and st["rel_res"] <= rtol * 1.01
    ah, vh = a.cpu().numpy(), v.cpu().numpy()
    ra, rv = O.av_apply(m, p, nu, sig, dt, ah, vh)
    res = np.sqrt(np.sum((ba - ra) ** 2) + np.sum((bv - rv) ** 2)) / np.sqrt(np.sum(ba ** 2) + np.sum(bv ** 2))
    assert res <= rtol * 1.05
    ess = O.ess_mask(m, p).astype(bool)
    G = O.grad_matrix(m, p)
    s_t = np.where(ess, 0.0, a_t) + G @ v_t
    s = np.where(ess, 0.0, ah) + G @ vh
    assert _rel(s, s_t) <= 1e-6
    assert np.abs(ah[ess] - a_t[ess]).max() <= 1e-9 * np.abs(a_t).max()
    # a second run from the same start is bitwise identical
    a2, v2 = torch.zeros_like(a), torch.zeros_like(v)
    st2 = op.av_solve(torch.from_numpy(ba).cuda(), torch.from_numpy(bv).cuda(), a2, v2, rtol=rtol, maxit=20000)
    assert st2["iterations"] == st["iterations"] and torch.equal(a2, a) and torch.equal(v2, v)


@pytest.mark.parametrize("p,n,tri", [(1, (6, 5, 4), False), (2, (4, 4, 3), True), (3, (3, 3, 4), False),
                                     (2, (12, 12, 12), False)])
def test_div_free_source(env, p, n, tri):
    """DivFreeSource::Apply on j = M x for a random (not divergence-free) field
    x: the mass solve recovers g + G Q = x, the projected g is discretely
    divergence-free (G^T M g = 0) while x is not, and lf gains M g; M (unit
    coefficient) and G from the oracle.  Box meshes precondition the
    projection with the node-space V-cycle (three levels at 12^3), the
    trilinear mesh with Jacobi."""
    torch, H, O = env
    print(f"case div-free source p={p} n={n} trilinear={tri}", flush=True)
    rng = np.random.default_rng(1500 + p)
    m0 = O.Mesh(*n)
    vert = None
    if tri:
        vert = m0.box_vertices() + 0.04 * rng.standard_normal(((n[0] + 1) * (n[1] + 1) * (n[2] + 1), 3)) / max(n)
    m = O.Mesh(*n, vert=vert)
    E = m.nelem
    N = O.ndofs(m, p)
    M = O.Csr(m, p, O.pa_setup(m, p, 0.0, 1.0), np.zeros(N, dtype=np.uint8)).to_scipy().tocsr()
    G = O.grad_matrix(m, p)
    x = rng.standard_normal(N)
    j = M @ x
    lf0 = rng.standard_normal(N)
    op = H.Operator(n, p, rng.uniform(0.5, 2.0, E), 10.0 ** rng.uniform(-2, 2, E), vert=vert).setup(0.3)
    lf = torch.from_numpy(lf0.copy()).cuda()
    g, Q, st = op.div_free_source(torch.from_numpy(j).cuda(), lf, project=True, rtol=1e-12, maxit=20000)
    print(f"  {st['iterations']} iterations (mass and projection solves)")
    assert st["converged"]
    gh, Qh, lfh = g.cpu().numpy(), Q.cpu().numpy(), lf.cpu().numpy()
    assert _rel(gh + G @ Qh, x) <= 1e-8
    div_x = G.T @ (M @ x)
    assert np.linalg.norm(G.T @ (M @ gh)) <= 1e-9 * np.linalg.norm(div_x)
    assert _rel(lfh - lf0, M @ gh) <= 1e-10
    # without the projection g is the mass solve alone
    lf2 = torch.from_numpy(lf0.copy()).cuda()
    g2, _, st2 = op.div_free_source(torch.from_numpy(j).cuda(), lf2, project=False, rtol=1e-12, maxit=20000)
    assert st2["converged"] and _rel(g2.cpu().numpy(), x) <= 1e-9
    # a solve that cannot converge reports an error and leaves lf unchanged
    lf3 = torch.from_numpy(lf0.copy()).cuda()
    with pytest.raises(H.HcError):
        op.div_free_source(torch.from_numpy(j).cuda(), lf3, project=True, rtol=1e-14, maxit=1)
    assert torch.equal(lf3.cpu(), torch.from_numpy(lf0))


@pytest.mark.parametrize("n,p", [(16, 2), (12, 3)])
def test_av_solve_multigrid_blocks(env, n, p):
    """On box meshes with multigrid hierarchies the AV solve uses the
    block-diagonal multigrid preconditioner (outer_iterations = 2): two-region
    sigma, b = AV x_true; it reaches rtol 1e-10 on the oracle's av_apply in at
    most 40 iterations, the gauge-invariant field s = a0 + G v matches x_true's,
    and a repeated solve is bitwise identical."""
    torch, H, O = env
    print(f"case av solve multigrid {n}^3 p={p}", flush=True)
    rng = np.random.default_rng(1300 + n + p)
    m = O.Mesh(n)
    c = (np.arange(n) + 0.5) / n
    ins = (c >= 0.25) & (c <= 0.75)
    Z, Y, X = np.meshgrid(ins, ins, ins, indexing="ij")
    inside = (X & Y & Z).ravel()
    nu = np.where(inside, 0.5, 1.0)
    sig = np.where(inside, 1e2, 1e-2)
    dt = 0.1
    a_t = rng.standard_normal(O.ndofs(m, p))
    v_t = rng.standard_normal(int(np.prod(O.node_dims(m, p))))
    ba, bv = O.av_apply(m, p, nu, sig, dt, a_t, v_t)
    op = H.Operator(n, p, nu, sig).setup(dt)
    a = torch.zeros(a_t.size, dtype=torch.float64, device="cuda")
    v = torch.zeros(v_t.size, dtype=torch.float64, device="cuda")
    st = op.av_solve(torch.from_numpy(ba).cuda(), torch.from_numpy(bv).cuda(), a, v, rtol=1e-10, maxit=2000)
    print(f"  {st['iterations']} iterations, preconditioner {'multigrid' if st['outer_iterations'] == 2 else 'Jacobi'}")
    assert st["outer_iterations"] == 2 and st["converged"] and st["iterations"] <= 40
    ah, vh = a.cpu().numpy(), v.cpu().numpy()
    ra, rv = O.av_apply(m, p, nu, sig, dt, ah, vh)
    res = np.sqrt(np.sum((ba - ra) ** 2) + np.sum((bv - rv) ** 2)) / np.sqrt(np.sum(ba ** 2) + np.sum(bv ** 2))
    assert res <= 1e-10 * 1.05
    ess = O.ess_mask(m, p).astype(bool)
    G = O.grad_matrix(m, p)
    s_t = np.where(ess, 0.0, a_t) + G @ v_t
    s = np.where(ess, 0.0, ah) + G @ vh
    assert _rel(s, s_t) <= 1e-6
    a2, v2 = torch.zeros_like(a), torch.zeros_like(v)
    st2 = op.av_solve(torch.from_numpy(ba).cuda(), torch.from_numpy(bv).cuda(), a2, v2, rtol=1e-10, maxit=2000)
    assert st2["iterations"] == st["iterations"] and torch.equal(a2, a) and torch.equal(v2, v)
    op.close()
