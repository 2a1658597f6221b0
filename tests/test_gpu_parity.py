"""HIP operator path vs the oracle and the golden fixtures (needs an MI355X).

Every call goes through the C ABI of libhcurl.so (include/hc_hcurl.h) via
hephaestus_amd.hcurl.  Tolerances: FP64 apply/diagonal relative L2 <= 1e-13
against the oracle and the golden fixtures (the bar is 1e-12); FP32 apply
relative L2 <= 2e-6; PCG solutions against a direct solve within what the
stopping tolerance allows; repeated applies bitwise identical.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.fixture(scope="module")
def hc():
    import hephaestus_amd as H
    H.load()
    return H


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _dev(torch, a, dtype=None):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype=dtype or torch.float64)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_golden_tiny_trilinear(torch_cuda, hc, p):
    torch = torch_cuda
    g = dict(np.load(os.path.join(GOLD, f"tiny_p{p}.npz")))
    n = tuple(int(v) for v in g["n"])
    print(f"case tiny p={p} n={n}", flush=True)
    op = hc.Operator(n, p, g["nu"], g["sig"], vert=g["vert"]).setup(1.0)
    x = _dev(torch, g["x"])
    y = op.apply(x).cpu().numpy()
    assert _rel(y, g["y"]) <= 1e-13
    d = op.diagonal().cpu().numpy()
    assert _rel(d, g["diag"]) <= 1e-13
    assert op.ess_count == int(g["ess"].sum())


def test_golden_config1(torch_cuda, hc):
    torch = torch_cuda
    g = dict(np.load(os.path.join(GOLD, "config1_16_p1.npz")))
    print("case config1 16^3 p=1", flush=True)
    op = hc.Operator(16, 1, 1.0, 1.0).setup(1.0)
    b = torch.empty(op.ndofs, dtype=torch.float64, device="cuda")
    op.fill_uniform(1, b)
    np.testing.assert_array_equal(b.cpu().numpy(), g["b"])
    y = op.apply(b).cpu().numpy()
    assert _rel(y, g["Ab"]) <= 1e-13
    assert _rel(op.diagonal().cpu().numpy(), g["diag"]) <= 1e-13
    x = torch.zeros_like(b)
    st = op.solve(b, x, rtol=1e-10, maxit=5000)
    assert st["converged"] and st["rel_res"] <= 1e-10
    assert _rel(x.cpu().numpy(), g["x_direct"]) <= 1e-8
    # FP32 operator + FP64 refinement reaches the same solution
    x2 = torch.zeros_like(b)
    st2 = op.solve(b, x2, rtol=1e-10, maxit=20000, mode=hc.SOLVE_FP32_IR)
    assert st2["converged"] and st2["rel_res"] <= 1e-10 and st2["outer_iterations"] >= 2
    assert _rel(x2.cpu().numpy(), g["x_direct"]) <= 1e-8


def test_golden_two_region(torch_cuda, hc):
    torch = torch_cuda
    g = dict(np.load(os.path.join(GOLD, "two_region_6_p3.npz")))
    print("case two-region 6^3 p=3", flush=True)
    op = hc.Operator(6, 3, g["nu"], g["sig"]).setup(1.0)
    y = op.apply(_dev(torch, g["x"])).cpu().numpy()
    assert _rel(y, g["y"]) <= 1e-13
    assert _rel(op.diagonal().cpu().numpy(), g["diag"]) <= 1e-13


# meshes with several tiles per direction, partial tiles, trilinear geometry,
# partial PEC
CASES = [
    (1, (9, 7, 5), 0x3F, True),
    (2, (9, 6, 5), 0x15, True),
    (3, (9, 5, 6), 0x3F, True),
    (3, (8, 8, 8), 0x2A, False),
    (2, (13, 4, 3), 0x00, False),
]


@pytest.mark.parametrize("p,n,pec,perturb", CASES)
def test_apply_vs_oracle(torch_cuda, hc, oracle, p, n, pec, perturb):
    torch = torch_cuda
    print(f"case apply p={p} n={n} pec={pec:#x} perturb={perturb}", flush=True)
    rng = np.random.default_rng(sum(n) * 10 + p)
    m0 = oracle.Mesh(*n)
    vert = None
    if perturb:
        vert = m0.box_vertices() + 0.06 * rng.standard_normal(((n[0] + 1) * (n[1] + 1) * (n[2] + 1), 3)) / max(n)
    m = oracle.Mesh(*n, vert=vert, pec_faces=pec)
    E = m.nelem
    nu = rng.uniform(0.5, 2.0, E)
    sig = 10.0 ** rng.uniform(-4, 4, E)
    dt = 0.37
    qd = oracle.pa_setup(m, p, dt * nu, sig)
    ess = oracle.ess_mask(m, p)
    x = oracle.rand_uniform(11, oracle.ndofs(m, p))
    y_ref = oracle.pa_apply(m, p, qd, ess, x)
    d_ref = oracle.pa_diagonal(m, p, qd, ess)
    op = hc.Operator(n, p, nu, sig, pec_faces=pec, vert=vert).setup(dt)
    assert op.ndofs == len(x)
    assert op.ess_count == int(ess.sum())
    xd = _dev(torch, x)
    y1 = op.apply(xd)
    y2 = op.apply(xd)
    assert torch.equal(y1, y2), "apply is not bitwise reproducible"
    assert _rel(y1.cpu().numpy(), y_ref) <= 1e-13
    assert _rel(op.diagonal().cpu().numpy(), d_ref) <= 1e-13
    y32 = op.apply_f32(xd.float()).double().cpu().numpy()
    assert _rel(y32, y_ref) <= 2e-6


@pytest.mark.parametrize("p", [1, 2, 3])
def test_every_tile_variant(torch_cuda, hc, oracle, p):
    """Each compiled tile shape of the apply kernel gives the oracle's result
    (several tiles per direction, partial tiles, trilinear geometry)."""
    torch = torch_cuda
    from hephaestus_amd import hcurl
    n = {1: (11, 9, 6), 2: (9, 7, 5), 3: (9, 5, 7)}[p]
    print(f"case tiles p={p} n={n}", flush=True)
    rng = np.random.default_rng(300 + p)
    m0 = oracle.Mesh(*n)
    vert = m0.box_vertices() + 0.05 * rng.standard_normal(((n[0] + 1) * (n[1] + 1) * (n[2] + 1), 3)) / max(n)
    m = oracle.Mesh(*n, vert=vert)
    E = m.nelem
    nu, sig = rng.uniform(0.5, 2.0, E), 10.0 ** rng.uniform(-3, 3, E)
    qd = oracle.pa_setup(m, p, nu, sig)
    ess = oracle.ess_mask(m, p)
    x = oracle.rand_uniform(21, oracle.ndofs(m, p))
    y_ref = oracle.pa_apply(m, p, qd, ess, x)
    op = hc.Operator(n, p, nu, sig, vert=vert).setup(1.0)
    xd = _dev(torch, x)
    tiles = hcurl.tile_list(p)
    assert len(tiles) >= 2
    for (tx, ty, tz, pf) in tiles:
        op.set_tile(64, tx, ty, tz, pf)
        op.set_tile(32, tx, ty, tz, pf)
        assert _rel(op.apply(xd).cpu().numpy(), y_ref) <= 1e-13, (tx, ty, tz, pf)
        assert _rel(op.apply_f32(xd.float()).double().cpu().numpy(), y_ref) <= 2e-6, (tx, ty, tz, pf)
    info = op.autotune(reps=2)
    assert info["tile64"] in [t for t in tiles]


@pytest.mark.parametrize("p", [1, 2, 3])
def test_many_tiles_multi_rhs(torch_cuda, hc, oracle, p):
    """Every tile variant on a mesh of several hundred tiles (more workgroups
    than the XCD renumbering's 8 groups times the CUs), three right-hand sides
    in one launch over (tile, rhs) work items: oracle parity, and the single
    rhs apply equals its row of the multi-rhs apply bitwise."""
    torch = torch_cuda
    from hephaestus_amd import hcurl
    n = {1: (40, 36, 30), 2: (30, 26, 22), 3: (24, 22, 20)}[p]
    print(f"case many tiles p={p} n={n}", flush=True)
    rng = np.random.default_rng(500 + p)
    m = oracle.Mesh(*n)
    E = m.nelem
    nu, sig = rng.uniform(0.5, 2.0, E), 10.0 ** rng.uniform(-3, 3, E)
    qd = oracle.pa_setup(m, p, nu, sig)
    ess = oracle.ess_mask(m, p)
    N = oracle.ndofs(m, p)
    X = np.stack([oracle.rand_uniform(31 + r, N) for r in range(3)])
    Y = np.stack([oracle.pa_apply(m, p, qd, ess, X[r]) for r in range(3)])
    op = hc.Operator(n, p, nu, sig).setup(1.0)
    op.set_rhs_block(1)
    xd = _dev(torch, X)
    for (tx, ty, tz, pf) in hcurl.tile_list(p):
        op.set_tile(64, tx, ty, tz, pf)
        y = op.apply(xd)
        for r in range(3):
            assert _rel(y[r].cpu().numpy(), Y[r]) <= 1e-13, (tx, ty, tz, pf, r)
        assert torch.equal(op.apply(xd[1]), y[1]), (tx, ty, tz, pf)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_tile_variants_repeatable(torch_cuda, hc, p):
    """Every tile variant gives bitwise identical results over repeated applies
    on a mesh of several hundred tiles (a race between the waves of a
    workgroup shows up as an occasional difference)."""
    torch = torch_cuda
    from hephaestus_amd import hcurl
    n = {1: (40, 36, 30), 2: (30, 26, 22), 3: (24, 22, 20)}[p]
    print(f"case repeatable tiles p={p} n={n}", flush=True)
    rng = np.random.default_rng(600 + p)
    E = n[0] * n[1] * n[2]
    op = hc.Operator(n, p, rng.uniform(0.5, 2.0, E), 10.0 ** rng.uniform(-3, 3, E)).setup(1.0)
    x = torch.empty(op.ndofs, dtype=torch.float64, device="cuda")
    op.fill_uniform(41, x)
    x32 = x.float()
    for t in hcurl.tile_list(p):
        op.set_tile(64, *t)
        op.set_tile(32, *t)
        y0, z0 = op.apply(x), op.apply_f32(x32)
        for _ in range(8):
            assert torch.equal(op.apply(x), y0), t
            assert torch.equal(op.apply_f32(x32), z0), t


def test_setup_dt_rule(torch_cuda, hc, oracle):
    """hc_op_setup re-assembles only when dt changes by more than 1e-12 dt
    (TimeDependentEquationSystem::SetTimeStep, equation_system.cpp:413-426)."""
    torch = torch_cuda
    n, p = (5, 4, 3), 2
    m = oracle.Mesh(*n)
    op = hc.Operator(n, p, 1.5, 2.0).setup(1.0)
    x = _dev(torch, oracle.rand_uniform(5, op.ndofs))
    y1 = op.apply(x).cpu().numpy()
    op.setup(2.0)
    y2 = op.apply(x).cpu().numpy()
    ess = oracle.ess_mask(m, p)
    ref2 = oracle.pa_apply(m, p, oracle.pa_setup(m, p, 2.0 * 1.5, 2.0), ess, x.cpu().numpy())
    assert _rel(y2, ref2) <= 1e-13
    assert _rel(y1, y2) > 1e-3
    op.set_coefficients(3.0, 1.0)
    op.setup(2.0)  # coefficients changed: re-assembled although dt is the same
    ref3 = oracle.pa_apply(m, p, oracle.pa_setup(m, p, 6.0, 1.0), ess, x.cpu().numpy())
    assert _rel(op.apply(x).cpu().numpy(), ref3) <= 1e-13


@pytest.mark.parametrize("p", [1, 2, 3])
def test_pcg_vs_oracle(torch_cuda, hc, oracle, p):
    torch = torch_cuda
    n = {1: 12, 2: 8, 3: 6}[p]
    print(f"case pcg p={p} n={n}", flush=True)
    m = oracle.Mesh(n)
    sig = oracle.two_region_coef(n, 10.0, 0.1)
    qd = oracle.pa_setup(m, p, 1.0, sig)
    ess = oracle.ess_mask(m, p)
    d = oracle.pa_diagonal(m, p, qd, ess)
    b = oracle.rhs_uniform(m, p, 2, ess)
    x_ref, st_ref = oracle.pcg_pa(m, p, qd, ess, d, b, 1e-11, 20000)
    op = hc.Operator(n, p, 1.0, sig).setup(1.0)
    bd = _dev(torch, b)
    x = torch.zeros_like(bd)
    st = op.solve(bd, x, rtol=1e-11, maxit=20000)
    assert st["converged"] and st["rel_res"] <= 1e-11
    assert abs(st["iterations"] - st_ref["iterations"]) <= max(5, 0.08 * st_ref["iterations"])
    assert _rel(x.cpu().numpy(), x_ref) <= 1e-9
    # FP32 apply + FP64 refinement: the same solution within the stated tolerance
    x2 = torch.zeros_like(bd)
    st2 = op.solve(bd, x2, rtol=1e-10, maxit=50000, mode=hc.SOLVE_FP32_IR)
    assert st2["converged"] and st2["rel_res"] <= 1e-10
    assert _rel(x2.cpu().numpy(), x_ref) <= 1e-7


def test_errors_are_reported(torch_cuda, hc):
    with pytest.raises(hc.HcError) as ei:
        hc.Operator(4, 5, 1.0, 1.0)
    assert ei.value.code == -1
    op = hc.Operator(4, 2, 1.0, 1.0, precision=hc.PREC_FP64)
    import torch
    x = torch.zeros(op.ndofs, dtype=torch.float64, device="cuda")
    with pytest.raises(hc.HcError):
        op.apply(x)  # no setup yet
    op.setup(1.0)
    with pytest.raises(hc.HcError):
        op.apply_f32(x.float())  # created without the FP32 copy
    # maxit too small -> not converged, status reported, no abort
    st = op.solve(torch.ones_like(x), torch.zeros_like(x), rtol=1e-12, maxit=2, raise_on_fail=False)
    assert st["rc"] == -4 and not st["converged"]
    # NaN in b -> detected in the first reduction
    b = torch.ones_like(x)
    b[3] = float("nan")
    st = op.solve(b, torch.zeros_like(x), rtol=1e-8, maxit=100, raise_on_fail=False)
    assert st["rc"] == -5


@pytest.mark.parametrize("p", [1, 2, 3])
def test_element_factors(torch_cuda, hc, oracle, p):
    """Box meshes apply from two factors per element (dt nu_e, sigma_e) and the
    axis constants: every tile variant against the oracle (FP64 1e-13, FP32
    2e-6) and against the quadrature-point factors of the same operator
    (1e-14), partial PEC, several right-hand sides in one launch bitwise equal
    to single applies, the curl-only term weights; trilinear meshes keep the
    quadrature-point factors."""
    torch = torch_cuda
    from hephaestus_amd import hcurl
    n = {1: (11, 9, 6), 2: (9, 7, 5), 3: (9, 5, 7)}[p]
    print(f"case element factors p={p} n={n}", flush=True)
    rng = np.random.default_rng(700 + p)
    pec = 0x19
    m = oracle.Mesh(*n, pec_faces=pec)
    E = m.nelem
    nu, sig = rng.uniform(0.5, 2.0, E), 10.0 ** rng.uniform(-4, 4, E)
    dt = 0.37
    qd = oracle.pa_setup(m, p, dt * nu, sig)
    ess = oracle.ess_mask(m, p)
    N = oracle.ndofs(m, p)
    X = np.stack([oracle.rand_uniform(51 + r, N) for r in range(3)])
    Y = np.stack([oracle.pa_apply(m, p, qd, ess, X[r]) for r in range(3)])
    op = hc.Operator(n, p, nu, sig, pec_faces=pec).setup(dt)
    assert op.factor_mode == 1
    xd = _dev(torch, X)
    for t in hcurl.tile_list(p):
        op.set_tile(64, *t)
        op.set_tile(32, *t)
        op.set_factor_mode(1)
        y = op.apply(xd)
        y32 = op.apply_f32(xd.float())
        k1 = op.apply_curl(xd[0])
        for r in range(3):
            assert _rel(y[r].cpu().numpy(), Y[r]) <= 1e-13, (t, r)
            assert _rel(y32[r].double().cpu().numpy(), Y[r]) <= 2e-6, (t, r)
            assert torch.equal(op.apply(xd[r]), y[r]), (t, r)
        op.set_factor_mode(0)
        assert op.factor_mode == 0
        yq = op.apply(xd)
        kq = op.apply_curl(xd[0])
        assert _rel(y.cpu().numpy(), yq.cpu().numpy()) <= 1e-14, t
        assert _rel(k1.cpu().numpy(), kq.cpu().numpy()) <= 1e-14, t
    op.set_factor_mode(1)
    op.close()
    vert = oracle.Mesh(*n).box_vertices() + 0.04 * rng.standard_normal(((n[0] + 1) * (n[1] + 1) * (n[2] + 1), 3)) / max(n)
    opt = hc.Operator(n, p, nu, sig, vert=vert).setup(dt)
    assert opt.factor_mode == 0
    opt.close()


def test_element_factors_mixed_batch(torch_cuda, hc, oracle):
    """A batch holding a box and a trilinear instance of one order runs both
    from their quadrature-point factors; each solution meets rtol on the
    oracle's operator and agrees with the same instance solved alone."""
    torch = torch_cuda
    p, n = 2, (7, 6, 5)
    print("case element factors mixed batch", flush=True)
    rng = np.random.default_rng(77)
    E = n[0] * n[1] * n[2]
    nu, sig = rng.uniform(0.5, 2.0, E), rng.uniform(0.5, 2.0, E)
    vert = oracle.Mesh(*n).box_vertices() + 0.04 * rng.standard_normal(((n[0] + 1) * (n[1] + 1) * (n[2] + 1), 3)) / max(n)
    ops = [hc.Operator(n, p, nu, sig).setup(1.0), hc.Operator(n, p, nu, sig, vert=vert).setup(1.0)]
    bs = [op.rhs_coil(90 + i) for i, op in enumerate(ops)]
    xs = [torch.zeros_like(b) for b in bs]
    bt = hc.Batch()
    for op, b, x in zip(ops, bs, xs):
        bt.add(op, b, x, rtol=1e-10, maxit=5000)
    _, st = bt.solve(mode=hc.SOLVE_FP64)
    assert all(s["converged"] for s in st), st
    for i, (op, b, x) in enumerate(zip(ops, bs, xs)):
        m = oracle.Mesh(*n, vert=vert if i else None)
        qd = oracle.pa_setup(m, p, nu, sig)
        ess = oracle.ess_mask(m, p)
        bh, xh = b.cpu().numpy(), x.cpu().numpy()
        r = bh - oracle.pa_apply(m, p, qd, ess, xh)
        assert np.linalg.norm(r) <= 1.05e-10 * np.linalg.norm(bh), i
        x1 = torch.zeros_like(b)
        op.solve(b, x1, rtol=1e-10, maxit=5000, mode=hc.SOLVE_FP64)
        assert _rel(x1.cpu().numpy(), xh) <= 1e-8, i
    bt.close()
    for op in ops:
        op.close()
