"""Geometric multigrid preconditioner HC_SOLVE_MG (needs an MI355X).

Transfers, pinned against the oracle: the ND spaces of the nested meshes are
nested (as is ND_1 in ND_p on one mesh) and both levels integrate exactly, so
with coefficients constant on every 2x2x2 block of fine elements (or on the
same mesh) the coarse operator is the Galerkin product P^T A_f P = A_c.  The check applies the GPU prolongation P and restriction
P^T between the oracle's assembled fine and coarse operators (scipy CSR on the
CPU) on random coarse vectors: R (A_f (P u)) = A_c u to 1e-12, for p = 1..3,
PEC on all faces and on a partial set, mesh-halving and order-dropping steps,
on every level pair of a three-level hierarchy.  A wrong 1-D node set, a missing h_f / h_c scaling, a transposed
axis or a wrong PEC treatment breaks it.  The restriction is the transpose of
the prolongation: (R x, y) = (x, P y) for y with zero PEC entries.

Level operators: a level's operator is the box operator of the halved mesh
with the coefficients averaged over the eight children and the same dt
(bitwise equal to one created directly), also after a re-assembly at another
dt (the hierarchy is rebuilt).

Solves: MG-PCG reaches rtol on the oracle's assembled operator and agrees
with a sparse direct solve on two-region coefficients (two and three levels,
an odd mesh through an order-1 level);
a mesh small enough for the coarsest level is solved directly (one
iteration); iteration counts stay flat from 16^3 to 32^3 (order 2); meshes
that cannot be coarsened down to a small level, and FP32 modes, are refused.
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
    from oracle import oracle as O
    H.load()
    O.build()
    return torch, H, O


def _two_region(n, inner, outer):
    c = (np.arange(n) + 0.5) / n
    ins = (c >= 0.25) & (c <= 0.75)
    Z, Y, X = np.meshgrid(ins, ins, ins, indexing="ij")
    return np.where((X & Y & Z).ravel(), inner, outer).astype(np.float64)


def _expand(coarse, nc):
    """Per-element values of an nc^3 mesh onto its 2nc^3 refinement."""
    c = coarse.reshape(nc, nc, nc)
    return np.repeat(np.repeat(np.repeat(c, 2, 0), 2, 1), 2, 2).ravel()


def _average(fine, nf):
    """Mean over the eight children, summed as the hierarchy sums it (children
    in z, y, x order, each weighted 1/8; not exact for three equal values)."""
    f = fine.reshape(nf // 2, 2, nf // 2, 2, nf // 2, 2)
    s = np.zeros((nf // 2,) * 3)
    for dz in range(2):
        for dy in range(2):
            for dx in range(2):
                s = s + 0.125 * f[:, dz, :, dy, :, dx]
    return s.ravel()


def _oracle_A(O, n, p, nu, sig, dt, pec):
    m = O.Mesh(n, pec_faces=pec)
    ess = O.ess_mask(m, p)
    qd = O.pa_setup(m, p, dt * np.broadcast_to(nu, (m.nelem,)).copy(), np.broadcast_to(sig, (m.nelem,)).copy())
    return O.Csr(m, p, qd, ess).to_scipy().tocsr(), ess.astype(bool)


@pytest.mark.parametrize("p,nc,pec,step", [(1, 8, 0x3F, "h"), (2, 4, 0x3F, "h"), (3, 3, 0x3F, "h"), (2, 4, 0x15, "h"),
                                          (1, 8, 0x26, "h"), (3, 5, 0x3F, "p"), (2, 6, 0x19, "p")])
def test_galerkin_transfer_vs_oracle(env, p, nc, pec, step):
    """step h: the fine mesh halves (coefficients constant on 2x2x2 blocks);
    step p: odd or non-aligned coefficients, the order drops to 1 on the same
    mesh (ND_1 inside ND_p, random per-element coefficients)."""
    torch, H, O = env
    rng = np.random.default_rng(10 * p + nc)
    dt = 0.3
    print(f"case mg galerkin p={p} coarse {nc}^3 pec {pec:#x} step {step}", flush=True)
    if step == "h":
        nf, pc = 2 * nc, p
        nu_c = rng.uniform(0.5, 2.0, nc ** 3)
        sig_c = 10.0 ** rng.uniform(-2, 2, nc ** 3)
        nu_f, sig_f = _expand(nu_c, nc), _expand(sig_c, nc)
    else:
        nf, pc = nc, 1
        nu_f = nu_c = rng.uniform(0.5, 2.0, nc ** 3)
        sig_f = sig_c = 10.0 ** rng.uniform(-2, 2, nc ** 3)
    op = H.Operator(nf, p, nu_f, sig_f, precision=H.PREC_FP64, pec_faces=pec).setup(dt)
    nl, ncoarse = op.mg_levels()
    assert nl == 2 and ncoarse == O.ndofs(O.Mesh(nc), pc)
    Af, ess_f = _oracle_A(O, nf, p, nu_f, sig_f, dt, pec)
    Ac, ess_c = _oracle_A(O, nc, pc, nu_c, sig_c, dt, pec)
    for k in range(3):
        u = rng.standard_normal(Ac.shape[0])
        u[ess_c] = 0.0
        Pu = op.mg_transfer(0, True, torch.from_numpy(u).cuda(), torch.empty(Af.shape[0], dtype=torch.float64,
                                                                              device="cuda"))
        Pu_h = Pu.cpu().numpy()
        assert np.all(Pu_h[ess_f] == 0.0)
        w = Af @ Pu_h
        Rw = op.mg_transfer(0, False, torch.from_numpy(w).cuda(), torch.empty(Ac.shape[0], dtype=torch.float64,
                                                                             device="cuda")).cpu().numpy()
        ref = Ac @ u
        err = np.linalg.norm(Rw - ref) / np.linalg.norm(ref)
        print(f"  P^T A_f P u vs A_c u: relative {err:.2e}")
        assert err <= 1e-12
        # the restriction is the transpose of the prolongation
        x = rng.standard_normal(Af.shape[0])
        Rx = op.mg_transfer(0, False, torch.from_numpy(x).cuda(), torch.empty(Ac.shape[0], dtype=torch.float64,
                                                                              device="cuda")).cpu().numpy()
        assert np.all(Rx[ess_c] == 0.0)
        lhs, rhs = float(Rx @ u), float(x @ Pu_h)
        assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(Rx) * np.linalg.norm(u) + 1e-13 * abs(rhs)
    op.close()


def test_three_levels_galerkin_and_level_ops(env):
    """16^3 p2 -> 8^3 -> 4^3: every level pair meets the Galerkin identity
    with the oracle's operators; level operators equal directly created ones,
    also after a re-assembly at another dt."""
    torch, H, O = env
    p, n = 2, 16
    rng = np.random.default_rng(7)
    nu4 = rng.uniform(0.5, 2.0, 4 ** 3)
    sig4 = 10.0 ** rng.uniform(-3, 3, 4 ** 3)
    nu = _expand(_expand(nu4, 4), 8)
    sig = _expand(_expand(sig4, 4), 8)
    print("case mg three levels 16^3 p2", flush=True)
    op = H.Operator(n, p, nu, sig, precision=H.PREC_FP64).setup(0.5)
    assert op.mg_levels() == (3, O.ndofs(O.Mesh(4), p))
    coefs = [(n, nu, sig), (8, _expand(nu4, 4), _expand(sig4, 4)), (4, nu4, sig4)]
    for lvl in range(2):
        nf, nuf, sgf = coefs[lvl]
        nc, nuc, sgc = coefs[lvl + 1]
        Af, _ = _oracle_A(O, nf, p, nuf, sgf, 0.5, 0x3F)
        Ac, ess_c = _oracle_A(O, nc, p, nuc, sgc, 0.5, 0x3F)
        u = rng.standard_normal(Ac.shape[0])
        u[ess_c] = 0.0
        Pu = op.mg_transfer(lvl, True, torch.from_numpy(u).cuda(),
                            torch.empty(Af.shape[0], dtype=torch.float64, device="cuda")).cpu().numpy()
        Rw = op.mg_transfer(lvl, False, torch.from_numpy(Af @ Pu).cuda(),
                            torch.empty(Ac.shape[0], dtype=torch.float64, device="cuda")).cpu().numpy()
        ref = Ac @ u
        assert np.linalg.norm(Rw - ref) <= 1e-12 * np.linalg.norm(ref), lvl
    # coefficients of each level: averages of the level above (as the hierarchy forms them)
    avg = [(nu, sig)]
    for nc in (8, 4):
        avg.append((_average(avg[-1][0], 2 * nc), _average(avg[-1][1], 2 * nc)))
    for dt in (0.5, 0.02):
        if dt != 0.5:
            op.setup(dt)
            assert op.mg_levels()[0] == 3
        for lvl in (1, 2):
            nc = coefs[lvl][0]
            direct = H.Operator(nc, p, avg[lvl][0], avg[lvl][1], precision=H.PREC_FP64).setup(dt)
            x = torch.empty(direct.ndofs, dtype=torch.float64, device="cuda")
            direct.fill_uniform(3 + lvl, x)
            lo = op.mg_level(lvl)
            assert lo.ndofs == direct.ndofs
            direct.set_tile(64, *lo.get_tile(64))  # (levels run their own autotuned tile variant)
            assert torch.equal(lo.apply(x), direct.apply(x)), (dt, lvl)
            direct.close()
    op.close()


@pytest.mark.parametrize("n,p,levels,direct", [(8, 2, 2, True), (16, 1, 2, True), (12, 2, 2, False),
                                               (16, 2, 3, False), (9, 3, 2, False)])
def test_mg_solve_vs_direct(env, n, p, levels, direct):
    import scipy.sparse.linalg as spla
    torch, H, O = env
    dt = 0.05
    nu = _two_region(n, 0.5, 1.0)
    sig = _two_region(n, 1e2, 1e-2)
    print(f"case mg solve vs direct {n}^3 p={p}", flush=True)
    op = H.Operator(n, p, nu, sig, precision=H.PREC_FP64).setup(dt)
    assert op.mg_levels()[0] == levels
    b = op.rhs_coil(11)
    x = torch.zeros_like(b)
    st = op.solve(b, x, rtol=1e-10, maxit=500, mode=H.SOLVE_FP64 | H.SOLVE_MG)
    A, ess = _oracle_A(O, n, p, nu, sig, dt, 0x3F)
    bh, xh = b.cpu().numpy(), x.cpu().numpy()
    res = np.linalg.norm(bh - A @ xh) / np.linalg.norm(bh)
    err = 0.0
    if direct:  # (the 45000-DOF case is checked through the oracle residual alone)
        xd = spla.spsolve(A.tocsc(), bh)
        err = np.linalg.norm(xh - xd) / np.linalg.norm(xd)
    print(f"  {st['iterations']} iterations, oracle residual {res:.2e}, vs direct {err:.2e}")
    assert st["converged"] and res <= 1e-10 * 1.05
    assert err <= 1e-7
    assert st["iterations"] <= 40
    op.close()


def test_single_level_is_direct(env):
    torch, H, O = env
    n, p = 6, 1
    print("case mg single level", flush=True)
    op = H.Operator(n, p, 1.0, _two_region(n, 1e3, 1e-3), precision=H.PREC_FP64).setup(1.0)
    assert op.mg_levels() == (1, op.ndofs)
    b = op.rhs_coil(3)
    x = torch.zeros_like(b)
    st = op.solve(b, x, rtol=1e-10, maxit=10, mode=H.SOLVE_FP64 | H.SOLVE_MG)
    assert st["converged"] and st["iterations"] <= 1
    A, _ = _oracle_A(O, n, p, 1.0, _two_region(n, 1e3, 1e-3), 1.0, 0x3F)
    bh = b.cpu().numpy()
    assert np.linalg.norm(bh - A @ x.cpu().numpy()) <= 1e-10 * np.linalg.norm(bh)
    op.close()


def test_mg_iterations_flat_in_h(env):
    """Order 2, nu = sigma = 1: MG-PCG iterations at 16^3 and 32^3 stay
    within 3 of each other and below a tenth of the Jacobi+gradient ones."""
    torch, H, O = env
    its = {}
    for n in (16, 32):
        print(f"case mg h-independence {n}^3", flush=True)
        op = H.Operator(n, 2, 1.0, 1.0, precision=H.PREC_FP64).setup(1.0)
        b = torch.empty(op.ndofs, dtype=torch.float64, device="cuda")
        op.fill_uniform(2, b)
        for name, mode in (("mg", H.SOLVE_FP64 | H.SOLVE_MG), ("aux", H.SOLVE_FP64 | H.SOLVE_AUX)):
            x = torch.zeros_like(b)
            st = op.solve(b, x, rtol=1e-8, maxit=5000, mode=mode)
            assert st["converged"]
            its[(name, n)] = st["iterations"]
        op.close()
    print(f"  iterations {its}")
    assert abs(its[("mg", 32)] - its[("mg", 16)]) <= 3
    assert its[("mg", 32)] * 10 <= its[("aux", 32)]


def test_mg_refusals(env):
    torch, H, O = env
    print("case mg refusals", flush=True)
    op = H.Operator(17, 1, 1.0, 1.0, precision=H.PREC_FP64).setup(1.0)  # odd, order 1, 16524 DOFs
    b = op.rhs_coil(1)
    x = torch.zeros_like(b)
    with pytest.raises(H.HcError) as e:
        op.solve(b, x, rtol=1e-8, maxit=10, mode=H.SOLVE_FP64 | H.SOLVE_MG)
    assert e.value.code == -1 and "coarsest level" in str(e.value)
    op.close()
    op = H.Operator(8, 1, 1.0, 1.0).setup(1.0)
    b = op.rhs_coil(1)
    with pytest.raises(H.HcError):
        op.solve(b, torch.zeros_like(b), rtol=1e-8, maxit=10, mode=H.SOLVE_FP32_IR | H.SOLVE_MG)
    with pytest.raises(H.HcError):
        op.mg_transfer(0, True, b, b)  # no hierarchy yet
    op.close()


def test_mg_graph_replay_matches_eager(env):
    """The finest-level V-cycle replayed from a captured hipGraph gives the
    bitwise same solutions and iteration counts as eager launches, over
    repeated solves of two operators interleaved with eager solves and other
    device work (the graph is captured on the second V-cycle of a hierarchy
    and replayed from then on)."""
    torch, H, O = env
    from hephaestus_amd import hcurl
    print("case mg graph replay vs eager", flush=True)
    specs = [(16, 2, _two_region(16, 1e2, 1e-2)), (12, 3, _two_region(12, 10.0, 0.1))]
    ops = [H.Operator(n, p, 1.0, sig, precision=H.PREC_FP64).setup(0.2) for n, p, sig in specs]
    bs = [op.rhs_coil(31 + i) for i, op in enumerate(ops)]
    ref = []
    try:
        hcurl.set_mg_graphs(False)
        for op, b in zip(ops, bs):
            x = torch.zeros_like(b)
            st = op.solve(b, x, rtol=1e-10, maxit=500, mode=H.SOLVE_FP64 | H.SOLVE_MG)
            ref.append((x, st["iterations"]))
        for rep in range(3):
            for i, (op, b) in enumerate(zip(ops, bs)):
                hcurl.set_mg_graphs(True)
                x = torch.zeros_like(b)
                st = op.solve(b, x, rtol=1e-10, maxit=500, mode=H.SOLVE_FP64 | H.SOLVE_MG)
                assert st["iterations"] == ref[i][1], (rep, i)
                assert torch.equal(x, ref[i][0]), (rep, i)
                # eager work between replays
                y = op.apply(x)
                assert torch.isfinite(y).all()
                hcurl.set_mg_graphs(False)
                x2 = torch.zeros_like(b)
                op.solve(b, x2, rtol=1e-10, maxit=500, mode=H.SOLVE_FP64 | H.SOLVE_MG)
                assert torch.equal(x2, ref[i][0]), (rep, i)
    finally:
        hcurl.set_mg_graphs(False)
    for op in ops:
        op.close()


def test_mg_multi_rhs(env):
    """Several right-hand sides through one V-cycle per iteration: every
    column meets rtol on the oracle's operator, equals the one-RHS multigrid
    solve to 1e-12 with the same iteration count, and is bitwise independent
    of its companions (the same column solved beside other right-hand sides
    in another order)."""
    torch, H, O = env
    n, p = 16, 2
    print("case mg multi-rhs 16^3 p2", flush=True)
    sig = _two_region(n, 1e2, 1e-2)
    op = H.Operator(n, p, 1.0, sig, precision=H.PREC_FP64).setup(0.2)
    B = torch.stack([op.rhs_coil(40 + r) for r in range(5)])
    X = torch.zeros_like(B)
    st = op.solve(B, X, rtol=1e-10, maxit=200, mode=H.SOLVE_FP64 | H.SOLVE_MG)
    assert st["converged"]
    A, _ = _oracle_A(O, n, p, 1.0, sig, 0.2, 0x3F)
    Bh, Xh = B.cpu().numpy(), X.cpu().numpy()
    for r in range(5):
        assert np.linalg.norm(Bh[r] - A @ Xh[r]) <= 1e-10 * 1.05 * np.linalg.norm(Bh[r]), r
        x1 = torch.zeros_like(B[r])
        s1 = op.solve(B[r].contiguous(), x1, rtol=1e-10, maxit=200, mode=H.SOLVE_FP64 | H.SOLVE_MG)
        assert np.linalg.norm(x1.cpu().numpy() - Xh[r]) <= 1e-12 * np.linalg.norm(Xh[r]), r
        assert s1["iterations"] <= st["iterations"]
    perm = [3, 0, 4]
    X2 = torch.zeros_like(B[perm])
    op.solve(B[perm].contiguous(), X2, rtol=1e-10, maxit=200, mode=H.SOLVE_FP64 | H.SOLVE_MG)
    for j, r in enumerate(perm):
        assert torch.equal(X2[j], X[r]), (j, r)
    op.close()
