"""Continuous-batching server of backward-Euler steps (needs an MI355X).

Each request computes u <- u + dt k with (dt K + M) k = b_coil - K u.  Checks:
one request against the oracle's step: the served k meets rtol 1e-10 on the
oracle's assembled operator and u agrees with a sparse direct solve to relative
L2 <= 1e-6 (rtol times the conditioning of this system), also with a server
tile variant that differs from the operator's own (autotuned) one and with the
multigrid workers (HC_SOLVE_MG); every
request of a mix of uneven instances (orders 1-3, sizes 5-9) is bitwise
identical to the same request served alone (mix invariance), both when all
arrive at once and when they arrive one after another (scheduled and Poisson
arrival times, at most three admitted at once) so that later requests join a
running batch; an HBM budget that holds only two requests at a time bounds the
admitted set and still completes every request; requests that can never fit
are refused.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


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


def _two_region(n, inner, outer):
    c = (np.arange(n) + 0.5) / n
    ins = (c >= 0.25) & (c <= 0.75)
    Z, Y, X = np.meshgrid(ins, ins, ins, indexing="ij")
    return np.where((X & Y & Z).ravel(), inner, outer).astype(np.float64)


MIX = [(6, 1, 10.0, 0.1), (5, 2, 1e2, 1e-2), (7, 3, 3.0, 1.0), (9, 1, 1.0, 1.0), (6, 2, 50.0, 0.5),
       (5, 3, 1e2, 1e-1), (8, 2, 2.0, 2.0)]
DT = 0.05


def _setup(torch, hc, specs):
    ops, bs, us0 = [], [], []
    for i, (n, p, s_in, s_out) in enumerate(specs):
        op = hc.Operator(n, p, 1.0, _two_region(n, s_in, s_out)).setup(DT)
        b = op.rhs_coil(40 + i)
        u = torch.empty_like(b)
        op.fill_uniform(70 + i, u)
        u.mul_(0.01)
        ops.append(op)
        bs.append(b)
        us0.append(u)
    return ops, bs, us0


def _serve(hc, torch, ops, bs, us, mode, budget=1 << 30, max_active=0, arrivals=None):
    srv = hc.Server(budget, mode=mode, rtol=1e-9, maxit=20000, max_active=max_active, round_iters=8)
    for i, (op, b, u) in enumerate(zip(ops, bs, us)):
        srv.submit(op, b, u, DT, tag=i, arrival_ms=-1.0 if arrivals is None else float(arrivals[i]))
    res = {}
    while len(res) < len(ops):
        for r in srv.poll(wait_ms=2000):
            res[r["tag"]] = r
    info = srv.info()
    srv.close()
    torch.cuda.synchronize()
    return res, info


@pytest.mark.parametrize("mode", ["fp32_ir", "fp64", "fp32_ir_aux", "fp64_mg"])
def test_server_mix_invariance(torch_cuda, hc, mode):
    torch = torch_cuda
    md = hc.SOLVE_FP32_IR if mode.startswith("fp32_ir") else hc.SOLVE_FP64
    if mode.endswith("_aux"):
        md |= hc.SOLVE_AUX
    if mode.endswith("_mg"):
        md |= hc.SOLVE_MG
    print(f"case server mix {mode} ({len(MIX)} requests)", flush=True)
    ops, bs, us0 = _setup(torch, hc, MIX)
    mixed = [u.clone() for u in us0]
    res, info = _serve(hc, torch, ops, bs, mixed, md)
    assert info["completed"] == len(MIX)
    assert all(r["converged"] and r["rel_res"] <= 1e-9 * (1 + 1e-6) for r in res.values())
    assert info["peak_active"] >= 2
    for i in range(len(MIX)):
        alone = us0[i].clone()
        r1, _ = _serve(hc, torch, [ops[i]], [bs[i]], [alone], md)
        assert r1[0]["iterations"] == res[i]["iterations"]
        assert torch.equal(alone, mixed[i]), f"request {i} differs alone vs in the mix"
        assert not torch.equal(alone, us0[i])


@pytest.mark.parametrize("arrivals,mode", [("staggered", "fp32_ir_aux"), ("poisson", "fp32_ir_aux"),
                                           ("poisson", "fp64_mg")])
def test_server_late_arrivals_join_bitwise(torch_cuda, hc, arrivals, mode):
    """Requests arrive while others are on the GPU (arrival times 4 ms apart, or
    exponential inter-arrival times of mean 4 ms from a seed; at most three
    admitted at once, three workers with multigrid): later requests are admitted
    as slots free up, and each result is bitwise the one of the same request
    served alone, with the same iteration count."""
    torch = torch_cuda
    md = hc.SOLVE_FP32_IR | hc.SOLVE_AUX if mode == "fp32_ir_aux" else hc.SOLVE_FP64 | hc.SOLVE_MG
    print(f"case server late arrivals {arrivals}", flush=True)
    ops, bs, us0 = _setup(torch, hc, MIX)
    if arrivals == "staggered":
        at = [4.0 * i for i in range(len(MIX))]
    else:
        at = np.cumsum(np.random.default_rng(5).exponential(4.0, len(MIX))).tolist()
    mixed = [u.clone() for u in us0]
    res, info = _serve(hc, torch, ops, bs, mixed, md, max_active=3, arrivals=at)
    assert info["completed"] == len(MIX)
    assert all(r["converged"] for r in res.values())
    # some request was admitted while another one was running
    span = {t: (r["arrival_ms"] + r["queue_ms"], r["arrival_ms"] + r["latency_ms"]) for t, r in res.items()}
    joined = [(i, j) for i in span for j in span if i != j and span[i][0] < span[j][0] < span[i][1]]
    assert joined, span
    for i in range(len(MIX)):
        assert res[i]["arrival_ms"] >= at[i] - 1e-6
        alone = us0[i].clone()
        r1, _ = _serve(hc, torch, [ops[i]], [bs[i]], [alone], md)
        assert r1[0]["iterations"] == res[i]["iterations"], i
        assert torch.equal(alone, mixed[i]), f"request {i} differs alone vs joining a running batch"


def test_server_budget_admission(torch_cuda, hc):
    torch = torch_cuda
    print("case server budget", flush=True)
    # five requests of equal size (one resident operator shared by all)
    op1, b1, _ = _setup(torch, hc, [MIX[2]])
    ops = op1 * 5
    bs = [op1[0].rhs_coil(90 + i) for i in range(5)]
    us = [torch.zeros_like(bs[0]) for _ in range(5)]
    need = hc.Server.request_bytes(op1[0].ndofs)
    res, info = _serve(hc, torch, ops, bs, us, hc.SOLVE_FP32_IR, budget=2 * need + 4096)
    assert info["completed"] == 5
    assert info["peak_active"] <= 2
    assert info["peak_bytes"] <= 2 * need + 4096
    assert max(r["queue_ms"] for r in res.values()) > 0.0
    # a request larger than the whole budget is refused at submission
    srv = hc.Server(1024, mode=hc.SOLVE_FP32_IR)
    with pytest.raises(hc.HcError):
        srv.submit(ops[0], bs[0], us[0], DT, tag=0)
    srv.close()


@pytest.mark.parametrize("tiles", ["operator", "server_tile", "multigrid"])
def test_server_step_vs_oracle(torch_cuda, hc, tiles):
    import scipy.sparse.linalg as spla
    from hephaestus_amd import hcurl
    from oracle import oracle as O
    torch = torch_cuda
    n, p = 6, 2
    print(f"case server step vs oracle ({tiles})", flush=True)
    nu = _two_region(n, 0.5, 1.0)
    sig = _two_region(n, 1e2, 1e-2)
    op = hc.Operator(n, p, nu, sig).setup(DT)
    b = op.rhs_coil(5)
    u = torch.empty_like(b)
    op.fill_uniform(6, u)
    u.mul_(0.1)
    u0 = u.cpu().numpy()
    srv = hc.Server(1 << 28, mode=hc.SOLVE_FP64 | hc.SOLVE_MG if tiles == "multigrid" else hc.SOLVE_FP32_IR,
                    rtol=1e-10, maxit=20000)
    if tiles == "server_tile":
        # the operator keeps its autotuned variant; the server runs another one
        op.autotune(reps=2)
        own = op.get_tile(64)
        other = [t for t in hcurl.tile_list(p) if t != own and t[3] < 8][-1]
        srv.set_tile(p, 64, *other)
        srv.set_tile(p, 32, *other)
        print(f"operator tile {own}, server tile {other}", flush=True)
    srv.submit(op, b, u, DT, tag=7)
    r = srv.poll(wait_ms=30000)
    srv.close()
    assert len(r) == 1 and r[0]["converged"] and r[0]["tag"] == 7
    O.build()
    m = O.Mesh(n)
    ess = O.ess_mask(m, p)
    u0[ess.astype(bool)] = 0.0  # PEC entries do not enter K u
    qK = O.pa_setup(m, p, nu, np.zeros(m.nelem))
    qA = O.pa_setup(m, p, DT * nu, sig)
    A = O.Csr(m, p, qA, ess).to_scipy().tocsc()
    f = b.cpu().numpy() - O.pa_apply(m, p, qK, ess, u0)
    f[ess.astype(bool)] = 0.0
    u_ref = u0 + DT * spla.spsolve(A, f)
    ug = u.cpu().numpy()
    free = ~ess.astype(bool)
    err = np.linalg.norm(ug[free] - u_ref[free]) / np.linalg.norm(u_ref[free])
    # the served k = (u1 - u0) / dt meets rtol on the oracle's own operator
    k = (ug - u0) / DT
    res = np.linalg.norm(f - A @ k) / np.linalg.norm(f)
    print(f"server step vs oracle: relative L2 {err:.3e}, oracle residual of k {res:.3e}")
    assert res <= 1e-10 * 1.05
    assert err <= 1e-6


MIX_MG = [(16, 2, 1e2, 1e-2), (12, 3, 10.0, 0.1), (17, 1, 1.0, 1.0), (9, 3, 50.0, 0.5), (20, 1, 2.0, 2.0),
          (16, 2, 1e-3, 1e3)]


def test_server_multigrid_multilevel_mix(torch_cuda, hc):
    """Multigrid workers on operators with real hierarchies (halved meshes, an
    order-1 level, a Chebyshev-smoothed coarsest level at 12^3 p3 and 9^3 p3, a 17^3
    order-1 mesh without a hierarchy served by the gradient-space PCG): every
    request converges, several run at once, and each is bitwise the same
    request served alone."""
    torch = torch_cuda
    md = hc.SOLVE_FP64 | hc.SOLVE_MG
    print("case server multigrid multilevel mix", flush=True)
    ops, bs, us0 = _setup(torch, hc, MIX_MG)
    mixed = [u.clone() for u in us0]
    res, info = _serve(hc, torch, ops, bs, mixed, md, max_active=3)
    assert info["completed"] == len(MIX_MG)
    assert all(r["converged"] and r["rel_res"] <= 1e-9 * (1 + 1e-6) for r in res.values()), res
    assert info["peak_active"] >= 2
    for i in range(len(MIX_MG)):
        alone = us0[i].clone()
        r1, _ = _serve(hc, torch, [ops[i]], [bs[i]], [alone], md)
        assert r1[0]["iterations"] == res[i]["iterations"]
        assert torch.equal(alone, mixed[i]), f"request {i} differs alone vs in the mix"
