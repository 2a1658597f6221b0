"""The uneven batch and the serving stream at their full size (needs an
MI355X): 64 cubes, edge U{8..40} elements, order U{1,2,3}, sigma/dt of the
inner region [0.25, 0.75]^3 and of the outer region each 10^U[-6,4] (instance
i draws from splitmix64 with seed 4 + i), coil right-hand sides (seed
1000 + i), 53.6 M DOFs in total; solved as one multigrid batch, and served as
128 backward-Euler step requests drawn from the 64 operators with Poisson
arrivals of mean 200 req/s (seed 5), by the multigrid server and by the
FP32 PCG + FP64 refinement server with the gradient-space correction.

Bars: every instance converges to rtol 1e-8 in its true residual, recomputed
here by a separate FP64 apply (the apply itself is pinned against the CPU
oracle by the parity tests); a fixed sample of instances solved alone through
the batch interface gives bitwise the batch's solution and iteration count,
and a fixed sample of requests served alone gives bitwise the result they had
in the stream.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-8
COUNT = 64


@pytest.fixture(scope="module")
def env():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import hephaestus_amd as H
    H.load()
    return torch, H


def _splitmix_u01(seed, i):
    z = (seed + (i + 1) * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return ((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0)


def _specs():
    out = []
    for i in range(COUNT):
        s = 4 + i
        n = 8 + int(_splitmix_u01(s, 0) * 33)
        p = 1 + int(_splitmix_u01(s, 1) * 3)
        s_in = 10.0 ** (-6 + 10 * _splitmix_u01(s, 2))
        s_out = 10.0 ** (-6 + 10 * _splitmix_u01(s, 3))
        out.append((n, p, s_in, s_out))
    return out


def _two_region(n, inner, outer):
    c = (np.arange(n) + 0.5) / n
    ins = (c >= 0.25) & (c <= 0.75)
    Z, Y, X = np.meshgrid(ins, ins, ins, indexing="ij")
    return np.where((X & Y & Z).ravel(), inner, outer).astype(np.float64)


@pytest.fixture(scope="module")
def mix(env):
    torch, H = env
    ops, bs = [], []
    for i, (n, p, s_in, s_out) in enumerate(_specs()):
        op = H.Operator(n, p, 1.0, _two_region(n, s_in, s_out)).setup(1.0)
        ops.append(op)
        bs.append(op.rhs_coil(1000 + i))
    torch.cuda.synchronize()
    return ops, bs


def test_uneven_64_multigrid_batch(env, mix):
    torch, H = env
    specs = _specs()
    sig = [s for (_, _, a, b) in specs for s in (a, b)]
    # the mix spans the stated range at both ends and every order
    assert min(sig) < 1e-5 and max(sig) > 1e3
    assert {p for (_, p, _, _) in specs} == {1, 2, 3}
    assert min(n for (n, _, _, _) in specs) <= 10 and max(n for (n, _, _, _) in specs) >= 38
    ops, bs = mix
    print(f"case uneven 64: {sum(o.ndofs for o in ops)} DOFs, multigrid batch", flush=True)
    mode = H.SOLVE_FP64 | H.SOLVE_MG
    xs = [torch.zeros_like(b) for b in bs]
    bt = H.Batch()
    for op, b, x in zip(ops, bs, xs):
        bt.add(op, b, x, rtol=RTOL, maxit=5000)
    _, st = bt.solve(mode, raise_on_fail=False)
    bt.close()
    for i, (op, b, x, s) in enumerate(zip(ops, bs, xs, st)):
        assert s["converged"], (i, specs[i], s)
        res = float(torch.linalg.norm(b - op.apply(x)) / torch.linalg.norm(b))
        assert res <= RTOL * (1 + 1e-6), (i, specs[i], res)
    print("case uneven 64: instances alone vs in the batch", flush=True)
    for i in range(0, COUNT, 9):
        x1 = torch.zeros_like(bs[i])
        b1 = H.Batch()
        b1.add(ops[i], bs[i], x1, rtol=RTOL, maxit=5000)
        _, s1 = b1.solve(mode, raise_on_fail=False)
        b1.close()
        assert s1[0]["iterations"] == st[i]["iterations"], (i, s1[0]["iterations"], st[i]["iterations"])
        assert torch.equal(x1, xs[i]), i


def _serve(H, torch, reqs, arrivals, mode):
    srv = H.Server(16 << 30, mode=mode, rtol=RTOL, maxit=20000, round_iters=16)
    for i, (op, b, u) in enumerate(reqs):
        srv.submit(op, b, u, 1.0, tag=i, arrival_ms=float(arrivals[i]))
    res = {}
    while len(res) < len(reqs):
        for r in srv.poll(wait_ms=2000):
            res[r["tag"]] = r
    info = srv.info()
    srv.close()
    torch.cuda.synchronize()
    return res, info


@pytest.mark.parametrize("mode", ["fp64_mg", "fp32_ir_aux"])
def test_serving_stream_128(env, mix, mode):
    torch, H = env
    ops, bs = mix
    md = H.SOLVE_FP64 | H.SOLVE_MG if mode == "fp64_mg" else H.SOLVE_FP32_IR | H.SOLVE_AUX
    nreq = 128
    rng = np.random.default_rng(5)
    arrivals = np.cumsum(rng.exponential(1.0 / 200.0, nreq)) * 1000.0
    pick = rng.integers(0, len(ops), nreq)
    us = [torch.zeros_like(bs[k]) for k in pick]
    print(f"case serving stream {mode}: {nreq} requests, Poisson 200 req/s", flush=True)
    res, info = _serve(H, torch, [(ops[k], bs[k], u) for k, u in zip(pick, us)], arrivals, md)
    assert info["completed"] == nreq
    for i in range(nreq):
        r = res[i]
        assert r["converged"] and r["arrival_ms"] >= arrivals[i] - 1e-6, (i, r)
        # u_n = 0, dt = 1: the served state solves A u = b
        op, b = ops[pick[i]], bs[pick[i]]
        rel = float(torch.linalg.norm(b - op.apply(us[i])) / torch.linalg.norm(b))
        assert rel <= RTOL * (1 + 1e-6), (i, rel)
    # requests overlapped on the device
    span = {t: (r["arrival_ms"] + r["queue_ms"], r["arrival_ms"] + r["latency_ms"]) for t, r in res.items()}
    assert any(span[i][0] < span[j][0] < span[i][1] for i in span for j in span if i != j)
    print("case serving stream: requests alone vs in the stream", flush=True)
    for i in range(0, nreq, 16):
        u1 = torch.zeros_like(bs[pick[i]])
        r1, _ = _serve(H, torch, [(ops[pick[i]], bs[pick[i]], u1)], [0.0], md)
        assert r1[0]["iterations"] == res[i]["iterations"], i
        assert torch.equal(u1, us[i]), f"request {i} differs alone vs in the stream"
