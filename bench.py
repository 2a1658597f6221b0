#!/usr/bin/env python3
"""Benchmark of the H(curl) curl-curl + sigma-mass operator path on one MI355X.

Workload (the single-GPU configuration the headline metric is quoted on):
48^3 hex unit cube, Nedelec order 3 (9,082,800 DOFs), nu = 1, two-region
sigma/dt (1e4 on elements with centroid in [0.25,0.75]^3, 1e-6 elsewhere), PEC
on all faces.  A "step" is one FP64 partial-assembly operator apply y = A x over
the whole mesh, with x, y and the quadrature data resident in HBM.

Printed: one JSON line with value = FP64 apply GDOF/s, a roofline object for the
apply (algorithmic bytes B = 8*2N + 8*12*E*Q^3 per launch over the measured
launch time), the CPU baseline (the oracle's OpenMP PA apply on the host cores,
bounded sample) and the solve figures, all to a true relative residual of 1e-8:
the 48^3 order-3 solve, 32 right-hand sides on 32^3 order 2, 64 uneven
operators (edge 8..40, order 1..3, random two-region sigma) batched vs one at a
time, and a Poisson stream of backward-Euler step requests over those 64
operators (p50/p99); each with plain Jacobi and with the gradient-space
correction of the preconditioner, FP64 vs FP32 apply + FP64 refinement.
The headline apply reads the quadrature-point factors; the element-factor
apply of box meshes (the operators' default, used by the solves) is reported
beside it under apply_element_factors.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD published
FP64_PEAK_TFLOPS = 78.6

METRIC = "ND curl-curl+mass apply GDOF/s (%HBM peak); solves/s + p99 at 1e-8 residual"


def two_region(n, inner=1e4, outer=1e-6):
    c = (np.arange(n) + 0.5) / n
    ins = (c >= 0.25) & (c <= 0.75)
    Z, Y, X = np.meshgrid(ins, ins, ins, indexing="ij")
    return np.where((X & Y & Z).ravel(), inner, outer).astype(np.float64)


def apply_bytes(N, E, p, s_vec=8, s_q=8, nrhs=1):
    Q = p + 1
    return s_vec * 2 * N * nrhs + s_q * 12 * E * Q ** 3


def measured_traffic(n, p):
    """Bytes per launch between the L2 caches and the memory fabric of the FP64
    tile kernel, from the committed rocprofv3 counter summary of this workload
    (TCC_EA0_RDREQ x 128 B + TCC_EA0_WRREQ_64B x 64 B + other write requests x
    32 B, summed over every L2 channel; the same sums give a 1 GiB device copy
    its 2 GiB exactly), or None when there is none."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"k_apply_tile_fp64_{n}_p{p}_hbm_counters.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(float(line.split(":", 1)[1].split()[0]))
    return None


def time_applies(torch, fn, steps, warmup):
    st = torch.cuda.current_stream()
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(steps):
        fn()
    e1.record(st)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps  # ms per apply


def cpu_baseline(n, p, sig, budget_s):
    """The oracle's OpenMP FP64 PA apply on this host (bounded sample)."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import oracle as O
    O.build()
    m = O.Mesh(n)
    qd = O.pa_setup(m, p, 1.0, sig)
    ess = O.ess_mask(m, p)
    x = O.rand_uniform(2, O.ndofs(m, p))
    O.pa_apply(m, p, qd, ess, x)  # warm-up
    times = []
    t_start = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        O.pa_apply(m, p, qd, ess, x)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_start > budget_s or len(times) >= 50:
            break
    t = float(np.median(times))
    N = O.ndofs(m, p)
    return {"value": N / t / 1e9, "unit": "GDOF/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"{len(times)} FP64 PA applies of the same 48^3 p=3 operator (median "
                      f"{t * 1e3:.1f} ms each, oracle/hc_oracle.c, OpenMP)",
            "ms_per_apply": t * 1e3}


def cpu_csr_path(torch, H, budget_s):
    """The reference's own per-iteration kernel on the host, restated: an
    assembled CSR matrix (full assembly with PEC elimination, oracle/hc_oracle.c,
    OpenMP) -- its SpMV at 32^3 order 2 (median of a bounded sample) beside the
    GPU FP64 apply of the same operator, and a CSR Jacobi-PCG solve at 16^3
    order 1 (sigma/dt = 1, nu = 1, uniform RHS seed 1, rtol 1e-10) and the same
    solve with the oracle's PA apply beside the GPU FP64 solve of the problem."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import oracle as O
    O.build()
    out = {"cores": O.num_threads(), "kind": "port"}
    n, p = 32, 2
    sig = two_region(n)
    m = O.Mesh(n)
    qd = O.pa_setup(m, p, 1.0, sig)
    ess = O.ess_mask(m, p)
    t0 = time.perf_counter()
    A = O.Csr(m, p, qd, ess)
    t_asm = time.perf_counter() - t0
    x = O.rand_uniform(2, A.n)
    A.apply(x)
    times = []
    t_start = time.perf_counter()
    while time.perf_counter() - t_start < budget_s and len(times) < 20:
        t0 = time.perf_counter()
        A.apply(x)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    op = H.Operator(n, p, 1.0, sig).setup(1.0)
    op.autotune(reps=3)
    xd = torch.empty(op.ndofs, dtype=torch.float64, device="cuda")
    op.fill_uniform(2, xd)
    yd = torch.empty_like(xd)
    ms = time_applies(torch, lambda: op.apply(xd, yd), 30, 3)
    out["spmv_32_p2"] = {"nnz": A.nnz, "assembly_s": round(t_asm, 2), "cpu_ms": round(t * 1e3, 2),
                         "cpu_GDOF_s": round(A.n / t / 1e9, 4), "gpu_apply_ms": round(ms, 4),
                         "gpu_GDOF_s": round(op.ndofs / ms / 1e6, 3), "sample": f"{len(times)} SpMVs"}
    del A, op
    n, p = 16, 1
    m = O.Mesh(n)
    qd = O.pa_setup(m, p, 1.0, 1.0)
    ess = O.ess_mask(m, p)
    A = O.Csr(m, p, qd, ess)
    b = O.rhs_uniform(m, p, 1, ess)
    d = O.pa_diagonal(m, p, qd, ess)
    t0 = time.perf_counter()
    _, st = A.pcg(d, b, 1e-10, 10000)
    t_cpu = time.perf_counter() - t0
    op = H.Operator(n, p, 1.0, 1.0).setup(1.0)
    bd = torch.from_numpy(b).cuda()
    xd = torch.zeros_like(bd)
    sg = op.solve(bd, xd, rtol=1e-10, maxit=10000, mode=H.SOLVE_FP64)
    # the same solve with the oracle's PA apply in the CG loop (MFEM-PA-equivalent path)
    t0 = time.perf_counter()
    _, st_pa = O.pcg_pa(m, p, qd, ess, d, b, 1e-10, 10000)
    t_pa = time.perf_counter() - t0
    out["pcg_16_p1"] = {"cpu_iterations": int(st["iterations"]), "cpu_ms": round(t_cpu * 1e3, 2),
                        "cpu_pa_iterations": int(st_pa["iterations"]), "cpu_pa_ms": round(t_pa * 1e3, 2),
                        "gpu_iterations": sg["iterations"], "gpu_ms": round(sg["solve_ms"], 2)}
    print(f"[bench] CPU CSR path {out}", file=sys.stderr, flush=True)
    return out


def rel_err(torch, x, x_ref):
    """Relative L2 difference to an FP64 reference solution (None without one)."""
    if x_ref is None:
        return None
    return float(torch.linalg.norm(x - x_ref) / torch.linalg.norm(x_ref))


def max_err(errs):
    errs = list(errs)
    return None if any(e is None for e in errs) else max(errs)


def mg_ok(op):
    """True when the operator's mesh halves down to a multigrid coarsest level."""
    try:
        op.mg_levels()
        return True
    except Exception:
        return False


def fp64_reference(H, op, b, args):
    """FP64 reference solution of A x = b to a true relative residual of 1e-11:
    multigrid PCG where the mesh coarsens, else PCG with the gradient-space
    correction (the operator itself is pinned against the CPU oracle by the
    tests)."""
    import torch
    x = torch.zeros_like(b)
    mode = H.SOLVE_FP64 | (H.SOLVE_MG if mg_ok(op) else H.SOLVE_AUX)
    st = op.solve(b, x, rtol=1e-11, maxit=args.solve_maxit, mode=mode, raise_on_fail=False)
    return x if st["converged"] else None


def ulp_floor(torch, op, b, x, draws=2):
    """Relative residual of x with every entry moved by up to one unit in the
    last place (x (1 + u 2^-53), u uniform in [-1, 1)): no FP64 vector near x
    has a smaller residual than about this, so a tolerance below it cannot be
    met by any FP64 solution."""
    g = torch.Generator(device="cuda").manual_seed(17)
    out = []
    for _ in range(draws):
        u = torch.rand(x.shape, generator=g, device="cuda", dtype=torch.float64) * 2.0 - 1.0
        xp = x * (1.0 + u * 2.0 ** -53)
        out.append(float(torch.linalg.norm(b - op.apply(xp)) / torch.linalg.norm(b)))
    return min(out)


def solve_entry(st, nsolves=1):
    return {"iterations": st["iterations"], "outer": st["outer_iterations"], "converged": bool(st["converged"]),
            "true_rel_res": st["rel_res"], "solve_ms": round(st["solve_ms"], 3),
            "solves_per_s": round(1000.0 * nsolves / st["solve_ms"], 4) if st["solve_ms"] > 0 else None}


def solves_single_48(torch, H, op, x, args):
    """48^3 p=3 two-region: FP64 Jacobi-PCG vs FP32 PCG + FP64 refinement to a
    true relative residual of 1e-8, each also with the gradient-space
    correction of the preconditioner (_aux), and FP64 PCG with the multigrid
    V-cycle (fp64_mg); with a coil current right-hand side (divergence-free,
    seed 2) and with a uniform random right-hand side (seed 2, capped at
    --c2-uniform-maxit iterations).  The uniform right-hand side puts a
    gradient field of size |b| / (sigma h^3) ~ 1e10 into the sigma = 1e-6
    region: its FP64 solution has a residual floor far above 1e-8 (the ulp
    floor below, the residual of the best solution with every entry moved by
    one unit in the last place), which every mode reaches and none can pass."""
    sol = {"rtol": 1e-8}
    for rhs, maxit in (("coil", args.solve_maxit), ("uniform", args.c2_uniform_maxit)):
        b = torch.empty_like(x)
        if rhs == "coil":
            op.rhs_coil(2, b)
        else:
            op.fill_uniform(2, b)
        entry = {"rhs": "coil current J = curl psi, seed 2" if rhs == "coil"
                 else "uniform [-1,1) per free DOF, seed 2", "maxit": maxit}
        x_ref = fp64_reference(H, op, b, args) if rhs == "coil" else None
        best = None
        for name, mode in (("fp64_mg", H.SOLVE_FP64 | H.SOLVE_MG), ("fp64", H.SOLVE_FP64),
                           ("fp32_ir", H.SOLVE_FP32_IR), ("fp64_aux", H.SOLVE_FP64 | H.SOLVE_AUX),
                           ("fp32_ir_aux", H.SOLVE_FP32_IR | H.SOLVE_AUX)):
            xs = torch.zeros_like(b)
            st = op.solve(b, xs, rtol=1e-8, maxit=args.solve_maxit if name.endswith(("aux", "mg")) else maxit,
                          mode=mode, raise_on_fail=False)
            entry[name] = solve_entry(st)
            entry[name]["err_vs_fp64_ref"] = rel_err(torch, xs, x_ref)
            entry[name]["fp64_continuations"] = st["fp64_fallbacks"] if (mode & 0xF) == H.SOLVE_FP32_IR else 0
            print(f"[bench] 48^3 p3 {rhs} solve {name}: {entry[name]}", file=sys.stderr, flush=True)
            if best is None or st["rel_res"] < best[0]:
                best = (st["rel_res"], xs)
        if rhs == "uniform":
            fl = ulp_floor(torch, op, b, best[1])
            entry["fp64_ulp_floor_rel_res"] = fl
            entry["best_true_rel_res_over_ulp_floor"] = round(best[0] / fl, 3)
            print(f"[bench] 48^3 p3 uniform: ulp floor {fl:.3e}, best {best[0]:.3e}", file=sys.stderr, flush=True)
        sol[rhs] = entry
    return sol


def solves_multi_rhs_32(torch, H, args):
    """32^3 p=2 two-region, 32 coil right-hand sides (seeds 3..34): plain FP64
    Jacobi-PCG one right-hand side at a time vs one batched FP32 PCG + FP64
    refinement solve of all 32, both to a true relative residual of 1e-8; each
    also with the gradient-space correction of the preconditioner (_aux), and
    multigrid one at a time and with all 32 through one V-cycle (_mg)."""
    n, p, nr = 32, 2, 32
    op = H.Operator(n, p, 1.0, two_region(n)).setup(1.0)
    op.autotune(reps=5)
    N = op.ndofs
    B = torch.empty((nr, N), dtype=torch.float64, device="cuda")
    for r in range(nr):
        op.rhs_coil(3 + r, B[r])
    res = {"workload": f"{n}^3 p={p} two-region 1e4/1e-6, {nr} coil RHS (J = curl psi, seeds 3..{2 + nr})",
           "rtol": 1e-8}
    X_ref = [fp64_reference(H, op, B[r].contiguous(), args) for r in range(nr)]
    # plain FP64 path: one solve per right-hand side (a sample of them)
    k = min(args.c3_fp64_sample, nr)
    for name, mode in (("fp64_single", H.SOLVE_FP64), ("fp64_aux_single", H.SOLVE_FP64 | H.SOLVE_AUX),
                       ("fp64_mg_single", H.SOLVE_FP64 | H.SOLVE_MG)):
        tot_ms, its, worst, err = 0.0, [], 0.0, 0.0
        for r in range(k):
            xs = torch.zeros(N, dtype=torch.float64, device="cuda")
            st = op.solve(B[r].contiguous(), xs, rtol=1e-8, maxit=args.solve_maxit, mode=mode, raise_on_fail=False)
            tot_ms += st["solve_ms"]
            its.append(st["iterations"])
            worst = max(worst, st["rel_res"])
            e = rel_err(torch, xs, X_ref[r])
            err = None if (e is None or err is None) else max(err, e)
        res[name] = {"sample_rhs": k, "iterations_mean": float(np.mean(its)), "max_true_rel_res": worst,
                     "max_err_vs_fp64_ref": err, "solves_per_s": round(1000.0 * k / tot_ms, 4)}
    for name, mode in (("fp64_batched", H.SOLVE_FP64), ("fp64_aux_batched", H.SOLVE_FP64 | H.SOLVE_AUX),
                       ("fp32_ir_batched", H.SOLVE_FP32_IR), ("fp32_ir_aux_batched", H.SOLVE_FP32_IR | H.SOLVE_AUX),
                       ("fp64_mg_batched", H.SOLVE_FP64 | H.SOLVE_MG)):
        X = torch.zeros_like(B)
        st = op.solve(B, X, rtol=1e-8, maxit=args.solve_maxit, mode=mode, raise_on_fail=False)
        res[name] = solve_entry(st, nr)
        res[name]["max_err_vs_fp64_ref"] = max_err(rel_err(torch, X[r], X_ref[r]) for r in range(nr))
    res["like_for_like"] = like_for_like(res, "batched")
    print(f"[bench] 32^3 p2 32-rhs: {res}", file=sys.stderr, flush=True)
    del op
    return res


def like_for_like(res, kind):
    """Speedups of the batched (or served) solves over one-at-a-time FP64 solves
    with the same preconditioner, and over the plain FP64 Jacobi path."""
    sps = lambda k: res[k]["solves_per_s"] if k in res and res[k].get("solves_per_s") else None
    out = {}
    for pre, base in (("", "fp64_single"), ("_aux", "fp64_aux_single"), ("_mg", "fp64_mg_single")):
        for prec in ("fp64", "fp32_ir"):
            key = f"{prec}{pre}_{kind}"
            if sps(key) and sps(base):
                out[f"{key}_vs_{base}"] = round(sps(key) / sps(base), 3)
    if sps(f"fp32_ir_aux_{kind}") and sps("fp64_single"):
        out[f"fp32_ir_aux_{kind}_vs_fp64_single (preconditioner differs)"] = round(
            sps(f"fp32_ir_aux_{kind}") / sps("fp64_single"), 3)
    if sps(f"fp32_ir_aux_{kind}") and sps("fp64_mg_single"):
        out[f"fp32_ir_aux_{kind}_vs_fp64_mg_single (fastest one-at-a-time FP64 path)"] = round(
            sps(f"fp32_ir_aux_{kind}") / sps("fp64_mg_single"), 3)
    return out


def _splitmix_u01(seed, i):
    z = (seed + (i + 1) * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return ((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0)


def uneven_specs(count=64):
    """Uneven instances: cube edge n ~ U{8..40}, order p ~ U{1,2,3}, sigma/dt of
    the inner region [0.25,0.75]^3 and of the outer region each 10^U[-6,4];
    instance i draws from splitmix64 with seed 4 + i."""
    out = []
    for i in range(count):
        s = 4 + i
        n = 8 + int(_splitmix_u01(s, 0) * 33)
        p = 1 + int(_splitmix_u01(s, 1) * 3)
        s_in = 10.0 ** (-6 + 10 * _splitmix_u01(s, 2))
        s_out = 10.0 ** (-6 + 10 * _splitmix_u01(s, 3))
        out.append((n, p, s_in, s_out))
    return out


def build_uneven(torch, H, count=64):
    ops, bs = [], []
    for i, (n, p, s_in, s_out) in enumerate(uneven_specs(count)):
        op = H.Operator(n, p, 1.0, two_region(n, s_in, s_out)).setup(1.0)
        ops.append(op)
        bs.append(op.rhs_coil(1000 + i))
    torch.cuda.synchronize()
    return ops, bs


def solves_uneven(torch, H, ops, bs, args):
    """64 uneven instances, coil right-hand sides, rtol 1e-8: one batched
    FP32 PCG + FP64 refinement solve (FP64 continuation per instance where FP32
    stalls) against the plain FP64 path solving the instances one at a time;
    each also with the gradient-space correction of the preconditioner (_aux);
    multigrid one at a time and as a batch (worker streams, _mg)."""
    res = {"workload": f"{len(ops)} cubes n~U{{8..40}}, p~U{{1,2,3}}, two regions sigma/dt=10^U[-6,4] "
                       f"(seed 4+i), coil RHS (seed 1000+i); {sum(o.ndofs for o in ops)} DOFs in total",
           "rtol": 1e-8}
    x_refs = [fp64_reference(H, op, b, args) for op, b in zip(ops, bs)]
    for name, mode in (("fp64_batched", H.SOLVE_FP64), ("fp64_aux_batched", H.SOLVE_FP64 | H.SOLVE_AUX),
                       ("fp32_ir_batched", H.SOLVE_FP32_IR), ("fp32_ir_aux_batched", H.SOLVE_FP32_IR | H.SOLVE_AUX),
                       ("fp64_mg_batched", H.SOLVE_FP64 | H.SOLVE_MG)):
        xs = [torch.zeros_like(b) for b in bs]
        bt = H.Batch()
        for op, b, x in zip(ops, bs, xs):
            bt.add(op, b, x, rtol=1e-8, maxit=args.solve_maxit)
        rc, st = bt.solve(mode, raise_on_fail=False)
        ms = st[0]["solve_ms"]
        res[name] = {"solve_ms": round(ms, 1), "solves_per_s": round(1000.0 * len(ops) / ms, 3),
                     "converged": int(sum(s["converged"] for s in st)),
                     "fp64_continuations": int(sum(s["fp64_fallbacks"] for s in st)),
                     "max_true_rel_res": max(s["rel_res"] for s in st),
                     "max_err_vs_fp64_ref": max_err(rel_err(torch, x, xr) for x, xr in zip(xs, x_refs)),
                     "iterations_mean": float(np.mean([s["iterations"] for s in st]))}
        del bt
    for name, mode in (("fp64_single", H.SOLVE_FP64), ("fp64_aux_single", H.SOLVE_FP64 | H.SOLVE_AUX),
                       ("fp64_mg_single", H.SOLVE_FP64 | H.SOLVE_MG)):
        tot, its, conv, worst, err = 0.0, [], 0, 0.0, 0.0
        for op, b, xr in zip(ops, bs, x_refs):
            x = torch.zeros_like(b)
            md = mode if (mode & H.SOLVE_MG) == 0 or mg_ok(op) else H.SOLVE_FP64 | H.SOLVE_AUX
            s1 = op.solve(b, x, rtol=1e-8, maxit=args.solve_maxit, mode=md, raise_on_fail=False)
            tot += s1["solve_ms"]
            its.append(s1["iterations"])
            conv += int(s1["converged"])
            worst = max(worst, s1["rel_res"])
            e = rel_err(torch, x, xr)
            err = None if (e is None or err is None) else max(err, e)
        res[name] = {"solve_ms_total": round(tot, 1), "solves_per_s": round(1000.0 * len(ops) / tot, 3),
                     "converged": conv, "max_true_rel_res": worst, "max_err_vs_fp64_ref": err,
                     "iterations_mean": float(np.mean(its))}
        if mode & H.SOLVE_MG:
            res[name]["instances_on_multigrid"] = int(sum(mg_ok(o) for o in ops))
            res[name]["note"] = "instances whose mesh does not halve down to a small level use fp64_aux"
    res["like_for_like"] = like_for_like(res, "batched")
    print(f"[bench] 64 uneven: {res}", file=sys.stderr, flush=True)
    return res, x_refs


def serve_poisson(torch, H, ops, bs, args, c4, x_refs, aux=True, rate=200.0, nreq=None, mg=False):
    """Poisson arrivals (mean `rate` req/s, seed 5) of backward-Euler step solves
    drawn uniformly from the 64 uneven operators (resident), FP32 PCG + FP64
    refinement to 1e-8 (preconditioner: Jacobi with the gradient-space
    correction when aux, else Jacobi), continuous batching under an HBM budget.
    With u_n = 0 and dt = 1 each result is the solution of the operator's coil
    system, compared with its FP64 reference."""
    rng = np.random.default_rng(5)
    nreq = nreq or args.serve_requests
    gaps = rng.exponential(1.0 / rate, nreq)
    arrivals_ms = np.cumsum(gaps) * 1000.0
    pick = rng.integers(0, len(ops), nreq)
    us = [torch.zeros_like(bs[k]) for k in pick]
    mode = H.SOLVE_FP64 | H.SOLVE_MG if mg else H.SOLVE_FP32_IR | (H.SOLVE_AUX if aux else 0)
    srv = H.Server(args.serve_budget_gb * (1 << 30), mode=mode, rtol=1e-8, maxit=args.solve_maxit, round_iters=16)
    t0 = time.perf_counter()
    for i in range(nreq):
        srv.submit(ops[pick[i]], bs[pick[i]], us[i], 1.0, tag=i, arrival_ms=float(arrivals_ms[i]))
    out = []
    while len(out) < nreq:
        out.extend(srv.poll(wait_ms=1000))
    wall = time.perf_counter() - t0
    info = srv.info()
    srv.close()
    lat = np.array([r["latency_ms"] for r in out])
    span_ms = max(r["arrival_ms"] + r["latency_ms"] for r in out)
    sps = nreq / (span_ms / 1000.0)
    err = max_err(rel_err(torch, us[i], x_refs[pick[i]]) for i in range(nreq))
    same = "fp64_mg_single" if mg else ("fp64_aux_single" if aux else "fp64_single")
    base = c4[same]["solves_per_s"]
    how = ("FP64 multigrid PCG, 4 worker streams (gradient-space PCG where a mesh has no hierarchy)" if mg else
           f"FP32 PCG + FP64 refinement, {'Jacobi + gradient-space correction' if aux else 'Jacobi'} preconditioner")
    res = {"workload": f"{nreq} requests, Poisson arrivals mean {rate:g} req/s (seed 5), operators drawn from the "
                       f"64-operator uneven mix, u_n = 0, rtol 1e-8, {how}",
           "arrival_rate_per_s": rate, "max_err_vs_fp64_ref": err,
           f"speedup_vs_{same} (same preconditioner, one at a time)": round(sps / base, 3),
           "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
           "mean_queue_ms": round(float(np.mean([r["queue_ms"] for r in out])), 2),
           "solves_per_s": round(sps, 3), "converged": int(sum(r["converged"] for r in out)),
           "fp64_continuations": int(sum(r["fp64_fallback"] for r in out)),
           "max_true_rel_res": max(r["rel_res"] for r in out), "peak_active": info["peak_active"],
           "budget_gb": args.serve_budget_gb, "peak_budget_use_gb": round(info["peak_bytes"] / 2 ** 30, 2),
           "rounds": info["rounds"], "wall_s": round(wall, 2),
           "speedup_vs_fp64_single (plain FP64 Jacobi path)": round(sps / c4["fp64_single"]["solves_per_s"], 3),
           "speedup_vs_fp64_mg_single (fastest one-at-a-time FP64 path)":
               round(sps / c4["fp64_mg_single"]["solves_per_s"], 3)}
    print(f"[bench] serving{' mg' if mg else (' aux' if aux else ' jacobi')}: {res}", file=sys.stderr, flush=True)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=48)
    ap.add_argument("--order", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--solve-maxit", type=int, default=20000)
    ap.add_argument("--c3-fp64-sample", type=int, default=32)
    ap.add_argument("--c2-uniform-maxit", type=int, default=5000)
    ap.add_argument("--no-batch", action="store_true", help="skip the uneven batch and serving workloads")
    ap.add_argument("--serve-requests", type=int, default=128)
    ap.add_argument("--serve-budget-gb", type=int, default=16)
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run a 1 GiB device copy (reference kernel for the HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU; --gpus must be 1",
                          "metric": METRIC, "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    import hephaestus_amd as H

    n, p = args.n, args.order
    sig = two_region(n)
    E = n ** 3
    t0 = time.perf_counter()
    op = H.Operator(n, p, 1.0, sig).setup(1.0)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    N = op.ndofs
    print(f"[bench] operator {n}^3 p={p} N={N} setup {setup_s:.2f}s device bytes "
          f"{op.device_bytes / 1e9:.2f} GB", file=sys.stderr, flush=True)
    # the headline apply reads the quadrature-point factors (partial assembly,
    # any hex mesh); the element-factor apply of box meshes is measured after it
    op.set_factor_mode(0)
    tune = op.autotune(reps=5)
    print(f"[bench] autotuned tiles: fp64 {tune['tile64']} ({tune['fp64_us']:.1f} us), "
          f"fp32 {tune['tile32']} ({tune['fp32_us']:.1f} us)", file=sys.stderr, flush=True)
    x = torch.empty(N, dtype=torch.float64, device="cuda")
    op.fill_uniform(2, x)
    y = torch.empty_like(x)
    ms = time_applies(torch, lambda: op.apply(x, y), args.steps, args.warmup)
    B = apply_bytes(N, E, p)
    gdofs = N / (ms * 1e-3) / 1e9
    achieved = B / (ms * 1e-3) / 1e9
    print(f"[bench] FP64 apply {ms * 1e3:.1f} us  {gdofs:.1f} GDOF/s  {achieved:.0f} GB/s",
          file=sys.stderr, flush=True)
    # per-kernel durations (HIP events around each launch, on the launch stream)
    prof = op.profile_apply(x, y, reps=max(args.steps, 5))
    tile_ach = B / (prof["tile_us"] * 1e-6) / 1e9
    print(f"[bench] FP64 tile kernel {prof['tile_us']:.1f} us ({tile_ach:.0f} GB/s algorithmic), "
          f"tile-face assembly {prof['planes_us']:.1f} us", file=sys.stderr, flush=True)
    # measured device copy of 1 GiB (read + write), the achievable-HBM reference
    ca = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
    ca.fill_(1.0)
    cb = torch.empty_like(ca)
    ms_copy = time_applies(torch, lambda: cb.copy_(ca), 10, 2)
    copy_gbs = 2.0 * ca.numel() * 8 / (ms_copy * 1e-3) / 1e9
    print(f"[bench] device copy 1 GiB: {ms_copy * 1e3:.1f} us = {copy_gbs:.0f} GB/s (read + write)",
          file=sys.stderr, flush=True)
    del ca, cb
    if args.copy_calib:
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
        c = torch.empty_like(a)
        for _ in range(3):
            c.copy_(a)
        torch.cuda.synchronize()
    # FP32 apply
    x32 = x.float()
    y32 = torch.empty_like(x32)
    ms32 = time_applies(torch, lambda: op.apply_f32(x32, y32), args.steps, args.warmup)
    B32 = apply_bytes(N, E, p, 4, 4)
    print(f"[bench] FP32 apply {ms32 * 1e3:.1f} us  {N / ms32 / 1e6:.1f} GDOF/s  "
          f"{B32 / ms32 / 1e6:.0f} GB/s", file=sys.stderr, flush=True)

    out = {
        "metric": METRIC,
        "value": round(gdofs, 3),
        "unit": "GDOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (splitmix64 x, uniform [-1,1) on free DOFs)",
        "config": {"workload": f"FP64 PA apply, {n}^3 hex unit cube, ND order {p}, N={N}, nu=1, "
                               f"sigma/dt two-region 1e4/1e-6, PEC all faces",
                   "n": n, "order": p, "ndofs": N, "elements": E,
                   "tile_fp64": list(tune["tile64"]), "tile_fp32": list(tune["tile32"])},
        "roofline": {"bound": "hbm", "achieved": round(tile_ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(tile_ach / HBM_PEAK_GBS, 4), "traffic": measured_traffic(n, p),
                     "kernel": "k_apply_tile (FP64)", "kernel_us": round(prof["tile_us"], 2),
                     "algorithmic_bytes": B,
                     "measured_copy_GB_s": round(copy_gbs, 1),
                     "frac_of_measured_copy": round(tile_ach / copy_gbs, 4),
                     "whole_apply": {"us": round(ms * 1e3, 2), "GB_s": round(achieved, 1),
                                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                                     "assembly_kernel_us": round(prof["planes_us"], 2)}},
        "fp32_apply": {"ms": round(ms32, 5), "GDOF_s": round(N / ms32 / 1e6, 3),
                       "GB_s": round(B32 / ms32 / 1e6, 1)},
    }
    # element factors (the default on box meshes, used by every solve below):
    # two factors per element instead of 12 (p+1)^3, same operator to round-off
    op.set_factor_mode(1)
    tune_ef = op.autotune(reps=10)
    y_q = y.clone()
    op.set_factor_mode(0)
    op.apply(x, y_q)
    op.set_factor_mode(1)
    ms_ef = time_applies(torch, lambda: op.apply(x, y), args.steps, args.warmup)
    diff_ef = ((y - y_q).norm() / y_q.norm()).item()
    prof_ef = op.profile_apply(x, y, reps=max(args.steps, 5))
    ms_ef32 = time_applies(torch, lambda: op.apply_f32(x32, y32), args.steps, args.warmup)
    B_ef = 16 * N + 16 * E
    out["apply_element_factors"] = {
        "workload": f"same operator, box-mesh element factors (dt nu_e, sigma_e per element; "
                    f"factor = coefficient x 1D weights x axis constant)",
        "tile_fp64": list(tune_ef["tile64"]), "tile_fp32": list(tune_ef["tile32"]),
        "fp64_us": round(ms_ef * 1e3, 2), "fp64_GDOF_s": round(N / ms_ef / 1e6, 3),
        "tile_kernel_us": round(prof_ef["tile_us"], 2), "assembly_kernel_us": round(prof_ef["planes_us"], 2),
        "fp32_us": round(ms_ef32 * 1e3, 2), "fp32_GDOF_s": round(N / ms_ef32 / 1e6, 3),
        "algorithmic_bytes": B_ef, "bytes_rule": "B_min = 16 N + 16 E (no per-quadrature-point data streamed)",
        "GB_s": round(B_ef / ms_ef / 1e6, 1), "frac_hbm": round(B_ef / ms_ef / 1e6 / HBM_PEAK_GBS, 4),
        "rel_diff_vs_quadrature_point_factors": diff_ef,
        "bound": "instruction issue (VALU + SALU per wave), not HBM"}
    print(f"[bench] element-factor apply: FP64 {ms_ef * 1e3:.1f} us = {N / ms_ef / 1e6:.1f} GDOF/s "
          f"(tile {prof_ef['tile_us']:.1f} us, assembly {prof_ef['planes_us']:.1f} us), FP32 "
          f"{ms_ef32 * 1e3:.1f} us; rel. diff vs quadrature-point factors {diff_ef:.1e}",
          file=sys.stderr, flush=True)
    del y_q
    # AV formulation block operator on the same mesh: (a, V) with V on the H1 nodes
    NN = op.nnodes
    v = torch.empty(NN, dtype=torch.float64, device="cuda")
    v.uniform_(-1.0, 1.0, generator=torch.Generator(device="cuda").manual_seed(6))
    ya, yv = torch.empty_like(x), torch.empty_like(v)
    ms_av = time_applies(torch, lambda: op.av_apply(x, v, ya, yv), args.steps, args.warmup)
    out["av_apply"] = {"workload": f"AV block operator (dt K + M, M G; G^T M, G^T M G), {n}^3 order {p}, "
                                   f"{N} ND DOFs + {NN} H1 nodes",
                       "ms": round(ms_av, 5), "GDOF_s": round((N + NN) / ms_av / 1e6, 3),
                       "kernels": "gather a0 + G v, mass-only tile apply, G^T, curl-only tile apply, finish"}
    print(f"[bench] AV block apply {ms_av * 1e3:.1f} us  {(N + NN) / ms_av / 1e6:.1f} GDOF/s",
          file=sys.stderr, flush=True)
    # AV block solve on a consistent right-hand side b = AV (a_t, v_t), random state
    if not args.no_solve:
        op.av_apply(x, v, ya, yv)
        a0, v0 = torch.zeros_like(x), torch.zeros_like(v)
        st = op.av_solve(ya, yv, a0, v0, rtol=1e-8, maxit=5000, raise_on_fail=False)
        out["av_solve"] = {"workload": f"AV block system, {n}^3 order {p}, b = AV (a_t, v_t) with a_t the "
                                       f"uniform-[-1,1) ND vector (seed 2), v_t uniform nodes (seed 6); "
                                       f"PCG with multigrid V-cycles on both blocks (block Jacobi where a "
                                       f"mesh has no hierarchy), rtol 1e-8, from zero",
                           "preconditioner": "multigrid blocks" if st["outer_iterations"] == 2 else "block Jacobi",
                           "iterations": st["iterations"], "converged": bool(st["converged"]),
                           "true_rel_res": st["rel_res"], "solve_ms": round(st["solve_ms"], 3)}
        print(f"[bench] AV block solve: {out['av_solve']}", file=sys.stderr, flush=True)
        del a0, v0
    del v, ya, yv
    if not args.no_solve:
        out["solves"] = {"single_48_p3": solves_single_48(torch, H, op, x, args),
                         "multi_rhs_32_p2": solves_multi_rhs_32(torch, H, args)}
        if not args.no_batch:
            del op
            torch.cuda.empty_cache()
            ops4, bs4 = build_uneven(torch, H)
            c4, x_refs = solves_uneven(torch, H, ops4, bs4, args)
            out["solves"]["uneven_64"] = c4
            out["serving"] = serve_poisson(torch, H, ops4, bs4, args, c4, x_refs, aux=True)
            out["serving_jacobi"] = serve_poisson(torch, H, ops4, bs4, args, c4, x_refs, aux=False)
            # latency at an arrival rate below the service rate (a load the server sustains)
            sust = 0.5 * out["serving"]["solves_per_s"]
            out["serving_sustained"] = serve_poisson(torch, H, ops4, bs4, args, c4, x_refs, aux=True,
                                                     rate=round(sust, 1), nreq=min(64, args.serve_requests))
            out["serving_mg"] = serve_poisson(torch, H, ops4, bs4, args, c4, x_refs, mg=True)
            sust = 0.5 * out["serving_mg"]["solves_per_s"]
            out["serving_mg_sustained"] = serve_poisson(torch, H, ops4, bs4, args, c4, x_refs, mg=True,
                                                        rate=round(sust, 1), nreq=min(64, args.serve_requests))
            del ops4, bs4, x_refs
            torch.cuda.empty_cache()
    if not args.no_cpu:
        op = None
        torch.cuda.empty_cache()
        cb = cpu_baseline(n, p, sig, args.cpu_budget)
        out["cpu_baseline"] = cb
        print(f"[bench] CPU baseline {cb}", file=sys.stderr, flush=True)
        if not args.no_solve:
            out["cpu_csr_path"] = cpu_csr_path(torch, H, min(args.cpu_budget, 5.0))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
