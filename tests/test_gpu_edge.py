"""Edge cases of the apply and the solvers (needs an MI355X), through the C ABI:
meshes of one or a few elements (every tile variant then covers a partial
tile), a space whose DOFs are all PEC, zero and single right-hand sides,
repeated operators in one batch, and argument errors returned as status codes.
"""
import ctypes as C

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
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("n,pec", [((1, 1, 1), 0x00), ((1, 1, 1), 0x3F), ((1, 2, 1), 0x15), ((2, 1, 3), 0x3F)])
def test_tiny_meshes_every_tile(env, p, n, pec):
    torch, H, O = env
    from hephaestus_amd import hcurl
    print(f"case tiny p={p} n={n} pec={pec:#x}", flush=True)
    rng = np.random.default_rng(77 + p)
    m = O.Mesh(*n, pec_faces=pec)
    E = m.nelem
    nu, sig = rng.uniform(0.5, 2.0, E), 10.0 ** rng.uniform(-2, 2, E)
    qd = O.pa_setup(m, p, nu, sig)
    ess = O.ess_mask(m, p)
    x = rng.standard_normal(O.ndofs(m, p))
    y_ref = O.pa_apply(m, p, qd, ess, x)
    op = H.Operator(n, p, nu, sig, pec_faces=pec).setup(1.0)
    xd = torch.from_numpy(x).cuda()
    for t in hcurl.tile_list(p):
        op.set_tile(64, *t)
        op.set_tile(32, *t)
        assert _rel(op.apply(xd).cpu().numpy(), y_ref) <= 1e-13, t
        assert _rel(op.apply_f32(xd.float()).double().cpu().numpy(), y_ref) <= 2e-6, t


def test_all_dofs_pec(env):
    """One element of order 1 with PEC on every face: all 12 DOFs are
    essential, A = I, and the solve returns b."""
    torch, H, O = env
    op = H.Operator(1, 1, 1.0, 1.0).setup(1.0)
    assert op.ndofs == 12
    b = torch.arange(1.0, 13.0, dtype=torch.float64, device="cuda")
    assert torch.equal(op.apply(b), b)
    x = torch.zeros_like(b)
    st = op.solve(b, x, rtol=1e-12, maxit=10)
    assert st["converged"] and torch.allclose(x, b, rtol=0, atol=1e-14)


@pytest.mark.parametrize("mode", ["fp64", "fp32_ir_aux"])
def test_zero_rhs_and_single_rhs_shapes(env, mode):
    torch, H, O = env
    md = H.SOLVE_FP64 if mode == "fp64" else H.SOLVE_FP32_IR | H.SOLVE_AUX
    op = H.Operator((5, 4, 3), 2, 1.0, 2.0).setup(1.0)
    z = torch.zeros(op.ndofs, dtype=torch.float64, device="cuda")
    x = torch.zeros_like(z)
    st = op.solve(z, x, rtol=1e-8, maxit=100, mode=md)
    assert st["converged"] and st["iterations"] == 0 and not x.any()
    # (1, N) and (N,) give the same bits
    b = torch.empty_like(z)
    op.fill_uniform(3, b)
    x1, x2 = torch.zeros_like(b), torch.zeros((1, op.ndofs), dtype=torch.float64, device="cuda")
    op.solve(b, x1, rtol=1e-9, maxit=5000, mode=md)
    op.solve(b.view(1, -1), x2, rtol=1e-9, maxit=5000, mode=md)
    assert torch.equal(x1, x2[0])


def test_batch_with_repeated_operator(env):
    """The same operator added twice with different right-hand sides, and once
    more with the first right-hand side: the repeated pair is bitwise equal."""
    torch, H, O = env
    op = H.Operator((6, 5, 4), 3, 1.0, 5.0).setup(1.0)
    bs = [torch.empty(op.ndofs, dtype=torch.float64, device="cuda") for _ in range(2)]
    op.fill_uniform(5, bs[0])
    op.fill_uniform(6, bs[1])
    xs = [torch.zeros_like(bs[0]) for _ in range(3)]
    bt = H.Batch()
    for b, x in zip([bs[0], bs[1], bs[0]], xs):
        bt.add(op, b, x, rtol=1e-9, maxit=10000)
    rc, st = bt.solve(H.SOLVE_FP64)
    assert rc == 0 and all(s["converged"] for s in st)
    assert torch.equal(xs[0], xs[2]) and not torch.equal(xs[0], xs[1])


def test_c_abi_argument_errors(env):
    torch, H, O = env
    from hephaestus_amd import hcurl
    lib = hcurl.load()
    op = H.Operator(3, 1, 1.0, 1.0).setup(1.0)
    vp = C.c_void_p
    assert lib.hc_op_apply(op._h, None, None, 1, None) == -1
    assert lib.hc_op_apply(None, None, None, 1, None) == -1
    x = torch.zeros(op.ndofs, dtype=torch.float64, device="cuda")
    assert lib.hc_op_apply(op._h, vp(x.data_ptr()), vp(x.data_ptr()), 0, None) == -1
    assert b"hc_op_apply" in lib.hc_last_error()
    v = torch.zeros(op.nnodes, dtype=torch.float64, device="cuda")
    # outputs aliasing inputs are refused
    assert lib.hc_op_av_apply(op._h, vp(x.data_ptr()), vp(v.data_ptr()), vp(x.data_ptr()), vp(v.data_ptr()),
                              None) == -1
