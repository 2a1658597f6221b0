"""hc_op_apply_noalloc (needs an MI355X): the entry point that allocates no
device memory gives the bits of hc_op_apply on two calls in a row (the second
on a smaller operator, so the shared static scratch holds the first call's
partials), for every compiled tile variant of the first operator's order, and
refuses an operator whose tile-face partials exceed the scratch."""
import ctypes

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
    return torch, H


def _op(H, n, p, seed):
    rng = np.random.default_rng(seed)
    E = int(np.prod(n))
    return H.Operator(n, p, rng.uniform(0.5, 2.0, E), 10.0 ** rng.uniform(-3, 3, E)).setup(0.7)


def _noalloc(H, torch, op, x):
    """Through the C ABI directly (ctypes), on torch's current stream."""
    lib = H.hcurl.load()
    y = torch.full_like(x, float("nan"))
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = lib.hc_op_apply_noalloc(op._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), st)
    assert rc == 0, lib.hc_last_error()
    return y


def test_noalloc_apply_matches_apply(env):
    torch, H = env
    big = _op(H, (13, 11, 9), 2, 1)
    small = _op(H, (5, 4, 6), 3, 2)
    for op, seed in ((big, 11), (small, 12)):
        x = torch.empty(op.ndofs, dtype=torch.float64, device="cuda")
        op.fill_uniform(seed, x)
        print(f"case noalloc n={op.n} p={op.order}", flush=True)
        tiles = H.hcurl.tile_list(op.order)
        for t in tiles if op is big else [None]:
            if t is not None:
                op.set_tile(64, *t)
            y_ref = op.apply(x)
            y = _noalloc(H, torch, op, x)
            torch.cuda.synchronize()
            assert torch.equal(y, y_ref), (op.n, op.order, t)


def test_noalloc_refuses_too_large(env):
    torch, H = env
    lib = H.hcurl.load()
    cap = lib.hc_noalloc_scratch_doubles()
    assert cap > 0
    # 40^3 order 3 with one-element-thick tile layers: far more tile-plane DOFs than the scratch holds
    op = H.Operator(40, 3, 1.0, 1.0, precision=H.PREC_FP64).setup(1.0)
    op.set_tile(64, 8, 2, 1, -1)
    x = torch.zeros(op.ndofs, dtype=torch.float64, device="cuda")
    y = torch.zeros_like(x)
    rc = lib.hc_op_apply_noalloc(op._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), None)
    assert rc == H.hcurl.HC_EINVAL
    assert b"static scratch" in lib.hc_last_error()
