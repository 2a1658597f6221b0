"""The oracle against the committed golden fixtures (tests/golden/), which were
written by the independent numpy/scipy route (tests/golden/make_golden.py)."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return dict(np.load(os.path.join(GOLD, name), allow_pickle=False))


@pytest.mark.parametrize("p", [1, 2, 3])
def test_tiny_trilinear(oracle, p):
    g = _load(f"tiny_p{p}.npz")
    n = tuple(int(v) for v in g["n"])
    m = oracle.Mesh(*n, vert=g["vert"])
    qd = oracle.pa_setup(m, p, g["nu"], g["sig"])
    ess = oracle.ess_mask(m, p)
    np.testing.assert_array_equal(ess, g["ess"])
    y = oracle.pa_apply(m, p, qd, ess, g["x"])
    assert np.linalg.norm(y - g["y"]) <= 1e-13 * np.linalg.norm(g["y"])
    d = oracle.pa_diagonal(m, p, qd, ess)
    np.testing.assert_allclose(d, g["diag"], rtol=1e-13)


def test_config1(oracle):
    g = _load("config1_16_p1.npz")
    m = oracle.Mesh(16)
    p = 1
    qd = oracle.pa_setup(m, p, 1.0, 1.0)
    ess = oracle.ess_mask(m, p)
    b = oracle.rhs_uniform(m, p, 1, ess)
    np.testing.assert_array_equal(b, g["b"])
    y = oracle.pa_apply(m, p, qd, ess, b)
    assert np.linalg.norm(y - g["Ab"]) <= 1e-14 * np.linalg.norm(g["Ab"])
    d = oracle.pa_diagonal(m, p, qd, ess)
    np.testing.assert_allclose(d, g["diag"], rtol=1e-14)
    x, st = oracle.pcg_pa(m, p, qd, ess, d, b, 1e-10, 2000)
    assert st["converged"]
    assert np.linalg.norm(x - g["x_direct"]) <= 1e-8 * np.linalg.norm(g["x_direct"])


def test_two_region(oracle):
    g = _load("two_region_6_p3.npz")
    m = oracle.Mesh(6)
    p = 3
    sig = oracle.two_region_coef(6, 1e4, 1e-6)
    np.testing.assert_array_equal(sig, g["sig"])
    qd = oracle.pa_setup(m, p, g["nu"], sig)
    ess = oracle.ess_mask(m, p)
    y = oracle.pa_apply(m, p, qd, ess, g["x"])
    assert np.linalg.norm(y - g["y"]) <= 1e-13 * np.linalg.norm(g["y"])
    np.testing.assert_allclose(oracle.pa_diagonal(m, p, qd, ess), g["diag"], rtol=1e-13)
