"""Pins of the oracle's 1D building blocks and input generators (CPU only).

The ND_p hex basis of MFEM's ND_HexahedronElement (used by the reference via
CurlCurlIntegrator / VectorFEMassIntegrator, src/kernels/curl_curl_kernel.cpp:23,
src/kernels/vector_fe_mass_kernel.cpp:23) is built from open Gauss-Legendre and
closed Gauss-Lobatto 1D Lagrange bases; these tests check the oracle's tables
against numpy's Legendre module and textbook closed forms.
"""
import numpy as np
import numpy.polynomial.legendre as L
import pytest

import brute as B


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_gauss_legendre_matches_numpy(oracle, n):
    x, w = oracle.gauss_legendre(n)
    xr, wr = L.leggauss(n)
    np.testing.assert_allclose(x, (xr + 1) / 2, atol=1e-15)
    np.testing.assert_allclose(w, wr / 2, atol=1e-15)
    # exactness: integrates t^k on [0,1] for k <= 2n-1
    for k in range(2 * n):
        assert abs(np.sum(w * x ** k) - 1.0 / (k + 1)) < 1e-14


def test_gauss_lobatto_closed_forms(oracle):
    np.testing.assert_allclose(oracle.gauss_lobatto(2), [0, 1], atol=0)
    np.testing.assert_allclose(oracle.gauss_lobatto(3), [0, 0.5, 1], atol=1e-16)
    s = 1 / np.sqrt(5)
    np.testing.assert_allclose(oracle.gauss_lobatto(4), [0, (1 - s) / 2, (1 + s) / 2, 1], atol=1e-15)
    s7 = np.sqrt(3 / 7)
    np.testing.assert_allclose(oracle.gauss_lobatto(5), [0, (1 - s7) / 2, 0.5, (1 + s7) / 2, 1],
                               atol=1e-15)


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_basis_tables(oracle, p):
    Bo, Bc, Gc, qp, qw = oracle.basis(p)
    b = B.Basis1D(p)
    np.testing.assert_allclose(Bo, b.open(qp).T, atol=1e-14)
    np.testing.assert_allclose(Bc, b.closed(qp).T, atol=1e-14)
    np.testing.assert_allclose(Gc, b.dclosed(qp).T, atol=1e-12)
    # partition of unity; derivative sums to zero
    np.testing.assert_allclose(Bo.sum(1), 1, atol=1e-14)
    np.testing.assert_allclose(Bc.sum(1), 1, atol=1e-14)
    np.testing.assert_allclose(Gc.sum(1), 0, atol=1e-12)


def _splitmix64_stream(seed, n):
    out = []
    s = seed
    M = (1 << 64) - 1
    for _ in range(n):
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        out.append(z ^ (z >> 31))
    return out


def test_splitmix64_published_vectors():
    # Published SplitMix64 outputs for seed 1234567 (Vigna's reference
    # generator, as quoted in the rand_xoshiro test-suite).
    assert _splitmix64_stream(1234567, 5) == [
        6457827717110365317, 3203168211198807973, 9817491932198370423,
        4593380528125082431, 16408922859458223821]


def test_rand_uniform_is_splitmix64(oracle):
    seed = 1234567
    ref = _splitmix64_stream(seed, 64)
    got = oracle.rand_uniform(seed, 64)
    exp = np.array([2.0 * ((u >> 11) * 2.0 ** -53) - 1.0 for u in ref])
    np.testing.assert_array_equal(got, exp)
    assert got.min() >= -1 and got.max() < 1
