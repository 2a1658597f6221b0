"""Writes the committed golden fixtures under tests/golden/.

The expected outputs come from the independent numpy route (tests/brute.py:
numpy-built Lagrange bases, finite-difference Jacobians, physical-space element
matrices, scipy sparse assembly) and scipy's direct solver -- not from the C
oracle -- so checking the oracle or the HIP path against these files is a
genuine cross-check.  The reference (hephaestus) holds no golden vectors for
this path, so these are the vector-level fixtures.

Inputs use the counter-based splitmix64 rule (b_i = 2*u01(splitmix64_i(seed))-1
on free DOFs, 0 on PEC DOFs), reimplemented here in Python.

Run:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import scipy.sparse.linalg as spl

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import brute as B  # noqa: E402

M64 = (1 << 64) - 1


def splitmix_uniform(seed, n):
    i = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return 2.0 * ((z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53) - 1.0


def tiny(p, seed):
    n = (3, 2, 2)
    rng = np.random.default_rng(seed)
    vert = B.box_vertices(n) + 0.08 * rng.standard_normal(((n[0] + 1) * (n[1] + 1) * (n[2] + 1), 3)) / max(n)
    E = n[0] * n[1] * n[2]
    nu = rng.uniform(0.5, 2.0, E)
    sig = 10.0 ** rng.uniform(-3, 3, E)
    A, Kg, Mg, ess = B.assemble(n, p, nu, sig, vert=vert)
    x = splitmix_uniform(seed, A.shape[0])
    return dict(n=np.array(n), p=p, vert=vert, nu=nu, sig=sig, ess=ess.astype(np.uint8),
                x=x, y=A @ x, diag=A.diagonal().copy())


def config1():
    n, p = (16, 16, 16), 1
    E = 16 ** 3
    A, _, _, ess = B.assemble(n, p, np.ones(E), np.ones(E))
    b = splitmix_uniform(1, A.shape[0])
    b[ess] = 0.0
    x = spl.spsolve(A.tocsc(), b)
    y = A @ b
    return dict(n=np.array(n), p=p, ess=ess.astype(np.uint8), b=b, Ab=y,
                diag=A.diagonal().copy(), x_direct=x)


def two_region(p, nn, seed):
    n = (nn, nn, nn)
    c = (np.arange(nn) + 0.5) / nn
    ins = (c >= 0.25) & (c <= 0.75)
    Z, Y, X = np.meshgrid(ins, ins, ins, indexing="ij")
    sig = np.where((X & Y & Z).ravel(), 1e4, 1e-6)
    nu = np.ones(nn ** 3)
    A, _, _, ess = B.assemble(n, p, nu, sig)
    x = splitmix_uniform(seed, A.shape[0])
    return dict(n=np.array(n), p=p, nu=nu, sig=sig, ess=ess.astype(np.uint8), x=x, y=A @ x,
                diag=A.diagonal().copy())


def main():
    for p in (1, 2, 3):
        np.savez_compressed(os.path.join(HERE, f"tiny_p{p}.npz"), **tiny(p, 40 + p))
    np.savez_compressed(os.path.join(HERE, "config1_16_p1.npz"), **config1())
    np.savez_compressed(os.path.join(HERE, "two_region_6_p3.npz"), **two_region(3, 6, 7))
    print("written to", HERE)


if __name__ == "__main__":
    main()
