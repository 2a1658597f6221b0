"""Pins of the oracle's AV formulation blocks (oracle.grad_matrix,
oracle.av_apply), CPU only.

The forms are those stated in the reference's kernel headers
(src/kernels/mixed_vector_gradient_kernel.hpp:8 (sigma grad V, u'),
src/kernels/vector_fe_weak_divergence_kernel.hpp:8 (sigma u, grad V'),
src/kernels/diffusion_kernel.hpp:8 (sigma grad V, grad V')).  Independent
routes, none of which shares code with the oracle:
* the discrete gradient against tests/brute.py's gradient (numpy Lagrange
  polynomials from their roots);
* the diffusion block G^T M G against an H1 stiffness matrix assembled by
  quadrature of the tensor Q_p Lagrange basis, and at p = 1 against the
  textbook trilinear brick stiffness (h/3 on the diagonal, 0 between edge
  neighbours, -h/12 across a face or body diagonal);
* the mixed block M G against (sigma grad psi_j, w_i) formed from the brute
  ND basis functions and the H1 basis gradients by quadrature;
* av_apply's gauge null space: (a, v) = (-G phi, phi) gives zero without PEC.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import brute
from oracle import oracle as O


@pytest.fixture(scope="module", autouse=True)
def _built():
    O.build()


def _h1_nodes(n, p, ex, ey, ez):
    M = [n[a] * p + 1 for a in range(3)]
    out = []
    for c in range(p + 1):
        for b in range(p + 1):
            for a in range(p + 1):
                out.append(((ez * p + c) * M[1] + (ey * p + b)) * M[0] + (ex * p + a))
    return np.array(out), M


def _brute_blocks(n, p, sig, length=(1.0, 1.0, 1.0)):
    """H1 stiffness S (NN x NN) and mixed matrix B (N x NN) on the box mesh."""
    b1 = brute.Basis1D(p)
    Q = p + 1
    pts, w = brute.quad_rule(Q)
    h = [length[a] / n[a] for a in range(3)]
    det = h[0] * h[1] * h[2]
    # H1 basis gradients (physical) at the points: (nodes, npts, 3), a fastest
    Cc = [b1.closed(pts[:, d]) for d in range(3)]
    dC = [b1.dclosed(pts[:, d]) for d in range(3)]
    grads = []
    for c in range(p + 1):
        for bb in range(p + 1):
            for a in range(p + 1):
                gx = dC[0][a] * Cc[1][bb] * Cc[2][c] / h[0]
                gy = Cc[0][a] * dC[1][bb] * Cc[2][c] / h[1]
                gz = Cc[0][a] * Cc[1][bb] * dC[2][c] / h[2]
                grads.append(np.stack([gx, gy, gz], axis=-1))
    grads = np.array(grads)
    phi, _ = brute.ref_basis(b1, pts)
    u = phi / np.array(h)[None, None, :]          # covariant map on a box: J^-T = diag(1/h)
    N = brute.ndofs(n, p)
    NN = int(np.prod([n[a] * p + 1 for a in range(3)]))
    Sr, Sc, Sv, Br, Bc, Bv = [], [], [], [], [], []
    e = 0
    for ez in range(n[2]):
        for ey in range(n[1]):
            for ex in range(n[0]):
                nodes, _ = _h1_nodes(n, p, ex, ey, ez)
                gids = brute.elem_gids(n, p, ex, ey, ez)
                wd = sig[e] * w * det
                Se = np.einsum("iqr,jqr,q->ij", grads, grads, wd)
                Be = np.einsum("iqr,jqr,q->ij", u, grads, wd)
                rr, cc = np.meshgrid(nodes, nodes, indexing="ij")
                Sr.append(rr.ravel()); Sc.append(cc.ravel()); Sv.append(Se.ravel())
                rr, cc = np.meshgrid(gids, nodes, indexing="ij")
                Br.append(rr.ravel()); Bc.append(cc.ravel()); Bv.append(Be.ravel())
                e += 1
    S = sp.csr_matrix((np.concatenate(Sv), (np.concatenate(Sr), np.concatenate(Sc))), shape=(NN, NN))
    B = sp.csr_matrix((np.concatenate(Bv), (np.concatenate(Br), np.concatenate(Bc))), shape=(N, NN))
    return S, B


def _oracle_mass(m, p, sig):
    ess = np.zeros(O.ndofs(m, p), dtype=np.uint8)
    return O.Csr(m, p, O.pa_setup(m, p, 0.0, sig), ess).to_scipy().tocsr()


@pytest.mark.parametrize("p", [1, 2, 3])
def test_grad_matrix_vs_brute(p):
    n = (3, 2, 4)
    G = O.grad_matrix(O.Mesh(*n), p)
    Gb, _ = brute.h1_gradient(n, p)
    assert G.shape == Gb.shape
    assert abs(G - Gb).max() <= 1e-13 * abs(Gb).max()


@pytest.mark.parametrize("p", [1, 2, 3])
def test_diffusion_and_mixed_blocks_vs_quadrature(p):
    n = (3, 2, 2)
    rng = np.random.default_rng(40 + p)
    m = O.Mesh(*n)
    sig = 10.0 ** rng.uniform(-2, 2, m.nelem)
    G = O.grad_matrix(m, p)
    M = _oracle_mass(m, p, sig)
    S, B = _brute_blocks(n, p, sig)
    MG = (M @ G).toarray()
    assert np.abs(MG - B.toarray()).max() <= 1e-12 * np.abs(B.toarray()).max()
    GMG = (G.T @ (M @ G)).toarray()
    assert np.abs(GMG - S.toarray()).max() <= 1e-12 * np.abs(S.toarray()).max()


def test_diffusion_p1_textbook_brick():
    """A one-element mesh, a cube of side h = 1/2, unit sigma."""
    m = O.Mesh(1, 1, 1, length=(0.5, 0.5, 0.5))
    h = 0.5
    G = O.grad_matrix(m, 1)
    S = (G.T @ (_oracle_mass(m, 1, np.ones(1)) @ G)).toarray()
    corners = [(a, b, c) for c in range(2) for b in range(2) for a in range(2)]
    for i, ci in enumerate(corners):
        for j, cj in enumerate(corners):
            d = sum(ci[k] != cj[k] for k in range(3))
            ref = {0: h / 3, 1: 0.0, 2: -h / 12, 3: -h / 12}[d]
            assert abs(S[i, j] - ref) <= 1e-14, (ci, cj, S[i, j], ref)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_av_apply_gauge_null_space_and_symmetry(p):
    n = (3, 3, 2)
    rng = np.random.default_rng(60 + p)
    m0 = O.Mesh(*n, pec_faces=0)
    E = m0.nelem
    nu, sig = rng.uniform(0.5, 2.0, E), 10.0 ** rng.uniform(-2, 2, E)
    G = O.grad_matrix(m0, p)
    phi = rng.standard_normal(G.shape[1])
    ya, yv = O.av_apply(m0, p, nu, sig, 0.3, -(G @ phi), phi)
    a1, v1 = rng.standard_normal(G.shape[0]), rng.standard_normal(G.shape[1])
    za, zv = O.av_apply(m0, p, nu, sig, 0.3, a1, v1)
    scale = max(np.abs(za).max(), np.abs(zv).max())
    assert np.abs(ya).max() <= 1e-12 * scale and np.abs(yv).max() <= 1e-12 * scale
    # symmetric on PEC-consistent vectors (PEC on all faces)
    m = O.Mesh(*n)
    ess = O.ess_mask(m, p).astype(bool)
    a2, v2 = rng.standard_normal(G.shape[0]), rng.standard_normal(G.shape[1])
    a1[ess] = 0.0
    a2[ess] = 0.0
    ya1, yv1 = O.av_apply(m, p, nu, sig, 0.3, a1, v1)
    ya2, yv2 = O.av_apply(m, p, nu, sig, 0.3, a2, v2)
    l, r = a1 @ ya2 + v1 @ yv2, ya1 @ a2 + yv1 @ v2
    assert abs(l - r) <= 1e-12 * (abs(l) + abs(r))
