"""Independent numpy brute-force route for the ND_p hex operator.

Used only by the tests to pin the oracle.  Nothing here shares code with
oracle/hc_oracle.c: the 1D points come from numpy's Legendre module, the
Lagrange polynomials are built from their roots with numpy.polynomial, the
element Jacobian is taken by central finite differences of the trilinear
vertex map, and element matrices are formed from physical-space basis vectors
(covariant Piola: u = J^{-T} u_hat, curl u = J curl_hat u_hat / det J)
rather than from the 6+6 symmetric quadrature factors.

Numbering convention (the same one the oracle documents): component c has the
open (Gauss-Legendre) direction c and closed (Gauss-Lobatto) directions
elsewhere; global grid per component (n_d*p or n_d*p+1 points per axis),
x fastest; element-local order is component x, y, z, each [dz][dy][dx].
"""
from __future__ import annotations

import numpy as np
import numpy.polynomial.legendre as L
import numpy.polynomial.polynomial as P
import scipy.sparse as sp


def gl_points(n):
    x, w = L.leggauss(n)
    return (x + 1) / 2, w / 2


def gll_points(n):
    p = n - 1
    inner = L.Legendre.basis(p).deriv().roots() if p >= 2 else np.array([])
    return np.concatenate([[0.0], (np.sort(inner.real) + 1) / 2, [1.0]])


def lagrange_polys(nodes):
    """Coefficient arrays (power basis) of the Lagrange polynomials."""
    polys = []
    for j, xj in enumerate(nodes):
        others = np.delete(nodes, j)
        c = P.polyfromroots(others) if len(others) else np.array([1.0])
        c = c / P.polyval(xj, c)
        polys.append(c)
    return polys


class Basis1D:
    def __init__(self, p):
        self.p = p
        self.open_nodes = gl_points(p)[0]
        self.closed_nodes = gll_points(p + 1)
        self.op = lagrange_polys(self.open_nodes)
        self.cp = lagrange_polys(self.closed_nodes)

    def open(self, t):
        return np.array([P.polyval(t, c) for c in self.op])  # (p, len t)

    def closed(self, t):
        return np.array([P.polyval(t, c) for c in self.cp])

    def dclosed(self, t):
        return np.array([P.polyval(t, P.polyder(c)) for c in self.cp])


def comp_dims(n, p, c):
    return tuple(n[d] * p if d == c else n[d] * p + 1 for d in range(3))


def ndofs(n, p):
    return sum(int(np.prod(comp_dims(n, p, c))) for c in range(3))


def elem_gids(n, p, ex, ey, ez):
    out = []
    off = 0
    for c in range(3):
        dims = comp_dims(n, p, c)
        nn = [p if d == c else p + 1 for d in range(3)]
        for dz in range(nn[2]):
            for dy in range(nn[1]):
                for dx in range(nn[0]):
                    ix, iy, iz = ex * p + dx, ey * p + dy, ez * p + dz
                    out.append(off + (iz * dims[1] + iy) * dims[0] + ix)
        off += int(np.prod(dims))
    return np.array(out)


def ess_mask(n, p, pec_faces=0x3F):
    out = []
    for c in range(3):
        dims = comp_dims(n, p, c)
        iz, iy, ix = np.meshgrid(*(np.arange(d) for d in dims[::-1]), indexing="ij")
        idx = (ix, iy, iz)
        e = np.zeros(dims[::-1], dtype=bool)
        for a in range(3):
            if a == c:
                continue
            if pec_faces >> (2 * a) & 1:
                e |= idx[a] == 0
            if pec_faces >> (2 * a + 1) & 1:
                e |= idx[a] == dims[a] - 1
        out.append(e.ravel())
    return np.concatenate(out)


def ref_basis(b: Basis1D, pts):
    """Reference-element basis values and curls at points pts (npts, 3).

    Returns phi (nd, npts, 3), curl (nd, npts, 3)."""
    p = b.p
    x, y, z = pts[:, 0], pts[:, 1], pts[:, 2]
    O = {0: b.open(x), 1: b.open(y), 2: b.open(z)}
    Cc = {0: b.closed(x), 1: b.closed(y), 2: b.closed(z)}
    dC = {0: b.dclosed(x), 1: b.dclosed(y), 2: b.dclosed(z)}
    phis, curls = [], []
    zero = np.zeros_like(x)
    for c in range(3):
        nn = [p if d == c else p + 1 for d in range(3)]
        for dz in range(nn[2]):
            for dy in range(nn[1]):
                for dx in range(nn[0]):
                    idx = (dx, dy, dz)
                    f = [O[d][idx[d]] if d == c else Cc[d][idx[d]] for d in range(3)]
                    g = [zero if d == c else dC[d][idx[d]] for d in range(3)]
                    val = f[0] * f[1] * f[2]
                    d_dx = g[0] * f[1] * f[2]
                    d_dy = f[0] * g[1] * f[2]
                    d_dz = f[0] * f[1] * g[2]
                    v = [zero, zero, zero]
                    v[c] = val
                    # curl of (f e_c)
                    if c == 0:
                        cu = [zero, d_dz, -d_dy]
                    elif c == 1:
                        cu = [-d_dz, zero, d_dx]
                    else:
                        cu = [d_dy, -d_dx, zero]
                    phis.append(np.stack(v, axis=-1))
                    curls.append(np.stack(cu, axis=-1))
    return np.array(phis), np.array(curls)


def box_vertices(n, x0=(0, 0, 0), length=(1, 1, 1)):
    xs = [x0[d] + length[d] * np.arange(n[d] + 1) / n[d] for d in range(3)]
    Z, Y, X = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)


def elem_map(vert, n, e_idx, ref):
    """Trilinear map of element (ex,ey,ez) at reference points ref (npts,3)."""
    ex, ey, ez = e_idx
    V = vert.reshape(n[2] + 1, n[1] + 1, n[0] + 1, 3)
    out = np.zeros((ref.shape[0], 3))
    for kc in (0, 1):
        for kb in (0, 1):
            for ka in (0, 1):
                la = ref[:, 0] if ka else 1 - ref[:, 0]
                lb = ref[:, 1] if kb else 1 - ref[:, 1]
                lc = ref[:, 2] if kc else 1 - ref[:, 2]
                out += (la * lb * lc)[:, None] * V[ez + kc, ey + kb, ex + ka][None, :]
    return out


def fd_jacobian(vert, n, e_idx, ref, h=0.5):
    """J[:, r, s] = d x_r / d xi_s by central differences.  The trilinear map is
    affine in each reference variable, so a central difference of any step is
    exact; a large step keeps the round-off at machine level."""
    J = np.zeros((ref.shape[0], 3, 3))
    for s in range(3):
        d = np.zeros(3)
        d[s] = h
        J[:, :, s] = (elem_map(vert, n, e_idx, ref + d) - elem_map(vert, n, e_idx, ref - d)) / (2 * h)
    return J


def quad_rule(Q):
    x, w = gl_points(Q)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    WZ, WY, WX = np.meshgrid(w, w, w, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1), (WX * WY * WZ).ravel()


def element_matrices(vert, n, p, nu, sig, e_idx, Q=None):
    """(K_e, M_e) of one element by physical-space quadrature."""
    b = Basis1D(p)
    Q = Q or p + 1
    pts, w = quad_rule(Q)
    phi, cphi = ref_basis(b, pts)
    J = fd_jacobian(vert, n, e_idx, pts)
    det = np.linalg.det(J)
    Jinv = np.linalg.inv(J)
    # physical vectors: u = J^{-T} phi_hat, curl u = J c_hat / det
    u = np.einsum("qsr,iqs->iqr", Jinv, phi)  # J^{-T} phi: (J^{-1})^T_{rs}=Jinv[s,r]
    cu = np.einsum("qrs,iqs->iqr", J, cphi) / det[None, :, None]
    wd = w * np.abs(det)
    M = sig * np.einsum("iqr,jqr,q->ij", u, u, wd)
    K = nu * np.einsum("iqr,jqr,q->ij", cu, cu, wd)
    return K, M


def assemble(n, p, nu_e, sig_e, vert=None, pec_faces=0x3F, Q=None):
    """Global (K+M) with PEC elimination as scipy CSR, plus the raw K, M."""
    n = tuple(n)
    if vert is None:
        vert = box_vertices(n)
    N = ndofs(n, p)
    rows, cols, kv, mv = [], [], [], []
    e = 0
    for ez in range(n[2]):
        for ey in range(n[1]):
            for ex in range(n[0]):
                K, M = element_matrices(vert, n, p, nu_e[e], sig_e[e], (ex, ey, ez), Q)
                g = elem_gids(n, p, ex, ey, ez)
                rr, cc = np.meshgrid(g, g, indexing="ij")
                rows.append(rr.ravel())
                cols.append(cc.ravel())
                kv.append(K.ravel())
                mv.append(M.ravel())
                e += 1
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    Kg = sp.csr_matrix((np.concatenate(kv), (rows, cols)), shape=(N, N))
    Mg = sp.csr_matrix((np.concatenate(mv), (rows, cols)), shape=(N, N))
    ess = ess_mask(n, p, pec_faces)
    A = eliminate(Kg + Mg, ess)
    return A, Kg, Mg, ess


def eliminate(A, ess):
    free = (~ess).astype(float)
    Df = sp.diags(free)
    return (Df @ A @ Df + sp.diags(ess.astype(float))).tocsr()


def h1_gradient(n, p):
    """Discrete gradient G: H1 Q_p (GLL nodal, global grid (n p+1)^3) -> ND_p,
    ND DOF = tangential derivative in reference coordinates at the ND point."""
    b = Basis1D(p)
    dL = b.dclosed(b.open_nodes)  # (p+1 closed fns, p open pts)
    h1dims = tuple(n[d] * p + 1 for d in range(3))
    Nh = int(np.prod(h1dims))
    N = ndofs(n, p)
    rows, cols, vals = [], [], []
    off = 0
    for c in range(3):
        dims = comp_dims(n, p, c)
        for iz in range(dims[2]):
            for iy in range(dims[1]):
                for ix in range(dims[0]):
                    idx = [ix, iy, iz]
                    r = off + (iz * dims[1] + iy) * dims[0] + ix
                    e = idx[c] // p
                    d_open = idx[c] % p
                    for m in range(p + 1):
                        hidx = list(idx)
                        hidx[c] = e * p + m
                        col = (hidx[2] * h1dims[1] + hidx[1]) * h1dims[0] + hidx[0]
                        rows.append(r)
                        cols.append(col)
                        vals.append(dL[m, d_open])
        off += int(np.prod(dims))
    return sp.csr_matrix((vals, (rows, cols)), shape=(N, Nh)), h1dims


def interpolate(n, p, field, length=(1, 1, 1)):
    """ND DOFs of a vector field on the box mesh: DOF = h_c * u_c(point)."""
    b = Basis1D(p)
    out = []
    for c in range(3):
        dims = comp_dims(n, p, c)
        coords = []
        for d in range(3):
            h = length[d] / n[d]
            if d == c:
                e = np.arange(dims[d]) // p
                loc = b.open_nodes[np.arange(dims[d]) % p]
            else:
                e = np.minimum(np.arange(dims[d]) // p, n[d] - 1)
                loc = b.closed_nodes[np.arange(dims[d]) - e * p]
            coords.append((e + loc) * h)
        Z, Y, X = np.meshgrid(coords[2], coords[1], coords[0], indexing="ij")
        u = field(X.ravel(), Y.ravel(), Z.ravel())[c]
        out.append(u * length[c] / n[c])
    return np.concatenate(out)
