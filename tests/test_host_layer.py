"""The C++ host layer (hephaestus_amd/host: CurlCurlKernel / VectorFEMassKernel /
WeakCurlCurlKernel / CurlCurlEquationSystem / DefaultJacobiPCGSolver /
TimeDomainEquationSystemProblemOperator over libhcurl.so).

CPU: the framework logic of the test program (parameters, coefficient algebra,
kernel registration, error paths).  GPU: backward-Euler steps of the two-region
curl-curl system driven entirely from C++, against the same steps computed by
the oracle with a sparse direct solve: u_{k+1} = u_k + dt k,
(dt K_nu + M_sigma) k = b_coil - K_nu u_k; relative L2 <= 1e-8 (PCG stops at
1e-12 relative residual).  With DivFreeSource in place of the coil source the
right-hand side is M g, g = M^-1 b_coil (unit mass), optionally Helmholtz
projected, restated for the oracle with scipy sparse solves.  The VTK field
output written after the steps holds the field at the element centres.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "hephaestus_amd", "host")
BIN = os.path.join(HOST, "build", "test_host")


def _built():
    if not os.path.exists(BIN):
        subprocess.run(["make", "-s", "-C", HOST], check=True)
    return BIN


def test_host_layer_cpu():
    r = subprocess.run([_built(), "cpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "cpu tests: ok" in r.stdout


def _two_region(n, inner, outer):
    c = (np.arange(n) + 0.5) / n
    ins = (c >= 0.25) & (c <= 0.75)
    Z, Y, X = np.meshgrid(ins, ins, ins, indexing="ij")
    return np.where((X & Y & Z).ravel(), inner, outer).astype(np.float64)


def _div_free_rhs(O, m, p, j, project):
    """DivFreeSource::Apply restated with scipy: g = M^-1 j (unit mass, no
    elimination), optionally g - G Q with (G^T M G) Q = G^T M g (Q pinned at
    node 0: the system is singular in constants), then M g."""
    import scipy.sparse.linalg as spla
    N = O.ndofs(m, p)
    M = O.Csr(m, p, O.pa_setup(m, p, 0.0, 1.0), np.zeros(N, dtype=np.uint8)).to_scipy().tocsc()
    g = spla.spsolve(M, j)
    if project:
        G = O.grad_matrix(m, p)
        S = (G.T @ M @ G).tocsc()
        rhs = G.T @ (M @ g)
        Q = np.zeros(S.shape[0])
        Q[1:] = spla.spsolve(S[1:, 1:], rhs[1:])
        g = g - G @ Q
    return M @ g


@pytest.mark.gpu
@pytest.mark.parametrize("precision,pc,source", [("fp64", "Jacobi", "Coil"), ("fp32_ir", "Jacobi", "Coil"),
                                                 ("fp32_ir", "AMS", "Coil"), ("fp64", "Jacobi", "DivFree"),
                                                 ("fp64", "Jacobi", "DivFreeProj")])
def test_host_layer_time_steps(tmp_path, precision, pc, source):
    import scipy.sparse.linalg as spla
    from oracle import oracle as O
    n, p, steps, dt = 6, 2, 3, 0.1
    print(f"case host layer {precision} {pc} {source} n={n} p={p} steps={steps}", flush=True)
    r = subprocess.run([_built(), "gpu", str(tmp_path), str(n), str(p), str(steps), str(dt), precision, pc, source],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "gpu tests: ok" in r.stdout
    O.build()
    m = O.Mesh(n)
    nu = _two_region(n, 0.5, 1.0)
    sig = _two_region(n, 1e2, 1e-2)
    ess = O.ess_mask(m, p)
    qK = O.pa_setup(m, p, nu, np.zeros(m.nelem))
    qA = O.pa_setup(m, p, dt * nu, sig)
    A = O.Csr(m, p, qA, ess).to_scipy().tocsc()
    solve = spla.factorized(A)
    b = O.rhs_coil(m, p, 11, ess)
    if source != "Coil":
        b = _div_free_rhs(O, m, p, b, source == "DivFreeProj")
    u = np.zeros(O.ndofs(m, p))
    for k in range(steps):
        f = b - O.pa_apply(m, p, qK, ess, u)
        f[ess.astype(bool)] = 0.0
        u = u + dt * solve(f)
        ug = np.fromfile(tmp_path / f"u_{k}.bin", dtype=np.float64)
        err = np.linalg.norm(ug - u) / np.linalg.norm(u)
        print(f"step {k}: relative L2 difference {err:.3e}")
        assert err <= 1e-8
    _check_vtk(tmp_path / "fields_000001.vtk", n, p, np.fromfile(tmp_path / f"u_{steps - 1}.bin", dtype=np.float64))


def _check_vtk(path, n, p, u):
    """The VTK output's cell vectors equal the field at the element centres,
    evaluated here from the DOFs with numpy Lagrange polynomials (box mesh:
    u_c = u_hat_c / h)."""
    import brute
    b = brute.Basis1D(p)
    bo, bc = b.open(np.array([0.5]))[:, 0], b.closed(np.array([0.5]))[:, 0]
    h = 1.0 / n
    ref = np.zeros((n, n, n, 3))
    off = 0
    for c in range(3):
        dims = brute.comp_dims((n, n, n), p, c)
        U = u[off:off + int(np.prod(dims))].reshape(dims[2], dims[1], dims[0])
        off += int(np.prod(dims))
        w = [bo if a == c else bc for a in range(3)]
        for ez in range(n):
            for ey in range(n):
                for ex in range(n):
                    blk = U[ez * p:ez * p + len(w[2]), ey * p:ey * p + len(w[1]), ex * p:ex * p + len(w[0])]
                    ref[ez, ey, ex, c] = np.einsum("k,j,i,kji->", w[2], w[1], w[0], blk) / h
    lines = path.read_text().splitlines()
    assert lines[0].startswith("# vtk DataFile") and "STRUCTURED_GRID" in lines[3]
    i = lines.index("VECTORS h double")
    vals = np.array([[float(t) for t in l.split()] for l in lines[i + 1:i + 1 + n ** 3]])
    assert np.abs(vals - ref.reshape(-1, 3)).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.gpu
def test_host_layer_av_steps(tmp_path):
    """AVEquationSystem steps from C++ (AddKernels as av_formulation.cpp:150-196,
    DivFreeSource with the Helmholtz projection) against the block system
    restated with scipy (PEC rows/columns of A eliminated; the system is
    singular in the gauge fields (-G phi, phi), so the reference solve adds a
    1e-12 shift on the V block and its residual is checked); the
    gauge-invariant parts agree: the field a0 + G v that sigma acts on, and
    K A (the curl content of A)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    from oracle import oracle as O
    n, p, steps, dt = 5, 2, 2, 0.1
    print(f"case host AV n={n} p={p} steps={steps}", flush=True)
    r = subprocess.run([_built(), "av", str(tmp_path), str(n), str(p), str(steps), str(dt)],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "av tests: ok" in r.stdout
    O.build()
    m = O.Mesh(n)
    N = O.ndofs(m, p)
    nu = _two_region(n, 0.5, 1.0)
    sig = _two_region(n, 1e2, 1e-2)
    ess8 = O.ess_mask(m, p)
    ess = ess8.astype(bool)
    none = np.zeros(N, dtype=np.uint8)
    K = O.Csr(m, p, O.pa_setup(m, p, nu, 0.0), none).to_scipy().tocsr()
    M = O.Csr(m, p, O.pa_setup(m, p, 0.0, sig), none).to_scipy().tocsr()
    G = O.grad_matrix(m, p)
    b_src = _div_free_rhs(O, m, p, O.rhs_coil(m, p, 11, ess8), True)
    F = sp.diags((~ess).astype(float))
    S = sp.bmat([[F @ (dt * K + M) @ F + sp.diags(ess.astype(float)), F @ M @ G],
                 [G.T @ M @ F, G.T @ M @ G]]).tocsr()
    NN = G.shape[1]
    solve = spla.factorized((S + sp.diags(np.concatenate([np.zeros(N), np.full(NN, 1e-12)]))).tocsc())
    A = np.zeros(N)
    for k in range(steps):
        ba = b_src - K @ A
        ba[ess] = 0.0
        rhs = np.concatenate([ba, np.zeros(NN)])
        x = solve(rhs)
        assert np.linalg.norm(rhs - S @ x) <= 1e-10 * np.linalg.norm(rhs)
        kk, vv = x[:N], x[N:]
        A = A + dt * kk
        kg = np.fromfile(tmp_path / f"k_{k}.bin", dtype=np.float64)
        vg = np.fromfile(tmp_path / f"v_{k}.bin", dtype=np.float64)
        ag = np.fromfile(tmp_path / f"a_{k}.bin", dtype=np.float64)
        s_ref = np.where(ess, 0.0, kk) + G @ vv
        s_g = np.where(ess, 0.0, kg) + G @ vg
        e1 = np.linalg.norm(s_g - s_ref) / np.linalg.norm(s_ref)
        e2 = np.linalg.norm(K @ ag - K @ A) / np.linalg.norm(K @ A)
        print(f"av step {k}: a0 + G v rel diff {e1:.3e}, K A rel diff {e2:.3e}")
        assert e1 <= 1e-5 and e2 <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("p", [1, 3])
def test_host_layer_mesh_file(tmp_path, p):
    """The operator on a mesh written to and read back from the MFEM v1.0 format
    applies as the operator on the original box mesh (relative L2 <= 1e-12)."""
    print(f"case host mesh file p={p}", flush=True)
    r = subprocess.run([_built(), "meshio", str(tmp_path), "5", str(p)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "meshio tests: ok" in r.stdout
