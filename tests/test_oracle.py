"""Pins the CPU oracle (oracle/msolve_oracle.c) before it is trusted as the
checker.  The reference holds no tests or fixtures for this path, so the pins
are: the golden Ke fixtures (sympy-exact bricks, 50-digit distorted hexes,
tests/golden/make_ke_golden.py), closed forms, invariants, brute force on
small meshes, and the integral equation the KL eigenpairs must satisfy."""
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ke_golden.json")
CASES = json.load(open(GOLDEN))["cases"]


def ke_oracle(xyz8, E, nu):
    gm = O.gauss_matrices(xyz8)
    return O.h8_k(np.tile(O.elastic_D(E, nu), (8, 1, 1)), gm["B"], gm["w"])


# ----------------------------------------------------------------- element
@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_ke_matches_golden(case):
    x = np.array(case["xyz"])
    gold = np.array(case["ke"])
    # exact bricks differ only by the reference's 13-digit abscissa (~4.5e-14)
    bar = 1e-13 if case["name"].startswith("exact") else 2e-15
    for K in (ke_oracle(x, case["E"], case["nu"]), O.ke_element_csharp(x, case["E"], case["nu"])):
        assert np.abs(K - gold).max() / np.abs(gold).max() < bar


def test_gauss_matrices_closed_form_brick():
    h = np.array([0.125, 0.25, 0.375])
    base = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                     [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    x = 0.3 + base * h
    gm = O.gauss_matrices(x)
    det = np.prod(h) / 8
    assert np.allclose(gm["detJ"], det, rtol=1e-14)
    assert np.allclose(gm["w"], det, rtol=1e-14)          # all 2-point weights are 1
    for M in range(8):
        # Fortran J(i,j) = dx_j/dnat_i, stored column-major
        Jf = gm["J"][M].reshape(3, 3).T
        assert np.allclose(Jf, np.diag(h / 2), atol=1e-15)
        assert abs(gm["S"][M].sum() - 1) < 1e-15        # partition of unity
        assert np.allclose(gm["DS"][M].reshape(3, 8).sum(axis=1), 0, atol=1e-15)
        B = gm["B"][M].T                                  # 6 x 24
        for d in range(3):                                # rigid translations
            t = np.zeros(24)
            t[d::3] = 1
            assert np.abs(B @ t).max() < 1e-12
    # Gauss points: xi outer, zeta inner, abscissa 0.5773502691896
    nat = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                    [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)
    g = 0.5773502691896
    pts = [(a, b, c) for a in (-g, g) for b in (-g, g) for c in (-g, g)]
    for M, (xi, eta, zeta) in enumerate(pts):
        N = (1 + nat[:, 0] * xi) * (1 + nat[:, 1] * eta) * (1 + nat[:, 2] * zeta) / 8
        assert np.allclose(gm["S"][M], N, atol=1e-15)


@pytest.mark.parametrize("case", CASES[3:], ids=[c["name"] for c in CASES[3:]])
def test_ke_invariants(case):
    x = np.array(case["xyz"])
    K = O.unpack300(ke_oracle(x, case["E"], case["nu"]))
    assert np.abs(K - K.T).max() == 0
    ev = np.linalg.eigvalsh(K)
    assert np.abs(ev[:6]).max() < 1e-12 * ev[-1]          # 6 rigid-body modes
    assert ev[6] > 1e-6 * ev[-1]                          # positive definite otherwise
    rigid = []
    for d in range(3):
        t = np.zeros(24)
        t[d::3] = 1
        rigid.append(t)
    for (i, j) in ((0, 1), (1, 2), (2, 0)):               # rotations
        r = np.zeros(24)
        r[i::3] = -x[:, j]
        r[j::3] = x[:, i]
        rigid.append(r)
    for r in rigid:
        assert np.abs(K @ r).max() < 1e-12 * np.abs(K).max() * np.abs(r).max()
    # linear in E, scales with length (3-D elasticity: K(sX) = s K(X))
    K2 = O.unpack300(ke_oracle(x, 3 * case["E"], case["nu"]))
    assert np.abs(K2 - 3 * K).max() <= 1e-14 * np.abs(K2).max()
    K3 = O.unpack300(ke_oracle(2.5 * x, case["E"], case["nu"]))
    assert np.allclose(K3, 2.5 * K, rtol=1e-12, atol=1e-12 * np.abs(K).max())


def test_elastic_d_closed_form():
    E, nu = 2.1e5, 0.3
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    D = O.elastic_D(E, nu)
    ref = np.zeros((6, 6))
    ref[:3, :3] = lam
    ref[np.arange(3), np.arange(3)] = lam + 2 * mu
    ref[np.arange(3, 6), np.arange(3, 6)] = mu
    assert np.allclose(D, ref, rtol=1e-15)


# ----------------------------------------------------------------- model
def test_mesh_numbering_and_loads():
    cl = O.Cantilever(3, 2, 4)
    nX, nY, nZ = 4, 3, 5
    xyz = cl.xyz.reshape(-1, 3)
    assert xyz.shape[0] == nX * nY * nZ
    # x fastest, then y, then z
    assert np.allclose(xyz[:nX, 0], np.arange(nX) / 3, atol=1e-15)
    assert np.allclose(xyz[::nX][:nY, 1], np.arange(nY) / 2, atol=1e-15)
    conn = cl.conn.reshape(-1, 8)
    assert conn.shape[0] == 24
    e = conn[0]
    # bottom face counter-clockwise then top face
    assert list(e) == [0, 1, 1 + nX, nX, nX * nY, nX * nY + 1, nX * nY + nX + 1, nX * nY + nX]
    fixed = xyz[:, 0] == 0
    assert cl.n == 3 * (~fixed).sum()
    assert (cl.dof.reshape(-1, 3)[fixed] == -1).all()
    assert abs(cl.f.sum() + 100) < 1e-12
    assert np.count_nonzero(cl.f) == nY * nZ
    # every element positively oriented
    for k in range(conn.shape[0]):
        assert (O.gauss_matrices(xyz[conn[k]])["detJ"] > 0).all()


def test_product_model_matches_oracle_mesh():
    import msolve_amd.model as PM
    for n in ((3, 3, 3), (5, 2, 4)):
        cl = O.Cantilever(*n)
        xyz, conn = PM.box_mesh(*n)
        assert np.array_equal(xyz.ravel(), cl.xyz)      # same accumulation, bitwise
        assert np.array_equal(conn.ravel(), cl.conn)


# ----------------------------------------------------------------- skyline
def small_system(n=3, seed=0, cov=0.3):
    cl = O.Cantilever(n)
    rng = np.random.default_rng(seed)
    E = 2.1e5 * np.exp(cov * rng.standard_normal(cl.n_elem))
    ke, bad = O.ke_mesh(cl.conn, cl.xyz, E, 0.3)
    assert bad == 0
    return cl, E, ke


def test_skyline_profile_bruteforce():
    cl, _, ke = small_system(3)
    ri, data = O.skyline_assemble(cl.conn, cl.dof, cl.n, ke)
    conn = cl.conn.reshape(-1, 8)
    low = np.arange(cl.n)
    for e in conn:
        ds = [d for d in cl.dof.reshape(-1, 3)[e].ravel() if d >= 0]
        for d in ds:
            low[d] = min(low[d], min(ds))
    heights = np.arange(cl.n) - low
    assert ri[0] == 0 and ri[1] == 1
    assert np.array_equal(np.diff(ri)[1:], heights[1:] + 1)


def test_skyline_assembly_equals_csr_and_coo():
    import scipy.sparse as sp
    cl, _, ke = small_system(3, 1)
    ri, data = O.skyline_assemble(cl.conn, cl.dof, cl.n, ke)
    Ks = O.skyline_to_dense(ri, data)
    rp, col, val = O.csr_assemble(cl.conn, cl.dof, cl.n, ke)
    Kc = sp.csr_matrix((val, col, rp), shape=(cl.n, cl.n)).toarray()
    assert np.array_equal(Ks, Kc)                          # same summation order
    rows, cols, vals = [], [], []
    conn = cl.conn.reshape(-1, 8)
    for e in range(conn.shape[0]):
        Ke = O.unpack300(ke[e])
        g = cl.dof.reshape(-1, 3)[conn[e]].ravel()
        m = g >= 0
        I, J = np.meshgrid(g[m], g[m], indexing="ij")
        rows.append(I.ravel()); cols.append(J.ravel()); vals.append(Ke[np.ix_(m, m)].ravel())
    Kcoo = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(cl.n, cl.n)).toarray()
    assert np.abs(Kcoo - Ks).max() <= 1e-14 * np.abs(Ks).max()
    assert np.abs(Ks - Ks.T).max() == 0
    # row sums vanish before constraints: full (unconstrained) assembly
    dof_all = np.arange(3 * cl.n_nodes, dtype=np.int32)
    rpa, cola, vala = O.csr_assemble(cl.conn, dof_all, 3 * cl.n_nodes, ke)
    Kfull = sp.csr_matrix((vala, cola, rpa)).toarray()
    for d in range(3):
        t = np.zeros(3 * cl.n_nodes)
        t[d::3] = 1
        assert np.abs(Kfull @ t).max() < 1e-9 * np.abs(Kfull).max()


def test_skyline_multiply_factorize_solve():
    cl, _, ke = small_system(4, 2)
    ri, data = O.skyline_assemble(cl.conn, cl.dof, cl.n, ke)
    K = O.skyline_to_dense(ri, data)
    x = np.random.default_rng(3).standard_normal(cl.n)
    assert np.abs(O.skyline_multiply(ri, data, x) - K @ x).max() < 1e-12 * np.abs(K @ x).max()
    fac, rc = O.skyline_factorize(ri, data, 1e-15)
    assert rc == 0
    u = O.skyline_solve(ri, fac, cl.f)
    uref = np.linalg.solve(K, cl.f)
    assert np.linalg.norm(u - uref) / np.linalg.norm(uref) < 1e-11
    assert np.linalg.norm(K @ u - cl.f) / np.linalg.norm(cl.f) < 1e-12
    # LDL^T: L D L^T reproduces K (unit lower L in the profile)
    n = cl.n
    F = O.skyline_to_dense(ri, fac)
    Lm = np.tril(F, -1) + np.eye(n)
    Dm = np.diag(np.diag(F))
    assert np.abs(Lm @ Dm @ Lm.T - K).max() < 1e-10 * np.abs(K).max()


def test_skyline_singular_detected():
    cl, _, ke = small_system(2, 4)
    dof_all = np.arange(3 * cl.n_nodes, dtype=np.int32)   # no constraints
    ri, data = O.skyline_assemble(cl.conn, dof_all, 3 * cl.n_nodes, ke)
    _, rc = O.skyline_factorize(ri, data, 1e-6)
    assert rc == 6                                          # six zero-energy modes


# ----------------------------------------------------------------- PCG
def test_pcg_reaches_direct_solution_and_true_residual():
    cl, _, ke = small_system(4, 5)
    ri, data = O.skyline_assemble(cl.conn, cl.dof, cl.n, ke)
    fac, _ = O.skyline_factorize(ri, data)
    ud = O.skyline_solve(ri, fac, cl.f)
    K = O.skyline_to_dense(ri, data)
    rp, col, val = O.csr_assemble(cl.conn, cl.dof, cl.n, ke)
    for tol in (1e-8, 1e-12):
        r1 = O.pcg_skyline(ri, data, cl.f, O.PRECOND_JACOBI, tol=tol, history=True)
        r2 = O.pcg_csr(rp, col, val, cl.f, O.PRECOND_JACOBI, tol=tol, history=True)
        for r in (r1, r2):
            assert r["converged"]
            true = np.linalg.norm(cl.f - K @ r["x"]) / np.linalg.norm(cl.f)
            assert true < 2 * tol and abs(true - r["history"][-1]) < 1e-3 * tol + 1e-14
        assert r1["iters"] == r2["iters"]
    r = O.pcg_csr(rp, col, val, cl.f, O.PRECOND_JACOBI, tol=1e-14)
    assert np.linalg.norm(r["x"] - ud) / np.linalg.norm(ud) < 1e-11
    # exact preconditioner of the same matrix: one product (MCPCGDirect with K0 = K)
    r = O.pcg_skyline(ri, data, cl.f, O.PRECOND_SKYLINE, ri, fac, tol=1e-10)
    assert r["iters"] == 1
    # unpreconditioned CG needs more products than Jacobi
    r0 = O.pcg_csr(rp, col, val, cl.f, O.PRECOND_NONE, tol=1e-8)
    rj = O.pcg_csr(rp, col, val, cl.f, O.PRECOND_JACOBI, tol=1e-8)
    assert r0["converged"] and r0["iters"] > rj["iters"]


def test_pcg_not_converged_flag():
    cl, _, ke = small_system(3, 6)
    rp, col, val = O.csr_assemble(cl.conn, cl.dof, cl.n, ke)
    r = O.pcg_csr(rp, col, val, cl.f, O.PRECOND_JACOBI, tol=1e-14, max_it=3)
    assert not r["converged"] and r["iters"] == 3


# ----------------------------------------------------------------- KL
def _phi_1d(w, a, idx, x):
    return a * (np.sin(w * x) if idx % 2 == 0 else np.cos(w * x))


def test_kl_eigenpairs_solve_integral_equation():
    L, b, kmax = 1.0, 0.25, 12
    tab = O.kl_table([L, L, L], b, 8, kmax)
    xg, wg = np.polynomial.legendre.leggauss(400)
    half = L / 2
    for k in range(kmax):
        w, lam, a = tab["w1d"][0][k], tab["lam1d"][0][k], tab["a1d"][0][k]
        idx = k + 1
        # split the kernel's kink: integrate on [-h, x] and [x, h]
        for x in (-0.37, 0.0, 0.21, 0.49):
            tot = 0.0
            for lo, hi in ((-half, x), (x, half)):
                y = 0.5 * (hi - lo) * xg + 0.5 * (hi + lo)
                tot += 0.5 * (hi - lo) * np.sum(wg * np.exp(-np.abs(x - y) / b) * _phi_1d(w, a, idx, y))
            assert abs(tot - lam * _phi_1d(w, a, idx, x)) < 1e-10
    # orthonormality
    y = half * xg
    Phi = np.array([_phi_1d(tab["w1d"][0][k], tab["a1d"][0][k], k + 1, y) for k in range(kmax)])
    G = (Phi * (half * wg)) @ Phi.T
    assert np.abs(G - np.eye(kmax)).max() < 1e-12
    # eigenvalues decrease and sum towards the variance times the length
    assert (np.diff(tab["lam1d"][0]) < 0).all()
    assert tab["lam1d"][0].sum() < L


def test_product_kl_table_matches_oracle_table():
    from msolve_amd.stochastic import kl_table
    a = kl_table([1.0, 1.0, 1.0], 0.25, 40)
    b = O.kl_table([1.0, 1.0, 1.0], 0.25, 40)
    for k in ("w1d", "lam1d", "a1d", "eigs"):
        assert np.allclose(a[k], b[k], rtol=1e-13, atol=0), k
    assert np.array_equal(a["evpPos"], b["evpPos"])


def test_kl_field_pins():
    cl = O.Cantilever(4)
    tab = O.kl_table([1.0, 1.0, 1.0], 0.25, 10)
    meanE, cov = 2.1e5, 0.3
    s = math.sqrt(math.log(cov * cov + 1))
    E0, bad = O.kl_field(cl.conn, cl.xyz, tab, np.zeros(10), cl.domain, meanE, cov)
    assert bad == 0 and np.allclose(E0, meanE * math.exp(-0.5 * s * s), rtol=1e-15)
    conn = cl.conn.reshape(-1, 8)
    xyz = cl.xyz.reshape(-1, 3)
    cent = np.array([xyz[c].mean(axis=0) for c in conn])   # brick centroid
    for i in (0, 3, 9):
        xi = np.zeros(10)
        xi[i] = 1.0
        E1, _ = O.kl_field(cl.conn, cl.xyz, tab, xi, cl.domain, meanE, cov)
        phi = s * math.sqrt(tab["eigs"][i]) * np.ones(len(cent))
        for d in range(3):
            idx = tab["evpPos"][i][d]
            phi *= _phi_1d(tab["w1d"][d][idx - 1], tab["a1d"][d][idx - 1], idx, cent[:, d] - 0.5)
        assert np.allclose(np.log(E1 / E0), phi, rtol=1e-12, atol=1e-14)
    # lognormal mean with the truncated variance: E[E] = meanE exp(-s^2/2 + v/2)
    rng = np.random.default_rng(0)
    xs = rng.standard_normal((20000, 10))
    Es = np.array([O.kl_field(cl.conn[:8], cl.xyz, tab, x, cl.domain, meanE, cov)[0][0] for x in xs[:4000]])
    one = np.zeros(10)
    v = 0.0
    for i in range(10):
        one[:] = 0
        one[i] = 1
        v += np.log(O.kl_field(cl.conn[:8], cl.xyz, tab, one, cl.domain, meanE, cov)[0][0] / E0[0]) ** 2
    expect = meanE * math.exp(-0.5 * s * s + 0.5 * v)
    assert abs(Es.mean() / expect - 1) < 4 * math.sqrt(math.exp(v) - 1) / math.sqrt(len(Es)) + 1e-3


# ----------------------------------------------------------------- MC loop
def test_mc_reference_loop_solutions():
    tab = O.kl_table([1.0, 1.0, 1.0], 0.25, 6)
    xi = np.random.default_rng(1).standard_normal((3, 6))
    out = O.mc_reference(3, tab, xi, cov=0.2, tol=1e-12)
    cl = O.Cantilever(3)
    assert out["iters"][0] == 1            # first sample preconditioned by its own factor
    for s in range(3):
        E, _ = O.kl_field(cl.conn, cl.xyz, tab, xi[s], cl.domain, 2.1e5, 0.2)
        ke, _ = O.ke_mesh(cl.conn, cl.xyz, E, 0.3)
        ri, data = O.skyline_assemble(cl.conn, cl.dof, cl.n, ke)
        fac, _ = O.skyline_factorize(ri, data)
        ud = O.skyline_solve(ri, fac, cl.f)
        assert np.linalg.norm(out["u"][s] - ud) / np.linalg.norm(ud) < 1e-10
    assert all(v >= 0 for v in out["times"].values())
