"""Pins the oracle's element results (CalcH8Strains, CalcH8Forces,
CalcH8MLumped, ElasticMaterial3D stresses): affine displacement fields give
exact constant strains on any trilinear hex, rigid motions give none,
internal forces equal Ke u, the lumped mass is rho * volume / 8 per node, and
a solved cantilever's element forces balance the loads and reactions."""
import numpy as np

from oracle import oracle as O

UNIT = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                 [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)


def hexes(rng, n=4):
    yield 0.2 + UNIT * np.array([0.5, 0.25, 1.5])
    for _ in range(n):
        yield UNIT * np.array([1.0, 0.9, 1.1]) + rng.uniform(-0.15, 0.15, (8, 3))


def test_strains_of_affine_fields_are_exact():
    rng = np.random.default_rng(5)
    for x in hexes(rng):
        gm = O.gauss_matrices(x)
        A = rng.standard_normal((3, 3))
        c = rng.standard_normal(3)
        u = (x @ A.T + c).ravel()
        eps = O.h8_strains(gm["B"], u)
        want = [A[0, 0], A[1, 1], A[2, 2], A[0, 1] + A[1, 0], A[1, 2] + A[2, 1], A[2, 0] + A[0, 2]]
        assert np.abs(eps - np.array(want)).max() < 1e-12 * np.abs(A).max()
        # rigid rotation + translation: no strain, no force
        W = np.array([[0, -0.3, 0.2], [0.3, 0, -0.1], [-0.2, 0.1, 0]])
        ur = (x @ W.T + c).ravel()
        assert np.abs(O.h8_strains(gm["B"], ur)).max() < 1e-13


def test_internal_forces_equal_ke_u():
    rng = np.random.default_rng(6)
    for x in hexes(rng):
        gm = O.gauss_matrices(x)
        E, nu = 2.1e5, 0.3
        u = rng.standard_normal(24) * 1e-3
        eps = O.h8_strains(gm["B"], u)
        sig = np.zeros((8, 6))
        for m in range(8):
            sig[m] = O.elastic_D(E, nu) @ eps[m]
        f = O.h8_forces(gm["B"], gm["w"], sig)
        K = O.unpack300(O.h8_k(np.tile(O.elastic_D(E, nu), (8, 1, 1)), gm["B"], gm["w"]))
        assert np.abs(f - K @ u).max() < 1e-11 * np.abs(K @ u).max()


def test_lumped_mass_closed_form():
    x = UNIT * np.array([0.5, 0.25, 2.0])
    gm = O.gauss_matrices(x)
    M = O.h8_mlumped(7.85e-9, gm["w"])
    Md = O.unpack300(M)
    assert np.allclose(np.diag(Md), 7.85e-9 * 0.25 / 8, rtol=1e-13)
    assert np.count_nonzero(Md - np.diag(np.diag(Md))) == 0


def test_cantilever_element_forces_balance():
    cl = O.Cantilever(4)
    rng = np.random.default_rng(8)
    E = 2.1e5 * np.exp(0.2 * rng.standard_normal(cl.n_elem))
    ke, _ = O.ke_mesh(cl.conn, cl.xyz, E, 0.3)
    ri, data = O.skyline_assemble(cl.conn, cl.dof, cl.n, ke)
    fac, rc = O.skyline_factorize(ri, data, 1e-15)
    u = O.skyline_solve(ri, fac, cl.f)
    eps, sig, f = O.element_results(cl, u, E)
    # assemble the element forces over all node DOFs
    N = 3 * cl.n_nodes
    g = np.zeros(N)
    conn = cl.conn.reshape(-1, 8)
    for e in range(cl.n_elem):
        idx = (3 * conn[e][:, None] + np.arange(3)).ravel()
        np.add.at(g, idx, f[e])
    free = cl.dof >= 0
    # free rows: K u = f; constrained rows: reactions, balancing the loads
    assert np.abs(g[free][np.argsort(cl.dof[free])] - cl.f).max() < 1e-9 * np.abs(cl.f).max()
    for d in range(3):
        react = g[~free][np.arange(N)[~free] % 3 == d].sum()
        load = cl.f[np.arange(cl.n) % 3 == d].sum()
        assert abs(react + load) < 1e-9 * 100
    # stresses are D(E_e) strains
    for e in (0, 17, cl.n_elem - 1):
        for m in range(8):
            assert np.allclose(sig[e, m], O.elastic_D(E[e], 0.3) @ eps[e, m], rtol=1e-14, atol=0)
