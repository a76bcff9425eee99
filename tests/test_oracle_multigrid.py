"""Pins of the multigrid oracle (oracle/multigrid.py) by closed forms and
invariants: the Galerkin identity of nested trilinear spaces, the Chebyshev
polynomial of the smoother, the Gershgorin bound, symmetry and positive
definiteness of the cycle, convergence of the preconditioned CG to the
direct solution."""
import numpy as np
import pytest
import scipy.sparse.linalg as spl

from oracle import multigrid as MG
from oracle import oracle as O


def level0(nex, E=None, warp=False, seed=0):
    m = O.Cantilever(nex)
    xyz = m.xyz.reshape(-1, 3).copy()
    if warp:  # graded tensor grid: x -> x^1.5, z -> (z + z^2)/2
        xyz[:, 0] = xyz[:, 0] ** 1.5
        xyz[:, 2] = 0.5 * (xyz[:, 2] + xyz[:, 2] ** 2)
    ne = m.conn.size // 8
    if E is None:
        E = 2.1e5 * np.exp(0.3 * np.random.default_rng(seed).standard_normal(ne))
    return m, MG.Level(xyz, m.conn, m.dof, m.n, np.broadcast_to(E, (ne,)).copy())


@pytest.mark.parametrize("warp", [False, True])
def test_rediscretized_coarse_matrix_is_galerkin_for_one_modulus(warp):
    """Nested trilinear spaces: with E constant, the coarse-mesh stiffness
    equals P^T K P exactly (up to rounding)."""
    _, L = level0(8, E=2.1e5, warp=warp)
    coarse, P = MG.coarsen(L)
    G = (P.T @ L.K @ P).toarray()
    Kc = coarse.K.toarray()
    assert coarse.n == 3 * 4 * 5 * 5
    assert np.abs(G - Kc).max() < 1e-11 * np.abs(Kc).max()
    # P reproduces the coarse coordinates' linear functions at free nodes
    assert np.allclose(P.sum(axis=1).A1[P.sum(axis=1).A1 > 0.999], 1.0)


def test_galerkin_identity_fails_for_wrong_weights():
    """A plausible mistake (midpoint weights on a graded grid) breaks it."""
    _, L = level0(8, E=2.1e5, warp=True)
    coarse, P = MG.coarsen(L)
    # every fractional weight snapped to a power of 1/2 (the uniform-grid values)
    wrong = P.copy()
    frac = (wrong.data > 0.01) & (wrong.data < 0.99)
    wrong.data[frac] = 0.5 ** np.round(np.log2(wrong.data[frac]))
    G = (wrong.T @ L.K @ wrong).toarray()
    assert np.abs(G - coarse.K.toarray()).max() > 1e-6 * np.abs(coarse.K.toarray()).max()


def test_chebyshev_smoother_is_scaled_T2():
    """From zero, two steps leave the error (I - S K) e = p(D^-1 K) e with
    p(l) = T2((theta - l)/delta) / T2(theta/delta)."""
    _, L = level0(4, seed=3)
    ratio = 12.0
    hi, lo = L.lam, L.lam / ratio
    theta, delta = (hi + lo) / 2, (hi - lo) / 2
    rng = np.random.default_rng(1)
    e = rng.standard_normal(L.n)
    Ke = L.K @ e
    err = e - MG.smooth(L, Ke, None, ratio)
    A = lambda v: L.dinv * (L.K @ v)
    u = (theta * e - A(e)) / delta
    t2 = 2 * (theta * u - A(u)) / delta - e      # 2 X^2 e - e, X = (theta - A)/delta
    sig = theta / delta
    ref = t2 / (2 * sig * sig - 1)
    assert np.linalg.norm(err - ref) < 1e-12 * np.linalg.norm(e)


def test_gershgorin_bounds_the_spectrum():
    _, L = level0(4, seed=5)
    Kd = L.K.toarray()
    Dh = np.sqrt(L.dinv)
    ev = np.linalg.eigvalsh(Dh[:, None] * Kd * Dh[None, :])
    assert ev.max() <= L.lam * (1 + 1e-12)
    assert L.lam < 2.5 * ev.max()


def test_vcycle_symmetric_positive_definite():
    m, L0 = level0(8, seed=7)
    levels, Ps = MG.hierarchy(L0.xyz, L0.conn, L0.dof, L0.n, L0.E)
    assert [L.n for L in levels] == [m.n, 300, 54]
    rng = np.random.default_rng(2)
    x, y = rng.standard_normal(m.n), rng.standard_normal(m.n)
    Bx, By = MG.vcycle(levels, Ps, x), MG.vcycle(levels, Ps, y)
    assert abs(x @ By - y @ Bx) < 1e-12 * abs(x @ Bx)
    for _ in range(3):
        v = rng.standard_normal(m.n)
        assert v @ MG.vcycle(levels, Ps, v) > 0


@pytest.mark.parametrize("nex,nlev", [(8, 2), (10, 1), (12, 2)])
def test_mg_pcg_converges_to_direct_solution(nex, nlev):
    m, L0 = level0(nex, seed=nex)
    levels, Ps = MG.hierarchy(L0.xyz, L0.conn, L0.dof, L0.n, L0.E)
    assert len(Ps) == nlev
    x, it = MG.mg_pcg(levels, Ps, m.f, tol=1e-11)
    xd = spl.spsolve(L0.K.tocsc(), m.f)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-9
    assert it < 25


def test_coarsening_refuses_odd_and_thin_grids():
    for nex in (9, 2):
        _, L = level0(nex, E=1.0)
        assert MG.coarsen(L) is None
