"""Pins the oracle's consistent mass and linear Newmark loop (the dynamics
path: consistent M, Rayleigh C, effective stiffness, mass-acceleration load,
Newmark-beta updates).  Pins: closed-form brick mass, total mass = rho * volume
by an independent high-order quadrature, the sinusoidal history load's closed
form and cut-off, and the whole time loop against an independently written
dense predictor-corrector Newmark (algebraically equal, different
arithmetic) with its own dense assembly and load vector."""
import numpy as np
import pytest

from oracle import oracle as O

UNIT = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                 [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)


def test_mass_brick_closed_form():
    h = np.array([0.5, 0.25, 2.0])
    rho = 7.85e-9
    M = O.unpack300(O.h8_mass_csharp(1.5 + UNIT * h, rho))
    sgn = 2 * UNIT - 1
    for a in range(8):
        for b in range(8):
            want = rho * np.prod([h[d] / 3 if sgn[a, d] == sgn[b, d] else h[d] / 6
                                  for d in range(3)])
            for p in range(3):
                for q in range(3):
                    got = M[3 * a + p, 3 * b + q]
                    if p == q:
                        # the reference's 13-digit Gauss abscissa: ~1e-13 relative
                        assert abs(got - want) <= 1e-12 * want
                    else:
                        assert got == 0.0


def hex_volume(x8, n=8):
    """Volume of a trilinear hex by an n-point Gauss-Legendre rule per axis
    (exact for the degree-2-per-axis Jacobian determinant)."""
    g, w = np.polynomial.legendre.leggauss(n)
    sgn = 2 * UNIT - 1
    vol = 0.0
    for i, xi in enumerate(g):
        for j, et in enumerate(g):
            for k, ze in enumerate(g):
                p = np.array([xi, et, ze])
                dN = np.zeros((3, 8))
                for a in range(8):
                    f = 1 + sgn[a] * p
                    dN[0, a] = sgn[a, 0] * f[1] * f[2] / 8
                    dN[1, a] = sgn[a, 1] * f[0] * f[2] / 8
                    dN[2, a] = sgn[a, 2] * f[0] * f[1] / 8
                vol += w[i] * w[j] * w[k] * np.linalg.det(dN @ x8)
    return vol


def test_mass_distorted_total_symmetric_positive():
    rng = np.random.default_rng(11)
    for _ in range(4):
        x = UNIT * np.array([1.0, 0.8, 1.2]) + rng.uniform(-0.15, 0.15, (8, 3))
        rho = 2.5
        M = O.unpack300(O.h8_mass_csharp(x, rho))
        assert np.allclose(M, M.T, rtol=0, atol=0)
        vol = hex_volume(x)
        for p in range(3):
            assert abs(M[p::3, p::3].sum() - rho * vol) <= 1e-12 * rho * vol
        assert np.linalg.eigvalsh(M).min() > 0


def test_sinusoidal_history_load():
    dt, period, tlen = 1e-4, 0.05, 0.05
    steps = int(tlen / dt)
    assert steps == 500
    for i in (0, 1, 7, 125, 499):
        want = 1.0 * np.sin(2 * np.pi / period * dt * i)
        assert O.sinusoidal_acc(i, 1.0, period, dt, tlen) == pytest.approx(want, abs=1e-15)
    assert O.sinusoidal_acc(500, 1.0, period, dt, tlen) == 0.0
    assert O.sinusoidal_acc(125, 2.0, period, dt, tlen) == pytest.approx(2.0, rel=1e-15)


def dense_system(model, E, nu, rho, alphaR, betaR, acc_dir):
    """Independent dense assembly: full (all-DOF) K, M from the element
    matrices, then free-free blocks and the mass-acceleration load vector
    M_full[free, dir] . 1."""
    nn = model.n_nodes
    N = 3 * nn
    Kf = np.zeros((N, N))
    Mf = np.zeros((N, N))
    ke, bad = O.ke_mesh(model.conn, model.xyz, E, nu)
    assert bad == 0
    conn = model.conn.reshape(-1, 8)
    xyz = model.xyz.reshape(-1, 3)
    for e in range(model.n_elem):
        g = (3 * conn[e][:, None] + np.arange(3)).ravel()
        Kf[np.ix_(g, g)] += O.unpack300(ke[e])
        Mf[np.ix_(g, g)] += O.unpack300(O.h8_mass_csharp(xyz[conn[e]], rho))
    free = np.where(model.dof >= 0)[0]
    free = free[np.argsort(model.dof[free])]
    K = Kf[np.ix_(free, free)]
    M = Mf[np.ix_(free, free)]
    C = alphaR * M + betaR * K
    unit = np.zeros(N)
    unit[acc_dir::3] = 1.0
    b = (Mf @ unit)[free]
    return K, M, C, b


def predictor_corrector(K, M, C, b, agfun, beta, gamma, dt, nsteps, probe):
    """Newmark-beta in acceleration form: predict, solve the effective mass
    equation for the new acceleration, correct (starting at rest with
    a0 = M^-1 (F0 - K u0 - C v0) = 0 since F0 = 0)."""
    n = K.shape[0]
    u, v, a = np.zeros(n), np.zeros(n), np.zeros(n)
    A = M + gamma * dt * C + beta * dt * dt * K
    hist = []
    for i in range(nsteps):
        F = b * agfun(i)
        up = u + dt * v + dt * dt / 2 * (1 - 2 * beta) * a
        vp = v + dt * (1 - gamma) * a
        a = np.linalg.solve(A, F - C @ vp - K @ up)
        u = up + beta * dt * dt * a
        v = vp + gamma * dt * a
        hist.append(u[probe])
    return np.array(hist), u, v, a


@pytest.mark.parametrize("alphaR,betaR", [(0.0, 0.0), (30.0, 2e-4)])
def test_newmark_loop_matches_independent_scheme(alphaR, betaR):
    model = O.Cantilever(2, 2, 1, lz=0.5)
    rng = np.random.default_rng(3)
    E = 2.1e5 * np.exp(0.2 * rng.standard_normal(model.n_elem))
    # a density that makes inertia matter at this dt (the history load's
    # period is a few natural periods of the block)
    rho, dt, nsteps, period = 50.0, 2e-3, 60, 0.05
    hist, u, v, a = O.newmark_skyline(model, E, rho=rho, alphaR=alphaR, betaR=betaR, dt=dt,
                                      nsteps=nsteps, acc_dir=2, acc_mag=1.0,
                                      acc_period=period, acc_tlen=nsteps * dt)
    K, M, C, b = dense_system(model, E, 0.3, rho, alphaR, betaR, 2)
    ag = lambda i: O.sinusoidal_acc(i, 1.0, period, dt, nsteps * dt)
    h2, u2, v2, a2 = predictor_corrector(K, M, C, b, ag, 0.25, 0.5, dt, nsteps, model.n - 1)
    scale = np.abs(h2).max()
    assert scale > 0
    assert np.abs(hist - h2).max() <= 1e-9 * scale
    assert np.linalg.norm(u - u2) <= 1e-9 * np.linalg.norm(u2)
    assert np.linalg.norm(v - v2) <= 1e-8 * np.linalg.norm(v2)
    assert np.linalg.norm(a - a2) <= 1e-7 * np.linalg.norm(a2)


def test_newmark_quasi_static_limit():
    """With a tiny density and a slow load, u(t_i) -> K^-1 F(t_i)."""
    model = O.Cantilever(2, 2, 2)
    E = np.full(model.n_elem, 2.1e5)
    f = np.zeros(model.n)
    f[-1] = -1.0
    hist, u, _, _ = O.newmark_skyline(model, E, rho=1e-12, dt=1e-3, nsteps=3, f_static=f,
                                      acc_dir=-1)
    K, _, _, _ = dense_system(model, E, 0.3, 1e-12, 0, 0, 2)
    us = np.linalg.solve(K, f)
    assert np.linalg.norm(u - us) <= 1e-8 * np.linalg.norm(us)


def test_newmark_rejects_unstable_parameters():
    model = O.Cantilever(1)
    with pytest.raises(RuntimeError):
        O.newmark_skyline(model, np.full(1, 2.1e5), nb_alpha=0.5, nb_delta=0.25)
