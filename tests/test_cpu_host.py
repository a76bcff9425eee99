"""CPU checks of the host-side product code: the cantilever model builder
against the oracle's mesh/numbering/loads, the KL table, the reference's KL
and randoms file layouts, and the C++ host layer's build."""
import os

import numpy as np

import msolve_amd as msa
from msolve_amd.stochastic import (kl_table, lognormal_sigma, read_kl_file, read_randoms_file,
                                   standard_normal, write_kl_file, write_randoms_file)
from oracle import oracle as O


def test_cantilever_matches_oracle_model():
    for n in (3, 7):
        m = msa.cantilever(n)
        o = O.Cantilever(n)
        assert np.array_equal(m.xyz.ravel(), o.xyz)
        assert np.array_equal(m.conn.ravel(), o.conn)
        assert np.array_equal(m.dof, o.dof) and m.n_free == o.n
        assert np.array_equal(m.f, o.f)


def test_lognormal_sigma():
    assert abs(lognormal_sigma(0.2) - 0.198042) < 1e-6
    assert abs(lognormal_sigma(0.3) - 0.293560) < 1e-6
    assert abs(lognormal_sigma(0.5) - np.sqrt(np.log(1.25))) < 1e-15


def test_kl_file_round_trip(tmp_path):
    t = kl_table([1.0, 1.0, 1.0], 0.25, 12)
    p = tmp_path / "kl.txt"
    write_kl_file(str(p), t, 1.0, 0.5)
    lines = open(p).read().splitlines()
    assert len(lines) == 5 * 12 + 2              # (3 axes + pos + eig) * M + mean + cov
    r = read_kl_file(str(p), 12)
    for k in ("w1d", "lam1d", "a1d", "eigs"):
        assert np.array_equal(r[k], t[k])
    assert np.array_equal(r["evpPos"], t["evpPos"])


def test_randoms_file_layout(tmp_path):
    xi = standard_normal(5, 7, 3)
    p = tmp_path / "randoms3.txt"
    write_randoms_file(str(p), xi)
    vals = np.loadtxt(p)
    assert vals.size == 21 and vals[0] == xi[0, 0] and vals[7] == xi[0, 1]   # term-major
    back = read_randoms_file(str(p), 3, n_file=7)
    assert np.array_equal(back, xi)


def test_standard_normal_is_seeded():
    a = standard_normal(20261202, 4, 40)
    b = standard_normal(20261202, 4, 40)
    assert np.array_equal(a, b) and a.shape == (4, 40)
    assert abs(a.mean()) < 0.3 and 0.7 < a.std() < 1.3


def test_kl_table_tie_order_and_limits():
    t = kl_table([1.0, 1.0, 1.0], 0.25, 40)
    assert (np.diff(t["eigs"]) <= 1e-15 * t["eigs"][0]).all()     # non-increasing
    assert t["evpPos"].min() >= 1 and t["evpPos"].max() <= t["kmax"]
    assert tuple(t["evpPos"][0]) == (1, 1, 1)
    # the three permutations of (2,1,1) follow, ordered by index
    assert [tuple(v) for v in t["evpPos"][1:4]] == [(1, 1, 2), (1, 2, 1), (2, 1, 1)]


def test_cpp_host_layer_compiles():
    from msolve_amd.build import build_host_check
    exe = build_host_check(verbose=False)
    assert os.path.exists(exe)


def test_sinusoidal_history_matches_oracle():
    h = msa.sinusoidal_history(0.05, 1e-4, 0.05, 1.0, 503)
    ref = np.array([O.sinusoidal_acc(i, 1.0, 0.05, 1e-4, 0.05) for i in range(503)])
    assert np.array_equal(h, ref)
    assert h[500:].tolist() == [0.0, 0.0, 0.0]


def test_newmark_analyzer_step_count():
    an = msa.NewmarkDynamicAnalyzer(None, [], [], timeStep=1e-4, totalTime=0.05)
    assert an.steps == 500   # (int)(0.05 / 1e-4), NewmarkDynamicAnalyzer.cs:284
