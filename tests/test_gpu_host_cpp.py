"""GPU: the C++ host layer (include/msolve_amd.hpp) driven end to end by
tests/cpp/host_api_check.cpp, compared with the oracle: SolverPCG (Jacobi
and multigrid), the Monte Carlo analyzer, Newmark, the element provider and
SolverSkyline with zero-energy modes."""
import json
import os
import subprocess

import numpy as np
import pytest

from msolve_amd.stochastic import kl_table, standard_normal
from oracle import oracle as O

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(__file__), "cpp", "_build", "host_api_check")


def direct(cl, E):
    ke, _ = O.ke_mesh(cl.conn, cl.xyz, E, 0.3)
    ri, data = O.skyline_assemble(cl.conn, cl.dof, cl.n, ke)
    fac, rc = O.skyline_factorize(ri, data)
    assert rc == 0
    return O.skyline_solve(ri, fac, cl.f)


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - b) / np.linalg.norm(b)


def test_cpp_host_layer(tmp_path):
    assert os.path.exists(EXE), "build() compiles tests/cpp/host_api_check"
    M = 10
    t = kl_table([1.0, 1.0, 1.0], 0.25, M)
    xi = standard_normal(99, 5, M)
    path = tmp_path / "kl.txt"
    with open(path, "w") as fh:
        fh.write(f"{M} {t['kmax']}\n")
        for arr in (t["w1d"].ravel(), t["a1d"].ravel(), t["evpPos"].ravel(), t["eigs"], xi.ravel()):
            fh.write(" ".join(repr(float(v)) if arr.dtype.kind == "f" else str(int(v)) for v in arr) + "\n")
    print("\n[case] host_api_check", flush=True)
    res = subprocess.run([EXE, str(path)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    out = json.loads(res.stdout)
    c6, c4 = O.Cantilever(6), O.Cantilever(4)
    assert out["status"] == 0 and out["current_iteration"] == out["products"] - 1
    assert rel(out["u_plain"], direct(c6, np.full(c6.n_elem, 2.1e5))) < 1e-10
    assert out["mc_bad"] == 0
    for s in range(3):
        E, _ = O.kl_field(c4.conn, c4.xyz, t, xi[s], c4.domain, 2.1e5, 0.3)
        assert rel(out[f"u4_{s}"], direct(c4, E)) < 1e-10
    for s in range(2):
        E, _ = O.kl_field(c6.conn, c6.xyz, t, xi[3 + s], c6.domain, 2.1e5, 0.3)
        assert rel(out[f"u6_{s}"], direct(c6, E)) < 1e-10
    nsteps = int(30 * 2e-3 / 2e-3)
    h_ref, _, _, _ = O.newmark_skyline(c4, np.full(c4.n_elem, 2.1e5), rho=50.0, dt=2e-3,
                                       nsteps=nsteps, f_static=c4.f, acc_dir=2, acc_mag=1.0,
                                       acc_period=0.05, acc_tlen=30 * 2e-3, probe=c4.n - 1)
    assert len(out["newmark_hist"]) == nsteps
    assert rel(out["newmark_hist"], h_ref) < 1e-9
    c8 = O.Cantilever(8)
    assert out["mg_status"] == 0 and out["mg_products"] < 25
    assert rel(out["u_mg8"], direct(c8, np.full(c8.n_elem, 2.1e5))) < 1e-10
    # SolverSkyline.Factorize(1e-3, zems, zemCols) on the floating 3^3 box
    # against the oracle's restatement (SkylineMatrix2D.Factorize + Solve)
    from msolve_amd.model import box_mesh
    xyz, conn = box_mesh(3, 3, 3)
    dof, nf = O.enumerate_dofs(np.zeros((xyz.shape[0], 3), np.uint8))   # no supports
    ke, _ = O.ke_mesh(conn, xyz, np.full(conn.shape[0], 2.1e5), 0.3)
    ri, data = O.skyline_assemble(conn, dof, nf, ke)
    fac, k, cols, Z = O.skyline_factorize_zem(ri, data, 1e-3, 16)
    assert out["zem_cols"] == cols.tolist() and k == 6 and out["sky_status"] == 0
    for j in range(6):
        assert np.abs(np.array(out[f"zem_{j}"]) - Z[j]).max() <= 1e-8 * np.abs(Z[j]).max()
    f = np.zeros(nf)
    f[nf - 3] = 1.0
    assert rel(out["u_sky"], O.skyline_solve(ri, fac, f)) < 1e-9
    ke, _ = O.ke_mesh(c4.conn[:8], c4.xyz, np.ones(1), 0.3)
    assert np.abs(np.array(out["ke0"]) - ke[0]).max() < 1e-13 * np.abs(ke[0]).max()
