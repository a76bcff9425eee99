"""The persistent batched-PCG launch (msa_batch_set_persistent): a Jacobi
PCG solve of a whole batch as ONE cooperative launch that iterates on the
device, against the graph-captured iteration blocks on the same inputs.
The per-item work and the summation order are the same, so solutions,
iteration counts and statuses are the same bits; the FP64 solutions also
meet the oracle's skyline direct solve (1e-10)."""
import numpy as np
import pytest

import msolve_amd as msa
from msolve_amd import _lib as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def lognormal(rng, shape, cov=0.3):
    s = np.sqrt(np.log(1 + cov * cov))
    return 2.1e5 * np.exp(-0.5 * s * s + s * rng.standard_normal(shape))


def solve(ctx, meshes, models, E, precision, persistent, unroll=3, refine=None):
    b = msa.Batch(ctx, meshes, [e.shape[0] for e in E], precision, 1)
    b.set_schedule(unroll)
    if refine is not None:
        b.set_refinement(refine, 8, 0.1)
    b.set_persistent(persistent)
    for c, m in enumerate(models):
        b.set_modulus(c, E[c])
        b.set_rhs(c, m.f)
    b.assemble(0.3)
    b.solve(tol=1e-11, max_it=5000, inner_tol=1e-5, graph_iters=8)
    out = [[b.solution(c, s) for s in range(E[c].shape[0])] for c in range(len(models))]
    st = [b.stats(c) for c in range(len(models))]
    info = b.persistent_info()
    b.close()
    return out, st, info


@pytest.mark.parametrize("unroll", [3, 1])
def test_persistent_equals_graph_fp64(ctx, unroll):
    print(f"\n[case] persistent PCG fp64 unroll={unroll}", flush=True)
    models = [msa.cantilever(6), msa.cantilever(8, 5, 4)]
    meshes = [msa.Mesh(ctx, m.xyz, m.conn, m.dof, m.n_free) for m in models]
    rng = np.random.default_rng(31)
    E = [lognormal(rng, (4, models[0].n_elem)), lognormal(rng, (3, models[1].n_elem))]
    ug, sg, ig = solve(ctx, meshes, models, E, L.MSA_FP64, False, unroll)
    up, sp, ip = solve(ctx, meshes, models, E, L.MSA_FP64, True, unroll)
    print(f"[persistent] {ip}; iterations {[s['iters'].tolist() for s in sp]}", flush=True)
    assert ig["launches"] == 0 and ip["launches"] == 1
    assert ip["workgroups"] > 0 and ip["workgroups"] % 8 == 0
    for c, m in enumerate(models):
        assert np.array_equal(sg[c]["iters"], sp[c]["iters"])
        assert np.array_equal(sg[c]["status"], sp[c]["status"])
        assert (sp[c]["status"] == msa.MSA_OK).all()
        for s in range(E[c].shape[0]):
            assert np.array_equal(ug[c][s], up[c][s])
            ke, _ = O.ke_mesh(m.conn, m.xyz, E[c][s], 0.3)
            ri, data = O.skyline_assemble(m.conn, m.dof, m.n_free, ke)
            fac, rc = O.skyline_factorize(ri, data, 1e-15)
            assert rel(up[c][s], O.skyline_solve(ri, fac, m.f)) < 1e-10
    for m in meshes:
        m.close()


def test_persistent_mixed_restart_inner_solves(ctx):
    """Mixed precision with restarted refinement: each inner FP32 solve is
    one persistent launch; the reliable-update scheme keeps the graphs."""
    print("\n[case] persistent PCG mixed restart", flush=True)
    model = msa.cantilever(6)
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    rng = np.random.default_rng(5)
    E = [lognormal(rng, (3, model.n_elem))]
    ug, sg, _ = solve(ctx, [mesh], [model], E, L.MSA_MIXED, False, refine=L.MSA_REFINE_RESTART)
    up, sp, ip = solve(ctx, [mesh], [model], E, L.MSA_MIXED, True, refine=L.MSA_REFINE_RESTART)
    assert ip["launches"] >= 2   # one per outer refinement step
    assert np.array_equal(sg[0]["iters"], sp[0]["iters"])
    for s in range(3):
        assert np.array_equal(ug[0][s], up[0][s])
    _, _, ir = solve(ctx, [mesh], [model], E, L.MSA_MIXED, True, refine=L.MSA_REFINE_RELIABLE)
    assert ir["launches"] == 0
    bt = msa.Batch(ctx, [mesh], [4], L.MSA_FP64, 4)
    with pytest.raises(msa.MsaError):
        bt.set_persistent(True)
    bt.close()
    mesh.close()
