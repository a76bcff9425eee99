"""The stored layout of box grids and the symmetric PCG product on the GPU.

Box grids (every block a point of a box, numbered lexicographically) keep
their block rows in stencil slot order, and one-sample tiles form the PCG
product q = K p from the stored upper triangle, the lower blocks being the
transposes of upper blocks of the rows above (the way SkylineMatrix2D.cs:
253-269 applies its upper profile to both triangles).  Checked here:

* which classes take the layout (msa_batch_layout_info): box grids at
  least 3 points wide in x and y, any lane tile; the symmetric product only
  with one-sample tiles; a perturbed grid, a 2-element-wide cantilever and
  a single element keep the rank order;
* the exported CSR of a stencil-ordered class is the oracle's skyline
  matrix (1e-13), including the partly constrained rows of a fixed face;
* FP64 solves with the symmetric product (one-sample tiles) equal the
  full product's (lane tiles of 4) to 1e-11, and the oracle's direct
  solve to 1e-9, on a box grid and on a mesh with partly fixed nodes;
* MIXED solves with the symmetric product reach the 1e-8 bar against
  the oracle's direct solve with the Jacobi and the multigrid
  preconditioners;
* msa_batch_set_symmetric_product switches a box-grid batch to the full
  image and back (assembly required, same bars, bit-equal on return)."""
from types import SimpleNamespace

import numpy as np
import pytest

import msolve_amd as msa
from msolve_amd import _lib as L
from msolve_amd.matrix_io import skyline_to_dense
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def oracle_K(m, E):
    ke, bad = O.ke_mesh(m.conn, m.xyz, E, 0.3)
    assert bad == 0
    ri, data = O.skyline_assemble(m.conn, m.dof, m.n_free, ke)
    return ri, data


def direct(m, E):
    ri, data = oracle_K(m, E)
    fac, rc = O.skyline_factorize(ri, data.copy(), 1e-15)
    assert rc == 0
    return O.skyline_solve(ri, fac, m.f)


def moduli(rng, ns, ne, cov=0.3):
    s = np.sqrt(np.log(1 + cov * cov))
    return 2.1e5 * np.exp(-0.5 * s * s + s * rng.standard_normal((ns, ne)))


def test_layout_info_selects_box_grids(ctx):
    print("\n[case] layout info", flush=True)
    cases = []
    for n, perturb in ((5, False), (5, True), (2, False), (1, False)):
        m = msa.cantilever(n)
        xyz = m.xyz.copy()
        if perturb:  # move one interior node: the cells are no longer equal
            k = [i for i in range(len(xyz)) if 0 < xyz[i, 0] < 1 and 0 < xyz[i, 1] < 1
                 and 0 < xyz[i, 2] < 1][0]
            xyz[k] += 0.01 / n
        cases.append((n, perturb, msa.Mesh(ctx, xyz, m.conn, m.dof, m.n_free)))
    for n, perturb, mesh in cases:
        for tile in (1, 4):
            b = msa.Batch(ctx, [mesh], [4], L.MSA_FP64, tile)
            li = b.layout_info(0)
            box = n >= 3 and not perturb
            assert li["stencil"] == box, (n, perturb, tile, li)
            assert li["upper_triangle_product"] == (box and tile == 1), (n, perturb, tile, li)
            R = 64 // tile
            nslices = (li["block_rows"] + R - 1) // R
            if box:
                assert li["slots"] == 27 * nslices
            b.close()
    for *_, mesh in cases:
        mesh.close()


@pytest.mark.parametrize("kind", ["cantilever", "partly_fixed"])
def test_stencil_export_is_oracle_matrix(ctx, kind):
    print(f"\n[case] stencil-ordered export ({kind})", flush=True)
    model = model_of(kind, 5)
    rng = np.random.default_rng(11)
    E = moduli(rng, 2, model.conn.shape[0])
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    for precision, bar in ((L.MSA_FP64, 1e-13), (L.MSA_MIXED, 1e-7)):
        b = msa.Batch(ctx, [mesh], [2], precision, 1)
        assert b.layout_info(0)["stencil"]
        b.set_modulus(0, E)
        b.assemble(0.3)
        for s in range(2):
            rp, col, val = b.matrix_csr(0, s)
            Kg = np.zeros((model.n_free, model.n_free))
            for r in range(model.n_free):
                Kg[r, col[rp[r]:rp[r + 1]]] = val[rp[r]:rp[r + 1]]
            ri, data = oracle_K(model, E[s])
            Ko = skyline_to_dense(ri, data)
            assert np.abs(Kg - Ko).max() <= bar * np.abs(Ko).max()
        b.close()
    mesh.close()


def model_of(kind, n):
    """The n^3 cantilever, or the same with its x = 1 face also held in z
    (rows with one or two free components inside the stencil layout)."""
    m = msa.cantilever(n)
    if kind == "cantilever":
        return SimpleNamespace(conn=m.conn, xyz=m.xyz, dof=m.dof, n_free=m.n_free, f=m.f)
    dof = m.dof.copy().reshape(-1, 3)
    face = np.isclose(m.xyz[:, 0], 1.0)
    dof[face, 2] = -1
    k = 0
    for i in range(dof.shape[0]):
        for d in range(3):
            if dof[i, d] >= 0:
                dof[i, d] = k
                k += 1
    f = np.zeros(k)
    for i in np.nonzero(face)[0]:
        f[dof[i, 1]] = -100.0 / face.sum()
    return SimpleNamespace(conn=m.conn, xyz=m.xyz, dof=dof.ravel(), n_free=k, f=f)


@pytest.mark.parametrize("kind", ["cantilever", "partly_fixed"])
def test_symmetric_product_equals_full_product(ctx, kind):
    print(f"\n[case] symmetric vs full product ({kind})", flush=True)
    mm = model_of(kind, 6)
    conn, xyz, dof, nf, f = mm.conn, mm.xyz, mm.dof, mm.n_free, mm.f
    rng = np.random.default_rng(12)
    ns = 3
    E = moduli(rng, ns, conn.shape[0])
    mesh = msa.Mesh(ctx, xyz, conn, dof, nf)
    sols, iters = {}, {}
    for tile in (1, 4):
        b = msa.Batch(ctx, [mesh], [ns], L.MSA_FP64, tile)
        assert b.layout_info(0)["upper_triangle_product"] == (tile == 1)
        b.set_modulus(0, E)
        b.set_rhs(0, f)
        b.assemble(0.3)
        b.solve(tol=1e-13, max_it=20000, graph_iters=8)
        st = b.stats(0)
        assert (st["status"] == msa.MSA_OK).all(), st
        sols[tile] = [b.solution(0, s) for s in range(ns)]
        iters[tile] = st["iters"].copy()
        b.close()
    for s in range(ns):
        d = rel(sols[1][s], sols[4][s])
        print(f"  sample {s}: symmetric vs full {d:.2e}, iterations {iters[1][s]} / {iters[4][s]}")
        assert d < 1e-11
        assert abs(int(iters[1][s]) - int(iters[4][s])) <= 2
        assert rel(sols[1][s], direct(mm, E[s])) < 1e-9
    mesh.close()


@pytest.mark.parametrize("precond", [L.MSA_PRECOND_JACOBI, L.MSA_PRECOND_MULTIGRID])
def test_mixed_symmetric_product_meets_bar(ctx, precond):
    print(f"\n[case] mixed symmetric product precond={precond}", flush=True)
    model = msa.cantilever(8)
    rng = np.random.default_rng(13)
    ns = 3
    E = moduli(rng, ns, model.n_elem)
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    b = msa.Batch(ctx, [mesh], [ns], L.MSA_MIXED, 1)
    assert b.layout_info(0)["upper_triangle_product"]
    if precond != L.MSA_PRECOND_JACOBI:
        b.set_preconditioner(precond)
    b.set_modulus(0, E)
    b.set_rhs(0, model.f)
    b.assemble(0.3)
    b.solve(tol=1e-9, max_it=20000, graph_iters=8)
    st = b.stats(0)
    assert (st["status"] == msa.MSA_OK).all(), st
    for s in range(ns):
        assert rel(b.solution(0, s), direct(model, E[s])) < 1e-8
    b.close()
    mesh.close()


@pytest.mark.parametrize("precision,precond", [(L.MSA_FP64, L.MSA_PRECOND_JACOBI),
                                               (L.MSA_MIXED, L.MSA_PRECOND_MULTIGRID)])
def test_full_storage_switch(ctx, precision, precond):
    """msa_batch_set_symmetric_product(b, 0) stores and streams both
    triangles on a box grid: the batch must be assembled again, its solves
    meet the same bars (FP64: to 1e-11 of the symmetric product's), its
    export is the oracle's matrix, and switching back reproduces the first
    solve bit for bit.  An unchanged setting keeps the batch assembled."""
    print(f"\n[case] full-storage switch precision={precision} precond={precond}", flush=True)
    model = msa.cantilever(6)
    rng = np.random.default_rng(14)
    ns = 3
    E = moduli(rng, ns, model.n_elem)
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    b = msa.Batch(ctx, [mesh], [ns], precision, 1)
    if precond != L.MSA_PRECOND_JACOBI:
        b.set_preconditioner(precond)
    b.set_modulus(0, E)
    b.set_rhs(0, model.f)
    tol = 1e-13 if precision == L.MSA_FP64 else 1e-9
    sols = {}
    for sym in (True, False, True):
        b.set_symmetric_product(sym)
        li = b.layout_info(0)
        assert li["stencil"] and li["upper_triangle_product"] == sym
        if sym is False:
            with pytest.raises(msa.MsaError, match="assembled"):
                b.solve(tol=tol, max_it=20000, graph_iters=8)
        b.assemble(0.3)
        b.set_symmetric_product(sym)  # unchanged: stays assembled
        b.solve(tol=tol, max_it=20000, graph_iters=8)
        st = b.stats(0)
        assert (st["status"] == msa.MSA_OK).all(), st
        u = [b.solution(0, s) for s in range(ns)]
        if sym in sols:
            for s in range(ns):
                assert np.array_equal(u[s], sols[sym][s])
        sols[sym] = u
        if precision == L.MSA_FP64:
            rp, col, val = b.matrix_csr(0, 0)
            Kg = np.zeros((model.n_free, model.n_free))
            for r in range(model.n_free):
                Kg[r, col[rp[r]:rp[r + 1]]] = val[rp[r]:rp[r + 1]]
            Ko = skyline_to_dense(*oracle_K(model, E[0]))
            assert np.abs(Kg - Ko).max() <= 1e-13 * np.abs(Ko).max()
    for s in range(ns):
        ud = direct(model, E[s])
        bar = 1e-9 if precision == L.MSA_FP64 else 1e-8
        assert rel(sols[True][s], ud) < bar and rel(sols[False][s], ud) < bar
        if precision == L.MSA_FP64:
            assert rel(sols[True][s], sols[False][s]) < 1e-11
    b.close()
    mesh.close()


@pytest.mark.parametrize("sizes", [(8,), (8, 12), ("partly_fixed", 8)])
def test_matrix_free_pcg_product(ctx, sizes):
    """MIXED multigrid batches of box grids take the matrix-free PCG product
    (msa_batch_product_form 2) by default, on the classes the multigrid
    coarsens (those hold its grid tables): the FP32 operator of the finest
    smoothing products, with the FP64 residual replacements fixing the
    accuracy.  Its solves meet the 1e-8 bar against the oracle's direct
    solve, as the stored upper-triangle product's do (same iteration counts
    within 2); a Jacobi batch keeps the stored product."""
    print(f"\n[case] matrix-free PCG product sizes={sizes}", flush=True)
    if sizes[0] == "partly_fixed":  # a box grid with partly constrained nodes
        models = [model_of("partly_fixed", sizes[1])]
        models[0].n_elem = models[0].conn.shape[0]
    else:
        models = [msa.cantilever(n) for n in sizes]
    meshes = [msa.Mesh(ctx, m.xyz, m.conn, m.dof, m.n_free) for m in models]
    rng = np.random.default_rng(15)
    ns = 3
    E = [moduli(rng, ns, m.n_elem) for m in models]
    sols, iters = {}, {}
    for mf in (True, False):
        b = msa.Batch(ctx, meshes, [ns] * len(models), L.MSA_MIXED, 1)
        b.set_preconditioner(L.MSA_PRECOND_MULTIGRID)
        b.set_matrix_free_product(mf)
        for c, m in enumerate(models):
            b.set_modulus(c, E[c])
            b.set_rhs(c, m.f)
        b.assemble(0.3)
        b.solve(tol=1e-9, max_it=20000, graph_iters=8)
        for c in range(len(models)):
            assert b.product_form(c) == ("matrix_free" if mf else "upper_triangle")
            st = b.stats(c)
            assert (st["status"] == msa.MSA_OK).all(), st
        sols[mf] = [[b.solution(c, s) for s in range(ns)] for c in range(len(models))]
        iters[mf] = [b.stats(c)["iters"].copy() for c in range(len(models))]
        b.close()
    for c, m in enumerate(models):
        for s in range(ns):
            ud = direct(m, E[c][s])
            print(f"  class {c} sample {s}: iterations {iters[True][c][s]} / {iters[False][c][s]}")
            assert rel(sols[True][c][s], ud) < 1e-8
            assert rel(sols[False][c][s], ud) < 1e-8
            assert abs(int(iters[True][c][s]) - int(iters[False][c][s])) <= 2
    b = msa.Batch(ctx, meshes[:1], [2], L.MSA_MIXED, 1)
    b.set_modulus(0, E[0][:2])
    b.set_rhs(0, models[0].f)
    b.assemble(0.3)
    b.solve(tol=1e-9, max_it=20000, graph_iters=8)
    assert b.product_form(0) == "upper_triangle"
    b.close()
    for mesh in meshes:
        mesh.close()


def test_matrix_free_product_switch_on_one_batch(ctx):
    """msa_batch_set_matrix_free_product on one batch between solves: the
    next solve rebuilds the product's work items and captures its graphs
    again; switching back reproduces the first solve bit for bit, and both
    forms meet the 1e-8 bar."""
    print("\n[case] matrix-free product switch on one batch", flush=True)
    model = msa.cantilever(8)
    rng = np.random.default_rng(16)
    ns = 3
    E = moduli(rng, ns, model.n_elem)
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    b = msa.Batch(ctx, [mesh], [ns], L.MSA_MIXED, 1)
    b.set_preconditioner(L.MSA_PRECOND_MULTIGRID)
    b.set_modulus(0, E)
    b.set_rhs(0, model.f)
    b.assemble(0.3)
    sols = []
    for mf in (True, False, True):
        b.set_matrix_free_product(mf)
        b.solve(tol=1e-9, max_it=20000, graph_iters=8)
        assert b.product_form(0) == ("matrix_free" if mf else "upper_triangle")
        assert (b.stats(0)["status"] == msa.MSA_OK).all()
        sols.append([b.solution(0, s) for s in range(ns)])
    for s in range(ns):
        assert np.array_equal(sols[0][s], sols[2][s])
        ud = direct(model, E[s])
        assert rel(sols[0][s], ud) < 1e-8 and rel(sols[1][s], ud) < 1e-8
    b.close()
    mesh.close()
