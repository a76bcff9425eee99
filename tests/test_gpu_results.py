"""GPU parity of the element results (CalcH8Strains / CalcH8Forces /
CalcH8MLumped drop-ins and the batched strains, stresses and internal
forces of solved samples) against the oracle on the same inputs: 1e-12
relative (FP64)."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import msolve_amd as msa
from msolve_amd import _lib as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ke_golden.json")


def p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def test_fortran_result_dropins(ctx):
    lib = L.lib()
    rng = np.random.default_rng(2)
    for case in json.load(open(GOLDEN))["cases"]:
        x = np.array(case["xyz"])
        gm = O.gauss_matrices(x)
        B = np.ascontiguousarray(gm["B"], np.float64).ravel()
        w = np.ascontiguousarray(gm["w"], np.float64)
        it = C.c_int(2)
        u = rng.standard_normal(24)
        eps = np.zeros(48)
        lib.CALCH8STRAINS(C.byref(it), p(B), p(u), p(eps))
        ref = O.h8_strains(gm["B"], u).ravel()
        assert np.abs(eps - ref).max() <= 1e-13 * np.abs(ref).max(), case["name"]
        sig = rng.standard_normal(48)
        f = np.zeros(24)
        lib.CALCH8FORCES(C.byref(it), p(B), p(w), p(sig), p(f))
        reff = O.h8_forces(gm["B"], gm["w"], sig)
        assert np.abs(f - reff).max() <= 1e-13 * np.abs(reff).max(), case["name"]
        M = np.zeros(300)
        rho = C.c_double(7.85e-9)
        lib.CALCH8MLUMPED(C.byref(it), C.byref(rho), p(w), p(M))
        assert np.array_equal(M, O.h8_mlumped(7.85e-9, gm["w"])), case["name"]


@pytest.mark.parametrize("precision", [L.MSA_FP64, L.MSA_MIXED])
def test_batched_element_results(ctx, precision):
    print(f"\n[case] element results precision={precision}", flush=True)
    model = msa.cantilever(6)
    om = O.Cantilever(6)
    rng = np.random.default_rng(12)
    s = np.sqrt(np.log(1.09))
    E = 2.1e5 * np.exp(-0.5 * s * s + s * rng.standard_normal((3, model.n_elem)))
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    b = msa.Batch(ctx, [mesh], [3], precision, 0)
    b.set_modulus(0, E)
    b.set_rhs(0, model.f)
    b.assemble(0.3)
    b.solve(tol=1e-11, max_it=5000, graph_iters=8)
    for smp in range(3):
        u = b.solution(0, smp)
        eps, sig, f = b.element_results(0, smp)
        reps, rsig, rf = O.element_results(om, u, E[smp])
        for got, want in ((eps, reps), (sig, rsig), (f, rf)):
            assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    b.close()
    mesh.close()


def test_skyline_export_matches_reference_assembly(ctx, tmp_path):
    """The GPU-assembled FP64 K exported in the reference's skyline layout
    equals the oracle's skyline assembly; the binary file round-trips."""
    from msolve_amd.matrix_io import read_skyline, write_skyline
    model = msa.cantilever(4)
    om = O.Cantilever(4)
    E = 2.1e5 * np.exp(0.25 * np.random.default_rng(3).standard_normal((2, model.n_elem)))
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    b = msa.Batch(ctx, [mesh], [2], L.MSA_FP64, 0)
    b.set_modulus(0, E)
    b.assemble(0.3)
    for smp in range(2):
        ri, data = b.matrix_skyline(0, smp)
        ke, _ = O.ke_mesh(om.conn, om.xyz, E[smp], 0.3)
        ri_ref, data_ref = O.skyline_assemble(om.conn, om.dof, om.n, ke)
        assert np.array_equal(ri, ri_ref)
        assert np.abs(data - data_ref).max() <= 1e-13 * np.abs(data_ref).max()
        name = str(tmp_path / f"K{smp}.bin")
        write_skyline(name, ri, data, binary=True)
        r2, d2 = read_skyline(name)
        assert np.array_equal(r2, ri) and np.array_equal(d2, data)
    b.close()
    mesh.close()
