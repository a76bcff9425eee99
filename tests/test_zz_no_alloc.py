"""msa_ke_packed_noalloc against msa_ke_packed, both called through the C ABI.

The no-allocation entry point stages into one fixed device array per GPU
instead of hipMalloc'ed buffers; it runs the same kernel, so its packed
element stiffness and per-element status must equal msa_ke_packed's bit for
bit.  Two calls in a row (the array keeps the first call's contents), the
second with fewer elements, and a call beyond the limit that must be refused
before any device work.
"""
import ctypes as C

import numpy as np
import pytest

from msolve_amd import _lib as L

LIMIT = 4096  # elements per call of msa_ke_packed_noalloc


def _hexes(n, seed):
    """n distorted hexahedra (corner perturbations of 15% of the edge), a few
    of them inverted so that the status channel carries MSA_NEG_DETJ."""
    rng = np.random.default_rng(seed)
    unit = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                     [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], np.float64)
    h = rng.uniform(0.01, 0.2, (n, 1, 3))
    x = unit[None] * h + rng.uniform(-0.15, 0.15, (n, 8, 3)) * h + rng.uniform(-1, 1, (n, 1, 3))
    bad = rng.choice(n, size=max(1, n // 500), replace=False)
    x[bad] = x[bad][:, [4, 5, 6, 7, 0, 1, 2, 3]]  # top and bottom faces swapped: detJ < 0
    E = rng.uniform(1.0e5, 3.0e5, n)
    return np.ascontiguousarray(x), E


def _call(fn, ctx, x, E, nu):
    n = x.shape[0]
    ke = np.full((n, 300), np.nan)
    st = np.full(n, -99, np.int32)
    rc = fn(ctx, n, x.ctypes.data_as(C.POINTER(C.c_double)),
            E.ctypes.data_as(C.POINTER(C.c_double)), nu,
            ke.ctypes.data_as(C.POINTER(C.c_double)), st.ctypes.data_as(C.POINTER(C.c_int)))
    return rc, ke, st


@pytest.mark.gpu
def test_noalloc_equals_allocating_entry_point():
    lib = L.lib()
    h = C.c_void_p()
    assert lib.msa_ctx_create(0, C.byref(h)) == L.MSA_OK, lib.msa_last_error()
    try:
        for n, seed, nu in ((LIMIT, 1, 0.3), (37, 2, 0.25)):
            print(f"[no-alloc] n_elem={n} seed={seed} nu={nu}", flush=True)
            x, E = _hexes(n, seed)
            rc_a, ke_a, st_a = _call(lib.msa_ke_packed, h, x, E, nu)
            assert rc_a == L.MSA_OK, lib.msa_last_error()
            rc_b, ke_b, st_b = _call(lib.msa_ke_packed_noalloc, h, x, E, nu)
            assert rc_b == L.MSA_OK, lib.msa_last_error()
            assert np.isfinite(ke_b).all()
            assert np.array_equal(ke_a, ke_b)            # same kernel: bit-identical
            assert np.array_equal(st_a, st_b)
            assert ((st_b == L.MSA_OK) | (st_b == L.MSA_NEG_DETJ)).all()
            assert (st_b == L.MSA_NEG_DETJ).sum() >= 1
        # beyond the limit: refused with an argument error, outputs untouched
        x, E = _hexes(LIMIT + 1, 3)
        rc, ke, st = _call(lib.msa_ke_packed_noalloc, h, x, E, 0.3)
        assert rc == L.MSA_ERR_ARG
        assert b"limit" in lib.msa_last_error()
        assert np.isnan(ke).all() and (st == -99).all()
    finally:
        lib.msa_ctx_destroy(h)
