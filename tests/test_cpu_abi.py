"""CPU-only checks of the C ABI library: it loads without a GPU, exports every
symbol include/msolve_amd.h declares, fails loudly (MSA_ERR_NO_DEVICE) when
no device is present, and its host-side numbering helper follows the
reference rule."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from msolve_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "msolve_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:msa_|CALCH8)[A-Za-z0-9_]+)\s*\(", text)))


def test_header_declares_exports_list():
    assert declared_symbols() == sorted(L.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(L.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_version_and_no_device_error():
    lib = L.lib()
    assert lib.msa_version() >= 100
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    rc = lib.msa_ctx_create(0, C.byref(h))
    assert rc == L.MSA_ERR_NO_DEVICE
    assert b"device" in lib.msa_last_error()


def test_enumerate_dofs_reference_rule():
    rng = np.random.default_rng(0)
    cons = (rng.random((50, 3)) < 0.3).astype(np.uint8)
    from msolve_amd import enumerate_dofs
    dof, n = enumerate_dofs(cons)
    # brute force: ascending node, X/Y/Z, constrained -> -1 (Model.cs:165-215)
    k, ref = 0, []
    for i in range(50):
        for d in range(3):
            if cons[i, d]:
                ref.append(-1)
            else:
                ref.append(k)
                k += 1
    assert n == k and dof.tolist() == ref
