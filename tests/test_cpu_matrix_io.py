"""Host-side storage formats: CSR -> skyline against the oracle's own
skyline assembly (profile rule of GlobalMatrixAssemblerSkyline), the
reference's ASCII (WriteToFile) and binary (ReadFromFile) file layouts, and
the DOFSLog text format."""
import struct

import numpy as np

from msolve_amd.matrix_io import (DOFSLog, csr_to_skyline, read_skyline, skyline_to_dense,
                                  write_skyline)
from oracle import oracle as O


def model_and_matrices(n=3, seed=1):
    cl = O.Cantilever(n)
    E = 2.1e5 * np.exp(0.3 * np.random.default_rng(seed).standard_normal(cl.n_elem))
    ke, _ = O.ke_mesh(cl.conn, cl.xyz, E, 0.3)
    ri, data = O.skyline_assemble(cl.conn, cl.dof, cl.n, ke)
    rp, col, val = O.csr_assemble(cl.conn, cl.dof, cl.n, ke)
    return cl, ri, data, rp, col, val


def test_csr_to_skyline_equals_reference_assembly():
    for n in (2, 3):
        cl, ri, data, rp, col, val = model_and_matrices(n)
        r2, d2 = csr_to_skyline(rp, col, val, cl.n)
        assert np.array_equal(r2, ri)
        # both sum the element contributions in element order: bit-equal
        assert np.array_equal(d2, data)
        K = skyline_to_dense(r2, d2)
        assert np.allclose(K, K.T)


def test_file_round_trips(tmp_path):
    cl, ri, data, *_ = model_and_matrices(2)
    name = str(tmp_path / "K.txt")
    ix, dn = write_skyline(name, ri, data)                 # WriteToFile layout
    assert ix.endswith("K-Ix.txt") and dn.endswith("K-Data.txt")
    lines = open(dn, newline="").read().split("\r\n")
    assert lines[0] == format(data[0], ".17g")
    r2, d2 = read_skyline(name, binary=False)
    assert np.array_equal(r2, ri) and np.array_equal(d2, data)
    write_skyline(name, ri, data, binary=True)              # ReadFromFile layout
    raw = open(ix, "rb").read()
    assert struct.unpack("<i", raw[:4])[0] == ri.size and len(raw) == 4 + 4 * ri.size
    r3, d3 = read_skyline(name)
    assert np.array_equal(r3, ri) and np.array_equal(d3, data)


def test_dofs_log_format():
    log = DOFSLog([0, 2])
    log.StoreResults(None, None, np.array([1.5e-5, 7.0, -2.25e3]))
    assert log.DOFValues == {0: 1.5e-5, 2: -2.25e3}
    assert str(log) == "(0): 1.500000e-005; (2): -2.250000e+003; "
