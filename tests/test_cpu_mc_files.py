"""Host side of the Monte Carlo analyzer's files: skyline <-> CSR, the two
stored-matrix layouts read back, the .NET value and TimeSpan formats, and
the analyzer's argument checks (no GPU)."""
import numpy as np
import pytest

import msolve_amd as msa
from msolve_amd import _lib as L
from msolve_amd.analyzers import StiffnessMatrixProductionMode as Mode
from msolve_amd.analyzers import _dotnet_double, _timespan
from msolve_amd.matrix_io import (csr_to_skyline, read_skyline, skyline_to_csr, skyline_to_dense,
                                  write_skyline)


def random_profile(n, seed):
    rng = np.random.default_rng(seed)
    A = np.zeros((n, n))
    for i in range(n):
        for j in range(max(0, i - int(rng.integers(0, 7))), i + 1):
            A[i, j] = A[j, i] = rng.standard_normal()
    A += 10 * np.eye(n)
    A[6, 0] = A[0, 6] = 1.0     # column 6's profile starts at row 0 ...
    A[6, 2] = A[2, 6] = 0.0     # ... so K[2, 6] is a stored zero
    return A


def dense_to_csr(A):
    rp, col, val = [0], [], []
    for i in range(A.shape[0]):
        nz = np.nonzero(A[i])[0]
        col += nz.tolist()
        val += A[i, nz].tolist()
        rp.append(len(col))
    return np.array(rp), np.array(col), np.array(val)


def test_skyline_csr_round_trip():
    A = random_profile(40, 1)
    ri, d = csr_to_skyline(*dense_to_csr(A), 40)
    assert np.array_equal(skyline_to_dense(ri, d), A)
    for drop in (False, True):
        rp, col, val = skyline_to_csr(ri, d, drop_zeros=drop)
        B = np.zeros_like(A)
        for i in range(40):
            seg = slice(rp[i], rp[i + 1])
            assert np.all(np.diff(col[seg]) > 0)
            B[i, col[seg]] = val[seg]
        assert np.array_equal(A, B)
        assert (2 in col[rp[6]:rp[7]]) != drop   # the profile zero K[6, 2]


@pytest.mark.parametrize("binary", [False, True])
def test_stored_matrix_layouts_read_back(tmp_path, binary):
    A = random_profile(25, 2)
    ri, d = csr_to_skyline(*dense_to_csr(A), 25)
    name = str(tmp_path / "KSub1Sim0.txt")
    write_skyline(name, ri, d, binary=binary)
    r2, d2 = read_skyline(name, None)
    assert np.array_equal(r2, ri) and np.array_equal(d2, d)   # g17 and binary are exact


def test_dotnet_formats():
    assert _timespan(3725.1234567) == "01:02:05.1234567"
    assert _timespan(2.0) == "00:00:02"
    assert _timespan(90061.5) == "1.01:01:01.5000000"
    assert _dotnet_double(0.5) == "0.5"
    assert _dotnet_double(1.2345e-5) == "1.2345E-05"
    assert _dotnet_double(-0.001234567890123456) == "-0.00123456789012346"
    assert _dotnet_double(1.5e17) == "1.5E+17"
    assert float(_dotnet_double(np.pi)) == float("%.15g" % np.pi)


def test_analyzer_argument_checks():
    m = msa.cantilever(2)
    table = {"M": 2}
    mk = msa.MonteCarloAnalyzerWithStochasticMaterial
    with pytest.raises(msa.MsaError):   # LoadFromDisk reads <fileNameForLogging>-Order.txt
        mk(None, [m], [2], table, 2, stiffnessMatrixProductionMode=Mode.LoadFromDiskAndCalculate)
    with pytest.raises(msa.MsaError):   # one model, FP64
        mk(None, [m, m], [2, 2], table, 2, fileNameForLogging="x",
           stiffnessMatrixProductionMode=Mode.LoadFromDiskAndCalculate)
    with pytest.raises(msa.MsaError):
        mk(None, [m], [2], table, 2, fileNameForLogging="x", precision=L.MSA_MIXED,
           stiffnessMatrixProductionMode=Mode.LoadFromDiskAndCalculate)
    with pytest.raises(msa.MsaError):
        mk(None, [m], [2], table, 2, outputDOF=m.n_free)
    an = mk(None, [m], [2], table, 2, outputDOF=1)
    assert an.outputDOF == [1] and an.stiffnessMatrixProductionMode == Mode.Normal
    assert [int(x) for x in Mode] == [0, 1, 2, 3]
