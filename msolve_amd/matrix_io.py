"""Stiffness storage and DOF logs in the reference's formats (host side).

* Skyline conversion of an assembled K (the FP64 image read back with
  Batch.matrix_csr) into SkylineMatrix2D's arrays: column j stores K[j, j],
  K[j-1, j], ..., K[j-h_j, j] at data[rowIndex[j] + (j - i)], with
  rowIndex[0] = 0 and rowIndex[j+1] = rowIndex[j] + h_j + 1
  (ISAAR.MSolve.Matrices/SkylineMatrix2D.cs:18-61; the heights are those of
  GlobalMatrixAssemblerSkyline.CalculateRowIndex, GMAS.cs:14-44, because a
  column's first stored row is the smallest DOF of the elements holding it).
* SkylineMatrix2D.WriteToFile (SkylineMatrix2D.cs:427-460) writes ASCII:
  "<name>-Ix<ext>" one rowIndex entry per line, "<name>-Data<ext>" one value
  per line in .NET "g17".  SkylineMatrix2D.ReadFromFile (:462-497) reads a
  different, binary layout: int32 count then int32 entries, int32 count then
  float64 values (little endian).  Both are provided, so a file written for
  the reference's reader round-trips.
* DOFSLog (ISAAR.MSolve.Logging/DOFsLog.cs:10-45): the values of chosen
  DOFs of a solution.
"""
from __future__ import annotations

import os
import struct

import numpy as np


def csr_to_skyline(rowptr, col, val, n: int):
    """Full symmetric CSR (rows sorted or not) -> (rowIndex int64[n+1],
    data float64[rowIndex[n]])."""
    rowptr = np.asarray(rowptr, np.int64)
    col = np.asarray(col, np.int64)
    val = np.asarray(val, np.float64)
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rowptr))
    upper = col >= rows            # entry K[r, c] with r <= c lives in column c
    r, c, v = rows[upper], col[upper], val[upper]
    # in the symmetric pattern the first stored row of column c is the
    # smallest column of row c
    first = np.full(n, np.iinfo(np.int64).max)
    np.minimum.at(first, rows, col)
    heights = np.arange(n, dtype=np.int64) - first
    rowIndex = np.zeros(n + 1, np.int64)
    np.cumsum(heights + 1, out=rowIndex[1:])
    data = np.zeros(int(rowIndex[n]))
    data[rowIndex[c] + (c - r)] = v
    return rowIndex, data


def skyline_to_csr(rowIndex, data, drop_zeros: bool = False):
    """SkylineMatrix2D arrays -> full symmetric CSR (rowptr int64[n+1],
    col int32, val float64), columns ascending per row.  Stored zeros of the
    profile stay entries unless drop_zeros (the diagonal always stays)."""
    rowIndex = np.asarray(rowIndex, np.int64)
    data = np.asarray(data, np.float64)
    n = rowIndex.size - 1
    h = np.diff(rowIndex)                      # entries of column j: K[j-k, j], k < h_j
    cols = np.repeat(np.arange(n, dtype=np.int64), h)
    k = np.arange(int(rowIndex[n]), dtype=np.int64) - np.repeat(rowIndex[:-1], h)
    if drop_zeros:
        keep = (data != 0) | (k == 0)
        cols, k, data = cols[keep], k[keep], data[keep]
    rows = cols - k
    off = k > 0                                # mirror the strictly upper entries
    r = np.concatenate([rows, cols[off]])
    c = np.concatenate([cols, rows[off]])
    v = np.concatenate([data, data[off]])
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    rowptr = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(r, minlength=n), out=rowptr[1:])
    return rowptr, c.astype(np.int32), v


def skyline_to_dense(rowIndex, data):
    n = len(rowIndex) - 1
    K = np.zeros((n, n))
    for j in range(n):
        h = rowIndex[j + 1] - rowIndex[j]
        for k in range(h):
            K[j - k, j] = K[j, j - k] = data[rowIndex[j] + k]
    return K


def _names(name: str):
    path = os.path.dirname(name)
    base, ext = os.path.splitext(os.path.basename(name))
    return (os.path.join(path, base + "-Ix" + ext), os.path.join(path, base + "-Data" + ext))


def _g17(x: float) -> str:
    """.NET Double.ToString("g17", en-US): 17 significant digits, C %g
    switching rules, lower-case exponent with at least two digits."""
    if x == 0.0:
        return "0"
    return format(float(x), ".17g")


def write_skyline(name: str, rowIndex, data, binary: bool = False):
    """WriteToFile's ASCII layout (default) or ReadFromFile's binary layout."""
    ix_name, data_name = _names(name)
    rowIndex = np.asarray(rowIndex, np.int64)
    data = np.asarray(data, np.float64)
    if binary:
        if rowIndex[-1] > np.iinfo(np.int32).max:
            raise ValueError("skyline larger than the reference's int32 rowIndex")
        with open(ix_name, "wb") as fh:
            fh.write(struct.pack("<i", rowIndex.size))
            fh.write(rowIndex.astype("<i4").tobytes())
        with open(data_name, "wb") as fh:
            fh.write(struct.pack("<i", data.size))
            fh.write(data.astype("<f8").tobytes())
        return ix_name, data_name
    with open(ix_name, "w", newline="\r\n") as fh:
        fh.write("\n".join(str(int(i)) for i in rowIndex) + "\n")
    with open(data_name, "w", newline="\r\n") as fh:
        fh.write("\n".join(_g17(d) for d in data) + "\n")
    return ix_name, data_name


def _is_binary(path: str) -> bool:
    """ReadFromFile's layout: an int32 count, then exactly that many items."""
    size = os.path.getsize(path)
    if size < 4:
        return False
    with open(path, "rb") as fh:
        (m,) = struct.unpack("<i", fh.read(4))
    return m >= 0 and size in (4 + 4 * m, 4 + 8 * m)


def read_skyline(name: str, binary: bool | None = True):
    """ReadFromFile's binary layout (default), the ASCII one WriteToFile
    produces, or (binary=None) whichever the files hold; returns (rowIndex
    int64, data float64).  The reference writes ASCII (WriteToFile) and
    reads binary (ReadFromFile), so its own stored matrices do not read
    back; None accepts both."""
    ix_name, data_name = _names(name)
    if binary is None:
        binary = _is_binary(ix_name)
    if binary:
        with open(ix_name, "rb") as fh:
            (m,) = struct.unpack("<i", fh.read(4))
            rowIndex = np.frombuffer(fh.read(4 * m), "<i4").astype(np.int64)
        with open(data_name, "rb") as fh:
            (m,) = struct.unpack("<i", fh.read(4))
            data = np.frombuffer(fh.read(8 * m), "<f8").astype(np.float64)
        return rowIndex, data
    rowIndex = np.array([int(l) for l in open(ix_name).read().split()], np.int64)
    data = np.array([float(l) for l in open(data_name).read().split()], np.float64)
    return rowIndex, data


def _dotnet_e(x: float) -> str:
    """.NET "{0:e}": six decimals, exponent sign and at least three digits."""
    m, e = f"{x:.6e}".split("e")
    return f"{m}e{e[0]}{abs(int(e)):03d}"


class DOFSLog:
    """Values of selected DOFs of each stored solution (DOFsLog.cs:10-45)."""

    def __init__(self, dofs):
        self.dofs = [int(d) for d in dofs]
        self.DOFValues: dict[int, float] = {}
        self.StartTime = None
        self.EndTime = None

    def StoreResults(self, startTime, endTime, solution):
        self.StartTime, self.EndTime = startTime, endTime
        for d in self.dofs:
            self.DOFValues[d] = float(solution[d])

    def __str__(self):
        return "".join(f"({d}): {_dotnet_e(v)}; " for d, v in self.DOFValues.items())
