"""Lognormal Karhunen-Loeve modulus field: eigen-table, the reference's KL and
random-variable file formats, and seeded standard-normal coefficients.

The reference reads a precomputed separable KL table from a text file
(LognormalFileStochasticCoefficientsProvider.cs:71-120) that is not part of
its repository.  `kl_table` solves the 1-D eigenproblem of the exponential
covariance exp(-|x - y| / b) on [-L/2, L/2] (Ghanem & Spanos, "Stochastic
Finite Elements: A Spectral Approach", 1991, sec. 2.3.3) and lays it out in
that file's fields; the field itself is evaluated on the GPU
(msa_batch_set_kl / msa_batch_eval_kl).
"""
from __future__ import annotations

import math

import numpy as np


def lognormal_sigma(cov: float, mean: float = 1.0) -> float:
    """sigma_G = sqrt(ln((cov/mean)^2 + 1)) (Lognormal...cs:115-117)."""
    return math.sqrt(math.log((cov / mean) ** 2 + 1.0))


def _root(fun, dfun, lo, hi, iters=200):
    """Safeguarded Newton on a bracketing interval (fun(lo), fun(hi) differ in sign)."""
    flo = fun(lo)
    x = 0.5 * (lo + hi)
    for _ in range(iters):
        fx = fun(x)
        if fx == 0.0:
            return x
        if (fx > 0) == (flo > 0):
            lo, flo = x, fx
        else:
            hi = x
        d = dfun(x)
        xn = x - fx / d if d != 0 else 0.5 * (lo + hi)
        if not (lo < xn < hi):
            xn = 0.5 * (lo + hi)
        if abs(xn - x) <= 1e-16 * max(1.0, abs(x)):
            return xn
        x = xn
    return x


def kl_modes_1d(length: float, corr_len: float, kmax: int):
    """Frequencies, eigenvalues and normalisations of the first kmax 1-D modes
    ordered by frequency; odd 1-based positions are cosine modes, even ones
    sine modes (the `index % 2 == 0 ? sin : cos` rule, Lognormal...cs:150)."""
    half = 0.5 * length
    c = 1.0 / corr_len
    w, lam, a = np.zeros(kmax), np.zeros(kmax), np.zeros(kmax)
    eps = 1e-13 / half
    for k in range(kmax):
        m = k // 2
        if k % 2 == 0:   # cos: c cos(w h) - w sin(w h) = 0 on (m pi, (m+1/2) pi)/h
            fun = lambda x: c * math.cos(x * half) - x * math.sin(x * half)
            dfun = lambda x: -c * half * math.sin(x * half) - math.sin(x * half) - x * half * math.cos(x * half)
            lo, hi = m * math.pi / half + eps, (m + 0.5) * math.pi / half - eps
        else:            # sin: w cos(w h) + c sin(w h) = 0 on ((m+1/2) pi, (m+1) pi)/h
            fun = lambda x: x * math.cos(x * half) + c * math.sin(x * half)
            dfun = lambda x: math.cos(x * half) - x * half * math.sin(x * half) + c * half * math.cos(x * half)
            lo, hi = (m + 0.5) * math.pi / half + eps, (m + 1.0) * math.pi / half - eps
        om = _root(fun, dfun, lo, hi)
        w[k] = om
        lam[k] = 2.0 * c / (om * om + c * c)
        s2 = math.sin(2.0 * om * half) / (2.0 * om)
        a[k] = 1.0 / math.sqrt(half + s2 if k % 2 == 0 else half - s2)
    return w, lam, a


def kl_table(domain, corr_len: float, M: int, kmax: int | None = None) -> dict:
    """Separable 3-D KL table with the M largest product eigenvalues.  Ties
    (cube symmetry) are ordered by the log-eigenvalue rounded to 1e-9, then by
    the 1-D indices (i, j, k).  Keys follow the reference's file fields:
    w1d/lam1d/a1d [3, kmax], evpPos [M, 3] (1-based), eigs [M]."""
    kmax = M if kmax is None else kmax
    W, Lm, A = [], [], []
    for d in range(3):
        w, l, a = kl_modes_1d(float(domain[d]), corr_len, kmax)
        W.append(w); Lm.append(l); A.append(a)
    cand = []
    for i in range(kmax):
        for j in range(kmax):
            for k in range(kmax):
                lam = (Lm[0][i] * Lm[1][j]) * Lm[2][k]
                cand.append((round(-math.log(lam) * 1e9), i, j, k, lam))
    cand.sort()
    sel = cand[:M]
    return dict(w1d=np.array(W), lam1d=np.array(Lm), a1d=np.array(A),
                evpPos=np.array([[c[1] + 1, c[2] + 1, c[3] + 1] for c in sel], np.int32),
                eigs=np.array([c[4] for c in sel]), M=M, kmax=kmax)


def write_kl_file(path: str, table: dict, mean: float = 1.0, cov: float = 0.5, sep: str = "\t"):
    """The reference's KL text layout (Lognormal...cs:90-113): per axis j and
    term i the line `w l a` at i + j*M, then evpPos lines, eigenvalue lines,
    a mean line and a covariance line.  Needs kmax == M."""
    M = table["M"]
    assert table["kmax"] == M, "the file format holds M 1-D modes per axis"
    lines = []
    for j in range(3):
        for i in range(M):
            lines.append(sep.join(repr(float(v)) for v in
                                  (table["w1d"][j][i], table["lam1d"][j][i], table["a1d"][j][i])))
    for i in range(M):
        lines.append(sep.join(str(int(v)) for v in table["evpPos"][i]))
    for i in range(M):
        lines.append(repr(float(table["eigs"][i])))
    lines.append(repr(float(mean)))
    lines.append(repr(float(cov)))
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def read_kl_file(path: str, M: int, sep: str = "\t") -> dict:
    """Parse the reference's KL text layout for 3 coordinates and M terms."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    W, Lm, A = np.zeros((3, M)), np.zeros((3, M)), np.zeros((3, M))
    pos, eigs = np.zeros((M, 3), np.int32), np.zeros(M)
    for i in range(M):
        for j in range(3):
            v = lines[i + j * M].split(sep)
            W[j, i], Lm[j, i], A[j, i] = float(v[0]), float(v[1]), float(v[2])
        pos[i] = [int(t) for t in lines[i + 3 * M].split(sep)[:3]]
        eigs[i] = float(lines[i + 4 * M].split(sep)[0])
    return dict(w1d=W, lam1d=Lm, a1d=A, evpPos=pos, eigs=eigs, M=M, kmax=M)


def standard_normal(seed: int, n_samples: int, M: int) -> np.ndarray:
    """Seeded N(0,1) coefficients [n_samples, M] (numpy Philox; the
    reference's Troschuetz NormalDistribution is unseeded)."""
    g = np.random.Generator(np.random.Philox(seed))
    return g.standard_normal((n_samples, M))


def write_randoms_file(path: str, xi: np.ndarray):
    """randoms{M}.txt layout (MC...Material.cs:199-204): term-major, one value
    per line."""
    n, M = xi.shape
    with open(path, "w") as fh:
        for j in range(M):
            for i in range(n):
                fh.write(repr(float(xi[i, j])) + "\n")


def read_randoms_file(path: str, M: int, n_file: int = 50000) -> np.ndarray:
    """Reader of MC...Material.cs:206-217 (n_file rows per term)."""
    vals = np.loadtxt(path, dtype=np.float64)
    return vals[: M * n_file].reshape(M, n_file).T.copy()
