"""CPU oracle: ctypes bindings over oracle/_build/libmsolve_oracle.so.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() and
by the cpu_baseline leg of bench.py, always as the checker (or the timed CPU
restatement), never by the msolve_amd product.

The C file restates the reference algorithm (citations in msolve_oracle.c).
This module adds numpy conveniences and `kl_table`, an independent solution of
the exponential-covariance Karhunen-Loeve eigenproblem (Ghanem & Spanos,
"Stochastic Finite Elements: A Spectral Approach", 1991, sec. 2.3.3) laid out
as the reference's KL file fields (LognormalFileStochasticCoefficientsProvider
.cs:90-109: per axis (w, lambda, a) per 1-D mode, evpPos, eigs).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libmsolve_oracle.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


def build() -> str:
    """Compile the oracle with its Makefile (gcc only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    I, D, V = C.c_int, C.c_double, None
    sig = {
        "orc_mesh_box": (I, [I, I, I, D, D, D, _f64p, _i32p]),
        "orc_enumerate_dofs": (I, [I, _u8p, _i32p]),
        "orc_cantilever_bc": (V, [I, I, I, _u8p]),
        "orc_cantilever_loads": (I, [I, I, I, _i32p, D, _f64p]),
        "orc_h8_gauss_matrices": (V, [I, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p]),
        "orc_h8_k": (V, [I, _f64p, _f64p, _f64p, _f64p]),
        "orc_elastic_D": (V, [D, D, _f64p]),
        "orc_ke_mesh": (I, [I, _i32p, _f64p, _f64p, D, _f64p]),
        "orc_ke_element_csharp": (V, [_f64p, D, D, _f64p]),
        "orc_h8_mass_csharp": (V, [_f64p, D, _f64p]),
        "orc_h8_strains": (V, [I, _f64p, _f64p, _f64p]),
        "orc_h8_forces": (V, [I, _f64p, _f64p, _f64p, _f64p]),
        "orc_h8_mlumped": (V, [I, D, _f64p, _f64p]),
        "orc_elastic_stress": (V, [D, D, _f64p, _f64p]),
        "orc_element_results": (V, [I, _i32p, _f64p, _i32p, _f64p, _f64p, D, _f64p, _f64p,
                                    _f64p]),
        "orc_sinusoidal_acc": (D, [I, D, D, D, D]),
        "orc_newmark_skyline": (I, [I, _i32p, _f64p, _i32p, I, _f64p, D, D, D, D, D, D, D, I,
                                    C.c_void_p, I, D, D, D, I, _f64p, _f64p, _f64p, _f64p]),
        "orc_skyline_rowindex": (C.c_int64, [I, _i32p, _i32p, I, _i64p]),
        "orc_skyline_assemble": (V, [I, _i32p, _i32p, _f64p, _i64p, _f64p]),
        "orc_skyline_multiply": (V, [I, _i64p, _f64p, _f64p, _f64p]),
        "orc_skyline_factorize": (I, [I, _i64p, _f64p, D]),
        "orc_skyline_factorize_zem": (I, [I, _i64p, _f64p, D, I, _i32p, _f64p]),
        "orc_skyline_solve": (V, [I, _i64p, _f64p, _f64p, _f64p]),
        "orc_csr_count": (C.c_int64, [I, _i32p, _i32p, I, _i64p]),
        "orc_csr_fill": (V, [I, _i32p, _i32p, I, _i64p, _i32p]),
        "orc_csr_assemble": (V, [I, _i32p, _i32p, I, _f64p, _i64p, _i32p, _f64p]),
        "orc_csr_multiply": (V, [I, _i64p, _i32p, _f64p, _f64p, _f64p]),
        "orc_pcg_skyline": (I, [I, _i64p, _f64p, I, C.c_void_p, C.c_void_p, _f64p, _f64p,
                                I, D, C.POINTER(C.c_int), C.c_void_p]),
        "orc_pcg_csr": (I, [I, _i64p, _i32p, _f64p, I, _f64p, _f64p, I, D,
                            C.POINTER(C.c_int), C.c_void_p]),
        "orc_element_centroid": (V, [_f64p, _f64p]),
        "orc_kl_field": (I, [I, _i32p, _f64p, I, I, _f64p, _f64p, _i32p, _f64p, _f64p,
                             _f64p, D, D, _f64p]),
        "orc_mc_reference": (I, [I, I, I, D, I, I, _f64p, _f64p, _i32p, _f64p, D, D, I,
                                 _f64p, D, I, C.c_void_p, C.c_void_p, _f64p]),
        "orc_mc_csr_jacobi": (I, [I, I, I, D, I, I, _f64p, _f64p, _i32p, _f64p, D, D, I,
                                  _f64p, D, I, C.c_void_p, C.c_void_p, _f64p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


PRECOND_NONE, PRECOND_JACOBI, PRECOND_SKYLINE = 0, 1, 2


# --------------------------------------------------------------------------
# model helpers
# --------------------------------------------------------------------------
class Cantilever:
    """Unit-cube (or box) cantilever: x = 0 face clamped, total load `load`
    in Z spread over the x = max face nodes (the workload of every config)."""

    def __init__(self, nex, ney=None, nez=None, lx=1.0, ly=1.0, lz=1.0, load=-100.0):
        ney = nex if ney is None else ney
        nez = nex if nez is None else nez
        self.nex, self.ney, self.nez = nex, ney, nez
        self.n_nodes = (nex + 1) * (ney + 1) * (nez + 1)
        self.n_elem = nex * ney * nez
        self.xyz = np.zeros(3 * self.n_nodes)
        self.conn = np.zeros(8 * self.n_elem, np.int32)
        lib().orc_mesh_box(nex, ney, nez, lx / nex, ly / ney, lz / nez, self.xyz, self.conn)
        self.constrained = np.zeros(3 * self.n_nodes, np.uint8)
        lib().orc_cantilever_bc(nex, ney, nez, self.constrained)
        self.dof = np.zeros(3 * self.n_nodes, np.int32)
        self.n = lib().orc_enumerate_dofs(self.n_nodes, self.constrained, self.dof)
        self.f = np.zeros(self.n)
        lib().orc_cantilever_loads(nex, ney, nez, self.dof, load, self.f)
        self.domain = np.array([lx, ly, lz])


def enumerate_dofs(constrained):
    c = np.ascontiguousarray(constrained, np.uint8).ravel()
    dof = np.zeros(c.size, np.int32)
    n = lib().orc_enumerate_dofs(c.size // 3, c, dof)
    return dof, n


def gauss_matrices(xyz8, iInt=2):
    n3 = iInt ** 3
    x = np.ascontiguousarray(xyz8, np.float64).ravel()
    w, S, DS = np.zeros(n3), np.zeros(8 * n3), np.zeros(24 * n3)
    J, dJ, B = np.zeros(9 * n3), np.zeros(n3), np.zeros(144 * n3)
    lib().orc_h8_gauss_matrices(iInt, x, w, S, DS, J, dJ, B)
    return dict(w=w, S=S.reshape(n3, 8), DS=DS.reshape(n3, 24), J=J.reshape(n3, 3, 3),
                detJ=dJ, B=B.reshape(n3, 24, 6))


def h8_k(E_gp, B, w, iInt=2):
    K = np.zeros(300)
    lib().orc_h8_k(iInt, np.ascontiguousarray(E_gp, np.float64).ravel(),
                   np.ascontiguousarray(B, np.float64).ravel(),
                   np.ascontiguousarray(w, np.float64), K)
    return K


def elastic_D(E, nu):
    D = np.zeros(36)
    lib().orc_elastic_D(E, nu, D)
    return D.reshape(6, 6)


def ke_mesh(conn, xyz, E_elem, nu):
    conn = np.ascontiguousarray(conn, np.int32).ravel()
    ne = conn.size // 8
    ke = np.zeros(300 * ne)
    bad = lib().orc_ke_mesh(ne, conn, np.ascontiguousarray(xyz, np.float64).ravel(),
                            np.ascontiguousarray(E_elem, np.float64), nu, ke)
    return ke.reshape(ne, 300), bad


def ke_element_csharp(xyz8, E, nu):
    K = np.zeros(300)
    lib().orc_ke_element_csharp(np.ascontiguousarray(xyz8, np.float64).ravel(), E, nu, K)
    return K


def h8_mass_csharp(xyz8, rho):
    """Consistent mass of one Hexa8 (Hexa8.cs:403-436), packed upper 300."""
    M = np.zeros(300)
    lib().orc_h8_mass_csharp(np.ascontiguousarray(xyz8, np.float64).ravel(), rho, M)
    return M


def h8_strains(B, u):
    out = np.zeros(48)
    lib().orc_h8_strains(2, np.ascontiguousarray(B, np.float64).ravel(),
                         np.ascontiguousarray(u, np.float64), out)
    return out.reshape(8, 6)


def h8_forces(B, w, stresses):
    out = np.zeros(24)
    lib().orc_h8_forces(2, np.ascontiguousarray(B, np.float64).ravel(),
                        np.ascontiguousarray(w, np.float64),
                        np.ascontiguousarray(stresses, np.float64).ravel(), out)
    return out


def h8_mlumped(density, w):
    out = np.zeros(300)
    lib().orc_h8_mlumped(2, density, np.ascontiguousarray(w, np.float64), out)
    return out


def element_results(model, u, E, nu=0.3):
    """strains/stresses [n_elem, 8, 6] and internal forces [n_elem, 24]."""
    ne = model.n_elem
    eps, sig, f = np.zeros(ne * 48), np.zeros(ne * 48), np.zeros(ne * 24)
    lib().orc_element_results(ne, model.conn, model.xyz, model.dof,
                              np.ascontiguousarray(u, np.float64),
                              np.ascontiguousarray(E, np.float64), nu, eps, sig, f)
    return eps.reshape(ne, 8, 6), sig.reshape(ne, 8, 6), f.reshape(ne, 24)


def sinusoidal_acc(i, mag, period, dt, tlen):
    return lib().orc_sinusoidal_acc(int(i), mag, period, dt, tlen)


def newmark_skyline(model, E, nu=0.3, rho=7.85e-9, alphaR=0.0, betaR=0.0, nb_alpha=0.25,
                    nb_delta=0.5, dt=1e-4, nsteps=10, f_static=None, acc_dir=2, acc_mag=1.0,
                    acc_period=0.05, acc_tlen=0.05, probe=None):
    """Reference Newmark loop with the skyline direct solver; returns
    (probe history, u, v, a) after nsteps."""
    n = model.n
    probe = n - 1 if probe is None else probe
    hist = np.zeros(max(nsteps, 1))
    u, v, a = np.zeros(n), np.zeros(n), np.zeros(n)
    fs = None
    if f_static is not None:
        f_static = np.ascontiguousarray(f_static, np.float64)
        fs = f_static.ctypes.data_as(C.c_void_p)
    rc = lib().orc_newmark_skyline(model.n_elem, model.conn, model.xyz, model.dof, n,
                                   np.ascontiguousarray(E, np.float64), nu, rho, alphaR, betaR,
                                   nb_alpha, nb_delta, dt, nsteps, fs, acc_dir, acc_mag,
                                   acc_period, acc_tlen, probe, hist, u, v, a)
    if rc != 0:
        raise RuntimeError(f"orc_newmark_skyline failed ({rc})")
    return hist[:nsteps], u, v, a


def unpack300(k300):
    """Packed upper-triangle row-major (SymmetricMatrix2D) -> dense 24x24."""
    K = np.zeros((24, 24))
    iu = np.triu_indices(24)
    K[iu] = k300
    K = K + np.triu(K, 1).T
    return K


def skyline_assemble(conn, dof, n, ke):
    conn = np.ascontiguousarray(conn, np.int32).ravel()
    ne = conn.size // 8
    ri = np.zeros(n + 1, np.int64)
    length = lib().orc_skyline_rowindex(ne, conn, dof, n, ri)
    data = np.zeros(length)
    lib().orc_skyline_assemble(ne, conn, dof, np.ascontiguousarray(ke).ravel(), ri, data)
    return ri, data


def skyline_multiply(ri, data, x):
    y = np.zeros(ri.size - 1)
    lib().orc_skyline_multiply(ri.size - 1, ri, data, np.ascontiguousarray(x, np.float64), y)
    return y


def skyline_factorize(ri, data, tol=1e-15):
    d = data.copy()
    rc = lib().orc_skyline_factorize(ri.size - 1, ri, d, tol)
    return d, rc


def skyline_factorize_zem(ri, data, tol, max_modes=16):
    """Factorize with zero-energy modes: (factor, n_modes, cols, zems[k][n])."""
    n = ri.size - 1
    d = data.copy()
    cols = np.zeros(max_modes, np.int32)
    zems = np.zeros((max_modes, n))
    k = lib().orc_skyline_factorize_zem(n, ri, d, tol, max_modes, cols, zems)
    kept = min(max(k, 0), max_modes)
    return d, k, cols[:kept].copy(), zems[:kept].copy()


def skyline_solve(ri, factored, f):
    x = np.zeros(ri.size - 1)
    lib().orc_skyline_solve(ri.size - 1, ri, factored, np.ascontiguousarray(f, np.float64), x)
    return x


def skyline_to_dense(ri, data):
    n = ri.size - 1
    K = np.zeros((n, n))
    for j in range(n):
        for h in range(ri[j + 1] - ri[j]):
            K[j - h, j] = K[j, j - h] = data[ri[j] + h]
    return K


def csr_assemble(conn, dof, n, ke):
    conn = np.ascontiguousarray(conn, np.int32).ravel()
    ne = conn.size // 8
    rp = np.zeros(n + 1, np.int64)
    nnz = lib().orc_csr_count(ne, conn, dof, n, rp)
    col = np.zeros(nnz, np.int32)
    lib().orc_csr_fill(ne, conn, dof, n, rp, col)
    val = np.zeros(nnz)
    lib().orc_csr_assemble(ne, conn, dof, n, np.ascontiguousarray(ke).ravel(), rp, col, val)
    return rp, col, val


def csr_multiply(rp, col, val, x):
    y = np.zeros(rp.size - 1)
    lib().orc_csr_multiply(rp.size - 1, rp, col, val, np.ascontiguousarray(x, np.float64), y)
    return y


def pcg_skyline(ri, data, f, precond=PRECOND_JACOBI, P_ri=None, P=None, tol=1e-10,
                max_it=20000, history=False):
    n = ri.size - 1
    x = np.zeros(n)
    it = C.c_int(0)
    hist = np.zeros(max_it) if history else None
    pri = P_ri.ctypes.data_as(C.c_void_p) if P_ri is not None else None
    pp = P.ctypes.data_as(C.c_void_p) if P is not None else None
    rc = lib().orc_pcg_skyline(n, ri, data, precond, pri, pp, np.ascontiguousarray(f), x,
                               max_it, tol, C.byref(it),
                               hist.ctypes.data_as(C.c_void_p) if history else None)
    out = dict(x=x, iters=it.value, converged=(rc == 0))
    if history:
        out["history"] = hist[: it.value]
    return out


def pcg_csr(rp, col, val, f, precond=PRECOND_JACOBI, tol=1e-10, max_it=20000, history=False):
    n = rp.size - 1
    x = np.zeros(n)
    it = C.c_int(0)
    hist = np.zeros(max_it) if history else None
    rc = lib().orc_pcg_csr(n, rp, col, val, precond, np.ascontiguousarray(f), x, max_it, tol,
                           C.byref(it), hist.ctypes.data_as(C.c_void_p) if history else None)
    out = dict(x=x, iters=it.value, converged=(rc == 0))
    if history:
        out["history"] = hist[: it.value]
    return out


# --------------------------------------------------------------------------
# Karhunen-Loeve table (independent solution, scipy brentq)
# --------------------------------------------------------------------------
def kl_1d_modes(length, corr_len, kmax):
    """1-D eigenpairs of exp(-|x-y|/b) on [-L/2, L/2], ordered by frequency:
    odd positions (1-based) are cosine modes, even positions sine modes, as
    the reference's `index % 2 == 0 ? sin : cos` rule requires
    (LognormalFileStochasticCoefficientsProvider.cs:150).  Returns arrays
    (w, lam, a) of length kmax with a the L2-normalisation factor."""
    from scipy.optimize import brentq

    half = 0.5 * length
    c = 1.0 / corr_len
    w, lam, a = [], [], []
    k = 0
    while len(w) < kmax:
        if k % 2 == 0:  # cosine: c - w tan(w half) = 0 on ((m) pi, (m + 1/2) pi)/half
            m = k // 2
            lo, hi = m * math.pi / half, (m + 0.5) * math.pi / half
            g = lambda x: c * math.cos(x * half) - x * math.sin(x * half)
            om = brentq(g, lo + 1e-14, hi - 1e-14, xtol=1e-15, rtol=1e-15, maxiter=500)
            nrm = half + math.sin(2 * om * half) / (2 * om)
        else:  # sine: w + c tan(w half) = 0 on ((m + 1/2) pi, (m + 1) pi)/half
            m = k // 2
            lo, hi = (m + 0.5) * math.pi / half, (m + 1.0) * math.pi / half
            g = lambda x: x * math.cos(x * half) + c * math.sin(x * half)
            om = brentq(g, lo + 1e-14, hi - 1e-14, xtol=1e-15, rtol=1e-15, maxiter=500)
            nrm = half - math.sin(2 * om * half) / (2 * om)
        w.append(om)
        lam.append(2 * c / (om * om + c * c))
        a.append(1.0 / math.sqrt(nrm))
        k += 1
    return np.array(w), np.array(lam), np.array(a)


def kl_table(domain, corr_len, M, kmax=None):
    """Separable 3-D table in the reference's KL-file layout: returns dict with
    w1d[3,kmax], lam1d[3,kmax], a1d[3,kmax], evpPos[M,3] (1-based), eigs[M].
    3-D terms are ranked by eigenvalue (descending); ties are ordered by the
    rounded log-eigenvalue, then by (i, j, k) ascending."""
    if kmax is None:
        kmax = M
    w1, l1, a1 = [], [], []
    for d in range(3):
        w, l, a = kl_1d_modes(domain[d], corr_len, kmax)
        w1.append(w); l1.append(l); a1.append(a)
    cand = []
    for i in range(kmax):
        for j in range(kmax):
            for k in range(kmax):
                lam = (l1[0][i] * l1[1][j]) * l1[2][k]
                cand.append((round(-math.log(lam) * 1e9), i, j, k, lam))
    cand.sort()
    sel = cand[:M]
    evp = np.array([[c[1] + 1, c[2] + 1, c[3] + 1] for c in sel], np.int32)
    eigs = np.array([c[4] for c in sel])
    return dict(w1d=np.array(w1), lam1d=np.array(l1), a1d=np.array(a1), evpPos=evp, eigs=eigs,
                kmax=kmax, M=M)


def kl_field(conn, xyz, table, xi, domain, meanE, cov):
    conn = np.ascontiguousarray(conn, np.int32).ravel()
    ne = conn.size // 8
    E = np.zeros(ne)
    bad = lib().orc_kl_field(ne, conn, np.ascontiguousarray(xyz, np.float64).ravel(),
                             table["M"], table["kmax"],
                             np.ascontiguousarray(table["w1d"]).ravel(),
                             np.ascontiguousarray(table["a1d"]).ravel(),
                             np.ascontiguousarray(table["evpPos"], np.int32).ravel(),
                             np.ascontiguousarray(table["eigs"]),
                             np.ascontiguousarray(domain, np.float64),
                             np.ascontiguousarray(xi, np.float64), meanE, cov, E)
    return E, bad


def mc_reference(nex, table, xi, meanE=2.1e5, cov=0.2, nu=0.3, tol=1e-10, max_it=20000,
                 keep_u=True, storage="skyline"):
    """Reference Monte Carlo loop (orc_mc_reference; storage="csr" runs
    orc_mc_csr_jacobi).  xi: [n_samples, M]."""
    xi = np.ascontiguousarray(xi, np.float64)
    ns = xi.shape[0]
    cl = Cantilever(nex)
    u = np.zeros((ns, cl.n)) if keep_u else None
    iters = np.zeros(ns, np.int32)
    times = np.zeros(4)
    fn = lib().orc_mc_reference if storage == "skyline" else lib().orc_mc_csr_jacobi
    rc = fn(nex, nex, nex, nu, table["M"], table["kmax"],
                                np.ascontiguousarray(table["w1d"]).ravel(),
                                np.ascontiguousarray(table["a1d"]).ravel(),
                                np.ascontiguousarray(table["evpPos"], np.int32).ravel(),
                                np.ascontiguousarray(table["eigs"]), meanE, cov, ns, xi.ravel(),
                                tol, max_it,
                                u.ctypes.data_as(C.c_void_p) if keep_u else None,
                                iters.ctypes.data_as(C.c_void_p), times)
    if rc < 0:
        raise RuntimeError(f"orc_mc_reference failed: {rc}")
    return dict(u=u, iters=iters, times=dict(element=times[0], factorize=times[1],
                                             solution=times[2], all=times[3]))
