"""Object wrappers over the C ABI: Context, Mesh, Batch."""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np

from . import _lib as L
from ._lib import check, f64, i32, lib, ptr


class Context:
    """One HIP device + stream (msa_ctx).  Meshes and batches created on it are
    closed before the context itself (the C ABI requires that order)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        self._children = weakref.WeakSet()
        check(lib().msa_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = device

    def _adopt(self, child):
        self._children.add(child)

    @property
    def stream(self) -> int:
        return lib().msa_ctx_stream(self.h)

    def synchronize(self):
        check(lib().msa_ctx_synchronize(self.h))

    def close(self):
        if getattr(self, "h", None):
            for child in sorted(list(self._children), key=lambda c: isinstance(c, Mesh)):
                child.close()          # batches first, then meshes
            lib().msa_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def enumerate_dofs(constrained):
    """Free-DOF numbering of Model.EnumerateGlobalDOFs (Model.cs:165-215):
    nodes ascending, X/Y/Z, constrained -> -1.  Returns (dof, n_free)."""
    c = np.ascontiguousarray(constrained, np.uint8).ravel()
    dof = np.zeros(c.size, np.int32)
    n = C.c_int()
    check(lib().msa_enumerate_dofs(c.size // 3, c.ctypes.data_as(C.c_void_p), ptr(dof, C.c_int),
                                   C.byref(n)))
    return dof, n.value


class Mesh:
    """Hexa8 mesh + free-DOF map uploaded to the device (msa_mesh)."""

    def __init__(self, ctx: Context, xyz, conn, dof, n_free: int):
        self.ctx = ctx
        self.xyz = f64(xyz).reshape(-1, 3)
        self.conn = i32(conn).reshape(-1, 8)
        self.dof = i32(dof).ravel()
        self.n_free = int(n_free)
        h = C.c_void_p()
        check(lib().msa_mesh_create(ctx.h, self.xyz.shape[0], ptr(self.xyz, C.c_double),
                                    self.conn.shape[0], ptr(self.conn, C.c_int),
                                    ptr(self.dof, C.c_int), self.n_free, C.byref(h)))
        self.h = h
        ctx._adopt(self)
        info = np.zeros(8, np.int64)
        check(lib().msa_mesh_info(self.h, ptr(info, C.c_longlong)))
        self.info = dict(blocks=int(info[0]), block_nnz=int(info[1]), nnz=int(info[2]),
                         colours=int(info[3]), max_row=int(info[4]), n_free=int(info[5]),
                         n_elem=int(info[6]), n_nodes=int(info[7]))

    @property
    def n_elem(self):
        return self.conn.shape[0]

    def close(self):
        if getattr(self, "h", None):
            lib().msa_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ke_packed(ctx: Context, xyz8, E, nu: float):
    """Batched element stiffness (packed 300 upper entries per element) and
    per-element status (MSA_OK / MSA_NEG_DETJ)."""
    x = f64(xyz8).reshape(-1, 24)
    E = f64(E).ravel()
    n = x.shape[0]
    ke = np.zeros((n, 300))
    st = np.zeros(n, np.int32)
    check(lib().msa_ke_packed(ctx.h, n, ptr(x, C.c_double), ptr(E, C.c_double), nu,
                              ptr(ke, C.c_double), ptr(st, C.c_int)))
    return ke, st


class Batch:
    """Many stochastic-modulus samples over one or more meshes (msa_batch)."""

    def __init__(self, ctx: Context, meshes, n_samples, precision=L.MSA_FP64, tile=0):
        self.ctx = ctx
        self.meshes = list(meshes)
        self.n_samples = [int(s) for s in n_samples]
        self.precision = precision
        arr = (C.c_void_p * len(self.meshes))(*[m.h for m in self.meshes])
        ns = np.array(self.n_samples, np.int32)
        h = C.c_void_p()
        check(lib().msa_batch_create(ctx.h, len(self.meshes), arr, ptr(ns, C.c_int), precision,
                                     tile, C.byref(h)))
        self.h = h
        ctx._adopt(self)

    @property
    def device_bytes(self) -> int:
        return int(lib().msa_batch_bytes(self.h))

    def set_modulus(self, cls: int, E):
        E = f64(E)
        if E.size != self.n_samples[cls] * self.meshes[cls].n_elem:
            raise L.MsaError(L.MSA_ERR_ARG, f"class {cls}: {E.size} moduli for "
                                            f"{self.n_samples[cls]} samples x "
                                            f"{self.meshes[cls].n_elem} elements")
        check(lib().msa_batch_set_modulus(self.h, cls, ptr(E, C.c_double)))

    def set_kl(self, cls: int, table, domain, meanE: float, cov: float):
        w = f64(table["w1d"])
        a = f64(table["a1d"])
        pos = i32(table["evpPos"])
        eig = f64(table["eigs"])
        dom = f64(domain)
        check(lib().msa_batch_set_kl(self.h, cls, int(table["M"]), int(table["kmax"]),
                                     ptr(w, C.c_double), ptr(a, C.c_double), ptr(pos, C.c_int),
                                     ptr(eig, C.c_double), ptr(dom, C.c_double), meanE, cov))

    def set_xi(self, cls: int, xi):
        xi = f64(xi)
        check(lib().msa_batch_set_xi(self.h, cls, ptr(xi, C.c_double)))

    def eval_kl(self):
        check(lib().msa_batch_eval_kl(self.h))

    def set_rhs(self, cls: int, f):
        f = f64(f)
        if f.size != self.meshes[cls].n_free:
            raise L.MsaError(L.MSA_ERR_ARG, f"class {cls}: load vector of {f.size} entries "
                                            f"for {self.meshes[cls].n_free} free DOFs")
        check(lib().msa_batch_set_rhs(self.h, cls, ptr(f, C.c_double)))

    def assemble(self, nu: float = 0.3):
        check(lib().msa_batch_assemble(self.h, nu))

    def solve(self, tol=1e-10, max_it=20000, inner_tol=1e-4, max_outer=20, graph_iters=0):
        check(lib().msa_batch_solve(self.h, tol, max_it, inner_tol, max_outer, graph_iters))

    def run(self, nu=0.3, tol=1e-10, max_it=20000, inner_tol=1e-4, max_outer=20, graph_iters=0):
        check(lib().msa_batch_run(self.h, nu, tol, max_it, inner_tol, max_outer, graph_iters))

    def solution(self, cls: int, sample: int):
        u = np.zeros(self.meshes[cls].n_free)
        check(lib().msa_batch_get_solution(self.h, cls, sample, ptr(u, C.c_double)))
        return u

    def stats(self, cls: int):
        n = self.n_samples[cls]
        it = np.zeros(n, np.int32)
        outer = np.zeros(n, np.int32)
        rel = np.zeros(n)
        st = np.zeros(n, np.int32)
        check(lib().msa_batch_get_stats(self.h, cls, ptr(it, C.c_int), ptr(outer, C.c_int),
                                        ptr(rel, C.c_double), ptr(st, C.c_int)))
        return dict(iters=it, outer=outer, relres=rel, status=st)

    def element_results(self, cls: int, sample: int):
        """strains [n_elem, 8, 6], stresses [n_elem, 8, 6], internal forces
        [n_elem, 24] of one solved sample (Hexa8.CalculateStresses /
        CalculateForces)."""
        ne = self.meshes[cls].n_elem
        eps, sig, f = np.zeros((ne, 8, 6)), np.zeros((ne, 8, 6)), np.zeros((ne, 24))
        check(lib().msa_batch_element_results(self.h, cls, sample, ptr(eps, C.c_double),
                                              ptr(sig, C.c_double), ptr(f, C.c_double)))
        return eps, sig, f

    def modulus(self, cls: int, sample: int):
        E = np.zeros(self.meshes[cls].n_elem)
        check(lib().msa_batch_get_modulus(self.h, cls, sample, ptr(E, C.c_double)))
        return E

    def matrix_csr(self, cls: int, sample: int):
        m = self.meshes[cls]
        nnz = m.info["nnz"]
        rp = np.zeros(m.n_free + 1, np.int64)
        col = np.zeros(nnz, np.int32)
        val = np.zeros(nnz)
        check(lib().msa_batch_get_matrix_csr(self.h, cls, sample, ptr(rp, C.c_longlong),
                                             ptr(col, C.c_int), ptr(val, C.c_double)))
        return rp, col, val

    def matrix_skyline(self, cls: int, sample: int):
        """Assembled K of one sample as SkylineMatrix2D arrays (rowIndex,
        data), e.g. for matrix_io.write_skyline."""
        from .matrix_io import csr_to_skyline
        rp, col, val = self.matrix_csr(cls, sample)
        return csr_to_skyline(rp, col, val, self.meshes[cls].n_free)

    def set_matrix_csr(self, cls: int, sample: int, rowptr, col, val):
        """Sets one sample's K from full symmetric CSR in the free-DOF
        numbering instead of assembling it (MSA_FP64 batches; every entry in
        the mesh's sparsity pattern)."""
        rp = np.ascontiguousarray(rowptr, np.int64)
        cc = np.ascontiguousarray(col, np.int32)
        vv = f64(val)
        n = self.meshes[cls].n_free
        if rp.size != n + 1 or cc.size < rp[-1] or vv.size < rp[-1]:
            raise L.MsaError(L.MSA_ERR_ARG, f"CSR of {rp.size - 1} rows / {cc.size} entries "
                                          f"for {n} free DOFs")
        check(lib().msa_batch_set_matrix_csr(self.h, cls, sample, ptr(rp, C.c_longlong),
                                             ptr(cc, C.c_int), ptr(vv, C.c_double)))

    def set_matrix_skyline(self, cls: int, sample: int, rowIndex, data):
        """Sets one sample's K from SkylineMatrix2D arrays (as read by
        matrix_io.read_skyline)."""
        from .matrix_io import skyline_to_csr
        self.set_matrix_csr(cls, sample, *skyline_to_csr(rowIndex, data, drop_zeros=True))

    def prepare_preconditioner(self):
        """Forms the preconditioner's values now instead of at the next solve."""
        check(lib().msa_batch_prepare_preconditioner(self.h))

    def spmv(self, cls: int, sample: int, x):
        x = f64(x)
        y = np.zeros(self.meshes[cls].n_free)
        check(lib().msa_batch_spmv(self.h, cls, sample, ptr(x, C.c_double), ptr(y, C.c_double)))
        return y

    def set_refinement(self, method=L.MSA_REFINE_RELIABLE, period: int = 8, drop: float = 0.1):
        check(lib().msa_batch_set_refinement(self.h, method, period, drop))

    def set_integration(self, precision=L.MSA_FP64):
        check(lib().msa_batch_set_integration(self.h, precision))

    # ------------------------------------------------------------ direct solver
    def factorize(self):
        """Band Cholesky of every sample's assembled FP64 K (SkylineMatrix2D.Factorize)."""
        check(lib().msa_batch_factorize(self.h))

    def set_zem(self, tol: float, max_modes: int = 16):
        """Zero-energy-mode extraction in the next factorization (tol 0: off)."""
        check(lib().msa_batch_set_zem(self.h, tol, max_modes))

    def zem(self, cls: int, sample: int, max_modes: int = 64):
        """(cols, modes[k][n_free]) of one sample's last factorization."""
        n = self.meshes[cls].n_free
        cnt = C.c_int(0)
        cols = np.zeros(max_modes, np.int32)
        modes = np.zeros((max_modes, n))
        check(lib().msa_batch_get_zem(self.h, cls, sample, max_modes, C.byref(cnt),
                                      ptr(cols, C.c_int), ptr(modes, C.c_double)))
        k = min(cnt.value, max_modes)
        return cols[:k].copy(), modes[:k].copy()

    def direct_solve(self):
        check(lib().msa_batch_direct_solve(self.h))

    def set_solver(self, solver):
        check(lib().msa_batch_set_solver(self.h, solver))

    def direct_info(self, cls: int):
        info = np.zeros(4, np.int64)
        check(lib().msa_batch_direct_info(self.h, cls, ptr(info, C.c_longlong)))
        return dict(n=int(info[0]), bandwidth=int(info[1]), width=int(info[2]),
                    bytes_per_sample=int(info[3]))

    # ------------------------------------------------------------ preconditioner
    def set_preconditioner(self, kind, max_levels: int = 0, ratio: float = 0.0):
        """MSA_PRECOND_JACOBI or MSA_PRECOND_MULTIGRID (geometric V-cycle on
        structured grids; classes whose mesh does not coarsen keep Jacobi)."""
        check(lib().msa_batch_set_preconditioner(self.h, kind, max_levels, ratio))

    def mg_info(self, cls: int):
        info = np.zeros(6, np.int64)
        check(lib().msa_batch_mg_info(self.h, cls, ptr(info, C.c_longlong)))
        return dict(levels=int(info[0]), coarsest_dofs=int(info[1]), coarse_dofs=int(info[2]),
                    coarsest_bandwidth=int(info[3]), matrix_free=bool(info[4]),
                    matrix_free_coarse_levels=int(info[5]))

    def set_smoothing_interval(self, lanczos_steps: int = 4, margin: float = 1.3):
        """Upper end of the Chebyshev smoothing interval: Lanczos estimate
        (lanczos_steps > 0) or the Gershgorin bound (0)."""
        check(lib().msa_batch_set_smoothing_interval(self.h, lanczos_steps, margin))

    def smoothing_interval(self):
        """(Lanczos steps, margin) of the smoothing interval."""
        k = C.c_int(0)
        m = C.c_double(0.0)
        check(lib().msa_batch_get_smoothing_interval(self.h, C.byref(k), C.byref(m)))
        return k.value, m.value

    def mg_lambda(self, cls: int, sample: int, max_levels: int = 8):
        lam = np.zeros(max_levels)
        check(lib().msa_batch_mg_lambda(self.h, cls, sample, max_levels, ptr(lam, C.c_double)))
        return lam[lam > 0]

    def mg_apply(self, cls: int, sample: int, r):
        """One V-cycle z = B r for one sample (reference DOF numbering)."""
        r = np.ascontiguousarray(r, np.float64)
        z = np.zeros_like(r)
        check(lib().msa_batch_mg_apply(self.h, cls, sample, ptr(r, C.c_double), ptr(z, C.c_double)))
        return z

    # ------------------------------------------------------------ dynamics
    def set_dynamics(self, rho: float, rayleigh_alpha: float = 0.0, rayleigh_beta: float = 0.0,
                     newmark_alpha: float = 0.25, newmark_delta: float = 0.5, dt: float = 1e-4):
        """Effective-stiffness mode of linear implicit Newmark
        (NewmarkDynamicAnalyzer.cs:53-77, ProblemStructural.cs:138-157)."""
        check(lib().msa_batch_set_dynamics(self.h, rho, rayleigh_alpha, rayleigh_beta,
                                           newmark_alpha, newmark_delta, dt))

    def set_probe(self, cls: int, dof: int):
        check(lib().msa_batch_set_probe(self.h, cls, dof))

    def newmark(self, acc, acc_dir: int = 2, tol: float = 1e-10, max_it: int = 20000,
                graph_iters: int = 16, warm_start: int = 1):
        """Runs len(acc) steps from rest; acc[i] is the base acceleration of
        step i (MassAccelerationSinusoidalLoad values).  warm_start: first
        iterate of each step's PCG, 0 zero, 1 previous displacement, 2 the
        predictor u + dt v + dt^2/2 a."""
        acc = f64(acc)
        self.nsteps = acc.size
        check(lib().msa_batch_newmark(self.h, acc.size, ptr(acc, C.c_double), acc_dir, tol,
                                      max_it, graph_iters, int(warm_start)))

    def history(self, cls: int, sample: int):
        h = np.zeros(self.nsteps)
        check(lib().msa_batch_get_history(self.h, cls, sample, ptr(h, C.c_double)))
        return h

    def dynamic_state(self, cls: int, sample: int):
        n = self.meshes[cls].n_free
        u, v, a = np.zeros(n), np.zeros(n), np.zeros(n)
        it = C.c_longlong()
        check(lib().msa_batch_get_dynamic_state(self.h, cls, sample, ptr(u, C.c_double),
                                                ptr(v, C.c_double), ptr(a, C.c_double),
                                                C.byref(it)))
        return u, v, a, it.value

    def set_kernel_timing(self, on: bool):
        check(lib().msa_batch_set_kernel_timing(self.h, 1 if on else 0))

    def kernel_timing(self):
        ms = np.zeros(6)
        n = np.zeros(6, np.int64)
        check(lib().msa_batch_get_kernel_timing(self.h, ptr(ms, C.c_double),
                                                ptr(n, C.c_longlong)))
        names = ["spmv_dot", "update", "assemble", "residual64", "other", "multigrid"]
        return {k: dict(ms=float(ms[i]), launches=int(n[i])) for i, k in enumerate(names)}

    def set_schedule(self, unroll: int):
        """Slots per group of the SpMV's rows (1..4 at lane tile 1; 3 otherwise)."""
        check(lib().msa_batch_set_schedule(self.h, int(unroll)))

    def set_matrix_cores(self, enable: bool):
        """Matrix-free smoothing on box grids: element products on the matrix
        cores (default) or on the vector ALU."""
        check(lib().msa_batch_set_matrix_cores(self.h, 1 if enable else 0))

    def graph_info(self):
        """Captured PCG iteration blocks: graphs captured (each validated on
        the host), kernel nodes of the current one, iterations per replay,
        replays so far."""
        info = np.zeros(4, np.int64)
        check(lib().msa_batch_graph_info(self.h, ptr(info, C.c_longlong)))
        return dict(captures=int(info[0]), kernel_nodes=int(info[1]),
                    iters_per_replay=int(info[2]), replays=int(info[3]))

    def layout_info(self, cls: int):
        """Stored layout of a class: stencil slot order (box grids), PCG
        product from the upper triangle, stored block slots per sample,
        block rows."""
        info = np.zeros(4, np.int64)
        check(lib().msa_batch_layout_info(self.h, cls, ptr(info, C.c_longlong)))
        return dict(stencil=bool(info[0]), upper_triangle_product=bool(info[1]),
                    slots=int(info[2]), block_rows=int(info[3]))

    def set_symmetric_product(self, enable: bool):
        """Upper-triangle storage and product on box grids (True, the
        default) or the full image (False); a change needs assembling
        again."""
        check(lib().msa_batch_set_symmetric_product(self.h, 1 if enable else 0))

    def set_matrix_free_product(self, enable: bool):
        """PCG product of box-grid classes matrix-free (True, the default:
        MIXED, one-sample tiles, multigrid preconditioner) or from the
        stored values (False); takes effect at the next solve."""
        check(lib().msa_batch_set_matrix_free_product(self.h, 1 if enable else 0))

    def product_form(self, cls: int) -> str:
        """The PCG product of a class at the last solve: "stored",
        "upper_triangle" or "matrix_free"."""
        f = C.c_int(0)
        check(lib().msa_batch_product_form(self.h, cls, C.byref(f)))
        return ("stored", "upper_triangle", "matrix_free")[f.value]

    def set_persistent(self, enable: bool):
        """Jacobi-PCG solves as one cooperative launch (one-sample tiles)."""
        check(lib().msa_batch_set_persistent(self.h, 1 if enable else 0))

    def persistent_info(self):
        info = np.zeros(2, np.int64)
        check(lib().msa_batch_persistent_info(self.h, ptr(info, C.c_longlong)))
        return dict(launches=int(info[0]), workgroups=int(info[1]))

    def time_spmv(self, reps: int = 20) -> float:
        """ms per launch of the SpMV + dot kernel over `reps` back-to-back
        launches with every sample active (overwrites the PCG iterates)."""
        ms = C.c_double()
        check(lib().msa_batch_time_spmv(self.h, int(reps), C.byref(ms)))
        return ms.value

    def spmv_bytes(self):
        b = np.zeros(2)
        check(lib().msa_batch_spmv_bytes(self.h, ptr(b, C.c_double)))
        return dict(kernel=float(b[0]), spmv=float(b[1]))

    def close(self):
        if getattr(self, "h", None):
            lib().msa_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune_tile(ctx: Context, mesh: Mesh, n_samples: int, precision=L.MSA_MIXED):
    best = C.c_int()
    ms = np.zeros(7)
    check(lib().msa_autotune_tile(ctx.h, mesh.h, n_samples, precision, C.byref(best),
                                  ptr(ms, C.c_double)))
    return best.value, dict(zip([1, 2, 4, 8, 16, 32, 64], ms.tolist()))


def autotune_schedule(ctx: Context, meshes, samples_per_class: int, precision=L.MSA_MIXED,
                      precond=L.MSA_PRECOND_MULTIGRID):
    """Times one assembly + solve per sample for every candidate (lane tile,
    slot unroll) on trial batches of `meshes`; returns dict(tile, unroll,
    table) with table[(tile, unroll)] = ms per sample (negative: skipped)."""
    arr = (C.c_void_p * len(meshes))(*[m.h for m in meshes])
    bt, bu = C.c_int(), C.c_int()
    ms = np.zeros(L.TUNE_CANDIDATES)
    check(lib().msa_autotune_schedule(ctx.h, len(meshes), arr, int(samples_per_class), precision,
                                      precond, C.byref(bt), C.byref(bu), ptr(ms, C.c_double)))
    tiles = np.zeros(L.TUNE_CANDIDATES, np.int32)
    unrolls = np.zeros(L.TUNE_CANDIDATES, np.int32)
    check(lib().msa_autotune_candidates(ptr(tiles, C.c_int), ptr(unrolls, C.c_int)))
    table = {(int(t), int(u)): float(v) for t, u, v in zip(tiles, unrolls, ms)}
    return dict(tile=bt.value, unroll=bu.value, table=table)
