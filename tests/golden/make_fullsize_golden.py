"""Writes tests/golden/fullsize_golden.npz: the CPU oracle's tightly
converged displacements of the full-size workloads, compressed to what the
GPU tests compare.

  c2      the 64^3 cantilever, E = 2.1e5 (811 200 free DOFs), oracle CSR
          Jacobi-PCG (orc_pcg_csr) to |r|/|f| < 1e-14
  c3_<s>  samples s = 0 and 255 of 256 lognormal-KL samples (COV 0.2,
          40 terms, correlation length 0.25, seed 20261205) of the 32^3
          cantilever, oracle per-sample loop (orc_mc_csr_jacobi) to 1e-14
  c4_<n>  the first sample of each mesh class n of the 1024-sample
          mixed-precision workload (sizes {16,24,32,40,48} drawn with seed
          20261202, COV 0.3), the same loop to 1e-12

Per displacement vector u (reference free-DOF numbering) the file holds:
  <key>_face   u at the free DOFs of the nodes on the loaded x = 1 face
  <key>_proj   g_k . u for k < 64, g_k ~ N(0, I) drawn in order from
               numpy Philox(<key seed>) (seed in <key>_seed)
  <key>_norm   |u|_2
The test regenerates g_k and estimates |u_gpu - u|_2 as
sqrt(mean_k (g_k . u_gpu - proj_k)^2), an unbiased estimate of the squared
norm with 64 degrees of freedom.

Run from the repository root (a few minutes of CPU, most of it the 64^3
solve):  python tests/golden/make_fullsize_golden.py
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fullsize_golden.npz")
NPROJ = 64
SEED_MC = 20261202
SEED_C3 = SEED_MC + 3


def face_dofs(cl):
    """Free DOFs of the nodes with x index == nex, ascending node, X/Y/Z."""
    nX = cl.nex + 1
    nodes = np.nonzero(np.arange(cl.n_nodes) % nX == cl.nex)[0]
    d = cl.dof.reshape(-1, 3)[nodes].ravel()
    return d[d >= 0]


def projections(u, seed):
    g = np.random.Generator(np.random.Philox(seed))
    return np.array([g.standard_normal(u.size) @ u for _ in range(NPROJ)])


def workload_counts(sizes=(16, 24, 32, 40, 48), samples=1024, seed=SEED_MC):
    rng = np.random.Generator(np.random.Philox(seed))
    draw = rng.integers(0, len(sizes), samples)
    return [int((draw == k).sum()) for k in range(len(sizes))]


def standard_normal(seed, n, M):
    return np.random.Generator(np.random.Philox(seed)).standard_normal((n, M))


def store(out, key, cl, u, seed):
    out[key + "_face"] = u[face_dofs(cl)]
    out[key + "_proj"] = projections(u, seed)
    out[key + "_norm"] = np.array([np.linalg.norm(u)])
    out[key + "_seed"] = np.array([seed])


def main():
    out = {}
    table = O.kl_table([1.0, 1.0, 1.0], 0.25, 40)
    # C4: first sample of each class, to 1e-12
    sizes = [16, 24, 32, 40, 48]
    counts = workload_counts()
    xi = standard_normal(SEED_MC, 1024, 40)
    off = 0
    for c, n in enumerate(sizes):
        t = time.time()
        r = O.mc_reference(n, table, xi[off:off + 1], cov=0.3, tol=1e-12, keep_u=True,
                           storage="csr")
        assert r["iters"][0] < 20000
        store(out, f"c4_{n}", O.Cantilever(n), r["u"][0], 1000 + n)
        print(f"c4 {n}^3 sample {off}: {r['iters'][0]} its, {time.time() - t:.1f} s", flush=True)
        off += counts[c]
    # C3: samples 0 and 255 of 256, to 1e-14
    xi3 = standard_normal(SEED_C3, 256, 40)
    for s in (0, 255):
        t = time.time()
        r = O.mc_reference(32, table, xi3[s:s + 1], cov=0.2, tol=1e-14, keep_u=True,
                           storage="csr")
        assert r["iters"][0] < 20000
        store(out, f"c3_{s}", O.Cantilever(32), r["u"][0], 3000 + s)
        print(f"c3 sample {s}: {r['iters'][0]} its, {time.time() - t:.1f} s", flush=True)
    # C2: 64^3, uniform modulus, to 1e-14
    t = time.time()
    cl = O.Cantilever(64)
    E = np.full(cl.n_elem, 2.1e5)
    ke, bad = O.ke_mesh(cl.conn, cl.xyz, E, 0.3)
    assert bad == 0
    rp, col, val = O.csr_assemble(cl.conn, cl.dof, cl.n, ke)
    r = O.pcg_csr(rp, col, val, cl.f, tol=1e-14, max_it=50000)
    assert r["converged"]
    store(out, "c2", cl, r["x"], 2064)
    print(f"c2 64^3: {r['iters']} its, {time.time() - t:.1f} s", flush=True)
    out["counts"] = np.array(counts)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
