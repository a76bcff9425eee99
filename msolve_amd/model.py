"""Model building on the host: the structured Hexa8 cantilever of the
workload, its free-DOF numbering and nodal loads.

Node numbering follows Hexa8MeshGenerator.BuildNodes
(ISAAR.MSolve.SimpleMeshGenerator/Hexa8MeshGenerator.cs:508-568: x fastest,
coordinates accumulated by repeated addition of the element size), element
node order Hexa8MeshGenerator.cs:1080-1089, DOF numbering Model.cs:165-215,
loads Model.cs:227-238 (each node's load is assigned, not added).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .core import enumerate_dofs


def _accumulated(n_nodes_axis: int, h: float) -> np.ndarray:
    """0, h, h+h, ... by repeated addition (the generator's `xyz += d`)."""
    out = np.zeros(n_nodes_axis)
    acc = 0.0
    for k in range(n_nodes_axis):
        out[k] = acc
        acc += h
    return out


def box_mesh(nex: int, ney: int, nez: int, lx=1.0, ly=1.0, lz=1.0):
    """Node coordinates [(nex+1)(ney+1)(nez+1), 3] and 0-based connectivity
    [nex*ney*nez, 8] of a structured brick mesh."""
    nX, nY, nZ = nex + 1, ney + 1, nez + 1
    x = _accumulated(nX, lx / nex)
    y = _accumulated(nY, ly / ney)
    z = _accumulated(nZ, lz / nez)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    xyz = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    i, j, k = np.meshgrid(np.arange(nez), np.arange(ney), np.arange(nex), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    n = lambda ii, jj, kk: nX * nY * ii + nX * jj + kk
    conn = np.stack([n(i, j, k), n(i, j, k + 1), n(i, j + 1, k + 1), n(i, j + 1, k),
                     n(i + 1, j, k), n(i + 1, j, k + 1), n(i + 1, j + 1, k + 1),
                     n(i + 1, j + 1, k)], axis=1).astype(np.int32)
    return xyz, conn


@dataclass
class CantileverModel:
    """Box clamped on x = 0 (all DOFs), total load `load` in Z spread over
    the nodes of the x = lx face."""
    nex: int
    ney: int
    nez: int
    xyz: np.ndarray
    conn: np.ndarray
    constrained: np.ndarray
    dof: np.ndarray
    n_free: int
    f: np.ndarray
    domain: np.ndarray

    @property
    def n_elem(self):
        return self.conn.shape[0]

    @property
    def n_nodes(self):
        return self.xyz.shape[0]

    def element_xyz(self):
        """[n_elem, 8, 3] node coordinates per element."""
        return self.xyz[self.conn]


def cantilever(nex: int, ney: int | None = None, nez: int | None = None, lx=1.0, ly=1.0,
               lz=1.0, load=-100.0) -> CantileverModel:
    ney = nex if ney is None else ney
    nez = nex if nez is None else nez
    xyz, conn = box_mesh(nex, ney, nez, lx, ly, lz)
    nX, nY, nZ = nex + 1, ney + 1, nez + 1
    constrained = np.zeros((nX * nY * nZ, 3), np.uint8)
    kk = np.arange(nX * nY * nZ) % nX
    constrained[kk == 0, :] = 1
    dof, n_free = enumerate_dofs(constrained)
    f = np.zeros(n_free)
    face = np.nonzero(kk == nex)[0]
    amount = load / face.size
    for node in face:
        g = dof[3 * node + 2]
        if g >= 0:
            f[g] = amount
    return CantileverModel(nex, ney, nez, xyz, conn, constrained.ravel(), dof, n_free, f,
                           np.array([lx, ly, lz]))
