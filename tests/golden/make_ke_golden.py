"""Writes tests/golden/ke_golden.json: element stiffness fixtures that pin the
Hexa8 Ke (oracle and HIP kernels) without running the reference.

Independent of the oracle's code: the textbook trilinear isoparametric brick
(N_a = (1 + xi_a xi)(1 + eta_a eta)(1 + zeta_a zeta)/8, node order of
Hexa8MeshGenerator.cs:1081-1088 = rElHexa8.f90:25-117), B with engineering
shear in Voigt order [xx, yy, zz, xy, yz, zx] (rElHexa8.f90:252-396), isotropic
D (ElasticMaterial3D.cs:20-44).

Cases:
  exact_*   : affine bricks integrated exactly by sympy (the 2x2x2 rule is
              exact there; the reference's 13-digit abscissa 0.5773502691896
              differs from 1/sqrt(3) by ~4.5e-14 relative, so these pin the
              result to ~1e-13).
  gauss50_* : distorted hexahedra, the 2x2x2 rule with the reference's
              13-digit constants evaluated in mpmath at 50 digits (pins the
              quadrature itself to ~1e-15).
Packed layout: upper triangle row-major, 300 entries (SymmetricMatrix2D.cs:
22-49).  Run: python tests/golden/make_ke_golden.py
"""
import json
import os

import mpmath as mp
import sympy as sp

NODES = [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1),
         (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)]


def d_matrix(E, nu):
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    D = [[0] * 6 for _ in range(6)]
    for i in range(3):
        for j in range(3):
            D[i][j] = lam + (2 * mu if i == j else 0)
    for i in range(3, 6):
        D[i][i] = mu
    return D


def b_matrix(grads):
    """grads[a] = (dNa/dx, dNa/dy, dNa/dz) -> 6x24."""
    B = [[0] * 24 for _ in range(6)]
    for a, (gx, gy, gz) in enumerate(grads):
        u, v, w = 3 * a, 3 * a + 1, 3 * a + 2
        B[0][u] = gx
        B[1][v] = gy
        B[2][w] = gz
        B[3][u], B[3][v] = gy, gx
        B[4][v], B[4][w] = gz, gy
        B[5][u], B[5][w] = gz, gx
    return B


def pack(K):
    return [K[i][j] for i in range(24) for j in range(i, 24)]


def exact_brick(origin, size, E, nu):
    x, y, z = sp.symbols("x y z")
    x0, y0, z0 = [sp.nsimplify(v) for v in origin]
    hx, hy, hz = [sp.nsimplify(v) for v in size]
    xi = 2 * (x - x0) / hx - 1
    eta = 2 * (y - y0) / hy - 1
    zeta = 2 * (z - z0) / hz - 1
    N = [(1 + a * xi) * (1 + b * eta) * (1 + c * zeta) / 8 for (a, b, c) in NODES]
    grads = [(sp.diff(n, x), sp.diff(n, y), sp.diff(n, z)) for n in N]
    B = b_matrix(grads)
    D = d_matrix(sp.nsimplify(E), sp.nsimplify(nu))
    DB = [[sum(D[m][k] * B[k][j] for k in range(6)) for j in range(24)] for m in range(6)]
    K = [[0] * 24 for _ in range(24)]
    for i in range(24):
        for j in range(i, 24):
            integrand = sp.expand(sum(B[m][i] * DB[m][j] for m in range(6)))
            val = sp.integrate(integrand, (x, x0, x0 + hx), (y, y0, y0 + hy), (z, z0, z0 + hz))
            K[i][j] = K[j][i] = val
    return [float(v) for v in pack(K)]


def gauss50(xyz, E, nu):
    mp.mp.dps = 50
    g = mp.mpf("0.5773502691896")
    pts = [-g, g]
    D = d_matrix(mp.mpf(E), mp.mpf(nu))
    K = [[mp.mpf(0)] * 24 for _ in range(24)]
    X = [[mp.mpf(repr(c)) for c in p] for p in xyz]
    for xi in pts:
        for eta in pts:
            for zeta in pts:
                dN = []
                for (a, b, c) in NODES:
                    dN.append((a * (1 + b * eta) * (1 + c * zeta) / 8,
                               b * (1 + a * xi) * (1 + c * zeta) / 8,
                               c * (1 + a * xi) * (1 + b * eta) / 8))
                # J[i][j] = d x_j / d natural_i
                J = mp.matrix(3, 3)
                for i in range(3):
                    for j in range(3):
                        J[i, j] = sum(dN[a][i] * X[a][j] for a in range(8))
                det = mp.det(J)
                Ji = J ** -1
                grads = []
                for a in range(8):
                    nat = mp.matrix([dN[a][0], dN[a][1], dN[a][2]])
                    gphys = Ji * nat  # dN/dx_j = sum_i Jinv[j][i] dN/dnat_i
                    grads.append((gphys[0], gphys[1], gphys[2]))
                B = b_matrix(grads)
                for i in range(24):
                    Di = [sum(D[m][k] * B[k][i] for k in range(6)) for m in range(6)]
                    for j in range(i, 24):
                        K[i][j] += sum(B[m][j] * Di[m] for m in range(6)) * det
    for i in range(24):
        for j in range(i):
            K[i][j] = K[j][i]
    return [float(v) for v in pack(K)]


def main():
    cases = []
    unit = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    cases.append(dict(name="exact_unit_cube", E=1.0, nu=0.3, xyz=[list(map(float, p)) for p in unit],
                      ke=exact_brick((0, 0, 0), (1, 1, 1), "1", "3/10")))
    o, s = (0.5, -0.25, 1.0), (0.125, 0.25, 0.375)
    brick = [(o[0] + (a > 0) * s[0], o[1] + (b > 0) * s[1], o[2] + (c > 0) * s[2]) for (a, b, c) in NODES]
    cases.append(dict(name="exact_brick", E=2.1e5, nu=0.3, xyz=[list(p) for p in brick],
                      ke=exact_brick(o, s, "210000", "3/10")))
    h = 1.0 / 16
    cube16 = [((a > 0) * h, (b > 0) * h, (c > 0) * h) for (a, b, c) in NODES]
    cases.append(dict(name="exact_cube_h16", E=2.1e5, nu=0.3, xyz=[list(p) for p in cube16],
                      ke=exact_brick((0, 0, 0), (h, h, h), "210000", "3/10")))
    dist1 = [(0.0, 0.0, 0.0), (1.1, 0.05, -0.02), (1.2, 0.9, 0.1), (-0.1, 1.05, 0.0),
             (0.05, -0.1, 0.95), (1.0, 0.0, 1.1), (1.15, 1.1, 0.9), (0.0, 0.95, 1.05)]
    cases.append(dict(name="gauss50_distorted_1", E=1.0, nu=0.25, xyz=[list(p) for p in dist1],
                      ke=gauss50(dist1, "1", "0.25")))
    dist2 = [(0.3, 0.1, 0.2), (0.55, 0.12, 0.18), (0.6, 0.4, 0.25), (0.28, 0.35, 0.22),
             (0.31, 0.08, 0.5), (0.58, 0.15, 0.52), (0.62, 0.42, 0.55), (0.27, 0.38, 0.49)]
    cases.append(dict(name="gauss50_distorted_2", E=2.1e5, nu=0.3, xyz=[list(p) for p in dist2],
                      ke=gauss50(dist2, "210000", "0.3")))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ke_golden.json")
    with open(out, "w") as fh:
        json.dump(dict(generator="tests/golden/make_ke_golden.py", cases=cases), fh, indent=0)
    print("wrote", out)


if __name__ == "__main__":
    main()
