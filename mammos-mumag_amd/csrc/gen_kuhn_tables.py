"""Generate kuhn_tables.h: exact P1 voxel matrices of the 6-tet Kuhn split of a unit cube.

Corner c = cx + 2 cy + 4 cz.  The six tetrahedra share the (0,0,0)-(1,1,1) diagonal and
follow the axis orders (x,y,z), (y,z,x), (z,x,y), (x,z,y), (z,y,x), (y,x,z), exactly as
mammos_mumag_amd.generate numbers them.  For a voxel of edge h:

  stiffness  S[a][c]    = sum_t int grad phi_a . grad phi_c          = h   * KS[a][c] / 6
  mass       M[a][c]    = sum_t int phi_a phi_c                      = h^3 * KM[a][c] / 120
  divergence B[a][c][d] = sum_t (vol_t/4) d_d phi_a  (t contains a, c) = h^2 * KB[a][c][d] / 24

All table entries are integers.  Run: ``python gen_kuhn_tables.py > kuhn_tables.h``.
"""
import itertools
from fractions import Fraction

PERMS = [(0, 1, 2), (1, 2, 0), (2, 0, 1), (0, 2, 1), (2, 1, 0), (1, 0, 2)]


def corner(c):
    return c[0] + 2 * c[1] + 4 * c[2]


def tets():
    out = []
    for p in PERMS:
        c = [0, 0, 0]
        nodes = [tuple(c)]
        for ax in p:
            c[ax] = 1
            nodes.append(tuple(c))
        out.append(nodes)
    return out


def inv3(m):
    a = [[Fraction(v) for v in row] for row in m]
    det = (a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0])
           + a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]))
    cof = [[(a[(i + 1) % 3][(j + 1) % 3] * a[(i + 2) % 3][(j + 2) % 3]
             - a[(i + 1) % 3][(j + 2) % 3] * a[(i + 2) % 3][(j + 1) % 3]) for j in range(3)] for i in range(3)]
    return [[cof[j][i] / det for j in range(3)] for i in range(3)], det


def tables():
    KS = [[Fraction(0)] * 8 for _ in range(8)]
    KM = [[Fraction(0)] * 8 for _ in range(8)]
    KB = [[[Fraction(0)] * 3 for _ in range(8)] for _ in range(8)]
    for t in tets():
        J = [[t[k + 1][a] - t[0][a] for a in range(3)] for k in range(3)]
        Jinv, det = inv3(J)
        vol = abs(det) / 6
        G = [[Jinv[a][k] for a in range(3)] for k in range(3)]  # grad phi_{k+1} = column k
        G = [[-(G[0][a] + G[1][a] + G[2][a]) for a in range(3)]] + G
        ids = [corner(c) for c in t]
        for k, l in itertools.product(range(4), range(4)):
            a, c = ids[k], ids[l]
            KS[a][c] += vol * sum(G[k][d] * G[l][d] for d in range(3))
            KM[a][c] += vol * (2 if k == l else 1) / 20
            for d in range(3):
                KB[a][c][d] += vol / 4 * G[k][d]
    return KS, KM, KB


def as_int(x, scale):
    v = x * scale
    assert v.denominator == 1, (x, scale)
    return int(v)


def main():
    KS, KM, KB = tables()
    print("// Generated by gen_kuhn_tables.py -- do not edit.")
    print("// Exact voxel matrices of the 6-tet Kuhn split (see the generator's docstring).")
    print("#pragma once")
    print("namespace mm {")
    print("constexpr int kKS[8][8] = {")
    for a in range(8):
        print("    {" + ", ".join(str(as_int(KS[a][c], 6)) for c in range(8)) + "},")
    print("};")
    print("constexpr int kKM[8][8] = {")
    for a in range(8):
        print("    {" + ", ".join(str(as_int(KM[a][c], 120)) for c in range(8)) + "},")
    print("};")
    print("constexpr int kKB[8][8][3] = {")
    for a in range(8):
        print("    {" + ", ".join("{" + ", ".join(str(as_int(KB[a][c][d], 24)) for d in range(3)) + "}"
                             for c in range(8)) + "},")
    print("};")
    print("}  // namespace mm")


if __name__ == "__main__":
    main()
