"""Generate exact element-tensor fixtures with sympy (rational arithmetic).

For a few affine tetrahedra with rational vertices, integrate symbolically
over the reference tetrahedron (iterated integrals of polynomials):

  E  int eps(phi_a e_i) : eps(phi_b e_j)   (30x30, local DOF 3a+i)
  G  int grad phi_a . grad phi_b            (10x10, P2 Laplacian)
  B  int psi_c div(phi_a e_i)               (30x4)
  M  int psi_c psi_d                        (4x4, P1 mass)
  S  int grad psi_c . grad psi_d            (4x4, P1 stiffness)
  Q3 int psi_c * phi3_m  (P1 test x P3 Lagrange node basis, 4x20)

These are the element tensors of the forms at
mpet/mpettotalpressuresolver.py:267-283 (coefficients stripped).  P2 local
nodes follow UFC: vertices 0..3, edges (2,3),(1,3),(1,2),(0,3),(0,2),(0,1).
Run:  python tests/golden/make_element_fixtures.py  (writes element_tensors.npz)
"""
import os
from itertools import product

import numpy as np
import sympy as S

X, Y, Z = S.symbols("X Y Z")
EDGES = ((2, 3), (1, 3), (1, 2), (0, 3), (0, 2), (0, 1))

TETS = [
    [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)],
    [(0, 0, 0), (S.Rational(1, 2), 0, 0), (S.Rational(1, 2), S.Rational(1, 2), 0),
     (S.Rational(1, 2), S.Rational(1, 2), S.Rational(1, 2))],
    [(S.Rational(1, 3), 0, S.Rational(1, 5)), (2, S.Rational(1, 7), 0),
     (S.Rational(1, 2), S.Rational(3, 2), S.Rational(1, 4)), (0, S.Rational(1, 3), 2)],
    [(1, 1, 1), (S.Rational(5, 4), 1, S.Rational(9, 8)), (1, S.Rational(3, 2), 1),
     (S.Rational(7, 8), S.Rational(9, 8), S.Rational(5, 3))],
]


def ref_integral(expr):
    return S.integrate(S.integrate(S.integrate(expr, (Z, 0, 1 - X - Y)), (Y, 0, 1 - X)), (X, 0, 1))


def bases():
    lam = [1 - X - Y - Z, X, Y, Z]
    p2 = [lam[i] * (2 * lam[i] - 1) for i in range(4)] + [4 * lam[i] * lam[j] for (i, j) in EDGES]
    p1 = lam
    # P3 Lagrange basis: nodes = barycentric multi-indices / 3, vertices first
    allidx = [a for a in product(range(4), repeat=4) if sum(a) == 3]
    vert = [tuple(3 if s == i else 0 for s in range(4)) for i in range(4)]
    idx = vert + sorted(a for a in allidx if a not in vert)
    p3 = []
    for a in idx:
        p = S.Integer(1)
        for s in range(4):
            for r in range(a[s]):
                p *= (3 * lam[s] - r) / S.Integer(r + 1)
        p3.append(S.expand(p))
    return p2, p1, p3


def tensors(verts):
    v = [S.Matrix(p) for p in verts]
    J = S.Matrix.hstack(v[1] - v[0], v[2] - v[0], v[3] - v[0])
    det = abs(J.det())
    JinvT = J.inv().T
    p2, p1, p3 = bases()
    ref = (X, Y, Z)

    def grad(f):
        return JinvT * S.Matrix([S.diff(f, r) for r in ref])

    g2 = [grad(f) for f in p2]
    g1 = [grad(f) for f in p1]
    G = S.zeros(10, 10)
    Gm = {}
    for a in range(10):
        for b in range(10):
            for m in range(3):
                for n in range(3):
                    Gm[a, b, m, n] = det * ref_integral(S.expand(g2[a][m] * g2[b][n]))
            G[a, b] = sum(Gm[a, b, m, m] for m in range(3))
    E = S.zeros(30, 30)
    for a, b in product(range(10), range(10)):
        for i, j in product(range(3), range(3)):
            E[3 * a + i, 3 * b + j] = S.Rational(1, 2) * ((1 if i == j else 0) * G[a, b]
                                                          + Gm[a, b, j, i])
    B = S.zeros(30, 4)
    for a, i, c in product(range(10), range(3), range(4)):
        B[3 * a + i, c] = det * ref_integral(S.expand(p1[c] * g2[a][i]))
    M = S.zeros(4, 4)
    St = S.zeros(4, 4)
    for c, d in product(range(4), range(4)):
        M[c, d] = det * ref_integral(S.expand(p1[c] * p1[d]))
        St[c, d] = det * ref_integral(S.expand((g1[c].T * g1[d])[0]))
    Q3 = S.zeros(4, 20)
    for c, m in product(range(4), range(20)):
        Q3[c, m] = det * ref_integral(S.expand(p1[c] * p3[m]))
    f = lambda Mx: np.array(Mx.tolist(), dtype=np.float64)  # noqa: E731
    return f(E), f(G), f(B), f(M), f(St), f(Q3)


def main():
    out = {}
    coords = []
    for k, tv in enumerate(TETS):
        E, G, B, M, St, Q3 = tensors(tv)
        out[f"E{k}"], out[f"G{k}"], out[f"B{k}"] = E, G, B
        out[f"M{k}"], out[f"S{k}"], out[f"Q3_{k}"] = M, St, Q3
        coords.append(np.array([[float(c) for c in p] for p in tv]))
    out["coords"] = np.stack(coords)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "element_tensors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path)


if __name__ == "__main__":
    main()
