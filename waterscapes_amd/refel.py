"""Exact reference-element tables for the P2 / P1 tetrahedron.

All element tensors of the total-pressure MPET forms
(mpet/mpettotalpressuresolver.py:267-283) are polynomial integrals over an
affine tetrahedron K.  Writing every basis function as a polynomial in the
barycentric coordinates lambda_0..lambda_3 turns each integral into

    |K| * sum_kl  W[a, b, k, l] * (grad lambda_k)_m (grad lambda_l)_n

with W a rational constant that depends only on the reference element.  The
tables are produced here in exact rational arithmetic (``fractions``) and
converted to float64 once; the GPU assembly kernel and the host code use the
same tables.

Conventions (UFC):
  * local vertices 0..3 sorted by global vertex index;
  * P2 local nodes 0..3 are the vertices, 4..9 the edges
    (2,3), (1,3), (1,2), (0,3), (0,2), (0,1);
  * face i of the tetrahedron is the triangle opposite vertex i.

Identity used for integrals:  (1/|K|) * int_K prod_s lambda_s^b_s
                              = prod_s b_s! * 3! / (sum_s b_s + 3)!
"""
from __future__ import annotations

from fractions import Fraction
from functools import lru_cache
from itertools import product
from math import factorial

import numpy as np

EDGES = ((2, 3), (1, 3), (1, 2), (0, 3), (0, 2), (0, 1))
FACES = ((1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2))

# ---------------------------------------------------------------------------
# tiny exact polynomial algebra in barycentric monomials
# ---------------------------------------------------------------------------


class BPoly:
    """Polynomial in n barycentric coordinates: {exponent tuple: Fraction}."""

    __slots__ = ("c", "n")

    def __init__(self, c=None, n=4):
        self.n = n
        self.c = {k: v for k, v in (c or {}).items() if v != 0}

    @classmethod
    def const(cls, v, n=4):
        return cls({(0,) * n: Fraction(v)}, n)

    @classmethod
    def lam(cls, i, n=4):
        e = [0] * n
        e[i] = 1
        return cls({tuple(e): Fraction(1)}, n)

    def __add__(self, o):
        if not isinstance(o, BPoly):
            o = BPoly.const(o, self.n)
        c = dict(self.c)
        for k, v in o.c.items():
            c[k] = c.get(k, 0) + v
        return BPoly(c, self.n)

    __radd__ = __add__

    def __neg__(self):
        return BPoly({k: -v for k, v in self.c.items()}, self.n)

    def __sub__(self, o):
        return self + (-o if isinstance(o, BPoly) else BPoly.const(-o, self.n))

    def __rsub__(self, o):
        return (-self) + o

    def __mul__(self, o):
        if not isinstance(o, BPoly):
            return BPoly({k: v * Fraction(o) for k, v in self.c.items()}, self.n)
        c = {}
        for k1, v1 in self.c.items():
            for k2, v2 in o.c.items():
                k = tuple(a + b for a, b in zip(k1, k2))
                c[k] = c.get(k, 0) + v1 * v2
        return BPoly(c, self.n)

    __rmul__ = __mul__

    def dlam(self, i):
        """Partial derivative with respect to lambda_i (lambdas independent)."""
        c = {}
        for k, v in self.c.items():
            if k[i] > 0:
                k2 = list(k)
                k2[i] -= 1
                c[tuple(k2)] = c.get(tuple(k2), 0) + v * k[i]
        return BPoly(c, self.n)

    def mean(self):
        """(1/|T|) int_T p over the simplex of dimension n-1."""
        d = self.n - 1
        s = Fraction(0)
        for k, v in self.c.items():
            num = 1
            for e in k:
                num *= factorial(e)
            s += v * Fraction(num * factorial(d), factorial(sum(k) + d))
        return s

    def __call__(self, lam):
        s = 0.0
        for k, v in self.c.items():
            t = float(v)
            for li, e in zip(lam, k):
                t *= li ** e
            s += t
        return s


def p1_basis(n=4):
    return [BPoly.lam(i, n) for i in range(n)]


def p2_basis():
    lam = p1_basis()
    vert = [lam[i] * (2 * lam[i] - 1) for i in range(4)]
    edge = [4 * lam[i] * lam[j] for (i, j) in EDGES]
    return vert + edge


def lagrange_multi_indices(k, n=4):
    """Multi-indices alpha (|alpha| = k) of the P_k Lagrange nodes, in a fixed
    order: vertices first, then lexicographic order of the rest."""
    if k == 0:
        return [tuple([0] * n)]
    allidx = [a for a in product(range(k + 1), repeat=n) if sum(a) == k]
    vert = [tuple(k if s == i else 0 for s in range(n)) for i in range(n)]
    rest = sorted(a for a in allidx if a not in vert)
    return vert + rest


def pk_basis(k, n=4):
    """P_k Lagrange basis on the simplex as barycentric polynomials.
    Node alpha sits at barycentric point alpha / k.  For k == 0 the single
    basis function is the constant 1 (node at the centroid)."""
    if k == 0:
        return [BPoly.const(1, n)]
    out = []
    for a in lagrange_multi_indices(k, n):
        p = BPoly.const(1, n)
        for s in range(n):
            for r in range(a[s]):
                p = p * (BPoly.lam(s, n) * k - r) * Fraction(1, r + 1)
        out.append(p)
    return out


def pk_nodes_bary(k, n=4):
    """Barycentric coordinates (nnodes, n) of the P_k nodes."""
    if k == 0:
        return np.full((1, n), 1.0 / n)
    return np.array(lagrange_multi_indices(k, n), dtype=np.float64) / k


# ---------------------------------------------------------------------------
# tables
# ---------------------------------------------------------------------------


@lru_cache(maxsize=None)
def grad_weights_p2p2():
    """W[a,b,k,l] = (1/|K|) int dphi_a/dlam_k * dphi_b/dlam_l  (P2 x P2).

    int_K d_m phi_a d_n phi_b = |K| sum_kl W[a,b,k,l] g_k[m] g_l[n]
    with g_k = grad lambda_k."""
    phi = p2_basis()
    dphi = [[p.dlam(k) for k in range(4)] for p in phi]
    W = np.zeros((10, 10, 4, 4))
    for a in range(10):
        for b in range(10):
            for k in range(4):
                for l in range(4):
                    W[a, b, k, l] = float((dphi[a][k] * dphi[b][l]).mean())
    return W


@lru_cache(maxsize=None)
def div_weights_p2p1():
    """V[b,c,k] = (1/|K|) int psi_c * dphi_b/dlam_k   (phi P2, psi P1).

    int_K psi_c d_m phi_b = |K| sum_k V[b,c,k] g_k[m]."""
    phi = p2_basis()
    psi = p1_basis()
    V = np.zeros((10, 4, 4))
    for b in range(10):
        for c in range(4):
            for k in range(4):
                V[b, c, k] = float((psi[c] * phi[b].dlam(k)).mean())
    return V


@lru_cache(maxsize=None)
def mass_p1():
    """(1/|K|) int psi_i psi_j = (1 + delta_ij) / 20."""
    psi = p1_basis()
    return np.array([[float((psi[i] * psi[j]).mean()) for j in range(4)]
                     for i in range(4)])


@lru_cache(maxsize=None)
def mass_p2():
    """(1/|K|) int phi_a phi_b for the P2 basis (10 x 10)."""
    phi = p2_basis()
    return np.array([[float((phi[a] * phi[b]).mean()) for b in range(10)]
                     for a in range(10)])


@lru_cache(maxsize=None)
def source_weights(test_degree, k):
    """Q[i, m] = (1/|K|) int v_i * phi^(k)_m  for the test space P_test_degree
    (1 or 2) and the P_k interpolant of a source (Expression(degree=k))."""
    test = p2_basis() if test_degree == 2 else p1_basis()
    src = pk_basis(k)
    return np.array([[float((t * s).mean()) for s in src] for t in test])


@lru_cache(maxsize=None)
def facet_source_weights(test_degree, k):
    """Facet analogue of source_weights on the reference triangle:
    Q[i, m] = (1/|F|) int_F v_i phi^(k)_m with the test functions restricted
    to the face (P2 face nodes: 3 vertices + 3 edges in FACE_P2 order)."""
    if test_degree == 2:
        lam = p1_basis(3)
        test = [lam[i] * (2 * lam[i] - 1) for i in range(3)] + \
               [4 * lam[i] * lam[j] for (i, j) in ((1, 2), (0, 2), (0, 1))]
    else:
        test = p1_basis(3)
    src = pk_basis(k, 3)
    return np.array([[float((t * s).mean()) for s in src] for t in test])


def face_local_p2_nodes(face):
    """Local P2 node ids (tet numbering) on face `face`: 3 vertices in face
    order then the 3 face edges in the triangle's (1,2),(0,2),(0,1) order."""
    vs = FACES[face]
    edge_ids = []
    for (i, j) in ((1, 2), (0, 2), (0, 1)):
        e = tuple(sorted((vs[i], vs[j])))
        edge_ids.append(4 + EDGES.index(e))
    return list(vs) + edge_ids


def bpoly_eval(p, bary):
    """Values of a barycentric polynomial at points (npts, n) of
    barycentric coordinates."""
    bary = np.asarray(bary, dtype=np.float64)
    out = np.zeros(bary.shape[0])
    for e, v in p.c.items():
        out += float(v) * np.prod(bary ** np.asarray(e, dtype=np.float64), axis=1)
    return out


@lru_cache(maxsize=None)
def gm_rule(degree):
    """Grundmann-Moeller rule on the tetrahedron (A. Grundmann, H. M.
    Moeller, SIAM J. Numer. Anal. 15 (1978) 282-290), exact for polynomials
    of degree <= 2 s + 1 >= degree.  Returns (bary (nq, 4), w (nq,)) with the
    weights normalised to sum 1, i.e. (1/|K|) int_K f ~ sum_q w_q f(x_q).
    Some weights are negative (as in the published rule)."""
    n = 3
    s = max(0, (int(degree) - 1 + 1) // 2)
    d = 2 * s + 1
    pts, wts = [], []
    for i in range(s + 1):
        w = Fraction((-1) ** i * (d + n - 2 * i) ** d,
                     2 ** (2 * s) * factorial(i) * factorial(d + n - i))
        for beta in product(range(s - i + 1), repeat=n + 1):
            if sum(beta) != s - i:
                continue
            pts.append([Fraction(2 * b + 1, d + n - 2 * i) for b in beta])
            wts.append(w)
    tot = sum(wts)
    return (np.array([[float(x) for x in p] for p in pts]),
            np.array([float(w / tot) for w in wts]))


def tables_float64():
    """All tables needed by the device kernels, packed in one float64 array
    with fixed offsets (see csrc/mpet_common.h)."""
    W = grad_weights_p2p2().reshape(-1)      # 1600
    V = div_weights_p2p1().reshape(-1)       # 160
    M2 = mass_p2().reshape(-1)               # 100
    return np.ascontiguousarray(np.concatenate([W, V, M2]))
