"""Independent dense assembly of the SIPG system (numpy).  TEST
INFRASTRUCTURE ONLY.

This module does not restate the reference kernels.  It builds the matrix
and right-hand side from the SIPG bilinear form itself (Thesis p.27 eq.
2.8-2.9 and p.31-33), so that a plausible slip in the matrix-free
restatement (a dropped term, a wrong sign, a transposed index) shows up as a
mismatch:

    a(u,v) = sum_K int_K grad u . grad v
           - sum_F int_F ({grad u}.[v] + {grad v}.[u])
           + sum_F int_F (sigma/h) [u].[v]
    l(v)   = sum_K int_K f v - sum_{F on boundary} int_F g (grad v . n)
           + sum_{F on boundary} int_F (sigma/h) g v

with GLL collocation on every element and face.  The nodes come from
numpy's Legendre module and the basis derivatives from barycentric
Lagrange weights, so nothing is shared with oracle/sipg_oracle.c.

Reference units: the 2-D stiffness is h-invariant, ds * d/dn is
h-invariant and ds*sigma/h = (sigma/2) w.  The reference kernel's `pen`
multiplies w_j [u] on every face (sipg_sem_2d_gpu_kernels.hpp:243-244,
308-309), so sigma = 2 * pen.

DoF numbering follows the reference: ((i1*(p+1)+i2)*noe + e), e = dimx*y+x,
y = 0 the top row, i2 = 0 an element's bottom edge (mode_vector.hpp:119,
mesh_info.hpp:231-309).
"""
import numpy as np
from numpy.polynomial import legendre as npleg


def gll_numpy(p):
    """GLL nodes/weights from numpy's Legendre roots (independent route)."""
    N = p
    c = np.zeros(N + 1)
    c[N] = 1.0
    interior = np.sort(npleg.legroots(npleg.legder(c))) if N >= 2 else np.array([])
    z = np.concatenate([[-1.0], interior, [1.0]])
    PN = npleg.legval(z, c)
    w = 2.0 / (N * (N + 1) * PN ** 2)
    return z, w


def lagrange_deriv(z):
    """Dl[j, i] = l_j'(z_i) by barycentric weights (Berrut & Trefethen 2004)."""
    n = len(z)
    lam = np.array([1.0 / np.prod([z[j] - z[k] for k in range(n) if k != j]) for j in range(n)])
    Dl = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if i != j:
                Dl[j, i] = (lam[j] / lam[i]) / (z[i] - z[j])
    for i in range(n):
        Dl[i, i] = -sum(Dl[k, i] for k in range(n) if k != i)
    return Dl


def assemble(dimx, dimy, p, pen):
    """Dense SIPG matrix in reference units and reference DoF order."""
    n1 = p + 1
    noe = dimx * dimy
    n = noe * n1 * n1
    z, w = gll_numpy(p)
    Dl = lagrange_deriv(z)          # Dl[mode, node]
    K1 = (Dl * w) @ Dl.T            # K1[a, c] = sum_k w_k l_a'(z_k) l_c'(z_k)
    sigma_half = pen                # (sigma/h) * ds = (sigma/2) w = pen * w

    def idx(a, b, e):
        return (a * n1 + b) * noe + e

    A = np.zeros((n, n))
    # volume: int grad u . grad v on each element
    for e in range(noe):
        for a in range(n1):
            for b in range(n1):
                r = idx(a, b, e)
                for c in range(n1):
                    A[r, idx(c, b, e)] += w[b] * K1[a, c]
                    A[r, idx(a, c, e)] += w[a] * K1[b, c]

    def face(J, G, wq):
        """Add -wq (J G^T + G J^T) + pen wq J J^T for one face point."""
        iJ = np.array(list(J.keys()), dtype=np.int64)
        vJ = np.array(list(J.values()))
        iG = np.array(list(G.keys()), dtype=np.int64)
        vG = np.array(list(G.values()))
        A[np.ix_(iJ, iG)] -= wq * np.outer(vJ, vG)
        A[np.ix_(iG, iJ)] -= wq * np.outer(vG, vJ)
        A[np.ix_(iJ, iJ)] += sigma_half * wq * np.outer(vJ, vJ)

    def add(d, k, v):
        d[k] = d.get(k, 0.0) + v

    for y in range(dimy):
        for x in range(dimx):
            e = dimx * y + x
            # vertical faces: right face of e, n = +x
            for j in range(n1):
                J, G = {}, {}
                if x + 1 < dimx:
                    er = e + 1
                    add(J, idx(p, j, e), 1.0)
                    add(J, idx(0, j, er), -1.0)
                    for a in range(n1):
                        add(G, idx(a, j, e), 0.5 * Dl[a, p])
                        add(G, idx(a, j, er), 0.5 * Dl[a, 0])
                else:                      # right boundary, outward n = +x
                    add(J, idx(p, j, e), 1.0)
                    for a in range(n1):
                        add(G, idx(a, j, e), Dl[a, p])
                face(J, G, w[j])
                if x == 0:                 # left boundary, outward n = -x
                    J, G = {}, {}
                    add(J, idx(0, j, e), 1.0)
                    for a in range(n1):
                        add(G, idx(a, j, e), -Dl[a, 0])
                    face(J, G, w[j])
            # horizontal faces: bottom face of e (row y+1 lies below), n = -y
            for i in range(n1):
                J, G = {}, {}
                if y + 1 < dimy:
                    eb = e + dimx
                    add(J, idx(i, 0, e), 1.0)
                    add(J, idx(i, p, eb), -1.0)
                    for b in range(n1):
                        add(G, idx(i, b, e), -0.5 * Dl[b, 0])
                        add(G, idx(i, b, eb), -0.5 * Dl[b, p])
                else:                      # bottom boundary, outward n = -y
                    add(J, idx(i, 0, e), 1.0)
                    for b in range(n1):
                        add(G, idx(i, b, e), -Dl[b, 0])
                face(J, G, w[i])
                if y == 0:                 # top boundary, outward n = +y
                    J, G = {}, {}
                    add(J, idx(i, p, e), 1.0)
                    for b in range(n1):
                        add(G, idx(i, b, e), Dl[b, p])
                    face(J, G, w[i])
    return A


def assemble_rhs(dim, p, pen, f, g):
    """l(v) for a dim x dim mesh on the unit square; f, g callables."""
    n1 = p + 1
    noe = dim * dim
    h = 1.0 / dim
    z, w = gll_numpy(p)
    Dl = lagrange_deriv(z)
    b = np.zeros(noe * n1 * n1)

    def idx(a, bb, e):
        return (a * n1 + bb) * noe + e

    for y in range(dim):
        for x in range(dim):
            e = dim * y + x
            x0, y0 = x * h, 1.0 - (y + 1) * h
            X = x0 + h * (z + 1) / 2
            Y = y0 + h * (z + 1) / 2
            for a in range(n1):
                for bb in range(n1):
                    b[idx(a, bb, e)] += (h / 2) ** 2 * w[a] * w[bb] * f(X[a], Y[bb])
            # boundary faces: -int g dv/dn + pen w g v
            if x == 0:
                for j in range(n1):
                    gv = g(x0, Y[j])
                    b[idx(0, j, e)] += pen * w[j] * gv
                    for a in range(n1):
                        b[idx(a, j, e)] -= w[j] * gv * (-Dl[a, 0])
            if x == dim - 1:
                for j in range(n1):
                    gv = g(x0 + h, Y[j])
                    b[idx(p, j, e)] += pen * w[j] * gv
                    for a in range(n1):
                        b[idx(a, j, e)] -= w[j] * gv * Dl[a, p]
            if y == dim - 1:
                for i in range(n1):
                    gv = g(X[i], y0)
                    b[idx(i, 0, e)] += pen * w[i] * gv
                    for bb in range(n1):
                        b[idx(i, bb, e)] -= w[i] * gv * (-Dl[bb, 0])
            if y == 0:
                for i in range(n1):
                    gv = g(X[i], y0 + h)
                    b[idx(i, p, e)] += pen * w[i] * gv
                    for bb in range(n1):
                        b[idx(i, bb, e)] -= w[i] * gv * Dl[bb, p]
    return b
