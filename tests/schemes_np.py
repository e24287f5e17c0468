"""Independent numpy restatements used to pin the oracle.

Nothing here is shared with the oracle's C code: the matrix-multiplication
tensor is built from its definition (C = A.B), and schemes such as Strassen's
are entered by hand from the textbook formulas.
"""
import numpy as np


def matmul_tensor(n1, n2, n3):
    """T[a, b, w] = 1 iff the product a*b contributes to output entry w, in the
    index layout the scheme files use: a = r*n2 + l (A is n1 x n2),
    b = l*n3 + q (B is n2 x n3), w = q*n1 + r (C^T, n3 x n1)."""
    T = np.zeros((n1 * n2, n2 * n3, n3 * n1), dtype=np.int64)
    for r in range(n1):
        for l in range(n2):
            for q in range(n3):
                T[r * n2 + l, l * n3 + q, q * n1 + r] = 1
    return T


def scheme_tensor(U, V, W):
    U = np.asarray(U, dtype=np.int64)
    V = np.asarray(V, dtype=np.int64)
    W = np.asarray(W, dtype=np.int64)
    return np.einsum("ta,tb,tc->abc", U, V, W)


def brent_ok(n, U, V, W, ring="Z2"):
    T = matmul_tensor(*n)
    S = scheme_tensor(U, V, W)
    if ring == "Z2":
        return bool(np.array_equal(S % 2, T))
    return bool(np.array_equal(S, T))


# Strassen (1969), entered from the textbook formulas.
# A index: a11=0 a12=1 a21=2 a22=3; B likewise; C^T index: c11=0 c21=1 c12=2 c22=3.
STRASSEN_U = [[1, 0, 0, 1], [0, 0, 1, 1], [1, 0, 0, 0], [0, 0, 0, 1], [1, 1, 0, 0], [-1, 0, 1, 0], [0, 1, 0, -1]]
STRASSEN_V = [[1, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, -1], [-1, 0, 1, 0], [0, 0, 0, 1], [1, 1, 0, 0], [0, 0, 1, 1]]
STRASSEN_W = [[1, 0, 0, 1], [0, 1, 0, -1], [0, 0, 1, 1], [1, 1, 0, 0], [-1, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0]]


def kron_scheme(s1, s2):
    """Tensor product of an (n1,n2,n3) scheme with an (p1,p2,p3) scheme:
    an (n1 p1, n2 p2, n3 p3) scheme of rank m1*m2.  Block index outer,
    inner index inner."""
    (n, U1, V1, W1), (p, U2, V2, W2) = s1, s2
    N = tuple(n[c] * p[c] for c in range(3))
    U, V, W = [], [], []
    for t1 in range(len(U1)):
        for t2 in range(len(U2)):
            u = [0] * (N[0] * N[1])
            v = [0] * (N[1] * N[2])
            w = [0] * (N[2] * N[0])
            for r1 in range(n[0]):
                for l1 in range(n[1]):
                    for r2 in range(p[0]):
                        for l2 in range(p[1]):
                            u[(r1 * p[0] + r2) * N[1] + l1 * p[1] + l2] = U1[t1][r1 * n[1] + l1] * U2[t2][r2 * p[1] + l2]
            for l1 in range(n[1]):
                for q1 in range(n[2]):
                    for l2 in range(p[1]):
                        for q2 in range(p[2]):
                            v[(l1 * p[1] + l2) * N[2] + q1 * p[2] + q2] = V1[t1][l1 * n[2] + q1] * V2[t2][l2 * p[2] + q2]
            for q1 in range(n[2]):
                for r1 in range(n[0]):
                    for q2 in range(p[2]):
                        for r2 in range(p[0]):
                            w[(q1 * p[2] + q2) * N[0] + r1 * p[0] + r2] = W1[t1][q1 * n[0] + r1] * W2[t2][q2 * p[0] + r2]
            U.append(u)
            V.append(v)
            W.append(w)
    return N, U, V, W


def strassen():
    return (2, 2, 2), STRASSEN_U, STRASSEN_V, STRASSEN_W


def strassen_squared():
    return kron_scheme(strassen(), strassen())


def to_blob(n, U, V, W):
    out = list(n) + [len(U)]
    for M in (U, V, W):
        for row in M:
            out.extend(int(x) for x in row)
    return np.array(out, dtype=np.int32)


def from_blob(blob):
    n1, n2, n3, m = (int(x) for x in blob[:4])
    nn = (n1 * n2, n2 * n3, n3 * n1)
    p = 4
    mats = []
    for c in range(3):
        mats.append(np.array(blob[p:p + m * nn[c]], dtype=np.int64).reshape(m, nn[c]))
        p += m * nn[c]
    return (n1, n2, n3), mats[0], mats[1], mats[2]


def equal_pairs(blob):
    """All (c, t1, t2) with t1 < t2 and equal component-c rows, in
    component / t1 / t2 order."""
    n, U, V, W = from_blob(blob)
    out = []
    for c, M in enumerate((U, V, W)):
        for a in range(M.shape[0]):
            for b in range(a + 1, M.shape[0]):
                if np.array_equal(M[a], M[b]):
                    out.append((c, a, b))
    return out
