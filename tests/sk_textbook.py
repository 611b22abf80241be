"""Textbook Slater-Koster two-centre table (Slater & Koster, Phys. Rev. 94,
1498 (1954), Table I) for s, p, d orbitals, written independently of the
oracle's rotation-matrix form. Orbital order: p = (y, z, x),
d = (xy, yz, 3z^2-r^2, zx, x^2-y^2). Used by the tests to pin the oracle.
"""
import numpy as np

S3 = np.sqrt(3.0)


def block(l1, l2, u, V):
    """[2l1+1, 2l2+1] block for l1 <= l2; u unit vector row->column atom;
    V = (sigma, pi, delta) integrals of the (l1, l2) pair."""
    l, m, n = u
    s = V[0]
    p = V[1] if len(V) > 1 else 0.0
    d = V[2] if len(V) > 2 else 0.0
    pc = {"x": l, "y": m, "z": n}
    P = ["y", "z", "x"]
    r = n * n - 0.5 * (l * l + m * m)
    q = l * l - m * m
    dsig = [S3 * l * m, S3 * m * n, r, S3 * l * n, 0.5 * S3 * q]
    if (l1, l2) == (0, 0):
        return np.array([[s]])
    if (l1, l2) == (0, 1):
        return np.array([[pc[a] * s for a in P]])
    if (l1, l2) == (0, 2):
        return np.array([[g * s for g in dsig]])
    if (l1, l2) == (1, 1):
        return np.array([[pc[a] * pc[b] * (s - p) + (p if a == b else 0.0) for b in P] for a in P])
    if (l1, l2) == (1, 2):
        def pd(a):
            c = pc[a]
            if a == "x":
                return [S3 * l * l * m * s + m * (1 - 2 * l * l) * p,
                        S3 * l * m * n * s - 2 * l * m * n * p,
                        l * r * s - S3 * l * n * n * p,
                        S3 * l * l * n * s + n * (1 - 2 * l * l) * p,
                        0.5 * S3 * l * q * s + l * (1 - q) * p]
            if a == "y":
                return [S3 * m * m * l * s + l * (1 - 2 * m * m) * p,
                        S3 * m * m * n * s + n * (1 - 2 * m * m) * p,
                        m * r * s - S3 * m * n * n * p,
                        S3 * l * m * n * s - 2 * l * m * n * p,
                        0.5 * S3 * m * q * s - m * (1 + q) * p]
            return [S3 * l * m * n * s - 2 * l * m * n * p,
                    S3 * n * n * m * s + m * (1 - 2 * n * n) * p,
                    n * r * s + S3 * n * (l * l + m * m) * p,
                    S3 * n * n * l * s + l * (1 - 2 * n * n) * p,
                    0.5 * S3 * n * q * s - n * q * p]
        return np.array([pd(a) for a in P])
    # d-d
    L2, M2, N2 = l * l, m * m, n * n
    t = np.zeros((5, 5))
    t[0, 0] = 3 * L2 * M2 * s + (L2 + M2 - 4 * L2 * M2) * p + (N2 + L2 * M2) * d
    t[1, 1] = 3 * M2 * N2 * s + (M2 + N2 - 4 * M2 * N2) * p + (L2 + M2 * N2) * d
    t[3, 3] = 3 * N2 * L2 * s + (N2 + L2 - 4 * N2 * L2) * p + (M2 + N2 * L2) * d
    t[0, 1] = 3 * l * M2 * n * s + l * n * (1 - 4 * M2) * p + l * n * (M2 - 1) * d
    t[1, 3] = 3 * m * N2 * l * s + m * l * (1 - 4 * N2) * p + m * l * (N2 - 1) * d
    t[0, 3] = 3 * L2 * m * n * s + m * n * (1 - 4 * L2) * p + m * n * (L2 - 1) * d
    t[0, 4] = 1.5 * l * m * q * s + 2 * l * m * (M2 - L2) * p + 0.5 * l * m * q * d
    t[1, 4] = 1.5 * m * n * q * s - m * n * (1 + 2 * q) * p + m * n * (1 + 0.5 * q) * d
    t[3, 4] = 1.5 * n * l * q * s + n * l * (1 - 2 * q) * p - n * l * (1 - 0.5 * q) * d
    t[0, 2] = S3 * l * m * r * s - 2 * S3 * l * m * N2 * p + 0.5 * S3 * l * m * (1 + N2) * d
    t[1, 2] = S3 * m * n * r * s + S3 * m * n * (L2 + M2 - N2) * p - 0.5 * S3 * m * n * (L2 + M2) * d
    t[3, 2] = S3 * l * n * r * s + S3 * l * n * (L2 + M2 - N2) * p - 0.5 * S3 * l * n * (L2 + M2) * d
    t[4, 4] = 0.75 * q * q * s + (L2 + M2 - q * q) * p + (N2 + 0.25 * q * q) * d
    t[4, 2] = 0.5 * S3 * q * r * s + S3 * N2 * (M2 - L2) * p + 0.25 * S3 * (1 + N2) * q * d
    t[2, 2] = r * r * s + 3 * N2 * (L2 + M2) * p + 0.75 * (L2 + M2) ** 2 * d
    for i in range(5):
        for j in range(i):
            if t[i, j] == 0.0 and t[j, i] != 0.0:
                t[i, j] = t[j, i]
            elif t[j, i] == 0.0:
                t[j, i] = t[i, j]
    return t
