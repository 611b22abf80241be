"""Generalised Hermitian eigensolve per k-point (oracle, test infrastructure only).

Restates eighb(a, b, scheme="chol") (slakonet/utils.py:585-816) as used by
SimpleDftb.solve_kpoint (slakonet/main.py:137-163):
zero rows/columns of S get a unit diagonal (:725-746), L = chol(S) (:752),
L^-1 = solve(L, I) (:762), C = L^-1 H L^-H (:776), (w, y) = eigh(C)
(:372), z = L^-H y (:783). Ghost sorting (:516-582) is a no-op for unpadded
systems and is not reproduced. Eigenvalues ascending, Hartree.
"""
import numpy as np


def eighb_chol(h, s, vectors=True):
    h = np.asarray(h)
    s = np.asarray(s)
    n = h.shape[-1]
    zero_s = np.all(s == 0, axis=-1) & np.all(s == 0, axis=-2)
    if zero_s.any():
        s = s + np.eye(n) * zero_s
    zero_h = np.all(h == 0, axis=-1) & np.all(h == 0, axis=-2)
    if zero_h.any():
        h = h + np.eye(n) * zero_h
    l = np.linalg.cholesky(s)
    eye = np.broadcast_to(np.eye(n, dtype=h.dtype), h.shape)
    linv = np.linalg.solve(l, eye)
    linv_h = np.conj(np.swapaxes(linv, -1, -2))
    c = linv @ h @ linv_h
    if vectors:
        w, y = np.linalg.eigh(c)
        return w, linv_h @ y
    return np.linalg.eigvalsh(c), None


def solve_all(hk, sk, vectors=False):
    """hk, sk [nk, n, n] -> eigenvalues [nk, n] (and vectors [nk, n, n])."""
    ws, vs = [], []
    for ik in range(hk.shape[0]):
        w, v = eighb_chol(hk[ik], sk[ik], vectors=vectors)
        ws.append(w)
        vs.append(v)
    w = np.stack(ws)
    return (w, np.stack(vs)) if vectors else (w, None)
