"""Eigenvalue adjoints -- TEST INFRASTRUCTURE ONLY (the checker, never shipped).

Two restatements used to check the device adjoint:

* ``eig_densities`` -- the exact first-order derivative of the generalized
  Hermitian eigenvalues, de_n = c_n^H (dH - e_n dS) c_n, as the weighted
  densities QH = C diag(g) C^H, QS = -C diag(g e) C^H (numpy / LAPACK).
  Pinned by central finite differences in the tests.
* ``reference_gradients`` -- the reference's autograd path for eigenvalue
  gradients: eighb(scheme="chol") (slakonet/utils.py:740-783: L = chol(S),
  L^-1 = solve(L, I), C = L^-1 H L^-H) with torch's own autograd for those
  ops, and _SymEigB.backward (slakonet/utils.py:386-473) whose
  eigenvalue-only part is a_bar = V diag(w_bar) V^T (transpose, no complex
  conjugate, :466-469). Run with torch on the CPU.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla


def eig_densities(H: np.ndarray, S: np.ndarray, g: np.ndarray):
    """QH, QS [n, n] and eigenvalues for dL = Re sum QH[nu,mu] dH[mu,nu] + QS[nu,mu] dS[mu,nu]."""
    e, C = sla.eigh(H, S)
    QH = (C * g[None, :]) @ C.conj().T
    QS = -(C * (g * e)[None, :]) @ C.conj().T
    return QH, QS, e


def directional(Q: np.ndarray, D: np.ndarray) -> float:
    """Re sum_{mu,nu} Q[nu,mu] D[mu,nu]."""
    return float(np.real(np.sum(Q.T * D)))


def reference_gradients(H, S, g):
    """torch .grad of H and S for L = sum g * eigenvalues via the reference's
    autograd path (CPU, float64 / complex128)."""
    import torch

    class _RefSymEig(torch.autograd.Function):
        @staticmethod
        def forward(ctx, a):
            w, v = torch.linalg.eigh(a)
            ctx.save_for_backward(v)
            return w

        @staticmethod
        def backward(ctx, w_bar):
            (v,) = ctx.saved_tensors
            lam = torch.diag_embed(w_bar).to(v.dtype)
            return v @ lam @ v.transpose(-2, -1)

    Ht = torch.as_tensor(H).clone().requires_grad_(True)
    St = torch.as_tensor(S).clone().requires_grad_(True)
    n = Ht.shape[-1]
    L = torch.linalg.cholesky(St)
    Linv = torch.linalg.solve(L, torch.eye(n, dtype=Ht.dtype))
    Lt = Linv.transpose(-1, -2)
    if Lt.is_complex():
        Lt = Lt.conj()
    w = _RefSymEig.apply(Linv @ Ht @ Lt)
    loss = (w * torch.as_tensor(g, dtype=w.dtype)).sum()
    loss.backward()
    return Ht.grad.numpy(), St.grad.numpy()


def torch_directional(grad: np.ndarray, D: np.ndarray) -> float:
    """First-order change of a real loss along D from a torch .grad
    (torch's convention for complex inputs: dL = Re sum conj(grad) * D)."""
    return float(np.real(np.sum(np.conj(grad) * D)))
