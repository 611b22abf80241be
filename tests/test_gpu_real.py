"""Real-symmetric batches (SLK_EIGH_REAL): H and S with zero imaginary parts,
as the build gives at k-points where every Bloch phase is +-1 (Gamma, the
1x1x1 Monkhorst-Pack point (1/2, 1/2, 1/2), ...). The real mode must give the
complex solve's eigenvalues up to rounding and LAPACK's real symmetric ones,
on every path of the solver, in both precisions; eigenvectors come out real.
"""
import numpy as np
import pytest
import scipy.linalg as sl
import torch

from slakonet_amd import _lib, engine

pytestmark = pytest.mark.gpu

FP64_ATOL = 1e-9    # Ha, O(1) spectra
MIXED_REL = 4e-6    # of the spectral radius (FP32 reduction and tridiagonalisation)


def _real_problem(n, batch, seed):
    rng = np.random.default_rng(seed)
    h = rng.normal(size=(batch, n, n))
    h = (h + h.transpose(0, 2, 1)) / (2 * np.sqrt(n))
    b = rng.normal(size=(batch, n, n))
    s = np.eye(n)[None] + 0.05 * (b @ b.transpose(0, 2, 1)) / n
    return h, s


def _solve(h, s, precision, real, mode=0, vectors=False):
    lib = _lib.lib()
    assert lib.slk_set_tridiag_mode(mode) == 0
    try:
        H = torch.from_numpy(h.astype(np.complex128)).to("cuda")
        S = torch.from_numpy(s.astype(np.complex128)).to("cuda")
        w, V, _ = engine.eigh(H, S, precision=precision, real=real, vectors=vectors)
        torch.cuda.synchronize()
        return w.cpu().numpy(), (V.cpu().numpy() if V is not None else None)
    finally:
        lib.slk_set_tridiag_mode(0)


# 33 and 97: the last trailing block has one row (a one-element reflector)
CASES = [(24, 6, 0), (33, 4, 0), (96, 4, 0), (97, 3, 0), (640, 2, 0), (1100, 2, 0), (300, 2, 1), (2432, 1, 0)]
IDS = ["small", "two-stage-33", "two-stage-lds-band", "two-stage-97", "two-stage-hbm-band", "two-stage-1100",
       "one-stage-forced", "one-stage"]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("n,batch,mode", CASES, ids=IDS)
def test_real_mode_matches_complex_and_lapack(n, batch, mode, precision):
    print(f"\ncase: real n={n} batch={batch} mode={mode} {precision}", flush=True)
    h, s = _real_problem(n, batch, seed=11 + n)
    wr, _ = _solve(h, s, precision, True, mode)
    wc, _ = _solve(h, s, precision, False, mode)
    ref = np.stack([sl.eigh(h[b], s[b], eigvals_only=True) for b in range(batch)])
    scale = float(np.abs(ref).max())
    e_ref = float(np.abs(wr[:, :n] - ref).max())
    e_cpx = float(np.abs(wr - wc).max())
    print(f"n={n}: |real - LAPACK| {e_ref:.2e}, |real - complex| {e_cpx:.2e} (radius {scale:.2f})")
    if precision == "fp64":
        assert e_ref < FP64_ATOL * max(1.0, scale)
        assert e_cpx < FP64_ATOL * max(1.0, scale)
    else:
        tol = MIXED_REL * max(1.0, (n / 1024) ** 0.5) * scale
        assert e_ref < tol and e_cpx < tol


@pytest.mark.parametrize("n", [200, 1300])
def test_real_mode_vectors(n):
    h, s = _real_problem(n, 2, seed=5 + n)
    w, V = _solve(h, s, "fp64", True, vectors=True)
    assert np.abs(V.imag).max() == 0.0
    for b in range(2):
        Z = V[b].real
        res = np.abs(h[b] @ Z - s[b] @ Z * w[b][None, :]).max()
        orth = np.abs(Z.T @ s[b] @ Z - np.eye(n)).max()
        print(f"n={n} b={b}: residual {res:.2e}, S-orthonormality {orth:.2e}")
        assert res < 1e-10 and orth < 1e-10
