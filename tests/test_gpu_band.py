"""Two-stage tridiagonalisation of the eigenvalue path (dense -> band of
width 16 -> tridiagonal by bulge chasing) against LAPACK on the same matrices
and against the blocked one-stage path (MI355X only).

Sizes cover every launch shape of the band reduction: fewer rows than one
panel (n = 33), ragged last panels, the register-resident panel QR tiers and
the global-memory panel QR (FP64 beyond 256 trailing rows, FP32 beyond 2048),
split-K products over several 256-column chunks, and identity-padded batches
(exact block structure through both stages).
"""
import numpy as np
import pytest
import torch

from oracle import solve as osolve
from slakonet_amd import _lib, engine

pytestmark = pytest.mark.gpu

FP64_ATOL = 1e-9          # Ha, eigenvalues of O(10) spectra
MIXED_REL = 2e-6          # of the spectral radius, FP32 reduction


def _problem(n, batch, seed):
    rng = np.random.default_rng(seed)
    h = rng.normal(size=(batch, n, n)) + 1j * rng.normal(size=(batch, n, n))
    h = (h + h.conj().transpose(0, 2, 1)) / (2 * np.sqrt(n))
    b = rng.normal(size=(batch, n, n)) + 1j * rng.normal(size=(batch, n, n))
    s = np.eye(n)[None] + 0.05 * (b @ b.conj().transpose(0, 2, 1)) / n
    return h, s


def _solve(h, s, precision, mode):
    lib = _lib.lib()
    assert lib.slk_set_tridiag_mode(mode) == 0
    try:
        H = torch.from_numpy(h).to("cuda")
        S = torch.from_numpy(s).to("cuda")
        w, _, info = engine.eigh(H, S, precision=precision)
        torch.cuda.synchronize()
        return w.cpu().numpy()
    finally:
        lib.slk_set_tridiag_mode(0)


def _lapack(h, s):
    w, _ = osolve.solve_all(h, s)
    return w


@pytest.mark.parametrize("n", [33, 47, 64, 100, 250, 300, 517, 1100, 1296])
def test_two_stage_fp64_vs_lapack(n):
    print(f"\ncase: two-stage fp64 n={n}", flush=True)
    h, s = _problem(n, 2, seed=n)
    ref = _lapack(h, s)
    w2 = _solve(h, s, "fp64", 2)
    w1 = _solve(h, s, "fp64", 1)
    err2 = float(np.abs(w2 - ref).max())
    err12 = float(np.abs(w2 - w1).max())
    print(f"n={n}: |two-stage - LAPACK| {err2:.2e} Ha, |two-stage - one-stage| {err12:.2e} Ha")
    assert err2 < FP64_ATOL
    assert err12 < FP64_ATOL


@pytest.mark.parametrize("n", [33, 100, 517, 1100, 2100, 2304])
def test_two_stage_mixed_vs_lapack(n):
    print(f"\ncase: two-stage mixed n={n}", flush=True)
    h, s = _problem(n, 1, seed=7 + n)
    ref = _lapack(h, s)
    w2 = _solve(h, s, "mixed", 2)
    scale = float(np.abs(ref).max())
    err = float(np.abs(w2 - ref).max())
    print(f"n={n}: |two-stage mixed - LAPACK| {err:.2e} Ha (spectral radius {scale:.2f})")
    # the FP32 rounding of the reduction accumulates ~ sqrt(n) on these random
    # matrices (measured 1.8e-6 of the radius at n = 2100, 2.1e-6 at 2304)
    assert err < MIXED_REL * max(1.0, (n / 1024) ** 0.5) * scale


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_two_stage_padded_batch(precision):
    """Identity-padded matrices of different sizes in one batch: the structure
    block's eigenvalues match LAPACK on the unpadded block, the padding gives
    the pad value to the bisection tolerance."""
    print(f"\ncase: two-stage padded batch {precision}", flush=True)
    ld, pad = 160, 4.0
    sizes = [160, 97, 130, 40]
    h = np.zeros((len(sizes), ld, ld), complex)
    s = np.zeros((len(sizes), ld, ld), complex)
    refs = []
    for b, n in enumerate(sizes):
        hb, sb = _problem(n, 1, seed=100 + n)
        h[b, :n, :n], s[b, :n, :n] = hb[0], sb[0]
        h[b, n:, n:] = pad * np.eye(ld - n)
        s[b, n:, n:] = np.eye(ld - n)
        refs.append(_lapack(hb, sb)[0])
    w = _solve(h, s, precision, 2)
    for b, n in enumerate(sizes):
        tol = FP64_ATOL if precision == "fp64" else MIXED_REL * float(np.abs(refs[b]).max())
        err = float(np.abs(w[b, :n] - refs[b]).max())
        assert err < tol, (b, n, err)
        if n < ld:
            # the pad eigenvalue is exact in the tridiagonal; bisection resolves
            # it to 4 ulp (FP64) or 2^-36 relative (mixed)
            pad_tol = 1e-12 if precision == "fp64" else 2.0 ** -36 * pad * 2
            assert float(np.abs(w[b, n:] - pad).max()) < pad_tol


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_degenerate_spectrum(precision):
    """Eigenvalues of multiplicity 1 to 4 (H = Q D Q^H with a random unitary
    Q, S = I): the Sturm-count grid puts several eigenvalues in one cell and
    the bisection must still resolve each index to its repeated value."""
    print(f"\ncase: degenerate spectrum {precision}", flush=True)
    rng = np.random.default_rng(11)
    n = 200
    vals = []
    lev = -3.0
    while len(vals) < n:
        vals += [lev] * int(rng.integers(1, 5))
        lev += float(rng.uniform(0.01, 0.2))
    d = np.array(vals[:n])
    z = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    q, _ = np.linalg.qr(z)
    h = (q * d) @ q.conj().T
    h = 0.5 * (h + h.conj().T)
    s = np.eye(n, dtype=complex)
    w = _solve(h[None], s[None], precision, 2)[0]
    ref = np.sort(d)
    scale = float(np.abs(ref).max())
    tol = FP64_ATOL if precision == "fp64" else MIXED_REL * scale
    err = float(np.abs(w[:n] - ref).max())
    print(f"|w - exact| {err:.2e} Ha (tol {tol:.1e})")
    assert err < tol


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_chase_give_up_is_reported(precision):
    """A bulge chase that gives up its bounded wait (forced from the host with
    a negative spin limit; every kernel still runs to its end, the waves
    merely stop waiting) flags the matrix with info[batch + b] = 2 and NaN
    eigenvalues, and the binding raises instead of returning them."""
    print(f"\ncase: chase give-up {precision}", flush=True)
    lib = _lib.lib()
    h, s = _problem(200, 2, seed=7)
    H = torch.from_numpy(h).to("cuda")
    S = torch.from_numpy(s).to("cuda")
    prev = lib.slk_set_chase_spin_limit(-1)
    try:
        w, _, info = engine.eigh(H, S, precision=precision, check=False)
        torch.cuda.synchronize()
        inf = info.cpu().numpy()
        assert (inf[:2] == 0).all() and (inf[2:] == 2).all()
        assert torch.isnan(w[:, :200]).all()
        with pytest.raises(engine.EigenFailure, match="bulge chase"):
            engine.eigh(H, S, precision=precision)
    finally:
        lib.slk_set_chase_spin_limit(prev)
    w, _, info = engine.eigh(H, S, precision=precision)
    assert np.abs(w.cpu().numpy() - _lapack(h, s)).max() < (FP64_ATOL if precision == "fp64" else 1e-5)
