"""slk_eigh_mixed_noalloc (workspace in the library's static device array)
against slk_eigh_mixed (caller workspace), both through the C-ABI."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import solve as osolve
from slakonet_amd import _lib, engine

pytestmark = pytest.mark.gpu


def _problem(n, batch, seed):
    rng = np.random.default_rng(seed)
    h = rng.normal(size=(batch, n, n)) + 1j * rng.normal(size=(batch, n, n))
    h = (h + h.conj().transpose(0, 2, 1)) / (2 * np.sqrt(n))
    b = rng.normal(size=(batch, n, n)) + 1j * rng.normal(size=(batch, n, n))
    s = np.eye(n)[None] + 0.05 * (b @ b.conj().transpose(0, 2, 1)) / n
    return h, s


def _noalloc(lib, H, S, n, batch):
    w = torch.full((batch, n), float("nan"), dtype=torch.float64, device="cuda")
    info = torch.full((2 * batch,), 7, dtype=torch.int32, device="cuda")     # cleared by the call
    rc = lib.slk_eigh_mixed_noalloc(C.c_void_p(H.data_ptr()), C.c_void_p(S.data_ptr()), n, n, batch,
                                    C.c_void_p(w.data_ptr()), n, C.c_void_p(info.data_ptr()),
                                    _lib.stream_handle())
    torch.cuda.synchronize()
    return rc, w, info


def test_noalloc_matches_workspace_entry_point():
    lib = _lib.lib()
    cap = int(lib.slk_noalloc_scratch_bytes())
    assert cap == 64 << 20
    # two calls in a row, the second smaller (n = 576 x 3, then 200 x 2)
    for n, batch, seed in ((576, 3, 1), (200, 2, 2)):
        print(f"\ncase: noalloc n={n} batch={batch}", flush=True)
        assert lib.slk_eigh_mixed_workspace_bytes(n, n, batch) <= cap
        h, s = _problem(n, batch, seed)
        H = torch.from_numpy(h).to("cuda")
        S = torch.from_numpy(s).to("cuda")
        w_ref, _, _ = engine.eigh(H, S, precision="mixed")
        rc, w, info = _noalloc(lib, H, S, n, batch)
        assert rc == 0
        assert (info.cpu().numpy() == 0).all()
        diff = float((w - w_ref).abs().max())
        print(f"noalloc vs workspace entry point: max |dw| {diff:.3e} Ha", flush=True)
        assert diff <= 1e-12 * float(w_ref.abs().max())
        lap, _ = osolve.solve_all(h, s)
        assert np.abs(w.cpu().numpy() - lap).max() < 2e-6 * np.abs(lap).max()


def test_noalloc_refuses_sizes_beyond_its_scratch():
    lib = _lib.lib()
    n, batch = 2048, 1
    assert lib.slk_eigh_mixed_workspace_bytes(n, n, batch) > lib.slk_noalloc_scratch_bytes()
    H = torch.zeros(batch, n, n, dtype=torch.complex128, device="cuda")
    S = torch.zeros_like(H)
    rc, w, info = _noalloc(lib, H, S, n, batch)
    assert rc == 4
    assert torch.isnan(w).all() and (info.cpu().numpy() == 7).all()   # nothing was launched
