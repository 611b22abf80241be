"""The batched eigensolve under HIP graph capture.

engine.eigh (every stage a kernel or a device-to-device copy on the caller's
stream, no host synchronisation with check=False) is captured once into a
graph with torch.cuda.graph and replayed: each replay, with new matrices
copied into the captured inputs and the program's own eager solves and
synchronisations in between, must give the eager eigenvalues bit for bit.
The replay and eager times are printed: a replay saves 5-15 % of a small
solve (ld 24 x 16, ld 96 x 8, where the launch chain is comparable to the
work) and nothing measurable from ld 512 up (ratio 1.000 +- 0.015).
"""
import time

import pytest
import torch

from slakonet_amd import engine

pytestmark = pytest.mark.gpu


def _pair(ld, b, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    x = torch.randn(b, ld, ld, dtype=torch.complex128, device="cuda", generator=g)
    H = ((x + x.conj().transpose(1, 2)) / (2 * ld ** 0.5)).contiguous()
    y = torch.randn(b, ld, ld, dtype=torch.complex128, device="cuda", generator=g) * (0.1 / ld ** 0.5)
    S = (torch.eye(ld, dtype=torch.complex128, device="cuda") + torch.matmul(y, y.conj().transpose(1, 2).contiguous()))
    return H.contiguous(), S.contiguous()


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("ld,batch", [(24, 16), (96, 8), (1280, 2), (2432, 1)],
                         ids=["small", "two-stage-lds-band", "two-stage-hbm-band", "one-stage"])
def test_eigh_graph_replays_match_eager(ld, batch, precision):
    st = torch.cuda.Stream()
    H, S = _pair(ld, batch, 1)
    torch.cuda.synchronize()
    with torch.cuda.stream(st):
        engine.eigh(H, S, precision=precision, check=False)   # warm-up: one-time kernel attributes
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=st):
        wg, _, ig = engine.eigh(H, S, precision=precision, check=False)
    torch.cuda.synchronize()
    for rep, seed in enumerate((1, 2, 3, 2)):
        H2, S2 = _pair(ld, batch, seed)
        torch.cuda.synchronize()
        H.copy_(H2)
        S.copy_(S2)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        we, _, ie = engine.eigh(H2, S2, precision=precision, check=False)   # eager, between replays
        torch.cuda.synchronize()
        assert int(ig.abs().sum()) == 0 and int(ie.abs().sum()) == 0
        assert torch.equal(wg, we), f"replay {rep}: max |dw| {float((wg - we).abs().max()):.3e}"
    # timing, for the record
    reps = 3
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        g.replay()
    torch.cuda.synchronize()
    tg = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    with torch.cuda.stream(st):
        for _ in range(reps):
            engine.eigh(H, S, precision=precision, check=False)
    torch.cuda.synchronize()
    te = (time.perf_counter() - t0) / reps
    print(f"ld {ld} x {batch} {precision}: eager {1e3 * te:.3f} ms, graph replay {1e3 * tg:.3f} ms ({tg / te:.3f})")
