"""Continuous-batching engine: results in the mix equal results alone."""
import numpy as np
import pytest
import torch

from slakonet_amd import engine, synth
from slakonet_amd.serving import ServingEngine, bucket_dim
from slakonet_amd.tables import TableSet
from slakonet_amd.units import AA_BOHR

pytestmark = pytest.mark.gpu

TOL_EV = 1e-7      # eigenvalue-derived scalars, eV (FP64 path)


def _stream(n, seed=31, n_max=48):
    sts = [synth.stream_structure(seed, i, n_max=n_max) for i in range(n)]
    species = sorted({int(z) for st in sts for z in st.z})
    tab = TableSet(synth.make_skf_set(species, seed=3), synth.shell_dict_spd(species), device="cuda")
    inputs = [engine.StructureInput(st.z, st.positions * AA_BOHR, st.cell * AA_BOHR, st.kpoints, st.k_weights)
              for st in sts]
    return tab, inputs


def test_bucket_dims():
    assert bucket_dim(18) == 32 and bucket_dim(257) == 320 and bucket_dim(2304) == 2304
    assert all(bucket_dim(n) >= n for n in range(1, 3000, 7))


def test_mix_equals_alone():
    tab, inputs = _stream(24)
    grid = torch.linspace(-20.0, 10.0, 2001, dtype=torch.float64, device="cuda")
    eng = ServingEngine(tab, dos_grid=grid, keep_outputs=True, precision="fp64")
    arrivals = np.zeros(len(inputs))
    arrivals[12:] = 0.002          # a second wave arrives while the first is on the device
    out = eng.replay(inputs, arrivals)
    assert len(out["results"]) == len(inputs)
    for i, inp in enumerate(inputs):
        res = out["results"][i]
        assert res["ok"]
        plan = engine.plan_build(tab, [inp])
        H, S = engine.run_build(tab, plan)
        w, _, _ = engine.eigh(H, S)
        edges, dos, stats = engine.post(plan, w, fermi_mode=0, dos_grid=grid)
        ed = engine.edges_dict(edges[0].cpu().numpy())
        for key in ("vbm", "cbm", "vb2", "cb2"):
            assert abs(res[key] - ed[key]) < TOL_EV, (i, key)
        assert abs(res["mu"] - ed["mu"]) * 27.211 < TOL_EV
        n = plan.n_orb[0]
        assert torch.allclose(res["eigenvalues"], w[:, :n], atol=1e-9, rtol=0)
        assert torch.allclose(res["dos"], dos[0], atol=1e-7, rtol=1e-7)


MIXED_TOL_EV = 1e-3   # gap / band-edge tolerance of the precision form (1 meV)


def test_mixed_precision_stream_within_1mev():
    """Precision form in the mix vs the FP64 path alone: band edges and gap
    within 1 meV, every eigenvalue within 1 meV."""
    tab, inputs = _stream(16, seed=41, n_max=96)
    eng = ServingEngine(tab, keep_outputs=True, precision="mixed")
    out = eng.replay(inputs, np.zeros(len(inputs)))
    worst = 0.0
    for i, inp in enumerate(inputs):
        res = out["results"][i]
        plan = engine.plan_build(tab, [inp])
        H, S = engine.run_build(tab, plan)
        w, _, _ = engine.eigh(H, S, precision="fp64")
        edges, _, _ = engine.post(plan, w, fermi_mode=0)
        ed = engine.edges_dict(edges[0].cpu().numpy())
        if ed["metal"] == 0 and res["metal"] == 0:
            assert abs(res["vbm"] - ed["vbm"]) < MIXED_TOL_EV
            assert abs(res["cbm"] - ed["cbm"]) < MIXED_TOL_EV
        n = plan.n_orb[0]
        err = float((res["eigenvalues"] - w[:, :n]).abs().max()) * 27.211
        worst = max(worst, err)
    assert worst < MIXED_TOL_EV, worst


REFINED_TOL_EV = 1e-6   # band edges of the precision form after FP64 Rayleigh refinement


def test_mixed_refined_band_edges_match_fp64():
    """Precision form: VBM/CBM after the FP64 Rayleigh-quotient refinement
    agree with the FP64 path to 1e-6 eV, alone and batched with padding."""
    tab, inputs = _stream(12, seed=43, n_max=128)
    worst_raw = worst_ref = 0.0
    refs = []
    for inp in inputs:
        plan = engine.plan_build(tab, [inp])
        H, S = engine.run_build(tab, plan)
        w64, _, _ = engine.eigh(H, S, precision="fp64")
        e64 = engine.edges_dict(engine.post(plan, w64)[0][0].cpu().numpy())
        refs.append(e64)
        wsl = []
        wm, _, _ = engine.eigh(H, S, precision="mixed", workspace=wsl)
        raw = engine.edges_dict(engine.post(plan, wm.clone())[0][0].cpu().numpy())
        ref = engine.edges_dict(engine.post(plan, wm, refine=(H, S, wsl[0]))[0][0].cpu().numpy())
        if e64["n_val"] > 0 and e64["n_con"] > 0:
            for key in ("vbm", "cbm"):
                worst_raw = max(worst_raw, abs(raw[key] - e64[key]))
                worst_ref = max(worst_ref, abs(ref[key] - e64[key]))
                assert abs(ref[key] - e64[key]) < REFINED_TOL_EV, (key, ref[key], e64[key])
    assert worst_ref <= worst_raw
    # batched with padding, through the serving engine
    eng = ServingEngine(tab, keep_outputs=True, precision="mixed", refine_edges=True)
    out = eng.replay(inputs, np.zeros(len(inputs)))
    for i, e64 in enumerate(refs):
        if e64["n_val"] > 0 and e64["n_con"] > 0:
            for key in ("vbm", "cbm"):
                assert abs(out["results"][i][key] - e64[key]) < REFINED_TOL_EV, (i, key)


def test_rocksalt_576_refined_edges():
    """64-atom rocksalt (576 orbitals), 8 of its k-points: refined edges within 1e-6 eV of FP64."""
    st = synth.rocksalt_64()
    skfs = synth.make_skf_set([12, 8], seed=1)
    tab = TableSet(skfs, synth.shell_dict_spd([8, 12]), device="cuda")
    k = st.kpoints[::8]
    inp = engine.StructureInput(st.z, st.positions * AA_BOHR, st.cell * AA_BOHR, k, np.full(len(k), 1 / len(k)))
    plan = engine.plan_build(tab, [inp])
    H, S = engine.run_build(tab, plan)
    w64, _, _ = engine.eigh(H, S, precision="fp64")
    e64 = engine.edges_dict(engine.post(plan, w64)[0][0].cpu().numpy())
    wsl = []
    wm, _, _ = engine.eigh(H, S, precision="mixed", workspace=wsl)
    raw_err = float((wm - w64).abs().max()) * 27.211
    ref = engine.edges_dict(engine.post(plan, wm, refine=(H, S, wsl[0]))[0][0].cpu().numpy())
    assert raw_err < 1e-3
    for key in ("vbm", "cbm"):
        assert abs(ref[key] - e64[key]) < REFINED_TOL_EV, (key, ref[key], e64[key])


def test_failed_requests_are_reported_not_ok():
    """A chase forced to give up its wait (negative spin limit) fails the
    requests of its batches; the others (n <= 32, solved without a chase) stay
    ok, and the replay finishes."""
    from slakonet_amd import _lib
    tab, inputs = _stream(10, seed=43, n_max=24)
    grid = torch.linspace(-20.0, 10.0, 2001, dtype=torch.float64, device="cuda")
    eng = ServingEngine(tab, dos_grid=grid, precision="mixed")
    lib = _lib.lib()
    prev = lib.slk_set_chase_spin_limit(-1)
    try:
        out = eng.replay(inputs, np.zeros(len(inputs)))
    finally:
        lib.slk_set_chase_spin_limit(prev)
    n_orb = [eng._orbitals(i) for i in inputs]
    assert any(n > 32 for n in n_orb) and any(n <= 32 for n in n_orb)
    for i, n in enumerate(n_orb):
        res = out["results"][i]
        if bucket_dim(n) > 32:
            assert not res["ok"] and "bulge chase" in res["error"], (i, n, res)
        else:
            assert res["ok"], (i, n)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_graph_replays_equal_eager_serving(precision):
    """Small buckets replay their eigensolve from a captured graph from the
    second batch of a shape on: every request's eigenvalues, edges and DOS
    equal those of the same stream served with graphs off, bit for bit."""
    tab, base = _stream(3, seed=57, n_max=24)
    inputs = [base[i % 3] for i in range(12)]
    arrivals = np.arange(12) * 0.015     # one request per batch (max_batch_matrices=1): shapes repeat
    grid = torch.linspace(-20.0, 10.0, 2001, dtype=torch.float64, device="cuda")
    outs = {}
    for graphs in (False, True):
        eng = ServingEngine(tab, dos_grid=grid, keep_outputs=True, precision=precision, n_streams=1,
                            max_batch_matrices=1, graphs=graphs)
        outs[graphs] = eng.replay(inputs, arrivals)["results"]
        eng.close()
        bad = [r.get("error") for r in outs[graphs].values() if not r["ok"]]
        assert not bad, bad
        if graphs:
            # per shape: eager, capture (+ replay), then two replays into the persistent buffers
            print(f"captures {eng.graph_captures}, replays {eng.graph_replays}")
            assert eng.graph_captures == 3 and eng.graph_replays == 9
    for i in range(len(inputs)):
        a, b = outs[False][i], outs[True][i]
        assert a["ok"] and b["ok"]
        assert torch.equal(a["eigenvalues"], b["eigenvalues"]), i
        assert torch.equal(a["dos"], b["dos"]), i
        for key in ("vbm", "cbm", "mu", "gap_eV"):
            assert a[key] == b[key], (i, key)
