"""Batched fitting (training loop) on the MI355X."""
import numpy as np
import pytest
import torch

from slakonet_amd import engine, fit, synth
from slakonet_amd.optim import MultiElementSkfParameterOptimizer
from slakonet_amd.tables import TableSet

pytestmark = pytest.mark.gpu

SPECIES = (6, 14, 31)


def _small_batch(n=4):
    sts = synth.fitting_batch(seed=3, n_cells=12, species=SPECIES)
    return sorted(sts, key=lambda s: len(s.z))[:n]


def test_batched_properties_equal_single_structure():
    sts = _small_batch(5)
    sp = sorted(SPECIES)
    tab = TableSet(synth.make_skf_set(sp, seed=7), synth.shell_dict_spd(sp), device="cuda")
    grid = torch.linspace(-20.0, 10.0, 2001, dtype=torch.float64, device="cuda")
    ds = fit.teacher_datasets(sts, tab, grid)
    for d in ds:
        plan = engine.plan_build(tab, [d.inp])
        H, S = engine.run_build(tab, plan)
        w, _, _ = engine.eigh(H, S)
        edges, dos, _ = engine.post(plan, w, dos_grid=grid)
        row = engine.edges_dict(edges[0].cpu().numpy())
        gap = 0.0 if row["metal"] > 0 else row["cbm"] - row["vbm"]
        assert torch.allclose(d.target_dos, dos[0], rtol=1e-8, atol=1e-8)
        assert abs(d.target_gap - gap) < 1e-8


def test_training_reduces_loss_and_saves(tmp_path):
    sts = _small_batch(4)
    sp = sorted(SPECIES)
    grid = torch.linspace(-20.0, 10.0, 2001, dtype=torch.float64, device="cuda")
    teacher = TableSet(synth.make_skf_set(sp, seed=7, teacher_seed=8), synth.shell_dict_spd(sp), device="cuda")
    ds = fit.teacher_datasets(sts, teacher, grid)
    model = MultiElementSkfParameterOptimizer(skf_dicts=synth.make_skf_set(sp, seed=7), device="cuda").to("cuda")
    model, hist = fit.train_skf_parameters(model, ds, synth.shell_dict_spd(sp), num_epochs=6, learning_rate=1e-3,
                                           loss="dos_gap", save_directory=str(tmp_path), log=None)
    losses = [h["total_loss"] for h in hist]
    assert len(losses) == 6 and all(np.isfinite(losses))
    assert min(losses[1:]) < losses[0]
    assert (tmp_path / "best_model" / "model_state.pt").exists()
    assert (tmp_path / "final_model.pt").exists()
    back = MultiElementSkfParameterOptimizer.load_model(tmp_path / "final_model", method="compact")
    a = {k: v.cpu() for k, v in model.state_dict().items()}
    assert all(torch.equal(a[k], v) for k, v in back.state_dict().items())
