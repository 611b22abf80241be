"""Multi-VASP training entry point on the MI355X: vasprun files with known
gaps, the default FCC k-path, the reference's natoms-weighted gap loss."""
import numpy as np
import pytest
import torch

from slakonet_amd import synth, vasp
from slakonet_amd.optim import MultiElementSkfParameterOptimizer
from vasp_fixture import write_vasprun

pytestmark = pytest.mark.gpu

SP = [6, 14]


def _write(path, species, frac, lattice, gap):
    eig = np.array([[[-5.0, 0.0, gap, gap + 3.0]]])
    occ = np.array([[[2.0, 2.0, 0.0, 0.0]]])
    e = np.linspace(-10, 10, 101)
    write_vasprun(path, lattice, species, frac, eig, occ, e, np.exp(-e[None] ** 2), [-10.0])


def test_multi_vasp_training_loss_and_saves(tmp_path):
    a_si, a_sic = 5.43, 4.36
    fcc = 0.5 * np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]])
    quarter = [[0, 0, 0], [0.25, 0.25, 0.25]]
    _write(tmp_path / "vasprun-si.xml", ["Si", "Si"], quarter, fcc * a_si, 1.1)
    _write(tmp_path / "vasprun-sic.xml", ["Si", "C"], quarter, fcc * a_sic, 2.3)
    shell = synth.shell_dict_spd(SP)
    model = MultiElementSkfParameterOptimizer(skf_dicts=synth.make_skf_set(SP, seed=7), device="cuda").to("cuda")
    kl = vasp.get_klines_example()
    # the loss of the first epoch, from the drop-in properties dict one structure at a time; every
    # epoch starts with the clamp, which moves tables that sit outside it (idempotent afterwards)
    model.apply_constraints()
    loader = vasp.MultiVaspDataLoader(str(tmp_path / "vasprun-*.xml"), log=None)
    with torch.no_grad():
        gaps = [model.compute_multi_element_properties(geometry=d["geometry"], shell_dict=shell, klines=kl)[0]
                ["band_gap_eV"] for d in loader.datasets]
    gaps = [float(torch.as_tensor(g).reshape(-1)[0]) for g in gaps]
    w = [d["metadata"]["natoms"] for d in loader.datasets]
    expect = sum(wi * abs(g - d["target_bandgap"]) for wi, g, d in zip(w, gaps, loader.datasets)) / sum(w)
    out = tmp_path / "run"
    model, hist, dl = vasp.train_multi_vasp_skf_parameters(model, str(tmp_path / "vasprun-*.xml"), num_epochs=4,
                                                           learning_rate=1e-3, save_directory=str(out),
                                                           shell_dict=shell, log=None)
    assert len(dl) == 2 and len(hist) == 4
    losses = [h["total_loss"] for h in hist]
    assert all(np.isfinite(losses))
    assert abs(hist[0]["batch_loss"] - expect) < 1e-6 * max(1.0, expect)
    assert (out / "best_model" / "model_state.pt").exists() and (out / "training_history.json").exists()
