"""Precision form against the FP64 oracle (MI355X only).

64-atom rocksalt (576 orbitals) on its full 4x4x4 Gamma-centred mesh: FP32
bulk (reduction + tridiagonalisation) with FP64 Cholesky, FP64 tridiagonal
eigenvalues and FP64 Rayleigh refinement of the band edges. The errors are
printed (pytest -s) and bounded here.
"""
import numpy as np
import pytest
import torch

from oracle import build as obuild, post as opost, solve as osolve
from slakonet_amd import engine, synth
from slakonet_amd.tables import TableSet
from slakonet_amd.units import AA_BOHR

pytestmark = pytest.mark.gpu

EIG_TOL_EV = 1e-3        # every eigenvalue, unrefined precision form
EDGE_TOL_EV = 1e-6       # refined VBM / CBM and gap
DOS_L1_REL_TOL = 1e-3    # integral |DOS_mixed - DOS_oracle| / integral DOS_oracle


def test_rocksalt_64k_precision_form_vs_oracle():
    st = synth.rocksalt_64()
    skfs = synth.make_skf_set([12, 8], seed=1)
    tab = TableSet(skfs, synth.shell_dict_spd([8, 12]), device="cuda")
    inp = engine.StructureInput(st.z, st.positions * AA_BOHR, st.cell * AA_BOHR, st.kpoints, st.k_weights)
    plan = engine.plan_build(tab, [inp])
    H, S = engine.run_build(tab, plan)
    grid = torch.linspace(-20.0, 10.0, 2001, dtype=torch.float64, device="cuda")
    wsl = []
    wm, _, _ = engine.eigh(H, S, precision="mixed", workspace=wsl)
    raw = wm.clone()
    edges, dos, _ = engine.post(plan, wm, dos_grid=grid, refine=(H, S, wsl[0]))
    ed = engine.edges_dict(edges[0].cpu().numpy())
    # oracle: FP64 numpy build + LAPACK on the CPU
    hs = obuild.build_hs(st.z, st.positions * AA_BOHR, st.cell * AA_BOHR, synth.shell_dict_spd([8, 12]), skfs,
                         st.kpoints)
    w_ref, _ = osolve.solve_all(hs["H"], hs["S"])
    ne = obuild.n_electrons(st.z, skfs)
    mu = opost.fermi_bisect(w_ref, ne, st.k_weights)
    gap_ref = opost.band_gap(w_ref, mu)
    _, dos_ref = opost.dos(w_ref, st.k_weights, mu, energy_range=(-20.0, 10.0), num_points=2001, sigma=0.1)
    eig_err = float(np.abs(raw.cpu().numpy() - w_ref).max()) * 27.211
    dos_np = dos[0].cpu().numpy()
    e = grid.cpu().numpy()
    l1 = float(np.trapezoid(np.abs(dos_np - dos_ref), e) / np.trapezoid(dos_ref, e))
    print(f"\nrocksalt 576x576, 64 k: max |eig err| {eig_err:.3e} eV (unrefined FP32 bulk); "
          f"VBM err {abs(ed['vbm'] - gap_ref['vbm']):.3e} eV, CBM err {abs(ed['cbm'] - gap_ref['cbm']):.3e} eV "
          f"(refined); gap {ed['cbm'] - ed['vbm']:.6f} eV vs {gap_ref['cbm'] - gap_ref['vbm']:.6f}; "
          f"DOS L1 rel {l1:.3e}")
    assert eig_err < EIG_TOL_EV
    if not gap_ref["metal"]:
        assert abs(ed["vbm"] - gap_ref["vbm"]) < EDGE_TOL_EV
        assert abs(ed["cbm"] - gap_ref["cbm"]) < EDGE_TOL_EV
    assert l1 < DOS_L1_REL_TOL


@pytest.mark.parametrize("n_atoms,length", [(320, 16.0), (1000, 25.0)])
def test_one_stage_edge_refinement_large(n_atoms, length):
    """Above 2560 orbitals the precision form takes the one-stage reduction, and
    the edge refinement's back-transform runs on 1024 threads per edge (2880
    orbitals: reflector prefetch; 9000: without). Refined VBM / CBM against the
    all-FP64 device path on the same H and S (Gamma only)."""
    st = synth.large_cubic(n_atoms=n_atoms, length=length)
    print(f"\ncase: cubic cell, {n_atoms} atoms, n = {9 * n_atoms}", flush=True)
    species = sorted({int(z) for z in st.z})
    tab = TableSet(synth.make_skf_set(species, seed=5), synth.shell_dict_spd(species), device="cuda")
    inp = engine.StructureInput(st.z, st.positions * AA_BOHR, st.cell * AA_BOHR, st.kpoints[:1], np.ones(1))
    plan = engine.plan_build(tab, [inp])
    H, S = engine.run_build(tab, plan)
    wsl = []
    wm, _, _ = engine.eigh(H, S, precision="mixed", workspace=wsl)
    raw = engine.edges_dict(engine.post(plan, wm.clone())[0][0].cpu().numpy())
    em, _, _ = engine.post(plan, wm, refine=(H, S, wsl[0]))
    wf, _, _ = engine.eigh(H, S, precision="fp64")
    ef, _, _ = engine.post(plan, wf)
    dm, df = engine.edges_dict(em[0].cpu().numpy()), engine.edges_dict(ef[0].cpu().numpy())
    err_v, err_c = abs(dm["vbm"] - df["vbm"]), abs(dm["cbm"] - df["cbm"])
    print(f"VBM err {err_v:.2e} eV, CBM err {err_c:.2e} eV (unrefined {abs(raw['vbm'] - df['vbm']):.2e}, "
          f"{abs(raw['cbm'] - df['cbm']):.2e})")
    assert err_v < 1e-8 and err_c < 1e-8
