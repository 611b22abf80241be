"""The drop-in properties dict and the interpolation tail, HIP path vs the oracle.

compute_multi_element_properties(geometry, shell_dict, kpoints, get_fermi=True)
(optim.py:949-1074; keys main.py:1703-1801, optim.py:1057-1073) is compared
key by key with oracle.post.properties evaluated on the oracle's own build and
LAPACK eigenvalues, in both Fermi modes, on the 2-atom diamond cell (8x8x8
Monkhorst-Pack) and a 3-species random cell.
"""
import numpy as np
import pytest
import torch

from oracle import build as obuild, post as opost, solve as osolve
from slakonet_amd import engine, synth
from slakonet_amd.geometry import Geometry
from slakonet_amd.optim import MultiElementSkfParameterOptimizer
from slakonet_amd.tables import TableSet
from slakonet_amd.units import AA_BOHR, H2E

pytestmark = pytest.mark.gpu

EIG_ATOL = 1e-9        # Hartree
E_ATOL = 1e-7          # eV, for edge / width / Fermi entries (EIG_ATOL x 27.211 plus bisection width)
DOS_RTOL = 1e-8


def _case(name):
    if name == "diamond":
        st = synth.diamond_si()
        return st.z, st.positions, st.cell, [14], [8, 8, 8], 0
    rng = np.random.default_rng(31)
    species = [1, 22, 31]
    z, pos, cell = synth.random_cell(rng, 12, species, 1.8)
    return z, pos, cell, species, [2, 2, 2], 5


def _oracle_props(z, pos, cell, species, mesh, seed, mode):
    from slakonet_amd.geometry import monkhorst_pack
    skfs = synth.make_skf_set(species, seed=seed)
    k, w = monkhorst_pack(mesh)
    ref = obuild.build_hs(z, pos * AA_BOHR, cell * AA_BOHR, synth.shell_dict_spd(species), skfs, k)
    wr, _ = osolve.solve_all(ref["H"], ref["S"])
    ne = obuild.n_electrons(z, skfs)
    return wr, w, opost.properties(wr, w, ne, get_fermi=True, fermi_method=mode), skfs


def _tie_ok(wr, k_index, value_ev, tol):
    """The chosen k-point holds a level at the reported energy (symmetric
    k-points are equal only to rounding, so either of them may be picked)."""
    return np.abs(wr[int(k_index)] * H2E - value_ev).min() < tol


@pytest.mark.parametrize("grad", [False, True], ids=["plain", "grad"])
@pytest.mark.parametrize("mode", ["bisection", "reference"])
@pytest.mark.parametrize("name", ["diamond", "random3"])
def test_properties_dict_matches_oracle(name, mode, grad):
    z, pos, cell, species, mesh, seed = _case(name)
    wr, kw, ref, skfs = _oracle_props(z, pos, cell, species, mesh, seed, mode)
    model = MultiElementSkfParameterOptimizer(skf_dicts=skfs, device="cuda")
    geom = Geometry(z, pos, cell=cell, units="angstrom")
    with torch.set_grad_enabled(grad):
        props, ok = model.compute_multi_element_properties(geometry=geom, shell_dict=synth.shell_dict_spd(species),
                                                           kpoints=mesh, get_fermi=True, device="cuda",
                                                           fermi_method=mode)
    assert ok is True
    eig = props["eigenvalues"].detach().cpu().numpy()[0]
    assert eig.shape == wr.shape
    assert np.abs(eig - wr).max() < EIG_ATOL
    for key in ("fermi_energy_Ha",):
        assert abs(props[key] - ref[key]) < EIG_ATOL, key
    for key in ("fermi_energy_eV", "vbm_eV", "cbm_eV", "valence_band_width_eV", "conduction_band_width_eV"):
        assert abs(float(props[key]) - float(ref[key])) < E_ATOL, (key, props[key], ref[key])
    assert abs(float(props["band_gap_eV"]) - float(ref["band_gap_eV"])) < E_ATOL
    if ref["minimum_direct_gap_eV"] is None:
        assert props["minimum_direct_gap_eV"] is None
    else:
        assert abs(props["minimum_direct_gap_eV"] - ref["minimum_direct_gap_eV"]) < E_ATOL
    assert bool(props["is_direct_semiconductor"]) == bool(ref["is_direct_semiconductor"])
    for key in ("nkpoints", "nbands", "nelectrons"):
        assert props[key] == ref[key], key
    metal_or_fallback = ref["vbm_eV"] == ref["cbm_eV"]
    if not metal_or_fallback:
        assert _tie_ok(wr, props["vbm_kpoint"], ref["vbm_eV"], E_ATOL)
        assert _tie_ok(wr, props["cbm_kpoint"], ref["cbm_eV"], E_ATOL)
        assert bool(props["is_direct_gap"]) == (int(props["vbm_kpoint"]) == int(props["cbm_kpoint"]))
    else:
        assert props["vbm_kpoint"] == ref["vbm_kpoint"] and props["cbm_kpoint"] == ref["cbm_kpoint"]
        assert bool(props["is_direct_gap"]) == bool(ref["is_direct_gap"])
    # DOS on the dict's (-10, 10) eV, 5000-point grid, Fermi-shifted
    grid = props["dos_energy_grid_tensor"].detach().cpu().numpy()
    d = props["dos_values_tensor"].detach().cpu().numpy()
    np.testing.assert_allclose(grid, ref["dos_energy_grid"], rtol=0, atol=1e-12)
    scale = max(1.0, np.abs(ref["dos_values"]).max())
    assert np.abs(d - ref["dos_values"]).max() < DOS_RTOL * scale
    assert abs(props["dos_at_fermi"] - ref["dos_at_fermi"]) < DOS_RTOL * scale
    assert abs(props["dos_total_states"] - ref["dos_total_states"]) < DOS_RTOL * scale * 20.0
    # gap centre: the arg-min of the DOS within 2 eV of mu; where the DOS is
    # flat (underflowed) any minimising grid point is a valid answer
    region = np.abs(grid) < 2.0
    gi = int(np.argmin(np.abs(grid - props["dos_gap_center_eV"])))
    assert region[gi] and abs(grid[gi] - props["dos_gap_center_eV"]) < 1e-12
    assert ref["dos_values"][gi] <= ref["dos_values"][region].min() + DOS_RTOL * scale
    # get_fermi block
    assert abs(float(props["efermi"].reshape(-1)[0]) - ref["efermi"]) < E_ATOL
    bgap = float(props["bandgap"].reshape(-1)[0])
    if np.isinf(ref["bandgap"]):
        assert np.isinf(bgap)
    else:
        assert abs(bgap - ref["bandgap"]) < E_ATOL
    assert props["calc"].eigenvalue is not None
    if grad:
        # the differentiable entries carry a graph back to the table parameters
        assert props["dos_values_tensor"].requires_grad


def _truncate(skfs, rows):
    out = {}
    for key, d in skfs.items():
        e = dict(d)
        e["grid"] = np.asarray(d["grid"])[:rows]
        e["hamiltonian"] = {k: np.asarray(v)[:, :rows] for k, v in d["hamiltonian"].items()}
        e["overlap"] = {k: np.asarray(v)[:, :rows] for k, v in d["overlap"].items()}
        out[key] = e
    return out


def test_build_through_interpolation_tail_matches_oracle():
    """Tables cut to 420 rows (r up to 8.4 Bohr): pair-images at 8.4-9.4 Bohr
    take the poly5 tail (interpolation.py:173-252) and those at 9.4-10 Bohr
    give zero; the HIP build must match the oracle there too."""
    st = synth.diamond_si()
    skfs = _truncate(synth.make_skf_set([14], seed=0), 420)
    tab = TableSet(skfs, synth.shell_dict_spd([14]), device="cuda")
    k = st.kpoints[:8]
    inp = engine.StructureInput(st.z, st.positions * AA_BOHR, st.cell * AA_BOHR, k, np.full(8, 1 / 8))
    plan = engine.plan_build(tab, [inp])
    H, S = engine.run_build(tab, plan)
    ref = obuild.build_hs(st.z, st.positions * AA_BOHR, st.cell * AA_BOHR, synth.shell_dict_spd([14]), skfs, k)
    # the tail region is populated: some pair-image distance falls in [8.4, 9.4) Bohr
    t, tc = obuild.reference_images(st.cell * AA_BOHR, st.positions * AA_BOHR)
    p = st.positions * AA_BOHR
    d = np.linalg.norm(tc[:, None, None, :] + p[None, None, :, :] - p[None, :, None, :], axis=-1)
    assert ((d >= 8.4) & (d < 9.4)).any()
    full = obuild.build_hs(st.z, st.positions * AA_BOHR, st.cell * AA_BOHR, synth.shell_dict_spd([14]),
                           synth.make_skf_set([14], seed=0), k)
    assert np.abs(full["H"] - ref["H"]).max() > 1e-8          # truncation changed H: the tail matters
    # The tail's y'' is the reference's central difference with delta = 1e-5
    # (interpolation.py:226-236): (y(x+d) + y(x-d) - 2 y(x)) / d^2 carries a
    # rounding error of order eps |y| / d^2 ~ 1e-6 |y| whose bits depend on how
    # the two interpolants are evaluated (Neville tableau in the oracle,
    # Lagrange weights on the device). It enters a tail value with weight
    # dx^2 C(xr) <= 1, so tail elements agree to ~1e-6 of the stencil values at
    # the table's end, not to the 1e-10 of the rest of the build.
    y_end = max(np.abs(np.asarray(v)[:, -9:-1]).max() for d in skfs.values()
                for part in ("hamiltonian", "overlap") for v in d[part].values())
    for got, want in ((H, ref["H"]), (S, ref["S"])):
        g = got.cpu().numpy()
        err = np.abs(g - want).max()
        print(f"tail build: max err {err:.3e} (|want| max {np.abs(want).max():.3e}, y_end {y_end:.3e})")
        assert err < 1e-10 * np.abs(want).max() + 1e-6 * y_end


@pytest.mark.parametrize("vectors", [False, True])
@pytest.mark.parametrize("name", ["diamond", "random3"])
def test_density_and_occupations_match_oracle(name, vectors):
    """calc.density [1, n, n, nk] and calc._occupations [1, nk, nb]
    (main.py:157-161, 187-189) against the oracle on LAPACK eigenvectors.
    rho is invariant to the phase of each eigenvector and to rotations inside
    fully occupied degenerate subspaces, so the comparison is basis-free."""
    from slakonet_amd.calculator import SimpleDftb
    z, pos, cell, species, mesh, seed = _case(name)
    skfs = synth.make_skf_set(species, seed=seed)
    tab = TableSet(skfs, synth.shell_dict_spd(species), device="cuda")
    k = np.array([[0.13, -0.21, 0.37], [0.31, 0.07, -0.19], [0.27, 0.41, 0.11]])
    calc = SimpleDftb(Geometry(z, pos, cell=cell, units="angstrom"), shell_dict=synth.shell_dict_spd(species),
                      tables=tab, kpoints_frac=k, with_eigenvectors=vectors)
    calc()
    ref = obuild.build_hs(z, pos * AA_BOHR, cell * AA_BOHR, synth.shell_dict_spd(species), skfs, k)
    wr, vr = osolve.solve_all(ref["H"], ref["S"], vectors=True)
    ne = obuild.n_electrons(z, skfs)
    n = wr.shape[-1]
    occ = opost.occupations(n, ne)
    # a partly filled degenerate level would make rho basis-dependent: the case must not have one
    top = int(ne // 2) - (0 if ne % 2 else 1)
    assert (np.abs(wr[:, top + 1] - wr[:, top]) > 1e-6).all()
    if ne % 2:
        assert (np.abs(wr[:, top] - wr[:, top - 1]) > 1e-6).all()
    rho_ref = opost.density(vr, occ)
    got_occ = calc._occupations.cpu().numpy()
    assert got_occ.shape == (1, len(k), n) and (got_occ[0] == occ[None]).all()
    rho = calc.density.cpu().numpy()
    assert rho.shape == (1, n, n, len(k))
    err = np.abs(rho[0].transpose(2, 0, 1) - rho_ref).max()
    print(f"density {name}: max err {err:.3e}")
    assert err < 1e-9 * max(1.0, np.abs(rho_ref).max())


@pytest.mark.parametrize("name", ["diamond", "random3"])
def test_electronic_energy_matches_oracle(name):
    """_calculate_electronic_energy (main.py:1430-1448) on a fresh calculator
    (it runs the solve itself, main.py:1433-1435) against
    oracle.post.electronic_energy on the oracle's build, LAPACK eigenvalues and
    the bisection Fermi level."""
    from slakonet_amd.calculator import SimpleDftb
    from slakonet_amd.geometry import monkhorst_pack
    z, pos, cell, species, mesh, seed = _case(name)
    skfs = synth.make_skf_set(species, seed=seed)
    k, w = monkhorst_pack(mesh)
    tab = TableSet(skfs, synth.shell_dict_spd(species), device="cuda")
    calc = SimpleDftb(Geometry(z, pos, cell=cell, units="angstrom"), shell_dict=synth.shell_dict_spd(species),
                      tables=tab, kpoints_frac=k, k_weights=w)
    e = float(calc._calculate_electronic_energy())
    ref = obuild.build_hs(z, pos * AA_BOHR, cell * AA_BOHR, synth.shell_dict_spd(species), skfs, k)
    wr, _ = osolve.solve_all(ref["H"], ref["S"])
    mu = opost.fermi_bisect(wr, obuild.n_electrons(z, skfs), w)
    want = opost.electronic_energy(wr, w, mu)
    # eigenvalues agree to 1e-9 Ha; each occupation moves by at most f' * 1e-9 < 1e-9 / (4 kT)
    assert abs(e - want) < 1e-7 * max(1.0, abs(want)), (e, want)
