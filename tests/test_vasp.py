"""Training-data path on the CPU: vasprun.xml reader, MultiVaspDataLoader and
the default k-path, against files written with known contents."""
import numpy as np
import pytest

from slakonet_amd import vasp
from slakonet_amd.geometry import klines_points
from slakonet_amd.units import AA_BOHR
from vasp_fixture import write_vasprun

SI_CELL = 0.5 * 5.43 * np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]])


def _si_run(path, gap=1.2, spins=1, relax=False, occ_full=2.0):
    rng = np.random.default_rng(1)
    nk, nb = 4, 8
    vb = rng.uniform(-6.0, 0.0, size=(spins, nk, 4))
    vb[..., -1] = 0.0 - rng.uniform(0.0, 0.3, size=(spins, nk))
    vb[0, 2, -1] = 0.0                                  # VBM at k 3
    cb = gap + rng.uniform(0.05, 4.0, size=(spins, nk, 4))
    cb[spins - 1, 1, 0] = gap                           # CBM at k 2
    eig = np.concatenate([vb, cb], axis=-1)
    occ = np.concatenate([np.full_like(vb, occ_full), np.zeros_like(cb)], axis=-1)
    occ[..., 3] = 0.98 * occ_full                       # nearly full, still occupied
    occ[..., 4] = 0.02 * occ_full                       # nearly empty
    dos_e = np.linspace(-10.0, 10.0, 301)
    dos = np.stack([np.exp(-(dos_e - s) ** 2) for s in range(spins)])
    frac = np.array([[0.0, 0.0, 0.0], [0.25, 0.25, 0.25]])
    steps = None
    energies = [-10.8]
    if relax:
        steps = [(SI_CELL * 1.01, frac + 0.01), (SI_CELL, frac)]
        energies = [-10.5, -10.8]
    write_vasprun(path, SI_CELL, ["Si", "Si"], frac, eig, occ, dos_e, dos, energies, steps=steps)
    return eig, occ, dos_e, dos


def test_vasprun_reader_fields(tmp_path):
    p = tmp_path / "vasprun.xml"
    eig, occ, dos_e, dos = _si_run(p, gap=1.2)
    vr = vasp.Vasprun(str(p))
    st = vr.all_structures[-1]
    assert st.num_atoms == 2 and st.species == ["Si", "Si"]
    assert np.allclose(st.lattice, SI_CELL, atol=1e-7)
    assert st.composition == {"Si": 2} and st.reduced_formula == "Si"
    assert abs(st.volume - 5.43 ** 3 / 4) < 1e-5
    assert vr.final_energy == pytest.approx(-10.8)
    e, o = vr.eigenvalues
    assert e.shape == (1, 4, 8) and np.allclose(e, eig, atol=6e-5) and np.allclose(o, occ, atol=6e-5)
    gap, cbm, vbm = vr.get_indir_gap
    assert vbm == pytest.approx(0.0, abs=1e-9) and cbm == pytest.approx(1.2, abs=1e-9)
    assert gap == pytest.approx(1.2, abs=1e-9)
    t = vr.total_dos
    assert len(t) == 2 and np.allclose(t[0], dos_e, atol=6e-5) and np.allclose(t[1], dos[0], atol=6e-5)
    assert vr.efermi == pytest.approx(0.5)


def test_vasprun_reader_relaxation_and_spin(tmp_path):
    """The final structure is the last ionic step's; the final energy is the
    last step's e_fr_energy (not an SCF step's); two spin channels give two
    total-DOS columns and the gap over both channels; occupations of 1 (spin
    polarised) use the same half-occupation split."""
    p = tmp_path / "vasprun.xml"
    _si_run(p, gap=0.7, spins=2, relax=True, occ_full=1.0)
    vr = vasp.Vasprun(str(p))
    assert len(vr.all_structures) == 2
    assert np.allclose(vr.all_structures[0].lattice, SI_CELL * 1.01, atol=1e-7)
    assert np.allclose(vr.all_structures[-1].frac_coords, [[0, 0, 0], [0.25, 0.25, 0.25]], atol=1e-7)
    assert vr.final_energy == pytest.approx(-10.8)
    assert len(vr.total_dos) == 3
    assert vr.get_indir_gap[0] == pytest.approx(0.7, abs=1e-9)


def test_vasprun_metal_gap_is_zero(tmp_path):
    p = tmp_path / "vasprun.xml"
    _si_run(p, gap=-0.4)            # CBM below VBM: overlapping bands
    assert vasp.Vasprun(str(p)).get_indir_gap[0] == 0.0


def test_loader_glob_geometry_and_bad_files(tmp_path):
    for i, g in enumerate((1.1, 2.3)):
        _si_run(tmp_path / f"vasprun-{i}.xml", gap=g)
    (tmp_path / "vasprun-bad.xml").write_text("<modeling><atominfo/></modeling>")
    msgs = []
    dl = vasp.MultiVaspDataLoader(str(tmp_path / "vasprun-*.xml"), log=msgs.append)
    assert len(dl) == 2 and any("vasprun-bad.xml" in m for m in msgs)
    assert dl.get_all_elements() == ["Si"]
    d = dl[0]
    assert d["target_bandgap"] == pytest.approx(1.1, abs=1e-9) and dl[1]["target_bandgap"] == pytest.approx(2.3)
    assert d["metadata"] == {"natoms": 2, "formula": "Si", "volume": pytest.approx(5.43 ** 3 / 4, abs=1e-5)}
    g = d["geometry"]
    assert np.allclose(g.cell_np, SI_CELL * AA_BOHR, atol=1e-6)
    assert np.allclose(g.positions_np[1], np.array([0.25, 0.25, 0.25]) @ SI_CELL * AA_BOHR, atol=1e-6)
    assert d["target_dos"].shape == d["dos_energies"].shape == (301,)
    assert len(dl.get_batch(batch_size=1)) == 1 and len(dl.get_batch()) == 2
    with pytest.raises(ValueError):
        vasp.MultiVaspDataLoader(str(tmp_path / "vasprun-bad.xml"), log=None)


def test_fcc_kpath_points_and_klines():
    """Segment ends are the tabulated points, in path order; the k-lines rows
    pair consecutive points as the reference pairs jarvis' list."""
    k = vasp.fcc_kpath(line_density=20)
    assert np.allclose(k[0], 0.0)
    for lbl in ("X", "W", "K", "L", "U"):
        assert np.any(np.all(np.isclose(k, vasp.FCC_POINTS[lbl]), axis=1)), lbl
    # |Gamma-X| of silicon = 2 pi / a: 20 points per inverse Angstrom
    assert np.all(np.isclose(k[:3], [[0, 0, 0], [0.5 / 23, 0, 0.5 / 23], [1.0 / 23, 0, 1.0 / 23]]))
    kl = vasp.get_klines_example()
    assert kl.shape == (len(k) // 2, 7) and np.all(kl[:, 6].numpy() == 2)
    kp, w = klines_points(kl)
    assert np.allclose(kp, k[: 2 * (len(k) // 2)]) and np.isclose(w.sum(), 1.0)
