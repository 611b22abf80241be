"""Host-side logic and the C-ABI library surface (no GPU calls)."""
import os
import re

import numpy as np
import pytest

from slakonet_amd import engine, synth
from slakonet_amd.geometry import Geometry, klines_points, lattice_translations, monkhorst_pack
from slakonet_amd.units import AA_BOHR

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "slakonet_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|size_t|void)\s+(slk_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_header_symbol():
    from slakonet_amd import _lib
    lib = _lib.load_library()
    names = header_functions()
    assert len(names) >= 8
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES), "binding signatures must cover the header exactly"


def test_ctypes_struct_sizes_match_numpy_layouts():
    import ctypes as C
    from slakonet_amd import _lib
    assert C.sizeof(_lib.SlkStructDesc) == engine.STRUCT_DT.itemsize == 40
    assert C.sizeof(_lib.SlkBuildWork) == engine.WORK_DT.itemsize == 32
    assert C.sizeof(_lib.SlkSpectrum) == engine.SPEC_DT.itemsize == 32
    assert C.sizeof(_lib.SlkDosJob) == engine.DOS_DT.itemsize == 40


def test_monkhorst_pack_matches_reference_rule():
    k, w = monkhorst_pack([2, 3, 1])
    assert k.shape == (6, 3) and np.allclose(w, 1 / 6)
    # (2i+1)/(2N), x slowest, z fastest; [1,1,1] -> (1/2,1/2,1/2)
    np.testing.assert_allclose(k[0], [0.25, 1 / 6, 0.5])
    np.testing.assert_allclose(k[1], [0.25, 0.5, 0.5])
    np.testing.assert_allclose(k[3], [0.75, 1 / 6, 0.5])
    np.testing.assert_allclose(monkhorst_pack([1, 1, 1])[0], [[0.5, 0.5, 0.5]])


def test_klines_inclusive_endpoints_and_error_on_single_point():
    k, w = klines_points([[0, 0, 0, 0.5, 0, 0, 3], [0.5, 0, 0, 0.5, 0.5, 0, 2]])
    np.testing.assert_allclose(k, [[0, 0, 0], [0.25, 0, 0], [0.5, 0, 0], [0.5, 0, 0], [0.5, 0.5, 0]])
    assert np.allclose(w, 0.2)
    with pytest.raises(ValueError):
        klines_points([[0, 0, 0, 1, 1, 1, 1]])


def test_translations_cover_all_pairs_within_cutoff():
    rng = np.random.default_rng(0)
    z, pos, cell = synth.random_cell(rng, 6, [6, 14], 1.8)
    pb, cb = pos * AA_BOHR, cell * AA_BOHR
    ti, tc = lattice_translations(cb, pb, 10.0)
    got = set()
    for a in range(6):
        for b in range(6):
            d = np.linalg.norm(tc + pb[b] - pb[a], axis=1)
            for t in np.nonzero((d < 10.0) & (d > 0))[0]:
                got.add((a, b, tuple(ti[t])))
    # brute force over a much larger box
    r = 8
    ax = np.arange(-r, r + 1)
    big = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    bc = big @ cb
    ref = set()
    for a in range(6):
        for b in range(6):
            d = np.linalg.norm(bc + pb[b] - pb[a], axis=1)
            for t in np.nonzero((d < 10.0) & (d > 0))[0]:
                ref.add((a, b, tuple(big[t])))
    assert got == ref


def test_geometry_units_and_fractional():
    cell = np.eye(3) * 5.0
    g = Geometry([14, 14], [[0.0, 0, 0], [0.5, 0.5, 0.5]], cell=cell, frac=[1, 1], units="angstrom")
    np.testing.assert_allclose(g.positions_np[1], [2.5 * AA_BOHR] * 3)
    np.testing.assert_allclose(g.cell_np, cell * AA_BOHR)


def test_column_tiles():
    t = engine.column_tiles(np.array([9] * 10 + [4, 1]))
    assert (t[:, 3] <= 64).all() and t[0].tolist() == [0, 7, 0, 63]
    assert t[:, 3].sum() == 95 and t[-1, 1] == 12


def test_synth_tables_deterministic_and_symmetric():
    a = synth.make_skf_set([6, 14], seed=3)
    b = synth.make_skf_set([6, 14], seed=3)
    np.testing.assert_array_equal(a[(6, 14)]["hamiltonian"]["1-2"], b[(6, 14)]["hamiltonian"]["1-2"])
    # l1 == l2 integrals symmetric in the pair, l1 < l2 independent
    np.testing.assert_array_equal(a[(6, 14)]["hamiltonian"]["1-1"], a[(14, 6)]["hamiltonian"]["1-1"])
    assert not np.array_equal(a[(6, 14)]["hamiltonian"]["0-1"], a[(14, 6)]["hamiltonian"]["0-1"])
    assert a[(6, 6)]["hamiltonian"]["2-2"].shape == (3, 661)


def test_stream_structures_respect_min_distance():
    for i in range(5):
        st = synth.stream_structure(7, i)
        pos, cell = st.positions, st.cell
        inv = np.linalg.inv(cell)
        for a in range(len(pos)):
            d = pos - pos[a]
            f = d @ inv
            f -= np.round(f)
            dd = np.linalg.norm(f @ cell, axis=1)
            dd[a] = 99
            assert dd.min() >= 1.8 - 1e-9
        assert 2 <= st.n_atoms <= 256


@pytest.mark.parametrize("method", ["state_dict", "compact", "universal_params"])
def test_model_save_load_round_trip(tmp_path, method):
    """Reference save/load layouts (slakonet/optim.py:204-625), read back with
    weights_only loads; trained values and constraint originals survive."""
    import torch
    from slakonet_amd import synth
    from slakonet_amd.optim import MultiElementSkfParameterOptimizer
    model = MultiElementSkfParameterOptimizer(skf_dicts=synth.make_skf_set([6, 14], seed=2), device="cpu")
    with torch.no_grad():
        for p in model.parameters():
            p.mul_(1.01)
    model.save_model(tmp_path / "m", method=method)
    back = MultiElementSkfParameterOptimizer.load_model(tmp_path / "m", method=method)
    a, b = model.state_dict(), back.state_dict()
    assert a.keys() == b.keys()
    assert all(torch.equal(a[k], b[k]) for k in a)
    if method == "state_dict":
        for key, opt in model.skf_optimizers.items():
            ob = back.skf_optimizers[key]
            assert all(torch.equal(opt.original_h_params[k], ob.original_h_params[k]) for k in opt.original_h_params)
    back.apply_constraints()


def test_skf_text_round_trip_and_repeat_notation(tmp_path):
    """DFTB+ simple-format .skf: write synthetic tables, read back; the
    "n*value" notation and comma separators expand."""
    from slakonet_amd import skf_io
    d = synth.make_skf(14, 14, seed=0)
    path = tmp_path / "Si-Si.skf"
    skf_io.write_skf(str(path), d)
    back = skf_io.read_skf(str(path))
    assert back["atom_pair"] == [14, 14]
    for part in ("hamiltonian", "overlap"):
        for key, v in d[part].items():
            k = key if isinstance(key, str) else f"{key[0]}-{key[1]}"
            assert np.allclose(np.asarray(back[part][k]), np.asarray(v), rtol=1e-14, atol=0)
    assert np.allclose(back["grid"], d["grid"], rtol=1e-12)
    assert np.allclose(back["atomic_data"]["on_sites"], d["atomic_data"]["on_sites"], rtol=1e-12)
    # hand-written heteronuclear file: 2 grid points, repeat notation, commas
    txt = "0.5, 2\n12.0 19*0.0\n" + "\n".join(
        " ".join(str(10 * g + c) for c in range(20)) for g in (1, 2)) + "\n"
    p2 = tmp_path / "C-Si.skf"
    p2.write_text(txt)
    h = skf_io.read_skf(str(p2))
    assert h["atom_pair"] == [6, 14] and not h["atomic"]
    assert h["grid"] == [0.5, 1.0]
    # file column 9 is Hss0, 0 is Hdd0, 10+9 is Sss0
    assert h["hamiltonian"]["0-0"][0] == [19.0, 29.0]
    assert h["hamiltonian"]["2-2"][0] == [10.0, 20.0]
    assert h["overlap"]["0-0"][0] == [29.0, 39.0]
    assert h["hamiltonian"]["1-2"][1] == [14.0, 24.0]      # Hpd1
    dd = skf_io.read_skf_directory(str(tmp_path), ["Si", "C"])
    assert set(dd) == {(14, 14), (6, 14)}
    assert skf_io._numbers("3*1.5, 2") .tolist() == [1.5, 1.5, 1.5, 2.0]


def test_native_planner_matches_python_bookkeeping():
    """slk_plan_size / slk_plan_fill (host C++) against the numpy statement of
    the same bookkeeping: translations (x slowest), column tiles, work items."""
    from slakonet_amd.tables import TableSet
    sts = [synth.stream_structure(77, i, n_max=40) for i in range(6)]
    species = sorted({int(z) for st in sts for z in st.z})
    tab = TableSet(synth.make_skf_set(species, seed=1), synth.shell_dict_spd(species), device="cpu")
    inputs = [engine.StructureInput(st.z, st.positions * AA_BOHR, st.cell * AA_BOHR, st.kpoints, st.k_weights)
              for st in sts]
    plan = engine.plan_build(tab, inputs)
    d = plan.dev
    structs = np.frombuffer(d["structs"].numpy().tobytes(), dtype=engine.STRUCT_DT)
    work = np.frombuffer(d["work"].numpy().tobytes(), dtype=engine.WORK_DT)
    tint = np.frombuffer(d["trans_int"].numpy().tobytes(), dtype=np.int32).reshape(-1, 3)
    tcart = np.frombuffer(d["trans_cart"].numpy().tobytes(), dtype=np.float64).reshape(-1, 3)
    w0 = 0
    for s, inp in enumerate(inputs):
        ti, tc = lattice_translations(inp.cell, inp.positions, 10.0)
        sd = structs[s]
        assert sd["nt"] == len(ti) and sd["nk"] == len(inp.kpoints)
        assert np.array_equal(tint[sd["t0"]:sd["t0"] + sd["nt"]], ti)
        assert np.allclose(tcart[sd["t0"]:sd["t0"] + sd["nt"]], tc, rtol=0, atol=1e-12)
        norb_atoms = tab.norb[tab.species_index(inp.z)]
        assert sd["n_orb"] == norb_atoms.sum() == plan.n_orb[s]
        tiles = engine.column_tiles(norb_atoms)
        nkc = (len(inp.kpoints) + engine.KC - 1) // engine.KC
        ii, tt, kk = np.meshgrid(np.arange(len(inp.z)), np.arange(len(tiles)), np.arange(nkc), indexing="ij")
        ws = work[w0:w0 + ii.size]
        assert np.array_equal(ws["i"], ii.ravel()) and np.array_equal(ws["kc0"], kk.ravel() * engine.KC)
        assert np.array_equal(ws["j0"], tiles[tt.ravel(), 0]) and np.array_equal(ws["ncol"], tiles[tt.ravel(), 3])
        w0 += ii.size
    assert w0 == plan.n_work
    # row-panel items follow the tiles: one per (structure, atom, k chunk), all columns
    rows = work[plan.n_work:]
    assert len(rows) == plan.n_row_work == sum(len(i.z) * ((len(i.kpoints) + engine.KC - 1) // engine.KC)
                                               for i in inputs)
    r0 = 0
    for s, inp in enumerate(inputs):
        nkc = (len(inp.kpoints) + engine.KC - 1) // engine.KC
        rs = rows[r0:r0 + len(inp.z) * nkc]
        assert np.all(rs["s"] == s) and np.all(rs["j0"] == 0) and np.all(rs["j1"] == len(inp.z))
        assert np.all(rs["ncol"] == plan.n_orb[s])
        assert np.array_equal(rs["i"], np.repeat(np.arange(len(inp.z)), nkc))
        assert np.array_equal(rs["kc0"], np.tile(np.arange(nkc) * engine.KC, len(inp.z)))
        r0 += len(rs)
    with pytest.raises(ValueError):
        engine.plan_build(tab, inputs, ld=8)


def test_kpts_to_klines_pairs():
    """kpts_to_klines (slakonet/optim.py:74-100): consecutive pairs become rows."""
    from slakonet_amd.geometry import kpts_to_klines
    kp = [[0, 0, 0], [0.5, 0, 0], [0.5, 0.5, 0], [0, 0, 0], [0.1, 0.1, 0.1]]
    kl = kpts_to_klines(kp, default_points=7).numpy()
    assert kl.shape == (2, 7)
    assert np.array_equal(kl[0], [0, 0, 0, 0.5, 0, 0, 7]) and np.array_equal(kl[1], [0.5, 0.5, 0, 0, 0, 0, 7])
    k, w = klines_points(kl)
    assert len(k) == 14 and np.allclose(k[6], [0.5, 0, 0]) and np.allclose(w, 1 / 14)


def test_skf_parameter_optimizer_from_path_and_clamp_rule(tmp_path):
    """SkfParameterOptimizer(skf_path) (slakonet/optim.py:1356-1363) reads the
    .skf file; apply_constraints (:1392-1434) clamps H to [c, 2 - c] x original
    with c = 0.9 (l1 == l2) / 0.7 (l1 != l2), S to [max(0.95 x, 0.1), 1.05 x]
    (l1 == l2) / [0.9, 1.1] x original."""
    import torch
    from slakonet_amd import skf_io
    from slakonet_amd.optim import SkfParameterOptimizer
    d = synth.make_skf(14, 14, seed=0)
    path = tmp_path / "Si-Si.skf"
    skf_io.write_skf(str(path), d)
    opt = SkfParameterOptimizer(str(path))
    assert set(opt.h_params) == {"0-0", "0-1", "0-2", "1-1", "1-2", "2-2"}
    for key, v in d["hamiltonian"].items():
        assert torch.allclose(opt.h_params[key].detach(), torch.as_tensor(np.asarray(v)), rtol=1e-14, atol=0)
    assert SkfParameterOptimizer(d).h_params.keys() == opt.h_params.keys()
    with torch.no_grad():
        for p in list(opt.h_params.values()) + list(opt.s_params.values()):
            p.mul_(3.0)
    opt.apply_constraints()
    for key, p in opt.h_params.items():
        o = opt.original_h_params[key]
        c = 0.9 if key[0] == key[2] else 0.7
        # original H integrals are negative: 3 x original lies below the bounds
        # [c o, (2 - c) o] -> swapped by sign; torch.clamp returns max(min(x, hi), lo)
        want = torch.clamp(3.0 * o, o * c, o * (2 - c))
        assert torch.equal(p.detach(), want), key
        assert (p.detach() <= 0).all()
    for key, p in opt.s_params.items():
        o = opt.original_s_params[key]
        if key[0] == key[2]:
            want = torch.minimum(torch.maximum(3.0 * o, torch.maximum(0.95 * o, torch.tensor(0.1))), 1.05 * o)
        else:
            want = 1.1 * o
        assert torch.allclose(p.detach(), want, rtol=1e-15, atol=0), key
    with torch.no_grad():
        for p in opt.s_params.values():
            p.zero_()
    opt.apply_constraints()
    for key, p in opt.s_params.items():
        o = opt.original_s_params[key]
        lo = torch.maximum(0.95 * o, torch.tensor(0.1)) if key[0] == key[2] else 0.9 * o
        assert torch.allclose(p.detach(), torch.minimum(lo, 1.05 * o if key[0] == key[2] else 1.1 * o)), key


def test_large_cubic_fixture_is_the_generator_output():
    """tests/data/large_cubic_1024.npz (read by the GPU tests of the
    1024-atom cell) is exactly synth.large_cubic(): seeded, 1024 atoms in a
    22 A cube, minimum distance 1.9 A, species 14/6/8."""
    d = np.load(os.path.join(ROOT, "tests", "data", "large_cubic_1024.npz"))
    st = synth.large_cubic()
    assert np.array_equal(d["z"], st.z) and np.array_equal(d["positions"], st.positions)
    assert np.array_equal(d["cell"], st.cell) and np.array_equal(d["kpoints"], st.kpoints)
    assert len(st.z) == 1024 and set(st.z.tolist()) == {6, 8, 14}


def test_padding_level_guard():
    """A structure whose highest band reaches the identity-padding level is
    flagged (its padded eigenvalues could sort among its own bands)."""
    base = dict(mu=0.1, n_con=5, s_cbm=0.5, cb_width=10.0, s_vbm=-0.3)
    assert not engine.reaches_padding(base)
    top_ev = (engine.PAD_H - 0.1) * 27.211 - 0.5 + 1.0
    assert engine.reaches_padding(dict(base, cb_width=top_ev))
    assert not engine.reaches_padding(dict(base, n_con=0))
