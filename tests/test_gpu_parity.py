"""HIP path vs the FP64 oracle on the same seeded inputs (MI355X only)."""
import numpy as np
import pytest
import torch

from oracle import build as obuild, post as opost, solve as osolve
from slakonet_amd import engine, synth
from slakonet_amd.tables import TableSet
from slakonet_amd.units import AA_BOHR, H2E

pytestmark = pytest.mark.gpu

# H/S relative tolerance in FP64 mode (fraction of max |H|, |S|)
HS_RTOL = 1e-10
EIG_ATOL = 1e-9      # Hartree, FP64 path
DOS_RTOL = 1e-8


def _inputs(st, skfs, kmax=None, kpts=None):
    k = st.kpoints if kpts is None else kpts
    w = st.k_weights if kpts is None else np.full(len(kpts), 1.0 / len(kpts))
    if kmax is not None:
        k, w = k[:kmax], np.full(min(kmax, len(k)), 1.0 / min(kmax, len(k)))
    return engine.StructureInput(st.z, st.positions * AA_BOHR, st.cell * AA_BOHR, k, w)


def _tables(skfs, species):
    return TableSet(skfs, synth.shell_dict_spd(species), device="cuda")


def _oracle_hs(st, skfs, species, k):
    return obuild.build_hs(st.z, st.positions * AA_BOHR, st.cell * AA_BOHR, synth.shell_dict_spd(species), skfs, k)


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def diamond_case():
    st = synth.diamond_si()
    skfs = synth.make_skf_set([14], seed=0)
    tab = _tables(skfs, [14])
    inp = _inputs(st, skfs)
    plan = engine.plan_build(tab, [inp])
    H, S = engine.run_build(tab, plan)
    torch.cuda.synchronize()
    return st, skfs, tab, inp, plan, H, S


def test_build_diamond_matches_oracle(diamond_case):
    st, skfs, tab, inp, plan, H, S = diamond_case
    ref = _oracle_hs(st, skfs, [14], st.kpoints)
    assert _rel(H.cpu().numpy(), ref["H"]) < HS_RTOL
    assert _rel(S.cpu().numpy(), ref["S"]) < HS_RTOL


def test_build_hetero_rocksalt_matches_oracle():
    st = synth.rocksalt_64()
    skfs = synth.make_skf_set([12, 8], seed=1)
    tab = _tables(skfs, [8, 12])
    k = st.kpoints[::9]                      # 8 of the 64 Gamma-centred points
    inp = _inputs(st, skfs, kpts=k)
    plan = engine.plan_build(tab, [inp])
    H, S = engine.run_build(tab, plan)
    ref = _oracle_hs(st, skfs, [8, 12], k)
    assert _rel(H.cpu().numpy(), ref["H"]) < HS_RTOL
    assert _rel(S.cpu().numpy(), ref["S"]) < HS_RTOL


def test_build_random_cell_three_species_matches_oracle():
    rng = np.random.default_rng(11)
    z, pos, cell = synth.random_cell(rng, 12, [1, 22, 31], 1.8)
    skfs = synth.make_skf_set([1, 22, 31], seed=5)
    tab = _tables(skfs, [1, 22, 31])
    k = np.array([[0.5, 0.5, 0.5], [0.1, -0.2, 0.3], [0.0, 0.0, 0.0]])
    inp = engine.StructureInput(z, pos * AA_BOHR, cell * AA_BOHR, k, np.full(3, 1 / 3))
    plan = engine.plan_build(tab, [inp])
    H, S = engine.run_build(tab, plan)
    ref = obuild.build_hs(z, pos * AA_BOHR, cell * AA_BOHR, synth.shell_dict_spd([1, 22, 31]), skfs, k)
    assert _rel(H.cpu().numpy(), ref["H"]) < HS_RTOL
    assert _rel(S.cpu().numpy(), ref["S"]) < HS_RTOL


def test_build_sp_shells_and_complex64_output():
    st = synth.diamond_si()
    skfs = synth.make_skf_set([14], seed=0)
    sd = {14: [0, 1]}
    tab = TableSet(skfs, sd, device="cuda")
    inp = _inputs(st, skfs, kmax=8)
    plan = engine.plan_build(tab, [inp])
    H, S = engine.run_build(tab, plan)
    Hf, Sf = engine.run_build(tab, plan, out_f32=True)
    ref = obuild.build_hs(st.z, st.positions * AA_BOHR, st.cell * AA_BOHR, sd, skfs, inp.kpoints)
    assert H.shape[-1] == 8
    assert _rel(H.cpu().numpy(), ref["H"]) < HS_RTOL
    assert _rel(Hf.cpu().numpy().astype(np.complex128), ref["H"]) < 1e-6


def test_eigh_values_diamond(diamond_case):
    st, skfs, tab, inp, plan, H, S = diamond_case
    w, _, _ = engine.eigh(H, S)
    ref, _ = osolve.solve_all(H.cpu().numpy(), S.cpu().numpy())
    assert np.abs(w.cpu().numpy() - ref).max() < EIG_ATOL


def test_eigh_vectors_residual_and_orthonormality():
    st = synth.rocksalt_64()
    skfs = synth.make_skf_set([12, 8], seed=1)
    tab = _tables(skfs, [8, 12])
    k = st.kpoints[[0, 21, 42]]
    inp = _inputs(st, skfs, kpts=k)
    plan = engine.plan_build(tab, [inp])
    H, S = engine.run_build(tab, plan)
    w, V, _ = engine.eigh(H, S, vectors=True)
    Hn, Sn, wn, Vn = (x.cpu().numpy() for x in (H, S, w, V))
    ref, _ = osolve.solve_all(Hn, Sn)
    assert np.abs(wn - ref).max() < EIG_ATOL
    for b in range(len(k)):
        r = Hn[b] @ Vn[b] - Sn[b] @ Vn[b] * wn[b][None, :]
        assert np.abs(r).max() < 1e-9
        g = Vn[b].conj().T @ Sn[b] @ Vn[b]
        assert np.abs(g - np.eye(576)).max() < 1e-9


@pytest.mark.parametrize("n", [1, 2, 5, 31, 32, 33, 64, 65, 100, 130])
def test_eigh_random_sizes_vs_lapack(n):
    """Fused (n <= 32) and blocked (panel edges at 32, 64) paths against
    LAPACK zhegv on random Hermitian / Hermitian positive definite pairs."""
    import scipy.linalg
    rng = np.random.default_rng(n)
    batch = 3
    Hs, Ss = [], []
    for _ in range(batch):
        x = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        Hs.append(0.5 * (x + x.conj().T))
        y = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        Ss.append(np.eye(n) + 0.2 * y @ y.conj().T / n)
    H = torch.from_numpy(np.stack(Hs)).cuda()
    S = torch.from_numpy(np.stack(Ss)).cuda()
    w, _, _ = engine.eigh(H, S)
    for b in range(batch):
        ref = scipy.linalg.eigh(Hs[b], Ss[b], eigvals_only=True)
        assert np.abs(w[b].cpu().numpy() - ref).max() < 1e-10 * max(1.0, np.abs(ref).max())
    if n >= 2:
        w2, V, _ = engine.eigh(H, S, vectors=True)
        Vn = V.cpu().numpy()
        for b in range(batch):
            r = Hs[b] @ Vn[b] - Ss[b] @ Vn[b] * w2[b].cpu().numpy()[None, :]
            assert np.abs(r).max() < 1e-9


def test_padded_batch_equals_alone():
    """Two structures of different size in one identity-padded batch give the
    same eigenvalues as each solved alone."""
    skfs = synth.make_skf_set([6, 14], seed=2)
    tab = _tables(skfs, [6, 14])
    rng = np.random.default_rng(3)
    ins = []
    for n in (3, 7):
        z, pos, cell = synth.random_cell(rng, n, [6, 14], 1.8)
        ins.append(engine.StructureInput(z, pos * AA_BOHR, cell * AA_BOHR, np.array([[0.5, 0.5, 0.5]]), np.ones(1)))
    plan = engine.plan_build(tab, ins, ld=72)
    H, S = engine.run_build(tab, plan)
    w, _, _ = engine.eigh(H, S)
    for s, inp in enumerate(ins):
        p1 = engine.plan_build(tab, [inp])
        h1, s1 = engine.run_build(tab, p1)
        w1, _, _ = engine.eigh(h1, s1)
        n = plan.n_orb[s]
        assert np.abs(w[plan.mat0[s], :n].cpu().numpy() - w1[0].cpu().numpy()).max() < EIG_ATOL
        assert (w[plan.mat0[s], n:].cpu().numpy() > 3.9).all()


def test_fermi_edges_dos_match_oracle(diamond_case):
    st, skfs, tab, inp, plan, H, S = diamond_case
    w, _, _ = engine.eigh(H, S)
    wn = w.cpu().numpy()
    ne = tab.n_electrons(st.z)
    grid = torch.linspace(-20.0, 10.0, 2001, dtype=torch.float64, device="cuda")
    for mode, ofn in ((0, opost.fermi_bisect), (1, opost.fermi_search)):
        edges, dos, stats = engine.post(plan, w, fermi_mode=mode, dos_grid=grid, sigma=0.1)
        ed = engine.edges_dict(edges[0].cpu().numpy())
        mu_ref = ofn(wn, ne, st.k_weights)
        assert abs(ed["mu"] - mu_ref) < 1e-9
        g, dref = opost.dos(wn, st.k_weights, ed["mu"], energy_range=(-20.0, 10.0), num_points=2001, sigma=0.1)
        d = dos[0].cpu().numpy()
        assert np.abs(d - dref).max() < DOS_RTOL * max(1.0, np.abs(dref).max())
        if mode == 0:
            bg = opost.band_gap(wn, ed["mu"])
            assert abs(ed["vbm"] - bg["vbm"]) < 1e-9 and abs(ed["cbm"] - bg["cbm"]) < 1e-9
            assert int(ed["vbm_k"]) == bg["vbm_kpoint"] and int(ed["cbm_k"]) == bg["cbm_kpoint"]
            bs = opost.band_structure(wn, ed["mu"])
            assert abs(ed["vb_width"] - bs["valence_band_width"]) < 1e-9
            assert abs(ed["min_direct_gap"] - bs["minimum_direct_gap"]) < 1e-9


def test_nonpositive_overlap_raises_with_k_index():
    """Inflated overlap integrals at short range make S(k) indefinite: the
    solve raises naming a k-point instead of dropping it."""
    st = synth.diamond_si()
    skfs = synth.make_skf_set([14], seed=0)
    bad = dict(skfs[(14, 14)])
    bad["overlap"] = {k: np.asarray(v) * 16.0 for k, v in bad["overlap"].items()}
    tab = TableSet({(14, 14): bad}, synth.shell_dict_spd([14]), device="cuda")
    inp = _inputs(st, skfs, kmax=4)
    plan = engine.plan_build(tab, [inp])
    H, S = engine.run_build(tab, plan)
    sref = obuild.build_hs(st.z, st.positions * AA_BOHR, st.cell * AA_BOHR, synth.shell_dict_spd([14]),
                           {(14, 14): bad}, inp.kpoints, kinds=("S",))["S"]
    bad_k = [i for i in range(4) if np.linalg.eigvalsh(sref[i]).min() <= 0]
    assert bad_k, "test input must make S indefinite"
    with pytest.raises(engine.EigenFailure, match=f"k-point index {bad_k[0]}"):
        engine.eigh(H, S)


def test_values_beyond_lds_limit_padded():
    """ld = 4224 > 4096 (Sturm counts from global memory): a structure padded
    to that dimension keeps its eigenvalues (FP64 and refined mixed edges)."""
    st = synth.diamond_si()
    skfs = synth.make_skf_set([14], seed=0)
    tab = _tables(skfs, [14])
    inp = _inputs(st, skfs, kmax=3)
    plan = engine.plan_build(tab, [inp])
    H, S = engine.run_build(tab, plan)
    w_ref, _, _ = engine.eigh(H, S)
    big = engine.plan_build(tab, [inp], ld=4224)
    Hb, Sb = engine.run_build(tab, big)
    w_big, _, _ = engine.eigh(Hb, Sb)
    n = plan.n_orb[0]
    assert torch.allclose(w_big[:, :n], w_ref[:, :n], atol=EIG_ATOL, rtol=0)
    assert torch.allclose(w_big[:, n:], torch.full_like(w_big[:, n:], engine.PAD_H), atol=1e-9, rtol=0)
    e_ref = engine.edges_dict(engine.post(plan, w_ref)[0][0].cpu().numpy())
    wsl = []
    w_mix, _, _ = engine.eigh(Hb, Sb, precision="mixed", workspace=wsl)
    e_mix = engine.edges_dict(engine.post(big, w_mix, refine=(Hb, Sb, wsl[0]))[0][0].cpu().numpy())
    for key in ("vbm", "cbm"):
        assert abs(e_mix[key] - e_ref[key]) < 1e-6, (key, e_mix[key], e_ref[key])


@pytest.mark.parametrize("case", ["diamond", "random3"])
def test_atom_projected_dos_matches_oracle(case):
    """Device atom-type PDOS (eigenvectors of the FP64 solve) vs the oracle on
    scipy eigenvectors; degenerate subspaces contribute through their projector,
    so the result does not depend on the basis chosen inside them."""
    from slakonet_amd.calculator import SimpleDftb
    from slakonet_amd.geometry import Geometry
    if case == "diamond":
        st = synth.diamond_si()
        z, pos, cell, species = st.z, st.positions, st.cell, [14]
        k = st.kpoints[:6]
    else:
        rng = np.random.default_rng(12)
        species = [1, 22, 31]
        z, pos, cell = synth.random_cell(rng, 7, species, 1.8)
        k = np.array([[0.1, -0.2, 0.3], [0.0, 0.0, 0.0], [0.5, 0.25, 0.0]])
    skfs = synth.make_skf_set(species, seed=5)
    tab = _tables(skfs, species)
    geom = Geometry(z, pos, cell=cell, units="angstrom")
    calc = SimpleDftb(geom, shell_dict=synth.shell_dict_spd(species), tables=tab, kpoints_frac=k,
                      with_eigenvectors=True)
    calc()
    grid, pd, uniq = calc.calculate_atom_projected_dos(sigma=0.1, energy_range=(-8, 6), num_points=700)
    ref = obuild.build_hs(z, pos * AA_BOHR, cell * AA_BOHR, synth.shell_dict_spd(species), skfs, k)
    w_ref, v_ref = osolve.solve_all(ref["H"], ref["S"], vectors=True)
    mu = calc._edge_values()["mu"]
    sp = tab.species_index(z)
    from slakonet_amd.units import CHEMICAL_SYMBOLS
    syms = [CHEMICAL_SYMBOLS[int(v)] for v in z]
    types = np.repeat([uniq.index(s) for s in syms], tab.norb[sp])
    _, pref = opost.atom_projected_dos(w_ref, v_ref, types, len(uniq), mu, sigma=0.1, energy_range=(-8, 6),
                                       n_points=700)
    for i, s in enumerate(uniq):
        assert np.allclose(pd[s], pref[i], rtol=1e-7, atol=1e-9 * pref.max()), s


def _residuals(H, S, w, V, n):
    """max |H V - S V diag(w)| and max |V^H S V - I| on the device (complex128)."""
    Hn, Sn, Vn = H[0, :n, :n], S[0, :n, :n], V[0, :n, :n]
    wn = w[0, :n]
    r = (Hn @ Vn - (Sn @ Vn) * wn[None, :]).abs().max().item()
    g = (Vn.conj().transpose(0, 1) @ (Sn @ Vn) - torch.eye(n, dtype=Vn.dtype, device=Vn.device)).abs().max().item()
    return r, g


@pytest.mark.parametrize("case", ["stream-2304", "random-5200"])
def test_eigh_vectors_large(case):
    """Eigenvectors beyond the former n <= 2048 limit: a 256-atom random cell
    as in the serving stream (n = 2304, complex k, QL arrays in LDS) and a random pair with n = 5200
    (beyond 5000, QL arrays in global memory). Eigenvalues against the
    values-only solve (two-stage / bisection, an independent path), residual
    and S-orthonormality of the vectors."""
    print(f"\ncase: vectors {case}", flush=True)
    if case == "stream-2304":
        species = [6, 14, 31]
        z, pos, cell = synth.random_cell(np.random.default_rng(2304), 256, species, 1.8)
        tab = _tables(synth.make_skf_set(species, seed=7), species)
        inp = engine.StructureInput(z, pos * AA_BOHR, cell * AA_BOHR, np.array([[0.5, 0.25, 0.0]]), np.ones(1))
        plan = engine.plan_build(tab, [inp])
        H, S = engine.run_build(tab, plan)
    else:
        n = 5200
        g = torch.Generator(device="cuda").manual_seed(5)
        x = torch.randn(n, n, dtype=torch.complex128, device="cuda", generator=g)
        H = ((x + x.conj().transpose(0, 1)) / (2 * n ** 0.5)).unsqueeze(0).contiguous()
        y = torch.randn(n, n, dtype=torch.complex128, device="cuda", generator=g)
        S = (torch.eye(n, dtype=torch.complex128, device="cuda") + 0.05 * (y @ y.conj().transpose(0, 1)) / n)
        S = S.unsqueeze(0).contiguous()
    n = H.shape[-1]
    w, V, _ = engine.eigh(H, S, vectors=True)
    w0, _, _ = engine.eigh(H, S)
    torch.cuda.synchronize()
    dw = (w - w0).abs().max().item()
    r, o = _residuals(H, S, w, V, n)
    scale = w.abs().max().item()
    print(f"n={n}: |w_vec - w_val| {dw:.2e}, residual {r:.2e}, orthonormality {o:.2e}", flush=True)
    assert dw < 1e-9 * max(1.0, scale)
    assert r < 1e-9 * max(1.0, scale)
    assert o < 1e-9
