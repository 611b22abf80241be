"""Differentiation form on the MI355X: the hand-written adjoints against central
finite differences (step 1e-4, 1e-6 relative) and against the reference's
autograd formula (MI355X only)."""
import numpy as np
import pytest
import torch

from oracle import adjoint as oadj
from slakonet_amd import adjoint, engine, synth
from slakonet_amd.geometry import Geometry
from slakonet_amd.optim import MultiElementSkfParameterOptimizer
from slakonet_amd.tables import TableSet
from slakonet_amd.units import AA_BOHR

pytestmark = pytest.mark.gpu

FD_STEP = 1e-4
FD_RTOL = 1e-6
DOT_RTOL = 1e-10     # linear identity of the table adjoint (exact up to rounding)
REF_RTOL = 1e-8      # against the reference autograd formula at Gamma

GEN_K = np.array([[0.0, 0.0, 0.0], [0.1, -0.2, 0.3], [0.5, 0.5, 0.5], [0.25, 0.1, -0.35]])


def _random_case(seed=11, n_atoms=8, species=(1, 22, 31), k=GEN_K, truncate=None):
    rng = np.random.default_rng(seed)
    z, pos, cell = synth.random_cell(rng, n_atoms, list(species), 1.8)
    sp = sorted(species)
    skfs = synth.make_skf_set(sp, seed=5)
    if truncate is not None:
        for d in skfs.values():
            for part in ("hamiltonian", "overlap"):
                d[part] = {key: np.asarray(v)[..., :truncate] for key, v in d[part].items()}
            d["grid"] = np.asarray(d["grid"])[:truncate]
    tab = TableSet(skfs, synth.shell_dict_spd(sp), device="cuda")
    inp = engine.StructureInput(z, pos * AA_BOHR, cell * AA_BOHR, k, np.full(len(k), 1.0 / len(k)))
    return tab, engine.plan_build(tab, [inp])


def _solve(tab, plan, t):
    H, S = engine.run_build(tab, plan, table_tensor=t)
    w, V, _ = engine.eigh(H, S, vectors=True)
    return w, V


def _smooth_loss(w, kw):
    # symmetric in degenerate levels, so the eigenvalue derivative is well defined
    return (kw[:, None] * torch.sin(2.0 * w + 0.3)).sum()


@pytest.mark.parametrize("case", ["random3", "diamond", "tail"])
def test_table_adjoint_dot_product(case):
    """<dtab, D> == sum_k Re sum Q[nu,mu] (M(tab + D) - M(tab))[mu,nu] for random Q, D."""
    if case == "random3":
        tab, plan = _random_case()
    elif case == "tail":
        # tables end at 8.38 Bohr: pairs between 8.38 and 9.38 Bohr use the poly5 tail
        tab, plan = _random_case(seed=4, n_atoms=6, species=(6, 14), truncate=420)
    else:
        st = synth.diamond_si()
        tab = TableSet(synth.make_skf_set([14], seed=0), synth.shell_dict_spd([14]), device="cuda")
        inp = engine.StructureInput(st.z, st.positions * AA_BOHR, st.cell * AA_BOHR, st.kpoints[:12],
                                    np.full(12, 1 / 12))
        plan = engine.plan_build(tab, [inp])
    g = torch.Generator(device="cuda").manual_seed(3)
    shape = (plan.n_mat, plan.ld, plan.ld)
    QH = torch.randn(shape, dtype=torch.complex128, device="cuda", generator=g)
    QS = torch.randn(shape, dtype=torch.complex128, device="cuda", generator=g)
    D = torch.randn(tab.tables.shape, dtype=torch.float64, device="cuda", generator=g)
    dtab = adjoint.table_gradient_from_densities(tab, plan, QH, QS)
    H0, S0 = engine.run_build(tab, plan)
    H1, S1 = engine.run_build(tab, plan, table_tensor=(tab.tables + D).contiguous())
    rhs = float((QH.transpose(-1, -2) * (H1 - H0)).real.sum() + (QS.transpose(-1, -2) * (S1 - S0)).real.sum())
    lhs = float((dtab * D).sum())
    assert abs(lhs - rhs) <= DOT_RTOL * abs(rhs), (lhs, rhs)


def test_eigenvalue_adjoint_matches_central_differences():
    tab, plan = _random_case()
    t0 = tab.tables.clone()
    w, V = _solve(tab, plan, t0)
    leaf = t0.clone().requires_grad_(True)
    wg = adjoint.TableEigenvalues.apply(leaf, tab, plan, V, w)
    loss = _smooth_loss(wg, plan.k_weights)
    loss.backward()
    gen = torch.Generator(device="cuda").manual_seed(7)
    D = torch.randn(t0.shape, dtype=torch.float64, device="cuda", generator=gen) * t0.abs().max()
    an = float((leaf.grad * D).sum())
    lp = float(_smooth_loss(_solve(tab, plan, (t0 + FD_STEP * D).contiguous())[0], plan.k_weights))
    lm = float(_smooth_loss(_solve(tab, plan, (t0 - FD_STEP * D).contiguous())[0], plan.k_weights))
    fd = (lp - lm) / (2 * FD_STEP)
    assert abs(an - fd) <= FD_RTOL * abs(fd), (an, fd)


def test_dos_adjoint_matches_central_differences_at_fixed_mu():
    tab, plan = _random_case(seed=12, n_atoms=6)
    t0 = tab.tables.clone()
    w, V = _solve(tab, plan, t0)
    edges, _, _ = engine.post(plan, w)
    grid = torch.linspace(-20.0, 10.0, 2001, dtype=torch.float64, device="cuda")
    target = torch.linspace(0.0, 1.0, 2001, dtype=torch.float64, device="cuda")

    def loss_of(wv):
        dos = adjoint.GaussianDos.apply(wv, plan, edges, grid, 0.1)
        return ((dos[0] - target) ** 2).mean()

    leaf = t0.clone().requires_grad_(True)
    loss = loss_of(adjoint.TableEigenvalues.apply(leaf, tab, plan, V, w))
    loss.backward()
    gen = torch.Generator(device="cuda").manual_seed(9)
    # the 0.1 eV Gaussians make the loss stiff: a direction of 1% of the table
    # scale keeps the O(step^2) truncation error of the differences below 1e-6
    D = torch.randn(t0.shape, dtype=torch.float64, device="cuda", generator=gen) * (0.01 * t0.abs().max())
    an = float((leaf.grad * D).sum())
    lp = float(loss_of(_solve(tab, plan, (t0 + FD_STEP * D).contiguous())[0]))
    lm = float(loss_of(_solve(tab, plan, (t0 - FD_STEP * D).contiguous())[0]))
    fd = (lp - lm) / (2 * FD_STEP)
    assert abs(an - fd) <= FD_RTOL * abs(fd), (an, fd)


def test_matches_reference_autograd_at_gamma():
    """At Gamma the reference's _SymEigB gradient equals ours (table gradient
    to 1e-8); at a complex k it does not (its missing conjugate)."""
    for k, agree in ((np.zeros((1, 3)), True), (np.array([[0.1, -0.2, 0.3]]), False)):
        tab, plan = _random_case(seed=13, n_atoms=5, k=k)
        w, V = _solve(tab, plan, tab.tables)
        n = plan.n_orb[0]
        gen = np.random.default_rng(2)
        g = gen.normal(size=n)
        H, S = engine.run_build(tab, plan)
        Hn, Sn = H[0].cpu().numpy(), S[0].cpu().numpy()
        if agree:
            Hn, Sn = Hn.real.copy(), Sn.real.copy()
        gh, gs = oadj.reference_gradients(Hn, Sn, g)
        QHr = torch.from_numpy(np.conj(gh).T.astype(np.complex128).copy()).cuda()[None]
        QSr = torch.from_numpy(np.conj(gs).T.astype(np.complex128).copy()).cuda()[None]
        d_ref = adjoint.table_gradient_from_densities(tab, plan, QHr, QSr)
        gt = torch.from_numpy(g).cuda()[None]
        d_ours = adjoint.table_gradient(tab, plan, V, w, gt)
        rel = float((d_ref - d_ours).abs().max() / d_ours.abs().max())
        if agree:
            assert rel < REF_RTOL, rel
        else:
            assert rel > 1e-3, rel


def test_optimizer_parameters_get_fd_gradients():
    """compute_multi_element_properties with trainable tables: parameter
    gradients of a smooth eigenvalue loss equal central differences."""
    rng = np.random.default_rng(21)
    z, pos, cell = synth.random_cell(rng, 6, [6, 14], 1.8)
    skfs = synth.make_skf_set([6, 14], seed=2)
    model = MultiElementSkfParameterOptimizer(skf_dicts=skfs, device="cuda").to("cuda")
    geom = Geometry(z, pos, cell=cell, units="angstrom")
    shells = synth.shell_dict_spd([6, 14])
    kf = GEN_K[:3]

    def loss_fn():
        props, ok = model.compute_multi_element_properties(geometry=geom, shell_dict=shells, kpoints_frac=kf,
                                                           device="cuda")
        return _smooth_loss(props["eigenvalues"][0], torch.full((3,), 1 / 3, dtype=torch.float64,
                                                                 device="cuda"))

    loss = loss_fn()
    loss.backward()
    params = [p for p in model.parameters()]
    gen = torch.Generator(device="cpu").manual_seed(4)
    dirs = [torch.randn(p.shape, dtype=torch.float64, generator=gen).to(p.device) * p.detach().abs().max()
            for p in params]
    an = sum(float((p.grad * d).sum()) for p, d in zip(params, dirs))
    vals = []
    with torch.no_grad():
        for sgn in (1.0, -1.0):
            for p, d in zip(params, dirs):
                p.add_(sgn * FD_STEP * d)
            vals.append(float(loss_fn()))
            for p, d in zip(params, dirs):
                p.sub_(sgn * FD_STEP * d)
    fd = (vals[0] - vals[1]) / (2 * FD_STEP)
    assert abs(an - fd) <= FD_RTOL * abs(fd), (an, fd)


def test_gap_gradient_matches_central_differences():
    """Insulating H2 molecular crystal (s shells): band_gap_eV carries the
    gradient of cbm - vbm."""
    z = np.array([1, 1])
    pos = np.array([[0.0, 0.0, 0.0], [0.62, 0.31, 0.18]])
    cell = np.array([[4.1, 0.0, 0.0], [0.6, 4.4, 0.0], [0.3, 0.5, 4.7]])
    skfs = synth.make_skf_set([1], seed=3)
    model = MultiElementSkfParameterOptimizer(skf_dicts=skfs, device="cuda").to("cuda")
    geom = Geometry(z, pos, cell=cell, units="angstrom")

    def gap():
        props, _ = model.compute_multi_element_properties(geometry=geom, shell_dict={1: [0]}, kpoints=[2, 2, 2],
                                                          device="cuda", get_fermi=True)
        return props["band_gap_eV"]

    g0 = gap()
    assert float(g0.detach()) > 0.5
    g0.backward()
    params = list(model.parameters())
    gen = torch.Generator(device="cpu").manual_seed(8)
    # 10% of the table scale keeps the O(step^2) difference error below 1e-6
    dirs = [torch.randn(p.shape, dtype=torch.float64, generator=gen).to(p.device) * (0.1 * p.detach().abs().max())
            for p in params]
    an = sum(float((p.grad * d).sum()) for p, d in zip(params, dirs))
    vals = []
    with torch.no_grad():
        for sgn in (1.0, -1.0):
            for p, d in zip(params, dirs):
                p.add_(sgn * FD_STEP * d)
            vals.append(float(gap()))
            for p, d in zip(params, dirs):
                p.sub_(sgn * FD_STEP * d)
    fd = (vals[0] - vals[1]) / (2 * FD_STEP)
    assert abs(an - fd) <= FD_RTOL * abs(fd), (an, fd)


def _band_energy(tab, z, pos_bohr, cell_bohr, k, mu, kT):
    inp = engine.StructureInput(z, pos_bohr, cell_bohr, k, np.full(len(k), 1.0 / len(k)))
    plan = engine.plan_build(tab, [inp])
    H, S = engine.run_build(tab, plan)
    w, _, _ = engine.eigh(H, S)
    e = w[:, :plan.n_orb[0]].double()
    f = 1.0 / (1.0 + torch.exp((e - mu) / kT))
    return float((plan.k_weights.reshape(-1, 1) * f * e).sum())


@pytest.mark.parametrize("case", ["random3", "carbon_si"])
def test_forces_match_central_differences(case):
    """Position adjoint (dual-number derivatives of the build) vs central
    differences (step 1e-4 Bohr) of the band energy at fixed Fermi level;
    the forces sum to zero (translation invariance)."""
    from slakonet_amd.calculator import SimpleDftb
    from slakonet_amd.units import H2E
    species = [1, 22, 31] if case == "random3" else [6, 14]
    rng = np.random.default_rng(31 if case == "random3" else 32)
    z, pos, cell = synth.random_cell(rng, 6, species, 1.8)
    tab = TableSet(synth.make_skf_set(species, seed=5), synth.shell_dict_spd(species), device="cuda")
    geom = Geometry(z, pos, cell=cell, units="angstrom")
    calc = SimpleDftb(geom, shell_dict=synth.shell_dict_spd(species), tables=tab, kpoints_frac=GEN_K,
                      with_eigenvectors=True)
    calc()
    F = calc.get_forces()[0].cpu().numpy()
    assert np.abs(F.sum(0)).max() < 1e-9 * max(1.0, np.abs(F).max())
    mu = calc._edge_values()["mu"]
    kT = 0.025 / H2E
    D = np.random.default_rng(3).normal(size=F.shape)
    pb, cb = pos * AA_BOHR, cell * AA_BOHR
    ep = _band_energy(tab, z, pb + FD_STEP * D, cb, GEN_K, mu, kT)
    em = _band_energy(tab, z, pb - FD_STEP * D, cb, GEN_K, mu, kT)
    fd = (ep - em) / (2 * FD_STEP)
    an = -float((F * D).sum())
    assert abs(an - fd) <= FD_RTOL * abs(fd), (an, fd)
