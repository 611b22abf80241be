"""Pins of the CPU oracle against what the mathematics fixes (no GPU).

Each test is chosen so that a plausible restatement mistake fails it: a wrong
stencil row, a swapped sign in a rotation, a dropped conjugate in the
Hermitian assembly, a missing k-weight or spin factor.
"""
import numpy as np
import pytest
import scipy.linalg

from oracle import build, interp, post, skrot, solve
from sk_textbook import block as sk_block
from slakonet_amd import synth
from slakonet_amd.units import AA_BOHR

LP = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]


# ---------------------------------------------------------------- interpolation
def test_interp_reproduces_degree7_polynomials():
    rng = np.random.default_rng(0)
    step, n = 0.02, 300
    x = step * (np.arange(n) + 1)
    coef = rng.normal(size=8)
    yy = np.polyval(coef, x)[:, None]
    r = rng.uniform(0.01, (n - 1) * step, 500)   # includes extrapolation below the first point
    got = interp.poly_interp_u(yy, step, r)[:, 0]
    np.testing.assert_allclose(got, np.polyval(coef, r), rtol=1e-9, atol=1e-9)


def test_interp_uses_rows_up_to_second_last_and_clamps():
    # the clamp ind_last <= n_grid means the last table row is never read
    step, n = 0.1, 40
    yy = np.zeros((n, 1))
    yy[-1] = 1e6
    r = np.linspace(0.05, (n - 1) * step - 1e-9, 100)
    assert np.abs(interp.poly_interp_u(yy, step, r)).max() == 0.0
    # a change in row n-2 is seen near the end of the table
    yy2 = np.zeros((n, 1))
    yy2[-2] = 1.0
    assert abs(interp.poly_interp_u(yy2, step, [(n - 1) * step])[0, 0] - 1.0) < 1e-12


def test_interp_exponential_table_accuracy():
    x = synth.GRID
    yy = np.exp(-1.3 * x * 0.529177249)[:, None]
    r = np.random.default_rng(1).uniform(0.5, 12.0, 1000)
    got = interp.poly_interp_u(yy, synth.GRID_STEP, r)[:, 0]
    np.testing.assert_allclose(got, np.exp(-1.3 * r * 0.529177249), rtol=1e-11)


def test_interp_tail_endpoints():
    step, n = 0.02, 400
    x = step * (np.arange(n) + 1)
    yy = np.exp(-x)[:, None]
    tail = 1.0
    rmax = (n - 1) * step + tail
    # tail region starts at ind == n_grid; value -> 0 at r_max and beyond -> 0
    r_in = np.array([n * step + 1e-9, rmax - 1e-9, rmax + 0.5])
    got = interp.poly_interp_u(yy, step, r_in)[:, 0]
    assert abs(got[1]) < 1e-12 and got[2] == 0.0
    # continuity of the tail polynomial with the table value at (n-1)*step
    y1 = yy[n - 2, 0]
    val = interp.poly5_zero(np.array([y1]), np.array([0.0]), np.array([0.0]), np.array([-tail]), -tail)
    assert abs(val[0] - y1) < 1e-15


# ---------------------------------------------------------------- SK rotation
@pytest.mark.parametrize("lp", LP)
def test_rotation_matches_textbook_sk_table(lp):
    l1, l2 = lp
    rng = np.random.default_rng(2)
    for _ in range(200):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        v = rng.normal(size=l1 + 1)
        a = skrot.sub_block_rot(l1, l2, u[None], v[None])[0]
        np.testing.assert_allclose(a, sk_block(l1, l2, u, v), atol=1e-13)


def test_rotation_axis_blocks():
    # bond along z: p_z-p_z is sigma, p_x-p_x is pi; s-d_z2 is sigma
    u = np.array([[0.0, 0.0, 1.0]])
    b = skrot.sub_block_rot(1, 1, u, np.array([[2.0, 0.5]]))[0]
    np.testing.assert_allclose(np.diag(b), [0.5, 2.0, 0.5], atol=1e-15)
    b = skrot.sub_block_rot(0, 2, u, np.array([[3.0]]))[0]
    np.testing.assert_allclose(b[0], [0, 0, 3.0, 0, 0], atol=1e-15)


@pytest.mark.parametrize("l", [1, 2])
def test_rotation_matrices_orthogonal(l):
    rng = np.random.default_rng(3)
    u = rng.normal(size=(50, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    for rot in ((skrot.rot_p_yz, skrot.rot_p_xy) if l == 1 else (skrot.rot_d_yz, skrot.rot_d_xy)):
        r = rot(u)
        np.testing.assert_allclose(np.einsum("mij,mkj->mik", r, r), np.broadcast_to(np.eye(2 * l + 1), r.shape),
                                   atol=1e-12)


# ---------------------------------------------------------------- build
@pytest.fixture(scope="module")
def diamond():
    st = synth.diamond_si()
    skfs = synth.make_skf_set([14], seed=0)
    k = st.kpoints[:40]
    hs = build.build_hs(st.z, st.positions * AA_BOHR, st.cell * AA_BOHR, synth.shell_dict_spd([14]), skfs, k)
    return st, skfs, k, hs


def test_build_hermitian_and_pd(diamond):
    _, _, _, hs = diamond
    for m in (hs["H"], hs["S"]):
        np.testing.assert_allclose(m, np.conj(m.transpose(0, 2, 1)), atol=1e-15)
    assert min(np.linalg.eigvalsh(s).min() for s in hs["S"]) > 0.5


def test_build_periodic_in_k_and_real_at_gamma():
    st = synth.diamond_si()
    skfs = synth.make_skf_set([14], seed=0)
    k = np.array([[0.1, 0.2, 0.3], [1.1, 0.2, -0.7], [0.0, 0.0, 0.0]])
    hs = build.build_hs(st.z, st.positions * AA_BOHR, st.cell * AA_BOHR, synth.shell_dict_spd([14]), skfs, k)
    np.testing.assert_allclose(hs["H"][0], hs["H"][1], atol=1e-13)
    assert np.abs(hs["H"][2].imag).max() < 1e-15


def test_build_dimer_block_equals_sk_table():
    """Two atoms 2.1 A apart in a 30 A cube: the off-diagonal block is the
    textbook SK block of the ordered table, the reverse block its transpose
    from the reversed table (hetero pair, asymmetric sp/sd/pd tables)."""
    za, zb = 6, 14
    skfs = synth.make_skf_set([za, zb], seed=4)
    cell = np.eye(3) * 30.0 * AA_BOHR
    u = np.array([0.3, -0.5, 0.8])
    u /= np.linalg.norm(u)
    d = 2.1 * AA_BOHR
    pos = np.array([[1.0, 1.0, 1.0], [1.0, 1.0, 1.0] + d * u])
    sd = synth.shell_dict_spd([za, zb])
    hs = build.build_hs(np.array([za, zb]), pos, cell, sd, skfs, np.zeros((1, 3)))
    H = hs["H"][0].real
    for kind, mat in (("hamiltonian", H), ("overlap", hs["S"][0].real)):
        def integ(a, b, l1, l2):
            yy = np.asarray(skfs[(a, b)][kind][f"{l1}-{l2}"]).T
            return interp.poly_interp_u(yy, synth.GRID_STEP, [d])[0]
        off = {0: 0, 1: 1, 2: 4}
        for (l1, l2) in LP:
            ref = sk_block(l1, l2, u, integ(za, zb, l1, l2))
            got = mat[off[l1]:off[l1] + 2 * l1 + 1, 9 + off[l2]:9 + off[l2] + 2 * l2 + 1]
            np.testing.assert_allclose(got, ref, atol=1e-14)
            if l1 != l2:
                # (l2 on atom a, l1 on atom b) comes from the reversed table, direction -u
                ref2 = sk_block(l1, l2, -u, integ(zb, za, l1, l2)).T
                got2 = mat[off[l2]:off[l2] + 2 * l2 + 1, 9 + off[l1]:9 + off[l1] + 2 * l1 + 1]
                np.testing.assert_allclose(got2, ref2, atol=1e-14)


# ---------------------------------------------------------------- solve
def test_solve_matches_scipy_generalized(diamond):
    _, _, _, hs = diamond
    w, v = solve.solve_all(hs["H"][:5], hs["S"][:5], vectors=True)
    for ik in range(5):
        ref = scipy.linalg.eigh(hs["H"][ik], hs["S"][ik], eigvals_only=True)
        np.testing.assert_allclose(w[ik], ref, atol=1e-12)
        z = v[ik]
        np.testing.assert_allclose(z.conj().T @ hs["S"][ik] @ z, np.eye(18), atol=1e-11)


# ---------------------------------------------------------------- post
def test_dos_closed_form_and_integral():
    eig = np.array([[-0.1, 0.0, 0.2], [-0.05, 0.1, 0.3]])
    w = np.array([0.25, 0.75])
    grid, d = post.dos(eig, w, 0.0, energy_range=(-20, 20), num_points=20001, sigma=0.1)
    # integral = sum of weights * bands = 3 (no spin factor)
    assert abs(np.trapezoid(d, grid) - 3.0) < 1e-8
    # peak value at an isolated level: w / (sigma sqrt(2 pi)) (grid centred on the level)
    lv = 0.3 * 27.211
    grid, d = post.dos(eig, w, 0.0, energy_range=(lv - 1.0, lv + 1.0), num_points=2001, sigma=0.1)
    assert abs(d[1000] - 0.75 / (0.1 * np.sqrt(2 * np.pi))) < 1e-9


def test_fermi_bisection_root():
    rng = np.random.default_rng(5)
    eig = np.sort(rng.uniform(-0.5, 0.5, (4, 10)), axis=1)
    w = np.full(4, 0.25)
    mu = post.fermi_bisect(eig, 7.0, w)
    assert abs(post.electron_count(eig, w, mu) - 7.0) < 1e-10


def test_fermi_reference_iteration_invariants():
    # symmetric two-level system: N(mean) = Ne exactly -> returns the mean unchanged
    eig = np.array([[-0.2, 0.2]])
    assert post.fermi_search(eig, 2.0, [1.0]) == 0.0
    # generic case: the returned mu is the first iterate with |dL/dmu| < tol
    rng = np.random.default_rng(6)
    eig = np.sort(rng.uniform(-0.3, 0.3, (3, 12)), axis=1)
    tr = []
    mu = post.fermi_search(eig, 11.0, np.full(3, 1 / 3), trace=tr)
    assert tr[-1][0] == mu and abs(tr[-1][2]) < 1e-6
    assert all(abs(g) >= 1e-6 for (_, _, g) in tr[:-1])
    assert abs(tr[1][0] - (tr[0][0] - tr[0][1] / (tr[0][2] + 1e-12))) < 1e-15


def test_gap_rules():
    eig = np.array([[-0.3, -0.1, 0.2], [-0.2, -0.05, 0.15]])
    mu = 0.0
    g = post.band_gap(eig, mu)
    assert abs(g["vbm"] - (-0.05 * 27.211)) < 1e-12 and abs(g["cbm"] - 0.15 * 27.211) < 1e-12
    assert g["vbm_kpoint"] == 1 and g["cbm_kpoint"] == 1 and g["direct"]
    bs = post.band_structure(eig, mu)
    assert abs(bs["minimum_direct_gap"] - 0.2 * 27.211) < 1e-12
    # metal rule: a level within 1e-3 eV of mu
    assert post.band_gap(np.array([[-0.1, 0.0]]), 0.0)["metal"]


# ---- eigenvalue adjoints (oracle/adjoint.py) --------------------------------

def _herm(rng, n, cplx):
    a = rng.normal(size=(n, n)) + (1j * rng.normal(size=(n, n)) if cplx else 0.0)
    return (a + np.conj(a.T)) / 2


def _spd(rng, n, cplx):
    b = rng.normal(size=(n, n)) + (1j * rng.normal(size=(n, n)) if cplx else 0.0)
    return b @ np.conj(b.T) + n * np.eye(n)


@pytest.mark.parametrize("cplx", [False, True])
def test_eig_densities_match_central_differences(cplx):
    """de = c^H (dH - e dS) c as densities vs central FD (step 1e-4)."""
    from oracle import adjoint
    rng = np.random.default_rng(11 + cplx)
    n = 7
    H, S = _herm(rng, n, cplx), _spd(rng, n, cplx)
    g = rng.normal(size=n)
    QH, QS, _ = adjoint.eig_densities(H, S, g)
    DH, DS = _herm(rng, n, cplx), 0.1 * _herm(rng, n, cplx)
    h = 1e-4
    loss = lambda t: g @ scipy.linalg.eigh(H + t * DH, S + t * DS, eigvals_only=True)
    fd = (loss(h) - loss(-h)) / (2 * h)
    an = adjoint.directional(QH, DH) + adjoint.directional(QS, DS)
    assert abs(an - fd) < 1e-6 * abs(fd)


def test_reference_autograd_agrees_at_gamma_only():
    """The reference's _SymEigB eigenvalue gradient (V diag(g) V^T, no conj)
    equals the exact one for real matrices and not for complex ones."""
    from oracle import adjoint
    rng = np.random.default_rng(5)
    n = 6
    for cplx in (False, True):
        H, S = _herm(rng, n, cplx), _spd(rng, n, cplx)
        g = rng.normal(size=n)
        QH, QS, _ = adjoint.eig_densities(H, S, g)
        gh, gs = adjoint.reference_gradients(H, S, g)
        DH, DS = _herm(rng, n, cplx), 0.1 * _herm(rng, n, cplx)
        ours = adjoint.directional(QH, DH) + adjoint.directional(QS, DS)
        ref = adjoint.torch_directional(gh, DH) + adjoint.torch_directional(gs, DS)
        if cplx:
            assert abs(ref - ours) > 1e-3 * abs(ours)
        else:
            assert abs(ref - ours) < 1e-10 * max(1.0, abs(ours))


def test_atom_projected_dos_sums_to_total_for_orthonormal_vectors():
    """With S = I the type weights of every state sum to 1, so the projected
    DOS summed over types is the total DOS at uniform k-weights."""
    rng = np.random.default_rng(9)
    nk, n = 3, 10
    eig, vecs = [], []
    for _ in range(nk):
        a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        h = (a + a.conj().T) / 20
        w, v = np.linalg.eigh(h)
        eig.append(w)
        vecs.append(v)
    eig = np.array(eig)
    types = np.array([0, 0, 1, 1, 1, 2, 2, 0, 1, 2])
    mu = 0.01
    grid, pd = post.atom_projected_dos(eig, vecs, types, 3, mu, sigma=0.1, energy_range=(-8, 6), n_points=500)
    _, tot = post.dos(eig, np.full(nk, 1.0 / nk), mu, energy_range=(-8, 6), num_points=500, sigma=0.1)
    assert np.allclose(pd.sum(0), tot, rtol=1e-12, atol=1e-12)
    # one state on a single type-1 orbital -> only type 1 gets its Gaussian
    vec = np.zeros((1, n, 1)); vec[0, 3, 0] = 1.0
    _, one = post.atom_projected_dos(np.array([[0.0]]), vec, types, 3, 0.0, n_points=11, energy_range=(-1, 1))
    assert one[0].max() == 0.0 and one[2].max() == 0.0 and abs(one[1][5] - 1 / (0.1 * np.sqrt(2 * np.pi))) < 1e-12


# ---- pins added for the self-image terms, the poly5 tail, the reference Fermi
# ---- iteration's gradient and the properties dict ---------------------------

def _analytic_integrals(za, zb, seed, r_bohr):
    """The synthetic tables' closed forms h = A e^{-b r}, s = B e^{-c r} (r in A)."""
    from slakonet_amd.units import BOHR_AA, HARTREE_EV
    a, b, bb, c = synth.pair_parameters(za, zb, seed)
    r = r_bohr * BOHR_AA
    return (a / HARTREE_EV) * np.exp(-b * r), bb * np.exp(-c * r)


def _lp_slice(l1, l2):
    """Positions of the (l1, l2) integrals in the canonical 10-integral order."""
    start = {(0, 0): 0, (0, 1): 1, (0, 2): 2, (1, 1): 3, (1, 2): 5, (2, 2): 7}[(l1, l2)]
    return slice(start, start + min(l1, l2) + 1)


def test_build_single_atom_bloch_sum_closed_form():
    """One atom in a simple-cubic cell (a = 2.5 A, spd): every off-site term is
    a self-image (i == j, T != 0). H(k) must equal the textbook Bloch sum
    eps + sum_{0<|T|<10 Bohr} E(T) exp(2 pi i k.n_T), with E the Slater-Koster
    element between the orbital at the origin and the one at T
    (slaterkoster.py:271,358-403). Dropping the self-images leaves a diagonal
    H; a flipped phase sign gives H(-k) (the s-p blocks are odd in T)."""
    z0, seed, a_aa = 14, 3, 2.5
    skfs = synth.make_skf_set([z0], seed=seed)
    a = a_aa * AA_BOHR
    cell = np.eye(3) * a
    pos = np.array([[0.3, -0.2, 0.1]])
    k = np.array([[0.13, -0.27, 0.41], [0.5, 0.25, 0.0], [0.0, 0.0, 0.0]])
    hs = build.build_hs(np.array([z0]), pos, cell, synth.shell_dict_spd([z0]), skfs, k)
    shells = [(0, 0), (1, 1), (2, 4)]
    eps = synth.onsite_energies(z0, seed)
    for kind, q in (("H", 0), ("S", 1)):
        ref = np.zeros((len(k), 9, 9), dtype=np.complex128)
        for nvec in np.ndindex(7, 7, 7):
            t = np.array(nvec) - 3
            r = np.linalg.norm(t) * a
            if not (0.0 < r < 10.0):
                continue
            u = t / np.linalg.norm(t)
            integ = _analytic_integrals(z0, z0, seed, r)[q]
            ph = np.exp(2j * np.pi * (k @ t))
            for (la, oa) in shells:
                for (lb, ob) in shells:
                    if la <= lb:
                        e = sk_block(la, lb, u, integ[_lp_slice(la, lb)])
                    else:
                        e = sk_block(lb, la, -u, integ[_lp_slice(lb, la)]).T
                    ref[:, oa:oa + 2 * la + 1, ob:ob + 2 * lb + 1] += ph[:, None, None] * e[None]
        diag = np.concatenate([[eps[0]], [eps[1]] * 3, [eps[2]] * 5]) if kind == "H" else np.ones(9)
        ref[:, np.arange(9), np.arange(9)] += diag
        got = hs[kind]
        offdiag = got - got * np.eye(9)[None]
        assert np.abs(offdiag).max() > 1e-3 * np.abs(got).max()     # self-images really contribute
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-10 * np.abs(ref).max())


def test_interp_tail_is_quintic_matching_value_and_derivatives():
    """The tail (ind >= n_grid) is the quintic that leaves (n_grid-1) dr with
    the table's value, first and second derivative and reaches zero with zero
    first and second derivative at r_max = (n_grid-1) dr + tail
    (interpolation.py:173-252). Checked with a table e^{-x}, whose y' and y''
    are non-zero, by fitting a quintic to tail samples."""
    step, n, tail = 0.02, 420, 1.0
    x = step * (np.arange(n) + 1)
    yy = np.exp(-x)[:, None]
    x0 = (n - 1) * step
    rmax = x0 + tail
    r = np.linspace(n * step + 1e-6, rmax - 1e-6, 40)
    vals = interp.poly_interp_u(yy, step, r)[:, 0]
    c = np.polynomial.polynomial.polyfit(r - x0, vals, 5)
    p = np.polynomial.Polynomial(c)
    assert np.abs(p(r - x0) - vals).max() < 1e-13          # the tail is one quintic
    y0 = np.exp(-x0)
    assert abs(p(0.0) - y0) < 1e-12
    assert abs(p.deriv(1)(0.0) - (-y0)) < 1e-6 * y0        # central difference, delta 1e-5
    assert abs(p.deriv(2)(0.0) - y0) < 1e-4 * y0
    for d in range(3):
        assert abs(p.deriv(d)(tail) if d else p(tail)) < 1e-10
    # a wrong end-point derivative would show here: compare with the closed form
    q = interp.poly5_zero(np.array([y0]), np.array([-y0]), np.array([y0]), (r - rmax)[:, None], -tail)[:, 0]
    assert np.abs(q - vals).max() < 1e-5 * y0


def test_fermi_search_gradient_is_derivative_of_loss():
    """dL/dmu of the reference iteration (autograd of L = (N(mu) - Ne)^2,
    fermi.py:176-188) against a central difference of L, at every iterate."""
    rng = np.random.default_rng(21)
    eig = np.sort(rng.uniform(-0.3, 0.3, (4, 10)), axis=1)
    w = np.full(4, 0.25)
    ne = 9.0
    tr = []
    post.fermi_search(eig, ne, w, trace=tr)
    assert len(tr) >= 3
    h = 1e-6
    for mu, loss, grad in tr:
        L = lambda m: (post.electron_count(eig, w, m) - ne) ** 2
        assert abs(loss - L(mu)) < 1e-12 * max(1.0, loss)
        fd = (L(mu + h) - L(mu - h)) / (2 * h)
        assert abs(grad - fd) < 1e-6 * max(1.0, abs(fd))


def test_properties_hand_made_spectrum():
    """get_properties_dict (main.py:1657-1819) and the get_fermi block
    (optim.py:1011-1059) on a spectrum whose answers are known by hand."""
    eig = np.array([[-0.10, -0.05, 0.06], [-0.12, -0.04, 0.08]])
    w = np.array([0.5, 0.5])
    p = post.properties(eig, w, 4.0)
    mu = p["fermi_energy_Ha"]
    assert abs(post.electron_count(eig, w, mu) - 4.0) < 1e-10
    assert -0.04 < mu < 0.06
    H = 27.211
    assert abs(p["vbm_eV"] - (-0.04 * H)) < 1e-12 and abs(p["cbm_eV"] - 0.06 * H) < 1e-12
    assert p["vbm_kpoint"] == 1 and p["cbm_kpoint"] == 0 and not p["is_direct_gap"]
    assert abs(p["band_gap_eV"] - 0.10 * H) < 1e-12
    assert abs(p["valence_band_width_eV"] - 0.08 * H) < 1e-12
    assert abs(p["conduction_band_width_eV"] - 0.02 * H) < 1e-12
    assert abs(p["minimum_direct_gap_eV"] - 0.11 * H) < 1e-12      # k0: 0.06 - (-0.05)
    assert not p["is_direct_semiconductor"]
    assert abs(p["bandgap"] - 0.10 * H) < 1e-12 and abs(p["efermi"] - mu * H) < 1e-12
    grid = p["dos_energy_grid"]
    assert len(grid) == 5000 and grid[0] == -10.0 and grid[-1] == 10.0
    # every level lies well inside [-10, 10] eV: trapz total = sum_k w_k * n_bands = 3
    assert abs(p["dos_total_states"] - 3.0) < 1e-6
    # DOS at the grid point nearest the Fermi level, as a closed-form Gaussian sum
    e0 = grid[np.argmin(np.abs(grid))]
    g = (w[:, None] * np.exp(-0.5 * ((e0 - (eig - mu) * H) / 0.1) ** 2)).sum() / (0.1 * np.sqrt(2 * np.pi))
    assert abs(p["dos_at_fermi"] - g) < 1e-12 * max(1.0, g)
    # the DOS minimum within 2 eV of mu lies between the two edges, at the grid
    # point where the two edge Gaussians balance (other levels are > 2 eV away)
    lo, hi = (-0.04 - mu) * H, (0.06 - mu) * H
    assert lo < p["dos_gap_center_eV"] < hi
    assert abs(p["dos_gap_center_eV"] - 0.5 * (lo + hi)) < 0.3
    # metal rule: a level within 1e-3 eV of mu
    m = post.properties(np.array([[-0.1, 0.0, 0.0]]), [1.0], 4.0)
    assert m["band_gap_eV"] == 0.0 and m["vbm_eV"] == m["cbm_eV"] == m["fermi_energy_eV"]
    assert abs(m["fermi_energy_Ha"]) < 1e-5      # 2 f(-mu) = 1 + f(0.1 + mu): mu just above 0


def test_properties_without_occupied_states_falls_back():
    """No state below mu (Ne = 0): calculate_band_gap fails and the reference
    reports a zero Fermi level and zero edges (main.py:1677-1687), while the
    get_fermi block keeps the real mu and an infinite gap (optim.py:1011-1059)."""
    eig = np.array([[-0.1, 0.05, 0.2]])
    p = post.properties(eig, [1.0], 0.0)
    assert p["fermi_energy_eV"] == 0.0 and p["vbm_eV"] == 0.0 and p["cbm_eV"] == 0.0
    assert p["band_gap_eV"] == 0.0 and p["is_direct_gap"]
    assert p["efermi"] < -0.1 * 27.211 and p["bandgap"] == np.inf
    assert p["valence_band_width_eV"] == 0.0 and p["minimum_direct_gap_eV"] is None


def test_occupations_and_density_invariants():
    """fermi occupations (slaterkoster.py:25-74) and the density of
    main.py:158-161 on the 2-atom diamond cell at complex k: P = conj(rho) is
    the physical density C f C^H, so Tr(P S) = Ne and P S P = 2 P (f in {0, 2});
    rho is Hermitian. An odd Ne puts one electron in the next state."""
    assert post.occupations(5, 7.0).tolist() == [2.0, 2.0, 2.0, 1.0, 0.0]
    assert post.occupations(3, 4.0).tolist() == [2.0, 2.0, 0.0]
    st = synth.diamond_si()
    skfs = synth.make_skf_set([14], seed=0)
    k = np.array([[0.13, -0.21, 0.37], [0.5, 0.25, 0.0]])
    hs = build.build_hs(st.z, st.positions * AA_BOHR, st.cell * AA_BOHR, synth.shell_dict_spd([14]), skfs, k)
    w, v = solve.solve_all(hs["H"], hs["S"], vectors=True)
    ne = build.n_electrons(st.z, skfs)
    occ = post.occupations(18, ne)
    rho = post.density(v, occ)
    for ik in range(len(k)):
        p = np.conj(rho[ik])
        s = hs["S"][ik]
        np.testing.assert_allclose(p, p.conj().T, atol=1e-12)
        assert abs(np.trace(p @ s) - ne) < 1e-10
        np.testing.assert_allclose(p @ s @ p, 2.0 * p, atol=1e-10)
        assert np.abs(p.imag).max() > 1e-6          # complex k: conj matters


def test_row_restricted_build_equals_full_build():
    """oracle.build.build_hs_rows (the rows of a few atoms of a large cell)
    reproduces the corresponding rows of the full oracle build on a 3-species
    cell at complex k, including hetero (l1 != l2) and i > j blocks."""
    rng = np.random.default_rng(17)
    sp = [1, 22, 31]
    z, pos, cell = synth.random_cell(rng, 7, sp, 1.8)
    skfs = synth.make_skf_set(sp, seed=5)
    k = np.array([[0.1, -0.2, 0.3], [0.0, 0.0, 0.0]])
    sd = synth.shell_dict_spd(sp)
    full = build.build_hs(z, pos * AA_BOHR, cell * AA_BOHR, sd, skfs, k)
    part = build.build_hs_rows(z, pos * AA_BOHR, cell * AA_BOHR, sd, skfs, k, [0, 3, 6])
    for kind in ("H", "S"):
        np.testing.assert_allclose(part[kind], full[kind][:, part["rows"], :], rtol=0,
                                   atol=1e-13 * np.abs(full[kind]).max())


def test_electronic_energy_closed_forms():
    """Two levels mu -/+ a per k-point: f(-a) + f(a) = 1 and
    f(-a) - f(a) = tanh(a / 2kT), so the energy is sum_k w_k (mu - a_k tanh(a_k / 2kT)).
    A flipped exponent sign gives mu + a tanh(...); dropped weights change the sum.
    At kT -> 0 it is the weighted sum of the levels below mu."""
    mu = -0.21
    a = np.array([0.001, 0.003, 0.02])
    w = np.array([0.5, 0.3, 0.2])
    eig = np.stack([mu - a, mu + a], axis=1)
    kT = 0.025 / post.H2E
    want = float((w * (mu - a * np.tanh(a / (2 * kT)))).sum())
    assert abs(post.electronic_energy(eig, w, mu) - want) < 1e-14
    rng = np.random.default_rng(3)
    eig = np.sort(rng.uniform(-1.0, 0.5, (4, 7)), axis=1)
    w = rng.uniform(0.1, 1.0, 4)
    mu = 0.5 * (eig[:, 3].max() + eig[:, 4].min()) if eig[:, 3].max() < eig[:, 4].min() else -0.3
    cold = float((np.where(eig < mu, eig, 0.0) * w[:, None]).sum())
    assert abs(post.electronic_energy(eig, w, mu, kT_eV=1e-6) - cold) < 1e-12
    # a level at exactly zero carries no occupation (fermi.py:95); its energy term is zero either way
    eig0 = np.array([[-0.5, 0.0, 0.4]])
    assert post.electronic_energy(eig0, [1.0], 0.1) == pytest.approx(-0.5 * 1.0 / (1 + np.exp(-0.6 / kT))
                                                                      + 0.4 / (1 + np.exp(0.3 / kT)), abs=1e-15)
