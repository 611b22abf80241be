"""Fermi level, band gap, DOS and the properties dict (oracle, test infra only).

* fermi_search: the reference's half-step Newton on L(mu) = (N(mu) - Ne)^2
  (slakonet/fermi.py:139-198): mu0 = mean of all eigenvalues, stop when
  |dL/dmu| < tol (checked before the update), mu <- mu - L / (dL/dmu + 1e-12),
  N(mu) = 2 sum_k w_k sum_b f((e_kb - mu)/kT), f = 1/(exp(x)+1). dL/dmu is
  written the way autograd evaluates it: df/dmu = f^2 exp(x) / kT.
* band_gap: calculate_band_gap (slakonet/main.py:637-707).
* band_structure: calculate_band_structure_properties (slakonet/main.py:1539-1642).
* dos: calculate_dos (slakonet/main.py:709-797), Fermi-shifted, eV.
* occupations / density: fermi (slakonet/slaterkoster.py:25-74) and the
  density of SimpleDftb.solve_kpoint (slakonet/main.py:157-161).
* electronic_energy: _calculate_electronic_energy (slakonet/main.py:1430-1448).
* properties: get_properties_dict(include_dos_data=True)
  (slakonet/main.py:1657-1819) plus the get_fermi block of
  compute_multi_element_properties (slakonet/optim.py:1011-1059).
"""
import numpy as np

H2E = 27.211


def fermi_search(eig, n_electrons, k_weights, kT=0.01, tol=1e-6, max_iter=100, trace=None):
    eig = np.asarray(eig, dtype=np.float64).reshape(-1, np.shape(eig)[-1])
    w = np.asarray(k_weights, dtype=np.float64).reshape(-1)
    mu = eig.mean()
    for it in range(max_iter):
        x = (eig - mu) / kT
        ex = np.exp(x)
        f = 1.0 / (ex + 1.0)
        total = 2.0 * (f * w[:, None]).sum()
        loss = (total - n_electrons) ** 2
        dn = 2.0 * ((f * f * ex / kT) * w[:, None]).sum()
        grad = 2.0 * (total - n_electrons) * dn
        if trace is not None:
            trace.append((mu, loss, grad))
        if abs(grad) < tol:
            break
        mu = mu - loss / (grad + 1e-12)
    return mu


def electron_count(eig, k_weights, mu, kT=0.01):
    eig = np.asarray(eig, dtype=np.float64)
    w = np.asarray(k_weights, dtype=np.float64).reshape(-1)
    f = 1.0 / (np.exp((eig - mu) / kT) + 1.0)
    return 2.0 * (f * w[:, None]).sum()


def band_gap(eig, mu):
    ev = np.asarray(eig) * H2E
    fev = mu * H2E
    if np.any(np.abs(ev - fev) < 1e-3):
        return dict(gap=0.0, vbm=fev, cbm=fev, direct=False, vbm_kpoint=0, cbm_kpoint=0, metal=True)
    occ = ev < fev
    un = ev >= fev
    vbm = ev[occ].max()
    cbm = ev[un].min()
    vk = int(np.argwhere(ev == vbm)[0][0])
    ck = int(np.argwhere(ev == cbm)[0][0])
    return dict(gap=cbm - vbm, vbm=vbm, cbm=cbm, direct=(vk == ck), vbm_kpoint=vk,
                cbm_kpoint=ck, metal=False)


def band_structure(eig, mu):
    e = (np.asarray(eig) - mu) * H2E
    val = e[e < 0]
    con = e[e >= 0]
    vbm = val.max() if len(val) else -np.inf
    vbw = (val.max() - val.min()) if len(val) else 0.0
    cbm = con.min() if len(con) else np.inf
    cbw = (con.max() - con.min()) if len(con) else 0.0
    gap = cbm - vbm if (np.isfinite(vbm) and np.isfinite(cbm)) else 0.0
    dg = []
    for ik in range(e.shape[0]):
        kb = e[ik]
        v = kb[kb < 0]
        c = kb[kb > 0]
        if len(v) and len(c):
            dg.append(c.min() - v.max())
    mdg = min(dg) if dg else np.inf
    return dict(band_gap=gap, valence_band_width=vbw, conduction_band_width=cbw,
                minimum_direct_gap=(mdg if np.isfinite(mdg) else None),
                is_direct_semiconductor=(abs(gap - mdg) < 0.01) if np.isfinite(mdg) else False)


def dos(eig, k_weights, mu, energy_range=(-10.0, 10.0), num_points=5000, sigma=0.1,
        fermi_shift=True):
    e = np.asarray(eig, dtype=np.float64)
    if fermi_shift:
        e = e - mu
    e = e * H2E
    grid = np.linspace(energy_range[0], energy_range[1], num_points)
    w = np.asarray(k_weights, dtype=np.float64).reshape(-1)
    norm = sigma * np.sqrt(2.0 * np.pi)
    out = np.zeros(num_points)
    for ik in range(e.shape[0]):
        g = np.exp(-0.5 * ((grid[None, :] - e[ik][:, None]) / sigma) ** 2) / norm
        out += w[ik] * g.sum(0)
    return grid, out


def properties(eig, k_weights, n_electrons, get_fermi=True, fermi_method="bisection"):
    """Numeric content of the reference properties dict for one structure."""
    eig = np.asarray(eig, dtype=np.float64)
    if fermi_method == "reference":
        mu = fermi_search(eig, n_electrons, k_weights)
    else:
        mu = fermi_bisect(eig, n_electrons, k_weights)
    try:
        bg = band_gap(eig, mu)
        mu_dict = mu
    except ValueError:
        # no occupied or no unoccupied state: the reference's except branch
        # (main.py:1677-1687) reports a zero Fermi level and zero edges
        bg = dict(gap=0.0, vbm=0.0, cbm=0.0, direct=True, vbm_kpoint=0, cbm_kpoint=0, metal=False)
        mu_dict = 0.0
    bs = band_structure(eig, mu)
    grid, d = dos(eig, k_weights, mu)
    fi = int(np.argmin(np.abs(grid)))
    region = np.abs(grid) < 2.0
    gc = grid[region][int(np.argmin(d[region]))]
    p = {
        "fermi_energy_eV": mu_dict * H2E,
        "fermi_energy_Ha": mu_dict,
        "band_gap_eV": bg["gap"],
        "vbm_eV": bg["vbm"],
        "cbm_eV": bg["cbm"],
        "is_direct_gap": bg["direct"],
        "vbm_kpoint": bg["vbm_kpoint"],
        "cbm_kpoint": bg["cbm_kpoint"],
        "valence_band_width_eV": bs["valence_band_width"],
        "conduction_band_width_eV": bs["conduction_band_width"],
        "minimum_direct_gap_eV": bs["minimum_direct_gap"],
        "is_direct_semiconductor": bs["is_direct_semiconductor"],
        "nkpoints": eig.shape[0],
        "nbands": eig.shape[-1],
        "nelectrons": int(n_electrons),
        "dos_at_fermi": d[fi],
        "dos_gap_center_eV": gc,
        "dos_total_states": np.trapezoid(d, grid),
        "dos_energy_grid": grid,
        "dos_values": d,
    }
    if get_fermi:
        occ = eig <= mu
        vb = np.where(occ, eig, -np.inf).max()
        cb = np.where(~occ, eig, np.inf).min()
        p["efermi"] = mu * H2E
        p["bandgap"] = max(cb - vb, 0.0) * H2E
    return p


def fermi_bisect(eig, n_electrons, k_weights, kT=0.01):
    """Root of N(mu) = Ne by bisection to full double resolution."""
    eig = np.asarray(eig, dtype=np.float64)
    lo = eig.min() - 50 * kT
    hi = eig.max() + 50 * kT
    for _ in range(400):
        mid = 0.5 * (lo + hi)
        if mid <= lo or mid >= hi:
            break
        if electron_count(eig, k_weights, mid, kT) > n_electrons:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


def atom_projected_dos(eig, vecs, orb_type, n_types, mu, sigma=0.1, energy_range=(-8.0, 6.0), n_points=1000):
    """compute_atom_projected_dos (slakonet/predict_slakonet.py:119-210): per
    state the weight of atom type t is sum |psi_mu|^2 over the type's orbitals
    (S-orthonormal eigenvector, no overlap), Gaussians on grid + mu (eV),
    divided by the number of k-points. eig [nk, nb] Hartree; vecs [nk, n, nb]."""
    eig = np.asarray(eig, dtype=np.float64)
    grid = np.linspace(energy_range[0], energy_range[1], n_points)
    grid_ev = grid + mu * H2E
    orb_type = np.asarray(orb_type)
    norm = 1.0 / (sigma * np.sqrt(2.0 * np.pi))
    out = np.zeros((n_types, n_points))
    nk, nb = eig.shape
    for k in range(nk):
        p = np.abs(np.asarray(vecs[k])) ** 2                       # [n, nb]
        w = np.stack([p[orb_type == t].sum(0) for t in range(n_types)])   # [T, nb]
        g = norm * np.exp(-0.5 * ((grid_ev[None, :] - eig[k][:, None] * H2E) / sigma) ** 2)   # [nb, E]
        out += w @ g
    return grid, out / nk


def occupations(n_states, n_electrons):
    """fermi(eigenvalue, nelectron) without spin or smearing
    (slakonet/slaterkoster.py:25-74): 2 for the floor(Ne/2) lowest states, 1
    for the next one when Ne is odd."""
    if n_electrons < 1:
        raise ValueError("at least one electron")
    n_pairs = int(n_electrons // 2)
    occ = np.zeros(n_states)
    occ[:min(n_pairs, n_states)] = 2.0
    if n_electrons % 2 > 0 and n_pairs < n_states:
        occ[n_pairs] = 1.0
    return occ


def electronic_energy(eig, k_weights, mu, kT_eV=0.025):
    """_calculate_electronic_energy (slakonet/main.py:1430-1448): sum over k
    and bands of w_k f(e_kb) e_kb with Fermi-Dirac occupations
    f = 1/(1+exp((e-mu)/kT)) at kT = 0.025 eV (fermi_smearing,
    slakonet/fermi.py:54-96, which zeroes f where e == 0); no spin factor.
    eig [nk, nb] Hartree, mu Hartree -> Hartree."""
    eig = np.asarray(eig, dtype=np.float64)
    w = np.asarray(k_weights, dtype=np.float64).reshape(-1)
    kT = kT_eV / H2E
    f = 1.0 / (1.0 + np.exp((eig - mu) / kT))
    f = np.where(eig == 0.0, 0.0, f)
    return float((f * eig * w[:, None]).sum())


def density(vecs, occ):
    """rho = conj(C sqrt(f)) (C sqrt(f))^T per k (slakonet/main.py:158-161);
    vecs [nk, n, nb] eigenvectors in columns, occ [nb] -> [nk, n, n]."""
    c = np.asarray(vecs) * np.sqrt(np.asarray(occ))[None, None, :]
    return np.conj(c) @ np.swapaxes(c, -1, -2)
