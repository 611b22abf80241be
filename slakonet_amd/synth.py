"""Seeded synthetic SK tables, structures and request streams.

These generate the five benchmark workloads (2-atom diamond cell, 64-atom
rocksalt supercell, the 2-256-atom serving stream, the fitting batch and the
1024-atom cubic cell). Tables use the reference Skf dict layout
(slakonet/skf.py:201-263): ``hamiltonian``/``overlap`` keyed "l1-l2" with
[n_bond, n_grid] arrays ordered sigma, pi, delta.

Decisions (stated once here and in the design notes):

* grid r_i = 0.02 (i + 1) Bohr, 661 points (to 7 Angstrom);
* h(r) = A exp(-b r), s(r) = B exp(-c r), r in Angstrom, A in eV converted
  with 27.2113845 eV/Ha (slakonet/units.py:66);
* integrals with l1 == l2 are symmetric in the species pair, those with
  l1 < l2 are drawn independently per ordered pair;
* every species carries s, p and d shells;
* valence electrons per species follow the main-group / transition-metal
  group count (``valence_electrons``), split over s, p, d occupations;
* random cells are scaled up to at least ``VOLUME_PER_ATOM`` A^3 per atom so
  that rejection sampling at the minimum distance terminates; k-meshes use
  N_a = max(1, ceil(|b*_a| / 0.2 A^-1)) with b* the crystallographic
  reciprocal vector (no 2 pi), Monkhorst-Pack (2m+1)/(2N) points.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from .units import AA_BOHR, BOHR_AA, HARTREE_EV

N_GRID = 661
GRID_STEP = 0.02
GRID = GRID_STEP * (np.arange(N_GRID) + 1.0)
# canonical integral order: (l1, l2, bond)
INTEGRALS = [(0, 0, 0), (0, 1, 0), (0, 2, 0), (1, 1, 0), (1, 1, 1),
             (1, 2, 0), (1, 2, 1), (2, 2, 0), (2, 2, 1), (2, 2, 2)]
LKEYS = ["0-0", "0-1", "0-2", "1-1", "1-2", "2-2"]
SYMMETRIC = [i for i, (l1, l2, _) in enumerate(INTEGRALS) if l1 == l2]
VOLUME_PER_ATOM = 15.0


def valence_electrons(z: int) -> int:
    """Valence electron count used for synthetic species."""
    if z <= 2:
        return z
    for lo, core in ((10, 2), (18, 10), (30, 18), (36, 28), (48, 36), (54, 46), (56, 54)):
        if z <= lo:
            return z - core
    return 3


def occupations(z: int) -> List[float]:
    v = valence_electrons(z)
    s = min(v, 2)
    p = min(max(v - 2, 0), 6)
    return [float(s), float(p), float(v - s - p)]


def _params(rng, n):
    a = rng.uniform(-3.0, -0.5, n)
    b = rng.uniform(0.8, 1.5, n)
    bb = rng.uniform(0.05, 0.3, n)
    c = rng.uniform(1.0, 1.8, n)
    return a, b, bb, c


def pair_parameters(za: int, zb: int, seed: int):
    """(A [eV], b [1/A], B, c [1/A]) per canonical integral for ordered pair."""
    lo, hi = min(za, zb), max(za, zb)
    sym = _params(np.random.default_rng([seed, lo, hi, 0]), 10)
    asym = _params(np.random.default_rng([seed, za, zb, 1]), 10)
    out = [np.array(sym[i]) for i in range(4)]
    for i in range(4):
        for k in range(10):
            if k not in SYMMETRIC:
                out[i][k] = asym[i][k]
    return tuple(out)


def onsite_energies(z: int, seed: int) -> np.ndarray:
    """[eps_s, eps_p, eps_d] in Hartree."""
    rng = np.random.default_rng([seed, z, 2])
    es = rng.uniform(-15.0, -5.0)
    ep = es + rng.uniform(3.0, 8.0)
    ed = ep + rng.uniform(3.0, 8.0)
    return np.array([es, ep, ed]) / HARTREE_EV


def tables_from_parameters(a, b, bb, c, grid=GRID):
    r = grid * BOHR_AA
    h = (a[:, None] / HARTREE_EV) * np.exp(-b[:, None] * r[None, :])
    s = bb[:, None] * np.exp(-c[:, None] * r[None, :])
    return h, s


def _split(tab):
    """[10, n_grid] canonical rows -> {"l1-l2": [n_bond, n_grid]}."""
    out = {}
    row = 0
    for key in LKEYS:
        l1 = int(key[0])
        out[key] = tab[row:row + l1 + 1].copy()
        row += l1 + 1
    return out


def make_skf(za: int, zb: int, seed: int, scale: Optional[Tuple[np.ndarray, ...]] = None) -> dict:
    """Skf dict for ordered pair (za, zb); ``scale`` multiplies (A, b, B, c)."""
    a, b, bb, c = pair_parameters(za, zb, seed)
    if scale is not None:
        a, b, bb, c = a * scale[0], b * scale[1], bb * scale[2], c * scale[3]
    h, s = tables_from_parameters(a, b, bb, c)
    d = {
        "atom_pair": [int(za), int(zb)],
        "hamiltonian": _split(h),
        "overlap": _split(s),
        "grid": GRID.copy(),
        "hs_cutoff": float(GRID[-1]),
        "atomic": za == zb,
    }
    if za == zb:
        d["atomic_data"] = {
            "on_sites": onsite_energies(za, seed).tolist(),
            "hubbard_us": [0.3, 0.3, 0.3],
            "mass": [2.0 * za],
            "occupations": occupations(za),
        }
    return d


def make_skf_set(species, seed: int, teacher_seed: Optional[int] = None) -> Dict[Tuple[int, int], dict]:
    """All ordered pairs over ``species``. With ``teacher_seed`` the (A,b,B,c)
    parameters are multiplied by 1 + N(0, 0.1) drawn from that seed."""
    out = {}
    for za in species:
        for zb in species:
            scale = None
            if teacher_seed is not None:
                rng = np.random.default_rng([teacher_seed, za, zb, 3])
                scale = tuple(1.0 + 0.1 * rng.standard_normal(10) for _ in range(4))
                lo, hi = min(za, zb), max(za, zb)
                rs = np.random.default_rng([teacher_seed, lo, hi, 4])
                sym = tuple(1.0 + 0.1 * rs.standard_normal(10) for _ in range(4))
                for i in range(4):
                    scale[i][SYMMETRIC] = sym[i][SYMMETRIC]
            out[(int(za), int(zb))] = make_skf(int(za), int(zb), seed, scale)
    return out


def shell_dict_spd(species):
    return {int(z): [0, 1, 2] for z in species}


# --------------------------------------------------------------------------
# structures
# --------------------------------------------------------------------------
@dataclass
class Structure:
    z: np.ndarray                 # [N] atomic numbers
    positions: np.ndarray         # [N, 3] Angstrom
    cell: np.ndarray              # [3, 3] Angstrom (rows)
    kpoints: np.ndarray           # [nk, 3] fractional
    k_weights: np.ndarray         # [nk]
    kmesh: Optional[Tuple[int, int, int]] = None
    meta: dict = field(default_factory=dict)

    @property
    def n_atoms(self):
        return len(self.z)

    def geometry(self):
        from .geometry import Geometry
        return Geometry(self.z, self.positions, cell=self.cell, units="angstrom")


def mp_mesh(n):
    from .geometry import monkhorst_pack
    return monkhorst_pack(n)


def gamma_mesh(n):
    ax = [np.arange(m) / m for m in n]
    k = np.stack(np.meshgrid(*ax, indexing="ij"), -1).reshape(-1, 3)
    return k, np.full(len(k), 1.0 / len(k))


def diamond_si():
    """2-atom diamond cell, a = 5.43 A, Z = 14, 8x8x8 MP mesh."""
    a = 5.43
    cell = 0.5 * a * np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]])
    pos = np.array([[0.0, 0.0, 0.0], [0.25 * a] * 3])
    k, w = mp_mesh((8, 8, 8))
    return Structure(np.array([14, 14]), pos, cell, k, w, (8, 8, 8), {"name": "diamond-2"})


def rocksalt_64(seed=1, za=12, zb=8, sigma=0.05):
    """2x2x2 supercell of the 8-atom rocksalt cube (a = 4.2 A) with N(0, 0.05 A)
    displacements; Gamma-centred 4x4x4 mesh."""
    a = 4.2
    fa = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
    fb = np.array([[.5, 0, 0], [0, .5, 0], [0, 0, .5], [.5, .5, .5]])
    zs, ps = [], []
    for i in range(2):
        for j in range(2):
            for k in range(2):
                sh = np.array([i, j, k], dtype=float)
                for f in fa:
                    zs.append(za)
                    ps.append((f + sh) * a)
                for f in fb:
                    zs.append(zb)
                    ps.append((f + sh) * a)
    rng = np.random.default_rng(seed)
    pos = np.array(ps) + rng.normal(0.0, sigma, (len(ps), 3))
    cell = np.eye(3) * 2 * a
    k, w = gamma_mesh((4, 4, 4))
    return Structure(np.array(zs), pos, cell, k, w, None, {"name": "rocksalt-64"})


def _cell_from_params(lengths, angles_deg):
    al, be, ga = np.radians(angles_deg)
    a, b, c = lengths
    va = np.array([a, 0.0, 0.0])
    vb = np.array([b * np.cos(ga), b * np.sin(ga), 0.0])
    cx = c * np.cos(be)
    cy = c * (np.cos(al) - np.cos(be) * np.cos(ga)) / np.sin(ga)
    cz2 = c * c - cx * cx - cy * cy
    if cz2 <= 1e-6 * c * c:
        return None
    return np.array([va, vb, [cx, cy, np.sqrt(cz2)]])


def kmesh_for_cell(cell, spacing=0.2):
    recip = np.linalg.inv(cell).T          # rows b*_a (no 2 pi), 1/A
    return tuple(max(1, int(math.ceil(np.linalg.norm(recip[a]) / spacing - 1e-9))) for a in range(3))


_SHIFTS = np.array([[i, j, k] for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)], dtype=float)
_BATCH = 64


def _min_image_d2(cands, pos, cell, inv, shifts_c):
    """Squared minimum-image distance of each candidate to each position
    (reduced vector plus the 27 neighbouring images, exact for skewed cells)."""
    d = cands[:, None, :] - pos[None, :, :]
    f = d @ inv
    f -= np.round(f)
    d = f @ cell
    dd = d[:, :, None, :] + shifts_c[None, None, :, :]
    return (dd * dd).sum(-1).min(-1)


def _place_atoms(rng, cell, n_atoms, dmin, max_tries):
    """Uniform positions with rejection at dmin; candidates are proposed in
    batches of 64 and accepted in proposal order."""
    inv = np.linalg.inv(cell)
    # with every lattice-plane spacing above 2 dmin the reduced vector
    # (fractional components in [-1/2, 1/2]) is the only image that can be
    # closer than dmin, so the 27-image scan is needed only for thin cells
    spacing = 1.0 / np.linalg.norm(inv, axis=0)
    shifts_c = (_SHIFTS if spacing.min() <= 2.0 * dmin else np.zeros((1, 3))) @ cell
    pos = np.zeros((n_atoms, 3))
    npl = 0
    tries = 0
    d2min = dmin * dmin
    while npl < n_atoms and tries < max_tries:
        cands = rng.uniform(0.0, 1.0, (_BATCH, 3)) @ cell
        tries += _BATCH
        if npl:
            ok = _min_image_d2(cands, pos[:npl], cell, inv, shifts_c).min(1) >= d2min
        else:
            ok = np.ones(_BATCH, dtype=bool)
        start = npl
        for c in np.nonzero(ok)[0]:
            if npl >= n_atoms:
                break
            if npl > start and _min_image_d2(cands[c:c + 1], pos[start:npl], cell, inv, shifts_c).min() < d2min:
                continue
            pos[npl] = cands[c]
            npl += 1
    return pos if npl == n_atoms else None


def _lattice_ok(cell, dmin):
    t = _SHIFTS[np.any(_SHIFTS != 0, axis=1)] @ cell
    return np.sqrt((t * t).sum(-1)).min() >= dmin


def random_cell(rng, n_atoms, species, dmin=1.8, cubic_length=None, max_tries=200000):
    """Random lattice (lengths U(3,12) A, angles U(60,120) deg) scaled to at
    least VOLUME_PER_ATOM A^3/atom, uniform positions with rejection at dmin
    under the minimum-image convention; species drawn uniformly per atom."""
    while True:
        if cubic_length is not None:
            cell = np.eye(3) * cubic_length
        else:
            cell = _cell_from_params(rng.uniform(3.0, 12.0, 3), rng.uniform(60.0, 120.0, 3))
            if cell is None:
                continue
            vol = abs(np.linalg.det(cell))
            need = n_atoms * VOLUME_PER_ATOM
            if vol < need:
                cell = cell * (need / vol) ** (1.0 / 3.0)
        if not _lattice_ok(cell, dmin):
            continue
        pos = _place_atoms(rng, cell, n_atoms, dmin, max_tries)
        if pos is not None:
            break
        if cubic_length is not None:
            raise RuntimeError("could not place atoms at the requested density")
    z = np.asarray(species)[rng.integers(0, len(species), n_atoms)]
    return z, pos, cell


def stream_structure(seed: int, index: int, n_min=2, n_max=256, z_max=65, dmin=1.8):
    """Request ``index`` of the serving stream."""
    rng = np.random.default_rng([seed, index])
    n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
    nsp = int(rng.integers(1, 5))
    species = rng.choice(np.arange(1, z_max + 1), size=nsp, replace=False)
    z, pos, cell = random_cell(rng, n, species, dmin)
    mesh = kmesh_for_cell(cell)
    k, w = mp_mesh(mesh)
    return Structure(z, pos, cell, k, w, mesh, {"index": index, "species": sorted(int(s) for s in species)})


def poisson_arrivals(n, rate, seed):
    rng = np.random.default_rng([seed, 99])
    return np.cumsum(rng.exponential(1.0 / rate, n))


def fitting_batch(seed=3, n_cells=32, species=(6, 14, 31)):
    """Fitting cells: 8-64 atoms (log-uniform), 3 species, random cells."""
    out = []
    for i in range(n_cells):
        rng = np.random.default_rng([seed, i, 7])
        n = int(round(math.exp(rng.uniform(math.log(8), math.log(64)))))
        z, pos, cell = random_cell(rng, n, list(species), 1.8)
        mesh = kmesh_for_cell(cell)
        k, w = mp_mesh(mesh)
        out.append(Structure(z, pos, cell, k, w, mesh, {"index": i}))
    return out


def large_cubic(seed=5, n_atoms=1024, length=22.0, species=(14, 6, 8), dmin=1.9):
    """1024-atom cubic cell, Gamma plus four complex k-points."""
    rng = np.random.default_rng([seed, 11])
    z, pos, cell = random_cell(rng, n_atoms, list(species), dmin, cubic_length=length,
                               max_tries=20_000_000)
    k = np.array([[0, 0, 0], [0.25, 0, 0], [0, 0.25, 0], [0, 0, 0.25], [0.25, 0.25, 0.25]], float)
    w = np.full(len(k), 1.0 / len(k))
    return Structure(z, pos, cell, k, w, None, {"name": "cubic-1024"})
