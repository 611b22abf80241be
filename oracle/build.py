"""Dense H(k)/S(k) build (oracle, test infrastructure only).

Restates hs_matrix for periodic systems (slakonet/slaterkoster.py:103-441)
together with the image/phase machinery of Periodic
(slakonet/atoms.py:165-243, 528-558):

* images: integer cell vectors in the box |T_a| <= floor(21 |b_a|/2pi) + 1
  (Periodic cutoff 20 + 1 Bohr, :80-84, :165-215), x slowest, keeping images
  with any pair within 21 Bohr (:309-334, :505-510);
* a pair-image (i, j, T) contributes when 0 < |R_j + T - R_i| < 10 Bohr
  (:250-271), with unit vector u = (R_j + T - R_i) / |.| (:201, atoms.py:222);
* ell-pairs (l1 <= l2) loop (:207-210): for l1 == l2 only atom pairs i <= j
  are evaluated and the strict lower blocks (i > j) are conjugate transposes
  (:358-401); for l1 < l2 every ordered pair is evaluated and the
  (l2, l1) block is the conjugate transpose (:403);
* integrals come from the ordered-pair table (Z_i, Z_j, l1, l2) (:900-935);
* Bloch sum with phase exp(i 2pi k . T) (atoms.py:543-558, slaterkoster.py
  :971-1010);
* on-site H = [eps_l per orbital], S = 1 on the diagonal for every k
  (:407-439, 781-831; skfeed.py:675-739).

Tables use the reference Skf.to_dict layout (slakonet/skf.py:201-263):
``skfs[(Za, Zb)] = {"grid": [...], "hamiltonian": {"l1-l2": [[...]...]},
"overlap": {...}, "atomic_data": {"on_sites": [...], "occupations": [...]}}``.
"""
import numpy as np

from .interp import poly_interp_u
from .skrot import sub_block_rot

LPAIRS = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]


def orbital_layout(z, shell_dict):
    """Per atom: list of (l, first orbital index); total orbital count."""
    layout = []
    off = 0
    for zi in z:
        shells = []
        for l in shell_dict[int(zi)]:
            if l > 2:
                raise NotImplementedError("f shells are not on the hot path")
            shells.append((l, off))
            off += 2 * l + 1
        layout.append(shells)
    return layout, off


def reference_images(cell, positions, periodic_cutoff=21.0):
    inv = np.linalg.inv(cell)
    recip = np.linalg.norm(inv, axis=0)
    tmp = np.floor(periodic_cutoff * recip)
    rng = (1 + tmp).astype(int)
    axes = [np.arange(-r, r + 1) for r in rng]
    tx = np.repeat(axes[0], len(axes[1]) * len(axes[2]))
    ty = np.tile(np.repeat(axes[1], len(axes[2])), len(axes[0]))
    tz = np.tile(axes[2], len(axes[0]) * len(axes[1]))
    t = np.stack([tx, ty, tz], -1).astype(np.float64)
    tcart = t @ cell
    # images with any pair within the periodic cutoff
    pe = tcart[:, None, :] + positions[None, :, :]            # [c, j, 3]
    d = np.sqrt(((pe[:, None, :, :] - positions[None, :, None, :]) ** 2).sum(-1))  # [c, i, j]
    keep = (d <= periodic_cutoff).any(-1).any(-1)
    return t[keep], tcart[keep]


def _table(skfs, za, zb, kind, l1, l2):
    skf = skfs[(int(za), int(zb))]
    data = skf["hamiltonian" if kind == "H" else "overlap"][f"{l1}-{l2}"]
    yy = np.asarray(data, dtype=np.float64).T   # [n_grid, n_bond]
    grid = np.asarray(skf["grid"], dtype=np.float64)
    return yy, grid[1] - grid[0]


def build_hs(z, positions, cell, shell_dict, skfs, kpoints, cutoff=10.0, kinds=("H", "S")):
    """Return dict kind -> [nk, n, n] complex128 (Bohr, Hartree)."""
    z = np.asarray(z)
    positions = np.asarray(positions, dtype=np.float64)
    cell = np.asarray(cell, dtype=np.float64)
    kpoints = np.asarray(kpoints, dtype=np.float64).reshape(-1, 3)
    layout, n = orbital_layout(z, shell_dict)
    t, tcart = reference_images(cell, positions)
    phase = np.exp(1j * ((2.0 * np.pi * kpoints) @ t.T))     # [nk, c]
    nk = len(kpoints)
    na = len(z)
    pe = tcart[:, None, :] + positions[None, :, :]            # [c, j, 3]
    vec = pe[:, None, :, :] - positions[None, :, None, :]     # [c, i, j, 3]
    dist = np.sqrt((vec ** 2).sum(-1))                         # [c, i, j]
    out = {}
    for kind in kinds:
        mat = np.zeros((nk, n, n), dtype=np.complex128)
        for (l1, l2) in LPAIRS:
            for i in range(na):
                for (la, oa) in layout[i]:
                    if la != l1:
                        continue
                    for j in range(na):
                        if l1 == l2 and i > j:
                            continue
                        for (lb, ob) in layout[j]:
                            if lb != l2:
                                continue
                            cs = np.nonzero((dist[:, i, j] < cutoff) & (dist[:, i, j] > 0))[0]
                            if len(cs) == 0:
                                continue
                            d = dist[cs, i, j]
                            u = vec[cs, i, j] / d[:, None]
                            yy, step = _table(skfs, z[i], z[j], kind, l1, l2)
                            integ = poly_interp_u(yy, step, d)
                            blk = sub_block_rot(l1, l2, u, integ)       # [m, a, b]
                            acc = np.einsum("kc,cab->kab", phase[:, cs], blk)
                            mat[:, oa:oa + 2 * l1 + 1, ob:ob + 2 * l2 + 1] = acc
                            if l1 != l2 or i < j:
                                mat[:, ob:ob + 2 * l2 + 1, oa:oa + 2 * l1 + 1] = \
                                    np.conj(acc).transpose(0, 2, 1)
        diag = np.zeros(n)
        for i in range(na):
            on = skfs[(int(z[i]), int(z[i]))]["atomic_data"]["on_sites"]
            for (l, o) in layout[i]:
                diag[o:o + 2 * l + 1] = on[l] if kind == "H" else 1.0
        idx = np.arange(n)
        mat[:, idx, idx] += diag
        out[kind] = mat
    return out


def n_electrons(z, skfs):
    """Sum of homonuclear occupations per atom (slakonet/optim.py:1132-1163)."""
    return float(sum(sum(skfs[(int(zi), int(zi))]["atomic_data"]["occupations"]) for zi in z))


def build_hs_rows(z, positions, cell, shell_dict, skfs, kpoints, row_atoms, cutoff=10.0, kinds=("H", "S")):
    """Rows of the selected atoms of H(k) / S(k), [nk, rows, n] (hs_matrix,
    slakonet/slaterkoster.py:103-441, restricted to row atoms). Each block
    (i, j) is evaluated the way the full build obtains it: directly from the
    ordered pair (i, j) when l_i < l_j, or l_i == l_j and i <= j; otherwise as
    the conjugate transpose of block (j, i) from the pair (j, i) (:358-403),
    i.e. the SK block of the reversed table and direction, transposed."""
    z = np.asarray(z)
    positions = np.asarray(positions, dtype=np.float64)
    cell = np.asarray(cell, dtype=np.float64)
    kpoints = np.asarray(kpoints, dtype=np.float64).reshape(-1, 3)
    layout, n = orbital_layout(z, shell_dict)
    t, tcart = reference_images(cell, positions)
    phase = np.exp(1j * ((2.0 * np.pi * kpoints) @ t.T))     # [nk, c]
    nk = len(kpoints)
    rows = []
    for i in row_atoms:
        rows.extend(range(layout[i][0][1], layout[i][-1][1] + 2 * layout[i][-1][0] + 1))
    rmap = {r: q for q, r in enumerate(rows)}
    out = {}
    for kind in kinds:
        mat = np.zeros((nk, len(rows), n), dtype=np.complex128)
        for i in row_atoms:
            vec = tcart[:, None, :] + positions[None, :, :] - positions[i][None, None, :]   # [c, j, 3]
            dist = np.sqrt((vec ** 2).sum(-1))                                              # [c, j]
            for j in range(len(z)):
                cs = np.nonzero((dist[:, j] < cutoff) & (dist[:, j] > 0))[0]
                if len(cs) == 0:
                    continue
                d = dist[cs, j]
                u = vec[cs, j] / d[:, None]
                for (la, oa) in layout[i]:
                    for (lb, ob) in layout[j]:
                        if la < lb or (la == lb and i <= j):
                            yy, step = _table(skfs, z[i], z[j], kind, la, lb)
                            blk = sub_block_rot(la, lb, u, poly_interp_u(yy, step, d))
                        else:
                            yy, step = _table(skfs, z[j], z[i], kind, lb, la)
                            blk = sub_block_rot(lb, la, -u, poly_interp_u(yy, step, d)).transpose(0, 2, 1)
                        acc = np.einsum("kc,cab->kab", phase[:, cs], blk)
                        r0 = rmap[oa]
                        mat[:, r0:r0 + 2 * la + 1, ob:ob + 2 * lb + 1] += acc
        for i in row_atoms:
            on = skfs[(int(z[i]), int(z[i]))]["atomic_data"]["on_sites"]
            for (l, o) in layout[i]:
                for q in range(2 * l + 1):
                    mat[:, rmap[o + q], o + q] += on[l] if kind == "H" else 1.0
        out[kind] = mat
    out["rows"] = np.array(rows)
    return out
