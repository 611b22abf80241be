"""DFTB+ Slater-Koster (.skf) text files: read and write (simple v1.0 format).

Layout of a file "A-B.skf" (DFTB+ "Format of the v1.0 Slater-Koster Files"):

* line 1: grid spacing (Bohr) and number of grid points; row i (1-based)
  of the table holds the integrals at r = i * spacing;
* line 2, homonuclear files only: Ed Ep Es SPE Ud Up Us fd fp fs
  (on-site energies, spin-polarisation error, Hubbard U, occupations);
* line 3: mass, repulsive polynomial c2..c9, cutoff and ten unused values;
* then one line per grid point with 20 columns
  Hdd0 Hdd1 Hdd2 Hpd0 Hpd1 Hpp0 Hpp1 Hsd0 Hsp0 Hss0 then the same for S
  (0 = sigma, 1 = pi, 2 = delta);
* an optional repulsive "Spline" section (not used on this path).

Entries may use the "n*value" repeat notation and commas as separators.
The column permutation to the (l1, l2) ordering used here is the one the
reference keeps in Skf._sorter (slakonet/skf.py:92). The extended format
with f orbitals (first line starting with "@") is not read: the device path
holds s, p and d shells only.
"""
from __future__ import annotations

import os
import re
from typing import Dict, Iterable, Optional, Tuple

import numpy as np

from .units import CHEMICAL_SYMBOLS, SYMBOL_TO_Z

# file column -> canonical position: ss, sp, sd, pp(s,p), pd(s,p), dd(s,p,d)
_FILE_TO_CANON = [9, 8, 7, 5, 6, 3, 4, 0, 1, 2]
_KEYS = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]

_REPEAT = re.compile(r"(\d+)\*([-+0-9.eEdD]+)")


def _numbers(line: str) -> np.ndarray:
    text = line.replace(",", " ")
    text = _REPEAT.sub(lambda m: " ".join([m.group(2)] * int(m.group(1))), text)
    text = text.replace("d", "e").replace("D", "e")
    return np.array([float(x) for x in text.split()], dtype=np.float64)


def read_skf(path: str, atom_pair: Optional[Tuple[int, int]] = None) -> dict:
    """Parse one .skf file into the reference's Skf dict layout."""
    if atom_pair is None:
        stem = os.path.splitext(os.path.basename(path))[0]
        a, b = stem.split("-")
        atom_pair = (SYMBOL_TO_Z[a], SYMBOL_TO_Z[b])
    za, zb = int(atom_pair[0]), int(atom_pair[1])
    with open(path) as f:
        lines = [ln for ln in f.read().splitlines()]
    if lines and lines[0].lstrip().startswith("@"):
        raise NotImplementedError(f"{path}: the extended (f-orbital) skf format is not supported")
    head = _numbers(lines[0])
    step, n_grid = float(head[0]), int(head[1])
    pos = 1
    atomic = None
    if za == zb:
        a2 = _numbers(lines[1])
        ed, ep, es, _spe, ud, up, us, fd, fp, fs = a2[:10]
        atomic = {"on_sites": [es, ep, ed], "hubbard_us": [us, up, ud], "occupations": [fs, fp, fd]}
        pos = 2
    a3 = _numbers(lines[pos])
    pos += 1
    rows = []
    while len(rows) < n_grid and pos < len(lines):
        vals = _numbers(lines[pos])
        pos += 1
        if vals.size == 0:
            continue
        if vals.size != 20:
            raise ValueError(f"{path}: integral row {len(rows) + 1} has {vals.size} values, expected 20")
        rows.append(vals)
    if len(rows) != n_grid:
        raise ValueError(f"{path}: {len(rows)} integral rows, header says {n_grid}")
    table = np.stack(rows)                       # [n_grid, 20]
    h = table[:, :10][:, _FILE_TO_CANON].T       # [10, n_grid] canonical
    s = table[:, 10:][:, _FILE_TO_CANON].T
    grid = step * np.arange(1, n_grid + 1, dtype=np.float64)

    def split(t):
        out, r = {}, 0
        for (l1, l2) in _KEYS:
            out[f"{l1}-{l2}"] = t[r:r + l1 + 1].tolist()
            r += l1 + 1
        return out

    d = {"atom_pair": [za, zb], "hamiltonian": split(h), "overlap": split(s), "grid": grid.tolist(),
         "hs_cutoff": float(grid[-1]), "atomic": atomic is not None}
    if atomic is not None:
        atomic["mass"] = [float(a3[0])]
        d["atomic_data"] = atomic
    return d


def write_skf(path: str, skf: dict):
    """Write an Skf dict (reference layout) as a simple-format .skf file."""
    grid = np.asarray(skf["grid"], dtype=np.float64)
    step = float(grid[1] - grid[0])
    n_grid = len(grid)

    def canon(part):
        rows = []
        for (l1, l2) in _KEYS:
            v = part.get(f"{l1}-{l2}", part.get((l1, l2)))
            v = np.zeros((l1 + 1, n_grid)) if v is None else np.asarray(v, dtype=np.float64)
            rows.append(v.reshape(l1 + 1, n_grid))
        return np.concatenate(rows, 0)            # [10, n_grid]

    h, s = canon(skf["hamiltonian"]), canon(skf["overlap"])
    inv = np.argsort(_FILE_TO_CANON)
    table = np.concatenate([h[inv].T, s[inv].T], 1)
    za, zb = skf["atom_pair"]
    out = [f"{step:.12f} {n_grid}"]
    atomic = skf.get("atomic_data")
    mass = 0.0
    if int(za) == int(zb):
        ons = list(atomic["on_sites"]) if atomic else [0.0, 0.0, 0.0]
        hub = list(atomic.get("hubbard_us") or [0.0, 0.0, 0.0]) if atomic else [0.0, 0.0, 0.0]
        occ = list(atomic["occupations"]) if atomic else [0.0, 0.0, 0.0]
        ons += [0.0] * (3 - len(ons))
        hub += [0.0] * (3 - len(hub))
        occ += [0.0] * (3 - len(occ))
        out.append(" ".join(f"{x:.12e}" for x in [ons[2], ons[1], ons[0], 0.0, hub[2], hub[1], hub[0],
                                                   occ[2], occ[1], occ[0]]))
        if atomic and atomic.get("mass") is not None:
            mass = float(np.ravel(atomic["mass"])[0])
    out.append(" ".join(f"{x:.6e}" for x in [mass] + [0.0] * 19))
    out += [" ".join(f"{x:.15e}" for x in row) for row in table]
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def read_skf_directory(directory: str, elements: Iterable[str]) -> Dict[Tuple[int, int], dict]:
    """All "A-B.skf" files for ordered pairs over ``elements`` present in ``directory``."""
    els = sorted(set(elements))
    out = {}
    for a in els:
        for b in els:
            path = os.path.join(directory, f"{a}-{b}.skf")
            if os.path.exists(path):
                out[(SYMBOL_TO_Z[a], SYMBOL_TO_Z[b])] = read_skf(path, (SYMBOL_TO_Z[a], SYMBOL_TO_Z[b]))
    if not out:
        raise FileNotFoundError(f"no .skf files for {els} in {directory}")
    return out


def pair_file_name(za: int, zb: int) -> str:
    return f"{CHEMICAL_SYMBOLS[za]}-{CHEMICAL_SYMBOLS[zb]}.skf"
