"""Slater-Koster table container with the reference dict format.

``Skf`` keeps the fields of slakonet/skf.py:152-199 and round-trips the
dictionary layout of Skf.to_dict / Skf.from_dict (slakonet/skf.py:201-353):
``hamiltonian`` / ``overlap`` keyed by (l1, l2) with [n_bond, n_grid]
tensors (sigma, pi, delta rows), a uniform ``grid`` (row j at (j+1)*step),
``hs_cutoff`` and optional atomic data (on_sites ordered s, p, d;
occupations).
"""
from __future__ import annotations

from typing import Any, Dict, Optional

import numpy as np
import torch


def _t(x, dtype=torch.float64):
    if isinstance(x, torch.Tensor):
        return x if x.dtype == dtype else x.to(dtype)
    return torch.as_tensor(np.asarray(x, dtype=np.float64), dtype=dtype)


def _key(k):
    if isinstance(k, str):
        a, b = k.split("-")
        return int(a), int(b)
    return int(k[0]), int(k[1])


class Skf:
    def __init__(self, atom_pair, hamiltonian, overlap, grid, hs_cut=None, on_sites=None,
                 occupations=None, hubbard_us=None, mass=None, **_):
        self.atom_pair = torch.as_tensor(atom_pair, dtype=torch.long)
        self.hamiltonian = {_key(k): _t(v) for k, v in hamiltonian.items()}
        self.overlap = {_key(k): _t(v) for k, v in overlap.items()}
        self.grid = _t(grid)
        steps = self.grid[1:] - self.grid[:-1]
        if len(steps) and float((steps - steps[0]).abs().max()) >= 1e-5:
            raise ValueError(f"SK table grid must be uniform (spacings vary from {float(steps.min())} "
                             f"to {float(steps.max())} Bohr)")
        self.hs_cutoff = hs_cut
        check = [v is not None for v in (on_sites, occupations)]
        if any(check) and not all(check):
            raise ValueError("on_sites and occupations must be given together")
        self.atomic = all(check)
        self.on_sites = _t(on_sites) if on_sites is not None else None
        self.occupations = _t(occupations) if occupations is not None else None
        self.hubbard_us = _t(hubbard_us) if hubbard_us is not None else None
        self.mass = _t(mass) if mass is not None else None

    @property
    def grid_step(self) -> float:
        return float(self.grid[1] - self.grid[0])

    @classmethod
    def from_dict(cls, data: Dict[str, Any], dtype=None, device=None) -> "Skf":
        atomic = data.get("atomic_data") if data.get("atomic", False) or "atomic_data" in data else None
        return cls(
            data["atom_pair"], data["hamiltonian"], data["overlap"], data["grid"], data.get("hs_cutoff"),
            on_sites=(atomic or {}).get("on_sites"), occupations=(atomic or {}).get("occupations"),
            hubbard_us=(atomic or {}).get("hubbard_us"), mass=(atomic or {}).get("mass"))

    def to_dict(self) -> Dict[str, Any]:
        def ser(v):
            return v.detach().cpu().tolist() if isinstance(v, torch.Tensor) else v
        d = {
            "atom_pair": ser(self.atom_pair),
            "hamiltonian": {f"{k[0]}-{k[1]}": ser(v) for k, v in self.hamiltonian.items()},
            "overlap": {f"{k[0]}-{k[1]}": ser(v) for k, v in self.overlap.items()},
            "grid": ser(self.grid),
            "hs_cutoff": ser(self.hs_cutoff),
            "atomic": self.atomic,
        }
        if self.atomic:
            d["atomic_data"] = {"on_sites": ser(self.on_sites), "occupations": ser(self.occupations),
                                "hubbard_us": ser(self.hubbard_us), "mass": ser(self.mass)}
        return d

    def __repr__(self):
        return f"Skf({int(self.atom_pair[0])}-{int(self.atom_pair[1])})"
