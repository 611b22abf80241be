"""Structures, lattice translations and k-point sampling (host side).

Mirrors the parts of slakonet/atoms.py that the H/S build consumes:

* ``Geometry`` keeps atomic numbers, Cartesian positions and the cell in Bohr
  (slakonet/atoms.py:643-707, Pbc at :1318-1381).
* ``monkhorst_pack`` reproduces ``Periodic._super_sampling``
  (slakonet/atoms.py:394-446): k_i = (2 m + 1) / (2 N_i), x slowest.
* ``klines_points`` reproduces ``Periodic._klines`` (slakonet/atoms.py:448-502).
* ``lattice_translations`` enumerates integer cell vectors in the same x-slowest
  order as ``Periodic.get_cell_translations`` (slakonet/atoms.py:165-215); the
  range is chosen so that every atom pair closer than the interaction cutoff is
  covered, which is all the build needs (pairs beyond 10 Bohr are dropped at
  slakonet/slaterkoster.py:271).

All of this is O(N + n_images) host bookkeeping; the O(N^2 n_images) distance
work happens inside the device build kernel.
"""
from __future__ import annotations

import math
from typing import Optional, Sequence, Union

import numpy as np
import torch

from .units import AA_BOHR, CHEMICAL_SYMBOLS

_LENGTH = {"bohr": 1.0, "au": 1.0, "angstrom": AA_BOHR, "aa": AA_BOHR, "a": AA_BOHR}


def _as_np(x, dtype):
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy().astype(dtype)
    return np.asarray(x, dtype=dtype)


class Geometry:
    """A single periodic structure.

    Arguments follow ``slakonet.atoms.Geometry`` (slakonet/atoms.py:643):
    ``atomic_numbers`` [N] (or [1, N]), ``positions`` [N, 3] (or [1, N, 3]),
    ``cell`` [3, 3] (rows are lattice vectors), ``units`` applies to both
    positions and cell, ``frac`` marks positions as fractional coordinates.
    Internally everything is Bohr / float64.
    """

    def __init__(self, atomic_numbers, positions, cell=None, frac=None,
                 units: Optional[str] = "bohr", **kwargs):
        z = _as_np(atomic_numbers, np.int64)
        pos = _as_np(positions, np.float64)
        if z.ndim == 2:
            if z.shape[0] != 1:
                raise ValueError("one structure per Geometry (batch of 1)")
            z = z[0]
        if pos.ndim == 3:
            if pos.shape[0] != 1:
                raise ValueError("one structure per Geometry (batch of 1)")
            pos = pos[0]
        if pos.shape != (len(z), 3):
            raise ValueError("`atomic_numbers` & `positions` size mismatch")
        if cell is None:
            raise ValueError("the hot path handles periodic structures only; pass `cell`")
        c = _as_np(cell, np.float64)
        if c.ndim == 3:
            c = c[0]
        if c.shape != (3, 3) or not np.any(c != 0):
            raise ValueError("cell must be three non-zero lattice vectors")
        scale = _LENGTH[(units or "bohr").lower()]
        c = c * scale
        is_frac = False
        if frac is not None:
            f = np.asarray(frac if not isinstance(frac, torch.Tensor) else frac.cpu().numpy())
            is_frac = bool(np.any(f != 0))
        if is_frac:
            # slakonet/atoms.py:127-145: wrap into [0, 1) then to Cartesian.
            if np.any(pos >= 1) or np.any(pos < 0):
                pos = np.abs(pos) - np.floor(np.abs(pos))
            pos = pos @ c
        else:
            pos = pos * scale
        self.atomic_numbers_np = z
        self.positions_np = pos
        self.cell_np = c
        self.is_periodic = True

    # Tensor views matching the reference attribute names.
    @property
    def atomic_numbers(self) -> torch.Tensor:
        return torch.from_numpy(self.atomic_numbers_np)

    @property
    def positions(self) -> torch.Tensor:
        return torch.from_numpy(self.positions_np)

    @property
    def cell(self) -> torch.Tensor:
        return torch.from_numpy(self.cell_np).unsqueeze(0)

    @property
    def n_atoms(self) -> int:
        return int(len(self.atomic_numbers_np))

    def unique_atomic_numbers(self):
        return torch.from_numpy(np.unique(self.atomic_numbers_np))

    def __repr__(self):
        counts = {}
        for z in self.atomic_numbers_np:
            counts[int(z)] = counts.get(int(z), 0) + 1
        formula = "".join(f"{CHEMICAL_SYMBOLS[z]}{n if n > 1 else ''}" for z, n in counts.items())
        return f"Geometry({formula})"


def monkhorst_pack(nkpts: Sequence[int]):
    """k-points of ``Periodic._super_sampling`` (slakonet/atoms.py:394-446).

    Returns fractional k [nk, 3] (x slowest, z fastest) and weights [nk] = 1/nk.
    """
    n = [int(v) for v in (nkpts.tolist() if isinstance(nkpts, torch.Tensor) else nkpts)]
    if len(n) != 3 or min(n) < 1:
        raise ValueError("kpoints must be three positive integers")
    ax = [(0.5 + np.arange(m)) / m for m in n]
    kx = np.repeat(ax[0], n[1] * n[2])
    ky = np.tile(np.repeat(ax[1], n[2]), n[0])
    kz = np.tile(ax[2], n[0] * n[1])
    k = np.stack([kx, ky, kz], -1)
    w = np.full(len(k), 1.0 / len(k))
    return k, w


def klines_points(klines):
    """k-points of ``Periodic._klines`` (slakonet/atoms.py:448-502).

    Each row is [k1x, k1y, k1z, k2x, k2y, k2z, N]: N points from k1 to k2
    inclusive with uniform weights 1/total. The reference divides by N-1, which
    turns an N == 1 row into NaN k-points; that case raises here instead.
    """
    kl = _as_np(klines, np.float64)
    if kl.ndim == 3:
        if kl.shape[0] != 1:
            raise ValueError("one structure per call")
        kl = kl[0]
    if kl.ndim != 2 or kl.shape[1] != 7:
        raise ValueError("klines rows must be [kx1, ky1, kz1, kx2, ky2, kz2, N]")
    pts = []
    for row in kl:
        npt = int(row[6])
        if npt < 1:
            continue
        if npt == 1:
            raise ValueError("a k-line with N == 1 has no defined step (NaN in the reference)")
        delta = (row[3:6] - row[0:3]) / (npt - 1)
        for i in range(npt):
            pts.append(row[0:3] + delta * i)
    k = np.asarray(pts, dtype=np.float64).reshape(-1, 3)
    w = np.full(len(k), 1.0 / len(k))
    return k, w


def kpts_to_klines(kpts, default_points=10):
    """Pairs (k[0], k[1]), (k[2], k[3]), ... of a k-point list as k-line rows
    [k1, k2, default_points] (slakonet/optim.py:74-100)."""
    k = _as_np(kpts, np.float64).reshape(-1, 3)
    rows = [np.concatenate([k[i], k[i + 1], [float(default_points)]]) for i in range(0, len(k) - 1, 2)]
    return torch.as_tensor(np.stack(rows) if rows else np.zeros((0, 7)), dtype=torch.float64)


def lattice_translations(cell: np.ndarray, positions: np.ndarray, cutoff: float):
    """Integer cell vectors covering every pair closer than ``cutoff`` (Bohr).

    A pair (i, j, T) has fractional component (f_j - f_i + T_a) along reciprocal
    axis a; its length is at least |f_j - f_i + T_a| * d_a with d_a the lattice
    plane spacing, so |T_a| <= cutoff / d_a + max|f_j - f_i| suffices.
    Order: x slowest, z fastest (slakonet/atoms.py:181-206).
    """
    inv = np.linalg.inv(cell)            # columns are b_a / (2 pi)
    recip_norm = np.linalg.norm(inv, axis=0)   # 1 / d_a
    frac = positions @ inv
    spread = frac.max(0) - frac.min(0) if len(frac) else np.zeros(3)
    rng = [int(math.floor(cutoff * recip_norm[a] + spread[a])) + 1 for a in range(3)]
    axes = [np.arange(-r, r + 1) for r in rng]
    tx = np.repeat(axes[0], len(axes[1]) * len(axes[2]))
    ty = np.tile(np.repeat(axes[1], len(axes[2])), len(axes[0]))
    tz = np.tile(axes[2], len(axes[0]) * len(axes[1]))
    t_int = np.stack([tx, ty, tz], -1).astype(np.int32)
    t_cart = t_int.astype(np.float64) @ cell
    return t_int, t_cart


def prune_translations(t_int, t_cart, positions, cutoff):
    """Keep translations for which at least one pair lies inside ``cutoff``.

    Mirrors the image filter of ``Periodic._neighbourlist``
    (slakonet/atoms.py:309-334) at the build cutoff; O(n_T * N) via bounding
    spheres so it stays cheap for 256-atom cells.
    """
    if len(positions) == 0:
        return t_int[:0], t_cart[:0]
    centre = positions.mean(0)
    radius = np.sqrt(((positions - centre) ** 2).sum(-1)).max()
    d = np.sqrt(((t_cart) ** 2).sum(-1))
    keep = d < cutoff + 2.0 * radius + 1e-9
    return t_int[keep], t_cart[keep]
