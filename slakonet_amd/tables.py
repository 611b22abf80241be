"""HBM-resident Slater-Koster table sets for the build kernel.

``TableSet`` packs every ordered species pair's H and S tables into one
[rows, 20] float64 device tensor (10 H integrals then 10 S integrals per grid
row, canonical order ss-s, sp-s, sd-s, pp-s, pp-p, pd-s, pd-p, dd-s, dd-p,
dd-d) plus the per-pair/per-species metadata the kernel reads. It replaces
the per-call feed rebuild of the reference (_create_comprehensive_feed,
slakonet/optim.py:1076-1130; _get_hs_dict / _get_onsite_dict,
slakonet/skfeed.py:560-739): tables are packed once and stay on the device;
when built from trainable tensors the packing is differentiable.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np
import torch

from .skf import Skf

LKEYS = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
N_INT = 10
SLOT_OF_L = {0: [0], 1: [1, 2, 3], 2: [4, 5, 6, 7, 8]}


def species_slots(shells: Sequence[int]) -> List[int]:
    slots = []
    for l in shells:
        if l not in SLOT_OF_L:
            raise NotImplementedError("only s, p and d shells are supported on the device path")
        slots.extend(SLOT_OF_L[l])
    if len(set(shells)) != len(shells):
        raise ValueError("repeated shells of one angular momentum are not supported")
    return slots


def pair_rows(h: Dict, s: Dict, n_grid: int, dtype=torch.float64, device=None) -> torch.Tensor:
    """[n_grid, 20] rows for one ordered pair from {(l1,l2): [n_bond, n_grid]} dicts."""
    cols = []
    for src in (h, s):
        for (l1, l2) in LKEYS:
            t = src.get((l1, l2))
            if t is None:
                cols.append(torch.zeros(l1 + 1, n_grid, dtype=dtype, device=device))
            else:
                t = t.to(device=device, dtype=dtype)
                if t.shape != (l1 + 1, n_grid):
                    raise ValueError(f"table ({l1},{l2}) has shape {tuple(t.shape)}, expected {(l1 + 1, n_grid)}")
                cols.append(t)
    return torch.cat(cols, 0).transpose(0, 1)


class TableSet:
    """Device tables for a set of ordered species pairs.

    skfs: {(Za, Zb): Skf or reference-format dict}; shell_dict: {Z: [l, ...]}.
    ``h_override`` / ``s_override`` ({(Za, Zb): {(l1, l2): tensor}}) replace
    the table values (trainable parameters) while keeping the metadata.
    """

    def __init__(self, skfs: Dict[Tuple[int, int], object], shell_dict: Dict[int, List[int]],
                 device="cuda", h_override=None, s_override=None, tail: float = 1.0):
        self.device = torch.device(device)
        self.skfs = {(int(k[0]), int(k[1])): (v if isinstance(v, Skf) else Skf.from_dict(v))
                     for k, v in skfs.items()}
        self.shell_dict = {int(k): list(v) for k, v in shell_dict.items()}
        zs = sorted({z for pair in self.skfs for z in pair})
        self.species = zs
        self.index = {z: i for i, z in enumerate(zs)}
        self._lut = -np.ones(max(zs + [0]) + 1, dtype=np.int32)
        for z, i in self.index.items():
            self._lut[z] = i
        self.tail = float(tail)
        ns = len(zs)
        self.pairs = sorted(self.skfs)
        pair_index = -np.ones(ns * ns, dtype=np.int32)
        row0, ngrid, step = [], [], []
        rows = 0
        for pid, (za, zb) in enumerate(self.pairs):
            skf = self.skfs[(za, zb)]
            pair_index[self.index[za] * ns + self.index[zb]] = pid
            ng = int(skf.grid.numel())
            if ng < 9:
                raise ValueError(f"table ({za},{zb}) has {ng} grid points; the 8-point stencil needs >= 9")
            row0.append(rows)
            ngrid.append(ng)
            step.append(skf.grid_step)
            rows += ng
        self.n_rows = rows
        norb = np.zeros(ns, dtype=np.int32)
        slot = np.zeros((ns, 9), dtype=np.int32)
        onsite = np.zeros((ns, 3), dtype=np.float64)
        self.occupation_sum = {}
        for z in zs:
            i = self.index[z]
            shells = self.shell_dict.get(z)
            if shells is None:
                raise KeyError(f"shell_dict has no entry for Z={z}")
            sl = species_slots(shells)
            norb[i] = len(sl)
            slot[i, :len(sl)] = sl
            homo = self.skfs.get((z, z))
            if homo is not None and homo.atomic:
                ons = homo.on_sites.detach().cpu().numpy()
                for l in range(min(3, len(ons))):
                    onsite[i, l] = ons[l]
                self.occupation_sum[z] = float(homo.occupations.sum())
        self.norb = norb
        self.meta = {
            "pair_row0": torch.tensor(row0, dtype=torch.int32, device=self.device),
            "pair_ngrid": torch.tensor(ngrid, dtype=torch.int32, device=self.device),
            "pair_step": torch.tensor(step, dtype=torch.float64, device=self.device),
            "pair_index": torch.from_numpy(pair_index).to(self.device),
            "spec_norb": torch.from_numpy(norb).to(self.device),
            "spec_slot": torch.from_numpy(slot.reshape(-1)).to(self.device),
            "spec_onsite": torch.from_numpy(onsite.reshape(-1)).to(self.device),
        }
        self.pair_index_np = pair_index
        self.update(h_override, s_override)

    def update(self, h_override=None, s_override=None):
        """(Re)pack the [rows, 20] table tensor; differentiable w.r.t. overrides."""
        if h_override is None and s_override is None:
            # constant tables: pack on the host and upload once
            host = np.zeros((self.n_rows, 20), dtype=np.float64)
            row = 0
            for (za, zb) in self.pairs:
                skf = self.skfs[(za, zb)]
                ng = int(skf.grid.numel())
                col = 0
                for src in (skf.hamiltonian, skf.overlap):
                    for (l1, l2) in LKEYS:
                        t = src.get((l1, l2))
                        if t is not None:
                            host[row:row + ng, col:col + l1 + 1] = t.detach().cpu().numpy().T
                        col += l1 + 1
                row += ng
            self.tables = torch.from_numpy(host).to(self.device)
            return self.tables
        parts = []
        for (za, zb) in self.pairs:
            skf = self.skfs[(za, zb)]
            ng = int(skf.grid.numel())
            h = dict(skf.hamiltonian)
            s = dict(skf.overlap)
            if h_override is not None and (za, zb) in h_override:
                h.update({tuple(k): v for k, v in h_override[(za, zb)].items()})
            if s_override is not None and (za, zb) in s_override:
                s.update({tuple(k): v for k, v in s_override[(za, zb)].items()})
            parts.append(pair_rows(h, s, ng, device=self.device))
        self.tables = torch.cat(parts, 0).contiguous()
        return self.tables

    def species_index(self, z: Iterable[int]) -> np.ndarray:
        z = np.asarray(z, dtype=np.int64).reshape(-1)
        if len(z) and (z.min() < 0 or z.max() >= len(self._lut)):
            bad = z[(z < 0) | (z >= len(self._lut))][0]
            raise KeyError(f"no SK tables for species Z={int(bad)}")
        idx = self._lut[z]
        if len(idx) and idx.min() < 0:
            raise KeyError(f"no SK tables for species Z={int(z[np.argmin(idx)])}")
        return idx

    def check_pairs(self, z: Iterable[int]):
        present = sorted({int(v) for v in z})
        for a in present:
            if (a, a) not in self.skfs or not self.skfs[(a, a)].atomic:
                raise KeyError(f"no homonuclear table with on-site data for Z={a}")
            for b in present:
                if (a, b) not in self.skfs:
                    raise KeyError(f"no SK table for the ordered pair ({a}, {b})")

    def n_electrons(self, z: Iterable[int]) -> float:
        """Sum of homonuclear occupations per atom (slakonet/optim.py:1132-1163)."""
        tot = 0.0
        for v in z:
            v = int(v)
            if v in self.occupation_sum:
                tot += self.occupation_sum[v]
            else:
                tot += float(v)
        return tot

    def c_struct(self, tables: Optional[torch.Tensor] = None):
        from ._lib import SlkTableSet
        t = self.tables if tables is None else tables
        m = self.meta
        return SlkTableSet(t.data_ptr(), m["pair_row0"].data_ptr(), m["pair_ngrid"].data_ptr(),
                           m["pair_step"].data_ptr(), m["pair_index"].data_ptr(), m["spec_norb"].data_ptr(),
                           m["spec_slot"].data_ptr(), m["spec_onsite"].data_ptr(), len(self.species), self.tail)
