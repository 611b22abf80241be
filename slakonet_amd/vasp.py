"""Training data from VASP runs: vasprun.xml reader, multi-file loader and the
multi-VASP training entry point.

Mirrors the data path of slakonet/optim.py:1446-1567 (MultiVaspDataLoader)
and :1570-1855 (train_multi_vasp_skf_parameters). The reference reads
vasprun.xml through jarvis-tools (``jarvis.io.vasp.outputs.Vasprun``,
requirements.txt: jarvis-tools>=2021.07.19), which is not part of this
package; ``Vasprun`` below is a reader of its own for the public vasprun.xml
layout (``<atominfo>``, ``<structure>``, ``<calculation>`` with ``<energy>``,
``<eigenvalues>`` and ``<dos>``), exposing the attributes the loader uses:

* ``all_structures`` -- every ``<structure>`` of the ionic steps, in file order
  (the ``finalpos`` block when the run has no ionic-step structures);
* ``final_energy`` -- ``e_fr_energy`` of the last ionic step (eV);
* ``get_indir_gap`` -- (gap, cbm, vbm) in eV: VBM is the highest eigenvalue
  with occupation above half the largest occupation of the file (1 or 2),
  CBM the lowest at or below it, over every spin channel and k-point;
  gap = max(cbm - vbm, 0);
* ``total_dos`` -- (energies, total spin 1[, total spin 2]) of the last
  ``<dos><total>`` block, energies as written (not shifted by E_F).

jarvis-tools' own rules for these quantities are not available here, so
their parity is unpinned; the tests pin this reader against vasprun files
written with known contents.

The training entry point solves every dataset of an epoch in one batched
device pass per size bucket (``fit.train_skf_parameters``) with the
reference's loss (|gap - target| per dataset, weighted by atom count,
:1701-1726; the reference computes DOS terms too, but overwrites them before
use, :1718).
"""
from __future__ import annotations

import glob
import os
import random
import xml.etree.ElementTree as ET
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from .units import AA_BOHR, CHEMICAL_SYMBOLS

SYMBOL_TO_Z = {s: z for z, s in enumerate(CHEMICAL_SYMBOLS) if z > 0}


def _floats(text: str) -> List[float]:
    return [float(x) for x in text.split()]


def _varray(node, name: str) -> np.ndarray:
    for va in node.iter("varray"):
        if va.get("name") == name:
            return np.array([_floats(v.text) for v in va.findall("v")], dtype=np.float64)
    raise ValueError(f"vasprun: varray {name!r} not found")


class VaspStructure:
    """Lattice (rows, Angstrom), fractional positions and species of one
    ``<structure>`` block."""

    def __init__(self, lattice: np.ndarray, frac: np.ndarray, species: Sequence[str]):
        self.lattice = np.asarray(lattice, dtype=np.float64)
        self.frac_coords = np.asarray(frac, dtype=np.float64)
        self.species = list(species)
        if self.lattice.shape != (3, 3) or self.frac_coords.shape != (len(self.species), 3):
            raise ValueError("vasprun: structure block does not match the atom list")

    @property
    def num_atoms(self) -> int:
        return len(self.species)

    @property
    def atomic_numbers(self) -> np.ndarray:
        return np.array([SYMBOL_TO_Z[s] for s in self.species], dtype=np.int64)

    @property
    def cart_coords(self) -> np.ndarray:
        return self.frac_coords @ self.lattice

    @property
    def volume(self) -> float:
        return float(abs(np.linalg.det(self.lattice)))

    @property
    def composition(self) -> Dict[str, int]:
        out: Dict[str, int] = {}
        for s in self.species:
            out[s] = out.get(s, 0) + 1
        return out

    @property
    def reduced_formula(self) -> str:
        comp = self.composition
        g = 0
        for n in comp.values():
            g = int(np.gcd(g, n))
        return "".join(f"{s}{n // g if n // g > 1 else ''}" for s, n in comp.items())

    def to_geometry(self):
        """A Geometry in Bohr (Geometry.from_ase_atoms in the reference converts
        Angstrom to Bohr the same way, slakonet/atoms.py:1335-1381)."""
        from .geometry import Geometry
        return Geometry(self.atomic_numbers, self.cart_coords, self.lattice, units="angstrom")


class Vasprun:
    """The parts of a vasprun.xml that the training-data path reads."""

    def __init__(self, path: str):
        self.path = path
        root = ET.parse(path).getroot()
        self.species = self._atoms(root)
        calcs = root.findall("calculation")
        structures = [c.find("structure") for c in calcs if c.find("structure") is not None]
        if not structures:
            structures = [s for s in root.findall("structure") if s.get("name") == "finalpos"]
        if not structures:
            structures = root.findall("structure")
        if not structures:
            raise ValueError(f"{path}: no <structure> block")
        self.all_structures = [self._structure(s) for s in structures]
        self._energies = []
        for c in calcs:
            en = [e for e in c.findall("energy")]
            if en:
                vals = {i.get("name"): float(i.text) for i in en[-1].findall("i")}
                self._energies.append(vals)
        self._eig_node = None
        self._dos_node = None
        for c in calcs:
            if c.find("eigenvalues") is not None:
                self._eig_node = c.find("eigenvalues")
            if c.find("dos") is not None:
                self._dos_node = c.find("dos")

    # ------------------------------------------------------------------
    @staticmethod
    def _atoms(root) -> List[str]:
        info = root.find("atominfo")
        if info is None:
            raise ValueError("vasprun: <atominfo> missing")
        for arr in info.findall("array"):
            if arr.get("name") == "atoms":
                return [rc.findall("c")[0].text.strip() for rc in arr.find("set").findall("rc")]
        raise ValueError("vasprun: atom list missing")

    def _structure(self, node) -> VaspStructure:
        crystal = node.find("crystal")
        lattice = _varray(crystal, "basis")
        frac = _varray(node, "positions")
        return VaspStructure(lattice, frac, self.species)

    @property
    def final_energy(self) -> float:
        if not self._energies:
            raise ValueError(f"{self.path}: no <energy> block")
        last = self._energies[-1]
        return float(last.get("e_fr_energy", last.get("e_0_energy")))

    @property
    def eigenvalues(self):
        """(energies [spin, k, band] eV, occupations [spin, k, band])."""
        if self._eig_node is None:
            raise ValueError(f"{self.path}: no <eigenvalues> block")
        top = self._eig_node.find("array").find("set")
        ens, occs = [], []
        for spin in top.findall("set"):
            e_s, o_s = [], []
            for kp in spin.findall("set"):
                rows = np.array([_floats(r.text) for r in kp.findall("r")], dtype=np.float64)
                e_s.append(rows[:, 0])
                o_s.append(rows[:, 1])
            ens.append(e_s)
            occs.append(o_s)
        return np.array(ens), np.array(occs)

    @property
    def efermi(self) -> float:
        if self._dos_node is None:
            raise ValueError(f"{self.path}: no <dos> block")
        for i in self._dos_node.findall("i"):
            if i.get("name") == "efermi":
                return float(i.text)
        raise ValueError(f"{self.path}: no efermi")

    @property
    def get_indir_gap(self):
        e, occ = self.eigenvalues
        half = 0.5 * float(occ.max()) if occ.size else 0.5
        occupied = occ > half
        if not occupied.any() or occupied.all():
            raise ValueError(f"{self.path}: no occupied/unoccupied split in the eigenvalues")
        vbm = float(e[occupied].max())
        cbm = float(e[~occupied].min())
        return max(cbm - vbm, 0.0), cbm, vbm

    @property
    def total_dos(self):
        if self._dos_node is None:
            raise ValueError(f"{self.path}: no <dos> block")
        total = self._dos_node.find("total")
        spins = total.find("array").find("set").findall("set")
        out = []
        for si, spin in enumerate(spins):
            rows = np.array([_floats(r.text) for r in spin.findall("r")], dtype=np.float64)
            if si == 0:
                out.append(rows[:, 0])
            out.append(rows[:, 1])
        return tuple(out)


class MultiVaspDataLoader:
    """Datasets from many vasprun.xml files (slakonet/optim.py:1446-1567).

    ``vasprun_paths``: a list of paths, one path, or a glob pattern. A file
    that fails to load is reported and skipped; no loadable file raises
    ValueError, as in the reference."""

    def __init__(self, vasprun_paths, log=print):
        if isinstance(vasprun_paths, str):
            self.vasprun_paths = sorted(glob.glob(vasprun_paths)) if "*" in vasprun_paths else [vasprun_paths]
        else:
            self.vasprun_paths = list(vasprun_paths)
        self.log = log
        self.datasets: List[dict] = []
        for i, p in enumerate(self.vasprun_paths):
            try:
                self.datasets.append(self._load_single_dataset(p, i))
            except Exception as e:  # the reference reports and skips a bad file (:1477-1484)
                if log:
                    log(f"failed to load {p}: {e}")
        if not self.datasets:
            raise ValueError("No valid datasets could be loaded!")
        if log:
            log(f"loaded {len(self.datasets)}/{len(self.vasprun_paths)} VASP datasets")

    def _load_single_dataset(self, vasprun_path: str, index: int) -> dict:
        vr = Vasprun(vasprun_path)
        st = vr.all_structures[-1]
        tdos = vr.total_dos
        comp = st.composition
        return {
            "index": index,
            "vasprun_path": vasprun_path,
            "geometry": st.to_geometry(),
            "target_energy": vr.final_energy,
            "target_bandgap": vr.get_indir_gap[0],
            "target_dos": torch.tensor(tdos[1]),
            "dos_energies": torch.tensor(tdos[0]),
            "elements": set(comp.keys()),
            "composition": comp,
            "metadata": {"natoms": st.num_atoms, "formula": st.reduced_formula, "volume": st.volume},
        }

    def get_all_elements(self):
        out = set()
        for d in self.datasets:
            out.update(d["elements"])
        return sorted(out)

    def get_batch(self, batch_size=None, shuffle=True):
        ds = self.datasets.copy()
        if shuffle:
            random.shuffle(ds)
        return ds if batch_size is None else ds[:batch_size]

    def __len__(self):
        return len(self.datasets)

    def __getitem__(self, idx):
        return self.datasets[idx]


# Face-centred cubic high-symmetry points in fractional coordinates of the
# primitive reciprocal lattice (Setyawan & Curtarolo, Comput. Mater. Sci. 49,
# 299 (2010), Table 2) and the path through them.
FCC_POINTS = {
    "G": (0.0, 0.0, 0.0), "X": (0.5, 0.0, 0.5), "W": (0.5, 0.25, 0.75), "K": (0.375, 0.375, 0.75),
    "L": (0.5, 0.5, 0.5), "U": (0.625, 0.25, 0.625),
}
FCC_PATH = [["G", "X", "W", "K", "G", "L", "U", "W", "L", "K"], ["U", "X"]]


def fcc_kpath(line_density: int = 20, cell=None) -> np.ndarray:
    """k-points along the FCC path, about ``line_density`` points per inverse
    Angstrom of path length (segment ends included once).

    The reference trains every dataset on the k-path of JVASP-1002 (silicon,
    FCC) from jarvis-tools (``get_klines_example``, slakonet/optim.py:103-136,
    :1655), fetched over the network. This restates that path from the
    published table; jarvis-tools' point spacing is not available here, so
    the exact point list is unpinned. ``cell`` (rows, Angstrom; default the
    silicon primitive cell, a = 5.43 A) sets the segment lengths."""
    if cell is None:
        a = 5.43
        cell = 0.5 * a * np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]])
    rec = np.linalg.inv(np.asarray(cell, dtype=np.float64)).T  # rows b_i / (2 pi)
    pts: List[np.ndarray] = []
    for branch in FCC_PATH:
        for a_lbl, b_lbl in zip(branch[:-1], branch[1:]):
            ka, kb = np.array(FCC_POINTS[a_lbl]), np.array(FCC_POINTS[b_lbl])
            length = np.linalg.norm((kb - ka) @ rec) * 2.0 * np.pi
            n = max(int(round(line_density * length)), 1)
            for i in range(n):
                pts.append(ka + (kb - ka) * i / n)
        pts.append(np.array(FCC_POINTS[branch[-1]]))
    return np.array(pts)


def get_klines_example(line_density: int = 20, default_points: int = 2):
    """k-line rows from the FCC path, paired as the reference pairs jarvis'
    k-point list (kpts_to_klines with default_points, slakonet/optim.py:135)."""
    from .geometry import kpts_to_klines
    return kpts_to_klines(fcc_kpath(line_density), default_points=default_points)


def train_multi_vasp_skf_parameters(multi_element_optimizer, vasprun_paths, num_epochs: int = 100,
                                    learning_rate: float = 0.00001, batch_size: Optional[int] = None,
                                    plot_frequency: int = 5, save_directory: str = "multi_vasp_optimization_all",
                                    weight_by_system_size: bool = True, early_stopping_patience: int = 20,
                                    klines=None, shell_dict=None, device="cuda", log=print):
    """Fit the tables to the band gaps of VASP runs (slakonet/optim.py:1570-1855).

    Every dataset is evaluated on ``klines`` (default: ``get_klines_example()``,
    the same k-lines for every structure, as the reference does). ``batch_size``
    picks that many datasets at random, once, for every epoch (the reference
    reshuffles per epoch; with the batched solve the build plans are made once
    per run). Returns (optimizer, loss_history, data_loader)."""
    from . import engine
    from .fit import FitDataset, train_skf_parameters
    from .geometry import klines_points
    from .units import generate_shell_dict_upto_Z65

    loader = MultiVaspDataLoader(vasprun_paths, log=log)
    missing = set(loader.get_all_elements()) - set(multi_element_optimizer.get_system_elements())
    if missing and log:
        log(f"missing SKF tables for elements: {sorted(missing)}")
    if shell_dict is None:
        shell_dict = generate_shell_dict_upto_Z65()
    kl = get_klines_example() if klines is None else klines
    kfrac, kw = klines_points(kl)
    chosen = loader.get_batch(batch_size=batch_size, shuffle=batch_size is not None)
    datasets = []
    for d in chosen:
        g = d["geometry"]
        inp = engine.StructureInput(g.atomic_numbers_np, g.positions_np, g.cell_np, kfrac, kw)
        datasets.append(FitDataset(inp, None, float(d["target_bandgap"]), d["metadata"]["natoms"],
                                   meta={"index": d["index"], "formula": d["metadata"]["formula"]}))
    model, history = train_skf_parameters(multi_element_optimizer, datasets, shell_dict, num_epochs=num_epochs,
                                          learning_rate=learning_rate, loss="reference",
                                          save_directory=save_directory,
                                          weight_by_system_size=weight_by_system_size,
                                          early_stopping_patience=early_stopping_patience, device=device,
                                          log=log, log_every=plot_frequency)
    return model, history, loader
