"""SimpleDftb-compatible calculator on the MI355X device path.

Mirrors slakonet/main.py:SimpleDftb for the hot path: construction
(:65-127), __call__ (:165-196; the per-k Python loop becomes one batched
device build + eigensolve), get_fermi_energy (:600-609), get_eigenvalues
(:611-635), calculate_band_gap (:637-707), calculate_dos (:709-797),
calculate_band_structure_properties (:1539-1642), get_properties_dict
(:1657-1819) and _calculate_electronic_energy (:1430-1448).

Differences from the reference, all deliberate:
* a k-point whose S(k) is not positive definite raises EigenFailure naming
  the k index instead of being printed and dropped (:182-184);
* fermi_method="bisection" (default) solves N(mu) = Ne exactly;
  fermi_method="reference" reproduces the half-step Newton of
  slakonet/fermi.py:176-196 including its stopping rule.
"""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from . import engine
from .geometry import Geometry, klines_points, monkhorst_pack
from .tables import TableSet
from .units import FERMI_KT_HA, H2E, HS_CUTOFF_BOHR

FERMI_MODES = {"bisection": 0, "reference": 1}


def _kpoints_from_args(kpoints, klines, kpoints_frac, k_weights):
    if kpoints_frac is not None:
        k = np.asarray(kpoints_frac.cpu() if isinstance(kpoints_frac, torch.Tensor) else kpoints_frac,
                       dtype=np.float64).reshape(-1, 3)
        w = (np.full(len(k), 1.0 / len(k)) if k_weights is None else
             np.asarray(k_weights.cpu() if isinstance(k_weights, torch.Tensor) else k_weights, dtype=np.float64))
        return k, w
    if klines is not None:
        return klines_points(klines)
    if kpoints is None:
        kpoints = [1, 1, 1]       # slakonet/atoms.py:391-392 default
    kp = kpoints.reshape(-1).tolist() if isinstance(kpoints, torch.Tensor) else list(np.ravel(kpoints))
    return monkhorst_pack([int(v) for v in kp])


class SimpleDftb:
    def __init__(self, geometry: Geometry, shell_dict=None, h_feed: Optional[TableSet] = None, s_feed=None,
                 nelectron=None, kpoints=None, klines=None, repulsive=False, device=None,
                 with_eigenvectors=False, tables: Optional[TableSet] = None, kpoints_frac=None,
                 k_weights=None, fermi_method: str = "bisection", table_tensor=None):
        self.tables = tables if tables is not None else h_feed
        if self.tables is None:
            raise ValueError("a TableSet (tables= or h_feed=) is required")
        self.device = self.tables.device if device is None else torch.device(device)
        self.geometry = geometry
        self.shell_dict = shell_dict
        self.with_eigenvectors = with_eigenvectors
        self.fermi_method = fermi_method
        if fermi_method not in FERMI_MODES:
            raise ValueError(f"fermi_method must be one of {list(FERMI_MODES)}")
        k, w = _kpoints_from_args(kpoints, klines, kpoints_frac, k_weights)
        self.kpoints_np = k
        self.k_weights = torch.from_numpy(w).to(self.device).unsqueeze(0)
        self.kpoints = torch.from_numpy(k).unsqueeze(0)
        self.max_nk = torch.tensor(len(k))
        z = geometry.atomic_numbers_np
        if nelectron is None:
            nelectron = self.tables.n_electrons(z)
        self.nelectron = torch.as_tensor(nelectron, dtype=torch.float64).reshape(-1)[:1].to(self.device)
        self._table_tensor = table_tensor
        self._inp = engine.StructureInput(z, geometry.positions_np, geometry.cell_np, k, w,
                                          float(self.nelectron.item()))
        self._edges = None
        self._band_gap = None
        self._density = None
        self._occ = None
        self._V = None
        self.eigenvalue = None

    # ------------------------------------------------------------------ solve
    def compute_hs_matrices(self):
        self._plan = engine.plan_build(self.tables, [self._inp])
        self.ham, self.overlap = engine.run_build(self.tables, self._plan, table_tensor=self._table_tensor)
        return self.ham, self.overlap

    def __call__(self):
        self.compute_hs_matrices()
        w, V, _ = engine.eigh(self.ham, self.overlap, vectors=self.with_eigenvectors,
                              real=engine.real_kpoints(self._inp.kpoints))
        self._w = w
        self._V = V
        self.eigenvalue = w.unsqueeze(0)                       # [1, nk, nb]
        if self.with_eigenvectors:
            self.eigenvectors = V.permute(1, 2, 0).unsqueeze(0)   # [1, n, nb, nk]
        self._edges = None
        self._band_gap = None
        self._density = None
        self._occ = None
        return self.eigenvalue

    # ------------------------------------------------------------ density
    @property
    def _occupations(self):
        """Aufbau occupations [1, nk, nb] (slakonet/slaterkoster.py:25-74, as
        called by SimpleDftb.solve_kpoint, slakonet/main.py:157): 2 for the
        floor(Ne/2) lowest states of every k-point, 1 for the next one when
        Ne is odd, 0 above (no smearing)."""
        if self._occ is None:
            ne = float(self.nelectron.item())
            if ne < 1:
                raise ValueError("occupations need at least one electron")
            nb = int(self.eigenvalue.shape[-1])
            nk = int(self.eigenvalue.shape[1])
            n_pairs = int(ne // 2)
            occ = torch.zeros(nb, dtype=torch.float64)
            occ[:min(n_pairs, nb)] = 2.0
            if ne % 2 > 0 and n_pairs < nb:
                occ[n_pairs] = 1.0
            self._occ = occ.to(self.device).expand(1, nk, nb).contiguous()
        return self._occ

    @property
    def density(self):
        """Density matrices [1, n, n, nk] (slakonet/main.py:158-161):
        rho_k = conj(C sqrt(f)) (C sqrt(f))^T, i.e. rho[mu][nu] =
        sum_j f_j conj(C[mu][j]) C[nu][j], with C the S-orthonormal
        eigenvectors of k in columns. Formed on the device on first access
        (column scaling, then one batched complex GEMM over all k); when the
        solve ran without eigenvectors, they are computed first."""
        if self._density is None:
            import ctypes as C
            from . import _lib
            if self.eigenvalue is None:
                self()
            V = self._V
            if V is None:
                _, V, _ = engine.eigh(self.ham, self.overlap, vectors=True, real=engine.real_kpoints(self._inp.kpoints))
            batch, ld, _ = V.shape
            n = self._plan.n_orb[0]
            occ = self._occupations[0]                                   # [nk, nb]
            g = torch.zeros(batch, ld, dtype=torch.float64, device=V.device)
            g[:, :occ.shape[-1]] = torch.sqrt(occ)
            lib = _lib.lib()
            sh = _lib.stream_handle()
            Y = torch.empty_like(V)
            R = torch.empty_like(V)
            st = ld * ld
            _lib.check(lib.slk_scale_columns(_lib.ptr(V), _lib.ptr(g), ld, ld, batch, ld, _lib.ptr(Y), sh),
                       "slk_scale_columns")
            # R = conj(Y) Y^T: op bit1 conjugates A, bit0 transposes B
            _lib.check(lib.slk_zgemm(batch, ld, ld, ld, 1.0, _lib.ptr(Y), ld, st, 2, _lib.ptr(Y), ld, st, 1,
                                     0.0, _lib.ptr(R), ld, st, sh), "slk_zgemm")
            self._density = R[:, :n, :n].permute(1, 2, 0).unsqueeze(0)
        return self._density

    # ------------------------------------------------------------------ fermi
    def _run_post(self, dos_grid=None, sigma=0.1):
        edges, dos, stats = engine.post(self._plan, self._w, FERMI_MODES[self.fermi_method], FERMI_KT_HA,
                                        dos_grid, sigma)
        return edges, dos, stats

    def _edge_values(self):
        if self._edges is None:
            edges, _, _ = self._run_post()
            self._edges_t = edges
            self._edges = engine.edges_dict(edges[0].cpu().numpy())
        return self._edges

    def get_fermi_energy(self, kT=0.025):
        """Fermi level [1, 1] in Hartree (kT of the search is 0.01 Ha as in the reference)."""
        self._edge_values()
        return self._edges_t[:, 0:1].clone()

    def get_eigenvalues(self, fermi_shift=True, unit="eV"):
        e = self.eigenvalue.clone()
        if fermi_shift:
            e = e - self.get_fermi_energy()
        if unit == "eV":
            e = e * H2E
        return e

    def calculate_band_gap(self, kT=0.025):
        if self._band_gap is None:
            ed = self._edge_values()
            mu = ed["mu"]
            fev = torch.tensor(mu * H2E, dtype=torch.float64)
            if ed["metal"] > 0:
                self._band_gap = {"gap": torch.tensor(0.0), "vbm": fev, "cbm": fev, "direct": False,
                                  "vbm_kpoint": 0, "cbm_kpoint": 0}
            elif ed["vbm"] < -1e299 or ed["cbm"] > 1e299:
                raise RuntimeError("no occupied or no unoccupied states at the Fermi level")
            else:
                vbm = torch.tensor(ed["vbm"], dtype=torch.float64)
                cbm = torch.tensor(ed["cbm"], dtype=torch.float64)
                self._band_gap = {"gap": cbm - vbm, "vbm": vbm, "cbm": cbm,
                                  "direct": int(ed["vbm_k"]) == int(ed["cbm_k"]),
                                  "vbm_kpoint": int(ed["vbm_k"]), "cbm_kpoint": int(ed["cbm_k"])}
        return self._band_gap

    def calculate_band_structure_properties(self):
        ed = self._edge_values()
        has_v, has_c = ed["n_val"] > 0, ed["n_con"] > 0
        vbm = ed["s_vbm"] if has_v else float("-inf")
        cbm = ed["s_cbm"] if has_c else float("inf")
        gap = cbm - vbm if (has_v and has_c) else 0.0
        mdg = ed["min_direct_gap"] if ed["min_direct_gap"] < 1e299 else float("inf")
        return {
            "band_gap": gap,
            "vbm": vbm if has_v else None,
            "cbm": cbm if has_c else None,
            "valence_band_width": ed["vb_width"] if has_v else 0.0,
            "conduction_band_width": ed["cb_width"] if has_c else 0.0,
            "minimum_direct_gap": mdg if np.isfinite(mdg) else None,
            "is_direct_semiconductor": (torch.tensor(abs(gap - mdg) < 0.01) if np.isfinite(mdg) else False),
            "nkpoints": int(self.eigenvalue.shape[1]),
            "nbands": int(self.eigenvalue.shape[2]),
            "fermi_energy_original": ed["mu"] * H2E,
        }

    # ------------------------------------------------------------------ DOS
    def calculate_dos(self, energy_range=(-10, 5), num_points=5000, sigma=0.1, fermi_shift=True, unit="eV"):
        grid = torch.linspace(energy_range[0], energy_range[1], num_points, dtype=torch.float64,
                              device=self.device)
        if unit != "eV":
            raise NotImplementedError("DOS is evaluated in eV as in the properties dict")
        plan = self._plan
        edges, dos, stats = engine.post(plan, self._w, FERMI_MODES[self.fermi_method], FERMI_KT_HA, grid, sigma)
        if not fermi_shift:
            # shift-free DOS: rerun with mu forced to 0 in the edges row
            edges[:, 0] = 0.0
            from . import _lib
            import ctypes as C
            lib = _lib.lib()
            jobs = np.zeros(1, dtype=engine.DOS_DT)
            jobs[0] = (0, plan.nk[0], plan.n_orb[0], self._w.shape[1], 0, 0, 0)
            jd = engine._upload(jobs, self.device)
            _lib.check(lib.slk_dos_launch(C.c_void_p(self._w.data_ptr()), C.c_void_p(plan.k_weights.data_ptr()),
                                          C.c_void_p(jd.data_ptr()), 1, C.c_void_p(grid.data_ptr()), num_points,
                                          float(sigma), C.c_void_p(edges.data_ptr()), C.c_void_p(dos.data_ptr()),
                                          None, _lib.stream_handle()), "slk_dos_launch")
        # the statistics (DOS at E_F, gap centre, total) belong to the
        # Fermi-shifted DOS only
        self._dos_stats = stats if fermi_shift else None
        return grid, dos[0]

    def calculate_atom_projected_dos(self, sigma=0.1, energy_range=(-8, 6), num_points=1000):
        """Atom-type projected DOS on the device (compute_atom_projected_dos,
        slakonet/predict_slakonet.py:119-210): per state, the weight of an atom
        type is the sum of |psi_mu|^2 over its atoms' orbitals; Gaussians of
        width ``sigma`` eV on a grid relative to the Fermi level; averaged over
        k-points (1/nk). Orbitals are attributed to atoms by their actual
        counts (the reference divides n_orb evenly, equal for uniform shells).
        Returns (energy grid [num_points] numpy, {symbol: pdos numpy}, symbols)."""
        import ctypes as C
        from . import _lib
        from .units import CHEMICAL_SYMBOLS
        if getattr(self, "_V", None) is None:
            raise RuntimeError("projected DOS needs eigenvectors: construct with with_eigenvectors=True")
        z = np.asarray(self.geometry.atomic_numbers_np)
        symbols = [CHEMICAL_SYMBOLS[int(v)] for v in z]
        unique = list(dict.fromkeys(symbols))
        sp = self.tables.species_index(z)
        norb_atoms = self.tables.norb[sp]
        orb_type = np.repeat(np.array([unique.index(sym) for sym in symbols], dtype=np.int32), norb_atoms)
        n = int(norb_atoms.sum())
        nk = self._w.shape[0]
        ed = self._edge_values()
        dev = self.device
        grid = torch.linspace(energy_range[0], energy_range[1], num_points, dtype=torch.float64, device=dev)
        ot = torch.from_numpy(orb_type).to(dev)
        wts = torch.empty(nk * n * len(unique), dtype=torch.float64, device=dev)
        out = torch.empty(len(unique), num_points, dtype=torch.float64, device=dev)
        lib = _lib.lib()
        _lib.check(lib.slk_pdos_launch(_lib.ptr(self._V), _lib.ptr(self._w), n, self._V.shape[1], nk,
                                       self._w.shape[1], _lib.ptr(ot), len(unique), _lib.ptr(grid), num_points,
                                       float(sigma), ed["mu"] * H2E, 1.0 / nk, _lib.ptr(wts), _lib.ptr(out),
                                       _lib.stream_handle()), "slk_pdos_launch")
        o = out.cpu().numpy()
        return grid.cpu().numpy(), {sym: o[i] for i, sym in enumerate(unique)}, unique

    def get_forces(self):
        """Forces [1, N, 3] (Hartree/Bohr) = -dE/dR of the band energy of
        _calculate_electronic_energy (Fermi-Dirac occupations at kT = 0.025 eV,
        Fermi level held fixed as in the reference), from the hand-written
        position adjoint of the build (slakonet/main.py:400-585 differentiates
        the same energy with autograd)."""
        from . import adjoint
        from .units import H2E
        if getattr(self, "_V", None) is None:
            raise RuntimeError("forces need eigenvectors: construct with with_eigenvectors=True")
        mu = self._edge_values()["mu"]
        n = self._plan.n_orb[0]
        g = adjoint.band_energy_weights(self._w, self._plan.k_weights, mu, 0.025 / H2E, n)
        QH, QS = adjoint.weighted_densities(self._V, self._w, g)
        dR = adjoint.position_gradient_from_densities(self.tables, self._plan, QH, QS)
        self._forces = (-dR).unsqueeze(0)
        return self._forces

    def _calculate_electronic_energy(self):
        if self.eigenvalue is None:
            self()                          # slakonet/main.py:1433-1435
        mu = self.get_fermi_energy()
        kT = 0.025 / H2E
        occ = 1.0 / (1.0 + torch.exp((self.eigenvalue - mu.view(1, 1, 1)) / kT))
        occ = torch.where(self.eigenvalue == 0, torch.zeros_like(occ), occ)
        return torch.sum(occ * self.eigenvalue * self.k_weights.unsqueeze(-1)).real

    # ------------------------------------------------------------------ dict
    def get_properties_dict(self, kT=0.025, include_bulk_modulus=False, include_dos_data=False):
        if include_bulk_modulus:
            raise NotImplementedError("bulk modulus is outside the electronic-structure hot path")
        try:
            fermi = self.get_fermi_energy(kT)
            bg = self.calculate_band_gap(kT)
        except RuntimeError:
            fermi = torch.tensor(0)
            bg = {"vbm": torch.tensor(0), "cbm": torch.tensor(0), "gap": torch.tensor(0), "direct": True,
                  "vbm_kpoint": 0, "cbm_kpoint": 0}
        bs = self.calculate_band_structure_properties()
        props = {
            "fermi_energy_eV": float((fermi * H2E).reshape(-1)[0].item()),
            "fermi_energy_Ha": float(fermi.reshape(-1)[0].item()),
            "band_gap_eV": bg["gap"],
            "vbm_eV": float(torch.as_tensor(bg["vbm"]).reshape(-1)[0].item()),
            "cbm_eV": float(torch.as_tensor(bg["cbm"]).reshape(-1)[0].item()),
            "is_direct_gap": bg["direct"],
            "vbm_kpoint": bg["vbm_kpoint"],
            "cbm_kpoint": bg["cbm_kpoint"],
            "valence_band_width_eV": bs["valence_band_width"],
            "conduction_band_width_eV": bs["conduction_band_width"],
            "minimum_direct_gap_eV": bs["minimum_direct_gap"],
            "is_direct_semiconductor": bs["is_direct_semiconductor"],
            "nkpoints": int(self.max_nk.item()),
            "nbands": int(self.eigenvalue.shape[-1]),
            "nelectrons": int(self.nelectron.item()),
        }
        if include_dos_data:
            grid, dos = self.calculate_dos(energy_range=(-10, 10), num_points=5000, sigma=0.1, fermi_shift=True)
            st = self._dos_stats[0]
            sth = st.cpu().numpy()
            props.update({
                "dos_at_fermi": float(sth[0]),
                "dos_gap_center_eV": float(sth[1]),
                "dos_total_states": float(sth[2]),
                "dos_at_fermi_tensor": st[0],
                "dos_gap_center_tensor": st[1],
                "dos_total_states_tensor": st[2],
                "dos_energy_grid_tensor": grid,
                "dos_values_tensor": dos,
            })
        return props


def compute_atom_projected_dos(properties, geometry=None, sigma=0.1, energy_range=(-8, 6), filename=None):
    """Drop-in for slakonet/predict_slakonet.py:119 (plotting excluded): the
    properties dict must hold "calc" (get_fermi=True) solved with eigenvectors."""
    return properties["calc"].calculate_atom_projected_dos(sigma=sigma, energy_range=energy_range)
