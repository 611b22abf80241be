"""Batched device pipeline: plan -> build H(k)/S(k) -> eigensolve -> Fermi/edges/DOS.

A batch holds any number of structures; every k-point of every structure is
one matrix of the batch. Host work is O(atoms + images + work items) numpy
bookkeeping packed into a few device uploads; all O(N^2) and O(n^3) work is in
the HIP kernels of libslakonet_amd.so.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .geometry import lattice_translations
from .tables import TableSet
from .units import FERMI_KT_HA, FERMI_MAX_ITER, FERMI_TOL, H2E, HS_CUTOFF_BOHR, METAL_THRESHOLD_EV

KC = 4          # k-points per build workgroup (slk_build.hip)
MAXCOL = 64     # columns per build workgroup
PAD_H = 4.0     # Hartree: H diagonal of identity-padded rows (above every band)

STRUCT_DT = np.dtype([("atom0", "<i4"), ("n_atoms", "<i4"), ("n_orb", "<i4"), ("ld", "<i4"), ("mat0", "<i8"),
                      ("k0", "<i4"), ("nk", "<i4"), ("t0", "<i4"), ("nt", "<i4")])
WORK_DT = np.dtype([(f, "<i4") for f in ("s", "i", "j0", "j1", "c0", "ncol", "kc0", "nkc")])
SPEC_DT = np.dtype([("w0", "<i8"), ("nk", "<i4"), ("nb", "<i4"), ("wld", "<i4"), ("kw0", "<i4"), ("nelec", "<f8")])
DOS_DT = np.dtype({"names": ["w0", "nk", "nb", "wld", "kw0", "edge", "out0"],
                   "formats": ["<i8", "<i4", "<i4", "<i4", "<i4", "<i4", "<i8"],
                   "offsets": [0, 8, 12, 16, 20, 24, 32], "itemsize": 40})

EDGE_FIELDS = ["mu", "metal", "vbm", "cbm", "vbm_k", "cbm_k", "s_vbm", "vb_width", "s_cbm", "cb_width",
               "min_direct_gap", "vb2", "cb2", "n_iter", "n_val", "n_con"]


@dataclass
class StructureInput:
    z: np.ndarray            # [N] atomic numbers
    positions: np.ndarray    # [N, 3] Bohr
    cell: np.ndarray         # [3, 3] Bohr, rows are lattice vectors
    kpoints: np.ndarray      # [nk, 3] fractional
    k_weights: np.ndarray    # [nk]
    nelec: Optional[float] = None


@dataclass
class BuildPlan:
    n_struct: int
    n_mat: int
    ld: int
    n_orb: List[int]
    nk: List[int]
    mat0: List[int]
    k_weights: torch.Tensor          # [n_mat] device float64
    nelec: List[float]
    dev: dict = field(default_factory=dict)
    n_work: int = 0
    real: bool = False               # every matrix real symmetric (set by callers that solve it so)
    n_row_work: int = 0              # row-panel build items after the n_work column tiles


def _upload(arr: np.ndarray, device) -> torch.Tensor:
    """Host -> device through pinned staging so the copy is stream-ordered and
    never blocks the host behind earlier work on the stream."""
    a = np.ascontiguousarray(arr)
    return _pinned(a.view(np.uint8).reshape(-1)).to(device, non_blocking=True)


def _pinned(a: np.ndarray) -> torch.Tensor:
    t = torch.from_numpy(np.ascontiguousarray(a))
    return t.pin_memory() if torch.cuda.is_available() else t


def column_tiles(norb_atoms: np.ndarray):
    """Runs of consecutive atoms with <= 64 orbitals: (j0, j1, c0, ncol)."""
    tiles = []
    j0, c0, acc = 0, 0, 0
    off = 0
    for j, no in enumerate(norb_atoms):
        if acc + no > MAXCOL and acc > 0:
            tiles.append((j0, j, c0, acc))
            j0, c0, acc = j, off, 0
        acc += int(no)
        off += int(no)
    if acc > 0:
        tiles.append((j0, len(norb_atoms), c0, acc))
    return np.array(tiles, dtype=np.int32).reshape(-1, 4)


def plan_build(tables: TableSet, inputs: Sequence[StructureInput], ld: Optional[int] = None,
               cutoff: float = HS_CUTOFF_BOHR) -> BuildPlan:
    """Plan one batch: the native planner (slk_plan_size / slk_plan_fill)
    computes translations, column tiles and work items and packs everything
    into one pinned blob, uploaded with a single stream-ordered copy."""
    device = tables.device
    zs = [np.asarray(inp.z, dtype=np.int64).reshape(-1) for inp in inputs]
    for z in zs:
        tables.check_pairs(z)
    spec = tables.species_index(np.concatenate(zs)) if zs else np.zeros(0, np.int32)
    n_atoms = np.array([len(z) for z in zs], dtype=np.int32)
    pos = np.ascontiguousarray(np.concatenate([np.asarray(i.positions, np.float64).reshape(-1, 3) for i in inputs]))
    cell = np.ascontiguousarray(np.stack([np.asarray(i.cell, np.float64).reshape(3, 3) for i in inputs]))
    kps = [np.asarray(i.kpoints, np.float64).reshape(-1, 3) for i in inputs]
    nk = np.array([len(k) for k in kps], dtype=np.int32)
    kfrac = np.ascontiguousarray(np.concatenate(kps))
    kw = np.ascontiguousarray(np.concatenate([np.asarray(i.k_weights, np.float64).reshape(-1) for i in inputs]))
    if len(kw) != len(kfrac):
        raise ValueError("k_weights and kpoints differ in length")
    spec = np.ascontiguousarray(spec, dtype=np.int32)
    norb = np.ascontiguousarray(tables.norb, dtype=np.int32)
    host = _lib.host_lib()
    pin = _lib.SlkPlanIn(len(inputs), n_atoms.ctypes.data, spec.ctypes.data, pos.ctypes.data, cell.ctypes.data,
                         nk.ctypes.data, kfrac.ctypes.data, kw.ctypes.data, norb.ctypes.data, len(norb),
                         float(cutoff), int(ld or 0))
    out = _lib.SlkPlanOut()
    rc = host.slk_plan_size(C.byref(pin), C.byref(out))
    if rc == 2:
        raise ValueError("ld smaller than the largest structure")
    if rc != 0:
        raise ValueError("invalid structure in the batch (no atoms, no k-points or a singular cell)")
    stage = _pinned_bytes(int(out.total))
    n_orb = np.zeros(len(inputs), dtype=np.int32)
    mat0 = np.zeros(len(inputs), dtype=np.int64)
    rc = host.slk_plan_fill(C.byref(pin), C.byref(out), C.c_void_p(stage.data_ptr()), n_orb.ctypes.data,
                            mat0.ctypes.data)
    if rc != 0:
        raise RuntimeError("slk_plan_fill failed")
    blob = stage.to(device, non_blocking=True)
    names = ["structs", "work", "atom_pos", "atom_spec", "atom_orb", "trans_cart", "trans_int", "kfrac", "kw"]
    views = {k: blob[out.off[q]:out.off[q] + out.bytes[q]] for q, k in enumerate(names)}
    dev = {k: v for k, v in views.items() if k != "kw"}
    dev["blob"] = blob
    dev["stage"] = stage
    kwd = views["kw"].view(torch.float64)
    nelec = [inp.nelec if inp.nelec is not None else tables.n_electrons(z) for inp, z in zip(inputs, zs)]
    return BuildPlan(len(inputs), int(out.n_mat), int(out.ld), n_orb.tolist(), nk.tolist(), mat0.tolist(), kwd,
                     nelec, dev, int(out.n_work), n_row_work=int(out.n_row_work))


def _pinned_bytes(n: int) -> torch.Tensor:
    if torch.cuda.is_available():
        return torch.empty(max(n, 1), dtype=torch.uint8, pin_memory=True)
    return torch.empty(max(n, 1), dtype=torch.uint8)


# set while the library's live kernel sampler is enabled (slk_sampler_reset)
SAMPLING = False


def build_args(tables: TableSet, plan: BuildPlan, H_ptr: int, S_ptr: int, out_f32: bool = False,
               cutoff: float = HS_CUTOFF_BOHR, table_tensor: Optional[torch.Tensor] = None):
    """SlkBuildArgs of a plan (shared by the build and its adjoint)."""
    d = plan.dev
    return _lib.SlkBuildArgs(tables.c_struct(table_tensor), d["structs"].data_ptr(), d["work"].data_ptr(),
                             plan.n_work, d["atom_pos"].data_ptr(), d["atom_spec"].data_ptr(),
                             d["atom_orb"].data_ptr(), d["trans_cart"].data_ptr(), d["trans_int"].data_ptr(),
                             d["kfrac"].data_ptr(), float(cutoff), H_ptr, S_ptr, int(out_f32), PAD_H,
                             plan.n_struct, plan.n_mat, plan.ld, plan.n_row_work)


def run_build(tables: TableSet, plan: BuildPlan, out_f32: bool = False, cutoff: float = HS_CUTOFF_BOHR,
              table_tensor: Optional[torch.Tensor] = None, stream=None, out=None):
    """H, S [n_mat, ld, ld] complex128 (or complex64). ``out`` = (H, S): write
    into these buffers (every element is written: the build or the padding)."""
    lib = _lib.lib()
    dt = torch.complex64 if out_f32 else torch.complex128
    if out is not None:
        H, S = out
        shape = (plan.n_mat, plan.ld, plan.ld)
        dev = torch.device(tables.device)
        if dev.index is None:
            dev = torch.device(dev.type, torch.cuda.current_device())
        for t in (H, S):
            if tuple(t.shape) != shape or t.dtype != dt or not t.is_contiguous() or t.device != dev:
                raise ValueError(f"out buffers must be contiguous {dt} {shape} on {dev}")
    else:
        H = torch.empty(plan.n_mat, plan.ld, plan.ld, dtype=dt, device=tables.device)
        S = torch.empty_like(H)
    d = plan.dev
    args = build_args(tables, plan, H.data_ptr(), S.data_ptr(), out_f32, cutoff, table_tensor)
    sh = _lib.stream_handle(stream)
    if SAMPLING:
        # algorithmic bytes of this launch for the live sampler: the H and S
        # elements the build writes (padding excluded)
        lib.slk_sampler_hint(1, float(sum(nk * 2 * n * n for nk, n in zip(plan.nk, plan.n_orb)) * H.element_size()),
                             0.0)
    # the whole-matrix form writes every element, identity padding included
    _lib.check(lib.slk_build_hs_launch(C.byref(args), sh), "slk_build_hs_launch")
    return H, S


class EigenFailure(RuntimeError):
    pass


EIGH_STAGES = ["potrf", "trinv", "reduce", "hetrd", "tridiag_solve", "back_transform"]


PRECISIONS = ("fp64", "mixed")


EIGH_MIXED, EIGH_REAL = 1, 2   # slk_eigh_ex flags


def real_kpoints(kfrac) -> bool:
    """True when every Bloch phase exp(2 pi i k.T) over integer T is +-1, i.e.
    2k is integer for every k-point (Gamma and the other time-reversal
    invariant points, such as the 1x1x1 Monkhorst-Pack point (1/2, 1/2, 1/2)):
    H(k) and S(k) are then real symmetric."""
    k2 = 2.0 * np.asarray(kfrac, dtype=np.float64)
    return bool(np.all(np.abs(k2 - np.round(k2)) < 1e-12))


def eigh(H: torch.Tensor, S: torch.Tensor, vectors: bool = False, stream=None, check: bool = True,
         stage_ms: Optional[list] = None, precision: str = "fp64", workspace: Optional[list] = None,
         real: bool = False):
    """Batched generalized eigensolve. Returns w [batch, ld], V or None, info.

    precision "fp64": every stage in FP64. "mixed": FP64 Cholesky and L^-1,
    FP32 reduction and tridiagonalisation, FP64 bisection (values only).
    real: every H and S is real symmetric (zero imaginary parts; k-points whose
    Bloch phases are all +-1): the solve skips the imaginary arithmetic.
    With ``stage_ms`` (a list) the per-stage milliseconds measured with HIP
    events on the launch stream are appended (one synchronisation). With
    ``workspace`` (a list) the solver workspace is appended (the band-edge
    refinement of the mixed path reads it)."""
    if precision not in PRECISIONS:
        raise ValueError(f"precision must be one of {PRECISIONS}")
    mixed = precision == "mixed"
    if mixed and vectors:
        raise ValueError("the mixed-precision solve returns eigenvalues only")
    lib = _lib.lib()
    check_pair(H, S)
    batch, ld, _ = H.shape
    n = ld
    ws_bytes = int(lib.slk_eigh_mixed_workspace_bytes(n, ld, batch) if mixed else
                   lib.slk_eigh_workspace_bytes(n, ld, batch, int(vectors)))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=H.device)
    if workspace is not None:
        workspace.append(ws)
    w = torch.empty(batch, ld, dtype=torch.float64, device=H.device)
    info = torch.zeros(2 * batch, dtype=torch.int32, device=H.device)
    V = torch.empty_like(H) if vectors else None
    flags = (EIGH_MIXED if mixed else 0) | (EIGH_REAL if real else 0)
    ms = (C.c_float * 6)() if stage_ms is not None else None
    rc = lib.slk_eigh_ex(C.c_void_p(H.data_ptr()), C.c_void_p(S.data_ptr()), n, ld, batch, C.c_void_p(w.data_ptr()),
                         ld, _lib.ptr(V) if V is not None else None, C.c_void_p(ws.data_ptr()), ws_bytes,
                         C.c_void_p(info.data_ptr()), flags, ms, _lib.stream_handle(stream))
    if stage_ms is not None:
        stage_ms.append(dict(zip(EIGH_STAGES, list(ms))))
    if rc == 3:
        raise NotImplementedError(f"matrix dimension {n} is beyond the solver's limit (n <= 16384)")
    _lib.check(rc, "slk_eigh_f64")
    if check:
        check_info(info, batch)
    return w, V, info


def check_pair(H: torch.Tensor, S: torch.Tensor):
    """The kernels read H and S as contiguous [batch, ld, ld] complex128 on
    the GPU; anything else is refused before a launch."""
    for name, t in (("H", H), ("S", S)):
        if t.dtype != torch.complex128:
            raise TypeError(f"{name} must be complex128, got {t.dtype}")
        if t.dim() != 3 or t.shape[1] != t.shape[2] or not t.is_contiguous() or not t.is_cuda:
            raise ValueError(f"{name} must be a contiguous [batch, ld, ld] GPU tensor, got {tuple(t.shape)}")
    if H.shape != S.shape or H.device != S.device:
        raise ValueError("H and S differ in shape or device")


def check_info(info: torch.Tensor, batch: int, k_offset: int = 0):
    inf = info.cpu().numpy()
    bad = np.nonzero(inf[:batch])[0]
    if len(bad):
        k = int(bad[0])
        raise EigenFailure(f"overlap matrix S(k) is not positive definite at k-point index {k + k_offset} "
                           f"(Cholesky pivot {int(inf[k])})")
    bad = np.nonzero(inf[batch:2 * batch])[0]
    if len(bad):
        raise EigenFailure(tridiag_failure(int(inf[batch + bad[0]]), int(bad[0]) + k_offset))


def tridiag_failure(code: int, k: int) -> str:
    """Message for a non-zero info[batch + b] of the eigensolver."""
    if code == 2:
        return f"bulge chase of the tridiagonalisation gave up a bounded wait at k-point index {k}"
    return f"tridiagonal stage failed (code {code}) at k-point index {k}"


def spectra(plan: BuildPlan, w_ld: int) -> np.ndarray:
    sp = np.zeros(plan.n_struct, dtype=SPEC_DT)
    for s in range(plan.n_struct):
        sp[s] = (plan.mat0[s] * w_ld, plan.nk[s], plan.n_orb[s], w_ld, plan.mat0[s], plan.nelec[s])
    return sp


def dos_jobs(plan: BuildPlan, w_ld: int, ne: int) -> np.ndarray:
    """One DOS job per structure: its eigenvalue rows, Fermi shift from its edges row."""
    jobs = np.zeros(plan.n_struct, dtype=DOS_DT)
    for s in range(plan.n_struct):
        jobs[s] = (plan.mat0[s] * w_ld, plan.nk[s], plan.n_orb[s], w_ld, plan.mat0[s], s, s * ne)
    return jobs


def refine_edges(H: torch.Tensor, S: torch.Tensor, ws: torch.Tensor, w: torch.Tensor, plan: BuildPlan,
                 spd: torch.Tensor, edges: torch.Tensor, stream=None):
    """FP64 Rayleigh-quotient refinement of every structure's VBM and CBM after
    a mixed-precision solve (see slk_refine_edges_mixed); updates w and edges."""
    lib = _lib.lib()
    batch, ld, _ = H.shape
    nbytes = int(lib.slk_refine_scratch_bytes(ld, plan.n_struct))
    scratch = torch.empty(nbytes // 8 + 1, dtype=torch.float64, device=H.device)
    rc = lib.slk_refine_edges_mixed(C.c_void_p(H.data_ptr()), C.c_void_p(S.data_ptr()), ld, ld, batch,
                                    C.c_void_p(ws.data_ptr()), ws.numel(), C.c_void_p(w.data_ptr()), w.shape[1],
                                    C.c_void_p(spd.data_ptr()), plan.n_struct, C.c_void_p(edges.data_ptr()),
                                    C.c_void_p(scratch.data_ptr()), scratch.numel() * 8, _lib.stream_handle(stream))
    if rc == 3:
        raise NotImplementedError(f"band-edge refinement supports n <= 16384 (got {ld})")
    _lib.check(rc, "slk_refine_edges_mixed")
    return scratch


def post(plan: BuildPlan, w: torch.Tensor, fermi_mode: int = 0, kT: float = FERMI_KT_HA,
         dos_grid: Optional[torch.Tensor] = None, sigma: float = 0.1, stream=None, refine=None):
    """Fermi level, band edges and (optionally) DOS for every structure of the plan.

    ``refine`` = (H, S, workspace) of a mixed-precision solve: the VBM and CBM
    of every structure are refined in FP64 before the DOS is formed.
    Returns edges [n_struct, F] (device), dos [n_struct, ne] or None, stats [n_struct, 3] or None.
    """
    lib = _lib.lib()
    dev = w.device
    nf = int(lib.slk_edge_fields())
    sp = spectra(plan, w.shape[1])
    spd = _upload(sp, dev)
    edges = torch.zeros(plan.n_struct, nf, dtype=torch.float64, device=dev)
    sh = _lib.stream_handle(stream)
    kw = plan.k_weights
    _lib.check(lib.slk_fermi_launch(C.c_void_p(w.data_ptr()), C.c_void_p(kw.data_ptr()), C.c_void_p(spd.data_ptr()),
                                    plan.n_struct, float(kT), int(fermi_mode), FERMI_TOL, FERMI_MAX_ITER,
                                    C.c_void_p(edges.data_ptr()), sh), "slk_fermi_launch")
    _lib.check(lib.slk_edges_launch(C.c_void_p(w.data_ptr()), C.c_void_p(spd.data_ptr()), plan.n_struct,
                                    METAL_THRESHOLD_EV, C.c_void_p(edges.data_ptr()), sh), "slk_edges_launch")
    if refine is not None:
        H, S, ws = refine
        refine_edges(H, S, ws, w, plan, spd, edges, stream)
    dos = stats = None
    if dos_grid is not None:
        ne = int(dos_grid.numel())
        jd = _upload(dos_jobs(plan, w.shape[1], ne), dev)
        dos = torch.empty(plan.n_struct, ne, dtype=torch.float64, device=dev)
        if SAMPLING:
            # algorithmic bytes (eigenvalues, k-weights and grid read, DOS
            # written) and Gaussians of the full sum, for the live sampler
            states = sum(nk * nb for nk, nb in zip(plan.nk, plan.n_orb))
            nbytes = 8.0 * sum(nk * nb + nk + 2 * ne for nk, nb in zip(plan.nk, plan.n_orb))
            lib.slk_sampler_hint(5, nbytes, float(states) * ne)
        stats = torch.empty(plan.n_struct, 3, dtype=torch.float64, device=dev)
        _lib.check(lib.slk_dos_launch(C.c_void_p(w.data_ptr()), C.c_void_p(kw.data_ptr()), C.c_void_p(jd.data_ptr()),
                                      plan.n_struct, C.c_void_p(dos_grid.data_ptr()), ne, float(sigma),
                                      C.c_void_p(edges.data_ptr()), C.c_void_p(dos.data_ptr()),
                                      C.c_void_p(stats.data_ptr()), sh), "slk_dos_launch")
    return edges, dos, stats


def reaches_padding(row: dict) -> bool:
    """True when a structure's highest band (mu + its highest shifted level)
    comes within 1e-6 Ha of the identity-padding level PAD_H: the padded
    eigenvalues could then sort among the structure's own bands."""
    if row.get("n_con", 0) <= 0:
        top = row["mu"] + row["s_vbm"] / H2E
    else:
        top = row["mu"] + (row["s_cbm"] + row["cb_width"]) / H2E
    return top >= PAD_H - 1e-6


def edges_dict(row: np.ndarray) -> dict:
    return {k: float(v) for k, v in zip(EDGE_FIELDS, row)}
