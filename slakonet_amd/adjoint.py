"""Differentiation form: hand-written adjoints behind torch.autograd.Function.

Eigenvalues. For H c = e S c with S-orthonormal eigenvectors c,
de_n = c_n^H (dH - e_n dS) c_n. For a loss with g = dL/de this gives
    dL = sum_k Re sum_{mu,nu} QH_k[nu][mu] dH_k[mu][nu] + QS_k[nu][mu] dS_k[mu][nu],
    QH_k = C_k diag(g_k) C_k^H,   QS_k = -C_k diag(g_k e_k) C_k^H,
formed on the device by a column scaling and one MFMA GEMM per matrix
(slk_scale_columns + slk_zgemm). slk_adjoint_tables_launch then carries Q
back through the Bloch sum, the Slater-Koster rotation and the interpolation
weights onto the packed table rows; torch's own graph maps the packed rows to
the trainable per-pair parameters (the packing is a concatenation).

This replaces the reference's autograd path: the CPU build (hs_matrix,
slakonet/slaterkoster.py:103-441, SKF interpolation,
slakonet/interpolation.py:69-311) and _SymEigB.backward
(slakonet/utils.py:386-473). The eigenvalue part of the latter is
V diag(g) V^T without a complex conjugate; it equals C diag(g) C^H where the
eigenvectors are real (Gamma) and differs from the true derivative at
complex k. The formula here is the true derivative everywhere.

Fermi level. As in the reference (fermi_search returns a detached leaf,
slakonet/fermi.py:139-215), mu carries no gradient by default: the DOS and
gap gradients are taken at fixed mu. GaussianDos(..., fermi_grad=True) adds
mu's own dependence on the eigenvalues (implicit function theorem on
N(mu) = Ne, slk_fermi_chain_launch), the derivative of the Fermi-shifted DOS
as evaluated; the fitting loop uses it for its DOS loss.

DOS. slk_dos_adjoint_launch is the adjoint of the Gaussian DOS kernel.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import torch

from . import _lib, engine
from .tables import TableSet
from .units import FERMI_KT_HA, H2E, HS_CUTOFF_BOHR

OP_N, OP_C = 0, 3


def weighted_densities(V: torch.Tensor, w: torch.Tensor, g: torch.Tensor, stream=None):
    """QH = C diag(g) C^H and QS = -C diag(g e) C^H for every matrix of a batch.

    V: [batch, ld, ld] complex128 eigenvectors (columns), w, g: [batch, ld]
    float64 (g zero on padded columns)."""
    lib = _lib.lib()
    batch, ld, _ = V.shape
    g = g.to(torch.float64).contiguous()
    gs = (-g * w).contiguous()
    Y = torch.empty_like(V)
    QH = torch.empty_like(V)
    QS = torch.empty_like(V)
    sh = _lib.stream_handle(stream)
    st = ld * ld
    for gv, Q in ((g, QH), (gs, QS)):
        _lib.check(lib.slk_scale_columns(_lib.ptr(V), _lib.ptr(gv), ld, ld, batch, ld, _lib.ptr(Y), sh),
                   "slk_scale_columns")
        _lib.check(lib.slk_zgemm(batch, ld, ld, ld, 1.0, _lib.ptr(Y), ld, st, OP_N, _lib.ptr(V), ld, st, OP_C,
                                 0.0, _lib.ptr(Q), ld, st, sh), "slk_zgemm")
    return QH, QS


def table_gradient_from_densities(tables: TableSet, plan: engine.BuildPlan, QH: torch.Tensor, QS: torch.Tensor,
                                  table_tensor: Optional[torch.Tensor] = None, cutoff: float = HS_CUTOFF_BOHR,
                                  stream=None) -> torch.Tensor:
    """dL/d(table rows) [rows, 20] given QH, QS laid out as the plan's H/S."""
    lib = _lib.lib()
    if QH.shape != (plan.n_mat, plan.ld, plan.ld) or QH.dtype != torch.complex128 or QS.shape != QH.shape:
        raise ValueError("QH/QS must be complex128 [n_mat, ld, ld] of the plan")
    dtab = torch.zeros(tables.n_rows, 20, dtype=torch.float64, device=tables.device)
    args = engine.build_args(tables, plan, 0, 0, False, cutoff, table_tensor)
    _lib.check(lib.slk_adjoint_tables_launch(C.byref(args), _lib.ptr(QH), _lib.ptr(QS), _lib.ptr(dtab),
                                             _lib.stream_handle(stream)), "slk_adjoint_tables_launch")
    return dtab


def position_gradient_from_densities(tables: TableSet, plan: engine.BuildPlan, QH: torch.Tensor, QS: torch.Tensor,
                                     cutoff: float = HS_CUTOFF_BOHR, stream=None) -> torch.Tensor:
    """dL/dR [atoms of the plan, 3] (Bohr) given QH, QS laid out as the plan's H/S."""
    lib = _lib.lib()
    n_atoms = int(plan.dev["atom_spec"].numel() // 4)
    dR = torch.zeros(n_atoms, 3, dtype=torch.float64, device=tables.device)
    args = engine.build_args(tables, plan, 0, 0, False, cutoff)
    _lib.check(lib.slk_forces_launch(C.byref(args), _lib.ptr(QH), _lib.ptr(QS), _lib.ptr(dR),
                                     _lib.stream_handle(stream)), "slk_forces_launch")
    return dR


def band_energy_weights(w: torch.Tensor, k_weights: torch.Tensor, mu: float, kT: float, n: int) -> torch.Tensor:
    """dE/de for E = sum_k w_k sum_b f(e_kb) e_kb at fixed mu (Fermi-Dirac f)."""
    e = w[:, :n]
    f = 1.0 / (1.0 + torch.exp((e - mu) / kT))
    f = torch.where(e == 0, torch.zeros_like(f), f)
    fp = -f * (1.0 - f) / kT
    g = torch.zeros_like(w)
    g[:, :n] = k_weights.reshape(-1, 1) * (f + e * fp)
    return g


def table_gradient(tables: TableSet, plan: engine.BuildPlan, V: torch.Tensor, w: torch.Tensor, g: torch.Tensor,
                   table_tensor: Optional[torch.Tensor] = None, stream=None) -> torch.Tensor:
    """dL/d(table rows) for eigenvalue gradient g [n_mat, ld] of the plan's solve."""
    QH, QS = weighted_densities(V, w, g, stream)
    return table_gradient_from_densities(tables, plan, QH, QS, table_tensor, stream=stream)


class TableEigenvalues(torch.autograd.Function):
    """Eigenvalues [n_mat, ld] of a solved plan as a function of the packed tables.

    The forward returns the already computed eigenvalues (the device solve
    ran outside autograd); the backward is the hand-written adjoint."""

    @staticmethod
    def forward(ctx, table_tensor, tables, plan, V, w):
        ctx.tables, ctx.plan, ctx.V, ctx.w = tables, plan, V, w
        return w.clone()

    @staticmethod
    def backward(ctx, g):
        dtab = table_gradient(ctx.tables, ctx.plan, ctx.V, ctx.w, g)
        return dtab, None, None, None, None


class GaussianDos(torch.autograd.Function):
    """DOS [n_struct, ne] of a plan's eigenvalues at the Fermi levels in ``edges``.

    fermi_grad False: the Fermi level is held fixed in the backward, as the
    reference's detached fermi_search (slakonet/fermi.py:198) does. True: the
    backward also carries the Fermi level's own dependence on the eigenvalues
    (slk_fermi_chain_launch; the edges must come from the bisection), which
    makes it the derivative of the Fermi-shifted DOS as evaluated."""

    @staticmethod
    def forward(ctx, w, plan, edges, grid, sigma, fermi_grad=False):
        lib = _lib.lib()
        ne = int(grid.numel())
        jobs = engine._upload(engine.dos_jobs(plan, w.shape[1], ne), w.device)
        dos = torch.empty(plan.n_struct, ne, dtype=torch.float64, device=w.device)
        wc = w.detach().contiguous()
        _lib.check(lib.slk_dos_launch(_lib.ptr(wc), _lib.ptr(plan.k_weights), _lib.ptr(jobs), plan.n_struct,
                                      _lib.ptr(grid), ne, float(sigma), _lib.ptr(edges), _lib.ptr(dos), None,
                                      _lib.stream_handle()), "slk_dos_launch")
        ctx.save_for_backward(wc, edges, grid)
        ctx.plan, ctx.jobs, ctx.sigma, ctx.fermi_grad = plan, jobs, float(sigma), bool(fermi_grad)
        return dos

    @staticmethod
    def backward(ctx, gdos):
        lib = _lib.lib()
        w, edges, grid = ctx.saved_tensors
        plan = ctx.plan
        gdos = gdos.to(torch.float64).contiguous()
        gw = torch.zeros_like(w)
        max_states = max(nk * n for nk, n in zip(plan.nk, plan.n_orb))
        _lib.check(lib.slk_dos_adjoint_launch(_lib.ptr(w), _lib.ptr(plan.k_weights), _lib.ptr(ctx.jobs),
                                              plan.n_struct, max_states, _lib.ptr(grid), int(grid.numel()),
                                              ctx.sigma, _lib.ptr(edges), _lib.ptr(gdos), _lib.ptr(gw),
                                              _lib.stream_handle()), "slk_dos_adjoint_launch")
        if ctx.fermi_grad:
            spd = engine._upload(engine.spectra(plan, w.shape[1]), w.device)
            _lib.check(lib.slk_fermi_chain_launch(_lib.ptr(w), _lib.ptr(plan.k_weights), _lib.ptr(spd), plan.n_struct,
                                                  FERMI_KT_HA, _lib.ptr(edges), _lib.ptr(gw), _lib.stream_handle()),
                       "slk_fermi_chain_launch")
        return gw, None, None, None, None, None


def attach_table_gradient(calc, tables: TableSet):
    """Eigenvalues [1, nk, nb] of a solved SimpleDftb carrying the table adjoint."""
    if getattr(calc, "_V", None) is None:
        raise RuntimeError("the calculator was solved without eigenvectors; the adjoint needs them")
    w_full = TableEigenvalues.apply(tables.tables, tables, calc._plan, calc._V, calc._w)
    calc._w_grad = w_full
    n = calc._plan.n_orb[0]
    return w_full[:, :n].unsqueeze(0)


def gap_tensor(w: torch.Tensor, edges_row: dict) -> torch.Tensor:
    """Band gap in eV as cbm - vbm of the two eigenvalues the edge kernel picked."""
    if edges_row["metal"] > 0 or edges_row["vbm"] < -1e299 or edges_row["cbm"] > 1e299:
        # metal, or no occupied / no unoccupied state (the properties dict then
        # reports a zero gap, slakonet/main.py:1677-1687)
        return torch.tensor(0.0, dtype=torch.float64, device=w.device)
    kv, kc = int(edges_row["vbm_k"]), int(edges_row["cbm_k"])
    wd = w.detach()
    bv = int(torch.argmin((wd[kv] * H2E - edges_row["vbm"]).abs()))
    bc = int(torch.argmin((wd[kc] * H2E - edges_row["cbm"]).abs()))
    return (w[kc, bc] - w[kv, bv]) * H2E


def differentiable_properties(calc, eigenvalues, energy_range=(-10.0, 10.0), num_points=5000, sigma=0.1):
    """band_gap_eV and dos_values_tensor of a solved calculator with gradients
    to the tables (the DOS grid and broadening of get_properties_dict,
    slakonet/main.py:1757-1762)."""
    w = calc._w_grad
    n = calc._plan.n_orb[0]
    calc._edge_values()
    grid = torch.linspace(energy_range[0], energy_range[1], num_points, dtype=torch.float64, device=w.device)
    dos = GaussianDos.apply(w, calc._plan, calc._edges_t, grid, sigma)[0]
    gap = gap_tensor(w[:, :n], calc._edges)
    return {"band_gap_eV": gap, "dos_values_tensor": dos, "dos_energy_grid_tensor": grid}
