"""ctypes binding of libslakonet_amd.so (declarations in include/slakonet_amd.h).

The library is the only compute path: ``lib()`` raises if it is missing or
no GPU is visible. Structures below mirror the header field by field.
"""
from __future__ import annotations

import ctypes as C
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libslakonet_amd.so")

c_int_p = C.POINTER(C.c_int)
c_double_p = C.POINTER(C.c_double)


class SlkStructDesc(C.Structure):
    _fields_ = [("atom0", C.c_int), ("n_atoms", C.c_int), ("n_orb", C.c_int), ("ld", C.c_int),
                ("mat0", C.c_longlong), ("k0", C.c_int), ("nk", C.c_int), ("t0", C.c_int), ("nt", C.c_int)]


class SlkBuildWork(C.Structure):
    _fields_ = [(f, C.c_int) for f in ("s", "i", "j0", "j1", "c0", "ncol", "kc0", "nkc")]


class SlkTableSet(C.Structure):
    _fields_ = [("tables", C.c_void_p), ("pair_row0", C.c_void_p), ("pair_ngrid", C.c_void_p),
                ("pair_step", C.c_void_p), ("pair_index", C.c_void_p), ("spec_norb", C.c_void_p),
                ("spec_slot", C.c_void_p), ("spec_onsite", C.c_void_p), ("nspec", C.c_int),
                ("tail", C.c_double)]


class SlkBuildArgs(C.Structure):
    _fields_ = [("tab", SlkTableSet), ("structs", C.c_void_p), ("work", C.c_void_p), ("n_work", C.c_int),
                ("atom_pos", C.c_void_p), ("atom_spec", C.c_void_p), ("atom_orb", C.c_void_p),
                ("trans_cart", C.c_void_p), ("trans_int", C.c_void_p), ("kfrac", C.c_void_p),
                ("cutoff", C.c_double), ("H", C.c_void_p), ("S", C.c_void_p), ("out_f32", C.c_int),
                ("pad_h", C.c_double), ("n_struct", C.c_int), ("n_mat", C.c_int), ("ld", C.c_int),
                ("n_row_work", C.c_int)]


class SlkSpectrum(C.Structure):
    _fields_ = [("w0", C.c_longlong), ("nk", C.c_int), ("nb", C.c_int), ("wld", C.c_int), ("kw0", C.c_int),
                ("nelec", C.c_double)]


class SlkDosJob(C.Structure):
    _fields_ = [("w0", C.c_longlong), ("nk", C.c_int), ("nb", C.c_int), ("wld", C.c_int), ("kw0", C.c_int),
                ("edge", C.c_int), ("out0", C.c_longlong)]


PLAN_SECTIONS = 9


class SlkPlanIn(C.Structure):
    _fields_ = [("n_struct", C.c_int), ("n_atoms", C.c_void_p), ("spec", C.c_void_p), ("pos", C.c_void_p),
                ("cell", C.c_void_p), ("nk", C.c_void_p), ("kfrac", C.c_void_p), ("kw", C.c_void_p),
                ("spec_norb", C.c_void_p), ("nspec", C.c_int), ("cutoff", C.c_double), ("ld", C.c_int)]


class SlkPlanOut(C.Structure):
    _fields_ = [("off", C.c_longlong * PLAN_SECTIONS), ("bytes", C.c_longlong * PLAN_SECTIONS),
                ("total", C.c_longlong), ("n_work", C.c_int), ("n_trans", C.c_int), ("n_mat", C.c_int),
                ("ld", C.c_int), ("n_row_work", C.c_int)]


# name -> (restype, argtypes)
SIGNATURES = {
    "slk_plan_size": (C.c_int, [C.POINTER(SlkPlanIn), C.POINTER(SlkPlanOut)]),
    "slk_plan_fill": (C.c_int, [C.POINTER(SlkPlanIn), C.POINTER(SlkPlanOut), C.c_void_p, C.c_void_p, C.c_void_p]),
    "slk_build_hs_launch": (C.c_int, [C.POINTER(SlkBuildArgs), C.c_void_p]),
    "slk_set_build_mode": (C.c_int, [C.c_int]),
    "slk_pad_launch": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                 C.c_double, C.c_int, C.c_void_p]),
    "slk_eigh_workspace_bytes": (C.c_size_t, [C.c_int, C.c_int, C.c_int, C.c_int]),
    "slk_eigh_f64": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int,
                               C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p]),
    "slk_eigh_mixed_workspace_bytes": (C.c_size_t, [C.c_int, C.c_int, C.c_int]),
    "slk_eigh_mixed": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                 C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p]),
    "slk_eigh_mixed_timed": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                       C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.POINTER(C.c_float)]),
    "slk_eigh_ex": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                              C.c_void_p, C.c_size_t, C.c_void_p, C.c_int, C.POINTER(C.c_float), C.c_void_p]),
    "slk_noalloc_scratch_bytes": (C.c_size_t, []),
    "slk_eigh_mixed_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                         C.c_void_p, C.c_void_p]),
    "slk_set_tridiag_mode": (C.c_int, [C.c_int]),
    "slk_set_chase_spin_limit": (C.c_int, [C.c_int]),
    "slk_eigh_small_f64": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                     C.c_void_p, C.c_void_p]),
    "slk_eigh_f64_timed": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                     C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p,
                                     C.POINTER(C.c_float)]),
    "slk_zgemm": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_int, C.c_longlong,
                            C.c_int, C.c_void_p, C.c_int, C.c_longlong, C.c_int, C.c_double, C.c_void_p, C.c_int,
                            C.c_longlong, C.c_void_p]),
    "slk_cgemm": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_void_p, C.c_int, C.c_longlong,
                            C.c_int, C.c_void_p, C.c_int, C.c_longlong, C.c_int, C.c_float, C.c_void_p, C.c_int,
                            C.c_longlong, C.c_void_p]),
    "slk_fermi_launch": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_int,
                                   C.c_double, C.c_int, C.c_void_p, C.c_void_p]),
    "slk_edges_launch": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_void_p]),
    "slk_edge_fields": (C.c_int, []),
    "slk_dos_launch": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_double,
                                 C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "slk_refine_scratch_bytes": (C.c_size_t, [C.c_int, C.c_int]),
    "slk_refine_edges_mixed": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_size_t,
                                         C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                         C.c_size_t, C.c_void_p]),
    "slk_pdos_launch": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                  C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                                  C.c_void_p]),
    "slk_sampler_reset": (C.c_int, [C.c_int, C.c_int, C.c_int]),
    "slk_sampler_hint": (C.c_int, [C.c_int, C.c_double, C.c_double]),
    "slk_sampler_read": (C.c_int, [C.c_int, C.c_void_p]),
    "slk_dos_adjoint_launch": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                         C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "slk_fermi_chain_launch": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_void_p,
                                         C.c_void_p, C.c_void_p]),
    "slk_scale_columns": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                    C.c_void_p]),
    "slk_adjoint_tables_launch": (C.c_int, [C.POINTER(SlkBuildArgs), C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_void_p]),
    "slk_forces_launch": (C.c_int, [C.POINTER(SlkBuildArgs), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
}

_LIB = None


def load_library(path: str = LIB_PATH):
    """Load the .so and bind signatures (no GPU needed)."""
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m slakonet_amd.build`; "
                           "there is no CPU fallback")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def host_lib():
    """The library for its host-only entry points (no GPU needed)."""
    global _HOST
    if _HOST is None:
        _HOST = load_library()
    return _HOST


_HOST = None


def lib():
    global _LIB
    if _LIB is None:
        if not torch.cuda.is_available():
            raise RuntimeError("slakonet_amd needs an MI355X GPU (HIP); no device is visible")
        _LIB = load_library()
    return _LIB


def check(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


def ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def stream_handle(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def struct_array(ctype, items, device):
    """Pack a list of ctypes structures into a device byte tensor."""
    n = len(items)
    arr = (ctype * max(n, 1))(*items) if n else (ctype * 1)()
    raw = bytes(arr)[: C.sizeof(ctype) * max(n, 1)]
    host = torch.frombuffer(bytearray(raw), dtype=torch.uint8)
    return host.to(device, non_blocking=False)
