"""ctypes binding of the HIP library (``libcunibs_amd.so``, declared in ``include/cunibs_amd.h``).

This replaces the reference's nanobind extension ``cunibs.solver._solver_ext``
(src/cunibs/solver/bindings.cpp:30-676). There is no fallback: if the library is missing or fails
to load, importing anything that needs it raises, naming the path and how to build it.
"""

from __future__ import annotations

import ctypes
import threading
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libcunibs_amd.so"

CN_OK, CN_EINVAL, CN_EHIP, CN_ENOCONV, CN_EINTERNAL = 0, 1, 2, 3, 4
RECOVERY_CODES = {"raw": 0, "spr_global": 1, "spr_tissue": 2, "harmonic": 3}
METRICS_LEN = 16
MAX_LEVELS = 32

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_D = ctypes.c_double
_F = ctypes.c_float


class SolverConvergenceError(RuntimeError):
    """The block PCG missed tolerance (fem/solve.py:368-392)."""

    def __init__(self, message: str, *, column: int | None = None) -> None:
        super().__init__(message)
        self.column = column


class Options(ctypes.Structure):
    _fields_ = [
        ("min_coarse_rows", _I32), ("max_levels", _I32), ("rounds", _I32), ("degree", _I32),
        ("lower_ratio", _F), ("keep_stiffness", _I32), ("conductivity", _P),
        ("n_conductivity", _I32), ("region_tag", _I32),
    ]


class Info(ctypes.Structure):
    _fields_ = [
        ("n_nodes", _I64), ("n_tets", _I64), ("n_skin", _I64),
        ("n_reduced", _I64), ("nnz_reduced", _I64), ("nnz_full", _I64),
        ("ground_node", _I32), ("n_levels", _I32),
        ("level_rows", _I64 * (MAX_LEVELS + 1)), ("level_nnz", _I64 * (MAX_LEVELS + 1)),
        ("n_region", _I64), ("region_volume_m3", _D), ("n_cu", _I32),
        ("level_p_nnz", _I64 * MAX_LEVELS), ("ell_width_full", _I32), ("ell_width_off", _I32),
        ("preconditioner_rebuilds", _I32), ("matched_preconditioners", _I32), ("dense_tail_level", _I32),
    ]


class Tuning(ctypes.Structure):
    _fields_ = [("degree", _I32), ("lower_ratio", _F), ("fine_sub", _I32), ("coarse_sub", _I32),
                ("coil_blocks", _I32), ("layout", _I32), ("dense_tail", _I32)]


class SolveOptions(ctypes.Structure):
    _fields_ = [
        ("tolerance", _D), ("max_iters", _I32), ("max_refine", _I32), ("drop_tol", _D),
        ("use_graph", _I32), ("check_every", _I32), ("recovery", _I32), ("warm_start", _I32), ("collect", _I32),
        ("sigma", _P), ("n_sigma", _I32), ("independent", _I32), ("recycle", _I32),
    ]


class BlockOutputs(ctypes.Structure):
    _fields_ = [
        ("transforms", _P), ("magnE", _P), ("E", _P), ("v", _P), ("metrics", _P),
        ("iterations", _P), ("relative_residual", _P), ("dadt_elm", _P), ("dadt_nodes", _P),
        ("rhs", _P), ("degenerate", _P), ("E_slots", _P), ("x0", _P), ("total_iterations", _I32), ("refinements", _I32),
        ("round_iterations", _I32 * 8), ("round_residual", _D * 8),
    ]


_SIGNATURES: dict[str, tuple[object, list[object]]] = {
    "cn_last_error": (ctypes.c_char_p, []),
    "cn_version": (_I32, []),
    "cn_block_sizes": (_I32, [_P, _I32]),
    "cn_max_block": (_I32, []),
    "cn_default_options": (None, [_P]),
    "cn_default_solve_options": (None, [_P]),
    "cn_context_create": (_I32, [_P, _I64, _P, _P, _I64, _P, _I64, _P, _P]),
    "cn_context_destroy": (None, [_P]),
    "cn_context_info": (_I32, [_P, _P]),
    "cn_context_get_tuning": (_I32, [_P, _P]),
    "cn_context_set_tuning": (_I32, [_P, _P]),
    "cn_context_mesh": (_I32, [_P, _P, _P, _P, _P, _P, _P]),
    "cn_context_geometry": (_I32, [_P, _P, _P, _P, _P]),
    "cn_context_stiffness": (_I32, [_P, _P, _P, _P]),
    "cn_context_reduced": (_I32, [_P, _P, _P, _P]),
    "cn_context_level": (_I32, [_P, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "cn_context_skin_normals": (_I32, [_P, _P]),
    "cn_solve_block": (_I32, [_P, _P, _P, _I64, _P, _P, _P, _I32, _D, _P, _P, _P]),
    "cn_solve_sweep": (_I32, [_P, _P, _P, _I64, _P, _P, _P, _I32, _I32, _D, _P, _P, _P, _P, _P, _P]),
    "cn_recovery_build": (_I32, [_P, _I32, _P, _P]),
    "cn_warm_reset": (_I32, [_P, _I32]),
    "cn_warm_count": (_I32, [_P, _P]),
    "cn_warm_gram": (_I32, [_P, _P]),
    "cn_warm_set_modes": (_I32, [_P, _P, _I32, _I32]),
    "cn_warm_mode_grams": (_I32, [_P, _P, _P]),
    "cn_warm_set_projector": (_I32, [_P, _P, _I32]),
    "cn_warm_set_coefficients": (_I32, [_P, _P, _I32, _I32]),
    "cn_warm_project_last": (_I32, [_P, _I32, _P]),
    "cn_warm_modes": (_I32, [_P, _P, _P]),
    "cn_uq_prepare": (_I32, [_P, _P, _P, _I64, _P, _P, _D, _D, _P, _P]),
    "cn_uq_solve": (_I32, [_P, _P, _I32, _P, _P]),
    "cn_uq_moments": (_I32, [_P, _P, _P, _P]),
    "cn_uq_guess_info": (_I32, [_P, _P]),
    "cn_mlp_create": (_I32, [_I32, _I32, _I32, ctypes.c_uint32, _P]),
    "cn_mlp_destroy": (None, [_P]),
    "cn_mlp_n_params": (_I32, [_P, _P]),
    "cn_mlp_set_params": (_I32, [_P, _P]),
    "cn_mlp_get_params": (_I32, [_P, _P]),
    "cn_mlp_gradients": (_I32, [_P, _P, _P, _I32, _P, _P]),
    "cn_mlp_train": (_I32, [_P, _P, _P, _I64, _I32, _I32, _F, ctypes.c_uint32, _P]),
    "cn_mlp_predict": (_I32, [_P, _P, _I32, _P]),
    "cn_recovery_operator": (_I32, [_P, _I32, _P, _P, _P, _P, _P, _P, _P, _P]),
    "cn_recovery_patches": (_I32, [_P, _P, _P, _P, _P, _P]),
    "cn_k_p1_gradients": (_I32, [_P, _I64, _P, _I64, _P, _P]),
    "cn_k_morton_keys": (_I32, [_P, _I64, _P]),
    "cn_k_tet_lowest_node": (_I32, [_P, _I64, _P, _I64, _P]),
    "cn_k_node2corner": (_I32, [_P, _I64, _I32, _I64, _P, _P]),
    "cn_k_l1_dinv": (_I32, [_I64, _P, _P, _P, _P]),
    "cn_k_select_size4": (_I32, [_I64, _P, _P, _P, _P, _P]),
    "cn_k_spgemm": (_I32, [_I64, _I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "cn_k_spmm": (_I32, [_P, _P, _P, _I32]),
    "cn_k_spmm_diff": (_I32, [_P, _P, _P, _I32]),
    "cn_k_gram": (_I32, [_P, _P, _P, _P, _P, _I32]),
    "cn_k_vcycle": (_I32, [_P, _P, _P, _I32]),
    "cn_k_dadt": (_I32, [_P, _P, _P, _I64, _P, _I64, _F, _F, _P]),
    "cn_k_place": (_I32, [_P, _P, _P, _P, _I32, _P, _P]),
    "cn_bench_spmm": (_I32, [_P, _I32, _I32, _I32, _I32, _P]),
    "cn_bench_vcycle": (_I32, [_P, _I32, _I32, _P]),
    "cn_bench_copy": (_I32, [_P, _I64, _I32, _P]),
    "cn_timing_enable": (_I32, [_P, _I32]),
    "cn_timing_read": (_I32, [_P, _P]),
    "cn_timing_phases": (_I32, [_P, _P]),
    "cn_bench_guess": (_I32, [_P, _I32, _I32, _P]),
    "cn_bench_dadt": (_I32, [_P, _P, _P, _I64, _P, _P, _D, _I32, _I32, _P]),
    "cn_bench_stages": (_I32, [_P, _P, _P, _I64, _P, _P, _D, _I32, _I32, _I32, _P]),
    "cn_k_gram_noalloc": (_I32, [_I64, _I32, _P, _P, _P, _P, _P]),
    "cn_context_setup_times": (_I32, [_P, _P]),
    "cn_debug_fail_captures": (_I32, [_P, _I32]),
    "cn_graph_execs": (_I32, [_P, _P]),
}

EXPORTED = tuple(_SIGNATURES)

_lib: ctypes.CDLL | None = None
_lock = threading.Lock()


def lib() -> ctypes.CDLL:
    """Load the library once; raise with a build hint if it is absent."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise ImportError(
                    f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                    "g.build()'` (or `make -C cunibs_amd/csrc`). There is no CPU fallback.")
            handle = ctypes.CDLL(str(LIB_PATH))
            for name, (res, args) in _SIGNATURES.items():
                fn = getattr(handle, name)
                fn.restype = res
                fn.argtypes = args
            _lib = handle
        return _lib


def last_error() -> str:
    return (lib().cn_last_error() or b"").decode(errors="replace")


def check(status: int, what: str) -> None:
    if status == CN_OK:
        return
    msg = last_error() or what
    if status == CN_EINVAL:
        raise ValueError(msg)
    if status == CN_ENOCONV:
        col = None
        marker = "(column "
        if marker in msg:
            try:
                col = int(msg.split(marker, 1)[1].split(" ", 1)[0])
            except ValueError:
                col = None
        raise SolverConvergenceError(msg, column=col)
    raise RuntimeError(f"{what}: {msg}")


def ptr(a: np.ndarray | None) -> int | None:
    if a is None:
        return None
    if not a.flags.c_contiguous:
        raise ValueError("arrays handed to the library must be C-contiguous")
    return a.ctypes.data


def block_sizes() -> tuple[int, ...]:
    buf = np.zeros(16, dtype=np.int32)
    n = lib().cn_block_sizes(ptr(buf), 16)
    return tuple(int(v) for v in buf[:n])


def max_block() -> int:
    return int(lib().cn_max_block())


def default_options() -> Options:
    o = Options()
    lib().cn_default_options(ctypes.byref(o))
    return o


def default_solve_options() -> SolveOptions:
    o = SolveOptions()
    lib().cn_default_solve_options(ctypes.byref(o))
    return o
