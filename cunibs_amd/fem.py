"""Device-side solver context and block solve, over the C ABI.

Mirrors the reference's FEM layer (src/cunibs/fem/solve.py): ``build_context`` (:690-746) becomes
``SolverContext(mesh)``, ``solve_placements_block`` (:865-979) keeps its name and its
(center_mm, pos_ydir_mm, distance_mm) site triples. All device state lives in the HIP library;
this module only moves host arrays across the boundary.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from cunibs_amd import _lib
from cunibs_amd._lib import SolverConvergenceError, check, lib, ptr

DEFAULT_TOLERANCE = 1e-6  # fem/solve.py:56
DEFAULT_MAX_ITERS = 2000  # fem/solve.py:57
# Compiled block widths, read from the library (solver/__init__.py BLOCK_SIZES there). Importing
# this module loads the HIP library; there is no CPU fallback.
BLOCK_SIZES: tuple[int, ...] = _lib.block_sizes()
MAX_BLOCK: int = _lib.max_block()
# Block width of the drop-in front end when the caller names none: the width the autotuner picks
# on the 1.5 mm head (throughput falls at 32, where the wider Gram and coefficient work outgrow
# the shared operator reads).
DEFAULT_BLOCK: int = 16
MU0_OVER_4PI = 1e-7  # fem/placement.py:16

TISSUE_CONDUCTIVITY: dict[int, float] = {
    1: 0.126, 2: 0.275, 3: 1.654, 4: 0.010, 5: 0.465,
    6: 0.5, 7: 0.008, 8: 0.025, 9: 0.6, 10: 0.16,
}  # fem/assembly.py:15-26
GM_TAG = 2

METRIC_FIELDS = (
    "peak_magnE", "peak_x", "peak_y", "peak_z", "cog_x", "cog_y", "cog_z", "region_volume_m3",
    "focality_0.5", "focality_0.75", "mean", "std", "p50", "p95", "p99", "p99.9",
)


def _blocks() -> tuple[int, ...]:
    return BLOCK_SIZES


@dataclass(frozen=True)
class AggregationParams:
    """fem/solve.py:137-150."""

    min_coarse_rows: int = 128
    max_levels: int = 50
    rounds: int = 2


@dataclass(frozen=True)
class SmootherParams:
    """fem/solve.py:153-168."""

    degree: int = 2
    lower_ratio: float = 12.0


@dataclass
class PlacementResult:
    transform: np.ndarray
    magnE: np.ndarray | None
    E: np.ndarray | None
    v: np.ndarray | None
    metrics: np.ndarray | None
    iterations: int
    relative_residual: float
    E_slots: np.ndarray | None = None


class SolverContext:
    """Per-mesh GPU state: reordered mesh, FP64 assembly, grounded operator, AMG hierarchy.

    The mesh is laid out along a Morton curve first (fem/solve.py:645-662); every per-node and
    per-element array a solve returns is indexed in that order, available as ``self.nodes_mm``,
    ``self.tet_nodes``, ``self.tet_tags`` and ``self.skin_tris``.
    """

    def __init__(
        self,
        nodes_mm: np.ndarray,
        tet_nodes: np.ndarray,
        tet_tags: np.ndarray,
        skin_tris: np.ndarray,
        *,
        aggregation: AggregationParams | None = None,
        smoother: SmootherParams | None = None,
        keep_stiffness: bool = False,
        conductivity: dict[int, float] | None = None,
        region_tag: int = GM_TAG,
    ) -> None:
        _blocks()
        agg = aggregation or AggregationParams()
        sm = smoother or SmootherParams()
        nodes = np.ascontiguousarray(nodes_mm, dtype=np.float64).reshape(-1, 3)
        tets = np.ascontiguousarray(tet_nodes, dtype=np.int32).reshape(-1, 4)
        tags = np.ascontiguousarray(tet_tags, dtype=np.int32).reshape(-1)
        skin = np.ascontiguousarray(skin_tris, dtype=np.int32).reshape(-1, 3)
        if tags.shape[0] != tets.shape[0]:
            raise ValueError("tet_tags must have one entry per tetrahedron")
        opts = _lib.default_options()
        opts.min_coarse_rows = agg.min_coarse_rows
        opts.max_levels = agg.max_levels
        opts.rounds = agg.rounds
        opts.degree = sm.degree
        opts.lower_ratio = sm.lower_ratio
        opts.keep_stiffness = int(keep_stiffness)
        opts.region_tag = int(region_tag)
        self._sigma = None
        if conductivity is not None:
            top = max(conductivity)
            lut = np.full(top + 1, np.nan)
            for t, s in conductivity.items():
                lut[t] = s
            self._sigma = lut
            opts.conductivity = lut.ctypes.data
            opts.n_conductivity = lut.shape[0]
        handle = ctypes.c_void_p()
        check(lib().cn_context_create(ptr(nodes), nodes.shape[0], ptr(tets), ptr(tags), tets.shape[0],
                                      ptr(skin), skin.shape[0], ctypes.byref(opts), ctypes.byref(handle)),
              "cn_context_create")
        self._h = handle
        info = self.info
        self.n_nodes = int(info.n_nodes)
        self.n_tets = int(info.n_tets)
        self.n_skin = int(info.n_skin)
        self.ground_node = int(info.ground_node)
        self.n_region = int(info.n_region)
        self.region_volume_m3 = float(info.region_volume_m3)
        self.region_tag = int(region_tag)
        self.hierarchy_sizes = [int(info.level_rows[i]) for i in range(info.n_levels + 1)]
        self._mesh: tuple[np.ndarray, ...] | None = None
        self._geom: tuple[np.ndarray, ...] | None = None

    # ---- lifetime ----
    @property
    def info(self) -> _lib.Info:
        """The context's sizes and counters (cn_context_info), read afresh on every access."""
        info = _lib.Info()
        check(lib().cn_context_info(self.handle, ctypes.byref(info)), "cn_context_info")
        return info

    def free(self) -> None:
        if getattr(self, "_h", None):
            lib().cn_context_destroy(self._h)
            self._h = None

    def __del__(self) -> None:
        try:
            self.free()
        except Exception:
            pass

    @property
    def handle(self) -> ctypes.c_void_p:
        if not self._h:
            raise RuntimeError("the solver context has been freed")
        return self._h

    # ---- host copies of device state (cached) ----
    def _mesh_arrays(self):
        if self._mesh is None:
            nodes = np.empty((self.n_nodes, 3), np.float64)
            tets = np.empty((self.n_tets, 4), np.int32)
            tags = np.empty(self.n_tets, np.int32)
            skin = np.empty((self.n_skin, 3), np.int32)
            nperm = np.empty(self.n_nodes, np.int32)
            tperm = np.empty(self.n_tets, np.int32)
            check(lib().cn_context_mesh(self.handle, ptr(nodes), ptr(tets), ptr(tags), ptr(skin), ptr(nperm),
                                        ptr(tperm)), "cn_context_mesh")
            self._mesh = (nodes, tets, tags, skin, nperm, tperm)
        return self._mesh

    @property
    def nodes_mm(self) -> np.ndarray:
        return self._mesh_arrays()[0]

    @property
    def tet_nodes(self) -> np.ndarray:
        return self._mesh_arrays()[1]

    @property
    def tet_tags(self) -> np.ndarray:
        return self._mesh_arrays()[2]

    @property
    def skin_tris(self) -> np.ndarray:
        return self._mesh_arrays()[3]

    @property
    def node_perm(self) -> np.ndarray:
        return self._mesh_arrays()[4]

    @property
    def tet_perm(self) -> np.ndarray:
        return self._mesh_arrays()[5]

    def _geometry(self):
        if self._geom is None:
            vols = np.empty(self.n_tets, np.float32)
            bary = np.empty((self.n_tets, 3), np.float64)
            g = np.empty((self.n_tets, 4, 3), np.float32)
            nvc = np.empty(self.n_tets, np.float32)
            check(lib().cn_context_geometry(self.handle, ptr(vols), ptr(bary), ptr(g), ptr(nvc)),
                  "cn_context_geometry")
            self._geom = (vols, bary, g, nvc)
        return self._geom

    @property
    def vols(self) -> np.ndarray:
        return self._geometry()[0]

    @property
    def tet_barycenters_mm(self) -> np.ndarray:
        return self._geometry()[1]

    @property
    def g(self) -> np.ndarray:
        return self._geometry()[2]

    @property
    def neg_vc(self) -> np.ndarray:
        return self._geometry()[3]

    def stiffness(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        nnz = int(self.info.nnz_full)
        indptr = np.empty(self.n_nodes + 1, np.int32)
        indices = np.empty(nnz, np.int32)
        data = np.empty(nnz, np.float64)
        check(lib().cn_context_stiffness(self.handle, ptr(indptr), ptr(indices), ptr(data)),
              "cn_context_stiffness")
        return indptr, indices, data

    def reduced(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        nnz = int(self.info.nnz_reduced)
        p = np.empty(self.n_nodes, np.int32)
        c = np.empty(nnz, np.int32)
        v = np.empty(nnz, np.float64)
        check(lib().cn_context_reduced(self.handle, ptr(p), ptr(c), ptr(v)), "cn_context_reduced")
        return p, c, v

    def level(self, i: int) -> dict[str, np.ndarray | int]:
        n = ctypes.c_int64()
        nnz = ctypes.c_int64()
        nc = ctypes.c_int64()
        pn = ctypes.c_int64()
        check(lib().cn_context_level(self.handle, i, ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(nc),
                                     ctypes.byref(pn), *([None] * 8)), "cn_context_level")
        n, nnz, nc, pn = n.value, nnz.value, nc.value, pn.value
        out = {"n": n, "nnz": nnz, "n_coarse": nc, "p_nnz": pn,
               "a_ptr": np.empty(n + 1, np.int32), "a_col": np.empty(nnz, np.int32),
               "a_val": np.empty(nnz, np.float32)}
        coarse = i == int(self.info.n_levels)
        if not coarse:
            out.update({"dinv": np.empty(n, np.float32), "agg": np.empty(n, np.int32),
                        "p_ptr": np.empty(n + 1, np.int32), "p_col": np.empty(pn, np.int32),
                        "p_val": np.empty(pn, np.float32)})
        check(lib().cn_context_level(
            self.handle, i, None, None, None, None, ptr(out["a_ptr"]), ptr(out["a_col"]), ptr(out["a_val"]),
            ptr(out.get("dinv")), ptr(out.get("agg")), ptr(out.get("p_ptr")), ptr(out.get("p_col")),
            ptr(out.get("p_val"))), "cn_context_level")
        return out

    def skin_normals(self) -> np.ndarray:
        out = np.empty((self.n_skin, 3), np.float64)
        check(lib().cn_context_skin_normals(self.handle, ptr(out)), "cn_context_skin_normals")
        return out

    # ---- kernels on the context ----
    def spmm(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        K = X.shape[1]
        Y = np.empty_like(X)
        check(lib().cn_k_spmm(self.handle, ptr(X), ptr(Y), K), "cn_k_spmm")
        return Y

    def spmm_diff(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.empty_like(X)
        check(lib().cn_k_spmm_diff(self.handle, ptr(X), ptr(Y), X.shape[1]), "cn_k_spmm_diff")
        return Y

    def gram(self, A: np.ndarray, B1: np.ndarray, B2: np.ndarray | None = None):
        """FP64 A^T B1 (and A^T B2) of (n_reduced, K) FP32 blocks, by the solver's Gram kernels."""
        A = np.ascontiguousarray(A, dtype=np.float32)
        B1 = np.ascontiguousarray(B1, dtype=np.float32)
        k = A.shape[1]
        if B1.shape != A.shape or (B2 is not None and np.shape(B2) != A.shape):
            raise ValueError("A, B1 and B2 must have the same shape")
        out = np.empty((1 if B2 is None else 2, k, k), np.float64)
        if B2 is not None:
            B2 = np.ascontiguousarray(B2, dtype=np.float32)
        check(lib().cn_k_gram(self.handle, ptr(A), ptr(B1), ptr(B2), ptr(out), k),
              "cn_k_gram")
        return out[0] if B2 is None else (out[0], out[1])

    def vcycle(self, B: np.ndarray) -> np.ndarray:
        B = np.ascontiguousarray(B, dtype=np.float32)
        X = np.empty_like(B)
        check(lib().cn_k_vcycle(self.handle, ptr(B), ptr(X), B.shape[1]), "cn_k_vcycle")
        return X

    def place(self, centers, handles, dists) -> tuple[np.ndarray, np.ndarray]:
        c = np.ascontiguousarray(centers, np.float64).reshape(-1, 3)
        h = None if handles is None else np.ascontiguousarray(handles, np.float64).reshape(-1, 3)
        d = np.ascontiguousarray(dists, np.float64).reshape(-1)
        out = np.empty((c.shape[0], 4, 4), np.float64)
        deg = np.empty(c.shape[0], np.int32)
        check(lib().cn_k_place(self.handle, ptr(c), ptr(h), ptr(d), c.shape[0], ptr(out), ptr(deg)), "cn_k_place")
        return out, deg

    # ---- reduced-basis warm start ----
    def warm_reset(self, capacity: int) -> None:
        check(lib().cn_warm_reset(self.handle, int(capacity)), "cn_warm_reset")

    def warm_count(self) -> int:
        c = ctypes.c_int32()
        check(lib().cn_warm_count(self.handle, ctypes.byref(c)), "cn_warm_count")
        return int(c.value)

    def warm_gram(self) -> np.ndarray:
        """G = W^T A W of the collected basis (FP64)."""
        s = self.warm_count()
        g = np.empty((s, s), np.float64)
        check(lib().cn_warm_gram(self.handle, ptr(g)), "cn_warm_gram")
        return g

    def warm_finalize(self, rcond: float = 1e-7, max_modes: int = 256) -> int:
        """Modes of the collected snapshots: the leading eigenpairs (V, lambda) of the Galerkin
        matrix G = W^T A W, at most max_modes and above rcond x the largest (near-duplicate
        snapshots drop out), give the A-orthonormal modes Phi = W V lambda^-1/2 (formed on the
        device in BF16); P = (Phi^T A Phi)^-1 removes their BF16 rounding. Returns the mode count."""
        g = self.warm_gram()
        g = 0.5 * (g + g.T)
        lam, v = np.linalg.eigh(g)
        order = np.argsort(lam)[::-1]
        lam, v = lam[order], v[:, order]
        r = int(min(max_modes, (lam > rcond * lam[0]).sum()))
        m = np.ascontiguousarray(v[:, :r] / np.sqrt(lam[:r]))
        check(lib().cn_warm_set_modes(self.handle, ptr(m), m.shape[0], r), "cn_warm_set_modes")
        gpp = np.empty((r, r), np.float64)
        gpw = np.empty((r, m.shape[0]), np.float64)
        check(lib().cn_warm_mode_grams(self.handle, ptr(gpp), ptr(gpw)), "cn_warm_mode_grams")
        p = np.ascontiguousarray(np.linalg.inv(0.5 * (gpp + gpp.T)))
        check(lib().cn_warm_set_projector(self.handle, ptr(p), r), "cn_warm_set_projector")
        self.warm_state = {"modes": r, "P": p, "G_modes": gpp, "G_cross": gpw}
        return r

    def warm_targets(self) -> np.ndarray:
        """Galerkin coefficients of every snapshot in the modes, (count, R): P Phi^T A W."""
        st = self.warm_state
        return np.ascontiguousarray((st["P"] @ st["G_cross"]).T)

    def warm_set_coefficients(self, c: np.ndarray) -> None:
        c = np.ascontiguousarray(c, np.float64)
        check(lib().cn_warm_set_coefficients(self.handle, ptr(c), c.shape[0], c.shape[1]),
              "cn_warm_set_coefficients")

    def warm_project_last(self, k: int) -> np.ndarray:
        """Phi^T b for the first k placements of the most recent block solve, (k, R)."""
        r = self.warm_state["modes"]
        out = np.empty((k, r), np.float64)
        check(lib().cn_warm_project_last(self.handle, int(k), ptr(out)), "cn_warm_project_last")
        return out

    def warm_guess_errors(self, k: int, energy: np.ndarray, coefficients: np.ndarray | None = None) -> np.ndarray:
        """Energy-norm relative error ||x - x0||_A / ||x||_A of the initial guesses of the most
        recent block's first k placements, given ||x||_A^2 = x^T b per placement: the Galerkin
        guess Phi P Phi^T b, or Phi c for supplied coefficients c (k, R). With g = Phi^T b and
        G = Phi^T A Phi, ||x - Phi c||_A^2 = x^T b - 2 c^T g + c^T G c."""
        st = self.warm_state
        g = self.warm_project_last(k)
        c = g @ st["P"].T if coefficients is None else np.asarray(coefficients, np.float64)
        gm = 0.5 * (st["G_modes"] + st["G_modes"].T)
        err2 = energy - 2.0 * np.einsum("ij,ij->i", c, g) + np.einsum("ij,jk,ik->i", c, gm, c)
        return np.sqrt(np.maximum(err2, 0.0) / energy)

    def warm_modes(self) -> np.ndarray:
        """The modes (n_reduced, width): hi + lo of their BF16 pairs, in FP64."""
        w = ctypes.c_int32()
        check(lib().cn_warm_modes(self.handle, None, ctypes.byref(w)), "cn_warm_modes")
        bits = np.empty((2, self.n_nodes - 1, w.value), np.uint16)
        check(lib().cn_warm_modes(self.handle, ptr(bits), ctypes.byref(w)), "cn_warm_modes")
        halves = (bits.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
        return halves[0] + halves[1]

    # ---- conductivity uncertainty ----
    def uq_prepare(self, dip_pos_m, dip_moment, site, didt: float) -> list[int]:
        """Per-tissue operators and RHS of one placement; returns the tissue tags (the column order
        of every sigma passed to uq_solve)."""
        pos = np.ascontiguousarray(dip_pos_m, np.float64).reshape(-1, 3)
        mom = np.ascontiguousarray(dip_moment, np.float64).reshape(-1, 3)
        c = np.ascontiguousarray(site[0], np.float64).reshape(3)
        h = None if site[1] is None else np.ascontiguousarray(site[1], np.float64).reshape(3)
        nt = ctypes.c_int32()
        tags = np.zeros(16, np.int32)
        check(lib().cn_uq_prepare(self.handle, ptr(pos), ptr(mom), pos.shape[0], ptr(c), ptr(h), float(site[2]),
                                  float(didt), ctypes.byref(nt), ptr(tags)), "cn_uq_prepare")
        self.uq_tags = [int(t) for t in tags[:nt.value]]
        return self.uq_tags

    def uq_solve(self, sigma: np.ndarray, *, magnitude: bool = True, vectors: bool = False, potential: bool = False,
                 metrics: bool = True, tolerance: float = DEFAULT_TOLERANCE,
                 max_iters: int = DEFAULT_MAX_ITERS, initial_guess: bool = True,
                 recovery: str = "raw") -> list[PlacementResult]:
        """k <= 16 conductivity samples (k x n_tissue, columns in uq_tags order). With
        ``initial_guess`` each draw starts from the nominal solution plus its Galerkin correction
        on the sensitivity span (uq/conductivity/run.py:154-200), else from zero. ``recovery``
        post-processes each draw's field before |E|, the metrics and the moments are taken."""
        if recovery not in _lib.RECOVERY_CODES:
            raise ValueError(f"unknown recovery {recovery!r}")
        sg = np.ascontiguousarray(sigma, np.float64)
        if sg.ndim != 2 or sg.shape[1] != len(self.uq_tags):
            raise ValueError("sigma must be (k, n_tissue) in the order uq_prepare returned")
        k, m, n = sg.shape[0], self.n_tets, self.n_nodes
        out = _lib.BlockOutputs()
        tf = np.empty((k, 4, 4), np.float64)
        it = np.empty(k, np.int32)
        rel = np.empty(k, np.float64)
        mag = np.empty((k, m), np.float32) if magnitude else None
        E = np.empty((k, m, 3), np.float32) if vectors else None
        v = np.empty((k, n), np.float64) if potential else None
        met = np.empty((k, _lib.METRICS_LEN), np.float64) if metrics else None
        out.transforms, out.iterations, out.relative_residual = ptr(tf), ptr(it), ptr(rel)
        out.magnE, out.E, out.v, out.metrics = ptr(mag), ptr(E), ptr(v), ptr(met)
        opt = _lib.default_solve_options()
        opt.tolerance = float(tolerance)
        opt.max_iters = int(max_iters)
        opt.warm_start = 1 if initial_guess else 0
        opt.recovery = _lib.RECOVERY_CODES[recovery]
        check(lib().cn_uq_solve(self.handle, ptr(sg), k, ctypes.byref(opt), ctypes.byref(out)), "cn_uq_solve")
        return [PlacementResult(transform=tf[i], magnE=None if mag is None else mag[i],
                                E=None if E is None else E[i], v=None if v is None else v[i],
                                metrics=None if met is None else met[i], iterations=int(it[i]),
                                relative_residual=float(rel[i])) for i in range(k)]

    def uq_guess_info(self) -> dict:
        """Iterations of the nominal and sensitivity solves behind the draws' initial guess, and
        the rank of its correction."""
        out = np.zeros(3, np.int32)
        check(lib().cn_uq_guess_info(self.handle, ptr(out)), "cn_uq_guess_info")
        return {"nominal_iterations": int(out[0]), "sensitivity_iterations": int(out[1]), "rank": int(out[2])}

    def uq_moments(self) -> tuple[np.ndarray, np.ndarray, int]:
        cnt = ctypes.c_int64()
        mean = np.empty(self.n_tets, np.float64)
        var = np.empty(self.n_tets, np.float64)
        check(lib().cn_uq_moments(self.handle, ptr(mean), ptr(var), ctypes.byref(cnt)), "cn_uq_moments")
        return mean, var, int(cnt.value)

    def recovery_build(self, mode: str) -> tuple[int, int]:
        """Build (once) the recovery operator of ``mode``; returns (n_slots, n_fallback)."""
        code = _lib.RECOVERY_CODES[mode]
        n_slots = ctypes.c_int64()
        n_fb = ctypes.c_int32()
        check(lib().cn_recovery_build(self.handle, code, ctypes.byref(n_slots), ctypes.byref(n_fb)),
              "cn_recovery_build")
        return int(n_slots.value), int(n_fb.value)

    def recovery_operator(self, mode: str) -> dict[str, np.ndarray]:
        """The recovery operator's arrays (built on first use), for inspection and tests."""
        n_slots, n_fb = self.recovery_build(mode)
        nc = 4 * self.n_tets
        spr = mode != "harmonic"
        out = {
            "ptr": np.empty(n_slots + 1, np.int32), "idx": np.empty(nc, np.int32),
            "w": np.empty(nc, np.float32) if spr else None,
            "slot_of_corner": np.empty(nc, np.int32), "slot_node": np.empty(n_slots, np.int32),
            "slot_tag": np.empty(n_slots, np.int32) if mode != "spr_global" else None,
            "boundary": np.empty(self.n_nodes, np.int32) if spr else None, "status": np.empty(n_slots, np.int32),
        }
        check(lib().cn_recovery_operator(self.handle, _lib.RECOVERY_CODES[mode], ptr(out["ptr"]), ptr(out["idx"]),
                                         ptr(out["w"]), ptr(out["slot_of_corner"]), ptr(out["slot_node"]),
                                         ptr(out["slot_tag"]), ptr(out["boundary"]), ptr(out["status"])),
              "cn_recovery_operator")
        out["n_fallback"] = n_fb
        if not spr:
            nnz = ctypes.c_int64()
            und = ctypes.c_int32()
            check(lib().cn_recovery_patches(self.handle, ctypes.byref(nnz), ctypes.byref(und), None, None, None),
                  "cn_recovery_patches")
            out["pptr"] = np.empty(n_slots + 1, np.int32)
            out["pidx"] = np.empty(nnz.value, np.int32)
            out["w3"] = np.empty((nnz.value, 3), np.float32)
            check(lib().cn_recovery_patches(self.handle, None, None, ptr(out["pptr"]), ptr(out["pidx"]),
                                            ptr(out["w3"])), "cn_recovery_patches")
            out["n_undetermined"] = int(und.value)
        return out

    @property
    def tuning(self) -> dict:
        """Smoother degree, Chebyshev ratio, fine-level SpMM tile and the coarser levels' tile
        (sub-lanes, 0 = automatic)."""
        t = _lib.Tuning()
        check(lib().cn_context_get_tuning(self.handle, ctypes.byref(t)), "cn_context_get_tuning")
        return {"degree": int(t.degree), "lower_ratio": float(t.lower_ratio), "fine_sub": int(t.fine_sub),
                "coarse_sub": int(t.coarse_sub), "coil_blocks": int(t.coil_blocks), "layout": int(t.layout),
                "dense_tail": int(t.dense_tail)}

    def set_tuning(self, degree: int | None = None, lower_ratio: float | None = None,
                   fine_sub: int | None = None, coarse_sub: int | None = None,
                   coil_blocks: int | None = None, layout: int | None = None,
                   dense_tail: int | None = None) -> None:
        """Change the settings of `tuning` between solves (None keeps the current value).
        ``layout`` 0 runs the finest level's row kernels on the sliced ELL copy (when its rows
        fit) for blocks of up to 4 columns and on the CSR arrays above, 1 always on the CSR
        arrays, 2 always on the ELL copy; results are the same. ``dense_tail`` 1 applies the
        V-cycle's bottom levels as one dense product formed at the first solve, 0 level by level."""
        cur = self.tuning
        t = _lib.Tuning(cur["degree"] if degree is None else int(degree),
                        cur["lower_ratio"] if lower_ratio is None else float(lower_ratio),
                        cur["fine_sub"] if fine_sub is None else int(fine_sub),
                        cur["coarse_sub"] if coarse_sub is None else int(coarse_sub),
                        cur["coil_blocks"] if coil_blocks is None else int(coil_blocks),
                        cur["layout"] if layout is None else int(layout),
                        cur["dense_tail"] if dense_tail is None else int(dense_tail))
        check(lib().cn_context_set_tuning(self.handle, ctypes.byref(t)), "cn_context_set_tuning")

    def bench_spmm(self, K: int, reps: int = 20, mode: int = 1, sub: int = 0) -> float:
        """Average ms of the fine-level product: mode 0 plain SpMM, 1 differenced SpMM, 2 with Grams."""
        ms = ctypes.c_double()
        check(lib().cn_bench_spmm(self.handle, K, reps, mode, sub, ctypes.byref(ms)), "cn_bench_spmm")
        return float(ms.value)

    def bench_vcycle(self, K: int, reps: int = 10) -> float:
        ms = ctypes.c_double()
        check(lib().cn_bench_vcycle(self.handle, K, reps, ctypes.byref(ms)), "cn_bench_vcycle")
        return float(ms.value)

    def timing_enable(self, on: bool = True) -> None:
        """Time the SpMM, the V-cycle and each iteration of later solves on the device clock."""
        check(lib().cn_timing_enable(self.handle, 1 if on else 0), "cn_timing_enable")

    def timing_read(self) -> dict:
        """Samples and mean ms of the in-solve differenced SpMM, V-cycle and iteration."""
        out = np.zeros(6)
        check(lib().cn_timing_read(self.handle, ptr(out)), "cn_timing_read")
        return {"samples": int(out[0]), "spmm_ms": float(out[1]), "vcycle_ms": float(out[2]),
                "iteration_ms": float(out[3]), "dadt_placements": int(out[4]), "dadt_ms": float(out[5])}

    def bench_guess(self, K: int = 16, reps: int = 5) -> dict:
        """Average ms of the warm-start guess's projection, coefficients and expansion at width K,
        on the right-hand sides of the last block solved at that width."""
        ms = np.zeros(3)
        check(lib().cn_bench_guess(self.handle, K, reps, ptr(ms)), "cn_bench_guess")
        return {"project_ms": float(ms[0]), "coefficients_ms": float(ms[1]), "expand_ms": float(ms[2])}

    PHASES = ("frames", "dadt", "rhs", "guess", "pcg", "field", "outputs")

    def timing_phases(self) -> dict:
        """Device ms per block of each phase of the blocks solved while timing is on."""
        out = np.zeros(9)
        check(lib().cn_timing_phases(self.handle, ptr(out)), "cn_timing_phases")
        blocks = int(out[7])
        per = {p: (float(out[i]) / blocks if blocks else 0.0) for i, p in enumerate(self.PHASES)}
        return {"blocks": blocks, "placements": int(out[8]), "ms_per_block": per}

    def bench_dadt(self, dip_pos_m, dip_moment, site, k: int = 8, reps: int = 5) -> float:
        """Average ms of the dA/dt N-body for k copies of one placement (repeated launches)."""
        pos = np.ascontiguousarray(dip_pos_m, np.float64).reshape(-1, 3)
        mom = np.ascontiguousarray(dip_moment, np.float64).reshape(-1, 3)
        c = np.ascontiguousarray(site[0], np.float64).reshape(3)
        h = np.ascontiguousarray(site[1], np.float64).reshape(3)
        ms = ctypes.c_double()
        check(lib().cn_bench_dadt(self.handle, ptr(pos), ptr(mom), pos.shape[0], ptr(c), ptr(h), float(site[2]), k,
                                  reps, ctypes.byref(ms)), "cn_bench_dadt")
        return float(ms.value)

    def bench_copy(self, nbytes: int, reps: int = 20) -> float:
        """Average ms of a streaming device copy of `nbytes` bytes (read and write)."""
        ms = ctypes.c_double()
        check(lib().cn_bench_copy(self.handle, int(nbytes), reps, ctypes.byref(ms)), "cn_bench_copy")
        return float(ms.value)

    def bench_stages(self, dip_pos_m, dip_moment, site, k: int = 16, recovery: str = "harmonic",
                     reps: int = 5) -> dict:
        """Average ms of the block stages around the solve for k copies of one placement: the coil
        stage (dA/dt, element pass, RHS), the field on a zero potential with `recovery`, and the
        region metrics (the block PCG does not run)."""
        pos = np.ascontiguousarray(dip_pos_m, np.float64).reshape(-1, 3)
        mom = np.ascontiguousarray(dip_moment, np.float64).reshape(-1, 3)
        c = np.ascontiguousarray(site[0], np.float64).reshape(3)
        h = np.ascontiguousarray(site[1], np.float64).reshape(3)
        ms = np.zeros(3)
        check(lib().cn_bench_stages(self.handle, ptr(pos), ptr(mom), pos.shape[0], ptr(c), ptr(h), float(site[2]),
                                    int(k), _lib.RECOVERY_CODES[recovery], int(reps), ptr(ms)), "cn_bench_stages")
        return {"coil_ms": float(ms[0]), "field_ms": float(ms[1]), "metrics_ms": float(ms[2])}

    def setup_times(self) -> dict:
        """Host wall time (ms) of the context build's phases."""
        ms = np.zeros(10)
        check(lib().cn_context_setup_times(self.handle, ptr(ms)), "cn_context_setup_times")
        names = ("checks", "upload", "order", "gradients_checks", "pattern", "stiffness_grounding", "hierarchy",
                 "solver_ell", "skin_region", "last_recovery_build")
        return {k: round(float(v), 3) for k, v in zip(names, ms)}

    def debug_fail_captures(self, n: int) -> None:
        """Test hook: the next `n` captures of the block iteration report failure, so the solves
        that need them launch their iterations directly."""
        check(lib().cn_debug_fail_captures(self.handle, int(n)), "cn_debug_fail_captures")

    def graph_execs(self) -> int:
        """Captured block-iteration executables alive in the solver (in use and retired)."""
        out = ctypes.c_int32()
        check(lib().cn_graph_execs(self.handle, ctypes.byref(out)), "cn_graph_execs")
        return int(out.value)


def solve_sweep(
    ctx: SolverContext,
    dip_pos_m: np.ndarray,
    dip_moment: np.ndarray,
    sites: Sequence[tuple[np.ndarray, np.ndarray, float]],
    didt: float,
    block_k: int,
    *,
    tolerance: float = DEFAULT_TOLERANCE,
    max_iters: int = DEFAULT_MAX_ITERS,
    use_graph: bool = True,
    recovery: str = "raw",
    warm_start: bool = False,
    recycle: int = 2,
) -> dict:
    """Every placement of `sites` in blocks of `block_k`, summaries only: the next block's dA/dt
    and right-hand side are computed while the current block solves. Returns "transforms"
    (n, 4, 4), "metrics" (n, METRICS_LEN), "iterations" and "relative_residual" (n) and
    "block_iterations" (per block). Each full block starts from the Galerkin projection of its
    solution onto the span of the previous `recycle` full blocks' solutions (at most 2; 0 starts
    every block from zero, and then each block's numbers are those of `solve_placements_block` on
    the same placements)."""
    _blocks()
    n_place = len(sites)
    if n_place < 1:
        raise ValueError("solve_sweep: no placements")
    if block_k < 1 or block_k > MAX_BLOCK:
        raise ValueError(f"solve_sweep: block_k={block_k} must be in [1, {MAX_BLOCK}]")
    if recovery not in _lib.RECOVERY_CODES:
        raise ValueError(f"unknown recovery {recovery!r}")
    pos = np.ascontiguousarray(dip_pos_m, np.float64).reshape(-1, 3)
    mom = np.ascontiguousarray(dip_moment, np.float64).reshape(-1, 3)
    if pos.shape != mom.shape:
        raise ValueError("dipole positions and moments must have the same shape")
    centers = np.ascontiguousarray([np.asarray(s[0], np.float64).reshape(3) for s in sites])
    handles = np.ascontiguousarray([np.asarray(s[1], np.float64).reshape(3) for s in sites])
    dists = np.ascontiguousarray([float(s[2]) for s in sites], np.float64)
    n_blocks = (n_place + block_k - 1) // block_k
    tf = np.empty((n_place, 4, 4), np.float64)
    met = np.empty((n_place, _lib.METRICS_LEN), np.float64)
    it = np.empty(n_place, np.int32)
    rel = np.empty(n_place, np.float64)
    bit = np.empty(n_blocks, np.int32)
    opt = _lib.default_solve_options()
    opt.recovery = _lib.RECOVERY_CODES[recovery]
    opt.warm_start = int(bool(warm_start))
    opt.tolerance = float(tolerance)
    opt.max_iters = int(max_iters)
    opt.use_graph = int(use_graph)
    if recycle < 0:
        raise ValueError(f"solve_sweep: recycle={recycle} must be >= 0")
    opt.recycle = int(recycle)
    check(lib().cn_solve_sweep(ctx.handle, ptr(pos), ptr(mom), pos.shape[0], ptr(centers), ptr(handles), ptr(dists),
                               n_place, int(block_k), float(didt), ctypes.byref(opt), ptr(tf), ptr(met), ptr(it),
                               ptr(rel), ptr(bit)), "cn_solve_sweep")
    return {"transforms": tf, "metrics": met, "iterations": it, "relative_residual": rel, "block_iterations": bit}


def solve_placements_block(
    ctx: SolverContext,
    dip_pos_m: np.ndarray,
    dip_moment: np.ndarray,
    sites: Sequence[tuple[np.ndarray, np.ndarray, float]],
    didt: float,
    *,
    magnitude: bool = True,
    vectors: bool = False,
    potential: bool = False,
    metrics: bool = True,
    tolerance: float = DEFAULT_TOLERANCE,
    max_iters: int = DEFAULT_MAX_ITERS,
    x0: np.ndarray | None = None,
    extra: dict | None = None,
    use_graph: bool = True,
    recovery: str = "raw",
    slots: bool = False,
    warm_start: bool = False,
    collect: bool = False,
    sigma: np.ndarray | None = None,
    independent: bool = False,
) -> list[PlacementResult]:
    """Solve up to MAX_BLOCK placements as one block (fem/solve.py:865-979). ``independent`` runs
    the block as k independent CG chains in lockstep, as the reference's solver does, instead of
    the true block method. ``recovery`` selects
    the field E, |E| and the metrics describe ("raw", "spr_global", "spr_tissue", "harmonic");
    ``slots`` also returns the recovered field on its slots. ``sigma`` (k x n_tags, S/m by tag,
    see `conductivity_rows`) gives every placement its own conductivities; the block then runs as
    independent CG chains, each on its own operator, under the context's AMG hierarchy."""
    _blocks()
    k = len(sites)
    if k < 1 or k > MAX_BLOCK:
        raise ValueError(f"solve_placements_block: k={k} must be in [1, {MAX_BLOCK}]")
    pos = np.ascontiguousarray(dip_pos_m, np.float64).reshape(-1, 3)
    mom = np.ascontiguousarray(dip_moment, np.float64).reshape(-1, 3)
    if pos.shape != mom.shape:
        raise ValueError("dipole positions and moments must have the same shape")
    centers = np.ascontiguousarray([np.asarray(s[0], np.float64).reshape(3) for s in sites])
    handles = np.ascontiguousarray([np.asarray(s[1], np.float64).reshape(3) for s in sites])
    dists = np.ascontiguousarray([float(s[2]) for s in sites], np.float64)
    m, n = ctx.n_tets, ctx.n_nodes
    out = _lib.BlockOutputs()
    tf = np.empty((k, 4, 4), np.float64)
    it = np.empty(k, np.int32)
    rel = np.empty(k, np.float64)
    mag = np.empty((k, m), np.float32) if magnitude else None
    E = np.empty((k, m, 3), np.float32) if vectors else None
    v = np.empty((k, n), np.float64) if potential else None
    met = np.empty((k, _lib.METRICS_LEN), np.float64) if metrics else None
    out.transforms = ptr(tf)
    out.iterations = ptr(it)
    out.relative_residual = ptr(rel)
    out.magnE = ptr(mag)
    out.E = ptr(E)
    out.v = ptr(v)
    out.metrics = ptr(met)
    ex = {}
    if extra is not None:
        for key, shape, dt in (("dadt_elm", (k, m, 3), np.float32), ("dadt_nodes", (k, n, 3), np.float32),
                               ("rhs", (k, n), np.float32), ("x0", (k, n - 1), np.float64)):
            if key in extra:
                ex[key] = np.empty(shape, dt)
                setattr(out, key, ptr(ex[key]))
    if recovery not in _lib.RECOVERY_CODES:
        raise ValueError(f"unknown recovery {recovery!r}")
    if slots and recovery == "raw":
        raise ValueError("slots=True needs a recovery mode")
    e_slots = None
    if slots:
        n_slots, _ = ctx.recovery_build(recovery)
        e_slots = np.empty((k, n_slots, 3), np.float32)
        out.E_slots = ptr(e_slots)
    opt = _lib.default_solve_options()
    opt.recovery = _lib.RECOVERY_CODES[recovery]
    opt.warm_start = int(warm_start)  # True: Galerkin projection; 2: coefficients set beforehand
    opt.collect = int(collect)
    opt.tolerance = float(tolerance)
    opt.max_iters = int(max_iters)
    opt.use_graph = int(use_graph)
    opt.independent = int(independent)
    sig = None
    if sigma is not None:
        sig = np.ascontiguousarray(sigma, np.float64)
        if sig.ndim != 2 or sig.shape[0] != k:
            raise ValueError("sigma must be (k, n_tags): one conductivity row per placement")
        opt.sigma = ptr(sig)
        opt.n_sigma = int(sig.shape[1])
    x0a = None
    if x0 is not None:
        x0a = np.ascontiguousarray(x0, np.float64).reshape(k, n - 1)
    status = lib().cn_solve_block(ctx.handle, ptr(pos), ptr(mom), pos.shape[0], ptr(centers), ptr(handles),
                                  ptr(dists), k, float(didt), ctypes.byref(opt), ptr(x0a), ctypes.byref(out))
    check(status, "cn_solve_block")
    if extra is not None:
        extra.update(ex)
        extra["total_iterations"] = int(out.total_iterations)
        extra["refinements"] = int(out.refinements)
        nr = min(int(out.refinements), 8)
        extra["round_iterations"] = [int(out.round_iterations[i]) for i in range(nr)]
        extra["round_residual"] = [float(out.round_residual[i]) for i in range(nr)]
    return [
        PlacementResult(
            transform=tf[i], magnE=None if mag is None else mag[i], E=None if E is None else E[i],
            v=None if v is None else v[i], metrics=None if met is None else met[i],
            iterations=int(it[i]), relative_residual=float(rel[i]),
            E_slots=None if e_slots is None else e_slots[i],
        )
        for i in range(k)
    ]


def conductivity_rows(overrides: Sequence[dict | None], table: dict | None = None) -> np.ndarray:
    """(k, 11) conductivity-by-tag rows for ``solve_placements_block(sigma=...)``: the table
    (fem/assembly.py:15-26) with each placement's {tag: S/m} overrides applied."""
    base = TISSUE_CONDUCTIVITY if table is None else table
    width = max(base) + 1
    rows = np.full((len(overrides), width), np.nan)
    for i, ov in enumerate(overrides):
        for t, v in base.items():
            rows[i, t] = v
        for t, v in (ov or {}).items():
            if not 0 <= int(t) < width:
                raise ValueError(f"tag {t} is outside the conductivity table")
            rows[i, int(t)] = float(v)
    return rows


def metrics_dict(rec: np.ndarray, region: str = "gray_matter") -> dict:
    """A metric record (CN_METRICS_LEN doubles) as the reference's FieldMetrics (metrics.py:28-37)."""
    return {
        "region": region,
        "peak_magnE": float(rec[0]),
        "peak_location_mm": np.array(rec[1:4], dtype=np.float64),
        "center_of_gravity_mm": np.array(rec[4:7], dtype=np.float64),
        "region_volume_m3": float(rec[7]),
        "focality_m3": {"0.5": float(rec[8]), "0.75": float(rec[9])},
        "distribution": {"mean": float(rec[10]), "std": float(rec[11]), "p50": float(rec[12]),
                         "p95": float(rec[13]), "p99": float(rec[14]), "p99.9": float(rec[15])},
    }


def dadt_nbody(s: np.ndarray, mp: np.ndarray, sn: np.ndarray, r: np.ndarray, didt: float,
               mu0_4pi: float = MU0_OVER_4PI) -> np.ndarray:
    """Raw dipole N-body kernel (bindings.cpp:194-206 dadt_nbody)."""
    s = np.ascontiguousarray(s, np.float32)
    mp = np.ascontiguousarray(mp, np.float32)
    sn = np.ascontiguousarray(sn, np.float32)
    r = np.ascontiguousarray(r, np.float32)
    out = np.empty_like(r)
    check(lib().cn_k_dadt(ptr(s), ptr(mp), ptr(sn), s.shape[0], ptr(r), r.shape[0], float(didt), float(mu0_4pi),
                          ptr(out)), "cn_k_dadt")
    return out


def p1_gradients(nodes_mm: np.ndarray, tet_nodes: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    nodes = np.ascontiguousarray(nodes_mm, np.float64)
    tets = np.ascontiguousarray(tet_nodes, np.int32)
    g = np.empty((tets.shape[0], 4, 3), np.float64)
    v = np.empty(tets.shape[0], np.float64)
    check(lib().cn_k_p1_gradients(ptr(nodes), nodes.shape[0], ptr(tets), tets.shape[0], ptr(g), ptr(v)),
          "cn_k_p1_gradients")
    return g, v


def morton_keys(points_mm: np.ndarray) -> np.ndarray:
    pts = np.ascontiguousarray(points_mm, np.float64)
    keys = np.empty(pts.shape[0], np.uint64)
    check(lib().cn_k_morton_keys(ptr(pts), pts.shape[0], ptr(keys)), "cn_k_morton_keys")
    return keys


def node2corner(elems: np.ndarray, n_nodes: int) -> tuple[np.ndarray, np.ndarray]:
    e = np.ascontiguousarray(elems, np.int32)
    ptr_ = np.empty(n_nodes + 1, np.int32)
    idx = np.empty(e.size, np.int32)
    check(lib().cn_k_node2corner(ptr(e), e.shape[0], e.shape[1], n_nodes, ptr(ptr_), ptr(idx)), "cn_k_node2corner")
    return ptr_, idx


def l1_dinv(n, a_ptr, a_col, a_val) -> np.ndarray:
    out = np.empty(n, np.float32)
    check(lib().cn_k_l1_dinv(n, ptr(np.ascontiguousarray(a_ptr, np.int32)), ptr(np.ascontiguousarray(a_col, np.int32)),
                             ptr(np.ascontiguousarray(a_val, np.float32)), ptr(out)), "cn_k_l1_dinv")
    return out


def select_size4(n, a_ptr, a_col, a_val) -> tuple[np.ndarray, int]:
    agg = np.empty(n, np.int32)
    nc = ctypes.c_int64()
    check(lib().cn_k_select_size4(n, ptr(np.ascontiguousarray(a_ptr, np.int32)),
                                  ptr(np.ascontiguousarray(a_col, np.int32)),
                                  ptr(np.ascontiguousarray(a_val, np.float32)), ptr(agg), ctypes.byref(nc)),
          "cn_k_select_size4")
    return agg, int(nc.value)


def spgemm(a_ptr, a_col, a_val, b_ptr, b_col, b_val, b_cols: int):
    a_ptr = np.ascontiguousarray(a_ptr, np.int32)
    a_col = np.ascontiguousarray(a_col, np.int32)
    a_val = np.ascontiguousarray(a_val, np.float32)
    b_ptr = np.ascontiguousarray(b_ptr, np.int32)
    b_col = np.ascontiguousarray(b_col, np.int32)
    b_val = np.ascontiguousarray(b_val, np.float32)
    n = a_ptr.shape[0] - 1
    c_ptr = np.empty(n + 1, np.int32)
    nnz = ctypes.c_int64()
    args = (n, b_ptr.shape[0] - 1, b_cols, ptr(a_ptr), ptr(a_col), ptr(a_val), ptr(b_ptr), ptr(b_col), ptr(b_val))
    check(lib().cn_k_spgemm(*args, ptr(c_ptr), None, None, ctypes.byref(nnz)), "cn_k_spgemm")
    c_col = np.empty(nnz.value, np.int32)
    c_val = np.empty(nnz.value, np.float32)
    check(lib().cn_k_spgemm(*args, ptr(c_ptr), ptr(c_col), ptr(c_val), ctypes.byref(nnz)), "cn_k_spgemm")
    return c_ptr, c_col, c_val
