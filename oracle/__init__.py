"""CPU oracle for the TMS E-field FEM path -- TEST INFRASTRUCTURE, never the product.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import this
package, and only to check the HIP library's results (or to time the CPU baseline). The product
path (``cunibs_amd``) never imports it and fails loudly when its HIP library is missing.

It restates the reference's algorithm on the CPU: the numerical kernels in ``oracle.c`` (built by
``oracle/Makefile``), the orchestration here in NumPy. Each function cites the reference lines it
follows. FP64 throughout except where the reference fixes FP32 (the AMG hierarchy's values), so a
solve to 1e-12 here is the FP64 oracle that GPU results are measured against.

Parity pins (what checks each piece, beyond the GPU comparison itself) are listed in the test
modules; pieces with none are marked "parity unpinned" in their docstrings.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "liboracle.so"

TISSUE_CONDUCTIVITY: dict[int, float] = {
    1: 0.126, 2: 0.275, 3: 1.654, 4: 0.010, 5: 0.465,
    6: 0.5, 7: 0.008, 8: 0.025, 9: 0.6, 10: 0.16,
}  # fem/assembly.py:15-26
GM_TAG = 2
MU0_OVER_4PI = 1e-7  # fem/placement.py:16 (line 178 of the module text)
MIN_COARSE_ROWS = 128  # fem/solve.py:146
MAX_LEVELS = 50  # fem/solve.py:147
ROUNDS = 2  # fem/solve.py:150
SA_DAMPING = np.float32(4.0 / 3.0)  # fem/solve.py:91
CHEB_DEGREE = 2  # fem/solve.py:167
CHEB_LOWER_RATIO = 12.0  # fem/solve.py:168


def build(force: bool = False) -> Path:
    if force or not _LIB_PATH.exists() or _LIB_PATH.stat().st_mtime < (_HERE / "oracle.c").stat().st_mtime:
        subprocess.run(["make", "-s", "-C", str(_HERE), "-B" if force else "liboracle.so"],
                       check=True)
    return _LIB_PATH


_lib: ctypes.CDLL | None = None

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_D = ctypes.c_double


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(str(_LIB_PATH))
        sig = {
            "or_morton_keys": (None, [_P, _I64, _P]),
            "or_p1_gradients": (None, [_P, _P, _I64, _P, _P]),
            "or_pattern_count": (_I64, [_P, _P, _P, _I64, _P]),
            "or_pattern_fill": (None, [_P, _P, _P, _I64, _P, _P]),
            "or_stiffness_values": (None, [_P, _P, _P, _P, _P, _I64, _P, _P, _P]),
            "or_l1_dinv": (None, [_I64, _P, _P, _P, _P]),
            "or_agg_edge_weights": (None, [_I64, _P, _P, _P, _P]),
            "or_select_size4": (_I64, [_I64, _P, _P, _P, _P]),
            "or_spgemm_count": (_I64, [_I64, _I64, _P, _P, _P, _P, _P]),
            "or_spgemm_fill": (None, [_I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
            "or_vcycle_apply": (None, [_P, _P, _P, _P]),
            "or_vcycle_work_size": (_I64, [_P]),
            "or_pcg": (ctypes.c_int32, [_P, _I64, _P, _P, _P, _P, _P, _D, ctypes.c_int32, _P]),
            "or_spmv_d": (None, [_I64, _P, _P, _P, _P, _P]),
            "or_dadt": (None, [_I64, _P, _P, _I64, _P, _D, _D, _P]),
            "or_place": (None, [_P, _P, _P, _I64, _P, _P, _P, _P, _I64, _P, _P]),
            "or_skin_normals": (None, [_P, _P, _I64, _I64, _P, _P, _P]),
            "or_rhs": (None, [_P, _P, _P, _P, _P, _I64, _P]),
            "or_reconstruct": (None, [_P, _P, _P, _P, _I64, _P, _P]),
            "or_num_threads": (ctypes.c_int32, []),
            "or_set_threads": (None, [ctypes.c_int32]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a: np.ndarray | None) -> int | None:
    if a is None:
        return None
    assert a.flags.c_contiguous
    return a.ctypes.data


def _c(a, dtype) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=dtype)


# ------------------------------------------------------------------------------------------------
# Mesh layout and assembly
# ------------------------------------------------------------------------------------------------

def morton_keys(points_mm: np.ndarray) -> np.ndarray:
    """63-bit Morton keys, fem/solve.py:299-323."""
    pts = _c(points_mm, np.float64)
    keys = np.empty(pts.shape[0], dtype=np.uint64)
    lib().or_morton_keys(_ptr(pts), pts.shape[0], _ptr(keys))
    return keys


def spatial_order(nodes_mm: np.ndarray, tet_nodes: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Node perm = stable argsort of Morton keys; tet perm = stable argsort of the lowest renumbered
    node. fem/solve.py:333-348, solver/fem/reorder.cu:14-24."""
    node_perm = np.argsort(morton_keys(nodes_mm), kind="stable").astype(np.int32)
    inverse = np.empty_like(node_perm)
    inverse[node_perm] = np.arange(node_perm.shape[0], dtype=np.int32)
    lowest = inverse[tet_nodes].min(axis=1)
    tet_perm = np.argsort(lowest, kind="stable").astype(np.int32)
    return node_perm, tet_perm


def reorder_mesh(nodes_mm, tet_nodes, tet_tags, skin_tris):
    """fem/solve.py:645-662 (_spatially_ordered)."""
    node_perm, tet_perm = spatial_order(nodes_mm, tet_nodes)
    inverse = np.empty_like(node_perm)
    inverse[node_perm] = np.arange(node_perm.shape[0], dtype=np.int32)
    nodes = _c(nodes_mm[node_perm], np.float64)
    tets = _c(inverse[tet_nodes[tet_perm]], np.int32)
    tags = _c(tet_tags[tet_perm], np.int32)
    skin = _c(inverse[skin_tris], np.int32) if len(skin_tris) else np.empty((0, 3), np.int32)
    return nodes, tets, tags, skin, node_perm, tet_perm


def p1_gradients(nodes_mm: np.ndarray, tet_nodes: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """solver/fem/gradient.cu:29-66."""
    nodes = _c(nodes_mm, np.float64)
    tets = _c(tet_nodes, np.int32)
    m = tets.shape[0]
    g = np.empty((m, 4, 3), dtype=np.float64)
    vols = np.empty(m, dtype=np.float64)
    lib().or_p1_gradients(_ptr(nodes), _ptr(tets), m, _ptr(g), _ptr(vols))
    return g, vols


def conductivity_per_tet(tet_tags: np.ndarray, table: dict[int, float] | None = None) -> np.ndarray:
    """fem/assembly.py:31-42: tag -> sigma, unknown tags raise."""
    table = TISSUE_CONDUCTIVITY if table is None else table
    max_tag = max(table)
    lut = np.full(max_tag + 1, np.nan)
    for t, s in table.items():
        lut[t] = s
    cond = lut[np.clip(tet_tags, 0, max_tag)]
    unknown = np.isnan(cond) | (tet_tags < 0) | (tet_tags > max_tag)
    if unknown.any():
        raise ValueError(f"Tetrahedra with unknown tags: {np.unique(tet_tags[unknown])}")
    return cond


def node2corner(tet_nodes: np.ndarray, n_nodes: int) -> tuple[np.ndarray, np.ndarray]:
    """fem/assembly.py:142-162: stable argsort of the flattened connectivity."""
    keys = np.ascontiguousarray(tet_nodes).ravel()
    ptr = np.zeros(n_nodes + 1, dtype=np.int32)
    if keys.size == 0:
        return ptr, np.empty(0, np.int32)
    order = np.argsort(keys, kind="stable").astype(np.int32)
    ptr[1:] = np.cumsum(np.bincount(keys, minlength=n_nodes)).astype(np.int32)
    return ptr, order


def stiffness_pattern(tet_nodes, ptr, idx, n):
    """solver/fem/pattern.cu:65-87,219-299: sorted distinct incident nodes per row."""
    tets = _c(tet_nodes, np.int32)
    indptr = np.empty(n + 1, dtype=np.int32)
    nnz = lib().or_pattern_count(_ptr(tets), _ptr(ptr), _ptr(idx), n, _ptr(indptr))
    indices = np.empty(nnz, dtype=np.int32)
    lib().or_pattern_fill(_ptr(tets), _ptr(ptr), _ptr(idx), n, _ptr(indptr), _ptr(indices))
    return indptr, indices


def stiffness_values(g, scale, tet_nodes, ptr, idx, indptr, indices):
    """solver/fem/stiffness.cu:32-66."""
    n = indptr.shape[0] - 1
    data = np.empty(indices.shape[0], dtype=np.float64)
    lib().or_stiffness_values(_ptr(_c(g, np.float64)), _ptr(_c(scale, np.float64)),
                              _ptr(_c(tet_nodes, np.int32)), _ptr(ptr), _ptr(idx), n,
                              _ptr(indptr), _ptr(indices), _ptr(data))
    return data


def ground_node_of(nodes_mm: np.ndarray) -> int:
    """fem/solve.py:287-289: lowest z, lowest index on ties."""
    return int(np.argmin(nodes_mm[:, 2]))


def reduce_csr(indptr, indices, data, ground):
    """fem/solve.py:292-296,361-365: drop row and column `ground`; columns above it shift down."""
    n = indptr.shape[0] - 1
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(indptr))
    keep = (rows != ground) & (indices != ground)
    r = rows[keep]
    r = r - (r > ground)
    c = indices[keep].astype(np.int64)
    c = c - (c > ground)
    new_ptr = np.zeros(n, dtype=np.int32)
    np.add.at(new_ptr, r + 1, 1) if n > 1 else None
    new_ptr = np.cumsum(new_ptr).astype(np.int32)
    return new_ptr, c.astype(np.int32), np.ascontiguousarray(data[keep])


# ------------------------------------------------------------------------------------------------
# Smoothed-aggregation hierarchy (FP32 values, fixed-order products)
# ------------------------------------------------------------------------------------------------

@dataclass
class Csr:
    ptr: np.ndarray
    col: np.ndarray
    val: np.ndarray
    n_cols: int

    @property
    def n(self) -> int:
        return self.ptr.shape[0] - 1

    def dense(self) -> np.ndarray:
        out = np.zeros((self.n, self.n_cols), dtype=np.float64)
        rows = np.repeat(np.arange(self.n), np.diff(self.ptr))
        out[rows, self.col] = self.val
        return out


def l1_dinv(a: Csr) -> np.ndarray:
    """solver/amg/l1.cu:15-35."""
    out = np.empty(a.n, dtype=np.float32)
    lib().or_l1_dinv(a.n, _ptr(a.ptr), _ptr(a.col), _ptr(_c(a.val, np.float32)), _ptr(out))
    return out


def edge_weights(a: Csr) -> np.ndarray:
    """solver/amg/aggregate.cu:108-134."""
    w = np.empty(a.col.shape[0], dtype=np.float32)
    lib().or_agg_edge_weights(a.n, _ptr(a.ptr), _ptr(a.col), _ptr(_c(a.val, np.float32)), _ptr(w))
    return w


def select_size4(a: Csr) -> tuple[np.ndarray, int]:
    """solver/amg/aggregate.cu:368-472: one SIZE_4 pass; returns (aggregate per row, count)."""
    agg = np.empty(a.n, dtype=np.int32)
    nc = lib().or_select_size4(a.n, _ptr(a.ptr), _ptr(a.col), _ptr(_c(a.val, np.float32)), _ptr(agg))
    return agg, int(nc)


def spgemm(a: Csr, b: Csr) -> Csr:
    """C = A B, FP32, fixed association (fmaf folds in A-row order), sorted rows.
    Stands in for cupyx SpGEMM (fem/solve.py:119-133); values: parity unpinned against the
    reference (its association order is unspecified), pinned against dense FP64 products."""
    cp_ = np.empty(a.n + 1, dtype=np.int32)
    nnz = lib().or_spgemm_count(a.n, b.n_cols, _ptr(a.ptr), _ptr(a.col), _ptr(b.ptr), _ptr(b.col),
                                _ptr(cp_))
    cc = np.empty(nnz, dtype=np.int32)
    cv = np.empty(nnz, dtype=np.float32)
    lib().or_spgemm_fill(a.n, b.n_cols, _ptr(a.ptr), _ptr(a.col), _ptr(_c(a.val, np.float32)),
                         _ptr(b.ptr), _ptr(b.col), _ptr(_c(b.val, np.float32)), _ptr(cp_),
                         _ptr(cc), _ptr(cv))
    return Csr(cp_, cc, cv, b.n_cols)


def transpose(a: Csr) -> Csr:
    """Counting-sort transpose; each output row lists its columns (the input rows) ascending."""
    counts = np.bincount(a.col, minlength=a.n_cols)
    ptr = np.zeros(a.n_cols + 1, dtype=np.int32)
    ptr[1:] = np.cumsum(counts)
    rows = np.repeat(np.arange(a.n, dtype=np.int32), np.diff(a.ptr))
    order = np.argsort(a.col, kind="stable")
    return Csr(ptr, rows[order].astype(np.int32), np.ascontiguousarray(a.val[order]), a.n)


def tentative_prolongator(agg: np.ndarray, n_coarse: int) -> Csr:
    """fem/solve.py:94-107: one entry per row, 1/sqrt(aggregate size)."""
    counts = np.bincount(agg, minlength=n_coarse).astype(np.float32)
    scale = np.float32(1.0) / np.sqrt(counts)
    n = agg.shape[0]
    return Csr(np.arange(n + 1, dtype=np.int32), agg.astype(np.int32), scale[agg].astype(np.float32),
               n_coarse)


def smooth_prolongator(a: Csr, p_tent: Csr, dinv: np.ndarray) -> Csr:
    """fem/solve.py:110-126: P = P_tent - (4/3) D^-1 A P_tent, FP32."""
    ap = spgemm(a, p_tent)
    rows = np.repeat(np.arange(ap.n), np.diff(ap.ptr))
    scaled = (ap.val * dinv[rows]).astype(np.float32)
    t = (SA_DAMPING * scaled).astype(np.float32)
    base = np.zeros_like(t)
    own = ap.col == p_tent.col[rows]
    base[own] = p_tent.val[rows[own]]
    return Csr(ap.ptr, ap.col, (base - t).astype(np.float32), ap.n_cols)


def galerkin(a: Csr, p: Csr) -> Csr:
    """fem/solve.py:129-134: A_c = P^T A P, here associated as P^T (A P), each product
    fixed-order FP32. (The reference evaluates (P^T A) P; the association only changes rounding,
    and this one keeps every intermediate row narrow.)"""
    return spgemm(transpose(p), spgemm(a, p))


def aggregate_level(a: Csr) -> tuple[np.ndarray | None, int]:
    """fem/solve.py:178-209 (_aggregate): compose ROUNDS pairwise passes."""
    cur = a
    agg = None
    n_coarse = a.n
    for round_i in range(ROUNDS):
        if cur.n <= MIN_COARSE_ROWS:
            break
        step, nc = select_size4(cur)
        if nc == cur.n or nc < MIN_COARSE_ROWS:
            break
        agg = step if agg is None else step[agg]
        n_coarse = nc
        if round_i + 1 < ROUNDS:
            cur = galerkin(cur, tentative_prolongator(step, nc))
    return agg, n_coarse


@dataclass
class Level:
    a: Csr
    agg: np.ndarray
    dinv: np.ndarray
    p: Csr
    r: Csr


@dataclass
class Hierarchy:
    levels: list[Level]
    coarse: Csr
    coarse_l: np.ndarray  # FP64 lower Cholesky factor
    cheb: dict = field(default_factory=dict)
    _keep: list = field(default_factory=list, repr=False)

    @property
    def sizes(self) -> list[int]:
        return [lv.a.n for lv in self.levels] + [self.coarse.n]


def chebyshev_coefficients(degree: int = CHEB_DEGREE, lower_ratio: float = CHEB_LOWER_RATIO) -> dict:
    """solver/amg/vcycle.cu:434-463, computed in FP32 as there."""
    f = np.float32
    hi = f(1.0)
    lo = f(hi / f(lower_ratio))
    theta = f(f(0.5) * f(hi + lo))
    delta = f(f(0.5) * f(hi - lo))
    sigma = f(theta / delta)
    alpha0 = f(f(1.0) / theta)
    rho = f(f(1.0) / sigma)
    c_cur, c_prev, alpha = [], [], []
    for _ in range(degree - 1):
        rho_next = f(f(1.0) / f(f(2.0) * sigma - rho))
        beta = f(rho_next * rho)
        c_cur.append(f(f(1.0) + beta))
        c_prev.append(f(-beta))
        alpha.append(f(f(f(2.0) * rho_next) / delta))
        rho = rho_next
    return {"degree": degree, "alpha0": alpha0, "c_cur": np.array(c_cur, np.float32),
            "c_prev": np.array(c_prev, np.float32), "alpha": np.array(alpha, np.float32)}


def build_hierarchy(a: Csr, degree: int = CHEB_DEGREE, lower_ratio: float = CHEB_LOWER_RATIO) -> Hierarchy:
    """fem/solve.py:224-256 (aggregation_levels) + :259-284 (build_native_vcycle)."""
    levels: list[Level] = []
    a = Csr(a.ptr, a.col, a.val.astype(np.float32), a.n_cols)
    for _ in range(MAX_LEVELS):
        if a.n <= MIN_COARSE_ROWS:
            break
        agg, nc = aggregate_level(a)
        if agg is None or nc == a.n or nc < MIN_COARSE_ROWS:
            break
        dinv = l1_dinv(a)
        p = smooth_prolongator(a, tentative_prolongator(agg, nc), dinv)
        levels.append(Level(a, agg, dinv, p, transpose(p)))
        a = galerkin(a, p)
    coarse_l = np.linalg.cholesky(a.dense())
    return Hierarchy(levels, a, np.ascontiguousarray(coarse_l), chebyshev_coefficients(degree, lower_ratio))


class _CHier(ctypes.Structure):
    _fields_ = [
        ("n_levels", ctypes.c_int32), ("n", _P),
        ("a_ptr", _P), ("a_col", _P), ("a_val", _P), ("dinv", _P),
        ("p_ptr", _P), ("p_col", _P), ("p_val", _P),
        ("r_ptr", _P), ("r_col", _P), ("r_val", _P),
        ("n_coarse", ctypes.c_int32), ("coarse_l", _P),
        ("degree", ctypes.c_int32), ("alpha0", _D),
        ("c_cur", _P), ("c_prev", _P), ("alpha", _P),
    ]


def _c_hier(h: Hierarchy) -> _CHier:
    keep = []

    def arr(xs):
        a = (ctypes.c_void_p * max(1, len(xs)))(*[x.ctypes.data for x in xs])
        keep.append(a)
        keep.extend(xs)
        return ctypes.cast(a, _P)

    def vec(x):
        keep.append(x)
        return x.ctypes.data

    lv = h.levels
    c = _CHier()
    c.n_levels = len(lv)
    c.n = vec(np.array([l.a.n for l in lv] or [0], dtype=np.int32))
    c.a_ptr = arr([l.a.ptr for l in lv]); c.a_col = arr([l.a.col for l in lv])
    c.a_val = arr([l.a.val for l in lv]); c.dinv = arr([l.dinv for l in lv])
    c.p_ptr = arr([l.p.ptr for l in lv]); c.p_col = arr([l.p.col for l in lv])
    c.p_val = arr([l.p.val for l in lv])
    c.r_ptr = arr([l.r.ptr for l in lv]); c.r_col = arr([l.r.col for l in lv])
    c.r_val = arr([l.r.val for l in lv])
    c.n_coarse = h.coarse.n
    c.coarse_l = vec(h.coarse_l)
    ch = h.cheb
    c.degree = ch["degree"]
    c.alpha0 = float(ch["alpha0"])
    c.c_cur = vec(np.ascontiguousarray(np.r_[ch["c_cur"], 0.0], dtype=np.float64))
    c.c_prev = vec(np.ascontiguousarray(np.r_[ch["c_prev"], 0.0], dtype=np.float64))
    c.alpha = vec(np.ascontiguousarray(np.r_[ch["alpha"], 0.0], dtype=np.float64))
    h._keep[:] = keep
    return c


def vcycle_apply(h: Hierarchy, b: np.ndarray) -> np.ndarray:
    """One zero-guess V-cycle, FP64 arithmetic, solver/amg/vcycle.cu:560-648."""
    c = _c_hier(h)
    b = _c(b, np.float64)
    x = np.empty_like(b)
    work = np.empty(lib().or_vcycle_work_size(ctypes.byref(c)), dtype=np.float64)
    lib().or_vcycle_apply(ctypes.byref(c), _ptr(b), _ptr(x), _ptr(work))
    return x


def pcg(h: Hierarchy, a: Csr, a_val64: np.ndarray, b: np.ndarray, tol: float, max_iters: int = 2000,
        x0: np.ndarray | None = None) -> tuple[np.ndarray, int, float]:
    """FP64 PCG with the AMG V-cycle, solver/amg/solver.cpp:200-318 (to ||r||/||b|| <= tol)."""
    c = _c_hier(h)
    b = _c(b, np.float64)
    x = np.zeros_like(b) if x0 is None else _c(x0, np.float64).copy()
    rel = np.zeros(1)
    it = lib().or_pcg(ctypes.byref(c), a.n, _ptr(a.ptr), _ptr(a.col), _ptr(_c(a_val64, np.float64)),
                      _ptr(b), _ptr(x), float(tol), int(max_iters), _ptr(rel))
    return x, int(it), float(rel[0])


def spmv(a: Csr, val64: np.ndarray, x: np.ndarray) -> np.ndarray:
    y = np.empty(a.n, dtype=np.float64)
    lib().or_spmv_d(a.n, _ptr(a.ptr), _ptr(a.col), _ptr(_c(val64, np.float64)), _ptr(_c(x, np.float64)),
                    _ptr(y))
    return y


# ------------------------------------------------------------------------------------------------
# Coil side, RHS, reconstruction
# ------------------------------------------------------------------------------------------------

def skin_normals(nodes_mm, skin_tris):
    """solver/fem/skin.cu:46-150 (two smoothing rounds)."""
    tris = _c(skin_tris, np.int32)
    out = np.empty((tris.shape[0], 3), dtype=np.float64)
    if tris.shape[0] == 0:
        return out
    n = nodes_mm.shape[0]
    ptr, idx = node2corner(tris, n)
    lib().or_skin_normals(_ptr(_c(nodes_mm, np.float64)), _ptr(tris), tris.shape[0], n, _ptr(ptr),
                          _ptr(idx), _ptr(out))
    return out


def place_transforms(centers, handles, dists, nodes_mm, skin_tris, normals):
    """solver/coil/place.cu:122-230; returns ((P,4,4) transforms, degenerate flags)."""
    c = _c(np.reshape(centers, (-1, 3)), np.float64)
    h = None if handles is None else _c(np.reshape(handles, (-1, 3)), np.float64)
    d = _c(np.reshape(dists, (-1,)), np.float64)
    ta = _c(nodes_mm[skin_tris[:, 0]], np.float64)
    tb = _c(nodes_mm[skin_tris[:, 1]], np.float64)
    tc = _c(nodes_mm[skin_tris[:, 2]], np.float64)
    out = np.empty((c.shape[0], 16), dtype=np.float64)
    deg = np.empty(c.shape[0], dtype=np.int32)
    lib().or_place(_ptr(c), _ptr(h), _ptr(d), c.shape[0], _ptr(ta), _ptr(tb), _ptr(tc),
                   _ptr(_c(normals, np.float64)), skin_tris.shape[0], _ptr(out), _ptr(deg))
    return out.reshape(-1, 4, 4), deg


def dadt_at_nodes(positions_m, moments, transform, didt, nodes_mm):
    """fem/placement.py:279-317: transform the dipoles, then the closed-form sum in FP64."""
    rot = transform[:3, :3]
    trans = transform[:3, 3]
    s = (np.asarray(positions_m) * 1e3 @ rot.T + trans) * 1e-3
    m = np.asarray(moments) @ rot.T
    o = trans * 1e-3
    s = _c(s - o, np.float64)
    r = _c(nodes_mm * 1e-3 - o, np.float64)
    out = np.empty_like(r)
    lib().or_dadt(s.shape[0], _ptr(s), _ptr(_c(m, np.float64)), r.shape[0], _ptr(r), float(didt),
                  MU0_OVER_4PI, _ptr(out))
    return out


def dadt_direct(s, m, r, didt):
    s = _c(s, np.float64); m = _c(m, np.float64); r = _c(r, np.float64)
    out = np.empty_like(r)
    lib().or_dadt(s.shape[0], _ptr(s), _ptr(m), r.shape[0], _ptr(r), float(didt), MU0_OVER_4PI,
                  _ptr(out))
    return out


def rhs(dadt_elm, g, neg_vc, ptr, idx):
    """solver/fem/rhs.cu:13-104 in FP64."""
    n = ptr.shape[0] - 1
    b = np.empty(n, dtype=np.float64)
    lib().or_rhs(_ptr(_c(dadt_elm, np.float64)), _ptr(_c(g, np.float64)), _ptr(_c(neg_vc, np.float64)),
                 _ptr(ptr), _ptr(idx), n, _ptr(b))
    return b


def reconstruct(v, tet_nodes, g, dadt_elm):
    """solver/fem/reconstruct.cu:12-31 in FP64."""
    m = tet_nodes.shape[0]
    e = np.empty((m, 3), dtype=np.float64)
    mag = np.empty(m, dtype=np.float64)
    lib().or_reconstruct(_ptr(_c(v, np.float64)), _ptr(_c(tet_nodes, np.int32)), _ptr(_c(g, np.float64)),
                         _ptr(_c(dadt_elm, np.float64)), m, _ptr(e), _ptr(mag))
    return e, mag


# ------------------------------------------------------------------------------------------------
# Metrics (metrics.py restated in NumPy)
# ------------------------------------------------------------------------------------------------

SUMMARY_PERCENTILES = (50.0, 95.0, 99.0, 99.9)
SUMMARY_FOCALITY_FRACS = (0.5, 0.75)


def weighted_quantiles(values, weights, qs):
    """metrics.py:82-95 (NumPy path: sequential cumsum)."""
    order = np.argsort(values, kind="stable")
    v = values[order]
    w = np.asarray(weights, dtype=np.float64)[order]
    cw = np.cumsum(w)
    pos = (cw - 0.5 * w) / cw[-1]
    return np.interp(np.asarray(qs, dtype=pos.dtype), pos, v)


def summarize(m, w, bary, region="gray_matter", volume_m3=None):
    """metrics.py:121-135,209-243 (RegionSlice.summarize) on a gathered region."""
    m = np.asarray(m)
    w = np.asarray(w)
    wsum = w.sum()
    mean = float((w * m).sum() / wsum)
    var = float((w * (m - mean) ** 2).sum() / wsum)
    qs = np.asarray([p / 100.0 for p in SUMMARY_PERCENTILES], dtype=m.dtype)
    pv = weighted_quantiles(m, w, qs)
    dist = {"mean": mean, "std": float(np.sqrt(var))}
    dist.update({f"p{p:g}": float(v) for p, v in zip(SUMMARY_PERCENTILES, pv)})
    anchor = dist["p99.9"]
    ww = w * m
    return {
        "region": region,
        "peak_magnE": float(m.max()),
        "peak_location_mm": np.asarray(bary[np.argmax(m)]),
        "center_of_gravity_mm": (ww[:, None] * bary).sum(0) / ww.sum(),
        "region_volume_m3": float(w.sum()) if volume_m3 is None else volume_m3,
        "focality_m3": {f"{f:g}": float(w[m >= f * anchor].sum()) for f in SUMMARY_FOCALITY_FRACS},
        "distribution": dist,
    }


# ------------------------------------------------------------------------------------------------
# End to end
# ------------------------------------------------------------------------------------------------

@dataclass
class OracleContext:
    nodes_mm: np.ndarray
    tet_nodes: np.ndarray
    tet_tags: np.ndarray
    skin_tris: np.ndarray
    g: np.ndarray
    vols: np.ndarray
    cond: np.ndarray
    n2c_ptr: np.ndarray
    n2c_idx: np.ndarray
    indptr: np.ndarray
    indices: np.ndarray
    data: np.ndarray
    ground: int
    reduced: Csr
    reduced_val64: np.ndarray
    hier: Hierarchy | None
    skin_normals: np.ndarray
    node_perm: np.ndarray
    tet_perm: np.ndarray


def build_context(nodes_mm, tet_nodes, tet_tags, skin_tris, *, hierarchy: bool = True,
                  conductivity: dict[int, float] | None = None) -> OracleContext:
    """fem/solve.py:690-746 (build_context), FP64 assembly."""
    nodes, tets, tags, skin, node_perm, tet_perm = reorder_mesh(
        np.asarray(nodes_mm, np.float64), np.asarray(tet_nodes, np.int32),
        np.asarray(tet_tags, np.int32), np.asarray(skin_tris, np.int32).reshape(-1, 3))
    n = nodes.shape[0]
    g, vols = p1_gradients(nodes, tets)
    bad = (vols <= 0.0) | ~np.isfinite(vols)
    if bad.any():
        raise ValueError(f"{int(bad.sum())} tetrahedra have zero or non-finite volume")
    cond = conductivity_per_tet(tags, conductivity)
    ptr, idx = node2corner(tets, n)
    indptr, indices = stiffness_pattern(tets, ptr, idx, n)
    data = stiffness_values(g, vols * cond, tets, ptr, idx, indptr, indices)
    ground = ground_node_of(nodes)
    rp, rc, rv = reduce_csr(indptr, indices, data, ground)
    red = Csr(rp, rc, rv.astype(np.float32), n - 1)
    hier = build_hierarchy(red) if hierarchy else None
    return OracleContext(nodes, tets, tags, skin, g, vols, cond, ptr, idx, indptr, indices, data,
                         ground, red, rv, hier, skin_normals(nodes, skin), node_perm, tet_perm)


def solve_site(ctx: OracleContext, positions_m, moments, site, didt, tol=1e-12, max_iters=5000):
    """One placement end to end (fem/solve.py:801-862, raw recovery), FP64."""
    center, handle, dist = site
    tf, deg = place_transforms(center, handle, [dist], ctx.nodes_mm, ctx.skin_tris, ctx.skin_normals)
    if deg[0]:
        raise ValueError("degenerate coil placement")
    tf = tf[0]
    dadt_nodes = dadt_at_nodes(positions_m, moments, tf, didt, ctx.nodes_mm)
    dadt_elm = dadt_nodes[ctx.tet_nodes].mean(axis=1)
    neg_vc = -(ctx.vols * ctx.cond)
    b = rhs(dadt_elm, ctx.g, neg_vc, ctx.n2c_ptr, ctx.n2c_idx)
    keep = np.ones(ctx.nodes_mm.shape[0], dtype=bool)
    keep[ctx.ground] = False
    b_red = b[keep]
    x, iters, rel = pcg(ctx.hier, ctx.reduced, ctx.reduced_val64, b_red, tol, max_iters)
    v = np.zeros(ctx.nodes_mm.shape[0])
    v[keep] = x
    e, mag = reconstruct(v, ctx.tet_nodes, ctx.g, dadt_elm)
    return {"transform": tf, "v": v, "E": e, "magnE": mag, "dadt_elm": dadt_elm, "b": b,
            "iterations": iters, "relative_residual": rel}


def gm_summary(ctx: OracleContext, magnE: np.ndarray, region_tag: int = GM_TAG):
    idx = np.nonzero(ctx.tet_tags == region_tag)[0]
    bary = ctx.nodes_mm[ctx.tet_nodes[idx]].mean(axis=1)
    w = ctx.vols[idx]
    return summarize(magnE[idx], w, bary, volume_m3=float(w.sum()))


def set_threads(n: int) -> None:
    lib().or_set_threads(int(n))


def num_threads() -> int:
    return int(lib().or_num_threads())


# ------------------------------------------------------------------------------------------------
# Superconvergent patch recovery (fem/recovery.py:120-453, solver/fem/recovery.cu:67-280,620-687)
# ------------------------------------------------------------------------------------------------

SPR_MAX_LEBESGUE = 8.0  # solver/fem/recovery.cu:36


def outer_boundary_nodes(tet_nodes: np.ndarray, n_nodes: int) -> np.ndarray:
    """Nodes of faces owned by exactly one tetrahedron (fem/recovery.py:120-141)."""
    t = np.asarray(tet_nodes, np.int64)
    faces = np.concatenate([t[:, [1, 2, 3]], t[:, [0, 2, 3]], t[:, [0, 1, 3]], t[:, [0, 1, 2]]])
    faces = np.sort(faces, axis=1)
    uniq, counts = np.unique(faces, axis=0, return_counts=True)
    mark = np.zeros(n_nodes, dtype=np.int32)
    mark[uniq[counts == 1].ravel()] = 1
    return mark


def tissue_slots(tet_nodes: np.ndarray, tet_tags: np.ndarray):
    """Corners grouped by (node, tag), stable (fem/recovery.py:187-228)."""
    t = np.asarray(tet_nodes, np.int64).ravel()
    tags = np.repeat(np.asarray(tet_tags, np.int64), 4)
    stride = int(tags.max()) + 1
    key = t * stride + tags
    order = np.argsort(key, kind="stable")
    sk = key[order]
    head = np.r_[True, sk[1:] != sk[:-1]]
    starts = np.flatnonzero(head)
    slot_sorted = np.cumsum(head) - 1
    ptr = np.r_[starts, sk.size].astype(np.int32)
    soc = np.empty(sk.size, np.int32)
    soc[order] = slot_sorted
    return (ptr, order.astype(np.int32), soc, (sk[starts] // stride).astype(np.int32),
            (sk[starts] % stride).astype(np.int32))


def spr_operator(nodes_mm, tet_nodes, tet_tags, vols, mode: str) -> dict:
    """One weight per corner: the linear fit over the slot's incident barycentres in the
    node-centred, radius-scaled basis, evaluated at the node; the volume-weighted average on the
    patch volume's boundary, for fewer than four elements, a pivot <= 1e-12 A_00 or a Lebesgue
    constant above 8 (solver/fem/recovery.cu:192-280)."""
    nodes = np.asarray(nodes_mm, np.float64)
    tets = np.asarray(tet_nodes, np.int32)
    n, m = nodes.shape[0], tets.shape[0]
    boundary = outer_boundary_nodes(tets, n)
    if mode == "spr_global":
        ptr, idx = node2corner(tets, n)
        soc = tets.ravel().astype(np.int32)
        slot_node = np.arange(n, dtype=np.int32)
        slot_tag = None
    elif mode == "spr_tissue":
        ptr, idx, soc, slot_node, slot_tag = tissue_slots(tets, tet_tags)
        rep = slot_node[:-1][slot_node[1:] == slot_node[:-1]]
        boundary[rep] = 1
    else:
        raise ValueError(f"not an SPR mode: {mode}")
    n_slots = slot_node.shape[0]
    counts = np.diff(ptr)
    owner = np.repeat(np.arange(n_slots), counts)
    e = idx >> 2
    p = nodes[tets[e]]
    b = (((p[:, 0] + p[:, 1]) + p[:, 2]) + p[:, 3]) * 0.25
    xn = nodes[slot_node[owner]]
    d = b - xn
    h = np.zeros(n_slots)
    np.maximum.at(h, owner, np.abs(d).max(axis=1))
    fit = (counts >= 4) & (boundary[slot_node] == 0) & (h > 0)
    hs = np.where(h > 0, h, 1.0)[owner]
    q = np.column_stack([np.ones(len(owner)), d[:, 0] / hs, d[:, 1] / hs, d[:, 2] / hs])
    a = np.zeros((n_slots, 4, 4))
    np.add.at(a, owner, q[:, :, None] * q[:, None, :])
    # Cholesky per slot with the pivot test, then A z = e_0.
    l = np.zeros_like(a)
    floor = 1e-12 * a[:, 0, 0]
    for j in range(4):
        dj = a[:, j, j] - (l[:, j, :j] ** 2).sum(axis=1)
        fit &= dj > floor
        l[:, j, j] = np.sqrt(np.where(dj > 0, dj, 1.0))
        for i in range(j + 1, 4):
            l[:, i, j] = (a[:, i, j] - (l[:, i, :j] * l[:, j, :j]).sum(axis=1)) / l[:, j, j]
    y = np.zeros((n_slots, 4))
    for i in range(4):
        y[:, i] = ((1.0 if i == 0 else 0.0) - (l[:, i, :i] * y[:, :i]).sum(axis=1)) / l[:, i, i]
    z = np.zeros((n_slots, 4))
    for i in range(3, -1, -1):
        z[:, i] = (y[:, i] - (l[:, i + 1:, i] * z[:, i + 1:]).sum(axis=1)) / l[:, i, i]
    w_fit = (z[owner] * q).sum(axis=1)
    leb = np.zeros(n_slots)
    np.add.at(leb, owner, np.abs(w_fit))
    fit &= leb <= SPR_MAX_LEBESGUE
    v = np.asarray(vols, np.float32).astype(np.float64)[e]
    tot = np.zeros(n_slots)
    np.add.at(tot, owner, v)
    by_vol = tot[owner] > 0
    w_avg = np.where(by_vol, v / np.where(tot[owner] > 0, tot[owner], 1.0), 1.0 / counts[owner])
    w = np.where(fit[owner], w_fit, w_avg).astype(np.float32)
    return {"ptr": ptr.astype(np.int32), "idx": idx.astype(np.int32), "w": w, "slot_of_corner": soc,
            "slot_node": slot_node, "slot_tag": slot_tag, "boundary": boundary,
            "status": (~fit).astype(np.int32)}


def spr_apply(op: dict, E: np.ndarray):
    """E*_s = sum_c w_c E[c >> 2]; E_out = mean of each element's four corner slots
    (solver/fem/recovery.cu:620-687), FP64 here."""
    counts = np.diff(op["ptr"])
    owner = np.repeat(np.arange(counts.size), counts)
    contrib = op["w"].astype(np.float64)[:, None] * np.asarray(E, np.float64)[op["idx"] >> 2]
    slots = np.zeros((counts.size, 3))
    np.add.at(slots, owner, contrib)
    soc = op["slot_of_corner"].reshape(-1, 4)
    e_out = slots[soc].mean(axis=1)
    return slots, e_out, np.linalg.norm(e_out, axis=1)


# ------------------------------------------------------------------------------------------------
# Harmonic recovery of the potential (fem/recovery.py:23-50,230-383; solver/fem/recovery.cu:290-610)
# ------------------------------------------------------------------------------------------------

HPR_MIN_PATCH_NODES = 16  # fem/recovery.py:237
HPR_MAX_AMPLIFICATION = 50.0  # solver/fem/recovery.cu:44
HPR_RUNGS = ((16, 20), (9, 9), (4, 4))  # (terms, smallest patch), widest first


def harmonic_basis(u, v, t, n_terms: int) -> np.ndarray:
    """1, u, v, t; the five quadratic and seven cubic harmonic polynomials (each solves
    Laplace's equation), truncated to n_terms."""
    cols = [np.ones_like(u), u, v, t]
    if n_terms >= 9:
        cols += [u * u - v * v, v * v - t * t, u * v, u * t, v * t]
    if n_terms >= 16:
        uu, vv, tt = u * u, v * v, t * t
        cols += [t * (2 * tt - 3 * uu - 3 * vv), u * (4 * tt - uu - vv), v * (4 * tt - uu - vv),
                 t * (uu - vv), u * v * t, u * (uu - 3 * vv), v * (3 * uu - vv)]
    return np.stack(cols, axis=-1)


def harmonic_patches(tet_nodes, tet_tags, n_nodes: int):
    """Per (node, tissue) slot: the distinct nodes of its own tetrahedra, grown to the union of its
    same-tissue neighbours' first rings when fewer than 16 (fem/recovery.py:230-295)."""
    tets = np.asarray(tet_nodes, np.int64)
    ptr, idx, soc, slot_node, slot_tag = tissue_slots(tets, tet_tags)
    ns = slot_node.shape[0]
    owner = np.repeat(np.arange(ns, dtype=np.int64), np.diff(ptr))
    pairs = np.unique(owner[:, None] * n_nodes + tets[idx >> 2])  # (slot, node) of every corner's tet
    r_slot, r_node = pairs // n_nodes, pairs % n_nodes
    rp = np.zeros(ns + 1, np.int64)
    rp[1:] = np.cumsum(np.bincount(r_slot, minlength=ns))
    stride = int(np.max(tet_tags)) + 1
    key = slot_node.astype(np.int64) * stride + slot_tag
    pp = [0]
    pi = []
    for s in range(ns):
        ring = r_node[rp[s]:rp[s + 1]]
        if ring.size >= HPR_MIN_PATCH_NODES:
            pi.append(ring)
        else:
            nb = np.searchsorted(key, ring * stride + slot_tag[s])
            pi.append(np.unique(np.concatenate([r_node[rp[b]:rp[b + 1]] for b in nb])))
        pp.append(pp[-1] + pi[-1].size)
    return (np.asarray(pp, np.int32), np.concatenate(pi).astype(np.int32), soc, slot_node, slot_tag)


def harmonic_operator(nodes_mm, tet_nodes, tet_tags) -> dict:
    """Three gradient weights per patch node: the widest harmonic rung whose normal matrix has all
    Cholesky pivots above 1e-12 A_00 and whose amplification h max_c sum|g_c| stays <= 50
    (solver/fem/recovery.cu:520-564). status: 0 cubic, 1 quadratic, 2 linear, 3 none."""
    from scipy.linalg import solve_triangular
    nodes = np.asarray(nodes_mm, np.float64) * 1e-3  # metres: E in V/m
    pptr, pidx, soc, slot_node, slot_tag = harmonic_patches(tet_nodes, tet_tags, nodes.shape[0])
    w3 = np.zeros((pidx.size, 3), np.float32)
    status = np.full(slot_node.size, 3, np.int32)
    for s in range(slot_node.size):
        b, e = pptr[s], pptr[s + 1]
        d = nodes[pidx[b:e]] - nodes[slot_node[s]]
        h = np.abs(d).max()
        if not h > 0:
            continue
        x = d / h
        for rung, (n_terms, min_patch) in enumerate(HPR_RUNGS):
            if e - b < min_patch:
                continue
            q = harmonic_basis(x[:, 0], x[:, 1], x[:, 2], n_terms)
            a = q.T @ q
            try:
                l = np.linalg.cholesky(a)
            except np.linalg.LinAlgError:
                continue
            if not (np.diag(l) ** 2 > 1e-12 * a[0, 0]).all():
                continue
            rhs = np.zeros((n_terms, 3))
            rhs[1, 0] = rhs[2, 1] = rhs[3, 2] = 1.0
            z = solve_triangular(l.T, solve_triangular(l, rhs, lower=True), lower=False)
            g = q @ z  # (patch, 3) gradient in the scaled frame
            if np.abs(g).sum(axis=0).max() <= HPR_MAX_AMPLIFICATION:
                w3[b:e] = (g / h).astype(np.float32)
                status[s] = rung
                break
    return {"pptr": pptr, "pidx": pidx, "w3": w3, "status": status, "slot_of_corner": soc,
            "slot_node": slot_node, "slot_tag": slot_tag}


def harmonic_apply(op: dict, v: np.ndarray, dadt_nodes: np.ndarray):
    """E[s] = -sum_m w[s, m] v[m] - dA/dt(x_node) (solver/fem/recovery.cu:571-610), then each
    element's mean over its four corner slots."""
    counts = np.diff(op["pptr"])
    owner = np.repeat(np.arange(counts.size), counts)
    grad = np.zeros((counts.size, 3))
    np.add.at(grad, owner, op["w3"].astype(np.float64) * np.asarray(v, np.float64)[op["pidx"], None])
    slots = -grad - np.asarray(dadt_nodes, np.float64)[op["slot_node"]]
    e_out = slots[op["slot_of_corner"].reshape(-1, 4)].mean(axis=1)
    return slots, e_out, np.linalg.norm(e_out, axis=1)


# ------------------------------------------------------------------------------------------------
# Warm-start network (FP64 restatement of cunibs_amd/csrc/mlp.hip's training arithmetic)
# ------------------------------------------------------------------------------------------------

MLP_IN_PAD = 16


def mlp_unpack(params: np.ndarray, n_hidden: int, n_out: int):
    p = np.asarray(params, np.float64)
    shapes = [(MLP_IN_PAD, n_hidden), (n_hidden, n_hidden), (n_hidden, n_out)]
    out, at = [], 0
    for r, c in shapes:
        w = p[at:at + r * c].reshape(r, c)
        at += r * c
        b = p[at:at + c]
        at += c
        out.append((w, b))
    assert at == p.size
    return out


def _gelu(x):
    from scipy.special import erf
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def _gelu_grad(x):
    from scipy.special import erf
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def mlp_forward(params, X, n_hidden: int, n_out: int):
    (w1, b1), (w2, b2), (w3, b3) = mlp_unpack(params, n_hidden, n_out)
    x = np.zeros((X.shape[0], MLP_IN_PAD))
    x[:, :X.shape[1]] = X
    z1 = x @ w1 + b1
    h1 = _gelu(z1)
    z2 = h1 @ w2 + b2
    h2 = _gelu(z2)
    return x, z1, h1, z2, h2, h2 @ w3 + b3


def mlp_loss_and_gradients(params, X, T, n_hidden: int, n_out: int):
    """Mean squared error over the batch and its gradient, in the flat parameter layout."""
    (w1, b1), (w2, b2), (w3, b3) = mlp_unpack(params, n_hidden, n_out)
    x, z1, h1, z2, h2, y = mlp_forward(params, X, n_hidden, n_out)
    d = y - T
    loss = float((d * d).mean())
    dy = 2.0 * d / d.size
    g3, gb3 = h2.T @ dy, dy.sum(0)
    dz2 = (dy @ w3.T) * _gelu_grad(z2)
    g2, gb2 = h1.T @ dz2, dz2.sum(0)
    dz1 = (dz2 @ w2.T) * _gelu_grad(z1)
    g1, gb1 = x.T @ dz1, dz1.sum(0)
    return loss, np.concatenate([g1.ravel(), gb1, g2.ravel(), gb2, g3.ravel(), gb3])
