"""The drop-in front end: ``Subject``, ``Placement``, ``FieldResult`` and ``Coil``
(src/cunibs/simulation.py:121-803 and src/cunibs/coil.py:52-98 of the reference).

``Subject.iter_simulate`` chunks placements into blocks of ``block_k`` and hands each block to the
HIP library (``cn_solve_block``): coil frames, dA/dt, RHS, block PCG, reconstruction and the
gray-matter metrics all run on the GPU; only the summaries and whichever full-volume arrays were
asked for come back. There is no CPU fallback.
"""

from __future__ import annotations

import threading
from collections.abc import Iterator, Mapping, Sequence
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from cunibs_amd import fem
from cunibs_amd.mesh import VOLUME_KEY_TO_LABEL, HeadMesh, load_mesh

RECOVERY_MODES: tuple[str, ...] = ("raw", "spr_global", "spr_tissue", "harmonic")  # fem/recovery.py:77
SUPPORTED_RECOVERY: tuple[str, ...] = RECOVERY_MODES
TISSUE_SLOT_MODES: tuple[str, ...] = ("spr_tissue", "harmonic")
REGIONS: dict[str, int] = {label: tag for tag, label in VOLUME_KEY_TO_LABEL.items()}
_DEFAULT_REGION = "gray_matter"
_SUMMARY_FOCALITY_FRACS = (0.5, 0.75)
_SUMMARY_PERCENTILES = (50.0, 95.0, 99.0, 99.9)
FOCALITY_ANCHOR_PERCENTILE = 99.9  # metrics.py
# Blocks per pipelined run when iter_simulate streams summaries only.
SWEEP_BLOCKS = 8


def _as_point(value) -> np.ndarray:
    p = np.ascontiguousarray(value, dtype=np.float64).reshape(-1)
    if p.shape != (3,):
        raise ValueError(f"Expected a 3-vector, got shape {np.shape(value)}.")
    return p


def validate_recovery(recovery: str) -> str:
    """fem/recovery.py:110-117, plus the modes this build has not implemented yet."""
    if recovery not in RECOVERY_MODES:
        raise ValueError(f"Unknown recovery {recovery!r}; use one of {', '.join(map(repr, RECOVERY_MODES))}.")
    if recovery not in SUPPORTED_RECOVERY:
        raise NotImplementedError(
            f"recovery={recovery!r} is not implemented in this build yet; use one of {SUPPORTED_RECOVERY}.")
    return recovery


@dataclass(frozen=True, init=False)
class Placement:
    """A coil placement on the scalp in mm (simulation.py:131-166): ``center_mm`` is the scalp
    target, ``handle_mm`` the positive handle direction, ``distance_mm`` the offset along the
    outward normal. Finite values, and a handle distinct from the centre, are required."""

    center_mm: np.ndarray
    handle_mm: np.ndarray
    distance_mm: float = 4.0

    def __init__(self, center_mm, handle_mm, distance_mm: float = 4.0) -> None:
        center, handle, distance = _as_point(center_mm), _as_point(handle_mm), float(distance_mm)
        if not (np.isfinite(center).all() and np.isfinite(handle).all()):
            raise ValueError("center_mm and handle_mm must be finite 3-vectors.")
        if not np.isfinite(distance):
            raise ValueError("distance_mm must be finite.")
        if np.array_equal(center, handle):
            raise ValueError("handle_mm must differ from center_mm; it defines the coil's handle direction.")
        object.__setattr__(self, "center_mm", center)
        object.__setattr__(self, "handle_mm", handle)
        object.__setattr__(self, "distance_mm", distance)


@dataclass(frozen=True)
class Coil:
    """Magnetic dipoles of a coil: positions in m, moments in A m^2 (coil.py:52-64)."""

    positions_m: np.ndarray
    moments: np.ndarray
    name: str = ""
    didt_max: float | None = None
    metadata: dict | None = None

    def __post_init__(self) -> None:
        pos = np.ascontiguousarray(self.positions_m, dtype=np.float64).reshape(-1, 3)
        mom = np.ascontiguousarray(self.moments, dtype=np.float64).reshape(-1, 3)
        if pos.shape != mom.shape or pos.shape[0] == 0:
            raise ValueError("a coil needs matching, non-empty (n, 3) positions and moments")
        if not (np.isfinite(pos).all() and np.isfinite(mom).all()):
            raise ValueError("coil dipoles must be finite")
        object.__setattr__(self, "positions_m", pos)
        object.__setattr__(self, "moments", mom)

    @classmethod
    def load(cls, source: str | Path) -> Coil:
        """An ``.h5`` coil (one ``dipoles`` dataset, n x 6 ``x y z mx my mz``, with the ``.ccd``
        header's keys as attributes; coil.py:66-83), read by :mod:`cunibs_amd.h5lite` so h5py is
        not needed, or a SimNIBS ``.ccd`` dipole file (header line ``# key=value; ...``, a count
        line, a column line, then the rows; coil.py:85-98). ``dIdtmax`` is stated in A/us and
        exposed in A/s; the raw attribute stays in ``metadata``."""
        path = Path(source)
        if path.suffix == ".h5":
            from . import h5lite
            dset = h5lite.open_file(path)["dipoles"]
            data = np.asarray(dset.read(), dtype=np.float64).reshape(-1, 6)
            attrs = {k: _decode_attr(v) for k, v in dset.attrs.items()}
        else:
            with path.open() as f:
                meta_line = f.readline().strip().lstrip("#").strip()
                count = int(f.readline().strip())
                data = np.loadtxt(f, dtype=np.float64, skiprows=1, max_rows=count).reshape(-1, 6)
            attrs = {k.strip(): v.strip() for k, v in (p.split("=", 1) for p in meta_line.split(";") if "=" in p)}
        didt = attrs.get("dIdtmax")
        return cls(data[:, 0:3], data[:, 3:6], name=str(attrs.get("coilname", path.stem)),
                   didt_max=None if didt is None else float(didt) * 1e6, metadata=attrs)


def _decode_attr(value: object) -> str | int | float:
    """An attribute as str, int or float (coil.py:38-45)."""
    if isinstance(value, bytes):
        return value.decode()
    if isinstance(value, np.generic):
        value = value.item()
    if isinstance(value, str | int | float):
        return value
    return str(value)


def _region_mask(tags: np.ndarray, region: str) -> np.ndarray:
    """metrics.py:50-55."""
    if region == "all":
        return np.ones(tags.shape[0], dtype=bool)
    if region not in REGIONS:
        raise ValueError(f"unknown region {region!r}")
    return tags == REGIONS[region]


def _weighted_quantiles(values: np.ndarray, weights: np.ndarray, qs: np.ndarray) -> np.ndarray:
    order = np.argsort(values, kind="stable")
    v = values[order]
    w = weights[order].astype(np.float64)
    cw = np.cumsum(w)
    pos = (cw - 0.5 * w) / cw[-1]
    return np.interp(np.asarray(qs, dtype=pos.dtype), pos, v)


def compute_metrics(magnE, vols, bary, tags, region: str = _DEFAULT_REGION,
                    focality_fracs=_SUMMARY_FOCALITY_FRACS, percentiles=_SUMMARY_PERCENTILES) -> dict:
    """Host-side metrics for a retained |E| (metrics.py:246-279): used for a region other than the
    one the device summarised, or for another focality fraction."""
    mask = _region_mask(tags, region)
    m = np.asarray(magnE)[mask]
    w = np.asarray(vols)[mask]
    b = np.asarray(bary)[mask]
    if m.size == 0:
        raise ValueError(f"region {region!r} has no elements")
    wsum = w.sum()
    mean = float((w * m).sum() / wsum)
    var = float((w * (m - mean) ** 2).sum() / wsum)
    qs = np.asarray([p / 100.0 for p in percentiles], dtype=m.dtype)
    pv = _weighted_quantiles(m, w, qs)
    dist = {"mean": mean, "std": float(np.sqrt(var))}
    dist.update({f"p{p:g}": float(v) for p, v in zip(percentiles, pv)})
    anchor = dist.get(f"p{FOCALITY_ANCHOR_PERCENTILE:g}")
    if anchor is None:
        anchor = float(_weighted_quantiles(m, w, np.asarray([FOCALITY_ANCHOR_PERCENTILE / 100.0], m.dtype))[0])
    ww = w * m
    return {
        "region": region,
        "peak_magnE": float(m.max()),
        "peak_location_mm": np.asarray(b[np.argmax(m)], dtype=np.float64),
        "center_of_gravity_mm": (ww[:, None] * b).sum(0) / ww.sum(),
        "region_volume_m3": float(w.astype(np.float64).sum()),
        "focality_m3": {f"{f:g}": float(w[m >= f * anchor].sum()) for f in focality_fracs},
        "distribution": dist,
    }


@dataclass
class FieldResult:
    """Metrics for one placement plus the full-volume arrays that were asked for
    (simulation.py:667-803). ``summary`` (gray matter) is always present; ``magnE``, ``E``, ``v``
    are None unless requested. Arrays are indexed in ``Subject.mesh`` order."""

    summary: dict
    magnE: np.ndarray | None
    E: np.ndarray | None
    v: np.ndarray | None
    transform: np.ndarray
    placement: Placement
    coil_name: str
    didt: float
    vols: np.ndarray | None
    tet_tags: np.ndarray | None
    barycenters_mm: np.ndarray | None
    recovery: str = "raw"
    iterations: int = 0
    relative_residual: float = 0.0
    n_nodes: int | None = None
    E_slots: np.ndarray | None = None
    slot_node: np.ndarray | None = None
    slot_tag: np.ndarray | None = None
    tet_nodes: np.ndarray | None = None
    _summaries: dict = field(default_factory=dict, repr=False)

    def nodal_field(self, region: str = _DEFAULT_REGION) -> np.ndarray:
        """The recovered field at the nodes, on ``region``'s side of any boundary, NaN where the
        region does not reach (simulation.py:709-747). Needs ``nodal=True``."""
        if self.E_slots is None or self.slot_node is None or self.n_nodes is None:
            raise ValueError("nodal_field needs the recovered nodal field, which this result did not retain. "
                             "Re-run with nodal=True.")
        if self.slot_tag is not None and region == "all":
            raise ValueError(f"nodal_field(region='all') is ambiguous for recovery={self.recovery!r}: the field is "
                             "two-valued at a tissue boundary. Name a tissue instead.")
        if self.slot_tag is not None:
            keep = self.slot_tag == REGIONS[region]
        elif region == "all":
            keep = np.ones(self.slot_node.shape[0], dtype=bool)
        else:
            if self.tet_nodes is None or self.tet_tags is None:
                raise ValueError("nodal_field for a region needs the connectivity; re-run with nodal=True and "
                                 "magnitude=True.")
            touched = np.zeros(self.n_nodes, dtype=bool)
            touched[self.tet_nodes[_region_mask(self.tet_tags, region)]] = True
            keep = touched[self.slot_node]
        out = np.full((self.n_nodes, 3), np.nan, dtype=np.float32)
        out[self.slot_node[keep]] = self.E_slots[keep]
        return out

    def _require_magnitude(self, what: str) -> np.ndarray:
        if self.magnE is None:
            raise ValueError(f"{what} needs the per-element |E|, which this result did not retain. "
                             "Re-run with magnitude=True.")
        return self.magnE

    def summary_for(self, region: str) -> dict:
        if region == self.summary["region"]:
            return self.summary
        if region not in self._summaries:
            magnE = self._require_magnitude(f"metrics for region {region!r}")
            self._summaries[region] = compute_metrics(magnE, self.vols, self.barycenters_mm, self.tet_tags, region)
        return self._summaries[region]

    def peak_magnE(self, region: str = _DEFAULT_REGION) -> float:
        return self.summary_for(region)["peak_magnE"]

    def peak_location_mm(self, region: str = _DEFAULT_REGION) -> np.ndarray:
        return self.summary_for(region)["peak_location_mm"]

    def center_of_gravity_mm(self, region: str = _DEFAULT_REGION) -> np.ndarray:
        return self.summary_for(region)["center_of_gravity_mm"]

    def focality(self, frac: float = 0.5, region: str = _DEFAULT_REGION) -> float:
        """Volume with |E| >= frac x the 99.9th percentile of |E| in the region."""
        key = f"{frac:g}"
        summary = self.summary_for(region)
        if key in summary["focality_m3"]:
            return summary["focality_m3"][key]
        magnE = self._require_magnitude(f"focality at frac={key}")
        return compute_metrics(magnE, self.vols, self.barycenters_mm, self.tet_tags, region,
                               focality_fracs=(frac,))["focality_m3"][key]

    def save(self, path: str | Path) -> None:
        """Write the result as an HDF5 file of the reference's format version 1
        (simulation.py:805-834): the retained arrays as datasets (absent ones are left out), the
        transform, a ``summary`` group, and the placement, coil, dI/dt and recovery as attributes."""
        from . import h5lite
        tree: dict = {name: getattr(self, name) for name in _SAVED_ARRAYS if getattr(self, name) is not None}
        tree["transform"] = np.asarray(self.transform, np.float64)
        tree["summary"] = _metrics_tree(self.summary)
        attrs = {"format_version": np.int64(_FORMAT_VERSION), "recovery": self.recovery}
        if self.n_nodes is not None:
            attrs["n_nodes"] = np.int64(self.n_nodes)
        attrs.update({"coil_name": self.coil_name, "didt": float(self.didt),
                      "placement_center_mm": self.placement.center_mm,
                      "placement_handle_mm": self.placement.handle_mm,
                      "placement_distance_mm": float(self.placement.distance_mm)})
        tree["__attrs__"] = attrs
        h5lite.write_file(path, tree, compression="gzip")

    @classmethod
    def load(cls, path: str | Path) -> FieldResult:
        """Read a result saved by ``save`` or by the reference (simulation.py:836-866)."""
        from . import h5lite
        f = h5lite.open_file(path)
        stored = int(f.attrs.get("format_version", 0))
        if stored != _FORMAT_VERSION:
            raise ValueError(f"{path}: file format version {stored} is not readable by this version "
                             f"(expected {_FORMAT_VERSION}).")
        opt = {name: (f[name].read() if name in f else None) for name in _SAVED_ARRAYS}
        a = f.attrs
        return cls(summary=_read_metrics_tree(f["summary"]), transform=f["transform"].read(),
                   placement=Placement(a["placement_center_mm"], a["placement_handle_mm"],
                                       float(a["placement_distance_mm"])),
                   coil_name=str(_decode_attr(a["coil_name"])), didt=float(a["didt"]),
                   recovery=str(_decode_attr(a.get("recovery", "raw"))),
                   n_nodes=int(a["n_nodes"]) if "n_nodes" in a else None, **opt)


_FORMAT_VERSION = 1
_SAVED_ARRAYS = ("magnE", "E", "v", "vols", "tet_tags", "barycenters_mm", "E_slots", "slot_node", "slot_tag",
                 "tet_nodes")


def _metrics_tree(m: dict) -> dict:
    """A FieldMetrics dict as the ``summary`` group (simulation.py:56-65)."""
    return {"__attrs__": {"region": m["region"], "peak_magnE": float(m["peak_magnE"]),
                          "region_volume_m3": float(m["region_volume_m3"])},
            "peak_location_mm": np.asarray(m["peak_location_mm"], np.float64),
            "center_of_gravity_mm": np.asarray(m["center_of_gravity_mm"], np.float64),
            "focality_m3": {"__attrs__": {k: float(v) for k, v in m["focality_m3"].items()}},
            "distribution": {"__attrs__": {k: float(v) for k, v in m["distribution"].items()}}}


def _read_metrics_tree(g) -> dict:
    """simulation.py:68-77."""
    return {"region": str(_decode_attr(g.attrs["region"])), "peak_magnE": float(g.attrs["peak_magnE"]),
            "peak_location_mm": g["peak_location_mm"].read(),
            "center_of_gravity_mm": g["center_of_gravity_mm"].read(),
            "region_volume_m3": float(g.attrs["region_volume_m3"]),
            "focality_m3": {k: float(v) for k, v in g["focality_m3"].attrs.items()},
            "distribution": {k: float(v) for k, v in g["distribution"].attrs.items()}}


def encode_ccd(ccd_file: str | Path, output_h5: str | Path) -> None:
    """A SimNIBS ``.ccd`` dipole file as the ``.h5`` coil ``Coil.load`` reads (coil.py:85-98): the
    rows as the ``dipoles`` dataset, the header's ``key=value`` pairs and ``num_elements`` as its
    attributes, gzip-compressed."""
    from . import h5lite
    with Path(ccd_file).open() as f:
        meta_line = f.readline().strip().lstrip("#").strip()
        count = int(f.readline().strip())
        data = np.loadtxt(f, dtype=np.float64, skiprows=1, max_rows=count).reshape(-1, 6)
    attrs: dict = {k.strip(): v.strip() for k, v in (p.split("=", 1) for p in meta_line.split(";") if "=" in p)}
    attrs["num_elements"] = np.int64(count)
    h5lite.write_file(output_h5, {"dipoles": h5lite.DatasetSpec(data, attrs)}, compression="gzip")


def _sites(placements, method: str) -> list[Placement]:
    if isinstance(placements, Placement):
        raise TypeError(f"{method}() takes a sequence of placements; wrap a single one in a list, "
                        "or use the matching simulate() method.")
    sites = list(placements)
    for s in sites:
        if not isinstance(s, Placement):
            raise TypeError(f"{method}() expects Placement objects, got {type(s).__name__}")
    return sites


class Subject:
    """A head mesh and its GPU solver state (simulation.py:169-231). The context (Morton layout,
    FP64 assembly, AMG hierarchy) is built on first use and kept until ``free()``."""

    def __init__(self, mesh: HeadMesh) -> None:
        self._mesh = mesh
        self._ctx: fem.SolverContext | None = None
        self._host: tuple[np.ndarray, np.ndarray, np.ndarray] | None = None
        self._lock = threading.RLock()

    @classmethod
    def from_mesh(cls, mesh_file: str | Path) -> Subject:
        return cls(load_mesh(mesh_file))

    def free(self) -> None:
        with self._lock:
            if self._ctx is not None:
                self._ctx.free()
            self._ctx = None
            self._host = None

    def __enter__(self) -> Subject:
        return self

    def __exit__(self, *exc) -> None:
        self.free()

    @property
    def context(self) -> fem.SolverContext:
        with self._lock:
            if self._ctx is None:
                m = self._mesh
                self._ctx = fem.SolverContext(m.nodes_mm, m.tet_nodes, m.tet_tags, m.skin_tris)
            return self._ctx

    @property
    def mesh(self) -> HeadMesh:
        """The mesh in the (Morton) order every result array is indexed by."""
        c = self.context
        return HeadMesh(c.nodes_mm, c.tet_nodes, c.tet_tags, c.skin_tris)

    def _host_metric_inputs(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        with self._lock:
            if self._host is None:
                c = self.context
                self._host = (c.vols, c.tet_tags, c.tet_barycenters_mm)
            return self._host

    def _slot_maps(self, recovery: str) -> tuple[np.ndarray, np.ndarray | None]:
        with self._lock:
            cache = self.__dict__.setdefault("_slot_cache", {})
            if recovery not in cache:
                op = self.context.recovery_operator(recovery)
                cache[recovery] = (op["slot_node"], op["slot_tag"])
            return cache[recovery]

    def _iter_blocks(self, coil: Coil, sites: list[Placement], didt: float, block_k: int | None,
                     magnitude: bool, vectors: bool, potential: bool, recovery: str,
                     nodal: bool = False) -> Iterator[FieldResult]:
        ctx = self.context
        chunk_k = fem.DEFAULT_BLOCK if block_k is None else max(1, min(fem.MAX_BLOCK, int(block_k)))
        if not (magnitude or vectors or potential or nodal) and len(sites) > chunk_k:
            # Summaries only: the pipelined sweep (the next block's dA/dt and RHS overlap the
            # current block's solve), in runs of SWEEP_BLOCKS blocks so results still stream.
            run = SWEEP_BLOCKS * chunk_k
            for start in range(0, len(sites), run):
                part = sites[start:start + run]
                args = [(s.center_mm, s.handle_mm, s.distance_mm) for s in part]
                out = fem.solve_sweep(ctx, coil.positions_m, coil.moments, args, didt, chunk_k, recovery=recovery)
                for i, site in enumerate(part):
                    yield FieldResult(
                        summary=fem.metrics_dict(out["metrics"][i]), magnE=None, E=None, v=None,
                        transform=out["transforms"][i], placement=site, coil_name=coil.name, didt=didt, vols=None,
                        tet_tags=None, barycenters_mm=None, recovery=recovery, iterations=int(out["iterations"][i]),
                        relative_residual=float(out["relative_residual"][i]), n_nodes=None, E_slots=None,
                        slot_node=None, slot_tag=None, tet_nodes=None)
            return
        host = self._host_metric_inputs() if (magnitude or nodal) else (None, None, None)
        slot_node, slot_tag = self._slot_maps(recovery) if nodal else (None, None)
        tets = ctx.tet_nodes if nodal and slot_tag is None else None
        for start in range(0, len(sites), chunk_k):
            chunk = sites[start:start + chunk_k]
            args = [(s.center_mm, s.handle_mm, s.distance_mm) for s in chunk]
            outs = fem.solve_placements_block(ctx, coil.positions_m, coil.moments, args, didt,
                                              magnitude=magnitude, vectors=vectors, potential=potential,
                                              metrics=True, recovery=recovery, slots=nodal)
            for site, out in zip(chunk, outs):
                yield FieldResult(
                    summary=fem.metrics_dict(out.metrics), magnE=out.magnE, E=out.E, v=out.v,
                    transform=out.transform, placement=site, coil_name=coil.name, didt=didt,
                    vols=host[0], tet_tags=host[1], barycenters_mm=host[2], recovery=recovery,
                    iterations=out.iterations, relative_residual=out.relative_residual,
                    n_nodes=ctx.n_nodes if nodal else None, E_slots=out.E_slots, slot_node=slot_node,
                    slot_tag=slot_tag, tet_nodes=tets)

    def iter_simulate(self, coil: Coil, placements: Sequence[Placement], didt: float = 1e6, *,
                      magnitude: bool = False, vectors: bool = False, potential: bool = False,
                      nodal: bool = False, recovery: str = "harmonic",
                      block_k: int | None = None) -> Iterator[FieldResult]:
        """Stream one FieldResult per placement, in order (simulation.py:454-509). Placements are
        solved in blocks of ``block_k`` (default fem.DEFAULT_BLOCK = 16, clamped to [1, MAX_BLOCK]) that share
        one read of the operator and hierarchy per iteration. Arguments are checked here, before
        the first result is requested."""
        sites = _sites(placements, "iter_simulate")
        recovery = validate_recovery(recovery)
        if nodal and recovery == "raw":
            raise ValueError("nodal=True needs a recovery mode; the raw per-tetrahedron field has no nodal form.")
        if not np.isfinite(didt):
            raise ValueError("didt must be finite")
        return self._iter_blocks(coil, sites, float(didt), block_k, magnitude, vectors, potential, recovery, nodal)

    def simulate(self, coil: Coil, placement: Placement, didt: float = 1e6, *, magnitude: bool = False,
                 vectors: bool = False, potential: bool = False, nodal: bool = False,
                 recovery: str = "harmonic", block_k: int | None = None) -> FieldResult:
        """One placement (simulation.py:511-546)."""
        if not isinstance(placement, Placement):
            raise TypeError("simulate() takes a single Placement. Use Subject.iter_simulate(...) to stream "
                            "a sequence.")
        (result,) = self.iter_simulate(coil, [placement], didt, magnitude=magnitude, vectors=vectors,
                                       potential=potential, nodal=nodal, recovery=recovery, block_k=block_k)
        return result

    def iter_simulate_conductivity_uq(self, coil: Coil, placements: Sequence[Placement], config, didt: float = 1e6,
                                      *, moments: bool = False, record_rois: Mapping | None = None,
                                      recovery: str = "harmonic"):
        """A conductivity Monte Carlo per placement, in the order given (simulation.py:593-632).
        Every result carries its gray-matter summary and the per-draw peak |E|, focality and peak
        location; ``moments`` keeps the per-element mean/std/cov of |E|; ``record_rois`` maps a
        name to an ROI (an object with ``elem_idx`` and ``weights``, or an ``(elem_idx, weights)``
        pair) whose weighted mean |E| each draw records."""
        from cunibs_amd import uq
        sites = _sites(placements, "iter_simulate_conductivity_uq")
        recovery = validate_recovery(recovery)
        if not isinstance(config, uq.ConductivityUQConfig):
            raise TypeError("config must be a ConductivityUQConfig")
        rois = dict(record_rois or {})

        def run():
            with self._lock:
                ctx = self.context
                host = self._host_metric_inputs()
                for site in sites:
                    yield uq.conductivity_uq(ctx, coil, site, config, didt, moments=moments, record_rois=rois,
                                             recovery=recovery, host_inputs=host)

        return run()

    def simulate_conductivity_uq(self, coil: Coil, placement: Placement, config, didt: float = 1e6, *,
                                 moments: bool = False, record_rois: Mapping | None = None,
                                 recovery: str = "harmonic"):
        """One placement's conductivity Monte Carlo (simulation.py:634-664)."""
        if not isinstance(placement, Placement):
            raise TypeError("simulate_conductivity_uq() takes a single Placement. Use "
                            "Subject.iter_simulate_conductivity_uq(...) to stream a sequence.")
        (result,) = self.iter_simulate_conductivity_uq(coil, [placement], config, didt, moments=moments,
                                                       record_rois=record_rois, recovery=recovery)
        return result
