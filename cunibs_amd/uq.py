"""Conductivity uncertainty for one placement (the reference's uq/conductivity, config.py:1-72,
run.py:203-387, result.py:1-256): tissue conductivities drawn around the table values, solved in
blocks of up to 16 samples that share the operator's pattern (one launch per kernel for the block),
each draw's field recovered before any statistic, the GM peak |E|, focality and peak location per
draw, and the per-element moments of |E| over all draws.

``ConductivityUQConfig``, ``sample_conductivities``, ``ConductivityUQResult`` (with ``summary``,
``tissue_sensitivity``) and ``Subject.simulate_conductivity_uq`` mirror the reference's public
surface. The draws come from NumPy's seeded generator, so they are not the reference's CuPy draws
for the same seed; their distribution is the same."""

from __future__ import annotations

from collections.abc import Mapping
from dataclasses import dataclass, field, replace

import numpy as np

from cunibs_amd import fem

# Lognormal coefficients of variation per tissue tag (config.py:12-25): the reference's
# illustrative defaults, overridable per study through ConductivityUQConfig.tissue_cov.
DEFAULT_TISSUE_COV: dict[int, float] = {
    1: 0.15, 2: 0.15, 3: 0.05, 4: 0.35, 5: 0.20, 6: 0.20, 7: 0.35, 8: 0.35, 9: 0.15, 10: 0.20,
}


@dataclass(frozen=True)
class ConductivityUQConfig:
    """Monte Carlo settings (config.py:28-53): each perturbed tissue's conductivity is an
    independent draw with the given coefficient of variation, log-normal with mean equal to the
    nominal value (or normal, clamped to 1e-6 x nominal). The preconditioner stays the nominal one
    for every draw; a draw that misses tolerance raises."""

    n_samples: int = 500
    seed: int = 0
    tissue_cov: dict[int, float] = field(default_factory=dict)
    perturbed_tags: tuple[int, ...] | None = None
    distribution: str = "lognormal"

    def cov_for(self, tag: int) -> float:
        if tag in self.tissue_cov:
            return self.tissue_cov[tag]
        return DEFAULT_TISSUE_COV[tag]


def sample_conductivities(config: ConductivityUQConfig, perturbed_tags) -> np.ndarray:
    """(n_samples, len(perturbed_tags)) conductivities (config.py:56-72): log-normal
    sigma_0 exp(s z - s^2/2) with s = sqrt(log(1 + cov^2)), so E[sigma] = sigma_0; or normal
    sigma_0 (1 + cov z) clamped to 1e-6 sigma_0."""
    if config.distribution not in ("lognormal", "normal"):
        raise ValueError(f"unknown distribution {config.distribution!r}")
    if config.n_samples < 1:
        raise ValueError("n_samples must be at least 1")
    tags = [int(t) for t in perturbed_tags]
    nominal = np.array([fem.TISSUE_CONDUCTIVITY[t] for t in tags], np.float64)
    covs = np.array([config.cov_for(t) for t in tags], np.float64)
    z = np.random.default_rng(config.seed).standard_normal((config.n_samples, len(tags)))
    if config.distribution == "lognormal":
        s = np.sqrt(np.log1p(covs * covs))
        return nominal * np.exp(z * s - 0.5 * s * s)
    return np.maximum(nominal * (1.0 + covs * z), 1e-6 * nominal)


@dataclass(frozen=True)
class ConductivityUQSummary:
    """result.py:64-76: metrics of the mean field over a region, and the largest per-element
    coefficient of variation of |E| there."""

    mean_field: dict
    max_local_cov: float


def _tissue_sensitivity(result, output: str) -> dict[int, float]:
    """First-order variance shares of a log-linear surrogate (the quantity of result.py:31-61).

    Model: log y = a + sum_j b_j u_j with u_j = log sigma_j, fitted by least squares over the
    draws. Under that model (and uncorrelated draws) Var(log y) = sum_j b_j^2 Var(u_j), so tissue
    j's share of the output's log-variance is b_j^2 Var(u_j) / Var(log y). The slopes solve the
    normal equations Cov(u) b = Cov(u, log y); a pseudo-inverse keeps a tissue whose draws do not
    vary (zero variance) at slope zero. Not a Sobol index."""
    named = {"peak": result.peak_samples, "focality": result.focality_samples, **result.roi_samples}
    if output not in named:
        raise ValueError(f"Unknown output {output!r}: choose one of {sorted(named)}")
    u = np.log(np.asarray(result.sigma_samples, np.float64))
    v = np.log(np.asarray(named[output], np.float64))
    n = v.shape[0]
    du = u - u.sum(axis=0) / n
    dv = v - v.sum() / n
    var_v = float(dv @ dv) / n
    tags = [int(t) for t in result.perturbed_tags]
    if var_v == 0.0:
        return dict.fromkeys(tags, 0.0)
    cov_uu = du.T @ du / n
    cov_uv = du.T @ dv / n
    slopes = np.linalg.pinv(cov_uu, rcond=1e-12, hermitian=True) @ cov_uv
    shares = slopes * slopes * np.diag(cov_uu) / var_v
    return {t: float(s) for t, s in zip(tags, shares)}


@dataclass
class ConductivityUQResult:
    """result.py:78-256: per-element moments of |E| over the draws (None unless ``moments``), the
    draws themselves and the per-draw gray-matter scalars, on the host."""

    mean_magnE: np.ndarray | None
    std_magnE: np.ndarray | None
    cov_magnE: np.ndarray | None
    n_samples: int
    perturbed_tags: tuple[int, ...]
    sigma_samples: np.ndarray
    vols: np.ndarray | None
    tet_tags: np.ndarray | None
    barycenters_mm: np.ndarray | None
    placement: object
    coil_name: str
    didt: float
    peak_samples: np.ndarray
    focality_samples: np.ndarray
    peak_location_samples: np.ndarray
    roi_samples: dict
    summary: ConductivityUQSummary | None = None
    recovery: str = "raw"
    iterations: np.ndarray | None = None

    def compute_summary(self, region: str = "gray_matter") -> ConductivityUQSummary:
        from cunibs_amd.simulation import _region_mask, compute_metrics
        if self.mean_magnE is None or self.cov_magnE is None:
            raise ValueError(f"Metrics for region {region!r} need the per-element moment arrays, which "
                             "this result did not retain. Re-run with moments=True.")
        mask = _region_mask(self.tet_tags, region)
        return ConductivityUQSummary(
            mean_field=compute_metrics(self.mean_magnE, self.vols, self.barycenters_mm, self.tet_tags, region=region),
            max_local_cov=float(np.max(self.cov_magnE[mask])) if mask.any() else 0.0)

    def _summary_for(self, region: str) -> ConductivityUQSummary:
        if self.summary is not None and region == self.summary.mean_field["region"]:
            return self.summary
        return self.compute_summary(region)

    def peak_mean_magnE(self, region: str = "gray_matter") -> float:
        return self._summary_for(region).mean_field["peak_magnE"]

    def max_local_cov(self, region: str = "gray_matter") -> float:
        return self._summary_for(region).max_local_cov

    def tissue_sensitivity(self, output: str = "peak") -> dict[int, float]:
        return _tissue_sensitivity(self, output)

    def to_numpy(self) -> ConductivityUQResult:
        return self

    def save(self, path) -> None:
        """An HDF5 file in the reference's layout (result.py:182-223): the retained moment arrays,
        the draws, the summary, the per-draw scalars and ROI records, the run's attributes."""
        from cunibs_amd import h5lite
        from cunibs_amd.simulation import _metrics_tree
        if self.summary is None:
            raise ValueError("This result has no summary to save; call compute_summary() first "
                             "(results from Subject already carry one).")
        tree: dict = {name: getattr(self, name) for name in _UQ_ARRAYS if getattr(self, name) is not None}
        tree["sigma_samples"] = np.asarray(self.sigma_samples, np.float64)
        tree["summary"] = {"mean_field": _metrics_tree(self.summary.mean_field),
                           "__attrs__": {"max_local_cov": float(self.summary.max_local_cov)}}
        tree["roi_samples"] = {k: np.asarray(v, np.float64) for k, v in self.roi_samples.items()}
        for name in ("peak_samples", "focality_samples", "peak_location_samples"):
            tree[name] = np.asarray(getattr(self, name), np.float64)
        tree["__attrs__"] = {
            "format_version": np.int64(1), "recovery": self.recovery, "n_samples": np.int64(self.n_samples),
            "perturbed_tags": np.asarray(self.perturbed_tags, np.int32), "coil_name": self.coil_name,
            "didt": float(self.didt), "placement_center_mm": self.placement.center_mm,
            "placement_handle_mm": self.placement.handle_mm,
            "placement_distance_mm": float(self.placement.distance_mm)}
        h5lite.write_file(path, tree, compression="gzip")

    @classmethod
    def load(cls, path) -> ConductivityUQResult:
        """result.py:225-256."""
        from cunibs_amd import h5lite
        from cunibs_amd.simulation import Placement, _decode_attr, _read_metrics_tree
        f = h5lite.open_file(path)
        stored = int(f.attrs.get("format_version", 0))
        if stored != 1:
            raise ValueError(f"{path}: file format version {stored} is not readable by this version (expected 1).")
        a = f.attrs
        opt = {name: (f[name].read() if name in f else None) for name in _UQ_ARRAYS}
        roi = f["roi_samples"]
        return cls(n_samples=int(a["n_samples"]), perturbed_tags=tuple(int(t) for t in np.atleast_1d(a["perturbed_tags"])),
                   sigma_samples=f["sigma_samples"].read(),
                   placement=Placement(a["placement_center_mm"], a["placement_handle_mm"],
                                       float(a["placement_distance_mm"])),
                   coil_name=str(_decode_attr(a["coil_name"])), didt=float(a["didt"]),
                   recovery=str(_decode_attr(a.get("recovery", "raw"))),
                   summary=ConductivityUQSummary(_read_metrics_tree(f["summary"]["mean_field"]),
                                                 float(f["summary"].attrs["max_local_cov"])),
                   roi_samples={k: roi[k].read() for k in roi.keys()},
                   peak_samples=f["peak_samples"].read(), focality_samples=f["focality_samples"].read(),
                   peak_location_samples=f["peak_location_samples"].read(), **opt)


_UQ_ARRAYS = ("mean_magnE", "std_magnE", "cov_magnE", "vols", "tet_tags", "barycenters_mm")


def _roi_arrays(target) -> tuple[np.ndarray, np.ndarray]:
    if isinstance(target, tuple):
        idx, w = target
    else:
        idx, w = target.elem_idx, target.weights
    return np.asarray(idx, np.int64), np.asarray(w, np.float64)


def conductivity_uq(ctx: fem.SolverContext, coil, placement, config: ConductivityUQConfig, didt: float, *,
                    moments: bool = False, record_rois: Mapping = {}, recovery: str = "harmonic",
                    host_inputs=None, block: int = 16) -> ConductivityUQResult:
    """One placement's Monte Carlo (run.py:203-387) on the GPU: every draw solved under the
    nominal preconditioner from the nominal solution plus its sensitivity correction, recovered,
    summarised on the device; moments accumulated on the device over all draws."""
    site = (placement.center_mm, placement.handle_mm, placement.distance_mm)
    tags = ctx.uq_prepare(coil.positions_m, coil.moments, site, didt)
    perturbed = tuple(sorted(int(t) for t in config.perturbed_tags)) if config.perturbed_tags is not None \
        else tuple(tags)
    missing = [t for t in perturbed if t not in tags]
    if missing:
        raise ValueError(f"perturbed tags {missing} are not in the mesh (tags {tags})")
    draws = sample_conductivities(config, perturbed)
    sigma = np.tile(np.array([fem.TISSUE_CONDUCTIVITY[t] for t in tags], np.float64), (config.n_samples, 1))
    for j, t in enumerate(perturbed):
        sigma[:, tags.index(t)] = draws[:, j]
    probes = {name: _roi_arrays(t) for name, t in record_rois.items()}
    n = config.n_samples
    peak, foc, loc, its = np.empty(n), np.empty(n), np.empty((n, 3)), np.empty(n, np.int32)
    roi = {name: np.empty(n) for name in probes}
    for i in range(0, n, block):
        res = ctx.uq_solve(sigma[i:i + block], magnitude=bool(probes), metrics=True, recovery=recovery)
        for j, r in enumerate(res):
            peak[i + j] = r.metrics[0]
            loc[i + j] = r.metrics[1:4]
            foc[i + j] = r.metrics[8]
            its[i + j] = r.iterations
            for name, (idx, w) in probes.items():
                roi[name][i + j] = float((r.magnE[idx].astype(np.float64) * w).sum())
    mean, var, _ = ctx.uq_moments()
    std = np.sqrt(np.clip(var, 0.0, None))
    cov = np.where(mean > 1e-12 * float(mean.max()), std / np.where(mean > 0, mean, 1.0), 0.0)
    vols, tet_tags, bary = host_inputs if host_inputs is not None else (None, None, None)
    result = ConductivityUQResult(
        mean_magnE=mean.astype(np.float32), std_magnE=std.astype(np.float32), cov_magnE=cov.astype(np.float32),
        n_samples=n, perturbed_tags=perturbed, sigma_samples=draws, vols=vols, tet_tags=tet_tags,
        barycenters_mm=bary, placement=placement, coil_name=coil.name, didt=float(didt), peak_samples=peak,
        focality_samples=foc, peak_location_samples=loc, roi_samples=roi, recovery=recovery, iterations=its)
    if vols is not None:
        result.summary = result.compute_summary()
    if not moments:  # the per-element arrays come as a set (simulation.py:572-581)
        result = replace(result, mean_magnE=None, std_magnE=None, cov_magnE=None, vols=None, tet_tags=None,
                         barycenters_mm=None)
    return result


def draw_lognormal(tags, n: int, log_sd: float = 0.2, seed: int = 0, table: dict | None = None,
                   mean_preserving: bool = True) -> np.ndarray:
    """(n, len(tags)) draws sigma = sigma_0 exp(s z - s^2/2) (mean sigma_0, the reference's
    parameterisation, config.py:63-69) or sigma_0 exp(s z) (median sigma_0), z ~ N(0, 1), seeded."""
    table = fem.TISSUE_CONDUCTIVITY if table is None else table
    s0 = np.array([table[t] for t in tags], np.float64)
    z = np.random.default_rng(seed).standard_normal((n, len(tags)))
    shift = 0.5 * log_sd * log_sd if mean_preserving else 0.0
    return s0 * np.exp(log_sd * z - shift)


@dataclass
class UQResult:
    sigma: np.ndarray          # (n, n_tissue)
    tags: list[int]
    gm_peak: np.ndarray        # (n,) GM peak |E| per sample
    iterations: np.ndarray     # (n,)
    magnE_mean: np.ndarray     # (m,) per-element mean of |E|
    magnE_var: np.ndarray      # (m,) per-element sample variance
    focality: np.ndarray | None = None       # (n,) GM volume (m^3) at |E| >= 0.5 x p99.9 per sample
    peak_location: np.ndarray | None = None  # (n, 3) mm
    guess: dict | None = None                # nominal / sensitivity iterations, correction rank

    @property
    def gm_peak_mean(self) -> float:
        return float(self.gm_peak.mean())

    @property
    def gm_peak_sd(self) -> float:
        return float(self.gm_peak.std(ddof=1)) if self.gm_peak.size > 1 else 0.0


def run_conductivity_uq(ctx: fem.SolverContext, dip_pos_m, dip_moment, site, didt: float, n_samples: int, *,
                        log_sd: float = 0.2, seed: int = 0, block: int = 16, sigma: np.ndarray | None = None,
                        initial_guess: bool = True) -> UQResult:
    tags = ctx.uq_prepare(dip_pos_m, dip_moment, site, didt)
    if sigma is None:
        sigma = draw_lognormal(tags, n_samples, log_sd, seed)
    peaks, iters, foc, loc = [], [], [], []
    for i in range(0, sigma.shape[0], block):
        res = ctx.uq_solve(sigma[i:i + block], magnitude=False, initial_guess=initial_guess)
        peaks.extend(float(r.metrics[0]) for r in res)
        loc.extend(np.asarray(r.metrics[1:4], np.float64) for r in res)
        foc.extend(float(r.metrics[8]) for r in res)
        iters.extend(r.iterations for r in res)
    mean, var, _ = ctx.uq_moments()
    return UQResult(sigma, tags, np.array(peaks), np.array(iters), mean, var, np.array(foc), np.array(loc),
                    ctx.uq_guess_info())
