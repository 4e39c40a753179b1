"""Conductivity uncertainty for one placement (the reference's uq/conductivity, config.py:56-72 and
run.py:203-387): tissue conductivities drawn log-normal around the table values, solved in blocks
of up to 16 samples that share the operator's pattern (one launch per kernel for the block), the
GM peak |E| per sample and the moments of |E| over all samples."""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from cunibs_amd import fem


def sample_conductivities(tags, n: int, log_sd: float = 0.2, seed: int = 0, table: dict | None = None,
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
        sigma = sample_conductivities(tags, n_samples, log_sd, seed)
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
