"""Autotuner for a placement sweep: block width, fine-level SpMM tile and Chebyshev degree.

The reference fixes all three: blocks of at most MAX_BLOCK = 8 placements (core/common.hpp:27-33),
Chebyshev degree 2 with ratio 12 (fem/solve.py:153-168) and the V-cycle's threads per row from its
`choose_tpr` rule (solver/amg/vcycle.cu:104-120). Here each is measured on the mesh at hand:

* the tile (sub-lanes per 4-column group of the fine level's row kernels) changes only the speed
  of one iteration, so it is chosen from kernel timings: the differenced SpMM with its Gram pass
  plus one V-cycle, HIP-event timed on the library's stream; the coarser levels' tile likewise,
  from V-cycle timings;
* the degree trades smoother cost against iterations, and the block width trades the operator
  stream's reuse and the faster convergence of wide blocks against their K^2 Gram work, so both
  are chosen by timing whole blocks over the same placements.

Every candidate is timed after one untimed block that captures that setting's iteration graph.
The context is left at the winning tile and degree; the winning width is returned.
"""

from __future__ import annotations

import time
from collections.abc import Sequence
from dataclasses import dataclass, field

from cunibs_amd import fem


@dataclass
class TuningResult:
    block_k: int
    degree: int
    fine_sub: int
    coarse_sub: int = 0
    seconds: float = 0.0
    tile_us: dict = field(default_factory=dict)          # fine_sub -> us per iteration's fine work
    coarse_tile_us: dict = field(default_factory=dict)   # coarse_sub -> us per V-cycle
    lower_ratio: float = 12.0
    degree_rates: dict = field(default_factory=dict)     # degree -> placements/s
    ratio_rates: dict = field(default_factory=dict)      # Chebyshev lower ratio -> placements/s
    width_rates: dict = field(default_factory=dict)      # block_k -> placements/s

    def as_dict(self) -> dict:
        return {"block_k": self.block_k, "degree": self.degree, "fine_sub": self.fine_sub,
                "coarse_sub": self.coarse_sub, "lower_ratio": self.lower_ratio, "seconds": round(self.seconds, 3),
                "tile_us": {str(k): round(v, 2) for k, v in self.tile_us.items()},
                "coarse_tile_us": {str(k): round(v, 2) for k, v in self.coarse_tile_us.items()},
                "degree_rates": {str(k): round(v, 2) for k, v in self.degree_rates.items()},
                "ratio_rates": {str(k): round(v, 2) for k, v in self.ratio_rates.items()},
                "width_rates": {str(k): round(v, 2) for k, v in self.width_rates.items()}}


WIDTH_TIE = 0.02  # relative rate difference below which two block widths count as tied


def pick_width(rates: dict) -> int:
    """The block width to solve at, from placements/s by width: widths within WIDTH_TIE of the
    best count as tied and the narrowest of them is taken (a narrower block's FP32 recursion drifts
    less from the FP64 true residual, and a sweep's recycled guess holds two earlier blocks of up to
    16 columns but one of 32)."""
    best = max(rates.values())
    return min(k for k, r in rates.items() if r >= (1.0 - WIDTH_TIE) * best)


def _rate(ctx, pos, mom, sites, didt, k: int, warm_start: bool) -> float:
    """Placements/s over `sites` in blocks of k, after one untimed block (graph capture)."""
    def run(chunk):
        fem.solve_placements_block(ctx, pos, mom, chunk, didt, magnitude=False, metrics=True, warm_start=warm_start)

    run(list(sites[:k]))
    t0 = time.perf_counter()
    for i in range(0, len(sites), k):
        run(list(sites[i:i + k]))
    return len(sites) / (time.perf_counter() - t0)


def autotune(ctx: fem.SolverContext, dip_pos_m, dip_moment, sites: Sequence, didt: float = 1e6, *,
             block_ks: Sequence[int] = (4, 8, 16, 32), degrees: Sequence[int] = (1, 2, 3, 4, 5, 6),
             subs: Sequence[int] = (0, 1, 2, 4, 8), coarse_subs: Sequence[int] = (0, 2, 4, 8, 16),
             ratios: Sequence[float] = (12.0,), base_k: int = 8, warm_start: bool = False) -> TuningResult:
    """Tune on `sites` (their count should be a multiple of every width, e.g. 32): the fine and the
    coarse tiles at width `base_k`, then the degree and the Chebyshev ratio at `base_k`, then the
    width."""
    t_start = time.perf_counter()
    widths = [k for k in block_ks if 1 <= k <= fem.MAX_BLOCK and k <= len(sites)]
    if not widths:
        raise ValueError("no candidate block width fits the placements given")
    if base_k not in widths:
        base_k = widths[len(widths) // 2]
    K = min(b for b in fem.BLOCK_SIZES if b >= base_k)
    res = TuningResult(block_k=base_k, degree=ctx.tuning["degree"], fine_sub=0)

    # 1. tile: kernel time of one iteration's fine-level work at the base width
    for sub in subs:
        try:
            ctx.set_tuning(fine_sub=sub)
            res.tile_us[sub] = 1e3 * (ctx.bench_spmm(K, 10, mode=2) + ctx.bench_vcycle(K, 5))
        except ValueError:
            continue  # more sub-lanes than a wavefront holds at this width
    res.fine_sub = min(res.tile_us, key=res.tile_us.get)
    ctx.set_tuning(fine_sub=res.fine_sub)
    for sub in coarse_subs:
        ctx.set_tuning(coarse_sub=sub)
        res.coarse_tile_us[sub] = 1e3 * ctx.bench_vcycle(K, 10)
    res.coarse_sub = min(res.coarse_tile_us, key=res.coarse_tile_us.get)
    ctx.set_tuning(coarse_sub=res.coarse_sub)

    # 2. degree: whole blocks at the base width
    for d in degrees:
        ctx.set_tuning(degree=d)
        res.degree_rates[d] = _rate(ctx, dip_pos_m, dip_moment, sites, didt, base_k, warm_start)
    res.degree = max(res.degree_rates, key=res.degree_rates.get)
    ctx.set_tuning(degree=res.degree)
    base_ratio = ctx.tuning["lower_ratio"]
    for r in ratios:
        if abs(r - base_ratio) < 1e-6:
            res.ratio_rates[r] = res.degree_rates[res.degree]
            continue
        ctx.set_tuning(lower_ratio=r)
        res.ratio_rates[r] = _rate(ctx, dip_pos_m, dip_moment, sites, didt, base_k, warm_start)
    if res.ratio_rates:
        res.lower_ratio = max(res.ratio_rates, key=res.ratio_rates.get)
        if res.ratio_rates[res.lower_ratio] > res.degree_rates[res.degree]:
            res.degree_rates[res.degree] = res.ratio_rates[res.lower_ratio]
    else:
        res.lower_ratio = base_ratio
    ctx.set_tuning(lower_ratio=res.lower_ratio)

    # 3. width: whole blocks over the same placements
    for k in widths:
        res.width_rates[k] = res.degree_rates[res.degree] if k == base_k else \
            _rate(ctx, dip_pos_m, dip_moment, sites, didt, k, warm_start)
    res.block_k = pick_width(res.width_rates)
    res.seconds = time.perf_counter() - t_start
    return res


def tune_overlap(ctx: fem.SolverContext, dip_pos_m, dip_moment, sites: Sequence, didt: float, block_k: int,
                 candidates: Sequence[int] = (0, 256, 512, 1024, 2048), warm_start: bool = False) -> dict:
    """The workgroup cap of the dA/dt that `fem.solve_sweep` overlaps with the previous block's
    solve: each candidate times a pipelined sweep of `sites` (after one untimed sweep of two
    blocks); the context is left at the fastest. Returns {"coil_blocks": best, "rates": {cap: placements/s}}."""
    rates = {}
    for cap in candidates:
        ctx.set_tuning(coil_blocks=cap)
        fem.solve_sweep(ctx, dip_pos_m, dip_moment, list(sites[:2 * block_k]), didt, block_k, warm_start=warm_start)
        t0 = time.perf_counter()
        fem.solve_sweep(ctx, dip_pos_m, dip_moment, list(sites), didt, block_k, warm_start=warm_start)
        rates[cap] = len(sites) / (time.perf_counter() - t0)
    best = max(rates, key=rates.get)
    ctx.set_tuning(coil_blocks=best)
    return {"coil_blocks": best, "rates": rates}
