"""Launch-shape autotuning for a sweep: the block width (placements per block PCG) that gives the
most placements per second on this mesh and GPU, measured, not modelled.

The reference fixes its block at MAX_BLOCK = 8 (core/common.hpp:27-33); here the block method's
iteration count and the cost of one iteration both change with the width (wider blocks converge in
fewer iterations and stream the operator once for more right-hand sides, but their FP64 Gram work
grows as K^2), so the best width depends on the mesh and is found by timing a block of each.
"""

from __future__ import annotations

import time
from collections.abc import Sequence

from cunibs_amd import fem


def choose_block_k(ctx: fem.SolverContext, dip_pos_m, dip_moment, sites: Sequence, didt: float = 1e6,
                   candidates: Sequence[int] = (8, 16), warm_start: bool = False, repeats: int = 1) -> tuple[int, dict]:
    """Time `repeats` blocks of each candidate width on the first placements of `sites` (after an
    untimed block that captures that width's iteration graph); returns the fastest width and the
    measured placements/s of every candidate."""
    rates: dict[int, float] = {}
    for k in candidates:
        if k < 1 or k > fem.MAX_BLOCK or len(sites) < k:
            continue
        chunk = list(sites[:k])
        fem.solve_placements_block(ctx, dip_pos_m, dip_moment, chunk, didt, magnitude=False, metrics=True,
                                   warm_start=warm_start)
        t0 = time.perf_counter()
        for _ in range(repeats):
            fem.solve_placements_block(ctx, dip_pos_m, dip_moment, chunk, didt, magnitude=False, metrics=True,
                                       warm_start=warm_start)
        rates[k] = repeats * k / (time.perf_counter() - t0)
    if not rates:
        raise ValueError("no candidate block width fits the placements given")
    best = max(rates, key=rates.get)
    return best, rates
