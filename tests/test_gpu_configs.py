"""The HIP path on the workloads the throughput metric is quoted on (the seeded perturbed five-shell
heads at 2.0, 1.5 and 1.25 mm, 0.43M to 1.7M nodes), against the FP64 oracle where it finishes in
seconds and through size-independent properties where it does not:

* 1.5 mm, a block of eight placements: GM |E| of two placements within 1e-3 of the oracle's CG to
  1e-12, and every placement's FP64 true residual, recomputed on the host from the returned
  potential, the solver's own right-hand side and the FP64 reduced operator, within 1e-6;
* 2.0 and 1.25 mm: pipelined sweeps with a ragged last block equal to block-by-block solves bit
  for bit; the true residual on the host; exact linearity in dI/dt (the power-of-two scaling
  commutes with every rounding); one oracle placement on 2.0 mm;
* conductivity uncertainty at 1.5 mm: one launch of 16 draws, two of them against per-draw oracle
  solves;
* the learned warm start trained on the 1.5 mm head: same answer as a cold start, fewer
  iterations on held-out placements.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import true_residuals

pytestmark = pytest.mark.gpu

GM_BAR = 1e-3


def rel(a, b) -> float:
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def _head(spacing):
    from cunibs_amd import synthetic
    return synthetic.shell_head(spacing, perturb_rms_mm=2.0, seed=1)


def _context(head):
    from cunibs_amd import fem
    return fem.SolverContext(head.nodes_mm, head.tet_nodes, head.tet_tags, head.skin_tris)


@pytest.fixture(scope="module")
def head15():
    return _head(1.5)


@pytest.fixture(scope="module")
def ctx15(head15):
    ctx = _context(head15)
    yield ctx
    ctx.free()


@pytest.fixture(scope="module")
def octx15(head15, oracle_mod):
    oracle_mod.set_threads(min(os.cpu_count() or 1, 16))
    return oracle_mod.build_context(head15.nodes_mm, head15.tet_nodes, head15.tet_tags, head15.skin_tris)


def test_config2_block_of_eight_at_1p5mm(ctx15, octx15, oracle_mod, coil):
    from cunibs_amd import fem, synthetic
    assert ctx15.n_nodes > 1_000_000 and ctx15.n_tets > 5_800_000
    sites = synthetic.cap_sites(8, seed=2)
    ex = {"rhs": 1}
    res = fem.solve_placements_block(ctx15, coil.positions_m, coil.moments, sites, 1e6, potential=True, extra=ex)
    assert all(r.relative_residual <= 1e-6 for r in res)
    assert max(true_residuals(ctx15, res, ex["rhs"])) <= 1e-6
    gm = octx15.tet_tags == fem.GM_TAG
    np.testing.assert_array_equal(ctx15.tet_tags, octx15.tet_tags)
    for i in (0, 5):
        ref = oracle_mod.solve_site(octx15, coil.positions_m, coil.moments, sites[i], 1e6, tol=1e-12)
        assert rel(res[i].magnE[gm], ref["magnE"][gm]) <= GM_BAR


@pytest.mark.parametrize("spacing", [2.0, 1.25])
def test_config3_heads_ragged_sweep_and_linearity(spacing, coil, oracle_mod):
    from cunibs_amd import fem, synthetic
    head = _head(spacing)
    ctx = _context(head)
    try:
        n_place = 13 if spacing == 2.0 else 9  # ragged last blocks of 5 (padded to 8) and 1
        sites = synthetic.grid_sites()[::256 // n_place][:n_place]
        sweep = fem.solve_sweep(ctx, coil.positions_m, coil.moments, sites, 1e6, block_k=8, recycle=0)
        ref, rhs = [], []
        for i in range(0, n_place, 8):
            ex = {"rhs": 1}
            ref.extend(fem.solve_placements_block(ctx, coil.positions_m, coil.moments, sites[i:i + 8], 1e6,
                                                  potential=True, extra=ex))
            rhs.extend(ex["rhs"])
        np.testing.assert_array_equal(sweep["metrics"], np.stack([r.metrics for r in ref]))
        np.testing.assert_array_equal(sweep["iterations"], [r.iterations for r in ref])
        assert max(true_residuals(ctx, ref, rhs)) <= 1e-6
        double = fem.solve_placements_block(ctx, coil.positions_m, coil.moments, sites[8:], 2e6)
        for a, b in zip(ref[8:], double):
            np.testing.assert_array_equal(2.0 * a.magnE, b.magnE)
        if spacing == 2.0:
            octx = oracle_mod.build_context(head.nodes_mm, head.tet_nodes, head.tet_tags, head.skin_tris)
            gm = octx.tet_tags == fem.GM_TAG
            o = oracle_mod.solve_site(octx, coil.positions_m, coil.moments, sites[12], 1e6, tol=1e-12)
            assert rel(ref[12].magnE[gm], o["magnE"][gm]) <= GM_BAR
    finally:
        ctx.free()


def test_config4_sixteen_draws_in_one_launch(ctx15, head15, coil, oracle_mod):
    from cunibs_amd import fem, synthetic, uq
    site = synthetic.cap_sites(1, seed=4)[0]
    tags = ctx15.uq_prepare(coil.positions_m, coil.moments, site, 1e6)
    sigma = uq.draw_lognormal(tags, 16, 0.2, seed=12)
    res = ctx15.uq_solve(sigma)
    assert len(res) == 16 and all(r.relative_residual <= 1e-6 for r in res)
    gm = ctx15.tet_tags == fem.GM_TAG
    for i in (3, 14):
        table = dict(fem.TISSUE_CONDUCTIVITY)
        table.update({t: float(s) for t, s in zip(tags, sigma[i])})
        octx = oracle_mod.build_context(head15.nodes_mm, head15.tet_nodes, head15.tet_tags, head15.skin_tris,
                                        conductivity=table)
        ref = oracle_mod.solve_site(octx, coil.positions_m, coil.moments, site, 1e6, tol=1e-12)
        assert rel(res[i].magnE[gm], ref["magnE"][gm]) <= GM_BAR
        assert float(res[i].metrics[0]) == pytest.approx(float(ref["magnE"][gm].max()), rel=1e-3)
    mean, var, count = ctx15.uq_moments()
    assert count == 16
    stack = np.stack([r.magnE.astype(np.float64) for r in res])
    np.testing.assert_allclose(mean, stack.mean(axis=0), rtol=1e-12, atol=1e-12)


def test_config5_learned_warm_start_at_1p5mm(ctx15, coil):
    from cunibs_amd import fem, learned, synthetic
    train = synthetic.cap_sites(256, seed=21, distance_range_mm=(2.0, 8.0))
    net = learned.train_warm_start(ctx15, coil.positions_m, coil.moments, train, modes=128, hidden=512,
                                   epochs=300)
    try:
        test = synthetic.cap_sites(16, seed=22, distance_range_mm=(2.0, 8.0))
        cold = fem.solve_placements_block(ctx15, coil.positions_m, coil.moments, test, 1e6)
        warm = learned.solve_with_network(ctx15, net, coil.positions_m, coil.moments, test)
        for a, b in zip(cold, warm):
            assert b.relative_residual <= 1e-6
            assert rel(b.magnE, a.magnE) <= 2e-5
        assert np.mean([r.iterations for r in warm]) < np.mean([r.iterations for r in cold])
    finally:
        net.free()
