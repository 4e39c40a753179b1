"""Tuning knobs (Chebyshev degree, fine-level SpMM tile) and the autotuner, on the 4 mm head.

Every setting must leave the numerics of its kernels where the oracle puts them: the V-cycle at
degree d against the oracle's FP64 V-cycle built with the same degree (fem/solve.py:153-168 fixes
d = 2; the coefficients follow the same recurrence for any d), and the products at any tile
against the FP64 product.
"""

from __future__ import annotations

import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b) -> float:
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def octx(head4, oracle_mod):
    return oracle_mod.build_context(head4.nodes_mm, head4.tet_nodes, head4.tet_tags, head4.skin_tris)


@pytest.fixture(scope="module")
def tctx(head4):
    from cunibs_amd import fem
    ctx = fem.SolverContext(head4.nodes_mm, head4.tet_nodes, head4.tet_tags, head4.skin_tris)
    yield ctx
    ctx.free()


def test_default_tuning_is_the_reference_setting(tctx):
    t = tctx.tuning
    assert t == {"degree": 2, "lower_ratio": pytest.approx(12.0), "fine_sub": 0, "coarse_sub": 0, "coil_blocks": 0,
                 "layout": 0, "dense_tail": 1}


@pytest.mark.parametrize("degree", [1, 3, 6])
def test_vcycle_at_another_degree_matches_the_oracle(tctx, octx, oracle_mod, degree):
    h = dataclasses.replace(octx.hier, cheb=oracle_mod.chebyshev_coefficients(degree, 12.0))
    B = np.random.default_rng(degree).standard_normal((octx.reduced.n, 8)).astype(np.float32)
    tctx.set_tuning(degree=degree)
    try:
        assert tctx.tuning["degree"] == degree
        got = tctx.vcycle(B)
    finally:
        tctx.set_tuning(degree=2)
    ref = np.stack([oracle_mod.vcycle_apply(h, B[:, c].astype(np.float64)) for c in range(8)], 1)
    assert rel(got, ref) < 1e-5


@pytest.mark.parametrize("sub", [1, 4, 16])
def test_products_at_any_tile_match_the_oracle(tctx, octx, oracle_mod, sub):
    n = octx.reduced.n
    X = np.random.default_rng(sub).standard_normal((n, 8)).astype(np.float32)
    ref = np.stack([oracle_mod.spmv(octx.reduced, octx.reduced.val.astype(np.float64), X[:, c].astype(np.float64))
                    for c in range(8)], 1)
    vref = np.stack([oracle_mod.vcycle_apply(octx.hier, X[:, c].astype(np.float64)) for c in range(8)], 1)
    tctx.set_tuning(fine_sub=sub)
    try:
        assert tctx.tuning["fine_sub"] == sub
        assert rel(tctx.spmm_diff(X), ref) < 1e-6
        assert rel(tctx.spmm(X), ref) < 1e-6
        assert rel(tctx.vcycle(X), vref) < 1e-5
    finally:
        tctx.set_tuning(fine_sub=0)


@pytest.mark.parametrize("sub", [1, 4, 16])
def test_vcycle_at_any_coarse_tile_matches_the_oracle(tctx, octx, oracle_mod, sub):
    """The coarser levels' tile changes only how the row kernels split a row's entries."""
    X = np.random.default_rng(50 + sub).standard_normal((octx.reduced.n, 8)).astype(np.float32)
    vref = np.stack([oracle_mod.vcycle_apply(octx.hier, X[:, c].astype(np.float64)) for c in range(8)], 1)
    tctx.set_tuning(coarse_sub=sub)
    try:
        assert tctx.tuning["coarse_sub"] == sub
        assert rel(tctx.vcycle(X), vref) < 1e-5
    finally:
        tctx.set_tuning(coarse_sub=0)


def test_bad_tuning_is_rejected(tctx):
    with pytest.raises(ValueError):
        tctx.set_tuning(degree=7)
    with pytest.raises(ValueError):
        tctx.set_tuning(fine_sub=3)
    with pytest.raises(ValueError):
        tctx.set_tuning(lower_ratio=1.0)
    with pytest.raises(ValueError):
        tctx.set_tuning(coarse_sub=6)
    assert tctx.tuning["degree"] == 2


def test_autotuned_context_gives_the_same_field(tctx, octx, oracle_mod, coil):
    """The tuner picks among the candidates, leaves the context at its choice, and a solve at the
    chosen setting meets the same GM |E| bar against the FP64 oracle."""
    from cunibs_amd import autotune, fem, synthetic
    sites = synthetic.cap_sites(8, seed=11)
    res = autotune.autotune(tctx, coil.positions_m, coil.moments, sites, block_ks=(2, 4, 8), degrees=(2, 3),
                            subs=(0, 2), coarse_subs=(0, 8), base_k=4)
    try:
        assert res.block_k in (2, 4, 8) and res.degree in (2, 3) and res.fine_sub in (0, 2)
        assert res.coarse_sub in (0, 8) and set(res.coarse_tile_us) == {0, 8}
        assert tctx.tuning["coarse_sub"] == res.coarse_sub
        assert set(res.width_rates) == {2, 4, 8} and set(res.degree_rates) == {2, 3}
        assert all(v > 0 for v in res.width_rates.values())
        assert tctx.tuning["degree"] == res.degree and tctx.tuning["fine_sub"] == res.fine_sub
        out = fem.solve_placements_block(tctx, coil.positions_m, coil.moments, sites[:res.block_k], 1e6)
        ref = oracle_mod.solve_site(octx, coil.positions_m, coil.moments, sites[0], 1e6, tol=1e-12)
        gm = octx.tet_tags == 2
        assert rel(out[0].magnE[gm], ref["magnE"][gm]) < 1e-4
        assert all(r.relative_residual <= 1e-6 for r in out)
    finally:
        tctx.set_tuning(degree=2, fine_sub=0, coarse_sub=0)


def test_in_solve_timing_leaves_the_answer_bitwise_unchanged(tctx, coil):
    """Device-clock stamps inside the captured iteration time the SpMM, the V-cycle and the
    iteration without touching the numerics."""
    from cunibs_amd import fem, synthetic
    sites = synthetic.cap_sites(8, seed=21)
    base = fem.solve_placements_block(tctx, coil.positions_m, coil.moments, sites, 1e6)
    tctx.timing_enable(True)
    try:
        timed = fem.solve_placements_block(tctx, coil.positions_m, coil.moments, sites, 1e6)
        t = tctx.timing_read()
        ph = tctx.timing_phases()
    finally:
        tctx.timing_enable(False)
    for a, b in zip(base, timed):
        np.testing.assert_array_equal(a.magnE, b.magnE)
        assert a.iterations == b.iterations
    assert t["samples"] >= 1
    assert 0 < t["spmm_ms"] < t["iteration_ms"]
    assert 0 < t["vcycle_ms"] < t["iteration_ms"]
    assert t["spmm_ms"] + t["vcycle_ms"] <= t["iteration_ms"]
    assert ph["blocks"] == 1 and ph["placements"] == 8
    per = ph["ms_per_block"]
    assert set(per) == set(fem.SolverContext.PHASES)
    assert all(v >= 0 for v in per.values()) and per["pcg"] > 0 and per["dadt"] > 0
    assert t["dadt_ms"] == pytest.approx(per["dadt"] / 8)
