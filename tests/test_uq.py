"""Conductivity uncertainty: the sampling law (the reference's tests/test_uq_sampling.py tier, CPU)
and the multi-matrix block solve against per-sample FP64 oracle solves (GPU)."""

from __future__ import annotations

import numpy as np
import pytest


def test_sampling_is_seeded_lognormal_with_the_table_mean():
    from cunibs_amd.uq import sample_conductivities
    tags = [1, 2, 3, 4, 5]
    a = sample_conductivities(tags, 20000, 0.2, seed=3)
    b = sample_conductivities(tags, 20000, 0.2, seed=3)
    np.testing.assert_array_equal(a, b)
    s0 = np.array([0.126, 0.275, 1.654, 0.010, 0.465])
    np.testing.assert_allclose(a.mean(axis=0), s0, rtol=0.01)  # mean-preserving form
    np.testing.assert_allclose(np.log(a).std(axis=0), 0.2, rtol=0.03)
    med = sample_conductivities(tags, 20000, 0.2, seed=4, mean_preserving=False)
    np.testing.assert_allclose(np.median(med, axis=0), s0, rtol=0.02)
    assert not np.array_equal(a, sample_conductivities(tags, 20000, 0.2, seed=5))


@pytest.mark.gpu
def test_uq_samples_match_per_sample_fp64_solves(head4, coil, oracle_mod):
    from cunibs_amd import fem, synthetic, uq
    ctx = fem.SolverContext(head4.nodes_mm, head4.tet_nodes, head4.tet_tags, head4.skin_tris)
    try:
        site = synthetic.SPHERE_SITE
        tags = ctx.uq_prepare(coil.positions_m, coil.moments, site, 1e6)
        assert tags == [1, 2, 3, 4, 5]
        sigma = uq.sample_conductivities(tags, 5, 0.2, seed=11)
        sigma[0] = [fem.TISSUE_CONDUCTIVITY[t] for t in tags]  # the nominal table
        res = ctx.uq_solve(sigma)
        nominal = fem.solve_placements_block(ctx, coil.positions_m, coil.moments, [site], 1e6)[0]
        assert np.linalg.norm(res[0].magnE - nominal.magnE) <= 1e-5 * np.linalg.norm(nominal.magnE)
        gm = ctx.tet_tags == 2
        for i in (1, 3, 4):
            table = dict(fem.TISSUE_CONDUCTIVITY)
            table.update({t: float(s) for t, s in zip(tags, sigma[i])})
            octx = oracle_mod.build_context(head4.nodes_mm, head4.tet_nodes, head4.tet_tags, head4.skin_tris,
                                            conductivity=table)
            ref = oracle_mod.solve_site(octx, coil.positions_m, coil.moments, site, 1e6, tol=1e-12)
            assert res[i].relative_residual <= 1e-6
            err = np.linalg.norm(res[i].magnE[gm] - ref["magnE"][gm]) / np.linalg.norm(ref["magnE"][gm])
            assert err < 1e-4
            assert float(res[i].metrics[0]) == pytest.approx(float(ref["magnE"][gm].max()), rel=1e-3)
        mean, var, count = ctx.uq_moments()
        assert count == 5
        stack = np.stack([r.magnE.astype(np.float64) for r in res])
        np.testing.assert_allclose(mean, stack.mean(axis=0), rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(var, stack.var(axis=0, ddof=1), rtol=1e-6, atol=1e-9 * stack.max() ** 2)
        # The initial guess (run.py:154-200) changes where a draw starts, not where it lands: from
        # zero the same draws reach the same field, in more iterations.
        info = ctx.uq_guess_info()
        assert info["rank"] == len(tags) and info["nominal_iterations"] > 0
        cold = ctx.uq_solve(sigma, initial_guess=False)
        for a, b in zip(res, cold):
            assert np.linalg.norm(a.magnE - b.magnE) <= 2e-5 * np.linalg.norm(b.magnE)
        assert sum(r.iterations for r in res) < sum(r.iterations for r in cold)
        assert res[0].iterations <= 2  # the nominal draw starts at x_nom
    finally:
        ctx.free()



@pytest.mark.gpu
def test_per_placement_conductivities_match_fp64_solves(head4, coil, oracle_mod):
    """solve_placements_block(sigma=...): every placement of a block with its own conductivities,
    against the FP64 oracle built with that placement's table; the nominal row reproduces the
    plain solve."""
    from cunibs_amd import fem, synthetic
    ctx = fem.SolverContext(head4.nodes_mm, head4.tet_nodes, head4.tet_tags, head4.skin_tris)
    try:
        sites = synthetic.cap_sites(3, seed=9)
        overrides = [None, {1: 0.2, 2: 0.4, 4: 0.006}, {3: 1.2, 5: 0.3}]
        rows = fem.conductivity_rows(overrides)
        res = fem.solve_placements_block(ctx, coil.positions_m, coil.moments, sites, 1e6, sigma=rows)
        plain = fem.solve_placements_block(ctx, coil.positions_m, coil.moments, sites[:1], 1e6)[0]
        assert np.linalg.norm(res[0].magnE - plain.magnE) <= 2e-5 * np.linalg.norm(plain.magnE)
        gm = ctx.tet_tags == 2
        for i in (1, 2):
            table = dict(fem.TISSUE_CONDUCTIVITY)
            table.update(overrides[i])
            octx = oracle_mod.build_context(head4.nodes_mm, head4.tet_nodes, head4.tet_tags, head4.skin_tris,
                                            conductivity=table)
            ref = oracle_mod.solve_site(octx, coil.positions_m, coil.moments, sites[i], 1e6, tol=1e-12)
            assert res[i].relative_residual <= 1e-6
            err = np.linalg.norm(res[i].magnE[gm] - ref["magnE"][gm]) / np.linalg.norm(ref["magnE"][gm])
            assert err < 1e-4
        with pytest.raises(ValueError):
            bad = rows.copy()
            bad[1, 3] = -1.0
            fem.solve_placements_block(ctx, coil.positions_m, coil.moments, sites, 1e6, sigma=bad)
    finally:
        ctx.free()
