"""Reduced-basis warm start: the initial guess is the Galerkin projection W P W^T b onto the span
of collected solutions (BF16 basis, MFMA expansion), PCG still iterates to the same tolerance,
and the iteration count drops."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bf16_round(x: np.ndarray) -> np.ndarray:
    """FP32 -> BF16 (round to nearest even) -> FP32, the basis' storage rounding."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


@pytest.fixture(scope="module")
def trained(head4, coil):
    from cunibs_amd import fem, synthetic
    ctx = fem.SolverContext(head4.nodes_mm, head4.tet_nodes, head4.tet_tags, head4.skin_tris)
    ctx.warm_reset(64)
    train = synthetic.cap_sites(48, seed=11)
    sols = []
    for i in range(0, 48, 16):
        res = fem.solve_placements_block(ctx, coil.positions_m, coil.moments, train[i:i + 16], 1e6,
                                         potential=True, collect=True)
        sols.extend(r.v for r in res)
    assert ctx.warm_count() == 48
    rank = ctx.warm_finalize()
    yield ctx, train, np.array(sols), rank
    ctx.free()


def test_modes_are_a_orthonormal_up_to_bf16(trained):
    """Phi^T A Phi = I up to the modes' rounding (BF16 hi + lo pairs), which P removes."""
    ctx, _, _, rank = trained
    g = ctx.warm_state["G_modes"]
    assert rank >= 40
    assert np.abs(g - np.eye(rank)).max() < 1e-2
    np.testing.assert_allclose(ctx.warm_state["P"] @ g, np.eye(rank), atol=1e-6)


@pytest.mark.parametrize("k", [4, 20])
def test_initial_guess_is_the_galerkin_projection(trained, coil, k):
    """x0 = Phi P Phi^T b, evaluated on the host from the device's modes, and the MFMA expansion of
    the coefficients agrees with it; a block of 20 pads to 32 and goes through in two halves of 16
    columns."""
    from cunibs_amd import fem, synthetic
    ctx, _, _, rank = trained
    keep = np.arange(ctx.n_nodes) != ctx.ground_node
    phi = ctx.warm_modes()[:, :rank].astype(np.float64)
    p = ctx.warm_state["P"]
    sites = synthetic.cap_sites(k, seed=12)
    extra = {"x0": 1, "rhs": 1}
    fem.solve_placements_block(ctx, coil.positions_m, coil.moments, sites, 1e6, warm_start=True, extra=extra)
    coefs = []
    for i in range(k):
        b = extra["rhs"][i][keep].astype(np.float64)
        c = p @ (phi.T @ b)
        coefs.append(c)
        want = phi @ c
        assert np.linalg.norm(extra["x0"][i] - want) <= 1e-4 * np.linalg.norm(want)
    # Supplied coefficients (warm_start = 2) expand to the same guess.
    ctx.warm_set_coefficients(np.array(coefs))
    extra2 = {"x0": 1}
    fem.solve_placements_block(ctx, coil.positions_m, coil.moments, sites, 1e6, warm_start=2, extra=extra2)
    assert np.linalg.norm(extra2["x0"] - extra["x0"]) <= 1e-5 * np.linalg.norm(extra["x0"])


def test_warm_start_keeps_the_answer_and_saves_iterations(trained, coil):
    from cunibs_amd import fem, synthetic
    ctx, train, _, _ = trained
    sites = synthetic.cap_sites(8, seed=13)
    cold = fem.solve_placements_block(ctx, coil.positions_m, coil.moments, sites, 1e6)
    warm = fem.solve_placements_block(ctx, coil.positions_m, coil.moments, sites, 1e6, warm_start=True)
    for a, b in zip(cold, warm):
        assert b.relative_residual <= 1e-6
        assert np.linalg.norm(a.magnE - b.magnE) <= 2e-5 * np.linalg.norm(a.magnE)
    assert np.mean([r.iterations for r in warm]) < np.mean([r.iterations for r in cold])
    # A placement whose solution is in the basis starts within the modes' rounding of the answer.
    again_c = fem.solve_placements_block(ctx, coil.positions_m, coil.moments, train[:4], 1e6)
    again_w = fem.solve_placements_block(ctx, coil.positions_m, coil.moments, train[:4], 1e6, warm_start=True)
    assert max(r.iterations for r in again_w) <= min(r.iterations for r in again_c) - 5


def test_guess_stage_timing(trained, coil):
    """cn_bench_guess times the projection, the coefficients and the expansion of the last
    block's right-hand sides, and leaves the warm-started answer unchanged."""
    from cunibs_amd import fem, synthetic
    ctx, _, _, _ = trained
    sites = synthetic.cap_sites(4, seed=14)
    before = fem.solve_placements_block(ctx, coil.positions_m, coil.moments, sites, 1e6, warm_start=True)
    t = ctx.bench_guess(4, reps=3)
    assert set(t) == {"project_ms", "coefficients_ms", "expand_ms"}
    assert all(v > 0 for v in t.values())
    after = fem.solve_placements_block(ctx, coil.positions_m, coil.moments, sites, 1e6, warm_start=True)
    for a, b in zip(before, after):
        np.testing.assert_array_equal(a.magnE, b.magnE)


def test_guess_energy_errors_match_a_direct_evaluation(trained, coil):
    """cn_warm_project_last gives Phi^T b of the last block; from it, P and Phi^T A Phi the
    energy-norm error of the Galerkin guess and of supplied coefficients follows without the
    guess itself. Checked against ||x - x0||_A / ||x||_A evaluated on the host with the FP64
    reduced operator, the solution and the device's own x0."""
    import scipy.sparse as sp
    from cunibs_amd import fem, synthetic
    ctx, _, _, rank = trained
    keep = np.arange(ctx.n_nodes) != ctx.ground_node
    rp, rc, rv = ctx.reduced()
    a = sp.csr_matrix((rv, rc, rp), shape=(ctx.n_nodes - 1,) * 2)
    sites = synthetic.cap_sites(6, seed=15)
    for mode, coef in (("galerkin", None), ("supplied", np.random.default_rng(3).normal(size=(6, rank)))):
        if coef is not None:
            ctx.warm_set_coefficients(coef)
        extra = {"x0": 1, "rhs": 1}
        res = fem.solve_placements_block(ctx, coil.positions_m, coil.moments, sites, 1e6, potential=True,
                                         warm_start=1 if coef is None else 2, extra=extra)
        energy = np.array([r.v[keep] @ b[keep].astype(np.float64) for r, b in zip(res, extra["rhs"])])
        got = ctx.warm_guess_errors(6, energy, coef)
        for i, r in enumerate(res):
            d = r.v[keep] - extra["x0"][i]
            want = np.sqrt(d @ (a @ d) / (r.v[keep] @ (a @ r.v[keep])))
            assert abs(got[i] - want) <= 2e-3 * want + 1e-4, (mode, i, got[i], want)
