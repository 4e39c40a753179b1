"""HIP library against the CPU oracle on the same seeded inputs (4 mm five-shell head, 56k nodes).

Integer and index work (layout, CSR pattern, aggregates) and the FP64 assembly must be bit-exact;
FP32 products within the stated tolerances; the end-to-end GM |E| within 1e-3 of the FP64 oracle
(CG to 1e-12), the bar the throughput metric is quoted at. Every call goes through the C ABI.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GM_ERR_BAR = 1e-3


def rel(a, b) -> float:
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def octx(head4, oracle_mod):
    return oracle_mod.build_context(head4.nodes_mm, head4.tet_nodes, head4.tet_tags, head4.skin_tris)


@pytest.fixture(scope="module")
def gctx(head4):
    from cunibs_amd import fem
    ctx = fem.SolverContext(head4.nodes_mm, head4.tet_nodes, head4.tet_tags, head4.skin_tris, keep_stiffness=True)
    yield ctx
    ctx.free()


def test_mesh_layout_is_bit_exact(gctx, octx):
    """fem/solve.py:645-662: Morton node order, tets by lowest renumbered node."""
    np.testing.assert_array_equal(gctx.node_perm, octx.node_perm)
    np.testing.assert_array_equal(gctx.tet_perm, octx.tet_perm)
    np.testing.assert_array_equal(gctx.nodes_mm, octx.nodes_mm)
    np.testing.assert_array_equal(gctx.tet_nodes, octx.tet_nodes)
    np.testing.assert_array_equal(gctx.tet_tags, octx.tet_tags)
    np.testing.assert_array_equal(gctx.skin_tris, octx.skin_tris)


def test_stiffness_is_bit_exact(gctx, octx):
    ip, ii, dd = gctx.stiffness()
    np.testing.assert_array_equal(ip, octx.indptr)
    np.testing.assert_array_equal(ii, octx.indices)
    np.testing.assert_array_equal(dd, octx.data)


def test_grounded_operator_is_bit_exact(gctx, octx):
    rp, rc, rv = gctx.reduced()
    assert gctx.ground_node == octx.ground
    np.testing.assert_array_equal(rp, octx.reduced.ptr)
    np.testing.assert_array_equal(rc, octx.reduced.col)
    np.testing.assert_array_equal(rv, octx.reduced_val64)


def test_amg_hierarchy_is_bit_exact(gctx, octx):
    """Aggregates, l1 diagonals, prolongators and Galerkin operators of every level."""
    assert gctx.hierarchy_sizes == octx.hier.sizes
    for i, lv in enumerate(octx.hier.levels):
        g = gctx.level(i)
        np.testing.assert_array_equal(g["agg"], lv.agg)
        np.testing.assert_array_equal(g["a_ptr"], lv.a.ptr)
        np.testing.assert_array_equal(g["a_col"], lv.a.col)
        np.testing.assert_array_equal(g["a_val"], lv.a.val.astype(np.float32))
        np.testing.assert_array_equal(g["dinv"], lv.dinv)
        np.testing.assert_array_equal(g["p_col"], lv.p.col)
        np.testing.assert_array_equal(g["p_val"], lv.p.val)
    gc = gctx.level(len(octx.hier.levels))
    np.testing.assert_array_equal(gc["a_col"], octx.hier.coarse.col)
    np.testing.assert_array_equal(gc["a_val"], octx.hier.coarse.val)


def test_skin_normals(gctx, octx):
    np.testing.assert_allclose(gctx.skin_normals(), octx.skin_normals, rtol=0, atol=1e-12)


@pytest.mark.parametrize("K", [1, 2, 4, 8, 16, 32])
def test_block_spmm_plain_and_differenced(gctx, octx, oracle_mod, K):
    """FP32 CSR x K products against the FP64 product of the same FP32 matrix."""
    n = octx.reduced.n
    X = np.random.default_rng(K).standard_normal((n, K)).astype(np.float32)
    ref = np.stack([oracle_mod.spmv(octx.reduced, octx.reduced.val.astype(np.float64), X[:, c].astype(np.float64))
                    for c in range(K)], 1)
    assert rel(gctx.spmm(X), ref) < 1e-6
    assert rel(gctx.spmm_diff(X), ref) < 1e-6


@pytest.mark.parametrize("K", [1, 2, 4, 8, 16, 32])
def test_block_grams_are_fp64_exact_products(gctx, octx, K):
    """The iteration's Gram kernels (vector lanes below 16 columns, FP64 matrix cores from 16):
    A^T B1 and A^T B2 of FP32 blocks against the FP64 product; the products of FP32 inputs are
    exact in FP64, so each entry is within 1e-12 of its sum of |terms|. B1 != B2 and A != B
    catch a transposed tile or a swapped operand."""
    n = octx.reduced.n
    rng = np.random.default_rng(7 + K)
    A, B1, B2 = (rng.standard_normal((n, K)).astype(np.float32) for _ in range(3))
    B1[::3] *= 3.0  # rows of unequal weight, so a dropped or repeated row shows
    g1, g2 = gctx.gram(A, B1, B2)
    g = gctx.gram(A, B2)
    for got, B in ((g1, B1), (g2, B2), (g, B2)):
        A64, B64 = A.astype(np.float64), B.astype(np.float64)
        ref = A64.T @ B64
        bound = 1e-12 * (np.abs(A64).T @ np.abs(B64))
        assert np.all(np.abs(got - ref) <= bound)
    assert K == 1 or not np.allclose(g1, g1.T)


@pytest.mark.parametrize("K", [1, 4, 8, 32])
def test_vcycle_against_the_fp64_vcycle(gctx, octx, oracle_mod, K):
    n = octx.reduced.n
    B = np.random.default_rng(100 + K).standard_normal((n, K)).astype(np.float32)
    got = gctx.vcycle(B)
    ref = np.stack([oracle_mod.vcycle_apply(octx.hier, B[:, c].astype(np.float64)) for c in range(K)], 1)
    assert rel(got, ref) < 1e-5


def test_placement_transforms(gctx, octx, oracle_mod):
    from cunibs_amd import synthetic
    sites = synthetic.cap_sites(12, seed=5)
    c = np.array([s[0] for s in sites]); h = np.array([s[1] for s in sites]); d = np.array([s[2] for s in sites])
    tf, deg = gctx.place(c, h, d)
    otf, odeg = oracle_mod.place_transforms(c, h, d, octx.nodes_mm, octx.skin_tris, octx.skin_normals)
    np.testing.assert_array_equal(deg, odeg)
    np.testing.assert_allclose(tf, otf, rtol=0, atol=1e-9)


@pytest.fixture(scope="module")
def single(gctx, octx, oracle_mod, coil):
    from cunibs_amd import fem, synthetic
    site = synthetic.SPHERE_SITE
    ref = oracle_mod.solve_site(octx, coil.positions_m, coil.moments, site, 1e6, tol=1e-12)
    extra = {"dadt_elm": 1, "rhs": 1}
    res = fem.solve_placements_block(gctx, coil.positions_m, coil.moments, [site], 1e6, potential=True,
                                     vectors=True, extra=extra)[0]
    return res, ref, extra


def test_single_placement_end_to_end(single, octx):
    """The 4 mm sphere-head case: dA/dt, RHS, solve, E, |E| against the FP64 oracle."""
    res, ref, extra = single
    assert rel(extra["dadt_elm"][0], ref["dadt_elm"]) < 1e-5
    assert rel(extra["rhs"][0], ref["b"]) < 1e-5
    assert res.relative_residual <= 1e-6
    np.testing.assert_allclose(res.transform, ref["transform"], atol=1e-9)
    gm = octx.tet_tags == 2
    assert rel(res.magnE[gm], ref["magnE"][gm]) < GM_ERR_BAR * 0.1
    assert rel(res.v, ref["v"]) < 1e-3
    assert rel(res.E, ref["E"]) < 1e-3
    np.testing.assert_allclose(np.linalg.norm(res.E.astype(np.float64), axis=1), res.magnE, rtol=1e-5)


def test_true_residual_in_fp64(single, gctx):
    """The reported residual is the true FP64 residual ||b - A x|| / ||b|| of the returned potential."""
    import scipy.sparse as sp
    res, _, extra = single
    rp, rc, rv = gctx.reduced()
    a = sp.csr_matrix((rv, rc, rp), shape=(len(rp) - 1,) * 2)
    keep = np.arange(gctx.n_nodes) != gctx.ground_node
    b = extra["rhs"][0][keep].astype(np.float64)
    x = res.v[keep] - res.v[gctx.ground_node]
    r = np.linalg.norm(b - a @ x) / np.linalg.norm(b)
    assert r <= 1e-6
    assert r == pytest.approx(res.relative_residual, rel=1e-3, abs=1e-12)


def test_gm_metrics_match_the_oracle_summary(single, gctx, oracle_mod):
    """metrics.py:209-243 on the GPU's own |E|: the device reduction equals the NumPy restatement."""
    from cunibs_amd import fem
    res, _, _ = single
    gm = np.nonzero(gctx.tet_tags == 2)[0]
    vols = gctx.vols[gm]
    bary = gctx.tet_barycenters_mm[gm]
    want = oracle_mod.summarize(res.magnE[gm], vols, bary, volume_m3=float(vols.astype(np.float64).sum()))
    got = fem.metrics_dict(res.metrics)
    assert got["peak_magnE"] == want["peak_magnE"]
    np.testing.assert_allclose(got["peak_location_mm"], want["peak_location_mm"], rtol=1e-12)
    # The restatement forms w * |E| in FP32 as the reference does; the device accumulates in FP64.
    np.testing.assert_allclose(got["center_of_gravity_mm"], want["center_of_gravity_mm"], rtol=1e-6)
    assert got["region_volume_m3"] == pytest.approx(want["region_volume_m3"], rel=1e-9)
    for key in ("mean", "std", "p50", "p95", "p99", "p99.9"):
        assert got["distribution"][key] == pytest.approx(want["distribution"][key], rel=1e-6), key
    for key in ("0.5", "0.75"):
        assert got["focality_m3"][key] == pytest.approx(want["focality_m3"][key], rel=1e-6), key


@pytest.fixture(scope="module")
def block8(gctx, coil):
    from cunibs_amd import fem, synthetic
    sites = synthetic.cap_sites(8, seed=3)
    return sites, fem.solve_placements_block(gctx, coil.positions_m, coil.moments, sites, 1e6)


def test_block_of_eight_against_the_oracle(block8, octx, oracle_mod, coil):
    sites, res = block8
    gm = octx.tet_tags == 2
    for i in (0, 5):
        ref = oracle_mod.solve_site(octx, coil.positions_m, coil.moments, sites[i], 1e6, tol=1e-12)
        assert rel(res[i].magnE[gm], ref["magnE"][gm]) < GM_ERR_BAR * 0.1
    assert all(r.relative_residual <= 1e-6 for r in res)


def test_block_metrics_match_the_oracle_per_placement(block8, gctx, oracle_mod):
    """The region summary of a whole block (every placement's |E| sorted together by placement,
    then value) equals the NumPy restatement run on each placement's own |E|."""
    from cunibs_amd import fem
    _, res = block8
    gm = np.nonzero(gctx.tet_tags == 2)[0]
    vols = gctx.vols[gm]
    bary = gctx.tet_barycenters_mm[gm]
    for r in res:
        want = oracle_mod.summarize(r.magnE[gm], vols, bary, volume_m3=float(vols.astype(np.float64).sum()))
        got = fem.metrics_dict(r.metrics)
        assert got["peak_magnE"] == want["peak_magnE"]
        np.testing.assert_allclose(got["peak_location_mm"], want["peak_location_mm"], rtol=1e-12)
        np.testing.assert_allclose(got["center_of_gravity_mm"], want["center_of_gravity_mm"], rtol=1e-6)
        for key in ("mean", "std", "p50", "p95", "p99", "p99.9"):
            assert got["distribution"][key] == pytest.approx(want["distribution"][key], rel=1e-6), key
        for key in ("0.5", "0.75"):
            assert got["focality_m3"][key] == pytest.approx(want["focality_m3"][key], rel=1e-6), key


def test_block_solve_is_bitwise_reproducible(block8, gctx, coil):
    """Fixed-order reductions: the same block gives the same bits."""
    from cunibs_amd import fem
    sites, res = block8
    again = fem.solve_placements_block(gctx, coil.positions_m, coil.moments, sites, 1e6)
    for a, b in zip(res, again):
        np.testing.assert_array_equal(a.magnE, b.magnE)
        np.testing.assert_array_equal(a.metrics, b.metrics)
        assert a.iterations == b.iterations


@pytest.mark.parametrize("k", [1, 3, 5])
def test_partial_blocks_agree_with_the_full_block(block8, gctx, coil, k):
    """Padded widths (3 -> 4, 5 -> 8) and single placements give the same field within the
    block method's tolerance (the reference's PARITY = 2e-5, tests/test_block_solve.py:20)."""
    from cunibs_amd import fem
    sites, res = block8
    sub = fem.solve_placements_block(gctx, coil.positions_m, coil.moments, sites[:k], 1e6)
    for i in range(k):
        assert rel(sub[i].magnE, res[i].magnE) < 2e-5


def test_repeated_placements_in_one_block(block8, gctx, coil):
    """A block that holds one placement three times and another twice has a singular Gram matrix
    from the first iteration: the coefficient step keeps one direction of each repeated set and
    drops the others (their Schur pivots vanish), and every copy still meets the tolerance with
    the field of the block of distinct placements, within the block method's tolerance."""
    from cunibs_amd import fem
    sites, res = block8
    order = [0, 1, 0, 2, 1, 3, 0, 4]
    rep = fem.solve_placements_block(gctx, coil.positions_m, coil.moments, [sites[i] for i in order], 1e6)
    for j, i in enumerate(order):
        assert rep[j].relative_residual <= 1e-6
        assert rel(rep[j].magnE, res[i].magnE) < 2e-5
    assert rel(rep[2].magnE, rep[0].magnE) < 1e-6 and rel(rep[4].magnE, rep[1].magnE) < 1e-6


def test_degenerate_placement_raises(gctx, coil):
    """place.cu:153-230: a handle along the scalp normal names the placement in a ValueError. The
    centre is a scalp node (its own projection) and the handle sits on the frame's normal there."""
    from cunibs_amd import fem
    c = gctx.nodes_mm[gctx.skin_tris[100, 0]]
    tf, deg = gctx.place([c], [c + np.array([0.0, 50.0, 0.0])], [0.0])
    assert deg[0] == 0
    handle = c - 10.0 * tf[0, :3, 2]  # the frame's z axis is the inward normal
    _, deg = gctx.place([c], [handle], [0.0])
    assert deg[0] == 1
    with pytest.raises(ValueError, match="index 0"):
        fem.solve_placements_block(gctx, coil.positions_m, coil.moments, [(c, handle, 4.0)], 1e6)


def test_missed_tolerance_raises(gctx, coil):
    """tests/test_block_solve.py:321-376: an unreachable tolerance within a small iteration cap
    raises SolverConvergenceError (forced from the host; every kernel exits normally)."""
    from cunibs_amd import fem, synthetic
    with pytest.raises(fem.SolverConvergenceError):
        fem.solve_placements_block(gctx, coil.positions_m, coil.moments, [synthetic.SPHERE_SITE], 1e6,
                                   tolerance=1e-14, max_iters=3)


# ------------------------------------------------------------------------------------------------
# Kernel-level entry points (one per reference binding)
# ------------------------------------------------------------------------------------------------

def test_kernel_gradients_morton_node2corner(head4, oracle_mod):
    from cunibs_amd import fem
    g, v = fem.p1_gradients(head4.nodes_mm, head4.tet_nodes)
    og, ov = oracle_mod.p1_gradients(head4.nodes_mm, head4.tet_nodes)
    np.testing.assert_array_equal(g, og)
    np.testing.assert_array_equal(v, ov)
    np.testing.assert_array_equal(fem.morton_keys(head4.nodes_mm), oracle_mod.morton_keys(head4.nodes_mm))
    p, i = fem.node2corner(head4.tet_nodes, head4.nodes_mm.shape[0])
    op, oi = oracle_mod.node2corner(head4.tet_nodes, head4.nodes_mm.shape[0])
    np.testing.assert_array_equal(p, op)
    np.testing.assert_array_equal(i, oi)


def test_kernel_l1_size4_spgemm(octx, oracle_mod):
    from cunibs_amd import fem
    a = octx.hier.levels[0].a
    np.testing.assert_array_equal(fem.l1_dinv(a.n, a.ptr, a.col, a.val), oracle_mod.l1_dinv(a))
    agg, nc = fem.select_size4(a.n, a.ptr, a.col, a.val)
    oagg, onc = oracle_mod.select_size4(a)
    assert nc == onc
    np.testing.assert_array_equal(agg, oagg)
    p = octx.hier.levels[0].p
    cp_, cc, cv = fem.spgemm(a.ptr, a.col, a.val, p.ptr, p.col, p.val, p.n_cols)
    want = oracle_mod.spgemm(a, p)
    np.testing.assert_array_equal(cp_, want.ptr)
    np.testing.assert_array_equal(cc, want.col)
    np.testing.assert_array_equal(cv, want.val)


def test_kernel_dadt_nbody(oracle_mod):
    """solver/coil/dadt.cu:27-104 against the FP64 closed form (tests/test_placement.py:334-346)."""
    from cunibs_amd import fem
    rng = np.random.default_rng(12)
    # Dipoles in a slab above the targets, as a coil sits outside the head (>= 25 mm apart here,
    # so the FP32 |r - s|^2 = |r|^2 + |s|^2 - 2 r.s stays well conditioned).
    s = np.column_stack([rng.normal(0, 0.02, (300, 2)), rng.uniform(0.0, 0.005, 300)])
    m = rng.normal(0, 1e-3, (300, 3))
    r = np.column_stack([rng.normal(0, 0.04, (5000, 2)), rng.uniform(-0.12, -0.025, 5000)])
    mp = np.concatenate([m, np.cross(m, s)], 1)
    sn = (s * s).sum(1)
    got = fem.dadt_nbody(s, mp, sn, r, 1e6)
    assert rel(got, oracle_mod.dadt_direct(s, m, r, 1e6)) < 1e-5


def test_dadt_general_and_parallel_moment_paths(gctx, coil, oracle_mod, octx):
    """The block solve uses the four-term N-body when every moment lies along one direction (the
    synthetic figure-8) and the general six-term one otherwise; both match the FP64 closed form.
    A 1e-6 tilt of the moments sends the same coil down the general path."""
    from cunibs_amd import fem, synthetic
    site = synthetic.SPHERE_SITE
    rng = np.random.default_rng(4)
    tilted = coil.moments + 1e-6 * np.abs(coil.moments).max() * rng.standard_normal(coil.moments.shape)
    out = {}
    for name, mom in (("parallel", coil.moments), ("general", tilted)):
        ex = {"dadt_elm": 1}
        fem.solve_placements_block(gctx, coil.positions_m, mom, [site], 1e6, magnitude=False, metrics=False,
                                   extra=ex)
        ref = oracle_mod.solve_site(octx, coil.positions_m, mom, site, 1e6, tol=1e-8)
        assert rel(ex["dadt_elm"][0], ref["dadt_elm"]) < 1e-5, name
        out[name] = ex["dadt_elm"][0]
    assert rel(out["general"], out["parallel"]) < 1e-5


@pytest.mark.parametrize("warm", [False, True])
def test_pipelined_sweep_equals_blockwise_solves(gctx, coil, warm):
    """cn_solve_sweep overlaps the next block's dA/dt and RHS with the current block's solve; with
    every block starting from zero (recycle=0) its summaries equal those of solving the same blocks
    one at a time, bit for bit (a ragged last block of one placement included, which the solver
    takes at its own narrower padded width)."""
    from cunibs_amd import fem, synthetic
    sites = synthetic.cap_sites(9, seed=17)
    if warm:
        gctx.warm_reset(16)
        fem.solve_placements_block(gctx, coil.positions_m, coil.moments, synthetic.cap_sites(16, seed=18), 1e6,
                                   metrics=False, magnitude=False, collect=True)
        gctx.warm_finalize()
    sweep = fem.solve_sweep(gctx, coil.positions_m, coil.moments, sites, 1e6, block_k=4, warm_start=warm, recycle=0)
    ref = []
    for i in range(0, len(sites), 4):
        ref.extend(fem.solve_placements_block(gctx, coil.positions_m, coil.moments, sites[i:i + 4], 1e6,
                                              magnitude=False, warm_start=warm))
    np.testing.assert_array_equal(sweep["metrics"], np.stack([r.metrics for r in ref]))
    np.testing.assert_array_equal(sweep["iterations"], [r.iterations for r in ref])
    np.testing.assert_array_equal(sweep["relative_residual"], [r.relative_residual for r in ref])
    assert len(sweep["block_iterations"]) == 3
    assert np.all(sweep["relative_residual"] <= 1e-6)


@pytest.mark.parametrize("block_k,n_sites", [(1, 12), (4, 16), (32, 64)])
def test_recycled_sweep_meets_the_tolerance_in_fewer_iterations(gctx, coil, block_k, n_sites):
    """Each full block of a sweep starts from the Galerkin projection of its solution onto the
    previous two blocks' solutions (one block at width 32, whose basis holds 32 vectors at most):
    every placement still meets the FP64 true-residual tolerance, its summary matches the block
    solved from zero to the block method's tolerance, and neighbouring placements (rows of the cap
    grid) need fewer iterations in all."""
    from cunibs_amd import fem, synthetic
    sites = synthetic.grid_sites()[:n_sites]
    rec = fem.solve_sweep(gctx, coil.positions_m, coil.moments, sites, 1e6, block_k=block_k)
    cold = fem.solve_sweep(gctx, coil.positions_m, coil.moments, sites, 1e6, block_k=block_k, recycle=0)
    again = fem.solve_sweep(gctx, coil.positions_m, coil.moments, sites, 1e6, block_k=block_k)
    assert np.all(rec["relative_residual"] <= 1e-6)
    np.testing.assert_array_equal(rec["metrics"], again["metrics"])  # deterministic
    np.testing.assert_array_equal(rec["transforms"], cold["transforms"])
    np.testing.assert_allclose(rec["metrics"][:, 0], cold["metrics"][:, 0], rtol=2e-5)  # peak |E|
    assert rec["block_iterations"][0] == cold["block_iterations"][0]  # the first block has no basis
    assert rec["block_iterations"][1:].sum() < cold["block_iterations"][1:].sum()


def test_graph_replays_equal_direct_launches(block8, gctx, coil):
    """The captured iteration and the same kernels launched directly give the same bits."""
    from cunibs_amd import fem
    sites, res = block8
    direct = fem.solve_placements_block(gctx, coil.positions_m, coil.moments, sites, 1e6, use_graph=False)
    for a, b in zip(res, direct):
        np.testing.assert_array_equal(a.magnE, b.magnE)
        assert a.iterations == b.iterations


def test_failed_capture_falls_back_to_direct_launches(block8, gctx, coil):
    """solver/amg/solver.hpp:88-122: a capture that fails (forced from the host) leaves the solve
    to launch its iterations directly, with the captured path's results; the next solve captures
    again."""
    from cunibs_amd import fem
    sites, res = block8
    gctx.set_tuning(fine_sub=gctx.tuning["fine_sub"])  # drop the captured iterations
    gctx.debug_fail_captures(1)
    fell_back = fem.solve_placements_block(gctx, coil.positions_m, coil.moments, sites, 1e6)
    before = gctx.graph_execs()
    again = fem.solve_placements_block(gctx, coil.positions_m, coil.moments, sites, 1e6)
    assert gctx.graph_execs() > before  # captured this time
    for a, b, c in zip(res, fell_back, again):
        np.testing.assert_array_equal(a.magnE, b.magnE)
        np.testing.assert_array_equal(a.magnE, c.magnE)


def test_graph_cache_is_bounded_across_operator_modes(gctx, coil):
    """Blocks with per-placement conductivities (multi-matrix mode) alternating with nominal
    blocks reuse one captured set per mode: after the first of each, no block captures again."""
    from cunibs_amd import fem, synthetic
    sites = synthetic.cap_sites(8, seed=23)
    rows = fem.conductivity_rows([{2: 0.3 + 0.02 * i} for i in range(4)])
    for rep in range(3):
        fem.solve_placements_block(gctx, coil.positions_m, coil.moments, sites[:4], 1e6, sigma=rows)
        fem.solve_placements_block(gctx, coil.positions_m, coil.moments, sites[4:], 1e6)
        if rep == 0:
            count = gctx.graph_execs()
    assert gctx.graph_execs() == count


def test_ell_layout_gives_the_csr_bits(block8, gctx, coil):
    """The finest level's sliced ELL copy and the CSR arrays give the same bits, at widths 1, 4 and
    16 and through a whole block solve: the slots follow the CSR entries in order and padding adds
    exact zeros. Kuhn rows fit (15 slots with the diagonal, 14 without)."""
    from cunibs_amd import fem
    assert gctx.info.ell_width_full == 15 and gctx.info.ell_width_off == 14
    assert gctx.tuning["layout"] == 0
    sites, res = block8
    rng = np.random.default_rng(8)
    Xs = [rng.standard_normal((gctx.n_nodes - 1, K)).astype(np.float32) for K in (1, 4, 16)]
    out = {}
    for layout in (2, 1, 0):  # ELL at every width, CSR, automatic
        gctx.set_tuning(layout=layout)
        out[layout] = ([(gctx.spmm_diff(X), gctx.vcycle(X)) for X in Xs],
                       fem.solve_placements_block(gctx, coil.positions_m, coil.moments, sites, 1e6))
    for layout in (2, 1):
        for (a1, a2), (b1, b2) in zip(out[layout][0], out[0][0]):
            np.testing.assert_array_equal(a1, b1)
            np.testing.assert_array_equal(a2, b2)
        for a, b in zip(out[layout][1], res):
            np.testing.assert_array_equal(a.magnE, b.magnE)
            assert a.iterations == b.iterations


def test_missed_columns_are_retried_singly_after_a_rebuild(gctx, coil):
    """fem/solve.py:508-552: a block whose columns miss tolerance (forced from the host with an
    iteration cap below what the block needs) rebuilds the hierarchy once and solves each missed
    column again on its own from its iterate; the answers meet the tolerance and agree with the
    uncapped block."""
    from cunibs_amd import fem, synthetic
    sites = synthetic.cap_sites(4, seed=31)
    full = fem.solve_placements_block(gctx, coil.positions_m, coil.moments, sites, 1e6)
    need = max(r.iterations for r in full)
    before = gctx.info.preconditioner_rebuilds
    capped = fem.solve_placements_block(gctx, coil.positions_m, coil.moments, sites, 1e6, max_iters=need - 4)
    assert gctx.info.preconditioner_rebuilds == before + 1
    for a, b in zip(full, capped):
        assert b.relative_residual <= 1e-6
        assert rel(b.magnE, a.magnE) < 2e-5
    assert max(r.iterations for r in capped) > need - 4


def test_unreachable_tolerance_rebuilds_then_raises(gctx, coil):
    """tests/test_block_solve.py:330-354: an unreachable tolerance rebuilds the hierarchy once, then
    raises, naming a column of the block."""
    from cunibs_amd import fem, synthetic
    before = gctx.info.preconditioner_rebuilds
    with pytest.raises(fem.SolverConvergenceError) as err:
        fem.solve_placements_block(gctx, coil.positions_m, coil.moments, synthetic.cap_sites(2, seed=32), 1e6,
                                   tolerance=0.0, max_iters=25)
    assert gctx.info.preconditioner_rebuilds == before + 1
    assert err.value.column in (0, 1)


def test_a_draw_that_misses_gets_its_own_preconditioner(head4, coil):
    """uq/conductivity/run.py:333-342: a conductivity draw that misses tolerance under the nominal
    preconditioner (forced with an iteration cap) is solved again from its iterate under a
    preconditioner built for its own operator (the nominal aggregates, its own prolongators and
    Galerkin products), which is then discarded; the answers agree with the uncapped draws."""
    from cunibs_amd import fem, synthetic, uq
    ctx = fem.SolverContext(head4.nodes_mm, head4.tet_nodes, head4.tet_tags, head4.skin_tris)
    try:
        tags = ctx.uq_prepare(coil.positions_m, coil.moments, synthetic.SPHERE_SITE, 1e6)
        sigma = uq.draw_lognormal(tags, 6, 0.5, seed=33)
        full = ctx.uq_solve(sigma)
        need = max(r.iterations for r in full)
        capped = ctx.uq_solve(sigma, max_iters=need - 3)
        assert ctx.info.matched_preconditioners >= 1
        for a, b in zip(full, capped):
            assert b.relative_residual <= 1e-6
            assert rel(b.magnE, a.magnE) < 2e-5
    finally:
        ctx.free()


@pytest.mark.parametrize("k", [13, 29])
def test_wide_ragged_blocks_against_the_oracle(gctx, octx, oracle_mod, coil, k):
    """Blocks padded to widths 16 and 32 (the drop-in default width and the widest): every
    placement's FP64 true residual, recomputed on the host, within the tolerance; two placements'
    GM |E| within the 1e-3 bar of the FP64 oracle."""
    from cunibs_amd import fem, synthetic
    from conftest import true_residuals
    sites = synthetic.cap_sites(k, seed=40 + k)
    ex = {"rhs": 1}
    res = fem.solve_placements_block(gctx, coil.positions_m, coil.moments, sites, 1e6, potential=True, extra=ex)
    assert max(true_residuals(gctx, res, ex["rhs"])) <= 1e-6
    gm = octx.tet_tags == 2
    for i in (0, k - 1):
        ref = oracle_mod.solve_site(octx, coil.positions_m, coil.moments, sites[i], 1e6, tol=1e-12)
        assert rel(res[i].magnE[gm], ref["magnE"][gm]) < GM_ERR_BAR


def test_degenerate_and_unknown_tetrahedra_are_rejected():
    """fem/solve.py:702-706 and fem/assembly.py:15-42: a tetrahedron of zero volume, and a tissue
    tag the conductivity table does not hold, are rejected with ValueError naming the problem."""
    from cunibs_amd import fem
    from conftest import CUBE_NODES, CUBE_TETS, boundary_tris
    skin = boundary_tris(CUBE_TETS, CUBE_NODES)
    flat = CUBE_TETS.copy()
    flat[2] = [0, 1, 2, 3]  # four coplanar corners of the bottom face
    with pytest.raises(ValueError, match="volume"):
        fem.SolverContext(CUBE_NODES, flat, np.ones(6, np.int32), skin)
    tags = np.ones(6, np.int32)
    tags[4] = 42
    with pytest.raises(ValueError, match="42|tag"):
        fem.SolverContext(CUBE_NODES, CUBE_TETS, tags, skin)


def test_dense_tail_matches_the_level_by_level_vcycle(block8, gctx, coil):
    """The V-cycle's bottom levels applied as one dense FP32 product (formed by running them on the
    identity's columns) agree with the level-by-level cycle to FP32 rounding, and a block solve
    with it converges alike."""
    from cunibs_amd import fem
    sites, res = block8
    assert gctx.tuning["dense_tail"] == 1
    t = gctx.info.dense_tail_level
    assert 0 <= t < gctx.info.n_levels and gctx.info.level_rows[t] <= 4096
    assert t == 0 or gctx.info.level_rows[t - 1] > 4096
    X = np.random.default_rng(21).standard_normal((gctx.n_nodes - 1, 16)).astype(np.float32)
    with_tail = gctx.vcycle(X)
    gctx.set_tuning(dense_tail=0)
    try:
        assert gctx.info.dense_tail_level == -1
        level_by_level = gctx.vcycle(X)
        plain = fem.solve_placements_block(gctx, coil.positions_m, coil.moments, sites, 1e6)
    finally:
        gctx.set_tuning(dense_tail=1)
    assert rel(with_tail, level_by_level) < 2e-6
    for a, b in zip(res, plain):
        assert abs(a.iterations - b.iterations) <= 1
        assert rel(a.magnE, b.magnE) < 2e-5


@pytest.mark.parametrize("K", [1, 4, 32])
def test_dense_tail_at_other_widths(gctx, K):
    """The dense tail's matrix-core product at widths below 16 (columns past K masked) and at 32
    (two 16-column tiles) agrees with the level-by-level cycle."""
    X = np.random.default_rng(40 + K).standard_normal((gctx.n_nodes - 1, K)).astype(np.float32)
    with_tail = gctx.vcycle(X)
    gctx.set_tuning(dense_tail=0)
    try:
        level_by_level = gctx.vcycle(X)
    finally:
        gctx.set_tuning(dense_tail=1)
    assert rel(with_tail, level_by_level) < 2e-6
