"""Harmonic recovery of the potential: the oracle pinned by exactness on harmonic polynomials
(the reference's tests/test_recovery.py:664-805 invariants), and the GPU operator and apply
against it."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import kuhn_grid


@pytest.fixture(scope="module")
def two_tissue(oracle_mod):
    nodes, tets = kuhn_grid(7, 6, 6, spacing=5.0, jitter=0.15, seed=1)
    tags = np.where(nodes[tets].mean(axis=1)[:, 0] < 17.5, 1, 2).astype(np.int32)
    return nodes, tets, tags, oracle_mod.harmonic_operator(nodes, tets, tags)


def potential_and_field(nodes_mm, kind):
    x, y, z = (nodes_mm * 1e-3).T
    if kind == "linear":
        return 3 * x - y + 2 * z, np.stack([-3 + 0 * x, 1 + 0 * x, -2 + 0 * x], 1)
    if kind == "harmonic_quadratic":
        return 50 * (x * x - y * y) + 30 * x * z, -np.stack([100 * x + 30 * z, -100 * y, 30 * x], 1)
    if kind == "harmonic_cubic":
        v = 4000 * x * y * z + 900 * x * (x * x - 3 * y * y)
        g = np.stack([4000 * y * z + 2700 * (x * x - y * y), 4000 * x * z - 5400 * x * y, 4000 * x * y], 1)
        return v, -g
    if kind == "nonharmonic":
        return 50 * x * x, -np.stack([100 * x, 0 * x, 0 * x], 1)
    raise ValueError(kind)


@pytest.mark.parametrize("kind, rungs", [("linear", (0, 1, 2)), ("harmonic_quadratic", (0, 1)),
                                          ("harmonic_cubic", (0,))])
def test_exact_on_harmonic_polynomials(two_tissue, oracle_mod, kind, rungs):
    """A harmonic polynomial of degree <= the rung's is in its fit space, so the recovered field
    at the node equals -grad v there (with dA/dt = 0)."""
    nodes, tets, tags, op = two_tissue
    v, e = potential_and_field(nodes, kind)
    slots, _, _ = oracle_mod.harmonic_apply(op, v, np.zeros((nodes.shape[0], 3)))
    sel = np.isin(op["status"], rungs)
    assert sel.sum() > 50
    want = e[op["slot_node"][sel]]
    np.testing.assert_allclose(slots[sel], want, atol=2e-5 * np.abs(want).max())


def test_not_exact_on_a_nonharmonic_quadratic(two_tissue, oracle_mod):
    """The harmonic restriction is real: x^2 is outside the space, so interior cubic-rung slots do
    not reproduce its gradient."""
    nodes, tets, tags, op = two_tissue
    v, e = potential_and_field(nodes, "nonharmonic")
    slots, _, _ = oracle_mod.harmonic_apply(op, v, np.zeros((nodes.shape[0], 3)))
    sel = op["status"] == 0
    err = np.abs(slots[sel] - e[op["slot_node"][sel]]).max()
    assert err > 1e-3 * np.abs(e).max()


def test_amplification_bound_and_patches(two_tissue):
    """solver/fem/recovery.cu:44: h max_c sum_m |w| <= 50 on every fitted slot; patches are sorted,
    distinct, contain their own node and hold at least 16 nodes or the grown 2-ring."""
    nodes, tets, tags, op = two_tissue
    pptr, pidx, w3 = op["pptr"], op["pidx"], op["w3"].astype(np.float64)
    for s in range(op["status"].size):
        b, e = pptr[s], pptr[s + 1]
        patch = pidx[b:e]
        assert np.all(np.diff(patch) > 0) and op["slot_node"][s] in patch
        if op["status"][s] < 3:
            h = np.abs(nodes[patch] * 1e-3 - nodes[op["slot_node"][s]] * 1e-3).max()
            assert h * np.abs(w3[b:e]).sum(axis=0).max() <= 50.0 * (1 + 1e-5)
    # Slots on the interface carry two tissues: one slot each.
    assert op["slot_node"].size > nodes.shape[0]


def test_dadt_enters_exactly(two_tissue, oracle_mod):
    """E = -grad v - dA/dt(x_node): the nodal dA/dt is subtracted as given."""
    nodes, tets, tags, op = two_tissue
    dadt = np.random.default_rng(0).normal(size=(nodes.shape[0], 3))
    slots, _, _ = oracle_mod.harmonic_apply(op, np.zeros(nodes.shape[0]), dadt)
    np.testing.assert_array_equal(slots, -dadt[op["slot_node"]])


# ------------------------------------------------------------------------------------------------
# GPU
# ------------------------------------------------------------------------------------------------

@pytest.fixture(scope="module")
def gctx8(head8):
    from cunibs_amd import fem
    ctx = fem.SolverContext(head8.nodes_mm, head8.tet_nodes, head8.tet_tags, head8.skin_tris)
    yield ctx
    ctx.free()


@pytest.mark.gpu
def test_gpu_harmonic_operator_matches_the_oracle(gctx8, oracle_mod):
    ctx = gctx8
    op = ctx.recovery_operator("harmonic")
    want = oracle_mod.harmonic_operator(ctx.nodes_mm, ctx.tet_nodes, ctx.tet_tags)
    np.testing.assert_array_equal(op["slot_node"], want["slot_node"])
    np.testing.assert_array_equal(op["slot_tag"], want["slot_tag"])
    np.testing.assert_array_equal(op["slot_of_corner"], want["slot_of_corner"])
    np.testing.assert_array_equal(op["pptr"], want["pptr"])
    np.testing.assert_array_equal(op["pidx"], want["pidx"])
    np.testing.assert_array_equal(op["status"], want["status"])
    assert np.linalg.norm(op["w3"] - want["w3"]) <= 1e-5 * np.linalg.norm(want["w3"])


@pytest.mark.gpu
@pytest.mark.parametrize("k", [3, 16])
def test_gpu_harmonic_field_matches_the_oracle(gctx8, oracle_mod, coil, k):
    """The block solve with harmonic recovery against the oracle's apply on the GPU's own
    potential and nodal dA/dt; at k = 3 each workgroup's patch entries exceed its LDS stage and
    are read from global memory, at k = 16 they are staged."""
    from cunibs_amd import fem, synthetic
    ctx = gctx8
    sites = synthetic.cap_sites(k, seed=6)
    extra = {"dadt_nodes": 1}
    res = fem.solve_placements_block(ctx, coil.positions_m, coil.moments, sites, 1e6, potential=True,
                                     vectors=True, recovery="harmonic", slots=True, extra=extra)
    op = oracle_mod.harmonic_operator(ctx.nodes_mm, ctx.tet_nodes, ctx.tet_tags)
    gm = ctx.tet_tags == 2
    for i, r in enumerate(res):
        slots, e_out, mag = oracle_mod.harmonic_apply(op, r.v, extra["dadt_nodes"][i])
        assert np.linalg.norm(r.E_slots - slots) <= 1e-5 * np.linalg.norm(slots)
        assert np.linalg.norm(r.magnE - mag) <= 1e-5 * np.linalg.norm(mag)
        assert float(r.metrics[0]) == pytest.approx(float(r.magnE[gm].max()))
