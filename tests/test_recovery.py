"""Superconvergent patch recovery: the oracle pinned on the CPU (the reference's
tests/test_recovery.py:17-145 invariants and its NumPy transcription of SimNIBS's
elm_data2node_data), then the HIP operator and apply against it."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import kuhn_grid


def simnibs_spr(nodes, tets, values, is_boundary, tags=None, slots=None):
    """SimNIBS's recovery in its absolute basis [1, x, y, z] (tests/test_recovery.py:17-48 of the
    reference, restated): a least-squares linear fit per patch evaluated at the node, the
    volume-weighted average on the boundary or for fewer than four elements."""
    bary = nodes[tets].mean(axis=1)
    vol = np.abs(np.linalg.det(nodes[tets][:, 1:] - nodes[tets][:, :1])) / 6.0
    incident = [[] for _ in range(nodes.shape[0])]
    for e, tet in enumerate(tets):
        for v in tet:
            incident[v].append(e)
    if slots is None:
        slots = [(v, None) for v in range(nodes.shape[0])]
    out = np.zeros((len(slots), 3))
    for s, (v, tag) in enumerate(slots):
        ets = np.array(incident[v], dtype=np.int64)
        if tag is not None:
            ets = ets[tags[ets] == tag]
        if ets.size == 0:
            continue
        if is_boundary[v] or ets.size < 4:
            out[s] = (values[ets] * vol[ets, None]).sum(0) / vol[ets].sum()
            continue
        design = np.hstack([np.ones((ets.size, 1)), bary[ets]])
        coef = np.linalg.solve(design.T @ design, design.T @ values[ets])
        out[s] = np.hstack([[1.0], nodes[v]]) @ coef
    return out


@pytest.fixture(scope="module")
def cube(oracle_mod):
    nodes, tets = kuhn_grid(5, 5, 5, spacing=8.0, jitter=0.15, seed=2)
    tags = np.where(nodes[tets].mean(axis=1)[:, 0] < 20.0, 1, 2).astype(np.int32)  # a flat interface
    _, vols = oracle_mod.p1_gradients(nodes, tets)
    return nodes, tets, tags, vols.astype(np.float32)


def test_outer_boundary_is_the_face_count(cube, oracle_mod):
    from cunibs_amd.synthetic import _boundary_faces
    nodes, tets, _, _ = cube
    want = np.zeros(nodes.shape[0], np.int32)
    want[np.unique(_boundary_faces(tets, nodes))] = 1
    np.testing.assert_array_equal(oracle_mod.outer_boundary_nodes(tets, nodes.shape[0]), want)
    grid = np.round(nodes / 8.0).astype(int)
    on_box = ((grid == 0) | (grid == 5)).any(axis=1)
    np.testing.assert_array_equal(want.astype(bool), on_box)


@pytest.mark.parametrize("mode", ["spr_global", "spr_tissue"])
def test_weights_are_a_partition_of_unity(cube, oracle_mod, mode):
    """tests/test_recovery.py:70-81: sum_c w_c = 1 for fits and averages alike."""
    nodes, tets, tags, vols = cube
    op = oracle_mod.spr_operator(nodes, tets, tags, vols, mode)
    sums = np.add.reduceat(op["w"].astype(np.float64), op["ptr"][:-1])
    np.testing.assert_allclose(sums, 1.0, atol=1e-6)


def test_recovery_reproduces_a_linear_field(cube, oracle_mod):
    """tests/test_recovery.py:84-110: a linear field is in the fit space, so interior slots and
    elements whose corners are all interior return it exactly."""
    nodes, tets, tags, vols = cube
    op = oracle_mod.spr_operator(nodes, tets, tags, vols, "spr_global")
    bary = nodes[tets].mean(axis=1)
    rng = np.random.default_rng(0)
    offset, slope = rng.normal(size=3), rng.normal(size=(3, 3))
    field = offset + bary @ slope.T
    slots, e_out, _ = oracle_mod.spr_apply(op, field.astype(np.float32))
    interior = op["boundary"] == 0
    scale = np.abs(field).max()
    np.testing.assert_allclose(slots[interior], offset + nodes[interior] @ slope.T, atol=1e-5 * scale)
    inner = interior[tets].all(axis=1)
    assert inner.sum() > 0
    np.testing.assert_allclose(e_out[inner], field[inner], atol=1e-5 * scale)


def test_global_mode_matches_the_simnibs_transcription(cube, oracle_mod):
    """The centred, scaled fit equals the absolute-basis one (tests/test_recovery.py:113-125)."""
    nodes, tets, tags, vols = cube
    op = oracle_mod.spr_operator(nodes, tets, tags, vols, "spr_global")
    field = np.random.default_rng(1).normal(size=(tets.shape[0], 3)).astype(np.float32)
    slots, _, _ = oracle_mod.spr_apply(op, field)
    want = simnibs_spr(nodes, tets, field.astype(np.float64), op["boundary"].astype(bool))
    np.testing.assert_allclose(slots, want, atol=2e-5, rtol=2e-5)


def test_tissue_mode_matches_the_cropped_transcription(cube, oracle_mod):
    """spr_tissue is SimNIBS's per-tissue path: patches cropped to one tag, interface nodes on each
    tissue's boundary (fem/recovery.py:144-176)."""
    nodes, tets, tags, vols = cube
    op = oracle_mod.spr_operator(nodes, tets, tags, vols, "spr_tissue")
    slots_list = list(zip(op["slot_node"].tolist(), op["slot_tag"].tolist()))
    field = np.random.default_rng(3).normal(size=(tets.shape[0], 3)).astype(np.float32)
    slots, _, _ = oracle_mod.spr_apply(op, field)
    want = simnibs_spr(nodes, tets, field.astype(np.float64), op["boundary"].astype(bool), tags=tags,
                       slots=slots_list)
    np.testing.assert_allclose(slots, want, atol=2e-5, rtol=2e-5)
    # Interface nodes carry two slots and are boundary nodes of both tissues.
    interface = np.unique(op["slot_node"][:-1][op["slot_node"][1:] == op["slot_node"][:-1]])
    assert interface.size > 0 and op["boundary"][interface].all()


# ------------------------------------------------------------------------------------------------
# GPU: the library's operator and apply against the oracle
# ------------------------------------------------------------------------------------------------

@pytest.fixture(scope="module")
def gpu_head(head4):
    from cunibs_amd import fem
    ctx = fem.SolverContext(head4.nodes_mm, head4.tet_nodes, head4.tet_tags, head4.skin_tris)
    yield ctx
    ctx.free()


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["spr_global", "spr_tissue"])
def test_gpu_operator_matches_the_oracle(gpu_head, oracle_mod, mode):
    ctx = gpu_head
    op = ctx.recovery_operator(mode)
    want = oracle_mod.spr_operator(ctx.nodes_mm, ctx.tet_nodes, ctx.tet_tags, ctx.vols, mode)
    np.testing.assert_array_equal(op["boundary"], want["boundary"])
    np.testing.assert_array_equal(op["ptr"], want["ptr"])
    np.testing.assert_array_equal(op["idx"], want["idx"])
    np.testing.assert_array_equal(op["slot_of_corner"], want["slot_of_corner"])
    np.testing.assert_array_equal(op["slot_node"], want["slot_node"])
    if mode == "spr_tissue":
        np.testing.assert_array_equal(op["slot_tag"], want["slot_tag"])
    np.testing.assert_array_equal(op["status"], want["status"])
    np.testing.assert_allclose(op["w"], want["w"], rtol=1e-5, atol=1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["spr_global", "spr_tissue"])
def test_gpu_recovered_field_matches_the_oracle(gpu_head, oracle_mod, coil, mode):
    """The block solve with recovery: E_slots and the recovered |E| against the oracle's apply on
    the GPU's own raw field, and the metrics summarise the recovered field."""
    from cunibs_amd import fem, synthetic
    ctx = gpu_head
    sites = synthetic.cap_sites(3, seed=4)
    raw = fem.solve_placements_block(ctx, coil.positions_m, coil.moments, sites, 1e6, vectors=True)
    rec = fem.solve_placements_block(ctx, coil.positions_m, coil.moments, sites, 1e6, vectors=True,
                                     recovery=mode, slots=True)
    op = oracle_mod.spr_operator(ctx.nodes_mm, ctx.tet_nodes, ctx.tet_tags, ctx.vols, mode)
    gm = ctx.tet_tags == 2
    for r0, r1 in zip(raw, rec):
        slots, e_out, mag = oracle_mod.spr_apply(op, r0.E)
        assert np.linalg.norm(r1.E_slots - slots) <= 1e-5 * np.linalg.norm(slots)
        assert np.linalg.norm(r1.E - e_out) <= 1e-5 * np.linalg.norm(e_out)
        assert np.linalg.norm(r1.magnE - mag) <= 1e-5 * np.linalg.norm(mag)
        assert float(r1.metrics[0]) == pytest.approx(float(r1.magnE[gm].max()))
