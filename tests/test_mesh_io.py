"""Gmsh 2.2 reader and the reference's real-head patch pins (CPU).

The patch (tests/data/head_patch_r25.msh.gz in the reference: 8404 nodes, 41139 tets, 1125 skin
triangles, a 25 mm ball cut from a CHARM head) and its manifest are data files of the reference's
tests, kept as fixtures under tests/golden/."""

from __future__ import annotations

import gzip
import json
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"
PATCH = GOLDEN / "head_patch_r25.msh.gz"
PATCH_JSON = GOLDEN / "head_patch_r25.json"
needs_patch = pytest.mark.skipif(not PATCH.exists(), reason="patch fixture missing")


@needs_patch
def test_patch_loads_with_the_recorded_counts():
    from cunibs_amd.mesh import load_mesh
    meta = json.loads(PATCH_JSON.read_text())
    m = load_mesh(PATCH)
    assert m.n_nodes == meta["n_nodes"]
    assert m.tet_nodes.shape[0] == meta["n_tets"]
    assert m.skin_tris.shape[0] == meta["n_skin_tris"]
    tags, counts = np.unique(m.tet_tags, return_counts=True)
    assert {str(t): int(c) for t, c in zip(tags, counts)} == meta["tet_tags"]


@needs_patch
def test_patch_hierarchy_sizes_match_the_reference(oracle_mod):
    """tests/test_solver_kernels.py:615-640 pins [8403, 465] with two pairwise rounds per level;
    tests/test_block_solve.py:389 pins 8403 -> 1936 -> 465 with one. This pins the FP32 reduced
    stiffness and the SIZE_4 aggregation the GPU reproduces bit for bit."""
    from cunibs_amd.mesh import load_mesh
    m = load_mesh(PATCH)
    ctx = oracle_mod.build_context(m.nodes_mm, m.tet_nodes, m.tet_tags, m.skin_tris)
    assert ctx.hier.sizes == [8403, 465]
    saved = oracle_mod.ROUNDS
    try:
        oracle_mod.ROUNDS = 1
        assert oracle_mod.build_hierarchy(ctx.reduced).sizes == [8403, 1936, 465]
    finally:
        oracle_mod.ROUNDS = saved


@pytest.mark.parametrize("binary", [True, False])
def test_write_then_load_round_trips(tmp_path, head8, binary):
    from cunibs_amd.mesh import HeadMesh, load_mesh, write_msh
    m = HeadMesh(head8.nodes_mm, head8.tet_nodes, head8.tet_tags, head8.skin_tris)
    path = tmp_path / "head.msh"
    write_msh(path, m, binary=binary)
    back = load_mesh(path)
    np.testing.assert_array_equal(back.nodes_mm, m.nodes_mm)
    np.testing.assert_array_equal(back.tet_nodes, m.tet_nodes)
    np.testing.assert_array_equal(back.tet_tags, m.tet_tags)
    np.testing.assert_array_equal(back.skin_tris, m.skin_tris)
    gz = tmp_path / "head.msh.gz"
    gz.write_bytes(gzip.compress(path.read_bytes()))
    np.testing.assert_array_equal(load_mesh(gz).tet_nodes, m.tet_nodes)


def test_unused_nodes_are_dropped_and_indices_rebased(tmp_path):
    from cunibs_amd.mesh import HeadMesh, load_mesh, write_msh
    nodes = np.array([[0.0, 0, 0], [9, 9, 9], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    m = HeadMesh(nodes, [[0, 2, 3, 4]], [2], [[0, 2, 3]])
    write_msh(tmp_path / "a.msh", m)
    back = load_mesh(tmp_path / "a.msh")
    np.testing.assert_array_equal(back.nodes_mm, nodes[[0, 2, 3, 4]])
    np.testing.assert_array_equal(back.tet_nodes, [[0, 1, 2, 3]])
    np.testing.assert_array_equal(back.skin_tris, [[0, 1, 2]])


def test_unknown_tissue_tags_and_truncation_are_rejected(tmp_path):
    """mesh.py:164-181: unknown volume tags raise instead of deleting a tissue; a cut file raises."""
    from cunibs_amd.mesh import HeadMesh, load_mesh, write_msh
    nodes = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    write_msh(tmp_path / "bad.msh", HeadMesh(nodes, [[0, 1, 2, 3]], [42], np.empty((0, 3))))
    with pytest.raises(ValueError, match="unknown volume tags"):
        load_mesh(tmp_path / "bad.msh")
    write_msh(tmp_path / "ok.msh", HeadMesh(nodes, [[0, 1, 2, 3]], [2], np.empty((0, 3))))
    data = (tmp_path / "ok.msh").read_bytes()
    (tmp_path / "cut.msh").write_bytes(data[: len(data) - 40])
    with pytest.raises(ValueError):
        load_mesh(tmp_path / "cut.msh")
    (tmp_path / "v4.msh").write_bytes(b"$MeshFormat\n4.1 0 8\n$EndMeshFormat\n")
    with pytest.raises(ValueError, match="2.2"):
        load_mesh(tmp_path / "v4.msh")
