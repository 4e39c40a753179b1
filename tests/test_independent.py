"""The independent-chains mode (the reference's block solver: k CG chains in lockstep, each column
with its own step lengths and its own freeze; solver/amg/solver.cpp:84-198). The reference's
batch-invariance tests (tests/test_block_solve.py:164-233) hold for it: a placement's field and
telemetry do not depend on which block it was solved in."""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def patch_ctx():
    from cunibs_amd import fem
    from cunibs_amd.mesh import load_mesh
    m = load_mesh(GOLDEN / "head_patch_r25.msh.gz")
    ctx = fem.SolverContext(m.nodes_mm, m.tet_nodes, m.tet_tags, m.skin_tris)
    yield ctx, m
    ctx.free()


@pytest.fixture(scope="module")
def patch_sites(patch_ctx):
    """Eight placements around the patch's top (tests/conftest.py patch_sites there: a ring of
    handles and small shifts of the centre)."""
    _, m = patch_ctx
    skin = np.unique(m.skin_tris)
    c = m.nodes_mm[skin]
    top = c[np.argmax(c[:, 2])]
    sites = []
    for i in range(8):
        a = 2 * np.pi * i / 8
        centre = top + np.array([3.0 * np.cos(a), 3.0 * np.sin(a), 0.0])
        sites.append((centre, centre + 50.0 * np.array([np.sin(a), np.cos(a), 0.0]), 4.0))
    return sites


def _solve(ctx, coil, sites, independent):
    from cunibs_amd import fem
    return fem.solve_placements_block(ctx, coil.positions_m, coil.moments, sites, 1e6, potential=True,
                                      independent=independent)


@pytest.mark.gpu
@pytest.mark.parametrize("k", [2, 3, 4, 5, 8])
def test_block_width_does_not_reach_the_answer(patch_ctx, patch_sites, k):
    """tests/test_block_solve.py:194-209 and :215-232 there: solved k at a time in independent mode,
    every placement's |E|, potential and iteration count equal its solve alone. Up to a padded
    width of 4 every solver kernel sums each column in the same order as at width 1, so the field
    is bitwise equal (the reported residual norm to 1e-9); the width-8 kernels use more lanes per
    row and agree within the solver's tolerance."""
    from cunibs_amd.simulation import Coil
    ctx, _ = patch_ctx
    coil = Coil.load(GOLDEN / "MagStim_D70.h5")
    ref = [_solve(ctx, coil, [s], True)[0] for s in patch_sites]
    got = [r for i in range(0, 8, k) for r in _solve(ctx, coil, patch_sites[i:i + k], True)]
    worst = 0.0
    for i, (g, r) in enumerate(zip(got, ref)):
        worst = max(worst, float(np.abs(g.magnE - r.magnE).max() / np.abs(r.magnE).max()))
        assert g.relative_residual <= 1e-6
        assert np.linalg.norm(g.magnE - r.magnE) <= 2e-5 * np.linalg.norm(r.magnE), f"placement {i}"
        assert abs(g.iterations - r.iterations) <= 1, f"placement {i}"
    print(f"k={k}: worst |E| difference against solving alone {worst:.3e}")
    if k <= 4:
        for i, (g, r) in enumerate(zip(got, ref)):
            np.testing.assert_array_equal(g.magnE, r.magnE, err_msg=f"k={k}, placement {i}")
            np.testing.assert_array_equal(g.v, r.v, err_msg=f"k={k}, placement {i}")
            assert g.iterations == r.iterations
            # the reported FP64 residual norm sums its rows in a width-dependent order
            assert g.relative_residual == pytest.approx(r.relative_residual, rel=1e-9)


@pytest.mark.gpu
def test_independent_and_block_modes_agree(patch_ctx, patch_sites):
    """Both methods meet the same FP64 true-residual bar, so their fields agree to it; the block
    method's shared Krylov space takes no more iterations."""
    from cunibs_amd.simulation import Coil
    ctx, _ = patch_ctx
    coil = Coil.load(GOLDEN / "MagStim_D70.h5")
    ind = _solve(ctx, coil, patch_sites, True)
    blk = _solve(ctx, coil, patch_sites, False)
    for a, b in zip(ind, blk):
        assert a.relative_residual <= 1e-6 and b.relative_residual <= 1e-6
        assert np.linalg.norm(a.magnE - b.magnE) <= 2e-5 * np.linalg.norm(b.magnE)
    assert max(b.iterations for b in blk) <= max(a.iterations for a in ind)
