"""End-to-end pins on the reference's real-head patch with the bundled MagStim D70 coil.

The reference pins (tests/test_realmesh.py:25-27, 116-141): at the topmost scalp node of the
patch, handle +50 mm along y, 4 mm off the scalp, dI/dt 1e6 A/s, gray-matter peak |E| 0.556405
V/m at (11.794, -1.125, 77.349) mm with raw recovery and 0.484461 V/m with harmonic recovery, both
to rel 2e-3 (location to 2 mm). The CPU tests hold the FP64 oracle to those numbers; the `gpu` tests
hold the HIP path to the same numbers through the public Subject API. Fixtures: tests/golden/."""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"
PATCH = GOLDEN / "head_patch_r25.msh.gz"
D70 = GOLDEN / "MagStim_D70.h5"

PATCH_PEAK_MAGNE = 0.556405
PATCH_PEAK_MAGNE_HARMONIC = 0.484461
PATCH_PEAK_LOCATION_MM = np.array([11.794, -1.125, 77.349])


@pytest.fixture(scope="module")
def patch_mesh():
    from cunibs_amd.mesh import load_mesh
    return load_mesh(PATCH)


@pytest.fixture(scope="module")
def d70():
    from cunibs_amd.simulation import Coil
    return Coil.load(D70)


@pytest.fixture(scope="module")
def patch_site(patch_mesh):
    """tests/conftest.py:337-349: the topmost scalp node, handle +50 mm in y, 4 mm offset."""
    skin = np.unique(patch_mesh.skin_tris)
    coords = patch_mesh.nodes_mm[skin]
    top = np.ascontiguousarray(coords[np.argmax(coords[:, 2])])
    return top, top + np.array([0.0, 50.0, 0.0]), 4.0


def test_d70_loads_without_h5py(d70):
    """tests/test_coil.py:54-65: the bundled D70 reads back, dIdtmax kept raw and exposed in A/s."""
    assert d70.name == "D70"
    assert d70.positions_m.shape == (964, 3) and d70.moments.shape == (964, 3)
    assert d70.positions_m.dtype == np.float64
    assert d70.metadata["dIdtmax"] == "114.73"
    assert d70.metadata["num_elements"] == 964 and isinstance(d70.metadata["num_elements"], int)
    assert d70.didt_max == pytest.approx(114.73e6)
    assert np.abs(d70.positions_m).max() < 1.0 and np.abs(d70.moments).max() > 0.0


@pytest.fixture(scope="module")
def oracle_patch(oracle_mod, patch_mesh, d70, patch_site):
    ctx = oracle_mod.build_context(patch_mesh.nodes_mm, patch_mesh.tet_nodes, patch_mesh.tet_tags,
                                   patch_mesh.skin_tris)
    r = oracle_mod.solve_site(ctx, d70.positions_m, d70.moments, patch_site, 1e6, tol=1e-10)
    return ctx, r


def test_oracle_raw_peak_matches_the_reference_pin(oracle_mod, oracle_patch):
    ctx, r = oracle_patch
    s = oracle_mod.gm_summary(ctx, r["magnE"])
    assert s["peak_magnE"] == pytest.approx(PATCH_PEAK_MAGNE, rel=2e-3)
    np.testing.assert_allclose(s["peak_location_mm"], PATCH_PEAK_LOCATION_MM, atol=2.0)
    # The oracle lands far inside the reference's tolerance: 6 digits and 1 micrometre.
    assert s["peak_magnE"] == pytest.approx(PATCH_PEAK_MAGNE, rel=2e-6)
    np.testing.assert_allclose(s["peak_location_mm"], PATCH_PEAK_LOCATION_MM, atol=1e-3)


def test_oracle_harmonic_peak_matches_the_reference_pin(oracle_mod, oracle_patch, d70):
    ctx, r = oracle_patch
    op = oracle_mod.harmonic_operator(ctx.nodes_mm, ctx.tet_nodes, ctx.tet_tags)
    dadt = oracle_mod.dadt_at_nodes(d70.positions_m, d70.moments, r["transform"], 1e6, ctx.nodes_mm)
    _, _, mag = oracle_mod.harmonic_apply(op, r["v"], dadt)
    s = oracle_mod.gm_summary(ctx, mag)
    assert s["peak_magnE"] == pytest.approx(PATCH_PEAK_MAGNE_HARMONIC, rel=2e-6)


def _smoothed_normals(nodes_mm, tris, rounds=2):
    """The skin smoothing of solver/fem/skin.cu as the reference's tests describe it
    (tests/test_solver_kernels.py:732-754), restated with a node-by-face incidence matrix C:
    face vectors f start as the unnormalised cross products (twice the area along the normal);
    each round spreads them to the nodes (u = C f) and gathers back normalised sums
    (f = C^T u / |C^T u|); the last node vectors are normalised, and each face takes the
    normalised sum of its three."""
    import scipy.sparse as sp
    tris = np.asarray(tris)
    used, local = np.unique(tris.ravel(), return_inverse=True)
    n_tri = tris.shape[0]
    inc = sp.csr_matrix((np.ones(3 * n_tri), (local, np.repeat(np.arange(n_tri), 3))), shape=(used.size, n_tri))
    p = np.asarray(nodes_mm, np.float64)[tris]
    f = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    unit = lambda v: v / np.linalg.norm(v, axis=1, keepdims=True)
    for _ in range(rounds):
        u = inc @ f
        f = unit(inc.T @ u)
    return unit(inc.T @ unit(u))


def test_oracle_skin_normals_match_the_smoothing_definition(oracle_mod, patch_mesh):
    """On the patch's scalp (tests/test_solver_kernels.py:757-770): the oracle's normals equal the
    restated definition to a few ulp; one smoothing round instead of two moves them by more than
    1e-3, so the tolerance separates a wrong round count, corner or weighting."""
    got = oracle_mod.skin_normals(patch_mesh.nodes_mm, patch_mesh.skin_tris)
    want = _smoothed_normals(patch_mesh.nodes_mm, patch_mesh.skin_tris)
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-15)
    assert np.abs(_smoothed_normals(patch_mesh.nodes_mm, patch_mesh.skin_tris, rounds=1) - want).max() > 1e-3


@pytest.fixture(scope="module")
def patch_sites8(patch_mesh):
    """tests/test_block_solve.py:46-54: eight scalp placements spread over the patch's top, each
    handle 40 mm out along its own azimuth, 4 mm off the scalp."""
    cen = patch_mesh.nodes_mm[patch_mesh.skin_tris].mean(axis=1)
    high = cen[np.argsort(cen[:, 2])[-400:]]
    chosen = high[::max(len(high) // 8, 1)][:8]
    ang = np.linspace(0.0, 2.0 * np.pi, 8, endpoint=False)
    return [(c, c + 40.0 * np.array([np.cos(a), np.sin(a), 0.0]), 4.0) for c, a in zip(chosen, ang)]


def _two_rhs(n):
    """tests/test_block_solve.py:289-293: a two-point source (+1 at n // 3, -1 at 2n // 3) and one
    period of a sine over the reduced unknowns in the solver's order."""
    pts = np.zeros(n)
    pts[n // 3], pts[2 * n // 3] = 1.0, -1.0
    return pts, np.sin(np.arange(n) * (2.0 * np.pi / n))


def test_oracle_pcg_takes_the_reference_iteration_counts(oracle_mod, oracle_patch, d70, patch_sites8):
    """The reference's counts at tolerance 1e-6, Chebyshev degree 2, ratio 12, two SIZE_4 rounds:
    every patch placement converges in 11 iterations (tests/test_block_solve.py:284), the sine in
    12 and the two-point source in 10 (:286-288). These pin the whole V-cycle composition (smoother
    polynomial, restriction, coarse correction, prolongation). Mistakes planted in the oracle's
    V-cycle, on the CPU, move them: one smoothing step fewer or more (degree 1 or 3), or the coarse
    correction left out."""
    import copy
    ctx, _ = oracle_patch
    for site in patch_sites8:
        r = oracle_mod.solve_site(ctx, d70.positions_m, d70.moments, site, 1e6, tol=1e-6)
        assert r["iterations"] == 11 and r["relative_residual"] <= 1e-6
    pts, sine = _two_rhs(ctx.reduced.n)

    def counts(h):
        return tuple(oracle_mod.pcg(h, ctx.reduced, ctx.reduced_val64, b, 1e-6)[1] for b in (pts, sine))

    assert counts(ctx.hier) == (10, 12)
    assert counts(oracle_mod.build_hierarchy(ctx.reduced, degree=1)) == (16, 18)
    assert counts(oracle_mod.build_hierarchy(ctx.reduced, degree=3)) == (9, 10)
    no_coarse = copy.copy(ctx.hier)
    no_coarse.coarse_l = ctx.hier.coarse_l * 1e8  # coarse solve returns ~0: correction dropped
    assert min(counts(no_coarse)) > 60


@pytest.fixture(scope="module")
def patch_subject(patch_mesh):
    from cunibs_amd.simulation import Subject
    s = Subject(patch_mesh)
    yield s
    s.free()


@pytest.mark.gpu
def test_gpu_raw_peak_matches_the_reference_pin(patch_subject, d70, patch_site):
    """tests/test_realmesh.py:116-133."""
    from cunibs_amd.simulation import Placement
    r = patch_subject.simulate(d70, Placement(*patch_site), 1e6, magnitude=True, vectors=True,
                               potential=True, recovery="raw")
    assert r.peak_magnE() == pytest.approx(PATCH_PEAK_MAGNE, rel=2e-3)
    np.testing.assert_allclose(r.peak_location_mm(), PATCH_PEAK_LOCATION_MM, atol=2.0)
    assert r.coil_name == d70.name
    assert r.relative_residual <= 1e-6 and 0 < r.iterations < 200


@pytest.mark.gpu
def test_gpu_harmonic_peak_matches_the_reference_pin(patch_subject, d70, patch_site):
    """tests/test_realmesh.py:136-141: the default recovery is harmonic."""
    from cunibs_amd.simulation import Placement
    r = patch_subject.simulate(d70, Placement(*patch_site), 1e6, magnitude=True)
    assert r.recovery == "harmonic"
    assert r.peak_magnE() == pytest.approx(PATCH_PEAK_MAGNE_HARMONIC, rel=2e-3)


@pytest.mark.gpu
def test_gpu_patch_field_matches_the_oracle(patch_subject, oracle_mod, oracle_patch, d70, patch_site):
    """Element |E| in gray matter against the FP64 oracle on the same placement, to the 1e-3 bar."""
    from cunibs_amd.simulation import Placement
    ctx, ref = oracle_patch
    r = patch_subject.simulate(d70, Placement(*patch_site), 1e6, magnitude=True, recovery="raw")
    # Results are indexed in the Morton order, which the oracle's context shares.
    np.testing.assert_array_equal(patch_subject.mesh.tet_tags, ctx.tet_tags)
    gm = ctx.tet_tags == 2
    ours = np.asarray(r.magnE)[gm]
    theirs = ref["magnE"][gm]
    assert np.abs(ours - theirs).max() <= 1e-3 * np.abs(theirs).max()


@pytest.mark.gpu
def test_gpu_field_result_round_trip(tmp_path, patch_subject, d70, patch_site):
    """tests/test_realmesh.py:262-278: save/load at 41k elements, compressed."""
    from cunibs_amd.simulation import FieldResult, Placement
    r = patch_subject.simulate(d70, Placement(*patch_site), 1e6, magnitude=True, vectors=True, potential=True)
    path = tmp_path / "patch.h5"
    r.save(path)
    loaded = FieldResult.load(path)
    for name in ("E", "magnE", "v", "transform", "vols", "tet_tags", "barycenters_mm"):
        np.testing.assert_array_equal(getattr(loaded, name), getattr(r, name), err_msg=name)
    assert loaded.summary["peak_magnE"] == r.summary["peak_magnE"]
    assert loaded.recovery == "harmonic" and loaded.coil_name == "D70"
    assert path.stat().st_size < 3_000_000


@pytest.mark.gpu
def test_gpu_iteration_counts_match_the_reference(patch_mesh, d70, patch_sites8):
    """tests/test_block_solve.py:284: at the reference's smoother (degree 2, ratio 12) every patch
    placement converges in 11 iterations on the HIP path too, alone and as the reference's
    independent chains in a block of eight."""
    from cunibs_amd import fem
    ctx = fem.SolverContext(patch_mesh.nodes_mm, patch_mesh.tet_nodes, patch_mesh.tet_tags, patch_mesh.skin_tris)
    try:
        ctx.set_tuning(degree=2, lower_ratio=12.0)
        alone = [fem.solve_placements_block(ctx, d70.positions_m, d70.moments, [s], 1e6)[0] for s in patch_sites8]
        chains = fem.solve_placements_block(ctx, d70.positions_m, d70.moments, patch_sites8, 1e6, independent=True)
        assert [r.iterations for r in alone] == [11] * 8
        assert [r.iterations for r in chains] == [11] * 8
        assert all(r.relative_residual <= 1e-6 for r in alone + chains)
    finally:
        ctx.free()
