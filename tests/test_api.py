"""Front-end validation that needs no GPU (the reference's tests/test_api.py:40-136 tier), and the
same front end end to end on the GPU."""

from __future__ import annotations

import numpy as np
import pytest

from cunibs_amd.simulation import (
    Coil,
    FieldResult,
    Placement,
    Subject,
    compute_metrics,
    validate_recovery,
)


def test_placement_validation():
    p = Placement([0, 20, 92], [0, 70, 92])
    assert p.distance_mm == 4.0
    assert p.center_mm.dtype == np.float64 and p.center_mm.shape == (3,)
    with pytest.raises(ValueError):
        Placement([0, 0, 1], [0, 0, 1])
    with pytest.raises(ValueError):
        Placement([0, np.nan, 1], [0, 0, 2])
    with pytest.raises(ValueError):
        Placement([0, 0, 1], [0, 0, 2], distance_mm=np.inf)
    with pytest.raises(ValueError):
        Placement([0, 0], [0, 0, 2])
    with pytest.raises(Exception):
        p.distance_mm = 3.0  # frozen


def test_autotuner_breaks_width_ties_toward_the_narrower_block():
    """Block widths whose rates lie within 2 % of the best count as tied; the narrowest is taken."""
    from cunibs_amd.autotune import pick_width
    assert pick_width({4: 311.2, 8: 394.0, 16: 475.2, 32: 476.4}) == 16
    assert pick_width({4: 300.0, 8: 390.0, 16: 455.0, 32: 476.4}) == 32
    assert pick_width({8: 100.0}) == 8


def test_recovery_modes():
    assert validate_recovery("raw") == "raw"
    with pytest.raises(ValueError):
        validate_recovery("bogus")
    assert validate_recovery("spr_tissue") == "spr_tissue"
    assert validate_recovery("harmonic") == "harmonic"


def test_iter_simulate_checks_arguments_before_any_device_work(head8, coil):
    from cunibs_amd.mesh import HeadMesh
    subj = Subject(HeadMesh(head8.nodes_mm, head8.tet_nodes, head8.tet_tags, head8.skin_tris))
    c = Coil(coil.positions_m, coil.moments, name="fig8")
    p = Placement([0, 20, 92], [0, 70, 92])
    with pytest.raises(TypeError):
        subj.iter_simulate(c, p)  # a single placement, not a sequence
    with pytest.raises(TypeError):
        subj.simulate(c, [p])
    with pytest.raises(ValueError):
        subj.iter_simulate(c, [p], recovery="nope")
    with pytest.raises(ValueError):
        subj.iter_simulate(c, [p], nodal=True, recovery="raw")
    with pytest.raises(TypeError):
        subj.iter_simulate(c, [(1, 2, 3)])
    assert subj._ctx is None


def test_coil_validation_and_ccd_round_trip(tmp_path, coil):
    with pytest.raises(ValueError):
        Coil(np.zeros((3, 3)), np.zeros((2, 3)))
    path = tmp_path / "fig8.ccd"
    data = np.column_stack([coil.positions_m, coil.moments])
    with path.open("w") as f:
        f.write("# coilname=figure8; dIdtmax=150\n")
        f.write(f"{data.shape[0]}\n")
        f.write("# centers and weighted directions\n")
        np.savetxt(f, data, fmt="%.17g")
    c = Coil.load(path)
    assert c.name == "figure8"
    assert c.didt_max == pytest.approx(150e6)
    np.testing.assert_array_equal(c.positions_m, coil.positions_m)
    np.testing.assert_array_equal(c.moments, coil.moments)


def test_host_metrics_golden_values():
    """tests/test_metrics.py:9-80 of the reference through the host metrics of FieldResult."""
    magn = np.array([1.0, 2.0, 3.0, 100.0])
    vols = np.ones(4)
    tags = np.array([2, 2, 2, 5], np.int32)
    bary = np.array([[0.0, 0, 0], [1, 0, 0], [2, 0, 0], [9, 9, 9]])
    m = compute_metrics(magn, vols, bary, tags, "gray_matter")
    assert m["peak_magnE"] == 3.0 and m["region_volume_m3"] == 3.0
    assert m["focality_m3"]["0.5"] == 2.0
    assert m["distribution"]["mean"] == pytest.approx(2.0)
    assert m["distribution"]["std"] == pytest.approx(np.sqrt(2 / 3))
    r = FieldResult(summary=m, magnE=magn, E=None, v=None, transform=np.eye(4), placement=None, coil_name="",
                    didt=1e6, vols=vols, tet_tags=tags, barycenters_mm=bary)
    assert r.peak_magnE() == 3.0
    assert r.peak_magnE("scalp") == 100.0
    assert r.focality(1.0) == 1.0
    assert r.focality(0.0) == 3.0
    bare = FieldResult(summary=m, magnE=None, E=None, v=None, transform=np.eye(4), placement=None,
                       coil_name="", didt=1e6, vols=None, tet_tags=None, barycenters_mm=None)
    with pytest.raises(ValueError, match="magnitude=True"):
        bare.peak_magnE("scalp")


@pytest.mark.gpu
def test_subject_end_to_end_matches_the_block_solver(head4, coil):
    """Subject.iter_simulate streams results in order, chunked by block_k, with summaries equal to
    the library's block solve of the same placements."""
    from cunibs_amd import fem, synthetic
    from cunibs_amd.mesh import HeadMesh
    sites = synthetic.cap_sites(5, seed=9)
    placements = [Placement(c, h, d) for c, h, d in sites]
    c = Coil(coil.positions_m, coil.moments, name="fig8")
    with Subject(HeadMesh(head4.nodes_mm, head4.tet_nodes, head4.tet_tags, head4.skin_tris)) as subj:
        res = list(subj.iter_simulate(c, placements, magnitude=True, block_k=2, recovery="raw"))
        assert [r.placement for r in res] == placements
        ref = fem.solve_placements_block(subj.context, c.positions_m, c.moments, sites[:2], 1e6)
        for a, b in zip(res[:2], ref):
            np.testing.assert_array_equal(a.magnE, b.magnE)
            assert a.summary["peak_magnE"] == pytest.approx(float(b.metrics[0]))
        for r in res:
            assert r.relative_residual <= 1e-6
            host = compute_metrics(r.magnE, r.vols, r.barycenters_mm, r.tet_tags, "gray_matter")
            assert r.summary["peak_magnE"] == host["peak_magnE"]
            assert r.summary["distribution"]["p99"] == pytest.approx(host["distribution"]["p99"], rel=1e-6)
            assert r.peak_magnE("csf") > 0
        one = subj.simulate(c, placements[0], magnitude=True, recovery="raw")
        assert np.linalg.norm(one.magnE - res[0].magnE) / np.linalg.norm(res[0].magnE) < 2e-5
        # The default is the reference's: harmonic recovery, with the nodal field on demand.
        h = subj.simulate(c, placements[0], magnitude=True, nodal=True)
        assert h.recovery == "harmonic" and h.E_slots is not None
        nf = h.nodal_field("gray_matter")
        assert nf.shape == (subj.context.n_nodes, 3) and np.isfinite(nf).any() and np.isnan(nf).any()
        with pytest.raises(ValueError):
            h.nodal_field("all")


@pytest.mark.gpu
def test_summary_only_streaming_uses_the_pipelined_sweep(head4, coil):
    """iter_simulate without field outputs streams through the pipelined sweep, whose blocks start
    from the Galerkin projection onto the previous blocks' solutions: each result's frame equals a
    block-by-block solve of the same placements, its summary matches it to the solve tolerance and
    its FP64 true residual meets that tolerance."""
    from cunibs_amd import fem, synthetic
    from cunibs_amd.mesh import HeadMesh
    sites = synthetic.cap_sites(5, seed=23)
    placements = [Placement(c, h, d) for c, h, d in sites]
    c = Coil(coil.positions_m, coil.moments, name="fig8")
    with Subject(HeadMesh(head4.nodes_mm, head4.tet_nodes, head4.tet_tags, head4.skin_tris)) as subj:
        res = list(subj.iter_simulate(c, placements, block_k=2, recovery="raw"))
        assert [r.placement for r in res] == placements
        ref = []
        for i in range(0, 5, 2):
            ref.extend(fem.solve_placements_block(subj.context, c.positions_m, c.moments, sites[i:i + 2], 1e6,
                                                  magnitude=False))
        for a, b in zip(res, ref):
            assert a.magnE is None
            want = fem.metrics_dict(b.metrics)
            assert a.summary.keys() == want.keys()
            for key, value in want.items():
                if isinstance(value, dict):
                    for sub_key, v in value.items():
                        assert a.summary[key][sub_key] == pytest.approx(v, rel=2e-5), (key, sub_key)
                elif isinstance(value, str):
                    assert a.summary[key] == value, key
                else:
                    np.testing.assert_allclose(a.summary[key], value, rtol=2e-5, err_msg=key)
            np.testing.assert_array_equal(a.transform, b.transform)
            assert a.relative_residual <= 1e-6
