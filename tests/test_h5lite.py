"""HDF5 without h5py: the coil files and FieldResult format version 1 (the reference's
tests/test_coil.py and tests/test_serialization.py tier, CPU). The reader is pinned on the
reference's h5py-written D70 coil (tests/golden/); the writer by round trips through the reader and
by reproducing the reference's own expectations for ``encode_ccd``."""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"
REF_COILS = Path("/root/reference/src/cunibs/coils")

_CCD = """\
# brand=Test;coilname=Synth;dIdtmax=100.0
2
# x y z mx my mz
0.0 0.0 0.0 0.0 0.0 1.0
0.01 0.0 0.0 0.0 0.0 -1.0
"""


def test_encode_ccd_roundtrips_to_coil(tmp_path):
    """tests/test_coil.py:28-41 and :94-106 there."""
    from cunibs_amd import h5lite
    from cunibs_amd.simulation import Coil, encode_ccd
    ccd = tmp_path / "synth.ccd"
    ccd.write_text(_CCD)
    h5 = tmp_path / "synth.h5"
    encode_ccd(ccd, h5)
    c = Coil.load(h5)
    assert c.name == "Synth"
    assert c.didt_max == 100.0e6
    assert c.metadata["dIdtmax"] == "100.0" and c.metadata["brand"] == "Test"
    np.testing.assert_allclose(c.positions_m, [[0.0, 0.0, 0.0], [0.01, 0.0, 0.0]])
    np.testing.assert_allclose(c.moments, [[0.0, 0.0, 1.0], [0.0, 0.0, -1.0]])
    f = h5lite.open_file(h5)
    assert f["dipoles"].shape == (2, 6) and int(f["dipoles"].attrs["num_elements"]) == 2


def test_coil_without_attributes_falls_back_to_the_stem(tmp_path):
    """tests/test_coil.py:109-115 there."""
    from cunibs_amd import h5lite
    from cunibs_amd.simulation import Coil
    h5 = tmp_path / "Nameless.h5"
    h5lite.write_file(h5, {"dipoles": np.zeros((3, 6))})
    with pytest.raises(ValueError):  # all-zero moments are still a valid array, but no dipoles move
        Coil(np.zeros((0, 3)), np.zeros((0, 3)))
    c = Coil.load(h5)
    assert c.didt_max is None and c.name == "Nameless"


@pytest.mark.parametrize("compression", [None, "gzip"])
def test_rewriting_the_d70_coil_reproduces_it(tmp_path, compression):
    """Read the h5py-written coil, write it with the writer, read that: same data, same attributes;
    the deflated chunk layout is the one the h5py file uses (class 2, rank + 1 dims, element size)."""
    from cunibs_amd import h5lite
    src = h5lite.open_file(GOLDEN / "MagStim_D70.h5")["dipoles"]
    h5 = tmp_path / "d70.h5"
    h5lite.write_file(h5, {"dipoles": h5lite.DatasetSpec(src.read(), dict(src.attrs))}, compression=compression)
    if compression:
        r = h5lite._Reader(h5.read_bytes(), "w")
        lay = next(b for t, b in r.messages(r.links(r.messages(r.root))["dipoles"]) if t == 0x08)
        assert lay[:3] == bytes([3, 2, 3]) and lay[11:23] == (964).to_bytes(4, "little") + \
            (6).to_bytes(4, "little") + (8).to_bytes(4, "little")
        assert h5.stat().st_size < 964 * 6 * 8
    back = h5lite.open_file(h5)["dipoles"]
    np.testing.assert_array_equal(back.read(), src.read())
    assert {k: (v.item() if isinstance(v, np.generic) else v) for k, v in back.attrs.items()} == \
        {k: (v.item() if isinstance(v, np.generic) else v) for k, v in src.attrs.items()}


def test_writer_encodings_match_the_h5py_written_file():
    """The writer's variable-length string type, scalar dataspace, float64 and int64 types are the
    bytes the HDF5 library put in the D70 file."""
    from cunibs_amd import h5lite
    r = h5lite._Reader((GOLDEN / "MagStim_D70.h5").read_bytes(), "d70")
    msgs = r.messages(r.links(r.messages(r.root))["dipoles"])
    dtype = next(b for t, b in msgs if t == 0x03)
    assert dtype[:len(h5lite._numeric_type(np.float64))] == h5lite._numeric_type(np.float64)
    attrs = [b for t, b in msgs if t == 0x0C]
    text = next(b for b in attrs if b[8:16].startswith(b"coilname"))
    num = next(b for b in attrs if b[8:21].startswith(b"num_elements"))
    assert text[24:24 + 20] == h5lite._VLEN_UTF8
    assert text[48:56] == h5lite._dataspace(())
    assert num[24:36] == h5lite._numeric_type(np.int64)


@pytest.mark.skipif(not REF_COILS.exists(), reason="reference coil folder not present")
def test_all_bundled_coils_load():
    """tests/test_coil.py:68-87 there, over the reference's 25 bundled files."""
    from cunibs_amd import coils
    from cunibs_amd.simulation import Coil
    assert len(coils.FILES) == 25
    for fname in coils.FILES.values():
        c = Coil.load(REF_COILS / fname)
        n = c.positions_m.shape[0]
        assert n > 0 and c.positions_m.shape == (n, 3) and c.moments.shape == (n, 3)
        assert np.isfinite(c.positions_m).all() and np.abs(c.positions_m).max() < 1.0
        assert np.abs(c.moments).max() > 0.0 and c.name
        assert c.didt_max is None or c.didt_max > 1e7
        assert "coilname" in c.metadata


def _field_result(with_arrays: bool):
    from cunibs_amd.simulation import FieldResult, Placement
    rng = np.random.default_rng(0)
    m, n = 50, 20
    summary = {"region": "gray_matter", "peak_magnE": 0.5, "peak_location_mm": np.array([1.0, 2.0, 3.0]),
               "center_of_gravity_mm": np.array([0.5, 0.25, 2.0]), "region_volume_m3": 1e-6,
               "focality_m3": {"0.5": 2e-7, "0.75": 1e-7},
               "distribution": {"mean": 0.1, "std": 0.05, "p50": 0.09, "p95": 0.3, "p99": 0.4, "p99.9": 0.45}}
    arr = (lambda a: a) if with_arrays else (lambda a: None)
    return FieldResult(summary=summary, magnE=arr(rng.random(m).astype(np.float32)),
                       E=arr(rng.random((m, 3)).astype(np.float32)), v=arr(rng.random(n)),
                       transform=np.eye(4), placement=Placement([0.0, 0.0, 90.0], [0.0, 50.0, 90.0], 4.0),
                       coil_name="D70", didt=1e6, vols=arr(rng.random(m).astype(np.float32)),
                       tet_tags=arr(rng.integers(1, 6, m).astype(np.int32)),
                       barycenters_mm=arr(rng.random((m, 3))), recovery="harmonic",
                       n_nodes=n if with_arrays else None)


@pytest.mark.parametrize("with_arrays", [True, False])
def test_field_result_save_load_round_trip(tmp_path, with_arrays):
    """tests/test_serialization.py there: every retained array and the summary come back; arrays
    the run did not keep load as None."""
    from cunibs_amd.simulation import FieldResult
    r = _field_result(with_arrays)
    path = tmp_path / "r.h5"
    r.save(path)
    back = FieldResult.load(path)
    for name in ("magnE", "E", "v", "vols", "tet_tags", "barycenters_mm"):
        a, b = getattr(r, name), getattr(back, name)
        if a is None:
            assert b is None
        else:
            np.testing.assert_array_equal(b, a)
            assert b.dtype == a.dtype
    np.testing.assert_array_equal(back.transform, r.transform)
    assert back.summary["peak_magnE"] == r.summary["peak_magnE"]
    np.testing.assert_array_equal(back.summary["peak_location_mm"], r.summary["peak_location_mm"])
    assert back.summary["focality_m3"] == r.summary["focality_m3"]
    assert back.summary["distribution"] == r.summary["distribution"]
    assert back.summary["region"] == "gray_matter"
    assert (back.coil_name, back.didt, back.recovery, back.n_nodes) == ("D70", 1e6, "harmonic", r.n_nodes)
    np.testing.assert_array_equal(back.placement.center_mm, r.placement.center_mm)
    np.testing.assert_array_equal(back.placement.handle_mm, r.placement.handle_mm)
    assert back.placement.distance_mm == 4.0


def test_field_result_rejects_other_format_versions(tmp_path):
    from cunibs_amd import h5lite
    from cunibs_amd.simulation import FieldResult
    h5lite.write_file(tmp_path / "v2.h5", {"__attrs__": {"format_version": np.int64(2)}, "transform": np.eye(4)})
    with pytest.raises(ValueError, match="format version 2"):
        FieldResult.load(tmp_path / "v2.h5")


def test_truncated_and_foreign_files_are_rejected(tmp_path):
    from cunibs_amd import h5lite
    (tmp_path / "x.h5").write_bytes(b"not an hdf5 file at all" * 10)
    with pytest.raises(ValueError, match="not an HDF5 file"):
        h5lite.open_file(tmp_path / "x.h5")
