"""The C-ABI library: it loads, exports every symbol the header declares with the signatures the
ctypes layer binds, and its host-only entry points and argument checks work without a GPU."""

from __future__ import annotations

import ctypes
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = ROOT / "include" / "cunibs_amd.h"


def header_functions() -> list[str]:
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(cn_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def clib():
    from cunibs_amd import _lib
    return _lib.lib()


def test_library_exports_every_declared_function(clib):
    names = header_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(clib, n)]
    assert not missing, missing


def test_ctypes_layer_binds_exactly_the_header(clib):
    from cunibs_amd import _lib
    assert sorted(_lib.EXPORTED) == header_functions()


def test_block_widths_and_version(clib):
    from cunibs_amd import fem
    buf = (ctypes.c_int32 * 16)()
    n = clib.cn_block_sizes(buf, 16)
    widths = list(buf[:n])
    assert widths == [1, 2, 4, 8, 16, 32]
    assert clib.cn_max_block() == 32
    assert fem.BLOCK_SIZES == tuple(widths)
    assert clib.cn_version() >= 1


def test_default_options_match_the_reference_defaults(clib):
    """fem/solve.py:137-168 (AggregationParams, Chebyshev degree 2 on [1/12, 1])."""
    from cunibs_amd import _lib
    o = _lib.Options()
    clib.cn_default_options(ctypes.byref(o))
    assert (o.min_coarse_rows, o.max_levels, o.rounds, o.degree) == (128, 50, 2, 2)
    assert o.lower_ratio == pytest.approx(12.0)
    assert o.region_tag == 2
    s = _lib.SolveOptions()
    clib.cn_default_solve_options(ctypes.byref(s))
    assert s.tolerance == pytest.approx(1e-6) and s.max_iters >= 1000 and s.use_graph == 1


def test_argument_errors_come_back_as_status_codes(clib):
    """No exception crosses the ABI: a bad argument is CN_EINVAL with a message, checked before
    any device work (so this runs without a GPU)."""
    from cunibs_amd import _lib
    out = ctypes.c_void_p()
    st = clib.cn_context_create(None, 0, None, None, 0, None, 0, None, ctypes.byref(out))
    assert st == _lib.CN_EINVAL
    assert "null" in _lib.last_error()
    o = _lib.Options()
    clib.cn_default_options(ctypes.byref(o))
    o.degree = 9
    nodes = np.zeros((4, 3))
    tets = np.zeros((1, 4), np.int32)
    tags = np.ones(1, np.int32)
    st = clib.cn_context_create(nodes.ctypes.data, 4, tets.ctypes.data, tags.ctypes.data, 1, None, 0,
                                ctypes.byref(o), ctypes.byref(out))
    assert st == _lib.CN_EINVAL
    assert "degree" in _lib.last_error()
    with pytest.raises(ValueError, match="degree"):
        _lib.check(st, "cn_context_create")


def test_status_mapping_to_python_exceptions():
    from cunibs_amd import _lib
    _lib.check(_lib.CN_OK, "ok")
    with pytest.raises(_lib.SolverConvergenceError):
        _lib.check(_lib.CN_ENOCONV, "solve")
    with pytest.raises(RuntimeError):
        _lib.check(_lib.CN_EHIP, "hip")


def test_product_path_never_imports_the_oracle():
    """The package must not reach the CPU oracle: no module under cunibs_amd/ imports it."""
    for path in (ROOT / "cunibs_amd").rglob("*.py"):
        text = path.read_text()
        assert not re.search(r"^\s*(import|from)\s+oracle\b", text, flags=re.M), path
