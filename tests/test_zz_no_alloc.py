"""cn_k_gram_noalloc (the Gram entry point that allocates nothing) against cn_k_gram on the same
inputs, bit for bit, on two calls in a row with the second smaller. The test's own device buffers
come from the HIP runtime the library is linked against (hipMalloc through ctypes), since the
entry point takes device pointers."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime the library loaded (same soname, one instance)."""

    def __init__(self):
        from cunibs_amd._lib import lib
        lib()  # the library first, so that its HIP runtime is the one resolved below
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.h.hipDeviceSynchronize.argtypes = []
        self.live = []

    def put(self, a: np.ndarray) -> int:
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        self.live.append(p.value)
        assert self.h.hipMemcpy(p.value, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p.value

    def empty(self, nbytes: int) -> int:
        return self.put(np.zeros(nbytes // 8 + 1, np.float64))

    def get(self, dev: int, shape, dtype) -> np.ndarray:
        out = np.empty(shape, dtype)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(out.ctypes.data, dev, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.live:
            self.h.hipFree(p)
        self.live.clear()


def _noalloc(hip, A, B1, B2):
    from cunibs_amd._lib import check, lib
    n, K = A.shape
    E = (1 if B2 is None else 2) * K * K
    out = hip.empty(8 * E)
    check(lib().cn_k_gram_noalloc(n, K, hip.put(A), hip.put(B1), None if B2 is None else hip.put(B2), out, None),
          "cn_k_gram_noalloc")
    return hip.get(out, (E,), np.float64)


def test_gram_noalloc_matches_the_allocating_entry(head8):
    from cunibs_amd import fem, synthetic
    hip = _Hip()
    small = synthetic.shell_head(16.0)
    ctxs = [fem.SolverContext(h.nodes_mm, h.tet_nodes, h.tet_tags, h.skin_tris) for h in (head8, small)]
    try:
        rng = np.random.default_rng(5)
        for ctx, K, two in ((ctxs[0], 16, True), (ctxs[1], 8, False)):  # the second call is smaller
            n = ctx.n_nodes - 1
            A, B1, B2 = (rng.standard_normal((n, K)).astype(np.float32) for _ in range(3))
            want = ctx.gram(A, B1, B2 if two else None)
            got = _noalloc(hip, A, B1, B2 if two else None)
            want = np.concatenate([w.ravel() for w in want]) if two else want.ravel()
            np.testing.assert_array_equal(got, want)
        assert ctxs[1].n_nodes < ctxs[0].n_nodes
    finally:
        for c in ctxs:
            c.free()
        hip.free()


def test_gram_noalloc_rejects_a_width_it_was_not_built_for():
    from cunibs_amd._lib import lib
    from cunibs_amd._lib import check
    with pytest.raises(ValueError):
        check(lib().cn_k_gram_noalloc(100, 12, 1, 1, None, 1, None), "cn_k_gram_noalloc")
