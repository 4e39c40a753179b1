"""Bitwise determinism across processes (the reference's tests/test_determinism_cross_process.py:93-119):
two fresh interpreters, each building its own context and solving the same block, produce the
same hierarchy and the same block results bit for bit. The children are started by this test's
process, which makes no HIP call itself (this module is collected before any other GPU test)."""

from __future__ import annotations

import json
import subprocess
import sys
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]

CHILD = r"""
import hashlib, json, sys
sys.path.insert(0, sys.argv[1])
import numpy as np
from cunibs_amd import fem, synthetic
head = synthetic.shell_head(4.0, perturb_rms_mm=2.0, seed=1)
coil = synthetic.figure8_coil(seed=0)
ctx = fem.SolverContext(head.nodes_mm, head.tet_nodes, head.tet_tags, head.skin_tris)
h = hashlib.sha256()
for i in range(len(ctx.hierarchy_sizes) - 1):
    lv = ctx.level(i)
    for key in ("agg", "a_val", "dinv", "p_val"):
        h.update(np.ascontiguousarray(lv[key]).tobytes())
hier = h.hexdigest()
sites = synthetic.cap_sites(8, seed=9)
res = fem.solve_placements_block(ctx, coil.positions_m, coil.moments, sites, 1e6, potential=True)
h = hashlib.sha256()
for r in res:
    for a in (r.magnE, r.v, r.metrics, np.array([r.iterations]), np.array([r.relative_residual])):
        h.update(np.ascontiguousarray(a).tobytes())
block = h.hexdigest()
sweep = fem.solve_sweep(ctx, coil.positions_m, coil.moments, synthetic.cap_sites(12, seed=10), 1e6, block_k=8)
h = hashlib.sha256()
for key in ("metrics", "iterations", "relative_residual"):
    h.update(np.ascontiguousarray(sweep[key]).tobytes())
ctx.free()
print(json.dumps({"hierarchy": hier, "block": block, "sweep": h.hexdigest()}), flush=True)
"""


def _child() -> dict:
    out = subprocess.run([sys.executable, "-c", CHILD, str(ROOT)], capture_output=True, text=True, timeout=280)
    assert out.returncode == 0, out.stderr[-2000:]
    return json.loads(out.stdout.strip().splitlines()[-1])


def test_two_processes_give_the_same_bits():
    a, b = _child(), _child()
    assert a == b
