"""Shared fixtures. CPU tests pin the oracle and the host logic; `gpu` tests run the HIP library
through its C ABI against the oracle (the reference's tiers, tests/conftest.py:82-95 there)."""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")


# 8-node, 6-tet cube of side 100 mm (Kuhn split along the main diagonal), the reference's smallest
# fixture shape (tests/conftest.py:24-64 there).
CUBE_NODES = np.array([[x, y, z] for z in (0.0, 100.0) for y in (0.0, 100.0) for x in (0.0, 100.0)])
CUBE_TETS = np.array([[0, 1, 3, 7], [0, 1, 5, 7], [0, 2, 3, 7], [0, 2, 6, 7], [0, 4, 5, 7], [0, 4, 6, 7]],
                     dtype=np.int32)


def kuhn_grid(nx: int, ny: int, nz: int, spacing: float = 10.0, jitter: float = 0.0, seed: int = 0):
    """Structured Kuhn mesh of a box: nodes (n, 3) mm, tets (m, 4) positively oriented."""
    g = np.stack(np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij"), -1)
    nodes = g.reshape(-1, 3).astype(np.float64) * spacing
    if jitter:
        rng = np.random.default_rng(seed)
        nodes += rng.uniform(-jitter, jitter, nodes.shape) * spacing
    idx = np.arange(nodes.shape[0]).reshape(nx + 1, ny + 1, nz + 1)
    corner = lambda dx, dy, dz: idx[dx:dx + nx, dy:dy + ny, dz:dz + nz].ravel()
    c = {(dx, dy, dz): corner(dx, dy, dz) for dx in (0, 1) for dy in (0, 1) for dz in (0, 1)}
    paths = [((1, 0, 0), (1, 1, 0)), ((1, 0, 0), (1, 0, 1)), ((0, 1, 0), (1, 1, 0)),
             ((0, 1, 0), (0, 1, 1)), ((0, 0, 1), (1, 0, 1)), ((0, 0, 1), (0, 1, 1))]
    tets = np.concatenate([np.stack([c[(0, 0, 0)], c[a], c[b], c[(1, 1, 1)]], 1) for a, b in paths])
    p = nodes[tets]
    det = np.einsum("ij,ij->i", np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), p[:, 3] - p[:, 0])
    flip = det < 0
    tets[flip] = tets[flip][:, [0, 2, 1, 3]]
    return nodes, tets.astype(np.int32)


def boundary_tris(tets: np.ndarray, nodes: np.ndarray) -> np.ndarray:
    from cunibs_amd.synthetic import _boundary_faces
    return _boundary_faces(tets, nodes)


def true_residuals(ctx, res, rhs):
    """||b - A v|| / ||b|| per placement in FP64 on the host: b the solver's FP32 right-hand side
    (the `rhs` extra output) with the ground row removed, A the FP64 reduced operator, v the
    returned potential."""
    import scipy.sparse as sp
    rp, rc, rv = ctx.reduced()
    a = sp.csr_matrix((rv, rc, rp), shape=(ctx.n_nodes - 1,) * 2)
    keep = np.ones(ctx.n_nodes, bool)
    keep[ctx.ground_node] = False
    out = []
    for r, b in zip(res, rhs):
        bb = b[keep].astype(np.float64)
        out.append(float(np.linalg.norm(bb - a @ r.v[keep]) / np.linalg.norm(bb)))
    return out


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def head8():
    """5-shell head on an 8 mm grid: small enough for dense checks on the CPU."""
    from cunibs_amd import synthetic
    return synthetic.shell_head(8.0)


@pytest.fixture(scope="session")
def head4():
    from cunibs_amd import synthetic
    return synthetic.shell_head(4.0)


@pytest.fixture(scope="session")
def coil():
    from cunibs_amd import synthetic
    return synthetic.figure8_coil(seed=0)
