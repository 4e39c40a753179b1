"""The warm-start network: FP64 oracle gradients pinned by finite differences (CPU); the GPU's
training gradients against them, its BF16 MFMA inference against a BF16 emulation, training
that converges, and the iterations the trained network saves over a cold start."""

from __future__ import annotations

import numpy as np
import pytest


def test_oracle_gradients_match_finite_differences(oracle_mod):
    rng = np.random.default_rng(0)
    H, O, B = 16, 16, 5
    n = 16 * H + H + H * H + H + H * O + O
    p = rng.normal(0, 0.3, n)
    p[:16 * H].reshape(16, H)[6:] = 0.0  # padded input rows
    X = rng.normal(size=(B, 6))
    T = rng.normal(size=(B, O))
    loss, g = oracle_mod.mlp_loss_and_gradients(p, X, T, H, O)
    for idx in rng.choice(n, 40, replace=False):
        if idx < 16 * H and (idx // H) >= 6:
            continue
        e = np.zeros(n)
        e[idx] = 1e-6
        lp, _ = oracle_mod.mlp_loss_and_gradients(p + e, X, T, H, O)
        lm, _ = oracle_mod.mlp_loss_and_gradients(p - e, X, T, H, O)
        assert g[idx] == pytest.approx((lp - lm) / 2e-6, rel=1e-5, abs=1e-9)


def bf16(x):
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32).view(np.float32)


@pytest.mark.gpu
def test_gpu_gradients_match_the_oracle(oracle_mod):
    from cunibs_amd.learned import WarmStartNetwork
    net = WarmStartNetwork(n_out=32, hidden=64, seed=3, n_in=14)
    rng = np.random.default_rng(1)
    X = rng.normal(size=(48, 14)).astype(np.float32)
    T = rng.normal(size=(48, 32)).astype(np.float32)
    p = net.params()
    loss, g = net.gradients(X, T)
    want_loss, want = oracle_mod.mlp_loss_and_gradients(p.astype(np.float64), X.astype(np.float64),
                                                        T.astype(np.float64), 64, 32)
    assert loss == pytest.approx(want_loss, rel=1e-5)
    assert np.linalg.norm(g - want) <= 1e-4 * np.linalg.norm(want)


@pytest.mark.gpu
def test_gpu_bf16_inference_matches_an_emulation(oracle_mod):
    from cunibs_amd.learned import WarmStartNetwork
    net = WarmStartNetwork(n_out=32, hidden=64, seed=4, n_in=14)
    X = np.random.default_rng(2).normal(size=(20, 14)).astype(np.float32)
    y = net.predict_raw(X)
    (w1, b1), (w2, b2), (w3, b3) = oracle_mod.mlp_unpack(net.params(), 64, 32)
    x = np.zeros((20, 16), np.float32)
    x[:, :14] = X
    h = bf16(oracle_mod._gelu(bf16(x).astype(np.float64) @ bf16(w1) + b1))
    h = bf16(oracle_mod._gelu(h.astype(np.float64) @ bf16(w2) + b2))
    want = h.astype(np.float64) @ bf16(w3) + b3
    assert np.linalg.norm(y - want) <= 1e-3 * np.linalg.norm(want)


@pytest.mark.gpu
def test_gpu_training_fits_a_smooth_map():
    from cunibs_amd.learned import WarmStartNetwork
    rng = np.random.default_rng(5)
    X = rng.uniform(-1, 1, size=(512, 4)).astype(np.float32)
    T = np.stack([np.sin(2 * X[:, 0]) + X[:, 1] * X[:, 2] for _ in range(16)], 1).astype(np.float32)
    T *= np.linspace(0.5, 2.0, 16, dtype=np.float32)
    net = WarmStartNetwork(n_out=16, hidden=128, seed=0, n_in=4)
    before = float(((net.predict(X) - T) ** 2).mean())
    net.fit(X, T, epochs=300, batch=128, lr=2e-3)
    after = float(((net.predict(X) - T) ** 2).mean())
    assert after < 0.02 * before


@pytest.mark.gpu
def test_learned_warm_start_saves_iterations(head4, coil):
    """Trained on solver-generated solutions, the network's coefficients start held-out placements
    closer than a cold start: fewer iterations, same answer."""
    from cunibs_amd import fem, learned, synthetic
    ctx = fem.SolverContext(head4.nodes_mm, head4.tet_nodes, head4.tet_tags, head4.skin_tris)
    try:
        train = synthetic.cap_sites(192, seed=21)
        net = learned.train_warm_start(ctx, coil.positions_m, coil.moments, train, modes=64, hidden=256,
                                       epochs=400)
        test = synthetic.cap_sites(8, seed=22)
        cold = fem.solve_placements_block(ctx, coil.positions_m, coil.moments, test, 1e6)
        warm = learned.solve_with_network(ctx, net, coil.positions_m, coil.moments, test)
        for a, b in zip(cold, warm):
            assert b.relative_residual <= 1e-6
            assert np.linalg.norm(a.magnE - b.magnE) <= 2e-5 * np.linalg.norm(a.magnE)
        assert np.mean([r.iterations for r in warm]) < np.mean([r.iterations for r in cold])
    finally:
        ctx.free()
