"""Placements/s of the TMS E-field path on one MI355X.

Workload: the 1.5 mm perturbed 5-shell head (~1.0M nodes, ~5.8M tetrahedra) and the 16 x 16 cap
grid of 256 placements, solved to relative residual 1e-6 in blocks of ``--block-k`` placements.
A step is one block: coil frames, dA/dt N-body, RHS, block PCG, E reconstruction and the
gray-matter metrics. Inputs are resident on the GPU (the context is built before the timed region).

Printed as one JSON line. Besides the contract fields it carries the plain single-placement GPU
PCG rate (k = 1) measured in the same run, the GM |E| error of sample placements against the FP64
oracle (CG to 1e-12), the SpMM roofline (``roofline``) and the CPU oracle timed on this host
(``cpu_baseline``).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "placements/s @ rel.res 1e-6, GM |E| err<=1e-3 vs FP64; batched+warm vs plain"
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the differenced SpMM, from two rocprofv3 counter passes over
# `bench.py --kernels-only --block-k 16 --fine-sub 2` (FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000, the
# factors that make the counters read 512 MiB each way for the streaming-copy launch), keyed by
# (grid spacing mm, block width). 198.4 MB fetched + 64.2 MB written, 1.07x the algorithmic bytes.
MEASURED_SPMM_TRAFFIC = {(1.5, 16): 262.6e6}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def spmm_diff_bytes(n: int, nnz: int, k: int) -> float:
    """Algorithmic bytes of the block iteration's differenced product S = A W (n rows, k columns):
    the off-diagonal CSR once (FP32 value + int32 column per entry, n + 1 row pointers), the FP32
    row sums, W read once and S written once (FP32)."""
    nnz_off = nnz - n
    return 8.0 * nnz_off + 4.0 * (n + 1) + 4.0 * n + 8.0 * n * k


def vcycle_bytes(info, K: int, degree: int) -> float:
    """Algorithmic bytes of one zero-guess V-cycle at width K (FP32 vectors, int32 indices): per
    smoothed level of n rows, nnz entries and p prolongator entries (n_c coarse rows) the first
    Jacobi step (b, dinv in; x out), degree - 1 further pre-smoothing and degree post-smoothing
    Chebyshev steps (matrix once, row pointers, dinv, x gathered once, x_prev, b, out), the
    residual (matrix, b, x, out), the restriction (P^T once, fine residual in, coarse rhs out) and
    the prolongation (P once, coarse x in, fine x read and written); then the coarsest FP64 inverse
    (read once) with its rhs and solution."""
    total = 0.0
    for i in range(info.n_levels):
        n, a, p = info.level_rows[i], info.level_nnz[i], info.level_p_nnz[i]
        nc = info.level_rows[i + 1]
        mat = 8.0 * a + 4.0 * (n + 1)
        cheby = mat + 4.0 * n + 16.0 * n * K
        total += 8.0 * n * K + 4.0 * n              # Jacobi step from zero
        total += (2 * degree - 1) * cheby            # pre (degree - 1) and post (degree) steps
        total += mat + 12.0 * n * K                  # residual
        total += 8.0 * p + 4.0 * (nc + 1) + 4.0 * K * (n + nc)        # restriction
        total += 8.0 * p + 4.0 * (n + 1) + 4.0 * K * nc + 8.0 * n * K  # prolongation
    nco = info.level_rows[info.n_levels]
    total += 8.0 * nco * nco + 8.0 * nco * K
    return total


COPY_BYTES = 512 << 20  # calibration copy: 512 MiB read + 512 MiB written per launch


def kernels_only(ctx, coil, sites, k: int) -> int:
    """The launches the HBM traffic counters are read on: 20 streaming copies of COPY_BYTES (the
    counter calibration), 20 differenced SpMMs at width k (each after one untimed launch), then one
    cold block of k placements with each recovery (raw, spr_global, harmonic), so that every kernel
    of the path appears once with the same sizes as in the sweep; then the sweep's warm-start basis
    (256 solves) and two warm-started raw blocks, the sweep's steady state."""
    from cunibs_amd import fem, synthetic

    K = k if k in fem.BLOCK_SIZES else min(b for b in fem.BLOCK_SIZES if b >= k)
    n_red, nnz = int(ctx.info.n_reduced), int(ctx.info.nnz_reduced)
    copy_ms = ctx.bench_copy(COPY_BYTES, 20)
    spmm_ms = ctx.bench_spmm(K, 20, mode=1)
    iters = {}
    for rec in ("raw", "spr_global", "harmonic"):
        res = fem.solve_placements_block(ctx, coil.positions_m, coil.moments, sites[:k], 1e6, magnitude=False,
                                         metrics=True, recovery=rec)
        iters[rec] = [r.iterations for r in res]
    train = synthetic.cap_sites(256, seed=7)
    ctx.warm_reset(len(train))
    for i in range(0, len(train), 16):
        fem.solve_placements_block(ctx, coil.positions_m, coil.moments, train[i:i + 16], 1e6, metrics=False,
                                   magnitude=False, collect=True)
    ctx.warm_finalize()
    for i in range(2):
        res = fem.solve_placements_block(ctx, coil.positions_m, coil.moments, sites[k * (i + 3):k * (i + 4)], 1e6,
                                         magnitude=False, metrics=True, warm_start=True)
        iters[f"warm{i}"] = [r.iterations for r in res]
    print(json.dumps({
        "copy_bytes_read": COPY_BYTES, "copy_bytes_written": COPY_BYTES, "copy_us": round(copy_ms * 1e3, 2),
        "copy_GBps": round(2 * COPY_BYTES / (copy_ms * 1e-3) / 1e9, 1),
        "spmm_K": K, "n_reduced": n_red, "nnz_reduced": nnz,
        "spmm_algorithmic_bytes": spmm_diff_bytes(n_red, nnz, K), "spmm_us": round(spmm_ms * 1e3, 2),
        "block_iterations": iters,
    }), flush=True)
    ctx.free()
    return 0


def sweep_only(ctx, coil, sites, args) -> int:
    """The pipelined cold sweep alone (warmup, then steps x block_k placements), with the library's
    in-sweep device-clock timing: a kernel trace of this command averages the differenced SpMM
    over exactly the launches the bench's in-sweep roofline averages."""
    from cunibs_amd import fem

    k = args.block_k
    ctx.set_tuning(degree=2, lower_ratio=8.0, fine_sub=max(args.fine_sub, 0), coil_blocks=args.coil_blocks)
    pos, mom = coil.positions_m, coil.moments
    run = lambda n: fem.solve_sweep(ctx, pos, mom, [sites[j % len(sites)] for j in range(n * k)], 1e6, k)
    ctx.timing_enable(True)
    run(args.warmup)
    ctx.timing_enable(True)  # reset the in-solve sums
    t0 = time.perf_counter()
    res = run(args.steps)
    elapsed = time.perf_counter() - t0
    timing = ctx.timing_read()
    n_red, nnz = int(ctx.info.n_reduced), int(ctx.info.nnz_reduced)
    K = k if k in fem.BLOCK_SIZES else min(b for b in fem.BLOCK_SIZES if b >= k)
    print(json.dumps({
        "sweep_placements_per_s": round(args.steps * k / elapsed, 2), "block_k": k, "tuning": ctx.tuning,
        "iterations_mean": round(float(np.mean(res["iterations"])), 2),
        "spmm_us": round(timing["spmm_ms"] * 1e3, 2), "spmm_samples": timing["samples"],
        "spmm_algorithmic_bytes": spmm_diff_bytes(n_red, nnz, K),
        "vcycle_us": round(timing["vcycle_ms"] * 1e3, 2), "iteration_us": round(timing["iteration_ms"] * 1e3, 2),
    }), flush=True)
    ctx.free()
    return 0


def learned_leg(ctx, pos, mom, n_train: int, n_test: int, log_sd: float = 0.2) -> dict:
    """The learned warm start on poses and conductivities (distance U[2, 8] mm, the five tissues of
    the head log-normal around the table): n_train solver-generated solutions, each with its own
    conductivities, form the snapshot basis and the training set; the network maps pose and log
    conductivities to the solution's coefficients in the 256 leading A-orthonormal modes; on
    n_test held-out placements it warm-starts PCG, which iterates to the same tolerance."""
    from cunibs_amd import fem, learned, synthetic, uq

    tags = [1, 2, 3, 4, 5]

    def draws(n, seed):
        sites = synthetic.cap_sites(n, seed=seed, distance_range_mm=(2.0, 8.0))
        sig = uq.draw_lognormal(tags, n, log_sd, seed=seed + 1)
        rows = fem.conductivity_rows([dict(zip(tags, r)) for r in sig])
        return sites, rows

    t0 = time.perf_counter()
    tr_sites, tr_rows = draws(n_train, 31)
    ctx.warm_reset(n_train)
    for i in range(0, n_train, 16):
        fem.solve_placements_block(ctx, pos, mom, tr_sites[i:i + 16], 1e6, metrics=False, magnitude=False,
                                   collect=True, sigma=tr_rows[i:i + 16])
    t_solve = time.perf_counter() - t0
    rank = ctx.warm_finalize(max_modes=256)
    t_basis = time.perf_counter() - t0 - t_solve
    feats = lambda sites, rows: np.stack([learned.pose_features(s, r[1:11]) for s, r in zip(sites, rows)])
    net = learned.WarmStartNetwork(n_out=(rank + 15) // 16 * 16, hidden=512, seed=0)
    T = np.zeros((n_train, net.n_out), np.float32)
    T[:, :rank] = ctx.warm_targets()[:, :rank]
    t1 = time.perf_counter()
    loss = net.fit(feats(tr_sites, tr_rows), T, epochs=300)
    net.modes = rank
    t_train = time.perf_counter() - t1

    te_sites, te_rows = draws(n_test, 37)
    X = feats(te_sites, te_rows)

    def run(mode):
        its, block_its, mags = [], [], []
        t = time.perf_counter()
        for i in range(0, n_test, 16):
            s, r = te_sites[i:i + 16], te_rows[i:i + 16]
            if mode == "learned":
                c = net.predict(X[i:i + 16])[:, :rank].astype(np.float64)
                ctx.warm_set_coefficients(c)
            res = fem.solve_placements_block(ctx, pos, mom, s, 1e6, metrics=True, magnitude=True, sigma=r,
                                             warm_start={"cold": 0, "galerkin": 1, "learned": 2}[mode])
            its.extend(x.iterations for x in res)
            block_its.append(max(x.iterations for x in res))
            mags.extend(x.magnE for x in res[:2])
        return n_test / (time.perf_counter() - t), its, block_its, mags

    out = {}
    ref = None
    for mode in ("cold", "learned", "galerkin"):
        rate, its, bits, mags = run(mode)
        out[mode] = {"placements_per_s": round(rate, 2), "iterations_mean": round(float(np.mean(its)), 2),
                     "block_iterations_mean": round(float(np.mean(bits)), 2)}
        if ref is None:
            ref = mags
        else:
            out[mode]["magnE_rel_diff_vs_cold"] = max(
                float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(mags, ref))
    out.update({"train_solves": n_train, "test_placements": n_test, "modes": rank, "train_loss": round(loss, 5),
                "solve_s": round(t_solve, 2), "basis_s": round(t_basis, 2), "network_train_s": round(t_train, 2)})
    net.free()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=32)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--block-k", type=int, default=8)
    ap.add_argument("--spacing", type=float, default=1.5)
    ap.add_argument("--plain", type=int, default=16, help="placements timed at k=1")
    ap.add_argument("--accuracy", type=int, default=2, help="placements checked against the FP64 oracle")
    ap.add_argument("--cpu-sample", type=int, default=2, help="placements the CPU oracle solves")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--warm-basis", type=int, default=256,
                    help="training placements whose solutions form the warm-start basis (0: cold only)")
    ap.add_argument("--autotune", type=int, default=1,
                    help="1: pick block width (4/8/16), SpMM tile and Chebyshev degree (2-6) on 32 training "
                         "placements before the sweeps; 0: use --block-k, the automatic tile and degree 2")
    ap.add_argument("--learned", type=int, default=512,
                    help="training solves of the learned warm start on random poses and conductivities, "
                         "evaluated on 256 held-out placements (0: skip)")
    ap.add_argument("--graphs", type=int, default=1,
                    help="0: launch every iteration's kernels directly instead of replaying its captured graph")
    ap.add_argument("--fine-sub", type=int, default=-1,
                    help="--kernels-only: fine-level tile (sub-lanes) to measure at (-1: automatic)")
    ap.add_argument("--isolated", type=int, default=0,
                    help="1: also time the SpMM and V-cycle in isolated repeated launches (warm caches)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the roofline launches (copy calibration + differenced SpMM), for counter runs")
    ap.add_argument("--sweep-only", action="store_true",
                    help="only the pipelined cold sweep at --block-k, --fine-sub and --coil-blocks (no tuning, no "
                         "blocks solved alone), so that a kernel trace's averages are the in-sweep ones")
    ap.add_argument("--coil-blocks", type=int, default=256,
                    help="--sweep-only: grid cap of the overlapped dA/dt")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the placement sweep shards "
              "over GPUs by placement (each GPU a contiguous share of the placements)", file=sys.stderr)
        return 2

    from cunibs_amd import fem, synthetic

    t0 = time.time()
    head = synthetic.shell_head(args.spacing, perturb_rms_mm=2.0, seed=1)
    coil = synthetic.figure8_coil(seed=0)
    sites = synthetic.grid_sites()
    log(f"mesh {head.nodes_mm.shape[0]} nodes {head.tet_nodes.shape[0]} tets in {time.time() - t0:.1f}s")
    t0 = time.time()
    ctx = fem.SolverContext(head.nodes_mm, head.tet_nodes, head.tet_tags, head.skin_tris)
    setup_s = time.time() - t0
    log(f"context {setup_s:.2f}s hierarchy {ctx.hierarchy_sizes}")
    k = args.block_k
    pos, mom = coil.positions_m, coil.moments
    if args.kernels_only:
        if args.fine_sub >= 0:
            ctx.set_tuning(fine_sub=args.fine_sub)
        return kernels_only(ctx, coil, sites, k)
    if args.sweep_only:
        return sweep_only(ctx, coil, sites, args)

    def block(i0: int, kk: int, warm: bool = False):
        chunk = [sites[(i0 + j) % len(sites)] for j in range(kk)]
        return fem.solve_placements_block(ctx, pos, mom, chunk, 1e6, magnitude=False, metrics=True,
                                          warm_start=warm, use_graph=bool(args.graphs))

    def sweep(warm: bool):
        """The timed sweep: steps x k placements through the pipelined sweep entry point (the next
        block's dA/dt and RHS overlap the current block's solve), after warmup blocks."""
        if args.warmup:
            fem.solve_sweep(ctx, pos, mom, [sites[j % len(sites)] for j in range(args.warmup * k)], 1e6, k,
                            warm_start=warm, use_graph=bool(args.graphs))
        chunk = [sites[j % len(sites)] for j in range(args.steps * k)]
        t_start = time.perf_counter()
        res = fem.solve_sweep(ctx, pos, mom, chunk, 1e6, k, warm_start=warm, use_graph=bool(args.graphs))
        elapsed = time.perf_counter() - t_start
        its = [int(i) for i in res["iterations"]]
        block_its = [max(its[b:b + k]) for b in range(0, len(its), k)]
        if not np.all(res["relative_residual"] <= 1e-6):
            raise RuntimeError("a placement of the sweep missed the tolerance")
        return elapsed, its, block_its

    def phase_run(warm: bool, n_blocks: int = 4) -> dict:
        """Device time per block by phase, measured on blocks solved one at a time (HIP events
        between the phases of the block solve; the pipelined sweep overlaps them)."""
        ctx.timing_enable(True)
        for b in range(n_blocks):
            block(b * k, k, warm)
        ph = ctx.timing_phases()
        nonlocal unshared_timing
        unshared_timing = ctx.timing_read()
        return {p: round(v, 3) for p, v in ph["ms_per_block"].items()}

    unshared_timing = None

    ctx.timing_enable(True)
    tune = None
    if args.autotune:
        from cunibs_amd import autotune

        tune = autotune.autotune(ctx, pos, mom, synthetic.cap_sites(32, seed=13), 1e6, block_ks=(4, 8, 16),
                                 ratios=(8.0, 12.0, 20.0, 30.0))
        k = tune.block_k
        log(f"autotune {tune.seconds:.2f}s: block_k {k}, degree {tune.degree}, ratio {tune.lower_ratio} "
            f"(rates {tune.ratio_rates}), fine_sub {tune.fine_sub}, "
            f"coarse_sub {tune.coarse_sub} (coarse tile us {tune.coarse_tile_us}); "
            f"tile us {tune.tile_us}, degree rates {tune.degree_rates}, width rates {tune.width_rates}")
    overlap = None
    if args.autotune:
        from cunibs_amd import autotune

        overlap = autotune.tune_overlap(ctx, pos, mom, synthetic.cap_sites(8 * k, seed=19), 1e6, k)
        log(f"overlap: coil_blocks {overlap['coil_blocks']}, rates "
            f"{ {c: round(r, 1) for c, r in overlap['rates'].items()} }")
    ctx.timing_enable(True)  # reset the in-solve sums
    cold_elapsed, cold_iters, cold_block_iters = sweep(False)
    cold_timing = ctx.timing_read()
    cold_value = args.steps * k / cold_elapsed
    log(f"batched cold k={k}: {cold_value:.1f} placements/s, {1e3 * cold_elapsed / args.steps:.1f} ms/step, "
        f"iterations mean {np.mean(cold_iters):.1f}")
    train_s = 0.0
    rank = 0
    if args.warm_basis > 0:
        # Warm-start basis: solutions of training placements drawn uniformly on the same cap
        # (a different seed from any sweep), solved once, then the Galerkin projector.
        t0 = time.perf_counter()
        ctx.warm_reset(args.warm_basis)
        train = synthetic.cap_sites(args.warm_basis, seed=7)
        for i in range(0, len(train), 16):
            fem.solve_placements_block(ctx, pos, mom, train[i:i + 16], 1e6, metrics=False, magnitude=False,
                                       collect=True)
        rank = ctx.warm_finalize()
        train_s = time.perf_counter() - t0
        log(f"warm basis: {args.warm_basis} training solves, rank {rank}, {train_s:.2f}s")
        if args.autotune:
            overlap = autotune.tune_overlap(ctx, pos, mom, synthetic.cap_sites(8 * k, seed=19), 1e6, k,
                                            candidates=(128, 256, 512, 768, 1024, 0), warm_start=True)
            log(f"overlap (warm): coil_blocks {overlap['coil_blocks']}, rates "
                f"{ {c: round(r, 1) for c, r in overlap['rates'].items()} }")
        ctx.timing_enable(True)
        elapsed, iters, block_iters = sweep(True)
        timing = ctx.timing_read()
    else:
        elapsed, iters, block_iters = cold_elapsed, cold_iters, cold_block_iters
        timing = cold_timing
    value = args.steps * k / elapsed
    log(f"batched{' warm' if args.warm_basis > 0 else ''} k={k}: {value:.1f} placements/s, "
        f"{1e3 * elapsed / args.steps:.1f} ms/step, iterations mean {np.mean(iters):.1f}")
    # Device time per block by phase, on blocks solved one at a time; their sum against the
    # pipelined step time shows what the overlap of the coil stage with the solve buys.
    block_phases = phase_run(args.warm_basis > 0)
    block_phases["sum"] = round(sum(block_phases.values()), 3)
    log(f"block phases, blocks one at a time (ms per block of {k}): {block_phases}; pipelined step "
        f"{1e3 * elapsed / args.steps:.1f} ms")
    dadt_phase_ms = block_phases["dadt"] / k
    guess_ms = None
    if args.warm_basis > 0:
        guess_ms = {key: round(v, 4) for key, v in ctx.bench_guess(k, 5).items()}
        log(f"warm-start guess stages: {guess_ms}")

    # Plain single-placement GPU PCG on the same sites, at the reference's smoother and the
    # automatic tile (no tuning, no warm start).
    plain_iters = [0]
    plain = float("nan")
    if args.plain > 0:
        tuned = ctx.tuning
        ctx.set_tuning(degree=2, lower_ratio=12.0, fine_sub=0, coarse_sub=0)
        plain_iters = []
        block(0, 1)
        t0 = time.perf_counter()
        for i in range(args.plain):
            res = block(i, 1)
            plain_iters.append(res[0].iterations)
        plain = args.plain / (time.perf_counter() - t0)
        log(f"plain k=1: {plain:.1f} placements/s, iterations mean {np.mean(plain_iters):.1f}")
        ctx.set_tuning(**tuned)

    rounds = {}
    for kk in ((k, 1) if args.plain > 0 else (k,)):
        ex: dict = {}
        fem.solve_placements_block(ctx, pos, mom, sites[:kk], 1e6, magnitude=False, metrics=False, extra=ex)
        rounds[kk] = {"iterations": ex["round_iterations"], "true_residual": ex["round_residual"]}
        log(f"rounds k={kk}: {rounds[kk]}")

    # Rooflines from the timed sweep itself: device wall-clock stamps inside the captured
    # iteration (the differenced SpMM's own blocks, and marks around the V-cycle), averaged over
    # every iteration of the sweep.
    n_red = int(ctx.info.n_reduced)
    nnz = int(ctx.info.nnz_reduced)
    K = k if k in fem.BLOCK_SIZES else min(b for b in fem.BLOCK_SIZES if b >= k)
    degree = ctx.tuning["degree"]
    spmm_ms, vc_ms, it_ms = timing["spmm_ms"], timing["vcycle_ms"], timing["iteration_ms"]
    achieved = spmm_diff_bytes(n_red, nnz, K) / (spmm_ms * 1e-3) / 1e9
    unshared_achieved = spmm_diff_bytes(n_red, nnz, K) / (unshared_timing["spmm_ms"] * 1e-3) / 1e9
    vc_bytes = vcycle_bytes(ctx.info, K, degree)
    vc_achieved = vc_bytes / (vc_ms * 1e-3) / 1e9
    log(f"in-sweep ({timing['samples']} samples) K={K}: differenced SpMM {spmm_ms * 1e3:.1f} us "
        f"({achieved:.0f} GB/s), V-cycle {vc_ms * 1e3:.1f} us ({vc_achieved:.0f} GB/s), "
        f"iteration {it_ms * 1e3:.1f} us")
    # dA/dt: FP32 on the vector ALUs, 20 flop per (node, dipole) pair (the distance as r.s, rn + sn
    # - 2 r.s, rsqrt, cube, six multiply-adds), timed with HIP events around each block's launch.
    n_dip = pos.shape[0]
    dadt_flop = 20.0 * ctx.n_nodes * n_dip
    dadt_ms = dadt_phase_ms
    dadt_tf = dadt_flop / (dadt_ms * 1e-3) / 1e12 if dadt_ms > 0 else float("nan")
    dadt_iso = round(ctx.bench_dadt(pos, mom, sites[0], k=8, reps=3) / 8, 4)
    log(f"dA/dt {dadt_ms:.3f} ms per placement (blocks one at a time) ({dadt_tf:.1f} TFLOP/s); isolated per placement "
        f"{dadt_iso}")
    isolated = None
    if args.isolated:
        isolated = {"spmm_us": round(ctx.bench_spmm(K, 50, mode=1) * 1e3, 2),
                    "spmm_with_gram_us": round(ctx.bench_spmm(K, 50, mode=2) * 1e3, 2),
                    "plain_spmm_us": round(ctx.bench_spmm(K, 50, mode=0) * 1e3, 2),
                    "vcycle_us": round(ctx.bench_vcycle(K, 10) * 1e3, 2)}
        log(f"isolated launches: {isolated}")

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "placements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "block_phases_ms": block_phases,
        "guess_stage_ms": guess_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (fp64 Gram/refinement)",
        "data": "synthetic (seeded perturbed 5-shell head, figure-8 coil)",
        "config": {
            "workload": f"{args.spacing} mm perturbed 5-shell head, 16x16 cap-grid sweep, block_k={k}",
            "n_nodes": ctx.n_nodes, "n_tets": ctx.n_tets, "nnz_reduced": nnz,
            "hierarchy": ctx.hierarchy_sizes, "block_k": k, "tolerance": 1e-6, "recovery": "raw",
            "warm_start": (f"Galerkin projection on {args.warm_basis} solver-generated solutions (FP32 snapshot "
                           f"basis, A-orthonormal modes as BF16 hi+lo pairs, MFMA expansion)")
            if args.warm_basis > 0 else "none",
        },
        "batched_cold_placements_per_s": round(cold_value, 3),
        "cold_iterations_mean": round(float(np.mean(cold_iters)), 2),
        "cold_block_iterations_mean": round(float(np.mean(cold_block_iters)), 2),
        "block_iterations_mean": round(float(np.mean(block_iters)), 2),
        "warm_basis_rank": rank,
        "warm_training_s": round(train_s, 3),
        "value_incl_training": round(args.steps * k / (elapsed + train_s), 3),
        "plain_placements_per_s": round(plain, 3),
        "speedup_vs_plain": round(value / plain, 3),
        "iterations_mean": round(float(np.mean(iters)), 2),
        "plain_iterations_mean": round(float(np.mean(plain_iters)), 2),
        "setup_s": round(setup_s, 3),
        "autotune": tune.as_dict() if tune else None,
        "overlap": ({"coil_blocks": overlap["coil_blocks"],
                     "rates": {str(c): round(r, 2) for c, r in overlap["rates"].items()}} if overlap else None),
        "refinement_rounds": rounds,
        "roofline": {
            "kernel": f"rowk_kernel<{K},{ctx.tuning['fine_sub'] or 'auto'},kSpmmDiff> (differenced SpMM S = A W)",
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": MEASURED_SPMM_TRAFFIC.get((args.spacing, K)),
            "avg_us": round(spmm_ms * 1e3, 2),
            "samples": timing["samples"],
            "algorithmic_bytes": spmm_diff_bytes(n_red, nnz, K),
            "note": "timed inside the pipelined sweep, sharing the device with the next block's dA/dt",
        },
        # the same kernel over the blocks solved one at a time (nothing else on the device)
        "roofline_unshared": {
            "achieved": round(unshared_achieved, 1),
            "frac": round(unshared_achieved / HBM_PEAK_GBS, 4),
            "avg_us": round(unshared_timing["spmm_ms"] * 1e3, 2),
            "samples": unshared_timing["samples"],
        },
        "vcycle_roofline": {
            "bound": "hbm",
            "achieved": round(vc_achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(vc_achieved / HBM_PEAK_GBS, 4),
            "avg_us": round(vc_ms * 1e3, 2),
            "algorithmic_bytes": vc_bytes,
            "degree": degree,
        },
        "dadt_roofline": {
            "bound": "valu",
            "achieved": round(dadt_tf, 2),
            "peak": 157.3,
            "unit": "TFLOP/s",
            "frac": round(dadt_tf / 157.3, 4),
            "avg_ms_per_placement": round(dadt_ms, 4),
            "flop_per_placement": dadt_flop,
            "isolated_ms_per_placement": dadt_iso,
        },
        "iteration_us": round(it_ms * 1e3, 2),
        "cold_in_sweep_us": {"spmm": round(cold_timing["spmm_ms"] * 1e3, 2),
                             "vcycle": round(cold_timing["vcycle_ms"] * 1e3, 2),
                             "iteration": round(cold_timing["iteration_ms"] * 1e3, 2)},
        "isolated_us": isolated,
    }

    if not args.no_cpu:
        import oracle

        ncores = min(os.cpu_count() or 1, 16)
        oracle.set_threads(ncores)
        t0 = time.time()
        octx = oracle.build_context(head.nodes_mm, head.tet_nodes, head.tet_tags, head.skin_tris)
        log(f"oracle context {time.time() - t0:.1f}s")
        # Accuracy: GM |E| of sample placements against the FP64 oracle solved to 1e-12.
        errs = []
        gm = octx.tet_tags == fem.GM_TAG
        for i in range(args.accuracy):
            ref = oracle.solve_site(octx, pos, mom, sites[(i * 37) % len(sites)], 1e6, tol=1e-12)
            g = fem.solve_placements_block(ctx, pos, mom, [sites[(i * 37) % len(sites)]], 1e6,
                                           warm_start=args.warm_basis > 0)[0]
            errs.append(float(np.linalg.norm(g.magnE[gm] - ref["magnE"][gm]) / np.linalg.norm(ref["magnE"][gm])))
        out["gm_magnE_rel_err_vs_fp64"] = max(errs) if errs else None
        # CPU baseline: the oracle's FP64 AMG-PCG path to 1e-6, per placement.
        t0 = time.time()
        for i in range(args.cpu_sample):
            oracle.solve_site(octx, pos, mom, sites[i], 1e6, tol=1e-6)
        cpu_s = time.time() - t0
        out["cpu_baseline"] = {
            "value": round(args.cpu_sample / cpu_s, 4),
            "unit": "placements/s",
            "cores": ncores,
            "kind": "port",
            "sample": f"{args.cpu_sample} placements of the same sweep (dA/dt, RHS, FP64 AMG-PCG to 1e-6, "
                      f"reconstruction) after an untimed oracle setup",
        }
        log(f"cpu baseline {out['cpu_baseline']['value']} placements/s on {ncores} threads; GM err {errs}")
    if args.learned > 0:
        out["learned"] = learned_leg(ctx, pos, mom, args.learned, 256)
        log(f"learned warm start: {out['learned']}")

    print(json.dumps(out), flush=True)
    ctx.free()
    return 0


if __name__ == "__main__":
    sys.exit(main())
