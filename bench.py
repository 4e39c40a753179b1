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
    """Launches for kernel traces and counter passes, with no block solve (the block PCG is not
    traced): the context build above (assembly, pattern, AMG setup), 20 streaming copies of
    COPY_BYTES (the counter calibration), 20 differenced SpMMs at width k on the sliced ELL copy and
    20 on the CSR arrays (each set after one untimed launch), 20 SpMMs with their Gram pass, 10
    V-cycles, the recovery operator builds, the dA/dt N-body for 8 placements (3 times) and the
    block stages around the solve (coil stage, field with harmonic and with raw recovery, region
    metrics; 5 times each)."""
    from cunibs_amd import fem

    K = k if k in fem.BLOCK_SIZES else min(b for b in fem.BLOCK_SIZES if b >= k)
    n_red, nnz = int(ctx.info.n_reduced), int(ctx.info.nnz_reduced)
    pos, mom = coil.positions_m, coil.moments
    copy_ms = ctx.bench_copy(COPY_BYTES, 20)
    spmm_ms = ctx.bench_spmm(K, 20, mode=1)
    ctx.set_tuning(layout=1)
    spmm_csr_ms = ctx.bench_spmm(K, 20, mode=1)
    ctx.set_tuning(layout=0)
    gram_ms = ctx.bench_spmm(K, 20, mode=2)
    vc_ms = ctx.bench_vcycle(K, 10)
    ctx.set_tuning(dense_tail=0)
    vc_lbl_ms = ctx.bench_vcycle(K, 10)
    ctx.set_tuning(dense_tail=1)
    t0 = time.perf_counter()
    ctx.recovery_build("harmonic")
    ctx.recovery_build("spr_global")
    rec_s = time.perf_counter() - t0
    dadt_ms = ctx.bench_dadt(pos, mom, sites[0], k=8, reps=3)
    stages_h = ctx.bench_stages(pos, mom, sites[0], k=K, recovery="harmonic", reps=5)
    stages_raw = ctx.bench_stages(pos, mom, sites[0], k=K, recovery="raw", reps=5)
    print(json.dumps({
        "copy_bytes_read": COPY_BYTES, "copy_bytes_written": COPY_BYTES, "copy_us": round(copy_ms * 1e3, 2),
        "copy_GBps": round(2 * COPY_BYTES / (copy_ms * 1e-3) / 1e9, 1),
        "spmm_K": K, "n_reduced": n_red, "nnz_reduced": nnz,
        "spmm_algorithmic_bytes": spmm_diff_bytes(n_red, nnz, K), "spmm_us": round(spmm_ms * 1e3, 2),
        "spmm_csr_us": round(spmm_csr_ms * 1e3, 2), "spmm_with_gram_us": round(gram_ms * 1e3, 2),
        "vcycle_us": round(vc_ms * 1e3, 2), "vcycle_level_by_level_us": round(vc_lbl_ms * 1e3, 2),
        "dense_tail_level": int(ctx.info.dense_tail_level), "vcycle_algorithmic_bytes": vcycle_bytes(ctx.info, K, ctx.tuning["degree"]),
        "recovery_build_s": round(rec_s, 3), "dadt_ms_per_8": round(dadt_ms, 3),
        "stages_harmonic_ms": stages_h, "stages_raw_ms": stages_raw,
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
    run = lambda n: fem.solve_sweep(ctx, pos, mom, [sites[j % len(sites)] for j in range(n * k)], 1e6, k,
                                    recovery=args.recovery, recycle=args.recycle)
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
        "sweep_placements_per_s": round(args.steps * k / elapsed, 2), "block_k": k, "recovery": args.recovery,
        "recycle": args.recycle,
        "tuning": ctx.tuning,
        "iterations_mean": round(float(np.mean(res["iterations"])), 2),
        "spmm_us": round(timing["spmm_ms"] * 1e3, 2), "spmm_samples": timing["samples"],
        "spmm_algorithmic_bytes": spmm_diff_bytes(n_red, nnz, K),
        "vcycle_us": round(timing["vcycle_ms"] * 1e3, 2), "iteration_us": round(timing["iteration_ms"] * 1e3, 2),
    }), flush=True)
    ctx.free()
    return 0


def learned_leg(ctx, pos, mom, n_train: int, n_test: int, log_sd: float = 0.2, n_diag: int = 32) -> dict:
    """The learned warm start on poses and conductivities (distance U[2, 8] mm, the five tissues of
    the head log-normal around the table): n_train solver-generated solutions, each with its own
    conductivities, form the snapshot basis and the training set; the network maps pose and log
    conductivities to the solution's coefficients in the 256 leading A-orthonormal modes; on
    n_test held-out placements it warm-starts PCG, which iterates to the same tolerance. Two
    networks are trained, on energy-scaled and on per-mode standardized targets. For the first
    n_diag held-out placements the energy-norm error ||x - x0||_A / ||x||_A of each initial guess
    is measured exactly (x^T b from the solution and right-hand side, Phi^T b from the device):
    the Galerkin guess is the best the modes allow, and a network predicting Galerkin
    coefficients cannot do better."""
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
    T = np.zeros((n_train, (rank + 15) // 16 * 16), np.float32)
    T[:, :rank] = ctx.warm_targets()[:, :rank]
    X_tr = feats(tr_sites, tr_rows)
    nets, fits = {}, {}
    for scale in ("energy", "per_output"):
        net = learned.WarmStartNetwork(n_out=T.shape[1], hidden=512, seed=0)
        t1 = time.perf_counter()
        loss = net.fit(X_tr, T, epochs=300, target_scale=scale)
        net.modes = rank
        pred = net.predict(X_tr)[:, :rank]
        fits[scale] = {"train_loss": round(loss, 5), "network_train_s": round(time.perf_counter() - t1, 2),
                       "train_coef_rel_err": round(float(np.linalg.norm(pred - T[:, :rank])
                                                         / np.linalg.norm(T[:, :rank])), 4)}
        nets[scale] = net

    te_sites, te_rows = draws(n_test, 37)
    X = feats(te_sites, te_rows)
    preds = {s: nets[s].predict(X)[:, :rank].astype(np.float64) for s in nets}

    def block(mode, i, **kw):
        s, r = te_sites[i:i + 16], te_rows[i:i + 16]
        if mode.startswith("learned"):
            ctx.warm_set_coefficients(preds["per_output" if mode == "learned_per_output" else "energy"][i:i + 16])
        return fem.solve_placements_block(ctx, pos, mom, s, 1e6, metrics=True, magnitude=True, sigma=r,
                                          warm_start={"cold": 0, "galerkin": 1}.get(mode, 2), **kw)

    def run(mode):
        """Timed: the held-out placements in blocks of 16 (after one untimed block that captures the
        mode's graphs). Untimed: the guess errors of the first n_diag placements."""
        block(mode, 0)
        its, block_its, mags, errs = [], [], [], []
        t = time.perf_counter()
        for i in range(0, n_test, 16):
            res = block(mode, i)
            its.extend(x.iterations for x in res)
            block_its.append(max(x.iterations for x in res))
            mags.extend(x.magnE for x in res[:2])
        rate = n_test / (time.perf_counter() - t)
        keep = np.ones(ctx.n_nodes, bool)
        keep[ctx.ground_node] = False
        for i in range(0, n_diag if mode != "cold" else 0, 16):
            extra = {"rhs": 1}
            res = block(mode, i, potential=True, extra=extra)
            energy = np.array([float(x.v[keep] @ b[keep].astype(np.float64)) for x, b in zip(res, extra["rhs"])])
            coef = None if mode == "galerkin" else \
                preds["per_output" if mode == "learned_per_output" else "energy"][i:i + len(res)]
            errs.extend(ctx.warm_guess_errors(len(res), energy, coef).tolist())
        return rate, its, block_its, mags, errs

    # Galerkin coefficients P Phi^T b of every held-out placement (untimed), against which the
    # networks' held-out predictions are scored.
    c_gal = []
    for i in range(0, n_test, 16):
        block("galerkin", i)
        c_gal.append(ctx.warm_project_last(len(te_sites[i:i + 16])) @ ctx.warm_state["P"].T)
    c_gal = np.concatenate(c_gal)
    for scale in nets:
        fits[scale]["test_coef_rel_err"] = round(float(np.linalg.norm(preds[scale] - c_gal) / np.linalg.norm(c_gal)), 4)

    out = {}
    ref = None
    for mode in ("cold", "galerkin", "learned", "learned_per_output"):
        rate, its, bits, mags, errs = run(mode)
        out[mode] = {"placements_per_s": round(rate, 2), "iterations_mean": round(float(np.mean(its)), 2),
                     "block_iterations_mean": round(float(np.mean(bits)), 2)}
        if errs:
            out[mode]["guess_energy_err_median"] = round(float(np.median(errs)), 4)
            out[mode]["guess_energy_err_max"] = round(float(np.max(errs)), 4)
        if mode.startswith("learned"):
            out[mode].update(fits["per_output" if mode == "learned_per_output" else "energy"])
        if ref is None:
            ref = mags
        else:
            out[mode]["magnE_rel_diff_vs_cold"] = max(
                float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(mags, ref))
    out.update({"train_solves": n_train, "test_placements": n_test, "modes": rank, "diagnosed_placements": n_diag,
                "solve_s": round(t_solve, 2), "basis_s": round(t_basis, 2)})
    for net in nets.values():
        net.free()
    return out


def spmm_layout(ctx, K: int) -> str:
    """The form of the finest differenced operator the block iteration's SpMM reads at width K."""
    layout = ctx.tuning.get("layout", 0)
    ell = ctx.info.ell_width_off > 0 and (layout == 2 or (layout == 0 and K <= 4))
    return "sliced ELL" if ell else "CSR"


def cpu_host() -> dict:
    """The host the CPU baseline runs on: logical CPUs and the model name (/proc/cpuinfo, as lscpu
    reports it)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "model": model}


def config3_leg(pos, mom, tuning: dict, block_k: int, recovery: str) -> dict:
    """The 256-placement sweep spread over three perturbed heads (2.0 / 1.5 / 1.25 mm grids, 86 / 85
    / 85 placements of the 16 x 16 cap grid, each head's share taking every third grid site, so
    every head ends on a partial block): per head the mesh size, the context setup (with the
    recovery operator) and the pipelined sweep at the tuning found on the 1.5 mm head; the whole sweep's placements/s with
    and without the three setups."""
    from cunibs_amd import fem, synthetic

    grid = synthetic.grid_sites()
    heads, total_solve, total_setup = [], 0.0, 0.0
    for j, spacing in enumerate((2.0, 1.5, 1.25)):
        sites = grid[j::3]
        head = synthetic.shell_head(spacing, perturb_rms_mm=2.0, seed=1)
        t0 = time.perf_counter()
        ctx = fem.SolverContext(head.nodes_mm, head.tet_nodes, head.tet_tags, head.skin_tris)
        ctx.set_tuning(**tuning)
        if recovery != "raw":
            ctx.recovery_build(recovery)
        setup = time.perf_counter() - t0
        fem.solve_sweep(ctx, pos, mom, sites[:block_k], 1e6, block_k, recovery=recovery)  # captures
        t0 = time.perf_counter()
        res = fem.solve_sweep(ctx, pos, mom, sites, 1e6, block_k, recovery=recovery)
        solve = time.perf_counter() - t0
        if not np.all(res["relative_residual"] <= 1e-6):
            raise RuntimeError(f"a placement of the {spacing} mm sweep missed the tolerance")
        heads.append({"spacing_mm": spacing, "n_nodes": ctx.n_nodes, "n_tets": ctx.n_tets,
                      "placements": len(sites), "setup_s": round(setup, 3), "sweep_s": round(solve, 3),
                      "placements_per_s": round(len(sites) / solve, 2),
                      "iterations_mean": round(float(np.mean(res["iterations"])), 2)})
        total_solve += solve
        total_setup += setup
        ctx.free()
        log(f"config 3 head {spacing} mm: {heads[-1]}")
    n = sum(h["placements"] for h in heads)
    return {"heads": heads, "placements": n, "placements_per_s": round(n / total_solve, 2),
            "placements_per_s_incl_setup": round(n / (total_solve + total_setup), 2), "recovery": recovery,
            "block_k": block_k}


def config4_leg(ctx, head, pos, mom, n_draws: int, oracle_draws: int) -> dict:
    """Conductivity uncertainty at one placement: n_draws draws of every tissue's conductivity
    (log-normal, log-sd 0.2, mean-preserving), 16 per launch as independent chains under the nominal
    preconditioner, each starting from the nominal solution's Galerkin correction; the GM peak |E|
    mean and sd over the draws, and, for `oracle_draws` of them, the peak against the FP64 oracle
    solved at that draw's conductivities."""
    from cunibs_amd import fem, synthetic, uq

    site = synthetic.grid_sites()[8]
    t0 = time.perf_counter()
    tags = ctx.uq_prepare(pos, mom, site, 1e6)
    prep = time.perf_counter() - t0
    sigma = uq.draw_lognormal(tags, n_draws, 0.2, seed=12)
    ctx.uq_solve(sigma[:16], magnitude=False)  # captures
    peaks, its = [], []
    t0 = time.perf_counter()
    for i in range(0, n_draws, 16):
        res = ctx.uq_solve(sigma[i:i + 16], magnitude=False)
        peaks.extend(float(r.metrics[0]) for r in res)
        its.extend(r.iterations for r in res)
    solve = time.perf_counter() - t0
    out = {"draws": n_draws, "per_launch": 16, "draws_per_s": round(n_draws / solve, 2),
           "prepare_s": round(prep, 3), "iterations_mean": round(float(np.mean(its)), 2),
           "gm_peak_mean": float(np.mean(peaks)), "gm_peak_sd": float(np.std(peaks, ddof=1)),
           "matched_preconditioners": int(ctx.info.matched_preconditioners)}
    if oracle_draws:
        import oracle

        errs = []
        for i in range(oracle_draws):
            table = dict(fem.TISSUE_CONDUCTIVITY)
            table.update({t: float(v) for t, v in zip(tags, sigma[i])})
            octx = oracle.build_context(head.nodes_mm, head.tet_nodes, head.tet_tags, head.skin_tris,
                                        conductivity=table)
            ref = oracle.solve_site(octx, pos, mom, site, 1e6, tol=1e-12)
            gm = octx.tet_tags == fem.GM_TAG
            errs.append(abs(peaks[i] - float(ref["magnE"][gm].max())) / float(ref["magnE"][gm].max()))
        out["gm_peak_rel_err_vs_fp64"] = max(errs)
        out["oracle_draws"] = oracle_draws
    return out


def cpu_baseline_leg(head, pos, mom, sites, sample: int) -> dict:
    """The FP64 CPU oracle on this host: `sample` placements of the sweep on min(nproc, 16)
    threads and one on a single thread (dA/dt, RHS, AMG-PCG to 1e-6, reconstruction, after an
    untimed setup), and the 4 mm sphere configuration end to end on the same threads (assembly,
    AMG setup and one placement to 1e-10)."""
    import oracle
    from cunibs_amd import synthetic

    host = cpu_host()
    threads = min(host["nproc"] or 1, 16)
    oracle.set_threads(threads)
    t0 = time.time()
    octx = oracle.build_context(head.nodes_mm, head.tet_nodes, head.tet_tags, head.skin_tris)
    setup = time.time() - t0
    t0 = time.time()
    for i in range(sample):
        oracle.solve_site(octx, pos, mom, sites[i], 1e6, tol=1e-6)
    multi = sample / (time.time() - t0)
    oracle.set_threads(1)
    t0 = time.time()
    oracle.solve_site(octx, pos, mom, sites[sample], 1e6, tol=1e-6)
    single = 1.0 / (time.time() - t0)
    oracle.set_threads(threads)
    sphere = synthetic.shell_head(4.0)
    t0 = time.time()
    sctx = oracle.build_context(sphere.nodes_mm, sphere.tet_nodes, sphere.tet_tags, sphere.skin_tris)
    t1 = time.time()
    oracle.solve_site(sctx, pos, mom, synthetic.SPHERE_SITE, 1e6, tol=1e-10)
    t2 = time.time()
    return {
        "value": round(multi, 4), "unit": "placements/s", "cores": threads, "kind": "port",
        "sample": f"{sample} placements of the 1.5 mm sweep (dA/dt, RHS, FP64 AMG-PCG to 1e-6, reconstruction) "
                  f"after an untimed oracle setup ({setup:.1f} s)",
        "nproc": host["nproc"], "model": host["model"],
        "single_thread_placements_per_s": round(single, 4),
        "config1_end_to_end_s": {"setup": round(t1 - t0, 3), "solve_1e-10": round(t2 - t1, 3),
                                 "total": round(t2 - t0, 3), "n_nodes": int(sphere.nodes_mm.shape[0])},
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16, help="blocks of the timed sweep (steps x block_k placements)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--block-k", type=int, default=16)
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
    ap.add_argument("--learned", type=int, default=2048,
                    help="training solves of the learned warm start on random poses and conductivities, "
                         "evaluated on 256 held-out placements (0: skip)")
    ap.add_argument("--config3", type=int, default=1, help="1: the 256-placement sweep over three heads")
    ap.add_argument("--uq-draws", type=int, default=512, help="conductivity draws of the UQ leg (0: skip)")
    ap.add_argument("--graphs", type=int, default=1,
                    help="0: launch every iteration's kernels directly instead of replaying its captured graph")
    ap.add_argument("--fine-sub", type=int, default=-1,
                    help="--kernels-only / --sweep-only: fine-level tile (sub-lanes) to measure at (-1: automatic)")
    ap.add_argument("--isolated", type=int, default=0,
                    help="1: also time the SpMM and V-cycle in isolated repeated launches (warm caches)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the roofline launches (copy calibration + differenced SpMM), for counter runs")
    ap.add_argument("--sweep-only", action="store_true",
                    help="only the pipelined cold sweep at --block-k, --fine-sub and --coil-blocks (no tuning, no "
                         "blocks solved alone), so that a kernel trace's averages are the in-sweep ones")
    ap.add_argument("--recovery", default="harmonic", choices=("raw", "spr_global", "spr_tissue", "harmonic"),
                    help="--sweep-only: the field recovery of the sweep (the headline's is harmonic)")
    ap.add_argument("--coil-blocks", type=int, default=256,
                    help="--sweep-only: grid cap of the overlapped dA/dt")
    ap.add_argument("--recycle", type=int, default=2,
                    help="--sweep-only: earlier blocks in each block's Galerkin initial guess (0: from zero)")
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
    setup_phases = ctx.setup_times()
    log(f"context {setup_s:.2f}s hierarchy {ctx.hierarchy_sizes}")
    k = args.block_k
    pos, mom = coil.positions_m, coil.moments
    if args.kernels_only:
        if args.fine_sub >= 0:
            ctx.set_tuning(fine_sub=args.fine_sub)
        return kernels_only(ctx, coil, sites, k)
    if args.sweep_only:
        return sweep_only(ctx, coil, sites, args)
    t0 = time.time()
    ctx.recovery_build("harmonic")
    harmonic_build_s = time.time() - t0

    def block(i0: int, kk: int, warm: bool = False, recovery: str = "raw"):
        chunk = [sites[(i0 + j) % len(sites)] for j in range(kk)]
        return fem.solve_placements_block(ctx, pos, mom, chunk, 1e6, magnitude=False, metrics=True,
                                          warm_start=warm, use_graph=bool(args.graphs), recovery=recovery)

    def sweep(warm: bool, recovery: str, stamps: bool = False, recycle: int = 0):
        """The timed sweep: steps x k placements of the 16 x 16 grid through the pipelined sweep entry
        point (the next block's dA/dt and RHS overlap the current block's solve), after warmup
        blocks; `recycle` earlier blocks' solutions in each block's Galerkin initial guess (0: every
        block from zero). With `stamps` the captured iterations carry the device-clock stamps the
        rooflines are read from (a few launches more per iteration); the headline sweeps run
        without them."""
        ctx.timing_enable(stamps)  # before the warmup: switching drops the captured iterations
        if args.warmup:
            fem.solve_sweep(ctx, pos, mom, [sites[j % len(sites)] for j in range(args.warmup * k)], 1e6, k,
                            warm_start=warm, use_graph=bool(args.graphs), recovery=recovery, recycle=recycle)
        chunk = [sites[j % len(sites)] for j in range(args.steps * k)]
        if stamps:
            ctx.timing_enable(True)  # reset the sums: they cover this sweep alone
        t_start = time.perf_counter()
        res = fem.solve_sweep(ctx, pos, mom, chunk, 1e6, k, warm_start=warm, use_graph=bool(args.graphs),
                              recovery=recovery, recycle=recycle)
        elapsed = time.perf_counter() - t_start
        timing = ctx.timing_read() if stamps else None
        its = [int(i) for i in res["iterations"]]
        block_its = [max(its[b:b + k]) for b in range(0, len(its), k)]
        if not np.all(res["relative_residual"] <= 1e-6):
            raise RuntimeError("a placement of the sweep missed the tolerance")
        return {"elapsed": elapsed, "iters": its, "block_iters": block_its, "timing": timing,
                "rate": args.steps * k / elapsed}

    ctx.timing_enable(True)
    tune = None
    if args.autotune:
        from cunibs_amd import autotune

        tune = autotune.autotune(ctx, pos, mom, synthetic.cap_sites(32, seed=13), 1e6, block_ks=(4, 8, 16, 32),
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
    # Sweeps at the reference's default recovery (harmonic) and at raw: every block from zero
    # (cold), and each block from the Galerkin projection onto the previous two blocks' solutions
    # (recycled, the sweep's default; nothing to build).
    cold = sweep(False, "harmonic")
    cold_raw = sweep(False, "raw")
    recycled = sweep(False, "harmonic", recycle=2)
    recycled_raw = sweep(False, "raw", recycle=2)
    log(f"batched cold k={k}: harmonic {cold['rate']:.1f} placements/s ({1e3 * cold['elapsed'] / args.steps:.1f} "
        f"ms/step), raw {cold_raw['rate']:.1f}; iterations mean {np.mean(cold['iters']):.1f}")
    log(f"batched recycled k={k}: harmonic {recycled['rate']:.1f} placements/s "
        f"({1e3 * recycled['elapsed'] / args.steps:.1f} ms/step), raw {recycled_raw['rate']:.1f}; "
        f"iterations mean {np.mean(recycled['iters']):.1f}")
    warm = warm_raw = None
    train_s = 0.0
    rank = 0
    cold_tuning = ctx.tuning
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
            overlap_warm = autotune.tune_overlap(ctx, pos, mom, synthetic.cap_sites(8 * k, seed=19), 1e6, k,
                                                 candidates=(128, 256, 512, 768, 1024, 0), warm_start=True)
            log(f"overlap (warm): coil_blocks {overlap_warm['coil_blocks']}")
        warm = sweep(True, "harmonic")
        warm_raw = sweep(True, "raw")
        log(f"batched warm k={k}: harmonic {warm['rate']:.1f} placements/s, raw {warm_raw['rate']:.1f}; "
            f"iterations mean {np.mean(warm['iters']):.1f}")
        ctx.set_tuning(**cold_tuning)
    n_sweep = args.steps * k
    warm_incl = n_sweep / (warm["elapsed"] + train_s) if warm else None
    # The headline: the sweep at the drop-in default recovery, by the method that is fastest for a
    # sweep of this size once the warm start's snapshot solves and mode build are charged to it.
    best_free = max((cold, "cold (every block from zero)"),
                    (recycled, "recycled (each block from the Galerkin projection onto the previous two "
                               "blocks' solutions)"), key=lambda r: r[0]["rate"])
    if warm and warm_incl > best_free[0]["rate"]:
        value, method, head_run = warm_incl, "warm-started, snapshot build included", warm
    else:
        value, method, head_run = best_free[0]["rate"], best_free[1], best_free[0]
    head_recycle = 2 if head_run is recycled else 0
    breakeven = None
    if warm and warm["rate"] > best_free[0]["rate"]:
        breakeven = int(np.ceil(train_s / (1.0 / best_free[0]["rate"] - 1.0 / warm["rate"])))
    # the headline method's sweep again with the device-clock stamps, for the in-solve rooflines
    head_stamped = sweep(head_run is warm, "harmonic", stamps=True, recycle=head_recycle)
    timing = head_stamped["timing"]
    # Device time per block by phase, on blocks solved one at a time.
    ctx.timing_enable(True)
    for b in range(4):
        block(b * k, k, False, "harmonic")
    ph = ctx.timing_phases()
    unshared_timing = ctx.timing_read()
    block_phases = {p: round(v, 3) for p, v in ph["ms_per_block"].items()}
    block_phases["sum"] = round(sum(block_phases.values()), 3)
    log(f"block phases, blocks one at a time (ms per block of {k}): {block_phases}; pipelined step "
        f"{1e3 * head_run['elapsed'] / args.steps:.1f} ms")
    dadt_phase_ms = block_phases["dadt"] / k

    # Plain single-placement GPU PCG on the same sites, at the reference's smoother and the
    # automatic tile (no tuning, no warm start), harmonic recovery as the headline.
    plain_iters = [0]
    plain = float("nan")
    if args.plain > 0:
        tuned = ctx.tuning
        ctx.set_tuning(degree=2, lower_ratio=12.0, fine_sub=0, coarse_sub=0)
        plain_iters = []
        block(0, 1, recovery="harmonic")
        t0 = time.perf_counter()
        for i in range(args.plain):
            res = block(i, 1, recovery="harmonic")
            plain_iters.append(res[0].iterations)
        plain = args.plain / (time.perf_counter() - t0)
        log(f"plain k=1: {plain:.1f} placements/s, iterations mean {np.mean(plain_iters):.1f}")
        ctx.set_tuning(**tuned)

    rounds = {}
    for kk in ((k, 1) if args.plain > 0 else (k,)):
        ex: dict = {}
        fem.solve_placements_block(ctx, pos, mom, sites[:kk], 1e6, magnitude=False, metrics=False, extra=ex)
        rounds[kk] = {"iterations": ex["round_iterations"], "true_residual": ex["round_residual"]}

    # Rooflines from the timed cold sweep itself: device wall-clock stamps inside the captured
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
    n_dip = pos.shape[0]
    dadt_flop = 20.0 * ctx.n_nodes * n_dip
    dadt_ms = dadt_phase_ms
    dadt_tf = dadt_flop / (dadt_ms * 1e-3) / 1e12 if dadt_ms > 0 else float("nan")
    dadt_iso = round(ctx.bench_dadt(pos, mom, sites[0], k=8, reps=3) / 8, 4)
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
        "ms_per_step": round(1e3 * head_run["elapsed"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (fp64 Gram/refinement)",
        "data": "synthetic (seeded perturbed 5-shell head, figure-8 coil)",
        "config": {
            "workload": f"{args.spacing} mm perturbed 5-shell head, {n_sweep} placements of the 16x16 cap grid, "
                        f"block_k={k}, harmonic recovery",
            "n_nodes": ctx.n_nodes, "n_tets": ctx.n_tets, "nnz_reduced": nnz,
            "hierarchy": ctx.hierarchy_sizes, "block_k": k, "tolerance": 1e-6, "recovery": "harmonic",
            "method": method,
        },
        "batched_cold_placements_per_s": round(cold["rate"], 3),
        "batched_cold_raw_placements_per_s": round(cold_raw["rate"], 3),
        "cold_iterations_mean": round(float(np.mean(cold["iters"])), 2),
        "cold_block_iterations_mean": round(float(np.mean(cold["block_iters"])), 2),
        "batched_recycled_placements_per_s": round(recycled["rate"], 3),
        "batched_recycled_raw_placements_per_s": round(recycled_raw["rate"], 3),
        "recycled_iterations_mean": round(float(np.mean(recycled["iters"])), 2),
        "recycled_block_iterations_mean": round(float(np.mean(recycled["block_iters"])), 2),
        "headline_stamped_placements_per_s": round(head_stamped["rate"], 3),
        "warm": None if not warm else {
            "placements_per_s_excl_build": round(warm["rate"], 3),
            "raw_placements_per_s_excl_build": round(warm_raw["rate"], 3),
            "placements_per_s_incl_build": round(warm_incl, 3),
            "snapshot_build_s": round(train_s, 3),
            "basis": f"Galerkin projection on {args.warm_basis} solver-generated solutions (FP32 snapshots, "
                     f"A-orthonormal modes as BF16 hi+lo pairs, MFMA expansion)",
            "rank": rank,
            "iterations_mean": round(float(np.mean(warm["iters"])), 2),
            "block_iterations_mean": round(float(np.mean(warm["block_iters"])), 2),
            "break_even_placements": breakeven,
        },
        "plain_placements_per_s": round(plain, 3),
        "speedup_vs_plain": round(value / plain, 3),
        "plain_iterations_mean": round(float(np.mean(plain_iters)), 2),
        "setup_s": round(setup_s, 3),
        "setup_phases_ms": setup_phases,
        "harmonic_build_s": round(harmonic_build_s, 3),
        "block_phases_ms": block_phases,
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
            "layout": spmm_layout(ctx, K),
            "note": "timed inside the headline method's pipelined sweep, sharing the device with the next block's "
                    "dA/dt",
        },
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
        "isolated_us": isolated,
    }

    if not args.no_cpu:
        import oracle

        oracle.set_threads(min(os.cpu_count() or 1, 16))
        octx = oracle.build_context(head.nodes_mm, head.tet_nodes, head.tet_tags, head.skin_tris)
        # Accuracy: GM |E| of sample placements against the FP64 oracle solved to 1e-12 (raw
        # field: the oracle's recovery is checked separately).
        errs = []
        gm = octx.tet_tags == fem.GM_TAG
        for i in range(args.accuracy):
            site = sites[(i * 37) % len(sites)]
            ref = oracle.solve_site(octx, pos, mom, site, 1e6, tol=1e-12)
            g = fem.solve_placements_block(ctx, pos, mom, [site], 1e6)[0]
            errs.append(float(np.linalg.norm(g.magnE[gm] - ref["magnE"][gm]) / np.linalg.norm(ref["magnE"][gm])))
        out["gm_magnE_rel_err_vs_fp64"] = max(errs) if errs else None
        del octx
        out["cpu_baseline"] = cpu_baseline_leg(head, pos, mom, sites, args.cpu_sample)
        log(f"cpu baseline {out['cpu_baseline']}; GM err {errs}")
    tuning = ctx.tuning
    if args.uq_draws > 0:
        out["conductivity_uq"] = config4_leg(ctx, head, pos, mom, args.uq_draws, 0 if args.no_cpu else 2)
        log(f"conductivity UQ: {out['conductivity_uq']}")
    if args.learned > 0:
        out["learned"] = learned_leg(ctx, pos, mom, args.learned, 256)
        log(f"learned warm start: {out['learned']}")
    ctx.free()
    if args.config3:
        out["three_head_sweep"] = config3_leg(pos, mom, tuning, k, "harmonic")
        log(f"config 3: {out['three_head_sweep']}")

    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
