"""Benchmark: PBE EXC+VXC grid points/s on one MI355X.

Workload (one "step" = one full EXC+VXC evaluation over every task):
  150-water cluster in a 17 A box, O 5s4p2d1f / H 3s2p even-tempered
  (beta 2.5) -> 7800 basis functions, 99 x 590 Mura-Knowles/Lebedev grid per
  atom (26.3M points, 133,238 tasks), SSF weights, 512-point batches, basis
  screening 1e-10, P = 2 C C^T with C of width 750: the configuration the
  mixed-vs-FP64 target is quoted on.  --config c1 runs the 24-water cluster
  (10 A box, 1248 functions, C width 120).  Inputs (tasks, weights, P,
  basis) are resident in HBM before the timed region; VXC stays in HBM.

Both precisions run on the same resident inputs.  Before timing, the device
autotuner picks the tile thresholds and queue count for each precision
(untimed); the timed steps replay the captured graph of the evaluation.
``value`` is the mixed-precision rate; the FP64 rate, the speed-up and the
mixed-vs-FP64 |dEXC| / max|dVXC| are reported beside it.

Prints one JSON line (see the driver contract) with a roofline entry for
the dominant stage (its kernel launches timed with HIP events on the
context's stream, in an eager step) and a CPU baseline: the oracle (a
restatement of GauXC's host integrator, its two per-task products through
a host BLAS when one loads) timed on a bounded sample of the same tasks
with the host threads this process may use.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_BF16_TFLOPS = 2500.0
# Matrix unit each stage's mixed-precision kernels run on (the FP64 path
# uses the FP64 MFMA throughout): the density stage's X GEMM is FP32, the
# B^T Z + Z^T B contraction BF16 slices.
MIXED_PEAK = {"density": PEAK_FP32_TFLOPS, "syr2k_scatter": PEAK_BF16_TFLOPS}
PEAK_HBM_GBS = 8000.0
STAGES = ["collocation", "density", "functional", "zmat", "syr2k_scatter", "symmetrize"]
# Kernels of each stage (the names rocprofv3 reports); the mixed path's
# kernels depend on its basis evaluation mode (0: stored planes, 1 / 3:
# evaluated in the density stage, 3 with Z from its own launch).


def stage_kernels(stage, prec, mode=0):
    t = "double" if prec == "fp64" else "float"
    if stage == "collocation":
        return [f"k_collocation<{'true' if mode == 0 else 'false'}, {t}, 3>"]
    if stage == "zmat":
        return [f"k_z_in_stage<true, 3, {t}>"] if mode == 3 else []
    if prec == "fp64":
        return {"density": [f"k_density_f64<{nt}, true, {mode}>" for nt in (32, 64)],
                "syr2k_scatter": ["k_syr2k<32>", "k_syr2k<64>"]}.get(stage, [])
    if stage == "density":
        return [f"k_density_mx<{nt}, true, {hi}, {mode}>"
                for nt, hi in ((32, "false"), (64, "false"), (64, "true"), (128, "false"))
                if not (hi == "true" and mode == 1)]
    if stage == "syr2k_scatter":
        return ["k_syr2k_mx<32, 1>", "k_syr2k_mx<64, 1>"]
    return []


# rocprofv3 PMC summaries (FETCH_SIZE / WRITE_SIZE per dispatch) of each
# workload, keyed by --config.
PMC_SUMMARY = {"c1": os.path.join(ROOT, "profiles", "pmc_c1_24water.json"),
               "c4": os.path.join(ROOT, "profiles", "pmc_c4_150water.json")}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic(stage, prec, ts, mode=0):
    """Algorithmic (bytes, flops) of one evaluation's launches of a stage.

    Per task t (npts, nbe).  Bytes per (point, function) element of the
    planes: FP64 stores B, dB/dx, dB/dy, dB/dz, Z at 8 B each; mixed stores
    the value plane B at 6 B (FP32 + bfloat16 residual) and the other four
    at 4 B.  P_sub is gathered once per task (8 B FP64, 4 B FP32 copy).
      collocation  writes B and grad B (+ 32 npts for points and weights)
      density      reads B and grad B, writes Z; gathers P_sub;
                   flops 2 npts nbe^2 (X = B P_sub)
      syr2k        reads B and Z; VXC lower triangle read + write
                   8 nbe (nbe + 1); flops 2 npts nbe^2 (B^T Z + Z^T B, lower)
    With the in-stage basis evaluation (mode 1 or 3; FP64 runs 3) the
    collocation stage writes only the value plane, the density stage reads
    only the X operand's part of it (FP32 for mixed, FP64) and in mode 3
    writes four FP64 Z coefficients per point instead of Z, which the Z
    launch ("zmat") reads before writing Z.
    """
    npts = ts["npts"].astype(np.float64)
    nbe = ts["nbe"].astype(np.float64)
    if prec == "fp64":
        bv, bg, bz, bp = 8.0, 8.0, 8.0, 8.0
    else:
        bv, bg, bz, bp = 6.0, 4.0, 4.0, 4.0
    pn = float(np.sum(npts * nbe))
    pnn = float(np.sum(npts * nbe * nbe))
    nn = float(np.sum(nbe * nbe))
    if mode in (1, 3):
        xop = 8.0 if prec == "fp64" else 4.0
        if stage == "collocation":
            return bv * pn + 32 * float(npts.sum()), 0.0
        if stage == "density":
            zout = 32 * float(npts.sum()) if mode == 3 else bz * pn
            return xop * pn + zout + bp * nn, 2 * pnn
        if stage == "zmat" and mode == 3:
            return bz * pn + 32 * float(npts.sum()), 0.0
    if stage == "collocation":
        return (bv + 3 * bg) * pn + 32 * float(npts.sum()), 0.0
    if stage == "density":
        return (bv + 3 * bg + bz) * pn + bp * nn, 2 * pnn
    if stage == "syr2k_scatter":
        return (bv + bz) * pn + 8 * float(np.sum(nbe * (nbe + 1))), 2 * pnn
    return 0.0, 0.0


def measured_traffic(config, stage, prec, mode=0):
    """HBM bytes of one evaluation's launches of the stage, from the committed
    rocprofv3 PMC summary of the workload: FETCH_SIZE and WRITE_SIZE per
    dispatch, scaled by the copy calibration factors it records, times the
    kernel's dispatches per evaluation (its dispatch count over that of
    k_reduce_final, which runs once per evaluation); None without a summary."""
    try:
        with open(PMC_SUMMARY[config]) as f:
            pmc = json.load(f)
    except (KeyError, OSError, ValueError):
        return None
    kern = pmc.get("kernels", {})
    evals = kern.get("k_reduce_final", {}).get("dispatches", 0)
    if not evals:
        return None
    total = 0.0
    found = False
    for k in stage_kernels(stage, prec, mode):
        e = kern.get(k)
        if e:
            per = (e["fetch_kib"] * pmc["fetch_scale"] + e["write_kib"] * pmc["write_scale"]) * 1024
            total += per * e["dispatches"] / evals
            found = True
    return total if found else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="auto", choices=["auto", "fp64", "mixed"])
    ap.add_argument("--config", default="c4", choices=["c1", "c3", "c4"],
                    help="c4: 150-water cluster in a 17 A box, 7800 bf (the workload the "
                         "metric's target is quoted on, default); c1: 24-water cluster; "
                         "c3: 64 independent clusters of 2-40 waters, batched")
    ap.add_argument("--learned", action="store_true",
                    help="also compare the learned cost model against GauXC's flop "
                         "heuristic (prediction error and queued wall time)")
    ap.add_argument("--n-mol", type=int, default=64, help="molecules of the c3 batch")
    ap.add_argument("--n-water", type=int, default=None)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--derivatives", action="store_true",
                    help="also time the FP64 EXC gradient (with and without the grid-weight "
                         "derivatives), the RKS FXC contraction and the UKS EXC/VXC on the "
                         "same workload")
    ap.add_argument("--tile-policy", default=None,
                    help="D,S,Q[,H[,W[,M]]]: fixed density / SYR2K 32-wide thresholds, "
                         "queue count, (mixed) density high-occupancy threshold, 128-wide "
                         "minimum nbe and basis evaluation mode for every precision "
                         "(skips the autotuner)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device copy (calibrates rocprofv3 "
                         "FETCH_SIZE/WRITE_SIZE against a known byte count)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = 1 << 28  # 2^28 float32 = 1 GiB
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "repeats": 3}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is described, not built)"}))
        sys.exit(2)

    import gauxc_amd as gx
    from gauxc_amd import workloads

    assert gx.device_count() > 0, "bench.py needs a HIP device"
    if args.config == "c3":
        return run_batched(args)
    # The device context first: the HIP runtime's one-time start-up then
    # stays out of the setup times below.
    ctx = gx.DeviceContext(0)
    t0 = time.time()
    mol, bs, P, cfg = workloads.make_config(args.config, seed=0, n_water=args.n_water)
    # Task generation with the batch screening on the device (GauXC's device
    # load balancer), then the SSF weights on the device.
    lb = workloads.build_tasks(mol, bs, grid=cfg["grid"], ex="Device")
    t1 = time.time()
    lb.task_counts()
    t2 = time.time()
    gx.MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
    t3 = time.time()
    t = lb.get_tasks()
    npts_total = int(t["npts"].sum())
    setup = {"task_build_s": round(t2 - t1, 3), "ssf_weights_s": round(t3 - t2, 3)}
    log(f"setup: nbf={cfg['nbf']} atoms={len(cfg['Z'])} tasks={len(t['npts'])} "
        f"points={npts_total} ({time.time() - t0:.1f}s; tasks {setup['task_build_s']} s with "
        f"device screening, SSF {setup['ssf_weights_s']} s)")

    ctx.upload_basis(bs.xcq_shells())
    # GauXC processes tasks in descending npts*nbe order.
    order = np.argsort(-(t["npts"].astype(np.int64) * t["nbe"].astype(np.int64)), kind="stable")
    ts = reorder_tasks(t, order)
    ctx.upload_tasks(ts)
    ctx.upload_density(P)
    precisions = ["fp64", "mixed"] if args.precision == "auto" else [args.precision]
    results = {}
    for prec in precisions:
        tuned = None
        if args.tile_policy:
            vals = [int(x) for x in args.tile_policy.split(",")]
            d_, s_, q_ = vals[:3]
            h_ = vals[3] if len(vals) > 3 else 0
            w_ = vals[4] if len(vals) > 4 else 1 << 30
            m_ = vals[5] if len(vals) > 5 else 3
            ctx.set_tile_policy(d_, s_)
            ctx.set_density_occupancy(h_)
            ctx.set_density_wide(w_)
            ctx.set_density_basis_eval(m_)
            ctx.set_schedule(q_)
            tuned = (d_, s_, q_, None, h_, w_, ctx.density_basis_eval())
        elif not args.no_autotune:
            tic = time.perf_counter()
            tuned = ctx.autotune("PBE", prec)
            pol = ctx.tile_policy()
            tuned = tuned + (pol[2], pol[4], ctx.density_basis_eval())
            log(f"{prec}: autotune -> dens32<=nbe {tuned[0]}, syr2k32<=nbe {tuned[1]}, "
                f"queues {tuned[2]}, density hi-occupancy<=nbe {pol[2]}, "
                f"128-wide>=nbe {pol[4]}, basis evaluation mode {tuned[6]} "
                f"({tuned[3]:.2f} ms; {time.perf_counter() - tic:.1f}s)")
        # Eager steps: per-stage HIP-event times of the kernels.
        ctx.use_graph(False)
        stage = np.zeros(7)
        for _ in range(3):
            ctx.exc_vxc("PBE", prec, True)
            stage += ctx.stage_times()
        stage /= 3
        # Timed steps: graph replays (or eager launches with --no-graph).
        ctx.use_graph(not args.no_graph)
        for _ in range(args.warmup):
            ctx.exc_vxc("PBE", prec, True)
        ctx.synchronize()
        tic = time.perf_counter()
        for _ in range(args.steps):
            ctx.exc_vxc("PBE", prec, True)
        ctx.synchronize()
        toc = time.perf_counter()
        exc, nel, vxc = ctx.results(want_vxc=True)
        ms = (toc - tic) * 1e3 / args.steps
        mode = ctx.density_basis_eval()
        if prec == "fp64" and mode:
            mode = 3  # the FP64 kernels' in-stage mode
        results[prec] = dict(ms=ms, stage=stage, exc=exc, nel=nel, vxc=vxc, tuned=tuned,
                             mode=mode)
        log(f"{prec}: {ms:.2f} ms/step, EXC={exc:.10f} NEL={nel:.6f} "
            f"stages(ms)={np.round(stage, 3).tolist()}")
    ctx.use_graph(False)
    derivs = None
    if args.derivatives:
        derivs = time_derivatives(ctx, cfg, args.steps)
        if not args.no_cpu:
            derivs["cpu"] = cpu_derivatives(bs, P, ts, cfg, args.cpu_seconds / 3)
    learned = None
    if args.learned:
        learned = learned_schedule(ctx, bs, ts)
    head = "mixed" if "mixed" in results else precisions[0]
    r = results[head]
    accuracy = None
    if "mixed" in results and "fp64" in results:
        accuracy = {
            "dEXC_mixed_vs_fp64": abs(results["mixed"]["exc"] - results["fp64"]["exc"]),
            "max_dVXC_mixed_vs_fp64": float(np.abs(results["mixed"]["vxc"] -
                                                   results["fp64"]["vxc"]).max()),
            "fp64_pts_per_s": npts_total / (results["fp64"]["ms"] * 1e-3),
            "fp64_ms_per_step": results["fp64"]["ms"],
            "speedup_mixed_vs_fp64": results["fp64"]["ms"] / results["mixed"]["ms"],
        }
        log(f"accuracy/speed: {accuracy}")

    # Roofline of the dominant stage: algorithmic bytes and flops of its
    # launches over their event-timed duration; bound by arithmetic intensity
    # against the ridge point of the unit it computes on.
    k_dom = int(np.argmax(r["stage"][:6]))
    name = STAGES[k_dom]
    dom_ms = float(r["stage"][k_dom])
    abytes, aflops = algorithmic(name, head, ts, r["mode"])
    peak_mfma = PEAK_FP64_TFLOPS if head == "fp64" else MIXED_PEAK.get(name, PEAK_BF16_TFLOPS)
    ridge = peak_mfma * 1e12 / (PEAK_HBM_GBS * 1e9)
    if aflops > 0 and aflops / abytes > ridge:
        roof = dict(bound="mfma", achieved=aflops / (dom_ms * 1e-3) / 1e12, peak=peak_mfma,
                    unit="TFLOP/s")
    else:
        roof = dict(bound="hbm", achieved=abytes / (dom_ms * 1e-3) / 1e9, peak=PEAK_HBM_GBS,
                    unit="GB/s")
    roof["frac"] = roof["achieved"] / roof["peak"]
    # PMC summaries exist per configuration at its stated size.
    on_pmc_workload = args.n_water is None
    roof["traffic"] = (measured_traffic(args.config, name, head, r["mode"])
                       if on_pmc_workload else None)
    roof["kernel"] = " + ".join(stage_kernels(name, head, r["mode"]) or [name])
    roof["basis_eval_mode"] = r["mode"]
    roof["algorithmic_bytes"] = abytes
    roof["algorithmic_flops"] = aflops
    roof["stage_ms"] = dom_ms

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(bs, P, ts, cfg, args.cpu_seconds)
    value = npts_total / (r["ms"] * 1e-3)
    line = {
        "metric": "PBE EXC+VXC grid pts/s, mixed vs FP64 on 1 GPU, at |dEXC|<=1e-6 Eh",
        "value": value, "unit": "grid points/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "bf16+f32+f64",
        "data": "synthetic",
        "config": {"workload": f"{cfg['n_water']}-water cluster, PBE, 99x590 grid, "
                               f"{cfg['nbf']} bf, {npts_total} points",
                   "precision": head, "nbf": cfg["nbf"], "points": npts_total,
                   "tasks": int(len(ts["npts"])), "graph": not args.no_graph,
                   "autotuned": {p: results[p]["tuned"] for p in results}},
        "roofline": roof,
        "cpu_baseline": cpu,
        "stage_ms": dict(zip(STAGES, [float(x) for x in r["stage"][:6]])),
        "exc": r["exc"],
        "mixed_vs_fp64": accuracy,
        "setup": setup,
    }
    if learned is not None:
        line["learned_schedule"] = learned
    if derivs is not None:
        line["derivatives"] = derivs
    print(json.dumps(line), flush=True)


def cpu_derivatives(bs, P, ts, cfg, budget_s):
    """The oracle's EXC gradient (with weight derivatives), RKS FXC
    contraction and UKS EXC/VXC on bounded samples of the same tasks
    (points/s; its X products through the host BLAS when one loads)."""
    import oracle
    nthreads = cpu_threads()
    blas = oracle.use_blas("auto")
    sh = oracle.shells_from(bs.xcq_shells())
    xyz = np.asarray(cfg["xyz"], dtype=np.float64)
    tP = workloads_density(cfg["nbf"])
    out = {"cores": nthreads, "blas": blas or "none"}
    for key, fn in (("exc_grad_pts_per_s",
                     lambda sub: oracle.exc_grad("PBE", cfg["nbf"], sh, xyz, P, sub,
                                                 nthreads=nthreads)),
                    ("fxc_contraction_pts_per_s",
                     lambda sub: oracle.fxc_contraction("PBE", cfg["nbf"], sh, P, tP, sub,
                                                        nthreads=nthreads)),
                    ("uks_exc_vxc_pts_per_s",
                     lambda sub: oracle.exc_vxc_uks("PBE", cfg["nbf"], sh, P, 0.05 * tP, sub,
                                                    nthreads=nthreads))):
        idx, pts, elapsed = sample_rate(ts, fn, budget_s, nthreads)
        out[key] = pts / elapsed
        out[key.replace("_pts_per_s", "_sample")] = (f"{len(idx)} tasks, {pts} points, "
                                                     f"{elapsed:.2f} s")
    oracle.use_blas(None)
    log(f"cpu derivatives: {out}")
    return out


def workloads_density(nbf):
    """The trial density the derivative timings use."""
    from gauxc_amd import workloads
    return workloads.density(nbf, 8, 7)


def time_derivatives(ctx, cfg, steps):
    """Wall ms per call (device work, results downloaded) of the FP64 EXC
    gradient with and without the grid-weight derivatives, of the RKS FXC
    contraction and of the UKS EXC/VXC, on the uploaded workload (PBE)."""
    xyz = np.asarray(cfg["xyz"], dtype=np.float64)
    out = {}
    for key, wd in (("exc_grad_ms", True), ("exc_grad_hf_ms", False)):
        ctx.exc_grad("PBE", xyz, weight_derivs=wd)
        ctx.results_grad()
        tic = time.perf_counter()
        for _ in range(steps):
            ctx.exc_grad("PBE", xyz, weight_derivs=wd)
            ctx.results_grad()
        out[key] = (time.perf_counter() - tic) * 1e3 / steps
    ctx.upload_trial_density(workloads_density(cfg["nbf"]))
    ctx.fxc_contraction("PBE")
    ctx.results_fxc()
    tic = time.perf_counter()
    for _ in range(steps):
        ctx.fxc_contraction("PBE")
        ctx.results_fxc()
    out["fxc_contraction_ms"] = (time.perf_counter() - tic) * 1e3 / steps
    # UKS EXC/VXC with P_s = P and a small spin density P_z.
    ctx.upload_density_z(0.05 * workloads_density(cfg["nbf"]))
    ctx.exc_vxc_uks("PBE", True)
    ctx.results_uks()
    tic = time.perf_counter()
    for _ in range(steps):
        ctx.exc_vxc_uks("PBE", True)
        ctx.results_uks()
    out["uks_exc_vxc_ms"] = (time.perf_counter() - tic) * 1e3 / steps
    log(f"derivatives (FP64, ms per call): {out}")
    return out


def time_eval(ctx, prec, reps=3):
    """Median wall time (ms) of `reps` synchronised evaluations."""
    ctx.exc_vxc("PBE", prec, True)
    ctx.synchronize()
    ts = []
    for _ in range(reps):
        tic = time.perf_counter()
        ctx.exc_vxc("PBE", prec, True)
        ctx.synchronize()
        ts.append((time.perf_counter() - tic) * 1e3)
    return float(np.median(ts))


def measure_task_times(ctx, prec="mixed", reps=3):
    """Per-task (density, B^T Z) times in microseconds, median of `reps`."""
    ctx.enable_task_timing(True)
    ctx.exc_vxc("PBE", prec, True)
    out = []
    for _ in range(reps):
        ctx.exc_vxc("PBE", prec, True)
        out.append(np.stack(ctx.task_times(), axis=1))
    ctx.enable_task_timing(False)
    return np.median(np.stack(out), axis=0)


def learned_schedule(ctx, bs, ts, prec="mixed"):
    """Learned cost model against GauXC's cost_exc_vxc heuristic on the
    uploaded workload: train on half of the tasks' measured times, report the
    prediction error on the other half, then the wall time of persistent
    queues packed by each model (and of plain one-workgroup-per-item
    dispatch)."""
    from gauxc_amd.costmodel import HeuristicModel, MLPCostModel, task_features
    tic = time.perf_counter()
    times = measure_task_times(ctx, prec)
    feats = task_features(ts, bs.xcq_shells())
    nt = len(times)
    perm = np.random.default_rng(0).permutation(nt)
    tr, te = perm[: nt // 2], perm[nt // 2:]
    sub = lambda idx: {"npts": ts["npts"][idx], "nbe": ts["nbe"][idx]}
    heur = HeuristicModel().fit(sub(tr), times[tr])
    mlp = MLPCostModel(0, seed=1)
    loss = mlp.fit(feats[tr], times[tr], epochs=300, batch=512, lr=1e-3)
    mae_h = np.abs(heur.predict(sub(te)) - times[te]).mean(axis=0)
    mae_m = np.abs(mlp.predict(feats[te]) - times[te]).mean(axis=0)
    nq = 4 * 256
    wall = {"direct": None, "queues_heuristic": None, "queues_mlp": None}
    ctx.set_schedule(0)
    wall["direct"] = time_eval(ctx, prec)
    hp = heur.predict(ts)
    ctx.set_schedule(nq, hp[:, 0], hp[:, 1])
    wall["queues_heuristic"] = time_eval(ctx, prec)
    mp = mlp.predict(feats)
    ctx.set_schedule(nq, mp[:, 0], mp[:, 1])
    wall["queues_mlp"] = time_eval(ctx, prec)
    ctx.set_schedule(0)
    res = {"tasks": int(nt), "train_loss": float(loss),
           "mae_us_density_syr2k": {"heuristic": mae_h.tolist(), "mlp": mae_m.tolist()},
           "mean_task_us": times.mean(axis=0).tolist(), "queues": nq,
           "wall_ms": wall, "seconds": time.perf_counter() - tic}
    log(f"learned schedule: {res}")
    return res


def run_batched(args):
    """64 independent water clusters (2-40 waters, seeded) with the 24-water
    basis/grid recipe, each with its own P, evaluated in one context: every
    task of every molecule in the same launches.  The autotuner and the cost
    model are trained on the first 8 molecules; the full batch then runs as
    persistent queues packed by the learned model, captured in a graph."""
    import gauxc_amd as gx
    from gauxc_amd import workloads
    rng = np.random.default_rng(20)
    nws = rng.integers(2, 41, size=args.n_mol)
    mols = []
    t0 = time.time()
    cfg = dict(workloads.CONFIGS["c1"])
    for m, nw in enumerate(nws):
        box = 10.0 * (nw / 24.0) ** (1.0 / 3.0)  # the 24-water cluster's density
        Z, xyz = workloads.water_cluster(int(nw), box, 100 + m)
        l, pure, alpha, coeff, origin = workloads.even_tempered_basis(Z, xyz, cfg["recipe"])
        mol = gx.Molecule(Z, xyz)
        bs = gx.BasisSet(l, pure, alpha, coeff, origin, normalize=True,
                         shell_tolerance=cfg["tol"])
        P = workloads.density(bs.nbf(), max(1, min(bs.nbf(), 5 * int(nw))), 100 + m)
        lb = workloads.build_tasks(mol, bs, grid=cfg["grid"])
        gx.MolecularWeightsFactory("Device", "Default").get_instance().modify_weights(lb)
        t = lb.get_tasks()
        order = np.argsort(-(t["npts"].astype(np.int64) * t["nbe"].astype(np.int64)),
                           kind="stable")
        mols.append((bs, P, reorder_tasks(t, order)))
        if (m + 1) % 8 == 0:
            log(f"c3 setup: {m + 1} of {len(nws)} molecules ({time.time() - t0:.1f}s)")
    npts_total = int(sum(int(t["npts"].sum()) for _, _, t in mols))
    ntasks = int(sum(len(t["npts"]) for _, _, t in mols))
    log(f"c3 setup: {len(mols)} molecules, waters {int(nws.min())}-{int(nws.max())} "
        f"(sum {int(nws.sum())}), {ntasks} tasks, {npts_total} points "
        f"({time.time() - t0:.1f}s)")

    def context(sel):
        b = gx.BatchedDeviceContext(0)
        b.upload_molecules([mols[i][0].xcq_shells() for i in sel])
        b.upload_tasks([mols[i][2] for i in sel])
        for k, i in enumerate(sel):
            b.upload_density(k, mols[i][1])
        return b

    # Tuning and training on a subset.
    from gauxc_amd.costmodel import MLPCostModel, task_features
    sub = list(range(min(8, len(mols))))
    sc = context(sub)
    tuned = {p: sc.autotune("PBE", p) for p in ("fp64", "mixed")}
    pol = sc.tile_policy()  # the mixed tuning's density settings
    mixed_extra = dict(hiocc=pol[2], wide=pol[4], basis_eval=sc.density_basis_eval())
    log(f"c3 autotune on {len(sub)} molecules: {tuned}, mixed density {mixed_extra}")
    sc.set_tile_policy(*tuned["mixed"][:2])
    times = measure_task_times(sc, "mixed")
    feats_sub = np.concatenate([task_features(mols[i][2], mols[i][0].xcq_shells()) for i in sub])
    mlp = MLPCostModel(0, seed=2)
    mlp.fit(feats_sub, times, epochs=300, batch=512, lr=1e-3)
    sc.close()

    full = context(range(len(mols)))
    feats = np.concatenate([task_features(t, bs.xcq_shells()) for bs, _, t in mols])
    pred = mlp.predict(feats)
    out = {}
    full.set_density_occupancy(mixed_extra["hiocc"])
    full.set_density_wide(mixed_extra["wide"])
    full.set_density_basis_eval(mixed_extra["basis_eval"])
    for prec in ("fp64", "mixed"):
        full.set_tile_policy(*tuned[prec][:2])
        full.use_graph(False)
        full.set_schedule(0)
        direct = time_eval(full, prec, reps=2)
        # Persistent workgroups per CU: the stage kernels keep different
        # numbers resident (more for FP64), so 4 and 8 per CU are both tried.
        qt = {}
        for q in (4 * 256, 8 * 256):
            full.set_schedule(q, pred[:, 0], pred[:, 1])
            qt[q] = time_eval(full, prec, reps=2)
        nq_p = min(qt, key=qt.get)
        full.set_schedule(nq_p, pred[:, 0], pred[:, 1])
        full.use_graph(True)
        for _ in range(args.warmup):
            full.exc_vxc("PBE", prec, True)
        full.synchronize()
        tic = time.perf_counter()
        for _ in range(args.steps):
            full.exc_vxc("PBE", prec, True)
        full.synchronize()
        ms = (time.perf_counter() - tic) * 1e3 / args.steps
        res = [full.results(i, want_vxc=True) for i in range(len(mols))]
        out[prec] = dict(ms=ms, direct_ms=direct, res=res, queues=nq_p)
        log(f"c3 {prec}: {ms:.1f} ms/step (queues {nq_p}, learned costs, graph; eager "
            f"{ {k: round(v, 1) for k, v in qt.items()} }); direct dispatch {direct:.1f} ms")
        full.use_graph(False)
    de = max(abs(a[0] - b[0]) for a, b in zip(out["mixed"]["res"], out["fp64"]["res"]))
    dv = max(float(np.abs(a[2] - b[2]).max()) for a, b in zip(out["mixed"]["res"],
                                                              out["fp64"]["res"]))
    line = {"metric": "PBE EXC+VXC grid pts/s, batched molecules, mixed vs FP64 on 1 GPU",
            "value": npts_total / (out["mixed"]["ms"] * 1e-3), "unit": "grid points/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": out["mixed"]["ms"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "bf16+f32+f64", "data": "synthetic",
            "config": {"workload": f"{len(mols)} water clusters of {int(nws.min())}-"
                                   f"{int(nws.max())} waters, PBE, 99x590 grid, "
                                   f"{ntasks} tasks, {npts_total} points",
                       "molecules": len(mols),
                       "queues": {p: out[p]["queues"] for p in out}, "graph": True,
                       "autotuned_on_subset": {k: list(v) for k, v in tuned.items()}},
            "fp64_ms_per_step": out["fp64"]["ms"],
            "direct_dispatch_ms": {p: out[p]["direct_ms"] for p in out},
            "speedup_mixed_vs_fp64": out["fp64"]["ms"] / out["mixed"]["ms"],
            "max_dEXC_mixed_vs_fp64": de, "max_dVXC_mixed_vs_fp64": dv}
    print(json.dumps(line), flush=True)


def reorder_tasks(t, order):
    out = dict(t)
    for k in ("iparent", "npts", "nshells", "nbe", "dist_nearest"):
        out[k] = t[k][order]
    pts, w, sl = [], [], []
    po, so = [], []
    p = s = 0
    for i in order:
        a, n = t["point_offset"][i], t["npts"][i]
        b, m = t["shell_offset"][i], t["nshells"][i]
        pts.append(t["points"][a:a + n])
        w.append(t["weights"][a:a + n])
        sl.append(t["shell_list"][b:b + m])
        po.append(p)
        so.append(s)
        p += n
        s += m
    out["points"] = np.concatenate(pts)
    out["weights"] = np.concatenate(w)
    out["shell_list"] = np.concatenate(sl)
    out["point_offset"] = np.array(po, dtype=np.int64)
    out["shell_offset"] = np.array(so, dtype=np.int64)
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """Host threads this process may use: the CPU affinity mask, limited by
    OMP_NUM_THREADS when the environment sets it (the share of the host
    that belongs to this GPU)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def sample_rate(ts, fn, budget_s, nthreads):
    """fn on evenly spaced tasks of ts, the sample grown until one run takes
    at least a third of budget_s (the time scales with the sample): (task
    indices, points, seconds) of the last run."""
    ntask = len(ts["npts"])

    def run(count):
        idx = np.unique(np.linspace(0, ntask - 1, max(1, min(count, ntask))).astype(np.int64))
        sub = reorder_tasks(ts, idx)
        tic = time.perf_counter()
        fn(sub)
        return idx, int(sub["npts"].sum()), time.perf_counter() - tic

    run(nthreads)  # first-call costs (library initialisation, thread pool)
    idx, pts, elapsed = run(4 * nthreads)
    while elapsed < budget_s / 3 and len(idx) < ntask:
        grow = min(20.0, max(2.0, budget_s / max(elapsed, 1e-3)))
        idx, pts, elapsed = run(int(len(idx) * grow))
    return idx, pts, elapsed


def cpu_baseline(bs, P, ts, cfg, budget_s):
    """Oracle (restated GauXC host integrator) on a bounded sample of tasks:
    OpenMP over tasks as GauXC's host integrator, the per-task DGEMM and
    DSYR2K through a host BLAS (MKL, OpenBLAS) when one loads."""
    import oracle
    nthreads = cpu_threads()
    blas = oracle.use_blas("auto")
    sh = oracle.shells_from(bs.xcq_shells())
    ntask = len(ts["npts"])
    idx, pts, elapsed = sample_rate(
        ts, lambda sub: oracle.exc_vxc("PBE", cfg["nbf"], sh, P, sub, nthreads=nthreads),
        budget_s, nthreads)
    oracle.use_blas(None)
    visible = os.cpu_count()
    log(f"cpu baseline: {len(idx)} tasks, {pts} points in {elapsed:.2f}s on {nthreads} threads "
        f"({cpu_model()}, {visible} visible CPUs), BLAS {blas or 'none (loops)'}")
    return {"value": pts / elapsed, "unit": "grid points/s", "cores": nthreads, "kind": "port",
            "cpu": cpu_model(), "visible_cpus": visible, "blas": blas or "none",
            "sample": f"{len(idx)} of {ntask} tasks (evenly spaced, {pts} points) of the "
                      f"same workload in {elapsed:.1f} s, oracle restating GauXC's host "
                      f"integrator, OpenMP over tasks"}


if __name__ == "__main__":
    main()
