#!/usr/bin/env python3
"""Throughput of the MI355X DPD engine on the polymer world (32x32x32 cells,
density 3, 98,304 beads, 10% of beads in H2T4 chains), one GPU.

bead-steps/s = beads x steps / wall time of the timed run, state resident in
HBM (the reference's benchmark_engine.cpp:128-147 rule with Attach excluded).
Prints ONE JSON line. Extra keys:
  roofline      pair-force kernel: algorithmic bytes / its average launch duration
                (HIP events on the engine's stream around back-to-back launches
                of the last timed step's pair-force kernel)
  cpu_baseline  the reference's NaiveDPDEngineHalfMerge compiled from its
                sources (-O3 -ffast-math), one thread, bounded sample
  cpu_port_mt   the repository's multi-threaded half-merge restatement
  batched       256-world mixed-size ensemble (8..32 cells per side), one
                launch set per step, vs a sample of the same worlds one at a time
  served        served sweep: Poisson arrivals, continuous batching, p50/p99
                latency and aggregate bead-steps/s
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "poets-dpd_amd"))

METRIC = "bead-steps/s on 1 MI355X, batched vs plain engine; p50/p99 request latency"
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, AMD public specification
FP32_PEAK_TFLOPS = 157.3
FORCE_BYTES_PER_BEAD = 40      # read x, v, key (28 B) + write f (12 B)
STEP_BYTES_PER_BEAD = 80       # algorithmic bytes of a whole step per bead
FLOP_PER_BEAD_STEP = 282       # 6.283 pairs x 41 + 24


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profiles(name="pair_force_polymer32_pmc.md"):
    """HBM bytes per pair-force launch from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration),
    or None when no such summary is present."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


def cpu_baseline(world, target_s):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O   # the checker's tree, used here only to time the CPU baseline
    out = {}
    if O.ref_available(fast=True):
        timing = {}
        steps = 1
        O.ref_run(world, "half_merge", steps, fast=True, timing=timing)
        per = max(timing["run_s"], 1e-6)
        steps = max(1, min(2000, int(target_s / per)))
        O.ref_run(world, "half_merge", steps, fast=True, timing=timing)
        out["cpu_baseline"] = {
            "value": world.n_beads * steps / timing["run_s"], "unit": "bead-steps/s", "cores": 1,
            "kind": "reference",
            "sample": f"NaiveDPDEngineHalfMerge built from the reference sources (-O3 -ffast-math), "
                      f"same polymer world, {steps} steps, Run timed alone"}
    nthreads = max(1, min(16, os.cpu_count() or 1))
    st = O.F64State(world)
    t0 = time.perf_counter()
    st.run_half_merge(1, nthreads=nthreads)
    per = max(time.perf_counter() - t0, 1e-6)
    steps = max(1, min(2000, int(target_s / per)))
    st = O.F64State(world)
    t0 = time.perf_counter()
    st.run_half_merge(steps, nthreads=nthreads)
    dt = time.perf_counter() - t0
    port = {"value": world.n_beads * steps / dt, "unit": "bead-steps/s", "cores": nthreads, "kind": "port",
            "sample": f"multi-threaded FP64 half-merge restatement, same world, {steps} steps"}
    if "cpu_baseline" not in out:
        out["cpu_baseline"] = port
    else:
        out["cpu_port_mt"] = port
    return out


def run_batched(args, capi, W):
    """Mixed-size ensemble: all worlds advance together (one launch set per step
    in each of the ensemble's contexts, whose streams overlap), vs the same
    worlds one at a time through the plain engine."""
    worlds = W.config_batch(n_worlds=args.batch, seed=1)
    total = sum(w.n_beads for w in worlds)
    steps = args.batch_steps
    parts = args.batch_parts if args.batch_parts else capi.suggested_parts(total)
    ctx = capi.Ensemble(0, parts)
    for w in worlds:
        ctx.add(w)
    tune = ctx.autotune(10) if not args.no_batch_autotune else "heuristic"
    ctx.run(2)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.run(steps)
    ctx.sync()
    tb = time.perf_counter() - t0
    batched = total * steps / tb
    ctx.close()
    # one at a time, a sample of the worlds, same steps
    k = min(len(worlds), args.batch_solo_sample)
    solo_time = 0.0
    solo_beads = 0
    for w in worlds[:k]:
        c = capi.Context(0)
        c.add(w)
        c.run(2)
        c.sync()
        t0 = time.perf_counter()
        c.run(steps)
        c.sync()
        solo_time += time.perf_counter() - t0
        solo_beads += w.n_beads
        c.close()
    solo = solo_beads * steps / solo_time
    return {"worlds": len(worlds), "beads": total, "steps": steps, "value": batched, "unit": "bead-steps/s",
            "one_at_a_time": solo, "one_at_a_time_sample": k, "speedup": batched / solo, "contexts": parts,
            "tuning": tune}


def run_served(args, capi, W):
    """Served sweep: Poisson arrivals (50/s), cubic sides log-uniform in
    [8, 48] cells, 1k-20k steps, per-request a_ij and bonded fraction, interval
    callbacks every 500 steps, admission by an HBM budget. Continuous batching:
    arrivals join the running launches at interval boundaries."""
    reqs = W.served_requests(args.serve, seed=1)
    srv = capi.Server(device=0, hbm_budget_bytes=int(args.serve_budget_gib * (1 << 30)), interval_steps=500,
                      time_scale=1.0, contexts=args.serve_contexts)
    seen = [0]

    def on_interval(rid, done, t, x, v, f):
        seen[0] += 1
        return True
    srv.set_callback(on_interval, copy=False)
    for r in reqs:
        w = W.make_world(r["L"], bonded_fraction=r["bonded_fraction"], seed=r["seed"], a_matrix=r["a_matrix"])
        srv.submit(w, r["arrival"], r["steps"])
    srv.run()
    st = srv.stats()
    srv.close()
    return {"requests": len(reqs), "completed": st["completed"], "p50_latency_s": st["p50_latency_s"],
            "p99_latency_s": st["p99_latency_s"], "mean_latency_s": st["mean_latency_s"],
            "value": st["bead_steps_per_s"], "unit": "bead-steps/s", "makespan_s": st["makespan_s"],
            "max_resident": st["max_resident"], "callbacks": seen[0], "interval_steps": 500,
            "arrival_rate_per_s": 50.0, "hbm_budget_gib": args.serve_budget_gib, "contexts": args.serve_contexts}


def copy_calibration(mib):
    """A plain device-to-device copy of known size: its FETCH_SIZE / WRITE_SIZE
    counters give the scale from counter KiB to bytes moved."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 5}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--profile-steps", type=int, default=200)
    ap.add_argument("--force-reps", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch", type=int, default=256, help="batched ensemble of this many worlds (0 = skip)")
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--batch-solo-sample", type=int, default=16)
    ap.add_argument("--batch-parts", type=int, default=0,
                    help="contexts (streams) the ensemble is spread over (0 = the engine's suggestion)")
    ap.add_argument("--autotune", action="store_true", help="autotune the plain engine's launch shape too")
    ap.add_argument("--no-batch-autotune", action="store_true",
                    help="batched ensemble with the default launch shape instead of the autotuner's")
    ap.add_argument("--serve", type=int, default=64, help="served sweep with this many requests (0 = skip)")
    ap.add_argument("--serve-budget-gib", type=float, default=32.0)
    ap.add_argument("--serve-contexts", type=int, default=1, help="contexts the served requests are spread over")
    ap.add_argument("--side", type=int, default=32)
    ap.add_argument("--force-band-rows", type=int, default=None,
                    help="force-block band height for worlds with large y-z planes (0 = layout order)")
    ap.add_argument("--large", action="store_true",
                    help="plain engine on the large world (128^3 cells, 6.3M beads, 5%% bonded) instead")
    ap.add_argument("--graph-steps", type=int, default=None,
                    help="steps per captured hipGraph (0 = direct launches, for per-dispatch counters)")
    ap.add_argument("--fused-integrate", type=int, default=None, choices=(0, 1, 2),
                    help="chained steps: 0 off, 1 on, 2 automatic (the engine's default)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of this many MiB, 5 times (HBM counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(independent worlds shard across GPUs with no communication; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    from pdpd_amd import capi
    from pdpd_amd import world as W
    capi.load()

    if args.large:
        args.side = 128
        world = W.config_large(seed=1, L=args.side)
    else:
        world = W.config_polymer(seed=1, L=args.side)
    n = world.n_beads
    bonded_pct = 5 if args.large else 10
    log(f"bench: polymer world {args.side}^3, {n} beads; warmup {args.warmup}, timed {args.steps}")
    ctx = capi.Context(0)
    slot = ctx.add(world)
    if args.graph_steps is not None:
        ctx.set_config("graph_steps", str(args.graph_steps))
    if args.fused_integrate is not None:
        ctx.set_config("fused_integrate", str(args.fused_integrate))
    if args.force_band_rows is not None:
        ctx.set_config("force_band_rows", str(args.force_band_rows))
    if args.autotune:
        log("bench: autotune " + ctx.autotune(20))
    ctx.run(args.warmup)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.run(args.steps)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    value = n * args.steps / elapsed
    log(f"bench: {value:.4e} bead-steps/s ({elapsed * 1e3 / args.steps:.4f} ms/step)")

    # the dominant kernel's average launch duration: HIP events on the engine's
    # stream around back-to-back launches of the last step's pair-force kernel
    force_ms = ctx.time_force(args.force_reps)
    ctx.run_profiled(args.profile_steps)
    st = ctx.stats()
    step_ms = elapsed * 1e3 / args.steps          # whole step, timed region
    achieved = n * FORCE_BYTES_PER_BEAD / (force_ms * 1e-3) / 1e9
    traffic = traffic_from_profiles("pair_force_large128_pmc.md" if args.large else "pair_force_polymer32_pmc.md")
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "kernel": "k_force", "kernel_ms": force_ms,
                "valu_fp32_frac": n * FLOP_PER_BEAD_STEP / (step_ms * 1e-3) / (FP32_PEAK_TFLOPS * 1e12),
                "step_hbm_frac": n * STEP_BYTES_PER_BEAD / (step_ms * 1e-3) / (HBM_PEAK_GBS * 1e9),
                "kernel_ms_per_step": {k[2:-3]: st[k] for k in ("k_integrate_ms", "k_scan_ms", "k_scatter_ms",
                                                               "k_reorder_ms", "k_force_ms", "k_stationary_ms")}}
    x, v, f, t = ctx.download(slot)
    fused = ctx.get_config("fused_integrate")
    ctx.close()

    out = {"metric": METRIC, "value": value, "unit": "bead-steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": elapsed * 1e3 / args.steps, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"plain engine, polymer world {args.side}x{args.side}x{args.side} cells, "
                                  f"density 3, {n} beads, {bonded_pct}% in H2T4 chains (Hookean k=128 r0=0.5, angle k=20)",
                      "beads": n, "dt": 0.01, "engine": f"HipDPDEngine (graph-captured steps, fused_integrate={fused})"},
           "roofline": roofline}
    if args.batch:
        log(f"bench: batched ensemble of {args.batch} worlds")
        out["batched"] = run_batched(args, capi, W)
    if args.serve:
        log(f"bench: served sweep of {args.serve} requests")
        out["served"] = run_served(args, capi, W)
    if not args.no_cpu_baseline:
        log("bench: CPU baseline")
        out.update(cpu_baseline(world, args.cpu_seconds))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
