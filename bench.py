#!/usr/bin/env python3
"""Throughput of the MI355X DPD engine, one GPU. Prints ONE JSON line.

Headline (`value`): the batched engine on the 256-world mixed ensemble
(sides drawn uniformly from 8..32 cells per axis, density 3, each world with
its own seed, a_ij in [25, 80] and 0..20% of beads in H2T4 chains; about 6.0M
beads), launch shape chosen by the engine's autotuner. bead-steps/s =
sum of beads x timed steps / wall time of the timed run, state resident in
HBM (the rule of the reference's benchmark_engine.cpp:128-147, Attach
excluded). The warm-up is rounded up to whole captured step graphs so the
timed steps replay a graph that already exists.

Extra keys:
  roofline      the pair-force kernel of the headline workload: algorithmic
                bytes per launch / its average launch duration, HIP events on
                each context's stream around each kernel in a profiled pass
                (each context alone, right after tuning, as --only
                ensemble-profile runs it); traffic = HBM bytes per launch from
                the committed rocprofv3 --pmc summary of that pass
  batched       the same ensemble for 5000 steps, against the same 256 worlds
                run one at a time through the plain engine for 5000 steps each;
                one_context: the ensemble in a single context (one launch set
                per step for all 256 worlds), autotuned, 5000 steps
  plain_polymer32 / plain_large128
                the plain engine on the polymer world (32^3 cells, 98,304
                beads, 10% bonded) and the large world (128^3 cells, 6.3M
                beads, 5% bonded), with per-kernel times
  served        served sweep: Poisson arrivals (50/s), cubic sides
                log-uniform in [8, 48], 1k..20k steps, interval callbacks every
                500 steps (each request's FP64 observables, reduced on the
                device), an HBM budget that holds only part of the requests;
                p50/p99 latency, admission wait and aggregate bead-steps/s
  served_sustained  the same kind of sweep, 64 requests at 3/s: an offered
                load (about 2.4e9 bead-steps/s) below the device's capacity, so
                latency is service time rather than queueing
  cpu_baseline  the reference's NaiveDPDEngineHalfMerge compiled from its
                sources (-O3 -ffast-math), one thread, bounded sample of the
                polymer world
  cpu_port_mt   the repository's multi-threaded FP64 half-merge restatement on
                every CPU this process may use, CPU model stated
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
KERNELS = ("k_integrate_ms", "k_scan_ms", "k_scatter_ms", "k_reorder_ms", "k_force_ms", "k_stationary_ms")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profiles(name):
    """HBM bytes per pair-force launch from a committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), or None."""
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


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def usable_cpus():
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def cpu_baseline(world, target_s):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O   # the checker's tree, used here only to time the CPU baseline
    out = {}
    model = cpu_model()
    if O.ref_available(fast=True):
        timing = {}
        O.ref_run(world, "half_merge", 1, fast=True, timing=timing)
        per = max(timing["run_s"], 1e-6)
        steps = max(1, min(2000, int(target_s / per)))
        O.ref_run(world, "half_merge", steps, fast=True, timing=timing)
        out["cpu_baseline"] = {
            "value": world.n_beads * steps / timing["run_s"], "unit": "bead-steps/s", "cores": 1,
            "kind": "reference", "cpu": model,
            "sample": f"NaiveDPDEngineHalfMerge built from the reference sources (-O3 -ffast-math), "
                      f"polymer world 32^3 ({world.n_beads} beads), {steps} steps, Run timed alone"}
    nthreads = usable_cpus()
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
            "cpu": model, "cpus_on_host": os.cpu_count(),
            "sample": f"multi-threaded FP64 half-merge restatement (8-colour waves, one thread per usable CPU), "
                      f"polymer world 32^3, {steps} steps"}
    if "cpu_baseline" not in out:
        out["cpu_baseline"] = port
    else:
        out["cpu_port_mt"] = port
    return out


def graph_steps_of(ctx):
    try:
        return max(0, int(ctx.get_config("graph_steps")))
    except Exception:
        return 0


def warm(ctx, warmup):
    """Warm-up steps, rounded up to whole step graphs (even graph length, so
    the buffers end where they started) so the timed run replays a graph that
    is already captured. Returns the steps actually run."""
    g = graph_steps_of(ctx) & ~1
    n = max(warmup, 1)
    if g >= 2:
        n = -(-n // g) * g
    ctx.run(n)
    ctx.sync()
    return n


def timed(ctx, steps):
    t0 = time.perf_counter()
    ctx.run(steps)
    ctx.sync()
    return time.perf_counter() - t0


def kernel_profile(contexts, steps):
    """Per-kernel HIP-event times (ms per step) of chained steps, summed over
    the contexts (each profiled on its own)."""
    acc = {k: 0.0 for k in KERNELS}
    for c in contexts:
        if not c.worlds:
            continue
        c.run_profiled(steps)
        st = c.stats()
        for k in KERNELS:
            acc[k] += st[k]
    return acc


def roofline_of(n_beads, launches, prof, step_ms, traffic):
    """The pair-force kernel: one launch per context per step, each context's
    launches timed alone (kernel_profile). Per launch: mean algorithmic bytes
    (40 B per bead) over mean duration."""
    force_ms = prof["k_force_ms"] / launches
    bytes_per_launch = n_beads * FORCE_BYTES_PER_BEAD / launches
    achieved = bytes_per_launch / (force_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": "k_force", "kernel_ms": force_ms,
            "launches_per_step": launches, "bytes_per_launch": bytes_per_launch,
            "valu_fp32_frac": n_beads * FLOP_PER_BEAD_STEP / (step_ms * 1e-3) / (FP32_PEAK_TFLOPS * 1e12),
            "step_hbm_frac": n_beads * STEP_BYTES_PER_BEAD / (step_ms * 1e-3) / (HBM_PEAK_GBS * 1e9),
            "kernel_ms_per_step": {k[2:-3]: prof[k] for k in KERNELS}}


def run_sizes(capi, W, sides=(8, 12, 16, 20, 24, 32, 48, 64, 96, 128), seconds=0.5):
    """The plain engine on water worlds of side L (density 3, L^3 cells) at its
    default launch shape: per side the step time over a run of about
    `seconds` after a warm-up (force_variant=auto picks the pair-force kernel
    by the number of resident beads)."""
    rows = []
    for L in sides:
        world = W.make_world((L, L, L), seed=L)
        ctx = capi.Context(0)
        ctx.add(world)
        warm(ctx, 40)
        probe = max(20, int(2e4 / max(1.0, world.n_beads / 1e4)))
        el = timed(ctx, probe)
        steps = max(20, int(seconds / (el / probe)))
        el = timed(ctx, steps)
        rows.append({"side": L, "beads": world.n_beads, "steps": steps, "us_per_step": el * 1e6 / steps,
                     "bead_steps_per_s": world.n_beads * steps / el})
        log(f"bench: sizes: {L}^3 ({world.n_beads} beads) {el * 1e6 / steps:.1f} us/step")
        ctx.close()
    return {"metric": "plain-engine step time by world size", "unit": "us/step", "data": "synthetic",
            "config": {"workload": "water worlds L^3 cells, density 3, a=25, gamma=4.5, dt=0.01, default launch shape"},
            "sizes": rows}


def make_ensemble(capi, worlds, parts):
    total = sum(w.n_beads for w in worlds)
    parts = parts if parts else capi.suggested_parts(total)
    ens = capi.Ensemble(0, parts)
    for w in worlds:
        ens.add(w)
    return ens, parts


def solo_leg(capi, worlds, steps, warmup):
    """The worlds one at a time through the plain engine, `steps` each, with
    the plain engine's default launch shape; wall time of the timed runs."""
    total_t = 0.0
    bead_steps = 0
    for w in worlds:
        c = capi.Context(0)
        c.add(w)
        warm(c, warmup)
        total_t += timed(c, steps)
        bead_steps += w.n_beads * steps
        c.close()
    return bead_steps / total_t, total_t


def run_plain(capi, world, steps, warmup, profile_steps, traffic_name):
    ctx = capi.Context(0)
    ctx.add(world)
    w_done = warm(ctx, warmup)
    el = timed(ctx, steps)
    prof = kernel_profile([ctx], profile_steps)
    step_ms = el * 1e3 / steps
    out = {"value": world.n_beads * steps / el, "unit": "bead-steps/s", "beads": world.n_beads, "steps": steps,
           "warmup": w_done, "ms_per_step": step_ms,
           "roofline": roofline_of(world.n_beads, 1, prof, step_ms, traffic_from_profiles(traffic_name))}
    ctx.close()
    return out


def active_parts(ens):
    return sum(1 for c in ens.parts if c.worlds)


def run_served(args, capi, W, n_requests, rate):
    """Served sweep with continuous batching: arrivals join the resident
    requests at interval boundaries while the HBM budget allows."""
    reqs = W.served_requests(n_requests, seed=1, rate=rate)
    srv = capi.Server(device=0, hbm_budget_bytes=int(args.serve_budget_gib * (1 << 30)), interval_steps=500,
                      time_scale=1.0, contexts=args.serve_contexts)
    seen = [0]
    temps = []

    def on_interval(rid, done, obs):
        seen[0] += 1
        temps.append(obs["temperature"])
        return True
    # per-interval statistics from sums reduced on the device (no state copy)
    srv.set_observables_callback(on_interval)
    ids = []
    for r in reqs:
        w = W.make_world(r["L"], bonded_fraction=r["bonded_fraction"], seed=r["seed"], a_matrix=r["a_matrix"])
        ids.append(srv.submit(w, r["arrival"], r["steps"]))
    srv.run()
    st = srv.stats()
    waits = sorted(max(0.0, res["admit_s"] - res["arrival_s"])
                   for res in (srv.result(i, with_state=False) for i in ids))
    srv.close()

    def pct(a, p):
        if not a:
            return 0.0
        pos = p * (len(a) - 1)
        lo = int(pos)
        hi = min(len(a) - 1, lo + 1)
        return a[lo] + (a[hi] - a[lo]) * (pos - lo)
    return {"requests": len(reqs), "completed": st["completed"], "p50_latency_s": st["p50_latency_s"],
            "p99_latency_s": st["p99_latency_s"], "mean_latency_s": st["mean_latency_s"],
            "p50_admission_wait_s": pct(waits, 0.5), "p99_admission_wait_s": pct(waits, 0.99),
            "value": st["bead_steps_per_s"], "unit": "bead-steps/s", "makespan_s": st["makespan_s"],
            "max_resident": st["max_resident"], "max_resident_beads": st["max_resident_beads"],
            "relayouts": st["relayouts"], "relayout_mean_ms": 1e3 * st["relayout_mean_s"],
            "relayout_max_ms": 1e3 * st["relayout_max_s"], "callbacks": seen[0],
            "callback": "observables (FP64 momentum, temperature, pressure per interval, reduced on the device)",
            "mean_interval_temperature": float(sum(temps) / len(temps)) if temps else None, "interval_steps": 500,
            "arrival_rate_per_s": rate,
            "offered_bead_steps_per_s": sum(3 * r["L"][0] ** 3 * r["steps"] for r in reqs) / reqs[-1]["arrival"]
            if reqs else 0.0, "last_arrival_s": reqs[-1]["arrival"] if reqs else 0.0,
            "hbm_budget_gib": args.serve_budget_gib, "contexts": args.serve_contexts}


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
    ap.add_argument("--steps", type=int, default=100, help="timed steps of the headline ensemble")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--profile-steps", type=int, default=40)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch", type=int, default=256, help="worlds in the ensemble")
    ap.add_argument("--batch-steps", type=int, default=5000, help="steps of the batched vs one-at-a-time comparison "
                                                                 "(0 = skip it)")
    ap.add_argument("--batch-solo-sample", type=int, default=0,
                    help="one-at-a-time leg over the first N worlds only (0 = all of them)")
    ap.add_argument("--batch-parts", type=int, default=0,
                    help="contexts (streams) the ensemble is spread over (0 = the engine's suggestion)")
    ap.add_argument("--no-autotune", action="store_true", help="default launch shape instead of the autotuner's")
    ap.add_argument("--no-one-context", action="store_true",
                    help="skip the batched run of the ensemble in a single context")
    ap.add_argument("--plain-steps", type=int, default=1000)
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-engine polymer and large worlds")
    ap.add_argument("--no-large", action="store_true", help="skip the plain-engine large world")
    ap.add_argument("--serve", type=int, default=256, help="served sweep with this many requests (0 = skip)")
    ap.add_argument("--serve-budget-gib", type=float, default=0.5)
    ap.add_argument("--serve-sustained", type=int, default=64,
                    help="a second served sweep at an arrival rate the device keeps up with (0 = skip)")
    ap.add_argument("--serve-sustained-rate", type=float, default=3.0)
    ap.add_argument("--serve-contexts", type=int, default=1, help="contexts the served requests are spread over")
    ap.add_argument("--only", choices=("ensemble", "ensemble-profile", "water16", "polymer32", "large128", "served",
                                       "sizes"),
                    default=None, help="run one workload only (profiling runs); ensemble-profile runs only the "
                                       "headline's profiled pass (no warm-up, no timed steps)")
    ap.add_argument("--graph-steps", type=int, default=None,
                    help="steps per captured hipGraph (0 = direct launches, for per-dispatch counters)")
    ap.add_argument("--config", action="append", default=[], metavar="KEY=VALUE",
                    help="engine setting for every context (pdpd_set_config), applied after autotuning")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of this many MiB, 5 times (HBM counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(independent worlds shard across GPUs with no communication)", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    from pdpd_amd import capi
    from pdpd_amd import world as W
    capi.load()

    if args.only in ("water16", "polymer32", "large128"):
        world = {"water16": W.config_water, "polymer32": W.config_polymer, "large128": W.config_large}[args.only](seed=1)
        ctx = capi.Context(0)
        ctx.add(world)
        if args.graph_steps is not None:
            ctx.set_config("graph_steps", str(args.graph_steps))
        for kv in args.config:
            ctx.set_config(*kv.split("=", 1))
        warm(ctx, args.warmup)
        el = timed(ctx, args.steps)
        log(f"bench: {args.only}: {world.n_beads * args.steps / el:.4e} bead-steps/s")
        ctx.close()
        return

    if args.only == "sizes":
        print(json.dumps(run_sizes(capi, W)), flush=True)
        return

    if args.only == "served":
        print(json.dumps(run_served(args, capi, W, args.serve, 50.0)), flush=True)
        return

    # ---- headline: the 256-world ensemble, batched engine
    worlds = W.config_batch(n_worlds=args.batch, seed=1)
    total = sum(w.n_beads for w in worlds)
    log(f"bench: ensemble of {len(worlds)} worlds, {total} beads")
    ens, parts = make_ensemble(capi, worlds, args.batch_parts)
    if args.graph_steps is not None:
        ens.set_config("graph_steps", str(args.graph_steps))
    tuning = "default launch shape" if args.no_autotune else ens.autotune(10)
    for kv in args.config:
        ens.set_config(*kv.split("=", 1))
        tuning += f" {kv}"
    log(f"bench: tuning: {tuning}")
    if args.only == "ensemble-profile":
        prof = kernel_profile(ens.parts, args.profile_steps)
        n = active_parts(ens)
        log(f"bench: ensemble profiled pass: {n} contexts, k_force {prof['k_force_ms']:.4f} ms per step, "
            f"{prof['k_force_ms'] / n:.4f} ms per launch")
        ens.close()
        return
    if args.only == "ensemble":
        warm(ens, args.warmup)
        el = timed(ens, args.steps)
        log(f"bench: ensemble: {total * args.steps / el:.4e} bead-steps/s")
        ens.close()
        return
    # the profiled pass first, from the freshly tuned state: exactly what
    # --only ensemble-profile runs, whose rocprofv3 trace is committed
    prof = kernel_profile(ens.parts, args.profile_steps)
    w_done = warm(ens, args.warmup)
    el = timed(ens, args.steps)
    value = total * args.steps / el
    step_ms = el * 1e3 / args.steps
    log(f"bench: {value:.4e} bead-steps/s ({step_ms:.4f} ms/step)")
    roofline = roofline_of(total, active_parts(ens), prof, step_ms, traffic_from_profiles("pair_force_ensemble256_pmc.md"))

    out = {"metric": METRIC, "value": value, "unit": "bead-steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": w_done, "ms_per_step": step_ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"batched engine, ensemble of {len(worlds)} worlds (sides uniform in 8..32 cells "
                                  f"per axis, density 3, own seed and a_ij in [25,80], 0-20% of beads in H2T4 chains)",
                      "worlds": len(worlds), "beads": total, "dt": 0.01, "contexts": parts, "tuning": tuning},
           "roofline": roofline}

    if args.batch_steps:
        log(f"bench: batched vs one at a time, {args.batch_steps} steps")
        tb = timed(ens, args.batch_steps)
        batched = total * args.batch_steps / tb
        sample = worlds[:args.batch_solo_sample] if args.batch_solo_sample else worlds
        solo, solo_t = solo_leg(capi, sample, args.batch_steps, args.warmup)
        out["batched"] = {"worlds": len(worlds), "beads": total, "steps": args.batch_steps, "value": batched,
                          "unit": "bead-steps/s", "seconds": tb, "one_at_a_time": solo,
                          "one_at_a_time_seconds": solo_t, "one_at_a_time_sample": len(sample),
                          "speedup": batched / solo, "contexts": parts, "tuning": tuning}
        log(f"bench: batched {batched:.4e}, one at a time {solo:.4e} -> {batched / solo:.2f}x")
    ens.close()
    if args.batch_steps and not args.no_one_context:
        # the same ensemble in one context (one launch set per step for all 256 worlds)
        one = capi.Ensemble(0, 1)
        for w in worlds:
            one.add(w)
        tuning1 = "default launch shape" if args.no_autotune else one.autotune(10)
        warm(one, args.warmup)
        t1 = timed(one, args.batch_steps)
        v1 = total * args.batch_steps / t1
        out["batched"]["one_context"] = {"value": v1, "seconds": t1, "speedup": v1 / out["batched"]["one_at_a_time"],
                                         "tuning": tuning1}
        log(f"bench: batched in one context {v1:.4e} -> {v1 / out['batched']['one_at_a_time']:.2f}x")
        one.close()

    if not args.no_plain:
        log("bench: plain engine, polymer world 32^3")
        out["plain_polymer32"] = run_plain(capi, W.config_polymer(seed=1), args.plain_steps, args.warmup,
                                           args.profile_steps, "pair_force_polymer32_pmc.md")
        if not args.no_large:
            log("bench: plain engine, large world 128^3")
            out["plain_large128"] = run_plain(capi, W.config_large(seed=1), max(20, args.plain_steps // 10),
                                              args.warmup, max(4, args.profile_steps // 4),
                                              "pair_force_large128_pmc.md")
    if args.serve:
        log(f"bench: served sweep of {args.serve} requests")
        out["served"] = run_served(args, capi, W, args.serve, 50.0)
    if args.serve_sustained:
        log(f"bench: served sweep at {args.serve_sustained_rate}/s, {args.serve_sustained} requests")
        out["served_sustained"] = run_served(args, capi, W, args.serve_sustained, args.serve_sustained_rate)
    if not args.no_cpu_baseline:
        log("bench: CPU baseline")
        out.update(cpu_baseline(W.config_polymer(seed=1), args.cpu_seconds))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
