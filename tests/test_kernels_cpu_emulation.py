"""The engine's kernels and host code built for the CPU against an emulation of
the HIP surface they use (tests/cpu_emulation): work-items run as fibers, one
workgroup at a time; __syncthreads and wave intrinsics are barriers among the
live work-items, and a wave intrinsic reached by only part of a wave aborts.

  * the AddressSanitizer + UBSan build runs worlds (water, chains, stationary
    wall with theta0 != 0, a batch with a frozen world, graph replay, a
    persistent grid, every pair-force variant) through the C ABI: no
    out-of-bounds access, no undefined behaviour, no divergent wave intrinsic;
  * the plain build's first step equals the FP32 oracle bit for bit (the same
    source as the gfx950 build, FP32 with -ffp-contract=off), including a dense
    world whose accepted pairs overflow the pair kernel's per-bead lists and
    per-wave queues many times (and whose bricks exceed the LDS capacity of
    the brick kernel), every brick shape, the per-slot kernel, a world whose
    strengths vary per cell, and that world batched with a constant-strength
    one.
This checks the kernel source on the CPU; GPU parity is tests/test_gpu_*.py."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU = os.path.join(ROOT, "tests", "cpu_emulation")
BUILD = os.path.join(EMU, "_build")


@pytest.fixture(scope="module")
def emu_build():
    r = subprocess.run(["make", "-s", "-j4", "-C", EMU], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stderr[-3000:]
    return BUILD


def test_kernels_bounds_checked_under_asan(emu_build):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([os.path.join(emu_build, "driver")], capture_output=True, text=True, timeout=900, env=env)
    err = "\n".join(l for l in r.stderr.splitlines() if "makecontext/swapcontext" not in l)
    assert r.returncode == 0, err[-4000:]
    assert "all cases done" in err
    assert "runtime error" not in err and "ERROR: AddressSanitizer" not in err, err[-4000:]


PROBE = r"""
import sys, numpy as np
sys.path[:0] = [%(pkg)r, %(oracle)r]
from pdpd_amd import capi, world as W
import oracle as O
capi.LIB_PATH = %(lib)r
rngA = np.random.default_rng(7)
worlds = {
    "water4": W.make_world((4, 4, 4), seed=11),
    "poly_6x7x8": W.make_world((6, 7, 8), bonded_fraction=0.2, seed=12, a_matrix=W.random_a_matrix(rngA)),
    "wall_bent6": W.add_wall(W.make_world((6, 6, 6), bonded_fraction=0.2, seed=15, theta0=1.0,
                                          a_matrix=W.random_a_matrix(rngA))),
    "lattice4": W.make_lattice_world(),
    "dense3": W.make_world((3, 4, 3), density=24.0, seed=16),    # pair lists and queues overflow
}
sp = W.make_world((6, 5, 7), bonded_fraction=0.2, seed=17)
sp.set_expression("conservative", 0, 0, "25+50*ux")
sp.set_expression("conservative", 0, 1, "30*uy+x/4")
sp.set_expression("dissipative", 1, 1, "3+3*uz")
sp.set_expression("dissipative", 0, 2, "4.5*(1+z/10)")
worlds["spatial"] = sp                                          # strengths per cell
for (name, w), var in zip(worlds.items(), (4, 1, 5, 3, 6, 2)):
    ctx = capi.Context(0)
    ctx.set_config("force_variant", var)
    ctx.set_config("scan_lookback_tiles", (var & 1) * 64)     # both cell-scan forms
    slot = ctx.add(w)
    ctx.run(1)
    x, ww, f, t = ctx.download_f32(slot)
    ref = O.F32State(w).step(1)
    same = (np.array_equal(x.view(np.uint32), ref.x.view(np.uint32)) and
            np.array_equal(ww.view(np.uint32), ref.w.view(np.uint32)) and
            np.array_equal(f.view(np.uint32), ref.f.view(np.uint32)))
    print(name, "bitwise" if same else "DIFFERENT", flush=True)
    ctx.close()
# the row-wave variant (one wave per stencil row; the dense world takes
# several rounds per row)
for var in (7,):
    same = True
    for name in ("poly_6x7x8", "wall_bent6", "dense3", "lattice4"):
        ctx = capi.Context(0)
        ctx.set_config("force_variant", var)
        slot = ctx.add(worlds[name])
        ctx.run(2)
        x, ww, f, t = ctx.download_f32(slot)
        ref = O.F32State(worlds[name]).step(2)
        same &= (np.array_equal(x.view(np.uint32), ref.x.view(np.uint32)) and
                 np.array_equal(ww.view(np.uint32), ref.w.view(np.uint32)) and
                 np.array_equal(f.view(np.uint32), ref.f.view(np.uint32)))
        ctx.close()
    print(f"variant{var}", "bitwise" if same else "DIFFERENT", flush=True)
# every brick shape of the LDS brick kernel (the dense world takes the path for
# bricks beyond the LDS capacity), and the per-slot kernel
for shape in range(int(capi.Context(0).get_config("force_brick_shapes"))):
    same = True
    for name in ("poly_6x7x8", "wall_bent6", "dense3", "lattice4"):
        ctx = capi.Context(0)
        ctx.set_config("force_kernel", 1)
        ctx.set_config("force_brick", shape)
        slot = ctx.add(worlds[name])
        ctx.run(2)
        x, ww, f, t = ctx.download_f32(slot)
        ref = O.F32State(worlds[name]).step(2)
        same &= (np.array_equal(x.view(np.uint32), ref.x.view(np.uint32)) and
                 np.array_equal(ww.view(np.uint32), ref.w.view(np.uint32)) and
                 np.array_equal(f.view(np.uint32), ref.f.view(np.uint32)))
        ctx.close()
    print(f"brick{shape}", "bitwise" if same else "DIFFERENT", flush=True)
ctx = capi.Context(0)
ctx.set_config("force_kernel", 0)
slot = ctx.add(worlds["poly_6x7x8"])
ctx.run(2)
x, ww, f, t = ctx.download_f32(slot)
ref = O.F32State(worlds["poly_6x7x8"]).step(2)
print("per_slot", "bitwise" if np.array_equal(f.view(np.uint32), ref.f.view(np.uint32)) else "DIFFERENT", flush=True)
ctx.close()
# one context with a spatial and a constant world: each world's own strengths
ctx = capi.Context(0)
slots = [ctx.add(worlds["spatial"]), ctx.add(worlds["water4"])]
ctx.run(2)
same = True
for slot, w in zip(slots, (worlds["spatial"], worlds["water4"])):
    x, ww, f, t = ctx.download_f32(slot)
    ref = O.F32State(w).step(2)
    same &= (np.array_equal(x.view(np.uint32), ref.x.view(np.uint32)) and
             np.array_equal(f.view(np.uint32), ref.f.view(np.uint32)))
print("mixed", "bitwise" if same else "DIFFERENT", flush=True)
ctx.close()
# chains of steps (each step's pair-force kernel integrating the next): a
# stationary wall whose budget ends mid-chain beside a world that runs on,
# direct launches and a replayed graph
for gs in (0, 4):
    ctx = capi.Context(0)
    ctx.set_config("graph_steps", gs)
    a = ctx.add(worlds["wall_bent6"]); b = ctx.add(worlds["poly_6x7x8"])
    ctx.set_steps(a, 7)
    ctx.run(11)                 # a graph of 4 (two chains of 2), a chain of 3, then 4 single steps
    same = True
    for slot, w, n in ((a, worlds["wall_bent6"], 7), (b, worlds["poly_6x7x8"], 11)):
        x, ww, f, t = ctx.download_f32(slot)
        ref = O.F32State(w).step(n)
        same &= (t == n and np.array_equal(x.view(np.uint32), ref.x.view(np.uint32)) and
                 np.array_equal(ww.view(np.uint32), ref.w.view(np.uint32)) and
                 np.array_equal(f.view(np.uint32), ref.f.view(np.uint32)))
    print(f"chain_graph{gs}", "bitwise" if same else "DIFFERENT", flush=True)
    ctx.close()
"""


def test_emulated_step_equals_oracle_bitwise(emu_build):
    code = PROBE % dict(pkg=os.path.join(ROOT, "poets-dpd_amd"), oracle=os.path.join(ROOT, "oracle"),
                        lib=os.path.join(emu_build, "libpdpd_emu.so"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stderr[-3000:]
    lines = r.stdout.split("\n")
    shapes = [l for l in lines if l.startswith("brick")]
    rows = [l for l in lines if l.startswith("variant")]
    assert len(shapes) == 10 and "DIFFERENT" not in r.stdout, r.stdout     # 6 bricks, 4 slabs
    assert len(rows) == 1, r.stdout
    assert sum(l.endswith("bitwise") for l in lines) == 10 + len(shapes) + len(rows), r.stdout


SERVE_PROBE = r"""
import sys, numpy as np
sys.path[:0] = [%(pkg)r]
from pdpd_amd import capi, world as W
capi.LIB_PATH = %(lib)r
rng = np.random.default_rng(5)
reqs = [(W.make_world((4, 4, 5), bonded_fraction=0.2, seed=600 + i), int(s)) for i, s in enumerate((23, 40, 17, 31))]
def solo(w, steps):
    c = capi.Context(0); s = c.add(w); c.run(steps); out = c.download(s); c.close(); return out
budget = sum(112 * w.n_beads + 8 * int(np.prod(w.box)) + 4096 for w, _ in reqs) // 2
srv = capi.Server(device=0, hbm_budget_bytes=budget, interval_steps=10, time_scale=0.0, contexts=2)
seen = {}
def cb(rid, done, t, x, v, f):
    seen.setdefault(rid, []).append((done, t))
    return not (rid == ids[3] and done >= 20)            # the last request is cancelled after 20 steps
srv.set_callback(cb)
ids = [srv.submit(w, 0.0, steps) for w, steps in reqs]
srv.run()
st = srv.stats()
ok = st["completed"] == 4 and 1 <= st["max_resident"] < 4
for rid, (w, steps) in zip(ids[:3], reqs[:3]):
    r = srv.result(rid)
    x, v, f, t = solo(w, steps)
    ok &= r["status"] == 0 and np.array_equal(r["x"], x) and np.array_equal(r["v"], v) and np.array_equal(r["f"], f)
    dones = [d for d, _ in seen[rid]]
    ok &= dones == sorted(set(dones)) and dones[-1] == steps and all(d == t for d, t in seen[rid])
ok &= srv.result(ids[3])["status"] == 3 and [d for d, _ in seen[ids[3]]] == [10, 20]
print("serving", "ok" if ok else "WRONG", st, flush=True)
srv.close()
# an ensemble over three contexts: each world as if run alone, budgets too
ens = capi.Ensemble(0, parts=3)
es = [ens.add(w) for w, _ in reqs]
ens.set_steps(es[1], 12)
ens.run(20)
ens.sync()
ok = True
for k, (w, _) in enumerate(reqs):
    x, v, f, t = solo(w, 12 if k == 1 else 20)
    gx, gv, gf, gt = ens.download(es[k])
    ok &= gt == t and np.array_equal(gx, x) and np.array_equal(gv, v) and np.array_equal(gf, f)
print("ensemble", "ok" if ok else "WRONG", flush=True)
ens.close()
# a request whose world fails (one bead thrown out of the box) is retired alone;
# the other two finish bit-identical to their solo runs
bad = W.make_world((4, 5, 4), bonded_fraction=0.2, seed=700)
bad.v = bad.v.copy()
bad.v[3] = [1e30, 0.0, 0.0]
good = [W.make_world((4, 4, 5), bonded_fraction=0.2, seed=701), W.make_world((5, 4, 4), seed=702)]
srv = capi.Server(device=0, interval_steps=10, time_scale=0.0)
srv.set_callback(lambda rid, done, t, x, v, f: True)
ids = [srv.submit(good[0], 0.0, 30), srv.submit(bad, 0.0, 30), srv.submit(good[1], 0.0, 30)]
srv.run()
st = srv.stats()
rb = srv.result(ids[1], with_state=False)
ok = st["failed"] == 1 and st["completed"] == 2 and rb["status"] == 4 and (rb["fail_flags"] & 1) == 1
for rid, w in ((ids[0], good[0]), (ids[2], good[1])):
    r = srv.result(rid)
    x, v, f, t = solo(w, 30)
    ok &= r["status"] == 0 and np.array_equal(r["x"], x) and np.array_equal(r["v"], v) and np.array_equal(r["f"], f)
print("failed_request", "ok" if ok else "WRONG", st, rb, flush=True)
srv.close()
# observables per interval from the device (no state copy) equal a solo run's
w = W.make_world((4, 5, 4), bonded_fraction=0.2, seed=703)
srv = capi.Server(device=0, interval_steps=10, time_scale=0.0)
obs = []
srv.set_observables_callback(lambda rid, done, o: obs.append((done, o)) or True)
rid = srv.submit(w, 0.0, 20)
srv.run()
srv.close()
c = capi.Context(0); s = c.add(w)
ok = [d for d, _ in obs] == [10, 20]
for done, o in obs:
    c.run(10)
    want = c.observables(s)
    ok &= all(o[k] == want[k] for k in ("t", "kinetic", "virial", "n_pairs", "momentum"))
c.close()
print("observables_callback", "ok" if ok else "WRONG", flush=True)
"""


def test_emulated_serving_pipeline(emu_build):
    """The serving loop with one interval in flight (snapshot, next interval,
    then downloads and callbacks) over two contexts: queueing under a memory
    budget, a callback per interval, a cancellation, and results equal to solo
    runs bit for bit; an ensemble of worlds over three contexts likewise; a
    request whose world fails (a bead thrown out of the box) retired alone,
    with its flags, while the others finish bit-identical to their solo runs."""
    code = SERVE_PROBE % dict(pkg=os.path.join(ROOT, "poets-dpd_amd"), lib=os.path.join(emu_build, "libpdpd_emu.so"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "serving ok" in r.stdout and "ensemble ok" in r.stdout, r.stdout
    assert "failed_request ok" in r.stdout, r.stdout
    assert "observables_callback ok" in r.stdout, r.stdout
