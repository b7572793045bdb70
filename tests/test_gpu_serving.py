"""Continuous-batching serving loop (pdpd_server_* in the C ABI).

Each served result must equal the same world run alone through the plain
engine for the same number of steps, bit for bit: requests share launches but
the per-bead arithmetic and its order do not depend on the batch.
"""
import numpy as np
import pytest

from pdpd_amd import capi
from pdpd_amd import world as W

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _requests(n, seed=3):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        L = tuple(int(v) for v in rng.integers(4, 11, 3))
        w = W.make_world(L, bonded_fraction=float(rng.uniform(0, 0.2)), seed=500 + i, a_matrix=W.random_a_matrix(rng))
        reqs.append((w, int(rng.integers(20, 140)), float(i) * 0.0))
    return reqs


def _solo(world, steps):
    ctx = capi.Context(0)
    slot = ctx.add(world)
    ctx.run(steps)
    ctx.sync()
    out = ctx.download(slot)
    ctx.close()
    return out


@pytest.mark.parametrize("contexts", [1, 3])
def test_served_results_equal_solo_runs(gpu_lib, contexts):
    reqs = _requests(10)
    # a budget that holds only part of the requests at once forces queueing
    budget = sum(112 * w.n_beads + 8 * int(np.prod(w.box)) + 4096 for w, _, _ in reqs) // 3
    srv = capi.Server(device=0, hbm_budget_bytes=budget, interval_steps=25, time_scale=0.0, contexts=contexts)
    ids = [srv.submit(w, arr, steps) for w, steps, arr in reqs]
    print(f"serving {len(ids)} requests, budget {budget} B", flush=True)
    srv.run()
    st = srv.stats()
    assert st["completed"] == len(reqs)
    assert 1 <= st["max_resident"] < len(reqs)        # the budget limited admission
    assert st["p99_latency_s"] >= st["p50_latency_s"] > 0
    for rid, (w, steps, _) in zip(ids, reqs):
        r = srv.result(rid)
        assert r["status"] == 0 and r["steps"] == steps
        x, v, f, t = _solo(w, steps)
        assert t == steps
        assert np.array_equal(r["x"], x) and np.array_equal(r["v"], v) and np.array_equal(r["f"], f)
    srv.close()


def test_interval_callback_sees_every_interval(gpu_lib):
    reqs = _requests(3, seed=9)
    srv = capi.Server(device=0, interval_steps=30, time_scale=0.0)
    seen = {}

    def cb(rid, done, t, x, v, f):
        seen.setdefault(rid, []).append((done, t))
        assert np.all(x >= 0)
        return True

    srv.set_callback(cb)
    ids = [srv.submit(w, 0.0, steps) for w, steps, _ in reqs]
    srv.run()
    for rid, (w, steps, _) in zip(ids, reqs):
        dones = [d for d, _ in seen[rid]]
        assert dones == sorted(dones) and dones[-1] == steps
        assert all(d == t for d, t in seen[rid])      # worlds start at t = 0
    srv.close()


def test_callback_cancels_request(gpu_lib):
    w = W.make_world((6, 6, 6), seed=77)
    srv = capi.Server(device=0, interval_steps=10, time_scale=0.0)
    srv.set_callback(lambda rid, done, t, x, v, f: done < 20)
    rid = srv.submit(w, 0.0, 100)
    srv.run()
    r = srv.result(rid)
    assert r["status"] == 3
    srv.close()


def test_failing_request_retired_alone(gpu_lib):
    """A request whose world fails (one bead thrown out of the box: the
    reference's engines throw on such a state, dpd_engine_avx2_half_merge_tbb.hpp:800-820)
    is retired with its flags; the requests sharing its launches finish
    bit-identical to their solo runs."""
    bad = W.make_world((6, 5, 7), bonded_fraction=0.2, seed=700)
    bad.v = bad.v.copy()
    bad.v[3] = [1e30, 0.0, 0.0]
    good = [W.make_world((7, 6, 5), bonded_fraction=0.2, seed=701), W.make_world((5, 6, 6), seed=702)]
    srv = capi.Server(device=0, interval_steps=20, time_scale=0.0)
    srv.set_callback(lambda rid, done, t, x, v, f: True)
    ids = [srv.submit(good[0], 0.0, 60), srv.submit(bad, 0.0, 60), srv.submit(good[1], 0.0, 60)]
    srv.run()
    st = srv.stats()
    rb = srv.result(ids[1], with_state=False)
    assert st["failed"] == 1 and st["completed"] == 2
    assert rb["status"] == 4 and rb["fail_flags"] & 0x1
    for rid, w in ((ids[0], good[0]), (ids[2], good[1])):
        r = srv.result(rid)
        x, v, f, t = _solo(w, 60)
        assert r["status"] == 0
        assert np.array_equal(r["x"], x) and np.array_equal(r["v"], v) and np.array_equal(r["f"], f)
    srv.close()


def test_world_flags_name_the_failing_world(gpu_lib):
    """pdpd_sync fails naming the world that raised a flag; a snapped bond (longer
    than half the box, naive_dpd_engine_half_step.hpp:435-437) and a bead out of
    the box are reported per world."""
    b = W.WorldBuilder((4, 4, 4), dt=0.01, seed=1)
    b.add_bead_type("A")
    b.set_interaction("A", "A", 25.0, 4.5)
    m = b.add_polymer_type("AA", ["A", "A"], bonds=[(0, 1, 128.0, 0.5)])
    b.add_polymer(m, [[0.5, 0.5, 0.5], [2.6, 2.6, 0.5]])       # bond of 2.97 > half of 4
    snapped = b.build()
    ok = W.make_world((5, 5, 5), seed=3)
    ctx = capi.Context(0)
    s_ok, s_bad = ctx.add(ok), ctx.add(snapped)
    ctx.run(2)
    with pytest.raises(capi.PdpdError) as e:
        ctx.sync()
    assert f"world {s_bad}" in str(e.value) and "bond" in str(e.value) and f"world {s_ok}" not in str(e.value)
    assert ctx.flags(s_bad) & 0x10 and ctx.flags(s_ok) == 0
    ctx.remove(s_bad)
    ctx.run(3)
    ctx.sync()                                                  # the failing world gone, the run goes on
    x, v, f, t = ctx.download(s_ok)
    assert t == 5
    ctx.close()


def test_observables_callback_matches_solo_runs(gpu_lib):
    """The serving loop's observables callback: after every interval each
    request's FP64 sums, reduced on the device from the interval's snapshot
    (no copy of the state), equal the sums of the same world run alone to the
    same step, bit for bit; results still equal solo runs."""
    rng = np.random.default_rng(23)
    reqs = [(W.make_world((6, 5, 7), bonded_fraction=0.2, seed=40 + i, a_matrix=W.random_a_matrix(rng)), st)
            for i, st in enumerate((30, 20, 40))]
    _say("observables callback: 3 requests, intervals of 10")
    srv = capi.Server(device=0, interval_steps=10, time_scale=0.0)
    seen = {}

    def on_obs(rid, done, obs):
        seen.setdefault(rid, []).append((done, obs))
        return True
    srv.set_observables_callback(on_obs)
    ids = [srv.submit(w, 0.0, st) for w, st in reqs]
    srv.run()
    results = {rid: srv.result(rid) for rid in ids}
    srv.close()
    for rid, (w, st) in zip(ids, reqs):
        assert results[rid]["status"] == 0
        got = seen[rid]
        assert [d for d, _ in got] == list(range(10, st + 1, 10))
        c = capi.Context(0)
        s = c.add(w)
        for done, obs in got:
            c.run(10)
            c.sync()
            want = c.observables(s)
            assert obs["t"] == want["t"] == done
            for k in ("kinetic", "virial", "n_pairs", "temperature", "pressure"):
                assert obs[k] == want[k], (rid, done, k)
            assert obs["momentum"] == want["momentum"]
        x, v, f, t = c.download(s)
        c.close()
        assert np.array_equal(results[rid]["x"], x) and np.array_equal(results[rid]["f"], f)
