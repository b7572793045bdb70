"""Parity at the benchmark workloads themselves (not only at small test worlds).

  polymer 32^3   (98,304 beads, 10% in H2T4 chains): one step bit-exact against
                 the FP32 oracle (positions, shifted velocities, cell starts, cell
                 order, forces), and 10 steps bit-exact as well
  large 128^3    (6.3M beads, 5% bonded): one step bit-exact against the oracle
  ensemble x256  (sides 8..32 per axis, own seeds and a_ij, 0..20% bonded): 20
                 steps through the batched engine; every world bitwise equal to
                 its solo run, and 16 worlds (smallest, largest, non-cubic)
                 bitwise equal to the oracle
  served sweep   (cubic sides log-uniform in [8, 48], Poisson arrivals, a budget
                 that holds only part of the requests, intervals of 100 steps,
                 200..1000 steps each): every completed result bitwise equal to
                 the oracle

Bars: cell keys, order and positions are integer / exact work (bit-exact);
forces carry the stated 1e-5 relative bar of the single-step comparison
(1e-5 of the per-bead sum of |F_ij|), asserted here together with equality,
since engine and oracle perform the same FP32 operations in the same order.
The reference's differential bar (test_differential.hpp:70-127) is far looser.
"""
import numpy as np
import pytest

import oracle as O
from pdpd_amd import capi
from pdpd_amd import world as W

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32 if a.dtype == np.float32 else np.uint64)


def _assert_step_matches(ctx, slot, ref, w):
    x, ww, f, t = ctx.download_f32(slot)
    order, cs = ctx.download_cells(slot)
    assert t == ref.t
    assert np.array_equal(_bits(x), _bits(ref.x)), "positions"
    assert np.array_equal(_bits(ww), _bits(ref.w)), "shifted velocities"
    assert np.array_equal(cs, ref.cell_start), "cell starts"
    assert np.array_equal(order, ref.order), "cell order"
    tol = 1e-5 * (ref.fabs[:, None] + np.abs(ref.f) + 1e-3)
    assert np.all(np.abs(f - ref.f) <= tol), "forces beyond 1e-5 of sum |F_ij|"
    assert np.array_equal(_bits(f), _bits(ref.f)), "forces"


def test_polymer32_one_and_ten_steps_bit_exact(gpu_lib):
    w = W.config_polymer(seed=1)
    _say(f"polymer 32^3: {w.n_beads} beads, 1 step then 9 more")
    ctx = capi.Context(0)
    slot = ctx.add(w)
    ref = O.F32State(w)
    ctx.run(1)
    ctx.sync()
    ref.step(1)
    _assert_step_matches(ctx, slot, ref, w)
    ctx.run(9)
    ctx.sync()
    ref.step(9)
    _assert_step_matches(ctx, slot, ref, w)
    ctx.close()


def test_large128_one_step_bit_exact(gpu_lib):
    w = W.config_large(seed=1)
    _say(f"large 128^3: {w.n_beads} beads, 1 step")
    ctx = capi.Context(0)
    slot = ctx.add(w)
    ctx.run(1)
    ctx.sync()
    ref = O.F32State(w).step(1)
    _assert_step_matches(ctx, slot, ref, w)
    ctx.close()


def _oracle_pick(worlds, k=16):
    """Indices of the k smallest and largest worlds and the most non-cubic ones."""
    n = np.array([w.n_beads for w in worlds])
    aspect = np.array([max(w.box) / min(w.box) for w in worlds])
    pick = list(np.argsort(n)[:5]) + list(np.argsort(n)[-5:])
    for i in np.argsort(-aspect):
        if len(pick) >= k:
            break
        if i not in pick:
            pick.append(i)
    return [int(i) for i in pick]


def test_ensemble256_bitwise_vs_solo_and_oracle(gpu_lib):
    worlds = W.config_batch(256, seed=1)
    total = sum(w.n_beads for w in worlds)
    parts = capi.suggested_parts(total)
    _say(f"ensemble: 256 worlds, {total} beads, {parts} contexts, 20 steps")
    ens = capi.Ensemble(0, parts)
    slots = [ens.add(w) for w in worlds]
    ens.run(20)
    ens.sync()
    got = [ens.download_f32(s) for s in slots]
    ens.close()
    _say("ensemble: solo runs")
    for i, w in enumerate(worlds):
        c = capi.Context(0)
        s = c.add(w)
        c.run(20)
        c.sync()
        alone = c.download_f32(s)
        c.close()
        assert alone[3] == got[i][3] == 20
        for a, b in zip(alone[:3], got[i][:3]):
            assert np.array_equal(_bits(a), _bits(b)), f"world {i} differs from its solo run"
    pick = _oracle_pick(worlds)
    _say(f"ensemble: oracle on worlds {pick}")
    for i in pick:
        ref = O.F32State(worlds[i]).step(20)
        x, ww, f, t = got[i]
        assert np.array_equal(_bits(x), _bits(ref.x)), f"world {i} x"
        assert np.array_equal(_bits(ww), _bits(ref.w)), f"world {i} w"
        assert np.array_equal(_bits(f), _bits(ref.f)), f"world {i} f"


def test_served_sweep_results_bitwise_vs_oracle(gpu_lib):
    reqs = W.served_requests(12, seed=4, steps_lo=200, steps_hi=1000)
    worlds = [W.make_world(r["L"], bonded_fraction=r["bonded_fraction"], seed=r["seed"], a_matrix=r["a_matrix"])
              for r in reqs]
    need = [112 * w.n_beads + 8 * int(np.prod(w.box)) + 4096 for w in worlds]
    budget = max(max(need), sum(need) // 3)      # holds about a third of the requests at once
    _say(f"served: {len(reqs)} requests, sides {[r['L'][0] for r in reqs]}, budget {budget} B")
    srv = capi.Server(device=0, hbm_budget_bytes=budget, interval_steps=100, time_scale=0.0)
    ids = [srv.submit(w, r["arrival"], r["steps"]) for w, r in zip(worlds, reqs)]
    srv.run()
    st = srv.stats()
    assert st["completed"] == len(reqs)
    assert 1 <= st["max_resident"] < len(reqs), "the budget should bind"
    for rid, w, r in zip(ids, worlds, reqs):
        res = srv.result(rid)
        assert res["status"] == 0 and res["steps"] == r["steps"]
        _say(f"served: request {rid} side {r['L'][0]} steps {r['steps']} vs oracle")
        x, v, f = O.F32State(w).step(r["steps"]).exported()
        assert np.array_equal(_bits(res["x"]), _bits(x)), f"request {rid} x"
        assert np.array_equal(_bits(res["v"]), _bits(v)), f"request {rid} v"
        assert np.array_equal(_bits(res["f"]), _bits(f)), f"request {rid} f"
    srv.close()
