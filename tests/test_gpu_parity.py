"""Parity of the HIP engine (through the C ABI) with the oracle and with the
reference's own engine outputs.

Bars (stated per test):
  RNG draws, cell keys, cell order, positions after the drift: bit-exact.
  Single-step forces: within 1e-5 of sum_j |F_ij| per bead against the FP32
  oracle (the engine and the oracle perform the same FP32 operations in the
  same order, so the tests also assert equality), and within FP32 error of the
  FP64 reference engine.
"""
import os

import numpy as np
import pytest

import oracle as O
from fixture_io import world_from_arrays
from pdpd_amd import capi
from pdpd_amd import world as W

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _say(msg):
    print(msg, flush=True)


def run_gpu(world, steps, **cfg):
    ctx = capi.Context(0)
    for k, v in cfg.items():
        ctx.set_config(k, v)
    slot = ctx.add(world)
    if steps:
        ctx.run(steps)
    ctx.sync()
    return ctx, slot


def small_worlds():
    rngA = np.random.default_rng(7)
    return {
        "water4": W.make_world((4, 4, 4), seed=11),
        "water3": W.make_world((3, 3, 3), seed=2),
        "poly_6x7x8": W.make_world((6, 7, 8), bonded_fraction=0.2, seed=12, a_matrix=W.random_a_matrix(rngA)),
        "poly_9x5x13": W.make_world((9, 5, 13), bonded_fraction=0.15, seed=14, a_matrix=W.random_a_matrix(rngA)),
        "water16": W.config_water(seed=3),
        # stationary wall slab + chains with theta0 = 1 rad
        "wall_bent8": W.add_wall(W.make_world((8, 8, 8), bonded_fraction=0.2, seed=15, theta0=1.0,
                                              a_matrix=W.random_a_matrix(rngA))),
        "lattice4": W.make_lattice_world(),
    }


def test_rng_bit_exact_vs_reference_vectors(gpu_lib):
    z = np.load(os.path.join(GOLDEN, "rng_golden.npz"))
    _say("rng: 1024 golden triples")
    u, f = capi.debug_rng(z["pair_t_hash"], z["a"], z["b"])
    assert np.array_equal(u, z["u32"])
    assert np.array_equal(f.view(np.uint32), z["f32"].view(np.uint32))
    r = np.random.default_rng(1)
    th = r.integers(0, 2**64, 100000, dtype=np.uint64)
    a = r.integers(0, 2**32, 100000, dtype=np.uint32)
    b = r.integers(0, 2**32, 100000, dtype=np.uint32)
    u, f = capi.debug_rng(th, a, b)
    uo, fo = O.hash_rng_sym(th, a, b)
    assert np.array_equal(u, uo) and np.array_equal(f.view(np.uint32), fo.view(np.uint32))


@pytest.mark.parametrize("name", list(small_worlds().keys()))
def test_one_step_vs_oracle_f32(gpu_lib, name):
    w = small_worlds()[name]
    _say(f"one step: {name} n={w.n_beads} box={tuple(w.box)}")
    ctx, slot = run_gpu(w, 1)
    x, ww, f, t = ctx.download_f32(slot)
    order, cs = ctx.download_cells(slot)
    ref = O.F32State(w).step(1)
    assert t == (int(w.t) + 1)
    # drift + wrap and the counting sort: bit-exact
    assert np.array_equal(x.view(np.uint32), ref.x.view(np.uint32))
    assert np.array_equal(ww.view(np.uint32), ref.w.view(np.uint32))
    assert np.array_equal(cs, ref.cell_start)
    assert np.array_equal(order, ref.order)
    # forces: 1e-5 of the per-bead sum of pair-force magnitudes
    tol = 1e-5 * (ref.fabs[:, None] + np.abs(ref.f) + 1e-3)
    assert np.all(np.abs(f - ref.f) <= tol)
    assert np.array_equal(f.view(np.uint32), ref.f.view(np.uint32))
    ctx.close()


@pytest.mark.parametrize("name", ["poly_6x7x8", "water16", "wall_bent8"])
def test_multi_step_vs_oracle_f32(gpu_lib, name):
    w = small_worlds()[name]
    _say(f"20 steps: {name}")
    ctx, slot = run_gpu(w, 20)
    x, ww, f, t = ctx.download_f32(slot)
    ref = O.F32State(w).step(20)
    # identical FP32 operations: the trajectories agree to the last bit while the
    # force sums do; allow FP32 noise growth in case they do not
    dx = x.astype(np.float64) - ref.x
    dx -= np.round(dx / w.box) * w.box
    assert np.abs(dx).max() < 1e-4
    assert np.abs(ww - ref.w).max() < 1e-3
    assert np.abs(f - ref.f).max() < 1e-2 * max(1.0, np.abs(ref.f).max()) * 1e-1
    _say(f"  bitwise x: {np.array_equal(x.view(np.uint32), ref.x.view(np.uint32))}")
    ctx.close()


CASES = sorted(f for f in os.listdir(GOLDEN) if f.startswith("naive_"))


@pytest.mark.parametrize("case", CASES)
def test_vs_reference_naive_engine(gpu_lib, case):
    """Against NaiveDPDEngine<false> (FP64) outputs written by the reference.
    Stated error for FP32 vs FP64 after s steps: |dx| <= 1e-4 sqrt(s),
    |dv| <= 1e-3 sqrt(s), |df| <= 2e-5 * max|f| * sqrt(s) + 1e-3 (the
    reference's own differential test allows 0.01 sqrt(s) for x and v)."""
    z = np.load(os.path.join(GOLDEN, case))
    w = world_from_arrays(z)
    for steps in (1, 10):
        _say(f"{case}: {steps} steps vs reference naive")
        ctx, slot = run_gpu(w, steps)
        x, v, f, t = ctx.download(slot)
        assert t == int(z[f"naive{steps}_t"])
        rx, rv, rf = z[f"naive{steps}_x"], z[f"naive{steps}_v"], z[f"naive{steps}_f"]
        dx = x - rx
        dx -= np.round(dx / w.box) * w.box
        s = np.sqrt(steps)
        assert np.abs(dx).max() <= 1e-4 * s
        assert np.abs(v - rv).max() <= 1e-3 * s
        assert np.abs(f - rf).max() <= 2e-5 * np.abs(rf).max() * s + 1e-3
        ctx.close()


def test_batch_alone_equals_in_mix_bitwise(gpu_lib):
    """Each world's result is the same alone and inside a batch of uneven worlds
    (same kernels, per-bead summation order independent of the batch)."""
    rng = np.random.default_rng(5)
    worlds = [W.make_world(tuple(int(v) for v in rng.integers(4, 12, 3)), bonded_fraction=float(rng.uniform(0, 0.2)),
                           seed=100 + i, a_matrix=W.random_a_matrix(rng)) for i in range(6)]
    _say("batch of 6 vs alone, 15 steps")
    ctx = capi.Context(0)
    slots = [ctx.add(w) for w in worlds]
    ctx.run(15)
    ctx.sync()
    mixed = [ctx.download_f32(s) for s in slots]
    ctx.close()
    for w, m in zip(worlds, mixed):
        c2, s2 = run_gpu(w, 15)
        alone = c2.download_f32(s2)
        for a, b in zip(alone[:3], m[:3]):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        assert alone[3] == m[3]
        c2.close()


def test_step_budgets_freeze_worlds(gpu_lib):
    """A world whose step budget is spent keeps its state while others advance."""
    wa = W.make_world((5, 6, 7), bonded_fraction=0.1, seed=31, a_matrix=W.random_a_matrix(np.random.default_rng(3)))
    wb = W.make_world((8, 8, 8), seed=32)
    ctx = capi.Context(0)
    sa, sb = ctx.add(wa), ctx.add(wb)
    ctx.set_steps(sa, 3)
    ctx.run(10)
    ctx.sync()
    ga = ctx.download_f32(sa)
    gb = ctx.download_f32(sb)
    assert ga[3] == 3 and gb[3] == 10
    ctx.close()
    for w, g, n in ((wa, ga, 3), (wb, gb, 10)):
        c2, s2 = run_gpu(w, n)
        alone = c2.download_f32(s2)
        for a, b in zip(alone[:3], g[:3]):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        c2.close()


def _same_f32(a, b):
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    assert a[3] == b[3]


def test_worlds_added_and_removed_midrun_bitwise(gpu_lib):
    """A resident bonded world keeps its trajectory bit for bit while other
    worlds join and leave (each change moves it to a new offset in the device
    layout, with its bonded beads' slots)."""
    rng = np.random.default_rng(8)
    wa = W.add_wall(W.make_world((7, 6, 8), bonded_fraction=0.2, seed=51, theta0=0.6, a_matrix=W.random_a_matrix(rng)))
    wb = W.make_world((5, 5, 6), bonded_fraction=0.15, seed=52, a_matrix=W.random_a_matrix(rng))
    wc = W.make_world((6, 4, 5), bonded_fraction=0.1, seed=53)
    _say("world joins and leaves around a resident world, 4 x 5 steps")
    ctx = capi.Context(0)
    sb = ctx.add(wb)             # wa lands after wb; when wb leaves, wa moves down
    sa = ctx.add(wa)
    ctx.run(5)
    sc = ctx.add(wc)
    ctx.run(5)
    ctx.remove(sb)
    ctx.run(5)
    ctx.remove(sc)
    ctx.run(5)
    ctx.sync()
    got = ctx.download_f32(sa)
    ctx.close()
    c2, s2 = run_gpu(wa, 20)
    _same_f32(c2.download_f32(s2), got)
    c2.close()


def test_autotune_leaves_state_unchanged(gpu_lib):
    """Autotuning runs trial steps on the attached worlds and restores them: the
    trajectory afterwards is the untuned one, bit for bit."""
    w = W.make_world((10, 9, 8), bonded_fraction=0.2, seed=61, theta0=0.4,
                     a_matrix=W.random_a_matrix(np.random.default_rng(6)))
    _say("autotune between steps 6 and 12")
    ctx = capi.Context(0)
    s = ctx.add(w)
    ctx.run(6)
    ctx.sync()
    ctx.autotune(4)
    ctx.run(6)
    ctx.sync()
    got = ctx.download_f32(s)
    ctx.close()
    c2, s2 = run_gpu(w, 12)
    _same_f32(c2.download_f32(s2), got)
    c2.close()


@pytest.mark.parametrize("cfg", [dict(graph_steps=0), dict(graph_steps=8), dict(force_grid=64),
                                 dict(force_grid=1024, graph_steps=6),
                                 dict(scan_lookback_tiles=0), dict(scan_lookback_tiles=0, graph_steps=4),
                                 dict(force_variant=1), dict(force_variant=2), dict(force_variant=3),
                                 dict(force_variant=4), dict(force_variant=5), dict(force_variant=6),
                                 dict(force_variant=4, force_grid=64, graph_steps=6),
                                 dict(force_variant=7), dict(force_variant=7, force_grid=64, graph_steps=6),
                                 dict(force_variant=7, fused_integrate=0),
                                 dict(fused_integrate=0), dict(fused_integrate=1, graph_steps=6),
                                 dict(fused_integrate=1, graph_steps=0, force_grid=64),
                                 dict(force_kernel=1), dict(force_kernel=1, force_brick=1),
                                 dict(force_kernel=1, force_brick=2), dict(force_kernel=1, force_brick=3),
                                 dict(force_kernel=1, force_brick=4), dict(force_kernel=1, force_brick=5),
                                 dict(force_kernel=1, force_grid=64, graph_steps=6),
                                 dict(force_kernel=1, fused_integrate=0),
                                 dict(force_kernel=1, force_brick=6), dict(force_kernel=1, force_brick=7),
                                 dict(force_kernel=1, force_brick=8), dict(force_kernel=1, force_brick=9),
                                 dict(force_kernel=1, force_brick=6, force_grid=64, graph_steps=6),
                                 dict(force_kernel=1, force_brick=7, fused_integrate=0)])
def test_launch_shapes_bitwise_identical(gpu_lib, cfg):
    w = W.make_world((12, 10, 9), bonded_fraction=0.1, seed=41, a_matrix=W.random_a_matrix(np.random.default_rng(4)))
    _say(f"launch shape {cfg}")
    c1, s1 = run_gpu(w, 24, graph_steps=0, force_grid=0)
    base = c1.download_f32(s1)
    c1.close()
    c2, s2 = run_gpu(w, 24, **cfg)
    got = c2.download_f32(s2)
    c2.close()
    for a, b in zip(base[:3], got[:3]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("bands", [(16,), (8,), (48,), (0, 16), (16, 0, 4)])
def test_banded_force_order_bitwise_identical(gpu_lib, bands):
    """Force blocks of a world with a large y-z plane (80 x 64 cells) listed in
    y-bands: same bits as the layout order, also when the order changes
    between runs on a resident world."""
    w = W.make_world((6, 80, 64), bonded_fraction=0.05, seed=43, a_matrix=W.random_a_matrix(np.random.default_rng(5)))
    _say(f"banded force order {bands}")
    c1, s1 = run_gpu(w, 24, force_band_rows=0)
    base = c1.download_f32(s1)
    c1.close()
    c2, s2 = run_gpu(w, 0, force_band_rows=bands[0])
    for b in bands[1:]:
        c2.run(24 // len(bands))
        c2.set_config("force_band_rows", b)
        assert c2.get_config("force_band_rows") == str(b)
    c2.run(24 - (24 // len(bands)) * (len(bands) - 1))
    c2.sync()
    got = c2.download_f32(s2)
    c2.close()
    for a, b in zip(base[:3], got[:3]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_export_round_trip_and_momentum(gpu_lib):
    """Attach/Run/export: t advances by n, positions stay in the box, total
    momentum stays at its initial value to FP32 rounding (pair forces are
    antisymmetric bit for bit, bonds conserve momentum)."""
    w = W.config_polymer(seed=5, L=16)
    p0 = w.v.sum(axis=0)
    _say("polymer 16^3, 200 steps")
    ctx, slot = run_gpu(w, 200)
    x, v, f, t = ctx.download(slot)
    assert t == 200
    assert np.all(x >= 0) and np.all(x < w.box)
    assert np.abs(v.sum(axis=0) - p0).max() < 1e-3 * w.n_beads ** 0.5
    ke = 0.5 * (v ** 2).sum() / w.n_beads
    assert 1.0 < ke < 2.5      # started at kT = 1 (1.5 per bead); DPD thermostat holds it near 1.5
    ctx.close()


def test_dense_cell_matches_oracle(gpu_lib):
    """1500 beads in one cell (the reference AVX2 engine would throw on more
    than 32, dpd_engine_avx2_half_merge_tbb.hpp:816-820): the pair lists flush
    as often as needed, the brick kernel stages the rows in chunks, and the
    step still matches the oracle bit for bit; the row-wave kernel finishes
    the dense rows over many rounds."""
    b = W.WorldBuilder((4, 4, 4), dt=0.001, seed=1)
    b.add_bead_type("A")
    b.set_interaction("A", "A", 25.0, 4.5)
    m = b.add_polymer_type("A", ["A"])
    r = np.random.default_rng(0)
    for i in range(1500):
        b.add_polymer(m, 1.0 + 0.9 * r.random(3))
    w = b.build()
    ref = O.F32State(w).step(1)
    # per-slot kernel; brick kernel (its path for bricks beyond the LDS capacity); row waves
    for cfg in (dict(force_kernel=0), dict(force_kernel=1), dict(force_kernel=0, force_variant=7)):
        _say(f"dense cell: 1500 beads in one cell, {cfg}")
        ctx, slot = run_gpu(w, 1, **cfg)
        x, ww, f, t = ctx.download_f32(slot)
        ctx.close()
        assert np.array_equal(x.view(np.uint32), ref.x.view(np.uint32))
        assert np.array_equal(f.view(np.uint32), ref.f.view(np.uint32))


def test_stationary_beads_stay_put(gpu_lib):
    """Beads of a stationary type keep x and end every step with v = f = 0
    (naive_dpd_engine_half_merge.hpp:327-331,387-390); their neighbours still
    feel them."""
    w = W.add_wall(W.make_world((6, 6, 6), seed=21))
    sel = w.bead_type == w.n_types - 1
    _say(f"stationary wall: {int(sel.sum())} of {w.n_beads} beads")
    ctx, slot = run_gpu(w, 7)
    x, v, f, t = ctx.download(slot)
    assert np.array_equal(x[sel], w.x[sel].astype(np.float32).astype(np.float64))
    assert np.all(v[sel] == 0) and np.all(f[sel] == 0)
    assert np.abs(f[~sel]).max() > 0
    ref = O.F32State(w).step(7)
    x32, w32, f32, _ = ctx.download_f32(slot)
    assert np.array_equal(x32.view(np.uint32), ref.x.view(np.uint32))
    assert np.array_equal(f32.view(np.uint32), ref.f.view(np.uint32))
    ctx.close()


def test_unsupported_world_reason(gpu_lib):
    w = W.make_world((4, 4, 4), seed=1)
    w.box = np.array([4.5, 4.0, 4.0])
    assert capi.can_support(w) == "World bounds must be integer."
    ctx = capi.Context(0)
    with pytest.raises(capi.PdpdError) as e:
        ctx.add(w)
    assert e.value.code == -2
    ctx.close()


def test_force_timing_hook_leaves_state_unchanged(gpu_lib):
    """pdpd_time_force re-launches the last step's pair-force kernel; the state
    and the trajectory that follows are bit-identical to a run without it."""
    w = small_worlds()["poly_6x7x8"]
    _say("time_force idempotence: poly_6x7x8")
    a, sa = run_gpu(w, 3)
    before = a.download_f32(sa)
    ms = a.time_force(5)
    after = a.download_f32(sa)
    assert ms > 0
    for u, v in zip(before[:3], after[:3]):
        assert np.array_equal(u.view(np.uint32), v.view(np.uint32))
    a.run(2)
    b, sb = run_gpu(w, 5)
    for u, v in zip(a.download_f32(sa)[:3], b.download_f32(sb)[:3]):
        assert np.array_equal(u.view(np.uint32), v.view(np.uint32))
    a.close()
    b.close()


def test_interval_snapshot_overlaps_later_steps(gpu_lib):
    """pdpd_snapshot: the state at the snapshot, bit for bit, while the
    steps enqueued behind it run; the state after them is unchanged by it."""
    w = W.make_world((10, 9, 8), bonded_fraction=0.2, seed=61)
    _say("snapshot after 5 steps, 7 more enqueued behind it")
    ctx = capi.Context(0)
    slot = ctx.add(w)
    ctx.run(5)
    ctx.snapshot()
    ctx.run(7)
    xs, vs, fs, ts = ctx.snapshot_download(slot)
    x12, v12, f12, t12 = ctx.download(slot)
    ctx.close()
    ref = capi.Context(0)
    rs = ref.add(w)
    ref.run(5)
    x5, v5, f5, t5 = ref.download(rs)
    ref.run(7)
    xr, vr, fr, tr = ref.download(rs)
    ref.close()
    assert ts == t5 == 5 and t12 == tr == 12
    for a, b in ((xs, x5), (vs, v5), (fs, f5), (x12, xr), (v12, vr), (f12, fr)):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    c = capi.Context(0)
    s1 = c.add(w)
    with pytest.raises(capi.PdpdError, match="no snapshot"):
        c.snapshot_download(s1)
    c.close()


def test_ensemble_over_contexts_bitwise_equal_to_solo(gpu_lib):
    """capi.Ensemble spreads worlds over contexts with their own streams; every
    world's result equals its solo run bit for bit, with step budgets too."""
    rng = np.random.default_rng(8)
    worlds = [W.make_world(tuple(int(v) for v in rng.integers(4, 12, 3)), bonded_fraction=0.15, seed=90 + i,
                           a_matrix=W.random_a_matrix(rng)) for i in range(7)]
    _say("ensemble of 7 worlds over 3 contexts, 17 steps, one budget of 9")
    ens = capi.Ensemble(0, parts=3)
    slots = [ens.add(w) for w in worlds]
    ens.set_steps(slots[2], 9)
    ens.run(17)
    ens.sync()
    got = [ens.download(s) for s in slots]
    ens.close()
    for i, w in enumerate(worlds):
        c, s = run_gpu(w, 9 if i == 2 else 17)
        ref = c.download(s)
        c.close()
        assert got[i][3] == ref[3]
        for a, b in zip(got[i][:3], ref[:3]):
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_ensemble_autotune_over_contexts_leaves_state_unchanged(gpu_lib):
    """The group autotuner times every trial with all contexts of an ensemble
    running side by side and restores every world: the trajectories after it
    are the untuned ones bit for bit, every context takes the chosen shape, and
    an empty context (more parts than worlds) takes it too."""
    rng = np.random.default_rng(9)
    worlds = [W.make_world(tuple(int(v) for v in rng.integers(5, 12, 3)), bonded_fraction=0.15, seed=120 + i,
                           a_matrix=W.random_a_matrix(rng)) for i in range(5)]
    _say("ensemble of 5 worlds over 6 contexts (one empty), autotune between steps 4 and 10")
    ens = capi.Ensemble(0, parts=6)
    slots = [ens.add(w) for w in worlds]
    ens.run(4)
    ens.sync()
    choice = ens.autotune(4)
    assert "contexts=6" in choice, choice
    keys = ("force_kernel", "force_brick", "force_grid", "force_variant", "graph_steps", "fused_integrate")
    cfgs = {tuple(c.get_config(k) for k in keys) for c in ens.parts}
    assert len(cfgs) == 1, cfgs
    ens.run(6)
    ens.sync()
    got = [ens.download(s) for s in slots]
    ens.close()
    for i, w in enumerate(worlds):
        c, s = run_gpu(w, 10)
        ref = c.download(s)
        c.close()
        assert got[i][3] == ref[3] == 10
        for a, b in zip(got[i][:3], ref[:3]):
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
