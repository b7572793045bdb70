"""FP64 observables reduced on the device (pdpd_world_observables) against the
same sums taken in FP64 on the host from the downloaded state, and long-run
statistics of the water box (16^3 cells, 12,288 beads) against the FP64
half-merge restatement of the reference engine.

Bars: the device sums equal the host sums to 1e-12 relative (both are FP64
sums of the same FP64 terms in different orders; the momentum, a sum with
cancellation, to 1e-12 of sum |v|). Long runs diverge (chaos, FP32 vs FP64),
so temperature and pressure are compared as sample means within 2% and g(r)
within 0.05 per bin beyond the core; the standard error of each mean is
several times smaller."""
import numpy as np
import pytest

import oracle as O
import physics as P
from pdpd_amd import capi
from pdpd_amd import world as W

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("name", ["water16", "polymer32", "ensemble_mix"])
def test_device_observables_match_host_fp64(gpu_lib, name):
    if name == "water16":
        worlds = [W.config_water(seed=3)]
    elif name == "polymer32":
        worlds = [W.config_polymer(seed=1)]
    else:
        worlds = W.config_batch(6, seed=7)
    _say(f"observables: {name}, {sum(w.n_beads for w in worlds)} beads, 7 steps")
    ctx = capi.Context(0)
    slots = [ctx.add(w) for w in worlds]
    ctx.run(7)
    ctx.sync()
    allobs = ctx.observables()
    for w, slot in zip(worlds, slots):
        o = ctx.observables(slot)
        assert o == allobs[slot]
        x, v, f, t = ctx.download(slot)
        mom, kin, vir, npairs = P.virial_sums(x, v, w.box, w.conservative, w.bead_type)
        assert o["t"] == t == 7 and o["n_beads"] == w.n_beads
        assert o["n_pairs"] == npairs
        scale = np.abs(v).sum()
        assert np.all(np.abs(np.array(o["momentum"]) - mom) <= 1e-12 * scale)
        assert abs(o["kinetic"] - kin) <= 1e-12 * kin
        assert abs(o["virial"] - vir) <= 1e-12 * abs(vir)
        vol = float(np.prod(w.box))
        assert abs(o["temperature"] - kin / (3 * w.n_beads)) <= 1e-12 * o["temperature"]
        assert abs(o["pressure"] - (kin / 3 + vir / 3) / vol) <= 1e-12 * abs(o["pressure"])
    ctx.close()


EQUIL, SAMPLES, EVERY = 200, 16, 50     # 1000 steps


def test_water16_long_run_statistics_vs_reference(gpu_lib):
    w = W.config_water(seed=5)
    _say(f"water 16^3: {EQUIL + SAMPLES * EVERY} steps on the GPU and in the FP64 reference restatement")
    ctx = capi.Context(0)
    slot = ctx.add(w)
    ctx.run(EQUIL)
    Tg, pg, gg = [], [], []
    for k in range(SAMPLES):
        ctx.run(EVERY)
        o = ctx.observables(slot)                     # device reductions, no state download
        Tg.append(o["temperature"])
        pg.append(o["pressure"])
        if k % 4 == 3:
            x, v, f, t = ctx.download(slot)
            r, g = P.rdf(x, w.box)
            gg.append(g)
    ctx.close()
    st = O.F64State(w)
    st.run_half_merge(EQUIL, nthreads=O.default_threads())
    Tr, pr, gr = [], [], []
    for k in range(SAMPLES):
        st.run_half_merge(EVERY, nthreads=O.default_threads())
        Tr.append(P.kinetic_temperature(st.v))
        pr.append(P.conservative_pressure(st.x, st.v, w.box, w.conservative, w.bead_type))
        if k % 4 == 3:
            gr.append(P.rdf(st.x, w.box)[1])
    Tg, pg, Tr, pr = map(np.array, (Tg, pg, Tr, pr))
    _say(f"  T gpu {Tg.mean():.4f} ref {Tr.mean():.4f}; p gpu {pg.mean():.3f} ref {pr.mean():.3f}")
    assert abs(Tg.mean() - Tr.mean()) / Tr.mean() < 0.02
    assert abs(pg.mean() - pr.mean()) / pr.mean() < 0.02
    gg, gr = np.mean(gg, axis=0), np.mean(gr, axis=0)
    m = r > 0.3
    assert np.abs(gg[m] - gr[m]).max() < 0.05
    assert abs(gg[r > 1.5].mean() - 1.0) < 0.03


def test_snapshot_observables_equal_synchronous_ones(gpu_lib):
    """pdpd_snapshot_observables reduces the snapshot's state behind the
    enqueued steps; read after more steps were enqueued, the sums equal those
    pdpd_world_observables takes of the same state in another context, bit for
    bit (same kernels, same per-world block partition)."""
    rng = np.random.default_rng(21)
    worlds = [W.make_world((7, 6, 8), bonded_fraction=0.2, seed=31, a_matrix=W.random_a_matrix(rng)),
              W.make_world((5, 9, 6), seed=32)]
    _say("snapshot observables of 2 worlds after 12 steps, 5 more enqueued behind")
    ctx = capi.Context(0)
    slots = [ctx.add(w) for w in worlds]
    ctx.run(12)
    ctx.snapshot_observables()
    ctx.run(5)
    got = [ctx.snapshot_world_observables(s) for s in slots]
    ctx.sync()
    ctx.close()
    ref = capi.Context(0)
    rs = [ref.add(w) for w in worlds]
    ref.run(12)
    ref.sync()
    want = [ref.observables(s) for s in rs]
    ref.close()
    for g, w_ in zip(got, want):
        assert g["t"] == w_["t"] == 12
        for k in ("kinetic", "virial", "n_pairs", "temperature", "pressure", "n_beads"):
            assert g[k] == w_[k], k
        assert g["momentum"] == w_["momentum"]
