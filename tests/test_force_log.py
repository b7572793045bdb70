"""Force logging: the reference's ForceLogging trace (logging.hpp,
logging_impl.hpp) from the HIP engine.

  * CPU: this repository's FileLogger writes the same file as the reference's
    FileLogger for a fixed call sequence (tests/cpp/force_log_sample.inc).
  * GPU: pdpd_trace_forces (the pair and bond records of the step that produced
    the current state) agrees with the forces the step computed, with the
    oracle's random draws bit for bit, and with a brute-force neighbour search;
    and run_world's force log (PDPD_LOG) agrees with the force log of the
    reference's own NaiveDPDEngine (compiled from the reference sources into
    oracle/_ref) for the same world: the same pairs, the same random draws, the
    terms, bond forces and total forces to FP32 accuracy."""
import collections
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from pdpd_amd import capi
from pdpd_amd import state_io
from pdpd_amd import world as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN_WORLD = os.path.join(ROOT, "poets-dpd_amd", "bin", "run_world")


def parse_log(path):
    """{(t, h0, h1, h2, name): [values]} and the per-line order of names."""
    out = {}
    for line in open(path):
        f = line.rstrip("\n").split(",")
        assert f[0].endswith("Prop") and len(f) == 9, line
        key = (int(f[1]), f[2], f[3], f[4], f[5])
        out[key] = [float(v) if v else None for v in f[6:9]]
    return out


def test_file_logger_matches_reference(tmp_path):
    if not O.ref_available():
        pytest.skip("oracle/_ref/libref_shim.so is not built")
    exe = tmp_path / "sample"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), "-I",
                    os.path.join(ROOT, "tests", "cpp"), os.path.join(ROOT, "tests", "cpp", "force_log_sample.cpp"),
                    "-o", str(exe)], check=True)
    mine = tmp_path / "mine.log"
    ref = tmp_path / "ref.log"
    subprocess.run([str(exe), str(mine)], check=True)
    O.ref_force_log_sample(ref)
    text = mine.read_text()
    assert text == ref.read_text()
    assert "Prop,7,,,,seed_lo,4294967295,," in text and "run1:Prop,7,1,2,268435455,tri," in text


def test_reference_log_parses(tmp_path):
    """The reference engine's own log has the structure the GPU comparison uses."""
    if not O.ref_available():
        pytest.skip("oracle/_ref/libref_shim.so is not built")
    w = W.make_world((4, 4, 5), bonded_fraction=0.2, seed=3)
    O.ref_run_logged(w, "naive", 1, tmp_path / "ref.log")
    log = parse_log(tmp_path / "ref.log")
    names = collections.Counter(k[4] for k in log)
    assert names["x"] == w.n_beads and names["f_next"] == w.n_beads
    assert names["dpd-rng"] == names["dr"] == names["f_next_dpd"] > 0
    assert names["f_next_hookean"] > 0 and names["f_next_hookean"] == names["f_next_hookean_dr"]


# ------------------------------------------------------------------ GPU ----

def _world():
    rng = np.random.default_rng(21)
    return W.make_world((6, 5, 7), bonded_fraction=0.25, seed=22, theta0=0.5, a_matrix=W.random_a_matrix(rng))


@pytest.mark.gpu
def test_trace_forces_matches_step(gpu_lib):
    w = _world()
    ctx = capi.Context(0)
    slot = ctx.add(w)
    ctx.run(3)
    ctx.sync()
    pairs, bonds = ctx.trace_forces(slot)
    x, wv, f, t = ctx.download_f32(slot)
    print(f"trace: {len(pairs)} pairs, {len(bonds)} bonds after {t} steps", flush=True)
    assert t == 3 and len(pairs) > 10 * w.n_beads
    # sorted by (home, other), each pair once per side, all within the cutoff
    key = pairs["home"].astype(np.uint64) << np.uint64(32) | pairs["other"]
    assert np.all(key[1:] > key[:-1])
    assert np.all((pairs["dr"] > 0) & (pairs["dr"] < 1))
    np.testing.assert_allclose(np.linalg.norm(pairs["dx"].astype(np.float64), axis=1), pairs["dr"], rtol=2e-7)
    # the random draws are the oracle's (hash.hpp) for step t - 1, bit for bit
    th = O.t_hash(np.full(len(pairs), t - 1), w.seed)
    _, u = O.hash_rng_sym(th, pairs["home"], pairs["other"])
    assert np.array_equal(u.view(np.uint32), pairs["rng"].view(np.uint32))
    # the other side of every pair: same draw and distance, opposite force
    where = {int(k): i for i, k in enumerate(key)}
    back = np.array([where[(int(o) << 32) | int(h)] for h, o in zip(pairs["home"], pairs["other"])])
    assert np.array_equal(pairs["rng"][back], pairs["rng"])
    # (through a periodic image the two sides round the shifted coordinate differently)
    np.testing.assert_allclose(pairs["dr"][back], pairs["dr"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(pairs["f"][back], -pairs["f"], atol=2e-5 * np.abs(pairs["f"]).max())
    # the pair forces add up to the forces the step stored (beads without bonds)
    hashes = np.array([W.bead_hash(w.bead_type[i], w.is_monomer[i], w.polymer_id[i], w.polymer_offset[i])
                       for i in range(w.n_beads)], np.uint32)
    order = np.argsort(hashes)
    idx = order[np.searchsorted(hashes[order], pairs["home"])]
    fsum = np.zeros((w.n_beads, 3))
    np.add.at(fsum, idx, pairs["f"].astype(np.float64))
    fabs = np.zeros(w.n_beads)
    np.add.at(fabs, idx, np.abs(pairs["f"].astype(np.float64)).sum(axis=1))
    free = w.is_monomer == 1
    err = np.abs(fsum[free] - f[free]).max(axis=1)
    assert np.all(err <= 1e-5 * (fabs[free] + 1.0))
    # brute force: every pair well inside the cutoff is traced, nothing outside it
    box = np.asarray(w.box, np.float64)
    d = x[:, None, :].astype(np.float64) - x[None, :, :]
    d -= box * np.round(d / box)
    r = np.linalg.norm(d, axis=2)
    np.fill_diagonal(r, 9.0)
    inside = set(zip(*np.nonzero(r < 1 - 1e-5)))
    traced = set(zip(idx, order[np.searchsorted(hashes[order], pairs["other"])]))
    assert inside <= traced
    assert all(r[i, j] < 1 + 1e-5 for i, j in traced)
    # Hookean bonds: Hooke's law on the traced separation
    assert len(bonds) == sum(len(w.polymer_types[w.polymer_type_of[p]].bonds) for p in range(len(w.polymer_type_of)))
    k, r0 = 128.0, 0.5          # make_world's chain bonds
    dr = bonds["dr"]
    np.testing.assert_allclose(np.linalg.norm(bonds["dx"], axis=1), dr, rtol=1e-12)
    want = -(k * (r0 - dr) / dr)[:, None] * bonds["dx"]
    np.testing.assert_allclose(bonds["f_head"], want, rtol=1e-12, atol=1e-12)
    ctx.close()


@pytest.mark.gpu
def test_run_world_log_matches_reference_log(gpu_lib, tmp_path):
    if not O.ref_available():
        pytest.skip("oracle/_ref/libref_shim.so is not built")
    if not os.path.exists(RUN_WORLD):
        pytest.skip("poets-dpd_amd/bin/run_world is not built")
    w = _world()
    src = tmp_path / "w.state"
    state_io.write(str(src), w, precise=True)
    w = state_io.read(str(src))
    steps = 2
    print("run_world hip_dpd_engine with PDPD_LOG, 2 steps", flush=True)
    r = subprocess.run([RUN_WORLD, "hip_dpd_engine", str(src), str(tmp_path / "out"), "1", str(steps), "--text-state"],
                       capture_output=True, text=True, timeout=120, env=dict(os.environ, PDPD_LOG=str(tmp_path / "gpu.log")))
    assert r.returncode == 0, r.stderr
    O.ref_run_logged(w, "naive", steps, tmp_path / "ref.log")
    mine, ref = parse_log(tmp_path / "gpu.log"), parse_log(tmp_path / "ref.log")
    n = w.n_beads
    names = collections.Counter(k[4] for k in mine)
    for nm in ("b_hash", "x", "v", "f", "v_next", "f_next"):
        assert names[nm] == steps * n, nm
    # the state the first step starts from is logged identically
    for k, v in ref.items():
        if k[0] == 0 and k[4] in ("x", "f"):
            assert mine[k] == v, k
    for t in range(steps):
        rp = {(k[1], k[2]) for k in ref if k[0] == t and k[4] == "dr"}
        mp = {(k[1], k[2]) for k in mine if k[0] == t and k[4] == "dr"}
        # the reference logs each pair once, from one side; this engine from both
        und = lambda s: {tuple(sorted(p)) for p in s}
        edge = {tuple(sorted((k[1], k[2]))) for k, v in ref.items()
                if k[0] == t and k[4] == "dr" and v[0] > 1 - 1e-5}
        assert und(rp) - edge <= und(mp) and und(mp) - und(rp) <= edge | {
            tuple(sorted((k[1], k[2]))) for k, v in mine.items() if k[0] == t and k[4] == "dr" and v[0] > 1 - 1e-5}
        for h0, h1 in rp:
            if (h0, h1) not in mp:
                continue
            get = lambda nm: (np.array(ref[(t, h0, h1, "", nm)], float), np.array(mine[(t, h0, h1, "", nm)], float))
            a, b = get("dpd-rng")
            assert abs(a[0] - b[0]) <= 2e-6                  # the same draw (FP64 there, FP32 here, 6 digits)
            a, b = get("dr")
            assert abs(a[0] - b[0]) < 1e-5
            for nm, tol in (("dpd-con", 1e-4), ("dpd-diss", 1e-3), ("dpd-rand", 1e-3), ("dpd-gammap", 1e-4),
                            ("dpd-rng-scale", 1e-3)):
                a, b = get(nm)
                assert abs(a[0] - b[0]) <= tol * (1 + abs(a[0])), (nm, a, b)
            a, b = get("f_next_dpd")
            assert np.abs(a - b).max() <= 2e-3 * (1 + np.abs(a).max()), (a, b)
        # Hookean bonds, both ways
        hk = [k for k in ref if k[0] == t and k[4] == "f_next_hookean"]
        assert hk and all(k in mine for k in hk)
        for k in hk:
            assert np.abs(np.array(ref[k]) - np.array(mine[k])).max() <= 1e-3 * (1 + np.abs(ref[k]).max()), k
        # the forces after the step
        fr = np.array([ref[k] for k in sorted(ref) if k[0] == t and k[4] == "f_next"])
        fm = np.array([mine[k] for k in sorted(mine) if k[0] == t and k[4] == "f_next"])
        assert np.abs(fr - fm).max() <= 2e-5 * np.abs(fr).max() + 1e-3
