"""FP64 oracle (NaiveDPDEngine<false> restatement) against fixtures written by the
compiled reference engine, plus closed-form physics checks of the oracle."""
import glob
import os

import numpy as np
import pytest

import oracle as O
from fixture_io import world_from_arrays
from pdpd_amd import world as W

CASES = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "naive_*.npz")))


@pytest.mark.parametrize("path", CASES, ids=[os.path.basename(c)[6:-4] for c in CASES])
@pytest.mark.parametrize("steps", [1, 10])
def test_naive_restatement_bit_exact_vs_reference(path, steps):
    z = np.load(path)
    w = world_from_arrays(z)
    s = O.F64State(w).run_naive(steps)
    assert s.t == int(z[f"naive{steps}_t"])
    # same operations in the same order as the reference: bit-identical FP64
    assert np.array_equal(s.x, z[f"naive{steps}_x"])
    assert np.array_equal(s.v, z[f"naive{steps}_v"])
    assert np.array_equal(s.f, z[f"naive{steps}_f"])


@pytest.mark.parametrize("path", CASES, ids=[os.path.basename(c)[6:-4] for c in CASES])
def test_half_merge_restatement_vs_reference(path):
    z = np.load(path)
    if "hm1_f" not in z:
        pytest.skip("half-merge engine rejects this box (sides not multiples of 4)")
    w = world_from_arrays(z)
    for threads in (1, 4):
        s = O.F64State(w).run_half_merge(1, nthreads=threads)
        # different pair visiting order than the reference's unordered_set stencil:
        # equal to FP64 rounding
        assert np.abs(s.f - z["hm1_f"]).max() < 1e-9 * max(1.0, np.abs(z["hm1_f"]).max())
        assert np.abs(s.x - z["hm1_x"]).max() < 1e-12


def _two_bead_world(x0, x1, v0=(0, 0, 0), v1=(0, 0, 0), a=25.0, gamma=0.0, box=(8, 8, 8)):
    b = W.WorldBuilder(box, dt=0.01, seed=1)
    b.add_bead_type("A")
    b.set_interaction("A", "A", a, gamma)
    m = b.add_polymer_type("A", ["A"])
    b.add_polymer(m, x0, v0)
    b.add_polymer(m, x1, v1)
    return b.build()


def test_conservative_pair_force_closed_form():
    """gamma = 0: f = a (1 - r) along the pair axis after one step (no random
    term since sigma = sqrt(2 gamma kT) = 0). A sign or index error fails."""
    w = _two_bead_world([2.0, 2.0, 2.0], [2.6, 2.0, 2.0])
    s = O.F64State(w).run_naive(1)
    r = s.x[1, 0] - s.x[0, 0]
    assert s.f[0, 0] == pytest.approx(-25.0 * (1 - r), rel=1e-12)
    assert s.f[1, 0] == pytest.approx(25.0 * (1 - r), rel=1e-12)
    assert np.all(s.f[:, 1:] == 0)


def test_pair_force_zero_beyond_cutoff_and_across_boundary():
    w = _two_bead_world([0.2, 4.0, 4.0], [7.5, 4.0, 4.0])   # 0.7 apart through the wall
    s = O.F64State(w).run_naive(1)
    assert s.f[0, 0] == pytest.approx(25.0 * 0.3, rel=1e-12)
    w = _two_bead_world([1.0, 1.0, 1.0], [2.5, 1.0, 1.0])
    s = O.F64State(w).run_naive(1)
    assert np.all(s.f == 0)


def test_isolated_moving_bead_closed_form():
    """test_isolated_moving_bead.hpp:102-126: x = x0 + v t, wrapped."""
    b = W.WorldBuilder((5, 6, 7), dt=0.05, seed=3)
    b.add_bead_type("A")
    m = b.add_polymer_type("A", ["A"])
    b.add_polymer(m, [4.5, 0.5, 6.5], [1.5, -2.0, 0.7])
    w = b.build()
    s = O.F64State(w).run_naive(40)
    expect = np.mod(np.array([4.5, 0.5, 6.5]) + np.array([1.5, -2.0, 0.7]) * 0.05 * 40, [5, 6, 7])
    assert np.allclose(s.x[0], expect, atol=1e-9)


def test_newton_third_law_sum_of_forces():
    w = W.make_world((6, 6, 6), bonded_fraction=0.2, seed=21, a_matrix=W.random_a_matrix(np.random.default_rng(2)))
    s = O.F64State(w).run_naive(1)
    assert np.abs(s.f.sum(axis=0)).max() < 1e-9 * np.abs(s.f).sum()


def test_hookean_bond_closed_form():
    """Two bonded beads, no DPD: f = kappa (r - r0) along the bond (toward each
    other when stretched)."""
    b = W.WorldBuilder((8, 8, 8), dt=0.01, seed=3)
    b.add_bead_type("A")
    b.set_interaction("A", "A", 0.0, 0.0)
    d = b.add_polymer_type("AA", ["A", "A"], bonds=[(0, 1, 128.0, 0.5)])
    b.add_polymer(d, [[7.9, 1.0, 1.0], [0.6, 1.0, 1.0]])   # 0.7 apart through the wall
    w = b.build()
    s = O.F64State(w).run_naive(1)
    r = (s.x[1, 0] + 8.0) - s.x[0, 0]
    assert s.f[0, 0] == pytest.approx(128.0 * (r - 0.5), rel=1e-12)
    assert s.f[1, 0] == pytest.approx(-128.0 * (r - 0.5), rel=1e-12)


def test_straight_angle_bond_has_no_force_when_straight():
    b = W.WorldBuilder((8, 8, 8), dt=0.01, seed=3)
    b.add_bead_type("A")
    b.set_interaction("A", "A", 0.0, 0.0)
    t = b.add_polymer_type("AAA", ["A", "A", "A"], bonds=[(0, 1, 1e-9, 0.5), (1, 2, 1e-9, 0.5)],
                           bond_pairs=[(0, 1, 20.0, 0.0)])
    b.add_polymer(t, [[1.0, 2.0, 2.0], [1.5, 2.0, 2.0], [2.0, 2.0, 2.0]])
    w = b.build()
    s = O.F64State(w).run_naive(1)
    assert np.abs(s.f).max() < 1e-6
