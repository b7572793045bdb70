"""Oracle RNG (hash.hpp:20-88, dpd_maths_core.hpp:55-79) against the reference's
golden vectors, and the invariants the algorithm fixes."""
import os

import numpy as np
import pytest

import oracle as O


@pytest.fixture(scope="module")
def rng(golden_dir):
    return np.load(os.path.join(golden_dir, "rng_golden.npz"))


def test_t_hash_matches_reference_vectors(rng):
    got = O.t_hash(rng["t"], rng["seed"][0])[:1]   # broadcasting form
    assert got[0] == rng["t_hash"][0]
    got = np.array([O.t_hash(np.array([t]), s)[0] for t, s in zip(rng["t"], rng["seed"])], np.uint64)
    assert np.array_equal(got, rng["t_hash"])


def test_hash_rng_sym_matches_reference_vectors(rng):
    u, f = O.hash_rng_sym(rng["pair_t_hash"], rng["a"], rng["b"])
    assert np.array_equal(u, rng["u32"])
    assert np.array_equal(f.view(np.uint32), rng["f32"].view(np.uint32))


def test_bead_hash_matches_reference_vectors(rng):
    from pdpd_amd.world import bead_hash
    got = bead_hash(rng["bh_type"], rng["bh_monomer"], rng["bh_polymer_id"], rng["bh_offset"])
    assert np.array_equal(got, rng["bh_hash"])
    lib = O.lib()
    for k in range(len(got)):
        assert lib.orc_bead_hash(int(rng["bh_type"][k]), int(rng["bh_monomer"][k]), int(rng["bh_polymer_id"][k]),
                                 int(rng["bh_offset"][k])) == rng["bh_hash"][k]


def test_bead_hash_closed_form():
    from pdpd_amd.world import bead_hash
    # construct(type=3, monomer, id=5) = (3<<28) | (1<<27) | 5
    assert int(bead_hash(3, 1, 5, 0)) == (3 << 28) | (1 << 27) | 5
    # polymer: type 2, id 7, offset 4 -> (2<<28) | (4<<20) | 7
    assert int(bead_hash(2, 0, 7, 4)) == (2 << 28) | (4 << 20) | 7


def _mix_py(x):
    M = (1 << 64) - 1
    C = 0xd392d2a7
    x ^= x >> 32
    x = (x * C) & M
    x ^= x >> 32
    x = (x * C) & M
    x ^= x >> 32
    return x


def test_rng_against_python_bigint_restatement():
    """Independent check by arbitrary-precision integers: a dropped xor-shift,
    a 32-bit instead of 64-bit product or a wrong operand order fails here."""
    r = np.random.default_rng(3)
    for _ in range(200):
        th = int(r.integers(0, 2**64, dtype=np.uint64))
        a = int(r.integers(0, 2**32, dtype=np.uint32))
        b = int(r.integers(0, 2**32, dtype=np.uint32))
        lo, hi = min(a, b), max(a, b)
        expect = _mix_py((lo | (hi << 32)) ^ th) & 0xFFFFFFFF
        u, f = O.hash_rng_sym(np.array([th], np.uint64), np.array([a], np.uint32), np.array([b], np.uint32))
        assert int(u[0]) == expect
        s = expect - (1 << 32) if expect >= (1 << 31) else expect
        assert float(f[0]) == float(np.float32(np.float32(s) * np.float32(2.0 ** -32)))
    t, seed = 77, 1234567
    assert int(O.t_hash(np.array([t]), seed)[0]) == _mix_py((seed + _mix_py(t)) & ((1 << 64) - 1))


def test_rng_symmetry_and_range():
    r = np.random.default_rng(4)
    th = r.integers(0, 2**64, 5000, dtype=np.uint64)
    a = r.integers(0, 2**32, 5000, dtype=np.uint32)
    b = r.integers(0, 2**32, 5000, dtype=np.uint32)
    u1, f1 = O.hash_rng_sym(th, a, b)
    u2, f2 = O.hash_rng_sym(th, b, a)
    assert np.array_equal(u1, u2)
    assert np.all(f1 >= -0.5) and np.all(f1 <= 0.5)
    # roughly uniform on [-0.5, 0.5): mean ~0, variance ~1/12
    assert abs(float(f1.mean())) < 0.02
    assert abs(float(f1.var()) - 1.0 / 12) < 0.005


def test_t_is_truncated_to_u32():
    # get_t_hash takes uint32_t: t and t + 2^32 give the same hash
    assert O.t_hash(np.array([5], np.uint64).astype(np.uint32), 9)[0] == \
        O.t_hash(np.array([(1 << 32) + 5], np.uint64).astype(np.uint32), 9)[0]
