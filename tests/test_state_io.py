"""WorldState files: text v0, v1binary and v2binary (dpd_state_io.hpp:109-588,
cvec3.hpp:10-300).

The Python writer (pdpd_amd.state_io), the C++ header include/pdpd/world_state_io.hpp
(through bin/pdpd_state) and the reference's own reader and writer (compiled from
its headers into oracle/_ref/ref_state_io where the reference tree is present)
must write the same bytes for the same state, and read what the others wrote.
The packed vector form is also pinned by closed forms (exact powers of two,
rounding ties, the carry into the next exponent, the zero and overflow limits)
and by the reference test's error bound (src/test/test_cvec3_half.cpp: error
below 1/128 of the largest component). The GPU test runs bin/run_world with the
HIP engine and checks its snapshots against the same world advanced through the
C ABI."""
import gzip
import os
import subprocess

import numpy as np
import pytest

from pdpd_amd import state_io
from pdpd_amd import world as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PDPD_STATE = os.path.join(ROOT, "poets-dpd_amd", "bin", "pdpd_state")
RUN_WORLD = os.path.join(ROOT, "poets-dpd_amd", "bin", "run_world")
REF_STATE_IO = os.path.join(ROOT, "oracle", "_ref", "ref_state_io")


def _worlds():
    rng = np.random.default_rng(4)
    wall = W.add_wall(W.make_world((6, 5, 7), bonded_fraction=0.3, seed=21, theta0=0.7,
                                   a_matrix=W.random_a_matrix(rng)))
    wall.origin = np.array([-3.0, 0.25, 10.5])
    wall.t = 123
    water = W.make_world((4, 4, 4), seed=2)
    water.f = np.random.default_rng(3).normal(size=water.f.shape)
    return {"water4": water, "wall_bent": wall, "lattice4": W.make_lattice_world()}


WORLDS = _worlds()


def _tool(path):
    if not os.path.exists(path):
        pytest.skip(f"{os.path.relpath(path, ROOT)} is not built")
    return path


def _same_world(a, b, exact, values=True):
    for k in ("box", "conservative", "dissipative", "polymer_type_of", "polymer_bead_start", "polymer_bead_ids",
              "bead_type", "is_monomer", "polymer_id", "polymer_offset"):
        np.testing.assert_array_equal(np.asarray(getattr(a, k)), np.asarray(getattr(b, k)), err_msg=k)
    assert (a.t, a.seed, a.bead_type_names) == (b.t, b.seed, b.bead_type_names)
    assert [(p.name, list(p.bead_types), [tuple(x) for x in p.bonds], [tuple(x) for x in p.bond_pairs])
            for p in a.polymer_types] == \
           [(p.name, list(p.bead_types), [tuple(x) for x in p.bonds], [tuple(x) for x in p.bond_pairs])
            for p in b.polymer_types]
    sa = a.stationary if a.stationary is not None else np.zeros(a.n_types, np.uint8)
    sb = b.stationary if b.stationary is not None else np.zeros(b.n_types, np.uint8)
    np.testing.assert_array_equal(sa, sb)
    o = np.zeros(3) if a.origin is None else np.asarray(a.origin, np.float64)
    for k in ("x", "v", "f") if values else ():
        if exact:   # 17 digits give the double back; positions are written as x + origin, read as that - origin
            want = (a.x + o) - o if k == "x" else getattr(a, k)
            np.testing.assert_array_equal(want, getattr(b, k), err_msg=k)
        else:   # 6 significant digits; positions are written with the origin added
            np.testing.assert_allclose(getattr(a, k), getattr(b, k), rtol=1e-5, atol=1e-4, err_msg=k)


@pytest.mark.parametrize("name", sorted(WORLDS))
@pytest.mark.parametrize("precise", [False, True])
def test_python_round_trip(name, precise):
    w = WORLDS[name]
    back = state_io.loads(state_io.dumps(w, precise=precise))
    _same_world(w, back, exact=precise)
    if precise:
        assert state_io.dumps(back, precise=True) == state_io.dumps(w, precise=True)


def test_header_lines():
    text = state_io.dumps(WORLDS["wall_bent"])
    lines = text.split("\n")
    w = WORLDS["wall_bent"]
    assert lines[0] == f"WorldState v0 4 {len(w.polymer_types)} {w.n_beads} {len(w.polymer_type_of)}"
    assert lines[1] == "T 123 0.01" and lines[2] == "Lambda 0.5"
    assert lines[3] == "Origin -3 0.25 10.5" and lines[4] == "Box 6 5 7"
    assert "BeadType 3 S 0.5 stationary" in lines
    assert any(l.startswith("BondPair 0 1 20.000000 0.700000") for l in lines)
    assert any(l.startswith("ConservativeStrength 0 ") and l.split()[2].endswith(".000000") for l in lines)


@pytest.mark.parametrize("name", sorted(WORLDS))
@pytest.mark.parametrize("precise", [False, True])
def test_cpp_writer_matches_python(name, precise, tmp_path):
    exe = _tool(PDPD_STATE)
    src = tmp_path / "in.state"
    out = tmp_path / "out.state"
    text = state_io.dumps(WORLDS[name], precise=precise)
    src.write_text(text)
    subprocess.run([exe, str(src), str(out)] + (["--precise"] if precise else []), check=True)
    assert out.read_text() == text     # 6-digit text read and written again is unchanged too


@pytest.mark.parametrize("name", sorted(WORLDS))
def test_reference_writer_matches(name, tmp_path):
    exe = _tool(REF_STATE_IO)
    src = tmp_path / "in.state"
    text = state_io.dumps(WORLDS[name])
    src.write_text(text)
    out = tmp_path / "ref.state"
    subprocess.run([exe, str(src), str(out)], check=True)
    assert out.read_text() == text
    # and the reference reads what the C++ tool compresses
    gz = tmp_path / "cpp.state.gz"
    subprocess.run([_tool(PDPD_STATE), str(src), str(gz)], check=True)
    out2 = tmp_path / "ref2.state"
    subprocess.run([exe, str(gz), str(out2)], check=True)
    assert out2.read_text() == text


def test_gzip_files(tmp_path):
    w = WORLDS["wall_bent"]
    p = tmp_path / "w.state.gz"
    state_io.write(str(p), w, precise=True)
    assert gzip.decompress(p.read_bytes()).decode() == state_io.dumps(w, precise=True)
    _same_world(w, state_io.read(str(p)), exact=True)
    if os.path.exists(PDPD_STATE):
        q = tmp_path / "c.state"
        subprocess.run([PDPD_STATE, str(p), str(q), "--precise"], check=True)
        assert q.read_text() == state_io.dumps(w, precise=True)


@pytest.mark.parametrize("edit,msg", [
    (lambda t: t.replace("WorldState v0", "WorldState v1", 1), "not supported"),
    (lambda t: t.replace("\nLambda", "\nLambada", 1), "expected 'Lambda'"),
    (lambda t: "\n".join(l for l in t.split("\n") if not l.startswith("B 5 ")), "expected 'B'"),
    (lambda t: t.replace(" stationary", " sticky", 1), "unknown BeadType flag"),
    (lambda t: t.replace("\nSeed ", "\nSeed 1 ", 1), "fields for 'Seed'"),
])
def test_reader_errors(edit, msg, tmp_path):
    text = edit(state_io.dumps(WORLDS["wall_bent"]))
    with pytest.raises(ValueError, match=msg):
        state_io.loads(text)
    if os.path.exists(PDPD_STATE):
        src = tmp_path / "bad.state"
        src.write_text(text)
        r = subprocess.run([PDPD_STATE, str(src), str(tmp_path / "o.state")], capture_output=True, text=True)
        assert r.returncode != 0 and r.stderr.startswith("pdpd_state:")


@pytest.mark.gpu
@pytest.mark.parametrize("text_state", [False, True])
def test_run_world_matches_capi(text_state, tmp_path):
    """bin/run_world with the HIP engine writes, every 20 steps, the state that
    the same world advanced through the C ABI downloads: to the bit with
    --text-state (17 digits), as the same packed records in v2binary."""
    from pdpd_amd import capi
    exe = _tool(RUN_WORLD)
    w = W.add_wall(W.make_world((6, 6, 6), bonded_fraction=0.3, seed=8, theta0=0.5))
    src = tmp_path / "w.state"
    state_io.write(str(src), w, precise=True)
    w = state_io.read(str(src))              # the exact doubles the tool reads
    cmd = [exe, "hip_dpd_engine", str(src), str(tmp_path / "out"), "6", "20", "10"] + \
          (["--text-state"] if text_state else [])
    print(" ".join(cmd[1:]), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = [l.split(",") for l in r.stdout.strip().split("\n")]
    assert [int(x[4]) for x in rows] == [10 * k for k in range(1, 7)] and all(x[0] == "hip_dpd_engine" for x in rows)
    assert sorted(os.listdir(tmp_path)) == sorted(["w.state"] + [f"out.{t:09d}.state.gz" for t in (20, 40, 60)])
    ctx = capi.Context(0)
    slot = ctx.add(w)
    for t_want in (20, 40, 60):
        ctx.run(20)
        ctx.sync()
        x, v, f, t = ctx.download(slot)
        name = tmp_path / f"out.{t_want:09d}.state.gz"
        assert gzip.decompress(name.read_bytes()).startswith(b"WorldState v0 " if text_state else b"WorldState v2binary ")
        got = state_io.read(str(name))
        assert got.t == t == t_want
        if not text_state:
            x = np.ldexp(state_io.pack_pos(x, w.box).astype(np.float64), -16)
            v = state_io.unpack_half3(state_io.pack_half3(v))
            f = state_io.unpack_half3(state_io.pack_half3(f))
        np.testing.assert_array_equal(got.x, x)
        np.testing.assert_array_equal(got.v, v)
        np.testing.assert_array_equal(got.f, f)
    ctx.close()


# ---- binary formats ---------------------------------------------------------

def _permuted(w, seed):
    """The same world with bead ids shuffled, so binary records need id deltas."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(w.n_beads)            # new id of old bead i
    inv = np.argsort(perm)
    u = w.copy()
    for k in ("x", "v", "f", "bead_type", "is_monomer", "polymer_id", "polymer_offset"):
        setattr(u, k, np.asarray(getattr(w, k))[inv].copy())
    u.polymer_bead_ids = perm[np.asarray(w.polymer_bead_ids)].astype(np.uint32)
    return u


def _edge_world():
    """Water with velocities and forces at the packing's edge cases, and
    positions that round up to the box edge."""
    w = W.make_world((4, 5, 6), seed=9)
    n = w.n_beads
    rng = np.random.default_rng(10)
    v = rng.normal(size=(n, 3)) * np.exp2(rng.integers(-16, 12, size=(n, 1)))
    special = np.array([
        [0, 0, 0], [1, 0, 0], [-1, 0.5, -0.25], [0.999, 0, 0], [255.5 / 256, 1e-3, -1e-3],
        [1e-6, 0, 0], [2 ** -15, 0, 0], [2 ** -14, 0, 0], [3e-310, 1.0, 0], [-0.0, 0, 2.0],
        [32000.0, -1, 1], [0.0497372, -0.0723867, 0], [129 / 256, 0, 0], [-(255.5 / 256), 0, 0],
    ])
    v[:len(special)] = special
    w.v = v
    w.f = rng.normal(size=(n, 3)) * 40.0
    w.f[:len(special)] = special[::-1]
    w.x[0] = [4 - 1e-7, 5 - 2 ** -18, 0.5 * 2 ** -16]    # rounds to the edge -> 0; exact tie
    return w


BIN_WORLDS = {"water_edges": _edge_world(), "wall_bent": WORLDS["wall_bent"],
              "wall_bent_shuffled": _permuted(WORLDS["wall_bent"], 3), "lattice4": WORLDS["lattice4"]}


def test_pack_half3_closed_forms():
    P, U = state_io.pack_half3, state_io.unpack_half3
    assert P([[0, 0, 0]])[0] == 0
    assert P([[1, 0, 0]])[0] == (1 + 15) | 128 << 5          # 1 = 128 * 2^(1-8)
    assert P([[-1, 0, 0]])[0] == (1 + 15) | (512 - 128) << 5  # 9-bit two's complement
    np.testing.assert_array_equal(U(P([[-1, 0.5, -0.25]])), [[-1, 0.5, -0.25]])
    np.testing.assert_array_equal(U(P([[0.999, 0, 0]])), [[1.0, 0, 0]])   # rounds to 256: carry
    assert P([[0.999, 0, 0]])[0] & 31 == 1 + 15
    assert P([[1e-6, 0, 0]])[0] == 0                           # below 2^-15: zero
    assert P([[2 ** -15, 0, 0]])[0] == (-14 + 15) | 128 << 5   # smallest kept exponent
    np.testing.assert_array_equal(U(P([[129.5 / 256, 0, 0]])), [[130 / 256, 0, 0]])   # tie away from zero
    np.testing.assert_array_equal(U(P([[-129.5 / 256, 0, 0]])), [[-130 / 256, 0, 0]])
    with pytest.raises(ValueError):
        P([[65536.0, 0, 0]])
    with pytest.raises(ValueError):
        P([[32767.0, 0, 0]])                                   # carries past 2^15
    with pytest.raises(ValueError):
        P([[np.nan, 0, 0]])
    np.testing.assert_array_equal(state_io._cround(np.array([2.5, -2.5, 0.49999999999999994, -0.5])),
                                  [3, -3, 0, -1])


def test_pack_half3_error_bound():
    """The reference test's bound: each component within 1/128 of the largest."""
    rng = np.random.default_rng(1)
    v = rng.normal(size=(20000, 3)) * np.exp2(rng.integers(-12, 14, size=(20000, 1)))
    got = state_io.unpack_half3(state_io.pack_half3(v))
    largest = np.abs(v).max(axis=1, keepdims=True)
    zero = (got == 0).all(axis=1)
    assert zero.any() and (largest[zero] < 2.0 ** -15).all()       # stored as the zero vector
    err = np.abs(got - v)[~zero] / largest[~zero]
    assert (err < 1 / 128).all()
    assert (err <= 2.0 ** -8 + 1e-15).all()      # half a step of 2^(e-8), with 2^(e-1) <= largest


@pytest.mark.parametrize("name", sorted(BIN_WORLDS))
def test_binary_python_round_trip(name):
    w = BIN_WORLDS[name]
    back = state_io.loads(state_io.dumps(w, binary=True))
    q = back.copy()
    q.x = np.ldexp(state_io.pack_pos(w.x, w.box).astype(np.float64), -16)
    q.v = state_io.unpack_half3(state_io.pack_half3(w.v))
    q.f = state_io.unpack_half3(state_io.pack_half3(w.f))
    for k in ("x", "v", "f"):
        np.testing.assert_array_equal(getattr(back, k), getattr(q, k), err_msg=k)
    if name == "water_edges":
        assert back.x[0, 0] == 0.0 and back.x[0, 2] == 2.0 ** -16        # wrapped edge; tie rounded up
    _same_world(w, back, exact=False, values=False)
    assert state_io.dumps(back, binary=True) == state_io.dumps(w, binary=True)


@pytest.mark.parametrize("name", sorted(BIN_WORLDS))
def test_binary_cpp_and_reference_bytes(name, tmp_path):
    """Python, C++ and the reference write the same v2binary bytes, and read
    each other's files to the same state."""
    exe = _tool(PDPD_STATE)
    w = BIN_WORLDS[name]
    src = tmp_path / "in.state"
    src.write_text(state_io.dumps(w, precise=True))
    mine = state_io.dumps(w, binary=True)
    cpp = tmp_path / "cpp.state"
    subprocess.run([exe, str(src), str(cpp), "--binary"], check=True)
    assert cpp.read_bytes() == mine
    back = tmp_path / "back.state"
    subprocess.run([exe, str(cpp), str(back), "--precise"], check=True)
    assert back.read_text() == state_io.dumps(state_io.loads(mine), precise=True)
    if os.path.exists(REF_STATE_IO):
        ref = tmp_path / "ref.state.gz"
        subprocess.run([REF_STATE_IO, str(src), str(ref), "v2"], check=True)
        assert gzip.decompress(ref.read_bytes()) == mine
        txt = tmp_path / "ref.txt"
        subprocess.run([REF_STATE_IO, str(cpp), str(txt), "text"], check=True)
        assert txt.read_text() == state_io.dumps(state_io.loads(mine))


@pytest.mark.parametrize("name", sorted(BIN_WORLDS))
def test_v1binary_from_reference(name, tmp_path):
    """v1binary (the reference reads it but writes only v2): files made by the
    reference's own v1 record writer read to the same state here as there."""
    ref = _tool(REF_STATE_IO)
    w = BIN_WORLDS[name]
    src = tmp_path / "in.state"
    src.write_text(state_io.dumps(w, precise=True))
    v1 = tmp_path / "v1.state"
    subprocess.run([ref, str(src), str(v1), "v1"], check=True)
    assert v1.read_bytes().startswith(b"WorldState v1binary ")
    got = state_io.loads(v1.read_bytes())
    ref_txt = tmp_path / "ref.txt"
    subprocess.run([ref, str(v1), str(ref_txt), "text"], check=True)
    assert ref_txt.read_text() == state_io.dumps(got)
    cpp_txt = tmp_path / "cpp.txt"
    subprocess.run([_tool(PDPD_STATE), str(v1), str(cpp_txt), "--precise"], check=True)
    assert cpp_txt.read_text() == state_io.dumps(got, precise=True)
    assert (np.abs(got.v - w.v) <= np.abs(w.v).max(axis=1, keepdims=True) / 64 + 2.0 ** -14).all()


def _reversed_polymers(w):
    """v2binary with the polymers in reverse id order (negative id deltas),
    written with the record layout directly."""
    out = bytearray(state_io.dumps(w, binary=True).split(b"BinaryData\n\0")[0] + b"BinaryData\n\0")
    pos = state_io.pack_pos(w.x, w.box)
    vb, fb = state_io.pack_half3(w.v), state_io.pack_half3(w.f)
    prev_p = prev_b = 0xFFFFFFFF
    for pid in reversed(range(len(w.polymer_type_of))):
        ids = [int(i) for i in w.polymer_bead_ids[w.polymer_bead_start[pid]:w.polymer_bead_start[pid + 1]]]
        out += struct_pack("<HI", int(w.polymer_type_of[pid]) | 7 << 13, (pid - prev_p) & 0xFFFFFFFF)
        prev_p = pid
        for b in ids:
            out += struct_pack("<I5I", (b - prev_b) & 0xFFFFFFFF, pos[b, 0], pos[b, 1], pos[b, 2], vb[b], fb[b])
            prev_b = b
    return bytes(out)


def struct_pack(fmt, *a):
    import struct
    return struct.pack(fmt, *a)


def test_binary_id_deltas_all_readers(tmp_path):
    w = BIN_WORLDS["wall_bent_shuffled"]
    data = _reversed_polymers(w)
    got = state_io.loads(data)
    np.testing.assert_array_equal(got.polymer_bead_ids, w.polymer_bead_ids)
    p = tmp_path / "rev.state"
    p.write_bytes(data)
    canon = state_io.dumps(w, binary=True)
    if os.path.exists(PDPD_STATE):
        q = tmp_path / "c.state"
        subprocess.run([PDPD_STATE, str(p), str(q), "--binary"], check=True)
        assert q.read_bytes() == canon
    if os.path.exists(REF_STATE_IO):
        r = tmp_path / "r.state"
        subprocess.run([REF_STATE_IO, str(p), str(r), "v2"], check=True)
        assert r.read_bytes() == canon


@pytest.mark.parametrize("edit,msg", [
    (lambda d: d[:-7], "ends early"),
    (lambda d: d.replace(b"BinaryData\n\0", b"BinaryData\nx\0"), "Unexpected char"),
    (lambda d: d.replace(b"BinaryData\n\0", b"BinaryData\n"), "Unexpected char|Missing null"),
])
def test_binary_reader_errors(edit, msg, tmp_path):
    data = edit(state_io.dumps(BIN_WORLDS["wall_bent"], binary=True))
    with pytest.raises(ValueError, match=msg):
        state_io.loads(data)
    if os.path.exists(PDPD_STATE):
        src = tmp_path / "bad.state"
        src.write_bytes(data)
        r = subprocess.run([PDPD_STATE, str(src), str(tmp_path / "o.state")], capture_output=True, text=True)
        assert r.returncode != 0 and r.stderr.startswith("pdpd_state:")


def test_bz2_and_validate(tmp_path):
    w = WORLDS["wall_bent"]
    p = tmp_path / "w.state.bz2"
    state_io.write(str(p), w, binary=True)
    assert state_io.read(str(p)).n_beads == w.n_beads
    if os.path.exists(PDPD_STATE):
        q = tmp_path / "w.state"
        subprocess.run([PDPD_STATE, str(p), str(q), "--binary", "--validate"], check=True)
        assert q.read_bytes() == state_io.dumps(w, binary=True)
        bad = w.copy()
        bad.v = bad.v.copy()
        bad.v[3, 1] = 100.0                         # above 0.8 / dt
        b = tmp_path / "bad.state"
        state_io.write(str(b), bad, precise=True)
        r = subprocess.run([PDPD_STATE, str(b), str(q), "--validate"], capture_output=True, text=True)
        assert r.returncode == 1 and "WorldState validation failed : |v| <= 0.8/dt" in r.stderr
