"""DPD strengths that vary with the cell position: expressions of x, y, z, ux,
uy, uz in the strength matrix (expression.hpp, naive_dpd_engine.hpp:145-188
and 452-468 of the reference). Each cell gets the strengths evaluated with its
bindings; a pair uses the mean of its two beads' cells.

  * expressions: parsing, precedence, the printed form and the errors;
  * files: a state with expressions is read and written back byte for byte by
    the reference's own reader and writer (oracle/_ref/ref_state_io), and by
    this repository's Python and C++ tools;
  * the FP32 oracle's per-cell path: a constant expression reduces it to the
    constant-strength path bit for bit; with varying strengths it agrees with
    the reference's NaiveDPDEngine run on the same file (the reference parses
    and evaluates the expressions itself), and differs from the constant world
    by far more than that agreement's tolerance;
  * GPU: the engine equals the FP32 oracle bit for bit (alone and batched with
    a constant world), agrees with the reference's naive engine, and run_world
    runs such a file; expressions of other variables are rejected with the
    reference's CanSupport message."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from pdpd_amd import capi
from pdpd_amd import expr
from pdpd_amd import state_io
from pdpd_amd import world as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_IO = os.path.join(ROOT, "oracle", "_ref", "ref_state_io")
RUN_WORLD = os.path.join(ROOT, "poets-dpd_amd", "bin", "run_world")
PDPD_STATE = os.path.join(ROOT, "poets-dpd_amd", "bin", "pdpd_state")


def spatial_world(L=(6, 5, 7), seed=17, bonded=0.2):
    w = W.make_world(L, bonded_fraction=bonded, seed=seed)
    w.set_expression("conservative", 0, 0, "25+50*ux")
    w.set_expression("conservative", 0, 1, "30*uy+x/4")
    w.set_expression("dissipative", 1, 1, "3+3*uz")
    w.set_expression("dissipative", 0, 2, "4.5*(1+z/10)")
    return w


# ------------------------------------------------------------ expressions ----

def test_expression_precedence_and_text():
    b = {"x": 2.0, "y": 3.0, "z": 5.0}
    cases = {"1+2*3": 7.0, "(1+2)*3": 9.0, "8/4/2": 1.0, "8-4-2": 2.0, "x*y+z": 11.0, "x*(y+z)": 16.0,
             "2.5e1": 25.0, " x  / y ": 2.0 / 3.0}
    for src, want in cases.items():
        assert expr.evaluate(expr.parse(src), b) == want, src
    assert expr.text(expr.parse("1+2*x")) == "(1.000000+(2.000000*x))"
    assert expr.text(expr.parse("8-4-2")) == "((8.000000-4.000000)-2.000000)"
    assert expr.variables(expr.parse("ux*(y+ux)/z")) == {"ux", "y", "z"}


def test_expression_errors():
    with pytest.raises(ValueError, match="Un-matched"):
        expr.parse("(1+2")
    with pytest.raises(ValueError, match="Unexpected token"):
        expr.parse("-1")                      # the language has no unary minus
    with pytest.raises(ValueError, match="Unconsumed"):
        expr.parse("1)")
    with pytest.raises(ValueError, match="binding"):
        expr.evaluate(expr.parse("q+1"), {})
    with pytest.raises(ValueError, match="Divide by 0"):
        expr.evaluate(expr.parse("1/(x-x)"), {"x": 1.0})


def test_cell_strengths_closed_form():
    w = spatial_world()
    cc, cd = w.cell_strengths()
    L = [int(b) for b in w.box]
    for cx, cy, cz in ((0, 0, 0), (5, 4, 6), (2, 3, 1)):
        c = (cx * L[1] + cy) * L[2] + cz
        assert cc[c, 0, 0] == 25 + 50 * ((cx + 0.5) / L[0])
        assert cc[c, 0, 1] == cc[c, 1, 0] == 30 * ((cy + 0.5) / L[1]) + cx / 4
        assert cd[c, 1, 1] == 3 + 3 * ((cz + 0.5) / L[2])
        assert cd[c, 0, 2] == cd[c, 2, 0] == 4.5 * (1 + cz / 10)
        assert cc[c, 2, 2] == w.conservative[2, 2] and cd[c, 0, 0] == w.dissipative[0, 0]
    with pytest.raises(ValueError, match="cell position"):
        w.set_expression("conservative", 1, 1, "t*2")
    w.set_expression("conservative", 1, 1, "5")              # a number is a constant
    assert w.conservative[1, 1] == 5.0 and ("conservative", 1, 1) not in w.exprs
    w.set_expression("conservative", 1, 1, "(2+3)")          # an expression, if of no variable (as the reference)
    assert w.exprs[("conservative", 1, 1)] == "(2.000000+3.000000)" and np.all(w.cell_strengths()[0][:, 1, 1] == 5)


# ----------------------------------------------------------------- files ----

def test_state_file_with_expressions_round_trips(tmp_path):
    w = spatial_world(bonded=0.1)
    text = state_io.dumps(w)
    assert "(25.000000+(50.000000*ux))" in text
    back = state_io.loads(text)
    assert back.exprs == w.exprs
    assert state_io.dumps(back) == text
    np.testing.assert_array_equal(back.cell_strengths()[0], w.cell_strengths()[0])
    src = tmp_path / "w.state"
    src.write_text(text)
    if os.path.exists(PDPD_STATE):
        out = tmp_path / "cpp.state"
        subprocess.run([PDPD_STATE, str(src), str(out)], check=True, capture_output=True)
        assert out.read_text() == text
    if not os.path.exists(REF_IO):
        pytest.skip("oracle/_ref/ref_state_io is not built")
    for mode in ("text", "v2"):
        out = tmp_path / f"ref.{mode}"
        subprocess.run([REF_IO, str(src), str(out), mode], check=True, capture_output=True)
        mine = state_io.dumps(back, binary=(mode == "v2"))     # the values the text file holds
        got = out.read_bytes()
        assert got == (mine if isinstance(mine, bytes) else mine.encode()), mode


# ---------------------------------------------------------------- oracle ----

def test_oracle_constant_expression_equals_constant_path():
    """Strengths written as expressions with the same value in every cell take
    the per-cell path and give the constant path's forces bit for bit (the mean
    of two equal floats is that float; sqrtf of the FP32 gamma equals the FP64
    square root rounded to FP32 for these values)."""
    w = W.make_world((5, 6, 7), bonded_fraction=0.2, seed=5, a_matrix=W.random_a_matrix(np.random.default_rng(2)))
    ws = w.copy()
    nt = w.n_types
    for i in range(nt):
        for j in range(nt):
            ws.set_expression("conservative", i, j, f"{float(w.conservative[i, j])!r}+0*x", symmetric=False)
            ws.set_expression("dissipative", i, j, f"{float(w.dissipative[i, j])!r}+0*uz", symmetric=False)
    assert ws.spatial and len(ws.exprs) == 2 * nt * nt
    a, b = O.F32State(w).step(3), O.F32State(ws).step(3)
    assert np.array_equal(a.f.view(np.uint32), b.f.view(np.uint32))
    assert np.array_equal(a.x.view(np.uint32), b.x.view(np.uint32))


def _ref_file_run(w, tmp_path, steps):
    path = tmp_path / "ref_in.state"
    state_io.write(str(path), w, precise=True)
    return O.ref_run_file(path, w.n_beads, "naive", steps), state_io.read(str(path))


def test_oracle_spatial_matches_reference_naive(tmp_path):
    if not O.ref_available():
        pytest.skip("oracle/_ref/libref_shim.so is not built")
    w = spatial_world()
    (xr, vr, fr, tr), w = _ref_file_run(w, tmp_path, 1)
    o = O.F32State(w).step(1)
    x, v, f = o.exported()
    scale = np.abs(fr).max()
    err = np.abs(f - fr).max()
    assert tr == 1 and err <= 2e-5 * scale + 1e-4, err
    # the variation matters far beyond that tolerance: the world with every
    # strength at its cell-0 value gives other forces
    flat = w.copy()
    cc, cd = w.cell_strengths()
    flat.conservative, flat.dissipative, flat.exprs = cc[0].copy(), cd[0].copy(), None
    fc = O.F32State(flat).step(1).exported()[2]
    assert np.abs(fc - fr).max() > 100 * (2e-5 * scale + 1e-4)


def test_oracle_spatial_bonded_wall_matches_reference(tmp_path):
    """Stationary wall, angle bonds with theta0 != 0, varying strengths, 3 steps."""
    if not O.ref_available():
        pytest.skip("oracle/_ref/libref_shim.so is not built")
    w = W.add_wall(W.make_world((6, 6, 6), bonded_fraction=0.2, seed=15, theta0=1.0))
    w.set_expression("conservative", 0, 3, "20+40*ux")
    w.set_expression("dissipative", 0, 0, "2+5*uy")
    (xr, vr, fr, tr), w = _ref_file_run(w, tmp_path, 3)
    x, v, f = O.F32State(w).step(3).exported()
    d = x - xr
    d -= np.asarray(w.box) * np.round(d / np.asarray(w.box))
    assert np.abs(d).max() < 1e-4
    assert np.abs(f - fr).max() <= 1e-4 * np.abs(fr).max() + 1e-3


# ----------------------------------------------------------- C++ engine -----

def test_can_support_rejects_non_spatial_variables(tmp_path):
    """A strength of a variable other than the cell position is rejected with
    the reference's message (dpd_engine.hpp:61-83), before any device work."""
    if not os.path.exists(RUN_WORLD):
        pytest.skip("poets-dpd_amd/bin/run_world is not built")
    w = W.make_world((4, 4, 4), seed=3)
    text = state_io.dumps(w).replace("ConservativeStrength 0 25.000000", "ConservativeStrength 0 (25.000000*t)", 1)
    assert "(25.000000*t)" in text
    src = tmp_path / "t.state"
    src.write_text(text)
    r = subprocess.run([RUN_WORLD, "hip_dpd_engine", str(src), str(tmp_path / "out"), "1", "1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 1
    assert "Engine does not support aribtrary varying DPD parameters." in r.stderr, r.stderr


# ------------------------------------------------------------------- GPU ----

@pytest.mark.gpu
def test_gpu_spatial_bitwise_vs_oracle(gpu_lib):
    w = spatial_world()
    wc = W.make_world((7, 5, 6), bonded_fraction=0.1, seed=23)
    ctx = capi.Context(0)
    s1, s2 = ctx.add(w), ctx.add(wc)
    print("spatial + constant world, 4 steps", flush=True)
    ctx.run(4)
    for slot, ww in ((s1, w), (s2, wc)):
        x, wv, f, t = ctx.download_f32(slot)
        ref = O.F32State(ww).step(4)
        assert t == 4
        assert np.array_equal(x.view(np.uint32), ref.x.view(np.uint32))
        assert np.array_equal(wv.view(np.uint32), ref.w.view(np.uint32))
        assert np.array_equal(f.view(np.uint32), ref.f.view(np.uint32))
    # the force trace reports the pair's mean strengths
    pairs, _ = ctx.trace_forces(s1)
    assert len(pairs) > 0 and np.all(np.isfinite(pairs["diss_strength"]))
    assert len(np.unique(pairs["diss_strength"])) > 3
    ctx.close()


@pytest.mark.gpu
def test_gpu_spatial_matches_reference_naive(gpu_lib, tmp_path):
    if not O.ref_available():
        pytest.skip("oracle/_ref/libref_shim.so is not built")
    w = spatial_world()
    (xr, vr, fr, tr), w = _ref_file_run(w, tmp_path, 2)
    ctx = capi.Context(0)
    slot = ctx.add(w)
    ctx.run(2)
    x, v, f, t = ctx.download(slot)
    ctx.close()
    assert t == tr == 2
    d = x - xr
    d -= np.asarray(w.box) * np.round(d / np.asarray(w.box))
    assert np.abs(d).max() < 1e-4
    assert np.abs(f - fr).max() <= 1e-4 * np.abs(fr).max() + 1e-3


@pytest.mark.gpu
def test_gpu_run_world_spatial_file(gpu_lib, tmp_path):
    if not os.path.exists(RUN_WORLD):
        pytest.skip("poets-dpd_amd/bin/run_world is not built")
    w = spatial_world(bonded=0.1)
    src = tmp_path / "s.state"
    state_io.write(str(src), w, precise=True)
    w = state_io.read(str(src))
    print("run_world on a spatial file, 3 steps", flush=True)
    r = subprocess.run([RUN_WORLD, "hip_dpd_engine", str(src), str(tmp_path / "out"), "1", "3", "--text-state"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = state_io.read(str(tmp_path / "out.000000003.state.gz"))
    assert out.exprs == w.exprs                  # the expressions are written back
    ref = O.F32State(w).step(3)
    x, v, f = ref.exported()
    np.testing.assert_array_equal(out.x, x + 0.0)
    np.testing.assert_array_equal(out.f, f)
