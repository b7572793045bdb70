"""ctypes front-end of the oracle (liboracle.so) and of the compiled reference
shim (oracle/_ref/libref_shim.so). TEST INFRASTRUCTURE ONLY: imported by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the product.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref_shim.so")
REF_FAST_LIB = os.path.join(HERE, "_ref", "libref_shim_fast.so")


class OrcWorld(C.Structure):
    _fields_ = [
        ("n_beads", C.c_int32), ("box", C.c_double * 3), ("dt", C.c_double), ("lambda_", C.c_double),
        ("seed", C.c_uint64), ("t", C.c_int64), ("n_types", C.c_int32),
        ("conservative", C.c_void_p), ("dissipative", C.c_void_p), ("stationary", C.c_void_p),
        ("x", C.c_void_p), ("v", C.c_void_p), ("f", C.c_void_p),
        ("hash", C.c_void_p), ("bead_type", C.c_void_p),
        ("n_polymers", C.c_int32), ("poly_bond_start", C.c_void_p), ("poly_angle_start", C.c_void_p),
        ("bond_head", C.c_void_p), ("bond_tail", C.c_void_p), ("bond_kappa", C.c_void_p), ("bond_r0", C.c_void_p),
        ("angle_head", C.c_void_p), ("angle_mid", C.c_void_p), ("angle_tail", C.c_void_p),
        ("angle_kappa", C.c_void_p), ("angle_theta0", C.c_void_p),
    ]


class OrcF32World(C.Structure):
    _fields_ = [
        ("n_beads", C.c_int32), ("L", C.c_int32 * 3), ("dt", C.c_float), ("seed", C.c_uint64),
        ("t", C.c_uint32), ("n_types", C.c_int32), ("a", C.c_void_p), ("sg", C.c_void_p),
        ("x", C.c_void_p), ("w", C.c_void_p), ("f", C.c_void_p), ("hash", C.c_void_p),
        ("n_polymers", C.c_int32), ("poly_bond_start", C.c_void_p), ("poly_angle_start", C.c_void_p),
        ("bond_head", C.c_void_p), ("bond_tail", C.c_void_p), ("bond_kappa", C.c_void_p), ("bond_r0", C.c_void_p),
        ("angle_head", C.c_void_p), ("angle_mid", C.c_void_p), ("angle_tail", C.c_void_p),
        ("angle_kappa", C.c_void_p), ("angle_cos0", C.c_void_p), ("angle_sin0", C.c_void_p),
        ("stationary", C.c_void_p), ("cell_par", C.c_void_p),
    ]


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None and a.size else None


_lib = None


def default_threads():
    """Threads for the FP32 pair-force pass: ORC_THREADS, else the CPUs this
    process may run on, at most 16 (results do not depend on the count)."""
    env = os.environ.get("ORC_THREADS")
    if env:
        return max(1, int(env))
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"oracle not built: {LIB} (make -C oracle)")
        L = C.CDLL(LIB)
        L.orc_mix64_m2.restype = C.c_uint64; L.orc_mix64_m2.argtypes = [C.c_uint64]
        L.orc_t_hash.restype = C.c_uint64; L.orc_t_hash.argtypes = [C.c_uint32, C.c_uint64]
        L.orc_hash_rng_sym.restype = C.c_uint32; L.orc_hash_rng_sym.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32]
        L.orc_default_hash.restype = C.c_float; L.orc_default_hash.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32]
        L.orc_rand_sym_f64.restype = C.c_double; L.orc_rand_sym_f64.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32]
        L.orc_bead_hash.restype = C.c_uint32
        L.orc_bead_hash.argtypes = [C.c_uint32, C.c_int, C.c_uint32, C.c_uint32]
        L.orc_t_hash_n.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_hash_rng_sym_n.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_naive_run.restype = C.c_int; L.orc_naive_run.argtypes = [C.POINTER(OrcWorld), C.c_int]
        L.orc_hm_run.restype = C.c_int; L.orc_hm_run.argtypes = [C.POINTER(OrcWorld), C.c_int, C.c_int]
        L.orc_f32_step.restype = C.c_int
        L.orc_f32_step.argtypes = [C.POINTER(OrcF32World), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_f32_integrate.argtypes = [C.POINTER(OrcF32World)]
        L.orc_f32_cells.restype = C.c_int
        L.orc_f32_cells.argtypes = [C.POINTER(OrcF32World), C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_f32_pair_forces.argtypes = [C.POINTER(OrcF32World), C.c_uint64, C.c_void_p, C.c_void_p]
        L.orc_f32_bonds.argtypes = [C.POINTER(OrcF32World), C.c_void_p]
        L.orc_f32_import.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float,
                                     C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_threads(default_threads())
        _lib = L
    return _lib


# ------------------------------------------------------------------ RNG ------

def t_hash(t, seed):
    t = np.ascontiguousarray(t, dtype=np.uint32).reshape(-1)
    seed = np.ascontiguousarray(np.broadcast_to(seed, t.shape), dtype=np.uint64)
    out = np.zeros(t.size, np.uint64)
    lib().orc_t_hash_n(t.size, _p(t), _p(seed), _p(out))
    return out


def hash_rng_sym(th, a, b):
    th = np.ascontiguousarray(th, dtype=np.uint64).reshape(-1)
    a = np.ascontiguousarray(a, dtype=np.uint32).reshape(-1)
    b = np.ascontiguousarray(b, dtype=np.uint32).reshape(-1)
    u = np.zeros(th.size, np.uint32)
    f = np.zeros(th.size, np.float32)
    lib().orc_hash_rng_sym_n(th.size, _p(th), _p(a), _p(b), _p(u), _p(f))
    return u, f


# ---------------------------------------------------------------- FP64 -------

class F64State:
    """FP64 copy of a World for the naive / half-merge restatements."""

    def __init__(self, world):
        self.world = world
        self.x = np.ascontiguousarray(world.x, dtype=np.float64).copy()
        self.v = np.ascontiguousarray(world.v, dtype=np.float64).copy()
        self.f = np.ascontiguousarray(world.f, dtype=np.float64).copy()
        self.t = int(world.t)
        if getattr(world, "spatial", False):
            raise NotImplementedError("the FP64 oracle takes constant strengths only (FP32 oracle: per-cell strengths)")
        topo = world.flat_topology()
        self.keep = dict(topo)
        self.keep["cons"] = np.ascontiguousarray(world.conservative.reshape(-1), np.float64)
        self.keep["diss"] = np.ascontiguousarray(world.dissipative.reshape(-1), np.float64)
        self.keep["hash"] = np.ascontiguousarray(world.hash, np.uint32)
        self.keep["btype"] = np.ascontiguousarray(world.bead_type, np.uint8)
        self.keep["stat"] = (np.ascontiguousarray(world.stationary, np.uint8)
                             if world.stationary is not None else None)

    def struct(self):
        w = self.world
        k = self.keep
        s = OrcWorld()
        s.n_beads = w.n_beads
        for d in range(3):
            s.box[d] = float(w.box[d])
        s.dt, s.lambda_, s.seed, s.t = float(w.dt), float(w.lambda_), int(w.seed) & (2**64 - 1), self.t
        s.n_types = w.n_types
        s.conservative, s.dissipative, s.stationary = _p(k["cons"]), _p(k["diss"]), _p(k["stat"])
        s.x, s.v, s.f = _p(self.x), _p(self.v), _p(self.f)
        s.hash, s.bead_type = _p(k["hash"]), _p(k["btype"])
        s.n_polymers = len(k["poly_bond_start"]) - 1
        for name in ("poly_bond_start", "poly_angle_start", "bond_head", "bond_tail", "bond_kappa", "bond_r0",
                     "angle_head", "angle_mid", "angle_tail", "angle_kappa", "angle_theta0"):
            setattr(s, name, _p(k[name]))
        return s

    def run_naive(self, nsteps):
        s = self.struct()
        rc = lib().orc_naive_run(C.byref(s), int(nsteps))
        if rc:
            raise RuntimeError("orc_naive_run failed")
        self.t = s.t
        return self

    def run_half_merge(self, nsteps, nthreads=1):
        s = self.struct()
        rc = lib().orc_hm_run(C.byref(s), int(nsteps), int(nthreads))
        if rc:
            raise RuntimeError("orc_hm_run failed")
        self.t = s.t
        return self


# ---------------------------------------------------------------- FP32 -------

def f32_import(world):
    """The engine's FP32 import (orc_f32_import): x, v, f rounded to float,
    w = fmaf(f, -dt/2, v), x >= L after rounding pulled below L."""
    n = world.n_beads
    x = np.zeros((n, 3), np.float32); w = np.zeros((n, 3), np.float32); f = np.zeros((n, 3), np.float32)
    L = np.ascontiguousarray(world.box, dtype=np.int32)
    xd = np.ascontiguousarray(world.x, np.float64); vd = np.ascontiguousarray(world.v, np.float64)
    fd = np.ascontiguousarray(world.f, np.float64)
    lib().orc_f32_import(n, _p(xd), _p(vd), _p(fd), _p(L), float(np.float32(world.dt)), _p(x), _p(w), _p(f))
    return x, w, f


class F32State:
    """FP32 restatement of the HIP engine on one world (bead-index order)."""

    def __init__(self, world):
        self.world = world
        self.L = np.asarray(world.box, dtype=np.int32)
        self.x, self.w, self.f = f32_import(world)
        self.t = int(world.t) & 0xFFFFFFFF
        nt = world.n_types
        self.a = np.ascontiguousarray(world.conservative.reshape(-1), np.float32)
        self.sg = np.ascontiguousarray(np.sqrt(world.dissipative.reshape(-1)), np.float32)
        self.n_types = nt
        topo = world.flat_topology()
        self.keep = {k: v for k, v in topo.items()}
        for k in ("bond_kappa", "bond_r0", "angle_kappa"):
            self.keep[k] = np.ascontiguousarray(topo[k], np.float64)
        # the C library's cos/sin, as the engine's host code uses
        th = np.asarray(topo["angle_theta0"], np.float64)
        self.keep["angle_cos0"] = np.array([math.cos(a) for a in th], np.float64)
        self.keep["angle_sin0"] = np.array([math.sin(a) for a in th], np.float64)
        self.keep["hash"] = np.ascontiguousarray(world.hash, np.uint32)
        self.keep["stat"] = (np.ascontiguousarray(world.stationary, np.uint8)
                             if world.stationary is not None else None)
        # strengths per cell (the world's tables rounded to FP32, (a, gamma) per
        # cell and ordered type pair), or None for constant strengths
        self.keep["cell_par"] = None
        if getattr(world, "spatial", False):
            cc, cd = world.cell_strengths()
            self.keep["cell_par"] = np.ascontiguousarray(np.stack([cc, cd], axis=-1).reshape(-1), np.float32)
        self.order = None
        self.cell_start = None
        self.key = None
        self.fabs = None

    def struct(self):
        w = self.world
        k = self.keep
        s = OrcF32World()
        s.n_beads = w.n_beads
        for d in range(3):
            s.L[d] = int(self.L[d])
        s.dt = float(np.float32(w.dt))
        s.seed = int(w.seed) & (2**64 - 1)
        s.t = self.t
        s.n_types = self.n_types
        s.a, s.sg = _p(self.a), _p(self.sg)
        s.x, s.w, s.f = _p(self.x), _p(self.w), _p(self.f)
        s.hash = _p(k["hash"])
        s.n_polymers = len(k["poly_bond_start"]) - 1
        for name in ("poly_bond_start", "poly_angle_start", "bond_head", "bond_tail", "bond_kappa", "bond_r0",
                     "angle_head", "angle_mid", "angle_tail", "angle_kappa", "angle_cos0", "angle_sin0"):
            setattr(s, name, _p(k[name]))
        s.stationary = _p(k["stat"])
        s.cell_par = _p(k["cell_par"])
        return s

    def step(self, nsteps=1):
        n = self.world.n_beads
        nc = int(np.prod(self.L))
        for _ in range(nsteps):
            self.order = np.zeros(n, np.int32)
            self.cell_start = np.zeros(nc + 1, np.int32)
            self.key = np.zeros(n, np.int32)
            self.fabs = np.zeros(n, np.float32)
            s = self.struct()
            rc = lib().orc_f32_step(C.byref(s), _p(self.order), _p(self.cell_start), _p(self.key), _p(self.fabs))
            if rc:
                raise RuntimeError("orc_f32_step: bead left the box")
            self.t = s.t
        return self

    def exported(self):
        """x, v, f as the engine exports them: v = w + f dt/2 in FP64."""
        hdt = float(np.float32(self.world.dt)) * 0.5
        return (self.x.astype(np.float64), self.w.astype(np.float64) + self.f.astype(np.float64) * hdt,
                self.f.astype(np.float64))


# ------------------------------------------------------------ reference -----

_ref = {}


def ref_available(fast=False):
    return os.path.exists(REF_FAST_LIB if fast else REF_LIB)


def ref_lib(fast=False):
    path = REF_FAST_LIB if fast else REF_LIB
    if path not in _ref:
        if not os.path.exists(path):
            raise RuntimeError(f"reference shim not built: {path} (make -C oracle ref)")
        L = C.CDLL(path)
        L.ref_last_error.restype = C.c_char_p
        L.ref_get_t_hash.restype = C.c_uint64; L.ref_get_t_hash.argtypes = [C.c_uint32, C.c_uint64]
        L.ref_hash_rng_sym.restype = C.c_uint32
        L.ref_hash_rng_sym.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32]
        L.ref_default_hash.restype = C.c_float
        L.ref_default_hash.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32]
        L.ref_bead_hash.restype = C.c_uint32
        L.ref_bead_hash.argtypes = [C.c_uint32, C.c_int, C.c_uint32, C.c_uint32]
        L.ref_run.restype = C.c_int
        L.ref_run.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                              C.POINTER(C.c_int64), C.POINTER(C.c_double)]
        L.ref_run_file.restype = C.c_int
        L.ref_run_file.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.POINTER(C.c_int64)]
        L.ref_run_logged.restype = C.c_int
        L.ref_run_logged.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_char_p]
        L.ref_force_log_sample.restype = C.c_int
        L.ref_force_log_sample.argtypes = [C.c_char_p]
        _ref[path] = L
    return _ref[path]


def ref_run(world, engine, nsteps, fast=False, timing=None):
    """Run a reference engine (compiled from the reference sources by
    oracle/Makefile) on a copy of the world; returns (x, v, f, t) or raises.
    engine: naive | naive_core | half_merge | half_step. If timing is a dict,
    timing["run_s"] receives the wall time of Run alone."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(HERE), "poets-dpd_amd"))
    from pdpd_amd.capi import world_desc
    d, keep = world_desc(world)
    n = world.n_beads
    x = np.zeros((n, 3)); v = np.zeros((n, 3)); f = np.zeros((n, 3))
    t = C.c_int64()
    secs = C.c_double()
    L = ref_lib(fast)
    rc = L.ref_run(C.addressof(d), engine.encode(), int(nsteps), _p(x), _p(v), _p(f), C.byref(t), C.byref(secs))
    if timing is not None:
        timing["run_s"] = secs.value
    if rc == 1:
        raise NotImplementedError(L.ref_last_error().decode())
    if rc:
        raise RuntimeError(L.ref_last_error().decode())
    return x, v, f, t.value


def ref_run_file(path, n_beads, engine, nsteps):
    """Run a reference engine (naive | naive_core) on the world the
    reference's own reader makes of a state file; returns (x, v, f, t)."""
    x = np.zeros((n_beads, 3)); v = np.zeros((n_beads, 3)); f = np.zeros((n_beads, 3))
    t = C.c_int64()
    L = ref_lib()
    rc = L.ref_run_file(str(path).encode(), engine.encode(), int(nsteps), _p(x), _p(v), _p(f), C.byref(t))
    if rc == 1:
        raise NotImplementedError(L.ref_last_error().decode())
    if rc:
        raise RuntimeError(L.ref_last_error().decode())
    return x, v, f, t.value


def ref_run_logged(world, engine, nsteps, log_path):
    """Run a reference engine with the reference's FileLogger writing its force
    log to log_path (the reference's force-logging format)."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(HERE), "poets-dpd_amd"))
    from pdpd_amd.capi import world_desc
    d, keep = world_desc(world)
    L = ref_lib()
    rc = L.ref_run_logged(C.addressof(d), engine.encode(), int(nsteps), str(log_path).encode())
    if rc:
        raise RuntimeError(L.ref_last_error().decode())


def ref_force_log_sample(path):
    """The fixed call sequence of tests/cpp/force_log_sample.inc through the
    reference's FileLogger."""
    ref_lib().ref_force_log_sample(str(path).encode())
