"""ctypes binding of libpdpd_hip.so (C ABI in include/pdpd_hip.h).

The library must be built in-tree (poets-dpd_amd/lib/libpdpd_hip.so, see
__graft_entry__.build); loading fails loudly otherwise -- there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libpdpd_hip.so")

PDPD_OK = 0
ERRORS = {-1: "invalid", -2: "unsupported", -3: "hip", -4: "capacity", -5: "state", -6: "nomem"}


class PdpdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pdpd error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


ABI_VERSION = 5          # PDPD_ABI_VERSION of the pdpd_world_desc layout below


class WorldDesc(C.Structure):
    _fields_ = [
        ("box", C.c_double * 3), ("dt", C.c_double), ("lambda_", C.c_double),
        ("seed", C.c_uint64), ("t", C.c_int64),
        ("n_bead_types", C.c_uint32), ("conservative", C.c_void_p), ("dissipative", C.c_void_p),
        ("stationary", C.c_void_p),
        ("n_beads", C.c_uint32), ("x", C.c_void_p), ("v", C.c_void_p), ("f", C.c_void_p),
        ("bead_type", C.c_void_p), ("is_monomer", C.c_void_p), ("polymer_id", C.c_void_p),
        ("polymer_offset", C.c_void_p),
        ("n_polymer_types", C.c_uint32), ("pt_bond_start", C.c_void_p), ("bond_head", C.c_void_p),
        ("bond_tail", C.c_void_p), ("bond_kappa", C.c_void_p), ("bond_r0", C.c_void_p),
        ("pt_angle_start", C.c_void_p), ("angle_head_bond", C.c_void_p), ("angle_tail_bond", C.c_void_p),
        ("angle_kappa", C.c_void_p), ("angle_theta0", C.c_void_p),
        ("n_polymers", C.c_uint32), ("polymer_type", C.c_void_p), ("polymer_bead_start", C.c_void_p),
        ("polymer_bead_ids", C.c_void_p),
        ("cell_conservative", C.c_void_p), ("cell_dissipative", C.c_void_p),
    ]


class Limits(C.Structure):
    _fields_ = [("max_bead_types", C.c_uint32), ("min_box", C.c_uint32), ("max_box", C.c_uint32),
                ("max_polymer_length", C.c_uint32), ("max_beads", C.c_uint32),
                ("max_bonds_per_bead", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("upload_s", C.c_double), ("run_s", C.c_double), ("steps", C.c_uint64),
                ("bead_steps", C.c_uint64), ("flags", C.c_uint32), ("n_worlds", C.c_uint32),
                ("n_beads", C.c_uint64), ("n_force_blocks", C.c_uint32),
                ("k_integrate_ms", C.c_double), ("k_scan_ms", C.c_double), ("k_scatter_ms", C.c_double),
                ("k_reorder_ms", C.c_double), ("k_force_ms", C.c_double), ("k_stationary_ms", C.c_double),
                ("profiled_steps", C.c_uint64), ("relayouts", C.c_uint64), ("relayout_s", C.c_double),
                ("relayout_max_s", C.c_double), ("relayout_max_beads", C.c_uint64)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        return d


class ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("hbm_budget_bytes", C.c_uint64), ("interval_steps", C.c_uint32),
                ("max_resident", C.c_uint32), ("time_scale", C.c_double), ("contexts", C.c_uint32)]


class RequestResult(C.Structure):
    _fields_ = [("id", C.c_uint64), ("n_beads", C.c_uint32), ("steps", C.c_uint32), ("arrival_s", C.c_double),
                ("admit_s", C.c_double), ("done_s", C.c_double), ("latency_s", C.c_double), ("status", C.c_int),
                ("fail_flags", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("requests", C.c_uint64), ("completed", C.c_uint64), ("intervals", C.c_uint64),
                ("max_resident", C.c_uint64), ("makespan_s", C.c_double), ("gpu_busy_s", C.c_double),
                ("bead_steps", C.c_double), ("bead_steps_per_s", C.c_double), ("p50_latency_s", C.c_double),
                ("p99_latency_s", C.c_double), ("mean_latency_s", C.c_double), ("max_latency_s", C.c_double),
                ("failed", C.c_uint64), ("relayouts", C.c_uint64), ("max_resident_beads", C.c_uint64),
                ("relayout_mean_s", C.c_double), ("relayout_max_s", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Observables(C.Structure):
    _fields_ = [("slot", C.c_int32), ("n_beads", C.c_uint32), ("t", C.c_int64), ("momentum", C.c_double * 3),
                ("kinetic", C.c_double), ("virial", C.c_double), ("n_pairs", C.c_uint64),
                ("temperature", C.c_double), ("pressure", C.c_double)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["momentum"] = list(self.momentum)
        return d


INTERVAL_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint64, C.c_uint32, C.c_int64, C.c_uint32,
                          C.c_void_p, C.c_void_p, C.c_void_p)
OBSERVABLES_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint64, C.c_uint32, C.POINTER(Observables))

# Every symbol include/pdpd_hip.h declares, with its signature.
# pdpd_pair_record / pdpd_bond_record of pdpd_hip.h
PAIR_RECORD = np.dtype([("home", "<u4"), ("other", "<u4"), ("dx", "<f4", 3), ("dr", "<f4"), ("dv", "<f4", 3),
                        ("diss_strength", "<f4"), ("inv_root_dt", "<f4"), ("gammap", "<f4"), ("rng", "<f4"),
                        ("con", "<f4"), ("diss", "<f4"), ("rng_scale", "<f4"), ("rand", "<f4"), ("f", "<f4", 3)])
BOND_RECORD = np.dtype([("head", "<u4"), ("tail", "<u4"), ("dx", "<f8", 3), ("dr", "<f8"), ("f_head", "<f8", 3)])
assert PAIR_RECORD.itemsize == 80 and BOND_RECORD.itemsize == 64

SIGNATURES = {
    "pdpd_last_error": (C.c_char_p, []),
    "pdpd_abi_version": (C.c_int, []),
    "pdpd_get_limits": (C.c_int, [C.POINTER(Limits)]),
    "pdpd_can_support": (C.c_char_p, [C.POINTER(WorldDesc)]),
    "pdpd_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "pdpd_ctx_destroy": (C.c_int, [C.c_void_p]),
    "pdpd_world_add": (C.c_int, [C.c_void_p, C.POINTER(WorldDesc), C.POINTER(C.c_int)]),
    "pdpd_world_remove": (C.c_int, [C.c_void_p, C.c_int]),
    "pdpd_world_set_steps": (C.c_int, [C.c_void_p, C.c_int, C.c_uint32]),
    "pdpd_run": (C.c_int, [C.c_void_p, C.c_uint32]),
    "pdpd_sync": (C.c_int, [C.c_void_p]),
    "pdpd_world_flags": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_uint32)]),
    "pdpd_world_download": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.POINTER(C.c_int64)]),
    "pdpd_snapshot": (C.c_int, [C.c_void_p]),
    "pdpd_world_snapshot_download": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                               C.POINTER(C.c_int64)]),
    "pdpd_world_download_f32": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.POINTER(C.c_uint32)]),
    "pdpd_world_download_cells": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "pdpd_set_config": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p]),
    "pdpd_get_config": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_size_t]),
    "pdpd_list_config": (C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t]),
    "pdpd_get_stats": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    "pdpd_run_profiled": (C.c_int, [C.c_void_p, C.c_uint32]),
    "pdpd_autotune": (C.c_int, [C.c_void_p, C.c_uint32, C.c_char_p, C.c_size_t]),
    "pdpd_autotune_group": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint32, C.c_uint32, C.c_char_p, C.c_size_t]),
    "pdpd_get_observables": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]),
    "pdpd_world_observables": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(Observables)]),
    "pdpd_observables_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "pdpd_snapshot_observables": (C.c_int, [C.c_void_p]),
    "pdpd_world_snapshot_observables": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(Observables)]),
    "pdpd_debug_rng": (C.c_int, [C.c_int, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p]),
    "pdpd_server_create": (C.c_int, [C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
    "pdpd_server_destroy": (C.c_int, [C.c_void_p]),
    "pdpd_server_set_config": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p]),
    "pdpd_server_set_callback": (C.c_int, [C.c_void_p, INTERVAL_FN, C.c_void_p]),
    "pdpd_server_set_observables_callback": (C.c_int, [C.c_void_p, OBSERVABLES_FN, C.c_void_p]),
    "pdpd_server_submit": (C.c_int, [C.c_void_p, C.POINTER(WorldDesc), C.c_double, C.c_uint32,
                                     C.POINTER(C.c_uint64)]),
    "pdpd_server_run": (C.c_int, [C.c_void_p]),
    "pdpd_server_result": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(RequestResult), C.c_void_p, C.c_void_p,
                                     C.c_void_p]),
    "pdpd_server_get_stats": (C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
    "pdpd_time_force": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_double)]),
    "pdpd_trace_forces": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32),
                                    C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]),
}

_lib = None


def load(path=None):
    """Load libpdpd_hip.so and bind every entry point; raises if missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise PdpdError(-3, f"HIP engine library not built: {p} (run __graft_entry__.build())")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.pdpd_abi_version() != ABI_VERSION:
        raise PdpdError(-3, f"{p}: ABI version {lib.pdpd_abi_version()}, these bindings are for {ABI_VERSION}")
    if path is None:
        _lib = lib
    return lib


def check(rc):
    if rc != PDPD_OK:
        raise PdpdError(rc, load().pdpd_last_error().decode())
    return rc


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None and a.size else None


class _Keep:
    """Holds the numpy arrays a WorldDesc points into."""

    def __init__(self):
        self.arrays = []

    def __call__(self, a, dtype):
        a = np.ascontiguousarray(a, dtype=dtype)
        self.arrays.append(a)
        return a


def world_desc(world):
    """Pack a World (pdpd_amd.world) into the C ABI struct; returns (desc, keepalive)."""
    keep = _Keep()
    d = WorldDesc()
    for k in range(3):
        d.box[k] = float(world.box[k])
    d.dt = float(world.dt)
    d.lambda_ = float(world.lambda_)
    d.seed = int(world.seed) & 0xFFFFFFFFFFFFFFFF
    d.t = int(world.t)
    nt = world.n_types
    d.n_bead_types = nt
    d.conservative = _ptr(keep(world.conservative.reshape(-1), np.float64))
    d.dissipative = _ptr(keep(world.dissipative.reshape(-1), np.float64))
    d.stationary = _ptr(keep(world.stationary, np.uint8)) if world.stationary is not None else None
    d.n_beads = world.n_beads
    d.x = _ptr(keep(world.x.reshape(-1), np.float64))
    d.v = _ptr(keep(world.v.reshape(-1), np.float64))
    d.f = _ptr(keep(world.f.reshape(-1), np.float64))
    d.bead_type = _ptr(keep(world.bead_type, np.uint8))
    d.is_monomer = _ptr(keep(world.is_monomer, np.uint8))
    d.polymer_id = _ptr(keep(world.polymer_id, np.uint32))
    d.polymer_offset = _ptr(keep(world.polymer_offset, np.uint8))
    pts = world.polymer_types
    d.n_polymer_types = len(pts)
    bs, bh, bt, bk, br = [0], [], [], [], []
    as_, ah, at, ak, a0 = [0], [], [], [], []
    for pt in pts:
        for (h, t, k, r0) in pt.bonds:
            bh.append(h); bt.append(t); bk.append(k); br.append(r0)
        bs.append(len(bh))
        for (hb, tb, k, th) in pt.bond_pairs:
            ah.append(hb); at.append(tb); ak.append(k); a0.append(th)
        as_.append(len(ah))
    d.pt_bond_start = _ptr(keep(bs, np.uint32))
    d.bond_head = _ptr(keep(bh, np.uint32))
    d.bond_tail = _ptr(keep(bt, np.uint32))
    d.bond_kappa = _ptr(keep(bk, np.float64))
    d.bond_r0 = _ptr(keep(br, np.float64))
    d.pt_angle_start = _ptr(keep(as_, np.uint32))
    d.angle_head_bond = _ptr(keep(ah, np.uint32))
    d.angle_tail_bond = _ptr(keep(at, np.uint32))
    d.angle_kappa = _ptr(keep(ak, np.float64))
    d.angle_theta0 = _ptr(keep(a0, np.float64))
    d.n_polymers = len(world.polymer_type_of)
    d.polymer_type = _ptr(keep(world.polymer_type_of, np.uint32))
    d.polymer_bead_start = _ptr(keep(world.polymer_bead_start, np.uint32))
    d.polymer_bead_ids = _ptr(keep(world.polymer_bead_ids, np.uint32))
    if world.spatial:
        cc, cd = world.cell_strengths()
        d.cell_conservative = _ptr(keep(cc.reshape(-1), np.float64))
        d.cell_dissipative = _ptr(keep(cd.reshape(-1), np.float64))
    return d, keep


def can_support(world):
    d, keep = world_desc(world)
    return load().pdpd_can_support(C.byref(d)).decode()


def limits():
    lim = Limits()
    check(load().pdpd_get_limits(C.byref(lim)))
    return {k: getattr(lim, k) for k, _ in Limits._fields_}


def debug_rng(t_hash, a, b, device=0):
    t_hash = np.ascontiguousarray(t_hash, dtype=np.uint64)
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    n = t_hash.size
    u = np.zeros(n, dtype=np.uint32)
    f = np.zeros(n, dtype=np.float32)
    check(load().pdpd_debug_rng(device, n, _ptr(t_hash), _ptr(a), _ptr(b), _ptr(u), _ptr(f)))
    return u, f


class Context:
    """A device context holding resident worlds (one world = the plain engine)."""

    def __init__(self, device=0):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.pdpd_ctx_create(device, C.byref(h)))
        self.h = h
        self.worlds = {}

    def close(self):
        if self.h:
            self.lib.pdpd_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add(self, world):
        d, keep = world_desc(world)
        slot = C.c_int()
        check(self.lib.pdpd_world_add(self.h, C.byref(d), C.byref(slot)))
        self.worlds[slot.value] = (world.n_beads, world.box.copy(), world.dt)
        return slot.value

    def remove(self, slot):
        check(self.lib.pdpd_world_remove(self.h, slot))
        self.worlds.pop(slot, None)

    def set_steps(self, slot, steps):
        check(self.lib.pdpd_world_set_steps(self.h, slot, int(steps)))

    def run(self, nsteps):
        check(self.lib.pdpd_run(self.h, int(nsteps)))

    def sync(self):
        check(self.lib.pdpd_sync(self.h))

    def run_profiled(self, nsteps):
        check(self.lib.pdpd_run_profiled(self.h, int(nsteps)))

    def flags(self, slot):
        """PDPD_FLAG_* bits the world has raised (pdpd_world_flags)."""
        f = C.c_uint32()
        check(self.lib.pdpd_world_flags(self.h, slot, C.byref(f)))
        return f.value

    def download(self, slot):
        n = self.worlds[slot][0]
        x = np.zeros((n, 3)); v = np.zeros((n, 3)); f = np.zeros((n, 3))
        t = C.c_int64()
        check(self.lib.pdpd_world_download(self.h, slot, _ptr(x), _ptr(v), _ptr(f), C.byref(t)))
        return x, v, f, t.value

    def snapshot(self):
        """Enqueue a copy of every resident world's state behind the work enqueued so far."""
        check(self.lib.pdpd_snapshot(self.h))

    def snapshot_observables(self):
        """snapshot() plus the FP64 observables of that state, reduced on the device."""
        check(self.lib.pdpd_snapshot_observables(self.h))

    def snapshot_world_observables(self, slot):
        """One world's observables from the last snapshot_observables() (waits for them only)."""
        o = Observables()
        check(self.lib.pdpd_world_snapshot_observables(self.h, slot, C.byref(o)))
        return o.as_dict()

    def snapshot_download(self, slot):
        """The last snapshot of the slot as (x, v, f, t), without waiting for later steps."""
        n = self.worlds[slot][0]
        x = np.zeros((n, 3)); v = np.zeros((n, 3)); f = np.zeros((n, 3))
        t = C.c_int64()
        check(self.lib.pdpd_world_snapshot_download(self.h, slot, _ptr(x), _ptr(v), _ptr(f), C.byref(t)))
        return x, v, f, t.value

    def download_f32(self, slot):
        n = self.worlds[slot][0]
        x = np.zeros((n, 3), np.float32); w = np.zeros((n, 3), np.float32); f = np.zeros((n, 3), np.float32)
        t = C.c_uint32()
        check(self.lib.pdpd_world_download_f32(self.h, slot, _ptr(x), _ptr(w), _ptr(f), C.byref(t)))
        return x, w, f, t.value

    def download_cells(self, slot):
        n, box, _ = self.worlds[slot]
        nc = int(np.prod(box))
        order = np.zeros(n, np.int32)
        cs = np.zeros(nc + 1, np.int32)
        check(self.lib.pdpd_world_download_cells(self.h, slot, _ptr(order), _ptr(cs)))
        return order, cs

    def set_config(self, key, value):
        check(self.lib.pdpd_set_config(self.h, key.encode(), str(value).encode()))

    def get_config(self, key):
        buf = C.create_string_buffer(256)
        check(self.lib.pdpd_get_config(self.h, key.encode(), buf, 256))
        return buf.value.decode()

    def list_config(self):
        buf = C.create_string_buffer(4096)
        check(self.lib.pdpd_list_config(self.h, buf, 4096))
        return buf.value.decode()

    def stats(self):
        s = Stats()
        check(self.lib.pdpd_get_stats(self.h, C.byref(s)))
        return s.as_dict()

    def time_force(self, reps=100):
        """Average duration (ms) of one pair-force launch, HIP events on the engine's stream."""
        ms = C.c_double()
        check(self.lib.pdpd_time_force(self.h, int(reps), C.byref(ms)))
        return ms.value

    def trace_forces(self, slot):
        """Force trace of the step that produced the world's current state
        (pdpd_trace_forces): (pair records, bond records) as numpy record arrays."""
        npair, nbond = C.c_uint32(), C.c_uint32()
        check(self.lib.pdpd_trace_forces(self.h, slot, None, 0, C.byref(npair), None, 0, C.byref(nbond)))
        pairs = np.zeros(npair.value, PAIR_RECORD)
        bonds = np.zeros(nbond.value, BOND_RECORD)
        check(self.lib.pdpd_trace_forces(self.h, slot, pairs.ctypes.data, len(pairs), C.byref(npair),
                                         bonds.ctypes.data, len(bonds), C.byref(nbond)))
        return pairs, bonds

    def autotune(self, steps_per_trial=20):
        buf = C.create_string_buffer(256)
        check(self.lib.pdpd_autotune(self.h, int(steps_per_trial), buf, 256))
        return buf.value.decode()

    def observables(self, slot=None):
        """FP64 observables reduced on the device (pdpd_world_observables /
        pdpd_get_observables): a dict for one slot, or {slot: dict} for all."""
        if slot is not None:
            o = Observables()
            check(self.lib.pdpd_world_observables(self.h, slot, C.byref(o)))
            return o.as_dict()
        arr = (Observables * max(1, len(self.worlds)))()
        n = C.c_uint32()
        check(self.lib.pdpd_get_observables(self.h, C.cast(arr, C.c_void_p), len(self.worlds), C.byref(n)))
        return {arr[i].slot: arr[i].as_dict() for i in range(min(n.value, len(self.worlds)))}


def suggested_parts(total_beads):
    """Contexts an ensemble of worlds is spread over (each with its own stream;
    their launches overlap on the GPU). Measured on the 256-world mixed
    ensemble (6.0M beads): 1 context 3.62e9 bead-steps/s, 4: 3.81e9, 8: 3.94e9,
    12: 4.10e9, 16: 3.58e9."""
    if total_beads >= 4_000_000:
        return 12
    if total_beads >= 1_000_000:
        return 4
    return 1


class Ensemble:
    """Many worlds advanced together, spread over `parts` contexts of one
    device, each with its own stream: a run enqueues every context's steps, so
    the launches of different contexts overlap. A world's result does not
    depend on the context it lands in (bit for bit the same as run alone)."""

    def __init__(self, device=0, parts=1):
        self.parts = [Context(device) for _ in range(max(1, int(parts)))]
        self.load = [0] * len(self.parts)
        self.where = {}           # ensemble slot -> (part, context slot)
        self.next_slot = 0

    def add(self, world):
        i = self.load.index(min(self.load))
        cs = self.parts[i].add(world)
        self.load[i] += world.n_beads
        slot = self.next_slot
        self.next_slot += 1
        self.where[slot] = (i, cs)
        return slot

    def remove(self, slot):
        i, cs = self.where.pop(slot)
        self.load[i] -= self.parts[i].worlds[cs][0]
        self.parts[i].remove(cs)

    def set_steps(self, slot, steps):
        i, cs = self.where[slot]
        self.parts[i].set_steps(cs, steps)

    def run(self, nsteps):
        for c in self.parts:
            if c.worlds:
                c.run(nsteps)

    def sync(self):
        for c in self.parts:
            c.sync()

    def download(self, slot):
        i, cs = self.where[slot]
        return self.parts[i].download(cs)

    def download_f32(self, slot):
        i, cs = self.where[slot]
        return self.parts[i].download_f32(cs)

    def set_config(self, key, value):
        for c in self.parts:
            c.set_config(key, value)

    def get_config(self, key):
        """The setting of the first context (set_config sets every context alike)."""
        return self.parts[0].get_config(key)

    def autotune(self, steps_per_trial=20):
        """The launch shape of the ensemble, timed as it runs: every trial
        advances all contexts side by side (pdpd_autotune_group), and every
        context takes the fastest shape."""
        if len(self.parts) == 1:
            return self.parts[0].autotune(steps_per_trial)
        lib = self.parts[0].lib
        arr = (C.c_void_p * len(self.parts))(*[c.h.value if isinstance(c.h, C.c_void_p) else c.h
                                                for c in self.parts])
        buf = C.create_string_buffer(256)
        check(lib.pdpd_autotune_group(arr, len(self.parts), int(steps_per_trial), buf, 256))
        return buf.value.decode()

    def close(self):
        for c in self.parts:
            c.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Server:
    """Continuous-batching serving loop (pdpd_server_* in the C ABI)."""

    def __init__(self, device=0, hbm_budget_bytes=0, interval_steps=500, max_resident=0, time_scale=1.0, contexts=1):
        self.lib = load()
        cfg = ServerConfig(device, int(hbm_budget_bytes), int(interval_steps), int(max_resident), float(time_scale),
                           int(contexts))
        h = C.c_void_p()
        check(self.lib.pdpd_server_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.sizes = {}
        self._cb = None
        self._ocb = None

    def close(self):
        if self.h:
            self.lib.pdpd_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_config(self, key, value):
        check(self.lib.pdpd_server_set_config(self.h, key.encode(), str(value).encode()))

    def set_callback(self, fn, copy=True):
        """fn(request_id, steps_done, t, x, v, f) -> bool (False cancels). With
        copy=False the arrays are views valid only during the call."""
        if fn is None:
            self._cb = None
            check(self.lib.pdpd_server_set_callback(self.h, INTERVAL_FN(), None))
            return

        def tramp(user, rid, done, t, n, x, v, f):
            def arr(p):
                a = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_double)), shape=(n, 3))
                return a.copy() if copy else a
            return 1 if fn(int(rid), int(done), int(t), arr(x), arr(v), arr(f)) else 0
        self._cb = INTERVAL_FN(tramp)
        check(self.lib.pdpd_server_set_callback(self.h, self._cb, None))

    def set_observables_callback(self, fn):
        """fn(request_id, steps_done, observables_dict) -> bool (False cancels),
        after every interval, from sums reduced on the device (no state copy)."""
        if fn is None:
            self._ocb = None
            check(self.lib.pdpd_server_set_observables_callback(self.h, OBSERVABLES_FN(), None))
            return

        def tramp(user, rid, done, obs):
            return 1 if fn(int(rid), int(done), obs.contents.as_dict()) else 0
        self._ocb = OBSERVABLES_FN(tramp)
        check(self.lib.pdpd_server_set_observables_callback(self.h, self._ocb, None))

    def submit(self, world, arrival_s, steps):
        d, keep = world_desc(world)
        rid = C.c_uint64()
        check(self.lib.pdpd_server_submit(self.h, C.byref(d), float(arrival_s), int(steps), C.byref(rid)))
        self.sizes[rid.value] = world.n_beads
        return rid.value

    def run(self):
        check(self.lib.pdpd_server_run(self.h))

    def result(self, rid, with_state=True):
        r = RequestResult()
        n = self.sizes[rid]
        x = np.zeros((n, 3)); v = np.zeros((n, 3)); f = np.zeros((n, 3))
        if with_state:
            check(self.lib.pdpd_server_result(self.h, rid, C.byref(r), _ptr(x), _ptr(v), _ptr(f)))
        else:
            check(self.lib.pdpd_server_result(self.h, rid, C.byref(r), None, None, None))
        res = {k: getattr(r, k) for k, _ in RequestResult._fields_}
        if with_state:
            res.update(x=x, v=v, f=f)
        return res

    def stats(self):
        s = ServerStats()
        check(self.lib.pdpd_server_get_stats(self.h, C.byref(s)))
        return s.as_dict()
