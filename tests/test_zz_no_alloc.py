"""pdpd_observables_noalloc (the observables without any allocation, partial
sums in the library's file-scope device array) against pdpd_get_observables
(partial sums in memory the library allocates), through the C ABI: the same
kernels in the same order, so the sums agree bit for bit. Two calls in a row,
the second on a smaller layout; the size limit and the rebuilt-layout
requirement are reported as errors."""
import ctypes as C

import numpy as np
import pytest

from pdpd_amd import capi
from pdpd_amd import world as W

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime the engine library itself links (already loaded with it):
    device memory for the caller-owned output buffer."""
    import os
    for name in ("libamdhip64.so.7", "libamdhip64.so.6", "libamdhip64.so"):
        try:
            hip = C.CDLL(name)
            break
        except OSError:
            continue
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return hip


def _noalloc(ctx, n_worlds):
    hip = _hip()
    dev = C.c_void_p()
    assert hip.hipMalloc(C.byref(dev), 8 * 8 * n_worlds) == 0
    try:
        init = np.full(8 * n_worlds, -1.0)
        assert hip.hipMemcpy(dev, init.ctypes.data, init.nbytes, 1) == 0          # host to device
        capi.check(ctx.lib.pdpd_observables_noalloc(ctx.h, dev, n_worlds))
        ctx.sync()
        out = np.zeros(8 * n_worlds)
        assert hip.hipMemcpy(out.ctypes.data, dev, out.nbytes, 2) == 0            # device to host
    finally:
        hip.hipFree(dev)
    return out.reshape(n_worlds, 8)


def _same(ctx, got):
    ref = ctx.observables()
    slots = sorted(ref)                      # layout order is slot order (worlds are kept by slot)
    assert len(slots) == got.shape[0]
    for row, slot in zip(got, slots):
        o = ref[slot]
        want = np.array(o["momentum"] + [o["kinetic"], o["virial"], float(o["n_pairs"]), float(o["n_beads"]), 0.0])
        assert np.array_equal(row.view(np.uint64), want.view(np.uint64)), (slot, row, want)


def test_noalloc_observables_equal_allocating_entry_point(gpu_lib):
    worlds = W.config_batch(4, seed=11)
    print(f"noalloc: 4 worlds, {sum(w.n_beads for w in worlds)} beads", flush=True)
    ctx = capi.Context(0)
    slots = [ctx.add(w) for w in worlds]
    ctx.run(5)
    ctx.sync()
    _same(ctx, _noalloc(ctx, 4))
    # second call, smaller input: two worlds leave
    ctx.remove(slots[0])
    ctx.remove(slots[2])
    with pytest.raises(capi.PdpdError) as e:          # the layout must be rebuilt first
        _noalloc(ctx, 2)
    assert e.value.code == -5
    ctx.run(3)
    ctx.sync()
    _same(ctx, _noalloc(ctx, 2))
    with pytest.raises(capi.PdpdError) as e:          # too small an output
        _noalloc(ctx, 1)
    assert e.value.code == -1
    ctx.close()


def test_noalloc_rejects_layouts_beyond_the_static_array(gpu_lib):
    w = W.make_world((96, 96, 96), seed=2)           # 2.65M beads: 10368 bead blocks > 8192
    print(f"noalloc: {w.n_beads} beads, beyond the array", flush=True)
    ctx = capi.Context(0)
    ctx.add(w)
    ctx.run(1)
    ctx.sync()
    with pytest.raises(capi.PdpdError) as e:
        _noalloc(ctx, 1)
    assert e.value.code == -4 and "8192" in str(e.value)
    ctx.close()
