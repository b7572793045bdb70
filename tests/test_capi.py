"""The C ABI library loads on a machine without a GPU and exports every symbol
include/pdpd_hip.h declares; host-only entry points work without a device."""
import os
import re

import numpy as np
import pytest

from pdpd_amd import capi
from pdpd_amd import world as W

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "pdpd_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(pdpd_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_functions():
    names = declared_functions()
    assert "pdpd_run" in names and "pdpd_world_add" in names and "pdpd_server_run" in names
    assert len(names) >= 25


@pytest.mark.parametrize("name", declared_functions())
def test_symbol_exported(name):
    lib = capi.load()
    assert hasattr(lib, name), f"{name} declared in pdpd_hip.h but not exported"


def test_python_binding_covers_header():
    missing = [n for n in declared_functions() if n not in capi.SIGNATURES]
    assert not missing, f"capi.SIGNATURES lacks {missing}"


def test_abi_version_and_limits():
    assert capi.load().pdpd_abi_version() == capi.ABI_VERSION == 5
    lim = capi.limits()
    assert lim["max_bead_types"] == 16 and lim["min_box"] == 3 and lim["max_polymer_length"] == 127


def test_can_support_reasons():
    w = W.make_world((4, 4, 4), seed=1)
    assert capi.can_support(w) == ""
    assert capi.can_support(W.add_wall(w)) == ""          # stationary bead types are supported
    w2 = w.copy()
    w2.conservative = np.full((17, 17), 25.0)
    w2.dissipative = np.full((17, 17), 4.5)
    assert capi.can_support(w2) == "Engine supports 1 to 16 bead types."
    w3 = w.copy()
    w3.box = np.array([4.5, 4, 4])
    assert capi.can_support(w3) == "World bounds must be integer."
    w4 = W.make_world((2, 4, 4), seed=1)
    assert capi.can_support(w4) == "All dimensions must be at least 3."


def test_context_without_device_fails_loudly():
    """No CPU fallback: creating a context needs a HIP device."""
    import subprocess
    import sys
    probe = ("import sys; sys.path.insert(0, %r); from pdpd_amd import capi\n"
             "try:\n    capi.Context(0)\nexcept capi.PdpdError as e:\n    print('ERR', e.code)\nelse:\n    print('OK')\n"
             % os.path.join(os.path.dirname(HEADER), "..", "poets-dpd_amd"))
    out = subprocess.run([sys.executable, "-c", probe], capture_output=True, text=True, timeout=120).stdout
    assert out.startswith("OK") or out.startswith("ERR -3") or out.startswith("ERR -1")


def test_new_entry_points_reject_null_arguments_without_a_device():
    """The group autotuner, snapshot observables and the server's observables
    callback check their arguments before any device call."""
    import ctypes as C
    lib = capi.load()
    buf = C.create_string_buffer(64)
    assert lib.pdpd_autotune_group(None, 0, 4, buf, 64) == -1
    arr = (C.c_void_p * 2)(None, None)
    assert lib.pdpd_autotune_group(arr, 2, 4, buf, 64) == -1
    assert lib.pdpd_snapshot_observables(None) == -1
    assert lib.pdpd_world_snapshot_observables(None, 0, None) == -1
    assert lib.pdpd_server_set_observables_callback(None, capi.OBSERVABLES_FN(), None) == -1
