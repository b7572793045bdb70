"""Python side of the MI355X DPD engine: ctypes binding of libpdpd_hip.so and the
WorldState-shaped world description + workload generators.

The engine itself is C++/HIP (poets-dpd_amd/csrc, include/pdpd); this package is
how tests and benchmarks reach it. Importing it does not load the library; the
first Context() or capi.load() does, and raises if it is not built.
"""
from .world import World, WorldBuilder, PolymerType, bead_hash, make_world  # noqa: F401
from . import world  # noqa: F401
from .capi import Context, PdpdError, can_support, limits, debug_rng, load  # noqa: F401
