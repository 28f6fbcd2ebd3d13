"""svxg: the fragment-graph elastic-equilibrium stress solve of structvox on one MI355X (gfx950).

The device library libsvxg.so (C ABI: include/svxg.h) does the work; `stress` mirrors the
reference's StressProblem over it, `gen` makes the synthetic workloads, `serving` runs the per-tick
request loop.
"""
from .abi import BOND, BOND_LOAD, MATERIAL, NODE, SVXG_FP64, SVXG_MIXED  # noqa: F401
