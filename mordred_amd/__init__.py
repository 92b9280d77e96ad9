"""mordred_amd -- MI355X (gfx950) query-execution path of the Mordred engine.

The kernels live in ``libmordred_hip.so`` (built from ``csrc/``) behind the C
ABI of ``include/mordred_hip.h``.  Python is host plumbing only: ``abi`` binds
the C ABI, ``datagen`` makes seeded synthetic star-schema tables, ``query``
assembles reference-style kernel arguments for the star queries.
There is no CPU fallback: calling a kernel without the built library raises.
"""
from . import abi  # noqa: F401

SEGMENT_SIZE = abi.SEGMENT_SIZE

__all__ = ["abi", "SEGMENT_SIZE"]
