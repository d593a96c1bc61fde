"""mray_amd: MI355X (gfx950) wavefront path-tracing hot path of MRay.

The compute path is libmray_amd.so (hand-written HIP kernels behind the C-ABI
in include/mray_amd.h); `native` binds it, `scenes` builds the seeded
procedural workloads, `abi` mirrors the C structs.
"""
from . import abi, scenes  # noqa: F401

__all__ = ["abi", "scenes", "native"]
