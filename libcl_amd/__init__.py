"""libcl_amd: MI355X (gfx950) implementation of libCL's 3D SPH step (oclFluid3D), key/value radix
sort (oclRadixSort) and triangle-mesh BVH with particle collision (oclBvhTrimesh, clCollideBVH), behind
the reference's program-object API.

Native code: libcl_amd/csrc/*.hip -> libcl_amd/libmfl.so (C ABI in include/mfl.h).
"""
from ._native import Params, default_params, derive, last_error, load, params_from  # noqa: F401
from .programs import (BVH_NODE, Adjoint, DeviceBuffer, FluidBatch, collide_bvh, oclBvhTrimesh,  # noqa: F401
                       oclFluid3D, oclFluid3Dnext, oclRadixSort)

__all__ = ["Adjoint", "Params", "default_params", "derive", "last_error", "load", "params_from", "DeviceBuffer",
           "FluidBatch", "oclFluid3D", "oclFluid3Dnext", "oclRadixSort", "oclBvhTrimesh", "collide_bvh", "BVH_NODE"]
