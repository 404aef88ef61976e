"""libcl_amd: MI355X (gfx950) implementation of libCL's 3D SPH step (oclFluid3D) and key/value
radix sort (oclRadixSort), behind the reference's program-object API.

Native code: libcl_amd/csrc/*.hip -> libcl_amd/libmfl.so (C ABI in include/mfl.h).
"""
from ._native import Params, default_params, derive, last_error, load, params_from  # noqa: F401
from .programs import Adjoint, DeviceBuffer, FluidBatch, oclFluid3D, oclRadixSort  # noqa: F401

__all__ = ["Adjoint", "Params", "default_params", "derive", "last_error", "load", "params_from", "DeviceBuffer",
           "FluidBatch", "oclFluid3D", "oclRadixSort"]
