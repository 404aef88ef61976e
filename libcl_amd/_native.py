"""ctypes binding of libmfl.so (the C ABI declared in include/mfl.h).

There is no fallback: if the library is missing or cannot be loaded, every program object raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmfl.so")

MFL_POSITION, MFL_VELOCITY, MFL_FORCE = 0, 1, 2
MFL_SORTED_POSITION, MFL_SORTED_VELOCITY, MFL_INDEX = 3, 4, 5
MFL_CELL, MFL_CELL_START, MFL_CELL_END, MFL_PARAMS = 6, 7, 8, 9
MFL_NEIGHBOUR_COUNT = 10

PARAM_FIELDS = (
    "deltaTime", "particleRadius", "cellSize", "mass", "viscosity", "pressure", "density", "spacing",
    "stiffness", "viscosityConstant", "pressureConstant", "kernelConstant", "velocitylimit",
)

# every symbol include/mfl.h declares
EXPORTS = (
    "mfl_last_error", "mfl_clear_error", "mfl_abi_version", "mfl_set_device", "mfl_device_finish", "mfl_memcpy",
    "mfl_device_alloc", "mfl_device_free",
    "mfl_device_alloc", "mfl_device_free",
    "mfl_params_default", "mfl_params_derive", "mfl_sqrt_cut",
    "mfl_radixsort_create", "mfl_radixsort_destroy", "mfl_radixsort_compile", "mfl_radixsort_compute",
    "mfl_radixsort_compute_noalloc",
    "mfl_fluid3d_create", "mfl_fluid3d_destroy", "mfl_fluid3d_compile", "mfl_fluid3d_compute",
    "mfl_fluid3d_set_particle_count", "mfl_fluid3d_get_particle_count", "mfl_fluid3d_set_buffer",
    "mfl_fluid3d_get_buffer", "mfl_fluid3d_write_buffer", "mfl_fluid3d_read_buffer", "mfl_fluid3d_params",
    "mfl_fluid3d_set_box", "mfl_fluid3d_set_gravity", "mfl_fluid3d_set_bucket_clear",
    "mfl_fluid3d_set_use_graph", "mfl_fluid3d_set_record_counts", "mfl_fluid3d_set_traversal",
    "mfl_fluid3d_get_traversal",
    "mfl_fluid3d_set_profiling",
    "mfl_fluid3d_get_tuning", "mfl_fluid3d_set_tuning", "mfl_fluid3d_autotune",
    "mfl_fluid3d_stage_time",
    "mfl_fluid3d_set_integrate_callback", "mfl_fluid3d_enqueue_kernel",
    "mfl_batch_create", "mfl_batch_destroy", "mfl_batch_scene_range", "mfl_batch_set_params",
    "mfl_batch_get_params", "mfl_batch_set_box", "mfl_batch_set_gravity", "mfl_batch_set_bucket_clear",
    "mfl_batch_set_use_graph", "mfl_batch_get_buffer", "mfl_batch_write_buffer", "mfl_batch_read_buffer",
    "mfl_batch_compute", "mfl_batch_get_tuning", "mfl_batch_set_tuning", "mfl_batch_autotune",
    "mfl_adjoint_create", "mfl_adjoint_destroy", "mfl_adjoint_forward", "mfl_adjoint_backward",
    "mfl_adjoint_get_buffer", "mfl_adjoint_param_grads",
    "mfl_bvh_trimesh_create", "mfl_bvh_trimesh_destroy", "mfl_bvh_trimesh_compile", "mfl_bvh_trimesh_compute",
    "mfl_bvh_trimesh_get_node_buffer", "mfl_bvh_trimesh_get_node_count", "mfl_bvh_trimesh_get_root_node",
    "mfl_bvh_trimesh_get_morton", "mfl_collide_bvh", "mfl_fluid3d_collide_bvh",
    "mfl_next_params_default", "mfl_fluid3dnext_create", "mfl_fluid3dnext_destroy", "mfl_fluid3dnext_compile",
    "mfl_fluid3dnext_set_particle_count", "mfl_fluid3dnext_get_particle_count", "mfl_fluid3dnext_params",
    "mfl_fluid3dnext_set_buffer", "mfl_fluid3dnext_get_buffer", "mfl_fluid3dnext_write_buffer",
    "mfl_fluid3dnext_read_buffer", "mfl_fluid3dnext_compute",
)
MFL_NEXT_POSITION, MFL_NEXT_SORTED_STATE, MFL_NEXT_INDEX, MFL_NEXT_CELL = 0, 1, 2, 3
MFL_NEXT_CELL_START, MFL_NEXT_CELL_END, MFL_NEXT_PRESSURE, MFL_NEXT_FORCE, MFL_NEXT_PARAMS = 4, 5, 6, 7, 8
NEXT_FIELDS = ("deltaTime", "particleRadius", "density", "stiffness0", "stiffness1", "viscosity0", "viscosity1",
               "radius", "cellSize", "cellCountX", "cellCountY")
MFL_ADJ_POSITION_GRAD, MFL_ADJ_VELOCITY_GRAD = 0, 1

STAGES = ("hash", "sort", "reorder", "bounds", "density", "forces", "order")
TRAVERSALS = ("sorted", "caller", "morton")


class NextParams(ctypes.Structure):
    """mfl_next_params == System of libCL/phys/oclFluid3Dnext.cl:15-30 (9 x fp32, 2 x int32)."""
    _fields_ = [(name, ctypes.c_float) for name in NEXT_FIELDS[:9]] + [("cellCountX", ctypes.c_int32),
                                                                        ("cellCountY", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k in NEXT_FIELDS}


class Params(ctypes.Structure):
    """mfl_params == libCL/phys/oclFluid3D.h:48-63 (13 x fp32)."""
    _fields_ = [(name, ctypes.c_float) for name in PARAM_FIELDS]

    def as_dict(self) -> dict:
        return {k: float(getattr(self, k)) for k in PARAM_FIELDS}

    def values(self) -> list:
        return [float(getattr(self, k)) for k in PARAM_FIELDS]


PASSES = ("list", "tile", "tile_list", "tile_kept")


class Tuning(ctypes.Structure):
    """mfl_tuning: launch shape of the neighbour passes and the sort."""
    _fields_ = [("neighbour_pass", ctypes.c_int32), ("density_lanes", ctypes.c_int32),
                ("tile_depth", ctypes.c_uint32), ("block_waves", ctypes.c_uint32), ("sort_items", ctypes.c_uint32),
                ("force_chunks", ctypes.c_uint32)]

    def as_dict(self) -> dict:
        return {"neighbour_pass": PASSES[self.neighbour_pass], "density_lanes": int(self.density_lanes),
                "tile_depth": int(self.tile_depth), "block_waves": int(self.block_waves),
                "sort_items": int(self.sort_items), "force_chunks": int(self.force_chunks)}


INTEGRATE_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)

_lib = None


def load() -> ctypes.CDLL:
    """Loads libmfl.so (built by libcl_amd.build); raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m libcl_amd.build` "
                           "(there is no CPU fallback)")
    # torch wheels bundle their own libamdhip64.so. Loading torch first makes libmfl.so bind to that
    # same HIP runtime, so device pointers and streams from torch tensors are valid in libmfl; the
    # other order leaves two HIP runtimes in the process and torch's device init fails.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, sz, u32p = ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint32)
    pp = ctypes.POINTER(Params)
    sig = {
        "mfl_last_error": (ctypes.c_char_p, []),
        "mfl_clear_error": (None, []),
        "mfl_abi_version": (ctypes.c_int, []),
        "mfl_set_device": (ctypes.c_int, [ctypes.c_int]),
        "mfl_device_finish": (ctypes.c_int, []),
        "mfl_memcpy": (ctypes.c_int, [vp, vp, sz]),
        "mfl_device_alloc": (vp, [sz]),
        "mfl_device_free": (None, [vp]),
        "mfl_params_default": (None, [pp]),
        "mfl_sqrt_cut": (ctypes.c_float, [ctypes.c_float]),
        "mfl_params_derive": (None, [pp]),
        "mfl_radixsort_create": (vp, []),
        "mfl_radixsort_destroy": (None, [vp]),
        "mfl_radixsort_compile": (ctypes.c_int, [vp]),
        "mfl_radixsort_compute": (ctypes.c_int, [vp, vp, vp, sz, ctypes.c_int, ctypes.c_int, vp]),
        "mfl_radixsort_compute_noalloc": (ctypes.c_int, [vp, vp, vp, sz, ctypes.c_int, ctypes.c_int, vp]),
        "mfl_fluid3d_create": (vp, []),
        "mfl_fluid3d_destroy": (None, [vp]),
        "mfl_fluid3d_compile": (ctypes.c_int, [vp]),
        "mfl_fluid3d_compute": (ctypes.c_int, [vp, vp]),
        "mfl_fluid3d_set_particle_count": (ctypes.c_int, [vp, sz]),
        "mfl_fluid3d_get_particle_count": (sz, [vp]),
        "mfl_fluid3d_set_buffer": (ctypes.c_int, [vp, ctypes.c_int, vp, sz]),
        "mfl_fluid3d_get_buffer": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(vp), ctypes.POINTER(sz)]),
        "mfl_fluid3d_write_buffer": (ctypes.c_int, [vp, ctypes.c_int, vp, sz]),
        "mfl_fluid3d_read_buffer": (ctypes.c_int, [vp, ctypes.c_int, vp, sz]),
        "mfl_fluid3d_params": (pp, [vp]),
        "mfl_fluid3d_set_box": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
        "mfl_fluid3d_set_gravity": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_float)]),
        "mfl_fluid3d_set_bucket_clear": (ctypes.c_int, [vp, ctypes.c_uint32]),
        "mfl_fluid3d_set_use_graph": (ctypes.c_int, [vp, ctypes.c_int]),
        "mfl_fluid3d_set_record_counts": (ctypes.c_int, [vp, ctypes.c_int]),
        "mfl_fluid3d_set_traversal": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_uint32]),
        "mfl_fluid3d_get_traversal": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int),
                                                     ctypes.POINTER(ctypes.c_uint32)]),
        "mfl_fluid3d_set_profiling": (ctypes.c_int, [vp, ctypes.c_uint32]),
        "mfl_fluid3d_get_tuning": (ctypes.c_int, [vp, ctypes.POINTER(Tuning)]),
        "mfl_fluid3d_set_tuning": (ctypes.c_int, [vp, ctypes.POINTER(Tuning)]),
        "mfl_fluid3d_autotune": (ctypes.c_int, [vp, vp, ctypes.c_uint32, ctypes.POINTER(Tuning),
                                                ctypes.POINTER(ctypes.c_double)]),
        "mfl_fluid3d_stage_time": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                                  ctypes.POINTER(ctypes.c_uint64), ctypes.c_int]),
        "mfl_fluid3d_set_integrate_callback": (ctypes.c_int, [vp, INTEGRATE_FN, vp]),
        "mfl_fluid3d_enqueue_kernel": (ctypes.c_int, [vp, ctypes.c_char_p, vp]),
        "mfl_batch_create": (vp, [ctypes.c_uint32, u32p]),
        "mfl_batch_destroy": (None, [vp]),
        "mfl_batch_scene_range": (ctypes.c_int, [vp, ctypes.c_uint32, u32p, u32p]),
        "mfl_batch_set_params": (ctypes.c_int, [vp, ctypes.c_int, pp]),
        "mfl_batch_get_params": (ctypes.c_int, [vp, ctypes.c_uint32, pp]),
        "mfl_batch_set_box": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                                             ctypes.POINTER(ctypes.c_float)]),
        "mfl_batch_set_gravity": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_float)]),
        "mfl_batch_set_bucket_clear": (ctypes.c_int, [vp, ctypes.c_uint32]),
        "mfl_batch_set_use_graph": (ctypes.c_int, [vp, ctypes.c_int]),
        "mfl_batch_get_buffer": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(vp), ctypes.POINTER(sz)]),
        "mfl_batch_write_buffer": (ctypes.c_int, [vp, ctypes.c_int, vp, sz]),
        "mfl_batch_read_buffer": (ctypes.c_int, [vp, ctypes.c_int, vp, sz]),
        "mfl_batch_compute": (ctypes.c_int, [vp, vp]),
        "mfl_batch_get_tuning": (ctypes.c_int, [vp, ctypes.POINTER(Tuning)]),
        "mfl_batch_set_tuning": (ctypes.c_int, [vp, ctypes.POINTER(Tuning)]),
        "mfl_batch_autotune": (ctypes.c_int, [vp, vp, ctypes.c_uint32, ctypes.POINTER(Tuning),
                                              ctypes.POINTER(ctypes.c_double)]),
        "mfl_adjoint_create": (vp, [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]),
        "mfl_adjoint_destroy": (None, [vp]),
        "mfl_adjoint_forward": (ctypes.c_int, [vp, vp]),
        "mfl_adjoint_backward": (ctypes.c_int, [vp, vp, vp, vp]),
        "mfl_adjoint_get_buffer": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(vp), ctypes.POINTER(sz)]),
        "mfl_adjoint_param_grads": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_double)]),
        "mfl_bvh_trimesh_create": (vp, []),
        "mfl_bvh_trimesh_destroy": (None, [vp]),
        "mfl_bvh_trimesh_compile": (ctypes.c_int, [vp]),
        "mfl_bvh_trimesh_compute": (ctypes.c_int, [vp, vp, sz, vp, sz, vp]),
        "mfl_bvh_trimesh_get_node_buffer": (ctypes.c_int, [vp, ctypes.POINTER(vp), ctypes.POINTER(sz)]),
        "mfl_bvh_trimesh_get_node_count": (ctypes.c_uint32, [vp]),
        "mfl_bvh_trimesh_get_root_node": (ctypes.c_uint32, [vp]),
        "mfl_bvh_trimesh_get_morton": (ctypes.c_int, [vp, ctypes.POINTER(vp), ctypes.POINTER(vp),
                                                      ctypes.POINTER(sz)]),
        "mfl_collide_bvh": (ctypes.c_int, [vp, vp, vp, sz, pp, vp, vp, sz, vp, sz, vp]),
        "mfl_fluid3d_collide_bvh": (ctypes.c_int, [vp, vp, vp, sz, vp, sz, vp]),
        "mfl_next_params_default": (None, [ctypes.POINTER(NextParams)]),
        "mfl_fluid3dnext_create": (vp, []),
        "mfl_fluid3dnext_destroy": (None, [vp]),
        "mfl_fluid3dnext_compile": (ctypes.c_int, [vp]),
        "mfl_fluid3dnext_set_particle_count": (ctypes.c_int, [vp, sz]),
        "mfl_fluid3dnext_get_particle_count": (sz, [vp]),
        "mfl_fluid3dnext_params": (ctypes.POINTER(NextParams), [vp]),
        "mfl_fluid3dnext_set_buffer": (ctypes.c_int, [vp, ctypes.c_int, vp, sz]),
        "mfl_fluid3dnext_get_buffer": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(vp), ctypes.POINTER(sz)]),
        "mfl_fluid3dnext_write_buffer": (ctypes.c_int, [vp, ctypes.c_int, vp, sz]),
        "mfl_fluid3dnext_read_buffer": (ctypes.c_int, [vp, ctypes.c_int, vp, sz]),
        "mfl_fluid3dnext_compute": (ctypes.c_int, [vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def last_error() -> str:
    return load().mfl_last_error().decode()


def default_params() -> Params:
    p = Params()
    load().mfl_params_default(ctypes.byref(p))
    return p


def derive(p: Params) -> Params:
    q = Params(*p.values())
    load().mfl_params_derive(ctypes.byref(q))
    return q


def params_from(values: dict) -> Params:
    """Params from the base fields in `values`; the three kernel constants are derived from cellSize
    and viscosity unless present in `values`."""
    p = default_params()
    for k, v in values.items():
        setattr(p, k, float(v))
    if not {"viscosityConstant", "pressureConstant", "kernelConstant"} & set(values):
        p = derive(p)
    return p
