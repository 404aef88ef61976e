"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, where it is the checker and the timed CPU baseline. The product (libcl_amd) never imports it.

The C side restates libCL/phys/oclFluid3D.cl:15-361, libCL/phys/oclFluid3D.cpp:213-342,
libCL/sort/oclRadixSort.{cl,cpp} (oracle/sph_oracle.c), and libCL/geom/oclBvhTrimesh.{cl,cpp} with
clCollideBVH, libCL/phys/oclFluid3D.cl:365-550 (oracle/bvh_oracle.c); each C function cites its lines.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

BUCKETS = 262144
EMPTY = 0xFFFFFFFF
REF_CLEAR = 65536
SORT_MAX = 4194304
PARAM_FIELDS = (
    "deltaTime", "particleRadius", "cellSize", "mass", "viscosity", "pressure", "density", "spacing",
    "stiffness", "viscosityConstant", "pressureConstant", "kernelConstant", "velocitylimit",
)


class Params(ctypes.Structure):
    """libCL/phys/oclFluid3D.h:48-63 (13 x fp32)."""
    _fields_ = [(name, ctypes.c_float) for name in PARAM_FIELDS]

    def as_dict(self):
        return {k: float(getattr(self, k)) for k in PARAM_FIELDS}


NEXT_FIELDS = ("deltaTime", "particleRadius", "density", "stiffness0", "stiffness1", "viscosity0", "viscosity1",
               "radius", "cellSize", "cellCountX", "cellCountY")


class NextParams(ctypes.Structure):
    """System of libCL/phys/oclFluid3Dnext.cl:15-30 (9 x fp32, 2 x int32)."""
    _fields_ = [(name, ctypes.c_float) for name in NEXT_FIELDS[:9]] + [("cellCountX", ctypes.c_int32),
                                                                        ("cellCountY", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k in NEXT_FIELDS}


_f32p = ctypes.POINTER(ctypes.c_float)
_u32p = ctypes.POINTER(ctypes.c_uint32)


class _Fluid(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_uint32), ("params", Params),
        ("box_min", ctypes.c_float * 4), ("box_max", ctypes.c_float * 4), ("gravity", ctypes.c_float * 4),
        ("clear_count", ctypes.c_uint32),
        ("pos", _f32p), ("vel", _f32p), ("force", _f32p),
        ("cell", _u32p), ("index", _u32p), ("spos", _f32p), ("svel", _f32p),
        ("cell_start", _u32p), ("cell_end", _u32p), ("counts", _u32p),
    ]


def build(force: bool = False) -> str:
    """Compile oracle/liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_init_params.argtypes = [ctypes.POINTER(Params)]
        L.orc_derive_constants.argtypes = [ctypes.POINTER(Params)]
        L.orc_hash_key.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_hash_key.restype = ctypes.c_uint32
        L.orc_grid_position.argtypes = [_f32p, ctypes.c_float, ctypes.POINTER(ctypes.c_int)]
        L.orc_hash.argtypes = [_f32p, ctypes.c_uint32, ctypes.c_float, _u32p, _u32p]
        for name in ("orc_radix_sort", "orc_radix_sort_unbounded"):
            fn = getattr(L, name)
            fn.argtypes = [_u32p, _u32p, ctypes.c_uint32, ctypes.c_int, ctypes.c_int]
            fn.restype = ctypes.c_int
        for name in ("orc_stable_sort", "orc_std_stable_sort"):
            fn = getattr(L, name)
            fn.argtypes = [_u32p, _u32p, ctypes.c_uint32, ctypes.c_int, ctypes.c_int]
            fn.restype = None
        L.orc_reorder.argtypes = [_u32p, _f32p, _f32p, _f32p, _f32p, ctypes.c_uint32]
        L.orc_clear_buckets.argtypes = [_u32p, _u32p, ctypes.c_uint32]
        L.orc_find_bounds.argtypes = [_u32p, ctypes.c_uint32, _u32p, _u32p]
        L.orc_density.argtypes = [_f32p, _f32p, _u32p, _u32p, ctypes.c_uint32, ctypes.POINTER(Params), _u32p]
        L.orc_forces.argtypes = [_f32p, _f32p, _f32p, _u32p, _u32p, _u32p, ctypes.c_uint32,
                                 ctypes.POINTER(Params)]
        L.orc_clip_box.argtypes = [_f32p, _f32p, _f32p, _u32p, ctypes.c_uint32, ctypes.POINTER(Params),
                                   _f32p, _f32p]
        L.orc_integrate_force.argtypes = [_f32p, _f32p, ctypes.c_uint32, ctypes.POINTER(Params)]
        L.orc_gravity.argtypes = [_f32p, ctypes.c_uint32, _f32p]
        L.orc_integrate_velocity.argtypes = [_f32p, _f32p, ctypes.c_uint32, ctypes.POINTER(Params)]
        L.orc_step.argtypes = [ctypes.POINTER(_Fluid), ctypes.c_int]
        L.orc_step.restype = ctypes.c_int
        L.orc_steps.argtypes = [ctypes.POINTER(_Fluid), ctypes.c_int]
        L.orc_steps.restype = ctypes.c_int
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_threads.restype = ctypes.c_int
        NP = ctypes.POINTER(NextParams)
        L.orc_next_init_params.argtypes = [NP]
        L.orc_next_w.argtypes = [ctypes.c_float, ctypes.c_float]
        L.orc_next_w.restype = ctypes.c_float
        L.orc_next_hash.argtypes = [_f32p, ctypes.c_uint32, NP, _u32p, _u32p]
        L.orc_next_reorder.argtypes = [_u32p, _f32p, _f32p, ctypes.c_uint32]
        L.orc_next_bounds.argtypes = [_u32p, ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p]
        L.orc_next_pressure.argtypes = [_f32p, _f32p, _u32p, _u32p, ctypes.c_uint32, NP]
        L.orc_next_forces.argtypes = [_f32p, _f32p, _f32p, _u32p, _u32p, ctypes.c_uint32, NP]
        L.orc_next_integrate.argtypes = [_f32p, _f32p, _u32p, ctypes.c_uint32, NP]
        L.orc_next_step.argtypes = [_f32p, ctypes.c_uint32, NP, _u32p, _u32p, _f32p, _u32p, _u32p, _f32p, _f32p]
        L.orc_morton_code.argtypes = [ctypes.c_uint32]
        L.orc_morton_code.restype = ctypes.c_uint32
        L.orc_bvh_aabb.argtypes = [_f32p, ctypes.c_uint32, _f32p, _f32p]
        L.orc_bvh_morton.argtypes = [_u32p, _f32p, ctypes.c_uint32, _f32p, _f32p, _u32p, _u32p]
        L.orc_bvh_build.argtypes = [_f32p, ctypes.c_uint32, _u32p, ctypes.c_uint32, ctypes.c_void_p, _u32p, _u32p,
                                    _u32p]
        L.orc_bvh_build.restype = ctypes.c_int
        L.orc_collide_bvh.argtypes = [_f32p, _f32p, _f32p, ctypes.c_uint32, ctypes.POINTER(Params), ctypes.c_void_p,
                                      ctypes.c_uint32, _u32p, _f32p, _u32p, _u32p]
        _lib = L
    return _lib


def _f(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_f32p)


def _u(a):
    assert a.dtype == np.uint32 and a.flags.c_contiguous
    return a.ctypes.data_as(_u32p)


def set_threads(n: int) -> int:
    return lib().orc_set_threads(n)


def default_params() -> Params:
    p = Params()
    lib().orc_init_params(ctypes.byref(p))
    return p


def make_params(**kw) -> Params:
    """Params with the given fields; the three derived constants are recomputed from cellSize and
    viscosity unless given explicitly."""
    p = default_params()
    for k, v in kw.items():
        setattr(p, k, v)
    derived = {"viscosityConstant", "pressureConstant", "kernelConstant"}
    if not derived & set(kw):
        lib().orc_derive_constants(ctypes.byref(p))
    return p


def derive_constants(p: Params) -> Params:
    q = Params(*[getattr(p, k) for k in PARAM_FIELDS])
    lib().orc_derive_constants(ctypes.byref(q))
    return q


def hash_key(gx: int, gy: int, gz: int) -> int:
    return int(lib().orc_hash_key(gx, gy, gz))


def hash_cells(pos: np.ndarray, cell_size: float):
    n = pos.shape[0]
    cell = np.empty(n, np.uint32)
    index = np.empty(n, np.uint32)
    lib().orc_hash(_f(pos), n, cell_size, _u(cell), _u(index))
    return cell, index


def radix_sort(key: np.ndarray, val: np.ndarray, start_bit: int, end_bit: int, bounded: bool = True):
    """In-place restated reference sort; returns True on success (False = reference rejection)."""
    fn = lib().orc_radix_sort if bounded else lib().orc_radix_sort_unbounded
    return bool(fn(_u(key), _u(val), key.shape[0], start_bit, end_bit))


def merge_sort(key, val, start_bit, end_bit):
    lib().orc_stable_sort(_u(key), _u(val), key.shape[0], start_bit, end_bit)


def std_stable_sort(key, val, start_bit, end_bit):
    lib().orc_std_stable_sort(_u(key), _u(val), key.shape[0], start_bit, end_bit)


class Fluid:
    """CPU state of one oclFluid3D program (user buffers + internal buffers + bucket tables)."""

    def __init__(self, pos, vel, force, params: Params, box_min=(-1, -1, -1, 0), box_max=(1, 1, 1, 0),
                 gravity=(0, 0, 1, 0), clear_count=REF_CLEAR, cell_start=None, cell_end=None):
        n = pos.shape[0]
        self.pos = np.ascontiguousarray(pos, np.float32).copy()
        self.vel = np.ascontiguousarray(vel, np.float32).copy()
        self.force = np.ascontiguousarray(force, np.float32).copy()
        self.cell = np.zeros(n, np.uint32)
        self.index = np.zeros(n, np.uint32)
        self.spos = np.zeros((n, 4), np.float32)
        self.svel = np.zeros((n, 4), np.float32)
        self.cell_start = (np.full(BUCKETS, EMPTY, np.uint32) if cell_start is None
                           else np.ascontiguousarray(cell_start, np.uint32).copy())
        self.cell_end = (np.full(BUCKETS, EMPTY, np.uint32) if cell_end is None
                         else np.ascontiguousarray(cell_end, np.uint32).copy())
        self.counts = np.zeros(n, np.uint32)  # accepted visits per sorted particle (last density pass)
        s = _Fluid()
        s.n = n
        s.params = params
        s.box_min[:] = [float(x) for x in box_min]
        s.box_max[:] = [float(x) for x in box_max]
        s.gravity[:] = [float(x) for x in gravity]
        s.clear_count = clear_count
        s.pos, s.vel, s.force = _f(self.pos), _f(self.vel), _f(self.force)
        s.cell, s.index = _u(self.cell), _u(self.index)
        s.spos, s.svel = _f(self.spos), _f(self.svel)
        s.cell_start, s.cell_end = _u(self.cell_start), _u(self.cell_end)
        s.counts = _u(self.counts)
        self._s = s

    @property
    def params(self) -> Params:
        return self._s.params

    def step(self, integrate: bool = True) -> None:
        if not lib().orc_step(ctypes.byref(self._s), 1 if integrate else 0):
            raise RuntimeError("oracle step rejected by the restated sort")

    def steps(self, k: int) -> None:
        if not lib().orc_steps(ctypes.byref(self._s), k):
            raise RuntimeError("oracle steps rejected by the restated sort")

    def set_state(self, pos, vel, force):
        self.pos[...] = pos
        self.vel[...] = vel
        self.force[...] = force


# ---- single-kernel entry points used by the unit tests --------------------------------------------

def reorder(index, pos, vel):
    n = index.shape[0]
    spos = np.zeros((n, 4), np.float32)
    svel = np.zeros((n, 4), np.float32)
    lib().orc_reorder(_u(index), _f(pos), _f(vel), _f(spos), _f(svel), n)
    return spos, svel


def clear_buckets(cell_start, cell_end, clear_count):
    lib().orc_clear_buckets(_u(cell_start), _u(cell_end), clear_count)


def find_bounds(cell, cell_start, cell_end):
    lib().orc_find_bounds(_u(cell), cell.shape[0], _u(cell_start), _u(cell_end))


def density(spos, svel, cell_start, cell_end, params: Params):
    """Writes density / pressure into the w lanes of spos / svel; returns the accepted visits per particle."""
    counts = np.zeros(spos.shape[0], np.uint32)
    lib().orc_density(_f(spos), _f(svel), _u(cell_start), _u(cell_end), spos.shape[0], ctypes.byref(params),
                      _u(counts))
    return counts


def forces(spos, svel, force, index, cell_start, cell_end, params: Params):
    lib().orc_forces(_f(spos), _f(svel), _f(force), _u(index), _u(cell_start), _u(cell_end), spos.shape[0],
                     ctypes.byref(params))


def clip_box(spos, svel, force, index, params: Params, box_min, box_max):
    mn = np.asarray(box_min, np.float32)
    mx = np.asarray(box_max, np.float32)
    lib().orc_clip_box(_f(spos), _f(svel), _f(force), _u(index), spos.shape[0], ctypes.byref(params), _f(mn),
                       _f(mx))


def integrate_force(vel, force, params: Params):
    lib().orc_integrate_force(_f(vel), _f(force), vel.shape[0], ctypes.byref(params))


def gravity(force, g=(0, 0, 1, 0)):
    gv = np.asarray(g, np.float32)
    lib().orc_gravity(_f(force), force.shape[0], _f(gv))


def integrate_velocity(pos, vel, params: Params):
    lib().orc_integrate_velocity(_f(pos), _f(vel), pos.shape[0], ctypes.byref(params))


# ---- triangle-mesh BVH and particle collision (oracle/bvh_oracle.c) --------------------------------

# BVHNode, libCL/geom/oclBvhTrimesh.h:33-41 (48 bytes)
BVH_NODE = np.dtype([("bbMin", "<f4", (4,)), ("bbMax", "<f4", (4,)), ("left", "<i4"), ("right", "<i4"),
                     ("bit", "<u4"), ("trav", "<u4")])
assert BVH_NODE.itemsize == 48


def morton_code(x: int) -> int:
    return int(lib().orc_morton_code(x))


def bvh_aabb(vertex):
    mn = np.zeros(4, np.float32)
    mx = np.zeros(4, np.float32)
    lib().orc_bvh_aabb(_f(vertex), vertex.shape[0], _f(mn), _f(mx))
    return mn, mx


def bvh_morton(index, vertex, mn, mx):
    t = index.shape[0] // 3
    key = np.zeros(t, np.uint32)
    val = np.zeros(t, np.uint32)
    lib().orc_bvh_morton(_u(index), _f(vertex), t, _f(np.asarray(mn, np.float32)), _f(np.asarray(mx, np.float32)),
                         _u(key), _u(val))
    return key, val


def bvh_build(vertex, index):
    """oclBvhTrimesh::compute: (nodes [2T-1] BVH_NODE, root, sorted keys, sorted triangle indices), or None
    when the build is rejected (fewer than 2 triangles, an index out of range)."""
    vertex = np.ascontiguousarray(vertex, np.float32)
    index = np.ascontiguousarray(index, np.uint32)
    t = index.shape[0] // 3
    if t < 2:
        return None
    nodes = np.zeros(2 * t - 1, BVH_NODE)
    root = np.zeros(1, np.uint32)
    key = np.zeros(t, np.uint32)
    val = np.zeros(t, np.uint32)
    if not lib().orc_bvh_build(_f(vertex), vertex.shape[0], _u(index), index.shape[0], nodes.ctypes.data, _u(root),
                               _u(key), _u(val)):
        return None
    return nodes, int(root[0]), key, val


def collide_bvh(pos, vel, force, params: Params, nodes, root, index, vertex, filled=None):
    """clCollideBVH over copies of pos / vel / force; returns them and the leaf that stopped each walk
    (0xFFFFFFFF where none did). `filled` (uint32 [n], optional) receives 1 where the ring filled up."""
    pos = np.ascontiguousarray(pos, np.float32).copy()
    vel = np.ascontiguousarray(vel, np.float32).copy()
    force = np.ascontiguousarray(force, np.float32).copy()
    nodes = np.ascontiguousarray(nodes, BVH_NODE)
    index = np.ascontiguousarray(index, np.uint32)
    vertex = np.ascontiguousarray(vertex, np.float32)
    hit = np.zeros(pos.shape[0], np.uint32)
    lib().orc_collide_bvh(_f(force), _f(pos), _f(vel), pos.shape[0], ctypes.byref(params), nodes.ctypes.data, root,
                          _u(index), _f(vertex), _u(hit), None if filled is None else _u(filled))
    return pos, vel, force, hit


# ---- oclFluid3Dnext (oracle/next_oracle.c) -------------------------------------------------------------

def next_default_params() -> NextParams:
    p = NextParams()
    lib().orc_next_init_params(ctypes.byref(p))
    return p


def next_w(h: float, q: float) -> float:
    return float(lib().orc_next_w(h, q))


class Next:
    """CPU state of one oclFluid3Dnext program: state (x, y, vx, vy) in caller order plus its scratch."""

    def __init__(self, state, params: NextParams):
        n = state.shape[0]
        self.params = params
        self.state = np.ascontiguousarray(state, np.float32).copy()
        self.cell = np.zeros(n, np.uint32)
        self.index = np.zeros(n, np.uint32)
        self.sorted = np.zeros((n, 4), np.float32)
        nc = int(params.cellCountX) * int(params.cellCountY)
        self.start = np.zeros(nc, np.uint32)
        self.end = np.zeros(nc, np.uint32)
        self.pressure = np.zeros((n, 2), np.float32)
        self.force = np.zeros((n, 2), np.float32)

    def step(self):
        lib().orc_next_step(_f(self.state), self.state.shape[0], ctypes.byref(self.params), _u(self.cell),
                            _u(self.index), _f(self.sorted), _u(self.start), _u(self.end), _f(self.pressure),
                            _f(self.force))

    def hash(self):
        lib().orc_next_hash(_f(self.state), self.state.shape[0], ctypes.byref(self.params), _u(self.cell),
                            _u(self.index))

    def bounds(self):
        lib().orc_next_bounds(_u(self.cell), self.cell.shape[0], self.start.shape[0], _u(self.start), _u(self.end))

    def pressure_pass(self):
        lib().orc_next_pressure(_f(self.sorted), _f(self.pressure), _u(self.start), _u(self.end),
                                self.sorted.shape[0], ctypes.byref(self.params))

    def force_pass(self):
        lib().orc_next_forces(_f(self.sorted), _f(self.pressure), _f(self.force), _u(self.start), _u(self.end),
                              self.sorted.shape[0], ctypes.byref(self.params))
