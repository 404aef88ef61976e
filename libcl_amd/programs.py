"""Python mirror of libCL's program objects on the HIP path: oclRadixSort, oclFluid3D, oclBvhTrimesh.

Names, argument meaning and the int (1 ok / 0 fail) return convention follow
libCL/sort/oclRadixSort.h:26-28, libCL/phys/oclFluid3D.h:25-74 and libCL/geom/oclBvhTrimesh.h:21-82; failures leave their text in
last_error() (the reference logs through Log(ERR), libCL/oclCommon.cpp:146-162). Device memory is
torch tensors (plumbing only); every call goes through libmfl.so's C ABI.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N


def _stream_ptr(stream) -> int | None:
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream  # torch.cuda.Stream


class DeviceBuffer:
    """A program buffer (oclBuffer, libCL/oclBuffer.h): device pointer + size, with blocking
    read()/write() of host numpy arrays (oclBuffer::read/write, libCL/oclBuffer.cpp:47-83)."""

    def __init__(self, program: "oclFluid3D", kind: int, name: str):
        self._program = program
        self.kind = kind
        self.name = name

    def _info(self):
        ptr, nbytes = ctypes.c_void_p(), ctypes.c_size_t()
        if not N.load().mfl_fluid3d_get_buffer(self._program._h, self.kind, ctypes.byref(ptr), ctypes.byref(nbytes)):
            raise RuntimeError(N.last_error())
        return ptr.value or 0, nbytes.value

    @property
    def ptr(self) -> int:
        return self._info()[0]

    @property
    def nbytes(self) -> int:
        return self._info()[1]

    def read(self, dtype=np.float32, shape=None) -> np.ndarray:
        nbytes = self.nbytes
        out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype)
        if not N.load().mfl_fluid3d_read_buffer(self._program._h, self.kind, out.ctypes.data, out.nbytes):
            raise RuntimeError(N.last_error())
        return out.reshape(shape) if shape is not None else out

    def write(self, host: np.ndarray) -> int:
        host = np.ascontiguousarray(host)
        ok = N.load().mfl_fluid3d_write_buffer(self._program._h, self.kind, host.ctypes.data, host.nbytes)
        return int(ok)


class oclRadixSort:
    """libCL/sort/oclRadixSort.h:26-28 over libmfl's gfx950 sort."""

    cBits = 4

    def __init__(self):
        L = N.load()
        self._h = L.mfl_radixsort_create()
        if not self._h:
            raise RuntimeError(N.last_error())

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and N is not None and N._lib is not None:
            N._lib.mfl_radixsort_destroy(h)
            self._h = None

    def compile(self) -> int:
        return int(N.load().mfl_radixsort_compile(self._h))

    def compute(self, key, val, start_bit: int, end_bit: int, stream=None) -> int:
        """Stable in-place sort of the uint32 (key, val) device tensors by bits [start_bit, end_bit)
        (libCL/sort/oclRadixSort.cpp:62-144). Rejects tensors of different sizes (cpp:64-68)."""
        import torch
        if key.numel() != val.numel():
            N.load().mfl_clear_error()
            self.error = f"key and value arrays are of different size ( {key.numel()},{val.numel()})"
            return 0
        for t in (key, val):
            if t.dtype not in (torch.int32, torch.uint32) or not t.is_contiguous() or not t.is_cuda:
                raise TypeError("key/val must be contiguous 32-bit integer device tensors")
        return int(N.load().mfl_radixsort_compute(self._h, key.data_ptr(), val.data_ptr(), key.numel(), start_bit,
                                                  end_bit, _stream_ptr(stream)))


class oclFluid3D:
    """libCL/phys/oclFluid3D.h:25-74 over libmfl's gfx950 SPH step."""

    cLocalSize = 256
    cBucketCount = 262144
    EVT_INTEGRATE = "OnIntegrate"

    def __init__(self):
        L = N.load()
        self._h = L.mfl_fluid3d_create()
        if not self._h:
            raise RuntimeError(N.last_error())
        self._user = {}
        self._cb = None
        self.bfPosition = DeviceBuffer(self, N.MFL_POSITION, "bfPosition")
        self.bfVelocity = DeviceBuffer(self, N.MFL_VELOCITY, "bfVelocity")
        self.bfForce = DeviceBuffer(self, N.MFL_FORCE, "bfForce")
        self.bfSortedPosition = DeviceBuffer(self, N.MFL_SORTED_POSITION, "bfSortedPosition")
        self.bfSortedVelocity = DeviceBuffer(self, N.MFL_SORTED_VELOCITY, "bfSortedVelocity")
        self.bfIndex = DeviceBuffer(self, N.MFL_INDEX, "bfIndex")
        self.bfCell = DeviceBuffer(self, N.MFL_CELL, "bfCell")
        self.bfCellStart = DeviceBuffer(self, N.MFL_CELL_START, "bfCellStart")
        self.bfCellEnd = DeviceBuffer(self, N.MFL_CELL_END, "bfCellEnd")
        self.bfParams = DeviceBuffer(self, N.MFL_PARAMS, "bfParams")
        self.bfNeighbourCount = DeviceBuffer(self, N.MFL_NEIGHBOUR_COUNT, "bfNeighbourCount")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and N is not None and N._lib is not None:
            N._lib.mfl_fluid3d_destroy(h)
            self._h = None

    # ---- program interface (oclFluid3D.h:28-46) ----
    def compile(self) -> int:
        return int(N.load().mfl_fluid3d_compile(self._h))

    def compute(self, stream=None) -> int:
        return int(N.load().mfl_fluid3d_compute(self._h, _stream_ptr(stream)))

    def setParticleCount(self, n: int) -> int:
        return int(N.load().mfl_fluid3d_set_particle_count(self._h, n))

    def getParticleCount(self) -> int:
        return int(N.load().mfl_fluid3d_get_particle_count(self._h))

    def _set(self, kind: int, tensor) -> int:
        import torch
        if tensor.dtype != torch.float32 or not tensor.is_contiguous() or not tensor.is_cuda:
            raise TypeError("buffers must be contiguous float32 device tensors of shape [N, 4]")
        ok = int(N.load().mfl_fluid3d_set_buffer(self._h, kind, tensor.data_ptr(), tensor.numel() * 4))
        if ok:
            self._user[kind] = tensor  # the caller keeps ownership; keep it alive while bound
        return ok

    def setPositionBuffer(self, tensor) -> int:
        return self._set(N.MFL_POSITION, tensor)

    def setVelocityBuffer(self, tensor) -> int:
        return self._set(N.MFL_VELOCITY, tensor)

    def setForceBuffer(self, tensor) -> int:
        return self._set(N.MFL_FORCE, tensor)

    def getPositionBuffer(self) -> DeviceBuffer:
        return self.bfPosition

    def getVelocityBuffer(self) -> DeviceBuffer:
        return self.bfVelocity

    def getForceBuffer(self) -> DeviceBuffer:
        return self.bfForce

    def getSortedPositionBuffer(self) -> DeviceBuffer:
        return self.bfSortedPosition

    def getSortedVelocityBuffer(self) -> DeviceBuffer:
        return self.bfSortedVelocity

    def getIndexBuffer(self) -> DeviceBuffer:
        return self.bfIndex

    def getParamBuffer(self) -> DeviceBuffer:
        return self.bfParams

    def getParameters(self) -> N.Params:
        """Live host Params (oclFluid3D::getParameters, cpp:204-207); edits apply at the next compute."""
        return N.load().mfl_fluid3d_params(self._h).contents

    def setParameters(self, values) -> None:
        p = self.getParameters()
        vals = values.values() if isinstance(values, N.Params) else [values[k] for k in N.PARAM_FIELDS]
        for k, v in zip(N.PARAM_FIELDS, vals):
            setattr(p, k, float(v))

    # ---- additions ----
    def setBox(self, box_min, box_max) -> int:
        mn = (ctypes.c_float * 4)(*[float(x) for x in box_min])
        mx = (ctypes.c_float * 4)(*[float(x) for x in box_max])
        return int(N.load().mfl_fluid3d_set_box(self._h, mn, mx))

    def setGravity(self, g) -> int:
        gv = (ctypes.c_float * 4)(*[float(x) for x in g])
        return int(N.load().mfl_fluid3d_set_gravity(self._h, gv))

    def setBucketClear(self, count: int) -> int:
        return int(N.load().mfl_fluid3d_set_bucket_clear(self._h, count))

    def setUseGraph(self, enable: bool) -> int:
        return int(N.load().mfl_fluid3d_set_use_graph(self._h, 1 if enable else 0))

    def setRecordCounts(self, enable: bool) -> int:
        """While enabled, each density pass writes every particle's accepted-visit count (by sorted slot)
        to bfNeighbourCount."""
        return int(N.load().mfl_fluid3d_set_record_counts(self._h, 1 if enable else 0))

    def setTraversal(self, order: str, refresh_steps: int = 16) -> int:
        """Visit order of the neighbour passes: "sorted", "caller" or "morton" (results unchanged)."""
        return int(N.load().mfl_fluid3d_set_traversal(self._h, N.TRAVERSALS.index(order), refresh_steps))

    def getTraversal(self) -> tuple:
        """(order name, refresh period) of the neighbour passes' visit order."""
        order, refresh = ctypes.c_int(), ctypes.c_uint32()
        if not N.load().mfl_fluid3d_get_traversal(self._h, ctypes.byref(order), ctypes.byref(refresh)):
            raise RuntimeError(N.load().mfl_last_error().decode())
        return N.TRAVERSALS[order.value], refresh.value

    def getTuning(self) -> dict:
        """Launch shape of the neighbour passes and the sort (mfl_tuning as a dict), and the traversal
        order."""
        t = N.Tuning()
        if not N.load().mfl_fluid3d_get_tuning(self._h, ctypes.byref(t)):
            raise RuntimeError(N.load().mfl_last_error().decode())
        return dict(t.as_dict(), traversal=self.getTraversal()[0])

    def setTuning(self, **kw) -> int:
        """Changes fields of the tuning: neighbour_pass ("list" / "tile" / "tile_list" / "tile_kept"),
        density_lanes,
        tile_depth, block_waves, sort_items, force_chunks, and the traversal order (its refresh period
        kept). Results differ between shapes only in FP32 summation order."""
        if "traversal" in kw:
            kw = dict(kw)
            if not self.setTraversal(kw.pop("traversal"), self.getTraversal()[1]):
                return 0
        cur = self.getTuning()
        cur.pop("traversal")
        if "density_lanes" not in kw and "neighbour_pass" in kw:
            # keep the lane count valid for the pass being selected (list: 1 or 8, tiles: 2, 4 or 8)
            ok = (1, 8) if kw["neighbour_pass"] == "list" else (2, 4, 8)
            if cur["density_lanes"] not in ok:
                kw = dict(kw, density_lanes=8)
        cur.update(kw)
        t = N.Tuning(N.PASSES.index(cur["neighbour_pass"]), cur["density_lanes"], cur["tile_depth"],
                     cur["block_waves"], cur["sort_items"], cur["force_chunks"])
        return int(N.load().mfl_fluid3d_set_tuning(self._h, ctypes.byref(t)))

    def autotune(self, steps: int = 3, stream=None):
        """Times each candidate shape on the current state (left unchanged) and keeps the fastest;
        returns (tuning dict, ms per step) or None on failure."""
        t = N.Tuning()
        ms = ctypes.c_double()
        if not N.load().mfl_fluid3d_autotune(self._h, _stream_ptr(stream), steps, ctypes.byref(t), ctypes.byref(ms)):
            return None
        return dict(t.as_dict(), traversal=self.getTraversal()[0]), ms.value

    def setProfiling(self, stages) -> int:
        """Times the named stages (subset of _native.STAGES) with HIP events (oclKernel::profile)."""
        mask = 0
        for s in stages:
            mask |= 1 << N.STAGES.index(s)
        return int(N.load().mfl_fluid3d_set_profiling(self._h, mask))

    def stageTime(self, stage: str, reset: bool = False):
        """(total_ms, launches) accumulated for a profiled stage."""
        ms, cnt = ctypes.c_double(), ctypes.c_uint64()
        if not N.load().mfl_fluid3d_stage_time(self._h, N.STAGES.index(stage), ctypes.byref(ms), ctypes.byref(cnt),
                                                1 if reset else 0):
            raise RuntimeError(N.last_error())
        return ms.value, cnt.value

    # ---- events (oclFluid3D.cpp:287-291, oclProgram.h:92-93) ----
    def addEventHandler(self, name: str, fn) -> int:
        """fn(program, stream_ptr) is called by compute() after the force pass instead of
        clip-box + integration (EVT_INTEGRATE)."""
        if name != self.EVT_INTEGRATE:
            return 0

        def tramp(_h, stream, _user):
            try:
                fn(self, stream)
            except Exception:  # the reference ignores the handler's result as well
                return 0
            return 1

        self._cb = N.INTEGRATE_FN(tramp)
        return int(N.load().mfl_fluid3d_set_integrate_callback(self._h, self._cb, None))

    def removeEventHandler(self) -> int:
        self._cb = None
        return int(N.load().mfl_fluid3d_set_integrate_callback(self._h, N.INTEGRATE_FN(), None))

    def enqueueKernel(self, name: str, stream=None) -> int:
        """getKernel(name) + clEnqueueNDRangeKernel with the bound arguments."""
        return int(N.load().mfl_fluid3d_enqueue_kernel(self._h, name.encode(), _stream_ptr(stream)))

    def collideBVH(self, bvh: "oclBvhTrimesh", triangle, vertex, stream=None) -> int:
        """clCollideBVH (cl:498-550) on this program's position / velocity / force buffers and Params,
        against `bvh` built over (triangle indices, vertices): what an EVT_INTEGRATE handler enqueues."""
        _f4_tensor(vertex, "vertex")
        _u32_tensor(triangle, "triangle")
        return int(N.load().mfl_fluid3d_collide_bvh(self._h, bvh._h, triangle.data_ptr(), triangle.numel(),
                                                    vertex.data_ptr(), vertex.numel() // 4, _stream_ptr(stream)))

    # ---- host helpers ----
    def upload(self, pos: np.ndarray, vel: np.ndarray, force: np.ndarray) -> int:
        ok = self.bfPosition.write(np.asarray(pos, np.float32))
        ok &= self.bfVelocity.write(np.asarray(vel, np.float32))
        ok &= self.bfForce.write(np.asarray(force, np.float32))
        return int(ok)

    def state(self) -> dict:
        n = self.getParticleCount()
        return dict(pos=self.bfPosition.read(np.float32, (-1, 4))[:n], vel=self.bfVelocity.read(np.float32, (-1, 4))[:n],
                    force=self.bfForce.read(np.float32, (-1, 4))[:n])

    def internals(self) -> dict:
        n = self.getParticleCount()
        return dict(cell=self.bfCell.read(np.uint32)[:n], index=self.bfIndex.read(np.uint32)[:n],
                    spos=self.bfSortedPosition.read(np.float32, (-1, 4))[:n],
                    svel=self.bfSortedVelocity.read(np.float32, (-1, 4))[:n],
                    cell_start=self.bfCellStart.read(np.uint32), cell_end=self.bfCellEnd.read(np.uint32))


# BVHNode, libCL/geom/oclBvhTrimesh.h:33-41 (mfl_bvh_node, 48 bytes)
BVH_NODE = np.dtype([("bbMin", "<f4", (4,)), ("bbMax", "<f4", (4,)), ("left", "<i4"), ("right", "<i4"),
                     ("bit", "<u4"), ("trav", "<u4")])


def _u32_tensor(t, what: str):
    import torch
    if t.dtype not in (torch.int32, torch.uint32) or not t.is_contiguous() or not t.is_cuda:
        raise TypeError(f"{what} must be a contiguous 32-bit integer device tensor")


def _f4_tensor(t, what: str):
    import torch
    if t.dtype != torch.float32 or not t.is_contiguous() or not t.is_cuda or t.numel() % 4:
        raise TypeError(f"{what} must be a contiguous float32 device tensor of float4 rows")


class oclBvhTrimesh:
    """libCL/geom/oclBvhTrimesh.h:21-82 over libmfl's gfx950 BVH build."""

    def __init__(self):
        L = N.load()
        self._h = L.mfl_bvh_trimesh_create()
        if not self._h:
            raise RuntimeError(N.last_error())
        self._mesh = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and N is not None and N._lib is not None:
            N._lib.mfl_bvh_trimesh_destroy(h)
            self._h = None

    def compile(self) -> int:
        return int(N.load().mfl_bvh_trimesh_compile(self._h))

    def compute(self, vertex, index, stream=None) -> int:
        """oclBvhTrimesh::compute(device, bfVertex, bfIndex) (cpp:53-162): vertex is a [V, 4] float32
        device tensor, index a uint32 device tensor of 3 vertex indices per triangle."""
        _f4_tensor(vertex, "vertex")
        _u32_tensor(index, "index")
        ok = int(N.load().mfl_bvh_trimesh_compute(self._h, vertex.data_ptr(), vertex.numel() // 4, index.data_ptr(),
                                                  index.numel(), _stream_ptr(stream)))
        if ok:
            self._mesh = (vertex, index)
        return ok

    def getNodeBuffer(self):
        """(device pointer, bytes) of the mfl_bvh_node array (getNodeBuffer, cpp:169-172)."""
        ptr, nbytes = ctypes.c_void_p(), ctypes.c_size_t()
        if not N.load().mfl_bvh_trimesh_get_node_buffer(self._h, ctypes.byref(ptr), ctypes.byref(nbytes)):
            raise RuntimeError(N.last_error())
        return ptr.value or 0, nbytes.value

    def getNodeCount(self) -> int:
        return int(N.load().mfl_bvh_trimesh_get_node_count(self._h))

    def getRootNode(self) -> int:
        return int(N.load().mfl_bvh_trimesh_get_root_node(self._h))

    def nodes(self) -> np.ndarray:
        """Host copy of the nodes as a BVH_NODE structured array."""
        ptr, nbytes = self.getNodeBuffer()
        out = np.zeros(nbytes // BVH_NODE.itemsize, BVH_NODE)
        if nbytes and not N.load().mfl_memcpy(out.ctypes.data, ptr, nbytes):
            raise RuntimeError(N.last_error())
        return out

    def morton(self):
        """Host copies of the sorted Morton keys and triangle indices (bfMortonKey, bfMortonVal)."""
        k, v, cnt = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_size_t()
        if not N.load().mfl_bvh_trimesh_get_morton(self._h, ctypes.byref(k), ctypes.byref(v), ctypes.byref(cnt)):
            raise RuntimeError(N.last_error())
        keys = np.zeros(cnt.value, np.uint32)
        vals = np.zeros(cnt.value, np.uint32)
        if cnt.value:
            N.load().mfl_memcpy(keys.ctypes.data, k.value, keys.nbytes)
            N.load().mfl_memcpy(vals.ctypes.data, v.value, vals.nbytes)
        return keys, vals


def collide_bvh(force, position, velocity, params: N.Params, bvh: oclBvhTrimesh, triangle, vertex,
                stream=None) -> int:
    """clCollideBVH (libCL/phys/oclFluid3D.cl:498-550) on float4 device tensors force / position /
    velocity with `bvh`'s nodes and root over the mesh (triangle indices, vertices)."""
    for t, w in ((force, "force"), (position, "position"), (velocity, "velocity"), (vertex, "vertex")):
        _f4_tensor(t, w)
    _u32_tensor(triangle, "triangle")
    n = position.numel() // 4
    if force.numel() // 4 < n or velocity.numel() // 4 < n:
        raise ValueError("force / velocity hold fewer particles than position")
    return int(N.load().mfl_collide_bvh(force.data_ptr(), position.data_ptr(), velocity.data_ptr(), n,
                                        ctypes.byref(params), bvh._h, triangle.data_ptr(), triangle.numel(),
                                        vertex.data_ptr(), vertex.numel() // 4, _stream_ptr(stream)))


class oclFluid3Dnext:
    """libCL/phys/oclFluid3Dnext.h:20-130 over libmfl's gfx950 kernels: a 2D particle state (x, y, vx, vy)
    per float4 on a dense cellCountX x cellCountY grid."""

    cLocalSize = 256

    def __init__(self):
        L = N.load()
        self._h = L.mfl_fluid3dnext_create()
        if not self._h:
            raise RuntimeError(N.last_error())
        self._user = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and N is not None and N._lib is not None:
            N._lib.mfl_fluid3dnext_destroy(h)
            self._h = None

    def compile(self) -> int:
        return int(N.load().mfl_fluid3dnext_compile(self._h))

    def compute(self, stream=None) -> int:
        return int(N.load().mfl_fluid3dnext_compute(self._h, _stream_ptr(stream)))

    def setParticleCount(self, n: int) -> int:
        return int(N.load().mfl_fluid3dnext_set_particle_count(self._h, n))

    def getParticleCount(self) -> int:
        return int(N.load().mfl_fluid3dnext_get_particle_count(self._h))

    def getParameters(self) -> N.NextParams:
        """Live host Params (getParameters, cpp:150-153); edits apply at the next compute."""
        return N.load().mfl_fluid3dnext_params(self._h).contents

    def setPositionBuffer(self, tensor) -> int:
        _f4_tensor(tensor, "state")
        ok = int(N.load().mfl_fluid3dnext_set_buffer(self._h, N.MFL_NEXT_POSITION, tensor.data_ptr(),
                                                     tensor.numel() * 4))
        if ok:
            self._user = tensor
        return ok

    def _read(self, kind: int, dtype, cols: int):
        ptr, nbytes = ctypes.c_void_p(), ctypes.c_size_t()
        if not N.load().mfl_fluid3dnext_get_buffer(self._h, kind, ctypes.byref(ptr), ctypes.byref(nbytes)):
            raise RuntimeError(N.last_error())
        out = np.empty(nbytes.value // np.dtype(dtype).itemsize, dtype)
        if not N.load().mfl_fluid3dnext_read_buffer(self._h, kind, out.ctypes.data, out.nbytes):
            raise RuntimeError(N.last_error())
        return out.reshape(-1, cols) if cols > 1 else out

    def upload(self, state: np.ndarray) -> int:
        state = np.ascontiguousarray(state, np.float32)
        return int(N.load().mfl_fluid3dnext_write_buffer(self._h, N.MFL_NEXT_POSITION, state.ctypes.data,
                                                         state.nbytes))

    def state(self) -> dict:
        n = self.getParticleCount()
        return dict(state=self._read(N.MFL_NEXT_POSITION, np.float32, 4)[:n])

    def internals(self) -> dict:
        n = self.getParticleCount()
        return dict(cell=self._read(N.MFL_NEXT_CELL, np.uint32, 1)[:n], index=self._read(N.MFL_NEXT_INDEX, np.uint32, 1)[:n],
                    sorted=self._read(N.MFL_NEXT_SORTED_STATE, np.float32, 4)[:n],
                    cell_start=self._read(N.MFL_NEXT_CELL_START, np.uint32, 1),
                    cell_end=self._read(N.MFL_NEXT_CELL_END, np.uint32, 1),
                    pressure=self._read(N.MFL_NEXT_PRESSURE, np.float32, 2)[:n],
                    force=self._read(N.MFL_NEXT_FORCE, np.float32, 2)[:n])


class FluidBatch:
    """Many independent oclFluid3D scenes advanced by one set of launches per step (an addition with no
    reference counterpart; each scene behaves as its own oclFluid3D, libCL/phys/oclFluid3D.cpp:293-342)."""

    def __init__(self, counts):
        counts = np.ascontiguousarray(counts, np.uint32)
        L = N.load()
        self._h = L.mfl_batch_create(len(counts), counts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
        if not self._h:
            raise RuntimeError(N.last_error())
        self.counts = counts.copy()
        self.offsets = np.concatenate([[0], np.cumsum(counts.astype(np.int64))])[:-1].astype(np.int64)
        self.total = int(counts.astype(np.int64).sum())

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and N is not None and N._lib is not None:
            N._lib.mfl_batch_destroy(h)
            self._h = None

    def setParameters(self, params, scene: int = -1) -> int:
        p = params if isinstance(params, N.Params) else N.params_from(params)
        return int(N.load().mfl_batch_set_params(self._h, scene, ctypes.byref(p)))

    def getParameters(self, scene: int) -> N.Params:
        p = N.Params()
        if not N.load().mfl_batch_get_params(self._h, scene, ctypes.byref(p)):
            raise RuntimeError(N.last_error())
        return p

    def setBox(self, box_min, box_max, scene: int = -1) -> int:
        mn = (ctypes.c_float * 4)(*[float(x) for x in box_min])
        mx = (ctypes.c_float * 4)(*[float(x) for x in box_max])
        return int(N.load().mfl_batch_set_box(self._h, scene, mn, mx))

    def setBucketClear(self, count: int) -> int:
        return int(N.load().mfl_batch_set_bucket_clear(self._h, count))

    def setUseGraph(self, enable: bool) -> int:
        return int(N.load().mfl_batch_set_use_graph(self._h, 1 if enable else 0))

    def compute(self, stream=None) -> int:
        return int(N.load().mfl_batch_compute(self._h, _stream_ptr(stream)))

    def getTuning(self) -> dict:
        t = N.Tuning()
        if not N.load().mfl_batch_get_tuning(self._h, ctypes.byref(t)):
            raise RuntimeError(N.load().mfl_last_error().decode())
        return t.as_dict()

    def setTuning(self, **kw) -> int:
        """As oclFluid3D.setTuning; batched scenes take the "list", "tile_list" and "tile_kept" passes."""
        cur = self.getTuning()
        if "density_lanes" not in kw and kw.get("neighbour_pass") == "list":
            kw = dict(kw, density_lanes=8)
        cur.update(kw)
        t = N.Tuning(N.PASSES.index(cur["neighbour_pass"]), cur["density_lanes"], cur["tile_depth"],
                     cur["block_waves"], cur["sort_items"], cur["force_chunks"])
        return int(N.load().mfl_batch_set_tuning(self._h, ctypes.byref(t)))

    def autotune(self, steps: int = 3, stream=None):
        t = N.Tuning()
        ms = ctypes.c_double()
        if not N.load().mfl_batch_autotune(self._h, _stream_ptr(stream), steps, ctypes.byref(t), ctypes.byref(ms)):
            return None
        return t.as_dict(), ms.value

    def write(self, kind: int, host: np.ndarray) -> int:
        host = np.ascontiguousarray(host)
        return int(N.load().mfl_batch_write_buffer(self._h, kind, host.ctypes.data, host.nbytes))

    def read(self, kind: int, dtype=np.float32, shape=None) -> np.ndarray:
        ptr, nbytes = ctypes.c_void_p(), ctypes.c_size_t()
        if not N.load().mfl_batch_get_buffer(self._h, kind, ctypes.byref(ptr), ctypes.byref(nbytes)):
            raise RuntimeError(N.last_error())
        out = np.empty(nbytes.value // np.dtype(dtype).itemsize, dtype)
        if not N.load().mfl_batch_read_buffer(self._h, kind, out.ctypes.data, out.nbytes):
            raise RuntimeError(N.last_error())
        return out.reshape(shape) if shape is not None else out

    def upload(self, pos, vel, force) -> int:
        ok = self.write(N.MFL_POSITION, np.asarray(pos, np.float32))
        ok &= self.write(N.MFL_VELOCITY, np.asarray(vel, np.float32))
        ok &= self.write(N.MFL_FORCE, np.asarray(force, np.float32))
        return int(ok)

    def state(self) -> dict:
        return {k: self.read(kind, np.float32, (-1, 4)) for k, kind in
                (("pos", N.MFL_POSITION), ("vel", N.MFL_VELOCITY), ("force", N.MFL_FORCE))}

    def scene_tables(self, scene: int):
        cs = self.read(N.MFL_CELL_START, np.uint32).reshape(len(self.counts), -1)
        ce = self.read(N.MFL_CELL_END, np.uint32).reshape(len(self.counts), -1)
        return cs[scene], ce[scene]


class Adjoint:
    """Reverse-mode derivative of `steps` oclFluid3D.compute() steps of a program (an addition with no
    reference counterpart): forward() runs the rollout with checkpoints every `checkpoint_every` steps,
    backward(pos_bar, vel_bar) returns dL/dx_0, dL/dv_0 and the viscosity / stiffness / rest-density
    gradients."""

    def __init__(self, program: "oclFluid3D", steps: int, checkpoint_every: int = 20, list_cap: int = 192):
        self.program = program
        self._h = N.load().mfl_adjoint_create(program._h, steps, checkpoint_every, list_cap)
        if not self._h:
            raise RuntimeError(N.last_error())
        self.steps = steps

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and N is not None and N._lib is not None:
            N._lib.mfl_adjoint_destroy(h)
            self._h = None

    def forward(self, stream=None) -> int:
        return int(N.load().mfl_adjoint_forward(self._h, _stream_ptr(stream)))

    def backward(self, pos_bar=None, vel_bar=None, stream=None) -> int:
        """pos_bar / vel_bar: contiguous float32 [N, 4] device tensors (dL/dx_T, dL/dv_T) or None."""
        pp = pos_bar.data_ptr() if pos_bar is not None else None
        vp = vel_bar.data_ptr() if vel_bar is not None else None
        return int(N.load().mfl_adjoint_backward(self._h, pp, vp, _stream_ptr(stream)))

    def velocity_grad(self):
        """dL/dv_0 as a float32 [N, 4] numpy array (caller order)."""
        return self._read(N.MFL_ADJ_VELOCITY_GRAD)

    def position_grad(self):
        return self._read(N.MFL_ADJ_POSITION_GRAD)

    def _read(self, which: int) -> np.ndarray:
        ptr, nbytes = ctypes.c_void_p(), ctypes.c_size_t()
        if not N.load().mfl_adjoint_get_buffer(self._h, which, ctypes.byref(ptr), ctypes.byref(nbytes)):
            raise RuntimeError(N.last_error())
        out = np.empty(nbytes.value // 4, np.float32)
        if not N.load().mfl_memcpy(out.ctypes.data, ptr.value, nbytes.value):
            raise RuntimeError(N.last_error())
        return out.reshape(-1, 4)

    def param_grads(self) -> dict:
        g = (ctypes.c_double * 4)()
        if not N.load().mfl_adjoint_param_grads(self._h, g):
            raise RuntimeError(N.last_error())
        return {"viscosity": g[0], "stiffness": g[1], "density": g[2], "viscosityConstant": g[3]}
