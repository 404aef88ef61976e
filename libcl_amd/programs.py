"""Python mirror of libCL's program objects on the HIP path: oclRadixSort and oclFluid3D.

Names, argument meaning and the int (1 ok / 0 fail) return convention follow
libCL/sort/oclRadixSort.h:26-28 and libCL/phys/oclFluid3D.h:25-74; failures leave their text in
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
        if h:
            N.load().mfl_radixsort_destroy(h)
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

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            N.load().mfl_fluid3d_destroy(h)
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

    def setTraversal(self, order: str, refresh_steps: int = 16) -> int:
        """Visit order of the neighbour passes: "sorted", "caller" or "morton" (results unchanged)."""
        return int(N.load().mfl_fluid3d_set_traversal(self._h, N.TRAVERSALS.index(order), refresh_steps))

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
