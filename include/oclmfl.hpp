/*
 * oclmfl.hpp - C++ program objects with libCL's class and method names over the libmfl C ABI (mfl.h),
 * so that a C++ caller of the reference's oclRadixSort / oclFluid3D / oclBvhTrimesh keeps its code
 * shape. Header only; needs no HIP header (streams are `void*` hipStream_t values, buffers are device
 * pointers). Link with -lmfl.
 *
 * What changes for a reference caller (paths relative to the reference root):
 *   - the oclContext& / oclDevice& arguments go: the programs run on the current HIP device, compute()
 *     takes an optional stream (null = the null stream) and mfl_device_finish() is clFinish;
 *   - get*Buffer() return an oclBuffer view (device pointer + size with blocking read / write) in place
 *     of the reference's mappable oclBuffer* (libCL/oclBuffer.h:60-70);
 *   - an EVT_INTEGRATE handler derives from oclFluid3D::srtEvent and enqueues named kernels with
 *     enqueueKernel(name, stream) instead of clEnqueueNDRangeKernel(getKernel(name)).
 */
#ifndef OCLMFL_HPP
#define OCLMFL_HPP

#include <stddef.h>
#include <stdint.h>

#include "mfl.h"

/* A program buffer: device pointer and byte size, blocking host copies (oclBuffer::read / write,
   libCL/oclBuffer.cpp:47-83, with the host side passed explicitly). */
class oclBuffer {
public:
    oclBuffer() : mPtr(0), mBytes(0) {}
    oclBuffer(void* ptr, size_t bytes) : mPtr(ptr), mBytes(bytes) {}
    void* ptr() const { return mPtr; }
    template <class T> T* ptr() const { return static_cast<T*>(mPtr); }
    size_t size() const { return mBytes; }
    int read(void* host, size_t bytes) const { return bytes <= mBytes && mfl_memcpy(host, mPtr, bytes); }
    int write(const void* host, size_t bytes) const { return bytes <= mBytes && mfl_memcpy(mPtr, host, bytes); }

private:
    void* mPtr;
    size_t mBytes;
};

/* libCL/sort/oclRadixSort.h:21-44 */
class oclRadixSort {
public:
    static const int cBits = 4;  // the reference's digit width; bit ranges stay multiples of it
    oclRadixSort() : s(mfl_radixsort_create()) {}
    ~oclRadixSort() { mfl_radixsort_destroy(s); }
    int compile() { return s ? mfl_radixsort_compile(s) : 0; }
    // oclRadixSort::compute(device, bfKey, bfVal, startBit, endBit) (cpp:62-144): n device uint32 pairs
    int compute(uint32_t* key, uint32_t* val, size_t n, int startBit, int endBit, void* stream = 0)
    {
        return mfl_radixsort_compute(s, key, val, n, startBit, endBit, stream);
    }

private:
    oclRadixSort(const oclRadixSort&);
    oclRadixSort& operator=(const oclRadixSort&);
    mfl_radixsort* s;
};

/* libCL/phys/oclFluid3D.h:21-123 */
class oclFluid3D {
public:
    static const size_t cLocalSize = 256;      // oclFluid3D.cpp:16
    static const size_t cBucketCount = 262144; // oclFluid3D.cpp:17
    typedef mfl_params Params;                 // oclFluid3D.h:48-63
    static const char* EVT_INTEGRATE() { return "OnIntegrate"; }  // oclFluid3D.cpp:19

    // srtEvent named EVT_INTEGRATE (libCL/oclProgram.h:146-150): called by compute() after the force
    // pass in place of clip box + integration, with the stream the step runs on.
    struct srtEvent {
        virtual ~srtEvent() {}
        virtual bool operator()(oclFluid3D& fluid, void* stream) = 0;
    };

    oclFluid3D() : f(mfl_fluid3d_create()), mHandler(0) {}
    ~oclFluid3D() { mfl_fluid3d_destroy(f); }

    int compile() { return f ? mfl_fluid3d_compile(f) : 0; }       // cpp:273-285
    int compute(void* stream = 0) { return mfl_fluid3d_compute(f, stream); }  // cpp:293-342
    int setParticleCount(size_t n) { return mfl_fluid3d_set_particle_count(f, n); }  // cpp:92-124
    size_t getParticleCount() { return mfl_fluid3d_get_particle_count(f); }          // cpp:126-129
    // set{Position,Velocity,Force}Buffer (cpp:131-165): caller-owned device memory
    int setPositionBuffer(void* dev, size_t bytes) { return mfl_fluid3d_set_buffer(f, MFL_POSITION, dev, bytes); }
    int setVelocityBuffer(void* dev, size_t bytes) { return mfl_fluid3d_set_buffer(f, MFL_VELOCITY, dev, bytes); }
    int setForceBuffer(void* dev, size_t bytes) { return mfl_fluid3d_set_buffer(f, MFL_FORCE, dev, bytes); }
    // get*Buffer (cpp:171-201)
    oclBuffer getPositionBuffer() { return buffer(MFL_POSITION); }
    oclBuffer getVelocityBuffer() { return buffer(MFL_VELOCITY); }
    oclBuffer getForceBuffer() { return buffer(MFL_FORCE); }
    oclBuffer getSortedPositionBuffer() { return buffer(MFL_SORTED_POSITION); }
    oclBuffer getSortedVelocityBuffer() { return buffer(MFL_SORTED_VELOCITY); }
    oclBuffer getIndexBuffer() { return buffer(MFL_INDEX); }
    oclBuffer getParamBuffer() { return buffer(MFL_PARAMS); }
    // the internal bfCell / bfCellStart / bfCellEnd (oclFluid3D.h:92-94), read-only
    oclBuffer getCellBuffer() { return buffer(MFL_CELL); }
    oclBuffer getCellStartBuffer() { return buffer(MFL_CELL_START); }
    oclBuffer getCellEndBuffer() { return buffer(MFL_CELL_END); }
    Params& getParameters() { return *mfl_fluid3d_params(f); }  // cpp:204-207

    // addEventHandler(srtEvent&) for EVT_INTEGRATE (cpp:287-291, 325-328); one handler, 0 removes it
    int addEventHandler(srtEvent* handler)
    {
        mHandler = handler;
        return mfl_fluid3d_set_integrate_callback(f, handler ? &oclFluid3D::integrate : 0, this);
    }
    // getKernel(name) + clEnqueueNDRangeKernel with the bound arguments (libCL/TEST/main.cpp:282-302)
    int enqueueKernel(const char* name, void* stream) { return mfl_fluid3d_enqueue_kernel(f, name, stream); }

    // additions: clip box and gravity (fixed in bindBuffers, cpp:256-268), bucket reset length
    int setBox(const float mn[4], const float mx[4]) { return mfl_fluid3d_set_box(f, mn, mx); }
    int setGravity(const float g[4]) { return mfl_fluid3d_set_gravity(f, g); }
    int setBucketClear(uint32_t count) { return mfl_fluid3d_set_bucket_clear(f, count); }
    mfl_fluid3d* handle() { return f; }

private:
    oclFluid3D(const oclFluid3D&);
    oclFluid3D& operator=(const oclFluid3D&);
    oclBuffer buffer(int kind)
    {
        void* p = 0;
        size_t bytes = 0;
        if (!mfl_fluid3d_get_buffer(f, kind, &p, &bytes)) return oclBuffer();
        return oclBuffer(p, bytes);
    }
    static int integrate(mfl_fluid3d*, void* stream, void* user)
    {
        oclFluid3D* self = static_cast<oclFluid3D*>(user);
        return self->mHandler && (*self->mHandler)(*self, stream) ? 1 : 0;
    }
    mfl_fluid3d* f;
    srtEvent* mHandler;
};

/* libCL/geom/oclBvhTrimesh.h:21-82 */
class oclBvhTrimesh {
public:
    typedef mfl_bvh_node BVHNode;  // oclBvhTrimesh.h:33-41
    oclBvhTrimesh() : b(mfl_bvh_trimesh_create()) {}
    ~oclBvhTrimesh() { mfl_bvh_trimesh_destroy(b); }
    int compile() { return b ? mfl_bvh_trimesh_compile(b) : 0; }
    // oclBvhTrimesh::compute(device, bfVertex, bfIndex) (cpp:53-162): float4 vertices, uint32 indices
    int compute(const void* vertex, size_t vertexCount, const uint32_t* index, size_t indexCount, void* stream = 0)
    {
        return mfl_bvh_trimesh_compute(b, vertex, vertexCount, index, indexCount, stream);
    }
    oclBuffer getNodeBuffer()
    {
        void* p = 0;
        size_t bytes = 0;
        if (!mfl_bvh_trimesh_get_node_buffer(b, &p, &bytes)) return oclBuffer();
        return oclBuffer(p, bytes);
    }
    uint32_t getNodeCount() { return mfl_bvh_trimesh_get_node_count(b); }
    uint32_t getRootNode() { return mfl_bvh_trimesh_get_root_node(b); }
    mfl_bvh_trimesh* handle() { return b; }

private:
    oclBvhTrimesh(const oclBvhTrimesh&);
    oclBvhTrimesh& operator=(const oclBvhTrimesh&);
    mfl_bvh_trimesh* b;
};

#endif
