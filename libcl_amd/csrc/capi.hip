// extern "C" entry points of libmfl.so (declared in include/mfl.h).
#include <cmath>
#include <cstring>
#include <new>

#include "adjoint.hpp"
#include "batch.hpp"
#include "bvh.hpp"
#include "fluid3dnext.hpp"
#include "fluid3d.hpp"
#include "radix_sort.hpp"

namespace mfl {
const char* last_error();
void clear_error();
}  // namespace mfl

struct mfl_fluid3d {
    mfl::Fluid3D impl;
};

struct mfl_radixsort {
    mfl::RadixSort impl;
};

struct mfl_bvh_trimesh {
    mfl::BvhTrimesh impl;
};

struct mfl_fluid3dnext {
    mfl::Fluid3DNext impl;
};

struct mfl_adjoint {
    mfl_adjoint(mfl::Fluid3D& f, uint32_t steps, uint32_t every, uint32_t cap) : impl(f, steps, every, cap) {}
    mfl::Adjoint impl;
};

struct mfl_fluid_batch {
    mfl_fluid_batch(uint32_t s, const uint32_t* c) : impl(s, c) {}
    mfl::FluidBatch impl;
};

namespace {

template <class F>
int guarded(F&& f)
{
    try {
        return f();
    } catch (const std::exception& e) {
        mfl::set_error(std::string("exception: ") + e.what());
    } catch (...) {
        mfl::set_error("unknown exception");
    }
    return 0;
}

bool need(const void* h, const char* what)
{
    if (h) return true;
    mfl::set_error(std::string("null ") + what + " handle");
    return false;
}

// M_PI_F of the OpenCL kernels (libCL/phys/oclFluid3D.cl:119-121): the float nearest pi.
constexpr float kPiF = 3.14159274101257324219f;

}  // namespace

extern "C" {

const char* mfl_last_error(void) { return mfl::last_error(); }

void mfl_clear_error(void) { mfl::clear_error(); }

int mfl_abi_version(void) { return MFL_ABI_VERSION; }

int mfl_set_device(int ordinal)
{
    MFL_HIP(hipSetDevice(ordinal));
    return 1;
}

int mfl_device_finish(void)
{
    MFL_HIP(hipDeviceSynchronize());
    return 1;
}

void* mfl_device_alloc(size_t bytes)
{
    void* p = nullptr;
    if (bytes == 0) bytes = 1;
    if (!mfl::hip_ok(hipMalloc(&p, bytes), "hipMalloc")) return nullptr;
    return p;
}

void mfl_device_free(void* ptr)
{
    if (ptr) (void)hipFree(ptr);
}

int mfl_memcpy(void* dst, const void* src, size_t bytes)
{
    if (bytes == 0) return 1;
    if (!dst || !src) {
        mfl::set_error("null copy address");
        return 0;
    }
    MFL_HIP(hipDeviceSynchronize());
    MFL_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDefault));
    return 1;
}

void mfl_params_derive(mfl_params* p)
{
    if (!p) return;
    // DensityConstant(h) = 315.0/(64.0*M_PI_F*pow(h, 9.0f)): pow in float, the rest in double (cl:124,145)
    const double h9 = (double)powf(p->cellSize, 9.0f);
    p->kernelConstant = (float)(315.0 / (64.0 * (double)kPiF * h9));
    // ViscosityConstant(h) = PressureConstant(h) = -45.0/(M_PI_F * pow(h, 6.0f)), the product in float (cl:126,128)
    const float pih6 = kPiF * powf(p->cellSize, 6.0f);
    const double c = -45.0 / (double)pih6;
    p->viscosityConstant = (float)(-(double)p->viscosity * c);  // cl:146
    p->pressureConstant = (float)(-c);                           // cl:147
}

float mfl_sqrt_cut(float h) { return mfl::sqrt_cut(h); }

void mfl_params_default(mfl_params* p)
{
    if (!p) return;
    // clInitFluid, libCL/phys/oclFluid3D.cl:132-147
    std::memset(p, 0, sizeof(*p));
    p->deltaTime = 0.005f;
    p->mass = 0.006f;
    p->density = 1000.0f;
    p->viscosity = 6.0f;
    p->pressure = 0.0f;
    p->velocitylimit = 500.0f;
    p->stiffness = 1.0f;
    p->spacing = (float)(0.87 * (double)powf(p->mass / p->density, 1.0f / 3.0f));
    p->particleRadius = (float)(0.5 * (double)p->spacing);
    p->cellSize = (float)(2.0 * (double)p->spacing);
    mfl_params_derive(p);
}

// ---- radix sort ---------------------------------------------------------------------------------

mfl_radixsort* mfl_radixsort_create(void)
{
    try {
        return new mfl_radixsort();
    } catch (...) {
        mfl::set_error("out of host memory");
        return nullptr;
    }
}

void mfl_radixsort_destroy(mfl_radixsort* s) { delete s; }

int mfl_radixsort_compile(mfl_radixsort* s)
{
    if (!need(s, "radix sort")) return 0;
    return guarded([&] { return s->impl.compile(); });
}

int mfl_radixsort_compute(mfl_radixsort* s, uint32_t* keys, uint32_t* vals, size_t n, int start_bit, int end_bit,
                          void* stream)
{
    if (!need(s, "radix sort")) return 0;
    return guarded([&] { return s->impl.compute(keys, vals, n, start_bit, end_bit, (hipStream_t)stream); });
}

int mfl_radixsort_compute_noalloc(mfl_radixsort* s, uint32_t* keys, uint32_t* vals, size_t n, int start_bit,
                                  int end_bit, void* stream)
{
    if (!need(s, "radix sort")) return 0;
    return guarded([&] { return s->impl.compute_noalloc(keys, vals, n, start_bit, end_bit, (hipStream_t)stream); });
}

// ---- fluid --------------------------------------------------------------------------------------

mfl_fluid3d* mfl_fluid3d_create(void)
{
    mfl::clear_error();
    mfl_fluid3d* f = nullptr;
    try {
        f = new mfl_fluid3d();
    } catch (...) {
        mfl::set_error("out of host memory");
        return nullptr;
    }
    if (mfl::last_error()[0] != '\0') {  // a device allocation in the constructor failed
        delete f;
        return nullptr;
    }
    f->impl.handle = f;
    return f;
}

void mfl_fluid3d_destroy(mfl_fluid3d* f) { delete f; }

int mfl_fluid3d_compile(mfl_fluid3d* f)
{
    if (!need(f, "fluid")) return 0;
    return guarded([&] { return f->impl.compile(); });
}

int mfl_fluid3d_compute(mfl_fluid3d* f, void* stream)
{
    if (!need(f, "fluid")) return 0;
    return guarded([&] { return f->impl.compute((hipStream_t)stream); });
}

int mfl_fluid3d_set_particle_count(mfl_fluid3d* f, size_t n)
{
    if (!need(f, "fluid")) return 0;
    return guarded([&] { return f->impl.set_particle_count(n); });
}

size_t mfl_fluid3d_get_particle_count(mfl_fluid3d* f) { return f ? f->impl.particle_count() : 0; }

int mfl_fluid3d_set_buffer(mfl_fluid3d* f, int kind, void* ptr, size_t bytes)
{
    if (!need(f, "fluid")) return 0;
    return guarded([&] { return f->impl.set_buffer(kind, ptr, bytes); });
}

int mfl_fluid3d_get_buffer(mfl_fluid3d* f, int kind, void** ptr, size_t* bytes)
{
    if (!need(f, "fluid")) return 0;
    return guarded([&] { return f->impl.get_buffer(kind, ptr, bytes); });
}

int mfl_fluid3d_write_buffer(mfl_fluid3d* f, int kind, const void* host, size_t bytes)
{
    if (!need(f, "fluid")) return 0;
    return guarded([&] { return f->impl.write_buffer(kind, host, bytes); });
}

int mfl_fluid3d_read_buffer(mfl_fluid3d* f, int kind, void* host, size_t bytes)
{
    if (!need(f, "fluid")) return 0;
    return guarded([&] { return f->impl.read_buffer(kind, host, bytes); });
}

mfl_params* mfl_fluid3d_params(mfl_fluid3d* f)
{
    if (!need(f, "fluid")) return nullptr;
    return &f->impl.params;
}

int mfl_fluid3d_set_box(mfl_fluid3d* f, const float* min4, const float* max4)
{
    if (!need(f, "fluid") || !min4 || !max4) {
        if (f) mfl::set_error("null box bound");
        return 0;
    }
    std::memcpy(f->impl.box_min, min4, 16);
    std::memcpy(f->impl.box_max, max4, 16);
    return 1;
}

int mfl_fluid3d_set_gravity(mfl_fluid3d* f, const float* dir4)
{
    if (!need(f, "fluid") || !dir4) {
        if (f) mfl::set_error("null gravity");
        return 0;
    }
    std::memcpy(f->impl.gravity, dir4, 16);
    return 1;
}

int mfl_fluid3d_set_bucket_clear(mfl_fluid3d* f, uint32_t count)
{
    if (!need(f, "fluid")) return 0;
    if (count > mfl::kBuckets) {
        mfl::set_error("bucket clear count exceeds 262144");
        return 0;
    }
    f->impl.clear_count = count;
    return 1;
}

int mfl_fluid3d_set_use_graph(mfl_fluid3d* f, int enable)
{
    if (!need(f, "fluid")) return 0;
    f->impl.use_graph = enable != 0;
    return 1;
}

int mfl_fluid3d_set_record_counts(mfl_fluid3d* f, int enable)
{
    if (!need(f, "fluid")) return 0;
    f->impl.record_counts = enable != 0;
    return 1;
}

int mfl_fluid3d_set_traversal(mfl_fluid3d* f, int order, uint32_t refresh_steps)
{
    if (!need(f, "fluid")) return 0;
    if (order < MFL_TRAVERSAL_SORTED || order > MFL_TRAVERSAL_MORTON) {
        mfl::set_error("unknown traversal order " + std::to_string(order));
        return 0;
    }
    f->impl.traversal = order;
    f->impl.refresh_interval = refresh_steps;
    f->impl.traversal_chosen = true;
    f->impl.invalidate_order();
    return 1;
}

int mfl_fluid3d_get_traversal(mfl_fluid3d* f, int* order, uint32_t* refresh_steps)
{
    if (!need(f, "fluid")) return 0;
    if (order) *order = f->impl.traversal;
    if (refresh_steps) *refresh_steps = f->impl.refresh_interval;
    return 1;
}

int mfl_fluid3d_get_tuning(mfl_fluid3d* f, mfl_tuning* out)
{
    if (!need(f, "fluid")) return 0;
    return f->impl.get_tuning(out);
}

int mfl_fluid3d_set_tuning(mfl_fluid3d* f, const mfl_tuning* t)
{
    if (!need(f, "fluid")) return 0;
    if (!t) {
        mfl::set_error("null tuning");
        return 0;
    }
    return f->impl.set_tuning(*t);
}

int mfl_fluid3d_autotune(mfl_fluid3d* f, void* stream, uint32_t steps, mfl_tuning* best, double* ms_per_step)
{
    if (!need(f, "fluid")) return 0;
    return guarded([&] { return f->impl.autotune((hipStream_t)stream, steps, best, ms_per_step); });
}

int mfl_fluid3d_set_profiling(mfl_fluid3d* f, uint32_t stage_mask)
{
    if (!need(f, "fluid")) return 0;
    f->impl.timer.mask = stage_mask & ((1u << MFL_STAGE_COUNT) - 1u);
    return 1;
}

int mfl_fluid3d_stage_time(mfl_fluid3d* f, int stage, double* total_ms, uint64_t* launches, int reset)
{
    if (!need(f, "fluid")) return 0;
    return guarded([&] { return f->impl.timer.read(stage, total_ms, launches, reset != 0); });
}

int mfl_fluid3d_set_integrate_callback(mfl_fluid3d* f, mfl_integrate_fn fn, void* user)
{
    if (!need(f, "fluid")) return 0;
    f->impl.integrate_cb = fn;
    f->impl.integrate_user = user;
    return 1;
}

int mfl_fluid3d_enqueue_kernel(mfl_fluid3d* f, const char* name, void* stream)
{
    if (!need(f, "fluid")) return 0;
    return guarded([&] { return f->impl.enqueue_kernel(name, (hipStream_t)stream); });
}

// ---- batched scenes ---------------------------------------------------------------------------

mfl_fluid_batch* mfl_batch_create(uint32_t num_scenes, const uint32_t* counts)
{
    mfl::clear_error();
    mfl_fluid_batch* b = nullptr;
    try {
        b = new mfl_fluid_batch(num_scenes, counts);
    } catch (...) {
        mfl::set_error("out of host memory");
        return nullptr;
    }
    if (!b->impl.ok()) {
        delete b;
        return nullptr;
    }
    return b;
}

void mfl_batch_destroy(mfl_fluid_batch* b) { delete b; }

int mfl_batch_scene_range(mfl_fluid_batch* b, uint32_t scene, uint32_t* offset, uint32_t* count)
{
    if (!need(b, "batch")) return 0;
    return b->impl.scene_range(scene, offset, count);
}

int mfl_batch_set_params(mfl_fluid_batch* b, int scene, const mfl_params* p)
{
    if (!need(b, "batch") || !need(p, "params")) return 0;
    return b->impl.set_params(scene, *p);
}

int mfl_batch_get_params(mfl_fluid_batch* b, uint32_t scene, mfl_params* p)
{
    if (!need(b, "batch")) return 0;
    return b->impl.get_params(scene, p);
}

int mfl_batch_set_box(mfl_fluid_batch* b, int scene, const float* min4, const float* max4)
{
    if (!need(b, "batch")) return 0;
    return b->impl.set_box(scene, min4, max4);
}

int mfl_batch_set_gravity(mfl_fluid_batch* b, const float* dir4)
{
    if (!need(b, "batch") || !need(dir4, "gravity")) return 0;
    std::memcpy(b->impl.gravity, dir4, 16);
    return 1;
}

int mfl_batch_set_bucket_clear(mfl_fluid_batch* b, uint32_t count)
{
    if (!need(b, "batch")) return 0;
    if (count > mfl::kBuckets) {
        mfl::set_error("bucket clear count exceeds 262144");
        return 0;
    }
    b->impl.clear_count = count;
    return 1;
}

int mfl_batch_set_use_graph(mfl_fluid_batch* b, int enable)
{
    if (!need(b, "batch")) return 0;
    b->impl.use_graph = enable != 0;
    return 1;
}

int mfl_batch_get_buffer(mfl_fluid_batch* b, int kind, void** ptr, size_t* bytes)
{
    if (!need(b, "batch")) return 0;
    return guarded([&] { return b->impl.get_buffer(kind, ptr, bytes); });
}

int mfl_batch_write_buffer(mfl_fluid_batch* b, int kind, const void* host, size_t bytes)
{
    if (!need(b, "batch")) return 0;
    return guarded([&] { return b->impl.write_buffer(kind, host, bytes); });
}

int mfl_batch_read_buffer(mfl_fluid_batch* b, int kind, void* host, size_t bytes)
{
    if (!need(b, "batch")) return 0;
    return guarded([&] { return b->impl.read_buffer(kind, host, bytes); });
}

int mfl_batch_compute(mfl_fluid_batch* b, void* stream)
{
    if (!need(b, "batch")) return 0;
    return guarded([&] { return b->impl.compute((hipStream_t)stream); });
}

int mfl_batch_get_tuning(mfl_fluid_batch* b, mfl_tuning* out)
{
    if (!need(b, "batch")) return 0;
    return b->impl.get_tuning(out);
}

int mfl_batch_set_tuning(mfl_fluid_batch* b, const mfl_tuning* t)
{
    if (!need(b, "batch")) return 0;
    if (!t) {
        mfl::set_error("null tuning");
        return 0;
    }
    return b->impl.set_tuning(*t);
}

int mfl_batch_autotune(mfl_fluid_batch* b, void* stream, uint32_t steps, mfl_tuning* best, double* ms_per_step)
{
    if (!need(b, "batch")) return 0;
    return guarded([&] { return b->impl.autotune((hipStream_t)stream, steps, best, ms_per_step); });
}

// ---- adjoint --------------------------------------------------------------------------------------

mfl_adjoint* mfl_adjoint_create(mfl_fluid3d* f, uint32_t steps, uint32_t checkpoint_every, uint32_t list_cap)
{
    mfl::clear_error();
    if (!need(f, "fluid")) return nullptr;
    mfl_adjoint* a = nullptr;
    try {
        a = new mfl_adjoint(f->impl, steps, checkpoint_every, list_cap);
    } catch (...) {
        mfl::set_error("out of host memory");
        return nullptr;
    }
    if (!a->impl.ok()) {
        delete a;
        return nullptr;
    }
    return a;
}

void mfl_adjoint_destroy(mfl_adjoint* a) { delete a; }

int mfl_adjoint_forward(mfl_adjoint* a, void* stream)
{
    if (!need(a, "adjoint")) return 0;
    return guarded([&] { return a->impl.forward((hipStream_t)stream); });
}

int mfl_adjoint_backward(mfl_adjoint* a, const void* pos_bar, const void* vel_bar, void* stream)
{
    if (!need(a, "adjoint")) return 0;
    return guarded([&] {
        return a->impl.backward(static_cast<const float4*>(pos_bar), static_cast<const float4*>(vel_bar),
                                (hipStream_t)stream);
    });
}

int mfl_adjoint_get_buffer(mfl_adjoint* a, int which, void** ptr, size_t* bytes)
{
    if (!need(a, "adjoint")) return 0;
    return a->impl.get_buffer(which, ptr, bytes);
}

int mfl_adjoint_param_grads(mfl_adjoint* a, double* out4)
{
    if (!need(a, "adjoint") || !need(out4, "output")) return 0;
    return guarded([&] { return a->impl.param_grads(out4); });
}

}  // extern "C"

// ---- triangle-mesh BVH and collision ------------------------------------------------------------------

extern "C" {

mfl_bvh_trimesh* mfl_bvh_trimesh_create(void)
{
    try {
        return new mfl_bvh_trimesh();
    } catch (...) {
        mfl::set_error("out of host memory");
        return nullptr;
    }
}

void mfl_bvh_trimesh_destroy(mfl_bvh_trimesh* b) { delete b; }

int mfl_bvh_trimesh_compile(mfl_bvh_trimesh* b)
{
    if (!need(b, "BVH")) return 0;
    return guarded([&] { return b->impl.compile(); });
}

int mfl_bvh_trimesh_compute(mfl_bvh_trimesh* b, const void* vertex, size_t vertex_count, const uint32_t* index,
                            size_t index_count, void* stream)
{
    if (!need(b, "BVH")) return 0;
    return guarded([&] {
        return b->impl.compute(static_cast<const float4*>(vertex), vertex_count, index, index_count,
                               (hipStream_t)stream);
    });
}

int mfl_bvh_trimesh_get_node_buffer(mfl_bvh_trimesh* b, void** ptr, size_t* bytes)
{
    if (!need(b, "BVH")) return 0;
    if (ptr) *ptr = b->impl.nodes();
    if (bytes) *bytes = (size_t)b->impl.node_count() * sizeof(mfl_bvh_node);
    return 1;
}

uint32_t mfl_bvh_trimesh_get_node_count(mfl_bvh_trimesh* b) { return b ? b->impl.node_count() : 0u; }

uint32_t mfl_bvh_trimesh_get_root_node(mfl_bvh_trimesh* b) { return b ? b->impl.root() : 0u; }

int mfl_bvh_trimesh_get_morton(mfl_bvh_trimesh* b, void** keys, void** vals, size_t* count)
{
    if (!need(b, "BVH")) return 0;
    if (keys) *keys = b->impl.morton_keys();
    if (vals) *vals = b->impl.morton_vals();
    if (count) *count = b->impl.node_count() ? (b->impl.node_count() + 1u) / 2u : 0u;
    return 1;
}

int mfl_collide_bvh(void* force, void* position, void* velocity, size_t n, const mfl_params* p, mfl_bvh_trimesh* bvh,
                    const uint32_t* triangle, size_t index_count, const void* vertex, size_t vertex_count,
                    void* stream)
{
    if (!need(bvh, "BVH")) return 0;
    if (!p) {
        mfl::set_error("null params");
        return 0;
    }
    return guarded([&] {
        return bvh->impl.collide(static_cast<float4*>(force), static_cast<float4*>(position),
                                 static_cast<float4*>(velocity), n, *p, triangle, index_count,
                                 static_cast<const float4*>(vertex), vertex_count, (hipStream_t)stream);
    });
}

int mfl_fluid3d_collide_bvh(mfl_fluid3d* f, mfl_bvh_trimesh* bvh, const uint32_t* triangle, size_t index_count,
                            const void* vertex, size_t vertex_count, void* stream)
{
    if (!need(f, "fluid") || !need(bvh, "BVH")) return 0;
    return guarded([&] {
        void *pos = nullptr, *vel = nullptr, *force = nullptr;
        if (!f->impl.get_buffer(MFL_POSITION, &pos, nullptr) || !f->impl.get_buffer(MFL_VELOCITY, &vel, nullptr) ||
            !f->impl.get_buffer(MFL_FORCE, &force, nullptr))
            return 0;
        return bvh->impl.collide(static_cast<float4*>(force), static_cast<float4*>(pos), static_cast<float4*>(vel),
                                 f->impl.particle_count(), f->impl.params, triangle, index_count,
                                 static_cast<const float4*>(vertex), vertex_count, (hipStream_t)stream);
    });
}

}  // extern "C"

// ---- oclFluid3Dnext --------------------------------------------------------------------------------

extern "C" {

void mfl_next_params_default(mfl_next_params* p)
{
    if (!p) return;
    // clInitFluid, libCL/phys/oclFluid3Dnext.cl:95-113 (double literals rounded into the float fields)
    std::memset(p, 0, sizeof(*p));
    p->deltaTime = (float)0.0333333333;
    p->density = 10000.0f;
    p->stiffness0 = (float)0.004;
    p->stiffness1 = (float)0.01;
    p->viscosity0 = 1.0f;
    p->viscosity1 = 1.0f;
    p->radius = 12.0f;
    p->cellSize = 2.0f;
    p->cellCountX = (int32_t)(1024 / p->cellSize);
    p->cellCountY = (int32_t)(768 / p->cellSize);
    p->particleRadius = (float)(0.87 * std::pow(1.0 / (double)p->density, (double)(1 / 3.0f)));
}

mfl_fluid3dnext* mfl_fluid3dnext_create(void)
{
    try {
        return new mfl_fluid3dnext();
    } catch (...) {
        mfl::set_error("out of host memory");
        return nullptr;
    }
}

void mfl_fluid3dnext_destroy(mfl_fluid3dnext* f) { delete f; }

int mfl_fluid3dnext_compile(mfl_fluid3dnext* f)
{
    if (!need(f, "oclFluid3Dnext")) return 0;
    return guarded([&] { return f->impl.compile(); });
}

int mfl_fluid3dnext_set_particle_count(mfl_fluid3dnext* f, size_t n)
{
    if (!need(f, "oclFluid3Dnext")) return 0;
    return guarded([&] { return f->impl.set_particle_count(n); });
}

size_t mfl_fluid3dnext_get_particle_count(mfl_fluid3dnext* f) { return f ? f->impl.particle_count() : 0; }

mfl_next_params* mfl_fluid3dnext_params(mfl_fluid3dnext* f) { return f ? &f->impl.params : nullptr; }

int mfl_fluid3dnext_set_buffer(mfl_fluid3dnext* f, int kind, void* ptr, size_t bytes)
{
    if (!need(f, "oclFluid3Dnext")) return 0;
    return guarded([&] { return f->impl.set_buffer(kind, ptr, bytes); });
}

int mfl_fluid3dnext_get_buffer(mfl_fluid3dnext* f, int kind, void** ptr, size_t* bytes)
{
    if (!need(f, "oclFluid3Dnext")) return 0;
    return guarded([&] { return f->impl.get_buffer(kind, ptr, bytes); });
}

int mfl_fluid3dnext_write_buffer(mfl_fluid3dnext* f, int kind, const void* host, size_t bytes)
{
    if (!need(f, "oclFluid3Dnext")) return 0;
    return guarded([&] { return f->impl.write_buffer(kind, host, bytes); });
}

int mfl_fluid3dnext_read_buffer(mfl_fluid3dnext* f, int kind, void* host, size_t bytes)
{
    if (!need(f, "oclFluid3Dnext")) return 0;
    return guarded([&] { return f->impl.read_buffer(kind, host, bytes); });
}

int mfl_fluid3dnext_compute(mfl_fluid3dnext* f, void* stream)
{
    if (!need(f, "oclFluid3Dnext")) return 0;
    return guarded([&] { return f->impl.compute((hipStream_t)stream); });
}

}  // extern "C"
