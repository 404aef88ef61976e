// extern "C" entry points of libmfl.so (declared in include/mfl.h).
#include <cmath>
#include <cstring>
#include <new>

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

int mfl_fluid3d_set_traversal(mfl_fluid3d* f, int order, uint32_t refresh_steps)
{
    if (!need(f, "fluid")) return 0;
    if (order < MFL_TRAVERSAL_SORTED || order > MFL_TRAVERSAL_MORTON) {
        mfl::set_error("unknown traversal order " + std::to_string(order));
        return 0;
    }
    f->impl.traversal = order;
    f->impl.refresh_interval = refresh_steps;
    return 1;
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

}  // extern "C"
