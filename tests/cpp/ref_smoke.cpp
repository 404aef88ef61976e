// The smoke flows of the reference's test program (libCL/TEST/main.cpp:133-343) written against
// include/oclmfl.hpp, the C++ drop-in over libmfl: testRadixSort, testFluid3D0 (plain compute loop)
// and testFluid3D1 (an EVT_INTEGRATE handler that enqueues clIntegrateForce, clGravity,
// clIntegrateVelocity). The reference fills its inputs with rand() and runs for seven seconds; here the
// inputs come from a seeded generator and each fluid flow runs a fixed number of steps, dumping the
// state around a few of them into `outdir` for a checker to replay.
//
// usage: ref_smoke <outdir>      exit status 0 when every call succeeded
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "oclmfl.hpp"

namespace {

std::string g_out;

// SplitMix64: the seeded stand-in for rand()
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    float unit() { return (float)(next() >> 40) * (1.0f / 16777216.0f); }  // [0, 1)
};

bool dump(const std::string& name, const void* data, size_t bytes)
{
    FILE* fp = std::fopen((g_out + "/" + name).c_str(), "wb");
    if (!fp) return false;
    const bool ok = std::fwrite(data, 1, bytes, fp) == bytes;
    return std::fclose(fp) == 0 && ok;
}

bool dump_buffer(const std::string& name, const oclBuffer& b, size_t bytes)
{
    std::vector<char> host(bytes);
    return b.read(host.data(), bytes) && dump(name, host.data(), bytes);
}

bool fail(const char* what)
{
    std::fprintf(stderr, "%s failed: %s\n", what, mfl_last_error());
    return false;
}

// testRadixSort (main.cpp:133-181): 1024 random keys, values 0..1023, sorted by bits [0, 32)
bool test_radix_sort()
{
    const size_t n = 1024;
    std::vector<uint32_t> key(n), val(n);
    Rng rng(11);
    for (size_t i = 0; i < n; i++) {
        key[i] = (uint32_t)(rng.next() >> 33);  // rand() range: [0, 2^31)
        val[i] = (uint32_t)i;
    }
    if (!dump("sort_keys_in.bin", key.data(), n * 4)) return false;
    // bfKey / bfVal (main.cpp:136-141)
    uint32_t* dk = static_cast<uint32_t*>(mfl_device_alloc(n * 4));
    uint32_t* dv = static_cast<uint32_t*>(mfl_device_alloc(n * 4));
    struct Free {
        void* p[2];
        ~Free()
        {
            mfl_device_free(p[0]);
            mfl_device_free(p[1]);
        }
    } release{{dk, dv}};
    if (!dk || !dv) return fail("mfl_device_alloc");
    if (!mfl_memcpy(dk, key.data(), n * 4) || !mfl_memcpy(dv, val.data(), n * 4)) return fail("upload");
    oclRadixSort clProgram;
    if (!clProgram.compile()) return fail("oclRadixSort::compile");
    if (!clProgram.compute(dk, dv, n, 0, 32)) return fail("oclRadixSort::compute");
    if (!mfl_device_finish()) return fail("finish");
    if (!mfl_memcpy(key.data(), dk, n * 4) || !mfl_memcpy(val.data(), dv, n * 4)) return fail("read back");
    for (size_t i = 1; i < n; i++)
        if (key[i] < key[i - 1]) {
            std::fprintf(stderr, "array not sorted\n");
            return false;
        }
    // the reference's rejections (oclRadixSort.cpp:76-80)
    if (clProgram.compute(dk, dv, n, 0, 30)) return false;
    return dump("sort_keys_out.bin", key.data(), n * 4) && dump("sort_vals_out.bin", val.data(), n * 4);
}

bool dump_state(oclFluid3D& fl, const std::string& tag, bool tables)
{
    const size_t n = fl.getParticleCount(), v4 = n * 16;
    bool ok = dump_buffer(tag + "_pos.bin", fl.getPositionBuffer(), v4) &&
              dump_buffer(tag + "_vel.bin", fl.getVelocityBuffer(), v4) &&
              dump_buffer(tag + "_force.bin", fl.getForceBuffer(), v4);
    if (tables)
        ok = ok && dump_buffer(tag + "_cell_start.bin", fl.getCellStartBuffer(), oclFluid3D::cBucketCount * 4) &&
             dump_buffer(tag + "_cell_end.bin", fl.getCellEndBuffer(), oclFluid3D::cBucketCount * 4);
    else
        ok = ok && dump_buffer(tag + "_cell.bin", fl.getCellBuffer(), n * 4) &&
             dump_buffer(tag + "_index.bin", fl.getIndexBuffer(), n * 4) &&
             dump_buffer(tag + "_spos.bin", fl.getSortedPositionBuffer(), v4) &&
             dump_buffer(tag + "_svel.bin", fl.getSortedVelocityBuffer(), v4) &&
             dump_buffer(tag + "_cell_start.bin", fl.getCellStartBuffer(), oclFluid3D::cBucketCount * 4) &&
             dump_buffer(tag + "_cell_end.bin", fl.getCellEndBuffer(), oclFluid3D::cBucketCount * 4);
    return ok;
}

// the EVT_INTEGRATE handler of testFluid3D1 (main.cpp:267-317): all the fluid kernels except clClipBox
struct srtFluid : public oclFluid3D::srtEvent {
    int calls = 0;
    bool operator()(oclFluid3D& fluid, void* stream) override
    {
        calls++;
        return fluid.enqueueKernel("clIntegrateForce", stream) && fluid.enqueueKernel("clGravity", stream) &&
               fluid.enqueueKernel("clIntegrateVelocity", stream);
    }
};

// testFluid3D0 / testFluid3D1 (main.cpp:188-343): the program's default particle count and clInitFluid
// Params, positions uniform in [-0.5, 0.5)^3, `steps` compute() + clFinish
bool test_fluid(const char* name, bool with_handler, int steps)
{
    oclFluid3D clProgram;
    if (!clProgram.compile()) return fail("oclFluid3D::compile");
    const size_t count = clProgram.getParticleCount();
    oclBuffer buf = clProgram.getPositionBuffer();
    std::vector<float> pos(count * 4);
    Rng rng(with_handler ? 23 : 17);
    for (size_t i = 0; i < count; i++) {
        pos[4 * i + 0] = rng.unit() - 0.5f;
        pos[4 * i + 1] = rng.unit() - 0.5f;
        pos[4 * i + 2] = rng.unit() - 0.5f;
        pos[4 * i + 3] = 0.f;
    }
    if (!buf.write(pos.data(), pos.size() * 4)) return fail("position upload");
    srtFluid handler;
    if (with_handler && !clProgram.addEventHandler(&handler)) return fail("addEventHandler");
    const int snaps[] = {0, 1, 10, steps - 1};
    for (int s = 0; s < steps; s++) {
        bool snap = false;
        for (int k : snaps) snap = snap || k == s;
        const std::string tag = std::string(name) + "_" + std::to_string(s);
        if (snap && !dump_state(clProgram, tag + "_before", true)) return fail("dump");
        if (!clProgram.compute()) return fail("oclFluid3D::compute");
        if (!mfl_device_finish()) return fail("clFinish");
        if (snap && !dump_state(clProgram, tag + "_after", false)) return fail("dump");
    }
    if (with_handler && handler.calls != steps) {
        std::fprintf(stderr, "EVT_INTEGRATE handler ran %d times in %d steps\n", handler.calls, steps);
        return false;
    }
    // average particle position (main.cpp:222-240)
    if (!buf.read(pos.data(), pos.size() * 4)) return fail("position read back");
    double avg[3] = {0, 0, 0};
    for (size_t i = 0; i < count; i++)
        for (int c = 0; c < 3; c++) avg[c] += pos[4 * i + c];
    std::printf("%s: %d steps, average particle position = (%g, %g, %g)\n", name, steps, avg[0] / count,
                avg[1] / count, avg[2] / count);
    return true;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <outdir>\n", argv[0]);
        return 2;
    }
    g_out = argv[1];
    if (!test_radix_sort()) return 1;
    std::printf("testRadixSort completed\n");
    if (!test_fluid("fluid0", false, 40)) return 1;
    if (!test_fluid("fluid1", true, 40)) return 1;
    std::printf("ref_smoke ok\n");
    return 0;
}
