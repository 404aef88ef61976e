// Error state and device-buffer helpers of libmfl.
#include "mfl_common.hpp"

namespace mfl {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }

const char* last_error() { return g_last_error.c_str(); }

void clear_error() { g_last_error.clear(); }

bool hip_ok(hipError_t e, const char* what)
{
    if (e == hipSuccess) return true;
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return false;
}

bool DeviceBuffer::allocate(size_t nbytes)
{
    if (!owned) {
        set_error("cannot re-allocate a caller-owned buffer");
        return false;
    }
    release();
    if (nbytes == 0) return true;
    if (!hip_ok(hipMalloc(&ptr, nbytes), "hipMalloc")) {
        ptr = nullptr;
        return false;
    }
    bytes = nbytes;
    return true;
}

bool DeviceBuffer::fit(size_t nbytes)
{
    if (bytes >= nbytes && ptr) return true;
    return allocate(nbytes);
}

void DeviceBuffer::release()
{
    if (owned && ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

}  // namespace mfl
