// Library state: thread-local error messages, device checks, the scratch
// areas (one per device and stream), and the generic fill kernel.
#include "common.hpp"

#include <cstdarg>
#include <cstring>
#include <map>
#include <mutex>
#include <utility>

namespace vkhr {

namespace {
thread_local char g_error[512] = {0};

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
bool g_checked[kMaxDevices] = {false};
struct Scratch {
    void* ptr = nullptr;
    size_t bytes = 0;
};
std::map<std::pair<int, hipStream_t>, Scratch> g_scratch;  // (device, stream) -> area
}  // namespace

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
}

void clear_error() { g_error[0] = '\0'; }

int ensure_device() {
    int dev = 0;
    VKHR_CHECK_HIP(hipGetDevice(&dev));
    VKHR_REQUIRE(dev >= 0 && dev < kMaxDevices, VKHR_ERR_NO_DEVICE, "device index %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_checked[dev]) return VKHR_OK;
    hipDeviceProp_t prop;
    VKHR_CHECK_HIP(hipGetDeviceProperties(&prop, dev));
    VKHR_REQUIRE(std::strncmp(prop.gcnArchName, "gfx950", 6) == 0, VKHR_ERR_NO_DEVICE,
                 "device %d is %s; this library is built for gfx950 only", dev, prop.gcnArchName);
    g_checked[dev] = true;
    return VKHR_OK;
}

int scratch(size_t bytes, hipStream_t stream, void** out) {
    int rc = ensure_device();
    if (rc) return rc;
    int dev = 0;
    VKHR_CHECK_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_mutex);
    Scratch& area = g_scratch[std::make_pair(dev, stream)];
    if (area.bytes < bytes) {
        // Grow: the old buffer may still be in use by queued work, so it is
        // kept (not freed) -- growth happens a handful of times per process.
        size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            set_error("scratch hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
            return VKHR_ERR_NOMEM;
        }
        area.ptr = p;
        area.bytes = want;
    }
    *out = area.ptr;
    return VKHR_OK;
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ dst, size_t n, uint32_t value) {
    size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = value;
}

int fill_u32(uint32_t* dst, size_t n, uint32_t value, hipStream_t stream) {
    if (n == 0) return VKHR_OK;
    fill_u32_kernel<<<grid_for(n, 256, 8192), 256, 0, stream>>>(dst, n, value);
    VKHR_CHECK_LAUNCH();
    return VKHR_OK;
}

}  // namespace vkhr

extern "C" {

int vkhr_abi_version(void) { return VKHR_ABI_VERSION; }

const char* vkhr_last_error(void) { return vkhr::g_error; }

int vkhr_init(void) {
    vkhr::clear_error();
    return vkhr::ensure_device();
}

}  // extern "C"
