// Library state: thread-local error messages, device checks, the scratch
// areas (one per device and stream), and the generic fill kernel.
#include "common.hpp"

#include <cstdarg>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

namespace vkhr {

namespace {
thread_local char g_error[512] = {0};

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
bool g_checked[kMaxDevices] = {false};
struct Scratch {
    void* ptr = nullptr;
    size_t bytes = 0;
    std::vector<void*> retired;  // outgrown buffers still owned by work captured into a graph
};
// (device, stream, thread) -> area.  The thread is part of the key only for
// hipStreamPerThread, one handle value that names a different stream on
// every thread.
using ScratchKey = std::tuple<int, hipStream_t, size_t>;
std::map<ScratchKey, Scratch> g_scratch;

// The fixed scratch of the allocation-free entry points: one copy per
// device, resident from module load until the process ends.  Host code
// never passes `scratch` itself (its host-side name is not a device
// address); ensure_device looks up each device's copy once, with that
// device current, and keeps the address in g_static_scratch.
static __device__ __align__(256) uint32_t scratch[kStaticScratchBytes / sizeof(uint32_t)];
void* g_static_scratch[kMaxDevices] = {nullptr};

ScratchKey scratch_key(int dev, hipStream_t stream) {
    const size_t thread = stream == hipStreamPerThread ? std::hash<std::thread::id>()(std::this_thread::get_id()) : 0;
    return ScratchKey(dev, stream, thread);
}
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
    void* fixed = nullptr;
    VKHR_CHECK_HIP(hipGetSymbolAddress(&fixed, HIP_SYMBOL(scratch)));
    VKHR_REQUIRE(fixed, VKHR_ERR_HIP, "device %d: no address for the static scratch", dev);
    g_static_scratch[dev] = fixed;
    g_checked[dev] = true;
    return VKHR_OK;
}

int static_scratch(size_t bytes, void** out) {
    int rc = ensure_device();
    if (rc) return rc;
    VKHR_REQUIRE(bytes <= kStaticScratchBytes, VKHR_ERR_OVERFLOW,
                 "needs %zu bytes of static scratch; the allocation-free entry points have %zu", bytes,
                 kStaticScratchBytes);
    int dev = 0;
    VKHR_CHECK_HIP(hipGetDevice(&dev));
    *out = g_static_scratch[dev];  // set under g_mutex before g_checked[dev]
    return VKHR_OK;
}

int check_volume(uint32_t nx, uint32_t ny, uint32_t nz, bool cells, const char* what) {
    VKHR_REQUIRE(nx && ny && nz, VKHR_ERR_INVALID, "%s: zero volume resolution %ux%ux%u", what, nx, ny, nz);
    VKHR_REQUIRE(nx <= kMaxVolumeSide && ny <= kMaxVolumeSide && nz <= kMaxVolumeSide, VKHR_ERR_OVERFLOW,
                 "%s: volume %ux%ux%u has a side above %u", what, nx, ny, nz, kMaxVolumeSide);
    const uint64_t x = nx, y = ny, z = nz;  // each <= 2^16: no product below wraps
    VKHR_REQUIRE(x * y * z <= (uint64_t(1) << 31), VKHR_ERR_OVERFLOW, "%s: volume %ux%ux%u exceeds 2^31 voxels",
                 what, nx, ny, nz);
    if (cells)
        VKHR_REQUIRE((x + 1) * (y + 1) * (z + 1) < (uint64_t(1) << 32) && (y + 1) * (z + 1) < (uint64_t(1) << 24),
                     VKHR_ERR_OVERFLOW, "%s: volume %ux%ux%u has too many cells for 32-bit cell indices", what, nx,
                     ny, nz);
    return VKHR_OK;
}

int stream_scratch(size_t bytes, hipStream_t stream, void** out) {
    int rc = ensure_device();
    if (rc) return rc;
    int dev = 0;
    VKHR_CHECK_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_mutex);
    Scratch& area = g_scratch[scratch_key(dev, stream)];
    if (area.bytes < bytes) {
        // Grow.  Work already queued on this stream may still read the old
        // buffer: it is freed once the stream has drained, or, while the
        // stream is being captured into a graph (which may replay it later),
        // kept until vkhr_release_scratch.
        if (area.ptr) {
            hipStreamCaptureStatus status = hipStreamCaptureStatusNone;
            VKHR_CHECK_HIP(hipStreamIsCapturing(stream, &status));
            if (status == hipStreamCaptureStatusNone) {
                VKHR_CHECK_HIP(hipStreamSynchronize(stream));
                VKHR_CHECK_HIP(hipFree(area.ptr));
            } else {
                area.retired.push_back(area.ptr);
            }
            area.ptr = nullptr;
            area.bytes = 0;
        }
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

int release_stream_scratch(hipStream_t stream) {
    int rc = ensure_device();
    if (rc) return rc;
    int dev = 0;
    VKHR_CHECK_HIP(hipGetDevice(&dev));
    Scratch area;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        auto it = g_scratch.find(scratch_key(dev, stream));
        if (it == g_scratch.end()) return VKHR_OK;
        area = it->second;
        g_scratch.erase(it);
    }
    VKHR_CHECK_HIP(hipStreamSynchronize(stream));
    if (area.ptr) VKHR_CHECK_HIP(hipFree(area.ptr));
    for (void* p : area.retired) VKHR_CHECK_HIP(hipFree(p));
    return VKHR_OK;
}

size_t stream_scratch_bytes(hipStream_t stream) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    std::lock_guard<std::mutex> lock(g_mutex);
    auto it = g_scratch.find(scratch_key(dev, stream));
    return it == g_scratch.end() ? 0 : it->second.bytes;
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

int vkhr_release_scratch(vkhr_stream stream) {
    vkhr::clear_error();
    return vkhr::release_stream_scratch(vkhr::as_stream(stream));
}

size_t vkhr_scratch_bytes(vkhr_stream stream) { return vkhr::stream_scratch_bytes(vkhr::as_stream(stream)); }

}  // extern "C"
