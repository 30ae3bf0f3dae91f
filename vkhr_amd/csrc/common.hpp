// Shared host/device helpers of libvkhr_hip: status plumbing, the per-device
// scratch area, and small device primitives.  gfx950 only (wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "vkhr_hip.h"

namespace vkhr {

constexpr int kWave = 64;

// ----------------------------------------------------------------- status

void set_error(const char* fmt, ...);
void clear_error();

#define VKHR_CHECK_HIP(expr)                                                   \
    do {                                                                       \
        hipError_t vkhr_e_ = (expr);                                           \
        if (vkhr_e_ != hipSuccess) {                                           \
            ::vkhr::set_error("%s failed: %s (%s:%d)", #expr,                  \
                              hipGetErrorString(vkhr_e_), __FILE__, __LINE__); \
            return VKHR_ERR_HIP;                                               \
        }                                                                      \
    } while (0)

#define VKHR_REQUIRE(cond, code, ...)                                          \
    do {                                                                       \
        if (!(cond)) {                                                         \
            ::vkhr::set_error(__VA_ARGS__);                                    \
            return (code);                                                     \
        }                                                                      \
    } while (0)

// Checks the launch that was just enqueued.
#define VKHR_CHECK_LAUNCH() VKHR_CHECK_HIP(hipGetLastError())

// Binds per-device state; returns VKHR_OK or an error code.
int ensure_device();

// Scratch of one device and stream: a device buffer of at least `bytes`
// bytes owned by the library until process end.  Calls on one stream are
// ordered by the stream; calls on different streams get different areas.
int scratch(size_t bytes, hipStream_t stream, void** out);

inline hipStream_t as_stream(vkhr_stream s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned grid_for(size_t n, unsigned block, unsigned cap = 1u << 20) {
    size_t g = (n + block - 1) / block;
    if (g == 0) g = 1;
    if (g > cap) g = cap;
    return static_cast<unsigned>(g);
}

// Fills n 32-bit words with a value (graph-capture safe, unlike memset nodes).
int fill_u32(uint32_t* dst, size_t n, uint32_t value, hipStream_t stream);

// ------------------------------------------------------------- device side

__device__ __forceinline__ float3 load_xyz(const float* __restrict__ p, uint32_t i) {
    return make_float3(p[3u * i + 0u], p[3u * i + 1u], p[3u * i + 2u]);
}

}  // namespace vkhr
