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

// Largest volume side any entry point accepts.  Every entry point that takes
// a volume validates its sides with check_volume before any size arithmetic,
// so no product of sides can wrap.
constexpr uint32_t kMaxVolumeSide = 1u << 16;

// Volume dimension check shared by every entry point that takes a volume:
// each side in [1, kMaxVolumeSide] (VKHR_ERR_INVALID for a zero side,
// VKHR_ERR_OVERFLOW above), at most 2^31 voxels, and with cells the cell
// lattice (nx+1)(ny+1)(nz+1) below 2^32 with (ny+1)(nz+1) below 2^24 (the
// kernels index cells with 24-bit multiplies and 32-bit sums).  All
// arithmetic is 64-bit.  `what` names the volume in the error message.
int check_volume(uint32_t nx, uint32_t ny, uint32_t nz, bool cells, const char* what);

// Scratch of one device and stream: a device buffer of at least `bytes`
// bytes owned by the library until vkhr_release_scratch on that stream (or
// process end).  Calls on one stream are ordered by the stream; calls on
// different streams (and hipStreamPerThread on different threads) get
// different areas.  Growing an area synchronizes the stream before freeing
// the old buffer (outside graph capture).
int stream_scratch(size_t bytes, hipStream_t stream, void** out);

// Bytes of the static per-device scratch of the allocation-free entry
// points (a file-scope __device__ array, no allocation at run time).
constexpr size_t kStaticScratchBytes = VKHR_STATIC_SCRATCH_BYTES;

// The current device's static scratch, after checking that `bytes` fit
// (VKHR_ERR_OVERFLOW otherwise).  Every call on a device shares it and it
// keeps the last call's contents: users clear what must start at zero, and
// calls that overlap on one device race.
int static_scratch(size_t bytes, void** out);

// Synchronizes the stream and frees its area (vkhr_release_scratch).
int release_stream_scratch(hipStream_t stream);

// Bytes held by the area of the calling thread's device and this stream.
size_t stream_scratch_bytes(hipStream_t stream);

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

// The launch's argument block T read in place from the kernarg segment
// (the explicit arguments start at its offset 0, so T must be the kernel's
// first parameter).  Read through this reference, each field is loaded
// from constant memory where it is used; taken from the by-value
// parameter, the compiler holds the block's fields in scalar registers
// across the kernel's loops, and a large block spills them into vector
// register lanes.
template <class T>
__device__ __forceinline__ const T& kernel_args() {
    return *(const T*)__builtin_amdgcn_kernarg_segment_ptr();
}

}  // namespace vkhr
