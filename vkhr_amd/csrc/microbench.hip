// Measurement kernels for the roofline figures (not part of the rendering
// path): the device-atomic throughput that bounds voxelization and the
// adjoint scatter, and a plain copy of known size that calibrates the HBM
// traffic counters.  Built into libvkhr_microbench.so, loaded by bench.py.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

// Address patterns of the atomic microbenchmark.
enum Pattern : int {
    kUniform = 0,   // uniform random over the whole buffer
    kBrick = 1,     // each wave inside one random 512-word (2 KiB) brick
    kSameWord = 2,  // all lanes of a wave on one random word
    kStream = 3,    // lane-consecutive words, no reuse within a launch
};

// Operand types.
enum Kind : int {
    kU32Device = 0,    // atomicAdd(uint32), device scope
    kF32Device = 1,    // float add without return, device scope
    kU64Device = 2,    // atomicAdd(uint64), device scope
    kF32Workgroup = 3, // float add without return, workgroup scope
};

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ uint64_t word_of(int pattern, uint64_t op, uint64_t n_words) {
    const uint64_t wave = op / 64, lane = op % 64;
    switch (pattern) {
    case kUniform:
        return (uint64_t(mix32(uint32_t(op))) * 0x9e3779b1ull + mix32(uint32_t(op >> 32) + 17u)) % n_words;
    case kBrick: {
        const uint64_t bricks = n_words / 512;
        const uint64_t b = uint64_t(mix32(uint32_t(wave))) % bricks;
        return b * 512 + (mix32(uint32_t(op) * 2654435761u) & 511u);
    }
    case kSameWord:
        return uint64_t(mix32(uint32_t(wave))) % n_words;
    default:
        return (wave * 64 + lane) % n_words;
    }
}

template <int K>
__global__ void __launch_bounds__(256) atomics_kernel(void* buf, uint64_t n_words, uint64_t n_ops, int pattern) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t op = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; op < n_ops; op += stride) {
        const uint64_t w = word_of(pattern, op, n_words);
        if constexpr (K == kU32Device) {
            atomicAdd(static_cast<uint32_t*>(buf) + w, 1u);
        } else if constexpr (K == kF32Device) {
            unsafeAtomicAdd(static_cast<float*>(buf) + w, 1.0f);
        } else if constexpr (K == kU64Device) {
            atomicAdd(static_cast<unsigned long long*>(buf) + w, 1ull);
        } else {
            __hip_atomic_fetch_add(static_cast<float*>(buf) + w, 1.0f, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
}

__global__ void __launch_bounds__(256) copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

}  // namespace

extern "C" {

// Enqueues n_ops atomic adds of the given kind over buf[0, n_words) (words
// of the kind's size) in the given pattern.  Returns 0, or -1 for invalid
// arguments (the brick pattern needs n_words >= 512).
int vkhr_mb_atomics(int kind, int pattern, void* buf, uint64_t n_words, uint64_t n_ops, hipStream_t stream) {
    if (!buf || n_words == 0 || pattern < 0 || pattern > 3 || kind < 0 || kind > 3) return -1;
    if (pattern == kBrick && n_words < 512) return -1;
    const unsigned grid = 256u * 64u;
    switch (kind) {
    case kU32Device: atomics_kernel<kU32Device><<<grid, 256, 0, stream>>>(buf, n_words, n_ops, pattern); break;
    case kF32Device: atomics_kernel<kF32Device><<<grid, 256, 0, stream>>>(buf, n_words, n_ops, pattern); break;
    case kU64Device: atomics_kernel<kU64Device><<<grid, 256, 0, stream>>>(buf, n_words, n_ops, pattern); break;
    default: atomics_kernel<kF32Workgroup><<<grid, 256, 0, stream>>>(buf, n_words, n_ops, pattern); break;
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

// Copies bytes (a multiple of 16) from src to dst: the reference kernel
// whose FETCH_SIZE / WRITE_SIZE counters calibrate the traffic figures.
int vkhr_mb_copy(const void* src, void* dst, uint64_t bytes, hipStream_t stream) {
    if (!src || !dst || bytes % 16) return -1;
    const size_t n = bytes / 16;
    copy_kernel<<<8192, 256, 0, stream>>>(static_cast<const float4*>(src), static_cast<float4*>(dst), n);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // extern "C"
