// Volume finalization: counts -> the density / tangent volumes the renderers
// sample (src/vkhr/scene_graph/hair_style.cc:283-291, 331-357;
// src/vkhr/rasterizer/hair_style.cc:75-77), plus the corner-packed density
// layout read by the raymarcher.
//
// Pass 1 (per voxel): the saturating uint8 view, tangent quantization, and
//                     the min/max of the uint8 view (one atomic pair per
//                     block).
// Pass 2 (per voxel): HairStyle::Volume::normalize in place and the float
//                     view.
// Pass 3 (per cell; blocks of 64 x 8 cells walking 32 cell levels up):
//                     corner packing, 8 texels -> one uint64 per cell, and
//                     the brick occupancy bits.
// Pass 4 (three launches over lines of bricks): the brick distance field.
// Compiled with -ffp-contract=off (quantization and normalize are float
// expressions whose rounding must match the reference's).
#include "volume.hpp"

namespace vkhr {
namespace {

__global__ void __launch_bounds__(256)
minmax_init_kernel(uint32_t* mm) {
    if (threadIdx.x < kMinMaxSlots) {
        mm[threadIdx.x * kMinMaxStride] = 255u;
        mm[threadIdx.x * kMinMaxStride + 1] = 0u;
    }
}

__global__ void __launch_bounds__(256)
finalize_counts_kernel(const uint32_t* __restrict__ counts, const long long* __restrict__ sums, size_t n,
                       uint8_t* __restrict__ density, int8_t* __restrict__ tangents,
                       uint32_t* __restrict__ minmax) {
    uint32_t lo = 255u, hi = 0u;
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    const size_t first = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    for (size_t q = first; q < n / 4; q += stride) finalize_quad(counts, sums, 4 * q, density, tangents, &lo, &hi);
    for (size_t i = (n & ~size_t(3)) + first; i < n; i += stride) {  // the tail of n % 4
        uint32_t d = finalize_voxel(counts, sums, i, density, tangents);
        lo = d < lo ? d : lo;
        hi = d > hi ? d : hi;
    }
    // wave min/max, then block min/max in LDS, then one atomic pair per block
    for (int o = 32; o > 0; o >>= 1) {
        uint32_t a = __shfl_xor(lo, o, kWave), b = __shfl_xor(hi, o, kWave);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    __shared__ uint32_t red[2][256 / kWave];
    int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    if (lane == 0) { red[0][wid] = lo; red[1][wid] = hi; }
    __syncthreads();
    if (minmax && threadIdx.x == 0) {
        for (int w = 1; w < int(blockDim.x / kWave); ++w) {
            lo = red[0][w] < lo ? red[0][w] : lo;
            hi = red[1][w] > hi ? red[1][w] : hi;
        }
        fold_minmax(minmax + (blockIdx.x % kMinMaxSlots) * kMinMaxStride, lo, hi);
    }
}

// One block per 64 cell columns x 8 cell rows x cell_slab() cell levels
// (CLAMP_TO_BORDER with an opaque-black border: texels outside the volume
// are 0), with the brick occupancy.
__global__ void __launch_bounds__(256)
cells_kernel(const uint8_t* __restrict__ density, uint32_t nx, uint32_t ny, uint32_t nz,
             unsigned long long* __restrict__ cells, uint32_t* __restrict__ occupancy) {
    const uint32_t n_blocks = cells_pass_blocks(nx, ny, nz);
    for (uint32_t b = blockIdx.x; b < n_blocks; b += gridDim.x) cells_pass(density, nx, ny, nz, b, cells, occupancy);
}

// The bitmask of the occupancy from the brick flags, one word per thread.
__global__ void __launch_bounds__(256)
pack_bits_kernel(uint32_t* __restrict__ occupancy, uint32_t nx, uint32_t ny, uint32_t nz) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w < occupancy_bit_words(nx, ny, nz)) pack_brick_bits(occupancy, nx, ny, nz, w);
}

// One pass of the brick distance field: four lines per block (one per wave).
__global__ void __launch_bounds__(256)
distance_kernel(uint32_t* __restrict__ occupancy, uint32_t nx, uint32_t ny, uint32_t nz, int axis) {
    __shared__ uint32_t lds[256 / kWave][kDistanceLineMax];
    const uint32_t wave = threadIdx.x / kWave;
    const uint32_t line = blockIdx.x * (256 / kWave) + wave;
    distance_line(occupancy, nx, ny, nz, axis, line, line < distance_lines(nx, ny, nz, axis), lds[wave]);
}

// Four voxels per thread (uchar4), the tail of n % 4 by the first threads.
__global__ void __launch_bounds__(256)
normalize_kernel(uint8_t* __restrict__ density, size_t n, const uint32_t* __restrict__ minmax, int normalize,
                 float* __restrict__ density_f32) {
    const DensityMap m = density_map_slots(minmax, normalize != 0);
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    const size_t first = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    uchar4* d4 = reinterpret_cast<uchar4*>(density);
    for (size_t q = first; q < n / 4; q += stride) {
        uchar4 v = d4[q];
        v.x = uint8_t(texel_density(v.x, m));
        v.y = uint8_t(texel_density(v.y, m));
        v.z = uint8_t(texel_density(v.z, m));
        v.w = uint8_t(texel_density(v.w, m));
        if (normalize) d4[q] = v;
        if (density_f32)  // R8_UNORM
            reinterpret_cast<float4*>(density_f32)[q] =
                make_float4(float(v.x) / 255.0f, float(v.y) / 255.0f, float(v.z) / 255.0f, float(v.w) / 255.0f);
    }
    for (size_t i = (n & ~size_t(3)) + first; i < n; i += stride) {
        const uint32_t d = texel_density(density[i], m);
        if (normalize) density[i] = uint8_t(d);
        if (density_f32) density_f32[i] = float(d) / 255.0f;
    }
}

// Volume::downsample: one thread per output voxel.
__global__ void __launch_bounds__(256)
downsample_kernel(const uint8_t* __restrict__ in, uint32_t nx, uint32_t ny, uint32_t ox, uint32_t oy, size_t n_out,
                  int filter, uint8_t* __restrict__ out) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t o = size_t(blockIdx.x) * blockDim.x + threadIdx.x; o < n_out; o += stride) {
        const size_t i = o % ox, jk = o / ox, j = jk % oy, k = jk / oy;
        uint32_t mx = 0u, mn = 255u, sum = 0u;
#pragma unroll
        for (int q = 0; q < 8; ++q) {  // neighbourhood[x + 2y + 4z]
            const size_t x = 2 * i + (q & 1), y = 2 * j + ((q >> 1) & 1), z = 2 * k + (q >> 2);
            const uint32_t v = in[x + nx * (y + size_t(ny) * z)];
            mx = v > mx ? v : mx;
            mn = v < mn ? v : mn;
            sum += v;
        }
        out[o] = uint8_t(filter == 0 ? mx : (filter == 1 ? mn : sum / 8u));
    }
}

// filter_volume at texel centres: the taps are texels, zero outside.
__global__ void __launch_bounds__(256)
filter_kernel(const float* __restrict__ in, int nx, int ny, int nz, int range, float sigma_squared,
              float* __restrict__ out) {
    const size_t n = size_t(nx) * ny * nz;
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t v = size_t(blockIdx.x) * blockDim.x + threadIdx.x; v < n; v += stride) {
        const int x0 = int(v % nx), y0 = int((v / nx) % ny), z0 = int(v / (size_t(nx) * ny));
        float density = 0.0f, total = 0.0f;
        for (int z = -range; z <= range; ++z)
            for (int y = -range; y <= range; ++y)
                for (int x = -range; x <= range; ++x) {
                    const float r2 = float(x * x + y * y + z * z);
                    const float exponent = -1.0f * r2 / 2.0f * sigma_squared;
                    const float w = 1.0f / (2.0f * 3.14159265358979f * sigma_squared) * expf(exponent);
                    const int tx = x0 + x, ty = y0 + y, tz = z0 + z;
                    const bool inside = tx >= 0 && ty >= 0 && tz >= 0 && tx < nx && ty < ny && tz < nz;
                    if (inside) density += in[size_t(tx) + size_t(nx) * (size_t(ty) + size_t(ny) * size_t(tz))] * w;
                    total += w;
                }
        out[v] = density / total;
    }
}

}  // namespace

}  // namespace vkhr

using namespace vkhr;

extern "C" int vkhr_volume_downsample(const uint8_t* in, uint32_t nx, uint32_t ny, uint32_t nz, int filter,
                                      uint8_t* out, vkhr_stream stream) {
    clear_error();
    int rc = ensure_device();
    if (rc) return rc;
    VKHR_REQUIRE(in && out, VKHR_ERR_INVALID, "in/out is null");
    if ((rc = check_volume(nx, ny, nz, false, "downsample input"))) return rc;
    VKHR_REQUIRE(nx >= 2 && ny >= 2 && nz >= 2 && nx % 2 == 0 && ny % 2 == 0 && nz % 2 == 0, VKHR_ERR_INVALID,
                 "downsample needs even sizes, got %ux%ux%u", nx, ny, nz);
    VKHR_REQUIRE(filter >= 0 && filter <= 2, VKHR_ERR_INVALID, "unknown filter %d", filter);
    const size_t n_out = size_t(nx / 2) * (ny / 2) * (nz / 2);
    downsample_kernel<<<grid_for(n_out, 256, 65536), 256, 0, as_stream(stream)>>>(in, nx, ny, nx / 2, ny / 2, n_out,
                                                                                 filter, out);
    VKHR_CHECK_LAUNCH();
    return VKHR_OK;
}

extern "C" int vkhr_volume_filter(const float* in, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t kernel_width,
                                  float* out, vkhr_stream stream) {
    clear_error();
    int rc = ensure_device();
    if (rc) return rc;
    VKHR_REQUIRE(in && out && in != out, VKHR_ERR_INVALID, "in/out is null or the same");
    if ((rc = check_volume(nx, ny, nz, false, "filter input"))) return rc;
    VKHR_REQUIRE(kernel_width % 2 == 1 && kernel_width <= 15, VKHR_ERR_INVALID,
                 "kernel_width %u: odd and at most 15", kernel_width);
    const float sigma = (float(kernel_width) / 2.0f) / 2.4f;
    const size_t n = size_t(nx) * ny * nz;
    filter_kernel<<<grid_for(n, 256, 65536), 256, 0, as_stream(stream)>>>(in, int(nx), int(ny), int(nz),
                                                                         int(kernel_width / 2), sigma * sigma, out);
    VKHR_CHECK_LAUNCH();
    return VKHR_OK;
}

namespace vkhr {
namespace {

// vkhr_volume_finalize with the min/max slots either from the per-stream
// scratch (minmax_static == nullptr) or from the static per-device scratch
// (vkhr_volume_finalize_noalloc), which it then neither allocates nor frees.
int finalize(const uint32_t* counts, const int64_t* tangent_acc, uint32_t nx, uint32_t ny, uint32_t nz,
             int normalize, uint8_t* density_u8, int8_t* tangent_i8x4, float* density_f32, uint64_t* cells,
             uint32_t* occupancy, hipStream_t s, uint32_t* minmax_static) {
    VKHR_REQUIRE(density_u8, VKHR_ERR_INVALID, "density_u8 is null");
    VKHR_REQUIRE(reinterpret_cast<uintptr_t>(density_u8) % 4 == 0 && reinterpret_cast<uintptr_t>(density_f32) % 16 == 0 &&
                     reinterpret_cast<uintptr_t>(counts) % 16 == 0 && reinterpret_cast<uintptr_t>(tangent_i8x4) % 16 == 0 &&
                     reinterpret_cast<uintptr_t>(tangent_acc) % 16 == 0,
                 VKHR_ERR_INVALID,
                 "density_u8 / counts, tangent_acc, density_f32, tangent_i8x4 must be 4 / 16-byte aligned");
    VKHR_REQUIRE(counts || tangent_acc, VKHR_ERR_INVALID, "neither counts nor tangent_acc");
    VKHR_REQUIRE(!occupancy || cells, VKHR_ERR_INVALID, "occupancy needs cells");
    int rc = check_volume(nx, ny, nz, cells != nullptr, "finalize volume");
    if (rc) return rc;
    VKHR_REQUIRE(!tangent_i8x4 || tangent_acc, VKHR_ERR_INVALID, "tangent_i8x4 needs tangent_acc");
    size_t n = size_t(nx) * ny * nz;
    uint32_t* mm = nullptr;
    if (normalize) {
        if (minmax_static) {
            mm = minmax_static;
        } else {
            void* scr = nullptr;
            if ((rc = stream_scratch(kMinMaxBytes, s, &scr))) return rc;
            mm = static_cast<uint32_t*>(scr);
        }
        minmax_init_kernel<<<1, 64, 0, s>>>(mm);  // every call starts from empty slots
        VKHR_CHECK_LAUNCH();
    }
    // One block per 1024 voxels (capped): one min/max atomic pair per block,
    // spread over the slots.
    finalize_counts_kernel<<<grid_for(n / 4 + 1, 256, 2048), 256, 0, s>>>(
        counts, reinterpret_cast<const long long*>(tangent_acc), n, density_u8, tangent_i8x4, mm);
    VKHR_CHECK_LAUNCH();
    return finish_volume(density_u8, nx, ny, nz, normalize, density_f32, cells, occupancy, mm, s);
}

}  // namespace

int finish_volume(uint8_t* density_u8, uint32_t nx, uint32_t ny, uint32_t nz, int normalize, float* density_f32,
                  uint64_t* cells, uint32_t* occupancy, const uint32_t* mm, hipStream_t s) {
    int rc = VKHR_OK;
    const size_t n = size_t(nx) * ny * nz;
    if (normalize || density_f32) {
        // capped: each wave first reduces the min/max slots
        normalize_kernel<<<grid_for(n / 4 + 1, 256, 2048), 256, 0, s>>>(density_u8, n, mm, normalize, density_f32);
        VKHR_CHECK_LAUNCH();
    }
    if (cells) {
        if (occupancy && (rc = fill_u32(occupancy, occupancy_words(nx, ny, nz), 0u, s))) return rc;
        cells_kernel<<<grid_for(cells_pass_blocks(nx, ny, nz), 1, 1u << 20), 256, 0, s>>>(
            density_u8, nx, ny, nz, reinterpret_cast<unsigned long long*>(cells), occupancy);
        VKHR_CHECK_LAUNCH();
        if (occupancy && (rc = finish_occupancy(occupancy, nx, ny, nz, s))) return rc;
    }
    return VKHR_OK;
}

int finish_occupancy(uint32_t* occupancy, uint32_t nx, uint32_t ny, uint32_t nz, hipStream_t s) {
    pack_bits_kernel<<<grid_for(occupancy_bit_words(nx, ny, nz), 256), 256, 0, s>>>(occupancy, nx, ny, nz);
    VKHR_CHECK_LAUNCH();
    if (!distance_field_fits(nx, ny, nz)) return VKHR_OK;
    for (int axis = 0; axis < 3; ++axis) {
        distance_kernel<<<grid_for(distance_lines(nx, ny, nz, axis), 256 / kWave), 256, 0, s>>>(occupancy, nx, ny,
                                                                                               nz, axis);
        VKHR_CHECK_LAUNCH();
    }
    return VKHR_OK;
}

}  // namespace vkhr

extern "C" int vkhr_volume_finalize(const uint32_t* counts, const int64_t* tangent_acc, uint32_t nx, uint32_t ny,
                                    uint32_t nz, int normalize, uint8_t* density_u8, int8_t* tangent_i8x4,
                                    float* density_f32, uint64_t* cells, uint32_t* occupancy,
                                    vkhr_stream stream) {
    clear_error();
    int rc = ensure_device();
    if (rc) return rc;
    return vkhr::finalize(counts, tangent_acc, nx, ny, nz, normalize, density_u8, tangent_i8x4, density_f32, cells,
                          occupancy, as_stream(stream), nullptr);
}

static_assert(kMinMaxBytes <= kStaticScratchBytes, "the min/max slots must fit the static scratch");

extern "C" int vkhr_volume_finalize_noalloc(const uint32_t* counts, const int64_t* tangent_acc, uint32_t nx,
                                            uint32_t ny, uint32_t nz, int normalize, uint8_t* density_u8,
                                            int8_t* tangent_i8x4, float* density_f32, uint64_t* cells,
                                            uint32_t* occupancy, vkhr_stream stream) {
    clear_error();
    void* slots = nullptr;
    int rc = static_scratch(kMinMaxBytes, &slots);  // binds the device and checks the size
    if (rc) return rc;
    // the min/max slots are re-initialised by a kernel at the start of every
    // call that normalizes, so the previous call's contents never leak in
    return vkhr::finalize(counts, tangent_acc, nx, ny, nz, normalize, density_u8, tangent_i8x4, density_f32, cells,
                          occupancy, as_stream(stream), static_cast<uint32_t*>(slots));
}

extern "C" size_t vkhr_occupancy_words(uint32_t nx, uint32_t ny, uint32_t nz) { return occupancy_words(nx, ny, nz); }
