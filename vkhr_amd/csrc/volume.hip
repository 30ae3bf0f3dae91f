// Volume finalization: counts -> the density / tangent volumes the renderers
// sample (src/vkhr/scene_graph/hair_style.cc:283-291, 331-357;
// src/vkhr/rasterizer/hair_style.cc:75-77), plus the corner-packed density
// layout read by the raymarcher.
//
// Pass 1 (per voxel): saturating uint8 view, tangent quantization, and the
//                     min/max of the uint8 view (one atomic pair per block).
// Pass 2 (per voxel): HairStyle::Volume::normalize and the float view.
// Pass 3 (per cell):  corner packing, 8 texels -> one uint64.
// Compiled with -ffp-contract=off (quantization and normalize are float
// expressions whose rounding must match the reference's).
#include "volume.hpp"

namespace vkhr {
namespace {

__global__ void __launch_bounds__(256)
minmax_init_kernel(uint32_t* mm) {
    if (threadIdx.x == 0) { mm[0] = 255u; mm[1] = 0u; }
}

__global__ void __launch_bounds__(256)
finalize_counts_kernel(const uint32_t* __restrict__ counts, const long long* __restrict__ sums, size_t n,
                       uint8_t* __restrict__ density, int8_t* __restrict__ tangents,
                       uint32_t* __restrict__ minmax) {
    uint32_t lo = 255u, hi = 0u;
    size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint32_t d = finalize_voxel(const_cast<uint32_t*>(counts), sums, i, density, tangents);
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
        atomicMin(&minmax[0], lo);
        atomicMax(&minmax[1], hi);
    }
}

__global__ void __launch_bounds__(256)
normalize_kernel(uint8_t* __restrict__ density, size_t n, const uint32_t* __restrict__ minmax, int normalize,
                 float* __restrict__ density_f32) {
    uint32_t lo = minmax ? minmax[0] : 0u, hi = minmax ? minmax[1] : 255u;
    // a constant volume (division by zero, NaN products) becomes all zero
    bool flat = (hi == lo);
    float scaling = normalize_scaling(lo, hi);
    size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint32_t d = density[i];
        if (normalize) {
            d = normalize_value(d, lo, scaling, flat);
            density[i] = uint8_t(d);
        }
        if (density_f32) density_f32[i] = float(d) / 255.0f;  // R8_UNORM
    }
}

// Cell (i,j,k) for i,j,k in [-1, n-1]: the 8 texels of the trilinear
// footprint whose lower corner is (i,j,k), texels outside the volume = 0
// (CLAMP_TO_BORDER with an opaque-black border).
__global__ void __launch_bounds__(256)
cells_kernel(const uint8_t* __restrict__ density, uint32_t nx, uint32_t ny, uint32_t nz,
             unsigned long long* __restrict__ cells) {
    uint32_t cx = nx + 1u, cy = ny + 1u, cz = nz + 1u;
    size_t total = size_t(cx) * cy * cz;
    size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x; c < total; c += stride) {
        cells[c] = pack_cell(density, nx, ny, nz, c);
    }
}

}  // namespace

int build_cells(const uint8_t* density, uint32_t nx, uint32_t ny, uint32_t nz, uint64_t* cells,
                hipStream_t s) {
    size_t total = size_t(nx + 1) * (ny + 1) * (nz + 1);
    cells_kernel<<<grid_for(total, 256, 65536), 256, 0, s>>>(density, nx, ny, nz,
                                                             reinterpret_cast<unsigned long long*>(cells));
    VKHR_CHECK_LAUNCH();
    return VKHR_OK;
}

}  // namespace vkhr

using namespace vkhr;

extern "C" int vkhr_volume_finalize(const uint32_t* counts, const int64_t* tangent_acc, uint32_t nx, uint32_t ny,
                                    uint32_t nz, int normalize, uint8_t* density_u8, int8_t* tangent_i8x4,
                                    float* density_f32, uint64_t* cells, vkhr_stream stream) {
    clear_error();
    int rc = ensure_device();
    if (rc) return rc;
    VKHR_REQUIRE(counts && density_u8, VKHR_ERR_INVALID, "counts/density_u8 is null");
    VKHR_REQUIRE(nx && ny && nz, VKHR_ERR_INVALID, "zero volume resolution");
    VKHR_REQUIRE(!tangent_i8x4 || tangent_acc, VKHR_ERR_INVALID, "tangent_i8x4 needs tangent_acc");
    size_t n = size_t(nx) * ny * nz;
    VKHR_REQUIRE(n <= (size_t(1) << 31), VKHR_ERR_OVERFLOW, "volume too large");
    hipStream_t s = as_stream(stream);
    uint32_t* mm = nullptr;
    if (normalize) {
        void* scr = nullptr;
        if ((rc = scratch(64, &scr))) return rc;
        mm = static_cast<uint32_t*>(scr);
        minmax_init_kernel<<<1, 64, 0, s>>>(mm);
        VKHR_CHECK_LAUNCH();
    }
    // A capped grid keeps the min/max atomics at one pair per block.
    unsigned grid = grid_for(n, 256, 2048);
    finalize_counts_kernel<<<grid, 256, 0, s>>>(counts, reinterpret_cast<const long long*>(tangent_acc), n, density_u8,
                                                tangent_i8x4, mm);
    VKHR_CHECK_LAUNCH();
    if (normalize || density_f32) {
        normalize_kernel<<<grid, 256, 0, s>>>(density_u8, n, mm, normalize, density_f32);
        VKHR_CHECK_LAUNCH();
    }
    if (cells && (rc = build_cells(density_u8, nx, ny, nz, cells, s))) return rc;
    return VKHR_OK;
}
