// Batched voxelization and finalization of many styles (compiled with
// -ffp-contract=off like the per-style path: the same DDA and finalize
// bodies, so counts, tangents, densities and cells are bit-identical to
// per-style launches).
#include "batch.hpp"
#include "volume.hpp"
#include "voxel.hpp"

namespace vkhr {
namespace {

__device__ __forceinline__ VoxelGrid grid_of(const BatchVolStyle& s) {
    VoxelGrid g;
    g.ox = s.ox; g.oy = s.oy; g.oz = s.oz;
    g.vx = s.vx; g.vy = s.vy; g.vz = s.vz;
    g.mx = s.mx; g.my = s.my; g.mz = s.mz;
    g.nx = s.nx; g.ny = s.ny; g.nz = s.nz;
    return g;
}

// Zeroes every style's counts or tangent accumulator and occupancy, resets the per-style
// min/max and the raymarch work queue (vector stores: 16 B per voxel of the
// accumulator, 4 counts per store).
__global__ void __launch_bounds__(256)
batch_clear_kernel(BatchDevice d) {
    if (blockIdx.x == 0) {
        for (uint32_t s = threadIdx.x; s < d.n_styles; s += blockDim.x) {
            d.minmax[2 * s] = 255u;
            d.minmax[2 * s + 1] = 0u;
        }
        if (threadIdx.x == 0) *d.work_counter = 0u;
    }
    for (uint32_t s = blockIdx.x; s < d.n_styles; s += gridDim.x) {
        const BatchVolStyle& st = d.styles[s];
        if (!st.occupancy) continue;
        const size_t words = occupancy_words(st.nx, st.ny, st.nz);
        for (size_t w = threadIdx.x; w < words; w += blockDim.x) st.occupancy[w] = 0u;
    }
    for (uint32_t c = blockIdx.x; c < d.n_voxel_chunks; c += gridDim.x) {
        const BatchChunk ch = d.voxel_chunks[c];
        const BatchVolStyle& st = d.styles[ch.style];
        // with the packed accumulator the counts are unpacked by finalize
        // (when they are asked for)
        if (st.acc) {
            longlong2* a = reinterpret_cast<longlong2*>(st.acc);
            for (uint64_t i = ch.begin + threadIdx.x; i < ch.end; i += blockDim.x) a[i] = make_longlong2(0, 0);
        } else {
            const uint64_t q1 = ch.end / 4;  // chunks start at multiples of 4
            for (uint64_t q = ch.begin / 4 + threadIdx.x; q < q1; q += blockDim.x)
                reinterpret_cast<uint4*>(st.counts)[q] = make_uint4(0u, 0u, 0u, 0u);
            for (uint64_t i = 4 * q1 + threadIdx.x; i < ch.end; i += blockDim.x) st.counts[i] = 0u;
        }
    }
}

// Grid-stride over wave-aligned ranges of the concatenated segments: the
// lanes of a wave may belong to different styles.
__global__ void __launch_bounds__(256)
batch_voxelize_kernel(BatchDevice d) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t base = uint64_t(blockIdx.x) * blockDim.x + (threadIdx.x & ~uint32_t(kWave - 1));
         base < d.total_pairs; base += stride) {
        const uint64_t g = base + (threadIdx.x & uint32_t(kWave - 1));
        const bool has = g < d.total_pairs;
        const uint32_t s = has ? style_of(d.pair_prefix, d.n_styles, g) : 0u;
        const BatchVolStyle& st = d.styles[s];
        const uint32_t p = has ? uint32_t(g - d.pair_prefix[s]) : 0u;
        voxelize_segment(grid_of(st), st.xyz, st.n_vertices, has, has ? st.indices[2u * p] : 0u,
                         has ? st.indices[2u * p + 1u] : 0u, st.tangents, st.counts, st.acc);
    }
}

__global__ void __launch_bounds__(256)
batch_finalize_kernel(BatchDevice d) {
    __shared__ uint32_t red[2][256 / kWave];
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    for (uint32_t c = blockIdx.x; c < d.n_voxel_chunks; c += gridDim.x) {
        const BatchChunk ch = d.voxel_chunks[c];
        const BatchVolStyle& st = d.styles[ch.style];
        const uint32_t* counts_in = st.acc ? nullptr : st.counts;
        uint32_t* counts_out = st.acc ? st.counts : nullptr;
        uint32_t lo = 255u, hi = 0u;
        const uint64_t q1 = ch.end / 4;  // chunks start at multiples of 4
        for (uint64_t q = ch.begin / 4 + threadIdx.x; q < q1; q += blockDim.x)
            finalize_quad(counts_in, st.acc, 4 * q, st.density, st.tangent_q, &lo, &hi, counts_out);
        for (uint64_t i = 4 * q1 + threadIdx.x; i < ch.end; i += blockDim.x) {
            const uint32_t v = finalize_voxel(counts_in, st.acc, i, st.density, st.tangent_q);
            if (counts_out) counts_out[i] = packed_count(st.acc, i);
            lo = v < lo ? v : lo;
            hi = v > hi ? v : hi;
        }
        for (int o = 32; o > 0; o >>= 1) {
            uint32_t a = __shfl_xor(lo, o, kWave), b = __shfl_xor(hi, o, kWave);
            lo = a < lo ? a : lo;
            hi = b > hi ? b : hi;
        }
        if (lane == 0) { red[0][wid] = lo; red[1][wid] = hi; }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < int(blockDim.x / kWave); ++w) {
                lo = red[0][w] < lo ? red[0][w] : lo;
                hi = red[1][w] > hi ? red[1][w] : hi;
            }
            atomicMin(&d.minmax[2 * ch.style], lo);
            atomicMax(&d.minmax[2 * ch.style + 1], hi);
        }
        __syncthreads();
    }
}

// Corner-packed cells of the normalized density and the brick occupancy:
// a cell chunk is one cells_block of one style.
__global__ void __launch_bounds__(256)
batch_cells_kernel(BatchDevice d) {
    for (uint32_t c = blockIdx.x; c < d.n_cell_chunks; c += gridDim.x) {
        const BatchChunk ch = d.cell_chunks[c];
        const BatchVolStyle& st = d.styles[ch.style];
        cells_block(st.density, st.nx, st.ny, st.nz, uint32_t(ch.begin), st.cells, st.occupancy);
    }
}

// One pass of every style's brick distance field: a line chunk is four
// lines (one per wave) of one style.
__global__ void __launch_bounds__(256)
batch_distance_kernel(BatchDevice d, int axis) {
    __shared__ uint32_t lds[256 / kWave][kDistanceLineMax];
    const uint32_t wave = threadIdx.x / kWave;
    for (uint32_t c = blockIdx.x; c < d.n_line_chunks[axis]; c += gridDim.x) {
        const BatchChunk ch = d.line_chunks[axis][c];
        const BatchVolStyle& st = d.styles[ch.style];
        const uint32_t line = uint32_t(ch.begin) + wave;
        distance_line(st.occupancy, st.nx, st.ny, st.nz, axis, line, line < uint32_t(ch.end), lds[wave]);
        __syncthreads();   // the LDS rows are reused by the next chunk
    }
}

__global__ void __launch_bounds__(256)
batch_normalize_kernel(BatchDevice d) {
    for (uint32_t c = blockIdx.x; c < d.n_voxel_chunks; c += gridDim.x) {
        const BatchChunk ch = d.voxel_chunks[c];
        const BatchVolStyle& st = d.styles[ch.style];
        const DensityMap m = density_map(&d.minmax[2 * ch.style], true);
        uchar4* d4 = reinterpret_cast<uchar4*>(st.density);
        const uint64_t q1 = ch.end / 4;  // chunks start at multiples of 4
        for (uint64_t q = ch.begin / 4 + threadIdx.x; q < q1; q += blockDim.x) {
            uchar4 v = d4[q];
            v.x = uint8_t(texel_density(v.x, m));
            v.y = uint8_t(texel_density(v.y, m));
            v.z = uint8_t(texel_density(v.z, m));
            v.w = uint8_t(texel_density(v.w, m));
            d4[q] = v;
        }
        for (uint64_t i = 4 * q1 + threadIdx.x; i < ch.end; i += blockDim.x)
            st.density[i] = uint8_t(texel_density(st.density[i], m));
    }
}

}  // namespace

void launch_batch_volume(const BatchDevice& d, hipStream_t stream) {
    batch_clear_kernel<<<4096, 256, 0, stream>>>(d);
    batch_voxelize_kernel<<<grid_for(d.total_pairs, 256, 65536), 256, 0, stream>>>(d);
    batch_finalize_kernel<<<4096, 256, 0, stream>>>(d);
    batch_normalize_kernel<<<4096, 256, 0, stream>>>(d);
    batch_cells_kernel<<<8192, 256, 0, stream>>>(d);
    for (int axis = 0; axis < 3; ++axis)
        if (d.n_line_chunks[axis]) batch_distance_kernel<<<grid_for(d.n_line_chunks[axis], 1, 4096), 256, 0, stream>>>(d, axis);
}

}  // namespace vkhr
