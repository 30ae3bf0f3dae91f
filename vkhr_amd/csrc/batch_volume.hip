// Batched volume build of many styles: the passes of the one-call build
// (build.hpp) -- init, bin count, scan, bin fill, bricks, the workspace
// finalize, cells -- each one launch over the concatenated workgroups of
// all styles, then the occupancy bitmask and brick distance field of every
// style.  Compiled with -ffp-contract=off like the per-style path: the same
// DDA, brick and cells bodies, so counts, tangents, densities and cells are
// bit-identical to per-style launches.
#include "batch.hpp"

namespace vkhr {
namespace {

// The style owning workgroup x of a pass (d.block_prefix[pass] holds the
// styles' workgroup offsets): each lane compares one offset, so one load
// round trip per 64 styles; styles with no workgroups are skipped.  The
// result is the number of styles that end at or before x, wave-uniform.
__device__ __forceinline__ uint32_t block_style(const uint32_t* __restrict__ prefix, uint32_t n, uint32_t x) {
    const uint32_t lane = threadIdx.x % kWave;
    uint32_t s = 0;
    for (uint32_t base = 1; base <= n; base += kWave) {
        const uint32_t i = base + lane;
        const unsigned long long ended = __ballot(i <= n && prefix[i] <= x);
        s += uint32_t(__popcll(ended));
        if (ended != ~0ull) break;  // the offsets are sorted: no later style has ended either
    }
    return uint32_t(__builtin_amdgcn_readfirstlane(int(s)));
}

// A workgroup's style, its index within that style's workgroups of the
// pass, and their number.
struct BlockOf {
    uint32_t style, block, blocks;
};

__device__ __forceinline__ BlockOf block_of(const BatchDevice& d, BatchPass pass) {
    const uint32_t* prefix = d.block_prefix[pass];
    const uint32_t s = block_style(prefix, d.n_styles, blockIdx.x);
    const uint32_t first = prefix[s];
    return {s, blockIdx.x - first, prefix[s + 1] - first};
}

__global__ void __launch_bounds__(256)
batch_init_kernel(BatchDevice d, const BatchVolStyle* __restrict__ styles) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *d.work_counter = 0u;  // the persistent raymarch's queue
    const BlockOf w = block_of(d, kPassInit);
    build_init_body(styles[w.style].a, w.block, w.blocks);
}

__global__ void __launch_bounds__(kBuildBlock)
batch_bin_count_kernel(BatchDevice d, const BatchVolStyle* __restrict__ styles) {
    __shared__ BinTable t;
    const BlockOf w = block_of(d, kPassBins);
    bin_count_body(styles[w.style].a, t, w.block);
}

// One workgroup per style.
__global__ void __launch_bounds__(kScanThreads)
batch_scan_kernel(const BatchVolStyle* __restrict__ styles) {
    __shared__ ScanLds l;
    scan_body(styles[blockIdx.x].a, l);
}

__global__ void __launch_bounds__(kBuildBlock)
batch_bin_fill_kernel(BatchDevice d, const BatchVolStyle* __restrict__ styles) {
    __shared__ BinTable t;
    const BlockOf w = block_of(d, kPassBins);
    const BuildArgs& a = styles[w.style].a;
    bin_fill_body(a, t, w.block, w.blocks, size_t(a.g.nx) * a.g.ny * a.g.nz);
}

// The styles' brick workgroups are multiples of 8, so a workgroup's index
// within its style keeps the XCD of its index in the launch.
__global__ void __launch_bounds__(kBrickBlock)
batch_brick_kernel(BatchDevice d, const BatchVolStyle* __restrict__ styles) {
    __shared__ BrickLds l;
    const BlockOf w = block_of(d, kPassBricks);
    brick_body(styles[w.style].a, l, w.block, w.blocks);
}

__global__ void __launch_bounds__(256)
batch_overflow_finalize_kernel(BatchDevice d, const BatchVolStyle* __restrict__ styles) {
    __shared__ uint32_t red[2][256 / kWave];
    const BlockOf w = block_of(d, kPassFinalize);
    const BuildArgs& a = styles[w.style].a;
    overflow_finalize_body(a, size_t(a.g.nx) * a.g.ny * a.g.nz, w.block, w.blocks, red);
}

// The corner-packed cells with the normalize applied as the uint8 view
// loads (the style's map from its min/max slots, built per workgroup),
// writing the normalized volume; one workgroup of cells_pass each.
__global__ void __launch_bounds__(256)
batch_cells_kernel(BatchDevice d, const BatchVolStyle* __restrict__ styles) {
    __shared__ uint8_t lut[256];
    const BlockOf w = block_of(d, kPassCells);
    const BatchVolStyle& st = styles[w.style];
    const BuildArgs& a = st.a;
    unsigned long long* cells = st.cells;
    uint8_t* density = st.density;
    if (build_lut(a.minmax, true, lut))
        cells_pass<1>(a.density, a.g.nx, a.g.ny, a.g.nz, w.block, cells, a.occupancy, nullptr, density);
    else
        cells_pass<2>(a.density, a.g.nx, a.g.ny, a.g.nz, w.block, cells, a.occupancy, lut, density);
}

// The bitmask of every style's occupancy from its brick flags: one
// workgroup per style.
__global__ void __launch_bounds__(256)
batch_pack_bits_kernel(const BatchVolStyle* __restrict__ styles) {
    const BuildArgs& a = styles[blockIdx.x].a;
    if (!a.occupancy) return;
    const uint32_t words = uint32_t(occupancy_bit_words(a.g.nx, a.g.ny, a.g.nz));
    for (uint32_t w = threadIdx.x; w < words; w += blockDim.x) pack_brick_bits(a.occupancy, a.g.nx, a.g.ny, a.g.nz, w);
}

// One pass of every style's brick distance field: a line chunk is four
// lines (one per wave) of one style.
__global__ void __launch_bounds__(256)
batch_distance_kernel(BatchDevice d, const BatchVolStyle* __restrict__ styles, int axis) {
    __shared__ uint32_t lds[256 / kWave][kDistanceLineMax];
    const uint32_t wave = threadIdx.x / kWave;
    for (uint32_t c = blockIdx.x; c < d.n_line_chunks[axis]; c += gridDim.x) {
        const BatchChunk ch = d.line_chunks[axis][c];
        const BuildArgs& a = styles[ch.style].a;
        const uint32_t line = uint32_t(ch.begin) + wave;
        distance_line(a.occupancy, a.g.nx, a.g.ny, a.g.nz, axis, line, line < uint32_t(ch.end), lds[wave]);
        __syncthreads();   // the LDS rows are reused by the next chunk
    }
}

}  // namespace

void launch_batch_volume(const BatchDevice& d, hipStream_t stream) {
    const BatchVolStyle* st = d.styles;
    batch_init_kernel<<<d.blocks[kPassInit], 256, 0, stream>>>(d, st);
    if (d.blocks[kPassBins]) batch_bin_count_kernel<<<d.blocks[kPassBins], kBuildBlock, 0, stream>>>(d, st);
    batch_scan_kernel<<<d.n_styles, kScanThreads, 0, stream>>>(st);
    if (d.blocks[kPassBins]) batch_bin_fill_kernel<<<d.blocks[kPassBins], kBuildBlock, 0, stream>>>(d, st);
    batch_brick_kernel<<<d.blocks[kPassBricks], kBrickBlock, 0, stream>>>(d, st);
    batch_overflow_finalize_kernel<<<d.blocks[kPassFinalize], 256, 0, stream>>>(d, st);
    batch_cells_kernel<<<d.blocks[kPassCells], 256, 0, stream>>>(d, st);
    batch_pack_bits_kernel<<<d.n_styles, 256, 0, stream>>>(st);
    for (int axis = 0; axis < 3; ++axis)
        if (d.n_line_chunks[axis])
            batch_distance_kernel<<<grid_for(d.n_line_chunks[axis], 1, 4096), 256, 0, stream>>>(d, st, axis);
}

}  // namespace vkhr
