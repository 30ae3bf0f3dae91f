// Device tables of the batched launch: many hair styles of uneven size,
// voxelized, finalized and raymarched by one graph of grid-stride kernels
// over chunk lists (and a raymarch over (style, tile) work items).
// The per-segment, per-voxel and per-pixel bodies are the per-style ones
// (voxel.hpp, volume.hpp, march.hpp), so the outputs are bit-identical to
// per-style launches.
#pragma once

#include "march.hpp"

namespace vkhr {

struct BatchVolStyle {
    float ox, oy, oz, vx, vy, vz, mx, my, mz;   // the VoxelGrid of the style
    uint32_t nx, ny, nz;
    const float* xyz;
    const uint32_t* indices;
    const float* tangents;
    uint32_t n_vertices, n_pairs;
    uint32_t* counts;
    long long* acc;
    uint8_t* density;
    int8_t* tangent_q;
    unsigned long long* cells;
    uint32_t* occupancy;                        // nullable
};

// A contiguous range of one style's voxels (or one cells_block of its cells).
struct BatchChunk {
    uint32_t style;
    uint32_t reserved;
    uint64_t begin, end;
};

struct BatchDevice {
    BatchVolStyle* styles;
    RaymarchArgs* args;             // per style
    uint64_t* pair_prefix;          // n_styles + 1
    uint32_t* tile_prefix;          // n_styles + 1
    BatchChunk* voxel_chunks;
    BatchChunk* cell_chunks;
    BatchChunk* line_chunks[3];     // four lines of bricks per chunk, per distance pass
    uint32_t n_line_chunks[3];
    uint32_t march_lds_bytes;       // the largest distance field the raymarch stages
    uint32_t* minmax;               // 2 per style
    uint32_t* work_counter;
    uint32_t n_styles;
    uint32_t n_voxel_chunks, n_cell_chunks;
    uint32_t total_tiles;
    uint64_t total_pairs;
    uint32_t tile_w, tile_h;        // pixels per wave tile (tile_w * tile_h == 64)
    uint32_t tiles_x;               // block tiles per image row
};

// Largest i with prefix[i] <= x (prefix[0] == 0, n + 1 entries).
template <class T>
__device__ __forceinline__ uint32_t style_of(const T* __restrict__ prefix, uint32_t n, T x) {
    uint32_t lo = 0, hi = n;  // answer in [lo, hi)
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) / 2;
        if (prefix[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

void launch_batch_volume(const BatchDevice& d, hipStream_t stream);
void launch_batch_march(const BatchDevice& d, unsigned grid, hipStream_t stream);

}  // namespace vkhr
