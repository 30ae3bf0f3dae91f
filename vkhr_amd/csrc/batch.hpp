// Device tables of the batched launch: many hair styles of uneven size,
// built into volumes by the LDS-binned passes of the one-call build (each
// pass one launch over the concatenated workgroups of all styles) and
// raymarched over (style, tile) work items, all in one graph.
// The per-pair, per-brick, per-cell and per-pixel bodies are the per-style
// ones (build.hpp, volume.hpp, march.hpp), so the outputs are bit-identical
// to per-style launches.
#pragma once

#include "build.hpp"
#include "march.hpp"

namespace vkhr {

// A style of the batch: the arguments of its build passes, and where its
// cells pass writes the corner-packed cells and the normalized volume.
struct BatchVolStyle {
    BuildArgs a;
    unsigned long long* cells;
    uint8_t* density;
};

// A contiguous range of one style's lines (of one distance pass).
struct BatchChunk {
    uint32_t style;
    uint32_t reserved;
    uint64_t begin, end;
};

// Workgroup offsets of one pass over the styles: style s owns workgroups
// [prefix[s], prefix[s + 1]) of the pass's launch (n_styles + 1 entries).
enum BatchPass { kPassInit, kPassBins, kPassBricks, kPassFinalize, kPassCells, kPassCount };

struct BatchDevice {
    const BatchVolStyle* styles;
    RaymarchArgs* args;             // per style
    uint32_t* tile_prefix;          // n_styles + 1
    uint32_t* block_prefix[kPassCount];
    uint32_t blocks[kPassCount];    // workgroups of each pass (its prefix's last entry)
    BatchChunk* line_chunks[3];     // four lines of bricks per chunk, per distance pass
    uint32_t n_line_chunks[3];
    uint32_t march_lds_bytes;       // the largest distance field the raymarch stages
    uint32_t* work_counter;
    uint32_t n_styles;
    uint32_t total_tiles;
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
