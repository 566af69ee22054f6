// Stable LSD radix sort of (uint64 or uint32 key, uint32 value) pairs (rtx_sort.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace rtx {

// Ping-pong buffers. The input is keys[0]/vals[0]; after radix_sort_pairs the sorted
// pairs are in keys[cur]/vals[cur]. K is uint64_t (BVH Morton codes) or uint32_t (rays).
template <class K>
struct SortBuffersT {
    K* keys[2];
    uint32_t* vals[2];
    uint32_t* hist;   // sort_hist_entries(n) entries: per-tile digit counts
    uint32_t* part;   // sort_part_entries(n) entries: digit totals
    int cur;
};
using SortBuffers = SortBuffersT<uint64_t>;
using SortBuffers32 = SortBuffersT<uint32_t>;

// Keys per tile of the passes; hist holds one count per digit and tile, part one total per
// digit.
constexpr uint32_t kSortTile = 4096;
constexpr size_t sort_hist_entries(uint32_t n) { return 256ull * ((n + (uint64_t)kSortTile - 1) / kSortTile); }
constexpr size_t sort_part_entries(uint32_t) { return 256; }

// Runs the 8-bit passes whose bit is set in pass_mask (pass p sorts bits 8p..8p+7).
// Skipping a pass whose digit is equal for every key leaves the result unchanged. Returns 0,
// or -1 when a launch failed.
int radix_sort_pairs(SortBuffers& S, uint32_t n, uint32_t pass_mask, hipStream_t st);
int radix_sort_pairs(SortBuffers32& S, uint32_t n, uint32_t pass_mask, hipStream_t st);

}  // namespace rtx
