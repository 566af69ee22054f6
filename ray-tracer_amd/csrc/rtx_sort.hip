// Stable LSD radix sort of (uint64 or uint32 key, uint32 value) pairs on 8-bit digits,
// shared by the builder (Morton codes, bvh_morton.cpp:109-159) and the ray reordering of
// traces.
//
// Each pass is three kernels over 4096-key tiles:
// - k_rs_count: the digit counts of every tile, digit-major (count[d][tile]);
// - k_rs_rowscan: one block per digit turns its row into exclusive prefixes over the tiles
//   and stores the row total;
// - k_rs_scatter: one block per tile ranks its keys stably (wave ballots and per-wave LDS
//   counts), places the tile in digit order in LDS and writes each digit's run to the start
//   of that digit (a scan of the row totals) plus the row prefix of its tile.
// No block waits on another. (A single-kernel pass chains each tile's offsets to its
// predecessors' through device-wide flags; with every tile of a 4M-key pass resident at once,
// that chain and the tile counter it needs cost more than the two extra launches.)
#include <hip/hip_runtime.h>

#include "rtx_sort.hpp"

namespace rtx {
namespace {

constexpr int kRsThreads = 256, kRsItems = 16, kRsTile = kRsThreads * kRsItems;
constexpr int kWaves = kRsThreads / 64;
constexpr int kScanItems = 4;   // row entries per thread and step of k_rs_rowscan
static_assert(kRsTile == (int)kSortTile, "scatter tile size");

// LDS words shared by the lanes of one wave: relaxed wave-scope atomics keep every access
// (a lane reads what another lane wrote) as a plain ds_* instruction; a wave's LDS operations
// complete in issue order.
template <class T>
__device__ __forceinline__ void wave_st(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}
template <class T>
__device__ __forceinline__ T wave_ld(T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

// Inclusive scan of one value per thread over the 256-thread block; wsum holds kWaves words,
// which after the call hold the waves' totals.
__device__ __forceinline__ uint32_t block_incl(uint32_t v, uint32_t* wsum) {
    const int lane = __lane_id(), w = threadIdx.x >> 6;
    uint32_t inc = v;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o);
        if (lane >= o) inc += y;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    for (int ww = 0; ww < w; ++ww) inc += wsum[ww];
    return inc;
}

// The keys of one 4096-key tile, kRsItems per lane: wave w holds the contiguous chunk
// [w * 1024, (w + 1) * 1024) of the tile, round r the 64 keys at r * 64. Every load is issued
// before any key is used.
template <class K>
__device__ __forceinline__ void load_chunk(const K* kin, uint32_t base, uint32_t n, K (&keys)[kRsItems]) {
    const int lane = __lane_id();
#pragma unroll
    for (int r = 0; r < kRsItems; ++r) {
        const uint32_t i = base + r * 64 + lane;
        keys[r] = i < n ? kin[i] : (K)0;
    }
}

// count[d * ntiles + tile] = keys of digit d in the tile.
template <class K>
__global__ void __launch_bounds__(kRsThreads) k_rs_count(const K* keys, uint32_t n, int shift, uint32_t* count,
                                                         uint32_t ntiles) {
    __shared__ uint32_t h[kWaves][256];   // per wave: fewer lanes collide on one counter
    const int w = threadIdx.x >> 6, d0 = threadIdx.x;
    for (int ww = 0; ww < kWaves; ++ww) h[ww][d0] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kRsTile + w * (kRsItems * 64);
    K k[kRsItems];
    load_chunk(keys, base, n, k);
#pragma unroll
    for (int r = 0; r < kRsItems; ++r)
        if (base + r * 64 + __lane_id() < n) atomicAdd(&h[w][(uint32_t)(k[r] >> shift) & 255u], 1u);
    __syncthreads();
    uint32_t c = 0;
    for (int ww = 0; ww < kWaves; ++ww) c += h[ww][d0];
    count[(size_t)d0 * ntiles + blockIdx.x] = c;
}

// Row d of count (one block per digit) to exclusive prefixes over the tiles; the row's total
// to total[d].
__global__ void __launch_bounds__(kRsThreads) k_rs_rowscan(uint32_t* count, uint32_t ntiles, uint32_t* total) {
    __shared__ uint32_t wsum[kWaves];
    uint32_t* row = count + (size_t)blockIdx.x * ntiles;
    uint32_t carry = 0;
    for (uint32_t b = 0; b < ntiles; b += kRsThreads * kScanItems) {
        const uint32_t i0 = b + threadIdx.x * kScanItems;
        uint32_t v[kScanItems], s = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            v[k] = i0 + k < ntiles ? row[i0 + k] : 0u;
            s += v[k];
        }
        uint32_t run = carry + block_incl(s, wsum) - s;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            if (i0 + k < ntiles) row[i0 + k] = run;
            run += v[k];
        }
        for (int ww = 0; ww < kWaves; ++ww) carry += wsum[ww];
        __syncthreads();   // wsum is rewritten by the next step
    }
    if (threadIdx.x == 0) total[blockIdx.x] = carry;
}

// Scatter of one pass (digit at bit `shift`), one block per tile. Each wave ranks its own
// contiguous chunk of the tile (kRsItems rounds of 64 keys): within a round, ballots give
// each lane its rank among the lanes holding the same digit, and a per-wave LDS count per
// digit carries the ranks from round to round, so the rounds need no block barrier. A key's
// place in the tile is the tile's start of its digit, plus the counts of that digit in the
// earlier waves' chunks, plus its rank in its own chunk: index order within each digit, so
// the sort is stable. The tile is placed in LDS in digit order and written out slot by slot,
// so consecutive lanes store consecutive elements of a digit's run.
template <class K>
__global__ void __launch_bounds__(kRsThreads) k_rs_scatter(const K* kin, const uint32_t* vin, K* kout,
                                                           uint32_t* vout, uint32_t n, int shift,
                                                           const uint32_t* count, const uint32_t* total,
                                                           uint32_t ntiles) {
    __shared__ uint32_t wcnt[kWaves][256];   // per wave: digit-d keys ranked so far, then the wave's offset in the digit's run
    __shared__ uint32_t gbase[256];          // global position of this tile's first digit-d element
    __shared__ uint32_t lstart[256];         // tile-local position of the first digit-d element
    __shared__ uint32_t wsum[2][kWaves];
    __shared__ K skey[kRsTile];
    __shared__ uint32_t sval[kRsTile];
    const int lane = __lane_id(), w = threadIdx.x >> 6, d0 = threadIdx.x;
    for (int ww = 0; ww < kWaves; ++ww) wcnt[ww][d0] = 0;
    // this tile's digit offsets do not depend on the ranking: load them first
    const uint32_t tot = total[d0], pre = count[(size_t)d0 * ntiles + blockIdx.x];
    __syncthreads();
    const uint64_t lt = (1ull << lane) - 1ull;
    const uint32_t base = blockIdx.x * kRsTile + w * (kRsItems * 64);   // this wave's chunk
    K keys[kRsItems];
    uint32_t vals[kRsItems], lrank[kRsItems];
    load_chunk(kin, base, n, keys);
#pragma unroll
    for (int r = 0; r < kRsItems; ++r) {
        const uint32_t i = base + r * 64 + lane;
        vals[r] = i < n ? vin[i] : 0u;
    }
#pragma unroll
    for (int r = 0; r < kRsItems; ++r) {
        const bool valid = base + r * 64 + lane < n;
        const uint32_t d = (uint32_t)(keys[r] >> shift) & 255u;
        uint64_t mask = __ballot(valid);
        for (int bit = 0; bit < 8; ++bit) {
            const uint64_t bb = __ballot(valid && ((d >> bit) & 1u));
            mask &= ((d >> bit) & 1u) ? bb : ~bb;
        }
        const uint32_t rank = __popcll(mask & lt);
        // every lane reads its digit's count before the first lane of each digit writes it (a
        // wave's LDS accesses run in issue order)
        const uint32_t before = valid ? wave_ld(&wcnt[w][d]) : 0u;
        if (valid && rank == 0) wave_st(&wcnt[w][d], before + (uint32_t)__popcll(mask));
        lrank[r] = before + rank;
    }
    __syncthreads();
    // digit d0 = this thread: the waves' counts become offsets, their sum the tile's count
    uint32_t c = 0;
    for (int ww = 0; ww < kWaves; ++ww) {
        const uint32_t x = wcnt[ww][d0];
        wcnt[ww][d0] = c;
        c += x;
    }
    // global digit starts and tile-local digit starts (exclusive scans over the digits)
    gbase[d0] = block_incl(tot, wsum[0]) - tot + pre;
    lstart[d0] = block_incl(c, wsum[1]) - c;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRsItems; ++r) {
        const uint32_t i = base + r * 64 + lane;
        if (i < n) {
            const uint32_t d = (uint32_t)(keys[r] >> shift) & 255u;
            const uint32_t slot = lstart[d] + wcnt[w][d] + lrank[r];
            skey[slot] = keys[r];
            sval[slot] = vals[r];
        }
    }
    __syncthreads();
    const uint32_t tb = blockIdx.x * kRsTile;
    const uint32_t m = n - tb < (uint32_t)kRsTile ? n - tb : (uint32_t)kRsTile;
    for (uint32_t j = threadIdx.x; j < m; j += kRsThreads) {
        const K k = skey[j];
        const uint32_t d = (uint32_t)(k >> shift) & 255u;
        const uint32_t pos = gbase[d] + (j - lstart[d]);
        kout[pos] = k;
        vout[pos] = sval[j];
    }
}

}  // namespace

template <class K>
int sort_impl(SortBuffersT<K>& S, uint32_t n, uint32_t pass_mask, hipStream_t st) {
    S.cur = 0;
    if (n == 0) return 0;
    const uint32_t ntiles = (n + kRsTile - 1) / kRsTile;
    for (int pass = 0; pass < (int)sizeof(K); ++pass) {
        if (!((pass_mask >> pass) & 1u)) continue;
        const int shift = 8 * pass;
        k_rs_count<K><<<ntiles, kRsThreads, 0, st>>>(S.keys[S.cur], n, shift, S.hist, ntiles);
        k_rs_rowscan<<<256, kRsThreads, 0, st>>>(S.hist, ntiles, S.part);
        k_rs_scatter<K><<<ntiles, kRsThreads, 0, st>>>(S.keys[S.cur], S.vals[S.cur], S.keys[S.cur ^ 1],
                                                       S.vals[S.cur ^ 1], n, shift, S.hist, S.part, ntiles);
        S.cur ^= 1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int radix_sort_pairs(SortBuffers& S, uint32_t n, uint32_t pass_mask, hipStream_t st) {
    return sort_impl(S, n, pass_mask, st);
}
int radix_sort_pairs(SortBuffers32& S, uint32_t n, uint32_t pass_mask, hipStream_t st) {
    return sort_impl(S, n, pass_mask, st);
}

}  // namespace rtx
