// Stable LSD radix sort of (uint64 key, uint32 value) pairs on 8-bit digits, shared by
// the builder (Morton codes, bvh_morton.cpp:109-159) and the ray reordering of traces.
//
// Per pass: k_rs_hist counts digits per 4096-key tile in LDS; a three-kernel scan turns
// the digit-major histogram into global offsets; k_rs_scatter ranks keys stably inside a
// tile with wave ballots (lanes holding the same digit) and per-wave LDS counts, then
// writes the tile out in digit order from LDS.
#include <hip/hip_runtime.h>

#include "rtx_sort.hpp"

namespace rtx {
namespace {

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

constexpr int kRsThreads = 256, kRsItems = 16, kRsTile = kRsThreads * kRsItems;
constexpr int kScanThreads = 1024, kScanItems = 4, kScanSeg = kScanThreads * kScanItems;

template <class K>
__global__ void k_rs_hist(const K* keys, uint32_t n, int shift, uint32_t* hist, uint32_t ntiles) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint32_t base = blockIdx.x * kRsTile;
    for (int r = 0; r < kRsItems; ++r) {
        uint32_t i = base + r * kRsThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Block-wide exclusive scan (1024 threads); returns the block total in *total.
__device__ uint32_t block_excl(uint32_t v, uint32_t* ws, uint32_t* total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t s = v;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(s, o);
        if (lane >= o) s += y;
    }
    if (lane == 63) ws[w] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int i = 0; i < kScanThreads / 64; ++i) { uint32_t x = ws[i]; ws[i] = run; run += x; }
        ws[kScanThreads / 64] = run;
    }
    __syncthreads();
    uint32_t ex = ws[w] + s - v;
    *total = ws[kScanThreads / 64];
    __syncthreads();
    return ex;
}

// Exclusive scan of segments of kScanSeg entries in place; segment totals to part[].
__global__ void __launch_bounds__(kScanThreads) k_scan_seg(uint32_t* data, uint32_t m, uint32_t* part) {
    __shared__ uint32_t ws[kScanThreads / 64 + 1];
    const uint32_t b = blockIdx.x * kScanSeg + threadIdx.x * kScanItems;
    uint32_t v[kScanItems], s = 0;
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (b + k < m) ? data[b + k] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t run = block_excl(s, ws, &tot);
    for (int k = 0; k < kScanItems; ++k) {
        if (b + k < m) data[b + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// Exclusive scan of the segment totals (any count, one block, serial per thread chunk).
__global__ void __launch_bounds__(kScanThreads) k_scan_parts(uint32_t* part, uint32_t np) {
    __shared__ uint32_t ws[kScanThreads / 64 + 1];
    const uint32_t chunk = (np + kScanThreads - 1) / kScanThreads;
    const uint32_t b = threadIdx.x * chunk, e = min(np, b + chunk);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += part[i];
    uint32_t tot;
    uint32_t run = block_excl(s, ws, &tot);
    for (uint32_t i = b; i < e; ++i) { uint32_t x = part[i]; part[i] = run; run += x; }
}

__global__ void __launch_bounds__(kScanThreads) k_scan_add(uint32_t* data, uint32_t m, const uint32_t* part) {
    const uint32_t add = part[blockIdx.x];
    const uint32_t b = blockIdx.x * kScanSeg + threadIdx.x * kScanItems;
    for (int k = 0; k < kScanItems; ++k)
        if (b + k < m) data[b + k] += add;
}

// Scatter of one pass. Each wave ranks its own contiguous chunk of the tile (kRsItems rounds
// of 64 keys): within a round, ballots give each lane its rank among the lanes holding the
// same digit, and a per-wave LDS count per digit carries the ranks from round to round, so
// the rounds need no block barrier. A key's place in the tile is then: the tile's start of
// its digit, plus the counts of that digit in the earlier waves' chunks, plus its rank in its
// own chunk; that is index order within each digit, so the sort is stable. The tile is placed
// in LDS in digit order and written out slot by slot, so consecutive lanes store consecutive
// elements of a digit's run (4-byte stores to unrelated addresses cost a memory transaction
// each). (The earlier form interleaved the waves' rounds and needed three block barriers per
// round.)
template <class K>
__global__ void __launch_bounds__(kRsThreads) k_rs_scatter(const K* kin, const uint32_t* vin, K* kout,
                                                           uint32_t* vout, uint32_t n, int shift,
                                                           const uint32_t* hist, uint32_t ntiles) {
    constexpr int kWaves = kRsThreads / 64;
    __shared__ uint32_t wcnt[kWaves][256];   // per wave: digit-d keys ranked so far, then the wave's offset in the digit's run
    __shared__ uint32_t gbase[256];          // global position of this tile's first digit-d element
    __shared__ uint32_t lstart[256];         // tile-local position of the first digit-d element
    __shared__ uint32_t wsum[kWaves];
    __shared__ K skey[kRsTile];
    __shared__ uint32_t sval[kRsTile];
    const int lane = __lane_id(), w = threadIdx.x >> 6;
    gbase[threadIdx.x] = hist[threadIdx.x * ntiles + blockIdx.x];
    for (int ww = 0; ww < kWaves; ++ww) wcnt[ww][threadIdx.x] = 0;
    __syncthreads();
    const uint64_t lt = (1ull << lane) - 1ull;
    const uint32_t base = blockIdx.x * kRsTile + w * (kRsItems * 64);   // this wave's chunk
    K keys[kRsItems];
    uint32_t vals[kRsItems], lrank[kRsItems];
#pragma unroll
    for (int r = 0; r < kRsItems; ++r) {
        const uint32_t i = base + r * 64 + lane;
        const bool valid = i < n;
        const K k = valid ? kin[i] : (K)0;
        const uint32_t v = valid ? vin[i] : 0u;
        const uint32_t d = (uint32_t)(k >> shift) & 255u;
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
        keys[r] = k;
        vals[r] = v;
        lrank[r] = before + rank;
    }
    __syncthreads();
    // digit d = this thread: the waves' counts become offsets, their sum the tile's count
    uint32_t c = 0;
    for (int ww = 0; ww < kWaves; ++ww) {
        const uint32_t x = wcnt[ww][threadIdx.x];
        wcnt[ww][threadIdx.x] = c;
        c += x;
    }
    // tile-local digit starts: exclusive scan of the tile's digit counts (one digit per thread)
    uint32_t inc = c;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o);
        if (lane >= o) inc += y;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t off = 0;
    for (int ww = 0; ww < w; ++ww) off += wsum[ww];
    lstart[threadIdx.x] = off + inc - c;
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

size_t sort_hist_entries(uint32_t n) { return 256ull * ((n + kRsTile - 1) / kRsTile); }
size_t sort_part_entries(uint32_t n) { return (sort_hist_entries(n) + kScanSeg - 1) / kScanSeg + 1; }

template <class K>
int sort_impl(SortBuffersT<K>& S, uint32_t n, uint32_t pass_mask, hipStream_t st) {
    S.cur = 0;
    if (n == 0) return 0;
    const uint32_t ntiles = (n + kRsTile - 1) / kRsTile;
    const uint32_t m = 256 * ntiles;
    const uint32_t nseg = (m + kScanSeg - 1) / kScanSeg;
    for (int pass = 0; pass < (int)sizeof(K); ++pass) {
        if (!((pass_mask >> pass) & 1u)) continue;
        const int shift = 8 * pass;
        k_rs_hist<K><<<ntiles, kRsThreads, 0, st>>>(S.keys[S.cur], n, shift, S.hist, ntiles);
        k_scan_seg<<<nseg, kScanThreads, 0, st>>>(S.hist, m, S.part);
        k_scan_parts<<<1, kScanThreads, 0, st>>>(S.part, nseg);
        k_scan_add<<<nseg, kScanThreads, 0, st>>>(S.hist, m, S.part);
        k_rs_scatter<K><<<ntiles, kRsThreads, 0, st>>>(S.keys[S.cur], S.vals[S.cur], S.keys[S.cur ^ 1],
                                                       S.vals[S.cur ^ 1], n, shift, S.hist, ntiles);
        S.cur ^= 1;
    }
    return 0;
}

int radix_sort_pairs(SortBuffers& S, uint32_t n, uint32_t pass_mask, hipStream_t st) {
    return sort_impl(S, n, pass_mask, st);
}
int radix_sort_pairs(SortBuffers32& S, uint32_t n, uint32_t pass_mask, hipStream_t st) {
    return sort_impl(S, n, pass_mask, st);
}

}  // namespace rtx
