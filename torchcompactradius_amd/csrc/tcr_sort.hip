// Stable LSD radix sort of (cell key, point index) pairs into cell order, one
// segment per cloud, wave64 multi-split ranking in LDS.  Replaces
// torch.argsort of the int32 cell keys (py/compactHashing/cellTable.py:33)
// with its stable contract: inside a cell, ascending original index.
//
// A tile is kSortTile keys of one cloud (tiles never straddle clouds).  The
// per-tile digit histograms are laid out cloud-major, then digit, then tile,
// so one exclusive scan over the whole histogram array yields the global
// destination of every (cloud, digit, tile) run: the batch stays segmented by
// cloud without any per-cloud launch.
#include "tcr_internal.h"

namespace tcr {

struct TileInfo {
    int64_t start;   // first global key of the tile
    int32_t len;     // keys in the tile
    int64_t hbase;   // histogram base of the cloud (kRadix * first tile)
    int64_t ntiles;  // tiles of the cloud
    int64_t lt;      // tile index inside the cloud
};

__device__ __forceinline__ TileInfo tile_info(int64_t t, int32_t B, const int64_t *__restrict__ t_start,
                                              const int64_t *__restrict__ p_start) {
    // Empty clouds own zero tiles; find the last cloud whose first tile <= t
    // and which owns tiles.
    const int c = find_segment(B, t, [&](int k) { return t_start[k]; });
    // clouds with zero tiles share t_start with their successor: advance.
    int cc = c;
    TileInfo ti;
    const int64_t lt = t - t_start[cc];
    const int64_t n = p_start[cc + 1] - p_start[cc];
    ti.start = p_start[cc] + lt * kSortTile;
    const int64_t rem = n - lt * kSortTile;
    ti.len = (int32_t)(rem < kSortTile ? rem : kSortTile);
    ti.hbase = (int64_t)kRadix * t_start[cc];
    ti.ntiles = t_start[cc + 1] - t_start[cc];
    ti.lt = lt;
    return ti;
}

__global__ void __launch_bounds__(kSortThreads)
k_radix_hist(const uint32_t *__restrict__ keys, int32_t B, const int64_t *__restrict__ t_start,
             const int64_t *__restrict__ p_start, int shift, uint32_t *__restrict__ hist) {
    __shared__ uint32_t h[kRadix];
    for (int k = threadIdx.x; k < kRadix; k += kSortThreads) h[k] = 0;
    __syncthreads();
    const TileInfo ti = tile_info(blockIdx.x, B, t_start, p_start);
    for (int k = threadIdx.x; k < ti.len; k += kSortThreads) {
        const uint32_t d = (keys[ti.start + k] >> shift) & (kRadix - 1);
        atomicAdd(&h[d], 1u);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < kRadix; d += kSortThreads)
        hist[ti.hbase + (int64_t)d * ti.ntiles + ti.lt] = h[d];
}

// Ranks each key among the equal-digit keys before it in tile order and
// scatters (key, val) to its global destination.  Wave w owns the contiguous
// key range [w * 1024, (w + 1) * 1024) of the tile and walks it in 16 rounds
// of 64 consecutive keys; peers with the same digit inside a round are found
// with kRadixBits ballots.
__global__ void __launch_bounds__(kSortThreads)
k_radix_scatter(const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                uint32_t *__restrict__ kout, uint32_t *__restrict__ vout, int32_t B,
                const int64_t *__restrict__ t_start, const int64_t *__restrict__ p_start, int shift,
                const uint32_t *__restrict__ hist_ex) {
    constexpr int kWaves = kSortThreads / kWave;
    constexpr int kPerWave = kSortTile / kWaves;  // 1024
    __shared__ uint32_t wcnt[kWaves][kRadix];
    __shared__ uint32_t gbase[kRadix];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    for (int k = threadIdx.x; k < kWaves * kRadix; k += kSortThreads) (&wcnt[0][0])[k] = 0;
    const TileInfo ti = tile_info(blockIdx.x, B, t_start, p_start);
    for (int d = threadIdx.x; d < kRadix; d += kSortThreads)
        gbase[d] = hist_ex[ti.hbase + (int64_t)d * ti.ntiles + ti.lt];
    __syncthreads();

    uint32_t key[kSortItems], val[kSortItems], rank[kSortItems];
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int pos = wave * kPerWave + r * kWave + lane;
        const bool valid = pos < ti.len;
        key[r] = valid ? kin[ti.start + pos] : 0u;
        val[r] = valid ? vin[ti.start + pos] : 0u;
        const uint32_t d = (key[r] >> shift) & (kRadix - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kRadixBits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t before = wcnt[wave][d];
        const uint32_t in_round = (uint32_t)__popcll(peers & lt_mask);
        rank[r] = before + in_round;
        // the lowest peer publishes the new running count of this digit
        if (valid && (peers & lt_mask) == 0) wcnt[wave][d] = before + (uint32_t)__popcll(peers);
        // LDS operations of one wavefront execute in program order, so the
        // next round's reads see this write; the empty asm keeps the
        // compiler from moving them across each other.
        __asm__ volatile("" ::: "memory");
    }
    __syncthreads();
    // exclusive prefix over waves per digit
    for (int d = threadIdx.x; d < kRadix; d += kSortThreads) {
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t c = wcnt[w][d];
            wcnt[w][d] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int pos = wave * kPerWave + r * kWave + lane;
        if (pos < ti.len) {
            const uint32_t d = (key[r] >> shift) & (kRadix - 1);
            const uint32_t dst = gbase[d] + wcnt[wave][d] + rank[r];
            kout[dst] = key[r];
            vout[dst] = val[r];
        }
    }
}

int radix_sort_ref(const Batch &b, Work &w, int passes, hipStream_t s) {
    if (b.rt_total == 0) return TCR_OK;
    const int64_t nh = (int64_t)kRadix * b.rt_total;
    uint32_t *ki = w.keys0, *vi = w.vals0, *ko = w.keys1, *vo = w.vals1;
    for (int p = 0; p < passes; ++p) {
        const int shift = p * kRadixBits;
        hipLaunchKernelGGL(k_radix_hist, dim3((unsigned)b.rt_total), dim3(kSortThreads), 0, s, ki, b.B,
                           (const int64_t *)w.rt_start, (const int64_t *)w.r_start, shift, w.hist);
        int rc = exclusive_scan<uint32_t, uint32_t>(w.hist, w.hist, nh, false, w.scan_part, s);
        if (rc) return rc;
        hipLaunchKernelGGL(k_radix_scatter, dim3((unsigned)b.rt_total), dim3(kSortThreads), 0, s, ki, vi, ko,
                           vo, b.B, (const int64_t *)w.rt_start, (const int64_t *)w.r_start, shift,
                           (const uint32_t *)w.hist);
        TCR_LAUNCH_CHECK();
        uint32_t *t = ki; ki = ko; ko = t;
        t = vi; vi = vo; vo = t;
    }
    return TCR_OK;
}

}  // namespace tcr
