// Stable LSD radix sort of (cell key, point index) pairs into cell order, one
// segment per cloud -- the contract of torch.argsort on the int32 cell keys
// (py/compactHashing/cellTable.py:33) on the reference's GPU path: inside a
// cell, ascending original index.
//
// Onesweep layout: the digits of every pass are counted per cloud first (by
// the compact path's keys kernel, which writes the keys a tile at a time, or
// by k_sort_hist here for the multi-level path's keys), then one kernel per
// pass ranks a tile of 4096 keys in LDS
// (16 wavefronts, 9-bit digits, wave64 multi-split with 9 ballots per
// round), finds the tile's global
// offset per digit by decoupled look-back over the earlier tiles of its
// cloud, and scatters.  Tiles are numbered by an atomic ticket, so a tile
// only ever waits for tiles that are already running.  The number of passes
// is decided on the device from the key range (regular clouds sort the
// linear index in [0, prod numCells); a cloud with points outside the grid
// or more than 2^31 cells sorts the sign-flipped 32-bit key), and the last
// pass gathers the sorted
// positions, supports, indices and stray bits directly.
#include "tcr_internal.h"

namespace tcr {

__global__ void __launch_bounds__(kSortThreads)
k_sort_hist(const CloudDev *__restrict__ desc, int32_t B, const int64_t *__restrict__ t_start,
            const int64_t *__restrict__ p_start, const uint32_t *__restrict__ keys, uint32_t *__restrict__ hist,
            uint32_t *__restrict__ misc, const int64_t *__restrict__ lt_start, uint64_t *__restrict__ lintab) {
    __shared__ uint32_t h[kMaxPasses][kRadix];
    for (int k = threadIdx.x; k < kMaxPasses * kRadix; k += kSortThreads) (&h[0][0])[k] = 0;
    const TileInfo ti = tile_info(blockIdx.x, B, t_start, p_start);
    const CloudDev &cd = desc[ti.c];
    // the tiles of a cloud empty its cell index for k_cells (the irregular
    // flag that picks the dense form is final once k_keys has run): the
    // whole open-addressing table, or the first (grid cells + 1) / 2 words of
    // a dense one.  A cloud without reference points has no tiles; its
    // index is never read (CellIndex answers "no cell").
    {
        const int64_t need = cell_index_dense(cd) ? (int64_t)((grid_cells(cd) + 1) / 2) : cd.lt_cap;
        const int64_t ntiles = t_start[ti.c + 1] - ti.first;
        uint64_t *tab = lintab + lt_start[ti.c];
        for (int64_t g = (blockIdx.x - ti.first) * kSortThreads + threadIdx.x; g < need; g += ntiles * kSortThreads)
            tab[g] = ~0ull;
    }
    __syncthreads();
    for (int base = 0; base < ti.len; base += kSortThreads) {  // uniform trip count: the runs span lanes
        const int k = base + threadIdx.x;
        const bool valid = k < ti.len;
        const uint32_t rk = valid ? keys[ti.start + k] : 0u;  // raw key digits (hist_digit maps a wide cloud's)
        hist_add_runs(h, rk, valid);
    }
    __syncthreads();
    uint32_t *hc = hist + (int64_t)ti.c * kMaxPasses * kRadix;
    for (int k = threadIdx.x; k < kMaxPasses * kRadix; k += kSortThreads) {
        const uint32_t v = (&h[0][0])[k];
        if (v) atomicAdd(&hc[k], v);
    }
    if (threadIdx.x == 0) atomicMax(&misc[kMiscPasses], cloud_passes(cd));
}

struct SortPassArgs {
    const CloudDev *desc;
    int32_t B;
    int64_t ntiles;
    const int64_t *t_start, *p_start;
    const uint32_t *kin, *vin;
    uint32_t *kout, *vout;
    const uint32_t *hist;
    uint64_t *lb;            // [ntiles * kRadix] of this pass
    uint32_t *misc;
    int npass_slot;          // misc word holding this sort's pass count
    int ticket_slot;         // first of its per-pass tile tickets
    int pass;
    // pass 0 empties the cell index when no k_sort_hist ran (the keys
    // kernel built the histograms before the irregular flag was final)
    int clear_index;
    const int64_t *lt_start;
    uint64_t *lintab;
    // last pass: gather
    const float4 *rwrap;
    const uint8_t *pflags;
    float4 *sorted;
    int32_t *sort_index;
    int32_t *skey;
    uint8_t *sflags;
};

__global__ void __launch_bounds__(kSortThreads) k_sort_pass(SortPassArgs a) {
    constexpr int kWaves = kSortThreads / kWave;
    constexpr int kPerWave = kSortTile / kWaves;  // 256
    static_assert(kRadix <= kSortThreads, "one thread per digit");
    static_assert(2 * kSortTile <= kWaves * kRadix, "the tile's keys and values fit the wave counters' LDS");
    __shared__ uint32_t wcnt[kWaves][kRadix];
    __shared__ uint32_t dbase[kRadix];      // digit base of the cloud in this pass
    __shared__ uint32_t tile_excl[kRadix];  // keys of this digit in earlier tiles of the cloud
    __shared__ uint32_t s_wtot[kRadix / kWave];
    __shared__ uint32_t s_ttot[kRadix / kWave];
    __shared__ uint32_t tcount[kRadix];     // the tile's keys per digit, then where each digit starts in the tile
    __shared__ int64_t s_tile;
    const uint32_t npass = a.misc[a.npass_slot];
    if ((uint32_t)a.pass >= npass) return;
    const bool last = (uint32_t)a.pass == npass - 1;
    const int shift = a.pass * kRadixBits;
    const int wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);

    if (threadIdx.x == 0) s_tile = atomicAdd(&a.misc[a.ticket_slot + a.pass], 1u);
    for (int k = threadIdx.x; k < kWaves * kRadix; k += kSortThreads) (&wcnt[0][0])[k] = 0;
    __syncthreads();
    const int64_t tile = s_tile;
    const TileInfo ti = tile_info(tile, a.B, a.t_start, a.p_start);
    const bool irr = cloud_wide(a.desc[ti.c]);
    if (a.clear_index && a.pass == 0) {
        const CloudDev &cd = a.desc[ti.c];
        const int64_t need = cell_index_dense(cd) ? (int64_t)((grid_cells(cd) + 1) / 2) : cd.lt_cap;
        const int64_t ntiles = a.t_start[ti.c + 1] - ti.first;
        uint64_t *tab = a.lintab + a.lt_start[ti.c];
        for (int64_t g = (tile - ti.first) * kSortThreads + threadIdx.x; g < need; g += ntiles * kSortThreads)
            tab[g] = ~0ull;
    }

    uint32_t key[kSortItems], val[kSortItems], rank[kSortItems];
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int pos = wave * kPerWave + r * kWave + lane;
        const bool valid = pos < ti.len;
        key[r] = valid ? a.kin[ti.start + pos] : 0u;
        val[r] = valid ? a.vin[ti.start + pos] : 0u;
        const uint32_t d = (eff_key(key[r], irr) >> shift) & (kRadix - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kRadixBits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t before = wcnt[wave][d];
        rank[r] = before + (uint32_t)__popcll(peers & lt_mask);
        // the lowest peer publishes the new running count of this digit;
        // LDS operations of one wavefront execute in program order, and the
        // empty asm keeps the compiler from reordering them
        if (valid && (peers & lt_mask) == 0) wcnt[wave][d] = before + (uint32_t)__popcll(peers);
        __asm__ volatile("" ::: "memory");
    }
    __syncthreads();
    // thread d < kRadix: tile count of digit d, wave prefixes, look-back,
    // digit base
    if (threadIdx.x < kRadix) {
        const int d = threadIdx.x;
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t c = wcnt[w][d];
            wcnt[w][d] = run;
            run += c;
        }
        tcount[d] = run;
        uint64_t *lb = a.lb + tile * kRadix + d;
        if (tile == ti.first) {
            lb_store(lb, kLbInc | run);
            tile_excl[d] = 0;
        } else {
            lb_store(lb, kLbAgg | run);
            const uint64_t excl = look_back(a.lb + d, kRadix, tile - 1, ti.first);
            lb_store(lb, kLbInc | (excl + run));
            tile_excl[d] = (uint32_t)excl;
        }
        dbase[d] = a.hist[((int64_t)ti.c * kMaxPasses + a.pass) * kRadix + hist_digit(d, a.pass, irr)];
    }
    __syncthreads();
    // exclusive scans over the digits: of the cloud's digit counts (its
    // digit bases) and of this tile's (where a digit's keys start in the
    // tile once ordered by digit): wavefront scans of 64 digits, then the
    // wavefront totals
    {
        const int d = threadIdx.x;
        uint32_t incl = 0, c = 0, tin = 0, tc = 0;
        if (d < kRadix) {
            c = dbase[d];
            tc = tcount[d];
            incl = c;
            tin = tc;
#pragma unroll
            for (int o = 1; o < kWave; o <<= 1) {
                const uint32_t v = __shfl_up(incl, o, kWave);
                const uint32_t tv = __shfl_up(tin, o, kWave);
                if (lane >= o) {
                    incl += v;
                    tin += tv;
                }
            }
            if (lane == kWave - 1) {
                s_wtot[wave] = incl;
                s_ttot[wave] = tin;
            }
        }
        __syncthreads();
        if (d < kRadix) {
            uint32_t pre = 0, tpre = 0;
            for (int w = 0; w < wave; ++w) {
                pre += s_wtot[w];
                tpre += s_ttot[w];
            }
            dbase[d] = pre + incl - c;
            tcount[d] = tpre + tin - tc;  // now the tile's digit start
        }
    }
    __syncthreads();
    // the tile ordered by digit in LDS (over the wave counters, which are
    // read into registers first), then written out in order: consecutive
    // lanes store consecutive slots of one digit's run
    uint32_t lpos[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const uint32_t d = (eff_key(key[r], irr) >> shift) & (kRadix - 1);
        lpos[r] = tcount[d] + wcnt[wave][d] + rank[r];
    }
    __syncthreads();
    uint32_t *const s_key = &wcnt[0][0];
    uint32_t *const s_val = s_key + kSortTile;
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int pos = wave * kPerWave + r * kWave + lane;
        if (pos < ti.len) {
            s_key[lpos[r]] = key[r];
            s_val[lpos[r]] = val[r];
        }
    }
    __syncthreads();
    const int64_t cbase = a.p_start[ti.c];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int i = r * kSortThreads + threadIdx.x;
        if (i < ti.len) {
            const uint32_t k = s_key[i], v = s_val[i];
            const uint32_t d = (eff_key(k, irr) >> shift) & (kRadix - 1);
            const int64_t dst = cbase + dbase[d] + tile_excl[d] + (i - tcount[d]);
            if (!last) {
                a.kout[dst] = k;
                a.vout[dst] = v;
            } else {
                a.sorted[dst] = a.rwrap[cbase + v];
                a.sort_index[dst] = (int32_t)v;
                a.skey[dst] = (int32_t)k;
                a.sflags[dst] = a.pflags[cbase + v];
            }
        }
    }
}

int radix_sort(const Batch &b, Work &w, hipStream_t s, bool hist_done) {
    if (b.rt_total == 0) return TCR_OK;
    // sort_hist and sort_lb were zeroed by build_prep's k_prep
    if (!hist_done)
        hipLaunchKernelGGL(k_sort_hist, dim3((unsigned)b.rt_total), dim3(kSortThreads), 0, s,
                           (const CloudDev *)w.desc, b.B, (const int64_t *)w.rt_start, (const int64_t *)w.r_start,
                           (const uint32_t *)w.keys0, w.sort_hist, w.misc, (const int64_t *)w.lt_start, w.lintab);
    SortPassArgs a;
    a.clear_index = hist_done ? 1 : 0;
    a.lt_start = w.lt_start;
    a.lintab = w.lintab;
    a.desc = w.desc;
    a.B = b.B;
    a.ntiles = b.rt_total;
    a.t_start = w.rt_start;
    a.p_start = w.r_start;
    a.hist = w.sort_hist;
    a.misc = w.misc;
    a.npass_slot = kMiscPasses;
    a.ticket_slot = kMiscSortTicket;
    a.rwrap = w.rwrap;
    a.pflags = w.pflags;
    a.sorted = w.sorted;
    a.sort_index = w.sort_index;
    a.skey = w.skey;
    a.sflags = w.sflags;
    for (int p = 0; p < b.sort_passes; ++p) {
        a.pass = p;
        a.kin = (p & 1) ? w.keys1 : w.keys0;
        a.vin = (p & 1) ? w.vals1 : w.vals0;
        a.kout = (p & 1) ? w.keys0 : w.keys1;
        a.vout = (p & 1) ? w.vals0 : w.vals1;
        a.lb = w.sort_lb + (int64_t)p * b.rt_total * kRadix;
        hipLaunchKernelGGL(k_sort_pass, dim3((unsigned)b.rt_total), dim3(kSortThreads), 0, s, a);
    }
    TCR_LAUNCH_CHECK();
    return TCR_OK;
}

// The same passes over the query keys of the query-cell clouds (histograms
// counted by k_qkeys): sorted query positions, indices, keys and flags.
int radix_sort_queries(const Batch &b, Work &w, hipStream_t s) {
    if (b.qt_total == 0) return TCR_OK;
    SortPassArgs a;
    a.clear_index = 0;
    a.lt_start = w.lt_start;
    a.lintab = w.lintab;
    a.desc = w.desc;
    a.B = b.B;
    a.ntiles = b.qt_total;
    a.t_start = w.qt_start;
    a.p_start = w.q_start;
    a.hist = w.qsort_hist;
    a.misc = w.misc;
    a.npass_slot = kMiscQPasses;
    a.ticket_slot = kMiscQSortTicket;
    a.rwrap = w.qwrap;
    a.pflags = w.qpflags;
    a.sorted = w.qsorted;
    a.sort_index = w.qsort_index;
    a.skey = w.qskey;
    a.sflags = w.qsflags;
    for (int p = 0; p < b.qsort_passes; ++p) {
        a.pass = p;
        a.kin = (p & 1) ? w.qkeys1 : w.qkeys0;
        a.vin = (p & 1) ? w.qvals1 : w.qvals0;
        a.kout = (p & 1) ? w.qkeys0 : w.qkeys1;
        a.vout = (p & 1) ? w.qvals0 : w.qvals1;
        a.lb = w.qsort_lb + (int64_t)p * b.qt_total * kRadix;
        hipLaunchKernelGGL(k_sort_pass, dim3((unsigned)b.qt_total), dim3(kSortThreads), 0, s, a);
    }
    TCR_LAUNCH_CHECK();
    return TCR_OK;
}

}  // namespace tcr
