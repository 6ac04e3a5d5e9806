// Structure build on the device: hMax, hCell, wrapped cell keys, the sort
// into cell order, the cell table and the lin -> cell table -- the GPU form
// of neighborSearch_cpp's preamble (py/compactHashing/cppSearch.py:215-224)
// and buildCompactHashMap_compat (py/compactHashing/hashTable.py:79-154).
// The reference's bucket-ordered tables (hashTable [L,2], sortedCellTable)
// are built only when a structure is exported (build_bucket_tables).
#include <algorithm>

#include "tcr_internal.h"

namespace tcr {

constexpr int kBlock = 256;

__device__ __forceinline__ int cloud_of(int32_t B, int64_t g, const int64_t *__restrict__ start) {
    return find_segment(B, g, [&](int k) { return start[k]; });
}

__global__ void __launch_bounds__(kBlock) k_fill_u32(uint32_t *__restrict__ p, uint32_t v, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) p[i] = v;
}

int fill_u32(void *p, uint32_t v, size_t n, hipStream_t s) {
    if (n == 0) return TCR_OK;
    const int64_t blocks = std::min<int64_t>(2048, ((int64_t)n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(kBlock), 0, s, (uint32_t *)p, v, (int64_t)n);
    TCR_LAUNCH_CHECK();
    return TCR_OK;
}

// hCell = compute_h(min, max, hMax) (py/util.py:28-79) in fp32, numCells =
// ceil(extent / hCell).int32 (py/compactHashing/cellTable.py:29).
__device__ void grid_of(CloudDev &cd, uint32_t hmax_ord) {
    const int dim = cd.h.dim;
    const float hmax = hmax_ord ? ordered_to_float(hmax_ord) : -INFINITY;
    float ext[3] = {0.f, 0.f, 0.f}, nc[3] = {0.f, 0.f, 0.f}, h[3] = {0.f, 0.f, 0.f};
    bool over = false, stretch = false;
    for (int d = 0; d < dim; ++d) {
        ext[d] = cd.h.domain_max[d] - cd.h.domain_min[d];
        nc[d] = floorf(ext[d] / hmax);
        h[d] = ext[d] / nc[d];
    }
    for (int d = 0; d < dim; ++d)
        if (ext[d] / h[d] - nc[d] > 0.0f) over = true;
    if (over)
        for (int d = 0; d < dim; ++d) {
            nc[d] = nc[d] - 1.0f;
            h[d] = ext[d] / nc[d];
        }
    for (int d = 0; d < dim; ++d)
        if (ceilf(ext[d] / h[d]) > ext[d] / h[d]) stretch = true;
    if (stretch)
        for (int d = 0; d < dim; ++d) h[d] = h[d] * 1.0001f;
    float hc = h[0];
    for (int d = 1; d < dim; ++d) hc = (h[d] > hc || isnan(h[d])) ? h[d] : hc;
    cd.hcell = hc;
    {
        // the filter geometry: candidates of a cell's stencil lie within
        // 1.5 cells of its centre, queries of the cell within 0.5 cells
        float coord_max = 0.0f, ext_max = 0.0f;
        for (int d = 0; d < 3; ++d) cd.f_ext[d] = 0.0f;
        for (int d = 0; d < dim; ++d) {
            const float mn = cd.h.domain_min[d], mx = cd.h.domain_max[d];
            cd.f_ext[d] = mx - mn;
            coord_max = fmaxf(coord_max, fmaxf(fabsf(mn), fabsf(mx)));
            if ((cd.h.periodic_mask >> d) & 1u) ext_max = fmaxf(ext_max, cd.f_ext[d]);
        }
        cd.f_box = 1.51f * hc;
        cd.f_geo = 3.0f * (coord_max + 2.0f * cd.f_box) + ext_max + 3.0f * hc;
        const float cell2 = hc * hc;
        cd.f_c80 = 80.0f * cell2;
    }
    uint32_t err = 0;
    // wavefront-per-cell eligibility: stencil radius 1, a finite positive
    // cell size, and >= 5 cells in every periodic dimension (the cells are
    // the reference set's for a self-search, the query cells otherwise).
    // A pair within support (<= hCell) has its minimal image within 1.5 cells
    // of the query's cell centre, and nc = ceil(ext / hCell) >= 5 gives
    // ext > 4 cells: the image nearest the centre (|offset| <= ext / 2 > 2
    // cells) is that minimal image, with half a cell to spare for rounding,
    // and the 3 stencil cells of a dimension are distinct.
    uint32_t fast = (cd.h.search_range == 1 && hc > 0.0f && !isinf(hc)) ? 1u : 0u;
    for (int d = 0; d < 3; ++d) {
        if (d < dim) {
            const float v = ceilf(ext[d] / hc);
            cd.nc[d] = (v >= -2147483648.0f && v < 2147483648.0f) ? (int32_t)v : INT32_MIN;
            if (((cd.h.periodic_mask >> d) & 1u) && cd.nc[d] <= 0 && cd.h.nq > 0 && cd.h.nr > 0) err = 1;
            if (cd.nc[d] <= 0 || (((cd.h.periodic_mask >> d) & 1u) && cd.nc[d] < 5)) fast = 0;
        } else {
            cd.nc[d] = 1;
        }
    }
    // query cells decode their coordinates from the linear index: a grid of
    // more than 2^31 cells (whose index wraps) leaves them to the per-query path
    if (!cd.h.self_search && fast) {
        uint64_t cells = 1;
        for (int d = 0; d < dim; ++d) cells *= (uint64_t)cd.nc[d];
        if (cells >= (1ull << 31)) fast = 0;
    }
    cd.err = err;
    cd.fast = fast;
}

// hMax per cloud (computeGridSupport, py/compactHashing/cellTable.py:40-81):
// gather reads query supports, scatter reference supports, the symmetric
// modes both.  Blocks [0, nbq) reduce slices of the concatenated query
// supports, the rest slices of the reference supports; a block reduces per
// cloud in registers and LDS and issues one atomicMax per cloud it touched.
__global__ void __launch_bounds__(kBlock) k_prep(CloudDev *__restrict__ desc, int32_t B,
                                                 const int64_t *__restrict__ q_start, int64_t nq, int nbq,
                                                 const int64_t *__restrict__ r_start, int64_t nr, int nbr,
                                                 ClearSpans clear) {
    __shared__ uint32_t s_max[kBlock / kWave];
    __shared__ int s_cloud[kBlock / kWave];
    // zero the sort and cell-extraction counters of this search (read only
    // by kernels that run after this one)
    for (int k = 0; k < ClearSpans::kMax; ++k)
        for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < clear.n[k]; i += (int64_t)gridDim.x * kBlock)
            clear.p[k][i] = 0u;
    const bool qside = (int)blockIdx.x < nbq;
    const int bi = qside ? (int)blockIdx.x : (int)blockIdx.x - nbq;
    const int nb = qside ? nbq : nbr;
    const int64_t n = qside ? nq : nr;
    const int64_t *start = qside ? q_start : r_start;
    if (nb > 0 && n > 0) {
        const int64_t per = ((n + nb - 1) / nb + kBlock - 1) / kBlock * kBlock;
        const int64_t g0 = (int64_t)bi * per, g1 = min(n, g0 + per);
        int c = -1;
        uint32_t m = 0u;  // ordered encoding: 0 is below every float
        bool use = false;
        const float *sup = nullptr;
        int64_t c_lo = 0, c_hi = -1;
        // a slice inside one cloud (always, for one cloud) reads without a
        // per-element cloud check, so the loads can be issued back to back
        const int cs = g0 < g1 ? cloud_of(B, g0, start) : 0;
        const bool one_cloud = g0 < g1 && start[cs + 1] >= g1;
        if (one_cloud) {
            const CloudDev &cd = desc[cs];
            const int mode = cd.h.mode;
            use = qside ? (mode != TCR_MODE_SCATTER || cd.h.self_search)
                        : (mode != TCR_MODE_GATHER && !cd.h.self_search);
            sup = qside ? cd.qs : cd.rs;
            c = cs;
            if (use && sup) {
                const float *p = sup - start[cs];
#pragma unroll 4
                for (int64_t g = g0 + threadIdx.x; g < g1; g += kBlock) m = max(m, float_to_ordered(p[g]));
            }
        }
        for (int64_t g = g0 + threadIdx.x; !one_cloud && g < g1; g += kBlock) {
            if (g >= c_hi) {
                if (c >= 0 && m) atomicMax(&desc[c].hmax_ord, m);
                c = cloud_of(B, g, start);
                c_lo = start[c];
                c_hi = start[c + 1];
                m = 0u;
                const CloudDev &cd = desc[c];
                const int mode = cd.h.mode;
                // a self-search cloud's query set is its reference set
                use = qside ? (mode != TCR_MODE_SCATTER || cd.h.self_search)
                            : (mode != TCR_MODE_GATHER && !cd.h.self_search);
                sup = qside ? cd.qs : cd.rs;
                if (!sup) use = false;
            }
            if (use) m = max(m, float_to_ordered(sup[g - c_lo]));
        }
        // lanes ending on the same cloud as lane 0 merge in the wave, then
        // in the block; others flush alone
        const int c0 = __shfl(c, 0, kWave);
        if (c != c0) {
            if (c >= 0 && m) atomicMax(&desc[c].hmax_ord, m);
            m = 0u;
        }
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, kWave));
        if ((threadIdx.x & (kWave - 1)) == 0) {
            s_max[threadIdx.x / kWave] = m;
            s_cloud[threadIdx.x / kWave] = c0;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 0; w < kBlock / kWave; ++w) {
                uint32_t mm = s_max[w];
                const int cw = s_cloud[w];
                if (cw < 0 || !mm) continue;
                for (int w2 = w + 1; w2 < kBlock / kWave; ++w2)
                    if (s_cloud[w2] == cw) {
                        mm = max(mm, s_max[w2]);
                        s_max[w2] = 0;
                    }
                atomicMax(&desc[cw].hmax_ord, mm);
            }
        }
    }
}

// Grids of every cloud from the reduced hMax (a kernel of its own: the
// launch boundary orders it after every block's atomicMax, where a
// last-block ticket would need an agent-scope release fence per block --
// an L2 write-back on this GPU).
__global__ void __launch_bounds__(kBlock) k_grid(CloudDev *__restrict__ desc, int32_t B,
                                                 uint32_t *__restrict__ misc) {
    for (int c = blockIdx.x * kBlock + threadIdx.x; c < B; c += gridDim.x * kBlock) {
        grid_of(desc[c], desc[c].hmax_ord);
        if (desc[c].err) atomicOr(&misc[kMiscErr], 1u);
        if (desc[c].h.nq > 0 && !owned_by_cells(desc[c])) atomicOr(&misc[kMiscGeneric], 1u);
    }
}

// Reference set: second wrap (cppSearch.py:221) on top of radiusSearch's
// first (neighborhood.py:203), cell coordinates and int32 linear key
// (cellTable.py:30-31).  Stray bits for the wavefront-per-cell path: bit 0
// -- the second wrap moved the point (its query copy, wrapped once, differs
// from its sorted copy); bit 1 -- its cell lies outside [0, numCells).
template <int DIM>
__device__ __forceinline__ void ref_key_one(const CloudDev &cd, int64_t local, float4 &wp, int32_t &lin,
                                            uint8_t &flags) {
    const float *p = cd.rp + local * DIM;
    float x[3] = {0.f, 0.f, 0.f};
    int32_t c[3];
    uint8_t f = 0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        float v = p[d];
        if ((cd.h.periodic_mask >> d) & 1u) {
            const float lo = cd.h.domain_min[d], ext = cd.h.domain_max[d] - cd.h.domain_min[d];
            const float w1 = wrap_coord(v, lo, ext);
            v = wrap_coord(w1, lo, ext);
            if (__float_as_uint(w1) != __float_as_uint(v)) f |= 1;
        }
        x[d] = v;
        c[d] = cell_coord(v, cd.h.domain_min[d], cd.hcell);
        if (c[d] < 0 || c[d] >= cd.nc[d]) f |= 2;
    }
    lin = linear_index<DIM>(c, cd.nc);
    wp = make_float4(x[0], x[1], x[2], cd.rs ? cd.rs[local] : 0.0f);
    flags = f;
}

// One sort tile of reference points per workgroup: per point the second wrap,
// cell key and stray bits (ref_key_one), and the tile's digit counts of
// every radix pass, folded into the cloud's histograms -- the work of
// k_sort_hist without its launch and its second read of the keys.
// Histograms count raw key digits (a wide cloud's are mapped by
// hist_digit); a tile that sees a point outside the grid marks the cloud
// irregular and asks for every pass.
__global__ void __launch_bounds__(kSortThreads) k_keys(CloudDev *__restrict__ desc, int32_t B,
                                                       const int64_t *__restrict__ t_start,
                                                       const int64_t *__restrict__ p_start,
                                                       float4 *__restrict__ rwrap, uint32_t *__restrict__ keys,
                                                       uint32_t *__restrict__ vals, uint8_t *__restrict__ pflags,
                                                       uint32_t *__restrict__ hist, uint32_t *__restrict__ misc) {
    __shared__ uint32_t h[kMaxPasses][kRadix];
    __shared__ uint32_t s_out;
    for (int k = threadIdx.x; k < kMaxPasses * kRadix; k += kSortThreads) (&h[0][0])[k] = 0;
    if (threadIdx.x == 0) s_out = 0;
    __syncthreads();
    const TileInfo ti = tile_info(blockIdx.x, B, t_start, p_start);
    CloudDev &cd = desc[ti.c];
    const int64_t cstart = p_start[ti.c];
    const int dim = cd.h.dim;
    bool out = false;
    for (int base = 0; base < ti.len; base += kSortThreads) {  // uniform trip count: the runs span lanes
        const int k = base + threadIdx.x;
        const bool valid = k < ti.len;
        int32_t lin = 0;
        if (valid) {
            const int64_t g = ti.start + k;
            const int64_t local = g - cstart;
            float4 wp;
            uint8_t f;
            switch (dim) {
                case 1: ref_key_one<1>(cd, local, wp, lin, f); break;
                case 2: ref_key_one<2>(cd, local, wp, lin, f); break;
                default: ref_key_one<3>(cd, local, wp, lin, f); break;
            }
            rwrap[g] = wp;
            keys[g] = (uint32_t)lin;
            vals[g] = (uint32_t)local;
            pflags[g] = f;
            out = out || (f & 2) != 0;
        }
        hist_add_runs(h, (uint32_t)lin, valid);
    }
    if (out) s_out = 1u;
    __syncthreads();
    uint32_t *hc = hist + (int64_t)ti.c * kMaxPasses * kRadix;
    for (int k = threadIdx.x; k < kMaxPasses * kRadix; k += kSortThreads) {
        const uint32_t v = (&h[0][0])[k];
        if (v) atomicAdd(&hc[k], v);
    }
    if (threadIdx.x == 0) {
        if (s_out) atomicOr(&cd.irregular, 1u);
        atomicMax(&misc[kMiscPasses], s_out ? (uint32_t)kMaxPasses : cloud_passes(cd));
    }
}

// Cell extraction (unique_consecutive of hashTable.py:113 with the cell
// coordinates of its first point, :121-122, and their bucket, :136):
// one tile of sorted slots per block, cell numbers by decoupled look-back
// over all earlier tiles (cells are numbered across the whole batch), and
// every cell inserted into its cloud's lin -> cell table.
template <int DIM>
__device__ __forceinline__ uint32_t cell_bucket(const CloudDev &cd, float4 p) {
    const float x[3] = {p.x, p.y, p.z};
    int32_t c[3];
#pragma unroll
    for (int d = 0; d < DIM; ++d) c[d] = cell_coord(x[d], cd.h.domain_min[d], cd.hcell);
    return hash_index<DIM>(c, cd.h.hash_map_length);
}

struct CellsArgs {
    CloudDev *desc;
    int32_t B;
    int64_t ntiles, n;
    const int64_t *t_start, *p_start;
    const int32_t *skey;
    const float4 *sorted;
    uint32_t *cell_begin;
    int32_t *cell_lin;
    uint32_t *cell_hash;
    uint32_t *cell_cloud;
    uint64_t *lintab;
    uint64_t *lb;
    uint32_t *misc;
};

__global__ void __launch_bounds__(kSortThreads) k_cells(CellsArgs a) {
    constexpr int kItems = kSortTile / kSortThreads;  // 16 consecutive slots per thread
    __shared__ uint32_t s_wsum[kSortThreads / kWave];
    __shared__ int64_t s_tile;
    __shared__ uint64_t s_excl;
    if (threadIdx.x == 0) s_tile = atomicAdd(&a.misc[kMiscCellTicket], 1u);
    __syncthreads();
    const int64_t tile = s_tile;
    const int c = find_segment(a.B, tile, [&](int k) { return a.t_start[k]; });
    const int64_t first_tile = a.t_start[c];
    const int64_t cstart = a.p_start[c], cn = a.p_start[c + 1] - cstart;
    const int64_t t0 = cstart + (tile - first_tile) * kSortTile;
    const int64_t len = min((int64_t)kSortTile, cn - (tile - first_tile) * kSortTile);
    // flags of this thread's slots
    const int64_t s0 = t0 + (int64_t)threadIdx.x * kItems;
    uint32_t fl = 0, cnt = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int64_t s = s0 + k;
        if (s - t0 < len) {
            const bool f = (s == cstart) || (a.skey[s] != a.skey[s - 1]);
            fl |= (uint32_t)f << k;
            cnt += f;
        }
    }
    // block scan of the per-thread counts
    uint32_t incl = cnt;
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t v = __shfl_up(incl, o, kWave);
        if (lane >= o) incl += v;
    }
    if (lane == kWave - 1) s_wsum[threadIdx.x / kWave] = incl;
    __syncthreads();
    uint32_t wpre = 0, ttot = 0;
    for (int w = 0; w < kSortThreads / kWave; ++w) {
        if (w < (int)threadIdx.x / kWave) wpre += s_wsum[w];
        ttot += s_wsum[w];
    }
    // look-back over all earlier tiles of the batch (wave 0, 64 tiles a step)
    if (threadIdx.x < kWave) {
        const uint64_t excl = tile_look_back(a.lb, tile, ttot);
        if (threadIdx.x == 0) {
            s_excl = excl;
            if (tile == first_tile) a.desc[c].cell_off = (int64_t)excl;
            if (tile == a.ntiles - 1) {
                a.misc[kMiscCells] = (uint32_t)(excl + ttot);
                a.cell_begin[excl + ttot] = (uint32_t)a.n;
            }
        }
    }
    __syncthreads();
    uint32_t cid = (uint32_t)(s_excl + wpre + incl - cnt);
    const CloudDev &cd = a.desc[c];
    uint64_t *tab = a.lintab + cd.lt_off;
    const uint64_t cap = (uint64_t)cd.lt_cap;
    const bool dense = cell_index_dense(cd);
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        if (!((fl >> k) & 1u)) continue;
        const int64_t s = s0 + k;
        const int32_t lin = a.skey[s];
        a.cell_begin[cid] = (uint32_t)s;
        a.cell_lin[cid] = lin;
        a.cell_cloud[cid] = (uint32_t)c;
        const float4 p = a.sorted[s];
        a.cell_hash[cid] = cd.h.dim == 1 ? cell_bucket<1>(cd, p) : cd.h.dim == 2 ? cell_bucket<2>(cd, p)
                                                                                 : cell_bucket<3>(cd, p);
        if (dense) {
            ((int32_t *)tab)[lin] = (int32_t)cid;   // lin in [0, cells) for a regular cloud
        } else {
            const uint64_t e = ((uint64_t)cid << 32) | (uint32_t)lin;
            uint64_t h = lintab_hash(lin) & (cap - 1);
            while (atomicCAS((unsigned long long *)&tab[h], ~0ull, (unsigned long long)e) != ~0ull)
                h = (h + 1) & (cap - 1);
        }
        ++cid;
    }
}

// Per-cloud cell counts from the first-cell numbers.
__global__ void k_cloud_cells(CloudDev *__restrict__ desc, int32_t B, const uint32_t *__restrict__ misc) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int64_t next = misc[kMiscCells];
    for (int c = B - 1; c >= 0; --c) {
        if (desc[c].h.nr == 0 && !(desc[c].h.self_search && desc[c].h.nq > 0)) {
            desc[c].cell_off = next;
            desc[c].ncells = 0;
            continue;
        }
        desc[c].ncells = next - desc[c].cell_off;
        next = desc[c].cell_off;
    }
}

static inline unsigned blocks_for(int64_t n, int t = kBlock) { return (unsigned)((n + t - 1) / t); }


int build_prep(const Batch &b, Work &w, hipStream_t s) {
    // hMax + grid: one launch
    bool ref_side = false;
    for (const CloudDev &d : b.dev)
        if (!d.h.self_search && d.h.mode != TCR_MODE_GATHER && d.h.nr > 0) ref_side = true;
    const int nbq = (int)std::min<int64_t>(512, blocks_for(b.nq_total));
    const int nbr = ref_side ? (int)std::min<int64_t>(512, blocks_for(b.nr_total)) : 0;
    ClearSpans clear{};
    if (b.nr_total > 0 && b.rt_total > 0) {
        clear.p[0] = w.sort_hist;
        clear.n[0] = (int64_t)b.B * kMaxPasses * kRadix;
        clear.p[1] = (uint32_t *)w.sort_lb;
        clear.n[1] = 2 * (int64_t)kMaxPasses * b.rt_total * kRadix;
        clear.p[2] = (uint32_t *)w.cell_lb;
        clear.n[2] = 2 * b.rt_total;
    }
    if (b.qt_total > 0) {
        clear.p[3] = w.qsort_hist;
        clear.n[3] = (int64_t)b.B * kMaxPasses * kRadix;
        clear.p[4] = (uint32_t *)w.qsort_lb;
        clear.n[4] = 2 * (int64_t)kMaxPasses * b.qt_total * kRadix;
        clear.p[5] = (uint32_t *)w.qcell_lb;
        clear.n[5] = 2 * b.qt_total;
    }
    clear.p[6] = (uint32_t *)w.scan_part;  // offsets-scan look-back status
    clear.n[6] = 2 * offsets_scan_tiles(b.nq_total);
    hipLaunchKernelGGL(k_prep, dim3((unsigned)std::max(1, nbq + nbr)), dim3(kBlock), 0, s, w.desc, b.B,
                       (const int64_t *)w.q_start, b.nq_total, nbq, (const int64_t *)w.r_start, b.nr_total, nbr,
                       clear);
    hipLaunchKernelGGL(k_grid, dim3((unsigned)std::min<int64_t>(64, (b.B + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                       w.desc, b.B, w.misc);
    TCR_LAUNCH_CHECK();
    return TCR_OK;
}

int build_sorted_cells(const Batch &b, Work &w, hipStream_t s, bool hist_done) {
    const int64_t n = b.nr_total;
    if (n == 0) return TCR_OK;
    // the sort's histogram kernel (or its first pass, after k_keys) also
    // empties the cell index; build_prep zeroed the look-back state of the
    // sort and of k_cells
    int rc = radix_sort(b, w, s, hist_done);
    if (rc) return rc;
    CellsArgs ca;
    ca.desc = w.desc;
    ca.B = b.B;
    ca.ntiles = b.rt_total;
    ca.n = n;
    ca.t_start = w.rt_start;
    ca.p_start = w.r_start;
    ca.skey = w.skey;
    ca.sorted = w.sorted;
    ca.cell_begin = w.cell_begin;
    ca.cell_lin = w.cell_lin;
    ca.cell_hash = w.cell_hash;
    ca.cell_cloud = w.cell_cloud;
    ca.lintab = w.lintab;
    ca.lb = w.cell_lb;
    ca.misc = w.misc;
    hipLaunchKernelGGL(k_cells, dim3((unsigned)b.rt_total), dim3(kSortThreads), 0, s, ca);
    TCR_LAUNCH_CHECK();
    return TCR_OK;
}

// Per-cloud cell counts (desc[c].ncells) -- read only by the structure
// export, which launches this first.
int cloud_cells_launch(const Batch &b, Work &w, hipStream_t s) {
    hipLaunchKernelGGL(k_cloud_cells, dim3(1), dim3(64), 0, s, w.desc, b.B, (const uint32_t *)w.misc);
    TCR_LAUNCH_CHECK();
    return TCR_OK;
}

int build_structure(const Batch &b, Work &w, hipStream_t s) {
    int rc = build_prep(b, w, s);
    if (rc) return rc;
    const int64_t n = b.nr_total;
    if (n == 0) return TCR_OK;
    hipLaunchKernelGGL(k_keys, dim3((unsigned)b.rt_total), dim3(kSortThreads), 0, s, w.desc, b.B,
                       (const int64_t *)w.rt_start, (const int64_t *)w.r_start, w.rwrap, w.keys0, w.vals0,
                       w.pflags, w.sort_hist, w.misc);
    TCR_LAUNCH_CHECK();
    rc = build_sorted_cells(b, w, s, true);
    if (rc) return rc;
    return build_queries(b, w, s);
}

// ------------------------------------------------------------ query cells --

// A query of a query-cell cloud: its once-wrapped position (the query set is
// wrapped once, neighborhood.py:202), the linear index of its cell in the
// reference grid as key, and bit 1 when that cell lies outside the grid
// (key 0 then: such a query goes to the per-query path, whatever cell it
// sorts into).
template <int DIM>
__device__ __forceinline__ void query_key_one(const CloudDev &cd, int64_t local, float4 &wp, int32_t &lin,
                                              uint8_t &flags) {
    const float *p = cd.qp + local * DIM;
    float x[3] = {0.f, 0.f, 0.f};
    int32_t c[3] = {0, 0, 0};
    bool out = false;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        float v = p[d];
        if ((cd.h.periodic_mask >> d) & 1u)
            v = wrap_coord(v, cd.h.domain_min[d], cd.h.domain_max[d] - cd.h.domain_min[d]);
        x[d] = v;
        c[d] = cell_coord(v, cd.h.domain_min[d], cd.hcell);
        if (c[d] < 0 || c[d] >= cd.nc[d]) out = true;
    }
    lin = out ? 0 : linear_index<DIM>(c, cd.nc);
    wp = make_float4(x[0], x[1], x[2], cd.qs ? cd.qs[local] : 0.0f);
    flags = out ? 2 : 0;
}

// One sort tile of queries per workgroup: keys, wrapped positions, flags and
// the tile's digit counts of every radix pass (as k_keys for the reference
// points); the query sort runs the passes the grid's cell count needs.
__global__ void __launch_bounds__(kSortThreads) k_qkeys(CloudDev *__restrict__ desc, int32_t B,
                                                        const int64_t *__restrict__ t_start,
                                                        const int64_t *__restrict__ p_start,
                                                        float4 *__restrict__ qwrap, uint32_t *__restrict__ keys,
                                                        uint32_t *__restrict__ vals, uint8_t *__restrict__ pflags,
                                                        uint32_t *__restrict__ hist, uint32_t *__restrict__ misc) {
    __shared__ uint32_t h[kMaxPasses][kRadix];
    for (int k = threadIdx.x; k < kMaxPasses * kRadix; k += kSortThreads) (&h[0][0])[k] = 0;
    __syncthreads();
    const TileInfo ti = tile_info(blockIdx.x, B, t_start, p_start);
    const CloudDev &cd = desc[ti.c];
    const int64_t cstart = p_start[ti.c];
    const int dim = cd.h.dim;
    for (int base = 0; base < ti.len; base += kSortThreads) {  // uniform trip count: the runs span lanes
        const int k = base + threadIdx.x;
        const bool valid = k < ti.len;
        int32_t lin = 0;
        if (valid) {
            const int64_t g = ti.start + k;
            const int64_t local = g - cstart;
            float4 wp;
            uint8_t f;
            switch (dim) {
                case 1: query_key_one<1>(cd, local, wp, lin, f); break;
                case 2: query_key_one<2>(cd, local, wp, lin, f); break;
                default: query_key_one<3>(cd, local, wp, lin, f); break;
            }
            qwrap[g] = wp;
            keys[g] = (uint32_t)lin;
            vals[g] = (uint32_t)local;
            pflags[g] = f;
        }
        hist_add_runs(h, (uint32_t)lin, valid);
    }
    __syncthreads();
    uint32_t *hc = hist + (int64_t)ti.c * kMaxPasses * kRadix;
    for (int k = threadIdx.x; k < kMaxPasses * kRadix; k += kSortThreads) {
        const uint32_t v = (&h[0][0])[k];
        if (v) atomicAdd(&hc[k], v);
    }
    if (threadIdx.x == 0) atomicMax(&misc[kMiscQPasses], cloud_passes(cd));
}

// Query cells: runs of equal keys in the sorted queries, numbered across the
// batch by decoupled look-back over the query tiles (as k_cells numbers the
// reference cells; no cell index is built for them).
__global__ void __launch_bounds__(kSortThreads) k_qcells(const int64_t *__restrict__ t_start,
                                                         const int64_t *__restrict__ p_start, int32_t B,
                                                         int64_t ntiles, const int32_t *__restrict__ skey,
                                                         uint32_t *__restrict__ cell_begin,
                                                         int32_t *__restrict__ cell_lin,
                                                         uint32_t *__restrict__ cell_cloud, uint64_t *__restrict__ lb,
                                                         uint32_t *__restrict__ misc, int64_t nq) {
    constexpr int kItems = kSortTile / kSortThreads;
    __shared__ uint32_t s_wsum[kSortThreads / kWave];
    __shared__ int64_t s_tile;
    __shared__ uint64_t s_excl;
    if (threadIdx.x == 0) s_tile = atomicAdd(&misc[kMiscQCellTicket], 1u);
    __syncthreads();
    const int64_t tile = s_tile;
    const int c = find_segment(B, tile, [&](int k) { return t_start[k]; });
    const int64_t first_tile = t_start[c];
    const int64_t cstart = p_start[c], cn = p_start[c + 1] - cstart;
    const int64_t t0 = cstart + (tile - first_tile) * kSortTile;
    const int64_t len = min((int64_t)kSortTile, cn - (tile - first_tile) * kSortTile);
    const int64_t s0 = t0 + (int64_t)threadIdx.x * kItems;
    uint32_t fl = 0, cnt = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int64_t s = s0 + k;
        if (s - t0 < len) {
            const bool f = (s == cstart) || (skey[s] != skey[s - 1]);
            fl |= (uint32_t)f << k;
            cnt += f;
        }
    }
    uint32_t incl = cnt;
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t v = __shfl_up(incl, o, kWave);
        if (lane >= o) incl += v;
    }
    if (lane == kWave - 1) s_wsum[threadIdx.x / kWave] = incl;
    __syncthreads();
    uint32_t wpre = 0, ttot = 0;
    for (int w = 0; w < kSortThreads / kWave; ++w) {
        if (w < (int)threadIdx.x / kWave) wpre += s_wsum[w];
        ttot += s_wsum[w];
    }
    if (threadIdx.x < kWave) {  // look-back: wave 0, 64 tiles a step
        const uint64_t excl = tile_look_back(lb, tile, ttot);
        if (threadIdx.x == 0) {
            s_excl = excl;
            if (tile == ntiles - 1) {
                misc[kMiscQCells] = (uint32_t)(excl + ttot);
                cell_begin[excl + ttot] = (uint32_t)nq;
            }
        }
    }
    __syncthreads();
    uint32_t cid = (uint32_t)(s_excl + wpre + incl - cnt);
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        if (!((fl >> k) & 1u)) continue;
        const int64_t s = s0 + k;
        cell_begin[cid] = (uint32_t)s;
        cell_lin[cid] = skey[s];
        cell_cloud[cid] = (uint32_t)c;
        ++cid;
    }
}

int build_queries(const Batch &b, Work &w, hipStream_t s) {
    if (b.qt_total == 0) return TCR_OK;
    hipLaunchKernelGGL(k_qkeys, dim3((unsigned)b.qt_total), dim3(kSortThreads), 0, s, w.desc, b.B,
                       (const int64_t *)w.qt_start, (const int64_t *)w.q_start, w.qwrap, w.qkeys0, w.qvals0,
                       w.qpflags, w.qsort_hist, w.misc);
    TCR_LAUNCH_CHECK();
    int rc = radix_sort_queries(b, w, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_qcells, dim3((unsigned)b.qt_total), dim3(kSortThreads), 0, s, (const int64_t *)w.qt_start,
                       (const int64_t *)w.q_start, b.B, b.qt_total, (const int32_t *)w.qskey, w.qcell_begin,
                       w.qcell_lin, w.qcell_cloud, w.qcell_lb, w.misc, b.nq_total);
    TCR_LAUNCH_CHECK();
    return TCR_OK;
}

// ------------------------------------------ reference-layout bucket tables --

__global__ void __launch_bounds__(kBlock) k_bucket_count(const CloudDev *__restrict__ desc, int64_t ncap,
                                                         const uint32_t *__restrict__ misc,
                                                         const uint32_t *__restrict__ cell_hash,
                                                         const uint32_t *__restrict__ cell_cloud,
                                                         uint32_t *__restrict__ bucket_cnt) {
    const int64_t cid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (cid >= ncap || cid >= (int64_t)misc[kMiscCells]) return;
    const CloudDev &cd = desc[cell_cloud[cid]];
    atomicAdd(&bucket_cnt[cd.hb_off + cell_hash[cid]], 1u);
}

// Cells grouped by hash bucket (hashTable.py:138-152).  Inside a bucket the
// order is that of the atomics: lookups match the unique linear index, so it
// never changes a search result.
__global__ void __launch_bounds__(kBlock) k_bucket_scatter(const CloudDev *__restrict__ desc, int64_t ncap,
                                                           const uint32_t *__restrict__ misc,
                                                           const uint32_t *__restrict__ cell_hash,
                                                           const uint32_t *__restrict__ cell_cloud,
                                                           const uint32_t *__restrict__ cell_begin,
                                                           const int32_t *__restrict__ cell_lin,
                                                           const int64_t *__restrict__ r_start,
                                                           const uint32_t *__restrict__ bucket_begin,
                                                           uint32_t *__restrict__ bucket_fill,
                                                           int32_t *__restrict__ cell_table) {
    const int64_t cid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (cid >= ncap || cid >= (int64_t)misc[kMiscCells]) return;
    const uint32_t c = cell_cloud[cid];
    const CloudDev &cd = desc[c];
    const int64_t b = cd.hb_off + cell_hash[cid];
    const uint32_t pos = bucket_begin[b] + atomicAdd(&bucket_fill[b], 1u);
    cell_table[3 * (int64_t)pos + 0] = cell_lin[cid];
    cell_table[3 * (int64_t)pos + 1] = (int32_t)((int64_t)cell_begin[cid] - r_start[c]);
    cell_table[3 * (int64_t)pos + 2] = (int32_t)(cell_begin[cid + 1] - cell_begin[cid]);
}

__global__ void __launch_bounds__(kBlock) k_hash_table(const CloudDev *__restrict__ desc, int32_t B,
                                                       const int64_t *__restrict__ hb_start, int64_t nb,
                                                       const uint32_t *__restrict__ bucket_cnt,
                                                       const uint32_t *__restrict__ bucket_begin,
                                                       int32_t *__restrict__ hash_table) {
    const int64_t b = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (b >= nb) return;
    const int c = cloud_of(B, b, hb_start);
    const uint32_t cnt = bucket_cnt[b];
    hash_table[2 * b + 0] = cnt ? (int32_t)((int64_t)bucket_begin[b] - desc[c].cell_off) : -1;
    hash_table[2 * b + 1] = (int32_t)cnt;
}

int build_bucket_tables(const Batch &b, Work &w, hipStream_t s) {
    const int64_t n = b.nr_total;
    int rc = fill_u32(w.bucket_cnt, 0u, (size_t)b.hb_total, s);
    if (rc) return rc;
    rc = fill_u32(w.bucket_fill, 0u, (size_t)b.hb_total, s);
    if (rc) return rc;
    if (n > 0)
        hipLaunchKernelGGL(k_bucket_count, dim3(blocks_for(n)), dim3(kBlock), 0, s, (const CloudDev *)w.desc, n,
                           (const uint32_t *)w.misc, (const uint32_t *)w.cell_hash, (const uint32_t *)w.cell_cloud,
                           w.bucket_cnt);
    rc = exclusive_scan<uint32_t, uint32_t>(w.bucket_cnt, w.bucket_begin, b.hb_total, false, w.scan_part, s);
    if (rc) return rc;
    if (n > 0)
        hipLaunchKernelGGL(k_bucket_scatter, dim3(blocks_for(n)), dim3(kBlock), 0, s, (const CloudDev *)w.desc, n,
                           (const uint32_t *)w.misc, (const uint32_t *)w.cell_hash, (const uint32_t *)w.cell_cloud,
                           (const uint32_t *)w.cell_begin, (const int32_t *)w.cell_lin, (const int64_t *)w.r_start,
                           (const uint32_t *)w.bucket_begin, w.bucket_fill, w.cell_table);
    hipLaunchKernelGGL(k_hash_table, dim3(blocks_for(b.hb_total)), dim3(kBlock), 0, s, (const CloudDev *)w.desc,
                       b.B, (const int64_t *)w.hb_start, b.hb_total, (const uint32_t *)w.bucket_cnt,
                       (const uint32_t *)w.bucket_begin, w.hash_table);
    TCR_LAUNCH_CHECK();
    return TCR_OK;
}

}  // namespace tcr
