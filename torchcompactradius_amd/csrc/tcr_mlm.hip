// Multi-level search for variable supports: the same pairs, in the same
// order, as the single-level compact search, found without scanning the
// hMax-sized cells of dense regions.
//
// The reference's multi-level variant (multiLevelMemory/wrapper.py:145-308,
// cpp/multiLevelMemory/mlmUtil.h:234-282) bins points into levels of cell
// size hCell / 2^l by their support and searches every level; its output is
// defined to equal the compact path's (a pair set plus the compact row
// order).  Here:
//
// 1. Every reference point gets a level l = min(L-1, floor(log2(hMax/h_j)))
//    and a cell on the level-l grid of cell size hCell * 2^-l.  The level
//    grids nest: the level-l cell of x, shifted right by l, is its hCell
//    cell, because division by a power-of-two multiple of hCell is the exact
//    scaling of the fp32 quotient.  The key is
//    (l << 29 | hCell-cell linear index << d*l | index of the sub-cell inside
//    the hCell cell), so the points of one (level, hCell cell) group are
//    contiguous after the same stable radix sort and cell extraction as the
//    single-level search; one lin -> cell table holds the sub-cells and,
//    under keys with bit 31 set, the groups.
// 2. A query visits, for each level and each of the 3^d stencil hCell cells
//    of its own hCell cell (what the compact search visits), the group of
//    that level in that cell: the sub-cells that meet the box within
//    R * (1 + 2^-16) of the query -- R the largest support any pair with a
//    point of the level can have (by mode, from the level's largest support)
//    -- or all the group's points when that is fewer probes.  A candidate is
//    a pair if the reference predicate holds (cpp/common.h:106-129,
//    cpp/neighborhoodDynamic.h:63-70).  Every pair the compact search emits
//    is found: dist < support <= R bounds every coordinate gap, and the cell
//    index is monotone in the coordinate.
// 3. The fill writes (stencil offset << 32 | original index) per hit into the
//    row's segment; one wavefront per row sorts the segment, which is the
//    compact order (stencil offset, then original index inside a cell).
//
// Inputs this path does not take (reported so the caller runs the
// single-level search instead, which gives the same output): gather mode
// (the radius depends only on the query), non-finite or negative supports,
// points outside the hCell grid, fewer than 3 cells in a dimension (the
// compact stencil would then revisit or alias cells), level grids beyond
// 2^29 cells.
#include <algorithm>

#include "tcr_internal.h"

namespace tcr {

constexpr int kMlmBlock = 256;
constexpr int kMlmLevelShift = 29;
constexpr uint32_t kMlmGroupBit = 0x80000000u;
constexpr int kMlmSortMax = 2048;           // rows up to this length sort in LDS (16 KiB per wavefront)
constexpr int kMlmKey32Shift = 27;          // 32-bit slot keys: stencil offset above 27 index bits

enum : uint32_t { kMlmBadSupport = 1, kMlmOutOfGrid = 2, kMlmQueryOutOfGrid = 4, kMlmSlotOverflow = 8 };

__device__ __forceinline__ float level_cell(float hcell, int l) { return ldexpf(hcell, -l); }

// Linear index of an hCell cell (dim 0 fastest, as the single-level search).
template <int DIM>
__device__ __forceinline__ uint32_t coarse_lin(const int32_t *c, const int32_t *nc) {
    uint32_t lin = 0, stride = 1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        lin += (uint32_t)c[d] * stride;
        stride *= (uint32_t)nc[d];
    }
    return lin;
}

// Sub-cell index inside an hCell cell at level l (l bits per dimension).
template <int DIM>
__device__ __forceinline__ uint32_t sub_index(const int32_t *sub, int l) {
    uint32_t s = 0;
#pragma unroll
    for (int d = DIM - 1; d >= 0; --d) s = (s << l) | (uint32_t)sub[d];
    return s;
}

template <int DIM>
__device__ __forceinline__ uint32_t fine_key(int l, uint32_t clin, uint32_t sub) {
    return ((uint32_t)l << kMlmLevelShift) | (clin << (DIM * l)) | sub;
}
template <int DIM>
__device__ __forceinline__ uint32_t group_key(int l, uint32_t clin) {
    return kMlmGroupBit | ((uint32_t)l << kMlmLevelShift) | clin;
}

template <int DIM>
__device__ __forceinline__ void mlm_key_one(const CloudDev &cd, int64_t g, int L, float hmax, uint32_t &key,
                                            float4 &wp, int &level, uint32_t &flags) {
    const float *p = cd.rp + g * DIM;
    float x[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        float v = p[d];
        if ((cd.h.periodic_mask >> d) & 1u) {
            const float lo = cd.h.domain_min[d], ext = cd.h.domain_max[d] - cd.h.domain_min[d];
            v = wrap_coord(wrap_coord(v, lo, ext), lo, ext);
        }
        x[d] = v;
    }
    const float h = cd.rs ? cd.rs[g] : 0.0f;
    flags = 0;
    if (!(h >= 0.0f) || isinf(h)) flags |= kMlmBadSupport;
    int l = 0;
    if (h > 0.0f) {
        float r = hmax / h;
        while (l < L - 1 && r >= 2.0f) {
            r *= 0.5f;
            ++l;
        }
    } else {
        l = L - 1;
    }
    const float c = level_cell(cd.hcell, l);
    int32_t cc[3] = {0, 0, 0}, sub[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const int32_t cf = cell_coord(x[d], cd.h.domain_min[d], c);
        if (cf < 0 || cf >= (cd.nc[d] << l)) {
            flags |= kMlmOutOfGrid;
            cc[d] = 0;
            sub[d] = 0;
        } else {
            cc[d] = cf >> l;
            sub[d] = cf & ((1 << l) - 1);
        }
    }
    key = fine_key<DIM>(l, coarse_lin<DIM>(cc, cd.nc), sub_index<DIM>(sub, l));
    wp = make_float4(x[0], x[1], x[2], h);
    level = l;
}

// Grid-stride over the points; each thread keeps its running maximum
// support per level and its flags, merged per block before one atomic per
// level (a per-wavefront atomic on the same four words serialised).
__global__ void __launch_bounds__(kMlmBlock) k_mlm_keys(CloudDev *__restrict__ desc, int64_t n, int L,
                                                        float4 *__restrict__ rwrap, uint32_t *__restrict__ keys,
                                                        uint32_t *__restrict__ vals, uint8_t *__restrict__ pflags,
                                                        uint32_t *__restrict__ misc) {
    constexpr int kMaxLevels = 4;
    __shared__ uint32_t s_max[kMlmBlock / kWave][kMaxLevels + 1];
    const CloudDev &cd = desc[0];
    const float hmax = ordered_to_float(cd.hmax_ord);
    uint32_t lmax[kMaxLevels] = {0u, 0u, 0u, 0u};
    uint32_t flags = 0;
    for (int64_t g = (int64_t)blockIdx.x * kMlmBlock + threadIdx.x; g < n; g += (int64_t)gridDim.x * kMlmBlock) {
        uint32_t key, f = 0;
        float4 wp;
        int level = -1;
        switch (cd.h.dim) {
            case 1: mlm_key_one<1>(cd, g, L, hmax, key, wp, level, f); break;
            case 2: mlm_key_one<2>(cd, g, L, hmax, key, wp, level, f); break;
            default: mlm_key_one<3>(cd, g, L, hmax, key, wp, level, f); break;
        }
        rwrap[g] = wp;
        keys[g] = key;
        vals[g] = (uint32_t)g;
        pflags[g] = 0;
        flags |= f;
        if (wp.w >= 0.0f) {
            const uint32_t o = float_to_ordered(wp.w);
#pragma unroll
            for (int l = 0; l < kMaxLevels; ++l)
                if (level == l) lmax[l] = max(lmax[l], o);
        }
    }
    // per-level maximum support and flags: wavefront, block, then one atomic each
    const int wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int l = 0; l < kMaxLevels; ++l) {
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) lmax[l] = max(lmax[l], __shfl_xor(lmax[l], o, kWave));
        if (lane == 0) s_max[wave][l] = lmax[l];
    }
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) flags |= __shfl_xor(flags, o, kWave);
    if (lane == 0) s_max[wave][kMaxLevels] = flags;
    __syncthreads();
    if (threadIdx.x <= kMaxLevels) {
        uint32_t m = 0;
        for (int w = 0; w < kMlmBlock / kWave; ++w) m = threadIdx.x < kMaxLevels ? max(m, s_max[w][threadIdx.x])
                                                                                 : (m | s_max[w][kMaxLevels]);
        if (threadIdx.x < kMaxLevels) {
            if (m && (int)threadIdx.x < L) atomicMax(&misc[kMiscLevelMax + threadIdx.x], m);
        } else if (m) {
            atomicOr(&misc[kMiscMlmFlags], m);
        }
    }
    // the level keys use 31 bits: every radix pass runs
    if (blockIdx.x == 0 && threadIdx.x == 0) desc[0].irregular = 1;
}

int mlm_keys_launch(const Batch &b, Work &w, int levels, hipStream_t s) {
    const int64_t n = b.nr_total;
    if (n == 0) return TCR_OK;
    const int64_t blocks = std::min<int64_t>(1024, (n + kMlmBlock - 1) / kMlmBlock);
    hipLaunchKernelGGL(k_mlm_keys, dim3((unsigned)blocks), dim3(kMlmBlock), 0, s, w.desc, n, levels, w.rwrap, w.keys0,
                       w.vals0, w.pflags, w.misc);
    TCR_LAUNCH_CHECK();
    return TCR_OK;
}

// Groups (level, hCell cell): the first sub-cell of each group enters the
// lin table under its group key; group_end[first] = one past its last cell.
__device__ __forceinline__ uint32_t group_of(uint32_t key, int dim) {
    const int l = (int)((key >> kMlmLevelShift) & 3u);
    return kMlmGroupBit | (key & (3u << kMlmLevelShift)) | ((key & ((1u << kMlmLevelShift) - 1)) >> (dim * l));
}

__global__ void __launch_bounds__(kMlmBlock) k_mlm_groups(const CloudDev *__restrict__ desc, int64_t ncap,
                                                          const uint32_t *__restrict__ misc,
                                                          const int32_t *__restrict__ cell_lin,
                                                          uint64_t *__restrict__ lintab,
                                                          uint32_t *__restrict__ group_end) {
    const int64_t cid = (int64_t)blockIdx.x * kMlmBlock + threadIdx.x;
    const int64_t ncells = misc[kMiscCells];
    if (cid >= ncap || cid >= ncells) return;
    const CloudDev &cd = desc[0];
    const int dim = cd.h.dim;
    const uint32_t gk = group_of((uint32_t)cell_lin[cid], dim);
    if (cid > 0 && group_of((uint32_t)cell_lin[cid - 1], dim) == gk) return;
    int64_t e = cid + 1;
    while (e < ncells && group_of((uint32_t)cell_lin[e], dim) == gk) ++e;
    group_end[cid] = (uint32_t)e;
    uint64_t *tab = lintab + cd.lt_off;
    const uint64_t cap = (uint64_t)cd.lt_cap;
    const uint64_t ent = ((uint64_t)cid << 32) | gk;
    uint64_t h = lintab_hash((int32_t)gk) & (cap - 1);
    while (atomicCAS((unsigned long long *)&tab[h], ~0ull, (unsigned long long)ent) != ~0ull) h = (h + 1) & (cap - 1);
}

int mlm_groups_launch(const Batch &b, Work &w, hipStream_t s) {
    const int64_t n = b.nr_total;
    if (n == 0) return TCR_OK;
    hipLaunchKernelGGL(k_mlm_groups, dim3((unsigned)((n + kMlmBlock - 1) / kMlmBlock)), dim3(kMlmBlock), 0, s,
                       (const CloudDev *)w.desc, n, (const uint32_t *)w.misc, (const int32_t *)w.cell_lin, w.lintab,
                       w.cell_hash);
    TCR_LAUNCH_CHECK();
    return TCR_OK;
}

struct MlmArgs {
    const CloudDev *desc;
    int64_t nq;
    int L;
    const uint32_t *misc;
    uint32_t *flags;
    const uint64_t *lintab;
    const uint32_t *cell_begin;
    const uint32_t *group_end;
    const float4 *sorted;
    const int32_t *sort_index;
    int32_t *counts;
    const int64_t *offsets;
    uint64_t *scratch;   // fill: (stencil offset << 32 | original index) per hit
    uint32_t *slots;     // emit: (stencil offset << 27 | original index), query g's at
    int64_t slot_cap;    // [g * slot_cap, (g + 1) * slot_cap); references < 2^27
};

// Pair support bound for a query of support hi against any point of a
// level whose largest support is hl.
__device__ __forceinline__ float level_radius(int mode, float hi, float hl) {
    switch (mode) {
        case TCR_MODE_SCATTER: return hl;
        case TCR_MODE_SYMMETRIC: return (hi + hl) / 2.f;
        default: return max_ref(hi, hl);
    }
}

// Passes of the traversal: count the hits; fill their keys at the row's
// offset; or emit -- count them and write their keys into the query's slot
// (rows longer than a slot raise kMlmSlotOverflow).
enum : int { kMlmCount = 0, kMlmFill = 1, kMlmEmit = 2 };
constexpr int kMlmBatch = 8;  // candidates loaded together

template <int DIM, int PASS>
__device__ void mlm_query(const MlmArgs &a, int64_t g) {
    constexpr bool FILL = PASS == kMlmFill;
    constexpr int kOff = DIM == 1 ? 3 : DIM == 2 ? 9 : 27;
    const CloudDev &cd = a.desc[0];
    float xi[3] = {0.f, 0.f, 0.f}, mn[3] = {0.f, 0.f, 0.f}, mx[3] = {0.f, 0.f, 0.f};
    int32_t ci[3] = {0, 0, 0}, nc[3] = {1, 1, 1};
    const uint32_t pmask = cd.h.periodic_mask;
    const float hcell = cd.hcell;
    bool in = true;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        mn[d] = cd.h.domain_min[d];
        mx[d] = cd.h.domain_max[d];
        nc[d] = cd.nc[d];
        float v = cd.qp[g * DIM + d];
        if ((pmask >> d) & 1u) v = wrap_coord(v, mn[d], mx[d] - mn[d]);
        xi[d] = v;
        ci[d] = cell_coord(v, mn[d], hcell);
        in = in && ci[d] >= 0 && ci[d] < nc[d];
    }
    const float hi = cd.qs ? cd.qs[g] : 0.0f;
    if (!in) {
        atomicOr(a.flags, kMlmQueryOutOfGrid);
        if (!FILL) a.counts[g] = 0;
        return;
    }
    const int mode = cd.h.mode;
    const uint64_t *tab = a.lintab + cd.lt_off;
    const uint64_t cap = (uint64_t)cd.lt_cap;
    int64_t k = 0;
    const int64_t base = FILL ? a.offsets[g] : 0;

    auto test_v = [&](const float4 &v, int32_t sidx, int o) {
        const float xj[3] = {v.x, v.y, v.z};
        const float dist = exact_distance<DIM>(xi, xj, mn, mx, pmask);
        if (!support_pred(dist, hi, v.w, mode)) return;
        const uint64_t key = ((uint64_t)(uint32_t)o << 32) | (uint32_t)sidx;
        if (FILL) a.scratch[base + k] = key;
        if (PASS == kMlmEmit && k < a.slot_cap)
            a.slots[g * a.slot_cap + k] = ((uint32_t)o << kMlmKey32Shift) | (uint32_t)sidx;
        ++k;
    };
    // candidates [jb, je) in batches whose loads are all in flight before
    // the first is tested (a hit's store would otherwise make every next
    // load wait); the tail re-reads the last candidate and skips it
    auto test_run = [&](uint32_t jb, uint32_t je, int o) {
        for (uint32_t j = jb; j < je; j += kMlmBatch) {
            float4 v[kMlmBatch];
            int32_t si[kMlmBatch];
#pragma unroll
            for (int u = 0; u < kMlmBatch; ++u) {
                const uint32_t jj = min(j + (uint32_t)u, je - 1);
                v[u] = a.sorted[jj];
                si[u] = a.sort_index[jj];
            }
#pragma unroll
            for (int u = 0; u < kMlmBatch; ++u)
                if (j + (uint32_t)u < je) test_v(v[u], si[u], o);
        }
    };

    for (int l = 0; l < a.L; ++l) {
        const uint32_t hl_ord = a.misc[kMiscLevelMax + l];
        if (!hl_ord) continue;
        const float R = level_radius(mode, hi, ordered_to_float(hl_ord));
        if (!(R > 0.0f)) continue;
        const float c = level_cell(hcell, l);
        const int32_t span = 1 << l;
        for (int o = 0; o < kOff; ++o) {
            int32_t cc[3] = {0, 0, 0}, slo[3] = {0, 0, 0}, shi[3] = {0, 0, 0};
            bool skip = false;
            int rem = o;
#pragma unroll
            for (int d = DIM - 1; d >= 0; --d) {
                const int32_t off = (rem % 3) - 1;
                rem /= 3;
                const int32_t u = ci[d] + off;   // unwrapped hCell coordinate
                float xq = xi[d];
                if ((pmask >> d) & 1u) {
                    cc[d] = pymod(u, nc[d]);
                    // the query's image next to the wrapped cell
                    const float ext = mx[d] - mn[d];
                    if (cc[d] > u) xq = xi[d] + ext;
                    else if (cc[d] < u) xq = xi[d] - ext;
                } else {
                    if (u < 0 || u >= nc[d]) skip = true;
                    cc[d] = u;
                }
                // box within R of the query image, with room for the rounding
                // of the image shift and of the cell quotients
                const float Rp = R * (1.0f + 0x1p-16f) +
                                 0x1p-20f * (fabsf(xq) + fabsf(mn[d]) + fabsf(mx[d]) + c);
                const float blo = floorf((xq - Rp - mn[d]) / c), bhi = floorf((xq + Rp - mn[d]) / c);
                // sub-cells of this hCell cell inside the box
                const float f0 = (float)cc[d] * (float)span;
                const float lo = fmaxf(blo - f0, 0.0f), hi2 = fminf(bhi - f0, (float)(span - 1));
                if (!(lo <= hi2)) skip = true;
                slo[d] = (int32_t)lo;
                shi[d] = (int32_t)hi2;
            }
            if (skip) continue;
            const uint32_t clin = coarse_lin<DIM>(cc, nc);
            const int32_t gid = lintab_find(tab, cap, (int32_t)group_key<DIM>(l, clin));
            if (gid < 0) continue;
            const uint32_t cb = a.cell_begin[gid], ce = a.cell_begin[a.group_end[gid]];
            int64_t nsub = 1;
#pragma unroll
            for (int d = 0; d < DIM; ++d) nsub *= (int64_t)(shi[d] - slo[d] + 1);
            if ((int64_t)(ce - cb) <= 2 * nsub) {
                test_run(cb, ce, o);
            } else {
                for (int32_t s2 = slo[DIM > 2 ? 2 : 0]; s2 <= (DIM > 2 ? shi[2] : slo[0]); ++s2)
                    for (int32_t s1 = slo[DIM > 1 ? 1 : 0]; s1 <= (DIM > 1 ? shi[1] : slo[0]); ++s1)
                        for (int32_t s0 = slo[0]; s0 <= shi[0]; ++s0) {
                            const int32_t sub[3] = {s0, s1, s2};
                            const int32_t cid = lintab_find(tab, cap, (int32_t)fine_key<DIM>(l, clin, sub_index<DIM>(sub, l)));
                            if (cid < 0) continue;
                            const uint32_t jb = a.cell_begin[cid], je = a.cell_begin[cid + 1];
                            if (jb < je) test_run(jb, je, o);
                        }
            }
        }
    }
    if (!FILL) a.counts[g] = (int32_t)k;
    if (PASS == kMlmEmit && k > a.slot_cap) atomicOr(a.flags, kMlmSlotOverflow);
}

// One thread per query.  A self-search takes its queries in sorted order
// (thread t: the point at sorted slot t), so that the lanes of a wavefront
// hold neighbouring points of one level and walk the same cells.
template <int PASS>
__global__ void __launch_bounds__(kMlmBlock) k_mlm_query(MlmArgs a) {
    const int64_t t = (int64_t)blockIdx.x * kMlmBlock + threadIdx.x;
    if (t >= a.nq) return;
    const int64_t g = a.desc[0].h.self_search ? (int64_t)a.sort_index[t] : t;
    switch (a.desc[0].h.dim) {
        case 1: mlm_query<1, PASS>(a, g); break;
        case 2: mlm_query<2, PASS>(a, g); break;
        default: mlm_query<3, PASS>(a, g); break;
    }
}

int mlm_query_launch(const Batch &b, Work &w, int levels, bool fill, const int64_t *offsets, uint64_t *scratch,
                     hipStream_t s, uint32_t *slots, int64_t slot_cap) {
    if (b.nq_total == 0) return TCR_OK;
    MlmArgs a;
    a.desc = w.desc;
    a.nq = b.nq_total;
    a.L = levels;
    a.misc = w.misc;
    a.flags = w.misc + kMiscMlmFlags;
    a.lintab = w.lintab;
    a.cell_begin = w.cell_begin;
    a.group_end = w.cell_hash;
    a.sorted = w.sorted;
    a.sort_index = w.sort_index;
    a.counts = w.counts;
    a.offsets = offsets;
    a.scratch = scratch;
    a.slots = slots;
    a.slot_cap = slot_cap;
    const dim3 grid((unsigned)((b.nq_total + kMlmBlock - 1) / kMlmBlock)), block(kMlmBlock);
    if (fill) hipLaunchKernelGGL(k_mlm_query<kMlmFill>, grid, block, 0, s, a);
    else if (slots) hipLaunchKernelGGL(k_mlm_query<kMlmEmit>, grid, block, 0, s, a);
    else hipLaunchKernelGGL(k_mlm_query<kMlmCount>, grid, block, 0, s, a);
    TCR_LAUNCH_CHECK();
    return TCR_OK;
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One wavefront per row: sort the row's keys and write the original
// indices.  Keys are (offset << 32 | index) (K = uint64_t: the second
// traversal's keys at the row's own offset) or (offset << 27 | index)
// (K = uint32_t: the emit pass's slots, query g's at g * src_cap; rows at
// most src_cap <= kMlmSortMax long); both orders are (stencil offset,
// original index).  Two launches: SHORT takes the rows of at most 64 keys and
// sorts them in registers (one key per lane, bitonic network over shuffles;
// no LDS, so full occupancy); the other takes the longer rows, up to
// kMlmSortMax in LDS (bitonic network over the next power of two, padded
// with all-ones keys), and has lane 0 sort still longer rows in place
// (insertion sort; 64-bit keys only).  The 64-bit `scratch` may alias the
// int64 column output: every entry is read before it is overwritten.
template <typename K>
__device__ __forceinline__ int32_t key_index(K key) {
    return sizeof(K) == 8 ? (int32_t)(uint32_t)key : (int32_t)((uint32_t)key & ((1u << kMlmKey32Shift) - 1u));
}
// v of lane (lane ^ j), j a power of two <= 32 known once the sorting
// network is unrolled: DPP quad permutes for 1 and 2 (no LDS), the swizzle's
// bit-mask mode for 4, 8 and 16 (no address register), a permute for 32.
__device__ __forceinline__ uint32_t xor_lane32(uint32_t v, int j) {
    switch (j) {
        case 1: return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
        case 2: return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
        case 4: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, (4 << 10) | 0x1F);
        case 8: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, (8 << 10) | 0x1F);
        case 16: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, (16 << 10) | 0x1F);
        default: return (uint32_t)__shfl_xor((int)v, j, kWave);
    }
}

template <typename K>
__device__ __forceinline__ K shfl_xor_key(K key, int j) {
    if (sizeof(K) == 4) return (K)xor_lane32((uint32_t)key, j);
    const uint32_t olo = xor_lane32((uint32_t)key, j);
    const uint32_t ohi = xor_lane32((uint32_t)((uint64_t)key >> 32), j);
    return (K)(((uint64_t)ohi << 32) | olo);
}

template <typename K>
__device__ __forceinline__ void sort_row_short(int64_t g, int64_t b, int64_t src, int64_t m, const K *keys,
                                               const PairOut &out) {
    const int lane = __lane_id();
    K key = lane < m ? keys[src + lane] : (K)~(K)0;
    // a network over the next power of two >= m: lanes beyond it never pair
    // with lanes below it
#pragma unroll
    for (int kk = 2; kk <= kWave; kk <<= 1) {
        if ((kk >> 1) >= m) break;  // uniform: the first kk/2 lanes hold the whole row
#pragma unroll
        for (int j = kk >> 1; j > 0; j >>= 1) {
            const K other = shfl_xor_key(key, j);
            // pair (i, i ^ j): the lower lane keeps the minimum in an
            // ascending block (lane & kk == 0), the maximum otherwise
            const bool keep_min = ((lane & j) == 0) == ((lane & kk) == 0);
            key = keep_min ? (other < key ? other : key) : (other > key ? other : key);
        }
    }
    if (lane < m) out.put(b + lane, g, key_index(key));
}

template <typename K>
__global__ void __launch_bounds__(kMlmBlock) k_mlm_sort_short(int64_t nq, const int64_t *__restrict__ offsets,
                                                              const K *keys, int64_t src_cap, PairOut out) {
    const int64_t wid = ((int64_t)blockIdx.x * kMlmBlock + threadIdx.x) / kWave;
    const int64_t nw = ((int64_t)gridDim.x * kMlmBlock) / kWave;
    for (int64_t g = wid; g < nq; g += nw) {
        const int64_t b = offsets[g], m = offsets[g + 1] - b;
        if (m > 0 && m <= kWave) sort_row_short<K>(g, b, src_cap > 0 ? g * src_cap : b, m, keys, out);
    }
}

template <typename K>
__global__ void __launch_bounds__(kMlmBlock) k_mlm_sort_rows(int64_t nq, const int64_t *__restrict__ offsets,
                                                             K *keys, int64_t src_cap, PairOut out) {
    __shared__ K s_keys[kMlmBlock / kWave][kMlmSortMax];
    K *sk = s_keys[threadIdx.x / kWave];
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wid = ((int64_t)blockIdx.x * kMlmBlock + threadIdx.x) / kWave;
    const int64_t nw = ((int64_t)gridDim.x * kMlmBlock) / kWave;
    for (int64_t g = wid; g < nq; g += nw) {
        const int64_t b = offsets[g], m = offsets[g + 1] - b;
        if (m <= kWave) continue;  // k_mlm_sort_short
        const int64_t src = src_cap > 0 ? g * src_cap : b;
        if (m <= kMlmSortMax) {
            int n2 = 1;
            while (n2 < m) n2 <<= 1;
            for (int t = lane; t < n2; t += kWave) sk[t] = t < m ? keys[src + t] : (K)~(K)0;
            wave_lds_sync();
            for (int kk = 2; kk <= n2; kk <<= 1)
                for (int j = kk >> 1; j > 0; j >>= 1) {
                    for (int t = lane; t < n2 / 2; t += kWave) {
                        const int i = ((t / j) * 2 * j) + (t % j), p = i + j;
                        const K x = sk[i], y = sk[p];
                        const bool asc = (i & kk) == 0;
                        if ((x > y) == asc) {
                            sk[i] = y;
                            sk[p] = x;
                        }
                    }
                    wave_lds_sync();
                }
            for (int t = lane; t < m; t += kWave) out.put(b + t, g, key_index(sk[t]));
            wave_lds_sync();
        } else if (sizeof(K) == 8 && lane == 0) {
            for (int64_t i = 1; i < m; ++i) {
                const K x = keys[b + i];
                int64_t j = i - 1;
                while (j >= 0 && keys[b + j] > x) {
                    keys[b + j + 1] = keys[b + j];
                    --j;
                }
                keys[b + j + 1] = x;
            }
            for (int64_t t = 0; t < m; ++t) out.put(b + t, g, key_index(keys[b + t]));
        }
    }
}

template <typename K>
static void sort_launch(const Batch &b, const int64_t *offsets, K *keys, int64_t src_cap, PairOut out,
                        hipStream_t s) {
    int64_t blocks = (b.nq_total + 3) / 4;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_mlm_sort_short<K>, dim3((unsigned)blocks), dim3(kMlmBlock), 0, s, b.nq_total, offsets,
                       (const K *)keys, src_cap, out);
    hipLaunchKernelGGL(k_mlm_sort_rows<K>, dim3((unsigned)blocks), dim3(kMlmBlock), 0, s, b.nq_total, offsets, keys,
                       src_cap, out);
}

int mlm_sort_rows_launch(const Batch &b, const int64_t *offsets, uint64_t *scratch, PairOut out, hipStream_t s,
                         const uint32_t *slots, int64_t slot_cap) {
    if (b.nq_total == 0) return TCR_OK;
    if (slots) {
        if (slot_cap < 1 || slot_cap > kMlmSortMax) {
            set_error("slot capacity beyond the in-LDS row sort");
            return TCR_ERR_ARG;
        }
        sort_launch<uint32_t>(b, offsets, const_cast<uint32_t *>(slots), slot_cap, out, s);
    } else {
        sort_launch<uint64_t>(b, offsets, scratch, 0, out, s);
    }
    TCR_LAUNCH_CHECK();
    return TCR_OK;
}

}  // namespace tcr
