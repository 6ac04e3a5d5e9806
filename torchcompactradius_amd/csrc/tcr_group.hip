// Wavefront-per-cell count and fill passes for self-search clouds (the query
// set is the reference set), searchRange 1.
//
// A wavefront owns one occupied cell ("group") at a time.  Its lanes look up
// the 3^d stencil cells once through the compact hash table (same lookup
// semantics as queryHashMap, cpp/neighborhood.h:35-58), lay the candidates
// out in the reference's order -- stencil offset order, then sorted position
// -- across lanes and rounds, and then loop over the group's query points
// with the query held uniform.  Per (query, round) the 64 candidate tests
// give one hit mask; the count pass sums its popcounts, the fill pass writes
// the hits with a lane prefix (mbcnt) so that a row comes out in exactly the
// reference order (cpp/neighborhoodDynamic.h:76-113).
//
// Predicate: a filtered-exact test.  An fp32 estimate of the squared
// distance with a rigorous error margin decides "in" or "out"; only pairs in
// the margin band (and candidates outside the stencil box, e.g. aliased
// out-of-domain points) take the exact fp64 modDistance + sqrtf + '<' path
// of the reference (cpp/common.h:106-129).  Queries whose once-wrapped
// position differs from their twice-wrapped sorted copy, or whose cell lies
// outside [0, numCells), are "strays": they go to the per-query kernel.
#include "tcr_internal.h"

namespace tcr {

constexpr int kGroupThreads = 256;           // 4 wavefronts per workgroup
constexpr int kRounds = 8;                   // candidate rounds per chunk
constexpr int kChunk = kRounds * kWave;      // 512 candidates per chunk

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

template <typename T>
__device__ __forceinline__ T uniform(T v) {
    return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ float uniform_f(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// Thresholds on the squared distance for "dist < h": in if s < lo, out if
// s >= hi, exact test in between.  rel covers the fp32 estimate's and the
// reference's rounding (<= 10 ulp relative) plus sqrtf rounding; mabs covers
// the absolute error of shifted periodic coordinates (delta per dim).
struct Thr {
    float lo, hi;
};
__device__ __forceinline__ Thr thresholds(float h, float delta, int dim) {
    Thr t;
    if (!(h > 0.0f)) {          // h <= 0 or NaN: no distance is below it
        t.lo = -1.0f;
        t.hi = -1.0f;
    } else if (isinf(h)) {      // every finite distance is below +inf
        t.lo = INFINITY;
        t.hi = INFINITY;
    } else {
        const float h2 = h * h;
        if (!(h2 > 1e-30f)) {   // tiny supports: decide every pair exactly
            t.lo = -1.0f;
            t.hi = INFINITY;
        } else {
            const float mabs = (float)dim * (2.0f * h * delta + delta * delta) * 1.01f;
            t.lo = h2 * (1.0f - 0x1p-18f) - mabs;
            t.hi = h2 * (1.0f + 0x1p-18f) + mabs;
        }
    }
    return t;
}

struct GroupArgs {
    const CloudDev *desc;
    const int64_t *r_start;
    const int64_t *q_start;
    const uint32_t *misc;        // [0] = occupied cells in the batch
    const uint32_t *cell_begin;  // [cells+1] first global sorted slot
    const uint32_t *cell_cloud;
    const int32_t *hash_table;
    const int32_t *cell_table;
    const float4 *sorted;
    const int32_t *sort_index;   // local original index per sorted slot
    const uint8_t *sflags;       // per sorted slot: nonzero = stray query
    int32_t *counts;             // [nq] per global query index
    const int64_t *offsets;      // [nq+1] (fill)
    int64_t *row, *col;          // fill outputs (row may be null)
    int32_t *cursor;             // [nq] per-query write cursor across chunks (fill)
    int32_t *stray_list;         // [nq]
    uint32_t *stray_count;
    int add_strays;              // count pass appends strays
};

template <int DIM, int MODE, bool FILL>
__device__ void group_one(const GroupArgs &a, uint32_t cid) {
    const int lane = __lane_id();
    const uint32_t c = uniform(a.cell_cloud[cid]);
    const CloudDev &cd = a.desc[c];
    const int64_t rs = uniform(a.r_start[c]);
    const int64_t qs = uniform(a.q_start[c]);
    const uint32_t gb = uniform(a.cell_begin[cid]), ge = uniform(a.cell_begin[cid + 1]);

    float mn[3], mx[3], ext[3];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        mn[d] = cd.h.domain_min[d];
        mx[d] = cd.h.domain_max[d];
        ext[d] = mx[d] - mn[d];
    }
    const uint32_t pmask = cd.h.periodic_mask;
    const float hcell = cd.hcell;
    const bool cloud_fast = cd.fast != 0;

    // group cell from its first point; a class whose first point is outside
    // the grid (aliased linear index) is handled entirely per query
    const float4 p0 = a.sorted[gb];
    const float p0v[3] = {p0.x, p0.y, p0.z};
    int32_t q[3];
    bool inrange = true;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        q[d] = cell_coord(p0v[d], mn[d], hcell);
        inrange = inrange && q[d] >= 0 && q[d] < cd.nc[d];
    }
    if (!cloud_fast || !inrange) {
        if (a.add_strays)
            for (uint32_t s = gb + lane; s < ge; s += kWave) {
                const uint32_t k = atomicAdd(a.stray_count, 1u);
                a.stray_list[k] = (int32_t)(qs + a.sort_index[s]);
            }
        return;
    }

    // ---- stencil lookup: lane o < 3^DIM resolves offset o (dim 0 slowest)
    constexpr int kOff = DIM == 1 ? 3 : DIM == 2 ? 9 : 27;
    int32_t cbeg = 0, ccnt = 0;
    if (lane < kOff) {
        int32_t cc[3];
        int32_t rem = lane;
#pragma unroll
        for (int d = DIM - 1; d >= 0; --d) {
            const int32_t off = (rem % 3) - 1;
            rem /= 3;
            cc[d] = ((pmask >> d) & 1u) ? pymod(q[d] + off, cd.nc[d]) : q[d] + off;
        }
        const int32_t lin = linear_index<DIM>(cc, cd.nc);
        const uint32_t hb = hash_index<DIM>(cc, cd.h.hash_map_length);
        const int32_t *ht = a.hash_table + 2 * cd.hb_off;
        const int32_t *ct = a.cell_table + 3 * cd.cell_off;
        const int32_t b0 = ht[2 * (int64_t)hb], bl = ht[2 * (int64_t)hb + 1];
        if (b0 != -1)
            for (int32_t i = b0; i < b0 + bl; ++i)
                if (ct[3 * (int64_t)i] == lin) {
                    cbeg = ct[3 * (int64_t)i + 1];
                    ccnt = ct[3 * (int64_t)i + 2];
                    break;
                }
    }
    // inclusive scan of the stencil cell sizes over lanes
    int32_t incl = ccnt;
#pragma unroll
    for (int o = 1; o < 32; o <<= 1) {
        const int32_t n = __shfl_up(incl, o, kWave);
        if (lane >= o) incl += n;
    }
    const int32_t pre = incl - ccnt;                 // exclusive prefix of cell `lane`
    const int32_t total = uniform(__shfl(incl, kOff - 1, kWave));

    // geometry of the group cell for image shifts and the stencil-box check
    float center[3], delta = 0.0f;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        center[d] = mn[d] + ((float)q[d] + 0.5f) * hcell;
        if ((pmask >> d) & 1u) {
            const float m = fmaxf(fabsf(mn[d]), fabsf(mx[d]));
            delta = fmaxf(delta, 0x1p-22f * (m + ext[d]));
        }
    }
    const float box = 1.51f * hcell;  // stencil box half-width around the centre

    for (int32_t cbase = 0; cbase < total; cbase += kChunk) {
        // ---- stage this chunk's candidates in registers
        float cx[kRounds][3];
        float chh[kRounds];
        int32_t cj[kRounds];
        uint64_t vmask[kRounds], exmask[kRounds];
#pragma unroll
        for (int r = 0; r < kRounds; ++r) {
            const int32_t k = cbase + r * kWave + lane;
            const bool valid = k < total;
            // stencil cell holding candidate k: last o with pre[o] <= k
            int lo = 0;
#pragma unroll
            for (int step = 16; step > 0; step >>= 1) {
                const int mid = lo + step;
                const int32_t pm = __shfl(pre, mid < kOff ? mid : kOff - 1, kWave);
                if (mid < kOff && pm <= k) lo = mid;
            }
            const int32_t ob = __shfl(cbeg, lo, kWave), op = __shfl(pre, lo, kWave);
            const int32_t j = valid ? (int32_t)rs + ob + (k - op) : (int32_t)rs;
            cj[r] = j;
            const float4 v = a.sorted[valid ? j : gb];
            const float xv[3] = {v.x, v.y, v.z};
            bool well = valid;
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                float x = xv[d];
                if ((pmask >> d) & 1u) {
                    const float e0 = x - center[d];
                    if (e0 > 0.5f * ext[d]) x = x - ext[d];
                    else if (e0 < -0.5f * ext[d]) x = x + ext[d];
                }
                cx[r][d] = x;
                well = well && fabsf(x - center[d]) <= box;
            }
            chh[r] = v.w;
            vmask[r] = __ballot(valid);
            exmask[r] = __ballot(valid && !well);
        }
        int32_t corig[kRounds];
        if (FILL) {
#pragma unroll
            for (int r = 0; r < kRounds; ++r) corig[r] = a.sort_index[cj[r]];
        }

        // ---- loop over the group's queries
        for (uint32_t s = gb; s < ge; ++s) {
            const uint8_t fl = a.sflags[s];
            const int32_t qloc = uniform(a.sort_index[s]);
            const int64_t qg = qs + qloc;
            if (fl) {
                if (a.add_strays && cbase == 0 && lane == 0) {
                    const uint32_t k = atomicAdd(a.stray_count, 1u);
                    a.stray_list[k] = (int32_t)qg;
                }
                continue;
            }
            const float4 qv = a.sorted[s];
            const float xi[3] = {uniform_f(qv.x), uniform_f(qv.y), uniform_f(qv.z)};
            const float hi = uniform_f(qv.w);
            Thr tq{0.f, 0.f};
            if (MODE == TCR_MODE_GATHER) tq = thresholds(hi, delta, DIM);
            int32_t hits = 0;
            int64_t base = 0;
            if (FILL) base = a.offsets[qg] + (cbase ? (int64_t)a.cursor[qg] : 0);
#pragma unroll
            for (int r = 0; r < kRounds; ++r) {
                if (cbase + r * kWave >= total) break;
                float s2 = 0.0f;
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    const float e = xi[d] - cx[r][d];
                    s2 = d == 0 ? e * e : fmaf(e, e, s2);
                }
                Thr t;
                if (MODE == TCR_MODE_GATHER) {
                    t = tq;
                } else if (MODE == TCR_MODE_SCATTER) {
                    t = thresholds(chh[r], delta, DIM);
                } else if (MODE == TCR_MODE_SYMMETRIC) {
                    t = thresholds((hi + chh[r]) / 2.f, delta, DIM);
                } else {
                    t = thresholds(max_ref(hi, chh[r]), delta, DIM);
                }
                const uint64_t inm = __ballot(s2 < t.lo);
                const uint64_t maybe = __ballot(s2 < t.hi);
                uint64_t hitm = inm & vmask[r] & ~exmask[r];
                const uint64_t band = ((maybe & ~inm) | exmask[r]) & vmask[r];
                if (band) {
                    bool ok = false;
                    if ((band >> lane) & 1ull) {
                        const float4 v = a.sorted[cj[r]];
                        const float xj[3] = {v.x, v.y, v.z};
                        const float dist = exact_distance<DIM>(xi, xj, mn, mx, pmask);
                        ok = support_pred(dist, hi, v.w, MODE);
                    }
                    hitm |= __ballot(ok) & band;
                }
                if (FILL) {
                    if ((hitm >> lane) & 1ull) {
                        const int64_t pos = base + __popcll(hitm & lanemask_lt());
                        a.col[pos] = (int64_t)corig[r];
                        if (a.row) a.row[pos] = qloc;
                    }
                    base += __popcll(hitm);
                }
                hits += __popcll(hitm);
            }
            if (lane == 0) {
                if (!FILL) {
                    a.counts[qg] = cbase ? a.counts[qg] + hits : hits;
                } else {
                    a.cursor[qg] = (cbase ? a.cursor[qg] : 0) + hits;
                }
            }
        }
    }
}

template <int DIM, int MODE, bool FILL>
__global__ void __launch_bounds__(kGroupThreads) k_group(GroupArgs a) {
    const uint32_t ncells = a.misc[0];
    const int wave = threadIdx.x / kWave;
    // XCD-aware split: workgroup b runs on XCD b % 8; each XCD takes a
    // contiguous range of cells so that neighbouring cells share its L2.
    const uint32_t nxcd = 8;
    const uint32_t xcd = blockIdx.x % nxcd;
    const uint32_t per_xcd_waves = (gridDim.x / nxcd) * (kGroupThreads / kWave);
    const uint32_t wix = (blockIdx.x / nxcd) * (kGroupThreads / kWave) + wave;
    const uint32_t c0 = (uint32_t)(((uint64_t)ncells * xcd) / nxcd);
    const uint32_t c1 = (uint32_t)(((uint64_t)ncells * (xcd + 1)) / nxcd);
    for (uint32_t cid = c0 + wix; cid < c1; cid += per_xcd_waves) {
        const uint32_t cl = uniform(a.cell_cloud[cid]);
        if (a.desc[cl].h.dim != DIM || a.desc[cl].h.mode != MODE || !a.desc[cl].h.self_search ||
            a.desc[cl].h.search_range != 1)
            continue;
        group_one<DIM, MODE, FILL>(a, cid);
    }
}

template <bool FILL>
int launch_group_t(const Batch &b, const GroupArgs &ga, hipStream_t s) {
    // workgroups: a multiple of 8 (XCDs), up to 8 wavefront groups per CU
    int64_t waves = b.nr_total;  // upper bound on occupied cells
    int64_t wgs = (waves + 3) / 4;
    if (wgs > 256 * 8) wgs = 256 * 8;
    wgs = ((wgs + 7) / 8) * 8;
    const dim3 grid((unsigned)wgs), block(kGroupThreads);
    bool seen[4][4] = {};
    for (const CloudDev &d : b.dev) {
        if (!d.h.self_search || d.h.search_range != 1) continue;
        seen[d.h.dim][d.h.mode] = true;
    }
#define TCR_LAUNCH_GROUP(D, M)                                                             \
    if (seen[D][M]) hipLaunchKernelGGL((k_group<D, M, FILL>), grid, block, 0, s, ga);
#define TCR_LAUNCH_GROUP_DIM(D)                                                            \
    TCR_LAUNCH_GROUP(D, TCR_MODE_SYMMETRIC) TCR_LAUNCH_GROUP(D, TCR_MODE_GATHER)          \
    TCR_LAUNCH_GROUP(D, TCR_MODE_SCATTER) TCR_LAUNCH_GROUP(D, TCR_MODE_SUPERSYMMETRIC)
    TCR_LAUNCH_GROUP_DIM(1)
    TCR_LAUNCH_GROUP_DIM(2)
    TCR_LAUNCH_GROUP_DIM(3)
#undef TCR_LAUNCH_GROUP_DIM
#undef TCR_LAUNCH_GROUP
    TCR_LAUNCH_CHECK();
    return TCR_OK;
}

int launch_generic(const Batch &b, Work &w, const int64_t *offsets, int64_t *row, int64_t *col, bool fill,
                   bool list, hipStream_t s);

// Count or fill pass over the batch: wavefront-per-cell kernels for the
// self-search clouds (their strays collected by the count pass), the
// per-query kernel for every other cloud and for the strays.
static int query_pass(const Batch &b, Work &w, const int64_t *offsets, int64_t *row, int64_t *col, bool fill,
                      hipStream_t s) {
    if (b.nq_total == 0) return TCR_OK;
    bool any_group = false, any_other = false;
    for (const CloudDev &d : b.dev) {
        if (d.h.nq == 0) continue;
        if (d.h.self_search && d.h.search_range == 1) any_group = true;
        else any_other = true;
    }
    if (any_group) {
        if (!fill) TCR_HIP(hipMemsetAsync(w.misc + 1, 0, sizeof(uint32_t), s));
        GroupArgs ga;
        ga.desc = w.desc;
        ga.r_start = w.r_start;
        ga.q_start = w.q_start;
        ga.misc = w.misc;
        ga.cell_begin = w.cell_begin;
        ga.cell_cloud = w.cell_cloud;
        ga.hash_table = w.hash_table;
        ga.cell_table = w.cell_table;
        ga.sorted = w.sorted;
        ga.sort_index = w.sort_index;
        ga.sflags = w.sflags;
        ga.counts = w.counts;
        ga.offsets = offsets;
        ga.row = row;
        ga.col = col;
        ga.cursor = w.cursor;
        ga.stray_list = w.stray_list;
        ga.stray_count = w.misc + 1;
        ga.add_strays = fill ? 0 : 1;
        int rc = fill ? launch_group_t<true>(b, ga, s) : launch_group_t<false>(b, ga, s);
        if (rc) return rc;
        rc = launch_generic(b, w, offsets, row, col, fill, true, s);
        if (rc) return rc;
    }
    if (any_other) return launch_generic(b, w, offsets, row, col, fill, false, s);
    return TCR_OK;
}

int query_count(const Batch &b, Work &w, hipStream_t s) { return query_pass(b, w, nullptr, nullptr, nullptr, false, s); }
int query_fill(const Batch &b, Work &w, const int64_t *offsets, int64_t *row, int64_t *col, hipStream_t s) {
    return query_pass(b, w, offsets, row, col, true, s);
}

}  // namespace tcr
