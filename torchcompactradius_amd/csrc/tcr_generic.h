// Device side of the per-query neighbour traversal (cpp/neighborhoodDynamic.h:37-113
// over cpp/neighborhood.h:35-94), shared by the per-query kernels of
// tcr_query.hip and the stray-list pass that rides on k_group_big
// (tcr_group.hip).
#pragma once
#include "tcr_internal.h"

namespace tcr {

constexpr int kQBlock = 256;
constexpr int kTraverseBatch = 4;  // candidates a per-query traversal loads together

// Candidate storage of the batched search: sorted float4 (x, y, z, h).
struct Float4Cands {
    const float4 *p;
    __device__ __forceinline__ void get(int64_t j, float *x, float &h) const {
        const float4 v = p[j];
        x[0] = v.x; x[1] = v.y; x[2] = v.z; h = v.w;
    }
};
// Candidate storage of the reference op layout: fp32 [nr, dim] + [nr].
template <int DIM>
struct RowCands {
    const float *p, *s;
    __device__ __forceinline__ void get(int64_t j, float *x, float &h) const {
#pragma unroll
        for (int d = 0; d < DIM; ++d) x[d] = p[j * DIM + d];
        h = s ? s[j] : 0.0f;
    }
};

// Reference layout: bucket hb of hashTable [L,2], scanned for the linear
// index (queryHashMap, cpp/neighborhood.h:35-58).  Returns [begin, end)
// relative to the candidate storage.
struct BucketLookup {
    const int32_t *hash_table, *cell_table;
    __device__ __forceinline__ bool find(int32_t lin, uint32_t hb, int64_t &b, int64_t &e) const {
        const int32_t b0 = hash_table[2 * (int64_t)hb], bl = hash_table[2 * (int64_t)hb + 1];
        if (b0 == -1) return false;
        for (int32_t i = b0; i < b0 + bl; ++i)
            if (cell_table[3 * (int64_t)i] == lin) {
                b = cell_table[3 * (int64_t)i + 1];
                e = b + cell_table[3 * (int64_t)i + 2];
                return true;
            }
        return false;
    }
};

// Search structure of this library: the cell with linear index lin, if it
// exists and lives in bucket hb -- the same answer as the bucket scan.
// Returns global sorted slots minus the cloud's first slot.
struct LinLookup {
    CellIndex index;
    const uint32_t *cell_hash, *cell_begin;
    int64_t rs;
    __device__ __forceinline__ bool find(int32_t lin, uint32_t hb, int64_t &b, int64_t &e) const {
        const int32_t cid = index.find(lin);
        if (cid < 0 || cell_hash[cid] != hb) return false;
        b = (int64_t)cell_begin[cid] - rs;
        e = (int64_t)cell_begin[cid + 1] - rs;
        return true;
    }
};

struct Geom {
    int32_t nc[3];
    uint32_t pmask, L;
    float mn[3], mx[3];
    float hcell;
    int32_t sr, mode;
};

// Visits the hits of query xi in reference order; emit(k, j) for the k-th
// hit (j relative to the candidate storage).  Returns the hit count.
template <int DIM, class Lookup, class Cands, class Emit>
__device__ __forceinline__ int64_t traverse(const Geom &g, const Lookup &lk, const Cands &cands, const float *xi,
                                            float hi, Emit emit) {
    int32_t q[3];
#pragma unroll
    for (int d = 0; d < DIM; ++d) q[d] = cell_coord(xi[d], g.mn[d], g.hcell);
    const int32_t w = 2 * g.sr + 1;
    int32_t noff = 1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) noff *= w;
    int64_t k = 0;
    for (int32_t o = 0; o < noff; ++o) {
        int32_t c[3];
        int32_t rem = o;
#pragma unroll
        for (int d = DIM - 1; d >= 0; --d) {
            const int32_t off = (rem % w) - g.sr;
            rem /= w;
            c[d] = ((g.pmask >> d) & 1u) ? pymod(q[d] + off, g.nc[d]) : q[d] + off;
        }
        int64_t cb, ce;
        if (!lk.find(linear_index<DIM>(c, g.nc), hash_index<DIM>(c, g.L), cb, ce)) continue;
        // candidates in batches whose loads are all in flight before the
        // first is tested (an emitted hit's store would otherwise make every
        // next load wait); the tail re-reads the cell's last candidate and
        // skips it
        for (int64_t j = cb; j < ce; j += kTraverseBatch) {
            float xj[kTraverseBatch][3], hj[kTraverseBatch];
#pragma unroll
            for (int u = 0; u < kTraverseBatch; ++u) cands.get(min(j + u, ce - 1), xj[u], hj[u]);
#pragma unroll
            for (int u = 0; u < kTraverseBatch; ++u) {
                if (j + u >= ce) break;
                const float dist = exact_distance<DIM>(xi, xj[u], g.mn, g.mx, g.pmask);
                if (support_pred(dist, hi, hj[u], g.mode)) {
                    emit(k, j + u);
                    ++k;
                }
            }
        }
    }
    return k;
}

// traverse() with one wavefront per query: the stencil walk is uniform and
// the lanes test 64 consecutive candidates of a cell at a time; hits are
// emitted in the same order (emit(k, j) for the k-th hit, from the lanes
// that hit).  Returns the hit count (uniform).  For queries with many
// candidates (strays of clouds with few, crowded cells).
template <int DIM, class Lookup, class Cands, class Emit>
__device__ __forceinline__ int64_t traverse_wave(const Geom &g, const Lookup &lk, const Cands &cands,
                                                 const float *xi, float hi, Emit emit) {
    const int lane = __lane_id();
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    int32_t q[3];
#pragma unroll
    for (int d = 0; d < DIM; ++d) q[d] = cell_coord(xi[d], g.mn[d], g.hcell);
    const int32_t w = 2 * g.sr + 1;
    int32_t noff = 1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) noff *= w;
    int64_t k = 0;
    for (int32_t ob = 0; ob < noff; ob += kWave) {
        // lane l resolves stencil cell ob + l: the lookups run in parallel
        int64_t lb = 0, le = 0;
        const int32_t ol = ob + lane;
        if (ol < noff) {
            int32_t c[3];
            int32_t rem = ol;
#pragma unroll
            for (int d = DIM - 1; d >= 0; --d) {
                const int32_t off = (rem % w) - g.sr;
                rem /= w;
                c[d] = ((g.pmask >> d) & 1u) ? pymod(q[d] + off, g.nc[d]) : q[d] + off;
            }
            if (!lk.find(linear_index<DIM>(c, g.nc), hash_index<DIM>(c, g.L), lb, le)) lb = le = 0;
        }
        const int32_t nob = min((int32_t)kWave, noff - ob);
        for (int32_t t = 0; t < nob; ++t) {  // stencil order
            const int64_t cb = __shfl(lb, t, kWave), ce = __shfl(le, t, kWave);
            for (int64_t jb = cb; jb < ce; jb += kWave) {
                const int64_t j = jb + lane;
                bool hit = false;
                if (j < ce) {
                    float xj[3], hj;
                    cands.get(j, xj, hj);
                    hit = support_pred(exact_distance<DIM>(xi, xj, g.mn, g.mx, g.pmask), hi, hj, g.mode);
                }
                const uint64_t m = __ballot(hit);
                if (hit) emit(k + __popcll(m & below), j);
                k += __popcll(m);
            }
        }
    }
    return k;
}

__device__ __forceinline__ Geom cloud_geom(const CloudDev &cd) {
    Geom g;
    for (int d = 0; d < 3; ++d) {
        g.nc[d] = cd.nc[d];
        g.mn[d] = cd.h.domain_min[d];
        g.mx[d] = cd.h.domain_max[d];
    }
    g.pmask = cd.h.periodic_mask;
    g.L = cd.h.hash_map_length;
    g.hcell = cd.hcell;
    g.sr = cd.h.search_range;
    g.mode = cd.h.mode;
    return g;
}

// Once-wrapped query position (py/neighborhood.py:202).
template <int DIM>
__device__ __forceinline__ void load_query(const CloudDev &cd, int64_t local, float *xi, float &hi) {
    const float *p = cd.qp + local * DIM;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        float v = p[d];
        if ((cd.h.periodic_mask >> d) & 1u)
            v = wrap_coord(v, cd.h.domain_min[d], cd.h.domain_max[d] - cd.h.domain_min[d]);
        xi[d] = v;
    }
    hi = cd.qs ? cd.qs[local] : 0.0f;
}

struct GenericArgs {
    const CloudDev *desc;
    int32_t B;
    const int64_t *q_start, *r_start;
    int64_t nq;
    const uint64_t *lintab;
    const uint32_t *cell_hash, *cell_begin;
    const float4 *sorted;
    const int32_t *sort_index;
    int32_t *counts;
    const int64_t *offsets;
    const int64_t *total;       // offsets + nq: the fill runs only if *total <= capacity
    int64_t capacity;
    PairOut out;
    const int32_t *list;        // stray list (list mode)
    const uint32_t *list_count;
};

template <int DIM, bool FILL>
__device__ __forceinline__ void query_one(const GenericArgs &a, const CloudDev &cd, int c, int64_t g,
                                          int64_t local) {
    float xi[3], hi;
    load_query<DIM>(cd, local, xi, hi);
    const Geom gm = cloud_geom(cd);
    const int64_t rs = a.r_start[c];
    const LinLookup lk{cell_index(cd, a.lintab), a.cell_hash, a.cell_begin, rs};
    const Float4Cands cands{a.sorted + rs};
    if (!FILL) {
        a.counts[g] = (int32_t)traverse<DIM>(gm, lk, cands, xi, hi, [](int64_t, int64_t) {});
    } else {
        const int64_t base = a.offsets[g];
        const int32_t *si = a.sort_index + rs;
        traverse<DIM>(gm, lk, cands, xi, hi, [&](int64_t k, int64_t j) { a.out.put(base + k, local, si[j]); });
    }
}

// One query per wavefront (every lane holds the same query).
template <int DIM, bool FILL>
__device__ __forceinline__ void query_one_wave(const GenericArgs &a, const CloudDev &cd, int c, int64_t g,
                                               int64_t local) {
    float xi[3], hi;
    load_query<DIM>(cd, local, xi, hi);
    const Geom gm = cloud_geom(cd);
    const int64_t rs = a.r_start[c];
    const LinLookup lk{cell_index(cd, a.lintab), a.cell_hash, a.cell_begin, rs};
    const Float4Cands cands{a.sorted + rs};
    if (!FILL) {
        const int64_t n = traverse_wave<DIM>(gm, lk, cands, xi, hi, [](int64_t, int64_t) {});
        if (__lane_id() == 0) a.counts[g] = (int32_t)n;
    } else {
        const int64_t base = a.offsets[g];
        const int32_t *si = a.sort_index + rs;
        traverse_wave<DIM>(gm, lk, cands, xi, hi, [&](int64_t k, int64_t j) { a.out.put(base + k, local, si[j]); });
    }
}

template <bool FILL>
__device__ __forceinline__ void generic_query_wave(const GenericArgs &a, int64_t g) {
    const int c = find_segment(a.B, g, [&](int k) { return a.q_start[k]; });
    const CloudDev &cd = a.desc[c];
    const int64_t local = g - a.q_start[c];
    switch (cd.h.dim) {
        case 1: query_one_wave<1, FILL>(a, cd, c, g, local); break;
        case 2: query_one_wave<2, FILL>(a, cd, c, g, local); break;
        default: query_one_wave<3, FILL>(a, cd, c, g, local); break;
    }
}

template <bool FILL>
__device__ __forceinline__ void generic_query(const GenericArgs &a, int64_t g) {
    const int c = find_segment(a.B, g, [&](int k) { return a.q_start[k]; });
    const CloudDev &cd = a.desc[c];
    const int64_t local = g - a.q_start[c];
    switch (cd.h.dim) {
        case 1: query_one<1, FILL>(a, cd, c, g, local); break;
        case 2: query_one<2, FILL>(a, cd, c, g, local); break;
        default: query_one<3, FILL>(a, cd, c, g, local); break;
    }
}

// Arguments of the per-query kernels for a batch.
inline GenericArgs generic_args(const Batch &b, const Work &w, const int64_t *offsets, PairOut out,
                                int64_t capacity) {
    GenericArgs a;
    a.desc = w.desc;
    a.B = b.B;
    a.q_start = w.q_start;
    a.r_start = w.r_start;
    a.nq = b.nq_total;
    a.lintab = w.lintab;
    a.cell_hash = w.cell_hash;
    a.cell_begin = w.cell_begin;
    a.sorted = w.sorted;
    a.sort_index = w.sort_index;
    a.counts = w.counts;
    a.offsets = offsets;
    a.total = offsets ? offsets + b.nq_total : nullptr;
    a.capacity = capacity;
    a.out = out;
    a.list = w.stray_list;
    a.list_count = w.misc + kMiscStrays;
    return a;
}

}  // namespace tcr
