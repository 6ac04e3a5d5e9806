// Per-query neighbour traversal: count and fill passes.
//
// Semantics are those of countNeighborsForParticle / buildNeighborhood
// (cpp/neighborhoodDynamic.h:37-113) over iterateOffsetCells and
// queryHashMap (cpp/neighborhood.h:35-94): query cell from the once-wrapped
// query position, (2r+1)^d stencil in lexicographic order (dim 0 slowest),
// periodic neighbour cells through pymod, non-periodic ones unclamped, the
// bucket scan matching only the linear index, the exact modDistance and the
// strict support predicate.  One lane per query; this is the general path
// that every input shape can take.
#include "tcr_internal.h"

namespace tcr {

constexpr int kQBlock = 256;

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

// Visits the hits of query xi in reference order; emit(k, j_sorted) for the
// k-th hit.  Returns the hit count.
template <int DIM, class Cands, class Emit>
__device__ __forceinline__ int64_t traverse(const GridView &g, const Cands &cands, const float *xi, float hi,
                                            Emit emit) {
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
        const int32_t lin = linear_index<DIM>(c, g.nc);
        const uint32_t hb = hash_index<DIM>(c, g.L);
        const int32_t b0 = g.hash_table[2 * (int64_t)hb], bl = g.hash_table[2 * (int64_t)hb + 1];
        if (b0 == -1) continue;
        int32_t cb = -1, ce = -1;
        for (int32_t i = b0; i < b0 + bl; ++i) {
            if (g.cell_table[3 * (int64_t)i] == lin) {
                cb = g.cell_table[3 * (int64_t)i + 1];
                ce = cb + g.cell_table[3 * (int64_t)i + 2];
                break;
            }
        }
        for (int32_t j = cb; j < ce; ++j) {
            float xj[3], hj;
            cands.get(j, xj, hj);
            const float dist = exact_distance<DIM>(xi, xj, g.mn, g.mx, g.pmask);
            if (support_pred(dist, hi, hj, g.mode)) {
                emit(k, j);
                ++k;
            }
        }
    }
    return k;
}

__device__ __forceinline__ GridView cloud_view(const CloudDev &cd, const int32_t *hash_table,
                                               const int32_t *cell_table) {
    GridView g;
    g.hash_table = hash_table + 2 * cd.hb_off;
    g.cell_table = cell_table + 3 * cd.cell_off;
    g.L = cd.h.hash_map_length;
    for (int d = 0; d < 3; ++d) {
        g.nc[d] = cd.nc[d];
        g.mn[d] = cd.h.domain_min[d];
        g.mx[d] = cd.h.domain_max[d];
    }
    g.pmask = cd.h.periodic_mask;
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

template <int DIM, bool FILL>
__device__ __forceinline__ void query_one(const CloudDev &cd, int64_t g, int64_t local,
                                          const int32_t *hash_table, const int32_t *cell_table,
                                          const float4 *sorted, const int64_t *r_start_c,
                                          const int32_t *sort_index, int32_t *counts,
                                          const int64_t *offsets, int64_t *row, int64_t *col) {
    float xi[3], hi;
    load_query<DIM>(cd, local, xi, hi);
    const GridView gv = cloud_view(cd, hash_table, cell_table);
    const int64_t rs = *r_start_c;
    Float4Cands cands{sorted + rs};
    if (!FILL) {
        counts[g] = (int32_t)traverse<DIM>(gv, cands, xi, hi, [](int64_t, int32_t) {});
    } else {
        const int64_t base = offsets[g];
        const int32_t *si = sort_index + rs;
        traverse<DIM>(gv, cands, xi, hi, [&](int64_t k, int32_t j) {
            col[base + k] = (int64_t)si[j];
            if (row) row[base + k] = local;
        });
    }
}

// Clouds whose queries the wavefront-per-cell kernels own (their strays
// come back through the stray list).
__device__ __forceinline__ bool group_owned(const CloudDev &cd) {
    return cd.h.self_search && cd.h.search_range == 1;
}

struct GenericArgs {
    const CloudDev *desc;
    int32_t B;
    const int64_t *q_start, *r_start;
    int64_t nq;
    const int32_t *hash_table, *cell_table;
    const float4 *sorted;
    const int32_t *sort_index;
    int32_t *counts;
    const int64_t *offsets;
    int64_t *row, *col;
    const int32_t *list;        // stray list (list mode)
    const uint32_t *list_count;
};

template <bool FILL>
__device__ __forceinline__ void generic_query(const GenericArgs &a, int64_t g) {
    const int c = find_segment(a.B, g, [&](int k) { return a.q_start[k]; });
    const CloudDev &cd = a.desc[c];
    const int64_t local = g - a.q_start[c];
    switch (cd.h.dim) {
        case 1: query_one<1, FILL>(cd, g, local, a.hash_table, a.cell_table, a.sorted, a.r_start + c, a.sort_index, a.counts, a.offsets, a.row, a.col); break;
        case 2: query_one<2, FILL>(cd, g, local, a.hash_table, a.cell_table, a.sorted, a.r_start + c, a.sort_index, a.counts, a.offsets, a.row, a.col); break;
        default: query_one<3, FILL>(cd, g, local, a.hash_table, a.cell_table, a.sorted, a.r_start + c, a.sort_index, a.counts, a.offsets, a.row, a.col); break;
    }
}

// Every query of the clouds the wavefront-per-cell kernels do not own.
template <bool FILL>
__global__ void __launch_bounds__(kQBlock) k_query_generic(GenericArgs a) {
    const int64_t g = (int64_t)blockIdx.x * kQBlock + threadIdx.x;
    if (g >= a.nq) return;
    const int c = find_segment(a.B, g, [&](int k) { return a.q_start[k]; });
    if (group_owned(a.desc[c])) return;
    generic_query<FILL>(a, g);
}

// The stray list left by the wavefront-per-cell count pass.
template <bool FILL>
__global__ void __launch_bounds__(kQBlock) k_query_list(GenericArgs a) {
    const int64_t n = *a.list_count;
    for (int64_t i = (int64_t)blockIdx.x * kQBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kQBlock)
        generic_query<FILL>(a, a.list[i]);
}

static GenericArgs generic_args(const Batch &b, Work &w, const int64_t *offsets, int64_t *row, int64_t *col) {
    GenericArgs a;
    a.desc = w.desc;
    a.B = b.B;
    a.q_start = w.q_start;
    a.r_start = w.r_start;
    a.nq = b.nq_total;
    a.hash_table = w.hash_table;
    a.cell_table = w.cell_table;
    a.sorted = w.sorted;
    a.sort_index = w.sort_index;
    a.counts = w.counts;
    a.offsets = offsets;
    a.row = row;
    a.col = col;
    a.list = w.stray_list;
    a.list_count = w.misc + 1;
    return a;
}

int launch_generic(const Batch &b, Work &w, const int64_t *offsets, int64_t *row, int64_t *col, bool fill,
                   bool list, hipStream_t s) {
    const GenericArgs a = generic_args(b, w, offsets, row, col);
    if (list) {
        const dim3 grid(1024), block(kQBlock);
        if (fill) hipLaunchKernelGGL(k_query_list<true>, grid, block, 0, s, a);
        else hipLaunchKernelGGL(k_query_list<false>, grid, block, 0, s, a);
    } else {
        if (b.nq_total == 0) return TCR_OK;
        const dim3 grid((unsigned)((b.nq_total + kQBlock - 1) / kQBlock)), block(kQBlock);
        if (fill) hipLaunchKernelGGL(k_query_generic<true>, grid, block, 0, s, a);
        else hipLaunchKernelGGL(k_query_generic<false>, grid, block, 0, s, a);
    }
    TCR_LAUNCH_CHECK();
    return TCR_OK;
}

// ------------------------------------------------ reference-layout ops ----

template <int DIM, bool FILL>
__global__ void __launch_bounds__(kQBlock) k_refop(RefOpArgs a) {
    const int64_t i = (int64_t)blockIdx.x * kQBlock + threadIdx.x;
    if (i >= a.nq) return;
    float xi[3];
#pragma unroll
    for (int d = 0; d < DIM; ++d) xi[d] = a.qpos[i * DIM + d];
    const float hi = a.qsup ? a.qsup[i] : 0.0f;
    RowCands<DIM> cands{a.spos, a.ssup};
    if (!FILL) {
        a.counts[i] = (int32_t)traverse<DIM>(a.g, cands, xi, hi, [](int64_t, int32_t) {});
    } else {
        const int64_t base = a.offsets32[i];
        traverse<DIM>(a.g, cands, xi, hi, [&](int64_t k, int32_t j) {
            a.li[base + k] = i;
            a.lj[base + k] = j;
        });
    }
}

int refop_launch(int dim, bool fill, const RefOpArgs &a, hipStream_t s) {
    if (a.nq == 0) return TCR_OK;
    const dim3 grid((unsigned)((a.nq + kQBlock - 1) / kQBlock)), block(kQBlock);
    if (!fill) {
        if (dim == 1) hipLaunchKernelGGL((k_refop<1, false>), grid, block, 0, s, a);
        else if (dim == 2) hipLaunchKernelGGL((k_refop<2, false>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((k_refop<3, false>), grid, block, 0, s, a);
    } else {
        if (dim == 1) hipLaunchKernelGGL((k_refop<1, true>), grid, block, 0, s, a);
        else if (dim == 2) hipLaunchKernelGGL((k_refop<2, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((k_refop<3, true>), grid, block, 0, s, a);
    }
    TCR_LAUNCH_CHECK();
    return TCR_OK;
}

__global__ void __launch_bounds__(kQBlock) k_hash_indices(const int32_t *__restrict__ cells, int64_t n, int dim,
                                                          uint32_t L, int32_t *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * kQBlock + threadIdx.x;
    if (i >= n) return;
    const int32_t *c = cells + i * dim;
    out[i] = dim == 1 ? (int32_t)hash_index<1>(c, L) : dim == 2 ? (int32_t)hash_index<2>(c, L)
                                                                : (int32_t)hash_index<3>(c, L);
}

int hash_indices_launch(const int32_t *cells, int64_t n, int dim, uint32_t L, int32_t *out, hipStream_t s) {
    if (n == 0) return TCR_OK;
    hipLaunchKernelGGL(k_hash_indices, dim3((unsigned)((n + kQBlock - 1) / kQBlock)), dim3(kQBlock), 0, s, cells,
                       n, dim, L, out);
    TCR_LAUNCH_CHECK();
    return TCR_OK;
}

}  // namespace tcr
