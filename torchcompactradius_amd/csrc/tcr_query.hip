// Per-query neighbour traversal: count and fill passes.
//
// Semantics are those of countNeighborsForParticle / buildNeighborhood
// (cpp/neighborhoodDynamic.h:37-113) over iterateOffsetCells and
// queryHashMap (cpp/neighborhood.h:35-94): query cell from the once-wrapped
// query position, (2r+1)^d stencil in lexicographic order (dim 0 slowest),
// periodic neighbour cells through pymod, non-periodic ones unclamped, the
// bucket lookup matching only the linear index, the exact modDistance and
// the strict support predicate.  One lane per query (k_query_generic: every
// query of non-self searches and of stencil radius > 1); the one-wavefront-
// per-query form (tcr_generic.h) takes the strays of the wavefront-per-cell
// path inside k_group_big.
#include "tcr_generic.h"

namespace tcr {

// Every query of the clouds the wavefront-per-cell kernels do not own.
template <bool FILL>
__global__ void __launch_bounds__(kQBlock) k_query_generic(GenericArgs a) {
    if (FILL && *a.total > a.capacity) return;
    const int64_t g = (int64_t)blockIdx.x * kQBlock + threadIdx.x;
    if (g >= a.nq) return;
    const int c = find_segment(a.B, g, [&](int k) { return a.q_start[k]; });
    if (owned_by_cells(a.desc[c])) return;
    generic_query<FILL>(a, g);
}

int launch_generic(const Batch &b, Work &w, const int64_t *offsets, PairOut out, int64_t capacity, bool fill,
                   hipStream_t s) {
    if (b.nq_total == 0) return TCR_OK;
    const GenericArgs a = generic_args(b, w, offsets, out, capacity);
    const dim3 grid((unsigned)((b.nq_total + kQBlock - 1) / kQBlock)), block(kQBlock);
    if (fill) {
        fill_kernel_begin(s);
        hipLaunchKernelGGL(k_query_generic<true>, grid, block, 0, s, a);
        fill_kernel_end(s);
    } else {
        hipLaunchKernelGGL(k_query_generic<false>, grid, block, 0, s, a);
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
    Geom g;
    for (int d = 0; d < 3; ++d) {
        g.nc[d] = a.g.nc[d];
        g.mn[d] = a.g.mn[d];
        g.mx[d] = a.g.mx[d];
    }
    g.pmask = a.g.pmask;
    g.L = a.g.L;
    g.hcell = a.g.hcell;
    g.sr = a.g.sr;
    g.mode = a.g.mode;
    const BucketLookup lk{a.g.hash_table, a.g.cell_table};
    const RowCands<DIM> cands{a.spos, a.ssup};
    if (!FILL) {
        a.counts[i] = (int32_t)traverse<DIM>(g, lk, cands, xi, hi, [](int64_t, int64_t) {});
    } else {
        const int64_t base = a.offsets32[i];
        traverse<DIM>(g, lk, cands, xi, hi, [&](int64_t k, int64_t j) {
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
