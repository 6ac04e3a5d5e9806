// Structure build on the device: hMax, hCell, wrapped cell keys, sorted
// positions, cell table and compact hash table -- the GPU form of
// neighborSearch_cpp's preamble (py/compactHashing/cppSearch.py:215-224) and
// buildCompactHashMap_compat (py/compactHashing/hashTable.py:79-154).
#include <algorithm>

#include "tcr_internal.h"

namespace tcr {

constexpr int kBlock = 256;

__device__ __forceinline__ int cloud_of(int32_t B, int64_t g, const int64_t *__restrict__ start) {
    return find_segment(B, g, [&](int k) { return start[k]; });
}

// hMax per cloud (computeGridSupport, py/compactHashing/cellTable.py:40-81):
// gather reads query supports, scatter reference supports, the symmetric
// modes both.  which = 0: query set, 1: reference set.
// Each workgroup reduces one contiguous slice of the concatenated supports;
// a thread keeps a running maximum for its current cloud and flushes it with
// one atomicMax when the cloud changes, and the wavefront merges lanes that
// end on the same cloud -- a few thousand atomics for any batch.
constexpr int kHmaxBlocks = 512;
__global__ void __launch_bounds__(kBlock) k_hmax(CloudDev *__restrict__ desc, int32_t B,
                                                 const int64_t *__restrict__ start, int64_t n, int which) {
    const int64_t per = ((n + gridDim.x - 1) / gridDim.x + kBlock - 1) / kBlock * kBlock;
    const int64_t g0 = (int64_t)blockIdx.x * per, g1 = min(n, g0 + per);
    int c = -1;
    uint32_t m = 0u;  // ordered encoding: 0 is below every float
    bool use = false;
    const float *sup = nullptr;
    int64_t c_lo = 0, c_hi = -1;  // [start[c], start[c+1])
    for (int64_t g = g0 + threadIdx.x; g < g1; g += kBlock) {
        if (g >= c_hi) {
            if (c >= 0 && m) atomicMax(&desc[c].hmax_ord, m);
            c = cloud_of(B, g, start);
            c_lo = start[c];
            c_hi = start[c + 1];
            m = 0u;
            const CloudDev &cd = desc[c];
            const int mode = cd.h.mode;
            // a self-search cloud's query set is its reference set
            use = which == 0 ? (mode != TCR_MODE_SCATTER || cd.h.self_search)
                             : (mode != TCR_MODE_GATHER && !cd.h.self_search);
            sup = which == 0 ? cd.qs : cd.rs;
            if (!sup) use = false;
        }
        if (use) m = max(m, float_to_ordered(sup[g - c_lo]));
    }
    const int c0 = __shfl(c, 0, kWave);
    if (__all(c == c0)) {
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, kWave));
        if ((threadIdx.x & (kWave - 1)) == 0 && c0 >= 0 && m) atomicMax(&desc[c0].hmax_ord, m);
    } else if (c >= 0 && m) {
        atomicMax(&desc[c].hmax_ord, m);
    }
}

// hCell = compute_h(min, max, hMax) (py/util.py:28-79) in fp32, numCells =
// ceil(extent / hCell).int32 (py/compactHashing/cellTable.py:29).
__global__ void k_grid(CloudDev *__restrict__ desc, int32_t B) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= B) return;
    CloudDev &cd = desc[c];
    const int dim = cd.h.dim;
    const float hmax = cd.hmax_ord ? ordered_to_float(cd.hmax_ord) : -INFINITY;
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
    uint32_t err = 0;
    // wavefront-per-cell eligibility: self-search, stencil radius 1, a
    // finite positive cell size, and >= 6 cells in every periodic dimension
    // (so that every candidate in the stencil box has one minimal image)
    uint32_t fast = (cd.h.self_search && cd.h.search_range == 1 && hc > 0.0f && !isinf(hc)) ? 1u : 0u;
    for (int d = 0; d < 3; ++d) {
        if (d < dim) {
            const float v = ceilf(ext[d] / hc);
            cd.nc[d] = (v >= -2147483648.0f && v < 2147483648.0f) ? (int32_t)v : INT32_MIN;
            if (((cd.h.periodic_mask >> d) & 1u) && cd.nc[d] <= 0 && cd.h.nq > 0 && cd.h.nr > 0) err = 1;
            if (cd.nc[d] <= 0 || (((cd.h.periodic_mask >> d) & 1u) && cd.nc[d] < 6)) fast = 0;
        } else {
            cd.nc[d] = 1;
        }
    }
    cd.err = err;
    cd.fast = fast;
}

// Reference set: second wrap (cppSearch.py:221) on top of radiusSearch's
// first (neighborhood.py:203), cell coordinates and int32 linear key
// (cellTable.py:30-31).  Keys are stored sign-flipped so that an unsigned
// radix sort yields the signed order of torch.argsort.
// Stray bits for the wavefront-per-cell path: bit 0 -- the second wrap moved
// the point (its query copy, wrapped once, differs from its sorted copy);
// bit 1 -- its cell lies outside [0, numCells) in some dimension.
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

__global__ void __launch_bounds__(kBlock) k_ref_keys(const CloudDev *__restrict__ desc, int32_t B,
                                                     const int64_t *__restrict__ r_start, int64_t n,
                                                     float4 *__restrict__ rwrap, uint32_t *__restrict__ keys,
                                                     uint32_t *__restrict__ vals, uint8_t *__restrict__ pflags) {
    const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (g >= n) return;
    const int c = cloud_of(B, g, r_start);
    const CloudDev &cd = desc[c];
    const int64_t local = g - r_start[c];
    float4 wp;
    int32_t lin;
    uint8_t f;
    switch (cd.h.dim) {
        case 1: ref_key_one<1>(cd, local, wp, lin, f); break;
        case 2: ref_key_one<2>(cd, local, wp, lin, f); break;
        default: ref_key_one<3>(cd, local, wp, lin, f); break;
    }
    rwrap[g] = wp;
    keys[g] = ((uint32_t)lin) ^ 0x80000000u;
    vals[g] = (uint32_t)local;
    pflags[g] = f;
}

// sortedPositions = x[sortIndex] (hashTable.py:109), sortedSupports
// (cppSearch.py:224), sorted keys, and the cell-start flags of
// unique_consecutive (hashTable.py:113).
__global__ void __launch_bounds__(kBlock) k_gather_sorted(const CloudDev *__restrict__ desc, int32_t B,
                                                          const int64_t *__restrict__ r_start, int64_t n,
                                                          const uint32_t *__restrict__ keys,
                                                          const uint32_t *__restrict__ vals,
                                                          const float4 *__restrict__ rwrap,
                                                          float4 *__restrict__ sorted, int32_t *__restrict__ sidx,
                                                          int32_t *__restrict__ skey, uint32_t *__restrict__ flags,
                                                          const uint8_t *__restrict__ pflags,
                                                          uint8_t *__restrict__ sflags) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const int c = cloud_of(B, s, r_start);
    const uint32_t v = vals[s];
    const int32_t k = (int32_t)(keys[s] ^ 0x80000000u);
    sorted[s] = rwrap[r_start[c] + v];
    sflags[s] = pflags[r_start[c] + v];
    sidx[s] = (int32_t)v;
    skey[s] = k;
    const bool first = (s == r_start[c]) || ((int32_t)(keys[s - 1] ^ 0x80000000u) != k);
    flags[s] = first ? 1u : 0u;
}

// One record per occupied cell: first slot, linear index, and the hash of
// the first point's cell coordinates (hashTable.py:121-136).
template <int DIM>
__device__ __forceinline__ uint32_t cell_hash_of(const CloudDev &cd, float4 p) {
    const float x[3] = {p.x, p.y, p.z};
    int32_t c[3];
#pragma unroll
    for (int d = 0; d < DIM; ++d) c[d] = cell_coord(x[d], cd.h.domain_min[d], cd.hcell);
    return hash_index<DIM>(c, cd.h.hash_map_length);
}

__global__ void __launch_bounds__(kBlock) k_cell_emit(const CloudDev *__restrict__ desc, int32_t B,
                                                      const int64_t *__restrict__ r_start, int64_t n,
                                                      const uint32_t *__restrict__ flags,
                                                      const uint32_t *__restrict__ cellpos,
                                                      const int32_t *__restrict__ skey,
                                                      const float4 *__restrict__ sorted,
                                                      uint32_t *__restrict__ cell_begin, int32_t *__restrict__ cell_lin,
                                                      uint32_t *__restrict__ cell_hash,
                                                      uint32_t *__restrict__ cell_cloud, uint32_t *__restrict__ misc) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    if (s == n - 1) {
        const uint32_t total = cellpos[s] + flags[s];
        cell_begin[total] = (uint32_t)n;
        misc[0] = total;
    }
    if (!flags[s]) return;
    const int c = cloud_of(B, s, r_start);
    const CloudDev &cd = desc[c];
    const uint32_t cid = cellpos[s];
    cell_begin[cid] = (uint32_t)s;
    cell_lin[cid] = skey[s];
    cell_cloud[cid] = (uint32_t)c;
    uint32_t h;
    switch (cd.h.dim) {
        case 1: h = cell_hash_of<1>(cd, sorted[s]); break;
        case 2: h = cell_hash_of<2>(cd, sorted[s]); break;
        default: h = cell_hash_of<3>(cd, sorted[s]); break;
    }
    cell_hash[cid] = h;
}

__global__ void k_cloud_cells(CloudDev *__restrict__ desc, int32_t B, const int64_t *__restrict__ r_start,
                              const uint32_t *__restrict__ flags, const uint32_t *__restrict__ cellpos) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= B) return;
    const int64_t nr = r_start[c + 1] - r_start[c];
    if (nr > 0) {
        const int64_t last = r_start[c + 1] - 1;
        desc[c].cell_off = cellpos[r_start[c]];
        desc[c].ncells = (int64_t)cellpos[last] + flags[last] - cellpos[r_start[c]];
    } else {
        desc[c].cell_off = 0;
        desc[c].ncells = 0;
    }
}

__global__ void __launch_bounds__(kBlock) k_bucket_count(const CloudDev *__restrict__ desc, int64_t ncap,
                                                         const uint32_t *__restrict__ misc,
                                                         const uint32_t *__restrict__ cell_hash,
                                                         const uint32_t *__restrict__ cell_cloud,
                                                         uint32_t *__restrict__ bucket_cnt) {
    const int64_t cid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (cid >= ncap || cid >= (int64_t)misc[0]) return;
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
    if (cid >= ncap || cid >= (int64_t)misc[0]) return;
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

static inline unsigned blocks_for(int64_t n, int t = kBlock) { return (unsigned)((n + t - 1) / t); }

int build_structure(const Batch &b, Work &w, hipStream_t s) {
    const int32_t B = b.B;
    if (b.nq_total > 0)
        hipLaunchKernelGGL(k_hmax, dim3((unsigned)std::min<int64_t>(kHmaxBlocks, blocks_for(b.nq_total))),
                           dim3(kBlock), 0, s, w.desc, B, (const int64_t *)w.q_start, b.nq_total, 0);
    if (b.nr_total > 0)
        hipLaunchKernelGGL(k_hmax, dim3((unsigned)std::min<int64_t>(kHmaxBlocks, blocks_for(b.nr_total))),
                           dim3(kBlock), 0, s, w.desc, B, (const int64_t *)w.r_start, b.nr_total, 1);
    hipLaunchKernelGGL(k_grid, dim3(blocks_for(B, 64)), dim3(64), 0, s, w.desc, B);
    TCR_LAUNCH_CHECK();
    const int64_t n = b.nr_total;
    if (n > 0) {
        hipLaunchKernelGGL(k_ref_keys, dim3(blocks_for(n)), dim3(kBlock), 0, s, (const CloudDev *)w.desc, B,
                           (const int64_t *)w.r_start, n, w.rwrap, w.keys0, w.vals0, w.pflags);
        TCR_LAUNCH_CHECK();
        int rc = radix_sort_ref(b, w, 4, s);
        if (rc) return rc;
        hipLaunchKernelGGL(k_gather_sorted, dim3(blocks_for(n)), dim3(kBlock), 0, s, (const CloudDev *)w.desc, B,
                           (const int64_t *)w.r_start, n, (const uint32_t *)w.keys0, (const uint32_t *)w.vals0,
                           (const float4 *)w.rwrap, w.sorted, w.sort_index, w.skey, w.flags,
                           (const uint8_t *)w.pflags, w.sflags);
        rc = exclusive_scan<uint32_t, uint32_t>(w.flags, w.cellpos, n, false, w.scan_part, s);
        if (rc) return rc;
        hipLaunchKernelGGL(k_cell_emit, dim3(blocks_for(n)), dim3(kBlock), 0, s, (const CloudDev *)w.desc, B,
                           (const int64_t *)w.r_start, n, (const uint32_t *)w.flags, (const uint32_t *)w.cellpos,
                           (const int32_t *)w.skey, (const float4 *)w.sorted, w.cell_begin, w.cell_lin,
                           w.cell_hash, w.cell_cloud, w.misc);
        TCR_LAUNCH_CHECK();
    } else {
        TCR_HIP(hipMemsetAsync(w.misc, 0, sizeof(uint32_t), s));
    }
    hipLaunchKernelGGL(k_cloud_cells, dim3(blocks_for(B, 64)), dim3(64), 0, s, w.desc, B,
                       (const int64_t *)w.r_start, (const uint32_t *)w.flags, (const uint32_t *)w.cellpos);
    TCR_HIP(hipMemsetAsync(w.bucket_cnt, 0, sizeof(uint32_t) * (size_t)b.hb_total, s));
    TCR_HIP(hipMemsetAsync(w.bucket_fill, 0, sizeof(uint32_t) * (size_t)b.hb_total, s));
    if (n > 0)
        hipLaunchKernelGGL(k_bucket_count, dim3(blocks_for(n)), dim3(kBlock), 0, s, (const CloudDev *)w.desc, n,
                           (const uint32_t *)w.misc, (const uint32_t *)w.cell_hash, (const uint32_t *)w.cell_cloud,
                           w.bucket_cnt);
    int rc = exclusive_scan<uint32_t, uint32_t>(w.bucket_cnt, w.bucket_begin, b.hb_total, false, w.scan_part, s);
    if (rc) return rc;
    if (n > 0)
        hipLaunchKernelGGL(k_bucket_scatter, dim3(blocks_for(n)), dim3(kBlock), 0, s, (const CloudDev *)w.desc, n,
                           (const uint32_t *)w.misc, (const uint32_t *)w.cell_hash, (const uint32_t *)w.cell_cloud,
                           (const uint32_t *)w.cell_begin, (const int32_t *)w.cell_lin, (const int64_t *)w.r_start,
                           (const uint32_t *)w.bucket_begin, w.bucket_fill, w.cell_table);
    hipLaunchKernelGGL(k_hash_table, dim3(blocks_for(b.hb_total)), dim3(kBlock), 0, s, (const CloudDev *)w.desc, B,
                       (const int64_t *)w.hb_start, b.hb_total, (const uint32_t *)w.bucket_cnt,
                       (const uint32_t *)w.bucket_begin, w.hash_table);
    TCR_LAUNCH_CHECK();
    return TCR_OK;
}

}  // namespace tcr
