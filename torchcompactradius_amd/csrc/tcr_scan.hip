// Device-wide exclusive scan (reduce -> scan partials -> apply), wave64.
// Used for the offsets scan of the counts (cppSearch.py:81-82 restated with
// int64 offsets), radix histograms, cell flags and hash buckets.
#include <algorithm>

#include "tcr_internal.h"

namespace tcr {

template <typename Tin>
__global__ void __launch_bounds__(kScanThreads) k_scan_reduce(const Tin *__restrict__ in, int64_t n,
                                                               uint64_t *__restrict__ part) {
    __shared__ uint64_t red[kScanThreads / kWave];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    uint64_t s = 0;
    for (int k = threadIdx.x; k < kScanTile; k += kScanThreads) {
        const int64_t i = base + k;
        if (i < n) s += (uint64_t)(int64_t)in[i];
    }
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int w = 0; w < kScanThreads / kWave; ++w) t += red[w];
        part[blockIdx.x] = t;
    }
}

// Single block: exclusive scan of nb partials in place; part[nb] = total.
__global__ void __launch_bounds__(kScanThreads) k_scan_partials(uint64_t *__restrict__ part, int64_t nb) {
    __shared__ uint64_t lds[kScanTile];
    __shared__ uint64_t wt[kScanThreads / kWave];
    uint64_t carry = 0;
    for (int64_t base = 0; base < nb; base += kScanTile) {
        for (int k = threadIdx.x; k < kScanTile; k += kScanThreads)
            lds[k] = (base + k < nb) ? part[base + k] : 0;
        __syncthreads();
        const uint64_t tot = block_exclusive_scan_tile<uint64_t>(lds, wt);
        for (int k = threadIdx.x; k < kScanTile; k += kScanThreads)
            if (base + k < nb) part[base + k] = lds[k] + carry;
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) part[nb] = carry;
}

// Applies the block prefixes; with int64 offsets it also writes the CSR
// arrays when asked for (indptr int32 [n+1] = the offsets, rowEntries int64
// [n] = the counts).
template <typename Tin, typename Tout>
__global__ void __launch_bounds__(kScanThreads) k_scan_apply(const Tin *__restrict__ in, Tout *__restrict__ out,
                                                              int64_t n, const uint64_t *__restrict__ part,
                                                              int64_t nb, int write_total,
                                                              int32_t *__restrict__ indptr,
                                                              int64_t *__restrict__ row_entries) {
    __shared__ uint64_t lds[kScanTile];
    __shared__ uint64_t wt[kScanThreads / kWave];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    for (int k = threadIdx.x; k < kScanTile; k += kScanThreads) {
        const int64_t i = base + k;
        lds[k] = (i < n) ? (uint64_t)(int64_t)in[i] : 0;
    }
    __syncthreads();
    block_exclusive_scan_tile<uint64_t>(lds, wt);
    const uint64_t pre = part[blockIdx.x];
    for (int k = threadIdx.x; k < kScanTile; k += kScanThreads) {
        const int64_t i = base + k;
        if (i < n) {
            out[i] = (Tout)(lds[k] + pre);
            if (indptr) indptr[i] = (int32_t)(int64_t)(lds[k] + pre);
            if (row_entries) row_entries[i] = (int64_t)in[i];
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (write_total) out[n] = (Tout)part[nb];
        if (indptr) indptr[n] = (int32_t)(int64_t)part[nb];
    }
}

// counts -> int64 offsets in one launch (k_scan_reduce / k_scan_partials /
// k_scan_apply in one): each tile scans its kOffsetsTile counts in LDS, takes
// its prefix by look-back, and writes offsets (plus the CSR arrays when
// asked for); the last tile writes the total.  Tiles of 8192 counts: a
// million queries make 123 tiles, whose look-back windows of 64 reach an
// inclusive prefix within two rounds.
constexpr int kOffsetsThreads = 1024;
constexpr int kOffsetsItems = 8;
constexpr int kOffsetsTile = kOffsetsThreads * kOffsetsItems;

__global__ void __launch_bounds__(kOffsetsThreads) k_offsets_scan(const int32_t *__restrict__ in,
                                                                   int64_t *__restrict__ out, int64_t n,
                                                                   uint64_t *__restrict__ status,
                                                                   uint32_t *__restrict__ ticket,
                                                                   int32_t *__restrict__ indptr,
                                                                   int64_t *__restrict__ row_entries) {
    __shared__ uint64_t lds[kOffsetsTile];
    __shared__ uint64_t wt[kOffsetsThreads / kWave];
    __shared__ uint64_t s_excl;
    __shared__ uint32_t s_tile;
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    const int64_t tile = s_tile;
    const int64_t base = tile * kOffsetsTile;
    for (int k = threadIdx.x; k < kOffsetsTile; k += kOffsetsThreads) {
        const int64_t i = base + k;
        lds[k] = (i < n) ? (uint64_t)(int64_t)in[i] : 0;
    }
    __syncthreads();
    const uint64_t agg = block_exclusive_scan_tile<uint64_t, kOffsetsThreads, kOffsetsItems>(lds, wt);
    if (threadIdx.x < kWave) {
        const uint64_t excl = tile_look_back(status, tile, agg);
        if (threadIdx.x == 0) s_excl = excl;
    }
    __syncthreads();
    const uint64_t pre = s_excl;
    for (int k = threadIdx.x; k < kOffsetsTile; k += kOffsetsThreads) {
        const int64_t i = base + k;
        if (i < n) {
            const int64_t o = (int64_t)(lds[k] + pre);
            out[i] = o;
            if (indptr) indptr[i] = (int32_t)o;
            if (row_entries) row_entries[i] = (int64_t)in[i];
        }
    }
    if (base + kOffsetsTile >= n && threadIdx.x == 0) {  // the last tile
        out[n] = (int64_t)(pre + agg);
        if (indptr) indptr[n] = (int32_t)(int64_t)(pre + agg);
    }
}

// Outputs of an empty query set: total 0 and indptr [0] -- a kernel, not a
// memset, so that a captured graph holds no memset node.
template <typename Tout>
__global__ void k_zero_empty(Tout *total, int32_t *indptr) {
    if (total) total[0] = 0;
    if (indptr) indptr[0] = 0;
}

size_t scan_partials_needed(int64_t n) { return (size_t)((n + kScanTile - 1) / kScanTile) + 2; }

template <typename Tin, typename Tout>
int exclusive_scan(const Tin *in, Tout *out, int64_t n, bool write_total, uint64_t *partials,
                   hipStream_t s, int32_t *indptr, int64_t *row_entries) {
    const int64_t nb = (n + kScanTile - 1) / kScanTile;
    if (nb == 0) {
        if (write_total || indptr) {
            hipLaunchKernelGGL(k_zero_empty<Tout>, dim3(1), dim3(1), 0, s, write_total ? out : (Tout *)nullptr,
                               indptr);
            TCR_LAUNCH_CHECK();
        }
        return TCR_OK;
    }
    hipLaunchKernelGGL(k_scan_reduce<Tin>, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, n, partials);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kScanThreads), 0, s, partials, nb);
    hipLaunchKernelGGL((k_scan_apply<Tin, Tout>), dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, out, n,
                       (const uint64_t *)partials, nb, write_total ? 1 : 0, indptr, row_entries);
    TCR_LAUNCH_CHECK();
    return TCR_OK;
}

template int exclusive_scan<uint32_t, uint32_t>(const uint32_t *, uint32_t *, int64_t, bool, uint64_t *,
                                                hipStream_t, int32_t *, int64_t *);
template int exclusive_scan<int32_t, int64_t>(const int32_t *, int64_t *, int64_t, bool, uint64_t *,
                                              hipStream_t, int32_t *, int64_t *);

int64_t offsets_scan_tiles(int64_t n) { return (n + kOffsetsTile - 1) / kOffsetsTile; }

int offsets_scan(const int32_t *counts, int64_t *offsets, int64_t n, uint64_t *status, uint32_t *ticket,
                 hipStream_t s, int32_t *indptr, int64_t *row_entries) {
    const int64_t nb = offsets_scan_tiles(n);
    if (nb == 0) {
        hipLaunchKernelGGL(k_zero_empty<int64_t>, dim3(1), dim3(1), 0, s, offsets, indptr);
    } else {
        hipLaunchKernelGGL(k_offsets_scan, dim3((unsigned)nb), dim3(kOffsetsThreads), 0, s, counts, offsets, n, status,
                           ticket, indptr, row_entries);
    }
    TCR_LAUNCH_CHECK();
    return TCR_OK;
}

// CSR arrays of the reference's SparseCSR from the int64 offsets
// (neighborhood.py:303-305, util.py:335-358): indptr int32[n+1] and
// rowEntries int64[n].
__global__ void __launch_bounds__(kScanThreads) k_csr_arrays(const int64_t *__restrict__ offsets, int64_t n,
                                                             int32_t *__restrict__ indptr,
                                                             int64_t *__restrict__ row_entries) {
    for (int64_t i = (int64_t)blockIdx.x * kScanThreads + threadIdx.x; i <= n; i += (int64_t)gridDim.x * kScanThreads) {
        const int64_t o = offsets[i];
        if (indptr) indptr[i] = (int32_t)o;
        if (row_entries && i < n) row_entries[i] = offsets[i + 1] - o;
    }
}

int csr_arrays_launch(const int64_t *offsets, int64_t n, int32_t *indptr, int64_t *row_entries, hipStream_t s) {
    const int64_t blocks = std::min<int64_t>(4096, (n + 1 + kScanThreads - 1) / kScanThreads);
    hipLaunchKernelGGL(k_csr_arrays, dim3((unsigned)blocks), dim3(kScanThreads), 0, s, offsets, n, indptr,
                       row_entries);
    TCR_LAUNCH_CHECK();
    return TCR_OK;
}

}  // namespace tcr
