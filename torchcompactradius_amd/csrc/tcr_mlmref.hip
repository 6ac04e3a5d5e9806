// algorithm='mlm' with the reference's own semantics (a separate mode; the
// default 'mlm' of this library returns the compact result): the multi-level
// search of torchCompactRadius 0.5.5's searchNeighbors_mlm
// (py/multiLevelMemory/wrapper.py:145-308) with its hash-map cell table,
// kept step for step:
//   * levels (py/multiLevelMemory/cellOp.py:5-46), computed on the host from
//     the reference supports' min and max (the reference's .item() reads);
//   * points sorted by the Morton code of a 2^16 grid (wrapper.py:68-77) --
//     a stable LSD radix sort here;
//   * per-level codes (morton(cell at hCell) >> 2*level).int()
//     (py/multiLevelMemory/indexing.py:71-78), cells as runs of equal codes
//     in sorted order merged over the levels (cellData.py:44-62,
//     cellOp.py:48-64);
//   * the traversal (cpp/multiLevelMemory/mlmUtil.h:234-282): level from the
//     point's support, C-remainder neighbour cells, cells of the hash bucket
//     matching level and Morton code.  A cell lies in bucket
//     hash(decode(code)) (hashData.py:21-39, torch's int32 decode), so the
//     bucket scan is a lookup of (level, code) in the cells sorted by that
//     key, taken when hash(decode(code)) equals the bucket of the offset
//     cell; matching cells come in cell order, as the bucket lists them;
//   * counts over the reference set at its unwrapped positions with the
//     superSymmetric support sync (countNeighborsHashmap.h:99-124), int32
//     offsets of length Nq (wrapper.py:255-270), the fill over the wrapped
//     query set (buildNeighborhoodHashmap.h:103-128), col = sortingIndices[col].
// A query that finds more pairs than its row holds (the reference writes
// past it) sets an error flag instead of writing.
#include <cmath>
#include <cstring>

#include "tcr_internal.h"

namespace tcr {
namespace {

constexpr int kT = 256;          // threads per block
constexpr int kRsTile = 1024;    // keys per radix tile (4 rounds of 256)
constexpr int kMaxLevels = 17;   // 2^16 cells per dimension at most

__device__ __forceinline__ int32_t trunc_x86(float v) {
    return (v >= -2147483648.0f && v < 2147483648.0f) ? (int32_t)v : INT32_MIN;
}

__device__ __forceinline__ uint64_t spread1(uint64_t x) {  // indexing.py:29-35
    x &= 0x00000000FFFFFFFFull;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}
__device__ __forceinline__ uint64_t spread2(uint64_t w) {  // indexing.py:3-10
    w &= 0x00000000001FFFFFull;
    w = (w | w << 32) & 0x001F00000000FFFFull;
    w = (w | w << 16) & 0x001F0000FF0000FFull;
    w = (w | w << 8) & 0x010F00F00F00F00Full;
    w = (w | w << 4) & 0x10C30C30C30C30C3ull;
    w = (w | w << 2) & 0x1249249249249249ull;
    return w;
}
// mortonEncode of int32 cells as int64 (the 1-D code is the cell itself).
template <int DIM>
__device__ __forceinline__ int64_t morton(const int32_t *c) {
    if (DIM == 1) return (int64_t)c[0];
    if (DIM == 2) return (int64_t)(spread1((uint64_t)(int64_t)c[0]) | (spread1((uint64_t)(int64_t)c[1]) << 1));
    return (int64_t)(spread2((uint64_t)(int64_t)c[0]) | (spread2((uint64_t)(int64_t)c[1]) << 1) |
                     (spread2((uint64_t)(int64_t)c[2]) << 2));
}
// torch's int32 decode (indexing.py:11-18, 36-43 on an int32 tensor: the
// 64-bit masks truncated to their low 32 bits, >> 32 the sign fill).
__device__ __forceinline__ int32_t compact1_i32(int32_t x) {
    x = x & (int32_t)0x55555555;
    x = (x | (x >> 1)) & (int32_t)0x33333333;
    x = (x | (x >> 2)) & (int32_t)0x0F0F0F0F;
    x = (x | (x >> 4)) & (int32_t)0x00FF00FF;
    x = (x | (x >> 8)) & (int32_t)0x0000FFFF;
    x = (x | (x >> 16)) & (int32_t)0xFFFFFFFF;
    return x;
}
__device__ __forceinline__ int32_t compact2_i32(int32_t w) {
    w = w & (int32_t)0x49249249;
    w = (w ^ (w >> 2)) & (int32_t)0xC30C30C3;
    w = (w ^ (w >> 4)) & (int32_t)0x0F00F00F;
    w = (w ^ (w >> 8)) & (int32_t)0xFF0000FF;
    w = (w ^ (w >> 16)) & (int32_t)0x0000FFFF;
    w = (w ^ (w < 0 ? -1 : 0)) & (int32_t)0x001FFFFF;
    return w;
}
template <int DIM>
__device__ __forceinline__ void decode_i32(int32_t code, int32_t *c) {
    if (DIM == 1) {
        c[0] = code;
    } else if (DIM == 2) {
        c[0] = compact1_i32(code);
        c[1] = compact1_i32(code >> 1);
    } else {
        c[0] = compact2_i32(code);
        c[1] = compact2_i32(code >> 1);
        c[2] = compact2_i32(code >> 2);
    }
}

__device__ __forceinline__ uint64_t lanemask_lt64() {
    const int lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// ------------------------------------------------ stable LSD radix sort ----
// 8-bit digits; per pass: tile histograms (digit-major, so that their
// exclusive scan is every (digit, tile)'s first output slot), the scan, and
// a scatter that ranks keys stably inside the tile with wave ballots.

__global__ void __launch_bounds__(kT) k_rs_hist(const uint64_t *__restrict__ keys, int64_t n, int shift,
                                                uint32_t *__restrict__ hist, int64_t ntiles) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kRsTile;
    for (int k = 0; k < kRsTile / kT; ++k) {
        const int64_t i = base + k * kT + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(kT) k_rs_scatter(const uint64_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                   int64_t n, int shift, const uint32_t *__restrict__ off,
                                                   int64_t ntiles, uint64_t *__restrict__ kout,
                                                   uint32_t *__restrict__ vout) {
    constexpr int kRounds = kRsTile / kT, kWaves = kT / kWave;
    __shared__ uint32_t cnt[kRounds * kWaves][256];
    for (int g = 0; g < kRounds * kWaves; ++g) cnt[g][threadIdx.x] = 0;
    __syncthreads();
    const int lane = __lane_id(), w = threadIdx.x / kWave;
    const int64_t base = (int64_t)blockIdx.x * kRsTile;
    uint64_t key[kRounds];
    uint32_t val[kRounds], dig[kRounds], rank[kRounds];
    bool valid[kRounds];
#pragma unroll
    for (int k = 0; k < kRounds; ++k) {
        const int64_t i = base + k * kT + w * kWave + lane;
        valid[k] = i < n;
        key[k] = valid[k] ? kin[i] : 0ull;
        val[k] = valid[k] ? vin[i] : 0u;
        dig[k] = (uint32_t)((key[k] >> shift) & 255u);
        uint64_t m = __ballot(valid[k]);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (dig[k] >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        rank[k] = (uint32_t)__popcll(m & lanemask_lt64());
        // the first lane of each digit group records the group's size
        if (valid[k] && rank[k] == 0) cnt[k * kWaves + w][dig[k]] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    {   // thread = digit: exclusive prefix over the groups in key order
        uint32_t run = 0;
        for (int g = 0; g < kRounds * kWaves; ++g) {
            const uint32_t c = cnt[g][threadIdx.x];
            cnt[g][threadIdx.x] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kRounds; ++k) {
        if (!valid[k]) continue;
        const uint64_t pos = (uint64_t)off[(int64_t)dig[k] * ntiles + blockIdx.x] + cnt[k * kWaves + w][dig[k]] + rank[k];
        kout[pos] = key[k];
        vout[pos] = val[k];
    }
}

struct RadixBufs {
    uint64_t *k0, *k1;
    uint32_t *v0, *v1;
    uint32_t *hist, *off;
    uint64_t *partials;
};

size_t radix_bytes(int64_t n) {
    const int64_t nt = (n + kRsTile - 1) / kRsTile;
    return (size_t)n * 24 + (size_t)(2 * 256 * nt) * 4 + scan_partials_needed(256 * nt) * 8 + 6 * 256;
}

RadixBufs carve_radix(Carve &cv, int64_t n) {
    const int64_t nt = (n + kRsTile - 1) / kRsTile;
    RadixBufs r;
    r.k0 = cv.take<uint64_t>(n);
    r.k1 = cv.take<uint64_t>(n);
    r.v0 = cv.take<uint32_t>(n);
    r.v1 = cv.take<uint32_t>(n);
    r.hist = cv.take<uint32_t>(256 * nt);
    r.off = cv.take<uint32_t>(256 * nt);
    r.partials = cv.take<uint64_t>(scan_partials_needed(256 * nt));
    return r;
}

// Digit passes of a sort whose keys differ only in the bits of `varying`.
int radix_passes(uint64_t varying) {
    int p = 0;
    for (int shift = 0; shift < 64; shift += 8) p += ((varying >> shift) & 255u) ? 1 : 0;
    return p;
}

// Sorts (k0, v0)[0, n) stably by the key bytes that vary (bits set in
// `varying`; bytes equal in every key need no pass); the result ends in
// (k0, v0) -- after an odd number of passes those are the carved k1, v1.
int radix_sort_u64(RadixBufs &r, int64_t n, uint64_t varying, hipStream_t s) {
    if (n <= 1) return TCR_OK;
    const int64_t nt = (n + kRsTile - 1) / kRsTile;
    for (int shift = 0; shift < 64; shift += 8) {
        if (!((varying >> shift) & 255u)) continue;
        hipLaunchKernelGGL(k_rs_hist, dim3((unsigned)nt), dim3(kT), 0, s, r.k0, n, shift, r.hist, nt);
        const int rc = exclusive_scan<uint32_t, uint32_t>(r.hist, r.off, 256 * nt, false, r.partials, s, nullptr,
                                                          nullptr);
        if (rc) return rc;
        hipLaunchKernelGGL(k_rs_scatter, dim3((unsigned)nt), dim3(kT), 0, s, r.k0, r.v0, n, shift, r.off, nt, r.k1,
                           r.v1);
        std::swap(r.k0, r.k1);
        std::swap(r.v0, r.v1);
    }
    TCR_LAUNCH_CHECK();
    return TCR_OK;
}

uint64_t low_bits(int64_t v) {  // mask of the bits of values in [0, v]
    uint64_t m = 0;
    while (m < (uint64_t)v) m = (m << 1) | 1u;
    return m;
}

// ------------------------------------------------------- search kernels ----

struct MrArgs {
    int32_t dim, mode, levels;
    uint32_t pmask, L;
    float hcell, hfine;
    float mn[3], mx[3];
    int32_t res[kMaxLevels * 3];
    int64_t nq, nr, ncells;
    const float *qpos, *qsup, *rpos, *rsup;
    float *rwrap, *qwrap;           // wrapped copies
    const uint32_t *si;             // sorted slot -> original reference index
    uint32_t *inv;                  // original -> sorted slot
    float *spos, *ssup, *nsup;      // sorted positions / supports, synced supports (sorted order)
    int32_t *sync, *counts, *tot;   // per sorted slot / per reference point
    int32_t *codes;                 // [levels][nr]
    uint32_t *flags, *cellid;       // [levels][nr]
    int32_t *cbegin, *cend, *cindex, *clevel;  // [cells]
    const uint64_t *ckey;           // cells sorted by (level << 32 | code)
    const uint32_t *cval;
    uint32_t *errflag;
};

__global__ void __launch_bounds__(kT) k_mr_minmax(const float *__restrict__ sup, int64_t n, uint32_t *mm) {
    uint32_t lo = 0xFFFFFFFFu, hi = 0u;
    for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
        const uint32_t o = float_to_ordered(sup[i]);
        lo = min(lo, o);
        hi = max(hi, o);
    }
    for (int o = 32; o >= 1; o >>= 1) {
        lo = min(lo, (uint32_t)__shfl_xor((int)lo, o, kWave));
        hi = max(hi, (uint32_t)__shfl_xor((int)hi, o, kWave));
    }
    if (__lane_id() == 0) {
        atomicMin(&mm[0], lo);
        atomicMax(&mm[1], hi);
    }
}

__global__ void __launch_bounds__(kT) k_mr_wrap(const float *__restrict__ in, int64_t n, MrArgs a,
                                                float *__restrict__ out) {
    for (int64_t k = (int64_t)blockIdx.x * kT + threadIdx.x; k < n * a.dim; k += (int64_t)gridDim.x * kT) {
        const int d = (int)(k % a.dim);
        const float p = in[k];
        out[k] = ((a.pmask >> d) & 1u) ? wrap_coord(p, a.mn[d], a.mx[d] - a.mn[d]) : p;
    }
}

// Fine-grid Morton key of every wrapped reference point (signed order as
// unsigned keys), value = the point's index.
template <int DIM>
__global__ void __launch_bounds__(kT) k_mr_keys(MrArgs a, uint64_t *__restrict__ keys, uint32_t *__restrict__ vals,
                                                unsigned long long *__restrict__ keyor) {
    uint64_t m = 0, z = ~0ull;
    for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < a.nr; i += (int64_t)gridDim.x * kT) {
        int32_t c[3] = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < DIM; ++d) c[d] = trunc_x86((a.rwrap[i * DIM + d] - a.mn[d]) / a.hfine);
        const uint64_t k = (uint64_t)morton<DIM>(c) ^ 0x8000000000000000ull;
        keys[i] = k;
        vals[i] = (uint32_t)i;
        m |= k;
        z &= k;
    }
    for (int o = 32; o >= 1; o >>= 1) {
        m |= __shfl_xor(m, o, kWave);
        z &= __shfl_xor(z, o, kWave);
    }
    if (__lane_id() == 0) {  // keyor[0]: OR of the keys, keyor[1]: AND
        atomicOr(&keyor[0], (unsigned long long)m);
        atomicAnd(&keyor[1], (unsigned long long)z);
    }
}

// Sorted positions and supports, the inverse order, and each sorted point's
// level codes with its run-start flags.
template <int DIM>
__global__ void __launch_bounds__(kT) k_mr_gather(MrArgs a) {
    for (int64_t s = (int64_t)blockIdx.x * kT + threadIdx.x; s < a.nr; s += (int64_t)gridDim.x * kT) {
        const uint32_t i = a.si[s];
        a.inv[i] = (uint32_t)s;
        float p[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            p[d] = a.rwrap[(int64_t)i * DIM + d];
            a.spos[s * DIM + d] = p[d];
        }
        a.ssup[s] = a.rsup[i];
        a.nsup[s] = a.rsup[i];
        a.sync[s] = 0;
        int32_t c[3] = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < DIM; ++d) c[d] = trunc_x86((p[d] - a.mn[d]) / a.hcell);
        const int64_t mo = morton<DIM>(c);
        for (int l = 0; l < a.levels; ++l) a.codes[(int64_t)l * a.nr + s] = (int32_t)(uint32_t)(uint64_t)(mo >> (2 * l));
    }
}

__global__ void __launch_bounds__(kT) k_mr_flags(MrArgs a) {
    const int64_t n = (int64_t)a.levels * a.nr;
    for (int64_t k = (int64_t)blockIdx.x * kT + threadIdx.x; k < n; k += (int64_t)gridDim.x * kT) {
        const int64_t s = k % a.nr;
        a.flags[k] = (s == 0 || a.codes[k] != a.codes[k - 1]) ? 1u : 0u;
    }
}

__global__ void __launch_bounds__(kT) k_mr_cells(MrArgs a, uint64_t *__restrict__ ckeys, uint32_t *__restrict__ cvals) {
    const int64_t n = (int64_t)a.levels * a.nr;
    for (int64_t k = (int64_t)blockIdx.x * kT + threadIdx.x; k < n; k += (int64_t)gridDim.x * kT) {
        if (!a.flags[k]) continue;
        const uint32_t c = a.cellid[k];
        const int32_t l = (int32_t)(k / a.nr);
        const int64_t s = k % a.nr;
        a.cbegin[c] = (int32_t)s;
        a.cindex[c] = a.codes[k];
        a.clevel[c] = l;
        // the run ends where the next one of this level starts
        int64_t e = s + 1;
        while (e < a.nr && !a.flags[(int64_t)l * a.nr + e]) ++e;
        a.cend[c] = (int32_t)e;
        ckeys[c] = ((uint64_t)(uint32_t)l << 32) | (uint32_t)a.codes[k];
        cvals[c] = c;
    }
}

__device__ __forceinline__ int level_of(float h, const MrArgs &a) {
    const float lg = (float)log2((double)(h / a.hcell));
    int64_t l = (int64_t)trunc_x86(ceilf(lg)) + 1;
    return (int)(l < 1 ? 1 : (l > a.levels ? a.levels : l));
}

// iterateCellHashed (mlmUtil.h:234-282) for a point at p with support h:
// calls visit(j) for every sorted slot in the reference's order.
template <int DIM, class F>
__device__ __forceinline__ void traverse(const MrArgs &a, const float *p, float h, F visit) {
    const int li = level_of(h, a);
    const float size = a.hcell * (float)(1 << (li - 1));
    int32_t cell[3] = {0, 0, 0}, res[3] = {1, 1, 1};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        cell[d] = trunc_x86(floorf((p[d] - a.mn[d]) / size));
        res[d] = a.res[(li - 1) * 3 + d];
    }
    constexpr int kOff = DIM == 1 ? 3 : DIM == 2 ? 9 : 27;
    for (int o = 0; o < kOff; ++o) {
        int32_t cur[3] = {0, 0, 0};
        int rem = o;
#pragma unroll
        for (int d = DIM - 1; d >= 0; --d) {  // dim 0 slowest
            cur[d] = (cell[d] + (rem % 3 - 1)) % res[d];
            rem /= 3;
        }
        const uint32_t bucket = hash_index<DIM>(cur, a.L);
        const int32_t mort = (int32_t)(uint32_t)(uint64_t)morton<DIM>(cur);
        int32_t dec[3] = {0, 0, 0};
        decode_i32<DIM>(mort, dec);
        if (hash_index<DIM>(dec, a.L) != bucket) continue;  // the code's cells live in another bucket
        const uint64_t key = ((uint64_t)(uint32_t)(li - 1) << 32) | (uint32_t)mort;
        int64_t lo = 0, hi = a.ncells;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (a.ckey[mid] < key) lo = mid + 1; else hi = mid;
        }
        for (int64_t k = lo; k < a.ncells && a.ckey[k] == key; ++k) {
            const uint32_t c = a.cval[k];
            for (int32_t j = a.cbegin[c]; j < a.cend[c]; ++j) visit(j);
        }
    }
}

__device__ __forceinline__ void atomic_max_float(float *addr, float val) {  // the CPU build's CAS loop
    float cur = *addr;
    while (val > cur) {
        const int old = atomicCAS((int *)addr, __float_as_int(cur), __float_as_int(val));
        if (old == __float_as_int(cur)) break;
        cur = __int_as_float(old);
    }
}

// countNeighborsHashmap_impl (countNeighborsHashmap.h:99-124) over the
// reference set at its unwrapped positions.
template <int DIM>
__global__ void __launch_bounds__(kT) k_mr_count(MrArgs a) {
    const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (i >= a.nr) return;
    float p[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int d = 0; d < DIM; ++d) p[d] = a.rpos[i * DIM + d];
    const float h = a.rsup[i];
    int32_t cnt = 0;
    traverse<DIM>(a, p, h, [&](int32_t j) {
        const float dist = exact_distance<DIM>(p, a.spos + (int64_t)j * DIM, a.mn, a.mx, a.pmask);
        if (dist < h) ++cnt;
        if (dist > a.ssup[j] && dist <= h) {
            atomicAdd(&a.sync[j], 1);
            atomic_max_float(&a.nsup[j], h);
        }
    });
    a.counts[i] = cnt;
}

// Row lengths: counts (+ sync counts of the point's sorted slot for
// superSymmetric), wrapper.py:255-270.
__global__ void __launch_bounds__(kT) k_mr_tot(MrArgs a) {
    const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (i >= a.nr) return;
    a.tot[i] = a.mode == TCR_MODE_SUPERSYMMETRIC ? a.counts[i] + a.sync[a.inv[i]] : a.counts[i];
}

__global__ void __launch_bounds__(kT) k_mr_offsets32(const int64_t *__restrict__ off64, int64_t n,
                                                     int32_t *__restrict__ off32) {
    const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (i < n) off32[i] = (int32_t)(uint32_t)(uint64_t)off64[i];
}

// buildNeighborhoodHashmap_impl (buildNeighborhoodHashmap.h:103-128) over
// the wrapped query set; the rows start at the reference-count offsets.
template <int DIM>
__global__ void __launch_bounds__(kT) k_mr_fill(MrArgs a, const int32_t *__restrict__ off32, int64_t total,
                                                int64_t *__restrict__ row, int64_t *__restrict__ col) {
    const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (i >= a.nq) return;
    float p[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int d = 0; d < DIM; ++d) p[d] = a.qwrap[i * DIM + d];
    const float hq = a.qsup[i];
    const bool sym = a.mode == TCR_MODE_SUPERSYMMETRIC;
    const float hs = sym ? a.nsup[a.inv[i]] : hq;
    int64_t o = off32[i];
    const int64_t end = i + 1 < a.nr ? (int64_t)off32[i + 1] : total;
    bool over = false;
    traverse<DIM>(a, p, hs, [&](int32_t j) {
        const float dist = exact_distance<DIM>(p, a.spos + (int64_t)j * DIM, a.mn, a.mx, a.pmask);
        const bool keep = sym ? (dist < hq || dist < a.ssup[j]) : (dist < hq);
        if (!keep) return;
        if (o >= end || o < 0) {
            over = true;
        } else {
            row[o] = i;
            col[o] = j;
        }
        ++o;
    });
    if (over) atomicOr(a.errflag, 1u);
}

__global__ void __launch_bounds__(kT) k_mr_remap(const uint32_t *__restrict__ si, int64_t n, int64_t *__restrict__ col) {
    for (int64_t k = (int64_t)blockIdx.x * kT + threadIdx.x; k < n; k += (int64_t)gridDim.x * kT)
        col[k] = si[col[k]];
}

__global__ void __launch_bounds__(kT) k_zero64(int64_t *__restrict__ p, int64_t n) {
    for (int64_t k = (int64_t)blockIdx.x * kT + threadIdx.x; k < n; k += (int64_t)gridDim.x * kT) p[k] = 0;
}

// Per-search device words: misc[0..15] zero, key OR 0, key AND all ones.
__global__ void k_mr_init(uint32_t *misc, unsigned long long *keyor) {
    if (threadIdx.x < 16) misc[threadIdx.x] = 0u;
    if (threadIdx.x == 0) {
        keyor[0] = 0ull;
        keyor[1] = ~0ull;
    }
}

// misc word k = v (records host-side facts for the fill call).
__global__ void k_mr_set(uint32_t *misc, int k, uint32_t v) { misc[k] = v; }

// ------------------------------------------------------------- host ------

unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(65536, (n + kT - 1) / kT)); }

struct MrWork {
    float *rwrap, *qwrap, *spos, *ssup, *nsup;
    uint32_t *inv, *flags, *cellid;
    int32_t *sync, *counts, *tot, *codes, *cbegin, *cend, *cindex, *clevel;
    int64_t *off64;
    uint64_t *partials;
    unsigned long long *keyor;
    uint32_t *misc;  // [0] error flag, [1] cells, [2] point-sort passes, [3] cell-sort passes
    RadixBufs pts, cells;
};

size_t carve_mr(const tcr_mlmref_grid &g, int64_t nq, int64_t nr, void *ws, MrWork &w) {
    Carve cv(ws);
    const int64_t dim = g.dim, lv = (int64_t)g.levels * nr;
    w.misc = cv.take<uint32_t>(16);
    w.keyor = cv.take<unsigned long long>(2);
    w.rwrap = cv.take<float>(nr * dim);
    w.qwrap = cv.take<float>(nq * dim);
    w.spos = cv.take<float>(nr * dim);
    w.ssup = cv.take<float>(nr);
    w.nsup = cv.take<float>(nr);
    w.inv = cv.take<uint32_t>(nr);
    w.sync = cv.take<int32_t>(nr);
    w.counts = cv.take<int32_t>(nr);
    w.tot = cv.take<int32_t>(nr);
    w.off64 = cv.take<int64_t>(nr + 1);
    w.codes = cv.take<int32_t>(lv);
    w.flags = cv.take<uint32_t>(lv);
    w.cellid = cv.take<uint32_t>(lv + 1);
    w.cbegin = cv.take<int32_t>(lv);
    w.cend = cv.take<int32_t>(lv);
    w.cindex = cv.take<int32_t>(lv);
    w.clevel = cv.take<int32_t>(lv);
    w.partials = cv.take<uint64_t>(scan_partials_needed(std::max<int64_t>(lv, nr) + 1));
    w.pts = carve_radix(cv, nr);
    w.cells = carve_radix(cv, lv);
    return cv.off + 256;
}

bool check_grid(const tcr_mlmref_grid *g) {
    if (!g || g->dim < 1 || g->dim > 3 || g->levels < 1 || g->levels > kMaxLevels) {
        set_error("invalid multi-level grid description");
        return false;
    }
    return true;
}

template <int DIM>
int mr_count_dim(MrArgs &a, MrWork &w, hipStream_t s, int64_t *ncells) {
    hipLaunchKernelGGL(k_mr_init, dim3(1), dim3(64), 0, s, w.misc, w.keyor);
    hipLaunchKernelGGL(k_mr_wrap, dim3(grid_for(a.nr * DIM)), dim3(kT), 0, s, a.rpos, a.nr, a, w.rwrap);
    a.rwrap = w.rwrap;
    hipLaunchKernelGGL(k_mr_keys<DIM>, dim3(grid_for(a.nr)), dim3(kT), 0, s, a, w.pts.k0, w.pts.v0, w.keyor);
    unsigned long long kb[2] = {0ull, 0ull};
    TCR_HIP(hipMemcpyAsync(kb, w.keyor, 16, hipMemcpyDeviceToHost, s));
    TCR_HIP(hipStreamSynchronize(s));
    const uint64_t varying = kb[0] ^ kb[1];
    int rc = radix_sort_u64(w.pts, a.nr, varying, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_mr_set, dim3(1), dim3(1), 0, s, w.misc, 2, (uint32_t)(a.nr > 1 ? radix_passes(varying) : 0));
    a.si = w.pts.v0;
    hipLaunchKernelGGL(k_mr_gather<DIM>, dim3(grid_for(a.nr)), dim3(kT), 0, s, a);
    const int64_t lv = (int64_t)a.levels * a.nr;
    hipLaunchKernelGGL(k_mr_flags, dim3(grid_for(lv)), dim3(kT), 0, s, a);
    rc = exclusive_scan<uint32_t, uint32_t>(w.flags, w.cellid, lv, true, w.partials, s, nullptr, nullptr);
    if (rc) return rc;
    uint32_t nc = 0;
    TCR_HIP(hipMemcpyAsync(&nc, w.cellid + lv, 4, hipMemcpyDeviceToHost, s));
    hipLaunchKernelGGL(k_mr_cells, dim3(grid_for(lv)), dim3(kT), 0, s, a, w.cells.k0, w.cells.v0);
    TCR_HIP(hipStreamSynchronize(s));
    a.ncells = nc;
    *ncells = nc;
    const uint64_t cvary = (low_bits(a.levels - 1) << 32) | 0xFFFFFFFFull;  // level << 32 | code
    rc = radix_sort_u64(w.cells, nc, cvary, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_mr_set, dim3(1), dim3(1), 0, s, w.misc, 3, (uint32_t)(nc > 1 ? radix_passes(cvary) : 0));
    a.ckey = w.cells.k0;
    a.cval = w.cells.v0;
    hipLaunchKernelGGL(k_mr_count<DIM>, dim3(grid_for(a.nr)), dim3(kT), 0, s, a);
    hipLaunchKernelGGL(k_mr_tot, dim3(grid_for(a.nr)), dim3(kT), 0, s, a);
    TCR_LAUNCH_CHECK();
    return TCR_OK;
}

}  // namespace
}  // namespace tcr

using namespace tcr;

extern "C" {

int tcr_mlmref_plan(const float *rsup, int64_t nr, int64_t nq, int32_t dim, const float *domain_min,
                    const float *domain_max, void *scratch, tcr_mlmref_grid *grid, size_t *workspace_bytes,
                    void *stream) {
    hipStream_t s = (hipStream_t)stream;
    if (!grid || !workspace_bytes || !scratch || !domain_min || !domain_max || dim < 1 || dim > 3 || nr < 1 ||
        nq < 0) {
        set_error("tcr_mlmref_plan: invalid arguments");
        return TCR_ERR_ARG;
    }
    if (nr >= ((int64_t)1 << 31) || nq >= ((int64_t)1 << 31)) {
        set_error("tcr_mlmref_plan: more than 2^31 points");
        return TCR_ERR_ARG;
    }
    uint32_t init[2] = {0xFFFFFFFFu, 0u};
    TCR_HIP(hipMemcpyAsync(scratch, init, 8, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_mr_minmax, dim3(grid_for(nr)), dim3(kT), 0, s, rsup, nr, (uint32_t *)scratch);
    TCR_LAUNCH_CHECK();
    uint32_t mm[2];
    TCR_HIP(hipMemcpyAsync(mm, scratch, 8, hipMemcpyDeviceToHost, s));
    TCR_HIP(hipStreamSynchronize(s));
    auto ord_to_f = [](uint32_t u) {
        const uint32_t b = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
        float f;
        std::memcpy(&f, &b, 4);
        return f;
    };
    const float hmin = ord_to_f(mm[0]), hmax = ord_to_f(mm[1]);
    // computeResolutionLevels (cellOp.py:5-46) in fp32 as torch evaluates it
    float ext[3] = {0.f, 0.f, 0.f}, hc = -INFINITY, hf = -INFINITY;
    int32_t q[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d) {
        ext[d] = domain_max[d] - domain_min[d];
        const float qf = ext[d] / hmin;
        const float lg = (float)std::log2((double)qf);
        const float pw = (float)std::ldexp(1.0, (int)std::floor(lg));
        q[d] = (pw >= -2147483648.0f && pw < 2147483648.0f && std::isfinite(lg)) ? (int32_t)pw : INT32_MIN;
        const float hd = ext[d] / (float)q[d];
        hc = (d == 0 || hd > hc) ? hd : hc;
        const float fd = ext[d] / 65536.0f;
        hf = (d == 0 || fd > hf) ? fd : hf;
    }
    const double ratio = (double)hmax / (double)hc;
    const int levels_hint = (int)std::ceil(std::log2(ratio));
    int L = 0;
    std::memset(grid, 0, sizeof(*grid));
    for (int i = 0; i <= levels_hint && L < kMaxLevels; ++i) {
        float mnq = INFINITY;
        int32_t r[3] = {1, 1, 1};
        for (int d = 0; d < dim; ++d) {
            const float v = (float)q[d] / (float)(1 << i);
            mnq = std::min(mnq, v);
            r[d] = (int32_t)v;
        }
        if (!(mnq > 1.0f)) break;
        for (int d = 0; d < 3; ++d) grid->res[L][d] = d < dim ? r[d] : 1;
        ++L;
    }
    if (L == 0 || !std::isfinite(hc) || !(hc > 0.0f)) {
        set_error("multi-level search: no resolution level for these supports and this domain (the reference "
                  "fails to merge an empty cell list)");
        return TCR_ERR_DOMAIN;
    }
    grid->dim = dim;
    grid->levels = L;
    grid->hcell = hc;
    grid->hfine = hf;
    grid->hmin = hmin;
    grid->hmax = hmax;
    MrWork w;
    *workspace_bytes = carve_mr(*grid, nq, nr, nullptr, w);
    return TCR_OK;
}

static int mr_setup(const tcr_mlmref_grid *grid, const float *qpos, const float *qsup, int64_t nq,
                    const float *rpos, const float *rsup, int64_t nr, const float *domain_min,
                    const float *domain_max, uint32_t periodic_mask, int32_t mode, int64_t L, void *ws,
                    size_t wsb, MrArgs &a, MrWork &w) {
    if (!check_grid(grid)) return TCR_ERR_ARG;
    if (!qpos || !qsup || !rpos || !rsup || !domain_min || !domain_max || nr < 1 || nq < 0 || L < 1 ||
        L > 0xFFFFFFFFll || mode < 0 || mode > 3 || (periodic_mask >> grid->dim)) {
        set_error("multi-level search: invalid arguments");
        return TCR_ERR_ARG;
    }
    const size_t need = carve_mr(*grid, nq, nr, ws, w);
    if (!ws || wsb < need) {
        set_error("workspace too small: need " + std::to_string(need) + " bytes");
        return TCR_ERR_SPACE;
    }
    std::memset(&a, 0, sizeof(a));
    a.dim = grid->dim;
    a.mode = mode;
    a.levels = grid->levels;
    a.pmask = periodic_mask;
    a.L = (uint32_t)L;
    a.hcell = grid->hcell;
    a.hfine = grid->hfine;
    for (int d = 0; d < 3; ++d) {
        a.mn[d] = d < grid->dim ? domain_min[d] : 0.f;
        a.mx[d] = d < grid->dim ? domain_max[d] : 0.f;
    }
    for (int l = 0; l < grid->levels; ++l)
        for (int d = 0; d < 3; ++d) a.res[l * 3 + d] = grid->res[l][d];
    a.nq = nq;
    a.nr = nr;
    a.qpos = qpos;
    a.qsup = qsup;
    a.rpos = rpos;
    a.rsup = rsup;
    a.rwrap = w.rwrap;
    a.qwrap = w.qwrap;
    a.si = w.pts.v0;
    a.inv = w.inv;
    a.spos = w.spos;
    a.ssup = w.ssup;
    a.nsup = w.nsup;
    a.sync = w.sync;
    a.counts = w.counts;
    a.tot = w.tot;
    a.codes = w.codes;
    a.flags = w.flags;
    a.cellid = w.cellid;
    a.cbegin = w.cbegin;
    a.cend = w.cend;
    a.cindex = w.cindex;
    a.clevel = w.clevel;
    a.errflag = w.misc;
    return TCR_OK;
}

int tcr_mlmref_count(const tcr_mlmref_grid *grid, const float *qpos, const float *qsup, int64_t nq,
                     const float *rpos, const float *rsup, int64_t nr, const float *domain_min,
                     const float *domain_max, uint32_t periodic_mask, int32_t mode, int64_t hash_map_length,
                     void *workspace, size_t workspace_bytes, int32_t *counts, int32_t *offsets, int64_t *total,
                     void *stream) {
    hipStream_t s = (hipStream_t)stream;
    MrArgs a;
    MrWork w;
    int rc = mr_setup(grid, qpos, qsup, nq, rpos, rsup, nr, domain_min, domain_max, periodic_mask, mode,
                      hash_map_length, workspace, workspace_bytes, a, w);
    if (rc) return rc;
    if (mode == TCR_MODE_SCATTER) {
        set_error("multi-level search: scatter runs as the transposed gather search");
        return TCR_ERR_ARG;
    }
    if (!counts || !offsets || !total) {
        set_error("counts, offsets and total are required");
        return TCR_ERR_ARG;
    }
    int64_t nc = 0;
    switch (grid->dim) {
        case 1: rc = mr_count_dim<1>(a, w, s, &nc); break;
        case 2: rc = mr_count_dim<2>(a, w, s, &nc); break;
        default: rc = mr_count_dim<3>(a, w, s, &nc); break;
    }
    if (rc) return rc;
    rc = exclusive_scan<int32_t, int64_t>(w.tot, w.off64, nr, true, w.partials, s, nullptr, nullptr);
    if (rc) return rc;
    hipLaunchKernelGGL(k_mr_offsets32, dim3(grid_for(nr)), dim3(kT), 0, s, w.off64, nr, offsets);
    TCR_HIP(hipMemcpyAsync(counts, w.tot, sizeof(int32_t) * nr, hipMemcpyDeviceToDevice, s));
    int64_t t = 0;
    TCR_HIP(hipMemcpyAsync(&t, w.off64 + nr, 8, hipMemcpyDeviceToHost, s));
    TCR_LAUNCH_CHECK();
    TCR_HIP(hipStreamSynchronize(s));
    *total = t;
    return TCR_OK;
}

int tcr_mlmref_fill(const tcr_mlmref_grid *grid, const float *qpos, const float *qsup, int64_t nq,
                    const float *rpos, const float *rsup, int64_t nr, const float *domain_min,
                    const float *domain_max, uint32_t periodic_mask, int32_t mode, int64_t hash_map_length,
                    void *workspace, size_t workspace_bytes, const int32_t *offsets, int64_t total, int64_t *row,
                    int64_t *col, void *stream) {
    hipStream_t s = (hipStream_t)stream;
    MrArgs a;
    MrWork w;
    int rc = mr_setup(grid, qpos, qsup, nq, rpos, rsup, nr, domain_min, domain_max, periodic_mask, mode,
                      hash_map_length, workspace, workspace_bytes, a, w);
    if (rc) return rc;
    if (nq > nr) {
        set_error("multi-level search: more query points than reference points (the reference reads its "
                  "offsets past their end)");
        return TCR_ERR_UNSUPPORTED;
    }
    if (total < 0 || (total > 0 && (!row || !col || !offsets))) {
        set_error("rows, columns and offsets are required");
        return TCR_ERR_ARG;
    }
    // the structure of tcr_mlmref_count is still in the workspace; the
    // cells' sorted keys are where its sort left them
    uint32_t nc = 0, passes[2] = {0u, 0u};
    TCR_HIP(hipMemcpyAsync(&nc, w.cellid + (int64_t)grid->levels * nr, 4, hipMemcpyDeviceToHost, s));
    TCR_HIP(hipMemcpyAsync(passes, w.misc + 2, 8, hipMemcpyDeviceToHost, s));
    TCR_HIP(hipStreamSynchronize(s));
    a.ncells = nc;
    // the sorts of tcr_mlmref_count left their results in the carved k1/v1
    // after an odd number of passes
    a.si = (passes[0] & 1) ? w.pts.v1 : w.pts.v0;
    a.ckey = (passes[1] & 1) ? w.cells.k1 : w.cells.k0;
    a.cval = (passes[1] & 1) ? w.cells.v1 : w.cells.v0;
    hipLaunchKernelGGL(k_mr_set, dim3(1), dim3(1), 0, s, w.misc, 0, 0u);
    hipLaunchKernelGGL(k_mr_wrap, dim3(grid_for(nq * grid->dim)), dim3(kT), 0, s, qpos, nq, a, w.qwrap);
    if (total > 0) {
        hipLaunchKernelGGL(k_zero64, dim3(grid_for(total)), dim3(kT), 0, s, row, total);
        hipLaunchKernelGGL(k_zero64, dim3(grid_for(total)), dim3(kT), 0, s, col, total);
        switch (grid->dim) {
            case 1: hipLaunchKernelGGL(k_mr_fill<1>, dim3(grid_for(nq)), dim3(kT), 0, s, a, offsets, total, row, col); break;
            case 2: hipLaunchKernelGGL(k_mr_fill<2>, dim3(grid_for(nq)), dim3(kT), 0, s, a, offsets, total, row, col); break;
            default: hipLaunchKernelGGL(k_mr_fill<3>, dim3(grid_for(nq)), dim3(kT), 0, s, a, offsets, total, row, col); break;
        }
        hipLaunchKernelGGL(k_mr_remap, dim3(grid_for(total)), dim3(kT), 0, s, a.si, total, col);
    }
    uint32_t err = 0;
    TCR_HIP(hipMemcpyAsync(&err, w.misc, 4, hipMemcpyDeviceToHost, s));
    TCR_LAUNCH_CHECK();
    TCR_HIP(hipStreamSynchronize(s));
    if (err) {
        set_error("multi-level search: a query finds more pairs than its row holds (the reference writes past "
                  "its rows; its output is undefined)");
        return TCR_ERR_UNSUPPORTED;
    }
    return TCR_OK;
}

size_t tcr_mlmref_transpose_workspace_size(int64_t n, int64_t n_rows) {
    if (n < 0 || n_rows < 0) return 0;
    Carve cv(nullptr);
    carve_radix(cv, n);
    cv.take<int32_t>(n_rows);
    cv.take<int64_t>(n_rows + 1);
    cv.take<uint64_t>(scan_partials_needed(n_rows + 1));
    return cv.off + 256;
}

}  // extern "C"

namespace tcr {
namespace {
__global__ void __launch_bounds__(kT) k_tr_keys(const int64_t *__restrict__ col, int64_t n, uint64_t *__restrict__ keys,
                                                uint32_t *__restrict__ vals, int32_t *__restrict__ cnt) {
    for (int64_t k = (int64_t)blockIdx.x * kT + threadIdx.x; k < n; k += (int64_t)gridDim.x * kT) {
        keys[k] = (uint64_t)col[k];
        vals[k] = (uint32_t)k;
        if (cnt) atomicAdd(&cnt[col[k]], 1);
    }
}
__global__ void __launch_bounds__(kT) k_tr_out(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ vals,
                                               const int64_t *__restrict__ row, int64_t n, int64_t *__restrict__ out_row,
                                               int64_t *__restrict__ out_col) {
    for (int64_t k = (int64_t)blockIdx.x * kT + threadIdx.x; k < n; k += (int64_t)gridDim.x * kT) {
        out_row[k] = (int64_t)keys[k];
        out_col[k] = row[vals[k]];
    }
}
__global__ void __launch_bounds__(kT) k_zero32(int32_t *__restrict__ p, int64_t n) {
    for (int64_t k = (int64_t)blockIdx.x * kT + threadIdx.x; k < n; k += (int64_t)gridDim.x * kT) p[k] = 0;
}
}  // namespace
}  // namespace tcr

extern "C" {

// transposeCOO(sorted=True) of the gather search (wrapper.py:119-142,
// 191-192): (row, col) -> (col, row) ordered by the new row, stably; with
// indptr / row_entries, the CSR arrays of coo_to_csr (py/util.py:335-358).
int tcr_mlmref_transpose(const int64_t *row, const int64_t *col, int64_t n, int64_t n_rows, void *workspace,
                         size_t workspace_bytes, int64_t *out_row, int64_t *out_col, int32_t *indptr,
                         int64_t *row_entries, void *stream) {
    hipStream_t s = (hipStream_t)stream;
    if (n < 0 || n_rows < 0 || (n > 0 && (!row || !col || !out_row || !out_col))) {
        set_error("tcr_mlmref_transpose: invalid arguments");
        return TCR_ERR_ARG;
    }
    Carve cv(workspace);
    RadixBufs r = carve_radix(cv, n);
    int32_t *cnt = cv.take<int32_t>(n_rows);
    int64_t *off = cv.take<int64_t>(n_rows + 1);
    uint64_t *partials = cv.take<uint64_t>(scan_partials_needed(n_rows + 1));
    if (!workspace || workspace_bytes < cv.off + 256) {
        set_error("workspace too small: need " + std::to_string(cv.off + 256) + " bytes");
        return TCR_ERR_SPACE;
    }
    const bool want_csr = indptr || row_entries;
    if (want_csr) hipLaunchKernelGGL(k_zero32, dim3(grid_for(n_rows)), dim3(kT), 0, s, cnt, n_rows);
    if (n > 0) {
        hipLaunchKernelGGL(k_tr_keys, dim3(grid_for(n)), dim3(kT), 0, s, col, n, r.k0, r.v0, want_csr ? cnt : nullptr);
        int rc = radix_sort_u64(r, n, low_bits(n_rows > 0 ? n_rows - 1 : 0), s);
        if (rc) return rc;
        hipLaunchKernelGGL(k_tr_out, dim3(grid_for(n)), dim3(kT), 0, s, r.k0, r.v0, row, n, out_row, out_col);
    }
    if (want_csr) {
        int rc = exclusive_scan<int32_t, int64_t>(cnt, off, n_rows, true, partials, s, nullptr, nullptr);
        if (rc) return rc;
        rc = csr_arrays_launch(off, n_rows, indptr, row_entries, s);
        if (rc) return rc;
    }
    TCR_LAUNCH_CHECK();
    return TCR_OK;
}

}  // extern "C"
