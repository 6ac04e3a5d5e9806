// Device-side building blocks of the compact-hashing radius search for gfx950.
//
// Arithmetic contract: every function marked "exact" reproduces the reference
// CPU path bit for bit (compile with -ffp-contract=off; HIP's default
// correctly rounded fp32 '/' and sqrtf; fp64 '/' is IEEE on gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/tcr.h"

namespace tcr {

constexpr int kWave = 64;

// Per-cloud descriptor in device memory: the caller's tcr_cloud plus what the
// device computes during the build.
struct CloudDev {
    tcr_cloud h;              // host-provided
    const float *qp, *qs;     // this cloud's query positions / supports
    const float *rp, *rs;     // this cloud's reference positions / supports
    // sort tiles
    int64_t rt_off;           // first radix tile of this cloud (reference set)
    int64_t rt_n;             // number of tiles
    int64_t qt_off, qt_n;     // query tiles (non-self-search grouping)
    int64_t hb_off;           // first hash bucket of this cloud in the batch
    // computed on device
    uint32_t hmax_ord;        // ordered-uint encoding of hMax (atomicMax target)
    float hcell;
    int32_t nc[3];
    uint32_t err;             // nonzero: degenerate grid
    uint32_t fast;            // nonzero: wavefront-per-cell path eligible
    uint32_t irregular;       // nonzero: some reference point's cell is outside the grid
    uint32_t pad0;
    int64_t lt_off, lt_cap;   // lin -> cell table slots of this cloud (cap: power of two)
    int64_t cell_off, ncells; // cells of this cloud in the batch cell arrays
    int64_t qcell_off, nqcells; // query groups (non-self) in the batch
    // filter geometry of the wavefront-per-cell kernels (grid_of): the
    // periodic extents, the stencil box half-width, the coordinate bound of
    // the error margin and its centring term 80 hCell^2
    float f_ext[3];
    float f_box, f_geo, f_c80;
    float f_pad[2];
};

// ------------------------------------------------------------------ utils --

__device__ __forceinline__ uint32_t float_to_ordered(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ordered_to_float(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// Index of the segment containing i in an ascending prefix array
// starts[0..n) (starts[k] <= i), picking the last such k.
template <typename T, typename F>
__device__ __forceinline__ int find_segment(int n, T i, F starts) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (starts(mid) <= i) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// ------------------------------------------------- exact reference pieces --

// torch.remainder(a, b) for fp32 (fmod, then shift into the sign of b); the
// wrap p' = remainder(p - min, max - min) + min (py/util.py:387,
// py/compactHashing/cppSearch.py:221).  exact.
__device__ __forceinline__ float wrap_coord(float p, float lo, float ext) {
    const float a = p - lo;
    // fmod(a, b) == a exactly whenever |a| < |b| (NaN and b == 0 fail the test)
    float r = fabsf(a) < fabsf(ext) ? a : fmodf(a, ext);
    if (r != 0.0f && ((r < 0.0f) != (ext < 0.0f))) r += ext;
    return r + lo;
}

// (int32) floorf((x - lo) / hcell) with the x86 conversion of the reference
// CPU build for values outside int32 (and NaN): INT32_MIN.  exact.
__device__ __forceinline__ int32_t cell_coord(float x, float lo, float hcell) {
    const float v = floorf((x - lo) / hcell);
    return (v >= -2147483648.0f && v < 2147483648.0f) ? (int32_t)v : INT32_MIN;
}

// Row-major linear index with int32 wrap (cpp/neighborhood.h:13-23).
template <int DIM>
__device__ __forceinline__ int32_t linear_index(const int32_t *c, const int32_t *nc) {
    uint32_t lin = 0, prod = 1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        lin += (uint32_t)c[d] * prod;
        prod *= (uint32_t)nc[d];
    }
    return (int32_t)lin;
}

// hashIndexing (cpp/hashing.h:20-35).
template <int DIM>
__device__ __forceinline__ uint32_t hash_index(const int32_t *c, uint32_t L) {
    if (DIM == 1) return ((uint32_t)c[0]) % L;
    const uint32_t primes[3] = {73856093u, 19349663u, 83492791u};
    uint32_t h = 0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) h += ((uint32_t)c[d]) * primes[d];
    return h % L;
}

__device__ __forceinline__ int32_t pymod(int32_t n, int32_t m) {  // cpp/common.h:95-97
    return n >= 0 ? n % m : ((n % m) + m) % m;
}

// modDistance2 (cpp/common.h:130-138) for float positions: fp32 difference,
// fp64 periodic modulo, fp64 square accumulated into an fp32 sum.  exact.
template <int DIM>
__device__ __forceinline__ float exact_distance2(const float *xi, const float *xj,
                                                 const float *mn, const float *mx, uint32_t pmask) {
    float sum = 0.0f;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const float pq = xi[d] - xj[d];
        double diff;
        if ((pmask >> d) & 1u) {
            const double h = (double)(mx[d] - mn[d]);
            const double t = (double)pq + h / 2.0;
            diff = (t - floor(t / h) * h) - h / 2.0;
        } else {
            diff = (double)pq;
        }
        sum = (float)((double)sum + diff * diff);
    }
    return sum;
}

// modDistance (cpp/common.h:106-129): sqrtf of modDistance2.  exact.
template <int DIM>
__device__ __forceinline__ float exact_distance(const float *xi, const float *xj,
                                                const float *mn, const float *mx, uint32_t pmask) {
    return sqrtf(exact_distance2<DIM>(xi, xj, mn, mx, pmask));
}

// std::max(a, b) as the reference calls it: (a < b) ? b : a.
__device__ __forceinline__ float max_ref(float a, float b) { return (a < b) ? b : a; }

// Support predicate (cpp/neighborhoodDynamic.h:63-70,105-108).  exact.
__device__ __forceinline__ bool support_pred(float dist, float hi, float hj, int mode) {
    switch (mode) {
        case TCR_MODE_SCATTER: return dist < hj;
        case TCR_MODE_GATHER: return dist < hi;
        case TCR_MODE_SYMMETRIC: return dist < (hi + hj) / 2.f;
        default: return dist < max_ref(hi, hj);
    }
}

}  // namespace tcr
