// Brute-force search with the semantics of the reference's default
// algorithm='naive' (searchAlgorithms/radiusNaive.py:8-56, called from
// neighborhood.py:226-229,268-271): every (query a, reference b) pair whose
// distance is <= the mode's support, ordered by reference index, then query
// index (the row-major flattening of radiusNaive's (ny, nx) mask, :17,36-37).
//
// Arithmetic, in the order the reference's fp32 torch ops take it:
//   * periodic coordinates wrapped by torch.remainder `wraps` times (the
//     caller's getPeriodicPointCloud, util.py:375-387, and radiusNaive's own
//     wrap, :12-13), into a workspace copy first;
//   * per dimension the fp32 difference x_a - y_b (:15), in periodic
//     dimensions through radiusNaive's `mod` (:4-6): where |x| > (max-min)/2,
//     sgn(x) * ((|x| + min) % (max - min) + min), with (max - min) and its half
//     formed in double (Python floats) and rounded to fp32 at the tensor op;
//   * squares summed in dimension order, sqrtf (correctly rounded here), then
//     `<=` against gather hx[a], scatter hy[b], symmetric (hx[a] + hy[b]) / 2,
//     superSymmetric torch.max(hx, hy)[a] = max(hx[a], hy[a]) (NaN-propagating);
//   * radiusNaiveFixed (:41-56): the squared distance <= h*h, h*h formed in
//     double and rounded to fp32.
//
// One wavefront per reference point walks the queries 64 at a time: the
// count pass stores the pairs per reference point, an exclusive scan gives
// each reference point's first slot, and the fill pass repeats the tests and
// writes the pairs at the ballot's lane ranks.  O(nx * ny) tests: a
// cross-check and small-cloud path, as in the reference.
//
// The same file holds algorithm='small', the brute-force op the reference
// keeps commented out (cpp/neighborhoodSmall_cpu.cpp:45-302, whose Python
// binding is None, so the reference raises there): for each query i, every
// reference j in index order with modDistance (the compact path's exact
// formula) < the mode's support -- superSymmetric std::max(h_i, h_j) -- and
// the fixed variant modDistance2 < h*h (h*h in fp32); positions as the
// caller passes them (radiusSearch has wrapped them once).  One wavefront per
// query walks the references.
#include <algorithm>
#include <cmath>

#include "tcr_internal.h"

namespace tcr {
namespace {

constexpr int kNaiveThreads = 256;
constexpr int64_t kNaiveMaxBlocks = 16384;

struct NaiveDomain {
    float lo[3];     // domain min
    float up[3];     // domain max
    float wext[3];   // max - min in fp32 (the tensor subtraction of the wrap)
    float mext[3];   // (max - min) formed in double, rounded to fp32 (mod's divisor)
    float mhalf[3];  // (max - min) / 2 formed in double, rounded to fp32 (mod's threshold)
    uint32_t pmask;
};

struct NaiveArgs {
    const float *xw, *yw;  // wrapped positions [n * dim] (workspace)
    const float *hx, *hy;
    int64_t nx, ny;
    NaiveDomain g;
    int32_t mode;
    int32_t fixed;         // radiusNaiveFixed: squared distance <= h2
    float h2;
    int32_t *ref_counts;   // count pass: pairs per reference point (small: per query)
    const int64_t *offsets;  // fill pass: first slot per reference point
    void *ii, *jj;         // fill pass: query / reference indices
    int32_t idx32;
};

__device__ __forceinline__ float torch_remainder(float a, float b) {
    float r = fmodf(a, b);
    if (r != 0.0f && ((r < 0.0f) != (b < 0.0f))) r += b;
    return r;
}

// radiusNaive.py:4-6
__device__ __forceinline__ float naive_mod(float x, float lo, float mext, float mhalf) {
    const float ax = fabsf(x);
    if (!(ax > mhalf)) return x;
    const float r = torch_remainder(ax + lo, mext) + lo;
    const float s = x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f);
    return s * r;
}

// torch.maximum: NaN if either operand is NaN
__device__ __forceinline__ float torch_max(float a, float b) {
    if (a != a || b != b) return __int_as_float(0x7fc00000);
    return a > b ? a : b;
}

__global__ void __launch_bounds__(kNaiveThreads) k_naive_wrap(const float *__restrict__ p, float *__restrict__ out,
                                                              int64_t n, int32_t dim, NaiveDomain g, int32_t wraps) {
    for (int64_t k = (int64_t)blockIdx.x * kNaiveThreads + threadIdx.x; k < n * dim;
         k += (int64_t)gridDim.x * kNaiveThreads) {
        const int d = (int)(k % dim);
        float v = p[k];
        if ((g.pmask >> d) & 1u)
            for (int w = 0; w < wraps; ++w) v = wrap_coord(v, g.lo[d], g.wext[d]);
        out[k] = v;
    }
}

template <int DIM, bool FILL>
__global__ void __launch_bounds__(kNaiveThreads) k_naive(NaiveArgs a) {
    const int lane = __lane_id();
    const int64_t w0 = ((int64_t)blockIdx.x * kNaiveThreads + threadIdx.x) / kWave;
    const int64_t nw = (int64_t)gridDim.x * (kNaiveThreads / kWave);
    for (int64_t b = w0; b < a.ny; b += nw) {  // one reference point per wavefront
        float yb[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int d = 0; d < DIM; ++d) yb[d] = a.yw[b * DIM + d];
        const bool use_hyb = !a.fixed && (a.mode == TCR_MODE_SCATTER || a.mode == TCR_MODE_SYMMETRIC);
        const float hyb = use_hyb ? a.hy[b] : 0.0f;
        int64_t pos = FILL ? a.offsets[b] : 0;
        int32_t cnt = 0;
        for (int64_t a0 = 0; a0 < a.nx; a0 += kWave) {
            const int64_t i = a0 + lane;
            bool hit = false;
            if (i < a.nx) {
                float s = 0.0f;
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    float diff = a.xw[i * DIM + d] - yb[d];
                    if ((a.g.pmask >> d) & 1u) diff = naive_mod(diff, a.g.lo[d], a.g.mext[d], a.g.mhalf[d]);
                    const float sq = diff * diff;
                    s = d == 0 ? sq : s + sq;
                }
                if (a.fixed) {
                    hit = s <= a.h2;
                } else {
                    float thr;
                    switch (a.mode) {
                        case TCR_MODE_GATHER: thr = a.hx[i]; break;
                        case TCR_MODE_SCATTER: thr = hyb; break;
                        case TCR_MODE_SYMMETRIC: thr = (a.hx[i] + hyb) / 2.0f; break;
                        default: thr = torch_max(a.hx[i], a.hy[i]); break;
                    }
                    hit = sqrtf(s) <= thr;
                }
            }
            const uint64_t m = __ballot(hit);
            if (FILL) {
                if (hit) {
                    const int64_t k = pos + __popcll(m & ((1ull << lane) - 1ull));
                    if (a.idx32) {
                        ((int32_t *)a.ii)[k] = (int32_t)i;
                        ((int32_t *)a.jj)[k] = (int32_t)b;
                    } else {
                        ((int64_t *)a.ii)[k] = i;
                        ((int64_t *)a.jj)[k] = b;
                    }
                }
                pos += __popcll(m);
            } else {
                cnt += __popcll(m);
            }
        }
        if (!FILL && lane == 0) a.ref_counts[b] = cnt;
    }
}

template <int DIM, bool FILL>
__global__ void __launch_bounds__(kNaiveThreads) k_small(NaiveArgs a) {
    const int lane = __lane_id();
    const int64_t w0 = ((int64_t)blockIdx.x * kNaiveThreads + threadIdx.x) / kWave;
    const int64_t nw = (int64_t)gridDim.x * (kNaiveThreads / kWave);
    const float mn[3] = {a.g.lo[0], a.g.lo[1], a.g.lo[2]};
    const float mx[3] = {a.g.up[0], a.g.up[1], a.g.up[2]};
    for (int64_t i = w0; i < a.nx; i += nw) {  // one query per wavefront
        float xi[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int d = 0; d < DIM; ++d) xi[d] = a.xw[i * DIM + d];
        const bool use_hi = !a.fixed && a.mode != TCR_MODE_SCATTER;
        const float hi = use_hi ? a.hx[i] : 0.0f;
        int64_t pos = FILL ? a.offsets[i] : 0;
        int32_t cnt = 0;
        for (int64_t j0 = 0; j0 < a.ny; j0 += kWave) {
            const int64_t j = j0 + lane;
            bool hit = false;
            if (j < a.ny) {
                float xj[3] = {0.f, 0.f, 0.f};
#pragma unroll
                for (int d = 0; d < DIM; ++d) xj[d] = a.yw[j * DIM + d];
                if (a.fixed) {
                    hit = exact_distance2<DIM>(xi, xj, mn, mx, a.g.pmask) < a.h2;
                } else {
                    const float hj = a.mode != TCR_MODE_GATHER ? a.hy[j] : 0.0f;
                    hit = support_pred(exact_distance<DIM>(xi, xj, mn, mx, a.g.pmask), hi, hj, a.mode);
                }
            }
            const uint64_t m = __ballot(hit);
            if (FILL) {
                if (hit) {
                    const int64_t k = pos + __popcll(m & ((1ull << lane) - 1ull));
                    if (a.idx32) {
                        ((int32_t *)a.ii)[k] = (int32_t)i;
                        ((int32_t *)a.jj)[k] = (int32_t)j;
                    } else {
                        ((int64_t *)a.ii)[k] = i;
                        ((int64_t *)a.jj)[k] = j;
                    }
                }
                pos += __popcll(m);
            } else {
                cnt += __popcll(m);
            }
        }
        if (!FILL && lane == 0) a.ref_counts[i] = cnt;
    }
}

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

struct NaiveWork {
    int32_t *ref_counts;
    uint64_t *partials;
    float *xw, *yw;
    size_t bytes;
};

NaiveWork carve_naive(void *ws, int64_t nx, int64_t ny, int32_t dim) {
    NaiveWork w{};
    char *p = (char *)ws;
    size_t o = 0;
    const int64_t nc = std::max(nx, ny);  // counts per reference point (naive) or query (small)
    w.ref_counts = (int32_t *)(p ? p + o : nullptr);
    o += align256(sizeof(int32_t) * (size_t)std::max<int64_t>(nc, 1));
    w.partials = (uint64_t *)(p ? p + o : nullptr);
    o += align256(sizeof(uint64_t) * scan_partials_needed(nc));
    w.xw = (float *)(p ? p + o : nullptr);
    o += align256(sizeof(float) * (size_t)std::max<int64_t>(nx * dim, 1));
    w.yw = (float *)(p ? p + o : nullptr);
    o += align256(sizeof(float) * (size_t)std::max<int64_t>(ny * dim, 1));
    w.bytes = o;
    return w;
}

// Validates the call and fills the kernel arguments; wraps the positions
// into the workspace.
int naive_setup(bool small, const float *x, const float *hx, int64_t nx, const float *y, const float *hy,
                int64_t ny, int32_t dim, const float *domain_min, const float *domain_max, uint32_t periodic_mask,
                int32_t wraps, int32_t mode, const double *fixed_support, void *workspace, size_t workspace_bytes,
                hipStream_t s, NaiveArgs &a, NaiveWork &w) {
    if (dim < 1 || dim > 3) { set_error("Unsupported dimension: " + std::to_string(dim)); return TCR_ERR_ARG; }
    if (mode < 0 || mode > 3) { set_error("mode must be one of gather, scatter, symmetric, or superSymmetric"); return TCR_ERR_ARG; }
    if (nx < 0 || ny < 0 || wraps < 0 || wraps > 2) { set_error("invalid sizes or wrap count"); return TCR_ERR_ARG; }
    if ((nx > 0 && !x) || (ny > 0 && !y) || !domain_min || !domain_max) {
        set_error("positions and domain are required");
        return TCR_ERR_ARG;
    }
    if (!fixed_support) {
        const bool need_hx = mode != TCR_MODE_SCATTER, need_hy = mode != TCR_MODE_GATHER;
        if ((need_hx && nx > 0 && !hx) || (need_hy && ny > 0 && !hy) ||
            (!small && mode == TCR_MODE_SUPERSYMMETRIC && nx > 0 && !hy)) {
            set_error("supports required by this mode are not defined");
            return TCR_ERR_ARG;
        }
        if (!small && mode == TCR_MODE_SUPERSYMMETRIC && nx != ny) {
            set_error("superSymmetric needs as many query as reference points (torch.max(hx, hy))");
            return TCR_ERR_ARG;
        }
    }
    w = carve_naive(workspace, nx, ny, dim);
    if (!workspace || workspace_bytes < w.bytes) {
        set_error("workspace too small: need " + std::to_string(w.bytes) + " bytes");
        return TCR_ERR_SPACE;
    }
    a = NaiveArgs{};
    for (int d = 0; d < 3; ++d) {
        const float mn = d < dim ? domain_min[d] : 0.0f, mx = d < dim ? domain_max[d] : 0.0f;
        a.g.lo[d] = mn;
        a.g.up[d] = mx;
        a.g.wext[d] = mx - mn;
        a.g.mext[d] = (float)((double)mx - (double)mn);
        a.g.mhalf[d] = (float)(((double)mx - (double)mn) / 2.0);
    }
    a.g.pmask = periodic_mask & ((1u << dim) - 1u);
    a.hx = hx;
    a.hy = hy;
    a.nx = nx;
    a.ny = ny;
    a.mode = mode;
    a.fixed = fixed_support ? 1 : 0;
    // radiusNaiveFixed: h**2 of a Python float; the small op: float support * support
    const float hf = fixed_support ? (float)*fixed_support : 0.0f;
    a.h2 = !fixed_support ? 0.0f : small ? hf * hf : (float)(*fixed_support * *fixed_support);
    a.xw = w.xw;
    a.yw = w.yw;
    a.ref_counts = w.ref_counts;
    const int64_t nmax = std::max(nx, ny) * dim;
    const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>(4096, (nmax + kNaiveThreads - 1) / kNaiveThreads));
    if (nx > 0)
        hipLaunchKernelGGL(k_naive_wrap, dim3(blocks), dim3(kNaiveThreads), 0, s, x, w.xw, nx, dim, a.g, wraps);
    if (ny > 0)
        hipLaunchKernelGGL(k_naive_wrap, dim3(blocks), dim3(kNaiveThreads), 0, s, y, w.yw, ny, dim, a.g, wraps);
    TCR_LAUNCH_CHECK();
    return TCR_OK;
}

// naive: one wavefront per reference point; small: one per query
template <bool FILL>
int naive_launch(bool small, int32_t dim, const NaiveArgs &a, hipStream_t s) {
    const int64_t outer = small ? a.nx : a.ny;
    if (outer == 0) return TCR_OK;
    constexpr int64_t wpb = kNaiveThreads / kWave;
    const dim3 grid((unsigned)std::min<int64_t>(kNaiveMaxBlocks, (outer + wpb - 1) / wpb));
    if (small) {
        switch (dim) {
            case 1: hipLaunchKernelGGL((k_small<1, FILL>), grid, dim3(kNaiveThreads), 0, s, a); break;
            case 2: hipLaunchKernelGGL((k_small<2, FILL>), grid, dim3(kNaiveThreads), 0, s, a); break;
            default: hipLaunchKernelGGL((k_small<3, FILL>), grid, dim3(kNaiveThreads), 0, s, a); break;
        }
    } else {
        switch (dim) {
            case 1: hipLaunchKernelGGL((k_naive<1, FILL>), grid, dim3(kNaiveThreads), 0, s, a); break;
            case 2: hipLaunchKernelGGL((k_naive<2, FILL>), grid, dim3(kNaiveThreads), 0, s, a); break;
            default: hipLaunchKernelGGL((k_naive<3, FILL>), grid, dim3(kNaiveThreads), 0, s, a); break;
        }
    }
    TCR_LAUNCH_CHECK();
    return TCR_OK;
}

int brute_count(bool small, const float *x, const float *hx, int64_t nx, const float *y, const float *hy, int64_t ny,
                int32_t dim, const float *domain_min, const float *domain_max, uint32_t periodic_mask, int32_t wraps,
                int32_t mode, const double *fixed_support, void *workspace, size_t workspace_bytes,
                int64_t *offsets, hipStream_t s) {
    if (!offsets) { set_error("offsets is required"); return TCR_ERR_ARG; }
    NaiveArgs a;
    NaiveWork w;
    int rc = naive_setup(small, x, hx, nx, y, hy, ny, dim, domain_min, domain_max, periodic_mask, wraps, mode,
                         fixed_support, workspace, workspace_bytes, s, a, w);
    if (rc) return rc;
    rc = naive_launch<false>(small, dim, a, s);
    if (rc) return rc;
    return exclusive_scan<int32_t, int64_t>(w.ref_counts, offsets, small ? nx : ny, true, w.partials, s);
}

int brute_fill(bool small, const float *x, const float *hx, int64_t nx, const float *y, const float *hy, int64_t ny,
               int32_t dim, const float *domain_min, const float *domain_max, uint32_t periodic_mask, int32_t wraps,
               int32_t mode, const double *fixed_support, void *workspace, size_t workspace_bytes,
               const int64_t *offsets, void *ii, void *jj, int32_t index_bytes, hipStream_t s) {
    if (!offsets || (index_bytes != 4 && index_bytes != 8)) {
        set_error("offsets and index_bytes of 4 or 8 are required");
        return TCR_ERR_ARG;
    }
    NaiveArgs a;
    NaiveWork w;
    int rc = naive_setup(small, x, hx, nx, y, hy, ny, dim, domain_min, domain_max, periodic_mask, wraps, mode,
                         fixed_support, workspace, workspace_bytes, s, a, w);
    if (rc) return rc;
    a.offsets = offsets;
    a.ii = ii;
    a.jj = jj;
    a.idx32 = index_bytes == 4;
    return naive_launch<true>(small, dim, a, s);
}

}  // namespace
}  // namespace tcr

extern "C" {

size_t tcr_naive_workspace_size(int64_t nx, int64_t ny, int32_t dim) {
    if (nx < 0 || ny < 0 || dim < 1 || dim > 3) return 0;
    return tcr::carve_naive(nullptr, nx, ny, dim).bytes;
}

int tcr_naive_count(const float *x, const float *hx, int64_t nx, const float *y, const float *hy, int64_t ny,
                    int32_t dim, const float *domain_min, const float *domain_max, uint32_t periodic_mask,
                    int32_t wraps, int32_t mode, const double *fixed_support, void *workspace,
                    size_t workspace_bytes, int64_t *offsets, void *stream) {
    return tcr::brute_count(false, x, hx, nx, y, hy, ny, dim, domain_min, domain_max, periodic_mask, wraps, mode,
                            fixed_support, workspace, workspace_bytes, offsets, (hipStream_t)stream);
}

int tcr_naive_fill(const float *x, const float *hx, int64_t nx, const float *y, const float *hy, int64_t ny,
                   int32_t dim, const float *domain_min, const float *domain_max, uint32_t periodic_mask,
                   int32_t wraps, int32_t mode, const double *fixed_support, void *workspace,
                   size_t workspace_bytes, const int64_t *offsets, void *ii, void *jj, int32_t index_bytes,
                   void *stream) {
    return tcr::brute_fill(false, x, hx, nx, y, hy, ny, dim, domain_min, domain_max, periodic_mask, wraps, mode,
                           fixed_support, workspace, workspace_bytes, offsets, ii, jj, index_bytes,
                           (hipStream_t)stream);
}

int tcr_small_count(const float *x, const float *hx, int64_t nx, const float *y, const float *hy, int64_t ny,
                    int32_t dim, const float *domain_min, const float *domain_max, uint32_t periodic_mask,
                    int32_t wraps, int32_t mode, const double *fixed_support, void *workspace,
                    size_t workspace_bytes, int64_t *offsets, void *stream) {
    return tcr::brute_count(true, x, hx, nx, y, hy, ny, dim, domain_min, domain_max, periodic_mask, wraps, mode,
                            fixed_support, workspace, workspace_bytes, offsets, (hipStream_t)stream);
}

int tcr_small_fill(const float *x, const float *hx, int64_t nx, const float *y, const float *hy, int64_t ny,
                   int32_t dim, const float *domain_min, const float *domain_max, uint32_t periodic_mask,
                   int32_t wraps, int32_t mode, const double *fixed_support, void *workspace,
                   size_t workspace_bytes, const int64_t *offsets, void *ii, void *jj, int32_t index_bytes,
                   void *stream) {
    return tcr::brute_fill(true, x, hx, nx, y, hy, ny, dim, domain_min, domain_max, periodic_mask, wraps, mode,
                           fixed_support, workspace, workspace_bytes, offsets, ii, jj, index_bytes,
                           (hipStream_t)stream);
}

}  // extern "C"
