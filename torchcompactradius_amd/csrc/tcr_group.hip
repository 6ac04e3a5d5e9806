// Wavefront-per-cell count and fill passes for self-search clouds (the query
// set is the reference set), searchRange 1.
//
// 1. k_stencil (one thread per occupied cell) resolves the 3^d neighbour
//    cells of every occupied cell through the cell index, with the lookup
//    semantics of queryHashMap (cpp/neighborhood.h:35-58): periodic
//    coordinates wrapped, non-periodic ones unclamped, a class matching only
//    its linear index and bucket.  It also packs a 32-byte record per cell
//    (points, cloud, state, centre).
// 2. k_count_small (one wavefront per cell with <= kSmallChunk candidates):
//    the lanes lay the candidates out in the reference's order -- stencil
//    offset order (dim 0 slowest), then sorted position -- over rounds of 64
//    and loop over the cell's query points with the query held uniform.  Per
//    (query, round) the 64 tests give one hit mask; the pass stores the
//    masks and the per-query counts.
// 3. k_fill_small replays the masks: each query's hits are packed in LDS at
//    the lane prefix count of the mask, so a row comes out in exactly the
//    reference order (cpp/neighborhoodDynamic.h:76-113), and leave in
//    coalesced stores -- no second distance evaluation.
// 4. Cells with more candidates go through k_group_big (count and fill, in
//    chunks, recomputing the tests in the fill).
// Both small-cell kernels walk their cells software-pipelined: the next
// cell's record and stencil ranges are in flight while the current cell
// computes.
//
// Predicate: a filtered-exact test.  Coordinates are centred on the cell and
// candidates moved to the periodic image nearest to it; the squared distance
// is estimated as |c|^2 - 2 q.c + |q|^2 with a rigorous error margin that
// covers the centring and image shift, the expansion's cancellation, the
// reference's own fp32 difference and fp64/fp32 accumulation, and the sqrtf
// rounding.  Pairs inside the margin band, and candidates outside the 3^d
// stencil box (e.g. aliased out-of-domain points), take the exact fp64
// modDistance + sqrtf + '<' path of the reference (cpp/common.h:106-129).
// Queries whose once-wrapped position differs from their twice-wrapped
// sorted copy, or whose cell lies outside [0, numCells), are "strays" and go
// to the per-query kernel, as do all points of a cell whose first point is
// outside the grid.
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

#include "tcr_generic.h"

namespace tcr {

constexpr int kGroupThreads = 256;           // 4 wavefronts per workgroup
constexpr int kSmallRounds = 7;              // candidate rounds of a small cell
constexpr int kSmallChunk = kSmallRounds * kWave;  // 448
constexpr int kBigRounds = 8;
constexpr int kBigChunk = kBigRounds * kWave;      // 512 candidates per chunk
constexpr int kStencil = 27;                 // stride of the per-cell stencil table
constexpr float kU = 0x1p-24f;               // fp32 unit roundoff

enum : uint32_t { kCellSkip = 0, kCellFast = 1, kCellStray = 2 };

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
// Read-only data through the constant address space: a wave-uniform address
// becomes a scalar load, which waits on the scalar counter instead of the
// vector-memory one that also covers the kernel's earlier stores.
template <typename T>
using cptr = const __attribute__((address_space(4))) T *;
template <typename T>
__device__ __forceinline__ cptr<T> as_const(const T *p) {
    return (cptr<T>)p;
}

// Lane m's value, m uniform (v_readlane).
__device__ __forceinline__ int32_t bcast(int32_t v, uint32_t m) { return __builtin_amdgcn_readlane(v, m); }
__device__ __forceinline__ uint32_t bcast(uint32_t v, uint32_t m) {
    return (uint32_t)__builtin_amdgcn_readlane((int32_t)v, m);
}
__device__ __forceinline__ float bcast(float v, uint32_t m) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), m));
}
__device__ __forceinline__ uint64_t bcast(uint64_t v, uint32_t m) {
    const uint32_t lo = bcast((uint32_t)v, m), hi = bcast((uint32_t)(v >> 32), m);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ int64_t bcast(int64_t v, uint32_t m) { return (int64_t)bcast((uint64_t)v, m); }
// The backend's writelane intrinsic (no clang builtin for it here); the
// backend sets up the lane select (m0) itself.
extern "C" __device__ int tcr_writelane_i32(int val, int lane, int old) __asm("llvm.amdgcn.writelane.i32");
// v with lane m replaced by the uniform value x (v_writelane_b32: one VALU
// operation, where a compare-and-select costs two moves and a select).
__device__ __forceinline__ uint32_t put_lane(uint32_t v, uint32_t x, uint32_t m) {
    return (uint32_t)tcr_writelane_i32((int)x, (int)m, (int)v);
}
// Position of this lane among the set lanes of uniform mask `bits`, plus base
// (v_mbcnt_lo/hi).
__device__ __forceinline__ int32_t lane_rank(uint64_t bits, int32_t base) {
    return (int32_t)__builtin_amdgcn_mbcnt_hi((uint32_t)(bits >> 32),
                                              __builtin_amdgcn_mbcnt_lo((uint32_t)bits, (uint32_t)base));
}

// Decision thresholds on the estimated squared distance for "dist < h":
// in if est < lo, out if est >= hi, exact test in between.
struct Thr {
    float lo, hi;
};
__device__ __forceinline__ Thr thresholds(float h, float geo, float c80, int dim) {
    Thr t;
    if (!(h > 0.0f)) {          // h <= 0 or NaN: no distance is below it
        t.lo = -INFINITY;
        t.hi = -INFINITY;
    } else if (isinf(h)) {      // every finite distance is below +inf
        t.lo = INFINITY;
        t.hi = INFINITY;
    } else {
        const float h2 = h * h;
        if (!(h2 > 0x1p-100f)) {  // tiny supports: decide every pair exactly
            t.lo = -INFINITY;
            t.hi = INFINITY;
        } else {
            // twice the bound: 80u cell^2 (centring + expansion), 4u h^2
            // (both accumulations, sqrtf, h*h), 2 dim h u geo (fp32
            // differences and image shifts of the reference and the estimate)
            const float m = 2.0f * kU * (c80 + 4.0f * h2 + 2.0f * (float)dim * h * geo);
            t.lo = h2 - m;
            t.hi = h2 + m;
        }
    }
    return t;
}

// Per-cell record written by k_stencil.
struct __align__(16) CellRec {
    uint32_t gb, ge;    // sorted slots [gb, ge) of the cell's points
    uint32_t cloud;
    uint32_t state;     // kCellSkip / kCellFast / kCellStray
    float4 center;      // w: 1 if a stencil neighbour lies across a periodic boundary, else 0
};

struct GroupArgs {
    const CloudDev *desc;
    const int64_t *r_start;
    const int64_t *q_start;
    const uint32_t *misc;        // device counters (kMisc* slots)
    const uint32_t *cell_begin;  // [cells+1] first global sorted slot
    const uint32_t *cell_cloud;
    const uint32_t *cell_hash;   // [cells] bucket of each cell
    const uint64_t *lintab;      // lin -> cell tables of all clouds
    const float4 *sorted;
    const int32_t *sort_index;   // local original index per sorted slot
    const uint8_t *sflags;       // per sorted slot: nonzero = stray query
    // the member queries of a cell: the sorted reference points for a
    // self-search, the sorted queries of a query-cell cloud
    const float4 *msorted;
    const int32_t *msort_index;
    const uint8_t *msflags;
    int2 *ranges;                // [cells * 27] (first global slot, count)
    CellRec *recs;               // [cells]
    uint64_t *masks;             // [kSmallRounds][nr]: round-major, so a batch's loads and stores coalesce
    int64_t mask_stride;         // nr (reference points of the batch)
    int32_t *counts;             // [nq] per global query index
    const int64_t *offsets;      // [nq+1] (fill)
    const int64_t *total;        // offsets + nq: the fill runs only if *total <= capacity
    int64_t capacity;
    PairOut out;                 // fill outputs
    int32_t *cursor;             // [nq] per-query write cursor across chunks (fill)
    int32_t *stray_list;         // [nq]
    uint32_t *stray_count;
    uint32_t *big_list;          // [cells]
    uint32_t *big_count;
    uint32_t cell_run;           // consecutive cells a wavefront takes in turn (tuned)
    const uint32_t *ncells;      // the cell count of this pass (misc[kMiscCells] or misc[kMiscQCells])
    // query cells (k_qstencil)
    const uint32_t *qcell_begin;
    const int32_t *qcell_lin;
    const uint32_t *qcell_cloud;
};

// ------------------------------------------------------------ stencil ----

__device__ __forceinline__ bool group_owned(const CloudDev &cd) {
    return cd.h.self_search && cd.h.search_range == 1;
}

template <int DIM>
__device__ __forceinline__ bool cell_coords(const CloudDev &cd, float4 p0, int32_t *q) {
    const float v[3] = {p0.x, p0.y, p0.z};
    bool inrange = true;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        q[d] = cell_coord(v[d], cd.h.domain_min[d], cd.hcell);
        inrange = inrange && q[d] >= 0 && q[d] < cd.nc[d];
    }
    return inrange;
}

// All 3^DIM stencil cells of one cell: the periodic neighbours of an in-grid
// coordinate are q - 1 + n or q + 1 - n (a fast cloud has >= 5 cells in every
// periodic dimension).  In a cloud whose points all lie in the grid, a
// neighbour whose coordinates all lie in the grid is its class's own cell,
// whose bucket is hash(c); only coordinates outside the grid (non-periodic
// borders, whose linear index may alias another cell), and every lookup of a
// cloud with points outside the grid (whose classes may start with such a
// point), need the bucket check of the reference lookup.
template <int DIM>
__device__ __forceinline__ void stencil_cell(const GroupArgs &a, const CloudDev &cd, const int32_t *q,
                                             int2 *__restrict__ out) {
    constexpr int kOff = DIM == 1 ? 3 : DIM == 2 ? 9 : 27;
    int32_t cc[3][3];
    bool in[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            cc[d][k] = 0;
            in[d][k] = true;
        }
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const int32_t n = cd.nc[d];
        const bool per = (cd.h.periodic_mask >> d) & 1u;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            int32_t u = q[d] + k - 1;
            if (per) u = u < 0 ? u + n : (u >= n ? u - n : u);
            cc[d][k] = u;
            in[d][k] = u >= 0 && u < n;
        }
    }
    const CellIndex idx = cell_index(cd, a.lintab);
    const uint32_t L = cd.h.hash_map_length;
#pragma unroll
    for (int o = 0; o < kOff; ++o) {
        // offset o: dim 0 slowest, as the reference's offset list
        int32_t c[3] = {0, 0, 0};
        bool all_in = true;
        int rem = o;
#pragma unroll
        for (int d = DIM - 1; d >= 0; --d) {
            const int k = rem % 3;
            rem /= 3;
            c[d] = cc[d][k];
            all_in = all_in && in[d][k];
        }
        const int32_t cid = idx.find(linear_index<DIM>(c, cd.nc));
        bool ok = cid >= 0;
        if (ok && (!all_in || cd.irregular)) ok = a.cell_hash[cid] == hash_index<DIM>(c, L);
        int2 r = make_int2(0, 0);
        if (ok) {
            const uint32_t b = a.cell_begin[cid];
            r = make_int2((int32_t)b, (int32_t)(a.cell_begin[cid + 1] - b));
        }
        out[o] = r;
    }
#pragma unroll
    for (int o = kOff; o < kStencil; ++o) out[o] = make_int2(0, 0);
}

// stencil_cell for a cloud with a dense cell index (every reference point in
// the grid): all lookups of the stencil are issued before any is used --
// 3^DIM index loads, then the cell ranges, then the bucket checks of the
// coordinates outside the grid -- instead of one chain of dependent loads
// per stencil cell.
template <int DIM>
__device__ __forceinline__ void stencil_cell_dense(const GroupArgs &a, const CloudDev &cd, const int32_t *q,
                                                   int2 *__restrict__ out) {
    constexpr int kOff = DIM == 1 ? 3 : DIM == 2 ? 9 : 27;
    int32_t cc[3][3];
    bool in[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            cc[d][k] = 0;
            in[d][k] = true;
        }
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const int32_t n = cd.nc[d];
        const bool per = (cd.h.periodic_mask >> d) & 1u;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            int32_t u = q[d] + k - 1;
            if (per) u = u < 0 ? u + n : (u >= n ? u - n : u);
            cc[d][k] = u;
            in[d][k] = u >= 0 && u < n;
        }
    }
    const int32_t *tab = (const int32_t *)(a.lintab + cd.lt_off);
    const uint32_t P = (uint32_t)grid_cells(cd);  // < 2^31 for a dense index
    int32_t cid[kOff];
#pragma unroll
    for (int o = 0; o < kOff; ++o) {
        int32_t c[3] = {0, 0, 0};
        int rem = o;
#pragma unroll
        for (int d = DIM - 1; d >= 0; --d) {
            c[d] = cc[d][rem % 3];
            rem /= 3;
        }
        const uint32_t lin = (uint32_t)linear_index<DIM>(c, cd.nc);
        const int32_t v = tab[lin < P ? lin : 0u];  // in-bounds load, its value used only when lin < P
        cid[o] = lin < P ? v : -1;
    }
    uint32_t beg[kOff], end[kOff];
#pragma unroll
    for (int o = 0; o < kOff; ++o) {
        const uint32_t k = cid[o] >= 0 ? (uint32_t)cid[o] : 0u;
        beg[o] = a.cell_begin[k];
        end[o] = a.cell_begin[k + 1];
    }
    const uint32_t L = cd.h.hash_map_length;
#pragma unroll
    for (int o = 0; o < kOff; ++o) {
        int32_t c[3] = {0, 0, 0};
        bool all_in = true;
        int rem = o;
#pragma unroll
        for (int d = DIM - 1; d >= 0; --d) {
            const int k = rem % 3;
            rem /= 3;
            c[d] = cc[d][k];
            all_in = all_in && in[d][k];
        }
        bool ok = cid[o] >= 0;
        if (ok && !all_in) ok = a.cell_hash[cid[o]] == hash_index<DIM>(c, L);
        out[o] = ok ? make_int2((int32_t)beg[o], (int32_t)(end[o] - beg[o])) : make_int2(0, 0);
    }
#pragma unroll
    for (int o = kOff; o < kStencil; ++o) out[o] = make_int2(0, 0);
}

// One thread per occupied cell: its record and its resolved stencil.
__global__ void __launch_bounds__(256) k_stencil(GroupArgs a) {
    const uint32_t ncells = a.misc[kMiscCells];
    for (int64_t cid = (int64_t)blockIdx.x * 256 + threadIdx.x; cid < ncells; cid += (int64_t)gridDim.x * 256) {
        const uint32_t c = a.cell_cloud[cid];
        const CloudDev &cd = a.desc[c];
        const int dim = cd.h.dim;
        const uint32_t gb = a.cell_begin[cid], ge = a.cell_begin[cid + 1];
        int2 *out = a.ranges + cid * kStencil;
        if (!group_owned(cd)) {
            a.recs[cid] = CellRec{gb, ge, c, kCellSkip, make_float4(0, 0, 0, 0)};
            continue;
        }
        const float4 p0 = a.sorted[gb];
        int32_t q[3] = {0, 0, 0};
        const bool inrange = dim == 1 ? cell_coords<1>(cd, p0, q) : dim == 2 ? cell_coords<2>(cd, p0, q)
                                                                             : cell_coords<3>(cd, p0, q);
        const bool fast = inrange && cd.fast;
        float cen[3] = {0.f, 0.f, 0.f};
        bool wrap = false;  // a periodic neighbour coordinate wraps around
        for (int d = 0; d < dim; ++d) {
            cen[d] = cd.h.domain_min[d] + ((float)q[d] + 0.5f) * cd.hcell;
            if (((cd.h.periodic_mask >> d) & 1u) && (q[d] == 0 || q[d] == cd.nc[d] - 1)) wrap = true;
        }
        a.recs[cid] = CellRec{gb, ge, c, fast ? kCellFast : kCellStray,
                              make_float4(cen[0], cen[1], cen[2], wrap ? 1.0f : 0.0f)};
        if (!fast) {
            for (int o = 0; o < kStencil; ++o) out[o] = make_int2(0, 0);
            continue;
        }
        if (cell_index_dense(cd) && cd.h.nr > 0) {
            switch (dim) {
                case 1: stencil_cell_dense<1>(a, cd, q, out); break;
                case 2: stencil_cell_dense<2>(a, cd, q, out); break;
                default: stencil_cell_dense<3>(a, cd, q, out); break;
            }
        } else {
            switch (dim) {
                case 1: stencil_cell<1>(a, cd, q, out); break;
                case 2: stencil_cell<2>(a, cd, q, out); break;
                default: stencil_cell<3>(a, cd, q, out); break;
            }
        }
    }
}

// One thread per query cell of the query-cell clouds: its record (query
// slots, cloud, state, centre) and its resolved stencil in the reference
// structure.  The cell's coordinates come from its key (the linear index of
// an in-grid cell); its queries outside the grid carry a stray flag and go
// to the per-query path one by one.  A cloud the cell kernels do not take
// (fewer than 5 cells in a periodic dimension, a degenerate grid) is left to
// the per-query kernel whole.
__global__ void __launch_bounds__(256) k_qstencil(GroupArgs a) {
    const uint32_t ncells = *a.ncells;
    for (int64_t cid = (int64_t)blockIdx.x * 256 + threadIdx.x; cid < ncells; cid += (int64_t)gridDim.x * 256) {
        const uint32_t c = a.qcell_cloud[cid];
        const CloudDev &cd = a.desc[c];
        const int dim = cd.h.dim;
        // a cloud's last cell ends with its queries (the next cloud's may
        // not be sorted into cells)
        const uint32_t gb = a.qcell_begin[cid];
        const uint32_t ge = (uint32_t)min((int64_t)a.qcell_begin[cid + 1], a.q_start[c + 1]);
        int2 *out = a.ranges + cid * kStencil;
        if (!cd.fast) {
            a.recs[cid] = CellRec{gb, ge, c, kCellSkip, make_float4(0, 0, 0, 0)};
            continue;
        }
        int32_t q[3] = {0, 0, 0};
        uint32_t lin = (uint32_t)a.qcell_lin[cid];
        float cen[3] = {0.f, 0.f, 0.f};
        bool wrap = false;
        for (int d = 0; d < dim; ++d) {
            q[d] = (int32_t)(lin % (uint32_t)cd.nc[d]);
            lin /= (uint32_t)cd.nc[d];
            cen[d] = cd.h.domain_min[d] + ((float)q[d] + 0.5f) * cd.hcell;
            if (((cd.h.periodic_mask >> d) & 1u) && (q[d] == 0 || q[d] == cd.nc[d] - 1)) wrap = true;
        }
        a.recs[cid] = CellRec{gb, ge, c, kCellFast, make_float4(cen[0], cen[1], cen[2], wrap ? 1.0f : 0.0f)};
        if (cell_index_dense(cd) && cd.h.nr > 0) {
            switch (dim) {
                case 1: stencil_cell_dense<1>(a, cd, q, out); break;
                case 2: stencil_cell_dense<2>(a, cd, q, out); break;
                default: stencil_cell_dense<3>(a, cd, q, out); break;
            }
        } else {
            switch (dim) {
                case 1: stencil_cell<1>(a, cd, q, out); break;
                case 2: stencil_cell<2>(a, cd, q, out); break;
                default: stencil_cell<3>(a, cd, q, out); break;
            }
        }
    }
}

// ------------------------------------------------------- per-cell setup ---

// One cell as a wavefront sees it: the record, and lane o holding stencil
// cell o's first global slot and exclusive candidate prefix.
struct CellView {
    uint32_t c, gb, ge, state;
    bool wrap;           // some stencil candidates are periodic images
    float4 center;
    int64_t qs;          // first global query index of the cloud
    int32_t rbeg, pre;   // per lane (stencil cell `lane`): first global slot, candidate prefix
    int32_t cnt;         // per lane: points of stencil cell `lane` (0 past the stencil)
    int32_t total;       // candidates
};

struct CellLoad {
    CellRec rec;
    int2 rng;
};

__device__ __forceinline__ CellLoad load_cell(const GroupArgs &a, uint32_t cid) {
    CellLoad l;
    const cptr<uint32_t> rw = as_const((const uint32_t *)(a.recs + uniform(cid)));
    l.rec.gb = rw[0];
    l.rec.ge = rw[1];
    l.rec.cloud = rw[2];
    l.rec.state = rw[3];
    l.rec.center = make_float4(__uint_as_float(rw[4]), __uint_as_float(rw[5]), __uint_as_float(rw[6]),
                               __uint_as_float(rw[7]));
    l.rng = __lane_id() < kStencil ? a.ranges[(int64_t)cid * kStencil + __lane_id()] : make_int2(0, 0);
    return l;
}

// Inclusive prefix sum over lanes 0..31 (the stencil cells) with DPP row
// shifts and the row-15 broadcast: no lane-address registers, which the
// shuffle form keeps live across the whole cell loop.
__device__ __forceinline__ int32_t scan_lanes32(int32_t x) {
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true);   // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true);   // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true);   // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true);   // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1 and 3
    return x;
}

__device__ __forceinline__ CellView make_view(const GroupArgs &a, const CellLoad &l, int noff) {
    const int lane = __lane_id();
    CellView v;
    v.gb = uniform(l.rec.gb);
    v.ge = uniform(l.rec.ge);
    v.c = uniform(l.rec.cloud);
    v.state = uniform(l.rec.state);
    v.center = make_float4(uniform_f(l.rec.center.x), uniform_f(l.rec.center.y), uniform_f(l.rec.center.z), 0.f);
    v.wrap = uniform_f(l.rec.center.w) != 0.0f;
    v.qs = as_const(a.q_start)[v.c];
    const int32_t cnt = lane < noff ? l.rng.y : 0;
    const int32_t incl = scan_lanes32(cnt);
    v.rbeg = l.rng.x;
    v.pre = incl - cnt;
    v.cnt = cnt;
    v.total = bcast(incl, (uint32_t)(noff - 1));
    return v;
}

// Global sorted slot of candidate k (k < total), by binary search over the
// stencil prefix (used where no LDS slot table is built).
__device__ __forceinline__ int32_t candidate_slot(const CellView &v, int32_t k, int noff) {
    int lo = 0;
#pragma unroll
    for (int step = 16; step > 0; step >>= 1) {
        const int mid = lo + step;
        const int32_t pm = __shfl(v.pre, mid < noff ? mid : noff - 1, kWave);
        if (mid < noff && pm <= k) lo = mid;
    }
    return __shfl(v.rbeg, lo, kWave) + (k - __shfl(v.pre, lo, kWave));
}

// Per-wavefront LDS table of a small cell: slot_off[k] = first global slot of
// the stencil cell holding candidate k minus its candidate prefix, so that
// candidate k lives at global slot k + slot_off[k].  Lane o writes the
// entries of stencil cell o.
__device__ __forceinline__ void expand_slots(const CellView &v, int32_t *slot_off, int noff) {
    const int32_t cnt = v.cnt;
    const int32_t d = v.rbeg - v.pre;
    for (int32_t t = 0; __any(t < cnt); ++t)
        if (t < cnt) slot_off[v.pre + t] = d;
    __asm__ volatile("" ::: "memory");
}

__device__ __forceinline__ void append_strays(const GroupArgs &a, const CellView &v) {
    for (uint32_t s = v.gb + __lane_id(); s < v.ge; s += kWave) {
        const uint32_t k = atomicAdd(a.stray_count, 1u);
        a.stray_list[k] = (int32_t)(v.qs + a.msort_index[s]);
    }
}

// The members of a fast cell whose stray bit is set (a query whose
// once-wrapped position differs from its twice-wrapped sorted copy) go to
// the per-query list; the cell kernels skip them.
__device__ __forceinline__ void append_flagged(const GroupArgs &a, const CellView &v) {
    for (uint32_t s = v.gb + __lane_id(); s < v.ge; s += kWave) {
        if (a.msflags[s]) {
            const uint32_t k = atomicAdd(a.stray_count, 1u);
            a.stray_list[k] = (int32_t)(v.qs + a.msort_index[s]);
        }
    }
}

// Geometry of the filter for one cell.
template <int DIM>
struct Geo {
    float ext[3], center[3];
    float box, geo, c80;  // c80 = 80 hCell^2, the centring term of the margin
    uint32_t pmask;
    __device__ __forceinline__ Geo(const CloudDev &cd, float4 cen) { init(&cd, cen); }
    __device__ __forceinline__ Geo(cptr<CloudDev> cd, float4 cen) { init(cd, cen); }
    // the per-cloud parts come from the descriptor (computed by grid_of)
    template <class P>
    __device__ __forceinline__ void init(P cd, float4 cen) {
        pmask = cd->h.periodic_mask;
        const float cv[3] = {cen.x, cen.y, cen.z};
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            ext[d] = cd->f_ext[d];
            center[d] = cv[d];
        }
        box = cd->f_box;
        geo = cd->f_geo;
        c80 = cd->f_c80;
    }
    // Candidate at its image nearest to the cell, centred, with |c|^2 in p:
    // +inf for a lane without a candidate, NaN for a candidate outside the
    // stencil box (then the pair is decided exactly).
    __device__ __forceinline__ void place(float4 v, bool valid, float *xc, float &p) const {
        const float xv[3] = {v.x, v.y, v.z};
        bool well = true;
        p = 0.0f;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            float x = xv[d];
            if ((pmask >> d) & 1u) {
                const float e0 = x - center[d];
                if (e0 > 0.5f * ext[d]) x = x - ext[d];
                else if (e0 < -0.5f * ext[d]) x = x + ext[d];
            }
            const float c = x - center[d];
            well = well && fabsf(c) <= box;
            xc[d] = c;
            p = d == 0 ? c * c : fmaf(c, c, p);
        }
        if (!valid) p = INFINITY;
        else if (!well) p = __int_as_float(0x7fc00000);
    }
    // place() for a cell none of whose stencil neighbours lies across a
    // periodic boundary: a candidate of such a stencil is within 1.5 cells of
    // the centre and a periodic extent spans > 4 cells, so no image shift
    // applies (an aliased candidate from elsewhere fails the box test as in
    // place() and is decided exactly)
    __device__ __forceinline__ void place_direct(float4 v, bool valid, float *xc, float &p) const {
        const float xv[3] = {v.x, v.y, v.z};
        bool well = true;
        p = 0.0f;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const float c = xv[d] - center[d];
            well = well && fabsf(c) <= box;
            xc[d] = c;
            p = d == 0 ? c * c : fmaf(c, c, p);
        }
        if (!valid) p = INFINITY;
        else if (!well) p = __int_as_float(0x7fc00000);
    }
};

// The reference's exact test for one pair; the domain is re-read from the
// descriptor so that nothing of it stays live in the hot loops (rare path).
template <int DIM, int MODE>
__device__ __attribute__((noinline)) bool exact_pair(const GroupArgs &a, const CloudDev *cd, uint32_t qs_slot,
                                                     int32_t j) {
    const float4 xq = a.msorted[qs_slot];
    const float xi[3] = {xq.x, xq.y, xq.z};
    const float4 v = a.sorted[j];
    const float xj[3] = {v.x, v.y, v.z};
    float mn[3] = {0.f, 0.f, 0.f}, mx[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        mn[d] = cd->h.domain_min[d];
        mx[d] = cd->h.domain_max[d];
    }
    const float dist = exact_distance<DIM>(xi, xj, mn, mx, cd->h.periodic_mask);
    return support_pred(dist, xq.w, v.w, MODE);
}

// One round of tests of query q against the 64 candidates in lanes:
// returns the hit mask.  In: est < lo; exact band: not in and not
// (est >= hi) -- NaN candidates land in the band, +inf lanes never hit.
template <int DIM, int MODE, class SlotOf>
__device__ __forceinline__ uint64_t round_hits(const GroupArgs &a, const CloudDev &cd, const Geo<DIM> &g,
                                               const float *cx, float cp, float ch, const Thr &tjr,
                                               const float *xi2, float pi, float hi, const Thr &tq,
                                               uint32_t q_slot, SlotOf slot_of) {
    float t = cp;
#pragma unroll
    for (int d = 0; d < DIM; ++d) t = fmaf(cx[d], xi2[d], t);
    Thr th;
    if (MODE == TCR_MODE_GATHER) {
        th = tq;
    } else {
        if (MODE == TCR_MODE_SCATTER) th = tjr;
        else if (MODE == TCR_MODE_SYMMETRIC) th = thresholds((hi + ch) / 2.f, g.geo, g.c80, DIM);
        else th = thresholds(max_ref(hi, ch), g.geo, g.c80, DIM);
        t += pi;
    }
    uint64_t hitm = __ballot(t < th.lo);
    const uint64_t band = __ballot(!(t >= th.hi)) & ~hitm;
    if (band) {
        // the candidate's slot is only needed here; evaluated with every
        // lane active (the lookup may read other lanes)
        const int32_t jslot = slot_of();
        bool ok = false;
        if ((band >> __lane_id()) & 1ull) ok = exact_pair<DIM, MODE>(a, &cd, q_slot, jslot);
        hitm |= __ballot(ok) & band;
    }
    return hitm;
}

// ------------------------------------------------------------ small cells --

// Tests the member queries in `todo` (lane m = query mb + m) against the
// cell's NR candidate rounds; lane m receives query m's hit masks.  NR is a
// template argument so that the rounds are unrolled without a per-round
// bound check.
template <int DIM, int MODE, int NR>
__device__ __forceinline__ void test_members(const GroupArgs &a, const CloudDev &cd, const Geo<DIM> &g,
                                             const float (&cx)[kSmallRounds][3], const float (&cp)[kSmallRounds],
                                             const float (&chh)[kSmallRounds], const Thr (&tj)[kSmallRounds],
                                             const float (&mxc)[3], float mpi, float mqw, const Thr &mt,
                                             uint64_t todo, uint32_t mb, const CellView &v,
                                             const int32_t *slot_off, uint32_t (&mk_lo)[kSmallRounds],
                                             uint32_t (&mk_hi)[kSmallRounds]) {
    const int lane = __lane_id();
    for (uint64_t rest = todo; rest; rest &= rest - 1) {
        const uint32_t m = (uint32_t)__builtin_ctzll(rest);
        float xi2[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int d = 0; d < DIM; ++d) xi2[d] = bcast(mxc[d], m);
        float pi = 0.0f, hi = 0.0f;
        Thr tq{0.f, 0.f};
        if (MODE == TCR_MODE_GATHER) {
            tq = Thr{bcast(mt.lo, m), bcast(mt.hi, m)};
        } else {
            pi = bcast(mpi, m);
            hi = bcast(mqw, m);
        }
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const uint64_t hitm = round_hits<DIM, MODE>(
                a, cd, g, cx[r], cp[r], chh[r], tj[r], xi2, pi, hi, tq, mb + m, [&]() {
                    const int32_t k = r * kWave + lane;
                    return k < v.total ? k + slot_off[k] : 0;
                });
            mk_lo[r] = put_lane(mk_lo[r], (uint32_t)hitm, m);
            mk_hi[r] = put_lane(mk_hi[r], (uint32_t)(hitm >> 32), m);
        }
    }
}

template <int DIM, int MODE>
__device__ void count_small(const GroupArgs &a, const CellView &v, uint32_t cid, int32_t *slot_off) {
    constexpr int kOff = DIM == 1 ? 3 : DIM == 2 ? 9 : 27;
    const int lane = __lane_id();
    const CloudDev &cd = a.desc[v.c];
    // first member batch: independent of the candidates, issue it first
    const uint32_t nm0 = min((uint32_t)kWave, v.ge - v.gb);
    const uint32_t ms0 = (uint32_t)lane < nm0 ? v.gb + lane : v.gb;
    uint32_t mfl = (uint32_t)lane < nm0 ? a.msflags[ms0] : 0u;
    int32_t mloc = a.msort_index[ms0];
    float4 mq = a.msorted[ms0];

    const Geo<DIM> g(as_const(a.desc) + v.c, v.center);
    const int nrounds = (v.total + kWave - 1) / kWave;  // 1..kSmallRounds
    expand_slots(v, slot_off, kOff);

    float cx[kSmallRounds][3], cp[kSmallRounds], chh[kSmallRounds];
    Thr tj[kSmallRounds];
#pragma unroll
    for (int r = 0; r < kSmallRounds; ++r) {
        cx[r][0] = cx[r][1] = cx[r][2] = 0.0f;
        cp[r] = INFINITY;
        chh[r] = 0.0f;
        tj[r] = Thr{0.f, 0.f};
        if (r >= nrounds) continue;  // uniform
        const int32_t k = r * kWave + lane;
        const bool valid = k < v.total;
        const int32_t j = valid ? k + slot_off[k] : 0;
        const float4 cv = a.sorted[j];
        if (v.wrap) g.place(cv, valid, cx[r], cp[r]);
        else g.place_direct(cv, valid, cx[r], cp[r]);
        chh[r] = cv.w;
        if (MODE == TCR_MODE_SCATTER) tj[r] = thresholds(cv.w, g.geo, g.c80, DIM);
    }

    for (uint32_t mb = v.gb; mb < v.ge; mb += kWave) {
        const uint32_t nm = min((uint32_t)kWave, v.ge - mb);
        const bool mvalid = (uint32_t)lane < nm;
        const uint32_t ms = mvalid ? mb + lane : v.gb;
        if (mb != v.gb) {
            mfl = mvalid ? a.msflags[ms] : 0u;
            mloc = a.msort_index[ms];
            mq = a.msorted[ms];
        }
        const int64_t mqg = v.qs + mloc;
        if (mvalid && mfl) {
            const uint32_t k = atomicAdd(a.stray_count, 1u);
            a.stray_list[k] = (int32_t)mqg;
        }
        const uint64_t todo = __ballot(mvalid && !mfl);
        if (!todo) continue;
        const float mqv[3] = {mq.x, mq.y, mq.z};
        float mxc[3] = {0.f, 0.f, 0.f}, mpi = 0.0f;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const float c = mqv[d] - g.center[d];
            mxc[d] = -2.0f * c;
            mpi = d == 0 ? c * c : fmaf(c, c, mpi);
        }
        Thr mt{0.f, 0.f};
        if (MODE == TCR_MODE_GATHER) {
            mt = thresholds(mq.w, g.geo, g.c80, DIM);
            mt.lo -= mpi;
            mt.hi -= mpi;
        }
        // lane m collects query m's masks
        uint32_t mk_lo[kSmallRounds], mk_hi[kSmallRounds];
#pragma unroll
        for (int r = 0; r < kSmallRounds; ++r) mk_lo[r] = mk_hi[r] = 0u;
#define TCR_TM(NR)                                                                                          \
    case NR:                                                                                                \
        test_members<DIM, MODE, NR>(a, cd, g, cx, cp, chh, tj, mxc, mpi, mq.w, mt, todo, mb, v, slot_off,  \
                                    mk_lo, mk_hi);                                                          \
        break;
        switch (nrounds) {
            TCR_TM(1) TCR_TM(2) TCR_TM(3) TCR_TM(4) TCR_TM(5) TCR_TM(6)
            default: test_members<DIM, MODE, kSmallRounds>(a, cd, g, cx, cp, chh, tj, mxc, mpi, mq.w, mt, todo,
                                                           mb, v, slot_off, mk_lo, mk_hi);
        }
#undef TCR_TM
        if (mvalid && !mfl) {
            // the count of a query is the population of its masks (rounds
            // past nrounds hold zero)
            uint32_t mhits = 0;
#pragma unroll
            for (int r = 0; r < kSmallRounds; ++r) mhits += __popc(mk_lo[r]) + __popc(mk_hi[r]);
            a.counts[mqg] = (int32_t)mhits;
#pragma unroll
            for (int r = 0; r < kSmallRounds; ++r)
                if (r < nrounds) a.masks[r * a.mask_stride + ms] = ((uint64_t)mk_hi[r] << 32) | mk_lo[r];
        }
    }
}

// Replays the stored masks.  Each query's row is assembled in LDS (rows of
// a small cell hold at most kSmallChunk entries) and leaves in coalesced
// stores.  Independent of mode and dimension except for the stencil size.
// Output kinds of the fill kernel: bit 1 = int32 indices (else int64),
// bit 0 = row indices written too (COO).
template <int KIND>
__device__ __forceinline__ void put_pair(const PairOut &o, int64_t pos, int64_t r, int32_t c) {
    if (KIND & 2) {
        ((int32_t *)o.col)[pos] = c;
        if (KIND & 1) ((int32_t *)o.row)[pos] = (int32_t)r;
    } else {
        ((int64_t *)o.col)[pos] = c;
        if (KIND & 1) ((int64_t *)o.row)[pos] = r;
    }
}
static int pair_kind(const PairOut &o) { return (o.idx32 ? 2 : 0) | (o.row ? 1 : 0); }

template <int KIND>
__device__ void fill_small(const GroupArgs &a, const CellView &v, int noff, int32_t *slot_off, int32_t *rowbuf) {
    const int lane = __lane_id();
    const int nrounds = (v.total + kWave - 1) / kWave;
    // first member batch: independent of the candidates, issue it first
    const uint32_t nm0 = min((uint32_t)kWave, v.ge - v.gb);
    const uint32_t ms0 = (uint32_t)lane < nm0 ? v.gb + lane : v.gb;
    // masks and row offsets are read for every member (a stray member's are
    // never used), so that their loads do not wait for the stray flags
    const bool mv0 = (uint32_t)lane < nm0;
    bool live = mv0 && !a.msflags[ms0];
    uint64_t mk[kSmallRounds];
#pragma unroll
    for (int r = 0; r < kSmallRounds; ++r)
        mk[r] = (r < nrounds && mv0) ? a.masks[r * a.mask_stride + ms0] : 0ull;
    int32_t mloc = a.msort_index[ms0];
    int64_t mbase = mv0 ? a.offsets[v.qs + mloc] : 0;

    expand_slots(v, slot_off, noff);
    int32_t corig[kSmallRounds];
#pragma unroll
    for (int r = 0; r < kSmallRounds; ++r) {
        const int32_t k = r * kWave + lane;
        corig[r] = k < v.total ? a.sort_index[k + slot_off[k]] : 0;
    }
    for (uint32_t mb = v.gb; mb < v.ge; mb += kWave) {
        const uint32_t nm = min((uint32_t)kWave, v.ge - mb);
        const bool mvalid = (uint32_t)lane < nm;
        const uint32_t ms = mvalid ? mb + lane : v.gb;
        if (mb != v.gb) {
            live = mvalid && !a.msflags[ms];
#pragma unroll
            for (int r = 0; r < kSmallRounds; ++r)
                mk[r] = (r < nrounds && mvalid) ? a.masks[r * a.mask_stride + ms] : 0ull;
            mloc = a.msort_index[ms];
            mbase = mvalid ? a.offsets[v.qs + mloc] : 0;
        }
        const uint64_t todo = __ballot(live);
        if (!todo) continue;
        for (uint32_t m = 0; m < nm; ++m) {
            if (!((todo >> m) & 1ull)) continue;
            int32_t len = 0;
#pragma unroll
            for (int r = 0; r < kSmallRounds; ++r) {
                if (r >= nrounds) break;
                const uint64_t hitm = bcast(mk[r], m);
                // uniform base (SGPR) + lane rank: one shift-add for the address
                int32_t *const rowr = rowbuf + len;
                if (__builtin_amdgcn_inverse_ballot_w64(hitm)) rowr[lane_rank(hitm, 0)] = corig[r];
                len += __popcll(hitm);
            }
            __asm__ volatile("" ::: "memory");
            const int64_t rbase = bcast(mbase, m);
            const int64_t qloc = (int64_t)bcast(mloc, m);
            if (lane < len) put_pair<KIND>(a.out, rbase + lane, qloc, rowbuf[lane]);
            for (int32_t k = lane + kWave; k < len; k += kWave)  // rows longer than a wavefront
                put_pair<KIND>(a.out, rbase + k, qloc, rowbuf[k]);
            __asm__ volatile("" ::: "memory");
        }
    }
}

// --------------------------------------------------------------- big cells --

template <int DIM, int MODE, bool FILL>
__device__ void group_big(const GroupArgs &a, uint32_t cid) {
    constexpr int kOff = DIM == 1 ? 3 : DIM == 2 ? 9 : 27;
    const int lane = __lane_id();
    const CellView v = make_view(a, load_cell(a, cid), kOff);
    const CloudDev &cd = a.desc[v.c];
    const Geo<DIM> g(cd, v.center);

    for (int32_t cbase = 0; cbase < v.total; cbase += kBigChunk) {
        const int nrounds = min(kBigRounds, (int)((v.total - cbase + kWave - 1) / kWave));
        float cx[kBigRounds][3], cp[kBigRounds], chh[kBigRounds];
        int32_t corig[kBigRounds];
        Thr tj[kBigRounds];
#pragma unroll
        for (int r = 0; r < kBigRounds; ++r) {
            const int32_t k = cbase + r * kWave + lane;
            const bool valid = r < nrounds && k < v.total;
            const int32_t j = candidate_slot(v, k, kOff);
            corig[r] = (FILL && valid) ? a.sort_index[j] : 0;
            const float4 cv = a.sorted[valid ? j : 0];
            g.place(cv, valid, cx[r], cp[r]);
            chh[r] = cv.w;
            if (MODE == TCR_MODE_SCATTER) tj[r] = thresholds(cv.w, g.geo, g.c80, DIM);
        }
        for (uint32_t s = v.gb; s < v.ge; ++s) {
            if (a.msflags[s]) continue;  // listed for the per-query kernel by k_count_small
            const int32_t qloc = uniform(a.msort_index[s]);
            const int64_t qg = v.qs + qloc;
            const float4 mq = a.msorted[s];
            const float mqv[3] = {mq.x, mq.y, mq.z};
            float xi2[3] = {0.f, 0.f, 0.f}, pi = 0.0f;
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                const float c = mqv[d] - g.center[d];
                xi2[d] = -2.0f * c;
                pi = d == 0 ? c * c : fmaf(c, c, pi);
            }
            const float hi = mq.w;
            Thr tq{0.f, 0.f};
            if (MODE == TCR_MODE_GATHER) {
                tq = thresholds(hi, g.geo, g.c80, DIM);
                tq.lo -= pi;
                tq.hi -= pi;
            }
            int32_t hits = 0;
            int64_t base = 0;
            if (FILL) base = a.offsets[qg] + (cbase ? (int64_t)a.cursor[qg] : 0);
#pragma unroll
            for (int r = 0; r < kBigRounds; ++r) {
                if (r >= nrounds) break;
                const uint64_t hitm = round_hits<DIM, MODE>(
                    a, cd, g, cx[r], cp[r], chh[r], tj[r], xi2, pi, hi, tq, s,
                    [&]() { return candidate_slot(v, cbase + r * kWave + lane, kOff); });
                if (FILL) {
                    if ((hitm >> lane) & 1ull) {
                        const int64_t pos = base + __popcll(hitm & lanemask_lt());
                        a.out.put(pos, qloc, corig[r]);
                    }
                    base += __popcll(hitm);
                }
                hits += __popcll(hitm);
            }
            if (lane == 0) {
                if (!FILL) a.counts[qg] = cbase ? a.counts[qg] + hits : hits;
                else a.cursor[qg] = (cbase ? a.cursor[qg] : 0) + hits;
            }
        }
    }
}

// ------------------------------------------------------------- kernels ----

// XCD-aware static schedule: workgroup b runs on XCD b % 8; each XCD takes a
// contiguous range of cells so that neighbouring cells share its L2.  Inside
// an XCD's range the wavefronts take runs of `run` consecutive cells in
// turn (run 1: wavefront w takes cells w, w + W, w + 2W, ...).
struct CellRange {
    uint32_t first, end, step, run;
    // the cell after cid; k counts the cells taken of the current run
    __device__ __forceinline__ uint32_t next(uint32_t cid, uint32_t &k) const {
        if (k < run) {
            ++k;
            return cid + 1;
        }
        k = 1;
        return cid + 1 + step - run;
    }
};
__device__ __forceinline__ CellRange my_cells(uint32_t ncells, uint32_t run) {
    const uint32_t nxcd = 8;
    const uint32_t wpb = blockDim.x / kWave;
    const uint32_t xcd = blockIdx.x % nxcd;
    const uint32_t per_xcd_waves = (gridDim.x / nxcd) * wpb;
    const uint32_t wix = (blockIdx.x / nxcd) * wpb + uniform(threadIdx.x / kWave);
    CellRange r;
    r.run = run > 0 ? run : 1u;
    const uint32_t c0 = (uint32_t)(((uint64_t)ncells * xcd) / nxcd);
    r.end = (uint32_t)(((uint64_t)ncells * (xcd + 1)) / nxcd);
    r.first = c0 + wix * r.run;
    r.step = per_xcd_waves * r.run;
    return r;
}

__device__ __forceinline__ bool cloud_is(const CloudDev &cd, int dim, int mode) {
    return cd.h.dim == dim && (mode < 0 || cd.h.mode == mode);
}
__device__ __forceinline__ bool cloud_is(cptr<CloudDev> cd, int dim, int mode) {
    return cd->h.dim == dim && (mode < 0 || cd->h.mode == mode);
}

// 5 waves per SIMD: the register budget (<= 102 VGPRs) spills only per-cell
// state, never the query loop
template <int DIM, int MODE>
__global__ void __launch_bounds__(kGroupThreads) __attribute__((amdgpu_waves_per_eu(5, 8))) k_count_small(GroupArgs a) {
    constexpr int kOff = DIM == 1 ? 3 : DIM == 2 ? 9 : 27;
    // per wavefront: the slot table (dynamic LDS: kSmallChunk words per
    // wavefront of the block)
    extern __shared__ int32_t s_dyn[];
    int32_t *const s_slot = s_dyn + uniform(threadIdx.x / kWave) * kSmallChunk;
    const CellRange cr = my_cells(*a.ncells, a.cell_run);
    if (cr.first >= cr.end) return;
    // no next-cell prefetch here: its registers would cost the occupancy
    // that hides the same latency
    for (uint32_t cid = cr.first, k = 1; cid < cr.end; cid = cr.next(cid, k)) {
        const CellView v = make_view(a, load_cell(a, cid), kOff);
        if (v.state != kCellSkip && cloud_is(as_const(a.desc) + v.c, DIM, MODE)) {
            if (v.state == kCellStray) {
                append_strays(a, v);
            } else if (v.total > kSmallChunk) {
                if (__lane_id() == 0) a.big_list[atomicAdd(a.big_count, 1u)] = cid;
                append_flagged(a, v);  // k_group_big skips them
            } else {
                count_small<DIM, MODE>(a, v, cid, s_slot);
            }
        }
    }
}

// 8 waves per SIMD (<= 64 VGPRs, no spills)
template <int KIND>
__global__ void __launch_bounds__(kGroupThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) k_fill_small(GroupArgs a) {
    // per wavefront: the slot table and the row buffer (dynamic LDS:
    // 2 * kSmallChunk words per wavefront of the block)
    extern __shared__ int32_t s_dyn[];
    int32_t *const s_slot = s_dyn + uniform(threadIdx.x / kWave) * (2 * kSmallChunk);
    int32_t *const s_row = s_slot + kSmallChunk;
    if (*a.total > a.capacity) return;
    const CellRange cr = my_cells(*a.ncells, a.cell_run);
    if (cr.first >= cr.end) return;
    for (uint32_t cid = cr.first, k = 1; cid < cr.end; cid = cr.next(cid, k)) {
        const CellLoad cur = load_cell(a, cid);
        const uint32_t state = uniform(cur.rec.state);
        if (state == kCellFast) {
            const int dim = as_const(a.desc)[uniform(cur.rec.cloud)].h.dim;
            const int noff = dim == 1 ? 3 : dim == 2 ? 9 : 27;
            const CellView v = make_view(a, cur, noff);
            if (v.total <= kSmallChunk)  // bigger cells: k_group_big
                fill_small<KIND>(a, v, noff, s_slot, s_row);
        }
    }
}

// Cells too big for the small-cell kernels, one wavefront per cell; with
// LIST (the last launch of a pass) the wavefronts then take the stray
// queries the count pass listed, one wavefront per query -- in the same
// launch, since both lists are usually short or empty.
template <int DIM, int MODE, bool FILL, bool LIST>
__global__ void __launch_bounds__(kGroupThreads) k_group_big(GroupArgs a, GenericArgs ga) {
    if (FILL && *a.total > a.capacity) return;
    const uint32_t n = *a.big_count;
    const uint32_t wpb = kGroupThreads / kWave;
    for (uint32_t i = blockIdx.x * wpb + threadIdx.x / kWave; i < n; i += gridDim.x * wpb) {
        const uint32_t cid = a.big_list[i];
        const CloudDev &cd = a.desc[uniform(a.cell_cloud[cid])];
        if (!cloud_is(cd, DIM, MODE)) continue;
        group_big<DIM, MODE, FILL>(a, cid);
    }
    if (LIST) {
        const int64_t ns = *ga.list_count;
        for (int64_t i = (int64_t)blockIdx.x * wpb + threadIdx.x / kWave; i < ns; i += (int64_t)gridDim.x * wpb)
            generic_query_wave<FILL>(ga, ga.list[i]);
    }
}

// ------------------------------------------------------------- driver -----

// Workgroups of `kernel` that the current device holds at once (all CUs),
// cached per kernel and device.
static int64_t resident_workgroups(const void *kernel, int threads, size_t lds) {
    static std::mutex mu;
    static std::map<std::tuple<const void *, int, int>, int64_t> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    std::lock_guard<std::mutex> lock(mu);
    const auto key = std::make_tuple(kernel, dev, threads);
    const auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int per_cu = 0, cus = 0;
    int64_t r = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
        r = (int64_t)per_cu * cus;
    cache[key] = r;
    return r;
}

// Workgroups of a wavefront-per-cell kernel: by default as many as the chip
// holds at once, so that every wave starts at launch and walks an equal
// share of the cells (a second, partial round of workgroups would run the
// kernel's tail at low occupancy), and no more than one wave per reference
// point (a bound on the occupied cells); a multiple of the 8 XCDs (my_cells splits the cells by
// blockIdx % 8).
// Threads per workgroup of the wavefront-per-cell kernels (the tuned block
// size: 64, 128 or 256) and their dynamic LDS (words per wavefront).
static int cell_threads() {
    const int t = tuning().cell_threads;
    return (t == 64 || t == 128) ? t : kGroupThreads;
}
static size_t cell_lds(int threads, int words_per_wave) {
    return (size_t)(threads / kWave) * (size_t)words_per_wave * sizeof(int32_t);
}

static unsigned group_grid(int64_t npts, const void *kernel, int threads, size_t lds) {
    int64_t wgs = tuning().cell_workgroups;
    if (wgs <= 0) {
        const int64_t resident = resident_workgroups(kernel, threads, lds);
        wgs = resident > 0 ? (resident / 8) * 8 : 256 * 8;
        wgs = std::min<int64_t>(wgs, (npts * (kGroupThreads / threads) + 3) / 4);
    }
    wgs = ((wgs + 7) / 8) * 8;
    return (unsigned)(wgs > 0 ? wgs : 8);
}

template <bool FILL>
static int launch_modes(const Batch &b, const GroupArgs &ga, const GenericArgs &gen, hipStream_t s, bool qcells) {
    bool seen[4][4] = {};
    int last = -1;  // (dim, mode) of the last big-cell launch, which also takes the stray list
    for (const CloudDev &d : b.dev) {
        const bool mine = qcells ? takes_query_cells(d.h) : (d.h.self_search && d.h.search_range == 1 && d.h.nq > 0);
        if (mine) seen[d.h.dim][d.h.mode] = true;
    }
    const int64_t npts = qcells ? b.nq_total : b.nr_total;
    for (int d = 1; d <= 3; ++d)
        for (int m = 0; m < 4; ++m)
            if (seen[d][m]) last = d * 4 + m;
    // 1024 workgroups: the stray list's wavefront-per-query pass wants more
    // waves than the big cells
    const dim3 block(kGroupThreads), bgrid(256), lgrid(1024);
    const int ct = cell_threads();
    const dim3 cblock(ct);
    const size_t count_lds = cell_lds(ct, kSmallChunk), fill_lds = cell_lds(ct, 2 * kSmallChunk);
#define TCR_L(D, M)                                                                                         \
    if (seen[D][M]) {                                                                                       \
        if (!FILL) {                                                                                        \
            kernel_begin(count_events(), s);                                                                \
            hipLaunchKernelGGL((k_count_small<D, M>),                                                       \
                               dim3(group_grid(npts, (const void *)k_count_small<D, M>, ct, count_lds)), cblock, \
                               count_lds, s, ga);                                                           \
            kernel_end(count_events(), s);                                                                  \
        }                                                                                                   \
        if (b.skip_lists)                                                                                   \
            ;                                                                                               \
        else if (last == D * 4 + M)                                                                         \
            hipLaunchKernelGGL((k_group_big<D, M, FILL, true>), lgrid, block, 0, s, ga, gen);               \
        else                                                                                                \
            hipLaunchKernelGGL((k_group_big<D, M, FILL, false>), bgrid, block, 0, s, ga, gen);              \
    }
#define TCR_LD(D) TCR_L(D, 0) TCR_L(D, 1) TCR_L(D, 2) TCR_L(D, 3)
    if (!FILL) {
        TCR_LD(1) TCR_LD(2) TCR_LD(3)
    } else {
        fill_kernel_begin(s);
        switch (pair_kind(ga.out)) {
            case 0: hipLaunchKernelGGL(k_fill_small<0>, dim3(group_grid(npts, (const void *)k_fill_small<0>, ct, fill_lds)), cblock, fill_lds, s, ga); break;
            case 1: hipLaunchKernelGGL(k_fill_small<1>, dim3(group_grid(npts, (const void *)k_fill_small<1>, ct, fill_lds)), cblock, fill_lds, s, ga); break;
            case 2: hipLaunchKernelGGL(k_fill_small<2>, dim3(group_grid(npts, (const void *)k_fill_small<2>, ct, fill_lds)), cblock, fill_lds, s, ga); break;
            default: hipLaunchKernelGGL(k_fill_small<3>, dim3(group_grid(npts, (const void *)k_fill_small<3>, ct, fill_lds)), cblock, fill_lds, s, ga); break;
        }
        fill_kernel_end(s);
        TCR_LD(1) TCR_LD(2) TCR_LD(3)
    }
#undef TCR_LD
#undef TCR_L
    TCR_LAUNCH_CHECK();
    return TCR_OK;
}

int launch_generic(const Batch &b, Work &w, const int64_t *offsets, PairOut out, int64_t capacity, bool fill,
                   hipStream_t s);

static GroupArgs group_args(const Batch &b, Work &w, const int64_t *offsets, PairOut out, int64_t capacity) {
    GroupArgs ga;
    ga.desc = w.desc;
    ga.r_start = w.r_start;
    ga.q_start = w.q_start;
    ga.misc = w.misc;
    ga.cell_begin = w.cell_begin;
    ga.cell_cloud = w.cell_cloud;
    ga.cell_hash = w.cell_hash;
    ga.lintab = w.lintab;
    ga.sorted = w.sorted;
    ga.sort_index = w.sort_index;
    ga.sflags = w.sflags;
    ga.ranges = w.ranges;
    ga.recs = (CellRec *)w.recs;
    ga.masks = w.masks;
    ga.mask_stride = b.nr_total;
    ga.counts = w.counts;
    ga.offsets = offsets;
    ga.total = offsets ? offsets + b.nq_total : nullptr;
    ga.capacity = capacity;
    ga.out = out;
    ga.cursor = w.cursor;
    ga.stray_list = w.stray_list;
    ga.stray_count = w.misc + kMiscStrays;
    ga.big_list = w.big_list;
    ga.big_count = w.misc + kMiscBig;
    ga.cell_run = (uint32_t)std::max(1, tuning().cells_per_wave);
    ga.ncells = w.misc + kMiscCells;
    ga.msorted = w.sorted;
    ga.msort_index = w.sort_index;
    ga.msflags = w.sflags;
    ga.qcell_begin = w.qcell_begin;
    ga.qcell_lin = w.qcell_lin;
    ga.qcell_cloud = w.qcell_cloud;
    return ga;
}

// The same for the query cells: members are the sorted queries; the cell
// records, stencil ranges, masks and lists are the query cells' own; the
// candidate lookups stay in the reference structure.
static GroupArgs query_cell_args(const Batch &b, Work &w, const int64_t *offsets, PairOut out, int64_t capacity) {
    GroupArgs ga = group_args(b, w, offsets, out, capacity);
    ga.cell_cloud = w.qcell_cloud;   // cloud of a (query) cell id, for the big-cell kernel
    ga.ranges = w.qranges;
    ga.recs = (CellRec *)w.qrecs;
    ga.masks = w.qmasks;
    ga.mask_stride = b.nq_total;
    ga.stray_list = w.qstray_list;
    ga.stray_count = w.misc + kMiscQStrays;
    ga.big_list = w.qbig_list;
    ga.big_count = w.misc + kMiscQBig;
    ga.ncells = w.misc + kMiscQCells;
    ga.msorted = w.qsorted;
    ga.msort_index = w.qsort_index;
    ga.msflags = w.qsflags;
    return ga;
}

// Count or fill pass over the batch: wavefront-per-cell kernels for the
// self-search clouds (their strays collected by the count pass), the
// per-query kernel for every other cloud and for the strays.
static int query_pass(const Batch &b, Work &w, const int64_t *offsets, PairOut out, int64_t capacity, bool fill,
                      hipStream_t s) {
    bool any_group = false, any_qcells = false, any_other = false;
    for (const CloudDev &d : b.dev) {
        if (d.h.nq == 0) continue;
        if (d.h.self_search && d.h.search_range == 1) any_group = true;
        else if (takes_query_cells(d.h)) any_qcells = true;
        else any_other = true;
    }
    if (b.nq_total == 0) return TCR_OK;
    if (any_group) {
        const GroupArgs ga = group_args(b, w, offsets, out, capacity);
        if (!fill) {
            // misc[kMiscStrays] and misc[kMiscBig] start at zero: the
            // header upload clears all of misc
            // one thread per cell: enough workgroups for every occupied cell
            // (at most one cell per reference point)
            const int32_t sw = tuning().stencil_workgroups;
            const unsigned sgrid = sw > 0 ? (unsigned)sw : (unsigned)std::max<int64_t>(1, (b.nr_total + 255) / 256);
            hipLaunchKernelGGL(k_stencil, dim3(sgrid), dim3(256), 0, s, ga);
        }
        const GenericArgs gen = generic_args(b, w, offsets, out, capacity);
        const int rc = fill ? launch_modes<true>(b, ga, gen, s, false) : launch_modes<false>(b, ga, gen, s, false);
        if (rc) return rc;
    }
    if (any_qcells) {
        const GroupArgs qa = query_cell_args(b, w, offsets, out, capacity);
        if (!fill) {
            // one thread per query cell (at most one per query)
            const int32_t sw = tuning().stencil_workgroups;
            const unsigned sgrid = sw > 0 ? (unsigned)sw : (unsigned)std::max<int64_t>(1, (b.nq_total + 255) / 256);
            hipLaunchKernelGGL(k_qstencil, dim3(sgrid), dim3(256), 0, s, qa);
        }
        GenericArgs gen = generic_args(b, w, offsets, out, capacity);
        gen.list = w.qstray_list;
        gen.list_count = w.misc + kMiscQStrays;
        const int rc = fill ? launch_modes<true>(b, qa, gen, s, true) : launch_modes<false>(b, qa, gen, s, true);
        if (rc) return rc;
    }
    // every other cloud, and query-cell clouds the cell kernels do not take
    // (none, when the launch history says so: skip_generic)
    if (any_other || (any_qcells && !b.skip_generic)) return launch_generic(b, w, offsets, out, capacity, fill, s);
    return TCR_OK;
}

int query_count(const Batch &b, Work &w, hipStream_t s) {
    return query_pass(b, w, nullptr, PairOut{nullptr, nullptr, 0}, -1, false, s);
}
int query_fill(const Batch &b, Work &w, const int64_t *offsets, PairOut out, int64_t capacity, hipStream_t s) {
    return query_pass(b, w, offsets, out, capacity, true, s);
}

}  // namespace tcr
