// Host-side internals shared by the translation units of libtcr_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "tcr_device.h"

namespace tcr {

constexpr int kSortThreads = 1024;         // 16 wavefronts per sort tile
constexpr int kSortItems = 4;              // keys per thread
constexpr int kSortTile = kSortThreads * kSortItems;  // 4096 keys per tile
constexpr int kRadixBits = 9;              // 2 passes for grids of up to 2^18 cells
constexpr int kRadix = 1 << kRadixBits;
constexpr int kMaxPasses = 4;              // 32-bit keys (36 bits of digits)

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;  // 2048

// misc[] slots (device counters of one search)
enum : int {
    kMiscCells = 0,      // occupied cells of the batch
    kMiscStrays = 1,     // stray queries listed by the count pass
    kMiscBig = 2,        // cells with more candidates than one register chunk
    kMiscPasses = 3,     // radix passes of this search
    kMiscCellTicket = 5, // tile ticket of k_cells
    kMiscErr = 6,        // nonzero: some cloud has a degenerate grid
    kMiscMlmFlags = 7,   // multi-level search: inputs it does not handle (see tcr_mlm.hip)
    kMiscSortTicket = 8, // tile tickets of the radix passes (kMaxPasses)
    kMiscLevelMax = 12,  // multi-level search: ordered max support per level (4 words)
    // the query cells of query != reference clouds (see build_queries)
    kMiscQCells = 16,      // occupied query cells of the batch
    kMiscQPasses = 17,     // radix passes of the query sort
    kMiscQCellTicket = 18, // tile ticket of k_qcells
    kMiscQStrays = 19,     // queries the query-cell count pass leaves to the per-query path
    kMiscQBig = 20,        // query cells with more candidates than one register chunk
    kMiscQSortTicket = 24, // tile tickets of the query radix passes (kMaxPasses)
    kMiscGeneric = 28,     // nonzero: some cloud's queries need the per-query kernel (k_grid)
    kMiscScanTicket = 29,  // tile ticket of the offsets scan
    kMiscWords = 32,
};

// Error reporting (thread-local message; see tcr_last_error()).
void set_error(const std::string &msg);
#define TCR_HIP(call)                                                              \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            ::tcr::set_error(std::string(#call) + ": " + hipGetErrorString(e_));   \
            return TCR_ERR_HIP;                                                    \
        }                                                                          \
    } while (0)
#define TCR_LAUNCH_CHECK()                                                         \
    do {                                                                           \
        hipError_t e_ = hipGetLastError();                                         \
        if (e_ != hipSuccess) {                                                    \
            ::tcr::set_error(std::string("kernel launch: ") + hipGetErrorString(e_)); \
            return TCR_ERR_HIP;                                                    \
        }                                                                          \
    } while (0)

// Bump allocator over a caller workspace; layout is a pure function of the
// clouds so that size query and use agree.
struct Carve {
    char *base;
    size_t off = 0;
    explicit Carve(void *b) : base((char *)b) {}
    template <typename T>
    T *take(size_t n) {
        off = (off + 255) & ~size_t(255);
        T *p = base ? (T *)(base + off) : nullptr;
        off += sizeof(T) * (n ? n : 1);
        return p;
    }
};

// Batch geometry derived on the host from the tcr_cloud array.
struct Batch {
    int32_t B = 0;
    int64_t nq_total = 0, nr_total = 0;   // points
    int64_t rt_total = 0;                 // sort tiles (reference points)
    int64_t qt_total = 0;                 // sort tiles of the queries of query-cell clouds
    int64_t hb_total = 0;                 // hash buckets (reference layout)
    int64_t lt_total = 0;                 // lin-table slots
    bool any_nonself = false;
    std::vector<CloudDev> dev;            // host copy of the descriptors
    // Launches of this search.  A repeated search on one workspace launches
    // only the radix passes and list kernels the previous search on it
    // needed; the device still counts what this one needs, and the host
    // repeats the search with every launch when that was more (finish_search).
    int sort_passes = kMaxPasses;         // radix passes over the reference keys
    int qsort_passes = kMaxPasses;        // over the query keys of query-cell clouds
    bool skip_lists = false;              // no big-cell / stray-query launches
    bool skip_generic = false;            // no per-query kernel for the query-cell clouds
};

// Workspace pointers.
struct Work {
    CloudDev *desc;
    int64_t *q_start;     // [B+1] query point prefix
    int64_t *r_start;     // [B+1] reference point prefix
    int64_t *rt_start;    // [B+1] reference tile prefix
    int64_t *hb_start;    // [B+1] bucket prefix (reference-layout tables)
    int64_t *lt_start;    // [B+1] lin-table slot prefix (power-of-two per cloud)
    int64_t *qt_start;    // [B+1] query tile prefix (query-cell clouds)
    uint32_t *misc;       // [kMiscWords]
    float4 *rwrap;        // [Nr] wrapped reference positions (+support in .w)
    uint32_t *keys0, *keys1;  // cell linear index (raw int32 bits)
    uint32_t *vals0, *vals1;  // local original index
    uint8_t *pflags;      // [Nr] stray bits per reference point (input order)
    uint32_t *sort_hist;  // [B * kMaxPasses * kRadix] digit counts per cloud
    uint64_t *sort_lb;    // [kMaxPasses * tiles * kRadix] look-back state
    uint64_t *cell_lb;    // [tiles] look-back state of the cell extraction
    float4 *sorted;       // [Nr] sorted positions (+support)
    int32_t *sort_index;  // [Nr] sorted slot -> local original index
    int32_t *skey;        // [Nr] sorted linear cell index
    uint8_t *sflags;      // [Nr] stray bits per sorted slot
    uint32_t *cell_begin; // [Nr+1] first global slot of each cell
    int32_t *cell_lin;    // [Nr]
    uint32_t *cell_hash;  // [Nr] bucket of the cell (hash of its first point's cell)
    uint32_t *cell_cloud; // [Nr]
    uint64_t *lintab;     // [LT] (cell id << 32 | lin), all ones = empty
    uint32_t *bucket_cnt; // [HB] (export)
    uint32_t *bucket_begin;
    uint32_t *bucket_fill;
    int32_t *cell_table;  // [Nr*3] reference layout, grouped by bucket (export)
    int32_t *hash_table;  // [HB*2] (export)
    uint64_t *scan_part;  // scan partials
    int32_t *counts;      // [Nq]
    int32_t *cursor;      // [Nq] per-query write cursor across chunks
    int32_t *stray_list;  // [Nq] global query indices for the per-query path
    uint64_t *masks;      // [Nr * 7] per sorted slot: hit mask of each candidate round
    int2 *ranges;         // [Nr * 27] per cell: stencil cell (first global slot, count)
    float4 *recs;         // [Nr * 2] per cell: 32-byte record (slots, cloud, state, centre)
    uint32_t *big_list;   // [Nr] cells with more candidates than one register chunk
    // Query cells of the clouds whose query set is not the reference set
    // (searchRange 1): the queries sorted by the cell of the reference grid
    // that holds them, so that a wavefront takes one cell's queries.  Arrays
    // of length Nq (all queries of the batch; only the query-cell clouds'
    // entries are used), empty when no cloud takes this path.
    float4 *qwrap;        // [Nq] once-wrapped query positions (+support)
    uint32_t *qkeys0, *qkeys1, *qvals0, *qvals1;
    uint8_t *qpflags;     // [Nq] bit 1: the query's cell is outside the grid
    uint32_t *qsort_hist; // [B * kMaxPasses * kRadix]
    uint64_t *qsort_lb;   // [kMaxPasses * query tiles * kRadix]
    uint64_t *qcell_lb;   // [query tiles]
    float4 *qsorted;      // [Nq] sorted query positions (+support)
    int32_t *qsort_index; // [Nq] sorted query slot -> local query index
    int32_t *qskey;       // [Nq] sorted cell key
    uint8_t *qsflags;     // [Nq] per sorted query slot
    uint32_t *qcell_begin;  // [Nq+1] first sorted query slot of each query cell
    int32_t *qcell_lin;   // [Nq] the cell's linear index
    uint32_t *qcell_cloud;  // [Nq]
    uint64_t *qmasks;     // [Nq * 7]
    int2 *qranges;        // [Nq * 27]
    float4 *qrecs;        // [Nq * 2]
    uint32_t *qbig_list;  // [Nq]
    int32_t *qstray_list; // [Nq]
};

// Bucket layout of the reference (hashTable [L,2] + cellTable [C,3]); used
// by the reference-layout ops.
struct GridView {
    const int32_t *hash_table;  // [L,2] of this cloud
    const int32_t *cell_table;  // [C,3] of this cloud
    uint32_t L;
    int32_t nc[3];
    uint32_t pmask;
    float mn[3], mx[3];
    float hcell;
    int32_t sr;
    int32_t mode;
};

// Arguments of the reference-layout countNeighbors / buildNeighborList ops.
struct RefOpArgs {
    const float *qpos, *qsup;
    int64_t nq;
    GridView g;
    const float *spos, *ssup;
    const int32_t *offsets32;
    int64_t *li, *lj;
    int32_t *counts;
};
int refop_launch(int dim, bool fill, const RefOpArgs &a, hipStream_t s);
int hash_indices_launch(const int32_t *cells, int64_t n, int dim, uint32_t L, int32_t *out, hipStream_t s);

// lt_mult: lin-table slots per reference point before rounding up to a power
// of two (2 for the single-level search; the multi-level search also stores
// its groups there and takes 4).
bool make_batch(const tcr_cloud *clouds, int32_t n, Batch &b, std::string &err, int lt_mult = 2);
size_t carve_work(const Batch &b, void *ws, Work &w);

// Fills n 32-bit words with v (a kernel, so it is safe inside graph capture).
int fill_u32(void *p, uint32_t v, size_t n, hipStream_t s);

// Optional HIP events around the fill pass's main kernel (benchmark timing):
// the launcher records `before` and `after` around the first fill kernel it
// launches and sets `recorded`.
struct KernelEvents {
    hipEvent_t before = nullptr, after = nullptr;
    bool recorded = false;
};
KernelEvents &fill_events();   // this thread's, around the fill pass's main kernel
KernelEvents &count_events();  // this thread's, around the first wavefront-per-cell count kernel
inline void kernel_begin(KernelEvents &k, hipStream_t s) {
    if (k.before && !k.recorded) (void)hipEventRecord(k.before, s);
}
inline void kernel_end(KernelEvents &k, hipStream_t s) {
    if (k.after && !k.recorded) {
        (void)hipEventRecord(k.after, s);
        k.recorded = true;
    }
}
inline void fill_kernel_begin(hipStream_t s) { kernel_begin(fill_events(), s); }
inline void fill_kernel_end(hipStream_t s) { kernel_end(fill_events(), s); }

// Word ranges a kernel zeroes on the side (k_prep: the sort, cell and
// offsets-scan look-back state of the search).
struct ClearSpans {
    static constexpr int kMax = 7;
    uint32_t *p[kMax];
    int64_t n[kMax];
};

// Generic device-wide exclusive scan: out[i] = sum(in[0..i)), and
// out[n] = total when write_total.  Tin: uint32_t or int32_t;
// Tout: uint32_t or int64_t.  The same launches write the CSR arrays of
// int64 offsets when asked for: indptr int32 [n+1], rowEntries int64 [n].
template <typename Tin, typename Tout>
int exclusive_scan(const Tin *in, Tout *out, int64_t n, bool write_total, uint64_t *partials,
                   hipStream_t s, int32_t *indptr = nullptr, int64_t *row_entries = nullptr);
size_t scan_partials_needed(int64_t n);
// Single-launch form for the counts -> int64 offsets scan of a search: tiles
// take tickets from *ticket and chain their prefixes by decoupled look-back
// over status[0 .. tiles); both must be zero when the launch starts (the
// status words fit in scan_partials_needed(n) words).
int offsets_scan(const int32_t *counts, int64_t *offsets, int64_t n, uint64_t *status, uint32_t *ticket,
                 hipStream_t s, int32_t *indptr = nullptr, int64_t *row_entries = nullptr);
int64_t offsets_scan_tiles(int64_t n);
int csr_arrays_launch(const int64_t *offsets, int64_t n, int32_t *indptr, int64_t *row_entries, hipStream_t s);

// Launch-shape parameters (tcr_set_tuning); 0 = automatic.
struct Tuning {
    int32_t cell_workgroups = 0;
    int32_t stencil_workgroups = 0;
    int32_t cell_threads = 0;     // 0: 256
    int32_t cells_per_wave = 0;   // 0: 1
};
Tuning &tuning();

// Pair outputs of the fill: row (may be null) and col, with int64 entries
// (the reference's dtype) or int32 entries.
struct PairOut {
    void *row, *col;
    int32_t idx32;
    __device__ __forceinline__ void put(int64_t pos, int64_t r, int32_t c) const {
        if (idx32) {
            ((int32_t *)col)[pos] = c;
            if (row) ((int32_t *)row)[pos] = (int32_t)r;
        } else {
            ((int64_t *)col)[pos] = c;
            if (row) ((int64_t *)row)[pos] = r;
        }
    }
};

// largest slot of the multi-level emit pass (rows sort in LDS up to this length)
constexpr int64_t kMlmSlotMax = 2048;
int mlm_sort_rows_launch(const Batch &b, const int64_t *offsets, uint64_t *scratch, PairOut out, hipStream_t s,
                         const uint32_t *slots = nullptr, int64_t slot_cap = 0);

// Structure build (hMax, grid, keys, sort, cells, lin table).  Expects the
// header block (descriptors, prefixes, zeroed misc) uploaded.
int build_structure(const Batch &b, Work &w, hipStream_t s);
// Clouds whose queries the query-cell path sorts and groups: a query set
// that is not the reference set, searchRange 1, points on both sides.
__host__ __device__ inline bool takes_query_cells(const tcr_cloud &h) {
    return !h.self_search && h.search_range == 1 && h.nq > 0 && h.nr > 0;
}
// Clouds whose queries the wavefront-per-cell kernels own (their strays
// come back through the stray lists): self-searches of searchRange 1, and
// the query-cell clouds whose grid the cell kernels take (k_grid's fast
// flag).
__device__ __forceinline__ bool owned_by_cells(const CloudDev &cd) {
    if (cd.h.self_search) return cd.h.search_range == 1;
    return takes_query_cells(cd.h) && cd.fast;
}
// Query keys, query sort and query cells of those clouds (after the
// reference structure).
int build_queries(const Batch &b, Work &w, hipStream_t s);
// Its stages: hMax + grid of every cloud (k_prep), and -- after the keys
// (keys0/vals0/rwrap/pflags) are written -- sort, cells and lin table.
int build_prep(const Batch &b, Work &w, hipStream_t s);
int build_sorted_cells(const Batch &b, Work &w, hipStream_t s, bool hist_done = false);
int cloud_cells_launch(const Batch &b, Work &w, hipStream_t s);

// Multi-level search of one cloud (tcr_mlm.hip): level keys, group table,
// count/fill traversal, per-row ordering.
int mlm_keys_launch(const Batch &b, Work &w, int levels, hipStream_t s);
int mlm_groups_launch(const Batch &b, Work &w, hipStream_t s);
int mlm_query_launch(const Batch &b, Work &w, int levels, bool fill, const int64_t *offsets, uint64_t *scratch,
                     hipStream_t s, uint32_t *slots = nullptr, int64_t slot_cap = 0);
// Count pass (counts[] and the fill's masks); fill pass into offsets,
// skipped on the device when offsets[nq_total] > capacity.
int query_count(const Batch &b, Work &w, hipStream_t s);
int query_fill(const Batch &b, Work &w, const int64_t *offsets, PairOut out, int64_t capacity, hipStream_t s);
// CSR arrays of a search (either may be null): indptr int32 [nq+1], rowEntries int64 [nq].
struct CsrOut {
    int32_t *indptr = nullptr;
    int64_t *row_entries = nullptr;
};
// The whole search of tcr_search_launch on the workspace `ws`, completed
// (one synchronisation) without allocating: pair total to *total_pairs,
// TCR_ERR_DOMAIN for a degenerate grid, TCR_ERR_SPACE when the pairs exceed
// the capacity (outputs not filled).
int search_complete(const tcr_cloud *clouds, int32_t n, const float *qpos, const float *qsup, const float *rpos,
                    const float *rsup, void *ws, size_t ws_bytes, int64_t *offsets, PairOut out, int64_t capacity,
                    int64_t *total_pairs, hipStream_t s);
// Pair outputs from the C ABI's (row, col, index_bytes); false (error set)
// for an index width other than 4 or 8.
bool pair_out(void *row, void *col, int32_t index_bytes, PairOut &out);
// Reference-layout bucket tables of every cloud (for export).
int build_bucket_tables(const Batch &b, Work &w, hipStream_t s);
// Segmented stable radix sort of the reference keys with the sorted
// positions gathered by its last pass.
// hist_done: the digit histograms (and the pass count) were built by the
// compact path's keys kernel; else k_sort_hist builds them.
int radix_sort(const Batch &b, Work &w, hipStream_t s, bool hist_done = false);
int radix_sort_queries(const Batch &b, Work &w, hipStream_t s);

// Decoupled look-back state: a 64-bit word per (tile[, digit]) holding a
// 32-bit count and a flag -- 0 not published, aggregate of the tile alone, or
// inclusive prefix up to and including the tile.
constexpr uint64_t kLbAgg = 1ull << 32;
constexpr uint64_t kLbInc = 2ull << 32;
__device__ __forceinline__ uint64_t lb_load(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void lb_store(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Exclusive prefix over tiles [first, t]: sums the published aggregates back
// to the nearest inclusive value, reading kLookBack predecessors per step
// (wider windows measured slower on the 1M-point sort); tiles that have not
// published yet are waited for.  Tile `first` always publishes an inclusive
// value, so the walk ends there at the latest.
constexpr int kLookBack = 1;
__device__ __forceinline__ uint64_t look_back(const uint64_t *lb, int64_t stride, int64_t t, int64_t first) {
    uint64_t excl = 0;
    while (true) {
        const int cnt = (int)min((int64_t)kLookBack, t - first + 1);
        uint64_t v[kLookBack];
#pragma unroll
        for (int k = 0; k < kLookBack; ++k) v[k] = k < cnt ? lb_load(lb + (t - k) * stride) : 0ull;
        int k = 0;
        bool done = false;
        for (; k < cnt; ++k) {
            const uint64_t flag = v[k] & ~0xffffffffull;
            if (flag == 0) break;  // not published yet: resume from this tile
            excl += (uint32_t)v[k];
            if (flag == kLbInc) {
                done = true;
                break;
            }
        }
        if (done) return excl;
        t -= k;
        if (k < cnt) __builtin_amdgcn_s_sleep(1);
    }
}

// Inclusive wave64 scan of v.
template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T v) {
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        T n = __shfl_up(v, o, kWave);
        if (lane >= o) v += n;
    }
    return v;
}

// Exclusive scan of the THREADS * ITEMS values in lds (in place); returns the
// block total.  Each thread owns ITEMS consecutive values.
template <typename T, int THREADS = kScanThreads, int ITEMS = kScanItems>
__device__ __forceinline__ T block_exclusive_scan_tile(T *lds, T *wave_tot) {
    const int t = threadIdx.x;
    T loc[ITEMS];
    T run = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        loc[k] = lds[t * ITEMS + k];
        run += loc[k];
    }
    T incl = wave_inclusive_scan<T>(run);
    const int wave = t / kWave, lane = t & (kWave - 1);
    if (lane == kWave - 1) wave_tot[wave] = incl;
    __syncthreads();
    T wave_pre = 0, total = 0;
#pragma unroll
    for (int w = 0; w < THREADS / kWave; ++w) {
        const T x = wave_tot[w];
        if (w < wave) wave_pre += x;
        total += x;
    }
    T ex = wave_pre + incl - run;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        lds[t * ITEMS + k] = ex;
        ex += loc[k];
    }
    __syncthreads();
    return total;
}

// Look-back over tiles that publish 62-bit sums (the offsets scan): 2 flag
// bits over the value.
constexpr uint64_t kScAgg = 1ull << 62;        // the tile's own sum
constexpr uint64_t kScInc = 2ull << 62;        // the sum of tiles [0, t]
constexpr uint64_t kScValue = (1ull << 62) - 1;

// Wave 0 of tile t: publishes the tile's sum, then reads the 64 nearest
// earlier tiles at once (one lane each; "tile -1" reads as inclusive 0),
// adds the sums back to the nearest inclusive one and repeats 64 tiles
// further back while none is inclusive.  Publishes and returns the
// exclusive prefix.  Tiles take tickets in launch order, so every earlier
// tile is running and publishes without waiting for this one.
__device__ __forceinline__ uint64_t tile_look_back(uint64_t *status, int64_t tile, uint64_t agg) {
    const int lane = threadIdx.x & (kWave - 1);
    if (tile == 0) {
        if (lane == 0) lb_store(status, kScInc | agg);
        return 0;
    }
    if (lane == 0) lb_store(status + tile, kScAgg | agg);
    uint64_t excl = 0;
    int64_t end = tile - 1;
    while (true) {
        const int64_t j = end - lane;
        const uint64_t v = j >= 0 ? lb_load(status + j) : kScInc;
        const uint64_t flag = v & ~kScValue;
        const uint64_t inc = __ballot(flag == kScInc);
        const uint64_t unready = __ballot(flag == 0);
        const int k = inc ? __builtin_ctzll(inc) : kWave - 1;  // last lane of this window
        const uint64_t upto = k == kWave - 1 ? ~0ull : ((2ull << k) - 1);
        if (unready & upto) {
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        uint64_t x = lane <= k ? (v & kScValue) : 0ull;
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, kWave);
        excl += x;
        if (inc) break;
        end -= kWave;
    }
    if (lane == 0) lb_store(status + tile, kScInc | (excl + agg));
    return excl;
}

// lin -> cell lookup with the reference's bucket semantics: the class with
// linear index `lin` answers a lookup for cell coordinates whose bucket is
// `bucket` only if the class lives in that bucket (its first point's cell
// hashes there) -- the outcome of queryHashMap (cpp/neighborhood.h:35-58).
__device__ __forceinline__ uint32_t lintab_hash(int32_t lin) { return (uint32_t)lin * 0x9E3779B1u; }
__device__ __forceinline__ int32_t lintab_find(const uint64_t *__restrict__ tab, uint64_t cap, int32_t lin) {
    uint64_t h = lintab_hash(lin) & (cap - 1);
    while (true) {
        const uint64_t e = tab[h];
        if (e == ~0ull) return -1;
        if ((int32_t)(uint32_t)e == lin) return (int32_t)(e >> 32);
        h = (h + 1) & (cap - 1);
    }
}

__device__ __forceinline__ uint64_t grid_cells(const CloudDev &cd) {
    uint64_t P = 1;
    for (int d = 0; d < cd.h.dim; ++d) P *= (uint64_t)(cd.nc[d] > 0 ? cd.nc[d] : 1);
    return P;
}

// A cloud's cell index: when every reference point lies in the grid and the
// grid has no more cells than the lin-table memory holds int32 entries, a
// dense int32 array indexed by the linear index (one load, L2-resident for
// dense clouds); the open-addressing table otherwise.  Both answer "the cell
// with linear index lin", so lookups keep the bucket check on top.
struct CellIndex {
    const uint64_t *tab;
    uint64_t cap, P;
    bool dense;
    bool empty;  // no reference points: no cells (the table is not cleared)
    __device__ __forceinline__ int32_t find(int32_t lin) const {
        if (empty) return -1;
        if (dense) return (uint32_t)lin < P ? ((const int32_t *)tab)[lin] : -1;
        return lintab_find(tab, cap, lin);
    }
};
__device__ __forceinline__ bool cell_index_dense(const CloudDev &cd) {
    const uint64_t P = grid_cells(cd);
    return !cd.irregular && P <= 2 * (uint64_t)cd.lt_cap && P < (1ull << 31);
}
__device__ __forceinline__ CellIndex cell_index(const CloudDev &cd, const uint64_t *lintab) {
    CellIndex ci;
    ci.tab = lintab + cd.lt_off;
    ci.cap = (uint64_t)cd.lt_cap;
    ci.P = grid_cells(cd);
    ci.dense = cell_index_dense(cd);
    ci.empty = cd.h.nr == 0;
    return ci;
}

// ---------------------------------------------------------- sort tiles --
// Sort tiles: kSortTile consecutive reference points of one cloud; a cloud's
// tiles are numbered consecutively from t_start[c].

struct TileInfo {
    int c;           // cloud
    int64_t start;   // first global key of the tile
    int32_t len;     // keys in the tile
    int64_t first;   // first tile of the cloud
};

__device__ __forceinline__ TileInfo tile_info(int64_t t, int32_t B, const int64_t *__restrict__ t_start,
                                              const int64_t *__restrict__ p_start) {
    // last cloud whose first tile <= t (clouds without tiles share their
    // successor's first tile and are skipped by taking the last match)
    const int c = find_segment(B, t, [&](int k) { return t_start[k]; });
    TileInfo ti;
    ti.c = c;
    ti.first = t_start[c];
    const int64_t lt = t - t_start[c];
    const int64_t n = p_start[c + 1] - p_start[c];
    ti.start = p_start[c] + lt * kSortTile;
    const int64_t rem = n - lt * kSortTile;
    ti.len = (int32_t)(rem < kSortTile ? rem : kSortTile);
    return ti;
}

// Sort key: the int32 linear index in signed order.  A cloud whose keys all
// lie in [0, prod numCells) < 2^31 sorts them unsigned (same order, fewer
// passes); any other ("wide") cloud flips the sign bit and takes all passes.
__device__ __forceinline__ uint32_t eff_key(uint32_t lin, bool wide) { return wide ? (lin ^ 0x80000000u) : lin; }

// Points outside the grid (INT32_MIN coordinates) or more than 2^31 cells
// (linear indices wrap to negative values).
__device__ __forceinline__ bool cloud_wide(const CloudDev &cd) {
    return cd.irregular || grid_cells(cd) > (1ull << 31);
}

// Radix passes a cloud needs.
__device__ __forceinline__ uint32_t cloud_passes(const CloudDev &cd) {
    if (cloud_wide(cd)) return kMaxPasses;
    const uint64_t P = grid_cells(cd);
    uint32_t bits = 1;
    while (bits < 32 && (1ull << bits) < P) ++bits;
    return (bits + kRadixBits - 1) / kRadixBits;
}

// Histograms count the digits of the raw linear key; a wide cloud sorts the
// sign-flipped key, whose digit in the pass holding bit 31 is the raw digit
// with that bit flipped.
constexpr int kFlipPass = 31 / kRadixBits;
constexpr uint32_t kFlipDigit = 1u << (31 % kRadixBits);
__device__ __forceinline__ int hist_digit(int d, int pass, bool wide) {
    return (wide && pass == kFlipPass) ? (int)((uint32_t)d ^ kFlipDigit) : d;
}

// Adds the raw digits of every pass of `key` to a tile's LDS histograms
// h[pass][digit] with one atomic per run of equal digits among the
// wavefront's lanes: the lanes hold consecutive keys, which share digits in
// long runs when the points come in spatial order (a single atomic per lane
// then serialises up to 64 updates of one word).  Every lane of the
// wavefront calls it; the valid lanes form a prefix.
__device__ __forceinline__ void hist_add_runs(uint32_t (*h)[kRadix], uint32_t key, bool valid) {
    const int lane = __lane_id();
    const uint32_t nv = (uint32_t)__popcll(__ballot(valid));
#pragma unroll
    for (int p = 0; p < kMaxPasses; ++p) {
        const uint32_t d = valid ? (key >> (p * kRadixBits)) & (kRadix - 1) : 0xffffffffu;
        const uint32_t prev = (uint32_t)__shfl_up((int)d, 1, kWave);
        const bool head = valid && (lane == 0 || prev != d);
        const uint64_t heads = __ballot(head);
        if (head) {
            const uint64_t after = lane == kWave - 1 ? 0ull : (heads >> (lane + 1));
            const uint32_t next = after ? (uint32_t)(lane + 1 + __builtin_ctzll(after)) : nv;
            atomicAdd(&h[p][d], next - (uint32_t)lane);
        }
    }
}

}  // namespace tcr
