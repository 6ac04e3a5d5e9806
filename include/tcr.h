/*
 * tcr.h -- C ABI of libtcr_hip.so, the MI355X (gfx950) compact-hashing radius
 * search.  Plain pointers and sizes only; every device pointer is HIP device
 * memory; every call is asynchronous on the given HIP stream (hipStream_t
 * passed as void*) unless its comment says it synchronises.
 *
 * Reference interfaces replaced (torchCompactRadius 0.5.5, paths relative to
 * its repository; cpp/ = src/torchCompactRadius/cppSrc/, py/ = src/torchCompactRadius/):
 *   tcr_compute_hash_indices   <- torch.ops.*.computeHashIndices
 *                                 (cpp/hashing.h:38, cpp/cppWrapper.cpp:45)
 *   tcr_count_neighbors        <- torch.ops.*.countNeighbors
 *                                 (cpp/neighborhood.h:102-108, cpp/cppWrapper.cpp:39)
 *   tcr_build_neighbor_list    <- torch.ops.*.buildNeighborList
 *                                 (cpp/neighborhood.h:117-124, cpp/cppWrapper.cpp:40)
 *   tcr_search_count/_fill     <- neighborSearch_cpp + searchNeighbors_cpp
 *                                 (py/compactHashing/cppSearch.py:12-98,179-244),
 *                                 i.e. the compact branch of radiusSearch
 *                                 (py/neighborhood.py:283) below the Python API,
 *                                 for one cloud or a batch of clouds
 *   tcr_search_export          <- the CompactHashMap record returned with
 *                                 returnStructure=True
 *                                 (py/searchAlgorithms/compactSearch.py:125-143)
 *   tcr_naive_count/_fill      <- radiusNaive / radiusNaiveFixed
 *                                 (py/searchAlgorithms/radiusNaive.py:8-56),
 *                                 radiusSearch's default algorithm='naive'
 *   tcr_small_count/_fill      <- neighborSearchSmall / neighborSearchSmallFixed
 *                                 (cpp/neighborhoodSmall_cpu.cpp:45-302, commented
 *                                 out of the reference build), algorithm='small'
 *
 * Errors: every entry point returns TCR_OK (0) or a negative code; the
 * message of the last failure on the calling thread is tcr_last_error().
 * No C++ exception crosses this boundary.
 */
#ifndef TCR_H
#define TCR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TCR_OK 0
#define TCR_ERR_ARG (-1)      /* invalid argument (shape, mode, dim, sizes)     */
#define TCR_ERR_HIP (-2)      /* a HIP runtime call failed                      */
#define TCR_ERR_DOMAIN (-3)   /* degenerate grid (e.g. zero cells in a periodic dim) */
#define TCR_ERR_SPACE (-4)    /* workspace or output smaller than required      */
#define TCR_ERR_UNSUPPORTED (-5) /* inputs the multi-level search does not take; the
                                    single-level search gives the same output    */

/* Support modes, same meaning as the reference's mode strings
 * (cpp/common.h:68-70, cpp/neighborhoodDynamic_cpu.cpp:77-88). */
#define TCR_MODE_SYMMETRIC 0
#define TCR_MODE_GATHER 1
#define TCR_MODE_SCATTER 2
#define TCR_MODE_SUPERSYMMETRIC 3

/* One point cloud pair of a (batched) search.  Positions are row-major
 * [n, dim] fp32; supports are [n] fp32 (may be absent when the mode does not
 * read them).  Offsets index the concatenated batch buffers passed to the
 * search calls: *_off count points, *_foff count floats. */
typedef struct tcr_cloud {
    int32_t dim;            /* 1, 2 or 3                                        */
    int32_t mode;           /* TCR_MODE_*                                       */
    uint32_t periodic_mask; /* bit d set: dimension d is periodic               */
    int32_t search_range;   /* cell stencil half-width, >= 1 (reference: 1)     */
    float domain_min[3];
    float domain_max[3];
    uint32_t hash_map_length;
    int32_t self_search;    /* 1: the query cloud is the reference cloud        */
    int64_t q_off, nq;      /* query points of this cloud in the batch          */
    int64_t r_off, nr;      /* reference points of this cloud in the batch      */
    int64_t q_foff, r_foff; /* float offsets of the position rows               */
} tcr_cloud;

/* Device bytes of workspace a search over these clouds needs; the same
 * workspace must be passed unchanged to tcr_search_count, _fill and _export. */
size_t tcr_search_workspace_size(const tcr_cloud *clouds, int32_t n_clouds);

/* Structure build (periodic wraps, hMax, hCell, cell keys, stable radix sort,
 * cell table, linear-index lookup table) and the neighbour count pass, then
 * the exclusive scan of the counts into offsets[nq_total + 1] (int64,
 * concatenated over the batch).  Synchronises the stream once at the end and
 * writes the total pair count of the batch to *total_pairs.  qsup/rsup may be
 * NULL when no cloud's mode reads them; for self-search clouds rpos/rsup may
 * be NULL. */
int tcr_search_count(const tcr_cloud *clouds, int32_t n_clouds,
                     const float *qpos, const float *qsup,
                     const float *rpos, const float *rsup,
                     void *workspace, size_t workspace_bytes,
                     int64_t *offsets, int64_t *total_pairs, void *stream);

/* Pair fill after tcr_search_count with the same arguments: row (may be NULL)
 * gets the query index, col the original reference index, both local to
 * their cloud, at offsets[q] .. offsets[q+1] for query q, in the reference's
 * order (cell stencil order, then sorted reference order).  index_bytes is
 * 8 (int64 entries, the reference's dtype) or 4 (int32 entries). */
int tcr_search_fill(const tcr_cloud *clouds, int32_t n_clouds,
                    const float *qpos, const float *qsup,
                    const float *rpos, const float *rsup,
                    void *workspace, size_t workspace_bytes,
                    const int64_t *offsets, void *row, void *col, int32_t index_bytes,
                    void *stream);

/* The whole search without a host round trip: what tcr_search_count
 * launches, then the fill of tcr_search_fill into row/col, which hold
 * `capacity` pairs; asynchronous (no synchronisation).  On the device the fill
 * is skipped when the total exceeds the capacity: read the total with
 * tcr_search_finish and, if it is larger, call tcr_search_fill with outputs of
 * that size.  Replaces the same reference chain as tcr_search_count + _fill
 * (py/compactHashing/cppSearch.py:179-244), whose host read of the pair total
 * (cppSearch.py:81-89) sits between the two passes. */
int tcr_search_launch(const tcr_cloud *clouds, int32_t n_clouds,
                      const float *qpos, const float *qsup,
                      const float *rpos, const float *rsup,
                      void *workspace, size_t workspace_bytes,
                      int64_t *offsets, void *row, void *col, int32_t index_bytes,
                      int64_t capacity, void *stream);

/* tcr_search_launch for the CSR format: no row indices; the reference's CSR
 * arrays, indptr int32[nq + 1] and rowEntries int64[nq] (either may be NULL;
 * py/neighborhood.py:303-305, py/util.py:335-358), are written from the
 * offsets before the fill.  Same completion (tcr_search_finish) and capacity
 * rule. */
int tcr_search_launch_csr(const tcr_cloud *clouds, int32_t n_clouds,
                          const float *qpos, const float *qsup,
                          const float *rpos, const float *rsup,
                          void *workspace, size_t workspace_bytes,
                          int64_t *offsets, void *col, int32_t index_bytes,
                          int64_t capacity, int32_t *indptr, int64_t *row_entries,
                          void *stream);

/* The launches of tcr_search_launch captured once into a HIP graph (on a
 * created stream, not the null stream) and replayed by tcr_graph_launch: the
 * replay reads the same input buffers, workspace and outputs, so a caller
 * that rewrites positions/supports in place between replays searches the new
 * values with one launch.  Complete each replay with tcr_search_finish.
 * *graph receives an opaque handle for tcr_graph_launch / _destroy.  No
 * counterpart in the reference (its chain is host-driven torch ops). */
int tcr_graph_capture(const tcr_cloud *clouds, int32_t n_clouds,
                      const float *qpos, const float *qsup,
                      const float *rpos, const float *rsup,
                      void *workspace, size_t workspace_bytes,
                      int64_t *offsets, void *row, void *col, int32_t index_bytes,
                      int64_t capacity, void **graph, void *stream);
int tcr_graph_launch(void *graph, void *stream);
int tcr_graph_destroy(void *graph);

/* Launch shape of the searches the calling thread launches (each host
 * thread has its own: threads that run differently tuned plans do not race);
 * 0 = automatic.  A captured graph keeps the shape it was captured with.
 * Picked per batch shape by the package's autotuner.  No counterpart in the
 * reference (its launch shapes are fixed). */
typedef struct tcr_tuning {
    int32_t cell_workgroups;     /* workgroups of the count/fill cell kernels (rounded up to a multiple of 8) */
    int32_t stencil_workgroups;  /* workgroups of the stencil resolution kernel */
    int32_t cell_threads;        /* threads per workgroup of the cell kernels: 64, 128 or 256 (0: 256) */
    int32_t cells_per_wave;      /* consecutive cells a wavefront takes in turn, 1..64 (0: 1) */
} tcr_tuning;
int tcr_set_tuning(const tcr_tuning *tuning);  /* NULL restores the defaults */
int tcr_get_tuning(tcr_tuning *tuning);

/* Completes tcr_search_launch: synchronises the stream once, reports a
 * degenerate grid (TCR_ERR_DOMAIN) and writes the total pair count. */
int tcr_search_finish(const tcr_cloud *clouds, int32_t n_clouds,
                      void *workspace, size_t workspace_bytes,
                      const int64_t *offsets, int64_t *total_pairs, void *stream);

/* Multi-level search of one cloud for variable supports: the same pairs in
 * the same order as tcr_search_count/_fill, with reference points binned
 * into up to max_levels (1..4) nested grids of cell size hCell * 2^-l by
 * their support.  Replaces the reference's algorithm='mlm' path
 * (py/neighborhood.py:293, py/multiLevelMemory/wrapper.py:145-308,
 * cpp/multiLevelMemory/countNeighborsHashmap.h:99-124,
 * buildNeighborhoodHashmap.h:103-128), whose output is defined as the compact
 * path's.  Returns TCR_ERR_UNSUPPORTED for inputs it does not take (gather
 * mode, search_range != 1, non-finite or negative supports, points outside
 * the grid, fewer than 3 cells in a dimension, grids beyond 2^29 cells): run
 * the single-level search then.  tcr_mlm_count synchronises twice and writes
 * offsets[nq + 1], the pair total and the level count to pass to
 * tcr_mlm_fill; scratch holds total_pairs uint64 and may be col itself when
 * index_bytes is 8.  With slots (uint32[nq * slot_cap], slot_cap <= 2048;
 * fewer than 2^27 reference points) the count pass also writes every row's
 * (stencil offset << 27 | index) keys into its query's slot and sets
 * *slots_complete when all rows fit; tcr_mlm_fill given those slots then
 * sorts the rows from there without a second traversal (scratch unused).
 * Otherwise pass slots = NULL to tcr_mlm_fill. */
size_t tcr_mlm_workspace_size(const tcr_cloud *cloud);
int tcr_mlm_count(const tcr_cloud *cloud, int32_t max_levels,
                  const float *qpos, const float *qsup, const float *rpos, const float *rsup,
                  void *workspace, size_t workspace_bytes, int64_t *offsets,
                  int64_t *total_pairs, int32_t *levels_used,
                  uint32_t *slots, int64_t slot_cap, int32_t *slots_complete, void *stream);
int tcr_mlm_fill(const tcr_cloud *cloud, int32_t levels,
                 const float *qpos, const float *qsup, const float *rpos, const float *rsup,
                 void *workspace, size_t workspace_bytes, const int64_t *offsets,
                 void *row, void *col, int32_t index_bytes, uint64_t *scratch,
                 const uint32_t *slots, int64_t slot_cap, void *stream);

/* algorithm='mlm' with the reference's own semantics (a separate mode; the
 * multi-level search above returns the compact result).  Replaces
 * searchNeighbors_mlm with the hash-map cell table (py/multiLevelMemory/
 * wrapper.py:145-308, cellOp.py:5-64, indexing.py:71-78, hashData.py:21-39,
 * cpp/multiLevelMemory/mlmUtil.h:234-282, countNeighborsHashmap.h:99-124,
 * buildNeighborhoodHashmap.h:103-128), quirks included: per-level codes
 * shifted by 2 bits per level in every dimension, C-remainder neighbour
 * cells with no periodic wrap, counts over the reference set at its
 * unwrapped positions, superSymmetric support synchronisation, int32
 * offsets of length nr.  Positions fp32 [n, dim] and supports fp32 [n] on
 * the device; domain arrays on the HOST.
 * 1. tcr_mlmref_plan reads the reference supports' range (synchronises),
 *    fills *grid (computeResolutionLevels) and the workspace size; scratch:
 *    >= 64 device bytes.  TCR_ERR_DOMAIN when no level exists (the
 *    reference fails there).
 * 2. tcr_mlmref_count builds the structure over the wrapped reference set,
 *    counts over the unwrapped reference set and writes counts int32[nr]
 *    (+ sync counts for superSymmetric), offsets int32[nr] (exclusive) and
 *    the pair total (synchronises).  Not for scatter: the reference runs it
 *    as the transposed gather search (tcr_mlmref_transpose).
 * 3. tcr_mlmref_fill with the same arguments and workspace writes int64
 *    (row, col) pairs, total entries (entries a row does not reach stay
 *    (0, sortingIndices[0]), as the reference's zero-initialised list);
 *    TCR_ERR_UNSUPPORTED where the reference's output is undefined (more
 *    queries than reference points, or a query with more pairs than its
 *    row). */
typedef struct tcr_mlmref_grid {
    int32_t dim, levels;
    float hcell, hfine, hmin, hmax;
    int32_t res[17][3];     /* cellResolutions per level */
} tcr_mlmref_grid;
int tcr_mlmref_plan(const float *rsup, int64_t nr, int64_t nq, int32_t dim, const float *domain_min,
                    const float *domain_max, void *scratch, tcr_mlmref_grid *grid, size_t *workspace_bytes,
                    void *stream);
int tcr_mlmref_count(const tcr_mlmref_grid *grid, const float *qpos, const float *qsup, int64_t nq,
                     const float *rpos, const float *rsup, int64_t nr, const float *domain_min,
                     const float *domain_max, uint32_t periodic_mask, int32_t mode, int64_t hash_map_length,
                     void *workspace, size_t workspace_bytes, int32_t *counts, int32_t *offsets, int64_t *total,
                     void *stream);
int tcr_mlmref_fill(const tcr_mlmref_grid *grid, const float *qpos, const float *qsup, int64_t nq,
                    const float *rpos, const float *rsup, int64_t nr, const float *domain_min,
                    const float *domain_max, uint32_t periodic_mask, int32_t mode, int64_t hash_map_length,
                    void *workspace, size_t workspace_bytes, const int32_t *offsets, int64_t total, int64_t *row,
                    int64_t *col, void *stream);
/* transposeCOO(sorted=True) (py/multiLevelMemory/wrapper.py:119-142): the
 * pairs (col, row) ordered stably by the new row; indptr int32[n_rows + 1]
 * and row_entries int64[n_rows] (either may be NULL) are coo_to_csr's
 * (py/util.py:335-358). */
size_t tcr_mlmref_transpose_workspace_size(int64_t n, int64_t n_rows);
int tcr_mlmref_transpose(const int64_t *row, const int64_t *col, int64_t n, int64_t n_rows, void *workspace,
                         size_t workspace_bytes, int64_t *out_row, int64_t *out_col, int32_t *indptr,
                         int64_t *row_entries, void *stream);

/* Copy one cloud's structure out of the workspace in the reference layout:
 * sort_index int64[nr], sorted_pos fp32[nr, dim], sorted_sup fp32[nr] (NULL to
 * skip), cell_table int32[ncells, 3] (lin, begin, count) ordered by hash
 * bucket, hash_table int32[L, 2] (begin, count | -1, 0), num_cells int32[dim]
 * and the scalars (device) hcell fp32[1], ncells int64[1]. */
int tcr_search_export(const tcr_cloud *clouds, int32_t n_clouds, int32_t cloud,
                      void *workspace, size_t workspace_bytes,
                      int64_t *sort_index, float *sorted_pos, float *sorted_sup,
                      int32_t *cell_table, int32_t *hash_table, int32_t *num_cells,
                      float *hcell, int64_t *ncells, void *stream);

/* computeHashIndices: hash[i] = hashIndexing(cells[i, :dim], L)
 * (cpp/hashing.h:20-35).  cells int32[n, dim], hash int32[n]. */
int tcr_compute_hash_indices(const int32_t *cells, int64_t n, int32_t dim,
                             int64_t hash_map_length, int32_t *hash, void *stream);

/* countNeighbors on the reference's own structure layout: hash_table
 * int32[L,2], cell_table int32[C,3], num_cells int32[dim], sorted positions
 * fp32[nr,dim], domain arrays fp32[dim], periodicity uint8[dim] (torch.bool).
 * counts int32[nq].  q_min is accepted and unused, as in the reference. */
int tcr_count_neighbors(const float *query_pos, const float *query_sup, int64_t nq,
                        int32_t search_range,
                        const float *sorted_pos, const float *sorted_sup, int64_t nr,
                        const int32_t *hash_table, int64_t hash_map_length,
                        const int32_t *num_cells, const int32_t *cell_table, int64_t n_cells,
                        const float *q_min, double h_cell,
                        const float *max_domain, const float *min_domain,
                        const uint8_t *periodicity, int32_t dim, int32_t mode,
                        int32_t *counts, void *stream);

/* buildNeighborList: writes (i, j_sorted) int64 pairs at offsets[i] (int32
 * offsets, as the reference), list_length entries in total. */
int tcr_build_neighbor_list(const int32_t *counts, const int32_t *offsets, int64_t list_length,
                            const float *query_pos, const float *query_sup, int64_t nq,
                            int32_t search_range,
                            const float *sorted_pos, const float *sorted_sup, int64_t nr,
                            const int32_t *hash_table, int64_t hash_map_length,
                            const int32_t *num_cells, const int32_t *cell_table, int64_t n_cells,
                            const float *q_min, double h_cell,
                            const float *max_domain, const float *min_domain,
                            const uint8_t *periodicity, int32_t dim, int32_t mode,
                            int64_t *list_i, int64_t *list_j, void *stream);

/* The reference's CSR arrays from int64 offsets[nq + 1] (stream-ordered, so
 * callable right after tcr_search_launch): indptr int32[nq + 1] and
 * rowEntries int64[nq] (either may be NULL) -- the SparseCSR of
 * py/neighborhood.py:303-305 / py/util.py:335-358. */
int tcr_csr_arrays(const int64_t *offsets, int64_t nq, int32_t *indptr, int64_t *row_entries, void *stream);

/* Brute-force search with the semantics of radiusSearch's default
 * algorithm='naive': radiusNaive / radiusNaiveFixed
 * (py/searchAlgorithms/radiusNaive.py:8-56, called at
 * py/neighborhood.py:226-229,268-271).  Every (query a, reference b) pair with
 * distance <= the mode's support (radiusNaive's own periodic `mod`, fp32
 * squares summed in dimension order, sqrtf; superSymmetric is
 * max(hx[a], hy[a]) and needs nx == ny), ordered by reference index, then
 * query index.  x fp32[nx, dim], y fp32[ny, dim], hx/hy fp32 (device; NULL
 * where the mode reads none); domain_min/max fp32[dim] on the HOST; bit d of
 * periodic_mask = dim d periodic; periodic coordinates are wrapped `wraps`
 * (0..2) times first (radiusSearch wraps once, radiusNaive once more).
 * fixed_support non-NULL selects radiusNaiveFixed: squared distance <=
 * (*fixed_support)^2, supports unused.
 * tcr_naive_count writes offsets int64[ny + 1] (first slot per reference
 * point, offsets[ny] = pair count); tcr_naive_fill, with the same inputs and
 * those offsets, writes ii (query) and jj (reference) indices, index_bytes 4
 * or 8 each.  Workspace: tcr_naive_workspace_size bytes (device). */
size_t tcr_naive_workspace_size(int64_t nx, int64_t ny, int32_t dim);
int tcr_naive_count(const float *x, const float *hx, int64_t nx, const float *y, const float *hy, int64_t ny,
                    int32_t dim, const float *domain_min, const float *domain_max, uint32_t periodic_mask,
                    int32_t wraps, int32_t mode, const double *fixed_support, void *workspace,
                    size_t workspace_bytes, int64_t *offsets, void *stream);
int tcr_naive_fill(const float *x, const float *hx, int64_t nx, const float *y, const float *hy, int64_t ny,
                   int32_t dim, const float *domain_min, const float *domain_max, uint32_t periodic_mask,
                   int32_t wraps, int32_t mode, const double *fixed_support, void *workspace,
                   size_t workspace_bytes, const int64_t *offsets, void *ii, void *jj, int32_t index_bytes,
                   void *stream);

/* Brute-force search with the semantics of algorithm='small', the op the
 * reference keeps commented out (cpp/neighborhoodSmall_cpu.cpp:45-302; its
 * Python binding is None at py/compactHashing/cppWrapper.py:87-88, so the
 * reference raises): for each query i, every reference j in index order with
 * modDistance (cpp/common.h:106-129) < the mode's support (superSymmetric
 * std::max(h_i, h_j)); fixed_support non-NULL: modDistance2 < h*h with h*h in
 * fp32.  Same arguments and workspace as tcr_naive_*, except that offsets is
 * int64[nx + 1] (first slot per query) and the pairs come in (query,
 * reference) order; radiusSearch passes positions wrapped once (wraps = 1). */
int tcr_small_count(const float *x, const float *hx, int64_t nx, const float *y, const float *hy, int64_t ny,
                    int32_t dim, const float *domain_min, const float *domain_max, uint32_t periodic_mask,
                    int32_t wraps, int32_t mode, const double *fixed_support, void *workspace,
                    size_t workspace_bytes, int64_t *offsets, void *stream);
int tcr_small_fill(const float *x, const float *hx, int64_t nx, const float *y, const float *hy, int64_t ny,
                   int32_t dim, const float *domain_min, const float *domain_max, uint32_t periodic_mask,
                   int32_t wraps, int32_t mode, const double *fixed_support, void *workspace,
                   size_t workspace_bytes, const int64_t *offsets, void *ii, void *jj, int32_t index_bytes,
                   void *stream);

/* The search of tcr_search_launch + tcr_search_finish without allocating:
 * no counterpart in the reference; for callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc).  The
 * workspace is a 64 MiB device array inside the library, one per GPU, held
 * until the process ends (enough for about 140,000 points of a 3-D
 * self-search); a search whose tcr_search_workspace_size exceeds it is
 * rejected with TCR_ERR_SPACE.  Synchronises the stream once: writes the
 * pair total to *total_pairs and, when it fits in `capacity`, the pairs as
 * tcr_search_launch does (else TCR_ERR_SPACE, outputs not filled: call again
 * with larger outputs).  Every call on a GPU shares that GPU's array: calls
 * that overlap on one GPU (other threads or streams) race. */
int tcr_search_noalloc(const tcr_cloud *clouds, int32_t n_clouds,
                       const float *qpos, const float *qsup,
                       const float *rpos, const float *rsup,
                       int64_t *offsets, void *row, void *col, int32_t index_bytes,
                       int64_t capacity, int64_t *total_pairs, void *stream);

/* Phase timing of tcr_search_launch / tcr_search_count on the calling thread
 * (no counterpart in the reference; a benchmark hook).  While enabled, each
 * launch records HIP events on its stream that delimit the count phase
 * (build, count, scan, up to the fill kernel), the first wavefront-per-cell
 * count kernel, and the fill pass's main kernel (the wavefront-per-cell
 * fill, else the per-query fill).  tcr_timing_collect waits for the recorded
 * events, writes up to max_entries milliseconds per launch into each
 * non-NULL array (-1 for a phase or kernel that did not run), the number
 * written to *n_entries, and forgets them. */
int tcr_timing_enable(int32_t on);
int tcr_timing_collect(float *count_ms, float *fill_ms, float *count_kernel_ms, int32_t max_entries,
                       int32_t *n_entries);

/* Message of the last failed call on this thread ("" if none). */
const char *tcr_last_error(void);

/* Library version string. */
const char *tcr_version(void);

#ifdef __cplusplus
}
#endif
#endif /* TCR_H */
