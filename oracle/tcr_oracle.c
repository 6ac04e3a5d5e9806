/*
 * tcr_oracle.c -- CPU restatement of torchCompactRadius 0.5.5's compact-hashing
 * radius search (radiusSearch(..., algorithm='compact')).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * library under torchcompactradius_amd/csrc and the "port" CPU baseline that
 * bench.py times.  Nothing in the product path links, loads or calls it; only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg do.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository; "cpp/" = src/torchCompactRadius/cppSrc/,
 * "py/" = src/torchCompactRadius/).  Arithmetic is kept IEEE-literal:
 *   - build with -O2 -ffp-contract=off and no -ffast-math / -march flags,
 *     like the reference CPU extension (x86-64 baseline, no FMA contraction);
 *   - fp32 where the reference computes in fp32 (torch fp32 tensor ops,
 *     float accessors), fp64 where C++ promotion makes it fp64
 *     (cpp/common.h:106-109 moduloOp, :121-129 modDistance).
 *
 * Pinning (what checks each function, see tests/test_oracle.py):
 *   orc_wrap          - torch.remainder on >1e6 values incl. boundary cases
 *   orc_compute_h     - torch fp32 replay of py/util.py:28-79 on many domains
 *   orc_hash_index    - hand-computed uint32 vectors
 *   orc_offsets       - the reference's own offset loop replayed in Python
 *   orc_search_*      - the reference's known answers (scripts/test.py:89-115:
 *                       41580 / 46080 / 11520 pairs, 15..45 / 45 per point),
 *                       an O(N^2) brute force, and closed-form lattice counts.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { ORC_SYMMETRIC = 0, ORC_GATHER = 1, ORC_SCATTER = 2, ORC_SUPERSYMMETRIC = 3 };

/* ---------------------------------------------------------------- wraps --- */

/* torch.remainder(a, b) for fp32 on CPU/CUDA: fmod, then shift into the sign
 * of b.  Used by py/util.py:387 (getPeriodicPointCloud) and
 * py/compactHashing/cppSearch.py:221 (second wrap of the reference set):
 *     remainder(p - min, max - min) + min                                     */
static inline float orc_remainder(float a, float b) {
    float r = fmodf(a, b);
    if (r != 0.0f && ((r < 0.0f) != (b < 0.0f))) r += b;
    return r;
}

/* One periodic wrap of every periodic dim, in place (py/util.py:375-387). */
void orc_wrap(float *pos, int64_t n, int dim, const float *mn, const float *mx,
              uint32_t periodic_mask) {
    for (int d = 0; d < dim; ++d) {
        if (!((periodic_mask >> d) & 1u)) continue;
        const float lo = mn[d];
        const float ext = mx[d] - mn[d];
        for (int64_t i = 0; i < n; ++i) {
            float *p = pos + i * dim + d;
            *p = orc_remainder(*p - lo, ext) + lo;
        }
    }
}

/* ------------------------------------------------------------- grid size --- */

/* compute_h (py/util.py:28-79) in fp32 tensor arithmetic; returns hCell.
 * numCells is decremented in every dim if any dim overshoots (:60-62), the
 * 1.0001 stretch is an fp32 multiply by (float)1.0001 (:68-69; torch casts
 * the Python scalar to the tensor dtype), and hCell = max over dims (:79). */
float orc_compute_h(const float *mn, const float *mx, float hmax, int dim) {
    float ext[3] = {0}, nc[3] = {0}, h[3] = {0};
    int over = 0;
    for (int d = 0; d < dim; ++d) {
        ext[d] = mx[d] - mn[d];
        nc[d] = floorf(ext[d] / hmax);
        h[d] = ext[d] / nc[d];
    }
    for (int d = 0; d < dim; ++d)
        if (ext[d] / h[d] - nc[d] > 0.0f) over = 1;
    if (over)
        for (int d = 0; d < dim; ++d) {
            nc[d] = nc[d] - 1.0f;
            h[d] = ext[d] / nc[d];
        }
    int stretch = 0;
    for (int d = 0; d < dim; ++d)
        if (ceilf(ext[d] / h[d]) > ext[d] / h[d]) stretch = 1;
    if (stretch)
        for (int d = 0; d < dim; ++d) h[d] = h[d] * 1.0001f;
    float hc = h[0];
    for (int d = 1; d < dim; ++d) hc = (h[d] > hc || isnan(h[d])) ? h[d] : hc;
    return hc;
}

/* cellCount = ceil(qExtent / hCell).int32 (py/compactHashing/cellTable.py:29). */
void orc_num_cells(const float *mn, const float *mx, float hcell, int dim, int32_t *nc) {
    for (int d = 0; d < dim; ++d) nc[d] = (int32_t)ceilf((mx[d] - mn[d]) / hcell);
}

/* floor((x - min) / hCell).int32 (cellTable.py:30, cpp/neighborhoodDynamic.h:52). */
static inline int32_t orc_cell_coord(float x, float lo, float hcell) {
    return (int32_t)floorf((x - lo) / hcell);
}

/* ------------------------------------------------------- index functions --- */

/* Row-major int32 linear index with wrap-around (py/util.py:176-193,
 * cpp/neighborhood.h:13-23).  Computed in uint32 to get defined wrap. */
int32_t orc_linear_index(const int32_t *c, const int32_t *nc, int dim) {
    uint32_t lin = 0, prod = 1;
    for (int d = 0; d < dim; ++d) {
        lin += (uint32_t)c[d] * prod;
        prod *= (uint32_t)nc[d];
    }
    return (int32_t)lin;
}

/* Spatial hash (cpp/hashing.h:20-35): uint32 sum of coord * prime, mod L. */
int32_t orc_hash_index(const int32_t *c, int dim, uint32_t L) {
    if (dim == 1) return (int32_t)(((uint32_t)c[0]) % L);
    static const uint32_t primes[3] = {73856093u, 19349663u, 83492791u};
    uint32_t h = 0;
    for (int d = 0; d < dim; ++d) h += ((uint32_t)c[d]) * primes[d];
    return (int32_t)(h % L);
}

static inline int32_t orc_ipow(int32_t b, int32_t e) {
    int32_t r = 1;
    for (int32_t i = 0; i < e; ++i) r *= b;
    return r;
}

/* Cell offset table (cpp/neighborhoodDynamic_cpu.cpp:153-178, py/util.py:249-272):
 * lexicographic over [-sr, sr]^dim with dim 0 slowest, dim-1 fastest.
 * out must hold (2sr+1)^dim * dim int32.  Restated as the closed form of the
 * reference's counter loop. */
int32_t orc_offsets(int32_t sr, int dim, int32_t *out) {
    const int32_t w = 2 * sr + 1;
    const int32_t n = orc_ipow(w, dim);
    for (int32_t o = 0; o < n; ++o) {
        int32_t rem = o;
        for (int d = dim - 1; d >= 0; --d) {
            out[o * dim + d] = (rem % w) - sr;
            rem /= w;
        }
    }
    return n;
}

static inline int32_t orc_pymod(int32_t n, int32_t m) { /* cpp/common.h:95-97 */
    return n >= 0 ? n % m : ((n % m) + m) % m;
}

/* ------------------------------------------------------------- distance --- */

/* modDistance2 (cpp/common.h:130-138) for float inputs: the fp32 difference,
 * fp64 periodic modulo, fp64 square added to the fp32 accumulator in fp64 and
 * rounded back to fp32 each step. */
static inline float orc_distance2(const float *xi, const float *xj, const float *mn,
                                  const float *mx, uint32_t pmask, int dim) {
    float sum = 0.0f;
    for (int d = 0; d < dim; ++d) {
        double diff;
        const float pq = xi[d] - xj[d];
        if ((pmask >> d) & 1u) {
            const float hf = mx[d] - mn[d];
            const double h = (double)hf;
            const double t = (double)pq + h / 2.0;
            diff = (t - floor(t / h) * h) - h / 2.0;
        } else {
            diff = (double)pq;
        }
        sum = (float)((double)sum + diff * diff);
    }
    return sum;
}

/* modDistance (cpp/common.h:106-129): sqrtf of modDistance2. */
static inline float orc_distance(const float *xi, const float *xj, const float *mn,
                                 const float *mx, uint32_t pmask, int dim) {
    return sqrtf(orc_distance2(xi, xj, mn, mx, pmask, dim));
}

float orc_pair_distance(const float *xi, const float *xj, const float *mn, const float *mx,
                        uint32_t pmask, int dim) {
    return orc_distance(xi, xj, mn, mx, pmask, dim);
}

/* Support predicate (cpp/neighborhoodDynamic.h:63-70, :105-108), strict '<'. */
static inline int orc_pred(float dist, float hi, float hj, int mode) {
    switch (mode) {
    case ORC_SCATTER: return dist < hj;
    case ORC_GATHER: return dist < hi;
    case ORC_SYMMETRIC: return dist < (hi + hj) / 2.f;
    default: return dist < ((hi < hj) ? hj : hi); /* std::max(hi, hj) */
    }
}

/* -------------------------------------------------------- stable sorting --- */

/* Stable argsort of int32 keys (signed order), the contract of
 * torch.argsort on the cell keys (py/compactHashing/cellTable.py:33) for the
 * reference's CUDA path and its CPU radix path (n >= 32768).  LSD radix over
 * the sign-flipped key, 4 passes of 8 bits. */
static void orc_stable_argsort_i32(const int32_t *key, int64_t n, int32_t *idx) {
    uint32_t *ka = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
    uint32_t *kb = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
    int32_t *ib = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
    for (int64_t i = 0; i < n; ++i) {
        ka[i] = ((uint32_t)key[i]) ^ 0x80000000u;
        idx[i] = (int32_t)i;
    }
    uint32_t *ks = ka, *kd = kb;
    int32_t *is = idx, *id = ib;
    for (int pass = 0; pass < 4; ++pass) {
        const int sh = pass * 8;
        int64_t cnt[257];
        memset(cnt, 0, sizeof(cnt));
        for (int64_t i = 0; i < n; ++i) cnt[((ks[i] >> sh) & 255u) + 1]++;
        for (int b = 0; b < 256; ++b) cnt[b + 1] += cnt[b];
        for (int64_t i = 0; i < n; ++i) {
            const int64_t p = cnt[(ks[i] >> sh) & 255u]++;
            kd[p] = ks[i];
            id[p] = is[i];
        }
        uint32_t *tk = ks; ks = kd; kd = tk;
        int32_t *ti = is; is = id; id = ti;
    }
    /* 4 passes: result back in the original buffers (ka / idx). */
    free(ka); free(kb); free(ib);
}

/* --------------------------------------------------- compact hash map ---- */

typedef struct orc_map {
    int dim;
    int64_t nr;
    uint32_t pmask;
    float mn[3], mx[3];
    float hcell;
    int32_t nc[3];
    uint32_t L;
    int32_t *sort_index;   /* [nr]   sorted position -> original index      */
    float *sorted_pos;     /* [nr,dim] twice-wrapped reference positions    */
    float *sorted_sup;     /* [nr] or NULL                                  */
    int64_t ncells;
    int32_t *cell_table;   /* [ncells,3] (lin, begin, count), hash order    */
    int32_t *hash_table;   /* [L,2] (begin, count) into cell_table, -1/0    */
} orc_map;

void orc_map_free(orc_map *m) {
    if (!m) return;
    free(m->sort_index); free(m->sorted_pos); free(m->sorted_sup);
    free(m->cell_table); free(m->hash_table);
    free(m);
}

/* buildCompactHashMap_compat (py/compactHashing/hashTable.py:79-154) preceded
 * by the reference-set wrap of neighborSearch_cpp (cppSearch.py:221).
 * rpos holds the positions as radiusSearch hands them on (i.e. already wrapped
 * once by py/neighborhood.py:203); they are wrapped once more here. */
orc_map *orc_map_build(const float *rpos, const float *rsup, int64_t nr, int dim,
                       const float *mn, const float *mx, uint32_t pmask, float hmax,
                       uint32_t L) {
    orc_map *m = (orc_map *)calloc(1, sizeof(orc_map));
    m->dim = dim; m->nr = nr; m->pmask = pmask; m->L = L;
    for (int d = 0; d < dim; ++d) { m->mn[d] = mn[d]; m->mx[d] = mx[d]; }
    size_t n1 = (size_t)(nr ? nr : 1);
    float *x = (float *)malloc(sizeof(float) * n1 * dim);
    memcpy(x, rpos, sizeof(float) * (size_t)nr * dim);
    orc_wrap(x, nr, dim, mn, mx, pmask);                      /* cppSearch.py:221 */
    m->hcell = orc_compute_h(mn, mx, hmax, dim);              /* cellTable.py:27 */
    orc_num_cells(mn, mx, m->hcell, dim, m->nc);              /* cellTable.py:29 */
    int32_t *lin = (int32_t *)malloc(sizeof(int32_t) * n1);
    for (int64_t i = 0; i < nr; ++i) {                        /* cellTable.py:30-31 */
        int32_t c[3];
        for (int d = 0; d < dim; ++d) c[d] = orc_cell_coord(x[i * dim + d], mn[d], m->hcell);
        lin[i] = orc_linear_index(c, m->nc, dim);
    }
    m->sort_index = (int32_t *)malloc(sizeof(int32_t) * n1);
    orc_stable_argsort_i32(lin, nr, m->sort_index);           /* cellTable.py:33 */
    m->sorted_pos = (float *)malloc(sizeof(float) * n1 * dim);
    for (int64_t i = 0; i < nr; ++i)                          /* hashTable.py:109 */
        memcpy(m->sorted_pos + i * dim, x + (int64_t)m->sort_index[i] * dim, sizeof(float) * dim);
    if (rsup) {                                               /* cppSearch.py:224 */
        m->sorted_sup = (float *)malloc(sizeof(float) * n1);
        for (int64_t i = 0; i < nr; ++i) m->sorted_sup[i] = rsup[m->sort_index[i]];
    }
    /* unique_consecutive over sorted keys (hashTable.py:113-117) */
    int32_t *cells = (int32_t *)malloc(sizeof(int32_t) * 3 * n1);
    int64_t nc = 0;
    for (int64_t i = 0; i < nr; ++i) {
        const int32_t k = lin[m->sort_index[i]];
        if (i == 0 || k != cells[(nc - 1) * 3]) {
            cells[nc * 3 + 0] = k;
            cells[nc * 3 + 1] = (int32_t)i;
            cells[nc * 3 + 2] = 0;
            ++nc;
        }
        cells[(nc - 1) * 3 + 2]++;
    }
    /* hash of the first point's cell coordinates (hashTable.py:121-136) */
    int32_t *hash = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nc ? nc : 1));
    for (int64_t c = 0; c < nc; ++c) {
        const float *p = m->sorted_pos + (int64_t)cells[c * 3 + 1] * dim;
        int32_t cc[3];
        for (int d = 0; d < dim; ++d) cc[d] = orc_cell_coord(p[d], mn[d], m->hcell);
        hash[c] = orc_hash_index(cc, dim, L);
    }
    /* cells ordered by hash bucket (hashTable.py:138-143); inside a bucket the
     * reference's argsort order is unspecified -- lookups match on the unique
     * linear index, so cell order inside a bucket never changes a result.
     * Counting sort by bucket, stable in cell order. */
    m->hash_table = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)L);
    int64_t *bcnt = (int64_t *)calloc((size_t)L + 1, sizeof(int64_t));
    for (int64_t c = 0; c < nc; ++c) bcnt[hash[c] + 1]++;
    for (uint32_t b = 0; b < L; ++b) bcnt[b + 1] += bcnt[b];
    m->cell_table = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)(nc ? nc : 1));
    for (uint32_t b = 0; b < L; ++b) {                        /* hashTable.py:147-152 */
        const int64_t beg = bcnt[b], cnt = bcnt[b + 1] - bcnt[b];
        m->hash_table[2 * b + 0] = cnt ? (int32_t)beg : -1;
        m->hash_table[2 * b + 1] = (int32_t)cnt;
    }
    int64_t *fillp = (int64_t *)malloc(sizeof(int64_t) * ((size_t)L + 1));
    memcpy(fillp, bcnt, sizeof(int64_t) * ((size_t)L + 1));
    for (int64_t c = 0; c < nc; ++c) {
        const int64_t p = fillp[hash[c]]++;
        memcpy(m->cell_table + p * 3, cells + c * 3, sizeof(int32_t) * 3);
    }
    m->ncells = nc;
    free(fillp); free(bcnt); free(hash); free(cells); free(lin); free(x);
    return m;
}

/* queryHashMap (cpp/neighborhood.h:35-58): bucket scan matching the linear
 * index only; first hit wins. */
static inline int orc_query(const orc_map *m, const int32_t *cell, int32_t *beg, int32_t *end) {
    const int32_t lin = orc_linear_index(cell, m->nc, m->dim);
    const int32_t h = orc_hash_index(cell, m->dim, m->L);
    const int32_t hb = m->hash_table[2 * h], hl = m->hash_table[2 * h + 1];
    if (hb == -1) return 0;
    for (int32_t i = hb; i < hb + hl; ++i) {
        if (m->cell_table[i * 3] == lin) {
            *beg = m->cell_table[i * 3 + 1];
            *end = *beg + m->cell_table[i * 3 + 2];
            return 1;
        }
    }
    return 0;
}

/* countNeighborsForParticle / buildNeighborhood (cpp/neighborhoodDynamic.h:37-113)
 * with iterateOffsetCells (cpp/neighborhood.h:73-94).  If out_j is NULL the
 * hits are only counted; otherwise sorted-space j are written in order. */
static int64_t orc_query_point(const orc_map *m, const float *xi, float hi, int mode,
                               const int32_t *offs, int32_t noff, int32_t *out_j) {
    const int dim = m->dim;
    int32_t q[3], c[3];
    for (int d = 0; d < dim; ++d) q[d] = orc_cell_coord(xi[d], m->mn[d], m->hcell);
    int64_t k = 0;
    for (int32_t o = 0; o < noff; ++o) {
        for (int d = 0; d < dim; ++d)
            c[d] = ((m->pmask >> d) & 1u) ? orc_pymod(q[d] + offs[o * dim + d], m->nc[d])
                                          : q[d] + offs[o * dim + d];
        int32_t b, e;
        if (!orc_query(m, c, &b, &e)) continue;
        for (int32_t j = b; j < e; ++j) {
            const float dist = orc_distance(xi, m->sorted_pos + (int64_t)j * dim, m->mn, m->mx,
                                            m->pmask, dim);
            const float hj = m->sorted_sup ? m->sorted_sup[j] : 0.0f;
            if (orc_pred(dist, hi, hj, mode)) {
                if (out_j) out_j[k] = j;
                ++k;
            }
        }
    }
    return k;
}

/* ---------------------------------------------------------- full search --- */

typedef struct orc_search {
    orc_map *map;
    int64_t nq;
    float *qpos;          /* once-wrapped query positions */
    float *qsup;
    int mode;
    int32_t noff;
    int32_t *offs;
    int32_t *counts;      /* [nq] */
    int64_t *offsets;     /* [nq+1] exclusive scan, int64 */
    int64_t total;
} orc_search;

static float orc_maxf(const float *v, int64_t n) {
    /* torch.max over a 1-D fp32 tensor; NaN propagates. */
    float r = -INFINITY;
    for (int64_t i = 0; i < n; ++i) {
        if (isnan(v[i])) return v[i];
        if (v[i] > r) r = v[i];
    }
    return r;
}

/* computeGridSupport (py/compactHashing/cellTable.py:40-81). */
float orc_grid_support(const float *qsup, int64_t nq, const float *rsup, int64_t nr, int mode) {
    if (mode == ORC_GATHER) return orc_maxf(qsup, nq);
    if (mode == ORC_SCATTER) return orc_maxf(rsup, nr);
    const float a = orc_maxf(rsup, nr), b = orc_maxf(qsup, nq);
    if (isnan(a) || isnan(b)) return isnan(a) ? a : b;
    return a > b ? a : b;
}

/* radiusSearch(..., algorithm='compact') up to the neighbour counts:
 * py/neighborhood.py:136-309 (once-wrap :202-203), compactSearch.py:12-171,
 * cppSearch.py:179-244 (hMax :215, 2nd wrap :221, build :223) and the count
 * pass + exclusive scan of cppSearch.py:74-89.  rpos == NULL means
 * self-search (reference = query cloud).  Returns NULL on bad arguments. */
orc_search *orc_search_create(const float *qpos, const float *qsup, int64_t nq,
                              const float *rpos, const float *rsup, int64_t nr, int dim,
                              const float *mn, const float *mx, uint32_t pmask, int mode,
                              uint32_t L, int32_t sr, int nthreads) {
    if (dim < 1 || dim > 3 || L == 0 || sr < 1) return NULL;
    if (!rpos) { rpos = qpos; rsup = qsup; nr = nq; }
    if ((mode != ORC_SCATTER && !qsup) || (mode != ORC_GATHER && !rsup)) return NULL;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    orc_search *s = (orc_search *)calloc(1, sizeof(orc_search));
    s->nq = nq; s->mode = mode;
    size_t q1 = (size_t)(nq ? nq : 1), r1 = (size_t)(nr ? nr : 1);
    s->qpos = (float *)malloc(sizeof(float) * q1 * dim);
    memcpy(s->qpos, qpos, sizeof(float) * (size_t)nq * dim);
    orc_wrap(s->qpos, nq, dim, mn, mx, pmask);                 /* neighborhood.py:202 */
    float *r = (float *)malloc(sizeof(float) * r1 * dim);
    memcpy(r, rpos, sizeof(float) * (size_t)nr * dim);
    orc_wrap(r, nr, dim, mn, mx, pmask);                       /* neighborhood.py:203 */
    if (qsup) {
        s->qsup = (float *)malloc(sizeof(float) * q1);
        memcpy(s->qsup, qsup, sizeof(float) * (size_t)nq);
    }
    const float hmax = orc_grid_support(qsup, nq, rsup, nr, mode);
    s->map = orc_map_build(r, rsup, nr, dim, mn, mx, pmask, hmax, L);
    free(r);
    const int32_t w = 2 * sr + 1;
    s->offs = (int32_t *)malloc(sizeof(int32_t) * (size_t)orc_ipow(w, dim) * dim);
    s->noff = orc_offsets(sr, dim, s->offs);
    s->counts = (int32_t *)malloc(sizeof(int32_t) * q1);
    s->offsets = (int64_t *)malloc(sizeof(int64_t) * (q1 + 1));
    const orc_map *m = s->map;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < nq; ++i) {
        const float hi = s->qsup ? s->qsup[i] : 0.0f;
        s->counts[i] = (int32_t)orc_query_point(m, s->qpos + i * dim, hi, mode, s->offs, s->noff, NULL);
    }
    s->offsets[0] = 0;
    for (int64_t i = 0; i < nq; ++i) s->offsets[i + 1] = s->offsets[i] + s->counts[i];
    s->total = s->offsets[nq];
    return s;
}

int64_t orc_search_total(const orc_search *s) { return s->total; }
const int32_t *orc_search_counts(const orc_search *s) { return s->counts; }
const orc_map *orc_search_map(const orc_search *s) { return s->map; }
float orc_search_hcell(const orc_search *s) { return s->map->hcell; }
int64_t orc_search_ncells(const orc_search *s) { return s->map->ncells; }

/* buildNeighborList pass + the sortIndex remap (cppSearch.py:88-97):
 * row[k] = i, col[k] = sortIndex[j_sorted].  row may be NULL. */
void orc_search_fill(const orc_search *s, int64_t *row, int64_t *col) {
    const orc_map *m = s->map;
    const int dim = m->dim;
#pragma omp parallel
    {
        int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m->nr ? m->nr : 1));
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < s->nq; ++i) {
            const float hi = s->qsup ? s->qsup[i] : 0.0f;
            const int64_t k = orc_query_point(m, s->qpos + i * dim, hi, s->mode, s->offs, s->noff, tmp);
            const int64_t base = s->offsets[i];
            for (int64_t t = 0; t < k; ++t) {
                if (row) row[base + t] = i;
                col[base + t] = m->sort_index[tmp[t]];
            }
        }
        free(tmp);
    }
}

/* Copies of the structure (returnStructure / CompactHashMap parity). */
void orc_map_export(const orc_map *m, int32_t *sort_index, float *sorted_pos,
                    int32_t *cell_table, int32_t *hash_table, int32_t *num_cells) {
    if (sort_index) memcpy(sort_index, m->sort_index, sizeof(int32_t) * (size_t)m->nr);
    if (sorted_pos) memcpy(sorted_pos, m->sorted_pos, sizeof(float) * (size_t)m->nr * m->dim);
    if (cell_table) memcpy(cell_table, m->cell_table, sizeof(int32_t) * 3 * (size_t)m->ncells);
    if (hash_table) memcpy(hash_table, m->hash_table, sizeof(int32_t) * 2 * (size_t)m->L);
    if (num_cells) memcpy(num_cells, m->nc, sizeof(int32_t) * (size_t)m->dim);
}

void orc_search_free(orc_search *s) {
    if (!s) return;
    orc_map_free(s->map);
    free(s->qpos); free(s->qsup); free(s->offs); free(s->counts); free(s->offsets);
    free(s);
}

/* O(Nq * Nr) brute force with the same distance and predicate, over the
 * once-wrapped query set and twice-wrapped reference set; counts only.
 * Independent of the cell grid and hash table: it checks traversal
 * completeness wherever the 3^d cells cover the support (cell >= hMax). */
void orc_brute_counts(const float *qpos, const float *qsup, int64_t nq, const float *rpos,
                      const float *rsup, int64_t nr, int dim, const float *mn, const float *mx,
                      uint32_t pmask, int mode, int32_t *counts) {
    float *q = (float *)malloc(sizeof(float) * (size_t)(nq ? nq : 1) * dim);
    float *r = (float *)malloc(sizeof(float) * (size_t)(nr ? nr : 1) * dim);
    memcpy(q, qpos, sizeof(float) * (size_t)nq * dim);
    memcpy(r, rpos, sizeof(float) * (size_t)nr * dim);
    orc_wrap(q, nq, dim, mn, mx, pmask);
    orc_wrap(r, nr, dim, mn, mx, pmask);
    orc_wrap(r, nr, dim, mn, mx, pmask);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nq; ++i) {
        int32_t k = 0;
        for (int64_t j = 0; j < nr; ++j) {
            const float dist = orc_distance(q + i * dim, r + j * dim, mn, mx, pmask, dim);
            if (orc_pred(dist, qsup ? qsup[i] : 0.0f, rsup ? rsup[j] : 0.0f, mode)) ++k;
        }
        counts[i] = k;
    }
    free(q); free(r);
}

/* --------------------------------------------------- small brute force --- */

/* neighborSearchSmall / neighborSearchSmallFixed, the brute-force op the
 * reference keeps commented out (cpp/neighborhoodSmall_cpu.cpp:45-160,
 * 162-302): for each query i, every reference j in index order with
 * modDistance < the mode's support (fixed: modDistance2 < h*h, h*h in fp32).
 * Positions as given (radiusSearch has wrapped them once).  offsets == NULL:
 * counts[nq] only; else pairs at offsets[i] into pi (query) / pj (reference). */
void orc_small(const float *q, const float *hq, int64_t nq, const float *r, const float *hr, int64_t nr,
               int dim, const float *mn, const float *mx, uint32_t pmask, int mode, int fixed, float h,
               int32_t *counts, const int64_t *offsets, int64_t *pi, int64_t *pj) {
    const float h2 = h * h;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nq; ++i) {
        int64_t k = offsets ? offsets[i] : 0;
        int32_t c = 0;
        for (int64_t j = 0; j < nr; ++j) {
            int hit;
            if (fixed) {
                hit = orc_distance2(q + i * dim, r + j * dim, mn, mx, pmask, dim) < h2;
            } else {
                const float dist = orc_distance(q + i * dim, r + j * dim, mn, mx, pmask, dim);
                hit = orc_pred(dist, hq ? hq[i] : 0.0f, hr ? hr[j] : 0.0f, mode);
            }
            if (!hit) continue;
            if (offsets) {
                pi[k] = i;
                pj[k] = j;
                ++k;
            }
            ++c;
        }
        if (counts) counts[i] = c;
    }
}
