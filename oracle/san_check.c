/*
 * san_check.c -- runs the CPU oracle (tcr_oracle.c, compiled into this
 * translation unit) under AddressSanitizer and UndefinedBehaviorSanitizer.
 *
 * TEST INFRASTRUCTURE ONLY (built by `make -C oracle sanitize`, run by
 * tests/test_oracle_sanitize.py).  It drives every entry point of the oracle
 * over the shapes that stress its indexing: 1-3 dimensions, all four modes,
 * periodic and non-periodic dimensions, hash-map lengths that put many cells
 * in one bucket (1, 7) and a roomy one, points outside the domain and exactly
 * on its bounds, empty query and reference sets, non-self searches, search
 * ranges 1 and 2, and the brute-force "small" search.  Besides the
 * sanitizers' own checks it compares the hashed counts with the O(N^2) brute
 * force and the filled rows with the counts, so a run that is clean but wrong
 * also fails.  Exit status 0 = every case clean and consistent.
 */
#include "tcr_oracle.c"

#include <stdio.h>

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;

static double urand(void) { /* xorshift64*, uniform in [0, 1) */
    rng_state ^= rng_state >> 12;
    rng_state ^= rng_state << 25;
    rng_state ^= rng_state >> 27;
    return (double)((rng_state * 2685821657736338717ull) >> 11) * (1.0 / 9007199254740992.0);
}

static int failures = 0;
static int brute_checked = 0;   /* cases whose counts met the brute force */

#define CHECK(cond, ...)                                                         \
    do {                                                                         \
        if (!(cond)) {                                                           \
            fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__);                 \
            fprintf(stderr, __VA_ARGS__);                                        \
            fputc('\n', stderr);                                                 \
            ++failures;                                                          \
        }                                                                        \
    } while (0)

/* points in [lo, hi)^dim; the first ones placed on the domain's edges: on
 * its minimum and one ulp below its maximum (inside), and with outside != 0
 * also on its maximum and one ulp beyond either bound */
static float *cloud(int64_t n, int dim, float lo, float hi, const float *mn, const float *mx, int outside) {
    float *p = (float *)malloc(sizeof(float) * (size_t)(n ? n : 1) * dim);
    for (int64_t i = 0; i < n * dim; ++i) p[i] = (float)(lo + (hi - lo) * urand());
    if (n > 5) {
        for (int d = 0; d < dim; ++d) {
            p[0 * dim + d] = mn[d];
            p[1 * dim + d] = nextafterf(mx[d], -INFINITY);
            if (outside) {
                p[2 * dim + d] = mx[d];
                p[3 * dim + d] = nextafterf(mn[d], -INFINITY);
                p[4 * dim + d] = nextafterf(mx[d], INFINITY);
            }
        }
    }
    return p;
}

static float *supports(int64_t n, float h, float spread) {
    float *s = (float *)malloc(sizeof(float) * (size_t)(n ? n : 1));
    for (int64_t i = 0; i < n; ++i) s[i] = h * (1.0f + spread * (float)urand());
    return s;
}

static void one_case(int dim, int mode, uint32_t pmask, uint32_t L, int32_t sr, int64_t nq, int64_t nr,
                     int self, float lo, float hi) {
    const int outside = lo < -1.0f || hi > 1.0f;
    float mn[3] = {-1.0f, -1.0f, -1.0f}, mx[3] = {1.0f, 1.0f, 1.0f};
    const float h = 0.3f;
    float *q = cloud(nq, dim, lo, hi, mn, mx, outside);
    float *hq = supports(nq, h, 0.5f);
    float *r = self ? NULL : cloud(nr, dim, lo, hi, mn, mx, outside);
    float *hr = self ? NULL : supports(nr, h, 0.5f);
    const int64_t nref = self ? nq : nr;
    orc_search *s = orc_search_create(q, hq, nq, r, hr, nr, dim, mn, mx, pmask, mode, L, sr, 1);
    CHECK(s != NULL, "create dim=%d mode=%d", dim, mode);
    if (!s) goto out;
    const int64_t total = orc_search_total(s);
    int64_t *row = (int64_t *)malloc(sizeof(int64_t) * (size_t)(total ? total : 1));
    int64_t *col = (int64_t *)malloc(sizeof(int64_t) * (size_t)(total ? total : 1));
    orc_search_fill(s, row, col);
    const int32_t *counts = orc_search_counts(s);
    int64_t k = 0;
    for (int64_t i = 0; i < nq; ++i)
        for (int32_t t = 0; t < counts[i]; ++t, ++k) {
            CHECK(row[k] == i, "row order dim=%d mode=%d", dim, mode);
            CHECK(col[k] >= 0 && col[k] < nref, "col range dim=%d mode=%d", dim, mode);
        }
    CHECK(k == total, "total dim=%d mode=%d", dim, mode);
    /* the 3^d cells of edge >= hMax hold every neighbour when no point lies
     * outside a non-periodic dimension (there a cell coordinate of -1 or
     * numCells aliases another cell's linear index, as in the reference) */
    if (sr == 1 && (pmask == (1u << dim) - 1u || !outside)) {
        int32_t *bc = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nq ? nq : 1));
        orc_brute_counts(q, hq, nq, self ? q : r, self ? hq : hr, nref, dim, mn, mx, pmask, mode, bc);
        for (int64_t i = 0; i < nq; ++i)
            CHECK(bc[i] == counts[i], "brute count dim=%d mode=%d L=%u i=%lld: %d vs %d", dim, mode, L,
                  (long long)i, bc[i], counts[i]);
        free(bc);
        ++brute_checked;
    }
    const orc_map *m = orc_search_map(s);
    int32_t *si = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nref ? nref : 1));
    float *sp = (float *)malloc(sizeof(float) * (size_t)(nref ? nref : 1) * dim);
    int32_t *ct = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)(m->ncells ? m->ncells : 1));
    int32_t *ht = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)L);
    int32_t nc[3];
    orc_map_export(m, si, sp, ct, ht, nc);
    free(si); free(sp); free(ct); free(ht);
    free(row); free(col);
    orc_search_free(s);
out:
    free(q); free(hq); free(r); free(hr);
}

static void small_case(int dim, int mode, int fixed) {
    float mn[3] = {-1.0f, -1.0f, -1.0f}, mx[3] = {1.0f, 1.0f, 1.0f};
    const int64_t nq = 150, nr = 170;
    float *q = cloud(nq, dim, -1.0f, 1.0f, mn, mx, 1), *r = cloud(nr, dim, -1.0f, 1.0f, mn, mx, 1);
    float *hq = supports(nq, 0.25f, 0.5f), *hr = supports(nr, 0.25f, 0.5f);
    int32_t *counts = (int32_t *)malloc(sizeof(int32_t) * nq);
    orc_small(q, hq, nq, r, hr, nr, dim, mn, mx, 1u, mode, fixed, 0.3f, counts, NULL, NULL, NULL);
    int64_t *off = (int64_t *)malloc(sizeof(int64_t) * (nq + 1));
    off[0] = 0;
    for (int64_t i = 0; i < nq; ++i) off[i + 1] = off[i] + counts[i];
    int64_t *pi = (int64_t *)malloc(sizeof(int64_t) * (size_t)(off[nq] ? off[nq] : 1));
    int64_t *pj = (int64_t *)malloc(sizeof(int64_t) * (size_t)(off[nq] ? off[nq] : 1));
    orc_small(q, hq, nq, r, hr, nr, dim, mn, mx, 1u, mode, fixed, 0.3f, NULL, off, pi, pj);
    for (int64_t i = 0; i < nq; ++i)
        for (int64_t t = off[i]; t < off[i + 1]; ++t) CHECK(pi[t] == i && pj[t] >= 0 && pj[t] < nr, "small pair");
    free(q); free(r); free(hq); free(hr); free(counts); free(off); free(pi); free(pj);
}

int main(void) {
    const uint32_t lengths[3] = {1, 7, 4096};
    int cases = 0;
    for (int dim = 1; dim <= 3; ++dim)
        for (int mode = 0; mode < 4; ++mode)
            for (uint32_t pm = 0; pm < 2; ++pm)
                for (int li = 0; li < 3; ++li) {
                    const uint32_t pmask = pm ? (1u << dim) - 1u : 0u;
                    one_case(dim, mode, pmask, lengths[li], 1, 400, 0, 1, -1.1f, 1.1f);
                    one_case(dim, mode, pmask, lengths[li], 1, 400, 0, 1, -1.0f, 1.0f);
                    one_case(dim, mode, pmask, lengths[li], 1, 300, 350, 0, -1.0f, 1.0f);
                    cases += 3;
                }
    for (int dim = 1; dim <= 3; ++dim) {
        one_case(dim, ORC_SYMMETRIC, 0u, 97u, 2, 300, 0, 1, -1.0f, 1.0f);     /* search range 2 */
        one_case(dim, ORC_GATHER, 1u, 97u, 1, 0, 0, 1, -1.0f, 1.0f);         /* empty self-search */
        one_case(dim, ORC_GATHER, 0u, 97u, 1, 0, 50, 0, -1.0f, 1.0f);        /* no queries */
        one_case(dim, ORC_SCATTER, 0u, 97u, 1, 50, 0, 0, -1.0f, 1.0f);       /* no references */
        cases += 4;
        for (int mode = 0; mode < 4; ++mode)
            for (int fixed = 0; fixed < 2; ++fixed) {
                small_case(dim, mode, fixed);
                ++cases;
            }
    }
    printf("%d cases (%d against the brute force), %d failures\n", cases, brute_checked, failures);
    return failures ? 1 : 0;
}
