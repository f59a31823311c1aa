/*
 * pp_oracle.c -- CPU oracle for the RDF / kNN spatial-statistics path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the GPU path is
 * compared against.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it; the product library never links or calls it.
 *
 * It restates, in plain C, the semantics of the reference path
 *   horderdist::compute_generic_spatstat3   (src/PARAPROBE_SolverHdl.cpp:3685-4296)
 * for the families RDF, 1NN, RIPK, MKNN and CNTNB:
 *   - neighbour set   { j != i (by index) : fl32((dx*dx + dy*dy) + dz*dz) <= fl32(R*R) },
 *     dx = fl32(x_center - x_j)          (src/PARAPROBE_KDTree.h:49-58,
 *                                          src/PARAPROBE_KDTree.cpp:563-573)
 *   - a center is considered iff its type is in some task's target set and
 *     ion2surf[i] >= fl32(R*R)           (S.cpp:3846-3860)
 *   - each neighbour distance is sqrtf(d2) (KDTree.cpp:571), binned with
 *     histogram::add                      (src/PARAPROBE_Histogram.h:54-116)
 *   - per task, a center counts once (the `break` at S.cpp:4010) ; RDF/RIPK add
 *     every env-type distance, 1NN adds the minimum or R+EPSILON, MKNN adds the
 *     k-th smallest (k-1 index into the descending candidate list) or R+EPSILON,
 *     CNTNB adds one count per env type      (S.cpp:3999-4101)
 * Neighbour search is either brute force (O(N^2), tiny inputs) or a uniform
 * grid with a 27-cell stencil of cell width >= 1.001 R computed in double, so
 * that the exact neighbour set is always enumerated; the two modes pin each
 * other in the tests.  The fp-contract-off FP32 distance expression is the
 * contract (the file is compiled with -ffp-contract=off).
 *
 * Output layout (int64): every histogram is [lowest, bins[nbins], highest]
 * (C = nbins + 2 entries); RDF/1NN/RIPK are [ntasks][C]; MKNN is
 * [ntasks][nk][C] with the k list in descending order (S.cpp:3759, 4086);
 * CNTNB is [ntasks][cntnb_cap] indexed by the neighbour count.
 *
 * Also: pp_oracle_binedge_report, the FP64 bin-edge oracle, which counts how
 * many in-ball pairs lie within one float rounding of a bin edge and how many
 * FP32-pipeline bins differ from the bin of the exact real distance.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_EPSILON (1.0e-6) /* src/PARAPROBE_Numerics.h:45 */

enum {
    ORC_RDF = 1u << 0,
    ORC_1NN = 1u << 1,
    ORC_RIPK = 1u << 2,
    ORC_MKNN = 1u << 3,
    ORC_CNTNB = 1u << 5
};

typedef struct {
    double start, width, end, tmp;
    uint32_t nbins;
    int valid;
} ohist;

/* histogram::histogram(start, width, end), src/PARAPROBE_Histogram.h:54-89.
 * The arguments are the float settings promoted to double. */
static void ohist_init(ohist *h, float s, float w, float e)
{
    memset(h, 0, sizeof(*h));
    double ds = s, dw = w, de = e;
    if (dw > ORC_EPSILON) {
        size_t nb = (size_t)((de - ds) / dw);
        if (nb <= 0xFFFFFFFFull) {
            h->start = ds;
            h->width = dw;
            h->end = de;
            h->nbins = (uint32_t)nb;
            h->tmp = 1.0 / (de - ds) * (double)nb;
            h->valid = 1;
        }
    }
}

/* histogram::add(float), src/PARAPROBE_Histogram.h:97-116, returning the
 * category: 0 = cnts_lowest, 1 + b = cnts[b], nbins + 1 = cnts_highest. */
static inline uint32_t ohist_cat(const ohist *h, float x)
{
    double xx = x;
    if (xx >= h->start && xx <= h->end) {
        float rb = (float)((xx - h->start) * h->tmp);
        unsigned int b = (unsigned int)floorf(rb);
        if (b < h->nbins)
            return 1u + b;
        return h->nbins + 1u;
    }
    if (xx < h->start)
        return 0u;
    return h->nbins + 1u;
}

int pp_oracle_hist_geometry(float rmin, float rincr, float rmax, int64_t *nbins, double *tmp)
{
    ohist h;
    ohist_init(&h, rmin, rincr, rmax);
    *nbins = h.nbins;
    *tmp = h.tmp;
    return h.valid ? 0 : 1;
}

/* Category of one float distance (the literal histogram::add). */
int64_t pp_oracle_category(float rmin, float rincr, float rmax, float d)
{
    ohist h;
    ohist_init(&h, rmin, rincr, rmax);
    return (int64_t)ohist_cat(&h, d);
}

/* Categories of many squared distances: cat(sqrtf(d2)). */
void pp_oracle_categories_d2(float rmin, float rincr, float rmax, const float *d2, int64_t n,
                             int32_t *out)
{
    ohist h;
    ohist_init(&h, rmin, rincr, rmax);
    for (int64_t i = 0; i < n; ++i)
        out[i] = (int32_t)ohist_cat(&h, sqrtf(d2[i]));
}

/* Every float d in [d_lo, d_hi] (non-negative), in bit-pattern order, through
 * the literal histogram::add: records each change of category (first[i] = bit
 * pattern of the first float of the new category, cat[i] = that category; at
 * most cap entries) and returns the number of changes, or -1 when the
 * category ever decreases. */
int64_t pp_oracle_category_sweep(float rmin, float rincr, float rmax, float d_lo, float d_hi,
                                 uint32_t *first, int32_t *cat, int64_t cap)
{
    ohist h;
    ohist_init(&h, rmin, rincr, rmax);
    if (!h.valid || !(d_lo >= 0.f) || !(d_hi >= d_lo))
        return -1;
    uint32_t u0, u1;
    memcpy(&u0, &d_lo, 4);
    memcpy(&u1, &d_hi, 4);
    enum { NCH = 256, LCAP = 4096 };
    /* chunks in parallel: each keeps its own changes, merged in order */
    static uint32_t lf[NCH][LCAP];
    static int32_t lc[NCH][LCAP];
    int64_t ln[NCH];
    uint32_t c_first[NCH], c_last[NCH];
    int bad = 0;
    const uint64_t span = (uint64_t)u1 - u0 + 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
#endif
    for (int k = 0; k < NCH; ++k) {
        const uint64_t a = u0 + span * (uint64_t)k / NCH, b = u0 + span * (uint64_t)(k + 1) / NCH;
        ln[k] = 0;
        c_first[k] = c_last[k] = 0;
        if (a >= b)
            continue;
        float x;
        uint32_t ua = (uint32_t)a;
        memcpy(&x, &ua, 4);
        uint32_t prev = ohist_cat(&h, x);
        c_first[k] = prev;
        for (uint64_t u = a + 1; u < b; ++u) {
            const uint32_t uu = (uint32_t)u;
            memcpy(&x, &uu, 4);
            const uint32_t c = ohist_cat(&h, x);
            if (c != prev) {
                if (c < prev)
                    bad = 1;
                if (ln[k] < LCAP) {
                    lf[k][ln[k]] = uu;
                    lc[k][ln[k]] = (int32_t)c;
                }
                ++ln[k];
                prev = c;
            }
        }
        c_last[k] = prev;
    }
    if (bad)
        return -1;
    int64_t n = 0;
    int have = 0;
    uint32_t prev = 0;
    for (int k = 0; k < NCH; ++k) {
        const uint64_t a = u0 + span * (uint64_t)k / NCH, b = u0 + span * (uint64_t)(k + 1) / NCH;
        if (a >= b)
            continue;
        if (have && c_first[k] != prev) {
            if (c_first[k] < prev)
                return -1;
            if (n < cap) {
                first[n] = (uint32_t)a;
                cat[n] = (int32_t)c_first[k];
            }
            ++n;
        }
        if (ln[k] > LCAP)
            return -1;
        for (int64_t i = 0; i < ln[k]; ++i) {
            if (n < cap) {
                first[n] = lf[k][i];
                cat[n] = lc[k][i];
            }
            ++n;
        }
        prev = c_last[k];
        have = 1;
    }
    return n;
}

static inline int has_type(const uint8_t *mask32, uint8_t t)
{
    return (mask32[t >> 3] >> (t & 7)) & 1;
}

static int cmp_float(const void *a, const void *b)
{
    float fa = *(const float *)a, fb = *(const float *)b;
    return (fa > fb) - (fa < fb);
}

/* Moves the kneed smallest of a[0..m) to the front, in ascending order
 * (kneed <= m): Hoare partitions around a median-of-three pivot until the
 * kneed-th position is settled (what std::nth_element guarantees, S.cpp:4088),
 * then a sort of the front.  Values are distances, never NaN. */
static void smallest_sorted(float *a, int64_t m, int64_t kneed)
{
    int64_t lo = 0, hi = m - 1;
    const int64_t k = kneed - 1;
    while (hi > lo) {
        const int64_t mid = lo + (hi - lo) / 2;
        float x = a[lo], y = a[mid], z = a[hi];
        const float piv = (x < y) ? ((y < z) ? y : (x < z ? z : x)) : ((x < z) ? x : (y < z ? z : y));
        int64_t i = lo, j = hi;
        while (i <= j) {
            while (a[i] < piv)
                ++i;
            while (a[j] > piv)
                --j;
            if (i <= j) {
                const float t = a[i];
                a[i] = a[j];
                a[j] = t;
                ++i;
                --j;
            }
        }
        /* a[lo..j] <= piv <= a[i..hi], and a[j+1..i-1] == piv */
        if (k <= j)
            hi = j;
        else if (k >= i)
            lo = i;
        else
            break;
    }
    qsort(a, (size_t)kneed, sizeof(float), cmp_float);
}

typedef struct {
    int64_t n;
    const float *xyz;
    double lo[3];
    double inv;
    int64_t dim[3];
    int64_t *start; /* ncells + 1 */
    int64_t *order; /* ions sorted by cell */
} ogrid;

static int64_t ogrid_coord(const ogrid *g, int a, float v)
{
    int64_t c = (int64_t)floor(((double)v - g->lo[a]) * g->inv);
    if (c < 0)
        c = 0;
    if (c >= g->dim[a])
        c = g->dim[a] - 1;
    return c;
}

static int ogrid_build(ogrid *g, int64_t n, const float *xyz, float R)
{
    memset(g, 0, sizeof(*g));
    g->n = n;
    g->xyz = xyz;
    double hi[3];
    for (int a = 0; a < 3; ++a) {
        g->lo[a] = 1e300;
        hi[a] = -1e300;
    }
    for (int64_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            double v = xyz[3 * i + a];
            if (v < g->lo[a])
                g->lo[a] = v;
            if (v > hi[a])
                hi[a] = v;
        }
    if (n == 0)
        for (int a = 0; a < 3; ++a)
            g->lo[a] = hi[a] = 0.0;
    double w = (double)R * 1.001;
    if (w <= 0.0)
        w = 1.0;
    /* cap the grid so that tiny radii on wide clouds stay bounded in memory */
    for (;;) {
        int64_t tot = 1;
        for (int a = 0; a < 3; ++a) {
            g->dim[a] = (int64_t)floor((hi[a] - g->lo[a]) / w) + 1;
            tot *= g->dim[a];
        }
        if (tot <= 4 * n + 64)
            break;
        w *= 1.5;
    }
    g->inv = 1.0 / w;
    int64_t nc = g->dim[0] * g->dim[1] * g->dim[2];
    g->start = (int64_t *)calloc((size_t)nc + 1, sizeof(int64_t));
    g->order = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    int64_t *cell = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    if (!g->start || !g->order || !cell) {
        free(cell);
        return 1;
    }
    for (int64_t i = 0; i < n; ++i) {
        int64_t cx = ogrid_coord(g, 0, xyz[3 * i]);
        int64_t cy = ogrid_coord(g, 1, xyz[3 * i + 1]);
        int64_t cz = ogrid_coord(g, 2, xyz[3 * i + 2]);
        cell[i] = (cz * g->dim[1] + cy) * g->dim[0] + cx;
        g->start[cell[i] + 1]++;
    }
    for (int64_t c = 0; c < nc; ++c)
        g->start[c + 1] += g->start[c];
    int64_t *fill = (int64_t *)malloc((size_t)nc * sizeof(int64_t) + 8);
    memcpy(fill, g->start, (size_t)nc * sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i)
        g->order[fill[cell[i]]++] = i;
    free(fill);
    free(cell);
    return 0;
}

static void ogrid_free(ogrid *g)
{
    free(g->start);
    free(g->order);
}

typedef struct {
    float d;
    uint8_t m;
} onbor;

/* Gather every neighbour of center i: d2 <= R2, j != i by index. */
static int64_t gather(const ogrid *g, int brute, int64_t i, const float *xyz, const uint8_t *type,
                      float R2, onbor **buf, int64_t *cap)
{
    int64_t cnt = 0;
    const float xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
#define ORC_VISIT(j)                                                   \
    do {                                                               \
        const float dx = xi - xyz[3 * (j)];                            \
        const float dy = yi - xyz[3 * (j) + 1];                        \
        const float dz = zi - xyz[3 * (j) + 2];                        \
        const float d2 = dx * dx + dy * dy + dz * dz;                  \
        if (!(d2 > R2) && (j) != i) {                                  \
            if (cnt == *cap) {                                         \
                *cap = *cap * 2 + 64;                                  \
                *buf = (onbor *)realloc(*buf, (size_t)*cap * sizeof(onbor)); \
            }                                                          \
            (*buf)[cnt].d = sqrtf(d2);                                 \
            (*buf)[cnt].m = type[(j)];                                 \
            ++cnt;                                                     \
        }                                                              \
    } while (0)
    if (brute) {
        for (int64_t j = 0; j < g->n; ++j)
            ORC_VISIT(j);
        return cnt;
    }
    int64_t c[3] = {ogrid_coord(g, 0, xi), ogrid_coord(g, 1, yi), ogrid_coord(g, 2, zi)};
    for (int64_t z = c[2] - 1; z <= c[2] + 1; ++z) {
        if (z < 0 || z >= g->dim[2])
            continue;
        for (int64_t y = c[1] - 1; y <= c[1] + 1; ++y) {
            if (y < 0 || y >= g->dim[1])
                continue;
            for (int64_t x = c[0] - 1; x <= c[0] + 1; ++x) {
                if (x < 0 || x >= g->dim[0])
                    continue;
                int64_t cc = (z * g->dim[1] + y) * g->dim[0] + x;
                for (int64_t p = g->start[cc]; p < g->start[cc + 1]; ++p) {
                    int64_t j = g->order[p];
                    ORC_VISIT(j);
                }
            }
        }
    }
#undef ORC_VISIT
    return cnt;
}

/*
 * The oracle proper.  stats[0] = considered centers, stats[1] = discarded
 * ions, stats[2] = sum over considered centers of |neighbours within R| (all
 * types, the size of the reference's neighbors1dm1), stats[3] = max CNTNB
 * count seen.  Returns 0 on success, 1 on bad arguments, 2 when a CNTNB count
 * exceeds cntnb_cap.
 */
int pp_oracle_spatstat(int64_t n, const float *xyz, const uint8_t *type, const float *ion2surf,
                       int32_t ntasks, const uint8_t *trg_mask, const uint8_t *env_mask,
                       float rmin, float rincr, float rmax, uint32_t families, int32_t nk,
                       const int64_t *kminus1_desc, int32_t brute, int32_t nthreads,
                       int64_t *rdf, int64_t *nn1, int64_t *ripk, int64_t *mknn, int64_t *cntnb,
                       int64_t cntnb_cap, int64_t *stats)
{
    ohist h;
    ohist_init(&h, rmin, rincr, rmax);
    if (!h.valid || ntasks < 0 || nk < 0)
        return 1;
    const int64_t C = (int64_t)h.nbins + 2;
    const float R = rmax;
    const float R2 = R * R;
    /* R + EPSILON: float R promoted to double, plus 1e-6, passed as float */
    const float sentinel = (float)((double)R + ORC_EPSILON);
    const uint32_t cat_sentinel = ohist_cat(&h, sentinel);

    uint8_t anytrg[32];
    memset(anytrg, 0, sizeof(anytrg));
    for (int t = 0; t < ntasks; ++t)
        for (int b = 0; b < 32; ++b)
            anytrg[b] |= trg_mask[32 * t + b];

    ogrid g;
    if (ogrid_build(&g, n, xyz, R))
        return 1;

    const int64_t len_rdf = (int64_t)ntasks * C;
    const int64_t len_mknn = (int64_t)ntasks * nk * C;
    const int64_t len_cnt = (int64_t)ntasks * cntnb_cap;
    if (families & ORC_RDF)
        memset(rdf, 0, (size_t)len_rdf * 8);
    if (families & ORC_1NN)
        memset(nn1, 0, (size_t)len_rdf * 8);
    if (families & ORC_RIPK)
        memset(ripk, 0, (size_t)len_rdf * 8);
    if (families & ORC_MKNN)
        memset(mknn, 0, (size_t)len_mknn * 8);
    if (families & ORC_CNTNB)
        memset(cntnb, 0, (size_t)len_cnt * 8);
    int64_t considered = 0, discarded = 0, inball = 0, maxcnt = 0;
    int overflow = 0;

#ifdef _OPENMP
    if (nthreads <= 0)
        nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads) reduction(+ : considered, discarded, inball) \
    reduction(max : maxcnt) reduction(| : overflow)
#endif
    {
        int64_t *l_rdf = (int64_t *)calloc((size_t)len_rdf + 1, 8);
        int64_t *l_nn1 = (int64_t *)calloc((size_t)len_rdf + 1, 8);
        int64_t *l_mknn = (int64_t *)calloc((size_t)len_mknn + 1, 8);
        int64_t *l_cnt = (int64_t *)calloc((size_t)len_cnt + 1, 8);
        onbor *nb = NULL;
        int64_t cap = 0;
        float *tmp = NULL;
        int64_t tcap = 0;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (int64_t i = 0; i < n; ++i) {
            const uint8_t mi = type[i];
            if (!(has_type(anytrg, mi) && ion2surf[i] >= R2)) {
                ++discarded;
                continue;
            }
            ++considered;
            int64_t cnt = gather(&g, brute, i, xyz, type, R2, &nb, &cap);
            inball += cnt;
            for (int t = 0; t < ntasks; ++t) {
                if (!has_type(trg_mask + 32 * t, mi))
                    continue; /* counted once per task */
                const uint8_t *env = env_mask + 32 * t;
                if (families & (ORC_RDF | ORC_RIPK)) {
                    for (int64_t q = 0; q < cnt; ++q)
                        if (has_type(env, nb[q].m))
                            l_rdf[t * C + ohist_cat(&h, nb[q].d)]++;
                }
                if (families & ORC_1NN) {
                    int found = 0;
                    float closest = 3.402823466e38f;
                    for (int64_t q = 0; q < cnt; ++q)
                        if (has_type(env, nb[q].m) && !(nb[q].d > closest)) {
                            closest = nb[q].d;
                            found = 1;
                        }
                    l_nn1[t * C + (found ? ohist_cat(&h, closest) : cat_sentinel)]++;
                }
                if (families & ORC_CNTNB) {
                    /* one count per env type id of the task (S.cpp:4057-4069) */
                    for (int e = 0; e < 256; ++e) {
                        if (!has_type(env, (uint8_t)e))
                            continue;
                        int64_t ce = 0;
                        for (int64_t q = 0; q < cnt; ++q)
                            ce += (nb[q].m == e);
                        if (ce > maxcnt)
                            maxcnt = ce;
                        if (ce < cntnb_cap)
                            l_cnt[t * cntnb_cap + ce]++;
                        else
                            overflow = 1;
                    }
                }
                if ((families & ORC_MKNN) && nk > 0) {
                    int64_t m = 0;
                    for (int64_t q = 0; q < cnt; ++q)
                        if (has_type(env, nb[q].m)) {
                            if (m == tcap) {
                                tcap = tcap * 2 + 64;
                                tmp = (float *)realloc(tmp, (size_t)tcap * sizeof(float));
                            }
                            tmp[m++] = nb[q].d;
                        }
                    /* k-th order statistic: the kneed smallest in order give
                     * the same values as nth_element (S.cpp:4088); a full sort
                     * when most of the list is needed */
                    int64_t kneed = 0;
                    for (int kth = 0; kth < nk; ++kth)
                        if (kminus1_desc[kth] + 1 > kneed)
                            kneed = kminus1_desc[kth] + 1;
                    if (kneed > m)
                        kneed = m;
                    if (m > 1 && 4 * kneed < m)
                        smallest_sorted(tmp, m, kneed);
                    else if (m > 1)
                        qsort(tmp, (size_t)m, sizeof(float), cmp_float);
                    for (int kth = 0; kth < nk; ++kth) {
                        int64_t k = kminus1_desc[kth];
                        uint32_t c = (m > k) ? ohist_cat(&h, tmp[k]) : cat_sentinel;
                        l_mknn[((int64_t)t * nk + kth) * C + c]++;
                    }
                }
            }
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        {
            if (families & ORC_RDF)
                for (int64_t q = 0; q < len_rdf; ++q)
                    rdf[q] += l_rdf[q];
            if (families & ORC_RIPK)
                for (int64_t q = 0; q < len_rdf; ++q)
                    ripk[q] += l_rdf[q];
            if (families & ORC_1NN)
                for (int64_t q = 0; q < len_rdf; ++q)
                    nn1[q] += l_nn1[q];
            if (families & ORC_MKNN)
                for (int64_t q = 0; q < len_mknn; ++q)
                    mknn[q] += l_mknn[q];
            if (families & ORC_CNTNB)
                for (int64_t q = 0; q < len_cnt; ++q)
                    cntnb[q] += l_cnt[q];
        }
        free(l_rdf);
        free(l_nn1);
        free(l_mknn);
        free(l_cnt);
        free(nb);
        free(tmp);
    }
    ogrid_free(&g);
    stats[0] = considered;
    stats[1] = discarded;
    stats[2] = inball;
    stats[3] = maxcnt;
    return overflow ? 2 : 0;
}

/*
 * FP64 bin-edge oracle.  For every considered center (same rules as above,
 * one task set given by trg/env masks of a single task) and every env-type
 * neighbour accepted by the FP32 pipeline, the exact real distance of the two
 * stored float points is evaluated in double and binned with exact real
 * arithmetic (floor((d - start) * tmp) on the double value).  Outputs:
 * out[0] = pairs examined, out[1] = pairs whose FP32-pipeline category
 * differs from the category of the exact distance, out[2] = pairs whose exact
 * distance lies within one float ulp of a bin edge, out[3] = pairs the FP64
 * test puts in the ball while FP32 does not, or the reverse.
 * center_stride > 1 samples every center_stride-th ion as a center.
 */
int pp_oracle_binedge_report(int64_t n, const float *xyz, const uint8_t *type, const float *ion2surf,
                             const uint8_t *trg_mask, const uint8_t *env_mask, float rmin,
                             float rincr, float rmax, int64_t center_stride, int32_t nthreads,
                             int64_t *out)
{
    ohist h;
    ohist_init(&h, rmin, rincr, rmax);
    if (!h.valid || center_stride < 1)
        return 1;
    const float R = rmax, R2 = R * R;
    ogrid g;
    if (ogrid_build(&g, n, xyz, R * 1.0001f))
        return 1;
    int64_t examined = 0, differ = 0, near_edge = 0, ball_flip = 0;
#ifdef _OPENMP
    if (nthreads <= 0)
        nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 16) \
    reduction(+ : examined, differ, near_edge, ball_flip)
#endif
    for (int64_t i = 0; i < n; i += center_stride) {
        if (!(has_type(trg_mask, type[i]) && ion2surf[i] >= R2))
            continue;
        const float xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
        int64_t c[3] = {ogrid_coord(&g, 0, xi), ogrid_coord(&g, 1, yi), ogrid_coord(&g, 2, zi)};
        for (int64_t z = c[2] - 1; z <= c[2] + 1; ++z) {
            if (z < 0 || z >= g.dim[2])
                continue;
            for (int64_t y = c[1] - 1; y <= c[1] + 1; ++y) {
                if (y < 0 || y >= g.dim[1])
                    continue;
                for (int64_t x = c[0] - 1; x <= c[0] + 1; ++x) {
                    if (x < 0 || x >= g.dim[0])
                        continue;
                    int64_t cc = (z * g.dim[1] + y) * g.dim[0] + x;
                    for (int64_t p = g.start[cc]; p < g.start[cc + 1]; ++p) {
                        int64_t j = g.order[p];
                        if (j == i || !has_type(env_mask, type[j]))
                            continue;
                        const float dx = xi - xyz[3 * j], dy = yi - xyz[3 * j + 1],
                                    dz = zi - xyz[3 * j + 2];
                        const float d2 = dx * dx + dy * dy + dz * dz;
                        /* the FP32 expression is within a few ulp of the exact
                         * value: far outside the ball neither test keeps it */
                        if (d2 > R2 * 1.001f)
                            continue;
                        const double ex = (double)xi - (double)xyz[3 * j];
                        const double ey = (double)yi - (double)xyz[3 * j + 1];
                        const double ez = (double)zi - (double)xyz[3 * j + 2];
                        const double dd = sqrt(ex * ex + ey * ey + ez * ez);
                        int in32 = !(d2 > R2), in64 = !(dd > (double)R);
                        if (in32 != in64)
                            ++ball_flip;
                        if (!in32)
                            continue;
                        ++examined;
                        uint32_t c32 = ohist_cat(&h, sqrtf(d2));
                        uint32_t c64;
                        if (dd < h.start)
                            c64 = 0;
                        else if (dd > h.end)
                            c64 = h.nbins + 1;
                        else {
                            double rb = (dd - h.start) * h.tmp;
                            double fb = floor(rb);
                            c64 = (fb < (double)h.nbins) ? 1u + (uint32_t)fb : h.nbins + 1u;
                        }
                        if (c32 != c64)
                            ++differ;
                        /* distance to the nearest ideal edge start + e/tmp */
                        double pos = (dd - h.start) * h.tmp;
                        double e = floor(pos + 0.5);
                        double edge_d = h.start + e / h.tmp;
                        double ulp = (double)nextafterf((float)dd, 3.4e38f) - (double)(float)dd;
                        if (fabs(dd - edge_d) <= ulp)
                            ++near_edge;
                    }
                }
            }
        }
    }
    ogrid_free(&g);
    out[0] = examined;
    out[1] = differ;
    out[2] = near_edge;
    out[3] = ball_flip;
    return 0;
}

/* ------------------------------------------------------------------------
 * Ion-to-surface squared distance (the ion2surf producer).
 *
 * pp_oracle_closest_tri restates mathHdl::closestPointOnTriangle
 * (src/PARAPROBE_Math.cpp:88-171): the FP32 squared distance from p to the
 * triangle t (9 floats, vertices 1..3), with the reference's region logic,
 * its CLAMP/MIN/MAX ternaries (src/PARAPROBE_Abinitio.h:47-49, a NaN clamps
 * to hi) and its operation order (1.f/det then two multiplies; a - 2*b + c).
 */
static float orc_clamp01(float x)
{
    const float m = (x < 0.f) ? 0.f : x;
    return (m < 1.f) ? m : 1.f;
}

float pp_oracle_closest_tri(const float *t, const float *p)
{
    const float e0u = t[3] - t[0], e0v = t[4] - t[1], e0w = t[5] - t[2];
    const float e1u = t[6] - t[0], e1v = t[7] - t[1], e1w = t[8] - t[2];
    const float v0u = t[0] - p[0], v0v = t[1] - p[1], v0w = t[2] - p[2];
    const float a = e0u * e0u + e0v * e0v + e0w * e0w;
    const float b = e0u * e1u + e0v * e1v + e0w * e1w;
    const float c = e1u * e1u + e1v * e1v + e1w * e1w;
    const float d = e0u * v0u + e0v * v0v + e0w * v0w;
    const float e = e1u * v0u + e1v * v0v + e1w * v0w;
    const float det = a * c - b * b;
    float s = b * e - c * d;
    float u = b * d - a * e;
    if (s + u < det) {
        if (s < 0.f) {
            if (u < 0.f) {
                if (d < 0.f) {
                    s = orc_clamp01(-d / a);
                    u = 0.f;
                } else {
                    s = 0.f;
                    u = orc_clamp01(-e / c);
                }
            } else {
                s = 0.f;
                u = orc_clamp01(-e / c);
            }
        } else if (u < 0.f) {
            s = orc_clamp01(-d / a);
            u = 0.f;
        } else {
            const float inv = 1.f / det;
            s *= inv;
            u *= inv;
        }
    } else if (s < 0.f) {
        const float t0 = b + d, t1 = c + e;
        if (t1 > t0) {
            s = orc_clamp01((t1 - t0) / (a - 2.f * b + c));
            u = 1.f - s;
        } else {
            u = orc_clamp01(-e / c);
            s = 0.f;
        }
    } else if (u < 0.f) {
        if (a + d > b + e) {
            s = orc_clamp01((c + e - b - d) / (a - 2.f * b + c));
            u = 1.f - s;
        } else {
            s = orc_clamp01(-e / c);
            u = 0.f;
        }
    } else {
        s = orc_clamp01((c + e - b - d) / (a - 2.f * b + c));
        u = 1.f - s;
    }
    const float qx = t[0] + s * e0u + u * e1u;
    const float qy = t[1] + s * e0v + u * e1v;
    const float qz = t[2] + s * e0w + u * e1w;
    const float dx = qx - p[0], dy = qy - p[1], dz = qz - p[2];
    return dx * dx + dy * dy + dz * dz;
}

/*
 * pp_oracle_ion2surf restates solver::surface_distancing2
 * (src/PARAPROBE_SolverHdl.cpp:312-484) by brute force over the triangles in
 * index order: start at fl32(R*R); a triangle is a candidate iff its vertex
 * box (fminf/fmaxf, SolverHdl.cpp:2716-2721) overlaps the float box
 * [p - R, p + R] with closed bounds (src/PARAPROBE_AABBTree.cpp:112-123);
 * a candidate's value replaces the running best unless it is larger
 * (SolverHdl.cpp:425-433).
 */
void pp_oracle_ion2surf(const float *xyz, int64_t n, const float *tri, int64_t ntri, float R,
                        float *out)
{
    const float RSQR = R * R;
    float *box = (float *)malloc(sizeof(float) * 6 * (size_t)(ntri > 0 ? ntri : 1));
    for (int64_t t = 0; t < ntri; ++t)
        for (int k = 0; k < 3; ++k) {
            const float *v = tri + 9 * t;
            box[6 * t + k] = fminf(fminf(v[k], v[3 + k]), v[6 + k]);
            box[6 * t + 3 + k] = fmaxf(fmaxf(v[k], v[3 + k]), v[6 + k]);
        }
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
        const float *p = xyz + 3 * i;
        const float lo[3] = {p[0] - R, p[1] - R, p[2] - R};
        const float hi[3] = {p[0] + R, p[1] + R, p[2] + R};
        float best = RSQR;
        for (int64_t t = 0; t < ntri; ++t) {
            const float *b = box + 6 * t;
            if (hi[0] < b[0] || lo[0] > b[3] || hi[1] < b[1] || lo[1] > b[4] || hi[2] < b[2] ||
                lo[2] > b[5])
                continue;
            const float v = pp_oracle_closest_tri(tri + 9 * t, p);
            if (v > best)
                continue;
            best = v;
        }
        out[i] = best;
    }
    free(box);
}

/* ------------------------------------------------------------------------
 * POS records and ranging.
 *
 * pp_oracle_pos_decode restates pos(pos_be_2_stl_le)
 * (src/PARAPROBE_EPOSEndianness.h:45-84: each of the four big-endian floats
 * of a 16-byte record has its bytes reversed) and PeriodicTable::mq2ionname
 * (src/PARAPROBE_PeriodicTable.cpp:65-91): the first range in the given order
 * with !(mq < lo) && !(mq > hi) gives the type, else 0.
 */
static void orc_records_decode(const uint8_t *rec, int64_t n, int64_t stride, const float *lo,
                               const float *hi, const uint8_t *rtype, int32_t nranges,
                               float *xyz, float *mq, uint8_t *type)
{
    for (int64_t i = 0; i < n; ++i) {
        float v[4];
        for (int k = 0; k < 4; ++k) {
            uint8_t b[4];
            for (int q = 0; q < 4; ++q)
                b[q] = rec[stride * i + 4 * k + 3 - q];
            memcpy(&v[k], b, 4);
        }
        xyz[3 * i] = v[0];
        xyz[3 * i + 1] = v[1];
        xyz[3 * i + 2] = v[2];
        if (mq)
            mq[i] = v[3];
        uint8_t t = 0;
        for (int32_t r = 0; r < nranges; ++r) {
            if (v[3] < lo[r])
                continue;
            if (v[3] > hi[r])
                continue;
            t = rtype[r];
            break;
        }
        type[i] = t;
    }
}

/* ------------------------------------------------------------------------
 * NPCORR (the n-point correlation family, S.cpp:3899-3960, npc3d::add
 * src/PARAPROBE_NPointCorrelations3D.cpp:165-207), by brute force.
 *
 * For every considered center (ion2surf >= fl32(R*R)) and every task whose
 * central types hold its type: when more than K ions of any type lie within R
 * (self excluded), the env-type neighbours within R are ordered by
 * d = sqrtf((u*u + v*v) + w*w), (u, v, w) = neighbour - center, equal d by
 * the squared length, then the smaller u, v, w (the reference leaves ties to
 * nth_element); the
 * K-th one is binned when d <= R: per component t = u * (1.f / width), index
 * (int)(t + 0.5) or (int)(t - 0.5) in double, plus half.  counts is
 * [ntasks][edge^3], x fastest.
 */
typedef struct {
    float d, d2, u, v, w;
} onpc;

static int onpc_cmp(const void *pa, const void *pb)
{
    const onpc *a = (const onpc *)pa, *b = (const onpc *)pb;
    if (a->d != b->d)
        return a->d < b->d ? -1 : 1;
    if (a->d2 != b->d2)
        return a->d2 < b->d2 ? -1 : 1;
    if (a->u != b->u)
        return a->u < b->u ? -1 : 1;
    if (a->v != b->v)
        return a->v < b->v ? -1 : 1;
    if (a->w != b->w)
        return a->w < b->w ? -1 : 1;
    return 0;
}

int pp_oracle_npcorr(int64_t n, const float *xyz, const uint8_t *type, const float *ion2surf,
                     int32_t ntasks, const uint8_t *trg, const uint8_t *env, float rmax,
                     float rincr, int64_t K, uint32_t *counts)
{
    const float realhalf = ceilf(rmax / rincr) + 1.f;
    const int64_t half = (int64_t)realhalf, edge = 1 + 2 * half;
    const int64_t nxyz = edge * edge * edge;
    const float R2 = rmax * rmax, inv = 1.f / rincr;
    memset(counts, 0, sizeof(uint32_t) * (size_t)(ntasks * nxyz));
    onpc *buf = (onpc *)malloc(sizeof(onpc) * (size_t)(n > 0 ? n : 1));
    if (!buf)
        return -1;
    for (int64_t i = 0; i < n; ++i) {
        const float *p = xyz + 3 * i;
        if (!(ion2surf[i] >= R2))
            continue;
        int64_t ntot = 0;
        for (int64_t j = 0; j < n; ++j) {
            if (j == i)
                continue;
            const float dx = p[0] - xyz[3 * j], dy = p[1] - xyz[3 * j + 1], dz = p[2] - xyz[3 * j + 2];
            const float d2 = dx * dx + dy * dy + dz * dz;
            ntot += d2 <= R2;
        }
        if (ntot <= K)
            continue;
        for (int32_t t = 0; t < ntasks; ++t) {
            if (!((trg[32 * t + type[i] / 8] >> (type[i] % 8)) & 1))
                continue;
            int64_t m = 0;
            for (int64_t j = 0; j < n; ++j) {
                if (j == i || !((env[32 * t + type[j] / 8] >> (type[j] % 8)) & 1))
                    continue;
                const float dx = p[0] - xyz[3 * j], dy = p[1] - xyz[3 * j + 1], dz = p[2] - xyz[3 * j + 2];
                const float d2 = dx * dx + dy * dy + dz * dz;
                if (!(d2 <= R2))
                    continue;
                onpc e;
                e.u = xyz[3 * j] - p[0];
                e.v = xyz[3 * j + 1] - p[1];
                e.w = xyz[3 * j + 2] - p[2];
                e.d2 = e.u * e.u + e.v * e.v + e.w * e.w;
                e.d = sqrtf(e.d2);
                buf[m++] = e;
            }
            if (m <= K)
                continue;
            qsort(buf, (size_t)m, sizeof(onpc), onpc_cmp);
            const onpc k = buf[K];
            if (!(k.d <= rmax))
                continue;
            const float c[3] = {k.u * inv, k.v * inv, k.w * inv};
            int64_t ix[3];
            for (int q = 0; q < 3; ++q) {
                const int disc = (c[q] >= 0.f) ? (int)((double)c[q] + 0.5) : (int)((double)c[q] - 0.5);
                ix[q] = disc + half;
            }
            counts[t * nxyz + ix[0] + ix[1] * edge + ix[2] * edge * edge] += 1;
        }
    }
    free(buf);
    return 0;
}

void pp_oracle_pos_decode(const uint8_t *rec, int64_t n, const float *lo, const float *hi,
                          const uint8_t *rtype, int32_t nranges, float *xyz, float *mq,
                          uint8_t *type)
{
    orc_records_decode(rec, n, 16, lo, hi, rtype, nranges, xyz, mq, type);
}

/* EPOS (src/PARAPROBE_EPOSEndianness.h:86-160): 44-byte records whose first
 * four big-endian floats are x, y, z and the mass-to-charge. */
void pp_oracle_epos_decode(const uint8_t *rec, int64_t n, const float *lo, const float *hi,
                           const uint8_t *rtype, int32_t nranges, float *xyz, float *mq,
                           uint8_t *type)
{
    orc_records_decode(rec, n, 44, lo, hi, rtype, nranges, xyz, mq, type);
}
