/*
 * CPU restatement of FastPGM's PC-stable CI stage (G^2 test, group size 1).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the product.
 * Compiled with -ffp-contract=off so FP64 arithmetic follows the reference's
 * x86-64 evaluation (no fused multiply-add) term by term.
 *
 * Each function names the reference lines (FastPGM tree) it restates.
 */
#include "pc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ counts */

/* dimz * dimx * dimy.  CellTable.cpp:39-51 (dimz = product of Z cardinalities). */
int64_t orc_table_cells(const orc_data *D, int x, int y, const int32_t *z, int d)
{
    int64_t dimz = 1;
    for (int j = 0; j < d; ++j) dimz *= D->card[z[j]];
    return dimz * D->card[x] * D->card[y];
}

/*
 * Dense table n[zi][xv][yv] with zi the mixed-radix index of Z, first Z
 * element most significant: zi = sum_j v_j * prod_{k>j} card(Z_k).
 * CellTable.cpp:254-291 (Counts3D::CountLevel1/CountLevelN) and
 * CellTable.cpp:430-441 (Counts2D::FillTable, d == 0).
 * Horner evaluation gives the same index as the cum_levels sum.
 */
void orc_count(const orc_data *D, int x, int y, const int32_t *z, int d, int32_t *table)
{
    const int64_t cells = orc_table_cells(D, x, y, z, d);
    memset(table, 0, (size_t)cells * sizeof(int32_t));
    const int cx = D->card[x], cy = D->card[y];
    const uint8_t *colx = D->cols + (int64_t)x * D->stride;
    const uint8_t *coly = D->cols + (int64_t)y * D->stride;
    enum { CH = 4096 };
    uint32_t key[CH];
    for (int64_t r0 = 0; r0 < D->n_rows; r0 += CH) {
        const int64_t m = (D->n_rows - r0 < CH) ? D->n_rows - r0 : CH;
        for (int64_t r = 0; r < m; ++r) key[r] = 0;
        for (int j = 0; j < d; ++j) {
            const uint8_t *cz = D->cols + (int64_t)z[j] * D->stride + r0;
            const uint32_t c = (uint32_t)D->card[z[j]];
            for (int64_t r = 0; r < m; ++r) key[r] = key[r] * c + cz[r];
        }
        for (int64_t r = 0; r < m; ++r) {
            const uint32_t k = (key[r] * (uint32_t)cx + colx[r0 + r]) * (uint32_t)cy + coly[r0 + r];
            table[k]++;
        }
    }
}

/*
 * G^2 and adjusted degrees of freedom from a dense table.
 * Marginals: CellTable.cpp:242-250.  df and G^2: IndependenceTest.cpp:94-138
 * (d >= 1) and :306-347 (d == 0, where total = num_instance = sum of the table).
 * Term order (z, x, y ascending), skips, and the expression
 * 2.0 * observed * log(observed / expected) with
 * expected = (double)sum_col * (double)sum_row / (double)total are kept as written.
 */
void orc_g2_table(const int32_t *n, int cx, int cy, int64_t dimz, double *g2_out, int32_t *df_out)
{
    int64_t ni[256], nj[256];
    double g2 = 0.0;
    int32_t df = 0;
    for (int64_t k = 0; k < dimz; ++k) {
        const int32_t *t = n + k * cx * cy;
        long nk = 0;
        for (int i = 0; i < cx; ++i) ni[i] = 0;
        for (int j = 0; j < cy; ++j) nj[j] = 0;
        for (int i = 0; i < cx; ++i)
            for (int j = 0; j < cy; ++j) {
                ni[i] += t[i * cy + j];
                nj[j] += t[i * cy + j];
                nk += t[i * cy + j];
            }
        int alx = 0, aly = 0;
        for (int i = 0; i < cx; ++i) alx += (ni[i] > 0);
        for (int j = 0; j < cy; ++j) aly += (nj[j] > 0);
        alx = (alx >= 1) ? alx : 1;
        aly = (aly >= 1) ? aly : 1;
        df += (alx - 1) * (aly - 1);
        const long total = nk;
        if (total == 0) continue;
        for (int i = 0; i < cx; ++i) {
            const long sum_row = ni[i];
            if (sum_row == 0) continue;
            for (int j = 0; j < cy; ++j) {
                const long sum_col = nj[j];
                const long observed = t[i * cy + j];
                if (sum_col == 0 || observed == 0) continue;
                const double expected = (double)sum_col * (double)sum_row / (double)total;
                g2 += 2.0 * observed * log(observed / expected);
            }
        }
    }
    *g2_out = g2;
    *df_out = df;
}

/* ------------------------------------------------------- chi-square CDF */

/*
 * Regularized lower incomplete gamma P(a, x), evaluated by the power series for
 * x < a + 1 and by the modified-Lentz continued fraction for Q = 1 - P otherwise
 * (Press et al., Numerical Recipes, 3rd ed., section 6.2).  Written from the
 * mathematics; stands in for the un-vendored gcem::incomplete_gamma.
 */
static double orc_gammp(double a, double x)
{
    if (!(x > 0.0)) return 0.0;
    const double eps = 1e-16;
    const double lpre = -x + a * log(x) - lgamma(a);
    if (x < a + 1.0) {
        double ap = a, del = 1.0 / a, sum = del;
        for (int it = 0; it < 1000000; ++it) {
            ap += 1.0;
            del *= x / ap;
            sum += del;
            if (fabs(del) < fabs(sum) * eps) break;
        }
        return sum * exp(lpre);
    }
    const double fpmin = 1e-300;
    double b = x + 1.0 - a, c = 1.0 / fpmin, dd = 1.0 / b, h = dd;
    for (int i = 1; i < 1000000; ++i) {
        const double an = -i * (i - a);
        b += 2.0;
        dd = an * dd + b;
        if (fabs(dd) < fpmin) dd = fpmin;
        c = b + an / c;
        if (fabs(c) < fpmin) c = fpmin;
        dd = 1.0 / dd;
        const double del = dd * c;
        h *= del;
        if (fabs(del - 1.0) < eps) break;
    }
    return 1.0 - exp(lpre) * h;
}

/* chi-square CDF with df degrees of freedom: stats::pchisq(x, df, false) = P(df/2, x/2). */
double orc_pchisq(double x, int32_t df)
{
    if (df <= 0) return (x > 0.0) ? 1.0 : 0.0;
    return orc_gammp(0.5 * df, 0.5 * x);
}

/* critical value c with 1 - pchisq(c, df) = alpha, by bisection to full precision */
double orc_chisq_crit(int32_t df, double alpha)
{
    double lo = 0.0, hi = (double)df + 10.0;
    while (1.0 - orc_pchisq(hi, df) > alpha) hi *= 2.0;
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (1.0 - orc_pchisq(mid, df) > alpha) lo = mid; else hi = mid;
    }
    return 0.5 * (lo + hi);
}

/* ------------------------------------------------------------- one test */

/*
 * IndependenceTest::IndependenceResult for "g square", c_size 1
 * (IndependenceTest.cpp:35-58 dispatch; :140-154 / :349-363 decision):
 * df == 0 -> independent with p = 1; else p = 1 - pchisq(G^2, df), independent iff p > alpha.
 */
static orc_result orc_ci_test_buf(const orc_data *D, int x, int y, const int32_t *z, int d,
                                  double alpha, int32_t *table)
{
    orc_result r;
    orc_count(D, x, y, z, d, table);
    int64_t dimz = 1;
    for (int j = 0; j < d; ++j) dimz *= D->card[z[j]];
    orc_g2_table(table, D->card[x], D->card[y], dimz, &r.g2, &r.df);
    if (r.df == 0) {
        r.p_value = 1.0;
        r.indep = 1;
        return r;
    }
    r.p_value = 1.0 - orc_pchisq(r.g2, r.df);
    r.indep = (r.p_value > alpha) ? 1 : 0;
    return r;
}

orc_result orc_ci_test(const orc_data *D, int x, int y, const int32_t *z, int d, double alpha)
{
    const int64_t cells = orc_table_cells(D, x, y, z, d);
    int32_t *table = (int32_t *)malloc((size_t)cells * sizeof(int32_t));
    orc_result r = orc_ci_test_buf(D, x, y, z, d, alpha, table);
    free(table);
    return r;
}

/* tests: n_tests rows of (x, y, d, z[0..max_d-1]) int32 */
int orc_ci_batch(const orc_data *D, const int32_t *tests, int64_t n_tests, int max_d,
                 double alpha, int n_threads, orc_result *out)
{
    const int w = 3 + max_d;
#pragma omp parallel num_threads(n_threads > 0 ? n_threads : 1)
    {
        int64_t cap = 0;
        int32_t *table = NULL;
#pragma omp for schedule(dynamic, 4)
        for (int64_t t = 0; t < n_tests; ++t) {
            const int32_t *tr = tests + t * w;
            const int64_t cells = orc_table_cells(D, tr[0], tr[1], tr + 3, tr[2]);
            if (cells > cap) {
                free(table);
                cap = cells;
                table = (int32_t *)malloc((size_t)cap * sizeof(int32_t));
            }
            out[t] = orc_ci_test_buf(D, tr[0], tr[1], tr + 3, tr[2], alpha, table);
        }
        free(table);
    }
    return 0;
}

/* ------------------------------------------------------------ enumeration */

int64_t orc_choose(int a, int b)
{
    if (b < 0 || a < b) return 0;
    if (b > a - b) b = a - b;
    int64_t r = 1;
    for (int i = 1; i <= b; ++i) r = r * (a - b + i) / i;
    return r;
}

/*
 * The ChoiceGenerator series (ChoiceGenerator.cpp:14-85): lexicographic
 * b-subsets of {0..a-1}.  Next(): scan from the right for the first index i with
 * choice[i] < i + (a - b), increment it and refill the tail with successors.
 */
static int orc_next_choice(int a, int b, int32_t *c)
{
    for (int i = b - 1; i >= 0; --i) {
        if (c[i] < i + (a - b)) {
            c[i]++;
            for (int j = i + 1; j < b; ++j) c[j] = c[j - 1] + 1;
            return 1;
        }
    }
    return 0;
}

int64_t orc_enumerate(int a, int b, int32_t *out, int64_t cap)
{
    if (b < 1 || a < b) return 0;
    int32_t c[256];
    for (int i = 0; i < b; ++i) c[i] = i;
    int64_t k = 0;
    do {
        if (k < cap) memcpy(out + k * b, c, (size_t)b * sizeof(int32_t));
        ++k;
    } while (orc_next_choice(a, b, c));
    return k;
}

/* rank -> rank-th lexicographic b-subset (combinatorial number system) */
void orc_unrank(int a, int b, int64_t rank, int32_t *out)
{
    int v = 0;
    for (int i = 0; i < b; ++i) {
        for (;;) {
            const int64_t cnt = orc_choose(a - v - 1, b - i - 1);
            if (rank < cnt) break;
            rank -= cnt;
            ++v;
        }
        out[i] = v++;
    }
}

/* ------------------------------------------------------------- skeleton */

typedef struct {
    int32_t removed;
    int32_t tests;
    int32_t z[64];
} orc_edge_res;

/*
 * One edge at level d (PCStable.cpp:399-494 CheckEdge, :500-515 FindAdjacencies,
 * :526-614 Testing with group_size 1): S_x = adj(x)\{y} ascending, every d-subset
 * in lexicographic order, stop at the first independent one; then, if the edge
 * survives, the same on S_y = adj(y)\{x}.  A side with fewer than d neighbours is
 * skipped.  Every test run counts toward num_ci_test (:561).
 */
typedef struct {
    int32_t *p;
    int64_t cap;
} orc_buf;

static int32_t *orc_buf_get(orc_buf *b, int64_t cells)
{
    if (cells > b->cap) {
        free(b->p);
        b->cap = cells;
        b->p = (int32_t *)malloc((size_t)cells * sizeof(int32_t));
    }
    return b->p;
}

static void orc_check_edge(const orc_data *D, double alpha, int d, int x, int y,
                           const int32_t *nb_off, const int32_t *nb, orc_buf *tb,
                           int32_t *sbuf, orc_edge_res *res)
{
    res->removed = 0;
    res->tests = 0;
    for (int side = 0; side < 2; ++side) {
        const int u = side == 0 ? x : y, other = side == 0 ? y : x;
        int a = 0;
        for (int32_t k = nb_off[u]; k < nb_off[u + 1]; ++k)
            if (nb[k] != other) sbuf[a++] = nb[k];
        if (a < d) continue;
        int32_t c[64], zz[64];
        for (int i = 0; i < d; ++i) c[i] = i;
        do {
            for (int i = 0; i < d; ++i) zz[i] = sbuf[c[i]];
            res->tests++;
            int32_t *tbl = orc_buf_get(tb, orc_table_cells(D, x, y, zz, d));
            orc_result r = orc_ci_test_buf(D, x, y, zz, d, alpha, tbl);
            if (r.indep) {
                res->removed = 1;
                memcpy(res->z, zz, (size_t)d * sizeof(int32_t));
                return;
            }
        } while (orc_next_choice(a, d, c));
    }
}

typedef struct {
    int32_t x, y, level;
    int64_t zoff;
    int32_t d;
} orc_sep_tmp;

static int orc_sep_cmp(const void *pa, const void *pb)
{
    const orc_sep_tmp *a = (const orc_sep_tmp *)pa, *b = (const orc_sep_tmp *)pb;
    if (a->x != b->x) return a->x < b->x ? -1 : 1;
    if (a->y != b->y) return a->y < b->y ? -1 : 1;
    return 0;
}

/*
 * StructLearnByPCStable's step 1 (PCStable.cpp:109-232) with SearchAtDepth
 * (:265-388): level 0 tests every pair of the complete graph (lexicographic edge
 * order, Network.cpp:270-283); level d >= 1 tests the surviving edges against a
 * frozen copy of the adjacency (:272); removals are applied after the level
 * (:369-386); the loop continues while FreeDegree (:620-626, max degree - 1 after
 * removal) exceeds d, for d < depth (:209).
 */
int orc_pc_skeleton(const orc_data *D, double alpha, int depth, int n_threads,
                    uint8_t *adj_out, int32_t *sep_xy, int32_t *sep_level,
                    int64_t *sep_off, int32_t *sep_z, int64_t sep_z_cap,
                    orc_pc_stats *st)
{
    const int n = D->n_vars;
    if (n < 0 || alpha < 0.0 || alpha > 1.0 || depth < 0) return -2;
    if (n_threads < 1) n_threads = 1;
    memset(st, 0, sizeof(*st));
    uint8_t *adj = adj_out;
    memset(adj, 0, (size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) adj[(int64_t)i * n + j] = (i != j);

    const int64_t e0 = (int64_t)n * (n - 1) / 2;
    int32_t *ex = (int32_t *)malloc(sizeof(int32_t) * (e0 > 0 ? e0 : 1));
    int32_t *ey = (int32_t *)malloc(sizeof(int32_t) * (e0 > 0 ? e0 : 1));
    orc_edge_res *res = (orc_edge_res *)malloc(sizeof(orc_edge_res) * (e0 > 0 ? e0 : 1));
    orc_sep_tmp *seps = (orc_sep_tmp *)malloc(sizeof(orc_sep_tmp) * (e0 > 0 ? e0 : 1));
    int32_t *zpool = (int32_t *)malloc(sizeof(int32_t) * (sep_z_cap > 0 ? sep_z_cap : 1));
    int64_t n_edges = 0, n_seps = 0, zused = 0;
    int rc = 0;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            ex[n_edges] = i;
            ey[n_edges] = j;
            ++n_edges;
        }

    int64_t max_cells = 1;
    for (int v = 0; v < n; ++v) max_cells = D->card[v] > max_cells ? D->card[v] : max_cells;
    max_cells = max_cells * max_cells;

    /* level 0: PCStable.cpp:133-195 */
#pragma omp parallel num_threads(n_threads)
    {
        int32_t *tbl = (int32_t *)malloc(sizeof(int32_t) * max_cells);
#pragma omp for schedule(dynamic, 8)
        for (int64_t e = 0; e < n_edges; ++e) {
            orc_result r = orc_ci_test_buf(D, ex[e], ey[e], NULL, 0, alpha, tbl);
            res[e].removed = r.indep;
            res[e].tests = 1;
        }
        free(tbl);
    }
    st->tests_per_level[0] = n_edges;
    {
        int64_t w = 0;
        for (int64_t e = 0; e < n_edges; ++e) {
            if (res[e].removed) {
                seps[n_seps].x = ex[e];
                seps[n_seps].y = ey[e];
                seps[n_seps].level = 0;
                seps[n_seps].d = 0;
                seps[n_seps].zoff = zused;
                ++n_seps;
                adj[(int64_t)ex[e] * n + ey[e]] = 0;
                adj[(int64_t)ey[e] * n + ex[e]] = 0;
            } else {
                ex[w] = ex[e];
                ey[w] = ey[e];
                ++w;
            }
        }
        n_edges = w;
    }
    st->edges_after_level[0] = n_edges;
    st->n_levels = 1;

    int32_t *nb_off = (int32_t *)malloc(sizeof(int32_t) * (n + 1));
    int32_t *nb = (int32_t *)malloc(sizeof(int32_t) * ((int64_t)n * n + 1));
    for (int d = 1; d < depth && d < 64; ++d) {
        /* frozen copy of the adjacency, ascending neighbour lists */
        nb_off[0] = 0;
        for (int i = 0; i < n; ++i) {
            int32_t k = nb_off[i];
            for (int j = 0; j < n; ++j)
                if (adj[(int64_t)i * n + j]) nb[k++] = j;
            nb_off[i + 1] = k;
        }
#pragma omp parallel num_threads(n_threads)
        {
            orc_buf tb = {NULL, 0};
            int32_t *sbuf = (int32_t *)malloc(sizeof(int32_t) * (n + 1));
#pragma omp for schedule(dynamic, 1)
            for (int64_t e = 0; e < n_edges; ++e)
                orc_check_edge(D, alpha, d, ex[e], ey[e], nb_off, nb, &tb, sbuf, &res[e]);
            free(sbuf);
            free(tb.p);
        }
        int64_t tests = 0, w = 0;
        for (int64_t e = 0; e < n_edges; ++e) {
            tests += res[e].tests;
            if (res[e].removed) {
                if (zused + d > sep_z_cap) {
                    rc = -1;
                } else {
                    memcpy(zpool + zused, res[e].z, sizeof(int32_t) * d);
                }
                seps[n_seps].x = ex[e];
                seps[n_seps].y = ey[e];
                seps[n_seps].level = d;
                seps[n_seps].d = d;
                seps[n_seps].zoff = zused;
                zused += d;
                ++n_seps;
                adj[(int64_t)ex[e] * n + ey[e]] = 0;
                adj[(int64_t)ey[e] * n + ex[e]] = 0;
            } else {
                ex[w] = ex[e];
                ey[w] = ey[e];
                ++w;
            }
        }
        n_edges = w;
        st->tests_per_level[d] = tests;
        st->edges_after_level[d] = n_edges;
        st->n_levels = d + 1;
        int maxdeg = 0;
        for (int i = 0; i < n; ++i) {
            int deg = 0;
            for (int j = 0; j < n; ++j) deg += adj[(int64_t)i * n + j];
            maxdeg = deg > maxdeg ? deg : maxdeg;
        }
        if (!(maxdeg - 1 > d)) break;
    }

    qsort(seps, (size_t)n_seps, sizeof(orc_sep_tmp), orc_sep_cmp);
    int64_t zo = 0;
    for (int64_t k = 0; k < n_seps; ++k) {
        sep_xy[2 * k] = seps[k].x;
        sep_xy[2 * k + 1] = seps[k].y;
        sep_level[k] = seps[k].level;
        sep_off[k] = zo;
        if (rc == 0) memcpy(sep_z + zo, zpool + seps[k].zoff, sizeof(int32_t) * seps[k].d);
        zo += seps[k].d;
    }
    sep_off[n_seps] = zo;
    st->n_sepsets = n_seps;
    st->sep_z_total = zo;
    int64_t total = 0;
    for (int l = 0; l < st->n_levels; ++l) total += st->tests_per_level[l];
    st->num_ci_test = total;

    free(nb);
    free(nb_off);
    free(zpool);
    free(seps);
    free(res);
    free(ey);
    free(ex);
    return rc;
}

int64_t orc_level0_sample(const orc_data *D, double alpha, int64_t n_pairs, int n_threads,
                          int32_t *indep_out)
{
    const int n = D->n_vars;
    const int64_t e0 = (int64_t)n * (n - 1) / 2;
    if (n_pairs > e0) n_pairs = e0;
    int maxc = 1;
    for (int v = 0; v < n; ++v) maxc = D->card[v] > maxc ? D->card[v] : maxc;
#pragma omp parallel num_threads(n_threads > 0 ? n_threads : 1)
    {
        int32_t *tbl = (int32_t *)malloc(sizeof(int32_t) * maxc * maxc);
#pragma omp for schedule(dynamic, 8)
        for (int64_t e = 0; e < n_pairs; ++e) {
            /* invert the lexicographic edge index (Network.cpp:278) */
            int i = 0;
            int64_t rem = e;
            while (rem >= n - 1 - i) {
                rem -= n - 1 - i;
                ++i;
            }
            const int j = i + 1 + (int)rem;
            orc_result r = orc_ci_test_buf(D, i, j, NULL, 0, alpha, tbl);
            if (indep_out) indep_out[e] = r.indep;
        }
        free(tbl);
    }
    return n_pairs;
}

/*
 * Bounded timing sample of one level: CheckEdge (PCStable.cpp:399-494) for the first
 * max_edges edges (lexicographic order) of the adjacency adj_before, which is frozen for
 * the level as at PCStable.cpp:272.  Level 0 tests the pairs themselves.  Returns the
 * number of CI tests run; *removed receives the number of edges found independent.
 */
int64_t orc_level_sample(const orc_data *D, double alpha, int d, const uint8_t *adj_before, int64_t max_edges,
                         int n_threads, int64_t *removed)
{
    const int n = D->n_vars;
    int64_t ne = 0;
    int32_t *ex = (int32_t *)malloc(sizeof(int32_t) * (max_edges > 0 ? max_edges : 1));
    int32_t *ey = (int32_t *)malloc(sizeof(int32_t) * (max_edges > 0 ? max_edges : 1));
    for (int i = 0; i < n && ne < max_edges; ++i)
        for (int j = i + 1; j < n && ne < max_edges; ++j)
            if (adj_before[(int64_t)i * n + j]) {
                ex[ne] = i;
                ey[ne] = j;
                ++ne;
            }
    int32_t *nb_off = (int32_t *)malloc(sizeof(int32_t) * (n + 1));
    int32_t *nb = (int32_t *)malloc(sizeof(int32_t) * ((int64_t)n * n + 1));
    nb_off[0] = 0;
    for (int i = 0; i < n; ++i) {
        int32_t k = nb_off[i];
        for (int j = 0; j < n; ++j)
            if (adj_before[(int64_t)i * n + j]) nb[k++] = j;
        nb_off[i + 1] = k;
    }
    int64_t tests = 0, rem = 0;
#pragma omp parallel num_threads(n_threads > 0 ? n_threads : 1) reduction(+ : tests, rem)
    {
        orc_buf tb = {NULL, 0};
        int32_t *sbuf = (int32_t *)malloc(sizeof(int32_t) * (n + 1));
        orc_edge_res r;
#pragma omp for schedule(dynamic, 1)
        for (int64_t e = 0; e < ne; ++e) {
            if (d == 0) {
                int32_t *tbl = orc_buf_get(&tb, (int64_t)D->card[ex[e]] * D->card[ey[e]]);
                orc_result q = orc_ci_test_buf(D, ex[e], ey[e], NULL, 0, alpha, tbl);
                tests += 1;
                rem += q.indep;
            } else {
                orc_check_edge(D, alpha, d, ex[e], ey[e], nb_off, nb, &tb, sbuf, &r);
                tests += r.tests;
                rem += r.removed;
            }
        }
        free(sbuf);
        free(tb.p);
    }
    free(nb);
    free(nb_off);
    free(ey);
    free(ex);
    if (removed) *removed = rem;
    return tests;
}

/*
 * The same bounded sample under the reference's own schedule of a level (SearchAtDepth,
 * PCStable.cpp:275-324, group size 1): the edges on a stack, 128 popped at a time and
 * checked under one `omp parallel for` (static schedule), each check running one CI test
 * of its edge -- the next conditioning set in CheckEdge's order (x side, then y side) --
 * and the unfinished edges pushed back, until the stack is empty; the last 1..127 edges
 * with min(size, threads) threads.  Test counts and removals equal orc_level_sample's;
 * only the schedule (a barrier every 128 tests) differs.  Level 0 has no such loop.
 */
typedef struct {
    int side;      /* 0: subsets of adj(x)\{y}; 1: of adj(y)\{x}; 2: exhausted */
    int started;
    int a;
    int32_t c[64];
    int32_t *s;    /* the current side's neighbour set */
} orc_walk;

static int orc_walk_side(orc_walk *w, int d, int x, int y, const int32_t *nb_off, const int32_t *nb, int side)
{
    for (; side < 2; ++side) {
        const int u = side == 0 ? x : y, other = side == 0 ? y : x;
        int a = 0;
        for (int32_t k = nb_off[u]; k < nb_off[u + 1]; ++k)
            if (nb[k] != other) w->s[a++] = nb[k];
        if (a >= d) {
            w->side = side;
            w->a = a;
            for (int i = 0; i < d; ++i) w->c[i] = i;
            return 1;
        }
    }
    w->side = 2;
    return 0;
}

int64_t orc_level_sample_ref128(const orc_data *D, double alpha, int d, const uint8_t *adj_before,
                                int64_t max_edges, int n_threads, int64_t *removed)
{
    if (d == 0) return orc_level_sample(D, alpha, d, adj_before, max_edges, n_threads, removed);
    const int n = D->n_vars, nt = n_threads > 0 ? n_threads : 1;
    const int64_t cap = max_edges > 0 ? max_edges : 1;
    int32_t *ex = (int32_t *)malloc(sizeof(int32_t) * cap), *ey = (int32_t *)malloc(sizeof(int32_t) * cap);
    int64_t ne = 0;
    for (int i = 0; i < n && ne < max_edges; ++i)
        for (int j = i + 1; j < n && ne < max_edges; ++j)
            if (adj_before[(int64_t)i * n + j]) {
                ex[ne] = i;
                ey[ne] = j;
                ++ne;
            }
    int32_t *nb_off = (int32_t *)malloc(sizeof(int32_t) * (n + 1));
    int32_t *nb = (int32_t *)malloc(sizeof(int32_t) * ((int64_t)n * n + 1));
    nb_off[0] = 0;
    for (int i = 0; i < n; ++i) {
        int32_t k = nb_off[i];
        for (int j = 0; j < n; ++j)
            if (adj_before[(int64_t)i * n + j]) nb[k++] = j;
        nb_off[i + 1] = k;
    }
    orc_walk *wk = (orc_walk *)calloc((size_t)cap, sizeof(orc_walk));
    for (int64_t e = 0; e < ne; ++e) wk[e].s = (int32_t *)malloc(sizeof(int32_t) * (n + 1));
    orc_buf *tb = (orc_buf *)calloc((size_t)nt, sizeof(orc_buf));
    int64_t *stack = (int64_t *)malloc(sizeof(int64_t) * cap), top = 0;
    for (int64_t e = ne - 1; e >= 0; --e) stack[top++] = e;  /* edge 0 on top */
    int64_t proc[128];
    int need[128];
    int64_t tests = 0, rem = 0;
    while (top > 0) {
        const int size = top >= 128 ? 128 : (int)top;
        for (int i = 0; i < size; ++i) proc[i] = stack[--top];
        const int p = (size < 128 && size < nt) ? size : nt;
#pragma omp parallel for num_threads(p) reduction(+ : tests, rem)
        for (int i = 0; i < size; ++i) {
            orc_walk *w = &wk[proc[i]];
            const int x = ex[proc[i]], y = ey[proc[i]];
            need[i] = 0;
            if (!w->started) {
                w->started = 1;
                if (!orc_walk_side(w, d, x, y, nb_off, nb, 0)) continue;
            }
            if (w->side >= 2) continue;
            int32_t zz[64];
            for (int k = 0; k < d; ++k) zz[k] = w->s[w->c[k]];
            orc_buf *b = &tb[omp_get_thread_num()];
            int32_t *tbl = orc_buf_get(b, orc_table_cells(D, x, y, zz, d));
            const orc_result r = orc_ci_test_buf(D, x, y, zz, d, alpha, tbl);
            tests += 1;
            if (r.indep) {
                rem += 1;
                w->side = 2;
                continue;
            }
            if (!orc_next_choice(w->a, d, w->c) && !orc_walk_side(w, d, x, y, nb_off, nb, w->side + 1)) continue;
            need[i] = 1;
        }
        for (int i = size - 1; i >= 0; --i)
            if (need[i]) stack[top++] = proc[i];
    }
    for (int i = 0; i < nt; ++i) free(tb[i].p);
    for (int64_t e = 0; e < ne; ++e) free(wk[e].s);
    free(tb);
    free(wk);
    free(stack);
    free(nb);
    free(nb_off);
    free(ey);
    free(ex);
    if (removed) *removed = rem;
    return tests;
}

/* ------------------------------------------------------------------ MLE counts */

/*
 * ParameterLearning::LearnParamsKnowStructCompData (ParameterLearning.cpp:24-82), the
 * counting and the Laplace-smoothed estimate, for nodes without the virtual ROOT parent:
 * N(AB)[index] over every row, N(B)[index % cum_levels[0]] with cum_levels[0] = the
 * parents' table size (ParameterLearning.cpp:57-68), potential = (N(AB) + alpha) /
 * (N(B) + alpha * c(node)) (:72-81).  A node without parents has N(B) = number of rows
 * (:35-40, :75-76).
 */
int64_t orc_mle(const orc_data *D, int n_nodes, const int32_t *par_off, const int32_t *parents, int alpha,
                int64_t *count_off, int32_t *counts, double *potentials)
{
    int64_t total = 0;
    for (int i = 0; i < n_nodes; ++i) {
        int64_t size = D->card[i];
        for (int32_t j = par_off[i]; j < par_off[i + 1]; ++j) size *= D->card[parents[j]];
        count_off[i] = total;
        total += size;
    }
    count_off[n_nodes] = total;
    if (!counts) return total;
    for (int i = 0; i < n_nodes; ++i) {
        int32_t *t = counts + count_off[i];
        const int64_t size = count_off[i + 1] - count_off[i];
        const int64_t cum0 = size / D->card[i];  /* cum_levels[0] */
        memset(t, 0, (size_t)size * sizeof(int32_t));
        for (int64_t r = 0; r < D->n_rows; ++r) {
            int64_t idx = D->cols[(int64_t)i * D->stride + r];
            for (int32_t j = par_off[i]; j < par_off[i + 1]; ++j)
                idx = idx * D->card[parents[j]] + D->cols[(int64_t)parents[j] * D->stride + r];
            t[idx]++;
        }
        if (!potentials) continue;
        int32_t *nb = (int32_t *)calloc((size_t)cum0, sizeof(int32_t));
        for (int64_t k = 0; k < size; ++k) nb[k % cum0] += t[k];
        for (int64_t k = 0; k < size; ++k)
            potentials[count_off[i] + k] =
                ((double)t[k] + alpha) / ((double)nb[k % cum0] + alpha * D->card[i]);
        free(nb);
    }
    return total;
}
