/*
 * CPU restatement of FastPGM's PC-stable conditional-independence stage.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links or calls it.
 *
 * Reference files restated (paths relative to the FastPGM source tree):
 *   counts      src/fastpgm/structure_learning/CellTable.cpp:23-91,226-291,430-455
 *   G^2 / df    src/fastpgm/structure_learning/IndependenceTest.cpp:65-155,295-364
 *   p-value     stats::pchisq (third-party StatsLib/gcem, un-vendored; restated
 *               from the mathematics as the regularized lower incomplete gamma)
 *   enumeration src/fastpgm/ChoiceGenerator.cpp:14-85
 *   skeleton    src/fastpgm/structure_learning/PCStable.cpp:84-232,265-626
 *   edges       src/fastpgm/Network.cpp:270-283, src/fastpgm/Edge.cpp:17-28
 *   MLE counts  src/fastpgm/parameter_learning/ParameterLearning.cpp:24-82
 *               (table index: src/fastpgm/PotentialTableBase.cpp:102-130)
 *
 * Data layout: uint8 column-major, column v occupies cols[v*stride .. v*stride+n_rows).
 * Value coding is the identity (value k of variable v is state k, k < card[v]).
 */
#ifndef FPGM_PC_ORACLE_H
#define FPGM_PC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    const uint8_t *cols;
    int64_t stride;
    int n_vars;
    int64_t n_rows;
    const int32_t *card;
} orc_data;

/* result of one CI test (IndependenceTest::Result + the statistic) */
typedef struct {
    double g2;
    double p_value;
    int32_t df;
    int32_t indep;
} orc_result;

typedef struct {
    int64_t tests_per_level[64];   /* reference-equivalent num_ci_test increments, gs=1 */
    int64_t edges_after_level[64]; /* network->num_edges after each level */
    int32_t n_levels;              /* number of levels executed (level 0 included) */
    int64_t num_ci_test;
    int64_t n_sepsets;
    int64_t sep_z_total;           /* total number of Z entries over all sepsets */
} orc_pc_stats;

int64_t orc_table_cells(const orc_data *D, int x, int y, const int32_t *z, int d);
void orc_count(const orc_data *D, int x, int y, const int32_t *z, int d, int32_t *table);
void orc_g2_table(const int32_t *n, int cx, int cy, int64_t dimz, double *g2, int32_t *df);
double orc_pchisq(double x, int32_t df);
double orc_chisq_crit(int32_t df, double alpha);
orc_result orc_ci_test(const orc_data *D, int x, int y, const int32_t *z, int d, double alpha);
int orc_ci_batch(const orc_data *D, const int32_t *tests, int64_t n_tests, int max_d,
                 double alpha, int n_threads, orc_result *out);

/* lexicographic enumeration: writes every b-subset of {0..a-1} in ChoiceGenerator order */
int64_t orc_choose(int a, int b);
int64_t orc_enumerate(int a, int b, int32_t *out, int64_t cap);
void orc_unrank(int a, int b, int64_t rank, int32_t *out);

/*
 * Whole CI stage.  adj_out: n*n uint8 adjacency after the stage.
 * Sepsets: one record per removed edge, in ascending (x,y) order:
 *   sep_xy[2*k], sep_xy[2*k+1] = (x,y) x<y; sep_level[k] = level it was removed at;
 *   sep_off[k]..sep_off[k+1] index into sep_z (ascending node ids).
 * The record arrays must hold n*(n-1)/2 entries; sep_z must hold sep_z_cap ints.
 * Returns 0, or -1 if sep_z_cap is too small, -2 on bad arguments.
 * max_tests_per_level (<=0 = unlimited) stops a level early, for bounded timing samples only.
 */
int orc_pc_skeleton(const orc_data *D, double alpha, int depth, int n_threads,
                    uint8_t *adj_out, int32_t *sep_xy, int32_t *sep_level,
                    int64_t *sep_off, int32_t *sep_z, int64_t sep_z_cap,
                    orc_pc_stats *st);

/* Bounded timing sample: level-0 tests over the first n_pairs lex pairs, on n_threads.
 * Returns the number of tests run. */
int64_t orc_level0_sample(const orc_data *D, double alpha, int64_t n_pairs, int n_threads,
                          int32_t *indep_out);

/* Bounded timing sample of level d over the first max_edges edges of adj_before. */
int64_t orc_level_sample(const orc_data *D, double alpha, int d, const uint8_t *adj_before, int64_t max_edges,
                         int n_threads, int64_t *removed);
/* The same sample under the reference's 128-edge batch schedule (PCStable.cpp:275-324). */
int64_t orc_level_sample_ref128(const orc_data *D, double alpha, int d, const uint8_t *adj_before,
                                int64_t max_edges, int n_threads, int64_t *removed);

/*
 * Maximum-likelihood tables of a known structure (ParameterLearning.cpp:24-69): node i
 * (i < n_nodes) with parents parents[par_off[i]..par_off[i+1]) (ascending ids); table
 * index value(i) * prod c(parents) + mixed-radix parent configuration, first parent most
 * significant (PotentialTableBase.cpp:102-130).  count_off (n_nodes+1) receives the
 * table offsets; counts (may be NULL for a size query) the tables back to back.
 * potentials (may be NULL) receives ((double)N(AB) + alpha) / ((double)N(B) + alpha * c(i))
 * per entry (ParameterLearning.cpp:72-81).  Returns the total number of cells.
 */
int64_t orc_mle(const orc_data *D, int n_nodes, const int32_t *par_off, const int32_t *parents, int alpha,
                int64_t *count_off, int32_t *counts, double *potentials);

#ifdef __cplusplus
}
#endif
#endif
