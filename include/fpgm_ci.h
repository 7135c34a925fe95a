/*
 * fpgm_ci.h -- C ABI of the MI355X (gfx950) PC-stable conditional-independence stage.
 *
 * A drop-in for the CI stage of FastPGM's PC-stable skeleton search.  Every entry
 * point below names the reference interface it replaces (paths relative to the
 * FastPGM source tree).  Plain C types only; no allocation crosses the ABI: the
 * caller owns every host buffer, the context owns every device buffer.
 *
 * Data: n_vars uint8 columns, column v at cols + v*row_stride, n_rows valid bytes
 * each (identity value coding, value < card[v] <= 255).  This is the reference's
 * Dataset::dataset_columns (Dataset.h:32) narrowed from int32 to uint8.
 *
 * Errors are returned as fpgm_status codes; nothing exits or throws across the ABI
 * (the reference calls exit(1): IndependenceTest.cpp:13-17, ChoiceGenerator.cpp:15-18).
 * Calls on one context are serialised by the context; distinct contexts may be used
 * from different threads.
 */
#ifndef FPGM_CI_H
#define FPGM_CI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t fpgm_status;
#define FPGM_OK 0
#define FPGM_EINVAL 1   /* bad argument (alpha outside [0,1], bad index, value >= card ...) */
#define FPGM_ENOMEM 2   /* device or host allocation failed                                */
#define FPGM_EDEVICE 3  /* HIP runtime error or no gfx950 device                           */
#define FPGM_ETOOBIG 4  /* a contingency table exceeds the dense-table limit               */
#define FPGM_ETIMEOUT 5 /* a level exceeded the caller's round limit                        */

typedef struct fpgm_ctx fpgm_ctx;

/* One CI test X _||_ Y | Z.  z points into the caller's array of d node ids, in the
 * order the reference passes them (ascending node id for PC-stable). */
typedef struct {
    int32_t x;
    int32_t y;
    int32_t d;
    int32_t z_off; /* offset of Z in the z pool passed with the batch */
} fpgm_test;

/* IndependenceTest::Result (IndependenceTest.h:38-49) plus the statistic. */
typedef struct {
    double g2;        /* G^2: FP64 when exact/rechecked, else the FP32 value widened */
    double p_value;   /* 1 - pchisq(g2, df); NaN when the FP32 path decided alone    */
    int32_t df;       /* adjusted degrees of freedom (IndependenceTest.cpp:94-110)    */
    uint8_t indep;    /* p_value > alpha, or df == 0                                  */
    uint8_t in_band;  /* FP32 statistic fell in the recheck band, FP64 decided       */
    uint8_t exact;    /* g2/p_value are FP64                                          */
    uint8_t pad;
} fpgm_result;

/* Options of the whole-stage entry point. */
typedef struct {
    int32_t mode;          /* FPGM_MODE_BATCHED (default) or FPGM_MODE_PLAIN             */
    int32_t n_threads;     /* accepted for call-shape parity; results never depend on it */
    int32_t group_size;    /* accepted for call-shape parity; results never depend on it */
    int32_t verbose;       /* 0 silent, 1 per-level lines on stderr                       */
    int64_t round_tests;   /* batched: target tests per speculative round (0 = auto)      */
    int64_t max_rounds;    /* batched: per-level round cap, 0 = none (FPGM_ETIMEOUT)      */
    int64_t plain_sample;  /* plain mode: > 0 runs only the first plain_sample tests of    */
                           /* each level one launch per test and the rest batched, so the */
                           /* plain rate is sampled while the results stay exact           */
    uint32_t flags;        /* FPGM_F_* below                                              */
    float band;            /* relative half-width of the FP64 recheck band (0 = 0.02)     */
} fpgm_pc_opts;

#define FPGM_MODE_BATCHED 0
#define FPGM_MODE_PLAIN 1

#define FPGM_F_EXACT 1u     /* compute FP64 G^2 and p-value for every test              */
#define FPGM_F_FP32ONLY 2u  /* diagnostics: decide every test from FP32 alone           */
#define FPGM_F_NO_GRAM 4u   /* level 0 by the counting kernel, not the MFMA Gram path    */

/* Per-level statistics of the last fpgm_pc_skeleton call. */
typedef struct {
    int64_t num_ci_test;            /* reference-equivalent tests (group size 1)          */
    int64_t tests_per_level[64];
    int64_t edges_after_level[64];
    int64_t executed_per_level[64]; /* tests the device ran, speculative ones included    */
    int64_t in_band_per_level[64];  /* tests decided by the FP64 recheck                  */
    int64_t rounds_per_level[64];
    double seconds_per_level[64];
    int64_t plain_tests_per_level[64];   /* tests run one launch per test (plain mode)     */
    double plain_seconds_per_level[64];  /* wall time of those single-test launches        */
    double kernel_seconds_per_level[64]; /* counting-kernel time (HIP events)               */
    int32_t n_levels;
    int32_t pad;
    int64_t n_sepsets;
    int64_t sep_z_total;
    double ci_seconds;              /* the "pc-stable step 1" region, PCStable.cpp:109-232 */
    double max_rel_err_g2_fp32;     /* over rechecked tests: |g2_32 - g2_64| / g2_64       */
} fpgm_stats;

/* Sepset record: PCStable::sepset entry (PCStable.h:35), key (x,y) with x<y. */
typedef struct {
    int32_t x;
    int32_t y;
    int32_t level;
    int32_t d;
    int64_t z_off; /* into the z array returned by fpgm_sepset_get */
} fpgm_sepset;

/*
 * Context.  Replaces handing a Dataset* to IndependenceTest / PCStable
 * (IndependenceTest.cpp:12-21, PCStable.cpp:25): uploads the columns once.
 * fpgm_ctx_create copies host columns to the device; fpgm_ctx_create_device
 * borrows columns already resident in device memory (the caller keeps them alive).
 */
fpgm_status fpgm_ctx_create(const uint8_t *cols, int32_t n_vars, int64_t n_rows,
                            int64_t row_stride, const int32_t *card, fpgm_ctx **out);
fpgm_status fpgm_ctx_create_device(const void *d_cols, int32_t n_vars, int64_t n_rows,
                                   int64_t row_stride, const int32_t *card, fpgm_ctx **out);
/*
 * Context over several datasets with the same variables (e.g. bootstrap replicates):
 * instance i has n_rows[i] rows at cols[i] with stride row_stride[i].  Used by
 * fpgm_pc_skeleton_multi; the single-test calls need a one-instance context.
 */
fpgm_status fpgm_ctx_create_multi(const uint8_t *const *cols, const int64_t *n_rows, const int64_t *row_stride,
                                  int32_t n_inst, int32_t n_vars, const int32_t *card, fpgm_ctx **out);
void fpgm_ctx_destroy(fpgm_ctx *ctx);
const char *fpgm_last_error(const fpgm_ctx *ctx);
const char *fpgm_status_str(fpgm_status s);

/*
 * One CI test, one launch, host-synchronous: replaces
 * IndependenceTest::IndependenceResult(x, y, z, "g square", timer, 1)
 * (IndependenceTest.cpp:35-58 -> ComputeGSquareXY :295-364 / ComputeGSquareXYZ :65-155).
 * Exact FP64 G^2 and p-value.
 */
fpgm_status fpgm_ci_test(fpgm_ctx *ctx, int32_t x, int32_t y, const int32_t *z, int32_t d,
                         double alpha, fpgm_result *out);

/*
 * n tests in one batched launch: replaces a sequence of IndependenceResult calls and
 * the grouped ComputeGSquareXYZGroup (IndependenceTest.cpp:163-287).  flags: FPGM_F_*.
 */
fpgm_status fpgm_ci_batch(fpgm_ctx *ctx, const fpgm_test *tests, int64_t n, const int32_t *z_pool,
                          int64_t z_pool_len, double alpha, uint32_t flags, fpgm_result *out);

/* IndependenceTest::Result (IndependenceTest.h:38-49) of a grouped call. */
typedef struct {
    double p_value;          /* as the reference returns it (see fpgm_ci_group)                 */
    uint8_t is_independent;
    uint8_t pad[3];
    int32_t first;           /* index of the first independent test of the group, or -1        */
} fpgm_group_result;

/*
 * A group of CI tests sharing (x, y): replaces
 * IndependenceTest::IndependenceResult(x, y, z, "g square", timer, c_size)
 * (IndependenceTest.cpp:35-58), the call PCStable::Testing makes with a group of
 * conditioning sets (PCStable.cpp:539-569).  z holds c_size sets of d node ids each, set m
 * at z[m*d .. m*d+d).  res receives the reference's Result for the dispatch it takes:
 *   d == 0            ComputeGSquareXY (:295-364): one test, {p, independent, first = -1};
 *   c_size == 1       ComputeGSquareXYZ (:65-155): {p, independent, first = 0};
 *   c_size > 1        ComputeGSquareXYZGroup (:163-287): the first independent set m of the
 *                     group gives {-1, true, m} (:280-284), none gives {0.0, false, -1}.
 * The reference holds at most 8 results (bool results[8], :170); here c_size is unbounded.
 * df == 0 counts as independent, as in ComputeGSquareXYZ (:140-142).  out (may be NULL;
 * c_size entries, 1 when d == 0) receives every test's record, FP64-exact.  Tests of the
 * group run in one launch sequence, so X and Y are read once into L2 for all of them.
 */
fpgm_status fpgm_ci_group(fpgm_ctx *ctx, int32_t x, int32_t y, const int32_t *z, int32_t d, int32_t c_size,
                          double alpha, fpgm_group_result *res, fpgm_result *out);

/*
 * fpgm_ci_batch without any allocation: no counterpart in the reference.  For callers
 * that cannot allocate device memory (serving loops, harnesses that forbid hipMalloc).
 * Same arguments, semantics and results as fpgm_ci_batch; the library's buffers for the
 * call -- descriptors, work items, results, counters, critical values, the workspace of
 * tables split across workgroups -- live in one static device array of 17.7 MB per GPU
 * (held by every GPU the library runs on until the process ends), and nothing is
 * allocated on the host or the device.  Limits: at most 4096 tests and 32768 z-pool
 * entries per call, tables of at most 8192 cells (FPGM_ETOOBIG beyond).  The context
 * (created beforehand) holds the columns.  Calls within one process are serialised;
 * calls from several processes on one GPU race on its array.
 */
fpgm_status fpgm_ci_batch_noalloc(fpgm_ctx *ctx, const fpgm_test *tests, int64_t n, const int32_t *z_pool,
                                  int64_t z_pool_len, double alpha, uint32_t flags, fpgm_result *out);

/*
 * Raw contingency table of one test, for parity checks: the dense
 * n[z][x][y] int32 table of Counts3D / Counts2D (CellTable.cpp:23-91, :174-205,
 * :226-291, :430-455), z the mixed-radix index with the first Z most significant.
 * cap = capacity of table in int32 entries; *cells receives the table size.
 */
fpgm_status fpgm_ci_counts(fpgm_ctx *ctx, int32_t x, int32_t y, const int32_t *z, int32_t d,
                           int32_t *table, int64_t cap, int64_t *cells);

/*
 * The whole CI stage on the device: replaces step 1 of
 * PCStable::StructLearnByPCStable (PCStable.cpp:109-232: complete graph, level 0,
 * SearchAtDepth per level :265-388).  Results equal the reference's with group size 1:
 * adj_out (n_vars*n_vars uint8, symmetric) is the skeleton handed to orientation,
 * sepsets are read back with fpgm_sepset_count / fpgm_sepset_get, stats (may be NULL)
 * receives num_ci_test and per-level figures.  depth as in PCStable(net, alpha, depth):
 * levels d = 1 .. depth-1 after level 0.
 */
fpgm_status fpgm_pc_skeleton(fpgm_ctx *ctx, double alpha, int32_t depth, const fpgm_pc_opts *opts,
                             uint8_t *adj_out, fpgm_stats *stats);

/*
 * The CI stage of every instance of a multi-instance context, each an independent
 * PC-stable run with the reference's results (PCStable.cpp:109-232 per instance), run
 * level-synchronously so that one level's tests of all instances share the launches.
 * edge_freq (n_vars*n_vars int32) receives, per ordered pair, the number of instances
 * whose skeleton keeps the edge; adj_out (n_inst*n_vars*n_vars uint8, may be NULL) the
 * per-instance skeletons; num_ci_test_out (n_inst, may be NULL) per-instance test counts;
 * stats the per-level sums over instances.  Sepsets are not kept.
 */
fpgm_status fpgm_pc_skeleton_multi(fpgm_ctx *ctx, double alpha, int32_t depth, const fpgm_pc_opts *opts,
                                   int32_t *edge_freq, uint8_t *adj_out, int64_t *num_ci_test_out,
                                   fpgm_stats *stats);

/* Sizes of the sepset map of the last fpgm_pc_skeleton call (PCStable::sepset). */
fpgm_status fpgm_sepset_count(const fpgm_ctx *ctx, int64_t *n_records, int64_t *n_z);
/* Records in ascending (x,y) order; z receives the sets (ascending node ids). */
fpgm_status fpgm_sepset_get(const fpgm_ctx *ctx, fpgm_sepset *records, int64_t cap_records,
                            int32_t *z, int64_t cap_z);

/*
 * Maximum-likelihood counts of a known structure: replaces the counting loops of
 * ParameterLearning::LearnParamsKnowStructCompData (ParameterLearning.cpp:24-69).  Node i
 * (= variable i, i < n_nodes) has the parents parents[par_off[i] .. par_off[i+1]) in
 * ascending id order (the reference keeps them in a std::set).  Its table N(AB) holds
 * card(i) * prod card(parents) int32 counts over every row, indexed like the reference's
 * PotentialTableBase (PotentialTableBase.cpp:102-130): value(i) most significant, then
 * the parents in the given order; N(B) (ParameterLearning.cpp:34,64-66) is its sum over
 * value(i).  count_off (n_nodes + 1 entries) receives the table offsets in counts; with
 * counts == NULL only count_off is filled (size query).  cap: capacity of counts.
 * The reference's virtual ROOT parent (random 0/1 per row) has no device counterpart.
 */
fpgm_status fpgm_mle_counts(fpgm_ctx *ctx, int32_t n_nodes, const int32_t *par_off, const int32_t *parents,
                            int64_t *count_off, int32_t *counts, int64_t cap);

/*
 * Autotuning of the launch shapes and the round schedule (no counterpart in the reference,
 * whose OpenMP schedule is fixed: PCStable.cpp:290-312).  Results never depend on these
 * settings; only the time does.  fpgm_autotune runs the CI stage once with the current
 * settings, keeping the state every level starts from, then replays each level alone under
 * candidate settings (coordinate descent, each candidate timed on the device) and keeps the
 * fastest per level in the context: later fpgm_pc_skeleton calls on the context use them.
 * out (cap entries) receives one record per level; fpgm_tune_set applies saved records.
 */
typedef struct {
    int32_t level;          /* 0: the MFMA Gram path; d >= 1: level d                       */
    int32_t pipeline;       /* 1: round r+1 built while round r runs (-1: default)            */
    int32_t concurrent;     /* 1: slice and device-memory sinks overlap the others on streams
                               of their own, 0: one stream (-1: default)                       */
    int32_t pad;
    int64_t items_per_cu;   /* work items a launch aims at per CU (level 0: Gram items)       */
    int64_t wg_per_cu[5];   /* grid workgroups per CU: lane, lane-wide, shared, slice, device */
    int64_t max_slices;     /* z slices of a table before the device-memory sink              */
    int64_t round_cap;      /* cap of the per-edge round ramp                                  */
    double default_ms;      /* the level's time with the settings before tuning                */
    double tuned_ms;        /* with the chosen settings                                        */
} fpgm_level_tune;

fpgm_status fpgm_autotune(fpgm_ctx *ctx, double alpha, int32_t depth, const fpgm_pc_opts *opts,
                          fpgm_level_tune *out, int32_t cap, int32_t *n_levels, double *seconds);
fpgm_status fpgm_tune_set(fpgm_ctx *ctx, const fpgm_level_tune *levels, int32_t n);
fpgm_status fpgm_tune_clear(fpgm_ctx *ctx);

/* chi-square helpers of the library's own FP64 CDF (the reference's stats::pchisq,
 * called at IndependenceTest.cpp:146,268,355), exposed for checks. */
double fpgm_pchisq(double x, int32_t df);
double fpgm_chisq_crit(int32_t df, double alpha);

/* Timing of the last batched launch sequence, for the roofline figure: the summed
 * duration (HIP events on the context stream) and algorithmic bytes of the counting
 * kernel launches of the last fpgm_pc_skeleton / fpgm_ci_batch call. */
fpgm_status fpgm_last_kernel_stats(const fpgm_ctx *ctx, double *kernel_ms, double *alg_bytes,
                                   int64_t *launches);

/* The same launches' bytes in the layout the counting kernels read: per test, N rows of
 * its packed columns (2, 4 or 8 bits per value by cardinality) instead of (|Z|+2) * N. */
fpgm_status fpgm_last_packed_bytes(const fpgm_ctx *ctx, double *packed_bytes);

/* Level-0 Gram path of the last fpgm_pc_skeleton call: MFMA kernel time, int8 operations
 * (2 * rows * 128 * 128 per upper-triangle tile), per-pair statistic kernel time. */
fpgm_status fpgm_last_gram_stats(const fpgm_ctx *ctx, double *gram_ms, double *int8_ops, double *pairs_ms,
                                 int64_t *launches);

/* Per counting sink of the last fpgm_pc_skeleton / fpgm_ci_batch call (no counterpart in the
 * reference; measurement only).  sink: 0 per-lane 16-bit LDS counters (tables <= 64 cells),
 * 1 the same with 48 or 64 KiB of counters (<= 96 / 128 cells), 2 shared LDS histogram
 * (<= 8192 cells), 3 z-sliced LDS histogram, 4 device-memory table.  ms: summed HIP-event
 * durations of its launches; alg_bytes: (|Z| + 2) * rows per test; packed_bytes: the same
 * tests' packed column bytes; row_tests: rows summed over its tests (each row of each test
 * is one histogram increment).  tests, items, bytes and row_tests count the tests the
 * device counted: those of a speculative round whose edge an earlier round decided are
 * dropped before counting and left out. */
typedef struct {
    int32_t sink;
    int32_t pad;
    int64_t launches;
    int64_t tests;
    int64_t items;
    double ms;
    double alg_bytes;
    double packed_bytes;
    double row_tests;
} fpgm_sink_stats;

fpgm_status fpgm_last_sink_stats(const fpgm_ctx *ctx, fpgm_sink_stats *out, int32_t cap, int32_t *n);

/* Grouped level-1 counting of the last fpgm_pc_skeleton call (tests sharing the walking
 * endpoint and the conditioning node, counted as one-hot int8 products on the MFMA units):
 * kernel time, int8 operations of the padded 32 x 32 x 32 fragments, tests counted there,
 * launch sequences that had such tests.  No counterpart in the reference. */
fpgm_status fpgm_last_group_stats(const fpgm_ctx *ctx, double *ms, double *int8_ops, int64_t *tests,
                                  int64_t *sequences);

#ifdef __cplusplus
}
#endif
#endif
