// C ABI and host-side level driver of the MI355X PC-stable CI stage.
//
// The driver keeps the reference's semantics (PCStable.cpp:84-626, group size 1):
//   * level 0 tests every pair of the complete graph in lexicographic edge order
//     (Network.cpp:270-283) in one batched launch sequence;
//   * level d >= 1 freezes the adjacency (PCStable.cpp:272); each surviving edge x<y
//     walks its conditioning sets in the reference's order -- every d-subset of
//     adj(x)\{y} in lexicographic order (ChoiceGenerator.cpp:55-85), then those of
//     adj(y)\{x} (CheckEdge, PCStable.cpp:412-460) -- and stops at the first
//     independent one, whose set becomes the sepset (PCStable.cpp:568-583);
//   * removals are applied after the level (PCStable.cpp:369-386) and the search
//     continues while max degree - 1 > d (FreeDegree, PCStable.cpp:620-626).
// The device evaluates each level in speculative rounds: every unfinished edge
// contributes its next W tests, all tests of a round run in one launch sequence,
// and the host keeps, per edge, the first independent test in enumeration order.
// A test's outcome does not depend on which others run, so the sepsets and the
// skeleton equal the sequential ones; tests after an edge's first independent one
// are counted as speculative work, never as reference-equivalent tests.
#include "ci_kernels.h"
#include "chisq.h"
#include "fpgm_ci.h"
#include "gram_kernels.h"
#include "round_kernels.h"
#include "mle_kernels.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include <omp.h>
#include <string>
#include <vector>

using namespace fpgm;

namespace {

constexpr int64_t kMaxCells = (int64_t)1 << 31;   // dense-table limit (the reference indexes with int)
constexpr int64_t kWsInit = (int64_t)1 << 24;     // workspace, int32 entries
constexpr int64_t kWsMax = (int64_t)1 << 30;      // 4 GiB of workspace at most per launch
constexpr int kMaxD = kMaxZLds;
constexpr int64_t kParallelTests = 16384;          // rounds this large are built on several host threads

template <typename T>
struct DevBuf {
    T *p = nullptr;
    int64_t cap = 0;
    hipError_t ensure(int64_t n, bool zero, hipStream_t s)
    {
        if (n <= cap) return hipSuccess;
        if (p) {
            (void)hipStreamSynchronize(s);
            (void)hipFree(p);
            p = nullptr;
            cap = 0;
        }
        int64_t want = std::max<int64_t>(n, 1024);
        want = want + want / 2;
        hipError_t e = hipMalloc((void **)&p, (size_t)want * sizeof(T));
        if (e != hipSuccess) {
            want = n;
            e = hipMalloc((void **)&p, (size_t)want * sizeof(T));
            if (e != hipSuccess) {
                p = nullptr;
                return e;
            }
        }
        cap = want;
        if (zero) {
            e = hipMemsetAsync(p, 0, (size_t)cap * sizeof(T), s);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Pinned host staging buffer (asynchronous copies to and from the device).
template <typename T>
struct HostBuf {
    T *p = nullptr;
    int64_t cap = 0;
    hipError_t ensure(int64_t n, hipStream_t s)
    {
        if (n <= cap) return hipSuccess;
        if (p) {
            (void)hipStreamSynchronize(s);  // no copy may still read or write the old buffer
            (void)hipHostFree(p);
            p = nullptr;
            cap = 0;
        }
        int64_t want = std::max<int64_t>(n, 1024);
        want = want + want / 2;
        hipError_t e = hipHostMalloc((void **)&p, (size_t)want * sizeof(T), hipHostMallocDefault);
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        cap = want;
        return hipSuccess;
    }
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Counting launches are grouped by sink kind, in this order.
constexpr int kGroups = 6;
constexpr int kGroupSink[kGroups] = {kSinkLane, kSinkLaneW, kSinkShared, kSinkSlice, kSinkGlobal, kSinkBits};
constexpr int kGroupBits = 5;   // the bit-plane launch (grid shape of the per-lane launch)
constexpr int kTuneGroups = 5;  // groups with a grid shape of their own (fpgm_level_tune::wg_per_cu)
constexpr int64_t kPackMinRows = 262144;  // mean rows per instance from which columns are packed
// instances of at least this many rows count 65..128-cell tables in the shared sink's column
// layout instead of the per-lane 16-bit counters (1000 variables x 10^6 rows: stage 175.7 ->
// 168.4 ms, the two sinks' time 84.1 -> 77.3 ms; 300 variables x 5 * 10^5 rows: unchanged)
constexpr int64_t kWideMaxRows = 262144;
constexpr int kGroupPending = kGroups;     // a 97..128-cell table whose group is decided per launch
static_assert(kGroupPending == kRoutePending && kGroupBits == kRouteBits, "routing groups");
constexpr int kGroupMfma = kGroups + 1;    // a level-1 test counted by the grouped MFMA kernel

// Buffers of one in-flight batch.  Two slots let the host build one batch while the
// device runs the other; the workspace and completion counters are shared, since both
// slots launch on the context's one stream and every launch leaves them zero.
struct Slot {
    DevBuf<DevTest> d_tests;
    DevBuf<int32_t> d_zpool;
    DevBuf<uint64_t> d_items, d_items_live;  // work items as built, and after compaction
    DevBuf<DevResult> d_res;
    DevBuf<uint32_t> d_queue;  // work-queue counters (kGroups x kXcds) and live-item counts (kGroups)
    DevBuf<WideItem> d_wide_items;
    DevBuf<WidePart> d_wide_parts;
    DevBuf<WideTest> d_wide_tests;
    DevBuf<GroupChunk> d_gchunks;  // grouped level-1 chunks, their tests and work items
    DevBuf<int32_t> d_gtest;
    DevBuf<uint64_t> d_gitems;
    DevBuf<uint64_t> d_pitems;  // partitioning work items (test << 32 | row chunk)
    HostBuf<uint64_t> h_pitems;
    HostBuf<GroupChunk> h_gchunks;
    HostBuf<int32_t> h_gtest;
    HostBuf<uint64_t> h_gitems;
    // device-generated rounds: the round's segments, each segment's outcome (seg_out), the
    // test -> segment map and the round's figures
    DevBuf<RoundSeg> d_segs;
    HostBuf<RoundSeg> h_segs;
    DevBuf<int32_t> d_seg_out, d_t2s;
    HostBuf<int32_t> h_seg_out;
    DevBuf<RoundStats> d_rstats;
    HostBuf<RoundStats> h_rstats;
    bool dev_round = false;  // the pending sequence is a device-generated round
    int64_t n_segs = 0;
    int dev_d = 0;
    std::vector<int32_t> g_bucket;  // grouping scratch: bucket starts, test order
    std::vector<int32_t> g_order;
    HostBuf<DevTest> h_tests;
    HostBuf<int32_t> h_zpool;
    HostBuf<uint64_t> h_items;  // test << 32 | (slice * n_chunks + chunk)
    HostBuf<DevResult> h_res;
    HostBuf<WideItem> h_wide_items;
    HostBuf<WideTest> h_wide_tests;
    std::vector<int32_t> mode_of;
    std::vector<int64_t> need;  // workspace int32 entries of each test of the batch being built
    std::vector<int64_t> pneed; // partition-buffer int32 entries of each (kTFPart) test
    std::vector<uint8_t> fit;   // table and marginals fit the workgroup's LDS scratch
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_done = nullptr;
    hipEvent_t ev_grp[kGroups + 1] = {};  // per-sink timing when tracing: start of each group
    hipEvent_t ev_gend[kGroups] = {};     // and its end (groups may run concurrently)
    hipEvent_t ev_fork = nullptr, ev_join[kGroups] = {};  // concurrent sink launches
    hipEvent_t ev_g0 = nullptr, ev_g1 = nullptr;           // grouped MFMA launches
    hipGraphExec_t gexec = nullptr;  // the captured sequence (FPGM_GRAPH=1), updated in place when it can be
    bool graphed = false;            // the pending sequence ran as a graph (no per-launch events)
    int64_t g_tests = 0, g_chunks = 0, g_items = 0;        // ... of the pending sequence
    double g_ops = 0.0;
    bool pending = false;
    int64_t n = 0;       // tests of the pending launch sequence
    double bytes = 0.0;  // their algorithmic bytes
    double pbytes = 0.0; // and the bytes of their packed columns
    struct GroupInfo {
        int64_t tests = 0, items = 0, dsum = 0;
        double csum = 0.0, cmax = 0.0, bytes = 0.0, pbytes = 0.0, rows = 0.0;
    } grp[kGroups];

    hipError_t init()
    {
        hipError_t e = hipSuccess;
        for (hipEvent_t *ev : {&ev0, &ev1, &ev_done})
            if ((e = hipEventCreate(ev)) != hipSuccess) return e;
        for (hipEvent_t &ev : ev_grp)
            if ((e = hipEventCreate(&ev)) != hipSuccess) return e;
        for (hipEvent_t &ev : ev_gend)
            if ((e = hipEventCreate(&ev)) != hipSuccess) return e;
        if ((e = hipEventCreate(&ev_g0)) != hipSuccess) return e;
        if ((e = hipEventCreate(&ev_g1)) != hipSuccess) return e;
        if ((e = hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming)) != hipSuccess) return e;
        for (hipEvent_t &ev : ev_join)
            if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) return e;
        return e;
    }
    void release()
    {
        d_tests.release();
        d_zpool.release();
        d_items.release();
        d_items_live.release();
        d_res.release();
        d_queue.release();
        d_wide_items.release();
        d_wide_parts.release();
        d_wide_tests.release();
        d_gchunks.release();
        d_gtest.release();
        d_gitems.release();
        d_pitems.release();
        h_pitems.release();
        h_gchunks.release();
        h_gtest.release();
        h_gitems.release();
        h_tests.release();
        h_items.release();
        h_zpool.release();
        h_res.release();
        h_wide_items.release();
        h_wide_tests.release();
        d_segs.release();
        h_segs.release();
        d_seg_out.release();
        d_t2s.release();
        h_seg_out.release();
        d_rstats.release();
        h_rstats.release();
        for (hipEvent_t ev : {ev0, ev1, ev_done})
            if (ev) (void)hipEventDestroy(ev);
        for (hipEvent_t ev : ev_grp)
            if (ev) (void)hipEventDestroy(ev);
        for (hipEvent_t ev : ev_gend)
            if (ev) (void)hipEventDestroy(ev);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        gexec = nullptr;
        if (ev_g0) (void)hipEventDestroy(ev_g0);
        if (ev_g1) (void)hipEventDestroy(ev_g1);
        for (hipEvent_t ev : ev_join)
            if (ev) (void)hipEventDestroy(ev);
    }
};

struct HostTest {
    int32_t x, y, d;
    int32_t inst;   // dataset instance
    int64_t z_off;  // into the host z pool
    int32_t edge = -1;  // index of the test's edge in the level (device skip checks), -1 for none
    int32_t side = 0;   // walk side that produced it: 0 subsets of adj(x), 1 of adj(y)
};

struct InstHost {
    const uint8_t *cols;  // device address
    int64_t stride, n_rows;
};

}  // namespace

struct fpgm_ctx {
    int dev = 0;
    hipStream_t stream = nullptr;
    hipStream_t side[6] = {};  // streams of a round's slice, device-memory and small LDS-sink launches
    const uint8_t *d_cols = nullptr;  // all instances' columns (owned unless borrowed)
    bool owns_cols = false;
    std::vector<InstHost> inst;       // dataset instances (1, or the bootstrap ensemble)
    DevInst *d_inst = nullptr;
    int32_t n_vars = 0;
    std::vector<int32_t> card;
    int32_t *d_card = nullptr;
    // packed columns read by the counting kernels: bits per value by cardinality (2 up to
    // 4 states, 4 up to 16, else 8), when the instances are large (kPackMinRows rows on
    // average); FPGM_PACK=1 packs always, FPGM_PACK=0 never
    int pack = -1;
    std::vector<uint8_t> colw;
    uint8_t *d_colw = nullptr;
    uint8_t *d_pcols = nullptr;  // owned; nullptr when every column has 8 bits (the uint8 columns are read)
    // bit planes of every instance (DevInst::planes), built with the packed copy for
    // variables of at most kBitsMaxCard states: |Z| = 1 tables of at most bits_cells cells
    // count from them.  FPGM_BITS=0 builds none; FPGM_BITS_CELLS sets the bound
    int bits = 1;
    int32_t bits_cells = 48;
    uint32_t *d_planes = nullptr;
    int32_t *d_poff = nullptr;
    int64_t planes_bytes = 0;
    std::string err;
    std::mutex mu;
    int n_cu = 256;
    int host_threads = 1;  // host threads for building large rounds
    struct {
        double base = 2.0;    // round ramp factor
        int64_t cap = 8;      // ramp cap (tests per edge per round); larger rounds come from the budget
        int64_t promote = 0;  // dependent tests after which an edge runs the rest of its walk
        double round_bytes = 2147483648.0;  // algorithmic bytes a round aims at (tests = bytes / ((d+2) N))
        int64_t items_per_cu = 16;          // work items a launch aims at per CU (row chunks per test)
        int64_t wpc[5] = {4, 3, 4, 2, 8};   // grid workgroups per CU: lane, lane-wide, shared, slice, device
        int pipeline = 1;                    // build round r+1 while the device counts round r
        int64_t gram_items_per_cu = 8;       // level-0 Gram work items (tile pairs x row splits) per CU
        int concurrent = 1;                  // slice and device-memory sinks on streams of their own
        int lds_own_chunks = 0;              // FPGM_LDS_OWN_CHUNKS: the shared sink's group chunked by its own test count
        int64_t side_cu_mult = 16;           // FPGM_SIDE_CU_MULT: LDS groups below this many items per CU on a side stream
        int64_t slice_min_rows = 0;          // FPGM_SLICE_MIN_ROWS: rows per z-slice work item at least this many
    } sched;
    // per-level scheduling and launch shapes chosen by fpgm_autotune (index = level); a
    // candidate being timed by the tuner overrides them
    std::vector<fpgm_level_tune> tune;
    const fpgm_level_tune *tune_try = nullptr;
    int64_t max_slices = kMaxSlices;  // z slices before a table goes to the device-memory sink
    int64_t wide_max_rows = kWideMaxRows;  // FPGM_WIDE_MAX_ROWS
    uint32_t sink_mask = 0;  // FPGM_SINKS=shared: bit 0 no wide per-lane sink, bit 1 no z slices; =nowide: bit 0; =nolane: bits 0, 2 (no per-lane sink); =nocol: bit 3 (no column layout in the per-lane launch)
    // grouped level-1 tests on the MFMA units (k_group_count) for instances of at least
    // group_min_rows rows: off unless FPGM_GROUP=1.  Measured slower than the LDS sinks on
    // level 1 at 1000 variables x 10^6 rows (1856 against 122 ms: in the best-grouped
    // round 22 tests per 128-column chunk, 0.71 POP/s on fragments whose A rows are half
    // padding, against 0.18 us per test on the lane sinks; later rounds' groups fragment as
    // edges reach their y side at different times) and on the bootstrap ensemble (2364
    // against 494 ms); FPGM_GROUP_MIN_ROWS sets the bound
    int group = 0;
    int64_t group_min_rows = 0;
    // FPGM_GRAPH=1: each launch sequence's device work (uploads, compaction, counting
    // launches, p-values, read-back) captured into a HIP graph and launched as one; off by
    // default (measured no faster)
    int graph = 0;
    int64_t graph_instantiations = 0, graph_updates = 0;
    // rounds generated on the device (the host plans per edge, the device expands the
    // tests, routes them and reduces each edge's outcome) at levels whose every table fits
    // LDS: FPGM_DEVROUNDS=1 always, 0 never, unset for instances of at most
    // devround_max_rows rows on average at levels of many tests (where the host's per-test
    // work sets the pace).  Measured: the 512-replicate bootstrap ensemble 4.70 -> 4.20 s
    // per stage; 37 variables x 5 000 rows 0.90 -> 0.93 ms, x 200 000 rows 1.50 -> 1.74 ms,
    // 300 variables x 500 000 rows 8.83 -> 9.64 ms (few tests per round: the host's rounds
    // cost less than the extra launches)
    // tables beyond LDS counted from rows partitioned by z slice first (k_part_keys, then
    // the z-slice launch over each slice's rows only), for instances of at least
    // part_min_rows rows: FPGM_PART=1 (default) the tables of more z slices than max_slices,
    // which would otherwise go to device memory (1000 variables x 10^6 rows: stage
    // 151.9 -> 149.7 ms, two runs each; the device-memory sink's 10.7 ms of atomics per
    // stage gone), 2 every table beyond LDS (166.7 ms: for tables of a few slices the
    // partitioning pass costs more than reading the rows once per slice), 0 none
    int part = 1;
    int64_t part_min_rows = 0;
    DevBuf<int32_t> d_part;  // partition buffers of a launch sequence's partitioned tests
    int devrounds = -1;
    int64_t devround_max_rows = 65536;
    int64_t devround_min_tests = (int64_t)1 << 20;  // ... and levels whose walks hold at least this many tests
    DevBuf<int32_t> d_nbr;    // the level's neighbour store (device rounds)
    DevBuf<int64_t> d_binom;  // C(n, k) for the level's subset sizes
    // host columns uploaded through pinned staging, chunk by chunk, overlapped with their
    // packing (FPGM_UPLOAD_PIPELINE=0: one pageable copy per instance, then the packing)
    int upload_pipeline = 1;
    fpgm_sink_stats sink[kGroups] = {};  // per counting sink of the last call
    double grp_ms = 0.0, grp_ops = 0.0;  // grouped launches of the last call: time, int8 operations
    int64_t grp_tests = 0, grp_launches = 0;

    double crit_alpha = -1.0;
    double z_alpha = 0.0;
    std::vector<double> crit_host;
    double *d_crit = nullptr;

    Slot slot[2];
    DevBuf<int32_t> d_ws;
    // speculative rounds of a level: round number of the launch being submitted (0 = no
    // skip checks) and, per edge of the level, the first round that found it independent
    int32_t skip_round = 0;
    DevBuf<int32_t> d_edge_done;
    DevBuf<uint32_t> d_done, d_flag;

    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int trace = 0;               // > 0: one line per counting launch on stderr
    double t_prep = 0.0, t_wait = 0.0;  // host seconds building launches / waiting for them (trace)
    double k_ms = 0.0, k_bytes = 0.0, k_pbytes = 0.0;
    int64_t k_launches = 0;

    // level-0 Gram path
    DevBuf<int32_t> d_C, d_rowvar, d_tiles, d_off;
    DevBuf<uint8_t> d_rowstate, d_l0_flags;
    DevBuf<PairSummary> d_l0_sum;
    HostBuf<uint8_t> h_l0_flags;
    HostBuf<PairSummary> h_l0_sum;
    double g_ms = 0.0, g_ops = 0.0, g_pair_ms = 0.0;
    int64_t g_launches = 0;

    // maximum-likelihood counts
    DevBuf<MleNode> d_mle_nodes;
    DevBuf<int32_t> d_mle_rel, d_mle_items, d_mle_out;

    std::vector<fpgm_sepset> seps;
    std::vector<int32_t> sep_z;
};

namespace {

fpgm_status fail(fpgm_ctx *c, fpgm_status s, const std::string &msg)
{
    if (c) c->err = msg;
    return s;
}

fpgm_status hip_fail(fpgm_ctx *c, hipError_t e, const char *what)
{
    const fpgm_status s = (e == hipErrorOutOfMemory) ? FPGM_ENOMEM : FPGM_EDEVICE;
    return fail(c, s, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr, what)                         \
    do {                                            \
        hipError_t e_ = (expr);                     \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, what); \
    } while (0)

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

fpgm_status ensure_crit(fpgm_ctx *ctx, double alpha)
{
    if (ctx->crit_alpha == alpha && ctx->d_crit) return FPGM_OK;
    ctx->crit_host.assign(kCritTab, 0.0);
    for (int df = 1; df <= kCritTab; ++df) ctx->crit_host[df - 1] = chisq_crit(df, alpha);
    if (!ctx->d_crit) HIP_TRY(hipMalloc((void **)&ctx->d_crit, sizeof(double) * kCritTab), "hipMalloc crit");
    HIP_TRY(hipMemcpyAsync(ctx->d_crit, ctx->crit_host.data(), sizeof(double) * kCritTab, hipMemcpyHostToDevice,
                           ctx->stream),
            "upload crit");
    ctx->z_alpha = (alpha > 0.0 && alpha < 1.0) ? norm_upper_quantile(alpha) : 0.0;
    ctx->crit_alpha = alpha;
    return FPGM_OK;
}

fpgm_status batch_wait(fpgm_ctx *ctx, Slot &S);

// Builds and enqueues one launch sequence for tests [i0, i1) of a call, as many as fit
// the workspace limit (*i1_out receives i1); nothing waits here.  The slot's pinned
// buffers stage the descriptors, and its results arrive in slot.h_res once
// Bytes per row of a test's packed columns (X, Y and Z at their widths).
double packed_row_bytes(const fpgm_ctx *ctx, const HostTest &h, const int32_t *zpool)
{
    int bits = ctx->colw[h.x] + ctx->colw[h.y];
    for (int j = 0; j < h.d; ++j) bits += ctx->colw[zpool[h.z_off + j]];
    return bits / 8.0;
}

// Sink group of a table of at most kLdsCells cells (marg: its marginal entries), the
// launch it joins: 0 per-lane (<= 64 cells), 1 lane-wide (<= 96), kGroupPending (97..128:
// lane-wide or shared, decided per launch), 2 shared (<= kLdsCells); -1 for larger tables.
// Sets the table's kTF* bits; *fits: table and marginals fit the workgroup's LDS scratch.
// grid workgroups per CU of a group's launch (the bit-plane launch takes the per-lane one's)
int64_t wpc_of(const fpgm_ctx *ctx, int g)
{
    return ctx->sched.wpc[g == kGroupBits ? 0 : g];
}

RouteParams route_params(const fpgm_ctx *ctx)
{
    RouteParams p{};
    p.sink_mask = ctx->sink_mask;
    p.packed = ctx->d_pcols != nullptr ? 1 : 0;
    p.wide_max_rows = ctx->wide_max_rows;
    p.bits_cells = ctx->d_planes ? ctx->bits_cells : 0;
    return p;
}

int lds_group(const fpgm_ctx *ctx, DevTest &T, int64_t cells, int64_t marg, int64_t rows, bool *fits)
{
    static_assert(kColCells == kLdsCells / 32, "column-layout cells");
    return route_lds(route_params(ctx), T.flags, cells, marg, rows, T.cx, T.cy, T.d, fits);
}

// batch_wait returns.  flags: FPGM_F_*, tflags: extra kTF* bits for every test.
// keep_ws receives the workspace offset of test i0 when kTFKeep is set.
fpgm_status batch_submit(fpgm_ctx *ctx, Slot &S, const HostTest *tests, int64_t n, const int32_t *zpool,
                         int64_t zlen, double alpha, uint32_t flags, float band, int32_t tflags,
                         int64_t rows_per_chunk_forced, int64_t i0, int64_t *i1_out, int64_t *keep_ws)
{
    fpgm_status st = batch_wait(ctx, S);  // a batch left pending by an earlier error
    if (st != FPGM_OK) return st;
    const double tp0 = now_s();
    st = ensure_crit(ctx, alpha);
    if (st != FPGM_OK) return st;
    int64_t N = 0;  // longest instance in the batch
    for (int64_t i = i0; i < n; ++i) N = std::max(N, ctx->inst[tests[i].inst].n_rows);
    const int64_t steps = std::max<int64_t>(1, (N + kRowsPerStep - 1) / kRowsPerStep);

    // rows per chunk, one value per sink group (each group is a launch of its own): the
    // LDS sinks share one, enough work items over all the launch's tests to fill every CU
    // (cutting a small group of them into more chunks measured slower: every chunk
    // clears, reduces and merges its LDS table); the z-slice and device-memory sinks,
    // which run beside the others on streams of their own, get enough items each
    const int64_t target_items = (int64_t)ctx->n_cu * ctx->sched.items_per_cu;
    int64_t rpc[kGroups];

    // descriptors go straight into the slot's pinned buffer.  Per test, in parallel: its
    // descriptor, sink group (gidx), work items and workspace need; then, in order, the
    // workspace offsets (a launch sequence ends where the workspace limit is reached) and
    // each test's first work item within its group.
    const int64_t m = n - i0;
    HIP_TRY(S.h_tests.ensure(m, ctx->stream), "alloc pinned tests");
    DevTest *dt = S.h_tests.p;
    std::vector<int32_t> &gidx = S.mode_of;
    std::vector<int64_t> &need = S.need;
    std::vector<int64_t> &pneed = S.pneed;
    std::vector<uint8_t> &fit = S.fit;
    gidx.resize((size_t)m);
    need.resize((size_t)m);
    pneed.resize((size_t)m);
    fit.resize((size_t)m);
    int too_big = 0;
    // grouped level-1 counting: instances long enough, no table kept for a parity read, and
    // the walk sides known (speculative rounds of the level driver)
    const bool grp_on = ctx->group && ctx->skip_round > 0 && !(tflags & kTFKeep) && N >= ctx->group_min_rows;
    int64_t units[kGroups] = {0, 0, 0, 0, 0};  // per group: the tests' z slices (work items per chunk)
    int64_t units_pending = 0;
    const int nth = (m >= kParallelTests) ? ctx->host_threads : 1;
#pragma omp parallel for if (nth > 1) num_threads(nth) schedule(static) reduction(| : too_big) \
    reduction(+ : units[:kGroups]) reduction(+ : units_pending)
    for (int64_t k = 0; k < m; ++k) {
        const HostTest &h = tests[i0 + k];
        DevTest T{};
        T.x = h.x;
        T.y = h.y;
        T.d = h.d;
        T.inst = h.inst;
        T.edge = h.edge;
        T.z_off = (int32_t)h.z_off;
        T.cx = ctx->card[h.x];
        T.cy = ctx->card[h.y];
        int64_t dimz = 1;
        bool narrow = ctx->colw[h.x] != 8 && ctx->colw[h.y] != 8;
        for (int j = 0; j < h.d; ++j) {
            dimz *= ctx->card[zpool[h.z_off + j]];
            narrow = narrow && ctx->colw[zpool[h.z_off + j]] != 8;
            if (dimz > kMaxCells) break;
        }
        const int64_t cells = dimz * T.cx * T.cy;
        if (dimz > kMaxCells || cells >= kMaxCells) {
            too_big = 1;
            continue;
        }
        T.dimz = (int32_t)dimz;
        T.n_slices = 1;
        T.slice_cells = (int32_t)cells;
        const int64_t marg = dimz * (T.cx + T.cy + 1);
        T.flags = tflags | (narrow ? kTFNarrow : 0);
        bool fits;
        int gi = lds_group(ctx, T, cells, marg, ctx->inst[h.inst].n_rows, &fits);
        if (grp_on && h.d == 1) {
            // a level-1 test whose (u, z) rows and other columns fit the grouped kernel
            const int32_t u = h.side ? h.y : h.x, o = h.side ? h.x : h.y;
            const int64_t cu = ctx->card[u], cz = ctx->card[zpool[h.z_off]];
            if (cu * cz <= kGrpRows && ctx->card[o] <= kGrpMaxOther) gi = kGroupMfma;
        }
        if (gi < 0) {
            gi = 4;  // kSinkGlobal
            fits = false;
            // large device-memory tables get their statistic from the grid-wide
            // finalize kernels (the counted table of a parity read stays in place)
            if (!(tflags & kTFKeep)) T.flags |= kTFWide;
            const int64_t cxy = (int64_t)T.cx * T.cy;
            // partitioned by z slice: the compile-time row loops (|Z| <= 8, every column of a
            // packed context packed narrow), no table kept for a parity read
            const bool part_ok = ctx->part && !(tflags & kTFKeep) && h.d <= 8 &&
                                 (ctx->d_pcols == nullptr || narrow) &&
                                 ctx->inst[h.inst].n_rows >= std::max<int64_t>(ctx->part_min_rows, 1);
            if (part_ok && cxy <= kSliceCells) {
                const int64_t sc = (kSliceCells / cxy) * cxy;
                const int64_t ns = (cells + sc - 1) / sc;
                // mode 1: the tables beyond the slice limit (else device memory); 2: all
                if (ns <= kPartMaxSlices && (ctx->part == 2 || ns > ctx->max_slices || (ctx->sink_mask & 2))) {
                    gi = 3;  // the z-slice launch, over partitioned rows
                    T.n_slices = (int32_t)ns;
                    T.slice_cells = (int32_t)sc;
                    T.flags = (T.flags & ~kTFWide) | kTFPart;
                }
            }
            if (!(T.flags & kTFPart) && cxy <= kSliceCells && !(ctx->sink_mask & 2)) {
                const int64_t sc = (kSliceCells / cxy) * cxy;
                const int64_t ns = (cells + sc - 1) / sc;
                if (ns <= ctx->max_slices) {
                    gi = 3;  // kSinkSlice
                    T.n_slices = (int32_t)ns;
                    T.slice_cells = (int32_t)sc;
                    T.flags &= ~kTFWide;
                }
            }
        }
        T.ws_off = -1;
        fit[k] = fits ? 1 : 0;
        gidx[k] = gi;
        if (gi == kGroupPending) ++units_pending;
        else if (gi != kGroupMfma) units[gi] += T.n_slices;
        dt[k] = T;
    }
    if (too_big) return fail(ctx, FPGM_ETOOBIG, "contingency table exceeds 2^31 cells");
    // Tables of 97..128 cells: with shared-sink tests in the launch anyway they join that
    // launch, and the lane-wide launch keeps its 48 KiB instance (3 workgroups per CU);
    // otherwise they stay per lane in the 64 KiB instance rather than add a launch.
    // Measured: 1000 variables, level 1 127 -> 123 ms with the first; 300 variables (tables
    // of at most 125 cells), level 1 6.1 ms with the second against 7.2 ms with a launch more.
    bool lane_w128 = false;
    if (units_pending > 0) {
        const bool to_shared = units[2] > 0;
        lane_w128 = !to_shared;
#pragma omp parallel for if (nth > 1) num_threads(nth) schedule(static)
        for (int64_t k = 0; k < m; ++k) {
            if (gidx[k] != kGroupPending) continue;
            DevTest &T = dt[k];
            if (to_shared) {
                const int64_t cells = (int64_t)T.dimz * T.cx * T.cy;
                const int64_t marg = (int64_t)T.dimz * (T.cx + T.cy + 1);
                T.flags |= kTFLdsTable;
                fit[k] = (cells + marg <= kLdsCells) ? 1 : 0;
                gidx[k] = 2;
            } else {
                gidx[k] = 1;
            }
        }
        units[lane_w128 ? 1 : 2] += units_pending;
    }
    for (int g = 0; g < kGroups; ++g) {
        int64_t r;
        if (rows_per_chunk_forced > 0) {
            r = (rows_per_chunk_forced + kRowsPerStep - 1) / kRowsPerStep * kRowsPerStep;
        } else {
            const int64_t u = (kGroupSink[g] == kSinkSlice || kGroupSink[g] == kSinkGlobal ||
                               (ctx->sched.lds_own_chunks && kGroupSink[g] == kSinkShared))
                                  ? std::max<int64_t>(units[g], 1)
                                  : m;
            const int64_t cpt = std::min(steps, std::max<int64_t>(1, (target_items + u - 1) / u));
            r = (steps + cpt - 1) / cpt * kRowsPerStep;
            if (kGroupSink[g] == kSinkSlice && ctx->sched.slice_min_rows > 0)
                r = std::max(r, (ctx->sched.slice_min_rows + kRowsPerStep - 1) / kRowsPerStep * kRowsPerStep);
        }
        rpc[g] = std::min<int64_t>(r, kLaneMaxRowsPerChunk);
    }
    // chunks and workspace of every test, by its group's chunk size
#pragma omp parallel for if (nth > 1) num_threads(nth) schedule(static)
    for (int64_t k = 0; k < m; ++k) {
        DevTest &T = dt[k];
        pneed[k] = 0;
        if (gidx[k] == kGroupMfma) {  // counted by the grouped kernel: no work items of its own
            T.n_chunks = 0;
            need[k] = 0;
            continue;
        }
        if (T.flags & kTFPart) {  // one work item per slice over the partitioned rows; no workspace
            T.n_chunks = 1;
            need[k] = 0;
            pneed[k] = part_layout(T.dimz, T.cx, T.cy, T.n_slices, ctx->inst[T.inst].n_rows).words;
            continue;
        }
        const int64_t r = rpc[gidx[k]];
        const int64_t n_chunks = std::max<int64_t>(1, (ctx->inst[T.inst].n_rows + r - 1) / r);
        T.n_chunks = (int32_t)n_chunks;
        const int64_t cells = (int64_t)T.dimz * T.cx * T.cy;
        const int64_t marg = (int64_t)T.dimz * (T.cx + T.cy + 1);
        const bool need_ws = (tflags & kTFKeep) || n_chunks > 1 || !fit[k];
        if (!need_ws) T.flags |= kTFLdsMarg;
        need[k] = need_ws ? (cells + marg + 63) / 64 * 64 : 0;
    }
    int64_t ws_used = 0, part_used = 0, n_items_total = 0, nb = 0, n_pitems = 0;
    int64_t grp_begin[kGroups] = {0, 0, 0, 0, 0}, grp_len[kGroups] = {0, 0, 0, 0, 0};
    S.bytes = 0.0;
    S.pbytes = 0.0;
    for (; nb < m; ++nb) {
        DevTest &T = dt[nb];
        if (gidx[nb] == kGroupMfma) {
            T.item0 = -1;
            continue;
        }
        if (need[nb] > 0) {
            if (ws_used + need[nb] > kWsMax && nb > 0) break;
            if (need[nb] > kWsMax) return fail(ctx, FPGM_ETOOBIG, "test workspace exceeds 4 GiB");
            T.ws_off = ws_used;
            ws_used += need[nb];
        }
        if (pneed[nb] > 0) {  // the partition buffer: the test's offset in ws_off
            if (part_used + pneed[nb] > kWsMax && nb > 0) break;
            if (pneed[nb] > kWsMax) return fail(ctx, FPGM_ETOOBIG, "test partition buffer exceeds 4 GiB");
            T.ws_off = part_used;
            part_used += pneed[nb];
            const int64_t ch = (ctx->inst[T.inst].n_rows + kPartRows - 1) / kPartRows;
            n_pitems += (ch + kPartItemChunks - 1) / kPartItemChunks;
        }
        const int64_t cnt = (int64_t)T.n_chunks * T.n_slices;
        T.item0 = (int32_t)grp_len[gidx[nb]];  // within the group; shifted by the group start below
        grp_len[gidx[nb]] += cnt;
        n_items_total += cnt;
        if (n_items_total > ((int64_t)1 << 30)) {
            ++nb;
            break;
        }
    }
    const int64_t i1 = i0 + nb;
    {  // algorithmic and packed bytes of the launch sequence's tests
        double b = 0.0, pb = 0.0;
        const int nth3 = (nb >= kParallelTests) ? ctx->host_threads : 1;
#pragma omp parallel for if (nth3 > 1) num_threads(nth3) schedule(static) reduction(+ : b, pb)
        for (int64_t k = 0; k < nb; ++k) {
            const double rows = (double)ctx->inst[dt[k].inst].n_rows;
            b += (double)(dt[k].d + 2) * rows;
            pb += packed_row_bytes(ctx, tests[i0 + k], zpool) * rows;
        }
        S.bytes = b;
        S.pbytes = pb;
    }
    // work items grouped by sink, one launch per group; within a group, tests keep the
    // caller's order (neighbouring tests share columns)
    HIP_TRY(S.h_items.ensure(std::max<int64_t>(n_items_total, 1), ctx->stream), "alloc pinned items");
    int64_t acc = 0;
    for (int gi = 0; gi < kGroups; ++gi) {
        grp_begin[gi] = acc;
        acc += grp_len[gi];
    }
    const int nth2 = (nb >= kParallelTests) ? ctx->host_threads : 1;
#pragma omp parallel for if (nth2 > 1) num_threads(nth2) schedule(static)
    for (int64_t k = 0; k < nb; ++k) {
        if (gidx[k] == kGroupMfma) continue;
        dt[k].item0 += (int32_t)grp_begin[gidx[k]];
        const int64_t cnt = (int64_t)dt[k].n_chunks * dt[k].n_slices;
        for (int64_t j = 0; j < cnt; ++j) S.h_items.p[dt[k].item0 + j] = ((uint64_t)k << 32) | (uint64_t)j;
    }
    const int64_t ni = n_items_total;
    // partitioning work items: every row chunk of every partitioned test
    if (n_pitems > 0) {
        HIP_TRY(S.h_pitems.ensure(n_pitems, ctx->stream), "alloc pinned partition items");
        int64_t q = 0;
        for (int64_t k = 0; k < nb; ++k) {
            if (!(dt[k].flags & kTFPart)) continue;
            const int64_t ch = (ctx->inst[dt[k].inst].n_rows + kPartRows - 1) / kPartRows;
            for (int64_t c = 0; c < ch; c += kPartItemChunks) S.h_pitems.p[q++] = ((uint64_t)k << 32) | (uint64_t)c;
        }
        HIP_TRY(S.d_pitems.ensure(n_pitems, false, ctx->stream), "alloc partition items");
        HIP_TRY(ctx->d_part.ensure(part_used, false, ctx->stream), "alloc partition buffer");
    }
    // device-memory tables: z ranges for the grid-wide finalize
    int64_t n_wi = 0, n_wt = 0;
    for (int64_t k = 0; k < nb; ++k)
        if (dt[k].flags & kTFWide) {
            n_wi += (dt[k].dimz + kWideZ - 1) / kWideZ;
            ++n_wt;
        }
    if (n_wt > 0) {
        HIP_TRY(S.h_wide_items.ensure(n_wi, ctx->stream), "alloc pinned finalize items");
        HIP_TRY(S.h_wide_tests.ensure(n_wt, ctx->stream), "alloc pinned finalize tests");
        int64_t wi = 0, wt = 0;
        for (int64_t k = 0; k < nb; ++k) {
            if (!(dt[k].flags & kTFWide)) continue;
            WideTest W{(int32_t)k, (int32_t)wi, 0, 0};
            for (int64_t z0 = 0; z0 < dt[k].dimz; z0 += kWideZ)
                S.h_wide_items.p[wi++] =
                    WideItem{(int32_t)k, (int32_t)z0, (int32_t)std::min<int64_t>(z0 + kWideZ, dt[k].dimz), 0};
            W.nparts = (int32_t)wi - W.part0;
            S.h_wide_tests.p[wt++] = W;
        }
        HIP_TRY(S.d_wide_items.ensure(n_wi, false, ctx->stream), "alloc finalize items");
        HIP_TRY(S.d_wide_parts.ensure(n_wi, false, ctx->stream), "alloc finalize parts");
        HIP_TRY(S.d_wide_tests.ensure(n_wt, false, ctx->stream), "alloc finalize tests");
        HIP_TRY(hipMemcpyAsync(S.d_wide_items.p, S.h_wide_items.p, sizeof(WideItem) * n_wi, hipMemcpyHostToDevice,
                               ctx->stream),
                "upload finalize items");
        HIP_TRY(hipMemcpyAsync(S.d_wide_tests.p, S.h_wide_tests.p, sizeof(WideTest) * n_wt, hipMemcpyHostToDevice,
                               ctx->stream),
                "upload finalize tests");
    }
    // grouped level-1 tests: (instance, u, z) groups cut into chunks of at most kGrpCols
    // one-hot columns and kGrpMaxTests tests, the tests of a group in launch order
    int64_t n_gc = 0, n_gi = 0, n_gi1 = 0, g_rps = 0;
    S.g_tests = S.g_chunks = S.g_items = 0;
    S.g_ops = 0.0;
    if (grp_on) {
        std::vector<int32_t> &order = S.g_order;
        order.clear();
        for (int64_t k = 0; k < nb; ++k)
            if (gidx[k] == kGroupMfma) order.push_back((int32_t)k);
        const int64_t ng = (int64_t)order.size();
        if (ng > 0) {
            auto u_of = [&](int32_t k) {
                const HostTest &h = tests[i0 + k];
                return h.side ? h.y : h.x;
            };
            auto o_of = [&](int32_t k) {
                const HostTest &h = tests[i0 + k];
                return h.side ? h.x : h.y;
            };
            auto z_of = [&](int32_t k) { return zpool[tests[i0 + k].z_off]; };
            auto bucket = [&](int32_t k) { return (int64_t)tests[i0 + k].inst * ctx->n_vars + u_of(k); };
            // counting sort by (instance, u), stable in launch order; then each bucket by z
            const int64_t nbk = (int64_t)ctx->inst.size() * ctx->n_vars;
            std::vector<int32_t> &bs = S.g_bucket;
            bs.assign((size_t)nbk + 1, 0);
            for (int32_t k : order) ++bs[bucket(k) + 1];
            for (int64_t b = 0; b < nbk; ++b) bs[b + 1] += bs[b];
            HIP_TRY(S.h_gtest.ensure(ng, ctx->stream), "alloc pinned group tests");
            int32_t *sorted = S.h_gtest.p;
            {
                std::vector<int32_t> cur(bs.begin(), bs.end() - 1);
                for (int32_t k : order) sorted[cur[bucket(k)]++] = k;
            }
            const int nthg = (ng >= kParallelTests) ? ctx->host_threads : 1;
#pragma omp parallel for if (nthg > 1) num_threads(nthg) schedule(dynamic, 64)
            for (int64_t b = 0; b < nbk; ++b)
                if (bs[b + 1] - bs[b] > 1)
                    std::stable_sort(sorted + bs[b], sorted + bs[b + 1],
                                     [&](int32_t p, int32_t q) { return z_of(p) < z_of(q); });
            // chunks, those with one A fragment (c(u) c(z) <= 32) first
            HIP_TRY(S.h_gchunks.ensure(ng, ctx->stream), "alloc pinned group chunks");
            GroupChunk *gc = S.h_gchunks.p;
            for (int pass = 0; pass < 2; ++pass)
                for (int64_t p = 0; p < ng;) {
                    const int32_t k0 = sorted[p];
                    const int64_t b = bucket(k0);
                    const int32_t z = z_of(k0), u = u_of(k0);
                    const int32_t cu = ctx->card[u], cz = ctx->card[z];
                    int64_t q = p;  // end of the (instance, u, z) group
                    while (q < ng && bucket(sorted[q]) == b && z_of(sorted[q]) == z) ++q;
                    if ((cu * cz > 32) != (pass == 1)) {
                        p = q;
                        continue;
                    }
                    while (p < q) {
                        GroupChunk G{};
                        G.inst = tests[i0 + k0].inst;
                        G.u = u;
                        G.z = z;
                        G.cu = cu;
                        G.cz = cz;
                        G.t0 = (int32_t)p;
                        G.ws_off = -1;
                        while (p < q && G.n_t < kGrpMaxTests && G.n_cols + ctx->card[o_of(sorted[p])] <= kGrpCols) {
                            G.n_cols += ctx->card[o_of(sorted[p])];
                            ++G.n_t;
                            ++p;
                        }
                        gc[n_gc++] = G;
                    }
                    if (pass == 0) n_gi1 = n_gc;  // chunks with one A fragment so far
                }
            // row splits: enough work items to fill the CUs, at least 4096 rows each; split
            // chunks keep a partial table in the workspace (one done counter each)
            const int64_t target = (int64_t)ctx->n_cu * ctx->sched.items_per_cu;
            int64_t splits = std::max<int64_t>(1, (target + n_gc - 1) / n_gc);
            splits = std::min<int64_t>(splits, std::max<int64_t>(1, N / 4096));
            if (splits > 1 && ws_used + n_gc * (int64_t)(kGrpRows * kGrpCols) > kWsMax) splits = 1;
            g_rps = std::max<int64_t>(128, ((N + splits - 1) / splits + 127) / 128 * 128);
            HIP_TRY(S.h_gitems.ensure(n_gc * splits + 1, ctx->stream), "alloc pinned group items");
            const int64_t chunks1 = n_gi1;
            for (int64_t c = 0; c < n_gc; ++c) {
                GroupChunk &G = gc[c];
                const int64_t rows = ctx->inst[G.inst].n_rows;
                G.n_splits = (int32_t)std::max<int64_t>(1, (rows + g_rps - 1) / g_rps);
                G.done = (int32_t)(nb + c);
                if (G.n_splits > 1) {
                    G.ws_off = ws_used;
                    ws_used += kGrpRows * kGrpCols;
                }
                for (int32_t sp = 0; sp < G.n_splits; ++sp) {
                    S.h_gitems.p[n_gi++] = ((uint64_t)c << 32) | (uint64_t)sp;
                    const int64_t r = std::min<int64_t>(g_rps, rows - (int64_t)sp * g_rps);
                    S.g_ops += 2.0 * (double)std::max<int64_t>(r, 0) * (G.cu * G.cz > 32 ? 64 : 32) * kGrpCols;
                }
                if (c + 1 == chunks1) n_gi1 = n_gi;
            }
            if (chunks1 == 0) n_gi1 = 0;
            S.g_tests = ng;
            S.g_chunks = n_gc;
            S.g_items = n_gi;
            HIP_TRY(S.d_gtest.ensure(ng, false, ctx->stream), "alloc group tests");
            HIP_TRY(S.d_gchunks.ensure(n_gc, false, ctx->stream), "alloc group chunks");
            HIP_TRY(S.d_gitems.ensure(n_gi, false, ctx->stream), "alloc group items");
            HIP_TRY(hipMemcpyAsync(S.d_gtest.p, S.h_gtest.p, sizeof(int32_t) * ng, hipMemcpyHostToDevice, ctx->stream),
                    "upload group tests");
            HIP_TRY(hipMemcpyAsync(S.d_gchunks.p, gc, sizeof(GroupChunk) * n_gc, hipMemcpyHostToDevice, ctx->stream),
                    "upload group chunks");
            HIP_TRY(hipMemcpyAsync(S.d_gitems.p, S.h_gitems.p, sizeof(uint64_t) * n_gi, hipMemcpyHostToDevice,
                                   ctx->stream),
                    "upload group items");
        }
    }
    HIP_TRY(S.h_zpool.ensure(std::max<int64_t>(zlen, 1), ctx->stream), "alloc pinned z pool");
    if (zlen > 0) {
        const int nthz = (zlen >= 4 * kParallelTests) ? ctx->host_threads : 1;
#pragma omp parallel for if (nthz > 1) num_threads(nthz) schedule(static)
        for (int ch = 0; ch < nthz; ++ch) {
            const int64_t a = zlen * ch / nthz, b = zlen * (ch + 1) / nthz;
            std::copy(zpool + a, zpool + b, S.h_zpool.p + a);
        }
    }
    HIP_TRY(S.d_zpool.ensure(std::max<int64_t>(zlen, 1), false, ctx->stream), "alloc z pool");
    HIP_TRY(S.d_tests.ensure(nb, false, ctx->stream), "alloc tests");
    HIP_TRY(S.d_items.ensure(std::max<int64_t>(ni, 1), false, ctx->stream), "alloc items");
    if (ctx->skip_round > 0) HIP_TRY(S.d_items_live.ensure(std::max<int64_t>(ni, 1), false, ctx->stream), "alloc items");
    HIP_TRY(ctx->d_ws.ensure(std::max<int64_t>(ws_used, kWsInit), true, ctx->stream), "alloc workspace");
    HIP_TRY(ctx->d_done.ensure(std::max<int64_t>(nb + n_gc, 1024), true, ctx->stream), "alloc counters");
    HIP_TRY(S.d_res.ensure(nb, false, ctx->stream), "alloc results");
    if (!S.d_queue.p) {  // zero once here; afterwards every sequence's k_pvalue re-zeroes them
        HIP_TRY(S.d_queue.ensure(kGroups * (kXcds + 1), false, ctx->stream), "alloc work queues");
        HIP_TRY(hipMemsetAsync(S.d_queue.p, 0, sizeof(uint32_t) * S.d_queue.cap, ctx->stream), "clear work queues");
    }
    HIP_TRY(S.h_res.ensure(nb, ctx->stream), "alloc pinned results");
    // graph mode: everything from here to the read-back is captured (every buffer is
    // allocated above; no memset is captured); an error ends the capture
    const bool graph = ctx->graph != 0;
    struct CaptureScope {
        hipStream_t s;
        bool on = false;
        ~CaptureScope()
        {
            if (!on) return;
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
        }
    } capture{ctx->stream};
    if (graph) {
        HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed), "begin capture");
        capture.on = true;
    }
    S.graphed = graph;
    if (zlen > 0)
        HIP_TRY(hipMemcpyAsync(S.d_zpool.p, S.h_zpool.p, sizeof(int32_t) * zlen, hipMemcpyHostToDevice, ctx->stream),
                "upload z pool");
    HIP_TRY(hipMemcpyAsync(S.d_tests.p, S.h_tests.p, sizeof(DevTest) * nb, hipMemcpyHostToDevice, ctx->stream),
            "upload tests");
    HIP_TRY(hipMemcpyAsync(S.d_items.p, S.h_items.p, sizeof(uint64_t) * ni, hipMemcpyHostToDevice, ctx->stream),
            "upload items");
    if (n_pitems > 0)
        HIP_TRY(hipMemcpyAsync(S.d_pitems.p, S.h_pitems.p, sizeof(uint64_t) * n_pitems, hipMemcpyHostToDevice,
                               ctx->stream),
                "upload partition items");
    LaunchArgs a{};
    a.inst = ctx->d_inst;
    a.tests = S.d_tests.p;
    a.zpool = S.d_zpool.p;
    a.card = ctx->d_card;
    a.colw = ctx->d_colw;
    a.poff = ctx->d_poff;
    a.items = S.d_items.p;
    a.ws = ctx->d_ws.p;
    a.done = ctx->d_done.p;
    a.crit = ctx->d_crit;
    a.alpha = alpha;
    a.z_alpha = ctx->z_alpha;
    a.band = band > 0.f ? band : 0.02f;
    a.flags = flags;
    a.res = S.d_res.p;
    a.edge_done = ctx->skip_round > 0 ? ctx->d_edge_done.p : nullptr;
    a.round = ctx->skip_round;
    a.part = ctx->d_part.p;
    ctx->t_prep += now_s() - tp0;
    if (!graph) HIP_TRY(hipEventRecord(S.ev0, ctx->stream), "event");
    uint32_t *live_count = S.d_queue.p + kGroups * kXcds;
    if (a.edge_done) {
        // a speculative round: drop the items of tests whose edge an earlier round decided
        CompactArgs c{};
        c.in = S.d_items.p;
        c.out = S.d_items_live.p;
        c.tests = S.d_tests.p;
        c.edge_done = a.edge_done;
        c.round = a.round;
        c.res = S.d_res.p;
        int64_t max_len = 0;
        for (int gi = 0; gi < kGroups; ++gi) {
            c.begin[gi] = grp_begin[gi];
            c.len[gi] = grp_len[gi];
            max_len = std::max(max_len, grp_len[gi]);
        }
        c.count = live_count;
        HIP_TRY(launch_compact(c, max_len, ctx->stream), "launch compact");
        a.items = S.d_items_live.p;
    }
    if (n_gc > 0) {
        GroupArgs ga{};
        ga.a = a;
        ga.chunks = S.d_gchunks.p;
        ga.gtest = S.d_gtest.p;
        ga.rows_per_split = g_rps;
        ga.max_card = *std::max_element(ctx->card.begin(), ctx->card.end());
        if (!graph) HIP_TRY(hipEventRecord(S.ev_g0, ctx->stream), "event");
        ga.items = S.d_gitems.p;
        ga.n_items = n_gi1;
        HIP_TRY(launch_group_count(ga, 1, ctx->stream), "launch group count");
        ga.items = S.d_gitems.p + n_gi1;
        ga.n_items = n_gi - n_gi1;
        HIP_TRY(launch_group_count(ga, 2, ctx->stream), "launch group count");
        if (!graph) HIP_TRY(hipEventRecord(S.ev_g1, ctx->stream), "event");
        ctx->k_launches += (n_gi1 > 0) + (n_gi > n_gi1);
    }
    // one launch per sink group.  The LDS sinks (lane, lane-wide, shared) run one after
    // another on the context stream: run together they compete for the same CUs' LDS and
    // measured slower.  The slice and device-memory sinks, bound by LDS capacity and device
    // atomics rather than by the CUs they occupy, get a stream each, so that they overlap
    // the rest, and so does an LDS group of fewer than 16 work items per CU (its last
    // items would leave the chip idle while they finish, and the next launch on the
    // context stream would wait for them; 300 variables x 500 K rows: 8.68 -> 8.08 ms,
    // against 1 item per CU); the context stream joins them before the p-value kernel.
    // Tests of different groups touch disjoint workspace, counters and results.
    int n_grp = 0;
    for (int gi = 0; gi < kGroups; ++gi) n_grp += (grp_len[gi] > 0);
    const bool concurrent = ctx->sched.concurrent && n_grp > 1;
    if (concurrent) HIP_TRY(hipEventRecord(S.ev_fork, ctx->stream), "event");
    bool joined[kGroups] = {false, false, false, false, false, false};
    for (int gi = 0; gi < kGroups; ++gi) {
        if (grp_len[gi] == 0) continue;
        hipStream_t gs = ctx->stream;
        int side = kGroupSink[gi] == kSinkSlice ? 0 : kGroupSink[gi] == kSinkGlobal ? 1 : -1;
        if (side < 0 && grp_len[gi] < ctx->n_cu * ctx->sched.side_cu_mult)
            side = gi == kGroupBits ? 5 : 2 + gi;  // lane, lane-wide, shared: 2, 3, 4; bit planes: 5
        if (concurrent && side >= 0) {
            gs = ctx->side[side];
            HIP_TRY(hipStreamWaitEvent(gs, S.ev_fork, 0), "stream wait");
            joined[gi] = true;
        }
        if (!graph) HIP_TRY(hipEventRecord(S.ev_grp[gi], gs), "event");
        a.item_begin = grp_begin[gi];
        a.n_items = grp_len[gi];
        a.rows_per_chunk = rpc[gi];
        a.queue = S.d_queue.p + gi * kXcds;
        a.n_items_dev = a.edge_done ? live_count + gi : nullptr;
        if (kGroupSink[gi] == kSinkSlice && n_pitems > 0) {
            // the partitioned tests' rows grouped by z slice, before the slice launch counts them
            PartArgs pa{};
            pa.a = a;
            pa.items = S.d_pitems.p;
            pa.n_items = n_pitems;
            const int pgrid = (int)std::min<int64_t>(n_pitems, (int64_t)ctx->n_cu * 3);
            HIP_TRY(launch_part_keys(pa, ctx->d_pcols != nullptr, pgrid, gs), "launch partition");
            ctx->k_launches += 1;
        }
        const int sink = (gi == 1 && lane_w128) ? kSinkLaneW128 : kGroupSink[gi];
        const int64_t wg_per_cu = wpc_of(ctx, gi);
        int64_t grid = std::min<int64_t>(a.n_items, (int64_t)ctx->n_cu * wg_per_cu);
        grid = std::max<int64_t>(8, (grid + 7) / 8 * 8);
        HIP_TRY(launch_count_stat(a, sink, ctx->d_pcols != nullptr, (int)grid, gs), "launch count_stat");
        ctx->k_launches += 1;
        if (sink == kSinkGlobal && n_wt > 0) {
            HIP_TRY(launch_finalize_wide(a, S.d_wide_items.p, n_wi, S.d_wide_parts.p, S.d_wide_tests.p, n_wt, gs),
                    "launch finalize");
            ctx->k_launches += 2;
        }
        if (!graph) HIP_TRY(hipEventRecord(S.ev_gend[gi], gs), "event");
        if (joined[gi]) HIP_TRY(hipEventRecord(S.ev_join[gi], gs), "event");
    }
    for (int gi = 0; gi < kGroups; ++gi)
        if (joined[gi]) HIP_TRY(hipStreamWaitEvent(ctx->stream, S.ev_join[gi], 0), "stream wait");
    if (!graph) HIP_TRY(hipEventRecord(S.ev1, ctx->stream), "event");
    HIP_TRY(launch_pvalue(S.d_res.p, nb, alpha, S.d_tests.p, a.edge_done, a.round, S.d_queue.p, kGroups * (kXcds + 1),
                          ctx->stream),
            "launch pvalue");
    HIP_TRY(hipMemcpyAsync(S.h_res.p, S.d_res.p, sizeof(DevResult) * nb, hipMemcpyDeviceToHost, ctx->stream),
            "download results");
    if (graph) {
        // the slot's executable graph is updated in place when the sequence has the same
        // shape as the last one, else instantiated anew; ev0..ev1 span the whole graph
        hipGraph_t gr = nullptr;
        capture.on = false;
        HIP_TRY(hipStreamEndCapture(ctx->stream, &gr), "end capture");
        hipError_t e = hipSuccess;
        if (S.gexec) {
            hipGraphNode_t err_node = nullptr;
            hipGraphExecUpdateResult ur = hipGraphExecUpdateSuccess;
            if (hipGraphExecUpdate(S.gexec, gr, &err_node, &ur) != hipSuccess || ur != hipGraphExecUpdateSuccess) {
                (void)hipGetLastError();
                (void)hipGraphExecDestroy(S.gexec);
                S.gexec = nullptr;
            } else {
                ++ctx->graph_updates;
            }
        }
        if (!S.gexec) {
            e = hipGraphInstantiate(&S.gexec, gr, nullptr, nullptr, 0);
            ++ctx->graph_instantiations;
        }
        (void)hipGraphDestroy(gr);
        HIP_TRY(e, "instantiate graph");
        HIP_TRY(hipEventRecord(S.ev0, ctx->stream), "event");
        HIP_TRY(hipGraphLaunch(S.gexec, ctx->stream), "launch graph");
        HIP_TRY(hipEventRecord(S.ev1, ctx->stream), "event");
    }
    HIP_TRY(hipEventRecord(S.ev_done, ctx->stream), "event");
    S.n = nb;
    // per-sink figures of the sequence (fpgm_last_sink_stats, trace lines)
    for (int gi = 0; gi < kGroups; ++gi) {
        Slot::GroupInfo &G = S.grp[gi];
        G = Slot::GroupInfo{};
        G.items = grp_len[gi];
    }
    for (int64_t k = 0; k < nb; ++k) {
        if (gidx[k] == kGroupMfma) continue;
        Slot::GroupInfo &G = S.grp[gidx[k]];
        const double c = (double)dt[k].dimz * dt[k].cx * dt[k].cy;
        const double rows = (double)ctx->inst[dt[k].inst].n_rows;
        G.tests += 1;
        G.dsum += dt[k].d;
        G.csum += c;
        G.cmax = std::max(G.cmax, c);
        G.bytes += (double)(dt[k].d + 2) * rows;
        G.pbytes += packed_row_bytes(ctx, tests[i0 + k], zpool) * rows;
        G.rows += rows;
    }
    S.dev_round = false;
    S.pending = true;
    if (keep_ws) *keep_ws = dt[0].ws_off;
    *i1_out = i1;
    return FPGM_OK;
}

// Waits for the slot's launch sequence; its results are then in S.h_res[0, S.n).
fpgm_status batch_wait(fpgm_ctx *ctx, Slot &S)
{
    if (!S.pending) return FPGM_OK;
    const double tw0 = now_s();
    HIP_TRY(hipEventSynchronize(S.ev_done), "synchronize");
    ctx->t_wait += now_s() - tw0;
    S.pending = false;
    if (S.dev_round) {  // a device-generated round: its figures come from the device
        const RoundStats &R = *S.h_rstats.p;
        S.bytes = 0.0;
        S.pbytes = 0.0;
        for (int gi = 0; gi < kGroups; ++gi) S.grp[gi] = Slot::GroupInfo{};
        for (int dg = 0; dg < kRouteGroups; ++dg) {
            Slot::GroupInfo &G = S.grp[dg == round_group(kRouteBits) ? kGroupBits : dg];
            G.tests = (int64_t)R.tests[dg];
            G.items = (int64_t)R.items[dg];
            G.dsum = G.tests * S.dev_d;
            G.rows = (double)R.rows[dg];
            G.bytes = (double)R.alg[dg];
            G.pbytes = (double)R.pbits[dg] / 8.0;
            S.bytes += G.bytes;
            S.pbytes += G.pbytes;
        }
    }
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, S.ev0, S.ev1), "event time");
    ctx->k_ms += ms;
    ctx->k_bytes += S.bytes;
    ctx->k_pbytes += S.pbytes;
    if (S.g_chunks > 0 && !S.graphed) {
        float gms = 0.f;
        HIP_TRY(hipEventElapsedTime(&gms, S.ev_g0, S.ev_g1), "event time");
        ctx->grp_ms += gms;
        ctx->grp_ops += S.g_ops;
        ctx->grp_tests += S.g_tests;
        ctx->grp_launches += 1;
        if (ctx->trace)
            fprintf(stderr, "[fpgm]   launch grouped   tests %7lld chunks %7lld items %8lld: %8.3f ms, %6.1f TOP/s int8\n",
                    (long long)S.g_tests, (long long)S.g_chunks, (long long)S.g_items, gms,
                    gms > 0 ? S.g_ops / (gms * 1e9) : 0.0);
    }
    {
        static const char *names[kGroups] = {"lane", "lane-wide", "shared", "slice", "device", "bits"};
        for (int gi = 0; gi < kGroups; ++gi) {
            const Slot::GroupInfo &G = S.grp[gi];
            if (G.items == 0) continue;
            float gms = 0.f;
            if (!S.graphed) HIP_TRY(hipEventElapsedTime(&gms, S.ev_grp[gi], S.ev_gend[gi]), "event time");
            fpgm_sink_stats &A = ctx->sink[gi];
            A.launches += 1;
            A.tests += G.tests;
            A.items += G.items;
            A.ms += gms;
            A.alg_bytes += G.bytes;
            A.packed_bytes += G.pbytes;
            A.row_tests += G.rows;
            if (!ctx->trace) continue;
            fprintf(stderr,
                    "[fpgm]   launch %-9s tests %7lld items %8lld mean |Z| %.2f cells mean %.0f max %.0f: "
                    "%8.3f ms, %7.0f GB/s\n",
                    names[gi], (long long)G.tests, (long long)G.items, (double)G.dsum / G.tests, G.csum / G.tests,
                    G.cmax, gms, gms > 0 ? G.bytes / (gms * 1e6) : 0.0);
        }
    }
    return FPGM_OK;
}

// A level run on device-generated rounds: what every round of it shares.
struct DevLevel {
    int d = 0;
    int64_t rows_max = 0;           // longest instance still running
    int64_t need_max = 0;           // largest table + marginals of the level (int32 entries, a multiple of 64)
    bool w128_shared = false;       // 97..128-cell tables go to the shared group
    bool lane_w128 = false;         // the lane-wide group runs the 64 KiB instance
    bool launch[kRouteGroups] = {true, true, true, true};  // sink groups the level's tables can reach
    int64_t walk_tests = 0;         // tests of every edge's whole walk (the level's most), saturated
    std::vector<int32_t> nbr;       // every running instance's frozen adjacency, concatenated
    std::vector<int64_t> nbr_base;  // instance -> its first entry in nbr
    std::vector<int64_t> binom;     // C(n, k), n <= amax, k <= d, saturated
    int64_t C(int64_t n, int k) const { return binom[(size_t)n * (size_t)(d + 1) + (size_t)k]; }
};

// Enqueues one device-generated round: segments up, expansion, one counting launch per
// sink group (item counts read on the device), p-values, the per-segment reduction, and
// the segments' outcomes and the round's figures back into the slot's pinned buffers.
fpgm_status dev_round_submit(fpgm_ctx *ctx, Slot &S, const DevLevel &L, const std::vector<RoundSeg> &segs,
                             int64_t n_tests, int32_t round, double alpha, uint32_t flags, float band)
{
    fpgm_status st = batch_wait(ctx, S);
    if (st != FPGM_OK) return st;
    const double tp0 = now_s();
    const int d = L.d;
    const int64_t n_segs = (int64_t)segs.size();
    if (n_tests * (int64_t)d >= (int64_t)INT32_MAX || n_tests >= (int64_t)INT32_MAX)
        return fail(ctx, FPGM_ETOOBIG, "device round too large");
    // rows per chunk as for the host's rounds: enough work items over the round's tests
    // to fill every CU
    const int64_t N = L.rows_max;
    const int64_t steps = std::max<int64_t>(1, (N + kRowsPerStep - 1) / kRowsPerStep);
    const int64_t target_items = (int64_t)ctx->n_cu * ctx->sched.items_per_cu;
    const int64_t cpt = std::min(steps, std::max<int64_t>(1, (target_items + n_tests - 1) / n_tests));
    const int64_t rpc = std::min<int64_t>((steps + cpt - 1) / cpt * kRowsPerStep, kLaneMaxRowsPerChunk);
    const int64_t chunks_max = std::max<int64_t>(1, (N + rpc - 1) / rpc);
    const int64_t cap = n_tests * chunks_max;  // items of one group at most
    // a test cut into chunks keeps its table in the workspace; single-chunk tests never
    // need it (the level's tables and marginals fit LDS)
    const int64_t ws_need = chunks_max > 1 ? n_tests * L.need_max : 0;
    if (ws_need > kWsMax) return fail(ctx, FPGM_ETOOBIG, "device round workspace exceeds 4 GiB");
    HIP_TRY(S.h_segs.ensure(n_segs, ctx->stream), "alloc pinned segments");
    HIP_TRY(S.h_seg_out.ensure(n_segs, ctx->stream), "alloc pinned segment results");
    HIP_TRY(S.h_rstats.ensure(1, ctx->stream), "alloc pinned round figures");
    HIP_TRY(S.d_segs.ensure(n_segs, false, ctx->stream), "alloc segments");
    HIP_TRY(S.d_seg_out.ensure(n_segs, false, ctx->stream), "alloc segment results");
    HIP_TRY(S.d_t2s.ensure(n_tests, false, ctx->stream), "alloc test segments");
    HIP_TRY(S.d_rstats.ensure(1, false, ctx->stream), "alloc round figures");
    HIP_TRY(S.d_tests.ensure(n_tests, false, ctx->stream), "alloc tests");
    HIP_TRY(S.d_zpool.ensure(std::max<int64_t>(n_tests * d, 1), false, ctx->stream), "alloc z pool");
    HIP_TRY(S.d_items.ensure(kRouteGroups * cap, false, ctx->stream), "alloc items");
    HIP_TRY(S.d_res.ensure(n_tests, false, ctx->stream), "alloc results");
    HIP_TRY(ctx->d_ws.ensure(std::max<int64_t>(ws_need, kWsInit), true, ctx->stream), "alloc workspace");
    HIP_TRY(ctx->d_done.ensure(std::max<int64_t>(n_tests, 1024), true, ctx->stream), "alloc counters");
    if (!S.d_queue.p) {  // zero once here; afterwards every sequence's k_pvalue re-zeroes them
        HIP_TRY(S.d_queue.ensure(kGroups * (kXcds + 1), false, ctx->stream), "alloc work queues");
        HIP_TRY(hipMemsetAsync(S.d_queue.p, 0, sizeof(uint32_t) * S.d_queue.cap, ctx->stream), "clear work queues");
    }
    std::copy(segs.begin(), segs.end(), S.h_segs.p);
    HIP_TRY(hipMemcpyAsync(S.d_segs.p, S.h_segs.p, sizeof(RoundSeg) * n_segs, hipMemcpyHostToDevice, ctx->stream),
            "upload segments");
    HIP_TRY(hipEventRecord(S.ev0, ctx->stream), "event");
    HIP_TRY(launch_zero(reinterpret_cast<uint32_t *>(S.d_rstats.p), sizeof(RoundStats) / sizeof(uint32_t), ctx->stream),
            "clear round figures");
    uint32_t *group_count = S.d_queue.p + kGroups * kXcds;  // zero: cleared by the last sequence's k_pvalue
    ExpandArgs e{};
    e.segs = S.d_segs.p;
    e.n_segs = n_segs;
    e.n_tests = n_tests;
    e.d = d;
    e.binom_stride = d + 1;
    e.nbr = ctx->d_nbr.p;
    e.binom = ctx->d_binom.p;
    e.card = ctx->d_card;
    e.colw = ctx->d_colw;
    e.inst = ctx->d_inst;
    e.edge_done = ctx->d_edge_done.p;
    e.round = round;
    e.w128_shared = L.w128_shared ? 1 : 0;
    e.launch_mask = 0u;
    for (int dg = 0; dg < kRouteGroups; ++dg) e.launch_mask |= L.launch[dg] ? 1u << dg : 0u;
    e.rp = route_params(ctx);
    e.rows_per_chunk = rpc;
    e.tests = S.d_tests.p;
    e.zpool = S.d_zpool.p;
    e.t2s = S.d_t2s.p;
    e.res = S.d_res.p;
    e.items = S.d_items.p;
    e.group_cap = cap;
    e.group_count = group_count;
    e.ws_cap = ws_need;
    e.seg_out = S.d_seg_out.p;
    e.stats = S.d_rstats.p;
    HIP_TRY(launch_expand(e, ctx->stream), "launch expand");
    LaunchArgs a{};
    a.inst = ctx->d_inst;
    a.tests = S.d_tests.p;
    a.zpool = S.d_zpool.p;
    a.card = ctx->d_card;
    a.colw = ctx->d_colw;
    a.poff = ctx->d_poff;
    a.items = S.d_items.p;
    a.ws = ctx->d_ws.p;
    a.done = ctx->d_done.p;
    a.crit = ctx->d_crit;
    a.alpha = alpha;
    a.z_alpha = ctx->z_alpha;
    a.band = band > 0.f ? band : 0.02f;
    a.flags = flags;
    a.res = S.d_res.p;
    a.edge_done = ctx->d_edge_done.p;
    a.round = round;
    a.rows_per_chunk = rpc;
    // every group the level's tables can reach is launched (a launch whose count reads
    // zero ends at once)
    for (int dg = 0; dg < kRouteGroups; ++dg) {
        if (!L.launch[dg]) continue;
        const int gi = dg == round_group(kRouteBits) ? kGroupBits : dg;  // the host's group
        HIP_TRY(hipEventRecord(S.ev_grp[gi], ctx->stream), "event");
        a.item_begin = dg * cap;
        a.n_items = cap;
        a.queue = S.d_queue.p + gi * kXcds;
        a.n_items_dev = group_count + dg;
        const int sink = (gi == 1 && L.lane_w128) ? kSinkLaneW128 : kGroupSink[gi];
        int64_t grid = std::min<int64_t>(cap, (int64_t)ctx->n_cu * wpc_of(ctx, gi));
        grid = std::max<int64_t>(8, (grid + 7) / 8 * 8);
        HIP_TRY(launch_count_stat(a, sink, ctx->d_pcols != nullptr, (int)grid, ctx->stream), "launch count_stat");
        HIP_TRY(hipEventRecord(S.ev_gend[gi], ctx->stream), "event");
        ctx->k_launches += 1;
    }
    HIP_TRY(hipEventRecord(S.ev1, ctx->stream), "event");
    HIP_TRY(launch_pvalue(S.d_res.p, n_tests, alpha, S.d_tests.p, a.edge_done, round, S.d_queue.p,
                          kGroups * (kXcds + 1), ctx->stream),
            "launch pvalue");
    ReduceArgs r{};
    r.segs = S.d_segs.p;
    r.t2s = S.d_t2s.p;
    r.res = S.d_res.p;
    r.n_tests = n_tests;
    r.seg_out = S.d_seg_out.p;
    r.stats = S.d_rstats.p;
    HIP_TRY(launch_round_reduce(r, ctx->stream), "launch round reduce");
    HIP_TRY(hipMemcpyAsync(S.h_seg_out.p, S.d_seg_out.p, sizeof(int32_t) * n_segs, hipMemcpyDeviceToHost, ctx->stream),
            "download segment results");
    HIP_TRY(hipMemcpyAsync(S.h_rstats.p, S.d_rstats.p, sizeof(RoundStats), hipMemcpyDeviceToHost, ctx->stream),
            "download round figures");
    HIP_TRY(hipEventRecord(S.ev_done, ctx->stream), "event");
    S.n = n_tests;
    S.n_segs = n_segs;
    S.dev_round = true;
    S.dev_d = d;
    S.graphed = false;
    S.g_tests = S.g_chunks = S.g_items = 0;
    S.g_ops = 0.0;
    S.pending = true;
    ctx->t_prep += now_s() - tp0;
    return FPGM_OK;
}

// Runs tests[0..n) on the device and waits; results land in out (host).
fpgm_status run_batch(fpgm_ctx *ctx, const HostTest *tests, int64_t n, const int32_t *zpool, int64_t zlen,
                      double alpha, uint32_t flags, float band, int32_t tflags, std::vector<DevResult> &out,
                      int64_t rows_per_chunk_forced = 0, int64_t *keep_ws = nullptr)
{
    out.resize((size_t)n);
    Slot &S = ctx->slot[0];
    fpgm_status st = batch_wait(ctx, S);
    if (st != FPGM_OK) return st;
    for (int64_t i0 = 0; i0 < n;) {
        int64_t i1 = i0;
        st = batch_submit(ctx, S, tests, n, zpool, zlen, alpha, flags, band, tflags, rows_per_chunk_forced, i0, &i1,
                          i0 == 0 ? keep_ws : nullptr);
        if (st != FPGM_OK) return st;
        st = batch_wait(ctx, S);
        if (st != FPGM_OK) return st;
        std::copy(S.h_res.p, S.h_res.p + (i1 - i0), out.begin() + i0);
        i0 = i1;
    }
    return FPGM_OK;
}

void reset_kernel_stats(fpgm_ctx *ctx)
{
    ctx->k_ms = 0.0;
    ctx->k_bytes = 0.0;
    ctx->k_pbytes = 0.0;
    ctx->k_launches = 0;
    ctx->g_ms = 0.0;
    ctx->g_ops = 0.0;
    ctx->g_pair_ms = 0.0;
    ctx->g_launches = 0;
    for (int g = 0; g < kGroups; ++g) {
        ctx->sink[g] = fpgm_sink_stats{};
        ctx->sink[g].sink = g;
    }
    ctx->grp_ms = 0.0;
    ctx->grp_ops = 0.0;
    ctx->grp_tests = 0;
    ctx->grp_launches = 0;
}

// Level 0 of every instance on the MFMA units: C_i = O_i^T O_i, then one thread per
// (instance, pair), every pair x < y in lexicographic order.  Returns false in *used when
// the one-hot width is too large for dense C blocks (the caller then counts pair by pair).
// Results: one flag byte per (instance, pair) (bit 0 independent, bit 1 in the recheck
// band) in ctx->h_l0_flags, instance i's pairs at i * n(n-1)/2, and the summary in
// ctx->h_l0_sum.
fpgm_status run_level0_gram(fpgm_ctx *ctx, double alpha, uint32_t flags, float band, bool *used)
{
    *used = false;
    Slot &S = ctx->slot[0];
    {
        fpgm_status st = batch_wait(ctx, S);
        if (st != FPGM_OK) return st;
    }
    const int n = ctx->n_vars;
    const int64_t n_inst = (int64_t)ctx->inst.size();
    std::vector<int32_t> off(n + 1, 0);
    for (int v = 0; v < n; ++v) off[v + 1] = off[v] + ctx->card[v];
    const int64_t W = off[n];
    const int64_t Wp = std::max<int64_t>(kGramTile, (W + kGramTile - 1) / kGramTile * kGramTile);
    // dense C blocks of every instance: at most 32768 one-hot columns and 8 GiB in all
    if (Wp > 32768 || (double)n_inst * Wp * Wp * 4.0 > 8.0 * (1 << 30)) return FPGM_OK;
    *used = true;
    fpgm_status st = ensure_crit(ctx, alpha);
    if (st != FPGM_OK) return st;
    std::vector<int32_t> rowvar((size_t)Wp, 0);
    std::vector<uint8_t> rowstate((size_t)Wp, 255);
    for (int v = 0; v < n; ++v)
        for (int a = 0; a < ctx->card[v]; ++a) {
            rowvar[off[v] + a] = v;
            rowstate[off[v] + a] = (uint8_t)a;
        }
    const int nt = (int)(Wp / kGramTile);
    std::vector<int32_t> tiles;
    for (int i = 0; i < nt; ++i)
        for (int j = i; j < nt; ++j) {
            tiles.push_back(i);
            tiles.push_back(j);
        }
    const int64_t n_tp = (int64_t)tiles.size() / 2;
    int64_t N = 0;
    double rows_all = 0.0;
    for (const InstHost &I : ctx->inst) {
        N = std::max(N, I.n_rows);
        rows_all += (double)I.n_rows;
    }
    // row splits of the longest instance: enough items to fill the CUs (splits past a
    // shorter instance's rows end at once)
    const int64_t ksteps = std::max<int64_t>(1, (N + 31) / 32);
    const int64_t target = (int64_t)ctx->n_cu * ctx->sched.gram_items_per_cu;
    const int64_t splits = std::min<int64_t>(ksteps, std::max<int64_t>(1, (target + n_tp * n_inst - 1) / (n_tp * n_inst)));
    const int64_t kps = (ksteps + splits - 1) / splits * 32;
    const int64_t n_items = n_inst * n_tp * splits;
    if (n_items > INT32_MAX) return fail(ctx, FPGM_ETOOBIG, "too many level-0 Gram work items");
    const int64_t np = (int64_t)n * (n - 1) / 2;
    HIP_TRY(ctx->d_C.ensure(n_inst * Wp * Wp, false, ctx->stream), "alloc gram");
    HIP_TRY(ctx->d_rowvar.ensure(Wp, false, ctx->stream), "alloc gram rows");
    HIP_TRY(ctx->d_rowstate.ensure(Wp, false, ctx->stream), "alloc gram rows");
    HIP_TRY(ctx->d_tiles.ensure((int64_t)tiles.size(), false, ctx->stream), "alloc gram tiles");
    HIP_TRY(ctx->d_off.ensure(n + 1, false, ctx->stream), "alloc gram offsets");
    HIP_TRY(S.d_res.ensure(np * n_inst, false, ctx->stream), "alloc results");
    HIP_TRY(ctx->d_l0_flags.ensure(np * n_inst, false, ctx->stream), "alloc level-0 flags");
    HIP_TRY(ctx->h_l0_flags.ensure(np * n_inst, ctx->stream), "alloc pinned level-0 flags");
    HIP_TRY(ctx->d_l0_sum.ensure(1, false, ctx->stream), "alloc level-0 summary");
    HIP_TRY(ctx->h_l0_sum.ensure(1, ctx->stream), "alloc pinned level-0 summary");
    HIP_TRY(hipMemcpyAsync(ctx->d_rowvar.p, rowvar.data(), sizeof(int32_t) * Wp, hipMemcpyHostToDevice, ctx->stream),
            "upload gram rows");
    HIP_TRY(hipMemcpyAsync(ctx->d_rowstate.p, rowstate.data(), Wp, hipMemcpyHostToDevice, ctx->stream),
            "upload gram rows");
    HIP_TRY(hipMemcpyAsync(ctx->d_tiles.p, tiles.data(), sizeof(int32_t) * tiles.size(), hipMemcpyHostToDevice,
                           ctx->stream),
            "upload gram tiles");
    HIP_TRY(hipMemcpyAsync(ctx->d_off.p, off.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, ctx->stream),
            "upload offsets");
    HIP_TRY(hipMemsetAsync(ctx->d_C.p, 0, sizeof(int32_t) * n_inst * Wp * Wp, ctx->stream), "clear gram");
    GramArgs g{};
    const InstHost &I0 = ctx->inst[0];
    g.cols = I0.cols;
    g.stride = I0.stride;
    g.n_rows = I0.n_rows;
    g.insts = n_inst > 1 ? ctx->d_inst : nullptr;
    g.n_tp = (int32_t)n_tp;
    g.c_stride = Wp * Wp;
    g.row_var = ctx->d_rowvar.p;
    g.row_state = ctx->d_rowstate.p;
    g.tiles = ctx->d_tiles.p;
    g.Wp = (int32_t)Wp;
    g.k_splits = (int32_t)splits;
    g.k_per_split = kps;
    g.C = ctx->d_C.p;
    g.max_card = *std::max_element(ctx->card.begin(), ctx->card.end());
    PairArgs p{};
    p.C = ctx->d_C.p;
    p.Wp = (int32_t)Wp;
    p.off = ctx->d_off.p;
    p.card = ctx->d_card;
    p.px = nullptr;
    p.py = nullptr;
    p.n_vars = n;
    p.n_pairs = np;
    p.n_inst = (int32_t)n_inst;
    p.c_stride = Wp * Wp;
    p.crit = ctx->d_crit;
    p.alpha = alpha;
    p.z_alpha = ctx->z_alpha;
    p.band = band > 0.f ? band : 0.02f;
    p.flags = flags;
    p.res = S.d_res.p;
    hipEvent_t e_mid = nullptr;
    HIP_TRY(hipEventCreate(&e_mid), "event");
    HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream), "event");
    HIP_TRY(launch_gram_l0(g, (int)n_items, ctx->stream), "launch gram");
    HIP_TRY(hipEventRecord(e_mid, ctx->stream), "event");
    HIP_TRY(launch_pairs_l0(p, ctx->stream), "launch pairs");
    HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream), "event");
    HIP_TRY(launch_pvalue(S.d_res.p, np * n_inst, alpha, nullptr, nullptr, 0, nullptr, 0, ctx->stream),
            "launch pvalue");
    HIP_TRY(hipMemsetAsync(ctx->d_l0_sum.p, 0, sizeof(PairSummary), ctx->stream), "clear level-0 summary");
    HIP_TRY(launch_pack_l0(S.d_res.p, np * n_inst, ctx->d_l0_flags.p, ctx->d_l0_sum.p, ctx->stream), "launch pack");
    HIP_TRY(hipMemcpyAsync(ctx->h_l0_flags.p, ctx->d_l0_flags.p, (size_t)(np * n_inst), hipMemcpyDeviceToHost,
                           ctx->stream),
            "download level-0 flags");
    HIP_TRY(hipMemcpyAsync(ctx->h_l0_sum.p, ctx->d_l0_sum.p, sizeof(PairSummary), hipMemcpyDeviceToHost, ctx->stream),
            "download level-0 summary");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "synchronize");
    float ms_g = 0.f, ms_p = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms_g, ctx->ev0, e_mid), "event time");
    HIP_TRY(hipEventElapsedTime(&ms_p, e_mid, ctx->ev1), "event time");
    (void)hipEventDestroy(e_mid);
    ctx->g_ms += ms_g;
    ctx->g_pair_ms += ms_p;
    ctx->g_ops += 2.0 * (double)n_tp * kGramTile * kGramTile * rows_all;
    ctx->g_launches += 1;
    return FPGM_OK;
}

fpgm_result to_result(const DevResult &r)
{
    fpgm_result o{};
    o.df = r.df;
    o.indep = r.indep;
    o.in_band = r.in_band;
    o.exact = r.exact;
    o.g2 = r.exact ? r.g2_64 : (double)r.g2_32;
    o.p_value = r.p64;
    return o;
}

// ---------------------------------------------------------------- skeleton

// Per-edge walk over the conditioning sets of one level in the reference's order.
struct EdgeWalk {
    int32_t x, y;
    int32_t side;      // 0: subsets of adj(x)\{y}; 1: subsets of adj(y)\{x}; 2: exhausted
    int32_t a;         // size of the current side's set
    int32_t skip;      // position of the excluded node in the neighbour list
    int32_t u;         // node whose neighbours are being enumerated
    int64_t tests;     // reference-equivalent tests so far
    int32_t removed;
    int64_t zsep;      // offset of the sepset in the level's z store
    int64_t rank;      // lexicographic rank of the side's next subset (device-generated rounds)
};

struct Frozen {
    std::vector<int32_t> off, nb;
    int32_t deg(int u) const { return off[u + 1] - off[u]; }
    int32_t pos(int u, int v) const
    {
        auto b = nb.begin() + off[u], e = nb.begin() + off[u + 1];
        return (int32_t)(std::lower_bound(b, e, v) - b);
    }
    // i-th element of adj(u) without the element at position skip
    int32_t at(int u, int skip, int i) const { return nb[off[u] + i + (i >= skip ? 1 : 0)]; }
};

bool start_side(EdgeWalk &w, const Frozen &F, int d, int side, int32_t *c)
{
    for (; side < 2; ++side) {
        const int u = side == 0 ? w.x : w.y, other = side == 0 ? w.y : w.x;
        const int a = F.deg(u) - 1;  // the other endpoint is a neighbour of u
        if (a >= d) {
            w.side = side;
            w.u = u;
            w.a = a;
            w.skip = F.pos(u, other);
            w.rank = 0;
            for (int i = 0; i < d; ++i) c[i] = i;
            return true;
        }
    }
    w.side = 2;
    return false;
}

// ChoiceGenerator::Next (ChoiceGenerator.cpp:55-85)
bool next_choice(int a, int b, int32_t *c)
{
    for (int i = b - 1; i >= 0; --i) {
        if (c[i] < i + (a - b)) {
            c[i]++;
            for (int j = i + 1; j < b; ++j) c[j] = c[j - 1] + 1;
            return true;
        }
    }
    return false;
}

// Runs one round of tests.  Plain form: one launch per test with a host
// synchronisation after each (the baseline), for at most the level's remaining
// plain_sample budget; the rest of the round (and the batched form) goes through
// batched launches.
fpgm_status run_tests(fpgm_ctx *ctx, bool plain, const fpgm_pc_opts &o, int level, const std::vector<HostTest> &tv,
                      const int32_t *zpool, int64_t zlen, double alpha, float band, std::vector<DevResult> &res,
                      fpgm_stats &S)
{
    const int64_t n = (int64_t)tv.size();
    res.resize((size_t)n);
    int64_t k = 0;
    const double ms0 = ctx->k_ms;
    if (plain) {
        const int64_t left = o.plain_sample > 0 ? o.plain_sample - S.plain_tests_per_level[level] : n;
        const int64_t m = std::min<int64_t>(n, std::max<int64_t>(0, left));
        std::vector<DevResult> one;
        const double t0 = now_s();
        for (; k < m; ++k) {
            fpgm_status st = run_batch(ctx, &tv[k], 1, zpool, zlen, alpha, o.flags | FPGM_F_EXACT, band, kTFExact, one);
            if (st != FPGM_OK) return st;
            res[k] = one[0];
        }
        S.plain_tests_per_level[level] += m;
        S.plain_seconds_per_level[level] += now_s() - t0;
    }
    if (k < n) {
        std::vector<DevResult> rest;
        fpgm_status st = run_batch(ctx, tv.data() + k, n - k, zpool, zlen, alpha, o.flags, band, 0, rest);
        if (st != FPGM_OK) return st;
        std::copy(rest.begin(), rest.end(), res.begin() + k);
    }
    S.kernel_seconds_per_level[level] += (ctx->k_ms - ms0) * 1e-3;
    return FPGM_OK;
}

void log_level(const fpgm_pc_opts &o, int d, int64_t tests, int64_t edges, double t)
{
    if (o.verbose > 0)
        fprintf(stderr, "[fpgm] level %d: %lld CI tests, %lld edges remain, %.4f s\n", d, (long long)tests,
                (long long)edges, t);
}

}  // namespace

extern "C" {

const char *fpgm_status_str(fpgm_status s)
{
    switch (s) {
    case FPGM_OK: return "ok";
    case FPGM_EINVAL: return "invalid argument";
    case FPGM_ENOMEM: return "out of memory";
    case FPGM_EDEVICE: return "device error";
    case FPGM_ETOOBIG: return "table too large";
    case FPGM_ETIMEOUT: return "round limit exceeded";
    default: return "unknown status";
    }
}

const char *fpgm_last_error(const fpgm_ctx *ctx)
{
    return ctx ? ctx->err.c_str() : "null context";
}

double fpgm_pchisq(double x, int32_t df)
{
    return pchisq(x, df);
}

double fpgm_chisq_crit(int32_t df, double alpha)
{
    return chisq_crit(df, alpha);
}

static fpgm_status ctx_init(fpgm_ctx *ctx, int32_t n_vars, const int32_t *card)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(ctx, FPGM_EDEVICE, std::string("device is not gfx950: ") + prop.gcnArchName);
    ctx->dev = dev;
    ctx->n_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    ctx->host_threads = std::max(1, std::min(16, omp_get_max_threads()));
    if (const char *e = std::getenv("FPGM_ROUND_BASE")) ctx->sched.base = std::max(1.0, std::atof(e));
    if (const char *e = std::getenv("FPGM_ROUND_CAP")) ctx->sched.cap = std::max<int64_t>(1, std::atoll(e));
    if (const char *e = std::getenv("FPGM_ROUND_PROMOTE")) ctx->sched.promote = std::max<int64_t>(0, std::atoll(e));
    if (const char *e = std::getenv("FPGM_MAX_SLICES")) ctx->max_slices = std::max<int64_t>(1, std::atoll(e));
    if (const char *e = std::getenv("FPGM_PIPELINE")) ctx->sched.pipeline = std::atoi(e) != 0;
    if (const char *e = std::getenv("FPGM_CONCURRENT")) ctx->sched.concurrent = std::atoi(e) != 0;
    if (const char *e = std::getenv("FPGM_ROUND_BYTES")) ctx->sched.round_bytes = std::max(1e6, std::atof(e));
    if (const char *e = std::getenv("FPGM_ITEMS_PER_CU"))
        ctx->sched.items_per_cu = std::max<int64_t>(1, std::atoll(e));
    if (const char *e = std::getenv("FPGM_WPC")) {  // five comma-separated values, in sink-group order
        int64_t v[kTuneGroups];
        if (std::sscanf(e, "%ld,%ld,%ld,%ld,%ld", &v[0], &v[1], &v[2], &v[3], &v[4]) == kTuneGroups)
            for (int g = 0; g < kTuneGroups; ++g) ctx->sched.wpc[g] = std::min<int64_t>(64, std::max<int64_t>(1, v[g]));
    }
    if (const char *e = std::getenv("FPGM_PACK")) ctx->pack = std::atoi(e) != 0 ? 1 : 0;
    if (const char *e = std::getenv("FPGM_GROUP")) ctx->group = std::atoi(e) != 0 ? 1 : 0;
    if (const char *e = std::getenv("FPGM_GRAPH")) ctx->graph = std::atoi(e) != 0 ? 1 : 0;
    if (const char *e = std::getenv("FPGM_DEVROUNDS")) ctx->devrounds = std::atoi(e) != 0 ? 1 : 0;
    if (const char *e = std::getenv("FPGM_PART")) ctx->part = std::min(2, std::max(0, std::atoi(e)));
    if (const char *e = std::getenv("FPGM_PART_MIN_ROWS")) ctx->part_min_rows = std::max<int64_t>(0, std::atoll(e));
    if (const char *e = std::getenv("FPGM_BITS")) ctx->bits = std::atoi(e) != 0 ? 1 : 0;
    if (const char *e = std::getenv("FPGM_BITS_CELLS")) ctx->bits_cells = std::min(kLaneCells, std::max(0, std::atoi(e)));
    if (const char *e = std::getenv("FPGM_UPLOAD_PIPELINE")) ctx->upload_pipeline = std::atoi(e) != 0 ? 1 : 0;
    if (const char *e = std::getenv("FPGM_GROUP_MIN_ROWS")) ctx->group_min_rows = std::max<int64_t>(0, std::atoll(e));
    if (const char *e = std::getenv("FPGM_LDS_OWN_CHUNKS")) ctx->sched.lds_own_chunks = std::atoi(e) != 0 ? 1 : 0;
    if (const char *e = std::getenv("FPGM_SIDE_CU_MULT")) ctx->sched.side_cu_mult = std::max<int64_t>(1, std::atoll(e));
    if (const char *e = std::getenv("FPGM_SLICE_MIN_ROWS")) ctx->sched.slice_min_rows = std::max<int64_t>(0, std::atoll(e));
    if (const char *e = std::getenv("FPGM_WIDE_MAX_ROWS")) ctx->wide_max_rows = std::max<int64_t>(0, std::atoll(e));
    if (const char *e = std::getenv("FPGM_SINKS")) {
        if (std::strcmp(e, "shared") == 0) ctx->sink_mask |= 3;
        if (std::strcmp(e, "nowide") == 0) ctx->sink_mask |= 1;
        if (std::strcmp(e, "nolane") == 0) ctx->sink_mask |= 5;
        if (std::strcmp(e, "nocol") == 0) ctx->sink_mask |= 8;
    }
    HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
    for (hipStream_t &sd : ctx->side) HIP_TRY(hipStreamCreateWithFlags(&sd, hipStreamNonBlocking), "hipStreamCreate");
    HIP_TRY(hipEventCreate(&ctx->ev0), "hipEventCreate");
    HIP_TRY(hipEventCreate(&ctx->ev1), "hipEventCreate");
    for (Slot &sl : ctx->slot) HIP_TRY(sl.init(), "hipEventCreate");
    ctx->n_vars = n_vars;
    ctx->card.assign(card, card + n_vars);
    HIP_TRY(hipMalloc((void **)&ctx->d_card, sizeof(int32_t) * std::max(1, n_vars)), "hipMalloc card");
    HIP_TRY(hipMemcpyAsync(ctx->d_card, card, sizeof(int32_t) * n_vars, hipMemcpyHostToDevice, ctx->stream),
            "upload card");
    return FPGM_OK;
}

static fpgm_status check_card(int32_t n_vars, const int32_t *card)
{
    if (n_vars < 2 || !card) return FPGM_EINVAL;
    for (int v = 0; v < n_vars; ++v)
        if (card[v] < 1 || card[v] > 255) return FPGM_EINVAL;
    return FPGM_OK;
}

static bool rows_ok(int64_t n_rows, int64_t row_stride)
{
    return n_rows >= 0 && row_stride >= n_rows && n_rows <= ((int64_t)1 << 31) - 1;
}

// Packing and value scan of columns [v0, v1) of instance I on the context stream.
static fpgm_status pack_and_scan(fpgm_ctx *ctx, const InstHost &I, uint8_t *pcols, int v0, int v1)
{
    const int64_t o = (int64_t)v0 * I.stride;
    if (pcols)
        HIP_TRY(launch_pack_cols(I.cols + o, pcols + o, I.stride, v1 - v0, ctx->d_colw + v0, ctx->stream),
                "launch pack");
    HIP_TRY(launch_check_values(I.cols + o, I.stride, I.n_rows, v1 - v0, ctx->d_card + v0, ctx->d_flag.p, ctx->stream),
            "launch check");
    return FPGM_OK;
}

// Upload of host columns (instance i at cols[i], row_stride[i] bytes apart) into the
// context's device columns, pipelined over chunks of consecutive columns in device order
// (the instances lie back to back): the host copies chunk k+1 into one of two pinned
// staging buffers -- each column's rows, zeros up to the device stride -- on several
// threads while chunk k's copy runs on a side stream, and the context stream packs and
// scans each chunk once its copy is done.  pcols: the packed copy's instance bases.
static fpgm_status upload_and_scan(fpgm_ctx *ctx, const uint8_t *const *cols, const int64_t *row_stride,
                                   const std::vector<uint8_t *> &pcols)
{
    const int n = ctx->n_vars;
    const int64_t n_inst = (int64_t)ctx->inst.size();
    int64_t widest = 0;
    for (const InstHost &I : ctx->inst) widest = std::max(widest, I.stride);
    const int64_t cap = std::max<int64_t>((int64_t)32 << 20, widest);
    HostBuf<uint8_t> stage[2];
    hipEvent_t copied[2] = {nullptr, nullptr};
    struct Cleanup {
        fpgm_ctx *c;
        HostBuf<uint8_t> *st;
        hipEvent_t *ev;
        ~Cleanup()
        {
            (void)hipStreamSynchronize(c->side[0]);
            (void)hipStreamSynchronize(c->stream);
            for (int b = 0; b < 2; ++b) {
                st[b].release();
                if (ev[b]) (void)hipEventDestroy(ev[b]);
            }
        }
    } cleanup{ctx, stage, copied};
    for (int b = 0; b < 2; ++b) {
        HIP_TRY(stage[b].ensure(cap, ctx->stream), "alloc pinned staging");
        HIP_TRY(hipEventCreateWithFlags(&copied[b], hipEventDisableTiming), "event");
    }
    struct Col {
        int32_t inst, v;
    };
    std::vector<Col> chunk;
    int64_t i = 0, v = 0;
    for (int64_t k = 0; i < n_inst; ++k) {
        HostBuf<uint8_t> &S = stage[k & 1];
        if (k >= 2) HIP_TRY(hipEventSynchronize(copied[k & 1]), "synchronize");  // its previous copy is done
        // the chunk: whole columns in device order while they fit the staging buffer
        chunk.clear();
        int64_t bytes = 0;
        uint8_t *dst = const_cast<uint8_t *>(ctx->inst[i].cols) + v * ctx->inst[i].stride;
        while (i < n_inst && bytes + ctx->inst[i].stride <= cap) {
            chunk.push_back(Col{(int32_t)i, (int32_t)v});
            bytes += ctx->inst[i].stride;
            if (++v == n) {
                v = 0;
                ++i;
            }
        }
        const int nth = (bytes >= ((int64_t)4 << 20)) ? ctx->host_threads : 1;
#pragma omp parallel for if (nth > 1) num_threads(nth) schedule(static)
        for (int64_t c = 0; c < (int64_t)chunk.size(); ++c) {
            const InstHost &I = ctx->inst[chunk[c].inst];
            // the chunk's columns are consecutive in device memory
            uint8_t *o = S.p + ((I.cols + (int64_t)chunk[c].v * I.stride) - dst);
            std::memcpy(o, cols[chunk[c].inst] + (int64_t)chunk[c].v * row_stride[chunk[c].inst], (size_t)I.n_rows);
            std::memset(o + I.n_rows, 0, (size_t)(I.stride - I.n_rows));
        }
        HIP_TRY(hipMemcpyAsync(dst, S.p, (size_t)bytes, hipMemcpyHostToDevice, ctx->side[0]), "upload columns");
        HIP_TRY(hipEventRecord(copied[k & 1], ctx->side[0]), "event");
        HIP_TRY(hipStreamWaitEvent(ctx->stream, copied[k & 1], 0), "stream wait");
        // packing and value scan, one launch pair per instance segment of the chunk
        for (size_t c0 = 0; c0 < chunk.size();) {
            size_t c1 = c0;
            while (c1 < chunk.size() && chunk[c1].inst == chunk[c0].inst) ++c1;
            const int32_t ii = chunk[c0].inst;
            fpgm_status st = pack_and_scan(ctx, ctx->inst[ii], pcols[ii], chunk[c0].v, chunk[c1 - 1].v + 1);
            if (st != FPGM_OK) return st;
            c0 = c1;
        }
    }
    return FPGM_OK;
}

// Chooses the packed widths, uploads host columns (cols != nullptr; else the instances'
// columns are already on the device), packs them for the counting kernels, uploads the
// instance table and scans every column for values >= card.
static fpgm_status finish_ctx(fpgm_ctx *ctx, const uint8_t *const *cols = nullptr, const int64_t *row_stride = nullptr)
{
    const int n = ctx->n_vars;
    ctx->colw.assign((size_t)n, 8);
    bool any_packed = false;
    // measured: at 10^6 rows (1000 variables) packing takes the stage from 250 to 208 ms;
    // with 2 000 - 50 000 rows per instance (bootstrap replicates, whose columns stay in
    // L2) the unpacking costs more than the bytes it saves (+5 %)
    double mean_rows = 0.0;
    for (const InstHost &I : ctx->inst) mean_rows += (double)I.n_rows / (double)ctx->inst.size();
    const bool pack = ctx->pack > 0 || (ctx->pack < 0 && mean_rows >= (double)kPackMinRows);
    if (pack)
        for (int v = 0; v < n; ++v) {
            const int c = ctx->card[v];
            ctx->colw[v] = (uint8_t)(c <= 4 ? 2 : c <= 16 ? 4 : 8);
            any_packed |= ctx->colw[v] != 8;
        }
    HIP_TRY(hipMalloc((void **)&ctx->d_colw, (size_t)std::max(1, n)), "hipMalloc widths");
    HIP_TRY(hipMemcpyAsync(ctx->d_colw, ctx->colw.data(), (size_t)n, hipMemcpyHostToDevice, ctx->stream),
            "upload widths");
    std::vector<uint8_t *> pcols(ctx->inst.size(), nullptr);
    if (any_packed) {
        // one allocation, each instance at the offset of its uint8 columns (same stride)
        int64_t total = 0;
        std::vector<int64_t> off(ctx->inst.size());
        for (size_t i = 0; i < off.size(); ++i) {
            off[i] = total;
            total += ctx->inst[i].stride * n;
        }
        HIP_TRY(hipMalloc((void **)&ctx->d_pcols, (size_t)std::max<int64_t>(total, 16)), "hipMalloc packed columns");
        for (size_t i = 0; i < off.size(); ++i) pcols[i] = ctx->d_pcols + off[i];
    }
    HIP_TRY(ctx->d_flag.ensure(1, true, ctx->stream), "alloc flag");
    HIP_TRY(hipMemsetAsync(ctx->d_flag.p, 0, sizeof(uint32_t), ctx->stream), "memset flag");
    if (cols && ctx->upload_pipeline) {
        fpgm_status st = upload_and_scan(ctx, cols, row_stride, pcols);
        if (st != FPGM_OK) return st;
    } else {
        for (size_t i = 0; cols && i < ctx->inst.size(); ++i) {
            const InstHost &I = ctx->inst[i];
            HIP_TRY(hipMemsetAsync(const_cast<uint8_t *>(I.cols), 0, (size_t)(I.stride * n), ctx->stream), "clear columns");
            if (I.n_rows > 0)
                HIP_TRY(hipMemcpy2DAsync(const_cast<uint8_t *>(I.cols), (size_t)I.stride, cols[i], (size_t)row_stride[i],
                                         (size_t)I.n_rows, (size_t)n, hipMemcpyHostToDevice, ctx->stream),
                        "upload columns");
        }
        for (size_t i = 0; i < ctx->inst.size(); ++i) {
            fpgm_status st = pack_and_scan(ctx, ctx->inst[i], pcols[i], 0, n);
            if (st != FPGM_OK) return st;
        }
    }
    // bit planes with the packed copy (large instances): every variable of at most
    // kBitsMaxCard states, nbits_of(card) planes each, each instance's planes kPlaneAlign-
    // word aligned and clear past its rows.  Without the memory the tables count as they
    // would without planes.
    std::vector<int32_t> poff((size_t)n, -1);
    int32_t n_planes = 0;
    for (int v = 0; v < n; ++v)
        if (ctx->card[v] <= kBitsMaxCard) {
            poff[(size_t)v] = n_planes;
            n_planes += nbits_of(ctx->card[v]);
        }
    std::vector<int64_t> pwords(ctx->inst.size(), 0), pbase(ctx->inst.size(), 0);
    if (pack && ctx->bits && ctx->bits_cells > 0 && n_planes > 0) {
        int64_t total = 0;
        for (size_t i = 0; i < ctx->inst.size(); ++i) {
            pwords[i] = ((ctx->inst[i].n_rows + 31) / 32 + kPlaneAlign - 1) / kPlaneAlign * kPlaneAlign;
            pbase[i] = total;
            total += pwords[i] * n_planes;
        }
        if (hipMalloc((void **)&ctx->d_planes, sizeof(uint32_t) * (size_t)total) != hipSuccess) {
            (void)hipGetLastError();
            ctx->d_planes = nullptr;
        } else {
            ctx->planes_bytes = (int64_t)sizeof(uint32_t) * total;
            HIP_TRY(hipMalloc((void **)&ctx->d_poff, sizeof(int32_t) * (size_t)n), "hipMalloc plane offsets");
            HIP_TRY(hipMemcpyAsync(ctx->d_poff, poff.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice,
                                   ctx->stream),
                    "upload plane offsets");
            for (size_t i = 0; i < ctx->inst.size(); ++i)
                HIP_TRY(launch_build_planes(ctx->inst[i].cols, ctx->inst[i].stride, ctx->inst[i].n_rows, n, ctx->d_card,
                                            ctx->d_poff, ctx->d_planes + pbase[i], pwords[i], ctx->stream),
                        "launch build planes");
        }
    }
    std::vector<DevInst> di(ctx->inst.size());
    for (size_t i = 0; i < di.size(); ++i)
        di[i] = DevInst{ctx->inst[i].cols, ctx->inst[i].stride, ctx->inst[i].n_rows,
                        pcols[i] ? pcols[i] : ctx->inst[i].cols, ctx->d_planes ? ctx->d_planes + pbase[i] : nullptr,
                        pwords[i]};
    HIP_TRY(hipMalloc((void **)&ctx->d_inst, sizeof(DevInst) * di.size()), "hipMalloc instances");
    HIP_TRY(hipMemcpyAsync(ctx->d_inst, di.data(), sizeof(DevInst) * di.size(), hipMemcpyHostToDevice, ctx->stream),
            "upload instances");
    uint32_t bad = 0;
    HIP_TRY(hipMemcpyAsync(&bad, ctx->d_flag.p, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream), "flag");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "synchronize");
    if (bad) return fail(ctx, FPGM_EINVAL, "a column holds a value >= its cardinality");
    return FPGM_OK;
}

fpgm_status fpgm_ctx_create_multi(const uint8_t *const *cols, const int64_t *n_rows, const int64_t *row_stride,
                                  int32_t n_inst, int32_t n_vars, const int32_t *card, fpgm_ctx **out)
{
    if (!out) return FPGM_EINVAL;
    *out = nullptr;
    if (!cols || !n_rows || !row_stride || n_inst < 1 || check_card(n_vars, card) != FPGM_OK) return FPGM_EINVAL;
    for (int32_t i = 0; i < n_inst; ++i)
        if (!cols[i] || !rows_ok(n_rows[i], row_stride[i])) return FPGM_EINVAL;
    fpgm_ctx *ctx = new (std::nothrow) fpgm_ctx();
    if (!ctx) return FPGM_ENOMEM;
    fpgm_status st = ctx_init(ctx, n_vars, card);
    if (st == FPGM_OK) {
        // one allocation holding every instance, each with a 256-byte-aligned stride
        std::vector<int64_t> off(n_inst + 1, 0);
        for (int32_t i = 0; i < n_inst; ++i) {
            const int64_t stride = std::max<int64_t>(256, (n_rows[i] + 255) / 256 * 256);
            off[i + 1] = off[i] + stride * n_vars;
        }
        uint8_t *d = nullptr;
        hipError_t e = hipMalloc((void **)&d, (size_t)off[n_inst]);
        if (e != hipSuccess) {
            st = hip_fail(ctx, e, "hipMalloc columns");
        } else {
            ctx->d_cols = d;
            ctx->owns_cols = true;
            for (int32_t i = 0; i < n_inst; ++i)
                ctx->inst.push_back(InstHost{d + off[i], (off[i + 1] - off[i]) / n_vars, n_rows[i]});
        }
    }
    if (st == FPGM_OK) st = finish_ctx(ctx, cols, row_stride);  // uploads, packs and scans the columns
    if (st != FPGM_OK) {
        fpgm_ctx_destroy(ctx);
        return st;
    }
    *out = ctx;
    return FPGM_OK;
}

fpgm_status fpgm_ctx_create(const uint8_t *cols, int32_t n_vars, int64_t n_rows, int64_t row_stride,
                            const int32_t *card, fpgm_ctx **out)
{
    return fpgm_ctx_create_multi(&cols, &n_rows, &row_stride, 1, n_vars, card, out);
}

fpgm_status fpgm_ctx_create_device(const void *d_cols, int32_t n_vars, int64_t n_rows, int64_t row_stride,
                                   const int32_t *card, fpgm_ctx **out)
{
    if (!out) return FPGM_EINVAL;
    *out = nullptr;
    if (!d_cols || ((uintptr_t)d_cols % 16) != 0 || row_stride % 16 != 0 || !rows_ok(n_rows, row_stride) ||
        check_card(n_vars, card) != FPGM_OK)
        return FPGM_EINVAL;
    fpgm_ctx *ctx = new (std::nothrow) fpgm_ctx();
    if (!ctx) return FPGM_ENOMEM;
    fpgm_status st = ctx_init(ctx, n_vars, card);
    if (st == FPGM_OK) {
        ctx->d_cols = (const uint8_t *)d_cols;
        ctx->owns_cols = false;
        ctx->inst.push_back(InstHost{(const uint8_t *)d_cols, row_stride, n_rows});
        st = finish_ctx(ctx);
    }
    if (st != FPGM_OK) {
        fpgm_ctx_destroy(ctx);
        return st;
    }
    *out = ctx;
    return FPGM_OK;
}

void fpgm_ctx_destroy(fpgm_ctx *ctx)
{
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->owns_cols && ctx->d_cols) (void)hipFree((void *)ctx->d_cols);
    if (ctx->d_inst) (void)hipFree(ctx->d_inst);
    if (ctx->d_card) (void)hipFree(ctx->d_card);
    if (ctx->d_colw) (void)hipFree(ctx->d_colw);
    if (ctx->d_pcols) (void)hipFree(ctx->d_pcols);
    if (ctx->d_planes) (void)hipFree(ctx->d_planes);
    if (ctx->d_poff) (void)hipFree(ctx->d_poff);
    if (ctx->d_crit) (void)hipFree(ctx->d_crit);
    for (Slot &sl : ctx->slot) sl.release();
    ctx->d_ws.release();
    ctx->d_part.release();
    ctx->d_done.release();
    ctx->d_flag.release();
    ctx->d_C.release();
    ctx->d_rowvar.release();
    ctx->d_rowstate.release();
    ctx->d_tiles.release();
    ctx->d_off.release();
    ctx->d_l0_flags.release();
    ctx->d_l0_sum.release();
    ctx->h_l0_flags.release();
    ctx->h_l0_sum.release();
    ctx->d_mle_nodes.release();
    ctx->d_mle_rel.release();
    ctx->d_mle_items.release();
    ctx->d_mle_out.release();
    ctx->d_edge_done.release();
    ctx->d_nbr.release();
    ctx->d_binom.release();
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    for (hipStream_t sd : ctx->side)
        if (sd) {
            (void)hipStreamSynchronize(sd);
            (void)hipStreamDestroy(sd);
        }
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

static fpgm_status check_test(fpgm_ctx *ctx, int32_t x, int32_t y, const int32_t *z, int32_t d, double alpha)
{
    if (ctx->inst.size() != 1) return fail(ctx, FPGM_EINVAL, "single-test calls need a one-instance context");
    if (!(alpha >= 0.0 && alpha <= 1.0)) return fail(ctx, FPGM_EINVAL, "alpha must lie in [0, 1]");
    if (x < 0 || x >= ctx->n_vars || y < 0 || y >= ctx->n_vars || x == y)
        return fail(ctx, FPGM_EINVAL, "x, y must be distinct node indices");
    if (d < 0 || d > kMaxD || (d > 0 && !z)) return fail(ctx, FPGM_EINVAL, "bad conditioning set size");
    for (int j = 0; j < d; ++j)
        if (z[j] < 0 || z[j] >= ctx->n_vars) return fail(ctx, FPGM_EINVAL, "bad conditioning node index");
    return FPGM_OK;
}

fpgm_status fpgm_ci_test(fpgm_ctx *ctx, int32_t x, int32_t y, const int32_t *z, int32_t d, double alpha,
                         fpgm_result *out)
{
    if (!ctx || !out) return FPGM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    fpgm_status st = check_test(ctx, x, y, z, d, alpha);
    if (st != FPGM_OK) return st;
    HostTest h{x, y, d, 0, 0};
    std::vector<DevResult> r;
    st = run_batch(ctx, &h, 1, z, d, alpha, FPGM_F_EXACT, 0.02f, kTFExact, r);
    if (st != FPGM_OK) return st;
    *out = to_result(r[0]);
    return FPGM_OK;
}

fpgm_status fpgm_ci_batch(fpgm_ctx *ctx, const fpgm_test *tests, int64_t n, const int32_t *z_pool,
                          int64_t z_pool_len, double alpha, uint32_t flags, fpgm_result *out)
{
    if (!ctx || (n > 0 && (!tests || !out)) || n < 0) return FPGM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    std::vector<HostTest> h((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        const fpgm_test &t = tests[i];
        if (t.d > 0 && (t.z_off < 0 || (int64_t)t.z_off + t.d > z_pool_len || !z_pool))
            return fail(ctx, FPGM_EINVAL, "z_off/d outside the z pool");
        fpgm_status st = check_test(ctx, t.x, t.y, t.d > 0 ? z_pool + t.z_off : nullptr, t.d, alpha);
        if (st != FPGM_OK) return st;
        h[i] = HostTest{t.x, t.y, t.d, 0, t.z_off};
    }
    reset_kernel_stats(ctx);
    std::vector<DevResult> r;
    fpgm_status st = run_batch(ctx, h.data(), n, z_pool, z_pool_len, alpha, flags, 0.02f, 0, r);
    if (st != FPGM_OK) return st;
    for (int64_t i = 0; i < n; ++i) out[i] = to_result(r[i]);
    return FPGM_OK;
}

fpgm_status fpgm_ci_group(fpgm_ctx *ctx, int32_t x, int32_t y, const int32_t *z, int32_t d, int32_t c_size,
                          double alpha, fpgm_group_result *res, fpgm_result *out)
{
    if (!ctx || !res || c_size < 1) return FPGM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    // IndependenceResult's dispatch (IndependenceTest.cpp:45-52): no Z -> one XY test, whatever
    // c_size; c_size 1 -> one XYZ test; else the group of c_size sets of d nodes each
    const int64_t n = (d == 0) ? 1 : c_size;
    if (d > 0 && (int64_t)d * n > INT32_MAX) return fail(ctx, FPGM_EINVAL, "group too large");
    std::vector<HostTest> h((size_t)n);
    for (int64_t m = 0; m < n; ++m) {
        fpgm_status st = check_test(ctx, x, y, d > 0 ? z + m * d : nullptr, d, alpha);
        if (st != FPGM_OK) return st;
        h[m] = HostTest{x, y, d, 0, m * d};
    }
    reset_kernel_stats(ctx);
    std::vector<DevResult> r;
    fpgm_status st = run_batch(ctx, h.data(), n, z, d * n, alpha, FPGM_F_EXACT, 0.02f, kTFExact, r);
    if (st != FPGM_OK) return st;
    if (out)
        for (int64_t m = 0; m < n; ++m) out[m] = to_result(r[m]);
    const fpgm_result r0 = to_result(r[0]);
    if (d == 0) {  // ComputeGSquareXY's Result(p, indep): first keeps its default -1 (:349-363)
        *res = fpgm_group_result{r0.p_value, r0.indep, {0, 0, 0}, -1};
    } else if (n == 1) {  // ComputeGSquareXYZ's Result(p, indep, 0) (:140-154)
        *res = fpgm_group_result{r0.p_value, r0.indep, {0, 0, 0}, 0};
    } else {  // ComputeGSquareXYZGroup: the first independent set of the group (:280-286)
        *res = fpgm_group_result{0.0, 0, {0, 0, 0}, -1};
        for (int64_t m = 0; m < n; ++m)
            if (r[m].indep) {
                *res = fpgm_group_result{-1.0, 1, {0, 0, 0}, (int32_t)m};
                break;
            }
    }
    return FPGM_OK;
}

fpgm_status fpgm_ci_counts(fpgm_ctx *ctx, int32_t x, int32_t y, const int32_t *z, int32_t d, int32_t *table,
                           int64_t cap, int64_t *cells)
{
    if (!ctx || !cells) return FPGM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    fpgm_status st = check_test(ctx, x, y, z, d, 0.05);
    if (st != FPGM_OK) return st;
    int64_t nc = (int64_t)ctx->card[x] * ctx->card[y];
    for (int j = 0; j < d; ++j) nc *= ctx->card[z[j]];
    *cells = nc;
    if (!table || cap < nc) return fail(ctx, FPGM_EINVAL, "table capacity too small");
    HostTest h{x, y, d, 0, 0};
    std::vector<DevResult> r;
    int64_t off = -1;
    st = run_batch(ctx, &h, 1, z, d, 0.05, 0, 0.02f, kTFKeep, r, 0, &off);
    if (st != FPGM_OK) return st;
    HIP_TRY(hipMemcpyAsync(table, ctx->d_ws.p + off, sizeof(int32_t) * nc, hipMemcpyDeviceToHost, ctx->stream),
            "download table");
    int64_t dimz = 1;
    for (int j = 0; j < d; ++j) dimz *= ctx->card[z[j]];
    const int64_t used = nc + dimz * (ctx->card[x] + ctx->card[y] + 1);
    HIP_TRY(hipMemsetAsync(ctx->d_ws.p + off, 0, sizeof(int32_t) * used, ctx->stream), "clear table");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "synchronize");
    return FPGM_OK;
}

fpgm_status fpgm_mle_counts(fpgm_ctx *ctx, int32_t n_nodes, const int32_t *par_off, const int32_t *parents,
                            int64_t *count_off, int32_t *counts, int64_t cap)
{
    if (!ctx || !par_off || !count_off) return FPGM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->inst.size() != 1) return fail(ctx, FPGM_EINVAL, "maximum-likelihood counts need a one-instance context");
    if (n_nodes < 1 || n_nodes > ctx->n_vars) return fail(ctx, FPGM_EINVAL, "n_nodes must lie in [1, n_vars]");
    if (par_off[0] != 0) return fail(ctx, FPGM_EINVAL, "par_off[0] must be 0");
    std::vector<MleNode> nodes((size_t)n_nodes);
    std::vector<int32_t> rel;
    int64_t total = 0;
    for (int32_t i = 0; i < n_nodes; ++i) {
        const int32_t k = par_off[i + 1] - par_off[i];
        if (k < 0 || k + 1 > kMleMaxRel || (k > 0 && !parents))
            return fail(ctx, FPGM_EINVAL, "bad parent list length");
        MleNode &N = nodes[i];
        N.n_rel = k + 1;
        N.rel_off = (int32_t)rel.size();
        rel.push_back(i);
        int64_t size = ctx->card[i];
        for (int32_t j = 0; j < k; ++j) {
            const int32_t p = parents[par_off[i] + j];
            if (p < 0 || p >= ctx->n_vars || p == i) return fail(ctx, FPGM_EINVAL, "bad parent index");
            if (j > 0 && p <= parents[par_off[i] + j - 1])
                return fail(ctx, FPGM_EINVAL, "parents must be in ascending order");
            rel.push_back(p);
            size *= ctx->card[p];
            if (size >= kMaxCells) return fail(ctx, FPGM_ETOOBIG, "probability table exceeds 2^31 cells");
        }
        N.size = (int32_t)size;
        N.out_off = total;
        count_off[i] = total;
        total += size;
    }
    count_off[n_nodes] = total;
    if (!counts) return FPGM_OK;  // size query
    if (cap < total) return fail(ctx, FPGM_EINVAL, "counts capacity too small");
    const InstHost &I = ctx->inst[0];
    const int64_t steps = std::max<int64_t>(1, (I.n_rows + kRowsPerStep - 1) / kRowsPerStep);
    // rows per chunk: enough work items for every CU, at least 16 steps per item
    const int64_t cpn = std::max<int64_t>(1, std::min<int64_t>(steps / 16, ((int64_t)ctx->n_cu * 8 + n_nodes - 1) / n_nodes));
    const int64_t rpc = (steps + cpn - 1) / cpn * kRowsPerStep;
    const int64_t n_chunks = std::max<int64_t>(1, (I.n_rows + rpc - 1) / rpc);
    std::vector<int32_t> items;
    items.reserve((size_t)(n_chunks * n_nodes));
    for (int32_t i = 0; i < n_nodes; ++i) {
        nodes[i].item0 = (int32_t)items.size();
        nodes[i].n_chunks = (int32_t)n_chunks;
        for (int64_t c = 0; c < n_chunks; ++c) items.push_back(i);
    }
    HIP_TRY(ctx->d_mle_nodes.ensure(n_nodes, false, ctx->stream), "alloc mle nodes");
    HIP_TRY(ctx->d_mle_rel.ensure((int64_t)rel.size(), false, ctx->stream), "alloc mle variables");
    HIP_TRY(ctx->d_mle_items.ensure((int64_t)items.size(), false, ctx->stream), "alloc mle items");
    HIP_TRY(ctx->d_mle_out.ensure(std::max<int64_t>(total, 1), false, ctx->stream), "alloc mle tables");
    HIP_TRY(hipMemcpyAsync(ctx->d_mle_nodes.p, nodes.data(), sizeof(MleNode) * n_nodes, hipMemcpyHostToDevice,
                           ctx->stream),
            "upload mle nodes");
    HIP_TRY(hipMemcpyAsync(ctx->d_mle_rel.p, rel.data(), sizeof(int32_t) * rel.size(), hipMemcpyHostToDevice,
                           ctx->stream),
            "upload mle variables");
    HIP_TRY(hipMemcpyAsync(ctx->d_mle_items.p, items.data(), sizeof(int32_t) * items.size(), hipMemcpyHostToDevice,
                           ctx->stream),
            "upload mle items");
    HIP_TRY(hipMemsetAsync(ctx->d_mle_out.p, 0, sizeof(int32_t) * total, ctx->stream), "clear mle tables");
    MleArgs a{};
    a.cols = I.cols;
    a.stride = I.stride;
    a.n_rows = I.n_rows;
    a.nodes = ctx->d_mle_nodes.p;
    a.rel = ctx->d_mle_rel.p;
    a.card = ctx->d_card;
    a.item_node = ctx->d_mle_items.p;
    a.n_items = (int64_t)items.size();
    a.rows_per_chunk = rpc;
    a.out = ctx->d_mle_out.p;
    const int grid = (int)std::min<int64_t>(a.n_items, (int64_t)ctx->n_cu * 4);
    HIP_TRY(launch_mle_counts(a, grid, ctx->stream), "launch mle counts");
    HIP_TRY(hipMemcpyAsync(counts, ctx->d_mle_out.p, sizeof(int32_t) * total, hipMemcpyDeviceToHost, ctx->stream),
            "download mle tables");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "synchronize");
    return FPGM_OK;
}

namespace {

// PC-stable state of one dataset instance across levels.
struct PcInst {
    std::vector<uint8_t> adj;       // n*n
    std::vector<int32_t> ex, ey;    // surviving edges, lexicographic
    struct Sep {
        int32_t x, y, level, d;
        int64_t zoff;
    };
    std::vector<Sep> seps;          // level 0's removals, then each later level's, each run ascending
    std::vector<int32_t> sepz;
    size_t n_sep0 = 0;              // level 0's removals (the first run of seps)
    int64_t num_ci_test = 0;
    bool cont = true;               // continues to the next level
    Frozen F;
    std::vector<EdgeWalk> walks;
    std::vector<int32_t> choice, zlevel;
};

// The subset of lexicographic rank r among the k-subsets of {0..a-1} (the order of
// ChoiceGenerator::Next, ChoiceGenerator.cpp:55-85); r < C(a, k).
void unrank_subset(const DevLevel &L, int a, int k, int64_t r, int32_t *c)
{
    int x = 0;
    for (int i = 0; i < k; ++i) {
        for (; x < a - k + i; ++x) {
            const int64_t cnt = L.C(a - 1 - x, k - 1 - i);
            if (r < cnt) break;
            r -= cnt;
        }
        c[i] = x++;
    }
}

// Whether level d runs on device-generated rounds, and what its rounds share.  Eligible:
// every table of the level, with its marginals, fits one workgroup's LDS (bounded per
// edge side by the largest d cardinalities of the side's neighbours), every side's subset
// count fits 31 bits, and the neighbour store 31-bit offsets.
bool dev_level_plan(const fpgm_ctx *ctx, int d, const std::vector<PcInst> &P,
                    const std::vector<std::pair<int32_t, int64_t>> &active, DevLevel &L)
{
    const int n = ctx->n_vars;
    L = DevLevel{};
    L.d = d;
    int64_t amax = 0;
    for (const auto &ie : active) {
        const PcInst &I = P[ie.first];
        const EdgeWalk &w = I.walks[ie.second];
        amax = std::max<int64_t>(amax, std::max(I.F.deg(w.x), I.F.deg(w.y)) - 1);
    }
    const int64_t kSat = INT64_MAX / 4;
    L.binom.assign((size_t)(amax + 1) * (size_t)(d + 1), 0);
    for (int64_t m = 0; m <= amax; ++m) {
        L.binom[(size_t)m * (d + 1)] = 1;
        for (int k = 1; k <= d && m > 0; ++k)
            L.binom[(size_t)m * (d + 1) + k] = std::min(kSat, L.C(m - 1, k - 1) + L.C(m - 1, k));
    }
    int64_t cells_max = 0, full_max = 0, rows_max = 0, rows_min = INT64_MAX;
    for (size_t i = 0; i < P.size(); ++i) {
        if (!P[i].cont) continue;
        rows_max = std::max(rows_max, ctx->inst[i].n_rows);
        rows_min = std::min(rows_min, ctx->inst[i].n_rows);
    }
    // per node, its neighbours' cardinalities in descending order
    std::vector<std::vector<int32_t>> desc;
    int32_t last_inst = -1;
    for (const auto &ie : active) {
        const PcInst &I = P[ie.first];
        if (ie.first != last_inst) {
            last_inst = ie.first;
            desc.assign((size_t)n, {});
            for (int u = 0; u < n; ++u) {
                auto &D = desc[(size_t)u];
                for (int32_t k = I.F.off[u]; k < I.F.off[u + 1]; ++k) D.push_back(ctx->card[I.F.nb[k]]);
                std::sort(D.begin(), D.end(), std::greater<int32_t>());
            }
        }
        const EdgeWalk &w = I.walks[ie.second];
        const int64_t cx = ctx->card[w.x], cy = ctx->card[w.y];
        for (int side = 0; side < 2; ++side) {
            const int u = side ? w.y : w.x, other = side ? w.x : w.y;
            const int64_t a = I.F.deg(u) - 1;
            if (a < d) continue;
            if (L.C(a, d) >= (int64_t)INT32_MAX) return false;
            L.walk_tests = std::min<int64_t>(L.walk_tests + L.C(a, d), (int64_t)1 << 62);
            // the largest product of d cardinalities among adj(u) \ {other}
            int64_t prod = 1;
            int taken = 0;
            bool removed = false;
            for (int32_t v : desc[(size_t)u]) {
                if (taken == d) break;
                if (!removed && v == ctx->card[other]) {
                    removed = true;
                    continue;
                }
                prod = std::min<int64_t>(prod * v, (int64_t)1 << 40);
                ++taken;
            }
            cells_max = std::max(cells_max, cx * cy * prod);
            full_max = std::max(full_max, (cx + 1) * (cy + 1) * prod);  // table + marginals
            if (full_max > kLdsCells) return false;
        }
    }
    L.rows_max = rows_max;
    L.need_max = (full_max + 63) / 64 * 64;
    L.w128_shared = cells_max > kLaneWide128Cells;
    L.lane_w128 = !L.w128_shared && cells_max > kLaneWideCells;
    if (ctx->sink_mask == 0) {
        // the lane-wide group takes 65..128 cells of instances below wide_max_rows; the
        // shared group 97..128 (w128_shared), larger tables, and 65..128 of long instances
        // outside the column layout
        L.launch[1] = cells_max > kLaneCells && rows_min < ctx->wide_max_rows;
        L.launch[2] = cells_max > kLaneWide128Cells || (cells_max > kLaneCells && rows_max >= ctx->wide_max_rows);
        L.launch[3] = d == 1 && route_params(ctx).bits_cells > 0;
    }
    L.nbr_base.assign(P.size(), -1);
    for (size_t i = 0; i < P.size(); ++i) {
        if (!P[i].cont) continue;
        L.nbr_base[i] = (int64_t)L.nbr.size();
        L.nbr.insert(L.nbr.end(), P[i].F.nb.begin(), P[i].F.nb.end());
    }
    return (int64_t)L.nbr.size() < (int64_t)INT32_MAX;
}

void apply_tune(fpgm_ctx *ctx, const fpgm_level_tune &t)
{
    if (t.items_per_cu > 0) ctx->sched.items_per_cu = t.items_per_cu;
    for (int g = 0; g < kTuneGroups; ++g)
        if (t.wg_per_cu[g] > 0) ctx->sched.wpc[g] = t.wg_per_cu[g];
    if (t.max_slices > 0) ctx->max_slices = t.max_slices;
    if (t.round_cap > 0) ctx->sched.cap = t.round_cap;
    if (t.pipeline >= 0) ctx->sched.pipeline = t.pipeline;
    if (t.concurrent >= 0) ctx->sched.concurrent = t.concurrent;
}

// Level d >= 1 of every instance still running (SearchAtDepth, PCStable.cpp:265-388): the
// adjacency frozen from the surviving edges, every edge's walk over its conditioning sets
// in speculative rounds, then the removals and FreeDegree.
fpgm_status run_level(fpgm_ctx *ctx, double alpha, int d, const fpgm_pc_opts &o, std::vector<PcInst> &P,
                      fpgm_stats &S, double t_start)
{
    struct SchedScope {  // the level's tuned settings, the defaults again afterwards
        fpgm_ctx *c;
        decltype(fpgm_ctx::sched) sched;
        int64_t max_slices;
        ~SchedScope()
        {
            c->sched = sched;
            c->max_slices = max_slices;
        }
    } sched_scope{ctx, ctx->sched, ctx->max_slices};
    const fpgm_level_tune *lt = ctx->tune_try;
    if (!lt && d < (int)ctx->tune.size() && ctx->tune[d].level == d) lt = &ctx->tune[d];
    if (lt) apply_tune(ctx, *lt);
    const float band = o.band > 0.f ? o.band : 0.02f;
    const bool plain = (o.mode == FPGM_MODE_PLAIN);
    const int n = ctx->n_vars;
    const int32_t n_inst = (int32_t)P.size();
    double avg_rows = 0.0;
    for (const InstHost &I : ctx->inst) avg_rows += (double)I.n_rows / (double)ctx->inst.size();
    fpgm_status st = FPGM_OK;
    if (d > kMaxD) return fail(ctx, FPGM_ETOOBIG, "conditioning depth exceeds the device limit");
    const double tl = now_s();
    const double prep0 = ctx->t_prep, wait0 = ctx->t_wait;
    std::vector<std::pair<int32_t, int64_t>> active;  // (instance, edge)
    for (int32_t i = 0; i < n_inst; ++i) {
        PcInst &I = P[i];
        if (!I.cont) continue;
        // frozen copy of the adjacency (PCStable.cpp:272), per node ascending, from the
        // surviving edges in lexicographic order: node u receives its neighbours v < u
        // (edges (v, u), ascending v) before its neighbours v > u (edges (u, v))
        const int64_t E = (int64_t)I.ex.size();
        I.F.off.assign(n + 1, 0);
        for (int64_t e = 0; e < E; ++e) {
            ++I.F.off[I.ex[e] + 1];
            ++I.F.off[I.ey[e] + 1];
        }
        for (int a = 0; a < n; ++a) I.F.off[a + 1] += I.F.off[a];
        I.F.nb.assign((size_t)(2 * E), 0);
        {
            std::vector<int32_t> cur(I.F.off.begin(), I.F.off.end() - 1);
            for (int64_t e = 0; e < E; ++e) {
                I.F.nb[cur[I.ex[e]]++] = I.ey[e];
                I.F.nb[cur[I.ey[e]]++] = I.ex[e];
            }
        }
        I.walks.assign((size_t)E, EdgeWalk{});
        I.choice.assign((size_t)E * d, 0);
        I.zlevel.clear();
        for (int64_t e = 0; e < E; ++e) {
            EdgeWalk &w = I.walks[e];
            w.x = I.ex[e];
            w.y = I.ey[e];
            w.tests = 0;
            w.removed = 0;
            w.zsep = -1;
            if (start_side(w, I.F, d, 0, &I.choice[(size_t)e * d])) active.emplace_back(i, e);
        }
    }
    // device-generated rounds (DevLevel): the host plans each round per edge, the device
    // expands, routes and reduces it
    DevLevel DL;
    bool devmode = false;
    if (!plain && !active.empty() && ctx->devrounds != 0 && ctx->graph == 0 && ctx->group == 0 && d <= kMaxDevD &&
        (ctx->devrounds == 1 || avg_rows <= (double)ctx->devround_max_rows))
        devmode = dev_level_plan(ctx, d, P, active, DL) &&
                  (ctx->devrounds == 1 || DL.walk_tests >= ctx->devround_min_tests);
    if (devmode) {
        HIP_TRY(ctx->d_nbr.ensure((int64_t)DL.nbr.size(), false, ctx->stream), "alloc neighbour store");
        HIP_TRY(ctx->d_binom.ensure((int64_t)DL.binom.size(), false, ctx->stream), "alloc binomials");
        HIP_TRY(hipMemcpyAsync(ctx->d_nbr.p, DL.nbr.data(), sizeof(int32_t) * DL.nbr.size(), hipMemcpyHostToDevice,
                               ctx->stream),
                "upload neighbour store");
        HIP_TRY(hipMemcpyAsync(ctx->d_binom.p, DL.binom.data(), sizeof(int64_t) * DL.binom.size(),
                               hipMemcpyHostToDevice, ctx->stream),
                "upload binomials");
    }
    int64_t executed = 0, inb = 0, rounds = 0;
    const int64_t budget =
        o.round_tests > 0
            ? o.round_tests
            : std::min<int64_t>((int64_t)1 << 20,
                                std::max<int64_t>(8192, (int64_t)(ctx->sched.round_bytes /
                                                                  std::max(1.0, (d + 2) * avg_rows))));
    const double kms_level0 = ctx->k_ms;
    // Edge g of the level is active[g].  done_round[g]: the first round in which a test of
    // edge g came out independent (INT32_MAX while none) -- the host's copy of the
    // device's d_edge_done, filled in round order from the results.
    const int64_t n_act = (int64_t)active.size();
    std::vector<int32_t> done_round((size_t)n_act, INT32_MAX);
    std::vector<int64_t> live((size_t)n_act);  // edges with tests left, not known to be done
    for (int64_t g = 0; g < n_act; ++g) live[g] = g;
    if (!plain && n_act > 0) {
        HIP_TRY(ctx->d_edge_done.ensure(n_act, false, ctx->stream), "alloc edge flags");
        // 0x7f7f7f7f: above every round number
        HIP_TRY(hipMemsetAsync(ctx->d_edge_done.p, 0x7f, sizeof(int32_t) * n_act, ctx->stream), "clear edge flags");
    }
    struct Round {
        int32_t id = 0;
        int slot = 0;
        std::vector<HostTest> tv;
        std::vector<int32_t> zpool;
        std::vector<DevResult> res;  // results of rounds run synchronously
        const DevResult *r = nullptr;
        bool pending = false, async = false;
        bool dev = false;             // a device-generated round: its segments instead of tests
        std::vector<RoundSeg> segs;
    };
    Round rd[2];
    rd[1].slot = 1;
    int32_t next_id = 1;
    struct Gen {  // one host thread's share of a round
        std::vector<HostTest> tv;
        std::vector<int32_t> z;
        std::vector<int64_t> still;
    };
    std::vector<Gen> gen;

    // Builds the next round from the live edges -- every live edge's next W tests in
    // enumeration order; the walks advance past them -- and submits it.  Batched rounds
    // run asynchronously: the host builds round r+1 while the device counts round r, and
    // the device skips round r+1's tests of edges that round r (or earlier) found
    // independent.  A test's outcome depends only on (x, y, Z) and the data, and tests of
    // one round never skip one another, so each edge's first independent test in
    // enumeration order -- its sepset -- is the sequential one.
    double t_build = 0.0, t_process = 0.0;
    auto build = [&](Round &R) -> fpgm_status {
        const double tb0 = now_s();
        struct BuildTimer {
            double t0, &acc;
            ~BuildTimer() { acc += now_s() - t0; }
        } build_timer{tb0, t_build};
        if (o.max_rounds > 0 && rounds >= o.max_rounds)
            return fail(ctx, FPGM_ETIMEOUT, "level exceeded the round limit");
        ++rounds;
        R.id = next_id++;
        const bool plain_round = plain && (o.plain_sample <= 0 || S.plain_tests_per_level[d] < o.plain_sample);
        // tests per edge this round: the budget shared by the live edges, at least a
        // geometric ramp (base^(round-1), capped) so long walks finish in few rounds
        int64_t W = plain_round ? 1 : std::max<int64_t>(1, budget / (int64_t)live.size());
        if (!plain_round) {
            const double ramp = std::pow(ctx->sched.base, (double)(R.id - 1));
            W = std::max<int64_t>(W, (int64_t)std::min<double>(ramp, (double)ctx->sched.cap));
        }
        if (devmode) {
            // per live edge, in live order: its next W subsets as one segment per walk side
            R.dev = true;
            R.segs.clear();
            R.tv.clear();
            int64_t T = 0;
            std::vector<int64_t> still;
            still.reserve(live.size());
            int32_t cbuf[kMaxDevD];
            for (const int64_t g : live) {
                PcInst &I = P[active[g].first];
                EdgeWalk &w = I.walks[active[g].second];
                int64_t We = W;
                if (ctx->sched.promote > 0 && w.tests >= ctx->sched.promote) We = INT64_MAX;
                while (We > 0 && w.side < 2) {
                    const int64_t total = DL.C(w.a, d);
                    const int64_t k = std::min(We, total - w.rank);
                    RoundSeg sg{};
                    sg.edge = (int32_t)g;
                    sg.inst = active[g].first;
                    sg.x = w.x;
                    sg.y = w.y;
                    sg.a = w.a;
                    sg.skip = w.skip;
                    sg.u_off = (int32_t)(DL.nbr_base[(size_t)sg.inst] + I.F.off[w.u]);
                    sg.count = (int32_t)k;
                    sg.rank0 = w.rank;
                    sg.t0 = T;
                    R.segs.push_back(sg);
                    T += k;
                    We -= k;
                    w.rank += k;
                    if (w.rank >= total) start_side(w, I.F, d, w.side + 1, cbuf);
                }
                if (w.side < 2) still.push_back(g);
            }
            live.swap(still);
            R.pending = true;
            R.async = true;
            return dev_round_submit(ctx, ctx->slot[R.slot], DL, R.segs, T, R.id, alpha, o.flags, band);
        }
        // every live edge's tests, in live order; large rounds are generated in
        // contiguous chunks of live edges on several host threads, then concatenated
        const int64_t L = (int64_t)live.size();
        const int nth = (L * W >= kParallelTests) ? ctx->host_threads : 1;
        gen.resize((size_t)nth);
#pragma omp parallel for if (nth > 1) num_threads(nth) schedule(static)
        for (int ch = 0; ch < nth; ++ch) {
            Gen &G = gen[ch];
            G.tv.clear();
            G.z.clear();
            G.still.clear();
            for (int64_t li = L * ch / nth; li < L * (ch + 1) / nth; ++li) {
                const int64_t g = live[li];
                PcInst &I = P[active[g].first];
                const int64_t e = active[g].second;
                EdgeWalk &w = I.walks[e];
                int32_t *c = &I.choice[(size_t)e * d];
                // an edge found dependent on its first `promote` sets rarely ends early:
                // give it the rest of its walk at once
                int64_t We = W;
                if (!plain_round && ctx->sched.promote > 0 && w.tests >= ctx->sched.promote) We = INT64_MAX;
                for (int64_t k = 0; k < We && w.side < 2; ++k) {
                    G.tv.push_back(HostTest{w.x, w.y, d, active[g].first, (int64_t)G.z.size(), (int32_t)g, w.side});
                    for (int i = 0; i < d; ++i) G.z.push_back(I.F.at(w.u, w.skip, c[i]));
                    if (!next_choice(w.a, d, c)) start_side(w, I.F, d, w.side + 1, c);
                }
                if (w.side < 2) G.still.push_back(g);
            }
        }
        // concatenated in chunk order (each chunk copied by its own thread)
        std::vector<int64_t> t_off((size_t)nth + 1, 0), z_off((size_t)nth + 1, 0);
        live.clear();
        for (int ch = 0; ch < nth; ++ch) {
            t_off[ch + 1] = t_off[ch] + (int64_t)gen[ch].tv.size();
            z_off[ch + 1] = z_off[ch] + (int64_t)gen[ch].z.size();
            live.insert(live.end(), gen[ch].still.begin(), gen[ch].still.end());
        }
        R.tv.resize((size_t)t_off[nth]);
        R.zpool.resize((size_t)z_off[nth]);
#pragma omp parallel for if (nth > 1) num_threads(nth) schedule(static)
        for (int ch = 0; ch < nth; ++ch) {
            const Gen &G = gen[ch];
            std::copy(G.z.begin(), G.z.end(), R.zpool.begin() + z_off[ch]);
            HostTest *dst = R.tv.data() + t_off[ch];
            for (size_t k = 0; k < G.tv.size(); ++k) {
                dst[k] = G.tv[k];
                dst[k].z_off += z_off[ch];
            }
        }
        R.pending = true;
        R.async = false;
        const int64_t nt = (int64_t)R.tv.size(), nz = (int64_t)R.zpool.size();
        if (plain) {  // synchronous: every earlier round is processed, nothing to skip
            fpgm_status st2 = run_tests(ctx, plain_round, o, d, R.tv, R.zpool.data(), nz, alpha, band, R.res, S);
            R.r = R.res.data();
            return st2;
        }
        struct SkipScope {  // the round number goes into the launch arguments
            fpgm_ctx *c;
            ~SkipScope() { c->skip_round = 0; }
        } skip_scope{ctx};
        ctx->skip_round = R.id;
        Slot &sl = ctx->slot[R.slot];
        int64_t i1 = 0;
        fpgm_status st2 =
            batch_submit(ctx, sl, R.tv.data(), nt, R.zpool.data(), nz, alpha, o.flags, band, 0, 0, 0, &i1, nullptr);
        if (st2 != FPGM_OK) return st2;
        if (i1 == nt) {
            R.async = true;
            return FPGM_OK;
        }
        // more than one launch sequence (workspace limit): finish the round here
        R.res.resize((size_t)nt);
        for (int64_t i0 = 0;;) {
            st2 = batch_wait(ctx, sl);
            if (st2 != FPGM_OK) return st2;
            std::copy(sl.h_res.p, sl.h_res.p + (i1 - i0), R.res.begin() + i0);
            i0 = i1;
            if (i0 >= nt) break;
            st2 = batch_submit(ctx, sl, R.tv.data(), nt, R.zpool.data(), nz, alpha, o.flags, band, 0, 0, i0, &i1,
                               nullptr);
            if (st2 != FPGM_OK) return st2;
        }
        R.r = R.res.data();
        return FPGM_OK;
    };

    // Takes a round's results in enumeration order: per edge, the tests up to its first
    // independent one count (reference-equivalent); later ones are speculative; the tests
    // of an edge done in an earlier round were skipped by the device and count nowhere.
    auto process = [&](Round &R) -> fpgm_status {
        struct ProcTimer {
            double t0, &acc;
            ~ProcTimer() { acc += now_s() - t0; }
        } proc_timer{now_s(), t_process};
        if (R.dev) {
            // per segment, in plan order: the first independent test decides its edge, the
            // tests up to it count; an edge decided in an earlier round was skipped whole
            Slot &sl = ctx->slot[R.slot];
            fpgm_status st2 = batch_wait(ctx, sl);
            if (st2 != FPGM_OK) return st2;
            const RoundStats &RS = *sl.h_rstats.p;
            if (RS.err)
                return fail(ctx, FPGM_EDEVICE,
                            "internal: device round expansion failed (code " + std::to_string(RS.err) + ")");
            const int32_t *so = sl.h_seg_out.p;
            bool any_done = false;
            int64_t host_skipped = 0;
            int32_t c[kMaxDevD];
            for (size_t si = 0; si < R.segs.size(); ++si) {
                const RoundSeg &sg = R.segs[si];
                const int32_t g = sg.edge, v = so[si];
                PcInst &I = P[active[g].first];
                EdgeWalk &w = I.walks[active[g].second];
                if (done_round[g] < R.id) {
                    if (v != kSegSkipped) return fail(ctx, FPGM_EDEVICE, "internal: the device ran a decided edge");
                    host_skipped += sg.count;
                    continue;
                }
                if (v == kSegSkipped || (v != kSegRun && (v < 0 || v >= sg.count)))
                    return fail(ctx, FPGM_EDEVICE, "internal: bad segment result of a live edge");
                executed += sg.count;
                if (done_round[g] == R.id) continue;  // an earlier segment of this round decided the edge
                if (v == kSegRun) {
                    w.tests += sg.count;
                    continue;
                }
                w.tests += (int64_t)v + 1;
                done_round[g] = R.id;
                w.removed = 1;
                any_done = true;
                unrank_subset(DL, sg.a, d, sg.rank0 + v, c);
                w.zsep = (int64_t)I.zlevel.size();
                for (int i = 0; i < d; ++i) I.zlevel.push_back(DL.nbr[(size_t)sg.u_off + c[i] + (c[i] >= sg.skip ? 1 : 0)]);
            }
            inb += RS.in_band;
            float me = 0.f;
            std::memcpy(&me, &RS.max_err_bits, sizeof(float));
            S.max_rel_err_g2_fp32 = std::max(S.max_rel_err_g2_fp32, (double)me);
            if (ctx->trace)
                fprintf(stderr, "[fpgm]   round %d (device): %zu segments, %lld tests, %lld skipped\n", R.id,
                        R.segs.size(), (long long)sl.n, (long long)host_skipped);
            if (any_done)
                live.erase(std::remove_if(live.begin(), live.end(),
                                          [&](int64_t g) { return done_round[g] != INT32_MAX; }),
                           live.end());
            R.pending = false;
            return FPGM_OK;
        }
        if (R.async) {
            fpgm_status st2 = batch_wait(ctx, ctx->slot[R.slot]);
            if (st2 != FPGM_OK) return st2;
            R.r = ctx->slot[R.slot].h_res.p;
        }
        const DevResult *r = R.r;
        // edges are independent: large rounds are taken in chunks of whole edges on several
        // host threads; the sepset appends (per-instance vectors) follow in test order
        const size_t ntv = R.tv.size();
        const int nthp = (ntv >= (size_t)kParallelTests) ? ctx->host_threads : 1;
        std::vector<size_t> cb((size_t)nthp + 1, ntv);
        cb[0] = 0;
        for (int c = 1; c < nthp; ++c) {
            size_t b = std::max(cb[c - 1], ntv * (size_t)c / (size_t)nthp);
            while (b > 0 && b < ntv && R.tv[b].edge == R.tv[b - 1].edge) ++b;
            cb[c] = b;
        }
        // the slot's descriptors and sink groups of a single-sequence round: the skipped
        // tests come off the per-sink figures batch_wait added
        const Slot &slr = ctx->slot[R.slot];
        const bool have_groups = R.async && slr.mode_of.size() == R.tv.size();
        struct Part {
            double skipped_bytes = 0.0, skipped_pbytes = 0.0, max_err = 0.0;
            double sk_rows[kGroups] = {}, sk_b[kGroups] = {}, sk_pb[kGroups] = {};
            int64_t sk_t[kGroups] = {}, sk_items[kGroups] = {};
            int64_t host_skipped = 0, dev_skipped = 0, executed = 0, in_band = 0;
            bool any_done = false, bad = false;
            std::vector<std::pair<int32_t, size_t>> sep;  // (edge, test) of each edge's first independent test
        };
        std::vector<Part> parts((size_t)nthp);
#pragma omp parallel for if (nthp > 1) num_threads(nthp) schedule(static)
        for (int c = 0; c < nthp; ++c) {
            Part &Q = parts[c];
            for (size_t k = cb[c]; k < cb[c + 1];) {
                const int32_t g = R.tv[k].edge;
                size_t k1 = k;
                while (k1 < cb[c + 1] && R.tv[k1].edge == g) ++k1;
                PcInst &I = P[active[g].first];
                EdgeWalk &w = I.walks[active[g].second];
                if (done_round[g] < R.id) {
                    const double rows = (double)ctx->inst[active[g].first].n_rows;
                    Q.skipped_bytes += (double)(k1 - k) * (d + 2) * rows;
                    for (size_t j = k; j < k1; ++j) {
                        const double pb = packed_row_bytes(ctx, R.tv[j], R.zpool.data()) * rows;
                        Q.skipped_pbytes += pb;
                        Q.dev_skipped += (r[j].df == kSkippedDf);
                        const int gi = have_groups ? slr.mode_of[j] : -1;
                        if (gi >= 0 && gi < kGroups) {
                            Q.sk_t[gi] += 1;
                            Q.sk_items[gi] += (int64_t)slr.h_tests.p[j].n_chunks * slr.h_tests.p[j].n_slices;
                            Q.sk_rows[gi] += rows;
                            Q.sk_b[gi] += (double)(d + 2) * rows;
                            Q.sk_pb[gi] += pb;
                        }
                    }
                    Q.host_skipped += (int64_t)(k1 - k);
                    k = k1;
                    continue;
                }
                bool done = false;
                for (; k < k1; ++k) {
                    if (r[k].df == kSkippedDf) Q.bad = true;
                    ++Q.executed;
                    Q.in_band += r[k].in_band;
                    if (r[k].exact && r[k].df > 0 && r[k].g2_64 > 0.0)
                        Q.max_err = std::max(Q.max_err, std::fabs((double)r[k].g2_32 - r[k].g2_64) / r[k].g2_64);
                    if (done) continue;
                    w.tests += 1;
                    if (r[k].indep) {
                        done = true;
                        Q.any_done = true;
                        done_round[g] = R.id;
                        w.removed = 1;
                        Q.sep.emplace_back(g, k);
                    }
                }
            }
        }
        double skipped_bytes = 0.0, skipped_pbytes = 0.0;
        int64_t host_skipped = 0, dev_skipped = 0;
        bool any_done = false;
        for (const Part &Q : parts) {
            if (Q.bad) return fail(ctx, FPGM_EDEVICE, "internal: the device skipped a test of a live edge");
            skipped_bytes += Q.skipped_bytes;
            skipped_pbytes += Q.skipped_pbytes;
            host_skipped += Q.host_skipped;
            dev_skipped += Q.dev_skipped;
            executed += Q.executed;
            inb += Q.in_band;
            S.max_rel_err_g2_fp32 = std::max(S.max_rel_err_g2_fp32, Q.max_err);
            any_done = any_done || Q.any_done;
            for (int gi = 0; gi < kGroups; ++gi) {
                fpgm_sink_stats &A = ctx->sink[gi];
                A.tests -= Q.sk_t[gi];
                A.items -= Q.sk_items[gi];
                A.row_tests -= Q.sk_rows[gi];
                A.alg_bytes -= Q.sk_b[gi];
                A.packed_bytes -= Q.sk_pb[gi];
            }
            for (const auto &gk : Q.sep) {
                PcInst &I = P[active[gk.first].first];
                EdgeWalk &w = I.walks[active[gk.first].second];
                w.zsep = (int64_t)I.zlevel.size();
                for (int i = 0; i < d; ++i) I.zlevel.push_back(R.zpool[R.tv[gk.second].z_off + i]);
            }
        }
        ctx->k_bytes -= skipped_bytes;  // batch_wait counted every submitted test
        ctx->k_pbytes -= skipped_pbytes;
        if (ctx->trace)
            fprintf(stderr, "[fpgm]   round %d: %zu tests, %lld skipped (device skipped %lld)\n", R.id, R.tv.size(),
                    (long long)host_skipped, (long long)dev_skipped);
        if (any_done)
            live.erase(std::remove_if(live.begin(), live.end(),
                                      [&](int64_t g) { return done_round[g] != INT32_MAX; }),
                       live.end());
        R.pending = false;
        return FPGM_OK;
    };

    int cur = 0;
    if (!live.empty()) {
        st = build(rd[0]);
        if (st != FPGM_OK) return st;
    }
    while (rd[cur].pending) {
        Round &A = rd[cur], &B = rd[cur ^ 1];
        const bool ahead = !plain && ctx->sched.pipeline;
        if (ahead && !live.empty()) {  // round r+1 goes to the device behind round r
            st = build(B);
            if (st != FPGM_OK) return st;
        }
        st = process(A);
        if (st != FPGM_OK) return st;
        if (!ahead && !live.empty()) {
            st = build(B);
            if (st != FPGM_OK) return st;
        }
        cur ^= 1;
    }
    if (!plain) S.kernel_seconds_per_level[d] += (ctx->k_ms - kms_level0) * 1e-3;
    int64_t tests = 0, left = 0;
    for (PcInst &I : P) {
        if (!I.cont) {
            left += (int64_t)I.ex.size();
            continue;
        }
        int64_t wkeep = 0, lt = 0;
        for (size_t e = 0; e < I.walks.size(); ++e) {
            const EdgeWalk &w = I.walks[e];
            lt += w.tests;
            if (w.removed) {
                I.seps.push_back(PcInst::Sep{w.x, w.y, d, d, (int64_t)I.sepz.size()});
                for (int i = 0; i < d; ++i) I.sepz.push_back(I.zlevel[w.zsep + i]);
            } else {
                I.ex[wkeep] = w.x;
                I.ey[wkeep] = w.y;
                ++wkeep;
            }
        }
        I.ex.resize(wkeep);
        I.ey.resize(wkeep);
        I.num_ci_test += lt;
        tests += lt;
        left += wkeep;
        // FreeDegree (PCStable.cpp:620-626): continue while max degree - 1 > d
        std::vector<int32_t> deg((size_t)n, 0);
        for (size_t e = 0; e < I.ex.size(); ++e) {
            ++deg[I.ex[e]];
            ++deg[I.ey[e]];
        }
        const int maxdeg = n > 0 ? *std::max_element(deg.begin(), deg.end()) : 0;
        I.cont = (maxdeg - 1 > d);
    }
    S.tests_per_level[d] = tests;
    S.executed_per_level[d] = executed;
    S.in_band_per_level[d] = inb;
    S.rounds_per_level[d] = rounds;
    S.edges_after_level[d] = left;
    S.seconds_per_level[d] = now_s() - tl;
    S.n_levels = d + 1;
    log_level(o, d, tests, left, now_s() - t_start);
    if (ctx->trace)
        fprintf(stderr, "[fpgm]   level %d host: %.4f s making rounds (%.4f s descriptors), %.4f s taking "
                "results (%.4f s waiting), %.4f s level total\n", d, t_build, ctx->t_prep - prep0, t_process,
                ctx->t_wait - wait0, now_s() - tl);
    return FPGM_OK;
}

// The CI stage of every instance, level-synchronously: level d's tests of all
// instances that reached level d share the same rounds and launches.
fpgm_status run_pc(fpgm_ctx *ctx, double alpha, int32_t depth, const fpgm_pc_opts &o, std::vector<PcInst> &P,
                   fpgm_stats &S, std::vector<std::vector<PcInst>> *snaps = nullptr)
{
    const float band = o.band > 0.f ? o.band : 0.02f;
    const bool plain = (o.mode == FPGM_MODE_PLAIN);
    const int n = ctx->n_vars;
    const int32_t n_inst = (int32_t)P.size();
    fpgm_status st = ensure_crit(ctx, alpha);
    if (st != FPGM_OK) return st;
    struct TraceScope {  // launch tracing for this stage only
        fpgm_ctx *c;
        ~TraceScope() { c->trace = 0; }
    } trace_scope{ctx};
    ctx->trace = (o.verbose >= 2) ? 1 : 0;
    HIP_TRY(hipStreamSynchronize(ctx->stream), "synchronize");
    double avg_rows = 0.0;
    for (const InstHost &I : ctx->inst) avg_rows += (double)I.n_rows / (double)ctx->inst.size();

    const double t_start = now_s();
    for (PcInst &I : P) {
        I.ex.clear();
        I.ey.clear();
        I.seps.clear();
        I.sepz.clear();
        I.num_ci_test = 0;
        I.cont = true;
    }
    std::vector<DevResult> res;
    std::vector<HostTest> tv;

    // ---- level 0: every pair of the complete graph (PCStable.cpp:133-195), pairs x < y in
    // the complete graph's lexicographic edge order (Network.cpp:270-283)
    {
        const double tl = now_s();
        const int64_t np = (int64_t)n * (n - 1) / 2;
        bool gram = false;
        if (!plain && !(o.flags & FPGM_F_NO_GRAM)) {
            const double ms0 = ctx->g_ms + ctx->g_pair_ms;
            const int64_t gram_default = ctx->sched.gram_items_per_cu;
            if (!ctx->tune.empty() && ctx->tune[0].level == 0 && ctx->tune[0].items_per_cu > 0)
                ctx->sched.gram_items_per_cu = ctx->tune[0].items_per_cu;
            st = run_level0_gram(ctx, alpha, o.flags, band, &gram);
            ctx->sched.gram_items_per_cu = gram_default;
            if (st != FPGM_OK) return st;
            S.kernel_seconds_per_level[0] += (ctx->g_ms + ctx->g_pair_ms - ms0) * 1e-3;
        }
        int64_t inb = 0, left = 0;
        if (gram) {
            for (int32_t i = 0; i < n_inst; ++i) {
                PcInst &I = P[i];
                const uint8_t *f = ctx->h_l0_flags.p + (int64_t)i * np;
                int64_t k = 0;
                for (int x = 0; x < n; ++x)
                    for (int y = x + 1; y < n; ++y, ++k) {
                        if (f[k] & 1) {
                            I.seps.push_back(PcInst::Sep{x, y, 0, 0, (int64_t)I.sepz.size()});
                        } else {
                            I.ex.push_back(x);
                            I.ey.push_back(y);
                        }
                    }
                I.num_ci_test += np;
                I.n_sep0 = I.seps.size();
                left += (int64_t)I.ex.size();
            }
            inb = ctx->h_l0_sum.p->in_band;
            float err;
            std::memcpy(&err, &ctx->h_l0_sum.p->max_rel_err_bits, sizeof(float));
            S.max_rel_err_g2_fp32 = std::max(S.max_rel_err_g2_fp32, (double)err);
        } else {
            tv.clear();
            for (int32_t i = 0; i < n_inst; ++i)
                for (int x = 0; x < n; ++x)
                    for (int y = x + 1; y < n; ++y) tv.push_back(HostTest{x, y, 0, i, 0});
            st = run_tests(ctx, plain, o, 0, tv, nullptr, 0, alpha, band, res, S);
            if (st != FPGM_OK) return st;
            const DevResult *r = res.data();
            int64_t k = 0;
            for (int32_t i = 0; i < n_inst; ++i) {
                PcInst &I = P[i];
                for (int x = 0; x < n; ++x)
                    for (int y = x + 1; y < n; ++y, ++k) {
                        inb += r[k].in_band;
                        if (r[k].exact && r[k].df > 0 && r[k].g2_64 > 0.0)
                            S.max_rel_err_g2_fp32 =
                                std::max(S.max_rel_err_g2_fp32, std::fabs((double)r[k].g2_32 - r[k].g2_64) / r[k].g2_64);
                        if (r[k].indep) {
                            I.seps.push_back(PcInst::Sep{x, y, 0, 0, (int64_t)I.sepz.size()});
                        } else {
                            I.ex.push_back(x);
                            I.ey.push_back(y);
                        }
                    }
                I.num_ci_test += np;
                I.n_sep0 = I.seps.size();
                left += (int64_t)I.ex.size();
            }
        }
        S.tests_per_level[0] = np * n_inst;
        S.executed_per_level[0] = np * n_inst;
        S.in_band_per_level[0] = inb;
        S.rounds_per_level[0] = 1;
        S.edges_after_level[0] = left;
        S.seconds_per_level[0] = now_s() - tl;
        S.n_levels = 1;
        log_level(o, 0, np * n_inst, left, now_s() - t_start);
        // level 1 follows unconditionally (PCStable.cpp:209-231: FreeDegree is checked after
        // each level d >= 1 only)
    }

    // ---- levels d >= 1 (PCStable.cpp:209-231, SearchAtDepth :265-388)
    for (int d = 1; d < depth && d < 64; ++d) {
        bool any = false;
        for (const PcInst &I : P) any = any || I.cont;
        if (!any) break;
        if (snaps) snaps->push_back(P);  // the state level d starts from (autotuning replays it)
        st = run_level(ctx, alpha, d, o, P, S, t_start);
        if (st != FPGM_OK) return st;
    }
    S.ci_seconds = now_s() - t_start;
    for (int l = 0; l < S.n_levels; ++l) S.num_ci_test += S.tests_per_level[l];
    for (PcInst &I : P) {  // the skeleton as an adjacency matrix (after the timed region)
        I.adj.assign((size_t)n * n, 0);
        for (size_t e = 0; e < I.ex.size(); ++e) {
            I.adj[(size_t)I.ex[e] * n + I.ey[e]] = 1;
            I.adj[(size_t)I.ey[e] * n + I.ex[e]] = 1;
        }
    }
    return FPGM_OK;
}

fpgm_status check_stage_args(fpgm_ctx *ctx, double alpha, int32_t depth)
{
    if (!(alpha >= 0.0 && alpha <= 1.0)) return fail(ctx, FPGM_EINVAL, "alpha must lie in [0, 1]");
    if (depth < 0) return fail(ctx, FPGM_EINVAL, "depth must be >= 0");
    return FPGM_OK;
}

}  // namespace

fpgm_status fpgm_pc_skeleton(fpgm_ctx *ctx, double alpha, int32_t depth, const fpgm_pc_opts *opts,
                             uint8_t *adj_out, fpgm_stats *stats)
{
    if (!ctx || !adj_out) return FPGM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->inst.size() != 1) return fail(ctx, FPGM_EINVAL, "fpgm_pc_skeleton needs a one-instance context");
    fpgm_status st = check_stage_args(ctx, alpha, depth);
    if (st != FPGM_OK) return st;
    fpgm_pc_opts o{};
    if (opts) o = *opts;
    fpgm_stats S{};
    reset_kernel_stats(ctx);
    std::vector<PcInst> P(1);
    st = run_pc(ctx, alpha, depth, o, P, S);
    if (st != FPGM_OK) return st;
    PcInst &I = P[0];
    // ascending (x, y): level 0's run is already in order, the later levels' removals (few)
    // are sorted and merged into it
    auto before = [](const PcInst::Sep &a, const PcInst::Sep &b) { return a.x != b.x ? a.x < b.x : a.y < b.y; };
    std::sort(I.seps.begin() + (std::ptrdiff_t)I.n_sep0, I.seps.end(), before);
    std::inplace_merge(I.seps.begin(), I.seps.begin() + (std::ptrdiff_t)I.n_sep0, I.seps.end(), before);
    ctx->seps.clear();
    ctx->sep_z.clear();
    ctx->seps.reserve(I.seps.size());
    ctx->sep_z.reserve(I.sepz.size());
    for (const PcInst::Sep &s : I.seps) {
        fpgm_sepset r{};
        r.x = s.x;
        r.y = s.y;
        r.level = s.level;
        r.d = s.d;
        r.z_off = (int64_t)ctx->sep_z.size();
        for (int i = 0; i < s.d; ++i) ctx->sep_z.push_back(I.sepz[s.zoff + i]);
        ctx->seps.push_back(r);
    }
    S.n_sepsets = (int64_t)ctx->seps.size();
    S.sep_z_total = (int64_t)ctx->sep_z.size();
    std::memcpy(adj_out, I.adj.data(), I.adj.size());
    if (stats) *stats = S;
    return FPGM_OK;
}

fpgm_status fpgm_autotune(fpgm_ctx *ctx, double alpha, int32_t depth, const fpgm_pc_opts *opts,
                          fpgm_level_tune *out, int32_t cap, int32_t *n_levels, double *seconds)
{
    if (!ctx || (cap > 0 && !out) || cap < 0) return FPGM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->inst.size() != 1) return fail(ctx, FPGM_EINVAL, "fpgm_autotune needs a one-instance context");
    fpgm_status st = check_stage_args(ctx, alpha, depth);
    if (st != FPGM_OK) return st;
    fpgm_pc_opts o{};
    if (opts) o = *opts;
    if (o.mode == FPGM_MODE_PLAIN) return fail(ctx, FPGM_EINVAL, "autotuning applies to the batched form");
    const double t0 = now_s();
    ctx->tune.clear();
    ctx->tune_try = nullptr;

    // one stage with the current settings, keeping the state each level starts from
    std::vector<std::vector<PcInst>> snaps;
    fpgm_stats S{};
    std::vector<PcInst> P(1);
    reset_kernel_stats(ctx);
    const int64_t gram0 = ctx->g_launches;
    st = run_pc(ctx, alpha, depth, o, P, S, &snaps);
    if (st != FPGM_OK) return st;
    const bool used_gram = ctx->g_launches > gram0;
    const int L = S.n_levels;
    auto current = [&](int lvl) {
        fpgm_level_tune t{};
        t.level = lvl;
        t.pipeline = ctx->sched.pipeline;
        t.concurrent = ctx->sched.concurrent;
        t.items_per_cu = lvl == 0 ? ctx->sched.gram_items_per_cu : ctx->sched.items_per_cu;
        for (int g = 0; g < kTuneGroups; ++g) t.wg_per_cu[g] = ctx->sched.wpc[g];
        t.max_slices = ctx->max_slices;
        t.round_cap = ctx->sched.cap;
        return t;
    };
    std::vector<fpgm_level_tune> res((size_t)L);
    for (int l = 0; l < L; ++l) res[l] = current(l);

    // level 0: the row split of the Gram tiles
    if (used_gram) {
        auto time_gram = [&](int64_t ipc) -> double {
            double best = 1e30;
            for (int r = 0; r < 3; ++r) {
                const int64_t keep = ctx->sched.gram_items_per_cu;
                ctx->sched.gram_items_per_cu = ipc;
                const double ms0 = ctx->g_ms + ctx->g_pair_ms;
                bool used = false;
                const fpgm_status s2 = run_level0_gram(ctx, alpha, o.flags, o.band > 0.f ? o.band : 0.02f, &used);
                ctx->sched.gram_items_per_cu = keep;
                if (s2 != FPGM_OK) return -1.0;
                best = std::min(best, ctx->g_ms + ctx->g_pair_ms - ms0);
            }
            return best;
        };
        double best = time_gram(res[0].items_per_cu);
        if (best < 0) return FPGM_EDEVICE;
        res[0].default_ms = best;
        for (int64_t ipc : {2, 4, 16, 32}) {
            const double ms = time_gram(ipc);
            if (ms < 0) return FPGM_EDEVICE;
            if (ms < best * 0.98) {
                best = ms;
                res[0].items_per_cu = ipc;
            }
        }
        res[0].tuned_ms = best;
    } else {
        res[0].level = -1;  // nothing to tune at level 0
    }

    // levels d >= 1: each replayed alone from its starting state under candidate settings;
    // coordinate descent, a candidate kept when it is at least 2 % faster
    fpgm_status fail_st = FPGM_OK;
    auto time_level = [&](int d, const fpgm_level_tune &t, int reps) -> double {
        double best = 1e30;
        for (int r = 0; r < reps; ++r) {
            std::vector<PcInst> Q = snaps[d - 1];
            fpgm_stats S2{};
            ctx->tune_try = &t;
            const double t1 = now_s();
            const fpgm_status s2 = run_level(ctx, alpha, d, o, Q, S2, t1);
            const double dt = now_s() - t1;
            ctx->tune_try = nullptr;
            if (s2 != FPGM_OK) {
                fail_st = s2;
                return -1.0;
            }
            best = std::min(best, dt * 1e3);
        }
        return best;
    };
    for (int d = 1; d < L && d <= (int)snaps.size(); ++d) {
        if (S.executed_per_level[d] == 0) {
            res[d].level = -1;
            continue;
        }
        const int reps = S.seconds_per_level[d] > 0.05 ? 1 : 3;
        fpgm_level_tune cur = res[d];
        double best = time_level(d, cur, reps);
        if (best < 0) return fail_st;
        cur.default_ms = best;
        std::vector<fpgm_level_tune> cands;
        auto try_all = [&]() -> fpgm_status {
            for (const fpgm_level_tune &c : cands) {
                const double ms = time_level(d, c, reps);
                if (ms < 0) return fail_st;
                if (ms < best * 0.98) {
                    best = ms;
                    cur = c;
                    cur.default_ms = res[d].default_ms;
                }
            }
            cands.clear();
            return FPGM_OK;
        };
        res[d].default_ms = cur.default_ms;
        {
            fpgm_level_tune c = cur;
            c.pipeline = 1 - cur.pipeline;
            cands.push_back(c);
        }
        if ((st = try_all()) != FPGM_OK) return st;
        {
            fpgm_level_tune c = cur;
            c.concurrent = 1 - cur.concurrent;
            cands.push_back(c);
        }
        if ((st = try_all()) != FPGM_OK) return st;
        for (int64_t ipc : {4, 8, 16, 32})
            if (ipc != cur.items_per_cu) {
                fpgm_level_tune c = cur;
                c.items_per_cu = ipc;
                cands.push_back(c);
            }
        if ((st = try_all()) != FPGM_OK) return st;
        for (int g = 0; g < kTuneGroups; ++g) {
            for (int64_t f : {-1, 1}) {
                fpgm_level_tune c = cur;
                c.wg_per_cu[g] = f < 0 ? std::max<int64_t>(1, cur.wg_per_cu[g] / 2) : std::min<int64_t>(16, cur.wg_per_cu[g] * 2);
                if (c.wg_per_cu[g] != cur.wg_per_cu[g]) cands.push_back(c);
            }
            if ((st = try_all()) != FPGM_OK) return st;
        }
        for (int64_t ms : {1, 2, 4, 16})
            if (ms != cur.max_slices) {
                fpgm_level_tune c = cur;
                c.max_slices = ms;
                cands.push_back(c);
            }
        if ((st = try_all()) != FPGM_OK) return st;
        if (S.rounds_per_level[d] > 1)
            for (int64_t rc : {4, 16})
                if (rc != cur.round_cap) {
                    fpgm_level_tune c = cur;
                    c.round_cap = rc;
                    cands.push_back(c);
                }
        if ((st = try_all()) != FPGM_OK) return st;
        cur.level = d;
        cur.tuned_ms = best;
        res[d] = cur;
    }
    ctx->tune = res;
    if (n_levels) *n_levels = L;
    for (int l = 0; l < L && l < cap; ++l) out[l] = res[l];
    if (seconds) *seconds = now_s() - t0;
    return FPGM_OK;
}

fpgm_status fpgm_tune_set(fpgm_ctx *ctx, const fpgm_level_tune *levels, int32_t n)
{
    if (!ctx || n < 0 || (n > 0 && !levels)) return FPGM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    std::vector<fpgm_level_tune> t;
    for (int32_t i = 0; i < n; ++i) {
        const fpgm_level_tune &r = levels[i];
        if (r.level < 0) continue;
        if (r.level >= 64 || r.pipeline > 1 || r.concurrent > 1 || r.items_per_cu < 0 || r.items_per_cu > 4096 || r.max_slices < 0 ||
            r.round_cap < 0)
            return fail(ctx, FPGM_EINVAL, "bad tuning record");
        for (int g = 0; g < kTuneGroups; ++g)
            if (r.wg_per_cu[g] < 0 || r.wg_per_cu[g] > 64) return fail(ctx, FPGM_EINVAL, "bad tuning record");
        if ((int)t.size() <= r.level) {
            fpgm_level_tune none{};
            none.level = -1;
            t.resize((size_t)r.level + 1, none);
        }
        t[r.level] = r;
    }
    ctx->tune = t;
    return FPGM_OK;
}

fpgm_status fpgm_tune_clear(fpgm_ctx *ctx)
{
    if (!ctx) return FPGM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->tune.clear();
    return FPGM_OK;
}

fpgm_status fpgm_pc_skeleton_multi(fpgm_ctx *ctx, double alpha, int32_t depth, const fpgm_pc_opts *opts,
                                   int32_t *edge_freq, uint8_t *adj_out, int64_t *num_ci_test_out,
                                   fpgm_stats *stats)
{
    if (!ctx || !edge_freq) return FPGM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    fpgm_status st = check_stage_args(ctx, alpha, depth);
    if (st != FPGM_OK) return st;
    fpgm_pc_opts o{};
    if (opts) o = *opts;
    fpgm_stats S{};
    reset_kernel_stats(ctx);
    std::vector<PcInst> P(ctx->inst.size());
    st = run_pc(ctx, alpha, depth, o, P, S);
    if (st != FPGM_OK) return st;
    const size_t nn = (size_t)ctx->n_vars * ctx->n_vars;
    std::memset(edge_freq, 0, sizeof(int32_t) * nn);
    int64_t nsep = 0;
    for (size_t i = 0; i < P.size(); ++i) {
        for (size_t k = 0; k < nn; ++k) edge_freq[k] += P[i].adj[k];
        if (adj_out) std::memcpy(adj_out + i * nn, P[i].adj.data(), nn);
        if (num_ci_test_out) num_ci_test_out[i] = P[i].num_ci_test;
        nsep += (int64_t)P[i].seps.size();
    }
    ctx->seps.clear();
    ctx->sep_z.clear();
    S.n_sepsets = nsep;
    if (stats) *stats = S;
    return FPGM_OK;
}

fpgm_status fpgm_sepset_count(const fpgm_ctx *ctx, int64_t *n_records, int64_t *n_z)
{
    if (!ctx || !n_records || !n_z) return FPGM_EINVAL;
    *n_records = (int64_t)ctx->seps.size();
    *n_z = (int64_t)ctx->sep_z.size();
    return FPGM_OK;
}

fpgm_status fpgm_sepset_get(const fpgm_ctx *ctx, fpgm_sepset *records, int64_t cap_records, int32_t *z,
                            int64_t cap_z)
{
    if (!ctx) return FPGM_EINVAL;
    if (cap_records < (int64_t)ctx->seps.size() || cap_z < (int64_t)ctx->sep_z.size()) return FPGM_EINVAL;
    if (!ctx->seps.empty()) {
        if (!records) return FPGM_EINVAL;
        std::memcpy(records, ctx->seps.data(), sizeof(fpgm_sepset) * ctx->seps.size());
    }
    if (!ctx->sep_z.empty()) {
        if (!z) return FPGM_EINVAL;
        std::memcpy(z, ctx->sep_z.data(), sizeof(int32_t) * ctx->sep_z.size());
    }
    return FPGM_OK;
}

fpgm_status fpgm_last_gram_stats(const fpgm_ctx *ctx, double *gram_ms, double *int8_ops, double *pairs_ms,
                                 int64_t *launches)
{
    if (!ctx) return FPGM_EINVAL;
    if (gram_ms) *gram_ms = ctx->g_ms;
    if (int8_ops) *int8_ops = ctx->g_ops;
    if (pairs_ms) *pairs_ms = ctx->g_pair_ms;
    if (launches) *launches = ctx->g_launches;
    return FPGM_OK;
}

fpgm_status fpgm_last_group_stats(const fpgm_ctx *ctx, double *ms, double *int8_ops, int64_t *tests,
                                  int64_t *sequences)
{
    if (!ctx) return FPGM_EINVAL;
    if (ms) *ms = ctx->grp_ms;
    if (int8_ops) *int8_ops = ctx->grp_ops;
    if (tests) *tests = ctx->grp_tests;
    if (sequences) *sequences = ctx->grp_launches;
    return FPGM_OK;
}

fpgm_status fpgm_last_sink_stats(const fpgm_ctx *ctx, fpgm_sink_stats *out, int32_t cap, int32_t *n)
{
    if (!ctx || !n || (cap > 0 && !out)) return FPGM_EINVAL;
    *n = kGroups;
    for (int g = 0; g < kGroups && g < cap; ++g) out[g] = ctx->sink[g];
    return FPGM_OK;
}

fpgm_status fpgm_last_packed_bytes(const fpgm_ctx *ctx, double *packed_bytes)
{
    if (!ctx) return FPGM_EINVAL;
    if (packed_bytes) *packed_bytes = ctx->k_pbytes;
    return FPGM_OK;
}

fpgm_status fpgm_last_kernel_stats(const fpgm_ctx *ctx, double *kernel_ms, double *alg_bytes, int64_t *launches)
{
    if (!ctx) return FPGM_EINVAL;
    if (kernel_ms) *kernel_ms = ctx->k_ms;
    if (alg_bytes) *alg_bytes = ctx->k_bytes;
    if (launches) *launches = ctx->k_launches;
    return FPGM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- allocation-free batch
//
// fpgm_ci_batch_noalloc: fpgm_ci_batch for callers that cannot allocate device memory.
// Every buffer the counting launches need beyond the context's columns -- descriptors,
// z pool, work items, results, completion and queue counters, critical values and the
// workspace of tables split across workgroups -- lives in `scratch`, one static device
// array per GPU, whose address is taken once per device with hipGetSymbolAddress.  The
// host stages descriptors in static arrays.  Calls are serialised by na_mu within a
// process; calls from several processes on one GPU would race on its scratch.

namespace {

constexpr int64_t kNaTests = 4096;          // tests per call
constexpr int64_t kNaZ = kNaTests * 8;      // z pool entries per call
constexpr int64_t kNaItems = 65536;         // work items per launch sequence
constexpr int64_t kNaWs = (int64_t)1 << 22; // workspace int32 entries (16 MiB)
constexpr int kNaDevices = 64;

constexpr int64_t na_align(int64_t b)
{
    return (b + 255) / 256 * 256;
}
// scratch layout (byte offsets)
constexpr int64_t kNaOffTests = 0;
constexpr int64_t kNaOffRes = kNaOffTests + na_align(kNaTests * (int64_t)sizeof(DevTest));
constexpr int64_t kNaOffZ = kNaOffRes + na_align(kNaTests * (int64_t)sizeof(DevResult));
constexpr int64_t kNaOffItems = kNaOffZ + na_align(kNaZ * 4);
constexpr int64_t kNaOffCrit = kNaOffItems + na_align(kNaItems * 8);
constexpr int64_t kNaOffDone = kNaOffCrit + na_align(kCritTab * 8);
constexpr int64_t kNaOffQueue = kNaOffDone + na_align(kNaTests * 4);
constexpr int64_t kNaOffWs = kNaOffQueue + na_align(kGroups * (kXcds + 1) * 4);
constexpr int64_t kNaBytes = kNaOffWs + kNaWs * 4;

}  // namespace

static __device__ __align__(256) uint8_t scratch[kNaBytes];

namespace {

std::mutex na_mu;
uint8_t *na_ptr[kNaDevices] = {};                 // scratch's device address, per device
double na_crit_alpha[kNaDevices];                 // alpha of the critical values in scratch
bool na_crit_valid[kNaDevices] = {};
// host staging (calls hold na_mu)
DevTest na_tests[kNaTests];
DevResult na_res[kNaTests];
uint64_t na_items[kNaItems];
int32_t na_gidx[kNaTests];
uint8_t na_fit[kNaTests];
double na_crit[kCritTab];

// scratch's address on the context's device, taken on the first call there (with that
// device current) and kept
fpgm_status na_scratch(fpgm_ctx *ctx, uint8_t **out)
{
    if (ctx->dev < 0 || ctx->dev >= kNaDevices) return fail(ctx, FPGM_EDEVICE, "device index beyond the scratch table");
    if (!na_ptr[ctx->dev]) {
        int cur = 0;
        HIP_TRY(hipGetDevice(&cur), "hipGetDevice");
        if (cur != ctx->dev) HIP_TRY(hipSetDevice(ctx->dev), "hipSetDevice");
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != ctx->dev) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p) return hip_fail(ctx, e != hipSuccess ? e : hipErrorInvalidSymbol, "scratch address");
        na_ptr[ctx->dev] = (uint8_t *)p;
    }
    *out = na_ptr[ctx->dev];
    return FPGM_OK;
}

// One launch sequence over tests [i0, i1) (descriptors already in na_tests[0, i1 - i0),
// groups in na_gidx): uploads, clears the counters and the workspace it uses, counts,
// takes the p-values and reads the results back into na_res.
fpgm_status na_sequence(fpgm_ctx *ctx, uint8_t *base, int64_t nb, const int32_t *zpool, int64_t zlen,
                        const int64_t *grp_begin, const int64_t *grp_len, const int64_t *rpc, bool lane_w128,
                        int64_t ws_used, double alpha, uint32_t flags)
{
    DevTest *d_tests = (DevTest *)(base + kNaOffTests);
    DevResult *d_res = (DevResult *)(base + kNaOffRes);
    int32_t *d_z = (int32_t *)(base + kNaOffZ);
    uint64_t *d_items = (uint64_t *)(base + kNaOffItems);
    double *d_crit = (double *)(base + kNaOffCrit);
    uint32_t *d_done = (uint32_t *)(base + kNaOffDone);
    uint32_t *d_queue = (uint32_t *)(base + kNaOffQueue);
    int32_t *d_ws = (int32_t *)(base + kNaOffWs);
    hipStream_t s = ctx->stream;
    int64_t ni = 0;
    for (int g = 0; g < kGroups; ++g) ni += grp_len[g];
    // the previous call's contents stay in scratch: counters and the workspace region this
    // sequence uses start from zero (done and queue counters are adjacent)
    HIP_TRY(launch_zero(d_done, (kNaOffWs - kNaOffDone) / 4, s), "clear counters");
    HIP_TRY(launch_zero((uint32_t *)d_ws, ws_used, s), "clear workspace");
    if (!na_crit_valid[ctx->dev] || na_crit_alpha[ctx->dev] != alpha) {
        for (int df = 1; df <= kCritTab; ++df) na_crit[df - 1] = chisq_crit(df, alpha);
        HIP_TRY(hipMemcpyAsync(d_crit, na_crit, sizeof(double) * kCritTab, hipMemcpyHostToDevice, s), "upload crit");
        na_crit_alpha[ctx->dev] = alpha;
        na_crit_valid[ctx->dev] = true;
    }
    if (zlen > 0) HIP_TRY(hipMemcpyAsync(d_z, zpool, sizeof(int32_t) * zlen, hipMemcpyHostToDevice, s), "upload z pool");
    HIP_TRY(hipMemcpyAsync(d_tests, na_tests, sizeof(DevTest) * nb, hipMemcpyHostToDevice, s), "upload tests");
    HIP_TRY(hipMemcpyAsync(d_items, na_items, sizeof(uint64_t) * ni, hipMemcpyHostToDevice, s), "upload items");
    LaunchArgs a{};
    a.inst = ctx->d_inst;
    a.tests = d_tests;
    a.zpool = d_z;
    a.card = ctx->d_card;
    a.colw = ctx->d_colw;
    a.poff = ctx->d_poff;
    a.items = d_items;
    a.ws = d_ws;
    a.done = d_done;
    a.crit = d_crit;
    a.alpha = alpha;
    a.z_alpha = (alpha > 0.0 && alpha < 1.0) ? norm_upper_quantile(alpha) : 0.0;
    a.band = 0.02f;
    a.flags = flags;
    a.res = d_res;
    a.edge_done = nullptr;
    a.round = 0;
    for (int gi = 0; gi < kGroups; ++gi) {
        if (grp_len[gi] == 0) continue;
        a.item_begin = grp_begin[gi];
        a.n_items = grp_len[gi];
        a.rows_per_chunk = rpc[gi];
        a.queue = d_queue + gi * kXcds;
        a.n_items_dev = nullptr;
        const int sink = (gi == 1 && lane_w128) ? kSinkLaneW128 : kGroupSink[gi];
        int64_t grid = std::min<int64_t>(a.n_items, (int64_t)ctx->n_cu * wpc_of(ctx, gi));
        grid = std::max<int64_t>(8, (grid + 7) / 8 * 8);
        HIP_TRY(launch_count_stat(a, sink, ctx->d_pcols != nullptr, (int)grid, s), "launch count_stat");
    }
    HIP_TRY(launch_pvalue(d_res, nb, alpha, d_tests, nullptr, 0, d_queue, kGroups * (kXcds + 1), s), "launch pvalue");
    HIP_TRY(hipMemcpyAsync(na_res, d_res, sizeof(DevResult) * nb, hipMemcpyDeviceToHost, s), "download results");
    HIP_TRY(hipStreamSynchronize(s), "synchronize");
    return FPGM_OK;
}

}  // namespace

extern "C" fpgm_status fpgm_ci_batch_noalloc(fpgm_ctx *ctx, const fpgm_test *tests, int64_t n,
                                             const int32_t *z_pool, int64_t z_pool_len, double alpha, uint32_t flags,
                                             fpgm_result *out)
{
    if (!ctx || (n > 0 && (!tests || !out)) || n < 0) return FPGM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    std::lock_guard<std::mutex> lk2(na_mu);
    if (n > kNaTests) return fail(ctx, FPGM_ETOOBIG, "more than 4096 tests in one allocation-free call");
    if (z_pool_len > kNaZ) return fail(ctx, FPGM_ETOOBIG, "z pool beyond 32768 entries in an allocation-free call");
    // checks and descriptors (fpgm_ci_batch's rules, tables of at most kLdsCells cells)
    int64_t N = 0;
    for (int64_t i = 0; i < n; ++i) {
        const fpgm_test &t = tests[i];
        if (t.d > 0 && (t.z_off < 0 || (int64_t)t.z_off + t.d > z_pool_len || !z_pool))
            return fail(ctx, FPGM_EINVAL, "z_off/d outside the z pool");
        fpgm_status st = check_test(ctx, t.x, t.y, t.d > 0 ? z_pool + t.z_off : nullptr, t.d, alpha);
        if (st != FPGM_OK) return st;
        DevTest T{};
        T.x = t.x;
        T.y = t.y;
        T.d = t.d;
        T.z_off = t.z_off;
        T.inst = 0;
        T.edge = -1;
        T.cx = ctx->card[t.x];
        T.cy = ctx->card[t.y];
        int64_t dimz = 1;
        bool narrow = ctx->colw[t.x] != 8 && ctx->colw[t.y] != 8;
        for (int j = 0; j < t.d; ++j) {
            dimz *= ctx->card[z_pool[t.z_off + j]];
            narrow = narrow && ctx->colw[z_pool[t.z_off + j]] != 8;
            if (dimz > kLdsCells) break;
        }
        const int64_t cells = dimz * T.cx * T.cy;
        bool fits = false;
        T.flags |= narrow ? kTFNarrow : 0;  // before the routing, which reads it
        const int gi =
            cells <= kLdsCells ? lds_group(ctx, T, cells, dimz * (T.cx + T.cy + 1), ctx->inst[0].n_rows, &fits) : -1;
        if (gi < 0) return fail(ctx, FPGM_ETOOBIG, "allocation-free calls take tables of at most 8192 cells");
        T.dimz = (int32_t)dimz;
        T.n_slices = 1;
        T.slice_cells = (int32_t)cells;
        T.ws_off = -1;
        na_tests[i] = T;
        na_gidx[i] = gi;
        na_fit[i] = fits ? 1 : 0;
    }
    if (n == 0) return FPGM_OK;
    uint8_t *base = nullptr;
    fpgm_status st = na_scratch(ctx, &base);
    if (st != FPGM_OK) return st;
    N = ctx->inst[0].n_rows;
    const int64_t steps = std::max<int64_t>(1, (N + kRowsPerStep - 1) / kRowsPerStep);
    const int64_t target_items = (int64_t)ctx->n_cu * ctx->sched.items_per_cu;
    // launch sequences: tests in order while their workspace and work items fit scratch
    for (int64_t i0 = 0; i0 < n;) {
        const int64_t m = n - i0;
        // rows per chunk: one value for the LDS sinks, enough items over all tests to fill the CUs
        const int64_t cpt = std::min(steps, std::max<int64_t>(1, (target_items + m - 1) / m));
        const int64_t r = std::min<int64_t>((steps + cpt - 1) / cpt * kRowsPerStep, kLaneMaxRowsPerChunk);
        const int64_t n_chunks = std::max<int64_t>(1, (N + r - 1) / r);
        // 97..128-cell tables join the shared launch when there is one (batch_submit's rule)
        bool any_shared = false;
        for (int64_t k = i0; k < n; ++k) any_shared = any_shared || na_gidx[k] == 2;
        int64_t ws_used = 0, items = 0, nb = 0;
        int64_t grp_len[kGroups] = {0, 0, 0, 0, 0, 0};
        for (; i0 + nb < n; ++nb) {
            const int64_t k = i0 + nb;
            DevTest T = na_tests[k];
            int gi = na_gidx[k];
            bool fit = na_fit[k] != 0;
            if (gi == kGroupPending) {
                if (any_shared) {
                    const int64_t cells = (int64_t)T.dimz * T.cx * T.cy;
                    T.flags |= kTFLdsTable;
                    fit = cells + (int64_t)T.dimz * (T.cx + T.cy + 1) <= kLdsCells;
                    gi = 2;
                } else {
                    gi = 1;
                }
            }
            const int64_t cells = (int64_t)T.dimz * T.cx * T.cy;
            const bool need_ws = n_chunks > 1 || !fit;
            const int64_t need = need_ws ? (cells + (int64_t)T.dimz * (T.cx + T.cy + 1) + 63) / 64 * 64 : 0;
            if (ws_used + need > kNaWs || items + n_chunks > kNaItems) break;
            T.n_chunks = (int32_t)n_chunks;
            if (need_ws) {
                T.ws_off = ws_used;
                ws_used += need;
            } else {
                T.flags |= kTFLdsMarg;
            }
            T.item0 = (int32_t)grp_len[gi];
            grp_len[gi] += n_chunks;
            items += n_chunks;
            na_tests[k] = T;
            na_gidx[k] = gi;
        }
        if (nb == 0) return fail(ctx, FPGM_ETOOBIG, "a test's workspace exceeds the allocation-free scratch");
        // descriptors of the sequence to the front of the staging arrays, items by group
        int64_t grp_begin[kGroups], acc = 0, rpc[kGroups];
        for (int g = 0; g < kGroups; ++g) {
            grp_begin[g] = acc;
            acc += grp_len[g];
            rpc[g] = r;
        }
        for (int64_t k = 0; k < nb; ++k) {
            DevTest &T = na_tests[i0 + k];
            T.item0 += (int32_t)grp_begin[na_gidx[i0 + k]];
            for (int64_t j = 0; j < T.n_chunks; ++j) na_items[T.item0 + j] = ((uint64_t)k << 32) | (uint64_t)j;
            if (i0 > 0) na_tests[k] = T;
        }
        st = na_sequence(ctx, base, nb, z_pool, z_pool_len, grp_begin, grp_len, rpc, !any_shared, ws_used, alpha,
                         flags);
        if (st != FPGM_OK) return st;
        for (int64_t k = 0; k < nb; ++k) out[i0 + k] = to_result(na_res[k]);
        i0 += nb;
    }
    return FPGM_OK;
}
