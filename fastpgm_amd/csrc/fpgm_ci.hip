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
#include "mle_kernels.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace fpgm;

namespace {

constexpr int64_t kMaxCells = (int64_t)1 << 31;   // dense-table limit (the reference indexes with int)
constexpr int64_t kWsInit = (int64_t)1 << 24;     // workspace, int32 entries
constexpr int64_t kWsMax = (int64_t)1 << 30;      // 4 GiB of workspace at most per launch
constexpr int kMaxD = kMaxZLds;

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
constexpr int kGroups = 5;
constexpr int kGroupSink[kGroups] = {kSinkLane, kSinkLaneW, kSinkShared, kSinkSlice, kSinkGlobal};

// Buffers of one in-flight batch.  Two slots let the host build one batch while the
// device runs the other; the workspace and completion counters are shared, since both
// slots launch on the context's one stream and every launch leaves them zero.
struct Slot {
    DevBuf<DevTest> d_tests;
    DevBuf<int32_t> d_zpool, d_item_test;
    DevBuf<DevResult> d_res;
    DevBuf<WideItem> d_wide_items;
    DevBuf<WidePart> d_wide_parts;
    DevBuf<WideTest> d_wide_tests;
    HostBuf<DevTest> h_tests;
    HostBuf<int32_t> h_items, h_zpool;
    HostBuf<DevResult> h_res;
    HostBuf<WideItem> h_wide_items;
    HostBuf<WideTest> h_wide_tests;
    std::vector<DevTest> dt;
    std::vector<int32_t> mode_of;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_done = nullptr;
    hipEvent_t ev_grp[kGroups + 1] = {};  // per-sink timing when tracing
    bool pending = false;
    int64_t n = 0;       // tests of the pending launch sequence
    double bytes = 0.0;  // their algorithmic bytes
    struct GroupInfo {
        int64_t tests = 0, items = 0, dsum = 0;
        double csum = 0.0, cmax = 0.0, bytes = 0.0;
    } grp[kGroups];

    hipError_t init()
    {
        hipError_t e = hipSuccess;
        for (hipEvent_t *ev : {&ev0, &ev1, &ev_done})
            if ((e = hipEventCreate(ev)) != hipSuccess) return e;
        for (hipEvent_t &ev : ev_grp)
            if ((e = hipEventCreate(&ev)) != hipSuccess) return e;
        return e;
    }
    void release()
    {
        d_tests.release();
        d_zpool.release();
        d_item_test.release();
        d_res.release();
        d_wide_items.release();
        d_wide_parts.release();
        d_wide_tests.release();
        h_tests.release();
        h_items.release();
        h_zpool.release();
        h_res.release();
        h_wide_items.release();
        h_wide_tests.release();
        for (hipEvent_t ev : {ev0, ev1, ev_done})
            if (ev) (void)hipEventDestroy(ev);
        for (hipEvent_t ev : ev_grp)
            if (ev) (void)hipEventDestroy(ev);
    }
};

struct HostTest {
    int32_t x, y, d;
    int32_t inst;   // dataset instance
    int64_t z_off;  // into the host z pool
};

struct InstHost {
    const uint8_t *cols;  // device address
    int64_t stride, n_rows;
};

}  // namespace

struct fpgm_ctx {
    int dev = 0;
    hipStream_t stream = nullptr;
    const uint8_t *d_cols = nullptr;  // all instances' columns (owned unless borrowed)
    bool owns_cols = false;
    std::vector<InstHost> inst;       // dataset instances (1, or the bootstrap ensemble)
    DevInst *d_inst = nullptr;
    int32_t n_vars = 0;
    std::vector<int32_t> card;
    int32_t *d_card = nullptr;
    std::string err;
    std::mutex mu;
    int n_cu = 256;
    struct {
        double base = 2.0;    // round ramp factor
        int64_t cap = 8;      // ramp cap (tests per edge per round); larger rounds come from the budget
        int64_t promote = 0;  // dependent tests after which an edge runs the rest of its walk
        int lanes = 1;        // edge lanes whose rounds alternate (host work overlaps device work)
    } sched;
    int64_t max_slices = kMaxSlices;  // z slices before a table goes to the device-memory sink
    uint32_t sink_mask = 0;  // FPGM_SINKS=shared: bit 0 no wide per-lane sink, bit 1 no z slices

    double crit_alpha = -1.0;
    double z_alpha = 0.0;
    std::vector<double> crit_host;
    double *d_crit = nullptr;

    Slot slot[2];
    DevBuf<int32_t> d_ws;
    DevBuf<uint32_t> d_done, d_flag;

    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int trace = 0;               // > 0: one line per counting launch on stderr
    double t_prep = 0.0, t_wait = 0.0;  // host seconds building launches / waiting for them (trace)
    double k_ms = 0.0, k_bytes = 0.0;
    int64_t k_launches = 0;

    // level-0 Gram path
    DevBuf<int32_t> d_C, d_rowvar, d_tiles, d_off, d_pairs;
    DevBuf<uint8_t> d_rowstate;
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

    // rows per chunk, common to the launch: enough work items to fill every CU
    const int64_t target_items = (int64_t)ctx->n_cu * 16;
    int64_t rpc;
    if (rows_per_chunk_forced > 0) {
        rpc = (rows_per_chunk_forced + kRowsPerStep - 1) / kRowsPerStep * kRowsPerStep;
    } else {
        int64_t cpt = std::max<int64_t>(1, (target_items + (n - i0) - 1) / (n - i0));
        cpt = std::min(cpt, steps);
        rpc = (steps + cpt - 1) / cpt * kRowsPerStep;
    }
    rpc = std::min<int64_t>(rpc, kLaneMaxRowsPerChunk);

    std::vector<DevTest> &dt = S.dt;
    std::vector<int32_t> &mode_of = S.mode_of;
    dt.clear();
    mode_of.clear();
    int64_t ws_used = 0, n_items_total = 0;
    int64_t i1 = i0;
    for (; i1 < n; ++i1) {
        const HostTest &h = tests[i1];
        DevTest T{};
        T.x = h.x;
        T.y = h.y;
        T.d = h.d;
        T.inst = h.inst;
        const int64_t n_chunks = std::max<int64_t>(1, (ctx->inst[h.inst].n_rows + rpc - 1) / rpc);
        T.z_off = (int32_t)h.z_off;
        T.cx = ctx->card[h.x];
        T.cy = ctx->card[h.y];
        int64_t dimz = 1;
        for (int j = 0; j < h.d; ++j) {
            dimz *= ctx->card[zpool[h.z_off + j]];
            if (dimz > kMaxCells) break;
        }
        const int64_t cells = dimz * T.cx * T.cy;
        if (dimz > kMaxCells || cells >= kMaxCells) return fail(ctx, FPGM_ETOOBIG, "contingency table exceeds 2^31 cells");
        T.dimz = (int32_t)dimz;
        T.n_chunks = (int32_t)n_chunks;
        T.n_slices = 1;
        T.slice_cells = (int32_t)cells;
        const int64_t marg = dimz * (T.cx + T.cy + 1);
        T.flags = tflags;
        bool fits;
        int mode;
        if (cells <= kLaneCells) {
            mode = kSinkLane;
            T.flags |= kTFLane;
            fits = cells + marg <= kSmallCells;
        } else if (cells <= kLaneWideCells && !(ctx->sink_mask & 1)) {
            mode = kSinkLaneW;
            fits = cells + marg <= kSmallCells;
        } else if (cells <= kLdsCells) {
            mode = kSinkShared;
            T.flags |= kTFLdsTable;
            fits = cells + marg <= kLdsCells;
        } else {
            mode = kSinkGlobal;
            fits = false;
            // large device-memory tables get their statistic from the grid-wide
            // finalize kernels (the counted table of a parity read stays in place)
            if (!(tflags & kTFKeep)) T.flags |= kTFWide;
            const int64_t cxy = (int64_t)T.cx * T.cy;
            if (cxy <= kSliceCells && !(ctx->sink_mask & 2)) {
                const int64_t sc = (kSliceCells / cxy) * cxy;
                const int64_t ns = (cells + sc - 1) / sc;
                if (ns <= ctx->max_slices) {
                    mode = kSinkSlice;
                    T.n_slices = (int32_t)ns;
                    T.slice_cells = (int32_t)sc;
                    T.flags &= ~kTFWide;
                }
            }
        }
        const bool need_ws = (tflags & kTFKeep) || n_chunks > 1 || !fits;
        if (!need_ws) T.flags |= kTFLdsMarg;
        T.ws_off = -1;
        if (need_ws) {
            const int64_t need = (cells + marg + 63) / 64 * 64;
            if (ws_used + need > kWsMax && i1 > i0) break;
            if (need > kWsMax) return fail(ctx, FPGM_ETOOBIG, "test workspace exceeds 4 GiB");
            T.ws_off = ws_used;
            ws_used += need;
        }
        dt.push_back(T);
        mode_of.push_back(mode);
        n_items_total += n_chunks * T.n_slices;
        if (n_items_total > ((int64_t)1 << 30)) {
            ++i1;
            break;
        }
    }
    const int64_t nb = (int64_t)dt.size();
    // work items grouped by sink, one launch per group; within a group, tests keep the
    // caller's order (neighbouring tests share columns)
    int64_t grp_begin[kGroups] = {0, 0, 0, 0, 0}, grp_len[kGroups] = {0, 0, 0, 0, 0};
    HIP_TRY(S.h_items.ensure(n_items_total, ctx->stream), "alloc pinned items");
    int64_t ni = 0;
    for (int gi = 0; gi < kGroups; ++gi) {
        grp_begin[gi] = ni;
        for (int64_t k = 0; k < nb; ++k) {
            if (mode_of[k] != kGroupSink[gi]) continue;
            dt[k].item0 = (int32_t)ni;
            for (int64_t c = 0; c < (int64_t)dt[k].n_chunks * dt[k].n_slices; ++c) S.h_items.p[ni++] = (int32_t)k;
        }
        grp_len[gi] = ni - grp_begin[gi];
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
    HIP_TRY(S.h_tests.ensure(nb, ctx->stream), "alloc pinned tests");
    std::copy(dt.begin(), dt.end(), S.h_tests.p);
    HIP_TRY(S.h_zpool.ensure(std::max<int64_t>(zlen, 1), ctx->stream), "alloc pinned z pool");
    if (zlen > 0) std::copy(zpool, zpool + zlen, S.h_zpool.p);
    HIP_TRY(S.d_zpool.ensure(std::max<int64_t>(zlen, 1), false, ctx->stream), "alloc z pool");
    HIP_TRY(S.d_tests.ensure(nb, false, ctx->stream), "alloc tests");
    HIP_TRY(S.d_item_test.ensure(std::max<int64_t>(ni, 1), false, ctx->stream), "alloc items");
    HIP_TRY(ctx->d_ws.ensure(std::max<int64_t>(ws_used, kWsInit), true, ctx->stream), "alloc workspace");
    HIP_TRY(ctx->d_done.ensure(std::max<int64_t>(nb, 1024), true, ctx->stream), "alloc counters");
    HIP_TRY(S.d_res.ensure(nb, false, ctx->stream), "alloc results");
    HIP_TRY(S.h_res.ensure(nb, ctx->stream), "alloc pinned results");
    if (zlen > 0)
        HIP_TRY(hipMemcpyAsync(S.d_zpool.p, S.h_zpool.p, sizeof(int32_t) * zlen, hipMemcpyHostToDevice, ctx->stream),
                "upload z pool");
    HIP_TRY(hipMemcpyAsync(S.d_tests.p, S.h_tests.p, sizeof(DevTest) * nb, hipMemcpyHostToDevice, ctx->stream),
            "upload tests");
    HIP_TRY(hipMemcpyAsync(S.d_item_test.p, S.h_items.p, sizeof(int32_t) * ni, hipMemcpyHostToDevice, ctx->stream),
            "upload items");
    LaunchArgs a{};
    a.inst = ctx->d_inst;
    a.tests = S.d_tests.p;
    a.zpool = S.d_zpool.p;
    a.card = ctx->d_card;
    a.item_test = S.d_item_test.p;
    a.rows_per_chunk = rpc;
    a.ws = ctx->d_ws.p;
    a.done = ctx->d_done.p;
    a.crit = ctx->d_crit;
    a.alpha = alpha;
    a.z_alpha = ctx->z_alpha;
    a.band = band > 0.f ? band : 0.02f;
    a.flags = flags;
    a.res = S.d_res.p;
    ctx->t_prep += now_s() - tp0;
    HIP_TRY(hipEventRecord(S.ev0, ctx->stream), "event");
    for (int gi = 0; gi < kGroups; ++gi) {
        if (ctx->trace) HIP_TRY(hipEventRecord(S.ev_grp[gi], ctx->stream), "event");
        if (grp_len[gi] == 0) continue;
        a.item_begin = grp_begin[gi];
        a.n_items = grp_len[gi];
        const int sink = kGroupSink[gi];
        const int64_t waves_per_cu = (sink == kSinkGlobal) ? 8 : (sink == kSinkLaneW || sink == kSinkSlice) ? 2 : 4;
        int64_t grid = std::min<int64_t>(a.n_items, (int64_t)ctx->n_cu * waves_per_cu);
        grid = std::max<int64_t>(8, (grid + 7) / 8 * 8);
        HIP_TRY(launch_count_stat(a, sink, (int)grid, ctx->stream), "launch count_stat");
        ctx->k_launches += 1;
        if (sink == kSinkGlobal && n_wt > 0) {
            HIP_TRY(launch_finalize_wide(a, S.d_wide_items.p, n_wi, S.d_wide_parts.p, S.d_wide_tests.p, n_wt,
                                         ctx->stream),
                    "launch finalize");
            ctx->k_launches += 2;
        }
    }
    if (ctx->trace) HIP_TRY(hipEventRecord(S.ev_grp[kGroups], ctx->stream), "event");
    HIP_TRY(hipEventRecord(S.ev1, ctx->stream), "event");
    HIP_TRY(launch_pvalue(S.d_res.p, nb, alpha, ctx->stream), "launch pvalue");
    HIP_TRY(hipMemcpyAsync(S.h_res.p, S.d_res.p, sizeof(DevResult) * nb, hipMemcpyDeviceToHost, ctx->stream),
            "download results");
    HIP_TRY(hipEventRecord(S.ev_done, ctx->stream), "event");
    S.n = nb;
    S.bytes = 0.0;
    for (int64_t k = i0; k < i1; ++k) S.bytes += (double)(tests[k].d + 2) * (double)ctx->inst[tests[k].inst].n_rows;
    if (ctx->trace) {
        for (int gi = 0; gi < kGroups; ++gi) {
            Slot::GroupInfo &G = S.grp[gi];
            G = Slot::GroupInfo{};
            G.items = grp_len[gi];
        }
        for (int64_t k = 0; k < nb; ++k) {
            int gi = 0;
            while (kGroupSink[gi] != mode_of[k]) ++gi;
            Slot::GroupInfo &G = S.grp[gi];
            const double c = (double)dt[k].dimz * dt[k].cx * dt[k].cy;
            G.tests += 1;
            G.dsum += dt[k].d;
            G.csum += c;
            G.cmax = std::max(G.cmax, c);
            G.bytes += (double)(dt[k].d + 2) * (double)ctx->inst[dt[k].inst].n_rows;
        }
    }
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
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, S.ev0, S.ev1), "event time");
    ctx->k_ms += ms;
    ctx->k_bytes += S.bytes;
    if (ctx->trace) {
        static const char *names[kGroups] = {"lane", "lane-wide", "shared", "slice", "device"};
        for (int gi = 0; gi < kGroups; ++gi) {
            const Slot::GroupInfo &G = S.grp[gi];
            if (G.items == 0) continue;
            float gms = 0.f;
            HIP_TRY(hipEventElapsedTime(&gms, S.ev_grp[gi], S.ev_grp[gi + 1]), "event time");
            fprintf(stderr,
                    "[fpgm]   launch %-9s tests %7lld items %8lld mean |Z| %.2f cells mean %.0f max %.0f: "
                    "%8.3f ms, %7.0f GB/s\n",
                    names[gi], (long long)G.tests, (long long)G.items, (double)G.dsum / G.tests, G.csum / G.tests,
                    G.cmax, gms, gms > 0 ? G.bytes / (gms * 1e6) : 0.0);
        }
    }
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
    ctx->k_launches = 0;
    ctx->g_ms = 0.0;
    ctx->g_ops = 0.0;
    ctx->g_pair_ms = 0.0;
    ctx->g_launches = 0;
}

// Level 0 of one instance on the MFMA units: C = O^T O, then one thread per pair
// (pairs in tv order, all of instance `inst`).  Returns false in *used when the
// one-hot width is too large for a dense C (the caller then counts pair by pair).
// Results land in the pinned buffer of slot 0; *out points at them.
fpgm_status run_level0_gram(fpgm_ctx *ctx, int inst, const std::vector<HostTest> &tv, double alpha, uint32_t flags,
                            float band, const DevResult **out, bool *used)
{
    *used = false;
    Slot &S = ctx->slot[0];
    {
        fpgm_status st = batch_wait(ctx, S);
        if (st != FPGM_OK) return st;
    }
    const int n = ctx->n_vars;
    std::vector<int32_t> off(n + 1, 0);
    for (int v = 0; v < n; ++v) off[v + 1] = off[v] + ctx->card[v];
    const int64_t W = off[n];
    const int64_t Wp = std::max<int64_t>(kGramTile, (W + kGramTile - 1) / kGramTile * kGramTile);
    if (Wp > 32768) return FPGM_OK;
    *used = true;
    fpgm_status st = ensure_crit(ctx, alpha);
    if (st != FPGM_OK) return st;
    const InstHost &I = ctx->inst[inst];
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
    const int64_t ksteps = std::max<int64_t>(1, (I.n_rows + 31) / 32);
    const int64_t target = (int64_t)ctx->n_cu * 8;
    const int64_t splits = std::min<int64_t>(ksteps, std::max<int64_t>(1, (target + n_tp - 1) / n_tp));
    const int64_t kps = (ksteps + splits - 1) / splits * 32;
    const int64_t n_items = n_tp * splits;
    const int64_t np = (int64_t)tv.size();
    std::vector<int32_t> pairs((size_t)2 * np);
    for (int64_t k = 0; k < np; ++k) {
        pairs[k] = tv[k].x;
        pairs[np + k] = tv[k].y;
    }
    HIP_TRY(ctx->d_C.ensure(Wp * Wp, false, ctx->stream), "alloc gram");
    HIP_TRY(ctx->d_rowvar.ensure(Wp, false, ctx->stream), "alloc gram rows");
    HIP_TRY(ctx->d_rowstate.ensure(Wp, false, ctx->stream), "alloc gram rows");
    HIP_TRY(ctx->d_tiles.ensure((int64_t)tiles.size(), false, ctx->stream), "alloc gram tiles");
    HIP_TRY(ctx->d_off.ensure(n + 1, false, ctx->stream), "alloc gram offsets");
    HIP_TRY(ctx->d_pairs.ensure(2 * np, false, ctx->stream), "alloc pairs");
    HIP_TRY(S.d_res.ensure(np, false, ctx->stream), "alloc results");
    HIP_TRY(S.h_res.ensure(np, ctx->stream), "alloc pinned results");
    HIP_TRY(hipMemcpyAsync(ctx->d_rowvar.p, rowvar.data(), sizeof(int32_t) * Wp, hipMemcpyHostToDevice, ctx->stream),
            "upload gram rows");
    HIP_TRY(hipMemcpyAsync(ctx->d_rowstate.p, rowstate.data(), Wp, hipMemcpyHostToDevice, ctx->stream),
            "upload gram rows");
    HIP_TRY(hipMemcpyAsync(ctx->d_tiles.p, tiles.data(), sizeof(int32_t) * tiles.size(), hipMemcpyHostToDevice,
                           ctx->stream),
            "upload gram tiles");
    HIP_TRY(hipMemcpyAsync(ctx->d_off.p, off.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, ctx->stream),
            "upload offsets");
    HIP_TRY(hipMemcpyAsync(ctx->d_pairs.p, pairs.data(), sizeof(int32_t) * 2 * np, hipMemcpyHostToDevice, ctx->stream),
            "upload pairs");
    HIP_TRY(hipMemsetAsync(ctx->d_C.p, 0, sizeof(int32_t) * Wp * Wp, ctx->stream), "clear gram");
    GramArgs g{};
    g.cols = I.cols;
    g.stride = I.stride;
    g.n_rows = I.n_rows;
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
    p.px = ctx->d_pairs.p;
    p.py = ctx->d_pairs.p + np;
    p.n_pairs = np;
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
    HIP_TRY(launch_pvalue(S.d_res.p, np, alpha, ctx->stream), "launch pvalue");
    *out = S.h_res.p;
    HIP_TRY(hipMemcpyAsync(S.h_res.p, S.d_res.p, sizeof(DevResult) * np, hipMemcpyDeviceToHost, ctx->stream),
            "download results");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "synchronize");
    float ms_g = 0.f, ms_p = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms_g, ctx->ev0, e_mid), "event time");
    HIP_TRY(hipEventElapsedTime(&ms_p, e_mid, ctx->ev1), "event time");
    (void)hipEventDestroy(e_mid);
    ctx->g_ms += ms_g;
    ctx->g_pair_ms += ms_p;
    ctx->g_ops += 2.0 * (double)n_tp * kGramTile * kGramTile * (double)I.n_rows;
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
    if (const char *e = std::getenv("FPGM_ROUND_BASE")) ctx->sched.base = std::max(1.0, std::atof(e));
    if (const char *e = std::getenv("FPGM_ROUND_CAP")) ctx->sched.cap = std::max<int64_t>(1, std::atoll(e));
    if (const char *e = std::getenv("FPGM_ROUND_PROMOTE")) ctx->sched.promote = std::max<int64_t>(0, std::atoll(e));
    if (const char *e = std::getenv("FPGM_LANES")) ctx->sched.lanes = std::atoi(e) == 2 ? 2 : 1;
    if (const char *e = std::getenv("FPGM_MAX_SLICES")) ctx->max_slices = std::max<int64_t>(1, std::atoll(e));
    if (const char *e = std::getenv("FPGM_SINKS"))
        if (std::strcmp(e, "shared") == 0) ctx->sink_mask |= 3;
    HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
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

// Uploads the instance table and scans every column for values >= card.
static fpgm_status finish_ctx(fpgm_ctx *ctx)
{
    std::vector<DevInst> di(ctx->inst.size());
    for (size_t i = 0; i < di.size(); ++i) di[i] = DevInst{ctx->inst[i].cols, ctx->inst[i].stride, ctx->inst[i].n_rows};
    HIP_TRY(hipMalloc((void **)&ctx->d_inst, sizeof(DevInst) * di.size()), "hipMalloc instances");
    HIP_TRY(hipMemcpyAsync(ctx->d_inst, di.data(), sizeof(DevInst) * di.size(), hipMemcpyHostToDevice, ctx->stream),
            "upload instances");
    HIP_TRY(ctx->d_flag.ensure(1, true, ctx->stream), "alloc flag");
    HIP_TRY(hipMemsetAsync(ctx->d_flag.p, 0, sizeof(uint32_t), ctx->stream), "memset flag");
    for (const InstHost &I : ctx->inst)
        HIP_TRY(launch_check_values(I.cols, I.stride, I.n_rows, ctx->n_vars, ctx->d_card, ctx->d_flag.p, ctx->stream),
                "launch check");
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
            e = hipMemsetAsync(d, 0, (size_t)off[n_inst], ctx->stream);
            for (int32_t i = 0; i < n_inst && e == hipSuccess; ++i) {
                const int64_t stride = (off[i + 1] - off[i]) / n_vars;
                ctx->inst.push_back(InstHost{d + off[i], stride, n_rows[i]});
                if (n_rows[i] > 0)
                    e = hipMemcpy2DAsync(d + off[i], (size_t)stride, cols[i], (size_t)row_stride[i], (size_t)n_rows[i],
                                         (size_t)n_vars, hipMemcpyHostToDevice, ctx->stream);
            }
            if (e != hipSuccess) st = hip_fail(ctx, e, "upload columns");
        }
    }
    if (st == FPGM_OK) st = finish_ctx(ctx);
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
    if (ctx->d_crit) (void)hipFree(ctx->d_crit);
    for (Slot &sl : ctx->slot) sl.release();
    ctx->d_ws.release();
    ctx->d_done.release();
    ctx->d_flag.release();
    ctx->d_C.release();
    ctx->d_rowvar.release();
    ctx->d_rowstate.release();
    ctx->d_tiles.release();
    ctx->d_off.release();
    ctx->d_pairs.release();
    ctx->d_mle_nodes.release();
    ctx->d_mle_rel.release();
    ctx->d_mle_items.release();
    ctx->d_mle_out.release();
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
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
    std::vector<Sep> seps;
    std::vector<int32_t> sepz;
    int64_t num_ci_test = 0;
    bool cont = true;               // continues to the next level
    Frozen F;
    std::vector<EdgeWalk> walks;
    std::vector<int32_t> choice, zlevel;
};

// The CI stage of every instance, level-synchronously: level d's tests of all
// instances that reached level d share the same rounds and launches.
fpgm_status run_pc(fpgm_ctx *ctx, double alpha, int32_t depth, const fpgm_pc_opts &o, std::vector<PcInst> &P,
                   fpgm_stats &S)
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
        I.adj.assign((size_t)n * n, 0);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) I.adj[(size_t)i * n + j] = (i != j);
        I.ex.clear();
        I.ey.clear();
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) {
                I.ex.push_back(i);
                I.ey.push_back(j);
            }
        I.seps.clear();
        I.sepz.clear();
        I.num_ci_test = 0;
        I.cont = true;
    }
    std::vector<DevResult> res;
    std::vector<HostTest> tv;

    // ---- level 0: every pair of the complete graph (PCStable.cpp:133-195)
    {
        const double tl = now_s();
        tv.clear();
        for (int32_t i = 0; i < n_inst; ++i)
            for (size_t e = 0; e < P[i].ex.size(); ++e) tv.push_back(HostTest{P[i].ex[e], P[i].ey[e], 0, i, 0});
        bool gram = false;
        const DevResult *r0 = nullptr;
        if (!plain && n_inst == 1 && !(o.flags & FPGM_F_NO_GRAM)) {
            const double ms0 = ctx->g_ms + ctx->g_pair_ms;
            st = run_level0_gram(ctx, 0, tv, alpha, o.flags, band, &r0, &gram);
            if (st != FPGM_OK) return st;
            S.kernel_seconds_per_level[0] += (ctx->g_ms + ctx->g_pair_ms - ms0) * 1e-3;
        }
        if (!gram) {
            st = run_tests(ctx, plain, o, 0, tv, nullptr, 0, alpha, band, res, S);
            if (st != FPGM_OK) return st;
            r0 = res.data();
        }
        const DevResult *res0 = r0;
        int64_t k = 0, inb = 0, left = 0;
        for (int32_t i = 0; i < n_inst; ++i) {
            PcInst &I = P[i];
            int64_t w = 0;
            for (size_t e = 0; e < I.ex.size(); ++e, ++k) {
                inb += res0[k].in_band;
                if (res0[k].exact && res0[k].df > 0 && res0[k].g2_64 > 0.0)
                    S.max_rel_err_g2_fp32 =
                        std::max(S.max_rel_err_g2_fp32, std::fabs((double)res0[k].g2_32 - res0[k].g2_64) / res0[k].g2_64);
                if (res0[k].indep) {
                    I.seps.push_back(PcInst::Sep{I.ex[e], I.ey[e], 0, 0, (int64_t)I.sepz.size()});
                    I.adj[(size_t)I.ex[e] * n + I.ey[e]] = 0;
                    I.adj[(size_t)I.ey[e] * n + I.ex[e]] = 0;
                } else {
                    I.ex[w] = I.ex[e];
                    I.ey[w] = I.ey[e];
                    ++w;
                }
            }
            I.num_ci_test += (int64_t)I.ex.size();
            I.ex.resize(w);
            I.ey.resize(w);
            left += w;
        }
        S.tests_per_level[0] = (int64_t)tv.size();
        S.executed_per_level[0] = (int64_t)tv.size();
        S.in_band_per_level[0] = inb;
        S.rounds_per_level[0] = 1;
        S.edges_after_level[0] = left;
        S.seconds_per_level[0] = now_s() - tl;
        S.n_levels = 1;
        log_level(o, 0, (int64_t)tv.size(), left, now_s() - t_start);
    }

    // ---- levels d >= 1 (PCStable.cpp:209-231, SearchAtDepth :265-388)
    for (int d = 1; d < depth && d < 64; ++d) {
        bool any = false;
        for (const PcInst &I : P) any = any || I.cont;
        if (!any) break;
        if (d > kMaxD) return fail(ctx, FPGM_ETOOBIG, "conditioning depth exceeds the device limit");
        const double tl = now_s();
        const double prep0 = ctx->t_prep, wait0 = ctx->t_wait;
        std::vector<std::pair<int32_t, int64_t>> active;  // (instance, edge)
        for (int32_t i = 0; i < n_inst; ++i) {
            PcInst &I = P[i];
            if (!I.cont) continue;
            // frozen copy of the adjacency (PCStable.cpp:272)
            I.F.off.assign(n + 1, 0);
            I.F.nb.clear();
            for (int a = 0; a < n; ++a) {
                for (int b = 0; b < n; ++b)
                    if (I.adj[(size_t)a * n + b]) I.F.nb.push_back(b);
                I.F.off[a + 1] = (int32_t)I.F.nb.size();
            }
            const int64_t E = (int64_t)I.ex.size();
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
        int64_t executed = 0, inb = 0, rounds = 0;
        const int64_t budget =
            o.round_tests > 0
                ? o.round_tests
                : std::min<int64_t>((int64_t)1 << 20,
                                    std::max<int64_t>(8192, (int64_t)(2147483648.0 / std::max(1.0, (d + 2) * avg_rows))));
        // With two lanes, the edges take turns: while the device counts one lane's round,
        // the host takes the other lane's results and builds its next round.  Edges are
        // independent within a level (frozen adjacency), so the split changes timing and
        // speculation only, never results.  One lane is the default: at 10^6 rows the
        // half-size launches of two lanes cost the device more than the overlap saves.
        // Plain mode runs one lane synchronously.
        struct Lane {
            std::vector<std::pair<int32_t, int64_t>> active;  // (instance, edge)
            std::vector<HostTest> tv;
            std::vector<int32_t> zpool;
            std::vector<int64_t> t_edge;  // active index of each test of the round
            std::vector<DevResult> res;   // results of rounds not run asynchronously
            const DevResult *r = nullptr;
            int64_t rounds = 0;
            int slot = 0;
            bool in_flight = false, async = false;
        };
        const int n_lanes = plain ? 1 : ctx->sched.lanes;
        Lane lanes[2];
        for (size_t ai = 0; ai < active.size(); ++ai) lanes[ai % n_lanes].active.push_back(active[ai]);
        for (int l = 0; l < n_lanes; ++l) lanes[l].slot = l;
        const int64_t lane_budget = std::max<int64_t>(1, budget / n_lanes);
        const double kms_level0 = ctx->k_ms;

        auto start = [&](Lane &ln) -> fpgm_status {
            if (o.max_rounds > 0 && ln.rounds >= o.max_rounds)
                return fail(ctx, FPGM_ETIMEOUT, "level exceeded the round limit");
            ++ln.rounds;
            const bool plain_round = plain && (o.plain_sample <= 0 || S.plain_tests_per_level[d] < o.plain_sample);
            // tests per edge this round: the budget shared by the active edges, at least a
            // geometric ramp (base^(round-1), capped) so long walks finish in few rounds
            int64_t W = plain_round ? 1 : std::max<int64_t>(1, lane_budget / (int64_t)ln.active.size());
            if (!plain_round) {
                const double ramp = std::pow(ctx->sched.base, (double)(ln.rounds - 1));
                W = std::max<int64_t>(W, (int64_t)std::min<double>(ramp, (double)ctx->sched.cap));
            }
            ln.tv.clear();
            ln.zpool.clear();
            ln.t_edge.clear();
            for (size_t ai = 0; ai < ln.active.size(); ++ai) {
                PcInst &I = P[ln.active[ai].first];
                const int64_t e = ln.active[ai].second;
                EdgeWalk &w = I.walks[e];
                int32_t *c = &I.choice[(size_t)e * d];
                // an edge found dependent on its first `promote` sets rarely ends early:
                // give it the rest of its walk at once
                int64_t We = W;
                if (!plain_round && ctx->sched.promote > 0 && w.tests >= ctx->sched.promote) We = INT64_MAX;
                for (int64_t k = 0; k < We && w.side < 2; ++k) {
                    ln.tv.push_back(HostTest{w.x, w.y, d, ln.active[ai].first, (int64_t)ln.zpool.size()});
                    for (int i = 0; i < d; ++i) ln.zpool.push_back(I.F.at(w.u, w.skip, c[i]));
                    ln.t_edge.push_back((int64_t)ai);
                    if (!next_choice(w.a, d, c)) start_side(w, I.F, d, w.side + 1, c);
                }
            }
            executed += (int64_t)ln.tv.size();
            ln.in_flight = true;
            ln.async = false;
            const int64_t nt = (int64_t)ln.tv.size(), nz = (int64_t)ln.zpool.size();
            if (plain) {
                fpgm_status st2 = run_tests(ctx, plain_round, o, d, ln.tv, ln.zpool.data(), nz, alpha, band, ln.res, S);
                ln.r = ln.res.data();
                return st2;
            }
            Slot &sl = ctx->slot[ln.slot];
            int64_t i1 = 0;
            fpgm_status st2 =
                batch_submit(ctx, sl, ln.tv.data(), nt, ln.zpool.data(), nz, alpha, o.flags, band, 0, 0, 0, &i1, nullptr);
            if (st2 != FPGM_OK) return st2;
            if (i1 == nt) {
                ln.async = true;
                return FPGM_OK;
            }
            // more than one launch sequence (workspace limit): finish the round here
            ln.res.resize((size_t)nt);
            for (int64_t i0 = 0;;) {
                st2 = batch_wait(ctx, sl);
                if (st2 != FPGM_OK) return st2;
                std::copy(sl.h_res.p, sl.h_res.p + (i1 - i0), ln.res.begin() + i0);
                i0 = i1;
                if (i0 >= nt) break;
                st2 = batch_submit(ctx, sl, ln.tv.data(), nt, ln.zpool.data(), nz, alpha, o.flags, band, 0, 0, i0, &i1,
                                   nullptr);
                if (st2 != FPGM_OK) return st2;
            }
            ln.r = ln.res.data();
            return FPGM_OK;
        };

        auto finish = [&](Lane &ln) -> fpgm_status {
            if (ln.async) {
                fpgm_status st2 = batch_wait(ctx, ctx->slot[ln.slot]);
                if (st2 != FPGM_OK) return st2;
                ln.r = ctx->slot[ln.slot].h_res.p;
            }
            const DevResult *r = ln.r;
            std::vector<std::pair<int32_t, int64_t>> still;
            size_t k = 0;
            for (size_t ai = 0; ai < ln.active.size(); ++ai) {
                PcInst &I = P[ln.active[ai].first];
                EdgeWalk &w = I.walks[ln.active[ai].second];
                bool done = false;
                for (; k < ln.tv.size() && ln.t_edge[k] == (int64_t)ai; ++k) {
                    inb += r[k].in_band;
                    if (r[k].exact && r[k].df > 0 && r[k].g2_64 > 0.0)
                        S.max_rel_err_g2_fp32 =
                            std::max(S.max_rel_err_g2_fp32, std::fabs((double)r[k].g2_32 - r[k].g2_64) / r[k].g2_64);
                    if (done) continue;
                    w.tests += 1;
                    if (r[k].indep) {
                        done = true;
                        w.removed = 1;
                        w.zsep = (int64_t)I.zlevel.size();
                        for (int i = 0; i < d; ++i) I.zlevel.push_back(ln.zpool[ln.tv[k].z_off + i]);
                    }
                }
                if (!done && w.side < 2) still.push_back(ln.active[ai]);
            }
            ln.active.swap(still);
            ln.in_flight = false;
            return FPGM_OK;
        };

        for (int l = 0; l < n_lanes; ++l)
            if (!lanes[l].active.empty()) {
                st = start(lanes[l]);
                if (st != FPGM_OK) return st;
            }
        for (bool any = true; any;) {
            any = false;
            for (int l = 0; l < n_lanes; ++l) {
                Lane &ln = lanes[l];
                if (!ln.in_flight) continue;
                any = true;
                st = finish(ln);
                if (st != FPGM_OK) return st;
                if (!ln.active.empty()) {
                    st = start(ln);
                    if (st != FPGM_OK) return st;
                }
            }
        }
        rounds = std::max(lanes[0].rounds, lanes[1].rounds);
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
                    I.adj[(size_t)w.x * n + w.y] = 0;
                    I.adj[(size_t)w.y * n + w.x] = 0;
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
            int maxdeg = 0;
            for (int a = 0; a < n; ++a) {
                int deg = 0;
                for (int b = 0; b < n; ++b) deg += I.adj[(size_t)a * n + b];
                maxdeg = std::max(maxdeg, deg);
            }
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
            fprintf(stderr, "[fpgm]   level %d host: %.4f s building launches, %.4f s launching and waiting, "
                    "%.4f s level total\n", d, ctx->t_prep - prep0, ctx->t_wait - wait0, now_s() - tl);
    }
    S.ci_seconds = now_s() - t_start;
    for (int l = 0; l < S.n_levels; ++l) S.num_ci_test += S.tests_per_level[l];
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
    std::sort(I.seps.begin(), I.seps.end(),
              [](const PcInst::Sep &a, const PcInst::Sep &b) { return a.x != b.x ? a.x < b.x : a.y < b.y; });
    ctx->seps.clear();
    ctx->sep_z.clear();
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

fpgm_status fpgm_last_kernel_stats(const fpgm_ctx *ctx, double *kernel_ms, double *alg_bytes, int64_t *launches)
{
    if (!ctx) return FPGM_EINVAL;
    if (kernel_ms) *kernel_ms = ctx->k_ms;
    if (alg_bytes) *alg_bytes = ctx->k_bytes;
    if (launches) *launches = ctx->k_launches;
    return FPGM_OK;
}

}  // extern "C"
