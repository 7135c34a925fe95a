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

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
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

struct HostTest {
    int32_t x, y, d;
    int64_t z_off;  // into the host z pool
};

}  // namespace

struct fpgm_ctx {
    int dev = 0;
    hipStream_t stream = nullptr;
    const uint8_t *d_cols = nullptr;
    bool owns_cols = false;
    int64_t stride = 0, n_rows = 0;
    int32_t n_vars = 0;
    std::vector<int32_t> card;
    int32_t *d_card = nullptr;
    std::string err;
    std::mutex mu;
    int n_cu = 256;

    double crit_alpha = -1.0;
    double z_alpha = 0.0;
    std::vector<double> crit_host;
    double *d_crit = nullptr;

    DevBuf<DevTest> d_tests;
    DevBuf<int32_t> d_zpool, d_item_test, d_ws;
    DevBuf<uint32_t> d_done, d_flag;
    DevBuf<DevResult> d_res;

    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double k_ms = 0.0, k_bytes = 0.0;
    int64_t k_launches = 0;

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

// Runs tests[0..n) on the device; results land in out (host).  flags: FPGM_F_*,
// tflags: extra kTF* bits for every test.  keep_ws receives the workspace offset
// of test 0 when kTFKeep is set.
fpgm_status run_batch(fpgm_ctx *ctx, const HostTest *tests, int64_t n, const int32_t *zpool, int64_t zlen,
                      double alpha, uint32_t flags, float band, int32_t tflags, std::vector<DevResult> &out,
                      int64_t rows_per_chunk_forced = 0, int64_t *keep_ws = nullptr)
{
    out.resize((size_t)n);
    if (n == 0) return FPGM_OK;
    fpgm_status st = ensure_crit(ctx, alpha);
    if (st != FPGM_OK) return st;
    const int64_t N = ctx->n_rows;
    const int64_t steps = std::max<int64_t>(1, (N + kRowsPerStep - 1) / kRowsPerStep);

    // rows per chunk, common to the launch: enough work items to fill every CU
    const int64_t target_items = (int64_t)ctx->n_cu * 16;
    int64_t rpc;
    if (rows_per_chunk_forced > 0) {
        rpc = (rows_per_chunk_forced + kRowsPerStep - 1) / kRowsPerStep * kRowsPerStep;
    } else {
        int64_t cpt = std::max<int64_t>(1, (target_items + n - 1) / n);
        cpt = std::min(cpt, steps);
        rpc = (steps + cpt - 1) / cpt * kRowsPerStep;
    }
    rpc = std::min<int64_t>(rpc, kLaneMaxRowsPerChunk);
    const int64_t n_chunks = std::max<int64_t>(1, (N + rpc - 1) / rpc);

    HIP_TRY(ctx->d_zpool.ensure(std::max<int64_t>(zlen, 1), false, ctx->stream), "alloc z pool");
    if (zlen > 0)
        HIP_TRY(hipMemcpyAsync(ctx->d_zpool.p, zpool, sizeof(int32_t) * zlen, hipMemcpyHostToDevice, ctx->stream),
                "upload z pool");

    std::vector<DevTest> dt;
    std::vector<int32_t> item_test, mode_of;
    int64_t i0 = 0;
    while (i0 < n) {
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
            T.z_off = (int32_t)h.z_off;
            T.cx = ctx->card[h.x];
            T.cy = ctx->card[h.y];
            int64_t dimz = 1;
            for (int j = 0; j < h.d; ++j) {
                dimz *= ctx->card[zpool[h.z_off + j]];
                if (dimz > kMaxCells) break;
            }
            const int64_t cells = dimz * T.cx * T.cy;
            if (dimz > kMaxCells || cells >= kMaxCells)
                return fail(ctx, FPGM_ETOOBIG, "contingency table exceeds 2^31 cells");
            T.dimz = (int32_t)dimz;
            T.n_chunks = (int32_t)n_chunks;
            const int64_t marg = dimz * (T.cx + T.cy + 1);
            T.flags = tflags;
            bool fits;
            int mode;
            if (cells <= kLaneCells) {
                mode = kSinkLane;
                T.flags |= kTFLane;
                fits = cells + marg <= kSmallCells;
            } else if (cells <= kLdsCells) {
                mode = kSinkShared;
                T.flags |= kTFLdsTable;
                fits = cells + marg <= kLdsCells;
            } else {
                mode = kSinkGlobal;
                fits = false;
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
            n_items_total += n_chunks;
            if (n_items_total > ((int64_t)1 << 30)) {
                ++i1;
                break;
            }
        }
        const int64_t nb = (int64_t)dt.size();
        // work items grouped by sink: [lane | shared | global], one launch per group
        int64_t grp_begin[3] = {0, 0, 0}, grp_len[3] = {0, 0, 0};
        const int order[3] = {kSinkLane, kSinkShared, kSinkGlobal};
        item_test.clear();
        item_test.reserve((size_t)n_items_total);
        for (int gi = 0; gi < 3; ++gi) {
            grp_begin[gi] = (int64_t)item_test.size();
            for (int64_t k = 0; k < nb; ++k) {
                if (mode_of[k] != order[gi]) continue;
                dt[k].item0 = (int32_t)item_test.size();
                for (int64_t c = 0; c < n_chunks; ++c) item_test.push_back((int32_t)k);
            }
            grp_len[gi] = (int64_t)item_test.size() - grp_begin[gi];
        }
        HIP_TRY(ctx->d_tests.ensure(nb, false, ctx->stream), "alloc tests");
        HIP_TRY(ctx->d_item_test.ensure((int64_t)item_test.size(), false, ctx->stream), "alloc items");
        HIP_TRY(ctx->d_ws.ensure(std::max<int64_t>(ws_used, kWsInit), true, ctx->stream), "alloc workspace");
        HIP_TRY(ctx->d_done.ensure(nb, true, ctx->stream), "alloc counters");
        HIP_TRY(ctx->d_res.ensure(nb, false, ctx->stream), "alloc results");
        HIP_TRY(hipMemcpyAsync(ctx->d_tests.p, dt.data(), sizeof(DevTest) * nb, hipMemcpyHostToDevice, ctx->stream),
                "upload tests");
        HIP_TRY(hipMemcpyAsync(ctx->d_item_test.p, item_test.data(), sizeof(int32_t) * item_test.size(),
                               hipMemcpyHostToDevice, ctx->stream),
                "upload items");
        LaunchArgs a{};
        a.cols = ctx->d_cols;
        a.stride = ctx->stride;
        a.n_rows = N;
        a.tests = ctx->d_tests.p;
        a.zpool = ctx->d_zpool.p;
        a.card = ctx->d_card;
        a.item_test = ctx->d_item_test.p;
        a.rows_per_chunk = rpc;
        a.ws = ctx->d_ws.p;
        a.done = ctx->d_done.p;
        a.crit = ctx->d_crit;
        a.alpha = alpha;
        a.z_alpha = ctx->z_alpha;
        a.band = band > 0.f ? band : 0.02f;
        a.flags = flags;
        a.res = ctx->d_res.p;
        HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream), "event");
        for (int gi = 0; gi < 3; ++gi) {
            if (grp_len[gi] == 0) continue;
            a.item_begin = grp_begin[gi];
            a.n_items = grp_len[gi];
            const int64_t waves_per_cu = (order[gi] == kSinkGlobal) ? 8 : 4;
            int64_t grid = std::min<int64_t>(a.n_items, (int64_t)ctx->n_cu * waves_per_cu);
            grid = std::max<int64_t>(8, (grid + 7) / 8 * 8);
            HIP_TRY(launch_count_stat(a, order[gi], (int)grid, ctx->stream), "launch count_stat");
            ctx->k_launches += 1;
        }
        HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream), "event");
        HIP_TRY(launch_pvalue(ctx->d_res.p, nb, alpha, ctx->stream), "launch pvalue");
        HIP_TRY(hipMemcpyAsync(out.data() + i0, ctx->d_res.p, sizeof(DevResult) * nb, hipMemcpyDeviceToHost,
                               ctx->stream),
                "download results");
        HIP_TRY(hipStreamSynchronize(ctx->stream), "synchronize");
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1), "event time");
        ctx->k_ms += ms;
        for (int64_t k = i0; k < i1; ++k) ctx->k_bytes += (double)(tests[k].d + 2) * (double)N;
        if (keep_ws) *keep_ws = dt[0].ws_off;
        i0 = i1;
    }
    return FPGM_OK;
}

void reset_kernel_stats(fpgm_ctx *ctx)
{
    ctx->k_ms = 0.0;
    ctx->k_bytes = 0.0;
    ctx->k_launches = 0;
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

static fpgm_status ctx_init(fpgm_ctx *ctx, int32_t n_vars, int64_t n_rows, int64_t row_stride, const int32_t *card)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(ctx, FPGM_EDEVICE, std::string("device is not gfx950: ") + prop.gcnArchName);
    ctx->dev = dev;
    ctx->n_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
    HIP_TRY(hipEventCreate(&ctx->ev0), "hipEventCreate");
    HIP_TRY(hipEventCreate(&ctx->ev1), "hipEventCreate");
    ctx->n_vars = n_vars;
    ctx->n_rows = n_rows;
    ctx->stride = row_stride;
    ctx->card.assign(card, card + n_vars);
    HIP_TRY(hipMalloc((void **)&ctx->d_card, sizeof(int32_t) * std::max(1, n_vars)), "hipMalloc card");
    HIP_TRY(hipMemcpyAsync(ctx->d_card, card, sizeof(int32_t) * n_vars, hipMemcpyHostToDevice, ctx->stream),
            "upload card");
    return FPGM_OK;
}

static fpgm_status check_args(int32_t n_vars, int64_t n_rows, int64_t row_stride, const int32_t *card,
                              std::string &msg)
{
    if (n_vars < 2 || n_rows < 0 || row_stride < n_rows || row_stride % 16 != 0 || !card) {
        msg = "need n_vars >= 2, 0 <= n_rows <= row_stride, row_stride % 16 == 0, card";
        return FPGM_EINVAL;
    }
    if (n_rows > ((int64_t)1 << 31) - 1) {
        msg = "n_rows must fit int32 counts";
        return FPGM_EINVAL;
    }
    for (int v = 0; v < n_vars; ++v)
        if (card[v] < 1 || card[v] > 255) {
            msg = "cardinalities must be in 1..255";
            return FPGM_EINVAL;
        }
    return FPGM_OK;
}

static fpgm_status validate_values(fpgm_ctx *ctx)
{
    HIP_TRY(ctx->d_flag.ensure(1, true, ctx->stream), "alloc flag");
    HIP_TRY(hipMemsetAsync(ctx->d_flag.p, 0, sizeof(uint32_t), ctx->stream), "memset flag");
    HIP_TRY(launch_check_values(ctx->d_cols, ctx->stride, ctx->n_rows, ctx->n_vars, ctx->d_card, ctx->d_flag.p,
                                ctx->stream),
            "launch check");
    uint32_t bad = 0;
    HIP_TRY(hipMemcpyAsync(&bad, ctx->d_flag.p, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream), "flag");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "synchronize");
    if (bad) return fail(ctx, FPGM_EINVAL, "a column holds a value >= its cardinality");
    return FPGM_OK;
}

fpgm_status fpgm_ctx_create(const uint8_t *cols, int32_t n_vars, int64_t n_rows, int64_t row_stride,
                            const int32_t *card, fpgm_ctx **out)
{
    if (!out) return FPGM_EINVAL;
    *out = nullptr;
    std::string msg;
    if (!cols || row_stride < n_rows || n_vars < 2) return FPGM_EINVAL;
    if (check_args(n_vars, n_rows, (row_stride + 255) / 256 * 256, card, msg) != FPGM_OK) return FPGM_EINVAL;
    fpgm_ctx *ctx = new (std::nothrow) fpgm_ctx();
    if (!ctx) return FPGM_ENOMEM;
    fpgm_status st = ctx_init(ctx, n_vars, n_rows, std::max<int64_t>(256, (n_rows + 255) / 256 * 256), card);
    if (st == FPGM_OK) {
        uint8_t *d = nullptr;
        const size_t bytes = (size_t)ctx->stride * n_vars;
        hipError_t e = hipMalloc((void **)&d, bytes);
        if (e != hipSuccess) {
            st = hip_fail(ctx, e, "hipMalloc columns");
        } else {
            ctx->d_cols = d;
            ctx->owns_cols = true;
            e = hipMemsetAsync(d, 0, bytes, ctx->stream);
            if (e == hipSuccess && n_rows > 0)
                e = hipMemcpy2DAsync(d, (size_t)ctx->stride, cols, (size_t)row_stride, (size_t)n_rows, (size_t)n_vars,
                                     hipMemcpyHostToDevice, ctx->stream);
            if (e != hipSuccess) st = hip_fail(ctx, e, "upload columns");
        }
    }
    if (st == FPGM_OK) st = validate_values(ctx);
    if (st != FPGM_OK) {
        fpgm_ctx_destroy(ctx);
        return st;
    }
    *out = ctx;
    return FPGM_OK;
}

fpgm_status fpgm_ctx_create_device(const void *d_cols, int32_t n_vars, int64_t n_rows, int64_t row_stride,
                                   const int32_t *card, fpgm_ctx **out)
{
    if (!out) return FPGM_EINVAL;
    *out = nullptr;
    std::string msg;
    if (!d_cols || ((uintptr_t)d_cols % 16) != 0) return FPGM_EINVAL;
    if (check_args(n_vars, n_rows, row_stride, card, msg) != FPGM_OK) return FPGM_EINVAL;
    fpgm_ctx *ctx = new (std::nothrow) fpgm_ctx();
    if (!ctx) return FPGM_ENOMEM;
    fpgm_status st = ctx_init(ctx, n_vars, n_rows, row_stride, card);
    if (st == FPGM_OK) {
        ctx->d_cols = (const uint8_t *)d_cols;
        ctx->owns_cols = false;
        st = validate_values(ctx);
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
    if (ctx->d_card) (void)hipFree(ctx->d_card);
    if (ctx->d_crit) (void)hipFree(ctx->d_crit);
    ctx->d_tests.release();
    ctx->d_zpool.release();
    ctx->d_item_test.release();
    ctx->d_ws.release();
    ctx->d_done.release();
    ctx->d_flag.release();
    ctx->d_res.release();
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

static fpgm_status check_test(fpgm_ctx *ctx, int32_t x, int32_t y, const int32_t *z, int32_t d, double alpha)
{
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
    HostTest h{x, y, d, 0};
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
        h[i] = HostTest{t.x, t.y, t.d, t.z_off};
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
    HostTest h{x, y, d, 0};
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

fpgm_status fpgm_pc_skeleton(fpgm_ctx *ctx, double alpha, int32_t depth, const fpgm_pc_opts *opts,
                             uint8_t *adj_out, fpgm_stats *stats)
{
    if (!ctx || !adj_out) return FPGM_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (!(alpha >= 0.0 && alpha <= 1.0)) return fail(ctx, FPGM_EINVAL, "alpha must lie in [0, 1]");
    if (depth < 0) return fail(ctx, FPGM_EINVAL, "depth must be >= 0");
    fpgm_pc_opts o{};
    if (opts) o = *opts;
    const float band = o.band > 0.f ? o.band : 0.02f;
    const bool plain = (o.mode == FPGM_MODE_PLAIN);
    const int n = ctx->n_vars;
    const int64_t N = ctx->n_rows;
    fpgm_stats S{};
    reset_kernel_stats(ctx);
    fpgm_status st = ensure_crit(ctx, alpha);
    if (st != FPGM_OK) return st;
    HIP_TRY(hipStreamSynchronize(ctx->stream), "synchronize");

    std::vector<uint8_t> adj((size_t)n * n, 0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) adj[(size_t)i * n + j] = (i != j);
    ctx->seps.clear();
    ctx->sep_z.clear();
    struct Sep {
        int32_t x, y, level, d;
        int64_t zoff;
    };
    std::vector<Sep> seps;
    std::vector<int32_t> sepz;

    const double t_start = now_s();
    // ---- level 0: every pair of the complete graph (PCStable.cpp:133-195)
    std::vector<int32_t> ex, ey;
    ex.reserve((size_t)n * (n - 1) / 2);
    ey.reserve((size_t)n * (n - 1) / 2);
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            ex.push_back(i);
            ey.push_back(j);
        }
    const int64_t E0 = (int64_t)ex.size();
    std::vector<DevResult> res;
    {
        const double tl = now_s();
        std::vector<HostTest> tv((size_t)E0);
        for (int64_t e = 0; e < E0; ++e) tv[e] = HostTest{ex[e], ey[e], 0, 0};
        st = run_tests(ctx, plain, o, 0, tv, nullptr, 0, alpha, band, res, S);
        if (st != FPGM_OK) return st;
        int64_t w = 0, inb = 0;
        for (int64_t e = 0; e < E0; ++e) {
            inb += res[e].in_band;
            if (res[e].exact && res[e].df > 0 && res[e].g2_64 > 0.0) {
                const double rel = std::fabs((double)res[e].g2_32 - res[e].g2_64) / res[e].g2_64;
                S.max_rel_err_g2_fp32 = std::max(S.max_rel_err_g2_fp32, rel);
            }
            if (res[e].indep) {
                seps.push_back(Sep{ex[e], ey[e], 0, 0, (int64_t)sepz.size()});
                adj[(size_t)ex[e] * n + ey[e]] = 0;
                adj[(size_t)ey[e] * n + ex[e]] = 0;
            } else {
                ex[w] = ex[e];
                ey[w] = ey[e];
                ++w;
            }
        }
        ex.resize(w);
        ey.resize(w);
        S.tests_per_level[0] = E0;
        S.executed_per_level[0] = E0;
        S.in_band_per_level[0] = inb;
        S.rounds_per_level[0] = 1;
        S.edges_after_level[0] = w;
        S.seconds_per_level[0] = now_s() - tl;
        S.n_levels = 1;
        log_level(o, 0, E0, w, now_s() - t_start);
    }

    // ---- levels d >= 1 (PCStable.cpp:209-231, SearchAtDepth :265-388)
    Frozen F;
    std::vector<EdgeWalk> walks;
    std::vector<int32_t> choice, zlevel;
    std::vector<HostTest> tv;
    std::vector<int32_t> zpool;
    std::vector<int64_t> t_edge;  // edge of each test in the round
    for (int d = 1; d < depth && d < 64; ++d) {
        if (d > kMaxD) return fail(ctx, FPGM_ETOOBIG, "conditioning depth exceeds the device limit");
        const double tl = now_s();
        F.off.assign(n + 1, 0);
        F.nb.clear();
        for (int i = 0; i < n; ++i) {
            for (int j = 0; j < n; ++j)
                if (adj[(size_t)i * n + j]) F.nb.push_back(j);
            F.off[i + 1] = (int32_t)F.nb.size();
        }
        const int64_t E = (int64_t)ex.size();
        walks.assign((size_t)E, EdgeWalk{});
        choice.assign((size_t)E * d, 0);
        zlevel.clear();
        std::vector<int64_t> active;
        for (int64_t e = 0; e < E; ++e) {
            EdgeWalk &w = walks[e];
            w.x = ex[e];
            w.y = ey[e];
            w.tests = 0;
            w.removed = 0;
            w.zsep = -1;
            if (start_side(w, F, d, 0, &choice[(size_t)e * d])) active.push_back(e);
        }
        int64_t executed = 0, inb = 0, rounds = 0;
        const int64_t budget = o.round_tests > 0
                                   ? o.round_tests
                                   : std::min<int64_t>((int64_t)1 << 20,
                                                       std::max<int64_t>(8192, ((int64_t)1 << 31) /
                                                                                   std::max<int64_t>(1, (d + 2) * N)));
        while (!active.empty()) {
            if (o.max_rounds > 0 && rounds >= o.max_rounds)
                return fail(ctx, FPGM_ETIMEOUT, "level exceeded the round limit");
            ++rounds;
            const bool plain_round = plain && (o.plain_sample <= 0 || S.plain_tests_per_level[d] < o.plain_sample);
            int64_t W = plain_round ? 1 : std::max<int64_t>(1, budget / (int64_t)active.size());
            if (!plain_round)
                W = std::max<int64_t>(W, std::min<int64_t>((int64_t)1 << std::min<int64_t>(rounds - 1, 20), 4096));
            tv.clear();
            zpool.clear();
            t_edge.clear();
            for (int64_t e : active) {
                EdgeWalk &w = walks[e];
                int32_t *c = &choice[(size_t)e * d];
                for (int64_t k = 0; k < W && w.side < 2; ++k) {
                    tv.push_back(HostTest{w.x, w.y, d, (int64_t)zpool.size()});
                    for (int i = 0; i < d; ++i) zpool.push_back(F.at(w.u, w.skip, c[i]));
                    t_edge.push_back(e);
                    if (!next_choice(w.a, d, c)) start_side(w, F, d, w.side + 1, c);
                }
            }
            st = run_tests(ctx, plain_round, o, d, tv, zpool.data(), (int64_t)zpool.size(), alpha, band, res, S);
            if (st != FPGM_OK) return st;
            executed += (int64_t)tv.size();
            std::vector<int64_t> still;
            size_t k = 0;
            for (int64_t e : active) {
                EdgeWalk &w = walks[e];
                bool done = false;
                for (; k < tv.size() && t_edge[k] == e; ++k) {
                    inb += res[k].in_band;
                    if (res[k].exact && res[k].df > 0 && res[k].g2_64 > 0.0) {
                        const double rel = std::fabs((double)res[k].g2_32 - res[k].g2_64) / res[k].g2_64;
                        S.max_rel_err_g2_fp32 = std::max(S.max_rel_err_g2_fp32, rel);
                    }
                    if (done) continue;
                    w.tests += 1;
                    if (res[k].indep) {
                        done = true;
                        w.removed = 1;
                        w.zsep = (int64_t)zlevel.size();
                        for (int i = 0; i < d; ++i) zlevel.push_back(zpool[tv[k].z_off + i]);
                    }
                }
                if (!done && w.side < 2) still.push_back(e);
            }
            active.swap(still);
        }
        int64_t tests = 0, wkeep = 0;
        for (int64_t e = 0; e < E; ++e) {
            const EdgeWalk &w = walks[e];
            tests += w.tests;
            if (w.removed) {
                seps.push_back(Sep{w.x, w.y, d, d, (int64_t)sepz.size()});
                for (int i = 0; i < d; ++i) sepz.push_back(zlevel[w.zsep + i]);
                adj[(size_t)w.x * n + w.y] = 0;
                adj[(size_t)w.y * n + w.x] = 0;
            } else {
                ex[wkeep] = w.x;
                ey[wkeep] = w.y;
                ++wkeep;
            }
        }
        ex.resize(wkeep);
        ey.resize(wkeep);
        S.tests_per_level[d] = tests;
        S.executed_per_level[d] = executed;
        S.in_band_per_level[d] = inb;
        S.rounds_per_level[d] = rounds;
        S.edges_after_level[d] = wkeep;
        S.seconds_per_level[d] = now_s() - tl;
        S.n_levels = d + 1;
        log_level(o, d, tests, wkeep, now_s() - t_start);
        int maxdeg = 0;
        for (int i = 0; i < n; ++i) {
            int deg = 0;
            for (int j = 0; j < n; ++j) deg += adj[(size_t)i * n + j];
            maxdeg = std::max(maxdeg, deg);
        }
        if (!(maxdeg - 1 > d)) break;
    }
    S.ci_seconds = now_s() - t_start;

    std::sort(seps.begin(), seps.end(),
              [](const Sep &a, const Sep &b) { return a.x != b.x ? a.x < b.x : a.y < b.y; });
    for (const Sep &s : seps) {
        fpgm_sepset r{};
        r.x = s.x;
        r.y = s.y;
        r.level = s.level;
        r.d = s.d;
        r.z_off = (int64_t)ctx->sep_z.size();
        for (int i = 0; i < s.d; ++i) ctx->sep_z.push_back(sepz[s.zoff + i]);
        ctx->seps.push_back(r);
    }
    S.n_sepsets = (int64_t)ctx->seps.size();
    S.sep_z_total = (int64_t)ctx->sep_z.size();
    for (int l = 0; l < S.n_levels; ++l) S.num_ci_test += S.tests_per_level[l];
    std::memcpy(adj_out, adj.data(), adj.size());
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

fpgm_status fpgm_last_kernel_stats(const fpgm_ctx *ctx, double *kernel_ms, double *alg_bytes, int64_t *launches)
{
    if (!ctx) return FPGM_EINVAL;
    if (kernel_ms) *kernel_ms = ctx->k_ms;
    if (alg_bytes) *alg_bytes = ctx->k_bytes;
    if (launches) *launches = ctx->k_launches;
    return FPGM_OK;
}

}  // extern "C"
