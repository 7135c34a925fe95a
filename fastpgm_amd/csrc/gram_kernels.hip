// Level 0 of the CI stage as an int8 one-hot Gram matrix on the MFMA units (gfx950).
//
// With O the N x W one-hot matrix of all variables (W = sum of cardinalities, column
// o_v + a is 1 where variable v takes state a), every level-0 contingency table is a
// block of C = O^T O: n_{vw}[a][b] = C[o_v + a][o_w + b] (the Counts2D table,
// CellTable.cpp:430-441).  C is computed with v_mfma_i32_32x32x32_i8 (int8 x int8 ->
// int32, exact) over 128 x 128 tiles of the upper triangle; the one-hot operands are
// formed in registers from 16-byte loads of the uint8 columns (one v_perm_b32 table
// lookup per 4 rows when every cardinality is <= 8).  A second kernel reads
// each pair's block and evaluates df / G^2 / decision like the counting kernels.
// The MFMA operand layout and the one-hot helpers are in onehot.h.
#include "ci_kernels.h"
#include "fpgm_ci.h"
#include "gram_kernels.h"
#include "onehot.h"

namespace fpgm {

namespace {

// One workgroup of 2 waves per (tile pair, row split): wave w computes the 128 x 64 block
// of rows ti*128 .. +128 and columns tj*128 + 64w .. +64 as 4 x 2 fragments of 32 x 32 --
// 6 operand loads for 8 MFMAs per 32-row step (a 64 x 64 block per wave needs 4 for 4).
constexpr int kGramWaves = 2;
constexpr int kGramA = 4, kGramB = 2;  // fragments per wave

struct Ops6 {
    uint4 a[kGramA], b[kGramB];
};

template <bool kPerm>
__global__ __launch_bounds__(kGramWaves * 64) __attribute__((amdgpu_waves_per_eu(2))) void k_gram_l0(GramArgs g)
{
    const int item = blockIdx.x;
    const int tpi = item / g.k_splits, ks = item - tpi * g.k_splits;
    int tp = tpi;
    const uint8_t *cols = g.cols;
    int64_t stride = g.stride, n_rows = g.n_rows;
    int32_t *C = g.C;
    if (g.insts) {  // instance tpi / n_tp
        const int inst = tpi / g.n_tp;
        tp = tpi - inst * g.n_tp;
        const DevInst I = g.insts[inst];
        cols = I.cols;
        stride = I.stride;
        n_rows = I.n_rows;
        C = g.C + (int64_t)inst * g.c_stride;
    }
    const int ti = g.tiles[2 * tp], tj = g.tiles[2 * tp + 1];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int h = lane >> 5, l32 = lane & 31;
    const int i_base = ti * kGramTile;
    const int j_base = tj * kGramTile + wave * 64;
    const uint8_t *pa[kGramA], *pb[kGramB];
    uint32_t la[kGramA], ha[kGramA], lb[kGramB], hb[kGramB];
#pragma unroll
    for (int f = 0; f < kGramA; ++f) {
        const int r = i_base + 32 * f + l32;
        pa[f] = cols + (int64_t)g.row_var[r] * stride + h * 16;
        onehot_lut<kPerm>(g.row_state[r], la[f], ha[f]);
    }
#pragma unroll
    for (int f = 0; f < kGramB; ++f) {
        const int r = j_base + 32 * f + l32;
        pb[f] = cols + (int64_t)g.row_var[r] * stride + h * 16;
        onehot_lut<kPerm>(g.row_state[r], lb[f], hb[f]);
    }

    v16i c[kGramA][kGramB];
#pragma unroll
    for (int f = 0; f < kGramA; ++f)
#pragma unroll
        for (int e = 0; e < kGramB; ++e) c[f][e] = v16i{0};
    const int64_t r_begin = (int64_t)ks * g.k_per_split;
    if (r_begin >= n_rows) return;  // a split past a shorter instance's rows
    const int64_t r_end = (r_begin + g.k_per_split < n_rows) ? r_begin + g.k_per_split : n_rows;
    // full 32-row steps: no masks; the loads of the next step are issued before the
    // current step's one-hot formation and MFMAs (two register buffers in turn)
    const int64_t full_end = r_begin + (r_end - r_begin) / 32 * 32;
    auto load = [&](int64_t r0, Ops6 &o) {
#pragma unroll
        for (int f = 0; f < kGramA; ++f) o.a[f] = *reinterpret_cast<const uint4 *>(pa[f] + r0);
#pragma unroll
        for (int f = 0; f < kGramB; ++f) o.b[f] = *reinterpret_cast<const uint4 *>(pb[f] + r0);
    };
    auto mma = [&](const Ops6 &o, const uint4 *mask) {
        v4i b[kGramB];
#pragma unroll
        for (int e = 0; e < kGramB; ++e) b[e] = onehot<kPerm>(o.b[e], lb[e], hb[e]);
#pragma unroll
        for (int f = 0; f < kGramA; ++f) {
            v4i a = onehot<kPerm>(o.a[f], la[f], ha[f]);
            if (mask) a = and4(a, *mask);
#pragma unroll
            for (int e = 0; e < kGramB; ++e) c[f][e] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b[e], c[f][e], 0, 0, 0);
        }
    };
    if (r_begin < full_end) {
        // the scheduling barriers keep the next step's loads ahead of this step's
        // arithmetic (left alone, the compiler sinks them to their first use)
        Ops6 o0, o1;
        load(r_begin, o0);
        int64_t r0 = r_begin;
        for (; r0 + 32 < full_end; r0 += 64) {
            load(r0 + 32, o1);
            __builtin_amdgcn_sched_barrier(0);
            mma(o0, nullptr);
            __builtin_amdgcn_sched_barrier(0);
            if (r0 + 64 < full_end) load(r0 + 64, o0);
            __builtin_amdgcn_sched_barrier(0);
            mma(o1, nullptr);
            __builtin_amdgcn_sched_barrier(0);
        }
        if (r0 < full_end) mma(o0, nullptr);
    }
    if (full_end < r_end) {
        // last partial step: bytes past r_end are masked out (their values are arbitrary)
        const int64_t valid = r_end - (full_end + h * 16);
        uint4 m;
        m.x = byte_mask_word(valid, 0);
        m.y = byte_mask_word(valid, 1);
        m.z = byte_mask_word(valid, 2);
        m.w = byte_mask_word(valid, 3);
        Ops6 o;
        load(full_end, o);
        mma(o, &m);
    }
    // accumulate the split's partial block into C
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int row = (q >> 2) * 8 + h * 4 + (q & 3);
#pragma unroll
        for (int f = 0; f < kGramA; ++f) {
            int32_t *cr = C + (int64_t)(i_base + 32 * f + row) * g.Wp + j_base + l32;
#pragma unroll
            for (int e = 0; e < kGramB; ++e)
                if (c[f][e][q]) atomicAdd(cr + 32 * e, c[f][e][q]);
        }
    }
}

__device__ __forceinline__ void nadd(float &s, float &c, float v)
{
    const float t = s + v;
    if (fabsf(s) >= fabsf(v)) c += (s - t) + v;
    else c += (v - t) + s;
    s = t;
}

__device__ __forceinline__ float term32(int32_t n, int32_t ni, int32_t nj, int32_t nk)
{
    const int64_t num = (int64_t)n * nk;
    const int64_t den = (int64_t)ni * nj;
    const int64_t diff = num - den;
    const int64_t ad = diff < 0 ? -diff : diff;
    float L;
    if (2 * ad <= den) L = log1pf((float)diff / (float)den);
    else L = logf((float)num / (float)den);
    return 2.0f * (float)n * L;
}

// One thread per pair (x < y): its block of C is the level-0 table n[a][b]; marginals,
// adjusted df (IndependenceTest.cpp:309-322), FP32 compensated G^2 in the reference's
// (a, b) order, FP32 decision outside the recheck band, else FP64 G^2 with the
// reference's expression (IndependenceTest.cpp:327-347) and k_pvalue decides.
__global__ __launch_bounds__(256) void k_pairs_l0(PairArgs p)
{
    const int64_t tg = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (tg >= p.n_pairs * p.n_inst) return;
    const int64_t inst = tg / p.n_pairs, t = tg - inst * p.n_pairs;
    int x, y;
    if (p.px) {
        x = p.px[t];
        y = p.py[t];
    } else {
        // t-th pair (x, y), x < y, in lexicographic order: row x starts at x (2n - x - 1) / 2
        const int64_t n = p.n_vars;
        const double b = 2.0 * (double)n - 1.0;
        int64_t xx = (int64_t)((b - sqrt(fmax(b * b - 8.0 * (double)t, 0.0))) * 0.5);
        auto row0 = [&](int64_t r) { return r * (2 * n - r - 1) / 2; };
        while (xx > 0 && row0(xx) > t) --xx;
        while (row0(xx + 1) <= t) ++xx;
        x = (int)xx;
        y = (int)(t - row0(xx) + xx + 1);
    }
    const int cx = p.card[x], cy = p.card[y];
    const int32_t *base = p.C + inst * p.c_stride + (int64_t)p.off[x] * p.Wp + p.off[y];
    int32_t nj[kPairMaxCard];
    int32_t nk = 0;
    int aly = 0;
    for (int b = 0; b < cy; ++b) {
        int32_t s = 0;
        for (int a = 0; a < cx; ++a) s += base[(int64_t)a * p.Wp + b];
        nj[b] = s;
        nk += s;
        aly += (s > 0);
    }
    int alx = 0;
    float s = 0.f, c = 0.f, as = 0.f;
    for (int a = 0; a < cx; ++a) {
        const int32_t *row = base + (int64_t)a * p.Wp;
        int32_t ni = 0;
        for (int b = 0; b < cy; ++b) ni += row[b];
        alx += (ni > 0);
        if (ni == 0) continue;
        for (int b = 0; b < cy; ++b) {
            const int32_t n = row[b];
            if (n == 0) continue;
            const float tm = term32(n, ni, nj[b], nk);
            nadd(s, c, tm);
            as += fabsf(tm);
        }
    }
    alx = alx >= 1 ? alx : 1;
    aly = aly >= 1 ? aly : 1;
    const int df = (alx - 1) * (aly - 1);
    const float g2 = s + c;
    DevResult R;
    R.g2_32 = g2;
    R.abs_sum = as;
    R.df = df;
    R.indep = 0;
    R.in_band = 0;
    R.exact = 0;
    R.need_p = 0;
    R.g2_64 = NAN;
    R.p64 = (df == 0) ? 1.0 : NAN;
    bool need64 = (p.flags & FPGM_F_EXACT) != 0;
    if (df == 0) {
        R.indep = 1;  // IndependenceTest.cpp:349-351
    } else {
        double crit;
        if (df <= kCritTab) {
            crit = p.crit[df - 1];
        } else {
            const double k = (double)df, q = 2.0 / (9.0 * k), u = 1.0 - q + p.z_alpha * sqrt(q);
            crit = k * u * u * u;
        }
        const double bound = 4e-6 * (double)as + (df > kCritTab ? 1e-5 * crit : 0.0) + 1e-30;
        const double width = fmax((double)p.band * crit, bound);
        R.in_band = (isnan(crit) || fabs((double)g2 - crit) <= width) ? 1 : 0;
        if ((p.flags & FPGM_F_FP32ONLY) && !isnan(crit)) R.indep = (double)g2 < crit;
        else if (R.in_band) need64 = true;
        else R.indep = (double)g2 < crit;
    }
    if (need64) {
        double G = 0.0;
        for (int a = 0; a < cx; ++a) {
            const int32_t *row = base + (int64_t)a * p.Wp;
            int32_t ni = 0;
            for (int b = 0; b < cy; ++b) ni += row[b];
            if (ni == 0) continue;
            for (int b = 0; b < cy; ++b) {
                const int32_t n = row[b];
                if (nj[b] == 0 || n == 0) continue;
                const double expected = (double)nj[b] * (double)ni / (double)nk;
                G += 2.0 * (double)n * log((double)n / expected);
            }
        }
        R.g2_64 = G;
        R.exact = 1;
        R.need_p = (df != 0) ? 1 : 0;
    }
    p.res[tg] = R;
}

__global__ __launch_bounds__(256) void k_pack_l0(const DevResult *res, int64_t n, uint8_t *flags, PairSummary *sum)
{
    __shared__ float s_err[4];
    __shared__ uint32_t s_inb[4];
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    float err = 0.f;
    uint32_t inb = 0;
    if (t < n) {
        const DevResult R = res[t];
        flags[t] = (uint8_t)((R.indep ? 1 : 0) | (R.in_band ? 2 : 0));
        inb = R.in_band ? 1u : 0u;
        if (R.exact && R.df > 0 && R.g2_64 > 0.0) err = (float)(fabs((double)R.g2_32 - R.g2_64) / R.g2_64);
    }
    for (int o = 32; o > 0; o >>= 1) {
        err = fmaxf(err, __shfl_xor(err, o, 64));
        inb += __shfl_xor(inb, o, 64);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_err[wave] = err;
        s_inb[wave] = inb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float e = s_err[0];
        uint32_t c = s_inb[0];
        for (int w = 1; w < 4; ++w) {
            e = fmaxf(e, s_err[w]);
            c += s_inb[w];
        }
        // non-negative floats order like their bit patterns
        if (e > 0.f) atomicMax(&sum->max_rel_err_bits, __float_as_uint(e));
        if (c) atomicAdd(&sum->in_band, c);
    }
}

}  // namespace

hipError_t launch_pack_l0(const DevResult *res, int64_t n, uint8_t *flags, PairSummary *sum, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_l0, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, res, n, flags, sum);
    return hipGetLastError();
}

hipError_t launch_gram_l0(const GramArgs &g, int n_items, hipStream_t s)
{
    if (n_items <= 0) return hipSuccess;
    if (g.max_card <= 8) hipLaunchKernelGGL(k_gram_l0<true>, dim3(n_items), dim3(kGramWaves * 64), 0, s, g);
    else hipLaunchKernelGGL(k_gram_l0<false>, dim3(n_items), dim3(kGramWaves * 64), 0, s, g);
    return hipGetLastError();
}

hipError_t launch_pairs_l0(const PairArgs &p, hipStream_t s)
{
    if (p.n_pairs <= 0 || p.n_inst <= 0) return hipSuccess;
    const int64_t blocks = (p.n_pairs * p.n_inst + 255) / 256;
    hipLaunchKernelGGL(k_pairs_l0, dim3((unsigned)blocks), dim3(256), 0, s, p);
    return hipGetLastError();
}

}  // namespace fpgm
