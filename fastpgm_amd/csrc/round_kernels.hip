// Device-generated rounds: expansion of the host's per-edge segments into tests, and the
// per-segment first-independent reduction.  See round_kernels.h.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>

#include "round_kernels.h"

namespace fpgm {

namespace {

constexpr int kRT = 256;  // threads per workgroup

__device__ __forceinline__ DevResult skipped_res()
{
    DevResult R;
    R.g2_32 = 0.f;
    R.abs_sum = 0.f;
    R.df = kSkippedDf;
    R.indep = 0;
    R.in_band = 0;
    R.exact = 0;
    R.need_p = 0;
    R.g2_64 = NAN;
    R.p64 = NAN;
    return R;
}

// Last segment whose first test is at or before t (segments are in test order).
__device__ __forceinline__ int64_t find_seg(const RoundSeg *segs, int64_t n_segs, int64_t t)
{
    int64_t lo = 0, hi = n_segs - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (segs[mid].t0 <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The subset of lexicographic rank r among the D-subsets of {0..a-1}: position i takes the
// smallest x whose subsets (C(a-1-x, D-1-i) of them continue from x) reach past r; the
// rank left over is zero exactly when r < C(a, D).  Positions stay in [0, a) either way.
template <int D>
__device__ __forceinline__ bool unrank(const ExpandArgs &e, int a, int64_t r, int32_t (&c)[D > 0 ? D : 1])
{
    int x = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
        const int last = a - D + i;  // the largest value position i can hold
        for (; x < last; ++x) {
            const int64_t cnt = e.binom[(int64_t)(a - 1 - x) * e.binom_stride + (D - 1 - i)];
            if (r < cnt) break;
            r -= cnt;
        }
        c[i] = x++;
    }
    return r == 0;
}

// ChoiceGenerator::Next (ChoiceGenerator.cpp:55-85) on the D-subsets of {0..a-1}.
template <int D>
__device__ __forceinline__ void next_subset(int a, int32_t (&c)[D > 0 ? D : 1])
{
#pragma unroll
    for (int i = D - 1; i >= 0; --i) {
        if (c[i] < i + (a - D)) {
            c[i]++;
            for (int j = i + 1; j < D; ++j) c[j] = c[j - 1] + 1;
            return;
        }
    }
}

template <int D>
__global__ __launch_bounds__(kRT) void k_expand(ExpandArgs e)
{
    __shared__ unsigned long long s_acc[5 * kRouteGroups];
    __shared__ uint32_t s_err;
    const int tid = threadIdx.x, lane = tid & 63;
    if (tid < 5 * kRouteGroups) s_acc[tid] = 0ull;
    if (tid == 0) s_err = 0u;
    __syncthreads();

    const int64_t tbeg = ((int64_t)blockIdx.x * kRT + tid) * kExpandPer;
    int grp[kExpandPer];
    int32_t nch[kExpandPer];
    uint32_t cnt[kRouteGroups] = {0u, 0u, 0u, 0u};
    uint32_t ntest[kRouteGroups] = {0u, 0u, 0u, 0u};
    unsigned long long rows_acc[kRouteGroups] = {0ull, 0ull, 0ull, 0ull};
    unsigned long long alg[kRouteGroups] = {0ull, 0ull, 0ull, 0ull};
    unsigned long long pbits[kRouteGroups] = {0ull, 0ull, 0ull, 0ull};
    uint32_t err = 0u;
    int32_t c[D > 0 ? D : 1];
    int64_t s = -1;
    RoundSeg S{};
    bool skip = false;
    int64_t rows = 0;
    const int64_t rpc = e.rows_per_chunk;
#pragma unroll
    for (int j = 0; j < kExpandPer; ++j) {
        grp[j] = -1;
        nch[j] = 0;
        const int64_t t = tbeg + j;
        if (t >= e.n_tests) continue;
        bool fresh = false;
        if (s < 0) {
            s = find_seg(e.segs, e.n_segs, t);
            fresh = true;
        } else if (t >= S.t0 + S.count) {
            ++s;
            fresh = true;
        }
        if (fresh) {
            S = e.segs[s];
            skip = e.edge_done[S.edge] < e.round;
            rows = e.inst[S.inst].n_rows;
            if (!unrank<D>(e, S.a, S.rank0 + (t - S.t0), c)) err |= 8u;
            if (t == S.t0) e.seg_out[s] = skip ? kSegSkipped : kSegRun;
        } else {
            next_subset<D>(S.a, c);
        }
        e.t2s[t] = (int32_t)s;
        if (skip) {
            e.res[t] = skipped_res();
            continue;
        }
        DevTest T{};
        T.x = S.x;
        T.y = S.y;
        T.d = D;
        T.z_off = (int32_t)(t * D);
        T.inst = S.inst;
        T.edge = S.edge;
        T.cx = e.card[S.x];
        T.cy = e.card[S.y];
        int64_t dimz = 1;
        bool narrow = e.colw[S.x] != 8 && e.colw[S.y] != 8;
        int bits = e.colw[S.x] + e.colw[S.y];
#pragma unroll
        for (int i = 0; i < D; ++i) {
            const int32_t z = e.nbr[S.u_off + c[i] + (c[i] >= S.skip ? 1 : 0)];
            e.zpool[t * D + i] = z;
            dimz *= e.card[z];
            const int w = e.colw[z];
            narrow = narrow && w != 8;
            bits += w;
        }
        const int64_t cells = dimz * T.cx * T.cy;
        const int64_t marg = dimz * (T.cx + T.cy + 1);
        int32_t flags = narrow ? kTFNarrow : 0;
        bool fits = false;
        int g = route_lds(e.rp, flags, cells, marg, rows, T.cx, T.cy, D, &fits);
        if (g == kRoutePending) {
            if (e.w128_shared) {
                flags |= kTFLdsTable;
                fits = cells + marg <= kLdsCells;
                g = 2;
            } else {
                g = 1;
            }
        }
        if (g >= 0) g = round_group(g);
        if (g < 0 || !((e.launch_mask >> g) & 1u)) {  // the host's bounds for the level exclude both
            err |= g < 0 ? 1u : 16u;
            e.res[t] = skipped_res();
            continue;
        }
        const int64_t n_chunks = rows > 0 ? (rows + rpc - 1) / rpc : 1;
        const bool need_ws = n_chunks > 1 || !fits;
        T.ws_off = -1;
        if (need_ws) {
            const unsigned long long need = (unsigned long long)((cells + marg + 63) / 64 * 64);
            const unsigned long long off = atomicAdd(&e.stats->ws_next, need);
            if ((int64_t)(off + need) > e.ws_cap) {
                err |= 4u;
                e.res[t] = skipped_res();
                continue;
            }
            T.ws_off = (int64_t)off;
        } else {
            flags |= kTFLdsMarg;
        }
        T.n_slices = 1;
        T.slice_cells = (int32_t)cells;
        T.dimz = (int32_t)dimz;
        T.n_chunks = (int32_t)n_chunks;
        T.item0 = (int32_t)s;
        T.flags = flags;
        e.tests[t] = T;
        grp[j] = g;
        nch[j] = (int32_t)n_chunks;
        cnt[g] += (uint32_t)n_chunks;
        ntest[g] += 1u;
        rows_acc[g] += (unsigned long long)rows;
        alg[g] += (unsigned long long)(D + 2) * (unsigned long long)rows;
        pbits[g] += (unsigned long long)bits * (unsigned long long)rows;
    }
    // work items: each wave reserves its items of every group with one atomic (the
    // lanes' counts scanned across the wave), so that items keep the tests' order
    // within a wave
    uint32_t base[kRouteGroups];
#pragma unroll
    for (int g = 0; g < kRouteGroups; ++g) {
        const uint32_t v = cnt[g];
        uint32_t incl = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(incl, (unsigned)o, 64);
            if (lane >= o) incl += y;
        }
        const uint32_t tot = __shfl(incl, 63, 64);
        uint32_t b = 0u;
        if (lane == 63 && tot) b = atomicAdd(&e.group_count[g], tot);
        b = __shfl(b, 63, 64);
        base[g] = b + incl - v;
    }
#pragma unroll
    for (int j = 0; j < kExpandPer; ++j) {
        const int g = grp[j];
        if (g < 0) continue;
        const int64_t t = tbeg + j;
        for (int32_t k = 0; k < nch[j]; ++k) {
            const int64_t pos = (int64_t)base[g]++;
            if (pos >= e.group_cap) {
                err |= 2u;
                break;
            }
            e.items[(int64_t)g * e.group_cap + pos] = ((uint64_t)t << 32) | (uint64_t)(uint32_t)k;
        }
    }
    // the workgroup's figures, then one device atomic per nonzero figure
#pragma unroll
    for (int g = 0; g < kRouteGroups; ++g) {
        if (ntest[g]) atomicAdd(&s_acc[0 * kRouteGroups + g], (unsigned long long)ntest[g]);
        if (cnt[g]) atomicAdd(&s_acc[1 * kRouteGroups + g], (unsigned long long)cnt[g]);
        if (rows_acc[g]) atomicAdd(&s_acc[2 * kRouteGroups + g], rows_acc[g]);
        if (alg[g]) atomicAdd(&s_acc[3 * kRouteGroups + g], alg[g]);
        if (pbits[g]) atomicAdd(&s_acc[4 * kRouteGroups + g], pbits[g]);
    }
    if (err) atomicOr(&s_err, err);
    __syncthreads();
    if (tid < 5 * kRouteGroups && s_acc[tid]) {
        unsigned long long *dst = tid < kRouteGroups         ? &e.stats->tests[tid]
                                  : tid < 2 * kRouteGroups ? &e.stats->items[tid - kRouteGroups]
                                  : tid < 3 * kRouteGroups ? &e.stats->rows[tid - 2 * kRouteGroups]
                                  : tid < 4 * kRouteGroups ? &e.stats->alg[tid - 3 * kRouteGroups]
                                                           : &e.stats->pbits[tid - 4 * kRouteGroups];
        atomicAdd(dst, s_acc[tid]);
    }
    if (tid == 0 && s_err) atomicOr(&e.stats->err, s_err);
}

// One thread per test: a segment's first independent test (lowest index within it),
// the executed tests' recheck-band count and largest FP32 G^2 error.
__global__ __launch_bounds__(kRT) void k_round_reduce(ReduceArgs r)
{
    const int64_t t = (int64_t)blockIdx.x * kRT + threadIdx.x;
    const int lane = threadIdx.x & 63;
    uint32_t inb = 0u;
    float err = 0.f;
    if (t < r.n_tests) {
        const DevResult R = r.res[t];
        if (R.df != kSkippedDf) {
            inb = R.in_band;
            if (R.exact && R.df > 0 && R.g2_64 > 0.0) err = (float)(fabs((double)R.g2_32 - R.g2_64) / R.g2_64);
            if (R.indep) {
                const int32_t s = r.t2s[t];
                atomicMin(&r.seg_out[s], (int32_t)(t - r.segs[s].t0));
            }
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        inb += __shfl_xor(inb, o, 64);
        err = fmaxf(err, __shfl_xor(err, o, 64));
    }
    if (lane == 0) {
        if (inb) atomicAdd(&r.stats->in_band, inb);
        if (err > 0.f) atomicMax(&r.stats->max_err_bits, __float_as_uint(err));
    }
}

}  // namespace

hipError_t launch_expand(const ExpandArgs &e, hipStream_t s)
{
    if (e.n_tests <= 0) return hipSuccess;
    const int64_t threads = (e.n_tests + kExpandPer - 1) / kExpandPer;
    const dim3 grid((unsigned)((threads + kRT - 1) / kRT)), block(kRT);
    switch (e.d) {
    case 1: hipLaunchKernelGGL(k_expand<1>, grid, block, 0, s, e); break;
    case 2: hipLaunchKernelGGL(k_expand<2>, grid, block, 0, s, e); break;
    case 3: hipLaunchKernelGGL(k_expand<3>, grid, block, 0, s, e); break;
    case 4: hipLaunchKernelGGL(k_expand<4>, grid, block, 0, s, e); break;
    case 5: hipLaunchKernelGGL(k_expand<5>, grid, block, 0, s, e); break;
    case 6: hipLaunchKernelGGL(k_expand<6>, grid, block, 0, s, e); break;
    case 7: hipLaunchKernelGGL(k_expand<7>, grid, block, 0, s, e); break;
    case 8: hipLaunchKernelGGL(k_expand<8>, grid, block, 0, s, e); break;
    case 9: hipLaunchKernelGGL(k_expand<9>, grid, block, 0, s, e); break;
    case 10: hipLaunchKernelGGL(k_expand<10>, grid, block, 0, s, e); break;
    case 11: hipLaunchKernelGGL(k_expand<11>, grid, block, 0, s, e); break;
    case 12: hipLaunchKernelGGL(k_expand<12>, grid, block, 0, s, e); break;
    case 13: hipLaunchKernelGGL(k_expand<13>, grid, block, 0, s, e); break;
    case 14: hipLaunchKernelGGL(k_expand<14>, grid, block, 0, s, e); break;
    case 15: hipLaunchKernelGGL(k_expand<15>, grid, block, 0, s, e); break;
    case 16: hipLaunchKernelGGL(k_expand<16>, grid, block, 0, s, e); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_round_reduce(const ReduceArgs &r, hipStream_t s)
{
    if (r.n_tests <= 0) return hipSuccess;
    const int64_t blocks = (r.n_tests + kRT - 1) / kRT;
    hipLaunchKernelGGL(k_round_reduce, dim3((unsigned)blocks), dim3(kRT), 0, s, r);
    return hipGetLastError();
}

}  // namespace fpgm
