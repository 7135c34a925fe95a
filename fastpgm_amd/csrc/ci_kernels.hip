// Fused count + G^2 kernels of the PC-stable CI stage, written for gfx950 (CDNA4).
//
// One work item = (test, row chunk).  A 256-thread workgroup streams the chunk's
// packed columns (16 rows per thread per step, 4096 rows per workgroup step: one
// 16-, 8- or 4-byte load per column as it holds 8, 4 or 2 bits per value), forms
// the contingency-table index of every row
//     key = ((z_1 * c(Z_2) + z_2) ... ) * c(X) + x) * c(Y) + y
// (Counts3D / Counts2D layout, CellTable.cpp:254-291,430-441), and increments an
// LDS histogram, one copy per wave when the table is small enough.  A test whose
// rows span several chunks merges the chunk histograms into a global table with
// integer atomics; the workgroup that retires the last chunk computes the
// statistic.  Counts are exact int32 throughout.
//
// Statistic (IndependenceTest.cpp:94-154): marginals, adjusted df, then G^2 in FP32
// with Neumaier-compensated summation of terms 2 n log(n n_z / (n_x n_y)), each
// term's ratio formed from exact int64 products.  When the FP32 value lies within
// max(band * crit, error bound) of the chi-square critical value (or df exceeds the
// tabulated range, or the caller asked for exact results) the workgroup recomputes
// G^2 in FP64 with the reference's expression and evaluates the FP64 p-value.
#include <algorithm>

#include "ci_kernels.h"
#include "chisq.h"
#include "fpgm_ci.h"
#include "onehot.h"

namespace fpgm {

namespace {

// 16-byte load of column bytes through the global address space: the column pointers
// come from descriptors in memory, and a generic (flat) load would also count against
// the LDS wait counter and serialise with the LDS sinks.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint4 ld16(const uint8_t *p)
{
    const u32x4 v = *(const __attribute__((address_space(1))) u32x4 *)(p);
    return make_uint4(v.x, v.y, v.z, v.w);
}

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// Packed columns (k_pack_cols): each group of 16 rows, rows 16g .. 16g+15, takes 16, 8
// or 4 bytes at byte offset 16g * w / 8 of the column, w = bits per value:
//   w = 8: byte b = row b;
//   w = 4: 32-bit word h (h = 0, 1) holds rows 8h+j (j = 0..3) in the low nibble of
//          byte j and rows 8h+4+j in its high nibble;
//   w = 2: the one 32-bit word holds row 4q+j in bits 2q..2q+1 of byte j.
// So that unpacking is one mask (and one shift) per 4 rows: word q of the unpacked
// group holds rows 4q..4q+3, one per byte, as an 8-bit column would.
// w is uniform over the wave (one test per workgroup), so the unpacking branches are
// scalar.  The loads themselves never branch: a group is read with one load at byte
// offset r >> sh (sh = 0, 1, 2 for w = 8, 4, 2) of the fixed size of the loop, 16
// bytes (NW = 4, any width) or 8 bytes (NW = 2, widths 4 and 2), the bytes past the
// group being the next group's (or the column's last bytes: never past the column).
// A load under a branch would need a register copy at the join, and that copy would
// wait for the load, which is still in flight to overlap the counting of the previous
// step.
template <int NW>
struct RawOf {
    typedef uint4 T;
};
template <>
struct RawOf<2> {
    typedef uint2 T;
};
template <>
struct RawOf<8> {  // uint8 columns (no column packed): 16 rows in 16 bytes, nothing to unpack
    typedef uint4 T;
};

typedef unsigned int u32x2a4 __attribute__((ext_vector_type(2), aligned(4)));
typedef unsigned int u32x4a4 __attribute__((ext_vector_type(4), aligned(4)));

template <int NW>
__device__ __forceinline__ typename RawOf<NW>::T ldw(const uint8_t *col, uint32_t sh, int64_t r);

template <>
__device__ __forceinline__ uint4 ldw<4>(const uint8_t *col, uint32_t sh, int64_t r)
{
    const u32x4a4 v = *(const __attribute__((address_space(1))) u32x4a4 *)(col + (r >> sh));
    return make_uint4(v.x, v.y, v.z, v.w);
}

template <>
__device__ __forceinline__ uint4 ldw<8>(const uint8_t *col, uint32_t, int64_t r)
{
    return ld16(col + r);
}

template <>
__device__ __forceinline__ uint2 ldw<2>(const uint8_t *col, uint32_t sh, int64_t r)
{
    const u32x2a4 v = *(const __attribute__((address_space(1))) u32x2a4 *)(col + (r >> sh));
    return make_uint2(v.x, v.y);
}

// load shift of a width
__device__ __forceinline__ uint32_t wshift(uint32_t w)
{
    return w == 8 ? 0u : w == 4 ? 1u : 2u;
}

// The unpacked bytes ORed with `orv` (0, or the per-lane sinks' bias in every byte: one
// v_and_or_b32 with the mask, inside the branch of the width).
__device__ __forceinline__ uint4 unpack4(uint32_t x, uint32_t y, uint32_t orv)
{
    const uint32_t m = 0x0F0F0F0Fu;
    return make_uint4((x & m) | orv, ((x >> 4) & m) | orv, (y & m) | orv, ((y >> 4) & m) | orv);
}

__device__ __forceinline__ uint4 unpack2(uint32_t x, uint32_t orv)
{
    const uint32_t m = 0x03030303u;
    return make_uint4((x & m) | orv, ((x >> 2) & m) | orv, ((x >> 4) & m) | orv, ((x >> 6) & m) | orv);
}

__device__ __forceinline__ uint4 unpackw(const uint4 &p, uint32_t w, uint32_t orv = 0u)
{
    if (w == 8) return make_uint4(p.x | orv, p.y | orv, p.z | orv, p.w | orv);
    if (w == 4) return unpack4(p.x, p.y, orv);
    return unpack2(p.x, orv);
}

__device__ __forceinline__ uint4 unpackw(const uint2 &p, uint32_t w, uint32_t orv = 0u)
{
    if (w == 4) return unpack4(p.x, p.y, orv);
    return unpack2(p.x, orv);
}

__device__ __forceinline__ uint32_t uniform(int32_t v)
{
    return (uint32_t)__builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ uint32_t byte_of(const uint4 &v, int b)
{
    const uint32_t w = (b < 4) ? v.x : (b < 8) ? v.y : (b < 12) ? v.z : v.w;
    return __builtin_amdgcn_ubfe(w, (b & 3) * 8, 8);
}

// a * b + c on the 24-bit multiplier (v_mad_u32_u24): table keys of LDS sinks stay
// below 2^24.  The explicit masks let the compiler prove the operand widths; without
// them it can select the 64-bit multiply-add for a 24-bit product plus a byte.
__device__ __forceinline__ uint32_t mad24(uint32_t a, uint32_t b, uint32_t c)
{
    return (a & 0xFFFFFFu) * (b & 0xFFFFFFu) + c;
}

__device__ __forceinline__ void nadd(float &s, float &c, float v)
{
    const float t = s + v;
    if (fabsf(s) >= fabsf(v)) c += (s - t) + v;
    else c += (v - t) + s;
    s = t;
}

__device__ __forceinline__ int wave_sum_i(int v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double wave_sum_d(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ float wave_sum_f(float v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ void wave_nsum(float &s, float &c)
{
    for (int o = 32; o > 0; o >>= 1) {
        const float s2 = __shfl_xor(s, o, 64);
        const float c2 = __shfl_xor(c, o, 64);
        nadd(s, c, s2);
        c += c2;
    }
}

static_assert(kMaxZLds <= kWG, "one staging thread per Z column");

struct Shared {
    int32_t zcol[kMaxZLds];
    int32_t zcard[kMaxZLds];
    int32_t zw[kMaxZLds];  // bits per value of the packed Z columns
    int32_t wx, wy;        // and of X and Y
    int32_t red_i[kWG / 64];
    float red_f[3 * (kWG / 64)];
    double red_d[kWG / 64];
    int32_t flag;
    int32_t slice_lo, slice_len;  // key range of this work item's slice (kSinkSlice)
    int32_t df;
    int32_t need64;
    int32_t indep;  // the FP32 decision (finalize)
    float g2_32;
    float abs_sum;
};

// FP32 term 2 n log(n nk / (ni nj)); the ratio comes from exact int64 products, and
// log1p is used when the ratio is near 1 so that small log values keep their
// relative accuracy.
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

// Counting sinks.  kSinkShared: one LDS histogram per wave (or shared by several
// waves), LDS atomics.  kSinkGlobal: the test's global table, device atomics (tables
// too large for LDS).  The per-lane sinks (kSinkLane, kSinkLaneW) have their own row
// loop, count_rows_lane below.

template <int kSink>
__device__ __forceinline__ void sink(int32_t *h, uint32_t key)
{
    atomicAdd(&h[key], 1);
}

template <int kSink>
__device__ __forceinline__ void put(int32_t *h, uint32_t key, const Shared &sh);

// kSinkSlice: only keys inside the work item's slice, relative to its first key
__device__ __forceinline__ void sink_slice(int32_t *h, uint32_t key, uint32_t lo, uint32_t len)
{
    const uint32_t k = key - lo;
    if (k < len) atomicAdd(&h[k], 1);
}

// ksh: the shared sink's key shift (5 in its column layout, h the lane's column; else 0)
template <int kSink>
__device__ __forceinline__ void put(int32_t *h, uint32_t key, const Shared &sh, uint32_t ksh)
{
    if (kSink == kSinkSlice) sink_slice(h, key, (uint32_t)sh.slice_lo, (uint32_t)sh.slice_len);
    else if (kSink == kSinkShared) atomicAdd(&h[key << ksh], 1);
    else sink<kSink>(h, key);
}

template <int D, int NW = 4>
struct Cols {
    typename RawOf<NW>::T v[(D > 0 ? D : 1) + 2];
};

// The packed columns of a test: Z (compile-time count D) at 0..D-1, X at X, Y at X+1,
// X = max(D, 1); slot 0 is unused when D = 0.
template <int D, int NW>
struct ColSet {
    static constexpr int X = (D > 0 ? D : 1);
    const uint8_t *p[X + 2];
    uint32_t w[X + 2], s[X + 2];
    __device__ __forceinline__ ColSet(const DevInst &I, const DevTest &T, const Shared &sh)
    {
#pragma unroll
        for (int j = 0; j < D; ++j) {
            p[j] = I.pcols + (int64_t)sh.zcol[j] * I.stride;
            w[j] = uniform(sh.zw[j]);
        }
        p[X] = I.pcols + (int64_t)T.x * I.stride;
        w[X] = uniform(sh.wx);
        p[X + 1] = I.pcols + (int64_t)T.y * I.stride;
        w[X + 1] = uniform(sh.wy);
#pragma unroll
        for (int j = 0; j < X + 2; ++j) s[j] = wshift(w[j]);
    }
    // the packed 16-row groups at row r
    __device__ __forceinline__ void load(int64_t r, Cols<D, NW> &c) const
    {
#pragma unroll
        for (int j = 0; j < D; ++j) c.v[j] = ldw<NW>(p[j], s[j], r);
        c.v[X] = ldw<NW>(p[X], s[X], r);
        c.v[X + 1] = ldw<NW>(p[X + 1], s[X + 1], r);
    }
    // ... as bytes, row b of the group in byte b (Y's bytes ORed with ybias)
    __device__ __forceinline__ void unpack(const Cols<D, NW> &c, Cols<D> &u, uint32_t ybias = 0u) const
    {
        if constexpr (NW == 8) {
#pragma unroll
            for (int j = 0; j < X + 2; ++j) u.v[j] = c.v[j];
            u.v[X + 1] = make_uint4(u.v[X + 1].x | ybias, u.v[X + 1].y | ybias, u.v[X + 1].z | ybias,
                                    u.v[X + 1].w | ybias);
        } else {
#pragma unroll
            for (int j = 0; j < D; ++j) u.v[j] = unpackw(c.v[j], w[j]);
            u.v[X] = unpackw(c.v[X], w[X]);
            u.v[X + 1] = unpackw(c.v[X + 1], w[X + 1], ybias);
        }
    }
};

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t pk_mad(uint32_t a, uint32_t b, uint32_t c)
{
    const u16x2 r = __builtin_bit_cast(u16x2, a) * __builtin_bit_cast(u16x2, b) + __builtin_bit_cast(u16x2, c);
    return __builtin_bit_cast(uint32_t, r);
}

__device__ __forceinline__ uint32_t word_of(const uint4 &v, int q)
{
    return q == 0 ? v.x : q == 1 ? v.y : q == 2 ? v.z : v.w;
}

// Rows [r_begin, r_end) of test T into the sink, |Z| = D known at compile time.
// Loads for the next step are issued before the current step is counted.
template <int D, int kSink, int NW>
__device__ __forceinline__ void count_rows_d(const DevInst &I, const DevTest &T, const Shared &sh,
                                             int64_t r_begin, int64_t r_end, int32_t *h, uint32_t ksh)
{
    const ColSet<D, NW> cs(I, T, sh);
    uint32_t zc[D > 0 ? D : 1];
#pragma unroll
    for (int j = 0; j < D; ++j) zc[j] = (uint32_t)sh.zcard[j];
    const uint32_t cx = (uint32_t)T.cx, cy = (uint32_t)T.cy;
    constexpr int X = (D > 0 ? D : 1);
    int64_t r = r_begin + (int64_t)threadIdx.x * kRowsPerThread;
    // two register buffers used in turn: the loads of the next step are in flight
    // while the current one is counted, and no register copy waits for them
    auto step = [&](const Cols<D, NW> &raw, int64_t r) {
        Cols<D> cur;
        cs.unpack(raw, cur);
        if constexpr (kSink == kSinkShared) {
            // every key is below kLdsCells, and so is its word offset key << ksh in the
            // column layouts: its byte offset stays below 2^15.  The byte offsets of two
            // rows are formed at once in the 16-bit halves of a word (v_pk_mad_u16): rows
            // 4q and 4q + 2 from the even bytes of column word q, rows 4q + 1 and 4q + 3
            // from the odd ones; partial keys only grow toward the final key, so no half
            // carries into the other
            const uint32_t m2 = (4u << ksh) * 0x10001u;
            uint32_t zc2[D > 0 ? D : 1];
#pragma unroll
            for (int j = 0; j < D; ++j) zc2[j] = zc[j] * 0x10001u;
            const uint32_t cx2 = cx * 0x10001u, cy2 = cy * 0x10001u;
            const int64_t nv = r_end - r;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                uint32_t k[2];
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    auto col = [&](int j) { return (word_of(cur.v[j], q) >> (8 * e)) & 0x00ff00ffu; };
                    uint32_t kk = (D > 0) ? col(0) : col(X);
#pragma unroll
                    for (int j = 1; j < D; ++j) kk = pk_mad(kk, zc2[j], col(j));
                    if (D > 0) kk = pk_mad(kk, cx2, col(X));
                    kk = pk_mad(kk, cy2, col(X + 1));
                    k[e] = pk_mad(kk, m2, 0u);
                }
                // rows 4q + 0, 4q + 1, 4q + 2, 4q + 3
                const uint32_t off[4] = {k[0] & 0xffffu, k[1] & 0xffffu, k[0] >> 16, k[1] >> 16};
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (nv >= kRowsPerThread || 4 * q + i < nv)
                        atomicAdd(reinterpret_cast<int32_t *>(reinterpret_cast<char *>(h) + off[i]), 1);
            }
            return;
        }
        uint32_t key[kRowsPerThread];
#pragma unroll
        for (int b = 0; b < kRowsPerThread; ++b) {
            uint32_t k = 0;
#pragma unroll
            for (int j = 0; j < D; ++j)
                k = (kSink == kSinkGlobal) ? k * zc[j] + byte_of(cur.v[j], b) : mad24(k, zc[j], byte_of(cur.v[j], b));
            if (kSink == kSinkGlobal) k = (k * cx + byte_of(cur.v[X], b)) * cy + byte_of(cur.v[X + 1], b);
            else k = mad24(mad24(k, cx, byte_of(cur.v[X], b)), cy, byte_of(cur.v[X + 1], b));
            key[b] = k;
        }
        const int64_t nv = r_end - r;
        if (nv >= kRowsPerThread) {
#pragma unroll
            for (int b = 0; b < kRowsPerThread; ++b) put<kSink>(h, key[b], sh, ksh);
        } else {
#pragma unroll
            for (int b = 0; b < kRowsPerThread; ++b)
                if (b < nv) put<kSink>(h, key[b], sh, ksh);
        }
    };
    Cols<D, NW> b0, b1;
    if (r < r_end) cs.load(r, b0);
    while (r < r_end) {
        // the loads are unconditional (clamped to a valid step): no branch around them
        const int64_t r1 = r + kRowsPerStep;
        cs.load((r1 < r_end) ? r1 : r, b1);
        step(b0, r);
        if (r1 >= r_end) break;
        const int64_t r2 = r1 + kRowsPerStep;
        cs.load((r2 < r_end) ? r2 : r1, b0);
        step(b1, r1);
        r = r2;
    }
}

// Runtime |Z| (d > 3).
template <int kSink>
__device__ __forceinline__ void count_rows_any(const DevInst &I, const DevTest &T, const Shared &sh,
                                               int64_t r_begin, int64_t r_end, int32_t *h, uint32_t ksh)
{
    const uint8_t *colx = I.pcols + (int64_t)T.x * I.stride;
    const uint8_t *coly = I.pcols + (int64_t)T.y * I.stride;
    const uint32_t wx = uniform(sh.wx), wy = uniform(sh.wy);
    const uint32_t cx = (uint32_t)T.cx, cy = (uint32_t)T.cy;
    const int d = T.d;
    for (int64_t r = r_begin + (int64_t)threadIdx.x * kRowsPerThread; r < r_end; r += kRowsPerStep) {
        uint32_t key[kRowsPerThread];
#pragma unroll
        for (int b = 0; b < kRowsPerThread; ++b) key[b] = 0;
        for (int j = 0; j < d; ++j) {
            const uint32_t wz = uniform(sh.zw[j]);
            const uint4 v = unpackw(ldw<4>(I.pcols + (int64_t)sh.zcol[j] * I.stride, wshift(wz), r), wz);
            const uint32_t c = (uint32_t)sh.zcard[j];
#pragma unroll
            for (int b = 0; b < kRowsPerThread; ++b)
                key[b] = (kSink == kSinkGlobal) ? key[b] * c + byte_of(v, b) : mad24(key[b], c, byte_of(v, b));
        }
        const uint4 vx = unpackw(ldw<4>(colx, wshift(wx), r), wx);
        const uint4 vy = unpackw(ldw<4>(coly, wshift(wy), r), wy);
#pragma unroll
        for (int b = 0; b < kRowsPerThread; ++b) {
            if (kSink == kSinkGlobal) key[b] = (key[b] * cx + byte_of(vx, b)) * cy + byte_of(vy, b);
            else key[b] = mad24(mad24(key[b], cx, byte_of(vx, b)), cy, byte_of(vy, b));
        }
        const int64_t nv = r_end - r;
        if (nv >= kRowsPerThread) {
#pragma unroll
            for (int b = 0; b < kRowsPerThread; ++b) put<kSink>(h, key[b], sh, ksh);
        } else {
#pragma unroll
            for (int b = 0; b < kRowsPerThread; ++b)
                if (b < nv) put<kSink>(h, key[b], sh, ksh);
        }
    }
}

// Per-lane sinks (kSinkLane: tables <= 64 cells, kSinkLaneW: <= 96, or <= 128 in its
// 64 KiB instance kSinkLaneW128): a private
// histogram per lane in 16-bit LDS counters.  The two waves of a wave pair share each
// 32-bit word, the even wave counting in the low half and the odd wave in the high half,
// so a wave's increment is a constant (1 or 1 << 16).  Layout [pair][cell][lane]: the
// 64 lanes of a wave always hit 64 distinct words in distinct banks, so neither bank
// conflicts nor atomic collisions occur however skewed the data.
//
// Every key is below kCells <= 128, so the keys of 4 rows are formed at once in one
// 32-bit word, one byte per row: the Horner steps run as packed 16-bit multiply-adds
// (v_pk_mad_u16) on the unpacked column words (one row per byte), and no byte ever
// carries into the next because every partial key stays below 256.  A row's LDS byte
// offset comes from that word with one byte permute (lane_put4).  Rows past the end
// of the data only ever carry into later rows of the same word, which are past the
// end as well and are not counted.
// Counts the (up to) 4 rows whose keys are the bytes of K; rows >= nvalid are skipped.
// A row's byte offset in the counters is pair * kCells * 256 + key * 256 + lane * 4: every
// key byte already carries the wave pair's bias (0, or kCells added to the key; see
// lane_bias), and one v_perm_b32 puts key byte j into byte 1 and lane * 4 (lane4, byte 0)
// into byte 0.
// kShr: the offset is shifted right by 1 (the shared sink's [cell][32] column layout: the
// key byte at byte 1 and (lane & 31) * 8 at byte 0, halved, give key * 128 + (lane & 31) * 4).
template <int kShr = 0>
__device__ __forceinline__ void lane_put4(char *lb, uint32_t lane4, uint32_t inc, uint32_t K, int nvalid)
{
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (j < nvalid) {
            const uint32_t off = __builtin_amdgcn_perm(K, lane4, 0x0C0C0000u | ((4u + (uint32_t)j) << 8)) >> kShr;
            atomicAdd(reinterpret_cast<uint32_t *>(lb + off), inc);
        }
}

// The [cell][16] column layout of 257..512-cell tables with |Z| = 1: K holds the keys of
// the 4 rows over (X, Y) only, x * c(Y) + y < 256, and Z is the page: a row's byte offset
// is (z * c(X) c(Y) + x * c(Y) + y) * 64 + (lane & 15) * 4, the permuted low key
// (lane16 = (lane & 15) * 16 in byte 0) shifted right by 2, plus z * page_bytes.
__device__ __forceinline__ void lane_put4_paged(char *lb, uint32_t lane16, uint32_t K, uint32_t Zw,
                                                uint32_t page_bytes, int nvalid)
{
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (j < nvalid) {
            const uint32_t lo = __builtin_amdgcn_perm(K, lane16, 0x0C0C0000u | ((4u + (uint32_t)j) << 8)) >> 2;
            const uint32_t off = __builtin_amdgcn_ubfe(Zw, 8u * (uint32_t)j, 8u) * page_bytes + lo;
            atomicAdd(reinterpret_cast<uint32_t *>(lb + off), 1u);
        }
}

// The wave pair's bias enters every key byte in one of two ways (the key of a valid row
// is below kCells, the biased key below 2 * kCells <= 256, so no byte carries into the
// next; rows past the end only carry into later rows, which are past the end as well):
//   kCells a power of two: ORed into the Y bytes at unpacking (fused with the unpacking
//     mask).  y < c(Y) <= kCells, so y | kCells == y + kCells;
//   otherwise (96 cells): added to the packed key word after the Horner steps.  An OR
//     would not do there: y | 96 != y + 96 once y >= 32.
template <int kCells>
struct lane_bias {
    static constexpr bool kOr = (kCells & (kCells - 1)) == 0;
};

// Packed keys of the 4 rows of word q, |Z| = D known at compile time.
// c's Y bytes carry the wave pair's bias when it is ORed in at unpacking: adding it to
// every y adds it to every key.
template <int D>
__device__ __forceinline__ uint32_t lane_key4(const Cols<D> &c, const uint32_t *zc2, uint32_t cx2, uint32_t cy2,
                                              int q)
{
    constexpr int X = (D > 0 ? D : 1);
    uint32_t K = 0;
#pragma unroll
    for (int j = 0; j < D; ++j) K = (j == 0) ? word_of(c.v[0], q) : pk_mad(K, zc2[j], word_of(c.v[j], q));
    K = (D == 0) ? word_of(c.v[X], q) : pk_mad(K, cx2, word_of(c.v[X], q));
    return pk_mad(K, cy2, word_of(c.v[X + 1], q));
}

// kOr: the bias is ORed into the Y bytes, else added to each key word (lane_bias).
// kPage (D = 1): the [cell][16] layout with Z as the page (lane_put4_paged; lane4 is
// lane16 there, inc 1 and bias 0).
template <int D, int NW, bool kOr, int kShr = 0, bool kPage = false>
__device__ __forceinline__ void count_rows_lane_d(const DevInst &I, const DevTest &T, const Shared &sh,
                                                  int64_t r_begin, int64_t r_end, char *lb, uint32_t lane4,
                                                  uint32_t inc, uint32_t bias, uint32_t page_bytes = 0u)
{
    static_assert(!kPage || D == 1, "paged keys take |Z| = 1");
    const uint32_t ybias = kOr ? bias : 0u, kbias = kOr ? 0u : bias;
    const ColSet<D, NW> cs(I, T, sh);
    uint32_t zc2[D > 0 ? D : 1];
#pragma unroll
    for (int j = 0; j < D; ++j) zc2[j] = (uint32_t)sh.zcard[j] * 0x10001u;
    const uint32_t cx2 = (uint32_t)T.cx * 0x10001u, cy2 = (uint32_t)T.cy * 0x10001u;
    int64_t r = r_begin + (int64_t)threadIdx.x * kRowsPerThread;
    auto step = [&](const Cols<D, NW> &raw, int64_t r) {
        Cols<D> cur;
        cs.unpack(raw, cur, ybias);
        const int64_t nv = r_end - r;
        if constexpr (kPage) {
            constexpr int X = 1;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                lane_put4_paged(lb, lane4, pk_mad(word_of(cur.v[X], q), cy2, word_of(cur.v[X + 1], q)),
                                word_of(cur.v[0], q), page_bytes, nv >= kRowsPerThread ? 4 : (int)nv - 4 * q);
        } else if (nv >= kRowsPerThread) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
                lane_put4<kShr>(lb, lane4, inc, lane_key4<D>(cur, zc2, cx2, cy2, q) + kbias, 4);
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q)
                lane_put4<kShr>(lb, lane4, inc, lane_key4<D>(cur, zc2, cx2, cy2, q) + kbias, (int)nv - 4 * q);
        }
    };
    Cols<D, NW> b0, b1;
    if (r < r_end) cs.load(r, b0);
    while (r < r_end) {
        const int64_t r1 = r + kRowsPerStep;
        cs.load((r1 < r_end) ? r1 : r, b1);
        step(b0, r);
        if (r1 >= r_end) break;
        const int64_t r2 = r1 + kRowsPerStep;
        cs.load((r2 < r_end) ? r2 : r1, b0);
        step(b1, r1);
        r = r2;
    }
}

// Runtime |Z|: d > 3, or a column of the test kept at 8 bits.
template <bool kOr, int kShr = 0>
__device__ __forceinline__ void count_rows_lane_any(const DevInst &I, const DevTest &T, const Shared &sh,
                                                    int64_t r_begin, int64_t r_end, char *lb, uint32_t lane4,
                                                    uint32_t inc, uint32_t bias)
{
    const uint32_t ybias = kOr ? bias : 0u, kbias = kOr ? 0u : bias;
    const uint8_t *colx = I.pcols + (int64_t)T.x * I.stride;
    const uint8_t *coly = I.pcols + (int64_t)T.y * I.stride;
    const uint32_t wx = uniform(sh.wx), wy = uniform(sh.wy);
    const uint32_t cx2 = (uint32_t)T.cx * 0x10001u, cy2 = (uint32_t)T.cy * 0x10001u;
    const int d = T.d;
    for (int64_t r = r_begin + (int64_t)threadIdx.x * kRowsPerThread; r < r_end; r += kRowsPerStep) {
        uint32_t K[4] = {0u, 0u, 0u, 0u};
        for (int j = 0; j < d; ++j) {
            const uint32_t wz = uniform(sh.zw[j]);
            const uint4 v = unpackw(ldw<4>(I.pcols + (int64_t)sh.zcol[j] * I.stride, wshift(wz), r), wz);
            const uint32_t c2 = (uint32_t)sh.zcard[j] * 0x10001u;
#pragma unroll
            for (int q = 0; q < 4; ++q) K[q] = pk_mad(K[q], c2, word_of(v, q));
        }
        const uint4 vx = unpackw(ldw<4>(colx, wshift(wx), r), wx);
        const uint4 vy = unpackw(ldw<4>(coly, wshift(wy), r), wy, ybias);
        const int64_t nv = r_end - r;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t k = pk_mad(pk_mad(K[q], cx2, word_of(vx, q)), cy2, word_of(vy, q)) + kbias;
            lane_put4<kShr>(lb, lane4, inc, k, nv >= kRowsPerThread ? 4 : (int)nv - 4 * q);
        }
    }
}

// blk: the workgroup's counters, [pair][cell][lane] 32-bit words, kCells cells a pair.
// |Z| <= 3 takes loops with |Z| known at compile time: with 8-byte loads when every
// column of the test is packed below 8 bits, 16-byte loads otherwise.
// Row loops by column form.  kPacked (some column of the context is packed): tests whose
// columns are all packed (kTFNarrow) with |Z| <= 3 take loops with |Z| known at compile
// time and 8-byte loads; the others (d > 3, or a column of more than 16 states) the
// runtime-|Z| loop.  Otherwise (uint8 columns) |Z| <= 3 takes compile-time loops with
// 16-byte loads and nothing to unpack.  One form per kernel keeps its registers low.
template <int kCells, bool kPacked>
__device__ __forceinline__ void count_rows_lane(const DevInst &I, const DevTest &T, const Shared &sh,
                                                int64_t r_begin, int64_t r_end, uint32_t *blk)
{
    char *lb = reinterpret_cast<char *>(blk);
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t inc = (wave & 1) ? 0x10000u : 1u;
    const uint32_t bias = (wave >> 1) ? (uint32_t)kCells * 0x01010101u : 0u;
    const uint32_t lane4 = lane * 4u;
    constexpr int NW = kPacked ? 2 : 8;
    constexpr bool kOr = lane_bias<kCells>::kOr;
    if (T.d <= 3 && (!kPacked || (uniform(T.flags) & kTFNarrow))) {
        switch (T.d) {
        case 0: count_rows_lane_d<0, NW, kOr>(I, T, sh, r_begin, r_end, lb, lane4, inc, bias); return;
        case 1: count_rows_lane_d<1, NW, kOr>(I, T, sh, r_begin, r_end, lb, lane4, inc, bias); return;
        case 2: count_rows_lane_d<2, NW, kOr>(I, T, sh, r_begin, r_end, lb, lane4, inc, bias); return;
        default: count_rows_lane_d<3, NW, kOr>(I, T, sh, r_begin, r_end, lb, lane4, inc, bias); return;
        }
    }
    count_rows_lane_any<kOr>(I, T, sh, r_begin, r_end, lb, lane4, inc, bias);
}

// The column layouts of the per-lane sink's launch: [cell][64] int32 words up to 128 cells,
// [cell][32] up to 256, [cell][16] up to 512, with the per-lane key arithmetic (keys below
// 256 formed four rows at a time, a row's LDS offset by one byte permute, and a shift for
// 32 columns; for 16 columns the key over (X, Y) and Z as the page), every wave adding 1.
// |Z| = 1 with every column in the launch's form and c(X) c(Y) <= 256 (the host routes only
// such tests here): with |Z| = 0, 2 and 3 as well the kernel exceeds 128 VGPRs.
template <bool kPacked>
__device__ __forceinline__ void count_rows_cols(const DevInst &I, const DevTest &T, const Shared &sh,
                                                int64_t r_begin, int64_t r_end, int32_t *lds, int ncol)
{
    char *lb = reinterpret_cast<char *>(lds);
    const uint32_t lane = threadIdx.x & 63;
    constexpr int NW = kPacked ? 2 : 8;
    if (ncol == 64) count_rows_lane_d<1, NW, true, 0>(I, T, sh, r_begin, r_end, lb, lane * 4u, 1u, 0u);
    else if (ncol == 32) count_rows_lane_d<1, NW, true, 1>(I, T, sh, r_begin, r_end, lb, (lane & 31u) * 8u, 1u, 0u);
    else
        count_rows_lane_d<1, NW, true, 0, true>(I, T, sh, r_begin, r_end, lb, (lane & 15u) * 16u, 1u, 0u,
                                                (uint32_t)(T.cx * T.cy) * 64u);
}

// |Z| <= 3 with compile-time loops of the context's column form (the caller checks that
// a packed context's test has all its columns packed)
// Largest |Z| with a compile-time loop: 3 for the LDS sinks held to 128 VGPRs (4 for the
// shared sink's packed form, whose narrow loads leave room); 8 for the z-slice and
// device-memory sinks, whose LDS (or none) leaves them twice the registers and whose
// deep-level tests would otherwise wait on every step's loads.
template <int kSink, bool kPacked>
constexpr int kMaxFixedD =
    (kSink == kSinkSlice || kSink == kSinkGlobal) ? 8 : (kSink == kSinkShared && kPacked) ? 4 : 3;

template <int kSink, bool kPacked>
__device__ __forceinline__ void count_rows_fixed(const DevInst &I, const DevTest &T, const Shared &sh,
                                                 int64_t r_begin, int64_t r_end, int32_t *h, uint32_t ksh)
{
    constexpr int NW = kPacked ? 2 : 8;
    switch (T.d) {
    case 0: count_rows_d<0, kSink, NW>(I, T, sh, r_begin, r_end, h, ksh); break;
    case 1: count_rows_d<1, kSink, NW>(I, T, sh, r_begin, r_end, h, ksh); break;
    case 2: count_rows_d<2, kSink, NW>(I, T, sh, r_begin, r_end, h, ksh); break;
    case 3: count_rows_d<3, kSink, NW>(I, T, sh, r_begin, r_end, h, ksh); break;
    default:
        if constexpr (kMaxFixedD<kSink, kPacked> == 4) {
            count_rows_d<4, kSink, NW>(I, T, sh, r_begin, r_end, h, ksh);
        } else if constexpr (kMaxFixedD<kSink, kPacked> > 4) {
            switch (T.d) {
            case 4: count_rows_d<4, kSink, NW>(I, T, sh, r_begin, r_end, h, ksh); break;
            case 5: count_rows_d<5, kSink, NW>(I, T, sh, r_begin, r_end, h, ksh); break;
            case 6: count_rows_d<6, kSink, NW>(I, T, sh, r_begin, r_end, h, ksh); break;
            case 7: count_rows_d<7, kSink, NW>(I, T, sh, r_begin, r_end, h, ksh); break;
            default: count_rows_d<8, kSink, NW>(I, T, sh, r_begin, r_end, h, ksh); break;
            }
        }
        break;
    }
}

template <bool kPacked, int kSink = kSinkShared>
__device__ __forceinline__ bool fixed_loop(const DevTest &T)
{
    return T.d <= kMaxFixedD<kSink, kPacked> && (!kPacked || (uniform(T.flags) & kTFNarrow));
}

template <int kSink, bool kPacked>
__device__ __forceinline__ void count_rows(const DevInst &I, const DevTest &T, const Shared &sh,
                                           int64_t r_begin, int64_t r_end, int32_t *h, uint32_t ksh = 0u)
{
    const uint32_t k = uniform((int32_t)ksh);
    if (fixed_loop<kPacked, kSink>(T)) count_rows_fixed<kSink, kPacked>(I, T, sh, r_begin, r_end, h, k);
    else count_rows_any<kSink>(I, T, sh, r_begin, r_end, h, k);
}

// ---------------------------------------------------------------- partitioned large tables

constexpr int kPartSteps = kPartRows / kRowsPerStep;  // workgroup steps of a partitioning item
static_assert(kPartRows % kRowsPerStep == 0, "whole steps per partitioning item");
static_assert(3 * kPartStage + 1 <= kSmallCells && kPartStage <= 128, "a slice item's staged segments");

// Slice of a key, key / sc, and the key within it: an FP32 reciprocal corrected by the
// remainder.  Keys stay below 2^26 (at most kPartMaxSlices slices of at most kSliceCells
// cells) and slices hold more than kSliceCells / 2 cells, so the FP32 quotient is within
// one of the true one.
__device__ __forceinline__ uint32_t part_slice_of(uint32_t key, uint32_t sc, float inv, uint32_t &rel)
{
    uint32_t q = (uint32_t)((float)key * inv);
    int32_t r = (int32_t)(key - q * sc);
    if (r < 0) {
        --q;
        r += (int32_t)sc;
    } else if (r >= (int32_t)sc) {
        ++q;
        r -= (int32_t)sc;
    }
    rel = (uint32_t)r;
    return q;
}

// Keys of the thread's rows of one partitioning item (rows [r0, r1), row r0 + s * 4096 +
// 16 * thread + b in key[16 s + b]; ~0u for rows at or past r1), |Z| = D at compile time;
// the next step's columns load while the current one is keyed.
template <int D, int NW>
__device__ __forceinline__ void part_keys_d(const DevInst &I, const DevTest &T, const Shared &sh, int64_t r0,
                                            int64_t r1, uint32_t (&key)[kPartSteps * kRowsPerThread])
{
    const ColSet<D, NW> cs(I, T, sh);
    uint32_t zc[D > 0 ? D : 1];
#pragma unroll
    for (int j = 0; j < D; ++j) zc[j] = (uint32_t)sh.zcard[j];
    const uint32_t cx = (uint32_t)T.cx, cy = (uint32_t)T.cy;
    constexpr int X = (D > 0 ? D : 1);
    const int64_t r = r0 + (int64_t)threadIdx.x * kRowsPerThread;
    Cols<D, NW> raw[2];
    cs.load((r < r1) ? r : r0, raw[0]);
#pragma unroll
    for (int st = 0; st < kPartSteps; ++st) {
        const int64_t rs = r + (int64_t)st * kRowsPerStep;
        if (st + 1 < kPartSteps) {
            const int64_t rn = rs + kRowsPerStep;
            cs.load((rn < r1) ? rn : r0, raw[(st + 1) & 1]);
        }
        Cols<D> cur;
        cs.unpack(raw[st & 1], cur);
#pragma unroll
        for (int b = 0; b < kRowsPerThread; ++b) {
            uint32_t k = 0;
#pragma unroll
            for (int j = 0; j < D; ++j) k = k * zc[j] + byte_of(cur.v[j], b);
            k = (k * cx + byte_of(cur.v[X], b)) * cy + byte_of(cur.v[X + 1], b);
            key[st * kRowsPerThread + b] = (rs + b < r1) ? k : ~0u;
        }
    }
}

// Exclusive scan of the item's per-slice row counts cnt[0, S) (LDS): each slice's
// (offset, count) segment goes to seg, and cnt[s] becomes the slice's fill position.
__device__ void part_scan(uint32_t *cnt, int S, uint32_t *seg, Shared &sh)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int per = (S + kWG - 1) / kWG;
    const int i0 = tid * per;
    uint32_t loc = 0;
    for (int j = 0; j < per; ++j)
        if (i0 + j < S) loc += cnt[i0 + j];
    uint32_t inc = loc;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t v = __shfl_up(inc, o, 64);
        if (lane >= o) inc += v;
    }
    if (lane == 63) sh.red_i[wave] = (int32_t)inc;
    __syncthreads();
    uint32_t off = inc - loc;
    for (int w = 0; w < wave; ++w) off += (uint32_t)sh.red_i[w];
    for (int j = 0; j < per; ++j) {
        const int i = i0 + j;
        if (i < S) {
            const uint32_t c = cnt[i];
            seg[2 * i] = off;
            seg[2 * i + 1] = c;
            cnt[i] = off;
            off += c;
        }
    }
}

// Exclusive scan, in place, of the S * kWG private slice counters cnt[s * kWG + thread]
// in that order; thread t sums words [t S, t S + S).  Leaves the total in
// sh.red_i[kWG / 64 - 1] (as the last wave's inclusive sum plus the waves before).
__device__ void part_scan_flat(uint32_t *cnt, int S, Shared &sh)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i0 = tid * S;
    uint32_t loc = 0;
    for (int j = 0; j < S; ++j) loc += cnt[i0 + j];
    uint32_t inc = loc;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t v = __shfl_up(inc, o, 64);
        if (lane >= o) inc += v;
    }
    if (lane == 63) sh.red_i[wave] = (int32_t)inc;
    __syncthreads();
    uint32_t off = inc - loc;
    for (int w = 0; w < wave; ++w) off += (uint32_t)sh.red_i[w];
    for (int j = 0; j < S; ++j) {
        const uint32_t c = cnt[i0 + j];
        cnt[i0 + j] = off;
        off += c;
    }
    __syncthreads();
    if (tid == kWG - 1) sh.red_i[kWG / 64 - 1] = (int32_t)off;  // the total
}

// One partitioning item: row chunk `chunk` of test T.  Every row's key once; its slice's
// row count (LDS cnt, zero on entry); the scan; then each row's slice-relative key placed
// at its slice's next position within the chunk's kPartRows keys.
template <bool kPacked>
__device__ void part_item(const LaunchArgs &a, const DevInst &I, const DevTest &T, Shared &sh, int64_t chunk,
                          uint32_t *cnt, uint16_t *stage)
{
    const int64_t r0 = chunk * kPartRows;
    const int64_t r1 = (r0 + kPartRows < I.n_rows) ? r0 + kPartRows : I.n_rows;
    constexpr int NW = kPacked ? 2 : 8;
    constexpr int K = kPartSteps * kRowsPerThread;
    uint32_t key[K];
    switch (T.d) {
    case 0: part_keys_d<0, NW>(I, T, sh, r0, r1, key); break;
    case 1: part_keys_d<1, NW>(I, T, sh, r0, r1, key); break;
    case 2: part_keys_d<2, NW>(I, T, sh, r0, r1, key); break;
    case 3: part_keys_d<3, NW>(I, T, sh, r0, r1, key); break;
    case 4: part_keys_d<4, NW>(I, T, sh, r0, r1, key); break;
    case 5: part_keys_d<5, NW>(I, T, sh, r0, r1, key); break;
    case 6: part_keys_d<6, NW>(I, T, sh, r0, r1, key); break;
    case 7: part_keys_d<7, NW>(I, T, sh, r0, r1, key); break;
    default: part_keys_d<8, NW>(I, T, sh, r0, r1, key); break;  // the host routes |Z| <= 8 only
    }
    const uint32_t sc = (uint32_t)T.slice_cells;
    const float inv = 1.0f / (float)sc;
    int32_t *base = a.part + T.ws_off;
    const PartLayout L = part_layout(T.dimz, T.cx, T.cy, T.n_slices, I.n_rows);
    uint32_t *seg = reinterpret_cast<uint32_t *>(base + L.seg) + chunk * T.n_slices * 2;
    const int S = T.n_slices;
    if (S <= kPartPrivSlices) {
        // few slices: a private counter per slice and thread, cnt[s * kWG + thread] (a wave's
        // 64 adds to one slice would otherwise serialise on one word); the scan over them in
        // that order places each thread's keys of a slice together, threads in order
        const int tid = threadIdx.x;
        for (int s2 = 0; s2 < S; ++s2) cnt[s2 * kWG + tid] = 0u;  // the thread's own words
#pragma unroll
        for (int i = 0; i < K; ++i)
            if (key[i] != ~0u) {
                uint32_t rel;
                ++cnt[part_slice_of(key[i], sc, inv, rel) * kWG + tid];
            }
        __syncthreads();
        part_scan_flat(cnt, S, sh);
        __syncthreads();
        uint32_t so = 0, sn = 0;  // slice tid's segment (read before the counters move on)
        if (tid < S) {
            so = cnt[tid * kWG];
            sn = ((tid + 1 < S) ? cnt[(tid + 1) * kWG] : (uint32_t)sh.red_i[kWG / 64 - 1]) - so;
        }
        __syncthreads();
        if (tid < S) {
            seg[2 * tid] = so;
            seg[2 * tid + 1] = sn;
        }
#pragma unroll
        for (int i = 0; i < K; ++i)
            if (key[i] != ~0u) {
                uint32_t rel;
                const uint32_t s2 = part_slice_of(key[i], sc, inv, rel);
                stage[cnt[s2 * kWG + tid]++] = (uint16_t)rel;
            }
    } else {
#pragma unroll
        for (int i = 0; i < K; ++i)
            if (key[i] != ~0u) {
                uint32_t rel;
                atomicAdd(&cnt[part_slice_of(key[i], sc, inv, rel)], 1u);
            }
        __syncthreads();
        part_scan(cnt, S, seg, sh);
        __syncthreads();
        // placed in LDS first (stage: kPartRows keys), then written out in 16-byte stores
#pragma unroll
        for (int i = 0; i < K; ++i)
            if (key[i] != ~0u) {
                uint32_t rel;
                const uint32_t s2 = part_slice_of(key[i], sc, inv, rel);
                stage[atomicAdd(&cnt[s2], 1u)] = (uint16_t)rel;
            }
    }
    __syncthreads();
    uint4 *out = reinterpret_cast<uint4 *>(reinterpret_cast<uint16_t *>(base + L.keys) + chunk * kPartRows);
    const int n16 = (int)((r1 - r0 + 7) / 8);  // 8 keys per store; keys past the rows are never read
    for (int i = threadIdx.x; i < n16; i += kWG) out[i] = reinterpret_cast<const uint4 *>(stage)[i];
}

// ---------------------------------------------------------------- bit-plane counting

// One-hot masks of the P values of a variable over words [w0, w0 + kBitsWords) from its
// nbits(P) planes: value p's mask is the AND of every plane or its complement, by p's bits.
template <int P>
__device__ __forceinline__ void onehot_words(const uint32_t *S, int64_t pw, int64_t w0, uint4 (&o)[P])
{
    constexpr int NB = nbits_of(P);
    uint4 b[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) b[i] = *reinterpret_cast<const uint4 *>(S + i * pw + w0);
#pragma unroll
    for (int p = 0; p < P; ++p) {
        uint4 m = make_uint4(~0u, ~0u, ~0u, ~0u);
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const bool one = (p >> i) & 1;
            m.x &= one ? b[i].x : ~b[i].x;
            m.y &= one ? b[i].y : ~b[i].y;
            m.z &= one ? b[i].z : ~b[i].z;
            m.w &= one ? b[i].w : ~b[i].w;
        }
        o[p] = m;
    }
}

// The roles of a bit-plane test's variables, sorted by cardinality: A (P states), B (Q),
// C (R), and each one's multiplier in the n[z][x][y] cell index (x: c(Y), y: 1, z:
// c(X) c(Y)).
struct BitsRoles {
    int vA, vB, vC;
    int P, Q, R;
    int mA, mB, mC;
};

__device__ __forceinline__ BitsRoles bits_roles(const DevTest &T, const Shared &sh)
{
    BitsRoles o;
    int v0 = T.x, c0 = T.cx, m0 = T.cy;
    int v1 = T.y, c1 = T.cy, m1 = 1;
    int v2 = sh.zcol[0], c2 = sh.zcard[0], m2 = T.cx * T.cy;
    auto swap3 = [](int &va, int &ca, int &ma, int &vb, int &cb, int &mb) {
        if (ca > cb) {
            int t = va; va = vb; vb = t;
            t = ca; ca = cb; cb = t;
            t = ma; ma = mb; mb = t;
        }
    };
    swap3(v0, c0, m0, v1, c1, m1);
    swap3(v1, c1, m1, v2, c2, m2);
    swap3(v0, c0, m0, v1, c1, m1);
    o.vA = v0; o.vB = v1; o.vC = v2;
    o.P = c0; o.Q = c1; o.R = c2;
    o.mA = m0; o.mB = m1; o.mC = m2;
    return o;
}

// A |Z| = 1 table from bit planes: cell (p, q, r) of the roles' table counts the rows
// where A = p, B = q and C = r.  Per 32-row word: one AND per (p, q), the value masks of C,
// and one AND and one population count per cell -- no per-row LDS operation.  Each thread
// takes kBitsWords consecutive words per step and adds each cell's count over them to its
// lane-private 16-bit counter of cell (r P + p) Q + q (the per-lane sink's [pair][cell][lane]
// layout, even wave low half, odd wave high; the offsets are compile-time, the reduction
// maps the cells to n[z][x][y]).  Rows at or past n_rows read as value 0 in every plane, so
// the words there are masked.
template <int P, int Q, int NBR>
__device__ __forceinline__ void count_bits_pq(const uint32_t *A, const uint32_t *B, const uint32_t *C, int64_t pw,
                                              int R, int64_t w_b, int64_t w_e, int64_t n_rows, uint32_t *pl)
{
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    uint32_t *base = pl + (wave >> 1) * (kLaneCells * 64) + lane;
    const uint32_t sh = (wave & 1) ? 16u : 0u;
    for (int64_t w0 = w_b + (int64_t)tid * kBitsWords; w0 < w_e; w0 += (int64_t)kWG * kBitsWords) {
        uint4 a[P], b[Q];
        onehot_words<P>(A, pw, w0, a);
        onehot_words<Q>(B, pw, w0, b);
        if ((w0 + kBitsWords) * 32 > n_rows) {  // the instance's last rows: words past them clear
            uint32_t v[kBitsWords];
#pragma unroll
            for (int j = 0; j < kBitsWords; ++j) {
                const int64_t left = n_rows - (w0 + j) * 32;
                v[j] = left >= 32 ? 0xffffffffu : left <= 0 ? 0u : ((1u << left) - 1u);
            }
#pragma unroll
            for (int p = 0; p < P; ++p) {
                a[p].x &= v[0];
                a[p].y &= v[1];
                a[p].z &= v[2];
                a[p].w &= v[3];
            }
        }
        uint4 ab[P * Q];
#pragma unroll
        for (int p = 0; p < P; ++p)
#pragma unroll
            for (int q = 0; q < Q; ++q)
                ab[p * Q + q] = make_uint4(a[p].x & b[q].x, a[p].y & b[q].y, a[p].z & b[q].z, a[p].w & b[q].w);
        uint4 cb[NBR];
#pragma unroll
        for (int i = 0; i < NBR; ++i) cb[i] = *reinterpret_cast<const uint4 *>(C + i * pw + w0);
#pragma unroll
        for (int r = 0; r < (1 << NBR); ++r) {
            if (r >= R) continue;  // uniform
            uint4 c = make_uint4(~0u, ~0u, ~0u, ~0u);
#pragma unroll
            for (int i = 0; i < NBR; ++i) {
                const bool one = (r >> i) & 1;
                c.x &= one ? cb[i].x : ~cb[i].x;
                c.y &= one ? cb[i].y : ~cb[i].y;
                c.z &= one ? cb[i].z : ~cb[i].z;
                c.w &= one ? cb[i].w : ~cb[i].w;
            }
#pragma unroll
            for (int pq = 0; pq < P * Q; ++pq) {
                const uint4 &u = ab[pq];
                const uint32_t n = __popc(u.x & c.x) + __popc(u.y & c.y) + __popc(u.z & c.z) + __popc(u.w & c.w);
                atomicAdd(base + (r * P * Q + pq) * 64, n << sh);
            }
        }
    }
}

// Rows [r_begin, r_end) of a bit-plane test (kTFBits: |Z| = 1, cardinalities accepted by
// bits_shape) into the per-lane counters; r_begin is a multiple of 4096, so whole words
// are counted.
__device__ __forceinline__ void count_rows_bits(const LaunchArgs &a, const DevInst &I, const DevTest &T,
                                                const Shared &sh, int64_t r_begin, int64_t r_end, uint32_t *pl)
{
    const BitsRoles ro = bits_roles(T, sh);
    const int64_t pw = I.pwords;
    const uint32_t *A = I.planes + (int64_t)a.poff[ro.vA] * pw;
    const uint32_t *B = I.planes + (int64_t)a.poff[ro.vB] * pw;
    const uint32_t *C = I.planes + (int64_t)a.poff[ro.vC] * pw;
    const int64_t w_b = r_begin >> 5, w_e = (r_end + 31) >> 5, nr = I.n_rows;
#define FPGM_BITS_CASE(P, Q, NBR)                                                                     \
    case (P * 16 + Q) * 8 + NBR:                                                                      \
        count_bits_pq<P, Q, NBR>(A, B, C, pw, ro.R, w_b, w_e, nr, pl);                                \
        break;
    switch (uniform((ro.P * 16 + ro.Q) * 8 + nbits_of(ro.R))) {
        // every (P, Q, planes of C) with P Q R <= kLaneCells and R >= Q
        FPGM_BITS_CASE(2, 2, 1)
        FPGM_BITS_CASE(2, 2, 2)
        FPGM_BITS_CASE(2, 2, 3)
        FPGM_BITS_CASE(2, 2, 4)
        FPGM_BITS_CASE(2, 3, 2)
        FPGM_BITS_CASE(2, 3, 3)
        FPGM_BITS_CASE(2, 3, 4)
        FPGM_BITS_CASE(2, 4, 2)
        FPGM_BITS_CASE(2, 4, 3)
        FPGM_BITS_CASE(2, 5, 3)
        FPGM_BITS_CASE(3, 3, 2)
        FPGM_BITS_CASE(3, 3, 3)
        FPGM_BITS_CASE(3, 4, 2)
        FPGM_BITS_CASE(3, 4, 3)
        FPGM_BITS_CASE(4, 4, 2)
    default: break;  // routing (bits_shape) sends no other shape here
    }
#undef FPGM_BITS_CASE
}

// Critical value of the chi-square test at the call's alpha: tabulated (exact FP64
// inversion on the host) for df <= kCritTab, Wilson-Hilferty beyond, whose relative
// error there is below 1e-6 (far inside the recheck band).
// crit_lds: the first n_lds entries of the table in LDS (the counting kernel's copy).
__device__ __forceinline__ double crit_of(const LaunchArgs &a, int df, const double *crit_lds, int n_lds)
{
    if (df <= n_lds) return crit_lds[df - 1];
    if (df <= kCritTab) return a.crit[df - 1];
    const double k = (double)df;
    const double q = 2.0 / (9.0 * k);
    const double t = 1.0 - q + a.z_alpha * sqrt(q);
    return k * t * t * t;
}

// FP64 term with the reference's expression (IndependenceTest.cpp:134-135).
__device__ __forceinline__ double term64(int32_t n, int32_t ni, int32_t nj, int32_t nk)
{
    const double expected = (double)nj * (double)ni / (double)nk;
    return 2.0 * (double)n * log((double)n / expected);
}

// One z slice: marginals (CellTable.cpp:242-250), df contribution
// (IndependenceTest.cpp:96-110) and compensated FP32 terms, by one thread.
__device__ __forceinline__ void zslice32(const int32_t *tz, int cx, int cy, int32_t *ni, int32_t *nj, int &df,
                                         float &s, float &c, float &as)
{
    int32_t nk = 0;
    int alx = 0, aly = 0;
    for (int i = 0; i < cx; ++i) {
        int32_t v = 0;
        for (int j = 0; j < cy; ++j) v += tz[i * cy + j];
        ni[i] = v;
        nk += v;
        alx += (v > 0);
    }
    for (int j = 0; j < cy; ++j) {
        int32_t v = 0;
        for (int i = 0; i < cx; ++i) v += tz[i * cy + j];
        nj[j] = v;
        aly += (v > 0);
    }
    alx = alx >= 1 ? alx : 1;
    aly = aly >= 1 ? aly : 1;
    df += (alx - 1) * (aly - 1);
    if (nk == 0) return;
    for (int i = 0; i < cx; ++i) {
        const int32_t ri = ni[i];
        if (ri == 0) continue;
        for (int j = 0; j < cy; ++j) {
            const int32_t n = tz[i * cy + j];
            if (n == 0) continue;
            const float tm = term32(n, ri, nj[j], nk);
            nadd(s, c, tm);
            as += fabsf(tm);
        }
    }
}

// FP32 decision of a test from its FP32 G^2 (IndependenceTest.cpp:140-154): df == 0 is
// independent; otherwise independent iff G^2 < chi-square critical value, unless G^2 lies
// within the recheck band of it (or the caller asked for exact results), in which case
// the FP64 statistic decides.  Writes the result record; returns 1 when FP64 is needed.
__device__ int decide32(const LaunchArgs &a, int32_t tflags, int df, float g2, float AS, DevResult &R,
                        const double *crit_lds = nullptr, int n_lds = 0)
{
    int need64 = 0;
    uint8_t indep = 0, in_band = 0;
    const bool exact = (tflags & kTFExact) || (a.flags & FPGM_F_EXACT);
    if (df == 0) {
        indep = 1;  // IndependenceTest.cpp:140-142
        need64 = exact;
    } else {
        const double crit = crit_of(a, df, crit_lds, n_lds);
        // FP32 error bound: a few ulp of every term, summed, plus the critical-value
        // approximation beyond the table
        const double bound = 4e-6 * (double)AS + (df > kCritTab ? 1e-5 * crit : 0.0) + 1e-30;
        const double width = fmax((double)a.band * crit, bound);
        // alpha outside (0, 1) has no critical value: the FP64 p-value decides
        in_band = (isnan(crit) || fabs((double)g2 - crit) <= width) ? 1 : 0;
        if ((a.flags & FPGM_F_FP32ONLY) && !isnan(crit)) {
            indep = (double)g2 < crit;
            need64 = exact;
        } else if (in_band || exact) {
            need64 = 1;
        } else {
            indep = (double)g2 < crit;  // p > alpha  <=>  G^2 < critical value
        }
    }
    R.g2_32 = g2;
    R.abs_sum = AS;
    R.df = df;
    R.indep = indep;
    R.in_band = in_band;
    R.exact = 0;
    R.need_p = 0;
    R.g2_64 = NAN;
    R.p64 = (df == 0) ? 1.0 : NAN;
    return need64;
}

// A test decided independent marks its edge with the launch's round: later rounds skip
// the edge's tests, tests of the same round never skip one another.
__device__ __forceinline__ void mark_edge(const LaunchArgs &a, const DevTest &T, bool indep, bool need_p)
{
    if (a.edge_done && T.edge >= 0 && indep && !need_p) atomicMin(&a.edge_done[T.edge], a.round);
}

__device__ __forceinline__ bool skipped(const LaunchArgs &a, const DevTest &T)
{
    return a.edge_done && T.edge >= 0 && a.edge_done[T.edge] < a.round;
}

__device__ __forceinline__ DevResult skipped_result()
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

// n / d for 0 <= n < 2^22 and 1 <= d < 2^22, rd = 1/d in FP32: the FP32 product is within
// n * 2^-23 < 0.5 of n / d, so its truncation is off by at most one, which the
// remainder's sign corrects.  (The small-table statistic divides cell and row indices
// below 63 * 65536 by cardinalities.)
__device__ __forceinline__ int32_t small_div(int32_t n, float rd, int32_t d)
{
    int32_t q = (int32_t)((float)n * rd);
    const int32_t r = n - (int32_t)__umul24((uint32_t)q, (uint32_t)d);
    q += (r >= d) ? 1 : 0;
    q -= (r < 0) ? 1 : 0;
    return q;
}

// Statistic of test t from its full table (LDS or global) with marginal scratch
// marg = [ni: dimz*cx | nj: dimz*cy | nk: dimz].  Tables with many z slices are
// processed one slice per thread in a single pass; small ones cell-parallel.
__device__ void finalize(const LaunchArgs &a, const DevTest &T, int t, const int32_t *tab,
                         int32_t *marg, Shared &sh, const double *crit_lds, int n_crit)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int cx = T.cx, cy = T.cy;
    const int32_t dimz = T.dimz, cxy = cx * cy, cells = dimz * cxy;
    int32_t *ni = marg;
    int32_t *nj = marg + (int64_t)dimz * cx;
    int32_t *nk = nj + (int64_t)dimz * cy;
    const bool zpar = dimz >= 64;

    int df_loc = 0;
    float s = 0.f, c = 0.f, as = 0.f;
    if (zpar) {
        for (int32_t z = tid; z < dimz; z += kWG)
            zslice32(tab + (int64_t)z * cxy, cx, cy, ni + (int64_t)z * cx, nj + (int64_t)z * cy, df_loc, s, c, as);
    } else {
        // one pass for the marginals: rows (z, i) sum their cy cells into n_{x+z}, columns
        // (z, j) their cx cells into n_{+yz}, slices z their cx*cy cells into n_{++z}
        const int32_t nr = dimz * cx, nc = dimz * cy;
        const float rcx = 1.0f / (float)cx, rcy = 1.0f / (float)cy;
        for (int32_t idx = tid; idx < nr + nc + dimz; idx += kWG) {
            int32_t v = 0;
            if (idx < nr) {
                const int32_t *row = tab + (int64_t)idx * cy;
                for (int j = 0; j < cy; ++j) v += row[j];
                ni[idx] = v;
            } else if (idx < nr + nc) {
                const int32_t k = idx - nr, z = small_div(k, rcy, cy), j = k - z * cy;
                const int32_t *col = tab + (int64_t)z * cxy + j;
                for (int i = 0; i < cx; ++i) v += col[i * cy];
                nj[k] = v;
            } else {
                const int32_t z = idx - nr - nc;
                const int32_t *tz = tab + (int64_t)z * cxy;
                for (int k = 0; k < cxy; ++k) v += tz[k];
                nk[z] = v;
            }
        }
        __syncthreads();
        // terms cell-parallel (row r = z * cx + i of cell r * cy + j)
        for (int32_t cell = tid; cell < cells; cell += kWG) {
            const int32_t n = tab[cell];
            if (n == 0) continue;
            const int32_t r = small_div(cell, rcy, cy), j = cell - r * cy;
            const int32_t z = small_div(r, rcx, cx);
            const float tm = term32(n, ni[r], nj[z * cy + j], nk[z]);
            nadd(s, c, tm);
            as += fabsf(tm);
        }
        // the adjusted degrees of freedom, one slice per thread from the last thread down
        // (those count one cell fewer when the cells do not fill the last pass)
        for (int32_t z = kWG - 1 - tid; z < dimz; z += kWG) {
            int alx = 0, aly = 0;
            for (int i = 0; i < cx; ++i) alx += (ni[z * cx + i] > 0);
            for (int j = 0; j < cy; ++j) aly += (nj[z * cy + j] > 0);
            alx = alx >= 1 ? alx : 1;
            aly = aly >= 1 ? aly : 1;
            df_loc += (alx - 1) * (aly - 1);
        }
    }
    df_loc = wave_sum_i(df_loc);
    wave_nsum(s, c);
    as = wave_sum_f(as);
    if (lane == 0) {
        sh.red_i[wave] = df_loc;
        sh.red_f[3 * wave] = s;
        sh.red_f[3 * wave + 1] = c;
        sh.red_f[3 * wave + 2] = as;
    }
    __syncthreads();
    if (tid == 0) {
        int df = 0;
        float S = 0.f, Cc = 0.f, AS = 0.f;
        for (int w = 0; w < kWG / 64; ++w) {
            df += sh.red_i[w];
            nadd(S, Cc, sh.red_f[3 * w]);
            Cc += sh.red_f[3 * w + 1];
            AS += sh.red_f[3 * w + 2];
        }
        sh.df = df;
        DevResult R;
        sh.need64 = decide32(a, T.flags, df, S + Cc, AS, R, crit_lds, n_crit);
        a.res[t] = R;
        sh.indep = R.indep;
    }
    __syncthreads();
    if (sh.need64) {
        // FP64 G^2, same term order within each z slice as IndependenceTest.cpp:112-137
        double g = 0.0;
        if (zpar) {
            for (int32_t z = tid; z < dimz; z += kWG) {
                const int32_t *tz = tab + (int64_t)z * cxy;
                const int32_t *zi = ni + (int64_t)z * cx, *zj = nj + (int64_t)z * cy;
                int32_t nkz = 0;
                for (int i = 0; i < cx; ++i) nkz += zi[i];
                if (nkz == 0) continue;
                for (int i = 0; i < cx; ++i) {
                    if (zi[i] == 0) continue;
                    for (int j = 0; j < cy; ++j) {
                        const int32_t n = tz[i * cy + j];
                        if (n) g += term64(n, zi[i], zj[j], nkz);
                    }
                }
            }
        } else {
            for (int32_t cell = tid; cell < cells; cell += kWG) {
                const int32_t n = tab[cell];
                if (n == 0) continue;
                const int32_t z = cell / cxy, rem = cell - z * cxy;
                const int32_t i = rem / cy, j = rem - i * cy;
                g += term64(n, ni[z * cx + i], nj[z * cy + j], nk[z]);
            }
        }
        g = wave_sum_d(g);
        if (lane == 0) sh.red_d[wave] = g;
        __syncthreads();
        if (tid == 0) {
            double G = 0.0;
            for (int w = 0; w < kWG / 64; ++w) G += sh.red_d[w];
            DevResult &R = a.res[t];
            R.g2_64 = G;
            R.exact = 1;
            R.need_p = (sh.df != 0) ? 1 : 0;  // k_pvalue decides from the FP64 p-value
        }
    }
    if (tid == 0) mark_edge(a, T, sh.indep != 0, sh.need64 && sh.df != 0);
    __syncthreads();
}

// Partial statistic of the z slices [z0, z1) of a table (slice z at tab + z c(X) c(Y), one
// per thread): marginals into ni + z c(X) and nj + z c(Y), df contribution, compensated
// FP32 terms and FP64 terms (the reference's order within a slice, IndependenceTest.cpp:
// 112-137), reduced over the workgroup; the result is thread 0's.  clear: the slices'
// cells and marginals are zeroed once read.  Every thread of the workgroup calls it.
__device__ WidePart range_partial(int32_t *tab, int32_t z0, int32_t z1, int cx, int cy, int32_t *ni, int32_t *nj,
                                  bool clear, Shared &sh)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int cxy = cx * cy;
    int df = 0;
    float s = 0.f, c = 0.f, as = 0.f;
    double g = 0.0;
    for (int32_t z = z0 + tid; z < z1; z += kWG) {
        int32_t *tz = tab + (int64_t)z * cxy;
        int32_t *zi = ni + (int64_t)z * cx, *zj = nj + (int64_t)z * cy;
        zslice32(tz, cx, cy, zi, zj, df, s, c, as);
        int32_t nkz = 0;
        for (int i = 0; i < cx; ++i) nkz += zi[i];
        if (nkz != 0) {
            for (int i = 0; i < cx; ++i) {
                if (zi[i] == 0) continue;
                for (int j = 0; j < cy; ++j) {
                    const int32_t n = tz[i * cy + j];
                    if (n) g += term64(n, zi[i], zj[j], nkz);
                }
            }
        }
        if (clear) {
            for (int k = 0; k < cxy; ++k) tz[k] = 0;
            for (int i = 0; i < cx; ++i) zi[i] = 0;
            for (int j = 0; j < cy; ++j) zj[j] = 0;
        }
    }
    df = wave_sum_i(df);
    wave_nsum(s, c);
    as = wave_sum_f(as);
    g = wave_sum_d(g);
    if (lane == 0) {
        sh.red_i[wave] = df;
        sh.red_f[3 * wave] = s;
        sh.red_f[3 * wave + 1] = c;
        sh.red_f[3 * wave + 2] = as;
        sh.red_d[wave] = g;
    }
    __syncthreads();
    WidePart P{0.0, 0.f, 0.f, 0.f, 0};
    if (tid == 0)
        for (int w = 0; w < kWG / 64; ++w) {
            P.df += sh.red_i[w];
            nadd(P.s, P.c, sh.red_f[3 * w]);
            P.c += sh.red_f[3 * w + 1];
            P.as += sh.red_f[3 * w + 2];
            P.g64 += sh.red_d[w];
        }
    __syncthreads();  // the reduction scratch is free again
    return P;
}

// Slice `slice` of a partitioned test (keys [lo, lo + len)) from its rows' stored keys.
// Its (offset, count) segment in each row chunk is staged, kPartStage chunks at a time, in
// seg_lds (3 kPartStage + 1 words: offset, count, and the prefix of the segments' 8-key
// groups); the workgroup's threads take the groups of all staged segments in turn, each
// group one 16-byte load of 8 keys (aligned: a segment's groups start at its offset
// rounded down to 8 keys; keys outside the segment are masked), four loads in flight per
// thread, and count them into tab (LDS).  Then the slice's strata give their partial
// statistic, and the last slice of the test combines the partials in slice order and
// decides as finalize does (FP64 G^2 at hand).
__device__ void part_slice_stat(const LaunchArgs &a, const DevInst &I, const DevTest &T, int t, int slice, int32_t lo,
                                int32_t len, int32_t *tab, int32_t *seg_lds, Shared &sh)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < len; i += kWG) tab[i] = 0;
    int32_t *base = a.part + T.ws_off;
    const PartLayout L = part_layout(T.dimz, T.cx, T.cy, T.n_slices, I.n_rows);
    const int64_t chunks = (I.n_rows + kPartRows - 1) / kPartRows;
    const uint32_t *seg = reinterpret_cast<const uint32_t *>(base + L.seg);
    const uint16_t *keys = reinterpret_cast<const uint16_t *>(base + L.keys);
    int32_t *s_off = seg_lds, *s_cnt = seg_lds + kPartStage, *s_pre = seg_lds + 2 * kPartStage;
    for (int64_t cb = 0; cb < chunks; cb += kPartStage) {
        const int nc = (int)((chunks - cb < kPartStage) ? chunks - cb : kPartStage);
        uint32_t ng = 0;
        if (tid < nc) {
            const int64_t e = ((cb + tid) * T.n_slices + slice) * 2;
            const int32_t off = (int32_t)seg[e], cnt = (int32_t)seg[e + 1];
            s_off[tid] = off;
            s_cnt[tid] = cnt;
            ng = cnt ? (uint32_t)((off + cnt - (off & ~7) + 7) >> 3) : 0u;
        }
        // exclusive prefix of the groups over the staged chunks (kPartStage <= 2 waves)
        uint32_t inc = ng;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t v = __shfl_up(inc, o, 64);
            if (lane >= o) inc += v;
        }
        if (lane == 63) sh.red_i[wave] = (int32_t)inc;
        __syncthreads();  // (the first pass: tab cleared as well)
        if (tid < nc) s_pre[tid] = (int32_t)(inc - ng) + (wave ? sh.red_i[0] : 0);
        if (tid == 0) s_pre[nc] = sh.red_i[0] + (nc > 64 ? sh.red_i[1] : 0);
        __syncthreads();
        const int32_t total = s_pre[nc];
        int j = 0;  // the staged chunk of this thread's current group (groups ascend)
        for (int32_t f0 = tid; f0 < total; f0 += 4 * kWG) {
            uint4 q[4];
            int32_t k0[4], o[4], e[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int32_t f = f0 + u * kWG;
                k0[u] = -1;
                if (f < total) {
                    while (s_pre[j + 1] <= f) ++j;
                    o[u] = s_off[j];
                    e[u] = o[u] + s_cnt[j];
                    k0[u] = (o[u] & ~7) + 8 * (f - s_pre[j]);
                    q[u] = *reinterpret_cast<const uint4 *>(keys + (cb + j) * kPartRows + k0[u]);
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (k0[u] < 0) continue;
                const uint32_t w[4] = {q[u].x, q[u].y, q[u].z, q[u].w};
#pragma unroll
                for (int h = 0; h < 8; ++h) {
                    const int32_t idx = k0[u] + h;
                    if (idx >= o[u] && idx < e[u]) atomicAdd(&tab[(w[h >> 1] >> (16 * (h & 1))) & 0xffffu], 1);
                }
            }
        }
        __syncthreads();
    }
    const int cxy = T.cx * T.cy;
    const int32_t zbase = lo / cxy;
    int32_t *ni = base + (int64_t)zbase * T.cx;
    int32_t *nj = base + (int64_t)T.dimz * T.cx + (int64_t)zbase * T.cy;
    const WidePart P = range_partial(tab, 0, len / cxy, T.cx, T.cy, ni, nj, false, sh);
    WidePart *parts = reinterpret_cast<WidePart *>(base + L.parts);
    if (tid == 0) parts[slice] = P;
    __threadfence();
    __syncthreads();
    if (tid == 0) {
        const uint32_t prev = atomicAdd(&a.done[t], 1u);
        sh.flag = (prev + 1 == (uint32_t)T.n_slices) ? 1 : 0;
    }
    __syncthreads();
    if (sh.flag && tid == 0) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        int df = 0;
        float S = 0.f, C = 0.f, AS = 0.f;
        double G = 0.0;
        for (int i = 0; i < T.n_slices; ++i) {
            const WidePart Q = parts[i];
            df += Q.df;
            nadd(S, C, Q.s);
            C += Q.c;
            AS += Q.as;
            G += Q.g64;
        }
        DevResult R;
        const int need64 = decide32(a, T.flags, df, S + C, AS, R);
        if (need64) {
            R.g2_64 = G;
            R.exact = 1;
            R.need_p = (df != 0) ? 1 : 0;  // k_pvalue decides from the FP64 p-value
        }
        a.res[t] = R;
        mark_edge(a, T, R.indep, R.need_p);
        a.done[t] = 0;
    }
    __syncthreads();
}

template <int kMode>
struct CritLds {
    static constexpr int value = (kMode == kSinkShared) ? kCritLdsShared : kCritLds;
};

template <int kMode>
struct LdsWords {
    static constexpr int value =
        (kMode == kSinkGlobal) ? 1
        : (kMode == kSinkLaneW) ? 2 * kLaneWideCells * 64
        : (kMode == kSinkLaneW128) ? 2 * kLaneWide128Cells * 64
        : (kMode == kSinkSlice) ? 2 * kLdsCells
                                : kLdsCells;
};

// Waves per SIMD the register allocation must allow: the LDS of the per-lane and shared
// sinks leaves room for 4 workgroups per CU (one wave each per SIMD), so those sinks are
// held to 128 VGPRs; the 64 KiB sinks fit 2 workgroups per CU whatever their registers.
template <int kMode, bool kPacked>
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(
    (kMode == kSinkLane || kMode == kSinkShared || kMode == kSinkBits) ? 4 : 1))) void k_count_stat(LaunchArgs a)
{
    __shared__ int32_t lds[LdsWords<kMode>::value];
    __shared__ int32_t small[kSmallCells];

    __shared__ Shared sh;
    __shared__ double s_crit[CritLds<kMode>::value];  // critical values for df = 1..CritLds
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;

    // XCD-aware dynamic order: workgroups are dispatched round-robin over the XCDs, so the
    // items are cut into one contiguous slab per XCD (neighbouring tests share columns
    // through that XCD's L2) and the workgroups of an XCD take its slab's items from an
    // atomic counter, then help the other slabs.  Item costs differ (table sinks, |Z|,
    // tests skipped because their edge is done), so a static split leaves the launch
    // waiting for its most loaded workgroup.
    //
    // Thread kQ (last wave; the Z staging below uses the first) keeps the queue one item
    // ahead of the workgroup, as a single reservation: at the top of each item it
    // publishes the work-item word it loaded during the previous item and draws the ticket
    // of the next item; it resolves that ticket once the item's rows are counted (the
    // atomic's latency hidden behind the counting) and starts loading its word, which the
    // statistic's time then hides.  Neither the queue atomic nor the work-item load sits
    // between one item and the next.
    constexpr int kQ = kWG - 1;
    constexpr uint64_t kNoItem = ~0ull;
    const int home = blockIdx.x % kXcds;
    const int64_t n_items = a.n_items_dev ? (int64_t)*a.n_items_dev : a.n_items;
    const int64_t slab = (n_items + kXcds - 1) / kXcds;
    __shared__ uint64_t s_pk[2];
    int tried = 0;         // slabs exhausted so far (kQ's state)
    uint32_t ticket = 0;   // kQ's ticket on slab (home + tried) % kXcds
    uint64_t pk_next = kNoItem;
    auto draw = [&]() {
        if (tried < kXcds) ticket = atomicAdd(&a.queue[(home + tried) % kXcds], 1u);
    };
    // the work-item word of `ticket`, drawing on the next slabs while the slab is exhausted
    // (every draw consumes a ticket or exhausts a slab, so the queue ends for every workgroup)
    auto resolve = [&]() -> uint64_t {
        while (tried < kXcds) {
            const int64_t lo = (int64_t)((home + tried) % kXcds) * slab;
            const int64_t hi = (lo + slab < n_items) ? lo + slab : n_items;
            if (lo + (int64_t)ticket < hi) return a.items[a.item_begin + lo + ticket];
            ++tried;
            draw();
        }
        return kNoItem;
    };
    if (tid == kQ) {
        draw();
        pk_next = resolve();
    }
    for (int i = tid; i < CritLds<kMode>::value; i += kWG) s_crit[i] = a.crit[i];
    for (int it = 0;; ++it) {
        if (tid == kQ) {
            s_pk[it & 1] = pk_next;
            if (pk_next != kNoItem) draw();  // the next item's ticket
        }
        __syncthreads();  // s_pk[it & 1] written; s_pk[(it + 1) & 1] read by everyone
        const uint64_t pk = s_pk[it & 1];
        if (pk == kNoItem) break;
        const int t = (int)(pk >> 32);
        const int64_t local = (int64_t)(uint32_t)pk;  // slice * n_chunks + chunk
        const DevTest T = a.tests[t];
        const DevInst I = a.inst[T.inst];
        // the Z columns are staged beside the edge flag's load, not after it (a skipped
        // test leaves them unused); T.d <= kMaxZLds <= kWG
        if (tid < T.d) {
            const int32_t zc = a.zpool[T.z_off + tid];
            sh.zcol[tid] = zc;
            sh.zcard[tid] = a.card[zc];
            sh.zw[tid] = a.colw[zc];
        }
        if (tid == 0) {
            sh.wx = a.colw[T.x];
            sh.wy = a.colw[T.y];
        }
        const bool skip = skipped(a, T);  // uniform: the flag cannot drop below round meanwhile
        __syncthreads();
        if (skip) {
            if (tid == 0 && local == 0) a.res[t] = skipped_result();
            if (tid == kQ) pk_next = resolve();  // published at the top of the next item
            continue;
        }
        const int64_t slice = (kMode == kSinkSlice) ? local / T.n_chunks : 0;
        const int64_t chunk = local - slice * T.n_chunks;
        const int64_t r_begin = chunk * a.rows_per_chunk;
        const int64_t r_end = (r_begin + a.rows_per_chunk < I.n_rows) ? r_begin + a.rows_per_chunk : I.n_rows;
        const int64_t cells = (int64_t)T.dimz * T.cx * T.cy;
        int32_t *g = (T.ws_off >= 0) ? a.ws + T.ws_off : nullptr;
        const int32_t *tab = nullptr;  // the chunk's table in LDS, when there is one
        if (kMode == kSinkLane && (uniform(T.flags) & kTFCol)) {
            // column layout: one [cell][ncol] int32 table for the workgroup, lane l counting
            // in column l (l & 31), so the lanes of a wave (of a half-wave) always hit
            // distinct banks however skewed the data (waves meet only through the atomics,
            // never within one instruction); then thread t sums cell t, reading its columns
            // from (t + j) on
            const int c32 = (int)cells;
            const int ncol = (c32 <= kColCells / 2) ? 64 : (c32 <= kColCells) ? 32 : 16;
            for (int i = tid; i < c32 * ncol; i += kWG) lds[i] = 0;
            __syncthreads();
            count_rows_cols<kPacked>(I, T, sh, r_begin, r_end, lds, ncol);
            __syncthreads();
            int32_t sum[2] = {0, 0};
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int c = tid + k * kWG;
                if (c < c32)
#pragma unroll 4
                    for (int j = 0; j < ncol; ++j) sum[k] += lds[c * ncol + ((c + j) & (ncol - 1))];
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < 2; ++k)
                if (tid + k * kWG < c32) lds[tid + k * kWG] = sum[k];
            __syncthreads();
            tab = lds;
        } else if (kMode == kSinkLane || kMode == kSinkLaneW || kMode == kSinkLaneW128 || kMode == kSinkBits) {
            // per-lane 16-bit counters: [pair][cell][lane], even wave low half, odd wave high
            constexpr int kCells = (kMode == kSinkLaneW)      ? kLaneWideCells
                                   : (kMode == kSinkLaneW128) ? kLaneWide128Cells
                                                              : kLaneCells;
            const int c32 = (int)cells;
            uint32_t *pl = reinterpret_cast<uint32_t *>(lds);
            // only the cells of this table (keys stay below c32: values are checked against
            // their cardinalities when the context is created, rows past the end are masked)
            for (int i = tid; i < c32 * 64; i += kWG) {
                pl[i] = 0;
                pl[kCells * 64 + i] = 0;
            }
            __syncthreads();
            if (kMode == kSinkBits) count_rows_bits(a, I, T, sh, r_begin, r_end, pl);
            else count_rows_lane<kCells, kPacked>(I, T, sh, r_begin, r_end, pl);
            __syncthreads();
            // sum the 256 private counters of every cell: 8 threads per cell, each over
            // 8 lanes of both pairs (both halves), then a 3-step shuffle reduction
            for (int idx = tid; idx < c32 * 8; idx += kWG) {
                const int k = idx >> 3, part = idx & 7;
                uint32_t sum = 0;
#pragma unroll
                for (int pr = 0; pr < 2; ++pr)
#pragma unroll
                    for (int l = 0; l < 8; ++l) {
                        const uint32_t v = pl[(pr * kCells + k) * 64 + part * 8 + l];
                        sum += (v & 0xffffu) + (v >> 16);
                    }
                for (int o = 1; o < 8; o <<= 1) sum += __shfl_xor(sum, o, 64);
                int dst = k;
                if (kMode == kSinkBits) {  // counter (r P + p) Q + q -> cell p mA + q mB + r mC
                    const BitsRoles ro = bits_roles(T, sh);
                    const int pq = k % (ro.P * ro.Q), r = k / (ro.P * ro.Q);
                    dst = (pq / ro.Q) * ro.mA + (pq % ro.Q) * ro.mB + r * ro.mC;
                }
                if (part == 0) small[dst] = (int32_t)sum;
            }
            __syncthreads();
            tab = small;
        } else if (kMode == kSinkShared) {
            // Tables of at most 2 * kColCells cells: column layout, one table of [cell][64]
            // words (up to 128 cells), [cell][32] (up to 256) or [cell][16] for the
            // workgroup, lane l counting in column l (l & 31, l & 15), key << 6 (5, 4) from
            // the lane's column, so the lanes of a wave (of a half-wave, of a quarter) always
            // hit distinct banks, however skewed the data (waves meet only through the
            // atomics, never within one instruction); then thread t sums cell t (and
            // t + 256), reading its columns from (t + j) on.  Larger
            // tables: k histogram copies shared by the waves (8 at most, 4 for items below
            // kColMinRows rows, whose clearing and reduction the counting does not repay;
            // fewer when they do not fit), lane l counting into copy l % k, so skewed data
            // collides on one address in at most 64 / k lanes of an instruction; copies are
            // offset by 4 banks from one another, so lanes of different copies hitting the
            // same cell use different banks.  One row loop serves both (the key shift).
            const int c32 = (int)cells;
            // (the column layout clears and reduces twice the words of the copies: it is
            // used for items of enough rows to repay that)
            const bool col = c32 <= 2 * kColCells && r_end - r_begin >= kColMinRows;
            // columns: every lane its own when they fit
            const int ncol = (c32 <= kColCells / 2) ? 64 : (c32 <= kColCells) ? 32 : 16;
            int copies = 1, my = 0, stride = c32, span = c32 * ncol;
            if (!col) {
                stride = ((c32 + 63) & ~63) + 4;
                const int kmax = (r_end - r_begin < kColMinRows) ? 4 : 8;
                while (copies < kmax && 2 * copies * stride <= kLdsCells) copies *= 2;
                my = lane & (copies - 1);
                span = (copies > 1) ? copies * stride : c32;
            }
            for (int i = tid; i < span; i += kWG) lds[i] = 0;
            __syncthreads();
            {
                // one row loop for every layout (a second inlined copy costs the kernel its
                // 128-VGPR budget)
                const uint32_t ksh = col ? (ncol == 64 ? 6u : ncol == 32 ? 5u : 4u) : 0u;
                int32_t *hp = col ? lds + (lane & (ncol - 1)) : lds + (copies > 1 ? my * stride : 0);
                count_rows<kSinkShared, kPacked>(I, T, sh, r_begin, r_end, hp, ksh);
            }
            __syncthreads();
            if (col) {
                int32_t sum[2] = {0, 0};
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const int c = tid + k * kWG;
                    if (c < c32)
#pragma unroll 4
                        for (int j = 0; j < ncol; ++j) sum[k] += lds[c * ncol + ((c + j) & (ncol - 1))];
                }
                __syncthreads();
#pragma unroll
                for (int k = 0; k < 2; ++k)
                    if (tid + k * kWG < c32) lds[tid + k * kWG] = sum[k];
                __syncthreads();
            } else if (copies > 1) {
                for (int i = tid; i < c32; i += kWG) {
                    int32_t s = lds[i];
                    for (int k = 1; k < copies; ++k) s += lds[k * stride + i];
                    lds[i] = s;
                }
                __syncthreads();
            }
            tab = lds;
        } else if (kMode == kSinkSlice) {
            const int32_t lo = (int32_t)slice * T.slice_cells;
            const int32_t len = (int32_t)((cells - lo < T.slice_cells) ? cells - lo : T.slice_cells);
            if (T.flags & kTFPart) {  // the slice's rows, partitioned beforehand
                if (tid == kQ) pk_next = resolve();
                part_slice_stat(a, I, T, t, (int)slice, lo, len, lds, small, sh);
                continue;
            }
            if (tid == 0) {
                sh.slice_lo = lo;
                sh.slice_len = len;
            }
            for (int i = tid; i < len; i += kWG) lds[i] = 0;
            __syncthreads();
            count_rows<kSinkSlice, kPacked>(I, T, sh, r_begin, r_end, lds);
            __syncthreads();
            if (T.n_slices == 1 && T.n_chunks == 1 && !(T.flags & kTFKeep)) {
                // the whole table counted in LDS by this item: the statistic straight from
                // it, marginals in the workspace, which then reads zero again
                if (tid == kQ) pk_next = resolve();
                const int64_t marg = (int64_t)T.dimz * (T.cx + T.cy + 1);
                finalize(a, T, t, lds, g + cells, sh, s_crit, CritLds<kMode>::value);
                for (int64_t i = tid; i < marg; i += kWG) g[cells + i] = 0;
                continue;
            }
            for (int i = tid; i < len; i += kWG) {
                const int32_t v = lds[i];
                if (v) atomicAdd(&g[lo + i], v);
            }
        } else {
            count_rows<kSinkGlobal, kPacked>(I, T, sh, r_begin, r_end, g);
        }
        if (tid == kQ) pk_next = resolve();  // published at the top of the next item
        if (tab) {
            if (T.flags & kTFLdsMarg) {
                finalize(a, T, t, tab, const_cast<int32_t *>(tab) + cells, sh, s_crit, CritLds<kMode>::value);
                continue;
            }
            for (int i = tid; i < (int)cells; i += kWG) {
                const int32_t v = tab[i];
                if (v) atomicAdd(&g[i], v);
            }
        }
        if (T.flags & kTFWide) {  // the grid-wide finalize kernels take over
            __syncthreads();       // the next item rewrites the staged Z columns
            continue;
        }
        // retire the chunk; the workgroup retiring the last chunk owns the statistic
        __threadfence();
        __syncthreads();
        if (tid == 0) {
            const uint32_t prev = atomicAdd(&a.done[t], 1u);
            sh.flag = (prev + 1 == (uint32_t)(T.n_chunks * T.n_slices)) ? 1 : 0;
        }
        __syncthreads();
        if (sh.flag) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            finalize(a, T, t, g, g + cells, sh, s_crit, CritLds<kMode>::value);
            if (!(T.flags & kTFKeep)) {
                // the workspace must read zero wherever the next launch places a table,
                // so the marginal scratch behind the table is cleared as well
                const int64_t used = cells + (int64_t)T.dimz * (T.cx + T.cy + 1);
                for (int64_t i = tid; i < used; i += kWG) g[i] = 0;
            }
            if (tid == 0) a.done[t] = 0;
        }
        __syncthreads();
    }
}

// Round compaction (one 256-item tile of one sink group per workgroup): a wave ballot of
// the live items, their ranks within the tile, one atomic reservation per tile.
__global__ __launch_bounds__(kWG) void k_compact(CompactArgs c)
{
    __shared__ uint32_t s_wave[kWG / 64];
    __shared__ uint32_t s_base;
    const int g = blockIdx.y, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t i = (int64_t)blockIdx.x * kWG + tid;
    bool live = false;
    uint64_t pk = 0;
    if (i < c.len[g]) {
        pk = c.in[c.begin[g] + i];
        const int t = (int)(pk >> 32);
        const int32_t e = c.tests[t].edge;
        live = !(e >= 0 && c.edge_done[e] < c.round);
        if (!live && (uint32_t)pk == 0u) c.res[t] = skipped_result();
    }
    const uint64_t mask = __ballot(live);
    const uint32_t rank = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    if (lane == 0) s_wave[wave] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (tid == 0) {
        uint32_t total = 0;
        for (int w = 0; w < kWG / 64; ++w) {
            const uint32_t v = s_wave[w];
            s_wave[w] = total;
            total += v;
        }
        s_base = total ? atomicAdd(&c.count[g], total) : 0u;
    }
    __syncthreads();
    if (live) c.out[c.begin[g] + s_base + s_wave[wave] + rank] = pk;
}

// Grid-wide finalize, part 1: the z slices [z0, z1) of one test's device-memory table,
// one slice per thread.  Marginals go to the test's scratch behind the table; FP32 and
// FP64 terms (same order within a slice as IndependenceTest.cpp:112-137) and df are
// reduced to one partial per workgroup; the slices read are cleared, so the workspace
// reads zero again for the next launch.
__global__ __launch_bounds__(kWG) void k_fin_wide_part(LaunchArgs a, const WideItem *items, WidePart *parts)
{
    __shared__ Shared sh;
    const WideItem it = items[blockIdx.x];
    const DevTest T = a.tests[it.t];
    if (skipped(a, T)) return;  // the table was never counted: it is still zero
    const int cx = T.cx, cy = T.cy;
    int32_t *tab = a.ws + T.ws_off;
    int32_t *ni = tab + (int64_t)T.dimz * cx * cy;
    int32_t *nj = ni + (int64_t)T.dimz * cx;
    const WidePart P = range_partial(tab, it.z0, it.z1, cx, cy, ni, nj, true, sh);
    if (threadIdx.x == 0) parts[blockIdx.x] = P;
}

// Grid-wide finalize, part 2: one thread per test sums its partials in order and decides
// like the counting kernels' finalize; the FP64 G^2 is always available here.
__global__ __launch_bounds__(kWG) void k_fin_wide_reduce(LaunchArgs a, const WidePart *parts, const WideTest *wt,
                                                         int64_t n_wide)
{
    const int64_t k = (int64_t)blockIdx.x * kWG + threadIdx.x;
    if (k >= n_wide) return;
    const WideTest W = wt[k];
    const DevTest T = a.tests[W.t];
    if (skipped(a, T)) {
        a.res[W.t] = skipped_result();
        return;
    }
    int df = 0;
    float S = 0.f, C = 0.f, AS = 0.f;
    double G = 0.0;
    for (int i = 0; i < W.nparts; ++i) {
        const WidePart P = parts[W.part0 + i];
        df += P.df;
        nadd(S, C, P.s);
        C += P.c;
        AS += P.as;
        G += P.g64;
    }
    DevResult R;
    const int need64 = decide32(a, T.flags, df, S + C, AS, R);
    if (need64) {
        R.g2_64 = G;
        R.exact = 1;
        R.need_p = (df != 0) ? 1 : 0;  // k_pvalue decides from the FP64 p-value
    }
    a.res[W.t] = R;
    mark_edge(a, T, R.indep, R.need_p);
}

__global__ __launch_bounds__(kWG) void k_check_values(const uint8_t *cols, int64_t stride, int64_t n_rows,
                                                     int n_vars, const int32_t *card, uint32_t *bad)
{
    const int v = blockIdx.y;
    if (v >= n_vars) return;
    const uint32_t c = (uint32_t)card[v];
    const uint8_t *col = cols + (int64_t)v * stride;
    uint32_t flag = 0;
    for (int64_t r = ((int64_t)blockIdx.x * kWG + threadIdx.x) * 16; r < n_rows;
         r += (int64_t)gridDim.x * kWG * 16) {
        const uint4 q = ld16(col + r);
        const int64_t nv = n_rows - r;
#pragma unroll
        for (int b = 0; b < 16; ++b)
            if (b < nv && byte_of(q, b) >= c) flag = 1;
    }
    if (flag) atomicOr(bad, 1u);
}

// Packed copy of the columns (layout at ldw): one thread per 16-row group, every group
// of the stride (rows past n_rows pack whatever the padding holds; the counting kernels
// never count them).  Each value is masked to w bits, so that no value can spill into
// a neighbour's bits.
__global__ __launch_bounds__(kWG) void k_pack_cols(const uint8_t *cols, uint8_t *pcols, int64_t stride, int n_vars,
                                                   const uint8_t *colw)
{
    const int v = blockIdx.y;
    if (v >= n_vars) return;
    const uint32_t w = colw[v];
    const uint8_t *col = cols + (int64_t)v * stride;
    uint8_t *out = pcols + (int64_t)v * stride;
    for (int64_t r = ((int64_t)blockIdx.x * kWG + threadIdx.x) * 16; r < stride; r += (int64_t)gridDim.x * kWG * 16) {
        const uint4 q = ld16(col + r);
        if (w == 8) {
            *reinterpret_cast<uint4 *>(out + r) = q;
        } else if (w == 4) {
            const uint32_t m = 0x0F0F0F0Fu;
            *reinterpret_cast<uint2 *>(out + (r >> 1)) =
                make_uint2((q.x & m) | ((q.y & m) << 4), (q.z & m) | ((q.w & m) << 4));
        } else {
            const uint32_t m = 0x03030303u;
            *reinterpret_cast<uint32_t *>(out + (r >> 2)) =
                (q.x & m) | ((q.y & m) << 2) | ((q.z & m) << 4) | ((q.w & m) << 6);
        }
    }
}

// FP64 p-value of the tests the statistic kernel rechecked (one thread per test):
// p = 1 - pchisq(G^2, df), independent iff p > alpha (IndependenceTest.cpp:146-154).
__global__ __launch_bounds__(kWG) void k_pvalue(DevResult *res, int64_t n, double alpha, const DevTest *tests,
                                                 int32_t *edge_done, int32_t round, uint32_t *clear, int n_clear)
{
    const int64_t t = (int64_t)blockIdx.x * kWG + threadIdx.x;
    if (clear && blockIdx.x == 0 && threadIdx.x < n_clear) clear[threadIdx.x] = 0u;
    if (t >= n) return;
    DevResult &R = res[t];
    if (!R.need_p) return;
    const double p = 1.0 - pchisq(R.g2_64, R.df);
    R.p64 = p;
    R.indep = (p > alpha) ? 1 : 0;
    R.need_p = 0;
    if (edge_done && tests && R.indep) {
        const int32_t e = tests[t].edge;
        if (e >= 0) atomicMin(&edge_done[e], round);
    }
}

// ---------------------------------------------------------------- grouped level 1 (MFMA)

// Statistic of one test of a group chunk (one thread per test) from the chunk's table
// tab (row stride kGrpCols): cell (c, i, j) of n[z][x][y] is tab[(c * cu + i) * kGrpCols
// + col0 + j] when u is x, tab[(c * cu + j) * kGrpCols + col0 + i] when u is y.  Same
// marginals, adjusted df, FP32 terms, decision and FP64 recheck (terms in the reference's
// z, x, y order, IndependenceTest.cpp:112-137) as finalize.  nj: the thread's marginal
// scratch, entry j at nj[j * kGrpMaxTests].
__device__ void group_stat(const LaunchArgs &a, int t, const DevTest &T, const int32_t *tab, int col0, int cu,
                           int cz, bool u_is_x, int32_t *nj)
{
    const int cx = T.cx, cy = T.cy;
    auto cell = [&](int c, int i, int j) -> int32_t {
        const int ur = u_is_x ? i : j, oc = u_is_x ? j : i;
        return tab[(c * cu + ur) * kGrpCols + col0 + oc];
    };
    // n_{+yz} of slice c into nj; returns n_{++z}, with the count of non-empty columns
    auto col_sums = [&](int c, int &aly) -> int32_t {
        int32_t nk = 0;
        aly = 0;
        for (int j = 0; j < cy; ++j) {
            int32_t v = 0;
            for (int i = 0; i < cx; ++i) v += cell(c, i, j);
            nj[j * kGrpMaxTests] = v;
            nk += v;
            aly += (v > 0);
        }
        return nk;
    };
    int df = 0;
    float s = 0.f, cc = 0.f, as = 0.f;
    for (int c = 0; c < cz; ++c) {
        int aly;
        const int32_t nk = col_sums(c, aly);
        int alx = 0;
        for (int i = 0; i < cx; ++i) {
            int32_t ni = 0;
            for (int j = 0; j < cy; ++j) ni += cell(c, i, j);
            alx += (ni > 0);
            if (ni == 0) continue;
            for (int j = 0; j < cy; ++j) {
                const int32_t n = cell(c, i, j);
                if (n == 0) continue;
                const float tm = term32(n, ni, nj[j * kGrpMaxTests], nk);
                nadd(s, cc, tm);
                as += fabsf(tm);
            }
        }
        alx = alx >= 1 ? alx : 1;
        aly = aly >= 1 ? aly : 1;
        df += (alx - 1) * (aly - 1);
    }
    DevResult R;
    if (decide32(a, T.flags, df, s + cc, as, R)) {
        double G = 0.0;
        for (int c = 0; c < cz; ++c) {
            int aly;
            const int32_t nk = col_sums(c, aly);
            if (nk == 0) continue;
            for (int i = 0; i < cx; ++i) {
                int32_t ni = 0;
                for (int j = 0; j < cy; ++j) ni += cell(c, i, j);
                if (ni == 0) continue;
                for (int j = 0; j < cy; ++j) {
                    const int32_t n = cell(c, i, j);
                    if (n) G += term64(n, ni, nj[j * kGrpMaxTests], nk);
                }
            }
        }
        R.g2_64 = G;
        R.exact = 1;
        R.need_p = (df != 0) ? 1 : 0;  // k_pvalue decides from the FP64 p-value
    }
    a.res[t] = R;
    mark_edge(a, T, R.indep != 0, R.need_p != 0);
}

// One work item = (chunk, row split) of grouped level-1 tests.  The four waves take the
// split's 32-row steps in turn; each forms the one-hot A rows of (u, z) -- the byte-wise
// AND of u's and z's one-hot bytes -- and the B columns of the chunk's others from
// 16-byte loads of the uint8 columns, and accumulates kNA x kGrpB fragments of 32 x 32
// with v_mfma_i32_32x32x32_i8 (exact int8 products, int32 sums).  The waves' fragments
// are summed in LDS; a split chunk adds its partial table to the workspace and the
// workgroup retiring the last split computes the statistics, one thread per test.
template <bool kPerm, int kNA>
__global__ __launch_bounds__(kWG) void k_group_count(GroupArgs g)
{
    __shared__ int32_t s_tab[kGrpRows * kGrpCols];     // 32 KiB
    __shared__ int32_t s_nj[kGrpRows * kGrpMaxTests];  // 16 KiB of marginal scratch
    __shared__ int32_t s_colvar[kGrpCols];
    __shared__ uint8_t s_colstate[kGrpCols];
    __shared__ int32_t s_col0[kGrpMaxTests + 1];
    __shared__ int32_t s_flag;
    const LaunchArgs &a = g.a;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, h = lane >> 5, l32 = lane & 31;
    // XCD-aware order: consecutive items (a chunk's splits, a group's chunks) run on one
    // XCD, whose L2 then serves their shared u and z columns
    const int64_t per = (g.n_items + kXcds - 1) / kXcds;
    const int64_t item = (int64_t)(blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    if (item >= g.n_items) return;
    const uint64_t w = g.items[item];
    const GroupChunk C = g.chunks[(int64_t)(w >> 32)];
    const int split = (int)(uint32_t)w;
    const DevInst I = a.inst[C.inst];
    if (tid == 0) {
        // a chunk whose tests were all decided in earlier rounds is not counted
        int live = 0, off = 0;
        for (int j = 0; j < C.n_t; ++j) {
            const DevTest T = a.tests[g.gtest[C.t0 + j]];
            live |= skipped(a, T) ? 0 : 1;
            s_col0[j] = off;
            off += (T.x == C.u) ? T.cy : T.cx;
        }
        s_col0[C.n_t] = off;
        s_flag = live;
    }
    for (int k = tid; k < kGrpCols; k += kWG) {
        s_colvar[k] = C.u;  // padding columns: a valid column with a state no byte holds
        s_colstate[k] = 255;
    }
    for (int i = tid; i < kGrpRows * kGrpCols; i += kWG) s_tab[i] = 0;
    __syncthreads();
    if (!s_flag) {
        if (split == 0 && tid < C.n_t) a.res[g.gtest[C.t0 + tid]] = skipped_result();
        return;
    }
    if (tid < C.n_t) {
        const DevTest T = a.tests[g.gtest[C.t0 + tid]];
        const int o = (T.x == C.u) ? T.y : T.x;
        for (int k = s_col0[tid], b = 0; k < s_col0[tid + 1]; ++k, ++b) {
            s_colvar[k] = o;
            s_colstate[k] = (uint8_t)b;
        }
    }
    __syncthreads();
    // this lane's A rows 32 f + l32 = (a, c) with row = c * cu + a, and B columns 32 e + l32
    uint32_t lau[kNA], hau[kNA], laz[kNA], haz[kNA];
#pragma unroll
    for (int f = 0; f < kNA; ++f) {
        const int r = 32 * f + l32;
        const bool real = r < C.cu * C.cz;
        onehot_lut<kPerm>(real ? (uint8_t)(r % C.cu) : (uint8_t)255, lau[f], hau[f]);
        onehot_lut<kPerm>(real ? (uint8_t)(r / C.cu) : (uint8_t)255, laz[f], haz[f]);
    }
    const uint8_t *pu = I.cols + (int64_t)C.u * I.stride + h * 16;
    const uint8_t *pz = I.cols + (int64_t)C.z * I.stride + h * 16;
    const uint8_t *pb[kGrpB];
    uint32_t lb[kGrpB], hb[kGrpB];
#pragma unroll
    for (int e = 0; e < kGrpB; ++e) {
        const int k = 32 * e + l32;
        pb[e] = I.cols + (int64_t)s_colvar[k] * I.stride + h * 16;
        onehot_lut<kPerm>(s_colstate[k], lb[e], hb[e]);
    }
    v16i acc[kNA][kGrpB];
#pragma unroll
    for (int f = 0; f < kNA; ++f)
#pragma unroll
        for (int e = 0; e < kGrpB; ++e) acc[f][e] = v16i{0};
    auto mma = [&](const uint4 &qu, const uint4 &qz, const uint4 *qb, const uint4 *mask) {
        v4i bo[kGrpB];
#pragma unroll
        for (int e = 0; e < kGrpB; ++e) bo[e] = onehot<kPerm>(qb[e], lb[e], hb[e]);
#pragma unroll
        for (int f = 0; f < kNA; ++f) {
            const v4i ou = onehot<kPerm>(qu, lau[f], hau[f]), oz = onehot<kPerm>(qz, laz[f], haz[f]);
            v4i ao = ou & oz;
            if (mask) ao = and4(ao, *mask);
#pragma unroll
            for (int e = 0; e < kGrpB; ++e) acc[f][e] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ao, bo[e], acc[f][e], 0, 0, 0);
        }
    };
    const int64_t r0 = (int64_t)split * g.rows_per_split;
    const int64_t r1 = (r0 + g.rows_per_split < I.n_rows) ? r0 + g.rows_per_split : I.n_rows;
    const int64_t full_end = r0 + (r1 - r0) / 32 * 32;
    for (int64_t k0 = r0 + 32 * wave; k0 < full_end; k0 += 128) {
        uint4 qb[kGrpB];
        const uint4 qu = ld16(pu + k0), qz = ld16(pz + k0);
#pragma unroll
        for (int e = 0; e < kGrpB; ++e) qb[e] = ld16(pb[e] + k0);
        mma(qu, qz, qb, nullptr);
    }
    if (full_end < r1 && wave == 0) {
        // last partial step: bytes past r1 are masked out (their values are arbitrary)
        const int64_t valid = r1 - (full_end + h * 16);
        uint4 m;
        m.x = byte_mask_word(valid, 0);
        m.y = byte_mask_word(valid, 1);
        m.z = byte_mask_word(valid, 2);
        m.w = byte_mask_word(valid, 3);
        uint4 qb[kGrpB];
        const uint4 qu = ld16(pu + full_end), qz = ld16(pz + full_end);
#pragma unroll
        for (int e = 0; e < kGrpB; ++e) qb[e] = ld16(pb[e] + full_end);
        mma(qu, qz, qb, &m);
    }
    // the waves' fragments summed in LDS
#pragma unroll
    for (int f = 0; f < kNA; ++f)
#pragma unroll
        for (int e = 0; e < kGrpB; ++e)
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int row = 32 * f + (q >> 2) * 8 + h * 4 + (q & 3);
                const int32_t v = acc[f][e][q];
                if (v) atomicAdd(&s_tab[row * kGrpCols + 32 * e + l32], v);
            }
    __syncthreads();
    const int32_t *tab = s_tab;
    int32_t *gw = (C.n_splits > 1) ? a.ws + C.ws_off : nullptr;
    if (gw) {
        for (int i = tid; i < kGrpRows * kGrpCols; i += kWG) {
            const int32_t v = s_tab[i];
            if (v) atomicAdd(&gw[i], v);
        }
        __threadfence();
        __syncthreads();
        if (tid == 0) {
            const uint32_t prev = atomicAdd(&a.done[C.done], 1u);
            s_flag = (prev + 1 == (uint32_t)C.n_splits) ? 1 : 0;
        }
        __syncthreads();
        if (!s_flag) return;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        tab = gw;
    }
    if (tid < C.n_t) {
        const int t = g.gtest[C.t0 + tid];
        const DevTest T = a.tests[t];
        if (skipped(a, T)) a.res[t] = skipped_result();
        else group_stat(a, t, T, tab, s_col0[tid], C.cu, C.cz, T.x == C.u, s_nj + tid);
    }
    if (gw) {
        // the workspace reads zero again for the next launch
        __syncthreads();
        for (int i = tid; i < kGrpRows * kGrpCols; i += kWG) gw[i] = 0;
        if (tid == 0) a.done[C.done] = 0;
    }
}

__global__ __launch_bounds__(kWG) void k_zero(uint32_t *p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x; i < n; i += (int64_t)gridDim.x * kWG) p[i] = 0u;
}

// Bit planes of one instance: one thread per (variable, 32-row word) reads the word's 32
// values and writes one word per plane of the variable (plane i: bit i of every value);
// words past the rows are written clear.  32 w < n_rows implies 32 w + 32 <= stride
// (strides are multiples of 256).
__global__ __launch_bounds__(kWG) void k_build_planes(const uint8_t *cols, int64_t stride, int64_t n_rows,
                                                      const int32_t *card, const int32_t *poff, uint32_t *planes,
                                                      int64_t pwords)
{
    const int v = blockIdx.y;
    const int32_t base = poff[v];
    if (base < 0) return;
    const int nb = nbits_of(card[v]);
    for (int64_t w = (int64_t)blockIdx.x * kWG + threadIdx.x; w < pwords; w += (int64_t)gridDim.x * kWG) {
        const int64_t r0 = w * 32;
        uint32_t val[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
        uint32_t valid = 0u;
        if (r0 < n_rows) {
            const uint4 q0 = ld16(cols + (int64_t)v * stride + r0);
            const uint4 q1 = ld16(cols + (int64_t)v * stride + r0 + 16);
            val[0] = q0.x; val[1] = q0.y; val[2] = q0.z; val[3] = q0.w;
            val[4] = q1.x; val[5] = q1.y; val[6] = q1.z; val[7] = q1.w;
            const int64_t nv = n_rows - r0;
            valid = nv >= 32 ? 0xffffffffu : ((1u << nv) - 1u);
        }
        for (int i = 0; i < nb; ++i) {
            uint32_t m = 0u;
#pragma unroll
            for (int j = 0; j < 32; ++j) m |= ((val[j >> 2] >> ((j & 3) * 8 + i)) & 1u) << j;
            planes[(int64_t)(base + i) * pwords + w] = m & valid;
        }
    }
}

// The partitioning launch: a fixed grid, each workgroup taking items blockIdx.x,
// blockIdx.x + gridDim.x, ... (items are alike: kPartItemChunks row chunks of one test
// each, the test's Z columns staged once for them).
template <bool kPacked>
__global__ __launch_bounds__(kWG) void k_part_keys(PartArgs p)
{
    __shared__ uint32_t s_cnt[kPartMaxSlices];
    __shared__ __align__(16) uint16_t s_stage[kPartRows];
    __shared__ Shared sh;
    const LaunchArgs &a = p.a;
    const int tid = threadIdx.x;
    for (int64_t it = blockIdx.x; it < p.n_items; it += gridDim.x) {
        const uint64_t pk = p.items[it];
        const int t = (int)(pk >> 32);
        const int64_t c0 = (int64_t)(uint32_t)pk;
        const DevTest T = a.tests[t];
        const DevInst I = a.inst[T.inst];
        const int64_t chunks = (I.n_rows + kPartRows - 1) / kPartRows;
        const int64_t c1 = (c0 + kPartItemChunks < chunks) ? c0 + kPartItemChunks : chunks;
        if (tid < T.d) {
            const int32_t zc = a.zpool[T.z_off + tid];
            sh.zcol[tid] = zc;
            sh.zcard[tid] = a.card[zc];
            sh.zw[tid] = a.colw[zc];
        }
        if (tid == 0) {
            sh.wx = a.colw[T.x];
            sh.wy = a.colw[T.y];
        }
        const bool skip = skipped(a, T);  // uniform: the flag cannot drop below round meanwhile
        for (int64_t chunk = c0; chunk < c1 && !skip; ++chunk) {
            for (int i = tid; i < T.n_slices; i += kWG) s_cnt[i] = 0u;
            __syncthreads();  // (the first chunk: the Z columns staged as well)
            part_item<kPacked>(a, I, T, sh, chunk, s_cnt, s_stage);
            __syncthreads();  // the next chunk clears the counters
        }
        __syncthreads();  // the next item restages the Z columns
    }
}

}  // namespace

hipError_t launch_part_keys(const PartArgs &p, bool packed, int grid, hipStream_t s)
{
    if (p.n_items <= 0) return hipSuccess;
    if (packed) hipLaunchKernelGGL(k_part_keys<true>, dim3((unsigned)grid), dim3(kWG), 0, s, p);
    else hipLaunchKernelGGL(k_part_keys<false>, dim3((unsigned)grid), dim3(kWG), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_group_count(const GroupArgs &g, int n_a, hipStream_t s)
{
    if (g.n_items <= 0) return hipSuccess;
    const int64_t grid = (g.n_items + kXcds - 1) / kXcds * kXcds;
    const bool perm = g.max_card <= 8;
    if (n_a <= 1) {
        if (perm) hipLaunchKernelGGL((k_group_count<true, 1>), dim3((unsigned)grid), dim3(kWG), 0, s, g);
        else hipLaunchKernelGGL((k_group_count<false, 1>), dim3((unsigned)grid), dim3(kWG), 0, s, g);
    } else {
        if (perm) hipLaunchKernelGGL((k_group_count<true, 2>), dim3((unsigned)grid), dim3(kWG), 0, s, g);
        else hipLaunchKernelGGL((k_group_count<false, 2>), dim3((unsigned)grid), dim3(kWG), 0, s, g);
    }
    return hipGetLastError();
}

hipError_t launch_build_planes(const uint8_t *cols, int64_t stride, int64_t n_rows, int n_vars, const int32_t *card,
                               const int32_t *poff, uint32_t *planes, int64_t pwords, hipStream_t s)
{
    if (n_vars <= 0 || pwords <= 0) return hipSuccess;
    const int64_t bx = std::min<int64_t>((pwords + kWG - 1) / kWG, 256);
    hipLaunchKernelGGL(k_build_planes, dim3((unsigned)bx, (unsigned)n_vars), dim3(kWG), 0, s, cols, stride, n_rows, card,
                       poff, planes, pwords);
    return hipGetLastError();
}

hipError_t launch_zero(uint32_t *p, int64_t n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n + kWG - 1) / kWG, 1024);
    hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(kWG), 0, s, p, n);
    return hipGetLastError();
}

template <bool kPacked>
void launch_count_stat_form(const LaunchArgs &a, int mode, int grid, hipStream_t s)
{
    switch (mode) {
    case kSinkLane: hipLaunchKernelGGL((k_count_stat<kSinkLane, kPacked>), dim3(grid), dim3(kWG), 0, s, a); break;
    case kSinkLaneW: hipLaunchKernelGGL((k_count_stat<kSinkLaneW, kPacked>), dim3(grid), dim3(kWG), 0, s, a); break;
    case kSinkLaneW128:
        hipLaunchKernelGGL((k_count_stat<kSinkLaneW128, kPacked>), dim3(grid), dim3(kWG), 0, s, a);
        break;
    case kSinkSlice: hipLaunchKernelGGL((k_count_stat<kSinkSlice, kPacked>), dim3(grid), dim3(kWG), 0, s, a); break;
    case kSinkShared: hipLaunchKernelGGL((k_count_stat<kSinkShared, kPacked>), dim3(grid), dim3(kWG), 0, s, a); break;
    case kSinkBits: hipLaunchKernelGGL((k_count_stat<kSinkBits, kPacked>), dim3(grid), dim3(kWG), 0, s, a); break;
    default: hipLaunchKernelGGL((k_count_stat<kSinkGlobal, kPacked>), dim3(grid), dim3(kWG), 0, s, a); break;
    }
}

hipError_t launch_count_stat(const LaunchArgs &a, int mode, bool packed, int grid, hipStream_t s)
{
    if (a.n_items <= 0) return hipSuccess;
    if (packed) launch_count_stat_form<true>(a, mode, grid, s);
    else launch_count_stat_form<false>(a, mode, grid, s);
    return hipGetLastError();
}

hipError_t launch_compact(const CompactArgs &c, int64_t max_len, hipStream_t s)
{
    if (max_len <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_compact, dim3((unsigned)((max_len + kWG - 1) / kWG), 6), dim3(kWG), 0, s, c);
    return hipGetLastError();
}

hipError_t launch_finalize_wide(const LaunchArgs &a, const WideItem *items, int64_t n_items, WidePart *parts,
                                const WideTest *wt, int64_t n_wide, hipStream_t s)
{
    if (n_items <= 0 || n_wide <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fin_wide_part, dim3((unsigned)n_items), dim3(kWG), 0, s, a, items, parts);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_fin_wide_reduce, dim3((unsigned)((n_wide + kWG - 1) / kWG)), dim3(kWG), 0, s, a, parts, wt,
                       n_wide);
    return hipGetLastError();
}

hipError_t launch_pvalue(DevResult *res, int64_t n, double alpha, const DevTest *tests, int32_t *edge_done,
                         int32_t round, uint32_t *clear, int n_clear, hipStream_t s)
{
    if (n <= 0 && !clear) return hipSuccess;
    const int64_t blocks = std::max<int64_t>(1, (n + kWG - 1) / kWG);
    hipLaunchKernelGGL(k_pvalue, dim3((unsigned)blocks), dim3(kWG), 0, s, res, n, alpha, tests, edge_done, round,
                       clear, n_clear);
    return hipGetLastError();
}

hipError_t launch_check_values(const uint8_t *cols, int64_t stride, int64_t n_rows, int n_vars,
                               const int32_t *card, uint32_t *bad, hipStream_t s)
{
    if (n_vars <= 0 || n_rows <= 0) return hipSuccess;
    int64_t steps = (n_rows + (int64_t)kWG * 16 - 1) / ((int64_t)kWG * 16);
    const int gx = (int)(steps < 64 ? steps : 64);
    hipLaunchKernelGGL(k_check_values, dim3(gx, n_vars), dim3(kWG), 0, s, cols, stride, n_rows, n_vars, card,
                       bad);
    return hipGetLastError();
}

hipError_t launch_pack_cols(const uint8_t *cols, uint8_t *pcols, int64_t stride, int n_vars, const uint8_t *colw,
                            hipStream_t s)
{
    if (n_vars <= 0 || stride <= 0) return hipSuccess;
    if (stride % 16 != 0) return hipErrorInvalidValue;
    const int64_t steps = (stride + (int64_t)kWG * 16 - 1) / ((int64_t)kWG * 16);
    const int gx = (int)(steps < 64 ? steps : 64);
    hipLaunchKernelGGL(k_pack_cols, dim3(gx, n_vars), dim3(kWG), 0, s, cols, pcols, stride, n_vars, colw);
    return hipGetLastError();
}

}  // namespace fpgm
