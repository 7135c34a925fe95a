// One-hot int8 operands of the MFMA counting kernels (gfx950), formed in registers from
// 16-byte loads of uint8 columns (16 rows, one per byte).
//
// v_mfma_i32_32x32x32_i8 operand layout (checked on the device): lane l supplies row
// (or column) l % 32 and k = (l / 32) * 16 + byte; accumulator q of lane l holds row
// (q / 4) * 8 + (l / 32) * 4 + q % 4, column l % 32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fpgm {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// 1 in every byte of w equal to the byte pattern pat (pat = state * 0x01010101), else 0.
__device__ __forceinline__ uint32_t eq_bytes(uint32_t w, uint32_t pat)
{
    const uint32_t t = w ^ pat;
    const uint32_t y = (t & 0x7F7F7F7Fu) + 0x7F7F7F7Fu;
    return (~(y | t) & 0x80808080u) >> 7;
}

__device__ __forceinline__ uint32_t byte_mask_word(int64_t valid, int word)
{
    // bytes [4*word, 4*word+4) of a 16-byte group, valid if their index < valid
    const int64_t lo = (int64_t)word * 4;
    if (valid >= lo + 4) return 0xFFFFFFFFu;
    if (valid <= lo) return 0u;
    return 0xFFFFFFFFu >> (8 * (4 - (int)(valid - lo)));
}

// One-hot operand of one row/column of O for 16 rows.  Every cardinality <= 8
// (kPerm): the column bytes are states 0..7 and index the 8-byte table (lo, hi) whose
// byte `state` is 1, so v_perm_b32 yields the 16 one-hot bytes in 4 instructions.
// Otherwise a SWAR byte comparison against state * 0x01010101 (lo).
template <bool kPerm>
__device__ __forceinline__ v4i onehot(const uint4 &q, uint32_t lo, uint32_t hi)
{
    v4i r;
    if (kPerm) {
        r[0] = (int)__builtin_amdgcn_perm(hi, lo, q.x);
        r[1] = (int)__builtin_amdgcn_perm(hi, lo, q.y);
        r[2] = (int)__builtin_amdgcn_perm(hi, lo, q.z);
        r[3] = (int)__builtin_amdgcn_perm(hi, lo, q.w);
    } else {
        r[0] = (int)eq_bytes(q.x, lo);
        r[1] = (int)eq_bytes(q.y, lo);
        r[2] = (int)eq_bytes(q.z, lo);
        r[3] = (int)eq_bytes(q.w, lo);
    }
    return r;
}

__device__ __forceinline__ v4i and4(v4i a, const uint4 &m)
{
    a[0] &= (int)m.x;
    a[1] &= (int)m.y;
    a[2] &= (int)m.z;
    a[3] &= (int)m.w;
    return a;
}

// Lookup words of one-hot row i: (lo, hi) for the v_perm_b32 form, (pattern, -) for
// the byte comparison.  Padding rows (state 255) never match in either form.
template <bool kPerm>
__device__ __forceinline__ void onehot_lut(uint8_t state, uint32_t &lo, uint32_t &hi)
{
    if (kPerm) {
        lo = (state < 4) ? 1u << (8 * state) : 0u;
        hi = (state >= 4 && state < 8) ? 1u << (8 * (state - 4)) : 0u;
    } else {
        lo = (uint32_t)state * 0x01010101u;
        hi = 0u;
    }
}

}  // namespace fpgm
