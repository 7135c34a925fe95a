// Maximum-likelihood counting kernel for gfx950: the tables N(AB) of
// ParameterLearning::LearnParamsKnowStructCompData (ParameterLearning.cpp:33-69).
//
// For node v with parents p_1 < ... < p_k (a std::set in the reference), the table index
// of a row is PotentialTableBase's (PotentialTableBase.cpp:102-130):
//     key = ((v * c(p_1) + p_1) * c(p_2) + p_2) ... * c(p_k) + p_k
// i.e. the node most significant, then the parents in ascending id order.  One work item
// = (node, row chunk); a 256-thread workgroup streams the chunk's uint8 columns with
// 16-byte loads (16 rows per thread per step), counts into an LDS histogram (one copy per
// wave when it fits, so that skewed data collides in at most one wave) and adds the
// non-zero cells to the node's table with integer atomics; tables beyond LDS are counted
// with device atomics directly.  Counts are exact int32, as the reference's int vectors.
#include "ci_kernels.h"
#include "mle_kernels.h"

namespace fpgm {

namespace {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint4 ld16g(const uint8_t *p)
{
    const u32x4 v = *(const __attribute__((address_space(1))) u32x4 *)(p);
    return make_uint4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ uint32_t byte16(const uint4 &v, int b)
{
    const uint32_t w = (b < 4) ? v.x : (b < 8) ? v.y : (b < 12) ? v.z : v.w;
    return __builtin_amdgcn_ubfe(w, (b & 3) * 8, 8);
}

__global__ __launch_bounds__(kWG) void k_mle_counts(MleArgs a)
{
    __shared__ int32_t lds[kLdsCells];
    __shared__ int32_t s_col[kMleMaxRel];
    __shared__ int32_t s_card[kMleMaxRel];
    const int tid = threadIdx.x, wave = tid >> 6;
    for (int64_t item = blockIdx.x; item < a.n_items; item += gridDim.x) {
        const MleNode N = a.nodes[a.item_node[item]];
        const int64_t chunk = item - N.item0;
        const int64_t r_begin = chunk * a.rows_per_chunk;
        const int64_t r_end = (r_begin + a.rows_per_chunk < a.n_rows) ? r_begin + a.rows_per_chunk : a.n_rows;
        for (int j = tid; j < N.n_rel; j += kWG) {
            const int32_t v = a.rel[N.rel_off + j];
            s_col[j] = v;
            s_card[j] = a.card[v];
        }
        const bool in_lds = N.size <= kLdsCells;
        // per-wave copies when four fit, else one copy for the workgroup
        const bool per_wave = 4 * N.size <= kLdsCells;
        if (in_lds) {
            const int span = per_wave ? 4 * N.size : N.size;
            for (int i = tid; i < span; i += kWG) lds[i] = 0;
        }
        __syncthreads();
        int32_t *h = in_lds ? lds + (per_wave ? wave * N.size : 0) : a.out + N.out_off;
        for (int64_t r = r_begin + (int64_t)tid * kRowsPerThread; r < r_end; r += kRowsPerStep) {
            uint32_t key[kRowsPerThread];
#pragma unroll
            for (int b = 0; b < kRowsPerThread; ++b) key[b] = 0;
            for (int j = 0; j < N.n_rel; ++j) {
                const uint4 v = ld16g(a.cols + (int64_t)s_col[j] * a.stride + r);
                const uint32_t c = (uint32_t)s_card[j];
#pragma unroll
                for (int b = 0; b < kRowsPerThread; ++b) key[b] = key[b] * c + byte16(v, b);
            }
            const int64_t nv = r_end - r;
#pragma unroll
            for (int b = 0; b < kRowsPerThread; ++b)
                if (b < nv) atomicAdd(&h[key[b]], 1);
        }
        __syncthreads();
        if (in_lds) {
            int32_t *out = a.out + N.out_off;
            for (int i = tid; i < N.size; i += kWG) {
                int32_t s = lds[i];
                if (per_wave)
                    for (int w = 1; w < kWG / 64; ++w) s += lds[w * N.size + i];
                if (s) atomicAdd(&out[i], s);
            }
        }
        __syncthreads();  // the next item rewrites the LDS table and the staged columns
    }
}

}  // namespace

hipError_t launch_mle_counts(const MleArgs &a, int grid, hipStream_t s)
{
    if (a.n_items <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mle_counts, dim3(grid), dim3(kWG), 0, s, a);
    return hipGetLastError();
}

}  // namespace fpgm
