// Level-0 one-hot Gram path: C = O^T O on the MFMA units, then per-pair statistics.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ci_kernels.h"

namespace fpgm {

constexpr int kGramTile = 128;     // output tile edge (2 waves x 128 x 64)
constexpr int kPairMaxCard = 256;  // per-thread marginal array of the pair kernel

struct GramArgs {
    const uint8_t *cols;
    int64_t stride;
    int64_t n_rows;
    const int32_t *row_var;    // [Wp] variable of one-hot row i
    const uint8_t *row_state;  // [Wp] state of row i; 255 for padding rows (never matches)
    const int32_t *tiles;      // upper-triangle tile pairs, interleaved (i, j), i <= j
    int32_t Wp;                // padded width, multiple of kGramTile
    int32_t k_splits;          // row splits per tile pair
    int64_t k_per_split;       // rows per split, multiple of 32
    int32_t *C;                // [Wp][Wp] int32, zero before the launch
    int32_t max_card;          // largest cardinality (<= 8: one-hot bytes by table lookup)
    // several instances (insts != nullptr): item = (instance * n_tp + tile pair) * k_splits
    // + split, instance i's columns from insts[i] and its C at C + i * c_stride
    const DevInst *insts;
    int32_t n_tp;              // tile pairs per instance
    int64_t c_stride;
};

struct PairArgs {
    const int32_t *C;
    int32_t Wp;
    const int32_t *off;        // one-hot offset of each variable
    const int32_t *card;
    const int32_t *px, *py;    // pairs x < y; nullptr: all pairs of n_vars in lexicographic order
    int32_t n_vars;
    int64_t n_pairs;           // pairs per instance
    int32_t n_inst;            // instances: thread t takes pair t % n_pairs of instance t / n_pairs
    int64_t c_stride;          // C of instance i at C + i * c_stride
    const double *crit;
    double alpha, z_alpha;
    float band;
    uint32_t flags;
    DevResult *res;
};

// Level-0 results in one byte per pair (bit 0 independent, bit 1 in the recheck band) and
// the summary the host keeps: the largest relative FP32 error of G^2 against FP64 over the
// rechecked pairs (as float bits) and the number of in-band pairs.
struct PairSummary {
    uint32_t max_rel_err_bits;
    uint32_t in_band;
};

hipError_t launch_gram_l0(const GramArgs &g, int n_items, hipStream_t s);
hipError_t launch_pairs_l0(const PairArgs &p, hipStream_t s);
hipError_t launch_pack_l0(const DevResult *res, int64_t n, uint8_t *flags, PairSummary *sum, hipStream_t s);

}  // namespace fpgm
