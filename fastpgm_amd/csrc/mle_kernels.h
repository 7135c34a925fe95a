// Maximum-likelihood counts of a known structure: one table per node over (node value,
// parent configuration), the N(AB) table of parameter learning.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fpgm {

// One node's table as the device sees it.
struct MleNode {
    int32_t n_rel;     // related variables: the node, then its parents (ascending)
    int32_t rel_off;   // into MleArgs::rel
    int64_t out_off;   // int32 offset of the table in MleArgs::out
    int32_t size;      // cells: card(node) * prod card(parents)
    int32_t item0;     // first work item of the node
    int32_t n_chunks;  // row chunks of the node (work items)
    int32_t pad;
};

struct MleArgs {
    const uint8_t *cols;
    int64_t stride;
    int64_t n_rows;
    const MleNode *nodes;
    const int32_t *rel;        // related variable ids, node first
    const int32_t *card;
    const int32_t *item_node;  // work item -> node
    int64_t n_items;
    int64_t rows_per_chunk;    // multiple of the rows a workgroup step covers (4096)
    int32_t *out;              // all tables back to back, zero before the launch
};

constexpr int kMleMaxRel = 64;  // node + parents staged in LDS per work item

hipError_t launch_mle_counts(const MleArgs &a, int grid, hipStream_t s);

}  // namespace fpgm
