// Rounds of a level generated on the device.  The host plans each round per edge -- a
// segment is one edge's next `count` conditioning sets on one side of its walk, from a
// lexicographic rank on -- and the device expands the segments into tests: it unranks
// each segment's first subset, steps through the rest in the order of
// ChoiceGenerator::Next (ChoiceGenerator.cpp:55-85), maps subset positions to neighbours
// as FindAdjacencies / Testing do (PCStable.cpp:500-518, 540-552), routes every table to
// its counting sink and appends the work items.  After the counting, the reduction leaves
// per segment the index of its first independent test (Testing's result.first,
// PCStable.cpp:557-585), so that the host reads one word per segment, never a test.
#pragma once

#include "ci_kernels.h"

namespace fpgm {

constexpr int kMaxDevD = 16;      // largest conditioning set of a device-generated round
constexpr int kRouteGroups = 4;   // sink groups of a device-generated round: per-lane, lane-wide, shared, bit planes
// a round's group index of a route_lds group (kRouteBits -> 3)
__host__ __device__ constexpr int round_group(int g) { return g == kRouteBits ? 3 : g; }
constexpr int kExpandPer = 8;     // consecutive tests per thread of the expansion
constexpr int32_t kSegRun = 0x7fffffff;   // seg_out: the segment ran, no test independent
constexpr int32_t kSegSkipped = -2;       // seg_out: the segment's edge was done before the round

// One edge's tests of a round on one side of its walk (48 bytes).
struct RoundSeg {
    int32_t edge;    // index of the edge in the level (edge_done)
    int32_t inst;    // dataset instance
    int32_t x, y;
    int32_t a;       // |adj(u) \ {other}|: the subsets are drawn from a elements
    int32_t skip;    // position of the other endpoint in adj(u)
    int32_t u_off;   // adj(u) in the level's neighbour store
    int32_t count;   // tests of the segment (> 0)
    int64_t rank0;   // lexicographic rank of its first subset
    int64_t t0;      // its first test in the round
};

// Per-round figures, cleared before the expansion.
struct RoundStats {
    unsigned long long tests[kRouteGroups];  // tests routed to each group (skipped ones excluded)
    unsigned long long items[kRouteGroups];  // their work items
    unsigned long long rows[kRouteGroups];   // rows scanned (sum over tests)
    unsigned long long alg[kRouteGroups];    // algorithmic bytes, (|Z| + 2) N per test
    unsigned long long pbits[kRouteGroups];  // packed-column bits x rows
    unsigned long long ws_next;              // workspace bump pointer (int32 entries)
    uint32_t in_band;                        // executed tests in the FP64 recheck band
    uint32_t max_err_bits;                   // largest |G2_32 - G2_64| / G2_64 (FP32 bits) of FP64-decided tests
    uint32_t err;                            // bits: 1 a table beyond LDS, 2 item overflow, 4 workspace overflow, 8 rank out of range,
                                             // 16 a table routed to a group without a launch
    uint32_t pad;
};

struct ExpandArgs {
    const RoundSeg *segs;
    int64_t n_segs, n_tests;
    int32_t d;
    int32_t binom_stride;      // d + 1
    const int32_t *nbr;        // the level's neighbour store (every instance's frozen adjacency)
    const int64_t *binom;      // binom[n * binom_stride + k] = C(n, k), saturated
    const int32_t *card;
    const uint8_t *colw;
    const DevInst *inst;
    const int32_t *edge_done;
    int32_t round;
    int32_t w128_shared;       // 97..128-cell tables to the shared group (else the lane-wide one)
    uint32_t launch_mask;      // bit g: group g has a counting launch this round
    RouteParams rp;
    int64_t rows_per_chunk;
    DevTest *tests;
    int32_t *zpool;            // test t's Z at t * d
    int32_t *t2s;              // segment of each test
    DevResult *res;
    uint64_t *items;           // group g's items at g * group_cap
    int64_t group_cap;
    uint32_t *group_count;     // [kRouteGroups], zero before the launch (the counting launches' item counts)
    int64_t ws_cap;            // workspace entries available to the round
    int32_t *seg_out;
    RoundStats *stats;
};

struct ReduceArgs {
    const RoundSeg *segs;
    const int32_t *t2s;
    const DevResult *res;
    int64_t n_tests;
    int32_t *seg_out;
    RoundStats *stats;
};

hipError_t launch_expand(const ExpandArgs &e, hipStream_t s);
hipError_t launch_round_reduce(const ReduceArgs &r, hipStream_t s);

}  // namespace fpgm
