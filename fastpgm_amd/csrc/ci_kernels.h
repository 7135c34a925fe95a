// Device-side data structures and launch wrappers of the CI counting kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fpgm {

constexpr int kWG = 256;                  // threads per workgroup (4 wave64)
constexpr int kRowsPerThread = 16;        // one 16-byte load per column per step
constexpr int kRowsPerStep = kWG * kRowsPerThread;  // 4096 rows per workgroup step
constexpr int kLdsCells = 8192;           // LDS table capacity in int32 cells (32 KiB)
constexpr int kLaneWords = kLdsCells / (4 * 64);    // 16-bit counter pairs per lane (4 waves)
constexpr int kLaneCells = 2 * kLaneWords;          // largest table counted per lane (64)
constexpr int kLaneWideCells = 96;  // per-lane sink with 48 KiB of counters: 3 workgroups per CU
constexpr int kLaneWide128Cells = 128;  // ... its 64 KiB instance (2 workgroups per CU)
constexpr int kSmallCells = 512;          // LDS table + marginals of per-lane tests
constexpr int64_t kLaneMaxRowsPerChunk = (int64_t)kRowsPerStep * 4000;  // 16-bit counters never wrap
constexpr int kMaxZLds = 64;              // Z columns staged in LDS per test
constexpr int kCritTab = 4096;            // critical values tabulated for df = 1..kCritTab
constexpr int kCritLds = 128;             // ... of which the counting kernel keeps df <= kCritLds in LDS
constexpr int kCritLdsShared = 640;       // ... (the shared sink's tables, up to kLdsCells cells, have larger df)

// One dataset instance: uint8 columns, column v at cols + v*stride, n_rows valid bytes.
// The counting kernels read the packed copy instead: column v at pcols + v*stride, its
// values at colw[v] bits each (8, 4 or 2; see k_pack_cols for the layout).
struct DevInst {
    const uint8_t *cols;
    int64_t stride;
    int64_t n_rows;
    const uint8_t *pcols;
    // bit planes (large instances): plane (v, i) -- bit r of word r / 32 is bit i of row
    // r's value -- at planes + (poff[v] + i) * pwords, nbits_of(card) planes per variable;
    // rows past n_rows read as value 0; nullptr when the instance has none
    const uint32_t *planes;
    int64_t pwords;
};

// One test as the device sees it (48 bytes).
struct DevTest {
    int32_t x, y, d, z_off;  // z_off into the device z pool
    int32_t inst;            // dataset instance
    int32_t n_slices;        // z slices of a sliced table (1 otherwise)
    int32_t slice_cells;     // cells per slice (whole z planes)
    int32_t edge;            // skeleton edge of the level (skip checks), -1 for none
    int32_t cx, cy;
    int32_t dimz;            // product of the Z cardinalities
    int32_t n_chunks;        // row chunks of this test (work items)
    int64_t ws_off;          // int32 offset of the test's workspace, -1 when none
    int32_t item0;           // first work item of the test
    int32_t flags;           // kTF* below
};

constexpr int32_t kTFLdsTable = 1;   // table counted in LDS
constexpr int32_t kTFLdsMarg = 2;    // single chunk and table+marginals fit in LDS
constexpr int32_t kTFExact = 4;      // FP64 statistic and p-value regardless of the band
constexpr int32_t kTFKeep = 8;       // leave the counted table in the workspace (parity reads)
constexpr int32_t kTFLane = 16;      // per-lane 16-bit LDS counters (cells <= kLaneCells)
constexpr int32_t kTFWide = 32;      // statistic by the grid-wide finalize kernels, not the counting kernel
constexpr int32_t kTFNarrow = 64;    // every column of the test packed at 4 or 2 bits per value
constexpr int32_t kTFCol = 128;      // per-lane sink launch, table in a column layout (<= 2 kColCells cells, |Z| = 1)
constexpr int32_t kTFBits = 256;     // per-lane sink launch, counted from bit planes (|Z| = 1, small tables)
constexpr int32_t kTFPart = 512;     // z-slice launch over rows partitioned by slice first (k_part_keys)
constexpr int kBitsMaxCard = 16;     // variables with planes: cardinality at most this
// planes of a variable of c states: bits of its largest value
__host__ __device__ constexpr int nbits_of(int c) { return c <= 1 ? 0 : c <= 2 ? 1 : c <= 4 ? 2 : c <= 8 ? 3 : 4; }
constexpr int kBitsWords = 4;        // plane words per thread and step (128 rows)
constexpr int64_t kPlaneAlign = 1024;  // plane length in words: a multiple of one workgroup step (kWG * kBitsWords)

// Grid-wide finalize of large device-memory tables: one workgroup per range of z slices
// writes a partial, then one thread per test sums its partials in order.
constexpr int kWideZ = 256;          // z slices per finalize workgroup
struct WideItem {
    int32_t t;       // test index within the launch
    int32_t z0, z1;  // z slices [z0, z1)
    int32_t pad;
};
struct WidePart {
    double g64;      // FP64 G^2 of the range
    float s, c, as;  // FP32 compensated sum, compensation, sum of |terms|
    int32_t df;
};
struct WideTest {
    int32_t t;              // test index within the launch
    int32_t part0, nparts;  // its partials
    int32_t pad;
};

// Partitioned large tables (kTFPart): a first launch (k_part_keys, one work item per
// kPartRows rows of a test) computes every row's key, counts its rows per z slice in LDS
// and stores each row's slice-relative key (uint16) grouped by slice within the item's
// rows; the z-slice launch then counts slice s of the test from those keys alone, every
// row once, and reduces the slice's strata to a partial statistic (its slice holds them
// whole), which the last slice combines.  Its partition buffer (int32 words from the
// test's offset): marginal scratch [dimz (cx + cy + 1)], the slices' partials, the
// (offset, count) of every (row chunk, slice) segment, then the keys, row chunk c at key
// c * kPartRows.
constexpr int kPartRows = 8192;       // rows per partitioning work item (2 workgroup steps)
constexpr int kPartStage = 128;       // row chunks whose segments a slice item stages at a time
constexpr int kPartItemChunks = 8;    // row chunks per partitioning work item
constexpr int kPartPrivSlices = 16;   // at most this many slices: a private counter per slice and thread
constexpr int kPartMaxSlices = 4096;  // z slices of a partitioned table at most
struct PartLayout {
    int64_t parts, seg, keys, words;
};
__host__ __device__ inline int64_t part_round64(int64_t w) { return (w + 63) / 64 * 64; }
__host__ __device__ inline PartLayout part_layout(int64_t dimz, int32_t cx, int32_t cy, int64_t n_slices,
                                                  int64_t n_rows)
{
    const int64_t chunks = (n_rows + kPartRows - 1) / kPartRows;
    PartLayout L;
    L.parts = part_round64(dimz * (cx + cy + 1));
    L.seg = L.parts + part_round64(n_slices * (int64_t)(sizeof(WidePart) / sizeof(int32_t)));
    L.keys = L.seg + part_round64(2 * chunks * n_slices);
    L.words = L.keys + part_round64((chunks * kPartRows + 1) / 2);
    return L;
}

// Result of one test on the device (32 bytes).
struct DevResult {
    float g2_32;      // FP32 compensated G^2
    float abs_sum;    // sum of |terms|, feeds the FP32 error bound
    int32_t df;       // kSkippedDf: not run, its edge found an independent test in an earlier round
    uint8_t indep, in_band, exact, need_p;
    double g2_64;     // FP64 G^2 (exact/in-band tests)
    double p64;       // FP64 p-value (exact/in-band tests), NaN otherwise
};

struct LaunchArgs {
    const DevInst *inst;
    const DevTest *tests;
    const int32_t *zpool;
    const int32_t *card;
    const uint8_t *colw;     // bits per value of each packed column: 8, 4 or 2
    const int32_t *poff;     // first bit plane of each variable (-1: none)
    const uint64_t *items;   // work items: test << 32 | (slice * n_chunks + chunk)
    int64_t item_begin;      // this launch's items: [item_begin, item_begin + n_items)
    int64_t n_items;         // (an upper bound when n_items_dev is set)
    const uint32_t *n_items_dev;  // item count written on the device (compacted rounds), or nullptr
    int64_t rows_per_chunk;
    int32_t *ws;
    uint32_t *done;
    const double *crit;      // crit[df-1] for df = 1..kCritTab at this alpha
    double alpha;
    double z_alpha;          // standard normal quantile of 1 - alpha (critical values beyond the table)
    float band;
    uint32_t flags;          // FPGM_F_* of the call
    DevResult *res;
    // speculative rounds: edge_done[e] = first round in which a test of edge e was found
    // independent (INT32_MAX while none); a test of this launch (round `round`) whose edge
    // was done in an earlier round is skipped.  edge_done == nullptr: no skipping.
    int32_t *edge_done;
    int32_t round;
    // work queue of this launch: kXcds counters, zero at launch; workgroups of XCD x take
    // items of slab x in turn, then of the other slabs
    uint32_t *queue;
    int32_t *part;           // partition buffer of kTFPart tests (their ws_off indexes it)
};

// The partitioning launch: work items test << 32 | first row chunk (kPartItemChunks chunks
// of kPartRows rows each).
struct PartArgs {
    LaunchArgs a;
    const uint64_t *items;
    int64_t n_items;
};

constexpr int kXcds = 8;  // MI355X: 8 XCDs, workgroups dispatched round-robin over them

constexpr int32_t kSkippedDf = -1;

// Counting sinks (one kernel each; a batch launches each sink over its items).
constexpr int kSinkShared = 0;  // LDS histogram shared by waves (copies per wave when small)
constexpr int kSinkGlobal = 1;  // device-memory table, device atomics (beyond LDS)
constexpr int kSinkLane = 2;    // per-lane 16-bit LDS counters (tiny tables)
constexpr int kSinkLaneW = 3;   // per-lane 16-bit LDS counters, 48 KiB (<= 96 cells)
constexpr int kSinkSlice = 4;   // large tables: one z slice (<= kSliceCells) per work item in LDS
constexpr int kSinkLaneW128 = 5;  // kSinkLaneW's instance for tables <= 128 cells (a launch mode, not a group)
constexpr int kSinkBits = 6;    // per-lane 16-bit LDS counters filled from bit planes (small |Z| = 1 tables)
constexpr int kSliceCells = 2 * kLdsCells;  // 64 KiB LDS slice table
constexpr int kColCells = kLdsCells / 32;  // shared sink column layouts: [cell][64] to 128 cells, [cell][32] to 256, [cell][16] to 512
constexpr int64_t kColMinRows = 65536;     // ... for work items of at least this many rows
constexpr int kMaxSlices = 8;              // more slices than this: device-memory sink (measured)

// Sink routing of one table of at most kLdsCells cells, shared by the host's launch
// builder and the device's round expansion.  Returns the group of the launch it joins:
// 0 per-lane (<= 64 cells, or a column layout up to 512), 1 lane-wide (<= 96),
// kRoutePending (97..128: lane-wide or shared, decided per launch), 2 shared
// (<= kLdsCells), kRouteBits (bit planes); -1 for larger tables.  Sets the table's kTF* bits in flags (kTFNarrow
// must be set already); *fits: table and marginals fit the workgroup's LDS scratch.
struct RouteParams {
    uint32_t sink_mask;     // bit 0: no lane-wide sink, bit 2: no per-lane sink, bit 3: no column layout
    int32_t packed;         // the counting kernels read packed columns
    int64_t wide_max_rows;  // instances of at least this many rows: column layout, no lane-wide sink
    int32_t bits_cells;     // |Z| = 1 tables of at most this many cells count from bit planes (0: none)
};
constexpr int kRouteBits = 5;
constexpr int kRoutePending = 6;

// Cardinalities of a bit-plane table of at most kLaneCells cells: sorted, the two smallest
// (P, Q) are one of the counting kernel's instances -- (2, 2..5), (3, 3..4), (4, 4) -- and
// every variable has planes.
__host__ __device__ inline bool bits_shape(int64_t c0, int64_t c1, int64_t c2)
{
    if (c0 > c1) { const int64_t t = c0; c0 = c1; c1 = t; }
    if (c1 > c2) { const int64_t t = c1; c1 = c2; c2 = t; }
    if (c0 > c1) { const int64_t t = c0; c0 = c1; c1 = t; }
    return c2 <= kBitsMaxCard && c0 * c1 * c2 <= kLaneCells &&
           ((c0 == 2 && c1 >= 2 && c1 <= 5) || (c0 == 3 && c1 >= 3 && c1 <= 4) || (c0 == 4 && c1 == 4));
}

__host__ __device__ inline int route_lds(const RouteParams &p, int32_t &flags, int64_t cells, int64_t marg,
                                         int64_t rows, int32_t cx, int32_t cy, int32_t d, bool *fits)
{
    // small |Z| = 1 tables of instances with bit planes: one AND and one population count
    // per 32 rows and cell, no per-row LDS operation
    if (p.bits_cells > 0 && d == 1 && cells <= p.bits_cells && cells <= kLaneCells &&
        bits_shape(cx, cy, cells / ((int64_t)cx * cy)) && !(p.sink_mask & 4)) {
        flags |= kTFBits | kTFLane;
        *fits = cells + marg <= kSmallCells;
        return kRouteBits;
    }
    if (cells <= kLaneCells && !(p.sink_mask & 4)) {
        flags |= kTFLane;
        *fits = cells + marg <= kSmallCells;
        return 0;
    }
    // 65..512-cell tables of long instances (|Z| = 1, columns of the launch's form) count
    // in a column layout -- [cell][64] int32 words up to 128 cells, [cell][32] up to 256,
    // [cell][16] up to 512 (then c(X) c(Y) <= 256, Z the page): 32 KiB, lane l in column
    // l (l & 31, l & 15) -- within the per-lane sink's launch, whose kernel holds its key
    // arithmetic at 128 VGPRs; short ones per lane (their items are too short for the
    // column layout to repay its clearing and reduction)
    if (cells <= 2 * (kLdsCells / 32) && (cells <= kLdsCells / 32 || (int64_t)cx * cy <= 256) &&
        rows >= p.wide_max_rows && d == 1 && (!p.packed || (flags & kTFNarrow)) && !(p.sink_mask & 8)) {
        flags |= kTFCol | kTFLdsTable;
        *fits = cells + marg <= kLdsCells;
        return 0;
    }
    const bool wide = !(p.sink_mask & 1) && rows < p.wide_max_rows;
    if (cells <= kLaneWideCells && wide) {
        *fits = cells + marg <= kSmallCells;
        return 1;
    }
    if (cells <= kLaneWide128Cells && wide) {
        *fits = cells + marg <= kSmallCells;
        return kRoutePending;
    }
    if (cells <= kLdsCells) {
        flags |= kTFLdsTable;
        *fits = cells + marg <= kLdsCells;
        return 2;
    }
    *fits = false;
    return -1;
}

// Compaction of a round's work items before its counting launches: items of tests whose
// edge was done in an earlier round are dropped (their result is marked skipped), the rest
// keep their order within each 256-item tile; count[g] receives group g's live items.
struct CompactArgs {
    const uint64_t *in;
    uint64_t *out;
    const DevTest *tests;
    const int32_t *edge_done;
    int32_t round;
    DevResult *res;
    int64_t begin[6];  // per sink group: first item (in and out use the same group regions)
    int64_t len[6];
    uint32_t *count;   // [6], zero before the launch
};

// Grouped level-1 counting on the MFMA units.  The tests (x, y | z) of a round that share
// the walking endpoint u (x for the x side of an edge's walk, y for its y side) and the
// conditioning node z form a group; with `other` the remaining endpoint of each test,
// the group's tables are the blocks of C = A^T B, A the one-hot of (u, z) over the rows
// (row index c * c(u) + a for u = a, z = c) and B the one-hot of every other (columns
// concatenated in test order): test j's table n[z][x][y] is C[z * c(u) + u][col_j + other]
// read with x, y in their order.  A chunk holds the tests of one group whose columns fit
// kGrpB 32-column fragments; one work item = (chunk, row split).
constexpr int kGrpB = 4;                    // B fragments (32 one-hot columns each) per chunk
constexpr int kGrpCols = 32 * kGrpB;        // one-hot columns per chunk
constexpr int kGrpRows = 64;                // A rows per chunk: c(u) * c(z) <= 64 (two fragments)
constexpr int kGrpMaxOther = 32;            // largest c(other)
constexpr int kGrpMaxTests = 64;            // tests per chunk

struct GroupChunk {
    int32_t inst, u, z;
    int32_t cu, cz;
    int32_t t0, n_t;        // its tests: gtest[t0 .. t0 + n_t), indices into the launch's tests
    int32_t n_cols;         // one-hot columns: the sum of the others' cardinalities (<= kGrpCols)
    int32_t n_splits;       // row splits (work items) of the chunk
    int32_t done;           // its completion counter in LaunchArgs::done when split
    int64_t ws_off;         // kGrpRows x kGrpCols int32 partial table in the workspace when split, else -1
};

struct GroupArgs {
    LaunchArgs a;           // tests, results, critical values, skip state, workspace, counters
    const GroupChunk *chunks;
    const int32_t *gtest;
    const uint64_t *items;  // chunk << 32 | split
    int64_t n_items;
    int64_t rows_per_split; // multiple of 128
    int32_t max_card;       // largest cardinality of the launch's columns (<= 8: table lookup)
};

// packed: some column of the context is packed (the kernels' packed-column form)
hipError_t launch_count_stat(const LaunchArgs &a, int mode, bool packed, int grid, hipStream_t s);
hipError_t launch_compact(const CompactArgs &c, int64_t max_len, hipStream_t s);
hipError_t launch_part_keys(const PartArgs &p, bool packed, int grid, hipStream_t s);
hipError_t launch_finalize_wide(const LaunchArgs &a, const WideItem *items, int64_t n_items, WidePart *parts,
                                const WideTest *wt, int64_t n_wide, hipStream_t s);
// p-values of the rechecked tests; with edge_done, an independent one marks its edge (round).
// It also zeroes `clear` (n_clear counters: the launch sequence's work queues) for the next
// sequence on the stream.
hipError_t launch_pvalue(DevResult *res, int64_t n, double alpha, const DevTest *tests, int32_t *edge_done,
                         int32_t round, uint32_t *clear, int n_clear, hipStream_t s);
// Packed copy of one instance's columns (16-row groups; colw[v] bits per value).
hipError_t launch_pack_cols(const uint8_t *cols, uint8_t *pcols, int64_t stride, int n_vars, const uint8_t *colw,
                            hipStream_t s);
hipError_t launch_check_values(const uint8_t *cols, int64_t stride, int64_t n_rows, int n_vars,
                               const int32_t *card, uint32_t *bad, hipStream_t s);
// Grouped level-1 chunks on the MFMA units (fused with each test's statistic).
// n_a: A fragments the launch's chunks need (1: c(u) * c(z) <= 32, else 2).
hipError_t launch_group_count(const GroupArgs &g, int n_a, hipStream_t s);
// Zeroes n 32-bit words at p (a fill kernel: no memset node in code a caller may capture).
hipError_t launch_zero(uint32_t *p, int64_t n, hipStream_t s);
// Bit planes of one instance (see DevInst): planes of variable v at poff[v] (none when
// poff[v] < 0), pwords words each.
hipError_t launch_build_planes(const uint8_t *cols, int64_t stride, int64_t n_rows, int n_vars, const int32_t *card,
                               const int32_t *poff, uint32_t *planes, int64_t pwords, hipStream_t s);

}  // namespace fpgm
