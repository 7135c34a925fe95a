// C entry points over FastPGM's own compiled classes, for cross-checking the
// CPU restatement.  TEST INFRASTRUCTURE ONLY, built by oracle/ref.mk from the
// reference sources where they lie (never copied into this repository) into
// oracle/_ref/libfpgm_ref.so.
//
// Exposed reference code paths:
//   Dataset::LoadCSVTrainingData  (Dataset.cpp:389-503, first-appearance value coding)
//   Counts2D / Counts3D / Counts3DGroup  (CellTable.cpp:23-455)
//   ChoiceGenerator::Next  (ChoiceGenerator.cpp:14-85)
// PCStable.cpp and IndependenceTest.cpp include the third-party stats/gcem headers,
// which are not in the source tree, so they are not built here.
#include "fastpgm/ChoiceGenerator.h"
#include "fastpgm/Dataset.h"
#include "fastpgm/structure_learning/CellTable.h"

#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct ColumnDataset {
    Dataset dts;
    std::vector<std::vector<int>> store;
    std::vector<int *> ptrs;

    ColumnDataset(const uint8_t *cols, int64_t stride, int n_vars, int n_rows, const int32_t *card)
    {
        dts.num_instance = n_rows;
        dts.num_vars = n_vars;
        dts.num_of_possible_values_of_disc_vars.assign(card, card + n_vars);
        store.resize(n_vars);
        ptrs.resize(n_vars);
        for (int v = 0; v < n_vars; ++v) {
            store[v].resize(n_rows > 0 ? n_rows : 1);
            for (int r = 0; r < n_rows; ++r) store[v][r] = cols[(int64_t)v * stride + r];
            ptrs[v] = store[v].data();
        }
        dts.dataset_columns = ptrs.data();
    }
    ~ColumnDataset() { dts.dataset_columns = nullptr; }
};

}  // namespace

extern "C" {

// Loads a CSV with the reference loader.  Returns a handle; sizes via ref_csv_dims.
void *ref_csv_load(const char *path)
{
    Dataset *d = new Dataset();
    d->LoadCSVTrainingData(path, true, -1);
    return d;
}

void ref_csv_dims(void *h, int32_t *n_vars, int32_t *n_rows)
{
    Dataset *d = (Dataset *)h;
    *n_vars = d->num_vars;
    *n_rows = d->num_instance;
}

// cols: n_vars x stride uint8; card: n_vars
void ref_csv_fetch(void *h, uint8_t *cols, int64_t stride, int32_t *card)
{
    Dataset *d = (Dataset *)h;
    for (int v = 0; v < d->num_vars; ++v) {
        card[v] = d->num_of_possible_values_of_disc_vars[v];
        for (int r = 0; r < d->num_instance; ++r) cols[(int64_t)v * stride + r] = (uint8_t)d->dataset_columns[v][r];
    }
}

void ref_csv_free(void *h)
{
    delete (Dataset *)h;
}

// Counts of X,Y|Z with the reference classes; n_out holds dimz*cx*cy ints,
// ni_out dimz*cx, nj_out dimz*cy, nk_out dimz (nk_out unused for d == 0).
void ref_counts(const uint8_t *cols, int64_t stride, int32_t n_vars, int32_t n_rows, const int32_t *card, int32_t x,
                int32_t y, const int32_t *z, int32_t d, int32_t *n_out, int32_t *ni_out, int32_t *nj_out,
                int32_t *nk_out)
{
    ColumnDataset cd(cols, stride, n_vars, n_rows, card);
    const int cx = card[x], cy = card[y];
    if (d == 0) {
        Counts2D t(cx, cy, x, y);
        t.FillTable(&cd.dts, nullptr);
        std::memcpy(n_out, t.n, sizeof(int) * cx * cy);
        std::memcpy(ni_out, t.ni, sizeof(int) * cx);
        std::memcpy(nj_out, t.nj, sizeof(int) * cy);
        return;
    }
    std::vector<int> dims, idx;
    for (int j = 0; j < d; ++j) {
        dims.push_back(card[z[j]]);
        idx.push_back(z[j]);
    }
    Counts3D t(cx, cy, x, y, dims, idx);
    t.FillTable(&cd.dts, nullptr);
    std::memcpy(n_out, t.n, sizeof(int) * t.dimz * cx * cy);
    std::memcpy(ni_out, t.ni, sizeof(int) * t.dimz * cx);
    std::memcpy(nj_out, t.nj, sizeof(int) * t.dimz * cy);
    std::memcpy(nk_out, t.nk, sizeof(int) * t.dimz);
}

// Grouped counting (Counts3DGroup) of g tests sharing x, y, each with d conditioning
// variables (z holds g*d ids); n_out receives the concatenated tables.
void ref_counts_group(const uint8_t *cols, int64_t stride, int32_t n_vars, int32_t n_rows, const int32_t *card,
                      int32_t x, int32_t y, const int32_t *z, int32_t d, int32_t g, int32_t *n_out)
{
    ColumnDataset cd(cols, stride, n_vars, n_rows, card);
    const int cx = card[x], cy = card[y];
    std::vector<int> dims, idx;
    for (int j = 0; j < g * d; ++j) {
        dims.push_back(card[z[j]]);
        idx.push_back(z[j]);
    }
    Counts3DGroup t(cx, cy, x, y, dims, idx, g);
    t.FillTableGroup(&cd.dts, g, nullptr);
    int total = 0;
    for (int m = 0; m < g; ++m) total += t.dimz[m];
    std::memcpy(n_out, t.n, sizeof(int) * total * cx * cy);
}

// Every combination the reference generator yields for a choose b; returns the count.
int64_t ref_choices(int32_t a, int32_t b, int32_t *out, int64_t cap)
{
    ChoiceGenerator cg(a, b);
    int64_t k = 0;
    for (;;) {
        std::vector<int> c = cg.Next();
        if (c.empty()) break;
        if (k < cap)
            for (int i = 0; i < b; ++i) out[k * b + i] = c[i];
        ++k;
    }
    return k;
}

}  // extern "C"
