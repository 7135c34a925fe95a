// C entry points over FastPGM's own compiled classes, for cross-checking the
// CPU restatement.  TEST INFRASTRUCTURE ONLY, built by oracle/ref.mk from the
// reference sources where they lie (never copied into this repository) into
// oracle/_ref/libfpgm_ref.so.
//
// Exposed reference code paths:
//   Dataset::LoadCSVTrainingData  (Dataset.cpp:389-503, first-appearance value coding)
//   Counts2D / Counts3D / Counts3DGroup  (CellTable.cpp:23-455)
//   ChoiceGenerator::Next  (ChoiceGenerator.cpp:14-85)
//   ParameterLearning::LearnParamsKnowStructCompData / SaveBNParameter
//                          (ParameterLearning.cpp:10-155)
//   BNSLComparison::GetSHD (BNSLComparison.cpp:12-73, with Network::OrderEdge,
//                           FindCompelled, IsDAG, GetEdge: Network.cpp:122-129, 547-699)
// PCStable.cpp and IndependenceTest.cpp include the third-party stats/gcem headers,
// which are not in the source tree, so they are not built here.
#include "fastpgm/ChoiceGenerator.h"
#include "fastpgm/Dataset.h"
#include "fastpgm/DiscreteNode.h"
#include "fastpgm/Network.h"
#include "fastpgm/parameter_learning/ParameterLearning.h"
#include "fastpgm/structure_learning/BNSLComparison.h"
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

// A reference Network with nodes 0..n-1 named as given (names may be null: "0", "1", ...),
// domain sizes card and value names "0".."c-1" unless the dataset gives its own.
Network *make_network(int n, const int32_t *card, const Dataset *dts)
{
    Network *net = new Network(true);
    net->num_nodes = n;
    for (int i = 0; i < n; ++i) {
        DiscreteNode *node = new DiscreteNode(i);
        if (dts && (int)dts->vec_var_names.size() == n) node->node_name = dts->vec_var_names[i];
        else node->node_name = std::to_string(i);
        if (dts && (int)dts->vars_possible_values_ids.size() == n) {
            node->possible_values_ids = dts->vars_possible_values_ids[i];
        } else {
            for (int k = 0; k < card[i]; ++k) node->possible_values_ids[std::to_string(k)] = k;
        }
        node->SetDomainSize(card[i]);
        net->map_idx_node_ptr[i] = node;
    }
    return net;
}

void add_arc(Network *net, int p, int c, bool edge)
{
    Node *np = net->FindNodePtrByIndex(p), *nc = net->FindNodePtrByIndex(c);
    net->SetParentChild(np, nc);
    if (edge) {
        Edge e(np, nc, TAIL, ARROW);
        e.label = COMPELLED;  // the reference leaves it unset; an edge FindCompelled skips keeps it
        net->vec_edges.push_back(e);
        ++net->num_edges;
    }
}

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

// ParameterLearning on a network whose node i has the parents parents[par_off[i] ..
// par_off[i+1]) over the columns (or over the CSV dataset h when non-null, with its
// names and value coding).  potentials receives every table back to back (cap entries);
// returns the total, or -1 when cap is too small.  param_file: SaveBNParameter output.
int64_t ref_mle(void *h, const uint8_t *cols, int64_t stride, int32_t n_vars, int32_t n_rows, const int32_t *card,
                const int32_t *par_off, const int32_t *parents, int32_t alpha, double *potentials, int64_t cap,
                const char *param_file)
{
    ColumnDataset *cd = nullptr;
    const Dataset *dts = (const Dataset *)h;
    if (!dts) {
        cd = new ColumnDataset(cols, stride, n_vars, n_rows, card);
        dts = &cd->dts;
    }
    std::vector<int32_t> cards(dts->num_of_possible_values_of_disc_vars.begin(),
                               dts->num_of_possible_values_of_disc_vars.end());
    Network *net = make_network(dts->num_vars, cards.data(), h ? dts : nullptr);
    for (int i = 0; i < dts->num_vars; ++i)
        for (int j = par_off[i]; j < par_off[i + 1]; ++j) add_arc(net, parents[j], i, true);
    ParameterLearning pl(net);
    pl.LearnParamsKnowStructCompData(dts, alpha, param_file != nullptr, param_file ? param_file : "", 0);
    int64_t total = 0;
    for (const PotentialTableBase &pt : net->pts) total += pt.table_size;
    if (total <= cap) {
        int64_t k = 0;
        for (const PotentialTableBase &pt : net->pts)
            for (int j = 0; j < pt.table_size; ++j) potentials[k++] = pt.potentials[j];
    }
    delete net;
    delete cd;
    return total <= cap ? total : -1;
}

// SHD of a learned graph against a true DAG (BNSLComparison::GetSHD).  true_arcs: n_true
// (parent, child) pairs in the reference BIF loader's order; learned: n_learned records
// (a, b, directed) in vec_edges order, directed = a -> b, else the undirected edge {a, b}.
int32_t ref_shd(int32_t n, const int32_t *true_arcs, int32_t n_true, const int32_t *learned, int32_t n_learned)
{
    std::vector<int32_t> card(n, 2);
    Network *t = make_network(n, card.data(), nullptr);
    Network *l = make_network(n, card.data(), nullptr);
    for (int k = 0; k < n_true; ++k) add_arc(t, true_arcs[2 * k], true_arcs[2 * k + 1], true);
    for (int k = 0; k < n_learned; ++k) {
        const int a = learned[3 * k], b = learned[3 * k + 1];
        if (learned[3 * k + 2]) {
            add_arc(l, a, b, true);
        } else {
            Edge e(l->FindNodePtrByIndex(a), l->FindNodePtrByIndex(b));
            e.label = COMPELLED;
            l->vec_edges.push_back(e);
            ++l->num_edges;
        }
    }
    BNSLComparison comp(t, l);
    const int shd = comp.GetSHD();
    delete t;
    delete l;
    return shd;
}

}  // extern "C"
