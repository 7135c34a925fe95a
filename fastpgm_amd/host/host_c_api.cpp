// C entry points over the host mirror's structure comparison (fpgm_host_c.h).
#include <string>

#include "fpgm_host.h"
#include "fpgm_host_c.h"

namespace {

thread_local std::string g_err;

// A network on n nodes holding the given arcs (and undirected edges) in order.
fpgm::Network make_net(int32_t n, const int32_t *rec, int32_t n_rec, int stride)
{
    fpgm::Network net(true);
    net.num_nodes = n;
    net.node_names.resize(n);
    net.node_cards.assign(n, 2);
    net.possible_values_ids.resize(n);
    net.set_parent_indexes.assign(n, {});
    net.set_children_indexes.assign(n, {});
    for (int32_t k = 0; k < n_rec; ++k) {
        const int a = rec[stride * k], b = rec[stride * k + 1];
        const bool directed = (stride == 2) || rec[stride * k + 2] != 0;
        if (a < 0 || b < 0 || a >= n || b >= n || a == b) throw fpgm::Error("edge endpoint out of range");
        if (directed) {
            net.set_children_indexes[a].insert(b);
            net.set_parent_indexes[b].insert(a);
            net.vec_edges.emplace_back(a, b, fpgm::TAIL, fpgm::ARROW);
        } else {
            net.vec_edges.emplace_back(a, b);
        }
        ++net.num_edges;
    }
    return net;
}

template <typename F>
int32_t guarded(F f)
{
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
        return -1;
    }
}

}  // namespace

extern "C" {

const char *fpgm_host_last_error(void)
{
    return g_err.c_str();
}

int32_t fpgm_host_shd(int32_t n, const int32_t *true_arcs, int32_t n_true, const int32_t *learned,
                      int32_t n_learned, int32_t *shd)
{
    return guarded([&] {
        if (n < 1 || !shd || (n_true > 0 && !true_arcs) || (n_learned > 0 && !learned)) throw fpgm::Error("bad argument");
        fpgm::Network t = make_net(n, true_arcs, n_true, 2);
        fpgm::Network l = make_net(n, learned, n_learned, 3);
        fpgm::BNSLComparison comp(&t, &l);
        *shd = comp.GetSHD();
    });
}

int32_t fpgm_host_shd_bif(const char *bif_path, int32_t n, const int32_t *learned, int32_t n_learned, int32_t *shd)
{
    return guarded([&] {
        if (!bif_path || n < 1 || !shd || (n_learned > 0 && !learned)) throw fpgm::Error("bad argument");
        fpgm::CustomNetwork t;
        t.LoadBIFFile(bif_path);
        fpgm::Network l = make_net(n, learned, n_learned, 3);
        fpgm::BNSLComparison comp(&t, &l);
        *shd = comp.GetSHD();
    });
}

int32_t fpgm_host_bif_arcs(const char *bif_path, int32_t *n_nodes, int32_t *cards, int32_t cap_nodes, int32_t *arcs,
                           int32_t cap_arcs, int32_t *n_arcs)
{
    return guarded([&] {
        if (!bif_path || !n_nodes || !n_arcs) throw fpgm::Error("bad argument");
        fpgm::CustomNetwork t;
        t.LoadBIFFile(bif_path);
        *n_nodes = t.num_nodes;
        *n_arcs = (int32_t)t.vec_edges.size();
        if (cards)
            for (int i = 0; i < t.num_nodes && i < cap_nodes; ++i) cards[i] = t.node_cards[i];
        if (arcs)
            for (int k = 0; k < *n_arcs && k < cap_arcs; ++k) {
                arcs[2 * k] = t.vec_edges[k].node1;
                arcs[2 * k + 1] = t.vec_edges[k].node2;
            }
    });
}

int32_t fpgm_host_compelled(int32_t n, const int32_t *arcs, int32_t n_arcs, int32_t *labels)
{
    return guarded([&] {
        if (n < 1 || !labels || (n_arcs > 0 && !arcs)) throw fpgm::Error("bad argument");
        fpgm::Network t = make_net(n, arcs, n_arcs, 2);
        if (t.ContainCircle()) throw fpgm::Error("the arcs do not form a DAG");
        std::vector<fpgm::Edge> order = t.OrderEdge();
        t.FindCompelled(order);
        for (int32_t k = 0; k < n_arcs; ++k) labels[k] = (t.vec_edges[k].label == fpgm::REVERSIBLE) ? 1 : 0;
    });
}

}  // extern "C"
