// Structure comparison of the host mirror: BIF structure loading, Chickering's DAG-to-CPDAG
// labelling and the structural Hamming distance (CustomNetwork.cpp:52-230,
// Network.cpp:95-129, 547-699, BNSLComparison.cpp:7-73).
#include <algorithm>
#include <fstream>

#include "fpgm_host.h"

namespace fpgm {

namespace {

// The reference's line tokenizer (common.cpp:115-141): leading characters below '!' are
// dropped, then the line is cut at every single space (consecutive spaces give empty
// tokens).
std::string trim_left(const std::string &s)
{
    size_t i = 0;
    while (i < s.size() && (unsigned char)s[i] < 33) ++i;
    return s.substr(i);
}

std::vector<std::string> split_spaces(const std::string &s)
{
    std::vector<std::string> out;
    size_t begin = 0, end;
    while ((end = s.find(' ', begin)) != std::string::npos) {
        out.push_back(s.substr(begin, end - begin));
        begin = end + 1;
    }
    out.push_back(s.substr(begin));
    return out;
}

std::string trim_right_comma(std::string s)
{
    if (!s.empty() && s.back() == ',') s.pop_back();
    return s;
}

}  // namespace

int Network::GetEdge(int a, int b) const
{
    for (size_t i = 0; i < vec_edges.size(); ++i) {
        const Edge &e = vec_edges[i];
        if ((e.node1 == a && e.node2 == b) || (e.node2 == a && e.node1 == b)) return (int)i;
    }
    return -1;
}

int Network::GetDirectedEdge(int a, int b) const
{
    const Edge e(a, b, TAIL, ARROW);
    auto it = std::find(vec_edges.begin(), vec_edges.end(), e);
    return it == vec_edges.end() ? -1 : (int)(it - vec_edges.begin());
}

bool Network::IsDAG() const
{
    // the reference reads vec_edges.at(0), which throws on an empty graph
    if (vec_edges.empty()) throw Error("IsDAG: the network has no edges");
    return vec_edges[0].IsDirected();
}

std::string Network::GetValueNameByIndex(int node, int index) const
{
    if (node >= 0 && node < (int)possible_values_ids.size())
        for (const auto &kv : possible_values_ids[node])
            if (kv.second == index) return kv.first;
    return std::to_string(index);  // the reference returns nothing here (undefined)
}

// Chickering (1995): y runs over the topological order; for each y, its parents x from
// the highest ordered down, each unordered x -> y appended (Network.cpp:547-584).
std::vector<Edge> Network::OrderEdge()
{
    std::vector<Edge> edge_order;
    const std::vector<int> topo = GetTopoOrd();
    for (size_t j = 0; j < topo.size(); ++j) {
        const int y = topo[j];
        if (set_parent_indexes[y].empty()) continue;
        for (int k = (int)j - 1; k >= 0; --k) {
            const int x = topo[k];
            if (!set_parent_indexes[y].count(x)) continue;
            const int pos = GetDirectedEdge(x, y);
            if (pos < 0) throw Error("OrderEdge: parent without a stored arc");
            if (!vec_edges[pos].is_ordered) {
                vec_edges[pos].is_ordered = true;
                edge_order.push_back(vec_edges[pos]);
            }
        }
    }
    if (edge_order.size() != vec_edges.size())
        throw Error("OrderEdge: num of edges = " + std::to_string(vec_edges.size()) +
                    " but num of ordered edges = " + std::to_string(edge_order.size()));
    return edge_order;
}

// Chickering's labelling (Network.cpp:590-684), step for step: the lowest unknown edge
// x -> y; every compelled w -> x either compels x -> y and all edges into y (w not a
// parent of y) or compels w -> y; failing that, a parent z of y other than x that is not a
// parent of x compels x -> y and the unknown edges into y, else they are all reversible.
// As in the reference, the second branch does not stop at the first such z: every further
// z removes the then-first unknown edge from the order again, unlabelled.
void Network::FindCompelled(std::vector<Edge> &edge_order)
{
    auto pos_in_order = [&](int a, int b) -> int {
        const Edge e(a, b, TAIL, ARROW);
        auto it = std::find(edge_order.begin(), edge_order.end(), e);
        return it == edge_order.end() ? -1 : (int)(it - edge_order.begin());
    };
    auto arc = [&](int a, int b) -> Edge & {
        const int p = GetDirectedEdge(a, b);
        if (p < 0) throw Error("FindCompelled: missing arc");
        return vec_edges[p];
    };
    auto erase_first = [&]() {
        if (!edge_order.empty()) edge_order.erase(edge_order.begin());  // reference: undefined when empty
    };
    while (!edge_order.empty()) {
        const int x = edge_order[0].node1, y = edge_order[0].node2;
        bool tmp1 = false;
        for (int w : set_parent_indexes[x]) {
            if (arc(w, x).label != COMPELLED) continue;
            if (!set_parent_indexes[y].count(w)) {
                tmp1 = true;
                arc(x, y).label = COMPELLED;
                erase_first();
                for (int p : set_parent_indexes[y]) {
                    const int pos = pos_in_order(p, y);
                    arc(p, y).label = COMPELLED;
                    if (pos != -1) edge_order.erase(edge_order.begin() + pos);
                }
                break;
            } else {
                const int pos = pos_in_order(w, y);
                if (pos != -1) {
                    arc(w, y).label = COMPELLED;
                    edge_order.erase(edge_order.begin() + pos);
                }
            }
        }
        if (tmp1) continue;
        bool tmp2 = false;
        for (int z : set_parent_indexes[y]) {
            if (z == x || set_parent_indexes[x].count(z)) continue;
            tmp2 = true;
            arc(x, y).label = COMPELLED;
            erase_first();
            for (int p : set_parent_indexes[y]) {
                const int pos = pos_in_order(p, y);
                if (pos != -1) {
                    arc(p, y).label = COMPELLED;
                    edge_order.erase(edge_order.begin() + pos);
                }
            }
        }
        if (!tmp2) {
            arc(x, y).label = REVERSIBLE;
            erase_first();
            for (int p : set_parent_indexes[y]) {
                const int pos = pos_in_order(p, y);
                if (pos != -1) {
                    arc(p, y).label = REVERSIBLE;
                    edge_order.erase(edge_order.begin() + pos);
                }
            }
        }
    }
}

// Variables in declaration order (name, domain size, value names), then one arc per
// parent of every probability block in the order the block lists them
// (CustomNetwork.cpp:52-160).  Probability values are not kept.
void CustomNetwork::LoadBIFFile(const std::string &path)
{
    std::ifstream in(path);
    if (!in.is_open()) throw Error("LoadBIFFile: unable to open file " + path);
    *this = CustomNetwork();
    std::string line;
    std::getline(in, line);  // network ... {
    std::getline(in, line);  // }
    std::map<std::string, int> index_of;
    bool in_variables = true;
    int cur = -1;
    std::vector<int> parents;
    while (std::getline(in, line)) {
        const std::vector<std::string> t = split_spaces(trim_left(line));
        if (t.empty()) continue;
        if (in_variables && t[0] != "probability") {
            if (t[0] == "variable" && t.size() > 1) {
                index_of[t[1]] = num_nodes;
                node_names.push_back(t[1]);
                node_cards.push_back(0);
                possible_values_ids.emplace_back();
                ++num_nodes;
            } else if (t[0] == "type" && !node_names.empty()) {
                if (t.size() < 4 || t[1] != "discrete")
                    throw Error("LoadBIFFile: contain continuous variable in BN");
                const int k = std::stoi(t[3]);
                node_cards.back() = k;
                for (int i = 6; i < 6 + k; ++i) {
                    if (i >= (int)t.size()) throw Error("LoadBIFFile: short value list");
                    possible_values_ids.back()[trim_right_comma(t[i])] = i - 6;
                }
            }
            continue;
        }
        if (in_variables) {
            in_variables = false;
            set_parent_indexes.assign(num_nodes, {});
            set_children_indexes.assign(num_nodes, {});
        }
        if (t[0] != "probability") continue;
        if (t.size() < 4 || t[1] != "(") throw Error("LoadBIFFile: error in format");
        auto node_of = [&](const std::string &name) {
            auto it = index_of.find(name);
            if (it == index_of.end()) throw Error("LoadBIFFile: unknown variable " + name);
            return it->second;
        };
        cur = node_of(t[2]);
        parents.clear();
        if (t[3] != ")") {
            for (size_t pos = 4; pos < t.size() && t[pos] != ")"; ++pos) parents.push_back(node_of(trim_right_comma(t[pos])));
            for (int p : parents) {
                SetParentChild(p, cur);
                vec_edges.emplace_back(p, cur, TAIL, ARROW);
                ++num_edges;
            }
        }
    }
    if (in_variables) {
        set_parent_indexes.assign(num_nodes, {});
        set_children_indexes.assign(num_nodes, {});
    }
    topo_ord.clear();
}

int BNSLComparison::GetSHD()
{
    if (!learned_graph->IsDAG()) {
        // compare against the true DAG's CPDAG: reversible arcs become undirected in place
        std::vector<Edge> order = true_graph->OrderEdge();
        true_graph->FindCompelled(order);
        for (int i = 0; i < true_graph->num_edges; ++i) {
            Edge &e = true_graph->vec_edges.at(i);
            if (e.label == REVERSIBLE) e = Edge(e.node1, e.node2);
        }
    }
    int error = 0;
    for (int i1 = 0; i1 < learned_graph->num_nodes; ++i1)
        for (int i2 = i1 + 1; i2 < learned_graph->num_nodes; ++i2) error += GetSHDOneEdge(i1, i2);
    return error;
}

// 0 when both graphs lack the pair or hold the same edge (same node order and the same
// second end point), 1 otherwise (BNSLComparison.cpp:44-73).
int BNSLComparison::GetSHDOneEdge(int index1, int index2)
{
    const int pos1 = true_graph->GetEdge(index1, index2);
    const int pos2 = learned_graph->GetEdge(index1, index2);
    if (pos1 == -1 && pos2 == -1) return 0;
    if (pos1 == -1 || pos2 == -1) return 1;
    const Edge &e1 = true_graph->vec_edges.at(pos1);
    const Edge &e2 = learned_graph->vec_edges.at(pos2);
    return (e1.node1 == e2.node1 && e1.node2 == e2.node2 && e1.end_point2 == e2.end_point2) ? 0 : 1;
}

}  // namespace fpgm
