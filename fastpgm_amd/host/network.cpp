// Graph operations of the host mirror (Network.cpp:90-130, 169-360, 371-411, 515-540;
// Edge.cpp:7-62; common.cpp:15-75).
#include <algorithm>
#include <queue>

#include "fpgm_host.h"

namespace fpgm {

Edge::Edge(int a, int b)
{
    node1 = std::min(a, b);
    node2 = std::max(a, b);
    end_point1 = TAIL;
    end_point2 = TAIL;
}

Edge::Edge(int a, int b, EndPoint e1, EndPoint e2) : node1(a), node2(b), end_point1(e1), end_point2(e2) {}

bool Edge::IsDirected() const
{
    return (end_point1 == ARROW && end_point2 == TAIL) || (end_point2 == ARROW && end_point1 == TAIL);
}

void Network::GenerateUndirectedCompleteGraph()
{
    // edge (i, j), i < j, at index (2n - i - 1) i / 2 + j - i - 1 (Network.cpp:270-283)
    vec_edges.clear();
    vec_edges.reserve((size_t)num_nodes * (num_nodes - 1) / 2);
    for (int i = 0; i < num_nodes; ++i)
        for (int j = i + 1; j < num_nodes; ++j) vec_edges.emplace_back(i, j);
    num_edges = (int)vec_edges.size();
    set_parent_indexes.assign(num_nodes, {});
    set_children_indexes.assign(num_nodes, {});
    topo_ord.clear();
}

void Network::SetParentChild(int p, int c)
{
    set_children_indexes[p].insert(c);
    set_parent_indexes[c].insert(p);
}

void Network::RemoveParentChild(int p, int c)
{
    set_children_indexes[p].erase(c);
    set_parent_indexes[c].erase(p);
}

// Kahn's algorithm over the parent/child sets: a cycle exists iff some node is never
// reached with in-degree 0 (the test of DirectedGraphContainsCircle, common.cpp:15-39).
bool Network::ContainCircle() const
{
    std::vector<int> indeg(num_nodes);
    for (int v = 0; v < num_nodes; ++v) indeg[v] = (int)set_parent_indexes[v].size();
    std::vector<int> q;
    q.reserve(num_nodes);
    for (int v = 0; v < num_nodes; ++v)
        if (indeg[v] == 0) q.push_back(v);
    for (size_t h = 0; h < q.size(); ++h)
        for (int c : set_children_indexes[q[h]])
            if (--indeg[c] == 0) q.push_back(c);
    return (int)q.size() != num_nodes;
}

bool Network::AddDirectedEdge(int p_index, int c_index)
{
    SetParentChild(p_index, c_index);
    if (ContainCircle()) {
        RemoveParentChild(p_index, c_index);
        return false;
    }
    vec_edges.emplace_back(p_index, c_index, TAIL, ARROW);
    return true;
}

bool Network::DeleteDirectedEdge(int p_index, int c_index)
{
    const Edge e(p_index, c_index, TAIL, ARROW);
    auto it = std::find(vec_edges.begin(), vec_edges.end(), e);
    if (it == vec_edges.end()) return false;
    RemoveParentChild(p_index, c_index);
    vec_edges.erase(it);
    return true;
}

void Network::AddUndirectedEdge(int p_index, int c_index)
{
    vec_edges.emplace_back(p_index, c_index);
}

bool Network::DeleteUndirectedEdge(int p_index, int c_index)
{
    const Edge e(p_index, c_index);
    auto it = std::find(vec_edges.begin(), vec_edges.end(), e);
    if (it == vec_edges.end()) return false;
    vec_edges.erase(it);
    return true;
}

bool Network::IsAdjacentTo(const Adjacencies &adj, int a, int b)
{
    return std::binary_search(adj[a].begin(), adj[a].end(), b);
}

bool Network::IsDirectedFromTo(int a, int b) const
{
    return set_parent_indexes[b].count(a) != 0;
}

bool Network::IsUndirected(const Adjacencies &adj, int a, int b) const
{
    return IsAdjacentTo(adj, a, b) && !IsDirectedFromTo(a, b) && !IsDirectedFromTo(b, a);
}

// Zero-in-degree-first order, children scanned by ascending index (common.cpp:46-72),
// computed once and kept (Network.cpp:371-376).
std::vector<int> Network::GetTopoOrd()
{
    if (!topo_ord.empty()) return topo_ord;
    std::vector<int> indeg(num_nodes);
    for (int v = 0; v < num_nodes; ++v) indeg[v] = (int)set_parent_indexes[v].size();
    std::queue<int> q;
    for (int v = 0; v < num_nodes; ++v)
        if (indeg[v] == 0) q.push(v);
    while (!q.empty()) {
        const int u = q.front();
        for (int c : set_children_indexes[u])
            if (--indeg[c] == 0) q.push(c);
        topo_ord.push_back(u);
        q.pop();
    }
    return topo_ord;
}

}  // namespace fpgm
