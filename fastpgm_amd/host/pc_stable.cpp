// PC-stable on the host mirror: step 1 (the CI stage) through the C ABI on the device,
// then v-structures, Meek's rules and the structure file on the host, in the
// reference's order (PCStable.cpp:25-255, 626-847; StructureLearning.cpp:11-34, 60-292).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <queue>
#include <random>
#include <sstream>

#include "fpgm_host.h"

namespace fpgm {

namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void check(fpgm_ctx *ctx, fpgm_status s, const char *what)
{
    if (s != FPGM_OK) {
        std::string msg = std::string(what) + ": " + fpgm_status_str(s);
        if (ctx) msg += std::string(" (") + fpgm_last_error(ctx) + ")";
        throw Error(msg);
    }
}

struct CtxGuard {
    fpgm_ctx *c = nullptr;
    ~CtxGuard() { fpgm_ctx_destroy(c); }
};

}  // namespace

PCStable::PCStable(Network *net, double a, int d) : depth(d), alpha(a), network(net)
{
    if (!(a >= 0.0 && a <= 1.0)) throw Error("alpha must lie in [0, 1]");  // IndependenceTest.cpp:12-21
}

void PCStable::AssignNodeInformation(const Dataset *dts)
{
    // one discrete node per variable, named and sized from the dataset (StructureLearning.cpp:11-34)
    network->num_nodes = dts->num_vars;
    network->node_names = dts->vec_var_names;
    network->node_cards = dts->num_of_possible_values_of_disc_vars;
    network->possible_values_ids = dts->vars_possible_values_ids;
    network->possible_values_ids.resize(dts->num_vars);
    network->pts.clear();
    network->vec_edges.clear();
    network->num_edges = 0;
    network->set_parent_indexes.assign(dts->num_vars, {});
    network->set_children_indexes.assign(dts->num_vars, {});
}

void PCStable::StructLearnCompData(Dataset *dts, int group_size, int num_threads, bool dag, bool add_root,
                                   bool save_struct, const std::string &struct_file, int verbose)
{
    if (verbose > 0) std::cout << "--------------------------------------------------\n"
                               << "Structure learning with PC-Stable..." << std::endl;
    const double t0 = now_s();
    AssignNodeInformation(dts);
    StructLearnByPCStable(dts, num_threads, group_size, dag, verbose);
    total_seconds = now_s() - t0;
    if (add_root) {
        const std::vector<int> roots = FindRootsInDAGForest();
        AddRootNode(roots);
        if (verbose > 0) {
            std::cout << "roots: ";
            for (int r : roots) std::cout << r << ", ";
            std::cout << std::endl;
        }
    }
    if (save_struct) SaveBNStructure(struct_file);
    if (verbose > 0) {
        std::cout << "--------------------------------------------------\n"
                  << "# of CI-tests = " << num_ci_test << "\n"
                  << "pc-stable step 1: " << step1_seconds << " s\n"
                  << "pc-stable: " << total_seconds << " s" << std::endl;
    }
}

void PCStable::StructLearnByPCStable(Dataset *dts, int num_threads, int group_size, bool dag, int verbose)
{
    LearnSkeleton(dts, num_threads, group_size, verbose);
    Orient(dag);
}

void PCStable::LearnSkeleton(Dataset *dts, int num_threads, int group_size, int verbose)
{
    const int n = dts->num_vars;
    if (n < 2) throw Error("structure learning needs at least two variables");
    for (int v = 0; v < n; ++v)
        if (dts->num_of_possible_values_of_disc_vars[v] < 1 || dts->num_of_possible_values_of_disc_vars[v] > 255)
            throw Error("variable " + std::to_string(v) + " has a cardinality outside 1..255");
    std::vector<int32_t> card(dts->num_of_possible_values_of_disc_vars.begin(),
                              dts->num_of_possible_values_of_disc_vars.end());
    CtxGuard g;
    check(nullptr,
          fpgm_ctx_create(dts->columns.data(), n, dts->num_instance, dts->stride, card.data(), &g.c),
          "fpgm_ctx_create");
    fpgm_pc_opts opts{};
    opts.mode = FPGM_MODE_BATCHED;
    opts.n_threads = num_threads;
    opts.group_size = group_size;
    opts.verbose = verbose;
    std::vector<uint8_t> adj((size_t)n * n);
    fpgm_stats st{};
    check(g.c, fpgm_pc_skeleton(g.c, alpha, depth, &opts, adj.data(), &st), "fpgm_pc_skeleton");
    int64_t nrec = 0, nz = 0;
    check(g.c, fpgm_sepset_count(g.c, &nrec, &nz), "fpgm_sepset_count");
    std::vector<fpgm_sepset> rec((size_t)nrec);
    std::vector<int32_t> z((size_t)std::max<int64_t>(nz, 1));
    check(g.c, fpgm_sepset_get(g.c, rec.data(), nrec, z.data(), nz), "fpgm_sepset_get");
    std::map<std::pair<int, int>, std::set<int>> seps;
    for (const fpgm_sepset &r : rec)
        seps[{r.x, r.y}] = std::set<int>(z.begin() + r.z_off, z.begin() + r.z_off + r.d);
    last_stats = st;
    step1_seconds = st.ci_seconds;
    SetSkeleton(adj, seps);
    num_ci_test = (int)st.num_ci_test;
}

void PCStable::SetSkeleton(const std::vector<uint8_t> &adj, const std::map<std::pair<int, int>, std::set<int>> &seps)
{
    const int n = network->num_nodes;
    if ((int64_t)adj.size() != (int64_t)n * n) throw Error("adjacency size does not match the network");
    // surviving edges keep the complete graph's lexicographic order: the reference erases
    // removed edges in place (PCStable.cpp:179-195, 352-379)
    network->vec_edges.clear();
    adjacencies.assign(n, {});
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            if (i != j && adj[(size_t)i * n + j]) adjacencies[i].push_back(j);
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            if (adj[(size_t)i * n + j]) network->vec_edges.emplace_back(i, j);
    network->num_edges = (int)network->vec_edges.size();
    network->set_parent_indexes.assign(n, {});
    network->set_children_indexes.assign(n, {});
    sepset = seps;
}

void PCStable::Orient(bool dag)
{
    OrientVStructure(adjacencies);
    OrientImplied(adjacencies);
    if (dag) DirectLeftEdges(adjacencies);
}

// For every b and every pair a < c of its neighbours with a, c not adjacent and b not in
// sepset(a, c): a -> b <- c.  Conflicting orientations are overwritten in this visiting
// order, and an arc that would close a cycle is not added: then what was deleted for this
// triple is put back (PCStable.cpp:639-735).
void PCStable::OrientVStructure(const Adjacencies &adj)
{
    Network &N = *network;
    for (int b = 0; b < N.num_nodes; ++b) {
        const std::vector<int> &nb = adj[b];
        if (nb.size() < 2) continue;
        for (size_t i = 0; i + 1 < nb.size(); ++i)
            for (size_t j = i + 1; j < nb.size(); ++j) {  // 2-subsets in lexicographic order
                const int a = nb[i], c = nb[j];
                if (Network::IsAdjacentTo(adj, a, c)) continue;
                const std::set<int> &s = sepset[std::make_pair(a, c)];
                if (s.count(b)) continue;
                bool del_dir1 = N.DeleteDirectedEdge(b, a), del_und1 = false;
                if (!del_dir1) del_und1 = N.DeleteUndirectedEdge(a, b);
                const bool to_add1 = del_dir1 || del_und1;
                bool del_dir2 = N.DeleteDirectedEdge(b, c), del_und2 = false;
                if (!del_dir2) del_und2 = N.DeleteUndirectedEdge(c, b);
                const bool to_add2 = del_dir2 || del_und2;
                bool added1 = false, added2 = false;
                if (to_add1) added1 = N.AddDirectedEdge(a, b);
                if (to_add2) added2 = N.AddDirectedEdge(c, b);
                if ((to_add1 && !added1) || (to_add2 && !added2)) {
                    if (to_add1) {
                        if (del_dir1) N.AddDirectedEdge(b, a);
                        else N.AddUndirectedEdge(a, b);
                    }
                    if (to_add2) {
                        if (del_dir2) N.AddDirectedEdge(b, c);
                        else N.AddUndirectedEdge(c, b);
                    }
                }
            }
    }
}

// Meek's rules over the undirected edges, front to back, until a pass orients nothing.
// A directed edge ends the pass: orienting moves an edge to the back of vec_edges, so the
// undirected edges come first (PCStable.cpp:746-780).
void PCStable::OrientImplied(const Adjacencies &adj)
{
    Network &N = *network;
    bool oriented = true;
    while (oriented) {
        oriented = false;
        for (size_t i = 0; i < N.vec_edges.size();) {
            const int x = N.vec_edges[i].node1, y = N.vec_edges[i].node2;
            if (!N.IsUndirected(adj, x, y)) break;
            if (Rule1(adj, x, y) || Rule1(adj, y, x) || Rule2(adj, x, y) || Rule2(adj, y, x) || Rule3(adj, x, y) ||
                Rule3(adj, y, x))
                oriented = true;  // the edge left position i; its successor moved in
            else
                ++i;
        }
    }
}

// R1: a -> b, b -- c, a not adjacent to c  =>  b -> c (PCStable.cpp:785-794)
bool PCStable::Rule1(const Adjacencies &adj, int b_idx, int c_idx)
{
    for (int a_idx : network->GetParentIdxesOfNode(b_idx)) {
        if (Network::IsAdjacentTo(adj, c_idx, a_idx)) continue;
        if (Direct(b_idx, c_idx)) return true;
    }
    return false;
}

// R2: a -> b -> c, a -- c  =>  a -> c (PCStable.cpp:799-813)
bool PCStable::Rule2(const Adjacencies &adj, int a_idx, int c_idx)
{
    for (int b_idx : GetCommonAdjacents(adj, a_idx, c_idx))
        if (network->IsDirectedFromTo(a_idx, b_idx) && network->IsDirectedFromTo(b_idx, c_idx))
            if (Direct(a_idx, c_idx)) return true;
    return false;
}

// R3: d -- a, d -- b, d -- c, b -> a, c -> a, b and c not adjacent  =>  d -> a
// (PCStable.cpp:818-847)
bool PCStable::Rule3(const Adjacencies &adj, int d_idx, int a_idx)
{
    const std::vector<int> common = GetCommonAdjacents(adj, a_idx, d_idx);
    if (common.size() < 2) return false;
    for (size_t i = 0; i < common.size(); ++i)
        for (size_t j = i + 1; j < common.size(); ++j) {
            const int b_idx = common[i], c_idx = common[j];
            if (!Network::IsAdjacentTo(adj, b_idx, c_idx))
                if (R3Helper(adj, a_idx, d_idx, b_idx, c_idx)) return true;
        }
    return false;
}

bool PCStable::R3Helper(const Adjacencies &adj, int a_idx, int d_idx, int b_idx, int c_idx)
{
    const Network &N = *network;
    if (N.IsUndirected(adj, d_idx, b_idx) && N.IsUndirected(adj, d_idx, c_idx) && N.IsDirectedFromTo(b_idx, a_idx) &&
        N.IsDirectedFromTo(c_idx, a_idx))
        return Direct(d_idx, a_idx);
    return false;
}

// a -- c becomes a -> c unless that closes a cycle (StructureLearning.cpp:72-83)
bool PCStable::Direct(int a, int c)
{
    const bool added = network->AddDirectedEdge(a, c);
    if (added) network->DeleteUndirectedEdge(a, c);
    return added;
}

std::vector<int> PCStable::GetCommonAdjacents(const Adjacencies &adj, int x_idx, int y_idx) const
{
    std::vector<int> out;
    std::set_intersection(adj[x_idx].begin(), adj[x_idx].end(), adj[y_idx].begin(), adj[y_idx].end(),
                          std::back_inserter(out));
    return out;
}

// CPDAG -> DAG: each remaining undirected edge gets a random direction, the other one when
// the first closes a cycle (StructureLearning.cpp:109-145).
void PCStable::DirectLeftEdges(const Adjacencies &adj)
{
    Network &N = *network;
    std::random_device rd;
    std::mt19937 gen(rd());
    std::uniform_int_distribution<int> coin(0, 1);
    for (size_t i = 0; i < N.vec_edges.size();) {
        const int x = N.vec_edges[i].node1, y = N.vec_edges[i].node2;
        if (!N.IsUndirected(adj, x, y)) break;
        bool ok;
        if (coin(gen) == 0) ok = Direct(x, y) || Direct(y, x);
        else ok = Direct(y, x) || Direct(x, y);
        if (!ok) throw Error("directing an undirected edge closes a cycle either way");
    }
}

// One root per weakly connected part of the DAG: repeatedly the unvisited node of
// in-degree 0 with most children, then a breadth-first sweep over children and parents
// (StructureLearning.cpp:152-208).
std::vector<int> PCStable::FindRootsInDAGForest()
{
    const Network &N = *network;
    std::vector<int> roots;
    std::vector<bool> visited(N.num_nodes, false);
    while (true) {
        int best = -1, root = -1;
        for (int i = 0; i < N.num_nodes; ++i)
            if (N.set_parent_indexes[i].empty() && !visited[i] && (int)N.set_children_indexes[i].size() > best) {
                best = (int)N.set_children_indexes[i].size();
                root = i;
            }
        if (best == -1) break;
        visited[root] = true;
        roots.push_back(root);
        std::queue<int> q;
        q.push(root);
        while (!q.empty()) {
            const int u = q.front();
            q.pop();
            for (int c : N.set_children_indexes[u])
                if (!visited[c]) {
                    visited[c] = true;
                    q.push(c);
                }
            for (int p : N.set_parent_indexes[u])
                if (!visited[p]) {
                    visited[p] = true;
                    q.push(p);
                }
        }
    }
    return roots;
}

// A binary node "ROOT" (index num_nodes) becomes the parent of every sub-graph root
// (StructureLearning.cpp:219-237).
void PCStable::AddRootNode(const std::vector<int> &sub_roots)
{
    Network &N = *network;
    const int r = N.num_nodes;
    N.set_parent_indexes.emplace_back();
    N.set_children_indexes.emplace_back();
    N.node_names.push_back("ROOT");
    N.node_cards.push_back(2);
    N.possible_values_ids.resize(r);
    N.possible_values_ids.push_back({{"0", 0}, {"1", 1}});
    for (int c : sub_roots) {
        N.set_children_indexes[r].insert(c);
        N.set_parent_indexes[c].insert(r);
        N.vec_edges.emplace_back(r, c, TAIL, ARROW);
        N.num_edges++;
    }
    N.num_nodes++;
}

// The structure file (StructureLearning.cpp:239-292): a header, then the CPDAG's edges in
// vec_edges order, or, when the first edge is directed, the DAG's arcs in topological
// order of their parents.
std::string PCStable::FormatBNStructure() const
{
    Network &N = *network;
    std::ostringstream out;
    out << "number of nodes = " << N.num_nodes << ", number of edges = " << N.num_edges << "\n";
    auto name = [&](int v) { return N.node_names[v] + "/v" + std::to_string(v); };
    if (!N.vec_edges.empty() && N.vec_edges[0].IsDirected()) {
        const std::vector<int> topo = N.GetTopoOrd();
        int k = 0;
        for (int i = 0; i < N.num_nodes && i < (int)topo.size(); ++i)
            for (int c : N.set_children_indexes[topo[i]])
                out << k++ << ": " << name(topo[i]) << " -> " << name(c) << "\n";
    } else {
        for (int i = 0; i < N.num_edges; ++i) {
            if (i >= (int)N.vec_edges.size()) throw Error("edge count exceeds the stored edges");
            const Edge &e = N.vec_edges[i];
            out << i << ": ";
            if (!e.IsDirected()) out << name(e.node1) << " -- " << name(e.node2);
            else if (e.end_point1 == TAIL) out << name(e.node1) << " -> " << name(e.node2);
            else out << name(e.node2) << " -> " << name(e.node1);
            out << "\n";
        }
    }
    return out.str();
}

void PCStable::SaveBNStructure(const std::string &struct_file) const
{
    const std::string text = FormatBNStructure();
    std::ofstream f(struct_file);
    if (!f.is_open()) throw Error("unable to open file " + struct_file);
    f << text;
}

}  // namespace fpgm
