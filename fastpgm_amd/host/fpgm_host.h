// C++ host side of the PC-stable drop-in: the reference's Dataset / Network / PCStable
// interface for structure learning with PC-stable, with step 1 (the CI stage) on the
// MI355X through the C ABI of include/fpgm_ci.h and the rest (orientation, CPDAG/DAG
// output) on the host.
//
// Names, argument order and meaning follow the reference:
//   PCStable(Network *net, double alpha, int depth = 1000)           PCStable.h:38
//   StructLearnCompData(dts, group_size, num_threads, dag, add_root,
//                       save_struct, struct_file, verbose)            PCStable.cpp:25-82
//   Dataset::LoadCSVTrainingData / LoadLIBSVMTrainingData            Dataset.cpp:389-503, 55-162
//   Network edge operations, Edge                                     Network.cpp:169-360, Edge.cpp:7-33
// Where the reference calls exit(1) (unreadable file, duplicate header names, alpha
// outside [0, 1], a random DAG completion that closes a cycle), these classes throw
// fpgm::Error under the same conditions.
#pragma once

#include <stdint.h>

#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "fpgm_ci.h"

namespace fpgm {

struct Error : std::runtime_error {
    explicit Error(const std::string &m) : std::runtime_error(m) {}
};

// Discrete dataset (Dataset.h:20-60, the members structure learning reads).  Values are
// stored as uint8 columns, column v at columns.data() + v * stride (stride = rows rounded
// up to 256), the layout the device path reads; values and cardinalities must therefore
// stay below 256.
class Dataset {
public:
    int num_instance = 0;
    int num_vars = 0;
    int class_var_index = 0;
    std::vector<std::string> vec_var_names;
    std::vector<int> num_of_possible_values_of_disc_vars;
    std::vector<std::map<std::string, int>> vars_possible_values_ids;
    std::vector<uint8_t> columns;
    int64_t stride = 0;

    // CSV with a header line of variable names (Dataset.cpp:389-503).  str_val: values are
    // strings coded by order of first appearance per column (0, 1, ...); otherwise integers
    // taken as they are.
    void LoadCSVTrainingData(const std::string &data_file_path, bool str_val = true, int cls_var_id = 0);
    // LIBSVM: "label index:value ..." lines; absent features read 0, cardinality =
    // largest value + 1, variable names "0", "1", ... (Dataset.cpp:55-162).
    void LoadLIBSVMTrainingData(const std::string &data_file_path, int cls_var_id = 0);

    int value(int var, int64_t row) const { return columns[(size_t)var * stride + row]; }
    // builds the column buffer from a row-major table (num_instance x num_vars)
    void SetRows(const std::vector<int> &rows, int n_rows, int n_vars);
};

enum EndPoint { TAIL, ARROW };

// One edge (Edge.h:15-45): undirected edges are stored with node1 < node2, directed
// edges p -> c as (p, c, TAIL, ARROW).
struct Edge {
    int node1 = 0, node2 = 0;
    EndPoint end_point1 = TAIL, end_point2 = TAIL;
    Edge() = default;
    Edge(int a, int b);                            // undirected
    Edge(int a, int b, EndPoint e1, EndPoint e2);  // as given
    bool IsDirected() const;
    bool operator==(const Edge &o) const
    {
        return node1 == o.node1 && node2 == o.node2 && end_point1 == o.end_point1 && end_point2 == o.end_point2;
    }
};

// Skeleton adjacency as structure learning leaves it after step 1: per node, its
// neighbours in ascending order (the reference's map<int, map<int, double>>).
using Adjacencies = std::vector<std::vector<int>>;

// The graph part of the reference Network (Network.h) that PC-stable uses.
class Network {
public:
    int num_nodes = 0;
    int num_edges = 0;
    bool pure_discrete = true;
    std::vector<std::string> node_names;
    std::vector<int> node_cards;
    std::vector<Edge> vec_edges;
    std::vector<std::set<int>> set_parent_indexes, set_children_indexes;

    explicit Network(bool pure_disc = true) : pure_discrete(pure_disc) {}

    void GenerateUndirectedCompleteGraph();
    // adds p -> c unless it closes a directed cycle; false (and no edge) when it would
    bool AddDirectedEdge(int p_index, int c_index);
    bool DeleteDirectedEdge(int p_index, int c_index);
    void AddUndirectedEdge(int p_index, int c_index);
    bool DeleteUndirectedEdge(int p_index, int c_index);
    static bool IsAdjacentTo(const Adjacencies &adj, int a, int b);
    bool IsDirectedFromTo(int a, int b) const;
    bool IsUndirected(const Adjacencies &adj, int a, int b) const;
    std::set<int> GetParentIdxesOfNode(int v) const { return set_parent_indexes[v]; }
    std::set<int> GetChildrenIdxesOfNode(int v) const { return set_children_indexes[v]; }
    bool ContainCircle() const;
    std::vector<int> GetTopoOrd();

private:
    std::vector<int> topo_ord;
    void SetParentChild(int p, int c);
    void RemoveParentChild(int p, int c);
};

// PC-stable structure learning (PCStable.h:24-62).
class PCStable {
public:
    int depth = 1000;
    int num_ci_test = 0;
    double alpha;
    // pairs (min, max) of non-adjacent nodes -> the conditioning set that separated them
    std::map<std::pair<int, int>, std::set<int>> sepset;
    Network *network;
    Adjacencies adjacencies;      // skeleton after step 1
    fpgm_stats last_stats{};      // device statistics of the last step 1
    double step1_seconds = 0.0;   // "pc-stable step 1" (PCStable.cpp:109-232)
    double total_seconds = 0.0;   // "pc-stable"

    PCStable(Network *net, double a, int d = 1000);

    virtual ~PCStable() = default;
    virtual void StructLearnCompData(Dataset *dts, int group_size, int num_threads, bool dag, bool add_root,
                                     bool save_struct, const std::string &struct_file, int verbose);
    void StructLearnByPCStable(Dataset *dts, int num_threads, int group_size, bool dag, int verbose);

    // step 1 on the device: skeleton, sepset, num_ci_test, vec_edges in lexicographic order
    void LearnSkeleton(Dataset *dts, int num_threads, int group_size, int verbose);
    // the same state from a skeleton computed elsewhere (adjacency n x n, sepsets)
    void SetSkeleton(const std::vector<uint8_t> &adj, const std::map<std::pair<int, int>, std::set<int>> &seps);
    // steps after the CI stage (PCStable.cpp:234-255)
    void Orient(bool dag);

    void AssignNodeInformation(const Dataset *dts);
    void OrientVStructure(const Adjacencies &adj);
    void OrientImplied(const Adjacencies &adj);
    bool Rule1(const Adjacencies &adj, int b_idx, int c_idx);
    bool Rule2(const Adjacencies &adj, int a_idx, int c_idx);
    bool Rule3(const Adjacencies &adj, int d_idx, int a_idx);
    bool R3Helper(const Adjacencies &adj, int a_idx, int d_idx, int b_idx, int c_idx);
    bool Direct(int a, int c);
    std::vector<int> GetCommonAdjacents(const Adjacencies &adj, int x_idx, int y_idx) const;
    void DirectLeftEdges(const Adjacencies &adj);
    std::vector<int> FindRootsInDAGForest();
    void AddRootNode(const std::vector<int> &sub_roots);
    void SaveBNStructure(const std::string &struct_file) const;
    std::string FormatBNStructure() const;
};

}  // namespace fpgm
