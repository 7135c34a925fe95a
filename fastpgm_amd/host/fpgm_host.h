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
//   Network::OrderEdge / FindCompelled / IsDAG / GetEdge               Network.cpp:95-129, 547-699
//   CustomNetwork::LoadBIFFile (structure and domains)                 CustomNetwork.cpp:52-230
//   BNSLComparison(true_graph, learned_graph).GetSHD()                BNSLComparison.cpp:7-73
//   ParameterLearning(net).LearnParamsKnowStructCompData(dts, alpha, save_param,
//                     param_file, verbose), SaveBNParameter           ParameterLearning.cpp:10-155
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
// Chickering's edge labels (Edge.h:13-31); an edge FindCompelled never labels keeps COMPELLED
// (the reference leaves the field uninitialised)
enum Label { COMPELLED, REVERSIBLE };

// One edge (Edge.h:15-45): undirected edges are stored with node1 < node2, directed
// edges p -> c as (p, c, TAIL, ARROW).
struct Edge {
    int node1 = 0, node2 = 0;
    EndPoint end_point1 = TAIL, end_point2 = TAIL;
    bool is_ordered = false;
    Label label = COMPELLED;
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

// Conditional probability table of one node (PotentialTableBase.h): variables = the node,
// then its parents ascending; entry index = sum config[k] * cum_levels[k].
struct PotentialTable {
    int num_variables = 0;
    std::vector<int> vec_related_variables, var_dims, cum_levels;
    int table_size = 0;
    std::vector<double> potentials;
};

// The graph part of the reference Network (Network.h) that PC-stable, SHD and parameter
// learning use.
class Network {
public:
    int num_nodes = 0;
    int num_edges = 0;
    bool pure_discrete = true;
    std::vector<std::string> node_names;
    std::vector<int> node_cards;
    std::vector<std::map<std::string, int>> possible_values_ids;  // value name -> state, per node
    std::vector<Edge> vec_edges;
    std::vector<std::set<int>> set_parent_indexes, set_children_indexes;
    std::vector<PotentialTable> pts;  // one per node after parameter learning

    explicit Network(bool pure_disc = true) : pure_discrete(pure_disc) {}
    virtual ~Network() = default;

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

    // position in vec_edges of the first edge joining a and b in either direction, or -1
    int GetEdge(int a, int b) const;
    // position of the directed edge a -> b, or -1
    int GetDirectedEdge(int a, int b) const;
    // edges of a DAG in Chickering's order (Network.cpp:547-584)
    std::vector<Edge> OrderEdge();
    // labels every edge COMPELLED or REVERSIBLE (Network.cpp:590-684)
    void FindCompelled(std::vector<Edge> &edge_order);
    // true when the first stored edge is directed (Network.cpp:692-698)
    bool IsDAG() const;
    std::string GetValueNameByIndex(int node, int index) const;

protected:
    std::vector<int> topo_ord;
    void SetParentChild(int p, int c);
    void RemoveParentChild(int p, int c);
};

// A network read from a file (CustomNetwork.h).  LoadBIFFile keeps what structure
// comparison and parameter learning read: nodes in declaration order with their names,
// domain sizes and value names, and one arc per parent in each probability block's order.
class CustomNetwork : public Network {
public:
    CustomNetwork() : Network(true) {}
    void LoadBIFFile(const std::string &path);
};

// Structural Hamming distance between a true DAG and a learned graph (BNSLComparison.h).
// A learned CPDAG is compared with the true DAG's CPDAG: the true graph's REVERSIBLE
// edges are replaced by undirected ones, in place, as the reference does.
class BNSLComparison {
public:
    Network *true_graph;
    Network *learned_graph;
    BNSLComparison(Network *t, Network *l) : true_graph(t), learned_graph(l) {}
    int GetSHD();
    int GetSHDOneEdge(int index1, int index2);
};

// Maximum-likelihood parameters of a known structure (ParameterLearning.h); the counts
// N(AB) of every node come from the device (fpgm_mle_counts).
class ParameterLearning {
public:
    Network *network;
    double seconds = 0.0;  // "parameter_learning" timer
    explicit ParameterLearning(Network *net) : network(net) {}
    void LearnParamsKnowStructCompData(const Dataset *dts, int alpha, bool save_param, const std::string &param_file,
                                       int verbose);
    void SaveBNParameter(const std::string &param_file) const;
    std::string FormatBNParameter() const;
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
