// Maximum-likelihood parameter learning of the host mirror (ParameterLearning.cpp:10-155),
// with the counting of every node's table N(AB) on the device (fpgm_mle_counts).
#include <chrono>
#include <fstream>
#include <iostream>
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
    if (s == FPGM_OK) return;
    std::string msg = std::string(what) + ": " + fpgm_status_str(s);
    if (ctx) msg += std::string(" (") + fpgm_last_error(ctx) + ")";
    throw Error(msg);
}

struct CtxGuard {
    fpgm_ctx *c = nullptr;
    ~CtxGuard() { fpgm_ctx_destroy(c); }
};

// The table layout of PotentialTableBase::ConstructPotentialTableWithEmptyPotentials
// (PotentialTableBase.cpp:102-130): the node, then its parents in ascending order;
// cum_levels[k] = product of the dimensions after k.
PotentialTable make_table(const Network &N, int node)
{
    PotentialTable pt;
    pt.vec_related_variables.push_back(node);
    for (int p : N.set_parent_indexes[node]) pt.vec_related_variables.push_back(p);
    pt.num_variables = (int)pt.vec_related_variables.size();
    for (int v : pt.vec_related_variables) pt.var_dims.push_back(N.node_cards[v]);
    pt.cum_levels.assign(pt.num_variables, 1);
    for (int k = pt.num_variables - 2; k >= 0; --k) pt.cum_levels[k] = pt.cum_levels[k + 1] * pt.var_dims[k + 1];
    pt.table_size = pt.cum_levels[0] * pt.var_dims[0];
    pt.potentials.assign(pt.table_size, 0.0);
    return pt;
}

}  // namespace

void ParameterLearning::LearnParamsKnowStructCompData(const Dataset *dts, int alpha, bool save_param,
                                                      const std::string &param_file, int verbose)
{
    if (verbose > 0)
        std::cout << "--------------------------------------------------\n"
                  << "Parameter learning with maximum likelihood estimation... \n"
                  << "Laplace smoothing param alpha = " << alpha << std::endl;
    const double t0 = now_s();
    Network &N = *network;
    const int nv = dts->num_vars;
    if (N.num_nodes < nv || (int)N.set_parent_indexes.size() < N.num_nodes)
        throw Error("ParameterLearning: the network has fewer nodes than the dataset has variables");
    // a virtual ROOT node follows the variables when the network has one more node
    const int root_idx = (N.num_nodes == nv) ? -1 : nv;
    N.pts.clear();
    N.pts.reserve(N.num_nodes);

    // device counts N(AB) of every node whose parents are variables of the dataset
    std::vector<int32_t> par_off(nv + 1, 0), parents;
    std::vector<bool> root_child(nv, false);
    for (int i = 0; i < nv; ++i) {
        const std::set<int> &ps = N.set_parent_indexes[i];
        if (!ps.empty() && *ps.begin() == root_idx) {
            root_child[i] = true;  // ROOT is then its only parent (AddRootNode)
        } else {
            for (int p : ps) {
                if (p >= nv) throw Error("ParameterLearning: a parent is not a dataset variable");
                parents.push_back(p);
            }
        }
        par_off[i + 1] = (int32_t)parents.size();
    }
    std::vector<int32_t> card(dts->num_of_possible_values_of_disc_vars.begin(),
                              dts->num_of_possible_values_of_disc_vars.end());
    CtxGuard g;
    check(nullptr, fpgm_ctx_create(dts->columns.data(), nv, dts->num_instance, dts->stride, card.data(), &g.c),
          "fpgm_ctx_create");
    std::vector<int64_t> off(nv + 1, 0);
    check(g.c, fpgm_mle_counts(g.c, nv, par_off.data(), parents.data(), off.data(), nullptr, 0), "fpgm_mle_counts");
    std::vector<int32_t> counts((size_t)std::max<int64_t>(off[nv], 1));
    check(g.c, fpgm_mle_counts(g.c, nv, par_off.data(), parents.data(), off.data(), counts.data(), off[nv]),
          "fpgm_mle_counts");

    std::mt19937 gen{std::random_device{}()};
    std::uniform_int_distribution<int> coin(0, 1);
    for (int i = 0; i < nv; ++i) {
        PotentialTable pt = make_table(N, i);
        std::vector<int> nab(pt.table_size, 0);           // N(AB)
        std::vector<int> nb(pt.cum_levels[0], 0);         // N(B)
        if (root_child[i]) {
            // ROOT has no column: the reference draws 0 or 1 for it per row
            // (ParameterLearning.cpp:41-55), so these counts are random there too
            for (int r = 0; r < dts->num_instance; ++r) {
                const int idx = dts->value(i, r) * pt.cum_levels[0] + coin(gen);
                nab[idx]++;
                nb[idx % pt.cum_levels[0]]++;
            }
        } else {
            const int32_t *t = counts.data() + off[i];
            for (int k = 0; k < pt.table_size; ++k) {
                nab[k] = t[k];
                nb[k % pt.cum_levels[0]] += t[k];
            }
        }
        for (int k = 0; k < pt.table_size; ++k) {
            const int num2 = (pt.num_variables == 1) ? nb[0] : nb[k % pt.cum_levels[0]];
            pt.potentials[k] = ((double)nab[k] + alpha) / ((double)num2 + alpha * pt.var_dims[0]);
        }
        N.pts.push_back(pt);
    }
    if (root_idx != -1) {
        // ROOT: two equally likely values (ParameterLearning.cpp:85-101)
        PotentialTable pt = make_table(N, root_idx);
        const int num1 = (int)(0.5 * dts->num_instance);
        const int num2 = dts->num_instance;
        for (int k = 0; k < 2 && k < pt.table_size; ++k)
            pt.potentials[k] = ((double)num1 + alpha) / ((double)num2 + alpha * pt.var_dims[0]);
        N.pts.push_back(pt);
    }
    seconds = now_s() - t0;
    if (verbose > 0) std::cout << "parameter_learning: " << seconds << " s" << std::endl;
    if (save_param) SaveBNParameter(param_file);
}

// One line per table entry (ParameterLearning.cpp:111-155):
//   P (NAME/vI=VALUE/k) = p                                     (no parents)
//   P (NAME/vI=VALUE/k | PAR/vJ=VALUE/m, ...) = p               (parents)
// probabilities printed like std::to_string(double).
std::string ParameterLearning::FormatBNParameter() const
{
    const Network &N = *network;
    std::ostringstream out;
    for (int i = 0; i < N.num_nodes && i < (int)N.pts.size(); ++i) {
        const std::string head = "P (" + N.node_names[i] + "/v" + std::to_string(i) + "=";
        const PotentialTable &pt = N.pts[i];
        for (int j = 0; j < pt.table_size; ++j) {
            std::string line;
            if (pt.num_variables == 1) {
                line = head + N.GetValueNameByIndex(i, j) + "/" + std::to_string(j) + ") = " +
                       std::to_string(pt.potentials[j]);
            } else {
                std::vector<int> config(pt.num_variables);
                for (int k = 0; k < pt.num_variables; ++k) config[k] = (j / pt.cum_levels[k]) % pt.var_dims[k];
                line = head + N.GetValueNameByIndex(i, config[0]) + "/" + std::to_string(config[0]) + " | ";
                for (int k = 1; k < pt.num_variables; ++k) {
                    const int p = pt.vec_related_variables[k];
                    line += N.node_names[p] + "/v" + std::to_string(p) + "=" + N.GetValueNameByIndex(p, config[k]) +
                            "/" + std::to_string(config[k]) + ", ";
                }
                line.pop_back();
                line.pop_back();
                line += ") = " + std::to_string(pt.potentials[j]);
            }
            out << line << "\n";
        }
    }
    return out.str();
}

void ParameterLearning::SaveBNParameter(const std::string &param_file) const
{
    std::ofstream f(param_file);
    if (!f.is_open()) throw Error("SaveBNParameter: unable to open file " + param_file);
    f << FormatBNParameter();
    std::cout << "Learned parameters have been saved in " << param_file << std::endl;
}

}  // namespace fpgm
