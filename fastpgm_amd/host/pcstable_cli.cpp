// Command-line driver of the host mirror, the flows of BNSL_PCStable and BNL_PCStable
// (fastpgm_api.cpp:18-75): load a training set, learn the structure with PC-stable
// (step 1 on the device), optionally save the structure file, compare it with a
// reference BIF network (SHD) and learn the parameters (counts on the device).
//
//   fpgm_pcstable [options] TRAIN_FILE
//     -alpha A         significance level (0.05)
//     -depth D         conditioning depth limit as in PCStable(net, alpha, D) (1000)
//     -g N, -t N       group size and thread count (accepted; results do not change)
//     -o FILE          save the structure (reference naming: TRAIN_FILE + "_struct")
//     -v N             verbosity
//     --libsvm         TRAIN_FILE is LIBSVM (default: CSV with header, string values)
//     --int-values     CSV values are integers (str_val = false)
//     --dag            complete the CPDAG to a DAG;  --add-root  add the ROOT node
//     --skeleton-in F  skip step 1: read "n", then "e x y" edges and "s x y k z..."
//                      separating sets from F (orientation and output only, no device)
//     --dump-data F    write the coded dataset: int32 n_vars, int32 n_rows, int32 cards,
//                      then the uint8 columns (n_rows bytes each), then the names, one a line
//     --load-only      load (and dump) the training set, learn nothing
//     --stats          print one JSON line: num_ci_test, edges, step-1 seconds
//     --ref-bif F      print "SHD = k" against the BIF network F (BNSLComparison)
//     --param [F]      learn the parameters (Laplace alpha 1) and save them to F
//                      (reference naming: TRAIN_FILE + "_param")
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>

#include "fpgm_host.h"

using namespace fpgm;

static void usage()
{
    std::fprintf(stderr,
                 "usage: fpgm_pcstable [-alpha A] [-depth D] [-g N] [-t N] [-o FILE] [-v N] [--libsvm] "
                 "[--int-values] [--dag] [--add-root] [--skeleton-in FILE] [--dump-data FILE] [--load-only] [--stats] "
                 "[--ref-bif FILE] [--param FILE] TRAIN_FILE\n");
}

static void read_skeleton(const std::string &path, int n_expect, std::vector<uint8_t> &adj,
                          std::map<std::pair<int, int>, std::set<int>> &seps)
{
    std::ifstream f(path);
    if (!f.is_open()) throw Error("unable to open file " + path);
    int n = 0;
    f >> n;
    if (n != n_expect) throw Error("skeleton has " + std::to_string(n) + " nodes, the dataset " +
                                   std::to_string(n_expect));
    adj.assign((size_t)n * n, 0);
    std::string tag;
    while (f >> tag) {
        int x, y;
        f >> x >> y;
        if (!f || x < 0 || y < 0 || x >= n || y >= n) throw Error("malformed skeleton line");
        if (tag == "e") {
            adj[(size_t)x * n + y] = adj[(size_t)y * n + x] = 1;
        } else if (tag == "s") {
            int k;
            f >> k;
            std::set<int> z;
            for (int i = 0; i < k; ++i) {
                int v;
                f >> v;
                z.insert(v);
            }
            seps[{std::min(x, y), std::max(x, y)}] = z;
        } else {
            throw Error("unknown skeleton line tag " + tag);
        }
    }
}

static void dump_data(const Dataset &D, const std::string &path)
{
    std::ofstream f(path, std::ios::binary);
    if (!f.is_open()) throw Error("unable to open file " + path);
    const int32_t hdr[2] = {D.num_vars, D.num_instance};
    f.write((const char *)hdr, sizeof(hdr));
    for (int v = 0; v < D.num_vars; ++v) {
        const int32_t c = D.num_of_possible_values_of_disc_vars[v];
        f.write((const char *)&c, sizeof(c));
    }
    for (int v = 0; v < D.num_vars; ++v) f.write((const char *)D.columns.data() + (size_t)v * D.stride, D.num_instance);
    for (const std::string &s : D.vec_var_names) f << s << "\n";
}

int main(int argc, char **argv)
{
    double alpha = 0.05;
    int depth = 1000, gs = 1, threads = 1, verbose = 0;
    bool libsvm = false, int_values = false, dag = false, add_root = false, stats = false, load_only = false;
    std::string out, skel, dump, train, ref_bif, param_file;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) {
                usage();
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "-alpha") alpha = std::stod(next());
        else if (a == "-depth") depth = std::stoi(next());
        else if (a == "-g") gs = std::stoi(next());
        else if (a == "-t") threads = std::stoi(next());
        else if (a == "-o") out = next();
        else if (a == "-v") verbose = std::stoi(next());
        else if (a == "--libsvm") libsvm = true;
        else if (a == "--int-values") int_values = true;
        else if (a == "--dag") dag = true;
        else if (a == "--add-root") add_root = true;
        else if (a == "--skeleton-in") skel = next();
        else if (a == "--dump-data") dump = next();
        else if (a == "--stats") stats = true;
        else if (a == "--load-only") load_only = true;
        else if (a == "--ref-bif") ref_bif = next();
        else if (a == "--param") param_file = next();
        else if (a == "-h" || a == "--help") {
            usage();
            return 0;
        } else if (!a.empty() && a[0] == '-') {
            usage();
            return 2;
        } else {
            train = a;
        }
    }
    if (train.empty()) {
        usage();
        return 2;
    }
    try {
        Dataset D;
        if (libsvm) D.LoadLIBSVMTrainingData(train, 0);
        else D.LoadCSVTrainingData(train, !int_values, 0);
        if (!dump.empty()) dump_data(D, dump);
        if (load_only) return 0;
        Network net(true);
        PCStable pc(&net, alpha, depth);
        if (skel.empty()) {
            pc.StructLearnCompData(&D, gs, threads, dag, add_root, !out.empty(), out, verbose);
        } else {
            std::vector<uint8_t> adj;
            std::map<std::pair<int, int>, std::set<int>> seps;
            pc.AssignNodeInformation(&D);
            read_skeleton(skel, D.num_vars, adj, seps);
            pc.SetSkeleton(adj, seps);
            pc.Orient(dag);
            if (add_root) pc.AddRootNode(pc.FindRootsInDAGForest());
            if (!out.empty()) pc.SaveBNStructure(out);
        }
        if (stats)
            std::printf("{\"num_ci_test\": %d, \"edges\": %d, \"step1_seconds\": %.6f, \"nodes\": %d}\n",
                        pc.num_ci_test, net.num_edges, pc.step1_seconds, net.num_nodes);
        if (!ref_bif.empty()) {
            CustomNetwork r_net;
            r_net.LoadBIFFile(ref_bif);
            BNSLComparison comp(&r_net, &net);
            std::cout << "SHD = " << comp.GetSHD() << std::endl;
        }
        if (!param_file.empty()) {
            ParameterLearning pl(&net);
            pl.LearnParamsKnowStructCompData(&D, 1, true, param_file, verbose);
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "fpgm_pcstable: %s\n", e.what());
        return 1;
    }
    return 0;
}
