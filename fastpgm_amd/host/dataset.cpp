// Dataset loaders of the host mirror (Dataset.cpp:55-162 LIBSVM, :389-503 CSV), storing
// uint8 columns with a 256-byte padded stride.
#include <algorithm>
#include <fstream>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "fpgm_host.h"

namespace fpgm {

namespace {

// trailing characters below '!' (whitespace, \r) are dropped (common.cpp TrimRight)
std::string trim_right(std::string s)
{
    while (!s.empty() && (unsigned char)s.back() < 33) s.pop_back();
    return s;
}

// every occurrence of the delimiter splits, empty fields included (common.cpp:132-141)
std::vector<std::string> split(const std::string &s, char delim)
{
    std::vector<std::string> out;
    size_t begin = 0, end;
    while ((end = s.find(delim, begin)) != std::string::npos) {
        out.push_back(s.substr(begin, end - begin));
        begin = end + 1;
    }
    out.push_back(s.substr(begin));
    return out;
}

int to_int(const std::string &s, const std::string &what)
{
    try {
        size_t used = 0;
        const int v = std::stoi(s, &used);
        return v;
    } catch (const std::exception &) {
        throw Error("not an integer " + what + ": '" + s + "'");
    }
}

}  // namespace

void Dataset::SetRows(const std::vector<int> &rows, int n_rows, int n_vars)
{
    num_instance = n_rows;
    num_vars = n_vars;
    stride = std::max<int64_t>(256, ((int64_t)n_rows + 255) / 256 * 256);
    columns.assign((size_t)n_vars * stride, 0);
    for (int r = 0; r < n_rows; ++r)
        for (int v = 0; v < n_vars; ++v) {
            const int x = rows[(size_t)r * n_vars + v];
            if (x < 0 || x > 255) throw Error("value outside 0..255 in variable " + std::to_string(v));
            columns[(size_t)v * stride + r] = (uint8_t)x;
        }
}

void Dataset::LoadCSVTrainingData(const std::string &data_file_path, bool str_val, int cls_var_id)
{
    std::ifstream in(data_file_path);
    if (!in.is_open()) throw Error("unable to open file " + data_file_path);
    class_var_index = cls_var_id;
    std::string line;
    std::getline(in, line);
    // the header line gives the variable count and names
    const int n_vars = (int)std::count(line.begin(), line.end(), ',') + 1;
    vec_var_names = split(trim_right(line), ',');
    if ((int)vec_var_names.size() != n_vars) throw Error("malformed header");
    if (std::set<std::string>(vec_var_names.begin(), vec_var_names.end()).size() != vec_var_names.size())
        throw Error("duplicate variable names in header");
    vars_possible_values_ids.assign(n_vars, {});
    std::vector<int> counter(n_vars, -1);
    std::vector<int> rows;
    int n_rows = 0;
    // the reference reads the line after the header unconditionally, then every further
    // line (Dataset.cpp:426-479): a file ending in a newline adds no empty row because
    // getline fails at the end
    bool have = (bool)std::getline(in, line);
    while (have) {
        const std::vector<std::string> f = split(trim_right(line), ',');
        if ((int)f.size() < n_vars) throw Error("row " + std::to_string(n_rows + 1) + " has too few values");
        for (int i = 0; i < n_vars; ++i) {
            int value;
            if (str_val) {
                auto it = vars_possible_values_ids[i].find(f[i]);
                if (it == vars_possible_values_ids[i].end()) {
                    value = ++counter[i];
                    vars_possible_values_ids[i].emplace(f[i], value);
                } else {
                    value = it->second;
                }
            } else {
                value = to_int(f[i], "in column " + vec_var_names[i]);
            }
            rows.push_back(value);
        }
        ++n_rows;
        have = (bool)std::getline(in, line);
    }
    SetRows(rows, n_rows, n_vars);
    // cardinality = number of distinct values seen (Dataset.cpp:484-488); with integer
    // values (str_val false) the reference leaves no value ids, so every cardinality is
    // the largest value + 1 here instead, which the device path needs
    num_of_possible_values_of_disc_vars.assign(n_vars, 0);
    for (int i = 0; i < n_vars; ++i) {
        if (str_val) {
            num_of_possible_values_of_disc_vars[i] = (int)vars_possible_values_ids[i].size();
        } else {
            int mx = -1;
            for (int r = 0; r < n_rows; ++r) mx = std::max(mx, value(i, r));
            num_of_possible_values_of_disc_vars[i] = mx + 1;
        }
    }
}

void Dataset::LoadLIBSVMTrainingData(const std::string &data_file_path, int cls_var_id)
{
    std::ifstream in(data_file_path);
    if (!in.is_open()) throw Error("unable to open file " + data_file_path);
    class_var_index = cls_var_id;
    std::vector<std::vector<std::pair<int, int>>> inst;
    int max_index = -1;
    std::string line;
    while (std::getline(in, line)) {
        const std::vector<std::string> f = split(trim_right(line), ' ');
        std::vector<std::pair<int, int>> one;
        one.emplace_back(class_var_index, to_int(f[0], "label"));
        for (size_t k = 1; k < f.size(); ++k) {
            const std::vector<std::string> iv = split(f[k], ':');
            if (iv.size() < 2) throw Error("malformed feature '" + f[k] + "'");
            const int idx = to_int(iv[0], "feature index");
            if (idx < 0) throw Error("negative feature index");
            max_index = std::max(max_index, idx);
            one.emplace_back(idx, to_int(iv[1], "feature value"));
        }
        inst.push_back(std::move(one));
    }
    const int n_vars = max_index + 1;
    const int n_rows = (int)inst.size();
    std::vector<int> rows((size_t)n_rows * std::max(n_vars, 0), 0);
    for (int r = 0; r < n_rows; ++r)
        for (const auto &iv : inst[r])
            if (iv.first < n_vars) rows[(size_t)r * n_vars + iv.first] = iv.second;
    SetRows(rows, n_rows, n_vars);
    vec_var_names.clear();
    for (int i = 0; i < n_vars; ++i) vec_var_names.push_back(std::to_string(i));
    vars_possible_values_ids.assign(n_vars, {});
    num_of_possible_values_of_disc_vars.assign(n_vars, 0);
    for (int i = 0; i < n_vars; ++i) {
        int mx = -1;
        for (int r = 0; r < n_rows; ++r) mx = std::max(mx, value(i, r));
        num_of_possible_values_of_disc_vars[i] = mx + 1;
    }
}

}  // namespace fpgm
