// cli_common.hpp -- shared pieces of the command-line tools (bna_cormat,
// bna_degree, bna_ec): directory scans, file names, option parsing.
#pragma once

#include "bna.h"

#include <dirent.h>
#include <sys/stat.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

namespace bnacli {

// Files of `dir` whose extension is exactly `ext` (".nii", ".csr"), as the
// reference tools select them (CorMat.cpp:72-82: a name with a '.', the last
// extension equal to ext), minus `skip`.  Sorted by name, so runs are repeatable.
inline std::vector<std::string> list_files(const std::string &dir, const std::string &ext,
                                           const std::string &skip = "")
{
    std::vector<std::string> out;
    DIR *dp = opendir(dir.c_str());
    if (!dp) return out;
    while (dirent *e = readdir(dp)) {
        const std::string f = e->d_name;
        const size_t dot = f.find_last_of('.');
        if (dot == std::string::npos || f.size() <= 4) continue;
        if (f.substr(dot) != ext) continue;
        if (!skip.empty() && f == skip) continue;
        out.push_back(f);
    }
    closedir(dp);
    std::sort(out.begin(), out.end());
    return out;
}

inline std::string join(const std::string &a, const std::string &b)
{
    if (a.empty()) return b;
    return a.back() == '/' ? a + b : a + "/" + b;
}

// Everything before the last '.', as the reference's a.substr(0, a.find_last_of('.')).
inline std::string stem(const std::string &path)
{
    const size_t dot = path.find_last_of('.');
    return dot == std::string::npos ? path : path.substr(0, dot);
}

inline bool is_dir(const std::string &p)
{
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

inline bool is_file(const std::string &p)
{
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

inline bool ensure_dir(const std::string &p) { return is_dir(p) || mkdir(p.c_str(), 0755) == 0; }

inline int die(bna_ctx *ctx, int rc, const std::string &what)
{
    fprintf(stderr, "error: %s: %s (%s)\n", what.c_str(), bna_last_error(ctx), bna_status_string(rc));
    return 1;
}

// Leading "--name[=value]" options, removed from argv.
struct Opts {
    int device = 0;
    int precision = BNA_BF16X3;
    int recheck = 1;
    int refine = 1;
    int compat_eps = 0;
    int binary = 0;
    bool seeded = false;  // --seed=S: the random networks' generator seed (else the clock)
    uint64_t seed = 0;
};

inline bool parse_opts(int &argc, char **&argv, Opts &o, std::string &err)
{
    int k = 1;
    for (; k < argc && strncmp(argv[k], "--", 2) == 0; k++) {
        const std::string a = argv[k];
        const size_t eq = a.find('=');
        const std::string key = a.substr(2, eq == std::string::npos ? std::string::npos : eq - 2);
        const std::string val = eq == std::string::npos ? "" : a.substr(eq + 1);
        if (key == "device") o.device = atoi(val.c_str());
        else if (key == "precision") {
            if (val == "fp32") o.precision = BNA_FP32;
            else if (val == "bf16") o.precision = BNA_BF16;
            else if (val == "bf16x3") o.precision = BNA_BF16X3;
            else { err = "unknown precision " + val; return false; }
        } else if (key == "no-recheck") o.recheck = 0;
        else if (key == "no-refine") o.refine = 0;
        else if (key == "compat-eps") o.compat_eps = 1;
        else if (key == "binary") o.binary = 1;
        else if (key == "seed") {
            o.seeded = true;
            o.seed = strtoull(val.c_str(), nullptr, 0);
        }
        else { err = "unknown option " + a; return false; }
    }
    argv[k - 1] = argv[0];
    argv += k - 1;
    argc -= k - 1;
    return true;
}

// Seed of random network l of a file: --seed=S gives S + l; without it the
// clock, as the reference's srand(time(0)) (Maslov.cpp:58), so runs differ.
inline uint64_t net_seed(const Opts &o, int l)
{
    static const uint64_t clock_seed = (uint64_t)time(nullptr) * 0x9E3779B97F4A7C15ull;
    return (o.seeded ? o.seed : clock_seed) + (uint64_t)l;
}

}  // namespace bnacli
