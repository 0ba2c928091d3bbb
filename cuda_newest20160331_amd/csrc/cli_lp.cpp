// bna_lp -- characteristic path length for every .csr of a directory, the
// CUBFS_Lp tool (B/src/Lp/CUBFS_Lp/CUBFS_Lp/CUBFS_LP_main.cpp) on libbna.
//
//   bna_lp [--device=D] [--seed=S] Dir [num_random_networks]
//
// <stem>_Lp.txt beside each .csr: the network's Lp with six significant digits,
// then one line per Maslov random network (bna_maslov, CUBFS_LP_main.cpp:
// 110-123).  The graph is taken as binary.  Random networks are seeded from the
// clock unless --seed is given.
#include "cli_common.hpp"

using namespace bnacli;

int main(int argc, char **argv)
{
    Opts o;
    std::string err;
    if (!parse_opts(argc, argv, o, err) || argc < 2) {
        fprintf(stderr, "%s%susage: bna_lp [--device=D] [--seed=S] dir_for_csr [num_random_networks]\n", err.c_str(),
                err.empty() ? "" : "\n");
        return 1;
    }
    std::string dir = argv[1];
    while (dir.size() > 1 && dir.back() == '/') dir.pop_back();
    if (!is_dir(dir)) {
        fprintf(stderr, "can't open %s\n", dir.c_str());
        return 1;
    }
    const int nrandom = argc > 2 ? std::max(0, atoi(argv[2])) : 0;
    const std::vector<std::string> files = list_files(dir, ".csr");
    printf("%zu files to be processed.\n", files.size());
    bna_ctx *ctx = nullptr;
    int rc = bna_create(&ctx, o.device, 0);
    if (rc) return die(nullptr, rc, "bna_create");
    int status = 0;
    for (const std::string &f : files) {
        const std::string in = join(dir, f);
        printf("\ncalculating Lp for %s ...\n", in.c_str());
        bna_csr g;
        if ((rc = bna_read_csr(in.c_str(), &g))) {
            status = die(ctx, rc, "read " + in);
            continue;
        }
        double lp = 0.0;
        if ((rc = bna_lp(ctx, &g, &lp, nullptr, 0))) {
            status = die(ctx, rc, "lp " + in);
            bna_csr_free(&g);
            continue;
        }
        printf("average Lp:\t%.6g\n", lp);
        std::vector<double> lps{lp};
        if (nrandom > 0) printf("Calculating Lp for random networks...\n");
        for (int l = 0; l < nrandom && status == 0; l++) {
            bna_csr r;
            double m = 0.0;
            if ((rc = bna_maslov(&g, 0, net_seed(o, l), &r))) {
                status = die(ctx, rc, "random network of " + in);
                break;
            }
            if ((rc = bna_lp(ctx, &r, &m, nullptr, 0))) status = die(ctx, rc, "lp of a random network of " + in);
            else lps.push_back(m);
            bna_csr_free(&r);
        }
        const std::string txt = stem(in) + "_Lp.txt";
        FILE *ft = fopen(txt.c_str(), "w");
        bool ok = ft != nullptr;
        for (double m : lps) ok = ok && fprintf(ft, "%.6g\n", m) >= 0;
        if (!ok) status = die(ctx, BNA_ERR_IO, "write " + txt);
        if (ft) fclose(ft);
        printf("Save Lp for the brain network and random networks as %s\n", txt.c_str());
        bna_csr_free(&g);
    }
    bna_destroy(ctx);
    return status;
}
