// bna_lpw -- weighted characteristic path length and nodal efficiency for every
// .csr of a directory, the CUBFW_Lp tool
// (W/src/Lp/CUBFW_Lp/CUBFW_Lp/CUBFW_Lp_main.cpp) on libbna.
//
//   bna_lpw [--device=D] [--seed=S] Dir [num_random_networks] [g|n|l|gn|gl|nl|gnl]
//
// The last argument picks the outputs as the reference's third argument does
// (CUBFW_Lp_main.cpp:67-89; default gn): g writes <stem>_Lp.txt (the network's
// Lp, then one line per random network, six significant digits, :236-245), n
// writes <stem>_eff.nm (int32 N, then the float nodal efficiency per vertex,
// :193-205), l writes <stem>_lambda.txt (the network's Lp over the random
// networks' mean, :246-253).  The edge costs are 1 / w (unit costs for a .csr
// without weights).  Random networks come from bna_maslov with the weights
// carried along (Maslov_weighted_1, :219-232), seeded from the clock unless
// --seed is given.
#include "cli_common.hpp"

#include <cstring>

using namespace bnacli;

int main(int argc, char **argv)
{
    Opts o;
    std::string err;
    if (!parse_opts(argc, argv, o, err) || argc < 2) {
        fprintf(stderr, "%s%susage: bna_lpw [--device=D] [--seed=S] dir_for_csr [num_random_networks] [g|n|l|gn|gl|nl|gnl]\n",
                err.c_str(), err.empty() ? "" : "\n");
        return 1;
    }
    std::string dir = argv[1];
    while (dir.size() > 1 && dir.back() == '/') dir.pop_back();
    if (!is_dir(dir)) {
        fprintf(stderr, "can't open %s\n", dir.c_str());
        return 1;
    }
    const char *kinds = argc > 3 ? argv[3] : "gn";
    static const char *const valid[] = {"g", "n", "l", "gn", "gl", "nl", "gnl"};
    bool known = false;
    for (const char *v : valid) known = known || strcmp(kinds, v) == 0;
    if (!known) {
        fprintf(stderr, "unknown output selection %s (g, n, l, gn, gl, nl or gnl)\n", kinds);
        return 1;
    }
    const bool global = strchr(kinds, 'g') != nullptr, nodal = strchr(kinds, 'n') != nullptr,
               lambda = strchr(kinds, 'l') != nullptr;
    const int nrandom = argc > 2 ? std::max(0, atoi(argv[2])) : 0;
    if (lambda && nrandom == 0) {
        fprintf(stderr, "lambda needs at least one random network\n");
        return 1;
    }
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
        int isolated = 0;
        std::vector<float> eff((size_t)g.N);
        if ((rc = bna_lp_weighted(ctx, &g, &lp, eff.data(), &isolated))) {
            status = die(ctx, rc, "lp " + in);
            bna_csr_free(&g);
            continue;
        }
        if (isolated > 0) printf("\nisolated vertices number : %d\n", isolated);
        printf("average Lp:\t%.6g\n", lp);
        if (nodal) {
            const std::string out = stem(in) + "_eff.nm";
            printf("Save nodal efficiency for each node as %s\n", out.c_str());
            if ((rc = bna_write_nm(out.c_str(), eff.data(), g.N))) status = die(ctx, rc, "write " + out);
        }
        std::vector<double> lps{lp};
        if (nrandom > 0) printf("Calculating Lp for random networks...\n");
        for (int l = 0; l < nrandom && status == 0; l++) {
            bna_csr r;
            double m = 0.0;
            int iso = 0;
            if ((rc = bna_maslov(&g, 0, net_seed(o, l), &r))) {
                status = die(ctx, rc, "random network of " + in);
                break;
            }
            if ((rc = bna_lp_weighted(ctx, &r, &m, eff.data(), &iso)))
                status = die(ctx, rc, "lp of a random network of " + in);
            else lps.push_back(m);
            bna_csr_free(&r);
        }
        if (global) {
            const std::string txt = stem(in) + "_Lp.txt";
            FILE *ft = fopen(txt.c_str(), "w");
            bool ok = ft != nullptr;
            for (double m : lps) ok = ok && fprintf(ft, "%.6g\n", m) >= 0;
            if (!ok) status = die(ctx, BNA_ERR_IO, "write " + txt);
            if (ft) fclose(ft);
            printf("Save Lp for the brain network and random networks as %s\n", txt.c_str());
        }
        if (lambda && lps.size() > 1) {
            double mr = 0.0;
            for (size_t q = 1; q < lps.size(); q++) mr += lps[q];
            mr /= (double)(lps.size() - 1);
            const std::string txt = stem(in) + "_lambda.txt";
            FILE *ft = fopen(txt.c_str(), "w");
            if (!ft || fprintf(ft, "%.6g\n", lp / mr) < 0) status = die(ctx, BNA_ERR_IO, "write " + txt);
            if (ft) fclose(ft);
            printf("Save corresponding lambda for the brain network and random networks as %s\n", txt.c_str());
        }
        bna_csr_free(&g);
    }
    bna_destroy(ctx);
    return status;
}
