// bna_lp -- characteristic path length for every .csr of a directory, the
// CUBFS_Lp tool (B/src/Lp/CUBFS_Lp/CUBFS_Lp/CUBFS_LP_main.cpp) on libbna.
//
//   bna_lp [--device=D] Dir [num_random_networks]
//
// <stem>_Lp.txt beside each .csr: the network's Lp with six significant digits
// (CUBFS_LP_main.cpp:117-122).  The graph is taken as binary.  The Maslov
// random networks the reference adds one line each for are not provided: a
// positive num_random_networks is reported and ignored.
#include "cli_common.hpp"

using namespace bnacli;

int main(int argc, char **argv)
{
    Opts o;
    std::string err;
    if (!parse_opts(argc, argv, o, err) || argc < 2) {
        fprintf(stderr, "%s%susage: bna_lp [--device=D] dir_for_csr [num_random_networks]\n", err.c_str(),
                err.empty() ? "" : "\n");
        return 1;
    }
    std::string dir = argv[1];
    while (dir.size() > 1 && dir.back() == '/') dir.pop_back();
    if (!is_dir(dir)) {
        fprintf(stderr, "can't open %s\n", dir.c_str());
        return 1;
    }
    if (argc > 2 && atoi(argv[2]) > 0)
        fprintf(stderr, "random-network (Maslov) comparison is not provided; only the network's Lp is written\n");
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
        const std::string txt = stem(in) + "_Lp.txt";
        FILE *ft = fopen(txt.c_str(), "w");
        if (!ft || fprintf(ft, "%.6g\n", lp) < 0) status = die(ctx, BNA_ERR_IO, "write " + txt);
        if (ft) fclose(ft);
        printf("Save Lp for the brain network as %s\n", txt.c_str());
        bna_csr_free(&g);
    }
    bna_destroy(ctx);
    return status;
}
