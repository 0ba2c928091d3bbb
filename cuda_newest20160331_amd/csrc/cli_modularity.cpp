// bna_modularity -- Newman modularity partition for every .csr of a directory,
// the CUmodularity tool (B/Newman_Modularity_GPU/CUmodularity/modularity_GPU.cu:531-726)
// on libbna.
//
//   bna_modularity [--device=D] Dir [num_random_networks]
//
// <stem>_modu.nm (int32 N, float[N] module labels from 1) and <stem>_modu.txt
// (module count and Q, the line the reference logs after the partition) beside
// each .csr.  The graph is taken as binary, as the reference reads it.  The
// Maslov random networks the reference partitions afterwards only add timing
// to its log: a positive num_random_networks is reported and ignored.
#include "cli_common.hpp"

using namespace bnacli;

int main(int argc, char **argv)
{
    Opts o;
    std::string err;
    if (!parse_opts(argc, argv, o, err) || argc < 2) {
        fprintf(stderr, "%s%susage: bna_modularity [--device=D] dir_for_csr [num_random_networks]\n", err.c_str(),
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
        fprintf(stderr, "random-network (Maslov) partitions are not provided; only the network's partition is written\n");
    const std::vector<std::string> files = list_files(dir, ".csr");
    printf("%zu files to be processed.\n", files.size());
    bna_ctx *ctx = nullptr;
    int rc = bna_create(&ctx, o.device, 0);
    if (rc) return die(nullptr, rc, "bna_create");
    int status = 0;
    for (const std::string &f : files) {
        const std::string in = join(dir, f);
        printf("\nModular analysis for %s ...\n", in.c_str());
        bna_csr g;
        if ((rc = bna_read_csr(in.c_str(), &g))) {
            status = die(ctx, rc, "read " + in);
            continue;
        }
        std::vector<int32_t> lab((size_t)g.N);
        double Q = 0.0;
        int mods = 0;
        if ((rc = bna_modularity(ctx, &g, lab.data(), &Q, &mods))) {
            status = die(ctx, rc, "modularity " + in);
            bna_csr_free(&g);
            continue;
        }
        printf("Number of Modules: %d,\tQ=%.6g\n", mods, Q);
        const std::string out = stem(in) + "_modu.nm";
        printf("Save partition results as %s\n", out.c_str());
        std::vector<float> fl(lab.begin(), lab.end());
        if ((rc = bna_write_nm(out.c_str(), fl.data(), g.N))) status = die(ctx, rc, "write " + out);
        const std::string txt = stem(in) + "_modu.txt";
        FILE *ft = fopen(txt.c_str(), "w");
        if (!ft || fprintf(ft, "Number of Modules: %d,\tQ=%.6g\n", mods, Q) < 0) status = die(ctx, BNA_ERR_IO, "write " + txt);
        if (ft) fclose(ft);
        bna_csr_free(&g);
    }
    bna_destroy(ctx);
    return status;
}
