// bna_cp -- clustering coefficient for every .csr of a directory, the CUCP tool
// (W/Cp_GPU/CUCP/Cp_main.cpp) on libbna.
//
//   bna_cp [--device=D] [--seed=S] Dir [num_random_networks]
//
// <stem>.cp (int32 N, float[N]) and <stem>_cp.txt (the mean |Cp|, six
// significant digits) beside each .csr.  A binary file (no V section) is taken
// with weight 1 on every edge.  Each random network (bna_maslov with the
// weights carried along, Maslov_weighted_1, then Cp; Cp_main.cpp:130-151) adds
// a line with its mean Cp to _cp.txt.  Random networks are seeded from the
// clock unless --seed is given.
#include "cli_common.hpp"

using namespace bnacli;

int main(int argc, char **argv)
{
    Opts o;
    std::string err;
    if (!parse_opts(argc, argv, o, err) || argc < 2) {
        fprintf(stderr, "%s%susage: bna_cp [--device=D] [--seed=S] dir_for_csr [num_random_networks]\n", err.c_str(),
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
        printf("\ncalculating Cp for %s ...\n", in.c_str());
        bna_csr g;
        if ((rc = bna_read_csr(in.c_str(), &g))) {
            status = die(ctx, rc, "read " + in);
            continue;
        }
        std::vector<float> cp((size_t)g.N);
        double mean = 0.0;
        if ((rc = bna_cp(ctx, &g, cp.data(), &mean))) {
            status = die(ctx, rc, "cp " + in);
            bna_csr_free(&g);
            continue;
        }
        printf("average Cp:\t%.6g\n", mean);
        const std::string out = stem(in) + ".cp";
        printf("Save Clustering Coefficient for each node as %s\n", out.c_str());
        if ((rc = bna_write_nm(out.c_str(), cp.data(), g.N))) status = die(ctx, rc, "write " + out);
        std::vector<double> means{mean};
        if (nrandom > 0) printf("Calculating Cp for random networks...\n");
        for (int l = 0; l < nrandom && status == 0; l++) {
            bna_csr r;
            double m = 0.0;
            if ((rc = bna_maslov(&g, 0, net_seed(o, l), &r))) {
                status = die(ctx, rc, "random network of " + in);
                break;
            }
            if ((rc = bna_cp(ctx, &r, cp.data(), &m))) status = die(ctx, rc, "cp of a random network of " + in);
            else means.push_back(m);
            bna_csr_free(&r);
        }
        const std::string txt = stem(in) + "_cp.txt";
        FILE *ft = fopen(txt.c_str(), "w");
        bool ok = ft != nullptr;
        for (double m : means) ok = ok && fprintf(ft, "%.6g\n", m) >= 0;
        if (!ok) status = die(ctx, BNA_ERR_IO, "write " + txt);
        if (ft) fclose(ft);
        printf("Save average Cp for the brain network and random networks as %s\n", txt.c_str());
        bna_csr_free(&g);
    }
    bna_destroy(ctx);
    return status;
}
