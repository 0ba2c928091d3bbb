// bna_ec -- eigenvector centrality for every .csr of a directory, the CUEC tool
// (W/CUEC/weighted-net.cu:222-332; binary B/CUEC/CUEC/CUEC_unweighted.cu) on libbna.
//
//   bna_ec [--device=D] Dir
//
// Beside each .csr: <stem>_ec.nm (int32 N, float[N]) and <stem>_evalue.txt (the
// last norm, printed with 6 significant digits as `setprecision(6) << x`).  A
// weighted file gets the weighted tool's output (A*y of the last iterate), a
// binary file the binary tool's (the unit-norm iterate).  Tolerance 1e-6,
// at most 10000 iterations.
#include "cli_common.hpp"

using namespace bnacli;

int main(int argc, char **argv)
{
    Opts o;
    std::string err;
    if (!parse_opts(argc, argv, o, err) || argc < 2) {
        fprintf(stderr, "%s%susage: bna_ec [--device=D] dir_for_csr\n", err.c_str(), err.empty() ? "" : "\n");
        return 1;
    }
    std::string dir = argv[1];
    while (dir.size() > 1 && dir.back() == '/') dir.pop_back();
    if (!is_dir(dir)) {
        fprintf(stderr, "can't open %s\n", dir.c_str());
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
        printf("\ncalculating eigenvalue centrality for %s ...\n", in.c_str());
        bna_csr g;
        if ((rc = bna_read_csr(in.c_str(), &g))) {
            status = die(ctx, rc, "read " + in);
            continue;
        }
        std::vector<double> ec((size_t)g.N);
        double lambda = 0.0;
        int iters = 0;
        rc = bna_eigcent(ctx, &g, 1e-6, 10000, g.V ? BNA_EC_WEIGHTED_REF : BNA_EC_BINARY_REF, ec.data(), &lambda,
                         &iters);
        if (rc) {
            status = die(ctx, rc, "eigcent " + in);
            bna_csr_free(&g);
            continue;
        }
        std::vector<float> ecf(ec.begin(), ec.end());
        const std::string out = stem(in) + "_ec.nm";
        const std::string ev = stem(in) + "_evalue.txt";
        printf("Save eigenvector centrality for each node as %s (%d iterations)\n", out.c_str(), iters);
        if ((rc = bna_write_nm(out.c_str(), ecf.data(), g.N))) status = die(ctx, rc, "write " + out);
        FILE *fp = fopen(ev.c_str(), "w");
        if (fp) {
            fprintf(fp, "%.6g\n", lambda);
            fclose(fp);
        } else {
            status = 1;
        }
        bna_csr_free(&g);
    }
    bna_destroy(ctx);
    return status;
}
