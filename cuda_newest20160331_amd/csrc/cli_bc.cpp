// bna_bc -- betweenness centrality for every .csr of a directory, the CUBC tool
// (B/src/BC/CUBC/CUBC_main.cpp) on libbna.
//
//   bna_bc [--device=D] Dir
//
// <stem>_bc.nm beside each .csr: int32 N, then float BC[N] (CUBC_main.cpp:95-103).
// The graph is taken as binary.
#include "cli_common.hpp"

using namespace bnacli;

int main(int argc, char **argv)
{
    Opts o;
    std::string err;
    if (!parse_opts(argc, argv, o, err) || argc < 2) {
        fprintf(stderr, "%s%susage: bna_bc [--device=D] dir_for_csr\n", err.c_str(), err.empty() ? "" : "\n");
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
        printf("\ncalculating Betweenness for %s ...\n", in.c_str());
        bna_csr g;
        if ((rc = bna_read_csr(in.c_str(), &g))) {
            status = die(ctx, rc, "read " + in);
            continue;
        }
        std::vector<float> bc((size_t)g.N);
        if ((rc = bna_bc(ctx, &g, bc.data()))) {
            status = die(ctx, rc, "bc " + in);
            bna_csr_free(&g);
            continue;
        }
        const std::string out = stem(in) + "_bc.nm";
        printf("Save betweenness data for each node as %s\n", out.c_str());
        if ((rc = bna_write_nm(out.c_str(), bc.data(), g.N))) status = die(ctx, rc, "write " + out);
        bna_csr_free(&g);
    }
    bna_destroy(ctx);
    return status;
}
