// bna_degree -- degree maps for every .csr of a directory, the Degree tool
// (W/Degree/Degree/Degree.cpp:13-110, binary B/Degree/Degree/Degree.cpp) on libbna.
//
//   bna_degree [--device=D] Dir
//
// <stem>_deg.nm (int32 N, float[N]) beside each .csr.  A weighted file gives the
// float sum of its row's V in CSR order; a binary file (no V section) gives the
// row length, and its count of isolated nodes is appended to
// Dir/isolated_v_mark.txt as the binary tool does (B/Degree/Degree/Degree.cpp:84-92).
#include "cli_common.hpp"

using namespace bnacli;

int main(int argc, char **argv)
{
    Opts o;
    std::string err;
    if (!parse_opts(argc, argv, o, err) || argc < 2) {
        fprintf(stderr, "%s%susage: bna_degree [--device=D] dir_for_csr\n", err.c_str(), err.empty() ? "" : "\n");
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
    FILE *iso = nullptr;
    int status = 0;
    for (const std::string &f : files) {
        const std::string in = join(dir, f);
        printf("\ncalculating Degree for %s ...\n", in.c_str());
        bna_csr g;
        if ((rc = bna_read_csr(in.c_str(), &g))) {
            status = die(ctx, rc, "read " + in);
            continue;
        }
        std::vector<float> deg((size_t)g.N);
        if ((rc = bna_degree(ctx, &g, deg.data()))) {
            status = die(ctx, rc, "degree " + in);
            bna_csr_free(&g);
            continue;
        }
        if (!g.V) {
            int isolated = 0;
            for (float d : deg) isolated += d == 0.0f;
            if (!iso) iso = fopen(join(dir, "isolated_v_mark.txt").c_str(), "w");
            if (iso) fprintf(iso, "%d\n", isolated);
        }
        const std::string out = stem(in) + "_deg.nm";
        printf("Save degree data for each node as %s\n", out.c_str());
        if ((rc = bna_write_nm(out.c_str(), deg.data(), g.N))) status = die(ctx, rc, "write " + out);
        bna_csr_free(&g);
    }
    if (iso) fclose(iso);
    bna_destroy(ctx);
    return status;
}
