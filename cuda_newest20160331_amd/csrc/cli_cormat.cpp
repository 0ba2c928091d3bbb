// bna_cormat -- network construction from NIfTI series, the CUCorMat command
// line (W/CorMat/CUCorMat/CUCorMat/CorMat.cpp:52-384) on libbna.
//
//   bna_cormat [--binary] [--precision=fp32|bf16|bf16x3] [--device=D] [--no-recheck]
//              [--no-refine] [--compat-eps]
//              Dir [mask.nii] thr_mask avg(n|yn|yf|bn|bf) save_cormat(y|n) type(r|s) thr [thr ...]
//
// Every .nii file of Dir except mask.nii is one subject.  The mask path may be
// left out, as W/bin/BNC.bat:34 does; then Dir/mask.nii is used.
//
// avg n (and b): each subject on its own (CorMat.cpp:196-384).  For each subject
// and threshold the CSR goes to
//   Dir/weighted/<stem>_<N>_spa<s>%_cor<r>_weighted.csr     (weighted, the default)
//   Dir/unweighted/<stem>_<N>_spa<s>%_cor<r>.csr            (--binary)
// with s and r printed "%.3f" as CorMat.cpp:345 does: for an r threshold s is the
// density reached (100 * Clength / M1 / 2), for a density threshold r is the
// k-th largest r.  save_cormat y also writes the subject's packed r matrix
// (invalid pairs 0) to <same dir>/<stem>_<N>.cormat (CorMat.cpp:301-310).
//
// avg y (and b): the group average of every subject's matrix (CorMat.cpp:418-670),
// second letter n for the mean of r, f for the mean in Fisher z mapped back.
// Files go to Dir itself, named after Dir's last component:
//   Dir/<last>_<subjects>_<N>_spa<s>%_cor<r>.csr   and, with save_cormat y,
//   Dir/<last>_<subjects>_<N>.cormat
#include "cli_common.hpp"

#include <chrono>

using namespace bnacli;

static void usage()
{
    fprintf(stderr,
            "usage: bna_cormat [--binary] [--precision=fp32|bf16|bf16x3] [--device=D] [--no-recheck]\n"
            "                  [--no-refine] [--compat-eps]\n"
            "                  Dir [mask.nii] thr_mask avg(n|yn|yf|bn|bf) save_cormat(y|n) type(r|s) thr [thr ...]\n"
            "example: bna_cormat /data/bold /data/bold/mask.nii 0.5 n n r 0.2 0.25 0.3\n");
}

int main(int argc, char **argv)
{
    Opts o;
    std::string err;
    if (!parse_opts(argc, argv, o, err)) {
        fprintf(stderr, "error: %s\n", err.c_str());
        usage();
        return 1;
    }
    if (argc < 7) {
        usage();
        return 1;
    }
    std::string dir = argv[1];
    while (dir.size() > 1 && dir.back() == '/') dir.pop_back();
    int a = 2;
    std::string mask_path;
    const std::string a2 = argv[2];
    if (argc >= 8 && (is_file(a2) || (a2.size() > 4 && a2.substr(a2.size() - 4) == ".nii"))) {
        mask_path = a2;
        a = 3;
    } else {
        mask_path = join(dir, "mask.nii");
    }
    if (argc - a < 5) {
        usage();
        return 1;
    }
    const float mask_thr = (float)atof(argv[a]);
    const char avg = argv[a + 1][0];
    const char avg2 = argv[a + 1][0] ? argv[a + 1][1] : 0;
    const char save = argv[a + 2][0];
    const char type = argv[a + 3][0];
    const int nthr = argc - (a + 4);
    if (nthr < 1) {
        usage();
        return 1;
    }
    const bool single = avg == 'n' || avg == 'N' || avg == 'b' || avg == 'B';
    const bool group = avg == 'y' || avg == 'Y' || avg == 'b' || avg == 'B';
    const bool fisher = avg2 == 'f' || avg2 == 'F';
    if (!single && !group) {
        fprintf(stderr, "avg type error! n: each subject, y: group average, b: both\n");
        return 1;
    }
    if (group && !fisher && avg2 != 'n' && avg2 != 'N') {
        fprintf(stderr, "avg type error! the group average needs a second letter: n (mean of r) or f (Fisher z)\n");
        return 1;
    }
    const bool save_cormat = save == 'y' || save == 'Y';
    if (!save_cormat && save != 'n' && save != 'N') {
        fprintf(stderr, "to_save_cor_matrix type error! y to save the whole correlation matrix, n to save only csr\n");
        return 1;
    }
    const bool rmode = type == 'r' || type == 'R';
    if (!rmode && type != 's' && type != 'S') {
        fprintf(stderr, "threshold type error! r for correlation threshold, s for sparsity threshold\n");
        return 1;
    }
    std::vector<float> thr(nthr);
    for (int k = 0; k < nthr; k++) thr[k] = (float)atof(argv[a + 4 + k]);
    if (rmode) {
        // CorMat.cpp:180-187 moves the smallest r threshold to the front
        int mi = 0;
        for (int k = 0; k < nthr; k++)
            if (thr[k] < thr[mi]) mi = k;
        std::swap(thr[0], thr[mi]);
    }

    const std::vector<std::string> files = list_files(dir, ".nii", "mask.nii");
    printf("%zu file(s) to be processed.\n", files.size());

    uint8_t *sel = nullptr;
    int64_t total = 0;
    int N = 0;
    int rc = bna_nifti_mask(mask_path.c_str(), mask_thr, &sel, &total, &N);
    if (rc) return die(nullptr, rc, "mask " + mask_path);
    printf("Data size: %lld voxels, Grey voxel count: %d.\n", (long long)total, N);

    bna_ctx *ctx = nullptr;
    if ((rc = bna_create(&ctx, o.device, 0))) return die(nullptr, rc, "bna_create");
    const std::string outdir = join(dir, o.binary ? "unweighted" : "weighted");
    if (single && !files.empty() && !ensure_dir(outdir)) {
        fprintf(stderr, "error: cannot create %s\n", outdir.c_str());
        return 1;
    }
    const double M1 = (double)N * (N - 1) / 2.0;
    int status = 0;
    auto opts_for = [&](int k) {
        bna_opts op;
        bna_opts_default(&op);
        op.thr_type = rmode ? BNA_THR_R : BNA_THR_DENSITY;
        op.thr = thr[k];
        op.weighted = !o.binary;
        op.precision = o.precision;
        op.recheck = o.recheck;
        op.refine = o.refine;
        op.ref_compat_eps = o.compat_eps;
        return op;
    };
    auto tag_for = [&](int k, const bna_csr &g) {
        float s_val, r_val;
        if (rmode) {
            s_val = (float)(100.0 * (double)g.nnz / M1 / 2.0);
            r_val = thr[k];
        } else {
            s_val = thr[k];
            r_val = (float)g.thr_used;
            printf("threshold for sparsity = %g\n", (double)thr[k]);
        }
        char tag[64];
        snprintf(tag, sizeof tag, "_spa%.3f%%_cor%.3f", (double)s_val, (double)r_val);
        return std::string(tag);
    };
    // the packed averages of a group to a .cormat file
    auto dump_cormat = [&](bna_group *grp, const std::string &path) {
        std::vector<float> packed((size_t)((int64_t)N * (N - 1) / 2));
        int r = bna_group_download(ctx, grp, packed.data());
        if (r) return die(ctx, r, "download " + path);
        printf("generating %s...\n", path.c_str());
        if ((r = bna_write_cormat(path.c_str(), packed.data(), N))) return die(ctx, r, "write " + path);
        return 0;
    };
    for (const std::string &f : (single ? files : std::vector<std::string>())) {
        const std::string in = join(dir, f);
        printf("\ncalculating correlation for %s ...\n", in.c_str());
        float *x = nullptr;
        int L = 0;
        if ((rc = bna_nifti_bold(in.c_str(), sel, total, N, &x, &L))) {
            status = die(ctx, rc, "read " + in);
            continue;
        }
        printf("BOLD length: %d\n", L);
        bna_series *s = nullptr;
        rc = bna_series_create(ctx, x, N, L, BNA_TIME_MAJOR, 0, &s);
        bna_free(x);
        if (rc) {
            status = die(ctx, rc, "upload " + in);
            continue;
        }
        const std::string base = stem(join(outdir, f)) + "_" + std::to_string(N);
        if (save_cormat) {  // CorMat.cpp:301-310
            std::vector<float> packed((size_t)((int64_t)N * (N - 1) / 2));
            const std::string path = base + ".cormat";
            if ((rc = bna_corr_packed(ctx, s, o.precision, packed.data()))) {
                status = die(ctx, rc, "r matrix of " + in);
            } else {
                printf("generating %s...\n", path.c_str());
                if ((rc = bna_write_cormat(path.c_str(), packed.data(), N))) status = die(ctx, rc, "write " + path);
            }
        }
        for (int k = 0; k < nthr; k++) {
            const bna_opts op = opts_for(k);
            const auto t0 = std::chrono::steady_clock::now();
            bna_csr dev, g;
            if ((rc = bna_build_csr_series(ctx, s, &op, &dev))) {
                status = die(ctx, rc, "build " + in);
                continue;
            }
            rc = bna_csr_to_host(ctx, &dev, &g);
            bna_csr_free(&dev);
            if (rc) {
                status = die(ctx, rc, "copy " + in);
                continue;
            }
            const double ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            const std::string out = base + tag_for(k, g) + (o.binary ? ".csr" : "_weighted.csr");
            printf("graph: %lld edges, %.1f ms; generating %s...\n", (long long)(g.nnz / 2), ms, out.c_str());
            if ((rc = bna_write_csr(out.c_str(), &g))) status = die(ctx, rc, "write " + out);
            bna_csr_free(&g);
        }
        bna_series_destroy(ctx, s);
    }

    if (group && !files.empty()) {
        // Dir/<last component>_<subjects>_<N> (CorMat.cpp:562-569)
        const size_t sl = dir.find_last_of('/');
        const std::string last = sl == std::string::npos ? std::string() : dir.substr(sl + 1);
        const std::string base = join(dir, last + "_" + std::to_string(files.size()) + "_" + std::to_string(N));
        bna_group *grp = nullptr;
        if ((rc = bna_group_create(ctx, N, fisher ? 1 : 0, &grp))) {
            status = die(ctx, rc, "group matrix");
        } else {
            bool ok = true;
            for (const std::string &f : files) {
                const std::string in = join(dir, f);
                printf("\ncalculating correlation for %s ...\n", in.c_str());
                float *x = nullptr;
                int L = 0;
                bna_series *s = nullptr;
                if ((rc = bna_nifti_bold(in.c_str(), sel, total, N, &x, &L)) ||
                    (rc = bna_series_create(ctx, x, N, L, BNA_TIME_MAJOR, 0, &s))) {
                    bna_free(x);
                    status = die(ctx, rc, "read " + in);
                    ok = false;
                    break;
                }
                bna_free(x);
                printf("BOLD length: %d\n", L);
                rc = bna_group_add(ctx, grp, s, o.precision);
                bna_series_destroy(ctx, s);
                if (rc) {
                    status = die(ctx, rc, "add " + in);
                    ok = false;
                    break;
                }
            }
            if (ok && (rc = bna_group_finish(ctx, grp))) {
                status = die(ctx, rc, "group average");
                ok = false;
            }
            if (ok && save_cormat && dump_cormat(grp, base + ".cormat")) status = 1;
            for (int k = 0; ok && k < nthr; k++) {
                bna_opts op = opts_for(k);
                bna_csr dev, g;
                if ((rc = bna_group_build_csr(ctx, grp, &op, &dev))) {
                    status = die(ctx, rc, "group graph");
                    continue;
                }
                rc = bna_csr_to_host(ctx, &dev, &g);
                bna_csr_release(ctx, &dev);
                if (rc) {
                    status = die(ctx, rc, "copy group graph");
                    continue;
                }
                const std::string out = base + tag_for(k, g) + ".csr";
                printf("group graph: %lld edges; generating %s...\n", (long long)(g.nnz / 2), out.c_str());
                if ((rc = bna_write_csr(out.c_str(), &g))) status = die(ctx, rc, "write " + out);
                bna_csr_free(&g);
            }
            bna_group_destroy(ctx, grp);
        }
    }
    bna_free(sel);
    bna_destroy(ctx);
    return status;
}
