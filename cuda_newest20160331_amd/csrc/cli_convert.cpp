// bna_convert_nii -- vertex maps (.nm) back onto the mask's voxel grid as NIfTI
// volumes, the ConvertNII tool (W/ConvertNII/ConvertNII/ConvertNII.cpp:16-179).
//
//   bna_convert_nii Dir mask.nii [thr_mask]
//
// Every file of Dir whose extension is exactly .nm (an int32 N, then N floats:
// _deg.nm, _ec.nm) becomes Dir/<stem>.nii.  thr_mask defaults to 0.5.  A map
// whose N differs from the number of voxels with mask >= thr_mask is skipped with
// a message, as the reference does, and the others are still converted.
#include "cli_common.hpp"

using namespace bnacli;

static bool read_nm(const std::string &path, std::vector<float> &v)
{
    FILE *fp = fopen(path.c_str(), "rb");
    if (!fp) return false;
    int32_t n = 0;
    bool ok = fread(&n, 4, 1, fp) == 1 && n >= 0;
    if (ok) {
        v.resize((size_t)n);
        ok = fread(v.data(), 4, (size_t)n, fp) == (size_t)n;
    }
    fclose(fp);
    return ok;
}

int main(int argc, char **argv)
{
    if (argc < 3 || argc > 4) {
        fprintf(stderr, "usage: bna_convert_nii Dir mask.nii [thr_mask]\n"
                        "example: bna_convert_nii /data/bold/weighted /data/bold/mask.nii 0.5\n");
        return 1;
    }
    std::string dir = argv[1];
    while (dir.size() > 1 && dir.back() == '/') dir.pop_back();
    const std::string mask = argv[2];
    const double thr = argc == 4 ? atof(argv[3]) : 0.5;
    const std::vector<std::string> files = list_files(dir, ".nm");
    printf("%zu files to be processed.\n", files.size());
    int status = 0;
    for (const std::string &f : files) {
        const std::string in = join(dir, f);
        printf("\nconvert analysis for %s ...\n", in.c_str());
        std::vector<float> v;
        if (!read_nm(in, v)) {
            fprintf(stderr, "error: cannot read %s\n", in.c_str());
            status = 1;
            continue;
        }
        const std::string out = stem(in) + ".nii";
        const int rc = bna_nm_to_nifti(mask.c_str(), thr, v.data(), (int)v.size(), out.c_str());
        if (rc == BNA_ERR_FORMAT) {
            printf("Mask file doesn't match to the file %s\nContinue to convert the next file\n", f.c_str());
            continue;
        }
        if (rc) {
            fprintf(stderr, "error: %s: %s\n", out.c_str(), bna_status_string(rc));
            status = 1;
            continue;
        }
        printf("Convert to .nii completed!\n");
    }
    return status;
}
