// host_util.cpp -- host-only parts of the C ABI: the density -> edge count rule,
// the seeded factor-model generator, and the .csr / .nm file formats.
#include "bna.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <unordered_set>
#include <vector>

extern "C" {

// CorMat.cpp:327-328: `(long long) 2*M1*s/100.0` multiplies the long long 2*M1
// by the float s in float arithmetic, divides by 100.0 in double and truncates;
// an odd count is then made even.  Returns undirected edges (Clength / 2).
int64_t bna_density_to_k(int N, float spa_percent)
{
    if (N < 2) return 0;
    const long long M1 = (long long)(N - 1) * N / 2;
    const long long twoM1 = 2LL * M1;
    const float prod = (float)twoM1 * spa_percent;
    long long Clength = (long long)((double)prod / 100.0);
    Clength += Clength % 2;
    return Clength / 2;
}

}  // extern "C"

// ---------------------------------------------------------------- generator
namespace {

inline uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// xoshiro256** seeded through splitmix64 from (seed, stream id).
struct Rng {
    uint64_t s[4];
    Rng(uint64_t seed, uint64_t stream)
    {
        uint64_t x = seed ^ (stream * 0xd1b54a32d192ed03ull + 0x8bb84b93962eacc9ull);
        for (auto &w : s) w = splitmix64(x);
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0, 1)
    // standard normal, Box-Muller (one value per call, the partner is dropped)
    double normal()
    {
        double u1 = uniform();
        while (u1 <= 0.0) u1 = uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    int below(int n) { return (int)((next() >> 33) % (uint64_t)n); }
};

}  // namespace

extern "C" int bna_synth_factor_model(float *x, int N, int L, int K, double phi, uint64_t seed, int layout)
{
    if (!x || N < 1 || L < 1 || K < 2 || !(phi > -1.0 && phi < 1.0)) return BNA_ERR_ARG;
    if (layout != BNA_TIME_MAJOR && layout != BNA_VOXEL_MAJOR) return BNA_ERR_ARG;
    // latent AR(1) signals, unit variance: f0 ~ N(0,1), f_t = phi f_{t-1} + sqrt(1-phi^2) e_t
    std::vector<double> f((size_t)K * L);
    const double c = std::sqrt(1.0 - phi * phi);
    for (int k = 0; k < K; k++) {
        Rng r(seed, 0x1000000000ull + (uint64_t)k);
        double v = r.normal();
        f[(size_t)k * L] = v;
        for (int t = 1; t < L; t++) {
            v = phi * v + c * r.normal();
            f[(size_t)k * L + t] = v;
        }
    }
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    auto work = [&](unsigned tid) {
        for (int v = (int)tid; v < N; v += (int)nt) {
            Rng r(seed, (uint64_t)v);
            const int k1 = r.below(K);
            int k2 = r.below(K - 1);
            if (k2 >= k1) k2++;
            const double a1 = 0.5 * r.normal(), a2 = 0.5 * r.normal();
            const double *f1 = &f[(size_t)k1 * L], *f2 = &f[(size_t)k2 * L];
            for (int t = 0; t < L; t++) {
                const double val = a1 * f1[t] + a2 * f2[t] + r.normal();
                if (layout == BNA_VOXEL_MAJOR)
                    x[(size_t)v * L + t] = (float)val;
                else
                    x[(size_t)t * N + v] = (float)val;
            }
        }
    };
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nt; t++) th.emplace_back(work, t);
    work(0);
    for (auto &t : th) t.join();
    return BNA_OK;
}

// ---------------------------------------------------------------- files
extern "C" int bna_write_csr(const char *path, const bna_csr *g)
{
    if (!path || !g || g->on_device || !g->R || (g->nnz > 0 && !g->C)) return BNA_ERR_ARG;
    if (g->nnz > 0x7fffffffLL || g->N >= 0x7fffffff) return BNA_ERR_OVERFLOW;
    FILE *fp = fopen(path, "wb");
    if (!fp) return BNA_ERR_IO;
    const int32_t rl = g->N + 1, cl = (int32_t)g->nnz;
    bool ok = fwrite(&rl, 4, 1, fp) == 1 && fwrite(g->R, 4, (size_t)rl, fp) == (size_t)rl &&
              fwrite(&cl, 4, 1, fp) == 1 && fwrite(g->C, 4, (size_t)cl, fp) == (size_t)cl;
    if (ok && g->V) ok = fwrite(&cl, 4, 1, fp) == 1 && fwrite(g->V, 4, (size_t)cl, fp) == (size_t)cl;
    ok = (fclose(fp) == 0) && ok;
    return ok ? BNA_OK : BNA_ERR_IO;
}

// .cormat (CorMat.cpp:301-310, 570-579): the pair count as a 4-byte int (the
// low 32 bits of M1, which the reference writes with sizeof(int)), then the
// packed upper triangle as floats.
extern "C" int bna_write_cormat(const char *path, const float *packed, int N)
{
    if (!path || !packed || N < 2) return BNA_ERR_ARG;
    const int64_t M1 = (int64_t)N * (N - 1) / 2;
    const uint32_t head = (uint32_t)((uint64_t)M1 & 0xffffffffu);
    FILE *fp = fopen(path, "wb");
    if (!fp) return BNA_ERR_IO;
    bool ok = fwrite(&head, 4, 1, fp) == 1 && fwrite(packed, 4, (size_t)M1, fp) == (size_t)M1;
    ok = (fclose(fp) == 0) && ok;
    return ok ? BNA_OK : BNA_ERR_IO;
}

extern "C" int bna_read_csr(const char *path, bna_csr *out)
{
    if (!path || !out) return BNA_ERR_ARG;
    memset(out, 0, sizeof(*out));
    FILE *fp = fopen(path, "rb");
    if (!fp) return BNA_ERR_IO;
    int32_t rl = 0, cl = 0, vl = 0;
    int rc = BNA_OK;
    int32_t *R = nullptr, *C = nullptr;
    float *V = nullptr;
    if (fread(&rl, 4, 1, fp) != 1 || rl < 1) { rc = BNA_ERR_FORMAT; goto done; }
    R = (int32_t *)malloc(4 * (size_t)rl);
    if (!R) { rc = BNA_ERR_NOMEM; goto done; }
    if (fread(R, 4, (size_t)rl, fp) != (size_t)rl) { rc = BNA_ERR_FORMAT; goto done; }
    if (fread(&cl, 4, 1, fp) != 1 || cl < 0) { rc = BNA_ERR_FORMAT; goto done; }
    C = (int32_t *)malloc(4 * (size_t)std::max(cl, 1));
    if (!C) { rc = BNA_ERR_NOMEM; goto done; }
    if (fread(C, 4, (size_t)cl, fp) != (size_t)cl) { rc = BNA_ERR_FORMAT; goto done; }
    if (fread(&vl, 4, 1, fp) == 1) {
        // weighted-net.cu:283-284 calls a file with Vlength != Clength damaged
        if (vl != cl) { rc = BNA_ERR_FORMAT; goto done; }
        V = (float *)malloc(4 * (size_t)std::max(vl, 1));
        if (!V) { rc = BNA_ERR_NOMEM; goto done; }
        if (fread(V, 4, (size_t)vl, fp) != (size_t)vl) { rc = BNA_ERR_FORMAT; goto done; }
    }
    // structural checks: R monotone from 0 to nnz, columns in range
    if (R[0] != 0 || R[rl - 1] != cl) { rc = BNA_ERR_FORMAT; goto done; }
    for (int32_t k = 0; k + 1 < rl; k++)
        if (R[k + 1] < R[k]) { rc = BNA_ERR_FORMAT; goto done; }
    for (int32_t k = 0; k < cl; k++)
        if (C[k] < 0 || C[k] >= rl - 1) { rc = BNA_ERR_FORMAT; goto done; }
done:
    fclose(fp);
    if (rc != BNA_OK) {
        free(R);
        free(C);
        free(V);
        return rc;
    }
    out->N = rl - 1;
    out->nnz = cl;
    out->R = R;
    out->C = C;
    out->V = V;
    out->on_device = 0;
    return BNA_OK;
}

extern "C" int bna_check_csr(const bna_csr *g, int ascending)
{
    if (!g || g->on_device || g->N < 0 || g->nnz < 0 || !g->R || (g->nnz > 0 && !g->C)) return BNA_ERR_ARG;
    if (g->R[0] != 0 || (int64_t)g->R[g->N] != g->nnz) return BNA_ERR_FORMAT;
    for (int32_t i = 0; i < g->N; i++) {
        if (g->R[i + 1] < g->R[i]) return BNA_ERR_FORMAT;
        for (int64_t k = g->R[i]; k < g->R[i + 1]; k++) {
            if (g->C[k] < 0 || g->C[k] >= g->N) return BNA_ERR_FORMAT;
            if (ascending && k > g->R[i] && g->C[k] <= g->C[k - 1]) return BNA_ERR_FORMAT;
        }
    }
    return BNA_OK;
}

extern "C" int bna_write_nm(const char *path, const float *v, int N)
{
    if (!path || !v || N < 0) return BNA_ERR_ARG;
    FILE *fp = fopen(path, "wb");
    if (!fp) return BNA_ERR_IO;
    const int32_t n = N;
    bool ok = fwrite(&n, 4, 1, fp) == 1 && fwrite(v, 4, (size_t)N, fp) == (size_t)N;
    ok = (fclose(fp) == 0) && ok;
    return ok ? BNA_OK : BNA_ERR_IO;
}

// ------------------------------------------------------------------ NIfTI
// NIfTI-1 the way the reference reads it (W/CorMat/CUCorMat/CUCorMat/CorMat.cpp:
// 141-238): the first 352 bytes taken as short[176], spatial dims hdr[21..23],
// time points hdr[24], datatype hdr[35] (mask: 2 = uint8, 16 = float), bits per
// value hdr[36] (series: 32 = float, 64 = double); voxel data from byte 352 on.
namespace {
constexpr int HDR_BYTES = 352;

int read_hdr(FILE *fp, int16_t h[HDR_BYTES / 2])
{
    return fread(h, 1, HDR_BYTES, fp) == (size_t)HDR_BYTES ? BNA_OK : BNA_ERR_FORMAT;
}

int64_t file_bytes(FILE *fp)
{
    if (fseek(fp, 0, SEEK_END) != 0) return -1;
    const long n = ftell(fp);
    if (fseek(fp, HDR_BYTES, SEEK_SET) != 0) return -1;
    return n;
}
}  // namespace

extern "C" {

int bna_nifti_mask(const char *mask_path, float mask_thr, uint8_t **selected, int64_t *total_voxels, int *N)
{
    if (!mask_path || !selected || !total_voxels || !N) return BNA_ERR_ARG;
    *selected = nullptr;
    FILE *fp = fopen(mask_path, "rb");
    if (!fp) return BNA_ERR_IO;
    int16_t h[HDR_BYTES / 2];
    int rc = read_hdr(fp, h);
    const int64_t total = (int64_t)h[21] * h[22] * h[23];
    const int dt = h[35];
    const int esize = dt == 2 ? 1 : (dt == 16 ? 4 : 0);
    uint8_t *sel = nullptr;
    int64_t n = 0;
    if (rc == BNA_OK && (total <= 0 || esize == 0)) rc = BNA_ERR_FORMAT;
    if (rc == BNA_OK && file_bytes(fp) < HDR_BYTES + total * esize) rc = BNA_ERR_FORMAT;
    if (rc == BNA_OK) {
        std::vector<unsigned char> raw((size_t)(total * esize));
        sel = (uint8_t *)malloc((size_t)total);
        if (!sel) rc = BNA_ERR_NOMEM;
        else if (fread(raw.data(), 1, raw.size(), fp) != raw.size()) rc = BNA_ERR_FORMAT;
        else
            for (int64_t k = 0; k < total; k++) {
                float m;
                if (dt == 2) m = (float)raw[(size_t)k];
                else memcpy(&m, &raw[(size_t)k * 4], 4);
                sel[k] = m >= mask_thr;  // CorMat.cpp:177
                n += sel[k];
            }
    }
    fclose(fp);
    if (rc == BNA_OK && n > 0x7fffffff) rc = BNA_ERR_OVERFLOW;
    if (rc != BNA_OK) {
        free(sel);
        return rc;
    }
    *selected = sel;
    *total_voxels = total;
    *N = (int)n;
    return BNA_OK;
}

int bna_nifti_bold(const char *path, const uint8_t *selected, int64_t total_voxels, int N, float **x, int *L)
{
    if (!path || !selected || !x || !L || total_voxels <= 0 || N < 0) return BNA_ERR_ARG;
    *x = nullptr;
    FILE *fp = fopen(path, "rb");
    if (!fp) return BNA_ERR_IO;
    int16_t h[HDR_BYTES / 2];
    int rc = read_hdr(fp, h);
    const int Lr = h[24];
    const int bits = h[36];
    const int esize = bits == 32 ? 4 : (bits == 64 ? 8 : 0);
    float *out = nullptr;
    if (rc == BNA_OK && (Lr < 1 || esize == 0)) rc = BNA_ERR_FORMAT;
    if (rc == BNA_OK && (int64_t)h[21] * h[22] * h[23] != total_voxels) rc = BNA_ERR_FORMAT;
    if (rc == BNA_OK && file_bytes(fp) < HDR_BYTES + (int64_t)Lr * total_voxels * esize) rc = BNA_ERR_FORMAT;
    if (rc == BNA_OK) {
        out = (float *)malloc(sizeof(float) * (size_t)std::max<int64_t>((int64_t)Lr * N, 1));
        if (!out) rc = BNA_ERR_NOMEM;
    }
    if (rc == BNA_OK) {
        // one time point (a whole volume) at a time: BOLD[l*N + i] = In[l*total + k]
        // for the selected voxels k in ascending order (CorMat.cpp:212-219)
        std::vector<unsigned char> vol((size_t)(total_voxels * esize));
        for (int l = 0; l < Lr && rc == BNA_OK; l++) {
            if (fread(vol.data(), 1, vol.size(), fp) != vol.size()) {
                rc = BNA_ERR_FORMAT;
                break;
            }
            float *row = out + (int64_t)l * N;
            int64_t i = 0;
            for (int64_t k = 0; k < total_voxels; k++)
                if (selected[k]) {
                    if (esize == 4) {
                        float v;
                        memcpy(&v, &vol[(size_t)k * 4], 4);
                        row[i++] = v;
                    } else {
                        double v;
                        memcpy(&v, &vol[(size_t)k * 8], 8);
                        row[i++] = (float)v;
                    }
                }
            if (i != N) rc = BNA_ERR_ARG;
        }
    }
    fclose(fp);
    if (rc != BNA_OK) {
        free(out);
        return rc;
    }
    *x = out;
    *L = Lr;
    return BNA_OK;
}

// ConvertNII (W/ConvertNII/ConvertNII/ConvertNII.cpp:65-167): the mask's own
// 352-byte header with datatype 16 and 32 bits per value, then one float per
// voxel of the mask volume: the map's next value where mask >= thr (ascending
// linear voxel index, the order of bna_nifti_bold), 0 elsewhere.  The threshold
// compares in double, as the tool's ProbCut does.
int bna_nm_to_nifti(const char *mask_path, double mask_thr, const float *values, int N, const char *out_path)
{
    if (!mask_path || !values || !out_path || N < 0) return BNA_ERR_ARG;
    FILE *fp = fopen(mask_path, "rb");
    if (!fp) return BNA_ERR_IO;
    int16_t h[HDR_BYTES / 2];
    int rc = read_hdr(fp, h);
    const int64_t total = (int64_t)h[21] * h[22] * h[23];
    const int dt = h[35];
    const int esize = dt == 2 ? 1 : (dt == 16 ? 4 : 0);
    if (rc == BNA_OK && (total <= 0 || esize == 0)) rc = BNA_ERR_FORMAT;
    if (rc == BNA_OK && file_bytes(fp) < HDR_BYTES + total * esize) rc = BNA_ERR_FORMAT;
    std::vector<unsigned char> raw;
    if (rc == BNA_OK) {
        raw.resize((size_t)(total * esize));
        if (fread(raw.data(), 1, raw.size(), fp) != raw.size()) rc = BNA_ERR_FORMAT;
    }
    fclose(fp);
    if (rc != BNA_OK) return rc;
    std::vector<float> vol((size_t)total, 0.0f);
    int64_t l = 0;
    for (int64_t k = 0; k < total; k++) {
        float m;
        if (dt == 2) m = (float)raw[(size_t)k];
        else memcpy(&m, &raw[(size_t)k * 4], 4);
        if ((double)m >= mask_thr) {
            if (l >= N) return BNA_ERR_FORMAT;  // the map is shorter than the mask
            vol[(size_t)k] = values[l++];
        }
    }
    if (l != N) return BNA_ERR_FORMAT;  // the mask selects fewer voxels than the map holds
    h[35] = 16;
    h[36] = 32;
    FILE *fo = fopen(out_path, "wb");
    if (!fo) return BNA_ERR_IO;
    bool ok = fwrite(h, 1, HDR_BYTES, fo) == (size_t)HDR_BYTES && fwrite(vol.data(), 4, vol.size(), fo) == vol.size();
    ok = (fclose(fo) == 0) && ok;
    return ok ? BNA_OK : BNA_ERR_IO;
}

void bna_free(void *p) { free(p); }

}  // extern "C"

// ------------------------------------------------------------ Maslov rewiring
namespace {

// splitmix64: the random networks' generator (any seed, every output 64 bits)
struct SplitMix {
    uint64_t x;
    uint64_t next()
    {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // uniform in [0, n), n > 0 (the high half of a 64 x 64-bit product)
    int64_t below(int64_t n) { return (int64_t)(((unsigned __int128)next() * (uint64_t)n) >> 64); }
};

struct MasEdge {
    int32_t a, b;  // the endpoints (either order)
    float v;
};

}  // namespace

// Degree-preserving rewiring by double edge swaps, the reference's Rewire
// (W/Cp_GPU/CUCP/Maslov.cpp:49-98): 2 M attempts over the M undirected edges;
// an attempt draws two edges a-b and c-d, gives up when they share a vertex,
// and otherwise replaces them by a-c, b-d or by a-d, c-b (one of the two at
// random) unless one of the new edges exists already.  Edge membership is a
// hash set of vertex pairs instead of the reference's dense N x N bool matrix.
extern "C" int bna_maslov(const bna_csr *src, int permute_weights, uint64_t seed, bna_csr *out)
{
    if (!src || !out || src->on_device || src->N < 0 || !src->R || (src->nnz > 0 && !src->C)) return BNA_ERR_ARG;
    memset(out, 0, sizeof(*out));
    int rc = bna_check_csr(src, 1);
    if (rc) return rc;
    const int N = src->N;
    const int64_t nnz = src->nnz;
    auto key = [N](int64_t p, int64_t q) -> uint64_t {
        return p < q ? (uint64_t)p * (uint64_t)N + (uint64_t)q : (uint64_t)q * (uint64_t)N + (uint64_t)p;
    };
    // the upper triangle as an edge list (weights along, Maslov.cpp:117-128)
    std::vector<MasEdge> E;
    E.reserve((size_t)(nnz / 2));
    std::unordered_set<uint64_t> S;
    S.reserve((size_t)(nnz / 2) * 2 + 16);
    for (int i = 0; i < N; i++)
        for (int64_t k = src->R[i]; k < src->R[i + 1]; k++) {
            const int j = src->C[k];
            if (j == i) return BNA_ERR_FORMAT;  // a self-loop has no swap partner semantics
            if (j > i) {
                E.push_back(MasEdge{i, j, src->V ? src->V[k] : 1.0f});
                S.insert(key(i, j));
            }
        }
    const int64_t M = (int64_t)E.size();
    if (2 * M != nnz) return BNA_ERR_FORMAT;  // not symmetric
    for (int i = 0; i < N; i++)
        for (int64_t k = src->R[i]; k < src->R[i + 1]; k++)
            if (src->C[k] < i && !S.count(key(i, src->C[k]))) return BNA_ERR_FORMAT;
    SplitMix rng{seed};
    if (M >= 2)
        for (int64_t t = 0; t < 2 * M; t++) {
            const int64_t e = rng.below(M), f = rng.below(M);
            const int v1 = E[e].a, v2 = E[e].b, v3 = E[f].a, v4 = E[f].b;
            if (v1 == v3 || v1 == v4 || v2 == v3 || v2 == v4) continue;
            const bool cross = rng.next() >> 63;
            const int p = cross ? v3 : v4, q = cross ? v4 : v3;  // new edges v1-p and v2-q
            if (S.count(key(v1, p)) || S.count(key(v2, q))) continue;
            S.erase(key(v1, v2));
            S.erase(key(v3, v4));
            S.insert(key(v1, p));
            S.insert(key(v2, q));
            E[e] = MasEdge{v1, p, E[e].v};
            E[f] = MasEdge{v2, q, E[f].v};
        }
    // Maslov_weighted_2 then permutes the weights over the edges (randperm,
    // Maslov.cpp:18-30)
    if (permute_weights)
        for (int64_t i = M - 1; i > 0; i--) std::swap(E[(size_t)rng.below(i + 1)].v, E[(size_t)i].v);
    // the CSR: R unchanged (every degree kept), each edge in both rows with its
    // weight, columns ascending (Maslov.cpp:161-174 scan the bool matrix row by
    // row and hand each entry its edge's weight)
    out->R = (int32_t *)malloc(4 * (size_t)(N + 1));
    out->C = (int32_t *)malloc(4 * (size_t)std::max<int64_t>(nnz, 1));
    out->V = src->V ? (float *)malloc(4 * (size_t)std::max<int64_t>(nnz, 1)) : nullptr;
    if (!out->R || !out->C || (src->V && !out->V)) {
        bna_csr_free(out);
        memset(out, 0, sizeof(*out));
        return BNA_ERR_NOMEM;
    }
    memcpy(out->R, src->R, 4 * (size_t)(N + 1));
    std::vector<int64_t> pos(src->R, src->R + N);
    std::vector<std::pair<int32_t, float>> ent((size_t)nnz);
    for (const MasEdge &x : E) {
        ent[(size_t)pos[x.a]++] = {x.b, x.v};
        ent[(size_t)pos[x.b]++] = {x.a, x.v};
    }
    for (int i = 0; i < N; i++) {
        if (pos[i] != src->R[i + 1]) {  // cannot happen: swaps keep every degree
            bna_csr_free(out);
            memset(out, 0, sizeof(*out));
            return BNA_ERR_FORMAT;
        }
        std::sort(ent.begin() + src->R[i], ent.begin() + src->R[i + 1],
                  [](const std::pair<int32_t, float> &p, const std::pair<int32_t, float> &q) { return p.first < q.first; });
    }
    for (int64_t k = 0; k < nnz; k++) {
        out->C[k] = ent[(size_t)k].first;
        if (out->V) out->V[k] = ent[(size_t)k].second;
    }
    out->N = N;
    out->nnz = nnz;
    out->on_device = 0;
    return BNA_OK;
}
