/*
 * bna.h -- C ABI of libbna, the MI355X (gfx950) network-construction library.
 *
 * Voxel time series in; CSR graphs, degree maps and eigenvector-centrality maps
 * out.  Plain C types only: pointers, sizes, status codes.  Every entry point
 * returns an int status (BNA_OK = 0, negative on error) unless noted, never
 * exits the process, and frees what it allocated on failure.
 *
 * Reference interfaces each entry point replaces are cited as
 *   W/ = BNAPlatform-weighted-win64-cuda7.0-20160331/src
 *   B/ = BNAPlatform-win64-cuda7.0-20160115/src
 *
 * Threading: a bna_ctx is driven by one host thread at a time; calls on one
 * context are not re-entrant.  Different contexts may be used concurrently.
 */
#ifndef BNA_H
#define BNA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BNA_ABI_VERSION 1

/* ------------------------------------------------------------ status codes */
enum bna_status {
    BNA_OK = 0,
    BNA_ERR_ARG = -1,          /* bad argument (NULL, N < 2, L < 2, bad enum) */
    BNA_ERR_HIP = -2,          /* HIP runtime error (message: bna_last_error) */
    BNA_ERR_NOMEM = -3,        /* host allocation failed */
    BNA_ERR_DEVICE_NOMEM = -4, /* device allocation failed or over the HBM budget */
    BNA_ERR_IO = -5,           /* file open/read/write failed */
    BNA_ERR_FORMAT = -6,       /* malformed .csr / NIfTI / mask */
    BNA_ERR_OVERFLOW = -7,     /* result exceeds the int32 .csr format */
    BNA_ERR_NOT_FOUND = -8,    /* unknown request id */
    BNA_ERR_BUSY = -9,         /* request not finished yet */
    BNA_ERR_CANCELLED = -10,   /* request was cancelled */
    BNA_ERR_DENSITY = -11      /* density target not reachable (fewer valid pairs than k) */
};

/* --------------------------------------------------------------- options */
enum bna_layout {
    BNA_TIME_MAJOR = 0, /* x[l*N + i], the reference's BOLD buffer (CorMat.cpp:218) */
    BNA_VOXEL_MAJOR = 1 /* x[i*L + l] */
};

enum bna_precision {
    BNA_FP32 = 0,  /* FP32 MFMA, FP32 accumulate */
    BNA_BF16 = 1,  /* one BF16 product per term, FP32 accumulate */
    BNA_BF16X3 = 2 /* split BF16: hi*hi + hi*lo + lo*hi, FP32 accumulate */
};

enum bna_thr_type {
    BNA_THR_R = 0,      /* edge iff r > thr (CorMat.cpp:696) */
    BNA_THR_DENSITY = 1 /* exactly k edges, k from a percentage (CorMat.cpp:327-328, 743-822) */
};

enum bna_ec_output {
    BNA_EC_WEIGHTED_REF = 0, /* last A*y, the weighted tool's output (W/CUEC/weighted-net.cu:190) */
    BNA_EC_BINARY_REF = 1,   /* last unit-norm y, the binary tool's output (B/CUEC/CUEC/CUEC_unweighted.cu:164) */
};

typedef struct bna_opts {
    int thr_type;       /* enum bna_thr_type */
    double thr;         /* r threshold, or density in percent (the CLI's s value) */
    int weighted;       /* 1: keep r values (weighted .csr); 0: binary */
    int precision;      /* enum bna_precision */
    int recheck;        /* 1: FP64 recheck of pairs within the error band (exact edge set) */
    int ref_compat_eps; /* 1: edge iff r > thr - 1e-6 (W/CuCorMat-2mm-fuse/CorMat-2mm-fuse.cu:13,158) */
    int tile;           /* reserved, 0 */
    int refine;         /* weighted: recompute every kept value in double after the
                           edge set is fixed, so |V - r64| <= half a float ulp + 1e-15 */
} bna_opts;

/* Fills the defaults: r threshold 0, weighted, BF16X3, recheck on, refine on. */
void bna_opts_default(bna_opts *o);

/* A CSR graph.  R/C/V are host arrays when on_device == 0 and device pointers
 * otherwise; the library allocates them (bna_csr_free releases).  V is NULL for
 * a binary graph.  The layout is the reference's: symmetric, no diagonal, rows
 * ascending, columns ascending within a row (CorMat.cpp:889-918). */
typedef struct bna_csr {
    int32_t N;
    int64_t nnz;   /* directed entries, 2 x undirected edges (the file's Clength) */
    int32_t *R;    /* N + 1 */
    int32_t *C;    /* nnz */
    float *V;      /* nnz or NULL */
    int on_device; /* 0 host, 1 device */
    double thr_used;      /* r threshold applied; density mode: the k-th largest r */
    double density_pct;   /* 100 * edges / M1, the reference's s value (CorMat.cpp:321) */
    int64_t n_candidates; /* pairs inside the error band, rechecked in FP64 */
    int64_t n_flips;      /* band pairs whose FP64 decision differs from the MFMA one */
    double band;          /* error band half-width used */
    double max_err;       /* max |r_mfma - r64| over rechecked pairs */
} bna_csr;

void bna_csr_free(bna_csr *g);

/* Per-call timings of the last build, milliseconds, measured with HIP events on
 * the context's stream. */
typedef struct bna_timings {
    double znorm_ms;
    double corr_ms;    /* MFMA tile kernel(s), all passes */
    double recheck_ms;
    double csr_ms;     /* fragment tables to rows, row lengths, scan and placement */
    double refine_ms;  /* value refinement */
    double select_ms;  /* density selection */
    double total_ms;   /* first to last event of the build */
    int64_t corr_launches;
    double degree_ms;  /* last bna_degree: the degree kernel */
    double ec_ms;      /* last bna_eigcent: the whole power iteration on the device */
    int64_t ec_iters;  /* last bna_eigcent: iterations */
    double cp_ms;      /* last bna_cp: weight pass and row kernel */
    double modularity_ms;      /* last bna_modularity: the whole partition */
    int64_t modularity_iters;  /* last bna_modularity: power iterations over all rounds */
    double lp_ms;              /* last bna_lp: every search on the device */
    double bc_ms;              /* last bna_bc: every source's search and dependencies */
    double lpw_ms;             /* last bna_lp_weighted: cost matrix, Floyd-Warshall and row sums */
} bna_timings;

/* ----------------------------------------------------------------- context */
typedef struct bna_ctx bna_ctx;

/* device: HIP ordinal; hbm_budget: bytes the context may hold (0 = 90% of free). */
int bna_create(bna_ctx **ctx, int device, size_t hbm_budget);
int bna_destroy(bna_ctx *ctx);
int bna_synchronize(bna_ctx *ctx);
/* Frees a device CSR on the context's stream, ordered after the context's work
 * (bna_csr_free orders device frees on the null stream instead). */
int bna_csr_release(bna_ctx *ctx, bna_csr *g);
const char *bna_status_string(int status);
/* Last error message recorded on this context (thread-local when ctx is NULL). */
const char *bna_last_error(bna_ctx *ctx);
int bna_get_timings(bna_ctx *ctx, bna_timings *t);
/* Device memory held by the context's workspace, bytes. */
int64_t bna_workspace_bytes(bna_ctx *ctx);

/* ------------------------------------------------------ construction path */
/* Host-buffer drop-in for one subject: normalise, correlate, threshold and
 * build the CSR.  Replaces CorMat_gpu (W/CorMat/CUCorMat/CUCorMat/CorMat.cu:17,
 * called at CorMat.cpp:259) together with FormFullGraph / FormFullGraph_s and
 * FormCSRGraph (CorMat.cpp:683-918); the result lands in host arrays. */
int bna_build_csr(bna_ctx *ctx, const float *x, int N, int L, int layout, const bna_opts *opts,
                  bna_csr *out);

/* Device-resident form: the series is uploaded (or taken from device memory)
 * and normalised once; builds leave the CSR in device memory. */
typedef struct bna_series bna_series;
int bna_series_create(bna_ctx *ctx, const float *x, int N, int L, int layout, int x_on_device,
                      bna_series **out);
int bna_series_destroy(bna_ctx *ctx, bna_series *s);
/* Normalisation only (z-score, kernel a).  Asynchronous: it is queued on the
 * context's stream and a build queued after it waits for it; its time is
 * read into bna_timings.znorm_ms at the next bna_get_timings. */
int bna_series_normalize(bna_ctx *ctx, bna_series *s, int precision);
/* Correlate + threshold + CSR on a normalised series; out->on_device = 1. */
int bna_build_csr_series(bna_ctx *ctx, bna_series *s, const bna_opts *opts, bna_csr *out);
/* Caller-owned buffers, the reference's calling convention (CorMat.cpp:250,
 * caller-allocated output): builds as bna_build_csr_series and copies the CSR
 * into R (N + 1 ints), C and V (cap entries each; V may be NULL for a binary
 * graph); *nnz gets the entry count.  R == NULL is a size query (nothing is
 * copied; a later call builds again).  cap < nnz: BNA_ERR_OVERFLOW, nothing
 * copied.  A density's nnz is known beforehand: 2 * bna_density_to_k(N, s). */
int bna_build_csr_into(bna_ctx *ctx, bna_series *s, const bna_opts *opts, int32_t *R, int32_t *C, float *V,
                       int64_t cap, int64_t *nnz);
/* Several subjects in one pass of batched launches (the batched form of the
 * reference's per-subject loop, CorMat.cpp:195-384): every subject's tiles go
 * through one histogram launch (density subjects) and one emission launch of
 * the tile kernel, then each subject's recheck, selection and CSR steps.
 * opts[q] per subject, one precision for the batch; out[q] gets a device CSR
 * and status[q] (may be NULL) that subject's status.  Returns BNA_OK when every
 * subject succeeded, else the first failure. */
int bna_build_csr_batch(bna_ctx *ctx, bna_series *const *series, const bna_opts *opts, int n, bna_csr *out,
                        int *status);
/* Releases the context's device workspace (it regrows on the next build). */
int bna_trim(bna_ctx *ctx);
/* Copies a device CSR to host arrays (out gets fresh host buffers). */
int bna_csr_to_host(bna_ctx *ctx, const bna_csr *dev, bna_csr *out);
/* Normalised series back to the host, for tests: z as double (N x L,
 * voxel-major), mean and norm per voxel (each may be NULL). */
int bna_series_download(bna_ctx *ctx, const bna_series *s, double *z64, double *mu, double *sd);

/* ------------------------------------------------------- group averaging */
/* The reference's group mode (avg = y or b, CorMat.cpp:418-670): every
 * subject's r, or its Fisher z (fishertrans, CorMat.cpp:36-42), is added to a
 * per-pair sum if it lies in [-1, 1] (post_block, :824-887; a NaN from an
 * invalid voxel adds nothing); the sums are divided by the subject count and
 * Fisher sums mapped back (inv_fishertrans, :44-50, :548-558); the averages are
 * thresholded like one subject's matrix (FormFullGraph strict r > thr, :683-703,
 * or the exact-k FormFullGraph_s with its tie rule, :743-822) and listed as CSR
 * rows in column order with V = the average (FormCSRGraph, :889-918).  The sums
 * are a packed upper triangle of M1 = N(N-1)/2 floats, row-major with i < j
 * (4 GB at N = 45000), held in device memory. */
typedef struct bna_group bna_group;
int bna_group_create(bna_ctx *ctx, int N, int fisher, bna_group **out);
int bna_group_destroy(bna_ctx *ctx, bna_group *g);
/* Exact averages (the default, on): the group keeps every subject's normalised
 * series (4 N L_pad bytes each), and bna_group_build_csr recomputes from FP64 r,
 * in the reference's float arithmetic, every pair whose average the MFMA error
 * bound leaves on either side of the threshold, and every kept edge's value;
 * the edge set and values are then the FP64 oracle's (plain means bit for bit,
 * Fisher means up to the ulp of logf / expf).  Off: averages carry the MFMA
 * error of their terms.  No counterpart in the reference, whose averages are
 * FP32 SGEMM sums; callers that cannot afford the copies turn it off.  Only
 * before the first bna_group_add. */
int bna_group_set_exact(bna_ctx *ctx, bna_group *g, int exact);
/* Adds one subject (the correlation pass of CorMat_gpu, CorMat.cpp:489, and
 * post_block, :501-504).  The series is normalised for `precision` if needed;
 * every subject must have the group's N, L may differ. */
int bna_group_add(bna_ctx *ctx, bna_group *g, bna_series *s, int precision);
/* Subjects added so far. */
int bna_group_count(const bna_group *g);
/* Pairs the last bna_group_build_csr recomputed at the threshold (exact groups). */
int64_t bna_group_band_pairs(const bna_group *g);
/* Averages the sums (CorMat.cpp:548-558); no subject can be added afterwards. */
int bna_group_finish(bna_ctx *ctx, bna_group *g);
/* The packed averages (or sums, before bna_group_finish): M1 floats, the
 * payload of the .cormat file (CorMat.cpp:570-579). */
int bna_group_download(bna_ctx *ctx, const bna_group *g, float *packed);
/* Thresholds the averages: opts->thr_type, thr and weighted are used.  out gets
 * a device CSR; thr_used is the r threshold or, for a density, the k-th
 * largest average.  The context keeps an N x N byte keep matrix (rows padded to
 * 64 bytes; 2 GB at N = 45000) in its workspace until bna_trim. */
int bna_group_build_csr(bna_ctx *ctx, bna_group *g, const bna_opts *opts, bna_csr *out);

/* The packed r matrix of one subject, M1 floats in row-major upper-triangle
 * order with 0 for pairs with an invalid voxel: what CorMat_gpu followed by
 * post_block leaves (CorMat.cu:17, CorMat.cpp:259-262), the payload of a
 * subject's .cormat.  r carries the precision's MFMA error; a value just outside
 * [-1, 1] (near-identical series) is clamped to it, where post_block would drop
 * the reference's own out-of-range SGEMM values. */
int bna_corr_packed(bna_ctx *ctx, bna_series *s, int precision, float *packed);
/* The r threshold of a density: the k-th largest r of the subject, with k from
 * bna_density_to_k, as CorMat_spa2rth returns it
 * (CorMat-2mm-transition-GPUhisto.cu:70-214) but exact (FP64 recheck of the
 * pairs at the boundary).  k may be NULL. */
int bna_select_density(bna_ctx *ctx, bna_series *s, int precision, double density_pct, double *r_star, int64_t *k);

/* ------------------------------------------------------------ graph maps */
/* Degree.  Weighted: float sum of V along each row in CSR order
 * (W/Degree/Degree/Degree.cpp:83-87); binary or V == NULL: row length as float
 * (B/Degree/Degree/Degree.cpp:84-91).  g may be host or device; deg is host, N floats. */
int bna_degree(bna_ctx *ctx, const bna_csr *g, float *deg);

/* Eigenvector centrality by FP64 power iteration with the reference's loop:
 * start x[i] = i, stop when ||x - x0|| <= tol or ||x + x0|| <= tol or after
 * max_iter iterations (W/CUEC/weighted-net.cu:55-221, Lead_Vector_GPU).
 * output: enum bna_ec_output.  ec is host, N doubles; lambda gets the last norm
 * (the value written to _evalue.txt); iters may be NULL.  A binary graph
 * (V == NULL) uses weight 1 for every entry. */
int bna_eigcent(bna_ctx *ctx, const bna_csr *g, double tol, int max_iter, int output, double *ec,
                double *lambda, int *iters);

/* Newman leading-eigenvector modularity, the CUmodularity tool
 * (B/Newman_Modularity_GPU/CUmodularity/modularity_GPU.cu:408-528 Partition_GPU,
 * :332-400 Sub_Partition_GPU, Modularity_CPU.cpp:176-272 Sub_Partition and
 * Lead_Vector): repeated bisection by the sign of the leading eigenvector of
 * the group's generalised modularity matrix (power iteration, eps 1e-6, at
 * most 10000 iterations; a negative leading eigenvalue is shifted out), a
 * division kept when the eigenvalue exceeds 0.01 and each side has two or more
 * nodes.  labels (host, N ints) gets each voxel's module, numbered from 1 in
 * the order the modules are found final (the _modu.nm values); Q the
 * modularity of the partition; n_modules (may be NULL) the module count.  The
 * graph is taken as binary (V is not used), as the reference tool reads it.
 * The power iteration starts from v_i = i (the weighted tree's Lead_Vector_GPU,
 * W/.../modularity_GPU.cu:243); the binary tool starts from rand() seeded by
 * the clock, which no run can repeat.  g may be host or device. */
int bna_modularity(bna_ctx *ctx, const bna_csr *g, int32_t *labels, double *Q, int *n_modules);

/* Characteristic path length of the binary graph, the CUBFS_Lp tool
 * (B/src/Lp/CUBFS_Lp/CUBFS_Lp/CUBFS_LP.cpp:14-53 CUBFS_Lp, BFS_APSP_GPU.cu:7-50
 * APSP_BFS_node_kernel): d(i, j) = hops of a shortest path from i to j along
 * the CSR rows, then Lp = 1 / (sum over ordered pairs with d > 0 of
 * (double)(1.0f / d), divided by N and by N - 1); pairs with no path add
 * nothing (an edgeless graph gives +inf).  hist (may be NULL, hist_len entries)
 * gets the number of ordered pairs at each distance d = 1 .. hist_len - 1, and
 * in hist[0] the ordered pairs with no path.  V is not used.  The reference
 * adds the N^2 reciprocals one by one; here they are added per distance, so Lp
 * agrees with it to the rounding of that sum (about 1e-15 relative).  g may be
 * host or device. */
int bna_lp(bna_ctx *ctx, const bna_csr *g, double *lp, int64_t *hist, int hist_len);
/* Betweenness centrality of the binary graph by Brandes' dependencies, the CUBC
 * tool (B/src/BC/CUBC/CUBC.cpp:13-125, BC_GPU_node_pred.cu): bc[v] = sum over
 * sources s != v of delta_s(v), delta_s(v) = sum over the successors w of v of
 * sigma_s(v) / sigma_s(w) * (1 + delta_s(w)), sigma the shortest-path counts.
 * Ordered source-target pairs are counted, as the reference does (an
 * undirected graph's values are twice the unordered ones); the _bc.nm values.
 * Path counts and dependencies are double (the reference's are float), summed
 * in source order; bc is rounded to float at the end.  V is not used.  g may be
 * host or device. */
int bna_bc(bna_ctx *ctx, const bna_csr *g, float *bc);
/* Characteristic path length and nodal efficiency of the weighted graph, the
 * CUBFW_Lp tool (W/src/Lp/CUBFW_Lp/CUBFW_Lp/CUBFW_Lp.cpp:13-162 with
 * init_block, FormBlock.cpp:45-75): the cost of edge (i, j) is 1.0f / w (1 when
 * g has no V), d(i, j) the float shortest-path cost by blocked Floyd-Warshall
 * over the dense N x N matrix (128-vertex blocks, the whole matrix in device
 * memory: about 4 N^2 bytes, BNA_ERR_DEVICE_NOMEM past the budget), and over
 * the ordered pairs i != j with d not below 0.1, 1.0f / d is summed (pairs with
 * no path add 0).  lp = 1 / (sum / N / (N - 1)) (the _Lp.txt value; row sums in
 * double, added in row order).  eff[i] (optional, N floats, the _eff.nm map) =
 * the float sum of row i in column order divided by N - 1, set to 0 when below
 * 1.0f / N; isolated (optional) = the number of such vertices.  Weights must be
 * non-negative (BNA_ERR_ARG for a negative or NaN weight; a zero weight is an
 * infinite cost).  Distances are the blocked recurrence's, so they can differ
 * from the reference's 3072-vertex blocking in the last bit of a float sum.
 * g may be host or device. */
int bna_lp_weighted(bna_ctx *ctx, const bna_csr *g, double *lp, float *eff, int *isolated);

/* No counterpart in the reference: bna_eigcent for callers that cannot
 * allocate device memory (serving loops, harnesses that forbid hipMalloc).
 * g must be a device CSR (on_device = 1, e.g. from bna_build_csr_series); ec is
 * a DEVICE array of N doubles; lambda and iters as bna_eigcent.  Its SpMV is
 * the plain L2-gather kernel (bna_eigcent copies a graph with 16 or more
 * entries per row into a blocked layout first, which needs memory), so the
 * results agree with bna_eigcent's to FP64 rounding, with the same stopping
 * rule.  The iterate vectors, partial sums and stopping state live in a static
 * device array of the library (about 16.4 MiB per GPU, held until the process
 * ends), so N is at most BNA_NOALLOC_EC_MAX_N (BNA_ERR_ARG
 * beyond it) and no device or pinned memory is allocated.  The array keeps the
 * last call's contents; every call initialises what it reads.  Calls on one GPU
 * share the array and would race: the library serialises them (a lock per GPU),
 * so overlapping calls wait.  Returns once ec is written (it synchronises the
 * context's stream). */
#define BNA_NOALLOC_EC_MAX_N 1048576
int bna_eigcent_noalloc(bna_ctx *ctx, const bna_csr *g, double tol, int max_iter, int output, double *ec,
                        double *lambda, int *iters);

/* Weighted clustering coefficient, the CUCP tool's Cp_2 (W/Cp_GPU/CUCP/Cp.cu:111-471):
 * w = cube root of V / V[m] (m the first entry of largest |V|), cci_i = (W^3)_ii,
 * Cp_i = cci_i / (t (t - 1)) with the reference's float t = (float)R[i+1] -
 * (float)R[i] (0 when t < 2).  A binary graph (V == NULL) uses w = 1.  cci is
 * summed in double.  cp is host, N floats; mean (may be NULL) gets the average
 * |Cp| as the float the tool prints to <stem>_cp.txt (Cp_main.cpp, sasum / N).
 * Columns must ascend strictly within each row (a host CSR that breaks this is
 * BNA_ERR_FORMAT, see bna_check_csr); N at most 524288. */
int bna_cp(bna_ctx *ctx, const bna_csr *g, float *cp, double *mean);

/* -------------------------------------------------------- host utilities */
/* Number of undirected edges for a density percentage, with the reference's
 * float arithmetic (CorMat.cpp:327-328). */
int64_t bna_density_to_k(int N, float spa_percent);

/* Seeded factor-model series: x[v,t] = sum_k a[v,k] f[k,t] + e[v,t] with K AR(1)
 * latent signals (coefficient phi, unit variance), each voxel loading on two
 * distinct factors with a ~ N(0, 0.5^2), e ~ N(0, 1).  Deterministic for a seed
 * whatever the thread count. */
int bna_synth_factor_model(float *x, int N, int L, int K, double phi, uint64_t seed, int layout);

/* .csr files (W/CorMat/CUCorMat/CUCorMat/CorMat.cpp:353-360; binary
 * B/CorMat/CUCorMat/CUCORMAT/CorMat.cpp:350-353): int32 N+1, int32 R[N+1],
 * int32 nnz, int32 C[nnz] and, weighted, int32 nnz, float V[nnz]. */
int bna_write_csr(const char *path, const bna_csr *g);
/* Reads a .csr (weighted when a V section follows); out gets host arrays. */
int bna_read_csr(const char *path, bna_csr *out);
/* Structural check of a host CSR: R from 0 to nnz without decreasing, columns
 * in [0, N), and with ascending != 0 columns strictly ascending within every
 * row (what FormCSRGraph writes, CorMat.cpp:889-918).  BNA_OK or
 * BNA_ERR_FORMAT; BNA_ERR_ARG for a device CSR.  Needs no device. */
int bna_check_csr(const bna_csr *g, int ascending);
/* Degree-preserving random network of a host CSR, the Maslov rewiring behind the
 * Cp and Lp tools' random-network comparisons (W/Cp_GPU/CUCP/Maslov.cpp:49-289,
 * W/Maslov/Maslov/Maslov.cpp; B/.../CUBFS_LP_main.cpp:110-116 for binary
 * graphs): 2 x edges attempts at a double edge swap (a-b, c-d -> a-c, b-d or
 * a-d, c-b, refused when the vertices are not distinct or a new edge exists).
 * Weights travel with their edges (Maslov_weighted_1); permute_weights != 0
 * then shuffles them over the edges (Maslov_weighted_2).  R is unchanged and
 * each row's columns ascend.  The reference seeds rand() from the clock; seed
 * drives a splitmix64 generator here, so a network can be made again.  out gets
 * host arrays (bna_csr_free).  The input must be symmetric without self-loops,
 * columns ascending (BNA_ERR_FORMAT otherwise).  Needs no device. */
int bna_maslov(const bna_csr *src, int permute_weights, uint64_t seed, bna_csr *out);
/* .nm node maps (Degree.cpp:95-99, weighted-net.cu:305-322): int32 N, float[N]. */
int bna_write_nm(const char *path, const float *v, int N);
/* A vertex map (N floats, e.g. a _deg.nm or _ec.nm payload) as a NIfTI volume
 * on the mask's grid, the ConvertNII tool (W/ConvertNII/ConvertNII/
 * ConvertNII.cpp:65-167): the mask's header with float data, the map's values at
 * the voxels with mask >= mask_thr in ascending order, 0 elsewhere.
 * BNA_ERR_FORMAT if the mask selects a different number of voxels than N. */
int bna_nm_to_nifti(const char *mask_path, double mask_thr, const float *values, int N, const char *out_path);
/* .cormat (CorMat.cpp:301-310): int32 low 32 bits of M1 = N(N-1)/2, then the
 * M1 packed floats (bna_group_download's layout). */
int bna_write_cormat(const char *path, const float *packed, int N);

/* NIfTI-1 ingest as CUCorMat does it (W/CorMat/CUCorMat/CUCorMat/CorMat.cpp:141-238):
 * the 352-byte header read as short[176] (dims hdr[21..23], time points hdr[24],
 * datatype hdr[35], bits hdr[36]), data from byte 352.
 * bna_nifti_mask: mask of uint8 (datatype 2) or float (16); voxel k is selected
 * iff mask[k] >= mask_thr (CorMat.cpp:177).  selected gets total_voxels bytes.
 * bna_nifti_bold: float (32-bit) or double (64-bit) series of the same volume
 * size; x gets the selected voxels in ascending order, time-major x[l*N + i]
 * (CorMat.cpp:212-219; doubles narrowed to float).  Stricter than the
 * reference: a short file or a volume-size mismatch is BNA_ERR_FORMAT.
 * Buffers are released with bna_free. */
int bna_nifti_mask(const char *mask_path, float mask_thr, uint8_t **selected, int64_t *total_voxels, int *N);
int bna_nifti_bold(const char *path, const uint8_t *selected, int64_t total_voxels, int N, float **x, int *L);
void bna_free(void *p);

/* ------------------------------------------------------------- serving */
/* A request loop for many subjects on one GPU (no counterpart in the
 * reference, which builds one subject at a time).  Requests are admitted in
 * arrival order while their estimated device footprint fits the budget and are
 * built in rounds (one bna_build_csr_batch per round) by worker threads that
 * each own a context; rounds of different workers overlap on the device. */
typedef struct bna_server bna_server;

typedef struct bna_server_config {
    int device;
    size_t hbm_budget; /* bytes for admitted requests; 0 = 80 % of free device memory */
    int workers;       /* rounds in flight at once (contexts), 1..8 */
    int max_batch;     /* requests per round */
    size_t reserve_per_worker; /* bytes each worker's pool takes at start (0: none) */
    double big_round_ms; /* rounds whose estimated tile-kernel time exceeds this run in
                            pieces that yield to small rounds (0: never) */
    double chunk_ms;     /* piece length for big rounds */
    int reserve_eighths; /* while a big round runs, its tile launches leave this many
                            eighths of the CUs (one XCD each on MI355X) to other rounds */
} bna_server_config;

typedef struct bna_request {
    const float *x; /* copied at submit */
    int N, L, layout;
    bna_opts opts;
    int want_degree, want_ec; /* EC: tol 1e-6, 10000 iterations, the reference output of the graph's kind */
    int device_result;        /* 1: the CSR stays in device memory (result.csr.on_device = 1) */
} bna_request;

typedef struct bna_result {
    int status;
    char error[256];
    bna_csr csr;      /* host arrays, or device arrays for device_result */
    float *degree;    /* N floats or NULL */
    double *ec;       /* N doubles or NULL */
    double lambda;
    int ec_iters;
    int batch_size;   /* requests in the round it ran in */
    double queue_ms;  /* submit -> round start */
    double run_ms;    /* round start -> result ready */
    double latency_ms;/* submit -> result ready */
    double upload_ms, build_ms, maps_ms, download_ms; /* phases of its round (build: the whole batch) */
    bna_timings build; /* the batch's device timings */
} bna_result;

typedef struct bna_server_stats {
    int64_t submitted, completed, failed, cancelled, rounds, queued;
    int64_t max_concurrent_rounds, max_batch_seen;
    int64_t max_reserved, budget; /* footprint bytes admitted at once, and the budget */
} bna_server_stats;

void bna_server_config_default(bna_server_config *c);
int bna_server_create(const bna_server_config *cfg, bna_server **out);
/* Cancels queued requests, waits for running rounds, frees uncollected results. */
int bna_server_destroy(bna_server *srv);
int bna_submit(bna_server *srv, const bna_request *req, int64_t *id);
/* BNA_ERR_BUSY until the request is finished; then moves the result to out
 * (release with bna_result_free) and forgets the id; returns its status. */
int bna_poll(bna_server *srv, int64_t id, bna_result *out);
int bna_wait(bna_server *srv, int64_t id, bna_result *out);
/* Removes a queued request (BNA_ERR_BUSY once it runs). */
int bna_cancel(bna_server *srv, int64_t id);
int bna_server_get_stats(bna_server *srv, bna_server_stats *out);
void bna_result_free(bna_result *r);
/* Estimated device bytes of one request: series, operand panel, tables, edge
 * lists and CSR (edges: the density target, or 2 % of pairs for an r threshold). */
int64_t bna_request_footprint(int N, int L, const bna_opts *o);
/* Long kernels of later builds (tile passes, value refinement) are issued in
 * pieces of about chunk_ms; with a gate, each piece waits for the previous one
 * to finish and then for gate(gate_arg) (which may block, e.g. while a small
 * request runs on another context); without one, two pieces stay in flight.
 * chunk_ms = 0 restores single launches. */
int bna_set_chunking(bna_ctx *ctx, double chunk_ms, void (*gate)(void *), void *gate_arg);
/* Caps the workgroups of later correlation launches (the persistent tile kernel
 * runs one per CU) at max_workgroups, rounded down to a multiple of 8; 0 lifts
 * the cap.  A server gives long builds fewer CUs than the device has, so short
 * requests on other contexts find CUs free.  No counterpart in the reference,
 * which runs one subject at a time. */
int bna_set_grid_limit(bna_ctx *ctx, int max_workgroups);
/* HIP graphs on (default) or off for later calls: the centrality's iteration
 * batches, and a batched build's emission launch with every subject's tails
 * (recheck, selection, row counts, scan, counter readbacks) captured as one
 * graph and replayed while its launch parameters stay the same.  No
 * counterpart in the reference, which launches kernels one by one. */
int bna_set_graphs(bna_ctx *ctx, int on);
/* How often the batched build captured its graph and replayed it. */
int bna_batch_graph_stats(bna_ctx *ctx, int64_t *captures, int64_t *replays);
/* ---------------------------------------------------------------- autotuner
 * No counterpart in the reference (its cuBLAS tiles are fixed at B = 2048).
 * The tile kernel walks its upper-triangle tile table in bands of G tile-rows
 * per tile-column, each XCD over its own contiguous range; G (the traversal
 * shape) decides which operand panels share an XCD's L2.  It is tuned per
 * precision and per series-length bucket (padded L <= 192, <= 320, <= 640,
 * longer): bna_autotune builds a synthetic N-voxel subject of length L with
 * G in {4, 8, 16, 32}, keeps the fastest tile-kernel time, and rewrites its line
 * of cache_path (NULL: no file).  The file holds "<precision> <bucket> <G> <ms>"
 * lines; bna_autotune_load applies one to a context.  Default G = 8. */
int bna_autotune(bna_ctx *ctx, int precision, int L, int N, const char *cache_path, int *band_rows,
                 double *best_ms);
int bna_autotune_load(bna_ctx *ctx, const char *path);
int bna_autotune_get(bna_ctx *ctx, int precision, int L, int *band_rows);
int bna_autotune_set(bna_ctx *ctx, int precision, int L, int band_rows);
/* Workgroups of an uncapped correlation launch: the device's CUs rounded up to a
 * multiple of 8 (-1 for a NULL context).  No counterpart in the reference. */
int bna_grid_size(bna_ctx *ctx);
/* Grows the context's memory pool by `bytes` now (allocated and returned to the
 * pool), so later builds draw on it without asking the driver for memory, and
 * sets up the pinned staging buffers; bytes = 0 does only the latter. */
int bna_reserve(bna_ctx *ctx, size_t bytes);
/* Free and total device memory of the context's device. */
int bna_memory_info(bna_ctx *ctx, int64_t *free_bytes, int64_t *total_bytes);

#ifdef __cplusplus
}
#endif

#endif /* BNA_H */
