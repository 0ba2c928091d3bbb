/*
 * bna_oracle.c -- CPU restatement of BNAPlatform's network-construction path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the GPU library is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg load it.  The product library never links or calls it.
 *
 * Reference paths below are relative to the reference tree:
 *   W/ = BNAPlatform-weighted-win64-cuda7.0-20160331/src
 *   B/ = BNAPlatform-win64-cuda7.0-20160115/src
 *
 * Contents and what each one follows:
 *   orc_znorm           W/CorMat/CUCorMat/CUCorMat/CorMat.cu:33-52 (mean and
 *                       norm in double, z = (x - mean) / norm), kept in double
 *                       (z64) and narrowed to float (z32) as the reference stores.
 *   orc_dot64           The FP64 oracle's Pearson r: sum_l z64_i[l] * z64_j[l]
 *                       in a fixed evaluation order (64 strided fused
 *                       multiply-add partial sums, then a halving tree).  The GPU
 *                       recheck evaluates the same order, so both give the same
 *                       double bit for bit.
 *   orc_ref_corr        The reference CPU path W/CorMat/CUCorMat/CUCorMat/
 *                       CorMat.cpp:954-965: float z products accumulated in
 *                       double, stored as float.
 *   orc_build_threshold FormFullGraph (CorMat.cpp:683-703, strict r > thr) followed
 *                       by FormCSRGraph (CorMat.cpp:889-918): rows in ascending
 *                       order, columns ascending, V = r of the pair.
 *   orc_build_density   FormFullGraph_s (CorMat.cpp:743-822): k-th largest r,
 *                       accept r >= r*, LIFO tie trimming in row-major
 *                       upper-triangle order.  Both the literal stack loop and the
 *                       closed form (first m ties in row-major order) are here so
 *                       tests can pin one against the other.
 *   orc_density_k       CorMat.cpp:327-328: Clength from the sparsity percentage,
 *                       with the reference's float arithmetic.
 *   orc_degree          W/Degree/Degree/Degree.cpp:83-87 (float sum of V in CSR
 *                       order); binary B/Degree/Degree/Degree.cpp:84-91.
 *   orc_eigcent         W/CUEC/weighted-net.cu:129-190 (weighted, output A*y) and
 *                       B/CUEC/CUEC/CUEC_unweighted.cu:156-195 (binary, output y).
 *   orc_cp              Cp_2 (W/Cp_GPU/CUCP/Cp.cu:111-471): cube-rooted weights
 *                       scaled by the largest |V|, (W^3)_ii over t (t - 1) with
 *                       the reference's float t = (float)R[i+1] - (float)R[i];
 *                       cci summed in double over a dense scatter of row i.
 *   orc_fisher,         fishertrans / inv_fishertrans (CorMat.cpp:36-50) in the
 *   orc_inv_fisher      reference's float arithmetic (float quotient, logf/expf).
 *   orc_group_add       post_block (CorMat.cpp:824-887): each pair's r, as float,
 *                       is added (or its Fisher z) when it lies in [-1, 1].
 *   orc_group_finish    CorMat.cpp:548-558: divide by the subject count in float,
 *                       then inv_fishertrans for Fisher sums.
 *   orc_threshold_packed FormFullGraph over packed values (strict v > thr) and
 *                       FormCSRGraph, for the averaged matrix.
 *   orc_modularity      Newman leading-eigenvector modularity: B/Newman_Modularity_GPU/
 *                       CUmodularity/modularity_GPU.cu:408-528 (Partition_GPU, the
 *                       round loop and Q), :332-400 (Sub_Partition_GPU) and
 *                       Modularity_CPU.cpp:176-320 (Sub_Partition, Lead_Vector: the
 *                       power iteration on the generalised modularity matrix).
 *                       The start vector is the node's index, as
 *                       W/Newman_Modularity_GPU/CUmodularity/modularity_GPU.cu:243
 *                       sets it (the binary tool draws rand() seeded by time(NULL),
 *                       which no run can reproduce).
 *   orc_lp              CUBFS_Lp (B/src/Lp/CUBFS_Lp/CUBFS_Lp/CUBFS_LP.cpp:14-53):
 *                       BFS hop distances from every vertex, then the N^2
 *                       reciprocals 1.0f / d (d > 0) added one by one in double in
 *                       row-major order, divided by N and N - 1, inverted; also
 *                       the count of ordered pairs per distance.
 *   orc_bc              CUBC's Brandes betweenness (B/src/BC/CUBC/BC_GPU_node_pred.cu):
 *                       per source, path counts sigma and dependencies delta
 *                       (pulled from each vertex's successors in CSR order), in
 *                       double, added over sources in order.
 *   orc_lpw             CUBFW_Lp (W/src/Lp/CUBFW_Lp/CUBFW_Lp/CUBFW_Lp.cpp:13-162,
 *                       init_block FormBlock.cpp:45-75): the dense cost matrix
 *                       (+inf, 1.0f / w per CSR entry in CSR order, 0 on the
 *                       diagonal), blocked Floyd-Warshall with block size B (the
 *                       diagonal block, then the row and column panels against
 *                       the finished diagonal block, then every other block as
 *                       a min-plus product of the two panels), then per row, in
 *                       column order, the float sum of 1.0f / d over d not below
 *                       0.1 (nodal efficiency, divided by N - 1 and zeroed below
 *                       1.0f / N) and the same terms in double; the rows' double
 *                       sums added in row order give Lp.  Also the distances.
 *   orc_sssp64          Dijkstra from every vertex over the same float costs
 *                       with path sums in double: the check of orc_lpw's
 *                       distances (equal on costs whose sums are exact).
 *   orc_ref_pipeline    The reference's CPU construction path end to end
 *                       (normalise, CorMat_cpu correlation, FormFullGraph(_s),
 *                       FormCSRGraph, degree, power iteration), timed by stage:
 *                       the CPU baseline, not a parity function.
 *
 * Pairs involving a voxel whose z has a non-finite entry (a constant series gives
 * 0/0) are never edges and take no part in density selection.
 *
 * Build: oracle/Makefile (gcc -O3 -mavx2 -mfma -ffp-contract=off).  Explicit
 * fma() calls are the only fused operations; nothing is contracted implicitly.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>
#include <time.h>

int orc_cpu_has_fma(void) { return __builtin_cpu_supports("fma") && __builtin_cpu_supports("avx2"); }

/* ------------------------------------------------------------------ z-norm */

/* layout 0: time-major x[l*N + i] (the reference's BOLD); 1: voxel-major x[i*L + l].
 * Outputs are voxel-major N x L.  mu/sd may be NULL; z32/valid may be NULL. */
void orc_znorm(const float *x, int N, int L, int layout, double *mu, double *sd,
               double *z64, float *z32, unsigned char *valid)
{
    for (int i = 0; i < N; i++) {
        double s1 = 0.0, s2 = 0.0;
        for (int l = 0; l < L; l++) {
            double v = layout ? x[(int64_t)i * L + l] : x[(int64_t)l * N + i];
            s1 += v;
        }
        s1 /= L;
        for (int l = 0; l < L; l++) {
            double v = layout ? x[(int64_t)i * L + l] : x[(int64_t)l * N + i];
            double d = v - s1;
            double p = d * d;
            s2 += p;
        }
        s2 = sqrt(s2);
        int ok = 1;
        for (int l = 0; l < L; l++) {
            double v = layout ? x[(int64_t)i * L + l] : x[(int64_t)l * N + i];
            double z = (v - s1) / s2;
            if (!isfinite(z)) ok = 0;
            z64[(int64_t)i * L + l] = z;
            if (z32) z32[(int64_t)i * L + l] = (float)z;
        }
        if (mu) mu[i] = s1;
        if (sd) sd[i] = s2;
        if (valid) valid[i] = (unsigned char)ok;
    }
}

/* ------------------------------------------------------------- FP64 r order */

double orc_dot64(const double *a, const double *b, int L)
{
    double p[64];
    for (int q = 0; q < 64; q++) p[q] = 0.0;
    int l0 = 0;
    for (; l0 + 64 <= L; l0 += 64)
        for (int q = 0; q < 64; q++) p[q] = fma(a[l0 + q], b[l0 + q], p[q]);
    for (int q = 0; l0 + q < L; q++) p[q] = fma(a[l0 + q], b[l0 + q], p[q]);
    for (int s = 32; s >= 1; s >>= 1)
        for (int q = 0; q < s; q++) p[q] = p[q] + p[q + s];
    return p[0];
}

void orc_corr_pairs(const double *z64, int L, const int *ii, const int *jj, int64_t n, double *out)
{
    for (int64_t t = 0; t < n; t++)
        out[t] = orc_dot64(z64 + (int64_t)ii[t] * L, z64 + (int64_t)jj[t] * L, L);
}

/* Reference CPU path, CorMat.cpp:954-965: sum3 (double) += float * float. */
static inline float ref_pair(const float *a, const float *b, int L)
{
    double s = 0.0;
    for (int l = 0; l < L; l++) {
        float p = a[l] * b[l];
        s += (double)p;
    }
    return (float)s;
}

/* Packed upper triangle of the reference-path r for rows [i0, i1), written in
 * row-major pair order starting at out (caller sizes it). */
void orc_ref_corr_rows(const float *z32, int N, int L, int i0, int i1, float *out)
{
    int64_t k = 0;
    for (int i = i0; i < i1; i++)
        for (int j = i + 1; j < N; j++) out[k++] = ref_pair(z32 + (int64_t)i * L, z32 + (int64_t)j * L, L);
}

/* -------------------------------------------------- threaded row partitions */

typedef struct {
    int64_t n, cap;
    int *i, *j;
    double *r;
} edge_vec;

static void ev_push(edge_vec *v, int i, int j, double r)
{
    if (v->n == v->cap) {
        v->cap = v->cap ? v->cap * 2 : 1024;
        v->i = (int *)realloc(v->i, sizeof(int) * v->cap);
        v->j = (int *)realloc(v->j, sizeof(int) * v->cap);
        v->r = (double *)realloc(v->r, sizeof(double) * v->cap);
    }
    v->i[v->n] = i;
    v->j[v->n] = j;
    v->r[v->n] = r;
    v->n++;
}

static void ev_free(edge_vec *v)
{
    free(v->i);
    free(v->j);
    free(v->r);
    memset(v, 0, sizeof(*v));
}

typedef struct {
    const double *z64;
    const unsigned char *valid;
    int N, L, tid, nthreads;
    double thr;
    edge_vec out;
} thr_arg;

/* Rows are dealt round-robin so the triangular work is balanced. */
static void *thr_worker(void *p)
{
    thr_arg *a = (thr_arg *)p;
    for (int i = a->tid; i < a->N; i += a->nthreads) {
        if (a->valid && !a->valid[i]) continue;
        const double *zi = a->z64 + (int64_t)i * a->L;
        for (int j = i + 1; j < a->N; j++) {
            if (a->valid && !a->valid[j]) continue;
            double r = orc_dot64(zi, a->z64 + (int64_t)j * a->L, a->L);
            if (r > a->thr) ev_push(&a->out, i, j, r);
        }
    }
    return NULL;
}

/* ----------------------------------------------------------------- graphs */

typedef struct {
    int N;
    int64_t nnz;       /* directed entries = 2 x undirected edges */
    int64_t *R;        /* N + 1 */
    int *C;            /* nnz */
    float *V;          /* nnz, (float) r64 */
    double *V64;       /* nnz, r64 */
    double r_star;     /* density mode: the k-th largest r */
    int64_t n_ties;    /* density mode: pairs with r == r_star */
} orc_graph;

/* FormCSRGraph order (CorMat.cpp:889-918): for each row ascending, columns
 * ascending; each undirected edge appears in both rows. */
static orc_graph *graph_from_edges(int N, edge_vec *e)
{
    orc_graph *g = (orc_graph *)calloc(1, sizeof(orc_graph));
    g->N = N;
    g->R = (int64_t *)calloc((size_t)N + 1, sizeof(int64_t));
    for (int64_t t = 0; t < e->n; t++) {
        g->R[e->i[t] + 1]++;
        g->R[e->j[t] + 1]++;
    }
    for (int i = 0; i < N; i++) g->R[i + 1] += g->R[i];
    g->nnz = g->R[N];
    g->C = (int *)malloc(sizeof(int) * (g->nnz ? g->nnz : 1));
    g->V = (float *)malloc(sizeof(float) * (g->nnz ? g->nnz : 1));
    g->V64 = (double *)malloc(sizeof(double) * (g->nnz ? g->nnz : 1));
    int64_t *cur = (int64_t *)malloc(sizeof(int64_t) * ((size_t)N + 1));
    memcpy(cur, g->R, sizeof(int64_t) * ((size_t)N + 1));
    /* Edges sorted by (i, j): put (i -> j) entries; lower (j -> i) entries of
     * row j come from rows i < j which are visited first, so a sweep over the
     * edges sorted by (i, j) fills every row in ascending column order when the
     * lower entries are placed before the upper ones.  Do it in two sweeps. */
    /* sweep 1: count lower part per row to know where upper entries start */
    int64_t *lowcnt = (int64_t *)calloc((size_t)N, sizeof(int64_t));
    for (int64_t t = 0; t < e->n; t++) lowcnt[e->j[t]]++;
    int64_t *upcur = (int64_t *)malloc(sizeof(int64_t) * (size_t)N);
    for (int i = 0; i < N; i++) upcur[i] = g->R[i] + lowcnt[i];
    /* edges must be in (i, j) order for both sweeps */
    for (int64_t t = 0; t < e->n; t++) {
        int i = e->i[t], j = e->j[t];
        int64_t pu = upcur[i]++;
        g->C[pu] = j;
        g->V[pu] = (float)e->r[t];
        g->V64[pu] = e->r[t];
        int64_t pl = cur[j]++;
        g->C[pl] = i;
        g->V[pl] = (float)e->r[t];
        g->V64[pl] = e->r[t];
    }
    free(cur);
    free(lowcnt);
    free(upcur);
    return g;
}

static int cmp_edge_idx(const void *a, const void *b, void *ctx)
{
    const edge_vec *e = (const edge_vec *)ctx;
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    if (e->i[x] != e->i[y]) return e->i[x] < e->i[y] ? -1 : 1;
    if (e->j[x] != e->j[y]) return e->j[x] < e->j[y] ? -1 : 1;
    return 0;
}

/* merge per-thread vectors into one (i, j)-sorted vector */
static void merge_sorted(edge_vec *parts, int np, edge_vec *out)
{
    int64_t n = 0;
    for (int p = 0; p < np; p++) n += parts[p].n;
    edge_vec all = {0};
    all.cap = n ? n : 1;
    all.i = (int *)malloc(sizeof(int) * all.cap);
    all.j = (int *)malloc(sizeof(int) * all.cap);
    all.r = (double *)malloc(sizeof(double) * all.cap);
    for (int p = 0; p < np; p++) {
        memcpy(all.i + all.n, parts[p].i, sizeof(int) * parts[p].n);
        memcpy(all.j + all.n, parts[p].j, sizeof(int) * parts[p].n);
        memcpy(all.r + all.n, parts[p].r, sizeof(double) * parts[p].n);
        all.n += parts[p].n;
    }
    int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * (n ? n : 1));
    for (int64_t t = 0; t < n; t++) idx[t] = t;
    qsort_r(idx, (size_t)n, sizeof(int64_t), cmp_edge_idx, &all);
    out->n = 0;
    out->cap = n ? n : 1;
    out->i = (int *)malloc(sizeof(int) * out->cap);
    out->j = (int *)malloc(sizeof(int) * out->cap);
    out->r = (double *)malloc(sizeof(double) * out->cap);
    for (int64_t t = 0; t < n; t++) {
        out->i[t] = all.i[idx[t]];
        out->j[t] = all.j[idx[t]];
        out->r[t] = all.r[idx[t]];
    }
    out->n = n;
    free(idx);
    ev_free(&all);
}

/* Binary or weighted graph of all valid pairs with r64 > thr (FormFullGraph's
 * strict comparison, CorMat.cpp:696). */
orc_graph *orc_build_threshold(const double *z64, const unsigned char *valid, int N, int L,
                               double thr, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
    thr_arg *args = (thr_arg *)calloc((size_t)nthreads, sizeof(thr_arg));
    for (int t = 0; t < nthreads; t++) {
        args[t].z64 = z64;
        args[t].valid = valid;
        args[t].N = N;
        args[t].L = L;
        args[t].tid = t;
        args[t].nthreads = nthreads;
        args[t].thr = thr;
        pthread_create(&th[t], NULL, thr_worker, &args[t]);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    edge_vec parts_sorted = {0};
    edge_vec *parts = (edge_vec *)malloc(sizeof(edge_vec) * nthreads);
    for (int t = 0; t < nthreads; t++) parts[t] = args[t].out;
    merge_sorted(parts, nthreads, &parts_sorted);
    for (int t = 0; t < nthreads; t++) ev_free(&parts[t]);
    orc_graph *g = graph_from_edges(N, &parts_sorted);
    g->r_star = thr;
    ev_free(&parts_sorted);
    free(parts);
    free(args);
    free(th);
    return g;
}

/* CorMat.cpp:327-328.  `(long long) 2*M1*s/100.0` multiplies the long long 2*M1
 * by the float s in float arithmetic, divides by 100.0 in double and truncates;
 * an odd count is rounded up to even.  Returns the number of undirected edges. */
int64_t orc_density_k(int N, float spa_percent)
{
    long long M1 = (long long)(N - 1) * N / 2;
    long long twoM1 = 2LL * M1;
    float prod = (float)twoM1 * spa_percent;
    long long Clength = (long long)((double)prod / 100.0);
    Clength += Clength % 2;
    return Clength / 2;
}

typedef struct {
    const double *z64;
    const unsigned char *valid;
    int N, L, tid, nthreads;
    double *packed; /* row-major upper triangle, NaN for invalid pairs */
} all_arg;

static inline int64_t packed_index(int N, int i, int j)
{
    return (int64_t)i * N - (int64_t)(1 + i) * i / 2 + j - (i + 1);
}

static void *all_worker(void *p)
{
    all_arg *a = (all_arg *)p;
    for (int i = a->tid; i < a->N; i += a->nthreads) {
        const double *zi = a->z64 + (int64_t)i * a->L;
        for (int j = i + 1; j < a->N; j++) {
            double r = NAN;
            if (!a->valid || (a->valid[i] && a->valid[j])) r = orc_dot64(zi, a->z64 + (int64_t)j * a->L, a->L);
            a->packed[packed_index(a->N, i, j)] = r;
        }
    }
    return NULL;
}

static int cmp_desc(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return (x < y) - (x > y);
}

/* All pairs' r64 in row-major upper-triangle order (NaN for invalid pairs). */
void orc_all_pairs(const double *z64, const unsigned char *valid, int N, int L, int nthreads, double *packed)
{
    if (nthreads < 1) nthreads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
    all_arg *args = (all_arg *)calloc((size_t)nthreads, sizeof(all_arg));
    for (int t = 0; t < nthreads; t++) {
        args[t] = (all_arg){z64, valid, N, L, t, nthreads, packed};
        pthread_create(&th[t], NULL, all_worker, &args[t]);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    free(args);
    free(th);
}

/* Exact-k selection over a packed triangle (NaN entries excluded).
 * literal = 1 runs the reference's stack loop (CorMat.cpp:784-812) verbatim in
 * behaviour; literal = 0 applies the closed form: every r > r*, then the first
 * m ties at r* in row-major order.  Returns NULL if fewer than k valid pairs. */
orc_graph *orc_select_packed(const double *packed, int N, int64_t k, int literal)
{
    int64_t M1 = (int64_t)N * (N - 1) / 2;
    int64_t nv = 0;
    for (int64_t t = 0; t < M1; t++) nv += !isnan(packed[t]);
    if (k < 1 || k > nv) return NULL;
    double *cp = (double *)malloc(sizeof(double) * (nv ? nv : 1));
    int64_t c = 0;
    for (int64_t t = 0; t < M1; t++)
        if (!isnan(packed[t])) cp[c++] = packed[t];
    qsort(cp, (size_t)nv, sizeof(double), cmp_desc);
    double rs = cp[k - 1];
    free(cp);

    edge_vec e = {0};
    int64_t n_ties = 0;
    if (literal) {
        unsigned char *adj = (unsigned char *)calloc((size_t)M1 ? (size_t)M1 : 1, 1);
        int64_t *stack = (int64_t *)malloc(sizeof(int64_t) * (M1 ? M1 : 1));
        int64_t sp = 0, cnt = 0, idx = 0;
        for (int i = 0; i < N; i++)
            for (int j = i + 1; j < N; j++, idx++) {
                double r = packed[idx];
                if (r >= rs) { adj[idx] = 1; cnt++; }
                if (r == rs) { stack[sp++] = idx; n_ties++; }
                if (cnt > k) { adj[stack[--sp]] = 0; cnt--; }
            }
        idx = 0;
        for (int i = 0; i < N; i++)
            for (int j = i + 1; j < N; j++, idx++)
                if (adj[idx]) ev_push(&e, i, j, packed[idx]);
        free(adj);
        free(stack);
    } else {
        int64_t above = 0, idx = 0;
        for (int64_t t = 0; t < M1; t++) {
            above += packed[t] > rs;
            n_ties += packed[t] == rs;
        }
        int64_t m = k - above;
        for (int i = 0; i < N; i++)
            for (int j = i + 1; j < N; j++, idx++) {
                double r = packed[idx];
                if (r > rs) ev_push(&e, i, j, r);
                else if (r == rs && m > 0) { ev_push(&e, i, j, r); m--; }
            }
    }
    orc_graph *g = graph_from_edges(N, &e);
    g->r_star = rs;
    g->n_ties = n_ties;
    ev_free(&e);
    return g;
}

orc_graph *orc_build_density(const double *z64, const unsigned char *valid, int N, int L,
                             int64_t k, int nthreads, int literal)
{
    int64_t M1 = (int64_t)N * (N - 1) / 2;
    double *packed = (double *)malloc(sizeof(double) * (M1 ? M1 : 1));
    orc_all_pairs(z64, valid, N, L, nthreads, packed);
    orc_graph *g = orc_select_packed(packed, N, k, literal);
    free(packed);
    return g;
}

/* fishertrans (CorMat.cpp:36-42): r == 1 is moved below 1 by 1e-6 (in double,
 * stored as float), the quotient is float, log of a float is float
 * (std::log(float)), the half is applied in double and the result stored as
 * float. */
float orc_fisher(float r)
{
    if (r == 1.0f) r = (float)((double)r - 1e-6);
    return (float)(0.5 * (double)logf((1.0f + r) / (1.0f - r)));
}

/* inv_fishertrans (CorMat.cpp:44-50): exp of a float is float, the rest float. */
float orc_inv_fisher(float z)
{
    float e = expf(2.0f * z);
    return (e - 1.0f) / (e + 1.0f);
}

/* post_block (CorMat.cpp:824-887) for one subject whose pair values are
 * packed_r (double, NaN for invalid pairs), taken as float the way the
 * reference's float matrix holds them. */
void orc_group_add(float *acc, const double *packed_r, int64_t M1, int fisher)
{
    for (int64_t t = 0; t < M1; t++) {
        float r = (float)packed_r[t];
        if (r >= -1.0f && r <= 1.0f) acc[t] += fisher ? orc_fisher(r) : r;
    }
}

/* CorMat.cpp:548-558. */
void orc_group_finish(float *acc, int64_t M1, int S, int fisher)
{
    for (int64_t t = 0; t < M1; t++) {
        acc[t] /= (float)S;
        if (fisher) acc[t] = orc_inv_fisher(acc[t]);
    }
}

/* FormFullGraph (CorMat.cpp:683-703) over packed values: pair (i, j) is an edge
 * iff v > thr; then FormCSRGraph with V = v. */
orc_graph *orc_threshold_packed(const double *packed, int N, double thr)
{
    edge_vec e = {0};
    int64_t idx = 0;
    for (int i = 0; i < N; i++)
        for (int j = i + 1; j < N; j++, idx++)
            if (packed[idx] > thr) ev_push(&e, i, j, packed[idx]);
    orc_graph *g = graph_from_edges(N, &e);
    g->r_star = thr;
    ev_free(&e);
    return g;
}

int64_t orc_graph_nnz(const orc_graph *g) { return g ? g->nnz : -1; }
double orc_graph_rstar(const orc_graph *g) { return g ? g->r_star : NAN; }
int64_t orc_graph_ties(const orc_graph *g) { return g ? g->n_ties : -1; }

void orc_graph_copy(const orc_graph *g, int64_t *R, int *C, float *V, double *V64)
{
    memcpy(R, g->R, sizeof(int64_t) * ((size_t)g->N + 1));
    if (C) memcpy(C, g->C, sizeof(int) * g->nnz);
    if (V) memcpy(V, g->V, sizeof(float) * g->nnz);
    if (V64) memcpy(V64, g->V64, sizeof(double) * g->nnz);
}

void orc_graph_free(orc_graph *g)
{
    if (!g) return;
    free(g->R);
    free(g->C);
    free(g->V);
    free(g->V64);
    free(g);
}

/* ------------------------------------------------------------------ degree */

/* weighted: float sum of V in CSR order (Degree.cpp:85-87); V == NULL gives the
 * binary degree (float)(R[k+1] - R[k]) (B/Degree/Degree/Degree.cpp:84-91). */
void orc_degree(const int64_t *R, const float *V, int N, float *deg)
{
    for (int k = 0; k < N; k++) {
        if (V) {
            float s = 0.0f;
            for (int64_t t = R[k]; t < R[k + 1]; t++) s += V[t];
            deg[k] = s;
        } else {
            deg[k] = (float)(R[k + 1] - R[k]);
        }
    }
}

/* -------------------------------------------------- clustering coefficient */

/* The CUCP tool's Cp_2 (W/Cp_GPU/CUCP/Cp.cu:111-471), restated on the host:
 *   alpha = (float)(1.0 / V[m]), m the first entry of largest |V| (cublasIsamax
 *   and the host division, :189-205); w = (float)pow((double)(V * alpha), 1.0/3)
 *   (cublasSscal, pow3_vector :33-46); cci_i = (W^3)_ii (the csrmm over dense
 *   row blocks and calc_cci_block, :48-92, :317-341), here over a dense scatter
 *   of row i, summed in double; t = (float)R[i+1] - (float)R[i], inv = 0 if
 *   t < 2 else (float)(1.0 / (t * (t - 1))) (cal_k_kernel :15-31); Cp_i =
 *   (float)cci_i * inv (dot_vv :96-109).  V == NULL: every w is 1.  Returns the
 *   float mean of |Cp| (cublasSasum / N, :455-457) as a double. */
double orc_cp(const int64_t *R, const int *C, const float *V, int N, float *cp)
{
    const int64_t nnz = R[N];
    float *w = (float *)malloc(sizeof(float) * (size_t)(nnz > 0 ? nnz : 1));
    double *row = (double *)calloc((size_t)N, sizeof(double));
    unsigned char *mark = (unsigned char *)calloc((size_t)N, 1);
    if (V && nnz > 0) {
        int64_t m = -1;
        float best = -1.0f;
        for (int64_t q = 0; q < nnz; q++) {
            const float a = fabsf(V[q]);
            if (a == a && a > best) { best = a; m = q; }
        }
        const float alpha = m >= 0 ? (float)(1.0 / (double)V[m]) : NAN;
        for (int64_t q = 0; q < nnz; q++) {
            const float s = V[q] * alpha;
            w[q] = (float)pow((double)s, 1.0 / 3);
        }
    } else {
        for (int64_t q = 0; q < nnz; q++) w[q] = 1.0f;
    }
    double sum = 0.0;
    for (int i = 0; i < N; i++) {
        for (int64_t q = R[i]; q < R[i + 1]; q++) { row[C[q]] = w[q]; mark[C[q]] = 1; }
        double cci = 0.0;
        for (int64_t q = R[i]; q < R[i + 1]; q++) {
            const int j = C[q];
            for (int64_t p = R[j]; p < R[j + 1]; p++) {
                const int k = C[p];
                if (mark[k]) cci += (double)w[q] * ((double)w[p] * row[k]);
            }
        }
        for (int64_t q = R[i]; q < R[i + 1]; q++) { row[C[q]] = 0.0; mark[C[q]] = 0; }
        float t = (float)R[i + 1];
        t -= (float)R[i];
        const float inv = t < 2.0f ? 0.0f : (float)(1.0 / (double)(t * (t - 1.0f)));
        cp[i] = (float)cci * inv;
        sum += fabs((double)cp[i]);
    }
    free(w);
    free(row);
    free(mark);
    return (double)((float)sum / (float)N);
}

/* ----------------------------------------------------- eigenvector centrality */

static double nrm2(const double *x, int N)
{
    double s = 0.0;
    for (int i = 0; i < N; i++) s += x[i] * x[i];
    return sqrt(s);
}

/* Power iteration as W/CUEC/weighted-net.cu:154-189:
 *   x0 <- x; n = ||x||; x <- x * (1/n); y = A x; x <- y;
 *   err1 = ||x - x0||, err2 = ||x + x0||; loop while both > tol and it < max_iter.
 * Start x[i] = i (init_AD, weighted-net.cu:38-43).  V == NULL is the binary graph
 * (init_unweightednet sets every value to 1, CUEC_unweighted.cu:47-52).
 * out_mode 0: weighted tree's output, the last A*y (weighted-net.cu:190);
 * out_mode 1: binary tree's output, the last unit-norm y (CUEC_unweighted.cu:164).
 * Returns n (the eigenvalue estimate written to _evalue.txt). */
double orc_eigcent(const int64_t *R, const int *C, const float *V, int N, double tol, int max_iter,
                   int out_mode, double *ec, int *iters)
{
    double *x = (double *)malloc(sizeof(double) * N);
    double *x0 = (double *)malloc(sizeof(double) * N);
    double *y = (double *)malloc(sizeof(double) * N);
    double *d = (double *)malloc(sizeof(double) * N);
    for (int i = 0; i < N; i++) x[i] = (double)i;
    double err1 = 1.0, err2 = 1.0, n = 0.0;
    int it = 0;
    while (err1 > tol && err2 > tol && it < max_iter) {
        memcpy(x0, x, sizeof(double) * N);
        n = nrm2(x, N);
        double inv = 1.0 / n;
        for (int i = 0; i < N; i++) x[i] *= inv;
        if (out_mode == 1) memcpy(ec, x, sizeof(double) * N);
        for (int i = 0; i < N; i++) {
            double s = 0.0;
            for (int64_t t = R[i]; t < R[i + 1]; t++) s += (V ? (double)V[t] : 1.0) * x[C[t]];
            y[i] = s;
        }
        memcpy(x, y, sizeof(double) * N);
        for (int i = 0; i < N; i++) d[i] = x[i] - x0[i];
        err1 = nrm2(d, N);
        for (int i = 0; i < N; i++) d[i] = x[i] + x0[i];
        err2 = nrm2(d, N);
        it++;
    }
    if (out_mode == 0) memcpy(ec, x, sizeof(double) * N);
    if (iters) *iters = it;
    free(x);
    free(x0);
    free(y);
    free(d);
    return n;
}

/* ----------------------------------------------------- Newman modularity */
#define NM_LAMBDA 0.01      /* a division needs an eigenvalue above this (modularity_GPU.cu:30) */
#define NM_MIN_GROUP 1      /* and more than this many nodes on each side (:31) */
#define NM_GPU_ROUND 25000  /* groups above this size take the GPU sub-partition (:494) */

/* Lead_Vector (Modularity_CPU.cpp:209-266): power iteration on B(g) - beta I
 * restricted to the group; v is left in nv[0, n), the previous iterate in nv0.
 * Returns the signed norm of the last iterate, as the reference does. */
static double nm_lead_vector(const int64_t *R, int n, const int *AD, const int64_t *SR, const int *SC,
                             const double *sumBG, double m, double beta, int max_iter, double eps, double *nv,
                             double *nv0, double *vv, int *iters)
{
    for (int i = 0; i < n; i++) nv[i] = (double)AD[i];
    double err1 = 1, err2 = 1, vnorm = 0;
    int it = 0;
    while (err1 > eps && err2 > eps && it < max_iter) {
        for (int i = 0; i < n; i++) {
            nv0[i] = nv[i];
            vv[AD[i]] = nv[i];
        }
        double vk = 0;
        for (int i = 0; i < n; i++) vk += nv0[i] * (double)(R[AD[i] + 1] - R[AD[i]]);
        for (int i = 0; i < n; i++) {
            double t = 0;
            for (int64_t j = SR[i]; j < SR[i + 1]; j++) t += vv[SC[j]];
            t -= vk / 2 / m * (double)(R[AD[i] + 1] - R[AD[i]]) + (sumBG[i] - beta) * nv0[i];
            nv[i] = t;
        }
        double s = 0;
        for (int i = 0; i < n; i++) s += nv[i] * nv[i];
        vnorm = sqrt(s);
        for (int i = 0; i < n; i++) nv[i] = nv[i] / vnorm;
        double e1 = 0, e2 = 0;
        for (int i = 0; i < n; i++) {
            const double a = nv[i] - nv0[i], b = nv[i] + nv0[i];
            e1 += a * a;
            e2 += b * b;
        }
        err1 = sqrt(e1);
        err2 = sqrt(e2);
        it++;
    }
    *iters = it;
    int mi = 0;
    for (int i = 0; i < n; i++)
        if (fabs(nv[i]) > fabs(nv[mi])) mi = i;
    return (n > 0 && nv[mi] * nv0[mi] > 0) ? vnorm : -vnorm;
}

/* Round 0 of the partition alone (Sub_Partition on the whole graph): the
 * leading vector v (n = N entries) and the eigenvalue after the shift. */
double orc_modularity_first(const int64_t *R, const int *C, int N, int max_iter, double eps, double *v)
{
    if (N < 1) return 0.0;
    const double m = (double)((R[N] - R[0]) / 2);
    int *AD = malloc(sizeof(int) * (size_t)(N + 1));
    double *sumBG = malloc(sizeof(double) * (size_t)(N + 1));
    double *nv0 = malloc(sizeof(double) * (size_t)(N + 1));
    double *vv = calloc((size_t)N + 1, sizeof(double));
    for (int i = 0; i < N; i++) {
        AD[i] = i;
        sumBG[i] = (double)(R[i + 1] - R[i]) - (double)(R[i + 1] - R[i]) * (double)(R[N] - R[0]) / 2 / m;
    }
    int it = 0;
    double lambda = nm_lead_vector(R, N, AD, R, C, sumBG, m, 0.0, max_iter, eps, v, nv0, vv, &it);
    const int shift = N > NM_GPU_ROUND ? (lambda < -NM_LAMBDA) : (lambda < 0);
    if (shift) lambda += nm_lead_vector(R, N, AD, R, C, sumBG, m, -lambda, max_iter, eps, v, nv0, vv, &it);
    free(AD);
    free(sumBG);
    free(nv0);
    free(vv);
    return lambda;
}

/* Partition_GPU (modularity_GPU.cu:408-528) with Sub_Partition(_GPU): repeated
 * bisection of the group labelled Round until no group divides; labels[i] is
 * the module of voxel i, numbered from 1 in the order the modules are found
 * final.  Q = (sum over entries of [same module] - sum over modules of
 * (sum of degrees)^2 / 2M) / 2M (the reference's double loop over all pairs,
 * summed per module).  Returns the number of modules; iters_total gets the
 * power iterations of all rounds.  R is int64 offsets of a binary CSR. */
int orc_modularity(const int64_t *R, const int *C, int N, int max_iter, double eps, int *labels, double *Q,
                   int *rounds_out, long long *iters_total)
{
    if (N < 1) return 0;
    const double m = (double)((R[N] - R[0]) / 2);
    int *res = calloc((size_t)N, sizeof(int));
    int *adjust = calloc((size_t)2 * N + 4, sizeof(int));
    int *AD = malloc(sizeof(int) * (size_t)(N + 1));
    char *G = malloc((size_t)N + 1);
    int64_t *SR = malloc(sizeof(int64_t) * (size_t)(N + 1));
    int *SC = malloc(sizeof(int) * (size_t)(R[N] + 1));
    double *sumBG = malloc(sizeof(double) * (size_t)(N + 1));
    double *nv = malloc(sizeof(double) * (size_t)(N + 1));
    double *nv0 = malloc(sizeof(double) * (size_t)(N + 1));
    double *vv = calloc((size_t)N + 1, sizeof(double));
    int round = 0, max_result = 0, index = 1, modules = 1;
    long long its = 0;
    while (round <= max_result) {
        long long inner = (long long)(R[N] - R[0]);
        int n = 0;
        for (int i = 0; i < N; i++) {
            G[i] = res[i] == round;
            if (G[i])
                AD[n++] = i;
            else
                inner -= R[i + 1] - R[i];
        }
        int64_t e = 0;
        SR[0] = 0;
        for (int a = 0; a < n; a++) {
            const int i = AD[a];
            for (int64_t j = R[i]; j < R[i + 1]; j++)
                if (G[C[j]]) SC[e++] = C[j];
            SR[a + 1] = e;
        }
        if (n) {
            for (int a = 0; a < n; a++)
                sumBG[a] = (double)(SR[a + 1] - SR[a]) - (double)(R[AD[a] + 1] - R[AD[a]]) * (double)inner / 2 / m;
            int it = 0;
            double lambda = nm_lead_vector(R, n, AD, SR, SC, sumBG, m, 0.0, max_iter, eps, nv, nv0, vv, &it);
            its += it;
            const int shift = n > NM_GPU_ROUND ? (lambda < -NM_LAMBDA) : (lambda < 0);
            if (shift) {
                lambda += nm_lead_vector(R, n, AD, SR, SC, sumBG, m, -lambda, max_iter, eps, nv, nv0, vv, &it);
                its += it;
            }
            long long sp = 0, sn = 0;
            for (int a = 0; a < n; a++) {
                sp += nv[a] > 0;
                sn += nv[a] <= 0;
            }
            const int sub = lambda > NM_LAMBDA && sp > NM_MIN_GROUP && sn > NM_MIN_GROUP;
            if (sub) {
                for (int a = 0; a < n; a++) res[AD[a]] = max_result + 1 + (nv[a] * ((double)(sp - sn) + 0.5) <= 0);
                max_result += 2;
            } else {
                adjust[round] = index++;
            }
            modules += sub;
        }
        round++;
    }
    /* Q: entries inside a module, minus the degree products per module */
    double within = 0;
    for (int i = 0; i < N; i++)
        for (int64_t j = R[i]; j < R[i + 1]; j++) within += 1.0 * (res[i] == res[C[j]]);
    double *ksum = calloc((size_t)2 * N + 4, sizeof(double));
    for (int i = 0; i < N; i++) ksum[res[i]] += (double)(R[i + 1] - R[i]);
    double expect = 0;
    for (int c = 0; c < 2 * N + 4; c++) expect += ksum[c] * ksum[c] / 2 / m;
    *Q = (within - expect) / (2 * m);
    for (int i = 0; i < N; i++) labels[i] = adjust[res[i]];
    if (rounds_out) *rounds_out = round;
    if (iters_total) *iters_total = its;
    free(ksum);
    free(res);
    free(adjust);
    free(AD);
    free(G);
    free(SR);
    free(SC);
    free(sumBG);
    free(nv);
    free(nv0);
    free(vv);
    return modules;
}

/* ------------------------------------------------ reference CPU path timing */

typedef struct {
    const float *z32;
    int N, L, i0, i1, tid, nthreads;
    float thr;
    int64_t edges;
    double checksum;
} ref_arg;

/* CorMat_cpu arithmetic (CorMat.cpp:954-965) with CorMat_cpu_MT's j-striping
 * over threads (B/CorMat/CorMat_CPU/CorMat/CorMat_CPU.cpp:371), counting the
 * pairs with r > thr (FormFullGraph's comparison).  Eight pairs run side by side
 * so the double accumulation chains overlap; each pair's sum keeps the
 * reference's order. */
static void *ref_worker(void *p)
{
    ref_arg *a = (ref_arg *)p;
    const int L = a->L;
    int64_t edges = 0;
    double cs = 0.0;
    for (int i = a->i0; i < a->i1; i++) {
        const float *zi = a->z32 + (int64_t)i * L;
        int j = i + 1 + a->tid;
        const int st = a->nthreads;
        for (; j + 7 * st < a->N; j += 8 * st) {
            double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            const float *zj[8];
            for (int u = 0; u < 8; u++) zj[u] = a->z32 + (int64_t)(j + u * st) * L;
            for (int l = 0; l < L; l++) {
                float zl = zi[l];
                for (int u = 0; u < 8; u++) {
                    float pr = zl * zj[u][l];
                    s[u] += (double)pr;
                }
            }
            for (int u = 0; u < 8; u++) {
                float r = (float)s[u];
                edges += r > a->thr;
                cs += r;
            }
        }
        for (; j < a->N; j += st) {
            float r = ref_pair(zi, a->z32 + (int64_t)j * L, L);
            edges += r > a->thr;
            cs += r;
        }
    }
    a->edges = edges;
    a->checksum = cs;
    return NULL;
}

/* Runs the reference CPU path over rows [i0, i1) with nthreads threads; returns
 * the number of pairs with r > thr and a checksum of the r values. */
int64_t orc_ref_cpu_rows(const float *z32, int N, int L, int i0, int i1, float thr, int nthreads,
                         double *checksum)
{
    if (nthreads < 1) nthreads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
    ref_arg *args = (ref_arg *)calloc((size_t)nthreads, sizeof(ref_arg));
    for (int t = 0; t < nthreads; t++) {
        args[t] = (ref_arg){z32, N, L, i0, i1, t, nthreads, thr, 0, 0.0};
        pthread_create(&th[t], NULL, ref_worker, &args[t]);
    }
    int64_t e = 0;
    double cs = 0.0;
    for (int t = 0; t < nthreads; t++) {
        pthread_join(th[t], NULL);
        e += args[t].edges;
        cs += args[t].checksum;
    }
    if (checksum) *checksum = cs;
    free(args);
    free(th);
    return e;
}

/* ------------------------------------------ reference CPU path, end to end */

/* The reference's CPU construction path, timed stage by stage, for the CPU
 * baseline (not a parity function: the GPU is checked against the FP64
 * functions above):
 *   normalise   CorMat.cu:33-52 (the host transpose and z in float, double sums)
 *   correlate   CorMat_cpu's arithmetic (CorMat.cpp:954-965: float z products
 *               summed in double, stored as float) into the packed upper
 *               triangle, j striped over threads as CorMat_cpu_MT does
 *               (B/CorMat/CorMat_CPU/CorMat/CorMat_CPU.cpp:283-383)
 *   threshold   FormFullGraph (CorMat.cpp:683-703, r > thr) or FormFullGraph_s
 *               (CorMat.cpp:743-822: the k-th largest packed value by
 *               selection, then r >= r* with the ties trimmed to exactly k in
 *               row-major order) into the N x N bool matrix
 *   csr         FormCSRGraph (CorMat.cpp:889-918): a row-major scan of the bool
 *               matrix, V from the packed triangle
 *   degree      Degree.cpp:83-87 (float sums in CSR order)
 *   ec          the power iteration of weighted-net.cu:154-189 on the CPU
 * Rows [row0, row1) < N: the correlation, threshold and CSR of those rows'
 * upper-triangle pairs only (a bounded sample of a large subject); degree and
 * centrality then do not run.  Returns the undirected edge count, or -1. */
typedef struct {
    double t_norm, t_corr, t_thr, t_csr, t_deg, t_ec;
    int64_t pairs, edges, nnz;
    int ec_iters;
    double r_star, lambda;
} orc_ref_times;

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

typedef struct {
    const float *z;
    int N, L, row0, row1, tid, nthreads;
    float *packed;  /* rows [row0, row1), pairs j > i, row-major */
    const int64_t *rowbase;
} rp_arg;

static void *rp_worker(void *p)
{
    rp_arg *a = (rp_arg *)p;
    const int L = a->L;
    for (int i = a->row0; i < a->row1; i++) {
        const float *zi = a->z + (int64_t)i * L;
        float *out = a->packed + a->rowbase[i - a->row0] - (i + 1);
        int j = i + 1 + a->tid;
        const int st = a->nthreads;
        for (; j + 7 * st < a->N; j += 8 * st) {
            double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            const float *zj[8];
            for (int u = 0; u < 8; u++) zj[u] = a->z + (int64_t)(j + u * st) * L;
            for (int l = 0; l < L; l++) {
                float zl = zi[l];
                for (int u = 0; u < 8; u++) {
                    float pr = zl * zj[u][l];
                    s[u] += (double)pr;
                }
            }
            for (int u = 0; u < 8; u++) out[j + u * st] = (float)s[u];
        }
        for (; j < a->N; j += st) out[j] = ref_pair(zi, a->z + (int64_t)j * L, L);
    }
    return NULL;
}

static void swapf(float *a, float *b) { float t = *a; *a = *b; *b = t; }

/* k-th largest (1-based) of v[0, n) by quickselect (the reference's select). */
static float kth_largest(float *v, int64_t n, int64_t k)
{
    int64_t lo = 0, hi = n - 1, want = k - 1;
    while (lo < hi) {
        float piv = v[lo + (hi - lo) / 2];
        int64_t i = lo, j = hi;
        while (i <= j) {
            while (v[i] > piv) i++;
            while (v[j] < piv) j--;
            if (i <= j) { swapf(&v[i], &v[j]); i++; j--; }
        }
        if (want <= j) hi = j;
        else if (want >= i) lo = i;
        else break;
    }
    return v[want];
}

int64_t orc_ref_pipeline(const float *x, int N, int L, int layout, int row0, int row1, double thr,
                         int64_t k, int nthreads, orc_ref_times *t)
{
    memset(t, 0, sizeof(*t));
    if (nthreads < 1) nthreads = 1;
    if (row0 < 0 || row1 > N || row0 >= row1) return -1;
    const int full = (row0 == 0 && row1 == N);
    double t0 = now_s();
    /* normalise: z in float, mean and norm in double, in time order */
    float *z = (float *)malloc(sizeof(float) * (size_t)N * L);
    for (int i = 0; i < N; i++) {
        double s1 = 0.0, s2 = 0.0;
        for (int l = 0; l < L; l++) s1 += layout ? x[(int64_t)i * L + l] : x[(int64_t)l * N + i];
        s1 /= L;
        for (int l = 0; l < L; l++) {
            double d = (layout ? x[(int64_t)i * L + l] : x[(int64_t)l * N + i]) - s1;
            double p = d * d;
            s2 += p;
        }
        s2 = sqrt(s2);
        for (int l = 0; l < L; l++)
            z[(int64_t)i * L + l] = (float)(((layout ? x[(int64_t)i * L + l] : x[(int64_t)l * N + i]) - s1) / s2);
    }
    double t1 = now_s();
    t->t_norm = t1 - t0;
    /* correlate into the packed rows */
    const int nr = row1 - row0;
    int64_t *rowbase = (int64_t *)malloc(sizeof(int64_t) * ((size_t)nr + 1));
    rowbase[0] = 0;
    for (int q = 0; q < nr; q++) rowbase[q + 1] = rowbase[q] + (N - 1 - (row0 + q));
    const int64_t np = rowbase[nr];
    t->pairs = np;
    float *packed = (float *)malloc(sizeof(float) * (size_t)(np ? np : 1));
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
    rp_arg *args = (rp_arg *)calloc((size_t)nthreads, sizeof(rp_arg));
    for (int q = 0; q < nthreads; q++) {
        args[q] = (rp_arg){z, N, L, row0, row1, q, nthreads, packed, rowbase};
        pthread_create(&th[q], NULL, rp_worker, &args[q]);
    }
    for (int q = 0; q < nthreads; q++) pthread_join(th[q], NULL);
    free(th);
    free(args);
    double t2 = now_s();
    t->t_corr = t2 - t1;
    /* threshold into the bool matrix (rows of the sample; full: N x N) */
    unsigned char *adj = (unsigned char *)calloc((size_t)nr * N, 1);
    int64_t edges = 0;
    if (k > 0) {
        if (!full) { free(z); free(rowbase); free(packed); free(adj); return -1; }
        float *cp = (float *)malloc(sizeof(float) * (size_t)np);
        memcpy(cp, packed, sizeof(float) * (size_t)np);
        const float rs = kth_largest(cp, np, k);
        free(cp);
        int64_t above = 0;
        for (int64_t q = 0; q < np; q++) above += packed[q] > rs;
        int64_t m = k - above, idx = 0;
        for (int i = 0; i < N; i++)
            for (int j = i + 1; j < N; j++, idx++) {
                const float r = packed[idx];
                if (r > rs || (r == rs && m-- > 0)) {
                    adj[(int64_t)i * N + j] = 1;
                    adj[(int64_t)j * N + i] = 1;
                    edges++;
                }
            }
        t->r_star = rs;
    } else {
        for (int q = 0; q < nr; q++) {
            const int i = row0 + q;
            const float *pr = packed + rowbase[q] - (i + 1);
            for (int j = i + 1; j < N; j++)
                if (pr[j] > (float)thr) {
                    adj[(int64_t)q * N + j] = 1;
                    if (full) adj[(int64_t)j * N + i] = 1;
                    edges++;
                }
        }
        t->r_star = thr;
    }
    double t3 = now_s();
    t->t_thr = t3 - t2;
    /* CSR: row-major scan, V from the packed triangle */
    int64_t *R = (int64_t *)malloc(sizeof(int64_t) * ((size_t)nr + 1));
    int64_t nnz = 0;
    R[0] = 0;
    for (int q = 0; q < nr; q++) {
        const unsigned char *ar = adj + (int64_t)q * N;
        for (int j = 0; j < N; j++) nnz += ar[j];
        R[q + 1] = nnz;
    }
    int *Cc = (int *)malloc(sizeof(int) * (size_t)(nnz ? nnz : 1));
    float *V = (float *)malloc(sizeof(float) * (size_t)(nnz ? nnz : 1));
    int64_t e = 0;
    for (int q = 0; q < nr; q++) {
        const int i = row0 + q;
        const unsigned char *ar = adj + (int64_t)q * N;
        for (int j = 0; j < N; j++)
            if (ar[j]) {
                const int a = i < j ? i : j, b = i < j ? j : i;
                Cc[e] = j;
                V[e] = packed[rowbase[a - row0] + (b - a - 1)];
                e++;
            }
    }
    free(adj);
    double t4 = now_s();
    t->t_csr = t4 - t3;
    t->edges = edges;
    t->nnz = nnz;
    if (full) {
        float *deg = (float *)malloc(sizeof(float) * (size_t)N);
        orc_degree(R, V, N, deg);
        double t5 = now_s();
        t->t_deg = t5 - t4;
        double *ec = (double *)malloc(sizeof(double) * (size_t)N);
        t->lambda = orc_eigcent(R, Cc, V, N, 1e-6, 10000, 0, ec, &t->ec_iters);
        t->t_ec = now_s() - t5;
        free(deg);
        free(ec);
    }
    free(R);
    free(Cc);
    free(V);
    free(z);
    free(rowbase);
    free(packed);
    return edges;
}


/* BFS hop distances from s along the CSR rows (dist -1: not reached); the
 * vertices in visiting order go to queue, level by level. Returns how many. */
static int bfs_order(const int64_t *R, const int *C, int N, int s, int *dist, int *queue)
{
    for (int v = 0; v < N; v++) dist[v] = -1;
    dist[s] = 0;
    queue[0] = s;
    int head = 0, tail = 1;
    while (head < tail) {
        const int u = queue[head++];
        for (int64_t e = R[u]; e < R[u + 1]; e++) {
            const int w = C[e];
            if (dist[w] < 0) {
                dist[w] = dist[u] + 1;
                queue[tail++] = w;
            }
        }
    }
    return tail;
}

/* CUBFS_LP.cpp:33-46.  hist[d] (d < hist_len) counts ordered pairs at distance
 * d, hist[0] the pairs with no path.  Returns Lp. */
double orc_lp(const int64_t *R, const int *C, int N, int64_t *hist, int hist_len)
{
    int *dist = (int *)malloc(sizeof(int) * (size_t)N);
    int *queue = (int *)malloc(sizeof(int) * (size_t)N);
    for (int d = 0; d < hist_len; d++) hist[d] = 0;
    double Lp_result = 0.0;
    int64_t reached = 0;
    for (int i = 0; i < N; i++) {
        bfs_order(R, C, N, i, dist, queue);
        for (int j = 0; j < N; j++) {
            const int d = dist[j];
            if (d > 0) {
                Lp_result += 1.0f / d;
                reached++;
                if (d < hist_len) hist[d]++;
            }
        }
    }
    if (hist_len > 0) hist[0] = (int64_t)N * (N - 1) - reached;
    Lp_result /= N;
    Lp_result /= N - 1;
    free(dist);
    free(queue);
    return 1 / Lp_result;
}

/* Brandes betweenness, ordered pairs, in double. */
void orc_bc(const int64_t *R, const int *C, int N, double *bc)
{
    int *dist = (int *)malloc(sizeof(int) * (size_t)N);
    int *queue = (int *)malloc(sizeof(int) * (size_t)N);
    double *sigma = (double *)malloc(sizeof(double) * (size_t)N);
    double *delta = (double *)malloc(sizeof(double) * (size_t)N);
    for (int v = 0; v < N; v++) bc[v] = 0.0;
    for (int s = 0; s < N; s++) {
        const int n = bfs_order(R, C, N, s, dist, queue);
        for (int v = 0; v < N; v++) {
            sigma[v] = 0.0;
            delta[v] = 0.0;
        }
        sigma[s] = 1.0;
        for (int k = 0; k < n; k++) {  /* predecessors are earlier in the queue */
            const int u = queue[k];
            for (int64_t e = R[u]; e < R[u + 1]; e++)
                if (dist[C[e]] == dist[u] + 1) sigma[C[e]] += sigma[u];
        }
        for (int k = n - 1; k >= 0; k--) {
            const int v = queue[k];
            double part = 0.0;
            for (int64_t e = R[v]; e < R[v + 1]; e++) {
                const int w = C[e];
                if (dist[w] == dist[v] + 1) part += sigma[v] / sigma[w] * (1.0 + delta[w]);
            }
            delta[v] = part;
        }
        for (int v = 0; v < N; v++)
            if (v != s) bc[v] += delta[v];
    }
    free(dist);
    free(queue);
    free(sigma);
    free(delta);
}


/* ------------------------------------------------------------ weighted Lp */

/* The cost matrix of FormBlock.cpp:45-75 on an Np x Np row-major array. */
static void fw_costs(const int64_t *R, const int *C, const float *V, int N, int Np, float *D)
{
    const float inf = INFINITY;
    for (int64_t e = 0; e < (int64_t)Np * Np; e++) D[e] = inf;
    for (int i = 0; i < N; i++)
        for (int64_t e = R[i]; e < R[i + 1]; e++) D[(int64_t)i * Np + C[e]] = V ? 1.0f / V[e] : 1.0f;
    for (int i = 0; i < Np; i++) D[(int64_t)i * Np + i] = 0.f;
}

static inline void fw_relax(float *d, float a, float b)
{
    const float t = a + b;
    if (t < *d) *d = t;
}

/* Blocked Floyd-Warshall with block size B (Np a multiple of B). */
static void fw_blocked(float *D, int Np, int B)
{
    const int nb = Np / B;
#define DD(i, j) D[(int64_t)(i) * Np + (j)]
    for (int K = 0; K < nb; K++) {
        const int k0 = K * B, k1 = k0 + B;
        for (int k = k0; k < k1; k++)
            for (int i = k0; i < k1; i++)
                for (int j = k0; j < k1; j++) fw_relax(&DD(i, j), DD(i, k), DD(k, j));
        for (int X = 0; X < nb; X++) {
            if (X == K) continue;
            const int x0 = X * B, x1 = x0 + B;
            for (int k = k0; k < k1; k++)
                for (int i = k0; i < k1; i++)
                    for (int j = x0; j < x1; j++) fw_relax(&DD(i, j), DD(i, k), DD(k, j));
            for (int k = k0; k < k1; k++)
                for (int i = x0; i < x1; i++)
                    for (int j = k0; j < k1; j++) fw_relax(&DD(i, j), DD(i, k), DD(k, j));
        }
        for (int I = 0; I < nb; I++) {
            if (I == K) continue;
            for (int J = 0; J < nb; J++) {
                if (J == K) continue;
                /* a min is exact: the k order of the terms does not matter */
                for (int i = I * B; i < (I + 1) * B; i++)
                    for (int k = k0; k < k1; k++) {
                        const float a = DD(i, k);
                        for (int j = J * B; j < (J + 1) * B; j++) fw_relax(&DD(i, j), a, DD(k, j));
                    }
            }
        }
    }
#undef DD
}

/* Returns Lp; eff (N floats), isolated and dist (N x N, row-major) optional.
 * B: the block size (a multiple-of-B padding of N is added). */
double orc_lpw(const int64_t *R, const int *C, const float *V, int N, int B, float *eff, int *isolated,
               float *dist)
{
    const int Np = (N + B - 1) / B * B;
    float *D = (float *)malloc(sizeof(float) * (size_t)Np * (size_t)Np);
    fw_costs(R, C, V, N, Np, D);
    fw_blocked(D, Np, B);
    double Lp_result = 0.0;
    int iso = 0;
    for (int u = 0; u < N; u++) {
        float li = 0.f;
        double rs = 0.0;
        for (int v = 0; v < N; v++) {
            const float d = D[(int64_t)u * Np + v];
            if (v != u && !((double)d < 0.1)) {
                li += 1.0f / d;
                rs += (double)(1.0f / d);
            }
        }
        Lp_result += rs;
        li /= (float)(N - 1);
        if (li < 1.0f / (float)N) {
            li = 0.f;
            iso++;
        }
        if (eff) eff[u] = li;
    }
    if (dist)
        for (int u = 0; u < N; u++) memcpy(dist + (int64_t)u * N, D + (int64_t)u * Np, sizeof(float) * (size_t)N);
    free(D);
    if (isolated) *isolated = iso;
    Lp_result /= N;
    Lp_result /= N - 1;
    return 1 / Lp_result;
}

/* Dijkstra (dense O(N^2) selection) from every vertex over the costs
 * (double)(1.0f / w) (1 without weights), path sums in double; dist N x N,
 * +inf where there is no path. */
void orc_sssp64(const int64_t *R, const int *C, const float *V, int N, double *dist)
{
    char *done = (char *)malloc((size_t)N);
    for (int s = 0; s < N; s++) {
        double *d = dist + (int64_t)s * N;
        for (int v = 0; v < N; v++) {
            d[v] = INFINITY;
            done[v] = 0;
        }
        d[s] = 0.0;
        for (;;) {
            int u = -1;
            for (int v = 0; v < N; v++)
                if (!done[v] && d[v] < INFINITY && (u < 0 || d[v] < d[u])) u = v;
            if (u < 0) break;
            done[u] = 1;
            for (int64_t e = R[u]; e < R[u + 1]; e++) {
                const int w = C[e];
                const double c = V ? (double)(1.0f / V[e]) : 1.0;
                if (w != u && d[u] + c < d[w]) d[w] = d[u] + c;
            }
        }
    }
    free(done);
}
