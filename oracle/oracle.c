/*
 * CPU oracle for the Louvain hot path: TEST INFRASTRUCTURE ONLY (see
 * oracle.h).  Sequential C99 restatement of olearczuk/gpu-louvain; optional
 * OpenMP parallelism is used only inside a degree group, where decisions are
 * independent (Jacobi), so results do not depend on the thread count.
 *
 * Build with -ffp-contract=off: the fp32 gain must be evaluated without
 * fused multiply-adds, in the reference's operation order.
 */
#include "oracle.h"

#include <limits.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static char g_err[512];

static void set_err(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

const char *orc_last_error(void) { return g_err; }

/* Degree-group bounds (modularity_optimisation.cuh:7-8): group b holds the
 * vertices with bounds[b] < deg <= bounds[b+1]; deg-0 vertices never move. */
static const int32_t k_bounds[ORC_NGROUPS + 1] = {0, 4, 8, 16, 32, 84, 319, INT_MAX};

int orc_group_of_degree(int32_t deg) {
    if (deg <= 0) return -1;
    for (int b = 0; b < ORC_NGROUPS; ++b)
        if (deg > k_bounds[b] && deg <= k_bounds[b + 1]) return b;
    return ORC_NGROUPS - 1;
}

/* ------------------------------------------------------------------ */
/* Input: utils.cu:9-62.  Lines whose first character is '%' are skipped;
 * the first other line holds "rows cols nnz" and V := cols (utils.cu:18).
 * Then nnz triples "i j w" (1-based).  Row i gets (j,w); if i != j row j
 * also gets (i,w) (utils.cu:32-45).  Rows keep insertion order; no dedup.
 * Extensions (documented): "pattern" banners read "i j" with w = 1; blank
 * lines before the header are skipped; indices outside [1,V] and short
 * files are errors instead of undefined behaviour. */
int orc_read_mtx(const char *path, orc_csr *out) {
    memset(out, 0, sizeof *out);
    FILE *f = fopen(path, "r");
    if (!f) {
        set_err("cannot open '%s'", path);
        return -1;
    }
    char line[4096];
    int pattern = 0, have_header = 0;
    long rows = 0, cols = 0, nnz = 0;
    while (fgets(line, sizeof line, f)) {
        if (line[0] == '%') {
            if (strncmp(line, "%%MatrixMarket", 14) == 0 && strstr(line, "pattern")) pattern = 1;
            continue;
        }
        if (line[0] == '\n' || line[0] == '\r' || line[0] == 0) continue;
        if (sscanf(line, "%ld %ld %ld", &rows, &cols, &nnz) != 3) {
            set_err("bad MatrixMarket size line in '%s'", path);
            fclose(f);
            return -2;
        }
        have_header = 1;
        break;
    }
    if (!have_header || cols < 0 || nnz < 0 || cols > INT_MAX) {
        set_err("missing or invalid size line in '%s'", path);
        fclose(f);
        return -2;
    }
    (void)rows;
    int32_t V = (int32_t)cols;
    int32_t *ti = malloc(sizeof(int32_t) * (size_t)(nnz ? nnz : 1));
    int32_t *tj = malloc(sizeof(int32_t) * (size_t)(nnz ? nnz : 1));
    float *tw = malloc(sizeof(float) * (size_t)(nnz ? nnz : 1));
    int64_t *deg = calloc((size_t)V + 1, sizeof(int64_t));
    if (!ti || !tj || !tw || !deg) {
        set_err("out of memory");
        fclose(f);
        free(ti); free(tj); free(tw); free(deg);
        return -3;
    }
    int64_t E = 0;
    for (long t = 0; t < nnz; ++t) {
        long a, b;
        float w = 1.0f;
        int got = pattern ? fscanf(f, "%ld %ld", &a, &b) : fscanf(f, "%ld %ld %f", &a, &b, &w);
        if (got != (pattern ? 2 : 3)) {
            set_err("'%s': expected %ld entries, got %ld", path, nnz, t);
            fclose(f);
            free(ti); free(tj); free(tw); free(deg);
            return -4;
        }
        if (a < 1 || a > V || b < 1 || b > V) {
            set_err("'%s': entry %ld index out of range", path, t + 1);
            fclose(f);
            free(ti); free(tj); free(tw); free(deg);
            return -5;
        }
        ti[t] = (int32_t)(a - 1);
        tj[t] = (int32_t)(b - 1);
        tw[t] = w;
        deg[ti[t]]++;
        E++;
        if (ti[t] != tj[t]) {
            deg[tj[t]]++;
            E++;
        }
    }
    fclose(f);
    if (E > INT_MAX) {
        set_err("'%s': %lld CSR entries exceed int32", path, (long long)E);
        free(ti); free(tj); free(tw); free(deg);
        return -6;
    }
    out->V = V;
    out->E = E;
    out->rowptr = malloc(sizeof(int32_t) * ((size_t)V + 1));
    out->col = malloc(sizeof(int32_t) * (size_t)(E ? E : 1));
    out->w = malloc(sizeof(float) * (size_t)(E ? E : 1));
    int64_t *fill = malloc(sizeof(int64_t) * ((size_t)V + 1));
    out->rowptr[0] = 0;
    for (int32_t v = 0; v < V; ++v) out->rowptr[v + 1] = out->rowptr[v] + (int32_t)deg[v];
    for (int32_t v = 0; v < V; ++v) fill[v] = out->rowptr[v];
    for (long t = 0; t < nnz; ++t) {
        int32_t a = ti[t], b = tj[t];
        out->col[fill[a]] = b;
        out->w[fill[a]++] = tw[t];
        if (a != b) {
            out->col[fill[b]] = a;
            out->w[fill[b]++] = tw[t];
        }
    }
    free(fill); free(ti); free(tj); free(tw); free(deg);
    return 0;
}

void orc_free_csr(orc_csr *g) {
    free(g->rowptr);
    free(g->col);
    free(g->w);
    memset(g, 0, sizeof *g);
}

/* M = (sum of all CSR weights)/2 (utils.cu:36-46: each non-loop triple adds
 * 2w, a self-loop adds w, then M /= 2).  FP64 instead of the reference's
 * fp32 accumulation, which stalls at 2^24. */
double orc_total_weight(const orc_csr *g) {
    double s = 0.0;
    for (int64_t e = 0; e < g->E; ++e) s += (double)g->w[e];
    return s / 2.0;
}

/* Weight grid.  No reference counterpart: the reference sums fp32 weights
 * with fp32 atomics in thread order (modularity_optimisation.cu:144-153,
 * community_aggregation.cu:146-159).  The HIP library rounds every weight to
 * a grid of 2^-s before a run, so that every FP64 sum of weights is exact;
 * the oracle applies the same rounding to its input:
 *   T = sum of |w|, FP64, in chunks of 65536 entries (in order inside a chunk,
 *       chunk sums added in order);
 *   s = 50 - c with c the smallest integer such that 2^c >= T * (1 + 2^-20)
 *       (s = 0 when T = 0);
 *   q(w) = rint(w * 2^s) * 2^-s, round half to even. */
int32_t orc_weight_shift(const float *w, int64_t n) {
    const int64_t chunk = 65536;
    double T = 0.0;
    for (int64_t lo = 0; lo < n; lo += chunk) {
        double part = 0.0;
        const int64_t hi = lo + chunk < n ? lo + chunk : n;
        for (int64_t e = lo; e < hi; ++e) part += fabs((double)w[e]);
        T += part;
    }
    if (!(T > 0.0)) return 0;
    const double target = T * (1.0 + 1.0 / 1048576.0);
    int c = 0;
    double p = 1.0;
    while (p < target) {
        p *= 2.0;
        ++c;
    }
    while (p / 2.0 >= target) {
        p /= 2.0;
        --c;
    }
    return 50 - c;
}

float orc_quantize_weight(float w, int32_t shift) {
    volatile double x = scalbn((double)w, shift);
    return (float)scalbn(rint(x), -shift);
}

int64_t orc_quantize(orc_csr *g) {
    const int32_t shift = orc_weight_shift(g->w, g->E);
    int64_t changed = 0;
    for (int64_t e = 0; e < g->E; ++e) {
        const float q = orc_quantize_weight(g->w[e], shift);
        if (memcmp(&q, &g->w[e], sizeof q) != 0) ++changed;
        g->w[e] = q;
    }
    return changed;
}

/* k_v = sum of row weights, self-loop counted once (modularity_optimisation.cu:23-40). */
void orc_strength(const orc_csr *g, double *k) {
    for (int32_t v = 0; v < g->V; ++v) {
        double s = 0.0;
        for (int32_t e = g->rowptr[v]; e < g->rowptr[v + 1]; ++e) s += (double)g->w[e];
        k[v] = s;
    }
}

/* Community totals rebuilt from scratch in vertex order
 * (modularity_optimisation.cu:49-55 with the fill at :328-330). */
void orc_sigma(int32_t V, const int32_t *comm, const double *k, double *sigma) {
    for (int32_t c = 0; c < V; ++c) sigma[c] = 0.0;
    for (int32_t v = 0; v < V; ++v) sigma[comm[v]] += k[v];
}

void orc_sizes(int32_t V, const int32_t *comm, int32_t *size) {
    for (int32_t c = 0; c < V; ++c) size[c] = 0;
    for (int32_t v = 0; v < V; ++v) size[comm[v]]++;
}

/* computeGain (modularity_optimisation.cu:96-102), fp32, same order:
 *   K/M + k*((S_cur - k) - S_c) / (2*M*M). */
float orc_gain(float K_c, float M32, float k_v, float sigma_cur, float sigma_c) {
    volatile float cs = sigma_cur - k_v;
    volatile float num = k_v * (cs - sigma_c);
    volatile float den = 2.0f * M32 * M32;
    volatile float a = K_c / M32;
    volatile float b = num / den;
    return a + b;
}

/* Scratch for decide(): dense accumulators indexed by community. */
typedef struct {
    double *acc;
    int32_t *touched;
    unsigned char *mark;
} orc_scratch;

static int scratch_init(orc_scratch *s, int32_t V) {
    s->acc = calloc((size_t)V + 1, sizeof(double));
    s->touched = malloc(sizeof(int32_t) * ((size_t)V + 1));
    s->mark = calloc((size_t)V + 1, 1);
    return (s->acc && s->touched && s->mark) ? 0 : -1;
}

static void scratch_free(orc_scratch *s) {
    free(s->acc);
    free(s->touched);
    free(s->mark);
}

/* computeMove (modularity_optimisation.cu:111-198) for one vertex:
 *   K[c] = sum of w over row entries (u,w), u != v, comm[u] = c   (:144-153)
 *   candidates c != cur with (c < cur || size[c] > 1 || size[cur] > 1) (:154-155)
 *   best = argmax g(c), ties -> lowest c, starting from (0, cur)  (:138,157,169)
 *   move iff bestGain - K[cur]/M > 0                            (:192-196) */
static int32_t decide_vertex(const orc_csr *g, float M32, int32_t v, const int32_t *comm,
                             const int32_t *size, const double *k, const double *sigma,
                             orc_scratch *s, float *margin, float *kown, double *gap) {
    const int32_t cur = comm[v];
    int32_t nt = 0;
    for (int32_t e = g->rowptr[v]; e < g->rowptr[v + 1]; ++e) {
        int32_t u = g->col[e];
        if (u == v) continue;
        int32_t c = comm[u];
        if (!s->mark[c]) {
            s->mark[c] = 1;
            s->touched[nt++] = c;
        }
        s->acc[c] += (double)g->w[e];
    }
    const double kcur64 = s->mark[cur] ? s->acc[cur] : 0.0;
    const float k32 = (float)k[v];
    const float scur = (float)sigma[cur];
    float best_gain = 0.0f;
    int32_t best = cur;
    /* the two largest of {K[cur]/M} and the eligible candidates' gains */
    const float stay = (float)kcur64 / M32;
    double top1 = stay, top2 = -INFINITY;
    for (int32_t i = 0; i < nt; ++i) {
        int32_t c = s->touched[i];
        if (c == cur) continue;
        if (!(c < cur || size[c] > 1 || size[cur] > 1)) continue;
        float gc = orc_gain((float)s->acc[c], M32, k32, scur, (float)sigma[c]);
        if (gc > best_gain || (gc == best_gain && c < best)) {
            best_gain = gc;
            best = c;
        }
        if ((double)gc > top1) {
            top2 = top1;
            top1 = gc;
        } else if ((double)gc > top2) {
            top2 = gc;
        }
    }
    if (gap) {
        const double mag = fmax(fabs(top1), fabs(top2));
        gap[v] = top2 == -INFINITY ? INFINITY : (mag > 0.0 ? (top1 - top2) / mag : 0.0);
    }
    for (int32_t i = 0; i < nt; ++i) {
        s->acc[s->touched[i]] = 0.0;
        s->mark[s->touched[i]] = 0;
    }
    volatile float t = (float)kcur64 / M32;
    volatile float d = best_gain - t;
    if (margin) margin[v] = d;
    if (kown) kown[v] = (float)kcur64;
    return d > 0.0f ? best : cur;
}

static int nthreads_eff(int nthreads) {
#ifdef _OPENMP
    return nthreads > 0 ? nthreads : 1;
#else
    (void)nthreads;
    return 1;
#endif
}

/* Decide for a list of vertices; parallel over the list (Jacobi within a group). */
static void decide_list(const orc_csr *g, float M32, const int32_t *list, int32_t n,
                        const int32_t *comm, const int32_t *size, const double *k,
                        const double *sigma, int32_t *new_comm, orc_scratch *scr, int nthr,
                        const uint8_t *active, float *margin, float *kown, double *gap) {
#ifdef _OPENMP
    if (nthr > 1 && n > 256) {
#pragma omp parallel num_threads(nthr)
        {
            orc_scratch *s = &scr[omp_get_thread_num()];
#pragma omp for schedule(dynamic, 256)
            for (int32_t i = 0; i < n; ++i)
                new_comm[list[i]] = (active && !active[list[i]])
                                        ? comm[list[i]]
                                        : decide_vertex(g, M32, list[i], comm, size, k, sigma, s,
                                                        margin, kown, gap);
        }
        return;
    }
#else
    (void)nthr;
#endif
    for (int32_t i = 0; i < n; ++i)
        new_comm[list[i]] = (active && !active[list[i]])
                                ? comm[list[i]]
                                : decide_vertex(g, M32, list[i], comm, size, k, sigma, &scr[0],
                                                margin, kown, gap);
}

/* Vertex lists per degree group, ascending vertex id. */
typedef struct {
    int32_t *ids;
    int32_t off[ORC_NGROUPS + 1];
} orc_groups;

static void build_groups(const orc_csr *g, orc_groups *gr) {
    int32_t cnt[ORC_NGROUPS] = {0};
    for (int32_t v = 0; v < g->V; ++v) {
        int b = orc_group_of_degree(g->rowptr[v + 1] - g->rowptr[v]);
        if (b >= 0) cnt[b]++;
    }
    gr->off[0] = 0;
    for (int b = 0; b < ORC_NGROUPS; ++b) gr->off[b + 1] = gr->off[b] + cnt[b];
    gr->ids = malloc(sizeof(int32_t) * ((size_t)gr->off[ORC_NGROUPS] + 1));
    int32_t pos[ORC_NGROUPS];
    for (int b = 0; b < ORC_NGROUPS; ++b) pos[b] = gr->off[b];
    for (int32_t v = 0; v < g->V; ++v) {
        int b = orc_group_of_degree(g->rowptr[v + 1] - g->rowptr[v]);
        if (b >= 0) gr->ids[pos[b]++] = v;
    }
}

void orc_decide_group(const orc_csr *g, double M, const int32_t *comm, int group,
                      int32_t *new_comm, int nthreads) {
    orc_decide_group_obs(g, M, comm, group, new_comm, NULL, NULL, nthreads);
}

void orc_decide_group_obs(const orc_csr *g, double M, const int32_t *comm, int group,
                          int32_t *new_comm, float *margin, float *kown, int nthreads) {
    orc_decide_group_gap(g, M, comm, group, new_comm, margin, kown, NULL, nthreads);
}

void orc_decide_group_gap(const orc_csr *g, double M, const int32_t *comm, int group,
                          int32_t *new_comm, float *margin, float *kown, double *gap,
                          int nthreads) {
    const int32_t V = g->V;
    int nthr = nthreads_eff(nthreads);
    double *k = malloc(sizeof(double) * ((size_t)V + 1));
    double *sigma = malloc(sizeof(double) * ((size_t)V + 1));
    int32_t *size = malloc(sizeof(int32_t) * ((size_t)V + 1));
    orc_strength(g, k);
    orc_sigma(V, comm, k, sigma);
    orc_sizes(V, comm, size);
    orc_scratch *scr = calloc((size_t)nthr, sizeof(orc_scratch));
    for (int t = 0; t < nthr; ++t) scratch_init(&scr[t], V);
    orc_groups gr;
    build_groups(g, &gr);
    for (int32_t v = 0; v < V; ++v) new_comm[v] = comm[v];
    const float M32 = (float)M;
    int lo = group < 0 ? 0 : group, hi = group < 0 ? ORC_NGROUPS - 1 : group;
    for (int b = lo; b <= hi; ++b)
        decide_list(g, M32, gr.ids + gr.off[b], gr.off[b + 1] - gr.off[b], comm, size, k, sigma,
                    new_comm, scr, nthr, NULL, margin, kown, gap);
    for (int t = 0; t < nthr; ++t) scratch_free(&scr[t]);
    free(scr);
    free(gr.ids);
    free(k);
    free(sigma);
    free(size);
}

/* Modularity (modularity_optimisation.cu:365-401), FP64:
 *   Q = (sum_v sum_{u in row v, comm[u]=comm[v]} w)/(2M) - (sum_c S_c^2)/(4M^2)
 * Self-loop entries count in the first sum. */
double orc_modularity(const orc_csr *g, double M, const int32_t *comm) {
    const int32_t V = g->V;
    double *k = malloc(sizeof(double) * ((size_t)V + 1));
    double *sigma = malloc(sizeof(double) * ((size_t)V + 1));
    orc_strength(g, k);
    orc_sigma(V, comm, k, sigma);
    double own = 0.0, sq = 0.0;
    for (int32_t v = 0; v < V; ++v) {
        const int32_t cv = comm[v];
        for (int32_t e = g->rowptr[v]; e < g->rowptr[v + 1]; ++e)
            if (comm[g->col[e]] == cv) own += (double)g->w[e];
    }
    for (int32_t c = 0; c < V; ++c) sq += sigma[c] * sigma[c];
    free(k);
    free(sigma);
    return own / (2.0 * M) - sq / (4.0 * M * M);
}

/* One sweep (modularity_optimisation.cu:309-350): groups 0..6 in order; each
 * group decides from the state at its start (Jacobi), then applies moves
 * (:237-249) and rebuilds the community totals (:328-330). */
static int64_t sweep_impl(const orc_csr *g, float M32, int32_t *comm, int32_t *size,
                          const double *k, double *sigma, int32_t *new_comm,
                          const orc_groups *gr, orc_scratch *scr, int nthr,
                          const uint8_t *active) {
    int64_t moved = 0;
    for (int b = 0; b < ORC_NGROUPS; ++b) {
        const int32_t *list = gr->ids + gr->off[b];
        const int32_t n = gr->off[b + 1] - gr->off[b];
        if (n == 0) continue;
        decide_list(g, M32, list, n, comm, size, k, sigma, new_comm, scr, nthr, active, NULL,
                    NULL, NULL);
        for (int32_t i = 0; i < n; ++i) {
            int32_t v = list[i];
            if (new_comm[v] != comm[v]) {
                size[comm[v]]--;
                size[new_comm[v]]++;
                comm[v] = new_comm[v];
                moved++;
            }
        }
        orc_sigma(g->V, comm, k, sigma);
    }
    return moved;
}

int64_t orc_sweep(const orc_csr *g, double M, int32_t *comm, int nthreads) {
    return orc_sweep_masked(g, M, comm, NULL, nthreads);
}

int64_t orc_sweep_masked(const orc_csr *g, double M, int32_t *comm, const uint8_t *active,
                         int nthreads) {
    const int32_t V = g->V;
    int nthr = nthreads_eff(nthreads);
    double *k = malloc(sizeof(double) * ((size_t)V + 1));
    double *sigma = malloc(sizeof(double) * ((size_t)V + 1));
    int32_t *size = malloc(sizeof(int32_t) * ((size_t)V + 1));
    int32_t *nc = malloc(sizeof(int32_t) * ((size_t)V + 1));
    orc_strength(g, k);
    orc_sigma(V, comm, k, sigma);
    orc_sizes(V, comm, size);
    orc_scratch *scr = calloc((size_t)nthr, sizeof(orc_scratch));
    for (int t = 0; t < nthr; ++t) scratch_init(&scr[t], V);
    orc_groups gr;
    build_groups(g, &gr);
    int64_t moved = sweep_impl(g, (float)M, comm, size, k, sigma, nc, &gr, scr, nthr, active);
    for (int t = 0; t < nthr; ++t) scratch_free(&scr[t]);
    free(scr);
    free(gr.ids);
    free(k); free(sigma); free(size); free(nc);
    return moved;
}

/* Phase (modularity_optimisation.cu:284-363):
 *   gain = Q_after - Q_before; repeat while gain >= min_gain (:308);
 *   changed |= gain > 0 (:353); moves of a negative-gain sweep are kept;
 *   at the end orig2comm[x] = comm[orig2comm[x]] (:361, :251-257).
 * Divergence: a sweep without positive gain also ends the phase.  For
 * min_gain > 0 this changes nothing; for min_gain <= 0 the reference can
 * loop forever (no move, or vertices swapping at zero gain), whereas here Q
 * strictly increases from sweep to sweep, so the phase terminates. */
int orc_phase(const orc_csr *g, double M, int32_t *comm, float min_gain,
              int32_t *orig2comm, int32_t V0, int32_t *sweeps_out, int nthreads) {
    const int32_t V = g->V;
    int nthr = nthreads_eff(nthreads);
    double *k = malloc(sizeof(double) * ((size_t)V + 1));
    double *sigma = malloc(sizeof(double) * ((size_t)V + 1));
    int32_t *size = malloc(sizeof(int32_t) * ((size_t)V + 1));
    int32_t *nc = malloc(sizeof(int32_t) * ((size_t)V + 1));
    orc_strength(g, k);
    orc_sigma(V, comm, k, sigma);
    orc_sizes(V, comm, size);
    orc_scratch *scr = calloc((size_t)nthr, sizeof(orc_scratch));
    for (int t = 0; t < nthr; ++t) scratch_init(&scr[t], V);
    orc_groups gr;
    build_groups(g, &gr);
    const float M32 = (float)M;
    int changed = 0;
    int32_t sweeps = 0;
    double q0 = orc_modularity(g, M, comm);
    for (;;) {
        int64_t moved = sweep_impl(g, M32, comm, size, k, sigma, nc, &gr, scr, nthr, NULL);
        sweeps++;
        double q1 = orc_modularity(g, M, comm);
        double gain = q1 - q0;
        if (gain > 0.0) changed = 1;
        if (moved == 0 || !(gain > 0.0)) break;
        if (!(gain >= (double)min_gain)) break;
        q0 = q1;
    }
    for (int32_t x = 0; x < V0; ++x) orig2comm[x] = comm[orig2comm[x]];
    if (sweeps_out) *sweeps_out = sweeps;
    for (int t = 0; t < nthr; ++t) scratch_free(&scr[t]);
    free(scr);
    free(gr.ids);
    free(k); free(sigma); free(size); free(nc);
    return changed;
}

static int cmp_i32(const void *a, const void *b) {
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/* Aggregation (community_aggregation.cu:248-347):
 *   present[c] = size[c] > 0; new_id = exclusive scan of present (:267-273);
 *   new row new_id[c] = distinct new_id[comm[u]] over all entries (u,w) of
 *   the rows of c's members, weights summed (:76-163); a community whose
 *   members have no edges becomes an isolated vertex; rows are emitted in
 *   ascending new id (the reference leaves the order unspecified);
 *   orig2comm[x] = new_id[orig2comm[x]] (:222-228). */
int orc_aggregate(const orc_csr *g, const int32_t *comm, orc_csr *out, int32_t *orig2comm,
                  int32_t V0, int32_t *new_id_out) {
    const int32_t V = g->V;
    int32_t *size = calloc((size_t)V + 1, sizeof(int32_t));
    int32_t *nid = malloc(sizeof(int32_t) * ((size_t)V + 1));
    int32_t *mstart = malloc(sizeof(int32_t) * ((size_t)V + 2));
    int32_t *members = malloc(sizeof(int32_t) * ((size_t)V + 1));
    for (int32_t v = 0; v < V; ++v) size[comm[v]]++;
    int32_t newV = 0;
    for (int32_t c = 0; c < V; ++c) nid[c] = size[c] > 0 ? newV++ : -1;
    mstart[0] = 0;
    for (int32_t c = 0; c < V; ++c) mstart[c + 1] = mstart[c] + size[c];
    {
        int32_t *pos = malloc(sizeof(int32_t) * ((size_t)V + 1));
        for (int32_t c = 0; c < V; ++c) pos[c] = mstart[c];
        for (int32_t v = 0; v < V; ++v) members[pos[comm[v]]++] = v;
        free(pos);
    }
    out->V = newV;
    out->rowptr = malloc(sizeof(int32_t) * ((size_t)newV + 1));
    out->col = malloc(sizeof(int32_t) * (size_t)(g->E ? g->E : 1));
    out->w = malloc(sizeof(float) * (size_t)(g->E ? g->E : 1));
    double *acc = calloc((size_t)newV + 1, sizeof(double));
    unsigned char *mark = calloc((size_t)newV + 1, 1);
    int32_t *touched = malloc(sizeof(int32_t) * ((size_t)newV + 1));
    int64_t pos = 0;
    out->rowptr[0] = 0;
    for (int32_t c = 0; c < V; ++c) {
        if (size[c] == 0) continue;
        int32_t nt = 0;
        for (int32_t m = mstart[c]; m < mstart[c + 1]; ++m) {
            int32_t v = members[m];
            for (int32_t e = g->rowptr[v]; e < g->rowptr[v + 1]; ++e) {
                int32_t key = nid[comm[g->col[e]]];
                if (!mark[key]) {
                    mark[key] = 1;
                    touched[nt++] = key;
                }
                acc[key] += (double)g->w[e];
            }
        }
        qsort(touched, (size_t)nt, sizeof(int32_t), cmp_i32); /* ascending new ids */
        for (int32_t i = 0; i < nt; ++i) {
            int32_t key = touched[i];
            out->col[pos] = key;
            out->w[pos] = (float)acc[key];
            pos++;
            acc[key] = 0.0;
            mark[key] = 0;
        }
        out->rowptr[nid[c] + 1] = (int32_t)pos;
    }
    out->E = pos;
    for (int32_t x = 0; x < V0; ++x) orig2comm[x] = nid[orig2comm[x]];
    if (new_id_out)
        for (int32_t c = 0; c < V; ++c) new_id_out[c] = nid[c];
    free(acc); free(mark); free(touched);
    free(size); free(nid); free(mstart); free(members);
    return 0;
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

/* Driver (gpulouvain.cu:30-44): loop { if !phase break; aggregate } then Q. */
int orc_run(const orc_csr *g0, float min_gain, int32_t *orig2comm, int32_t *level_V,
            orc_stats *stats, int nthreads) {
    double t0 = now_s();
    const int32_t V0 = g0->V;
    const double M = orc_total_weight(g0);
    orc_csr cur;
    cur.V = g0->V;
    cur.E = g0->E;
    cur.rowptr = malloc(sizeof(int32_t) * ((size_t)cur.V + 1));
    cur.col = malloc(sizeof(int32_t) * (size_t)(cur.E ? cur.E : 1));
    cur.w = malloc(sizeof(float) * (size_t)(cur.E ? cur.E : 1));
    memcpy(cur.rowptr, g0->rowptr, sizeof(int32_t) * ((size_t)cur.V + 1));
    memcpy(cur.col, g0->col, sizeof(int32_t) * (size_t)cur.E);
    memcpy(cur.w, g0->w, sizeof(float) * (size_t)cur.E);
    for (int32_t x = 0; x < V0; ++x) orig2comm[x] = x;
    int32_t *comm = malloc(sizeof(int32_t) * ((size_t)cur.V + 1));
    int32_t levels = 0, sweeps = 0;
    for (;;) {
        for (int32_t v = 0; v < cur.V; ++v) comm[v] = v;
        int32_t s = 0;
        int changed = orc_phase(&cur, M, comm, min_gain, orig2comm, V0, &s, nthreads);
        levels++;
        sweeps += s;
        if (!changed) break;
        orc_csr next;
        orc_aggregate(&cur, comm, &next, orig2comm, V0, NULL);
        orc_free_csr(&cur);
        cur = next;
    }
    double q = orc_modularity(&cur, M, comm);
    if (level_V) *level_V = cur.V;
    if (stats) {
        stats->levels = levels;
        stats->sweeps = sweeps;
        stats->final_V = cur.V;
        stats->modularity = q;
        stats->seconds = now_s() - t0;
    }
    free(comm);
    orc_free_csr(&cur);
    return 0;
}
