/*
 * CPU oracle for the Louvain hot path: TEST INFRASTRUCTURE ONLY.
 *
 * A sequential restatement (plain C99) of olearczuk/gpu-louvain's two-phase
 * method: modularity optimisation (bucket-synchronous sweeps) and community
 * aggregation, plus the modularity evaluation and the level driver.  It is the
 * checker the HIP path is compared against and the CPU baseline timed beside
 * it.  Only tests, the smoke check and the benchmark's CPU-baseline leg load
 * it; nothing on the product path calls into it.
 *
 * Each function cites the reference file:line it restates.  Arithmetic
 * conventions (shared with the HIP path):
 *   - M, vertex strengths k, community totals S, per-candidate weight sums K,
 *     merged aggregation weights and modularity are accumulated in FP64.
 *     With weights on a 2^-s grid and totals below 2^(53-s) these sums are
 *     exact, hence independent of summation order.  orc_quantize rounds a
 *     graph's weights to such a grid, as the HIP library does before a run.
 *   - The gain of a candidate community is evaluated in fp32 with the
 *     reference's operation order (modularity_optimisation.cu:96-102) from
 *     fp32 roundings of the FP64 quantities; build with -ffp-contract=off.
 *   - Aggregated rows are emitted in ascending new community id.
 */
#ifndef LOUVAIN_ORACLE_H
#define LOUVAIN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NGROUPS 7

typedef struct {
    int32_t V;
    int64_t E;
    int32_t *rowptr; /* [V+1] */
    int32_t *col;    /* [E]   */
    float *w;        /* [E]   */
} orc_csr;

typedef struct {
    int32_t levels;        /* phases run */
    int32_t sweeps;        /* sweeps over all levels */
    int32_t final_V;       /* vertices at the last level */
    double modularity;     /* final Q (FP64) */
    double seconds;        /* wall time of the run (excludes reading) */
} orc_stats;

/* --- input (utils.cu:9-62) ------------------------------------------- */
/* Returns 0 on success, negative on error (message in orc_last_error()). */
int orc_read_mtx(const char *path, orc_csr *out);
void orc_free_csr(orc_csr *g);
const char *orc_last_error(void);

/* --- weight grid (no reference counterpart; the HIP library's canonical
 * input rounding, under which every FP64 sum of weights is exact) ------ */
/* s = 50 - c, c the smallest integer with 2^c >= T*(1+2^-20), T = sum |w|
 * in FP64 over chunks of 65536 entries added in order (0 when T = 0). */
int32_t orc_weight_shift(const float *w, int64_t n);
/* rint(w * 2^shift) * 2^-shift (ties to even); exact in fp32. */
float orc_quantize_weight(float w, int32_t shift);
/* Rounds g's weights in place with the shift of g; returns how many changed. */
int64_t orc_quantize(orc_csr *g);

/* --- per-level quantities ------------------------------------------- */
double orc_total_weight(const orc_csr *g);                 /* M = (sum w)/2 */
void orc_strength(const orc_csr *g, double *k);            /* k_v = row sum */
int orc_group_of_degree(int32_t deg);                      /* -1 for deg 0 */
void orc_sigma(int32_t V, const int32_t *comm, const double *k, double *sigma);
void orc_sizes(int32_t V, const int32_t *comm, int32_t *size);

/* --- decision R3 (modularity_optimisation.cu:111-198) ----------------- */
/* Decide the best community for every vertex of degree group `group`
 * (0..6; -1 = every vertex of degree > 0) given the state `comm`, writing
 * new_comm[v] for those vertices and comm[v] for all others. */
void orc_decide_group(const orc_csr *g, double M, const int32_t *comm, int group,
                      int32_t *new_comm, int nthreads);
/* As orc_decide_group, also recording for each decided vertex the move
 * margin (bestGain - K[cur]/M in fp32, bestGain starting at 0 as in
 * computeMove) and K[cur] as fp32 (margin/kown may be NULL). */
void orc_decide_group_obs(const orc_csr *g, double M, const int32_t *comm, int group,
                          int32_t *new_comm, float *margin, float *kown, int nthreads);

/* As orc_decide_group_obs, also recording for each decided vertex how close
 * its decision was: with v1 >= v2 the two largest of {K[cur]/M (fp32)} and
 * the fp32 gains of its eligible candidates, gap[v] = (v1 - v2) /
 * max(|v1|, |v2|) (0 when both are 0, +inf without candidates).  A decision
 * can change under a perturbation of the sums only where gap is small. */
void orc_decide_group_gap(const orc_csr *g, double M, const int32_t *comm, int group,
                          int32_t *new_comm, float *margin, float *kown, double *gap,
                          int nthreads);

/* Gain of moving v to community c (fp32, reference operation order). */
float orc_gain(float K_c, float M32, float k_v, float sigma_cur, float sigma_c);

/* --- modularity R6 (modularity_optimisation.cu:365-401) --------------- */
double orc_modularity(const orc_csr *g, double M, const int32_t *comm);

/* --- one sweep / one phase R5 (modularity_optimisation.cu:284-363) --- */
/* One sweep over groups 0..6 in place; returns the number of moved
 * vertices.  size/sigma/k are maintained for the state. */
int64_t orc_sweep(const orc_csr *g, double M, int32_t *comm, int nthreads);
/* A sweep in which only vertices with active[v] != 0 are decided (the
 * others keep their community); active == NULL decides every vertex. */
int64_t orc_sweep_masked(const orc_csr *g, double M, int32_t *comm, const uint8_t *active,
                         int nthreads);

/* Runs sweeps while gain >= min_gain (and at least one vertex moved).
 * Updates orig2comm[x] = comm[orig2comm[x]] at the end and returns 1 if
 * any sweep had positive gain.  `sweeps_out` may be NULL. */
int orc_phase(const orc_csr *g, double M, int32_t *comm, float min_gain,
              int32_t *orig2comm, int32_t V0, int32_t *sweeps_out, int nthreads);

/* --- aggregation R7 (community_aggregation.cu:248-347) --------------- */
/* Builds the next level from (g, comm).  `out` receives freshly malloc'd
 * arrays; orig2comm is renumbered in place.  new_id (optional, [V]) gets
 * the renumbering (-1 for empty communities). */
int orc_aggregate(const orc_csr *g, const int32_t *comm, orc_csr *out,
                  int32_t *orig2comm, int32_t V0, int32_t *new_id);

/* --- driver R8 (gpulouvain.cu:30-44) --------------------------------- */
/* Full multi-level run.  orig2comm [V0] receives the final assignment in
 * the last level's vertex ids; *level_V the last level's vertex count. */
int orc_run(const orc_csr *g, float min_gain, int32_t *orig2comm, int32_t *level_V,
            orc_stats *stats, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
