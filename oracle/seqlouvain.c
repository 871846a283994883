/*
 * seqlouvain: command-line front end of the CPU oracle (TEST INFRASTRUCTURE
 * ONLY).  Same flags and stdout as gpulouvain (gpulouvain.cu:6-46,
 * utils.cu:161-186, modularity_optimisation.cu:403-418):
 *   seqlouvain -f file.mtx -g min_gain [-v] [-t threads]
 * prints Q ("%f"), then "algo_ms total_ms", then with -v the level vertex
 * count and one line per community: id+1 followed by its original vertices+1.
 */
#include "oracle.h"

#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static void usage(void) {
    printf("Usage: ./gpulouvain -f mtx-matrix-file -g min-gain [-v]\n");
    exit(1);
}

int main(int argc, char **argv) {
    int opt, verbose = 0, have_f = 0, have_g = 0, threads = 1;
    float min_gain = 0.0f;
    const char *path = NULL;
    while ((opt = getopt(argc, argv, "f:g:vt:")) != -1) {
        switch (opt) {
        case 'g': have_g = 1; min_gain = strtof(optarg, NULL); break;
        case 'v': verbose = 1; break;
        case 'f': have_f = 1; path = optarg; break;
        case 't': threads = atoi(optarg); break;
        default: usage();
        }
    }
    if (!have_f || !have_g) usage();
    orc_csr g;
    if (orc_read_mtx(path, &g) != 0) {
        fprintf(stderr, "error: %s\n", orc_last_error());
        return 2;
    }
    orc_quantize(&g); /* the library's weight grid (oracle.h) */
    int32_t *o2c = malloc(sizeof(int32_t) * ((size_t)g.V + 1));
    int32_t level_V = 0;
    orc_stats st;
    orc_run(&g, min_gain, o2c, &level_V, &st, threads);
    printf("%f\n", st.modularity);
    printf("%f %f\n", st.seconds * 1e3, st.seconds * 1e3);
    if (verbose) {
        /* group original vertices by final community, ascending */
        int32_t *cnt = calloc((size_t)level_V + 1, sizeof(int32_t));
        int32_t *start = calloc((size_t)level_V + 2, sizeof(int32_t));
        int32_t *items = malloc(sizeof(int32_t) * ((size_t)g.V + 1));
        for (int32_t x = 0; x < g.V; ++x) cnt[o2c[x]]++;
        for (int32_t c = 0; c < level_V; ++c) start[c + 1] = start[c] + cnt[c];
        for (int32_t c = 0; c < level_V; ++c) cnt[c] = start[c];
        for (int32_t x = 0; x < g.V; ++x) items[cnt[o2c[x]]++] = x;
        printf("%d\n", level_V);
        for (int32_t c = 0; c < level_V; ++c) {
            printf("%d", c + 1);
            for (int32_t i = start[c]; i < start[c + 1]; ++i) printf(" %d", items[i] + 1);
            printf("\n");
        }
        free(cnt); free(start); free(items);
    }
    free(o2c);
    orc_free_csr(&g);
    return 0;
}
