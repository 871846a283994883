// gpulouvain: command-line driver over the C-ABI (include/louvain.h), with
// the reference's flags and stdout (gpulouvain.cu:6-46, utils.cu:161-186,
// modularity_optimisation.cu:403-418):
//   gpulouvain -f mtx-matrix-file -g min-gain [-v]
// prints the final modularity ("%f"), then "algorithm_ms total_ms" (total
// adds the upload time), then with -v the last level's vertex count and one
// line per community: its id+1 followed by its original vertices (+1).
// Flags default to off (the reference leaves them uninitialised); errors go
// to stderr with a non-zero exit status instead of undefined behaviour.
#include <getopt.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/louvain.h"

static void usage() {
    std::printf("Usage: ./gpulouvain -f mtx-matrix-file -g min-gain [-v]\n");
    std::exit(1);
}

int main(int argc, char **argv) {
    bool have_f = false, have_g = false, verbose = false;
    float min_gain = 0.0f;
    const char *path = nullptr;
    int opt;
    while ((opt = getopt(argc, argv, "f:g:v")) != -1) {
        switch (opt) {
        case 'g':
            have_g = true;
            min_gain = std::strtof(optarg, nullptr);
            break;
        case 'v': verbose = true; break;
        case 'f':
            have_f = true;
            path = optarg;
            break;
        default: usage();
        }
    }
    if (!have_f || !have_g) usage();

    int32_t *rowptr = nullptr, *col = nullptr;
    float *w = nullptr;
    int32_t V = 0;
    int64_t E = 0;
    // a MatrixMarket file, or a binary CSR file written by lv_write_csr
    if (lv_read_graph(path, &rowptr, &col, &w, &V, &E) != LV_OK) {
        std::fprintf(stderr, "gpulouvain: %s\n", lv_last_error());
        return 2;
    }
    lv_options o{};
    lv_graph *g = nullptr;
    if (lv_create(rowptr, col, w, V, E, &o, &g) != LV_OK) {
        std::fprintf(stderr, "gpulouvain: %s\n", lv_last_error());
        return 3;
    }
    lv_stats st{};
    if (lv_run(g, min_gain, &st) != LV_OK) {
        std::fprintf(stderr, "gpulouvain: %s\n", lv_last_error());
        lv_destroy(g);
        return 3;
    }
    std::printf("%f\n", st.modularity);
    std::printf("%f %f\n", st.algo_ms, st.algo_ms + st.memory_ms);
    if (verbose) {
        std::vector<int32_t> o2c((size_t)V + 1);
        int32_t levelV = 0;
        if (lv_assignment(g, o2c.data(), &levelV) != LV_OK) {
            std::fprintf(stderr, "gpulouvain: %s\n", lv_last_error());
            lv_destroy(g);
            return 3;
        }
        std::vector<int32_t> start((size_t)levelV + 2, 0), items((size_t)V + 1);
        for (int32_t x = 0; x < V; ++x) start[o2c[x] + 1]++;
        for (int32_t c = 0; c < levelV; ++c) start[c + 1] += start[c];
        std::vector<int32_t> pos(start.begin(), start.end());
        for (int32_t x = 0; x < V; ++x) items[pos[o2c[x]]++] = x;
        std::printf("%d\n", levelV);
        for (int32_t c = 0; c < levelV; ++c) {
            std::printf("%d", c + 1);
            for (int32_t i = start[c]; i < start[c + 1]; ++i) std::printf(" %d", items[i] + 1);
            std::printf("\n");
        }
    }
    lv_destroy(g);
    lv_free_host(rowptr);
    lv_free_host(col);
    lv_free_host(w);
    return 0;
}
