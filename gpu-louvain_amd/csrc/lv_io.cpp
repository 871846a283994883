// Host-side input and synthetic graphs of the Louvain library.
//
// lv_read_mtx restates readInputData (utils.cu:9-62): lines starting with
// '%' are skipped, the first other line is "rows cols nnz" with V := cols,
// then nnz triples "i j w" (1-based); row i receives (j,w) and, when i != j,
// row j receives (i,w); rows keep insertion order, nothing is deduplicated.
// Additions: "pattern" files read w = 1; out-of-range indices and short
// files are reported instead of read past.
//
// Generators are seeded and deterministic for any thread count: randomness
// is drawn per fixed work unit (a chunk of RMAT samples, a block pair of the
// planted partition) from xoshiro256** streams seeded by SplitMix64.  Real
// weights are (1 + floor(r * 2^24)) * 2^-24 in (0,1], a function of the
// seed and the unordered vertex pair, so every FP64 sum over them is exact.
#include <parallel/algorithm>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <omp.h>

#include "../../include/louvain.h"
#include "lv_error.h"

namespace {

lv_status io_fail(lv_status st, const std::string &m) {
    lv::set_error(m);
    return st;
}

inline uint64_t splitmix64(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Xoshiro {
    uint64_t s[4];
    explicit Xoshiro(uint64_t seed) {
        uint64_t x = seed;
        for (auto &v : s) v = splitmix64(x);
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    inline uint64_t next() {
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
    inline double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0,1)
};

inline uint64_t mix_seed(uint64_t seed, uint64_t a, uint64_t b) {
    uint64_t x = seed ^ (a * 0xD1B54A32D192ED03ull) ^ (b * 0xABC98388FB8FAC03ull);
    splitmix64(x);
    return splitmix64(x);
}

inline float pair_weight(uint64_t seed, uint32_t x, uint32_t y) {
    const uint32_t lo = x < y ? x : y, hi = x < y ? y : x;
    const uint64_t h = mix_seed(seed ^ 0x5bd1e995ull, lo, hi);
    return (float)((double)(1 + (h >> 40)) * 0x1.0p-24);
}

// Symmetric CSR from unique undirected pairs (lo << 32 | hi, lo < hi), rows
// sorted by neighbour id.
lv_status csr_from_pairs(const std::vector<uint64_t> &pairs, int32_t n, uint64_t seed, int unit,
                         int32_t **rowptr, int32_t **col, float **w, int32_t *V, int64_t *E) {
    const int64_t m = (int64_t)pairs.size();
    if (2 * m > 0x7fffffffLL) return io_fail(LV_ERR_ARG, "graph exceeds int32 CSR entries");
    int32_t *rp = (int32_t *)std::malloc(sizeof(int32_t) * ((size_t)n + 1));
    int32_t *cl = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(2 * m, 1));
    float *wt = (float *)std::malloc(sizeof(float) * (size_t)std::max<int64_t>(2 * m, 1));
    if (!rp || !cl || !wt) {
        std::free(rp);
        std::free(cl);
        std::free(wt);
        return io_fail(LV_ERR_OOM, "host allocation failed");
    }
    std::vector<int32_t> deg((size_t)n + 1, 0);
    for (int64_t i = 0; i < m; ++i) {
        deg[(uint32_t)(pairs[i] >> 32)]++;
        deg[(uint32_t)pairs[i]]++;
    }
    rp[0] = 0;
    for (int32_t v = 0; v < n; ++v) rp[v + 1] = rp[v] + deg[v];
    std::vector<int32_t> fill(rp, rp + n);
    for (int64_t i = 0; i < m; ++i) {
        const uint32_t a = (uint32_t)(pairs[i] >> 32), b = (uint32_t)pairs[i];
        const float x = unit ? 1.0f : pair_weight(seed, a, b);
        cl[fill[a]] = (int32_t)b;
        wt[fill[a]++] = x;
        cl[fill[b]] = (int32_t)a;
        wt[fill[b]++] = x;
    }
#pragma omp parallel for schedule(dynamic, 4096)
    for (int32_t v = 0; v < n; ++v) {
        const int32_t lo = rp[v], hi = rp[v + 1];
        if (hi - lo < 2) continue;
        std::vector<std::pair<int32_t, float>> row(hi - lo);
        for (int32_t e = lo; e < hi; ++e) row[e - lo] = {cl[e], wt[e]};
        std::sort(row.begin(), row.end(),
                  [](const std::pair<int32_t, float> &p, const std::pair<int32_t, float> &q) {
                      return p.first < q.first;
                  });
        for (int32_t e = lo; e < hi; ++e) {
            cl[e] = row[e - lo].first;
            wt[e] = row[e - lo].second;
        }
    }
    *rowptr = rp;
    *col = cl;
    *w = wt;
    *V = n;
    *E = 2 * m;
    return LV_OK;
}

}  // namespace

extern "C" {

void lv_free_host(void *p) { std::free(p); }

lv_status lv_read_mtx(const char *path, int32_t **rowptr, int32_t **col, float **w, int32_t *V,
                      int64_t *E) {
    if (!path || !rowptr || !col || !w || !V || !E) return io_fail(LV_ERR_ARG, "null argument");
    FILE *f = std::fopen(path, "rb");
    if (!f) return io_fail(LV_ERR_IO, std::string("cannot open '") + path + "'");
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    std::string buf;
    buf.resize(sz > 0 ? (size_t)sz : 0);
    if (sz > 0 && std::fread(&buf[0], 1, (size_t)sz, f) != (size_t)sz) {
        std::fclose(f);
        return io_fail(LV_ERR_IO, std::string("cannot read '") + path + "'");
    }
    std::fclose(f);
    const char *p = buf.c_str();
    const char *end = p + buf.size();
    bool pattern = false, have = false;
    long long rows = 0, cols = 0, nnz = 0;
    while (p < end) {
        const char *nl = (const char *)std::memchr(p, '\n', (size_t)(end - p));
        const char *le = nl ? nl : end;
        if (*p == '%') {
            std::string line(p, le);
            if (line.rfind("%%MatrixMarket", 0) == 0 && line.find("pattern") != std::string::npos)
                pattern = true;
        } else if (p < le && *p != '\r') {
            std::string line(p, le);
            if (std::sscanf(line.c_str(), "%lld %lld %lld", &rows, &cols, &nnz) != 3)
                return io_fail(LV_ERR_ARG, std::string("bad MatrixMarket size line in '") + path + "'");
            have = true;
            p = nl ? nl + 1 : end;
            break;
        }
        p = nl ? nl + 1 : end;
    }
    (void)rows;
    if (!have || cols < 0 || nnz < 0 || cols > 0x7fffffffLL)
        return io_fail(LV_ERR_ARG, std::string("missing or invalid size line in '") + path + "'");
    const int32_t n = (int32_t)cols;
    std::vector<int32_t> ti((size_t)nnz), tj((size_t)nnz);
    std::vector<float> tw((size_t)nnz);
    std::vector<int64_t> deg((size_t)n + 1, 0);
    char *q = const_cast<char *>(p);
    int64_t Ecount = 0;
    for (long long t = 0; t < nnz; ++t) {
        char *e1 = nullptr;
        errno = 0;
        long long a = std::strtoll(q, &e1, 10);
        if (e1 == q) return io_fail(LV_ERR_ARG, "'" + std::string(path) + "': expected " + std::to_string(nnz) + " entries, got " + std::to_string(t));
        q = e1;
        long long b = std::strtoll(q, &e1, 10);
        if (e1 == q) return io_fail(LV_ERR_ARG, "'" + std::string(path) + "': truncated entry " + std::to_string(t + 1));
        q = e1;
        float x = 1.0f;
        if (!pattern) {
            x = std::strtof(q, &e1);
            if (e1 == q) return io_fail(LV_ERR_ARG, "'" + std::string(path) + "': missing weight in entry " + std::to_string(t + 1));
            q = e1;
        }
        if (a < 1 || a > n || b < 1 || b > n)
            return io_fail(LV_ERR_ARG, "'" + std::string(path) + "': entry " + std::to_string(t + 1) + " index out of range");
        ti[t] = (int32_t)(a - 1);
        tj[t] = (int32_t)(b - 1);
        tw[t] = x;
        deg[ti[t]]++;
        Ecount++;
        if (ti[t] != tj[t]) {
            deg[tj[t]]++;
            Ecount++;
        }
    }
    if (Ecount > 0x7fffffffLL) return io_fail(LV_ERR_ARG, "graph exceeds int32 CSR entries");
    int32_t *rp = (int32_t *)std::malloc(sizeof(int32_t) * ((size_t)n + 1));
    int32_t *cl = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(Ecount, 1));
    float *wt = (float *)std::malloc(sizeof(float) * (size_t)std::max<int64_t>(Ecount, 1));
    if (!rp || !cl || !wt) {
        std::free(rp);
        std::free(cl);
        std::free(wt);
        return io_fail(LV_ERR_OOM, "host allocation failed");
    }
    rp[0] = 0;
    for (int32_t v = 0; v < n; ++v) rp[v + 1] = rp[v] + (int32_t)deg[v];
    std::vector<int64_t> fill(rp, rp + n);
    for (long long t = 0; t < nnz; ++t) {
        const int32_t a = ti[t], b = tj[t];
        cl[fill[a]] = b;
        wt[fill[a]++] = tw[t];
        if (a != b) {
            cl[fill[b]] = a;
            wt[fill[b]++] = tw[t];
        }
    }
    *rowptr = rp;
    *col = cl;
    *w = wt;
    *V = n;
    *E = Ecount;
    return LV_OK;
}

lv_status lv_write_mtx(const char *path, const int32_t *rowptr, const int32_t *col, const float *w,
                       int32_t V) {
    if (!path || !rowptr || V < 0) return io_fail(LV_ERR_ARG, "bad argument");
    FILE *f = std::fopen(path, "wb");
    if (!f) return io_fail(LV_ERR_IO, std::string("cannot create '") + path + "'");
    int64_t nnz = 0;
    for (int32_t v = 0; v < V; ++v)
        for (int32_t e = rowptr[v]; e < rowptr[v + 1]; ++e)
            if (col[e] <= v) nnz++;
    std::fprintf(f, "%%%%MatrixMarket matrix coordinate real symmetric\n");
    std::fprintf(f, "%d %d %lld\n", V, V, (long long)nnz);
    std::vector<char> line(64);
    for (int32_t v = 0; v < V; ++v)
        for (int32_t e = rowptr[v]; e < rowptr[v + 1]; ++e) {
            if (col[e] > v) continue;
            const float x = w[e];
            if (x == std::floor(x) && std::fabs(x) < 1e9f)
                std::fprintf(f, "%d %d %lld\n", v + 1, col[e] + 1, (long long)x);
            else
                std::fprintf(f, "%d %d %.9g\n", v + 1, col[e] + 1, (double)x);
        }
    if (std::fclose(f) != 0) return io_fail(LV_ERR_IO, std::string("cannot write '") + path + "'");
    return LV_OK;
}

lv_status lv_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                      uint64_t seed, int32_t unit_weights, int32_t **rowptr, int32_t **col,
                      float **w, int32_t *V, int64_t *E) {
    if (scale < 1 || scale > 30 || edge_factor < 1 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return io_fail(LV_ERR_ARG, "bad RMAT parameters");
    const int64_t n = (int64_t)1 << scale;
    const int64_t m = (int64_t)edge_factor * n;
    std::vector<uint64_t> keys((size_t)m);
    const int64_t kChunk = 1 << 16;
    const int64_t nchunks = (m + kChunk - 1) / kChunk;
    const double ab = a + b, abc = a + b + c;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t ch = 0; ch < nchunks; ++ch) {
        Xoshiro rng(mix_seed(seed, 0x524d4154ull, (uint64_t)ch));
        const int64_t lo = ch * kChunk, hi = std::min<int64_t>(m, lo + kChunk);
        for (int64_t i = lo; i < hi; ++i) {
            uint64_t u = 0, v = 0;
            for (int bit = scale - 1; bit >= 0; --bit) {
                const double r = rng.uniform();
                uint64_t ub = 0, vb = 0;
                if (r < a) {
                } else if (r < ab) {
                    vb = 1;
                } else if (r < abc) {
                    ub = 1;
                } else {
                    ub = 1;
                    vb = 1;
                }
                u |= ub << bit;
                v |= vb << bit;
            }
            keys[i] = (u == v) ? ~0ull : ((std::min(u, v) << 32) | std::max(u, v));
        }
    }
    __gnu_parallel::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    while (!keys.empty() && keys.back() == ~0ull) keys.pop_back();
    // seeded vertex permutation (hubs away from low ids)
    std::vector<uint32_t> perm((size_t)n);
    for (int64_t i = 0; i < n; ++i) perm[i] = (uint32_t)i;
    {
        Xoshiro rng(mix_seed(seed, 0x5045524dull, 0));
        for (int64_t i = n - 1; i > 0; --i) {
            const uint64_t j = rng.next() % (uint64_t)(i + 1);
            std::swap(perm[i], perm[j]);
        }
    }
    const int64_t mm = (int64_t)keys.size();
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < mm; ++i) {
        const uint32_t x = perm[(uint32_t)(keys[i] >> 32)], y = perm[(uint32_t)keys[i]];
        keys[i] = ((uint64_t)std::min(x, y) << 32) | std::max(x, y);
    }
    return csr_from_pairs(keys, (int32_t)n, seed, unit_weights, rowptr, col, w, V, E);
}

lv_status lv_gen_planted(int32_t n, int32_t blocks, double p_in, double p_out, uint64_t seed,
                         int32_t unit_weights, int32_t **rowptr, int32_t **col, float **w,
                         int32_t *V, int64_t *E) {
    if (n < 0 || blocks < 1 || blocks > std::max(n, 1) || p_in < 0 || p_in > 1 || p_out < 0 ||
        p_out > 1)
        return io_fail(LV_ERR_ARG, "bad planted-partition parameters");
    auto bstart = [&](int64_t bb) { return (int32_t)(bb * (int64_t)n / blocks); };
    const int64_t npairs = (int64_t)blocks * (blocks + 1) / 2;
    std::vector<std::vector<uint64_t>> parts((size_t)npairs);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t pi = 0; pi < npairs; ++pi) {
        // pair index -> (b1, b2), b1 <= b2, row-major over the upper triangle
        int64_t b1 = 0, rem = pi;
        while (rem >= blocks - b1) {
            rem -= blocks - b1;
            ++b1;
        }
        const int64_t b2 = b1 + rem;
        const int32_t s1 = bstart(b1), e1 = bstart(b1 + 1);
        const int32_t s2 = bstart(b2), e2 = bstart(b2 + 1);
        const double p = (b1 == b2) ? p_in : p_out;
        const int64_t n1 = e1 - s1, n2 = e2 - s2;
        const int64_t total = (b1 == b2) ? n1 * (n1 - 1) / 2 : n1 * n2;
        if (p <= 0.0 || total <= 0) continue;
        Xoshiro rng(mix_seed(seed, (uint64_t)b1 + 1, (uint64_t)b2 + 1));
        std::vector<uint64_t> &out = parts[pi];
        const double lq = p < 1.0 ? std::log1p(-p) : 0.0;
        int64_t k = -1;
        int64_t row = 0, row_start = 0;  // triangle walk state (b1 == b2)
        for (;;) {
            if (p >= 1.0) {
                k += 1;
            } else {
                double u = 1.0 - rng.uniform();  // (0,1]
                k += 1 + (int64_t)std::floor(std::log(u) / lq);
            }
            if (k >= total || k < 0) break;
            int64_t x, y;
            if (b1 == b2) {
                while (k >= row_start + (n1 - 1 - row)) {
                    row_start += n1 - 1 - row;
                    ++row;
                }
                x = s1 + row;
                y = s1 + row + 1 + (k - row_start);
            } else {
                x = s1 + k / n2;
                y = s2 + k % n2;
            }
            out.push_back(((uint64_t)std::min(x, y) << 32) | (uint64_t)std::max(x, y));
        }
    }
    std::vector<uint64_t> pairs;
    size_t tot = 0;
    for (auto &pp : parts) tot += pp.size();
    pairs.reserve(tot);
    for (auto &pp : parts) pairs.insert(pairs.end(), pp.begin(), pp.end());
    return csr_from_pairs(pairs, n, seed, unit_weights, rowptr, col, w, V, E);
}

}  // extern "C"
