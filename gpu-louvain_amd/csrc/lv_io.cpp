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
#include <climits>
#include <cstring>
#include <string>
#include <vector>

#include <fcntl.h>
#include <omp.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

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


// Token parsers of the reader.  parse_int: an optional sign and decimal
// digits only.  parse_float: the correctly rounded float of a decimal token;
// tokens of at most 19 significant digits and a decimal exponent within
// +-22 take one exact double operation (Clinger's fast path), then the
// double's rounding to float, which is exact unless the double lies on a
// midpoint between two floats (a midpoint between the token and the double
// would have been the nearer double); every other token goes to strtof.
inline bool parse_int(const char *t, size_t len, long long &out) {
    size_t i = 0;
    bool neg = false;
    if (i < len && (t[i] == '+' || t[i] == '-')) neg = t[i++] == '-';
    if (i >= len || len - i > 18) return false;
    long long v = 0;
    for (; i < len; ++i) {
        const unsigned d = (unsigned)(t[i] - '0');
        if (d > 9) return false;
        v = v * 10 + d;
    }
    out = neg ? -v : v;
    return true;
}

inline bool parse_float(char *t, size_t len, float &out) {
    static const double p10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,
                                   1e8,  1e9,  1e10, 1e11, 1e12, 1e13, 1e14, 1e15,
                                   1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
    size_t i = 0;
    bool neg = false;
    if (i < len && (t[i] == '+' || t[i] == '-')) neg = t[i++] == '-';
    uint64_t m = 0;
    int digits = 0, e10 = 0;
    bool any = false, fast = true;
    for (; i < len && (unsigned)(t[i] - '0') <= 9; ++i) {
        any = true;
        if (m == 0 && t[i] == '0') continue;
        if (digits < 19) {
            m = m * 10 + (unsigned)(t[i] - '0');
            ++digits;
        } else {
            fast = false;
        }
    }
    if (i < len && t[i] == '.') {
        for (++i; i < len && (unsigned)(t[i] - '0') <= 9; ++i) {
            any = true;
            if (m == 0 && t[i] == '0') {
                --e10;
                continue;
            }
            if (digits < 19) {
                m = m * 10 + (unsigned)(t[i] - '0');
                ++digits;
                --e10;
            } else {
                fast = false;
            }
        }
    }
    if (any && i < len && (t[i] == 'e' || t[i] == 'E')) {
        size_t j = i + 1;
        bool eneg = false;
        if (j < len && (t[j] == '+' || t[j] == '-')) eneg = t[j++] == '-';
        int ev = 0;
        bool ed = false;
        for (; j < len && (unsigned)(t[j] - '0') <= 9; ++j) {
            ed = true;
            if (ev < 10000) ev = ev * 10 + (t[j] - '0');
        }
        if (ed) {
            e10 += eneg ? -ev : ev;
            i = j;
        }
    }
    if (fast && any && i == len && m < (1ull << 53) && e10 >= -22 && e10 <= 22) {
        double d = (double)m;
        d = e10 < 0 ? d / p10[-e10] : d * p10[e10];
        // a float midpoint has zeros in the 28 low mantissa bits below the
        // half-ulp bit of float precision
        uint64_t bits;
        std::memcpy(&bits, &d, sizeof bits);
        const uint64_t low = bits & ((1ull << 29) - 1);
        if (d == 0.0 || low != (1ull << 28)) {
            out = (float)(neg ? -d : d);
            return true;
        }
    }
    char *e1 = nullptr;
    const char save = t[len];
    t[len] = '\0';
    out = std::strtof(t, &e1);
    t[len] = save;
    return e1 == t + len;
}

inline bool is_space(char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

// Read-only mapping of a whole file (empty files map to an empty range).
struct MappedFile {
    const char *data = nullptr;
    size_t size = 0;
    void *base = nullptr;
    bool open(const char *path) {
        const int fd = ::open(path, O_RDONLY);
        if (fd < 0) return false;
        struct stat st;
        if (::fstat(fd, &st) != 0) {
            ::close(fd);
            return false;
        }
        size = (size_t)st.st_size;
        if (size > 0) {
            base = ::mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
            if (base == MAP_FAILED) {
                base = nullptr;
                ::close(fd);
                return false;
            }
            ::madvise(base, size, MADV_SEQUENTIAL);
            data = (const char *)base;
        } else {
            static const char empty = 0;
            data = &empty;
        }
        ::close(fd);
        return true;
    }
    ~MappedFile() {
        if (base) ::munmap(base, size);
    }
};

// CSR of entries (i, j, w) in the reference's order: row i receives (j, w)
// and, when i != j, row j receives (i, w), each row in entry order
// (readInputData, utils.cu:29-52).  Threads fill contiguous entry ranges;
// per-range row counts give each range its place inside every row.
lv_status csr_in_entry_order(const std::vector<int32_t> &ti, const std::vector<int32_t> &tj,
                             const std::vector<float> &tw, int32_t n, int32_t **rowptr,
                             int32_t **col, float **w, int32_t *V, int64_t *E) {
    const int64_t m = (int64_t)ti.size();
    // up to 8 ranges of >= 64Ki entries, the counts within 1 GiB
    int T = (int)std::min<int64_t>(std::min(omp_get_max_threads(), 8), m / 65536);
    T = std::max(T, 1);
    while (T > 1 && (int64_t)T * ((int64_t)n + 1) * 4 > (int64_t)1 << 30) --T;
    std::vector<int32_t> cnt((size_t)T * ((size_t)n + 1), 0);
    std::vector<int64_t> lo(T + 1);
    for (int t = 0; t <= T; ++t) lo[t] = m * t / T;
#pragma omp parallel for num_threads(T) schedule(static, 1)
    for (int t = 0; t < T; ++t) {
        int32_t *c = cnt.data() + (size_t)t * ((size_t)n + 1);
        for (int64_t e = lo[t]; e < lo[t + 1]; ++e) {
            c[ti[e]]++;
            if (ti[e] != tj[e]) c[tj[e]]++;
        }
    }
    // rowptr, then each range's start inside every row (cnt turns into offsets)
    int32_t *rp = (int32_t *)std::malloc(sizeof(int32_t) * ((size_t)n + 1));
    if (!rp) return io_fail(LV_ERR_OOM, "host allocation failed");
    int64_t total = 0;
    for (int32_t v = 0; v < n; ++v) {
        if (total > 0x7fffffffLL) break;
        rp[v] = (int32_t)total;
        for (int t = 0; t < T; ++t) {
            int32_t &c = cnt[(size_t)t * ((size_t)n + 1) + v];
            const int32_t k = c;
            c = (int32_t)total;
            total += k;
        }
    }
    if (total > 0x7fffffffLL) {
        std::free(rp);
        return io_fail(LV_ERR_ARG, "graph exceeds int32 CSR entries");
    }
    rp[n] = (int32_t)total;
    int32_t *cl = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(total, 1));
    float *wt = (float *)std::malloc(sizeof(float) * (size_t)std::max<int64_t>(total, 1));
    if (!cl || !wt) {
        std::free(rp);
        std::free(cl);
        std::free(wt);
        return io_fail(LV_ERR_OOM, "host allocation failed");
    }
#pragma omp parallel for num_threads(T) schedule(static, 1)
    for (int t = 0; t < T; ++t) {
        int32_t *pos = cnt.data() + (size_t)t * ((size_t)n + 1);
        for (int64_t e = lo[t]; e < lo[t + 1]; ++e) {
            const int32_t a = ti[e], b = tj[e];
            cl[pos[a]] = b;
            wt[pos[a]++] = tw[e];
            if (a != b) {
                cl[pos[b]] = a;
                wt[pos[b]++] = tw[e];
            }
        }
    }
    *rowptr = rp;
    *col = cl;
    *w = wt;
    *V = n;
    *E = total;
    return LV_OK;
}

}  // namespace

extern "C" {

void lv_free_host(void *p) { std::free(p); }

lv_status lv_read_mtx(const char *path, int32_t **rowptr, int32_t **col, float **w, int32_t *V,
                      int64_t *E) {
    if (!path || !rowptr || !col || !w || !V || !E) return io_fail(LV_ERR_ARG, "null argument");
    const std::string where = std::string("'") + path + "'";
    MappedFile mf;
    if (!mf.open(path)) return io_fail(LV_ERR_IO, std::string("cannot open ") + where);
    const char *p = mf.data;
    const char *end = mf.data + mf.size;
    // header: '%' lines (a "pattern" banner means w = 1), then "rows cols nnz"
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
                return io_fail(LV_ERR_ARG, "bad MatrixMarket size line in " + where);
            have = true;
            p = nl ? nl + 1 : end;
            break;
        }
        p = nl ? nl + 1 : end;
    }
    (void)rows;
    if (!have || cols < 0 || nnz < 0 || cols > 0x7fffffffLL)
        return io_fail(LV_ERR_ARG, "missing or invalid size line in " + where);
    const int32_t n = (int32_t)cols;
    const int K = pattern ? 2 : 3;  // tokens per entry
    const int64_t need = (int64_t)nnz * K;

    // Entries are whitespace-separated tokens, as the reference's stream
    // extraction reads them.  Pass 1 counts the tokens of byte chunks cut at
    // whitespace; their prefix gives every token its entry and field, so
    // pass 2 parses the chunks independently.
    const int64_t region = (int64_t)(end - p);
    const int nt = std::max(1, omp_get_max_threads());
    int64_t chunk = std::max<int64_t>(region / (8 * (int64_t)nt) + 1, 1 << 16);
    if (const char *env = std::getenv("LV_MTX_CHUNK_BYTES")) chunk = std::max(1LL, std::atoll(env));
    std::vector<const char *> cut;
    cut.push_back(p);
    for (const char *c = p + chunk; c < end; c += chunk) {
        const char *x = std::max(c, cut.back());
        while (x < end && !is_space(*x)) ++x;
        if (x >= end) break;
        cut.push_back(x);
    }
    cut.push_back(end);
    const int64_t nch = (int64_t)cut.size() - 1;
    std::vector<int64_t> ntok(nch + 1, 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t c = 0; c < nch; ++c) {
        int64_t k = 0;
        bool in = false;
        for (const char *x = cut[c]; x < cut[c + 1]; ++x) {
            const bool sp = is_space(*x);
            if (!sp && !in) ++k;
            in = !sp;
        }
        ntok[c + 1] = k;
    }
    for (int64_t c = 0; c < nch; ++c) ntok[c + 1] += ntok[c];
    const int64_t have_tok = ntok[nch];
    if (have_tok < need) {
        const int64_t t = have_tok / K, r = have_tok % K;
        if (r == 0)
            return io_fail(LV_ERR_ARG, where + ": expected " + std::to_string(nnz) + " entries, got " +
                                           std::to_string(t));
        if (r == 1) return io_fail(LV_ERR_ARG, where + ": truncated entry " + std::to_string(t + 1));
        return io_fail(LV_ERR_ARG, where + ": missing weight in entry " + std::to_string(t + 1));
    }
    std::vector<int32_t> ti((size_t)nnz), tj((size_t)nnz);
    std::vector<float> tw((size_t)nnz, 1.0f);
    // first bad entry (1-based) and its kind: 1 = not a number, 2 = index out of range
    int64_t bad_entry = INT64_MAX;
    int bad_kind = 0;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t c = 0; c < nch; ++c) {
        int64_t g = ntok[c];
        if (g >= need) continue;
        int64_t my_bad = INT64_MAX;
        int my_kind = 0;
        const char *x = cut[c];
        const char *ce = cut[c + 1];
        char tok[64];
        while (x < ce && g < need) {
            while (x < ce && is_space(*x)) ++x;
            if (x >= ce) break;
            const char *t0 = x;
            while (x < ce && !is_space(*x)) ++x;
            const size_t len = (size_t)(x - t0);
            const int64_t entry = g / K;
            const int field = (int)(g % K);
            ++g;
            if (entry + 1 >= my_bad) continue;
            // a token is copied and terminated (long ones into a string)
            std::string big;
            char *tk = tok;
            if (len >= sizeof tok) {
                big.assign(t0, len);
                tk = &big[0];
            } else {
                std::memcpy(tok, t0, len);
                tok[len] = '\0';
            }
            bool ok = true;
            {
                if (field < 2) {
                    long long v = 0;
                    ok = parse_int(tk, len, v);
                    if (ok && (v < 1 || v > n)) {
                        my_bad = entry + 1;
                        my_kind = 2;
                        continue;
                    }
                    if (ok) (field == 0 ? ti : tj)[(size_t)entry] = (int32_t)(v - 1);
                } else {
                    float v = 0.0f;
                    ok = parse_float(tk, len, v);
                    if (ok) tw[(size_t)entry] = v;
                }
            }
            if (!ok) {
                my_bad = entry + 1;
                my_kind = 1;
            }
        }
        if (my_bad != INT64_MAX) {
#pragma omp critical(lv_mtx_bad)
            if (my_bad < bad_entry) {
                bad_entry = my_bad;
                bad_kind = my_kind;
            }
        }
    }
    if (bad_entry != INT64_MAX)
        return io_fail(LV_ERR_ARG, where + ": entry " + std::to_string(bad_entry) +
                                       (bad_kind == 2 ? " index out of range" : " is not a number"));
    return csr_in_entry_order(ti, tj, tw, n, rowptr, col, w, V, E);
}

static const char kCsrMagic[8] = {'L', 'V', 'C', 'S', 'R', '0', '1', '\0'};

lv_status lv_write_csr(const char *path, const int32_t *rowptr, const int32_t *col,
                       const float *w, int32_t V, int64_t E) {
    if (!path || !rowptr || V < 0 || E < 0 || (E > 0 && (!col || !w)))
        return io_fail(LV_ERR_ARG, "bad argument");
    if (rowptr[0] != 0 || rowptr[V] != E) return io_fail(LV_ERR_ARG, "rowptr does not span E entries");
    FILE *f = std::fopen(path, "wb");
    if (!f) return io_fail(LV_ERR_IO, std::string("cannot create '") + path + "'");
    const int64_t hdr[2] = {(int64_t)V, E};
    bool ok = std::fwrite(kCsrMagic, 1, 8, f) == 8 && std::fwrite(hdr, sizeof hdr, 1, f) == 1 &&
              std::fwrite(rowptr, sizeof(int32_t), (size_t)V + 1, f) == (size_t)V + 1 &&
              (E == 0 || (std::fwrite(col, sizeof(int32_t), (size_t)E, f) == (size_t)E &&
                          std::fwrite(w, sizeof(float), (size_t)E, f) == (size_t)E));
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return io_fail(LV_ERR_IO, std::string("cannot write '") + path + "'");
    return LV_OK;
}

lv_status lv_read_csr(const char *path, int32_t **rowptr, int32_t **col, float **w, int32_t *V,
                      int64_t *E) {
    if (!path || !rowptr || !col || !w || !V || !E) return io_fail(LV_ERR_ARG, "null argument");
    const std::string where = std::string("'") + path + "'";
    MappedFile mf;
    if (!mf.open(path)) return io_fail(LV_ERR_IO, std::string("cannot open ") + where);
    if (mf.size < 24 || std::memcmp(mf.data, kCsrMagic, 8) != 0)
        return io_fail(LV_ERR_ARG, where + " is not a binary CSR file");
    int64_t hdr[2];
    std::memcpy(hdr, mf.data + 8, sizeof hdr);
    const int64_t n = hdr[0], m = hdr[1];
    if (n < 0 || n > 0x7fffffffLL || m < 0 || m > 0x7fffffffLL ||
        (uint64_t)mf.size != 24 + 4 * ((uint64_t)n + 1) + 8 * (uint64_t)m)
        return io_fail(LV_ERR_ARG, where + ": sizes do not match the file length");
    int32_t *rp = (int32_t *)std::malloc(sizeof(int32_t) * ((size_t)n + 1));
    int32_t *cl = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(m, 1));
    float *wt = (float *)std::malloc(sizeof(float) * (size_t)std::max<int64_t>(m, 1));
    if (!rp || !cl || !wt) {
        std::free(rp);
        std::free(cl);
        std::free(wt);
        return io_fail(LV_ERR_OOM, "host allocation failed");
    }
    const char *q = mf.data + 24;
    std::memcpy(rp, q, sizeof(int32_t) * ((size_t)n + 1));
    q += sizeof(int32_t) * ((size_t)n + 1);
    // large arrays copied by all threads
    const int64_t kPiece = 1 << 22;
#pragma omp parallel for schedule(static)
    for (int64_t s0 = 0; s0 < m; s0 += kPiece) {
        const int64_t k = std::min(kPiece, m - s0);
        std::memcpy(cl + s0, q + 4 * s0, 4 * (size_t)k);
        std::memcpy(wt + s0, q + 4 * m + 4 * s0, 4 * (size_t)k);
    }
    bool ok = rp[0] == 0 && rp[n] == m;
    for (int64_t v = 0; ok && v < n; ++v) ok = rp[v] <= rp[v + 1];
    if (!ok) {
        std::free(rp);
        std::free(cl);
        std::free(wt);
        return io_fail(LV_ERR_ARG, where + ": rowptr is not monotone from 0 to E");
    }
    *rowptr = rp;
    *col = cl;
    *w = wt;
    *V = (int32_t)n;
    *E = m;
    return LV_OK;
}

lv_status lv_read_graph(const char *path, int32_t **rowptr, int32_t **col, float **w, int32_t *V,
                        int64_t *E) {
    if (!path) return io_fail(LV_ERR_ARG, "null argument");
    char head[8] = {0};
    FILE *f = std::fopen(path, "rb");
    if (!f) return io_fail(LV_ERR_IO, std::string("cannot open '") + path + "'");
    const size_t got = std::fread(head, 1, 8, f);
    std::fclose(f);
    if (got == 8 && std::memcmp(head, kCsrMagic, 8) == 0)
        return lv_read_csr(path, rowptr, col, w, V, E);
    return lv_read_mtx(path, rowptr, col, w, V, E);
}

lv_status lv_write_mtx(const char *path, const int32_t *rowptr, const int32_t *col, const float *w,
                       int32_t V) {
    if (!path || !rowptr || V < 0) return io_fail(LV_ERR_ARG, "bad argument");
    FILE *f = std::fopen(path, "wb");
    if (!f) return io_fail(LV_ERR_IO, std::string("cannot create '") + path + "'");
    // rows are formatted by all threads in blocks, written in row order
    const int32_t kRows = 1 << 14;
    const int32_t nblk = (V + kRows - 1) / kRows;
    std::vector<std::string> text((size_t)nblk);
    std::vector<int64_t> cnt((size_t)nblk, 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int32_t b = 0; b < nblk; ++b) {
        std::string &out = text[(size_t)b];
        char line[80];
        int64_t k = 0;
        for (int32_t v = b * kRows; v < std::min(V, (b + 1) * kRows); ++v)
            for (int32_t e = rowptr[v]; e < rowptr[v + 1]; ++e) {
                if (col[e] > v) continue;
                const float x = w[e];
                int len;
                if (x == std::floor(x) && std::fabs(x) < 1e9f)
                    len = std::snprintf(line, sizeof line, "%d %d %lld\n", v + 1, col[e] + 1,
                                        (long long)x);
                else
                    len = std::snprintf(line, sizeof line, "%d %d %.9g\n", v + 1, col[e] + 1,
                                        (double)x);
                out.append(line, (size_t)len);
                ++k;
            }
        cnt[(size_t)b] = k;
    }
    int64_t nnz = 0;
    for (int64_t k : cnt) nnz += k;
    bool ok = std::fprintf(f, "%%%%MatrixMarket matrix coordinate real symmetric\n") > 0 &&
              std::fprintf(f, "%d %d %lld\n", V, V, (long long)nnz) > 0;
    for (int32_t b = 0; ok && b < nblk; ++b)
        ok = std::fwrite(text[(size_t)b].data(), 1, text[(size_t)b].size(), f) ==
             text[(size_t)b].size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return io_fail(LV_ERR_IO, std::string("cannot write '") + path + "'");
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
