// fgpu_datagen.cpp — deterministic synthetic tables for the filter / GROUP BY /
// join workloads (host C++, multi-threaded, bit-identical for any thread count).
//
// Every column is produced in fixed chunks of 2^20 rows.  Chunk c of a column
// with seed s draws from xoshiro256** seeded through SplitMix64 with
// s ^ (c * 0x9E3779B97F4A7C15), so the bytes do not depend on how many threads
// generated them.  Column seeds are 0xFA1C0DB500000000 + 16*config + column.
//
// Tables (sizes are arguments):
//   fgen_c1: id = shuffled 0..n-1, dept = u32 code in [0,12) uniform,
//            score = 100*U[0,1), salary = uniform int [20000, 200000].
//   fgen_keys: n distinct int64 drawn from [0, 2^bits), in random order.
//   fgen_c2: key = K[rank] with rank ~ Zipf(s) over |K| ranks (inverse CDF),
//            v1 = uniform int [-1e6, 1e6], v2 ~ N(0,1) (Box-Muller),
//            f = uniform int32 [0, 100).
//   fgen_c3_build: bk = K in a fresh random order, bpay = U[0,1).
//   fgen_c3_probe: pk = uniform member of K with probability `match_frac`,
//            otherwise a uniform value of [0, 2^bits) that is not in K;
//            ppay = uniform int [-1e6, 1e6].
//   fgen_c4: key = uniform int32 [0, card), val = uniform int [-1e6, 1e6],
//            score = 100*U[0,1).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <unordered_set>
#include <vector>

namespace {

constexpr int64_t kChunk = int64_t(1) << 20;

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
        for (auto &w : s) w = splitmix64(x);
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
    // U[0,1) with 53 random bits.
    inline double uniform() { return double(next() >> 11) * 0x1.0p-53; }
    // Uniform integer in [lo, hi] (multiply-shift; hi - lo < 2^63).
    inline int64_t range(int64_t lo, int64_t hi) {
        const unsigned __int128 span = (unsigned __int128)(uint64_t)(hi - lo) + 1;
        return lo + (int64_t)(uint64_t)(((unsigned __int128)next() * span) >> 64);
    }
};

inline uint64_t chunk_seed(uint64_t col_seed, int64_t chunk) {
    return col_seed ^ (uint64_t(chunk) * 0x9E3779B97F4A7C15ull);
}

// Runs f(chunk_index, begin, end) over all chunks of n rows on a few threads.
template <class F>
void for_chunks(int64_t n, F f) {
    const int64_t n_chunks = (n + kChunk - 1) / kChunk;
    int t = (int)std::min<int64_t>(std::max(1u, std::thread::hardware_concurrency()), 32);
    t = (int)std::min<int64_t>(t, std::max<int64_t>(n_chunks, 1));
    std::vector<std::thread> pool;
    for (int w = 0; w < t; ++w) {
        pool.emplace_back([=, &f] {
            for (int64_t c = w; c < n_chunks; c += t) {
                const int64_t b = c * kChunk;
                f(c, b, std::min(n, b + kChunk));
            }
        });
    }
    for (auto &th : pool) th.join();
}

constexpr uint64_t seed_of(int config, int column) {
    return 0xFA1C0DB500000000ull + 16ull * (uint64_t)config + (uint64_t)column;
}

}  // namespace

extern "C" {

uint64_t fgen_seed(int config, int column) { return seed_of(config, column); }

// n distinct keys in [0, 2^bits), random order.  Deterministic, serial.
int fgen_keys(int64_t n, int bits, uint64_t seed, int64_t *out) {
    if (n < 0 || bits < 1 || bits > 62) return -1;
    if (bits < 62 && n > (int64_t(1) << bits)) return -1;
    Xoshiro rng(seed);
    std::unordered_set<int64_t> seen;
    seen.reserve((size_t)n * 2);
    const int64_t hi = (int64_t(1) << bits) - 1;
    int64_t k = 0;
    while (k < n) {
        const int64_t v = rng.range(0, hi);
        if (seen.insert(v).second) out[k++] = v;
    }
    for (int64_t i = n - 1; i > 0; --i) {
        const int64_t j = rng.range(0, i);
        std::swap(out[i], out[j]);
    }
    return 0;
}

int fgen_c1(int64_t n, int64_t *id, uint32_t *dept, double *score, int64_t *salary) {
    if (n < 0) return -1;
    {
        Xoshiro rng(seed_of(1, 0));
        for (int64_t i = 0; i < n; ++i) id[i] = i;
        for (int64_t i = n - 1; i > 0; --i) {
            const int64_t j = rng.range(0, i);
            std::swap(id[i], id[j]);
        }
    }
    for_chunks(n, [&](int64_t c, int64_t b, int64_t e) {
        Xoshiro rd(chunk_seed(seed_of(1, 1), c)), rs(chunk_seed(seed_of(1, 2), c)),
            rp(chunk_seed(seed_of(1, 3), c));
        for (int64_t i = b; i < e; ++i) {
            dept[i] = (uint32_t)rd.range(0, 11);
            score[i] = 100.0 * rs.uniform();
            salary[i] = rp.range(20000, 200000);
        }
    });
    return 0;
}

// Zipf(s) over ranks 1..nk: cdf[k] = sum_{i<=k} i^-s / H.
static void zipf_cdf(int64_t nk, double s, std::vector<double> &cdf) {
    cdf.resize((size_t)nk);
    double acc = 0.0;
    for (int64_t i = 0; i < nk; ++i) {
        acc += std::pow((double)(i + 1), -s);
        cdf[(size_t)i] = acc;
    }
    for (auto &x : cdf) x /= acc;
    cdf.back() = 1.0;
}

int fgen_c2(int64_t n, const int64_t *K, int64_t nk, double s, int64_t *key, int64_t *v1,
            double *v2, int32_t *f) {
    if (n < 0 || nk < 1) return -1;
    std::vector<double> cdf;
    zipf_cdf(nk, s, cdf);
    for_chunks(n, [&](int64_t c, int64_t b, int64_t e) {
        Xoshiro rk(chunk_seed(seed_of(2, 0), c)), r1(chunk_seed(seed_of(2, 1), c)),
            r2(chunk_seed(seed_of(2, 2), c)), rf(chunk_seed(seed_of(2, 3), c));
        for (int64_t i = b; i < e; ++i) {
            const double u = rk.uniform();
            const int64_t r = std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin();
            key[i] = K[std::min(r, nk - 1)];
            v1[i] = r1.range(-1000000, 1000000);
            const double u1 = (double)((r2.next() >> 11) + 1) * 0x1.0p-53;  // (0, 1]
            const double u2 = r2.uniform();
            v2[i] = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
            f[i] = (int32_t)rf.range(0, 99);
        }
    });
    return 0;
}

int fgen_c3_build(int64_t nb, const int64_t *K, int64_t *bk, double *bpay) {
    if (nb < 0) return -1;
    std::memcpy(bk, K, sizeof(int64_t) * (size_t)nb);
    Xoshiro rng(seed_of(3, 0));
    for (int64_t i = nb - 1; i > 0; --i) {
        const int64_t j = rng.range(0, i);
        std::swap(bk[i], bk[j]);
    }
    for_chunks(nb, [&](int64_t c, int64_t b, int64_t e) {
        Xoshiro rp(chunk_seed(seed_of(3, 1), c));
        for (int64_t i = b; i < e; ++i) bpay[i] = rp.uniform();
    });
    return 0;
}

int fgen_c3_probe(int64_t np, const int64_t *K, int64_t nk, int bits, double match_frac,
                  int64_t *pk, int64_t *ppay) {
    if (np < 0 || nk < 1 || bits < 1 || bits > 62) return -1;
    std::vector<int64_t> sorted(K, K + nk);
    std::sort(sorted.begin(), sorted.end());
    const int64_t hi = (int64_t(1) << bits) - 1;
    for_chunks(np, [&](int64_t c, int64_t b, int64_t e) {
        Xoshiro rm(chunk_seed(seed_of(3, 2), c)), rk(chunk_seed(seed_of(3, 3), c)),
            rp(chunk_seed(seed_of(3, 4), c));
        for (int64_t i = b; i < e; ++i) {
            if (rm.uniform() < match_frac) {
                pk[i] = K[rk.range(0, nk - 1)];
            } else {
                int64_t v;
                do {
                    v = rk.range(0, hi);
                } while (std::binary_search(sorted.begin(), sorted.end(), v));
                pk[i] = v;
            }
            ppay[i] = rp.range(-1000000, 1000000);
        }
    });
    return 0;
}

int fgen_c4(int64_t n, uint64_t seed, int32_t card, int32_t *key, int64_t *val, double *score) {
    if (n < 0 || card < 1) return -1;
    for_chunks(n, [&](int64_t c, int64_t b, int64_t e) {
        Xoshiro rk(chunk_seed(seed + 0, c)), rv(chunk_seed(seed + 1, c)),
            rs(chunk_seed(seed + 2, c));
        for (int64_t i = b; i < e; ++i) {
            key[i] = (int32_t)rk.range(0, card - 1);
            val[i] = rv.range(-1000000, 1000000);
            score[i] = 100.0 * rs.uniform();
        }
    });
    return 0;
}

// Uniform doubles / ints for tests (one stream, chunked like the rest).
int fgen_uniform_f64(int64_t n, uint64_t seed, double lo, double hi, double *out) {
    for_chunks(n, [&](int64_t c, int64_t b, int64_t e) {
        Xoshiro r(chunk_seed(seed, c));
        for (int64_t i = b; i < e; ++i) out[i] = lo + (hi - lo) * r.uniform();
    });
    return 0;
}

int fgen_uniform_i64(int64_t n, uint64_t seed, int64_t lo, int64_t hi, int64_t *out) {
    for_chunks(n, [&](int64_t c, int64_t b, int64_t e) {
        Xoshiro r(chunk_seed(seed, c));
        for (int64_t i = b; i < e; ++i) out[i] = r.range(lo, hi);
    });
    return 0;
}

}  // extern "C"
