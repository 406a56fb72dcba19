// falcon_oracle.cpp — CPU restatement of FalconDB's WHERE filter, hash GROUP BY
// and IntHashTable hash join.
//
// TEST INFRASTRUCTURE.  This file is the correctness oracle for the HIP backend
// and the timed CPU baseline beside it.  Only tests, the smoke check and the
// benchmark's CPU leg load it; the product library never links or calls it.
//
// It keeps the reference's algorithm choices (rayon over 65,536-row batches,
// packed LSB-first masks, materialised filter output, per-batch hash maps with a
// serial merge in batch order, a serial chained-table join build with a Bloom
// filter and parallel probe units) so that its timing stands for the reference's
// CPU path, and it adds exact results (integer sums in 128 bits, float sums
// exactly rounded from a wide integer accumulator) that the parity tests compare
// against.
//
// Reference files followed (paths relative to the FalconDB tree):
//   filter       src/query/expr.rs:635-1175, src/storage/table.rs:154-242,
//                src/storage/column.rs:1077-1123, src/query/select.rs:105-160, :457-487
//   COUNT(*)     src/db.rs:1439-1476
//   GROUP BY     src/query/aggregate.rs:535-1420 (fast_utf8_gb :619-835,
//                fast_i64_gb :840-912, accumulate_batch :916-1169,
//                merge_state_entry :1351-1390, build_agg_col :3390-3420)
//   scalar aggs  src/query/aggregate.rs:1394-1528, sum_lanes :1557-1590,
//                checked_sum_i64 / par_sum_* :3039-3119
//   join         src/query/join.rs:49-132 (semi/anti), :184-502 (inner/left),
//                :607-723 (build / probe_one), :1054-1107, :1318-1350
//                (IntHashTable), :1441-1482 (BuildFilter)
//   partitioning src/query/tuning.rs:28-53, src/storage/table.rs:339
//
// Parity status of each exported function is listed in the header comment of
// tests/test_oracle.py (which pins them against the reference's own known answers,
// closed forms and brute force).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <limits>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../include/falcon_gpu.h"

namespace {

constexpr uint32_t NO_ROW = 0xFFFFFFFFu;  // src/query/join.rs:1050
constexpr uint32_t NULL_CODE = 0xFFFFFFFFu;  // src/storage/column.rs Utf8 NULL code

// ─── thread fan-out (stands in for rayon's global pool) ─────────────────────
int g_threads = 0;
int threads() {
    if (g_threads > 0) return g_threads;
    unsigned h = std::thread::hardware_concurrency();
    return h ? (int)h : 1;
}

void parallel_for(int64_t n_units, const std::function<void(int64_t)> &fn) {
    const int t = (int)std::min<int64_t>(threads(), std::max<int64_t>(n_units, 1));
    if (t <= 1) {
        for (int64_t u = 0; u < n_units; ++u) fn(u);
        return;
    }
    std::atomic<int64_t> next{0};
    std::vector<std::thread> pool;
    for (int w = 0; w < t; ++w)
        pool.emplace_back([&] {
            for (;;) {
                const int64_t u = next.fetch_add(1);
                if (u >= n_units) break;
                fn(u);
            }
        });
    for (auto &th : pool) th.join();
}

// src/query/tuning.rs:39-53: per-thread thresholds scaled by the thread count.
int64_t par_min_rows(int64_t per_thread) { return per_thread * threads(); }

// ─── columns ───────────────────────────────────────────────────────────────
struct Col {
    int dtype = 0;
    const void *values = nullptr;
    const uint64_t *validity = nullptr;
    int dict_len = 0;
    const char *const *dict = nullptr;

    bool valid(int64_t i) const {
        return validity ? ((validity[i >> 6] >> (i & 63)) & 1) : true;
    }
    bool has_nulls(int64_t n) const {
        if (dtype == FGPU_UTF8) {
            const uint32_t *c = (const uint32_t *)values;
            for (int64_t i = 0; i < n; ++i)
                if (c[i] == NULL_CODE) return true;
            return false;
        }
        if (!validity) return false;
        for (int64_t i = 0; i < n; ++i)
            if (!valid(i)) return true;
        return false;
    }
    bool is_null(int64_t i) const {
        if (dtype == FGPU_UTF8) return ((const uint32_t *)values)[i] == NULL_CODE;
        return !valid(i);
    }
    int64_t i64(int64_t i) const {
        switch (dtype) {
            case FGPU_INT32: return ((const int32_t *)values)[i];
            case FGPU_INT64: return ((const int64_t *)values)[i];
            case FGPU_UTF8: return ((const uint32_t *)values)[i];
            default: return (int64_t)((const double *)values)[i];
        }
    }
    double f64(int64_t i) const {
        switch (dtype) {
            case FGPU_INT32: return (double)((const int32_t *)values)[i];
            case FGPU_INT64: return (double)((const int64_t *)values)[i];
            case FGPU_FLOAT64: return ((const double *)values)[i];
            default: return (double)((const uint32_t *)values)[i];
        }
    }
};

std::vector<Col> make_cols(const fgpu_column_desc *d, int n) {
    std::vector<Col> v((size_t)n);
    for (int i = 0; i < n; ++i) {
        v[i].dtype = d[i].dtype;
        v[i].values = d[i].values;
        v[i].validity = d[i].validity;
        v[i].dict_len = d[i].dict_len;
        v[i].dict = d[i].dict;
    }
    return v;
}

// ─── exact float summation ──────────────────────────────────────────────────
// Every finite double is an integer multiple of 2^-1074.  The sum is kept as
// two unsigned big integers (positive and negative terms) in units of 2^-1074,
// 34 limbs of 64 bits each (2176 bits: the largest double is below 2^2098 in
// these units, which leaves 78 bits of room for the number of terms).  Nothing
// can overflow or round until result(), which rounds the exact difference to
// nearest-even once.
struct ExactSum {
    static constexpr int NL = 34;
    uint64_t pos[NL] = {}, neg[NL] = {};
    double special = 0.0;   // running sum of non-finite inputs (IEEE semantics)
    bool any_special = false;

    static void add_at(uint64_t *a, uint64_t m, int bit) {
        int l = bit >> 6, s = bit & 63;
        const uint64_t lo = m << s, hi = s ? (m >> (64 - s)) : 0;
        unsigned __int128 t = (unsigned __int128)a[l] + lo;
        a[l] = (uint64_t)t;
        uint64_t carry = (uint64_t)(t >> 64) + hi;
        while (carry && ++l < NL) {
            t = (unsigned __int128)a[l] + carry;
            a[l] = (uint64_t)t;
            carry = (uint64_t)(t >> 64);
        }
    }
    void add(double x) {
        if (!std::isfinite(x)) {
            special += x;
            any_special = true;
            return;
        }
        uint64_t bits;
        std::memcpy(&bits, &x, 8);
        const int be = (int)((bits >> 52) & 0x7ff);
        uint64_t m = bits & 0x000fffffffffffffull;
        if (be) m |= 0x0010000000000000ull;
        if (!m) return;
        const int bit = be ? be - 1 : 0;  // x = m * 2^(bit - 1074)
        add_at((bits >> 63) ? neg : pos, m, bit);
    }
    // a >= b over NL limbs
    static bool ge(const uint64_t *a, const uint64_t *b) {
        for (int l = NL - 1; l >= 0; --l)
            if (a[l] != b[l]) return a[l] > b[l];
        return true;
    }
    double result() const {
        if (any_special) return special;  // NaN / +-inf dominate
        const bool negative = !ge(pos, neg);
        const uint64_t *a = negative ? neg : pos, *b = negative ? pos : neg;
        uint64_t d[NL];
        uint64_t borrow = 0;
        for (int l = 0; l < NL; ++l) {
            const unsigned __int128 t = (unsigned __int128)a[l] - b[l] - borrow;
            d[l] = (uint64_t)t;
            borrow = (uint64_t)(t >> 64) ? 1 : 0;
        }
        int top = NL - 1;
        while (top >= 0 && d[top] == 0) --top;
        if (top < 0) return 0.0;
        const int msb = top * 64 + 63 - __builtin_clzll(d[top]);
        auto bit_at = [&](int i) -> uint64_t { return (d[i >> 6] >> (i & 63)) & 1; };
        double r;
        if (msb <= 52) {  // below 2^53 units: exact (a subnormal or small normal)
            r = std::ldexp((double)d[0], -1074);
        } else {
            const int shift = msb - 52;  // drop `shift` low bits
            // bits [shift, shift + 52] of d
            const int l = shift >> 6, o = shift & 63;
            uint64_t mant = d[l] >> o;
            if (o && l + 1 < NL) mant |= d[l + 1] << (64 - o);
            mant &= (1ull << 53) - 1;
            const bool round = bit_at(shift - 1);
            // sticky: any set bit below shift - 1
            const int sb = shift - 1;
            bool sticky = (d[sb >> 6] & ((1ull << (sb & 63)) - 1)) != 0;
            for (int i = 0; i < (sb >> 6) && !sticky; ++i) sticky = d[i] != 0;
            int e = shift;
            if (round && (sticky || (mant & 1))) {
                if (++mant == (1ull << 53)) {
                    mant >>= 1;
                    ++e;
                }
            }
            r = std::ldexp((double)mant, e - 1074);  // inf when it leaves the range
        }
        return negative ? -r : r;
    }
};

// ─── predicates: fast_cmp_mask semantics (src/query/expr.rs:1076-1175) ──────
struct Pred {
    std::vector<fgpu_pred_node> nodes;
    // For UTF8 comparisons: per-dictionary-entry pass table (decided once per
    // distinct string, src/query/expr.rs:1139-1166).
    std::vector<std::vector<uint8_t>> dict_pass;
};

// like_recursive (src/query/expr.rs:1186-1199) over bytes, memoised:
// m[i][j] = s[i..] matches p[j..].  '%' matches zero or more bytes, '_' one
// byte, any other byte itself.
bool like_matches(const std::string &s, const std::string &p) {
    const size_t ns = s.size(), np = p.size();
    std::vector<uint8_t> m((ns + 1) * (np + 1), 0);
    auto at = [&](size_t i, size_t j) -> uint8_t & { return m[i * (np + 1) + j]; };
    at(ns, np) = 1;
    for (size_t j = np; j-- > 0;) {
        for (size_t i = ns + 1; i-- > 0;) {
            bool r;
            if (p[j] == '%') r = at(i, j + 1) || (i < ns && at(i + 1, j));
            else if (i == ns) r = false;
            else if (p[j] == '_') r = at(i + 1, j + 1);
            else r = s[i] == p[j] && at(i + 1, j + 1);
            at(i, j) = r;
        }
    }
    return at(0, 0);
}

// ILIKE's lower-casing, ASCII letters only (str::to_lowercase folds every
// Unicode letter; non-ASCII letters are compared as they are here).
std::string lower_ascii(const char *x) {
    std::string s = x ? x : "";
    for (char &ch : s)
        if (ch >= 'A' && ch <= 'Z') ch = (char)(ch + 32);
    return s;
}

bool build_pred(const fgpu_pred *p, const std::vector<Col> &cols, Pred &out) {
    out.nodes.clear();
    out.dict_pass.clear();
    if (!p) return true;
    out.nodes.assign(p->nodes, p->nodes + p->n_nodes);
    out.dict_pass.resize(out.nodes.size());
    int depth = 0;
    for (size_t k = 0; k < out.nodes.size(); ++k) {
        const auto &nd = out.nodes[k];
        if (nd.kind == FGPU_P_AND || nd.kind == FGPU_P_OR) {
            if (depth < 2) return false;
            --depth;
            continue;
        }
        if (nd.col < 0 || nd.col >= (int)cols.size()) return false;
        ++depth;
        if (nd.kind == FGPU_P_LIKE) {  // decided once per dictionary entry
            const Col &c = cols[nd.col];
            if (c.dtype != FGPU_UTF8 || nd.op < 0 || nd.op > 3) return false;
            const bool neg = nd.op & 1, fold = nd.op & 2;
            const std::string pat = fold ? lower_ascii(nd.lit_str) : std::string(nd.lit_str ? nd.lit_str : "");
            auto &pass = out.dict_pass[k];
            pass.resize((size_t)c.dict_len);
            for (int i = 0; i < c.dict_len; ++i) {
                const bool hit = like_matches(fold ? lower_ascii(c.dict[i]) : std::string(c.dict[i]), pat);
                pass[i] = hit != neg;
            }
            continue;
        }
        if (nd.kind == FGPU_P_IN && (nd.op < 0 || nd.op > 1 || nd.n_list < 0 || (nd.n_list > 0 && !nd.list)))
            return false;
        if (nd.kind == FGPU_P_CMP && cols[nd.col].dtype == FGPU_UTF8 &&
            nd.lit_type == FGPU_UTF8) {
            const Col &c = cols[nd.col];
            auto &pass = out.dict_pass[k];
            pass.resize((size_t)c.dict_len);
            for (int i = 0; i < c.dict_len; ++i) {
                const int r = std::strcmp(c.dict[i], nd.lit_str ? nd.lit_str : "");
                bool h = false;
                switch (nd.op) {
                    case FGPU_EQ: h = r == 0; break;
                    case FGPU_NE: h = r != 0; break;
                    case FGPU_LT: h = r < 0; break;
                    case FGPU_LE: h = r <= 0; break;
                    case FGPU_GT: h = r > 0; break;
                    case FGPU_GE: h = r >= 0; break;
                }
                pass[i] = h;
            }
        }
    }
    return depth == 1 || out.nodes.empty();
}

template <class T>
inline bool cmp(T x, T v, int op) {
    switch (op) {
        case FGPU_EQ: return x == v;
        case FGPU_NE: return x != v;
        case FGPU_LT: return x < v;
        case FGPU_LE: return x <= v;
        case FGPU_GT: return x > v;
        case FGPU_GE: return x >= v;
    }
    return false;
}

// One comparison leaf over rows [b, e): bit (i - b) of words.
void leaf_mask(const fgpu_pred_node &nd, const Col &c, const std::vector<uint8_t> &pass,
               int64_t b, int64_t e, uint64_t *w) {
    const int64_t n = e - b;
    const int64_t nw = (n + 63) / 64;
    std::memset(w, 0, sizeof(uint64_t) * (size_t)nw);
    auto set = [&](int64_t i) { w[i >> 6] |= 1ull << (i & 63); };
    if (nd.kind == FGPU_P_LIKE) {
        // a NULL row never matched, and NOT LIKE negates that (src/query/expr.rs:732-736)
        const bool null_passes = nd.op & 1;
        const uint32_t *x = (const uint32_t *)c.values + b;
        for (int64_t i = 0; i < n; ++i)
            if (x[i] == NULL_CODE ? null_passes : pass[x[i]] != 0) set(i);
        return;
    }
    if (nd.kind == FGPU_P_IN) {
        // the union of each literal's `=` mask (src/query/expr.rs:776-789)
        std::vector<uint64_t> one((size_t)nw);
        for (int m = 0; m < nd.n_list; ++m) {
            fgpu_pred_node eq = nd;
            eq.kind = FGPU_P_CMP;
            eq.op = FGPU_EQ;
            if (nd.lit_type == FGPU_UTF8) eq.lit_str = ((const char *const *)nd.list)[m];
            else if (nd.lit_type == FGPU_FLOAT64) eq.lit_f64 = ((const double *)nd.list)[m];
            else eq.lit_i64 = ((const int64_t *)nd.list)[m];
            std::vector<uint8_t> eq_pass;
            if (c.dtype == FGPU_UTF8 && nd.lit_type == FGPU_UTF8) {
                eq_pass.resize((size_t)c.dict_len);
                for (int i = 0; i < c.dict_len; ++i)
                    eq_pass[i] = std::strcmp(c.dict[i], eq.lit_str ? eq.lit_str : "") == 0;
            }
            leaf_mask(eq, c, eq_pass, b, e, one.data());
            for (int64_t i = 0; i < nw; ++i) w[i] |= one[i];
        }
        if (nd.op == 1) {  // NOT IN: the complement, without padding bits or NULL rows (:790-800)
            for (int64_t i = 0; i < nw; ++i) w[i] = ~w[i];
            if (n & 63) w[nw - 1] &= (1ull << (n & 63)) - 1;
            for (int64_t i = 0; i < n; ++i)
                if (c.is_null(b + i)) w[i >> 6] &= ~(1ull << (i & 63));
        }
        return;
    }
    if (nd.kind == FGPU_P_IS_NULL || nd.kind == FGPU_P_IS_NOT_NULL) {
        const bool want_null = nd.kind == FGPU_P_IS_NULL;
        for (int64_t i = 0; i < n; ++i)
            if (c.is_null(b + i) == want_null) set(i);
        return;
    }
    const int op = nd.op;
    const int lt = nd.lit_type;
    if (c.dtype == FGPU_FLOAT64 && (lt == FGPU_FLOAT64 || lt == FGPU_INT64 || lt == FGPU_INT32)) {
        const double v = lt == FGPU_FLOAT64 ? nd.lit_f64 : (double)nd.lit_i64;
        const double *x = (const double *)c.values + b;
        for (int64_t i = 0; i < n; ++i)
            if (cmp(x[i], v, op)) set(i);
    } else if (c.dtype == FGPU_INT64 && (lt == FGPU_INT64 || lt == FGPU_INT32)) {
        const int64_t v = nd.lit_i64;
        const int64_t *x = (const int64_t *)c.values + b;
        for (int64_t i = 0; i < n; ++i)
            if (cmp(x[i], v, op)) set(i);
    } else if (c.dtype == FGPU_INT64 && lt == FGPU_FLOAT64) {
        // widened to f64 (src/query/expr.rs:1122-1131)
        const int64_t *x = (const int64_t *)c.values + b;
        for (int64_t i = 0; i < n; ++i)
            if (cmp((double)x[i], nd.lit_f64, op)) set(i);
    } else if (c.dtype == FGPU_INT32 && (lt == FGPU_INT32 || lt == FGPU_INT64)) {
        const int32_t v = (int32_t)nd.lit_i64;  // `as i32` (src/query/expr.rs:1134)
        const int32_t *x = (const int32_t *)c.values + b;
        for (int64_t i = 0; i < n; ++i)
            if (cmp(x[i], v, op)) set(i);
    } else if (c.dtype == FGPU_UTF8 && lt == FGPU_UTF8) {
        const uint32_t *x = (const uint32_t *)c.values + b;
        for (int64_t i = 0; i < n; ++i)
            if (x[i] != NULL_CODE && x[i] < pass.size() && pass[x[i]]) set(i);
        return;
    } else {
        return;  // unsupported pairing: no rows pass (src/query/expr.rs:1171)
    }
    // clear_null_bits (src/query/expr.rs:961-973)
    if (c.validity)
        for (int64_t i = 0; i < n; ++i)
            if (!c.valid(b + i)) w[i >> 6] &= ~(1ull << (i & 63));
}

// eval_mask_bits over one batch [b, e): postfix evaluation of word masks.
void eval_mask(const Pred &p, const std::vector<Col> &cols, int64_t b, int64_t e,
               std::vector<uint64_t> &out) {
    const int64_t nw = (e - b + 63) / 64;
    out.assign((size_t)nw, 0);
    if (p.nodes.empty()) {
        for (int64_t i = 0; i < e - b; ++i) out[i >> 6] |= 1ull << (i & 63);
        return;
    }
    std::vector<std::vector<uint64_t>> st;
    for (size_t k = 0; k < p.nodes.size(); ++k) {
        const auto &nd = p.nodes[k];
        if (nd.kind == FGPU_P_AND || nd.kind == FGPU_P_OR) {
            auto r = std::move(st.back());
            st.pop_back();
            auto &l = st.back();
            for (int64_t i = 0; i < nw; ++i) l[i] = nd.kind == FGPU_P_AND ? (l[i] & r[i]) : (l[i] | r[i]);
            continue;
        }
        std::vector<uint64_t> w((size_t)nw);
        leaf_mask(nd, cols[nd.col], p.dict_pass[k], b, e, w.data());
        st.push_back(std::move(w));
    }
    out = std::move(st.back());
}

// ─── thread-local error text ────────────────────────────────────────────────
thread_local std::string g_err;
int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

// ─── results ────────────────────────────────────────────────────────────────
struct OCol {
    std::string name;
    int dtype = FGPU_INT64;        // exact representation
    std::vector<int64_t> i64;
    std::vector<double> f64;
    std::vector<uint8_t> valid;
    int ref_dtype = FGPU_INT64;    // reference representation (Float64 for SUM/MIN/MAX/AVG)
    std::vector<int64_t> ref_i64;
    std::vector<double> ref_f64;
    int dict_len = 0;
    const char *const *dict = nullptr;
};

}  // namespace

struct ora_result {
    int64_t rows = 0;
    std::vector<OCol> cols;
};

namespace {

std::string agg_alias(const fgpu_agg &a, int i) {
    if (a.alias && *a.alias) return a.alias;
    return "col" + std::to_string(i + 1);
}

// ─── materialised filter output for one batch (H2) ──────────────────────────
struct FBatch {
    int64_t rows = 0;
    // Gathered needed columns (raw 8-byte slots) + validity; or views when no filter.
    std::vector<std::vector<int64_t>> data;   // per needed column, bit patterns
    std::vector<std::vector<uint8_t>> valid;  // per needed column (empty = all valid)
};

// Gathers the rows whose bit is set (gather_by_bitmask, src/storage/column.rs:1077-1123).
void gather(const Col &c, int64_t b, const std::vector<uint64_t> &mask, int64_t cnt,
            std::vector<int64_t> &out, std::vector<uint8_t> &ovalid) {
    out.resize((size_t)cnt);
    const bool nulls = c.validity != nullptr;
    if (nulls) ovalid.resize((size_t)cnt);
    int64_t k = 0;
    for (size_t w = 0; w < mask.size(); ++w) {
        uint64_t bits = mask[w];
        while (bits) {
            const int j = __builtin_ctzll(bits);
            bits &= bits - 1;
            const int64_t i = b + (int64_t)w * 64 + j;
            int64_t v;
            switch (c.dtype) {
                case FGPU_INT32: v = ((const int32_t *)c.values)[i]; break;
                case FGPU_UTF8: v = ((const uint32_t *)c.values)[i]; break;
                default: std::memcpy(&v, (const char *)c.values + 8 * i, 8);
            }
            out[k] = v;
            if (nulls) ovalid[k] = c.valid(i);
            ++k;
        }
    }
}

// execute_select with a filter over the needed columns; batches with no
// surviving rows are dropped (src/query/select.rs:145-149).
std::vector<FBatch> scan_filter(const std::vector<Col> &cols, int64_t n, int64_t batch_rows,
                                const Pred &pred, const std::vector<int> &needed) {
    const int64_t nb = (n + batch_rows - 1) / batch_rows;
    std::vector<FBatch> out((size_t)nb);
    parallel_for(nb, [&](int64_t bi) {
        const int64_t b = bi * batch_rows, e = std::min(n, b + batch_rows);
        std::vector<uint64_t> mask;
        eval_mask(pred, cols, b, e, mask);
        int64_t cnt = 0;
        for (auto w : mask) cnt += __builtin_popcountll(w);
        FBatch &fb = out[(size_t)bi];
        fb.rows = cnt;
        fb.data.resize(needed.size());
        fb.valid.resize(needed.size());
        for (size_t k = 0; k < needed.size(); ++k)
            gather(cols[needed[k]], b, mask, cnt, fb.data[k], fb.valid[k]);
    });
    std::vector<FBatch> kept;
    for (auto &fb : out)
        if (fb.rows > 0) kept.push_back(std::move(fb));
    return kept;
}

inline double bits_to_f64(int64_t v) {
    double d;
    std::memcpy(&d, &v, 8);
    return d;
}

// Per-aggregate reference state (BatchAggState, src/query/aggregate.rs:489-520).
struct AggState {
    std::vector<int64_t> count;
    std::vector<double> sum;
    std::vector<uint8_t> has_val;
    void resize(size_t n) {
        count.resize(n, 0);
        sum.resize(n, 0.0);
        has_val.resize(n, 0);
    }
};

struct AggPlan {
    int func;
    int input;     // index into needed[] or -1 for COUNT(*)
    int dtype;     // input column type
};

// accumulate_batch (src/query/aggregate.rs:916-1169) for rows [s, e) of a batch.
void accumulate(const std::vector<AggPlan> &plan, const FBatch &fb, const uint32_t *gids,
                int64_t s, int64_t e, std::vector<AggState> &st) {
    for (size_t a = 0; a < plan.size(); ++a) {
        const AggPlan &ap = plan[a];
        AggState &S = st[a];
        if (ap.input < 0) {
            for (int64_t r = s; r < e; ++r) S.count[gids[r - s]] += 1;
            continue;
        }
        const auto &d = fb.data[ap.input];
        const auto &vd = fb.valid[ap.input];
        const bool isf = ap.dtype == FGPU_FLOAT64;
        for (int64_t r = s; r < e; ++r) {
            if (!vd.empty() && !vd[r]) continue;
            const uint32_t g = gids[r - s];
            const double v = isf ? bits_to_f64(d[r]) : (double)d[r];
            switch (ap.func) {
                case FGPU_COUNT: S.count[g] += 1; break;
                case FGPU_SUM: S.sum[g] += v; S.has_val[g] = 1; break;
                case FGPU_AVG: S.sum[g] += v; S.count[g] += 1; break;
                case FGPU_MIN:
                    if (!S.has_val[g] || v < S.sum[g]) { S.sum[g] = v; S.has_val[g] = 1; }
                    break;
                case FGPU_MAX:
                    if (!S.has_val[g] || v > S.sum[g]) { S.sum[g] = v; S.has_val[g] = 1; }
                    break;
            }
        }
    }
}

// merge_state_entry (src/query/aggregate.rs:1351-1390)
void merge_entry(int func, const AggState &src, size_t si, AggState &dst, size_t di) {
    switch (func) {
        case FGPU_COUNT: dst.count[di] += src.count[si]; break;
        case FGPU_SUM:
            if (src.has_val[si]) { dst.sum[di] += src.sum[si]; dst.has_val[di] = 1; }
            break;
        case FGPU_AVG:
            dst.sum[di] += src.sum[si];
            dst.count[di] += src.count[si];
            break;
        case FGPU_MIN:
            if (src.has_val[si] && (!dst.has_val[di] || src.sum[si] < dst.sum[di])) {
                dst.sum[di] = src.sum[si]; dst.has_val[di] = 1;
            }
            break;
        case FGPU_MAX:
            if (src.has_val[si] && (!dst.has_val[di] || src.sum[si] > dst.sum[di])) {
                dst.sum[di] = src.sum[si]; dst.has_val[di] = 1;
            }
            break;
    }
}

// Open-addressing map i64 -> u32 with FxHash-style mixing: stands in for the
// reference's FastMap<i64, u32> (hashbrown + FxHasher, src/query/hash.rs:12-77).
struct I64Map {
    std::vector<int64_t> keys;
    std::vector<uint32_t> vals;
    std::vector<uint8_t> used;
    size_t mask = 0, n = 0;
    I64Map() { rehash(16); }
    static inline size_t h(int64_t k) {
        return (size_t)(((uint64_t)k * 0x517cc1b727220a95ull) >> 7);
    }
    void rehash(size_t cap) {
        std::vector<int64_t> ok = std::move(keys);
        std::vector<uint32_t> ov = std::move(vals);
        std::vector<uint8_t> ou = std::move(used);
        keys.assign(cap, 0);
        vals.assign(cap, 0);
        used.assign(cap, 0);
        mask = cap - 1;
        n = 0;
        for (size_t i = 0; i < ou.size(); ++i)
            if (ou[i]) insert_new(ok[i], ov[i]);
    }
    void insert_new(int64_t k, uint32_t v) {
        size_t i = h(k) & mask;
        while (used[i]) i = (i + 1) & mask;
        used[i] = 1;
        keys[i] = k;
        vals[i] = v;
        ++n;
    }
    // Returns the value for k, inserting `fresh` if absent; sets `inserted`.
    uint32_t get_or_insert(int64_t k, uint32_t fresh, bool &inserted) {
        size_t i = h(k) & mask;
        while (used[i]) {
            if (keys[i] == k) { inserted = false; return vals[i]; }
            i = (i + 1) & mask;
        }
        if ((n + 1) * 8 > (mask + 1) * 7) {
            rehash((mask + 1) * 2);
            return get_or_insert(k, fresh, inserted);
        }
        used[i] = 1;
        keys[i] = k;
        vals[i] = fresh;
        ++n;
        inserted = true;
        return fresh;
    }
};

// Exact integer totals.  The reference never wraps: its GROUP BY sums `v as
// f64` (src/query/aggregate.rs:1020-1036) and its scalar SUM falls back to f64
// lanes when a checked i64 sum overflows (checked_sum_i64 / sum_i64, :3039-3110;
// known answer falls_back_instead_of_wrapping_on_overflow, :1975-1982).  The
// exact form keeps the total in 128 bits: an int64 when it fits, else its
// correctly rounded double (`is_more_accurate_past_the_float_limit`, :1959-1970).
inline bool i128_fits(__int128 t) { return t >= (__int128)INT64_MIN && t <= (__int128)INT64_MAX; }

// Fills an exact SUM column from 128-bit totals: Int64 when every total fits,
// else Float64 of every total.
void int_sum_column(const std::vector<__int128> &tot, OCol &oc) {
    bool all_fit = true;
    for (auto t : tot) all_fit = all_fit && i128_fits(t);
    if (all_fit) {
        oc.dtype = FGPU_INT64;
        oc.i64.resize(tot.size());
        for (size_t g = 0; g < tot.size(); ++g) oc.i64[g] = (int64_t)tot[g];
    } else {
        oc.dtype = FGPU_FLOAT64;
        oc.f64.resize(tot.size());
        for (size_t g = 0; g < tot.size(); ++g) oc.f64[g] = (double)tot[g];
    }
}

// Exact per-group values, computed by a separate pass over the filtered rows.
void exact_values(const std::vector<AggPlan> &plan, const std::vector<FBatch> &fbs,
                  const std::vector<std::vector<uint32_t>> &batch_gids, size_t n_groups,
                  ora_result &res, size_t first_agg_col) {
    for (size_t a = 0; a < plan.size(); ++a) {
        const AggPlan &ap = plan[a];
        OCol &oc = res.cols[first_agg_col + a];
        const bool isf = ap.dtype == FGPU_FLOAT64;
        std::vector<int64_t> cnt(n_groups, 0), mn(n_groups, INT64_MAX), mx(n_groups, INT64_MIN);
        std::vector<__int128> isum(n_groups, 0);
        std::vector<double> fmn(n_groups, 0.0), fmx(n_groups, 0.0);
        std::vector<uint8_t> fhas(n_groups, 0);
        std::vector<ExactSum> fs;
        if (isf && (ap.func == FGPU_SUM || ap.func == FGPU_AVG)) fs.resize(n_groups);
        for (size_t bi = 0; bi < fbs.size(); ++bi) {
            const FBatch &fb = fbs[bi];
            const auto &gids = batch_gids[bi];
            for (int64_t r = 0; r < fb.rows; ++r) {
                const uint32_t g = gids[r];
                if (ap.input < 0) { cnt[g]++; continue; }
                const auto &vd = fb.valid[ap.input];
                if (!vd.empty() && !vd[r]) continue;
                const int64_t raw = fb.data[ap.input][r];
                cnt[g]++;
                if (isf) {
                    const double v = bits_to_f64(raw);
                    if (!fs.empty()) fs[g].add(v);
                    if (!std::isnan(v)) {
                        if (!fhas[g] || v < fmn[g]) fmn[g] = v;
                        if (!fhas[g] || v > fmx[g]) fmx[g] = v;
                        fhas[g] = 1;
                    }
                } else {
                    isum[g] += raw;
                    mn[g] = std::min(mn[g], raw);
                    mx[g] = std::max(mx[g], raw);
                }
            }
        }
        oc.valid.assign(n_groups, 1);
        if (ap.func == FGPU_COUNT) {
            oc.dtype = FGPU_INT64;
            oc.i64 = cnt;
            continue;
        }
        for (size_t g = 0; g < n_groups; ++g) oc.valid[g] = cnt[g] > 0;
        if (ap.func == FGPU_AVG) {
            oc.dtype = FGPU_FLOAT64;
            oc.f64.resize(n_groups);
            for (size_t g = 0; g < n_groups; ++g)
                oc.f64[g] = cnt[g] ? (isf ? fs[g].result() : (double)isum[g]) / (double)cnt[g] : 0.0;  // rounded total / count
            continue;
        }
        if (isf) {
            oc.dtype = FGPU_FLOAT64;
            oc.f64.resize(n_groups);
            for (size_t g = 0; g < n_groups; ++g) {
                if (ap.func == FGPU_SUM) oc.f64[g] = cnt[g] ? fs[g].result() : 0.0;
                else {  // f64::min / f64::max folding: NaN skipped, all-NaN -> NaN
                    oc.f64[g] = !fhas[g] ? std::numeric_limits<double>::quiet_NaN()
                                : ap.func == FGPU_MIN ? fmn[g] : fmx[g];
                }
            }
        } else if (ap.func == FGPU_SUM) {
            int_sum_column(isum, oc);
        } else {
            oc.dtype = FGPU_INT64;
            oc.i64.resize(n_groups);
            for (size_t g = 0; g < n_groups; ++g) oc.i64[g] = ap.func == FGPU_MIN ? mn[g] : mx[g];
        }
    }
}

void ref_values(const std::vector<AggPlan> &plan, const std::vector<AggState> &st,
                size_t n_groups, ora_result &res, size_t first_agg_col) {
    // build_agg_col (src/query/aggregate.rs:3390-3420)
    for (size_t a = 0; a < plan.size(); ++a) {
        OCol &oc = res.cols[first_agg_col + a];
        const AggState &S = st[a];
        if (plan[a].func == FGPU_COUNT) {
            oc.ref_dtype = FGPU_INT64;
            oc.ref_i64.assign(S.count.begin(), S.count.begin() + n_groups);
            continue;
        }
        oc.ref_dtype = FGPU_FLOAT64;
        oc.ref_f64.resize(n_groups);
        for (size_t g = 0; g < n_groups; ++g)
            oc.ref_f64[g] = plan[a].func == FGPU_AVG
                                ? (S.count[g] > 0 ? S.sum[g] / (double)S.count[g] : 0.0)
                                : S.sum[g];
    }
}

}  // namespace

extern "C" {

const char *ora_last_error(void) { return g_err.c_str(); }
void ora_set_threads(int t) { g_threads = t; }
int ora_threads(void) { return threads(); }

double ora_exact_sum(const double *x, int64_t n) {
    ExactSum s;
    for (int64_t i = 0; i < n; ++i) s.add(x[i]);
    return s.result();
}

// Reference-order f64 sum of one slice (sum_lanes, src/query/aggregate.rs:1557-1590).
double ora_sum_lanes(const double *x, int64_t n) {
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int64_t i = 0;
    for (; i + 8 <= n; i += 8)
        for (int k = 0; k < 8; ++k) acc[k] += x[i + k];
    double t = 0.0;
    for (int k = 0; k < 8; ++k) t += acc[k];
    for (; i < n; ++i) t += x[i];
    return t;
}

// checked_sum_i64 (src/query/aggregate.rs:3087-3110): four lanes of checked
// adds, then the lanes, then the tail.  Returns 0 with *out set, or 1 when an
// add leaves i64 (the reference's None).
int ora_checked_sum_i64(const int64_t *x, int64_t n, int64_t *out) {
    int64_t acc[4] = {0, 0, 0, 0};
    int64_t i = 0;
    for (; i + 4 <= n; i += 4)
        for (int k = 0; k < 4; ++k)
            if (__builtin_add_overflow(acc[k], x[i + k], &acc[k])) return 1;
    int64_t total = 0;
    for (int k = 0; k < 4; ++k)
        if (__builtin_add_overflow(total, acc[k], &total)) return 1;
    for (; i < n; ++i)
        if (__builtin_add_overflow(total, x[i], &total)) return 1;
    *out = total;
    return 0;
}

// sum_i64 (src/query/aggregate.rs:3039-3044): the checked total as f64, else
// the f64 lanes (sum_lanes over the values widened with `as f64`).
double ora_sum_i64(const int64_t *x, int64_t n) {
    int64_t t;
    if (ora_checked_sum_i64(x, n, &t) == 0) return (double)t;
    std::vector<double> v((size_t)n);
    for (int64_t i = 0; i < n; ++i) v[i] = (double)x[i];
    return ora_sum_lanes(v.data(), n);
}

// Packed mask, one batch of `batch_rows` at a time, concatenated (batch_rows is
// a multiple of 64 so the concatenation is the table-wide mask).
int ora_filter_mask(const fgpu_column_desc *cd, int ncols, int64_t n, int32_t batch_rows,
                    const fgpu_pred *p, uint64_t *words) {
    if (batch_rows <= 0 || batch_rows % 64) return fail(FGPU_E_INVALID_QUERY, "batch_rows must be a positive multiple of 64");
    auto cols = make_cols(cd, ncols);
    Pred pred;
    if (!build_pred(p, cols, pred)) return fail(FGPU_E_INVALID_QUERY, "malformed predicate");
    const int64_t nb = (n + batch_rows - 1) / batch_rows;
    parallel_for(nb, [&](int64_t bi) {
        const int64_t b = bi * batch_rows, e = std::min(n, b + (int64_t)batch_rows);
        std::vector<uint64_t> m;
        eval_mask(pred, cols, b, e, m);
        std::memcpy(words + b / 64, m.data(), m.size() * 8);
    });
    return FGPU_OK;
}

// COUNT(*) WHERE p: serial over batches, sum of popcounts (src/db.rs:1439-1476).
int ora_filter_count(const fgpu_column_desc *cd, int ncols, int64_t n, int32_t batch_rows,
                     const fgpu_pred *p, int64_t *out) {
    if (batch_rows <= 0 || batch_rows % 64) return fail(FGPU_E_INVALID_QUERY, "batch_rows must be a positive multiple of 64");
    auto cols = make_cols(cd, ncols);
    Pred pred;
    if (!build_pred(p, cols, pred)) return fail(FGPU_E_INVALID_QUERY, "malformed predicate");
    int64_t total = 0;
    std::vector<uint64_t> m;
    for (int64_t b = 0; b < n; b += batch_rows) {
        eval_mask(pred, cols, b, std::min(n, b + (int64_t)batch_rows), m);
        for (auto w : m) total += __builtin_popcountll(w);
    }
    *out = total;
    return FGPU_OK;
}

// Materialised WHERE output of `sel` columns, in row order.
int ora_filter_select(const fgpu_column_desc *cd, int ncols, int64_t n, int32_t batch_rows,
                      const fgpu_pred *p, const int32_t *sel, int nsel, ora_result **out) {
    auto cols = make_cols(cd, ncols);
    Pred pred;
    if (!build_pred(p, cols, pred)) return fail(FGPU_E_INVALID_QUERY, "malformed predicate");
    std::vector<int> needed(sel, sel + nsel);
    for (int c : needed)
        if (c < 0 || c >= ncols) return fail(FGPU_E_COLUMN_NOT_FOUND, "column index out of range");
    auto fbs = scan_filter(cols, n, batch_rows, pred, needed);
    auto *r = new ora_result();
    for (auto &fb : fbs) r->rows += fb.rows;
    r->cols.resize((size_t)nsel);
    for (int k = 0; k < nsel; ++k) {
        OCol &oc = r->cols[k];
        const Col &c = cols[needed[k]];
        oc.name = cd[needed[k]].name ? cd[needed[k]].name : "";
        oc.dtype = c.dtype == FGPU_FLOAT64 ? FGPU_FLOAT64 : c.dtype == FGPU_UTF8 ? FGPU_UTF8 : FGPU_INT64;
        oc.dict_len = c.dict_len;
        oc.dict = c.dict;
        oc.valid.reserve((size_t)r->rows);
        for (auto &fb : fbs) {
            for (int64_t i = 0; i < fb.rows; ++i) {
                const int64_t v = fb.data[k][i];
                if (oc.dtype == FGPU_FLOAT64) oc.f64.push_back(bits_to_f64(v));
                else oc.i64.push_back(v);
                oc.valid.push_back(fb.valid[k].empty() ? 1 : fb.valid[k][i]);
            }
        }
    }
    *out = r;
    return FGPU_OK;
}

// SELECT key, aggs FROM t [WHERE p] GROUP BY key  (key_col >= 0), or the scalar
// form (key_col = -1).  want_exact = 0 skips the exact columns (timed baseline).
int ora_groupby(const fgpu_column_desc *cd, int ncols, int64_t n, int32_t batch_rows,
                const fgpu_pred *p, const fgpu_groupby_spec *spec, int want_exact,
                ora_result **out) {
    if (batch_rows <= 0 || batch_rows % 64) return fail(FGPU_E_INVALID_QUERY, "batch_rows must be a positive multiple of 64");
    auto cols = make_cols(cd, ncols);
    Pred pred;
    if (!build_pred(p, cols, pred)) return fail(FGPU_E_INVALID_QUERY, "malformed predicate");
    const int key = spec->key_col;
    if (key >= ncols) return fail(FGPU_E_COLUMN_NOT_FOUND, "group key column out of range");
    if (key >= 0 && cols[key].dtype == FGPU_FLOAT64)
        return fail(FGPU_E_INVALID_QUERY, "GROUP BY on a Float64 column is not on the fast path");
    // needed columns: key first, then each distinct aggregate input
    std::vector<int> needed;
    if (key >= 0) needed.push_back(key);
    std::vector<AggPlan> plan;
    for (int a = 0; a < spec->n_aggs; ++a) {
        const fgpu_agg &g = spec->aggs[a];
        AggPlan ap{g.func, -1, 0};
        if (g.col >= 0) {
            if (g.col >= ncols) return fail(FGPU_E_COLUMN_NOT_FOUND, "aggregate column out of range");
            auto it = std::find(needed.begin(), needed.end(), g.col);
            if (it == needed.end()) {
                needed.push_back(g.col);
                it = needed.end() - 1;
            }
            ap.input = (int)(it - needed.begin());
            ap.dtype = cols[g.col].dtype;
            if (ap.dtype == FGPU_UTF8 && g.func != FGPU_COUNT)
                return fail(FGPU_E_INVALID_QUERY, "only COUNT over a Utf8 column is on the fast path");
        } else if (g.func != FGPU_COUNT) {
            return fail(FGPU_E_INVALID_QUERY, "only COUNT may take *");
        }
        plan.push_back(ap);
    }
    auto fbs = scan_filter(cols, n, batch_rows, pred, needed);
    auto *r = new ora_result();
    const size_t na = plan.size();
    const size_t first = key >= 0 ? 1 : 0;
    r->cols.resize(first + na);
    for (size_t a = 0; a < na; ++a) r->cols[first + a].name = agg_alias(spec->aggs[a], (int)a);

    std::vector<std::vector<uint32_t>> batch_gids(fbs.size());
    size_t n_groups = 0;
    std::vector<AggState> final_st(na);

    if (key < 0) {
        // execute_agg_only: one group; fold over batches (src/query/aggregate.rs:1394-1528)
        n_groups = 1;
        for (auto &s : final_st) s.resize(1);
        std::vector<std::vector<AggState>> part(fbs.size(), std::vector<AggState>(na));
        parallel_for((int64_t)fbs.size(), [&](int64_t bi) {
            for (auto &s : part[bi]) s.resize(1);
            batch_gids[bi].assign((size_t)fbs[bi].rows, 0);
            accumulate(plan, fbs[bi], batch_gids[bi].data(), 0, fbs[bi].rows, part[bi]);
        });
        for (auto &ps : part)
            for (size_t a = 0; a < na; ++a) merge_entry(plan[a].func, ps[a], 0, final_st[a], 0);
        r->rows = 1;
        ref_values(plan, final_st, 1, *r, first);
        // COUNT with no rows still yields 0, SUM/AVG/MIN/MAX yield NULL.
        if (want_exact) exact_values(plan, fbs, batch_gids, 1, *r, first);
        *out = r;
        return FGPU_OK;
    }

    OCol &kc = r->cols[0];
    kc.name = cd[key].name ? cd[key].name : "";
    const Col &kcol = cols[key];
    if (kcol.dtype == FGPU_UTF8) {
        kc.dtype = kc.ref_dtype = FGPU_UTF8;
        kc.dict_len = kcol.dict_len;
        kc.dict = kcol.dict;
        bool any_null = false;
        for (auto &fb : fbs)
            for (int64_t i = 0; i < fb.rows && !any_null; ++i)
                if ((uint32_t)fb.data[0][i] == NULL_CODE) any_null = true;
        // Group id = dictionary code; the NULL key joins the "" group (Q4) -- here
        // an extra group after the dictionary entries when "" is absent.
        int64_t empty_code = -1;
        for (int i = 0; i < kcol.dict_len; ++i)
            if (kcol.dict[i][0] == 0) { empty_code = i; break; }
        const size_t ng = (size_t)kcol.dict_len + ((any_null && empty_code < 0) ? 1 : 0);
        const uint32_t null_gid = empty_code >= 0 ? (uint32_t)empty_code : (uint32_t)kcol.dict_len;
        for (size_t bi = 0; bi < fbs.size(); ++bi) {
            auto &g = batch_gids[bi];
            g.resize((size_t)fbs[bi].rows);
            for (int64_t i = 0; i < fbs[bi].rows; ++i) {
                const uint32_t c = (uint32_t)fbs[bi].data[0][i];
                g[i] = c == NULL_CODE ? null_gid : c;
            }
        }
        // work units of max(total/(2T), 4096) rows (src/query/aggregate.rs:675-690)
        int64_t total = 0;
        for (auto &fb : fbs) total += fb.rows;
        const int64_t target = std::max<int64_t>(total / std::max(2 * threads(), 1), 4096);
        struct Unit { size_t bi; int64_t s, e; };
        std::vector<Unit> units;
        for (size_t bi = 0; bi < fbs.size(); ++bi)
            for (int64_t s = 0; s < fbs[bi].rows; s += target)
                units.push_back({bi, s, std::min(fbs[bi].rows, s + target)});
        std::vector<std::vector<AggState>> part(units.size(), std::vector<AggState>(na));
        parallel_for((int64_t)units.size(), [&](int64_t u) {
            for (auto &s : part[u]) s.resize(ng);
            const Unit &U = units[u];
            accumulate(plan, fbs[U.bi], batch_gids[U.bi].data() + U.s, U.s, U.e, part[u]);
        });
        for (auto &s : final_st) s.resize(ng);
        for (auto &ps : part)
            for (size_t a = 0; a < na; ++a)
                for (size_t g = 0; g < ng; ++g) merge_entry(plan[a].func, ps[a], g, final_st[a], g);
        // Rows per group decide which groups exist (empty groups omitted, SQL semantics).
        std::vector<int64_t> rows_per(ng, 0);
        for (auto &g : batch_gids)
            for (auto x : g) rows_per[x]++;
        std::vector<uint32_t> keep;
        for (size_t g = 0; g < ng; ++g)
            if (rows_per[g] > 0) keep.push_back((uint32_t)g);
        n_groups = ng;
        ref_values(plan, final_st, ng, *r, 1);
        if (want_exact) exact_values(plan, fbs, batch_gids, ng, *r, 1);
        // compact to kept groups
        auto compact = [&](auto &v) {
            if (v.empty()) return;
            std::remove_reference_t<decltype(v)> o;
            o.reserve(keep.size());
            for (auto g : keep) o.push_back(v[g]);
            v.swap(o);
        };
        for (size_t c = 1; c < r->cols.size(); ++c) {
            compact(r->cols[c].i64); compact(r->cols[c].f64); compact(r->cols[c].valid);
            compact(r->cols[c].ref_i64); compact(r->cols[c].ref_f64);
        }
        kc.i64.clear();
        kc.valid.clear();
        for (auto g : keep) {
            kc.i64.push_back(g < (uint32_t)kcol.dict_len ? (int64_t)g : (int64_t)empty_code);
            kc.valid.push_back(1);
        }
        // the synthetic NULL group prints as "" : mark its code as dict_len
        for (size_t i = 0; i < keep.size(); ++i)
            if (keep[i] >= (uint32_t)kcol.dict_len) kc.i64[i] = kcol.dict_len;
        r->rows = (int64_t)keep.size();
        *out = r;
        return FGPU_OK;
    }

    // Int64 / Int32 key: fast_i64_gb (src/query/aggregate.rs:840-912)
    kc.dtype = kc.ref_dtype = FGPU_INT64;
    std::vector<std::vector<int64_t>> local_keys(fbs.size());
    std::vector<std::vector<AggState>> part(fbs.size(), std::vector<AggState>(na));
    parallel_for((int64_t)fbs.size(), [&](int64_t bi) {
        const FBatch &fb = fbs[bi];
        // as_int64_values: Vec<Option<i64>> copy (src/storage/column.rs:198-204)
        std::vector<std::pair<bool, int64_t>> opt((size_t)fb.rows);
        for (int64_t i = 0; i < fb.rows; ++i)
            opt[i] = {fb.valid[0].empty() || fb.valid[0][i], fb.data[0][i]};
        I64Map m;
        auto &lk = local_keys[bi];
        auto &g = batch_gids[bi];
        g.resize((size_t)fb.rows);
        for (int64_t i = 0; i < fb.rows; ++i) {
            const int64_t k = opt[i].first ? opt[i].second : INT64_MIN;  // NULL -> i64::MIN
            bool ins;
            const uint32_t id = m.get_or_insert(k, (uint32_t)lk.size(), ins);
            if (ins) lk.push_back(k);
            g[i] = id;
        }
        for (auto &s : part[bi]) s.resize(lk.size());
        accumulate(plan, fb, g.data(), 0, fb.rows, part[bi]);
    });
    // single-threaded merge in batch order (src/query/aggregate.rs:889-909)
    I64Map fm;
    std::vector<int64_t> final_keys;
    std::vector<std::vector<uint32_t>> remap(fbs.size());
    for (size_t bi = 0; bi < fbs.size(); ++bi) {
        auto &lk = local_keys[bi];
        remap[bi].resize(lk.size());
        for (size_t li = 0; li < lk.size(); ++li) {
            bool ins;
            const uint32_t fid = fm.get_or_insert(lk[li], (uint32_t)final_keys.size(), ins);
            if (ins) {
                final_keys.push_back(lk[li]);
                for (auto &s : final_st) { s.count.push_back(0); s.sum.push_back(0.0); s.has_val.push_back(0); }
            }
            remap[bi][li] = fid;
            for (size_t a = 0; a < na; ++a) merge_entry(plan[a].func, part[bi][a], li, final_st[a], fid);
        }
    }
    n_groups = final_keys.size();
    ref_values(plan, final_st, n_groups, *r, 1);
    if (want_exact) {
        for (size_t bi = 0; bi < fbs.size(); ++bi)
            for (auto &g : batch_gids[bi]) g = remap[bi][g];
        exact_values(plan, fbs, batch_gids, n_groups, *r, 1);
    }
    kc.i64 = final_keys;
    kc.valid.resize(n_groups);
    for (size_t g = 0; g < n_groups; ++g) kc.valid[g] = final_keys[g] != INT64_MIN;  // :1204-1209
    r->rows = (int64_t)n_groups;
    *out = r;
    return FGPU_OK;
}

}  // extern "C"

// ─── hash join ──────────────────────────────────────────────────────────────
namespace {

inline uint64_t mix33(int64_t k) {  // bucket_with / int_key_hash (src/query/join.rs:1368-1374, :1423-1429)
    uint64_t h = (uint64_t)k;
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    return h;
}

struct IntHashTable {  // src/query/join.rs:1054-1107, :1318-1350
    size_t mask = 0;
    std::vector<uint32_t> heads, next;
    const Col *key = nullptr;
    void init(size_t n, const Col *k) {
        size_t b = 1;
        while (b < std::max<size_t>(n, 1) * 2) b <<= 1;
        mask = b - 1;
        heads.assign(b, NO_ROW);
        next.assign(n, NO_ROW);
        key = k;
    }
    void insert(int64_t k, size_t row) {
        const size_t b = mix33(k) & mask;
        next[row] = heads[b];
        heads[b] = (uint32_t)row;
    }
    template <class F>
    void matches(int64_t k, F f) const {
        for (uint32_t cur = heads[mix33(k) & mask]; cur != NO_ROW; cur = next[cur])
            if (!key->is_null(cur) && key->i64(cur) == k) f(cur);
    }
};

struct BuildFilter {  // src/query/join.rs:1441-1482
    std::vector<uint64_t> bits;
    size_t mask = 0;
    void init(size_t n) {
        size_t want = 1;
        while (want < std::max<size_t>(n, 1) * 16) want <<= 1;
        want = std::max<size_t>(want, 64);
        bits.assign(want / 64, 0);
        mask = want / 64 - 1;
    }
    void insert(uint64_t h) {
        const uint64_t h2 = ((h << 32) | (h >> 32)) | 1;
        bits[(h >> 6) & mask] |= 1ull << (h & 63);
        bits[(h2 >> 6) & mask] |= 1ull << (h2 & 63);
    }
    bool might_contain(uint64_t h) const {
        const uint64_t h2 = ((h << 32) | (h >> 32)) | 1;
        return (bits[(h >> 6) & mask] >> (h & 63) & 1) && (bits[(h2 >> 6) & mask] >> (h2 & 63) & 1);
    }
};

struct JoinInputs {
    std::vector<Col> probe, build;
    int64_t np = 0, nb = 0;
    int pk = 0, bk = 0;
};

}  // namespace

extern "C" {

// SELECT aggs FROM probe [LEFT] JOIN build ON ... / [NOT] EXISTS forms.
// probe_pred filters the probe (FROM) table at scan time (Appendix Q9).
// want_pairs != 0 additionally returns the (probe, build) row pairs.
int ora_join(const fgpu_column_desc *pcd, int pncols, int64_t np,
             const fgpu_column_desc *bcd, int bncols, int64_t nb, int32_t batch_rows,
             const fgpu_pred *probe_pred, const fgpu_join_spec *spec, int want_exact,
             ora_result **out, uint32_t **pairs_l, uint32_t **pairs_r, int64_t *n_pairs) {
    auto pc = make_cols(pcd, pncols);
    auto bc = make_cols(bcd, bncols);
    if (spec->probe_key < 0 || spec->probe_key >= pncols || spec->build_key < 0 ||
        spec->build_key >= bncols)
        return fail(FGPU_E_COLUMN_NOT_FOUND, "join key column out of range");
    const Col &pkc = pc[spec->probe_key];
    const Col &bkc = bc[spec->build_key];
    if ((pkc.dtype != FGPU_INT64 && pkc.dtype != FGPU_INT32) || pkc.dtype != bkc.dtype)
        return fail(FGPU_E_TYPE_MISMATCH, "join keys must be integer columns of one type");
    const int jt = spec->type;
    for (int a = 0; a < spec->n_aggs; ++a) {
        const fgpu_agg &g = spec->aggs[a];
        if (g.col >= 0 && ((g.side == 1 && g.col >= bncols) || (g.side == 0 && g.col >= pncols)))
            return fail(FGPU_E_COLUMN_NOT_FOUND, "aggregate column out of range");
        if (g.side == 1 && (jt == FGPU_SEMI || jt == FGPU_ANTI))
            return fail(FGPU_E_COLUMN_NOT_FOUND, "SEMI/ANTI output has only probe columns");
    }
    Pred pred;
    if (!build_pred(probe_pred, pc, pred)) return fail(FGPU_E_INVALID_QUERY, "malformed predicate");

    // Scan the probe (FROM) table; with a WHERE the surviving rows' ids are kept
    // (materialisation of the needed columns, src/query/select.rs:457-487).
    std::vector<int64_t> prow;  // probe row ids in order
    const bool filtered = probe_pred && probe_pred->n_nodes > 0;
    if (filtered) {
        std::vector<int64_t> cnt;
        const int64_t nbt = (np + batch_rows - 1) / batch_rows;
        std::vector<std::vector<int64_t>> ids((size_t)nbt);
        parallel_for(nbt, [&](int64_t bi) {
            const int64_t b = bi * batch_rows, e = std::min(np, b + (int64_t)batch_rows);
            std::vector<uint64_t> m;
            eval_mask(pred, pc, b, e, m);
            for (size_t w = 0; w < m.size(); ++w)
                for (uint64_t bits = m[w]; bits; bits &= bits - 1)
                    ids[bi].push_back(b + (int64_t)w * 64 + __builtin_ctzll(bits));
        });
        for (auto &v : ids) prow.insert(prow.end(), v.begin(), v.end());
    }
    const int64_t nl = filtered ? (int64_t)prow.size() : np;
    auto lrow = [&](int64_t i) { return filtered ? prow[(size_t)i] : i; };

    // SEMI/ANTI merge both sides first: a full copy of the key column of each
    // (merge_batches, src/query/join.rs:69-70).
    std::vector<int64_t> lkeys_copy;
    std::vector<uint8_t> lnull_copy;
    if (jt == FGPU_SEMI || jt == FGPU_ANTI) {
        lkeys_copy.resize((size_t)nl);
        lnull_copy.resize((size_t)nl);
        parallel_for((nl + 65535) / 65536, [&](int64_t u) {
            for (int64_t i = u * 65536; i < std::min(nl, (u + 1) * 65536); ++i) {
                lkeys_copy[i] = pkc.i64(lrow(i));
                lnull_copy[i] = pkc.is_null(lrow(i));
            }
        });
    }

    // Build: serial insert loop + Bloom filter (src/query/join.rs:607-645).
    IntHashTable ht;
    ht.init((size_t)nb, &bkc);
    BuildFilter bloom;
    const bool use_bloom = nb >= 1024;  // BLOOM_MIN_ROWS
    if (use_bloom) bloom.init((size_t)nb);
    for (int64_t r = 0; r < nb; ++r) {
        if (bkc.is_null(r)) continue;  // a NULL key joins nothing
        const int64_t k = bkc.i64(r);
        ht.insert(k, (size_t)r);
        if (use_bloom) bloom.insert(mix33(k));
    }

    // Probe in parallel units of max(n/T, 8192) rows (src/query/join.rs:311-358, :115-126).
    const int64_t unit = std::max<int64_t>(nl / std::max(threads(), 1), 8192);
    const int64_t n_units = nl ? (nl + unit - 1) / unit : 0;
    std::vector<std::vector<uint32_t>> ul((size_t)n_units), ur((size_t)n_units);
    std::vector<uint8_t> keep;
    if (jt == FGPU_SEMI || jt == FGPU_ANTI) keep.resize((size_t)nl);
    parallel_for(n_units, [&](int64_t u) {
        const int64_t s = u * unit, e = std::min(nl, s + unit);
        std::vector<uint32_t> scratch;
        for (int64_t i = s; i < e; ++i) {
            const bool is_null = (jt == FGPU_SEMI || jt == FGPU_ANTI) ? lnull_copy[i] : pkc.is_null(lrow(i));
            const int64_t k = (jt == FGPU_SEMI || jt == FGPU_ANTI) ? lkeys_copy[i] : pkc.i64(lrow(i));
            scratch.clear();
            if (!is_null && (!use_bloom || bloom.might_contain(mix33(k))))
                ht.matches(k, [&](uint32_t r) { scratch.push_back(r); });
            if (jt == FGPU_SEMI || jt == FGPU_ANTI) {
                keep[i] = (jt == FGPU_ANTI) ? scratch.empty() : !scratch.empty();
                continue;
            }
            if (scratch.empty()) {
                if (jt == FGPU_LEFT) { ul[u].push_back((uint32_t)i); ur[u].push_back(NO_ROW); }
                continue;
            }
            for (uint32_t r : scratch) { ul[u].push_back((uint32_t)i); ur[u].push_back(r); }
        }
    });

    // Output rows as (probe row, build row) pairs over the ORIGINAL probe rows.
    std::vector<uint32_t> L, R;
    if (jt == FGPU_SEMI || jt == FGPU_ANTI) {
        // filter_by_mask builds a usize index list (src/storage/table.rs:214-242)
        std::vector<size_t> idx;
        for (int64_t i = 0; i < nl; ++i)
            if (keep[i]) idx.push_back((size_t)i);
        L.resize(idx.size());
        R.assign(idx.size(), NO_ROW);
        for (size_t i = 0; i < idx.size(); ++i) L[i] = (uint32_t)lrow((int64_t)idx[i]);
    } else {
        size_t tot = 0;
        for (auto &v : ul) tot += v.size();
        L.resize(tot);
        R.resize(tot);
        size_t o = 0;
        for (int64_t u = 0; u < n_units; ++u) {
            for (size_t j = 0; j < ul[u].size(); ++j) {
                L[o] = (uint32_t)lrow(ul[u][j]);
                R[o] = ur[u][j];
                ++o;
            }
        }
    }

    // Aggregate over the join output (execute_agg_only).  Gather each needed
    // column (build_output_dense, src/query/join.rs:403-502), then reduce.
    auto *res = new ora_result();
    res->rows = 1;
    res->cols.resize((size_t)spec->n_aggs);
    const int64_t nout = (int64_t)L.size();
    for (int a = 0; a < spec->n_aggs; ++a) {
        const fgpu_agg &g = spec->aggs[a];
        OCol &oc = res->cols[a];
        oc.name = agg_alias(g, a);
        oc.valid.assign(1, 1);
        if (g.col < 0) {  // COUNT(*): row count (src/query/aggregate.rs:1399-1408)
            oc.dtype = oc.ref_dtype = FGPU_INT64;
            oc.i64.assign(1, nout);
            oc.ref_i64.assign(1, nout);
            continue;
        }
        const Col &c = g.side == 1 ? bc[g.col] : pc[g.col];
        const std::vector<uint32_t> &rows = g.side == 1 ? R : L;
        // gathered column with NULLs (NO_ROW or invalid)
        std::vector<double> vals;
        std::vector<int64_t> ivals;
        vals.reserve((size_t)nout);
        const bool isf = c.dtype == FGPU_FLOAT64;
        for (int64_t i = 0; i < nout; ++i) {
            const uint32_t r = rows[i];
            if (r == NO_ROW || c.is_null(r)) continue;
            vals.push_back(c.f64(r));
            if (!isf) ivals.push_back(c.i64(r));
        }
        const int64_t nv = (int64_t)vals.size();
        oc.ref_dtype = g.func == FGPU_COUNT ? FGPU_INT64 : FGPU_FLOAT64;
        oc.ref_i64.assign(1, nv);
        oc.ref_f64.assign(1, 0.0);
        // reference accumulation (Accumulator::{Sum,Avg,Min,Max}; sum_lanes over
        // 16,384-row chunks when one batch is split, src/query/aggregate.rs:3113-3119)
        if (g.func == FGPU_SUM || g.func == FGPU_AVG) {
            double t = 0.0;
            if (!isf && g.func == FGPU_SUM) {
                // par_sum_i64: per 16,384-value chunk a checked i64 sum, or the f64
                // lanes when it overflows; chunks added in f64 (:3087-3119)
                for (int64_t s = 0; s < nv; s += 16384) {
                    const int64_t e = std::min(nv, s + 16384);
                    __int128 acc = 0;
                    for (int64_t i = s; i < e; ++i) acc += ivals[i];
                    t += i128_fits(acc) ? (double)(int64_t)acc : ora_sum_lanes(vals.data() + s, e - s);
                }
            } else {
                for (int64_t s = 0; s < nv; s += 16384)
                    t += ora_sum_lanes(vals.data() + s, std::min<int64_t>(16384, nv - s));
            }
            oc.ref_f64[0] = g.func == FGPU_AVG ? (nv ? t / (double)nv : 0.0) : t;
        } else if (g.func == FGPU_MIN || g.func == FGPU_MAX) {
            double m = 0.0;
            bool has = false;
            for (double v : vals)
                if (!has || (g.func == FGPU_MIN ? v < m : v > m)) { m = v; has = true; }
            oc.ref_f64[0] = m;
        }
        if (!want_exact) continue;
        oc.valid[0] = (g.func == FGPU_COUNT) || nv > 0;
        if (g.func == FGPU_COUNT) {
            oc.dtype = FGPU_INT64;
            oc.i64.assign(1, nv);
        } else if (!isf && g.func == FGPU_SUM) {
            __int128 s = 0;
            for (int64_t x : ivals) s += x;
            int_sum_column({s}, oc);
        } else if (isf || g.func == FGPU_AVG) {
            oc.dtype = FGPU_FLOAT64;
            double v = 0.0;
            if (g.func == FGPU_SUM || g.func == FGPU_AVG) {
                if (isf) {
                    ExactSum es;
                    for (double x : vals) es.add(x);
                    v = es.result();
                } else {
                    __int128 s = 0;
                    for (int64_t x : ivals) s += x;
                    v = (double)s;
                }
                if (g.func == FGPU_AVG) v = nv ? v / (double)nv : 0.0;
            } else {
                // f64::min / f64::max folding (src/query/aggregate.rs:283-302): NaN
                // skipped, all-NaN -> NaN
                bool has = false;
                for (double x : vals)
                    if (!std::isnan(x) && (!has || (g.func == FGPU_MIN ? x < v : x > v))) { v = x; has = true; }
                if (!has) v = std::numeric_limits<double>::quiet_NaN();
            }
            oc.f64.assign(1, v);
        } else {
            oc.dtype = FGPU_INT64;
            int64_t v = g.func == FGPU_MIN ? INT64_MAX : INT64_MIN;
            for (int64_t x : ivals) v = g.func == FGPU_MIN ? std::min(v, x) : std::max(v, x);
            oc.i64.assign(1, v);
        }
    }
    if (pairs_l && pairs_r && n_pairs) {
        *n_pairs = (int64_t)L.size();
        *pairs_l = (uint32_t *)std::malloc(sizeof(uint32_t) * std::max<size_t>(L.size(), 1));
        *pairs_r = (uint32_t *)std::malloc(sizeof(uint32_t) * std::max<size_t>(R.size(), 1));
        std::memcpy(*pairs_l, L.data(), sizeof(uint32_t) * L.size());
        std::memcpy(*pairs_r, R.data(), sizeof(uint32_t) * R.size());
    }
    *out = res;
    return FGPU_OK;
}

// ORDER BY col [ASC | DESC] LIMIT k over the rows that pass p
// (execute_sort_limit, src/query/sort.rs:59-115; try_typed_sort :209-330).
// Keys: the column's total order (integers with the sign bit flipped, floats
// through f64_ord_key :452-454), inverted for DESC; ties by row; NULL rows
// last for DESC and first for ASC (SortExpr::asc / desc, :28-34), the first
// NULL rows in row order (NullSink :627-641; assemble :773-790).
int ora_topk(const fgpu_column_desc *cd, int ncols, int64_t n, int32_t batch_rows, const fgpu_pred *p,
             int col, int desc, int64_t k, int64_t *rows_out, int64_t *n_out) {
    if (batch_rows <= 0 || batch_rows % 64) return fail(FGPU_E_INVALID_QUERY, "batch_rows must be a positive multiple of 64");
    auto cols = make_cols(cd, ncols);
    if (col < 0 || col >= (int)cols.size()) return fail(FGPU_E_COLUMN_NOT_FOUND, "sort column not found");
    const Col &c = cols[col];
    if (c.dtype == FGPU_UTF8) return fail(FGPU_E_UNSUPPORTED, "ORDER BY a Utf8 column");
    Pred pred;
    if (!build_pred(p, cols, pred)) return fail(FGPU_E_INVALID_QUERY, "malformed predicate");
    std::vector<std::pair<uint64_t, int64_t>> pairs;
    std::vector<int64_t> nulls;
    std::vector<uint64_t> m;
    for (int64_t b = 0; b < n; b += batch_rows) {
        const int64_t e = std::min(n, b + (int64_t)batch_rows);
        eval_mask(pred, cols, b, e, m);
        for (int64_t i = b; i < e; ++i) {
            if (!((m[(i - b) >> 6] >> ((i - b) & 63)) & 1)) continue;
            if (!c.valid(i)) {
                nulls.push_back(i);
                continue;
            }
            uint64_t key;
            if (c.dtype == FGPU_FLOAT64) {
                uint64_t bits;
                std::memcpy(&bits, (const double *)c.values + i, 8);
                key = (bits >> 63) == 0 ? bits ^ (1ull << 63) : ~bits;
            } else if (c.dtype == FGPU_INT32) {
                key = (uint64_t)(int64_t)((const int32_t *)c.values)[i] ^ (1ull << 63);
            } else {
                key = (uint64_t)((const int64_t *)c.values)[i] ^ (1ull << 63);
            }
            pairs.emplace_back(desc ? ~key : key, i);
        }
    }
    std::sort(pairs.begin(), pairs.end());
    std::vector<int64_t> out;
    if (!desc) out.insert(out.end(), nulls.begin(), nulls.end());
    for (const auto &pr : pairs) out.push_back(pr.second);
    if (desc) out.insert(out.end(), nulls.begin(), nulls.end());
    if ((int64_t)out.size() > k) out.resize((size_t)k);
    std::copy(out.begin(), out.end(), rows_out);
    *n_out = (int64_t)out.size();
    return FGPU_OK;
}

// ORDER BY k1 [ASC | DESC], k2 ... [LIMIT limit] over the rows that pass p
// (execute_sort_limit, src/query/sort.rs:59-115): each key compares rows by
// its column's total order -- integers exactly, doubles by f64_ord_key
// (:451-454), Utf8 by the string (Rust's byte-wise String order) -- reversed
// for DESC, with NULL rows after the others when nulls_last and before them
// otherwise (compare_scalar :807-815); rows equal on every key keep row order
// (the typed single-key path's (key, row) pairs, :209-330).
int ora_sort_limit(const fgpu_column_desc *cd, int ncols, int64_t n, int32_t batch_rows, const fgpu_pred *p,
                   const int32_t *key_cols, const int32_t *desc, const int32_t *nulls_last, int n_keys, int64_t limit,
                   int64_t *rows_out, int64_t *n_out) {
    if (batch_rows <= 0 || batch_rows % 64) return fail(FGPU_E_INVALID_QUERY, "batch_rows must be a positive multiple of 64");
    auto cols = make_cols(cd, ncols);
    for (int j = 0; j < n_keys; ++j)
        if (key_cols[j] < 0 || key_cols[j] >= ncols) return fail(FGPU_E_COLUMN_NOT_FOUND, "sort column not found");
    Pred pred;
    if (!build_pred(p, cols, pred)) return fail(FGPU_E_INVALID_QUERY, "malformed predicate");
    std::vector<int64_t> rows;
    std::vector<uint64_t> m;
    for (int64_t b = 0; b < n; b += batch_rows) {
        const int64_t e = std::min(n, b + (int64_t)batch_rows);
        eval_mask(pred, cols, b, e, m);
        for (int64_t i = b; i < e; ++i)
            if ((m[(i - b) >> 6] >> ((i - b) & 63)) & 1) rows.push_back(i);
    }
    // compares row a with row b on key j: <0, 0, >0 before DESC / NULL placement
    auto cmp_key = [&](int j, int64_t a, int64_t b) -> int {
        const Col &c = cols[key_cols[j]];
        const bool na = c.is_null(a), nb = c.is_null(b);
        if (na || nb) {
            if (na && nb) return 0;
            const int r = na ? 1 : -1;  // NULL after the other row
            return nulls_last[j] ? r : -r;
        }
        int r = 0;
        if (c.dtype == FGPU_UTF8) {
            const uint32_t ca = ((const uint32_t *)c.values)[a], cb = ((const uint32_t *)c.values)[b];
            r = std::strcmp(c.dict[ca], c.dict[cb]);  // byte order: strcmp compares as unsigned char
            r = r < 0 ? -1 : r > 0 ? 1 : 0;
        } else if (c.dtype == FGPU_FLOAT64) {
            uint64_t xa, xb;
            std::memcpy(&xa, (const double *)c.values + a, 8);
            std::memcpy(&xb, (const double *)c.values + b, 8);
            xa = (xa >> 63) == 0 ? xa ^ (1ull << 63) : ~xa;
            xb = (xb >> 63) == 0 ? xb ^ (1ull << 63) : ~xb;
            r = xa < xb ? -1 : xa > xb ? 1 : 0;
        } else {
            const int64_t va = c.i64(a), vb = c.i64(b);
            r = va < vb ? -1 : va > vb ? 1 : 0;
        }
        return desc[j] ? -r : r;
    };
    std::stable_sort(rows.begin(), rows.end(), [&](int64_t a, int64_t b) {
        for (int j = 0; j < n_keys; ++j) {
            const int r = cmp_key(j, a, b);
            if (r) return r < 0;
        }
        return false;
    });
    const int64_t k = limit < 0 ? (int64_t)rows.size() : std::min<int64_t>(limit, (int64_t)rows.size());
    std::copy(rows.begin(), rows.begin() + k, rows_out);
    *n_out = k;
    return FGPU_OK;
}

void ora_free(void *p) { std::free(p); }

// ─── result accessors ───────────────────────────────────────────────────────
int64_t ora_result_rows(const ora_result *r) { return r->rows; }
int ora_result_cols(const ora_result *r) { return (int)r->cols.size(); }
// which = 0: exact representation, 1: reference representation.
int ora_result_column(const ora_result *r, int i, int which, fgpu_column_view *v) {
    if (i < 0 || i >= (int)r->cols.size()) return fail(FGPU_E_INDEX_OUT_OF_BOUNDS, "column index");
    const OCol &c = r->cols[i];
    v->name = c.name.c_str();
    const bool has_ref = c.ref_dtype == FGPU_FLOAT64 ? !c.ref_f64.empty() : !c.ref_i64.empty();
    const bool has_exact = c.dtype == FGPU_FLOAT64 ? !c.f64.empty() : !c.i64.empty();
    // which = 0 falls back to the reference representation when exact values
    // were not requested
    const bool ref = has_ref && (which == 1 || !has_exact);
    v->dtype = ref ? c.ref_dtype : c.dtype;
    if (v->dtype == FGPU_FLOAT64) v->values = ref ? (const void *)c.ref_f64.data() : (const void *)c.f64.data();
    else v->values = ref ? (const void *)c.ref_i64.data() : (const void *)c.i64.data();
    v->valid = c.valid.empty() ? nullptr : c.valid.data();
    v->dict_len = c.dict_len;
    v->dict = c.dict;
    return FGPU_OK;
}
void ora_result_free(ora_result *r) { delete r; }

}  // extern "C"
