// fgpu_internal.h — definitions shared by the host runtime and the gfx950 kernels.
//
// Data layout in HBM (one device, one table):
//   column      contiguous array of n_rows values (int32 / int64 / double / u32 code)
//   validity    optional ceil(n/64) LSB-first u64 words (bit set = valid)
//   group table keys[C] (int64, EMPTY = INT64_MIN) + state[(C+1) * S] u64 words,
//               row C of state is the NULL-key group; S words per group
//   join index  keys[Cb] (int64, EMPTY = INT64_MIN) + rows[Cb] (u32 build row)
//
// Exact float sums: a double x = m * 2^e (m a 53-bit integer) is added into a
// fixed-point number with least significant bit 2^L, where L is the smallest
// mantissa-LSB exponent of the column (computed at upload).  The fixed-point
// number is kept in carry-save form: W signed 64-bit words, word k holding
// digits of weight 2^(L + 32k); every addend contributes < 2^32 per word, so up
// to 2^31 additions cannot overflow a word.  The final value is normalised and
// rounded to nearest-even once, so the result is the exactly rounded sum,
// independent of the order of the additions.
#pragma once

#include <stdint.h>

#include <hip/hip_runtime.h>

#include "../../include/falcon_gpu.h"

#define FGPU_DEV __device__ __forceinline__
#define FGPU_HD __host__ __device__ __forceinline__

namespace fgpu {

constexpr int64_t EMPTY_KEY = INT64_MIN;     // empty hash slot / NULL int key sentinel
constexpr uint32_t NULL_CODE = 0xFFFFFFFFu;  // Utf8 NULL code
constexpr uint32_t NO_ROW = 0xFFFFFFFFu;
constexpr uint64_t CANON_NAN = 0x7FF8000000000000ull;  // the quiet NaN results carry
constexpr int MAX_COLS = 16;    // columns one kernel can read
constexpr int MAX_OPS = 16;     // state operations per group
constexpr int MAX_FX_WORDS = 15;  // fixed-point words of one exact float sum (a packed op's W field)
constexpr int BLOCK = 256;      // threads per workgroup for the scan kernels

// Loads and stores through the global address space.  A pointer read from a
// descriptor in memory is generic to the compiler, and a generic (flat)
// access counts against both the vector-memory and the LDS wait counters, so
// every LDS wait after it also waits for HBM.  The pointee must be device
// memory.  Vector types are clang vectors (a class type would be copied
// through a generic reference).
typedef long long ll2 __attribute__((ext_vector_type(2)));
template <class T>
FGPU_DEV T gld(const T *p) {
    return *(const __attribute__((address_space(1))) T *)p;
}
template <class T>
FGPU_DEV void gst(T *p, T v) {
    *(__attribute__((address_space(1))) T *)p = v;
}

// ─── hashing ────────────────────────────────────────────────────────────────
// MurmurHash3 64-bit finaliser (a bijection on 64-bit values).
FGPU_HD uint64_t hash64(uint64_t h) {
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ull;
    h ^= h >> 33;
    return h;
}

// Bijective mix of b-bit values (1 <= b <= 63; mask = 2^b - 1): a xorshift, a
// multiplication by an odd constant modulo 2^b and a xorshift, each a
// bijection of [0, 2^b).  The compact join records key on cmix(key - key_min)
// over the build key range, so equal mixes are equal keys within that range.
// Up to 32 bits the product needs only its low word.
FGPU_HD uint64_t cmix(uint64_t x, int b, uint64_t mask) {
    const int s = (b + 1) >> 1;
    x ^= x >> s;
    if (b <= 32) x = (uint64_t)((uint32_t)x * 0x7F4A7C15u) & mask;
    else x = (x * 0x9E3779B97F4A7C15ull) & mask;
    x ^= x >> s;
    return x;
}

// ─── order-preserving encoding of doubles for atomic min / max ──────────────
FGPU_HD uint64_t f64_order(uint64_t bits) {
    return (bits >> 63) ? ~bits : (bits | 0x8000000000000000ull);
}
FGPU_HD uint64_t f64_unorder(uint64_t enc) {
    return (enc >> 63) ? (enc & 0x7fffffffffffffffull) : ~enc;
}

// ─── exact fixed-point digits of a double ───────────────────────────────────
// Splits finite nonzero x into three signed digits (< 2^32 in magnitude except
// the top one, < 2^22) added at words w0, w0+1, w0+2 of a fixed point number
// with LSB 2^L.  Returns false for zero (nothing to add).  A value that is not
// a multiple of 2^L (a column whose exponents span more than the accumulator
// holds, so L sits above its smallest mantissa LSB) is first rounded to the
// nearest multiple of 2^L, ties to even; for a multiple of 2^L that rounding
// changes nothing.
// m >> s rounded to nearest, ties to even (0 < s; m < 2^53).
FGPU_HD uint64_t rne_shift(uint64_t m, int s) {
    if (s >= 64) return 0;
    const uint64_t q = m >> s;
    return (m + ((1ull << (s - 1)) - 1) + (q & 1)) >> s;
}
// Branch-free form of fx_digits for a finite value (zero yields three zero
// digits).  The exact-sum layout leaves room for digits w0..w0+2 of every
// value of the column, so callers add all three without tests.
FGPU_DEV void fx_digits3(uint64_t bits, int L, int &w0, int64_t &d0, int64_t &d1, int64_t &d2) {
    const int be = (int)((bits >> 52) & 0x7ff);
    uint64_t m = (bits & 0x000fffffffffffffull) | (be ? 0x0010000000000000ull : 0ull);
    const int e = (be ? be : 1) - 1075;
    const int s = L - e;  // > 0: the value's low s bits are below the accumulator's LSB
    m = s > 0 ? rne_shift(m, s) : m;
    const int shift = s > 0 ? 0 : -s;
    w0 = shift >> 5;
    const int off = shift & 31;
    const uint64_t lo = m << off;
    const uint64_t hi = off ? (m >> (64 - off)) : 0ull;
    const bool neg = bits >> 63;
    const int64_t a0 = (int64_t)(lo & 0xffffffffull), a1 = (int64_t)(lo >> 32), a2 = (int64_t)hi;
    d0 = neg ? -a0 : a0;
    d1 = neg ? -a1 : a1;
    d2 = neg ? -a2 : a2;
}

FGPU_HD bool fx_digits(uint64_t bits, int L, int &w0, int64_t &d0, int64_t &d1, int64_t &d2) {
    uint64_t m = bits & 0x000fffffffffffffull;
    const int be = (int)((bits >> 52) & 0x7ff);
    int e;
    if (be == 0) {
        if (m == 0) return false;
        e = -1074;
    } else {
        m |= 0x0010000000000000ull;
        e = be - 1075;
    }
    if (e < L) {  // bits below the accumulator's LSB: round to a multiple of 2^L
        m = rne_shift(m, L - e);
        if (m == 0) return false;
        e = L;
    }
    const int shift = e - L;
    w0 = shift >> 5;
    const int off = shift & 31;
    const uint64_t lo = m << off;
    const uint64_t hi = off ? (m >> (64 - off)) : 0;
    d0 = (int64_t)(lo & 0xffffffffull);
    d1 = (int64_t)(lo >> 32);
    d2 = (int64_t)hi;
    if (bits >> 63) {
        d0 = -d0;
        d1 = -d1;
        d2 = -d2;
    }
    return true;
}

// Converts W carry-save words (LSB weight 2^L) to the nearest double
// (round half to even).  `w` is read-only.
// The words are first read into registers (NW >= W of them, the rest zero),
// then two fully unrolled passes run over the digits: the first finds the
// sign of the total, the second produces the base-2^32 digits of its
// magnitude bottom up and keeps the three digits from the highest nonzero
// one down, plus whether any digit below those is nonzero.  No loop trip
// depends on a loaded value, so nothing waits on a chain of loads.
template <int NW>
FGPU_HD double fx_to_double_n(const int64_t *w, int W, int L) {
    constexpr int ND = NW + 3;  // the total's digits: the words' span plus carries and sign
    int64_t x[NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) x[k] = k < W ? w[k] : 0;
    // pass 1: the top digit of the normalised two's complement integer
    int64_t carry = 0;
    uint32_t top = 0;
#pragma unroll
    for (int k = 0; k < ND; ++k) {
        // t = x[k] + carry may exceed int64 only if |x[k]| ~ 2^63; words stay
        // below 2^63 by construction (see header comment).
        const int64_t t = (k < NW ? x[k < NW ? k : 0] : 0) + carry;
        top = (uint32_t)(t & 0xffffffffll);
        carry = t >> 32;  // arithmetic shift
    }
    const bool neg = (top >> 31) & 1;
    // pass 2: the magnitude's digits (two's complement negated on the fly)
    carry = 0;
    uint64_t nc = 1;                  // carry of the negation's +1
    uint32_t l0 = 0, l1 = 0, l2 = 0;  // the last three digits, l2 the newest
    bool below = false;               // a nonzero digit below l0
    uint32_t x0 = 0, x1 = 0, x2 = 0;  // l0..l2 when the newest nonzero digit came
    bool x_below = false;
    int x_pos = -1;  // index of the highest nonzero digit
#pragma unroll
    for (int k = 0; k < ND; ++k) {
        const int64_t t = (k < NW ? x[k < NW ? k : 0] : 0) + carry;
        carry = t >> 32;
        uint32_t d = (uint32_t)(t & 0xffffffffll);
        if (neg) {
            const uint64_t v = (uint64_t)(uint32_t)~d + nc;
            d = (uint32_t)v;
            nc = v >> 32;
        }
        below = below || l0 != 0;
        l0 = l1;
        l1 = l2;
        l2 = d;
        if (d) {
            x0 = l0;
            x1 = l1;
            x2 = l2;
            x_below = below;
            x_pos = k;
        }
    }
    if (x_pos < 0) return 0.0;
    const int lz = __builtin_clz(x2);
    const int msb = x_pos * 32 + (31 - lz);  // bit index of the leading one
    // the 64 bits from the leading one down, and whether anything below them is set
    const uint64_t hi = ((uint64_t)x2 << 32) | x1;
    const uint64_t win = lz ? (hi << lz) | ((uint64_t)x0 >> (32 - lz)) : hi;
    const bool sticky = x_below || (uint32_t)(x0 << lz) != 0;
    uint64_t mant = win >> 11;             // 53 bits
    const uint64_t rest = win & 0x7ffull;  // 11 bits below the mantissa
    const uint64_t half = 0x400ull;
    int exp2 = msb - 52 + L;  // value = mant * 2^exp2
    if (rest > half || (rest == half && (sticky || (mant & 1)))) {
        ++mant;
        if (mant == (1ull << 53)) {
            mant >>= 1;
            ++exp2;
        }
    }
    // mant < 2^53 exactly representable; scaling by 2^exp2 is exact unless the
    // result overflows (-> inf) or is subnormal (then mant had no bits to round).
    double r = (double)mant;
#ifdef __HIP_DEVICE_COMPILE__
    r = ldexp(r, exp2);
#else
    r = __builtin_ldexp(r, exp2);
#endif
    return neg ? -r : r;
}
FGPU_HD double fx_to_double(const int64_t *w, int W, int L) {
    if (W <= 4) return fx_to_double_n<4>(w, W, L);
    if (W <= 8) return fx_to_double_n<8>(w, W, L);
    return fx_to_double_n<MAX_FX_WORDS>(w, W, L);
}

// Error-free addition (Knuth TwoSum): a + b = s + err exactly.
FGPU_HD void two_sum(double a, double b, double &s, double &err) {
    s = a + b;
    const double bb = s - a;
    err = (a - (s - bb)) + (b - bb);
}

// ─── kernel-side column / predicate / aggregate descriptors ─────────────────
struct DCols {
    const void *v[MAX_COLS];
    const uint64_t *valid[MAX_COLS];
    int32_t dtype[MAX_COLS];
    int n;
};

// Loads value of column c at row r as a 64-bit pattern (int32 sign-extended,
// Utf8 code zero-extended, double bits).
FGPU_DEV int64_t load_raw(const DCols &c, int k, int64_t r) {
    switch (c.dtype[k]) {
        case FGPU_INT32: return (int64_t)gld((const int32_t *)c.v[k] + r);
        case FGPU_UTF8: return (int64_t)gld((const uint32_t *)c.v[k] + r);
        default: return gld((const int64_t *)c.v[k] + r);
    }
}
FGPU_DEV bool load_valid(const DCols &c, int k, int64_t r) {
    if (c.dtype[k] == FGPU_UTF8) return gld((const uint32_t *)c.v[k] + r) != NULL_CODE;
    const uint64_t *vb = c.valid[k];
    return vb ? ((gld(vb + (r >> 6)) >> (r & 63)) & 1) : true;
}

// Comparison classes resolved on the host from (column type, literal type),
// following fast_cmp_mask (src/query/expr.rs:1076-1175).
enum : int8_t {
    CMP_F64 = 0,     // double column vs double literal
    CMP_I64 = 1,     // int64 column vs int64 literal
    CMP_I64_AS_F64 = 2,  // int64 column widened to double
    CMP_I32 = 3,     // int32 column vs literal narrowed with `as i32`
    CMP_DICT = 4,    // Utf8 code -> per-dictionary pass table
    CMP_FALSE = 5,   // unsupported pairing: no row passes
    CMP_IS_NULL = 6,
    CMP_IS_NOT_NULL = 7,
    CMP_IN = 8       // [NOT] IN over a sorted list of lit_i values at `pass`
};
// CMP_IN: op = flavor | IN_NOT.  The list holds int64 values (IN_I64, and
// IN_I32 with the literals narrowed `as i32`) or doubles (IN_F64 for a double
// column, IN_I64_AS_F64 for an int64 column widened to double).
enum : int32_t { IN_I64 = 0, IN_F64 = 1, IN_I64_AS_F64 = 2, IN_I32 = 3, IN_NOT = 4 };

constexpr int MAX_PCOLS = 4;  // distinct columns one predicate may read

struct DPredNode {
    int32_t kind;  // FGPU_P_*
    int32_t cls;   // CMP_*
    int32_t op;    // FGPU_EQ..FGPU_GE
    int32_t col;   // column slot in DCols
    int32_t pc;    // index of that column in DPred::pcol (tile evaluation)
    int64_t lit_i;
    double lit_f;
    const uint8_t *pass;  // CMP_DICT pass table (op = 1: a NULL code passes); CMP_IN list
};
struct DPred {
    int n;        // 0 = no predicate (all rows pass)
    int n_pcols;  // distinct columns the predicate reads
    int32_t pcol[MAX_PCOLS];  // their DCols slots
    DPredNode nodes[FGPU_MAX_PRED_NODES];
    // zone-map pruning (the filter kernels): live[b] = 0 when batch b's bounds
    // prove no row passes; nullptr = every batch is read
    const uint8_t *live;
    int64_t batch_rows;
};

// Zone-map pruning in the scans: true when p.live rules out every batch that
// rows [r0, r1) touch.  Uniform arguments give a uniform answer, so a loop can
// skip the whole tile.
FGPU_DEV bool zone_dead(const uint8_t *live, int64_t batch_rows, int64_t r0, int64_t r1) {
    if (!live || r1 <= r0) return false;
    for (int64_t bi = r0 / batch_rows; bi <= (r1 - 1) / batch_rows; ++bi)
        if (live[bi]) return false;
    return true;
}
FGPU_DEV bool tile_dead(const DPred &p, int64_t r0, int64_t r1) {
    return zone_dead(p.live, p.batch_rows, r0, r1);
}

// One batch's bounds for one column (upload-time zone map): integers as int64;
// floats as f64_order keys over the non-NaN values; valid and NaN counts.
struct BatchZone {
    int64_t lo, hi;
    int64_t n_valid, n_nan;
};

template <class T>
FGPU_DEV bool cmp_op(T x, T v, int op) {
    switch (op) {
        case FGPU_EQ: return x == v;
        case FGPU_NE: return x != v;
        case FGPU_LT: return x < v;
        case FGPU_LE: return x <= v;
        case FGPU_GT: return x > v;
        default: return x >= v;
    }
}

// Whether value v (an integer, or a double's bits for IN_F64) is in the
// sorted list of a CMP_IN node: a lower-bound binary search.
FGPU_DEV bool in_list(const DPredNode &nd, int64_t v) {
    const int fl = nd.op & 3;
    const int n = (int)nd.lit_i;
    int lo = 0, hi = n;
    if (fl == IN_I64 || fl == IN_I32) {
        const int64_t *a = (const int64_t *)nd.pass;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (a[mid] < v) lo = mid + 1;
            else hi = mid;
        }
        return lo < n && a[lo] == v;
    }
    const double d = fl == IN_F64 ? __builtin_bit_cast(double, v) : (double)v;
    const double *a = (const double *)nd.pass;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < d) lo = mid + 1;
        else hi = mid;
    }
    return lo < n && a[lo] == d;
}
// A CMP_IN leaf for a row whose value is v (ok = not NULL).
FGPU_DEV bool in_leaf(const DPredNode &nd, int64_t v, bool ok) {
    return ok && (in_list(nd, v) != ((nd.op & IN_NOT) != 0));
}

FGPU_DEV bool eval_leaf(const DPredNode &nd, const DCols &c, int64_t r) {
    const int k = nd.col;
    if (nd.cls == CMP_IS_NULL) return !load_valid(c, k, r);
    if (nd.cls == CMP_IS_NOT_NULL) return load_valid(c, k, r);
    if (nd.cls == CMP_FALSE) return false;
    bool h;
    switch (nd.cls) {
        case CMP_F64: {
            const double x = ((const double *)c.v[k])[r];
            h = cmp_op(x, nd.lit_f, nd.op);
            break;
        }
        case CMP_I64: h = cmp_op(((const int64_t *)c.v[k])[r], nd.lit_i, nd.op); break;
        case CMP_I64_AS_F64: h = cmp_op((double)((const int64_t *)c.v[k])[r], nd.lit_f, nd.op); break;
        case CMP_I32: h = cmp_op(((const int32_t *)c.v[k])[r], (int32_t)nd.lit_i, nd.op); break;
        case CMP_IN: {
            const int64_t v = c.dtype[k] == FGPU_INT32 ? (int64_t)((const int32_t *)c.v[k])[r]
                                                      : ((const int64_t *)c.v[k])[r];
            return in_leaf(nd, v, load_valid(c, k, r));
        }
        default: {  // CMP_DICT: the pass table has one byte per entry
            const uint32_t code = ((const uint32_t *)c.v[k])[r];
            return code != NULL_CODE ? nd.pass[code] != 0 : nd.op != 0;
        }
    }
    // clear_null_bits: a comparison against NULL is never true
    if (c.valid[k] && !((c.valid[k][r >> 6] >> (r & 63)) & 1)) h = false;
    return h;
}

FGPU_DEV bool eval_pred(const DPred &p, const DCols &c, int64_t r) {
    if (p.n == 0) return true;
    if (p.n == 1) return eval_leaf(p.nodes[0], c, r);
    uint32_t st = 0;  // bit stack
    int sp = 0;
    for (int i = 0; i < p.n; ++i) {
        const DPredNode &nd = p.nodes[i];
        if (nd.kind == FGPU_P_AND || nd.kind == FGPU_P_OR) {
            const uint32_t b = (st >> (sp - 1)) & 1, a = (st >> (sp - 2)) & 1;
            const uint32_t v = nd.kind == FGPU_P_AND ? (a & b) : (a | b);
            sp -= 2;
            st = (st & ~(3u << sp)) | (v << sp);
            ++sp;
        } else {
            const uint32_t v = eval_leaf(nd, c, r) ? 1u : 0u;
            st = (st & ~(1u << sp)) | (v << sp);
            ++sp;
        }
    }
    return (st >> (sp - 1)) & 1;
}

// ─── tile evaluation: all loads of a tile issued before any use ─────────────
// Loads the value (64-bit pattern) and validity of column k at row r.
FGPU_DEV void load_cell(const DCols &c, int k, int64_t r, int64_t &v, bool &ok) {
    switch (c.dtype[k]) {
        case FGPU_INT32: v = (int64_t)gld((const int32_t *)c.v[k] + r); break;
        case FGPU_UTF8: v = (int64_t)gld((const uint32_t *)c.v[k] + r); break;
        default: v = gld((const int64_t *)c.v[k] + r); break;
    }
    const uint64_t *vb = c.valid[k];
    ok = vb ? ((gld(vb + (r >> 6)) >> (r & 63)) & 1) : (c.dtype[k] != FGPU_UTF8 || (uint32_t)v != NULL_CODE);
}

// ─── two-phase column reads ─────────────────────────────────────────────────
// A column resolved once per kernel (uniform: lives in SGPRs).
struct ColRef {
    const uint8_t *base;
    const uint32_t *valid;  // validity bits as 32-bit words, or null
    int dtype;
};
FGPU_DEV ColRef col_ref(const DCols &c, int k) {
    return ColRef{(const uint8_t *)c.v[k], (const uint32_t *)c.valid[k], c.dtype[k]};
}
// The raw loads of one cell.  cell_issue only issues them; nothing reads the
// results until cell_finish, so a caller can keep many loads in flight.
struct RawCell {
    uint64_t bits;
    uint32_t vword;
};
FGPU_DEV void cell_issue(const ColRef &c, int64_t r, RawCell &o) {
    if (c.dtype == FGPU_INT32 || c.dtype == FGPU_UTF8) o.bits = (uint64_t)gld((const uint32_t *)(c.base + r * 4));
    else o.bits = gld((const uint64_t *)(c.base + r * 8));
    o.vword = c.valid ? gld(c.valid + (r >> 5)) : 0xffffffffu;
}
// Issue for row tile0 + idx where tile0 is a multiple of 32: the column and
// validity bases are uniform (SGPR) and the lane offset is 32-bit.
FGPU_DEV void cell_issue_tile(const ColRef &c, int64_t tile0, uint32_t idx, RawCell &o) {
    if (c.dtype == FGPU_INT32 || c.dtype == FGPU_UTF8) o.bits = (uint64_t)gld((const uint32_t *)c.base + tile0 + idx);
    else o.bits = gld((const uint64_t *)c.base + tile0 + idx);
    o.vword = c.valid ? gld(c.valid + (tile0 >> 5) + (idx >> 5)) : 0xffffffffu;
}
FGPU_DEV void cell_finish(const ColRef &c, int64_t r, const RawCell &o, int64_t &v, bool &ok) {
    v = c.dtype == FGPU_INT32 ? (int64_t)(int32_t)(uint32_t)o.bits : (int64_t)o.bits;
    // bitwise, not short-circuit: no branch around the code test
    ok = (((o.vword >> (r & 31)) & 1u) != 0) & ((c.dtype != FGPU_UTF8) | ((uint32_t)o.bits != NULL_CODE));
}

// Two-phase cell reads with a fixed shape per column: one value load (64-bit
// for 8-byte columns, 32-bit otherwise; hi repeats lo for 4-byte columns) and,
// only for a column with NULLs, its validity word.  The column's type and
// validity are uniform, so a loop that keeps several cells in flight has a
// fixed number of loads outstanding and the compiler's waits stay exact.
struct RawCell3 {
    uint32_t lo, hi, vw;
};
FGPU_DEV void cell_issue3(const ColRef &c, int64_t r, RawCell3 &o) {
    if (c.dtype == FGPU_INT64 || c.dtype == FGPU_FLOAT64) {
        const uint64_t x = gld((const uint64_t *)c.base + r);
        o.lo = (uint32_t)x;
        o.hi = (uint32_t)(x >> 32);
    } else {
        o.lo = o.hi = gld((const uint32_t *)c.base + r);
    }
    o.vw = c.valid ? gld(c.valid + (r >> 5)) : 0xffffffffu;
}
// Row tile0 + idx, tile0 a multiple of 32: the bases stay uniform (SGPR) and
// the lane offsets 32-bit.
FGPU_DEV void cell_issue3_tile(const ColRef &c, int64_t tile0, uint32_t idx, RawCell3 &o) {
    if (c.dtype == FGPU_INT64 || c.dtype == FGPU_FLOAT64) {
        const uint64_t x = gld((const uint64_t *)c.base + tile0 + idx);
        o.lo = (uint32_t)x;
        o.hi = (uint32_t)(x >> 32);
    } else {
        o.lo = o.hi = gld((const uint32_t *)c.base + tile0 + idx);
    }
    o.vw = c.valid ? gld(c.valid + (tile0 >> 5) + (idx >> 5)) : 0xffffffffu;
}
FGPU_DEV void cell_finish3(const ColRef &c, int64_t r, const RawCell3 &o, int64_t &v, bool &ok) {
    const bool wide = c.dtype == FGPU_INT64 || c.dtype == FGPU_FLOAT64;
    const uint64_t bits = wide ? ((uint64_t)o.hi << 32 | o.lo) : (uint64_t)o.lo;
    v = c.dtype == FGPU_INT32 ? (int64_t)(int32_t)o.lo : (int64_t)bits;
    const uint32_t vw = c.valid ? o.vw : 0xffffffffu;
    ok = (((vw >> (r & 31)) & 1u) != 0) & ((c.dtype != FGPU_UTF8) | ((uint32_t)bits != NULL_CODE));
}

template <int NPC>
FGPU_DEV int64_t pickn(const int64_t (&v)[NPC], int j) {
    int64_t r = v[0];
#pragma unroll
    for (int k = 1; k < NPC; ++k) r = j == k ? v[k] : r;
    return r;
}

FGPU_DEV bool leaf_reg(const DPredNode &nd, int64_t v, bool ok) {
    switch (nd.cls) {
        case CMP_IS_NULL: return !ok;
        case CMP_IS_NOT_NULL: return ok;
        case CMP_FALSE: return false;
        case CMP_F64: return ok && cmp_op(__builtin_bit_cast(double, v), nd.lit_f, nd.op);
        case CMP_I64: return ok && cmp_op(v, nd.lit_i, nd.op);
        case CMP_I64_AS_F64: return ok && cmp_op((double)v, nd.lit_f, nd.op);
        case CMP_I32: return ok && cmp_op((int32_t)v, (int32_t)nd.lit_i, nd.op);
        case CMP_IN: return in_leaf(nd, v, ok);
        default: {
            const uint32_t code = (uint32_t)v;
            return code != NULL_CODE ? nd.pass[code] != 0 : nd.op != 0;
        }
    }
}

template <int NPC>
FGPU_DEV bool eval_reg(const DPred &p, const int64_t (&v)[NPC], uint32_t ok) {
    if (p.n == 1) return leaf_reg(p.nodes[0], pickn<NPC>(v, p.nodes[0].pc), (ok >> p.nodes[0].pc) & 1);
    uint32_t st = 0;
    int sp = 0;
    for (int i = 0; i < p.n; ++i) {
        const DPredNode &nd = p.nodes[i];
        if (nd.kind == FGPU_P_AND || nd.kind == FGPU_P_OR) {
            const uint32_t b = (st >> (sp - 1)) & 1, a = (st >> (sp - 2)) & 1;
            const uint32_t r = nd.kind == FGPU_P_AND ? (a & b) : (a | b);
            sp -= 2;
            st = (st & ~(3u << sp)) | (r << sp);
            ++sp;
        } else {
            const uint32_t r = leaf_reg(nd, pickn<NPC>(v, nd.pc), (ok >> nd.pc) & 1) ? 1u : 0u;
            st = (st & ~(1u << sp)) | (r << sp);
            ++sp;
        }
    }
    return (st >> (sp - 1)) & 1;
}

// pass[u] = row rows[u] exists (< n) and satisfies p.  Every predicate column
// of every row of the tile is loaded before the first comparison, so a wave
// keeps U * n_pcols independent loads in flight.
// NPC (1, 2 or 4) bounds the predicate columns the caller was compiled for;
// the host launches an instantiation with NPC >= p.n_pcols.
template <int U, int NPC = MAX_PCOLS>
FGPU_DEV void pred_tile(const DPred &p, const DCols &c, const int64_t (&rows)[U], int64_t n,
                        bool (&pass)[U]) {
    if (p.n == 0 || n <= 0) {
#pragma unroll
        for (int u = 0; u < U; ++u) pass[u] = p.n == 0 && rows[u] < n;
        return;
    }
    // uniform column bases; rows past n re-read row n - 1
    ColRef pc[NPC];
#pragma unroll
    for (int j = 0; j < NPC; ++j) pc[j] = col_ref(c, j < p.n_pcols ? p.pcol[j] : 0);
    RawCell3 raw[U][NPC];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int64_t r = rows[u] < n ? rows[u] : n - 1;
#pragma unroll
        for (int j = 0; j < NPC; ++j)
            if (j < p.n_pcols) cell_issue3(pc[j], r, raw[u][j]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int64_t r = rows[u] < n ? rows[u] : n - 1;
        int64_t v[NPC];
        uint32_t ok = 0;
#pragma unroll
        for (int j = 0; j < NPC; ++j) {
            v[j] = 0;
            if (j < p.n_pcols) {
                bool okj;
                cell_finish3(pc[j], r, raw[u][j], v[j], okj);
                ok |= (okj ? 1u : 0u) << j;
            }
        }
        pass[u] = rows[u] < n && eval_reg<NPC>(p, v, ok);
    }
}

// ─── group state operations ─────────────────────────────────────────────────
// Word 0 of every group state is its row count.
enum : int8_t {
    OP_CNT = 1,    // non-NULL count of a column
    OP_ISUM = 2,   // exact integer sum, W words (see isum_total)
    OP_IMIN = 3,
    OP_IMAX = 4,
    OP_FSUM = 5,   // exact fixed-point sum: W words at `word`, then a flags word
    OP_FMIN = 6,   // order-encoded double min (NaN skipped)
    OP_FMAX = 7
};

struct DOp {
    int32_t kind;
    int32_t col;    // column slot
    int32_t W;      // OP_FSUM words
    int32_t has_flags;  // OP_FSUM: column holds NaN/inf, flags word follows the W words
    int32_t word;   // first state word
    int32_t L;      // OP_FSUM fixed-point LSB exponent
    // Register accumulators only (scalar aggregates, joins): OP_FSUM sums the
    // values times rscale = 2^-sc, so that no partial sum overflows.
    int32_t sc;
    double rscale;
};
struct DOps {
    int n;
    int S;  // words per group state
    DOp op[MAX_OPS];
    // op[i] packed into one dword (op_pack): the per-row loops unroll over
    // these with constant indices, so the compiler keeps them in SGPRs.  A
    // scalar load inside such a loop would need s_waitcnt lgkmcnt(0), which
    // also waits for every LDS atomic in flight.
    uint32_t pk[MAX_OPS];
    uint64_t ident[64];  // identity value of each state word (S <= 64)
};

// Packed op: kind 0..2, input column 3..5, W 6..9, has_flags 10, word 11..17,
// L (signed) 18..31.
FGPU_HD uint32_t op_pack(const DOp &o) {
    return (uint32_t)(o.kind & 7) | (uint32_t)(o.col & 7) << 3 | (uint32_t)(o.W & 15) << 6 |
           (uint32_t)(o.has_flags & 1) << 10 | (uint32_t)(o.word & 127) << 11 | (uint32_t)o.L << 18;
}
FGPU_HD bool op_packable(const DOp &o) {
    return o.kind >= 0 && o.kind < 8 && o.col >= 0 && o.col < 8 && o.W >= 0 && o.W < 16 && o.word >= 0 &&
           o.word < 128 && o.L >= -8192 && o.L < 8192;
}
FGPU_DEV int opk_kind(uint32_t d) { return (int)(d & 7u); }
FGPU_DEV int opk_col(uint32_t d) { return (int)((d >> 3) & 7u); }
FGPU_DEV int opk_W(uint32_t d) { return (int)((d >> 6) & 15u); }
FGPU_DEV bool opk_flags(uint32_t d) { return (d >> 10) & 1u; }
FGPU_DEV int opk_word(uint32_t d) { return (int)((d >> 11) & 127u); }
FGPU_DEV int opk_L(uint32_t d) { return (int32_t)d >> 18; }

// Flags word bits for non-finite float inputs.
constexpr uint64_t FL_NAN = 1, FL_PINF = 2, FL_NINF = 4;

// ─── exact integer sums ─────────────────────────────────────────────────────
// An OP_ISUM state has W = 1 or 2 words.
//   W = 1: one word of wrapping 64-bit adds.  The host picks it only when the
//          column's bounds prove |total| <= rows * max|v| < 2^63; the sum mod
//          2^64 is then the total itself.
//   W = 2: carry-save in two words: word 0 adds the low 32 bits of each value
//          (unsigned), word 1 adds the high 32 bits (arithmetic shift).  For up
//          to 2^32 - 1 values neither word can overflow, and the total is
//          word1 * 2^32 + word0 exactly.
// isum_digits gives the two addends of a value for W = 2.
FGPU_HD void isum_digits(int64_t v, uint64_t &lo, int64_t &hi) {
    lo = (uint64_t)(uint32_t)(uint64_t)v;
    hi = v >> 32;
}
// The 128-bit total (t_hi * 2^64 + t_lo) of a W = 2 state.
FGPU_HD void isum_total(uint64_t w0, int64_t w1, uint64_t &t_lo, int64_t &t_hi) {
    const uint64_t s = (uint64_t)w1 << 32;
    t_lo = w0 + s;
    t_hi = (w1 >> 32) + (t_lo < w0 ? 1 : 0);
}
// True when the 128-bit total fits int64 (its value is then (int64_t)t_lo).
FGPU_HD bool isum_fits(uint64_t t_lo, int64_t t_hi) {
    return t_hi == ((int64_t)t_lo >> 63);
}
// The correctly rounded double (round half to even) of a W = 2 state.
FGPU_HD double isum_to_double(uint64_t w0, int64_t w1) {
    // three base-2^32 words: w0 low, w1 low + w0 high, w1 high (signed)
    const int64_t w[3] = {(int64_t)(w0 & 0xffffffffull),
                          (int64_t)((uint64_t)w1 & 0xffffffffull) + (int64_t)(w0 >> 32), w1 >> 32};
    return fx_to_double(w, 3, 0);
}

FGPU_HD double fx_finish(const int64_t *w, int W, int L, uint64_t flags) {
    if (flags) {
        if ((flags & FL_NAN) || ((flags & FL_PINF) && (flags & FL_NINF))) return __builtin_nan("");
        return (flags & FL_PINF) ? __builtin_inf() : -__builtin_inf();
    }
    return fx_to_double(w, W, L);
}

}  // namespace fgpu
