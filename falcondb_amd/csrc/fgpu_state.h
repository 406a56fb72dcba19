// fgpu_state.h — device helpers that update and merge group states
// (shared by the GROUP BY and join kernels).
#pragma once

#include "fgpu_launch.h"

namespace fgpu {

// ─── scoped atomics ────────────────────────────────────────────────────────
template <int SCOPE>
FGPU_DEV void a_add(uint64_t *p, uint64_t v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, SCOPE);
}
template <int SCOPE>
FGPU_DEV void a_smin(uint64_t *p, int64_t v) {
    __hip_atomic_fetch_min((int64_t *)p, v, __ATOMIC_RELAXED, SCOPE);
}
template <int SCOPE>
FGPU_DEV void a_smax(uint64_t *p, int64_t v) {
    __hip_atomic_fetch_max((int64_t *)p, v, __ATOMIC_RELAXED, SCOPE);
}
template <int SCOPE>
FGPU_DEV void a_umin(uint64_t *p, uint64_t v) {
    __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, SCOPE);
}
template <int SCOPE>
FGPU_DEV void a_umax(uint64_t *p, uint64_t v) {
    __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, SCOPE);
}
template <int SCOPE>
FGPU_DEV void a_or(uint64_t *p, uint64_t v) {
    __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, SCOPE);
}

constexpr int LDS_SCOPE = __HIP_MEMORY_SCOPE_WORKGROUP;
constexpr int DEV_SCOPE = __HIP_MEMORY_SCOPE_AGENT;

// The aggregate inputs of one row, as separate registers (an array indexed by a
// runtime op field would be placed in scratch memory).
struct Inputs {
    int64_t v0, v1, v2, v3, v4, v5;
    uint32_t valid;  // bit j set = input j is not NULL
};
static_assert(MAX_IN == 6, "Inputs holds MAX_IN registers");

FGPU_DEV int64_t pick(const Inputs &in, int j) {
    switch (j) {
        case 0: return in.v0;
        case 1: return in.v1;
        case 2: return in.v2;
        case 3: return in.v3;
        case 4: return in.v4;
        default: return in.v5;
    }
}

// The packed ops of a kernel, copied into registers once (kernels that run
// update_row in a loop build it before the loop).
struct OpList {
    int n;
    int check_valid;  // some input may be NULL (else the per-op validity test is skipped)
    uint32_t pk[MAX_OPS];
};
FGPU_DEV OpList op_list(const DOps &ops, bool check_valid = true) {
    OpList l;
    l.n = ops.n;
    l.check_valid = check_valid;
#pragma unroll
    for (int i = 0; i < MAX_OPS; ++i) l.pk[i] = ops.pk[i];
    return l;
}

// Adds one row (input values and validity bits in `in`) into a group state.
template <int SCOPE>
FGPU_DEV void update_row(uint64_t *st, const OpList &ops, const Inputs &in) {
    a_add<SCOPE>(st, 1ull);
#pragma unroll
    for (int i = 0; i < MAX_OPS; ++i) {  // no early exit: fully unrolled, pk stays in SGPRs
        if (i >= ops.n) continue;
        const uint32_t d = ops.pk[i];
        const int col = opk_col(d);
        if (ops.check_valid && !((in.valid >> col) & 1)) continue;
        const int64_t v = pick(in, col);
        uint64_t *w = st + opk_word(d);
        switch (opk_kind(d)) {
            case OP_CNT: a_add<SCOPE>(w, 1ull); break;
            case OP_ISUM:
                if (opk_W(d) == 2) {
                    uint64_t lo;
                    int64_t hi;
                    isum_digits(v, lo, hi);
                    a_add<SCOPE>(w, lo);
                    a_add<SCOPE>(w + 1, (uint64_t)hi);
                } else {
                    a_add<SCOPE>(w, (uint64_t)v);
                }
                break;
            case OP_IMIN: a_smin<SCOPE>(w, v); break;
            case OP_IMAX: a_smax<SCOPE>(w, v); break;
            case OP_FSUM: {
                const uint64_t bits = (uint64_t)v;
                // the flags word exists only for columns holding NaN / inf
                if (opk_flags(d) && ((bits >> 52) & 0x7ff) == 0x7ff) {
                    const uint64_t fl = (bits & 0x000fffffffffffffull) ? FL_NAN
                                        : (bits >> 63)                ? FL_NINF
                                                                      : FL_PINF;
                    a_or<SCOPE>(w + opk_W(d), fl);
                    break;
                }
                int w0;
                int64_t d0, d1, d2;
                fx_digits3(bits, opk_L(d), w0, d0, d1, d2);
                a_add<SCOPE>(w + w0, (uint64_t)d0);
                a_add<SCOPE>(w + w0 + 1, (uint64_t)d1);
                a_add<SCOPE>(w + w0 + 2, (uint64_t)d2);
                break;
            }
            case OP_FMIN:
            case OP_FMAX: {
                const uint64_t bits = (uint64_t)v;
                if (((bits >> 52) & 0x7ff) == 0x7ff && (bits & 0x000fffffffffffffull)) break;  // NaN
                const uint64_t e = f64_order(bits);
                if (opk_kind(d) == OP_FMIN) a_umin<SCOPE>(w, e);
                else a_umax<SCOPE>(w, e);
                break;
            }
        }
    }
}

// update_row for R rows at once, op-major: each op's packed descriptor is
// read once (made uniform) for the R rows, and their atomics issue back to
// back.  Same effect as update_row on each row with ok[r] set (every update is
// an integer add, min, max or or: order-free).
template <int SCOPE, int R>
FGPU_DEV void update_rows(uint64_t *const (&st)[R], const bool (&ok)[R], const DOps &ops, const Inputs (&in)[R]) {
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (ok[r]) a_add<SCOPE>(st[r], 1ull);
    for (int i = 0; i < ops.n; ++i) {
        const uint32_t d = __builtin_amdgcn_readfirstlane(ops.pk[i]);
        const int col = opk_col(d), word = opk_word(d);
        bool on[R];
        int64_t v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            on[r] = ok[r] && ((in[r].valid >> col) & 1);
            v[r] = pick(in[r], col);
        }
        switch (opk_kind(d)) {
            case OP_CNT:
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (on[r]) a_add<SCOPE>(st[r] + word, 1ull);
                break;
            case OP_ISUM:
                if (opk_W(d) == 2) {
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        if (on[r]) {
                            uint64_t lo;
                            int64_t hi;
                            isum_digits(v[r], lo, hi);
                            a_add<SCOPE>(st[r] + word, lo);
                            a_add<SCOPE>(st[r] + word + 1, (uint64_t)hi);
                        }
                } else {
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        if (on[r]) a_add<SCOPE>(st[r] + word, (uint64_t)v[r]);
                }
                break;
            case OP_IMIN:
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (on[r]) a_smin<SCOPE>(st[r] + word, v[r]);
                break;
            case OP_IMAX:
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (on[r]) a_smax<SCOPE>(st[r] + word, v[r]);
                break;
            case OP_FSUM: {
                const bool has_flags = opk_flags(d);
                const int W = opk_W(d), L = opk_L(d);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (!on[r]) continue;
                    const uint64_t bits = (uint64_t)v[r];
                    // the flags word exists only for columns holding NaN / inf
                    if (has_flags && ((bits >> 52) & 0x7ff) == 0x7ff) {
                        const uint64_t fl = (bits & 0x000fffffffffffffull) ? FL_NAN
                                            : (bits >> 63)                ? FL_NINF
                                                                          : FL_PINF;
                        a_or<SCOPE>(st[r] + word + W, fl);
                        continue;
                    }
                    int w0;
                    int64_t d0, d1, d2;
                    fx_digits3(bits, L, w0, d0, d1, d2);
                    a_add<SCOPE>(st[r] + word + w0, (uint64_t)d0);
                    a_add<SCOPE>(st[r] + word + w0 + 1, (uint64_t)d1);
                    a_add<SCOPE>(st[r] + word + w0 + 2, (uint64_t)d2);
                }
                break;
            }
            case OP_FMIN:
            case OP_FMAX: {
                const bool is_min = opk_kind(d) == OP_FMIN;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const uint64_t bits = (uint64_t)v[r];
                    if (!on[r] || (((bits >> 52) & 0x7ff) == 0x7ff && (bits & 0x000fffffffffffffull))) continue;
                    const uint64_t e = f64_order(bits);
                    if (is_min) a_umin<SCOPE>(st[r] + word, e);
                    else a_umax<SCOPE>(st[r] + word, e);
                }
                break;
            }
        }
    }
}

// Two-phase reads of a row's key and inputs: gb_issue issues the loads of
// row r from uniform column bases (one value load per cell, a validity word
// only for a column with NULLs) into a GbRow; gb_finish turns them into what
// load_key / load_inputs return, so the loads of several rows are in flight
// together.
struct GbCols {
    ColRef key;
    ColRef in[MAX_IN];
};
FGPU_DEV GbCols gb_cols(const GbArgs &a) {
    GbCols g;
    g.key = col_ref(a.cols, a.key_col >= 0 ? a.key_col : 0);
#pragma unroll
    for (int j = 0; j < MAX_IN; ++j) g.in[j] = col_ref(a.cols, j < a.n_in ? a.in_col[j] : 0);
    return g;
}
struct GbRow {
    int64_t key;
    uint32_t kvw;
    int64_t v[MAX_IN];
    uint32_t vw[MAX_IN];
};
FGPU_DEV void gb_cell_issue(const ColRef &c, int64_t r, int64_t &v, uint32_t &vw) {
    if (c.dtype == FGPU_INT64 || c.dtype == FGPU_FLOAT64) v = gld((const int64_t *)c.base + r);
    else v = (int64_t)gld((const uint32_t *)c.base + r);
    vw = c.valid ? gld(c.valid + (r >> 5)) : 0xffffffffu;
}
// load_raw / load_valid of an issued cell
FGPU_DEV bool gb_cell_finish(const ColRef &c, int64_t r, int64_t &v, uint32_t vw) {
    if (c.dtype == FGPU_INT32) v = (int64_t)(int32_t)(uint32_t)v;
    return (((vw >> (r & 31)) & 1u) != 0) && (c.dtype != FGPU_UTF8 || (uint32_t)v != NULL_CODE);
}
FGPU_DEV void gb_issue(const GbArgs &a, const GbCols &g, int64_t r, GbRow &x) {
    x.key = 0;
    x.kvw = 0;
    if (a.key_col >= 0) gb_cell_issue(g.key, r, x.key, x.kvw);
#pragma unroll
    for (int j = 0; j < MAX_IN; ++j) {
        x.v[j] = 0;
        x.vw[j] = 0;
        if (j < a.n_in) gb_cell_issue(g.in[j], r, x.v[j], x.vw[j]);
    }
}
// The inputs of an issued row (load_inputs' result).
FGPU_DEV Inputs gb_finish_inputs(const GbArgs &a, const GbCols &g, int64_t r, const GbRow &x) {
    Inputs in{0, 0, 0, 0, 0, 0, 0};
    int64_t v[MAX_IN];
#pragma unroll
    for (int j = 0; j < MAX_IN; ++j) {
        v[j] = x.v[j];
        if (j < a.n_in && gb_cell_finish(g.in[j], r, v[j], x.vw[j])) in.valid |= 1u << j;
    }
    in.v0 = v[0];
    in.v1 = v[1];
    in.v2 = v[2];
    in.v3 = v[3];
    in.v4 = v[4];
    in.v5 = v[5];
    return in;
}
// Returns has_key (load_key's result; false without a key column).
FGPU_DEV bool gb_finish(const GbArgs &a, const GbCols &g, int64_t r, GbRow &x, int64_t &key, Inputs &in) {
    bool has_key = false;
    key = 0;
    if (a.key_col >= 0) {
        if (a.key_is_utf8) {
            const uint32_t code = (uint32_t)x.key;
            key = code == NULL_CODE ? (int64_t)a.utf8_null_code : (int64_t)code;
            has_key = true;
        } else {
            int64_t v = x.key;
            const bool ok = gb_cell_finish(g.key, r, v, x.kvw);
            key = v;
            has_key = ok && v != EMPTY_KEY;
        }
    }
    in = gb_finish_inputs(a, g, r, x);
    return has_key;
}

// Fills n_rows group states of S words with the ops' identity values.
// Consecutive threads write consecutive words; each thread steps its word
// index instead of dividing per word (a 64-bit modulo is a long sequence).
FGPU_DEV void fill_ident(uint64_t *dst, int64_t n_rows, int S, const uint64_t *ident) {
    const int64_t total = n_rows * S;
    const int step = (int)(blockDim.x % (unsigned)S);
    int w = (int)(threadIdx.x % (unsigned)S);
    for (int64_t i = threadIdx.x; i < total; i += blockDim.x) {
        dst[i] = ident[w];
        w += step;
        if (w >= S) w -= S;
    }
}
FGPU_DEV void fill_ident(uint64_t *dst, int64_t n_rows, int S, const DOps &ops) {
    fill_ident(dst, n_rows, S, ops.ident);
}

// Exclusive prefix of v over the workgroup and the workgroup total.  Every
// thread of the block calls it (blockDim.x a multiple of 64, at most 1024).
FGPU_DEV unsigned block_excl_scan(unsigned v, unsigned &total) {
    __shared__ unsigned wsum[17];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    unsigned x = v;
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned t = __shfl_up(x, d, 64);
        if (lane >= d) x += t;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned run = 0;
        for (int w = 0; w < nw; ++w) {
            const unsigned c = wsum[w];
            wsum[w] = run;
            run += c;
        }
        wsum[16] = run;
    }
    __syncthreads();
    total = wsum[16];
    const unsigned r = wsum[wave] + x - v;
    __syncthreads();
    return r;
}

// Reserves `total` output rows for the workgroup with one device atomic (one
// atomic per wave on a shared counter serialises across thousands of waves).
FGPU_DEV unsigned long long block_reserve(unsigned long long *counter, unsigned total) {
    __shared__ unsigned long long base;
    if (threadIdx.x == 0) base = total ? atomicAdd(counter, (unsigned long long)total) : 0ull;
    __syncthreads();
    const unsigned long long b = base;
    __syncthreads();
    return b;
}

template <int SCOPE>
FGPU_DEV void update_row(uint64_t *st, const DOps &ops, const Inputs &in) {
    update_row<SCOPE>(st, op_list(ops), in);
}

// Merges a whole state (src, e.g. an LDS partial) into dst (HBM by default).
template <int SCOPE = DEV_SCOPE>
FGPU_DEV void merge_state(uint64_t *dst, const uint64_t *src, const DOps &ops) {
    a_add<SCOPE>(dst, src[0]);
#pragma unroll
    for (int i = 0; i < MAX_OPS; ++i) {  // no early exit: fully unrolled, pk stays in SGPRs
        if (i >= ops.n) continue;
        const uint32_t d = ops.pk[i];
        uint64_t *w = dst + opk_word(d);
        const uint64_t *s = src + opk_word(d);
        switch (opk_kind(d)) {
            case OP_CNT:
                if (s[0]) a_add<SCOPE>(w, s[0]);
                break;
            case OP_ISUM:
                if (s[0]) a_add<SCOPE>(w, s[0]);
                if (opk_W(d) == 2 && s[1]) a_add<SCOPE>(w + 1, s[1]);
                break;
            case OP_IMIN: a_smin<SCOPE>(w, (int64_t)s[0]); break;
            case OP_IMAX: a_smax<SCOPE>(w, (int64_t)s[0]); break;
            case OP_FSUM: {
                const int W = opk_W(d);
                for (int k = 0; k < W; ++k)
                    if (s[k]) a_add<SCOPE>(w + k, s[k]);
                if (opk_flags(d) && s[W]) a_or<SCOPE>(w + W, s[W]);
                break;
            }
            case OP_FMIN: a_umin<SCOPE>(w, s[0]); break;
            case OP_FMAX: a_umax<SCOPE>(w, s[0]); break;
        }
    }
}


// ─── per-thread register accumulators (scalar aggregates, joins) ────────────
// One op's accumulator is an int64 `acc` and a double pair hi + lo:
//   COUNT, MIN, MAX   acc (MIN/MAX of doubles as f64_order keys; NaN skipped)
//   ISUM, W = 1       acc (wrapping; the host proved the total fits)
//   ISUM, W = 2       acc = sum of the values' low 32 bits, and the bits of
//                     `hi` hold the sum of their high 32 bits (isum_digits)
//   FSUM              hi + lo: a double-double sum of the values times
//                     2^-sc (sc > 0 only for columns near the top of the double
//                     range, so that no partial sum can overflow); acc ORs the
//                     FL_* flags of non-finite inputs
// A thread's register sums are then folded exactly into the workgroup's state
// row.  The double-double sum is exact while each error term adds into `lo`
// without rounding; otherwise its error is below 2^-104 of the sum of the
// thread's absolute values.
FGPU_DEV int64_t dbits(double x) { return __double_as_longlong(x); }
FGPU_DEV double bitsd(int64_t x) { return __longlong_as_double(x); }

// Identity of a register accumulator for op kind k.
FGPU_DEV int64_t reg_ident(int k) {
    return k == OP_IMIN ? INT64_MAX : k == OP_IMAX ? INT64_MIN : k == OP_FMIN ? (int64_t)~0ull : 0;
}

// Adds value v (raw 64-bit pattern) into the register accumulator of an op of
// kind k (W words; rscale = 2^-sc for FSUM).
FGPU_DEV void reg_add(int k, int W, double rscale, int64_t v, int64_t &acc, double &hi, double &lo) {
    switch (k) {
        case OP_CNT: acc += 1; break;
        case OP_ISUM:
            if (W == 2) {
                uint64_t dl;
                int64_t dh;
                isum_digits(v, dl, dh);
                acc = (int64_t)((uint64_t)acc + dl);
                hi = bitsd(dbits(hi) + dh);
            } else {
                acc = (int64_t)((uint64_t)acc + (uint64_t)v);
            }
            break;
        case OP_IMIN: acc = v < acc ? v : acc; break;
        case OP_IMAX: acc = v > acc ? v : acc; break;
        case OP_FSUM: {
            const uint64_t bits = (uint64_t)v;
            if (((bits >> 52) & 0x7ff) == 0x7ff) {
                acc |= (bits & 0x000fffffffffffffull) ? FL_NAN : (bits >> 63) ? FL_NINF : FL_PINF;
            } else {
                double s, e;
                two_sum(hi, bitsd(v) * rscale, s, e);
                hi = s;
                lo += e;
            }
            break;
        }
        case OP_FMIN:
        case OP_FMAX: {
            const uint64_t bits = (uint64_t)v;
            if (((bits >> 52) & 0x7ff) == 0x7ff && (bits & 0x000fffffffffffffull)) break;
            const uint64_t e = f64_order(bits);
            if (k == OP_FMIN) acc = (int64_t)(e < (uint64_t)acc ? e : (uint64_t)acc);
            else acc = (int64_t)(e > (uint64_t)acc ? e : (uint64_t)acc);
            break;
        }
    }
}
FGPU_DEV void reg_add(const DOp &o, int64_t v, int64_t &acc, double &hi, double &lo) {
    reg_add(o.kind, o.W, o.rscale, v, acc, hi, lo);
}

// Folds one thread's register accumulators for op `o` into an LDS state row.
FGPU_DEV void fold_reg(uint64_t *row, const DOp &o, int64_t acc, double hi, double lo) {
    uint64_t *w = row + o.word;
    switch (o.kind) {
        case OP_CNT:
            if (acc) a_add<LDS_SCOPE>(w, (uint64_t)acc);
            break;
        case OP_ISUM:
            if (acc) a_add<LDS_SCOPE>(w, (uint64_t)acc);
            if (o.W == 2 && dbits(hi)) a_add<LDS_SCOPE>(w + 1, (uint64_t)dbits(hi));
            break;
        case OP_IMIN: a_smin<LDS_SCOPE>(w, acc); break;
        case OP_IMAX: a_smax<LDS_SCOPE>(w, acc); break;
        case OP_FMIN: a_umin<LDS_SCOPE>(w, (uint64_t)acc); break;
        case OP_FMAX: a_umax<LDS_SCOPE>(w, (uint64_t)acc); break;
        case OP_FSUM: {
            if (acc) a_or<LDS_SCOPE>(w + o.W, (uint64_t)acc);
            // hi and lo are finite multiples of 2^(L - sc): digits relative to
            // 2^(L - sc) of the scaled sums are digits relative to 2^L of the sums
            const double parts[2] = {hi, lo};
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                int w0;
                int64_t d0, d1, d2;
                if (!fx_digits((uint64_t)dbits(parts[q]), o.L - o.sc, w0, d0, d1, d2)) continue;
                if (d0) a_add<LDS_SCOPE>(w + w0, (uint64_t)d0);
                if (d1) a_add<LDS_SCOPE>(w + w0 + 1, (uint64_t)d1);
                if (d2) a_add<LDS_SCOPE>(w + w0 + 2, (uint64_t)d2);
            }
            break;
        }
    }
}

// One output of a group state: its value bits and validity byte.
FGPU_DEV void emit_one(const DOut &d, const uint64_t *st, int64_t &out, uint8_t &valid) {
    const int64_t cnt = (int64_t)st[d.cnt_word];
    const uint64_t w = st[d.word];
    out = 0;
    valid = cnt > 0;
    switch (d.kind) {
        case OUT_COUNT_ROWS: out = (int64_t)st[0]; valid = 1; break;
        case OUT_COUNT: out = cnt; valid = 1; break;
        case OUT_SUM_I:
            out = (int64_t)w;
            if (d.W == 2) {  // the exact total, or (valid = 2) its rounded double
                uint64_t t_lo;
                int64_t t_hi;
                isum_total(w, (int64_t)st[d.word + 1], t_lo, t_hi);
                out = (int64_t)t_lo;
                if (!isum_fits(t_lo, t_hi)) {
                    out = dbits(isum_to_double(w, (int64_t)st[d.word + 1]));
                    valid = valid ? 2 : 0;
                }
            }
            break;
        case OUT_MIN_I:
        case OUT_MAX_I: out = (int64_t)w; break;
        case OUT_AVG_I: {
            const double s2 = d.W == 2 ? isum_to_double(w, (int64_t)st[d.word + 1]) : (double)(int64_t)w;
            out = dbits(cnt ? s2 / (double)cnt : 0.0);
            break;
        }
        case OUT_SUM_F:
        case OUT_AVG_F: {
            double s2 = fx_finish((const int64_t *)(st + d.word), d.W, d.L, d.has_flags ? st[d.word + d.W] : 0ull);
            if (d.kind == OUT_AVG_F) s2 = cnt ? s2 / (double)cnt : 0.0;
            out = __double_as_longlong(s2);
            break;
        }
        // MIN / MAX fold like f64::min / f64::max: NaN values are skipped, a
        // group whose values are all NaN yields NaN (the state still holds its
        // identity), a group without values NULL
        case OUT_MIN_F: out = w == ~0ull ? (int64_t)CANON_NAN : (int64_t)f64_unorder(w); break;
        case OUT_MAX_F: out = w == 0ull ? (int64_t)CANON_NAN : (int64_t)f64_unorder(w); break;
    }
}

// Converts one group state into its output row o (one value + validity byte
// per aggregate).
FGPU_DEV void emit_row(const DOuts &outs, const uint64_t *st, int64_t o, int64_t *const *out_val,
                       uint8_t *const *out_valid) {
    for (int k = 0; k < outs.n; ++k) {
        int64_t out;
        uint8_t valid;
        emit_one(outs.o[k], st, out, valid);
        gst(out_val[k] + o, out);
        gst(out_valid[k] + o, valid);
    }
}

}  // namespace fgpu
