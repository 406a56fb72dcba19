// fgpu_gbpart.hip — radix-partitioned GROUP BY for high-cardinality keys on
// gfx950 (see GbPartArgs).  Every update on the hot path is an LDS operation:
// device-scope atomics are limited to one output reservation per workgroup.
//
// Traffic per input row: the scanned columns once; per cold row (a key the
// LDS cache did not hold) one record written in pass 1 and read in pass 2.
#include <type_traits>

#include "fgpu_state.h"

namespace fgpu {

FGPU_DEV int gbp_part(uint64_t h, int log_p) { return log_p ? (int)(h >> (64 - log_p)) : 0; }

// Index in [0, n) from 32 hash bits (multiply-high: n need not be a power of two).
FGPU_DEV int gbp_range(uint32_t bits, int n) { return (int)(((uint64_t)bits * (uint32_t)n) >> 32); }

// Find-or-insert in an LDS table of n slots with linear probing (at most
// `limit` probes); the start slot comes from hash bits 0..31, independent of
// the partition bits at the top.
FGPU_DEV int gbp_slot(int64_t *keys, int n, int limit, int64_t key, uint64_t h) {
    int i = gbp_range((uint32_t)h, n);
    for (int p = 0; p < limit; ++p) {
        const int64_t k = __hip_atomic_load(keys + i, __ATOMIC_RELAXED, LDS_SCOPE);
        if (k == key) return i;
        if (k == EMPTY_KEY) {
            int64_t expect = EMPTY_KEY;
            if (__hip_atomic_compare_exchange_strong(keys + i, &expect, key, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, LDS_SCOPE))
                return i;
            if (expect == key) return i;
        }
        i = i + 1 == n ? 0 : i + 1;
    }
    return -1;
}

// ─── pass 1 ─────────────────────────────────────────────────────────────────
// Each wave streams tiles of 64*U rows through a three-stage pipeline:
// predicate columns of tile t+2 and the key/inputs of tile t+1 (loaded only
// for rows that passed) are in flight while tile t is aggregated.  Passing
// rows are compacted across tiles into dense 64-lane chunks with ds_permute,
// so the hash / LDS work runs on full waves whatever the selectivity.
//
// The hot-key cache is 8-way set associative: one bucket (64 bytes) is read
// per lookup.  Ways fill in order and are never emptied, so a key occupies at
// most one way.  A key enters the cache on its second sighting (doorkeeper
// bitmap), which keeps one-off keys from taking the ways of heavy hitters.

constexpr int GBP_DOOR_WORDS = 2048;  // doorkeeper bitmap: 65,536 bits of LDS (hash bits 0..15)
constexpr int GBP_WAYS = 8;
constexpr uint32_t GBP_HAS_KEY = 1u << 30;  // record validity word: key is not NULL
typedef long long gbp_ll2 __attribute__((ext_vector_type(2)));

// Bit u of the result = operand rows u of the tile satisfy comparison leaf nd.
template <int U, int NPC>
FGPU_DEV uint32_t leaf_mask(const DPredNode nd, const int64_t (&v)[U][NPC], const uint32_t (&ok)[U]) {
    uint32_t m = 0;
    const int pc = nd.pc;
    switch (nd.cls) {
        case CMP_IS_NULL:
#pragma unroll
            for (int u = 0; u < U; ++u) m |= ((ok[u] >> pc) & 1u) ? 0u : 1u << u;
            return m;
        case CMP_IS_NOT_NULL:
#pragma unroll
            for (int u = 0; u < U; ++u) m |= ((ok[u] >> pc) & 1u) << u;
            return m;
        case CMP_FALSE: return 0;
        case CMP_DICT:
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t code = (uint32_t)pickn<NPC>(v[u], pc);
                m |= (code != NULL_CODE ? nd.pass[code] != 0 : nd.op != 0) ? 1u << u : 0u;
            }
            return m;
        case CMP_IN:
#pragma unroll
            for (int u = 0; u < U; ++u) m |= in_leaf(nd, pickn<NPC>(v[u], pc), (ok[u] >> pc) & 1u) ? 1u << u : 0u;
            return m;
        default: break;
    }
    // (lt, eq, gt) selected by the operator; NE is the complement of EQ, which
    // keeps NaN semantics (NaN != x holds, every other comparison fails).
    const int op = nd.op;
    const bool s_lt = op == FGPU_LT || op == FGPU_LE, s_gt = op == FGPU_GT || op == FGPU_GE;
    const bool s_eq = op == FGPU_EQ || op == FGPU_NE || op == FGPU_LE || op == FGPU_GE;
    const bool inv = op == FGPU_NE;
    if (nd.cls == CMP_F64 || nd.cls == CMP_I64_AS_F64) {
        const bool bits = nd.cls == CMP_F64;
        const double lf = nd.lit_f;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t x = pickn<NPC>(v[u], pc);
            const double d = bits ? __builtin_bit_cast(double, x) : (double)x;
            const bool r = ((s_lt && d < lf) || (s_eq && d == lf) || (s_gt && d > lf)) != inv;
            m |= (((ok[u] >> pc) & 1u) && r) ? 1u << u : 0u;
        }
    } else {  // CMP_I64, CMP_I32 (the literal narrowed with `as i32`)
        const int64_t li = nd.cls == CMP_I32 ? (int64_t)(int32_t)nd.lit_i : nd.lit_i;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t x = pickn<NPC>(v[u], pc);
            const bool r = ((s_lt && x < li) || (s_eq && x == li) || (s_gt && x > li)) != inv;
            m |= (((ok[u] >> pc) & 1u) && r) ? 1u << u : 0u;
        }
    }
    return m;
}

// Bit u = row u of the tile satisfies the postfix predicate.  Each node is
// decoded once per tile.  The stack holds U-bit masks, 8 bits per entry: a
// postfix of at most 16 binary nodes has at most 8 leaves, so depth <= 8.
template <int U, int NPC>
FGPU_DEV uint32_t pred_mask_tile(const DPred &p, const int64_t (&v)[U][NPC], const uint32_t (&ok)[U]) {
    static_assert(U <= 8, "masks are 8 bits wide");
    if (p.n == 0) return (1u << U) - 1u;
    if (p.n == 1) return leaf_mask<U, NPC>(p.nodes[0], v, ok);
    uint64_t st = 0;
    for (int i = 0; i < p.n; ++i) {
        const DPredNode nd = p.nodes[i];
        uint32_t m;
        if (nd.kind == FGPU_P_AND || nd.kind == FGPU_P_OR) {
            const uint32_t b = (uint32_t)st & 0xffu;
            const uint32_t a = (uint32_t)(st >> 8) & 0xffu;
            st >>= 16;
            m = nd.kind == FGPU_P_AND ? (a & b) : (a | b);
        } else {
            m = leaf_mask<U, NPC>(nd, v, ok);
        }
        st = (st << 8) | m;
    }
    return (uint32_t)st & 0xffu;
}

// One lane's row as it travels through the compaction.
template <int NIN>
struct GbpRow {
    int64_t key;
    int64_t in[NIN > 0 ? NIN : 1];
    uint32_t vw;  // bit j: input j not NULL; GBP_HAS_KEY: key not NULL
};

// Columns and parameters of one scan, resolved into registers at kernel start
// so that the tile loop issues no scalar loads (their waits would also wait
// for the LDS atomics in flight).
template <int NPC, int NIN>
struct GbpScan {
    ColRef pc[NPC];
    int n_pcols;
    DPredNode leaf0;  // the predicate's only node when it has one
    ColRef key;
    ColRef in[NIN > 0 ? NIN : 1];
    OpList ops;
    int S, H, NB, log_p, Cw, RW, has_valid_word;
    int pack, pk_kbits;  // packed row records (GbPartArgs)
    int64_t pk_kmin, pk_vmin;
    unsigned long long *flags;
    int64_t *my;  // this workgroup's cells
    const uint8_t *live;  // zone maps: batches that may hold a passing row (null: all)
    int64_t batch_rows;
};

template <int U, int NPC>
struct GbpPredRaw {
    RawCell3 c[U][NPC];
    bool dead;  // the zone maps rule the tile out: nothing was loaded
};

template <int U, int NIN>
struct GbpDataRaw {
    RawCell3 key[U];
    RawCell3 in[U][NIN > 0 ? NIN : 1];
    uint32_t pass;
};

template <int U, int NIN>
struct GbpDataTile {
    GbpRow<NIN> r[U];
    uint32_t pass;  // bit u: row u passed
};

FGPU_DEV int64_t perm64(int addr, int64_t v) {
    const int lo = __builtin_amdgcn_ds_permute(addr, (int)(uint32_t)(uint64_t)v);
    const int hi = __builtin_amdgcn_ds_permute(addr, (int)(uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint64_t)(uint32_t)lo);
}

// Tiles start at multiples of 128 rows (workgroup ranges are 256-aligned), so
// a row's validity bit is bit (lane & 31) of its 32-bit validity word.

// Row index within the tile for predicate loads, clamped into the table.
FGPU_DEV uint32_t gbp_pred_idx(int64_t base, int u, int64_t r1) {
    const uint32_t i = (uint32_t)(u * 64 + (threadIdx.x & 63));
    const uint32_t last = (uint32_t)(r1 - base - 1);
    return i < last ? i : last;
}

template <int U, int NPC, int NIN>
FGPU_DEV void gbp_issue_pred(const GbpScan<NPC, NIN> &g, int64_t base, int64_t r1, GbpPredRaw<U, NPC> &t) {
    t.dead = zone_dead(g.live, g.batch_rows, base, base + 64 * U < r1 ? base + 64 * U : r1);
    if (t.dead) return;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t i = gbp_pred_idx(base, u, r1);
#pragma unroll
        for (int j = 0; j < NPC; ++j)
            cell_issue3_tile(g.pc[j], base, i, t.c[u][j]);  // unused slots re-read column 0: loads stay fixed
    }
}

// Bit u = row u of the tile exists and satisfies the predicate.
template <int U, int NPC, int NIN>
FGPU_DEV uint32_t gbp_finish_pred(const GbArgs &a, const GbpScan<NPC, NIN> &g, int64_t base, int64_t r1,
                                  const GbpPredRaw<U, NPC> &t) {
    if (t.dead) return 0;
    int64_t v[U][NPC];
    uint32_t ok[U];
    uint32_t exists = 0;
    const int64_t r = threadIdx.x & 31;  // validity bit of every row of the lane
#pragma unroll
    for (int u = 0; u < U; ++u) {
        exists |= base + u * 64 + (threadIdx.x & 63) < r1 ? 1u << u : 0u;
        ok[u] = 0;
#pragma unroll
        for (int j = 0; j < NPC; ++j) {
            v[u][j] = 0;
            if (j < g.n_pcols) {
                bool okj;
                cell_finish3(g.pc[j], r, t.c[u][j], v[u][j], okj);
                ok[u] |= (okj ? 1u : 0u) << j;
            }
        }
    }
    const uint32_t m = a.pred.n == 1 ? leaf_mask<U, NPC>(g.leaf0, v, ok) : pred_mask_tile<U, NPC>(a.pred, v, ok);
    return m & exists;
}

// Key and inputs of the rows in `pass`; a lane whose row did not pass reads
// the tile's first row instead, so lines holding no passing row stay unread.
template <int U, int NPC, int NIN>
FGPU_DEV void gbp_issue_data(const GbpScan<NPC, NIN> &g, int64_t base, uint32_t pass, GbpDataRaw<U, NIN> &t) {
    t.pass = pass;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t i = ((pass >> u) & 1u) ? (uint32_t)(u * 64 + (threadIdx.x & 63)) : 0u;
        cell_issue3_tile(g.key, base, i, t.key[u]);
#pragma unroll
        for (int j = 0; j < NIN; ++j) cell_issue3_tile(g.in[j], base, i, t.in[u][j]);
    }
}

template <int U, int NPC, int NIN>
FGPU_DEV void gbp_finish_data(const GbpScan<NPC, NIN> &g, int64_t base, const GbpDataRaw<U, NIN> &t,
                              GbpDataTile<U, NIN> &d) {
    d.pass = t.pass;
    const int64_t r = threadIdx.x & 31;  // validity bit of every passing row of the lane
#pragma unroll
    for (int u = 0; u < U; ++u) {
        bool ok;
        cell_finish3(g.key, r, t.key[u], d.r[u].key, ok);
        d.r[u].vw = (ok && d.r[u].key != EMPTY_KEY) ? GBP_HAS_KEY : 0u;
#pragma unroll
        for (int j = 0; j < NIN; ++j) {
            cell_finish3(g.in[j], r, t.in[u][j], d.r[u].in[j], ok);
            d.r[u].vw |= (ok ? 1u : 0u) << j;
        }
    }
}

// ─── specialised tile loads ──────────────────────────────────────────────────
// The common shape -- one comparison on a column without NULLs, an INT64 key
// and 64-bit inputs, none with NULLs -- needs none of the generic cell
// machinery: each cell is one load from a wave-uniform base plus a 32-bit
// lane offset (the scalar-base addressing form), with no per-dtype finish.
// FP = the predicate column's width in bytes (4: INT32, 8: INT64 / FLOAT64).
template <int U, int FP>
struct GbpPredFast {
    uint64_t x[U];  // INT32 values sign-extended, 64-bit values as bits
    bool dead;      // the zone maps rule the tile out: nothing was loaded
};

template <int U, int NIN>
struct GbpDataFast {
    uint64_t key[U];
    uint64_t in[U][NIN > 0 ? NIN : 1];
    uint32_t pass;
};

// Column loads of the specialised tiles are non-temporal: every column word is
// read once, and keeping the stream out of L2 leaves it to the cold-key
// records, whose partly written lines otherwise leave early.
template <class T>
FGPU_DEV T gbp_ld(const uint8_t *base, uint32_t off) {
    return __builtin_nontemporal_load((const T *)(base + off));
}

template <int U, int FP, int NPC, int NIN>
FGPU_DEV void gbp_issue_pred(const GbpScan<NPC, NIN> &g, int64_t base, int64_t r1, GbpPredFast<U, FP> &t) {
    t.dead = zone_dead(g.live, g.batch_rows, base, base + 64 * U < r1 ? base + 64 * U : r1);
    if (t.dead) return;
    const uint8_t *b = g.pc[0].base + base * FP;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t off = gbp_pred_idx(base, u, r1) * (uint32_t)FP;
        t.x[u] = FP == 4 ? (uint64_t)(int64_t)gbp_ld<int32_t>(b, off) : gbp_ld<uint64_t>(b, off);
    }
}

template <int U, int FP, int NPC, int NIN>
FGPU_DEV uint32_t gbp_finish_pred(const GbArgs &, const GbpScan<NPC, NIN> &g, int64_t base, int64_t r1,
                                  const GbpPredFast<U, FP> &t) {
    if (t.dead) return 0;
    int64_t v[U][1];
    uint32_t ok[U];
    uint32_t exists = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        exists |= base + u * 64 + (threadIdx.x & 63) < r1 ? 1u << u : 0u;
        v[u][0] = (int64_t)t.x[u];
        ok[u] = 1u;
    }
    return leaf_mask<U, 1>(g.leaf0, v, ok) & exists;
}

template <int U, int NPC, int NIN>
FGPU_DEV void gbp_issue_data(const GbpScan<NPC, NIN> &g, int64_t base, uint32_t pass, GbpDataFast<U, NIN> &t) {
    t.pass = pass;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t off = ((pass >> u) & 1u) ? (uint32_t)(u * 64 + (threadIdx.x & 63)) * 8u : 0u;
        t.key[u] = gbp_ld<uint64_t>(g.key.base + base * 8, off);
#pragma unroll
        for (int j = 0; j < NIN; ++j) t.in[u][j] = gbp_ld<uint64_t>(g.in[j].base + base * 8, off);
    }
}

// The validity word is rebuilt after compaction from the key (FP != 0).
template <int U, int NPC, int NIN>
FGPU_DEV void gbp_finish_data(const GbpScan<NPC, NIN> &, int64_t, const GbpDataFast<U, NIN> &t,
                              GbpDataTile<U, NIN> &d) {
    d.pass = t.pass;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        d.r[u].key = (int64_t)t.key[u];
#pragma unroll
        for (int j = 0; j < NIN; ++j) d.r[u].in[j] = (int64_t)t.in[u][j];
        d.r[u].vw = 0;
    }
}

template <int NIN>
FGPU_DEV Inputs gbp_inputs(const GbpRow<NIN> &r) {
    Inputs in{0, 0, 0, 0, 0, 0, r.vw & 0x3fu};
    if (NIN > 0) in.v0 = r.in[0];
    if (NIN > 1) in.v1 = r.in[NIN > 1 ? 1 : 0];
    if (NIN > 2) in.v2 = r.in[NIN > 2 ? 2 : 0];
    if (NIN > 3) in.v3 = r.in[NIN > 3 ? 3 : 0];
    if (NIN > 4) in.v4 = r.in[NIN > 4 ? 4 : 0];
    if (NIN > 5) in.v5 = r.in[NIN > 5 ? 5 : 0];
    return in;
}

struct GbpLds {
    int64_t *keys;
    uint64_t *state, *null_state;
    unsigned int *front, *back, *door;
};

// Aggregates one compacted row (lanes with active == false do nothing).
template <int NPC, int NIN>
FGPU_DEV void gbp_row(const GbpScan<NPC, NIN> &g, const GbpLds &L, bool active, const GbpRow<NIN> &row) {
    const Inputs in = gbp_inputs<NIN>(row);
    if (!active) return;
    const int64_t key = row.key;
    const uint64_t h = hash64((uint64_t)key);
    uint64_t *st = nullptr;  // the NULL group's or a cached key's state
    if (!(row.vw & GBP_HAS_KEY)) {
        st = L.null_state;
    } else if (g.H) {
        const int b = gbp_range((uint32_t)(h >> 16), g.NB);  // bits 16..47
        int64_t *bk = L.keys + b * GBP_WAYS;
        int64_t k[GBP_WAYS];
#pragma unroll
        for (int w = 0; w < GBP_WAYS; w += 2) {
            const gbp_ll2 kk = *(const gbp_ll2 *)(bk + w);
            k[w] = kk.x;
            k[w + 1] = kk.y;
        }
        int way = -1, empty = -1;
#pragma unroll
        for (int w = GBP_WAYS - 1; w >= 0; --w) {
            way = k[w] == key ? w : way;
            empty = k[w] == EMPTY_KEY ? w : empty;
        }
        if (way < 0 && empty >= 0) {
            const unsigned bit = 1u << (h & 31);
            const unsigned seen = atomicOr(L.door + ((h >> 5) & (GBP_DOOR_WORDS - 1)), bit) & bit;
            // Insert at the first empty way; a way lost to another key sends the
            // insert to the next one, so ways still fill strictly in order.
            for (int w = empty; seen && w < GBP_WAYS; ++w) {
                int64_t expect = EMPTY_KEY;
                if (__hip_atomic_compare_exchange_strong(bk + w, &expect, key, __ATOMIC_RELAXED,
                                                         __ATOMIC_RELAXED, LDS_SCOPE) ||
                    expect == key) {
                    way = w;
                    break;
                }
            }
        }
        if (way >= 0) st = L.state + (int64_t)(b * GBP_WAYS + way) * g.S;
    }
    if (st) {
        update_row<LDS_SCOPE>(st, g.ops, in);
        return;
    }
    const int p = gbp_part(h, g.log_p);
    const unsigned pos = atomicAdd(L.front + p, 1u);
    if (pos >= (unsigned)g.Cw) {
        atomicExch(g.flags, 1ull);
        return;
    }
    int64_t *rec = g.my + ((int64_t)p * g.Cw + pos) * g.RW;
    if (NIN > 0 && g.pack) {  // key and input 0 in one word (no NULLs: no validity word)
        rec[0] = (int64_t)(((uint64_t)key - (uint64_t)g.pk_kmin) |
                           (((uint64_t)row.in[0] - (uint64_t)g.pk_vmin) << g.pk_kbits));
#pragma unroll
        for (int j = 1; j < NIN; ++j) rec[j] = row.in[j];
        return;
    }
    rec[0] = key;
#pragma unroll
    for (int j = 0; j < NIN; ++j) rec[1 + j] = row.in[j];
    if (g.has_valid_word) rec[1 + NIN] = (int64_t)(row.vw & 0x3fu);
}

template <int U, int NPC, int NIN, int FP>
__global__ __launch_bounds__(1024) void k_gbp_scan(GbPartArgs A) {
    extern __shared__ uint64_t lds[];
    const GbArgs &a = A.g;
    GbpScan<NPC, NIN> g;
#pragma unroll
    for (int j = 0; j < NPC; ++j) g.pc[j] = col_ref(a.cols, j < a.pred.n_pcols ? a.pred.pcol[j] : 0);
    g.n_pcols = a.pred.n_pcols;
    g.leaf0 = a.pred.nodes[0];
    g.live = a.pred.live;
    g.batch_rows = a.pred.batch_rows;
    g.key = col_ref(a.cols, a.key_col);
#pragma unroll
    for (int j = 0; j < NIN; ++j) g.in[j] = col_ref(a.cols, a.in_col[j]);
    g.ops = op_list(a.ops, A.has_valid_word != 0);
    g.S = a.ops.S;
    g.H = A.H;
    g.NB = A.H / GBP_WAYS;
    g.log_p = A.log_p;
    g.Cw = A.Cw;
    g.RW = A.RW;
    g.has_valid_word = A.has_valid_word;
    g.pack = A.pack;
    g.pk_kbits = A.pk_kbits;
    g.pk_kmin = A.pk_kmin;
    g.pk_vmin = A.pk_vmin;
    g.flags = A.flags;
    const int S = g.S, H = g.H, P = 1 << g.log_p;
    GbpLds L;
    L.keys = (int64_t *)lds;
    L.state = lds + H;
    L.null_state = L.state + (int64_t)H * S;
    L.front = (unsigned int *)(L.null_state + S);
    L.back = L.front + P;
    L.door = L.back + P;
    for (int i = threadIdx.x; i < H; i += blockDim.x) L.keys[i] = EMPTY_KEY;
    fill_ident(L.state, (int64_t)H + 1, S, a.ops);  // H cache rows, then the NULL group's
    for (int i = threadIdx.x; i < 2 * P + GBP_DOOR_WORDS; i += blockDim.x) L.front[i] = 0;
    __syncthreads();
    const int64_t r0 = (int64_t)blockIdx.x * A.rows_per_wg;
    const int64_t r1 = r0 + A.rows_per_wg < a.n_rows ? r0 + A.rows_per_wg : a.n_rows;
    g.my = A.cells + (int64_t)blockIdx.x * P * g.Cw * g.RW;
    // wave, tile and chunk-fill counters are wave-uniform: readfirstlane lets
    // the compiler keep them in SGPRs and branch on them with scalar branches.
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int nw = __builtin_amdgcn_readfirstlane((int)(blockDim.x >> 6));
    constexpr int TR = 64 * U;  // rows per tile
    const int n_tiles = r1 > r0 ? (int)((r1 - r0 + TR - 1) / TR) : 0;

    GbpRow<NIN> cur, nxt;  // the chunk being filled and its overflow
    cur.key = nxt.key = 0;
    cur.vw = nxt.vw = 0;
#pragma unroll
    for (int j = 0; j < (NIN > 0 ? NIN : 1); ++j) cur.in[j] = nxt.in[j] = 0;
    int fill = 0;

    // Pipeline: while tile t is aggregated, the data loads of tile t + nw and
    // the predicate loads of tile t + 2 nw are in flight.
    int t = wave;
    std::conditional_t<FP != 0, GbpPredFast<U, FP>, GbpPredRaw<U, NPC>> pr;
    std::conditional_t<FP != 0, GbpDataFast<U, NIN>, GbpDataRaw<U, NIN>> draw;
    draw.pass = 0;
    if (t < n_tiles) {
        gbp_issue_pred(g, r0 + (int64_t)t * TR, r1, pr);
        const uint32_t m = gbp_finish_pred(a, g, r0 + (int64_t)t * TR, r1, pr);
        gbp_issue_data(g, r0 + (int64_t)t * TR, m, draw);
        if (t + nw < n_tiles) gbp_issue_pred(g, r0 + (int64_t)(t + nw) * TR, r1, pr);
    }
    // One pass per tile, then one more (have == false) that aggregates the last,
    // partial chunk; gbp_row has a single call site to keep the loop compact.
    for (;; t += nw) {
        const bool have = t < n_tiles;
        GbpDataTile<U, NIN> dt;
        dt.pass = 0;
        if (have) {
            gbp_finish_data(g, r0 + (int64_t)t * TR, draw, dt);
            if (t + nw < n_tiles) {
                const uint32_t m = gbp_finish_pred(a, g, r0 + (int64_t)(t + nw) * TR, r1, pr);
                gbp_issue_data(g, r0 + (int64_t)(t + nw) * TR, m, draw);
                if (t + 2 * nw < n_tiles) gbp_issue_pred(g, r0 + (int64_t)(t + 2 * nw) * TR, r1, pr);
            }
        }
        // compact tile t's passing rows into chunks and aggregate full ones
#pragma unroll 1
        for (int u = 0; u < U; ++u) {
            GbpRow<NIN> src = dt.r[0];
#pragma unroll
            for (int q = 1; q < U; ++q)
                if (u == q) src = dt.r[q];
            const bool pass = (dt.pass >> u) & 1u;
            const uint64_t bal = __ballot(pass);
            const int pc = __popcll(bal);
            if (pc) {
                const int rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                           __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                const int addr = ((fill + (pass ? rank : pc + (lane - rank))) & 63) << 2;
                const bool recv = ((lane - fill) & 63) < pc;
                const bool to_cur = recv && lane >= fill, to_nxt = recv && lane < fill;
                const int64_t k = perm64(addr, src.key);
                cur.key = to_cur ? k : cur.key;
                nxt.key = to_nxt ? k : nxt.key;
#pragma unroll
                for (int j = 0; j < NIN; ++j) {
                    const int64_t x = perm64(addr, src.in[j]);
                    cur.in[j] = to_cur ? x : cur.in[j];
                    nxt.in[j] = to_nxt ? x : nxt.in[j];
                }
                uint32_t vw;
                if constexpr (FP != 0) {
                    vw = (k != EMPTY_KEY ? GBP_HAS_KEY : 0u) | ((1u << NIN) - 1u);
                } else {
                    vw = (uint32_t)__builtin_amdgcn_ds_permute(addr, (int)src.vw);
                }
                cur.vw = to_cur ? vw : cur.vw;
                nxt.vw = to_nxt ? vw : nxt.vw;
                fill = __builtin_amdgcn_readfirstlane(fill + pc);
            }
            if (fill >= 64 || (!have && fill > 0)) {
                gbp_row<NPC, NIN>(g, L, lane < fill, cur);
                // a full chunk continues with its overflow; after the final
                // partial chunk the loop ends, so the copy is harmless there
                fill = __builtin_amdgcn_readfirstlane(fill >= 64 ? fill - 64 : 0);
                cur = nxt;
            }
        }
        if (!have) break;
    }
    __syncthreads();
    // flush the cache: state records fill each cell from its back end
    for (int i = threadIdx.x; i < H; i += blockDim.x) {
        const int64_t k = L.keys[i];
        if (k == EMPTY_KEY || L.state[(int64_t)i * S] == 0) continue;
        const int p = gbp_part(hash64((uint64_t)k), g.log_p);
        const unsigned b = atomicAdd(L.back + p, 1u);
        const int64_t slot = (int64_t)g.Cw - (int64_t)(b + 1) * A.SR;
        const int64_t used = L.front[p] < (unsigned)g.Cw ? L.front[p] : g.Cw;
        if (slot < used) {
            atomicExch(A.flags, 1ull);
            continue;
        }
        int64_t *dst = g.my + ((int64_t)p * g.Cw + slot) * g.RW;
        dst[0] = k;
        for (int w = 0; w < S; ++w) dst[1 + w] = (int64_t)L.state[(int64_t)i * S + w];
    }
    __syncthreads();
    for (int p = threadIdx.x; p < P; p += blockDim.x) {
        A.cell_cnt[((int64_t)blockIdx.x * P + p) * 2] = L.front[p];
        A.cell_cnt[((int64_t)blockIdx.x * P + p) * 2 + 1] = L.back[p];
    }
    for (int w = threadIdx.x; w < S; w += blockDim.x) A.nullbuf[(int64_t)blockIdx.x * S + w] = L.null_state[w];
}

// ─── pass 2 ─────────────────────────────────────────────────────────────────
// One workgroup per partition.  The cell counts of the partition are scanned
// into LDS prefix arrays first, so every record has a flat index and all
// record loads are independent: each thread keeps two batches of U records in
// flight (the next batch loads while the current one is folded).  The records
// aggregate into an LDS table that is emitted at the end.
constexpr int GBP_AGG_UNROLL = 4;

// Largest g in [0, G) with pre[g] <= i (pre[0] = 0 <= i < pre[G]).
FGPU_DEV int gbp_cell_of(const unsigned *pre, int G, unsigned i) {
    int lo = 0, hi = G;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Records i0, i0 + blockDim.x, ... (U of them) of a partition: the cell of
// each flat index, then its loads; have[u] = the index exists.
template <int NIN, int U>
FGPU_DEV void gbp_agg_fetch(const GbPartArgs &A, const unsigned *pre, int G, const int64_t *part_cells,
                            int64_t cell_stride, unsigned n_rows, unsigned i0, int64_t (&key)[U], Inputs (&in)[U],
                            bool (&have)[U]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const unsigned i = i0 + u * blockDim.x;
        have[u] = i < n_rows;
        key[u] = 0;
        in[u] = Inputs{0, 0, 0, 0, 0, 0, 0};
        if (have[u]) {
            const int g = gbp_cell_of(pre, G, i);
            const int64_t *rec = part_cells + g * cell_stride + (int64_t)(i - pre[g]) * A.RW;
            if (NIN > 0 && A.pack) {  // packed: key and input 0 in the first word
                const uint64_t w = (uint64_t)rec[0];
                key[u] = (int64_t)((uint64_t)A.pk_kmin + (w & ((1ull << A.pk_kbits) - 1)));
                in[u].v0 = (int64_t)((uint64_t)A.pk_vmin + (w >> A.pk_kbits));
                if (NIN > 1) in[u].v1 = rec[1];
                if (NIN > 2) in[u].v2 = rec[2];
                if (NIN > 3) in[u].v3 = rec[3];
                if (NIN > 4) in[u].v4 = rec[4];
                if (NIN > 5) in[u].v5 = rec[5];
                in[u].valid = 0xffffffffu;
                continue;
            }
            key[u] = rec[0];
            if (NIN > 0) in[u].v0 = rec[1];
            if (NIN > 1) in[u].v1 = rec[2];
            if (NIN > 2) in[u].v2 = rec[3];
            if (NIN > 3) in[u].v3 = rec[4];
            if (NIN > 4) in[u].v4 = rec[5];
            if (NIN > 5) in[u].v5 = rec[6];
            in[u].valid = A.has_valid_word ? (uint32_t)rec[1 + NIN] : 0xffffffffu;
        }
    }
}

template <int NIN>
__global__ __launch_bounds__(1024) void k_gbp_agg(GbPartArgs A) {
    extern __shared__ uint64_t lds[];
    const GbArgs &a = A.g;
    const int S = a.ops.S, T2 = A.T2, P = 1 << A.log_p, G = A.G;
    int64_t *tkeys = (int64_t *)lds;
    uint64_t *tstate = lds + T2;
    unsigned *pre = (unsigned *)(tstate + (int64_t)T2 * S);  // G + 1: row records
    unsigned *spre = pre + G + 1;                             // G + 1: state records
    unsigned *list = spre + G + 1;                            // T2: the occupied slots, in output order
    const int p = blockIdx.x;
    const int lane = threadIdx.x & 63;
    // pass 1 overflowed a cell: the host resizes and runs both passes again
    if (__hip_atomic_load(A.flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    // the cell counts' loads are issued first, so they land during the table init
    unsigned nf = 0, ns = 0;
    if ((int)threadIdx.x < G) {
        const int64_t c = (int64_t)threadIdx.x * P + p;
        nf = A.cell_cnt[2 * c];
        ns = A.cell_cnt[2 * c + 1];
    }
    for (int i = threadIdx.x; i < T2; i += blockDim.x) tkeys[i] = EMPTY_KEY;
    fill_ident(tstate, T2, S, a.ops);
    for (int g = threadIdx.x; g < G; g += blockDim.x) {
        if (g != (int)threadIdx.x) {  // G > blockDim.x
            const int64_t c = (int64_t)g * P + p;
            nf = A.cell_cnt[2 * c];
            ns = A.cell_cnt[2 * c + 1];
        }
        pre[g] = nf < (unsigned)A.Cw ? nf : (unsigned)A.Cw;
        spre[g] = ns;
    }
    __syncthreads();
    if (threadIdx.x < 64) {  // exclusive scans of both count arrays by one wave
        const int per = (G + 63) / 64;
        unsigned s1 = 0, s2 = 0;
        for (int k = 0; k < per; ++k) {
            const int g = lane * per + k;
            if (g < G) {
                s1 += pre[g];
                s2 += spre[g];
            }
        }
        unsigned i1 = s1, i2 = s2;
        for (int d = 1; d < 64; d <<= 1) {
            const unsigned t1 = __shfl_up(i1, d, 64), t2 = __shfl_up(i2, d, 64);
            if (lane >= d) {
                i1 += t1;
                i2 += t2;
            }
        }
        unsigned e1 = i1 - s1, e2 = i2 - s2;
        for (int k = 0; k < per; ++k) {
            const int g = lane * per + k;
            if (g < G) {
                const unsigned c1 = pre[g], c2 = spre[g];
                pre[g] = e1;
                spre[g] = e2;
                e1 += c1;
                e2 += c2;
            }
        }
        if (lane == 63) {
            pre[G] = i1;
            spre[G] = i2;
        }
    }
    __syncthreads();
    const unsigned n_rows = pre[G], n_states = spre[G];
    const int64_t *part_cells = A.cells + (int64_t)p * A.Cw * A.RW;  // cell (g, p) = part_cells + g * P * Cw * RW
    const int64_t cell_stride = (int64_t)P * A.Cw * A.RW;
    {
    // Two batches in flight: the cell searches and record loads of batch
    // i0 + step are issued before batch i0 is folded into the table.
    constexpr int U = NIN >= 5 ? 2 : GBP_AGG_UNROLL;  // two batches of U records within 128 VGPRs
    const unsigned step = blockDim.x * U;
    int64_t key[U], nkey[U];
    Inputs in[U], nin[U];
    bool have[U], nhave[U];
    gbp_agg_fetch<NIN, U>(A, pre, G, part_cells, cell_stride, n_rows, threadIdx.x, key, in, have);
    for (unsigned i0 = threadIdx.x; i0 < n_rows; i0 += step) {
        gbp_agg_fetch<NIN, U>(A, pre, G, part_cells, cell_stride, n_rows, i0 + step, nkey, nin, nhave);
        uint64_t *st[U];
        bool ok[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int li = have[u] ? gbp_slot(tkeys, T2, T2, key[u], hash64((uint64_t)key[u])) : 0;
            if (li < 0) atomicExch(A.flags + 1, 1ull);
            ok[u] = have[u] && li >= 0;
            st[u] = tstate + (int64_t)(ok[u] ? li : 0) * S;
        }
        update_rows<LDS_SCOPE, U>(st, ok, a.ops, in);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            key[u] = nkey[u];
            in[u] = nin[u];
            have[u] = nhave[u];
        }
    }
    }
    for (unsigned i = threadIdx.x; i < n_states; i += blockDim.x) {
        const int g = gbp_cell_of(spre, G, i);
        const int64_t slot = (int64_t)A.Cw - (int64_t)(i - spre[g] + 1) * A.SR;
        const int64_t nf = (int64_t)(pre[g + 1] - pre[g]);
        if (slot < nf) continue;  // not written (cell full; flagged in pass 1)
        const int64_t *rec = part_cells + g * cell_stride + slot * A.RW;
        const int64_t key = rec[0];
        const int li = gbp_slot(tkeys, T2, T2, key, hash64((uint64_t)key));
        if (li < 0) {
            atomicExch(A.flags + 1, 1ull);
            continue;
        }
        merge_state<LDS_SCOPE>(tstate + (int64_t)li * S, (const uint64_t *)(rec + 1), a.ops);
    }
    __syncthreads();
    // emit: the occupied slots are listed in output order first, so that the
    // emitting lanes are all busy and their stores contiguous; one output
    // reservation for the whole partition
    unsigned n_have = 0;
    for (int i = threadIdx.x; i < T2; i += blockDim.x)
        n_have += tkeys[i] != EMPTY_KEY && tstate[(int64_t)i * S] > 0 ? 1u : 0u;
    unsigned total;
    unsigned pos = block_excl_scan(n_have, total);
    for (int i = threadIdx.x; i < T2; i += blockDim.x)
        if (tkeys[i] != EMPTY_KEY && tstate[(int64_t)i * S] > 0) list[pos++] = (unsigned)i;
    const int64_t base = (int64_t)block_reserve(A.out_count, total);  // its barriers publish the list
    for (unsigned j = threadIdx.x; j < total; j += blockDim.x) {
        const int64_t o = base + j;
        if (o >= A.out_cap) {
            atomicExch(A.flags + 2, 1ull);
            break;
        }
        const unsigned i = list[j];
        gst(A.out_key + o, tkeys[i]);
        gst(A.out_key_valid + o, (uint8_t)1);
        emit_row(A.outs, tstate + (int64_t)i * S, o, A.out_val, A.out_valid);
    }
}

// The NULL group: merge every workgroup's partial state, emit one row.
__global__ void k_gbp_null(GbPartArgs A) {
    __shared__ uint64_t row[64];
    if (__hip_atomic_load(A.flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    const DOps &ops = A.g.ops;
    for (int w = threadIdx.x; w < ops.S; w += blockDim.x) row[w] = ops.ident[w];
    __syncthreads();
    for (int g = threadIdx.x; g < A.G; g += blockDim.x)
        merge_state<LDS_SCOPE>(row, A.nullbuf + (int64_t)g * ops.S, ops);
    __syncthreads();
    if (threadIdx.x == 0 && row[0] > 0) {
        const int64_t o = (int64_t)atomicAdd(A.out_count, 1ull);
        if (o >= A.out_cap) {
            atomicExch(A.flags + 2, 1ull);
            return;
        }
        A.out_key[o] = 0;
        A.out_key_valid[o] = 0;
        emit_row(A.outs, row, o, A.out_val, A.out_valid);
    }
}

template <int NPC, int NIN, int FP>
static hipError_t gbp_scan_go(hipStream_t s, const GbPartArgs &a, int block, size_t lds) {
    constexpr int U = 2;  // 128-row tiles: two in flight per wave within 128 VGPRs
    const void *fn = (const void *)k_gbp_scan<U, NPC, NIN, FP>;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_gbp_scan<U, NPC, NIN, FP>), dim3(a.G), dim3(block), lds, s, a);
    return hipGetLastError();
}

template <int NPC, int FP>
static hipError_t gbp_scan_nin(hipStream_t s, const GbPartArgs &a, int block, size_t lds) {
    switch (a.g.n_in) {
        case 0: return gbp_scan_go<NPC, 0, FP>(s, a, block, lds);
        case 1: return gbp_scan_go<NPC, 1, FP>(s, a, block, lds);
        case 2: return gbp_scan_go<NPC, 2, FP>(s, a, block, lds);
        case 3: return gbp_scan_go<NPC, 3, FP>(s, a, block, lds);
        case 4: return gbp_scan_go<NPC, 4, FP>(s, a, block, lds);
        case 5: return gbp_scan_go<NPC, 5, FP>(s, a, block, lds);
        default: return gbp_scan_go<NPC, 6, FP>(s, a, block, lds);
    }
}

// Width in bytes of the predicate column when the scan can take the
// specialised tile loads (see GbpPredFast), else 0.
int gbp_fast_width(const GbPartArgs &a) {
    const GbArgs &g = a.g;
    if (a.has_valid_word || g.pred.n != 1 || g.pred.n_pcols != 1) return 0;
    const DPredNode &nd = g.pred.nodes[0];
    if (nd.kind != FGPU_P_CMP || nd.pc != 0) return 0;
    if (nd.cls != CMP_F64 && nd.cls != CMP_I64 && nd.cls != CMP_I64_AS_F64 && nd.cls != CMP_I32) return 0;
    const int pcol = g.pred.pcol[0];
    if (g.cols.valid[pcol] || g.cols.valid[g.key_col] || g.cols.dtype[g.key_col] != FGPU_INT64) return 0;
    for (int j = 0; j < g.n_in; ++j) {
        const int c = g.in_col[j];
        if (g.cols.valid[c] || (g.cols.dtype[c] != FGPU_INT64 && g.cols.dtype[c] != FGPU_FLOAT64)) return 0;
    }
    const int dt = g.cols.dtype[pcol];
    if (dt == FGPU_INT32 && (nd.cls == CMP_I32 || nd.cls == CMP_I64_AS_F64)) return 4;
    if ((dt == FGPU_INT64 || dt == FGPU_FLOAT64) && nd.cls != CMP_I32) return 8;
    return 0;
}

hipError_t launch_gbp_scan(hipStream_t s, const GbPartArgs &a, int block, size_t lds) {
    if (a.H % GBP_WAYS) return hipErrorInvalidValue;
    switch (gbp_fast_width(a)) {
        case 4: return gbp_scan_nin<1, 4>(s, a, block, lds);
        case 8: return gbp_scan_nin<1, 8>(s, a, block, lds);
        default: break;
    }
    if (a.g.pred.n_pcols <= 1) return gbp_scan_nin<1, 0>(s, a, block, lds);
    if (a.g.pred.n_pcols <= 2) return gbp_scan_nin<2, 0>(s, a, block, lds);
    return gbp_scan_nin<MAX_PCOLS, 0>(s, a, block, lds);
}

template <int NIN>
static hipError_t gbp_agg_go(hipStream_t s, const GbPartArgs &a, int block, size_t lds) {
    hipError_t e = hipFuncSetAttribute((const void *)k_gbp_agg<NIN>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_gbp_agg<NIN>), dim3(1 << a.log_p), dim3(block), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_gbp_agg(hipStream_t s, const GbPartArgs &a, int block, size_t lds) {
    if (a.T2 > 32 * block) return hipErrorInvalidValue;  // the emit mask holds 32 slots per thread
    hipError_t e;
    switch (a.g.n_in) {
        case 0: e = gbp_agg_go<0>(s, a, block, lds); break;
        case 1: e = gbp_agg_go<1>(s, a, block, lds); break;
        case 2: e = gbp_agg_go<2>(s, a, block, lds); break;
        case 3: e = gbp_agg_go<3>(s, a, block, lds); break;
        case 4: e = gbp_agg_go<4>(s, a, block, lds); break;
        case 5: e = gbp_agg_go<5>(s, a, block, lds); break;
        default: e = gbp_agg_go<6>(s, a, block, lds); break;
    }
    if (e != hipSuccess) return e;
    if (!a.null_group) return hipGetLastError();  // no row can be in the NULL group
    hipLaunchKernelGGL(k_gbp_null, dim3(1), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace fgpu
