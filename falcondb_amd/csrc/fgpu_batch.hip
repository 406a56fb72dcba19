// fgpu_batch.hip — many GROUP BY queries in two persistent kernels (see
// BatchQuery).  Each query is the same operator as fgpu_groupby
// (GroupByExec::execute, src/query/aggregate.rs:535-1420) over its own table;
// the batch only changes how the work is scheduled:
//
//   pass 1 (k_batch_scan): workgroups pull (query, row range) units from one
//     counter.  A query whose whole key range fits an LDS table aggregates
//     there and writes one state block per unit (BQ_DENSE); a wider query
//     writes (key, inputs) records into key-range partitions through LDS
//     cursors, so no global atomic is taken per row (BQ_PART).
//   pass 2 (k_batch_merge): workgroups pull (query, slot range) units, fold
//     the blocks or the partition's records into a dense LDS table (slot =
//     key - key_min: no hashing) and emit the groups with one output
//     reservation per unit.
//
// Both kernels run a grid of resident workgroups that exit when the unit
// counter passes the last unit, and the host captures the resets plus both
// launches in one hipGraph that it replays per run.
#include "fgpu_state.h"

namespace fgpu {

constexpr int BATCH_UNROLL = 4;  // 64-row steps per wave tile (queries of at most two inputs)

FGPU_DEV int64_t rfl64(unsigned long long v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Next unit index for the whole workgroup (thread 0 takes it from the counter).
FGPU_DEV int64_t batch_next(unsigned long long *counter, unsigned long long *slot) {
    if (threadIdx.x == 0) *slot = atomicAdd(counter, 1ull);
    __syncthreads();
    const int64_t u = rfl64(*slot);
    __syncthreads();
    return u;
}

// A BQ_PART record: word 0 is the key slot, or (packed) the slot with input 0
// above it; then the remaining inputs and the validity word.  A two-word
// record goes out as one 16-byte store (records are 16-byte aligned then).
template <int NINC>
FGPU_DEV void batch_put(int64_t *rec, const BatchQuery &Q, int64_t slot, const int64_t (&v)[NINC],
                        uint32_t valid) {
    const int ni = Q.g.n_in;
    int64_t w[NINC + 2];  // at most the slot, NINC inputs and the validity word
    w[0] = Q.pack ? (int64_t)((uint64_t)slot | ((uint64_t)(v[0] - Q.pk_vmin) << Q.pk_kbits)) : slot;
    const int j0 = Q.pack ? 1 : 0;  // first input after word 0
#pragma unroll
    for (int k = 1; k <= NINC + 1; ++k) {
        const int j = j0 + k - 1;  // input j, or the validity word after the last input
        int64_t x = (int64_t)valid;
#pragma unroll
        for (int q = 0; q < NINC; ++q) x = (j == q && q < ni) ? v[q] : x;
        w[k] = x;
    }
    if (Q.RW == 2) {
        gst((ll2 *)rec, ll2{w[0], w[1]});
        return;
    }
#pragma unroll
    for (int k = 0; k <= NINC + 1; ++k)
        if (k < Q.RW) gst(rec + k, w[k]);
}

constexpr int FAST_UNROLL = 4;  // rows per lane per tile of the specialised pass-1 tile

// A cell of a NULL-free column of 4 (INT32, sign-extended) or 8 bytes.  The
// load is non-temporal: every column word is read once, and keeping the
// stream out of L2 leaves L2 to the partly written record lines.
FGPU_DEV int64_t fast_cell(const void *base, int w, int64_t r) {
    typedef const __attribute__((address_space(1))) int32_t *g32;
    typedef const __attribute__((address_space(1))) int64_t *g64;
    return w == 4 ? (int64_t)__builtin_nontemporal_load((g32)base + r) : __builtin_nontemporal_load((g64)base + r);
}

// Pass-1 rows of a unit of a query with BatchQuery::fast set: each row is a
// key, one or two inputs and at most one predicate cell, all NULL-free, so a
// tile is FAST_UNROLL loads per column issued together, one comparison and
// one record of one or two words per passing row.
FGPU_DEV void batch_fast_rows(const BatchQuery &Q, const BatchUnit &U, int64_t *my_cells, unsigned *cursor,
                              unsigned long long *flags, int wave, int nw) {
    const GbArgs &a = Q.g;
    const int lane = threadIdx.x & 63;
    const void *kb = a.cols.v[a.key_col];
    const void *ib0 = a.cols.v[a.in_col[0]];
    const void *ib1 = a.cols.v[a.in_col[a.n_in > 1 ? 1 : 0]];
    const void *pb = a.pred.n ? a.cols.v[a.pred.pcol[0]] : kb;
    const DPredNode nd = a.pred.nodes[0];
    const int kw = Q.fast_kw, pw = Q.fast_pw, iw0 = Q.fast_iw0, iw1 = Q.fast_iw1;
    const bool two = a.n_in > 1;
    const int64_t key_min = Q.key_min, vmin = Q.pk_vmin;
    const int shift = Q.shift, kbits = Q.pk_kbits;
    const bool pack = Q.pack != 0;
    const unsigned Cw = (unsigned)Q.Cw;
    const int RW = Q.RW;  // 1 (packed, one input) or 2
    for (int64_t base = U.row0 + (int64_t)wave * 64 * FAST_UNROLL; base < U.row1;
         base += (int64_t)nw * 64 * FAST_UNROLL) {
        int64_t k[FAST_UNROLL], x0[FAST_UNROLL], x1[FAST_UNROLL], pv[FAST_UNROLL];
        // every cell of the tile issued before any is examined (rows past the
        // unit re-read its last row)
#pragma unroll
        for (int u = 0; u < FAST_UNROLL; ++u) {
            const int64_t r0 = base + u * 64 + lane;
            const int64_t r = r0 < U.row1 ? r0 : U.row1 - 1;
            k[u] = fast_cell(kb, kw, r);
            x0[u] = fast_cell(ib0, iw0, r);
            x1[u] = two ? fast_cell(ib1, iw1, r) : 0;
            pv[u] = pw ? fast_cell(pb, pw, r) : 0;
        }
        unsigned pos[FAST_UNROLL];
        int part[FAST_UNROLL];
        bool pass[FAST_UNROLL];
#pragma unroll
        for (int u = 0; u < FAST_UNROLL; ++u) {
            pass[u] = base + u * 64 + lane < U.row1 && (pw == 0 || leaf_reg(nd, pv[u], true));
            part[u] = (int)((k[u] - key_min) >> shift);
            pos[u] = pass[u] ? atomicAdd(cursor + part[u], 1u) : 0u;
        }
#pragma unroll
        for (int u = 0; u < FAST_UNROLL; ++u) {
            if (!pass[u]) continue;
            if (pos[u] >= Cw) {
                atomicExch(flags, 1ull);
                continue;
            }
            const int64_t slot = k[u] - key_min;
            // packed: slot | input 0 << kbits, then input 1 if any; else slot, input 0
            const int64_t w0 = pack ? (int64_t)((uint64_t)slot | ((uint64_t)(x0[u] - vmin) << kbits)) : slot;
            int64_t *rec = my_cells + ((int64_t)part[u] * Cw + pos[u]) * RW;
            if (RW == 2) gst((ll2 *)rec, ll2{w0, pack ? x1[u] : x0[u]});
            else gst(rec, w0);
        }
    }
}

// Pass 1.  NINC bounds the aggregate inputs of the batch's queries (2 or
// MAX_IN: register arrays sized for it), NPC the predicate columns.
template <int NPC, int NINC>
__global__ __launch_bounds__(1024) void k_batch_scan(const BatchQuery *__restrict__ qs,
                                                     const BatchUnit *__restrict__ units, BatchArgs A) {
    extern __shared__ uint64_t lds[];
    __shared__ unsigned long long unit_slot;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int nw = __builtin_amdgcn_readfirstlane((int)(blockDim.x >> 6));
    constexpr int UR = NINC <= 2 ? BATCH_UNROLL : 2;  // rows per lane per tile (fits 128 VGPRs)
    for (;;) {
        const int64_t ui = batch_next(A.work, &unit_slot);
        if (ui >= A.n_units) return;
        const BatchUnit U = units[ui];
        const BatchQuery &Q = qs[U.q];
        const GbArgs &a = Q.g;
        const int S = Q.S;
        const int ni = a.n_in;
        const bool dense = Q.mode == BQ_DENSE;
        const OpList ops = op_list(a.ops, Q.has_valid_word != 0);
        uint64_t *null_st = lds;
        uint64_t *table = lds + S;                          // BQ_DENSE: range * S words
        unsigned int *cursor = (unsigned int *)(lds + S);   // BQ_PART: P cursors
        for (int i = threadIdx.x; i < S; i += blockDim.x) null_st[i] = a.ops.ident[i];
        if (dense) {
            fill_ident(table, Q.range, S, a.ops);
        } else {
            for (int i = threadIdx.x; i < Q.P; i += blockDim.x) cursor[i] = 0;
        }
        __syncthreads();
        int64_t *my_cells = dense ? nullptr : Q.cells + (int64_t)U.local * Q.P * Q.Cw * Q.RW;
        if (Q.fast) {
            batch_fast_rows(Q, U, my_cells, cursor, A.flags, wave, nw);
        } else {
        const ColRef kc = col_ref(a.cols, a.key_col);
        ColRef ic[NINC];
#pragma unroll
        for (int j = 0; j < NINC; ++j) ic[j] = col_ref(a.cols, j < ni ? a.in_col[j] : 0);
        for (int64_t base = U.row0 + (int64_t)wave * 64 * UR; base < U.row1;
             base += (int64_t)nw * 64 * UR) {
            int64_t rows[UR];
#pragma unroll
            for (int u = 0; u < UR; ++u) rows[u] = base + u * 64 + lane;
            // key and inputs of every row of the tile issued first (rows past
            // the unit re-read its last row), then the predicate columns: one
            // round trip for all of them.  At the selectivities these scans
            // see, a row that fails still shares its cache lines with rows
            // that pass.
            RawCell3 kr[UR], vr[UR][NINC];
#pragma unroll
            for (int u = 0; u < UR; ++u) {
                const int64_t r = rows[u] < U.row1 ? rows[u] : U.row1 - 1;
                cell_issue3(kc, r, kr[u]);
#pragma unroll
                for (int j = 0; j < NINC; ++j)
                    if (j < ni) cell_issue3(ic[j], r, vr[u][j]);
            }
            bool pass[UR];
            pred_tile<UR, NPC>(a.pred, a.cols, rows, U.row1, pass);
            int64_t key[UR];
            bool has_key[UR];
            int64_t v[UR][NINC];
            uint32_t valid[UR];
#pragma unroll
            for (int u = 0; u < UR; ++u) {
                bool ok;
                cell_finish3(kc, rows[u], kr[u], key[u], ok);
                has_key[u] = ok && key[u] != EMPTY_KEY;  // NULL and the sentinel form the NULL group
                valid[u] = 0;
#pragma unroll
                for (int j = 0; j < NINC; ++j) {
                    v[u][j] = 0;
                    if (j < ni) {
                        bool okj;
                        cell_finish3(ic[j], rows[u], vr[u][j], v[u][j], okj);
                        valid[u] |= okj ? 1u << j : 0u;
                    }
                }
            }
            if (!dense) {
                // BQ_PART: every keyed row's cell position first (the cursor
                // atomics in flight together), then its record
                unsigned pos[UR];
                int part[UR];
#pragma unroll
                for (int u = 0; u < UR; ++u) {
                    // slot in [0, range): the range comes from the column's min / max
                    part[u] = (int)((key[u] - Q.key_min) >> Q.shift);
                    pos[u] = pass[u] && has_key[u] ? atomicAdd(cursor + part[u], 1u) : 0u;
                }
#pragma unroll
                for (int u = 0; u < UR; ++u) {
                    if (!pass[u] || !has_key[u]) continue;
                    if (pos[u] >= (unsigned)Q.Cw) {
                        atomicExch(A.flags, 1ull);
                        continue;
                    }
                    batch_put<NINC>(my_cells + ((int64_t)part[u] * Q.Cw + pos[u]) * Q.RW, Q, key[u] - Q.key_min,
                                    v[u], valid[u]);
                }
            }
            // rows whose state is in LDS: the NULL group, and every row of a
            // BQ_DENSE query.  One row per iteration, picked with selects, so
            // that update_row has a single call site (the loop stays rolled).
#pragma unroll 1
            for (int u = 0; u < UR; ++u) {
                bool ps = pass[0], hk = has_key[0];
                int64_t k = key[0];
                int64_t vu[NINC];
                uint32_t vl = valid[0];
#pragma unroll
                for (int j = 0; j < NINC; ++j) vu[j] = v[0][j];
#pragma unroll
                for (int q = 1; q < UR; ++q)
                    if (u == q) {
                        ps = pass[q];
                        hk = has_key[q];
                        k = key[q];
                        vl = valid[q];
#pragma unroll
                        for (int j = 0; j < NINC; ++j) vu[j] = v[q][j];
                    }
                if (!ps || (hk && !dense)) continue;
                int64_t w6[MAX_IN];
#pragma unroll
                for (int j = 0; j < MAX_IN; ++j) w6[j] = j < NINC ? vu[j < NINC ? j : 0] : 0;
                const Inputs iu{w6[0], w6[1], w6[2], w6[3], w6[4], w6[5], vl};
                update_row<LDS_SCOPE>(!hk ? null_st : table + (k - Q.key_min) * S, ops, iu);
            }
        }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < S; i += blockDim.x) Q.null_blocks[(int64_t)U.local * S + i] = null_st[i];
        if (dense) {
            uint64_t *blk = Q.blocks + (int64_t)U.local * Q.range * S;
            for (int64_t i = threadIdx.x; i < Q.range * S; i += blockDim.x) blk[i] = table[i];
        } else {
            for (int i = threadIdx.x; i < Q.P; i += blockDim.x) Q.cell_cnt[(int64_t)i * Q.n_units + U.local] = cursor[i];
        }
        __syncthreads();  // the next unit re-initialises the LDS
    }
}

constexpr int MERGE_INFLIGHT = 4;  // records per lane whose loads are in flight together

// The unpacked fields of a BQ_PART record's words (NW of them; words past
// the record's RW are zero): input j of the query, the slot, the validity.
struct RecFmt {
    int pack, kbits;
    int64_t vmin;
    int valid_word;  // index of the validity word, or -1 (inputs all valid)
};
template <int NW>
FGPU_DEV int64_t rec_word(const int64_t (&w)[NW], int k) {
    int64_t x = w[0];
#pragma unroll
    for (int q = 1; q < NW; ++q) x = k == q ? w[q] : x;
    return x;
}
template <int NW>
FGPU_DEV int64_t rec_input(const int64_t (&w)[NW], int j, const RecFmt &f) {
    if (f.pack && j == 0) return (int64_t)((uint64_t)w[0] >> f.kbits) + f.vmin;
    return rec_word<NW>(w, f.pack ? j : j + 1);
}
template <int NW>
FGPU_DEV int64_t rec_slot(const int64_t (&w)[NW], const RecFmt &f) {
    return f.pack ? (int64_t)((uint64_t)w[0] & ((uint64_t(1) << f.kbits) - 1)) : w[0];
}

// update_row for R records at once, op-major: each op's packed descriptor is
// read once (from LDS, made uniform) for all R records, and their atomics
// issue back to back.  Same effect as update_row on each record (every
// update is an integer add, min, max or or: order-free).
template <int R, int NW>
FGPU_DEV void batch_update(uint64_t *table, int S, const uint32_t *pk_s, int n_ops, const RecFmt &f,
                           const int64_t (&w)[R][NW], const int64_t (&sl)[R], const bool (&ok)[R]) {
    uint64_t *st[R];
    uint32_t valid[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        st[r] = table + sl[r] * S;
        valid[r] = f.valid_word >= 0 ? (uint32_t)rec_word<NW>(w[r], f.valid_word) : 0xffffffffu;
        if (ok[r]) a_add<LDS_SCOPE>(st[r], 1ull);
    }
    for (int i = 0; i < n_ops; ++i) {
        const uint32_t d = __builtin_amdgcn_readfirstlane(pk_s[i]);
        const int col = opk_col(d), word = opk_word(d);
        bool on[R];
        int64_t v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            on[r] = ok[r] && ((valid[r] >> col) & 1);
            v[r] = rec_input<NW>(w[r], col, f);
        }
        switch (opk_kind(d)) {
            case OP_CNT:
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (on[r]) a_add<LDS_SCOPE>(st[r] + word, 1ull);
                break;
            case OP_ISUM:
                if (opk_W(d) == 2) {
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        if (on[r]) {
                            uint64_t lo;
                            int64_t hi;
                            isum_digits(v[r], lo, hi);
                            a_add<LDS_SCOPE>(st[r] + word, lo);
                            a_add<LDS_SCOPE>(st[r] + word + 1, (uint64_t)hi);
                        }
                } else {
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        if (on[r]) a_add<LDS_SCOPE>(st[r] + word, (uint64_t)v[r]);
                }
                break;
            case OP_IMIN:
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (on[r]) a_smin<LDS_SCOPE>(st[r] + word, v[r]);
                break;
            case OP_IMAX:
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (on[r]) a_smax<LDS_SCOPE>(st[r] + word, v[r]);
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
                        a_or<LDS_SCOPE>(st[r] + word + W, fl);
                        continue;
                    }
                    int w0;
                    int64_t d0, d1, d2;
                    fx_digits3(bits, L, w0, d0, d1, d2);
                    a_add<LDS_SCOPE>(st[r] + word + w0, (uint64_t)d0);
                    a_add<LDS_SCOPE>(st[r] + word + w0 + 1, (uint64_t)d1);
                    a_add<LDS_SCOPE>(st[r] + word + w0 + 2, (uint64_t)d2);
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
                    if (is_min) a_umin<LDS_SCOPE>(st[r] + word, e);
                    else a_umax<LDS_SCOPE>(st[r] + word, e);
                }
                break;
            }
        }
    }
}

// One round of a merge unit's records: R per lane, NW words each.  The
// cells are found first, then every load is issued, and only then does the
// first update wait for its data.
// pre_lane: this lane's entry of the cell prefix (pre[lane] for lane <= G,
// else all ones), used when G < 64.
template <int R, int NW>
FGPU_DEV void merge_round(uint64_t *table, int S, const uint32_t *pk_s, int n_ops, const RecFmt &f,
                          const unsigned *pre, unsigned pre_lane, int G, unsigned n_rec, unsigned i0, unsigned step,
                          const int64_t *part_cells, int64_t cell_stride, int RW, int64_t slot0) {
    const int lane = threadIdx.x & 63;
    const int64_t *addr[R];
    bool ok[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        // the cell of the wave's first record (one search per wave: the bound
        // is uniform), then each lane steps forward over the few cells its
        // 64-record window spans
        const unsigned ib = __builtin_amdgcn_readfirstlane(i0 + r * step);
        const unsigned i = ib + lane;
        ok[r] = i < n_rec;
        addr[r] = part_cells;  // a lane without a record reads the first one
        if (ib >= n_rec) continue;
        int lo = 0, hi = G;
        if (G < 64) {
            // the entries <= ib, counted across the wave's lanes (pre[0] = 0
            // counts, pre[G] = n_rec > ib does not)
            lo = __popcll(__ballot(pre_lane <= ib)) - 1;
            hi = lo + 1;
        }
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (pre[mid] <= ib) lo = mid;
            else hi = mid;
        }
        if (!ok[r]) continue;
        while (pre[lo + 1] <= i) ++lo;  // pre[G] = n_rec > i ends the walk
        addr[r] = part_cells + lo * cell_stride + (int64_t)(i - pre[lo]) * RW;
    }
    int64_t w[R][NW];
    if (NW == 2) {  // one 16-byte load per record
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const ll2 x = gld((const ll2 *)addr[r]);
            w[r][0] = x.x;
            w[r][NW - 1] = x.y;
        }
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int k = 0; k < NW; ++k) w[r][k] = k < RW ? gld(addr[r] + k) : 0;
    }
    int64_t sl[R];
#pragma unroll
    for (int r = 0; r < R; ++r) sl[r] = ok[r] ? rec_slot<NW>(w[r], f) - slot0 : 0;  // 0: never dereferenced
    batch_update<R, NW>(table, S, pk_s, n_ops, f, w, sl, ok);
}

__global__ __launch_bounds__(1024) void k_batch_merge(const BatchQuery *__restrict__ qs,
                                                      const MergeUnit *__restrict__ units2, BatchArgs A) {
    extern __shared__ uint64_t lds[];
    __shared__ unsigned long long unit_slot;
    __shared__ uint64_t null_row[64];
    __shared__ uint64_t ident_row[64];
    __shared__ uint32_t pk_s[MAX_OPS];
    __shared__ unsigned long long out_base;
    const int lane = threadIdx.x & 63;
    if (threadIdx.x == 0) unit_slot = atomicAdd(A.work + 1, 1ull);
    __syncthreads();
    for (;;) {
        const int64_t mi = rfl64(unit_slot);
        if (mi >= A.n_units2) return;
        // the next unit's index is taken now: the atomic's round trip overlaps
        // this unit, and thread 0 publishes it before the closing barrier
        unsigned long long next = 0;
        if (threadIdx.x == 0) next = atomicAdd(A.work + 1, 1ull);
        const MergeUnit M = units2[mi];
        const BatchQuery &Q = qs[M.q];
        const GbArgs &a = Q.g;
        const int S = Q.S;
        const int n = (int)(M.k1 - M.k0);
        const int G = (int)Q.n_units;
        const int n_ops = a.ops.n;
        uint64_t *table = lds;                                       // n * S words
        unsigned int *list = (unsigned int *)(lds + (int64_t)n * S);  // n: the occupied slots, in output order
        unsigned int *pre = list + n;                                  // BQ_PART: G + 1 prefix
        if (threadIdx.x < S) ident_row[threadIdx.x] = a.ops.ident[threadIdx.x];
        if (threadIdx.x < n_ops) pk_s[threadIdx.x] = a.ops.pk[threadIdx.x];
        if (M.part >= 0)
            for (int g = threadIdx.x; g < G; g += blockDim.x) {
                const unsigned c = gld(Q.cell_cnt + (int64_t)M.part * G + g);  // partition-major: contiguous
                pre[g] = c < (unsigned)Q.Cw ? c : (unsigned)Q.Cw;
            }
        if (M.null_group && threadIdx.x < S) null_row[threadIdx.x] = a.ops.ident[threadIdx.x];
        __syncthreads();
        if (S <= 16) {
            // whole rows per thread, the identity row held in registers
            uint64_t iv[16];
#pragma unroll
            for (int w = 0; w < 16; ++w) iv[w] = w < S ? ident_row[w] : 0;
            for (int i = threadIdx.x; i < n; i += blockDim.x) {
                uint64_t *row = table + (int64_t)i * S;
#pragma unroll
                for (int w = 0; w < 16; ++w)
                    if (w < S) row[w] = iv[w];
            }
        } else {
            fill_ident(table, n, S, ident_row);
        }
        if (M.part >= 0 && threadIdx.x < 64) {  // exclusive scan of the partition's cell counts
            const int per = (G + 63) / 64;
            unsigned s1 = 0;
            for (int k = 0; k < per; ++k) {
                const int g = lane * per + k;
                if (g < G) s1 += pre[g];
            }
            unsigned i1 = s1;
            for (int d = 1; d < 64; d <<= 1) {
                const unsigned t = __shfl_up(i1, d, 64);
                if (lane >= d) i1 += t;
            }
            unsigned e1 = i1 - s1;
            for (int k = 0; k < per; ++k) {
                const int g = lane * per + k;
                if (g < G) {
                    const unsigned c1 = pre[g];
                    pre[g] = e1;
                    e1 += c1;
                }
            }
            if (lane == 63) pre[G] = i1;
        }
        if (M.null_group)  // LDS atomics: the units' NULL-group states merge in parallel
            for (int g = threadIdx.x; g < G; g += blockDim.x)
                merge_state<LDS_SCOPE>(null_row, Q.null_blocks + (int64_t)g * S, a.ops);
        __syncthreads();
        if (M.part < 0) {
            // BQ_DENSE: every unit's block over slots [k0, k1); a thread owns its slots
            for (int g = 0; g < G; ++g) {
                const uint64_t *blk = Q.blocks + ((int64_t)g * Q.range + M.k0) * S;
                for (int i = threadIdx.x; i < n; i += blockDim.x)
                    merge_state<LDS_SCOPE>(table + (int64_t)i * S, blk + (int64_t)i * S, a.ops);
            }
        } else {
            const unsigned n_rec = pre[G];
            const int64_t slot0 = (int64_t)M.part << Q.shift;
            const int64_t cell_stride = (int64_t)Q.P * Q.Cw * Q.RW;
            const int64_t *part_cells = Q.cells + (int64_t)M.part * Q.Cw * Q.RW;
            const unsigned step = blockDim.x;
            const int RW = Q.RW;
            RecFmt f;
            f.pack = Q.pack;
            f.kbits = Q.pk_kbits;
            f.vmin = Q.pk_vmin;
            f.valid_word = Q.has_valid_word ? a.n_in + (Q.pack ? 0 : 1) : -1;
            const unsigned first = threadIdx.x - lane;
            const unsigned pre_lane = lane <= G ? pre[lane] : ~0u;
            if (RW == 2) {  // the common shape: MERGE_INFLIGHT 16-byte records per lane in flight
                for (unsigned i0 = first; i0 < n_rec; i0 += MERGE_INFLIGHT * step)
                    merge_round<MERGE_INFLIGHT, 2>(table, S, pk_s, n_ops, f, pre, pre_lane, G, n_rec, i0, step,
                                                   part_cells, cell_stride, RW, slot0);
            } else {
                for (unsigned i0 = first; i0 < n_rec; i0 += step)
                    merge_round<1, MAX_IN + 2>(table, S, pk_s, n_ops, f, pre, pre_lane, G, n_rec, i0, step,
                                               part_cells, cell_stride, RW, slot0);
            }
        }
        __syncthreads();
        // emit the occupied slots (and the NULL group) with one reservation:
        // the occupied slots are listed in output order first, so that the
        // emitting lanes are all busy and their stores are contiguous
        unsigned n_have = 0;
        for (int i = threadIdx.x; i < n; i += blockDim.x) n_have += table[(int64_t)i * S] > 0 ? 1u : 0u;
        unsigned total;
        unsigned pos = block_excl_scan(n_have, total);
        const bool emit_null = M.null_group && null_row[0] > 0;  // uniform
        // the output reservation is issued before the list is built and read
        // after it, so the list hides the atomic's round trip
        unsigned long long res = 0;
        if (threadIdx.x == 0 && total + (emit_null ? 1u : 0u))
            res = atomicAdd(Q.out_count, (unsigned long long)(total + (emit_null ? 1u : 0u)));
        for (int i = threadIdx.x; i < n; i += blockDim.x)
            if (table[(int64_t)i * S] > 0) list[pos++] = (unsigned)i;
        if (threadIdx.x == 0) out_base = res;
        __syncthreads();
        const int64_t base = (int64_t)out_base;
        const int64_t key0 = Q.key_min + (M.part >= 0 ? ((int64_t)M.part << Q.shift) : M.k0);
        for (unsigned j = threadIdx.x; j < total; j += blockDim.x) {
            const int64_t o = base + j;
            if (o >= Q.out_cap) {
                atomicExch(A.flags + 1, 1ull);
                break;
            }
            const unsigned i = list[j];
            gst(Q.out_key + o, key0 + i);
            gst(Q.out_key_valid + o, (uint8_t)1);
            emit_row(Q.outs, table + (int64_t)i * S, o, Q.out_val, Q.out_valid);
        }
        if (emit_null && threadIdx.x == 0) {
            const int64_t o = base + total;
            if (o < Q.out_cap) {
                gst(Q.out_key + o, (int64_t)0);
                gst(Q.out_key_valid + o, (uint8_t)0);
                emit_row(Q.outs, null_row, o, Q.out_val, Q.out_valid);
            } else {
                atomicExch(A.flags + 1, 1ull);
            }
        }
        if (threadIdx.x == 0) unit_slot = next;
        __syncthreads();  // the next unit re-initialises the LDS
    }
}

// The pass-1 instantiation for a batch's predicate columns and inputs.
const void *batch_scan_fn(int npc, int nin) {
    if (nin <= 2) {
        if (npc <= 1) return (const void *)k_batch_scan<1, 2>;
        if (npc <= 2) return (const void *)k_batch_scan<2, 2>;
        return (const void *)k_batch_scan<MAX_PCOLS, 2>;
    }
    if (npc <= 1) return (const void *)k_batch_scan<1, MAX_IN>;
    if (npc <= 2) return (const void *)k_batch_scan<2, MAX_IN>;
    return (const void *)k_batch_scan<MAX_PCOLS, MAX_IN>;
}

hipError_t launch_batch_scan(hipStream_t s, const BatchQuery *qs, const BatchArgs &a, int npc, int nin, int grid,
                             int block, size_t lds) {
    const BatchUnit *units = a.units;
    void *args[] = {(void *)&qs, (void *)&units, (void *)&a};
    return hipLaunchKernel(batch_scan_fn(npc, nin), dim3(grid), dim3(block), args, lds, s);
}

hipError_t launch_batch_merge(hipStream_t s, const BatchQuery *qs, const BatchArgs &a, int grid, int block,
                              size_t lds) {
    hipLaunchKernelGGL(k_batch_merge, dim3(grid), dim3(block), lds, s, qs, a.units2, a);
    return hipGetLastError();
}

// Resident workgroups per CU of each pass at this block size and LDS use:
// the persistent grids launch that many per CU, so a small pass-2 table lets
// several workgroups share a CU and hide each other's unit-start latency.
hipError_t batch_occupancy(int npc, int nin, int block1, int block2, size_t lds1, size_t lds2, int &per_cu1,
                           int &per_cu2) {
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu1, batch_scan_fn(npc, nin), block1, lds1);
    if (e != hipSuccess) return e;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu2, (const void *)k_batch_merge, block2, lds2);
    per_cu1 = std::max(1, per_cu1);
    per_cu2 = std::max(1, per_cu2);
    return e;
}

// Raises the dynamic LDS limit of the kernels (outside any graph capture).
hipError_t batch_set_lds(size_t lds1, size_t lds2) {
    for (int npc : {1, 2, MAX_PCOLS})
        for (int nin : {2, MAX_IN}) {
            const hipError_t e =
                hipFuncSetAttribute(batch_scan_fn(npc, nin), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds1);
            if (e != hipSuccess) return e;
        }
    return hipFuncSetAttribute((const void *)k_batch_merge, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds2);
}

}  // namespace fgpu
