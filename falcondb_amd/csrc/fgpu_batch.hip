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

constexpr int BATCH_UNROLL = 4;  // 64-row steps per wave tile

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

FGPU_DEV int64_t batch_input(const GbArgs &a, int j, int64_t r, uint32_t &valid) {
    int64_t v = 0;
    if (j < a.n_in) {
        bool ok;
        load_cell(a.cols, a.in_col[j], r, v, ok);
        valid |= ok ? 1u << j : 0u;
    }
    return v;
}

FGPU_DEV Inputs batch_inputs(const GbArgs &a, int64_t r) {
    Inputs in;
    in.valid = 0;
    in.v0 = batch_input(a, 0, r, in.valid);
    in.v1 = batch_input(a, 1, r, in.valid);
    in.v2 = batch_input(a, 2, r, in.valid);
    in.v3 = batch_input(a, 3, r, in.valid);
    in.v4 = batch_input(a, 4, r, in.valid);
    in.v5 = batch_input(a, 5, r, in.valid);
    return in;
}

// A BQ_PART record: word 0 is the key slot, or (packed) the slot with input 0
// above it; then the remaining inputs and the validity word.  A two-word
// record goes out as one 16-byte store (records are 16-byte aligned then).
FGPU_DEV void batch_put(int64_t *rec, const BatchQuery &Q, int64_t slot, const Inputs &iu) {
    const int ni = Q.g.n_in;
    int64_t w[MAX_IN + 2];  // at most the slot, MAX_IN inputs and the validity word
    w[0] = Q.pack ? (int64_t)((uint64_t)slot | ((uint64_t)(iu.v0 - Q.pk_vmin) << Q.pk_kbits)) : slot;
    const int j0 = Q.pack ? 1 : 0;  // first input after word 0
#pragma unroll
    for (int k = 1; k <= MAX_IN + 1; ++k) {
        const int j = j0 + k - 1;
        w[k] = j < ni ? pick(iu, j) : (int64_t)iu.valid;
    }
    if (Q.RW == 2) {
        gst((ll2 *)rec, ll2{w[0], w[1]});
        return;
    }
#pragma unroll
    for (int k = 0; k <= MAX_IN + 1; ++k)
        if (k < Q.RW) gst(rec + k, w[k]);
}

template <int NPC>
__global__ __launch_bounds__(1024) void k_batch_scan(const BatchQuery *__restrict__ qs,
                                                     const BatchUnit *__restrict__ units, BatchArgs A) {
    extern __shared__ uint64_t lds[];
    __shared__ unsigned long long unit_slot;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int nw = __builtin_amdgcn_readfirstlane((int)(blockDim.x >> 6));
    for (;;) {
        const int64_t ui = batch_next(A.work, &unit_slot);
        if (ui >= A.n_units) return;
        const BatchUnit U = units[ui];
        const BatchQuery &Q = qs[U.q];
        const GbArgs &a = Q.g;
        const int S = Q.S;
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
        for (int64_t base = U.row0 + (int64_t)wave * 64 * BATCH_UNROLL; base < U.row1;
             base += (int64_t)nw * 64 * BATCH_UNROLL) {
            int64_t rows[BATCH_UNROLL];
            bool pass[BATCH_UNROLL];
#pragma unroll
            for (int u = 0; u < BATCH_UNROLL; ++u) rows[u] = base + u * 64 + lane;
            pred_tile<BATCH_UNROLL, NPC>(a.pred, a.cols, rows, U.row1, pass);
            int64_t key[BATCH_UNROLL];
            bool has_key[BATCH_UNROLL];
            Inputs in[BATCH_UNROLL];
#pragma unroll
            for (int u = 0; u < BATCH_UNROLL; ++u) {
                key[u] = 0;
                has_key[u] = false;
                in[u] = Inputs{0, 0, 0, 0, 0, 0, 0};
                if (pass[u]) {
                    bool ok;
                    load_cell(a.cols, a.key_col, rows[u], key[u], ok);
                    has_key[u] = ok && key[u] != EMPTY_KEY;  // NULL and the sentinel form the NULL group
                    in[u] = batch_inputs(a, rows[u]);
                }
            }
            // one row per iteration; the u-th values are picked with selects so
            // that update_row has a single call site (the loop stays rolled)
#pragma unroll 1
            for (int u = 0; u < BATCH_UNROLL; ++u) {
                bool ps = pass[0], hk = has_key[0];
                int64_t k = key[0];
                Inputs iu = in[0];
#pragma unroll
                for (int q = 1; q < BATCH_UNROLL; ++q)
                    if (u == q) {
                        ps = pass[q];
                        hk = has_key[q];
                        k = key[q];
                        iu = in[q];
                    }
                if (!ps) continue;
                const int64_t slot = k - Q.key_min;  // in [0, range): the range comes from the column's min / max
                uint64_t *st = !hk ? null_st : dense ? table + slot * S : nullptr;
                if (st) {
                    update_row<LDS_SCOPE>(st, ops, iu);
                    continue;
                }
                const int p = (int)(slot >> Q.shift);
                const unsigned pos = atomicAdd(cursor + p, 1u);
                if (pos >= (unsigned)Q.Cw) {
                    atomicExch(A.flags, 1ull);
                    continue;
                }
                batch_put(my_cells + ((int64_t)p * Q.Cw + pos) * Q.RW, Q, slot, iu);
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < S; i += blockDim.x) Q.null_blocks[(int64_t)U.local * S + i] = null_st[i];
        if (dense) {
            uint64_t *blk = Q.blocks + (int64_t)U.local * Q.range * S;
            for (int64_t i = threadIdx.x; i < Q.range * S; i += blockDim.x) blk[i] = table[i];
        } else {
            for (int i = threadIdx.x; i < Q.P; i += blockDim.x) Q.cell_cnt[(int64_t)U.local * Q.P + i] = cursor[i];
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
template <int R, int NW>
FGPU_DEV void merge_round(uint64_t *table, int S, const uint32_t *pk_s, int n_ops, const RecFmt &f,
                          const unsigned *pre, int G, unsigned n_rec, unsigned i0, unsigned step,
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
                const unsigned c = gld(Q.cell_cnt + (int64_t)g * Q.P + M.part);
                pre[g] = c < (unsigned)Q.Cw ? c : (unsigned)Q.Cw;
            }
        if (M.null_group && threadIdx.x < S) null_row[threadIdx.x] = a.ops.ident[threadIdx.x];
        __syncthreads();
        fill_ident(table, n, S, ident_row);
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
            if (RW == 2) {  // the common shape: MERGE_INFLIGHT 16-byte records per lane in flight
                for (unsigned i0 = first; i0 < n_rec; i0 += MERGE_INFLIGHT * step)
                    merge_round<MERGE_INFLIGHT, 2>(table, S, pk_s, n_ops, f, pre, G, n_rec, i0, step, part_cells,
                                                   cell_stride, RW, slot0);
            } else {
                for (unsigned i0 = first; i0 < n_rec; i0 += step)
                    merge_round<1, MAX_IN + 2>(table, S, pk_s, n_ops, f, pre, G, n_rec, i0, step, part_cells,
                                               cell_stride, RW, slot0);
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
        for (int i = threadIdx.x; i < n; i += blockDim.x)
            if (table[(int64_t)i * S] > 0) list[pos++] = (unsigned)i;
        const bool emit_null = M.null_group && null_row[0] > 0;  // uniform
        const int64_t base = (int64_t)block_reserve(Q.out_count, total + (emit_null ? 1u : 0u));  // has barriers
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

hipError_t launch_batch_scan(hipStream_t s, const BatchQuery *qs, const BatchArgs &a, int npc, int grid,
                             int block, size_t lds) {
    if (npc <= 1) hipLaunchKernelGGL(k_batch_scan<1>, dim3(grid), dim3(block), lds, s, qs, a.units, a);
    else if (npc <= 2) hipLaunchKernelGGL(k_batch_scan<2>, dim3(grid), dim3(block), lds, s, qs, a.units, a);
    else hipLaunchKernelGGL(k_batch_scan<MAX_PCOLS>, dim3(grid), dim3(block), lds, s, qs, a.units, a);
    return hipGetLastError();
}

hipError_t launch_batch_merge(hipStream_t s, const BatchQuery *qs, const BatchArgs &a, int grid, int block,
                              size_t lds) {
    hipLaunchKernelGGL(k_batch_merge, dim3(grid), dim3(block), lds, s, qs, a.units2, a);
    return hipGetLastError();
}

// Resident workgroups per CU of each pass at this block size and LDS use:
// the persistent grids launch that many per CU, so a small pass-2 table lets
// several workgroups share a CU and hide each other's unit-start latency.
hipError_t batch_occupancy(int npc, int block, size_t lds1, size_t lds2, int &per_cu1, int &per_cu2) {
    const void *scan = npc <= 1 ? (const void *)k_batch_scan<1>
                       : npc <= 2 ? (const void *)k_batch_scan<2>
                                  : (const void *)k_batch_scan<MAX_PCOLS>;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu1, scan, block, lds1);
    if (e != hipSuccess) return e;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu2, (const void *)k_batch_merge, block, lds2);
    per_cu1 = std::max(1, per_cu1);
    per_cu2 = std::max(1, per_cu2);
    return e;
}

// Raises the dynamic LDS limit of the kernels (outside any graph capture).
hipError_t batch_set_lds(size_t lds1, size_t lds2) {
    const void *scans[3] = {(const void *)k_batch_scan<1>, (const void *)k_batch_scan<2>,
                            (const void *)k_batch_scan<MAX_PCOLS>};
    for (const void *f : scans) {
        const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds1);
        if (e != hipSuccess) return e;
    }
    return hipFuncSetAttribute((const void *)k_batch_merge, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds2);
}

}  // namespace fgpu
