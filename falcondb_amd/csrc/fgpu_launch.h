// fgpu_launch.h — kernel argument blocks and host-side launchers (one set per
// kernel source file).  Everything here is internal to libfalcon_gpu.so.
#pragma once

#include "fgpu_internal.h"

namespace fgpu {

constexpr int MAX_IN = 6;  // aggregate input columns per query

// Output descriptor of one aggregate, resolved from the state layout.
enum : int8_t {
    OUT_COUNT_ROWS = 0,  // COUNT(*): word 0
    OUT_COUNT = 1,       // COUNT(col): a count word
    OUT_SUM_I = 2,
    OUT_AVG_I = 3,
    OUT_SUM_F = 4,
    OUT_AVG_F = 5,
    OUT_MIN_I = 6,
    OUT_MAX_I = 7,
    OUT_MIN_F = 8,
    OUT_MAX_F = 9
};
struct DOut {
    int32_t kind;
    int32_t W;
    int32_t has_flags;
    int32_t word;      // value word (first fixed-point word for float sums)
    int32_t cnt_word;  // non-NULL count word (0 = row count)
    int32_t L;
};
struct DOuts {
    int n;
    DOut o[FGPU_MAX_AGGS];
};

// ─── GROUP BY ───────────────────────────────────────────────────────────────
enum : int { GB_DENSE_LDS = 0, GB_DENSE_GLOBAL = 1, GB_HASH = 2, GB_SCALAR = 3 };

struct GbArgs {
    DCols cols;
    DPred pred;
    DOps ops;
    int n_in;
    int32_t in_col[MAX_IN];  // column slot of each aggregate input
    int key_col;            // column slot of the key (-1 = scalar)
    int key_is_utf8;
    uint32_t utf8_null_code;  // code NULL Utf8 keys are mapped to
    int64_t key_min;          // dense modes: index = key - key_min
    int64_t range;            // dense modes: key slots; slot `range` = NULL group
    int64_t n_rows;
    int64_t *keys;            // hash mode: cap keys
    uint64_t *state;          // (cap + 1) * S words (dense: (range + 1) * S)
    uint64_t cap_mask;        // hash mode: cap - 1
    unsigned long long *counters;  // [0] groups inserted, [1] overflow flag
    int lds_slots;            // hash mode: LDS cache slots (power of two, 0 = off)
    int probe_limit;          // hash mode: LDS probes before falling back to HBM
    // dense-LDS and scalar modes: workgroup b merges its states into replica
    // b % replicas (rows rep_stride words apart), so that fewer workgroups
    // contend for each HBM word; launch_gb_fold folds them into replica 0
    int replicas;
    int64_t rep_stride;
};

struct EmitArgs {
    DOuts outs;
    const int64_t *keys;      // hash mode (else NULL)
    const uint64_t *state;
    int S;
    int64_t n_slots;          // hash: cap + 1; dense: range + 1; scalar: 1
    int mode;
    int64_t key_min;
    int key_is_utf8;
    int64_t *out_key;
    uint8_t *out_key_valid;
    int64_t *out_val[FGPU_MAX_AGGS];
    uint8_t *out_valid[FGPU_MAX_AGGS];
    unsigned long long *out_count;
};

// Radix-partitioned GROUP BY for high-cardinality keys.  Pass 1 streams the
// rows once: keys found in (or admitted to) a per-workgroup LDS cache are folded
// there; every other row becomes a record {key, inputs[, validity]} appended to
// cell (workgroup, partition) from the front; at the end the cache is written
// as state records {key, S words} into the back of the same cells.  Pass 2
// aggregates one partition per workgroup entirely in an LDS table and emits its
// groups; NULL-key partial states are merged by a last one-workgroup kernel.
struct GbPartArgs {
    GbArgs g;                 // columns, predicate, state layout, key
    int log_p;                // P = 1 << log_p partitions
    int H;                    // pass-1 LDS cache slots (power of two)
    int probe_limit;
    int G;                    // pass-1 workgroups
    int64_t rows_per_wg;
    int Cw;                   // record slots per cell
    int RW;                   // words per record slot
    int has_valid_word;       // records carry the inputs' validity bits
    int SR;                   // record slots per state record
    int64_t *cells;           // (P + 2) * G * Cw * RW words, partition-major: partitions 0..P-1,
                              // NULL keys, sentinel keys; cell (w, p) at ((p * G + w) * Cw) * RW
    unsigned int *cell_cnt;   // G * (P + 2) * 2: row records, state records
    uint64_t *nullbuf;        // G * S
    unsigned long long *flags;  // [0] cell full, [1] pass-2 table full, [2] output full
    int null_group;           // the key column holds NULL or INT64_MIN rows: merge the NULL group
    int T2;                   // pass-2 LDS table slots (power of two)
    // Packed row records: when pack is set, a row record's first word holds
    // (key - pk_kmin) | ((input 0 - pk_vmin) << pk_kbits) (the key column
    // has no NULL, input 0 is an integer column without NULLs, and the two
    // ranges fit 64 bits), then inputs 1 .. n_in - 1; RW is one word shorter.
    int pack, pk_kbits;
    int64_t pk_kmin, pk_vmin;
    DOuts outs;
    int64_t out_cap;
    int64_t *out_key;
    uint8_t *out_key_valid;
    int64_t *out_val[FGPU_MAX_AGGS];
    uint8_t *out_valid[FGPU_MAX_AGGS];
    unsigned long long *out_count;
};
hipError_t launch_gbp_scan(hipStream_t s, const GbPartArgs &a, int block, size_t lds);
// Predicate width (4 / 8) when the scan takes its specialised tile loads, else 0.
int gbp_fast_width(const GbPartArgs &a);
hipError_t launch_gbp_agg(hipStream_t s, const GbPartArgs &a, int block, size_t lds);

hipError_t launch_state_init(hipStream_t s, int64_t *keys, int64_t n_keys, uint64_t *state,
                             int64_t n_groups, const DOps &ops);
hipError_t launch_gb_fold(hipStream_t s, const GbArgs &a, int64_t n_groups);
hipError_t launch_groupby(hipStream_t s, int mode, const GbArgs &a, int grid, int block,
                          size_t lds_bytes);
hipError_t launch_emit(hipStream_t s, const EmitArgs &a);

// One int64 GROUP BY key per row from several key columns or one Float64 key
// (see k_composite_key).
constexpr int MAX_GROUP_KEYS = 4;
struct CompositeKeyArgs {
    int n;                    // key columns
    int float_key;            // 1: one Float64 column, out = canonical bits + 1 (NULL: INT64_MIN)
    const void *v[MAX_GROUP_KEYS];
    const uint64_t *valid[MAX_GROUP_KEYS];
    int dtype[MAX_GROUP_KEYS];
    int64_t kmin[MAX_GROUP_KEYS];       // integer columns: the value coded 1
    int64_t mul[MAX_GROUP_KEYS];        // weight of the column's code
    int64_t null_code[MAX_GROUP_KEYS];  // text columns: code of a NULL key
    int64_t n_rows;
    int64_t *out;
};
hipError_t launch_composite_key(hipStream_t s, const CompositeKeyArgs &a, int grid);

// ─── WHERE ──────────────────────────────────────────────────────────────────
hipError_t launch_filter_count(hipStream_t s, const DCols &c, const DPred &p, int64_t n,
                               unsigned long long *d_count, int grid);
hipError_t launch_filter_mask(hipStream_t s, const DCols &c, const DPred &p, int64_t n,
                              uint64_t *d_words, int grid);
// Shared COUNT(*) scan: q single-comparison predicates on one column.
constexpr int SHARED_MAX_Q = 32;
struct SharedCount {
    int q;
    int dtype;                 // FGPU_INT32 / FGPU_INT64 / FGPU_FLOAT64
    const void *values;
    const uint32_t *valid;     // validity bits as 32-bit words, or null
    int64_t n;
    unsigned long long *out;   // q counters (zeroed by the caller)
    DPredNode leaf[SHARED_MAX_Q];
};
hipError_t launch_filter_count_shared(hipStream_t s, const SharedCount &a, int grid);
// Exclusive prefix sum of per-word popcounts: offs[i] = sum_{j<i} popcount(words[j]);
// offs[n_words] = total.  `block_tmp` holds ceil(n_words / 4096) + 1 entries.
hipError_t launch_word_offsets(hipStream_t s, const uint64_t *words, int64_t n_words,
                               int64_t *offs, int64_t *block_tmp);
hipError_t launch_gather(hipStream_t s, const uint64_t *words, const int64_t *offs,
                         int64_t n_words, const void *src, int elem_bytes,
                         const uint64_t *src_valid, void *dst, uint8_t *dst_valid, int grid);
// dst[i] = src[idx[i]] with validity (NO_ROW = NULL); see k_gather_idx.
hipError_t launch_gather_idx(hipStream_t s, const void *src, int elem_bytes, int utf8, const uint64_t *src_valid,
                             const uint32_t *idx, int64_t n, void *dst, uint64_t *dst_valid, int grid);
// Exclusive prefix sum of u32 counts into int64 offsets (offs[n] = total).
hipError_t launch_scan_u32(hipStream_t s, const uint32_t *cnt, int64_t n, int64_t *offs,
                           int64_t *block_tmp);

// ─── column statistics (upload) ─────────────────────────────────────────────
struct ColStatsDev {
    long long imin, imax;               // integers / codes
    unsigned long long fmin_enc, fmax_enc;  // doubles (order-encoded, NaN skipped)
    int emin, emax;                     // mantissa-LSB exponents of finite nonzero doubles
    unsigned long long nonfinite;       // flags FL_*
    unsigned long long nulls;
    unsigned int hll[4096];             // HyperLogLog registers (p = 12)
};
hipError_t launch_dict_remap(hipStream_t s, uint32_t *codes, int64_t n, const int64_t *row_off, int n_batches,
                             const uint32_t *lut, const int64_t *lut_off, unsigned long long *bad, int grid);
hipError_t launch_batch_zones(hipStream_t s, const void *values, int dtype, const uint64_t *valid, int64_t n,
                              int64_t batch_rows, BatchZone *out);
hipError_t launch_col_stats(hipStream_t s, const void *values, int dtype, const uint64_t *valid,
                            int64_t n, ColStatsDev *d_stats, int grid);

// ─── hash join ──────────────────────────────────────────────────────────────
struct JoinBuildArgs {
    const void *keys_col;   // build key column
    int key_dtype;          // FGPU_INT32 / FGPU_INT64
    const uint64_t *key_valid;
    int64_t n;
    int64_t *slots;         // 2 * cap words: slot i = {key (EMPTY = INT64_MIN), build row}
    uint64_t cap_mask;
    unsigned long long *counters;  // [1] overflow, [2] rows with key INT64_MIN, [3] duplicate keys
    uint32_t *min_key_rows;        // build rows whose key is INT64_MIN (list)
    int64_t min_key_cap;
};

// Per-thread register aggregate ops of a join / scalar aggregate.
struct JoinOp {
    int32_t kind;    // OP_CNT .. OP_FMAX (OP_CNT with col < 0 = COUNT(*))
    int32_t side;    // 0 probe, 1 build
    int32_t col;     // column slot within that side's DCols
    int32_t word;    // destination word in the global state row
    int32_t L;
    int32_t W;
    int32_t has_flags;
    int32_t sc;      // OP_FSUM: values are summed times rscale = 2^-sc
    double rscale;
};
struct JoinProbeArgs {
    DCols pcols;      // probe columns
    DCols bcols;      // build columns
    DPred pred;       // probe-side WHERE
    int pkey;         // probe key slot in pcols
    int type;         // FGPU_INNER..FGPU_ANTI
    int64_t n_probe;
    const int64_t *slots;  // 2 * cap words: {key, build row}
    uint64_t cap_mask;
    const uint32_t *min_key_rows;  // build rows with key INT64_MIN
    int64_t n_min_key_rows;
    int unique_keys;  // build keys are distinct: a probe stops at its first match
    int n_ops;
    JoinOp ops[FGPU_MAX_AGGS];
    int S;
    uint64_t *state;  // one state row (S words), identity-initialised
    // pair mode
    uint32_t *counts;     // pass 1: output rows per probe row
    const int64_t *offs;  // pass 2: exclusive offsets
    uint32_t *out_l;
    uint32_t *out_r;
};
// Radix-partitioned join.  The build side is split by the top `log_p` hash
// bits into P sub-tables of T 16-byte slots {key, payload} (payload = the one
// build column the aggregates read, as raw bits, or the build row); pass 1
// streams the probe side once and appends records {key, probe payload...} to
// per-(workgroup, partition) cells; pass 2 probes each partition's records
// against its sub-table with every workgroup of one XCD on the same partition,
// so the sub-table stays in that XCD's L2.
struct JoinPartArgs {
    JoinProbeArgs pa;         // columns, predicate, type, ops, state row, side list
    int log_p;                // P = 1 << log_p
    int T;                    // slots per sub-table (power of two)
    int64_t *sub;             // P * T * 2 words
    unsigned int *sub_cnt;    // keys per sub-table
    int build_pay_col;        // bcols slot of the inlined build column (-1 = build row)
    int G;                    // pass-1 workgroups
    int64_t rows_per_wg;
    int Cw;                   // records per cell
    int RW;                   // words per record (key + probe payloads)
    int n_pay;                // probe payload columns carried in records (<= 2)
    int32_t pay_col[2];       // their pcols slots
    int32_t field_of_op[FGPU_MAX_AGGS];  // probe-side op -> record field (1..n_pay)
    int64_t *cells;           // G * P * Cw * RW words
    unsigned int *cell_cnt;   // G * P
    unsigned long long *flags;  // [0] cell overflow, [1] sub-table overflow, [2] side-list rows,
                                // [3] duplicate build keys seen
    uint32_t *min_key_rows;   // build rows with key INT64_MIN (pass-1 build output)
    int64_t min_key_cap;
    int split;                // pass-2 workgroups per partition
    int need_vals;            // an aggregate reads a build column (pass 2 keeps slot values in LDS)
    const int64_t *xoff;      // exact partitioning: record offset of (w, p) at [p * G + w]
    const int64_t *chunks;    // exact partitioning: pass-2 work items {p, first record, records}
    // Record format.  kw = 8: a record is RW int64 words {hash64(key), payloads}
    // in `cells`.  kw = 4 (compact; build keys span kbits <= 31 + log_p bits):
    // h = cmix(key - kmin) over kbits bits, the partition is h's top log_p
    // bits and the record key its low rbits bits; record i is the uint32 key
    // and the RW - 1 payloads packed at byte 4 * crec_words(RW, pw) * i of
    // `cells`: each payload as two uint32 words (low first), or, with pw = 4
    // (one integer payload whose range fits 32 bits), as one uint32 offset
    // from pmin.  Probe keys outside [kmin, kmax] go to the NULL-key side
    // partition.
    int kw;
    int kbits, rbits;
    int64_t kmin, kmax;
    int pw;
    int64_t pmin;
    int lds_budget;           // pass-1 LDS bytes per workgroup (the G workgroups share the CUs' LDS)
};
// The empty slot of a compact sub-table in LDS (record keys are < 2^31).
constexpr uint32_t EMPTY32 = 0xFFFFFFFFu;
// uint32 words of a compact record with RW - 1 payloads of pw bytes.
FGPU_HD int crec_words(int RW, int pw) { return 1 + (RW - 1) * (pw == 4 ? 1 : 2); }
// Partitioned join pass 1: JS_TPB threads per workgroup, join_scatter_rows()
// rows each per tile.
constexpr int JS_TPB = 1024;
constexpr size_t JOIN_LDS_MAX = 159 * 1024;  // dynamic LDS a join workgroup may take (static use < 1 KB)
// LDS bytes of one pass-1 / pass-2 workgroup for these arguments.
int join_scatter_rows(const JoinPartArgs &a, int mode);
size_t join_scatter_lds(const JoinPartArgs &a, int mode);
size_t join_probe_lds(const JoinPartArgs &a);
hipError_t launch_join_part_count(hipStream_t s, const JoinPartArgs &a, const JoinBuildArgs &b, int grid);
hipError_t launch_join_part_build(hipStream_t s, const JoinPartArgs &a, const JoinBuildArgs &b, int grid);
// mode: 0 cells (optimistic), 1 count, 2 exact (see k_jp_scatter)
hipError_t launch_join_part_scatter(hipStream_t s, const JoinPartArgs &a, int mode);
// n_chunks > 0: chunk mode over a.chunks (grid = n_chunks); else cell mode.
hipError_t launch_join_part_probe(hipStream_t s, const JoinPartArgs &a, int n_chunks);
hipError_t launch_join_build(hipStream_t s, const JoinBuildArgs &a, int grid);
hipError_t launch_join_probe_agg(hipStream_t s, const JoinProbeArgs &a, int grid, int block);
hipError_t launch_join_pairs(hipStream_t s, const JoinProbeArgs &a, int pass, int grid);

// Fill helpers
hipError_t launch_fill_u64(hipStream_t s, uint64_t *p, int64_t n, uint64_t v);

// ─── batched GROUP BY ─────────────────────────────────────────────────────
// Many GROUP BY queries (each over its own table) run as one persistent pass 1
// over (query, row range) units and one persistent pass 2 over (query, key
// range) units.  Keys are integers with a dense range: slot = key - key_min.
enum : int32_t {
    BQ_DENSE = 0,  // pass 1 aggregates the whole key range in LDS, one state block per unit
    BQ_PART = 1    // pass 1 writes (key, inputs) records into key-range partitions
};

struct BatchQuery {
    GbArgs g;                 // columns, predicate, ops, key slot, n_rows
    DOuts outs;
    int mode;                 // BQ_*
    int64_t key_min;
    int64_t range;            // dense key slots [0, range); the NULL group is separate
    int shift;                // BQ_PART: partition = slot >> shift (2^shift slots each)
    int P;                    // BQ_PART partitions
    int S, RW, Cw, has_valid_word;
    int pack, pk_kbits;       // BQ_PART records: slot | (input 0 - pk_vmin) << pk_kbits in word 0
    int64_t pk_vmin;
    // Specialised pass-1 tile (fast != 0): a BQ_PART query of one- or two-word records
    // whose key, inputs and predicate column have no NULLs and whose
    // predicate is one comparison (or none).  Widths in bytes: key 4 / 8,
    // predicate column 0 (none) / 4 / 8, inputs 4 / 8 (INT32 sign-extended).
    int fast, fast_kw, fast_pw, fast_iw0, fast_iw1;
    int64_t unit0, n_units;   // this query's pass-1 units
    int64_t *cells;           // BQ_PART: n_units * P * Cw * RW words
    unsigned int *cell_cnt;   // BQ_PART: P * n_units record counts (partition-major)
    uint64_t *blocks;         // BQ_DENSE: n_units * range * S words
    uint64_t *null_blocks;    // n_units * S words: each unit's NULL-key group state
    int64_t *out_key;
    uint8_t *out_key_valid;
    int64_t *out_val[FGPU_MAX_AGGS];
    uint8_t *out_valid[FGPU_MAX_AGGS];
    unsigned long long *out_count;
    int64_t out_cap;
};

struct BatchUnit {            // pass 1: rows [row0, row1) of query q (unit `local` of it)
    int32_t q;
    int32_t local;
    int64_t row0, row1;
};
struct MergeUnit {            // pass 2: dense slots [k0, k1) or partition `part` of query q
    int32_t q;
    int32_t part;             // BQ_PART partition, or -1 for a BQ_DENSE slot range
    int64_t k0, k1;
    int32_t null_group;       // this unit also merges and emits the NULL-key group
    int32_t pad;
};

struct BatchArgs {
    const BatchUnit *units;
    int64_t n_units;
    const MergeUnit *units2;
    int64_t n_units2;
    unsigned long long *work;   // [0] pass-1 cursor, [1] pass-2 cursor (zeroed before each run)
    unsigned long long *flags;  // [0] cell full, [1] output full
    int lds_slots;              // LDS table slots (pass 1 BQ_DENSE range cap, pass 2 partition size)
};
hipError_t launch_batch_scan(hipStream_t s, const BatchQuery *qs, const BatchArgs &a, int npc, int nin, int grid,
                             int block, size_t lds);
hipError_t launch_batch_merge(hipStream_t s, const BatchQuery *qs, const BatchArgs &a, int grid, int block,
                              size_t lds);
hipError_t batch_set_lds(size_t lds1, size_t lds2);
hipError_t batch_occupancy(int npc, int nin, int block1, int block2, size_t lds1, size_t lds2, int &per_cu1,
                           int &per_cu2);

// ─── ORDER BY with any LIMIT and any number of keys (fgpu_sort.hip) ────────
hipError_t launch_mask_rows(hipStream_t s, const uint64_t *words, const int64_t *offs, int64_t n_words,
                            uint32_t *out);
hipError_t launch_sort_key(hipStream_t s, const void *values, int dtype, const uint64_t *valid, const uint32_t *rank,
                           const uint32_t *rows, int64_t n, int desc, int nulls_last, uint64_t *key, uint8_t *flag);
hipError_t launch_take_u64(hipStream_t s, const uint64_t *src, const uint32_t *idx, int64_t n, uint64_t *dst);
hipError_t launch_take_u8(hipStream_t s, const uint8_t *src, const uint32_t *idx, int64_t n, uint8_t *dst);
hipError_t launch_take_u32(hipStream_t s, const uint32_t *src, const uint32_t *idx, int64_t n, uint32_t *dst);
hipError_t launch_iota_u32(hipStream_t s, uint32_t *p, int64_t n);
hipError_t sort_pairs_u64(void *temp, size_t *temp_bytes, const uint64_t *k_in, uint64_t *k_out, const uint32_t *v_in,
                          uint32_t *v_out, int64_t n, int end_bit, hipStream_t s);
hipError_t sort_pairs_u8(void *temp, size_t *temp_bytes, const uint8_t *k_in, uint8_t *k_out, const uint32_t *v_in,
                         uint32_t *v_out, int64_t n, hipStream_t s);

// ─── ORDER BY ... LIMIT k ──────────────────────────────────────────────────
// cands: grid * K pairs of 16 bytes; out: K pairs (best first).
hipError_t launch_topk(hipStream_t s, const DCols &c, const DPred &p, int col, int desc, int null_mode, int64_t n,
                       int K, int grid, void *cands, void *out);

// Kernel arguments travel in the 4 KiB kernarg segment; descriptor fields are
// 32-bit so that the kernels read them with scalar loads.
static_assert(sizeof(GbPartArgs) <= 4096, "GbPartArgs exceeds the kernarg segment");
static_assert(sizeof(JoinProbeArgs) <= 4096, "JoinProbeArgs exceeds the kernarg segment");
static_assert(sizeof(JoinPartArgs) <= 4096, "JoinPartArgs exceeds the kernarg segment");

}  // namespace fgpu
