/*
 * falcon_gpu.h — C ABI of the MI355X (gfx950) execution backend for FalconDB's
 * WHERE filter, hash GROUP BY aggregation and IntHashTable hash join.
 *
 * Every entry point takes plain pointers and sizes; no HIP or torch type crosses
 * the boundary (a stream, where one is accepted, is passed as `void*`).  Each
 * declaration names the reference routine it replaces (paths are relative to the
 * FalconDB source tree).  A Rust `extern "C"` block declaring exactly these
 * functions is the binding the reference's query layer would add.
 *
 * Conventions (mirroring src/api/c_ffi.rs:26-144 and src/storage/types.rs:303-333):
 *   - Functions return an `int` status: FGPU_OK (0) or one of the FGPU_E_* codes.
 *     Codes 1..9 are the FalconError variants in declaration order; 100+ are
 *     device-side failures.  fgpu_last_error() returns the thread-local message
 *     of the most recent failure on the calling thread.
 *   - Handles are opaque; whoever receives one frees it with the matching *_free.
 *   - Host input buffers are borrowed only for the duration of a call.  Device
 *     copies (tables, join indexes) are owned by their handle.
 *   - A table is immutable after upload and may be shared by every context and
 *     thread on the same device (like the reference's Arc<Column>).  One context
 *     serialises its own calls on its own stream; use one context per concurrent
 *     caller.
 *   - Result rows of GROUP BY come back in unspecified order (SQL gives none;
 *     the reference's first-seen order is restated by the host layer on request).
 */
#ifndef FALCON_GPU_H
#define FALCON_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FGPU_ABI_VERSION 1

/* ── status codes ─────────────────────────────────────────────────────────── */
enum {
    FGPU_OK = 0,
    FGPU_E_COLUMN_NOT_FOUND = 1,     /* FalconError::ColumnNotFound */
    FGPU_E_TYPE_MISMATCH = 2,        /* FalconError::TypeMismatch */
    FGPU_E_INDEX_OUT_OF_BOUNDS = 3,  /* FalconError::IndexOutOfBounds */
    FGPU_E_PARSE = 4,                /* FalconError::ParseError */
    FGPU_E_IO = 5,                   /* FalconError::Io */
    FGPU_E_INVALID_QUERY = 6,        /* FalconError::InvalidQuery */
    FGPU_E_DIVISION_BY_ZERO = 7,     /* FalconError::DivisionByZero */
    FGPU_E_OVERFLOW = 8,             /* FalconError::Overflow */
    FGPU_E_CONSTRAINT = 9,           /* FalconError::ConstraintViolation */
    FGPU_E_DEVICE = 100,             /* HIP runtime failure */
    FGPU_E_OUT_OF_MEMORY = 101,      /* device allocation failed / HBM budget exhausted */
    FGPU_E_UNSUPPORTED = 102         /* input shape the kernels do not accept (message says why) */
};

/* ── column types (src/storage/types.rs:4-19 subset on this path) ─────────── */
enum {
    FGPU_INT32 = 1,
    FGPU_INT64 = 2,
    FGPU_FLOAT64 = 3,
    FGPU_UTF8 = 4   /* dictionary-encoded: u32 codes, code 0xFFFFFFFF = NULL
                       (src/storage/column.rs:1272-1277) */
};

typedef struct fgpu_ctx fgpu_ctx;
typedef struct fgpu_table fgpu_table;
typedef struct fgpu_result fgpu_result;
typedef struct fgpu_join_index fgpu_join_index;

/* One column of a table to upload.  Mirrors PrimitiveColumn{values, validity}
 * (src/storage/column.rs:842-845) and Utf8Column{dict, codes}
 * (src/storage/column.rs:1272-1277). */
typedef struct {
    const char *name;
    int32_t dtype;              /* FGPU_INT32 / FGPU_INT64 / FGPU_FLOAT64 / FGPU_UTF8 */
    const void *values;         /* n_rows elements: int32 / int64 / double / uint32 codes */
    const uint64_t *validity;   /* NULL = no NULLs; else ceil(n/64) words, LSB-first,
                                   bit set = valid (src/storage/column.rs:132-133) */
    int32_t dict_len;           /* FGPU_UTF8 only */
    const char *const *dict;    /* FGPU_UTF8 only: dict_len NUL-terminated strings */
} fgpu_column_desc;

/* ── predicates (Expr::eval_mask_bits, src/query/expr.rs:635-808) ────────── */
enum { FGPU_EQ = 0, FGPU_NE = 1, FGPU_LT = 2, FGPU_LE = 3, FGPU_GT = 4, FGPU_GE = 5 };
enum {
    FGPU_P_CMP = 0,         /* column op literal */
    FGPU_P_AND = 1,         /* pops two, pushes a AND b */
    FGPU_P_OR = 2,          /* pops two, pushes a OR b */
    FGPU_P_IS_NULL = 3,
    FGPU_P_IS_NOT_NULL = 4,
    /* col [NOT] IN (list): the union of the list's `=` comparisons, each
     * paired with the column as `=` is; NOT IN is its complement without the
     * NULL rows (Expr::InList in try_fast_mask, src/query/expr.rs:756-803).
     * op = 0 (IN) or 1 (NOT IN); the n_list literals at `list` are all of
     * lit_type: int64_t, double, or const char * for a Utf8 column. */
    FGPU_P_IN = 5,
    /* Utf8 col [NOT] [I]LIKE lit_str, decided once per dictionary entry with
     * '%' (any bytes) and '_' (one byte), no escape character; a NULL row
     * passes NOT LIKE only (Expr::Like in try_fast_mask, src/query/expr.rs:
     * 697-754; like_matches :1180-1199).  op bit 0: NOT; bit 1: ILIKE
     * (ASCII letters fold to lower case). */
    FGPU_P_LIKE = 6
};
/* A predicate is a postfix program of at most FGPU_MAX_PRED_NODES nodes. */
#define FGPU_MAX_PRED_NODES 16
typedef struct {
    int32_t kind;        /* FGPU_P_* */
    int32_t col;         /* CMP / IS_NULL / IS_NOT_NULL: column index in the table */
    int32_t op;          /* CMP: FGPU_EQ..FGPU_GE */
    int32_t lit_type;    /* CMP: FGPU_INT64, FGPU_INT32, FGPU_FLOAT64 or FGPU_UTF8 */
    int64_t lit_i64;
    double lit_f64;
    const char *lit_str; /* FGPU_UTF8 literal; LIKE pattern */
    int32_t n_list;      /* IN: number of literals */
    const void *list;    /* IN: n_list literals of lit_type */
} fgpu_pred_node;
typedef struct {
    int32_t n_nodes;
    const fgpu_pred_node *nodes;
} fgpu_pred;

/* ── aggregates (AggFunc, src/query/aggregate.rs:489-1420) ────────────────── */
enum { FGPU_COUNT = 0, FGPU_SUM = 1, FGPU_AVG = 2, FGPU_MIN = 3, FGPU_MAX = 4 };
#define FGPU_MAX_AGGS 8
typedef struct {
    int32_t func;        /* FGPU_COUNT..FGPU_MAX */
    int32_t col;         /* column index; -1 = COUNT(*) */
    int32_t side;        /* joins only: 0 = probe (FROM) table, 1 = build table */
    const char *alias;   /* output name; NULL = "col<i>" like src/query/sql.rs:780-788 */
} fgpu_agg;

typedef struct {
    int32_t key_col;     /* GROUP BY column; -1 = no GROUP BY (execute_agg_only) */
    int32_t n_aggs;
    const fgpu_agg *aggs;
} fgpu_groupby_spec;

enum { FGPU_INNER = 0, FGPU_LEFT = 1, FGPU_SEMI = 2, FGPU_ANTI = 3 };
typedef struct {
    int32_t type;        /* FGPU_INNER / FGPU_LEFT / FGPU_SEMI / FGPU_ANTI */
    int32_t probe_key;   /* key column of the probe (left, FROM) table */
    int32_t build_key;   /* key column of the build (right) table */
    int32_t n_aggs;
    const fgpu_agg *aggs;
} fgpu_join_spec;

/* A result column.  Host memory, owned by the fgpu_result. */
typedef struct {
    const char *name;
    int32_t dtype;            /* FGPU_INT64 / FGPU_FLOAT64 / FGPU_UTF8 (codes) */
    const void *values;       /* n_rows elements (int64 / double / uint32 codes) */
    const uint8_t *valid;     /* n_rows bytes, 1 = valid; NULL = all valid */
    int32_t dict_len;         /* FGPU_UTF8 key column: the table's dictionary */
    const char *const *dict;
} fgpu_column_view;

/* Per-kernel timing of the last call on a context (hipEvents on its stream). */
typedef struct {
    float total_ms;           /* first kernel start .. last kernel end */
    float main_ms;            /* the dominant (scan / probe) kernel */
    float build_ms;           /* join build / table clear */
    float finalize_ms;        /* compaction of the result */
    int64_t bytes_main;       /* algorithmic bytes read+written by the dominant kernel */
} fgpu_timing;

/* ── context ─────────────────────────────────────────────────────────────── */
/* Opens device `device` and creates a stream and a scratch arena for this
 * context.  (Database::in_memory, src/db.rs:93; falcondb_open, src/api/c_ffi.rs:26) */
int fgpu_open(int device, fgpu_ctx **out);
/* Batches prepared on the context (fgpu_batch_prepare) use its stream: free
 * them before closing it. */
void fgpu_close(fgpu_ctx *ctx);
/* Thread-local message of the last failure on this thread ("" if none). */
const char *fgpu_last_error(void);
int fgpu_abi_version(void);
/* Returns the context's hipStream_t as void*. */
void *fgpu_stream(fgpu_ctx *ctx);
/* Turn per-kernel hipEvent timing on/off for this context (off by default). */
int fgpu_set_timing(fgpu_ctx *ctx, int on);
int fgpu_last_timing(const fgpu_ctx *ctx, fgpu_timing *out);
/* Device bytes currently held by the context's scratch arena. */
int64_t fgpu_scratch_bytes(const fgpu_ctx *ctx);

/* ── tables (Database::create_table_from_batches, src/db.rs:180-204) ─────── */
/* Copies `n_cols` host columns of `n_rows` rows to HBM.  `batch_rows` records the
 * reference's batch size (65,536, src/storage/table.rs:339) for batch-relative
 * semantics; the device layout is one contiguous array per column.  Upload also
 * computes per-column statistics (min/max, NULL count, float exponent range,
 * approximate distinct count) used to choose kernels. */
int fgpu_table_upload(fgpu_ctx *ctx, const fgpu_column_desc *cols, int n_cols,
                      int64_t n_rows, int32_t batch_rows, fgpu_table **out);
/* A Utf8 column delivered as record batches with one dictionary each, as the
 * CSV and Parquet loaders build them (src/storage/csv.rs:232-237,
 * src/storage/parquet_io.rs:194-207). */
typedef struct {
    int32_t n_batches;
    const int64_t *batch_rows;        /* rows of each batch, in row order; they sum to n_rows */
    const int32_t *dict_len;          /* entries of each batch's dictionary */
    const char *const *const *dicts;  /* each batch's dict_len[b] strings */
} fgpu_utf8_batches;

/* fgpu_table_upload for tables whose Utf8 columns carry one dictionary per
 * batch.  For a FGPU_UTF8 column i with utf8[i].n_batches > 0, cols[i].values
 * holds every batch's local codes in row order and cols[i].dict is ignored.
 * The dictionaries are merged the way Utf8Column::concat_all merges them
 * (src/storage/column.rs:1539-1595): batches that share one dictionary (the
 * same array, or equal entries) keep it; otherwise entries are numbered in
 * first-seen order over the batches' dictionaries in batch order, and each
 * batch's codes are remapped through its local-to-global map on the device.
 * A local code past its batch's dictionary fails with
 * FGPU_E_INDEX_OUT_OF_BOUNDS.  utf8 == NULL makes this fgpu_table_upload. */
int fgpu_table_upload_batches(fgpu_ctx *ctx, const fgpu_column_desc *cols, const fgpu_utf8_batches *utf8,
                              int n_cols, int64_t n_rows, int32_t batch_rows, fgpu_table **out);
void fgpu_table_free(fgpu_table *t);
int64_t fgpu_table_rows(const fgpu_table *t);
/* The table's schema (Schema, src/storage/schema.rs): column count, and per
 * column its name, dtype and (Utf8) dictionary; values / valid are NULL. */
int fgpu_table_cols(const fgpu_table *t);
int fgpu_table_column_info(const fgpu_table *t, int col, fgpu_column_view *out);
int64_t fgpu_table_device_bytes(const fgpu_table *t);
/* Column statistics computed at upload: min/max as int64 (integers, codes) or
 * double bits (floats); approximate distinct count. */
int fgpu_table_column_stats(const fgpu_table *t, int col, int64_t *min_i64,
                            int64_t *max_i64, int64_t *null_count, int64_t *distinct_est);

/* ── WHERE ───────────────────────────────────────────────────────────────── */
/* COUNT(*) ... WHERE p.  (db.rs:1439-1476 COUNT(*)-with-WHERE fast path:
 * Σ popcount(eval_mask_bits)) */
int fgpu_filter_count(fgpu_ctx *ctx, const fgpu_table *t, const fgpu_pred *p,
                      int64_t *out_count);
/* The packed mask: ceil(n/64) LSB-first words into host memory.
 * (Expr::eval_mask_bits, src/query/expr.rs:635; fast_cmp_mask :1076-1175;
 *  clear_null_bits :961-973) */
int fgpu_filter_mask(fgpu_ctx *ctx, const fgpu_table *t, const fgpu_pred *p,
                     uint64_t *host_words);
/* Materialise the selected rows of `cols` in row order (execute_select +
 * RecordBatch::filter_by_bitmask_cols, src/query/select.rs:105-160, :457-487;
 * src/storage/table.rs:172-189; gather_by_bitmask, src/storage/column.rs:1077-1123). */
int fgpu_filter_select(fgpu_ctx *ctx, const fgpu_table *t, const fgpu_pred *p,
                       const int32_t *cols, int n_cols, fgpu_result **out);

/* ── GROUP BY / scalar aggregates ────────────────────────────────────────── */
/* SELECT key, aggs... FROM t [WHERE p] GROUP BY key, fused in one pass.
 * (GroupByExec::execute, src/query/aggregate.rs:535-1420, preceded by the
 *  execute_select filter; key_col = -1 is execute_agg_only, :1394-1528).
 * Integer SUM/MIN/MAX are returned exactly as INT64 (the reference returns
 * Float64, Appendix semantics restated by the host layer); float SUM/AVG are
 * exactly rounded sums.  p may be NULL. */
int fgpu_groupby(fgpu_ctx *ctx, const fgpu_table *t, const fgpu_pred *p,
                 const fgpu_groupby_spec *spec, fgpu_result **out);

/* GROUP BY over key_cols[0..n_keys) (1..FGPU_MAX_GROUP_KEYS columns), the
 * general form of fgpu_groupby (spec->key_col is ignored).  One Int32 / Int64 /
 * Utf8 key is GroupByExec's fast path and runs exactly as fgpu_groupby.  Several
 * keys, or one Float64 key, are its slow path (GroupByExec::execute_slow /
 * accumulate_slow, src/query/aggregate.rs:1232-1346), which keys a group on the
 * display strings of its key values (GroupKey::from_row, :79-83): equal values
 * group, NULL is a group of its own (a Utf8 NULL joins the text "NULL"), every
 * NaN is one group and -0.0 is apart from 0.0.  Output: one column per key
 * (Int32 / Int64 as Int64, Float64, Utf8; NULL keys invalid), then the
 * aggregates.  The key columns' combinations must number at most 2^62 (Float64
 * keys only alone); otherwise FGPU_E_UNSUPPORTED. */
#define FGPU_MAX_GROUP_KEYS 4
int fgpu_groupby_keys(fgpu_ctx *ctx, const fgpu_table *t, const fgpu_pred *p, const int32_t *key_cols,
                      int32_t n_keys, const fgpu_groupby_spec *spec, fgpu_result **out);

/* GROUP BY that allocates nothing.  No counterpart in the reference: this is
 * for callers that cannot allocate device or pinned memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same result as fgpu_groupby for the strategies that need a bounded state:
 * no GROUP BY, or a key whose range takes a dense table (Utf8 keys by
 * dictionary code).  Every scratch buffer (counters, the row count, Utf8
 * predicate pass tables, the group state) lives in a static device array of
 * FGPU_STATIC_SCRATCH_BYTES per device; a larger state, or a key that needs a
 * hash table, fails with FGPU_E_UNSUPPORTED.  The outputs are the caller's
 * DEVICE buffers of `capacity` rows: out_key / out_key_valid (int64 keys,
 * Utf8 keys as codes; unused without a key) and, per aggregate i,
 * out_vals[i] (int64, or the bits of a double for SUM / AVG / MIN / MAX of a
 * Float64 column and AVG of an integer column) and out_valid[i] (1 = valid,
 * 0 = NULL; 2 marks an integer SUM whose total left int64, its value then
 * being the bits of the correctly rounded double).  If `capacity` is below
 * the rows the call may emit, *n_out receives that bound and
 * FGPU_E_INDEX_OUT_OF_BOUNDS is returned.  Calls on one device share its
 * array and are serialised by the library; the call returns with its stream
 * synchronised and *n_out set to the rows written. */
#define FGPU_STATIC_SCRATCH_BYTES (16 << 20)
int fgpu_groupby_static(fgpu_ctx *ctx, const fgpu_table *t, const fgpu_pred *p,
                        const fgpu_groupby_spec *spec, int64_t *out_key, uint8_t *out_key_valid,
                        int64_t *const *out_vals, uint8_t *const *out_valid, int64_t capacity,
                        int64_t *n_out);

/* ── hash join ───────────────────────────────────────────────────────────── */
/* Builds a device hash index over `key_col` of the build table
 * (build_hash_table / IntHashTable::insert, src/query/join.rs:607-645,
 *  :1054-1107, :1318-1324).  Reusable across probes and contexts (also
 *  concurrently).  The first partitioned join through an index adds the
 *  build side's partitioned sub-tables to it (P * T * 16 bytes, 64 MB for
 *  1M build keys), so later joins skip that build;
 *  fgpu_join_index_device_bytes counts them. */
int fgpu_join_build(fgpu_ctx *ctx, const fgpu_table *build, int32_t key_col,
                    fgpu_join_index **out);
void fgpu_join_index_free(fgpu_join_index *idx);
int64_t fgpu_join_index_device_bytes(const fgpu_join_index *idx);
/* SELECT aggs FROM probe [LEFT] JOIN build ON probe.k = build.k [WHERE probe_pred]
 * and the EXISTS / NOT EXISTS forms (SEMI / ANTI), fused with the aggregates.
 * (execute_hash_join_projected, src/query/join.rs:184-282 + execute_agg_only,
 *  src/query/aggregate.rs:1394; execute_hash_semi_anti_join, src/query/join.rs:49-132)
 * `idx` may be NULL (the index is then built and dropped inside the call); it
 * must have been built on spec->build_key of `build`.  Any number of
 * aggregates: past what one probe pass holds (four accumulators, or
 * FGPU_MAX_AGGS outputs) the call runs them in parts, one join each, and
 * returns their columns in order. */
int fgpu_join_agg(fgpu_ctx *ctx, const fgpu_table *probe, const fgpu_pred *probe_pred,
                  const fgpu_table *build, const fgpu_join_index *idx,
                  const fgpu_join_spec *spec, fgpu_result **out);
/* Pair mode: the (probe row, build row) pairs the join produces, NO_ROW =
 * 0xFFFFFFFF for a LEFT row without a match; SEMI/ANTI give the kept probe rows
 * with build row NO_ROW.  Pairs are ordered by probe row; equal probe rows list
 * build rows in ascending order.  (probe_one / collect_matches,
 * src/query/join.rs:657-723, :1328-1337; output_plan :828-899)  If `capacity` is
 * too small, *n_out receives the required count and FGPU_E_INDEX_OUT_OF_BOUNDS is
 * returned. */
int fgpu_join_pairs(fgpu_ctx *ctx, const fgpu_table *probe, const fgpu_pred *probe_pred,
                    const fgpu_table *build, const fgpu_join_index *idx,
                    const fgpu_join_spec *spec, uint32_t *probe_rows,
                    uint32_t *build_rows, int64_t capacity, int64_t *n_out);

/* The join's output itself: SELECT <probe_cols>, <build_cols> FROM probe
 * [LEFT] JOIN build ON probe.k = build.k [WHERE probe_pred], or for SEMI /
 * ANTI the kept probe rows (probe columns only), materialised as a new
 * device-resident table (execute_hash_join_projected, src/query/join.rs:184-282;
 * output_plan / build_output_dense :828-899, :403-502).  Columns are the kept
 * probe columns, then the kept build columns; a build column whose name the
 * probe table also has is named "right.<name>" (src/query/join.rs:877-886);
 * build columns of a LEFT join are NULL where no build row matched (NO_ROW).
 * With no columns kept the table still has the join's row count, so
 * COUNT(*) over it works (:891-897).  Rows come in probe order.  The table is
 * an ordinary fgpu_table (GROUP BY, ORDER BY, WHERE over the join output);
 * the caller frees it with fgpu_table_free.  The spec's aggregates are
 * ignored. */
int fgpu_join_materialize(fgpu_ctx *ctx, const fgpu_table *probe, const fgpu_pred *probe_pred,
                          const fgpu_table *build, const fgpu_join_index *idx, const fgpu_join_spec *spec,
                          const int32_t *probe_cols, int n_probe_cols, const int32_t *build_cols,
                          int n_build_cols, fgpu_table **out);

/* ── ORDER BY col [ASC | DESC] LIMIT k (execute_sort_limit,
 * src/query/sort.rs:59-115; typed top-k :209-330) ───────────────────────────
 * rows_out receives the indices of the k rows in output order: the column's
 * total order (f64_ord_key, :452-454) reversed for DESC, ties by row, NULL
 * rows last for DESC and first for ASC.  vals_out (may be NULL) receives the
 * sort column's values as int64 / double bit patterns (0 for NULL rows).
 * k <= 32 on the device; larger limits return FGPU_E_UNSUPPORTED. */
int fgpu_topk(fgpu_ctx *ctx, const fgpu_table *t, const fgpu_pred *p, int32_t col, int desc,
              int32_t k, int64_t *rows_out, int64_t *vals_out, int64_t *n_out);

/* ORDER BY key [ASC | DESC], ... [LIMIT limit] for any limit and any number
 * of keys (execute_sort_limit, src/query/sort.rs:59-115; try_typed_sort
 * :209-330; compare_scalar :807-840).  Each key orders by its column's total
 * order (integers, doubles by f64_ord_key :451-454, Utf8 by string, i.e. by
 * the code's rank in the sorted dictionary), reversed for DESC, with NULL rows
 * last when nulls_last is set and first otherwise (SortExpr::desc sets it,
 * ::asc does not, :28-34).  Rows equal on every key keep row order.  The
 * first `limit` rows (all when limit < 0) of the rows passing p come back
 * with columns `cols` in order; rows_out (may be NULL) receives their row
 * numbers.  Single-key LIMIT k <= 32 is also served by fgpu_topk. */
typedef struct {
    int32_t col;
    int32_t desc;
    int32_t nulls_last;
} fgpu_sort_key;
int fgpu_sort_limit(fgpu_ctx *ctx, const fgpu_table *t, const fgpu_pred *p, const fgpu_sort_key *keys,
                    int n_keys, int64_t limit, const int32_t *cols, int n_cols, int64_t *rows_out,
                    fgpu_result **out);

/* ── batched GROUP BY ─────────────────────────────────────────────────────
 * Many independent GROUP BY queries, each over its own device table, prepared
 * once and then run as one replay of a captured hipGraph: a persistent
 * pass-1 kernel over (query, row range) units and a persistent pass-2 kernel
 * over (query, key range) units.  Each query's result is the one
 * fgpu_groupby gives for it (GroupByExec::execute,
 * src/query/aggregate.rs:535-1420); the prepare / run split follows the
 * reference's streaming operator shape (pipeline::Operator process / flush,
 * src/query/pipeline.rs:46-72).  Integer keys with a dense range take the
 * batch kernels; other queries run as plain fgpu_groupby calls inside
 * fgpu_batch_run (fgpu_batch_fallbacks counts them).  The batch copies the
 * predicate and spec arrays; the tables must outlive the batch. */
typedef struct fgpu_batch fgpu_batch;
typedef struct {
    const fgpu_table *table;
    const fgpu_pred *pred;           /* NULL = no WHERE */
    const fgpu_groupby_spec *spec;
} fgpu_batch_query;

int fgpu_batch_prepare(fgpu_ctx *ctx, const fgpu_batch_query *queries, int n, fgpu_batch **out);
/* Times up to max_trials configurations (workgroup size, rows per unit, LDS
 * table size = key-range fan-out) on the batch itself and keeps the fastest. */
int fgpu_batch_tune(fgpu_batch *b, int max_trials);
/* outs[i] receives query i's result (free each with fgpu_result_free). */
int fgpu_batch_run(fgpu_batch *b, fgpu_result **outs);
const char *fgpu_batch_info(const fgpu_batch *b);  /* chosen configuration, as text */
int fgpu_batch_fallbacks(const fgpu_batch *b);
void fgpu_batch_free(fgpu_batch *b);  /* before fgpu_close of its context */

/* ── serving: continuous batching of mixed requests ───────────────────────
 * A request server over device-resident tables: the concurrent form of
 * Database::sql (src/db.rs:242; Database is Sync and the REST server runs one
 * sql call per worker, src/bin/server.rs:181-203).  The server owns n_slots
 * contexts (one HIP stream each) and one worker thread per slot.  Requests
 * queue in arrival order; a free slot admits the oldest request whose scratch
 * memory fits what the running requests leave of the HBM budget (smaller ones
 * may pass a waiting large one), so requests that arrive while others run
 * start as soon as a slot and memory free up, and their kernels overlap on
 * the device.  COUNT(*) requests that test one numeric column of the same
 * table with one comparison ride a single shared scan (up to 32 per launch).
 * A request's scratch need is learned per request shape from its first run
 * (first guess: a bound from the row counts); a run that outgrows its
 * reservation is re-queued with a larger one.  Tables and join indexes are
 * the caller's, outside the budget, and must outlive the requests on them.
 * Results equal what the plain entry points give for the same request. */
typedef struct fgpu_server fgpu_server;
enum {
    FGPU_REQ_COUNT = 0,    /* COUNT(*) FROM table WHERE pred (src/db.rs:1439-1476) */
    FGPU_REQ_GROUPBY = 1,  /* fgpu_groupby(table, pred, groupby) */
    FGPU_REQ_JOIN = 2,     /* fgpu_join_agg(table = probe, pred, build, index, join); index required */
    FGPU_REQ_TOPK = 3      /* fgpu_topk(table, pred, topk_col, topk_desc, topk_k) */
};
typedef struct {
    int32_t kind;                       /* FGPU_REQ_* */
    const fgpu_table *table;            /* FROM (probe) table */
    const fgpu_pred *pred;              /* WHERE on `table`; NULL = none */
    const fgpu_groupby_spec *groupby;   /* FGPU_REQ_GROUPBY */
    const fgpu_table *build;            /* FGPU_REQ_JOIN */
    const fgpu_join_index *index;       /* FGPU_REQ_JOIN */
    const fgpu_join_spec *join;         /* FGPU_REQ_JOIN */
    int32_t topk_col, topk_desc, topk_k;  /* FGPU_REQ_TOPK */
    const char *alias;                  /* FGPU_REQ_COUNT output name; NULL = "col0" */
} fgpu_request;

/* Per-request record, filled by fgpu_server_wait. */
typedef struct {
    double queue_ms;      /* submit -> admitted to a slot */
    double run_ms;        /* admitted -> result ready */
    double latency_ms;    /* submit -> result ready */
    int64_t rows_in;      /* input rows the request reads (probe + build rows for joins) */
    int64_t reserved_bytes;  /* HBM reserved for it at admission */
    int32_t slot;         /* slot it ran on */
    int32_t shared;       /* requests in the launch it rode in (1 = alone) */
    int32_t retries;      /* re-admissions after outgrowing its reservation */
} fgpu_request_stats;

/* Server-wide counters. */
typedef struct {
    int64_t submitted, completed, failed;
    int64_t shared_launches;    /* shared COUNT(*) scans */
    int64_t shared_requests;    /* requests answered by them */
    int64_t retries;            /* budget re-queues */
    int64_t trims;              /* idle slot arenas released to admit a request */
    int64_t peak_held_bytes;    /* most HBM the slots held at once (<= budget) */
    int32_t peak_running;       /* most requests running at once */
    int32_t n_slots;
    int64_t budget_bytes;
} fgpu_server_stats;

/* flags for fgpu_server_open.  FGPU_SERVE_SHARE_COUNTS: COUNT(*) requests that
 * test one numeric column of the same table with one comparison ride one
 * shared scan.  Without FGPU_SERVE_OVERLAP_ALWAYS a request starts beside a
 * running one only while at least two requests wait (under light load the
 * slots run requests one after the other, which gives the lower tail
 * latency); with it, any free slot starts the oldest request that fits. */
#define FGPU_SERVE_SHARE_COUNTS 1
#define FGPU_SERVE_OVERLAP_ALWAYS 2
int fgpu_server_open(int device, int n_slots, int64_t hbm_budget_bytes, int flags,
                     fgpu_server **out);
/* Copies the request (predicate, specs and strings) and queues it; returns at once. */
int fgpu_server_submit(fgpu_server *s, const fgpu_request *req, uint64_t *ticket);
/* Blocks until the request is done.  Returns its status; on success *out is the
 * result (free with fgpu_result_free).  stats may be NULL.  Each ticket is
 * waited for once. */
int fgpu_server_wait(fgpu_server *s, uint64_t ticket, fgpu_result **out, fgpu_request_stats *stats);
/* on = 1 holds admissions (queued requests wait; running ones finish); 0 resumes.
 * Closing a paused server still finishes its queue. */
int fgpu_server_pause(fgpu_server *s, int on);
int fgpu_server_get_stats(fgpu_server *s, fgpu_server_stats *out);
/* Finishes the queued requests, then stops the workers and frees the slots.
 * Results nobody waited for are freed. */
void fgpu_server_close(fgpu_server *s);

/* ── results (QueryResult{batches}, src/db.rs:25-27) ─────────────────────── */
int64_t fgpu_result_rows(const fgpu_result *r);
int fgpu_result_cols(const fgpu_result *r);
int fgpu_result_column(const fgpu_result *r, int i, fgpu_column_view *out);
void fgpu_result_free(fgpu_result *r);

#ifdef __cplusplus
}
#endif

#endif /* FALCON_GPU_H */
