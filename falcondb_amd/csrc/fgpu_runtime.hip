// fgpu_runtime.hip — host side of libfalcon_gpu.so: contexts, device-resident
// tables, plan compilation (predicates, aggregate state layouts, strategy
// choice) and the C ABI declared in include/falcon_gpu.h.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <strings.h>
#include <list>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "fgpu_host.h"
#include "fgpu_launch.h"

using namespace fgpu;

// ─── static scratch of fgpu_groupby_static ──────────────────────────────────
// fgpu_groupby_static keeps every buffer its kernels need beyond the inputs
// and outputs (counters, the output row count, Utf8 predicate pass tables and
// the group state table) in this array, so that it allocates nothing.  Each
// device holds its own copy for the life of the process.  In host code the
// name `scratch` is a stand-in, not a device address: host code only ever
// uses the address hipGetSymbolAddress returns for the current device
// (g_static[device].ptr, looked up once per device by fgpu_open).
constexpr size_t STATIC_SCRATCH_BYTES = size_t(16) << 20;
static __device__ __align__(256) uint64_t scratch[STATIC_SCRATCH_BYTES / 8];

namespace {
constexpr int STATIC_MAX_DEVICES = 64;
struct StaticScratch {
    void *ptr = nullptr;       // this device's copy of `scratch`
    hipError_t lookup = hipSuccess;
    bool looked_up = false;
    std::mutex use;            // held for a whole fgpu_groupby_static call
};
StaticScratch g_static[STATIC_MAX_DEVICES];
std::mutex g_static_init;

// Looks up the device address of `scratch` for `device` (current) once.
void static_scratch_init(int device) {
    if (device < 0 || device >= STATIC_MAX_DEVICES) return;
    std::lock_guard<std::mutex> lk(g_static_init);
    StaticScratch &st = g_static[device];
    if (st.looked_up) return;
    void *ptr = nullptr;
    st.lookup = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    st.ptr = st.lookup == hipSuccess ? ptr : nullptr;
    st.looked_up = true;
}
}  // namespace

// ─── errors ─────────────────────────────────────────────────────────────────
namespace {
thread_local std::string g_err;

int set_err(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) {                                                        \
            const int c_ = e_ == hipErrorOutOfMemory ? FGPU_E_OUT_OF_MEMORY : FGPU_E_DEVICE; \
            return set_err(c_, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
        }                                                                              \
    } while (0)

int64_t next_pow2(int64_t x) {
    int64_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

// Inverse of hash64 (each step of the finaliser is invertible: xorshifts by
// 33 >= 32 bits undo themselves, and odd multipliers have inverses mod 2^64).
uint64_t inv_odd(uint64_t c) {
    uint64_t x = c;  // Newton's iteration doubles the correct low bits each step
    for (int i = 0; i < 6; ++i) x *= 2 - c * x;
    return x;
}
uint64_t unhash64(uint64_t h) {
    h ^= h >> 33;
    h *= inv_odd(0xc4ceb9fe1a85ec53ull);
    h ^= h >> 33;
    h *= inv_odd(0xff51afd7ed558ccdull);
    h ^= h >> 33;
    return h;
}

int env_int(const char *name, int dflt) {
    const char *v = std::getenv(name);
    return v && *v ? std::atoi(v) : dflt;
}
}  // namespace

// ─── device column / table ──────────────────────────────────────────────────
struct DevCol {
    std::string name;
    int dtype = 0;
    void *d_values = nullptr;
    uint64_t *d_valid = nullptr;
    int64_t null_count = 0;
    std::vector<std::string> dict;
    std::vector<const char *> dictp;
    int64_t empty_code = -1;  // dictionary index of "" (Utf8)
    // statistics
    int64_t imin = 0, imax = -1;
    double fmin = 0, fmax = 0;
    int emin = 0, emax = -1;  // float mantissa-LSB exponents (emin > emax: no finite nonzero)
    uint64_t nonfinite = 0;
    int64_t distinct = 0;
    std::vector<BatchZone> zones;  // per batch_rows batch (numeric columns): the zone map
    size_t elem_bytes() const { return dtype == FGPU_INT64 || dtype == FGPU_FLOAT64 ? 8 : 4; }
};

struct fgpu_table {
    int device = 0;
    int64_t n_rows = 0;
    int32_t batch_rows = 65536;
    std::vector<DevCol> cols;
    int64_t bytes = 0;
};

// Radix-partitioned sub-tables of the build side for one (fan-out, table
// size, build column held in the slots), built by the first partitioned join
// that needs them and kept by the index.  sub == nullptr records that these
// keys cannot be partitioned so (a sub-table would overflow).
struct JoinSubTables {
    int log_p = 0, T = 0;
    int kw = 8;                     // record format the slots hold (hash64 or compact keys)
    const void *val_col = nullptr;  // build column whose values the slots hold; nullptr: build rows
    int64_t *sub = nullptr;         // P * T * 2 words: {record key, value or build row}
    uint32_t *min_key_rows = nullptr;
    int64_t n_min_key_rows = 0;
    int unique_keys = 0;
};

struct fgpu_join_index {
    int device = 0;
    const fgpu_table *build = nullptr;
    int key_col = -1;
    int64_t cap = 0;
    int64_t *slots = nullptr;  // 2 * cap words: {key, build row}
    uint32_t *min_key_rows = nullptr;
    int64_t n_min_key_rows = 0;
    int unique_keys = 0;
    mutable int64_t bytes = 0;
    mutable std::mutex sub_mu;                 // guards subs (joins on many contexts share an index)
    mutable std::list<JoinSubTables> subs;     // a list: entries stay put while others are added
};

// ─── results ────────────────────────────────────────────────────────────────
// Vectors that leave new elements uninitialised: result columns are filled by
// copies right away, so value-initialising them first is wasted host time.
template <class T>
struct NoInitAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = NoInitAlloc<U>;
    };
    NoInitAlloc() = default;
    template <class U>
    NoInitAlloc(const NoInitAlloc<U> &) {}
    template <class U>
    void construct(U *p) noexcept {
        ::new ((void *)p) U;
    }
    template <class U, class... Args>
    void construct(U *p, Args &&...args) {
        ::new ((void *)p) U(std::forward<Args>(args)...);
    }
};
template <class T>
using RVec = std::vector<T, NoInitAlloc<T>>;

struct RCol {
    std::string name;
    int dtype = FGPU_INT64;
    RVec<int64_t> i64;
    RVec<double> f64;
    RVec<uint32_t> u32;
    RVec<uint8_t> valid;
    const DevCol *dict_src = nullptr;
    // large results: values and validity live in the result's pinned block
    const void *xv = nullptr;
    const uint8_t *xvalid = nullptr;
};
// Pinned host blocks for result columns, kept by the process after a result
// is freed: the next result of a similar size is copied by DMA into memory
// that is already pinned and faulted in (a fresh allocation costs a page fault
// per 4 KB on first touch, more than the copy itself).
namespace {
std::mutex g_pin_mu;
std::multimap<size_t, void *> g_pin_free;
size_t g_pin_free_bytes = 0;
constexpr size_t PIN_KEEP_BYTES = size_t(1) << 30;  // at most this much is kept unused

void *pin_get(size_t bytes, size_t *got) {
    {
        std::lock_guard<std::mutex> lk(g_pin_mu);
        auto it = g_pin_free.lower_bound(bytes);
        if (it != g_pin_free.end() && it->first <= 2 * bytes + (size_t(4) << 20)) {
            void *p = it->second;
            *got = it->first;
            g_pin_free_bytes -= it->first;
            g_pin_free.erase(it);
            return p;
        }
    }
    const size_t n = (bytes + (size_t(2) << 20) - 1) / (size_t(2) << 20) * (size_t(2) << 20);
    void *p = nullptr;
    if (hipHostMalloc(&p, n, hipHostMallocDefault) != hipSuccess) return nullptr;
    *got = n;
    return p;
}
void pin_put(void *p, size_t n) {
    if (!p) return;
    std::unique_lock<std::mutex> lk(g_pin_mu);
    if (g_pin_free_bytes + n > PIN_KEEP_BYTES) {
        lk.unlock();
        (void)hipHostFree(p);
        return;
    }
    g_pin_free.emplace(n, p);
    g_pin_free_bytes += n;
}
}  // namespace

struct fgpu_result {
    int64_t rows = 0;
    std::vector<RCol> cols;
    void *block = nullptr;  // pinned storage of the columns that point into it
    size_t block_n = 0;
    ~fgpu_result() { pin_put(block, block_n); }
};

// ─── context ────────────────────────────────────────────────────────────────
enum BufSlot {
    B_KEYS, B_STATE, B_COUNTERS, B_OUT, B_MASK, B_OFFS, B_TMP, B_PASS, B_PAIRS, B_CNT,
    B_CELLS, B_CELLCNT, B_SUB, B_SUBCNT, B_MINKEY, B_NULLBUF, B_TOPK, B_LIVE, B_SORT, B_GKEY, B_NSLOTS
};

struct Tuning {
    int zone_maps = 1;             // WHERE scans skip batches whose bounds rule them out
    int gb_block = 256;
    int gb_waves_per_cu = 16;      // target resident waves per CU for streaming kernels
    int gb_replicas = 4;           // dense-LDS / scalar GROUP BY: HBM state replicas
    int hash_lds_slots = 1024;
    int probe_limit = 8;
    int dense_lds_bytes = 48 * 1024;
    int join_block = 256;
    int join_waves_per_cu = 16;
    int filter_waves_per_cu = 32;
    int keep_on_device = 0;  // 1: results stay in device scratch (only the row count is copied)
    int join_partition = 1;        // radix-partitioned join for large probes
    int join_part_min_rows = 1 << 21;
    int gb_partition = 1;          // radix-partitioned GROUP BY for high-cardinality keys
    int gb_part_min_rows = 1 << 20;
    int gbp_cache_slots = 8192;    // cap on the pass-1 LDS cache slots (8-way buckets)
    int gbp_probe_limit = 8;
    int gbp_table_slots = 4096;    // cap on the pass-2 LDS table slots
    int gbp_partitions = 0;        // radix fan-out; 0 = from the key's distinct estimate
    int gbp_cells_max_mb = 16384;  // bound on the pass-1 cell buffer
    int gbp_scan_block = 1024;
    int gbp_agg_block = 1024;
    int gbp_scan_wgs_per_cu = 1;
    int lds_limit = 160 * 1024;    // LDS bytes one workgroup may use
    int join_partitions = 0;       // least radix partitions of the build / probe sides (0: from the build rows)
    int join_part_wgs_per_cu = 2;  // pass-1 workgroups per CU (two fit by registers)
    int join_probe_wgs_per_cu = 4; // pass-2 workgroups per CU
    int join_slots_per_key = 4;    // pass-2 LDS sub-table slots per build key (load factor 1 / this)
    int join_compact = 1;          // compact record keys: 0 never, 1 for key-only records, 2 always
    int gbp_pack = 1;              // partitioned and batched GROUP BY: key and an integer input in one record word
    int batch_unit_rows = 0;       // batched GROUP BY: rows per pass-1 unit before tuning (0: the default)
    int batch_scan_block = 0;      // batched GROUP BY: pass-1 workgroup size (0: the batch's tuned block)
};

struct fgpu_ctx {
    int device = 0;
    int num_cus = 256;
    hipStream_t stream = nullptr;
    struct Buf {
        void *p = nullptr;
        size_t n = 0;
    } bufs[B_NSLOTS];
    bool timing = false;
    hipEvent_t ev[6] = {};
    fgpu_timing last{};
    Tuning tune;
    std::string plan;  // strategy of the last call (diagnostics)
    std::vector<uint8_t> live_host;  // zone-map flags being copied to B_LIVE (calls end synchronised)
    uint64_t *mail = nullptr;        // MAIL_BYTES of pinned host memory for small read-backs
    std::mutex mu;
    // HBM budget of a serving slot: the arena may not grow past arena_limit
    // bytes (0 = unbounded).  call_req[s] is the largest size buffer s was
    // asked for since the tally was last reset: the exact scratch a request
    // needs, which the server learns per request shape.
    int64_t arena_limit = 0;
    size_t call_req[B_NSLOTS] = {};
    bool limit_hit = false;
};

// Small device-to-host reads (flags, counts) go through the context's pinned
// mailbox: a copy into pageable memory is staged by the runtime, one into
// pinned memory is a direct DMA.  mail_copy issues the copy on the context's
// stream; mail_take waits for it and hands the bytes over.
constexpr size_t MAIL_BYTES = 512;
static hipError_t mail_copy(fgpu_ctx *c, const void *src, size_t bytes) {
    if (bytes > MAIL_BYTES) return hipErrorInvalidValue;
    return hipMemcpyAsync(c->mail, src, bytes, hipMemcpyDeviceToHost, c->stream);
}
static hipError_t mail_take(fgpu_ctx *c, void *dst, size_t bytes) {
    const hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) memcpy(dst, c->mail, bytes);
    return e;
}

namespace {

size_t arena_want(size_t bytes) { return std::max<size_t>(bytes + bytes / 4, 4096); }

int buf(fgpu_ctx *c, BufSlot s, size_t bytes, void **out) {
    auto &b = c->bufs[s];
    c->call_req[s] = std::max(c->call_req[s], bytes);
    if (b.n < bytes) {
        const size_t want = arena_want(bytes);
        if (c->arena_limit) {
            // Buffers this call has not asked for go first; then the limit holds.
            size_t others = 0;
            for (int i = 0; i < B_NSLOTS; ++i)
                if (i != s) others += c->bufs[i].n;
            if (others + want > (size_t)c->arena_limit) {
                HIPCHK(hipStreamSynchronize(c->stream));
                for (int i = 0; i < B_NSLOTS; ++i) {
                    auto &o = c->bufs[i];
                    if (i == s || c->call_req[i] || !o.p) continue;
                    HIPCHK(hipFree(o.p));
                    others -= o.n;
                    o.p = nullptr;
                    o.n = 0;
                }
            }
            if (others + want > (size_t)c->arena_limit) {
                c->limit_hit = true;
                return set_err(FGPU_E_OUT_OF_MEMORY, "HBM budget: %zu scratch bytes needed, slot limit %lld",
                               others + want, (long long)c->arena_limit);
            }
        }
        if (b.p) {
            HIPCHK(hipStreamSynchronize(c->stream));
            HIPCHK(hipFree(b.p));
            b.p = nullptr;
            b.n = 0;
        }
        HIPCHK(hipMalloc(&b.p, want));
        b.n = want;
    }
    *out = b.p;
    return FGPU_OK;
}

int set_device(int dev) {
    HIPCHK(hipSetDevice(dev));
    return FGPU_OK;
}

// Events: 0 start, 1 main begin, 2 main end, 3 end, 4 build begin, 5 build end.
void mark(fgpu_ctx *c, int i) {
    if (c->timing) (void)hipEventRecord(c->ev[i], c->stream);
}
void finish_timing(fgpu_ctx *c, int64_t bytes_main, bool has_build) {
    if (!c->timing) return;
    (void)hipEventSynchronize(c->ev[3]);
    float t = 0;
    c->last = fgpu_timing{};
    if (hipEventElapsedTime(&t, c->ev[0], c->ev[3]) == hipSuccess) c->last.total_ms = t;
    if (hipEventElapsedTime(&t, c->ev[1], c->ev[2]) == hipSuccess) c->last.main_ms = t;
    if (has_build && hipEventElapsedTime(&t, c->ev[4], c->ev[5]) == hipSuccess) c->last.build_ms = t;
    if (hipEventElapsedTime(&t, c->ev[2], c->ev[3]) == hipSuccess) c->last.finalize_ms = t;
    c->last.bytes_main = bytes_main;
}

int grid_for(const fgpu_ctx *c, int64_t work_items, int block, int waves_per_cu) {
    const int64_t need = std::max<int64_t>(1, (work_items + block - 1) / block);
    const int64_t resident = (int64_t)c->num_cus * std::max(1, waves_per_cu * 64 / block);
    return (int)std::max<int64_t>(1, std::min(need, resident));
}

// Maps table columns to kernel column slots.
struct Binder {
    const fgpu_table *t;
    DCols dc{};
    int slot_of[256];
    explicit Binder(const fgpu_table *tt) : t(tt) {
        dc.n = 0;
        for (int &x : slot_of) x = -1;
    }
    int bind(int col) {
        if (col < 0 || col >= (int)t->cols.size()) return -1;
        if (col < 256 && slot_of[col] >= 0) return slot_of[col];
        if (dc.n >= MAX_COLS) return -2;
        const DevCol &c = t->cols[col];
        const int s = dc.n++;
        dc.v[s] = c.d_values;
        dc.valid[s] = c.d_valid;
        dc.dtype[s] = (int8_t)c.dtype;
        if (col < 256) slot_of[col] = s;
        return s;
    }
};

// SQL LIKE over bytes: '%' matches any run of bytes, '_' exactly one byte,
// every other byte itself; no escape character (like_matches,
// src/query/expr.rs:1180-1199).  Greedy with one backtrack point, which
// decides the same language as the recursive definition.
static bool like_match(const std::string &s, const std::string &p) {
    size_t i = 0, j = 0, star = std::string::npos, mark = 0;
    while (i < s.size()) {
        if (j < p.size() && p[j] == '%') {
            star = j++;
            mark = i;
        } else if (j < p.size() && (p[j] == '_' || p[j] == s[i])) {
            ++i;
            ++j;
        } else if (star != std::string::npos) {
            j = star + 1;
            i = ++mark;
        } else {
            return false;
        }
    }
    while (j < p.size() && p[j] == '%') ++j;
    return j == p.size();
}

static std::string ascii_lower(std::string s) {
    for (char &ch : s)
        if (ch >= 'A' && ch <= 'Z') ch = (char)(ch + 32);
    return s;
}

// Bytes a predicate node's device table takes (a Utf8 pass table, one byte
// per dictionary entry, or a numeric IN list, eight bytes per literal), or 0.
static size_t pred_table_bytes(const fgpu_table *t, const fgpu_pred_node &nd) {
    if (nd.col < 0 || nd.col >= (int)t->cols.size()) return 0;
    const DevCol &col = t->cols[nd.col];
    size_t b = 0;
    if (col.dtype == FGPU_UTF8 && ((nd.kind == FGPU_P_CMP && nd.lit_type == FGPU_UTF8) || nd.kind == FGPU_P_LIKE ||
                                   (nd.kind == FGPU_P_IN && nd.lit_type == FGPU_UTF8)))
        b = col.dict.size();
    else if (nd.kind == FGPU_P_IN && col.dtype != FGPU_UTF8 && nd.n_list > 0)
        b = (size_t)nd.n_list * 8;
    return (b + 255) & ~size_t(255);
}

// Compiles a predicate (postfix program) into device form.  Utf8 pass tables
// and IN lists go to the context's arena, or, when pass_dst is given, to
// pass_dst (at most pass_cap bytes; *pass_used receives the bytes taken).
int compile_pred(fgpu_ctx *c, const fgpu_table *t, const fgpu_pred *p, Binder &b, DPred &out,
                 uint8_t *pass_dst = nullptr, size_t pass_cap = 0, size_t *pass_used = nullptr) {
    if (pass_used) *pass_used = 0;
    out = DPred{};
    out.n = 0;
    out.n_pcols = 0;
    if (!p || p->n_nodes == 0) return FGPU_OK;
    if (p->n_nodes > FGPU_MAX_PRED_NODES)
        return set_err(FGPU_E_INVALID_QUERY, "predicate has %d nodes (max %d)", p->n_nodes, FGPU_MAX_PRED_NODES);
    // dictionary pass tables live in one device buffer
    size_t pass_bytes = 0;
    for (int i = 0; i < p->n_nodes; ++i) pass_bytes += pred_table_bytes(t, p->nodes[i]);
    uint8_t *d_pass = nullptr;
    std::vector<uint8_t> h_pass(pass_bytes);
    if (pass_bytes && pass_dst) {
        if (pass_bytes > pass_cap)
            return set_err(FGPU_E_UNSUPPORTED, "predicate pass tables need %zu bytes (%zu available)", pass_bytes,
                           pass_cap);
        d_pass = pass_dst;
        if (pass_used) *pass_used = pass_bytes;
    } else if (pass_bytes) {
        int rc = buf(c, B_PASS, pass_bytes, (void **)&d_pass);
        if (rc) return rc;
    }
    size_t pass_off = 0;
    int depth = 0;
    for (int i = 0; i < p->n_nodes; ++i) {
        const fgpu_pred_node &nd = p->nodes[i];
        DPredNode &o = out.nodes[i];
        o = DPredNode{};
        o.kind = (int8_t)nd.kind;
        if (nd.kind == FGPU_P_AND || nd.kind == FGPU_P_OR) {
            if (depth < 2) return set_err(FGPU_E_INVALID_QUERY, "AND/OR needs two operands");
            --depth;
            continue;
        }
        if (nd.kind != FGPU_P_CMP && nd.kind != FGPU_P_IS_NULL && nd.kind != FGPU_P_IS_NOT_NULL &&
            nd.kind != FGPU_P_IN && nd.kind != FGPU_P_LIKE)
            return set_err(FGPU_E_INVALID_QUERY, "unknown predicate node kind %d", nd.kind);
        if (nd.col < 0 || nd.col >= (int)t->cols.size())
            return set_err(FGPU_E_COLUMN_NOT_FOUND, "predicate column %d not in table", nd.col);
        const int slot = b.bind(nd.col);
        if (slot < 0) return set_err(FGPU_E_UNSUPPORTED, "query reads more than %d columns", MAX_COLS);
        o.col = (int8_t)slot;
        int pc = 0;
        while (pc < out.n_pcols && out.pcol[pc] != slot) ++pc;
        if (pc == out.n_pcols) {
            if (out.n_pcols == MAX_PCOLS)
                return set_err(FGPU_E_UNSUPPORTED, "a predicate may read at most %d distinct columns", MAX_PCOLS);
            out.pcol[out.n_pcols++] = (int8_t)slot;
        }
        o.pc = (int8_t)pc;
        ++depth;
        if (nd.kind == FGPU_P_IS_NULL) { o.cls = CMP_IS_NULL; continue; }
        if (nd.kind == FGPU_P_IS_NOT_NULL) { o.cls = CMP_IS_NOT_NULL; continue; }
        if (nd.kind == FGPU_P_LIKE) {
            const DevCol &col = t->cols[nd.col];
            if (col.dtype != FGPU_UTF8) return set_err(FGPU_E_UNSUPPORTED, "LIKE on a column that is not Utf8");
            if (nd.op < 0 || nd.op > 3) return set_err(FGPU_E_INVALID_QUERY, "bad LIKE flags %d", nd.op);
            const bool neg = nd.op & 1, fold = nd.op & 2;
            const std::string pat = fold ? ascii_lower(nd.lit_str ? nd.lit_str : "") : std::string(nd.lit_str ? nd.lit_str : "");
            for (size_t k = 0; k < col.dict.size(); ++k)
                h_pass[pass_off + k] = like_match(fold ? ascii_lower(col.dict[k]) : col.dict[k], pat) != neg;
            o.cls = CMP_DICT;
            o.op = neg ? 1 : 0;  // a NULL row passes NOT LIKE (src/query/expr.rs:732-736)
            o.pass = d_pass + pass_off;
            pass_off += pred_table_bytes(t, nd);
            continue;
        }
        if (nd.kind == FGPU_P_IN) {
            if (nd.op != 0 && nd.op != 1) return set_err(FGPU_E_INVALID_QUERY, "bad IN flags %d", nd.op);
            if (nd.n_list < 0 || (nd.n_list > 0 && !nd.list))
                return set_err(FGPU_E_INVALID_QUERY, "IN list of %d literals at NULL", nd.n_list);
            const bool neg = nd.op == 1;
            const DevCol &col = t->cols[nd.col];
            const int lt = nd.lit_type;
            const bool lit_int = lt == FGPU_INT64 || lt == FGPU_INT32;
            if (col.dtype == FGPU_UTF8 && lt == FGPU_UTF8) {
                std::vector<uint8_t> member(col.dict.size(), 0);
                const char *const *lits = (const char *const *)nd.list;
                for (size_t k = 0; k < col.dict.size(); ++k)
                    for (int m = 0; m < nd.n_list && !member[k]; ++m)
                        member[k] = std::strcmp(col.dict[k].c_str(), lits[m] ? lits[m] : "") == 0;
                for (size_t k = 0; k < col.dict.size(); ++k) h_pass[pass_off + k] = member[k] != neg;
                o.cls = CMP_DICT;
                o.op = 0;  // NULL never passes IN, and NOT IN drops the NULL rows
                o.pass = d_pass + pass_off;
                pass_off += pred_table_bytes(t, nd);
                continue;
            }
            // the `=` pairing of each literal with the column (fast_cmp_mask)
            int flavor = -1;
            if (col.dtype == FGPU_FLOAT64 && (lt == FGPU_FLOAT64 || lit_int)) flavor = IN_F64;
            else if (col.dtype == FGPU_INT64 && lit_int) flavor = IN_I64;
            else if (col.dtype == FGPU_INT64 && lt == FGPU_FLOAT64) flavor = IN_I64_AS_F64;
            else if (col.dtype == FGPU_INT32 && lit_int) flavor = IN_I32;
            if (flavor < 0 || nd.n_list == 0) {
                // no member can match: IN passes nothing, NOT IN every non-NULL row
                o.cls = neg ? CMP_IS_NOT_NULL : CMP_FALSE;
                pass_off += pred_table_bytes(t, nd);
                continue;
            }
            int64_t n_vals = 0;
            int64_t *dst = (int64_t *)(h_pass.data() + pass_off);
            if (flavor == IN_I64 || flavor == IN_I32) {
                std::vector<int64_t> v((size_t)nd.n_list);
                for (int m = 0; m < nd.n_list; ++m) {
                    const int64_t x = ((const int64_t *)nd.list)[m];
                    v[m] = flavor == IN_I32 ? (int64_t)(int32_t)x : x;  // `as i32`
                }
                std::sort(v.begin(), v.end());
                v.erase(std::unique(v.begin(), v.end()), v.end());
                n_vals = (int64_t)v.size();
                std::memcpy(dst, v.data(), v.size() * 8);
            } else {
                std::vector<double> v;
                for (int m = 0; m < nd.n_list; ++m) {
                    const double x = lit_int ? (double)((const int64_t *)nd.list)[m] : ((const double *)nd.list)[m];
                    if (x == x) v.push_back(x == 0.0 ? 0.0 : x);  // NaN equals nothing; -0.0 == 0.0
                }
                std::sort(v.begin(), v.end());
                v.erase(std::unique(v.begin(), v.end()), v.end());
                n_vals = (int64_t)v.size();
                std::memcpy(dst, v.data(), v.size() * 8);
            }
            if (n_vals == 0) {
                o.cls = neg ? CMP_IS_NOT_NULL : CMP_FALSE;
            } else {
                o.cls = CMP_IN;
                o.op = flavor | (neg ? IN_NOT : 0);
                o.lit_i = n_vals;
                o.pass = d_pass + pass_off;
            }
            pass_off += pred_table_bytes(t, nd);
            continue;
        }
        if (nd.op < FGPU_EQ || nd.op > FGPU_GE) return set_err(FGPU_E_INVALID_QUERY, "bad comparison op %d", nd.op);
        o.op = (int8_t)nd.op;
        const DevCol &col = t->cols[nd.col];
        const int lt = nd.lit_type;
        const bool lit_int = lt == FGPU_INT64 || lt == FGPU_INT32;
        // fast_cmp_mask pairings (src/query/expr.rs:1111-1171)
        if (col.dtype == FGPU_FLOAT64 && (lt == FGPU_FLOAT64 || lit_int)) {
            o.cls = CMP_F64;
            o.lit_f = lt == FGPU_FLOAT64 ? nd.lit_f64 : (double)nd.lit_i64;
        } else if (col.dtype == FGPU_INT64 && lit_int) {
            o.cls = CMP_I64;
            o.lit_i = nd.lit_i64;
        } else if (col.dtype == FGPU_INT64 && lt == FGPU_FLOAT64) {
            o.cls = CMP_I64_AS_F64;
            o.lit_f = nd.lit_f64;
        } else if (col.dtype == FGPU_INT32 && lit_int) {
            o.cls = CMP_I32;
            o.lit_i = (int64_t)(int32_t)nd.lit_i64;  // `as i32`
        } else if (col.dtype == FGPU_UTF8 && lt == FGPU_UTF8) {
            o.cls = CMP_DICT;
            const char *lit = nd.lit_str ? nd.lit_str : "";
            for (size_t k = 0; k < col.dict.size(); ++k) {
                const int r = std::strcmp(col.dict[k].c_str(), lit);
                bool h = false;
                switch (nd.op) {
                    case FGPU_EQ: h = r == 0; break;
                    case FGPU_NE: h = r != 0; break;
                    case FGPU_LT: h = r < 0; break;
                    case FGPU_LE: h = r <= 0; break;
                    case FGPU_GT: h = r > 0; break;
                    default: h = r >= 0; break;
                }
                h_pass[pass_off + k] = h;
            }
            o.op = 0;  // a NULL code fails
            o.pass = d_pass + pass_off;
            pass_off += pred_table_bytes(t, nd);
        } else {
            o.cls = CMP_FALSE;  // unsupported pairing: no rows pass
        }
    }
    if (depth != 1) return set_err(FGPU_E_INVALID_QUERY, "predicate program does not reduce to one value");
    out.n = p->n_nodes;
    if (pass_bytes) {  // h_pass is a local: wait for the copy before it goes away
        HIPCHK(hipMemcpyAsync(d_pass, h_pass.data(), pass_bytes, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return FGPU_OK;
}

// ─── aggregate state layout ─────────────────────────────────────────────────
struct OpKey {
    int kind, side, col;
};

struct AggLayout {
    DOps ops{};
    std::vector<OpKey> keys;  // per op
    DOuts outs{};
    std::vector<std::string> names;
};

// Exact-sum layout of float column c: LSB exponent L and W words.  Exact for
// columns whose mantissa-LSB exponents span at most 364 binades (15 words).
// A wider column is refused unless `one_row` (the aggregate has one output
// row: no GROUP BY key, or a join), where the accumulator's LSB moves up to
// the largest L that MAX_FX_WORDS words still cover and every value is
// rounded to a multiple of 2^L before it is added (fx_digits): the sum is
// the exact sum of the rounded values, within n * 2^(L-1) of the exact sum,
// far below 1e-12 of any sum that does not cancel against the column's
// largest values.  Per group such a bound would be relative to the whole
// column, not to the group, so GROUP BY refuses those columns.
int fx_words(const DevCol &c, int &L, int &W, bool one_row) {
    if (c.emin > c.emax) {  // no finite nonzero value
        L = 0;
        W = 1;
        return FGPU_OK;
    }
    // digits of one value reach bit (emax - emin) + 52 + 31 (in-word offset);
    // a per-thread partial sum (< 2^31 addends) may be 31 bits larger still.
    L = c.emin;
    W = ((c.emax - c.emin) + 53 + 31 + 31) / 32 + 1;
    if (W > MAX_FX_WORDS) {
        if (!one_row)
            return set_err(FGPU_E_UNSUPPORTED,
                           "column '%s': float exponent range 2^%d..2^%d is too wide for an exact SUM per group "
                           "(%d words > %d)", c.name.c_str(), c.emin, c.emax + 52, W, MAX_FX_WORDS);
        L = c.emax - ((MAX_FX_WORDS - 1) * 32 - (53 + 31 + 31) + 31);
        W = ((c.emax - L) + 53 + 31 + 31) / 32 + 1;
        if (W > MAX_FX_WORDS) return set_err(FGPU_E_DEVICE, "exact SUM layout of '%s'", c.name.c_str());
    }
    return FGPU_OK;
}

// ceil(log2(n)) for n >= 1.
int ceil_log2(int64_t n) {
    int b = 0;
    while (b < 63 && (int64_t(1) << b) < n) ++b;
    return b;
}

// Words of an exact integer SUM of at most `rows` values of column c: one
// wrapping word when rows * max|v| < 2^63 proves that the total fits int64,
// else the two-word carry-save form (isum_total).
int isum_words(const DevCol &c, int64_t rows) {
    if (c.imin > c.imax || rows <= 0) return 1;  // no valid value
    const unsigned __int128 mag = (unsigned __int128)std::max<__int128>(-(__int128)c.imin, (__int128)c.imax);
    const unsigned __int128 bound = mag * (unsigned __int128)(uint64_t)rows;
    return bound < ((unsigned __int128)1 << 63) ? 1 : 2;
}

// Power-of-two prescale of a register float sum over at most `rows` values of
// column c: every value is below 2^(emax + 53), so the partial sums stay below
// 2^(emax + 53 + ceil(log2(rows))), which must stay below 2^1023.
int fsum_scale(const DevCol &c, int64_t rows) {
    if (c.emin > c.emax) return 0;
    const int top = c.emax + 53 + ceil_log2(std::max<int64_t>(rows, 1)) + 1;
    return top > 1023 ? top - 1023 : 0;
}

// Adds (or finds) an op; returns its index.  `col` is a table column index;
// `rows` bounds the number of values any one group's aggregate takes in.
int add_op(AggLayout &L, int kind, int side, int col, const DevCol &dc, int slot, int64_t rows, bool one_row) {
    for (size_t i = 0; i < L.keys.size(); ++i)
        if (L.keys[i].kind == kind && L.keys[i].side == side && L.keys[i].col == col) return (int)i;
    if (L.ops.n >= MAX_OPS) return set_err(FGPU_E_UNSUPPORTED, "too many aggregate states (max %d)", MAX_OPS) * 0 - 1;
    DOp o{};
    o.kind = kind;
    o.col = slot;
    o.word = L.ops.S;
    int nwords = 1;
    o.rscale = 1.0;
    if (kind == OP_FSUM) {
        int l, w;
        if (fx_words(dc, l, w, one_row)) return -1;
        o.L = l;
        o.W = w;
        o.has_flags = dc.nonfinite ? 1 : 0;
        o.sc = fsum_scale(dc, rows);
        o.rscale = std::ldexp(1.0, -o.sc);
        nwords = w + o.has_flags;
    } else if (kind == OP_ISUM) {
        o.W = isum_words(dc, rows);
        if (o.W == 2 && rows >= (int64_t(1) << 32)) {
            set_err(FGPU_E_UNSUPPORTED, "exact SUM of '%s' over %lld rows (the two-word sum takes < 2^32)",
                    dc.name.c_str(), (long long)rows);
            return -1;
        }
        nwords = o.W;
    }
    if (L.ops.S + nwords > 64) {
        set_err(FGPU_E_UNSUPPORTED, "aggregate state exceeds 64 words");
        return -1;
    }
    for (int k = 0; k < nwords; ++k) {
        uint64_t id = 0;
        if (kind == OP_IMIN) id = (uint64_t)INT64_MAX;
        else if (kind == OP_IMAX) id = (uint64_t)INT64_MIN;
        else if (kind == OP_FMIN) id = ~0ull;
        L.ops.ident[L.ops.S + k] = id;
    }
    if (!op_packable(o)) {
        set_err(FGPU_E_UNSUPPORTED, "aggregate state layout out of range");
        return -1;
    }
    L.ops.S += nwords;
    L.ops.pk[L.ops.n] = op_pack(o);
    L.ops.op[L.ops.n++] = o;
    L.keys.push_back({kind, side, col});
    return L.ops.n - 1;
}

std::string alias_of(const fgpu_agg &a, int i) {
    if (a.alias && *a.alias) return a.alias;
    return "col" + std::to_string(i + 1);
}

// Builds ops + outputs for aggregates over one or two tables.  `slot_of(side,
// col)` binds a column and returns its slot (or <0).
// `rows` bounds the number of values one group's aggregate takes in (-1: the
// rows of t0).
template <class SlotFn>
// one_row: the aggregates have a single output row (no GROUP BY key, or a
// join), which lets an exact SUM take a rounded window (fx_words).
int build_layout(const fgpu_agg *aggs, int n_aggs, const fgpu_table *t0, const fgpu_table *t1,
                 SlotFn slot_of, AggLayout &L, int64_t rows = -1, bool one_row = false) {
    if (rows < 0) rows = t0->n_rows;
    L = AggLayout{};
    L.ops.n = 0;
    L.ops.S = 1;  // word 0 = rows
    L.ops.ident[0] = 0;
    if (n_aggs > FGPU_MAX_AGGS) return set_err(FGPU_E_UNSUPPORTED, "at most %d aggregates", FGPU_MAX_AGGS);
    L.outs.n = n_aggs;
    for (int i = 0; i < n_aggs; ++i) {
        const fgpu_agg &a = aggs[i];
        L.names.push_back(alias_of(a, i));
        DOut &o = L.outs.o[i];
        o = DOut{};
        if (a.func < FGPU_COUNT || a.func > FGPU_MAX) return set_err(FGPU_E_INVALID_QUERY, "unknown aggregate %d", a.func);
        if (a.col < 0) {
            if (a.func != FGPU_COUNT) return set_err(FGPU_E_INVALID_QUERY, "only COUNT may take *");
            o.kind = OUT_COUNT_ROWS;
            continue;
        }
        const int side = t1 ? a.side : 0;
        const fgpu_table *t = side == 1 ? t1 : t0;
        if (!t || a.col >= (int)t->cols.size())
            return set_err(FGPU_E_COLUMN_NOT_FOUND, "aggregate column %d not found", a.col);
        const DevCol &dc = t->cols[a.col];
        const int slot = slot_of(side, a.col);
        if (slot < 0) return set_err(FGPU_E_UNSUPPORTED, "query reads too many columns");
        const bool nullable = dc.null_count > 0;
        const bool isf = dc.dtype == FGPU_FLOAT64;
        if (dc.dtype == FGPU_UTF8 && a.func != FGPU_COUNT)
            return set_err(FGPU_E_INVALID_QUERY, "Cannot aggregate '%s' as float", L.names.back().c_str());
        int cnt_word = 0;
        // A NULL build side (LEFT) makes every build column nullable.
        const bool need_cnt = nullable || side == 1 || a.func == FGPU_AVG || a.func == FGPU_COUNT;
        if (need_cnt && (nullable || side == 1)) {
            const int ci = add_op(L, OP_CNT, side, a.col, dc, slot, rows, one_row);
            if (ci < 0) return FGPU_E_UNSUPPORTED;
            cnt_word = L.ops.op[ci].word;
        }
        o.cnt_word = (int16_t)cnt_word;
        int kind = 0;
        switch (a.func) {
            case FGPU_COUNT:
                o.kind = cnt_word ? OUT_COUNT : OUT_COUNT_ROWS;
                o.word = (int16_t)cnt_word;
                continue;
            case FGPU_SUM: kind = isf ? OP_FSUM : OP_ISUM; o.kind = isf ? OUT_SUM_F : OUT_SUM_I; break;
            case FGPU_AVG: kind = isf ? OP_FSUM : OP_ISUM; o.kind = isf ? OUT_AVG_F : OUT_AVG_I; break;
            case FGPU_MIN: kind = isf ? OP_FMIN : OP_IMIN; o.kind = isf ? OUT_MIN_F : OUT_MIN_I; break;
            default: kind = isf ? OP_FMAX : OP_IMAX; o.kind = isf ? OUT_MAX_F : OUT_MAX_I; break;
        }
        const int oi = add_op(L, kind, side, a.col, dc, slot, rows, one_row);
        if (oi < 0) return g_err.empty() ? set_err(FGPU_E_UNSUPPORTED, "aggregate layout") : FGPU_E_UNSUPPORTED;
        const DOp &op = L.ops.op[oi];
        o.word = op.word;
        o.W = op.W;
        o.L = op.L;
        o.has_flags = op.has_flags;
    }
    return FGPU_OK;
}

// Copies the emitted rows of the last GROUP BY / aggregate into a result.
// `known` >= 0: the row count is already on the host (a batch copies all of
// them at once); otherwise it is read from d_count.
// Copies the emitted groups into a result.  The columns' values and validity
// bytes are copied by DMA straight into the result's pinned block (see
// pin_get); Utf8 key codes are narrowed to u32 on the host.
int collect(fgpu_ctx *c, const AggLayout &L, int64_t n_slots, const DevCol *key, int key_dtype,
            const std::string &key_name, int64_t *d_key, uint8_t *d_key_valid,
            int64_t **d_val, uint8_t **d_valid, unsigned long long *d_count, fgpu_result **out,
            int64_t known = -1) {
    unsigned long long n = (unsigned long long)known;
    if (known < 0) {
        HIPCHK(mail_copy(c, d_count, sizeof n));
        HIPCHK(mail_take(c, &n, sizeof n));
    }
    if ((int64_t)n > n_slots) return set_err(FGPU_E_DEVICE, "emit counter %llu exceeds %lld slots", n, (long long)n_slots);
    auto *r = new fgpu_result();
    r->rows = (int64_t)n;
    if (c->tune.keep_on_device) n = 0;  // columns stay in device scratch
    const int n_cols = (key ? 1 : 0) + L.outs.n;
    char *blk = nullptr;
    if (n) {
        blk = (char *)pin_get((size_t)n * 9 * n_cols, &r->block_n);
        if (!blk) {
            delete r;
            return set_err(FGPU_E_OUT_OF_MEMORY, "pinned host memory for %llu result rows", n);
        }
        r->block = blk;
    }
    auto col_at = [&](int ci) { return blk + (size_t)ci * n * 9; };
    hipError_t e = hipSuccess;
    auto copy_in = [&](int ci, const void *d_v, const void *d_ok) {
        if (!n || e != hipSuccess) return;
        e = hipMemcpyAsync(col_at(ci), d_v, n * 8, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(col_at(ci) + n * 8, d_ok, n, hipMemcpyDeviceToHost, c->stream);
    };
    int ci = 0;
    if (key) copy_in(ci++, d_key, d_key_valid);
    for (int i = 0; i < L.outs.n; ++i) copy_in(ci++, d_val[i], d_valid[i]);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        delete r;
        return set_err(FGPU_E_DEVICE, "result copy: %s", hipGetErrorString(e));
    }
    ci = 0;
    if (key) {
        RCol kc;
        kc.name = key_name;
        kc.dtype = key_dtype == FGPU_UTF8 ? FGPU_UTF8 : FGPU_INT64;
        const char *v = col_at(ci++);
        if (key_dtype == FGPU_UTF8) {  // Utf8 keys: codes as u32
            kc.u32.resize(n);
            for (size_t i = 0; i < n; ++i) kc.u32[i] = (uint32_t)((const int64_t *)v)[i];
            kc.valid.assign((const uint8_t *)v + n * 8, (const uint8_t *)v + n * 9);
        } else {
            kc.xv = v;
            kc.xvalid = (const uint8_t *)v + n * 8;
        }
        kc.dict_src = key_dtype == FGPU_UTF8 ? key : nullptr;
        r->cols.push_back(std::move(kc));
    }
    for (int i = 0; i < L.outs.n; ++i) {
        RCol rc;
        rc.name = L.names[i];
        const int k = L.outs.o[i].kind;
        const bool is_f = k == OUT_AVG_I || k == OUT_SUM_F || k == OUT_AVG_F || k == OUT_MIN_F || k == OUT_MAX_F;
        rc.dtype = is_f ? FGPU_FLOAT64 : FGPU_INT64;
        char *v = col_at(ci++);
        rc.xv = v;
        rc.xvalid = (const uint8_t *)v + n * 8;
        if (k == OUT_SUM_I && L.outs.o[i].W == 2 && n) {
            // A two-word integer sum emits the exact int64 total, or, where the
            // total leaves int64, its correctly rounded double with validity 2.
            // A column with any such group is returned as Float64 (the type the
            // reference gives every SUM, src/query/aggregate.rs:3396-3413).
            int64_t *vals = (int64_t *)v;
            uint8_t *ok = (uint8_t *)v + n * 8;
            bool any_wide = false;
            for (size_t g = 0; g < n && !any_wide; ++g) any_wide = ok[g] == 2;
            if (any_wide) {
                for (size_t g = 0; g < n; ++g) {
                    if (ok[g] == 1) {
                        const double d = (double)vals[g];
                        std::memcpy(&vals[g], &d, 8);
                    } else if (ok[g] == 2) {
                        ok[g] = 1;
                    }
                }
                rc.dtype = FGPU_FLOAT64;
            }
        }
        r->cols.push_back(std::move(rc));
    }
    *out = r;
    return FGPU_OK;
}

// Emit-stage buffers carved from the B_OUT slot.
struct OutBufs {
    int64_t *key;
    uint8_t *key_valid;
    int64_t *val[FGPU_MAX_AGGS];
    uint8_t *valid[FGPU_MAX_AGGS];
    unsigned long long *count;
};
int out_bufs(fgpu_ctx *c, int64_t n_slots, int n_aggs, OutBufs &ob) {
    const size_t per = (size_t)((n_slots * 9 + 255) & ~int64_t(255));
    void *p;
    int rc = buf(c, B_OUT, per * (size_t)(n_aggs + 1) + 256, &p);
    if (rc) return rc;
    char *b = (char *)p;
    ob.count = (unsigned long long *)b;
    b += 256;
    ob.key = (int64_t *)b;
    ob.key_valid = (uint8_t *)(b + n_slots * 8);
    b += per;
    for (int i = 0; i < n_aggs; ++i) {
        ob.val[i] = (int64_t *)b;
        ob.valid[i] = (uint8_t *)(b + n_slots * 8);
        b += per;
    }
    for (int i = n_aggs; i < FGPU_MAX_AGGS; ++i) {
        ob.val[i] = nullptr;
        ob.valid[i] = nullptr;
    }
    return FGPU_OK;
}

double hll_estimate(const unsigned int *reg) {
    const int m = 4096;
    double sum = 0;
    int zeros = 0;
    for (int i = 0; i < m; ++i) {
        sum += std::ldexp(1.0, -(int)reg[i]);
        zeros += reg[i] == 0;
    }
    const double alpha = 0.7213 / (1.0 + 1.079 / m);
    double e = alpha * m * (double)m / sum;
    if (e <= 2.5 * m && zeros) e = m * std::log((double)m / zeros);  // linear counting
    return e;
}

}  // namespace

// ─── C ABI ──────────────────────────────────────────────────────────────────
extern "C" {

const char *fgpu_last_error(void) { return g_err.c_str(); }
int fgpu_abi_version(void) { return FGPU_ABI_VERSION; }

int fgpu_open(int device, fgpu_ctx **out) {
    if (!out) return set_err(FGPU_E_INVALID_QUERY, "out is NULL");
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return set_err(FGPU_E_DEVICE, "device %d not present (%d devices)", device, n);
    HIPCHK(hipSetDevice(device));
    auto *c = new fgpu_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
    const hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return set_err(FGPU_E_DEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    if (hipHostMalloc((void **)&c->mail, MAIL_BYTES, hipHostMallocDefault) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return set_err(FGPU_E_OUT_OF_MEMORY, "pinned host memory for the context");
    }
    for (auto &ev : c->ev) (void)hipEventCreate(&ev);
    static_scratch_init(device);
    Tuning &t = c->tune;
    t.gb_block = env_int("FGPU_GB_BLOCK", t.gb_block);
    t.gb_waves_per_cu = env_int("FGPU_GB_WAVES", t.gb_waves_per_cu);
    t.gb_replicas = env_int("FGPU_GB_REPLICAS", t.gb_replicas);
    t.hash_lds_slots = env_int("FGPU_HASH_LDS_SLOTS", t.hash_lds_slots);
    t.probe_limit = env_int("FGPU_PROBE_LIMIT", t.probe_limit);
    t.join_block = env_int("FGPU_JOIN_BLOCK", t.join_block);
    t.join_waves_per_cu = env_int("FGPU_JOIN_WAVES", t.join_waves_per_cu);
    t.filter_waves_per_cu = env_int("FGPU_FILTER_WAVES", t.filter_waves_per_cu);
    t.join_partition = env_int("FGPU_JOIN_PARTITION", t.join_partition);
    t.join_partitions = env_int("FGPU_JOIN_PARTITIONS", t.join_partitions);
    t.join_slots_per_key = env_int("FGPU_JOIN_SLOTS_PER_KEY", t.join_slots_per_key);
    t.join_part_wgs_per_cu = env_int("FGPU_JOIN_PART_WGS", t.join_part_wgs_per_cu);
    t.join_probe_wgs_per_cu = env_int("FGPU_JOIN_PROBE_WGS", t.join_probe_wgs_per_cu);
    t.gb_partition = env_int("FGPU_GB_PARTITION", t.gb_partition);
    *out = c;
    return FGPU_OK;
}

void fgpu_close(fgpu_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto &b : c->bufs)
        if (b.p) (void)hipFree(b.p);
    for (auto &ev : c->ev) (void)hipEventDestroy(ev);
    (void)hipStreamDestroy(c->stream);
    if (c->mail) (void)hipHostFree(c->mail);
    delete c;
}

void *fgpu_stream(fgpu_ctx *c) { return c ? (void *)c->stream : nullptr; }

int fgpu_set_timing(fgpu_ctx *c, int on) {
    if (!c) return set_err(FGPU_E_INVALID_QUERY, "ctx is NULL");
    c->timing = on != 0;
    return FGPU_OK;
}
int fgpu_last_timing(const fgpu_ctx *c, fgpu_timing *out) {
    if (!c || !out) return set_err(FGPU_E_INVALID_QUERY, "NULL argument");
    *out = c->last;
    return FGPU_OK;
}
int64_t fgpu_scratch_bytes(const fgpu_ctx *c) {
    int64_t s = 0;
    if (c)
        for (auto &b : c->bufs) s += (int64_t)b.n;
    return s;
}

// The strategy the last call on this context used (diagnostics).
const char *fgpu_last_plan(const fgpu_ctx *c) { return c ? c->plan.c_str() : ""; }

// Tuning knobs (the autotuner sets these): name = value; unknown name -> error.
int fgpu_set_tuning(fgpu_ctx *c, const char *name, int value) {
    if (!c || !name) return set_err(FGPU_E_INVALID_QUERY, "NULL argument");
    Tuning &t = c->tune;
    const std::string n = name;
    if (n == "gb_block") t.gb_block = value;
    else if (n == "gb_waves_per_cu") t.gb_waves_per_cu = value;
    else if (n == "gb_replicas") t.gb_replicas = std::max(1, std::min(value, 256));
    else if (n == "hash_lds_slots") t.hash_lds_slots = value;
    else if (n == "probe_limit") t.probe_limit = value;
    else if (n == "dense_lds_bytes") t.dense_lds_bytes = value;
    else if (n == "join_block") t.join_block = value;
    else if (n == "join_waves_per_cu") t.join_waves_per_cu = value;
    else if (n == "filter_waves_per_cu") t.filter_waves_per_cu = value;
    else if (n == "keep_on_device") t.keep_on_device = value;
    else if (n == "zone_maps") t.zone_maps = value;
    else if (n == "join_partition") t.join_partition = value;
    else if (n == "gb_partition") t.gb_partition = value;
    else if (n == "gb_part_min_rows") t.gb_part_min_rows = value;
    else if (n == "gbp_cache_slots") t.gbp_cache_slots = value;
    else if (n == "gbp_probe_limit") t.gbp_probe_limit = value;
    else if (n == "gbp_table_slots") t.gbp_table_slots = value;
    else if (n == "gbp_partitions") t.gbp_partitions = value;
    else if (n == "gbp_cells_max_mb") t.gbp_cells_max_mb = value;
    else if (n == "gbp_scan_block") t.gbp_scan_block = value;
    else if (n == "gbp_agg_block") t.gbp_agg_block = value;
    else if (n == "gbp_scan_wgs_per_cu") t.gbp_scan_wgs_per_cu = value;
    else if (n == "join_part_min_rows") t.join_part_min_rows = value;
    else if (n == "join_partitions") t.join_partitions = value;
    else if (n == "join_part_wgs_per_cu") t.join_part_wgs_per_cu = value;
    else if (n == "join_probe_wgs_per_cu") t.join_probe_wgs_per_cu = value;
    else if (n == "join_slots_per_key") t.join_slots_per_key = value;
    else if (n == "join_compact") t.join_compact = value;
    else if (n == "gbp_pack") t.gbp_pack = value;
    else if (n == "batch_unit_rows") t.batch_unit_rows = value;
    else if (n == "batch_scan_block") {
        if (value != 0 && (value < 64 || value > 1024 || value % 64))
            return set_err(FGPU_E_INVALID_QUERY, "batch_scan_block must be 0 or a multiple of 64 in [64, 1024]");
        t.batch_scan_block = value;
    }
    else return set_err(FGPU_E_INVALID_QUERY, "unknown tuning knob '%s'", name);
    if (t.gb_block < 64 || t.gb_block > 1024 || t.gb_block % 64 || t.join_block < 64 ||
        t.join_block > 1024 || t.join_block % 64)
        return set_err(FGPU_E_INVALID_QUERY, "block sizes must be multiples of 64 in [64, 1024]");
    if (t.hash_lds_slots & (t.hash_lds_slots - 1))
        return set_err(FGPU_E_INVALID_QUERY, "hash_lds_slots must be a power of two");
    if (t.gbp_cache_slots < 0 || t.gbp_cache_slots % 8 || t.gbp_table_slots < 64 ||
        t.gbp_partitions < 0 || t.gbp_partitions > 16384 || (t.gbp_partitions & (t.gbp_partitions - 1)) ||
        t.gbp_scan_block % 64 || t.gbp_agg_block % 64 ||
        t.gbp_scan_block < 64 || t.gbp_scan_block > 1024 || t.gbp_agg_block < 64 ||
        t.gbp_agg_block > 1024 || t.gbp_scan_wgs_per_cu < 1 || t.gbp_probe_limit < 1)
        return set_err(FGPU_E_INVALID_QUERY, "partitioned GROUP BY knobs out of range");
    if (t.join_partitions && (t.join_partitions < 8 || t.join_partitions > 4096 ||
                              (t.join_partitions & (t.join_partitions - 1))))
        return set_err(FGPU_E_INVALID_QUERY, "join_partitions must be 0 or a power of two in [8, 4096]");
    if (t.join_part_wgs_per_cu < 1 || t.join_probe_wgs_per_cu < 1)
        return set_err(FGPU_E_INVALID_QUERY, "join workgroup knobs out of range");
    return FGPU_OK;
}

// ─── tables ─────────────────────────────────────────────────────────────────
void fgpu_table_free(fgpu_table *t) {
    if (!t) return;
    (void)hipSetDevice(t->device);
    for (auto &c : t->cols) {
        if (c.d_values) (void)hipFree(c.d_values);
        if (c.d_valid) (void)hipFree(c.d_valid);
    }
    delete t;
}

// A Utf8 column's per-batch dictionaries merged into one, following
// Utf8Column::concat_all (src/storage/column.rs:1539-1595).
struct DictMerge {
    std::vector<std::string> dict;  // the table-wide dictionary
    bool keep = false;              // every batch has the first batch's dictionary: codes stay
    std::vector<uint32_t> lut;      // local code -> global code, batch after batch
    std::vector<int64_t> lut_off;   // n_batches + 1
    std::vector<int64_t> row_off;   // n_batches + 1
};

static int merge_dicts(const fgpu_utf8_batches &ub, int64_t n_rows, const std::string &name, DictMerge &m) {
    const int nb = ub.n_batches;
    if (!ub.batch_rows || !ub.dict_len || !ub.dicts)
        return set_err(FGPU_E_INVALID_QUERY, "column '%s': batch arrays are NULL", name.c_str());
    m.row_off.assign((size_t)nb + 1, 0);
    for (int b = 0; b < nb; ++b) {
        if (ub.batch_rows[b] < 0 || ub.dict_len[b] < 0 || (ub.dict_len[b] > 0 && !ub.dicts[b]))
            return set_err(FGPU_E_INVALID_QUERY, "column '%s': bad batch %d", name.c_str(), b);
        m.row_off[(size_t)b + 1] = m.row_off[(size_t)b] + ub.batch_rows[b];
    }
    if (m.row_off[(size_t)nb] != n_rows)
        return set_err(FGPU_E_INVALID_QUERY, "column '%s': batches hold %lld rows, the table %lld", name.c_str(),
                       (long long)m.row_off[(size_t)nb], (long long)n_rows);
    auto str = [&](int b, int k) { return std::string(ub.dicts[b][k] ? ub.dicts[b][k] : ""); };
    // fast paths: one shared dictionary array, or equal entries everywhere
    bool same = true;
    for (int b = 1; b < nb && same; ++b) {
        if (ub.dict_len[b] != ub.dict_len[0]) same = false;
        else if (ub.dicts[b] != ub.dicts[0])
            for (int k = 0; k < ub.dict_len[0] && same; ++k) same = str(b, k) == str(0, k);
    }
    if (same) {
        m.keep = true;
        for (int k = 0; k < ub.dict_len[0]; ++k) m.dict.push_back(str(0, k));
        return FGPU_OK;
    }
    // merge: first-seen order over the dictionaries in batch order
    std::unordered_map<std::string, uint32_t> seen;
    m.lut_off.assign((size_t)nb + 1, 0);
    for (int b = 0; b < nb; ++b) {
        for (int k = 0; k < ub.dict_len[b]; ++k) {
            std::string t = str(b, k);
            auto it = seen.find(t);
            uint32_t g;
            if (it != seen.end()) {
                g = it->second;
            } else {
                g = (uint32_t)m.dict.size();
                if (g == NULL_CODE) return set_err(FGPU_E_UNSUPPORTED, "column '%s': dictionary too large", name.c_str());
                seen.emplace(t, g);
                m.dict.push_back(std::move(t));
            }
            m.lut.push_back(g);
        }
        m.lut_off[(size_t)b + 1] = (int64_t)m.lut.size();
    }
    return FGPU_OK;
}

// Applies a merge's local-to-global maps to the column's device codes.
static int remap_codes(fgpu_ctx *ctx, DevCol &c, const DictMerge &m) {
    const int nb = (int)m.row_off.size() - 1;
    const size_t lut_bytes = std::max<size_t>(m.lut.size(), 1) * 4, off_bytes = m.row_off.size() * 8;
    char *mem = nullptr;
    if (hipMalloc(&mem, lut_bytes + 2 * off_bytes + 8) != hipSuccess)
        return set_err(FGPU_E_OUT_OF_MEMORY, "column '%s': dictionary maps", c.name.c_str());
    int64_t *d_row_off = (int64_t *)mem, *d_lut_off = (int64_t *)(mem + off_bytes);
    unsigned long long *d_bad = (unsigned long long *)(mem + 2 * off_bytes);
    uint32_t *d_lut = (uint32_t *)(mem + 2 * off_bytes + 8);
    unsigned long long bad = 0;
    hipError_t e = hipMemcpy(d_row_off, m.row_off.data(), off_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_lut_off, m.lut_off.data(), off_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_bad, &bad, 8, hipMemcpyHostToDevice);
    if (e == hipSuccess && !m.lut.empty()) e = hipMemcpy(d_lut, m.lut.data(), m.lut.size() * 4, hipMemcpyHostToDevice);
    const int64_t n = m.row_off.back();
    if (e == hipSuccess)
        e = launch_dict_remap(nullptr, (uint32_t *)c.d_values, n, d_row_off, nb, d_lut, d_lut_off, d_bad,
                              grid_for(ctx, n, BLOCK, 8));
    if (e == hipSuccess) e = hipMemcpy(&bad, d_bad, 8, hipMemcpyDeviceToHost);
    (void)hipFree(mem);
    if (e != hipSuccess) return set_err(FGPU_E_DEVICE, "dictionary remap: %s", hipGetErrorString(e));
    if (bad) return set_err(FGPU_E_INDEX_OUT_OF_BOUNDS, "column '%s': a code is past its batch's dictionary",
                            c.name.c_str());
    return FGPU_OK;
}

// Statistics of a device column (upload, materialised joins): min / max, NULL
// count, float exponent range and non-finite flags, the distinct estimate and
// the per-batch zone map.  d_st is scratch for one ColStatsDev.
static int column_stats(fgpu_ctx *ctx, DevCol &c, int64_t n_rows, int32_t batch_rows, ColStatsDev *d_st) {
    ColStatsDev h{};
    h.imin = INT64_MAX;
    h.imax = INT64_MIN;
    h.fmin_enc = ~0ull;
    h.fmax_enc = 0;
    h.emin = 1 << 20;
    h.emax = -(1 << 20);
    hipError_t e = hipMemcpyAsync(d_st, &h, sizeof h, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && n_rows) {
        const int grid = grid_for(ctx, n_rows, BLOCK, 8);
        e = launch_col_stats(ctx->stream, c.d_values, c.dtype, c.d_valid, n_rows, d_st, grid);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&h, d_st, sizeof h, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return set_err(FGPU_E_DEVICE, "column stats: %s", hipGetErrorString(e));
    c.null_count = (int64_t)h.nulls;
    c.imin = h.imin;
    c.imax = h.imax;
    c.fmin = __builtin_bit_cast(double, f64_unorder(h.fmin_enc));
    c.fmax = __builtin_bit_cast(double, f64_unorder(h.fmax_enc));
    c.emin = h.emin;
    c.emax = h.emax;
    c.nonfinite = h.nonfinite;
    c.distinct = (int64_t)std::llround(hll_estimate(h.hll));
    c.zones.clear();
    if (n_rows && c.dtype != FGPU_UTF8) {  // zone map
        const int64_t nb = (n_rows + batch_rows - 1) / batch_rows;
        BatchZone *d_z = nullptr;
        e = hipMalloc(&d_z, (size_t)nb * sizeof(BatchZone));
        if (e == hipSuccess) e = launch_batch_zones(ctx->stream, c.d_values, c.dtype, c.d_valid, n_rows, batch_rows, d_z);
        if (e == hipSuccess) {
            c.zones.resize((size_t)nb);
            e = hipMemcpyAsync(c.zones.data(), d_z, (size_t)nb * sizeof(BatchZone), hipMemcpyDeviceToHost, ctx->stream);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (d_z) (void)hipFree(d_z);
        if (e != hipSuccess) return set_err(FGPU_E_DEVICE, "zone map: %s", hipGetErrorString(e));
    }
    return FGPU_OK;
}

static int table_upload_impl(fgpu_ctx *ctx, const fgpu_column_desc *cols, const fgpu_utf8_batches *utf8,
                             int n_cols, int64_t n_rows, int32_t batch_rows, fgpu_table **out) {
    if (!ctx || !out || (n_cols > 0 && !cols)) return set_err(FGPU_E_INVALID_QUERY, "NULL argument");
    if (n_rows < 0 || n_cols < 0) return set_err(FGPU_E_INVALID_QUERY, "negative size");
    if (n_rows >= (int64_t(1) << 31))
        return set_err(FGPU_E_UNSUPPORTED, "tables are limited to 2^31-1 rows (u32 row ids)");
    if (batch_rows <= 0 || batch_rows % 64) return set_err(FGPU_E_INVALID_QUERY, "batch_rows must be a positive multiple of 64");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (int rc = set_device(ctx->device)) return rc;
    auto *t = new fgpu_table();
    t->device = ctx->device;
    t->n_rows = n_rows;
    t->batch_rows = batch_rows;
    t->cols.resize(n_cols);
    auto fail = [&](int rc) {
        fgpu_table_free(t);
        return rc;
    };
    ColStatsDev *d_st = nullptr;
    if (hipMalloc(&d_st, sizeof(ColStatsDev)) != hipSuccess)
        return fail(set_err(FGPU_E_OUT_OF_MEMORY, "stats buffer"));
    int rc = FGPU_OK;
    for (int i = 0; i < n_cols && rc == FGPU_OK; ++i) {
        const fgpu_column_desc &d = cols[i];
        DevCol &c = t->cols[i];
        c.name = d.name ? d.name : "";
        c.dtype = d.dtype;
        if (d.dtype != FGPU_INT32 && d.dtype != FGPU_INT64 && d.dtype != FGPU_FLOAT64 && d.dtype != FGPU_UTF8) {
            rc = set_err(FGPU_E_TYPE_MISMATCH, "column '%s': unsupported dtype %d", c.name.c_str(), d.dtype);
            break;
        }
        if (n_rows > 0 && !d.values) {
            rc = set_err(FGPU_E_INVALID_QUERY, "column '%s': values is NULL", c.name.c_str());
            break;
        }
        DictMerge merge;
        const bool batched = d.dtype == FGPU_UTF8 && utf8 && utf8[i].n_batches > 0;
        if (batched && (rc = merge_dicts(utf8[i], n_rows, c.name, merge))) break;
        if (d.dtype == FGPU_UTF8) {
            if (batched) {
                c.dict = merge.dict;
            } else {
                if (d.dict_len < 0 || (d.dict_len > 0 && !d.dict)) {
                    rc = set_err(FGPU_E_INVALID_QUERY, "column '%s': bad dictionary", c.name.c_str());
                    break;
                }
                for (int k = 0; k < d.dict_len; ++k) c.dict.push_back(d.dict[k] ? d.dict[k] : "");
            }
            for (size_t k = 0; k < c.dict.size(); ++k)
                if (c.empty_code < 0 && c.dict[k].empty()) c.empty_code = (int)k;
            for (auto &s : c.dict) c.dictp.push_back(s.c_str());
            // every non-NULL code must index the dictionary (merged columns: on the device)
            if (!batched || merge.keep) {
                const uint32_t *codes = (const uint32_t *)d.values;
                const uint32_t len = (uint32_t)c.dict.size();
                for (int64_t r = 0; r < n_rows; ++r)
                    if (codes[r] != NULL_CODE && codes[r] >= len) {
                        rc = set_err(FGPU_E_INDEX_OUT_OF_BOUNDS,
                                     "column '%s': code %u at row %lld >= dictionary size %u", c.name.c_str(),
                                     codes[r], (long long)r, len);
                        break;
                    }
                if (rc) break;
            }
        }
        const size_t bytes = (size_t)std::max<int64_t>(n_rows, 1) * c.elem_bytes();
        hipError_t e = hipMalloc(&c.d_values, bytes);
        if (e != hipSuccess) { rc = set_err(FGPU_E_OUT_OF_MEMORY, "column '%s': %zu bytes", c.name.c_str(), bytes); break; }
        t->bytes += (int64_t)bytes;
        if (n_rows) e = hipMemcpy(c.d_values, d.values, (size_t)n_rows * c.elem_bytes(), hipMemcpyHostToDevice);
        if (e != hipSuccess) { rc = set_err(FGPU_E_DEVICE, "upload: %s", hipGetErrorString(e)); break; }
        if (batched && !merge.keep && n_rows && (rc = remap_codes(ctx, c, merge))) break;
        if (d.validity && d.dtype != FGPU_UTF8) {
            const size_t vb = (size_t)((n_rows + 63) / 64) * 8;
            e = hipMalloc(&c.d_valid, std::max<size_t>(vb, 8));
            if (e == hipSuccess && vb) e = hipMemcpy(c.d_valid, d.validity, vb, hipMemcpyHostToDevice);
            if (e != hipSuccess) { rc = set_err(FGPU_E_DEVICE, "validity upload: %s", hipGetErrorString(e)); break; }
            t->bytes += (int64_t)vb;
        }
        if ((rc = column_stats(ctx, c, n_rows, batch_rows, d_st))) break;
    }
    (void)hipFree(d_st);
    if (rc) return fail(rc);
    *out = t;
    return FGPU_OK;
}

int fgpu_table_upload(fgpu_ctx *ctx, const fgpu_column_desc *cols, int n_cols, int64_t n_rows,
                      int32_t batch_rows, fgpu_table **out) {
    return table_upload_impl(ctx, cols, nullptr, n_cols, n_rows, batch_rows, out);
}

int fgpu_table_upload_batches(fgpu_ctx *ctx, const fgpu_column_desc *cols, const fgpu_utf8_batches *utf8,
                              int n_cols, int64_t n_rows, int32_t batch_rows, fgpu_table **out) {
    return table_upload_impl(ctx, cols, utf8, n_cols, n_rows, batch_rows, out);
}

int64_t fgpu_table_rows(const fgpu_table *t) { return t ? t->n_rows : -1; }
int fgpu_table_cols(const fgpu_table *t) { return t ? (int)t->cols.size() : -1; }
int fgpu_table_column_info(const fgpu_table *t, int col, fgpu_column_view *out) {
    if (!t || !out) return set_err(FGPU_E_INVALID_QUERY, "NULL argument");
    if (col < 0 || col >= (int)t->cols.size())
        return set_err(FGPU_E_INDEX_OUT_OF_BOUNDS, "column %d of %zu", col, t->cols.size());
    const DevCol &c = t->cols[col];
    out->name = c.name.c_str();
    out->dtype = c.dtype;
    out->values = nullptr;
    out->valid = nullptr;
    out->dict_len = (int32_t)c.dict.size();
    out->dict = c.dictp.empty() ? nullptr : c.dictp.data();
    return FGPU_OK;
}
int64_t fgpu_table_device_bytes(const fgpu_table *t) { return t ? t->bytes : 0; }

int fgpu_table_column_stats(const fgpu_table *t, int col, int64_t *min_i64, int64_t *max_i64,
                            int64_t *null_count, int64_t *distinct_est) {
    if (!t || col < 0 || col >= (int)t->cols.size()) return set_err(FGPU_E_COLUMN_NOT_FOUND, "column %d", col);
    const DevCol &c = t->cols[col];
    if (min_i64) *min_i64 = c.dtype == FGPU_FLOAT64 ? __builtin_bit_cast(int64_t, c.fmin) : c.imin;
    if (max_i64) *max_i64 = c.dtype == FGPU_FLOAT64 ? __builtin_bit_cast(int64_t, c.fmax) : c.imax;
    if (null_count) *null_count = c.null_count;
    if (distinct_est) *distinct_est = c.distinct;
    return FGPU_OK;
}

// Float exponent range recorded at upload (mantissa-LSB exponents).
int fgpu_table_float_range(const fgpu_table *t, int col, int *emin, int *emax, uint64_t *nonfinite) {
    if (!t || col < 0 || col >= (int)t->cols.size()) return set_err(FGPU_E_COLUMN_NOT_FOUND, "column %d", col);
    const DevCol &c = t->cols[col];
    if (emin) *emin = c.emin;
    if (emax) *emax = c.emax;
    if (nonfinite) *nonfinite = c.nonfinite;
    return FGPU_OK;
}

// ─── WHERE ──────────────────────────────────────────────────────────────────
// Zone-map pruning (BatchStats::can_skip, src/storage/stats.rs:57-113): may a
// row of batch b pass leaf i of the predicate?  The reference compares f64
// images of the bounds; here each comparison class uses the evaluator's own
// arithmetic (int64 / int32 exactly, (double)v for an integer column against a
// float literal, doubles with NaN apart), so a batch is skipped exactly when no
// row of it can pass.  IS [NOT] NULL prunes on the counts; dictionary
// comparisons never prune.
static bool zone_may(const DevCol &col, const BatchZone &z, int64_t rows, const DPredNode &nd) {
    const int op = nd.op;
    auto test = [&](double lo, double hi, double lit) {
        switch (op) {
            case FGPU_LT: return lo < lit;
            case FGPU_LE: return lo <= lit;
            case FGPU_GT: return hi > lit;
            case FGPU_GE: return hi >= lit;
            case FGPU_EQ: return lo <= lit && lit <= hi;
            default: return !(lo == hi && lo == lit);  // NE
        }
    };
    switch (nd.cls) {
        case CMP_FALSE: return false;
        case CMP_IS_NULL: return rows - z.n_valid > 0;
        case CMP_IS_NOT_NULL: return z.n_valid > 0;
        case CMP_DICT: return true;
        case CMP_I64:
        case CMP_I32: {
            if (z.n_valid == 0) return false;
            const int64_t lo = z.lo, hi = z.hi, lit = nd.lit_i;
            switch (op) {
                case FGPU_LT: return lo < lit;
                case FGPU_LE: return lo <= lit;
                case FGPU_GT: return hi > lit;
                case FGPU_GE: return hi >= lit;
                case FGPU_EQ: return lo <= lit && lit <= hi;
                default: return !(lo == hi && lo == lit);
            }
        }
        case CMP_I64_AS_F64:
            if (z.n_valid == 0) return false;
            return test((double)z.lo, (double)z.hi, nd.lit_f);
        case CMP_F64: {
            if (op == FGPU_NE && z.n_nan > 0) return true;  // NaN != x holds
            if (z.n_valid - z.n_nan == 0) return false;
            const double lo = __builtin_bit_cast(double, f64_unorder((uint64_t)z.lo));
            const double hi = __builtin_bit_cast(double, f64_unorder((uint64_t)z.hi));
            return test(lo, hi, nd.lit_f);
        }
        default: return true;
    }
    (void)col;
}

// Sets dp.live to the batches that may hold a passing row, when the zone maps
// rule some out (uploaded into the context's scratch); else leaves it null.
// The plan note goes to *note when given, else onto c->plan.
static int prune_batches(fgpu_ctx *c, const fgpu_table *t, const fgpu_pred *p, DPred &dp,
                         std::string *note = nullptr) {
    dp.live = nullptr;
    dp.batch_rows = t->batch_rows;
    if (!c->tune.zone_maps || !p || dp.n == 0 || t->n_rows == 0) return FGPU_OK;
    const int64_t nb = (t->n_rows + t->batch_rows - 1) / t->batch_rows;
    std::vector<uint8_t> &live = c->live_host;
    live.assign((size_t)nb, 0);
    int64_t n_live = 0;
    for (int64_t b = 0; b < nb; ++b) {
        const int64_t rows = std::min<int64_t>(t->batch_rows, t->n_rows - b * t->batch_rows);
        bool st[FGPU_MAX_PRED_NODES];
        int sp = 0;
        for (int i = 0; i < dp.n; ++i) {
            const fgpu_pred_node &nd = p->nodes[i];
            if (nd.kind == FGPU_P_AND || nd.kind == FGPU_P_OR) {
                const bool y = st[--sp], x = st[--sp];
                st[sp++] = nd.kind == FGPU_P_AND ? x && y : x || y;
                continue;
            }
            const DevCol &col = t->cols[nd.col];
            st[sp++] = col.zones.empty() ? true : zone_may(col, col.zones[(size_t)b], rows, dp.nodes[i]);
        }
        live[(size_t)b] = st[0] ? 1 : 0;
        n_live += live[(size_t)b];
    }
    char msg[96];
    snprintf(msg, sizeof msg, " zone maps: %lld of %lld batches live", (long long)n_live, (long long)nb);
    (note ? *note : c->plan) += msg;
    if (n_live == nb) return FGPU_OK;
    void *d;
    if (int rc = buf(c, B_LIVE, (size_t)nb, &d)) return rc;
    HIPCHK(hipMemcpyAsync(d, live.data(), (size_t)nb, hipMemcpyHostToDevice, c->stream));
    dp.live = (const uint8_t *)d;
    return FGPU_OK;
}

int fgpu_filter_count(fgpu_ctx *c, const fgpu_table *t, const fgpu_pred *p, int64_t *out_count) {
    if (!c || !t || !out_count) return set_err(FGPU_E_INVALID_QUERY, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    if (int rc = set_device(c->device)) return rc;
    Binder b(t);
    DPred dp;
    if (int rc = compile_pred(c, t, p, b, dp)) return rc;
    c->plan = "filter:count";
    if (int rc = prune_batches(c, t, p, dp)) return rc;
    unsigned long long *d_cnt;
    if (int rc = buf(c, B_COUNTERS, 64 * 8, (void **)&d_cnt)) return rc;
    mark(c, 0);
    HIPCHK(launch_fill_u64(c->stream, (uint64_t *)d_cnt, 1, 0));
    const int grid = grid_for(c, (t->n_rows + 63) / 64 * 64, BLOCK, c->tune.filter_waves_per_cu);
    mark(c, 1);
    if (t->n_rows) HIPCHK(launch_filter_count(c->stream, b.dc, dp, t->n_rows, d_cnt, grid));
    mark(c, 2);
    unsigned long long n = 0;
    HIPCHK(mail_copy(c, d_cnt, 8));
    mark(c, 3);
    HIPCHK(mail_take(c, &n, 8));
    int64_t bytes = 0;
    for (int s = 0; s < b.dc.n; ++s) bytes += t->n_rows * (b.dc.dtype[s] == FGPU_INT64 || b.dc.dtype[s] == FGPU_FLOAT64 ? 8 : 4);
    finish_timing(c, bytes, false);
    *out_count = (int64_t)n;
    return FGPU_OK;
}

static int filter_mask_dev(fgpu_ctx *c, const fgpu_table *t, const fgpu_pred *p, uint64_t **d_words) {
    Binder b(t);
    DPred dp;
    if (int rc = compile_pred(c, t, p, b, dp)) return rc;
    c->plan = "filter:mask";
    if (int rc = prune_batches(c, t, p, dp)) return rc;
    const int64_t nw = (t->n_rows + 63) / 64;
    if (int rc = buf(c, B_MASK, (size_t)std::max<int64_t>(nw, 1) * 8, (void **)d_words)) return rc;
    const int grid = grid_for(c, nw * 64, BLOCK, c->tune.filter_waves_per_cu);
    mark(c, 1);
    if (nw) HIPCHK(launch_filter_mask(c->stream, b.dc, dp, t->n_rows, *d_words, grid));
    mark(c, 2);
    return FGPU_OK;
}

int fgpu_filter_mask(fgpu_ctx *c, const fgpu_table *t, const fgpu_pred *p, uint64_t *host_words) {
    if (!c || !t || !host_words) return set_err(FGPU_E_INVALID_QUERY, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    if (int rc = set_device(c->device)) return rc;
    uint64_t *d_words;
    mark(c, 0);
    if (int rc = filter_mask_dev(c, t, p, &d_words)) return rc;
    const int64_t nw = (t->n_rows + 63) / 64;
    if (nw) HIPCHK(hipMemcpyAsync(host_words, d_words, nw * 8, hipMemcpyDeviceToHost, c->stream));
    mark(c, 3);
    HIPCHK(hipStreamSynchronize(c->stream));
    finish_timing(c, 0, false);
    return FGPU_OK;
}

int fgpu_filter_select(fgpu_ctx *c, const fgpu_table *t, const fgpu_pred *p, const int32_t *sel,
                       int n_sel, fgpu_result **out) {
    if (!c || !t || !out || (n_sel > 0 && !sel)) return set_err(FGPU_E_INVALID_QUERY, "NULL argument");
    for (int i = 0; i < n_sel; ++i)
        if (sel[i] < 0 || sel[i] >= (int)t->cols.size())
            return set_err(FGPU_E_COLUMN_NOT_FOUND, "column %d not in table", sel[i]);
    std::lock_guard<std::mutex> lk(c->mu);
    if (int rc = set_device(c->device)) return rc;
    uint64_t *d_words;
    mark(c, 0);
    if (int rc = filter_mask_dev(c, t, p, &d_words)) return rc;
    const int64_t nw = (t->n_rows + 63) / 64;
    int64_t *d_offs, *d_tmp;
    if (int rc = buf(c, B_OFFS, (size_t)(nw + 1) * 8, (void **)&d_offs)) return rc;
    if (int rc = buf(c, B_TMP, (size_t)(nw / 4096 + 2) * 8, (void **)&d_tmp)) return rc;
    HIPCHK(launch_word_offsets(c->stream, d_words, nw, d_offs, d_tmp));
    int64_t total = 0;
    HIPCHK(hipMemcpyAsync(&total, d_offs + nw, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    auto *r = new fgpu_result();
    r->rows = total;
    void *d_out;
    int rc = buf(c, B_OUT, (size_t)std::max<int64_t>(total, 1) * 9 + 64, &d_out);
    if (rc) {
        delete r;
        return rc;
    }
    const int grid = grid_for(c, nw * 64, BLOCK, 16);
    for (int i = 0; i < n_sel; ++i) {
        const DevCol &dc = t->cols[sel[i]];
        RCol rc2;
        rc2.name = dc.name;
        rc2.dtype = dc.dtype == FGPU_FLOAT64 ? FGPU_FLOAT64 : dc.dtype == FGPU_UTF8 ? FGPU_UTF8 : FGPU_INT64;
        const int eb = (int)dc.elem_bytes();
        uint8_t *d_valid = (uint8_t *)d_out + (size_t)std::max<int64_t>(total, 1) * 8;
        hipError_t e = launch_gather(c->stream, d_words, d_offs, nw, dc.d_values, eb, dc.d_valid, d_out, d_valid, grid);
        rc2.valid.resize(total);
        std::vector<int32_t> tmp32;
        if (e == hipSuccess && total) {
            if (eb == 8) {
                if (rc2.dtype == FGPU_FLOAT64) { rc2.f64.resize(total); e = hipMemcpyAsync(rc2.f64.data(), d_out, total * 8, hipMemcpyDeviceToHost, c->stream); }
                else { rc2.i64.resize(total); e = hipMemcpyAsync(rc2.i64.data(), d_out, total * 8, hipMemcpyDeviceToHost, c->stream); }
            } else {
                tmp32.resize(total);
                e = hipMemcpyAsync(tmp32.data(), d_out, total * 4, hipMemcpyDeviceToHost, c->stream);
            }
            if (e == hipSuccess) e = hipMemcpyAsync(rc2.valid.data(), d_valid, total, hipMemcpyDeviceToHost, c->stream);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            delete r;
            return set_err(FGPU_E_DEVICE, "filter_select: %s", hipGetErrorString(e));
        }
        if (eb == 4) {
            if (dc.dtype == FGPU_UTF8) {
                rc2.u32.resize(total);
                for (int64_t k = 0; k < total; ++k) {
                    rc2.u32[k] = (uint32_t)tmp32[k];
                    rc2.valid[k] = rc2.u32[k] != NULL_CODE;
                }
                rc2.dict_src = &dc;
            } else {
                rc2.i64.resize(total);
                for (int64_t k = 0; k < total; ++k) rc2.i64[k] = tmp32[k];
            }
        }
        r->cols.push_back(std::move(rc2));
    }
    mark(c, 3);
    finish_timing(c, 0, false);
    *out = r;
    return FGPU_OK;
}

// ─── GROUP BY ───────────────────────────────────────────────────────────────
// Radix-partitioned GROUP BY (GbPartArgs).  Returns FGPU_OK with *out set, a
// negative value when the query should take the single-table hash path
// instead (inputs unsuitable, or a cell / table / output bound was exceeded),
// or an error code.
static int groupby_partitioned(fgpu_ctx *c, const fgpu_table *t, const DevCol *key, const GbArgs &a,
                               const AggLayout &L, const Binder &b, fgpu_result **out) {
    const Tuning &tu = c->tune;
    if (!tu.gb_partition || t->n_rows < tu.gb_part_min_rows) return -1;
    const int S = L.ops.S;
    GbPartArgs A{};
    A.g = a;
    A.outs = L.outs;
    bool nullable_inputs = false;
    for (int j = 0; j < a.n_in; ++j)
        for (size_t col = 0; col < t->cols.size(); ++col)
            if (t->cols[col].d_values == b.dc.v[a.in_col[j]] && t->cols[col].null_count) nullable_inputs = true;
    A.has_valid_word = nullable_inputs ? 1 : 0;
    // rows with a NULL key, or the INT64_MIN key that shares the NULL group
    A.null_group = key->null_count > 0 || key->imin == INT64_MIN ? 1 : 0;
    // Packed row records: key and an integer input 0 in one word when their
    // ranges (from the upload statistics) fit 64 bits together.
    A.pack = 0;
    if (tu.gbp_pack && a.n_in > 0 && !A.has_valid_word && !A.null_group && key->dtype != FGPU_UTF8 &&
        key->imin <= key->imax) {
        const DevCol *in0 = nullptr;
        for (const DevCol &dc : t->cols)
            if (dc.d_values == b.dc.v[a.in_col[0]]) in0 = &dc;
        if (in0 && (in0->dtype == FGPU_INT64 || in0->dtype == FGPU_INT32) && in0->imin <= in0->imax) {
            auto bits = [](uint64_t span) { return span ? 64 - __builtin_clzll(span) : 1; };
            const int kb = bits((uint64_t)key->imax - (uint64_t)key->imin);
            const int vb = bits((uint64_t)in0->imax - (uint64_t)in0->imin);
            if (kb + vb <= 64) {
                A.pack = 1;
                A.pk_kbits = kb;
                A.pk_kmin = key->imin;
                A.pk_vmin = in0->imin;
            }
        }
    }
    A.RW = (A.pack ? 0 : 1) + a.n_in + A.has_valid_word;
    A.SR = (1 + S + A.RW - 1) / A.RW;
    A.G = c->num_cus * tu.gbp_scan_wgs_per_cu;
    // Pass-2 table: the most slots (a multiple of 64, at most the tuned cap)
    // that fit in LDS beside the two cell-count prefix arrays and the emit
    // scan.  The fan-out keeps a partition's expected keys at <= 0.6 of that,
    // then the table shrinks to ~2x the expected keys (more workgroups per CU).
    const int64_t lds_fixed2 = 8 * (int64_t)(A.G + 1) + 256;
    // (a slot is its key, its state and a 4-byte entry of the emit list)
    const int64_t t2max =
        std::min<int64_t>(tu.gbp_table_slots, ((int64_t)tu.lds_limit - lds_fixed2) / (8 * (S + 1) + 4)) / 64 * 64;
    if (t2max < 64) return -1;
    const int64_t distinct = std::max<int64_t>(key->distinct, 1);
    int64_t P = tu.gbp_partitions ? tu.gbp_partitions
                                  : next_pow2(std::max<int64_t>(8, (int64_t)std::ceil((double)distinct / (0.6 * t2max))));
    if (P > 16384) return -1;
    const int T2 = (int)std::max<int64_t>(64, std::min<int64_t>(t2max, ((int64_t)std::ceil(2.0 * distinct / P) + 63) / 64 * 64));
    A.T2 = T2;
    auto lds2 = [&](int t2) { return (size_t)t2 * (12 + 8 * (size_t)S) + (size_t)lds_fixed2; };
    A.log_p = 0;
    while ((int64_t(1) << A.log_p) < P) ++A.log_p;
    // Pass-1 cache: the most slots (a multiple of 8, the ways) beside the cell
    // cursors and the doorkeeper bitmap, at most the tuned cap.
    auto lds1 = [&](int64_t h) { return (size_t)(h + h * S + S) * 8 + (size_t)P * 8 + 2048 * 4 + 256; };
    const int64_t hmax = ((int64_t)tu.lds_limit - (int64_t)lds1(0)) / (8 * (S + 1));
    int H = (int)(std::min<int64_t>(tu.gbp_cache_slots, std::max<int64_t>(hmax, 0)) / 8 * 8);
    if (lds1(H) > (size_t)tu.lds_limit) return -1;
    A.H = H;
    A.probe_limit = tu.gbp_probe_limit;
    A.rows_per_wg = ((t->n_rows + A.G - 1) / A.G + 255) / 256 * 256;  // tiles stay 32-row aligned
    A.Cw = (int)std::ceil((double)A.rows_per_wg / (double)P * 1.25) +
           A.SR * ((int)((4 * (int64_t)H + P - 1) / P) + 4) + 32;
    const int64_t out_cap = std::min<int64_t>(t->n_rows + 1, std::max<int64_t>(65536, 2 * key->distinct + 1024));
    A.out_cap = out_cap;
    int rc;
    void *p;
    if ((rc = buf(c, B_CELLCNT, (size_t)A.G * P * 8, &p))) return rc;
    A.cell_cnt = (unsigned int *)p;
    if ((rc = buf(c, B_NULLBUF, (size_t)A.G * S * 8, &p))) return rc;
    A.nullbuf = (uint64_t *)p;
    unsigned long long *d_flags;
    if ((rc = buf(c, B_COUNTERS, 64 * 8, (void **)&d_flags))) return rc;
    A.flags = d_flags;
    OutBufs ob;
    if ((rc = out_bufs(c, out_cap, L.outs.n, ob))) return rc;
    A.out_key = ob.key;
    A.out_key_valid = ob.key_valid;
    for (int i = 0; i < FGPU_MAX_AGGS; ++i) {
        A.out_val[i] = ob.val[i];
        A.out_valid[i] = ob.valid[i];
    }
    A.out_count = d_flags + 3;  // zeroed and read back with the flags
    char msg[256];
    const int fw = gbp_fast_width(A);
    snprintf(msg, sizeof msg, "groupby:partitioned P=%lld H=%d T2=%d G=%d RW=%d%s tiles=%s", (long long)P, H, T2,
             A.G, A.RW, A.pack ? " packed" : "", fw == 4 ? "fast4" : fw == 8 ? "fast8" : "generic");
    c->plan = msg;
    unsigned long long h[4] = {0, 0, 0, 0};  // the three flags, then the output row count
    mark(c, 0);
    // A cell that overflows (a heavy key some workgroup could not cache) is
    // retried once with cells sized from the counts the first attempt kept.
    for (int attempt = 0; attempt < 2; ++attempt) {
        if ((size_t)A.G * P * A.Cw * A.RW * 8 > (size_t)tu.gbp_cells_max_mb << 20) {
            c->plan += " -> fallback: cells over budget";
            return -1;
        }
        if ((rc = buf(c, B_CELLS, (size_t)A.G * P * A.Cw * A.RW * 8, &p))) return rc;
        A.cells = (int64_t *)p;
        HIPCHK(launch_fill_u64(c->stream, (uint64_t *)d_flags, 4, 0));
        mark(c, 1);
        HIPCHK(launch_gbp_scan(c->stream, A, tu.gbp_scan_block, lds1(H)));
        mark(c, 2);
        HIPCHK(launch_gbp_agg(c->stream, A, tu.gbp_agg_block, lds2(T2)));  // returns at once after a cell overflow
        mark(c, 3);
        HIPCHK(mail_copy(c, d_flags, sizeof h));
        HIPCHK(mail_take(c, h, sizeof h));
        snprintf(msg, sizeof msg, " Cw=%d", A.Cw);
        c->plan += msg;
        const bool diag = env_int("FGPU_GBP_DIAG", 0) != 0;
        if (h[0] || diag) {  // cell occupancy: rows that missed the cache, state records
            std::vector<unsigned int> cc((size_t)A.G * P * 2);
            HIPCHK(hipMemcpy(cc.data(), A.cell_cnt, cc.size() * 4, hipMemcpyDeviceToHost));
            uint64_t cold = 0, states = 0, mx = 0, need = 0;
            for (size_t i = 0; i < cc.size(); i += 2) {
                cold += cc[i];
                states += cc[i + 1];
                mx = std::max<uint64_t>(mx, cc[i]);
                need = std::max<uint64_t>(need, cc[i] + (uint64_t)cc[i + 1] * A.SR);
            }
            if (diag) {
                snprintf(msg, sizeof msg, " [cold=%llu states=%llu max_cell=%llu]", (unsigned long long)cold,
                         (unsigned long long)states, (unsigned long long)mx);
                c->plan += msg;
            }
            if (h[0] && attempt == 0) {
                c->plan += " cell full, resized";
                A.Cw = (int)std::min<uint64_t>((uint64_t)need * 5 / 4 + 32, (uint64_t)INT32_MAX / 2);
                continue;
            }
        }
        break;
    }
    if (h[0] || h[1] || h[2]) {  // a bound was exceeded: single-table hash path
        c->plan += h[0] ? " -> fallback: cell full" : h[1] ? " -> fallback: pass-2 table full"
                                                       : " -> fallback: output full";
        return -1;
    }
    if ((rc = collect(c, L, out_cap, key, key->dtype, key->name, ob.key, ob.key_valid, ob.val, ob.valid,
                      ob.count, out, (int64_t)h[3])))
        return rc;
    int64_t bytes = 0;
    for (int s2 = 0; s2 < b.dc.n; ++s2)
        bytes += t->n_rows * (b.dc.dtype[s2] == FGPU_INT64 || b.dc.dtype[s2] == FGPU_FLOAT64 ? 8 : 4);
    finish_timing(c, bytes, false);
    return FGPU_OK;
}

// Resolves a GROUP BY call: the key column, the compiled predicate (Utf8 pass
// tables in the arena, or at pass_dst when given), the bound columns and the
// aggregate state layout.
static int gb_prepare(fgpu_ctx *c, const fgpu_table *t, const fgpu_pred *p, const fgpu_groupby_spec *spec,
                      Binder &b, GbArgs &a, AggLayout &L, const DevCol *&key, uint8_t *pass_dst = nullptr,
                      size_t pass_cap = 0, size_t *pass_used = nullptr) {
    const int kc = spec->key_col;
    if (kc >= (int)t->cols.size()) return set_err(FGPU_E_COLUMN_NOT_FOUND, "group key column %d not in table", kc);
    key = kc >= 0 ? &t->cols[kc] : nullptr;
    if (key && key->dtype == FGPU_FLOAT64)
        return set_err(FGPU_E_INVALID_QUERY, "a Float64 GROUP BY key goes through fgpu_groupby_keys");
    DPred dp;
    if (int rc = compile_pred(c, t, p, b, dp, pass_dst, pass_cap, pass_used)) return rc;
    a.pred = dp;
    a.key_col = key ? b.bind(kc) : -1;
    if (key && a.key_col < 0) return set_err(FGPU_E_UNSUPPORTED, "query reads too many columns");
    // aggregate inputs
    int in_of_col[256];
    for (int &x : in_of_col) x = -1;
    a.n_in = 0;
    bool too_many = false;
    auto slot_of = [&](int /*side*/, int col) -> int {
        if (col < 256 && in_of_col[col] >= 0) return in_of_col[col];
        if (a.n_in >= MAX_IN) { too_many = true; return -1; }
        const int s = b.bind(col);
        if (s < 0) return -1;
        a.in_col[a.n_in] = (int8_t)s;
        if (col < 256) in_of_col[col] = a.n_in;
        return a.n_in++;
    };
    if (int rc = build_layout(spec->aggs, spec->n_aggs, t, nullptr, slot_of, L, -1, kc < 0))
        return too_many ? set_err(FGPU_E_UNSUPPORTED, "at most %d distinct aggregate input columns", MAX_IN) : rc;
    a.cols = b.dc;
    a.ops = L.ops;
    a.n_rows = t->n_rows;
    return FGPU_OK;
}

// Picks the GROUP BY strategy from the key's statistics: scalar registers (no
// key), a dense state table in LDS or in HBM (slot = key - key_min; Utf8 keys
// by dictionary code), or a hash table.  Fills a.key_min / range / key_is_utf8
// / utf8_null_code and the state rows a dense or scalar mode needs.
static int gb_pick_mode(const fgpu_ctx *c, const DevCol *key, const AggLayout &L, GbArgs &a,
                        int64_t &n_state_groups) {
    const int S = L.ops.S;
    int mode;
    if (!key && L.ops.n <= 8) {  // register accumulators of the scalar kernel
        mode = GB_SCALAR;
        a.range = 0;
        n_state_groups = 1;
    } else if (!key) {  // many aggregates, no GROUP BY: one dense group
        a.key_min = 0;
        a.range = 0;
        mode = GB_DENSE_LDS;
        n_state_groups = 1;
    } else {
        a.key_is_utf8 = key->dtype == FGPU_UTF8;
        int64_t range;
        if (a.key_is_utf8) {
            a.key_min = 0;
            range = (int64_t)key->dict.size() + 1;  // + synthetic "" group for NULL keys
            a.utf8_null_code = key->empty_code >= 0 ? (uint32_t)key->empty_code : (uint32_t)key->dict.size();
        } else if (key->imin > key->imax) {  // all NULL / empty
            a.key_min = 0;
            range = 1;
        } else {
            a.key_min = key->imin;
            const unsigned __int128 span = (unsigned __int128)((__int128)key->imax - (__int128)key->imin) + 1;
            range = span > (unsigned __int128)(int64_t(1) << 40) ? (int64_t(1) << 40) : (int64_t)span;
        }
        a.range = range;
        const int64_t dense_bytes = (range + 1) * S * 8;
        if (dense_bytes <= c->tune.dense_lds_bytes) {
            mode = GB_DENSE_LDS;
            n_state_groups = range + 1;
        } else if (a.key_is_utf8 || (range <= (int64_t(1) << 24) && range <= std::max<int64_t>(4 * key->distinct, 65536))) {
            mode = GB_DENSE_GLOBAL;
            n_state_groups = range + 1;
        } else {
            mode = GB_HASH;
            n_state_groups = 0;
        }
    }
    return mode;
}

static int groupby_impl(fgpu_ctx *c, const fgpu_table *t, const fgpu_pred *p, const fgpu_groupby_spec *spec,
                        fgpu_result **out);

int fgpu_groupby(fgpu_ctx *c, const fgpu_table *t, const fgpu_pred *p,
                 const fgpu_groupby_spec *spec, fgpu_result **out) {
    if (!c || !t || !spec || !out || (spec->n_aggs > 0 && !spec->aggs))
        return set_err(FGPU_E_INVALID_QUERY, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    if (int rc = set_device(c->device)) return rc;
    return groupby_impl(c, t, p, spec, out);
}

// GROUP BY over several key columns or one Float64 key (the reference's slow
// path): one int64 key per row (k_composite_key) appended to a view of the
// table, the ordinary GROUP BY over it, then the group keys decoded back into
// one typed column per key.
int fgpu_groupby_keys(fgpu_ctx *c, const fgpu_table *t, const fgpu_pred *p, const int32_t *key_cols,
                      int32_t n_keys, const fgpu_groupby_spec *spec, fgpu_result **out) {
    if (!c || !t || !spec || !out || !key_cols || (spec->n_aggs > 0 && !spec->aggs))
        return set_err(FGPU_E_INVALID_QUERY, "NULL argument");
    if (n_keys < 1 || n_keys > FGPU_MAX_GROUP_KEYS)
        return set_err(FGPU_E_INVALID_QUERY, "GROUP BY takes 1 to %d key columns, not %d", FGPU_MAX_GROUP_KEYS, n_keys);
    for (int i = 0; i < n_keys; ++i)
        if (key_cols[i] < 0 || key_cols[i] >= (int)t->cols.size())
            return set_err(FGPU_E_COLUMN_NOT_FOUND, "group key column %d not in table", key_cols[i]);
    std::lock_guard<std::mutex> lk(c->mu);
    if (int rc = set_device(c->device)) return rc;
    fgpu_groupby_spec s2 = *spec;
    const DevCol &k0 = t->cols[key_cols[0]];
    if (n_keys == 1 && k0.dtype != FGPU_FLOAT64) {  // the fast path
        s2.key_col = key_cols[0];
        return groupby_impl(c, t, p, &s2, out);
    }
    CompositeKeyArgs ka{};
    ka.n = n_keys;
    ka.n_rows = t->n_rows;
    DevCol kc;
    kc.name = "__group_key";
    kc.dtype = FGPU_INT64;
    int64_t card[FGPU_MAX_GROUP_KEYS] = {0};
    if (n_keys == 1) {  // one Float64 key: its canonical bits + 1 (NULL: INT64_MIN)
        ka.float_key = 1;
        ka.v[0] = k0.d_values;
        ka.valid[0] = k0.d_valid;
        ka.dtype[0] = FGPU_FLOAT64;
        kc.imin = -(int64_t(1) << 62);  // a range no dense table covers: the hash strategies
        kc.imax = int64_t(1) << 62;
        kc.null_count = k0.null_count;
        kc.distinct = std::max<int64_t>(1, k0.distinct);
    } else {
        unsigned __int128 combos = 1;
        double dist = 1.0;
        for (int i = 0; i < n_keys; ++i) {
            const DevCol &d = t->cols[key_cols[i]];
            if (d.dtype == FGPU_FLOAT64)
                return set_err(FGPU_E_UNSUPPORTED, "a Float64 GROUP BY key ('%s') must be the only key", d.name.c_str());
            if (d.dtype == FGPU_UTF8) {
                card[i] = (int64_t)d.dict.size() + 1;
                ka.null_code[i] = 0;
                for (size_t e = 0; e < d.dict.size(); ++e)
                    if (d.dict[e] == "NULL") ka.null_code[i] = (int64_t)e + 1;  // NULL and "NULL" print alike
            } else if (d.imin > d.imax) {  // no value: NULL only
                card[i] = 1;
                ka.kmin[i] = 0;
            } else {
                const unsigned __int128 span = (unsigned __int128)((__int128)d.imax - (__int128)d.imin) + 2;
                if (span > ((unsigned __int128)1 << 62))
                    return set_err(FGPU_E_UNSUPPORTED, "GROUP BY key '%s' spans more than 2^62 values", d.name.c_str());
                card[i] = (int64_t)span;
                ka.kmin[i] = d.imin;
            }
            ka.v[i] = d.d_values;
            ka.valid[i] = d.d_valid;
            ka.dtype[i] = d.dtype;
            ka.mul[i] = (int64_t)combos;
            combos *= (unsigned __int128)card[i];
            if (combos > ((unsigned __int128)1 << 62))
                return set_err(FGPU_E_UNSUPPORTED, "the GROUP BY keys have more than 2^62 combinations");
            dist *= (double)std::max<int64_t>(1, d.distinct) + (d.null_count ? 1.0 : 0.0);
        }
        kc.imin = 0;
        kc.imax = (int64_t)combos - 1;
        kc.distinct = (int64_t)std::max(1.0, std::min({dist, (double)std::max<int64_t>(1, t->n_rows), (double)combos}));
    }
    void *kbuf;
    if (int rc = buf(c, B_GKEY, (size_t)std::max<int64_t>(1, t->n_rows) * 8, &kbuf)) return rc;
    ka.out = (int64_t *)kbuf;
    kc.d_values = kbuf;
    if (t->n_rows) HIPCHK(launch_composite_key(c->stream, ka, grid_for(c, t->n_rows, 256, 8)));
    fgpu_table view;
    view.device = t->device;
    view.n_rows = t->n_rows;
    view.batch_rows = t->batch_rows;
    view.bytes = t->bytes;
    view.cols = t->cols;
    view.cols.push_back(kc);
    s2.key_col = (int)view.cols.size() - 1;
    fgpu_result *r = nullptr;
    if (int rc = groupby_impl(c, &view, p, &s2, &r)) return rc;
    // one column per key, decoded from the group's int64 key
    std::vector<RCol> keys((size_t)n_keys);
    const RCol &gk = r->cols[0];
    const int64_t n = gk.xv ? r->rows : 0;  // 0: the columns stayed on the device
    const int64_t *kv = (const int64_t *)gk.xv;
    const uint8_t *kok = gk.xvalid;
    for (int i = 0; i < n_keys; ++i) {
        const DevCol &d = t->cols[key_cols[i]];
        RCol &rc = keys[(size_t)i];
        rc.name = d.name;
        rc.dtype = d.dtype == FGPU_UTF8 ? FGPU_UTF8 : d.dtype == FGPU_FLOAT64 ? FGPU_FLOAT64 : FGPU_INT64;
        rc.dict_src = d.dtype == FGPU_UTF8 ? &d : nullptr;
        rc.valid.resize((size_t)n);
        if (rc.dtype == FGPU_UTF8) rc.u32.resize((size_t)n);
        else if (rc.dtype == FGPU_FLOAT64) rc.f64.resize((size_t)n);
        else rc.i64.resize((size_t)n);
    }
    for (int64_t g = 0; g < n; ++g) {
        if (ka.float_key) {
            const uint64_t bits = (uint64_t)kv[g] - 1;
            double x;
            std::memcpy(&x, &bits, 8);
            keys[0].f64[(size_t)g] = kok[g] ? x : 0.0;
            keys[0].valid[(size_t)g] = kok[g] ? 1 : 0;
            continue;
        }
        int64_t comp = kv[g];
        for (int i = 0; i < n_keys; ++i) {
            const int64_t code = comp % card[i];
            comp /= card[i];
            RCol &rc = keys[(size_t)i];
            rc.valid[(size_t)g] = code != 0;
            if (rc.dtype == FGPU_UTF8) rc.u32[(size_t)g] = code ? (uint32_t)(code - 1) : NULL_CODE;
            else rc.i64[(size_t)g] = code ? ka.kmin[i] + code - 1 : 0;
        }
    }
    std::vector<RCol> cols = std::move(keys);
    for (size_t i = 1; i < r->cols.size(); ++i) cols.push_back(std::move(r->cols[i]));
    r->cols = std::move(cols);
    *out = r;
    return FGPU_OK;
}

static int groupby_run(fgpu_ctx *c, const fgpu_table *t, const DevCol *key, Binder &b, GbArgs &a, AggLayout &L,
                       fgpu_result **out);

static int groupby_impl(fgpu_ctx *c, const fgpu_table *t, const fgpu_pred *p, const fgpu_groupby_spec *spec,
                        fgpu_result **out) {
    const DevCol *key = nullptr;
    Binder b(t);
    GbArgs a{};
    AggLayout L;
    if (int rc = gb_prepare(c, t, p, spec, b, a, L, key)) return rc;
    // zone maps: every strategy's scan skips the batches that cannot pass
    std::string zone_note;
    if (int rc = prune_batches(c, t, p, a.pred, &zone_note)) return rc;
    const int rc = groupby_run(c, t, key, b, a, L, out);
    c->plan += zone_note;
    return rc;
}

// State replicas of the dense-LDS and scalar GROUP BY: every workgroup merges
// its states into HBM with device atomics at its end, and about a thousand
// workgroups merging into the same few words queue behind one another; with
// R replicas each word takes grid / R merges and k_gb_fold adds R - 1 rows per
// group.  Never more replicas than workgroups.
static int gb_replicas(const fgpu_ctx *c, int mode, int64_t n_rows) {
    if (mode != GB_DENSE_LDS && mode != GB_SCALAR) return 1;
    const int block = mode == GB_SCALAR ? 256 : c->tune.gb_block;
    const int grid = grid_for(c, (n_rows + 3) / 4, block, c->tune.gb_waves_per_cu);
    return std::max(1, std::min(c->tune.gb_replicas, grid));
}

// Runs a prepared GROUP BY with the strategy the key and state size call for.
static int groupby_run(fgpu_ctx *c, const fgpu_table *t, const DevCol *key, Binder &b, GbArgs &a, AggLayout &L,
                       fgpu_result **out) {
    const int S = L.ops.S;
    int64_t n_state_groups;
    const int mode = gb_pick_mode(c, key, L, a, n_state_groups);
    c->plan.clear();
    if (mode == GB_HASH && !a.key_is_utf8) {
        const int prc = groupby_partitioned(c, t, key, a, L, b, out);
        if (prc >= 0) return prc;
    }
    {
        static const char *names[] = {"dense_lds", "dense_global", "hash", "scalar"};
        if (!c->plan.empty()) c->plan += "; ";
        c->plan += std::string("groupby:") + names[mode];
    }
    int64_t cap = 0;
    if (mode == GB_HASH)
        cap = next_pow2(std::max<int64_t>(1024, (int64_t)(2.2 * (double)std::max<int64_t>(key->distinct, 1))));
    unsigned long long *d_cnt;
    if (int rc = buf(c, B_COUNTERS, 64 * 8, (void **)&d_cnt)) return rc;
    const int reps = gb_replicas(c, mode, t->n_rows);
    for (int attempt = 0; attempt < 4; ++attempt) {
        if (mode == GB_HASH) n_state_groups = cap + 1;
        int64_t *d_keys = nullptr;
        uint64_t *d_state;
        if (mode == GB_HASH)
            if (int rc = buf(c, B_KEYS, (size_t)cap * 8, (void **)&d_keys)) return rc;
        if (int rc = buf(c, B_STATE, (size_t)n_state_groups * S * 8 * reps, (void **)&d_state)) return rc;
        a.keys = d_keys;
        a.state = d_state;
        a.cap_mask = cap ? (uint64_t)(cap - 1) : 0;
        a.counters = d_cnt;
        a.replicas = reps;
        a.rep_stride = n_state_groups * S;
        mark(c, 0);
        HIPCHK(launch_fill_u64(c->stream, (uint64_t *)d_cnt, 8, 0));
        mark(c, 4);
        HIPCHK(launch_state_init(c->stream, d_keys, cap, d_state, n_state_groups * reps, L.ops));
        mark(c, 5);
        // launch shape
        int block = c->tune.gb_block;
        size_t lds = 0;
        if (mode == GB_DENSE_LDS) lds = (size_t)n_state_groups * S * 8;
        if (mode == GB_HASH) {
            int slots = c->tune.hash_lds_slots;
            while (slots > 0 && (size_t)slots * (8 + (size_t)S * 8) > 64 * 1024) slots >>= 1;
            a.lds_slots = slots;
            a.probe_limit = std::max(1, c->tune.probe_limit);
            lds = (size_t)slots * (8 + (size_t)S * 8);
        }
        if (mode == GB_SCALAR) block = 256;
        const int grid = grid_for(c, (t->n_rows + 3) / 4, block, c->tune.gb_waves_per_cu);
        mark(c, 1);
        if (t->n_rows) HIPCHK(launch_groupby(c->stream, mode, a, grid, block, lds));
        if (t->n_rows) HIPCHK(launch_gb_fold(c->stream, a, n_state_groups));
        mark(c, 2);
        unsigned long long h_cnt[2] = {0, 0};
        if (mode == GB_HASH) {
            HIPCHK(mail_copy(c, d_cnt, 16));
            HIPCHK(mail_take(c, h_cnt, 16));
            if (h_cnt[1]) {  // table overflowed: grow and rerun
                cap *= 4;
                continue;
            }
        }
        // emit
        OutBufs ob;
        if (int rc = out_bufs(c, n_state_groups, L.outs.n, ob)) return rc;
        EmitArgs e{};
        e.outs = L.outs;
        e.keys = d_keys;
        e.state = d_state;
        e.S = S;
        e.n_slots = n_state_groups;
        e.mode = key ? mode : GB_SCALAR;  // no GROUP BY: always one row
        e.key_min = a.key_min;
        e.key_is_utf8 = a.key_is_utf8;
        e.out_key = key ? ob.key : nullptr;
        e.out_key_valid = ob.key_valid;
        for (int i = 0; i < FGPU_MAX_AGGS; ++i) {
            e.out_val[i] = ob.val[i];
            e.out_valid[i] = ob.valid[i];
        }
        e.out_count = ob.count;
        HIPCHK(launch_fill_u64(c->stream, (uint64_t *)ob.count, 1, 0));
        HIPCHK(launch_emit(c->stream, e));
        mark(c, 3);
        int64_t bytes = 0;
        for (int s = 0; s < b.dc.n; ++s) bytes += t->n_rows * (b.dc.dtype[s] == FGPU_INT64 || b.dc.dtype[s] == FGPU_FLOAT64 ? 8 : 4);
        int rc = collect(c, L, n_state_groups, key, key ? key->dtype : 0, key ? key->name : "", ob.key,
                         ob.key_valid, ob.val, ob.valid, ob.count, out);
        if (rc) return rc;
        finish_timing(c, bytes, true);
        return FGPU_OK;
    }
    return set_err(FGPU_E_OUT_OF_MEMORY, "group table kept overflowing (last capacity %lld)", (long long)cap);
}

int fgpu_groupby_static(fgpu_ctx *c, const fgpu_table *t, const fgpu_pred *p, const fgpu_groupby_spec *spec,
                        int64_t *out_key, uint8_t *out_key_valid, int64_t *const *out_vals,
                        uint8_t *const *out_valid, int64_t capacity, int64_t *n_out) {
    if (!c || !t || !spec || !n_out || (spec->n_aggs > 0 && (!spec->aggs || !out_vals || !out_valid)))
        return set_err(FGPU_E_INVALID_QUERY, "NULL argument");
    *n_out = 0;
    std::lock_guard<std::mutex> lk(c->mu);
    if (int rc = set_device(c->device)) return rc;
    if (c->device >= STATIC_MAX_DEVICES) return set_err(FGPU_E_UNSUPPORTED, "device %d has no static scratch", c->device);
    StaticScratch &st = g_static[c->device];
    if (!st.ptr)
        return set_err(FGPU_E_DEVICE, "static scratch of device %d: %s", c->device,
                       st.looked_up ? hipGetErrorString(st.lookup) : "not looked up");
    // One call at a time per device: every call shares this device's array.
    std::lock_guard<std::mutex> use(st.use);
    // scratch layout: [0, 512) counters, [512, 520) output row count,
    // [4096, 4096 + pass) Utf8 pass tables, then the state table (256-aligned)
    uint8_t *base = (uint8_t *)st.ptr;
    unsigned long long *d_cnt = (unsigned long long *)base;
    unsigned long long *d_rows = (unsigned long long *)(base + 512);
    const size_t pass_off = 4096;
    size_t pass_used = 0;
    const DevCol *key = nullptr;
    Binder b(t);
    GbArgs a{};
    AggLayout L;
    if (int rc = gb_prepare(c, t, p, spec, b, a, L, key, base + pass_off, STATIC_SCRATCH_BYTES - pass_off,
                            &pass_used))
        return rc;
    const int S = L.ops.S;
    int64_t n_state_groups;
    const int mode = gb_pick_mode(c, key, L, a, n_state_groups);
    if (mode == GB_HASH)
        return set_err(FGPU_E_UNSUPPORTED, "key '%s' spans %lld values: a hash GROUP BY allocates (use fgpu_groupby)",
                       key->name.c_str(), (long long)a.range);
    const size_t state_off = (pass_off + pass_used + 255) & ~size_t(255);
    const size_t state_bytes = (size_t)n_state_groups * S * 8;
    // replicas only while they fit beside the one table the query needs
    const int reps = (int)std::max<size_t>(1, std::min<size_t>(gb_replicas(c, mode, t->n_rows),
        state_off + state_bytes > STATIC_SCRATCH_BYTES ? 1 : (STATIC_SCRATCH_BYTES - state_off) / std::max<size_t>(state_bytes, 1)));
    if (state_off + state_bytes > STATIC_SCRATCH_BYTES)
        return set_err(FGPU_E_UNSUPPORTED, "group state of %lld slots x %d words needs %zu bytes; the static scratch "
                       "holds %zu", (long long)n_state_groups, S, state_off + state_bytes, STATIC_SCRATCH_BYTES);
    const int64_t max_rows = key ? n_state_groups : 1;
    if (capacity < max_rows) {
        *n_out = max_rows;
        return set_err(FGPU_E_INDEX_OUT_OF_BOUNDS, "output capacity %lld below the %lld rows this GROUP BY may emit",
                       (long long)capacity, (long long)max_rows);
    }
    if (key && (!out_key || !out_key_valid)) return set_err(FGPU_E_INVALID_QUERY, "NULL key output");
    for (int i = 0; i < spec->n_aggs; ++i)
        if (!out_vals[i] || !out_valid[i]) return set_err(FGPU_E_INVALID_QUERY, "NULL output of aggregate %d", i);
    uint64_t *d_state = (uint64_t *)(base + state_off);
    a.keys = nullptr;
    a.state = d_state;
    a.cap_mask = 0;
    a.counters = d_cnt;
    a.replicas = reps;
    a.rep_stride = n_state_groups * S;
    // the array keeps the last call's contents: clear counters and the row count
    HIPCHK(launch_fill_u64(c->stream, (uint64_t *)d_cnt, 65, 0));
    HIPCHK(launch_state_init(c->stream, nullptr, 0, d_state, n_state_groups * reps, L.ops));
    const int block = mode == GB_SCALAR ? 256 : c->tune.gb_block;
    const size_t lds = mode == GB_DENSE_LDS ? (size_t)n_state_groups * S * 8 : 0;
    const int grid = grid_for(c, (t->n_rows + 3) / 4, block, c->tune.gb_waves_per_cu);
    if (t->n_rows) HIPCHK(launch_groupby(c->stream, mode, a, grid, block, lds));
    if (t->n_rows) HIPCHK(launch_gb_fold(c->stream, a, n_state_groups));
    EmitArgs e{};
    e.outs = L.outs;
    e.keys = nullptr;
    e.state = d_state;
    e.S = S;
    e.n_slots = n_state_groups;
    e.mode = key ? mode : GB_SCALAR;
    e.key_min = a.key_min;
    e.key_is_utf8 = a.key_is_utf8;
    e.out_key = key ? out_key : nullptr;
    e.out_key_valid = key ? out_key_valid : nullptr;
    for (int i = 0; i < FGPU_MAX_AGGS; ++i) {
        e.out_val[i] = i < spec->n_aggs ? out_vals[i] : nullptr;
        e.out_valid[i] = i < spec->n_aggs ? out_valid[i] : nullptr;
    }
    e.out_count = d_rows;
    HIPCHK(launch_emit(c->stream, e));
    unsigned long long rows = 0;
    HIPCHK(mail_copy(c, d_rows, sizeof rows));
    HIPCHK(mail_take(c, &rows, sizeof rows));
    if ((int64_t)rows > max_rows) return set_err(FGPU_E_DEVICE, "emit counter %llu exceeds %lld", rows, (long long)max_rows);
    *n_out = (int64_t)rows;
    c->plan = std::string("groupby_static:") + (mode == GB_SCALAR ? "scalar" : mode == GB_DENSE_LDS ? "dense_lds"
                                                                                                  : "dense_global");
    return FGPU_OK;
}

// ─── hash join ──────────────────────────────────────────────────────────────
void fgpu_join_index_free(fgpu_join_index *idx) {
    if (!idx) return;
    (void)hipSetDevice(idx->device);
    if (idx->slots) (void)hipFree(idx->slots);
    if (idx->min_key_rows) (void)hipFree(idx->min_key_rows);
    for (auto &st : idx->subs) {
        if (st.sub) (void)hipFree(st.sub);
        if (st.min_key_rows) (void)hipFree(st.min_key_rows);
    }
    delete idx;
}

int64_t fgpu_join_index_device_bytes(const fgpu_join_index *idx) { return idx ? idx->bytes : 0; }

static int join_build_impl(fgpu_ctx *c, const fgpu_table *build, int32_t key_col, fgpu_join_index **out) {
    if (key_col < 0 || key_col >= (int)build->cols.size())
        return set_err(FGPU_E_COLUMN_NOT_FOUND, "build key column %d not in table", key_col);
    const DevCol &kc = build->cols[key_col];
    if (kc.dtype != FGPU_INT64 && kc.dtype != FGPU_INT32)
        return set_err(FGPU_E_TYPE_MISMATCH, "join key '%s' must be Int64 or Int32", kc.name.c_str());
    auto *idx = new fgpu_join_index();
    idx->device = c->device;
    idx->build = build;
    idx->key_col = key_col;
    const int64_t n = build->n_rows;
    idx->cap = next_pow2(std::max<int64_t>(1024, 2 * n));
    auto fail = [&](int rc) {
        fgpu_join_index_free(idx);
        return rc;
    };
    if (hipMalloc(&idx->slots, idx->cap * 16) != hipSuccess)
        return fail(set_err(FGPU_E_OUT_OF_MEMORY, "join index of %lld slots", (long long)idx->cap));
    idx->bytes = idx->cap * 16;
    const bool has_min = kc.dtype == FGPU_INT64 && kc.imin == INT64_MIN;
    const int64_t min_cap = has_min ? std::max<int64_t>(n, 1) : 0;
    if (min_cap && hipMalloc(&idx->min_key_rows, min_cap * 4) != hipSuccess)
        return fail(set_err(FGPU_E_OUT_OF_MEMORY, "join index side list"));
    unsigned long long *d_cnt;
    if (int rc = buf(c, B_COUNTERS, 64 * 8, (void **)&d_cnt)) return fail(rc);
    mark(c, 4);
    hipError_t e = launch_fill_u64(c->stream, (uint64_t *)d_cnt, 8, 0);
    if (e == hipSuccess) e = launch_fill_u64(c->stream, (uint64_t *)idx->slots, 2 * idx->cap, (uint64_t)EMPTY_KEY);
    JoinBuildArgs ba{};
    ba.keys_col = kc.d_values;
    ba.key_dtype = kc.dtype;
    ba.key_valid = kc.d_valid;
    ba.n = n;
    ba.slots = idx->slots;
    ba.cap_mask = (uint64_t)(idx->cap - 1);
    ba.counters = d_cnt;
    ba.min_key_rows = idx->min_key_rows;
    ba.min_key_cap = min_cap;
    if (e == hipSuccess && n) e = launch_join_build(c->stream, ba, grid_for(c, n, BLOCK, 16));
    mark(c, 5);
    unsigned long long h[4] = {0, 0, 0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(h, d_cnt, 32, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(set_err(FGPU_E_DEVICE, "join build: %s", hipGetErrorString(e)));
    if (h[1]) return fail(set_err(FGPU_E_DEVICE, "join index overflow"));
    idx->n_min_key_rows = (int64_t)h[2];
    idx->unique_keys = !h[3] && h[2] <= 1;
    if (idx->min_key_rows) {
        // slot order of the side list is arbitrary; sort it so probes see ascending rows
        std::vector<uint32_t> v(h[2]);
        if (h[2]) {
            (void)hipMemcpy(v.data(), idx->min_key_rows, h[2] * 4, hipMemcpyDeviceToHost);
            std::sort(v.begin(), v.end());
            (void)hipMemcpy(idx->min_key_rows, v.data(), h[2] * 4, hipMemcpyHostToDevice);
        }
    }
    *out = idx;
    return FGPU_OK;
}

int fgpu_join_build(fgpu_ctx *c, const fgpu_table *build, int32_t key_col, fgpu_join_index **out) {
    if (!c || !build || !out) return set_err(FGPU_E_INVALID_QUERY, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    if (int rc = set_device(c->device)) return rc;
    return join_build_impl(c, build, key_col, out);
}

struct JoinPlan {
    JoinProbeArgs a{};
    AggLayout L;
    Binder *pb = nullptr, *bb = nullptr;
};

// Set when join_prepare refused a spec for needing more register accumulators
// than the probe kernels hold (JOIN_MAX_OPS).
static thread_local bool t_join_ops_over = false;

static int join_prepare(fgpu_ctx *c, const fgpu_table *probe, const fgpu_pred *pred,
                        const fgpu_table *build, const fgpu_join_index *idx,
                        const fgpu_join_spec *spec, Binder &pb, Binder &bb, JoinProbeArgs &a,
                        AggLayout &L) {
    if (spec->type < FGPU_INNER || spec->type > FGPU_ANTI)
        return set_err(FGPU_E_INVALID_QUERY, "unknown join type %d", spec->type);
    if (spec->probe_key < 0 || spec->probe_key >= (int)probe->cols.size())
        return set_err(FGPU_E_COLUMN_NOT_FOUND, "probe key column %d not in table", spec->probe_key);
    const DevCol &pk = probe->cols[spec->probe_key];
    if (pk.dtype != FGPU_INT64 && pk.dtype != FGPU_INT32)
        return set_err(FGPU_E_TYPE_MISMATCH, "join key '%s' must be Int64 or Int32", pk.name.c_str());
    if (idx && (idx->build != build || idx->key_col != spec->build_key))
        return set_err(FGPU_E_INVALID_QUERY, "join index was built on another table or key");
    for (int i = 0; i < spec->n_aggs; ++i)
        if (spec->aggs[i].col >= 0 && spec->aggs[i].side == 1 &&
            (spec->type == FGPU_SEMI || spec->type == FGPU_ANTI))
            return set_err(FGPU_E_COLUMN_NOT_FOUND, "SEMI/ANTI output has only probe columns");
    DPred dp;
    if (int rc = compile_pred(c, probe, pred, pb, dp)) return rc;
    a = JoinProbeArgs{};
    a.pred = dp;
    a.pkey = pb.bind(spec->probe_key);
    a.type = spec->type;
    a.n_probe = probe->n_rows;
    auto slot_of = [&](int side, int col) -> int { return side == 1 ? bb.bind(col) : pb.bind(col); };
    // Output rows of the join: at most one per probe row when the build keys are
    // unique (SEMI / ANTI always), else up to one per (probe, build) pair.
    int64_t out_rows = probe->n_rows;
    if ((spec->type == FGPU_INNER || spec->type == FGPU_LEFT) && !(idx && idx->unique_keys)) {
        const __int128 pairs = (__int128)probe->n_rows * std::max<int64_t>(build->n_rows, 1);
        out_rows = pairs > (__int128)INT64_MAX ? INT64_MAX : (int64_t)pairs;
    }
    if (int rc = build_layout(spec->aggs, spec->n_aggs, probe, build, slot_of, L, out_rows, true)) return rc;
    if (L.ops.n > 4) {
        t_join_ops_over = true;  // fgpu_join_agg then runs the aggregates in parts
        return set_err(FGPU_E_UNSUPPORTED, "join aggregates need %d accumulators (max 4)", L.ops.n);
    }
    a.pcols = pb.dc;
    a.bcols = bb.dc;
    a.n_ops = L.ops.n;
    for (int i = 0; i < L.ops.n; ++i) {
        const DOp &o = L.ops.op[i];
        JoinOp &j = a.ops[i];
        j.kind = o.kind;
        j.side = (int8_t)L.keys[i].side;
        j.col = o.col;
        j.word = o.word;
        j.L = o.L;
        j.W = o.W;
        j.has_flags = o.has_flags;
        j.sc = o.sc;
        j.rscale = o.rscale;
    }
    a.S = L.ops.S;
    return FGPU_OK;
}

// Radix-partitioned probe (see JoinPartArgs).  Returns FGPU_OK with the state
// row filled, a negative value when the inputs do not suit it (the caller then
// runs the direct probe), or an error code.
// Builds the P sub-tables into jp.sub (and the sentinel key's build rows into
// jp.min_key_rows) on the context's stream; fits = false when a sub-table
// would be too full (skewed keys) or an insert found no slot.  Uses jp.flags
// and jp.sub_cnt as scratch; st receives the side-list length and whether the
// build keys are unique.  Returns with the stream synchronised.
static int join_build_sub_tables(fgpu_ctx *c, JoinPartArgs &jp, const JoinBuildArgs &ba, JoinSubTables &st,
                                 bool &fits) {
    fits = false;
    const int64_t P = int64_t(1) << jp.log_p;
    HIPCHK(launch_fill_u64(c->stream, (uint64_t *)jp.flags, 8, 0));
    HIPCHK(launch_fill_u64(c->stream, (uint64_t *)jp.sub, P * jp.T * 2, (uint64_t)EMPTY_KEY));
    HIPCHK(launch_fill_u64(c->stream, (uint64_t *)jp.sub_cnt, (P + 1) / 2, 0));
    if (ba.n) {
        std::vector<unsigned int> counts((size_t)P);
        HIPCHK(launch_join_part_count(c->stream, jp, ba, grid_for(c, ba.n, BLOCK, 8)));
        HIPCHK(hipMemcpyAsync(counts.data(), jp.sub_cnt, P * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        const unsigned mx = *std::max_element(counts.begin(), counts.end());
        if (mx > (unsigned)(jp.T - jp.T / 4)) return FGPU_OK;  // skewed keys
        HIPCHK(launch_join_part_build(c->stream, jp, ba, grid_for(c, ba.n, BLOCK, 16)));
    }
    unsigned long long h[4] = {0, 0, 0, 0};
    HIPCHK(mail_copy(c, jp.flags, 32));
    HIPCHK(mail_take(c, h, 32));
    if (h[1]) return FGPU_OK;  // an insert found no slot
    st.n_min_key_rows = (int64_t)h[2];
    st.unique_keys = !h[3] && h[2] <= 1;
    fits = true;
    return FGPU_OK;
}

// Pass-1 record storage for n records in jp's format: hash64 records are RW
// int64 words each, compact records crec_words(RW, pw) uint32 words.
static int join_cells(fgpu_ctx *c, JoinPartArgs &jp, int64_t n) {
    void *p;
    const size_t bytes = jp.kw == 4 ? (size_t)n * 4 * crec_words(jp.RW, jp.pw) : (size_t)n * jp.RW * 8;
    if (int rc = buf(c, B_CELLS, bytes + 64, &p)) return rc;
    jp.cells = (int64_t *)p;
    return FGPU_OK;
}

static int join_agg_partitioned(fgpu_ctx *c, const fgpu_table *probe, const fgpu_table *build,
                                const fgpu_join_index *idx, const fgpu_join_spec *spec, JoinProbeArgs &a,
                                const AggLayout &L) {
    const Tuning &tu = c->tune;
    if (!tu.join_partition || probe->n_rows < tu.join_part_min_rows) return -1;
    // probe payload columns (non-NULL, at most 2) and build columns the aggregates read
    std::vector<int> pcols_used, bcols_used;
    for (int i = 0; i < spec->n_aggs; ++i) {
        const fgpu_agg &g = spec->aggs[i];
        if (g.col < 0) continue;
        auto &v = g.side == 1 ? bcols_used : pcols_used;
        if (std::find(v.begin(), v.end(), g.col) == v.end()) v.push_back(g.col);
    }
    if (pcols_used.size() > 2) return -1;
    for (int col : pcols_used)
        if (probe->cols[col].null_count) return -1;
    JoinPartArgs jp{};
    jp.pa = a;
    jp.n_pay = (int)pcols_used.size();
    Binder pb(probe), bb(build);  // slots must match those in `a`: rebind identically
    (void)pb;
    (void)bb;
    for (int k = 0; k < jp.n_pay; ++k) {
        // find the slot `a` uses for this column
        int slot = -1;
        for (int s2 = 0; s2 < a.pcols.n; ++s2)
            if (a.pcols.v[s2] == probe->cols[pcols_used[k]].d_values) slot = s2;
        if (slot < 0) return -1;
        jp.pay_col[k] = (int8_t)slot;
    }
    for (int i = 0; i < L.ops.n; ++i) {
        jp.field_of_op[i] = 0;
        if (L.keys[i].side == 0)
            for (int k = 0; k < jp.n_pay; ++k)
                if (L.keys[i].col == pcols_used[k]) jp.field_of_op[i] = (int8_t)(k + 1);
    }
    jp.build_pay_col = -1;
    if (bcols_used.size() == 1 && !build->cols[bcols_used[0]].null_count) {
        for (int s2 = 0; s2 < a.bcols.n; ++s2)
            if (a.bcols.v[s2] == build->cols[bcols_used[0]].d_values) jp.build_pay_col = s2;
    }
    const DevCol &bkc = build->cols[spec->build_key];
    const int64_t nb = build->n_rows;
    // Record format: compact (4-byte keys, cmix over the build key range) when
    // the build side has a non-NULL key and its range spans at most
    // 31 + log_p bits, else hash64 records.  One integer probe payload whose
    // range fits 32 bits rides as a 4-byte offset (8-byte records).  By
    // default records are compact when they hold the key alone or such a
    // payload: 12-byte compact records over twice the partitions lose to
    // 16-byte hash64 records over half of them.
    int kbits = 0;
    const bool has_keys = nb > 0 && bkc.null_count < nb && bkc.imin <= bkc.imax;
    bool narrow_pay = false;
    int64_t pay_min = 0;
    if (pcols_used.size() == 1) {
        const DevCol &pc = probe->cols[pcols_used[0]];
        narrow_pay = (pc.dtype == FGPU_INT64 || pc.dtype == FGPU_INT32) && pc.imin <= pc.imax &&
                     (uint64_t)pc.imax - (uint64_t)pc.imin <= 0xFFFFFFFFull;
        pay_min = pc.imin;
    }
    const bool compact_ok = tu.join_compact >= 2 || (tu.join_compact == 1 && (pcols_used.empty() || narrow_pay));
    if (compact_ok && has_keys) {
        const uint64_t range = (uint64_t)bkc.imax - (uint64_t)bkc.imin;
        kbits = range ? 64 - __builtin_clzll(range) : 1;
    }
    // P partitions, each with a sub-table that pass 2 holds in LDS: at most
    // tmax slots (a power of two: 8,192 of key + build value, 16,384 of key
    // alone, so that form needs half the partitions and pass 1 writes longer
    // runs), at load <= 0.5 on average (checked <= 0.75 after counting).
    const int64_t slot_bytes = bcols_used.empty() ? 8 : 16;
    const int64_t tmax = std::min<int64_t>(131072 / slot_bytes,
                                           next_pow2(((int64_t)tu.lds_limit - 1024) / slot_bytes + 1) / 2);
    const int64_t spk = std::max(2, tu.join_slots_per_key);  // slots per build key: load <= 1 / spk
    int64_t P = std::max<int64_t>(8, next_pow2((spk * nb + tmax - 1) / tmax));
    if (tu.join_partitions) P = std::max<int64_t>(P, next_pow2(tu.join_partitions));
    jp.log_p = 0;
    while ((int64_t(1) << jp.log_p) < P) ++jp.log_p;
    jp.kw = 8;
    jp.pw = 8;
    if (kbits > 0) {
        // compact records keep kbits - log_p <= 31 bits: more partitions if needed
        const int lp = std::max(jp.log_p, kbits - 31);
        if (lp < kbits && lp <= 12) {
            jp.kw = 4;
            jp.log_p = lp;
            P = int64_t(1) << lp;
            jp.kbits = kbits;
            jp.rbits = kbits - lp;
            jp.kmin = bkc.imin;
            jp.kmax = bkc.imax;
            jp.pw = narrow_pay ? 4 : 8;
            jp.pmin = pay_min;
        }
    }
    if (P > 4096) return -1;
    jp.T = (int)std::max<int64_t>(256, std::min<int64_t>(tmax, next_pow2((spk * nb + P - 1) / P)));
    jp.RW = 1 + jp.n_pay;
    jp.need_vals = bcols_used.empty() ? 0 : 1;
    // pass 1: as many workgroups per CU as LDS allows (at most the tuned count);
    // cells of P partitions plus the NULL-key and sentinel-key side partitions
    if (join_scatter_lds(jp, 0) > (size_t)tu.lds_limit) return -1;
    const int64_t fit1 = std::max<int64_t>(1, std::min<int64_t>(tu.join_part_wgs_per_cu,
                                                                 (int64_t)tu.lds_limit / (int64_t)join_scatter_lds(jp, 0)));
    jp.G = (int)(c->num_cus * fit1);
    jp.lds_budget = (int)(tu.lds_limit / fit1);
    if (join_probe_lds(jp) > (size_t)tu.lds_limit) return -1;
    jp.rows_per_wg = (probe->n_rows + jp.G - 1) / jp.G;
    jp.Cw = ((int)std::ceil((double)jp.rows_per_wg / (double)P * 1.15) + 64 + 15) / 16 * 16;  // whole lines
    if (jp.Cw >= (1 << 17)) return -1;  // a record's position in its cell is packed in 17 bits
    const int64_t fit2 = std::max<int64_t>(1, std::min<int64_t>(tu.join_probe_wgs_per_cu,
                                                                 (int64_t)tu.lds_limit / (int64_t)join_probe_lds(jp)));
    jp.split = (int)std::max<int64_t>(1, (int64_t)c->num_cus * fit2 / P);
    const int64_t P2 = P + 2;
    int rc = FGPU_OK;
    void *p;
    if ((rc = join_cells(c, jp, (int64_t)jp.G * P2 * jp.Cw))) return rc;
    if ((rc = buf(c, B_CELLCNT, (size_t)jp.G * P2 * 4, &p))) return rc;
    jp.cell_cnt = (unsigned int *)p;
    unsigned long long *d_flags;
    if ((rc = buf(c, B_COUNTERS, 64 * 8, (void **)&d_flags))) return rc;
    jp.flags = d_flags;
    if ((rc = buf(c, B_SUBCNT, (size_t)P * 4 + 64, &p))) return rc;
    jp.sub_cnt = (unsigned int *)p;
    // the partitioned path keys on hash64(key): the one key whose hash is the
    // empty-slot marker keeps its build rows in a side list
    const int64_t sentinel = (int64_t)unhash64((uint64_t)EMPTY_KEY);
    const bool has_min = jp.kw == 8 && bkc.imin <= sentinel && sentinel <= bkc.imax;
    jp.min_key_cap = has_min ? std::max<int64_t>(nb, 1) : 0;
    JoinBuildArgs ba{};
    ba.keys_col = bkc.d_values;
    ba.key_dtype = bkc.dtype;
    ba.key_valid = bkc.d_valid;
    ba.n = nb;
    const void *val_col = jp.build_pay_col >= 0 ? a.bcols.v[jp.build_pay_col] : nullptr;
    mark(c, 4);
    JoinSubTables built;
    bool ok = false;
    if (idx) {  // built once per shape and kept by the index
        std::lock_guard<std::mutex> lk(idx->sub_mu);
        const JoinSubTables *hit = nullptr;
        for (const auto &st : idx->subs)
            if (st.log_p == jp.log_p && st.T == jp.T && st.kw == jp.kw && st.val_col == val_col) hit = &st;
        if (!hit) {
            JoinSubTables st;
            st.log_p = jp.log_p;
            st.T = jp.T;
            st.kw = jp.kw;
            st.val_col = val_col;
            const size_t sub_bytes = (size_t)P * jp.T * 16, min_bytes = (size_t)jp.min_key_cap * 4;
            HIPCHK(hipMalloc(&st.sub, sub_bytes));
            if (min_bytes && hipMalloc(&st.min_key_rows, min_bytes) != hipSuccess) {
                (void)hipFree(st.sub);
                return set_err(FGPU_E_OUT_OF_MEMORY, "join sub-tables: %zu bytes", min_bytes);
            }
            jp.sub = st.sub;
            jp.min_key_rows = st.min_key_rows;
            bool fits = false;
            rc = join_build_sub_tables(c, jp, ba, st, fits);
            if (rc || !fits) {
                (void)hipFree(st.sub);
                if (st.min_key_rows) (void)hipFree(st.min_key_rows);
                st.sub = nullptr;
                st.min_key_rows = nullptr;
                if (rc) return rc;
            } else {
                idx->bytes += (int64_t)(sub_bytes + min_bytes);
            }
            idx->subs.push_back(st);
            hit = &idx->subs.back();
        }
        ok = hit->sub != nullptr;
        built = *hit;
    } else {  // no index: the context's scratch, rebuilt per call
        if ((rc = buf(c, B_SUB, (size_t)P * jp.T * 16, &p))) return rc;
        built.sub = (int64_t *)p;
        if (jp.min_key_cap) {
            if ((rc = buf(c, B_MINKEY, (size_t)jp.min_key_cap * 4, &p))) return rc;
            built.min_key_rows = (uint32_t *)p;
        }
        jp.sub = built.sub;
        jp.min_key_rows = built.min_key_rows;
        if ((rc = join_build_sub_tables(c, jp, ba, built, ok))) return rc;
    }
    mark(c, 5);
    if (!ok) return -1;  // a sub-table would overflow: direct probe
    jp.sub = built.sub;
    jp.min_key_rows = built.min_key_rows;
    jp.pa.min_key_rows = built.min_key_rows;
    jp.pa.n_min_key_rows = built.n_min_key_rows;
    jp.pa.unique_keys = built.unique_keys;
    unsigned long long h[4] = {0, 0, 0, 0};
    HIPCHK(launch_state_init(c->stream, nullptr, 0, a.state, 1, L.ops));
    jp.xoff = nullptr;
    jp.chunks = nullptr;
    c->plan += jp.kw == 4 ? (jp.RW == 2 && jp.pw == 4 ? " records=compact pay4" : " records=compact") : " records=hash64";
    mark(c, 1);
    HIPCHK(launch_join_part_scatter(c->stream, jp, 0));
    HIPCHK(launch_join_part_probe(c->stream, jp, 0));
    mark(c, 2);
    HIPCHK(mail_copy(c, d_flags, 8));
    HIPCHK(mail_take(c, h, 8));
    if (!h[0]) return FGPU_OK;
    // A cell overflowed: the probe keys are skewed (a hot key fills its
    // partition).  Partition exactly instead: count the records of every
    // (workgroup, partition), prefix-sum them into partition-major offsets,
    // scatter to those offsets (each partition's records contiguous), and run
    // pass 2 over chunks of at most `chunk` records, so a hot partition is
    // shared by many workgroups.
    c->plan += " (skewed probe keys: exact partitions)";
    HIPCHK(launch_state_init(c->stream, nullptr, 0, a.state, 1, L.ops));
    const int64_t nc = (int64_t)jp.G * P2;
    if ((rc = buf(c, B_CELLCNT, (size_t)nc * 4, &p))) return rc;
    jp.cell_cnt = (unsigned int *)p;
    int64_t *d_off, *d_tmp;
    if ((rc = buf(c, B_OFFS, (size_t)(nc + 1) * 8, (void **)&d_off))) return rc;
    if ((rc = buf(c, B_TMP, (size_t)((nc + 4095) / 4096 + 1) * 8, (void **)&d_tmp))) return rc;
    HIPCHK(launch_join_part_scatter(c->stream, jp, 1));
    HIPCHK(launch_scan_u32(c->stream, jp.cell_cnt, nc, d_off, d_tmp));
    std::vector<int64_t> offs((size_t)nc + 1);
    HIPCHK(hipMemcpyAsync(offs.data(), d_off, (size_t)(nc + 1) * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const int64_t total = offs[(size_t)nc];
    if ((rc = join_cells(c, jp, std::max<int64_t>(total, 1)))) return rc;
    jp.xoff = d_off;
    HIPCHK(launch_join_part_scatter(c->stream, jp, 2));
    const int64_t chunk = std::max<int64_t>(16384, (total + 2 * c->num_cus - 1) / (2 * c->num_cus));
    std::vector<int64_t> ch;
    for (int64_t q = 0; q < P2; ++q) {
        const int64_t base = offs[(size_t)(q * jp.G)], tot = offs[(size_t)((q + 1) * jp.G)] - base;
        for (int64_t i = 0; i < tot; i += chunk) {
            ch.push_back(q);
            ch.push_back(base + i);
            ch.push_back(std::min(chunk, tot - i));
        }
    }
    const int n_chunks = (int)(ch.size() / 3);
    if (n_chunks) {
        int64_t *d_ch;
        if ((rc = buf(c, B_PAIRS, ch.size() * 8, (void **)&d_ch))) return rc;
        HIPCHK(hipMemcpyAsync(d_ch, ch.data(), ch.size() * 8, hipMemcpyHostToDevice, c->stream));
        jp.chunks = d_ch;
        HIPCHK(launch_join_part_probe(c->stream, jp, n_chunks));
        HIPCHK(hipStreamSynchronize(c->stream));  // ch is a local
    }
    mark(c, 2);
    return FGPU_OK;
}

static int join_agg_impl(fgpu_ctx *c, const fgpu_table *probe, const fgpu_pred *pred, const fgpu_table *build,
                         const fgpu_join_index *idx, const fgpu_join_spec *spec, fgpu_result **out);

// A join's aggregates in as few calls as the probe kernels' register
// accumulators allow: a spec that needs more is cut in two, each part run as
// its own join, and the parts' columns (one output row each) joined in order.
static int join_agg_parts(fgpu_ctx *c, const fgpu_table *probe, const fgpu_pred *pred, const fgpu_table *build,
                          const fgpu_join_index *idx, const fgpu_join_spec *spec, fgpu_result **out, int &parts) {
    int rc;
    if (spec->n_aggs <= FGPU_MAX_AGGS) {  // more aggregates than one output layout holds: split at once
        t_join_ops_over = false;
        rc = join_agg_impl(c, probe, pred, build, idx, spec, out);
        if (rc != FGPU_E_UNSUPPORTED || !t_join_ops_over || spec->n_aggs < 2) {
            ++parts;
            return rc;
        }
        g_err.clear();
    }
    const int h = spec->n_aggs / 2;
    fgpu_join_spec s1 = *spec, s2 = *spec;
    s1.n_aggs = h;
    s2.aggs = spec->aggs + h;
    s2.n_aggs = spec->n_aggs - h;
    fgpu_result *r1 = nullptr, *r2 = nullptr;
    if ((rc = join_agg_parts(c, probe, pred, build, idx, &s1, &r1, parts))) return rc;
    if ((rc = join_agg_parts(c, probe, pred, build, idx, &s2, &r2, parts))) {
        delete r1;
        return rc;
    }
    for (size_t j = 0; j < r2->cols.size(); ++j) {
        RCol &col = r2->cols[j];
        col.name = alias_of(spec->aggs[h + j], h + (int)j);  // the output position in the whole spec
        if (col.xv) {  // values in r2's pinned block: copy the one row out
            if (col.dtype == FGPU_FLOAT64) col.f64.assign(1, *(const double *)col.xv);
            else if (col.dtype == FGPU_UTF8) col.u32.assign(1, *(const uint32_t *)col.xv);
            else col.i64.assign(1, *(const int64_t *)col.xv);
            if (col.xvalid) col.valid.assign(1, *col.xvalid);
            col.xv = nullptr;
            col.xvalid = nullptr;
        }
        r1->cols.push_back(std::move(col));
    }
    delete r2;
    *out = r1;
    return FGPU_OK;
}

int fgpu_join_agg(fgpu_ctx *c, const fgpu_table *probe, const fgpu_pred *pred, const fgpu_table *build,
                  const fgpu_join_index *idx, const fgpu_join_spec *spec, fgpu_result **out) {
    if (!c || !probe || !build || !spec || !out || (spec->n_aggs > 0 && !spec->aggs))
        return set_err(FGPU_E_INVALID_QUERY, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    if (int rc = set_device(c->device)) return rc;
    int parts = 0;
    const int rc = join_agg_parts(c, probe, pred, build, idx, spec, out, parts);
    if (rc == FGPU_OK && parts > 1) c->plan += "; aggregates in " + std::to_string(parts) + " parts";
    return rc;
}

static int join_agg_impl(fgpu_ctx *c, const fgpu_table *probe, const fgpu_pred *pred, const fgpu_table *build,
                         const fgpu_join_index *idx, const fgpu_join_spec *spec, fgpu_result **out) {
    Binder pb(probe), bb(build);
    JoinProbeArgs a;
    AggLayout L;
    struct OwnIndex {  // an index this call builds, freed on every return
        fgpu_join_index *p = nullptr;
        ~OwnIndex() { fgpu_join_index_free(p); }
    } own_index;
    fgpu_join_index *&own = own_index.p;
    // An exact integer SUM is sized by the join's output rows.  Without an
    // index the build keys may repeat, bounding INNER / LEFT output by
    // probe x build rows; past 2^32 rows the sum could not be laid out.  An
    // index built here tells whether the keys are unique (then the output
    // has at most one row per probe row), and the call then uses it.
    if (!idx && (spec->type == FGPU_INNER || spec->type == FGPU_LEFT) &&
        (__int128)probe->n_rows * std::max<int64_t>(build->n_rows, 1) >= ((__int128)1 << 32)) {
        bool int_sum = false;
        for (int i = 0; i < spec->n_aggs; ++i) {
            const fgpu_agg &g = spec->aggs[i];
            const fgpu_table *tt = g.side == 1 ? build : probe;
            if ((g.func == FGPU_SUM || g.func == FGPU_AVG) && g.col >= 0 && g.col < (int)tt->cols.size() &&
                tt->cols[g.col].dtype != FGPU_FLOAT64)
                int_sum = true;
        }
        if (int_sum) {
            if (int rc = join_build_impl(c, build, spec->build_key, &own)) return rc;
            idx = own;
        }
    }
    if (int rc = join_prepare(c, probe, pred, build, idx, spec, pb, bb, a, L)) return rc;
    std::string zone_note;  // the probe scans skip the batches that cannot pass
    if (int rc = prune_batches(c, probe, pred, a.pred, &zone_note)) return rc;
    mark(c, 0);
    uint64_t *d_state;
    int rc = buf(c, B_STATE, (size_t)a.S * 8, (void **)&d_state);
    if (rc) return rc;
    a.state = d_state;
    hipError_t e = hipSuccess;
    c->plan.clear();
    const int prc = join_agg_partitioned(c, probe, build, idx, spec, a, L);
    if (prc > 0) return prc;
    c->plan = (prc == 0 ? "join:partitioned" + c->plan : "join:direct") + zone_note;
    if (prc < 0) {  // direct probe of the open-addressing index
        if (!idx) {
            if (int rc2 = join_build_impl(c, build, spec->build_key, &own)) return rc2;
            idx = own;
        }
        a.slots = idx->slots;
        a.cap_mask = (uint64_t)(idx->cap - 1);
        a.min_key_rows = idx->min_key_rows;
        a.n_min_key_rows = idx->n_min_key_rows;
        a.unique_keys = idx->unique_keys;
        e = launch_state_init(c->stream, nullptr, 0, d_state, 1, L.ops);
        const int grid = grid_for(c, (probe->n_rows + 3) / 4, c->tune.join_block, c->tune.join_waves_per_cu);
        mark(c, 1);
        if (e == hipSuccess && probe->n_rows) e = launch_join_probe_agg(c->stream, a, grid, c->tune.join_block);
        mark(c, 2);
    }
    if (e != hipSuccess) return set_err(FGPU_E_DEVICE, "join probe: %s", hipGetErrorString(e));
    OutBufs ob;
    rc = out_bufs(c, 1, L.outs.n, ob);
    if (rc) return rc;
    EmitArgs em{};
    em.outs = L.outs;
    em.state = d_state;
    em.S = a.S;
    em.n_slots = 1;
    em.mode = GB_SCALAR;
    em.out_key = nullptr;
    for (int i = 0; i < FGPU_MAX_AGGS; ++i) {
        em.out_val[i] = ob.val[i];
        em.out_valid[i] = ob.valid[i];
    }
    em.out_count = ob.count;
    e = launch_fill_u64(c->stream, (uint64_t *)ob.count, 1, 0);
    if (e == hipSuccess) e = launch_emit(c->stream, em);
    mark(c, 3);
    if (e != hipSuccess) return set_err(FGPU_E_DEVICE, "join emit: %s", hipGetErrorString(e));
    rc = collect(c, L, 1, nullptr, 0, "", nullptr, nullptr, ob.val, ob.valid, ob.count, out);
    int64_t bytes = 0;
    for (int s = 0; s < pb.dc.n; ++s) bytes += probe->n_rows * (pb.dc.dtype[s] == FGPU_INT64 || pb.dc.dtype[s] == FGPU_FLOAT64 ? 8 : 4);
    for (int s = 0; s < bb.dc.n; ++s) bytes += build->n_rows * (bb.dc.dtype[s] == FGPU_INT64 || bb.dc.dtype[s] == FGPU_FLOAT64 ? 8 : 4);
    if (!rc) finish_timing(c, bytes, own != nullptr);
    return rc;
}

// The (probe row, build row) pairs of a join on the device, in the
// context's B_PAIRS slot: *d_l and *d_r hold *total rows each (NO_ROW for the
// NULL-extended side).  Rows come in probe order.  The caller holds c->mu.
static int join_pairs_dev(fgpu_ctx *c, const fgpu_table *probe, const fgpu_pred *pred, const fgpu_table *build,
                          const fgpu_join_index *idx, const fgpu_join_spec *spec, uint32_t **d_l, uint32_t **d_r,
                          int64_t *total_out) {
    Binder pb(probe), bb(build);
    JoinProbeArgs a;
    AggLayout L;
    fgpu_join_spec s2 = *spec;
    s2.n_aggs = 0;
    if (int rc = join_prepare(c, probe, pred, build, idx, &s2, pb, bb, a, L)) return rc;
    fgpu_join_index *own = nullptr;
    if (!idx) {
        if (int rc = join_build_impl(c, build, spec->build_key, &own)) return rc;
        idx = own;
    }
    a.slots = idx->slots;
    a.cap_mask = (uint64_t)(idx->cap - 1);
    a.min_key_rows = idx->min_key_rows;
    a.n_min_key_rows = idx->n_min_key_rows;
    const int64_t n = probe->n_rows;
    int rc = FGPU_OK;
    uint32_t *d_cnt = nullptr;
    int64_t *d_offs = nullptr, *d_tmp = nullptr;
    if (!rc) rc = buf(c, B_CNT, (size_t)std::max<int64_t>(n, 1) * 4, (void **)&d_cnt);
    if (!rc) rc = buf(c, B_OFFS, (size_t)(n + 1) * 8, (void **)&d_offs);
    if (!rc) rc = buf(c, B_TMP, (size_t)(n / 4096 + 2) * 8, (void **)&d_tmp);
    if (rc) {
        fgpu_join_index_free(own);
        return rc;
    }
    a.counts = d_cnt;
    const int grid = grid_for(c, n, BLOCK, 16);
    hipError_t e = n ? launch_join_pairs(c->stream, a, 0, grid) : hipSuccess;
    if (e == hipSuccess) e = launch_scan_u32(c->stream, d_cnt, n, d_offs, d_tmp);
    int64_t total = 0;
    if (e == hipSuccess) e = mail_copy(c, d_offs + n, 8);
    if (e == hipSuccess) e = mail_take(c, &total, 8);
    if (e != hipSuccess) {
        fgpu_join_index_free(own);
        return set_err(FGPU_E_DEVICE, "join pairs: %s", hipGetErrorString(e));
    }
    uint32_t *d_pairs;
    rc = buf(c, B_PAIRS, (size_t)std::max<int64_t>(total, 1) * 8, (void **)&d_pairs);
    if (rc) {
        fgpu_join_index_free(own);
        return rc;
    }
    a.offs = d_offs;
    a.out_l = d_pairs;
    a.out_r = d_pairs + std::max<int64_t>(total, 1);
    e = (n && total) ? launch_join_pairs(c->stream, a, 1, grid) : hipSuccess;
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    fgpu_join_index_free(own);
    if (e != hipSuccess) return set_err(FGPU_E_DEVICE, "join pairs: %s", hipGetErrorString(e));
    *d_l = a.out_l;
    *d_r = a.out_r;
    *total_out = total;
    return FGPU_OK;
}

int fgpu_join_pairs(fgpu_ctx *c, const fgpu_table *probe, const fgpu_pred *pred, const fgpu_table *build,
                    const fgpu_join_index *idx, const fgpu_join_spec *spec, uint32_t *probe_rows,
                    uint32_t *build_rows, int64_t capacity, int64_t *n_out) {
    if (!c || !probe || !build || !spec || !n_out) return set_err(FGPU_E_INVALID_QUERY, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    if (int rc = set_device(c->device)) return rc;
    uint32_t *d_l = nullptr, *d_r = nullptr;
    int64_t total = 0;
    if (int rc = join_pairs_dev(c, probe, pred, build, idx, spec, &d_l, &d_r, &total)) return rc;
    *n_out = total;
    if (total > capacity || (total > 0 && (!probe_rows || !build_rows)))
        return set_err(FGPU_E_INDEX_OUT_OF_BOUNDS, "join produces %lld pairs, capacity %lld", (long long)total,
                       (long long)capacity);
    std::vector<uint32_t> l(total), r(total);
    hipError_t e = hipSuccess;
    if (total) e = hipMemcpyAsync(l.data(), d_l, total * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && total) e = hipMemcpyAsync(r.data(), d_r, total * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return set_err(FGPU_E_DEVICE, "join pairs: %s", hipGetErrorString(e));
    // within one probe row, list build rows in ascending order
    int64_t i = 0;
    while (i < total) {
        int64_t j = i + 1;
        while (j < total && l[j] == l[i]) ++j;
        if (j - i > 1) std::sort(r.begin() + i, r.begin() + j);
        i = j;
    }
    if (total) {
        std::memcpy(probe_rows, l.data(), total * 4);
        std::memcpy(build_rows, r.data(), total * 4);
    }
    return FGPU_OK;
}

// Copies `m` rows of table columns, gathered at device row numbers d_rows,
// into a host result (values, and validity bytes from validity words).
static int result_from_rows(fgpu_ctx *c, const fgpu_table *t, const uint32_t *d_rows, int64_t m, const int32_t *cols,
                            int n_cols, fgpu_result **out) {
    void *d_out;
    const int64_t nw = (m + 63) / 64;
    if (int rc = buf(c, B_OUT, (size_t)std::max<int64_t>(m, 1) * 8 + (size_t)std::max<int64_t>(nw, 1) * 8 + 64, &d_out))
        return rc;
    uint64_t *d_vw = (uint64_t *)((char *)d_out + (size_t)std::max<int64_t>(m, 1) * 8);
    auto *r = new fgpu_result();
    r->rows = m;
    const int grid = grid_for(c, (m + 63) / 64 * 64, BLOCK, 8);
    std::vector<uint64_t> vw((size_t)nw);
    for (int i = 0; i < n_cols; ++i) {
        const DevCol &dc = t->cols[cols[i]];
        RCol rc2;
        rc2.name = dc.name;
        rc2.dtype = dc.dtype == FGPU_FLOAT64 ? FGPU_FLOAT64 : dc.dtype == FGPU_UTF8 ? FGPU_UTF8 : FGPU_INT64;
        const int eb = (int)dc.elem_bytes();
        hipError_t e = m ? launch_gather_idx(c->stream, dc.d_values, eb, dc.dtype == FGPU_UTF8, dc.d_valid, d_rows, m,
                                             d_out, d_vw, grid)
                         : hipSuccess;
        std::vector<uint32_t> tmp32;
        if (e == hipSuccess && m) {
            if (eb == 8) {
                if (rc2.dtype == FGPU_FLOAT64) {
                    rc2.f64.resize(m);
                    e = hipMemcpyAsync(rc2.f64.data(), d_out, m * 8, hipMemcpyDeviceToHost, c->stream);
                } else {
                    rc2.i64.resize(m);
                    e = hipMemcpyAsync(rc2.i64.data(), d_out, m * 8, hipMemcpyDeviceToHost, c->stream);
                }
            } else {
                tmp32.resize(m);
                e = hipMemcpyAsync(tmp32.data(), d_out, m * 4, hipMemcpyDeviceToHost, c->stream);
            }
            if (e == hipSuccess) e = hipMemcpyAsync(vw.data(), d_vw, nw * 8, hipMemcpyDeviceToHost, c->stream);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            delete r;
            return set_err(FGPU_E_DEVICE, "row gather: %s", hipGetErrorString(e));
        }
        rc2.valid.resize(m);
        for (int64_t k = 0; k < m; ++k) rc2.valid[k] = (vw[k >> 6] >> (k & 63)) & 1;
        if (eb == 4) {
            if (dc.dtype == FGPU_UTF8) {
                rc2.u32.resize(m);
                for (int64_t k = 0; k < m; ++k) rc2.u32[k] = tmp32[k];
                rc2.dict_src = &dc;
            } else {
                rc2.i64.resize(m);
                for (int64_t k = 0; k < m; ++k) rc2.i64[k] = (int32_t)tmp32[k];
            }
        }
        r->cols.push_back(std::move(rc2));
    }
    *out = r;
    return FGPU_OK;
}

int fgpu_sort_limit(fgpu_ctx *c, const fgpu_table *t, const fgpu_pred *p, const fgpu_sort_key *keys, int n_keys,
                    int64_t limit, const int32_t *cols, int n_cols, int64_t *rows_out, fgpu_result **out) {
    if (!c || !t || !out || n_keys < 1 || !keys || (n_cols > 0 && !cols) || n_cols < 0)
        return set_err(FGPU_E_INVALID_QUERY, "NULL argument");
    for (int j = 0; j < n_keys; ++j)
        if (keys[j].col < 0 || keys[j].col >= (int)t->cols.size())
            return set_err(FGPU_E_COLUMN_NOT_FOUND, "sort column %d not in table", keys[j].col);
    for (int i = 0; i < n_cols; ++i)
        if (cols[i] < 0 || cols[i] >= (int)t->cols.size())
            return set_err(FGPU_E_COLUMN_NOT_FOUND, "column %d not in table", cols[i]);
    std::lock_guard<std::mutex> lk(c->mu);
    if (int rc = set_device(c->device)) return rc;
    mark(c, 0);
    // the rows that pass p, in row order
    uint64_t *d_words;
    if (int rc = filter_mask_dev(c, t, p, &d_words)) return rc;
    const int64_t nw = (t->n_rows + 63) / 64;
    int64_t *d_offs, *d_tmp;
    if (int rc = buf(c, B_OFFS, (size_t)(nw + 1) * 8, (void **)&d_offs)) return rc;
    if (int rc = buf(c, B_TMP, (size_t)(nw / 4096 + 2) * 8, (void **)&d_tmp)) return rc;
    HIPCHK(launch_word_offsets(c->stream, d_words, nw, d_offs, d_tmp));
    int64_t n = 0;
    HIPCHK(mail_copy(c, d_offs + nw, 8));
    HIPCHK(mail_take(c, &n, 8));
    // Utf8 sort keys order by string: each code's rank in the sorted dictionary
    std::vector<uint32_t> ranks;
    std::vector<size_t> rank_off((size_t)n_keys, 0);
    for (int j = 0; j < n_keys; ++j) {
        const DevCol &dc = t->cols[keys[j].col];
        if (dc.dtype != FGPU_UTF8) continue;
        std::vector<uint32_t> order(dc.dict.size());
        for (size_t k = 0; k < order.size(); ++k) order[k] = (uint32_t)k;
        std::sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return dc.dict[x] < dc.dict[y]; });
        rank_off[j] = ranks.size();
        ranks.resize(ranks.size() + std::max<size_t>(order.size(), 1));
        uint32_t rk = 0;  // equal strings (a dictionary may repeat one) share a rank
        for (size_t pos = 0; pos < order.size(); ++pos) {
            if (pos && dc.dict[order[pos]] != dc.dict[order[pos - 1]]) rk = (uint32_t)pos;
            ranks[rank_off[j] + order[pos]] = rk;
        }
    }
    // scratch: rows, two orders, two key arrays, two flag arrays, ranks, sort temp
    size_t temp_u64 = 0, temp_u32 = 0, temp_u8 = 0;
    HIPCHK(sort_pairs_u64(nullptr, &temp_u64, nullptr, nullptr, nullptr, nullptr, std::max<int64_t>(n, 1), 64, c->stream));
    HIPCHK(sort_pairs_u64(nullptr, &temp_u32, nullptr, nullptr, nullptr, nullptr, std::max<int64_t>(n, 1), 32, c->stream));
    temp_u64 = std::max(temp_u64, temp_u32);
    HIPCHK(sort_pairs_u8(nullptr, &temp_u8, nullptr, nullptr, nullptr, nullptr, std::max<int64_t>(n, 1), c->stream));
    const size_t N = (size_t)std::max<int64_t>(n, 1);
    auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t o_rows = 0, o_ia = up(o_rows + N * 4), o_ib = up(o_ia + N * 4), o_ka = up(o_ib + N * 4),
                 o_kb = up(o_ka + N * 8), o_fa = up(o_kb + N * 8), o_fb = up(o_fa + N), o_rank = up(o_fb + N),
                 o_temp = up(o_rank + std::max<size_t>(ranks.size(), 1) * 4),
                 total = up(o_temp + std::max(temp_u64, temp_u8));
    char *base;
    if (int rc = buf(c, B_SORT, total, (void **)&base)) return rc;
    uint32_t *rows = (uint32_t *)(base + o_rows), *ia = (uint32_t *)(base + o_ia), *ib = (uint32_t *)(base + o_ib);
    uint64_t *ka = (uint64_t *)(base + o_ka), *kb = (uint64_t *)(base + o_kb);
    uint8_t *fa = (uint8_t *)(base + o_fa), *fb = (uint8_t *)(base + o_fb);
    uint32_t *d_rank = (uint32_t *)(base + o_rank);
    void *temp = base + o_temp;
    if (!ranks.empty())
        HIPCHK(hipMemcpyAsync(d_rank, ranks.data(), ranks.size() * 4, hipMemcpyHostToDevice, c->stream));
    mark(c, 1);
    if (n) {
        HIPCHK(launch_mask_rows(c->stream, d_words, d_offs, nw, rows));
        HIPCHK(launch_iota_u32(c->stream, ia, n));
        for (int j = n_keys - 1; j >= 0; --j) {  // stable passes, last key first
            const fgpu_sort_key &sk = keys[j];
            const DevCol &dc = t->cols[sk.col];
            HIPCHK(launch_sort_key(c->stream, dc.d_values, dc.dtype, dc.d_valid, d_rank + rank_off[j], rows, n,
                                   sk.desc != 0, sk.nulls_last != 0, ka, fa));
            HIPCHK(launch_take_u64(c->stream, ka, ia, n, kb));
            size_t tb = temp_u64;
            const int bits = (dc.dtype == FGPU_UTF8 || dc.dtype == FGPU_INT32) ? 32 : 64;  // see k_sort_key
            HIPCHK(sort_pairs_u64(temp, &tb, kb, ka, ia, ib, n, bits, c->stream));
            std::swap(ia, ib);
            if (dc.null_count > 0) {  // NULL rows first or last
                HIPCHK(launch_take_u8(c->stream, fa, ia, n, fb));
                tb = temp_u8;
                HIPCHK(sort_pairs_u8(temp, &tb, fb, fa, ia, ib, n, c->stream));
                std::swap(ia, ib);
            }
        }
        HIPCHK(launch_take_u32(c->stream, rows, ia, n, ib));  // row numbers in sorted order
    }
    mark(c, 2);
    const int64_t m = limit < 0 ? n : std::min(limit, n);
    if (rows_out && m) {
        std::vector<uint32_t> h((size_t)m);
        HIPCHK(hipMemcpyAsync(h.data(), ib, (size_t)m * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int64_t i = 0; i < m; ++i) rows_out[i] = h[i];
    }
    int rc = result_from_rows(c, t, ib, m, cols, n_cols, out);
    mark(c, 3);
    char msg[96];
    snprintf(msg, sizeof msg, "sort:radix keys=%d rows=%lld", n_keys, (long long)n);
    c->plan = msg;
    finish_timing(c, 0, false);
    return rc;
}

int fgpu_join_materialize(fgpu_ctx *c, const fgpu_table *probe, const fgpu_pred *pred, const fgpu_table *build,
                          const fgpu_join_index *idx, const fgpu_join_spec *spec, const int32_t *probe_cols,
                          int n_probe_cols, const int32_t *build_cols, int n_build_cols, fgpu_table **out) {
    if (!c || !probe || !build || !spec || !out || (n_probe_cols > 0 && !probe_cols) ||
        (n_build_cols > 0 && !build_cols) || n_probe_cols < 0 || n_build_cols < 0)
        return set_err(FGPU_E_INVALID_QUERY, "NULL argument");
    for (int i = 0; i < n_probe_cols; ++i)
        if (probe_cols[i] < 0 || probe_cols[i] >= (int)probe->cols.size())
            return set_err(FGPU_E_COLUMN_NOT_FOUND, "probe column %d not in table", probe_cols[i]);
    for (int i = 0; i < n_build_cols; ++i)
        if (build_cols[i] < 0 || build_cols[i] >= (int)build->cols.size())
            return set_err(FGPU_E_COLUMN_NOT_FOUND, "build column %d not in table", build_cols[i]);
    if (n_build_cols && (spec->type == FGPU_SEMI || spec->type == FGPU_ANTI))
        return set_err(FGPU_E_COLUMN_NOT_FOUND, "SEMI/ANTI output has only probe columns");
    std::lock_guard<std::mutex> lk(c->mu);
    if (int rc = set_device(c->device)) return rc;
    uint32_t *d_l = nullptr, *d_r = nullptr;
    int64_t total = 0;
    if (int rc = join_pairs_dev(c, probe, pred, build, idx, spec, &d_l, &d_r, &total)) return rc;
    if (total >= (int64_t(1) << 31))
        return set_err(FGPU_E_UNSUPPORTED, "join output of %lld rows (tables are limited to 2^31-1 rows)",
                       (long long)total);
    auto *t = new fgpu_table();
    t->device = c->device;
    t->n_rows = total;
    t->batch_rows = probe->batch_rows;
    t->cols.reserve((size_t)(n_probe_cols + n_build_cols));
    auto fail = [&](int rc) {
        fgpu_table_free(t);
        return rc;
    };
    ColStatsDev *d_st = nullptr;
    if (hipMalloc(&d_st, sizeof(ColStatsDev)) != hipSuccess) return fail(set_err(FGPU_E_OUT_OF_MEMORY, "stats buffer"));
    const int grid = grid_for(c, (total + 63) / 64 * 64, BLOCK, 8);
    int rc = FGPU_OK;
    for (int k = 0; k < n_probe_cols + n_build_cols && rc == FGPU_OK; ++k) {
        const bool right = k >= n_probe_cols;
        const DevCol &src = right ? build->cols[build_cols[k - n_probe_cols]] : probe->cols[probe_cols[k]];
        const uint32_t *rows = right ? d_r : d_l;
        DevCol col;
        col.name = src.name;
        if (right)  // output_plan: a right column whose name the left side has becomes right.<name>
            for (const DevCol &lc : probe->cols)
                if (strcasecmp(lc.name.c_str(), src.name.c_str()) == 0) col.name = "right." + src.name;
        col.dtype = src.dtype;
        col.dict = src.dict;
        col.empty_code = src.empty_code;
        for (auto &str : col.dict) col.dictp.push_back(str.c_str());
        t->cols.push_back(std::move(col));
        DevCol &d = t->cols.back();
        const size_t bytes = (size_t)std::max<int64_t>(total, 1) * d.elem_bytes();
        if (hipMalloc(&d.d_values, bytes) != hipSuccess) {
            rc = set_err(FGPU_E_OUT_OF_MEMORY, "join output column '%s': %zu bytes", d.name.c_str(), bytes);
            break;
        }
        t->bytes += (int64_t)bytes;
        // a build column of a LEFT join is NULL where no build row matched
        const bool nullable = src.dtype != FGPU_UTF8 && (src.d_valid || (right && spec->type == FGPU_LEFT));
        if (nullable) {
            const size_t vb = (size_t)std::max<int64_t>((total + 63) / 64, 1) * 8;
            if (hipMalloc(&d.d_valid, vb) != hipSuccess) {
                rc = set_err(FGPU_E_OUT_OF_MEMORY, "join output validity");
                break;
            }
            t->bytes += (int64_t)vb;
        }
        hipError_t e = total ? launch_gather_idx(c->stream, src.d_values, (int)src.elem_bytes(), src.dtype == FGPU_UTF8,
                                                 src.d_valid, rows, total, d.d_values, d.d_valid, grid)
                             : hipSuccess;
        if (e != hipSuccess) {
            rc = set_err(FGPU_E_DEVICE, "join output gather: %s", hipGetErrorString(e));
            break;
        }
        rc = column_stats(c, d, total, t->batch_rows, d_st);
    }
    (void)hipFree(d_st);
    if (rc) return fail(rc);
    for (DevCol &d : t->cols) {  // dictionary views of the columns in their final place
        d.dictp.clear();
        for (auto &str : d.dict) d.dictp.push_back(str.c_str());
    }
    c->plan = "join:materialize";
    *out = t;
    return FGPU_OK;
}

// ─── results ────────────────────────────────────────────────────────────────
int64_t fgpu_result_rows(const fgpu_result *r) { return r ? r->rows : -1; }
int fgpu_result_cols(const fgpu_result *r) { return r ? (int)r->cols.size() : -1; }
int fgpu_result_column(const fgpu_result *r, int i, fgpu_column_view *v) {
    if (!r || !v) return set_err(FGPU_E_INVALID_QUERY, "NULL argument");
    if (i < 0 || i >= (int)r->cols.size()) return set_err(FGPU_E_INDEX_OUT_OF_BOUNDS, "column %d of %zu", i, r->cols.size());
    const RCol &c = r->cols[i];
    v->name = c.name.c_str();
    v->dtype = c.dtype;
    if (c.xv) {
        v->values = c.xv;
        v->valid = c.xvalid;
    } else {
        v->values = c.dtype == FGPU_FLOAT64 ? (const void *)c.f64.data()
                  : c.dtype == FGPU_UTF8   ? (const void *)c.u32.data()
                                           : (const void *)c.i64.data();
        v->valid = c.valid.empty() ? nullptr : c.valid.data();
    }
    v->dict_len = c.dict_src ? (int32_t)c.dict_src->dictp.size() : 0;
    v->dict = c.dict_src ? c.dict_src->dictp.data() : nullptr;
    return FGPU_OK;
}
void fgpu_result_free(fgpu_result *r) { delete r; }

// ─── host-side check of the exact float accumulator ─────────────────────────
// Sums x[0..n) with the same digit decomposition and normalisation the kernels
// use (LSB exponent L and word count W derived from the data like an upload
// would), entirely on the host.  Returns the exactly rounded sum in *out.
int fgpu_host_exact_sum(const double *x, int64_t n, double *out) {
    if (!out || (n > 0 && !x)) return set_err(FGPU_E_INVALID_QUERY, "NULL argument");
    int emin = 1 << 20, emax = -(1 << 20);
    uint64_t flags = 0;
    for (int64_t i = 0; i < n; ++i) {
        const uint64_t b = __builtin_bit_cast(uint64_t, x[i]);
        const int be = (int)((b >> 52) & 0x7ff);
        if (be == 0x7ff) {
            flags |= (b & 0x000fffffffffffffull) ? FL_NAN : (b >> 63) ? FL_NINF : FL_PINF;
            continue;
        }
        if ((b & 0x7fffffffffffffffull) == 0) continue;
        const int e = be == 0 ? -1074 : be - 1075;
        emin = std::min(emin, e);
        emax = std::max(emax, e);
    }
    DevCol dc;
    dc.name = "x";
    dc.emin = emin;
    dc.emax = emax;
    int L, W;
    if (int rc = fx_words(dc, L, W, true)) return rc;
    int64_t w[MAX_FX_WORDS] = {0};
    for (int64_t i = 0; i < n; ++i) {
        const uint64_t b = __builtin_bit_cast(uint64_t, x[i]);
        if (((b >> 52) & 0x7ff) == 0x7ff) continue;
        int w0;
        int64_t d0, d1, d2;
        if (!fx_digits(b, L, w0, d0, d1, d2)) continue;
        w[w0] += d0;
        if (w0 + 1 < W) w[w0 + 1] += d1;
        if (w0 + 2 < W) w[w0 + 2] += d2;
    }
    *out = fx_finish(w, W, L, flags);
    return FGPU_OK;
}

}  // extern "C"

// ─── batched GROUP BY ───────────────────────────────────────────────────────
// The knobs the autotuner sweeps: workgroup size, rows per pass-1 unit and the
// LDS table size of pass 2 (which is also the key-range fan-out of BQ_PART).
struct BatchCfg {
    int block = 1024;
    int64_t unit_rows = int64_t(1) << 18;
    int slots = 2048;
    int wgs_per_cu = 8;  // cap on the resident workgroups per CU of each pass
};

struct BatchQ {
    const fgpu_table *t = nullptr;
    int fallback = 0;  // a shape the batch kernels do not take: runs through fgpu_groupby
    std::vector<fgpu_pred_node> nodes;
    std::vector<std::string> strs;  // aliases the copies point at
    fgpu_host::PredCopies pred_copies;  // literal strings and IN lists
    std::vector<fgpu_agg> aggs;
    fgpu_pred pred{};
    bool has_pred = false;
    fgpu_groupby_spec spec{};
    AggLayout L;
    const DevCol *key = nullptr;
    GbArgs g{};
    int has_valid_word = 0;
    int pack = 0, pk_kbits = 0;  // BQ_PART packed records (BatchQuery::pack)
    int64_t pk_vmin = 0;
    int64_t Cw_need = 0;  // cell size seen by an overflowing run (0 = from the row count)
    BatchQuery q{};
};

struct fgpu_batch {
    fgpu_ctx *c = nullptr;
    std::vector<BatchQ> qs;
    BatchCfg cfg;
    int npc = 1;
    int nin = 0;  // the most aggregate inputs of a batched query
    void *mem = nullptr;
    BatchQuery *d_qs = nullptr;
    BatchUnit *d_units = nullptr;
    MergeUnit *d_units2 = nullptr;
    unsigned long long *d_work = nullptr, *d_flags = nullptr, *d_counts = nullptr;
    int64_t n_units = 0, n_units2 = 0;
    size_t lds1 = 0, lds2 = 0;
    int per_cu1 = 1, per_cu2 = 1;  // persistent workgroups per CU of pass 1 / pass 2
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string info;
    float last_ms = 0;
};

namespace {

void batch_release(fgpu_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    if (b->mem) (void)hipFree(b->mem);
    b->mem = nullptr;
}

// Byte width of a column's cells for the specialised pass-1 tile (no NULLs;
// INT32 4, INT64 / FLOAT64 8), else 0.
int fast_width(const DCols &c, int k) {
    if (c.valid[k]) return 0;
    if (c.dtype[k] == FGPU_INT32) return 4;
    if (c.dtype[k] == FGPU_INT64 || c.dtype[k] == FGPU_FLOAT64) return 8;
    return 0;
}

// Sets q.fast and the widths when a BQ_PART query takes the specialised tile.
void batch_fast_tile(BatchQuery &q) {
    const GbArgs &g = q.g;
    q.fast = 0;
    if (q.mode != BQ_PART || q.has_valid_word || q.RW < 1 || q.RW > 2 || g.n_in < 1 || g.n_in > 2) return;
    const int kw = fast_width(g.cols, g.key_col);
    if (kw == 0 || g.cols.dtype[g.key_col] == FGPU_FLOAT64) return;
    int pw = 0;
    if (g.pred.n != 0) {
        const DPredNode &nd = g.pred.nodes[0];
        if (g.pred.n != 1 || g.pred.n_pcols != 1 || nd.kind != FGPU_P_CMP || nd.pc != 0) return;
        pw = fast_width(g.cols, g.pred.pcol[0]);
        const int dt = g.cols.dtype[g.pred.pcol[0]];
        const bool cls_ok = (nd.cls == CMP_F64 && dt == FGPU_FLOAT64) || (nd.cls == CMP_I64 && dt == FGPU_INT64) ||
                            (nd.cls == CMP_I64_AS_F64 && (dt == FGPU_INT64 || dt == FGPU_INT32)) ||
                            (nd.cls == CMP_I32 && dt == FGPU_INT32);
        if (pw == 0 || !cls_ok) return;
    }
    const int iw0 = fast_width(g.cols, g.in_col[0]);
    const int iw1 = g.n_in > 1 ? fast_width(g.cols, g.in_col[1]) : 8;
    if (iw0 == 0 || iw1 == 0) return;
    q.fast = 1;
    q.fast_kw = kw;
    q.fast_pw = pw;
    q.fast_iw0 = iw0;
    q.fast_iw1 = iw1;
}

// Lays out every batched query for b->cfg, allocates the plan's device memory
// in one piece, uploads the descriptors and captures the run as a graph.
int batch_plan(fgpu_batch *b) {
    fgpu_ctx *c = b->c;
    batch_release(b);
    const BatchCfg &cfg = b->cfg;
    const size_t lim = (size_t)c->tune.lds_limit - 1024;  // static shared of the kernels
    struct Piece {
        size_t off = 0, bytes = 0;
    };
    size_t total = 0;
    auto take = [&](size_t bytes) {
        Piece pc;
        pc.off = total;
        pc.bytes = bytes;
        total += (bytes + 255) & ~size_t(255);
        return pc;
    };
    std::vector<BatchUnit> units;
    std::vector<MergeUnit> units2;
    struct QPieces {
        Piece cells, cnt, blocks, nulls, key, kvalid, val[FGPU_MAX_AGGS], valid[FGPU_MAX_AGGS];
    };
    std::vector<QPieces> pieces(b->qs.size());
    b->lds1 = b->lds2 = 8;
    b->npc = 1;
    b->nin = 0;
    for (size_t qi = 0; qi < b->qs.size(); ++qi) {
        BatchQ &Q = b->qs[qi];
        if (Q.fallback) continue;
        BatchQuery &q = Q.q;
        q = BatchQuery{};
        q.g = Q.g;
        q.outs = Q.L.outs;
        const int S = Q.L.ops.S;
        q.S = S;
        q.has_valid_word = Q.has_valid_word;
        q.pack = Q.pack;
        q.pk_kbits = Q.pk_kbits;
        q.pk_vmin = Q.pk_vmin;
        q.RW = (Q.pack ? 0 : 1) + Q.g.n_in + Q.has_valid_word;
        q.fast = 0;
        const DevCol &key = *Q.key;
        q.key_min = key.imin <= key.imax ? key.imin : 0;
        q.range = key.imin <= key.imax ? key.imax - key.imin + 1 : 1;
        const int64_t n_rows = Q.t->n_rows;
        q.n_units = std::max<int64_t>(1, (n_rows + cfg.unit_rows - 1) / cfg.unit_rows);
        const int64_t per_unit = (n_rows + q.n_units - 1) / q.n_units;
        // pass-2 slots: a multiple-of-two table that fits next to the G + 1 prefix
        int64_t slots = cfg.slots;
        // (each slot also has a 4-byte entry in the emit list)
        while (slots > 64 && (size_t)slots * (S * 8 + 4) + (size_t)(q.n_units + 1) * 4 > lim) slots >>= 1;
        if ((size_t)(S + q.range * S) * 8 <= std::min<size_t>(lim, 64 * 1024)) {
            q.mode = BQ_DENSE;
            b->lds1 = std::max(b->lds1, (size_t)(S + q.range * S) * 8);
            for (int64_t k0 = 0; k0 < q.range; k0 += slots)
                units2.push_back(MergeUnit{(int32_t)qi, -1, k0, std::min(k0 + slots, q.range), k0 == 0, 0});
            b->lds2 = std::max(b->lds2, (size_t)std::min(slots, q.range) * (S * 8 + 4));
        } else {
            q.mode = BQ_PART;
            int shift = 0;
            while ((int64_t(2) << shift) <= slots) ++shift;
            q.shift = shift;
            q.P = (int)((q.range + (int64_t(1) << shift) - 1) >> shift);
            if ((size_t)S * 8 + (size_t)q.P * 4 > lim) {  // the pass-1 cursors do not fit
                Q.fallback = 1;
                continue;
            }
            const int64_t cw_rows = std::min<int64_t>(per_unit, n_rows);
            q.Cw = (int)std::max<int64_t>(Q.Cw_need, (int64_t)std::ceil((double)cw_rows / q.P * 1.25) + 64);
            b->lds1 = std::max(b->lds1, (size_t)S * 8 + (size_t)q.P * 4);
            for (int p = 0; p < q.P; ++p) {
                const int64_t k0 = (int64_t)p << shift;
                units2.push_back(MergeUnit{(int32_t)qi, p, k0, std::min(k0 + (int64_t(1) << shift), q.range), p == 0, 0});
            }
            b->lds2 = std::max(b->lds2, ((size_t)1 << shift) * (S * 8 + 4) + (size_t)(q.n_units + 1) * 4);
            batch_fast_tile(q);
        }
        q.unit0 = (int64_t)units.size();
        for (int64_t u = 0; u < q.n_units; ++u)
            units.push_back(BatchUnit{(int32_t)qi, (int32_t)u, std::min(u * per_unit, n_rows),
                                      std::min((u + 1) * per_unit, n_rows)});
        q.out_cap = std::min(q.range, n_rows) + 1;
        QPieces &pc = pieces[qi];
        if (q.mode == BQ_PART) {
            pc.cells = take((size_t)q.n_units * q.P * q.Cw * q.RW * 8);
            pc.cnt = take((size_t)q.n_units * q.P * 4);
        } else {
            pc.blocks = take((size_t)q.n_units * q.range * S * 8);
        }
        pc.nulls = take((size_t)q.n_units * S * 8);
        pc.key = take((size_t)q.out_cap * 8);
        pc.kvalid = take((size_t)q.out_cap);
        for (int k = 0; k < Q.L.outs.n; ++k) {
            pc.val[k] = take((size_t)q.out_cap * 8);
            pc.valid[k] = take((size_t)q.out_cap);
        }
        b->npc = std::max(b->npc, Q.g.pred.n_pcols);
        b->nin = std::max(b->nin, Q.g.n_in);
    }
    // larger units first: the persistent workgroups finish together
    std::stable_sort(units.begin(), units.end(), [](const BatchUnit &x, const BatchUnit &y) {
        return x.row1 - x.row0 > y.row1 - y.row0;
    });
    std::stable_sort(units2.begin(), units2.end(), [](const MergeUnit &x, const MergeUnit &y) {
        return x.k1 - x.k0 > y.k1 - y.k0;
    });
    b->n_units = (int64_t)units.size();
    b->n_units2 = (int64_t)units2.size();
    const Piece p_qs = take(b->qs.size() * sizeof(BatchQuery));
    const Piece p_u = take(std::max<size_t>(1, units.size()) * sizeof(BatchUnit));
    const Piece p_u2 = take(std::max<size_t>(1, units2.size()) * sizeof(MergeUnit));
    const Piece p_ctl = take((4 + b->qs.size()) * 8);
    if (b->lds1 > lim || b->lds2 > lim) return set_err(FGPU_E_UNSUPPORTED, "batch tables exceed the LDS");
    if (hipMalloc(&b->mem, std::max<size_t>(total, 256)) != hipSuccess)
        return set_err(FGPU_E_OUT_OF_MEMORY, "batch plan of %zu bytes", total);
    char *m = (char *)b->mem;
    b->d_qs = (BatchQuery *)(m + p_qs.off);
    b->d_units = (BatchUnit *)(m + p_u.off);
    b->d_units2 = (MergeUnit *)(m + p_u2.off);
    b->d_work = (unsigned long long *)(m + p_ctl.off);
    b->d_flags = b->d_work + 2;
    b->d_counts = b->d_work + 4;
    std::vector<BatchQuery> hq(b->qs.size());
    for (size_t qi = 0; qi < b->qs.size(); ++qi) {
        BatchQ &Q = b->qs[qi];
        if (Q.fallback) continue;
        BatchQuery &q = Q.q;
        const QPieces &pc = pieces[qi];
        q.cells = q.mode == BQ_PART ? (int64_t *)(m + pc.cells.off) : nullptr;
        q.cell_cnt = q.mode == BQ_PART ? (unsigned int *)(m + pc.cnt.off) : nullptr;
        q.blocks = q.mode == BQ_DENSE ? (uint64_t *)(m + pc.blocks.off) : nullptr;
        q.null_blocks = (uint64_t *)(m + pc.nulls.off);
        q.out_key = (int64_t *)(m + pc.key.off);
        q.out_key_valid = (uint8_t *)(m + pc.kvalid.off);
        for (int k = 0; k < Q.L.outs.n; ++k) {
            q.out_val[k] = (int64_t *)(m + pc.val[k].off);
            q.out_valid[k] = (uint8_t *)(m + pc.valid[k].off);
        }
        q.out_count = b->d_counts + qi;
        hq[qi] = q;
    }
    HIPCHK(hipMemcpyAsync(b->d_qs, hq.data(), hq.size() * sizeof(BatchQuery), hipMemcpyHostToDevice, c->stream));
    if (!units.empty())
        HIPCHK(hipMemcpyAsync(b->d_units, units.data(), units.size() * sizeof(BatchUnit), hipMemcpyHostToDevice,
                              c->stream));
    if (!units2.empty())
        HIPCHK(hipMemcpyAsync(b->d_units2, units2.data(), units2.size() * sizeof(MergeUnit),
                              hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(batch_set_lds(b->lds1, b->lds2));
    // the run: resets by fill kernels (no memset nodes), then the two passes
    BatchArgs A{};
    A.units = b->d_units;
    A.n_units = b->n_units;
    A.units2 = b->d_units2;
    A.n_units2 = b->n_units2;
    A.work = b->d_work;
    A.flags = b->d_flags;
    A.lds_slots = cfg.slots;
    // persistent grids: as many workgroups per CU as stay resident together
    int per_cu1 = 1, per_cu2 = 1;
    // pass 1 may run at its own block size (batch_scan_block), pass 2 at cfg.block
    const int block1 = c->tune.batch_scan_block ? c->tune.batch_scan_block : cfg.block;
    HIPCHK(batch_occupancy(b->npc, b->nin, block1, cfg.block, b->lds1, b->lds2, per_cu1, per_cu2));
    b->per_cu1 = std::min(per_cu1, cfg.wgs_per_cu);
    b->per_cu2 = std::min(per_cu2, cfg.wgs_per_cu);
    const int grid1 = c->num_cus * b->per_cu1, grid2 = c->num_cus * b->per_cu2;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = launch_fill_u64(c->stream, (uint64_t *)b->d_work, (int64_t)(4 + b->qs.size()), 0);
    if (e == hipSuccess && b->n_units)
        e = launch_batch_scan(c->stream, b->d_qs, A, b->npc, b->nin, grid1, block1, b->lds1);
    if (e == hipSuccess && b->n_units2) e = launch_batch_merge(c->stream, b->d_qs, A, grid2, cfg.block, b->lds2);
    const hipError_t e2 = hipStreamEndCapture(c->stream, &b->graph);
    if (e != hipSuccess || e2 != hipSuccess)
        return set_err(FGPU_E_DEVICE, "batch graph capture: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    int n_packed = 0, n_fast = 0;
    for (const BatchQ &Q : b->qs) {
        n_packed += !Q.fallback && Q.q.mode == BQ_PART && Q.q.pack;
        n_fast += !Q.fallback && Q.q.fast;
    }
    char msg[200];
    snprintf(msg, sizeof msg,
             "block=%d unit_rows=%lld slots=%d units=%lld/%lld wgs_per_cu=%d/%d plan_bytes=%zu packed=%d fast=%d",
             cfg.block, (long long)cfg.unit_rows, cfg.slots, (long long)b->n_units, (long long)b->n_units2, b->per_cu1,
             b->per_cu2, total, n_packed, n_fast);
    b->info = msg;
    return FGPU_OK;
}

// One replay of the graph; `flags` gets the overflow flags, `counts` the rows
// each query emitted.
int batch_launch(fgpu_batch *b, unsigned long long flags[2], std::vector<unsigned long long> &counts) {
    fgpu_ctx *c = b->c;
    counts.assign(b->qs.size(), 0);
    HIPCHK(hipEventRecord(b->ev0, c->stream));
    HIPCHK(hipGraphLaunch(b->exec, c->stream));
    HIPCHK(hipEventRecord(b->ev1, c->stream));
    HIPCHK(hipMemcpyAsync(flags, b->d_flags, 16, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(counts.data(), b->d_counts, counts.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    float ms = 0;
    if (hipEventElapsedTime(&ms, b->ev0, b->ev1) == hipSuccess) b->last_ms = ms;
    return FGPU_OK;
}

// Runs the graph; a cell overflow re-plans once with cells sized from the
// counts the run kept.
int batch_execute(fgpu_batch *b, std::vector<unsigned long long> &counts) {
    unsigned long long flags[2] = {0, 0};
    for (int attempt = 0; attempt < 2; ++attempt) {
        if (int rc = batch_launch(b, flags, counts)) return rc;
        if (!flags[0]) break;
        if (attempt == 1) return set_err(FGPU_E_UNSUPPORTED, "batch cells overflowed after resizing");
        for (BatchQ &Q : b->qs) {
            if (Q.fallback || Q.q.mode != BQ_PART) continue;
            std::vector<unsigned int> cc((size_t)Q.q.n_units * Q.q.P);
            HIPCHK(hipMemcpy(cc.data(), Q.q.cell_cnt, cc.size() * 4, hipMemcpyDeviceToHost));
            const unsigned mx = cc.empty() ? 0 : *std::max_element(cc.begin(), cc.end());
            if ((int64_t)mx > Q.q.Cw) Q.Cw_need = (int64_t)mx * 5 / 4 + 64;
        }
        if (int rc = batch_plan(b)) return rc;
    }
    if (flags[1]) return set_err(FGPU_E_DEVICE, "batch output overflow");
    return FGPU_OK;
}

}  // namespace

extern "C" {

int fgpu_topk(fgpu_ctx *c, const fgpu_table *t, const fgpu_pred *p, int32_t col, int desc, int32_t k,
              int64_t *rows_out, int64_t *vals_out, int64_t *n_out) {
    if (!c || !t || !n_out || (k > 0 && !rows_out)) return set_err(FGPU_E_INVALID_QUERY, "NULL argument");
    if (k < 0) return set_err(FGPU_E_INVALID_QUERY, "negative LIMIT");
    if (k > 32) return set_err(FGPU_E_UNSUPPORTED, "LIMIT %d: the device top-k takes at most 32 rows", k);
    if (col < 0 || col >= (int)t->cols.size()) return set_err(FGPU_E_COLUMN_NOT_FOUND, "sort column %d not in table", col);
    const DevCol &dc = t->cols[col];
    if (dc.dtype == FGPU_UTF8) return set_err(FGPU_E_UNSUPPORTED, "ORDER BY a Utf8 column");
    std::lock_guard<std::mutex> lk(c->mu);
    if (int rc = set_device(c->device)) return rc;
    *n_out = 0;
    if (k == 0 || t->n_rows == 0) return FGPU_OK;
    Binder b(t);
    DPred dp;
    if (int rc = compile_pred(c, t, p, b, dp)) return rc;
    std::string zone_note;  // the scan skips the batches that cannot pass
    if (int rc = prune_batches(c, t, p, dp, &zone_note)) return rc;
    const int slot = b.bind(col);
    if (slot < 0) return set_err(FGPU_E_UNSUPPORTED, "query reads too many columns");
    const int K = k <= 8 ? 8 : k <= 16 ? 16 : 32;
    const int grid = c->num_cus * 4;
    void *mem;
    const size_t cand_bytes = (size_t)grid * K * 16;
    if (int rc = buf(c, B_TOPK, 2 * (cand_bytes + (size_t)K * 16), &mem)) return rc;
    char *m = (char *)mem;
    void *cands = m, *out_keys = m + cand_bytes, *cands_n = m + cand_bytes + K * 16,
         *out_nulls = m + 2 * cand_bytes + K * 16;
    const bool nulls = dc.null_count > 0;
    mark(c, 0);
    mark(c, 1);
    HIPCHK(launch_topk(c->stream, b.dc, dp, slot, desc, 0, t->n_rows, K, grid, cands, out_keys));
    if (nulls) HIPCHK(launch_topk(c->stream, b.dc, dp, slot, desc, 1, t->n_rows, K, grid, cands_n, out_nulls));
    mark(c, 2);
    std::vector<int64_t> h(4 * (size_t)K);  // (key, row) pairs
    HIPCHK(hipMemcpyAsync(h.data(), out_keys, (size_t)K * 16, hipMemcpyDeviceToHost, c->stream));
    if (nulls) HIPCHK(hipMemcpyAsync(h.data() + 2 * K, out_nulls, (size_t)K * 16, hipMemcpyDeviceToHost, c->stream));
    mark(c, 3);
    HIPCHK(hipStreamSynchronize(c->stream));
    // NULL rows last for DESC, first for ASC (SortExpr::desc / asc, src/query/sort.rs:28-34)
    // a key is a bijection of the value: undo the order transform
    auto value_of = [&](uint64_t key) -> int64_t {
        const uint64_t o = desc ? ~key : key;
        if (dc.dtype == FGPU_FLOAT64) return (int64_t)((o >> 63) ? o ^ (1ull << 63) : ~o);
        return (int64_t)(o ^ (1ull << 63));
    };
    std::vector<std::pair<int64_t, int64_t>> keyed, nullr;  // (row, value)
    for (int j = 0; j < K; ++j) {
        if (h[2 * j + 1] != INT64_MAX) keyed.emplace_back(h[2 * j + 1], value_of((uint64_t)h[2 * j]));
        if (nulls && h[2 * K + 2 * j + 1] != INT64_MAX) nullr.emplace_back(h[2 * K + 2 * j + 1], 0);
    }
    std::vector<std::pair<int64_t, int64_t>> rows;
    if (!desc) rows.insert(rows.end(), nullr.begin(), nullr.end());
    rows.insert(rows.end(), keyed.begin(), keyed.end());
    if (desc) rows.insert(rows.end(), nullr.begin(), nullr.end());
    if ((int)rows.size() > k) rows.resize((size_t)k);
    for (size_t j = 0; j < rows.size(); ++j) {
        rows_out[j] = rows[j].first;
        if (vals_out) vals_out[j] = rows[j].second;
    }
    *n_out = (int64_t)rows.size();
    finish_timing(c, t->n_rows * (int64_t)dc.elem_bytes(), false);
    c->plan = "topk" + zone_note;
    return FGPU_OK;
}

int fgpu_batch_prepare(fgpu_ctx *c, const fgpu_batch_query *queries, int n, fgpu_batch **out) {
    if (!c || !out || n < 0 || (n > 0 && !queries)) return set_err(FGPU_E_INVALID_QUERY, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    if (int rc = set_device(c->device)) return rc;
    auto *b = new fgpu_batch();
    b->c = c;
    if (c->tune.batch_unit_rows >= 64) b->cfg.unit_rows = c->tune.batch_unit_rows;
    b->qs.resize((size_t)n);
    for (int i = 0; i < n; ++i) {
        const fgpu_batch_query &bq = queries[i];
        BatchQ &Q = b->qs[(size_t)i];
        if (!bq.table || !bq.spec || (bq.spec->n_aggs > 0 && !bq.spec->aggs)) {
            delete b;
            return set_err(FGPU_E_INVALID_QUERY, "batch query %d: NULL table or spec", i);
        }
        Q.t = bq.table;
        // copies of the query (the caller's arrays need not outlive this call)
        Q.strs.reserve((size_t)(bq.pred ? bq.pred->n_nodes : 0) + (size_t)bq.spec->n_aggs);
        if (bq.pred) {
            Q.nodes.assign(bq.pred->nodes, bq.pred->nodes + bq.pred->n_nodes);
            for (auto &nd : Q.nodes) Q.pred_copies.keep(nd);
            Q.pred.n_nodes = (int32_t)Q.nodes.size();
            Q.pred.nodes = Q.nodes.data();
            Q.has_pred = true;
        }
        Q.aggs.assign(bq.spec->aggs, bq.spec->aggs + bq.spec->n_aggs);
        for (auto &ag : Q.aggs)
            if (ag.alias) {
                Q.strs.emplace_back(ag.alias);
                ag.alias = Q.strs.back().c_str();
            }
        Q.spec = *bq.spec;
        Q.spec.aggs = Q.aggs.data();
        // the batch kernels take integer keys with a dense range and no
        // dictionary predicates; anything else runs through fgpu_groupby
        const fgpu_table *t = Q.t;
        const int kc = Q.spec.key_col;
        if (kc < 0 || kc >= (int)t->cols.size()) {
            Q.fallback = 1;
            continue;
        }
        const DevCol &key = t->cols[kc];
        bool dict_pred = false;
        for (auto &nd : Q.nodes)
            if ((nd.kind == FGPU_P_CMP && nd.col >= 0 && nd.col < (int)t->cols.size() &&
                 t->cols[nd.col].dtype == FGPU_UTF8) ||
                nd.kind == FGPU_P_IN || nd.kind == FGPU_P_LIKE)
                dict_pred = true;
        const bool int_key = key.dtype == FGPU_INT32 || key.dtype == FGPU_INT64;
        const int64_t range = key.imin <= key.imax ? key.imax - key.imin + 1 : 1;
        if (!int_key || dict_pred || key.imin == INT64_MIN || range <= 0 || range > (int64_t(1) << 26)) {
            Q.fallback = 1;
            continue;
        }
        Q.key = &key;
        Binder bd(t);
        DPred dp;
        if (int rc = compile_pred(c, t, Q.has_pred ? &Q.pred : nullptr, bd, dp)) {
            delete b;
            return rc;
        }
        GbArgs &a = Q.g;
        a = GbArgs{};
        a.pred = dp;
        a.key_col = bd.bind(kc);
        int in_of_col[256];
        for (int &x : in_of_col) x = -1;
        a.n_in = 0;
        bool too_many = false;
        auto slot_of = [&](int /*side*/, int col) -> int {
            if (col < 256 && in_of_col[col] >= 0) return in_of_col[col];
            if (a.n_in >= MAX_IN) {
                too_many = true;
                return -1;
            }
            const int sl = bd.bind(col);
            if (sl < 0) return -1;
            a.in_col[a.n_in] = sl;
            if (col < 256) in_of_col[col] = a.n_in;
            return a.n_in++;
        };
        if (a.key_col < 0 || build_layout(Q.spec.aggs, Q.spec.n_aggs, t, nullptr, slot_of, Q.L) || too_many) {
            g_err.clear();
            Q.fallback = 1;  // fgpu_groupby reports the error at run time
            continue;
        }
        a.cols = bd.dc;
        a.ops = Q.L.ops;
        a.n_rows = t->n_rows;
        for (int j = 0; j < a.n_in; ++j)
            for (const DevCol &dc : t->cols)
                if (dc.d_values == a.cols.v[a.in_col[j]] && dc.null_count) Q.has_valid_word = 1;
        // Packed records: the key slot and an integer input 0 in one word when
        // their ranges (from the upload statistics) fit 64 bits together.
        if (c->tune.gbp_pack && a.n_in > 0 && !Q.has_valid_word) {
            const DevCol *in0 = nullptr;
            for (const DevCol &dc : t->cols)
                if (dc.d_values == a.cols.v[a.in_col[0]]) in0 = &dc;
            if (in0 && (in0->dtype == FGPU_INT64 || in0->dtype == FGPU_INT32) && in0->imin <= in0->imax) {
                auto bits = [](uint64_t span) { return span ? 64 - __builtin_clzll(span) : 1; };
                const int kb = bits((uint64_t)range - 1);
                const int vb = bits((uint64_t)in0->imax - (uint64_t)in0->imin);
                if (kb + vb <= 64) {
                    Q.pack = 1;
                    Q.pk_kbits = kb;
                    Q.pk_vmin = in0->imin;
                }
            }
        }
    }
    if (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
        if (b->ev0) (void)hipEventDestroy(b->ev0);
        delete b;
        return set_err(FGPU_E_DEVICE, "batch events");
    }
    if (int rc = batch_plan(b)) {
        (void)hipEventDestroy(b->ev0);
        (void)hipEventDestroy(b->ev1);
        batch_release(b);
        delete b;
        return rc;
    }
    *out = b;
    return FGPU_OK;
}

int fgpu_batch_tune(fgpu_batch *b, int max_trials) {
    if (!b) return set_err(FGPU_E_INVALID_QUERY, "NULL batch");
    fgpu_ctx *c = b->c;
    std::lock_guard<std::mutex> lk(c->mu);
    if (int rc = set_device(c->device)) return rc;
    std::vector<BatchCfg> cands;
    // slots: 4096 is capped at the largest table the state width lets fit
    // (one pass-2 workgroup per CU); smaller tables let several share a CU
    for (int block : {1024, 512, 256})
        for (int64_t ur : {int64_t(1) << 18, int64_t(1) << 16, int64_t(1) << 20})
            for (int slots : {4096, 1024, 512, 256}) {
                BatchCfg k;
                k.block = block;
                k.unit_rows = ur;
                k.slots = slots;
                cands.push_back(k);
            }
    if (max_trials > 0 && (int)cands.size() > max_trials) cands.resize((size_t)max_trials);
    BatchCfg best = b->cfg;
    float best_ms = 1e30f;
    std::vector<unsigned long long> counts;
    int tried = 0;
    std::string trials;
    for (const BatchCfg &k : cands) {
        b->cfg = k;
        if (batch_plan(b) != FGPU_OK) {
            g_err.clear();
            continue;
        }
        if (batch_execute(b, counts) != FGPU_OK) {
            g_err.clear();
            continue;
        }
        float ms = 0;
        for (int r = 0; r < 2; ++r) {  // after the warm-up above
            if (int rc = batch_execute(b, counts)) return rc;
            ms += b->last_ms;
        }
        ++tried;
        char t[64];
        snprintf(t, sizeof t, " %d/%lld/%d:%.3f", k.block, (long long)k.unit_rows, k.slots, ms / 2);
        trials += t;
        if (ms < best_ms) {
            best_ms = ms;
            best = b->cfg;
        }
    }
    b->cfg = best;
    if (int rc = batch_plan(b)) return rc;
    char msg[96];
    snprintf(msg, sizeof msg, " tuned=%d trials best=%.3f ms;", tried, best_ms / 2);
    b->info += msg;
    b->info += " block/unit_rows/slots:ms" + trials;
    return FGPU_OK;
}

int fgpu_batch_run(fgpu_batch *b, fgpu_result **outs) {
    if (!b || (!outs && !b->qs.empty())) return set_err(FGPU_E_INVALID_QUERY, "NULL argument");
    fgpu_ctx *c = b->c;
    std::vector<unsigned long long> counts;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        if (int rc = set_device(c->device)) return rc;
        if (int rc = batch_execute(b, counts)) return rc;
        c->last = fgpu_timing{};
        c->last.total_ms = c->last.main_ms = b->last_ms;
        for (size_t qi = 0; qi < b->qs.size(); ++qi) {
            BatchQ &Q = b->qs[qi];
            outs[qi] = nullptr;
            if (Q.fallback) continue;
            const BatchQuery &q = Q.q;
            int64_t *vals[FGPU_MAX_AGGS];
            uint8_t *valid[FGPU_MAX_AGGS];
            for (int k = 0; k < FGPU_MAX_AGGS; ++k) {
                vals[k] = q.out_val[k];
                valid[k] = q.out_valid[k];
            }
            if (int rc = collect(c, Q.L, q.out_cap, Q.key, Q.key->dtype, Q.key->name, q.out_key, q.out_key_valid,
                                 vals, valid, q.out_count, &outs[qi], (int64_t)counts[qi]))
                return rc;
        }
    }
    for (size_t qi = 0; qi < b->qs.size(); ++qi) {  // the rest, one plain call each
        BatchQ &Q = b->qs[qi];
        if (!Q.fallback) continue;
        if (int rc = fgpu_groupby(c, Q.t, Q.has_pred ? &Q.pred : nullptr, &Q.spec, &outs[qi])) return rc;
    }
    return FGPU_OK;
}

const char *fgpu_batch_info(const fgpu_batch *b) { return b ? b->info.c_str() : ""; }

int fgpu_batch_fallbacks(const fgpu_batch *b) {
    int n = 0;
    if (b)
        for (const BatchQ &Q : b->qs) n += Q.fallback;
    return n;
}

void fgpu_batch_free(fgpu_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->c->device);
    (void)hipStreamSynchronize(b->c->stream);
    batch_release(b);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    delete b;
}

}  // extern "C"


// ─── internals shared with the request server (fgpu_host.h) ─────────────────
namespace fgpu_host {

int set_error(int code, const char *msg) { return set_err(code, "%s", msg); }

void ctx_budget_begin(fgpu_ctx *c, int64_t limit) {
    std::lock_guard<std::mutex> lk(c->mu);
    c->arena_limit = limit;
    for (auto &r : c->call_req) r = 0;
    c->limit_hit = false;
}

int64_t ctx_call_need(fgpu_ctx *c) {
    std::lock_guard<std::mutex> lk(c->mu);
    int64_t s = 0;
    for (size_t r : c->call_req)
        if (r) s += (int64_t)arena_want(r);
    return s;
}

bool ctx_limit_hit(fgpu_ctx *c) {
    std::lock_guard<std::mutex> lk(c->mu);
    return c->limit_hit;
}

int ctx_trim(fgpu_ctx *c) {
    std::lock_guard<std::mutex> lk(c->mu);
    if (int rc = set_device(c->device)) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    for (auto &b : c->bufs) {
        if (b.p) HIPCHK(hipFree(b.p));
        b.p = nullptr;
        b.n = 0;
    }
    return FGPU_OK;
}

fgpu_result *result_int64(const char *name, int64_t v) {
    auto *r = new fgpu_result();
    r->rows = 1;
    r->cols.emplace_back();
    r->cols[0].name = name ? name : "col0";
    r->cols[0].dtype = FGPU_INT64;
    r->cols[0].i64 = {v};
    return r;
}

fgpu_result *result_topk(const int64_t *rows, const int64_t *vals, int64_t n, bool is_f64,
                         const uint8_t *valid) {
    auto *r = new fgpu_result();
    r->rows = n;
    r->cols.resize(2);
    r->cols[0].name = "row";
    r->cols[0].dtype = FGPU_INT64;
    r->cols[0].i64.assign(rows, rows + n);
    r->cols[1].name = "value";
    r->cols[1].dtype = is_f64 ? FGPU_FLOAT64 : FGPU_INT64;
    if (is_f64) {
        r->cols[1].f64.resize((size_t)n);
        for (int64_t i = 0; i < n; ++i) r->cols[1].f64[i] = __builtin_bit_cast(double, vals[i]);
    } else {
        r->cols[1].i64.assign(vals, vals + n);
    }
    if (valid) r->cols[1].valid.assign(valid, valid + n);
    return r;
}

bool column_is_f64(const fgpu_table *t, int col) {
    return t && col >= 0 && col < (int)t->cols.size() && t->cols[col].dtype == FGPU_FLOAT64;
}
bool column_has_nulls(const fgpu_table *t, int col) {
    return t && col >= 0 && col < (int)t->cols.size() && t->cols[col].null_count > 0;
}

int shared_count_column(const fgpu_table *t, const fgpu_pred *p) {
    if (!t || !p || p->n_nodes != 1 || !p->nodes) return -1;
    const fgpu_pred_node &nd = p->nodes[0];
    if (nd.col < 0 || nd.col >= (int)t->cols.size()) return -1;
    const int dt = t->cols[nd.col].dtype;
    if (dt != FGPU_INT32 && dt != FGPU_INT64 && dt != FGPU_FLOAT64) return -1;
    if (nd.kind == FGPU_P_IS_NULL || nd.kind == FGPU_P_IS_NOT_NULL) return nd.col;
    if (nd.kind != FGPU_P_CMP || nd.lit_type == FGPU_UTF8) return -1;
    return nd.col;
}

int filter_count_shared(fgpu_ctx *c, const fgpu_table *t, int col, const fgpu_pred *const *preds, int q,
                        int64_t *counts) {
    if (!c || !t || !preds || !counts || q < 1 || q > SHARED_MAX_Q)
        return set_err(FGPU_E_INVALID_QUERY, "shared COUNT: bad arguments");
    std::lock_guard<std::mutex> lk(c->mu);
    if (int rc = set_device(c->device)) return rc;
    SharedCount a{};
    a.q = q;
    for (int i = 0; i < q; ++i) {
        if (shared_count_column(t, preds[i]) != col)
            return set_err(FGPU_E_INVALID_QUERY, "shared COUNT: predicate %d does not read column %d alone", i, col);
        Binder b(t);
        DPred dp;
        if (int rc = compile_pred(c, t, preds[i], b, dp)) return rc;
        a.leaf[i] = dp.nodes[0];
        a.leaf[i].col = 0;
        a.leaf[i].pc = 0;
    }
    const DevCol &dc = t->cols[col];
    a.dtype = dc.dtype;
    a.values = dc.d_values;
    a.valid = (const uint32_t *)dc.d_valid;
    a.n = t->n_rows;
    unsigned long long *d_cnt;
    if (int rc = buf(c, B_COUNTERS, 64 * 8, (void **)&d_cnt)) return rc;
    a.out = d_cnt;
    mark(c, 0);
    HIPCHK(launch_fill_u64(c->stream, (uint64_t *)d_cnt, q, 0));
    mark(c, 1);
    const int grid = grid_for(c, (t->n_rows + 63) / 64 * 64, BLOCK, c->tune.filter_waves_per_cu);
    if (t->n_rows) HIPCHK(launch_filter_count_shared(c->stream, a, grid));
    mark(c, 2);
    unsigned long long h[SHARED_MAX_Q];
    HIPCHK(mail_copy(c, d_cnt, (size_t)q * 8));
    mark(c, 3);
    HIPCHK(mail_take(c, h, (size_t)q * 8));
    finish_timing(c, t->n_rows * (int64_t)dc.elem_bytes(), false);
    for (int i = 0; i < q; ++i) counts[i] = (int64_t)h[i];
    c->plan = "count:shared q=" + std::to_string(q);
    return FGPU_OK;
}

}  // namespace fgpu_host
