// fgpu_host.h — host-side internals the runtime shares with the request
// server (fgpu_server.hip).  Nothing here crosses the C ABI.
#ifndef FGPU_HOST_H
#define FGPU_HOST_H

#include <cstdint>
#include <deque>
#include <string>
#include <vector>

#include "falcon_gpu.h"

namespace fgpu_host {

// Deep copies of what predicate nodes point at (literal strings, IN lists),
// for a query that outlives its caller's arrays.  Deques keep element
// addresses stable as copies are added.
struct PredCopies {
    std::deque<std::string> strs;
    std::deque<std::vector<int64_t>> nums;
    std::deque<std::vector<const char *>> ptrs;
    const char *str(const char *p) {
        if (!p) return nullptr;
        strs.emplace_back(p);
        return strs.back().c_str();
    }
    void keep(fgpu_pred_node &nd) {
        nd.lit_str = str(nd.lit_str);
        if (nd.kind != FGPU_P_IN || nd.n_list <= 0 || !nd.list) return;
        if (nd.lit_type == FGPU_UTF8) {
            const char *const *src = (const char *const *)nd.list;
            ptrs.emplace_back();
            for (int i = 0; i < nd.n_list; ++i) ptrs.back().push_back(str(src[i]));
            nd.list = ptrs.back().data();
        } else {  // int64 or double literals, copied as 8-byte patterns
            const int64_t *src = (const int64_t *)nd.list;
            nums.emplace_back(src, src + nd.n_list);
            nd.list = nums.back().data();
        }
    }
};

// Sets the calling thread's fgpu_last_error() message; returns `code`.
int set_error(int code, const char *msg);

// Serving-slot budget of a context: from now on its scratch arena may not grow
// past `limit` bytes (0 = unbounded; buffers the current request has not
// asked for are released first), and the per-request tally restarts.
void ctx_budget_begin(fgpu_ctx *c, int64_t limit);
// Scratch bytes the requests since ctx_budget_begin asked for (with the
// arena's growth headroom): what a request of that shape needs.
int64_t ctx_call_need(fgpu_ctx *c);
// Whether a call since ctx_budget_begin failed because of the arena limit.
bool ctx_limit_hit(fgpu_ctx *c);
// Releases every arena buffer of an idle context.
int ctx_trim(fgpu_ctx *c);

// A one-row result of a single Int64 column.
fgpu_result *result_int64(const char *name, int64_t v);
// ORDER BY ... LIMIT k result: columns "row" (Int64) and "value" (Int64, or
// Float64 when is_f64; NULL rows carry valid = 0).
fgpu_result *result_topk(const int64_t *rows, const int64_t *vals, int64_t n, bool is_f64,
                         const uint8_t *valid);

// Column that a COUNT(*) predicate reads if it can ride a shared scan: one
// comparison (or IS [NOT] NULL) on an Int32 / Int64 / Float64 column with a
// numeric literal.  -1 otherwise.
int shared_count_column(const fgpu_table *t, const fgpu_pred *p);
// counts[i] = COUNT(*) WHERE preds[i], all on column `col` of t, in one scan.
int filter_count_shared(fgpu_ctx *c, const fgpu_table *t, int col, const fgpu_pred *const *preds,
                        int q, int64_t *counts);
// Whether column `col` of t is a Float64 column; whether it holds NULLs.
bool column_is_f64(const fgpu_table *t, int col);
bool column_has_nulls(const fgpu_table *t, int col);

}  // namespace fgpu_host

#endif
