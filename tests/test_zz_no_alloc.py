"""fgpu_groupby_static, the GROUP BY that allocates nothing, called through the
C ABI with caller-owned device outputs (torch tensors) and checked against
fgpu_groupby on the same inputs: two calls in a row, the second on a smaller
table, for a Utf8 key with a dictionary predicate (pass tables in the static
scratch), a dense int32 key whose state lives in HBM, and no GROUP BY."""
import ctypes as C

import numpy as np
import pytest
import torch

import falcondb_amd as fa
from falcondb_amd import _abi
from falcondb_amd import datagen as dg
from falcondb_amd._abi import AVG, COUNT, MAX, MIN, SUM
from falcondb_amd.operators import AggExpr, DictColumn, _aggs_struct, _pred, col

pytestmark = pytest.mark.gpu


def run_static(ctx, table, key, aggs, filt, capacity):
    """Returns (rc, n_out, keys, key_valid, [values], [valid]) as numpy arrays."""
    lib = _abi.load()
    p, keep = _pred(filt, table)
    arr, keep2 = _aggs_struct(aggs, lambda side, name: table.column_index(name))
    spec = _abi.GroupBySpec(table.column_index(key) if key else -1, len(aggs),
                            C.cast(arr, C.POINTER(_abi.Agg)))
    dev = torch.device("cuda", 0)
    cap = max(capacity, 1)
    kbuf = torch.full((cap,), -7, dtype=torch.int64, device=dev)
    kval = torch.full((cap,), 9, dtype=torch.uint8, device=dev)
    vals = [torch.full((cap,), -7, dtype=torch.int64, device=dev) for _ in aggs]
    valid = [torch.full((cap,), 9, dtype=torch.uint8, device=dev) for _ in aggs]
    vp = (C.c_void_p * len(aggs))(*[v.data_ptr() for v in vals])
    vv = (C.c_void_p * len(aggs))(*[v.data_ptr() for v in valid])
    n = C.c_int64(-1)
    torch.cuda.synchronize()  # the fills above run on torch's stream
    rc = lib.fgpu_groupby_static(ctx.handle, table.handle, p, C.byref(spec), kbuf.data_ptr(),
                                 kval.data_ptr(), vp, vv, capacity, C.byref(n))
    k = n.value if rc == 0 else 0
    return (rc, n.value, kbuf[:k].cpu().numpy(), kval[:k].cpu().numpy(),
            [v[:k].cpu().numpy() for v in vals], [v[:k].cpu().numpy() for v in valid])


def check_against_plain(ctx, table, key, aggs, filt, dictionary=None):
    exp = fa.GroupByExec([key], aggs).execute(ctx, table, filt) if key else \
        fa.execute_agg_only(ctx, table, aggs, filt)
    rc, n, keys, kvalid, vals, valid = run_static(ctx, table, key, aggs, filt, 1 << 20)
    assert rc == 0, fa.operators._lib().fgpu_last_error()
    assert n == exp.num_rows
    order = np.arange(n)
    first = 1 if key else 0
    if key:
        if dictionary is not None:  # Utf8 keys come back as codes
            got_keys = np.array([dictionary[c] for c in keys], dtype=object)
        else:
            got_keys = keys
        order = np.argsort(got_keys.astype(str) if dictionary is not None else got_keys, kind="stable")
        exp = exp.sort_by(key)
        assert list(got_keys[order]) == list(exp.columns[0])
        assert np.all(kvalid == 1)
    for i in range(len(aggs)):
        e = exp.columns[first + i]
        ev = exp.validity[first + i]
        ev = np.ones(n, bool) if ev is None else ev
        g = vals[i][order]
        gv = valid[i][order]
        assert np.array_equal(gv.astype(bool), ev), i
        if exp.dtypes[first + i] == "Float64":
            g = g.view(np.float64)
        assert np.array_equal(g[ev], e[ev]), (i, g[ev][:5], e[ev][:5])


def test_static_groupby_matches_plain_two_calls(ctx):
    # Utf8 key with a dictionary predicate: dense in LDS, pass tables in scratch
    aggs = [AggExpr(COUNT), AggExpr(SUM, "score"), AggExpr(AVG, "salary"), AggExpr(MIN, "salary"),
            AggExpr(MAX, "score")]
    filt = (col("score") > 50) & (col("dept") != "dept03")
    for n in (1_000_000, 100_000):
        c1 = dg.c1_table(n)
        t = fa.Table(ctx, {"id": c1["id"], "dept": DictColumn(c1["dept"], dg.DEPTS),
                           "score": c1["score"], "salary": c1["salary"]})
        check_against_plain(ctx, t, "dept", aggs, filt, dictionary=dg.DEPTS)
        t.free()
    # int32 key over a 100k range: the state table lives in the static scratch
    aggs = [AggExpr(COUNT), AggExpr(SUM, "val"), AggExpr(MIN, "val"), AggExpr(MAX, "val"),
            AggExpr(SUM, "score")]
    for n, card in ((2_000_000, 100_000), (50_000, 30_000)):
        t = fa.Table(ctx, dg.c4_batch(n, card, 11 + n))
        check_against_plain(ctx, t, "key", aggs, col("score") < 60)
        t.free()
    # no GROUP BY
    for n in (300_000, 1000):
        rng = np.random.default_rng(n)
        t = fa.Table(ctx, {"a": rng.normal(size=n), "b": rng.integers(-10**6, 10**6, n)})
        check_against_plain(ctx, t, None, [AggExpr(SUM, "a"), AggExpr(COUNT), AggExpr(MIN, "b"),
                                           AggExpr(AVG, "b")], col("b") > 0)
        t.free()


def test_static_groupby_limits(ctx):
    # an integer SUM past int64 comes back with validity 2 and the double's bits
    v = np.array([2**63 - 1] * 4 + [1, 2, 3], np.int64)
    t = fa.Table(ctx, {"v": v})
    rc, n, _, _, vals, valid = run_static(ctx, t, None, [AggExpr(SUM, "v")], None, 1)
    assert rc == 0 and n == 1 and valid[0][0] == 2
    assert vals[0].view(np.float64)[0] == float(sum(v.tolist()))
    # capacity below what the call may emit: the bound comes back
    c1 = dg.c1_table(10_000)
    t = fa.Table(ctx, {"dept": DictColumn(c1["dept"], dg.DEPTS), "score": c1["score"]})
    rc, n, *_ = run_static(ctx, t, "dept", [AggExpr(COUNT)], None, 5)
    assert rc == _abi.E_INDEX_OUT_OF_BOUNDS and n == len(dg.DEPTS) + 2
    # a key that needs a hash table allocates: refused
    k = np.array([0, 10**12, 5, 7], np.int64)
    t = fa.Table(ctx, {"k": k, "x": np.ones(4)})
    rc, *_ = run_static(ctx, t, "k", [AggExpr(SUM, "x")], None, 1 << 20)
    assert rc == _abi.E_UNSUPPORTED
