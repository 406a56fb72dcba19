"""ORDER BY col [DESC] LIMIT k on the device (fgpu_topk) against the oracle."""
import numpy as np
import pytest

import falcondb_amd as fa
import oracle as ora
from falcondb_amd.operators import FalconError, NullableColumn, col, execute_sort_limit

from test_gpu_parity import _table

pytestmark = pytest.mark.gpu


def _cols(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=n)
    x[rng.integers(0, n, 20)] = np.nan
    x[rng.integers(0, n, 5)] = np.inf
    x[::1000] = x[17]  # ties
    xv = rng.random(n) > 0.01
    return {"x": NullableColumn(x, xv), "i": rng.integers(-10**12, 10**12, n),
            "j": rng.integers(-1000, 1000, n).astype(np.int32), "f": rng.integers(0, 100, n).astype(np.int32)}


@pytest.mark.parametrize("k", [1, 10, 32])
@pytest.mark.parametrize("desc", [False, True])
def test_topk_matches_oracle(ctx, k, desc):
    cols = _cols(1_000_000, 5)
    t = _table(ctx, cols)
    for name in ("x", "i", "j"):
        for filt in (None, col("f") < 30):
            rows, vals = execute_sort_limit(ctx, t, name, k, desc, filt)
            exp = ora.topk(cols, filt, name, k, desc)
            assert rows.tolist() == exp.tolist(), (name, filt)
            v = cols[name].values if isinstance(cols[name], NullableColumn) else cols[name]
            ok = cols[name].valid[rows] if isinstance(cols[name], NullableColumn) else np.ones(len(rows), bool)
            if v.dtype == np.float64:
                assert np.array_equal(vals[ok].view(np.int64), v[rows][ok].view(np.int64))
            else:
                assert np.array_equal(vals[ok], v[rows][ok])


def test_topk_nulls_and_small_inputs(ctx):
    n = 1000
    s = np.arange(n, dtype=np.float64)
    cols = {"s": NullableColumn(s, np.arange(n) >= n - 3), "f": np.arange(n, dtype=np.int64)}
    t = _table(ctx, cols)
    for desc in (False, True):
        for k in (2, 5, 8):
            rows, _ = execute_sort_limit(ctx, t, "s", k, desc)
            assert rows.tolist() == ora.topk(cols, None, "s", k, desc).tolist()
    rows, _ = execute_sort_limit(ctx, t, "s", 5, True, col("f") > 10**6)  # nothing passes
    assert rows.size == 0
    with pytest.raises(FalconError):
        execute_sort_limit(ctx, t, "s", 33, True)
