"""CPU-only checks of the native boundary: the C-ABI library loads and exports
every symbol include/falcon_gpu.h declares, the host copy of the kernels'
exact float accumulator agrees with math.fsum, the synthetic generators are
deterministic, and the host-side helpers behave.  Nothing here touches a GPU."""
import ctypes as C
import math
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from falcondb_amd import _abi, datagen as dg
from falcondb_amd.operators import host_exact_sum, pack_validity

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "falcon_gpu.h"


def header_symbols():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fgpu_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(_abi.GPU_LIB_PATH))
    syms = header_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # the Python declarations cover the header exactly
    assert sorted(_abi.DECLARED) == syms


def test_abi_version_and_error_string_without_gpu():
    lib = _abi.load()
    assert lib.fgpu_abi_version() == 1
    assert isinstance(lib.fgpu_last_error(), bytes)


@pytest.mark.parametrize("seed", range(5))
def test_host_exact_sum_matches_fsum(seed):
    rng = np.random.default_rng(seed)
    for _ in range(200):
        n = int(rng.integers(0, 80))
        scale = 10.0 ** rng.integers(-12, 12, size=n)
        x = rng.normal(size=n) * scale
        if seed % 2 and n:
            x = np.concatenate([x, -x[: n // 2]])
        assert host_exact_sum(x) == math.fsum(x)


def test_host_exact_sum_special_values():
    assert host_exact_sum(np.array([])) == 0.0
    assert host_exact_sum(np.array([0.1] * 10)) == math.fsum([0.1] * 10)
    assert math.isnan(host_exact_sum(np.array([1.0, np.nan])))
    assert host_exact_sum(np.array([np.inf, 1.0])) == np.inf
    assert math.isnan(host_exact_sum(np.array([np.inf, -np.inf])))
    assert host_exact_sum(np.array([5e-324, 5e-324, 1e-320])) == math.fsum([5e-324, 5e-324, 1e-320])
    # ties round to even
    assert host_exact_sum(np.array([1.0, 2.0 ** -53])) == 1.0
    assert host_exact_sum(np.array([1.0, 2.0 ** -53, 2.0 ** -105])) == 1.0 + 2.0 ** -52


def window_lsb(x) -> int | None:
    """The exact-sum accumulator's LSB exponent L for a column: its smallest
    mantissa-LSB exponent, or, when its exponents span more than 15 words
    hold, 364 below its largest (values below 2^L are then rounded to it;
    only aggregates with one output row take such a window)."""
    e = [(int(b) >> 52 & 0x7FF) for b in np.asarray(x, np.float64).view(np.uint64)
         if (int(b) >> 52 & 0x7FF) != 0x7FF and (int(b) & 0x7FFFFFFFFFFFFFFF)]
    if not e:
        return None
    lsb = [max(v, 1) - 1075 for v in e]
    emin, emax = min(lsb), max(lsb)
    return emin if (emax - emin + 115) // 32 + 1 <= 15 else emax - 364


def window_sum(x, L) -> float:
    """The correctly rounded exact sum of the values, each first rounded to a
    multiple of 2^L (ties to even): what the exact accumulator computes."""
    from fractions import Fraction
    q = Fraction(2) ** L
    return float(sum((round(Fraction(float(v)) / q) * q for v in x), Fraction(0)))


def test_host_exact_sum_wide_exponents():
    # 1e-25 next to 1e25 spans ~166 binades: exact
    x = np.array([1e25, 1e-25, -1e25])
    assert host_exact_sum(x) == math.fsum(x) == 1e-25
    # spans up to 364 binades stay exact
    rng = np.random.default_rng(12)
    for _ in range(50):
        n = int(rng.integers(1, 60))
        x = rng.choice([-1.0, 1.0], n) * 10.0 ** rng.uniform(-50, 50, n)
        assert host_exact_sum(x) == math.fsum(x)
    # wider columns (one output row only): values below the window's LSB are
    # rounded to it; the sum is the exact sum of the rounded values, within
    # n * 2^(L-1) of the exact one
    for _ in range(50):
        n = int(rng.integers(1, 60))
        x = rng.choice([-1.0, 1.0], n) * 10.0 ** rng.uniform(-300, 300, n)
        L = window_lsb(x)
        got = host_exact_sum(x)
        assert got == window_sum(x, L)
        assert abs(got - math.fsum(x)) <= n * 2.0 ** (L - 1) + abs(math.fsum(x)) * 2.0 ** -52
        y = np.abs(x)  # no cancellation: within 1e-12 relative of the exact sum
        assert abs(host_exact_sum(y) - math.fsum(y)) <= 1e-12 * math.fsum(y)
    # 1.0 next to a subnormal: the subnormal rounds away, ties to even
    assert host_exact_sum(np.array([1.0, 5e-324])) == 1.0


def test_host_exact_sum_cancellation_and_signs():
    # the conversion of the accumulator to a double: negative totals, totals
    # far below the terms (carries through many digits), ties at every
    # position of the mantissa, and subnormal totals
    rng = np.random.default_rng(77)
    for _ in range(300):
        n = int(rng.integers(2, 40))
        big = rng.choice([-1.0, 1.0], n) * 2.0 ** rng.integers(-60, 90, n) * (1 + rng.random(n))
        x = np.concatenate([big, -big[: n // 2], rng.normal(size=3) * 2.0 ** int(rng.integers(-80, 80))])
        rng.shuffle(x)
        assert host_exact_sum(x) == math.fsum(x), x.tolist()
    for k in range(1, 60):  # 1 + 2^-53 * (odd/even) ties and near-ties, both signs
        for s in (1.0, -1.0):
            for extra in (0.0, 2.0 ** (-53 - k), -(2.0 ** (-53 - k))):
                x = np.array([s, s * 2.0 ** -53, s * extra, s * 2.0 ** -52])
                assert host_exact_sum(x) == math.fsum(x), (k, s, extra)
    x = np.array([3e-308, -2.9e-308, 1e-320])
    assert host_exact_sum(x) == math.fsum(x)
    assert host_exact_sum(np.array([-5e-324])) == -5e-324
    assert host_exact_sum(np.array([2.0 ** 60, -2.0 ** 60])) == 0.0


def test_generators_are_deterministic():
    K = dg.keys(1000, 40, seed_=5)
    assert len(set(K.tolist())) == 1000 and K.min() >= 0 and K.max() < 2 ** 40
    a, b = dg.c2_table(50_000, K), dg.c2_table(50_000, K)
    for k in a:
        assert np.array_equal(a[k], b[k])
    assert set(np.unique(a["key"]).tolist()) <= set(K.tolist())
    assert a["f"].min() >= 0 and a["f"].max() < 100
    assert abs(a["v2"].mean()) < 0.05 and abs(a["v2"].std() - 1) < 0.05
    # Zipf(1.1) over 1000 ranks: the top rank holds 1 / sum_k k^-1.1 of the rows
    top = np.bincount(np.searchsorted(np.sort(K), a["key"])).max() / a["key"].size
    expect = 1.0 / np.sum(np.arange(1, 1001, dtype=np.float64) ** -1.1)
    assert abs(top - expect) < 0.01
    p = dg.c3_probe(100_000, K)
    assert 0.88 < np.isin(p["pk"], K).mean() < 0.92
    c1 = dg.c1_table(10_000)
    assert sorted(c1["id"].tolist()) == list(range(10_000))
    assert c1["dept"].max() == 11 and 20000 <= c1["salary"].min() and c1["salary"].max() <= 200000


def test_pack_validity_lsb_first():
    v = np.zeros(130, bool)
    v[[0, 63, 64, 129]] = True
    w = pack_validity(v)
    assert w.tolist() == [1 | (1 << 63), 1, 2]


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_server_open_refuses_unknown_flags_without_gpu():
    lib = _abi.load()
    h = C.c_void_p()
    assert lib.fgpu_server_open(0, 4, 1 << 30, 4, C.byref(h)) != 0
    assert b"flags" in lib.fgpu_last_error()
    assert not h.value
