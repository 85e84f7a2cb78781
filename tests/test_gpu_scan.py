"""GPU: the step's device-wide prefix sum (rbh_prefix_sum, include/rbh_testing.h).

The step scans per-row pair counts, per-pair slot and block counts, per-side
candidate counts, the Frank-Wolfe keep flags and the dedup flags (bytes), and
island / unit / body counts.  Capacities below 32768 take one workgroup; from
there on the single-pass look-back scan runs, with tiles of 8192 ints or 16384
bytes.  These cases put the live length on and beside every tile edge, at 0,
at the capacity and over thousands of tiles, and check every output against
NumPy's cumulative sum (exact integers).
"""
import numpy as np
import pytest

from rainbow_amd import native

pytestmark = pytest.mark.gpu

INT_TILE, BYTE_TILE, SINGLE = 8192, 16384, 32768


@pytest.fixture(scope="module")
def ctx():
    c = native.Context(0, "fp32")
    yield c
    c.close()


def expected(a):
    out = np.zeros(a.size + 1, dtype=np.int64)
    np.cumsum(a, dtype=np.int64, out=out[1:])
    return out


def check(ctx, a, cap):
    out, total = ctx.prefix_sum(a, cap)
    exp = expected(a)
    assert total == exp[-1]
    bad = np.nonzero(out.astype(np.int64) != exp)[0]
    assert bad.size == 0, "n=%d cap=%d: first mismatch at %d (%d vs %d)" % (
        a.size, cap, bad[0], out[bad[0]], exp[bad[0]])


BYTE_CASES = [(0, 1), (1, 1), (4095, 4096), (4096, 4096), (4097, 8192), (SINGLE - 1, SINGLE - 1),
              (0, SINGLE), (1, SINGLE), (BYTE_TILE - 1, SINGLE), (BYTE_TILE, SINGLE), (BYTE_TILE + 1, SINGLE),
              (SINGLE, SINGLE), (5 * BYTE_TILE, 5 * BYTE_TILE + 100), (3_000_017, 4_000_000),
              ((1 << 22) + 7, (1 << 22) + 7)]
INT_CASES = [(0, SINGLE), (INT_TILE - 1, SINGLE), (INT_TILE, SINGLE), (INT_TILE + 1, SINGLE), (SINGLE, SINGLE),
             (65536, 100_000), (300 * INT_TILE, 300 * INT_TILE), (4_000_001, 5_000_000), (2000, 2000)]


@pytest.mark.parametrize("n,cap", BYTE_CASES)
def test_byte_flags(ctx, n, cap):
    rng = np.random.default_rng(n + 7 * cap)
    check(ctx, (rng.random(n) < 0.3).astype(np.uint8), cap)


def test_byte_values_full_range(ctx):
    rng = np.random.default_rng(3)
    for n, cap in [(BYTE_TILE * 3 + 5, BYTE_TILE * 4), (100_000, 100_000)]:
        check(ctx, rng.integers(0, 256, n).astype(np.uint8), cap)


@pytest.mark.parametrize("n,cap", INT_CASES)
def test_int_counts(ctx, n, cap):
    rng = np.random.default_rng(n + 11 * cap)
    check(ctx, rng.integers(0, 300, n).astype(np.int32), cap)


def test_all_zero_and_all_ones(ctx):
    for n in (SINGLE + 3, 10 * BYTE_TILE):
        check(ctx, np.zeros(n, dtype=np.uint8), n)
        check(ctx, np.ones(n, dtype=np.uint8), n)
        check(ctx, np.ones(n, dtype=np.int32), n + INT_TILE)


def test_repeated_calls(ctx):
    # every call clears its own look-back state: a second, shorter call is not
    # confused by the first one's published tiles
    rng = np.random.default_rng(5)
    for n in (40 * INT_TILE + 3, 7 * INT_TILE, 40 * INT_TILE + 3):
        check(ctx, rng.integers(0, 50, n).astype(np.int32), 41 * INT_TILE)


def test_arguments_checked(ctx):
    with pytest.raises(native.RbhError):
        ctx.prefix_sum(np.ones(10, dtype=np.int32), cap=5)
    with pytest.raises(TypeError):
        ctx.prefix_sum(np.ones(10, dtype=np.float32))
