"""The allocation-free entry point lal_exclusive_scan_i64_noalloc against
lal_exclusive_scan_i64, through the C ABI: bit-identical results on two calls
in a row (the second smaller, so stale tile sums of the first would show),
n = 0, a single tile, a partial last tile, and the refusal of n beyond
LAL_SCAN_NOALLOC_MAX_N (checked on the host before any launch)."""
import numpy as np
import pytest
import torch

from linalglib_amd import lal
from tests.gpu_util import host, log

pytestmark = pytest.mark.gpu

MAX_N = 67108864  # LAL_SCAN_NOALLOC_MAX_N in include/lal.h


@pytest.fixture(scope="module")
def h():
    return lal.Handle()


def _pair(h, x):
    a = torch.from_numpy(x).cuda()
    b = a.clone()
    lal.call("lal_exclusive_scan_i64", h.h, len(x), lal.ptr(a))
    lal.call("lal_exclusive_scan_i64_noalloc", h.h, len(x), lal.ptr(b))
    return host(a), host(b)


def test_noalloc_scan_matches_allocating_scan_twice(h):
    rng = np.random.default_rng(11)
    for n in (5_000_001, 1_234_567):  # second call smaller than the first
        log(f"noalloc scan n={n}")
        x = rng.integers(-(1 << 40), 1 << 40, n, dtype=np.int64)
        ref, got = _pair(h, x)
        assert np.array_equal(got, ref), n
        want = np.concatenate(([0], np.cumsum(x)[:-1]))
        assert np.array_equal(got, want), n


def test_noalloc_scan_small_and_edge_sizes(h):
    rng = np.random.default_rng(12)
    for n in (1, 2047, 2048, 2049, 100_000):
        x = rng.integers(0, 1000, n, dtype=np.int64)
        ref, got = _pair(h, x)
        assert np.array_equal(got, ref), n
        assert got[0] == 0 and (n == 1 or got[-1] == x[:-1].sum()), n
    # n = 0 is a no-op and passes no pointer
    lal.call("lal_exclusive_scan_i64_noalloc", h.h, 0, None)


def test_noalloc_scan_at_the_limit_and_refusal(h):
    x = torch.ones(MAX_N, dtype=torch.int64, device="cuda")
    lal.call("lal_exclusive_scan_i64_noalloc", h.h, MAX_N, lal.ptr(x))
    torch.cuda.synchronize()
    assert x[0].item() == 0 and x[-1].item() == MAX_N - 1
    assert bool((x[1:] - x[:-1] == 1).all().item())
    del x
    # one entry too many: refused on the host, nothing launched
    st = lal.load().lal_exclusive_scan_i64_noalloc(h.h, MAX_N + 1, lal.ptr(
        torch.zeros(1, dtype=torch.int64, device="cuda")))
    assert st == 1  # LAL_ERR_INVALID
    assert b"LAL_SCAN_NOALLOC_MAX_N" in lal.load().lal_last_error()
