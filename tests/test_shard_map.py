"""The shard splitter (pga_destination_map) against the reference's S-08
cases (tests/unit/test_dsort.c:134-258): structural invariants for several
histogram shapes and world sizes, and its two pinned expectations.  Host code
only: runs without a GPU."""
import numpy as np
import pytest

from distributedpanga_amd import pga


def check_map(hist, ws):
    ks, sel = pga.destination_map(hist, ws)
    ks2, sel2 = pga.destination_map(hist, ws)
    assert np.array_equal(ks, ks2) and np.array_equal(sel, sel2)      # deterministic
    assert sel.min() >= 0 and sel.max() < ws                           # in range
    assert np.all(np.diff(sel) >= 0)                                   # monotone
    assert ks[0] == 0 and ks[ws] == 1024 and np.all(np.diff(ks) >= 0)  # consistent
    return ks, sel


@pytest.mark.parametrize("ws", [1, 3, 4, 1024])
def test_uniform(ws):
    check_map(np.ones(1024, np.int64), ws)


def test_uniform_four_quarters():
    _, sel = check_map(np.ones(1024, np.int64), 4)
    assert np.array_equal(sel, np.arange(1024) // 256)


@pytest.mark.parametrize("ws", [3, 4])
def test_single_bucket(ws):
    h = np.zeros(1024, np.int64)
    h[500] = 1000
    check_map(h, ws)


@pytest.mark.parametrize("ws", [1, 4])
def test_empty(ws):
    check_map(np.zeros(1024, np.int64), ws)


def test_two_buckets_r01_fixture():
    h = np.zeros(1024, np.int64)
    h[768] = 4
    h[770] = 4
    _, sel = check_map(h, 4)
    assert (sel[768], sel[769], sel[770]) == (0, 1, 2)


def test_random_histograms_balance():
    rng = np.random.default_rng(4)
    for ws in (2, 5, 8, 16):
        h = rng.integers(0, 10_000, 1024).astype(np.int64)
        ks, sel = check_map(h, ws)
        per = np.bincount(sel, weights=h, minlength=ws)
        # every shard within one bucket of its share
        assert np.all(np.abs(per - h.sum() / ws) <= h.max())
        for r in range(ws):
            assert np.all(sel[ks[r]:ks[r + 1]] == r)
