"""The oracle (CPU) and the HIP path (GPU) against fixtures written by the
compiled reference (tests/golden/make_golden.py)."""
import glob
import os

import numpy as np
import pytest

import oracle
from distributedpanga_amd import synth

# the fixtures tests/golden/make_golden.py writes (other .npz files there belong
# to other tests)
GOLD = sorted(p for p in glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz"))
              if not os.path.basename(p).startswith("ref_pair_1gdb"))


def load(path):
    z = np.load(path)          # plain arrays only (allow_pickle is off)
    g = synth.Gdb(bps=z["bps"], boff=z["boff"], clen=z["clen"], n_a=int(z["n_a"]))
    cum = np.zeros(1 << 24, np.int64)
    idx, val = z["cum_idx"], z["cum_val"]
    # the table is a step function: value val[k] from idx[k] up to idx[k+1]
    bounds = np.concatenate([idx, [1 << 24]])
    for k in range(len(idx)):
        cum[bounds[k]:bounds[k + 1]] = val[k]
    return g, z, cum


def test_fixtures_present():
    assert len(GOLD) >= 3


@pytest.mark.parametrize("path", GOLD, ids=[os.path.basename(p) for p in GOLD])
def test_oracle_matches_reference_fixture(path):
    g, z, cum = load(path)
    rec, _, (P, Cb) = oracle.extract(g.bps, g.boff[:-1], g.clen, mode=0)
    assert (P, Cb) == (int(z["P"]), int(z["C"]))
    assert np.array_equal(rec, z["extracted"])
    rec1, _, _ = oracle.extract(g.bps, g.boff[:-1], g.clen, mode=1)
    assert np.array_equal(rec1, z["extracted"])
    o = oracle.pipeline(g.bps, g.boff[:-1], g.clen)
    assert np.array_equal(o["records"], z["sorted_canon"])
    assert np.array_equal(o["records"][:, 0], z["lcp"])
    assert np.array_equal(o["entries"], z["entries"])
    assert np.array_equal(o["cum"], cum)
    assert o["maxpre"] == int(z["maxpre"])


@pytest.mark.gpu
@pytest.mark.parametrize("path", GOLD, ids=[os.path.basename(p) for p in GOLD])
def test_gpu_matches_reference_fixture(path):
    from distributedpanga_amd import pga
    g, z, cum = load(path)
    print(f"[golden] {os.path.basename(path)}", flush=True)
    got = pga.run_pair(g)
    assert (got["info"]["P"], got["info"]["C"]) == (int(z["P"]), int(z["C"]))
    assert np.array_equal(got["entries"], z["entries"])
    assert np.array_equal(got["cum"], cum)
    assert got["info"]["maxpre"] == int(z["maxpre"])
