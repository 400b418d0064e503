"""GPU: the diagonal-sorted seed array (pga_seed_sort) against the oracle's
reimport + per-panel sort, byte for byte, for both seed streams."""
import numpy as np
import pytest

import oracle
from distributedpanga_amd import pga, synth

pytestmark = pytest.mark.gpu


def _check(g, freq=10, **opts):
    want = oracle.pipeline(g.bps, g.boff[:-1], g.clen, freq=freq)
    with pga.Context(freq=freq, **opts) as ctx:
        ctx.load_gdb(g)
        ctx.build_index()
        ctx.merge()
        for comp, stream in ((0, want["nseeds"]), (1, want["cseeds"])):
            print(f"[seed-sort] ncontig={g.ncontig} bases={g.seqtot} comp={comp} opts={opts}", flush=True)
            recs, sw = oracle.seed_sort(stream, want["P"], want["C"], g.ncontig, g.maxctg, comp)
            got = ctx.seed_sort(0, comp)
            assert got.shape == recs.shape
            assert np.array_equal(got, recs)


def test_workload1_small():
    _check(synth.workload(1, 0.05))


@pytest.mark.parametrize("tile", [2048, 8192])
def test_workload1_tiles(tile):
    _check(synth.workload(1, 0.3), tile_records=tile)


def test_workload1_wide_second_digit():
    _check(synth.workload(1, 0.3), tile_records=2048, digit2_bits=10)


def test_many_contigs():
    rng = np.random.default_rng(21)
    contigs = [rng.integers(0, 4, int(n)).astype(np.uint8) for n in rng.integers(50, 3000, 400)]
    _check(synth.from_codes(contigs))


def test_low_complexity_equal_records():
    contigs = [np.tile(np.array([0, 1, 2], np.uint8), 4000), np.zeros(6000, np.uint8),
               np.random.default_rng(2).integers(0, 4, 8000).astype(np.uint8)]
    _check(synth.from_codes(contigs), freq=24)


def test_workload2_scaled():
    _check(synth.workload(2, 0.05))


def test_no_seeds():
    g = synth.from_codes([np.random.default_rng(5).integers(0, 4, 30).astype(np.uint8)])
    with pga.Context() as ctx:
        ctx.load_gdb(g)
        ctx.build_index()
        ctx.merge()
        assert ctx.seed_sort(0, 0).size == 0
