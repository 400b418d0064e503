"""pga_run_graph (the device-driven build + merge, captured in a HIP graph and
replayed) against the oracle and against the eager calls, through the C-ABI."""
import numpy as np
import pytest

import oracle
from distributedpanga_amd import pga, synth

pytestmark = pytest.mark.gpu


def _want(g, freq=10):
    return oracle.pipeline(g.bps, g.boff[:-1], g.clen, freq=freq)


def _check(ctx, p, want):
    ent, cum, info = ctx.index_export(p)
    assert info["nrec"] == want["nrec"] and (info["P"], info["C"]) == (want["P"], want["C"])
    assert np.array_equal(ent.reshape(-1), want["entries"].reshape(-1)), f"pair {p}: entries"
    assert np.array_equal(cum, want["cum"]) and info["maxpre"] == want["maxpre"], f"pair {p}: prefix table"
    N, Cs, nb, cb = ctx.seeds_export(p)
    assert np.array_equal(N, want["nseeds"]) and np.array_equal(Cs, want["cseeds"]), f"pair {p}: seeds"
    assert np.array_equal(nb, want["nbuck"]) and np.array_equal(cb, want["cbuck"]), f"pair {p}: bucks"


def _batch():
    gs = [synth.workload(1, 0.05)] + synth.batch(5, scale=0.03)
    gs.append(synth.from_codes([np.zeros(0, np.uint8)]))                 # a pair with no records
    gs.append(synth.from_codes([np.random.default_rng(3).integers(0, 4, 60).astype(np.uint8)]))
    return gs


def test_graph_replays_match_the_oracle_and_the_eager_path():
    gs = _batch()
    wants = [_want(g) for g in gs]
    with pga.Context(tile_records=2048, digit2_bits=8) as ctx:
        ctx.load_batch(gs)
        for rep in range(3):
            print(f"[graph] replay {rep}", flush=True)
            ctx.run_graph()
            assert ctx.timings()["graph_fallbacks"] == 0
            for p, w in enumerate(wants):
                _check(ctx, p, w)
            # the program's own eager work between replays
            ctx.build_index()
            ctx.merge()
            ctx.sync()
            for p, w in enumerate(wants):
                _check(ctx, p, w)


def test_graph_recaptured_after_a_new_batch_and_new_options():
    with pga.Context(tile_records=4096, digit2_bits=10) as ctx:
        a = synth.batch(4, scale=0.02, seed=0x5EED0101)
        ctx.load_batch(a)
        ctx.run_graph()
        for p, g in enumerate(a):
            _check(ctx, p, _want(g))
        b = [synth.workload(1, 0.1), synth.workload(1, 0.02)]
        ctx.load_batch(b)                                   # a new batch: a new graph
        ctx.run_graph()
        ctx.run_graph()
        for p, g in enumerate(b):
            _check(ctx, p, _want(g))
        ctx.set_opts(tile_records=2048, merge_tile=256, digit2_bits=9)
        ctx.run_graph()
        for p, g in enumerate(b):
            _check(ctx, p, _want(g))
        assert ctx.timings()["graph_fallbacks"] == 0


def test_graph_falls_back_on_a_deep_partition():
    """Poly-A runs longer than a tile need more partition levels than the
    graph runs: the call redoes the work on the eager path."""
    g = synth.from_codes(["A" * 60000, "ACGT" * 3000, "A" * 20000])
    with pga.Context() as ctx:
        ctx.load_gdb(g)
        ctx.run_graph()
        assert ctx.timings()["graph_fallbacks"] == 1
        _check(ctx, 0, _want(g))


def test_graph_frequency_cut():
    gs = synth.batch(3, scale=0.02, seed=0x5EED0202)
    with pga.Context(freq=25) as ctx:
        ctx.load_batch(gs)
        ctx.run_graph()
        for p, g in enumerate(gs):
            _check(ctx, p, _want(g, freq=25))
