"""Parity at full size: the 100 Mbp pair and the 64-pair batch byte for byte
against the oracle's outputs, and the human-scale pair (6.2 Gbp, 4.84 G
records) through size-independent properties plus exact checks of sampled
24-bit prefix groups.

The checker at these sizes is oracle/pga_cpu.c: the oracle's own code run on
several host threads, which tests/test_cpu_baseline.py holds equal to the
single-threaded oracle."""
import os
import time

import numpy as np
import pytest

import oracle
from distributedpanga_amd import pga, synth

pytestmark = pytest.mark.gpu

THREADS = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)


def _check_pair(ctx, p, g, freq=10, want=None):
    if want is None:
        want = oracle.pipeline(g.bps, g.boff[:-1], g.clen, freq=freq, threads=THREADS)
    ent, cum, info = ctx.index_export(p)
    assert info["nrec"] == want["nrec"] and (info["P"], info["C"]) == (want["P"], want["C"])
    assert np.array_equal(ent.reshape(-1), want["entries"].reshape(-1)), f"pair {p}: entries"
    assert np.array_equal(cum, want["cum"]) and info["maxpre"] == want["maxpre"], f"pair {p}: prefix table"
    N, Cs, nb, cb = ctx.seeds_export(p)
    assert np.array_equal(N, want["nseeds"]) and np.array_equal(Cs, want["cseeds"]), f"pair {p}: seeds"
    assert np.array_equal(nb, want["nbuck"]) and np.array_equal(cb, want["cbuck"]), f"pair {p}: bucks"
    return want


def test_100mbp_pair_full_size():
    g = synth.workload(2, 1.0)
    print(f"[fullsize] 100 Mbp pair: {g.seqtot} bases, checker on {THREADS} threads", flush=True)
    with pga.Context(digit2_bits=10) as ctx:
        ctx.load_gdb(g)
        ctx.build_index()
        ctx.merge()
        _check_pair(ctx, 0, g)


def test_64_pair_batch_full_size():
    gs = synth.batch(64, scale=1.0)
    print(f"[fullsize] 64-pair batch: {sum(x.seqtot for x in gs)} bases", flush=True)
    with pga.Context(tile_records=2048, chunk_records=8192, digit2_bits=8) as ctx:
        ctx.load_batch(gs)
        ctx.build_index()
        ctx.merge()
        wants = [_check_pair(ctx, p, g) for p, g in enumerate(gs)]
        # the same batch through the captured device-driven pipeline
        ctx.run_graph()
        ctx.run_graph()
        assert ctx.timings()["graph_fallbacks"] == 0
        for p, g in enumerate(gs):
            _check_pair(ctx, p, g, want=wants[p])


def _groups(cum, rng, n):
    """Sampled 24-bit groups [g0, g1) with two or more entries, plus the largest."""
    ends = cum
    starts = np.concatenate([[0], cum[:-1]])
    size = ends - starts
    multi = np.nonzero(size >= 2)[0]
    pick = rng.choice(multi, size=min(n, len(multi)), replace=False)
    pick = np.concatenate([pick, np.argsort(size)[-20:]])          # the largest groups too
    return [(int(starts[p]), int(ends[p])) for p in np.unique(pick)]


def test_human_scale_pair_full_size():
    t = time.perf_counter()
    g = synth.workload(4, 1.0)
    print(f"[fullsize] human-scale pair: {g.seqtot} bases generated in {time.perf_counter() - t:.1f}s", flush=True)
    with pga.Context(tile_records=4096, chunk_records=8192, merge_tile=512, digit2_bits=8) as ctx:
        ctx.load_gdb(g)
        ctx.build_index()
        ctx.merge()
        info = ctx.index_info(0)
        sinfo = ctx.seeds_info(0)
        # the device check: order and every LCP
        assert ctx.verify(0) == (0, -1, 0)
        # every record counted: the CPU's bucket tally of the extraction, from
        # the GPU's prefix table
        t = time.perf_counter()
        want_b = oracle.bucket_counts(g.bps, g.boff[:-1], g.clen, threads=THREADS)
        print(f"[fullsize] CPU bucket counts in {time.perf_counter() - t:.1f}s", flush=True)
        got_b = ctx.bucket_counts(0)
        assert np.array_equal(got_b, want_b) and got_b.sum() == info["nrec"] == 4839854061
        # every seed counted once in its anchor contig's buck
        s = info["s"]
        nbuck, cbuck = ctx.bucks(0)
        assert nbuck.sum() * s == sinfo["nlen"] and cbuck.sum() * s == sinfo["clen"]
        # exact seeds of sampled prefix groups: the merge never crosses a group
        # (src/fastga_pipeline.c:304-321), so the oracle's merge of a group's
        # entries alone gives that group's seeds
        cum = ctx.prefix_table(0)
        rng = np.random.default_rng(44)
        groups = _groups(cum, rng, 3000)
        t = time.perf_counter()
        nseed = 0
        for g0, g1 in groups:
            ent = ctx.index_export_range(0, g0, g1 - g0)
            wn, wc, _, _ = oracle.merge(ent, info["P"], info["C"], info["ncontig"], freq=10)
            n0, c0 = ctx.seeds_locate(0, g0)
            n1, c1 = ctx.seeds_locate(0, g1)
            assert (n1 - n0) * s == len(wn) and (c1 - c0) * s == len(wc), (g0, g1)
            assert np.array_equal(ctx.seeds_export_range(0, 0, n0, n1 - n0), wn), (g0, g1)
            assert np.array_equal(ctx.seeds_export_range(0, 1, c0, c1 - c0), wc), (g0, g1)
            nseed += n1 - n0 + c1 - c0
        print(f"[fullsize] {len(groups)} groups, {nseed} seeds checked in {time.perf_counter() - t:.1f}s", flush=True)
        assert nseed > 0
