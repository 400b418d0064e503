"""GPU parity: the HIP path through the C-ABI against the oracle, byte for byte.

Outputs compared: GIX entries (k-mer suffix, mask, LCP, payload), the 2^24
prefix table, the largest prefix group, the N/C seed streams in emission order
and the per-anchor-contig seed counts.  Integer work: the bar is bit-exact.
"""
import numpy as np
import pytest

import oracle
from distributedpanga_amd import pga, synth

pytestmark = pytest.mark.gpu


def _compare(g, freq=10, **opts):
    print(f"[parity] ncontig={g.ncontig} bases={g.seqtot} opts={opts} freq={freq}", flush=True)
    got = pga.run_pair(g, freq=freq, **opts)
    want = oracle.pipeline(g.bps, g.boff[:-1], g.clen, freq=freq)
    assert got["info"]["nrec"] == want["nrec"]
    assert (got["info"]["P"], got["info"]["C"]) == (want["P"], want["C"])
    assert np.array_equal(got["entries"], want["entries"])
    if got["cum"] is not None:
        assert np.array_equal(got["cum"], want["cum"])
        assert got["info"]["maxpre"] == want["maxpre"]
    assert np.array_equal(got["nseeds"], want["nseeds"])
    assert np.array_equal(got["cseeds"], want["cseeds"])
    assert np.array_equal(got["nbuck"], want["nbuck"])
    assert np.array_equal(got["cbuck"], want["cbuck"])
    return got, want


def xorshift_seq(n, seed):
    x = seed or 1
    out = np.zeros(n, np.uint8)
    for i in range(n):
        x ^= (x << 13) & 0xFFFFFFFF
        x ^= x >> 17
        x ^= (x << 5) & 0xFFFFFFFF
        out[i] = x & 3
    return out


def test_e06_contig():
    g = synth.from_codes([xorshift_seq(4000, 20260910)])
    got, want = _compare(g)
    assert got["info"]["nrec"] == 3106


def test_workload1_small():
    _compare(synth.workload(1, 0.05))


@pytest.mark.parametrize("tile", [2048, 4096, 8192])
def test_tile_sizes(tile):
    _compare(synth.workload(1, 0.2), tile_records=tile)


@pytest.mark.parametrize("tile,chunk,digit2", [(2048, 4096, 9), (2048, 2048, 10), (4096, 4096, 10),
                                               (4096, 8192, 10), (2048, 8192, 9), (2048, 8192, 11),
                                               (4096, 4096, 11), (2048, 2048, 11)])
def test_second_digit_widths(tile, chunk, digit2):
    """A 9- or 10-bit second digit: buckets of 2^(8+w) at the second level,
    bins of tiles split there cut from a wider digit span."""
    got, _ = _compare(synth.workload(1, 0.2), tile_records=tile, chunk_records=chunk, digit2_bits=digit2)
    assert got["timings"]["partition_levels"] >= 2


@pytest.mark.parametrize("tile,chunk", [(2048, 2048), (2048, 8192), (4096, 8192)])
def test_second_digit_bytes_from_extraction(tile, chunk):
    """An 8-bit second digit: the extraction also writes every record's
    second digit (whole 32-byte blocks a bucket run, single bytes at a run's
    ends), and the second level counts those bytes instead of the records."""
    got, _ = _compare(synth.workload(1, 1.0), tile_records=tile, chunk_records=chunk, digit2_bits=8)
    assert got["timings"]["partition_levels"] >= 2


def test_second_digit_bytes_low_complexity():
    """The same where rounds of the extraction overflow its stage (poly-A and
    short tandem repeats): those rounds write every record and its digit
    byte straight to its slot."""
    contigs = [np.zeros(30000, np.uint8), np.tile(np.array([0, 1], np.uint8), 12000),
               np.tile(np.array([2, 2, 3], np.uint8), 5000)]
    rng = np.random.default_rng(5)
    contigs += [rng.integers(0, 4, 20000).astype(np.uint8) for _ in range(3)]
    _compare(synth.from_codes(contigs), tile_records=2048, digit2_bits=8)


@pytest.mark.parametrize("chunk,mt", [(2048, 512), (8192, 1024), (4096, 256)])
def test_chunk_and_merge_tiles(chunk, mt):
    _compare(synth.workload(1, 0.2), chunk_records=chunk, merge_tile=mt)


@pytest.mark.parametrize("freq", [1, 2, 4, 25, 64])
def test_frequency_cuts(freq):
    _compare(synth.workload(1, 0.1), freq=freq)


def test_edge_lengths():
    rng = np.random.default_rng(17)
    contigs = [rng.integers(0, 4, n).astype(np.uint8) for n in
               [1, 5, 11, 12, 13, 39, 40, 41, 52, 67, 68, 100, 255, 256, 257, 1023, 4099]]
    _compare(synth.from_codes(contigs))


def test_no_records():
    _compare(synth.from_codes([np.zeros(5, np.uint8), np.zeros(11, np.uint8)]))


def test_low_complexity_deep_partition():
    """Identical k-mers beyond one tile force every partition level down to the payload bits.

    The poly-A contig selects both strands at almost every position (1.97
    records per position), so its extraction blocks emit more records than the
    emit pass stages in LDS and the rest take the direct-store path."""
    contigs = [np.zeros(30000, np.uint8), np.tile(np.array([0, 1], np.uint8), 12000),
               np.tile(np.array([2, 2, 3], np.uint8), 5000)]
    rng = np.random.default_rng(3)
    contigs += [rng.integers(0, 4, 5000).astype(np.uint8) for _ in range(3)]
    _compare(synth.from_codes(contigs), tile_records=2048)


@pytest.mark.parametrize("digit2", [8, 10])
def test_identical_kmers_stay_in_place(digit2):
    """30 kb of poly-A: the forward records share one k-mer, the reverse ones
    another, and each group is beyond a tile.  Below the first digit every
    segment has one non-empty bucket until the payload digits, so those levels
    leave the records where they are: only levels that split move records."""
    g = synth.from_codes([np.zeros(30000, np.uint8)])
    got, _ = _compare(g, tile_records=2048, digit2_bits=digit2)
    t, nrec = got["timings"], got["info"]["nrec"]
    assert t["partition_levels"] >= 12
    assert t["scatter_bytes"] <= 4 * 2 * 16 * nrec


def test_many_contigs_two_byte_ids():
    rng = np.random.default_rng(8)
    contigs = [rng.integers(0, 4, int(n)).astype(np.uint8) for n in rng.integers(12, 400, 300)]
    got, _ = _compare(synth.from_codes(contigs))
    assert got["info"]["C"] == 2


def test_workload1_full():
    _compare(synth.workload(1, 1.0))


def test_workload2_scaled():
    _compare(synth.workload(2, 0.05))


def test_batch_equals_single():
    gdbs = synth.batch(6, scale=0.02)
    with pga.Context() as ctx:
        ctx.load_batch(gdbs)
        ctx.build_index()
        ctx.merge()
        for p, g in enumerate(gdbs):
            print(f"[batch] pair {p}", flush=True)
            ent, cum, info = ctx.index_export(p)
            N, Cs, nb, cb = ctx.seeds_export(p)
            want = oracle.pipeline(g.bps, g.boff[:-1], g.clen)
            assert np.array_equal(ent, want["entries"])
            assert np.array_equal(cum, want["cum"]) and info["maxpre"] == want["maxpre"]
            assert np.array_equal(N, want["nseeds"]) and np.array_equal(Cs, want["cseeds"])
            assert np.array_equal(nb, want["nbuck"]) and np.array_equal(cb, want["cbuck"])


def test_context_reuse_smaller_second_call():
    a = synth.workload(1, 0.3)
    b = synth.workload(1, 0.05)
    with pga.Context() as ctx:
        for g in (a, b, a):
            ctx.load_gdb(g)
            ctx.build_index()
            ctx.merge()
            ent, cum, _ = ctx.index_export(0)
            N, Cs, nb, cb = ctx.seeds_export(0)
            want = oracle.pipeline(g.bps, g.boff[:-1], g.clen)
            assert np.array_equal(ent, want["entries"]) and np.array_equal(cum, want["cum"])
            assert np.array_equal(N, want["nseeds"]) and np.array_equal(Cs, want["cseeds"])


def test_verify_and_bucket_counts():
    """Device sortedness/LCP check is clean on a built index, and the 5-base
    bucket counts equal the oracle's record_bucket tally (src/dsort.h:30-33)."""
    g = synth.workload(1, 0.2)
    want = oracle.pipeline(g.bps, g.boff[:-1], g.clen)
    rec = want["records"]
    wb = np.bincount((rec[:, 1].astype(np.int64) << 2) | (rec[:, 2].astype(np.int64) >> 6), minlength=1024)
    with pga.Context() as ctx:
        ctx.load_gdb(g)
        ctx.build_index()
        assert ctx.verify(0) == (0, -1, 0)
        assert np.array_equal(ctx.bucket_counts(0), wb)


def test_human_scale_workload_reduced():
    """The human-scale pair's shape (long contigs, 45% repeats from 500
    families, satellites) at 1/100 of its length."""
    got, _ = _compare(synth.workload(4, 0.01))
    assert got["info"]["nrec"] > 40_000_000


def test_batch_mixed_table_paths():
    """One pair with many records per prefix (its table filled from group
    ends) batched with small pairs (tables filled prefix by prefix)."""
    gdbs = [synth.workload(1, 0.05), synth.workload(4, 0.01), synth.workload(1, 0.02)]
    with pga.Context() as ctx:
        ctx.load_batch(gdbs)
        ctx.build_index()
        ctx.merge()
        for p, g in enumerate(gdbs):
            print(f"[mixed] pair {p}", flush=True)
            ent, cum, info = ctx.index_export(p)
            N, Cs, nb, cb = ctx.seeds_export(p)
            want = oracle.pipeline(g.bps, g.boff[:-1], g.clen)
            assert np.array_equal(ent, want["entries"])
            assert np.array_equal(cum, want["cum"]) and info["maxpre"] == want["maxpre"]
            assert np.array_equal(N, want["nseeds"]) and np.array_equal(Cs, want["cseeds"])
            assert np.array_equal(nb, want["nbuck"]) and np.array_equal(cb, want["cbuck"])


def test_batch_with_empty_and_tiny_pairs():
    """Pairs with blocks but no records (30 bases: candidate offsets, no
    40-mer), with no blocks at all (10 bases), and with a handful of records,
    batched between ordinary pairs: their extraction ranges, digit counts and
    bucket starts are empty or tiny next to the others'."""
    rng = np.random.default_rng(44)
    gdbs = [synth.workload(1, 0.02),
            synth.from_codes([rng.integers(0, 4, 30).astype(np.uint8)]),
            synth.from_codes([rng.integers(0, 4, 10).astype(np.uint8)]),
            synth.from_codes([rng.integers(0, 4, 90).astype(np.uint8), rng.integers(0, 4, 45).astype(np.uint8)]),
            synth.workload(1, 0.01)]
    with pga.Context(tile_records=2048) as ctx:
        ctx.load_batch(gdbs)
        ctx.build_index()
        ctx.merge()
        for p, g in enumerate(gdbs):
            print(f"[empty-tiny] pair {p} bases={g.seqtot}", flush=True)
            ent, cum, info = ctx.index_export(p)
            N, Cs, nb, cb = ctx.seeds_export(p)
            want = oracle.pipeline(g.bps, g.boff[:-1], g.clen)
            assert info["nrec"] == want["nrec"]
            assert np.array_equal(ent, want["entries"])
            assert np.array_equal(cum, want["cum"]) and info["maxpre"] == want["maxpre"]
            assert np.array_equal(N, want["nseeds"]) and np.array_equal(Cs, want["cseeds"])
            assert np.array_equal(nb, want["nbuck"]) and np.array_equal(cb, want["cbuck"])
