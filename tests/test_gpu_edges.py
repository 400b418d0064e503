"""GPU edge cases: contigs without sequence, more contigs than the merge's LDS
counters hold, the device check flagging a broken index, and failures forced
from the host (call order, bad arguments, allocation failure)."""
import ctypes as C

import numpy as np
import pytest

import oracle
from distributedpanga_amd import pga, synth

from test_gpu_parity import _compare

pytestmark = pytest.mark.gpu

ENOMEM, EINVAL, EOVERFLOW = -1, -3, -4


def test_contig_without_sequence():
    """boff < 0: the reference skips the contig and keeps its id
    (src/extract.c:97-98); its length still counts for the sizing."""
    g = synth.workload(1, 0.05)
    boff = g.boff.copy()
    boff[3] = -1
    boff[g.ncontig - 1] = -1
    g2 = synth.Gdb(bps=g.bps, boff=boff, clen=g.clen, n_a=g.n_a)
    got, want = _compare(g2)
    assert got["nbuck"][3] == 0 and got["cbuck"][3] == 0


def _many_contigs(n_a, seed):
    rng = np.random.default_rng(seed)
    a = [rng.integers(0, 4, int(n)).astype(np.uint8) for n in rng.integers(150, 2500, n_a)]
    b = []
    for x in a:
        y = x.copy()
        snp = rng.random(len(y)) < 0.03
        y[snp] = (y[snp] + rng.integers(1, 4, int(snp.sum()))) % 4
        b.append(y)
    return synth.from_codes(a + b, n_a=n_a)


def test_more_contigs_than_lds_counters():
    """2600 contigs (C = 2): the merge counts seeds per anchor contig with
    global atomics instead of its 1024 LDS counters."""
    g = _many_contigs(1300, 5)
    got, want = _compare(g)
    assert got["info"]["C"] == 2 and g.ncontig > 1024
    assert want["nbuck"][1024:].sum() > 0 and want["cbuck"][1024:].sum() > 0


def test_more_contigs_than_lds_counters_in_a_batch():
    gs = [_many_contigs(700, 6), synth.workload(1, 0.02), _many_contigs(600, 7)]
    with pga.Context() as ctx:
        ctx.load_batch(gs)
        ctx.build_index()
        ctx.merge()
        for p, g in enumerate(gs):
            want = oracle.pipeline(g.bps, g.boff[:-1], g.clen)
            N, Cs, nb, cb = ctx.seeds_export(p)
            assert np.array_equal(N, want["nseeds"]) and np.array_equal(Cs, want["cseeds"])
            assert np.array_equal(nb, want["nbuck"]) and np.array_equal(cb, want["cbuck"])


def test_verify_flags_an_order_break():
    g = synth.workload(1, 0.05)
    with pga.Context() as ctx:
        ctx.load_gdb(g)
        ctx.build_index()
        assert ctx.verify(0) == (0, -1, 0)
        ent, _, info = ctx.index_export(0, cum=False)
        # an entry inside a prefix group whose predecessor's suffix is not all zero
        lcp = ent[:, 8]
        cand = np.nonzero((lcp[1:] >= 12) & (ent[:-1, :7].max(axis=1) > 0))[0] + 1
        i = int(cand[len(cand) // 2])
        for b in range(7):
            ctx.poke(0, i, b, 0)                   # its k-mer now sorts before the predecessor's
        bad, first, _ = ctx.verify(0)
        assert bad >= 1 and first == i


def test_verify_flags_a_wrong_lcp():
    g = synth.workload(1, 0.05)
    with pga.Context() as ctx:
        ctx.load_gdb(g)
        ctx.build_index()
        ent, _, info = ctx.index_export(0, cum=False)
        j = info["nrec"] // 3
        ctx.poke(0, j, 8, (int(ent[j, 8]) + 1) % 41)
        assert ctx.verify(0) == (0, -1, 1)


def test_call_order_and_arguments():
    L = pga._setup(pga.hip_lib())
    with pga.Context() as ctx:
        h = ctx._h
        assert L.pga_build_index(h) == EINVAL                       # nothing loaded
        g = synth.workload(1, 0.02)
        ctx.load_gdb(g)
        assert L.pga_merge(h) == EINVAL                             # merge before the build
        m = pga.SeedsMeta()
        assert L.pga_seeds_info(h, 0, C.byref(m)) == EINVAL
        n, sw = C.c_int64(), C.c_int()
        assert L.pga_seed_sort(h, 0, 0, C.byref(n), C.byref(sw)) == EINVAL
        ctx.build_index()
        im = pga.IndexMeta()
        assert L.pga_index_info(h, 1, C.byref(im)) == EINVAL       # no pair 1
        assert L.pga_index_poke(h, 0, 1 << 40, 0, 0) == EINVAL
        ctx.merge()
        assert L.pga_seed_sort(h, 0, 2, C.byref(n), C.byref(sw)) == EINVAL
        assert b"pair" in L.pga_last_error(h) or b"stream" in L.pga_last_error(h)


def test_load_failures_forced_from_the_host():
    L = pga._setup(pga.hip_lib())
    with pga.Context() as ctx:
        h = ctx._h
        buf = np.zeros(64, np.uint8)
        u8 = buf.ctypes.data_as(C.POINTER(C.c_uint8))
        boff = np.zeros(1, np.int64)
        i64 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int64))
        # a contig longer than the int32 field of the device contig table
        clen = np.array([1 << 31], np.int64)
        assert L.pga_load_gdb(h, u8, 1 << 40, i64(boff), i64(clen), 1) == EOVERFLOW
        # sequence outside the .bps image
        clen = np.array([1000], np.int64)
        assert L.pga_load_gdb(h, u8, 64, i64(boff), i64(clen), 1) == EINVAL
        # a .bps image larger than the device: the allocation fails before any copy
        assert L.pga_load_gdb(h, u8, 1 << 50, i64(boff), i64(clen), 1) == ENOMEM
        assert b"allocation" in L.pga_last_error(h)
        # the context still works afterwards
        g = synth.workload(1, 0.02)
        ctx.load_gdb(g)
        ctx.build_index()
        assert ctx.verify(0) == (0, -1, 0)


def _repeat_genome(copies, unit_len, div, seed, sat=0):
    """Copies of one element at `div` per-base divergence between random
    spacers (plus, optionally, a satellite of a 171 bp unit): repeated k-mers
    that crowd the tile sort's bins."""
    rng = np.random.default_rng(seed)
    unit = rng.integers(0, 4, unit_len).astype(np.uint8)
    parts = []
    for _ in range(copies):
        c = unit.copy()
        m = rng.random(unit_len) < div
        c[m] = (c[m] + rng.integers(1, 4, int(m.sum()))) % 4
        parts += [rng.integers(0, 4, int(rng.integers(20, 400))).astype(np.uint8), c]
    if sat:
        s171 = rng.integers(0, 4, 171).astype(np.uint8)
        parts.append(np.tile(s171, sat))
    a = np.concatenate(parts)
    b = a.copy()
    m = rng.random(len(b)) < 0.02
    b[m] = (b[m] + rng.integers(1, 4, int(m.sum()))) % 4
    return synth.from_codes([a[: len(a) // 2], a[len(a) // 2:], b], n_a=2)


@pytest.mark.parametrize("tile", [2048, 4096, 8192])
def test_crowded_bins_of_repeats(tile):
    """Hundreds of near-identical element copies and a satellite: bins of more
    than 256 records, sorted by sub-binning (or the merge sort)."""
    g = _repeat_genome(600, 500, 0.01, 21, sat=300)
    got, _ = _compare(g, tile_records=tile, digit2_bits=8)
    t = got["timings"]
    assert t["crowded_tiles"] + t["fallback_tiles"] > 0


@pytest.mark.parametrize("freq", [10, 64])
def test_crowded_bins_divergent_copies(freq):
    g = _repeat_genome(2000, 300, 0.08, 22)
    _compare(g, freq=freq)
