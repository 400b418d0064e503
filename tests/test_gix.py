"""The on-disk GIX writer/reader (include/pga_gix.h): round trip, the byte
offsets the reference's reader uses (src/fastga_order_core.c:21-35), the
.ktab part layout (notes/GIXmake_dump.c:1632-1650) and prefix-aligned parts.
CPU: written from the oracle's index."""
import numpy as np
import pytest

import oracle
from distributedpanga_amd import gix, synth


@pytest.fixture(scope="module")
def index():
    g = synth.workload(1, 0.05)
    o = oracle.pipeline(g.bps, g.boff[:-1], g.clen)
    return g, o


@pytest.mark.parametrize("nparts", [1, 3])
def test_round_trip(tmp_path, index, nparts):
    g, o = index
    perm = np.arange(g.ncontig, dtype=np.int32)[::-1].copy()
    gix.write(tmp_path, "pair", o["entries"], o["cum"], o["P"], o["C"], o["maxpre"], g.ncontig, perm=perm,
              nparts=nparts)
    r = gix.read(tmp_path, "pair")
    assert (r.P, r.C, r.maxpre, r.nparts, r.freq) == (o["P"], o["C"], o["maxpre"], nparts, 0)
    assert np.array_equal(r.entries, o["entries"])
    assert np.array_equal(r.cum, o["cum"])
    assert np.array_equal(r.perm, perm)


def test_reference_reader_offsets(tmp_path, index):
    """Parse the files exactly where the reference's reader looks."""
    g, o = index
    gix.write(tmp_path, "x", o["entries"], o["cum"], o["P"], o["C"], o["maxpre"], g.ncontig, nparts=2)
    raw = (tmp_path / "x.gix").read_bytes()
    head = np.frombuffer(raw[:16], np.int32)
    assert head.tolist() == [40, 2, 1, 3]
    cum = np.frombuffer(raw[16:16 + 8 * (1 << 24)], np.int64)
    assert np.array_equal(cum, o["cum"])
    at = 16 + 8 * (1 << 24)                              # lseek(4*sizeof(int)+0x1000000*sizeof(int64))
    pbyte, cbyte, nfile = np.frombuffer(raw[at:at + 12], np.int32)
    maxp = np.frombuffer(raw[at + 12:at + 20], np.int64)[0]
    freq, nctg = np.frombuffer(raw[at + 20:at + 28], np.int32)
    perm = np.frombuffer(raw[at + 28:at + 28 + 4 * nctg], np.int32)
    tail = np.frombuffer(raw[at + 28 + 4 * nctg:], np.int64)
    assert (pbyte, cbyte, nfile, maxp, freq, nctg) == (o["P"], o["C"], 2, o["maxpre"], 0, g.ncontig)
    assert np.array_equal(perm, np.arange(g.ncontig)) and tail.tolist() == [-1]
    e = 9 + o["P"] + o["C"]
    parts = []
    for k in (1, 2):
        b = (tmp_path / f".x.ktab.{k}").read_bytes()
        kmer = np.frombuffer(b[:4], np.int32)[0]
        n = np.frombuffer(b[4:12], np.int64)[0]
        assert kmer == 40 and len(b) == 12 + n * e
        parts.append(np.frombuffer(b[12:], np.uint8).reshape(n, e))
    assert np.array_equal(np.concatenate(parts), o["entries"])
    # the split falls on a 24-bit prefix boundary: the first entry of part 2
    # starts a prefix group (its index is a value of the cumulative table)
    assert len(parts[0]) in set(o["cum"].tolist())


def test_rejects_inconsistent_table(tmp_path, index):
    g, o = index
    bad = o["cum"].copy()
    bad[-1] += 1
    with pytest.raises(OSError):
        gix.write(tmp_path, "bad", o["entries"], bad, o["P"], o["C"], o["maxpre"], g.ncontig)
