"""GDB genome files (`<root>.gdb` ONE text skeleton + `.<root>.bps`): the
package's reader and writer against each other and against the reference's
own Read_GDB / Write_GDB (lib/GDB.c:1171-1400, 1519-1607), compiled into
oracle/_ref.  A genome read from disk then runs through the device path."""
import numpy as np
import pytest

import oracle
from distributedpanga_amd import gdbfile, pga, synth

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (no reference sources)")


def _toy():
    """Three scaffolds: gaps before, between and after contigs, an id with
    spaces, and a scaffold that is all gap."""
    clen = np.array([1000, 300, 77, 5, 12], np.int64)
    sbeg = np.array([0, 1050, 20, 100, 110], np.int64)
    fctg = np.array([0, 2, 2, 5], np.int32)
    slen = np.array([1350, 40, 150], np.int64)
    boff = np.concatenate([[0], np.cumsum((clen + 3) // 4)]).astype(np.int64)
    rng = np.random.default_rng(5)
    bps = rng.integers(0, 256, int(boff[-1])).astype(np.uint8)
    return gdbfile.GdbGenome(ids=["chr1 test scaffold", "gap_only", "s3"], slen=slen, fctg=fctg, sbeg=sbeg,
                             clen=clen, boff=boff, bps=bps, freq=(0.3, 0.2, 0.2, 0.3))


def _same(a, b, bases=True):
    assert a.ids == b.ids
    for k in ("slen", "fctg", "sbeg", "clen", "boff"):
        assert np.array_equal(np.asarray(getattr(a, k)), np.asarray(getattr(b, k))), k
    assert np.allclose(a.freq, b.freq, atol=1e-6)
    if bases:
        assert np.array_equal(a.bps, b.bps)


def test_round_trip(tmp_path):
    g = _toy()
    gdbfile.write(tmp_path / "toy.gdb", g)
    assert (tmp_path / ".toy.bps").stat().st_size == int(g.boff[-1])
    _same(gdbfile.read(tmp_path / "toy.gdb"), g)


def test_genome_through_files(tmp_path):
    """A synthetic pair written and read back loads the same bases and offsets."""
    g = synth.workload(1, 0.02)
    assert np.array_equal(np.diff(g.boff[:g.ncontig + 1]), (np.asarray(g.clen) + 3) // 4)
    gdbfile.write(tmp_path / "pair.gdb", gdbfile.from_genome(g))
    h = gdbfile.read(tmp_path / "pair.gdb")
    assert h.ncontig == g.ncontig and h.seqtot == g.seqtot
    assert np.array_equal(h.boff, np.asarray(g.boff)[:g.ncontig + 1])
    assert np.array_equal(h.bps, np.asarray(g.bps)[:int(h.boff[-1])])


@needs_ref
def test_reference_reads_ours(tmp_path):
    g = _toy()
    gdbfile.write(tmp_path / "ours.gdb", g)
    r = oracle.ref_gdb_read(tmp_path / "ours.gdb")
    assert r["ids"] == g.ids
    for k in ("slen", "fctg", "sbeg", "clen", "boff"):
        assert np.array_equal(r[k], getattr(g, k)), k
    assert np.allclose(r["freq"], g.freq, atol=1e-6)


@needs_ref
def test_we_read_the_references(tmp_path):
    g = _toy()
    (tmp_path / ".ref.bps").write_bytes(g.bps.tobytes())
    oracle.ref_gdb_write(tmp_path / "ref.gdb", g.ids, g.slen, g.fctg, g.sbeg, g.clen, g.freq)
    _same(gdbfile.read(tmp_path / "ref.gdb"), g)


@needs_ref
def test_we_read_the_references_binary_file(tmp_path):
    """The reference's Write_GDB saving `<root>.1gdb` (ONE binary); the package
    reads the skeleton and the .bps next to it."""
    g = _toy()
    (tmp_path / ".bin.bps").write_bytes(g.bps.tobytes())
    oracle.ref_gdb_write(tmp_path / "bin.1gdb", g.ids, g.slen, g.fctg, g.sbeg, g.clen, g.freq)
    _same(gdbfile.read(tmp_path / "bin.1gdb"), g)


def _big(nscaff, ids_len, seed):
    """Scaffolds with many contigs of lengths and gaps over every integer width
    of the binary encoding (1, 2, 3, 4 bytes)."""
    rng = np.random.default_rng(seed)
    ids, slen, fctg, sbeg, clen = [], [], [0], [], []
    for s in range(nscaff):
        ids.append(f"scaffold_{s}_" + "x" * int(rng.integers(0, ids_len)))
        pos = int(rng.choice([0, 5, 100, 9000, 70000]))
        for c in range(int(rng.integers(1, 6))):
            n = int(rng.choice([1, 50, 63, 64, 8191, 8192, 65535, 65536, 300000, 2 ** 24 + 3]))
            sbeg.append(pos)
            clen.append(n)
            pos += n + int(rng.choice([1, 63, 64, 8192, 100000]))
        slen.append(pos + int(rng.integers(0, 500)))
        fctg.append(len(clen))
    clen = np.array(clen, np.int64)
    boff = np.concatenate([[0], np.cumsum((clen + 3) // 4)]).astype(np.int64)
    return gdbfile.GdbGenome(ids=ids, slen=np.array(slen, np.int64), fctg=np.array(fctg, np.int32),
                             sbeg=np.array(sbeg, np.int64), clen=clen, boff=boff,
                             bps=np.zeros(int(boff[-1]), np.uint8), freq=(0.31, 0.19, 0.21, 0.29))


@needs_ref
def test_binary_file_with_wide_integers(tmp_path):
    g = _big(40, 30, 8)
    (tmp_path / ".big.bps").write_bytes(g.bps.tobytes())
    oracle.ref_gdb_write(tmp_path / "big.1gdb", g.ids, g.slen, g.fctg, g.sbeg, g.clen, g.freq)
    got = gdbfile.read(tmp_path / "big.1gdb", with_bases=False)
    _same(got, g, bases=False)
    # and the reference's own reader agrees on the same binary file
    r = oracle.ref_gdb_read(tmp_path / "big.1gdb", cap=1 << 12)
    assert r["ids"] == got.ids and np.array_equal(r["clen"], got.clen) and np.array_equal(r["sbeg"], got.sbeg)


@needs_ref
def test_binary_file_with_codec_compressed_ids_is_refused(tmp_path):
    """Past 100 KB of scaffold ids ONE compresses them with a trained codec,
    which this reader does not decode: a clear error, not wrong ids."""
    g = _big(1500, 120, 9)
    (tmp_path / ".many.bps").write_bytes(g.bps.tobytes())
    oracle.ref_gdb_write(tmp_path / "many.1gdb", g.ids, g.slen, g.fctg, g.sbeg, g.clen, g.freq)
    with pytest.raises(OSError, match="codec-compressed"):
        gdbfile.read(tmp_path / "many.1gdb", with_bases=False)


def test_binary_fixture_written_by_the_reference():
    """tests/golden/ref_pair.1gdb: a 2 Mbp-pair-shaped genome at 1/50 saved by
    the reference's Write_GDB as ONE binary (tests/golden/make_gdb1.py)."""
    import os
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    g = gdbfile.read(os.path.join(here, "ref_pair.1gdb"))
    z = np.load(os.path.join(here, "ref_pair_1gdb.npz"))
    assert np.array_equal(g.clen, z["clen"]) and np.array_equal(g.boff, z["boff"])
    assert np.array_equal(g.bps, z["bps"])


@pytest.mark.parametrize("text", [
    "1 3 seq 2 1\n.\nS 1 a\nC 10\n",          # not a gdb file
    "1 3 gdb 2 1\n.\nC 10\n",                  # contig before any scaffold
    "1 3 gdb 2 1\n.\nS 9 ab\nC 10\n",          # scaffold id shorter than its length
    "1 3 gdb 2 1\n.\nS 1 a\nC x\n",            # not a number
    "1 3 gdb 2 1\n.\nS 1 a\nQ 3\n",            # not a gdb line type
    "",                                         # empty
])
def test_bad_files(tmp_path, text):
    (tmp_path / "bad.gdb").write_text(text)
    with pytest.raises(OSError, match="bad argument or file contents"):
        gdbfile.read(tmp_path / "bad.gdb", with_bases=False)


def test_missing_file(tmp_path):
    with pytest.raises(OSError, match="I/O error"):
        gdbfile.read(tmp_path / "none.gdb")


@pytest.mark.gpu
def test_gpu_index_and_merge_of_a_gdb_from_disk(tmp_path):
    g = synth.workload(1, 0.05)
    gdbfile.write(tmp_path / "pair.gdb", gdbfile.from_genome(g))
    h = gdbfile.read(tmp_path / "pair.gdb")
    print(f"[gdbfile] ncontig={h.ncontig} bases={h.seqtot}", flush=True)
    got = pga.run_pair(h)
    want = oracle.pipeline(h.bps, h.boff[:-1], h.clen, freq=10)
    assert np.array_equal(got["entries"], want["entries"])
    assert np.array_equal(got["cum"], want["cum"])
    assert np.array_equal(got["nseeds"], want["nseeds"])
    assert np.array_equal(got["cseeds"], want["cseeds"])
    assert np.array_equal(got["nbuck"], want["nbuck"])
    assert np.array_equal(got["cbuck"], want["cbuck"])


@pytest.mark.gpu
def test_gpu_index_and_merge_of_a_binary_gdb_written_by_the_reference():
    """The reference's own binary `.1gdb` (tests/golden/ref_pair.1gdb) read from
    disk, indexed and merged on the device, against the oracle."""
    import os
    h = gdbfile.read(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_pair.1gdb"))
    print(f"[gdbfile] binary: ncontig={h.ncontig} bases={h.seqtot}", flush=True)
    got = pga.run_pair(h)
    want = oracle.pipeline(h.bps, h.boff[:-1], h.clen, freq=10)
    assert want["nrec"] > 0
    for k in ("entries", "cum", "nseeds", "cseeds", "nbuck", "cbuck"):
        assert np.array_equal(got[k], want[k]), k


@pytest.mark.parametrize("old,new", [("S 18 chr1 test scaffold", "S 9223372036854775000 chr1 test scaffold"),
                                     ("S 18 chr1 test scaffold", "S 40 chr1 test scaffold"),
                                     ("S 2 s3", "S 3 s3"),
                                     ("C 1000", "C 2147483648"),
                                     ("C 1000", "C 99999999999999999"),
                                     ("G 50", "G -5")])
def test_malformed_lines_are_refused(tmp_path, old, new):
    """Lengths read from the file are bounded before they are used: a scaffold
    id longer than its line, a contig past INT32_MAX, a negative gap."""
    g = _toy()
    gdbfile.write(tmp_path / "toy.gdb", g)
    text = (tmp_path / "toy.gdb").read_text()
    assert old in text
    (tmp_path / "toy.gdb").write_text(text.replace(old, new, 1))
    with pytest.raises(OSError):
        gdbfile.read(tmp_path / "toy.gdb")
