"""One GPU's share of a multi-GPU index: the pair's 5-base buckets split into
G contiguous ranges by the reference's splitter (pga_destination_map,
src/dsort.c:120-167), each range indexed and merged on its own
(pga_set_shard), and the shards' outputs, concatenated in bucket order, equal
the unsharded run -- prefix groups never cross a bucket, so the merge is local
(report/sec/harshaan/algorithm-mpi.tex:12-16)."""
import numpy as np
import pytest

import oracle
from distributedpanga_amd import pga, synth

pytestmark = pytest.mark.gpu


def _run(ctx, g):
    ctx.load_gdb(g)
    ctx.build_index()
    ent, cum, info = ctx.index_export(0)
    ctx.merge()
    N, Cs, nb, cb = ctx.seeds_export(0)
    return dict(ent=ent.copy(), cum=cum.copy(), info=info, N=N, C=Cs, nb=nb, cb=cb)


def _lcp(a, b):
    """Leading equal bases of two 10-byte MSB-first k-mers."""
    for k in range(10):
        x = int(a[k]) ^ int(b[k])
        if x:
            l = 4 * k
            while not (x & 0xC0):
                x <<= 2
                l += 1
            return l
    return 40


def _kmer(ent, cum, i):
    """The 80-bit k-mer of entry i: its 24-bit prefix (from the cumulative
    prefix counts) then the entry's 7 suffix bytes."""
    p = int(np.searchsorted(cum, i, side="right"))
    return bytes([p >> 16, (p >> 8) & 255, p & 255]) + bytes(ent[i, :7])


@pytest.mark.parametrize("workload,scale,world", [(1, 1.0, 4), (2, 0.05, 3), (4, 0.005, 8)])
def test_shards_concatenate_to_the_whole_index(workload, scale, world):
    g = synth.workload(workload, scale)
    with pga.Context() as ctx:
        whole = _run(ctx, g)
        want = oracle.pipeline(g.bps, g.boff[:-1], g.clen)
        assert np.array_equal(whole["ent"].reshape(-1), want["entries"].reshape(-1))
        counts = ctx.bucket_counts(0)
        ksplit, select = pga.destination_map(counts, world)
        parts = []
        for s in range(world):
            ctx.set_shard(int(ksplit[s]), int(ksplit[s + 1]))
            part = _run(ctx, g)
            assert part["info"]["nrec"] == counts[ksplit[s]:ksplit[s + 1]].sum()
            assert (part["info"]["P"], part["info"]["C"]) == (whole["info"]["P"], whole["info"]["C"])
            parts.append(part)
    # entries: concatenation, with each shard's first LCP set across the seam
    ents = []
    prev_kmer = None
    for part in parts:
        e = part["ent"].copy()
        if len(e):
            assert e[0, 8] == 0                             # a shard's first entry has no predecessor
            k0 = _kmer(e, part["cum"], 0)
            e[0, 8] = _lcp(prev_kmer, k0) if prev_kmer is not None else 0
            prev_kmer = _kmer(e, part["cum"], len(e) - 1)
        ents.append(e)
    assert np.array_equal(np.concatenate(ents).reshape(-1), whole["ent"].reshape(-1))
    # prefix counts add up; the largest group is the largest of the shards'
    assert np.array_equal(sum(p["cum"] for p in parts), whole["cum"])
    assert max(p["info"]["maxpre"] for p in parts) == whole["info"]["maxpre"]
    # seed streams concatenate in bucket order; per-contig counts add up
    assert np.array_equal(np.concatenate([p["N"] for p in parts]), whole["N"])
    assert np.array_equal(np.concatenate([p["C"] for p in parts]), whole["C"])
    assert np.array_equal(sum(p["nb"] for p in parts), whole["nb"])
    assert np.array_equal(sum(p["cb"] for p in parts), whole["cb"])
