"""CPU: the oracle's seed reimport + per-panel sort (the diagonal-sorted seed
array pair_sort_search builds, src/fastga_pipeline.c:1391-1450, 1711-1822)
pinned by closed forms on hand-built seeds and by the reference's own
rmsd_sort (lib/RSDsort.c, compiled into oracle/_ref) on the same records."""
import numpy as np
import pytest

import oracle
from distributedpanga_amd import synth


def seed(plen, ipos, icont, jpos, jcont, jstrand, P, Cb):
    b = [plen]
    b += [(ipos >> (8 * k)) & 255 for k in range(P)]
    b += [(icont >> (8 * k)) & 255 for k in range(Cb)]
    b += [(jpos >> (8 * k)) & 255 for k in range(P)]
    jc = jcont | (jstrand << (8 * Cb - 1))
    b += [(jc >> (8 * k)) & 255 for k in range(Cb)]
    return b


def test_dbyte():
    # DBYTE: smallest d with 256^d >= 2*maxctg (src/pga-mpi-merge.c:415-425)
    assert oracle.dbyte(100) == 1 and oracle.dbyte(128) == 1 and oracle.dbyte(129) == 2
    assert oracle.dbyte(32768) == 2 and oracle.dbyte(32769) == 3 and oracle.dbyte(250_000_000) == 4


@pytest.mark.parametrize("comp", [0, 1])
def test_reimport_closed_form(comp):
    P, Cb, maxctg = 2, 1, 1000                 # DBYTE 2, swide 2*2+1+2 = 7
    raw = np.array(seed(33, 700, 1, 100, 2, 1, P, Cb) + seed(20, 5, 0, 900, 1, 0, P, Cb), np.uint8)
    recs, sw = oracle.seed_sort(raw, P, Cb, 3, maxctg, comp, sort=False)
    assert sw == 7
    # panel of contig 0 first (the second seed), then contig 1
    def rec(lcp, i, j, jc):
        if comp:
            diag, anti = 2 * maxctg - (i + j), maxctg - (i - j)
        else:
            diag, anti = maxctg + (i - j), i + j
        band = diag >> 6
        return [lcp, diag & 63, anti & 255, anti >> 8, band & 255, band >> 8, jc]
    assert recs[0].tolist() == rec(20, 5, 900, 1)
    assert recs[1].tolist() == rec(33, 700, 100, 2)      # jcont's strand bit dropped


def test_sort_order_is_last_byte_first():
    P, Cb = 1, 1
    raw = []
    for (l, i, j, jc) in [(12, 10, 3, 1), (12, 10, 3, 0), (13, 10, 3, 0), (12, 11, 4, 0), (40, 60, 2, 1)]:
        raw += seed(l, i, 0, j, jc, 0, P, Cb)
    recs, sw = oracle.seed_sort(np.array(raw, np.uint8), P, Cb, 1, 100, 0)
    keys = [int.from_bytes(bytes(r), "little") for r in recs]
    assert keys == sorted(keys)


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.parametrize("scale,comp", [(0.05, 0), (0.05, 1), (0.2, 0), (0.2, 1)])
def test_panel_sort_matches_reference_rmsd_sort(scale, comp):
    g = synth.workload(1, scale)
    o = oracle.pipeline(g.bps, g.boff[:-1], g.clen)
    stream = o["nseeds"] if comp == 0 else o["cseeds"]
    buck = o["nbuck"] if comp == 0 else o["cbuck"]
    unsorted, sw = oracle.seed_sort(stream, o["P"], o["C"], g.ncontig, g.maxctg, comp, sort=False)
    mine, _ = oracle.seed_sort(stream, o["P"], o["C"], g.ncontig, g.maxctg, comp, sort=True)
    theirs = oracle.ref_rmsd_sort(unsorted, buck)
    assert len(mine) == buck.sum() > 100
    assert np.array_equal(mine, theirs)
