"""Write tests/golden/ref_pair.1gdb (+ .ref_pair.bps): a genome saved by the
reference's own Write_GDB (lib/GDB.c:1519-1607, compiled into oracle/_ref) in
the binary ONE form, for the binary reader's fixture test and the GPU test that
indexes a genome read from it.  Also ref_pair_1gdb.npz with the contig lengths,
.bps offsets and bases the file must give back.

The genome is the 2 Mbp pair's generator at 1/50 (40 kbp, 20 contigs), one
scaffold per contig with small leading gaps.  Run from the repository root:
    python tests/golden/make_gdb1.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from distributedpanga_amd import synth  # noqa: E402


def main():
    g = synth.workload(1, 0.02)
    n = g.ncontig
    bps = g.bps[: int(g.boff[n])]
    with open(os.path.join(HERE, ".ref_pair.bps"), "wb") as f:
        f.write(bps.tobytes())
    ids = [f"seq{i} pair-member {'A' if i < g.n_a else 'B'}" for i in range(n)]
    gap = np.arange(n, dtype=np.int64) * 7 % 50
    oracle.ref_gdb_write(os.path.join(HERE, "ref_pair.1gdb"), ids, g.clen + gap, np.arange(n + 1, dtype=np.int32),
                         gap, g.clen.astype(np.int64), (0.295, 0.205, 0.205, 0.295))
    np.savez_compressed(os.path.join(HERE, "ref_pair_1gdb.npz"), clen=g.clen.astype(np.int64),
                        boff=g.boff.astype(np.int64), bps=bps, n_a=g.n_a)
    print("wrote ref_pair.1gdb, .ref_pair.bps, ref_pair_1gdb.npz")


if __name__ == "__main__":
    main()
