"""The C host driver (distributedpanga_amd/pga_drive): a GDB genome file in,
GIX files and seed streams out, through the C-ABI; its outputs against the
oracle, byte for byte.  The driver runs as a child process."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from distributedpanga_amd import gdbfile, gix, synth

pytestmark = pytest.mark.gpu

DRIVE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "distributedpanga_amd", "pga_drive")


def _as_file(g, path):
    n = g.ncontig
    gen = gdbfile.GdbGenome(ids=[f"contig_{i}" for i in range(n)], slen=g.clen.copy(),
                            fctg=np.arange(n + 1, dtype=np.int32), sbeg=np.zeros(n, np.int64), clen=g.clen.copy(),
                            boff=g.boff.copy(), bps=g.bps[: int(g.boff[n])].copy(), freq=(0.3, 0.2, 0.2, 0.3))
    gdbfile.write(path, gen)


@pytest.mark.parametrize("flags,freq", [([], 10), (["-g", "-p", "3"], 10), (["-f", "4"], 4)])
def test_driver_outputs_equal_the_oracle(tmp_path, flags, freq):
    assert os.path.exists(DRIVE), "pga_drive is not built"
    g = synth.workload(1, 0.3)
    _as_file(g, tmp_path / "pair.gdb")
    out = tmp_path / "out"
    out.mkdir()
    cmd = [DRIVE] + flags + (["-f", str(freq)] if "-f" not in flags else []) + [str(tmp_path / "pair.gdb"), str(out)]
    print("[drive]", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "records" in r.stdout
    want = oracle.pipeline(g.bps, g.boff[:-1], g.clen, freq=freq)
    idx = gix.read(out, "pair")
    assert idx.P == want["P"] and idx.C == want["C"] and idx.maxpre == want["maxpre"]
    assert np.array_equal(idx.entries.reshape(-1), want["entries"].reshape(-1))
    assert np.array_equal(idx.cum, want["cum"])
    assert np.array_equal(np.fromfile(out / "pair.N.seeds", np.uint8), want["nseeds"])
    assert np.array_equal(np.fromfile(out / "pair.C.seeds", np.uint8), want["cseeds"])
    buck = np.fromfile(out / "pair.buck", np.int64)
    assert np.array_equal(buck[: g.ncontig], want["nbuck"]) and np.array_equal(buck[g.ncontig:], want["cbuck"])


def test_driver_usage_and_missing_input(tmp_path):
    assert subprocess.run([DRIVE], stdout=subprocess.PIPE, stderr=subprocess.PIPE).returncode == 1
    r = subprocess.run([DRIVE, str(tmp_path / "none.gdb"), str(tmp_path)], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True)
    assert r.returncode == 1 and "cannot read" in r.stderr
