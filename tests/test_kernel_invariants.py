"""Host checks of index arithmetic the kernels rely on, compiled from the same
header the kernels include (distributedpanga_amd/csrc/pga_device.h)."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "distributedpanga_amd", "csrc")

PROG = r"""
#include <stdio.h>
#include "pga_device.h"
using namespace pga;
int main() {
    int bad = 0, good = 0;
    for (int nt = 1; nt <= 2048; nt++) {
        const bool want = nt > 64;               // any thread past the scanning wave clears
        if (emit_clear_covers(nt) != want) { printf("nt=%d\n", nt); bad++; } else good++;
    }
    // an independent restatement: the digits thread tid clears are tid-64 + k(nt-64)
    for (int nt = 65; nt <= 2048; nt++) {
        int hits[256] = {0};
        for (int tid = 64; tid < nt; tid++)
            for (int d = tid - 64; d < 256; d += nt - 64) hits[d]++;
        for (int d = 0; d < 256; d++) if (hits[d] != 1) { printf("nt=%d d=%d hits=%d\n", nt, d, hits[d]); bad++; break; }
    }
    printf("good=%d bad=%d\n", good, bad);
    return bad != 0;
}
"""


def test_extraction_emit_clears_every_digit_counter_for_every_block_size():
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "t.cpp"), os.path.join(d, "t")
        with open(src, "w") as f:
            f.write(PROG)
        subprocess.run(["g++", "-std=c++17", "-O1", "-I" + CSRC, "-o", exe, src], check=True)
        r = subprocess.run([exe], stdout=subprocess.PIPE, text=True)
        assert r.returncode == 0, r.stdout
        assert "bad=0" in r.stdout
