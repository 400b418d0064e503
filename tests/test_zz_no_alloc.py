"""pga_merge_gix_noalloc (the merge entry point that allocates nothing) against
pga_merge + pga_seeds_export on the same index, through the C-ABI.

Device buffers for the caller's side (entries in, streams and bucks out) come
from hipMalloc in the HIP runtime libpga_hip.so itself links, so the pointers
belong to the same runtime as the library's kernels."""
import ctypes as C

import numpy as np
import pytest

from distributedpanga_amd import pga, synth
from distributedpanga_amd._lib import hip_lib

PGA_ELIMIT = -5
PGA_EOVERFLOW = -4
H2D, D2H = 1, 2


class Hip:
    def __init__(self):
        hip_lib()                                  # loads the runtime the library links
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []
        self.live = []

    def alloc(self, n):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(int(n), 16)) == 0
        self.live.append(p)
        return p

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes + 8)
        assert self.L.hipMemcpy(p, arr.ctypes.data_as(C.c_void_p), arr.nbytes, H2D) == 0
        return p

    def get(self, p, n, dtype=np.uint8):
        out = np.zeros(n, dtype=dtype)
        if n:
            assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.live:
            self.L.hipFree(p)
        self.live = []


def _bind():
    L = hip_lib()
    f = L.pga_merge_gix_noalloc
    f.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int,
                  C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                  C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.c_void_p]
    f.restype = C.c_int
    return f


def _existing(g, freq):
    """Index + seeds from the context entry points (pga_merge, pga_seeds_export)."""
    with pga.Context(freq=freq) as ctx:
        ctx.load_gdb(g)
        ctx.build_index()
        ent, _, info = ctx.index_export(0, cum=False)
        ctx.merge()
        N, Cs, nb, cb = ctx.seeds_export(0)
    return info, np.ascontiguousarray(ent).reshape(-1), dict(nseeds=N, cseeds=Cs, nbuck=nb, cbuck=cb)


def _noalloc(f, hip, info, ent, freq):
    e, s, ncontig = info["e"], info["s"], info["ncontig"]
    d_ent = hip.put(ent)
    nlen, clen = C.c_int64(), C.c_int64()
    d_nb, d_cb = hip.alloc(8 * ncontig), hip.alloc(8 * ncontig)
    # sizes first: zero capacity reports the stream lengths, writes nothing
    rc = f(d_ent, ent.nbytes + 8, info["nrec"], info["P"], info["C"], ncontig, freq,
           None, 0, None, 0, d_nb, d_cb, C.byref(nlen), C.byref(clen), None)
    assert rc in (0, PGA_EOVERFLOW), rc
    d_n, d_c = hip.alloc(nlen.value), hip.alloc(clen.value)
    rc = f(d_ent, ent.nbytes + 8, info["nrec"], info["P"], info["C"], ncontig, freq,
           d_n, nlen.value, d_c, clen.value, d_nb, d_cb, C.byref(nlen), C.byref(clen), None)
    assert rc == 0, rc
    assert e * info["nrec"] == ent.nbytes
    return dict(nseeds=hip.get(d_n, nlen.value), cseeds=hip.get(d_c, clen.value),
                nbuck=hip.get(d_nb, ncontig, np.int64), cbuck=hip.get(d_cb, ncontig, np.int64), s=s)


@pytest.mark.gpu
def test_noalloc_merge_equals_context_merge_twice_smaller_second():
    f = _bind()
    hip = Hip()
    try:
        for scale in (0.5, 0.1):                               # the second call is smaller
            g = synth.workload(1, scale)
            print(f"noalloc case: workload 1 at scale {scale}", flush=True)
            info, ent, want = _existing(g, 10)
            assert 0 < info["nrec"] <= (1 << 23)
            got = _noalloc(f, hip, info, ent, 10)
            assert np.array_equal(got["nseeds"], want["nseeds"]), "N stream differs"
            assert np.array_equal(got["cseeds"], want["cseeds"]), "C stream differs"
            assert np.array_equal(got["nbuck"], want["nbuck"]) and np.array_equal(got["cbuck"], want["cbuck"])
            assert len(want["nseeds"]) > 0 and len(want["cseeds"]) > 0
    finally:
        hip.free()


@pytest.mark.gpu
def test_noalloc_merge_limits():
    f = _bind()
    hip = Hip()
    try:
        nlen, clen = C.c_int64(), C.c_int64()
        d = hip.alloc(64)
        # beyond the scratch limit: rejected on the host, before any launch
        rc = f(d, 1 << 40, (1 << 23) + 1, 3, 1, 20, 10, d, 0, d, 0, d, d, C.byref(nlen), C.byref(clen), None)
        assert rc == PGA_ELIMIT
        # an entries buffer without its 8 readable bytes of padding
        rc = f(d, 13 * 4, 4, 3, 1, 20, 10, d, 0, d, 0, d, d, C.byref(nlen), C.byref(clen), None)
        assert rc == -3
    finally:
        hip.free()
