"""GDB genome files: the `<root>.gdb` skeleton (ONE text form) and the
`.<root>.bps` sequence file (layout in include/pga_gdbfile.h; the reference
reads them with Read_GDB, lib/GDB.c:1171-1400, and writes them with
Write_GDB, lib/GDB.c:1519-1607).  ctypes binding of csrc/pga_gdbfile.c.

A read genome has the attributes Context.load_gdb takes (bps, boff, clen,
ncontig), so a GDB on disk goes to the device as it is.
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._lib import host_lib


class _File(C.Structure):
    _fields_ = [("nscaff", C.c_int), ("ncontig", C.c_int), ("seqtot", C.c_int64), ("maxctg", C.c_int64),
                ("freq", C.c_float * 4),
                ("slen", C.POINTER(C.c_int64)), ("fctg", C.POINTER(C.c_int)), ("hoff", C.POINTER(C.c_int64)),
                ("headers", C.c_void_p), ("hdrtot", C.c_int64),
                ("sbeg", C.POINTER(C.c_int64)), ("clen", C.POINTER(C.c_int64)), ("boff", C.POINTER(C.c_int64))]


_ERR = {-1: "I/O error", -2: "binary ONE file with codec-compressed scaffold ids (not decoded)",
        -3: "bad argument or file contents"}


@dataclass
class GdbGenome:
    """A GDB: scaffolds (ids, lengths with gaps, first contigs) and contigs
    (lengths, starts within their scaffold, byte offsets in the .bps image)."""
    ids: list
    slen: np.ndarray          # int64 [nscaff]
    fctg: np.ndarray          # int32 [nscaff + 1]
    sbeg: np.ndarray          # int64 [ncontig]
    clen: np.ndarray          # int64 [ncontig]
    boff: np.ndarray          # int64 [ncontig + 1]
    bps: np.ndarray           # uint8 [boff[-1]]: 2-bit bases, LSB-first
    freq: tuple = field(default=(0.25, 0.25, 0.25, 0.25))

    @property
    def ncontig(self):
        return int(self.clen.size)

    @property
    def nscaff(self):
        return len(self.ids)

    @property
    def seqtot(self):
        return int(self.clen.sum())


def _setup(L):
    if getattr(L, "_gdbfile_ready", False):
        return L
    L.pga_gdbfile_read.argtypes = [C.c_char_p, C.POINTER(_File)]
    L.pga_gdbfile_read_bps.argtypes = [C.c_char_p, C.POINTER(_File), C.POINTER(C.c_uint8)]
    L.pga_gdbfile_write.argtypes = [C.c_char_p, C.POINTER(_File), C.POINTER(C.c_uint8)]
    L.pga_gdbfile_free.argtypes = [C.POINTER(_File)]
    L.pga_gdbfile_free.restype = None
    L._gdbfile_ready = True
    return L


def _check(rc, what):
    if rc:
        raise OSError(f"{what}: {_ERR.get(rc, rc)}")


def read(path, with_bases=True):
    """Read `<root>.gdb` (and, with_bases, `.<root>.bps`)."""
    L = _setup(host_lib())
    f = _File()
    _check(L.pga_gdbfile_read(str(path).encode(), C.byref(f)), f"pga_gdbfile_read {path}")
    try:
        ns, nc = f.nscaff, f.ncontig
        hdr = C.string_at(f.headers, f.hdrtot) if f.hdrtot else b""
        hoff = np.ctypeslib.as_array(f.hoff, (ns,)).copy() if ns else np.zeros(0, np.int64)
        ids = [hdr[o:hdr.index(b"\0", o)].decode() for o in hoff]
        g = GdbGenome(ids=ids,
                      slen=np.ctypeslib.as_array(f.slen, (ns,)).copy() if ns else np.zeros(0, np.int64),
                      fctg=np.ctypeslib.as_array(f.fctg, (ns + 1,)).astype(np.int32),
                      sbeg=np.ctypeslib.as_array(f.sbeg, (nc,)).copy() if nc else np.zeros(0, np.int64),
                      clen=np.ctypeslib.as_array(f.clen, (nc,)).copy() if nc else np.zeros(0, np.int64),
                      boff=np.ctypeslib.as_array(f.boff, (nc + 1,)).copy(),
                      bps=np.zeros(0, np.uint8), freq=tuple(float(x) for x in f.freq))
        if with_bases:
            g.bps = np.zeros(int(g.boff[-1]), np.uint8)
            _check(L.pga_gdbfile_read_bps(str(path).encode(), C.byref(f),
                                          g.bps.ctypes.data_as(C.POINTER(C.c_uint8))),
                   f"pga_gdbfile_read_bps {path}")
        return g
    finally:
        L.pga_gdbfile_free(C.byref(f))


def write(path, g):
    """Write g (a GdbGenome) as `<root>.gdb` and `.<root>.bps`."""
    L = _setup(host_lib())
    hdr = b"".join(s.encode() + b"\0" for s in g.ids)
    hoff = np.zeros(len(g.ids), np.int64)
    o = 0
    for i, s in enumerate(g.ids):
        hoff[i] = o
        o += len(s.encode()) + 1
    slen = np.ascontiguousarray(g.slen, np.int64)
    fctg = np.ascontiguousarray(g.fctg, np.int32)
    sbeg = np.ascontiguousarray(g.sbeg, np.int64)
    clen = np.ascontiguousarray(g.clen, np.int64)
    bps = np.ascontiguousarray(g.bps, np.uint8)
    hbuf = C.create_string_buffer(hdr, max(len(hdr), 1))
    f = _File(nscaff=len(g.ids), ncontig=int(clen.size), seqtot=0, maxctg=0,
              freq=(C.c_float * 4)(*g.freq),
              slen=slen.ctypes.data_as(C.POINTER(C.c_int64)), fctg=fctg.ctypes.data_as(C.POINTER(C.c_int)),
              hoff=hoff.ctypes.data_as(C.POINTER(C.c_int64)), headers=C.cast(hbuf, C.c_void_p).value,
              hdrtot=len(hdr), sbeg=sbeg.ctypes.data_as(C.POINTER(C.c_int64)),
              clen=clen.ctypes.data_as(C.POINTER(C.c_int64)), boff=None)
    _check(L.pga_gdbfile_write(str(path).encode(), C.byref(f),
                               bps.ctypes.data_as(C.POINTER(C.c_uint8)) if bps.size else None),
           f"pga_gdbfile_write {path}")


def from_genome(g, ids=None):
    """A GdbGenome of an in-memory genome (bps, boff, clen; e.g. synth.workload),
    one contig per scaffold."""
    n = int(g.ncontig)
    clen = np.asarray(g.clen, np.int64)[:n]
    return GdbGenome(ids=list(ids) if ids is not None else [f"contig_{i}" for i in range(n)],
                     slen=clen.copy(), fctg=np.arange(n + 1, dtype=np.int32), sbeg=np.zeros(n, np.int64),
                     clen=clen.copy(), boff=np.asarray(g.boff, np.int64)[:n + 1].copy(),
                     bps=np.asarray(g.bps, np.uint8).copy())
