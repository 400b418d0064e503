"""The on-disk GIX index: `<root>.gix` + `.<root>.ktab.<part>` (layout in
include/pga_gix.h; notes/GIXmake_dump.c:1632-1709).  ctypes binding of the host
C writer/reader in csrc/pga_gix.c."""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import host_lib

NUM_PREFIXES = 1 << 24
_u8p = C.POINTER(C.c_uint8)
_i64p = C.POINTER(C.c_int64)


class _Header(C.Structure):
    _fields_ = [("kmer", C.c_int), ("nparts", C.c_int), ("P", C.c_int), ("C", C.c_int),
                ("maxpre", C.c_int64), ("freq", C.c_int), ("ncontig", C.c_int), ("nents", C.c_int64)]


@dataclass
class GixIndex:
    entries: np.ndarray      # uint8 [nents, 9+P+C]
    cum: np.ndarray          # int64 [2^24]
    perm: np.ndarray         # int32 [ncontig]
    P: int
    C: int
    maxpre: int
    nparts: int
    freq: int


def _setup(L):
    if getattr(L, "_gix_ready", False):
        return L
    L.pga_gix_write.argtypes = [C.c_char_p, C.c_char_p, C.c_int, _u8p, C.c_int64, C.c_int, C.c_int, _i64p,
                                C.c_int64, C.c_int, C.POINTER(C.c_int)]
    L.pga_gix_read_header.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(_Header), _i64p, C.POINTER(C.c_int)]
    L.pga_gix_read_entries.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(_Header), _u8p, C.c_int64]
    L._gix_ready = True
    return L


def write(directory, root, entries, cum, P, Cb, maxpre, ncontig, perm=None, nparts=1):
    L = _setup(host_lib())
    ent = np.ascontiguousarray(entries, np.uint8).reshape(-1)
    nrec = ent.size // (9 + P + Cb)
    cum = np.ascontiguousarray(cum, np.int64)
    pp = None
    if perm is not None:
        perm = np.ascontiguousarray(perm, np.int32)
        pp = perm.ctypes.data_as(C.POINTER(C.c_int))
    rc = L.pga_gix_write(str(directory).encode(), root.encode(), int(nparts),
                         ent.ctypes.data_as(_u8p) if ent.size else None, int(nrec), int(P), int(Cb),
                         cum.ctypes.data_as(_i64p), int(maxpre), int(ncontig), pp)
    if rc:
        raise OSError(f"pga_gix_write failed ({rc})")


def read(directory, root):
    L = _setup(host_lib())
    h = _Header()
    cum = np.zeros(NUM_PREFIXES, np.int64)
    rc = L.pga_gix_read_header(str(directory).encode(), root.encode(), C.byref(h), cum.ctypes.data_as(_i64p), None)
    if rc:
        raise OSError(f"pga_gix_read_header failed ({rc})")
    perm = np.zeros(max(h.ncontig, 1), np.int32)
    rc = L.pga_gix_read_header(str(directory).encode(), root.encode(), C.byref(h), None,
                               perm.ctypes.data_as(C.POINTER(C.c_int)))
    if rc:
        raise OSError(f"pga_gix_read_header failed ({rc})")
    e = 9 + h.P + h.C
    ent = np.zeros(max(h.nents * e, 1), np.uint8)
    rc = L.pga_gix_read_entries(str(directory).encode(), root.encode(), C.byref(h), ent.ctypes.data_as(_u8p),
                                int(ent.size))
    if rc:
        raise OSError(f"pga_gix_read_entries failed ({rc})")
    return GixIndex(entries=ent[: h.nents * e].reshape(h.nents, e), cum=cum, perm=perm[: h.ncontig], P=h.P,
                    C=h.C, maxpre=int(h.maxpre), nparts=h.nparts, freq=h.freq)


def write_from_context(ctx, directory, root, pair=0, perm=None, nparts=1):
    """Export pair's index from a device context and write it as GIX files."""
    ent, cum, info = ctx.index_export(pair, cum=True)
    write(directory, root, ent, cum, info["P"], info["C"], info["maxpre"], info["ncontig"], perm=perm,
          nparts=nparts)
    return info
