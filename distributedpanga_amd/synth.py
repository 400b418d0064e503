"""Synthetic genome pairs (binding of the C generator in csrc/pga_synth.c).

A pair is one GDB holding genome A's contigs then genome B's, in the `.bps`
layout (base i of a contig in bits 2*(i%4) of byte boff+i/4, lib/gene_core.c:336-357).
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import host_lib


class _Gdb(C.Structure):
    _fields_ = [("ncontig", C.c_int), ("n_a", C.c_int),
                ("clen", C.POINTER(C.c_int64)), ("boff", C.POINTER(C.c_int64)),
                ("bps", C.POINTER(C.c_uint8)), ("bps_bytes", C.c_int64),
                ("seqtot", C.c_int64), ("maxctg", C.c_int64)]


class PairSpec(C.Structure):
    _fields_ = [("a_bases", C.c_int64), ("a_contigs", C.c_int), ("len_sigma", C.c_double),
                ("len_min", C.c_int64), ("len_max", C.c_int64), ("gc", C.c_double),
                ("rep_families", C.c_int), ("rep_copies", C.c_int), ("rep_fraction", C.c_double),
                ("rep_len_min", C.c_int), ("rep_len_max", C.c_int),
                ("rep_div_min", C.c_double), ("rep_div_max", C.c_double),
                ("tandem_arrays", C.c_int), ("tan_unit_min", C.c_int), ("tan_unit_max", C.c_int),
                ("tan_copies_min", C.c_int), ("tan_copies_max", C.c_int),
                ("sat_bases", C.c_int64), ("sat_unit", C.c_int), ("sat_div", C.c_double),
                ("snp_rate", C.c_double), ("indel_rate", C.c_double), ("indel_mean", C.c_double),
                ("indel_max", C.c_int), ("inversions", C.c_int),
                ("inv_min", C.c_int64), ("inv_max", C.c_int64), ("moves", C.c_int)]


@dataclass
class Gdb:
    """Host view of one GDB: packed bases and contig table (numpy arrays)."""
    bps: np.ndarray      # uint8, packed 2-bit bases
    boff: np.ndarray     # int64 [ncontig+1]
    clen: np.ndarray     # int64 [ncontig]
    n_a: int             # contigs of genome A (the rest are genome B's)

    @property
    def ncontig(self):
        return int(self.clen.shape[0])

    @property
    def seqtot(self):
        return int(self.clen.sum())

    @property
    def maxctg(self):
        return int(self.clen.max()) if self.ncontig else 0

    def contig_codes(self, c):
        """Numeric bases (0..3) of contig c."""
        n = int(self.clen[c])
        b = self.bps[int(self.boff[c]): int(self.boff[c]) + (n + 3) // 4]
        codes = np.stack([(b >> (2 * k)) & 3 for k in range(4)], axis=1).reshape(-1)
        return codes[:n].astype(np.uint8)


def _setup(L):
    if getattr(L, "_synth_ready", False):
        return L
    L.pga_synth_preset.argtypes = [C.c_int, C.c_double, C.POINTER(PairSpec), C.POINTER(C.c_uint64)]
    L.pga_synth_pair.argtypes = [C.POINTER(PairSpec), C.c_uint64, C.POINTER(_Gdb)]
    L.pga_synth_batch_specs.argtypes = [C.c_int, C.c_uint64, C.c_double, C.POINTER(PairSpec),
                                        C.POINTER(C.c_uint64)]
    L.pga_gdb_from_codes.argtypes = [C.POINTER(C.c_uint8), C.POINTER(C.c_int64), C.c_int, C.c_int,
                                     C.POINTER(_Gdb)]
    L.pga_gdb_free.argtypes = [C.POINTER(_Gdb)]
    L.pga_synth_query.argtypes = [C.POINTER(_Gdb), C.c_uint64, C.c_int64, C.c_double, C.c_int64, C.c_int64,
                                  C.POINTER(_Gdb)]
    L._synth_ready = True
    return L


def _adopt(g):
    n = g.ncontig
    clen = np.ctypeslib.as_array(g.clen, shape=(n,)).copy() if n else np.zeros(0, np.int64)
    boff = np.ctypeslib.as_array(g.boff, shape=(n + 1,)).copy()
    nb = int(g.bps_bytes)
    bps = np.ctypeslib.as_array(g.bps, shape=(max(nb, 1),))[:nb].copy()
    return Gdb(bps=bps, boff=boff, clen=clen, n_a=int(g.n_a))


def preset(workload, scale=1.0):
    """(PairSpec, seed) of a named workload (1, 2 or 4), lengths scaled by `scale`."""
    L = _setup(host_lib())
    sp = PairSpec()
    seed = C.c_uint64()
    if L.pga_synth_preset(int(workload), float(scale), C.byref(sp), C.byref(seed)):
        raise ValueError(f"unknown workload {workload}")
    return sp, seed.value


def pair(spec, seed):
    L = _setup(host_lib())
    g = _Gdb()
    if L.pga_synth_pair(C.byref(spec), C.c_uint64(seed), C.byref(g)):
        raise MemoryError("pga_synth_pair failed")
    try:
        return _adopt(g)
    finally:
        L.pga_gdb_free(C.byref(g))


def workload(number, scale=1.0):
    sp, seed = preset(number, scale)
    return pair(sp, seed)


def batch_specs(npairs=64, seed=0x5EED0003, scale=1.0):
    L = _setup(host_lib())
    specs = (PairSpec * npairs)()
    seeds = (C.c_uint64 * npairs)()
    if L.pga_synth_batch_specs(npairs, C.c_uint64(seed), float(scale), specs, seeds):
        raise ValueError("bad batch spec")
    return [specs[i] for i in range(npairs)], [int(seeds[i]) for i in range(npairs)]


def batch(npairs=64, seed=0x5EED0003, scale=1.0):
    specs, seeds = batch_specs(npairs, seed, scale)
    return [pair(s, sd) for s, sd in zip(specs, seeds)]


def _as_c(g):
    """A ctypes view of a Gdb (arrays must stay alive while it is used)."""
    c = _Gdb()
    c.ncontig = g.ncontig
    c.n_a = g.n_a
    c.clen = g.clen.ctypes.data_as(C.POINTER(C.c_int64))
    c.boff = g.boff.ctypes.data_as(C.POINTER(C.c_int64))
    c.bps = g.bps.ctypes.data_as(C.POINTER(C.c_uint8))
    c.bps_bytes = int(g.bps.size)
    c.seqtot = g.seqtot
    c.maxctg = g.maxctg
    return c


def query(target, seed, qbases, div, seg_min=100_000, seg_max=10_000_000):
    """A query genome from random target segments (see pga_synth_query)."""
    L = _setup(host_lib())
    tgt = _Gdb()
    keep = (np.ascontiguousarray(target.bps), np.ascontiguousarray(target.boff, np.int64),
            np.ascontiguousarray(target.clen, np.int64))
    tgt = _as_c(synth_view(target, keep))
    g = _Gdb()
    if L.pga_synth_query(C.byref(tgt), C.c_uint64(seed), int(qbases), float(div), int(seg_min), int(seg_max),
                         C.byref(g)):
        raise ValueError("pga_synth_query failed")
    try:
        out = _adopt(g)
        out.n_a = out.ncontig
        return out
    finally:
        L.pga_gdb_free(C.byref(g))


def synth_view(g, keep):
    return Gdb(bps=keep[0], boff=keep[1], clen=keep[2], n_a=g.n_a)


def concat(a, b):
    """GDB holding a's contigs then b's (e.g. target then query)."""
    bps = np.concatenate([a.bps[: int(a.boff[a.ncontig])], b.bps[: int(b.boff[b.ncontig])]])
    boff = np.concatenate([a.boff[: a.ncontig], b.boff[: b.ncontig] + int(a.boff[a.ncontig]),
                           [int(a.boff[a.ncontig]) + int(b.boff[b.ncontig])]]).astype(np.int64)
    clen = np.concatenate([a.clen, b.clen]).astype(np.int64)
    return Gdb(bps=bps, boff=boff, clen=clen, n_a=a.ncontig)


def take(g, lo, hi):
    """GDB of contigs [lo, hi) of g (e.g. genome A of a pair: take(g, 0, g.n_a))."""
    b0, b1 = int(g.boff[lo]), int(g.boff[hi])
    return Gdb(bps=g.bps[b0:b1].copy(), boff=(g.boff[lo:hi + 1] - b0).astype(np.int64),
               clen=g.clen[lo:hi].copy(), n_a=hi - lo)


def from_codes(contigs, n_a=None):
    """GDB from a list of numeric-base arrays (0..3) or ACGT strings."""
    L = _setup(host_lib())
    arrs = []
    for c in contigs:
        if isinstance(c, str):
            c = np.frombuffer(c.upper().encode(), dtype=np.uint8)
            lut = np.full(256, 255, np.uint8)
            for v, ch in enumerate(b"ACGT"):
                lut[ch] = v
            c = lut[c]
            if (c > 3).any():
                raise ValueError("non-ACGT base")
        arrs.append(np.ascontiguousarray(c, dtype=np.uint8))
    clen = np.array([len(a) for a in arrs], dtype=np.int64)
    codes = np.concatenate(arrs) if arrs else np.zeros(1, np.uint8)
    if codes.size == 0:
        codes = np.zeros(1, np.uint8)
    g = _Gdb()
    if L.pga_gdb_from_codes(codes.ctypes.data_as(C.POINTER(C.c_uint8)),
                            clen.ctypes.data_as(C.POINTER(C.c_int64)), len(arrs),
                            len(arrs) if n_a is None else n_a, C.byref(g)):
        raise MemoryError("pga_gdb_from_codes failed")
    try:
        return _adopt(g)
    finally:
        L.pga_gdb_free(C.byref(g))
