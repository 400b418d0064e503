"""CPU oracle for the index build + adaptamer merge path.

TEST INFRASTRUCTURE ONLY.  `tests/`, `__graft_entry__.smoke()` and the
CPU-baseline leg of `bench.py` use it as the checker; the product
(`distributedpanga_amd`) never imports it.

Two shared libraries are bound here:
  * ``oracle/liboracle.so`` -- the C restatement in ``oracle/pga_oracle.c``
    (built by ``__graft_entry__.build()``);
  * ``oracle/_ref/libpgaref.so`` -- the reference's own extraction, sort, LCP
    and GIX-adapter sources compiled by ``oracle/ref_build.sh`` (present only
    where the reference sources were available at build time).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(_HERE, "liboracle.so")
REF_SO = os.path.join(_HERE, "_ref", "libpgaref.so")
NPRE = 1 << 24

_u8p = C.POINTER(C.c_uint8)
_i64p = C.POINTER(C.c_int64)


class ExtractStats(C.Structure):
    _fields_ = [("n_sel", C.c_int64), ("n_right", C.c_int64),
                ("n_left", C.c_int64), ("n_tie", C.c_int64)]


class OracleResult(C.Structure):
    _fields_ = [("P", C.c_int), ("C", C.c_int), ("rsize", C.c_int), ("e", C.c_int), ("s", C.c_int),
                ("nrec", C.c_int64), ("records", _u8p), ("entries", _u8p), ("cum", _i64p),
                ("maxpre", C.c_int64), ("nseeds", _u8p), ("nlen", C.c_int64),
                ("cseeds", _u8p), ("clen", C.c_int64), ("nbuck", _i64p), ("cbuck", _i64p),
                ("ncontig", C.c_int), ("xs", ExtractStats),
                ("t_extract", C.c_double), ("t_sort", C.c_double),
                ("t_adapter", C.c_double), ("t_merge", C.c_double)]


class RefResult(C.Structure):
    _fields_ = [("P", C.c_int), ("C", C.c_int), ("rsize", C.c_int), ("e", C.c_int),
                ("nrec", C.c_int64), ("extracted", _u8p), ("sorted", _u8p), ("entries", _u8p),
                ("cum", _i64p), ("maxpre", C.c_int64),
                ("t_extract", C.c_double), ("t_sort", C.c_double), ("t_adapter", C.c_double)]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle library missing: {ORACLE_SO} (run __graft_entry__.build())")
        L = C.CDLL(ORACLE_SO)
        L.oracle_record_sizing.argtypes = [C.c_int64, C.c_int64, C.POINTER(C.c_int),
                                           C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.oracle_pack_forward.argtypes = [_u8p, _u8p]
        L.oracle_pack_rc.argtypes = [_u8p, _u8p]
        L.oracle_comp_table.argtypes = [C.POINTER(C.c_int)]
        L.oracle_tmap.argtypes = [C.c_int]
        L.oracle_mz.argtypes = [_u8p, C.c_int64]
        L.oracle_extract.argtypes = [_u8p, _i64p, _i64p, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.POINTER(_u8p), _i64p, C.POINTER(ExtractStats)]
        L.oracle_sort_records.argtypes = [_u8p, C.c_int64, C.c_int]
        L.oracle_sort_records_qsort.argtypes = [_u8p, C.c_int64, C.c_int, C.c_int, C.c_int]
        L.oracle_lcp_bases.argtypes = [_u8p, _u8p]
        L.oracle_recalc_lcps.argtypes = [_u8p, C.c_int64, C.c_int]
        L.oracle_build_adapter.argtypes = [_u8p, C.c_int64, C.c_int, C.c_int, _u8p, _i64p, _i64p]
        L.oracle_merge.argtypes = [_u8p, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.POINTER(_u8p), _i64p, C.POINTER(_u8p), _i64p, _i64p, _i64p]
        L.oracle_pipeline.argtypes = [_u8p, _i64p, _i64p, C.c_int, C.c_int, C.POINTER(OracleResult)]
        L.oracle_result_free.argtypes = [C.POINTER(OracleResult)]
        L.pga_cpu_bucket_counts.argtypes = [_u8p, _i64p, _i64p, C.c_int, C.c_int, _i64p]
        L.pga_cpu_pipeline.argtypes = [_u8p, _i64p, _i64p, C.c_int, C.c_int, C.c_int, C.POINTER(OracleResult)]
        L.oracle_free.argtypes = [C.c_void_p]
        L.oracle_dbyte.argtypes = [C.c_int64]
        L.oracle_seed_sort.argtypes = [_u8p, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int64, C.c_int, C.c_int,
                                       C.POINTER(_u8p), C.POINTER(C.c_int)]
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"reference build missing: {REF_SO} (run oracle/ref_build.sh)")
        L = C.CDLL(REF_SO)
        L.ref_run.argtypes = [_u8p, C.c_int64, _i64p, _i64p, C.c_int, C.c_int, C.POINTER(RefResult)]
        L.ref_result_free.argtypes = [C.POINTER(RefResult)]
        L.ref_lcp_bases.argtypes = [_u8p, _u8p]
        L.ref_rmsd_sort.argtypes = [_u8p, C.c_int64, C.c_int, C.c_int, _i64p]
        _ip, _fp = C.POINTER(C.c_int), C.POINTER(C.c_float)
        L.ref_gdb_write.argtypes = [C.c_char_p, C.c_int, C.c_char_p, _i64p, _i64p, _ip, C.c_int, _i64p, _i64p, _fp]
        L.ref_gdb_read.argtypes = [C.c_char_p, _ip, _ip, _i64p, _ip, _i64p, _i64p, _i64p, C.c_char_p, C.c_int64,
                                   _fp, C.c_int, C.c_int]
        _ref = L
    return _ref


def _u8(a):
    return a.ctypes.data_as(_u8p)


def _i64(a):
    return a.ctypes.data_as(_i64p)


def _take(ptr, n, dtype=np.uint8):
    if n <= 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def sizing(maxctg, ncontig):
    P, Cb, rs = C.c_int(), C.c_int(), C.c_int()
    lib().oracle_record_sizing(maxctg, ncontig, C.byref(P), C.byref(Cb), C.byref(rs))
    return P.value, Cb.value, rs.value


def extract(bps, boff, clen, mode=1):
    """Records (nrec x rsize uint8) in extraction order, plus the selection counters."""
    bps = np.ascontiguousarray(bps, dtype=np.uint8)
    boff = np.ascontiguousarray(boff, dtype=np.int64)
    clen = np.ascontiguousarray(clen, dtype=np.int64)
    n = len(clen)
    P, Cb, rs = sizing(int(clen.max()) if n else 0, n)
    out = _u8p()
    nrec = C.c_int64()
    st = ExtractStats()
    rc = lib().oracle_extract(_u8(bps), _i64(boff), _i64(clen), n, mode, P, Cb,
                              C.byref(out), C.byref(nrec), C.byref(st))
    if rc:
        raise MemoryError("oracle_extract failed")
    recs = _take(out, nrec.value * rs).reshape(-1, rs)
    lib().oracle_free(C.cast(out, C.c_void_p))
    return recs, dict(sel=st.n_sel, right=st.n_right, left=st.n_left, tie=st.n_tie), (P, Cb)


def sort_records(recs):
    recs = np.ascontiguousarray(recs, dtype=np.uint8).copy()
    lib().oracle_sort_records(_u8(recs), recs.shape[0], recs.shape[1])
    return recs


def sort_records_qsort(recs, P, Cb):
    recs = np.ascontiguousarray(recs, dtype=np.uint8).copy()
    lib().oracle_sort_records_qsort(_u8(recs), recs.shape[0], recs.shape[1], P, Cb)
    return recs


def recalc_lcps(recs):
    recs = np.ascontiguousarray(recs, dtype=np.uint8).copy()
    lib().oracle_recalc_lcps(_u8(recs), recs.shape[0], recs.shape[1])
    return recs


def build_adapter(recs, P, Cb):
    recs = np.ascontiguousarray(recs, dtype=np.uint8)
    n = recs.shape[0]
    e = 9 + P + Cb
    ent = np.zeros(max(n * e, 1), dtype=np.uint8)
    cum = np.zeros(NPRE, dtype=np.int64)
    mp = C.c_int64()
    lib().oracle_build_adapter(_u8(recs), n, P, Cb, _u8(ent), _i64(cum), C.byref(mp))
    return ent[: n * e].reshape(n, e), cum, mp.value


def merge(entries, P, Cb, ncontig, freq=10, mode=0):
    entries = np.ascontiguousarray(entries, dtype=np.uint8)
    n = entries.shape[0]
    nb = np.zeros(max(ncontig, 1), dtype=np.int64)
    cb = np.zeros(max(ncontig, 1), dtype=np.int64)
    ns, cs = _u8p(), _u8p()
    nl, cl = C.c_int64(), C.c_int64()
    rc = lib().oracle_merge(_u8(entries), n, P, Cb, ncontig, freq, mode,
                            C.byref(ns), C.byref(nl), C.byref(cs), C.byref(cl), _i64(nb), _i64(cb))
    if rc:
        raise MemoryError("oracle_merge failed")
    N = _take(ns, nl.value)
    Cs = _take(cs, cl.value)
    lib().oracle_free(C.cast(ns, C.c_void_p))
    lib().oracle_free(C.cast(cs, C.c_void_p))
    return N, Cs, nb[:ncontig], cb[:ncontig]


def pipeline(bps, boff, clen, freq=10, threads=0):
    """The whole path on the CPU.  Returns a dict of numpy arrays and timings.
    threads=0: the single-threaded oracle (oracle_pipeline); threads >= 1: the
    CPU baseline of oracle/pga_cpu.c on that many threads (no sorted records)."""
    bps = np.ascontiguousarray(bps, dtype=np.uint8)
    boff = np.ascontiguousarray(boff, dtype=np.int64)
    clen = np.ascontiguousarray(clen, dtype=np.int64)
    r = OracleResult()
    if threads:
        rc = lib().pga_cpu_pipeline(_u8(bps), _i64(boff), _i64(clen), len(clen), freq, int(threads), C.byref(r))
    else:
        rc = lib().oracle_pipeline(_u8(bps), _i64(boff), _i64(clen), len(clen), freq, C.byref(r))
    if rc:
        lib().oracle_result_free(C.byref(r))
        raise MemoryError("oracle pipeline failed")
    out = dict(P=r.P, C=r.C, rsize=r.rsize, e=r.e, s=r.s, nrec=r.nrec,
               records=_take(r.records, r.nrec * r.rsize).reshape(-1, r.rsize),
               entries=_take(r.entries, r.nrec * r.e).reshape(-1, r.e),
               cum=_take(r.cum, NPRE, np.int64), maxpre=r.maxpre,
               nseeds=_take(r.nseeds, r.nlen), cseeds=_take(r.cseeds, r.clen),
               nbuck=_take(r.nbuck, r.ncontig, np.int64), cbuck=_take(r.cbuck, r.ncontig, np.int64),
               stats=dict(sel=r.xs.n_sel),
               t_extract=r.t_extract, t_sort=r.t_sort, t_adapter=r.t_adapter, t_merge=r.t_merge)
    lib().oracle_result_free(C.byref(r))
    return out


def bucket_counts(bps, boff, clen, threads=1):
    """Records per 5-base bucket of a GDB (record_bucket, src/dsort.h:30-33),
    counted by the rolling scan without storing the records."""
    bps = np.ascontiguousarray(bps, dtype=np.uint8)
    boff = np.ascontiguousarray(boff, dtype=np.int64)
    clen = np.ascontiguousarray(clen, dtype=np.int64)
    out = np.zeros(1024, np.int64)
    if lib().pga_cpu_bucket_counts(_u8(bps), _i64(boff), _i64(clen), len(clen), int(threads), _i64(out)):
        raise MemoryError("pga_cpu_bucket_counts failed")
    return out


def ref_run(bps, boff, clen, do_index=True):
    """The reference's own extraction (and sort + LCP + adapter) on the same GDB."""
    bps = np.ascontiguousarray(bps, dtype=np.uint8)
    boff = np.ascontiguousarray(boff, dtype=np.int64)
    clen = np.ascontiguousarray(clen, dtype=np.int64)
    r = RefResult()
    rc = ref().ref_run(_u8(bps), len(bps), _i64(boff), _i64(clen), len(clen), int(do_index), C.byref(r))
    if rc:
        raise MemoryError("ref_run failed")
    rs = r.rsize
    out = dict(P=r.P, C=r.C, rsize=rs, e=r.e, nrec=r.nrec,
               extracted=_take(r.extracted, r.nrec * rs).reshape(-1, rs),
               t_extract=r.t_extract, t_sort=r.t_sort, t_adapter=r.t_adapter)
    if do_index:
        out.update(sorted=_take(r.sorted, r.nrec * rs).reshape(-1, rs),
                   entries=_take(r.entries, r.nrec * r.e).reshape(-1, r.e),
                   cum=_take(r.cum, NPRE, np.int64), maxpre=r.maxpre)
    ref().ref_result_free(C.byref(r))
    return out


def dbyte(maxctg):
    """DBYTE of a GDB whose longest contig is maxctg (src/pga-mpi-merge.c:415-425)."""
    return lib().oracle_dbyte(int(maxctg))


def seed_sort(seeds, P, Cb, ncontig, maxctg, comp, sort=True):
    """Diagonal records of one seed stream, panels in contig order, each sorted
    on the whole record from its last byte.  Returns (records [n, swide], swide)."""
    L = lib()
    s = 1 + 2 * (P + Cb)
    seeds = np.ascontiguousarray(seeds, np.uint8)
    n = seeds.size // s
    out = _u8p()
    sw = C.c_int()
    buf = seeds if seeds.size else np.zeros(1, np.uint8)
    if L.oracle_seed_sort(_u8(buf), n, P, Cb, ncontig, int(maxctg), int(comp), int(bool(sort)), C.byref(out),
                          C.byref(sw)):
        raise MemoryError("oracle_seed_sort failed")
    recs = _take(out, n * sw.value).reshape(n, sw.value)
    L.oracle_free(out)
    return recs, sw.value


def ref_rmsd_sort(records, panel_sizes):
    """The reference's own rmsd_sort (lib/RSDsort.c) over reimported records,
    one panel per contig (panel_sizes in records)."""
    R = ref()
    recs = np.ascontiguousarray(records, np.uint8).copy()
    n, sw = recs.shape
    panel = np.ascontiguousarray(np.asarray(panel_sizes, np.int64) * sw)
    if n:
        R.ref_rmsd_sort(_u8(recs), n, sw, len(panel), _i64(panel))
    return recs


def ref_gdb_write(path, ids, slen, fctg, sbeg, clen, freq=(0.25, 0.25, 0.25, 0.25)):
    """The reference's Write_GDB (lib/GDB.c:1519-1607): path ends in ".gdb"
    (ONE text) or ".1gdb" (ONE binary); dir/.<root>.bps must exist."""
    R = ref()
    hdr = b"".join(s.encode() + b"\0" for s in ids)
    hoff = np.zeros(max(len(ids), 1), np.int64)
    o = 0
    for i, s in enumerate(ids):
        hoff[i] = o
        o += len(s.encode()) + 1
    slen = np.ascontiguousarray(slen, np.int64)
    fctg = np.ascontiguousarray(fctg, np.int32)
    sbeg = np.ascontiguousarray(sbeg, np.int64)
    clen = np.ascontiguousarray(clen, np.int64)
    fr = np.ascontiguousarray(freq, np.float32)
    rc = R.ref_gdb_write(str(path).encode(), len(ids), hdr, _i64(hoff), _i64(slen),
                         fctg.ctypes.data_as(C.POINTER(C.c_int)), int(clen.size), _i64(sbeg), _i64(clen),
                         fr.ctypes.data_as(C.POINTER(C.c_float)))
    if rc:
        raise OSError(f"Write_GDB failed ({rc})")


def ref_gdb_read(path, cap=1 << 16):
    """The reference's Read_GDB (lib/GDB.c:1171-1400): scaffold ids, lengths and
    first contigs, contig starts, lengths and .bps offsets, base frequencies."""
    R = ref()
    ns, nc = C.c_int(), C.c_int()
    slen = np.zeros(cap, np.int64)
    fctg = np.zeros(cap + 1, np.int32)
    sbeg = np.zeros(cap, np.int64)
    clen = np.zeros(cap, np.int64)
    boff = np.zeros(cap + 1, np.int64)
    hdr = C.create_string_buffer(cap * 64)
    fr = np.zeros(4, np.float32)
    rc = R.ref_gdb_read(str(path).encode(), C.byref(ns), C.byref(nc), _i64(slen),
                        fctg.ctypes.data_as(C.POINTER(C.c_int)), _i64(sbeg), _i64(clen), _i64(boff), hdr,
                        cap * 64, fr.ctypes.data_as(C.POINTER(C.c_float)), cap, cap)
    if rc:
        raise OSError(f"Read_GDB failed ({rc})")
    ns, nc = ns.value, nc.value
    raw = hdr.raw
    ids, o = [], 0
    for _ in range(ns):
        e = raw.index(b"\0", o)
        ids.append(raw[o:e].decode())
        o = e + 1
    return dict(ids=ids, slen=slen[:ns].copy(), fctg=fctg[:ns + 1].copy(), sbeg=sbeg[:nc].copy(),
                clen=clen[:nc].copy(), boff=boff[:nc + 1].copy(), freq=tuple(float(x) for x in fr))
