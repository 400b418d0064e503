"""Python binding of the C-ABI in include/libpga_hip.h (ctypes over libpga_hip.so).

The device path has no CPU fallback: every call goes to the HIP kernels, and a
missing library or a HIP failure raises.

Mirrors the reference's stage interface (SakshamDiwan/DistributedPanGA):
    build_index()   extract_to_records + run_stage2 + build_kmer_adapter
                    (src/extract.h:40, src/dsort.h:82, src/kmer_adapter.h:46)
    merge()         self_adaptamer_merge (src/fastga_pipeline.h:18)
"""
import ctypes as C

import numpy as np

from ._lib import hip_lib

NUM_PREFIXES = 1 << 24
_ERR = {-1: "out of memory", -2: "HIP error", -3: "invalid argument", -4: "field overflow",
        -5: "beyond a fixed limit"}


class Opts(C.Structure):
    _fields_ = [("device", C.c_int), ("freq", C.c_int), ("tile_records", C.c_int),
                ("chunk_records", C.c_int), ("merge_tile", C.c_int), ("build_cum", C.c_int),
                ("digit2_bits", C.c_int)]


class IndexMeta(C.Structure):
    _fields_ = [("nrec", C.c_int64), ("P", C.c_int), ("C", C.c_int), ("e", C.c_int), ("s", C.c_int),
                ("maxpre", C.c_int64), ("ncontig", C.c_int), ("seqtot", C.c_int64)]


class SeedsMeta(C.Structure):
    _fields_ = [("nlen", C.c_int64), ("clen", C.c_int64), ("s", C.c_int), ("ncontig", C.c_int)]


class Timings(C.Structure):
    _fields_ = [("extract_ms", C.c_double), ("partition_ms", C.c_double), ("tile_sort_ms", C.c_double),
                ("table_ms", C.c_double), ("merge_ms", C.c_double), ("total_ms", C.c_double),
                ("partition_levels", C.c_int), ("partition_launches", C.c_int), ("tiles", C.c_int64),
                ("scatter_ms", C.c_double), ("scatter_bytes", C.c_int64), ("scatter_launches", C.c_int),
                ("merge_emit_ms", C.c_double), ("fallback_tiles", C.c_int64),
                ("crowded_tiles", C.c_int64), ("graph_fallbacks", C.c_int64)]


_u8p = C.POINTER(C.c_uint8)
_i64p = C.POINTER(C.c_int64)


def _setup(L):
    if getattr(L, "_pga_ready", False):
        return L
    L.pga_default_opts.argtypes = [C.POINTER(Opts)]
    L.pga_default_opts.restype = None
    L.pga_create.argtypes = [C.POINTER(Opts), C.POINTER(C.c_void_p)]
    L.pga_set_opts.argtypes = [C.c_void_p, C.POINTER(Opts)]
    L.pga_destroy.argtypes = [C.c_void_p]
    L.pga_destroy.restype = None
    L.pga_last_error.argtypes = [C.c_void_p]
    L.pga_last_error.restype = C.c_char_p
    L.pga_load_gdb.argtypes = [C.c_void_p, _u8p, C.c_int64, _i64p, _i64p, C.c_int]
    L.pga_load_batch.argtypes = [C.c_void_p, C.c_int, C.POINTER(_u8p), _i64p, C.POINTER(_i64p),
                                 C.POINTER(_i64p), C.POINTER(C.c_int)]
    L.pga_build_index.argtypes = [C.c_void_p]
    L.pga_index_info.argtypes = [C.c_void_p, C.c_int, C.POINTER(IndexMeta)]
    L.pga_index_export.argtypes = [C.c_void_p, C.c_int, _u8p, _i64p]
    L.pga_index_verify.argtypes = [C.c_void_p, C.c_int, _i64p, _i64p, _i64p]
    L.pga_bucket_counts.argtypes = [C.c_void_p, C.c_int, _i64p]
    L.pga_index_export_range.argtypes = [C.c_void_p, C.c_int, C.c_int64, C.c_int64, _u8p]
    L.pga_seeds_locate.argtypes = [C.c_void_p, C.c_int, C.c_int64, _i64p, _i64p]
    L.pga_seeds_export_range.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int64, C.c_int64, _u8p]
    L.pga_index_poke.argtypes = [C.c_void_p, C.c_int, C.c_int64, C.c_int, C.c_int]
    L.pga_destination_map.argtypes = [_i64p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.pga_merge.argtypes = [C.c_void_p]
    L.pga_run_graph.argtypes = [C.c_void_p]
    L.pga_set_shard.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.pga_seeds_info.argtypes = [C.c_void_p, C.c_int, C.POINTER(SeedsMeta)]
    L.pga_seeds_export.argtypes = [C.c_void_p, C.c_int, _u8p, _u8p, _i64p, _i64p]
    L.pga_timings_get.argtypes = [C.c_void_p, C.POINTER(Timings)]
    L.pga_seed_sort.argtypes = [C.c_void_p, C.c_int, C.c_int, _i64p, C.POINTER(C.c_int)]
    L.pga_seed_sort_export.argtypes = [C.c_void_p, _u8p]
    L.pga_sync.argtypes = [C.c_void_p]
    L._pga_ready = True
    return L


class PgaError(RuntimeError):
    pass


class Context:
    """One device context: one HIP stream and a pool of device buffers."""

    def __init__(self, device=0, freq=10, tile_records=4096, chunk_records=8192, merge_tile=1024,
                 build_cum=True, digit2_bits=None):
        self._L = _setup(hip_lib())
        o = Opts()
        self._L.pga_default_opts(C.byref(o))
        o.device, o.freq, o.tile_records = device, freq, tile_records
        o.chunk_records, o.merge_tile, o.build_cum = chunk_records, merge_tile, int(bool(build_cum))
        if digit2_bits is not None:
            o.digit2_bits = digit2_bits
        self.opts = o
        h = C.c_void_p()
        rc = self._L.pga_create(C.byref(o), C.byref(h))
        if rc:
            raise PgaError(f"pga_create failed: {_ERR.get(rc, rc)}")
        self._h = h
        self._keep = None
        self.npairs = 0

    def _check(self, rc, what):
        if rc:
            msg = self._L.pga_last_error(self._h).decode(errors="replace")
            raise PgaError(f"{what}: {_ERR.get(rc, rc)}: {msg}")

    def close(self):
        if getattr(self, "_h", None):
            self._L.pga_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_opts(self, **kw):
        """Change tunables (tile_records, chunk_records, merge_tile, digit2_bits, freq,
        build_cum) in place."""
        o = Opts()
        C.memmove(C.byref(o), C.byref(self.opts), C.sizeof(Opts))
        for k, v in kw.items():
            setattr(o, k, int(v))
        self._check(self._L.pga_set_opts(self._h, C.byref(o)), "pga_set_opts")
        self.opts = o

    def set_shard(self, bucket_lo=0, bucket_hi=1024):
        """Index only records whose 5-base bucket is in [bucket_lo, bucket_hi)."""
        self._check(self._L.pga_set_shard(self._h, int(bucket_lo), int(bucket_hi)), "pga_set_shard")

    # -- input
    def load_batch(self, gdbs):
        n = len(gdbs)
        # an empty .bps image still needs a valid host pointer
        bps = [np.ascontiguousarray(g.bps, np.uint8) if g.bps.size else np.zeros(1, np.uint8) for g in gdbs]
        boff = [np.ascontiguousarray(g.boff[:g.ncontig], np.int64) for g in gdbs]
        clen = [np.ascontiguousarray(g.clen, np.int64) for g in gdbs]
        self._keep = (bps, boff, clen)
        pb = (_u8p * n)(*[b.ctypes.data_as(_u8p) for b in bps])
        nb = (C.c_int64 * n)(*[int(g.bps.size) for g in gdbs])
        po = (_i64p * n)(*[b.ctypes.data_as(_i64p) for b in boff])
        pl = (_i64p * n)(*[b.ctypes.data_as(_i64p) for b in clen])
        nc = (C.c_int * n)(*[int(g.ncontig) for g in gdbs])
        self._check(self._L.pga_load_batch(self._h, n, pb, nb, po, pl, nc), "pga_load_batch")
        self.npairs = n

    def load_gdb(self, gdb):
        self.load_batch([gdb])

    # -- index
    def build_index(self):
        self._check(self._L.pga_build_index(self._h), "pga_build_index")

    def index_info(self, pair=0):
        m = IndexMeta()
        self._check(self._L.pga_index_info(self._h, pair, C.byref(m)), "pga_index_info")
        return {k: getattr(m, k) for k, _ in IndexMeta._fields_}

    def index_export(self, pair=0, cum=True):
        info = self.index_info(pair)
        ent = np.zeros(max(info["nrec"] * info["e"], 1), np.uint8)
        cp = np.zeros(NUM_PREFIXES, np.int64) if cum else None
        self._check(self._L.pga_index_export(self._h, pair, ent.ctypes.data_as(_u8p),
                                             cp.ctypes.data_as(_i64p) if cum else None),
                    "pga_index_export")
        return ent[: info["nrec"] * info["e"]].reshape(info["nrec"], info["e"]), cp, info

    def prefix_table(self, pair=0):
        """The cumulative 2^24 prefix counts alone (no entries copied)."""
        cp = np.zeros(NUM_PREFIXES, np.int64)
        self._check(self._L.pga_index_export(self._h, pair, None, cp.ctypes.data_as(_i64p)), "pga_index_export")
        return cp

    def verify(self, pair=0):
        """(bad, first_bad, lcp_bad) of the pair's index, checked on the device."""
        v = [C.c_int64(), C.c_int64(), C.c_int64()]
        self._check(self._L.pga_index_verify(self._h, pair, *[C.byref(x) for x in v]), "pga_index_verify")
        return tuple(x.value for x in v)

    def index_export_range(self, pair, first, count):
        """Entries [first, first+count) of pair: array [count, e]."""
        e = self.index_info(pair)["e"]
        out = np.zeros(max(count * e, 1), np.uint8)
        self._check(self._L.pga_index_export_range(self._h, pair, int(first), int(count), out.ctypes.data_as(_u8p)),
                    "pga_index_export_range")
        return out[: count * e].reshape(count, e)

    def seeds_locate(self, pair, anchor):
        """(N, C) index of the first seed of anchor entry `anchor`."""
        n, c = C.c_int64(), C.c_int64()
        self._check(self._L.pga_seeds_locate(self._h, pair, int(anchor), C.byref(n), C.byref(c)), "pga_seeds_locate")
        return n.value, c.value

    def seeds_export_range(self, pair, comp, first, count):
        """Seeds [first, first+count) of stream comp (0: N, 1: C) as bytes."""
        s = self.seeds_info(pair)["s"]
        out = np.zeros(max(count * s, 1), np.uint8)
        self._check(self._L.pga_seeds_export_range(self._h, pair, comp, int(first), int(count),
                                                   out.ctypes.data_as(_u8p)), "pga_seeds_export_range")
        return out[: count * s]

    def poke(self, pair, entry, byte, value):
        """Overwrite one byte of a GIX entry on the device (tests of the check)."""
        self._check(self._L.pga_index_poke(self._h, pair, int(entry), int(byte), int(value)), "pga_index_poke")

    def bucket_counts(self, pair=0):
        out = np.zeros(1024, np.int64)
        self._check(self._L.pga_bucket_counts(self._h, pair, out.ctypes.data_as(_i64p)), "pga_bucket_counts")
        return out

    def run_graph(self):
        """build_index + merge as one device-driven pipeline, captured in a HIP
        graph on the first call after a load and replayed afterwards."""
        self._check(self._L.pga_run_graph(self._h), "pga_run_graph")

    # -- merge
    def merge(self):
        self._check(self._L.pga_merge(self._h), "pga_merge")

    def seeds_info(self, pair=0):
        m = SeedsMeta()
        self._check(self._L.pga_seeds_info(self._h, pair, C.byref(m)), "pga_seeds_info")
        return {k: getattr(m, k) for k, _ in SeedsMeta._fields_}

    def seeds_export(self, pair=0):
        info = self.seeds_info(pair)
        N = np.zeros(max(info["nlen"], 1), np.uint8)
        Cs = np.zeros(max(info["clen"], 1), np.uint8)
        nb = np.zeros(max(info["ncontig"], 1), np.int64)
        cb = np.zeros(max(info["ncontig"], 1), np.int64)
        self._check(self._L.pga_seeds_export(self._h, pair, N.ctypes.data_as(_u8p), Cs.ctypes.data_as(_u8p),
                                             nb.ctypes.data_as(_i64p), cb.ctypes.data_as(_i64p)),
                    "pga_seeds_export")
        k = info["ncontig"]
        return N[: info["nlen"]], Cs[: info["clen"]], nb[:k], cb[:k]

    def bucks(self, pair=0):
        """(nbuck, cbuck): seeds per anchor contig, without copying the streams."""
        k = self.seeds_info(pair)["ncontig"]
        nb = np.zeros(max(k, 1), np.int64)
        cb = np.zeros(max(k, 1), np.int64)
        self._check(self._L.pga_seeds_export(self._h, pair, None, None, nb.ctypes.data_as(_i64p),
                                             cb.ctypes.data_as(_i64p)), "pga_seeds_export")
        return nb[:k], cb[:k]

    def seed_sort(self, pair=0, comp=0):
        """Diagonal-sorted seed records of one stream (0: N, 1: C): array [n, swide]."""
        n = C.c_int64()
        sw = C.c_int()
        self._check(self._L.pga_seed_sort(self._h, pair, comp, C.byref(n), C.byref(sw)), "pga_seed_sort")
        out = np.zeros(max(n.value * sw.value, 1), np.uint8)
        self._check(self._L.pga_seed_sort_export(self._h, out.ctypes.data_as(_u8p)), "pga_seed_sort_export")
        return out[: n.value * sw.value].reshape(n.value, sw.value)

    def _seed_sort_only(self, pair=0, comp=0):
        """pga_seed_sort without the copy to the host (timing); returns the record count."""
        n = C.c_int64()
        sw = C.c_int()
        self._check(self._L.pga_seed_sort(self._h, pair, comp, C.byref(n), C.byref(sw)), "pga_seed_sort")
        return n.value

    def timings(self):
        t = Timings()
        self._check(self._L.pga_timings_get(self._h, C.byref(t)), "pga_timings_get")
        return {k: getattr(t, k) for k, _ in Timings._fields_}

    def sync(self):
        self._check(self._L.pga_sync(self._h), "pga_sync")


def destination_map(global_counts, world_size):
    """(ksplit[world_size+1], select[1024]): contiguous bucket ranges per shard."""
    L = _setup(hip_lib())
    g = np.ascontiguousarray(global_counts, np.int64)
    if g.shape != (1024,):
        raise ValueError("global_counts must have 1024 entries")
    ks = (C.c_int * (world_size + 1))()
    sel = (C.c_int * 1024)()
    if L.pga_destination_map(g.ctypes.data_as(_i64p), int(world_size), ks, sel):
        raise PgaError("pga_destination_map: invalid argument")
    return np.array(ks[:], np.int64), np.array(sel[:], np.int64)


def run_pair(gdb, **opts):
    """Index + self-merge one GDB on the device; returns host copies of every output."""
    with Context(**opts) as ctx:
        ctx.load_gdb(gdb)
        ctx.build_index()
        ent, cum, info = ctx.index_export(0, cum=bool(opts.get("build_cum", True)))
        ctx.merge()
        N, Cs, nb, cb = ctx.seeds_export(0)
        return dict(entries=ent, cum=cum, info=info, nseeds=N, cseeds=Cs, nbuck=nb, cbuck=cb,
                    timings=ctx.timings())
