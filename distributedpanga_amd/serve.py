"""Serving: a resident target index answering query genomes as they arrive.

`Server` binds the pga_server_* entry points of include/libpga_hip.h.  A
query's answer is the part of the reference's self-merge of (target contigs,
then the query's contigs) -- src/pga-mpi-merge.c:509, src/fastga_pipeline.c:
245-412 -- in which the anchor or the partner lies on a query contig.

`serve()` is the request loop (continuous batching): whenever one of the
server's slots is free, every request that has arrived is admitted into that
slot's next batch as long as the projected device footprint of everything in
flight stays within the budget; the batch starts on the device at once, beside
the batches still running on the other slots.  Latency is measured from a
request's arrival to the end of the batch that answered it.
"""
import ctypes as C
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from ._lib import hip_lib
from .pga import _ERR, IndexMeta, Opts, PgaError, SeedsMeta, _i64p, _u8p


def _setup(L):
    if getattr(L, "_srv_ready", False):
        return L
    L.pga_server_create.argtypes = [C.POINTER(Opts), C.c_int, C.POINTER(C.c_void_p)]
    L.pga_server_destroy.argtypes = [C.c_void_p]
    L.pga_server_destroy.restype = None
    L.pga_server_slots.argtypes = [C.c_void_p]
    L.pga_server_error.argtypes = [C.c_void_p, C.c_int]
    L.pga_server_error.restype = C.c_char_p
    L.pga_server_set_target_gix.argtypes = [C.c_void_p, _u8p, C.c_int64, C.c_int, C.c_int, _i64p, C.c_int64, _i64p,
                                            C.c_int]
    L.pga_server_set_target.argtypes = [C.c_void_p, _u8p, C.c_int64, _i64p, _i64p, C.c_int]
    L.pga_server_target_info.argtypes = [C.c_void_p, C.POINTER(IndexMeta)]
    L.pga_server_footprint.argtypes = [C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]
    L.pga_server_run.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(_u8p), _i64p, C.POINTER(_i64p),
                                 C.POINTER(_i64p), C.POINTER(C.c_int)]
    L.pga_server_seeds_info.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(SeedsMeta)]
    L.pga_server_seeds_export.argtypes = [C.c_void_p, C.c_int, C.c_int, _u8p, _u8p, _i64p, _i64p]
    L.pga_server_last_ms.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double)]
    L._srv_ready = True
    return L


def _gdb_arrays(gdbs):
    bps = [np.ascontiguousarray(g.bps, np.uint8) for g in gdbs]
    boff = [np.ascontiguousarray(g.boff[: g.ncontig], np.int64) for g in gdbs]
    clen = [np.ascontiguousarray(g.clen, np.int64) for g in gdbs]
    for b in bps:
        if b.size == 0:
            b.resize(1, refcheck=False)
    return bps, boff, clen


class Server:
    """A target GDB indexed once on the device, answering query GDBs on `slots`
    independent slots (one batch in flight per slot; slots run concurrently
    when driven from different threads -- ctypes releases the GIL)."""

    def __init__(self, target, device=0, freq=10, slots=1, tile_records=4096, chunk_records=4096,
                 merge_tile=256, digit2_bits=10):
        self._L = _setup(hip_lib())
        o = Opts()
        self._L.pga_default_opts(C.byref(o))
        o.device, o.freq, o.tile_records = device, freq, tile_records
        o.chunk_records, o.merge_tile, o.build_cum = chunk_records, merge_tile, 1
        o.digit2_bits = digit2_bits
        h = C.c_void_p()
        rc = self._L.pga_server_create(C.byref(o), int(slots), C.byref(h))
        if rc:
            raise PgaError(f"pga_server_create failed: {_ERR.get(rc, rc)}")
        self._h = h
        self.freq = freq
        self.slots = int(slots)
        self.n_target = 0
        if target is not None:
            self.set_target(target)

    def set_target(self, target):
        """Index `target` (a GDB) on the device; it stays resident."""
        self.n_target = target.ncontig
        bps, boff, clen = _gdb_arrays([target])
        self._check(self._L.pga_server_set_target(self._h, bps[0].ctypes.data_as(_u8p), int(target.bps.size),
                                                  boff[0].ctypes.data_as(_i64p), clen[0].ctypes.data_as(_i64p),
                                                  target.ncontig), "pga_server_set_target", -1)

    def set_target_gix(self, index, clen):
        """The target from a persisted GIX index (gix.read's GixIndex: entries,
        prefix counts, P, C, maxpre) and its contig lengths, without its
        sequence: a restart of the serving loop without re-indexing."""
        ent = np.ascontiguousarray(index.entries, np.uint8).reshape(-1)
        if ent.size == 0:
            ent = np.zeros(1, np.uint8)
        cum = np.ascontiguousarray(index.cum, np.int64)
        cl = np.ascontiguousarray(clen, np.int64)
        nrec = int(index.entries.shape[0])
        self.n_target = int(cl.size)
        self._check(self._L.pga_server_set_target_gix(self._h, ent.ctypes.data_as(_u8p), nrec, int(index.P),
                                                      int(index.C), cum.ctypes.data_as(_i64p), int(index.maxpre),
                                                      cl.ctypes.data_as(_i64p), int(cl.size)),
                    "pga_server_set_target_gix", -1)

    @classmethod
    def from_gix(cls, directory, root, clen, **kw):
        """A server whose target is the GIX index <directory>/<root>.gix (+ .ktab parts)."""
        from . import gix
        srv = cls(None, **kw)
        try:
            srv.set_target_gix(gix.read(directory, root), clen)
        except Exception:
            srv.close()
            raise
        return srv

    def _check(self, rc, what, slot):
        if rc:
            msg = self._L.pga_server_error(self._h, slot).decode(errors="replace")
            raise PgaError(f"{what}: {_ERR.get(rc, rc)}: {msg}")

    def close(self):
        if getattr(self, "_h", None):
            self._L.pga_server_destroy(self._h)
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

    def target_info(self):
        m = IndexMeta()
        self._check(self._L.pga_server_target_info(self._h, C.byref(m)), "pga_server_target_info", -1)
        return {k: getattr(m, k) for k, _ in IndexMeta._fields_}

    def footprint(self, query_bases):
        b = C.c_int64()
        self._check(self._L.pga_server_footprint(self._h, int(query_bases), C.byref(b)), "pga_server_footprint", -1)
        return b.value

    def run(self, queries, slot=0):
        """Answer a batch of query GDBs on `slot` (answers stay on the device until exported)."""
        n = len(queries)
        bps, boff, clen = _gdb_arrays(queries)
        pb = (_u8p * n)(*[b.ctypes.data_as(_u8p) for b in bps])
        nb = (C.c_int64 * n)(*[int(g.bps.size) for g in queries])
        po = (_i64p * n)(*[b.ctypes.data_as(_i64p) for b in boff])
        pl = (_i64p * n)(*[b.ctypes.data_as(_i64p) for b in clen])
        nc = (C.c_int * n)(*[int(g.ncontig) for g in queries])
        self._check(self._L.pga_server_run(self._h, slot, n, pb, nb, po, pl, nc), "pga_server_run", slot)

    def last_ms(self, slot=0):
        v = C.c_double()
        self._check(self._L.pga_server_last_ms(self._h, slot, C.byref(v)), "pga_server_last_ms", slot)
        return v.value

    def seeds_info(self, q, slot=0):
        m = SeedsMeta()
        self._check(self._L.pga_server_seeds_info(self._h, slot, q, C.byref(m)), "pga_server_seeds_info", slot)
        return {k: getattr(m, k) for k, _ in SeedsMeta._fields_}

    def seeds_export(self, q, slot=0):
        info = self.seeds_info(q, slot)
        N = np.zeros(max(info["nlen"], 1), np.uint8)
        Cs = np.zeros(max(info["clen"], 1), np.uint8)
        nb = np.zeros(max(info["ncontig"], 1), np.int64)
        cb = np.zeros(max(info["ncontig"], 1), np.int64)
        self._check(self._L.pga_server_seeds_export(self._h, slot, q, N.ctypes.data_as(_u8p),
                                                    Cs.ctypes.data_as(_u8p), nb.ctypes.data_as(_i64p),
                                                    cb.ctypes.data_as(_i64p)), "pga_server_seeds_export", slot)
        k = info["ncontig"]
        return dict(nseeds=N[: info["nlen"]], cseeds=Cs[: info["clen"]], nbuck=nb[:k], cbuck=cb[:k],
                    s=info["s"])

    def answer(self, queries, slot=0):
        """run() then export every answer."""
        self.run(queries, slot)
        return [self.seeds_export(q, slot) for q in range(len(queries))]


@dataclass
class Request:
    rid: int
    arrival: float           # seconds from the start of the loop
    query: object            # synth.Gdb
    bases: int = 0
    start: float = 0.0       # when its batch was dispatched
    done: float = 0.0
    seeds: int = 0
    slot: int = -1


@dataclass
class ServeStats:
    requests: int = 0
    bases: int = 0
    seeds: int = 0
    wall_s: float = 0.0
    batches: int = 0
    max_batch: int = 0
    max_inflight: int = 0            # batches on the device at once
    max_inflight_bytes: int = 0      # summed footprint admitted at once
    latencies_ms: list = field(default_factory=list)
    device_ms: float = 0.0

    def summary(self):
        lat = np.asarray(self.latencies_ms, dtype=np.float64)
        return dict(requests=self.requests, batches=self.batches, max_batch=self.max_batch,
                    max_inflight=self.max_inflight, max_inflight_bytes=self.max_inflight_bytes,
                    bases_per_s=self.bases / self.wall_s if self.wall_s > 0 else 0.0,
                    seeds=self.seeds, wall_s=self.wall_s, device_ms=self.device_ms,
                    p50_ms=float(np.percentile(lat, 50)) if lat.size else 0.0,
                    p99_ms=float(np.percentile(lat, 99)) if lat.size else 0.0,
                    mean_ms=float(lat.mean()) if lat.size else 0.0)


def poisson_arrivals(n, rate_per_s, seed=1):
    """Arrival times (seconds) of n requests of a Poisson process."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_per_s, size=n))


def serve(server, requests, budget_bytes, max_batch=64, export=False, clock=time.perf_counter):
    """Answer `requests` (sorted by arrival) as they arrive; returns (stats, answers).

    Continuous batching over the server's slots: whenever a slot is free, every
    request that has arrived is admitted into that slot's next batch while the
    footprint of everything in flight stays within budget_bytes (a request
    larger than the budget is admitted alone once nothing else is in flight).
    Each slot is driven by its own thread, so a batch is admitted and put on the
    device while earlier batches are still running there.  Answers are exported
    to the host when `export` is set.
    """
    nslots = getattr(server, "slots", 1)
    stats = ServeStats()
    answers = {}
    lock = threading.Condition()
    free = list(range(nslots))
    inflight = {"n": 0, "bytes": 0}
    errors = []
    t0 = clock()

    def work(slot, batch, fp):
        try:
            server.run([r.query for r in batch], slot)
            ms = server.last_ms(slot)
            res = []
            for i, r in enumerate(batch):
                info = server.seeds_info(i, slot)
                r.seeds = (info["nlen"] + info["clen"]) // info["s"]
                if export:
                    res.append((r.rid, server.seeds_export(i, slot)))
            t = clock() - t0
        except Exception as e:       # surfaced by the dispatcher
            with lock:
                errors.append(e)
                free.append(slot)
                inflight["n"] -= 1
                inflight["bytes"] -= fp
                lock.notify_all()
            return
        with lock:
            for r in batch:
                r.done = t
                stats.latencies_ms.append(1000.0 * (t - r.arrival))
                stats.bases += r.bases or r.query.seqtot
                stats.seeds += r.seeds
            answers.update(res)
            stats.device_ms += ms
            free.append(slot)
            inflight["n"] -= 1
            inflight["bytes"] -= fp
            lock.notify_all()

    threads = []
    k = 0
    n = len(requests)
    with lock:
        while k < n and not errors:
            now = clock() - t0
            if requests[k].arrival > now:
                lock.wait(timeout=requests[k].arrival - now)
                continue
            if not free:
                lock.wait()
                continue
            batch, used = [], 0
            while k < n and len(batch) < max_batch and requests[k].arrival <= clock() - t0:
                r = requests[k]
                fp = server.footprint(r.bases or r.query.seqtot)
                if inflight["bytes"] + used + fp > budget_bytes and (batch or inflight["n"] > 0):
                    break
                batch.append(r)
                used += fp
                k += 1
            if not batch:                 # the budget is full: wait for a batch to finish
                lock.wait()
                continue
            slot = free.pop()
            inflight["n"] += 1
            inflight["bytes"] += used
            stats.max_inflight = max(stats.max_inflight, inflight["n"])
            stats.max_inflight_bytes = max(stats.max_inflight_bytes, inflight["bytes"])
            stats.batches += 1
            stats.requests += len(batch)
            stats.max_batch = max(stats.max_batch, len(batch))
            tstart = clock() - t0
            for r in batch:
                r.start, r.slot = tstart, slot
            th = threading.Thread(target=work, args=(slot, batch, used), daemon=True)
            th.start()
            threads.append(th)
    for th in threads:
        th.join()
    if errors:
        raise errors[0]
    stats.wall_s = clock() - t0
    return stats, answers


def filter_query_seeds(nseeds, cseeds, s, P, Cb, n_target, ncontig):
    """Host-side reduction of a full self-merge's streams to a query's answer:
    keep seeds whose anchor or partner contig is >= n_target, recount bucks.
    (Used to compare against the self-merge of the combined GDB.)"""
    out = []
    for stream in (nseeds, cseeds):
        a = stream.reshape(-1, s)
        pc = P + Cb
        w = 1 << np.arange(0, 8 * Cb, 8, dtype=np.uint64)
        ci = (a[:, 1 + P: 1 + pc].astype(np.uint64) * w).sum(axis=1) & np.uint64((1 << (8 * Cb - 1)) - 1)
        cp = (a[:, 1 + pc + P: 1 + 2 * pc].astype(np.uint64) * w).sum(axis=1) & np.uint64((1 << (8 * Cb - 1)) - 1)
        keep = (ci >= n_target) | (cp >= n_target)
        buck = np.bincount(ci[keep].astype(np.int64), minlength=ncontig).astype(np.int64)
        out.append((a[keep].reshape(-1), buck))
    return out[0][0], out[1][0], out[0][1], out[1][1]
