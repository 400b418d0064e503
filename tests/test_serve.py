"""Serving: a resident target answering query genomes.

GPU: each query's answer (pga_server_*) equals the oracle's self-merge of the
combined GDB (target contigs, then the query's) reduced to the seeds with a
query-side anchor or partner, byte for byte, and the same whether the query is
answered alone or in a batch.  CPU: the request loop's admission and batching,
and the host-side reduction.
"""
import threading
import time

import numpy as np
import pytest

import oracle
from distributedpanga_amd import serve, synth


# ---------------------------------------------------------------- CPU: host logic

class FakeServer:
    """Stands in for Server in the request loop: each batch takes a fixed time."""

    def __init__(self, slots=1, per_batch_s=0.02, fp=100):
        self.slots, self.per_batch_s, self.fp = slots, per_batch_s, fp
        self.batches = []
        self.lock = threading.Lock()
        self.running = 0
        self.max_running = 0
        self.sizes = {}

    def footprint(self, bases):
        return self.fp * bases

    def run(self, queries, slot=0):
        with self.lock:
            self.batches.append(len(queries))
            self.running += 1
            self.max_running = max(self.max_running, self.running)
            self.sizes[slot] = len(queries)
        time.sleep(self.per_batch_s)
        with self.lock:
            self.running -= 1

    def last_ms(self, slot=0):
        return 1000.0 * self.per_batch_s

    def seeds_info(self, q, slot=0):
        return dict(nlen=10 * 7, clen=0, s=7)


def _reqs(arrivals, bases=1000):
    q = synth.Gdb(bps=np.zeros(1, np.uint8), boff=np.zeros(2, np.int64), clen=np.array([bases], np.int64), n_a=1)
    return [serve.Request(rid=i, arrival=float(a), query=q, bases=bases) for i, a in enumerate(arrivals)]


def test_loop_batches_what_has_arrived():
    srv = FakeServer(slots=1, per_batch_s=0.05)
    reqs = _reqs([0.0, 0.005, 0.01, 0.2])
    st, _ = serve.serve(srv, reqs, budget_bytes=10 ** 9)
    # request 0 runs alone; 1 and 2 arrive while it runs and form the next batch; 3 comes later
    assert srv.batches == [1, 2, 1]
    assert st.requests == 4 and st.batches == 3 and st.max_batch == 2 and st.seeds == 40
    assert all(r.done >= r.arrival + 0.05 for r in reqs)
    assert reqs[1].start >= 0.05 and reqs[3].start >= 0.2


def test_loop_overlaps_batches_on_free_slots():
    srv = FakeServer(slots=3, per_batch_s=0.1)
    reqs = _reqs([0.0, 0.01, 0.02, 0.03])
    st, _ = serve.serve(srv, reqs, budget_bytes=10 ** 9)
    # each arrival finds a free slot and starts at once; the fourth waits for one
    assert srv.batches[:3] == [1, 1, 1]
    assert srv.max_running == 3 and st.max_inflight == 3
    assert reqs[3].start >= 0.1 - 1e-3


def test_loop_respects_the_budget():
    srv = FakeServer(slots=4, fp=100)
    reqs = _reqs([0.0] * 7, bases=1000)          # 100 kB each
    st, _ = serve.serve(srv, reqs, budget_bytes=250_000)
    assert st.max_inflight_bytes <= 250_000
    assert sum(srv.batches) == 7 and max(srv.batches) <= 2
    srv2 = FakeServer(slots=2, fp=100)
    st2, _ = serve.serve(srv2, _reqs([0.0] * 3, bases=10_000), budget_bytes=250_000)
    assert srv2.batches == [1, 1, 1] and st2.max_inflight == 1   # each alone exceeds the budget


def test_loop_max_batch():
    srv = FakeServer(slots=1, per_batch_s=0.01)
    serve.serve(srv, _reqs([0.0] * 10), budget_bytes=10 ** 12, max_batch=4)
    assert srv.batches == [4, 4, 2]


def test_loop_surfaces_errors():
    class Broken(FakeServer):
        def run(self, queries, slot=0):
            raise RuntimeError("device failure")
    with pytest.raises(RuntimeError):
        serve.serve(Broken(slots=2), _reqs([0.0, 0.0]), budget_bytes=10 ** 9)


def test_poisson_arrivals_rate():
    a = serve.poisson_arrivals(20000, 50.0, seed=3)
    assert np.all(np.diff(a) > 0)
    assert abs(20000 / a[-1] - 50.0) < 2.0


def test_filter_query_seeds():
    # P=1, C=1, s=5: [plen][pos][ctg|strand<<7][pos][ctg|strand<<7]; target contigs 0,1; query 2,3
    rows = np.array([[20, 5, 0, 9, 1],        # target-target: dropped
                     [21, 5, 0, 7, 2],        # target anchor, query partner
                     [22, 6, 3, 1, 0x80 | 1],  # query anchor
                     [23, 4, 1, 2, 0x80 | 0]], np.uint8)
    N, Cs, nb, cb = serve.filter_query_seeds(rows.reshape(-1), np.zeros(0, np.uint8), 5, 1, 1, 2, 4)
    assert np.array_equal(N.reshape(-1, 5), rows[[1, 2]])
    assert Cs.size == 0
    assert nb.tolist() == [1, 0, 0, 1] and cb.tolist() == [0, 0, 0, 0]


# ---------------------------------------------------------------- GPU parity

def _expected(target, q, freq):
    comb = synth.concat(target, q)
    o = oracle.pipeline(comb.bps, comb.boff[:-1], comb.clen, freq=freq)
    s = 1 + 2 * (o["P"] + o["C"])
    N, Cs, nb, cb = serve.filter_query_seeds(o["nseeds"], o["cseeds"], s, o["P"], o["C"], target.ncontig,
                                             comb.ncontig)
    return dict(nseeds=N, cseeds=Cs, nbuck=nb, cbuck=cb, s=s)


def _same(got, want, what):
    assert got["s"] == want["s"], what
    for k in ("nseeds", "cseeds", "nbuck", "cbuck"):
        assert np.array_equal(got[k], want[k]), f"{what}: {k}"


def _queries(target):
    rng = np.random.default_rng(11)
    qs = [synth.query(target, 101, 60_000, 0.03, seg_min=2_000, seg_max=20_000),
          synth.query(target, 102, 30_000, 0.0, seg_min=5_000, seg_max=15_000),
          synth.query(target, 103, 40_000, 0.12, seg_min=1_000, seg_max=8_000),
          synth.from_codes([rng.integers(0, 4, 20_000).astype(np.uint8)]),      # unrelated
          synth.from_codes([np.zeros(30, np.uint8)])]                          # no records
    return qs


@pytest.mark.gpu
@pytest.mark.parametrize("freq", [10, 1, 24])
def test_server_matches_oracle_alone_and_in_batch(freq):
    target = synth.workload(1, 0.1)
    qs = _queries(target)
    with serve.Server(target, freq=freq) as srv:
        mix = srv.answer(qs)
        for i, q in enumerate(qs):
            print(f"[serve] freq={freq} query {i} contigs={q.ncontig} bases={q.seqtot}", flush=True)
            want = _expected(target, q, freq)
            _same(mix[i], want, f"query {i} in batch")
            alone = srv.answer([q])[0]
            _same(alone, want, f"query {i} alone")


@pytest.mark.gpu
def test_server_low_complexity_groups():
    """Large 24-bit groups on both sides: segments cut by the SEG_Q rule and
    cores far wider than the LDS window of any one group."""
    rng = np.random.default_rng(5)
    unit = rng.integers(0, 4, 37).astype(np.uint8)
    tandem = np.tile(unit, 400)
    tandem[rng.integers(0, tandem.size, 60)] = rng.integers(0, 4, 60)
    contigs = [tandem, np.tile(np.array([0, 1, 1], np.uint8), 3000), rng.integers(0, 4, 30_000).astype(np.uint8)]
    target = synth.from_codes(contigs)
    q1 = synth.from_codes([np.tile(unit, 150), tandem[1000:5000].copy()])
    q2 = synth.from_codes([np.tile(np.array([0, 1, 1], np.uint8), 700), contigs[2][100:9000].copy()])
    with serve.Server(target, freq=24, tile_records=2048) as srv:
        got = srv.answer([q1, q2])
        for i, q in enumerate((q1, q2)):
            print(f"[serve-lc] query {i}", flush=True)
            _same(got[i], _expected(target, q, 24), f"query {i}")


@pytest.mark.gpu
def test_server_many_contigs():
    """Target plus query past 1024 contigs: per-contig seed counts of contigs
    beyond the join's LDS counters go through the global-atomic branch."""
    rng = np.random.default_rng(17)
    tcont = [rng.integers(0, 4, int(n)).astype(np.uint8) for n in rng.integers(300, 1500, 900)]
    target = synth.from_codes(tcont)

    def pieces(n, seed):
        r = np.random.default_rng(seed)
        out = []
        for _ in range(n):
            c = tcont[int(r.integers(0, len(tcont)))]
            a = int(r.integers(0, c.size // 2))
            p = c[a:a + int(r.integers(150, c.size - a + 1))].copy()
            flip = r.random(p.size) < 0.02
            p[flip] = r.integers(0, 4, int(flip.sum()))
            out.append(p if r.random() < 0.5 else (3 - p[::-1]).astype(np.uint8))
        return synth.from_codes(out)

    qs = [pieces(700, 1), pieces(300, 2)]
    assert target.ncontig + qs[0].ncontig > 1024
    with serve.Server(target) as srv:
        got = srv.answer(qs)
        for i, q in enumerate(qs):
            print(f"[serve-many] query {i} contigs={q.ncontig}", flush=True)
            want = _expected(target, q, 10)
            assert want["nbuck"][1024:].sum() > 0
            _same(got[i], want, f"query {i}")


@pytest.mark.gpu
def test_server_many_batches_reuse():
    target = synth.workload(1, 0.05)
    qs = [synth.query(target, 200 + i, 10_000 + 5_000 * i, 0.02 * i, seg_min=1_000, seg_max=6_000)
          for i in range(6)]
    with serve.Server(target) as srv:
        for batch in (qs[:4], qs[4:], qs[::-1], qs[2:3]):
            got = srv.answer(batch)
            for g, q in zip(got, batch):
                _same(g, _expected(target, q, 10), "batch")


@pytest.mark.gpu
def test_server_rejects_large_freq():
    target = synth.workload(1, 0.01)
    with pytest.raises(Exception):
        serve.Server(target, freq=25)


@pytest.mark.gpu
def test_serve_loop_overlapping_slots_match_alone():
    """The request loop with three slots driven concurrently: every answer
    equals the oracle's and the query's answer when run alone."""
    target = synth.workload(1, 0.1)
    qs = [synth.query(target, 300 + i, 20_000 + 7_000 * (i % 5), 0.01 + 0.03 * (i % 4), seg_min=1_000,
                      seg_max=9_000) for i in range(14)]
    reqs = [serve.Request(rid=i, arrival=0.002 * i, query=q) for i, q in enumerate(qs)]
    with serve.Server(target, slots=3) as srv:
        st, answers = serve.serve(srv, reqs, budget_bytes=1 << 40, max_batch=4, export=True)
        print(f"[serve-loop] {st.summary()}", flush=True)
        assert st.requests == len(qs) and len(answers) == len(qs)
        for i, q in enumerate(qs):
            want = _expected(target, q, 10)
            _same(answers[i], want, f"request {i} in the loop")
            _same(srv.answer([q], slot=i % 3)[0], want, f"request {i} alone")


@pytest.mark.gpu
@pytest.mark.parametrize("freq", [10, 24])
def test_server_target_from_persisted_gix(tmp_path, freq):
    """A serving target restarted from its GIX files (pga_server_set_target_gix):
    every query's answer equals the one from the target indexed from its
    sequence, and the oracle's."""
    from distributedpanga_amd import gix, pga
    target = synth.workload(1, 0.1)
    qs = _queries(target)
    with pga.Context() as ctx:
        ctx.load_gdb(target)
        ctx.build_index()
        gix.write_from_context(ctx, tmp_path, "target", nparts=2)
    with serve.Server(target, freq=freq) as srv:
        want = srv.answer(qs)
        tinfo = srv.target_info()
    with serve.Server.from_gix(tmp_path, "target", target.clen, freq=freq) as srv2:
        info2 = srv2.target_info()
        assert info2["nrec"] == tinfo["nrec"] and info2["maxpre"] == tinfo["maxpre"]
        assert (info2["P"], info2["C"]) == (tinfo["P"], tinfo["C"])
        got = srv2.answer(qs)
        for i, q in enumerate(qs):
            _same(got[i], want[i], f"query {i} from the GIX target")
            _same(got[i], _expected(target, q, freq), f"query {i} against the oracle")


@pytest.mark.gpu
def test_server_gix_target_rejects_mismatched_contigs(tmp_path):
    from distributedpanga_amd import gix, pga
    target = synth.workload(1, 0.05)
    with pga.Context() as ctx:
        ctx.load_gdb(target)
        ctx.build_index()
        gix.write_from_context(ctx, tmp_path, "t")
    idx = gix.read(tmp_path, "t")
    with serve.Server(None) as srv:
        with pytest.raises(pga.PgaError, match="do not fit"):
            srv.set_target_gix(idx, np.full(target.ncontig, 1 << 26, np.int64))   # P would be 4, not 3
