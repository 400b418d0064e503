"""Benchmark: index build + adaptamer seed merge on one MI355X.

One step = one pass of the hot path over the HBM-resident input: syncmer
extraction -> radix partition -> tile sort with LCP and GIX entries -> 2^24
prefix table -> self-merge into N/C seed streams.  The `.bps` bytes are
copied to the device once before timing (the PCIe-inclusive rate is reported
separately).

Printed (one JSON line): `value` is the rate of the human-scale pair (3.1 Gbp
A + B, 4.84 G records, fixed launch shape) with the dominant kernel's roofline
and the device check of its index; beside it the serving loop with genome A resident, the CPU baseline on a
bounded sample of the same pair, the 100 Mbp pair (autotuned), and the
64-pair workload batched and pair by pair, and the diagonal seed sort of the
100 Mbp pair's streams.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from distributedpanga_amd import autotune, pga, serve, synth  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "index+seed Gbases/s and % HBM peak, batched vs plain GPU vs CPU, 1 MI355X"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_steps(ctx, steps, warmup):
    for _ in range(warmup):
        ctx.build_index()
        ctx.merge()
    ctx.sync()
    stage = {k: 0.0 for k in ("extract_ms", "partition_ms", "tile_sort_ms", "table_ms", "merge_ms",
                              "scatter_ms", "merge_emit_ms")}
    scat_bytes = 0
    scat_launches = 0
    last = {}
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.build_index()
        ctx.merge()
        t = ctx.timings()
        for k in stage:
            stage[k] += t[k]
        scat_bytes += t["scatter_bytes"]
        scat_launches += t["scatter_launches"]
        last = t
    ctx.sync()
    dt = time.perf_counter() - t0
    per = {k: v / steps for k, v in stage.items()}
    return dt, per, scat_bytes / steps, scat_launches / steps, last


def algorithmic_bytes(seqtot, nrec, e, s, nseed):
    """Fixed byte model of the path: read .bps, write records, a two-pass sort
    (2 x read+write 16 B), LCP read, GIX entries written and read by the merge,
    seeds written."""
    return seqtot / 4 + 16 * nrec + 64 * nrec + 16 * nrec + 2 * e * nrec + s * nseed


TRAFFIC_PROFILE_HUMAN = "index_merge_human_scale_hbm_traffic.md"
TRAFFIC_PROFILE_PLAIN = "index_merge_workload2_hbm_traffic.md"


def measured_traffic(kernel, profile):
    """HBM bytes per launch of one kernel from a committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (of the full-size workload it names)."""
    if profile is None:
        return None, None
    path = os.path.join(ROOT, "profiles", profile)
    if not os.path.exists(path):
        return None, None
    key = f"traffic_bytes_per_launch[{kernel}]:"
    with open(path) as f:
        for line in f:
            if line.startswith(key):
                return int(line.split(":")[1]), "profiles/" + profile
    return None, None


def subsample(g, per_contig):
    """The first per_contig bases (a multiple of 4) of every contig of g, as one
    GDB: a bounded sample that keeps the workload's contig count, so the
    reference's split of whole contigs over workers stays as wide as on the
    full pair."""
    keep = np.minimum(g.clen, per_contig).astype(np.int64)
    nbytes = (keep + 3) // 4
    parts = [g.bps[int(g.boff[c]): int(g.boff[c]) + int(nbytes[c])] for c in range(g.ncontig)]
    boff = np.zeros(g.ncontig + 1, np.int64)
    boff[1:] = np.cumsum(nbytes)
    return synth.Gdb(bps=np.concatenate(parts), boff=boff, clen=keep, n_a=g.n_a)


def cpu_threads(requested):
    if requested > 0:
        return requested
    env = os.environ.get("OMP_NUM_THREADS")
    return int(env) if env and env.isdigit() and int(env) > 0 else (os.cpu_count() or 1)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(g, target_bases, threads):
    """oracle/pga_cpu.c (the reference's pga-mpi decomposition on `threads`
    host threads) on a bounded sample of the human-scale pair."""
    import oracle                                     # the checker library, timed as the CPU baseline
    T = cpu_threads(threads)
    per = max(4096, (target_bases // max(g.ncontig, 1)) // 4 * 4)
    s = subsample(g, per)
    t0 = time.perf_counter()
    o = oracle.pipeline(s.bps, s.boff[:-1], s.clen, threads=T)
    dt = time.perf_counter() - t0
    return dict(value=round(s.seqtot / dt / 1e9, 6), unit="Gbases/s", cores=T, kind="port",
                host_cpus=os.cpu_count(), cpu_model=cpu_model(),
                sample=f"the first {per} bases of each of the {s.ncontig} contigs of the human-scale pair: "
                       f"{s.seqtot} bases, {o['nrec']} records; pga-mpi's decomposition on {T} threads "
                       f"(extract {o['t_extract']:.2f}s, exchange+sort+LCP {o['t_sort']:.2f}s, "
                       f"adapter {o['t_adapter']:.2f}s, merge {o['t_merge']:.2f}s)",
                seconds=round(dt, 3))


def serving_leg(args, target=None):
    """Serving: genome A of the human-scale pair workload (at --serve-scale;
    `target` when the caller already has it) resident as the target; query genomes (lognormal 0.5-50 Mbp, from random
    target segments at 1-20% divergence, segment order shuffled) arrive as a
    Poisson process at 70% of the measured capacity and are answered by the
    continuous-batching loop over --serve-slots slots."""
    t = time.perf_counter()
    if target is None:
        pair = synth.workload(4, args.serve_scale)
        target = synth.take(pair, 0, pair.n_a)
        del pair
    rng = np.random.default_rng(0x5EED0005)
    sizes = np.clip(np.exp(rng.normal(np.log(5e6), 1.0, args.serve_pool)), 5e5, 5e7).astype(np.int64)
    divs = rng.uniform(0.01, 0.20, args.serve_pool)
    pool = [synth.query(target, 0x5EED0500 + i, int(sizes[i]), float(divs[i])) for i in range(args.serve_pool)]
    log(f"[bench] serving: target {target.seqtot} bases in {target.ncontig} contigs, pool of {len(pool)} queries "
        f"({sum(q.seqtot for q in pool)} bases) generated ({time.perf_counter() - t:.1f}s)")
    budget = int(args.serve_budget_gb * 1e9)
    with serve.Server(target, slots=args.serve_slots) as srv:
        tinfo = srv.target_info()
        for i in range(0, len(pool), args.serve_max_batch):            # warm every slot's buffers
            srv.run(pool[i:i + args.serve_max_batch], slot=(i // args.serve_max_batch) % args.serve_slots)
        # capacity: every request present at t=0, all slots busy
        order = rng.integers(0, len(pool), args.serve_requests)
        burst = [serve.Request(rid=i, arrival=0.0, query=pool[j]) for i, j in enumerate(order)]
        cap, _ = serve.serve(srv, burst, budget, max_batch=args.serve_max_batch)
        cap_bps = cap.bases / cap.wall_s
        mean_bases = float(np.mean([pool[j].seqtot for j in order]))
        rate = 0.7 * cap_bps / mean_bases
        arr = serve.poisson_arrivals(args.serve_requests, rate, seed=5)
        reqs = [serve.Request(rid=i, arrival=float(a), query=pool[j]) for i, (a, j) in enumerate(zip(arr, order))]
        st, _ = serve.serve(srv, reqs, budget, max_batch=args.serve_max_batch)
    sm = st.summary()
    log(f"[bench] serving: capacity {cap_bps / 1e9:.2f} Gbases/s; at 70%: {sm['bases_per_s'] / 1e9:.2f} Gbases/s, "
        f"p50 {sm['p50_ms']:.1f} ms, p99 {sm['p99_ms']:.1f} ms ({time.perf_counter() - t:.1f}s)")
    return dict(workload=f"target = genome A of the human-scale pair workload at scale {args.serve_scale} "
                         f"({target.seqtot} bases, {target.ncontig} contigs, {tinfo['nrec']} resident records); "
                         f"{args.serve_requests} Poisson requests drawn from {len(pool)} query genomes "
                         "(lognormal 0.5-50 Mbp, median 5 Mbp, 1-20% divergence), FREQ=10",
                slots=args.serve_slots, max_batch=args.serve_max_batch, budget_gb=args.serve_budget_gb,
                capacity_gbases_s=round(cap_bps / 1e9, 4), arrival_rate_per_s=round(rate, 3),
                gbases_s=round(sm["bases_per_s"] / 1e9, 4), p50_ms=round(sm["p50_ms"], 3),
                p99_ms=round(sm["p99_ms"], 3), mean_ms=round(sm["mean_ms"], 3), batches=sm["batches"],
                max_batch_seen=sm["max_batch"], max_inflight=sm["max_inflight"],
                slot_busy_sum=round(sm["device_ms"] / 1e3 / sm["wall_s"], 3),
                slot_busy_mean=round(sm["device_ms"] / 1e3 / sm["wall_s"] / args.serve_slots, 3),
                seeds=sm["seeds"],
                pcie_inclusive=True)


def copy_calibration(gib=1, reps=5):
    """Device-to-device copies of a known size through the HIP runtime (its
    copy kernel): the FETCH_SIZE / WRITE_SIZE counters of these dispatches
    against their bytes give the factors that turn the counters into HBM
    bytes.  Prints one line per copy."""
    import ctypes as C
    from distributedpanga_amd._lib import hip_lib
    hip_lib()                                              # the runtime the library links
    L = C.CDLL("libamdhip64.so.7")
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    n = gib << 30
    a, b = C.c_void_p(), C.c_void_p()
    if L.hipMalloc(C.byref(a), n) or L.hipMalloc(C.byref(b), n):
        raise RuntimeError("hipMalloc failed")
    L.hipMemset(a, 1, n)
    L.hipDeviceSynchronize()
    for _ in range(reps):
        t0 = time.perf_counter()
        if L.hipMemcpy(b, a, n, 3):                        # device to device
            raise RuntimeError("hipMemcpy failed")
        L.hipDeviceSynchronize()
        dt = time.perf_counter() - t0
        print(f"copy {n} bytes: {dt * 1e3:.3f} ms host-timed, {2 * n / dt / 1e9:.1f} GB/s read+write", flush=True)
    L.hipFree(a)
    L.hipFree(b)


def human_pair(scale, cache=None):
    """The human-scale pair at `scale`; with `cache` (a .npz path) it is read
    from there when present and written there otherwise, so repeated runs
    (profiler passes) skip the generation."""
    if cache and os.path.exists(cache):
        z = np.load(cache)
        if float(z["scale"]) == scale:
            return synth.Gdb(bps=z["bps"], boff=z["boff"], clen=z["clen"], n_a=int(z["n_a"]))
    g = synth.workload(4, scale)
    if cache:
        np.savez(cache, bps=g.bps, boff=g.boff, clen=g.clen, n_a=g.n_a, scale=scale)
    return g


def kernel_lines(info, nseed, per, traffic_profile):
    """The step's two longest launches, each with its algorithmic bytes per
    launch.  Tile sort: every record read once (16 B); its GIX entry (e B) and
    LCP/strand byte written once (the prefix table gets group-end marks, a few
    per tile, not counted).  Merge emit: every GIX entry read once (e B),
    every seed (s B) and the two buck vectors written once.  Returns
    (dominant, other)."""
    def line(name, nbytes, ms):
        gbs = nbytes / (ms * 1e-3) / 1e9
        traffic, src = measured_traffic(name, traffic_profile)
        return dict(kernel=name, achieved=round(gbs, 1), frac=round(gbs / HBM_PEAK_GBS, 4), traffic=traffic,
                    traffic_source=src, bytes_per_launch=int(nbytes), avg_launch_ms=round(ms, 4))
    sort_k = line("k_tile_sort", info["nrec"] * (16 + info["e"] + 1), per["tile_sort_ms"])
    emit_k = line("k_merge_emit", info["nrec"] * info["e"] + nseed * info["s"] + 16 * info["ncontig"],
                  per["merge_emit_ms"])
    return (emit_k, sort_k) if per["merge_emit_ms"] > per["tile_sort_ms"] else (sort_k, emit_k)


def roofline_of(info, nseed, per, scat_bytes, seqtot, ms_step, traffic_profile):
    dom, other = kernel_lines(info, nseed, per, traffic_profile)
    scat_gbs = scat_bytes / (per["scatter_ms"] * 1e-3) / 1e9 if per["scatter_ms"] > 0 else 0.0
    balg = algorithmic_bytes(seqtot, info["nrec"], info["e"], info["s"], nseed)
    path_gbs = balg / (ms_step * 1e-3) / 1e9
    roofline = dict(bound="hbm", kernel=dom["kernel"], achieved=dom["achieved"], peak=HBM_PEAK_GBS,
                    unit="GB/s", frac=dom["frac"], traffic=dom["traffic"], traffic_source=dom["traffic_source"],
                    bytes_per_launch=dom["bytes_per_launch"], avg_launch_ms=dom["avg_launch_ms"])
    roofline[other["kernel"][2:]] = {k: v for k, v in other.items() if k != "kernel"}
    roofline["radix_scatter"] = dict(achieved=round(scat_gbs, 1), frac=round(scat_gbs / HBM_PEAK_GBS, 4),
                                     bytes_per_step=int(scat_bytes), ms_per_step=round(per["scatter_ms"], 4))
    roofline["path"] = dict(algorithmic_bytes=int(balg), achieved=round(path_gbs, 1),
                            frac=round(path_gbs / HBM_PEAK_GBS, 4))
    return roofline


HUMAN_WORKLOAD = ("human-scale pair: A 3.1 Gbp in 24 contigs of 50-250 Mbp, GC 0.41, 45% repeats from 500 "
                  "families (100-6000 bp, 5-25% divergence) + 2 Mbp of 171 bp satellites; B = A with 2% SNP, "
                  "0.3% indel, 200 rearrangements of 10 kbp-5 Mbp; one GDB, k=40, FREQ=10")
PLAIN_WORKLOAD = ("plain single pair: 100 Mbp A (500 lognormal contigs, GC 0.41, 2000x300 bp repeat at 10%, "
                  "50 tandem arrays) + B (5% SNP, 0.5% indel, 10 inversions), one GDB, k=40, FREQ=10")


def plain_leg(args):
    """The 100 Mbp pair (autotuned, cached per shape): a side leg beside the
    human-scale headline, with the diagonal seed sort of its two streams."""
    opts = dict(tile_records=4096, chunk_records=8192, merge_tile=1024, digit2_bits=10)
    t = time.perf_counter()
    g = synth.workload(2, 1.0)
    log(f"[bench] 100 Mbp pair generated: {g.ncontig} contigs, {g.seqtot} bases ({time.perf_counter() - t:.1f}s)")
    ctx = pga.Context(**opts)
    ctx.load_gdb(g)
    tuned = None
    if not args.no_tune:
        tuned, _ = autotune.tune(ctx, [g], use_cache=not args.retune, log=None)
        opts.update(tuned)
    steps = 10
    dt, per, scat_bytes, _, last = time_steps(ctx, steps, 2)
    info = ctx.index_info(0)
    sinfo = ctx.seeds_info(0)
    nseed = (sinfo["nlen"] + sinfo["clen"]) // sinfo["s"]
    # the next step downstream: the diagonal-sorted seed array of both streams
    ss = {}
    for comp, name in (() if args.no_seed_sort else ((0, "N"), (1, "C"))):
        ctx.seed_sort(0, comp)                      # warm the buffers
        t0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            n = ctx._seed_sort_only(0, comp)
        ss[name] = dict(records=n, ms=round((time.perf_counter() - t0) / reps * 1e3, 3))
    ctx.close()
    ms = dt / steps * 1e3
    roof = roofline_of(info, nseed, per, scat_bytes, g.seqtot, ms, TRAFFIC_PROFILE_PLAIN)
    log(f"[bench] 100 Mbp pair: {ms:.2f} ms/step, {g.seqtot / (ms * 1e-3) / 1e9:.2f} Gbases/s")
    return dict(workload=PLAIN_WORKLOAD, bases=g.seqtot, records=info["nrec"], seeds=int(nseed),
                ms_per_step=round(ms, 3), gbases_s=round(g.seqtot / (ms * 1e-3) / 1e9, 4), steps=steps,
                opts=opts, autotuned=tuned is not None, stages_ms={k: round(v, 4) for k, v in per.items()},
                tiles=last.get("tiles"), crowded_tiles=last.get("crowded_tiles"),
                fallback_tiles=last.get("fallback_tiles"),
                roofline={k: roof[k] for k in ("kernel", "achieved", "frac", "radix_scatter", "path")},
                seed_sort=seed_sort_line(ss, sinfo["s"]))


def seed_sort_line(ss, s):
    """The seed sort's figures: both streams' records and time, and a roofline
    on the sort's own input and output (each seed read once, s bytes, and its
    sorted record written once, s bytes); the passes over 16-byte keys in
    between are the sort's cost, not counted as useful bytes."""
    n = sum(v["records"] for v in ss.values())
    ms = sum(v["ms"] for v in ss.values())
    alg = 2 * s * n
    ach = alg / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    return dict(streams=ss, records=n, ms=round(ms, 3),
                roofline=dict(bound="hbm", algorithmic_bytes=alg, achieved=round(ach, 1), peak=HBM_PEAK_GBS,
                              unit="GB/s", frac=round(ach / HBM_PEAK_GBS, 4)),
                note="reimport + per-panel sort of both seed streams, records left on the device")


def batch_leg(args):
    t = time.perf_counter()
    gdbs = synth.batch(args.batch_pairs, scale=1.0)
    tot = sum(x.seqtot for x in gdbs)
    log(f"[bench] batch workload generated: {len(gdbs)} pairs, {tot} bases ({time.perf_counter() - t:.1f}s)")
    base = dict(tile_records=4096, chunk_records=8192, merge_tile=1024, digit2_bits=10)
    bopts = dict(base)
    bctx = pga.Context(**bopts)
    bctx.load_batch(gdbs)
    if not args.no_tune:
        btuned, _ = autotune.tune(bctx, gdbs, use_cache=not args.retune, log=None)
        bopts.update(btuned)
    bsteps = 5
    bdt, bper, _, _, _ = time_steps(bctx, bsteps, 1)
    # the same batch as one device-driven pipeline replayed from a HIP graph
    bctx.run_graph()                               # capture
    bctx.sync()
    t0 = time.perf_counter()
    for _ in range(bsteps):
        bctx.run_graph()
    bctx.sync()
    gdt = time.perf_counter() - t0
    gfall = bctx.timings()["graph_fallbacks"]
    bctx.close()
    batched = tot * bsteps / bdt / 1e9
    graphed = tot * bsteps / gdt / 1e9
    # plain form: the same pairs one at a time (input resident, one build + merge
    # per pair), tuned on the median pair
    popts = dict(base)
    pctx = pga.Context(**popts)
    if not args.no_tune:
        med = sorted(gdbs, key=lambda x: x.seqtot)[len(gdbs) // 2]
        pctx.load_gdb(med)
        ptuned, _ = autotune.tune(pctx, [med], use_cache=not args.retune, log=None)
        popts.update(ptuned)
    plain_s = 0.0
    for x in gdbs:
        pctx.load_gdb(x)
        pctx.build_index()
        pctx.merge()
        pctx.sync()
        t0 = time.perf_counter()
        pctx.build_index()
        pctx.merge()
        pctx.sync()
        plain_s += time.perf_counter() - t0
    pctx.close()
    plain = tot / plain_s / 1e9
    log(f"[bench] batched {batched:.2f} Gbases/s (graph {graphed:.2f}) vs plain per pair {plain:.2f} Gbases/s")
    return dict(workload=f"{len(gdbs)} independent pairs, A lognormal 1-12 Mbp, 1-50 contigs, "
                         "GC 0.30-0.65, B at 1-15% SNP and 0.1-1% indel",
                bases=tot, batched_gbases_s=round(batched, 4), batched_ms=round(bdt / bsteps * 1e3, 3),
                plain_gbases_s=round(plain, 4), plain_ms=round(plain_s * 1e3, 3),
                speedup=round(batched / plain, 3), batched_opts=bopts, plain_opts=popts,
                graph_gbases_s=round(graphed, 4), graph_ms=round(gdt / bsteps * 1e3, 3),
                graph_speedup=round(graphed / plain, 3), graph_fallbacks=gfall,
                batched_stages_ms={k: round(v, 3) for k, v in bper.items()})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=float, default=1.0,
                    help="scale of the human-scale pair (1.0 = 3.1 Gbp A + B)")
    ap.add_argument("--tile", type=int, default=4096)
    ap.add_argument("--chunk", type=int, default=8192)
    ap.add_argument("--merge-tile", type=int, default=1024)
    ap.add_argument("--digit2", type=int, default=8, help="bits of the partition's second digit (8..10)")
    ap.add_argument("--no-plain", action="store_true", help="skip the 100 Mbp pair leg")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-seed-sort", action="store_true")
    ap.add_argument("--only-headline", action="store_true", help="the human-scale pair only (profiling)")
    ap.add_argument("--only-plain", action="store_true", help="the 100 Mbp pair leg only (profiling)")
    ap.add_argument("--only-serve", action="store_true", help="the serving leg only (profiling)")
    ap.add_argument("--gdb-cache", default=None,
                    help="a .npz file holding the generated human-scale pair (read if present, else written)")
    ap.add_argument("--cpu-bases", type=int, default=500_000_000)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the CPU baseline (0: OMP_NUM_THREADS, else the host's cores)")
    ap.add_argument("--batch-pairs", type=int, default=64)
    ap.add_argument("--serve-scale", type=float, default=1.0,
                    help="scale of the resident serving target (genome A of the human-scale pair)")
    ap.add_argument("--serve-pool", type=int, default=48)
    ap.add_argument("--serve-requests", type=int, default=240)
    ap.add_argument("--serve-slots", type=int, default=2)
    ap.add_argument("--serve-max-batch", type=int, default=8)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--no-tune", action="store_true", help="side legs use fixed shapes, not the autotuner")
    ap.add_argument("--retune", action="store_true", help="ignore the autotuner cache")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 1 GiB device copies (the counter calibration for the traffic profiles)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU exchange is a design, not a run)", file=sys.stderr)
        sys.exit(2)
    if args.only_headline:
        args.no_plain = args.no_batch = args.no_cpu = args.no_serve = args.no_seed_sort = True
    if args.only_plain or args.only_serve:          # one side leg alone, printed as its JSON object
        print(json.dumps(plain_leg(args) if args.only_plain else serving_leg(args)), flush=True)
        return

    # ---- headline: the human-scale pair, input resident in HBM
    opts = dict(tile_records=args.tile, chunk_records=args.chunk, merge_tile=args.merge_tile,
                digit2_bits=args.digit2)
    t = time.perf_counter()
    g = human_pair(args.scale, args.gdb_cache)
    log(f"[bench] human-scale pair at scale {args.scale}: {g.ncontig} contigs, {g.seqtot} bases ready "
        f"({time.perf_counter() - t:.1f}s)")
    ctx = pga.Context(**opts)
    t = time.perf_counter()
    ctx.load_gdb(g)
    h2d_s = time.perf_counter() - t
    dt, per, scat_bytes, scat_launches, last = time_steps(ctx, args.steps, args.warmup)
    info = ctx.index_info(0)
    sinfo = ctx.seeds_info(0)
    nseed = (sinfo["nlen"] + sinfo["clen"]) // sinfo["s"]
    ms_step = dt / args.steps * 1e3
    value = g.seqtot * args.steps / dt / 1e9
    log(f"[bench] human-scale pair: {ms_step:.2f} ms/step, {value:.2f} Gbases/s, {info['nrec']} records, "
        f"{nseed} seeds, stages {json.dumps({k: round(v, 3) for k, v in per.items()})}")
    bad, first_bad, lcp_bad = ctx.verify(0)
    log(f"[bench] device check: {bad} order violations, {lcp_bad} LCP mismatches")
    roofline = roofline_of(info, nseed, per, scat_bytes, g.seqtot, ms_step,
                           TRAFFIC_PROFILE_HUMAN if args.scale == 1.0 else None)
    roofline["radix_scatter"]["launches_per_step"] = scat_launches
    ctx.close()
    out = {"metric": METRIC, "value": round(value, 4), "unit": "Gbases/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
           "data": "synthetic",
           "config": {"workload": HUMAN_WORKLOAD if args.scale == 1.0 else
                      f"{HUMAN_WORKLOAD} (at scale {args.scale})",
                      "bases": g.seqtot, "contigs": g.ncontig, "records": info["nrec"], "seeds": int(nseed),
                      "P": info["P"], "C": info["C"], "scale": args.scale, "tile_records": opts["tile_records"],
                      "chunk_records": opts["chunk_records"], "merge_tile": opts["merge_tile"],
                      "digit2_bits": opts["digit2_bits"], "parallelism": "1 GPU"},
           "stages_ms": {k: round(v, 4) for k, v in per.items()},
           "tiles": last.get("tiles"), "crowded_tiles": last.get("crowded_tiles"),
           "fallback_tiles": last.get("fallback_tiles"),
           "partition_levels": last.get("partition_levels"),
           "device_check": dict(kmer_order_violations=bad, first_bad=first_bad, lcp_mismatches=lcp_bad),
           "h2d_load_s": round(h2d_s, 4),
           "pcie_inclusive_gbases_s": round(g.seqtot / (ms_step * 1e-3 + h2d_s) / 1e9, 4),
           "roofline": roofline}
    if not args.no_serve:
        target = synth.take(g, 0, g.n_a) if args.serve_scale == args.scale else None
        out["serving"] = serving_leg(args, target)
    if not args.no_cpu:
        t = time.perf_counter()
        out["cpu_baseline"] = cpu_baseline(g, args.cpu_bases, args.cpu_threads)
        log(f"[bench] cpu baseline {out['cpu_baseline']['value']} Gbases/s ({time.perf_counter() - t:.1f}s)")
    del g
    if not args.no_plain:
        out["plain_100mbp"] = plain_leg(args)
    if not args.no_batch:
        out["batched"] = batch_leg(args)

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
