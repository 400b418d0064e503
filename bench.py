"""Benchmark: index build + adaptamer seed merge on one MI355X.

One step = one pass of the hot path over the HBM-resident input: syncmer
extraction -> radix partition -> tile sort with LCP and GIX entries -> 2^24
prefix table -> self-merge into N/C seed streams.  The `.bps` bytes are
copied to the device once before timing (the PCIe-inclusive rate is reported
separately).

Printed (one JSON line): the plain single-pair rate on the 100 Mbp pair
workload (`value`), the batched 64-pair workload in one batch and pair by pair,
the dominant kernel's roofline, and the CPU baseline (the C oracle on a
bounded sample of the same workload, one core).
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
    tiles = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.build_index()
        ctx.merge()
        t = ctx.timings()
        for k in stage:
            stage[k] += t[k]
        scat_bytes += t["scatter_bytes"]
        scat_launches += t["scatter_launches"]
        tiles = t["tiles"]
    ctx.sync()
    dt = time.perf_counter() - t0
    per = {k: v / steps for k, v in stage.items()}
    return dt, per, scat_bytes / steps, scat_launches / steps, tiles


def algorithmic_bytes(seqtot, nrec, e, s, nseed):
    """Fixed byte model of the path: read .bps, write records, a two-pass sort
    (2 x read+write 16 B), LCP read, GIX entries written and read by the merge,
    seeds written."""
    return seqtot / 4 + 16 * nrec + 64 * nrec + 16 * nrec + 2 * e * nrec + s * nseed


TRAFFIC_PROFILE = "index_merge_workload2_hbm_traffic.md"


def measured_traffic(kernel, scale):
    """HBM bytes per launch of one kernel from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (valid for the full-size workload only)."""
    path = os.path.join(ROOT, "profiles", TRAFFIC_PROFILE)
    if scale != 1.0 or not os.path.exists(path):
        return None, None
    key = f"traffic_bytes_per_launch[{kernel}]:"
    with open(path) as f:
        for line in f:
            if line.startswith(key):
                return int(line.split(":")[1]), "profiles/" + TRAFFIC_PROFILE
    return None, None


def subsample(g, target_bases):
    """The first A contigs and their B counterparts until target_bases."""
    n_a = g.n_a
    pick, tot = [], 0
    for i in range(n_a):
        if tot >= target_bases:
            break
        pick += [i, n_a + i]
        tot += int(g.clen[i] + g.clen[n_a + i])
    codes = [g.contig_codes(c) for c in pick]
    return synth.from_codes(codes, n_a=len(pick) // 2)


def cpu_baseline(g, target_bases):
    import oracle                                     # the checker, timed as the CPU baseline
    s = subsample(g, target_bases)
    t0 = time.perf_counter()
    o = oracle.pipeline(s.bps, s.boff[:-1], s.clen)
    dt = time.perf_counter() - t0
    return dict(value=round(s.seqtot / dt / 1e9, 6), unit="Gbases/s", cores=1, kind="port",
                sample=f"first {len(s.clen) // 2} A contigs + their B counterparts of the 100 Mbp pair "
                       f"workload: {s.seqtot} bases, {o['nrec']} records, single-threaded C oracle "
                       f"(extract {o['t_extract']:.2f}s, sort+LCP {o['t_sort']:.2f}s, "
                       f"adapter {o['t_adapter']:.2f}s, merge {o['t_merge']:.2f}s)",
                seconds=round(dt, 3))


def serving_leg(args):
    """Serving: genome A of the human-scale pair workload (at --serve-scale)
    resident as the target; query genomes (lognormal 0.5-50 Mbp, from random
    target segments at 1-20% divergence, segment order shuffled) arrive as a
    Poisson process at 70% of the measured capacity and are answered by the
    continuous-batching loop over --serve-slots slots."""
    t = time.perf_counter()
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
                device_busy_frac=round(sm["device_ms"] / 1e3 / sm["wall_s"], 3), seeds=sm["seeds"],
                pcie_inclusive=True)


def human_leg(args, opts):
    """The human-scale pair (3.1 Gbp A + B, 24 contigs of 50-250 Mbp, 45%
    repeats, satellites; B with 2% SNP, 0.3% indel, 200 rearrangements) at
    --human-scale: index build + merge with the input resident in HBM, the
    device sortedness/LCP check, and the memory the context holds.  It runs
    with the plain leg's shape options (tuned or given)."""
    t = time.perf_counter()
    log(f"[bench] generating the human-scale pair at scale {args.human_scale}")
    g = synth.workload(4, args.human_scale)
    log(f"[bench] human-scale pair at scale {args.human_scale}: {g.ncontig} contigs, {g.seqtot} bases "
        f"generated ({time.perf_counter() - t:.1f}s)")
    ctx = pga.Context(**opts)
    ctx.load_gdb(g)
    steps = max(1, args.human_steps)
    dt, per, scat_bytes, _, _ = time_steps(ctx, steps, 1)
    info = ctx.index_info(0)
    sinfo = ctx.seeds_info(0)
    nseed = (sinfo["nlen"] + sinfo["clen"]) // sinfo["s"]
    bad, first_bad, lcp_bad = ctx.verify(0)
    ms = dt / steps * 1e3
    scat_gbs = scat_bytes / (per["scatter_ms"] * 1e-3) / 1e9 if per["scatter_ms"] > 0 else 0.0
    out = dict(workload=f"human-scale pair at scale {args.human_scale}", bases=g.seqtot, contigs=g.ncontig,
               records=info["nrec"], seeds=int(nseed), ms_per_step=round(ms, 3),
               gbases_s=round(g.seqtot / (ms * 1e-3) / 1e9, 4), steps=steps,
               stages_ms={k: round(v, 3) for k, v in per.items()},
               radix_scatter_gbs=round(scat_gbs, 1), radix_scatter_frac=round(scat_gbs / HBM_PEAK_GBS, 4),
               verify=dict(kmer_order_violations=bad, first_bad=first_bad, lcp_mismatches=lcp_bad),
               P=info["P"], C=info["C"], maxpre=info["maxpre"], opts=dict(opts))
    ctx.close()
    log(f"[bench] human-scale: {ms:.1f} ms/step, {out['gbases_s']} Gbases/s, {info['nrec']} records, "
        f"{nseed} seeds, verify {out['verify']}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=float, default=1.0, help="workload scale (1.0 = full size)")
    ap.add_argument("--tile", type=int, default=4096)
    ap.add_argument("--chunk", type=int, default=8192)
    ap.add_argument("--merge-tile", type=int, default=512)
    ap.add_argument("--digit2", type=int, default=10, help="bits of the partition's second digit (8..10)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-bases", type=int, default=150_000_000)
    ap.add_argument("--batch-pairs", type=int, default=64)
    ap.add_argument("--human-scale", type=float, default=0.0,
                    help="also run the human-scale pair at this scale (1.0 = 3.1 Gbp A + B); 0 skips it")
    ap.add_argument("--human-steps", type=int, default=3)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-seed-sort", action="store_true", help="skip the seed-sort leg (keeps profiles to the step)")
    ap.add_argument("--serve-scale", type=float, default=0.1, help="scale of the resident target (1.0 = 3.1 Gbp)")
    ap.add_argument("--serve-pool", type=int, default=48)
    ap.add_argument("--serve-requests", type=int, default=240)
    ap.add_argument("--serve-slots", type=int, default=2)
    ap.add_argument("--serve-max-batch", type=int, default=8)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--no-tune", action="store_true", help="use the fixed --tile/--chunk/--merge-tile/--digit2")
    ap.add_argument("--retune", action="store_true", help="ignore the autotuner cache")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU exchange is a design, not a run)", file=sys.stderr)
        sys.exit(2)

    opts = dict(tile_records=args.tile, chunk_records=args.chunk, merge_tile=args.merge_tile,
                digit2_bits=args.digit2)
    t = time.perf_counter()
    g = synth.workload(2, args.scale)
    log(f"[bench] workload 2 generated: {g.ncontig} contigs, {g.seqtot} bases ({time.perf_counter() - t:.1f}s)")

    ctx = pga.Context(**opts)
    t = time.perf_counter()
    ctx.load_gdb(g)
    h2d_s = time.perf_counter() - t
    tuned = None
    if not args.no_tune:
        t = time.perf_counter()
        tuned, _ = autotune.tune(ctx, [g], use_cache=not args.retune, log=log)
        opts.update(tuned)
        log(f"[bench] tuned plain: {tuned} ({time.perf_counter() - t:.1f}s)")
    dt, per, scat_bytes, scat_launches, tiles = time_steps(ctx, args.steps, args.warmup)
    info = ctx.index_info(0)
    sinfo = ctx.seeds_info(0)
    nseed = (sinfo["nlen"] + sinfo["clen"]) // sinfo["s"]
    ms_step = dt / args.steps * 1e3
    value = g.seqtot * args.steps / dt / 1e9
    log(f"[bench] plain: {ms_step:.2f} ms/step, {value:.2f} Gbases/s, {info['nrec']} records, {nseed} seeds, "
        f"stages {json.dumps({k: round(v, 3) for k, v in per.items()})}")

    # The two longest launches of the step, each with its algorithmic bytes per
    # launch.  Tile sort: every record read once (16 B); its GIX entry (e B),
    # LCP/strand byte and 24-bit prefix (4 B) written once.  Merge emit: every
    # GIX entry read once (e B), every seed (s B) and the two buck vectors
    # written once.  The longer one is the roofline kernel, the other is
    # reported beside it.
    def kernel_line(name, nbytes, ms):
        gbs = nbytes / (ms * 1e-3) / 1e9
        traffic, src = measured_traffic(name, args.scale)
        return dict(kernel=name, achieved=round(gbs, 1), frac=round(gbs / HBM_PEAK_GBS, 4), traffic=traffic,
                    traffic_source=src, bytes_per_launch=int(nbytes), avg_launch_ms=round(ms, 4))
    sort_k = kernel_line("k_tile_sort", info["nrec"] * (16 + info["e"] + 1 + 4), per["tile_sort_ms"])
    emit_k = kernel_line("k_merge_emit", info["nrec"] * info["e"] + nseed * info["s"] + 16 * info["ncontig"],
                         per["merge_emit_ms"])
    dom, other = (emit_k, sort_k) if per["merge_emit_ms"] > per["tile_sort_ms"] else (sort_k, emit_k)
    scat_gbs = scat_bytes / (per["scatter_ms"] * 1e-3) / 1e9 if per["scatter_ms"] > 0 else 0.0
    balg = algorithmic_bytes(g.seqtot, info["nrec"], info["e"], info["s"], nseed)
    path_gbs = balg / (ms_step * 1e-3) / 1e9
    roofline = dict(bound="hbm", kernel=dom["kernel"], achieved=dom["achieved"], peak=HBM_PEAK_GBS,
                    unit="GB/s", frac=dom["frac"], traffic=dom["traffic"], traffic_source=dom["traffic_source"],
                    bytes_per_launch=dom["bytes_per_launch"], avg_launch_ms=dom["avg_launch_ms"])
    roofline[other["kernel"][2:]] = {k: v for k, v in other.items() if k not in ("kernel", "traffic_source")}
    roofline["radix_scatter"] = dict(achieved=round(scat_gbs, 1), frac=round(scat_gbs / HBM_PEAK_GBS, 4),
                                     bytes_per_step=int(scat_bytes), launches_per_step=scat_launches,
                                     ms_per_step=round(per["scatter_ms"], 4))
    roofline["path"] = dict(algorithmic_bytes=int(balg), achieved=round(path_gbs, 1),
                            frac=round(path_gbs / HBM_PEAK_GBS, 4))
    # the next step downstream: the diagonal-sorted seed array of both streams
    ss = {}
    for comp, name in (() if args.no_seed_sort else ((0, "N"), (1, "C"))):
        ctx.seed_sort(0, comp)                      # warm the buffers
        t0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            n = ctx._seed_sort_only(0, comp)
        ss[name] = dict(records=n, ms=round((time.perf_counter() - t0) / reps * 1e3, 3))
    ss_seeds = sum(v["records"] for v in ss.values())
    ss_ms = sum(v["ms"] for v in ss.values())
    seed_sort = dict(streams=ss, records=ss_seeds, ms=round(ss_ms, 3),
                     mrecords_s=round(ss_seeds / (ss_ms * 1e-3) / 1e6, 1) if ss_ms > 0 else None,
                     note="reimport + per-panel sort of both seed streams, records left on the device")
    if ss:
        log(f"[bench] seed sort: {ss_seeds} records in {ss_ms:.2f} ms")
    ctx.close()

    out = {"metric": METRIC, "value": round(value, 4), "unit": "Gbases/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
           "data": "synthetic",
           "config": {"workload": "plain single pair, 100 Mbp A (500 lognormal contigs, GC 0.41, "
                                  "2000x300 bp repeat at 10%, 50 tandem arrays) + B (5% SNP, 0.5% indel, "
                                  "10 inversions), one GDB, k=40, FREQ=10",
                      "bases": g.seqtot, "contigs": g.ncontig, "records": info["nrec"], "seeds": int(nseed),
                      "scale": args.scale, "tile_records": opts["tile_records"],
                      "chunk_records": opts["chunk_records"], "merge_tile": opts["merge_tile"],
                      "digit2_bits": opts["digit2_bits"],
                      "autotuned": tuned is not None, "parallelism": "1 GPU"},
           "stages_ms": {k: round(v, 4) for k, v in per.items()},
           "h2d_load_s": round(h2d_s, 4),
           "pcie_inclusive_gbases_s": round(g.seqtot / (ms_step * 1e-3 + h2d_s) / 1e9, 4),
           "roofline": roofline,
           "seed_sort": seed_sort}

    if not args.no_batch:
        t = time.perf_counter()
        gdbs = synth.batch(args.batch_pairs, scale=args.scale)
        tot = sum(x.seqtot for x in gdbs)
        log(f"[bench] batch workload generated: {len(gdbs)} pairs, {tot} bases ({time.perf_counter() - t:.1f}s)")
        bopts = dict(tile_records=args.tile, chunk_records=args.chunk, merge_tile=args.merge_tile,
                     digit2_bits=args.digit2)
        bctx = pga.Context(**bopts)
        bctx.load_batch(gdbs)
        if not args.no_tune:
            btuned, _ = autotune.tune(bctx, gdbs, use_cache=not args.retune, log=log)
            bopts.update(btuned)
        bsteps = max(2, args.steps // 2)
        bdt, bper, _, _, _ = time_steps(bctx, bsteps, 1)
        bctx.close()
        batched = tot * bsteps / bdt / 1e9
        # plain form: the same pairs one at a time (input resident, one build + merge per pair),
        # tuned on the median pair
        popts = dict(tile_records=args.tile, chunk_records=args.chunk, merge_tile=args.merge_tile,
                     digit2_bits=args.digit2)
        pctx = pga.Context(**popts)
        if not args.no_tune:
            med = sorted(gdbs, key=lambda x: x.seqtot)[len(gdbs) // 2]
            pctx.load_gdb(med)
            ptuned, _ = autotune.tune(pctx, [med], use_cache=not args.retune, log=log)
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
        out["batched"] = dict(workload=f"{len(gdbs)} independent pairs, A lognormal 1-12 Mbp, 1-50 contigs, "
                                       "GC 0.30-0.65, B at 1-15% SNP and 0.1-1% indel",
                              bases=tot, batched_gbases_s=round(batched, 4), batched_ms=round(bdt / bsteps * 1e3, 3),
                              plain_gbases_s=round(plain, 4), plain_ms=round(plain_s * 1e3, 3),
                              speedup=round(batched / plain, 3), batched_opts=bopts, plain_opts=popts,
                              batched_stages_ms={k: round(v, 3) for k, v in bper.items()})
        log(f"[bench] batched {batched:.2f} Gbases/s vs plain per pair {plain:.2f} Gbases/s")

    if not args.no_serve:
        out["serving"] = serving_leg(args)

    if args.human_scale > 0:
        out["human_scale"] = human_leg(args, opts)

    if not args.no_cpu:
        t = time.perf_counter()
        out["cpu_baseline"] = cpu_baseline(g, args.cpu_bases)
        log(f"[bench] cpu baseline {out['cpu_baseline']['value']} Gbases/s ({time.perf_counter() - t:.1f}s)")

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
