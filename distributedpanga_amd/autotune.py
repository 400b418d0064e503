"""Autotuner for the device path's launch shapes, cached per workload shape.

Tunables (all bit-exact-neutral: every setting gives the same output bytes):
  tile_records   records per sort workgroup (2K / 4K / 8K): also the bucket size
                 at which the radix partition stops, so it sets the pass count
  chunk_records  records per radix-scatter workgroup (2K / 4K / 8K): the run
                 length per digit, i.e. write coalescing, against occupancy
  merge_tile     anchors per merge workgroup (256 / 512 / 1024)
  digit2_bits    width of the partition's second digit (8 / 9 / 10 / 11): wider
                 digits split the second level finer, so with smaller tiles
                 fewer records need a third pass

The search is exhaustive (108 points, a couple of steps each, on the loaded
batch) and the winner is cached in a JSON file keyed by the batch shape.
"""
import itertools
import json
import math
import os
import time

SPACE = dict(tile_records=(2048, 4096, 8192), chunk_records=(2048, 4096, 8192), merge_tile=(256, 512, 1024),
             digit2_bits=(8, 9, 10, 11))


def cache_path():
    return os.environ.get("PGA_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "distributedpanga_amd", "tune.json"))


def shape_key(gdbs):
    """npairs, log2 of total bases, log2 of the largest pair: shapes that tune alike."""
    tot = sum(g.seqtot for g in gdbs)
    big = max(g.seqtot for g in gdbs)
    return f"p{len(gdbs)}-t{int(math.log2(max(tot, 1)))}-m{int(math.log2(max(big, 1)))}"


def _load_cache():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(d):
    p = cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        with open(p + ".tmp", "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(p + ".tmp", p)
    except OSError:
        pass


def _time(ctx, steps):
    ctx.build_index()
    ctx.merge()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.build_index()
        ctx.merge()
    ctx.sync()
    return (time.perf_counter() - t0) / steps


def tune(ctx, gdbs, steps=2, use_cache=True, log=None):
    """Pick the fastest launch shape for the batch already loaded in `ctx`
    (a pga.Context); leaves ctx set to it and returns (opts dict, table)."""
    key = shape_key(gdbs)
    cache = _load_cache() if use_cache else {}
    if use_cache and key in cache:
        best = cache[key]["opts"]
        ctx.set_opts(**best)
        return best, cache[key].get("table", [])
    table = []
    for point in itertools.product(*SPACE.values()):
        o = dict(zip(SPACE, point))
        ctx.set_opts(**o)
        dt = _time(ctx, steps)
        table.append(dict(o, ms=round(dt * 1e3, 3)))
        if log:
            log(f"[tune] {key} tile={o['tile_records']} chunk={o['chunk_records']} merge={o['merge_tile']} "
                f"digit2={o['digit2_bits']}: {dt * 1e3:.2f} ms")
    best = min(table, key=lambda r: r["ms"])
    opts = {k: best[k] for k in SPACE}
    ctx.set_opts(**opts)
    if use_cache:
        cache[key] = dict(opts=opts, table=table)
        _store_cache(cache)
    return opts, table
