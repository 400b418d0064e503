/* libpga_hip.h -- C-ABI of the MI355X index build + adaptamer seed merge.
 *
 * Plain C, plain pointers and sizes; every call returns an int status
 * (PGA_OK or a negative PGA_E* code) and never exits the process.
 *
 * One context = one HIP stream on one device plus a grow-only pool of device
 * buffers.  A context holds ONE batch at a time: a batch is one or more
 * independent GDBs ("pairs": genome A's contigs then genome B's), each indexed
 * and self-merged on its own, in the same kernel launches.  The plain path is a
 * batch of one.  Calls on one context must not overlap; contexts on different
 * host threads are independent.
 *
 * What each entry point replaces in the reference (SakshamDiwan/DistributedPanGA):
 *   pga_load_gdb / pga_load_batch   the GDB the driver reads, lib/GDB.c:1171 (Read_GDB)
 *                                   and :1831-1930 (Get_Contig_Piece) -- `.bps` bytes in HBM
 *   pga_build_index                 extract_to_records   src/extract.h:40
 *                                   run_stage2           src/dsort.h:82 (sort + LCP, 1 rank)
 *                                   build_kmer_adapter   src/kmer_adapter.h:46
 *   pga_index_info / _export        KmerStreamAdapter + PayloadLayout fields,
 *                                   src/kmer_adapter.h:16-44 (entries, prefix counts, maxpre)
 *   pga_merge                       self_adaptamer_merge src/fastga_pipeline.h:18
 *   pga_seeds_info / _export        Recv_N_Data/Recv_C_Data + N/C_Units[0].buck,
 *                                   src/pga-mpi-merge.c:523-547 (1-rank shortcut)
 *   pga_seed_sort / _export         reimport_thread + rmsd_sort as pair_sort_search
 *                                   drives them, src/fastga_pipeline.c:1391-1450, 1711-1822
 *   pga_index_verify                verify_per_rank_sorted, src/dsort.c:229-246
 *   pga_bucket_counts               bucket_local_counts, src/dsort.c:108-118
 *   pga_destination_map             compute_destination_map, src/dsort.c:120-167
 */
#ifndef LIBPGA_HIP_H
#define LIBPGA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PGA_OK          0
#define PGA_ENOMEM     -1   /* device or host allocation failed          */
#define PGA_EHIP       -2   /* a HIP runtime call failed                 */
#define PGA_EINVAL     -3   /* bad argument or call order                */
#define PGA_EOVERFLOW  -4   /* a field does not fit its byte width       */

#define PGA_NUM_PREFIXES (1 << 24)   /* 24-bit k-mer prefix table length */

typedef struct pga_ctx pga_ctx;

typedef struct {
    int device;          /* HIP device ordinal                                  */
    int freq;            /* frequency cut: skip groups of >= freq entries (10)  */
    int tile_records;    /* records per sort workgroup: 2048, 4096 or 8192      */
    int chunk_records;   /* records per radix-scatter workgroup: 2048..8192     */
    int merge_tile;      /* entries per merge workgroup: 256, 512 or 1024       */
    int build_cum;       /* 1: build the 2^24 prefix table (GIX layout)         */
    int digit2_bits;     /* width of the radix partition's second digit: 8 to 11
                            (the others are 8); wider digits give smaller
                            buckets, so fewer records need a third level      */
} pga_opts;

typedef struct {
    int64_t nrec;          /* entries (== records) of the index                    */
    int     P, C;          /* position / contig+strand bytes (src/record.c:18-36)   */
    int     e;             /* entry bytes = 9+P+C (src/kmer_adapter.c:16-20)        */
    int     s;             /* seed bytes  = 1+2(P+C) (src/pga-mpi-merge.c:68-76)    */
    int64_t maxpre;        /* largest 24-bit prefix group                           */
    int     ncontig;
    int64_t seqtot;
} pga_index_meta;

typedef struct {
    int64_t nlen, clen;    /* bytes in the same-strand (N) / cross-strand (C) stream */
    int     s;             /* seed bytes                                             */
    int     ncontig;       /* length of each buck vector                             */
} pga_seeds_meta;

/* Milliseconds of device time per stage of the last build / merge
 * (HIP events on the context's stream). */
typedef struct {
    double extract_ms;     /* syncmer scan: count + emit                      */
    double partition_ms;   /* MSD radix partition passes                      */
    double tile_sort_ms;   /* LDS sort of final buckets + LCP + GIX entries   */
    double table_ms;       /* boundary LCP fix + 2^24 prefix table + maxpre   */
    double merge_ms;       /* adaptamer merge: count + scan + emit            */
    double total_ms;       /* first event to last event, wall on the device   */
    int    partition_levels;
    int    partition_launches;
    int64_t tiles;
    /* the radix partition scatter (dominant HBM kernel): summed kernel time
       and bytes moved (records read + written, and the next level's digit
       byte written beside each record bound for a next-level segment) */
    double scatter_ms;
    int64_t scatter_bytes;
    int    scatter_launches;
    /* the seed emission kernel of the merge alone (its own event pair) */
    double merge_emit_ms;
    /* tiles with bins of repeated k-mers too crowded to rank record by
       record: sorted by sub-binning those bins (crowded) or, when that does
       not resolve them, by the merge-sort path (fallback) */
    int64_t fallback_tiles;
    int64_t crowded_tiles;
    /* pga_run_graph calls that had to redo the work on the eager path (a
       partition deeper than the captured levels, or seeds past the buffers) */
    int64_t graph_fallbacks;
} pga_timings;

void        pga_default_opts(pga_opts *o);
int         pga_create(const pga_opts *o, pga_ctx **out);
/* Change the tunables of a context (the device stays); buffers are reused.
 * Takes effect from the next pga_build_index / pga_merge. */
int         pga_set_opts(pga_ctx *ctx, const pga_opts *o);
void        pga_destroy(pga_ctx *ctx);
const char *pga_last_error(const pga_ctx *ctx);

/* One GDB from host memory: `.bps` bytes (LSB-first 2-bit, contigs byte-aligned
 * at boff[i]) and contig lengths.  Copied to the device. */
int pga_load_gdb(pga_ctx *ctx, const uint8_t *bps, int64_t bps_bytes,
                 const int64_t *boff, const int64_t *clen, int ncontig);

/* A batch of npairs GDBs (arrays of per-pair pointers). */
int pga_load_batch(pga_ctx *ctx, int npairs, const uint8_t *const *bps, const int64_t *bps_bytes,
                   const int64_t *const *boff, const int64_t *const *clen, const int *ncontig);

/* Index every GDB of the loaded batch on the device. */
int pga_build_index(pga_ctx *ctx);

int pga_index_info(pga_ctx *ctx, int pair, pga_index_meta *info);

/* Copy pair's GIX entries (nrec*e bytes) and, if cum != NULL, the cumulative
 * prefix counts (PGA_NUM_PREFIXES int64: entries with prefix <= p) to the host.
 * Byte-identical to build_kmer_adapter's entries / cumulative_count_by_prefix
 * (src/kmer_adapter.c:55-80) under the canonical tie order (equal k-mers in
 * (contig, offset, forward-before-reverse) order). */
int pga_index_export(pga_ctx *ctx, int pair, uint8_t *entries, int64_t *cum);

/* Index only the records whose 5-base bucket (record_bucket, src/dsort.h:30-33:
 * the first 10 k-mer bits) lies in [bucket_lo, bucket_hi), from the next build
 * on; [0, 1024) (the default) keeps all.  This is one shard of a multi-GPU
 * index built from the whole (replicated) `.bps` image: with the ranges of
 * pga_destination_map, the shards' entries, prefix counts and seed streams,
 * concatenated in bucket order, are the whole index's (prefix groups never
 * cross a bucket), except the LCP byte of each shard's first entry, which a
 * shard computes without its predecessor (0) -- the seam the reference fixes
 * across ranks (src/dsort.c:507-516). */
int pga_set_shard(pga_ctx *ctx, int bucket_lo, int bucket_hi);

/* Check pair's index on the device (verify_per_rank_sorted, src/dsort.c:229-246):
 * *bad = entries whose k-mer is smaller than the previous entry's, *first_bad =
 * the first such entry (-1 if none), *lcp_bad = entries whose LCP byte differs
 * from the LCP recomputed from the k-mers (0 for the first entry). */
int pga_index_verify(pga_ctx *ctx, int pair, int64_t *bad, int64_t *first_bad, int64_t *lcp_bad);

/* Test support, no counterpart in the reference: overwrite byte `byte` of GIX
 * entry `entry` of pair on the device (e.g. to show that pga_index_verify
 * flags a broken order or LCP).  Leaves the prefix table as it was. */
int pga_index_poke(pga_ctx *ctx, int pair, int64_t entry, int byte, int value);

/* Records of pair per 5-base bucket (the first 10 k-mer bits; record_bucket,
 * src/dsort.h:30-33), counts[1024]: bucket_local_counts, src/dsort.c:108-118,
 * read off the prefix table (needs opts.build_cum). */
int pga_bucket_counts(pga_ctx *ctx, int pair, int64_t *counts);

/* Host-only: contiguous bucket ranges per shard from summed bucket counts
 * (compute_destination_map, src/dsort.c:120-167): ksplit[world_size+1] bucket
 * boundaries, select[1024] shard of each bucket. */
int pga_destination_map(const int64_t *global_counts, int world_size, int *ksplit, int *select);

/* Self-merge every indexed GDB of the batch (frequency cut = opts.freq). */
int pga_merge(pga_ctx *ctx);

int pga_seeds_info(pga_ctx *ctx, int pair, pga_seeds_meta *info);

/* Copy pair's seed streams and per-anchor-contig counts to the host.  N/C are
 * the streams new_self_merge_thread writes (src/fastga_pipeline.c:357-404):
 * anchors in index order, partners ascending, 1+2(P+C) bytes per seed. */
int pga_seeds_export(pga_ctx *ctx, int pair, uint8_t *nseeds, uint8_t *cseeds,
                     int64_t *nbuck, int64_t *cbuck);

/* Pieces of an index or a seed stream, for consumers that read a large index
 * in parts (the GIX .ktab writer of one part, a check of sampled groups):
 *   pga_index_export_range   entries [first, first+count) of pair
 *   pga_seeds_locate         the index, in seeds, of the first N and C seed of
 *                            anchor entry `anchor` (anchor == nrec: the
 *                            streams' lengths); an anchor's seeds, and those
 *                            of a run of anchors, are contiguous in each stream
 *   pga_seeds_export_range   seeds [first, first+count) of stream comp
 *                            (0: N, 1: C) of pair */
int pga_index_export_range(pga_ctx *ctx, int pair, int64_t first, int64_t count, uint8_t *entries);
int pga_seeds_locate(pga_ctx *ctx, int pair, int64_t anchor, int64_t *nfirst, int64_t *cfirst);
int pga_seeds_export_range(pga_ctx *ctx, int pair, int comp, int64_t first, int64_t count, uint8_t *out);

/* The seed array pair_sort_search sorts before its alignment search
 * (src/fastga_pipeline.c:1711-1822): every seed of one stream of pair (comp 0:
 * the N stream, 1: the C stream) reimported as a diagonal record
 * (reimport_thread, src/fastga_pipeline.c:1391-1450)
 *     [lcp][diag & 63][anti, DBYTE bytes LE][band = diag >> 6, DBYTE LE][jcont, C bytes LE]
 * (N: diag = M + i - j, anti = i + j; C: diag = 2M - (i + j), anti = M - (i - j);
 * M = the longest contig, DBYTE = bytes for 2M, src/pga-mpi-merge.c:415-425),
 * placed in anchor-contig panels (contig order, panel sizes = the buck vector)
 * and each panel sorted on the whole record read from its last byte (rmsd_sort,
 * lib/RSDsort.c:292-377).  *nels = records, *swide = 2*DBYTE + C + 2 bytes each.
 * Needs pga_merge; the records stay on the device until pga_seed_sort_export. */
int pga_seed_sort(pga_ctx *ctx, int pair, int comp, int64_t *nels, int *swide);
int pga_seed_sort_export(pga_ctx *ctx, uint8_t *records);

int pga_timings_get(pga_ctx *ctx, pga_timings *t);

/* ------------------------------------------------------------ no allocation
 *
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the self-merge of pga_merge + pga_seeds_export for one GIX
 * index that the caller already holds on the device, with no context and no
 * allocation.  Its intermediates (LCP/strand bytes, interval records, tile
 * counts and offsets) live in one static device array of
 * PGA_NOALLOC_SCRATCH_BYTES that each GPU the library runs on keeps until the
 * process ends, which limits an index to PGA_NOALLOC_MAX_ENTRIES entries
 * (PGA_ELIMIT beyond).  All calls on one GPU share that array: calls that
 * overlap on one GPU race (serialise them).
 *
 *   entries       device, nrec GIX entries of 9+P+C bytes (pga_index_export's
 *                 layout), in a buffer of entries_bytes >= nrec*(9+P+C) + 8
 *                 readable bytes (payloads are read 8 bytes at a time)
 *   nseeds/ncap   device N stream and its capacity in bytes (cseeds/ccap: C)
 *   nbuck, cbuck  device, ncontig int64 each: seeds per anchor contig
 *   nlen, clen    host: the streams' lengths in bytes
 *   stream        a hipStream_t (NULL: the null stream); returns once the
 *                 outputs are complete on the device
 * When a stream does not fit its capacity, PGA_EOVERFLOW is returned with
 * *nlen and *clen set to the sizes needed and nothing written to the streams. */
#define PGA_ELIMIT               -5        /* input beyond a fixed limit          */
#define PGA_NOALLOC_MAX_ENTRIES  (1 << 23)
#define PGA_NOALLOC_SCRATCH_BYTES (41u << 20)
int pga_merge_gix_noalloc(const uint8_t *entries, int64_t entries_bytes, int64_t nrec, int P, int C,
                          int ncontig, int freq, uint8_t *nseeds, int64_t ncap, uint8_t *cseeds, int64_t ccap,
                          int64_t *nbuck, int64_t *cbuck, int64_t *nlen, int64_t *clen, void *stream);

/* pga_build_index + pga_merge of the loaded batch as one device-driven
 * pipeline (no host round trip between kernels; every count stays on the
 * device) captured in a HIP graph on the first call after a load and replayed
 * on later calls.  Outputs and exports are those of the eager calls.  Buffers
 * are sized from load-time bounds (two records a candidate position), so it
 * takes more device memory than the eager path.  When a batch needs more than
 * 8 partition levels (long low-complexity runs) or its seeds outgrow the
 * buffers, the call redoes the work on the eager path (counted in
 * pga_timings.graph_fallbacks).  Only the total time is timed. */
int pga_run_graph(pga_ctx *ctx);

/* Wait for all work queued on the context's stream. */
int pga_sync(pga_ctx *ctx);

/* ------------------------------------------------------------------ serving
 *
 * A server keeps one target GDB indexed on the device and answers batches of
 * query GDBs against it.  A query's answer is exactly the part of the
 * reference's self-merge of one GDB holding the target's contigs followed by
 * the query's (src/pga-mpi-merge.c:509, src/fastga_pipeline.c:245-412) in which
 * the anchor or the partner lies on a query contig: same bytes, same order,
 * with the query's contig ids numbered after the target's, and P/C widths of
 * that combined GDB.  The buck vectors count those seeds per anchor contig
 * (target contigs first).  Each query is answered independently of the others
 * in its batch.  Limits: freq <= 24; target + query contigs <= 32767; contig
 * lengths < 2^31.
 *
 * A server has 1..8 slots.  A slot answers one batch at a time on its own HIP
 * stream; different slots may be driven from different host threads at once
 * (their batches overlap on the device, all reading the one resident target).
 * Calls on one slot must not overlap, and set_target must not overlap any run. */
typedef struct pga_server pga_server;

int         pga_server_create(const pga_opts *o, int nslots, pga_server **out);
void        pga_server_destroy(pga_server *s);
int         pga_server_slots(pga_server *s);
/* Message of the last failed call on `slot` (slot < 0: the target calls). */
const char *pga_server_error(const pga_server *s, int slot);

/* Index the target (same arguments as pga_load_gdb); it stays resident. */
int pga_server_set_target(pga_server *s, const uint8_t *bps, int64_t bps_bytes,
                          const int64_t *boff, const int64_t *clen, int ncontig);
/* The target from a persisted GIX index instead of its sequence (a restart of
 * the serving loop without re-indexing): nrec entries of 9+P+C bytes in the
 * adapter layout, the cumulative 2^24 prefix counts and the largest group (the
 * .ktab / .gix contents, notes/GIXmake_dump.c:1632-1709, as pga_gix_read_*
 * return them), and the target's contig lengths (P and C must be the widths
 * those lengths give).  Answers equal those of a target set from its
 * sequence. */
int pga_server_set_target_gix(pga_server *s, const uint8_t *entries, int64_t nrec, int P, int C,
                              const int64_t *cum, int64_t maxpre, const int64_t *clen, int ncontig);
int pga_server_target_info(pga_server *s, pga_index_meta *info);

/* Device bytes a query of query_bases bases may take while it is answered
 * (an upper estimate, for admission control). */
int pga_server_footprint(pga_server *s, int64_t query_bases, int64_t *bytes);

/* Answer nq queries on `slot` (arrays of per-query pointers, as pga_load_batch).
 * Returns when the answers are on the device. */
int pga_server_run(pga_server *s, int slot, int nq, const uint8_t *const *bps, const int64_t *bps_bytes,
                   const int64_t *const *boff, const int64_t *const *clen, const int *ncontig);

/* Sizes / copies of query q's answer from the slot's last run
 * (ncontig = target + query contigs = length of each buck vector). */
int pga_server_seeds_info(pga_server *s, int slot, int q, pga_seeds_meta *info);
int pga_server_seeds_export(pga_server *s, int slot, int q, uint8_t *nseeds, uint8_t *cseeds,
                            int64_t *nbuck, int64_t *cbuck);

/* Device milliseconds of the slot's last run (load, index, join). */
int pga_server_last_ms(pga_server *s, int slot, double *ms);

#ifdef __cplusplus
}
#endif
#endif
