/* ref_harness.c -- drives the reference's own extraction, sort, LCP and
 * GIX-adapter code (compiled from the reference sources by oracle/ref_build.sh
 * into oracle/_ref/) on an in-memory GDB, so the oracle restatement can be
 * compared with the real thing.  TEST INFRASTRUCTURE ONLY.
 *
 * The GDB is filled the way the reference's tests do it
 * (tests/support/synth_gdb.c:41-107): seqstate EXTERNAL over an fmemopen'd
 * `.bps` byte image, so Get_Contig_Piece takes the same path as with a file.
 * The merge (src/fastga_pipeline.c) is not driven: it needs src/seed_buffer.c,
 * which needs MPI, and MPI is not installed.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "GDB.h"
#include "record.h"
#include "extract.h"
#include "sort.h"
#include "kmer_adapter.h"

typedef struct {
    int      P, C, rsize, e;
    int64_t  nrec;
    uint8_t *extracted;   /* nrec*rsize, extraction order, as extract_to_records left them */
    uint8_t *sorted;      /* nrec*rsize after sort_records + recalc_all_lcps                */
    uint8_t *entries;     /* nrec*e from build_kmer_adapter                                 */
    int64_t *cum;         /* 1<<24                                                          */
    int64_t  maxpre;
    double   t_extract, t_sort, t_adapter;
} ref_result;

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double) ts.tv_sec + 1e-9 * (double) ts.tv_nsec;
}

int ref_run(const uint8_t *bps, int64_t bps_bytes, const int64_t *boff, const int64_t *clen,
            int ncontig, int do_index, ref_result *out)
{
    GDB g;
    RecordBuffer buf;
    double t0;
    memset(&g, 0, sizeof(g));
    memset(out, 0, sizeof(*out));
    g.contigs = (GDB_CONTIG *) calloc((size_t) ncontig + 1, sizeof(GDB_CONTIG));
    if (!g.contigs) return -1;
    for (int i = 0; i < ncontig; i++) {
        g.contigs[i].clen = clen[i];
        g.contigs[i].boff = boff[i];
        g.seqtot += clen[i];
        if (clen[i] > g.maxctg) g.maxctg = clen[i];
    }
    g.contigs[ncontig].boff = bps_bytes;
    g.ncontig = ncontig;
    g.seqstate = EXTERNAL;
    g.seqs = fmemopen((void *) bps, (size_t) (bps_bytes > 0 ? bps_bytes : 1), "rb");
    if (!g.seqs) { free(g.contigs); return -1; }

    RecordSizing rs = compute_record_sizing(&g);
    out->P = rs.position_bytes;
    out->C = rs.contig_strand_bytes;
    out->rsize = rs.record_size;
    out->e = 9 + rs.position_bytes + rs.contig_strand_bytes;

    record_buffer_init(&buf, rs);
    t0 = now_s();
    extract_to_records(&g, 1, 0, &buf);
    out->t_extract = now_s() - t0;
    out->nrec = buf.count;
    size_t bytes = (size_t) (buf.count * rs.record_size);
    out->extracted = (uint8_t *) malloc(bytes + 1);
    if (!out->extracted) return -1;
    memcpy(out->extracted, buf.data, bytes);

    if (do_index) {
        KmerStreamAdapter T;
        PayloadLayout L;
        t0 = now_s();
        sort_records(&buf);
        recalc_all_lcps(&buf);
        out->t_sort = now_s() - t0;
        out->sorted = (uint8_t *) malloc(bytes + 1);
        if (!out->sorted) return -1;
        memcpy(out->sorted, buf.data, bytes);
        memset(&T, 0, sizeof(T));
        memset(&L, 0, sizeof(L));
        t0 = now_s();
        if (build_kmer_adapter(&buf, &T, &L) < 0) return -1;
        out->t_adapter = now_s() - t0;
        out->entries = (uint8_t *) malloc((size_t) (T.num_entries * T.entry_bytes) + 1);
        out->cum = (int64_t *) malloc(sizeof(int64_t) * (size_t) T.num_prefixes);
        if (!out->entries || !out->cum) return -1;
        memcpy(out->entries, T.entries, (size_t) (T.num_entries * T.entry_bytes));
        memcpy(out->cum, T.cumulative_count_by_prefix, sizeof(int64_t) * (size_t) T.num_prefixes);
        out->maxpre = L.largest_prefix_group_num_entries;
        free_kmer_adapter(&T, &L);
    }
    record_buffer_free(&buf);
    fclose((FILE *) g.seqs);
    free(g.contigs);
    return 0;
}

void ref_result_free(ref_result *r)
{
    free(r->extracted); free(r->sorted); free(r->entries); free(r->cum);
    memset(r, 0, sizeof(*r));
}

/* The reference's packers and LCP on raw bytes, for the known-answer tests. */
int ref_lcp_bases(const uint8_t *a, const uint8_t *b) { return simple_kmer_lcp_bases(a, b, 10); }


/* The reference's seed sort (lib/RSDsort.c), declared the way its own caller
 * declares it (src/fastga_pipeline.c:27-28, 124-128), run on records the
 * oracle reimported.  panel_bytes[c] = bytes of contig c's panel. */
typedef struct {
    int   beg;
    int   end;
    int64 off;
} Range;

extern int rmsd_sort(uint8 *array, int64 nelem, int rsize, int ksize, int nparts, int64 *part,
                     int nthreads, Range *parms);

int ref_rmsd_sort(uint8_t *arr, int64_t nelem, int swide, int nparts, int64_t *panel_bytes)
{
    Range r[1];
    if (nelem <= 0) return 0;
    return rmsd_sort(arr, nelem, swide, swide, nparts, panel_bytes, 1, r);
}
