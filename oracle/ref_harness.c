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

/* The reference's own GDB skeleton writer and reader (Write_GDB
 * lib/GDB.c:1519-1607, Read_GDB lib/GDB.c:1171-1400), for pinning the
 * package's `.gdb` text reader/writer.  path ends in ".gdb" (ONE text) or
 * ".1gdb" (ONE binary); Write_GDB needs dir/.<root>.bps to exist.  Scaffold s
 * holds contigs [fctg[s], fctg[s+1]); sbeg/clen give each contig's place. */
int ref_gdb_write(const char *path, int nscaff, const char *headers, const int64_t *hoff,
                  const int64_t *slen, const int *fctg, int ncontig, const int64_t *sbeg,
                  const int64_t *clen, const float *freq)
{
    GDB g;
    int64_t boff = 0, hdrtot = 0;
    memset(&g, 0, sizeof(g));
    if (Prog_Name == NULL) Prog_Name = strdup("ref_harness");
    if (Command_Line == NULL) Command_Line = strdup("ref_harness");
    for (int s = 0; s < nscaff; s++) hdrtot = hoff[s] + (int64_t) strlen(headers + hoff[s]) + 1;
    g.nscaff = nscaff;
    g.scaffolds = (GDB_SCAFFOLD *) calloc((size_t) nscaff + 1, sizeof(GDB_SCAFFOLD));
    g.contigs = (GDB_CONTIG *) calloc((size_t) ncontig + 1, sizeof(GDB_CONTIG));
    g.headers = (char *) malloc((size_t) hdrtot + 1);
    if (!g.scaffolds || !g.contigs || !g.headers) return -1;
    memcpy(g.headers, headers, (size_t) hdrtot);
    g.hdrtot = (int) hdrtot;
    for (int s = 0; s < nscaff; s++) {
        g.scaffolds[s].slen = slen[s];
        g.scaffolds[s].fctg = fctg[s];
        g.scaffolds[s].ectg = fctg[s + 1];
        g.scaffolds[s].hoff = hoff[s];
        for (int c = fctg[s]; c < fctg[s + 1]; c++) g.contigs[c].scaf = s;
    }
    for (int c = 0; c < ncontig; c++) {
        g.contigs[c].clen = clen[c];
        g.contigs[c].sbeg = sbeg[c];
        g.contigs[c].boff = boff;
        boff += COMPRESSED_LEN(clen[c]);
        g.seqtot += clen[c];
        if (clen[c] > g.maxctg) g.maxctg = clen[c];
    }
    g.contigs[ncontig].boff = boff;
    g.ncontig = ncontig;
    g.seqstate = EXTERNAL;
    g.srcpath = strdup("synthetic.fa");
    for (int i = 0; i < 4; i++) g.freq[i] = freq[i];
    char *p = strdup(path);
    const int rc = Write_GDB(&g, p);
    free(p);
    free(g.srcpath);
    free(g.headers);
    free(g.contigs);
    free(g.scaffolds);
    return rc;
}

/* Read_GDB on path; copies out what fits (nscaff/ncontig report the totals). */
int ref_gdb_read(const char *path, int *nscaff, int *ncontig, int64_t *slen, int *fctg, int64_t *sbeg,
                 int64_t *clen, int64_t *boff, char *headers, int64_t hcap, float *freq, int cap_s, int cap_c)
{
    GDB g;
    char *p = strdup(path);
    memset(&g, 0, sizeof(g));
    if (Read_GDB(&g, p)) return -1;
    *nscaff = g.nscaff;
    *ncontig = g.ncontig;
    for (int s = 0; s < g.nscaff && s < cap_s; s++) {
        slen[s] = g.scaffolds[s].slen;
        fctg[s] = g.scaffolds[s].fctg;
    }
    if (g.nscaff < cap_s) fctg[g.nscaff] = g.ncontig;
    for (int c = 0; c < g.ncontig && c < cap_c; c++) {
        sbeg[c] = g.contigs[c].sbeg;
        clen[c] = g.contigs[c].clen;
        boff[c] = g.contigs[c].boff;
    }
    if (g.ncontig < cap_c) boff[g.ncontig] = g.contigs[g.ncontig].boff;
    if (headers && hcap > 0) {
        const int64_t n = g.hdrtot < hcap ? g.hdrtot : hcap;
        memcpy(headers, g.headers, (size_t) n);
    }
    for (int i = 0; i < 4; i++) freq[i] = g.freq[i];
    Close_GDB(&g);
    free(p);
    return 0;
}
