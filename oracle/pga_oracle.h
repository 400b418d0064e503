/* pga_oracle.h -- CPU restatement of the reference's index-build + adaptamer
 * merge path.  TEST INFRASTRUCTURE ONLY: the parity tests, the smoke check and
 * the CPU-baseline leg of the benchmark call it as the checker.  No product
 * code links it.
 *
 * Each function names the reference file:line it restates.  Byte layouts are
 * the reference's host layouts:
 *   record (rsize = 12+P+C, src/record.h:8-30):
 *     [0] LCP  [1..10] k-mer MSB-first  [11] mask=0  [12..12+P) pos LE
 *     [12+P..rsize) contig LE, strand in bit 8C-1
 *   GIX entry (e = 9+P+C, src/kmer_adapter.c:55-62):
 *     [0..7) k-mer bytes 3..9  [7] mask=0  [8] LCP  [9..e) pos LE, contig|strand LE
 *   seed (s = 1+2(P+C), src/pga-mpi-merge.c:68-76):
 *     [0] plen  [1..1+P+C) payload(i), strand bit cleared  [..s) payload(p)
 *
 * How each function is pinned (tests/test_oracle.py, tests/test_golden.py):
 *   record_sizing     R-03 width table and E-13 (tests/CATALOG.md:415-423)
 *   pack_forward/_rc  E-01, E-02 closed forms; RC == reverse-complement string
 *   extract           E-06 known answer (3,106 records, right/left/tie 762/801/2),
 *                     naive == rolling over many lengths, E-08/E-09/E-10, R-03;
 *                     byte-equal to the compiled reference (oracle/_ref) at two
 *                     scales and to the reference-written fixtures
 *   sort_records      == qsort brute force with the full comparator, conservation;
 *                     == the compiled reference after canonicalising equal k-mers
 *   lcp_bases/recalc  S-01 hand table, S-02, R-04; == compiled reference
 *   build_adapter     K-01..K-03; entries, prefix table and maxpre == the
 *                     compiled reference and the fixtures
 *   merge             closed forms on hand-built groups (plen, strand split,
 *                     strand-bit clearing, FREQ cut at n = FREQ-1/FREQ/FREQ+1,
 *                     no crossing of 24-bit groups) and two independent
 *                     formulations (the reference's walk, mode 0, and the
 *                     LCP-interval definition, mode 1) equal on random and
 *                     workload inputs.  The compiled reference merge is not
 *                     available here (its translation units need MPI), so no
 *                     reference-run vector pins it.
 *   seed_sort         reimport closed forms on hand-built seeds (both streams);
 *                     the per-panel sort == the compiled reference rmsd_sort
 *                     (lib/RSDsort.c) run on the same reimported records.
 */
#ifndef PGA_ORACLE_H
#define PGA_ORACLE_H

#include <stdint.h>

typedef struct {
    int64_t n_sel;      /* selected 12-mers                                   */
    int64_t n_right;    /* index 4 is the strict unique minimum               */
    int64_t n_left;     /* earliest argmin is index 0                          */
    int64_t n_tie;      /* index 4 ties the minimum without being the earliest */
} oracle_extract_stats;

typedef struct {
    int      P, C, rsize, e, s;
    int64_t  nrec;
    uint8_t *records;          /* nrec*rsize, sorted, semantic LCPs            */
    uint8_t *entries;          /* nrec*e                                        */
    int64_t *cum;              /* [1<<24], count of entries with prefix <= p    */
    int64_t  maxpre;           /* largest 24-bit prefix group                   */
    uint8_t *nseeds;           /* same-strand seed stream                       */
    int64_t  nlen;             /* bytes                                         */
    uint8_t *cseeds;           /* cross-strand seed stream                      */
    int64_t  clen;             /* bytes                                         */
    int64_t *nbuck, *cbuck;    /* [ncontig] seeds per anchor contig             */
    int      ncontig;
    oracle_extract_stats xs;
    double   t_extract, t_sort, t_adapter, t_merge;   /* seconds                */
} oracle_result;

/* src/record.c:6-36 */
void oracle_record_sizing(int64_t maxctg, int64_t ncontig, int *P, int *C, int *rsize);

/* Packing helpers on numeric bases (one 0..3 per byte): src/pack.c:29-39, 64-78. */
void oracle_pack_forward(const uint8_t *bases40, uint8_t *out10);
void oracle_pack_rc(const uint8_t *bases40, uint8_t *out10);   /* bases of [j-28, j+12) */
void oracle_comp_table(int comp[256]);                          /* src/tables.c:24-33 */
int  oracle_tmap(int i);                                        /* src/tables.c:5-22  */

/* Canonical 8-mer hash at offset i of numeric bases (tests/support/ref_syncmer.c:151-158). */
int  oracle_mz(const uint8_t *codes, int64_t i);

/* Extraction of one GDB into host records, in extraction order.
 * mode 0: the naive per-offset definition (tests/support/ref_syncmer.c:196-227);
 * mode 1: the rolling scan (src/extract.c:85-249).
 * *out is malloc'd (nrec*rsize bytes); returns 0 or -1. */
int  oracle_extract(const uint8_t *bps, const int64_t *boff, const int64_t *clen, int ncontig,
                    int mode, int P, int C, uint8_t **out, int64_t *nrec,
                    oracle_extract_stats *st);

/* Sort records by the 10 k-mer bytes, ties in extraction order (contig, j, strand).
 * Input must be in extraction order (the stable passes keep it). */
int  oracle_sort_records(uint8_t *rec, int64_t n, int rsize);
/* Brute force: qsort with the full (k-mer, contig, j, strand) comparator. */
int  oracle_sort_records_qsort(uint8_t *rec, int64_t n, int rsize, int P, int C);

/* src/sort.c:16-32 */
int  oracle_lcp_bases(const uint8_t *a, const uint8_t *b);
/* src/sort.c:38-57 */
void oracle_recalc_lcps(uint8_t *rec, int64_t n, int rsize);

/* src/kmer_adapter.c:6-96.  entries: n*e bytes; cum: 1<<24 int64. */
int  oracle_build_adapter(const uint8_t *rec, int64_t n, int P, int C,
                          uint8_t *entries, int64_t *cum, int64_t *maxpre);

/* Adaptamer self-merge over GIX entries.
 * mode 0: the walk of src/fastga_pipeline.c:245-412 (plen / eorun / vlcp state);
 * mode 1: the LCP-interval definition, written independently.
 * Streams are malloc'd; nbuck/cbuck are caller arrays of ncontig. */
int  oracle_merge(const uint8_t *entries, int64_t n, int P, int C, int ncontig,
                  int freq, int mode, uint8_t **nseeds, int64_t *nlen,
                  uint8_t **cseeds, int64_t *clen, int64_t *nbuck, int64_t *cbuck);

/* Diagonal records of one seed stream, sorted per anchor-contig panel: the
 * reimport of src/fastga_pipeline.c:1391-1450 (comp 0 for the N stream,
 * comp 1 for the C stream) with AMXPOS = BMXPOS = maxctg, MAXDAG = 2*maxctg
 * and DBYTE from src/pga-mpi-merge.c:415-425, then each panel sorted on the
 * whole record read from its last byte (rmsd_sort, lib/RSDsort.c:292-377, as
 * pair_sort_search drives it, src/fastga_pipeline.c:1779-1822).  Records:
 * [lcp][diag&63][anti DBYTE LE][band DBYTE LE][jcont C LE], swide = 2*DBYTE+C+2.
 * *out is malloc'd (nseeds*swide bytes, panels in contig order). */
int  oracle_dbyte(int64_t maxctg);
int  oracle_seed_sort(const uint8_t *seeds, int64_t nseeds, int P, int C, int ncontig, int64_t maxctg,
                      int comp, int sort, uint8_t **out, int *swide);

/* Whole path: extract (rolling) -> sort -> LCP -> adapter -> merge. */
int  oracle_pipeline(const uint8_t *bps, const int64_t *boff, const int64_t *clen, int ncontig,
                     int freq, oracle_result *r);
void oracle_result_free(oracle_result *r);

void oracle_free(void *p);

#endif
