/* pga_synth.h -- deterministic synthetic genome pairs and the in-memory GDB view.
 *
 * A genome pair is held the way the reference's driver sees one GDB that holds
 * genome A's contigs followed by genome B's (the reference self-merges one GDB;
 * src/pga-mpi-merge.c:314,509).  Sequence bytes follow the `.bps` layout: each
 * contig starts at a byte-aligned offset `boff` and base i of a contig sits in
 * bits 2*(i%4) of byte boff + i/4, A=0 C=1 G=2 T=3 (lib/gene_core.c:336-357,
 * lib/GDB.c:1831-1930).  Only these fields of a GDB are read by extraction
 * (tests/support/synth_gdb.h:8-33), so they are all this struct carries.
 *
 * Everything here is host C.  The generator draws from splitmix64 only, so the
 * same (preset, seed, scale) gives byte-identical contigs on every machine.
 */
#ifndef PGA_SYNTH_H
#define PGA_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int      ncontig;     /* contigs in the GDB: genome A's, then genome B's        */
    int      n_a;         /* how many of them belong to genome A                     */
    int64_t *clen;        /* [ncontig] contig lengths in bases                       */
    int64_t *boff;        /* [ncontig+1] byte offset of each contig in bps           */
    uint8_t *bps;         /* packed 2-bit bases, LSB-first, contigs byte-aligned     */
    int64_t  bps_bytes;   /* == boff[ncontig]                                        */
    int64_t  seqtot;      /* total bases                                             */
    int64_t  maxctg;      /* longest contig                                          */
} pga_gdb;

/* Parameters of one synthetic pair: genome A, then B derived from A. */
typedef struct {
    int64_t a_bases;        /* total bases of A                                        */
    int     a_contigs;      /* contigs of A                                            */
    double  len_sigma;      /* lognormal sigma of contig lengths (rescaled to a_bases) */
    int64_t len_min;        /* uniform contig lengths in [len_min,len_max] when > 0     */
    int64_t len_max;        /*   (used instead of the lognormal when len_max > 0)       */
    double  gc;             /* P(G)+P(C)                                               */
    /* interspersed repeats: rep_families families, rep_copies copies each (or
       rep_fraction of A when rep_copies == 0), element length U[rep_len_min,
       rep_len_max], per-copy divergence U[rep_div_min, rep_div_max]           */
    int     rep_families;
    int     rep_copies;
    double  rep_fraction;
    int     rep_len_min, rep_len_max;
    double  rep_div_min, rep_div_max;
    /* tandem arrays: unit U[tan_unit_min,max] repeated U[tan_copies_min,max] times */
    int     tandem_arrays;
    int     tan_unit_min, tan_unit_max;
    int     tan_copies_min, tan_copies_max;
    /* satellite arrays: sat_bases of sat_unit-bp units at sat_div divergence */
    int64_t sat_bases;
    int     sat_unit;
    double  sat_div;
    /* genome B */
    double  snp_rate;       /* per-base substitution probability                       */
    double  indel_rate;     /* per-base indel start probability (ins:del 1:1)          */
    double  indel_mean;     /* geometric mean length                                   */
    int     indel_max;      /* cap                                                     */
    int     inversions;     /* in-place reverse complements of U[inv_min,inv_max] bp   */
    int64_t inv_min, inv_max;
    int     moves;          /* segment swaps of U[inv_min,inv_max] bp inside a contig  */
} pga_pair_spec;

/* Named workloads:
 *   1  2 Mbp A (10 lognormal contigs, sigma 0.8, GC 0.41), B = A + 2% SNP + 0.2% indel
 *   2  100 Mbp A (500 contigs) + 2000 x 300 bp repeat at 10% + 50 tandem arrays;
 *      B = A + 5% SNP + 0.5% indel + 10 inversions of 10-500 kbp
 *   4  3.1 Gbp A (24 contigs of 50-250 Mbp), 45% repeats from 500 families,
 *      2 Mbp of 171 bp satellites; B = A + 2% SNP + 0.3% indel + 200 rearrangements
 * `scale` multiplies every length and copy count (1.0 = the full workload). */
int  pga_synth_preset(int workload, double scale, pga_pair_spec *spec, uint64_t *seed);

/* Generate one pair into *out (A contigs then B contigs).  0 or -1 (OOM / bad spec). */
int  pga_synth_pair(const pga_pair_spec *spec, uint64_t seed, pga_gdb *out);

/* The 64-pair batch workload: pair i gets A lognormal in [1,12] Mbp (median 4),
 * 1-50 contigs, GC U[0.30,0.65], B at SNP U[1,15]% and indel U[0.1,1]%.
 * Writes the per-pair specs and seeds (arrays of npairs). */
int  pga_synth_batch_specs(int npairs, uint64_t seed, double scale,
                           pga_pair_spec *specs, uint64_t *seeds);

/* A query genome for the serving workload: qbases taken from random segments
 * of the target (segments of U[seg_min, seg_max] bases, at random offsets of
 * random target contigs), each copied with per-base substitution probability
 * div and indels at div/10, half of them reverse-complemented, in shuffled
 * order.  Each segment is one contig of the query. */
int  pga_synth_query(const pga_gdb *target, uint64_t seed, int64_t qbases, double div,
                     int64_t seg_min, int64_t seg_max, pga_gdb *out);

/* Build a GDB from numeric bases (0..3, one per byte), contig after contig. */
int  pga_gdb_from_codes(const uint8_t *codes, const int64_t *clen, int ncontig,
                        int n_a, pga_gdb *out);

/* Unpack contig c into codes[0..clen) (0..3 per byte). */
void pga_gdb_contig_codes(const pga_gdb *g, int c, uint8_t *codes);

void pga_gdb_free(pga_gdb *g);

/* splitmix64 step, exported so tests can reproduce draws. */
uint64_t pga_splitmix64(uint64_t *state);

#ifdef __cplusplus
}
#endif
#endif
