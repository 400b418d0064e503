/* pga_synth.c -- deterministic synthetic genome pairs (host C).
 *
 * Genome A is i.i.d. bases at a given GC content, cut into contigs, then
 * overwritten in place with interspersed repeat copies, tandem arrays and
 * satellite arrays.  Genome B is a copy of A with rearrangements (in-place
 * reverse complements and segment swaps) followed by substitutions and indels.
 * The pair is packed into one GDB, A's contigs first, in the `.bps` byte layout
 * (lib/gene_core.c:336-357: base i of a contig in bits 2*(i%4) of byte i/4).
 *
 * Soft-masked (lowercase) runs are not generated: the 2-bit layout carries no
 * case, and the record's mask byte is always 0 (src/extract.c:368).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "pga_synth.h"

uint64_t pga_splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double unif(uint64_t *s) { return (double) (pga_splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

static int64_t unif_int(uint64_t *s, int64_t lo, int64_t hi)   /* inclusive */
{
    if (hi <= lo) return lo;
    return lo + (int64_t) (pga_splitmix64(s) % (uint64_t) (hi - lo + 1));
}

static double normal(uint64_t *s)
{
    double u1 = unif(s), u2 = unif(s);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* One base at GC content gc: C and G each gc/2, A and T each (1-gc)/2. */
static uint8_t draw_base(uint64_t *s, double gc)
{
    double u = unif(s);
    if (u < gc * 0.5) return 1;
    if (u < gc)       return 2;
    if (u < gc + (1.0 - gc) * 0.5) return 0;
    return 3;
}

static uint8_t other_base(uint64_t *s, uint8_t b)
{
    return (uint8_t) ((b + 1 + (pga_splitmix64(s) % 3)) & 3);
}

static void fill_random(uint64_t *s, uint8_t *dst, int64_t n, double gc)
{
    for (int64_t i = 0; i < n; i++) dst[i] = draw_base(s, gc);
}

/* Copy src[0..n) to dst with each base substituted with probability div. */
static void diverged_copy(uint64_t *s, uint8_t *dst, const uint8_t *src, int64_t n, double div)
{
    for (int64_t i = 0; i < n; i++)
        dst[i] = (unif(s) < div) ? other_base(s, src[i]) : src[i];
}

static void revcomp_inplace(uint8_t *a, int64_t n)
{
    for (int64_t i = 0, j = n - 1; i <= j; i++, j--) {
        uint8_t x = (uint8_t) (3 - a[i]), y = (uint8_t) (3 - a[j]);
        a[i] = y; a[j] = x;
    }
}

typedef struct { uint8_t **seq; int64_t *len; int n; int64_t total; } contig_set;

/* A random (contig, offset) for a piece of length n, weighted by contig length. */
static int place(uint64_t *s, const contig_set *cs, int64_t n, int *c, int64_t *off)
{
    for (int tries = 0; tries < 64; tries++) {
        int64_t x = unif_int(s, 0, cs->total - 1);
        int k = 0;
        while (k < cs->n - 1 && x >= cs->len[k]) { x -= cs->len[k]; k++; }
        if (cs->len[k] < n) continue;
        *c = k;
        *off = unif_int(s, 0, cs->len[k] - n);
        return 1;
    }
    return 0;
}

static void contig_lengths(uint64_t *s, const pga_pair_spec *sp, int64_t *len)
{
    int n = sp->a_contigs;
    double *w = (double *) malloc((size_t) n * sizeof(double)), tot = 0.0;
    for (int i = 0; i < n; i++) {
        w[i] = (sp->len_max > 0) ? (double) unif_int(s, sp->len_min, sp->len_max)
                                 : exp(sp->len_sigma * normal(s));
        tot += w[i];
    }
    int64_t used = 0;
    for (int i = 0; i < n; i++) {
        int64_t l = (int64_t) floor(w[i] / tot * (double) sp->a_bases);
        if (l < 64) l = 64;
        len[i] = l;
        used += l;
    }
    /* put the rounding remainder on the longest contig */
    int big = 0;
    for (int i = 1; i < n; i++) if (len[i] > len[big]) big = i;
    len[big] += sp->a_bases - used;
    if (len[big] < 64) len[big] = 64;
    free(w);
}

static int make_a(uint64_t *s, const pga_pair_spec *sp, contig_set *A)
{
    A->n = sp->a_contigs;
    A->len = (int64_t *) calloc((size_t) A->n, sizeof(int64_t));
    A->seq = (uint8_t **) calloc((size_t) A->n, sizeof(uint8_t *));
    if (!A->len || !A->seq) return -1;
    contig_lengths(s, sp, A->len);
    A->total = 0;
    for (int i = 0; i < A->n; i++) {
        A->seq[i] = (uint8_t *) malloc((size_t) A->len[i]);
        if (!A->seq[i]) return -1;
        fill_random(s, A->seq[i], A->len[i], sp->gc);
        A->total += A->len[i];
    }

    /* interspersed repeat families */
    if (sp->rep_families > 0) {
        int maxlen = sp->rep_len_max > 0 ? sp->rep_len_max : 1;
        uint8_t *elem = (uint8_t *) malloc((size_t) maxlen);
        if (!elem) return -1;
        int64_t budget = (int64_t) (sp->rep_fraction * (double) A->total);
        int64_t per_family = sp->rep_families ? budget / sp->rep_families : 0;
        for (int f = 0; f < sp->rep_families; f++) {
            int el = (int) unif_int(s, sp->rep_len_min, sp->rep_len_max);
            double dv = sp->rep_div_min + (sp->rep_div_max - sp->rep_div_min) * unif(s);
            fill_random(s, elem, el, sp->gc);
            int64_t copies = sp->rep_copies > 0 ? sp->rep_copies : (per_family / el);
            for (int64_t k = 0; k < copies; k++) {
                int c; int64_t off;
                if (!place(s, A, el, &c, &off)) break;
                diverged_copy(s, A->seq[c] + off, elem, el, dv);
                if (unif(s) < 0.5) revcomp_inplace(A->seq[c] + off, el);
            }
        }
        free(elem);
    }

    /* tandem arrays */
    for (int t = 0; t < sp->tandem_arrays; t++) {
        int unit = (int) unif_int(s, sp->tan_unit_min, sp->tan_unit_max);
        int copies = (int) unif_int(s, sp->tan_copies_min, sp->tan_copies_max);
        int64_t n = (int64_t) unit * copies;
        uint8_t u[64];
        int c; int64_t off;
        if (unit > 64) unit = 64;
        fill_random(s, u, unit, sp->gc);
        if (!place(s, A, n, &c, &off)) continue;
        for (int64_t i = 0; i < n; i++) A->seq[c][off + i] = u[i % unit];
    }

    /* satellite arrays: arrays of 20-200 kbp built from one unit, lightly diverged */
    if (sp->sat_bases > 0 && sp->sat_unit > 0) {
        uint8_t *u = (uint8_t *) malloc((size_t) sp->sat_unit);
        if (!u) return -1;
        fill_random(s, u, sp->sat_unit, sp->gc);
        int64_t left = sp->sat_bases;
        while (left > 0) {
            int64_t n = unif_int(s, 20000, 200000);
            if (n > left) n = left;
            int c; int64_t off;
            if (!place(s, A, n, &c, &off)) break;
            for (int64_t i = 0; i < n; i++) {
                uint8_t b = u[i % sp->sat_unit];
                A->seq[c][off + i] = (unif(s) < sp->sat_div) ? other_base(s, b) : b;
            }
            left -= n;
        }
        free(u);
    }
    return 0;
}

/* B contig from A contig: rearrangements in a scratch copy, then SNP + indels. */
static int make_b(uint64_t *s, const pga_pair_spec *sp, const contig_set *A, contig_set *B)
{
    B->n = A->n;
    B->len = (int64_t *) calloc((size_t) B->n, sizeof(int64_t));
    B->seq = (uint8_t **) calloc((size_t) B->n, sizeof(uint8_t *));
    if (!B->len || !B->seq) return -1;

    /* rearrangements are placed on A's coordinates, applied to a copy */
    uint8_t **tmp = (uint8_t **) calloc((size_t) A->n, sizeof(uint8_t *));
    if (!tmp) return -1;
    for (int i = 0; i < A->n; i++) {
        tmp[i] = (uint8_t *) malloc((size_t) A->len[i]);
        if (!tmp[i]) return -1;
        memcpy(tmp[i], A->seq[i], (size_t) A->len[i]);
    }
    for (int k = 0; k < sp->inversions; k++) {
        int64_t n = unif_int(s, sp->inv_min, sp->inv_max);
        int c; int64_t off;
        if (!place(s, A, n, &c, &off)) continue;
        revcomp_inplace(tmp[c] + off, n);
    }
    for (int k = 0; k < sp->moves; k++) {
        int64_t n = unif_int(s, sp->inv_min, sp->inv_max);
        int c; int64_t off;
        if (!place(s, A, 2 * n, &c, &off)) continue;
        /* swap [off, off+n) with [off+n, off+2n) */
        uint8_t *x = tmp[c] + off;
        for (int64_t i = 0; i < n; i++) { uint8_t t = x[i]; x[i] = x[n + i]; x[n + i] = t; }
    }

    double q = 1.0 / (sp->indel_mean > 1.0 ? sp->indel_mean : 1.0);
    for (int i = 0; i < A->n; i++) {
        int64_t n = A->len[i];
        int64_t cap = n + n / 8 + 1024;
        uint8_t *out = (uint8_t *) malloc((size_t) cap);
        if (!out) return -1;
        int64_t o = 0;
        for (int64_t p = 0; p < n; p++) {
            if (o + 64 > cap) {
                cap = cap + cap / 2 + 1024;
                out = (uint8_t *) realloc(out, (size_t) cap);
                if (!out) return -1;
            }
            if (sp->indel_rate > 0.0 && unif(s) < sp->indel_rate) {
                int l = 1;
                while (l < sp->indel_max && unif(s) >= q) l++;
                if (pga_splitmix64(s) & 1) {          /* insertion before base p */
                    for (int k = 0; k < l; k++) out[o++] = draw_base(s, sp->gc);
                } else {                               /* deletion of bases p..p+l-1 */
                    p += l - 1;
                    continue;
                }
            }
            uint8_t b = tmp[i][p];
            out[o++] = (sp->snp_rate > 0.0 && unif(s) < sp->snp_rate) ? other_base(s, b) : b;
        }
        if (o < 64) {                                  /* keep every contig scannable */
            while (o < 64) out[o++] = draw_base(s, sp->gc);
        }
        B->seq[i] = out;
        B->len[i] = o;
        B->total += o;
        free(tmp[i]);
    }
    free(tmp);
    return 0;
}

static void free_set(contig_set *c)
{
    if (c->seq) for (int i = 0; i < c->n; i++) free(c->seq[i]);
    free(c->seq);
    free(c->len);
    memset(c, 0, sizeof(*c));
}

static void pack_contig(uint8_t *dst, const uint8_t *codes, int64_t n)
{
    int64_t full = n / 4;
    for (int64_t k = 0; k < full; k++) {
        const uint8_t *c = codes + 4 * k;
        dst[k] = (uint8_t) (c[0] | (c[1] << 2) | (c[2] << 4) | (c[3] << 6));
    }
    if (n & 3) {
        uint8_t v = 0;
        for (int64_t i = full * 4; i < n; i++) v |= (uint8_t) (codes[i] << (2 * (i & 3)));
        dst[full] = v;
    }
}

static int pack_sets(const contig_set *A, const contig_set *B, pga_gdb *out)
{
    int n = A->n + (B ? B->n : 0);
    memset(out, 0, sizeof(*out));
    out->ncontig = n;
    out->n_a = A->n;
    out->clen = (int64_t *) calloc((size_t) n, sizeof(int64_t));
    out->boff = (int64_t *) calloc((size_t) n + 1, sizeof(int64_t));
    if (!out->clen || !out->boff) return -1;
    int64_t off = 0;
    for (int i = 0; i < n; i++) {
        int64_t l = (i < A->n) ? A->len[i] : B->len[i - A->n];
        out->clen[i] = l;
        out->boff[i] = off;
        off += (l + 3) / 4;
        out->seqtot += l;
        if (l > out->maxctg) out->maxctg = l;
    }
    out->boff[n] = off;
    out->bps_bytes = off;
    out->bps = (uint8_t *) malloc((size_t) (off ? off : 1));
    if (!out->bps) return -1;
    for (int i = 0; i < n; i++) {
        const uint8_t *codes = (i < A->n) ? A->seq[i] : B->seq[i - A->n];
        pack_contig(out->bps + out->boff[i], codes, out->clen[i]);
    }
    return 0;
}

int pga_synth_pair(const pga_pair_spec *sp, uint64_t seed, pga_gdb *out)
{
    contig_set A, B;
    uint64_t s = seed;
    int rc = -1;
    memset(&A, 0, sizeof(A));
    memset(&B, 0, sizeof(B));
    memset(out, 0, sizeof(*out));
    if (sp->a_contigs <= 0 || sp->a_bases < 64 * (int64_t) sp->a_contigs) return -1;
    if (make_a(&s, sp, &A) == 0 && make_b(&s, sp, &A, &B) == 0 && pack_sets(&A, &B, out) == 0)
        rc = 0;
    free_set(&A);
    free_set(&B);
    if (rc) pga_gdb_free(out);
    return rc;
}

static int64_t scaled(int64_t v, double scale, int64_t lo)
{
    int64_t x = (int64_t) llround((double) v * scale);
    return x < lo ? lo : x;
}

int pga_synth_preset(int workload, double scale, pga_pair_spec *sp, uint64_t *seed)
{
    memset(sp, 0, sizeof(*sp));
    if (!(scale > 0.0)) return -1;
    sp->indel_mean = 4.0;
    sp->indel_max = 20;
    switch (workload) {
    case 1:
        *seed = 0x5EED0001ull;
        sp->a_bases = scaled(2000000, scale, 2000);
        sp->a_contigs = 10;
        sp->len_sigma = 0.8;
        sp->gc = 0.41;
        sp->snp_rate = 0.02;
        sp->indel_rate = 0.002;
        return 0;
    case 2:
        *seed = 0x5EED0002ull;
        sp->a_bases = scaled(100000000, scale, 64000);
        sp->a_contigs = (int) scaled(500, scale, 4);
        sp->len_sigma = 0.8;
        sp->gc = 0.41;
        sp->rep_families = 1;
        sp->rep_copies = (int) scaled(2000, scale, 4);
        sp->rep_len_min = sp->rep_len_max = 300;
        sp->rep_div_min = sp->rep_div_max = 0.10;
        sp->tandem_arrays = (int) scaled(50, scale, 2);
        sp->tan_unit_min = 2;  sp->tan_unit_max = 50;
        sp->tan_copies_min = 20; sp->tan_copies_max = 500;
        sp->snp_rate = 0.05;
        sp->indel_rate = 0.005;
        sp->inversions = (int) scaled(10, scale, 1);
        sp->inv_min = scaled(10000, scale, 100);
        sp->inv_max = scaled(500000, scale, 1000);
        return 0;
    case 4:
        *seed = 0x5EED0004ull;
        sp->a_bases = scaled(3100000000ll, scale, 1000000);
        sp->a_contigs = 24;
        sp->len_min = scaled(50000000, scale, 10000);
        sp->len_max = scaled(250000000, scale, 50000);
        sp->gc = 0.41;
        sp->rep_families = (int) scaled(500, scale, 8);
        sp->rep_copies = 0;
        sp->rep_fraction = 0.45;
        sp->rep_len_min = 100; sp->rep_len_max = 6000;
        sp->rep_div_min = 0.05; sp->rep_div_max = 0.25;
        sp->sat_bases = scaled(2000000, scale, 2000);
        sp->sat_unit = 171;
        sp->sat_div = 0.02;
        sp->snp_rate = 0.02;
        sp->indel_rate = 0.003;
        sp->moves = (int) scaled(100, scale, 1);
        sp->inversions = (int) scaled(100, scale, 1);
        sp->inv_min = scaled(10000, scale, 200);
        sp->inv_max = scaled(5000000, scale, 2000);
        return 0;
    default:
        return -1;
    }
}

int pga_synth_batch_specs(int npairs, uint64_t seed, double scale,
                          pga_pair_spec *specs, uint64_t *seeds)
{
    uint64_t s = seed;
    if (npairs <= 0 || !(scale > 0.0)) return -1;
    for (int i = 0; i < npairs; i++) {
        pga_pair_spec *sp = specs + i;
        memset(sp, 0, sizeof(*sp));
        /* lognormal around a 4 Mbp median, clipped to [1,12] Mbp */
        double mb = 4.0 * exp(0.6 * normal(&s));
        if (mb < 1.0) mb = 1.0;
        if (mb > 12.0) mb = 12.0;
        sp->a_bases = scaled((int64_t) (mb * 1e6), scale, 4000);
        sp->a_contigs = (int) unif_int(&s, 1, 50);
        while (sp->a_contigs > 1 && sp->a_bases < 1000 * (int64_t) sp->a_contigs) sp->a_contigs--;
        sp->len_sigma = 0.8;
        sp->gc = 0.30 + 0.35 * unif(&s);
        sp->snp_rate = 0.01 + 0.14 * unif(&s);
        sp->indel_rate = 0.001 + 0.009 * unif(&s);
        sp->indel_mean = 4.0;
        sp->indel_max = 20;
        seeds[i] = pga_splitmix64(&s);
    }
    return 0;
}

int pga_synth_query(const pga_gdb *target, uint64_t seed, int64_t qbases, double div,
                    int64_t seg_min, int64_t seg_max, pga_gdb *out)
{
    uint64_t s = seed;
    contig_set Q;
    int64_t got = 0;
    int cap = 16;
    memset(&Q, 0, sizeof(Q));
    memset(out, 0, sizeof(*out));
    if (target->ncontig <= 0 || qbases <= 0 || seg_min < 64 || seg_max < seg_min) return -1;
    Q.seq = (uint8_t **) calloc((size_t) cap, sizeof(uint8_t *));
    Q.len = (int64_t *) calloc((size_t) cap, sizeof(int64_t));
    if (!Q.seq || !Q.len) { free_set(&Q); return -1; }
    int64_t longest = 0;
    for (int c = 0; c < target->ncontig; c++) if (target->clen[c] > longest) longest = target->clen[c];
    if (longest < 64) { free_set(&Q); return -1; }
    uint8_t *src = (uint8_t *) malloc((size_t) longest);
    if (!src) { free_set(&Q); return -1; }
    while (got < qbases) {
        int64_t n = unif_int(&s, seg_min, seg_max);
        if (n > qbases - got) n = qbases - got;
        if (n < 64) n = 64;
        int c;
        int tries = 0;
        do { c = (int) unif_int(&s, 0, target->ncontig - 1); } while (target->clen[c] < n && ++tries < 64);
        if (target->clen[c] < n) n = target->clen[c];
        int64_t off = unif_int(&s, 0, target->clen[c] - n);
        pga_gdb_contig_codes(target, c, src);
        uint8_t *seg = (uint8_t *) malloc((size_t) (n + n / 4 + 64));
        if (!seg) { free(src); free_set(&Q); return -1; }
        int64_t o = 0;
        for (int64_t i = 0; i < n; i++) {
            double u = unif(&s);
            if (u < div * 0.05) {                      /* deletion of one base */
                continue;
            } else if (u < div * 0.1) {                /* insertion of one base */
                seg[o++] = draw_base(&s, 0.41);
            }
            uint8_t b = src[off + i];
            seg[o++] = (unif(&s) < div) ? other_base(&s, b) : b;
        }
        while (o < 64) seg[o++] = draw_base(&s, 0.41);
        if (pga_splitmix64(&s) & 1) revcomp_inplace(seg, o);
        if (Q.n == cap) {
            cap *= 2;
            Q.seq = (uint8_t **) realloc(Q.seq, (size_t) cap * sizeof(uint8_t *));
            Q.len = (int64_t *) realloc(Q.len, (size_t) cap * sizeof(int64_t));
            if (!Q.seq || !Q.len) { free(seg); free(src); return -1; }
        }
        Q.seq[Q.n] = seg;
        Q.len[Q.n] = o;
        Q.n++;
        got += n;
    }
    free(src);
    /* shuffled contig order (Fisher-Yates) */
    for (int i = Q.n - 1; i > 0; i--) {
        int j = (int) unif_int(&s, 0, i);
        uint8_t *ts = Q.seq[i]; Q.seq[i] = Q.seq[j]; Q.seq[j] = ts;
        int64_t tl = Q.len[i]; Q.len[i] = Q.len[j]; Q.len[j] = tl;
    }
    int rc = pack_sets(&Q, NULL, out);
    free_set(&Q);
    if (rc) pga_gdb_free(out);
    return rc;
}

int pga_gdb_from_codes(const uint8_t *codes, const int64_t *clen, int ncontig,
                       int n_a, pga_gdb *out)
{
    contig_set A;
    memset(&A, 0, sizeof(A));
    A.n = ncontig;
    A.len = (int64_t *) clen;
    A.seq = (uint8_t **) malloc((size_t) (ncontig > 0 ? ncontig : 1) * sizeof(uint8_t *));
    if (!A.seq) return -1;
    int64_t off = 0;
    for (int i = 0; i < ncontig; i++) {
        A.seq[i] = (uint8_t *) codes + off;
        off += clen[i];
    }
    int rc = pack_sets(&A, NULL, out);
    out->n_a = n_a;
    free(A.seq);
    if (rc) pga_gdb_free(out);
    return rc;
}

void pga_gdb_contig_codes(const pga_gdb *g, int c, uint8_t *codes)
{
    const uint8_t *b = g->bps + g->boff[c];
    for (int64_t i = 0; i < g->clen[c]; i++)
        codes[i] = (uint8_t) ((b[i >> 2] >> (2 * (i & 3))) & 3);
}

void pga_gdb_free(pga_gdb *g)
{
    free(g->clen);
    free(g->boff);
    free(g->bps);
    memset(g, 0, sizeof(*g));
}
