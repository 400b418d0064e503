/* pga_oracle.c -- CPU restatement of the reference path, used only as the
 * checker by the tests, the smoke check and the benchmark's CPU-baseline leg.
 *
 * Parity pins (what each function is checked against, in tests/):
 *   oracle_record_sizing   R-03 / E-13 sizing tables (tests/CATALOG.md:415-423,131)
 *   oracle_pack_*          E-01 / E-02 closed forms, first-principles RC
 *   oracle_extract mode 0  E-06 known answer (1564 selected, 3106 records,
 *                          right/left/tie 762/801/2), naive == rolling, and the
 *                          compiled reference extraction (oracle/_ref)
 *   oracle_sort_records    brute-force qsort, compiled reference sort
 *                          (canonicalised), order + conservation invariants
 *   oracle_lcp_bases       S-01 hand table, S-02, R-04
 *   oracle_build_adapter   K-01..K-03, compiled reference adapter
 *   oracle_merge           mode 0 (the reference's walk) == mode 1 (interval
 *                          definition) on adversarial LCP arrays; closed forms
 *                          for single runs; symmetry and count invariants.
 *                          The reference merge itself does not build here
 *                          (src/seed_buffer.c needs MPI), so no compiled
 *                          reference run pins it.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "pga_oracle.h"

#define O_KMER   40
#define O_KBYTES 10
#define O_TMER   12
#define O_SMER    8
#define O_SOFF    4
#define O_NPRE   (1 << 24)

/* The 4-mer scramble table of the syncmer hash (src/tables.c:5-22).  It is
 * part of the selection rule's definition, so it is data, not logic. */
static const uint8_t TMAP[256] = {
    0xff,0xd4,0xf5,0xfd,0xe4,0xad,0x21,0xa5,0xed,0x64,0xbf,0xa9,0xf3,0x70,0xd6,0xf0,
    0xca,0x89,0xcb,0xc9,0x82,0x9d,0x13,0x79,0x0a,0x0f,0x25,0x19,0x3e,0x47,0xa3,0xa8,
    0xf9,0x5e,0xe8,0xa1,0xb0,0x71,0x1d,0x8c,0xde,0x69,0xe7,0x7c,0x56,0x3f,0x90,0xa4,
    0xeb,0x45,0x59,0xf1,0x97,0x4c,0x08,0xa0,0xb8,0x4a,0x86,0xc8,0xcd,0x98,0x7d,0xfc,
    0xef,0x4d,0x83,0x7e,0xdc,0x66,0x2b,0x8e,0xe0,0xa7,0xd0,0xa2,0x88,0x5f,0x7f,0xd9,
    0x9b,0x78,0xd1,0x8b,0xc3,0x8f,0x2d,0xe6,0x18,0x27,0x2c,0x24,0x94,0xb7,0xce,0xbd,
    0x0d,0x04,0x1c,0x09,0x16,0x23,0x00,0x1e,0x1a,0x29,0x2e,0x15,0x01,0x10,0x2a,0x20,
    0xbe,0x31,0x43,0x58,0xc2,0xaa,0x1f,0xe5,0xc5,0x9e,0xcf,0xc6,0x68,0xb2,0x80,0xf4,
    0xf8,0x53,0xb6,0x93,0x76,0x37,0x11,0x40,0xda,0x51,0xba,0x46,0x42,0x30,0x60,0x6d,
    0x5c,0x39,0x9f,0x48,0x6c,0x62,0x28,0x67,0x06,0x12,0x26,0x0e,0x33,0x50,0xa6,0x63,
    0xdd,0x3b,0xab,0x4b,0x72,0x5b,0x22,0x6f,0xb4,0x61,0x92,0x99,0x36,0x38,0x65,0xac,
    0x4f,0x2f,0x32,0x44,0x54,0x3c,0x03,0x5d,0x73,0x3a,0x77,0x84,0x8d,0x4e,0x49,0xd2,
    0xfb,0x91,0x6a,0xcc,0x8a,0x35,0x02,0x55,0x7a,0x34,0x96,0x3d,0xd3,0x41,0x85,0xf2,
    0xb1,0x75,0xc4,0xb5,0xbb,0xb3,0x1b,0xd5,0x07,0x05,0x17,0x0b,0x7b,0xd7,0xdf,0xea,
    0xe3,0x57,0xc0,0x95,0x9c,0x6e,0x14,0xae,0xb9,0x6b,0xc1,0x81,0x87,0x74,0xd8,0xe2,
    0xec,0x52,0xbc,0xe9,0xe1,0xdb,0x0c,0xf7,0xaf,0x5a,0x9a,0xc7,0xfa,0xf6,0xee,0xfe
};

int oracle_tmap(int i) { return TMAP[i & 255]; }

void oracle_free(void *p) { free(p); }

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double) ts.tv_sec + 1e-9 * (double) ts.tv_nsec;
}

/* ------------------------------------------------------------------ sizing */

/* Smallest b with 256^b >= v+1 (src/record.c:6-16). */
static int bytes_for(int64_t v)
{
    int b = 0;
    int64_t cap = 1;
    while (cap < v + 1) { cap *= 256; b++; }
    return b;
}

void oracle_record_sizing(int64_t maxctg, int64_t ncontig, int *P, int *C, int *rsize)
{
    *P = bytes_for(maxctg);                 /* positions run [0, maxctg] */
    *C = bytes_for(2 * ncontig - 1);        /* contig id + strand bit    */
    *rsize = 12 + *P + *C;
}

/* ----------------------------------------------------------------- packing */

/* 4-mer reverse complement: reverse the four 2-bit fields, complement each
 * (A<->T, C<->G is v -> 3-v).  Equal to Comp[] of src/tables.c:24-33. */
static int rc4(int b)
{
    int out = 0;
    for (int f = 0; f < 4; f++) out |= (3 - ((b >> (2 * f)) & 3)) << (2 * (3 - f));
    return out & 0xff;
}

void oracle_comp_table(int comp[256])
{
    for (int i = 0; i < 256; i++) comp[i] = rc4(i);
}

static int pack4(const uint8_t *b) { return (b[0] << 6) | (b[1] << 4) | (b[2] << 2) | b[3]; }

void oracle_pack_forward(const uint8_t *bases40, uint8_t *out10)
{
    for (int k = 0; k < O_KBYTES; k++) out10[k] = (uint8_t) pack4(bases40 + 4 * k);
}

/* RC 40-mer of forward bases [j-28, j+12): output byte k is the RC of the
 * forward 4-mer at local offset 36-4k (src/pack.c:64-78). */
void oracle_pack_rc(const uint8_t *bases40, uint8_t *out10)
{
    for (int k = 0; k < O_KBYTES; k++) out10[k] = (uint8_t) rc4(pack4(bases40 + 36 - 4 * k));
}

int oracle_mz(const uint8_t *codes, int64_t i)
{
    int lo = pack4(codes + i), hi = pack4(codes + i + 4);
    int fwd = (TMAP[lo] << 8) | TMAP[hi];
    int rev = (TMAP[rc4(hi)] << 8) | TMAP[rc4(lo)];
    return fwd < rev ? fwd : rev;
}

/* -------------------------------------------------------------- extraction */

typedef struct {
    uint8_t *buf;
    int64_t  n, cap;
    int      P, C, rsize;
} recbuf;

static int rb_push(recbuf *b, int contig, int64_t pos, int strand, const uint8_t *kmer)
{
    if (b->n == b->cap) {
        int64_t nc = b->cap ? 2 * b->cap : 1 << 16;
        uint8_t *p = (uint8_t *) realloc(b->buf, (size_t) (nc * b->rsize));
        if (!p) return -1;
        b->buf = p;
        b->cap = nc;
    }
    uint8_t *r = b->buf + b->n * b->rsize;
    r[0] = 0;
    memcpy(r + 1, kmer, O_KBYTES);
    r[11] = 0;
    for (int k = 0; k < b->P; k++) r[12 + k] = (uint8_t) (pos >> (8 * k));
    uint64_t cv = (uint64_t) contig;
    if (strand) cv |= 1ull << (8 * b->C - 1);
    for (int k = 0; k < b->C; k++) r[12 + b->P + k] = (uint8_t) (cv >> (8 * k));
    b->n++;
    return 0;
}

/* Emit the records of a selected offset j (src/extract.c:200-228). */
static int emit_j(recbuf *b, const uint8_t *s, int64_t len, int contig, int64_t j)
{
    uint8_t k[O_KBYTES];
    if (j <= len - O_KMER) {
        oracle_pack_forward(s + j, k);
        if (rb_push(b, contig, j, 0, k)) return -1;
    }
    if (j >= O_KMER - O_TMER) {
        oracle_pack_rc(s + j - (O_KMER - O_TMER), k);
        if (rb_push(b, contig, j + O_TMER, 1, k)) return -1;
    }
    return 0;
}

/* Naive definition: a 12-mer at j is a closed syncmer iff the minimum of its
 * five canonical 8-mer hashes is at index 0 or index 4; ties select
 * (tests/support/ref_syncmer.c:196-227). */
static int scan_naive(recbuf *b, const uint8_t *s, int64_t len, int contig, oracle_extract_stats *st)
{
    if (len < O_TMER) return 0;
    for (int64_t j = 0; j <= len - O_TMER; j++) {
        int h[5], m, a = 0;
        for (int t = 0; t < 5; t++) h[t] = oracle_mz(s, j + t);
        m = h[0];
        for (int t = 1; t < 5; t++) if (h[t] < m) { m = h[t]; a = t; }
        if (a == 0) st->n_left++;
        if (a == 4) st->n_right++;
        if (h[4] == m && a < 4) st->n_tie++;
        if (!(a == 0 || h[4] == m)) continue;
        st->n_sel++;
        if (emit_j(b, s, len, contig, j)) return -1;
    }
    return 0;
}

/* Rolling scan with a 4-slot ring of 8-mer hashes and an incrementally kept
 * earliest minimum over the last four (src/extract.c:115-229), over the whole
 * contig held in memory (the window refills of :230-248 change nothing). */
static int scan_rolling(recbuf *b, const uint8_t *s, int64_t len, int contig)
{
    int ring[4], mn, mpos;
    if (len < O_TMER) return 0;
    for (int t = 0; t < 4; t++) ring[t] = oracle_mz(s, t);
    mn = ring[0]; mpos = 0;
    for (int t = 1; t < 4; t++) if (ring[t] < mn) { mn = ring[t]; mpos = t; }
    for (int64_t i = O_SOFF; i <= len - O_SMER; i++) {
        int h = oracle_mz(s, i);
        int sel = 0;
        ring[i & 3] = h;
        if (h < mn) {                         /* new strict minimum: right end */
            mn = h; mpos = (int) i; sel = 1;
        } else if (mpos == i - O_SOFF) {      /* minimum at the left end: select, rescan */
            sel = 1;
            mpos++;
            mn = ring[mpos & 3];
            for (int64_t q = mpos + 1; q <= i; q++)
                if (ring[q & 3] < mn) { mn = ring[q & 3]; mpos = (int) q; }
        } else if (h == mn) {                 /* tie with the minimum at the right end */
            sel = 1;
        }
        if (sel && emit_j(b, s, len, contig, i - O_SOFF)) return -1;
    }
    return 0;
}

static void unpack_contig(const uint8_t *bps, int64_t boff, int64_t n, uint8_t *codes)
{
    const uint8_t *p = bps + boff;
    for (int64_t i = 0; i < n; i++) codes[i] = (uint8_t) ((p[i >> 2] >> (2 * (i & 3))) & 3);
}

int oracle_extract(const uint8_t *bps, const int64_t *boff, const int64_t *clen, int ncontig,
                   int mode, int P, int C, uint8_t **out, int64_t *nrec,
                   oracle_extract_stats *st)
{
    recbuf b;
    int64_t maxl = 0;
    memset(&b, 0, sizeof(b));
    memset(st, 0, sizeof(*st));
    b.P = P; b.C = C; b.rsize = 12 + P + C;
    for (int c = 0; c < ncontig; c++) if (clen[c] > maxl) maxl = clen[c];
    uint8_t *codes = (uint8_t *) malloc((size_t) (maxl + 8));
    if (!codes) return -1;
    for (int c = 0; c < ncontig; c++) {
        if (boff[c] < 0) continue;                   /* src/extract.c:97-98 */
        unpack_contig(bps, boff[c], clen[c], codes);
        int rc = mode == 0 ? scan_naive(&b, codes, clen[c], c, st)
                           : scan_rolling(&b, codes, clen[c], c);
        if (rc) { free(codes); free(b.buf); return -1; }
    }
    free(codes);
    *out = b.buf;
    *nrec = b.n;
    return 0;
}

/* --------------------------------------------------------------------- sort */

/* Stable LSD radix sort on k-mer bytes 10..1 (one byte per pass).  With the
 * input in extraction order the equal-k-mer runs keep (contig, j, strand)
 * order: the canonical form of the reference's unstable msd_sort
 * (src/sort.c:59-115, lib/MSDsort.c:404-507). */
int oracle_sort_records(uint8_t *rec, int64_t n, int rsize)
{
    if (n <= 1) return 0;
    uint8_t *tmp = (uint8_t *) malloc((size_t) (n * rsize));
    if (!tmp) return -1;
    uint8_t *src = rec, *dst = tmp;
    for (int byte = O_KBYTES; byte >= 1; byte--) {
        int64_t cnt[256];
        memset(cnt, 0, sizeof(cnt));
        for (int64_t i = 0; i < n; i++) cnt[src[i * rsize + byte]]++;
        int64_t sum = 0;
        for (int d = 0; d < 256; d++) { int64_t c = cnt[d]; cnt[d] = sum; sum += c; }
        for (int64_t i = 0; i < n; i++) {
            const uint8_t *r = src + i * rsize;
            memcpy(dst + (cnt[r[byte]]++) * rsize, r, (size_t) rsize);
        }
        uint8_t *t = src; src = dst; dst = t;
    }
    if (src != rec) memcpy(rec, src, (size_t) (n * rsize));   /* 10 passes: even, never taken */
    free(tmp);
    return 0;
}

static int g_qP, g_qC, g_qrs;

static uint64_t le_field(const uint8_t *p, int nb)
{
    uint64_t v = 0;
    for (int k = 0; k < nb; k++) v |= (uint64_t) p[k] << (8 * k);
    return v;
}

static int qcmp(const void *x, const void *y)
{
    const uint8_t *a = (const uint8_t *) x, *b = (const uint8_t *) y;
    int c = memcmp(a + 1, b + 1, O_KBYTES);
    if (c) return c;
    uint64_t ca = le_field(a + 12 + g_qP, g_qC), cb = le_field(b + 12 + g_qP, g_qC);
    uint64_t smask = 1ull << (8 * g_qC - 1);
    int sa = (ca & smask) != 0, sb = (cb & smask) != 0;
    ca &= ~smask; cb &= ~smask;
    if (ca != cb) return ca < cb ? -1 : 1;
    int64_t ja = (int64_t) le_field(a + 12, g_qP) - O_TMER * sa;
    int64_t jb = (int64_t) le_field(b + 12, g_qP) - O_TMER * sb;
    if (ja != jb) return ja < jb ? -1 : 1;
    return sa - sb;
}

int oracle_sort_records_qsort(uint8_t *rec, int64_t n, int rsize, int P, int C)
{
    g_qP = P; g_qC = C; g_qrs = rsize;
    qsort(rec, (size_t) n, (size_t) rsize, qcmp);
    return 0;
}

/* ---------------------------------------------------------------------- LCP */

/* Leading bases two packed k-mers share, in [0,40] (src/sort.c:16-32). */
int oracle_lcp_bases(const uint8_t *a, const uint8_t *b)
{
    for (int k = 0; k < O_KBYTES; k++) {
        int x = a[k] ^ b[k];
        if (x) {
            int l = 4 * k;
            while (!(x & 0xc0)) { x <<= 2; l++; }
            return l;
        }
    }
    return O_KMER;
}

/* First record 0, every other one its LCP with its predecessor (src/sort.c:38-57). */
void oracle_recalc_lcps(uint8_t *rec, int64_t n, int rsize)
{
    if (n <= 0) return;
    rec[0] = 0;
    for (int64_t i = 1; i < n; i++)
        rec[i * rsize] = (uint8_t) oracle_lcp_bases(rec + (i - 1) * rsize + 1, rec + i * rsize + 1);
}

/* ------------------------------------------------------------------ adapter */

/* GIX-layout entries and the cumulative 24-bit prefix counts
 * (src/kmer_adapter.c:6-96). */
int oracle_build_adapter(const uint8_t *rec, int64_t n, int P, int C,
                         uint8_t *entries, int64_t *cum, int64_t *maxpre)
{
    int rsize = 12 + P + C, e = 9 + P + C;
    int64_t run = 0, best = 0;
    int prev = -1;
    memset(cum, 0, sizeof(int64_t) * O_NPRE);
    for (int64_t i = 0; i < n; i++) {
        const uint8_t *r = rec + i * rsize;
        uint8_t *d = entries + i * e;
        int pre = (r[1] << 16) | (r[2] << 8) | r[3];
        if (pre == prev) run++;
        else { if (run > best) best = run; run = 1; prev = pre; }
        memcpy(d, r + 4, 7);
        d[7] = 0;
        d[8] = r[0];
        memcpy(d + 9, r + 12, (size_t) (P + C));
        cum[pre]++;
    }
    if (run > best) best = run;
    for (int64_t p = 1; p < O_NPRE; p++) cum[p] += cum[p - 1];
    *maxpre = best;
    return 0;
}

/* -------------------------------------------------------------------- merge */

typedef struct {
    uint8_t *buf;
    int64_t  len, cap;
} stream;

static int st_put(stream *s, const uint8_t *src, int n)
{
    if (s->len + n > s->cap) {
        int64_t nc = s->cap ? s->cap : 1 << 16;
        while (nc < s->len + n) nc *= 2;
        uint8_t *p = (uint8_t *) realloc(s->buf, (size_t) nc);
        if (!p) return -1;
        s->buf = p;
        s->cap = nc;
    }
    memcpy(s->buf + s->len, src, (size_t) n);
    s->len += n;
    return 0;
}

typedef struct {
    const uint8_t *ent;
    int e, P, C, pc;
    stream N, Cs;
    int64_t *nbuck, *cbuck;
} mctx;

/* Seeds of anchor i against every other entry of [lo, hi). */
static int emit_group(mctx *m, int64_t i, int64_t lo, int64_t hi, int plen)
{
    const uint8_t *pi = m->ent + i * m->e + 9;
    uint8_t seed[1 + 2 * 16];
    int si = (pi[m->pc - 1] & 0x80) != 0;
    uint64_t icont = le_field(pi + m->P, m->C) & ~(1ull << (8 * m->C - 1));
    seed[0] = (uint8_t) plen;
    memcpy(seed + 1, pi, (size_t) m->pc);
    seed[m->pc] &= 0x7f;                         /* anchor's strand bit cleared */
    for (int64_t p = lo; p < hi; p++) {
        if (p == i) continue;
        const uint8_t *pp = m->ent + p * m->e + 9;
        int sp = (pp[m->pc - 1] & 0x80) != 0;
        memcpy(seed + 1 + m->pc, pp, (size_t) m->pc);
        if (si == sp) {
            if (st_put(&m->N, seed, 1 + 2 * m->pc)) return -1;
            m->nbuck[icont]++;
        } else {
            if (st_put(&m->Cs, seed, 1 + 2 * m->pc)) return -1;
            m->cbuck[icont]++;
        }
    }
    return 0;
}

#define LCPAT(k) (m->ent[(k) * m->e + 8])

/* The walk of src/fastga_pipeline.c:287-407 over one prefix group [g0, g1):
 * plen follows max(LCP[i], LCP[i+1]) through the eorun flag, vlcp[x] holds the
 * start of the current run of LCP >= x, and the forward scan stops at the
 * group end (an LCP of 11) or once FREQ entries are covered. */
static int walk_group_fastga(mctx *m, int64_t g0, int64_t g1, int freq)
{
    int64_t vlcp[O_KMER + 1];
    int64_t rend, low, hgh, top;
    int plen, eorun;
#define L(k) ((k) >= g1 ? 11 : (int) LCPAT(k))
    for (int x = 0; x <= O_KMER; x++) vlcp[x] = g0;
    rend = g0 + 1;
    plen = L(rend);
    for (int x = L(g0); x <= plen; x++) vlcp[x] = g0;
    eorun = plen <= 11;
    for (int64_t i = g0; i < g1; i++) {
        if (i > g0) {
            if (eorun) plen = L(i);
            rend++;
            if (L(rend) < plen) eorun = 1;
            else if (L(rend) == plen) eorun = 0;
            else {
                for (int x = plen + 1; x <= L(rend); x++) vlcp[x] = rend - 1;
                eorun = 0;
                plen = L(rend);
            }
        }
        low = vlcp[plen];
        hgh = rend;
        top = low + freq;
        if (!eorun) {
            do {
                hgh++;
                if (hgh > top) break;
            } while (L(hgh) >= plen);
        }
        if (hgh >= top) continue;
        if (emit_group(m, i, low, hgh, plen)) return -1;
    }
#undef L
    return 0;
}

/* The same seeds from the definition: anchor i of a group of two or more
 * entries gets plen = max(LCP[i], LCP[i+1]) (a group edge counts as < 12);
 * its partners are the maximal run [a, b) around i whose internal LCPs are all
 * >= plen; the anchor is skipped when b - a >= FREQ. */
static int walk_group_interval(mctx *m, int64_t g0, int64_t g1, int freq)
{
    if (g1 - g0 < 2) return 0;
    for (int64_t i = g0; i < g1; i++) {
        int li = (i == g0) ? 0 : LCPAT(i);
        int ln = (i + 1 < g1) ? LCPAT(i + 1) : 0;
        int plen = li > ln ? li : ln;
        int64_t a = i, b = i + 1;
        while (a > g0 && LCPAT(a) >= plen) a--;
        while (b < g1 && LCPAT(b) >= plen) b++;
        if (b - a >= freq) continue;
        if (emit_group(m, i, a, b, plen)) return -1;
    }
    return 0;
}

int oracle_merge(const uint8_t *entries, int64_t n, int P, int C, int ncontig,
                 int freq, int mode, uint8_t **nseeds, int64_t *nlen,
                 uint8_t **cseeds, int64_t *clen, int64_t *nbuck, int64_t *cbuck)
{
    mctx mm, *m = &mm;
    memset(m, 0, sizeof(*m));
    m->ent = entries; m->e = 9 + P + C; m->P = P; m->C = C; m->pc = P + C;
    m->nbuck = nbuck; m->cbuck = cbuck;
    memset(nbuck, 0, sizeof(int64_t) * (size_t) ncontig);
    memset(cbuck, 0, sizeof(int64_t) * (size_t) ncontig);
    int64_t g0 = 0;
    while (g0 < n) {
        int64_t g1 = g0 + 1;
        while (g1 < n && LCPAT(g1) >= 12) g1++;     /* a 24-bit prefix group */
        int rc = mode == 0 ? walk_group_fastga(m, g0, g1, freq)
                           : walk_group_interval(m, g0, g1, freq);
        if (rc) { free(m->N.buf); free(m->Cs.buf); return -1; }
        g0 = g1;
    }
    *nseeds = m->N.buf;  *nlen = m->N.len;
    *cseeds = m->Cs.buf; *clen = m->Cs.len;
    return 0;
}
#undef LCPAT

/* ----------------------------------------------------------------- pipeline */

int oracle_pipeline(const uint8_t *bps, const int64_t *boff, const int64_t *clen, int ncontig,
                    int freq, oracle_result *r)
{
    int64_t maxctg = 0;
    double t0;
    memset(r, 0, sizeof(*r));
    for (int c = 0; c < ncontig; c++) if (clen[c] > maxctg) maxctg = clen[c];
    oracle_record_sizing(maxctg, ncontig, &r->P, &r->C, &r->rsize);
    r->e = 9 + r->P + r->C;
    r->s = 1 + 2 * (r->P + r->C);
    r->ncontig = ncontig;

    t0 = now_s();
    if (oracle_extract(bps, boff, clen, ncontig, 1, r->P, r->C, &r->records, &r->nrec, &r->xs))
        return -1;
    r->t_extract = now_s() - t0;

    t0 = now_s();
    if (oracle_sort_records(r->records, r->nrec, r->rsize)) return -1;
    oracle_recalc_lcps(r->records, r->nrec, r->rsize);
    r->t_sort = now_s() - t0;

    t0 = now_s();
    r->entries = (uint8_t *) malloc((size_t) (r->nrec * r->e + 1));
    r->cum = (int64_t *) malloc(sizeof(int64_t) * O_NPRE);
    if (!r->entries || !r->cum) return -1;
    oracle_build_adapter(r->records, r->nrec, r->P, r->C, r->entries, r->cum, &r->maxpre);
    r->t_adapter = now_s() - t0;

    t0 = now_s();
    r->nbuck = (int64_t *) calloc((size_t) (ncontig > 0 ? ncontig : 1), sizeof(int64_t));
    r->cbuck = (int64_t *) calloc((size_t) (ncontig > 0 ? ncontig : 1), sizeof(int64_t));
    if (!r->nbuck || !r->cbuck) return -1;
    if (oracle_merge(r->entries, r->nrec, r->P, r->C, ncontig, freq, 0,
                     &r->nseeds, &r->nlen, &r->cseeds, &r->clen, r->nbuck, r->cbuck))
        return -1;
    r->t_merge = now_s() - t0;
    return 0;
}

void oracle_result_free(oracle_result *r)
{
    free(r->records); free(r->entries); free(r->cum);
    free(r->nseeds); free(r->cseeds); free(r->nbuck); free(r->cbuck);
    memset(r, 0, sizeof(*r));
}


/* ---------------------------------------------------------------- seed sort */

/* src/pga-mpi-merge.c:415-425: DBYTE = bytes for MAXDAG = AMXPOS + BMXPOS. */
int oracle_dbyte(int64_t maxctg)
{
    const int64_t maxdag = 2 * maxctg;
    int d = 0;
    int64_t cum = 1;
    while (cum < maxdag) { cum *= 256; d++; }
    return d;
}

static int O_SWIDE;

/* rmsd_sort's order: the whole record as a number, last byte most significant. */
static int cmp_from_end(const void *a, const void *b)
{
    const uint8_t *x = (const uint8_t *) a, *y = (const uint8_t *) b;
    for (int k = O_SWIDE - 1; k >= 0; k--)
        if (x[k] != y[k]) return x[k] < y[k] ? -1 : 1;
    return 0;
}

static int64_t le_get(const uint8_t *p, int n)
{
    int64_t v = 0;
    for (int k = n - 1; k >= 0; k--) v = (v << 8) | p[k];
    return v;
}

static void le_put(uint8_t *p, int64_t v, int n)
{
    for (int k = 0; k < n; k++) p[k] = (uint8_t) (v >> (8 * k));
}

int oracle_seed_sort(const uint8_t *seeds, int64_t nseeds, int P, int C, int ncontig, int64_t maxctg,
                     int comp, int sort, uint8_t **out, int *swide_out)
{
    const int s = 1 + 2 * (P + C);
    const int dbyte = oracle_dbyte(maxctg);
    const int swide = 2 * dbyte + C + 2;
    const int64_t amx = maxctg, bmx = maxctg, maxdag = amx + bmx;
    const int64_t mask = ((int64_t) 1 << (8 * C - 1)) - 1;
    *swide_out = swide;
    int64_t *start = (int64_t *) calloc((size_t) ncontig + 1, sizeof(int64_t));
    uint8_t *arr = (uint8_t *) malloc((size_t) (nseeds * swide + 1));
    if (!start || !arr) { free(start); free(arr); return -1; }
    /* panel sizes, then offsets (the bucket prefix of pair_sort_search) */
    for (int64_t i = 0; i < nseeds; i++) start[le_get(seeds + i * s + 1 + P, C) + 1]++;
    for (int c = 0; c < ncontig; c++) start[c + 1] += start[c];
    int64_t *fill = (int64_t *) malloc(sizeof(int64_t) * (size_t) (ncontig + 1));
    if (!fill) { free(start); free(arr); return -1; }
    memcpy(fill, start, sizeof(int64_t) * (size_t) (ncontig + 1));
    for (int64_t i = 0; i < nseeds; i++) {
        const uint8_t *b = seeds + i * s;
        const int lcp = b[0];
        const int64_t ipost = le_get(b + 1, P), icont = le_get(b + 1 + P, C);
        const int64_t jpost = le_get(b + 1 + P + C, P), jcont = le_get(b + 1 + 2 * P + C, C) & mask;
        int64_t diag, anti;
        if (comp) { diag = maxdag - (ipost + jpost); anti = amx - (ipost - jpost); }
        else      { diag = bmx + (ipost - jpost);    anti = ipost + jpost; }
        const int64_t band = diag >> 6;                       /* BUCK_SHIFT */
        uint8_t *x = arr + swide * fill[icont]++;
        x[0] = (uint8_t) lcp;
        x[1] = (uint8_t) (diag - (band << 6));
        le_put(x + 2, anti, dbyte);
        le_put(x + 2 + dbyte, band, dbyte);
        le_put(x + 2 + 2 * dbyte, jcont, C);
    }
    if (sort) {
        O_SWIDE = swide;
        for (int c = 0; c < ncontig; c++)
            if (start[c + 1] - start[c] > 1)
                qsort(arr + start[c] * swide, (size_t) (start[c + 1] - start[c]), (size_t) swide, cmp_from_end);
    }
    free(start);
    free(fill);
    *out = arr;
    return 0;
}
