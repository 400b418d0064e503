/* pga_gdbfile.c -- reader and writer of GDB genome files: the `<root>.gdb`
 * skeleton in ONE text form and the `.<root>.bps` sequence file (layout in
 * pga_gdbfile.h; lib/GDB.c:1171-1400 reads them, lib/GDB.c:1519-1607 writes them). */
#include "pga_gdbfile.h"

#include <ctype.h>
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* The skeleton's schema as the reference writes it into every file
 * (lib/GDB.c:37-47): kept byte-alike so that ONE readers accept the files. */
static const char *const SCHEMA_LINES[] = {
    "~ D f 4 4 REAL 4 REAL 4 REAL 4 REAL   global: base frequency vector",
    "~ D u 0                               global: upper case when displayed (Deprecated)",
    "~ O S 1 6 STRING                      id for a scaffold",
    "~ D G 1 3 INT                         gap of given length",
    "~ D C 1 3 INT                         contig of given length",
    "~ D M 1 8 INT_LIST                    mask pair list for a contig",
};

/* ".../<root>.gdb" or ".../<root>.1gdb" -> ".../.<root>.bps" (NULL for any
 * other name). */
static char *bps_path(const char *path)
{
    const size_t n = strlen(path);
    size_t ext;
    if (n >= 6 && strcmp(path + n - 5, ".1gdb") == 0) ext = 5;
    else if (n >= 5 && strcmp(path + n - 4, ".gdb") == 0) ext = 4;
    else return NULL;
    const char *slash = strrchr(path, '/');
    const size_t dir = slash ? (size_t) (slash - path) + 1 : 0;
    if (n - ext <= dir) return NULL;
    const size_t root = n - ext - dir;
    char *out = (char *) malloc(dir + root + 6);
    if (!out) return NULL;
    memcpy(out, path, dir);
    out[dir] = '.';
    memcpy(out + dir + 1, path + dir, root);
    memcpy(out + dir + 1 + root, ".bps", 5);
    return out;
}

static int slurp(const char *path, char **buf, size_t *len)
{
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    size_t cap = 1 << 16, n = 0;
    char *b = (char *) malloc(cap + 1);
    while (b) {
        const size_t got = fread(b + n, 1, cap - n, f);
        n += got;
        if (n < cap) break;
        cap *= 2;
        char *nb = (char *) realloc(b, cap + 1);
        if (!nb) { free(b); b = NULL; }
        else b = nb;
    }
    const int bad = !b || ferror(f);
    fclose(f);
    if (bad) { free(b); return -1; }
    b[n] = '\0';
    *buf = b;
    *len = n;
    return 0;
}

/* A decimal integer at *p (after optional blanks); advances *p. */
static int get_int(const char **p, const char *end, int64_t *v)
{
    const char *s = *p;
    while (s < end && *s == ' ') s++;
    if (s >= end || !(isdigit((unsigned char) *s) || *s == '-')) return -1;
    char *e;
    errno = 0;
    const long long x = strtoll(s, &e, 10);
    if (errno || e == s || e > end) return -1;
    *v = x;
    *p = e;
    return 0;
}

/* Grow-only arrays: *a holds at least need elements afterwards. */
#define DEFINE_GROW(name, T)                                                    \
    static int name(T **a, size_t *cap, size_t need)                            \
    {                                                                           \
        if (need <= *cap) return 0;                                             \
        size_t c = *cap ? *cap : 64;                                            \
        while (c < need) c *= 2;                                                \
        T *q = (T *) realloc(*a, c * sizeof(T));                                \
        if (!q) return -1;                                                      \
        *a = q;                                                                 \
        *cap = c;                                                               \
        return 0;                                                               \
    }
DEFINE_GROW(grow_i64, int64_t)
DEFINE_GROW(grow_int, int)
DEFINE_GROW(grow_chr, char)

void pga_gdbfile_free(pga_gdbfile *g)
{
    if (!g) return;
    free(g->slen); free(g->fctg); free(g->hoff); free(g->headers);
    free(g->sbeg); free(g->clen); free(g->boff);
    memset(g, 0, sizeof(*g));
}

/* Reader state: the arrays' capacities and the running scaffold position and
 * .bps offset. */
typedef struct {
    pga_gdbfile *g;
    size_t cs, cf, ch, cb, cc, co, ct;
    int64_t spos, boff;
} rd_state;

/* A scaffold line with id[0..n). */
static int rd_scaffold(rd_state *r, const char *id, int64_t n)
{
    pga_gdbfile *g = r->g;
    if (g->nscaff > 0) g->slen[g->nscaff - 1] = r->spos;
    if (grow_i64(&g->slen, &r->cs, (size_t) g->nscaff + 1) || grow_int(&g->fctg, &r->cf, (size_t) g->nscaff + 2) ||
        grow_i64(&g->hoff, &r->ch, (size_t) g->nscaff + 1) ||
        grow_chr(&g->headers, &r->ct, (size_t) (g->hdrtot + n + 1)))
        return -1;
    g->fctg[g->nscaff] = g->ncontig;
    g->hoff[g->nscaff] = g->hdrtot;
    memcpy(g->headers + g->hdrtot, id, (size_t) n);
    g->hdrtot += n;
    g->headers[g->hdrtot++] = '\0';
    g->nscaff++;
    r->spos = 0;
    return 0;
}

/* A gap ('G') or contig ('C') line of length v inside the current scaffold.
 * Lengths past INT32_MAX are refused (the device loader takes int32 contigs). */
static int rd_piece(rd_state *r, char t, int64_t v)
{
    pga_gdbfile *g = r->g;
    if (g->nscaff == 0 || v < 0 || v > INT32_MAX) return -3;
    if (t == 'C') {
        if (grow_i64(&g->sbeg, &r->cb, (size_t) g->ncontig + 1) || grow_i64(&g->clen, &r->cc, (size_t) g->ncontig + 1) ||
            grow_i64(&g->boff, &r->co, (size_t) g->ncontig + 2))
            return -1;
        g->sbeg[g->ncontig] = r->spos;
        g->clen[g->ncontig] = v;
        g->boff[g->ncontig] = r->boff;
        g->ncontig++;
        g->seqtot += v;
        if (v > g->maxctg) g->maxctg = v;
        r->boff += (v + 3) >> 2;
    }
    r->spos += v;
    return 0;
}

/* One integer in ONE's binary encoding (lib/ONElib.c:3730-3830): a first byte
 * 01xxxxxx holds 0..63 (11xxxxxx: -64..-1), 001xxxxx and one more byte hold
 * 0..8191, 0000 0kkk (1000 0kkk: negative) is followed by the value's k+1
 * low-order bytes, little-endian. */
static int one_int(const unsigned char **pp, const unsigned char *end, int64_t *v)
{
    const unsigned char *p = *pp;
    if (p >= end) return -3;
    const unsigned u0 = p[0];
    if (u0 & 0x40) {
        *v = (u0 & 0x80) ? (int64_t) (int8_t) u0 : (int64_t) (u0 & 0x3f);
        *pp = p + 1;
        return 0;
    }
    if (u0 & 0x20) {
        if (u0 & 0x80 || end - p < 2) return -3;
        *v = ((int64_t) (u0 & 0x1f) << 8) | p[1];
        *pp = p + 2;
        return 0;
    }
    const int k = (int) (u0 & 0x07);
    if (k == 0 || (u0 & 0x78) || end - p < k + 2) return -3;
    uint64_t x = 0;
    for (int b = 0; b <= k; b++) x |= (uint64_t) p[1 + b] << (8 * b);
    if ((u0 & 0x80) && k < 7) x |= ~(uint64_t) 0 << (8 * (k + 1));
    *v = (int64_t) x;
    *pp = p + 2 + k;
    return 0;
}

/* The binary records of a .1gdb file (after its "$ 0" header line) up to the
 * blank record that ends them: each starts with a byte 1ccccccz (c: the line
 * type, A-Z as 0-25 and a-z as 26-51; z: its list is codec-compressed), then
 * the fields (a REAL as 8 bytes, an INT or a list length as one_int), then the
 * list (a STRING's bytes; an INT_LIST's first element, element width and
 * packed rest).  Compressed mask lists are skipped; compressed scaffold ids
 * need ONE's codecs and are refused (-2). */
static int rd_binary(rd_state *r, const unsigned char *p, const unsigned char *end)
{
    pga_gdbfile *g = r->g;
    while (p < end) {
        const unsigned x = *p++;
        if (x == '\n') return 0;                  /* end of the records; the footer follows */
        if (!(x & 0x80)) return -3;
        const int code = (int) ((x & 0x7f) >> 1), comp = (int) (x & 1);
        const char t = code < 26 ? (char) ('A' + code) : code < 52 ? (char) ('a' + code - 26) : code == 54 ? '/' : 0;
        int64_t v, n;
        int rc;
        switch (t) {
        case 'f':
            if (end - p < 32) return -3;
            for (int i = 0; i < 4; i++) {
                double d;
                memcpy(&d, p + 8 * i, 8);
                g->freq[i] = (float) d;
            }
            p += 32;
            break;
        case 'u':
            break;
        case 'S':
            if (one_int(&p, end, &n) || n < 0) return -3;
            if (comp) return -2;
            if (n > end - p) return -3;
            if ((rc = rd_scaffold(r, (const char *) p, n))) return rc;
            p += n;
            break;
        case 'G':
        case 'C':
            if (one_int(&p, end, &v)) return -3;
            if ((rc = rd_piece(r, t, v))) return rc;
            break;
        case 'M':                                  /* mask intervals: not used, skipped */
            if (one_int(&p, end, &n) || n < 0) return -3;
            if (n == 0) break;
            if (one_int(&p, end, &v)) return -3;
            if (n == 1) break;
            if (p >= end) return -3;
            {
                const int64_t width = *p++;
                int64_t skip;
                if (comp) {
                    int64_t bits;
                    if (one_int(&p, end, &bits) || bits < 0) return -3;
                    skip = (bits + 7) >> 3;
                } else {
                    if (width < 1 || width > 8 || n - 1 > (end - p) / width) return -3;
                    skip = (n - 1) * width;
                }
                if (skip > end - p) return -3;
                p += skip;
            }
            break;
        case '/':                                  /* a comment */
            if (one_int(&p, end, &n) || n < 0) return -3;
            if (comp) {
                int64_t bits;
                if (one_int(&p, end, &bits) || bits < 0) return -3;
                n = (bits + 7) >> 3;
            }
            if (n > end - p) return -3;
            p += n;
            break;
        default:
            return -3;
        }
    }
    return 0;
}

int pga_gdbfile_read(const char *path, pga_gdbfile *g)
{
    if (!path || !g) return -3;
    memset(g, 0, sizeof(*g));
    char *buf;
    size_t len;
    if (slurp(path, &buf, &len)) return -1;
    const char *p = buf, *end = buf + len;
    int rc = -3;
    rd_state st;
    memset(&st, 0, sizeof(st));
    st.g = g;
    int first = 1;
    while (p < end) {
        const char *eol = memchr(p, '\n', (size_t) (end - p));
        if (!eol) eol = end;
        const char t = *p;
        const char *q = p + 1;
        if (first) {
            /* "1 <n> gdb <major> <minor>" */
            int64_t n;
            if (t != '1' || get_int(&q, eol, &n) || n != 3 || q + 4 > eol || memcmp(q, " gdb", 4) != 0) goto done;
            first = 0;
        } else if (t == '$') {
            /* "$ <big-endian>": the binary records follow (little-endian only) */
            int64_t big;
            if (get_int(&q, eol, &big) || big != 0 || eol >= end) goto done;
            const int b = rd_binary(&st, (const unsigned char *) eol + 1, (const unsigned char *) end);
            if (b) { rc = b; goto done; }
            break;                                 /* the footer needs no reading */
        } else if (t == 'S') {
            int64_t n;
            /* the bound is checked before any pointer arithmetic with n */
            if (get_int(&q, eol, &n) || n < 0 || n > (int64_t) (eol - q) - 1 || *q != ' ') goto done;
            q++;
            const int r2 = rd_scaffold(&st, q, n);
            if (r2) { rc = r2; goto done; }
        } else if (t == 'G' || t == 'C') {
            int64_t v;
            if (get_int(&q, eol, &v)) goto done;
            const int r2 = rd_piece(&st, t, v);
            if (r2) { rc = r2; goto done; }
        } else if (t == 'f') {
            for (int i = 0; i < 4; i++) {
                while (q < eol && *q == ' ') q++;
                char *e;
                const double x = strtod(q, &e);
                if (e == q || e > eol) goto done;
                g->freq[i] = (float) x;
                q = e;
            }
        } else if (t == 'u' || t == 'M') {
            /* ignored, as Read_GDB does */
        } else if (isalpha((unsigned char) t)) {
            goto done;                             /* not a gdb line type */
        }
        /* anything else is a header, schema or spacer line */
        p = eol + 1;
    }
    if (first) goto done;
    if (g->nscaff > 0) g->slen[g->nscaff - 1] = st.spos;
    if (grow_int(&g->fctg, &st.cf, (size_t) g->nscaff + 1) || grow_i64(&g->boff, &st.co, (size_t) g->ncontig + 1)) {
        rc = -1;
        goto done;
    }
    g->fctg[g->nscaff] = g->ncontig;
    g->boff[g->ncontig] = st.boff;
    rc = 0;
done:
    free(buf);
    if (rc) pga_gdbfile_free(g);
    return rc;
}

int pga_gdbfile_read_bps(const char *path, const pga_gdbfile *g, uint8_t *bps)
{
    if (!path || !g || (!bps && g->ncontig > 0 && g->boff[g->ncontig] > 0)) return -3;
    char *bp = bps_path(path);
    if (!bp) return -3;
    FILE *f = fopen(bp, "rb");
    free(bp);
    if (!f) return -1;
    const int64_t n = g->boff ? g->boff[g->ncontig] : 0;
    const int bad = n > 0 && fread(bps, 1, (size_t) n, f) != (size_t) n;
    fclose(f);
    return bad ? -1 : 0;
}

int pga_gdbfile_write(const char *path, const pga_gdbfile *g, const uint8_t *bps)
{
    if (!path || !g || g->nscaff < 0 || g->ncontig < 0 || (g->nscaff > 0 && (!g->fctg || !g->slen || !g->hoff || !g->headers)))
        return -3;
    if (g->ncontig > 0 && (!g->clen || !g->sbeg)) return -3;
    if (g->nscaff == 0 ? g->ncontig != 0 : (g->fctg[0] != 0 || g->fctg[g->nscaff] != g->ncontig)) return -3;
    int64_t nbytes = 0;
    for (int c = 0; c < g->ncontig; c++) {
        if (g->clen[c] < 0) return -3;
        nbytes += (g->clen[c] + 3) >> 2;
    }
    /* contigs sit in order inside their scaffold */
    for (int s = 0; s < g->nscaff; s++) {
        int64_t spos = 0;
        for (int c = g->fctg[s]; c < g->fctg[s + 1]; c++) {
            if (g->sbeg[c] < spos) return -3;
            spos = g->sbeg[c] + g->clen[c];
        }
        if (g->fctg[s + 1] < g->fctg[s] || g->slen[s] < spos) return -3;
    }
    if (nbytes > 0 && !bps) return -3;
    char *bp = bps_path(path);
    if (!bp) return -3;
    FILE *f = fopen(bp, "wb");
    free(bp);
    if (!f) return -1;
    int bad = nbytes > 0 && fwrite(bps, 1, (size_t) nbytes, f) != (size_t) nbytes;
    if (fclose(f)) bad = 1;
    if (bad) return -1;

    f = fopen(path, "w");
    if (!f) return -1;
    char date[64];
    const time_t now = time(NULL);
    struct tm tmv;
    localtime_r(&now, &tmv);
    strftime(date, sizeof(date), "%Y-%m-%d_%H:%M:%S", &tmv);
    const char *prog = "distributedpanga_amd", *ver = "1.0", *cmd = "pga_gdbfile_write", *src = "unknown";
    fprintf(f, "1 3 gdb 2 1\n");
    fprintf(f, "! 4 %zu %s %zu %s %zu %s %zu %s\n", strlen(prog), prog, strlen(ver), ver, strlen(cmd), cmd,
            strlen(date), date);
    fprintf(f, ".\n< %zu %s 1\n.\n", strlen(src), src);
    for (size_t i = 0; i < sizeof(SCHEMA_LINES) / sizeof(SCHEMA_LINES[0]); i++) fprintf(f, "%s\n", SCHEMA_LINES[i]);
    /* line counts ('#'), longest ('@') and total ('+') scaffold id: ONE
     * readers size their arrays and string buffers from them (Read_GDB does,
     * lib/GDB.c:1260-1262) */
    int64_t ngap = 0, hmax = 0, htot = 0;
    for (int s = 0; s < g->nscaff; s++) {
        const int64_t h = (int64_t) strlen(g->headers + g->hoff[s]);
        hmax = h > hmax ? h : hmax;
        htot += h;
        int64_t spos = 0;
        for (int c = g->fctg[s]; c < g->fctg[s + 1]; c++) {
            ngap += g->sbeg[c] > spos;
            spos = g->sbeg[c] + g->clen[c];
        }
        ngap += g->slen[s] > spos;
    }
    fprintf(f, ".\n# f 1\n");
    if (g->nscaff > 0) {
        fprintf(f, "# S %d\n", g->nscaff);
        if (hmax > 0) fprintf(f, "@ S %lld\n", (long long) hmax);
        if (htot > 0) fprintf(f, "+ S %lld\n", (long long) htot);
    }
    if (ngap > 0) fprintf(f, "# G %lld\n", (long long) ngap);
    if (g->ncontig > 0) fprintf(f, "# C %d\n", g->ncontig);
    fprintf(f, ".\n");
    fprintf(f, "f %f %f %f %f\n", g->freq[0], g->freq[1], g->freq[2], g->freq[3]);
    for (int s = 0; s < g->nscaff; s++) {
        const char *h = g->headers + g->hoff[s];
        fprintf(f, "S %zu %s\n", strlen(h), h);
        int64_t spos = 0;
        for (int c = g->fctg[s]; c < g->fctg[s + 1]; c++) {
            if (g->sbeg[c] > spos) fprintf(f, "G %lld\n", (long long) (g->sbeg[c] - spos));
            fprintf(f, "C %lld\n", (long long) g->clen[c]);
            spos = g->sbeg[c] + g->clen[c];
        }
        if (g->slen[s] > spos) fprintf(f, "G %lld\n", (long long) (g->slen[s] - spos));
    }
    bad = ferror(f);
    if (fclose(f)) bad = 1;
    return bad ? -1 : 0;
}
