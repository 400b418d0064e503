/* pga_gix.c -- writer and reader of the on-disk GIX index (layout in pga_gix.h,
 * notes/GIXmake_dump.c:1632-1709; reader offsets src/fastga_order_core.c:21-35). */
#include "pga_gix.h"

#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NPRE (1 << 24)
#define KMER 40

static void part_path(char *out, size_t n, const char *dir, const char *root, int part)
{
    snprintf(out, n, "%s/.%s.ktab.%d", dir, root, part);
}

static void gix_path(char *out, size_t n, const char *dir, const char *root)
{
    snprintf(out, n, "%s/%s.gix", dir, root);
}

static int put(FILE *f, const void *p, size_t bytes)
{
    return fwrite(p, 1, bytes, f) == bytes ? 0 : -1;
}

static int get(FILE *f, void *p, size_t bytes)
{
    return fread(p, 1, bytes, f) == bytes ? 0 : -1;
}

/* Part boundaries: nparts runs of consecutive entries split at the prefix
 * group boundary nearest to each equal share. */
static void split_parts(const int64_t *cum, int64_t nrec, int nparts, int64_t *bound)
{
    bound[0] = 0;
    for (int k = 1; k < nparts; k++) {
        const int64_t want = nrec * k / nparts;
        int lo = 0, hi = NPRE - 1;                 /* first prefix with cum >= want */
        while (lo < hi) {
            const int m = lo + (hi - lo) / 2;
            if (cum[m] >= want) hi = m; else lo = m + 1;
        }
        const int64_t after = cum[lo], before = lo ? cum[lo - 1] : 0;
        int64_t b = (after - want < want - before) ? after : before;
        if (b < bound[k - 1]) b = bound[k - 1];
        bound[k] = b;
    }
    bound[nparts] = nrec;
}

int pga_gix_write(const char *dir, const char *root, int nparts, const uint8_t *entries, int64_t nrec,
                  int P, int C, const int64_t *cum, int64_t maxpre, int ncontig, const int *perm)
{
    if (!dir || !root || nparts < 1 || nrec < 0 || (nrec > 0 && !entries) || !cum || P < 1 || C < 1 || ncontig < 0)
        return -3;
    if (cum[NPRE - 1] != nrec) return -3;
    const int e = 9 + P + C;
    char path[4096];
    int64_t *bound = malloc(sizeof(int64_t) * (size_t) (nparts + 1));
    if (!bound) return -1;
    split_parts(cum, nrec, nparts, bound);
    int rc = 0;
    for (int k = 0; k < nparts && !rc; k++) {
        part_path(path, sizeof(path), dir, root, k + 1);
        FILE *f = fopen(path, "wb");
        if (!f) { rc = -1; break; }
        const int kmer = KMER;
        const int64_t n = bound[k + 1] - bound[k];
        rc = put(f, &kmer, sizeof(int));
        if (!rc) rc = put(f, &n, sizeof(int64_t));
        if (!rc && n > 0) rc = put(f, entries + bound[k] * (int64_t) e, (size_t) (n * e));
        if (fclose(f) && !rc) rc = -1;
    }
    free(bound);
    if (rc) return rc;

    gix_path(path, sizeof(path), dir, root);
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    const int head[4] = {KMER, nparts, 1, 3};
    const int zero = 0;
    const int64_t end = -1;
    rc = put(f, head, sizeof(head));
    if (!rc) rc = put(f, cum, sizeof(int64_t) * NPRE);
    if (!rc) rc = put(f, &P, sizeof(int));
    if (!rc) rc = put(f, &C, sizeof(int));
    if (!rc) rc = put(f, &nparts, sizeof(int));
    if (!rc) rc = put(f, &maxpre, sizeof(int64_t));
    if (!rc) rc = put(f, &zero, sizeof(int));
    if (!rc) rc = put(f, &ncontig, sizeof(int));
    for (int i = 0; i < ncontig && !rc; i++) {
        const int v = perm ? perm[i] : i;
        rc = put(f, &v, sizeof(int));
    }
    if (!rc) rc = put(f, &end, sizeof(int64_t));
    if (fclose(f) && !rc) rc = -1;
    return rc;
}

int pga_gix_read_header(const char *dir, const char *root, pga_gix_header *h, int64_t *cum, int *perm)
{
    if (!dir || !root || !h) return -3;
    char path[4096];
    gix_path(path, sizeof(path), dir, root);
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    int head[4];
    int rc = get(f, head, sizeof(head));
    if (!rc && (head[0] != KMER || head[1] < 1 || head[2] != 1 || head[3] != 3)) rc = -3;
    if (!rc) {
        if (cum) rc = get(f, cum, sizeof(int64_t) * NPRE);
        else rc = fseek(f, (long) (sizeof(int64_t) * NPRE), SEEK_CUR) ? -1 : 0;
    }
    int nfile = 0;
    if (!rc) rc = get(f, &h->P, sizeof(int));
    if (!rc) rc = get(f, &h->C, sizeof(int));
    if (!rc) rc = get(f, &nfile, sizeof(int));
    if (!rc) rc = get(f, &h->maxpre, sizeof(int64_t));
    if (!rc) rc = get(f, &h->freq, sizeof(int));
    if (!rc) rc = get(f, &h->ncontig, sizeof(int));
    if (!rc && (nfile != head[1] || h->ncontig < 0 || h->P < 1 || h->C < 1)) rc = -3;
    if (!rc && perm) rc = get(f, perm, sizeof(int) * (size_t) h->ncontig);
    fclose(f);
    if (rc) return rc;
    h->kmer = head[0];
    h->nparts = head[1];
    h->nents = 0;
    for (int k = 0; k < h->nparts; k++) {
        part_path(path, sizeof(path), dir, root, k + 1);
        FILE *g = fopen(path, "rb");
        if (!g) return -1;
        int kmer = 0;
        int64_t n = 0;
        rc = get(g, &kmer, sizeof(int));
        if (!rc) rc = get(g, &n, sizeof(int64_t));
        fclose(g);
        if (rc) return rc;
        if (kmer != KMER || n < 0) return -3;
        h->nents += n;
    }
    return 0;
}

int pga_gix_read_entries(const char *dir, const char *root, const pga_gix_header *h, uint8_t *entries,
                         int64_t cap)
{
    if (!dir || !root || !h || (!entries && h->nents > 0)) return -3;
    const int e = 9 + h->P + h->C;
    if (h->nents * (int64_t) e > cap) return -3;
    char path[4096];
    int64_t at = 0;
    for (int k = 0; k < h->nparts; k++) {
        part_path(path, sizeof(path), dir, root, k + 1);
        FILE *f = fopen(path, "rb");
        if (!f) return -1;
        int kmer = 0;
        int64_t n = 0;
        int rc = get(f, &kmer, sizeof(int));
        if (!rc) rc = get(f, &n, sizeof(int64_t));
        if (!rc && (at + n) * (int64_t) e > cap) rc = -3;
        if (!rc && n > 0) rc = get(f, entries + at * (int64_t) e, (size_t) (n * e));
        fclose(f);
        if (rc) return rc;
        at += n;
    }
    return at == h->nents ? 0 : -3;
}
