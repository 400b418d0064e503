/* pga_gix.h -- the on-disk GIX index (`<root>.gix` + `.<root>.ktab.<part>`).
 *
 * Layout of the files GIXmake writes (notes/GIXmake_dump.c:1632-1709), which
 * the reference's readers take apart (src/fastga_order_core.c:11-49):
 *
 *   .<root>.ktab.<p>   (p = 1..nparts)
 *       int     kmer (40)
 *       int64   nents
 *       nents x entry      entry = 7 k-mer suffix bytes, mask 0, LCP, pos LE (P),
 *                          contig|strand LE (C): the GIX entry of
 *                          src/kmer_adapter.c:55-62, e = 9+P+C bytes
 *   <root>.gix
 *       int     kmer, nparts, 1, 3
 *       int64   prefix[2^24]      entries with 24-bit prefix <= p (cumulative)
 *       int     P (PostBytes), C (ContBytes), nparts
 *       int64   maxpre            largest prefix group
 *       int     freq (0), ncontig
 *       int     perm[ncontig]     contig order
 *       int64   -1
 *
 * Parts hold consecutive entries; every part boundary falls on a 24-bit prefix
 * boundary.  All integers are little-endian host ints (as GIXmake writes them).
 * Host C; every call returns 0, or -1 (I/O or allocation, errno set) or -3
 * (bad argument or file contents).
 */
#ifndef PGA_GIX_H
#define PGA_GIX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int     kmer;        /* 40                                       */
    int     nparts;
    int     P, C;        /* position / contig+strand bytes            */
    int64_t maxpre;
    int     freq;
    int     ncontig;
    int64_t nents;       /* sum over the parts (filled by the reader) */
} pga_gix_header;

/* Write an index: entries[nrec*(9+P+C)], cum[2^24], perm[ncontig] (NULL =
 * identity) into dir/<root>.gix and dir/.<root>.ktab.1..nparts. */
int pga_gix_write(const char *dir, const char *root, int nparts, const uint8_t *entries, int64_t nrec,
                  int P, int C, const int64_t *cum, int64_t maxpre, int ncontig, const int *perm);

/* Read the header (and, when non-NULL, the prefix table [2^24] and perm
 * [ncontig]); h->nents is the total over the part files. */
int pga_gix_read_header(const char *dir, const char *root, pga_gix_header *h, int64_t *cum, int *perm);

/* Read every part's entries, concatenated, into entries (capacity cap bytes). */
int pga_gix_read_entries(const char *dir, const char *root, const pga_gix_header *h, uint8_t *entries,
                         int64_t cap);

#ifdef __cplusplus
}
#endif
#endif
