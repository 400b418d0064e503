/* pga_gdbfile.h -- GDB genome files: the `<root>.gdb` skeleton (ONE text format)
 * and the `.<root>.bps` sequence file.
 *
 * What the reference's Read_GDB / Write_GDB keep in them (lib/GDB.c:1171-1400,
 * 1519-1607; schema lib/GDB.c:37-47):
 *
 *   <root>.gdb      ONE text file of type "gdb": a header (file-type line,
 *                   '!' provenance, '<' reference, '~' schema and '.' lines),
 *                   then data lines
 *                     f <4 reals>        base frequencies A, C, G, T
 *                     S <len> <chars>    a scaffold and its id
 *                     G <int>            a gap of that length inside the scaffold
 *                     C <int>            a contig of that length
 *                   (u and M lines are accepted and ignored, as Read_GDB does)
 *   .<root>.bps     every contig's bases, 2 bits each, LSB-first, each contig
 *                   starting on a byte: contig c occupies (clen[c]+3)/4 bytes
 *                   from boff[c]
 *
 * The same skeleton saved as `<root>.1gdb` is a binary ONE file (the same
 * header, a "$ 0" line, then one binary record per data line; lib/ONElib.c):
 * pga_gdbfile_read reads it too, with its `.<root>.bps`.  Binary mask lists
 * are skipped as their text form is; scaffold ids that ONE compressed with a
 * trained codec (it starts after 100 KB of ids) are not decoded and the read
 * returns -2.  Host C; every call returns 0, or -1 (I/O or allocation, errno
 * set), -2 (codec-compressed ids) or -3 (bad argument or file contents).  The
 * tables feed pga_load_gdb (libpga_hip.h) directly: the .bps bytes, boff and
 * clen.
 */
#ifndef PGA_GDBFILE_H
#define PGA_GDBFILE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int      nscaff, ncontig;
    int64_t  seqtot;     /* bases over all contigs                        */
    int64_t  maxctg;     /* longest contig                                */
    float    freq[4];    /* A, C, G, T                                    */
    int64_t *slen;       /* [nscaff]      scaffold length, gaps included  */
    int     *fctg;       /* [nscaff + 1]  first contig of each scaffold   */
    int64_t *hoff;       /* [nscaff]      offset of its id in headers     */
    char    *headers;    /* the ids, each NUL-terminated                  */
    int64_t  hdrtot;     /* bytes in headers                              */
    int64_t *sbeg;       /* [ncontig]     start of the contig in its scaffold */
    int64_t *clen;       /* [ncontig]                                     */
    int64_t *boff;       /* [ncontig + 1] byte offset in the .bps file    */
} pga_gdbfile;

/* Read the skeleton of path (".../<root>.gdb" or ".../<root>.1gdb") into g
 * (arrays malloc'd). */
int pga_gdbfile_read(const char *path, pga_gdbfile *g);

/* Read the .bps file next to path (".../.<root>.bps") into bps, which holds
 * g->boff[g->ncontig] bytes. */
int pga_gdbfile_read_bps(const char *path, const pga_gdbfile *g, uint8_t *bps);

/* Write g as path (".../<root>.gdb", ONE text) and its sequence bps
 * (g->boff[g->ncontig] bytes) as ".../.<root>.bps".  g->boff, seqtot and
 * maxctg are not read: offsets follow from the contig lengths. */
int pga_gdbfile_write(const char *path, const pga_gdbfile *g, const uint8_t *bps);

/* Free the arrays pga_gdbfile_read allocated. */
void pga_gdbfile_free(pga_gdbfile *g);

#ifdef __cplusplus
}
#endif
#endif
