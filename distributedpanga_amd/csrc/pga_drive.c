/* pga_drive.c -- the C host driver of the device path: one GDB genome file in,
 * its GIX index and seed streams out, through the C-ABI of libpga_hip.h the way
 * the reference's driver sequences its stages (src/pga-mpi-merge.c:331-548:
 * Read_GDB, extract_to_records, run_stage2, build_kmer_adapter,
 * self_adaptamer_merge, the 1-rank seed hand-over).
 *
 *   pga_drive [-f FREQ] [-p NPARTS] [-g] <in.gdb> <out_dir>
 *
 * reads <in.gdb> and the .bps file next to it; writes into out_dir:
 *   <root>.gix, .<root>.ktab.<k>     the GIX index (notes/GIXmake_dump.c:1632-1709)
 *   <root>.N.seeds, <root>.C.seeds   the same-strand and cross-strand seed
 *                                    streams, 1+2(P+C) bytes a seed
 *                                    (src/pga-mpi-merge.c:68-76)
 *   <root>.buck                      nbuck then cbuck, ncontig int64 each
 * and prints one line: records, seeds, P, C and the device milliseconds.
 * -g runs the build and merge as the captured device-driven pipeline.
 * Exit status 0 on success, 1 on a usage or file error, 2 on a device error. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "libpga_hip.h"
#include "pga_gdbfile.h"
#include "pga_gix.h"

static int write_file(const char *dir, const char *root, const char *ext, const void *p, size_t n)
{
    char path[4096];
    snprintf(path, sizeof(path), "%s/%s.%s", dir, root, ext);
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    const int ok = n == 0 || fwrite(p, 1, n, f) == n;
    return fclose(f) == 0 && ok ? 0 : -1;
}

/* "<dir>/<root>.gdb" -> "<root>" */
static void root_of(const char *path, char *root, size_t n)
{
    const char *b = strrchr(path, '/');
    b = b ? b + 1 : path;
    snprintf(root, n, "%s", b);
    char *dot = strrchr(root, '.');
    if (dot) *dot = '\0';
}

static int device_error(pga_ctx *ctx, const char *what, int rc)
{
    fprintf(stderr, "pga_drive: %s failed (%d): %s\n", what, rc, pga_last_error(ctx));
    pga_destroy(ctx);
    return 2;
}

int main(int argc, char **argv)
{
    int freq = 10, nparts = 1, graph = 0, a = 1;
    for (; a < argc && argv[a][0] == '-'; a++) {
        if (!strcmp(argv[a], "-f") && a + 1 < argc) freq = atoi(argv[++a]);
        else if (!strcmp(argv[a], "-p") && a + 1 < argc) nparts = atoi(argv[++a]);
        else if (!strcmp(argv[a], "-g")) graph = 1;
        else break;
    }
    if (argc - a != 2 || nparts < 1) {
        fprintf(stderr, "usage: pga_drive [-f FREQ] [-p NPARTS] [-g] <in.gdb> <out_dir>\n");
        return 1;
    }
    const char *in = argv[a], *out = argv[a + 1];
    char root[1024];
    root_of(in, root, sizeof(root));

    pga_gdbfile g;
    int rc = pga_gdbfile_read(in, &g);
    if (rc) {
        fprintf(stderr, "pga_drive: cannot read %s (%d)\n", in, rc);
        return 1;
    }
    const int64_t bps_bytes = g.boff[g.ncontig];
    uint8_t *bps = (uint8_t *) malloc((size_t) bps_bytes + 1);
    if (!bps || pga_gdbfile_read_bps(in, &g, bps)) {
        fprintf(stderr, "pga_drive: cannot read the .bps file of %s\n", in);
        free(bps);
        pga_gdbfile_free(&g);
        return 1;
    }

    pga_opts o;
    pga_default_opts(&o);
    o.freq = freq;
    pga_ctx *ctx = NULL;
    if ((rc = pga_create(&o, &ctx))) {
        fprintf(stderr, "pga_drive: pga_create failed (%d)\n", rc);
        free(bps);
        pga_gdbfile_free(&g);
        return 2;
    }
    rc = pga_load_gdb(ctx, bps, bps_bytes, g.boff, g.clen, g.ncontig);
    free(bps);
    if (rc) { pga_gdbfile_free(&g); return device_error(ctx, "pga_load_gdb", rc); }
    if (graph) {
        if ((rc = pga_run_graph(ctx))) { pga_gdbfile_free(&g); return device_error(ctx, "pga_run_graph", rc); }
    } else {
        if ((rc = pga_build_index(ctx))) { pga_gdbfile_free(&g); return device_error(ctx, "pga_build_index", rc); }
        if ((rc = pga_merge(ctx))) { pga_gdbfile_free(&g); return device_error(ctx, "pga_merge", rc); }
    }

    pga_index_meta im;
    pga_seeds_meta sm;
    if ((rc = pga_index_info(ctx, 0, &im)) || (rc = pga_seeds_info(ctx, 0, &sm))) {
        pga_gdbfile_free(&g);
        return device_error(ctx, "info", rc);
    }
    uint8_t *ent = (uint8_t *) malloc((size_t) (im.nrec * im.e) + 1);
    int64_t *cum = (int64_t *) malloc(sizeof(int64_t) * PGA_NUM_PREFIXES);
    uint8_t *ns = (uint8_t *) malloc((size_t) sm.nlen + 1), *cs = (uint8_t *) malloc((size_t) sm.clen + 1);
    int64_t *buck = (int64_t *) malloc(sizeof(int64_t) * 2 * (size_t) (sm.ncontig > 0 ? sm.ncontig : 1));
    int *perm = (int *) malloc(sizeof(int) * (size_t) (g.ncontig > 0 ? g.ncontig : 1));
    int status = 0;
    if (!ent || !cum || !ns || !cs || !buck || !perm) {
        fprintf(stderr, "pga_drive: out of host memory\n");
        status = 1;
    } else if ((rc = pga_index_export(ctx, 0, ent, cum)) ||
               (rc = pga_seeds_export(ctx, 0, ns, cs, buck, buck + sm.ncontig))) {
        status = device_error(ctx, "export", rc);
        ctx = NULL;
    } else {
        for (int i = 0; i < g.ncontig; i++) perm[i] = i;      /* contig ids as the GDB numbers them */
        if (pga_gix_write(out, root, nparts, ent, im.nrec, im.P, im.C, cum, im.maxpre, im.ncontig, perm) ||
            write_file(out, root, "N.seeds", ns, (size_t) sm.nlen) ||
            write_file(out, root, "C.seeds", cs, (size_t) sm.clen) ||
            write_file(out, root, "buck", buck, sizeof(int64_t) * 2 * (size_t) sm.ncontig)) {
            fprintf(stderr, "pga_drive: cannot write into %s\n", out);
            status = 1;
        } else {
            pga_timings t;
            pga_timings_get(ctx, &t);
            printf("pga_drive: %s: %lld bases, %lld records, %lld N + %lld C seeds, P=%d C=%d, %.3f ms on the device\n",
                   root, (long long) im.seqtot, (long long) im.nrec, (long long) (sm.nlen / sm.s),
                   (long long) (sm.clen / sm.s), im.P, im.C,
                   graph ? t.total_ms : t.total_ms + t.merge_ms);
        }
    }
    free(ent); free(cum); free(ns); free(cs); free(buck); free(perm);
    pga_gdbfile_free(&g);
    if (ctx) pga_destroy(ctx);
    return status;
}
