#!/bin/sh
# Build the reference's extraction / sort / LCP / GIX-adapter path from its own
# sources where they lie under /root/reference, plus oracle/ref_harness.c, into
# oracle/_ref/libpgaref.so.  Test infrastructure only; the output is git-ignored.
#
# Files: src/{extract,pack,tables,work_split,record,sort,kmer_adapter}.c and the
# vendored lib/{GDB,gene_core,ONElib,ANO,MSDsort,RSDsort}.c, with the reference
# Makefile's flags (-O3 -fno-strict-aliasing -DLCPs).  The merge
# (src/fastga_pipeline.c) is left out: it links src/seed_buffer.c, which
# includes <mpi.h>, and no MPI is installed.
set -e
REF=${PGA_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -f "$REF/src/extract.c" ]; then
    echo "ref_build: reference sources not found at $REF; skipping" >&2
    exit 0
fi
mkdir -p "$OUT"
gcc -O3 -fPIC -shared -w -fno-strict-aliasing -DLCPs \
    -I"$REF/lib" -I"$REF/src" \
    -o "$OUT/libpgaref.so.tmp" \
    "$HERE/ref_harness.c" \
    "$REF/src/extract.c" "$REF/src/pack.c" "$REF/src/tables.c" "$REF/src/work_split.c" \
    "$REF/src/record.c" "$REF/src/sort.c" "$REF/src/kmer_adapter.c" \
    "$REF/lib/GDB.c" "$REF/lib/gene_core.c" "$REF/lib/ONElib.c" "$REF/lib/ANO.c" \
    "$REF/lib/MSDsort.c" "$REF/lib/RSDsort.c" \
    -lpthread -lm -lz
mv "$OUT/libpgaref.so.tmp" "$OUT/libpgaref.so"
echo "ref_build: $OUT/libpgaref.so"
