// pga_device.h -- types shared by the kernels and the C-ABI host code.
#pragma once
#include <stdint.h>

namespace pga {

constexpr int MAX_FREQ = 64;     // largest frequency cut the merge window supports

// Digit counters cleared by the extraction emit between rounds: with NT threads,
// threads 64..NT-1 clear digits tid-64, tid-64+(NT-64), ... below 256 while the
// first wave scans.  True iff that covers all 256 digits (checked at compile
// time for every instantiated NT, and on the host for every NT by the tests).
constexpr bool emit_clear_covers(int nt)
{
    if (nt <= 64) return false;
    bool seen[256] = {};
    for (int tid = 64; tid < nt; tid++)
        for (int d = tid - 64; d < 256; d += nt - 64) seen[d] = true;
    for (int d = 0; d < 256; d++)
        if (!seen[d]) return false;
    return true;
}

struct Rec {                     // 16-byte device record (see pga_kernels.hip)
    uint64_t w0, w1;
};

struct ContigDev {
    int64_t bitoff;              // bit offset of base 0 in the device .bps image
    int64_t pos0;                // first flat candidate position of the contig
    int64_t ncand;               // candidate offsets: max(0, clen - 11)
    int32_t clen;
    int32_t local;               // contig id inside its GDB
    int32_t pair;
    int32_t pad;
};

// An extraction block's (XT*XU positions) first contig, as the extraction
// kernels need it: the contig (the last whose first candidate position is at
// or before the block's), the .bps bit and contig offset of the block's first
// position, the contig's local id, and the block-relative positions [lo, hi)
// that lie inside the contig at offset 28 or more (zero when none do).
struct BlkDev {
    int64_t bbit;
    uint32_t jblk;
    uint32_t local;
    uint32_t frange;             // lo | hi << 16
    int32_t c;
};

struct PairDev {
    int64_t rec0, nrec;          // the pair's records in the global record space
    int64_t ent_off;             // byte offset of its GIX entries
    int64_t cum_off;             // element offset of its prefix table
    int64_t buck_off;            // offset of its contigs in the buck vectors
    int64_t nseed_off, cseed_off;// byte offsets of its seed streams
    int64_t mblk0;               // first merge block
    int32_t P, C, e, s;
    int32_t ncontig;
    int32_t pad;
};

struct Seg {                     // a key range still to be partitioned
    int64_t start, count;
    int32_t pair, level;         // level = digit (1..16) to split on next
    int32_t buf;                 // record buffer (0/1) holding its records
    int32_t skip;                // bit 0, set by the plan: one digit only, left in place;
                                 // bit 1: the previous level's scatter moved it and wrote its digits
};

struct Tile {                    // a bucket run small enough for one workgroup sort
    int64_t start;
    int32_t count, pair, buf;
    int32_t level;               // radix digit (1..16) its buckets differ in; all share the digits above
    int32_t drange;              // its smallest and largest digit at that level: lo | hi << 16
};

constexpr int32_t drange_all(int bits) { return ((1 << bits) - 1) << 16; }   // every digit of a level

struct XRange {                  // a run of extraction blocks inside one pair, one workgroup's share
    int64_t lb0, lb1;            // logical blocks [lb0, lb1)
    int32_t pair, pad;
};

constexpr int GRAPH_LEVELS = 8;  // partition levels the device-driven pipeline runs

// Status of a device-driven build + merge, written on the device and read by
// the host after it (pga_run_graph).
struct GStat {
    int nsegs[GRAPH_LEVELS + 2]; // segments split at level L (nsegs[GRAPH_LEVELS+1] > 0: not converged)
    int ntiles;                  // sort tiles
    int overflow;                // seed streams larger than their buffers: nothing emitted
    int bad;                     // the extraction passes disagreed on a first digit
    int capped;                  // a segment or tile count reached its list's bound
    int pad[2];
    unsigned long long kept;     // records the partition left in place (unused here)
    int64_t R, nmb;              // records, merge tiles
};

} // namespace pga
