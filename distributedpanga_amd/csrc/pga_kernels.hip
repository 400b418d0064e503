// pga_kernels.hip -- CDNA4 (gfx950) kernels of the index build and adaptamer merge.
//
// Device record (16 B, one per emitted k-mer occurrence):
//   w0 = k-mer bases 0..31, MSB-first (base 0 in bits 63..62)
//   w1 = bases 32..39 in bits 63..48 | payload in bits 47..0,
//        payload = contig << (jbits+1) | j << 1 | strand
// where j is the syncmer offset (the stored position is j + 12*strand,
// src/extract.c:200-228).  Comparing (w0, w1) as one 128-bit number orders by
// k-mer, then contig, offset and forward-before-reverse: the extraction order,
// which is the canonical tie order for equal k-mers.  Keys are unique, so any
// correct sort gives the same array; no stage needs to be stable.
//
// Pipeline: syncmer scan (count with first-digit counts per range of blocks,
// scan, emit straight into the first-level buckets) -> MSD radix partition
// (the first level only planned; a second digit of 8..10 bits, then 8-bit
// digits, until every bucket fits one workgroup's tile) -> per-tile binned LDS
// sort that also writes LCPs and GIX entries -> prefix table -> merge.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
#include "pga_device.h"

namespace pga {

__constant__ uint8_t c_tmap[256];

// ---------------------------------------------------------------- helpers

__device__ __forceinline__ uint64_t field_rev64(uint64_t x)
{
    uint64_t y = __builtin_bitreverse64(x);
    return ((y >> 1) & 0x5555555555555555ull) | ((y & 0x5555555555555555ull) << 1);
}

__device__ __forceinline__ uint32_t field_rev16(uint32_t x)
{
    uint32_t y = __builtin_bitreverse32(x) >> 16;
    return ((y >> 1) & 0x5555u) | ((y & 0x5555u) << 1);
}

// 128 bits (64 bases, LSB-first as in .bps) starting at bit offset `bit`.
__device__ __forceinline__ void window128(const uint64_t *__restrict__ b, int64_t bit, uint64_t &lo, uint64_t &hi)
{
    const int64_t w = bit >> 6;
    const int s = (int) (bit & 63);
    const uint64_t x = b[w], y = b[w + 1];
    if (s) {
        const uint64_t z = b[w + 2];
        lo = (x >> s) | (y << (64 - s));
        hi = (y >> s) | (z << (64 - s));
    } else {
        lo = x;
        hi = y;
    }
}

// 80 bits (40 bases, LSB-first as in .bps) starting at bit offset `bit`.
__device__ __forceinline__ void window80(const uint64_t *__restrict__ b, int64_t bit,
                                         uint64_t &lo, uint32_t &hi)
{
    int64_t w = bit >> 6;
    int s = (int) (bit & 63);
    uint64_t x = b[w], y = b[w + 1];
    if (s) {
        uint64_t z = b[w + 2];
        lo = (x >> s) | (y << (64 - s));
        hi = (uint32_t) (((y >> s) | (z << (64 - s))) & 0xFFFFu);
    } else {
        lo = x;
        hi = (uint32_t) (y & 0xFFFFu);
    }
}

// Canonical hash of the 8-mer at base t of W (LSB-first), src/extract.c:85-249:
// min(forward, reverse complement) of (TMap[4-mer] << 8 | TMap[4-mer]).
// tl[x] = TMap[field-reversed x], tm[x] = TMap[x]; the reverse-complement
// 4-mer of an LSB-first byte x, packed MSB-first, is ~x.
__device__ __forceinline__ uint32_t hash8(uint64_t W, int t, const uint8_t *tl, const uint8_t *tm)
{
    const uint32_t lo = (uint32_t) (W >> (2 * t)) & 0xFFu;
    const uint32_t hi = (uint32_t) (W >> (2 * t + 8)) & 0xFFu;
    const uint32_t fw = ((uint32_t) tl[lo] << 8) | tl[hi];
    const uint32_t rv = ((uint32_t) tm[hi ^ 0xFFu] << 8) | tm[lo ^ 0xFFu];
    return fw < rv ? fw : rv;
}

// Closed (12,8)-syncmer test on the 12 bases in bits 0..23 of W (LSB-first):
// the earliest or the last of the five 8-mer hashes is the minimum.
__device__ __forceinline__ bool syncmer(uint64_t W, const uint8_t *tl, const uint8_t *tm)
{
    uint32_t h[5];
#pragma unroll
    for (int t = 0; t < 5; t++) h[t] = hash8(W, t, tl, tm);
    uint32_t m = min(min(min(h[0], h[1]), min(h[2], h[3])), h[4]);
    return h[0] == m || h[4] == m;
}

// byte tables: 256 entries in 64 LDS words, so random lookups meet few bank
// conflicts
__device__ __forceinline__ void load_tables(uint8_t *tl, uint8_t *tm)
{
    for (int t = threadIdx.x; t < 256; t += blockDim.x) {
        uint32_t r = ((t & 3) << 6) | (((t >> 2) & 3) << 4) | (((t >> 4) & 3) << 2) | ((t >> 6) & 3);
        tl[t] = c_tmap[r];
        tm[t] = c_tmap[t];
    }
}

template <int NT>
__device__ __forceinline__ uint64_t block_sum_u64(uint64_t v, uint64_t *red)
{
    // wave64 reduction through LDS
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0) red[wv] = v;
    __syncthreads();
    uint64_t s = 0;
    if (threadIdx.x == 0) {
        for (int k = 0; k < NT / 64; k++) s += red[k];
        red[NT / 64] = s;
    }
    __syncthreads();
    s = red[NT / 64];
    __syncthreads();
    return s;
}

// Exclusive scan of one value per thread over the block (NT threads); returns
// the thread's exclusive prefix, *total gets the block total.
template <int NT, typename T>
__device__ __forceinline__ T block_excl_scan(T v, T *sh, T *total)
{
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    T x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wv] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        T run = 0;
        for (int k = 0; k < NT / 64; k++) { T t = sh[k]; sh[k] = run; run += t; }
        sh[NT / 64] = run;
    }
    __syncthreads();
    T r = x - v + sh[wv];
    *total = sh[NT / 64];
    __syncthreads();
    return r;
}

typedef unsigned __int128 u128;

// A record through a pointer that was itself read from device memory (the
// compiler cannot tell its address space, and would use a flat load, which
// also counts against the LDS wait counter): load it as global memory.
__device__ __forceinline__ Rec load_rec_global(const Rec *p)
{
    typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
    const u64x2 v = *(const __attribute__((address_space(1))) u64x2 *) p;
    Rec r;
    r.w0 = v.x;
    r.w1 = v.y;
    return r;
}

// Write n fixed-width items (s <= 16 bytes each, item j held little-endian in
// slot[j]) as one packed byte run starting at g: every lane assembles one
// aligned 16-byte chunk of the run from the (at most three) slots it
// overlaps; the partial chunks at the two ends are written byte by byte.
template <int NT>
__device__ __forceinline__ void write_packed(uint8_t *__restrict__ g, const u128 *__restrict__ slot, int n, int s)
{
    if (n <= 0) return;
    const int nb = n * s;
    const int lead = (int) ((uintptr_t) g & 15);        // the run starts at this offset of chunk 0
    // Addresses derive from g by pointer arithmetic (never from an integer) so
    // that the stores stay global ones: flat stores would also count against
    // the LDS wait counter and make every LDS wait wait for them.
    uint8_t *const a0 = g - lead;
    const int nchunks = (lead + nb + 15) >> 4;
    for (int c = threadIdx.x; c < nchunks; c += NT) {
        const int o0 = 16 * c - lead;                   // run offset of the chunk's first byte
        const int o = o0 < 0 ? 0 : o0;
        int j = o / s;
        const int t = o - j * s;
        u128 acc = slot[j] >> (8 * t);
        int filled = s - t;
        for (j++; filled < 16 && j < n; j++) {
            acc |= slot[j] << (8 * filled);
            filled += s;
        }
        if (o0 < 0) acc <<= 8 * (-o0);                  // chunk 0 starts before the run
        uint8_t *dst = a0 + 16 * c;
        if (o0 >= 0 && o0 + 16 <= nb) {
            uint4 w;
            const uint64_t lo = (uint64_t) acc, hi = (uint64_t) (acc >> 64);
            w.x = (uint32_t) lo; w.y = (uint32_t) (lo >> 32); w.z = (uint32_t) hi; w.w = (uint32_t) (hi >> 32);
            *reinterpret_cast<uint4 *>(dst) = w;
        } else {
#pragma unroll
            for (int b = 0; b < 16; b++)
                if (o0 + b >= 0 && o0 + b < nb) dst[b] = (uint8_t) (acc >> (8 * b));
        }
    }
}

// ---------------------------------------------------------------- extraction

// Positions per thread and threads per block of the syncmer scan.  Flat
// position f enumerates every offset j in [0, clen-12] of every contig, pair
// after pair; each pair's range starts on a block boundary.
constexpr int XT = 256;
constexpr int XU = 4;

// First radix digit (the key's top byte) of the records at an offset whose
// LSB-first window starts at W: the forward key begins with bases j..j+3,
// MSB-first; the reverse-complement key with the complements of j+11..j+8.
__device__ __forceinline__ uint32_t digit_fwd(uint64_t W)
{
    const uint32_t x = (uint32_t) W & 0xFFu;
    return ((x & 3u) << 6) | ((x & 0xCu) << 2) | ((x >> 2) & 0xCu) | (x >> 6);
}

__device__ __forceinline__ uint32_t digit_rev(uint64_t W) { return ~(uint32_t) (W >> 16) & 0xFFu; }

// 5-base bucket (record_bucket, src/dsort.h:30-33: the key's top 10 bits) of
// the same records: the first digit and the next base (j+4 forward; the
// complement of j+7 for the reverse complement).
__device__ __forceinline__ uint32_t bucket_fwd(uint64_t W) { return (digit_fwd(W) << 2) | ((uint32_t) (W >> 8) & 3u); }
__device__ __forceinline__ uint32_t bucket_rev(uint64_t W) { return (digit_rev(W) << 2) | (~(uint32_t) (W >> 14) & 3u); }

struct XEmit {
    uint64_t f0, f1, r0, r1;     // forward and reverse-complement record keys
    uint32_t dg;                 // first radix digits: forward (bits 0-7), reverse complement (8-15)
    uint32_t bk;                 // 5-base buckets: forward (bits 0-9), reverse complement (10-19)
    bool fwd, rev;
};

// Syncmer test at offset j of contig c; with keys, also the records' keys.
__device__ __forceinline__ int scan_position(const uint64_t *__restrict__ bps64,
                                             const ContigDev &c, int64_t j, int jbits,
                                             const uint8_t *tl, const uint8_t *tm,
                                             bool keys, XEmit &out)
{
    uint64_t W0;
    uint32_t W1;
    out.fwd = out.rev = false;
    window80(bps64, c.bitoff + 2 * j, W0, W1);
    if (!syncmer(W0, tl, tm)) return 0;
    out.dg = digit_fwd(W0) | (digit_rev(W0) << 8);
    out.bk = bucket_fwd(W0) | (bucket_rev(W0) << 10);
    out.fwd = j <= (int64_t) c.clen - 40;          // forward 40-mer fits
    out.rev = j >= 28;                             // reverse complement of [j-28, j+12) fits
    if (keys) {
        const uint64_t pay = ((uint64_t) c.local << (jbits + 1)) | ((uint64_t) j << 1);
        // forward: MSB-first packing of [j, j+40) = field-reversed LSB-first window
        out.f0 = field_rev64(W0);
        out.f1 = ((uint64_t) field_rev16(W1) << 48) | pay;
        if (out.rev) {
            // reverse complement: the MSB-first RC 40-mer is the bitwise NOT of
            // the LSB-first window of [j-28, j+12)
            uint64_t V0;
            uint32_t V1;
            window80(bps64, c.bitoff + 2 * (j - 28), V0, V1);
            out.r0 = ~(((uint64_t) V1 << 48) | (V0 >> 16));
            out.r1 = ((uint64_t) ((~V0) & 0xFFFFu) << 48) | pay | 1ull;
        }
    }
    return (int) out.fwd + (int) out.rev;
}

// Keys of the records at an offset already known to be selected (e.fwd / e.rev set).
__device__ __forceinline__ void record_keys(const uint64_t *__restrict__ bps64, const ContigDev &c, int64_t j,
                                            int jbits, XEmit &out)
{
    const uint64_t pay = ((uint64_t) c.local << (jbits + 1)) | ((uint64_t) j << 1);
    if (out.fwd) {
        uint64_t W0;
        uint32_t W1;
        window80(bps64, c.bitoff + 2 * j, W0, W1);
        out.f0 = field_rev64(W0);
        out.f1 = ((uint64_t) field_rev16(W1) << 48) | pay;
    }
    if (out.rev) {
        uint64_t V0;
        uint32_t V1;
        window80(bps64, c.bitoff + 2 * (j - 28), V0, V1);
        out.r0 = ~(((uint64_t) V1 << 48) | (V0 >> 16));
        out.r1 = ((uint64_t) ((~V0) & 0xFFFFu) << 48) | pay | 1ull;
    }
}

// Per range of logical blocks (a workgroup's share, inside one pair): the
// selection bits of every position (forward / reverse-complement record), the
// record count of every block, and the range's count of each first radix
// digit.  blk[LB].c is the contig of block LB's first position; every thread
// steps forward from it (a thread of XU positions crosses few contig starts).
// Only records whose 5-base bucket lies in [bk_lo, bk_hi) are selected (a
// shard of the index; [0, 1024) keeps all).
__global__ __launch_bounds__(XT) void k_extract_count(const uint64_t *__restrict__ bps64,
                                                      const ContigDev *__restrict__ ctg,
                                                      const BlkDev *__restrict__ blk, int nctg,
                                                      int64_t F, const XRange *__restrict__ ranges, int nrange,
                                                      uint8_t *__restrict__ sel,
                                                      uint32_t *__restrict__ rangehist, uint32_t bk_lo, uint32_t bk_hi)
{
    const bool shard = bk_lo > 0 || bk_hi < 1024;
    static_assert(XT == 256, "one thread per digit");
    __shared__ uint8_t tl[256], tm[256];
    __shared__ uint32_t hist[256];
    load_tables(tl, tm);
    for (int r = blockIdx.x; r < nrange; r += gridDim.x) {
    hist[threadIdx.x] = 0;
    __syncthreads();
    const XRange xr = ranges[r];
    // the next block's first contig is loaded a block ahead, and a contig
    // together with the next one's start
    int cnext = xr.lb0 < xr.lb1 ? blk[xr.lb0].c : 0;
    for (int64_t LB = xr.lb0; LB < xr.lb1; LB++) {
    const int64_t fb = LB * XT * XU;
    int c = cnext;
    if (LB + 1 < xr.lb1) cnext = blk[LB + 1].c;
    const int64_t f0 = fb + (int64_t) threadIdx.x * XU;
    uint32_t bits = 0;
    ContigDev c0 = ctg[c];
    while (c + 1 < nctg) {
        const ContigDev nx = ctg[c + 1];
        if (nx.pos0 > f0) break;
        c++;
        c0 = nx;
    }
    if (f0 >= c0.pos0 && f0 + XU <= c0.pos0 + c0.ncand) {
        // the thread's XU positions lie in one contig: the 8-mers at
        // j0 .. j0+XU+3 are hashed once and shared by the XU windows
        const int64_t j0 = f0 - c0.pos0;
        uint64_t W0;
        uint32_t W1;
        window80(bps64, c0.bitoff + 2 * j0, W0, W1);
        uint32_t h[XU + 4];
#pragma unroll
        for (int t = 0; t < XU + 4; t++) h[t] = hash8(W0, t, tl, tm);
#pragma unroll
        for (int u = 0; u < XU; u++) {
            const uint32_t m = min(min(min(h[u], h[u + 1]), min(h[u + 2], h[u + 3])), h[u + 4]);
            if (h[u] != m && h[u + 4] != m) continue;
            const int64_t j = j0 + u;
            bool fwd = j <= (int64_t) c0.clen - 40, rev = j >= 28;
            if (shard) {
                const uint32_t bf = bucket_fwd(W0 >> (2 * u)), br = bucket_rev(W0 >> (2 * u));
                fwd = fwd && bf >= bk_lo && bf < bk_hi;
                rev = rev && br >= bk_lo && br < bk_hi;
            }
            bits |= ((fwd ? 1u : 0u) | (rev ? 2u : 0u)) << (2 * u);
            if (fwd) atomicAdd(&hist[digit_fwd(W0 >> (2 * u))], 1u);
            if (rev) atomicAdd(&hist[digit_rev(W0 >> (2 * u))], 1u);
        }
    } else {
        XEmit x;
#pragma unroll
        for (int u = 0; u < XU; u++) {
            const int64_t f = f0 + u;
            if (f >= F) break;
            while (c + 1 < nctg && ctg[c + 1].pos0 <= f) c++;
            const ContigDev cc = ctg[c];
            if (f >= cc.pos0 + cc.ncand) continue;         // padding before the next pair
            scan_position(bps64, cc, f - cc.pos0, 0, tl, tm, false, x);
            if (shard) {
                const uint32_t bf = x.bk & 1023u, br = x.bk >> 10;
                x.fwd = x.fwd && bf >= bk_lo && bf < bk_hi;
                x.rev = x.rev && br >= bk_lo && br < bk_hi;
            }
            bits |= ((x.fwd ? 1u : 0u) | (x.rev ? 2u : 0u)) << (2 * u);
            if (x.fwd) atomicAdd(&hist[x.dg & 0xFFu], 1u);
            if (x.rev) atomicAdd(&hist[x.dg >> 8], 1u);
        }
    }
    sel[fb / XU + threadIdx.x] = (uint8_t) bits;       // the emit pass reads the selection back
    }
    __syncthreads();
    rangehist[(size_t) r * 256 + threadIdx.x] = hist[threadIdx.x];
    __syncthreads();
    }
}

// Per pair and 16 digits (blockIdx.y), one thread per (digit, lane of 16):
// where each range's records of a digit start inside the pair's bucket of
// that digit (records of the digit in the pair's earlier ranges), and the
// pair's digit totals (hist1).  Each lane takes a contiguous share of the
// pair's ranges.
__global__ __launch_bounds__(256) void k_xrange_scan(const int *__restrict__ prange,
                                                     const uint32_t *__restrict__ rangehist,
                                                     int64_t *__restrict__ rangeoff,
                                                     unsigned long long *__restrict__ hist1)
{
    __shared__ int64_t part[16][17];
    const int p = blockIdx.x;
    const int dg = threadIdx.x & 15, lane = threadIdx.x >> 4;
    const int d = blockIdx.y * 16 + dg;
    const int r0 = prange[p], n = prange[p + 1] - r0, K = (n + 15) / 16;
    const int a = r0 + min(n, lane * K), b = r0 + min(n, lane * K + K);
    int64_t s = 0;
#pragma unroll 8
    for (int r = a; r < b; r++) s += rangehist[(size_t) r * 256 + d];
    part[lane][dg] = s;
    __syncthreads();
    int64_t run = 0;
    for (int l = 0; l < lane; l++) run += part[l][dg];
    if (lane == 15) hist1[(size_t) p * 256 + d] = (unsigned long long) (run + s);
    for (int r = a; r < b; r++) {
        rangeoff[(size_t) r * 256 + d] = run;
        run += rangehist[(size_t) r * 256 + d];
    }
}

// Per pair: the start of each digit's bucket, pr[p] + the pair's records of
// smaller digits.
__global__ __launch_bounds__(256) void k_xrange_base(const int64_t *__restrict__ pr,
                                                     const unsigned long long *__restrict__ hist1,
                                                     int64_t *__restrict__ bstart)
{
    __shared__ int64_t sh[256 / 64 + 1];
    const int p = blockIdx.x;
    int64_t tot;
    const int64_t o = block_excl_scan<256, int64_t>((int64_t) hist1[(size_t) p * 256 + threadIdx.x], sh, &tot);
    bstart[(size_t) p * 256 + threadIdx.x] = pr[p] + o;
}

// Emit every selected record straight into its first-level bucket.  Range r's
// records of digit d fill the run of rangehist[r][d] slots at bstart[pair][d]
// + rangeoff[r][d].  A workgroup of 256*QB threads takes QB blocks per round:
// each thread makes its (at most 2*XU) records' first digits and takes a rank
// in each digit (an LDS counter); a scan of the round's digit counts lays the
// records out in digit order in an LDS stage, and the stage goes out as one
// contiguous run per digit at the digit's next free slots.  A round's run of a
// digit is short (about six records), so its last HOLD-aligned segment is
// usually incomplete: thread d keeps digit d's records past the run's last
// whole segment in LDS and puts them in front of the digit's records in the
// next round's stage, so that the bucket is written in whole aligned segments
// (the range's last round writes everything; partial lines written at
// different times reach HBM as separate partial writes).  With dig (an 8-bit
// second digit), each record's second digit goes into dig at the record's
// slot, for the second partition level's counts.  A record past its digit's
// run (the two passes disagreeing on a digit) is not written and sets *bad.
constexpr int EMIT_QB = 2;           // blocks per emit round (512 threads, 48 KB stage, two workgroups a CU)
constexpr int EMIT_HOLD = 2;         // records per aligned bucket segment (32 B: measured faster than
                                     // 64 B, whose longer carries cost more than the writes they save)

template <int QB>
__global__ __launch_bounds__(256 * QB, 4) void k_extract_emit(const uint64_t *__restrict__ bps64,
                                                           const ContigDev *__restrict__ ctg,
                                                           const BlkDev *__restrict__ blk, int nctg,
                                                           int jbits, const uint8_t *__restrict__ sel,
                                                           const XRange *__restrict__ ranges, int nrange,
                                                           const uint32_t *__restrict__ rangehist,
                                                           const int64_t *__restrict__ rangeoff,
                                                           const int64_t *__restrict__ bstart,
                                                           Rec *__restrict__ out, uint8_t *__restrict__ dig,
                                                           int *__restrict__ bad)
{
    constexpr int NT = 256 * QB;
    constexpr int NS = 2 * XU;                         // record slots per thread
    constexpr int HOLD = EMIT_HOLD;
    static_assert(XT == 256 && NT >= 256, "one thread per digit");
    static_assert(emit_clear_covers(NT), "the counter clearing must reach every digit");
    static_assert((HOLD & (HOLD - 1)) == 0 && HOLD >= 1 && HOLD <= 8, "segment of 1-8 records");
    constexpr int SCAP = 3 * 1024;                     // stage records (a round holds about 1600)
    __shared__ Rec stage[SCAP];
    __shared__ Rec held[256][HOLD > 1 ? HOLD - 1 : 1];  // a digit's records past its last whole segment
    __shared__ int64_t base[256], gpos[256];           // a digit's run in the range; this round's slot (less the stage index)
    __shared__ uint32_t cnt[2][256], loff[256], lim[256], done[256];
    __shared__ uint32_t hcnt[256], wcnt[256], wout[256];   // held records; records out this round (in bounds)
    __shared__ uint32_t stot, spill;
    const int tid = threadIdx.x, q = tid >> 8, t = tid & 255;
    uint32_t hn = 0;                                   // threads below 256: digit t's held records
    // Inputs of a round.  Two rounds ahead: the selection bits and the
    // block's first contig (fetch_s).  At the start of the round before, from
    // those: whether the thread's XU positions lie in that contig past base
    // 28, and where their forward and reverse-complement 128-bit spans start
    // (plan_w).  In the round before, once its own words are used: the words
    // under the spans (fetch_w).  Every load is issued unconditionally (a
    // clamped block, a harmless address), so no load waits where it is issued;
    // they are in flight while the round before writes its records.
    uint32_t bits_n = 0;
    int64_t lb_n = 0;
    BlkDev bk_n{};
    uint32_t ninfo = 0, njl = 0, nloc = 0;             // the next round's: bits | fast << 8 | shift << 9
    int64_t nwb = 0;                                   // and the first of its four words
    uint64_t w0 = 0, w1 = 0, w2 = 0, w3 = 0;
    // the addresses of the loads in flight, kept until the loads are used:
    // the compiler waits for an outstanding load before it reuses the
    // registers of its address
    const uint8_t *as = sel;
    const BlkDev *ab = blk;
    const uint64_t *aw = bps64;
    auto fetch_s = [&](int64_t LB, int64_t lb1) {
        const int64_t L = LB < lb1 ? LB : lb1 - 1;     // lb1 > lb0: in range
        lb_n = LB;
        as = sel + L * XT + t;
        ab = blk + L;
        bits_n = *as;
        bk_n = *ab;
    };
    auto plan_w = [&](int64_t lb1) {
        const uint32_t b = lb_n < lb1 ? bits_n : 0u;
        const uint32_t p = (uint32_t) t * XU;
        const bool f = b && p >= (bk_n.frange & 0xFFFFu) && p + XU <= (bk_n.frange >> 16);
        // the reverse-complement span starts 28 bases (56 bits) before the
        // forward one; the four words from the former's hold both
        const int64_t bb = f ? bk_n.bbit + 2 * (int64_t) p - 56 : 0;
        nwb = bb >> 6;
        ninfo = b | (f ? 1u << 8 : 0u) | ((uint32_t) (bb & 63) << 9);
        njl = bk_n.jblk + p;
        nloc = bk_n.local;
    };
    auto fetch_w = [&]() {
        aw = bps64 + nwb;
        w0 = aw[0];
        w1 = aw[1];
        w2 = aw[2];
        w3 = aw[3];
    };
    int par = 0;                                       // which digit counters this round uses
    if (tid < 256) cnt[0][tid] = cnt[1][tid] = hcnt[tid] = 0;
    for (int r = blockIdx.x; r < nrange; r += gridDim.x) {
    const XRange xr = ranges[r];
    __syncthreads();                                   // the previous range is done with the runs
    if (tid < 256) {
        base[tid] = bstart[(size_t) xr.pair * 256 + tid] + rangeoff[(size_t) r * 256 + tid];
        lim[tid] = rangehist[(size_t) r * 256 + tid];
        done[tid] = 0;
    }
    if (xr.lb0 >= xr.lb1) continue;                    // (ranges are never empty)
    fetch_s(xr.lb0 + q, xr.lb1);
    plan_w(xr.lb1);
    fetch_w();
    fetch_s(xr.lb0 + QB + q, xr.lb1);
    for (int64_t lb = xr.lb0; lb < xr.lb1; lb += QB) {
    const bool last = lb + QB >= xr.lb1;
    const uint32_t bits = ninfo & 0xFFu, sb = ninfo >> 9, jl = njl, loc = nloc;
    const bool fast = (ninfo >> 8) & 1u;
    if (!last) plan_w(xr.lb1);                         // the next round's spans
    asm volatile("" ::"v"(as), "v"(ab), "v"(aw));      // (ALU only: the addresses' registers stay live)
    // The thread's records are made twice: here only their first digits, for
    // their ranks; after the layout in full, into the stage (16-byte records
    // held in registers across the layout would take 32 VGPRs).
    auto funnel = [](uint64_t x, uint64_t y, uint32_t s) { return s ? (x >> s) | (y << (64 - s)) : x; };
    // every record of the thread, by slot (forward 2u, reverse complement 2u+1)
    auto records = [&](auto &&sink) {
        if (fast) {
            // the spans from their words: reverse complement at bit sb of the
            // first word, forward 56 bits on
            const uint64_t B0 = funnel(w0, w1, sb), B1 = funnel(w1, w2, sb);
            const bool one = sb < 8;                   // the forward span starts in the first word too
            const uint32_t sa = (sb + 56u) & 63u;
            const uint64_t A0 = funnel(one ? w0 : w1, one ? w1 : w2, sa);
            const uint64_t A1 = funnel(one ? w1 : w2, one ? w2 : w3, sa);
#pragma unroll
            for (int u = 0; u < XU; u++) {
                const uint64_t pay = ((uint64_t) loc << (jbits + 1)) | ((uint64_t) (jl + u) << 1);
                const int sh = 2 * u;
                if ((bits >> (2 * u)) & 1u) {
                    const uint64_t W0 = sh ? (A0 >> sh) | (A1 << (64 - sh)) : A0;
                    const uint32_t W1 = (uint32_t) (A1 >> sh) & 0xFFFFu;
                    sink(2 * u, field_rev64(W0), ((uint64_t) field_rev16(W1) << 48) | pay);
                }
                if ((bits >> (2 * u + 1)) & 1u) {
                    const uint64_t V0 = sh ? (B0 >> sh) | (B1 << (64 - sh)) : B0;
                    const uint32_t V1 = (uint32_t) (B1 >> sh) & 0xFFFFu;
                    sink(2 * u + 1, ~(((uint64_t) V1 << 48) | (V0 >> 16)),
                         ((uint64_t) ((~V0) & 0xFFFFu) << 48) | pay | 1ull);
                }
            }
        } else if (bits) {
            // a block crossing a contig start, or positions near a contig's ends
            const int64_t f0 = (lb + q) * XT * XU + (int64_t) t * XU;
            int c = blk[lb + q].c;
#pragma unroll
            for (int u = 0; u < XU; u++) {
                const int64_t f = f0 + u;
                if (!((bits >> (2 * u)) & 3u)) continue;
                while (c + 1 < nctg && ctg[c + 1].pos0 <= f) c++;
                const ContigDev cc = ctg[c];
                XEmit e;
                e.fwd = (bits >> (2 * u)) & 1u;
                e.rev = (bits >> (2 * u + 1)) & 1u;
                record_keys(bps64, cc, f - cc.pos0, jbits, e);
                if (e.fwd) sink(2 * u, e.f0, e.f1);
                if (e.rev) sink(2 * u + 1, e.r0, e.r1);
            }
            // nothing of this path stays in flight: otherwise the compiler,
            // which tracks outstanding loads per register and not per path,
            // makes the copy-out wait for the prefetch
            __builtin_amdgcn_s_waitcnt(0x0F70);        // vmcnt(0)
        }
    };
    // the next rounds' inputs, once this round's words are used
    auto prefetch = [&]() {
        if (!last) {
            fetch_w();
            fetch_s(lb + 2 * QB + q, xr.lb1);
        }
    };
    uint32_t rk[NS];                                   // each record's rank in its digit this round
    if (fast) {
        // first digits from the forward span's low bits alone
        const bool one = sb < 8;
        const uint64_t A0 = funnel(one ? w0 : w1, one ? w1 : w2, (sb + 56u) & 63u);
#pragma unroll
        for (int u = 0; u < XU; u++) {
            const uint64_t W = A0 >> (2 * u);
            if ((bits >> (2 * u)) & 1u) rk[2 * u] = atomicAdd(&cnt[par][digit_fwd(W)], 1u);
            if ((bits >> (2 * u + 1)) & 1u) rk[2 * u + 1] = atomicAdd(&cnt[par][digit_rev(W)], 1u);
        }
    } else if (bits) {
        records([&](int k, uint64_t a, uint64_t) { rk[k] = atomicAdd(&cnt[par][(uint32_t) (a >> 56)], 1u); });
    }
    __syncthreads();
    // the round's digit runs (one wave: four digits a lane): each digit's
    // held records, then its new ones, in stage order.  A round whose records
    // overflow the stage writes every record straight to its slot instead
    // (spill; low-complexity sequence only).  The other counters are cleared
    // for the next round.
    if (tid < 64) {
        uint32_t cd[4], sum = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            cd[k] = cnt[par][4 * tid + k] + hcnt[4 * tid + k];
            sum += cd[k];
        }
        uint32_t x = sum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if (tid >= o) x += y;
        }
        uint32_t run = x - sum;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            loff[4 * tid + k] = run;
            run += cd[k];
        }
        if (tid == 63) { stot = x; spill = x > (uint32_t) SCAP; }
    } else {
        for (int d = tid - 64; d < 256; d += NT - 64) cnt[par ^ 1][d] = 0;   // emit_clear_covers(NT)
    }
    __syncthreads();
    uint32_t reg = 0;
    if (spill) {
        if (tid < 256) {
            const int64_t g = base[t] + done[t];
            gpos[t] = g;
            done[t] += cnt[par][t];
            for (uint32_t k = 0; k < hn; k++) {
                const int64_t dst = g - hn + k;
                if (dst >= base[t] && dst < base[t] + lim[t]) {
                    out[dst] = held[t][k];
                    if (dig) dig[dst] = (uint8_t) (held[t][k].w0 >> 48);
                } else {
                    *bad = 1;
                }
            }
            hcnt[t] = hn = 0;
        }
        __syncthreads();
        records([&](int k, uint64_t a, uint64_t b) {
            const uint32_t d = (uint32_t) (a >> 56);
            const int64_t dst = gpos[d] + rk[k];
            if (dst >= base[d] && dst < base[d] + lim[d]) {
                out[dst] = Rec{a, b};
                if (dig) dig[dst] = (uint8_t) (a >> 48);
            } else {
                *bad = 1;
            }
        });
    } else {
        records([&](int k, uint64_t a, uint64_t b) {
            const uint32_t d = (uint32_t) (a >> 56);
            stage[loff[d] + hcnt[d] + rk[k]] = Rec{a, b};
        });
        if (tid < 256) {
            // digit t: its held records go in front; the run's first slot and
            // how much of it goes out now (whole segments, or all of it in the
            // range's last round)
            const uint32_t lo = loff[t];
            for (uint32_t k = 0; k < hn; k++) stage[lo + k] = held[t][k];
            const uint32_t cn = cnt[par][t];
            const int64_t g = base[t] + done[t] - hn;
            reg = hn + cn;
            const int64_t e = g + reg;
            const int64_t w = last ? e : (e & ~(int64_t) (HOLD - 1));
            const uint32_t wc = w > g ? (uint32_t) (w - g) : 0u;
            // the run must lie in the digit's slots (the two passes agreeing
            // on its digit counts); what does not is not written
            const int64_t room = base[t] + lim[t] - g;
            if ((int64_t) wc > room) *bad = 1;
            wcnt[t] = wc;
            wout[t] = (int64_t) wc > room ? (uint32_t) max<int64_t>(room, 0) : wc;
            gpos[t] = g;
            done[t] += cn;
        }
    }
    prefetch();                                        // in flight during the copy-out
    __syncthreads();
    const int tot = spill ? 0 : (int) stot;
    // unrolled over the stage's capacity: as a loop, the compiler would wait
    // for the prefetch before entering it
#pragma unroll
    for (int k = 0; k < (SCAP + NT - 1) / NT; k++) {
        const int i = tid + k * NT;
        if (i < tot) {
            const Rec x = stage[i];
            const uint32_t d = (uint32_t) (x.w0 >> 56);
            const uint32_t o = (uint32_t) i - loff[d];
            if (o < wout[d]) {
                out[gpos[d] + o] = x;
                if (dig) dig[gpos[d] + o] = (uint8_t) (x.w0 >> 48);
            }
        }
    }
    if (tot && tid < 256) {
        // the run's incomplete last segment stays
        const uint32_t w = wcnt[t];
        hn = reg - w;
        const uint32_t h0 = loff[t] + w;
        for (uint32_t k = 0; k < hn; k++) held[t][k] = stage[h0 + k];
        hcnt[t] = hn;
    }
    par ^= 1;
    }
    }
}

template __global__ void k_extract_emit<EMIT_QB>(const uint64_t *, const ContigDev *, const BlkDev *, int, int,
                                                 const uint8_t *, const XRange *, int, const uint32_t *,
                                                 const int64_t *, const int64_t *, Rec *, uint8_t *, int *);

// ---------------------------------------------------------------- scans

constexpr int SCAN_T = 256;
constexpr int SCAN_PER = 16;
constexpr int SCAN_SPAN = SCAN_T * SCAN_PER;

template <typename TIn>
__global__ __launch_bounds__(SCAN_T) void k_scan_reduce(const TIn *__restrict__ in, int64_t n,
                                                        int64_t *__restrict__ bsum)
{
    __shared__ uint64_t red[SCAN_T / 64 + 1];
    int64_t base = (int64_t) blockIdx.x * SCAN_SPAN;
    uint64_t s = 0;
    for (int k = 0; k < SCAN_PER; k++) {
        int64_t i = base + (int64_t) k * SCAN_T + threadIdx.x;
        if (i < n) s += (uint64_t) in[i];
    }
    s = block_sum_u64<SCAN_T>(s, red);
    if (threadIdx.x == 0) bsum[blockIdx.x] = (int64_t) s;
}

// In-place exclusive scan of bsum[0..nb), total into bsum[nb].  One block.
__global__ __launch_bounds__(1024) void k_scan_top(int64_t *__restrict__ bsum, int64_t nb)
{
    __shared__ int64_t sh[1024 / 64 + 1];
    int64_t per = (nb + 1023) / 1024;
    int64_t b0 = threadIdx.x * per, b1 = b0 + per < nb ? b0 + per : nb;
    int64_t s = 0;
    for (int64_t i = b0; i < b1; i++) s += bsum[i];
    int64_t tot;
    int64_t run = block_excl_scan<1024, int64_t>(s, sh, &tot);
    for (int64_t i = b0; i < b1; i++) { int64_t t = bsum[i]; bsum[i] = run; run += t; }
    if (threadIdx.x == 0) bsum[nb] = tot;
}

template <typename TIn>
__global__ __launch_bounds__(SCAN_T) void k_scan_down(const TIn *__restrict__ in, int64_t n,
                                                      const int64_t *__restrict__ bsum,
                                                      int64_t *__restrict__ out)
{
    __shared__ int64_t sh[SCAN_T / 64 + 1];
    int64_t base = (int64_t) blockIdx.x * SCAN_SPAN + (int64_t) threadIdx.x * SCAN_PER;
    int64_t v[SCAN_PER];
    int64_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER; k++) {
        int64_t i = base + k;
        v[k] = i < n ? (int64_t) in[i] : 0;
        s += v[k];
    }
    int64_t tot;
    int64_t run = block_excl_scan<SCAN_T, int64_t>(s, sh, &tot) + bsum[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_PER; k++) {
        int64_t i = base + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = bsum[gridDim.x];
}

template __global__ void k_scan_reduce<uint32_t>(const uint32_t *, int64_t, int64_t *);
template __global__ void k_scan_down<uint32_t>(const uint32_t *, int64_t, const int64_t *, int64_t *);

// ---------------------------------------------------------------- small utilities

__global__ void k_fill_u64(unsigned long long *__restrict__ p, int64_t n, unsigned long long v)
{
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t) gridDim.x * blockDim.x)
        p[i] = v;
}

// p[0, *cnt * mult) = v: a fill whose length is a device-side count.
__global__ void k_fill_u64_dev(unsigned long long *__restrict__ p, const int *__restrict__ cnt, int64_t mult,
                               unsigned long long v)
{
    const int64_t n = (int64_t) *cnt * mult;
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x)
        p[i] = v;
}

// Entries of a scanned array at the given indexes: out[p] = ofs[idx[p]], p <= npairs.
__global__ void k_pair_ranges(const int64_t *__restrict__ ofs, const int64_t *__restrict__ idx,
                              int npairs, int64_t *__restrict__ out)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p <= npairs) out[p] = ofs[idx[p]];
}

// Record ranges of the pairs (one block): pr[p+1] - pr[p] = pair p's records,
// the sum of its first-digit totals.
__global__ __launch_bounds__(256) void k_pair_totals(const unsigned long long *__restrict__ hist1, int npairs,
                                                     int64_t *__restrict__ pr)
{
    __shared__ int64_t sh[256 / 64 + 1];
    int64_t run = 0;
    for (int p0 = 0; p0 < npairs; p0 += 256) {
        const int p = p0 + (int) threadIdx.x;
        int64_t n = 0;
        if (p < npairs)
            for (int d = 0; d < 256; d++) n += (int64_t) hist1[(size_t) p * 256 + d];
        int64_t tot;
        const int64_t x = block_excl_scan<256, int64_t>(n, sh, &tot);
        if (p < npairs) pr[p] = run + x;
        run += tot;
    }
    if (threadIdx.x == 0) pr[npairs] = run;
}

// ---------------------------------------------------------------- radix partition

// Radix digits: 8 bits at every level except the second, which is w2 bits
// (8..10) wide, so that one more level's worth of buckets fits the tiles.
// Key bit 0 is the top bit of w0; bits past 127 read as zero.
constexpr int ND_MAX = 1 << 11;      // largest digit count of a level

__device__ __forceinline__ int digit_pos(int level, int w2)
{
    return level == 1 ? 0 : level == 2 ? 8 : w2 + 8 * (level - 2);
}

__device__ __forceinline__ int digit_bits(int level, int w2) { return level == 2 ? w2 : 8; }

// the 64 key bits from key bit p on
__device__ __forceinline__ uint64_t key_bits_from(uint64_t w0, uint64_t w1, int p)
{
    if (p == 0) return w0;
    if (p < 64) return (w0 << p) | (w1 >> (64 - p));
    return p < 128 ? w1 << (p - 64) : 0ull;
}

__device__ __forceinline__ uint32_t digit_at(uint64_t w0, uint64_t w1, int level, int w2)
{
    return (uint32_t) (key_bits_from(w0, w1, digit_pos(level, w2)) >> (64 - digit_bits(level, w2)));
}

__device__ __forceinline__ int seg_of_chunk(const int64_t *__restrict__ chk0, int nseg, int64_t b)
{
    int lo = 0, hi = nseg;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (chk0[mid] <= b) lo = mid; else hi = mid;
    }
    return lo;
}

// Exclusive prefix of per-segment chunk counts (one block); chk0[nseg] = total
// chunks, chk0[nseg+1] = total records of the segments.
// dn (may be null): the segment count in device memory, read instead of nseg
// (the device-driven pipeline; the same holds for the kernels below).
__global__ __launch_bounds__(1024) void k_chunk_prefix(const Seg *__restrict__ segs, int nseg, int chunk,
                                                       int64_t *__restrict__ chk0, const int *__restrict__ dn)
{
    __shared__ int64_t sh[1024 / 64 + 1];
    if (dn) nseg = *dn;
    int per = (nseg + 1023) / 1024;
    int b0 = threadIdx.x * per, b1 = b0 + per < nseg ? b0 + per : nseg;
    int64_t s = 0, r = 0;
    for (int i = b0; i < b1; i++) { s += (segs[i].count + chunk - 1) / chunk; r += segs[i].count; }
    int64_t tot, rtot;
    int64_t run = block_excl_scan<1024, int64_t>(s, sh, &tot);
    (void) block_excl_scan<1024, int64_t>(r, sh, &rtot);
    for (int i = b0; i < b1; i++) { chk0[i] = run; run += (segs[i].count + chunk - 1) / chunk; }
    if (threadIdx.x == 0) { chk0[nseg] = tot; chk0[nseg + 1] = rtot; }
}

constexpr int PT = 256;   // threads of the partition kernels
constexpr unsigned long long CURSOR_DIG = 1ull << 62;   // digit cursor flag: its bucket is a next-level segment

// Digit counts of the segments at one level (all segments of a pass share
// it): one row of nd = 2^digit_bits counters per segment.  The chunk count
// comes from chk0[nseg], so the grid is sized without a round trip.  A
// segment whose records the previous level's scatter moved (skip bit 1) has
// its digits in the byte array dig, which that scatter wrote beside the
// records: its counts read one byte a record instead of the 16-byte record.
template <int PER>
__global__ __launch_bounds__(PT) void k_seg_hist(const Rec *__restrict__ buf0, const Rec *__restrict__ buf1,
                                                 const Seg *__restrict__ segs,
                                                 int nseg, const int64_t *__restrict__ chk0, int level, int w2,
                                                 unsigned long long *__restrict__ seghist,
                                                 const uint8_t *__restrict__ dig, const int *__restrict__ dn)
{
    __shared__ uint32_t hist[ND_MAX];
    if (dn) nseg = *dn;
    const int CH = PT * PER;
    const int64_t nlog = chk0[nseg];
    const int nd = 1 << digit_bits(level, w2);
    // a contiguous run of chunks a block, counted into one LDS histogram that
    // goes out (global atomics) only when the segment changes: a chunk of
    // digit bytes is 8 KB, too little to pay for 2^bits global atomics
    const int64_t per = (nlog + gridDim.x - 1) / gridDim.x;
    const int64_t c0 = (int64_t) blockIdx.x * per, c1 = c0 + per < nlog ? c0 + per : nlog;
    for (int d = threadIdx.x; d < nd; d += PT) hist[d] = 0;
    __syncthreads();
    auto flush = [&](int s) {
        __syncthreads();
        for (int d = threadIdx.x; d < nd; d += PT) {
            const uint32_t h = hist[d];
            hist[d] = 0;
            if (h) atomicAdd(&seghist[(size_t) s * nd + d], (unsigned long long) h);
        }
        __syncthreads();
    };
    int cur = -1;
    for (int64_t LB = c0; LB < c1; LB++) {
        const int s = cur >= 0 && LB < chk0[cur + 1] ? cur : seg_of_chunk(chk0, nseg, LB);
        if (s != cur) {
            if (cur >= 0) flush(cur);
            cur = s;
        }
        const Seg sg = segs[s];
        const int64_t off = (LB - chk0[s]) * (int64_t) CH;
        const int n = (int) min((int64_t) CH, sg.count - off);
        if (dig && (sg.skip & 2) && PER % 16 == 0) {
            // a thread's PER consecutive bytes, sixteen at a time (unaligned
            // 16-byte loads; a run cut by the chunk's end byte by byte)
            constexpr int V = PER >= 16 ? PER / 16 : 1;
            const uint8_t *q = dig + sg.start + off + (int64_t) threadIdx.x * PER;
            const int m = n - (int) threadIdx.x * PER;           // bytes of this thread's run
            if (m >= PER) {
                uint4 v[V];
#pragma unroll
                for (int k = 0; k < V; k++) __builtin_memcpy(&v[k], q + 16 * k, 16);
#pragma unroll
                for (int k = 0; k < V; k++) {
                    const uint32_t w[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
#pragma unroll
                    for (int b = 0; b < 16; b++) atomicAdd(&hist[(w[b >> 2] >> (8 * (b & 3))) & 0xFFu], 1u);
                }
            } else {
                for (int k = 0; k < m; k++) atomicAdd(&hist[q[k]], 1u);
            }
        } else if (dig && (sg.skip & 2)) {
            const uint8_t *q = dig + sg.start + off;
#pragma unroll 4
            for (int k = 0; k < PER; k++) {
                const int i = k * PT + threadIdx.x;
                if (i < n) atomicAdd(&hist[q[i]], 1u);
            }
        } else {
            const Rec *p = (sg.buf ? buf1 : buf0) + sg.start + off;
#pragma unroll 4
            for (int k = 0; k < PER; k++) {
                const int i = k * PT + threadIdx.x;
                if (i < n) {
                    const Rec r = p[i];
                    atomicAdd(&hist[digit_at(r.w0, r.w1, level, w2)], 1u);
                }
            }
        }
    }
    if (cur >= 0) flush(cur);
}

// Per segment: digit offsets (cursors) and the children.  Children of at most
// `cap` records are grouped into sort tiles (runs of consecutive non-empty
// buckets); larger ones become segments of the next level.  Children go to the
// other record buffer, except when the segment has one non-empty bucket: then
// its records are already in bucket order, the scatter skips it, and its child
// (segment or tile) stays where it is.  With in_place the whole level is
// already in bucket order (the extraction wrote it so) and is not scattered.
// dig_next: this level's scatter writes the next level's digits into the
// digit byte array for the children that become next-level segments (their
// cursors carry CURSOR_DIG), and those children are marked (skip bit 1) so
// that the next level counts them from there.
__global__ __launch_bounds__(256) void k_seg_plan(Seg *__restrict__ segs,
                                                  const unsigned long long *__restrict__ seghist,
                                                  int hist_by_pair, int level, int w2, int in_place, int dig_next,
                                                  int cap,
                                                  unsigned long long *__restrict__ cursor,
                                                  Seg *__restrict__ next, int *__restrict__ nnext, int next_cap,
                                                  Tile *__restrict__ tiles, int *__restrict__ ntiles, int tile_cap,
                                                  unsigned long long *__restrict__ kept, int nseg,
                                                  const int *__restrict__ dn)
{
    __shared__ int64_t sh[256 / 64 + 1];
    __shared__ int64_t cnt[ND_MAX], off[ND_MAX];
    __shared__ uint16_t tfirst[ND_MAX], tlast[ND_MAX], lnext[ND_MAX];
    __shared__ int nt, nn, tbase, nbase, one_digit;
    const int nd = 1 << digit_bits(level, w2);
    const int pd = nd / 256;                        // digits per thread, contiguous
    if (dn) nseg = *dn;
    for (int sb = blockIdx.x; sb < nseg; sb += gridDim.x) {
    __syncthreads();                                // the previous segment is done with the LDS lists
    const Seg sg = segs[sb];
    const size_t row = (hist_by_pair ? (size_t) sg.pair : (size_t) sb) * (size_t) nd;
    int64_t sum = 0;
    int nz = 0;
    for (int q = 0; q < pd; q++) {
        const int d = threadIdx.x * pd + q;
        const int64_t c = (int64_t) seghist[row + d];
        cnt[d] = c;
        sum += c;
        nz += c != 0;
    }
    if (nz == 1)                                    // the digit, should it be the segment's only one
        for (int q = 0; q < pd; q++)
            if (cnt[threadIdx.x * pd + q]) one_digit = threadIdx.x * pd + q;
    const int holders = __syncthreads_count(nz != 0);
    if (holders == 1 && __syncthreads_count(nz > 1) == 0) {
        // a single bucket: no scatter; the child keeps the buffer and the range
        if (threadIdx.x == 0) {
            segs[sb].skip = 1;
            if (sg.count > cap) {
                const int q = atomicAdd(nnext, 1);
                if (q < next_cap) next[q] = Seg{sg.start, sg.count, sg.pair, level + 1, sg.buf, 0};
            } else {
                const int t = atomicAdd(ntiles, 1);
                if (t < tile_cap) tiles[t] = Tile{sg.start, (int32_t) sg.count, sg.pair, sg.buf, level,
                                                  one_digit | one_digit << 16};
            }
            atomicAdd(kept, (unsigned long long) sg.count);
        }
        continue;
    }
    const int dstbuf = in_place ? sg.buf : sg.buf ^ 1;   // in_place: already in bucket order, no scatter
    int64_t tot;
    int64_t run = block_excl_scan<256, int64_t>(sum, sh, &tot);
    for (int q = 0; q < pd; q++) {
        const int d = threadIdx.x * pd + q;
        const int64_t c = cnt[d];
        off[d] = run;
        cursor[(size_t) sb * nd + d] = (unsigned long long) (sg.start + run) |
                                       (dig_next && c > cap ? CURSOR_DIG : 0ull);
        run += cnt[d];
    }
    __syncthreads();
    // One lane groups the digits (a serial greedy over the counts) into LDS
    // lists of digit indices: a tile is a run of digits [tfirst, tlast], a
    // next-level segment one digit.  The segment then reserves its slots with
    // one atomic per list.
    if (threadIdx.x == 0) {
        int t = 0, q = 0, tf = 0, tl = 0;
        int64_t tc = 0;
        for (int d = 0; d < nd; d++) {
            const int64_t k = cnt[d];
            if (k == 0) continue;
            if (k > cap) {
                if (tc) {
                    tfirst[t] = (uint16_t) tf;
                    tlast[t++] = (uint16_t) tl;
                    tc = 0;
                }
                lnext[q++] = (uint16_t) d;
            } else {
                if (tc + k > cap) {
                    tfirst[t] = (uint16_t) tf;
                    tlast[t++] = (uint16_t) tl;
                    tc = 0;
                }
                if (tc == 0) tf = d;
                tl = d;
                tc += k;
            }
        }
        if (tc) {
            tfirst[t] = (uint16_t) tf;
            tlast[t++] = (uint16_t) tl;
        }
        nt = t;
        nn = q;
        tbase = t ? atomicAdd(ntiles, t) : 0;
        nbase = q ? atomicAdd(nnext, q) : 0;
    }
    __syncthreads();
    // slots past the capacity are counted but not written: the caller sees the
    // count above the capacity and reports an internal error (the lists are
    // sized from a bound that a correct count never exceeds)
    for (int i = threadIdx.x; i < nt && tbase + i < tile_cap; i += 256) {
        const int f = tfirst[i], l = tlast[i];
        tiles[tbase + i] = Tile{sg.start + off[f], (int32_t) (off[l] + cnt[l] - off[f]), sg.pair, dstbuf, level,
                                f | l << 16};
    }
    for (int i = threadIdx.x; i < nn && nbase + i < next_cap; i += 256) {
        const int d = lnext[i];
        next[nbase + i] = Seg{sg.start + off[d], cnt[d], sg.pair, level + 1, dstbuf, dig_next ? 2 : 0};
    }
    }
}

// Scatter one level: each chunk of CH records is ranked by its digit in LDS,
// staged in digit order, and written as one run per digit at the digit's
// cursor (reserved with one atomic per digit and chunk).  The next chunk's
// records are loaded into registers before the current chunk is written out,
// so its reads overlap the writes.  With dig, a moved record whose bucket is a
// next-level segment (its cursor carries CURSOR_DIG) also writes its digit at
// the next level (an 8-bit one) to dig at its new position.
template <int T, int PER>
__global__ __launch_bounds__(T) void k_seg_scatter(Rec *__restrict__ buf0, Rec *__restrict__ buf1,
                                                   const Seg *__restrict__ segs, int nseg,
                                                   const int64_t *__restrict__ chk0, int level, int w2,
                                                   unsigned long long *__restrict__ cursor,
                                                   uint8_t *__restrict__ dig, const int *__restrict__ dn)
{
    static_assert(T >= 256, "one thread per digit at least");
    if (dn) nseg = *dn;
    constexpr int CH = T * PER;
    const int64_t nlog = chk0[nseg];
    __shared__ Rec stage[CH];
    __shared__ uint32_t hist[ND_MAX];               // digit counts, then the digits' runs in the stage
    __shared__ int64_t gb[ND_MAX];                  // a digit's destination minus its stage offset
    __shared__ uint8_t gdig[ND_MAX];                // the digit's bucket is a next-level segment: write dig
    __shared__ uint32_t sh[T / 64 + 1];
    const int tid = threadIdx.x;
    const int nd = 1 << digit_bits(level, w2);
    const int pd = nd > T ? nd / T : 1;             // digits per thread in the scan, contiguous
    Rec r[PER];
    int s = 0, n = 0, from = 0;
    // records of chunk lb into r[] (segment s, n records, in buffer `from`);
    // chunks of skipped segments have no records to move
    auto fetch = [&](int64_t lb) {
        s = seg_of_chunk(chk0, nseg, lb);
        const Seg sg = segs[s];
        const int64_t off = (lb - chk0[s]) * (int64_t) CH;
        n = (sg.skip & 1) ? 0 : (int) min((int64_t) CH, sg.count - off);
        from = sg.buf;
        const Rec *p = (from ? buf1 : buf0) + sg.start + off;
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int i = k * T + tid;
            if (i < n) r[k] = p[i];
        }
    };
    if ((int64_t) blockIdx.x < nlog) fetch(blockIdx.x);
    for (int64_t LB = blockIdx.x; LB < nlog; LB += gridDim.x) {
    for (int d = tid; d < nd; d += T) hist[d] = 0;
    __syncthreads();
    uint32_t rk[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
        int i = k * T + tid;
        if (i < n) rk[k] = atomicAdd(&hist[digit_at(r[k].w0, r[k].w1, level, w2)], 1u);
    }
    __syncthreads();
    // a thread's digits: their slots are reserved before the block scan, so the
    // atomics' round trips overlap the scan and end in one wait
    constexpr int PDM = ND_MAX / T > 1 ? ND_MAX / T : 1;
    const int d0 = tid * pd;
    uint32_t hq[PDM], hs = 0;
    unsigned long long gq[PDM];
#pragma unroll
    for (int q = 0; q < PDM; q++) {
        hq[q] = q < pd && d0 + q < nd ? hist[d0 + q] : 0u;
        hs += hq[q];
        gq[q] = 0;
    }
#pragma unroll
    for (int q = 0; q < PDM; q++)
        if (hq[q]) gq[q] = atomicAdd(&cursor[(size_t) s * nd + d0 + q], (unsigned long long) hq[q]);
    uint32_t tot;
    uint32_t lo = block_excl_scan<T, uint32_t>(hs, sh, &tot);
#pragma unroll
    for (int q = 0; q < PDM; q++) {
        if (q < pd && d0 + q < nd) {
            hist[d0 + q] = lo;
            if (hq[q]) {
                gb[d0 + q] = (int64_t) (gq[q] & ~CURSOR_DIG) - lo;
                gdig[d0 + q] = (gq[q] & CURSOR_DIG) ? 1 : 0;
            }
            lo += hq[q];
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; k++) {
        int i = k * T + tid;
        if (i < n) stage[hist[digit_at(r[k].w0, r[k].w1, level, w2)] + rk[k]] = r[k];
    }
    __syncthreads();
    const int ncur = n;
    Rec *const dst = from ? buf0 : buf1;
    if (LB + gridDim.x < nlog) fetch(LB + gridDim.x);
#pragma unroll
    for (int k = 0; k < PER; k++) {
        int i = k * T + tid;
        if (i < ncur) {
            Rec x = stage[i];
            const uint32_t d = digit_at(x.w0, x.w1, level, w2);
            const int64_t to = gb[d] + i;
            dst[to] = x;
            if (dig && gdig[d]) dig[to] = (uint8_t) digit_at(x.w0, x.w1, level + 1, w2);
        }
    }
    __syncthreads();
    }
}

template __global__ void k_seg_hist<16>(const Rec *, const Rec *, const Seg *, int, const int64_t *, int, int, unsigned long long *, const uint8_t *, const int *);
template __global__ void k_seg_hist<8>(const Rec *, const Rec *, const Seg *, int, const int64_t *, int, int, unsigned long long *, const uint8_t *, const int *);
template __global__ void k_seg_hist<32>(const Rec *, const Rec *, const Seg *, int, const int64_t *, int, int, unsigned long long *, const uint8_t *, const int *);
template __global__ void k_seg_scatter<256, 8>(Rec *, Rec *, const Seg *, int, const int64_t *, int, int, unsigned long long *, uint8_t *, const int *);
template __global__ void k_seg_scatter<256, 16>(Rec *, Rec *, const Seg *, int, const int64_t *, int, int, unsigned long long *, uint8_t *, const int *);
template __global__ void k_seg_scatter<1024, 8>(Rec *, Rec *, const Seg *, int, const int64_t *, int, int, unsigned long long *, uint8_t *, const int *);

// ---------------------------------------------------------------- tile sort

__device__ __forceinline__ bool key_le(uint64_t a0, uint64_t a1, uint64_t b0, uint64_t b1)
{
    return a0 < b0 || (a0 == b0 && a1 <= b1);
}

// 1 if (y, tie) comes before x: y < x as 128-bit keys, or y == x and tie --
// the borrow out of y - x - tie, with no branches.
__device__ __forceinline__ uint32_t comes_before(const Rec &y, const Rec &x, bool tie)
{
    unsigned long long b0, b1;
    (void) __builtin_subcll(y.w1, x.w1, tie ? 1ull : 0ull, &b0);
    (void) __builtin_subcll(y.w0, x.w0, b0, &b1);
    return (uint32_t) b1;
}

// rank + 1 if (y, i < me) comes before x (the borrow of y - x - (i < me) over
// the four 32-bit words), else rank: five VALU instructions chained through
// VCC.  Written out because the compiler turns the borrow chain of
// comes_before into 64-bit compares combined in scalar registers.
__device__ __forceinline__ uint32_t rank_step(uint32_t rank, const Rec &y, const Rec &x, uint32_t i, uint32_t me)
{
    uint32_t t0, t1, t2, t3;
    asm("v_cmp_lt_u32 vcc, %[i], %[me]\n\t"
        "v_subb_co_u32 %[t0], vcc, %[y1l], %[x1l], vcc\n\t"
        "v_subb_co_u32 %[t1], vcc, %[y1h], %[x1h], vcc\n\t"
        "v_subb_co_u32 %[t2], vcc, %[y0l], %[x0l], vcc\n\t"
        "v_subb_co_u32 %[t3], vcc, %[y0h], %[x0h], vcc\n\t"
        "v_addc_co_u32 %[r], vcc, 0, %[r], vcc"
        : [r] "+v"(rank), [t0] "=&v"(t0), [t1] "=&v"(t1), [t2] "=&v"(t2), [t3] "=&v"(t3)
        : [i] "v"(i), [me] "v"(me),
          [y1l] "v"((uint32_t) y.w1), [y1h] "v"((uint32_t) (y.w1 >> 32)),
          [y0l] "v"((uint32_t) y.w0), [y0h] "v"((uint32_t) (y.w0 >> 32)),
          [x1l] "v"((uint32_t) x.w1), [x1h] "v"((uint32_t) (x.w1 >> 32)),
          [x0l] "v"((uint32_t) x.w0), [x0h] "v"((uint32_t) (x.w0 >> 32))
        : "vcc");
    return rank;
}

// Leading equal bases of two keys (40 if the k-mers are equal), with no
// branches: the first differing bit in w0, else in the 16 k-mer bits of w1
// (bit 15 set below them stands for "none", giving 32 + 8).
__device__ __forceinline__ int lcp_bases(uint64_t a0, uint64_t a1, uint64_t b0, uint64_t b1)
{
    const uint64_t x = a0 ^ b0;
    const uint32_t y = (uint32_t) ((a1 ^ b1) >> 48);
    const int lx = __builtin_clzll(x | 1ull) >> 1;
    const int ly = 32 + (__builtin_clz((y << 16) | 0x8000u) >> 1);
    return x ? lx : ly;
}

constexpr int FB_BIN = 256;          // largest bin the binned tile sort ranks directly
constexpr int MAXC = 32;             // crowded ranges a tile sorts by sub-bins (more: merge sort)
constexpr int CROWD_ROUNDS = 6;      // rounds of sub-binning before the merge sort takes over

#define PGA_CE(i, j)                                                             \
    do {                                                                          \
        bool sw_ = (a0[j] < a0[i]) || (a0[j] == a0[i] && a1[j] < a1[i]);          \
        uint64_t x0_ = sw_ ? a0[j] : a0[i], x1_ = sw_ ? a1[j] : a1[i];            \
        uint64_t y0_ = sw_ ? a0[i] : a0[j], y1_ = sw_ ? a1[i] : a1[j];            \
        a0[i] = x0_; a1[i] = x1_; a0[j] = y0_; a1[j] = y1_;                       \
    } while (0)

// One workgroup sorts one tile (<= T*8 records) with an in-register 8-sort and
// merge-path merges through LDS, then writes, per record: the GIX entry
// (src/kmer_adapter.c:55-62 layout), the LCP/strand byte and the 24-bit prefix.
// The LCP of a tile's first record is written by k_tile_fix.
template <int T>
__global__ __launch_bounds__(T, 4) void k_tile_sort(Rec *const *__restrict__ bufs,
                                                 const Tile *__restrict__ tiles,
                                                 const PairDev *__restrict__ pairs, int jbits, int w2,
                                                 uint8_t *__restrict__ entries,
                                                 uint8_t *__restrict__ ls, uint32_t *__restrict__ pre, int pre_all,
                                                 unsigned long long *__restrict__ cum,
                                                 Rec *__restrict__ skeys, int64_t nlog, int *__restrict__ nfall,
                                                 const int *__restrict__ dn)
{
    constexpr int E = 8;
    if (dn) nlog = *dn;
    constexpr int CAP = T * E;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Rec *sk = reinterpret_cast<Rec *>(smem_raw);
    const int tid = threadIdx.x;
    // a tile's records (striped, coalesced) into registers; the next tile's
    // are loaded while the current one writes its output
    Rec r[E];
    auto load_tile = [&](int64_t lb) {
        const Tile t = tiles[lb];
        const Rec *src = bufs[t.buf] + t.start;
#pragma unroll
        for (int k = 0; k < E; k++)
            if (k * T + tid < t.count) r[k] = load_rec_global(src + k * T + tid);
    };
    if ((int64_t) blockIdx.x < nlog) load_tile(blockIdx.x);
    for (int64_t LB = blockIdx.x; LB < nlog; LB += gridDim.x) {
    const Tile tl = tiles[LB];
    const PairDev pd = pairs[tl.pair];
    const int n = tl.count;

    // bins of the binned sort: one per record, 16-bit counters (later bin
    // starts), two to a word; counts and starts are <= CAP < 2^16
    constexpr int NB = CAP;
    uint32_t *bh = reinterpret_cast<uint32_t *>(smem_raw + (size_t) CAP * sizeof(Rec));
    auto bin_start = [&](uint32_t b) -> uint32_t { return (bh[b >> 1] >> ((b & 1u) << 4)) & 0xFFFFu; };
    __shared__ uint32_t scan_sh[T / 64 + 1];
    __shared__ int fallback;
    // crowded ranges (bins, then sub-bins, of more than FB_BIN records): start
    // in the tile and size, two lists (this round's and the next's)
    __shared__ int crst[2][MAXC], crm[2][MAXC], ncrow[2];
    __shared__ int mo[MAXC + 1], cq[MAXC], csb[MAXC + 1], cp[MAXC];
    __shared__ unsigned long long cx[MAXC][2];

    // smallest and largest digit at the tile's level (its records share the
    // digits above it), as the plan listed them
    const int L = tl.level;
    const int dp = digit_pos(L, w2), dbits = digit_bits(L, w2);
    const uint32_t dmin = (uint32_t) tl.drange & 0xFFFFu, dmax = (uint32_t) tl.drange >> 16;
    for (int c = tid; c < NB / 2; c += T) bh[c] = 0;
    if (tid == 0) { fallback = 0; ncrow[0] = 0; }
    __syncthreads();
    // Bin of a record: the top BB bits of (its digit at the tile's level - the
    // smallest such digit, then the key bits right below that digit).  A tile
    // is a run of buckets of one parent, so buckets map to disjoint bin ranges
    // (or, when the span is wider than BB, to ranges in key order) and the bits
    // inside a bucket are near-uniform: bins hold ~1 record unless the k-mers
    // repeat.
    constexpr int BB = __builtin_ctz(NB);
    const int span = dmax > dmin ? 32 - __builtin_clz(dmax - dmin) : 0;     // <= the digit's bits
    // The 64 key bits from the digit on, minus dmin in the digit's field, are
    // (digit - dmin) then the bits below it, with the top (bits - span) bits
    // zero: the bin is their next BB bits.
    const uint64_t dsub = (uint64_t) dmin << (64 - dbits);
    const int bsh = 64 - BB - dbits + span;                                 // 41..53
    auto bin_of = [&](uint64_t w0, uint64_t w1) -> uint32_t {
        return (uint32_t) ((key_bits_from(w0, w1, dp) - dsub) >> bsh);
    };
    uint32_t rk[E];
    // all the atomics issued before the first result is taken apart: which
    // half of its counter word a record's bin is goes to bit k of odd
    uint32_t odd = 0;
#pragma unroll
    for (int k = 0; k < E; k++)
        if (k * T + tid < n) {
            const uint32_t b = bin_of(r[k].w0, r[k].w1);
            odd |= (b & 1u) << k;
            rk[k] = atomicAdd(&bh[b >> 1], 1u << ((b & 1u) << 4));
        }
#pragma unroll
    for (int k = 0; k < E; k++) rk[k] = (rk[k] >> (((odd >> k) & 1u) << 4)) & 0xFFFFu;
    __syncthreads();
    // Bin counts -> bin starts (16-bit, two to a word).  A bin of more than
    // FB_BIN records (repeated k-mers) is listed as a crowded range through
    // `crowded(start, count)` instead of being ranked record by record.
    constexpr int WPT = NB / 2 / T;                 // counter words per thread
    auto scan_bins = [&](auto crowded) {
        uint32_t cw[WPT], sum = 0;
#pragma unroll
        for (int q = 0; q < WPT; q++) {
            cw[q] = bh[tid * WPT + q];
            sum += (cw[q] & 0xFFFFu) + (cw[q] >> 16);
        }
        uint32_t tot;
        uint32_t run = block_excl_scan<T, uint32_t>(sum, scan_sh, &tot);
#pragma unroll
        for (int q = 0; q < WPT; q++) {
            const uint32_t lo = cw[q] & 0xFFFFu, hi = cw[q] >> 16;
            if (lo > FB_BIN) crowded(run, lo);
            if (hi > FB_BIN) crowded(run + lo, hi);
            bh[tid * WPT + q] = run | ((run + lo) << 16);
            run += lo + hi;
        }
    };
    auto add_crowd = [&](int list, int st, int m) {
        const int x = atomicAdd(&ncrow[list], 1);
        if (x < MAXC) { crst[list][x] = st; crm[list][x] = m; }
        else fallback = 1;                          // too many: merge sort the tile
    };
    scan_bins([&](uint32_t st, uint32_t m) { add_crowd(0, (int) st, (int) m); });
    __syncthreads();
    const bool binned = !fallback;
    if (binned) {
        // scatter into bins, then every record counts the keys of its bin that
        // come before it (smaller, or equal and scattered earlier: the seed sort
        // has equal keys) and moves to bin start + that rank.  The ranking goes
        // in bin order (lane p takes the record scattered to p), so a wave's
        // lanes share bins and only the waves over a crowded bin loop long.
#pragma unroll
        for (int k = 0; k < E; k++)
            if (k * T + tid < n) sk[bin_start(bin_of(r[k].w0, r[k].w1)) + rk[k]] = r[k];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < E; k++) {
            const uint32_t me = (uint32_t) (k * T + tid);
            if (me >= (uint32_t) n) continue;
            r[k] = sk[me];
            const uint32_t b = bin_of(r[k].w0, r[k].w1);
            const uint32_t st = bin_start(b);
            const uint32_t en = b + 1 < (uint32_t) NB ? bin_start(b + 1) : (uint32_t) n;
            if (en - st > (uint32_t) FB_BIN) { rk[k] = me; continue; }   // crowded: sorted below
            // four reads in flight a step over a long bin, then two
            uint32_t rank = 0, i = st;
            for (; i + 4 <= en; i += 4) {
                const Rec y0 = sk[i], y1 = sk[i + 1], y2 = sk[i + 2], y3 = sk[i + 3];
                rank = rank_step(rank_step(rank, y0, r[k], i, me), y1, r[k], i + 1, me);
                rank = rank_step(rank_step(rank, y2, r[k], i + 2, me), y3, r[k], i + 3, me);
            }
            for (; i + 2 <= en; i += 2) {
                const Rec y0 = sk[i], y1 = sk[i + 1];
                rank = rank_step(rank_step(rank, y0, r[k], i, me), y1, r[k], i + 1, me);
            }
            if (i < en) rank = rank_step(rank, sk[i], r[k], i, me);
            rk[k] = st + rank;                          // the record's place in the tile
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < E; k++)
            if (k * T + tid < n) sk[rk[k]] = r[k];
        __syncthreads();
        // Crowded ranges, sorted in place round by round: each range's records
        // share the key bits above the first bit where any two of them differ
        // (the OR of their XORs with the range's first record), so their
        // sub-bin is the next q bits from there: q = floor(log2 m) (<= m
        // sub-bins a range, <= NB in all), or all of the bin bits for a round's
        // only range.  Sub-bins of at most FB_BIN records are
        // ranked as above; larger ones are the next round's ranges.  Ranges of
        // identical keys (the seed sort's duplicates) are sorted already.
        int cur = 0;
        for (int round = 0; round < CROWD_ROUNDS; round++) {
            const int nc = min(ncrow[cur], MAXC);
            if (nc == 0 || fallback) break;
            if (tid == 0) {
                int acc = 0;
                for (int j = 0; j < nc; j++) { mo[j] = acc; acc += crm[cur][j]; }
                mo[nc] = acc;
                ncrow[cur ^ 1] = 0;
            }
            if (tid < nc) { cx[tid][0] = 0; cx[tid][1] = 0; }
            if (tid == 0 && round == 0 && nfall) atomicAdd(nfall + 1, 1);
            __syncthreads();
            const int M = mo[nc];
            auto range_of = [&](int i) {                 // the range holding item i of the concatenation
                int lo = 0, hi = nc;
                while (hi - lo > 1) { const int mid = (lo + hi) >> 1; if (mo[mid] <= i) lo = mid; else hi = mid; }
                return lo;
            };
#pragma unroll
            for (int k = 0; k < E; k++) {
                const int i = k * T + tid;
                if (i >= M) continue;
                const int j = range_of(i);
                r[k] = sk[crst[cur][j] + i - mo[j]];
                const Rec ref = sk[crst[cur][j]];
                const unsigned long long d0 = r[k].w0 ^ ref.w0, d1 = r[k].w1 ^ ref.w1;
                if (d0) atomicOr(&cx[j][0], d0);
                if (d1) atomicOr(&cx[j][1], d1);
            }
            for (int c = tid; c < NB / 2; c += T) bh[c] = 0;
            __syncthreads();
            if (tid == 0) {
                int acc = 0;
                for (int j = 0; j < nc; j++) {
                    const unsigned long long x0 = cx[j][0], x1 = cx[j][1];
                    const int p = x0 ? __builtin_clzll(x0) : x1 ? 64 + __builtin_clzll(x1) : 128;
                    const int lg = 31 - __builtin_clz((uint32_t) crm[cur][j]);
                    // a lone range (the usual case) takes all NB sub-bins
                    const int q = p >= 128 ? 0 : min(min(nc == 1 ? BB : lg, 12), 128 - p);
                    cp[j] = p;
                    cq[j] = q;
                    csb[j] = acc;
                    acc += 1 << q;
                }
                csb[nc] = acc;
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < E; k++) {
                const int i = k * T + tid;
                if (i >= M) continue;
                const int j = range_of(i);
                const int q = cq[j];
                const uint32_t sb = (uint32_t) csb[j] +
                                    (q ? (uint32_t) (key_bits_from(r[k].w0, r[k].w1, cp[j]) >> (64 - q)) : 0u);
                const uint32_t sh = (sb & 1u) << 4;
                rk[k] = (atomicAdd(&bh[sb >> 1], 1u << sh) >> sh) & 0xFFFFu;
            }
            __syncthreads();
            // sub-bin starts in the concatenation; crowded sub-bins of ranges
            // that can still split become the next round's ranges (in tile
            // positions)
            scan_bins([&](uint32_t st, uint32_t m) {
                const int j = range_of((int) st);
                if (cq[j] > 0) add_crowd(cur ^ 1, crst[cur][j] + (int) st - mo[j], (int) m);
            });
            __syncthreads();
            if (fallback) break;
            // lay each range out by sub-bin
#pragma unroll
            for (int k = 0; k < E; k++) {
                const int i = k * T + tid;
                if (i >= M) continue;
                const int j = range_of(i);
                if (!cq[j]) continue;
                const uint32_t sb = (uint32_t) csb[j] + (uint32_t) (key_bits_from(r[k].w0, r[k].w1, cp[j]) >> (64 - cq[j]));
                rk[k] += bin_start(sb) - (uint32_t) mo[j] + (uint32_t) crst[cur][j];   // tile position
                sk[rk[k]] = r[k];
            }
            __syncthreads();
            // rank inside sub-bins of at most FB_BIN records
#pragma unroll
            for (int k = 0; k < E; k++) {
                const int i = k * T + tid;
                if (i >= M) continue;
                const int j = range_of(i);
                if (!cq[j]) { rk[k] = ~0u; continue; }
                const uint32_t sb = (uint32_t) csb[j] + (uint32_t) (key_bits_from(r[k].w0, r[k].w1, cp[j]) >> (64 - cq[j]));
                const uint32_t base = (uint32_t) crst[cur][j] - (uint32_t) mo[j];
                const uint32_t st = bin_start(sb) + base;
                const uint32_t en = (sb + 1 < (uint32_t) csb[nc] ? bin_start(sb + 1) : (uint32_t) M) + base;
                if (en - st > (uint32_t) FB_BIN) { rk[k] = ~0u; continue; }   // the next round's
                // four reads in flight a step: a round's few lanes each walk a
                // sub-bin of up to FB_BIN records, so the walk is LDS latency
                const uint32_t me = rk[k];
                uint32_t rank = 0, x = st;
                for (; x + 4 <= en; x += 4) {
                    const Rec y0 = sk[x], y1 = sk[x + 1], y2 = sk[x + 2], y3 = sk[x + 3];
                    rank = rank_step(rank_step(rank, y0, r[k], x, me), y1, r[k], x + 1, me);
                    rank = rank_step(rank_step(rank, y2, r[k], x + 2, me), y3, r[k], x + 3, me);
                }
                for (; x < en; x++) rank = rank_step(rank, sk[x], r[k], x, me);
                rk[k] = st + rank;
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < E; k++)
                if (k * T + tid < M && rk[k] != ~0u) sk[rk[k]] = r[k];
            __syncthreads();
            cur ^= 1;
        }
        if (tid == 0 && ncrow[cur] > 0) fallback = 1;   // clusters left after the last round
        __syncthreads();
    } else {
#pragma unroll
        for (int k = 0; k < E; k++) {
            const int i = k * T + tid;
            if (i >= n) { r[k].w0 = ~0ull; r[k].w1 = ~0ull; }
            sk[i] = r[k];
        }
    }
    __syncthreads();
    const bool merge = fallback != 0;
    if (merge && binned) {
        // every record is in sk[0, n) already: pad the rest for the merge sort
        for (int i = n + tid; i < CAP; i += T) { Rec x; x.w0 = ~0ull; x.w1 = ~0ull; sk[i] = x; }
        __syncthreads();
    }
    if (merge && tid == 0 && nfall) atomicAdd(nfall, 1);
    if (merge) {
    uint64_t a0[E], a1[E];
#pragma unroll
    for (int k = 0; k < E; k++) { Rec x = sk[tid * E + k]; a0[k] = x.w0; a1[k] = x.w1; }
    // optimal 19-comparator network for 8 inputs
    PGA_CE(0, 2); PGA_CE(1, 3); PGA_CE(4, 6); PGA_CE(5, 7);
    PGA_CE(0, 4); PGA_CE(1, 5); PGA_CE(2, 6); PGA_CE(3, 7);
    PGA_CE(0, 1); PGA_CE(2, 3); PGA_CE(4, 5); PGA_CE(6, 7);
    PGA_CE(2, 4); PGA_CE(3, 5);
    PGA_CE(1, 4); PGA_CE(3, 6);
    PGA_CE(1, 2); PGA_CE(3, 4); PGA_CE(5, 6);

    for (int w = E; w < CAP; w <<= 1) {
        // runs of w are sorted; skip levels where the whole tile is one run already
        if (n <= w) break;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < E; k++) { Rec r; r.w0 = a0[k]; r.w1 = a1[k]; sk[tid * E + k] = r; }
        __syncthreads();
        const int o = tid * E;
        const int base = o & ~(2 * w - 1);
        const int diag = o - base;
        const Rec *A = sk + base;
        const Rec *B = sk + base + w;
        int lo = diag - w > 0 ? diag - w : 0;
        int hi = diag < w ? diag : w;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            Rec x = A[mid], y = B[diag - 1 - mid];
            if (key_le(x.w0, x.w1, y.w0, y.w1)) lo = mid + 1; else hi = mid;
        }
        int i = lo, j = diag - lo;
        Rec ca = A[i < w ? i : w - 1], cb = B[j < w ? j : w - 1];
#pragma unroll
        for (int k = 0; k < E; k++) {
            bool takeA = (j >= w) || (i < w && key_le(ca.w0, ca.w1, cb.w0, cb.w1));
            a0[k] = takeA ? ca.w0 : cb.w0;
            a1[k] = takeA ? ca.w1 : cb.w1;
            if (takeA) { i++; ca = A[i < w ? i : w - 1]; }
            else       { j++; cb = B[j < w ? j : w - 1]; }
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < E; k++) { Rec x; x.w0 = a0[k]; x.w1 = a1[k]; sk[tid * E + k] = x; }
    __syncthreads();
    }
    // the tile is sorted in LDS and r[] is free: start loading the next tile
    if (LB + gridDim.x < nlog) load_tile(LB + gridDim.x);
    if (!entries) {
        // keys only (the seed sort): sorted records to skeys at their positions
#pragma unroll
        for (int k = 0; k < E; k++) {
            const int idx = k * T + tid;
            if (idx < n) skeys[tl.start + idx] = sk[idx];
        }
        __syncthreads();
        continue;
    }
    // Output, record k*T + tid of the sorted tile: its LCP with its predecessor
    // (the tile's first record's is set by k_tile_fix) and its GIX entry
    // (src/kmer_adapter.c:55-62 layout): k-mer bytes 3..9, mask 0, LCP, pos LE,
    // contig|strand LE, stored at its place.
    const int e = pd.e, P = pd.P, C = pd.C;
    const uint64_t jmask = (1ull << jbits) - 1ull;
#pragma unroll
    for (int k = 0; k < E; k++) {
        const int idx = k * T + tid;
        if (idx < n) {
            const Rec x = sk[idx];
            const int64_t li = tl.start - pd.rec0 + idx;      // the record's place in its pair
            uint64_t lcp = 0;
            if (idx > 0) {
                const Rec y = sk[idx - 1];
                lcp = (uint64_t) lcp_bases(y.w0, y.w1, x.w0, x.w1);
                // the record before ends its prefix group: mark the table
                if (cum && (y.w0 >> 40) != (x.w0 >> 40))
                    atomicMax(&cum[pd.cum_off + (y.w0 >> 40)], (unsigned long long) li);
            }
            if (cum && idx == n - 1) atomicMax(&cum[pd.cum_off + (x.w0 >> 40)], (unsigned long long) (li + 1));
            const uint64_t strand = x.w1 & 1ull;
            const uint64_t j = (x.w1 >> 1) & jmask;
            const uint64_t contig = (x.w1 & 0xFFFFFFFFFFFFull) >> (jbits + 1);
            const uint64_t pos = j + 12ull * strand;
            const uint64_t cs = contig | (strand << (8 * C - 1));
            // the last 7 k-mer bytes, big-endian -> little-endian bytes 0..6, byte 7 (mask) 0
            const uint64_t low56 = ((x.w0 & 0xFFFFFFFFFFull) << 16) | (x.w1 >> 48);
            Rec ent;
            ent.w0 = __builtin_bswap64(low56 << 8);
            ent.w1 = lcp | (pos << 8) | (cs << (8 + 8 * P));
            ls[tl.start + idx] = (uint8_t) (lcp | (strand << 7));
            // the 24-bit prefix: every record's when asked for (kept keys),
            // else only the tile's ends (the LCP fix-up across tiles reads them)
            if (pre_all || idx == 0 || idx == n - 1) pre[tl.start + idx] = (uint32_t) (x.w0 >> 40);
            if (skeys) skeys[tl.start + idx] = x;
            // the entry's e bytes as two overlapping 8-byte stores (a wave's
            // lanes write one contiguous run of entries)
            uint8_t *dst = entries + pd.ent_off + li * (int64_t) e;
            const uint64_t tail = e == 16 ? ent.w1 : (ent.w0 >> (8 * (e - 8))) | (ent.w1 << (8 * (16 - e)));
            __builtin_memcpy(dst, &ent.w0, 8);
            __builtin_memcpy(dst + e - 8, &tail, 8);
        }
    }
    __syncthreads();
    }
}

template __global__ void k_tile_sort<256>(Rec *const *, const Tile *, const PairDev *, int, int, uint8_t *, uint8_t *, uint32_t *, int, unsigned long long *, Rec *, int64_t, int *, const int *);
template __global__ void k_tile_sort<512>(Rec *const *, const Tile *, const PairDev *, int, int, uint8_t *, uint8_t *, uint32_t *, int, unsigned long long *, Rec *, int64_t, int *, const int *);
template __global__ void k_tile_sort<1024>(Rec *const *, const Tile *, const PairDev *, int, int, uint8_t *, uint8_t *, uint32_t *, int, unsigned long long *, Rec *, int64_t, int *, const int *);

// Pair of a global record index.
// LCP of each tile's first record with its predecessor (0 at a pair's start),
// from the 24-bit prefixes and the 7 suffix bytes of the two entries.
__global__ void k_tile_fix(const Tile *__restrict__ tiles, int ntiles, const PairDev *__restrict__ pairs,
                           uint8_t *__restrict__ entries, uint8_t *__restrict__ ls,
                           const uint32_t *__restrict__ pre, const int *__restrict__ dn)
{
    if (dn) ntiles = *dn;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < ntiles; t += gridDim.x * blockDim.x) {
    const Tile tl = tiles[t];
    const PairDev pd = pairs[tl.pair];
    int lcp = 0;
    const int64_t li = tl.start - pd.rec0;
    uint8_t *cur = entries + pd.ent_off + li * (int64_t) pd.e;
    if (li > 0) {
        const uint8_t *prv = cur - pd.e;
        uint64_t x0 = ((uint64_t) pre[tl.start - 1] << 40), y0 = ((uint64_t) pre[tl.start] << 40);
        for (int b = 0; b < 5; b++) {
            x0 |= (uint64_t) prv[b] << (32 - 8 * b);
            y0 |= (uint64_t) cur[b] << (32 - 8 * b);
        }
        uint64_t x1 = ((uint64_t) prv[5] << 56) | ((uint64_t) prv[6] << 48);
        uint64_t y1 = ((uint64_t) cur[5] << 56) | ((uint64_t) cur[6] << 48);
        lcp = lcp_bases(x0, x1, y0, y1);
    }
    cur[8] = (uint8_t) lcp;
    ls[tl.start] = (uint8_t) ((ls[tl.start] & 0x80) | lcp);
    }
}

// ---------------------------------------------------------------- prefix table

// Prefix table of a pair (src/kmer_adapter.c:64-80: cum[p] = entries with
// prefix <= p) from the marks the tile sort leaves in it: the last record of
// every prefix group inside a tile, and a tile's last record, raise the table
// at their prefix to their count (a group split over tiles is marked from
// each, so the largest mark wins); every other prefix holds 0.  Prefix ranges
// of CR_SPAN: k_cum_bmax takes each range's largest mark, k_cum_carry the
// largest mark before each range, and k_cum_fill carries that over the
// range's empty prefixes (a block max-scan) and finds the range's largest
// group (the largest step of the table), which k_cum_max reduces per pair.
constexpr int CR_SPAN = 4096;
constexpr int CR_PER = CR_SPAN / 256;

// block (x, y): range x of pair y
__global__ __launch_bounds__(256) void k_cum_bmax(const PairDev *__restrict__ pairs, const int64_t *__restrict__ cum,
                                                  unsigned long long *__restrict__ bmax)
{
    __shared__ uint64_t red[256 / 64];
    const PairDev pd = pairs[blockIdx.y];
    const int64_t *cp = cum + pd.cum_off + (int64_t) blockIdx.x * CR_SPAN;
    uint64_t m = 0;
#pragma unroll
    for (int k = 0; k < CR_PER; k++) {
        const uint64_t x = (uint64_t) cp[k * 256 + threadIdx.x];
        m = x > m ? x : m;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { const uint64_t x = __shfl_xor(m, o, 64); m = x > m ? x : m; }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; w++) m = red[w] > m ? red[w] : m;
        bmax[(size_t) blockIdx.y * gridDim.x + blockIdx.x] = m;
    }
}

// block y: exclusive max-scan of pair y's nrange range maxima (in place)
__global__ __launch_bounds__(256) void k_cum_carry(unsigned long long *__restrict__ bmax, int nrange)
{
    __shared__ uint64_t sh[256 / 64];
    unsigned long long *b = bmax + (size_t) blockIdx.x * nrange;
    const int per = (nrange + 255) / 256, k0 = threadIdx.x * per;
    uint64_t m = 0;
    for (int k = k0; k < k0 + per && k < nrange; k++) m = b[k] > m ? b[k] : m;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint64_t inc = m;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t y = __shfl_up(inc, o, 64);
        if (lane >= o) inc = y > inc ? y : inc;
    }
    if (lane == 63) sh[wv] = inc;
    __syncthreads();
    uint64_t run = __shfl_up(inc, 1, 64);
    if (lane == 0) run = 0;
    for (int w = 0; w < wv; w++) run = sh[w] > run ? sh[w] : run;
    for (int k = k0; k < k0 + per && k < nrange; k++) {
        const uint64_t x = b[k];
        b[k] = run;
        run = x > run ? x : run;
    }
}

// block (x, y): range x of pair y filled; its largest step to gmax
__global__ __launch_bounds__(256) void k_cum_fill(const PairDev *__restrict__ pairs, int64_t *__restrict__ cum,
                                                  const unsigned long long *__restrict__ carry,
                                                  unsigned long long *__restrict__ gmax)
{
    __shared__ uint64_t sh[256 / 64];
    __shared__ unsigned long long gm;
    const PairDev pd = pairs[blockIdx.y];
    int64_t *cp = cum + pd.cum_off + (int64_t) blockIdx.x * CR_SPAN + threadIdx.x * CR_PER;
    const uint64_t cin = carry[(size_t) blockIdx.y * gridDim.x + blockIdx.x];
    if (threadIdx.x == 0) gm = 0;
    uint64_t v[CR_PER], m = cin;
#pragma unroll
    for (int k = 0; k < CR_PER; k++) { const uint64_t x = (uint64_t) cp[k]; m = x > m ? x : m; v[k] = m; }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint64_t inc = m;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t y = __shfl_up(inc, o, 64);
        if (lane >= o) inc = y > inc ? y : inc;
    }
    if (lane == 63) sh[wv] = inc;
    __syncthreads();
    uint64_t before = __shfl_up(inc, 1, 64);          // the count before this thread's prefixes
    if (lane == 0) before = cin;
    for (int w = 0; w < wv; w++) before = sh[w] > before ? sh[w] : before;
    uint64_t prev = before, big = 0;
#pragma unroll
    for (int k = 0; k < CR_PER; k++) {
        const uint64_t x = v[k] > before ? v[k] : before;
        big = x - prev > big ? x - prev : big;
        prev = x;
        cp[k] = (int64_t) x;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { const uint64_t x = __shfl_xor(big, o, 64); big = x > big ? x : big; }
    if (lane == 0) atomicMax(&gm, (unsigned long long) big);
    __syncthreads();
    if (threadIdx.x == 0) gmax[(size_t) blockIdx.y * gridDim.x + blockIdx.x] = gm;
}

// Largest prefix group of each pair: the maximum of its ranges' maxima.
__global__ __launch_bounds__(256) void k_cum_max(const int *__restrict__ plist, const unsigned long long *__restrict__ gmax,
                                                 int nrange, unsigned long long *__restrict__ maxpre)
{
    __shared__ uint64_t red[4];
    uint64_t m = 0;
    for (int i = threadIdx.x; i < nrange; i += 256) {
        const uint64_t x = gmax[(size_t) blockIdx.x * nrange + i];
        m = x > m ? x : m;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { const uint64_t x = __shfl_xor(m, o, 64); m = x > m ? x : m; }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t x = red[0];
        for (int w = 1; w < 4; w++) x = red[w] > x ? red[w] : x;
        maxpre[plist[blockIdx.x]] = x;
    }
}

// ---------------------------------------------------------------- adaptamer merge

// The interval of anchor g (global index) in a window of LCP/strand bytes.
// Returns the partner counts (same strand, other strand), or -1 if the anchor
// emits nothing.  L(k) < 12 marks a prefix-group start; outside the pair the
// window holds 0.  Follows src/fastga_pipeline.c:315-353: plen = max(LCP[g],
// LCP[g+1]) with group edges < 12, [a,b) the maximal run sharing >= plen
// bases, skipped when b - a >= freq.
__device__ __forceinline__ bool anchor_interval(const uint8_t *__restrict__ win, int gi, int freq,
                                                int &plen, int &a, int &b)
{
    int Li = win[gi] & 63, Ln = win[gi + 1] & 63;
    int li = Li >= 12 ? Li : 0, ln = Ln >= 12 ? Ln : 0;
    plen = li > ln ? li : ln;
    if (plen < 12) return false;
    a = gi;
    int steps = 0;
    while (steps < freq && (win[a] & 63) >= plen) { a--; steps++; }
    if ((win[a] & 63) >= plen) return false;      // the run reaches past FREQ entries back
    b = gi + 1;
    while (b - a < freq && (win[b] & 63) >= plen) b++;
    return b - a < freq;
}

__device__ __forceinline__ int merge_block_pair(const PairDev *__restrict__ pairs, int npairs, int64_t b)
{
    int lo = 0, hi = npairs;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (pairs[mid].mblk0 <= b) lo = mid; else hi = mid;
    }
    return lo;
}

constexpr int MTH = 256;             // threads of the merge kernels
constexpr int BUCK_LDS = 1024;       // per-pair contigs counted in LDS by the merge

// LCP/strand bytes of window [w0, w0+n) of the pair into win (0 outside it),
// n <= NMAX.  Every load of the thread is issued before the first is used.
template <int NMAX>
__device__ __forceinline__ void load_win(uint8_t *__restrict__ win, const uint8_t *__restrict__ ls,
                                         int64_t w0, int n, int64_t rec0, int64_t end)
{
    constexpr int WI = (NMAX + MTH - 1) / MTH;
    uint8_t x[WI];
#pragma unroll
    for (int j = 0; j < WI; j++) {
        const int k = (int) threadIdx.x + j * MTH;
        const int64_t g = w0 + k;
        x[j] = k < n && g >= rec0 && g < end ? ls[g] : 0;
    }
#pragma unroll
    for (int j = 0; j < WI; j++) {
        const int k = (int) threadIdx.x + j * MTH;
        if (k < n) win[k] = x[j];
    }
}

// Per-anchor interval record handed from the count pass to the emit pass, six
// bits a field (FREQ <= 64 keeps every field below 64): plen, the interval's
// start and end relative to the anchor (gi - a, b - gi), and the anchor's
// same-strand and cross-strand partner counts.  0 = the anchor emits nothing.
__device__ __forceinline__ uint32_t minfo_pack(int plen, int aoff, int boff, uint32_t nn, uint32_t nc)
{
    return (uint32_t) plen | (uint32_t) aoff << 6 | (uint32_t) boff << 12 | nn << 18 | nc << 24;
}

// 8 window bytes win[i..i+8) as a little-endian word (LDS reads need no alignment).
__device__ __forceinline__ uint64_t win8(const uint8_t *__restrict__ win, int i)
{
    uint64_t w;
    __builtin_memcpy(&w, win + i, 8);
    return w;
}

// Bit 7 of byte k is set iff the LCP (bits 0..5) of byte k of w is below plen
// (12 <= plen <= 63): (LCP | 0x80) - plen keeps bit 7 exactly when LCP >= plen,
// and no byte borrows from its neighbour.
__device__ __forceinline__ uint64_t lcp_below(uint64_t w, uint32_t plen)
{
    const uint64_t t = ((w & 0x3F3F3F3F3F3F3F3Full) | 0x8080808080808080ull) - 0x0101010101010101ull * plen;
    return ~t & 0x8080808080808080ull;
}

// anchor_interval plus the partner counts, for FREQ <= 16, without loops: the
// 16 window bytes on each side of the anchor are compared with plen eight at a
// time; the run lengths sharing >= plen bases come from leading/trailing zero
// counts and the strand split from popcounts of the strand bits (bit 7).
// Needs 15 window bytes before gi and 16 after it.  Returns the interval
// record (0 = no seeds).
__device__ __forceinline__ uint32_t anchor_info16(const uint8_t *__restrict__ win, int gi, int freq,
                                                  uint32_t &xn, uint32_t &xc)
{
    const uint64_t b0 = win8(win, gi - 7), f0 = win8(win, gi + 1);    // gi-7..gi, gi+1..gi+8
    const uint32_t Li = (uint32_t) (b0 >> 56) & 63u, Ln = (uint32_t) f0 & 63u;
    const uint32_t li = Li >= 12 ? Li : 0, ln = Ln >= 12 ? Ln : 0;
    const uint32_t plen = li > ln ? li : ln;
    xn = xc = 0;
    if (plen < 12) return 0;
    const uint64_t lb0 = lcp_below(b0, plen), lf0 = lcp_below(f0, plen);
    // entries gi, gi-1, ... sharing >= plen (cb), and gi+1, gi+2, ... (cf):
    // the interval is [gi - cb, gi + cf + 1).  The next 8 bytes on a side
    // (gi-15..gi-8, gi+9..gi+16) are read only when its run reaches them.
    uint64_t b1 = 0, f1 = 0;
    int cb = lb0 ? __builtin_clzll(lb0) >> 3 : 8;
    if (!lb0) {
        b1 = win8(win, gi - 15);
        const uint64_t lb1 = lcp_below(b1, plen);
        cb = 8 + (lb1 ? __builtin_clzll(lb1) >> 3 : 8);
    }
    int cf = lf0 ? __builtin_ctzll(lf0) >> 3 : 8;
    if (!lf0) {
        f1 = win8(win, gi + 9);
        const uint64_t lf1 = lcp_below(f1, plen);
        cf = 8 + (lf1 ? __builtin_ctzll(lf1) >> 3 : 8);
    }
    const int m = cb + 1 + cf;
    if (m >= freq) return 0;                     // so cb, cf <= 14 below
    // strand bits of the partners gi-cb..gi-1 and gi+1..gi+cf
    const uint64_t HI = 0x8080808080808080ull;
    const uint64_t mb0 = 0x00FFFFFFFFFFFFFFull & ~((1ull << (cb >= 7 ? 0 : 8 * (7 - cb))) - 1);
    const uint64_t mb1 = cb >= 8 ? ~((1ull << (8 * (15 - cb))) - 1) : 0ull;
    const uint64_t mf0 = cf >= 8 ? ~0ull : (1ull << (8 * cf)) - 1;
    const uint64_t mf1 = cf > 8 ? (1ull << (8 * (cf - 8))) - 1 : 0ull;
    const uint32_t ones = __popcll(b0 & mb0 & HI) + __popcll(b1 & mb1 & HI) +
                          __popcll(f0 & mf0 & HI) + __popcll(f1 & mf1 & HI);
    const uint32_t part = (uint32_t) m - 1;
    xn = (b0 >> 63) ? ones : part - ones;
    xc = part - xn;
    return minfo_pack((int) plen, cb, cf + 1, xn, xc);
}

// anchor_info16 with the partners restricted to query entries (the serving
// join: a target anchor pairs only with query entries).  qf points at the
// query-flag bytes (bit 0) of the window, indexed like win: qf[x] for window
// entry x; the 8-byte reads around the anchor may touch bytes outside the
// interval, which the interval masks drop.
__device__ __forceinline__ uint32_t anchor_info16_q(const uint8_t *__restrict__ win, const uint8_t *__restrict__ qf,
                                                    int gi, int freq, uint32_t &xn, uint32_t &xc)
{
    const uint64_t b0 = win8(win, gi - 7), f0 = win8(win, gi + 1);
    const uint32_t Li = (uint32_t) (b0 >> 56) & 63u, Ln = (uint32_t) f0 & 63u;
    const uint32_t li = Li >= 12 ? Li : 0, ln = Ln >= 12 ? Ln : 0;
    const uint32_t plen = li > ln ? li : ln;
    xn = xc = 0;
    if (plen < 12) return 0;
    const uint64_t lb0 = lcp_below(b0, plen), lf0 = lcp_below(f0, plen);
    uint64_t b1 = 0, f1 = 0;
    int cb = lb0 ? __builtin_clzll(lb0) >> 3 : 8;
    if (!lb0) {
        b1 = win8(win, gi - 15);
        const uint64_t lb1 = lcp_below(b1, plen);
        cb = 8 + (lb1 ? __builtin_clzll(lb1) >> 3 : 8);
    }
    int cf = lf0 ? __builtin_ctzll(lf0) >> 3 : 8;
    if (!lf0) {
        f1 = win8(win, gi + 9);
        const uint64_t lf1 = lcp_below(f1, plen);
        cf = 8 + (lf1 ? __builtin_ctzll(lf1) >> 3 : 8);
    }
    const int m = cb + 1 + cf;
    if (m >= freq) return 0;
    const uint64_t HI = 0x8080808080808080ull;
    const uint64_t mb0 = 0x00FFFFFFFFFFFFFFull & ~((1ull << (cb >= 7 ? 0 : 8 * (7 - cb))) - 1);
    const uint64_t mb1 = cb >= 8 ? ~((1ull << (8 * (15 - cb))) - 1) : 0ull;
    const uint64_t mf0 = cf >= 8 ? ~0ull : (1ull << (8 * cf)) - 1;
    const uint64_t mf1 = cf > 8 ? (1ull << (8 * (cf - 8))) - 1 : 0ull;
    // query flags moved to bit 7 of each byte, like the strand bits
    auto q7 = [&](int i) -> uint64_t { return (win8(qf, i) & 0x0101010101010101ull) << 7; };
    const uint64_t qb0 = mb0 & HI & q7(gi - 7), qb1 = mb1 & HI & (mb1 ? q7(gi - 15) : 0ull);
    const uint64_t qf0 = mf0 & HI & q7(gi + 1), qf1 = mf1 & HI & (mf1 ? q7(gi + 9) : 0ull);
    const uint32_t part = __popcll(qb0) + __popcll(qb1) + __popcll(qf0) + __popcll(qf1);
    const uint32_t ones = __popcll(b0 & qb0) + __popcll(b1 & qb1) + __popcll(f0 & qf0) + __popcll(f1 & qf1);
    xn = (b0 >> 63) ? ones : part - ones;
    xc = part - xn;
    return minfo_pack((int) plen, cb, cf + 1, xn, xc);
}

// Seed counts of the tiles of A*MTH anchors (A consecutive anchors per thread).
template <int A>
__global__ __launch_bounds__(MTH) void k_merge_count(const PairDev *__restrict__ pairs, int npairs,
                                                     const uint8_t *__restrict__ ls, int freq,
                                                     uint32_t *__restrict__ cntN, uint32_t *__restrict__ cntC,
                                                     uint32_t *__restrict__ minfo, int64_t nlog,
                                                     const int64_t *__restrict__ dn)
{
    constexpr int TA = MTH * A;
    constexpr int NW = TA + 2 * MAX_FREQ + 4;
    constexpr int WI = (NW + MTH - 1) / MTH;
    if (dn) nlog = *dn;
    __shared__ uint8_t win[NW];
    __shared__ uint64_t red[MTH / 64 + 1];
    const bool swar = freq <= 16;
    const int H = swar ? 17 : freq + 1;          // window halo on each side
    // A tile's window bytes (its LCP/strand bytes [lo - H, lo + TA + H), 0
    // outside the pair) are loaded into registers one tile ahead, so their
    // latency overlaps the previous tile's evaluation.
    uint8_t x[WI];
    auto fetch = [&](int64_t lb, int64_t &lo, int64_t &end) {
        const PairDev &pd = pairs[merge_block_pair(pairs, npairs, lb)];
        lo = pd.rec0 + (lb - pd.mblk0) * (int64_t) TA;
        end = pd.rec0 + pd.nrec;
        // window bytes [kmin, kmax) lie inside the pair (the tile's first
        // anchor does, so kmin <= H < kmax)
        const int64_t w0 = lo - H;
        const int kmin = (int) (pd.rec0 > w0 ? pd.rec0 - w0 : 0);
        const int kmax = (int) (end - w0 < TA + 2 * H ? end - w0 : TA + 2 * H);
#pragma unroll
        for (int j = 0; j < WI; j++) {
            const int k = (int) threadIdx.x + j * MTH;
            x[j] = k >= kmin && k < kmax ? ls[w0 + k] : 0;
        }
    };
    int64_t nlo = 0, nend = 0;
    if ((int64_t) blockIdx.x < nlog) fetch(blockIdx.x, nlo, nend);
    for (int64_t LB = blockIdx.x; LB < nlog; LB += gridDim.x) {
    const int64_t lo = nlo, end = nend;
#pragma unroll
    for (int j = 0; j < WI; j++) {
        const int k = (int) threadIdx.x + j * MTH;
        if (k < TA + 2 * H) win[k] = x[j];
    }
    __syncthreads();
    if (LB + gridDim.x < nlog) fetch(LB + gridDim.x, nlo, nend);
    uint32_t nN = 0, nC = 0;
    // the thread's anchors, with the search chosen once per tile (loop-free
    // for FREQ <= 16)
    auto anchors = [&](auto swar_c) {
#pragma unroll
        for (int k = 0; k < A; k++) {
            const int gi = H + threadIdx.x * A + k;
            const int64_t g = lo + threadIdx.x * A + k;
            uint32_t info = 0, xn = 0, xc = 0;
            if (decltype(swar_c)::value) {
                info = anchor_info16(win, gi, freq, xn, xc);
            } else {
                int plen, a, b;
                if (anchor_interval(win, gi, freq, plen, a, b)) {
                    const uint8_t si = win[gi] & 0x80;
                    for (int q = a; q < b; q++) {
                        if (q == gi) continue;
                        if ((win[q] & 0x80) == si) xn++; else xc++;
                    }
                    info = minfo_pack(plen, gi - a, b - gi, xn, xc);
                }
            }
            if (g < end) {
                minfo[g] = info;
                nN += xn;
                nC += xc;
            }
        }
    };
    if (swar) anchors(std::integral_constant<bool, true>());
    else anchors(std::integral_constant<bool, false>());
    const uint64_t sum = block_sum_u64<MTH>((uint64_t) nN | (uint64_t) nC << 32, red);
    if (threadIdx.x == 0) { cntN[LB] = (uint32_t) sum; cntC[LB] = (uint32_t) (sum >> 32); }
    __syncthreads();
    }
}

__device__ __forceinline__ void flush_buck(uint32_t *__restrict__ lb, int n, int64_t off,
                                           unsigned long long *__restrict__ nbuck,
                                           unsigned long long *__restrict__ cbuck)
{
    for (int c = threadIdx.x; c < n; c += MTH) {
        if (lb[c]) { atomicAdd(&nbuck[off + c], (unsigned long long) lb[c]); lb[c] = 0; }
        if (lb[BUCK_LDS + c]) { atomicAdd(&cbuck[off + c], (unsigned long long) lb[BUCK_LDS + c]); lb[BUCK_LDS + c] = 0; }
    }
}

// Bytes [s-8, s) and [0, 8) of a seed whose first 1+pc bytes are hd (plen,
// then the anchor payload) and whose last pc bytes are pq (the partner
// payload), for pc >= 4.  Both 8-byte words lie inside the seed, so two
// stores of them write exactly its s = 1 + 2 pc bytes.
__device__ __forceinline__ void seed_words(uint64_t hd, uint64_t pq, int pc, uint64_t &w0, uint64_t &w1)
{
    w0 = hd | ((pq << (8 * pc)) << 8);
    w1 = (hd >> (8 * (2 * pc - 7))) | (pq << (8 * (8 - pc)));
}

constexpr int SMAP = 2048;           // seeds of each stream a merge-emit round maps in LDS

// Seeds of a run of merge tiles (A*MTH anchors each, A consecutive per
// thread), from the interval records of the count pass.  Per tile, LDS holds
// one payload per window entry (the tile's anchors plus a FREQ halo each side;
// the strand is the payload's top bit) and the window's strand bits; a block
// scan places each anchor's seeds in the tile's slices of the N and C streams.
// Each round of up to SMAP seeds a stream:
//  - every anchor's thread lists its seeds in an LDS map, one aligned 32-bit
//    word a seed (anchor index | partner window index << 11 | plen << 22),
//    walking the set bits of its same-strand and other-strand partner masks
//    (a short loop per anchor: the walk is the only lane-divergent part);
//  - then one lane takes one seed: its map word and the two payloads give its
//    bytes, stored straight to HBM as two overlapping 8-byte stores (4-byte
//    for payloads under 4 bytes).  Consecutive lanes hold consecutive seeds,
//    so a wave's stores cover one contiguous run of the stream.
// Per-anchor-contig counts accumulate in LDS while the pair stays the same.
template <int A>
__global__ __launch_bounds__(MTH) void k_merge_emit(const PairDev *__restrict__ pairs, int npairs,
                                                    const uint8_t *__restrict__ entries, int freq,
                                                    const int64_t *__restrict__ ofsN,
                                                    const int64_t *__restrict__ ofsC,
                                                    uint8_t *__restrict__ seedN, uint8_t *__restrict__ seedC,
                                                    unsigned long long *__restrict__ nbuck,
                                                    unsigned long long *__restrict__ cbuck,
                                                    const uint32_t *__restrict__ minfo,
                                                    int64_t nmb, int64_t per_block,
                                                    const int64_t *__restrict__ dn, const int *__restrict__ skip)
{
    // device-driven form: the tile count from dn, an even share per block,
    // and nothing written when the streams overflow their buffers (*skip)
    if (skip && *skip) return;
    if (dn) {
        nmb = *dn;
        per_block = (nmb + gridDim.x - 1) / gridDim.x;
    }
    constexpr int TA = MTH * A;
    constexpr int W = TA + 2 * MAX_FREQ + 4;
    constexpr int NSW = (W + 63) / 64;
    static_assert(W <= 2048, "anchor and partner window indexes fit 11 bits of a map word");
    __shared__ uint64_t pay[W];
    __shared__ uint64_t sbits[NSW + 1];             // strand bit of window entry k: bit k & 63 of word k >> 6
    __shared__ uint32_t smap[2][SMAP];              // the round's N and C seeds: anchor | partner << 11 | plen << 22
    __shared__ uint32_t lb[2 * BUCK_LDS];
    __shared__ uint32_t sh[MTH / 64 + 1];
    for (int c = threadIdx.x; c < 2 * BUCK_LDS; c += MTH) lb[c] = 0;
    if (threadIdx.x == 0) sbits[NSW] = 0;
    const int64_t t0 = (int64_t) blockIdx.x * per_block;
    const int64_t t1 = t0 + per_block < nmb ? t0 + per_block : nmb;
    const int lane = (int) threadIdx.x & 63, wv = (int) threadIdx.x >> 6;
    int cur_pair = -1;
    for (int64_t tile = t0; tile < t1; tile++) {
        const int p = merge_block_pair(pairs, npairs, tile);
        const PairDev pd = pairs[p];
        if (p != cur_pair) {
            __syncthreads();
            if (cur_pair >= 0 && pairs[cur_pair].ncontig <= BUCK_LDS)
                flush_buck(lb, pairs[cur_pair].ncontig, pairs[cur_pair].buck_off, nbuck, cbuck);
            cur_pair = p;
        }
        const bool lds_buck = pd.ncontig <= BUCK_LDS;
        const int64_t lo = pd.rec0 + (tile - pd.mblk0) * (int64_t) TA;
        const int64_t end = pd.rec0 + pd.nrec;
        const int H = freq + 1;
        const int pc = pd.P + pd.C, e = pd.e, s = pd.s;
        const int sbit = 8 * pc - 1;
        const uint64_t pmask = (1ull << (8 * pc)) - 1;
        // interval records of this thread's anchors (0 past the pair's end)
        uint32_t v[A];
#pragma unroll
        for (int k = 0; k < A; k++) {
            const int64_t g = lo + threadIdx.x * A + k;
            v[k] = g < end ? minfo[g] : 0u;
        }
        __syncthreads();                                 // the previous tile is done with the LDS arrays
        {
            // window entry k is global entry lo-H+k; only those inside the pair
            // are read (intervals never leave the pair).  The 8-byte load at
            // the payload may run past an entry: the entries buffer is padded.
            const int64_t wlo = lo - H > pd.rec0 ? lo - H : pd.rec0;
            const int64_t whi = lo + TA + H < end ? lo + TA + H : end;
            const uint8_t *ew = entries + pd.ent_off + (lo - H - pd.rec0) * (int64_t) e + 9;
            const int k0 = (int) (wlo - (lo - H)), k1 = (int) (whi - (lo - H));
            // every load of the thread is issued before the first one is used
            constexpr int WI = (W + MTH - 1) / MTH;
            uint64_t x[WI];
#pragma unroll
            for (int j = 0; j < WI; j++) {
                const int k = k0 + (int) threadIdx.x + j * MTH;
                if (k < k1) __builtin_memcpy(&x[j], ew + (int64_t) k * e, 8);
            }
#pragma unroll
            for (int j = 0; j < WI; j++) {
                const int k = k0 + (int) threadIdx.x + j * MTH;
                if (k < k1) pay[k] = x[j] & pmask;
            }
        }
        uint32_t sN = 0, sC = 0;
#pragma unroll
        for (int k = 0; k < A; k++) {
            sN += (v[k] >> 18) & 63u;
            sC += (v[k] >> 24) & 63u;
        }
        uint32_t tot;
        const uint32_t off = block_excl_scan<MTH, uint32_t>(sN | sC << 16, sh, &tot);   // also orders pay[]
        const int totN = (int) (tot & 0xFFFFu), totC = (int) (tot >> 16);
        // strand bits of the window (entries outside the pair read as garbage;
        // no interval reaches them)
        for (int w = wv; w < NSW; w += MTH / 64) {
            const int k = 64 * w + lane;
            const uint64_t b = __ballot(k < W && ((pay[k < W ? k : 0] >> sbit) & 1ull));
            if (lane == 0) sbits[w] = b;
        }
        // per-anchor-contig counts (src/fastga_pipeline.c:403)
#pragma unroll
        for (int k = 0; k < A; k++) {
            if (!(v[k] >> 18)) continue;
            const uint32_t nn = (v[k] >> 18) & 63u, nc = (v[k] >> 24) & 63u;
            const int gi = H + (int) threadIdx.x * A + k;
            const uint64_t icont = (pay[gi] >> (8 * pd.P)) & ((1ull << (8 * pd.C - 1)) - 1);
            if (lds_buck) {
                if (nn) atomicAdd(&lb[icont], nn);
                if (nc) atomicAdd(&lb[BUCK_LDS + icont], nc);
            } else {
                if (nn) atomicAdd(&nbuck[pd.buck_off + icont], (unsigned long long) nn);
                if (nc) atomicAdd(&cbuck[pd.buck_off + icont], (unsigned long long) nc);
            }
        }
        __syncthreads();
        // each anchor's partners over its interval [qa, qb): same strand (N)
        // and other strand (C), bit j = window entry qa + j
        uint64_t mn[A], mc[A];
#pragma unroll
        for (int k = 0; k < A; k++) {
            mn[k] = mc[k] = 0;
            if (!(v[k] >> 18)) continue;
            const int aoff = (int) ((v[k] >> 6) & 63u), len = aoff + (int) ((v[k] >> 12) & 63u);
            const int qa = H + (int) threadIdx.x * A + k - aoff;
            const int w = qa >> 6, o = qa & 63;
            const uint64_t bits = o ? (sbits[w] >> o) | (sbits[w + 1] << (64 - o)) : sbits[w];
            const uint64_t range = len >= 64 ? ~0ull : (1ull << len) - 1;
            const uint64_t me = 1ull << aoff;
            mn[k] = ((bits & me) ? bits : ~bits) & range & ~me;
            mc[k] = range & ~mn[k] & ~me;
        }
        uint8_t *const gN = seedN + pd.nseed_off + (ofsN[tile] - ofsN[pd.mblk0]) * (int64_t) s;
        uint8_t *const gC = seedC + pd.cseed_off + (ofsC[tile] - ofsC[pd.mblk0]) * (int64_t) s;
        const uint64_t clear = ~(1ull << sbit);
        const int rounds = (max(totN, totC) + SMAP - 1) / SMAP;
        for (int r = 0; r < rounds; r++) {
            // (a stream with fewer rounds than the other has none left: n1 = n0)
            const int n0 = r * SMAP, n1 = max(n0, min(n0 + SMAP, totN));
            const int c0 = r * SMAP, c1 = max(c0, min(c0 + SMAP, totC));
            // map the round's seeds: anchors in order, partners ascending
            {
                int kn = (int) (off & 0xFFFFu), kc = (int) (off >> 16);
#pragma unroll
                for (int k = 0; k < A; k++) {
                    const int i = (int) threadIdx.x * A + k;
                    const uint32_t qa = (uint32_t) (H + i) - ((v[k] >> 6) & 63u);
                    const uint32_t tag = (uint32_t) i | qa << 11 | (v[k] & 63u) << 22;
                    if (rounds == 1) {
                        for (uint64_t m = mn[k]; m; m &= m - 1) smap[0][kn++] = tag + ((uint32_t) __builtin_ctzll(m) << 11);
                        for (uint64_t m = mc[k]; m; m &= m - 1) smap[1][kc++] = tag + ((uint32_t) __builtin_ctzll(m) << 11);
                    } else {
                        for (uint64_t m = mn[k]; m; m &= m - 1, kn++)
                            if (kn >= n0 && kn < n1) smap[0][kn - n0] = tag + ((uint32_t) __builtin_ctzll(m) << 11);
                        for (uint64_t m = mc[k]; m; m &= m - 1, kc++)
                            if (kc >= c0 && kc < c1) smap[1][kc - c0] = tag + ((uint32_t) __builtin_ctzll(m) << 11);
                    }
                }
            }
            __syncthreads();
            // the round's seeds, one lane a seed: consecutive lanes write
            // consecutive seeds, so a wave's stores cover one contiguous run
            const int nN = n1 - n0, nR = nN + c1 - c0;
            for (int u = (int) threadIdx.x; u < nR; u += MTH) {
                const int st = u >= nN;                  // 0: N stream, 1: C stream
                const int t = st ? u - nN : u;
                const uint32_t m = smap[st][t];
                const uint64_t pa = pay[H + (int) (m & 0x7FFu)], pq = pay[(m >> 11) & 0x7FFu];
                const uint64_t hd = (uint64_t) (m >> 22) | ((pa & clear) << 8);
                uint8_t *const dst = st ? gC + (int64_t) (c0 + t) * s : gN + (int64_t) (n0 + t) * s;
                if (pc >= 4) {
                    uint64_t w0, w1;
                    seed_words(hd, pq, pc, w0, w1);
                    __builtin_memcpy(dst, &w0, 8);
                    __builtin_memcpy(dst + s - 8, &w1, 8);
                } else {
                    const uint64_t w = hd | (pq << (8 * (1 + pc)));
                    const uint32_t x0 = (uint32_t) w, x1 = (uint32_t) (w >> (8 * (s - 4)));
                    __builtin_memcpy(dst, &x0, 4);
                    __builtin_memcpy(dst + s - 4, &x1, 4);
                }
            }
            __syncthreads();
        }
    }
    __syncthreads();
    if (cur_pair >= 0 && pairs[cur_pair].ncontig <= BUCK_LDS)
        flush_buck(lb, pairs[cur_pair].ncontig, pairs[cur_pair].buck_off, nbuck, cbuck);
}

// ---------------------------------------------------------------- device-driven pipeline

// Per-pair layout from the record ranges pr[0..npairs]: records, GIX entry
// offsets, merge tiles, and the first partition level's segments (pairs of
// more than cap records) and tiles (the others); also clears the status.
// The load-time fields of pairs[] (P, C, e, s, ncontig, cum_off, buck_off)
// are already set.
__global__ __launch_bounds__(256) void k_graph_layout(const int64_t *__restrict__ pr, int npairs, int mt, int cap,
                                                      PairDev *__restrict__ pairs, int64_t *__restrict__ mblk,
                                                      Seg *__restrict__ seg0, Tile *__restrict__ tile0,
                                                      GStat *__restrict__ st)
{
    if (threadIdx.x == 0) {
        int64_t ent = 0, mb = 0;
        int nseg = 0, nt = 0;
        for (int p = 0; p < npairs; p++) {
            PairDev d = pairs[p];
            d.rec0 = pr[p];
            d.nrec = pr[p + 1] - pr[p];
            d.ent_off = ent;
            ent += d.nrec * d.e;
            d.mblk0 = mb;
            mblk[p] = mb;
            mb += (d.nrec + mt - 1) / mt;
            pairs[p] = d;
            if (d.nrec > cap) seg0[nseg++] = Seg{d.rec0, d.nrec, p, 1, 1, 0};
            else if (d.nrec > 0) tile0[nt++] = Tile{d.rec0, (int32_t) d.nrec, p, 1, 1, drange_all(8)};
        }
        mblk[npairs] = mb;
        st->R = pr[npairs];
        st->nmb = mb;
        st->nsegs[0] = 0;
        st->nsegs[1] = nseg;
        for (int l = 2; l < GRAPH_LEVELS + 2; l++) st->nsegs[l] = 0;
        st->ntiles = nt;
        st->overflow = 0;
        st->bad = 0;
        st->capped = 0;
        st->kept = 0;
    }
}

// A device-side count held to its list's capacity before kernels read it
// (the lists are sized from a bound no correct count exceeds; *capped tells
// the host to redo the work).
__global__ void k_clamp_count(int *__restrict__ cnt, int cap, int *__restrict__ capped)
{
    if (threadIdx.x == 0 && *cnt > cap) {
        *cnt = cap;
        *capped = 1;
    }
}

// Seed stream offsets of the pairs from the per-pair seed ranges (N in
// pofs[0..npairs], C in pofs[npairs+1..]); flags an overflow of either
// stream's buffer.
__global__ __launch_bounds__(64) void k_graph_seeds(const int64_t *__restrict__ pofs, int npairs,
                                                    PairDev *__restrict__ pairs, int64_t capN, int64_t capC,
                                                    GStat *__restrict__ st)
{
    if (threadIdx.x == 0) {
        int64_t nb = 0, cb = 0;
        for (int p = 0; p < npairs; p++) {
            const int s = pairs[p].s;
            pairs[p].nseed_off = nb;
            pairs[p].cseed_off = cb;
            nb += (pofs[p + 1] - pofs[p]) * s;
            cb += (pofs[npairs + 1 + p + 1] - pofs[npairs + 1 + p]) * s;
        }
        st->overflow = nb + 16 > capN || cb + 16 > capC;
    }
}

// ---------------------------------------------------------------- an index from GIX entries

// pre[i] = the 24-bit prefix of entry i, from the cumulative prefix counts
// (entries [cum[p-1], cum[p]) have prefix p).
__global__ void k_gix_prefix(const int64_t *__restrict__ cum, uint32_t *__restrict__ pre)
{
    for (int64_t p = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; p < (1 << 24);
         p += (int64_t) gridDim.x * blockDim.x) {
        const int64_t lo = p ? cum[p - 1] : 0, hi = cum[p];
        for (int64_t i = lo; i < hi; i++) pre[i] = (uint32_t) p;
    }
}

// The sorted 16-byte keys of GIX entries (the layout k_tile_sort writes): the
// k-mer from the prefix and the 7 suffix bytes, the offset j = pos - 12 strand,
// the contig and the strand from the payload; and each entry's LCP/strand byte.
__global__ void k_gix_keys(const uint8_t *__restrict__ ent, int e, int P, int C, const uint32_t *__restrict__ pre,
                           int64_t n, int jbits, Rec *__restrict__ keys, uint8_t *__restrict__ ls)
{
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x) {
        const uint8_t *d = ent + i * e;
        uint64_t suf = 0;                          // bases 12..39, MSB-first (56 bits)
        for (int k = 0; k < 7; k++) suf = (suf << 8) | d[k];
        uint64_t pos = 0, cs = 0;
        for (int k = 0; k < P; k++) pos |= (uint64_t) d[9 + k] << (8 * k);
        for (int k = 0; k < C; k++) cs |= (uint64_t) d[9 + P + k] << (8 * k);
        const uint64_t strand = cs >> (8 * C - 1);
        const uint64_t contig = cs & ((1ull << (8 * C - 1)) - 1);
        const uint64_t j = pos - 12 * strand;
        Rec r;
        r.w0 = ((uint64_t) pre[i] << 40) | (suf >> 16);
        r.w1 = ((suf & 0xFFFFull) << 48) | (contig << (jbits + 1)) | (j << 1) | strand;
        keys[i] = r;
        ls[i] = (uint8_t) ((d[8] & 63) | (strand << 7));
    }
}

// ---------------------------------------------------------------- checks

// 80-bit k-mer of GIX entry i: 24-bit prefix, then the entry's 7 suffix bytes.
__device__ __forceinline__ void entry_kmer(const uint8_t *__restrict__ ent, int e, const uint32_t *__restrict__ pre,
                                           int64_t g, int64_t li, uint64_t &k0, uint64_t &k1)
{
    const uint8_t *x = ent + li * e;
    uint64_t suf = 0;
#pragma unroll
    for (int b = 0; b < 7; b++) suf = (suf << 8) | x[b];
    k0 = (pre ? (uint64_t) pre[g] << 40 : 0ull) | (suf >> 16);
    k1 = (suf & 0xFFFFull) << 48;
}

// Prefix of pair entry i from the pair's prefix table: the first p in
// [lo, hi] with cum[p] > i (cum[hi] > i).
__device__ __forceinline__ uint32_t prefix_of(const int64_t *__restrict__ cum, int64_t i, uint32_t lo, uint32_t hi)
{
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (cum[mid] > i) hi = mid; else lo = mid + 1;
    }
    return lo;
}

// Sortedness and LCP check of one pair's index (src/dsort.c:229-246): bad
// counts entries whose k-mer is smaller than its predecessor's; lcp_bad
// counts entries whose LCP byte differs from the recomputed LCP with the
// predecessor (0 for the first entry).  out = {bad, first_bad, lcp_bad}.
// Prefixes come from pre (every record's) or, when cum is given, from the
// prefix table: a block's 256 consecutive entries search only between the
// prefixes of the entries before and after them.
__global__ void k_verify(const uint8_t *__restrict__ ent, int e, const uint32_t *__restrict__ pre,
                         const int64_t *__restrict__ cum, int64_t rec0, int64_t nrec,
                         unsigned long long *__restrict__ out)
{
    __shared__ uint64_t red[256 / 64 + 1];
    __shared__ uint32_t span[2];
    uint64_t bad = 0, lbad = 0, first = ~0ull;
    // grid-stride over blocks of 256 consecutive entries (uniform trip count)
    for (int64_t b = (int64_t) blockIdx.x * blockDim.x; b < nrec; b += (int64_t) gridDim.x * blockDim.x) {
        const int64_t i = b + threadIdx.x;
        uint32_t plo = 0, phi = 0;
        if (cum) {
            __syncthreads();                            // the previous round is done with span
            if (threadIdx.x == 0) span[0] = prefix_of(cum, b > 0 ? b - 1 : 0, 0, (1u << 24) - 1);
            if (threadIdx.x == 1) span[1] = prefix_of(cum, min(b + 255, nrec - 1), 0, (1u << 24) - 1);
            __syncthreads();
            plo = span[0];
            phi = span[1];
        }
        if (i >= nrec) continue;
        auto kmer = [&](int64_t x, uint64_t &k0, uint64_t &k1) {
            if (cum) {
                const uint64_t p = prefix_of(cum, x, plo, phi);
                entry_kmer(ent, e, nullptr, 0, x, k0, k1);
                k0 |= p << 40;
            } else {
                entry_kmer(ent, e, pre, rec0 + x, x, k0, k1);
            }
        };
        uint64_t b0, b1;
        kmer(i, b0, b1);
        int l = 0;
        if (i > 0) {
            uint64_t a0, a1;
            kmer(i - 1, a0, a1);
            if (a0 > b0 || (a0 == b0 && a1 > b1)) { bad++; if ((uint64_t) i < first) first = (uint64_t) i; }
            l = lcp_bases(a0, a1, b0, b1);
        }
        if (ent[i * e + 8] != (uint8_t) l) lbad++;
    }
    bad = block_sum_u64<256>(bad, red);
    lbad = block_sum_u64<256>(lbad, red);
    if (threadIdx.x == 0) {
        if (bad) atomicAdd(&out[0], (unsigned long long) bad);
        if (lbad) atomicAdd(&out[2], (unsigned long long) lbad);
    }
    if (first != ~0ull) atomicMin(&out[1], (unsigned long long) first);
}

// Records per 5-base bucket (src/dsort.h:30-33: the first 10 bits of the
// k-mer) from a pair's cumulative 24-bit prefix table.
__global__ void k_bucket_counts(const int64_t *__restrict__ cum, long long *__restrict__ out)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < 1024) {
        const int64_t hi = cum[((b + 1) << 14) - 1];
        const int64_t lo = b ? cum[(b << 14) - 1] : 0;
        out[b] = (long long) (hi - lo);
    }
}

// ---------------------------------------------------------------- seed sort
//
// The diagonal records of one seed stream (reimport_thread,
// src/fastga_pipeline.c:1391-1450): [lcp][diag&63][anti DBYTE][band DBYTE]
// [jcont C], all little-endian.  rmsd_sort orders each anchor-contig panel on
// the whole record read from its last byte (lib/RSDsort.c:292-377), i.e. on
// the record as one little-endian number; prefixing that number with the
// anchor contig gives one key whose plain sort is the panel-by-panel order.
// The key holds every field at its exact width (leading zero bits dropped,
// the order unchanged): anchor contig, partner contig (bc bits each), band
// (bb), anti (ba), diag&63 (6), lcp (6, at most 40); left-aligned in 128 bits
// so that every MSD digit carries information.

struct SeedKeyFields {
    int bc, bb, ba;            // bits of the contig ids, the band and the anti-diagonal
    int shift;                 // 128 - key bits
};

// (also counts the keys' first radix digit, the top byte, into hrow: the
// partition's first level needs no counting pass of its own)
__global__ __launch_bounds__(256) void k_seed_keys(const uint8_t *__restrict__ seeds, int64_t n, int P, int C,
                                                   int64_t maxctg, int dbyte, int comp, SeedKeyFields kf,
                                                   Rec *__restrict__ keys, unsigned long long *__restrict__ hrow)
{
    __shared__ uint32_t hist[256];
    hist[threadIdx.x] = 0;
    __syncthreads();
    const int s = 1 + 2 * (P + C);
    const int64_t amx = maxctg, bmx = maxctg, maxdag = amx + bmx;
    const uint64_t cmask = (1ull << (8 * C - 1)) - 1ull;
    const uint64_t dmask = dbyte >= 8 ? ~0ull : (1ull << (8 * dbyte)) - 1ull;
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x) {
        // the seed's s <= 15 bytes as one little-endian 128-bit value (two
        // unaligned 8-byte loads; the streams carry 16 bytes of padding)
        const uint8_t *b = seeds + i * s;
        uint64_t lo, hi;
        __builtin_memcpy(&lo, b, 8);
        __builtin_memcpy(&hi, b + 8, 8);
        const u128 v = ((u128) hi << 64) | lo;
        auto field = [&](int off, int w) -> uint64_t {
            return (uint64_t) (v >> (8 * off)) & (w >= 8 ? ~0ull : (1ull << (8 * w)) - 1);
        };
        const uint64_t lcp = lo & 0xFFu;
        const int64_t ipost = (int64_t) field(1, P);
        const uint64_t icont = field(1 + P, C);
        const int64_t jpost = (int64_t) field(1 + P + C, P);
        const uint64_t jcont = field(1 + 2 * P + C, C) & cmask;
        int64_t diag, anti;
        if (comp) { diag = maxdag - (ipost + jpost); anti = amx - (ipost - jpost); }
        else      { diag = bmx + (ipost - jpost);    anti = ipost + jpost; }
        const int64_t band = diag >> 6;                               // BUCK_SHIFT
        const u128 low = (u128) lcp | ((u128) (uint64_t) (diag & 63) << 6) | ((u128) ((uint64_t) anti & dmask) << 12);
        const u128 k = ((((((u128) icont << kf.bc) | jcont) << kf.bb | ((uint64_t) band & dmask)) << (kf.ba + 12))
                        | low) << kf.shift;
        Rec x;
        x.w0 = (uint64_t) (k >> 64);
        x.w1 = (uint64_t) k;
        keys[i] = x;
        atomicAdd(&hist[(uint32_t) (x.w0 >> 56)], 1u);
    }
    __syncthreads();
    if (hist[threadIdx.x]) atomicAdd(&hrow[threadIdx.x], (unsigned long long) hist[threadIdx.x]);
}

// Sorted keys back to packed swide-byte records: the fields of each key are
// laid out again as [lcp][diag&63][anti DBYTE][band DBYTE][jcont C], LE.
constexpr int SP_T = 256, SP_PER = 16;

__global__ __launch_bounds__(SP_T) void k_seed_pack(const Rec *__restrict__ keys, int64_t n, SeedKeyFields kf,
                                                    int dbyte, int swide, uint8_t *__restrict__ out)
{
    __shared__ u128 slot[SP_T * SP_PER];
    for (int64_t b0 = (int64_t) blockIdx.x * SP_T * SP_PER; b0 < n; b0 += (int64_t) gridDim.x * SP_T * SP_PER) {
        const int cnt = (int) min((int64_t) SP_T * SP_PER, n - b0);
        for (int i = threadIdx.x; i < cnt; i += SP_T) {
            const Rec x = keys[b0 + i];
            const u128 k = (((u128) x.w0 << 64) | x.w1) >> kf.shift;
            const uint64_t lcp = (uint64_t) k & 63u;
            const uint64_t d6 = (uint64_t) (k >> 6) & 63u;
            const uint64_t anti = (uint64_t) (k >> 12) & ((1ull << kf.ba) - 1);
            const uint64_t band = (uint64_t) (k >> (12 + kf.ba)) & ((1ull << kf.bb) - 1);
            const uint64_t jcont = (uint64_t) (k >> (12 + kf.ba + kf.bb)) & ((1ull << kf.bc) - 1);
            slot[i] = (u128) lcp | ((u128) d6 << 8) | ((u128) anti << 16) | ((u128) band << (16 + 8 * dbyte))
                    | ((u128) jcont << (16 + 16 * dbyte));
        }
        __syncthreads();
        write_packed<SP_T>(out + b0 * swide, slot, cnt, swide);
        __syncthreads();
    }
}

#define PGA_MERGE_INST(A_)                                                                              \
    template __global__ void k_merge_count<A_>(const PairDev *, int, const uint8_t *, int, uint32_t *, \
                                               uint32_t *, uint32_t *, int64_t, const int64_t *);       \
    template __global__ void k_merge_emit<A_>(const PairDev *, int, const uint8_t *,                    \
                                              int, const int64_t *, const int64_t *, uint8_t *,         \
                                              uint8_t *, unsigned long long *, unsigned long long *,    \
                                              const uint32_t *, int64_t, int64_t, const int64_t *,      \
                                              const int *);
PGA_MERGE_INST(1)
PGA_MERGE_INST(2)
PGA_MERGE_INST(4)
#undef PGA_MERGE_INST

} // namespace pga
