// pga_api.hip -- the C-ABI of include/libpga_hip.h: batch bookkeeping, device
// buffer pool, kernel launch sequence, exports.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <mutex>
#include <vector>

#include "libpga_hip.h"
#include "pga_device.h"

#include "pga_kernels.hip"

using namespace pga;

namespace {

// The 4-mer scramble of the syncmer hash (src/tables.c:5-22): data of the
// selection rule's definition.
const uint8_t TMAP[256] = {
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

constexpr int XBLOCK = 256 * 4;           // positions per extraction block (XT * XU)
constexpr int NPRE = 1 << 24;
constexpr int MAX_LEVELS = 16;

struct DBuf {
    void *p = nullptr;
    size_t cap = 0;
};

struct HostPair {
    int ncontig = 0;
    int64_t seqtot = 0, maxctg = 0;
    int64_t blk0 = 0, nblk = 0;
    int ctg0 = 0;
};

int bytes_for(int64_t v)          // smallest b with 256^b >= v+1 (src/record.c:6-16)
{
    int b = 0;
    int64_t cap = 1;
    while (cap < v + 1) { cap *= 256; b++; }
    return b;
}

int bits_for(int64_t v)           // bits to hold every value in [0, v]
{
    int b = 0;
    while (b < 63 && (int64_t(1) << b) <= v) b++;
    return b;
}

} // namespace

struct pga_ctx {
    pga_opts o{};
    hipStream_t st = nullptr;
    int ncu = 256;
    char err[512] = {0};

    // loaded batch
    int npairs = 0;
    std::vector<HostPair> hp;
    int nctg = 0;
    int64_t F = 0;
    int64_t nblk = 0;
    int jbits = 0;
    bool loaded = false, built = false, merged = false;

    // host mirrors
    std::vector<PairDev> pairs;
    int64_t R = 0;
    int ntiles = 0;
    std::vector<int64_t> nseeds, cseeds;

    // device buffers
    DBuf bps, ctg, blk, bsum, pr, hist1, rec[2], bufs, entries, ls, pre, cum,
        counters, segs[2], seghist, cursor, chk0, tiles, dpairs, maxpre, mcntN, mcntC, mofsN, mofsC,
        mblk, pofs, seedN, seedC, nbuck, cbuck, sel, plist, gmax, cbmax;
    bool pre_full = true;             // pre holds every record's prefix (else the tile ends' only)
    // extraction ranges (one workgroup's run of blocks inside a pair), their
    // first-digit counts and bucket offsets, and each pair's bucket starts
    DBuf xrange, prange, rangehist, rangeoff, bstart;
    int nrange = 0;
    int xbad = 0;                      // set by the emit pass when a digit overflows its run

    DBuf vout;                         // results of the small check kernels
    // seed sort: keys (ping-pong), their pointer pair, the packed records
    DBuf ssk[2], ssbufs, ssout;
    int64_t ss_n = 0;
    int ss_swide = 0;
    bool ss_ready = false;
    DBuf minfo;                        // per-anchor interval + seed counts (count pass -> emit pass)
    // where the last merge put its outputs: inside the record buffers (dead once
    // the index is built) when they fit there, else in minfo / seedN / seedC
    uint32_t *minfo_p = nullptr;
    uint8_t *seedN_p = nullptr, *seedC_p = nullptr;
    // sorted device keys kept after the build (serving)
    bool keep_keys = false;
    DBuf skeys;

    // timing
    hipEvent_t ev[8] = {};
    std::vector<hipEvent_t> sev;       // scatter launch start/stop pairs
    int nscatter = 0;
    int64_t scatter_bytes = 0;
    pga_timings tm{};
    // event times not read yet: the build and the merge return without
    // waiting for the device; the times are read when asked for
    bool tm_build = false, tm_merge = false;

    // device-driven build + merge captured in a HIP graph (pga_run_graph):
    // valid for the loaded batch while no pool buffer is reallocated
    uint64_t pool_gen = 0, graph_gen = ~0ull;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    DBuf gstat, plist_all;
    uint8_t *g_minfo = nullptr, *g_seedN = nullptr, *g_seedC = nullptr;
    int64_t g_nmb_ub = 0, g_seg_cap = 0, g_tile_cap = 0;
    int64_t graph_fallbacks = 0;
    // 5-base bucket range of the records this context indexes (all by default)
    int shard_lo = 0, shard_hi = 1024;
};

namespace {
void graph_drop(pga_ctx *c);
}

namespace {

int fail(pga_ctx *c, int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(c->err, sizeof(c->err), fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(ctx, call)                                                                      \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail((ctx), PGA_EHIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                   \
    } while (0)

#define LAUNCHCHK(ctx)                                                                         \
    do {                                                                                       \
        hipError_t e_ = hipGetLastError();                                                     \
        if (e_ != hipSuccess)                                                                  \
            return fail((ctx), PGA_EHIP, "kernel launch failed: %s (%s:%d)",                  \
                        hipGetErrorString(e_), __FILE__, __LINE__);                            \
    } while (0)

#define TRY(x)                                                                                 \
    do {                                                                                       \
        int r_ = (x);                                                                          \
        if (r_) return r_;                                                                     \
    } while (0)

int ensure(pga_ctx *c, DBuf &b, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    if (b.cap >= bytes) return PGA_OK;
    c->pool_gen++;                                  // a captured graph's pointers are stale now
    if (b.p) {
        hipError_t e = hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
        if (e != hipSuccess) return fail(c, PGA_EHIP, "hipFree: %s", hipGetErrorString(e));
    }
    // headroom so small growth does not reallocate (bounded: buffers can be tens of GB)
    size_t want = bytes + std::min(bytes / 8, (size_t) 1 << 30);
    if (hipMalloc(&b.p, want) != hipSuccess) {
        (void) hipGetLastError();
        if (hipMalloc(&b.p, bytes) != hipSuccess) {
            (void) hipGetLastError();
            b.p = nullptr;
            return fail(c, PGA_ENOMEM, "device allocation of %zu bytes failed", bytes);
        }
        want = bytes;
    }
    b.cap = want;
    return PGA_OK;
}

template <typename T> T *P(DBuf &b) { return reinterpret_cast<T *>(b.p); }

int fill_u64(pga_ctx *c, void *p, int64_t n, unsigned long long v)
{
    if (n <= 0) return PGA_OK;
    int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned) blocks), dim3(256), 0, c->st,
                       (unsigned long long *) p, n, v);
    LAUNCHCHK(c);
    return PGA_OK;
}

// Exclusive scan of a uint32 array into int64 out[0..n], out[n] = total.
int scan_u32(pga_ctx *c, const uint32_t *in, int64_t n, int64_t *out)
{
    const int64_t span = 256 * 16;
    int64_t nb = (n + span - 1) / span;
    if (nb == 0) nb = 1;
    TRY(ensure(c, c->bsum, (size_t) (nb + 1) * sizeof(int64_t)));
    hipLaunchKernelGGL(k_scan_reduce<uint32_t>, dim3((unsigned) nb), dim3(256), 0, c->st, in, n,
                       P<int64_t>(c->bsum));
    LAUNCHCHK(c);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, c->st, P<int64_t>(c->bsum), nb);
    LAUNCHCHK(c);
    hipLaunchKernelGGL(k_scan_down<uint32_t>, dim3((unsigned) nb), dim3(256), 0, c->st, in, n,
                       (const int64_t *) P<int64_t>(c->bsum), out);
    LAUNCHCHK(c);
    return PGA_OK;
}

int cap_of(const pga_ctx *c) { return c->o.tile_records; }

// Persistent grid: as many blocks as fit on the device at once (or fewer when
// there is less work); each block loops over logical blocks.
unsigned grid_for(const pga_ctx *c, const void *fn, int threads, size_t lds, int64_t nlog)
{
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, threads, lds) != hipSuccess || occ < 1) {
        (void) hipGetLastError();
        occ = 1;
    }
    const int64_t g = std::min<int64_t>(nlog, (int64_t) c->ncu * occ);
    return (unsigned) std::max<int64_t>(g, 1);
}

int chunk_per(const pga_ctx *c) { return c->o.chunk_records / 256; }

// The partition kernels loop over the chunks counted in chk0[nseg] on a grid
// of one resident wave of blocks; blocks past the chunk count exit at once.
constexpr int64_t ANY_CHUNKS = (int64_t) 1 << 40;

int launch_seg_hist(pga_ctx *c, int per, const Rec *buf0, const Rec *buf1, const Seg *segs, int nseg,
                    const int64_t *chk0, int level, unsigned long long *hist, const uint8_t *dig,
                    const int *dn = nullptr)
{
#define PGA_HIST(PER)                                                                                  \
    hipLaunchKernelGGL(k_seg_hist<PER>, dim3(grid_for(c, (const void *) k_seg_hist<PER>, 256, 0, ANY_CHUNKS)), \
                       dim3(256), 0, c->st, buf0, buf1, segs, nseg, chk0, level, c->o.digit2_bits, hist, dig, dn)
    switch (per) {
    case 8:  PGA_HIST(8); break;
    case 32: PGA_HIST(32); break;
    default: PGA_HIST(16); break;
    }
#undef PGA_HIST
    LAUNCHCHK(c);
    return PGA_OK;
}

int launch_seg_scatter(pga_ctx *c, int per, Rec *buf0, Rec *buf1, const Seg *segs,
                       int nseg, const int64_t *chk0, int level, unsigned long long *cursor, uint8_t *dig,
                       const int *dn = nullptr)
{
    // chunks of 2048 / 4096 records: 256 threads; 8192: 1024 threads of 8
    // records (the 128 KB stage allows one workgroup a CU: sixteen waves hide
    // more of the memory latency than eight, measured 77 -> 73 ms at human scale)
#define PGA_SCAT(TT, PER)                                                                              \
    hipLaunchKernelGGL((k_seg_scatter<TT, PER>),                                                        \
                       dim3(grid_for(c, (const void *) k_seg_scatter<TT, PER>, TT, 0, ANY_CHUNKS)),      \
                       dim3(TT), 0, c->st, buf0, buf1, segs, nseg, chk0, level, c->o.digit2_bits, cursor, dig, dn)
    switch (per) {
    case 8:  PGA_SCAT(256, 8); break;
    case 32: PGA_SCAT(1024, 8); break;
    default: PGA_SCAT(256, 16); break;
    }
#undef PGA_SCAT
    LAUNCHCHK(c);
    return PGA_OK;
}

// bufs: device array {buffer 0, buffer 1} the tiles point into; entries ==
// nullptr sorts keys only (into skeys).
int launch_tile_sort(pga_ctx *c, unsigned nlog, Rec *const *bufs, uint8_t *entries, uint8_t *ls, uint32_t *pre,
                     unsigned long long *cum, Rec *skeys, int *nfall, const int *dn = nullptr)
{
    // every record's prefix only where something reads it (kept keys, no
    // prefix table); otherwise the tile ends' for the LCP fix-up
    const int pre_all = (!cum || skeys) ? 1 : 0;
    if (entries) c->pre_full = pre_all != 0;
    const int T = c->o.tile_records / 8;
    const size_t lds = (size_t) c->o.tile_records * sizeof(Rec) + (size_t) c->o.tile_records / 2 * 4 + 16;
    const Tile *tiles = P<Tile>(c->tiles);
    const PairDev *pairs = P<PairDev>(c->dpairs);
#define PGA_TSORT(TV)                                                                                   \
    hipLaunchKernelGGL(k_tile_sort<TV>, dim3(grid_for(c, (const void *) k_tile_sort<TV>, TV, lds, nlog)), \
                       dim3(TV), lds, c->st, bufs, tiles, pairs, c->jbits, c->o.digit2_bits, entries, ls,   \
                       pre, pre_all, cum, skeys, (int64_t) nlog, nfall, dn)
    switch (T) {
    case 256: PGA_TSORT(256); break;
    case 1024: PGA_TSORT(1024); break;
    default: PGA_TSORT(512); break;
    }
#undef PGA_TSORT
    LAUNCHCHK(c);
    return PGA_OK;
}

// The prefix tables of all pairs from the tile sort's group-end marks (the
// tables were cleared before it): range maxima, carries, fill; the largest
// prefix group per pair to maxpre.
int fill_prefix_tables(pga_ctx *c, int npairs, const int *plist_all)
{
    const int nrange = NPRE / CR_SPAN;
    TRY(ensure(c, c->cbmax, (size_t) npairs * nrange * sizeof(uint64_t)));
    TRY(ensure(c, c->gmax, (size_t) npairs * nrange * sizeof(uint64_t)));
    hipLaunchKernelGGL(k_cum_bmax, dim3(nrange, (unsigned) npairs), dim3(256), 0, c->st,
                       (const PairDev *) P<PairDev>(c->dpairs), (const int64_t *) P<int64_t>(c->cum),
                       P<unsigned long long>(c->cbmax));
    hipLaunchKernelGGL(k_cum_carry, dim3((unsigned) npairs), dim3(256), 0, c->st, P<unsigned long long>(c->cbmax), nrange);
    hipLaunchKernelGGL(k_cum_fill, dim3(nrange, (unsigned) npairs), dim3(256), 0, c->st,
                       (const PairDev *) P<PairDev>(c->dpairs), P<int64_t>(c->cum),
                       (const unsigned long long *) P<unsigned long long>(c->cbmax), P<unsigned long long>(c->gmax));
    hipLaunchKernelGGL(k_cum_max, dim3((unsigned) npairs), dim3(256), 0, c->st, plist_all,
                       (const unsigned long long *) P<unsigned long long>(c->gmax), nrange,
                       P<unsigned long long>(c->maxpre));
    LAUNCHCHK(c);
    return PGA_OK;
}

float ev_ms(hipEvent_t a, hipEvent_t b)
{
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        (void) hipGetLastError();
        return -1.f;
    }
    return ms;
}

int record(pga_ctx *c, int i)
{
    HIPCHK(c, hipEventRecord(c->ev[i], c->st));
    return PGA_OK;
}

int scatter_event(pga_ctx *c, hipEvent_t *out)
{
    if ((size_t) c->nscatter * 2 + 2 > c->sev.size()) {
        hipEvent_t a, b;
        HIPCHK(c, hipEventCreate(&a));
        HIPCHK(c, hipEventCreate(&b));
        c->sev.push_back(a);
        c->sev.push_back(b);
    }
    out[0] = c->sev[2 * c->nscatter];
    out[1] = c->sev[2 * c->nscatter + 1];
    return PGA_OK;
}

} // namespace

extern "C" {

void pga_default_opts(pga_opts *o)
{
    o->device = 0;
    o->freq = 10;
    o->tile_records = 4096;
    o->chunk_records = 8192;
    o->merge_tile = 1024;
    o->build_cum = 1;
    o->digit2_bits = 10;
}

const char *pga_last_error(const pga_ctx *c) { return c ? c->err : "no context"; }

static bool opts_valid(const pga_opts &q)
{
    return q.freq >= 1 && q.freq <= MAX_FREQ &&
           (q.tile_records == 2048 || q.tile_records == 4096 || q.tile_records == 8192) &&
           (q.chunk_records == 2048 || q.chunk_records == 4096 || q.chunk_records == 8192) &&
           (q.merge_tile == 256 || q.merge_tile == 512 || q.merge_tile == 1024) &&
           q.digit2_bits >= 8 && q.digit2_bits <= 11;
}

int pga_set_opts(pga_ctx *c, const pga_opts *o)
{
    if (!c || !o) return PGA_EINVAL;
    if (!opts_valid(*o) || o->device != c->o.device) return fail(c, PGA_EINVAL, "invalid options");
    c->o = *o;
    c->built = c->merged = false;
    c->graph_gen = ~0ull;                           // the captured launch shapes changed
    return PGA_OK;
}

int pga_create(const pga_opts *o, pga_ctx **out)
{
    *out = nullptr;
    pga_ctx *c = new (std::nothrow) pga_ctx();
    if (!c) return PGA_ENOMEM;
    if (o) c->o = *o; else pga_default_opts(&c->o);
    const pga_opts &q = c->o;
    if (!opts_valid(q)) {
        delete c;
        return PGA_EINVAL;
    }
    if (hipSetDevice(q.device) != hipSuccess || hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
        (void) hipGetLastError();
        delete c;
        return PGA_EHIP;
    }
    for (auto &e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) { delete c; return PGA_EHIP; }
    if (hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, q.device) != hipSuccess || c->ncu < 1) {
        (void) hipGetLastError();
        c->ncu = 256;
    }
    if (hipMemcpyToSymbol(HIP_SYMBOL(c_tmap), TMAP, sizeof(TMAP)) != hipSuccess) { delete c; return PGA_EHIP; }
    const size_t big = (size_t) 8192 * sizeof(Rec) + 4096 * 4 + 16;
    (void) hipFuncSetAttribute((const void *) k_tile_sort<1024>, hipFuncAttributeMaxDynamicSharedMemorySize, (int) big);
    (void) hipFuncSetAttribute((const void *) k_tile_sort<512>, hipFuncAttributeMaxDynamicSharedMemorySize, (int) big);
    (void) hipFuncSetAttribute((const void *) k_tile_sort<256>, hipFuncAttributeMaxDynamicSharedMemorySize, (int) big);
    (void) hipGetLastError();
    *out = c;
    return PGA_OK;
}

void pga_destroy(pga_ctx *c)
{
    if (!c) return;
    (void) hipSetDevice(c->o.device);
    if (c->st) (void) hipStreamSynchronize(c->st);
    DBuf *all[] = {&c->bps, &c->ctg, &c->blk, &c->bsum, &c->pr, &c->hist1,
                   &c->rec[0], &c->rec[1], &c->bufs, &c->entries, &c->ls, &c->pre, &c->cum,
                   &c->counters, &c->segs[0], &c->segs[1], &c->seghist, &c->cursor, &c->chk0, &c->tiles,
                   &c->dpairs, &c->maxpre, &c->mcntN, &c->mcntC, &c->mofsN, &c->mofsC, &c->mblk, &c->pofs, &c->vout, &c->minfo, &c->sel, &c->plist, &c->gmax, &c->cbmax, &c->ssk[0], &c->ssk[1], &c->ssbufs, &c->ssout,
                   &c->seedN, &c->seedC, &c->nbuck, &c->cbuck, &c->skeys,
                   &c->xrange, &c->prange, &c->rangehist, &c->rangeoff, &c->bstart};
    for (DBuf *b : all)
        if (b->p) (void) hipFree(b->p);
    graph_drop(c);
    for (DBuf *b : {&c->gstat, &c->plist_all})
        if (b->p) (void) hipFree(b->p);
    for (auto &e : c->ev) if (e) (void) hipEventDestroy(e);
    for (auto &e : c->sev) (void) hipEventDestroy(e);
    if (c->st) (void) hipStreamDestroy(c->st);
    delete c;
}

int pga_set_shard(pga_ctx *c, int bucket_lo, int bucket_hi)
{
    if (!c) return PGA_EINVAL;
    if (bucket_lo < 0 || bucket_hi > 1024 || bucket_lo > bucket_hi)
        return fail(c, PGA_EINVAL, "bucket range [%d, %d) outside [0, 1024)", bucket_lo, bucket_hi);
    c->shard_lo = bucket_lo;
    c->shard_hi = bucket_hi;
    c->built = c->merged = false;
    c->graph_gen = ~0ull;                           // the captured extraction filters differently
    return PGA_OK;
}

int pga_sync(pga_ctx *c)
{
    HIPCHK(c, hipStreamSynchronize(c->st));
    return PGA_OK;
}

} // extern "C"

// Load a batch.  contig_base (may be NULL) offsets each pair's contig ids in
// the device keys (a query's contigs follow the target's when it is served);
// jbits_min widens the offset field so keys of different batches compare.
static int load_batch_impl(pga_ctx *c, int npairs, const uint8_t *const *bps, const int64_t *bps_bytes,
                           const int64_t *const *boff, const int64_t *const *clen, const int *ncontig,
                           const int *contig_base, int jbits_min)
{
    if (!c || npairs <= 0) return c ? fail(c, PGA_EINVAL, "npairs must be > 0") : PGA_EINVAL;
    HIPCHK(c, hipSetDevice(c->o.device));
    c->loaded = c->built = c->merged = false;
    c->npairs = npairs;
    c->hp.assign(npairs, HostPair());
    int64_t total_bytes = 0;
    std::vector<int64_t> byte0(npairs);
    int nctg = 0;
    int64_t maxctg = 0, maxn = 1;
    for (int p = 0; p < npairs; p++) {
        if (ncontig[p] < 0 || bps_bytes[p] < 0) return fail(c, PGA_EINVAL, "pair %d: bad sizes", p);
        byte0[p] = total_bytes;
        total_bytes += (bps_bytes[p] + 7) & ~(int64_t) 7;
        HostPair &h = c->hp[p];
        h.ncontig = ncontig[p];
        h.ctg0 = nctg;
        nctg += ncontig[p];
        for (int i = 0; i < ncontig[p]; i++) {
            if (clen[p][i] < 0 || clen[p][i] > INT32_MAX)
                return fail(c, PGA_EOVERFLOW, "pair %d contig %d: length %lld", p, i, (long long) clen[p][i]);
            // a negative offset marks a contig without sequence: it keeps its id
            // and its length counts for the sizing, but it has no records
            // (src/extract.c:97-98)
            if (boff[p][i] >= 0 && boff[p][i] + (clen[p][i] + 3) / 4 > bps_bytes[p])
                return fail(c, PGA_EINVAL, "pair %d contig %d: sequence outside the .bps image", p, i);
            h.seqtot += clen[p][i];
            h.maxctg = std::max<int64_t>(h.maxctg, clen[p][i]);
        }
        maxctg = std::max(maxctg, h.maxctg);
        maxn = std::max<int64_t>(maxn, (int64_t) ncontig[p] + (contig_base ? contig_base[p] : 0));
        int Pb = bytes_for(h.maxctg), Cb = bytes_for(2 * (int64_t) h.ncontig - 1);
        if (Pb + Cb > 7) return fail(c, PGA_EOVERFLOW, "pair %d: payload of %d bytes exceeds 7", p, Pb + Cb);
    }
    c->jbits = std::max(bits_for(maxctg), jbits_min);
    int cbits = bits_for(maxn - 1);
    if (c->jbits + cbits + 1 > 48)
        return fail(c, PGA_EOVERFLOW, "offset (%d bits) + contig (%d bits) + strand exceed 48 bits", c->jbits, cbits);

    // contig table over the flat candidate-position space (pairs block-aligned)
    std::vector<ContigDev> tab(std::max(nctg, 1));
    int64_t pos = 0;
    for (int p = 0; p < npairs; p++) {
        HostPair &h = c->hp[p];
        pos = (pos + XBLOCK - 1) / XBLOCK * XBLOCK;
        h.blk0 = pos / XBLOCK;
        int64_t start = pos;
        for (int i = 0; i < h.ncontig; i++) {
            ContigDev &d = tab[h.ctg0 + i];
            const bool seq = boff[p][i] >= 0;
            d.bitoff = seq ? 8 * (byte0[p] + boff[p][i]) : 8 * byte0[p];
            d.pos0 = pos;
            d.ncand = seq && clen[p][i] >= 12 ? clen[p][i] - 11 : 0;
            d.clen = (int32_t) clen[p][i];
            d.local = i + (contig_base ? contig_base[p] : 0);
            d.pair = p;
            d.pad = 0;
            pos += d.ncand;
        }
        h.nblk = (pos - start + XBLOCK - 1) / XBLOCK;
    }
    c->nctg = nctg;
    c->F = pos;
    c->nblk = (pos + XBLOCK - 1) / XBLOCK;

    TRY(ensure(c, c->bps, (size_t) total_bytes + 64));
    HIPCHK(c, hipMemsetAsync(c->bps.p, 0, (size_t) total_bytes + 64, c->st));
    for (int p = 0; p < npairs; p++)
        if (bps_bytes[p] > 0)
            HIPCHK(c, hipMemcpyAsync(P<uint8_t>(c->bps) + byte0[p], bps[p], (size_t) bps_bytes[p],
                                     hipMemcpyHostToDevice, c->st));
    TRY(ensure(c, c->ctg, tab.size() * sizeof(ContigDev)));
    HIPCHK(c, hipMemcpyAsync(c->ctg.p, tab.data(), tab.size() * sizeof(ContigDev), hipMemcpyHostToDevice, c->st));
    // contig of every extraction block's first position: the last contig whose
    // first candidate position is <= it (what a binary search over pos0 gives),
    // and where the block lies in it
    std::vector<BlkDev> blk((size_t) std::max<int64_t>(c->nblk, 1), BlkDev{});
    for (int64_t b = 0, i = 0; b < c->nblk; b++) {
        const int64_t fb = b * XBLOCK;
        while (i + 1 < nctg && tab[i + 1].pos0 <= fb) i++;
        BlkDev &d = blk[b];
        d.c = (int32_t) i;
        if (nctg == 0 || fb < tab[i].pos0) continue;
        const ContigDev &cc = tab[i];
        const int64_t lo = std::max<int64_t>(0, cc.pos0 + 28 - fb);
        const int64_t hi = std::min<int64_t>(XBLOCK, cc.pos0 + cc.ncand - fb);
        if (hi <= lo) continue;
        d.bbit = cc.bitoff + 2 * (fb - cc.pos0);
        d.jblk = (uint32_t) (fb - cc.pos0);
        d.local = (uint32_t) cc.local;
        d.frange = (uint32_t) lo | (uint32_t) hi << 16;
    }
    TRY(ensure(c, c->blk, blk.size() * sizeof(BlkDev)));
    HIPCHK(c, hipMemcpyAsync(c->blk.p, blk.data(), blk.size() * sizeof(BlkDev), hipMemcpyHostToDevice, c->st));
    // extraction ranges: each pair's blocks cut into runs of about nblk / (one
    // resident wave of count workgroups), a multiple of the emit's blocks per
    // round; every range lies inside one pair
    const int64_t xg = grid_for(c, (const void *) k_extract_count, 256, 0, std::max<int64_t>(c->nblk, 1));
    int64_t per = std::max<int64_t>(1, (c->nblk + xg - 1) / xg);
    per = (per + EMIT_QB - 1) / EMIT_QB * EMIT_QB;
    std::vector<XRange> xr;
    std::vector<int32_t> prange(npairs + 1);
    for (int p = 0; p < npairs; p++) {
        prange[p] = (int32_t) xr.size();
        const int64_t b0 = c->hp[p].blk0, b1 = b0 + c->hp[p].nblk;
        for (int64_t b = b0; b < b1; b += per) xr.push_back(XRange{b, std::min(b + per, b1), p, 0});
    }
    prange[npairs] = (int32_t) xr.size();
    c->nrange = (int) xr.size();
    TRY(ensure(c, c->xrange, std::max<size_t>(xr.size(), 1) * sizeof(XRange)));
    if (!xr.empty())
        HIPCHK(c, hipMemcpyAsync(c->xrange.p, xr.data(), xr.size() * sizeof(XRange), hipMemcpyHostToDevice, c->st));
    TRY(ensure(c, c->prange, prange.size() * sizeof(int32_t)));
    HIPCHK(c, hipMemcpyAsync(c->prange.p, prange.data(), prange.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->st));
    TRY(ensure(c, c->rangehist, std::max<size_t>(xr.size(), 1) * 256 * sizeof(uint32_t)));
    TRY(ensure(c, c->rangeoff, std::max<size_t>(xr.size(), 1) * 256 * sizeof(int64_t)));
    TRY(ensure(c, c->bstart, (size_t) npairs * 256 * sizeof(int64_t)));
    HIPCHK(c, hipStreamSynchronize(c->st));     // host arrays above are stack/vector owned
    c->loaded = true;
    c->graph_gen = ~0ull;                           // a new batch needs a new graph
    return PGA_OK;
}

namespace {

// MSD radix partition: the segments seg0 (records in buf0) are split on
// digits (8 bits; opts.digit2_bits at the second level), level after level,
// until every bucket fits a sort tile; tiles go to c->tiles after the tile0
// ones (c->ntiles is set).  Level-1 digit counts come from hist1 (one 256-row
// per pair) when given, else they are counted here.  level1_in_place: the
// segments' records are already in level-1 bucket order (hist1's layout), so
// the first level only plans and does not scatter.
// The tiles' `buf` index refers to {buf0, buf1}.  dig (may be null): a byte
// per record position, free during the partition; each scatter from the
// second level on writes the moved records' next digits there, and the next
// level counts those segments from it.  dig1: the level-1 records' second
// digits are there already (level1_in_place only).
int partition(pga_ctx *c, Rec *buf0, Rec *buf1, const std::vector<Seg> &seg0, const std::vector<Tile> &tile0,
              const unsigned long long *hist1, bool level1_in_place, int64_t seg_cap, int64_t tile_cap,
              uint8_t *dig, bool dig1, int *levels_out)
{
    const int CAP = cap_of(c);
    const int per = chunk_per(c);
    const int CH = per * 256;
    Rec *buf[2] = {buf0, buf1};
    // [0] next segs, [1] tiles, [2..3] records left in place by the level (u64)
    int *dcount = P<int>(c->counters);
    int hcount[4] = {0, (int) tile0.size(), 0, 0};
    HIPCHK(c, hipMemcpyAsync(dcount, hcount, sizeof(hcount), hipMemcpyHostToDevice, c->st));
    if (!tile0.empty())
        HIPCHK(c, hipMemcpyAsync(c->tiles.p, tile0.data(), tile0.size() * sizeof(Tile), hipMemcpyHostToDevice, c->st));
    int nseg = (int) seg0.size();
    if (nseg)
        HIPCHK(c, hipMemcpyAsync(c->segs[0].p, seg0.data(), seg0.size() * sizeof(Seg), hipMemcpyHostToDevice, c->st));
    int cur = 0, level = 1;
    // the previous level's scatter wrote digit bytes, and the records it left in place
    bool prev_dig = false;
    int64_t prev_kept = 0;
    while (nseg > 0) {
        if (level > MAX_LEVELS) return fail(c, PGA_EINVAL, "radix partition did not converge");
        hipLaunchKernelGGL(k_chunk_prefix, dim3(1), dim3(1024), 0, c->st, (const Seg *) P<Seg>(c->segs[cur]),
                           nseg, CH, P<int64_t>(c->chk0), (const int *) nullptr);
        LAUNCHCHK(c);
        const bool counted = level == 1 && hist1;
        const bool in_place = level == 1 && level1_in_place;
        // the next level's digits are in dig: written by the level's scatter
        // (8-bit digits from the third level on), or by the extraction for
        // level 2 (dig1)
        const bool dig_next = dig && (in_place ? dig1 : level >= 2);
        const int64_t nd = level == 2 ? (int64_t) 1 << c->o.digit2_bits : 256;   // digit rows of this level
        TRY(ensure(c, c->seghist, (size_t) nseg * nd * sizeof(uint64_t)));
        TRY(ensure(c, c->cursor, (size_t) nseg * nd * sizeof(uint64_t)));
        if (!counted) {
            TRY(fill_u64(c, c->seghist.p, (int64_t) nseg * nd, 0));
            TRY(launch_seg_hist(c, per, buf[0], buf[1], P<Seg>(c->segs[cur]), nseg, P<int64_t>(c->chk0), level,
                                P<unsigned long long>(c->seghist), dig));
        }
        HIPCHK(c, hipMemsetAsync(dcount, 0, sizeof(int), c->st));
        HIPCHK(c, hipMemsetAsync(dcount + 2, 0, 2 * sizeof(int), c->st));
        hipLaunchKernelGGL(k_seg_plan, dim3(nseg), dim3(256), 0, c->st, P<Seg>(c->segs[cur]),
                           counted ? hist1 : (const unsigned long long *) P<unsigned long long>(c->seghist),
                           counted ? 1 : 0, level, c->o.digit2_bits, in_place ? 1 : 0, dig_next ? 1 : 0, CAP,
                           P<unsigned long long>(c->cursor),
                           P<Seg>(c->segs[cur ^ 1]), dcount, (int) seg_cap, P<Tile>(c->tiles), dcount + 1,
                           (int) tile_cap, reinterpret_cast<unsigned long long *>(dcount + 2), nseg,
                           (const int *) nullptr);
        LAUNCHCHK(c);
        if (!in_place) {
            hipEvent_t se[2];
            TRY(scatter_event(c, se));
            HIPCHK(c, hipEventRecord(se[0], c->st));
            TRY(launch_seg_scatter(c, per, buf[0], buf[1], P<Seg>(c->segs[cur]), nseg, P<int64_t>(c->chk0), level,
                                   P<unsigned long long>(c->cursor), dig_next ? dig : nullptr));
            HIPCHK(c, hipEventRecord(se[1], c->st));
            c->nscatter++;
        }
        // chunk and record totals of the level: chk0[nseg], chk0[nseg + 1]
        int64_t ck[2] = {0, 0};
        HIPCHK(c, hipMemcpyAsync(ck, P<int64_t>(c->chk0) + nseg, sizeof(ck), hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipMemcpyAsync(hcount, dcount, 4 * sizeof(int), hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        uint64_t kept;
        memcpy(&kept, &hcount[2], sizeof(kept));
        if (!in_place)
            c->scatter_bytes += 2 * (int64_t) sizeof(Rec) * (ck[1] - (int64_t) kept);   // moved records, read + write
        // the previous scatter's digit bytes: one for each record of this
        // level's segments that it moved (the others it left in place)
        if (prev_dig) c->scatter_bytes += ck[1] - prev_kept;
        prev_dig = dig_next && !in_place;
        prev_kept = (int64_t) kept;
        nseg = hcount[0];
        cur ^= 1;
        level++;
        if (nseg > seg_cap) return fail(c, PGA_EOVERFLOW, "segment list overflow");
        if (hcount[1] > tile_cap) return fail(c, PGA_EOVERFLOW, "tile list overflow");
    }
    // hcount[1] is current: the last level read the counters back (or, with no
    // segments, it still holds the tile0 count written above)
    c->ntiles = hcount[1];
    if (c->ntiles > tile_cap) return fail(c, PGA_EOVERFLOW, "tile list overflow");
    *levels_out = level - 1;
    return PGA_OK;
}

} // namespace

extern "C" {

int pga_load_batch(pga_ctx *c, int npairs, const uint8_t *const *bps, const int64_t *bps_bytes,
                   const int64_t *const *boff, const int64_t *const *clen, const int *ncontig)
{
    return load_batch_impl(c, npairs, bps, bps_bytes, boff, clen, ncontig, nullptr, 0);
}

int pga_load_gdb(pga_ctx *c, const uint8_t *bps, int64_t bps_bytes, const int64_t *boff,
                 const int64_t *clen, int ncontig)
{
    return pga_load_batch(c, 1, &bps, &bps_bytes, &boff, &clen, &ncontig);
}

int pga_build_index(pga_ctx *c)
{
    if (!c || !c->loaded) return c ? fail(c, PGA_EINVAL, "no GDB loaded") : PGA_EINVAL;
    HIPCHK(c, hipSetDevice(c->o.device));
    c->built = c->merged = c->ss_ready = false;
    c->nscatter = 0;
    c->scatter_bytes = 0;
    memset(&c->tm, 0, sizeof(c->tm));
    c->tm_build = c->tm_merge = false;
    const int npairs = c->npairs;
    const int CAP = cap_of(c);

    TRY(record(c, 0));
    // ---- extraction: count; per range and first digit, the range's offset in
    // the digit's bucket; per pair, the digit totals and the record range
    const int64_t nblk = c->nblk;
    TRY(ensure(c, c->pr, (size_t) (npairs + 1) * sizeof(int64_t)));
    TRY(ensure(c, c->sel, (size_t) std::max<int64_t>(nblk, 1) * 256));    // 2 bits per position
    TRY(ensure(c, c->hist1, (size_t) npairs * 256 * sizeof(uint64_t)));
    if (nblk > 0) {
        hipLaunchKernelGGL(k_extract_count, dim3(grid_for(c, (const void *) k_extract_count, 256, 0, c->nrange)),
                           dim3(256), 0, c->st, P<uint64_t>(c->bps), P<ContigDev>(c->ctg),
                           (const BlkDev *) P<BlkDev>(c->blk), c->nctg, c->F, (const XRange *) P<XRange>(c->xrange),
                           c->nrange, P<uint8_t>(c->sel), P<uint32_t>(c->rangehist),
                           (uint32_t) c->shard_lo, (uint32_t) c->shard_hi);
        LAUNCHCHK(c);
    }
    if (c->nrange > 0) {
        hipLaunchKernelGGL(k_xrange_scan, dim3(npairs, 16), dim3(256), 0, c->st, (const int *) P<int>(c->prange),
                           (const uint32_t *) P<uint32_t>(c->rangehist), P<int64_t>(c->rangeoff),
                           P<unsigned long long>(c->hist1));
        LAUNCHCHK(c);
    } else {
        TRY(fill_u64(c, c->hist1.p, (int64_t) npairs * 256, 0));
    }
    hipLaunchKernelGGL(k_pair_totals, dim3(1), dim3(256), 0, c->st,
                       (const unsigned long long *) P<unsigned long long>(c->hist1), npairs, P<int64_t>(c->pr));
    LAUNCHCHK(c);
    std::vector<int64_t> pr(npairs + 1);
    HIPCHK(c, hipMemcpyAsync(pr.data(), c->pr.p, pr.size() * sizeof(int64_t), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    const int64_t R = pr[npairs];
    c->R = R;

    // ---- per-pair layout
    c->pairs.assign(npairs, PairDev());
    int64_t ent = 0;
    for (int p = 0; p < npairs; p++) {
        PairDev &d = c->pairs[p];
        const HostPair &h = c->hp[p];
        d.rec0 = pr[p];
        d.nrec = pr[p + 1] - pr[p];
        d.P = bytes_for(h.maxctg);
        d.C = bytes_for(2 * (int64_t) h.ncontig - 1);
        d.e = 9 + d.P + d.C;
        d.s = 1 + 2 * (d.P + d.C);
        d.ncontig = h.ncontig;
        d.ent_off = ent;
        ent += d.nrec * d.e;
        d.cum_off = (int64_t) p * NPRE;
        d.buck_off = h.ctg0;
    }
    const size_t nrec_alloc = (size_t) std::max<int64_t>(R, 1);
    TRY(ensure(c, c->rec[0], nrec_alloc * sizeof(Rec)));
    TRY(ensure(c, c->rec[1], nrec_alloc * sizeof(Rec)));
    TRY(ensure(c, c->bufs, 2 * sizeof(Rec *)));
    {
        Rec *bp[2] = {P<Rec>(c->rec[0]), P<Rec>(c->rec[1])};
        HIPCHK(c, hipMemcpyAsync(c->bufs.p, bp, sizeof(bp), hipMemcpyHostToDevice, c->st));
    }
    TRY(ensure(c, c->entries, (size_t) ent + 32));
    TRY(ensure(c, c->ls, nrec_alloc + 64));
    TRY(ensure(c, c->pre, (nrec_alloc + 1) * sizeof(uint32_t)));
    if (c->keep_keys) TRY(ensure(c, c->skeys, nrec_alloc * sizeof(Rec)));
    TRY(ensure(c, c->hist1, (size_t) npairs * 256 * sizeof(uint64_t)));
    TRY(ensure(c, c->dpairs, (size_t) npairs * sizeof(PairDev)));
    TRY(ensure(c, c->counters, 16 * sizeof(int)));
    // List bounds (a count past them is an internal error, reported as
    // PGA_EOVERFLOW).  Segments: a segment holds more than CAP records and the
    // segments of one level are disjoint, so a level has at most R/CAP of them
    // (plus the pairs at level 1).  Tiles: the plan closes a tile only when the
    // next bucket would overflow it, so two consecutive tiles of a segment hold
    // more than CAP records together -- at most 2R/CAP such tiles -- plus, over
    // at most 16 levels, one partial tile per segment and one tile break per
    // bucket that becomes a segment (each at most R/CAP a level): 34 R/CAP.
    const int64_t seg_cap = R / CAP + npairs + 16;
    const int64_t tile_cap = 34 * (R / CAP) + npairs + 1024;
    TRY(ensure(c, c->segs[0], (size_t) seg_cap * sizeof(Seg)));
    TRY(ensure(c, c->segs[1], (size_t) seg_cap * sizeof(Seg)));
    TRY(ensure(c, c->seghist, (size_t) seg_cap * 256 * sizeof(uint64_t)));
    TRY(ensure(c, c->cursor, (size_t) seg_cap * 256 * sizeof(uint64_t)));
    TRY(ensure(c, c->chk0, (size_t) (seg_cap + 2) * sizeof(int64_t)));
    TRY(ensure(c, c->tiles, (size_t) tile_cap * sizeof(Tile)));
    HIPCHK(c, hipMemcpyAsync(c->dpairs.p, c->pairs.data(), (size_t) npairs * sizeof(PairDev),
                             hipMemcpyHostToDevice, c->st));

    // ---- extraction: level-1 digit totals and bucket offsets per range, then
    // every record straight into its level-1 bucket in record buffer 1
    int *dbad = P<int>(c->counters) + 8;
    HIPCHK(c, hipMemsetAsync(dbad, 0, 4 * sizeof(int), c->st));     // [8] digit mismatch, [10..11] tile sort paths
    hipLaunchKernelGGL(k_xrange_base, dim3(npairs), dim3(256), 0, c->st, (const int64_t *) P<int64_t>(c->pr),
                       (const unsigned long long *) P<unsigned long long>(c->hist1), P<int64_t>(c->bstart));
    LAUNCHCHK(c);
    // with an 8-bit second digit, the extraction also writes each record's
    // second digit to the LCP/strand buffer (free until the tile sort), and
    // the second level counts those bytes instead of the records
    const bool dig2 = c->o.digit2_bits == 8;
    if (c->nrange > 0 && R > 0) {
        hipLaunchKernelGGL(k_extract_emit<EMIT_QB>,
                           dim3(grid_for(c, (const void *) k_extract_emit<EMIT_QB>, 256 * EMIT_QB, 0, c->nrange)),
                           dim3(256 * EMIT_QB), 0, c->st, P<uint64_t>(c->bps), P<ContigDev>(c->ctg),
                           (const BlkDev *) P<BlkDev>(c->blk), c->nctg, c->jbits, (const uint8_t *) P<uint8_t>(c->sel),
                           (const XRange *) P<XRange>(c->xrange), c->nrange,
                           (const uint32_t *) P<uint32_t>(c->rangehist), (const int64_t *) P<int64_t>(c->rangeoff),
                           (const int64_t *) P<int64_t>(c->bstart), P<Rec>(c->rec[1]),
                           dig2 ? P<uint8_t>(c->ls) : nullptr, dbad);
        LAUNCHCHK(c);
    }
    c->xbad = 0;
    HIPCHK(c, hipMemcpyAsync(&c->xbad, dbad, sizeof(int), hipMemcpyDeviceToHost, c->st));   // read after the partition's sync
    TRY(record(c, 1));

    // ---- MSD radix partition, from level-1 buckets already in record buffer 1
    std::vector<Seg> seg0;
    std::vector<Tile> tile0;
    for (int p = 0; p < npairs; p++) {
        const PairDev &d = c->pairs[p];
        if (d.nrec > CAP) seg0.push_back(Seg{d.rec0, d.nrec, p, 1, 1, 0});
        else if (d.nrec > 0) tile0.push_back(Tile{d.rec0, (int32_t) d.nrec, p, 1, 1, drange_all(8)});
    }
    int levels = 0;
    // the LCP/strand bytes are written by the tile sort: until then their
    // buffer holds the partition's digit bytes
    TRY(partition(c, P<Rec>(c->rec[0]), P<Rec>(c->rec[1]), seg0, tile0, P<unsigned long long>(c->hist1), true,
                  seg_cap, tile_cap, P<uint8_t>(c->ls), dig2, &levels));
    if (c->xbad) return fail(c, PGA_EINVAL, "internal: extraction passes disagree on a record's first digit");
    c->tm.partition_levels = levels;
    c->tm.partition_launches = c->nscatter;
    TRY(record(c, 2));

    // ---- per-tile LDS sort, LCP, GIX entries; group ends marked in the
    // cleared prefix tables
    if (c->o.build_cum) {
        TRY(ensure(c, c->cum, (size_t) npairs * NPRE * sizeof(int64_t)));
        TRY(fill_u64(c, c->cum.p, (int64_t) npairs * NPRE, 0));
    }
    unsigned long long *cum_marks = c->o.build_cum ? P<unsigned long long>(c->cum) : nullptr;
    if (c->ntiles > 0)
        TRY(launch_tile_sort(c, (unsigned) c->ntiles, (Rec *const *) c->bufs.p, P<uint8_t>(c->entries),
                             P<uint8_t>(c->ls), P<uint32_t>(c->pre), cum_marks,
                             c->keep_keys ? P<Rec>(c->skeys) : nullptr, P<int>(c->counters) + 10));
    TRY(record(c, 3));
    if (c->ntiles > 0) {
        hipLaunchKernelGGL(k_tile_fix, dim3(grid_for(c, (const void *) k_tile_fix, 256, 0, (c->ntiles + 255) / 256)),
                           dim3(256), 0, c->st,
                           (const Tile *) P<Tile>(c->tiles), c->ntiles, (const PairDev *) P<PairDev>(c->dpairs),
                           P<uint8_t>(c->entries), P<uint8_t>(c->ls), (const uint32_t *) P<uint32_t>(c->pre),
                           (const int *) nullptr);
        LAUNCHCHK(c);
    }

    // ---- 2^24 prefix table and largest prefix group
    TRY(ensure(c, c->maxpre, (size_t) npairs * sizeof(uint64_t)));
    TRY(fill_u64(c, c->maxpre.p, npairs, 0));
    if (c->o.build_cum) {
        std::vector<int> all(npairs);
        for (int p = 0; p < npairs; p++) all[p] = p;
        TRY(ensure(c, c->plist, (size_t) npairs * sizeof(int)));
        HIPCHK(c, hipMemcpyAsync(c->plist.p, all.data(), all.size() * sizeof(int), hipMemcpyHostToDevice, c->st));
        TRY(fill_prefix_tables(c, npairs, (const int *) P<int>(c->plist)));
        HIPCHK(c, hipStreamSynchronize(c->st));       // the pair list is a local
    }
    TRY(record(c, 4));
    // no wait here: the merge can queue behind the build; the stage times are
    // read from the events when asked for (pga_timings_get)
    c->tm.tiles = c->ntiles;
    c->tm.scatter_launches = c->nscatter;
    c->tm.scatter_bytes = c->scatter_bytes;
    c->tm_build = true;
    c->tm_merge = false;
    c->built = true;
    return PGA_OK;
}

int pga_index_info(pga_ctx *c, int pair, pga_index_meta *info)
{
    if (!c || !c->built || pair < 0 || pair >= c->npairs) return c ? fail(c, PGA_EINVAL, "no index for pair %d", pair) : PGA_EINVAL;
    const PairDev &d = c->pairs[pair];
    uint64_t mp = 0;
    HIPCHK(c, hipMemcpyAsync(&mp, P<uint64_t>(c->maxpre) + pair, sizeof(mp), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    info->nrec = d.nrec;
    info->P = d.P;
    info->C = d.C;
    info->e = d.e;
    info->s = d.s;
    info->maxpre = c->o.build_cum ? (int64_t) mp : -1;
    info->ncontig = d.ncontig;
    info->seqtot = c->hp[pair].seqtot;
    return PGA_OK;
}

int pga_index_export(pga_ctx *c, int pair, uint8_t *entries, int64_t *cum)
{
    if (!c || !c->built || pair < 0 || pair >= c->npairs) return c ? fail(c, PGA_EINVAL, "no index for pair %d", pair) : PGA_EINVAL;
    const PairDev &d = c->pairs[pair];
    if (entries && d.nrec > 0)
        HIPCHK(c, hipMemcpyAsync(entries, P<uint8_t>(c->entries) + d.ent_off, (size_t) (d.nrec * d.e),
                                 hipMemcpyDeviceToHost, c->st));
    if (cum) {
        if (!c->o.build_cum) return fail(c, PGA_EINVAL, "prefix table not built (opts.build_cum = 0)");
        HIPCHK(c, hipMemcpyAsync(cum, P<int64_t>(c->cum) + d.cum_off, (size_t) NPRE * sizeof(int64_t),
                                 hipMemcpyDeviceToHost, c->st));
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    return PGA_OK;
}

int pga_index_verify(pga_ctx *c, int pair, int64_t *bad, int64_t *first_bad, int64_t *lcp_bad)
{
    if (!c || !c->built || pair < 0 || pair >= c->npairs) return c ? fail(c, PGA_EINVAL, "no index for pair %d", pair) : PGA_EINVAL;
    HIPCHK(c, hipSetDevice(c->o.device));
    const PairDev &d = c->pairs[pair];
    TRY(ensure(c, c->vout, 4 * sizeof(uint64_t)));
    uint64_t init[3] = {0, ~0ull, 0}, res[3] = {0, 0, 0};
    HIPCHK(c, hipMemcpyAsync(c->vout.p, init, sizeof(init), hipMemcpyHostToDevice, c->st));
    if (d.nrec > 0) {
        hipLaunchKernelGGL(k_verify, dim3(grid_for(c, (const void *) k_verify, 256, 0, (d.nrec + 255) / 256)), dim3(256),
                           0, c->st, (const uint8_t *) P<uint8_t>(c->entries) + d.ent_off, d.e,
                           (const uint32_t *) P<uint32_t>(c->pre),
                           c->pre_full ? (const int64_t *) nullptr : (const int64_t *) P<int64_t>(c->cum) + d.cum_off,
                           d.rec0, d.nrec, P<unsigned long long>(c->vout));
        LAUNCHCHK(c);
    }
    HIPCHK(c, hipMemcpyAsync(res, c->vout.p, sizeof(res), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (bad) *bad = (int64_t) res[0];
    if (first_bad) *first_bad = res[1] == ~0ull ? -1 : (int64_t) res[1];
    if (lcp_bad) *lcp_bad = (int64_t) res[2];
    return PGA_OK;
}

int pga_index_poke(pga_ctx *c, int pair, int64_t entry, int byte, int value)
{
    if (!c || !c->built || pair < 0 || pair >= c->npairs) return c ? fail(c, PGA_EINVAL, "no index for pair %d", pair) : PGA_EINVAL;
    const PairDev &d = c->pairs[pair];
    if (entry < 0 || entry >= d.nrec || byte < 0 || byte >= d.e || value < 0 || value > 255)
        return fail(c, PGA_EINVAL, "entry %lld byte %d outside the index", (long long) entry, byte);
    HIPCHK(c, hipSetDevice(c->o.device));
    const uint8_t v = (uint8_t) value;
    HIPCHK(c, hipMemcpyAsync(P<uint8_t>(c->entries) + d.ent_off + entry * d.e + byte, &v, 1, hipMemcpyHostToDevice,
                             c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return PGA_OK;
}

int pga_bucket_counts(pga_ctx *c, int pair, int64_t *counts)
{
    if (!c || !c->built || pair < 0 || pair >= c->npairs || !counts)
        return c ? fail(c, PGA_EINVAL, "no index for pair %d", pair) : PGA_EINVAL;
    if (!c->o.build_cum) return fail(c, PGA_EINVAL, "prefix table not built (opts.build_cum = 0)");
    HIPCHK(c, hipSetDevice(c->o.device));
    TRY(ensure(c, c->vout, 1024 * sizeof(int64_t)));
    hipLaunchKernelGGL(k_bucket_counts, dim3(4), dim3(256), 0, c->st,
                       (const int64_t *) P<int64_t>(c->cum) + c->pairs[pair].cum_off, P<long long>(c->vout));
    LAUNCHCHK(c);
    HIPCHK(c, hipMemcpyAsync(counts, c->vout.p, 1024 * sizeof(int64_t), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return PGA_OK;
}

// The shard splitter of a multi-GPU build: FastGA's distribute() rule as the
// reference states it (src/dsort.c:120-167) -- walk the cumulative bucket
// counts; at each multiple t of total/world_size the shard boundary goes
// before or after the crossing bucket, whichever is nearer to t.
int pga_destination_map(const int64_t *global_counts, int world_size, int *ksplit, int *select)
{
    if (!global_counts || !ksplit || !select || world_size < 1) return PGA_EINVAL;
    const int NB = 1024;
    int64_t total = 0;
    for (int b = 0; b < NB; b++) total += global_counts[b];
    ksplit[0] = 0;
    int shard = 0;                      // shards closed so far
    int64_t target = total / world_size, run = 0;
    for (int b = 0; b < NB; b++) {
        const int64_t before = run;
        run += global_counts[b];
        if (run >= target) {
            const bool next = run - target > target - before;   // nearer to close before b
            select[b] = next ? shard + 1 : shard;
            ksplit[shard + 1] = next ? b : b + 1;
            shard = shard + 1 < world_size - 1 ? shard + 1 : world_size - 1;
            target = (int64_t) (shard + 1) * total / world_size;
        } else {
            select[b] = shard;
        }
    }
    ksplit[world_size] = NB;
    for (int r = shard + 1; r < world_size; r++) ksplit[r] = NB;
    return PGA_OK;
}

int pga_merge(pga_ctx *c)
{
    if (!c || !c->built) return c ? fail(c, PGA_EINVAL, "pga_merge before pga_build_index") : PGA_EINVAL;
    HIPCHK(c, hipSetDevice(c->o.device));
    const int npairs = c->npairs;
    const int MT = c->o.merge_tile;
    std::vector<int64_t> mblk(npairs + 1);
    int64_t nmb = 0;
    for (int p = 0; p < npairs; p++) {
        c->pairs[p].mblk0 = nmb;
        mblk[p] = nmb;
        nmb += (c->pairs[p].nrec + MT - 1) / MT;
    }
    mblk[npairs] = nmb;
    TRY(record(c, 5));
    TRY(ensure(c, c->mblk, mblk.size() * sizeof(int64_t)));
    TRY(ensure(c, c->mcntN, (size_t) std::max<int64_t>(nmb, 1) * sizeof(uint32_t)));
    TRY(ensure(c, c->mcntC, (size_t) std::max<int64_t>(nmb, 1) * sizeof(uint32_t)));
    TRY(ensure(c, c->mofsN, (size_t) (nmb + 1) * sizeof(int64_t)));
    TRY(ensure(c, c->mofsC, (size_t) (nmb + 1) * sizeof(int64_t)));
    TRY(ensure(c, c->pofs, (size_t) 2 * (npairs + 1) * sizeof(int64_t)));
    HIPCHK(c, hipMemcpyAsync(c->mblk.p, mblk.data(), mblk.size() * sizeof(int64_t), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->dpairs.p, c->pairs.data(), (size_t) npairs * sizeof(PairDev), hipMemcpyHostToDevice, c->st));
    std::vector<int64_t> po(2 * (npairs + 1), 0);
    // Merge outputs go into the two record buffers when they fit: after the
    // build the records live on only as GIX entries (and, for serving, as the
    // kept keys), so their memory is free.  That keeps the human-scale pair
    // (2 x 77 GB of records) within one GPU.
    struct Arena {
        uint8_t *p[2];
        size_t cap[2], used[2];
        void *take(size_t bytes) {
            bytes = (bytes + 255) & ~(size_t) 255;
            for (int k = 0; k < 2; k++)
                if (p[k] && cap[k] - used[k] >= bytes) { void *r = p[k] + used[k]; used[k] += bytes; return r; }
            return nullptr;
        }
    } arena{{P<uint8_t>(c->rec[0]), P<uint8_t>(c->rec[1])}, {c->rec[0].cap, c->rec[1].cap}, {0, 0}};
    // the count pass hands each anchor's interval and seed counts to the emit
    // pass in a 32-bit record per anchor
    uint32_t *minfo = nullptr;
    if (c->R > 0) {
        minfo = static_cast<uint32_t *>(arena.take((size_t) c->R * sizeof(uint32_t)));
        if (!minfo) {
            TRY(ensure(c, c->minfo, (size_t) c->R * sizeof(uint32_t)));
            minfo = P<uint32_t>(c->minfo);
        }
    }
    c->minfo_p = minfo;
    if (nmb > 0) {
#define PGA_COUNT(A_)                                                                                   \
    hipLaunchKernelGGL(k_merge_count<A_>, dim3(grid_for(c, (const void *) k_merge_count<A_>, MTH, 0, nmb)), \
                       dim3(MTH), 0, c->st,                                                             \
                       (const PairDev *) P<PairDev>(c->dpairs), npairs, (const uint8_t *) P<uint8_t>(c->ls), \
                       c->o.freq, P<uint32_t>(c->mcntN), P<uint32_t>(c->mcntC), minfo, nmb,            \
                       (const int64_t *) nullptr)
        switch (MT) {
        case 512: PGA_COUNT(2); break;
        case 1024: PGA_COUNT(4); break;
        default: PGA_COUNT(1); break;
        }
#undef PGA_COUNT
        LAUNCHCHK(c);
        TRY(scan_u32(c, P<uint32_t>(c->mcntN), nmb, P<int64_t>(c->mofsN)));
        TRY(scan_u32(c, P<uint32_t>(c->mcntC), nmb, P<int64_t>(c->mofsC)));
        hipLaunchKernelGGL(k_pair_ranges, dim3((npairs + 256) / 256), dim3(256), 0, c->st,
                           (const int64_t *) P<int64_t>(c->mofsN), (const int64_t *) P<int64_t>(c->mblk), npairs,
                           P<int64_t>(c->pofs));
        hipLaunchKernelGGL(k_pair_ranges, dim3((npairs + 256) / 256), dim3(256), 0, c->st,
                           (const int64_t *) P<int64_t>(c->mofsC), (const int64_t *) P<int64_t>(c->mblk), npairs,
                           P<int64_t>(c->pofs) + (npairs + 1));
        LAUNCHCHK(c);
        HIPCHK(c, hipMemcpyAsync(po.data(), c->pofs.p, po.size() * sizeof(int64_t), hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
    }
    c->nseeds.assign(npairs, 0);
    c->cseeds.assign(npairs, 0);
    int64_t nbytes = 0, cbytes = 0;
    for (int p = 0; p < npairs; p++) {
        PairDev &d = c->pairs[p];
        c->nseeds[p] = po[p + 1] - po[p];
        c->cseeds[p] = po[npairs + 1 + p + 1] - po[npairs + 1 + p];
        d.nseed_off = nbytes;
        d.cseed_off = cbytes;
        nbytes += c->nseeds[p] * d.s;
        cbytes += c->cseeds[p] * d.s;
    }
    c->seedN_p = static_cast<uint8_t *>(arena.take((size_t) nbytes + 16));
    if (!c->seedN_p) {
        TRY(ensure(c, c->seedN, (size_t) nbytes + 16));
        c->seedN_p = P<uint8_t>(c->seedN);
    }
    c->seedC_p = static_cast<uint8_t *>(arena.take((size_t) cbytes + 16));
    if (!c->seedC_p) {
        TRY(ensure(c, c->seedC, (size_t) cbytes + 16));
        c->seedC_p = P<uint8_t>(c->seedC);
    }
    const int64_t nb_elems = std::max(c->nctg, 1);
    TRY(ensure(c, c->nbuck, (size_t) nb_elems * sizeof(uint64_t)));
    TRY(ensure(c, c->cbuck, (size_t) nb_elems * sizeof(uint64_t)));
    TRY(fill_u64(c, c->nbuck.p, nb_elems, 0));
    TRY(fill_u64(c, c->cbuck.p, nb_elems, 0));
    HIPCHK(c, hipMemcpyAsync(c->dpairs.p, c->pairs.data(), (size_t) npairs * sizeof(PairDev), hipMemcpyHostToDevice, c->st));
    // a whole number of resident waves of blocks, four rounds deep
    int occ = 1;
    {
        const void *fn = MT == 512 ? (const void *) k_merge_emit<2>
                       : MT == 1024 ? (const void *) k_merge_emit<4> : (const void *) k_merge_emit<1>;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, MTH, 0) != hipSuccess || occ < 1) {
            (void) hipGetLastError();
            occ = 1;
        }
    }
    const int64_t egrid = std::min<int64_t>(nmb, (int64_t) c->ncu * occ * 4);
    const int64_t eper = egrid > 0 ? (nmb + egrid - 1) / egrid : 0;
    TRY(record(c, 7));
    if (nmb > 0) {
#define PGA_EMIT(A_)                                                                                    \
    hipLaunchKernelGGL(k_merge_emit<A_>, dim3((unsigned) egrid), dim3(MTH), 0, c->st,                   \
                       (const PairDev *) P<PairDev>(c->dpairs), npairs,                                 \
                       (const uint8_t *) P<uint8_t>(c->entries), c->o.freq,                            \
                       (const int64_t *) P<int64_t>(c->mofsN), (const int64_t *) P<int64_t>(c->mofsC),  \
                       c->seedN_p, c->seedC_p, P<unsigned long long>(c->nbuck),                         \
                       P<unsigned long long>(c->cbuck), (const uint32_t *) minfo, nmb, eper,            \
                       (const int64_t *) nullptr, (const int *) nullptr)
        switch (MT) {
        case 512: PGA_EMIT(2); break;
        case 1024: PGA_EMIT(4); break;
        default: PGA_EMIT(1); break;
        }
#undef PGA_EMIT
        LAUNCHCHK(c);
    }

    TRY(record(c, 6));
    c->tm_merge = true;                  // times read when asked for
    c->merged = true;
    return PGA_OK;
}

int pga_seeds_info(pga_ctx *c, int pair, pga_seeds_meta *info)
{
    if (!c || !c->merged || pair < 0 || pair >= c->npairs) return c ? fail(c, PGA_EINVAL, "no seeds for pair %d", pair) : PGA_EINVAL;
    const PairDev &d = c->pairs[pair];
    info->nlen = c->nseeds[pair] * d.s;
    info->clen = c->cseeds[pair] * d.s;
    info->s = d.s;
    info->ncontig = d.ncontig;
    return PGA_OK;
}

int pga_seeds_export(pga_ctx *c, int pair, uint8_t *nseeds, uint8_t *cseeds, int64_t *nbuck, int64_t *cbuck)
{
    if (!c || !c->merged || pair < 0 || pair >= c->npairs) return c ? fail(c, PGA_EINVAL, "no seeds for pair %d", pair) : PGA_EINVAL;
    const PairDev &d = c->pairs[pair];
    if (nseeds && c->nseeds[pair] > 0)
        HIPCHK(c, hipMemcpyAsync(nseeds, c->seedN_p + d.nseed_off, (size_t) (c->nseeds[pair] * d.s),
                                 hipMemcpyDeviceToHost, c->st));
    if (cseeds && c->cseeds[pair] > 0)
        HIPCHK(c, hipMemcpyAsync(cseeds, c->seedC_p + d.cseed_off, (size_t) (c->cseeds[pair] * d.s),
                                 hipMemcpyDeviceToHost, c->st));
    const int64_t row = 0;
    if (nbuck && d.ncontig > 0)
        HIPCHK(c, hipMemcpyAsync(nbuck, P<uint64_t>(c->nbuck) + row + d.buck_off, (size_t) d.ncontig * sizeof(int64_t),
                                 hipMemcpyDeviceToHost, c->st));
    if (cbuck && d.ncontig > 0)
        HIPCHK(c, hipMemcpyAsync(cbuck, P<uint64_t>(c->cbuck) + row + d.buck_off, (size_t) d.ncontig * sizeof(int64_t),
                                 hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return PGA_OK;
}

int pga_index_export_range(pga_ctx *c, int pair, int64_t first, int64_t count, uint8_t *entries)
{
    if (!c || !c->built || pair < 0 || pair >= c->npairs) return c ? fail(c, PGA_EINVAL, "no index for pair %d", pair) : PGA_EINVAL;
    const PairDev &d = c->pairs[pair];
    if (first < 0 || count < 0 || first + count > d.nrec || (count > 0 && !entries))
        return fail(c, PGA_EINVAL, "entries [%lld, +%lld) outside the index", (long long) first, (long long) count);
    if (count > 0)
        HIPCHK(c, hipMemcpyAsync(entries, P<uint8_t>(c->entries) + d.ent_off + first * d.e, (size_t) (count * d.e),
                                 hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return PGA_OK;
}

// The seeds of anchor a start, in each stream, at its merge tile's offset
// plus the seeds of the tile's anchors before it (their interval records).
int pga_seeds_locate(pga_ctx *c, int pair, int64_t anchor, int64_t *nfirst, int64_t *cfirst)
{
    if (!c || !c->merged || pair < 0 || pair >= c->npairs) return c ? fail(c, PGA_EINVAL, "no seeds for pair %d", pair) : PGA_EINVAL;
    const PairDev &d = c->pairs[pair];
    if (anchor < 0 || anchor > d.nrec || !nfirst || !cfirst)
        return fail(c, PGA_EINVAL, "anchor %lld outside the index", (long long) anchor);
    if (anchor == d.nrec) {
        *nfirst = c->nseeds[pair];
        *cfirst = c->cseeds[pair];
        return PGA_OK;
    }
    HIPCHK(c, hipSetDevice(c->o.device));
    const int64_t MT = c->o.merge_tile;
    const int64_t tile = d.mblk0 + anchor / MT, t0 = anchor / MT * MT;
    int64_t ofs[3] = {0, 0, 0};
    std::vector<uint32_t> mi((size_t) (anchor - t0) + 1);
    HIPCHK(c, hipMemcpyAsync(&ofs[0], P<int64_t>(c->mofsN) + tile, sizeof(int64_t), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(&ofs[1], P<int64_t>(c->mofsN) + d.mblk0, sizeof(int64_t), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(&ofs[2], P<int64_t>(c->mofsC) + tile, sizeof(int64_t), hipMemcpyDeviceToHost, c->st));
    int64_t cofs0 = 0;
    HIPCHK(c, hipMemcpyAsync(&cofs0, P<int64_t>(c->mofsC) + d.mblk0, sizeof(int64_t), hipMemcpyDeviceToHost, c->st));
    if (anchor > t0)
        HIPCHK(c, hipMemcpyAsync(mi.data(), c->minfo_p + d.rec0 + t0, (size_t) (anchor - t0) * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    int64_t n = ofs[0] - ofs[1], cc = ofs[2] - cofs0;
    for (int64_t k = 0; k < anchor - t0; k++) {
        n += (mi[k] >> 18) & 63u;
        cc += (mi[k] >> 24) & 63u;
    }
    *nfirst = n;
    *cfirst = cc;
    return PGA_OK;
}

int pga_seeds_export_range(pga_ctx *c, int pair, int comp, int64_t first, int64_t count, uint8_t *out)
{
    if (!c || !c->merged || pair < 0 || pair >= c->npairs || (comp != 0 && comp != 1))
        return c ? fail(c, PGA_EINVAL, "no seeds for pair %d / stream %d", pair, comp) : PGA_EINVAL;
    const PairDev &d = c->pairs[pair];
    const int64_t n = comp ? c->cseeds[pair] : c->nseeds[pair];
    if (first < 0 || count < 0 || first + count > n || (count > 0 && !out))
        return fail(c, PGA_EINVAL, "seeds [%lld, +%lld) outside stream %d", (long long) first, (long long) count, comp);
    const uint8_t *src = comp ? c->seedC_p + d.cseed_off : c->seedN_p + d.nseed_off;
    if (count > 0)
        HIPCHK(c, hipMemcpyAsync(out, src + first * d.s, (size_t) (count * d.s), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return PGA_OK;
}

// Diagonal-sorted seed records of one stream of one pair (see k_seed_keys).
int pga_seed_sort(pga_ctx *c, int pair, int comp, int64_t *nels, int *swide_out)
{
    if (!c || !c->merged || pair < 0 || pair >= c->npairs || (comp != 0 && comp != 1))
        return c ? fail(c, PGA_EINVAL, "no seeds for pair %d / stream %d", pair, comp) : PGA_EINVAL;
    HIPCHK(c, hipSetDevice(c->o.device));
    const PairDev &d = c->pairs[pair];
    const int64_t n = comp ? c->cseeds[pair] : c->nseeds[pair];
    const uint8_t *stream = comp ? c->seedC_p + d.cseed_off : c->seedN_p + d.nseed_off;
    const int64_t maxctg = c->hp[pair].maxctg;
    int dbyte = 0;                                       // src/pga-mpi-merge.c:415-425
    for (int64_t cum = 1; cum < 2 * maxctg; cum *= 256) dbyte++;
    const int swide = 2 * dbyte + d.C + 2;
    // key fields at their exact widths: contig ids, band (diag >> 6), anti
    // (stored in dbyte bytes), then diag&63 and lcp in 6 bits each
    const int64_t dmax = dbyte >= 8 ? INT64_MAX : ((int64_t) 1 << (8 * dbyte)) - 1;
    SeedKeyFields kf;
    kf.bc = bits_for(std::max(d.ncontig - 1, 0));
    kf.bb = bits_for(std::min<int64_t>((2 * maxctg) >> 6, dmax));
    kf.ba = bits_for(std::min<int64_t>(2 * maxctg, dmax));
    const int kbits = 2 * kf.bc + kf.bb + kf.ba + 12;
    if (kbits > 128) return fail(c, PGA_EOVERFLOW, "seed sort key of %d bits exceeds 128", kbits);
    kf.shift = 128 - kbits;
    c->ss_n = n;
    c->ss_swide = swide;
    c->ss_ready = false;
    const size_t n1 = (size_t) std::max<int64_t>(n, 1);
    TRY(ensure(c, c->ssk[0], n1 * sizeof(Rec)));
    TRY(ensure(c, c->ssk[1], n1 * sizeof(Rec)));
    TRY(ensure(c, c->ssbufs, 2 * sizeof(Rec *)));
    TRY(ensure(c, c->ssout, n1 * (size_t) swide + 16));
    {
        Rec *bp[2] = {P<Rec>(c->ssk[0]), P<Rec>(c->ssk[1])};
        HIPCHK(c, hipMemcpyAsync(c->ssbufs.p, bp, sizeof(bp), hipMemcpyHostToDevice, c->st));
    }
    if (n > 0) {
        // the keys' first-digit counts go to the pair's row of hist1 (the
        // build's counts are no longer needed once it is done)
        unsigned long long *hrow = P<unsigned long long>(c->hist1) + (size_t) pair * 256;
        TRY(fill_u64(c, hrow, 256, 0));
        hipLaunchKernelGGL(k_seed_keys, dim3(grid_for(c, (const void *) k_seed_keys, 256, 0, (n + 255) / 256)), dim3(256),
                           0, c->st, stream, n, d.P, d.C, maxctg, dbyte, comp, kf, P<Rec>(c->ssk[0]), hrow);
        LAUNCHCHK(c);
        // partition + tile sort, the keys ending in ssk[0]; the build's statistics are kept
        const int CAP = cap_of(c);
        std::vector<Seg> seg0;
        std::vector<Tile> tile0;
        if (n > CAP) seg0.push_back(Seg{0, n, pair, 1});
        else tile0.push_back(Tile{0, (int32_t) n, pair, 0, 1, drange_all(8)});
        const int64_t seg_cap = n / CAP + 16;
        const int64_t tile_cap = 34 * (n / CAP) + 1024;
        TRY(ensure(c, c->segs[0], (size_t) seg_cap * sizeof(Seg)));
        TRY(ensure(c, c->segs[1], (size_t) seg_cap * sizeof(Seg)));
        TRY(ensure(c, c->seghist, (size_t) seg_cap * 256 * sizeof(uint64_t)));
        TRY(ensure(c, c->cursor, (size_t) seg_cap * 256 * sizeof(uint64_t)));
        TRY(ensure(c, c->chk0, (size_t) (seg_cap + 2) * sizeof(int64_t)));
        TRY(ensure(c, c->tiles, (size_t) tile_cap * sizeof(Tile)));
        const int ns = c->nscatter;
        const int64_t sb = c->scatter_bytes;
        const int64_t nt = c->ntiles;
        int levels = 0;
        const int rc = partition(c, P<Rec>(c->ssk[0]), P<Rec>(c->ssk[1]), seg0, tile0,
                                 P<unsigned long long>(c->hist1), false, seg_cap, tile_cap, nullptr, false, &levels);
        const int64_t sort_tiles = c->ntiles;
        c->nscatter = ns;
        c->scatter_bytes = sb;
        c->ntiles = nt;
        TRY(rc);
        if (sort_tiles > 0)
            TRY(launch_tile_sort(c, (unsigned) sort_tiles, (Rec *const *) c->ssbufs.p, nullptr, nullptr, nullptr,
                                 nullptr, P<Rec>(c->ssk[0]), nullptr));
        const int64_t nb = (n + SP_T * SP_PER - 1) / (SP_T * SP_PER);
        hipLaunchKernelGGL(k_seed_pack, dim3(grid_for(c, (const void *) k_seed_pack, SP_T, 0, nb)), dim3(SP_T), 0,
                           c->st, (const Rec *) P<Rec>(c->ssk[0]), n, kf, dbyte, swide, P<uint8_t>(c->ssout));
        LAUNCHCHK(c);
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->ss_ready = true;
    if (nels) *nels = n;
    if (swide_out) *swide_out = swide;
    return PGA_OK;
}

int pga_seed_sort_export(pga_ctx *c, uint8_t *records)
{
    if (!c || !c->ss_ready) return c ? fail(c, PGA_EINVAL, "no seed sort to export") : PGA_EINVAL;
    if (records && c->ss_n > 0)
        HIPCHK(c, hipMemcpyAsync(records, c->ssout.p, (size_t) (c->ss_n * c->ss_swide), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return PGA_OK;
}

int pga_timings_get(pga_ctx *c, pga_timings *t)
{
    if (!c) return PGA_EINVAL;
    if (c->tm_build || c->tm_merge) {
        HIPCHK(c, hipStreamSynchronize(c->st));
        if (c->tm_build) {
            c->tm.extract_ms = ev_ms(c->ev[0], c->ev[1]);
            c->tm.partition_ms = ev_ms(c->ev[1], c->ev[2]);
            c->tm.tile_sort_ms = ev_ms(c->ev[2], c->ev[3]);
            c->tm.table_ms = ev_ms(c->ev[3], c->ev[4]);
            c->tm.total_ms = ev_ms(c->ev[0], c->ev[4]);
            double sc = 0;
            for (int i = 0; i < (int) c->tm.scatter_launches; i++) sc += ev_ms(c->sev[2 * i], c->sev[2 * i + 1]);
            c->tm.scatter_ms = sc;
            int nf[2] = {0, 0};
            HIPCHK(c, hipMemcpyAsync(nf, P<int>(c->counters) + 10, sizeof(nf), hipMemcpyDeviceToHost, c->st));
            HIPCHK(c, hipStreamSynchronize(c->st));
            c->tm.fallback_tiles = nf[0];
            c->tm.crowded_tiles = nf[1];
        }
        if (c->tm_merge) {
            c->tm.merge_ms = ev_ms(c->ev[5], c->ev[6]);
            c->tm.merge_emit_ms = ev_ms(c->ev[7], c->ev[6]);
        }
        c->tm_build = c->tm_merge = false;
    }
    c->tm.graph_fallbacks = c->graph_fallbacks;
    *t = c->tm;
    return PGA_OK;
}

} // extern "C"

// ======================================================================== device-driven pipeline
//
// pga_run_graph: the build and the merge of the loaded batch with no host
// round trip.  Every count a kernel needs -- records, segments of each
// partition level, sort tiles, merge tiles, seed totals -- stays in device
// memory (GStat, the per-pair layout) and is read there; buffers are sized
// from bounds known at load time (at most two records a candidate position).
// The launch sequence is captured once per loaded batch into a HIP graph and
// replayed; the host reads one status block afterwards.  When the status
// says the partition needed more than GRAPH_LEVELS levels, a list outgrew its
// bound, or the seeds outgrew their buffers, the eager path redoes the work.

namespace {

constexpr int PLAN_GRID = 2048;      // persistent blocks of the plan kernel

int graph_sizes(pga_ctx *c, int64_t &R_ub, int64_t &ent_ub, int64_t &nmb_ub)
{
    const int MT = c->o.merge_tile;
    R_ub = 0;
    ent_ub = 0;
    nmb_ub = 0;
    for (int p = 0; p < c->npairs; p++) {
        const HostPair &h = c->hp[p];
        const int64_t r = 2 * h.nblk * (int64_t) XBLOCK;        // two records a candidate position at most
        const int e = 9 + bytes_for(h.maxctg) + bytes_for(2 * (int64_t) h.ncontig - 1);
        R_ub += r;
        ent_ub += r * e;
        nmb_ub += (r + MT - 1) / MT;
    }
    return PGA_OK;
}

// Every buffer of the device-driven pipeline, at its bound; the load-time
// fields of the pair layout uploaded.  No allocation happens after this.
int graph_prepare(pga_ctx *c)
{
    const int npairs = c->npairs;
    const int CAP = cap_of(c);
    int64_t R_ub, ent_ub, nmb_ub;
    TRY(graph_sizes(c, R_ub, ent_ub, nmb_ub));
    const size_t r1 = (size_t) std::max<int64_t>(R_ub, 1);
    const int64_t nblk = c->nblk;
    TRY(ensure(c, c->pr, (size_t) (npairs + 1) * sizeof(int64_t)));
    TRY(ensure(c, c->sel, (size_t) std::max<int64_t>(nblk, 1) * 256));
    TRY(ensure(c, c->rec[0], r1 * sizeof(Rec)));
    TRY(ensure(c, c->rec[1], r1 * sizeof(Rec)));
    TRY(ensure(c, c->bufs, 2 * sizeof(Rec *)));
    TRY(ensure(c, c->entries, (size_t) ent_ub + 32));
    TRY(ensure(c, c->ls, r1 + 64));
    TRY(ensure(c, c->pre, (r1 + 1) * sizeof(uint32_t)));
    TRY(ensure(c, c->hist1, (size_t) npairs * 256 * sizeof(uint64_t)));
    TRY(ensure(c, c->dpairs, (size_t) npairs * sizeof(PairDev)));
    TRY(ensure(c, c->counters, 16 * sizeof(int)));
    const int64_t seg_cap = R_ub / CAP + npairs + 16;
    const int64_t tile_cap = 34 * (R_ub / CAP) + npairs + 1024;
    const int64_t rows = std::max<int64_t>(std::min<int64_t>((int64_t) 256 * npairs, seg_cap) << c->o.digit2_bits,
                                           seg_cap * 256);
    TRY(ensure(c, c->segs[0], (size_t) seg_cap * sizeof(Seg)));
    TRY(ensure(c, c->segs[1], (size_t) seg_cap * sizeof(Seg)));
    TRY(ensure(c, c->seghist, (size_t) rows * sizeof(uint64_t)));
    TRY(ensure(c, c->cursor, (size_t) rows * sizeof(uint64_t)));
    TRY(ensure(c, c->chk0, (size_t) (seg_cap + 2) * sizeof(int64_t)));
    TRY(ensure(c, c->tiles, (size_t) tile_cap * sizeof(Tile)));
    TRY(ensure(c, c->maxpre, (size_t) npairs * sizeof(uint64_t)));
    TRY(ensure(c, c->cum, (size_t) npairs * NPRE * sizeof(int64_t)));
    TRY(ensure(c, c->plist_all, (size_t) npairs * sizeof(int)));
    TRY(ensure(c, c->gmax, (size_t) npairs * (NPRE / CR_SPAN) * sizeof(uint64_t)));
    TRY(ensure(c, c->cbmax, (size_t) npairs * (NPRE / CR_SPAN) * sizeof(uint64_t)));
    TRY(ensure(c, c->mblk, (size_t) (npairs + 1) * sizeof(int64_t)));
    TRY(ensure(c, c->mcntN, (size_t) (nmb_ub + 2) * sizeof(uint32_t)));
    TRY(ensure(c, c->mcntC, (size_t) (nmb_ub + 2) * sizeof(uint32_t)));
    TRY(ensure(c, c->mofsN, (size_t) (nmb_ub + 1) * sizeof(int64_t)));
    TRY(ensure(c, c->mofsC, (size_t) (nmb_ub + 1) * sizeof(int64_t)));
    TRY(ensure(c, c->pofs, (size_t) 2 * (npairs + 1) * sizeof(int64_t)));
    TRY(ensure(c, c->nbuck, (size_t) std::max(c->nctg, 1) * sizeof(uint64_t)));
    TRY(ensure(c, c->cbuck, (size_t) std::max(c->nctg, 1) * sizeof(uint64_t)));
    TRY(ensure(c, c->gstat, sizeof(GStat)));
    const int64_t nscan = std::max<int64_t>(std::max<int64_t>(nblk, nmb_ub), 1);
    TRY(ensure(c, c->bsum, (size_t) ((nscan + 4095) / 4096 + 2) * sizeof(int64_t)));
    // load-time fields of the pair layout (the rest is written on the device)
    c->pairs.assign(npairs, PairDev());
    for (int p = 0; p < npairs; p++) {
        PairDev &d = c->pairs[p];
        const HostPair &h = c->hp[p];
        d.P = bytes_for(h.maxctg);
        d.C = bytes_for(2 * (int64_t) h.ncontig - 1);
        d.e = 9 + d.P + d.C;
        d.s = 1 + 2 * (d.P + d.C);
        d.ncontig = h.ncontig;
        d.cum_off = (int64_t) p * NPRE;
        d.buck_off = h.ctg0;
    }
    HIPCHK(c, hipMemcpyAsync(c->dpairs.p, c->pairs.data(), (size_t) npairs * sizeof(PairDev), hipMemcpyHostToDevice,
                             c->st));
    std::vector<int> all(npairs);
    for (int p = 0; p < npairs; p++) all[p] = p;
    HIPCHK(c, hipMemcpyAsync(c->plist_all.p, all.data(), all.size() * sizeof(int), hipMemcpyHostToDevice, c->st));
    Rec *bp[2] = {P<Rec>(c->rec[0]), P<Rec>(c->rec[1])};
    HIPCHK(c, hipMemcpyAsync(c->bufs.p, bp, sizeof(bp), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));          // host arrays above are locals
    // merge outputs inside the record buffers (dead after the tile sort):
    // interval records then the N stream in buffer 0, the C stream in buffer 1
    c->g_minfo = P<uint8_t>(c->rec[0]);
    c->g_seedN = P<uint8_t>(c->rec[0]) + ((r1 * sizeof(uint32_t) + 255) & ~(size_t) 255);
    c->g_seedC = P<uint8_t>(c->rec[1]);
    c->g_nmb_ub = nmb_ub;
    c->g_seg_cap = seg_cap;
    c->g_tile_cap = tile_cap;
    return PGA_OK;
}

// The whole launch sequence on the context's stream: no host read, no
// allocation, no memset (buffers are cleared by kernels).
int graph_enqueue(pga_ctx *c)
{
    const int npairs = c->npairs;
    const int CAP = cap_of(c);
    const int MT = c->o.merge_tile;
    const int per = chunk_per(c);
    const int CH = per * 256;
    GStat *gs = P<GStat>(c->gstat);
    const int64_t nblk = c->nblk;
    // ---- extraction
    if (nblk > 0) {
        hipLaunchKernelGGL(k_extract_count, dim3(grid_for(c, (const void *) k_extract_count, 256, 0, c->nrange)),
                           dim3(256), 0, c->st, P<uint64_t>(c->bps), P<ContigDev>(c->ctg),
                           (const BlkDev *) P<BlkDev>(c->blk), c->nctg, c->F, (const XRange *) P<XRange>(c->xrange),
                           c->nrange, P<uint8_t>(c->sel), P<uint32_t>(c->rangehist),
                           (uint32_t) c->shard_lo, (uint32_t) c->shard_hi);
        LAUNCHCHK(c);
    }
    if (c->nrange > 0) {
        hipLaunchKernelGGL(k_xrange_scan, dim3(npairs, 16), dim3(256), 0, c->st, (const int *) P<int>(c->prange),
                           (const uint32_t *) P<uint32_t>(c->rangehist), P<int64_t>(c->rangeoff),
                           P<unsigned long long>(c->hist1));
        LAUNCHCHK(c);
    } else {
        TRY(fill_u64(c, c->hist1.p, (int64_t) npairs * 256, 0));
    }
    hipLaunchKernelGGL(k_pair_totals, dim3(1), dim3(256), 0, c->st,
                       (const unsigned long long *) P<unsigned long long>(c->hist1), npairs, P<int64_t>(c->pr));
    hipLaunchKernelGGL(k_graph_layout, dim3(1), dim3(256), 0, c->st, (const int64_t *) P<int64_t>(c->pr), npairs, MT,
                       CAP, P<PairDev>(c->dpairs), P<int64_t>(c->mblk), P<Seg>(c->segs[0]), P<Tile>(c->tiles), gs);
    LAUNCHCHK(c);
    TRY(fill_u64(c, P<int>(c->counters) + 10, 1, 0));                       // tile sort path counters
    hipLaunchKernelGGL(k_xrange_base, dim3(npairs), dim3(256), 0, c->st, (const int64_t *) P<int64_t>(c->pr),
                       (const unsigned long long *) P<unsigned long long>(c->hist1), P<int64_t>(c->bstart));
    LAUNCHCHK(c);
    if (c->nrange > 0 && nblk > 0) {
        hipLaunchKernelGGL(k_extract_emit<EMIT_QB>,
                           dim3(grid_for(c, (const void *) k_extract_emit<EMIT_QB>, 256 * EMIT_QB, 0, c->nrange)),
                           dim3(256 * EMIT_QB), 0, c->st, P<uint64_t>(c->bps), P<ContigDev>(c->ctg),
                           (const BlkDev *) P<BlkDev>(c->blk), c->nctg, c->jbits, (const uint8_t *) P<uint8_t>(c->sel),
                           (const XRange *) P<XRange>(c->xrange), c->nrange,
                           (const uint32_t *) P<uint32_t>(c->rangehist), (const int64_t *) P<int64_t>(c->rangeoff),
                           (const int64_t *) P<int64_t>(c->bstart), P<Rec>(c->rec[1]), (uint8_t *) nullptr,
                           &gs->bad);
        LAUNCHCHK(c);
    }
    // ---- MSD partition, GRAPH_LEVELS levels, counts on the device
    Rec *buf[2] = {P<Rec>(c->rec[0]), P<Rec>(c->rec[1])};
    int cur = 0;
    for (int level = 1; level <= GRAPH_LEVELS; level++) {
        const int *dn = &gs->nsegs[level];
        hipLaunchKernelGGL(k_clamp_count, dim3(1), dim3(64), 0, c->st, &gs->nsegs[level], (int) c->g_seg_cap,
                           &gs->capped);
        hipLaunchKernelGGL(k_chunk_prefix, dim3(1), dim3(1024), 0, c->st, (const Seg *) P<Seg>(c->segs[cur]), 0, CH,
                           P<int64_t>(c->chk0), dn);
        LAUNCHCHK(c);
        const bool first = level == 1;                 // counted and in place: the extraction wrote it
        const int64_t nd = level == 2 ? (int64_t) 1 << c->o.digit2_bits : 256;
        if (!first) {
            hipLaunchKernelGGL(k_fill_u64_dev, dim3(1024), dim3(256), 0, c->st, P<unsigned long long>(c->seghist), dn,
                               nd, 0ull);
            TRY(launch_seg_hist(c, per, buf[0], buf[1], P<Seg>(c->segs[cur]), 0, P<int64_t>(c->chk0), level,
                                P<unsigned long long>(c->seghist), P<uint8_t>(c->ls), dn));
        }
        hipLaunchKernelGGL(k_seg_plan, dim3(PLAN_GRID), dim3(256), 0, c->st, P<Seg>(c->segs[cur]),
                           first ? (const unsigned long long *) P<unsigned long long>(c->hist1)
                                 : (const unsigned long long *) P<unsigned long long>(c->seghist),
                           first ? 1 : 0, level, c->o.digit2_bits, first ? 1 : 0, first ? 0 : 1, CAP,
                           P<unsigned long long>(c->cursor), P<Seg>(c->segs[cur ^ 1]), &gs->nsegs[level + 1],
                           (int) c->g_seg_cap, P<Tile>(c->tiles), &gs->ntiles, (int) c->g_tile_cap, &gs->kept, 0, dn);
        LAUNCHCHK(c);
        if (!first)
            TRY(launch_seg_scatter(c, per, buf[0], buf[1], P<Seg>(c->segs[cur]), 0, P<int64_t>(c->chk0), level,
                                   P<unsigned long long>(c->cursor), P<uint8_t>(c->ls), dn));
        cur ^= 1;
    }
    // ---- tile sort (group ends marked in the cleared prefix tables), LCP
    // fix-up, prefix tables
    hipLaunchKernelGGL(k_clamp_count, dim3(1), dim3(64), 0, c->st, &gs->ntiles, (int) c->g_tile_cap, &gs->capped);
    if (c->o.build_cum) TRY(fill_u64(c, c->cum.p, (int64_t) npairs * NPRE, 0));
    TRY(launch_tile_sort(c, (unsigned) c->g_tile_cap, (Rec *const *) c->bufs.p, P<uint8_t>(c->entries),
                         P<uint8_t>(c->ls), P<uint32_t>(c->pre),
                         c->o.build_cum ? P<unsigned long long>(c->cum) : nullptr, nullptr,
                         P<int>(c->counters) + 10, &gs->ntiles));
    hipLaunchKernelGGL(k_tile_fix, dim3(grid_for(c, (const void *) k_tile_fix, 256, 0, (c->g_tile_cap + 255) / 256)),
                       dim3(256), 0, c->st, (const Tile *) P<Tile>(c->tiles), 0, (const PairDev *) P<PairDev>(c->dpairs),
                       P<uint8_t>(c->entries), P<uint8_t>(c->ls), (const uint32_t *) P<uint32_t>(c->pre),
                       (const int *) &gs->ntiles);
    LAUNCHCHK(c);
    TRY(fill_u64(c, c->maxpre.p, npairs, 0));
    if (c->o.build_cum) TRY(fill_prefix_tables(c, npairs, (const int *) P<int>(c->plist_all)));
    // ---- merge
    const int64_t nmb_ub = c->g_nmb_ub;
    uint32_t *minfo = reinterpret_cast<uint32_t *>(c->g_minfo);
    TRY(fill_u64(c, c->mcntN.p, (nmb_ub + 1) / 2, 0));
    TRY(fill_u64(c, c->mcntC.p, (nmb_ub + 1) / 2, 0));
    const int64_t *dnmb = &gs->nmb;
#define PGA_GCOUNT(A_)                                                                                  \
    hipLaunchKernelGGL(k_merge_count<A_>, dim3(grid_for(c, (const void *) k_merge_count<A_>, MTH, 0, nmb_ub)), \
                       dim3(MTH), 0, c->st, (const PairDev *) P<PairDev>(c->dpairs), npairs,             \
                       (const uint8_t *) P<uint8_t>(c->ls), c->o.freq, P<uint32_t>(c->mcntN),           \
                       P<uint32_t>(c->mcntC), minfo, (int64_t) 0, dnmb)
    switch (MT) {
    case 512: PGA_GCOUNT(2); break;
    case 1024: PGA_GCOUNT(4); break;
    default: PGA_GCOUNT(1); break;
    }
#undef PGA_GCOUNT
    LAUNCHCHK(c);
    TRY(scan_u32(c, P<uint32_t>(c->mcntN), nmb_ub, P<int64_t>(c->mofsN)));
    TRY(scan_u32(c, P<uint32_t>(c->mcntC), nmb_ub, P<int64_t>(c->mofsC)));
    hipLaunchKernelGGL(k_pair_ranges, dim3((npairs + 256) / 256), dim3(256), 0, c->st,
                       (const int64_t *) P<int64_t>(c->mofsN), (const int64_t *) P<int64_t>(c->mblk), npairs,
                       P<int64_t>(c->pofs));
    hipLaunchKernelGGL(k_pair_ranges, dim3((npairs + 256) / 256), dim3(256), 0, c->st,
                       (const int64_t *) P<int64_t>(c->mofsC), (const int64_t *) P<int64_t>(c->mblk), npairs,
                       P<int64_t>(c->pofs) + (npairs + 1));
    const int64_t capN = (int64_t) (P<uint8_t>(c->rec[0]) + c->rec[0].cap - c->g_seedN);
    const int64_t capC = (int64_t) c->rec[1].cap;
    hipLaunchKernelGGL(k_graph_seeds, dim3(1), dim3(64), 0, c->st, (const int64_t *) P<int64_t>(c->pofs), npairs,
                       P<PairDev>(c->dpairs), capN, capC, gs);
    LAUNCHCHK(c);
    TRY(fill_u64(c, c->nbuck.p, std::max(c->nctg, 1), 0));
    TRY(fill_u64(c, c->cbuck.p, std::max(c->nctg, 1), 0));
    int occ = 1;
    {
        const void *fn = MT == 512 ? (const void *) k_merge_emit<2>
                       : MT == 1024 ? (const void *) k_merge_emit<4> : (const void *) k_merge_emit<1>;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, MTH, 0) != hipSuccess || occ < 1) {
            (void) hipGetLastError();
            occ = 1;
        }
    }
    const int64_t egrid = std::max<int64_t>(1, std::min<int64_t>(nmb_ub, (int64_t) c->ncu * occ * 4));
#define PGA_GEMIT(A_)                                                                                   \
    hipLaunchKernelGGL(k_merge_emit<A_>, dim3((unsigned) egrid), dim3(MTH), 0, c->st,                   \
                       (const PairDev *) P<PairDev>(c->dpairs), npairs,                                 \
                       (const uint8_t *) P<uint8_t>(c->entries), c->o.freq,                            \
                       (const int64_t *) P<int64_t>(c->mofsN), (const int64_t *) P<int64_t>(c->mofsC),  \
                       c->g_seedN, c->g_seedC, P<unsigned long long>(c->nbuck),                         \
                       P<unsigned long long>(c->cbuck), (const uint32_t *) minfo, (int64_t) 0, (int64_t) 0, \
                       dnmb, (const int *) &gs->overflow)
    switch (MT) {
    case 512: PGA_GEMIT(2); break;
    case 1024: PGA_GEMIT(4); break;
    default: PGA_GEMIT(1); break;
    }
#undef PGA_GEMIT
    LAUNCHCHK(c);
    return PGA_OK;
}

void graph_drop(pga_ctx *c)
{
    if (c->gexec) (void) hipGraphExecDestroy(c->gexec);
    if (c->graph) (void) hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
    c->graph_gen = ~0ull;
}

int graph_capture(pga_ctx *c)
{
    graph_drop(c);
    TRY(graph_prepare(c));
    HIPCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    const int rc = graph_enqueue(c);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(c->st, &g);
    if (rc) {
        if (g) (void) hipGraphDestroy(g);
        return rc;
    }
    if (ec != hipSuccess || !g) return fail(c, PGA_EHIP, "stream capture failed: %s", hipGetErrorString(ec));
    c->graph = g;
    HIPCHK(c, hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
    c->graph_gen = c->pool_gen;
    return PGA_OK;
}

} // namespace

extern "C" int pga_run_graph(pga_ctx *c)
{
    if (!c || !c->loaded) return c ? fail(c, PGA_EINVAL, "no GDB loaded") : PGA_EINVAL;
    HIPCHK(c, hipSetDevice(c->o.device));
    c->built = c->merged = c->ss_ready = false;
    if (!c->gexec || c->graph_gen != c->pool_gen) TRY(graph_capture(c));
    memset(&c->tm, 0, sizeof(c->tm));
    c->tm_build = c->tm_merge = false;
    TRY(record(c, 0));
    HIPCHK(c, hipGraphLaunch(c->gexec, c->st));
    TRY(record(c, 4));
    // the status, the seed ranges and the device-written pair layout
    const int npairs = c->npairs;
    GStat st;
    std::vector<int64_t> po(2 * (npairs + 1));
    std::vector<PairDev> pd(npairs);
    HIPCHK(c, hipMemcpyAsync(&st, c->gstat.p, sizeof(st), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(po.data(), c->pofs.p, po.size() * sizeof(int64_t), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(pd.data(), c->dpairs.p, pd.size() * sizeof(PairDev), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    const bool redo = st.overflow || st.bad || st.capped || st.nsegs[GRAPH_LEVELS + 1] > 0;
    if (redo) {
        // beyond the captured form: the eager path, which sizes as it goes
        c->graph_fallbacks++;
        TRY(pga_build_index(c));
        TRY(pga_merge(c));
        return PGA_OK;
    }
    c->R = st.R;
    c->pairs = pd;
    c->ntiles = st.ntiles;
    c->nseeds.assign(npairs, 0);
    c->cseeds.assign(npairs, 0);
    for (int p = 0; p < npairs; p++) {
        c->nseeds[p] = po[p + 1] - po[p];
        c->cseeds[p] = po[npairs + 1 + p + 1] - po[npairs + 1 + p];
    }
    c->minfo_p = reinterpret_cast<uint32_t *>(c->g_minfo);
    c->seedN_p = c->g_seedN;
    c->seedC_p = c->g_seedC;
    c->built = c->merged = true;
    int nf[2] = {0, 0};
    HIPCHK(c, hipMemcpyAsync(nf, P<int>(c->counters) + 10, sizeof(nf), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->tm.total_ms = ev_ms(c->ev[0], c->ev[4]);
    c->tm.tiles = st.ntiles;
    c->tm.fallback_tiles = nf[0];
    c->tm.crowded_tiles = nf[1];
    int levels = 0;
    for (int l = 1; l <= GRAPH_LEVELS; l++) if (st.nsegs[l] > 0) levels = l;
    c->tm.partition_levels = levels;
    return PGA_OK;
}

// ======================================================================== merge without allocation
//
// pga_merge_gix_noalloc: the self-merge of one GIX index that the caller holds
// on the device, with every intermediate in one file-scope device array
// (`scratch`) instead of pool buffers.  The kernels are the index merge's own
// (k_merge_count, the scans, k_merge_emit); a small kernel first derives the
// per-entry LCP/strand byte from the entries.

namespace {

constexpr int64_t NA_MAX = PGA_NOALLOC_MAX_ENTRIES;
constexpr int NA_A = 2;                                   // anchors per thread: 512-entry merge tiles
constexpr int64_t NA_TILE = (int64_t) MTH * NA_A;
constexpr int64_t NA_MAXTILES = (NA_MAX + NA_TILE - 1) / NA_TILE;
constexpr int64_t NA_SCANBLK = (NA_MAXTILES + SCAN_SPAN - 1) / SCAN_SPAN + 1;
constexpr int64_t na_round(int64_t x) { return (x + 255) & ~(int64_t) 255; }
// scratch layout (byte offsets, each 256-aligned)
constexpr int64_t NA_OFF_PAIR = 0;
constexpr int64_t NA_OFF_LS = NA_OFF_PAIR + na_round(sizeof(PairDev));
constexpr int64_t NA_OFF_MINFO = NA_OFF_LS + na_round(NA_MAX + 64);
constexpr int64_t NA_OFF_CNTN = NA_OFF_MINFO + na_round(NA_MAX * 4);
constexpr int64_t NA_OFF_CNTC = NA_OFF_CNTN + na_round(NA_MAXTILES * 4);
constexpr int64_t NA_OFF_OFSN = NA_OFF_CNTC + na_round(NA_MAXTILES * 4);
constexpr int64_t NA_OFF_OFSC = NA_OFF_OFSN + na_round((NA_MAXTILES + 1) * 8);
constexpr int64_t NA_OFF_BSUM = NA_OFF_OFSC + na_round((NA_MAXTILES + 1) * 8);
constexpr int64_t NA_BYTES = NA_OFF_BSUM + na_round((NA_SCANBLK + 1) * 8);
static_assert(NA_BYTES <= PGA_NOALLOC_SCRATCH_BYTES, "no-alloc scratch layout exceeds its stated size");

} // namespace

// One copy per GPU, alive until the process ends.  Host code never uses this
// name as an address: the device address comes from hipGetSymbolAddress.
static __device__ __align__(256) uint8_t scratch[PGA_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int NA_MAXDEV = 64;
std::mutex na_mu;
uint8_t *na_ptr[NA_MAXDEV] = {};

// The calling thread's current device's copy of `scratch` (looked up once per device).
int na_scratch(uint8_t **out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) { (void) hipGetLastError(); return PGA_EHIP; }
    if (dev < 0 || dev >= NA_MAXDEV) return PGA_EINVAL;
    std::lock_guard<std::mutex> g(na_mu);
    if (!na_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void) hipGetLastError();
            return PGA_EHIP;
        }
        na_ptr[dev] = static_cast<uint8_t *>(p);
    }
    *out = na_ptr[dev];
    return PGA_OK;
}

__global__ void k_na_pair(PairDev *dst, PairDev v) { *dst = v; }

// LCP | strand << 7 of every entry (the merge count's input): the LCP byte at
// offset 8, the strand as the top bit of the entry's last byte.
__global__ void k_na_ls(const uint8_t *__restrict__ ent, int e, int64_t n, uint8_t *__restrict__ ls)
{
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x)
        ls[i] = (uint8_t) ((ent[i * e + 8] & 63) | (ent[i * e + e - 1] & 0x80));
}

unsigned na_grid(const void *fn, int threads, int64_t nlog)
{
    int dev = 0, ncu = 256, occ = 1;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1) {
        (void) hipGetLastError();
        ncu = 256;
    }
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, threads, 0) != hipSuccess || occ < 1) {
        (void) hipGetLastError();
        occ = 1;
    }
    return (unsigned) std::max<int64_t>(1, std::min<int64_t>(nlog, (int64_t) ncu * occ));
}

} // namespace

extern "C" int pga_merge_gix_noalloc(const uint8_t *entries, int64_t entries_bytes, int64_t nrec, int P, int C,
                                     int ncontig, int freq, uint8_t *nseeds, int64_t ncap, uint8_t *cseeds,
                                     int64_t ccap, int64_t *nbuck, int64_t *cbuck, int64_t *nlen, int64_t *clen,
                                     void *stream)
{
    if (nlen) *nlen = 0;
    if (clen) *clen = 0;
    if (nrec < 0 || P < 1 || C < 1 || P + C > 7 || ncontig < 1 || freq < 1 || freq > MAX_FREQ ||
        (nrec > 0 && !entries) || !nbuck || !cbuck || !nlen || !clen)
        return PGA_EINVAL;
    if (nrec > NA_MAX) return PGA_ELIMIT;
    const int e = 9 + P + C, s = 1 + 2 * (P + C);
    if (entries_bytes < nrec * e + 8) return PGA_EINVAL;      // the payload loads read 8 bytes at offset 9
    hipStream_t st = static_cast<hipStream_t>(stream);
    uint8_t *sc = nullptr;
    TRY(na_scratch(&sc));
#define NA_CHK(call)                                                      \
    do {                                                                  \
        if ((call) != hipSuccess) { (void) hipGetLastError(); return PGA_EHIP; } \
    } while (0)
    PairDev pd{};
    pd.rec0 = 0;
    pd.nrec = nrec;
    pd.P = P;
    pd.C = C;
    pd.e = e;
    pd.s = s;
    pd.ncontig = ncontig;
    PairDev *dpair = reinterpret_cast<PairDev *>(sc + NA_OFF_PAIR);
    uint8_t *ls = sc + NA_OFF_LS;
    uint32_t *minfo = reinterpret_cast<uint32_t *>(sc + NA_OFF_MINFO);
    uint32_t *cntN = reinterpret_cast<uint32_t *>(sc + NA_OFF_CNTN), *cntC = reinterpret_cast<uint32_t *>(sc + NA_OFF_CNTC);
    int64_t *ofsN = reinterpret_cast<int64_t *>(sc + NA_OFF_OFSN), *ofsC = reinterpret_cast<int64_t *>(sc + NA_OFF_OFSC);
    int64_t *bsum = reinterpret_cast<int64_t *>(sc + NA_OFF_BSUM);
    const int64_t nmb = (nrec + NA_TILE - 1) / NA_TILE;

    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned) std::min<int64_t>((ncontig + 255) / 256, 4096)), dim3(256), 0, st,
                       reinterpret_cast<unsigned long long *>(nbuck), (int64_t) ncontig, 0ull);
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned) std::min<int64_t>((ncontig + 255) / 256, 4096)), dim3(256), 0, st,
                       reinterpret_cast<unsigned long long *>(cbuck), (int64_t) ncontig, 0ull);
    NA_CHK(hipGetLastError());
    if (nmb == 0) {
        NA_CHK(hipStreamSynchronize(st));
        return PGA_OK;
    }
    hipLaunchKernelGGL(k_na_pair, dim3(1), dim3(1), 0, st, dpair, pd);
    hipLaunchKernelGGL(k_na_ls, dim3(na_grid((const void *) k_na_ls, 256, (nrec + 255) / 256)), dim3(256), 0, st,
                       entries, e, nrec, ls);
    hipLaunchKernelGGL(k_merge_count<NA_A>, dim3(na_grid((const void *) k_merge_count<NA_A>, MTH, nmb)), dim3(MTH), 0,
                       st, (const PairDev *) dpair, 1, (const uint8_t *) ls, freq, cntN, cntC, minfo, nmb,
                       (const int64_t *) nullptr);
    NA_CHK(hipGetLastError());
    const int64_t nb = (nmb + SCAN_SPAN - 1) / SCAN_SPAN;
    for (int k = 0; k < 2; k++) {
        const uint32_t *in = k ? cntC : cntN;
        int64_t *out = k ? ofsC : ofsN;
        hipLaunchKernelGGL(k_scan_reduce<uint32_t>, dim3((unsigned) nb), dim3(SCAN_T), 0, st, in, nmb, bsum);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, bsum, nb);
        hipLaunchKernelGGL(k_scan_down<uint32_t>, dim3((unsigned) nb), dim3(SCAN_T), 0, st, in, nmb,
                           (const int64_t *) bsum, out);
        NA_CHK(hipGetLastError());
    }
    int64_t tot[2] = {0, 0};
    NA_CHK(hipMemcpyAsync(&tot[0], ofsN + nmb, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    NA_CHK(hipMemcpyAsync(&tot[1], ofsC + nmb, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    NA_CHK(hipStreamSynchronize(st));
    *nlen = tot[0] * s;
    *clen = tot[1] * s;
    if (*nlen > ncap || *clen > ccap) return PGA_EOVERFLOW;   // sizes of the streams reported, nothing written
    if ((*nlen > 0 && !nseeds) || (*clen > 0 && !cseeds)) return PGA_EINVAL;
    // a whole number of resident waves of blocks, four rounds deep (as pga_merge)
    const int64_t egrid = std::min<int64_t>(nmb, (int64_t) na_grid((const void *) k_merge_emit<NA_A>, MTH, nmb) * 4);
    const int64_t eper = (nmb + egrid - 1) / egrid;
    hipLaunchKernelGGL(k_merge_emit<NA_A>, dim3((unsigned) egrid), dim3(MTH), 0, st, (const PairDev *) dpair, 1,
                       entries, freq, (const int64_t *) ofsN, (const int64_t *) ofsC, nseeds, cseeds,
                       reinterpret_cast<unsigned long long *>(nbuck), reinterpret_cast<unsigned long long *>(cbuck),
                       (const uint32_t *) minfo, nmb, eper, (const int64_t *) nullptr, (const int *) nullptr);
    NA_CHK(hipGetLastError());
    NA_CHK(hipStreamSynchronize(st));
#undef NA_CHK
    return PGA_OK;
}

// ======================================================================== serving

#include "pga_serve.hip"

// One slot answers one batch of queries at a time on its own stream; slots of
// one server share the resident target and may run concurrently (one host
// thread per slot), so batches overlap on the device.
struct pga_slot {
    pga_ctx *Q = nullptr;
    char err_msg[512] = {0};
    bool have_result = false;
    int nq = 0;
    std::vector<QDev> qd;
    std::vector<int64_t> nseeds, cseeds;
    DBuf tpos, flag, segidx, segstart, segs, qdev, cntN, cntC, ofsN, ofsC, gath, qrec, seedN, seedC, nbuck, cbuck;
    DBuf jseg, cost, soff, pstart, err;
    DBuf tofN, tofC, tmpN, tmpC, used;    // the join's staging streams and each segment's place in them
    unsigned long long capN = 0, capC = 0;
    double rateN = 0, rateC = 0;          // largest staging bytes per query record seen so far
    hipEvent_t ev[2] = {};
    double last_ms = 0.0;
};

struct pga_server {
    pga_opts o{};
    pga_ctx *T = nullptr;
    std::vector<pga_slot *> slots;
    char err[512] = {0};
    bool have_target = false;
    int nT = 0;
    int64_t maxctgT = 0;
};

namespace {

constexpr int SERVE_JBITS = 32;    // offset bits of served keys: contigs up to 2^32 bases, 32767 contigs
constexpr int MAX_SLOTS = 8;

int sfail(char *err, int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, 512, fmt, ap);
    va_end(ap);
    return code;
}

// Copy a context's message into the caller-visible one.
int from_ctx(char *err, pga_ctx *c, int rc)
{
    if (rc) snprintf(err, 512, "%s", c ? c->err : "context");
    return rc;
}

void slot_free(pga_slot *q, int device)
{
    if (!q) return;
    (void) hipSetDevice(device);
    if (q->Q && q->Q->st) (void) hipStreamSynchronize(q->Q->st);
    DBuf *all[] = {&q->tpos, &q->flag, &q->segidx, &q->segstart, &q->segs, &q->qdev, &q->cntN, &q->cntC,
                   &q->ofsN, &q->ofsC, &q->gath, &q->qrec, &q->seedN, &q->seedC, &q->nbuck, &q->cbuck,
                   &q->jseg, &q->cost, &q->soff, &q->pstart, &q->err, &q->tofN, &q->tofC, &q->tmpN,
                   &q->tmpC, &q->used};
    for (DBuf *b : all)
        if (b->p) (void) hipFree(b->p);
    for (auto &e : q->ev) if (e) (void) hipEventDestroy(e);
    pga_destroy(q->Q);
    delete q;
}

int slot_run(pga_server *s, pga_slot *sl, int nq, const uint8_t *const *bps, const int64_t *bps_bytes,
             const int64_t *const *boff, const int64_t *const *clen, const int *ncontig)
{
    pga_ctx *Q = sl->Q, *T = s->T;
    std::vector<int> base(nq, s->nT);
    for (int q = 0; q < nq; q++)
        if ((int64_t) s->nT + ncontig[q] > 32767) return sfail(sl->err_msg, PGA_EOVERFLOW, "query %d: too many contigs", q);
    HIPCHK(Q, hipSetDevice(s->o.device));
    HIPCHK(Q, hipEventRecord(sl->ev[0], Q->st));
    TRY(load_batch_impl(Q, nq, bps, bps_bytes, boff, clen, ncontig, base.data(), T->jbits));
    if (Q->jbits != T->jbits) return sfail(sl->err_msg, PGA_EOVERFLOW, "a query contig is longer than 2^%d bases", T->jbits);
    TRY(pga_build_index(Q));
    const int64_t RQ = Q->R;
    const int W = Q->o.freq + 1;
    sl->nq = nq;
    sl->qd.assign(nq, QDev());
    int64_t nb = 0;
    std::vector<int64_t> qrec(nq + 1);
    for (int q = 0; q < nq; q++) {
        QDev &d = sl->qd[q];
        d.rec0 = Q->pairs[q].rec0;
        d.nrec = Q->pairs[q].nrec;
        d.ncomb = s->nT + ncontig[q];
        const int64_t mc = std::max(s->maxctgT, Q->hp[q].maxctg);
        d.P = bytes_for(mc);
        d.C = bytes_for(2 * (int64_t) d.ncomb - 1);
        d.s = 1 + 2 * (d.P + d.C);
        d.buck_off = nb;
        nb += d.ncomb;
        qrec[q] = d.rec0;
    }
    qrec[nq] = RQ;
    const size_t rq1 = (size_t) std::max<int64_t>(RQ, 1);
    TRY(ensure(Q, sl->tpos, rq1 * sizeof(int64_t)));
    TRY(ensure(Q, sl->flag, rq1 * sizeof(uint32_t)));
    TRY(ensure(Q, sl->segidx, (rq1 + 1) * sizeof(int64_t)));
    TRY(ensure(Q, sl->segstart, rq1 * sizeof(int64_t)));
    TRY(ensure(Q, sl->qdev, (size_t) nq * sizeof(QDev)));
    TRY(ensure(Q, sl->qrec, (size_t) (nq + 1) * sizeof(int64_t)));
    TRY(ensure(Q, sl->gath, (size_t) 2 * (nq + 1) * sizeof(int64_t)));
    HIPCHK(Q, hipMemcpyAsync(sl->qrec.p, qrec.data(), qrec.size() * sizeof(int64_t), hipMemcpyHostToDevice, Q->st));
    HIPCHK(Q, hipMemcpyAsync(sl->qdev.p, sl->qd.data(), (size_t) nq * sizeof(QDev), hipMemcpyHostToDevice, Q->st));
    // the host vectors stay untouched until the stream has consumed them (synchronised below)
    int64_t nseg = 0;
    std::vector<int64_t> seg0(nq + 1, 0);
    const int64_t *tcum = P<int64_t>(T->cum);
    if (RQ > 0) {
        const unsigned g = grid_for(Q, (const void *) k_join_locate, 256, 0, (RQ + 255) / 256);
        hipLaunchKernelGGL(k_join_locate, dim3(g), dim3(256), 0, Q->st, (const Rec *) P<Rec>(T->skeys), tcum,
                           (const Rec *) P<Rec>(Q->skeys), (const uint32_t *) P<uint32_t>(Q->pre), RQ,
                           P<int64_t>(sl->tpos));
        LAUNCHCHK(Q);
        hipLaunchKernelGGL(k_join_flags, dim3(g), dim3(256), 0, Q->st, (const QDev *) P<QDev>(sl->qdev), nq,
                           (const uint32_t *) P<uint32_t>(Q->pre), (const int64_t *) P<int64_t>(sl->tpos), RQ, W,
                           P<uint32_t>(sl->flag));
        LAUNCHCHK(Q);
        TRY(scan_u32(Q, P<uint32_t>(sl->flag), RQ, P<int64_t>(sl->segidx)));
        hipLaunchKernelGGL(k_pair_ranges, dim3((nq + 256) / 256), dim3(256), 0, Q->st,
                           (const int64_t *) P<int64_t>(sl->segidx), (const int64_t *) P<int64_t>(sl->qrec), nq,
                           P<int64_t>(sl->gath));
        LAUNCHCHK(Q);
        HIPCHK(Q, hipMemcpyAsync(seg0.data(), sl->gath.p, (size_t) (nq + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, Q->st));
    }
    HIPCHK(Q, hipStreamSynchronize(Q->st));
    nseg = seg0[nq];
    for (int q = 0; q < nq; q++) sl->qd[q].seg0 = seg0[q];
    std::vector<int64_t> po(2 * (nq + 1), 0), sidx(nq + 1);
    int64_t npack = 0;
    if (nseg > 0) {
        TRY(ensure(Q, sl->segs, (size_t) nseg * sizeof(QSeg)));
        TRY(ensure(Q, sl->cntN, (size_t) nseg * sizeof(uint32_t)));
        TRY(ensure(Q, sl->cntC, (size_t) nseg * sizeof(uint32_t)));
        TRY(ensure(Q, sl->ofsN, (size_t) (nseg + 1) * sizeof(int64_t)));
        TRY(ensure(Q, sl->ofsC, (size_t) (nseg + 1) * sizeof(int64_t)));
        TRY(ensure(Q, sl->jseg, (size_t) nseg * sizeof(JSeg)));
        TRY(ensure(Q, sl->cost, (size_t) nseg * sizeof(uint32_t)));
        TRY(ensure(Q, sl->soff, (size_t) (nseg + 1) * sizeof(int64_t)));
        TRY(ensure(Q, sl->err, 16));
        TRY(fill_u64(Q, sl->err.p, 1, 0));
        HIPCHK(Q, hipMemcpyAsync(sl->qdev.p, sl->qd.data(), (size_t) nq * sizeof(QDev), hipMemcpyHostToDevice, Q->st));
        const unsigned g = grid_for(Q, (const void *) k_join_segs, 256, 0, (RQ + 255) / 256);
        hipLaunchKernelGGL(k_join_segs, dim3(g), dim3(256), 0, Q->st, (const uint32_t *) P<uint32_t>(sl->flag),
                           (const int64_t *) P<int64_t>(sl->segidx), RQ, P<int64_t>(sl->segstart));
        LAUNCHCHK(Q);
        hipLaunchKernelGGL(k_join_segs2, dim3(grid_for(Q, (const void *) k_join_segs2, 256, 0, (nseg + 255) / 256)),
                           dim3(256), 0, Q->st, (const QDev *) P<QDev>(sl->qdev), nq,
                           (const uint32_t *) P<uint32_t>(Q->pre), (const int64_t *) P<int64_t>(sl->tpos), tcum, RQ, W,
                           (const int64_t *) P<int64_t>(sl->segstart), P<QSeg>(sl->segs), nseg);
        LAUNCHCHK(Q);
        // slot ranges of the segments, then packs of consecutive segments
        hipLaunchKernelGGL(k_join_size, dim3(grid_for(Q, (const void *) k_join_size, 256, 0, (nseg + 255) / 256)),
                           dim3(256), 0, Q->st, (const QSeg *) P<QSeg>(sl->segs), nseg, (const QDev *) P<QDev>(sl->qdev),
                           (const uint32_t *) P<uint32_t>(Q->pre), Q->o.freq, P<JSeg>(sl->jseg), P<uint32_t>(sl->cost),
                           P<unsigned int>(sl->err));
        LAUNCHCHK(Q);
        TRY(scan_u32(Q, P<uint32_t>(sl->cost), nseg, P<int64_t>(sl->soff)));
        int64_t head[2] = {0, 0};
        HIPCHK(Q, hipMemcpyAsync(&head[0], P<int64_t>(sl->soff) + nseg, sizeof(int64_t), hipMemcpyDeviceToHost, Q->st));
        HIPCHK(Q, hipMemcpyAsync(&head[1], sl->err.p, sizeof(unsigned int), hipMemcpyDeviceToHost, Q->st));
        HIPCHK(Q, hipStreamSynchronize(Q->st));
        if ((unsigned int) head[1]) return sfail(sl->err_msg, PGA_EINVAL, "query records out of order (%u segments)", (unsigned) head[1]);
        const int64_t hp = PK_CAP - max_segment(Q->o.freq);
        npack = (head[0] + hp - 1) / hp;
        TRY(ensure(Q, sl->pstart, (size_t) (npack + 1) * sizeof(int64_t)));
        hipLaunchKernelGGL(k_join_packs, dim3(grid_for(Q, (const void *) k_join_packs, 256, 0, (npack + 256) / 256)),
                           dim3(256), 0, Q->st, (const int64_t *) P<int64_t>(sl->soff), nseg, hp, npack,
                           P<int64_t>(sl->pstart));
        LAUNCHCHK(Q);
        // one pass per pack: seed counts, bucks and the seeds themselves, in
        // staging streams; a pass that outgrew them is run again with room
        TRY(ensure(Q, sl->nbuck, (size_t) std::max<int64_t>(nb, 1) * sizeof(uint64_t)));
        TRY(ensure(Q, sl->cbuck, (size_t) std::max<int64_t>(nb, 1) * sizeof(uint64_t)));
        TRY(ensure(Q, sl->tofN, (size_t) nseg * sizeof(int64_t)));
        TRY(ensure(Q, sl->tofC, (size_t) nseg * sizeof(int64_t)));
        TRY(ensure(Q, sl->used, 2 * sizeof(uint64_t)));
        // sized from the largest rate seen (a slot's first batch with seeds
        // runs the pass twice: once to learn it)
        sl->capN = std::max<unsigned long long>(sl->capN, (unsigned long long) (sl->rateN * 1.125 * (double) RQ) + 256);
        sl->capC = std::max<unsigned long long>(sl->capC, (unsigned long long) (sl->rateC * 1.125 * (double) RQ) + 256);
        const unsigned gj = grid_for(Q, (const void *) k_join_pack, PK_T, 0, npack);
        for (int attempt = 0;; attempt++) {
            TRY(ensure(Q, sl->tmpN, (size_t) sl->capN + 16));
            TRY(ensure(Q, sl->tmpC, (size_t) sl->capC + 16));
            TRY(fill_u64(Q, sl->nbuck.p, nb, 0));
            TRY(fill_u64(Q, sl->cbuck.p, nb, 0));
            TRY(fill_u64(Q, sl->used.p, 2, 0));
            hipLaunchKernelGGL(k_join_pack, dim3(gj), dim3(PK_T), 0, Q->st, (const int64_t *) P<int64_t>(sl->pstart),
                               npack, (const int64_t *) P<int64_t>(sl->soff), (const JSeg *) P<JSeg>(sl->jseg),
                               (const QSeg *) P<QSeg>(sl->segs), (const QDev *) P<QDev>(sl->qdev),
                               (const Rec *) P<Rec>(T->skeys), (const Rec *) P<Rec>(Q->skeys),
                               (const int64_t *) P<int64_t>(sl->tpos), T->jbits, Q->o.freq, P<uint32_t>(sl->cntN),
                               P<uint32_t>(sl->cntC), P<int64_t>(sl->tofN), P<int64_t>(sl->tofC),
                               P<uint8_t>(sl->tmpN), P<uint8_t>(sl->tmpC), sl->capN, sl->capC,
                               P<unsigned long long>(sl->used), P<unsigned long long>(sl->nbuck),
                               P<unsigned long long>(sl->cbuck));
            LAUNCHCHK(Q);
            unsigned long long u[2] = {0, 0};
            HIPCHK(Q, hipMemcpyAsync(u, sl->used.p, sizeof(u), hipMemcpyDeviceToHost, Q->st));
            HIPCHK(Q, hipStreamSynchronize(Q->st));
            if (RQ > 0) {
                sl->rateN = std::max(sl->rateN, (double) u[0] / (double) RQ);
                sl->rateC = std::max(sl->rateC, (double) u[1] / (double) RQ);
            }
            if (u[0] <= sl->capN && u[1] <= sl->capC) break;
            if (attempt > 0) return sfail(sl->err_msg, PGA_EINVAL, "internal: join staging outgrew its exact size");
            sl->capN = std::max(sl->capN, u[0] + (u[0] >> 3));
            sl->capC = std::max(sl->capC, u[1] + (u[1] >> 3));
        }
        TRY(scan_u32(Q, P<uint32_t>(sl->cntN), nseg, P<int64_t>(sl->ofsN)));
        TRY(scan_u32(Q, P<uint32_t>(sl->cntC), nseg, P<int64_t>(sl->ofsC)));
        // per-query seed totals: the scanned offsets at each query's first segment
        for (int q = 0; q <= nq; q++) sidx[q] = seg0[q];
        HIPCHK(Q, hipMemcpyAsync(sl->qrec.p, sidx.data(), sidx.size() * sizeof(int64_t), hipMemcpyHostToDevice, Q->st));
        hipLaunchKernelGGL(k_pair_ranges, dim3((nq + 256) / 256), dim3(256), 0, Q->st,
                           (const int64_t *) P<int64_t>(sl->ofsN), (const int64_t *) P<int64_t>(sl->qrec), nq,
                           P<int64_t>(sl->gath));
        hipLaunchKernelGGL(k_pair_ranges, dim3((nq + 256) / 256), dim3(256), 0, Q->st,
                           (const int64_t *) P<int64_t>(sl->ofsC), (const int64_t *) P<int64_t>(sl->qrec), nq,
                           P<int64_t>(sl->gath) + (nq + 1));
        LAUNCHCHK(Q);
        HIPCHK(Q, hipMemcpyAsync(po.data(), sl->gath.p, po.size() * sizeof(int64_t), hipMemcpyDeviceToHost, Q->st));
        HIPCHK(Q, hipStreamSynchronize(Q->st));
    }
    sl->nseeds.assign(nq, 0);
    sl->cseeds.assign(nq, 0);
    int64_t nbytes = 0, cbytes = 0;
    for (int q = 0; q < nq; q++) {
        QDev &d = sl->qd[q];
        sl->nseeds[q] = po[q + 1] - po[q];
        sl->cseeds[q] = po[nq + 1 + q + 1] - po[nq + 1 + q];
        d.nseed_off = nbytes;
        d.cseed_off = cbytes;
        nbytes += sl->nseeds[q] * d.s;
        cbytes += sl->cseeds[q] * d.s;
    }
    TRY(ensure(Q, sl->seedN, (size_t) nbytes + 16));
    TRY(ensure(Q, sl->seedC, (size_t) cbytes + 16));
    if (nseg == 0) {                                 // no pass ran: no seeds, zero bucks
        TRY(ensure(Q, sl->nbuck, (size_t) std::max<int64_t>(nb, 1) * sizeof(uint64_t)));
        TRY(ensure(Q, sl->cbuck, (size_t) std::max<int64_t>(nb, 1) * sizeof(uint64_t)));
        TRY(fill_u64(Q, sl->nbuck.p, nb, 0));
        TRY(fill_u64(Q, sl->cbuck.p, nb, 0));
    }
    HIPCHK(Q, hipMemcpyAsync(sl->qdev.p, sl->qd.data(), (size_t) nq * sizeof(QDev), hipMemcpyHostToDevice, Q->st));
    if (nseg > 0) {
        hipLaunchKernelGGL(k_join_copy, dim3(grid_for(Q, (const void *) k_join_copy, 256, 0, (nseg + 255) / 256)),
                           dim3(256), 0, Q->st, (const QSeg *) P<QSeg>(sl->segs), nseg,
                           (const QDev *) P<QDev>(sl->qdev), (const uint32_t *) P<uint32_t>(sl->cntN),
                           (const uint32_t *) P<uint32_t>(sl->cntC), (const int64_t *) P<int64_t>(sl->tofN),
                           (const int64_t *) P<int64_t>(sl->tofC), (const int64_t *) P<int64_t>(sl->ofsN),
                           (const int64_t *) P<int64_t>(sl->ofsC), (const uint8_t *) P<uint8_t>(sl->tmpN),
                           (const uint8_t *) P<uint8_t>(sl->tmpC), P<uint8_t>(sl->seedN), P<uint8_t>(sl->seedC));
        LAUNCHCHK(Q);
    }
    HIPCHK(Q, hipEventRecord(sl->ev[1], Q->st));
    HIPCHK(Q, hipStreamSynchronize(Q->st));
    sl->last_ms = ev_ms(sl->ev[0], sl->ev[1]);
    sl->have_result = true;
    return PGA_OK;
}

pga_slot *get_slot(pga_server *s, int slot)
{
    if (!s || slot < 0 || slot >= (int) s->slots.size()) return nullptr;
    return s->slots[slot];
}

} // namespace

extern "C" {

int pga_server_create(const pga_opts *o, int nslots, pga_server **out)
{
    *out = nullptr;
    if (nslots < 1 || nslots > MAX_SLOTS) return PGA_EINVAL;
    pga_server *s = new (std::nothrow) pga_server();
    if (!s) return PGA_ENOMEM;
    if (o) s->o = *o; else pga_default_opts(&s->o);
    if (s->o.freq > MAX_SERVE_FREQ) { delete s; return PGA_EINVAL; }
    pga_opts to = s->o;
    to.build_cum = 1;
    int rc = pga_create(&to, &s->T);
    if (rc) { delete s; return rc; }
    s->T->keep_keys = true;
    pga_opts qo = s->o;
    qo.build_cum = 0;                  // query groups are located through the target's table
    for (int k = 0; k < nslots && !rc; k++) {
        pga_slot *sl = new (std::nothrow) pga_slot();
        if (!sl) { rc = PGA_ENOMEM; break; }
        s->slots.push_back(sl);
        rc = pga_create(&qo, &sl->Q);
        if (!rc) {
            sl->Q->keep_keys = true;
            for (auto &e : sl->ev)
                if (hipEventCreate(&e) != hipSuccess) rc = PGA_EHIP;
        }
    }
    if (rc) { pga_server_destroy(s); return rc; }
    *out = s;
    return PGA_OK;
}

void pga_server_destroy(pga_server *s)
{
    if (!s) return;
    for (pga_slot *sl : s->slots) slot_free(sl, s->o.device);
    pga_destroy(s->T);
    delete s;
}

const char *pga_server_error(const pga_server *s, int slot)
{
    if (!s) return "no server";
    if (slot >= 0 && slot < (int) s->slots.size()) {
        const pga_slot *sl = s->slots[slot];
        if (sl->err_msg[0]) return sl->err_msg;
        if (sl->Q && sl->Q->err[0]) return sl->Q->err;
        return "";
    }
    if (s->err[0]) return s->err;
    if (s->T && s->T->err[0]) return s->T->err;
    return "";
}

int pga_server_set_target(pga_server *s, const uint8_t *bps, int64_t bps_bytes, const int64_t *boff,
                          const int64_t *clen, int ncontig)
{
    if (!s) return PGA_EINVAL;
    s->err[0] = 0;
    s->T->err[0] = 0;
    s->have_target = false;
    for (pga_slot *sl : s->slots) sl->have_result = false;
    TRY(from_ctx(s->err, s->T, load_batch_impl(s->T, 1, &bps, &bps_bytes, &boff, &clen, &ncontig, nullptr, SERVE_JBITS)));
    TRY(from_ctx(s->err, s->T, pga_build_index(s->T)));
    if (s->T->jbits != SERVE_JBITS) return sfail(s->err, PGA_EOVERFLOW, "a target contig is longer than 2^%d bases", SERVE_JBITS);
    // every slot reads the target from its own stream
    if (hipStreamSynchronize(s->T->st) != hipSuccess) return sfail(s->err, PGA_EHIP, "target build failed");
    s->nT = ncontig;
    s->maxctgT = s->T->hp[0].maxctg;
    s->have_target = true;
    return PGA_OK;
}

int pga_server_set_target_gix(pga_server *s, const uint8_t *entries, int64_t nrec, int Pb, int Cb, const int64_t *cum,
                              int64_t maxpre, const int64_t *clen, int ncontig)
{
    if (!s) return PGA_EINVAL;
    s->err[0] = 0;
    s->have_target = false;
    for (pga_slot *sl : s->slots) sl->have_result = false;
    if (nrec < 0 || (nrec > 0 && !entries) || !cum || !clen || ncontig < 1 || maxpre < 0)
        return sfail(s->err, PGA_EINVAL, "bad GIX target arguments");
    int64_t maxctg = 0, seqtot = 0;
    for (int i = 0; i < ncontig; i++) {
        if (clen[i] < 0) return sfail(s->err, PGA_EINVAL, "contig %d: negative length", i);
        maxctg = std::max(maxctg, clen[i]);
        seqtot += clen[i];
    }
    if (bytes_for(maxctg) != Pb || bytes_for(2 * (int64_t) ncontig - 1) != Cb)
        return sfail(s->err, PGA_EINVAL, "the index's P=%d C=%d do not fit %d contigs of at most %lld bases", Pb, Cb,
                     ncontig, (long long) maxctg);
    if (cum[NPRE - 1] != nrec) return sfail(s->err, PGA_EINVAL, "prefix counts end at %lld, not %lld entries",
                                            (long long) cum[NPRE - 1], (long long) nrec);
    if (bits_for(maxctg) > SERVE_JBITS) return sfail(s->err, PGA_EOVERFLOW, "a target contig is longer than 2^%d bases", SERVE_JBITS);
    if (ncontig > 32767) return sfail(s->err, PGA_EOVERFLOW, "too many target contigs");
    pga_ctx *T = s->T;
    T->err[0] = 0;
    if (hipSetDevice(s->o.device) != hipSuccess) return sfail(s->err, PGA_EHIP, "hipSetDevice failed");
    T->loaded = T->built = T->merged = T->ss_ready = false;
    T->graph_gen = ~0ull;
    T->npairs = 1;
    T->hp.assign(1, HostPair());
    T->hp[0].ncontig = ncontig;
    T->hp[0].seqtot = seqtot;
    T->hp[0].maxctg = maxctg;
    T->nctg = ncontig;
    T->jbits = SERVE_JBITS;
    T->R = nrec;
    T->pairs.assign(1, PairDev());
    PairDev &d = T->pairs[0];
    d.rec0 = 0;
    d.nrec = nrec;
    d.P = Pb;
    d.C = Cb;
    d.e = 9 + Pb + Cb;
    d.s = 1 + 2 * (Pb + Cb);
    d.ncontig = ncontig;
    const size_t n1 = (size_t) std::max<int64_t>(nrec, 1);
    TRY(from_ctx(s->err, T, ensure(T, T->entries, n1 * d.e + 32)));
    TRY(from_ctx(s->err, T, ensure(T, T->skeys, n1 * sizeof(Rec))));
    TRY(from_ctx(s->err, T, ensure(T, T->pre, (n1 + 1) * sizeof(uint32_t))));
    TRY(from_ctx(s->err, T, ensure(T, T->ls, n1 + 64)));
    TRY(from_ctx(s->err, T, ensure(T, T->cum, (size_t) NPRE * sizeof(int64_t))));
    TRY(from_ctx(s->err, T, ensure(T, T->maxpre, sizeof(uint64_t))));
    TRY(from_ctx(s->err, T, ensure(T, T->dpairs, sizeof(PairDev))));
    const uint64_t mp = (uint64_t) maxpre;
    HIPCHK(T, hipMemcpyAsync(T->cum.p, cum, (size_t) NPRE * sizeof(int64_t), hipMemcpyHostToDevice, T->st));
    HIPCHK(T, hipMemcpyAsync(T->maxpre.p, &mp, sizeof(mp), hipMemcpyHostToDevice, T->st));
    HIPCHK(T, hipMemcpyAsync(T->dpairs.p, &d, sizeof(PairDev), hipMemcpyHostToDevice, T->st));
    if (nrec > 0) {
        HIPCHK(T, hipMemcpyAsync(T->entries.p, entries, (size_t) (nrec * d.e), hipMemcpyHostToDevice, T->st));
        hipLaunchKernelGGL(k_gix_prefix, dim3(grid_for(T, (const void *) k_gix_prefix, 256, 0, NPRE / 256)), dim3(256), 0,
                           T->st, (const int64_t *) P<int64_t>(T->cum), P<uint32_t>(T->pre));
        hipLaunchKernelGGL(k_gix_keys, dim3(grid_for(T, (const void *) k_gix_keys, 256, 0, (nrec + 255) / 256)), dim3(256),
                           0, T->st, (const uint8_t *) P<uint8_t>(T->entries), d.e, Pb, Cb,
                           (const uint32_t *) P<uint32_t>(T->pre), nrec, T->jbits, P<Rec>(T->skeys), P<uint8_t>(T->ls));
        if (hipGetLastError() != hipSuccess) return sfail(s->err, PGA_EHIP, "kernel launch failed");
    }
    if (hipStreamSynchronize(T->st) != hipSuccess) return sfail(s->err, PGA_EHIP, "target load failed");
    T->built = true;
    s->nT = ncontig;
    s->maxctgT = maxctg;
    s->have_target = true;
    return PGA_OK;
}

int pga_server_target_info(pga_server *s, pga_index_meta *info)
{
    if (!s) return PGA_EINVAL;
    if (!s->have_target) return sfail(s->err, PGA_EINVAL, "no target");
    return from_ctx(s->err, s->T, pga_index_info(s->T, 0, info));
}

int pga_server_footprint(pga_server *s, int64_t query_bases, int64_t *bytes)
{
    if (!s || query_bases < 0 || !bytes) return PGA_EINVAL;
    // Per base: at most 2 records (one per strand) in practice 0.8; per record two
    // 16 B partition buffers, 16 B kept keys, e <= 18 B entry, LCP, prefix, join
    // position and flags (~70 B); seeds (target anchors pair with query
    // entries: ~6 seeds a query base at FREQ 10, s = 11 B) twice, in the
    // join's staging streams and in the answer.
    *bytes = query_bases * 200 + ((int64_t) 64 << 20);
    return PGA_OK;
}

int pga_server_slots(pga_server *s) { return s ? (int) s->slots.size() : 0; }

int pga_server_run(pga_server *s, int slot, int nq, const uint8_t *const *bps, const int64_t *bps_bytes,
                   const int64_t *const *boff, const int64_t *const *clen, const int *ncontig)
{
    pga_slot *sl = get_slot(s, slot);
    if (!sl) return PGA_EINVAL;
    sl->err_msg[0] = 0;
    sl->Q->err[0] = 0;
    sl->have_result = false;
    if (!s->have_target) return sfail(sl->err_msg, PGA_EINVAL, "no target");
    if (nq <= 0) return sfail(sl->err_msg, PGA_EINVAL, "no queries");
    return slot_run(s, sl, nq, bps, bps_bytes, boff, clen, ncontig);
}

int pga_server_seeds_info(pga_server *s, int slot, int q, pga_seeds_meta *info)
{
    pga_slot *sl = get_slot(s, slot);
    if (!sl) return PGA_EINVAL;
    if (!sl->have_result || q < 0 || q >= sl->nq) return sfail(sl->err_msg, PGA_EINVAL, "no result for query %d", q);
    info->nlen = sl->nseeds[q] * sl->qd[q].s;
    info->clen = sl->cseeds[q] * sl->qd[q].s;
    info->s = sl->qd[q].s;
    info->ncontig = sl->qd[q].ncomb;
    return PGA_OK;
}

int pga_server_seeds_export(pga_server *s, int slot, int q, uint8_t *nseeds, uint8_t *cseeds, int64_t *nbuck,
                            int64_t *cbuck)
{
    pga_slot *sl = get_slot(s, slot);
    if (!sl) return PGA_EINVAL;
    if (!sl->have_result || q < 0 || q >= sl->nq) return sfail(sl->err_msg, PGA_EINVAL, "no result for query %d", q);
    pga_ctx *Q = sl->Q;
    const QDev &d = sl->qd[q];
    HIPCHK(Q, hipSetDevice(s->o.device));
    if (nseeds && sl->nseeds[q] > 0)
        HIPCHK(Q, hipMemcpyAsync(nseeds, P<uint8_t>(sl->seedN) + d.nseed_off, (size_t) (sl->nseeds[q] * d.s),
                                 hipMemcpyDeviceToHost, Q->st));
    if (cseeds && sl->cseeds[q] > 0)
        HIPCHK(Q, hipMemcpyAsync(cseeds, P<uint8_t>(sl->seedC) + d.cseed_off, (size_t) (sl->cseeds[q] * d.s),
                                 hipMemcpyDeviceToHost, Q->st));
    if (nbuck)
        HIPCHK(Q, hipMemcpyAsync(nbuck, P<uint64_t>(sl->nbuck) + d.buck_off, (size_t) d.ncomb * 8,
                                 hipMemcpyDeviceToHost, Q->st));
    if (cbuck)
        HIPCHK(Q, hipMemcpyAsync(cbuck, P<uint64_t>(sl->cbuck) + d.buck_off, (size_t) d.ncomb * 8,
                                 hipMemcpyDeviceToHost, Q->st));
    HIPCHK(Q, hipStreamSynchronize(Q->st));
    return PGA_OK;
}

int pga_server_last_ms(pga_server *s, int slot, double *ms)
{
    pga_slot *sl = get_slot(s, slot);
    if (!sl || !ms) return PGA_EINVAL;
    *ms = sl->last_ms;
    return PGA_OK;
}

} // extern "C"
