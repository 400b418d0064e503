// pga_serve.hip -- serving kernels: queries merged against a resident target index.
//
// A query Q's seeds are those of the self-merge of the GDB (target contigs,
// then Q's contigs) (src/fastga_pipeline.c:245-412) that have at least one
// Q-side entry (anchor or partner).  Seeds never leave a 24-bit prefix group
// and an emitted anchor's partners lie within FREQ-1 positions of it, so only
// target entries near a Q entry matter.  Per query:
//   1. every Q record finds its insertion position in the target's sorted keys
//      (its 24-bit group range from the target's prefix table, then a binary
//      search inside the group);
//   2. Q records are cut into segments whose cores (target entries within W of
//      a Q record, plus the Q records) are disjoint and ordered;
//   3. consecutive segments are packed into workgroup-sized packs; each
//      segment's target and Q keys (core plus FREQ+2 entries of context on
//      each side, or up to the group edge) are merged in LDS, LCPs recomputed,
//      and every core anchor's interval evaluated exactly as in the index merge;
//      seeds with a Q side are counted, then emitted in anchor order.
// Device keys of target and queries share one encoding (contig ids of a query
// start after the target's; one offset width), so they compare directly.

namespace pga {

constexpr int SEG_Q = 16;              // Q records per segment at most
constexpr int MAX_SERVE_FREQ = 24;     // largest FREQ a server accepts

struct QSeg {
    int64_t ct_lo, ct_hi;              // core target range
    int64_t cq_lo, cq_hi;              // core query range (indices in the query batch)
    int64_t glo, ghi;                  // the target's 24-bit group
    int32_t q, pad;
};

struct QDev {
    int64_t rec0, nrec;                // the query's records in the batch
    int64_t nseed_off, cseed_off;      // byte offsets of its seed streams
    int64_t seg0;                      // its first segment
    int64_t buck_off;                  // offset of its (target + query) contigs in the buck vectors
    int32_t P, C, s, ncomb;            // payload widths of the combined GDB; its contig count
};

__device__ __forceinline__ bool key_lt(const Rec &a, const Rec &b)
{
    return a.w0 < b.w0 || (a.w0 == b.w0 && a.w1 < b.w1);
}

__device__ __forceinline__ int query_of(const QDev *__restrict__ qd, int nq, int64_t g)
{
    int lo = 0, hi = nq;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (qd[mid].rec0 <= g) lo = mid; else hi = mid;
    }
    return lo;
}

// 1. Insertion position of every query record among the target's keys.
__global__ void k_join_locate(const Rec *__restrict__ tkeys, const int64_t *__restrict__ tcum,
                              const Rec *__restrict__ qkeys, const uint32_t *__restrict__ qpre, int64_t RQ,
                              int64_t *__restrict__ tpos)
{
    for (int64_t g = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; g < RQ;
         g += (int64_t) gridDim.x * blockDim.x) {
        const Rec k = qkeys[g];
        const uint32_t p = qpre[g];
        int64_t lo = p ? tcum[p - 1] : 0, hi = tcum[p];
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (key_lt(tkeys[mid], k)) lo = mid + 1; else hi = mid;
        }
        tpos[g] = lo;
    }
}

__device__ __forceinline__ bool natural_start(const QDev &d, const uint32_t *qpre, const int64_t *tpos,
                                              int64_t g, int W)
{
    return g == d.rec0 || qpre[g] != qpre[g - 1] || tpos[g] - W >= tpos[g - 1] + W;
}

// 2a. Segment start flags: a new group, non-overlapping windows, or SEG_Q records.
__global__ void k_join_flags(const QDev *__restrict__ qd, int nq, const uint32_t *__restrict__ qpre,
                             const int64_t *__restrict__ tpos, int64_t RQ, int W, uint32_t *__restrict__ flag)
{
    for (int64_t g = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; g < RQ;
         g += (int64_t) gridDim.x * blockDim.x) {
        const QDev d = qd[query_of(qd, nq, g)];
        flag[g] = natural_start(d, qpre, tpos, g, W) || ((g - d.rec0) % SEG_Q) == 0;
    }
}

// 2b. Segment starts from the scanned flags, then the descriptors.
__global__ void k_join_segs(const uint32_t *__restrict__ flag, const int64_t *__restrict__ segidx, int64_t RQ,
                            int64_t *__restrict__ segstart)
{
    for (int64_t g = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; g < RQ;
         g += (int64_t) gridDim.x * blockDim.x)
        if (flag[g]) segstart[segidx[g]] = g;
}

__global__ void k_join_segs2(const QDev *__restrict__ qd, int nq, const uint32_t *__restrict__ qpre,
                             const int64_t *__restrict__ tpos, const int64_t *__restrict__ tcum, int64_t RQ,
                             int W, const int64_t *__restrict__ segstart, QSeg *__restrict__ segs, int64_t nseg)
{
    for (int64_t s = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; s < nseg;
         s += (int64_t) gridDim.x * blockDim.x) {
        const int64_t g0 = segstart[s];
        const int64_t g1 = s + 1 < nseg ? segstart[s + 1] : RQ;
        const int q = query_of(qd, nq, g0);
        const QDev d = qd[q];
        const uint32_t p = qpre[g0];
        const int64_t glo = p ? tcum[p - 1] : 0, ghi = tcum[p];
        const bool forced_start = !natural_start(d, qpre, tpos, g0, W);
        const bool forced_end = g1 < d.rec0 + d.nrec && !natural_start(d, qpre, tpos, g1, W);
        QSeg sg;
        sg.ct_lo = forced_start ? tpos[g0] : max(glo, tpos[g0] - W);
        sg.ct_hi = forced_end ? tpos[g1] : min(ghi, tpos[g1 - 1] + W);
        sg.cq_lo = g0;
        sg.cq_hi = g1;
        sg.glo = glo;
        sg.ghi = ghi;
        sg.q = q;
        sg.pad = 0;
        segs[s] = sg;
    }
}

// Payload (pos LE in P bytes, contig|strand LE in C bytes) of a device key.
__device__ __forceinline__ uint64_t key_payload(const Rec &k, int jbits, int P, int C)
{
    const uint64_t strand = k.w1 & 1ull;
    const uint64_t j = (k.w1 >> 1) & ((1ull << jbits) - 1ull);
    const uint64_t contig = (k.w1 & 0xFFFFFFFFFFFFull) >> (jbits + 1);
    const uint64_t pos = j + 12ull * strand;
    const uint64_t cs = contig | (strand << (8 * C - 1));
    return pos | (cs << (8 * P));
}

// ---------------------------------------------------------------- packed join
//
// Most segments hold one or two query records and a few dozen target entries,
// so one workgroup works on a "pack" of consecutive segments at a time: up to
// PK_CAP combined entries in LDS, each segment merged in place inside its own
// slot range.  Segment slot ranges are padded to at least PK_MINC entries, so a
// pack holds at most PK_CAP / PK_MINC segments.  Packs cut the running sum of
// slot sizes at multiples of Hp = PK_CAP - (largest segment): pack p takes the
// segments whose slot range starts in [p*Hp, (p+1)*Hp).

constexpr int PK_CAP = 2048;           // combined entries of one pack
constexpr int PK_MINC = 8;             // slots per segment at least
constexpr int PK_SEG = PK_CAP / PK_MINC;
constexpr int PK_T = 512;              // threads of the pack kernels
constexpr int PK_PER = PK_CAP / PK_T;  // anchor slots per thread
constexpr int PK_BUCK = 1024;          // anchor contigs counted in LDS (the pack's first query)
constexpr int PK_MAP = 1024;           // seeds of each stream an emit round maps in LDS
static_assert(PK_CAP <= 2048, "slot indexes fit 11 bits of a map word");
static_assert(PK_CAP / PK_MINC <= PK_T, "one thread per segment of a pack");

// Largest segment in combined entries for a FREQ: 33W core target entries,
// FREQ+2 of context each side, SEG_Q + 2(FREQ+2) query records.
__host__ __device__ constexpr int max_segment(int freq)
{
    return 33 * (freq + 1) + 4 * (freq + 2) + SEG_Q;
}
static_assert(max_segment(MAX_SERVE_FREQ) < PK_CAP / 2, "pack too small for the largest segment");

// Loaded ranges of one segment (relative offsets fit 16 bits: a segment is
// smaller than PK_CAP).
struct JSeg {
    int64_t tlo, qlo;                  // first loaded target entry / query record
    uint16_t nt, nq;                   // loaded counts
    uint16_t ct0, ct1;                 // core target range, relative to tlo
    uint16_t cq0, cq1;                 // core query range, relative to qlo
    uint16_t trunc;                    // bit0 t_lo, bit1 t_hi, bit2 q_lo, bit3 q_hi truncated
    uint16_t pad;
};

// S1. Loaded ranges and slot cost of every segment.
__global__ void k_join_size(const QSeg *__restrict__ segs, int64_t nseg, const QDev *__restrict__ qd,
                            const uint32_t *__restrict__ qpre, int freq, JSeg *__restrict__ js,
                            uint32_t *__restrict__ cost, unsigned int *__restrict__ err)
{
    const int MG = freq + 2;
    const uint32_t cap = (uint32_t) max_segment(freq);
    for (int64_t s = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; s < nseg;
         s += (int64_t) gridDim.x * blockDim.x) {
        const QSeg sg = segs[s];
        const QDev d = qd[sg.q];
        const uint32_t p = qpre[sg.cq_lo];
        const int64_t tlo = max(sg.glo, sg.ct_lo - MG), thi = min(sg.ghi, sg.ct_hi + MG);
        int64_t qlo = max(d.rec0, sg.cq_lo - MG), qhi = min(d.rec0 + d.nrec, sg.cq_hi + MG);
        while (qlo < sg.cq_lo && qpre[qlo] != p) qlo++;
        while (qhi > sg.cq_hi && qpre[qhi - 1] != p) qhi--;
        JSeg j;
        j.tlo = tlo;
        j.qlo = qlo;
        j.nt = (uint16_t) (thi - tlo);
        j.nq = (uint16_t) (qhi - qlo);
        j.ct0 = (uint16_t) (sg.ct_lo - tlo);
        j.ct1 = (uint16_t) (sg.ct_hi - tlo);
        j.cq0 = (uint16_t) (sg.cq_lo - qlo);
        j.cq1 = (uint16_t) (sg.cq_hi - qlo);
        j.trunc = (uint16_t) ((tlo > sg.glo ? 1 : 0) | (thi < sg.ghi ? 2 : 0) |
                              ((qlo > d.rec0 && qpre[qlo - 1] == p) ? 4 : 0) |
                              ((qhi < d.rec0 + d.nrec && qpre[qhi] == p) ? 8 : 0));
        j.pad = 0;
        js[s] = j;
        uint32_t n = (uint32_t) (thi - tlo + qhi - qlo);
        if (n > cap) {                     // cannot happen for sorted inputs; never overrun LDS
            atomicAdd(err, 1u);
            n = cap;
            j.nt = 0;
            j.nq = 0;
            js[s] = j;
        }
        cost[s] = n > (uint32_t) PK_MINC ? n : (uint32_t) PK_MINC;
    }
}

// S2. First segment of every pack (segments whose slot start lies in [p*Hp, (p+1)*Hp)).
__global__ void k_join_packs(const int64_t *__restrict__ soff, int64_t nseg, int64_t hp, int64_t npack,
                             int64_t *__restrict__ pstart)
{
    for (int64_t p = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; p <= npack;
         p += (int64_t) gridDim.x * blockDim.x) {
        if (p == npack) { pstart[p] = nseg; continue; }
        const int64_t x = p * hp;
        int64_t lo = 0, hi = nseg;
        while (lo < hi) {
            const int64_t m = (lo + hi) >> 1;
            if (soff[m] < x) lo = m + 1; else hi = m;
        }
        pstart[p] = lo;
    }
}

struct PackLds {
    Rec key[PK_CAP];
    uint8_t win[PK_CAP + 2 * (MAX_SERVE_FREQ + 1) + 8];
    uint8_t flg[PK_CAP];               // bit0 query record, bit1 core anchor
    uint8_t sidx[PK_CAP];              // segment of each slot
    uint16_t qtp[PK_CAP];              // query slots: insertion point relative to the segment's tlo
    // per segment of the pack
    int64_t tlo[PK_SEG], qlo[PK_SEG];
    uint32_t sN0[PK_SEG], sC0[PK_SEG]; // pack-local number of the segment's first N / C seed
    uint32_t bN0[PK_SEG], bC0[PK_SEG]; // pack-local byte offset of the segment's N / C seeds
    uint16_t so[PK_SEG + 1], nt[PK_SEG], nq[PK_SEG], lo[PK_SEG], hi[PK_SEG];
    uint16_t ct0[PK_SEG], ct1[PK_SEG], cq0[PK_SEG], cq1[PK_SEG];
    int32_t q[PK_SEG];
    uint32_t cN[PK_SEG], cC[PK_SEG];   // seeds per segment (count pass)
    uint8_t sbytes[PK_SEG], Pb[PK_SEG], Cb[PK_SEG];
    uint32_t bk[2][PK_BUCK];           // buck counters of the pack's first query (emit pass)
    uint32_t smap[2][PK_MAP];          // the emit round's N and C seeds: anchor slot | partner slot << 11 | plen << 22
    uint32_t scan[PK_T / 64 + 1];
    uint64_t scanw[PK_T / 64 + 1];
    int ns, npos, qa, ovf;
    unsigned long long baseN, baseC;   // the pack's byte ranges in the staging streams
};

__device__ __forceinline__ int pack_seg_of(const PackLds &L, int pos)
{
    int lo = 0, hi = L.ns;             // largest t with so[t] <= pos
    while (hi - lo > 1) {
        const int m = (lo + hi) >> 1;
        if (L.so[m] <= pos) lo = m; else hi = m;
    }
    return lo;
}

// Load pack p into LDS: segments, merged keys, flags, valid windows, LCP/strand window.
__device__ void pack_load(PackLds &L, int64_t p, const int64_t *__restrict__ pstart,
                          const int64_t *__restrict__ soff, const JSeg *__restrict__ js,
                          const QSeg *__restrict__ segs, const QDev *__restrict__ qd,
                          const Rec *__restrict__ tkeys, const Rec *__restrict__ qkeys,
                          const int64_t *__restrict__ tpos, int H)
{
    const int64_t s0 = pstart[p], s1 = pstart[p + 1];
    const int ns = (int) (s1 - s0);
    const int64_t base = s0 < s1 ? soff[s0] : 0;
    for (int t = threadIdx.x; t < ns; t += PK_T) {
        const JSeg j = js[s0 + t];
        const int q = segs[s0 + t].q;
        L.so[t] = (uint16_t) (soff[s0 + t] - base);
        L.tlo[t] = j.tlo;
        L.qlo[t] = j.qlo;
        L.nt[t] = j.nt;
        L.nq[t] = j.nq;
        L.ct0[t] = j.ct0;
        L.ct1[t] = j.ct1;
        L.cq0[t] = j.cq0;
        L.cq1[t] = j.cq1;
        L.q[t] = q;
        L.cN[t] = 0;
        L.cC[t] = 0;
        const QDev d = qd[q];
        L.sbytes[t] = (uint8_t) d.s;
        L.Pb[t] = (uint8_t) d.P;
        L.Cb[t] = (uint8_t) d.C;
        // valid window needs the query insertion points: filled below
        L.lo[t] = (uint16_t) j.trunc;            // trunc bits parked here until then
    }
    if (threadIdx.x == 0) {
        L.ns = ns;
        L.npos = ns > 0 ? (int) (soff[s1] - base) : 0;
        L.so[ns] = (uint16_t) L.npos;
        L.qa = ns > 0 ? segs[s0].q : -1;
    }
    __syncthreads();
    const int npos = L.npos;
    // query slots: insertion points relative to tlo, clipped to [0, nt]
    for (int pos = threadIdx.x; pos < npos; pos += PK_T) {
        const int s = pack_seg_of(L, pos);
        L.sidx[pos] = (uint8_t) s;
        L.flg[pos] = 0;
        const int k = pos - L.so[s], nt = L.nt[s];
        if (k >= nt && k < nt + L.nq[s]) {
            int64_t r = tpos[L.qlo[s] + (k - nt)] - L.tlo[s];
            r = r < 0 ? 0 : (r > nt ? nt : r);
            L.qtp[pos] = (uint16_t) r;
        }
    }
    __syncthreads();
    // place target and query keys at their merged slot inside the segment
    for (int pos = threadIdx.x; pos < npos; pos += PK_T) {
        const int s = L.sidx[pos];
        const int b = L.so[s], k = pos - b, nt = L.nt[s], nq = L.nq[s];
        if (k < nt) {
            int a = 0, e = nq;                           // query records placed before T[tlo+k]
            while (a < e) { const int m = (a + e) >> 1; if (L.qtp[b + nt + m] <= k) a = m + 1; else e = m; }
            L.key[b + k + a] = tkeys[L.tlo[s] + k];
            L.flg[b + k + a] = (k >= L.ct0[s] && k < L.ct1[s]) ? 2 : 0;
        } else if (k < nt + nq) {
            const int i = k - nt;
            const int m = L.qtp[pos] + i;
            L.key[b + m] = qkeys[L.qlo[s] + i];
            L.flg[b + m] = (uint8_t) (1 | ((i >= L.cq0[s] && i < L.cq1[s]) ? 2 : 0));
        }
    }
    // valid window of each segment, from its truncation bits and query insertion points
    for (int t = threadIdx.x; t < ns; t += PK_T) {
        const int tr = L.lo[t];
        const int b = L.so[t], nt = L.nt[t], nq = L.nq[t], n = nt + nq;
        const uint16_t *qt = L.qtp + b + nt;
        int lo = 0, hi = n - 1;
        if ((tr & 1) && nq) {                        // query records at or before tlo come first
            int a = 0, e = nq;
            while (a < e) { const int m = (a + e) >> 1; if (qt[m] <= 0) a = m + 1; else e = m; }
            lo = max(lo, a);
        }
        if ((tr & 4) && nq) lo = max(lo, (int) qt[0]);
        if ((tr & 2) && nt) {
            int a = 0, e = nq;
            while (a < e) { const int m = (a + e) >> 1; if (qt[m] <= nt - 1) a = m + 1; else e = m; }
            hi = min(hi, nt - 1 + a);
        }
        if ((tr & 8) && nq) hi = min(hi, (int) qt[nq - 1] + nq - 1);
        L.lo[t] = (uint16_t) lo;
        L.hi[t] = (uint16_t) (hi < lo ? 0 : hi);
        if (hi < lo) L.lo[t] = 1;                    // empty window: lo > hi
    }
    for (int i = threadIdx.x; i < H; i += PK_T) { L.win[i] = 0; L.win[H + npos + i] = 0; }
    if (threadIdx.x == 0) { L.win[2 * H + npos] = 0; L.win[2 * H + npos + 1] = 0; }
    __syncthreads();
    // LCP | strand<<7 inside each valid window, 0 elsewhere
    for (int pos = threadIdx.x; pos < npos; pos += PK_T) {
        const int s = L.sidx[pos];
        const int i = pos - L.so[s];
        uint8_t w = 0;
        if (i >= L.lo[s] && i <= L.hi[s] && i < L.nt[s] + L.nq[s]) {
            const Rec k = L.key[pos];
            int l = 0;
            if (i > L.lo[s]) { const Rec pk = L.key[pos - 1]; l = lcp_bases(pk.w0, pk.w1, k.w0, k.w1); }
            w = (uint8_t) (l | ((k.w1 & 1ull) << 7));
        } else {
            L.flg[pos] = 0;                          // outside the window: never an anchor
        }
        L.win[H + pos] = w;
    }
    __syncthreads();
}

// S3. One pass per pack: seeds counted per segment and written, in stream
// order, to the pack's own byte ranges of two staging streams (reserved with
// one atomic each), with per-anchor-contig counts.  A pack whose range would
// pass a staging stream's capacity writes no seeds; the host sees the
// reserved total above the capacity, grows the streams and runs the pass
// again.  k_join_copy then moves every segment's seeds to their place.
__global__ __launch_bounds__(PK_T, 4) void k_join_pack(const int64_t *__restrict__ pstart, int64_t npack,
                                                    const int64_t *__restrict__ soff, const JSeg *__restrict__ js,
                                                    const QSeg *__restrict__ segs, const QDev *__restrict__ qd,
                                                    const Rec *__restrict__ tkeys, const Rec *__restrict__ qkeys,
                                                    const int64_t *__restrict__ tpos, int jbits, int freq,
                                                    uint32_t *__restrict__ cntN, uint32_t *__restrict__ cntC,
                                                    int64_t *__restrict__ tofN, int64_t *__restrict__ tofC,
                                                    uint8_t *__restrict__ tmpN, uint8_t *__restrict__ tmpC,
                                                    unsigned long long capN, unsigned long long capC,
                                                    unsigned long long *__restrict__ used,
                                                    unsigned long long *__restrict__ nbuck,
                                                    unsigned long long *__restrict__ cbuck)
{
    __shared__ PackLds L;
    // FREQ <= 16: loop-free interval search, which reads 15 window bytes before
    // an anchor and 16 after it, so the zeroed halo is 17 bytes
    const bool swar = freq <= 16;
    const int H = swar ? 17 : freq + 1;
    for (int64_t p = blockIdx.x; p < npack; p += gridDim.x) {
        pack_load(L, p, pstart, soff, js, segs, qd, tkeys, qkeys, tpos, H);
        for (int c = threadIdx.x; c < PK_BUCK; c += PK_T) { L.bk[0][c] = 0; L.bk[1][c] = 0; }
        const int npos = L.npos;
        const int pos0 = threadIdx.x * PK_PER;
        const int qa = L.qa;
        uint32_t tn = 0, tc = 0;
        uint32_t mE = 0;                     // anchors of this thread: bit k set when slot k emits
        // the anchors' intervals [ia, ib) and plen, kept for the emit
        int ia[PK_PER], ib[PK_PER], ipl[PK_PER];
        __syncthreads();                     // bk cleared
#pragma unroll
        for (int k = 0; k < PK_PER; k++) {
            ia[k] = ib[k] = ipl[k] = 0;
            const int pos = pos0 + k;
            if (pos >= npos || !(L.flg[pos] & 2)) continue;
            const int gi = H + pos;
            const bool iq = L.flg[pos] & 1;
            uint32_t ln = 0, lc = 0;
            if (swar) {
                // a query anchor pairs with every entry of its interval, a
                // target anchor only with query entries
                const uint32_t info = iq ? anchor_info16(L.win, gi, freq, ln, lc)
                                         : anchor_info16_q(L.win, L.flg - H, gi, freq, ln, lc);
                ipl[k] = (int) (info & 63u);
                ia[k] = gi - (int) ((info >> 6) & 63u);
                ib[k] = gi + (int) ((info >> 12) & 63u);
            } else {
                int plen, a, b;
                if (!anchor_interval(L.win, gi, freq, plen, a, b)) continue;
                const uint8_t si = L.win[gi] & 0x80;
                for (int x = a; x < b; x++) {
                    if (x == gi) continue;
                    if (!iq && !(L.flg[x - H] & 1)) continue;
                    if ((L.win[x] & 0x80) == si) ln++; else lc++;
                }
                ipl[k] = plen;
                ia[k] = a;
                ib[k] = b;
            }
            if (!(ln | lc)) continue;
            tn += ln;
            tc += lc;
            mE |= 1u << k;
            const int s = L.sidx[pos];
            if (ln) atomicAdd(&L.cN[s], ln);
            if (lc) atomicAdd(&L.cC[s], lc);
            // per-anchor-contig counts (src/fastga_pipeline.c:403)
            const int P = L.Pb[s], Cb = L.Cb[s];
            const uint64_t pi = key_payload(L.key[pos], jbits, P, Cb);
            const int icont = (int) ((pi >> (8 * P)) & ((1ull << (8 * Cb - 1)) - 1));
            if (L.q[s] == qa && icont < PK_BUCK) {
                if (ln) atomicAdd(&L.bk[0][icont], ln);
                if (lc) atomicAdd(&L.bk[1][icont], lc);
            } else {
                const int64_t bo = qd[L.q[s]].buck_off + icont;
                if (ln) atomicAdd(&nbuck[bo], (unsigned long long) ln);
                if (lc) atomicAdd(&cbuck[bo], (unsigned long long) lc);
            }
        }
        // pack-local seed numbers of this thread's first N / C seed (seeds are
        // numbered in stream order: segments, anchors, partners ascending)
        uint32_t totN, totC;
        const uint32_t rn = block_excl_scan<PK_T, uint32_t>(tn, L.scan, &totN);
        const uint32_t rc = block_excl_scan<PK_T, uint32_t>(tc, L.scan, &totC);
        // the segments' first seeds and byte offsets inside the pack (thread t:
        // segment t; a pack holds at most PK_SEG <= PK_T segments)
        const int ns = L.ns;
        const int t = threadIdx.x;
        const uint32_t cn = t < ns ? L.cN[t] : 0u, cc = t < ns ? L.cC[t] : 0u;
        const uint32_t sb = t < ns ? L.sbytes[t] : 0u;
        uint64_t totb;
        const uint64_t on = block_excl_scan<PK_T, uint64_t>((uint64_t) cn | (uint64_t) (cn * sb) << 32, L.scanw, &totb);
        const uint64_t totNb = totb >> 32;
        const uint64_t oc = block_excl_scan<PK_T, uint64_t>((uint64_t) cc | (uint64_t) (cc * sb) << 32, L.scanw, &totb);
        const uint64_t totCb = totb >> 32;
        if (t < ns) {
            L.sN0[t] = (uint32_t) on;
            L.bN0[t] = (uint32_t) (on >> 32);
            L.sC0[t] = (uint32_t) oc;
            L.bC0[t] = (uint32_t) (oc >> 32);
        }
        if (t == 0) {
            const unsigned long long bn = totNb ? atomicAdd(&used[0], (unsigned long long) totNb) : 0ull;
            const unsigned long long bc = totCb ? atomicAdd(&used[1], (unsigned long long) totCb) : 0ull;
            L.baseN = bn;
            L.baseC = bc;
            L.ovf = bn + totNb > capN || bc + totCb > capC;
        }
        __syncthreads();
        const int64_t s0 = pstart[p];
        if (t < ns) {
            cntN[s0 + t] = cn;
            cntC[s0 + t] = cc;
            tofN[s0 + t] = (int64_t) (L.baseN + L.bN0[t]);
            tofC[s0 + t] = (int64_t) (L.baseC + L.bC0[t]);
        }
        // Rounds of up to PK_MAP seeds a stream: every anchor lists its seeds
        // in an LDS map (anchor slot | partner slot << 11 | plen << 22), then
        // one lane takes one seed, so a wave's stores cover a contiguous run.
        const int rounds = L.ovf ? 0 : (int) ((max(totN, totC) + PK_MAP - 1) / PK_MAP);
        for (int r = 0; r < rounds; r++) {
            const int n0 = r * PK_MAP, n1 = max(n0, min(n0 + PK_MAP, (int) totN));
            const int c0 = r * PK_MAP, c1 = max(c0, min(c0 + PK_MAP, (int) totC));
            int kn = (int) rn, kc = (int) rc;
#pragma unroll 1
            for (int k = 0; k < PK_PER; k++) {
                if (!((mE >> k) & 1u)) continue;
                const int pos = pos0 + k;
                const int gi = H + pos;
                const bool iq = L.flg[pos] & 1;
                const uint8_t si = L.win[gi] & 0x80;
                const uint32_t tag = (uint32_t) pos | (uint32_t) ipl[k] << 22;
                for (int x = ia[k]; x < ib[k]; x++) {
                    const int j = x - H;
                    if (x == gi || (!iq && !(L.flg[j] & 1))) continue;
                    const uint32_t m = tag | (uint32_t) j << 11;
                    if ((L.win[x] & 0x80) == si) {
                        if (kn >= n0 && kn < n1) L.smap[0][kn - n0] = m;
                        kn++;
                    } else {
                        if (kc >= c0 && kc < c1) L.smap[1][kc - c0] = m;
                        kc++;
                    }
                }
            }
            __syncthreads();
            const int nN = n1 - n0, nR = nN + c1 - c0;
            for (int u = threadIdx.x; u < nR; u += PK_T) {
                const int st = u >= nN;
                const int i = st ? u - nN + c0 : u + n0;          // pack-local seed number
                const uint32_t m = L.smap[st][st ? i - c0 : i - n0];
                const int pos = (int) (m & 0x7FFu), j = (int) ((m >> 11) & 0x7FFu);
                const int s = L.sidx[pos];
                const int P = L.Pb[s], Cb = L.Cb[s], sbs = L.sbytes[s], pc = P + Cb;
                const uint64_t pi = key_payload(L.key[pos], jbits, P, Cb) & ~(0x80ull << (8 * (pc - 1)));
                const uint64_t pq = key_payload(L.key[j], jbits, P, Cb);
                uint8_t *dst = st ? tmpC + L.baseC + L.bC0[s] + (uint64_t) (i - (int) L.sC0[s]) * sbs
                                  : tmpN + L.baseN + L.bN0[s] + (uint64_t) (i - (int) L.sN0[s]) * sbs;
                // the seed's exact sbs = 1 + 2 pc bytes in two overlapping stores
                const uint64_t hd = (uint64_t) (m >> 22) | (pi << 8);
                if (pc >= 4) {
                    uint64_t w0, w1;
                    seed_words(hd, pq, pc, w0, w1);
                    __builtin_memcpy(dst, &w0, 8);
                    __builtin_memcpy(dst + sbs - 8, &w1, 8);
                } else {
                    const uint64_t w = hd | (pq << (8 * (1 + pc)));
                    const uint32_t x0 = (uint32_t) w, x1 = (uint32_t) (w >> (8 * (sbs - 4)));
                    __builtin_memcpy(dst, &x0, 4);
                    __builtin_memcpy(dst + sbs - 4, &x1, 4);
                }
            }
            __syncthreads();
        }
        __syncthreads();
        if (qa >= 0) {
            const QDev d = qd[qa];
            const int lim = d.ncomb < PK_BUCK ? d.ncomb : PK_BUCK;
            for (int c = threadIdx.x; c < lim; c += PK_T) {
                if (L.bk[0][c]) atomicAdd(&nbuck[d.buck_off + c], (unsigned long long) L.bk[0][c]);
                if (L.bk[1][c]) atomicAdd(&cbuck[d.buck_off + c], (unsigned long long) L.bk[1][c]);
            }
        }
        __syncthreads();
    }
}

// S4. Every segment's seeds from the staging streams to their place in their
// query's streams (global seed index ofs[s] onward); one thread a segment,
// 8 bytes at a time (the last 8 bytes overlapping).
__device__ __forceinline__ void copy_bytes(uint8_t *__restrict__ dst, const uint8_t *__restrict__ src, int64_t n)
{
    if (n < 8) {
        for (int64_t b = 0; b < n; b++) dst[b] = src[b];
        return;
    }
    for (int64_t b = 0; b < n; b += 8) {
        const int64_t o = b + 8 <= n ? b : n - 8;
        uint64_t w;
        __builtin_memcpy(&w, src + o, 8);
        __builtin_memcpy(dst + o, &w, 8);
    }
}

__global__ void k_join_copy(const QSeg *__restrict__ segs, int64_t nseg, const QDev *__restrict__ qd,
                            const uint32_t *__restrict__ cntN, const uint32_t *__restrict__ cntC,
                            const int64_t *__restrict__ tofN, const int64_t *__restrict__ tofC,
                            const int64_t *__restrict__ ofsN, const int64_t *__restrict__ ofsC,
                            const uint8_t *__restrict__ tmpN, const uint8_t *__restrict__ tmpC,
                            uint8_t *__restrict__ seedN, uint8_t *__restrict__ seedC)
{
    for (int64_t s = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; s < nseg;
         s += (int64_t) gridDim.x * blockDim.x) {
        const QDev d = qd[segs[s].q];
        if (cntN[s])
            copy_bytes(seedN + d.nseed_off + (ofsN[s] - ofsN[d.seg0]) * d.s, tmpN + tofN[s], (int64_t) cntN[s] * d.s);
        if (cntC[s])
            copy_bytes(seedC + d.cseed_off + (ofsC[s] - ofsC[d.seg0]) * d.s, tmpC + tofC[s], (int64_t) cntC[s] * d.s);
    }
}

} // namespace pga
