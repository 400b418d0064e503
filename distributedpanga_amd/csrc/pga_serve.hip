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
//   3. a segment's target and Q keys (core plus FREQ+2 entries of context on
//      each side, or up to the group edge) are merged in LDS, LCPs recomputed,
//      and every core anchor's interval evaluated exactly as in the index merge;
//      seeds with a Q side are counted, then emitted in anchor order.
// Device keys of target and queries share one encoding (contig ids of a query
// start after the target's; one offset width), so they compare directly.

namespace pga {

constexpr int SEG_Q = 16;              // Q records per segment at most
constexpr int SEGCAP = 1024;           // combined entries a segment can hold in LDS
constexpr int JT = 256;                // threads of the join kernels
constexpr int MAX_SERVE_FREQ = 24;     // largest FREQ a server accepts
// A segment holds < 33W core target entries (W = FREQ+1: consecutive query
// records of one segment sit < 2W target positions apart, the core reaches W
// before the first and at most 2W past the last), FREQ+2 target entries of
// context on each side, and at most SEG_Q + 2(FREQ+2) query records.
static_assert(33 * (MAX_SERVE_FREQ + 1) + 4 * (MAX_SERVE_FREQ + 2) + SEG_Q <= SEGCAP, "segment exceeds LDS");

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

// Shared state of one segment in LDS.
struct JoinLds {
    Rec key[SEGCAP];
    uint8_t win[SEGCAP + 2 * MAX_FREQ + 8];    // LCP | strand<<7, offset by H, 0 outside
    uint8_t isq[SEGCAP];                        // bit0: query record, bit1: core anchor
    int64_t qt[SEG_Q + 2 * (MAX_SERVE_FREQ + 2)]; // tpos of the loaded query records
    int n, lo_idx, hi_idx;                      // combined entries; valid window [lo_idx, hi_idx]
};

// Load and merge one segment; returns the number of combined entries (all threads).
__device__ int join_load(JoinLds &L, const QSeg &sg, const QDev &d, const Rec *__restrict__ tkeys,
                         const Rec *__restrict__ qkeys, const uint32_t *__restrict__ qpre,
                         const int64_t *__restrict__ tpos, int freq, int H)
{
    const int MG = freq + 2;
    const uint32_t p = qpre[sg.cq_lo];
    const int64_t tlo = max(sg.glo, sg.ct_lo - MG), thi = min(sg.ghi, sg.ct_hi + MG);
    int64_t qlo = max(d.rec0, sg.cq_lo - MG), qhi = min(d.rec0 + d.nrec, sg.cq_hi + MG);
    // clip the query context to the group (at most MG steps each way)
    while (qlo < sg.cq_lo && qpre[qlo] != p) qlo++;
    while (qhi > sg.cq_hi && qpre[qhi - 1] != p) qhi--;
    const int nt = (int) (thi - tlo), nqq = (int) (qhi - qlo);
    const bool t_lo_trunc = tlo > sg.glo, t_hi_trunc = thi < sg.ghi;
    const bool q_lo_trunc = qlo > d.rec0 && qpre[qlo - 1] == p;
    const bool q_hi_trunc = qhi < d.rec0 + d.nrec && qpre[qhi] == p;
    const int n = nt + nqq;
    // query records: combined index = (records before it in T) + (rank among loaded Q)
    for (int i = threadIdx.x; i < nqq; i += JT) L.qt[i] = tpos[qlo + i];
    __syncthreads();
    for (int i = threadIdx.x; i < nqq; i += JT) {
        int64_t t = L.qt[i] - tlo;
        t = t < 0 ? 0 : (t > nt ? nt : t);
        const int ci = (int) t + i;
        L.key[ci] = qkeys[qlo + i];
        L.isq[ci] = 1 | ((qlo + i >= sg.cq_lo && qlo + i < sg.cq_hi) ? 2 : 0);
    }
    for (int i = threadIdx.x; i < nt; i += JT) {
        const int64_t k = tlo + i;
        // loaded query records placed before T[k]: those with tpos <= k
        int a = 0, b = nqq;
        while (a < b) { int m = (a + b) >> 1; if (L.qt[m] <= k) a = m + 1; else b = m; }
        const int ci = i + a;
        L.key[ci] = tkeys[k];
        L.isq[ci] = (k >= sg.ct_lo && k < sg.ct_hi) ? 2 : 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // the contiguous part of the combined order: clipped where a list was truncated
        int lo = 0, hi = n - 1;
        if (t_lo_trunc) { int a = 0, b = nqq; while (a < b) { int m = (a + b) >> 1; if (L.qt[m] <= tlo) a = m + 1; else b = m; } lo = max(lo, a); }
        if (q_lo_trunc && nqq) { int64_t t = L.qt[0] - tlo; t = t < 0 ? 0 : (t > nt ? nt : t); lo = max(lo, (int) t); }
        if (t_hi_trunc && nt) { int a = 0, b = nqq; while (a < b) { int m = (a + b) >> 1; if (L.qt[m] <= thi - 1) a = m + 1; else b = m; } hi = min(hi, nt - 1 + a); }
        if (q_hi_trunc && nqq) { int64_t t = L.qt[nqq - 1] - tlo; t = t < 0 ? 0 : (t > nt ? nt : t); hi = min(hi, (int) t + nqq - 1); }
        L.n = n;
        L.lo_idx = lo;
        L.hi_idx = hi;
    }
    __syncthreads();
    // LCP / strand window; 0 outside [lo_idx, hi_idx] and at the group start
    const int lo = L.lo_idx, hi = L.hi_idx;
    for (int i = threadIdx.x; i < n + 2 * H + 2; i += JT) L.win[i] = 0;
    __syncthreads();
    for (int i = lo + threadIdx.x; i <= hi; i += JT) {
        int l = 0;
        if (i > lo) l = lcp_bases(L.key[i - 1].w0, L.key[i - 1].w1, L.key[i].w0, L.key[i].w1);
        L.win[H + i] = (uint8_t) (l | ((L.key[i].w1 & 1ull) << 7));
    }
    __syncthreads();
    return n;
}

// Counts (count pass) or seeds (emit pass) of one segment's core anchors.
// A seed is kept when its anchor or its partner is a query record.
template <bool EMIT>
__device__ void join_anchors(JoinLds &L, const QDev &d, int jbits, int freq, int H,
                             uint32_t *nN, uint32_t *nC, uint8_t *__restrict__ outN, uint8_t *__restrict__ outC,
                             int64_t oN, int64_t oC, unsigned long long *__restrict__ nbuck,
                             unsigned long long *__restrict__ cbuck)
{
    const int pc = d.P + d.C, s = d.s;
    uint32_t cn = 0, cc = 0;
    for (int i = threadIdx.x; i < L.n; i += JT) {
        if (!(L.isq[i] & 2) || i < L.lo_idx || i > L.hi_idx) continue;
        const int gi = H + i;
        int plen, a, b;
        if (!anchor_interval(L.win, gi, freq, plen, a, b)) continue;
        const bool iq = L.isq[i] & 1;
        const uint8_t si = L.win[gi] & 0x80;
        uint32_t ln = 0, lc = 0;
        for (int q = a; q < b; q++) {
            if (q == gi) continue;
            const int j = q - H;
            if (!iq && !(L.isq[j] & 1)) continue;
            if ((L.win[q] & 0x80) == si) ln++; else lc++;
        }
        if (EMIT && (ln | lc)) {
            // offsets of this anchor's seeds: the anchors before it in the segment
            // were counted by the same loop order in a prefix pass (see caller)
            uint64_t pi = key_payload(L.key[i], jbits, d.P, d.C) & ~(0x80ull << (8 * (pc - 1)));
            int64_t kn = oN + nN[i], kc = oC + nC[i];
            for (int q = a; q < b; q++) {
                if (q == gi) continue;
                const int j = q - H;
                if (!iq && !(L.isq[j] & 1)) continue;
                const bool same = (L.win[q] & 0x80) == si;
                uint8_t *dst = same ? outN + (kn++) * s : outC + (kc++) * s;
                const uint64_t pq = key_payload(L.key[j], jbits, d.P, d.C);
                dst[0] = (uint8_t) plen;
                for (int t = 0; t < pc; t++) { dst[1 + t] = (uint8_t) (pi >> (8 * t)); dst[1 + pc + t] = (uint8_t) (pq >> (8 * t)); }
            }
            const uint64_t icont = (pi >> (8 * d.P)) & ((1ull << (8 * d.C - 1)) - 1);
            if (ln) atomicAdd(&nbuck[d.buck_off + icont], (unsigned long long) ln);
            if (lc) atomicAdd(&cbuck[d.buck_off + icont], (unsigned long long) lc);
        }
        if (!EMIT) { nN[i] = ln; nC[i] = lc; }
        cn += ln;
        cc += lc;
    }
    if (!EMIT) { (void) outN; }
    (void) cn; (void) cc;
}

// 3a. Seed counts per segment; per-entry counts kept in LDS for the exclusive
// prefix within the segment.
__global__ __launch_bounds__(JT) void k_join_count(const QSeg *__restrict__ segs, int64_t nseg,
                                                   const QDev *__restrict__ qd, const Rec *__restrict__ tkeys,
                                                   const Rec *__restrict__ qkeys, const uint32_t *__restrict__ qpre,
                                                   const int64_t *__restrict__ tpos, int jbits, int freq,
                                                   uint32_t *__restrict__ cntN, uint32_t *__restrict__ cntC)
{
    __shared__ JoinLds L;
    __shared__ uint32_t nN[SEGCAP], nC[SEGCAP];
    __shared__ uint64_t red[JT / 64 + 1];
    const int H = freq + 1;
    for (int64_t s = blockIdx.x; s < nseg; s += gridDim.x) {
        const QSeg sg = segs[s];
        const QDev d = qd[sg.q];
        const int n = join_load(L, sg, d, tkeys, qkeys, qpre, tpos, freq, H);
        for (int i = threadIdx.x; i < n; i += JT) { nN[i] = 0; nC[i] = 0; }
        __syncthreads();
        join_anchors<false>(L, d, jbits, freq, H, nN, nC, nullptr, nullptr, 0, 0, nullptr, nullptr);
        __syncthreads();
        uint64_t a = 0, b = 0;
        for (int i = threadIdx.x; i < n; i += JT) { a += nN[i]; b += nC[i]; }
        a = block_sum_u64<JT>(a, red);
        b = block_sum_u64<JT>(b, red);
        if (threadIdx.x == 0) { cntN[s] = (uint32_t) a; cntC[s] = (uint32_t) b; }
        __syncthreads();
    }
}

// 3b. Seeds: recount, exclusive prefix over the segment's entries, write.
__global__ __launch_bounds__(JT) void k_join_emit(const QSeg *__restrict__ segs, int64_t nseg,
                                                  const QDev *__restrict__ qd, const Rec *__restrict__ tkeys,
                                                  const Rec *__restrict__ qkeys, const uint32_t *__restrict__ qpre,
                                                  const int64_t *__restrict__ tpos, int jbits, int freq,
                                                  const int64_t *__restrict__ ofsN, const int64_t *__restrict__ ofsC,
                                                  uint8_t *__restrict__ seedN, uint8_t *__restrict__ seedC,
                                                  unsigned long long *__restrict__ nbuck,
                                                  unsigned long long *__restrict__ cbuck)
{
    __shared__ JoinLds L;
    __shared__ uint32_t nN[SEGCAP], nC[SEGCAP];
    __shared__ uint32_t sh[JT / 64 + 1];
    const int H = freq + 1;
    for (int64_t s = blockIdx.x; s < nseg; s += gridDim.x) {
        const QSeg sg = segs[s];
        const QDev d = qd[sg.q];
        (void) join_load(L, sg, d, tkeys, qkeys, qpre, tpos, freq, H);
        for (int i = threadIdx.x; i < SEGCAP; i += JT) { nN[i] = 0; nC[i] = 0; }
        __syncthreads();
        join_anchors<false>(L, d, jbits, freq, H, nN, nC, nullptr, nullptr, 0, 0, nullptr, nullptr);
        __syncthreads();
        // exclusive prefix of the per-entry counts in entry order (JT threads x 4 entries)
        constexpr int PER = SEGCAP / JT;
        uint32_t vn[PER], vc[PER], tn = 0, tc = 0;
#pragma unroll
        for (int k = 0; k < PER; k++) { vn[k] = nN[threadIdx.x * PER + k]; vc[k] = nC[threadIdx.x * PER + k]; tn += vn[k]; tc += vc[k]; }
        uint32_t totN, totC;
        uint32_t rn = block_excl_scan<JT, uint32_t>(tn, sh, &totN);
        uint32_t rc = block_excl_scan<JT, uint32_t>(tc, sh, &totC);
#pragma unroll
        for (int k = 0; k < PER; k++) { nN[threadIdx.x * PER + k] = rn; nC[threadIdx.x * PER + k] = rc; rn += vn[k]; rc += vc[k]; }
        __syncthreads();
        uint8_t *outN = seedN + d.nseed_off + (ofsN[s] - ofsN[d.seg0]) * (int64_t) d.s;
        uint8_t *outC = seedC + d.cseed_off + (ofsC[s] - ofsC[d.seg0]) * (int64_t) d.s;
        join_anchors<true>(L, d, jbits, freq, H, nN, nC, outN, outC, 0, 0, nbuck, cbuck);
        __syncthreads();
    }
}

} // namespace pga
