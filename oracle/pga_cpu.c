/* pga_cpu.c -- the CPU baseline timed beside the GPU path: the reference's
 * pga-mpi decomposition of extraction, bucket exchange, sort, LCP, adapter
 * and merge, run on T threads of one host (ranks become threads and the
 * all-to-all an in-memory copy).  Test infrastructure: bench.py's
 * cpu_baseline leg times it and tests/test_cpu_baseline.py checks that its
 * outputs equal the oracle's.  It builds into liboracle.so together with the
 * oracle, whose scan, LCP, adapter and merge-walk code it reuses per thread.
 *
 * Stages (rank r of the reference = thread r here):
 *   1. contig work split by bases, compute_work_split (src/work_split.c:3-27);
 *      each thread extracts its contigs (rolling closed-syncmer scan,
 *      src/extract.c:85-249) into its own records;
 *   2. 1024-bucket counts per thread, summed, and the shard splitter
 *      (bucket_local_counts / compute_destination_map, src/dsort.c:108-167);
 *   3. each thread scatters its records by owner (organize_send_buffer,
 *      src/dsort.c:172-224) and every owner gathers its slices from all
 *      threads in thread order (the all-to-all of run_stage2, :369-518);
 *   4. each owner sorts its records by k-mer (sort_records, src/sort.c:59-115;
 *      here a stable MSD radix sort, so equal k-mers keep extraction order)
 *      and computes LCPs, the first against the previous owner's last record
 *      (the seam fix-up, src/dsort.c:507-516);
 *   5. GIX entries at each owner's global offset and the 2^24 prefix counts of
 *      its bucket range (build_kmer_adapter, src/kmer_adapter.c:6-96);
 *   6. each owner runs the adaptamer walk over its prefix groups, which never
 *      cross owners (src/fastga_pipeline.c:245-412); streams are concatenated
 *      in owner order and the per-contig counts summed.
 */
#include <pthread.h>

#include "pga_oracle.c"

#define CPU_NB 1024                        /* 5-base buckets (src/dsort.h:30-33) */

typedef struct {
    /* shared input */
    const uint8_t *bps;
    const int64_t *boff, *clen;
    int ncontig, T, P, C, rsize, e, freq;
    int *split;                            /* contig split [T+1]          */
    int ksplit[CPU_NB + 1];                /* bucket boundary per owner    */
    int select[CPU_NB];
    /* per thread */
    recbuf *rb;                            /* extracted records            */
    int64_t (*bcount)[CPU_NB];             /* bucket counts                */
    int64_t *sendcnt;                      /* [T][T]: thread t -> owner o  */
    uint8_t **sendbuf;                     /* thread t's records by owner  */
    uint8_t **own;                         /* owner o's records            */
    int64_t *nown;
    int64_t *goff;                         /* owner's first global record  */
    uint8_t *entries;
    int64_t *cum;
    int64_t *maxpre_o;
    uint8_t *seam;                         /* owner's first LCP (vs the previous owner) */
    mctx *mm;                              /* per-owner merge streams      */
    int err;
} cpu_job;

typedef struct {
    cpu_job *j;
    int t;
    void (*fn)(cpu_job *, int);
} cpu_arg;

static void *cpu_thread(void *p)
{
    cpu_arg *a = (cpu_arg *) p;
    a->fn(a->j, a->t);
    return NULL;
}

static void cpu_parallel(cpu_job *j, void (*fn)(cpu_job *, int))
{
    pthread_t th[256];
    cpu_arg args[256];
    for (int t = 0; t < j->T; t++) {
        args[t].j = j;
        args[t].t = t;
        args[t].fn = fn;
        if (t > 0) pthread_create(&th[t], NULL, cpu_thread, &args[t]);
    }
    fn(j, 0);
    for (int t = 1; t < j->T; t++) pthread_join(th[t], NULL);
}

/* ------------------------------------------------------------ 1. extraction */

static void st_extract(cpu_job *j, int t)
{
    recbuf *b = &j->rb[t];
    int64_t maxl = 0, bases = 0;
    for (int c = j->split[t]; c < j->split[t + 1]; c++) {
        if (j->clen[c] > maxl) maxl = j->clen[c];
        bases += j->clen[c];
    }
    /* room for the usual ~0.8 records a base up front (the buffer still grows) */
    b->cap = bases * 0.85 > (1 << 16) ? (int64_t) (bases * 0.85) : (1 << 16);
    b->buf = (uint8_t *) malloc((size_t) (b->cap * b->rsize));
    if (!b->buf) { j->err = 1; return; }
    uint8_t *codes = (uint8_t *) malloc((size_t) (maxl + 8));
    if (!codes) { j->err = 1; return; }
    for (int c = j->split[t]; c < j->split[t + 1]; c++) {
        if (j->boff[c] < 0) continue;                     /* src/extract.c:97-98 */
        unpack_contig(j->bps, j->boff[c], j->clen[c], codes);
        if (scan_rolling(b, codes, j->clen[c], c)) { j->err = 1; break; }
    }
    free(codes);
    /* 2. bucket counts of the thread's records (the first 10 k-mer bits) */
    int64_t *bc = j->bcount[t];
    memset(bc, 0, sizeof(int64_t) * CPU_NB);
    for (int64_t i = 0; i < b->n; i++) {
        const uint8_t *r = b->buf + i * j->rsize;
        bc[(r[1] << 2) | (r[2] >> 6)]++;
    }
}

/* ------------------------------------------------------------ 3. exchange */

static void st_scatter(cpu_job *j, int t)
{
    recbuf *b = &j->rb[t];
    const int T = j->T, rs = j->rsize;
    int64_t *cnt = j->sendcnt + (int64_t) t * T;
    int64_t cur[256];
    memset(cnt, 0, sizeof(int64_t) * T);
    for (int k = 0; k < CPU_NB; k++) cnt[j->select[k]] += j->bcount[t][k];
    int64_t run = 0;
    for (int o = 0; o < T; o++) { cur[o] = run; run += cnt[o]; }
    j->sendbuf[t] = (uint8_t *) malloc((size_t) (b->n * rs + 1));
    if (!j->sendbuf[t]) { j->err = 1; return; }
    for (int64_t i = 0; i < b->n; i++) {
        const uint8_t *r = b->buf + i * rs;
        const int o = j->select[(r[1] << 2) | (r[2] >> 6)];
        memcpy(j->sendbuf[t] + (cur[o]++) * rs, r, (size_t) rs);
    }
    free(b->buf);
    b->buf = NULL;
}

static void st_gather(cpu_job *j, int o)
{
    const int T = j->T, rs = j->rsize;
    int64_t n = 0;
    for (int t = 0; t < T; t++) n += j->sendcnt[(int64_t) t * T + o];
    j->nown[o] = n;
    j->own[o] = (uint8_t *) malloc((size_t) (n * rs + 1));
    if (!j->own[o]) { j->err = 1; return; }
    int64_t at = 0;
    for (int t = 0; t < T; t++) {               /* thread order = contig order */
        int64_t before = 0;
        for (int q = 0; q < o; q++) before += j->sendcnt[(int64_t) t * T + q];
        const int64_t c = j->sendcnt[(int64_t) t * T + o];
        memcpy(j->own[o] + at * rs, j->sendbuf[t] + before * rs, (size_t) (c * rs));
        at += c;
    }
}

/* ------------------------------------------------------------ 4. sort + LCP */

/* Stable MSD radix sort of n records on k-mer bytes [byte, 10]; the result is
 * left in a (tmp is scratch of the same size). */
static void msd_stable(uint8_t *a, uint8_t *tmp, int64_t n, int rs, int byte)
{
    while (n > 1 && byte <= O_KBYTES) {
        if (n <= 32) {                           /* stable insertion sort */
            uint8_t rec[64];
            for (int64_t i = 1; i < n; i++) {
                memcpy(rec, a + i * rs, (size_t) rs);
                int64_t k = i;
                while (k > 0 && memcmp(a + (k - 1) * rs + byte, rec + byte, (size_t) (O_KBYTES + 1 - byte)) > 0) {
                    memcpy(a + k * rs, a + (k - 1) * rs, (size_t) rs);
                    k--;
                }
                memcpy(a + k * rs, rec, (size_t) rs);
            }
            return;
        }
        int64_t cnt[256], start[256];
        memset(cnt, 0, sizeof(cnt));
        for (int64_t i = 0; i < n; i++) cnt[a[i * rs + byte]]++;
        int64_t run = 0, nz = 0;
        for (int d = 0; d < 256; d++) {
            start[d] = run;
            run += cnt[d];
            if (cnt[d]) nz++;
        }
        if (nz == 1) { byte++; continue; }      /* one digit: nothing moves */
        int64_t cur[256];
        memcpy(cur, start, sizeof(cur));
        for (int64_t i = 0; i < n; i++) {
            const uint8_t *r = a + i * rs;
            memcpy(tmp + (cur[r[byte]]++) * rs, r, (size_t) rs);
        }
        memcpy(a, tmp, (size_t) (n * rs));
        for (int d = 0; d < 256; d++)
            if (cnt[d] > 1) msd_stable(a + start[d] * rs, tmp + start[d] * rs, cnt[d], rs, byte + 1);
        return;
    }
}

static void st_sort(cpu_job *j, int o)
{
    const int rs = j->rsize;
    const int64_t n = j->nown[o];
    uint8_t *tmp = (uint8_t *) malloc((size_t) (n * rs + 1));
    if (!tmp) { j->err = 1; return; }
    msd_stable(j->own[o], tmp, n, rs, 1);
    free(tmp);
    oracle_recalc_lcps(j->own[o], n, rs);
}

/* ------------------------------------------------------------ 5. adapter */

static void st_adapter(cpu_job *j, int o)
{
    const int rs = j->rsize, e = j->e;
    const int64_t n = j->nown[o], g0 = j->goff[o];
    const uint8_t *own = j->own[o];
    uint8_t *ent = j->entries + g0 * e;
    const int64_t p0 = (int64_t) j->ksplit[o] << 14, p1 = (int64_t) j->ksplit[o + 1] << 14;
    for (int64_t p = p0; p < p1; p++) j->cum[p] = 0;
    int64_t run = 0, best = 0;
    int last = -1;
    for (int64_t i = 0; i < n; i++) {
        const uint8_t *r = own + i * rs;
        uint8_t *d = ent + i * e;
        int lcp = r[0];
        if (i == 0) lcp = j->seam[o];
        const int pre = (r[1] << 16) | (r[2] << 8) | r[3];
        if (pre == last) run++;
        else { if (run > best) best = run; run = 1; last = pre; }
        memcpy(d, r + 4, 7);
        d[7] = 0;
        d[8] = (uint8_t) lcp;
        memcpy(d + 9, r + 12, (size_t) (j->P + j->C));
        j->cum[pre]++;
    }
    if (run > best) best = run;
    j->maxpre_o[o] = best;
    int64_t acc = g0;                           /* entries of the owners before */
    for (int64_t p = p0; p < p1; p++) { acc += j->cum[p]; j->cum[p] = acc; }
}

/* ------------------------------------------------------------ 6. merge */

static void st_merge(cpu_job *j, int o)
{
    mctx *m = &j->mm[o];
    memset(m, 0, sizeof(*m));
    m->ent = j->entries; m->e = j->e; m->P = j->P; m->C = j->C; m->pc = j->P + j->C;
    m->nbuck = (int64_t *) calloc((size_t) (j->ncontig > 0 ? j->ncontig : 1), sizeof(int64_t));
    m->cbuck = (int64_t *) calloc((size_t) (j->ncontig > 0 ? j->ncontig : 1), sizeof(int64_t));
    if (!m->nbuck || !m->cbuck) { j->err = 1; return; }
    int64_t g0 = j->goff[o];
    const int64_t end = j->goff[o] + j->nown[o];
    while (g0 < end) {
        int64_t g1 = g0 + 1;
        while (g1 < end && j->entries[g1 * j->e + 8] >= 12) g1++;
        if (walk_group_fastga(m, g0, g1, j->freq)) { j->err = 1; return; }
        g0 = g1;
    }
}

static double cpu_now(void) { return now_s(); }

int pga_cpu_pipeline(const uint8_t *bps, const int64_t *boff, const int64_t *clen, int ncontig,
                     int freq, int nthreads, oracle_result *r)
{
    cpu_job jj, *j = &jj;
    memset(j, 0, sizeof(*j));
    memset(r, 0, sizeof(*r));
    if (nthreads < 1 || nthreads > 256) return -1;
    const int T = nthreads;
    int64_t maxctg = 0, seqtot = 0;
    for (int c = 0; c < ncontig; c++) {
        if (clen[c] > maxctg) maxctg = clen[c];
        seqtot += clen[c];
    }
    oracle_record_sizing(maxctg, ncontig, &r->P, &r->C, &r->rsize);
    r->e = 9 + r->P + r->C;
    r->s = 1 + 2 * (r->P + r->C);
    r->ncontig = ncontig;
    j->bps = bps; j->boff = boff; j->clen = clen; j->ncontig = ncontig; j->T = T;
    j->P = r->P; j->C = r->C; j->rsize = r->rsize; j->e = r->e; j->freq = freq;

    double t0 = cpu_now();
    /* 1. work split by bases (src/work_split.c:3-27) */
    int split[257];
    {
        int64_t so_far = 0, next = seqtot / T;
        int w = 1;
        split[0] = 0;
        for (int c = 0; c < ncontig; c++) {
            so_far += clen[c];
            while (so_far >= next && w < T) {
                split[w] = c + 1;
                w++;
                next = seqtot * w / T;
            }
        }
        for (; w <= T; w++) split[w] = ncontig;
    }
    j->split = split;
    j->rb = (recbuf *) calloc((size_t) T, sizeof(recbuf));
    j->bcount = (int64_t (*)[CPU_NB]) calloc((size_t) T, sizeof(*j->bcount));
    j->sendcnt = (int64_t *) calloc((size_t) T * T, sizeof(int64_t));
    j->sendbuf = (uint8_t **) calloc((size_t) T, sizeof(uint8_t *));
    j->own = (uint8_t **) calloc((size_t) T, sizeof(uint8_t *));
    j->nown = (int64_t *) calloc((size_t) T, sizeof(int64_t));
    j->goff = (int64_t *) calloc((size_t) T + 1, sizeof(int64_t));
    j->maxpre_o = (int64_t *) calloc((size_t) T, sizeof(int64_t));
    j->seam = (uint8_t *) calloc((size_t) T, 1);
    j->mm = (mctx *) calloc((size_t) T, sizeof(mctx));
    if (!j->rb || !j->bcount || !j->sendcnt || !j->sendbuf || !j->own || !j->nown || !j->goff ||
        !j->maxpre_o || !j->seam || !j->mm)
        return -1;
    for (int t = 0; t < T; t++) { j->rb[t].P = r->P; j->rb[t].C = r->C; j->rb[t].rsize = r->rsize; }
    cpu_parallel(j, st_extract);
    r->t_extract = cpu_now() - t0;
    if (j->err) return -1;

    t0 = cpu_now();
    /* 2. global bucket counts and the owners' bucket ranges (src/dsort.c:120-167) */
    int64_t global[CPU_NB];
    memset(global, 0, sizeof(global));
    for (int t = 0; t < T; t++)
        for (int k = 0; k < CPU_NB; k++) global[k] += j->bcount[t][k];
    {
        int64_t cum[CPU_NB];
        cum[0] = global[0];
        for (int i = 1; i < CPU_NB; i++) cum[i] = cum[i - 1] + global[i];
        const int64_t total = cum[CPU_NB - 1];
        int n = 1;
        int64_t tt = total / T;
        j->ksplit[0] = 0;
        for (int i = 0; i < CPU_NB; i++) {
            if (cum[i] >= tt) {
                const int64_t prev = i > 0 ? cum[i - 1] : 0;
                if (cum[i] - tt > tt - prev) { j->select[i] = n; j->ksplit[n] = i; }
                else { j->select[i] = n - 1; j->ksplit[n] = i + 1; }
                n++;
                if (n > T) n = T;
                tt = (int64_t) n * total / T;
            } else {
                j->select[i] = n - 1;
            }
        }
        j->ksplit[T] = CPU_NB;
        for (int q = n; q < T; q++) j->ksplit[q] = CPU_NB;
    }
    cpu_parallel(j, st_scatter);
    if (j->err) return -1;
    cpu_parallel(j, st_gather);
    for (int t = 0; t < T; t++) { free(j->sendbuf[t]); j->sendbuf[t] = NULL; }
    if (j->err) return -1;
    cpu_parallel(j, st_sort);
    r->t_sort = cpu_now() - t0;
    if (j->err) return -1;

    t0 = cpu_now();
    for (int o = 0; o < T; o++) j->goff[o + 1] = j->goff[o] + j->nown[o];
    r->nrec = j->goff[T];
    r->entries = (uint8_t *) malloc((size_t) (r->nrec * r->e + 1));
    r->cum = (int64_t *) malloc(sizeof(int64_t) * O_NPRE);
    if (!r->entries || !r->cum) return -1;
    j->entries = r->entries;
    j->cum = r->cum;
    /* seam LCPs (src/dsort.c:507-516): each owner's first record against the
       previous non-empty owner's last */
    for (int o = 0, prev = -1; o < T; o++) {
        if (j->nown[o] == 0) continue;
        j->seam[o] = prev >= 0 ? (uint8_t) oracle_lcp_bases(j->own[prev] + (j->nown[prev] - 1) * r->rsize + 1,
                                                              j->own[o] + 1) : 0;
        prev = o;
    }
    cpu_parallel(j, st_adapter);
    for (int o = 0; o < T; o++) { free(j->own[o]); j->own[o] = NULL; }
    r->maxpre = 0;
    for (int o = 0; o < T; o++) if (j->maxpre_o[o] > r->maxpre) r->maxpre = j->maxpre_o[o];
    r->t_adapter = cpu_now() - t0;

    t0 = cpu_now();
    cpu_parallel(j, st_merge);
    if (j->err) return -1;
    r->nbuck = (int64_t *) calloc((size_t) (ncontig > 0 ? ncontig : 1), sizeof(int64_t));
    r->cbuck = (int64_t *) calloc((size_t) (ncontig > 0 ? ncontig : 1), sizeof(int64_t));
    int64_t nl = 0, cl = 0;
    for (int o = 0; o < T; o++) { nl += j->mm[o].N.len; cl += j->mm[o].Cs.len; }
    r->nseeds = (uint8_t *) malloc((size_t) nl + 1);
    r->cseeds = (uint8_t *) malloc((size_t) cl + 1);
    if (!r->nbuck || !r->cbuck || !r->nseeds || !r->cseeds) return -1;
    for (int o = 0; o < T; o++) {
        mctx *m = &j->mm[o];
        if (m->N.len) memcpy(r->nseeds + r->nlen, m->N.buf, (size_t) m->N.len);
        if (m->Cs.len) memcpy(r->cseeds + r->clen, m->Cs.buf, (size_t) m->Cs.len);
        r->nlen += m->N.len;
        r->clen += m->Cs.len;
        for (int c = 0; c < ncontig; c++) { r->nbuck[c] += m->nbuck[c]; r->cbuck[c] += m->cbuck[c]; }
        free(m->N.buf); free(m->Cs.buf); free(m->nbuck); free(m->cbuck);
    }
    r->t_merge = cpu_now() - t0;
    free(j->rb); free(j->bcount); free(j->sendcnt); free(j->sendbuf); free(j->own);
    free(j->nown); free(j->goff); free(j->maxpre_o); free(j->seam); free(j->mm);
    return 0;
}

/* ------------------------------------------------------ bucket counts only */

/* Records per 5-base bucket (record_bucket, src/dsort.h:30-33) of the whole
 * GDB without storing records: the rolling scan's selections, with each
 * record's first five bases read off the contig (forward record at j: bases
 * j..j+4; reverse-complement record: the complements of bases j+11..j+7).
 * A size-independent checksum of the extraction for inputs whose records do
 * not fit host memory. */
typedef struct {
    const uint8_t *bps;
    const int64_t *boff, *clen;
    int *split;
    int64_t (*cnt)[CPU_NB];
    int err;
} cnt_job;

static void count_contig(const uint8_t *s, int64_t len, int64_t *cnt)
{
    int ring[4], mn, mpos;
    if (len < O_TMER) return;
    for (int t = 0; t < 4; t++) ring[t] = oracle_mz(s, t);
    mn = ring[0]; mpos = 0;
    for (int t = 1; t < 4; t++) if (ring[t] < mn) { mn = ring[t]; mpos = t; }
    for (int64_t i = O_SOFF; i <= len - O_SMER; i++) {
        int h = oracle_mz(s, i), sel = 0;
        ring[i & 3] = h;
        if (h < mn) { mn = h; mpos = (int) i; sel = 1; }
        else if (mpos == i - O_SOFF) {
            sel = 1;
            mpos++;
            mn = ring[mpos & 3];
            for (int64_t q = mpos + 1; q <= i; q++)
                if (ring[q & 3] < mn) { mn = ring[q & 3]; mpos = (int) q; }
        } else if (h == mn) sel = 1;
        if (!sel) continue;
        const int64_t j = i - O_SOFF;
        if (j <= len - O_KMER)
            cnt[(s[j] << 8) | (s[j + 1] << 6) | (s[j + 2] << 4) | (s[j + 3] << 2) | s[j + 4]]++;
        if (j >= O_KMER - O_TMER)
            cnt[((3 - s[j + 11]) << 8) | ((3 - s[j + 10]) << 6) | ((3 - s[j + 9]) << 4) |
                ((3 - s[j + 8]) << 2) | (3 - s[j + 7])]++;
    }
}

typedef struct { cnt_job *j; int t; } cnt_arg;

static void *cnt_thread(void *p)
{
    cnt_arg *a = (cnt_arg *) p;
    cnt_job *j = a->j;
    int64_t maxl = 0;
    for (int c = j->split[a->t]; c < j->split[a->t + 1]; c++) if (j->clen[c] > maxl) maxl = j->clen[c];
    uint8_t *codes = (uint8_t *) malloc((size_t) (maxl + 8));
    if (!codes) { j->err = 1; return NULL; }
    for (int c = j->split[a->t]; c < j->split[a->t + 1]; c++) {
        if (j->boff[c] < 0) continue;
        unpack_contig(j->bps, j->boff[c], j->clen[c], codes);
        count_contig(codes, j->clen[c], j->cnt[a->t]);
    }
    free(codes);
    return NULL;
}

int pga_cpu_bucket_counts(const uint8_t *bps, const int64_t *boff, const int64_t *clen, int ncontig,
                          int nthreads, int64_t *counts)
{
    if (nthreads < 1 || nthreads > 256) return -1;
    /* one contig per task, handed out in order over the threads' split */
    const int T = nthreads < ncontig ? nthreads : (ncontig > 0 ? ncontig : 1);
    int split[257];
    int64_t seqtot = 0;
    for (int c = 0; c < ncontig; c++) seqtot += clen[c];
    {
        int64_t so_far = 0, next = seqtot / T;
        int w = 1;
        split[0] = 0;
        for (int c = 0; c < ncontig; c++) {
            so_far += clen[c];
            while (so_far >= next && w < T) { split[w] = c + 1; w++; next = seqtot * w / T; }
        }
        for (; w <= T; w++) split[w] = ncontig;
    }
    cnt_job j;
    memset(&j, 0, sizeof(j));
    j.bps = bps; j.boff = boff; j.clen = clen; j.split = split;
    j.cnt = (int64_t (*)[CPU_NB]) calloc((size_t) T, sizeof(*j.cnt));
    if (!j.cnt) return -1;
    pthread_t th[256];
    cnt_arg args[256];
    for (int t = 0; t < T; t++) {
        args[t].j = &j;
        args[t].t = t;
        pthread_create(&th[t], NULL, cnt_thread, &args[t]);
    }
    for (int t = 0; t < T; t++) pthread_join(th[t], NULL);
    memset(counts, 0, sizeof(int64_t) * CPU_NB);
    for (int t = 0; t < T; t++)
        for (int k = 0; k < CPU_NB; k++) counts[k] += j.cnt[t][k];
    free(j.cnt);
    return j.err ? -1 : 0;
}
