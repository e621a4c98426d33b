/* gb_oracle — CPU oracle of the dense verification path (TEST INFRASTRUCTURE:
 * only tests/, the smoke check and bench.py's CPU-baseline leg run it; the
 * product path never does).
 *
 * Restates, from the definitions, what the reference's CPU oracle computes
 * (tests/stats_ref.cpp:64-190) and what its verifier records in one run
 * (include/gstats.hpp GoldbachStats, src/verifier.cpp segmentation):
 *
 *   gb_oracle ref START LIMIT [THREADS]
 *       the stats_ref output line: evens pmax primes twins primes30 twins30
 *       first hist gaps gmom gmax gtail ghist (stats_ref.cpp:104-189), every
 *       value by Miller-Rabin per even / per odd q;
 *   gb_oracle gbstats START LIMIT SEGMENT SAMPLE GAPS [THREADS]
 *       the GBSTATS v2 statistics fields of a fresh-state run over the same
 *       range with segment span SEGMENT, GB_STATS_SAMPLE = SAMPLE and
 *       GB_GAPS = GAPS (exact | sampled | 0), then the state records
 *       "sampled", "tailcls" and "tailagg" (gstats.hpp:913-969, 1267-1315);
 *       everything after "version=..." so the product's line can be compared
 *       byte for byte;
 *   gb_oracle pn N
 *       p(n) of one even (gstats.hpp:139-144);
 *   gb_oracle ridges START LIMIT FLOOR K [THREADS]
 *       the K largest Goldbach ridges (n, p(n)) with p(n) >= FLOOR, ordered
 *       by p then n descending (include/ridge.hpp:52-60, ridge_before), one
 *       "n p" line each.
 *
 * Parity pins (tests/test_oracle.py): the reference's own stats_ref output on
 * the golden windows (tests/golden/ref_*.txt), its --known table
 * (stats_ref.cpp:35-42), the base-2 pseudoprimes below 2^32 for the
 * Miller-Rabin, and the published prime counts / records the reference's
 * tests and README quote.
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gb_mr.h"

#define MAX_GB_PRIME 9781u
#define TAIL_T 3000u          /* CHECK30_TAIL, include/check30.hpp:42 */
#define GAP_TAIL 700u         /* GAPS30_TAIL, include/gstats.hpp:74 */
#define GAP_BINS 1024u        /* GAPS30_HBINS */

static uint32_t g_pr[1300];   /* odd primes 3..9781 */
static uint32_t g_np;
static int16_t g_pidx[MAX_GB_PRIME / 2 + 1];   /* p >> 1 -> index, -1 none */

static void init_primes(void)
{
    static char comp[MAX_GB_PRIME + 1];
    for (uint32_t p = 3; p * p <= MAX_GB_PRIME; p += 2)
        if (!comp[p])
            for (uint32_t m = p * p; m <= MAX_GB_PRIME; m += 2 * p) comp[m] = 1;
    memset(g_pidx, 0xFF, sizeof g_pidx);
    for (uint32_t p = 3; p <= MAX_GB_PRIME; p += 2)
        if (!comp[p]) { g_pidx[p >> 1] = (int16_t)g_np; g_pr[g_np++] = p; }
}

/* p(n) by the ascending scan of stats_ref.cpp:80-83: the candidates <= 9781
 * by "n - c prime", then every odd c >= 9783 with c and n - c prime. */
static uint64_t pn_scan(uint64_t n)
{
    for (uint32_t i = 0; i < g_np; ++i)
        if (gb_is_prime_u64(n - g_pr[i])) return g_pr[i];
    for (uint64_t c = MAX_GB_PRIME + 2;; c += 2)
        if (gb_is_prime_u64(c) && gb_is_prime_u64(n - c)) return c;
}

/* p(n) as pn_exact (gstats.hpp:139-144): smallest odd prime p <= n/2 with
 * n - p prime, 0 if none. */
static uint64_t pn_exact(uint64_t n)
{
    for (uint64_t p = 3; p <= n / 2; p += 2)
        if (gb_is_prime_u64(p) && gb_is_prime_u64(n - p)) return p;
    return 0;
}


static void u128_str(gb_u128 v, char* out)
{
    char b[48];
    int n = 0;
    do { b[n++] = (char)('0' + (int)(v % 10)); v /= 10; } while (v);
    for (int i = 0; i < n; ++i) out[i] = b[n - 1 - i];
    out[n] = 0;
}

/* ---------------------------------------------------------------- per-even pass */

typedef struct {
    uint64_t start, limit, evens;
    unsigned nt, t;
    /* segmentation (gbstats mode): even index ranges of the sampled segments */
    const uint64_t* samp_lo;
    const uint64_t* samp_hi;
    size_t nsamp;
    /* outputs */
    uint64_t* hist;          /* [15 * np]: p(n) by class n mod 30 / 2 and prime index, sampled evens only (gbstats) or all (ref) */
    uint64_t* tail_all;      /* [np] tail counts by prime index over all evens */
    uint64_t first[1300];    /* min n per prime index (0 none) */
    uint64_t primes30[30], twins30[30];
    uint64_t* beyond_n;      /* p(n) > 9781: n, p (rare) */
    uint64_t* beyond_p;
    size_t nbeyond, cbeyond;
    uint64_t sampled_evens;
    int gbstats;
} EvenTask;

static int in_sampled(const EvenTask* a, uint64_t i)
{
    for (size_t k = 0; k < a->nsamp; ++k)
        if (i >= a->samp_lo[k] && i <= a->samp_hi[k]) return 1;
    return 0;
}

static void* even_worker(void* arg)
{
    EvenTask* a = (EvenTask*)arg;
    for (uint64_t i = a->t; i < a->evens; i += a->nt) {
        const uint64_t n = a->start + 2 * i;
        const uint64_t p = pn_scan(n);
        const uint32_t cr = (uint32_t)(n % 30) / 2;
        const int sampled = a->gbstats ? in_sampled(a, i) : 1;
        if (p <= MAX_GB_PRIME) {
            const uint32_t pi = (uint32_t)g_pidx[p >> 1];
            if (sampled) { a->hist[cr * g_np + pi]++; a->sampled_evens++; }
            if (p >= TAIL_T) a->tail_all[pi]++;
            if (!a->first[pi] || n < a->first[pi]) a->first[pi] = n;
        } else {
            if (sampled) a->sampled_evens++;
            if (a->nbeyond == a->cbeyond) {
                a->cbeyond = a->cbeyond ? 2 * a->cbeyond : 16;
                a->beyond_n = realloc(a->beyond_n, a->cbeyond * 8);
                a->beyond_p = realloc(a->beyond_p, a->cbeyond * 8);
            }
            a->beyond_n[a->nbeyond] = n;
            a->beyond_p[a->nbeyond++] = p;
        }
        /* primes and twins of q = n + 1 in (start, limit) (stats_ref.cpp:87-92) */
        const uint64_t q = n + 1;
        if (q < a->limit && gb_is_prime_u64(q)) {
            a->primes30[q % 30]++;
            if (q + 2 < a->limit && gb_is_prime_u64(q + 2)) a->twins30[q % 30]++;
        }
    }
    return NULL;
}

/* ---------------------------------------------------------------- gap pass */

typedef struct {
    uint64_t a, e;                /* odd q in [a, e) */
    uint64_t first, last, n;
    gb_u128 d[4];
    uint64_t* h;                  /* [GAPHMAX] by g / 2 */
    uint64_t tail_cnt[4096], tail_lo[4096];   /* g >= 700 by g / 2 (g < 8192) */
    uint64_t gmax, gmax_lo;
    /* per-gap filter for the walked-only histogram: gaps whose upper prime
     * lies in a walked segment (gbstats GB_GAPS=sampled) */
    const uint64_t* w_lo;
    const uint64_t* w_hi;
    size_t nw;
    int filter;
    uint64_t wn;
    gb_u128 wd[4];
    uint64_t* wh;
} GapTask;

#define GAPHMAX 8192

static int walked_q(const GapTask* c, uint64_t q)
{
    if (!c->filter) return 1;
    for (size_t k = 0; k < c->nw; ++k)
        if (q >= c->w_lo[k] && q <= c->w_hi[k]) return 1;
    return 0;
}

static void gap_add(GapTask* c, uint64_t g, uint64_t up)
{
    ++c->n;
    const gb_u128 g1 = g, g2 = g1 * g1;
    c->d[0] += g1; c->d[1] += g2; c->d[2] += g2 * g1; c->d[3] += g2 * g2;
    c->h[g / 2]++;
    if (g >= GAP_TAIL) {
        if (!c->tail_cnt[g / 2] || up - g < c->tail_lo[g / 2]) c->tail_lo[g / 2] = up - g;
        c->tail_cnt[g / 2]++;
    }
    if (g > c->gmax || (g == c->gmax && up - g < c->gmax_lo)) { c->gmax = g; c->gmax_lo = up - g; }
    if (walked_q(c, up)) {
        ++c->wn;
        c->wd[0] += g1; c->wd[1] += g2; c->wd[2] += g2 * g1; c->wd[3] += g2 * g2;
        c->wh[g / 2]++;
    }
}

static void* gap_worker(void* arg)
{
    GapTask* c = (GapTask*)arg;
    for (uint64_t q = c->a; q < c->e; q += 2) {
        if (!gb_is_prime_u64(q)) continue;
        if (c->last) gap_add(c, q - c->last, q); else c->first = q;
        c->last = q;
    }
    return NULL;
}

/* ---------------------------------------------------------------- output helpers */

static void dense(const uint64_t* v, size_t n, size_t from, char** out, size_t* cap, size_t* len)
{
    size_t last = from;
    for (size_t i = from; i < n; ++i) if (v[i]) last = i + 1;
    for (size_t i = from; i < last; ++i) {
        char b[32];
        int k = 0;
        if (i > from) b[k++] = ',';
        if (v[i]) k += sprintf(b + k, "%llu", (unsigned long long)v[i]);
        if (*len + (size_t)k + 1 > *cap) { *cap = 2 * (*cap + (size_t)k + 64); *out = realloc(*out, *cap); }
        memcpy(*out + *len, b, (size_t)k);
        *len += (size_t)k;
        (*out)[*len] = 0;
    }
}

static void print_dense(const uint64_t* v, size_t n, size_t from)
{
    char* s = malloc(64);
    size_t cap = 64, len = 0;
    s[0] = 0;
    dense(v, n, from, &s, &cap, &len);
    fputs(s, stdout);
    free(s);
}

static int dense_empty(const uint64_t* v, size_t n, size_t from)
{
    for (size_t i = from; i < n; ++i) if (v[i]) return 0;
    return 1;
}

/* ---------------------------------------------------------------- main runs */

typedef struct {
    EvenTask* ev;
    GapTask* gp;
    unsigned nt;
    uint64_t start, limit, evens;
    /* merged */
    uint64_t hist_c[15 * 1300], tail_all[1300], first[1300], primes30[30], twins30[30];
    uint64_t sampled_evens;
    uint64_t *bn, *bp;
    size_t nb;
    /* gaps merged */
    uint64_t gn, gmax, gmax_lo, wn;
    gb_u128 gd[4], wd[4];
    uint64_t gh[GAPHMAX], wh[GAPHMAX], tail_cnt[4096], tail_lo[4096];
    int have_gaps;
} Run;

static void run_pass(Run* R, const uint64_t* slo, const uint64_t* shi, size_t ns,
                     const uint64_t* wlo, const uint64_t* whi, size_t nw, int filter, int gbstats)
{
    const unsigned nt = R->nt;
    R->ev = calloc(nt, sizeof(EvenTask));
    pthread_t* th = calloc(nt, sizeof(pthread_t));
    for (unsigned t = 0; t < nt; ++t) {
        EvenTask* a = &R->ev[t];
        a->start = R->start; a->limit = R->limit; a->evens = R->evens; a->nt = nt; a->t = t;
        a->samp_lo = slo; a->samp_hi = shi; a->nsamp = ns; a->gbstats = gbstats;
        a->hist = calloc(15 * g_np, 8);
        a->tail_all = calloc(g_np, 8);
        pthread_create(&th[t], NULL, even_worker, a);
    }
    for (unsigned t = 0; t < nt; ++t) pthread_join(th[t], NULL);
    for (unsigned t = 0; t < nt; ++t) {
        EvenTask* a = &R->ev[t];
        for (uint32_t i = 0; i < 15 * g_np; ++i) R->hist_c[i] += a->hist[i];
        for (uint32_t i = 0; i < g_np; ++i) {
            R->tail_all[i] += a->tail_all[i];
            if (a->first[i] && (!R->first[i] || a->first[i] < R->first[i])) R->first[i] = a->first[i];
        }
        for (int i = 0; i < 30; ++i) { R->primes30[i] += a->primes30[i]; R->twins30[i] += a->twins30[i]; }
        R->sampled_evens += a->sampled_evens;
        for (size_t k = 0; k < a->nbeyond; ++k) {
            R->bn = realloc(R->bn, (R->nb + 1) * 8);
            R->bp = realloc(R->bp, (R->nb + 1) * 8);
            R->bn[R->nb] = a->beyond_n[k];
            R->bp[R->nb++] = a->beyond_p[k];
        }
    }
    /* beyond in ascending n (deterministic) */
    for (size_t i = 1; i < R->nb; ++i)
        for (size_t j = i; j > 0 && R->bn[j - 1] > R->bn[j]; --j) {
            uint64_t x = R->bn[j]; R->bn[j] = R->bn[j - 1]; R->bn[j - 1] = x;
            x = R->bp[j]; R->bp[j] = R->bp[j - 1]; R->bp[j - 1] = x;
        }
    /* gaps: every prime q in [start + 1, limit - 1], chunks stitched in order
     * (stats_ref.cpp:139-170) */
    const uint64_t qlo = R->start + 1, qhi = R->limit - 1;
    R->have_gaps = qhi >= qlo;
    if (!R->have_gaps) { free(th); return; }
    R->gp = calloc(nt, sizeof(GapTask));
    const uint64_t span = (qhi - qlo) / 2 + 1;
    for (unsigned t = 0; t < nt; ++t) {
        GapTask* c = &R->gp[t];
        c->a = qlo + 2 * (uint64_t)((gb_u128)span * t / nt);
        c->e = qlo + 2 * (uint64_t)((gb_u128)span * (t + 1) / nt);
        c->h = calloc(GAPHMAX, 8);
        c->wh = calloc(GAPHMAX, 8);
        c->w_lo = wlo; c->w_hi = whi; c->nw = nw; c->filter = filter;
        pthread_create(&th[t], NULL, gap_worker, c);
    }
    for (unsigned t = 0; t < nt; ++t) pthread_join(th[t], NULL);
    GapTask G;
    memset(&G, 0, sizeof G);
    G.h = R->gh; G.wh = R->wh; G.w_lo = wlo; G.w_hi = whi; G.nw = nw; G.filter = filter;
    uint64_t glast = 0;
    for (unsigned t = 0; t < nt; ++t) {
        GapTask* c = &R->gp[t];
        if (!c->first) continue;
        uint64_t prev = glast;
        if (!prev) { prev = c->first - 2; while (!gb_is_prime_u64(prev)) prev -= 2; }
        gap_add(&G, c->first - prev, c->first);
        G.n += c->n;
        G.wn += c->wn;
        for (int k = 0; k < 4; ++k) { G.d[k] += c->d[k]; G.wd[k] += c->wd[k]; }
        for (int i = 0; i < GAPHMAX; ++i) { G.h[i] += c->h[i]; G.wh[i] += c->wh[i]; }
        for (int i = 0; i < 4096; ++i)
            if (c->tail_cnt[i]) {
                if (!G.tail_cnt[i] || c->tail_lo[i] < G.tail_lo[i]) G.tail_lo[i] = c->tail_lo[i];
                G.tail_cnt[i] += c->tail_cnt[i];
            }
        if (c->gmax > G.gmax || (c->gmax == G.gmax && c->gmax_lo < G.gmax_lo)) { G.gmax = c->gmax; G.gmax_lo = c->gmax_lo; }
        glast = c->last;
    }
    R->gn = G.n; R->wn = G.wn; R->gmax = G.gmax; R->gmax_lo = G.gmax_lo;
    for (int k = 0; k < 4; ++k) { R->gd[k] = G.d[k]; R->wd[k] = G.wd[k]; }
    memcpy(R->tail_cnt, G.tail_cnt, sizeof G.tail_cnt);
    memcpy(R->tail_lo, G.tail_lo, sizeof G.tail_lo);
    free(th);
}

static const int U8[8] = {1, 7, 11, 13, 17, 19, 23, 29};

static int cmd_ref(uint64_t start, uint64_t limit, unsigned nt)
{
    Run* R = calloc(1, sizeof(Run));
    R->nt = nt; R->start = start; R->limit = limit; R->evens = (limit - start) / 2 + 1;
    run_pass(R, NULL, NULL, 0, NULL, NULL, 0, 0, 0);
    uint64_t hist[1300] = {0}, pmax = 0, np = 0, ntw = 0;
    for (uint32_t cr = 0; cr < 15; ++cr)
        for (uint32_t i = 0; i < g_np; ++i) hist[i] += R->hist_c[cr * g_np + i];
    for (uint32_t i = 0; i < g_np; ++i) if (hist[i]) pmax = g_pr[i];
    for (size_t k = 0; k < R->nb; ++k) if (R->bp[k] > pmax) pmax = R->bp[k];
    for (int i = 0; i < 30; ++i) { np += R->primes30[i]; ntw += R->twins30[i]; }
    printf("evens=%llu pmax=%llu primes=%llu twins=%llu primes30=", (unsigned long long)R->evens,
           (unsigned long long)pmax, (unsigned long long)np, (unsigned long long)ntw);
    for (int i = 0; i < 8; ++i) printf("%s%llu", i ? "," : "", (unsigned long long)R->primes30[U8[i]]);
    printf(" twins30=%llu,%llu,%llu", (unsigned long long)R->twins30[11], (unsigned long long)R->twins30[17],
           (unsigned long long)R->twins30[29]);
    /* first occurrences by p ascending, beyond entries (p > 9781) after the
     * candidates, as the reference's std::map orders them */
    printf(" first=");
    int f1 = 1;
    for (uint32_t i = 0; i < g_np; ++i)
        if (R->first[i]) { printf("%s%u:%llu", f1 ? "" : ";", g_pr[i], (unsigned long long)R->first[i]); f1 = 0; }
    {
        /* beyond: per distinct p the smallest n, ascending p */
        for (size_t k = 0; k < R->nb; ++k) {
            int dup = 0;
            for (size_t j = 0; j < k; ++j) if (R->bp[j] == R->bp[k]) dup = 1;
            if (dup) continue;
            uint64_t mn = R->bn[k];
            for (size_t j = 0; j < R->nb; ++j) if (R->bp[j] == R->bp[k] && R->bn[j] < mn) mn = R->bn[j];
            (void)mn;
        }
        /* sorted output of distinct beyond p */
        uint64_t lastp = 0;
        for (;;) {
            uint64_t bp = 0, bn = 0;
            for (size_t k = 0; k < R->nb; ++k)
                if (R->bp[k] > lastp && (!bp || R->bp[k] < bp || (R->bp[k] == bp && R->bn[k] < bn))) { bp = R->bp[k]; bn = R->bn[k]; }
            if (!bp) break;
            printf("%s%llu:%llu", f1 ? "" : ";", (unsigned long long)bp, (unsigned long long)bn);
            f1 = 0;
            lastp = bp;
        }
    }
    printf(" hist=");
    print_dense(hist, g_np, 0);
    if (R->have_gaps) {
        char m[4][48];
        for (int k = 0; k < 4; ++k) u128_str(R->gd[k], m[k]);
        printf(" gaps=%llu gmom=%s,%s,%s,%s gmax=%llu:%llu gtail=", (unsigned long long)R->gn, m[0], m[1], m[2], m[3],
               (unsigned long long)R->gmax, (unsigned long long)R->gmax_lo);
        int t1 = 1;
        for (int i = 0; i < 4096; ++i)
            if (R->tail_cnt[i]) {
                printf("%s%d:%llu:%llu", t1 ? "" : ";", 2 * i, (unsigned long long)R->tail_cnt[i],
                       (unsigned long long)R->tail_lo[i]);
                t1 = 0;
            }
        if (t1) printf("-");
        printf(" ghist=");
        print_dense(R->gh, GAPHMAX, 0);
    } else {
        printf(" gaps=0 gmom=0,0,0,0 gmax=0:0 gtail=- ghist=");
    }
    printf("\n");
    return 0;
}

/* Segmentation of a run (src/verifier.cpp:597-629 resolve_range, 2489-2512
 * segment_loop; src/host_math.cpp:13-25 segment_geometry). Returns the
 * number of segments; lo/hi get each segment's first / last even. */
static size_t segments_of(uint64_t* start_io, uint64_t* limit_io, uint64_t seg,
                          uint64_t** lo, uint64_t** hi)
{
    uint64_t limit = *limit_io & ~(uint64_t)1, start = *start_io;
    if (start < 4) start = 4;
    if (start & 1) start += 1;
    if (seg < 2) seg = 2;
    if (seg & 1) seg += 1;
    const uint64_t total = (limit - start) / 2 + 1;
    if (seg > total) seg = (total & ~(uint64_t)1) > 2 ? (total & ~(uint64_t)1) : 2;
    if (start == 4 && seg < 4) seg = 4;
    size_t n = 0, cap = 16;
    *lo = malloc(cap * 8); *hi = malloc(cap * 8);
    for (uint64_t s = start; s <= limit; s += seg) {
        const uint64_t e = (limit - s <= seg) ? limit : s + seg - 2;
        if (n == cap) { cap *= 2; *lo = realloc(*lo, cap * 8); *hi = realloc(*hi, cap * 8); }
        (*lo)[n] = s; (*hi)[n++] = e;
        if (e == limit) break;
    }
    *start_io = start; *limit_io = limit;
    return n;
}

static int cmd_gbstats(uint64_t start, uint64_t limit, uint64_t seg, uint32_t sample, const char* gaps, unsigned nt)
{
    uint64_t *lo, *hi;
    const size_t ns = segments_of(&start, &limit, seg, &lo, &hi);
    const int gaps_on = strcmp(gaps, "0") != 0;
    const int gaps_exact = gaps_on && strcmp(gaps, "sampled") != 0;
    /* sampled segments (segment index % sample == 0) as even-index ranges;
     * walked segments as q ranges [seg_start - 3, seg_end - 3] (count30's
     * counting range, gstats.hpp:262-266): a gap belongs to the segment that
     * counts its upper prime; limit - 1 belongs to the last segment */
    uint64_t* slo = malloc(ns * 8); uint64_t* shi = malloc(ns * 8); size_t nsamp = 0;
    uint64_t* wlo = malloc(ns * 8); uint64_t* whi = malloc(ns * 8); size_t nw = 0;
    uint64_t sampled_total = 0, walk_evens = 0;
    for (size_t k = 0; k < ns; ++k) {
        const int sampled = (k % sample) == 0;
        if (sampled) {
            slo[nsamp] = (lo[k] - start) / 2; shi[nsamp++] = (hi[k] - start) / 2;
            sampled_total += (hi[k] - lo[k]) / 2 + 1;
        }
        if (gaps_on && (sampled || gaps_exact)) {
            wlo[nw] = lo[k] - 3;
            whi[nw++] = k + 1 == ns ? limit - 1 : hi[k] - 3;
            walk_evens += (hi[k] - lo[k]) / 2 + 1;
        }
    }
    Run* R = calloc(1, sizeof(Run));
    R->nt = nt; R->start = start; R->limit = limit; R->evens = (limit - start) / 2 + 1;
    run_pass(R, slo, shi, nsamp, wlo, whi, nw, 1, 1);
    uint64_t np = 0, ntw = 0, tail = 0, pmax = 0;
    for (int i = 0; i < 30; ++i) { np += R->primes30[i]; ntw += R->twins30[i]; }
    for (uint32_t i = 0; i < g_np; ++i) tail += R->tail_all[i];
    tail += R->nb;
    for (uint32_t i = 0; i < g_np; ++i) if (R->tail_all[i]) pmax = g_pr[i];
    {
        uint64_t bm = 0;
        for (size_t k = 0; k < R->nb; ++k) if (R->bp[k] > bm) bm = R->bp[k];
        if (bm) pmax = bm;
    }
    printf("start=%llu limit=%llu evens=%llu evens_sampled=%llu tail_t=%u tail_evens=%llu pmax=%llu primes=%llu twins=%llu",
           (unsigned long long)start, (unsigned long long)limit, (unsigned long long)R->evens,
           (unsigned long long)R->sampled_evens, TAIL_T, (unsigned long long)tail, (unsigned long long)pmax,
           (unsigned long long)np, (unsigned long long)ntw);
    printf(" primes30=");
    for (int i = 0; i < 8; ++i) printf("%s%llu", i ? "," : "", (unsigned long long)R->primes30[U8[i]]);
    printf(" twins30=%llu,%llu,%llu", (unsigned long long)R->twins30[11], (unsigned long long)R->twins30[17],
           (unsigned long long)R->twins30[29]);
    printf(" new_first=");
    int f1 = 1;
    for (uint32_t i = 0; i < g_np; ++i)
        if (g_pr[i] >= TAIL_T && R->first[i]) {
            printf("%s%u:%llu", f1 ? "" : ";", g_pr[i], (unsigned long long)R->first[i]);
            f1 = 0;
        }
    if (f1) printf("-");
    printf(" beyond=");
    for (size_t k = 0; k < R->nb; ++k)
        printf("%s%llu:%llu", k ? ";" : "", (unsigned long long)R->bp[k], (unsigned long long)R->bn[k]);
    if (!R->nb) printf("-");
    if (gaps_on) {
        /* gd1 = sum of the run's gaps; gmax exact when walked everywhere or
         * >= 700 (gstats.hpp:434, 974-989) */
        char d1[48], m[4][48];
        u128_str(R->have_gaps ? R->gd[0] : 0, d1);
        for (int k = 0; k < 4; ++k) u128_str(R->wd[k], m[k]);
        printf(" gaps=%llu gd1=%s gmax=", (unsigned long long)np, d1);
        /* the largest gap (first lower prime among equals) is reported when
         * every segment was walked, else only when it is a tail gap (>= 700:
         * every such gap is seen exactly) */
        const uint64_t gm = R->have_gaps ? R->gmax : 0, gl = R->gmax_lo;
        if (gaps_exact ? gm != 0 : gm >= GAP_TAIL) printf("%llu:%llu", (unsigned long long)gm, (unsigned long long)gl);
        else printf("-");
        printf(" gtail=");
        int t1 = 1;
        for (int i = 0; i < 4096; ++i)
            if (R->tail_cnt[i]) {
                printf("%s%d:%llu:%llu", t1 ? "" : ";", 2 * i, (unsigned long long)R->tail_cnt[i],
                       (unsigned long long)R->tail_lo[i]);
                t1 = 0;
            }
        if (t1) printf("-");
        printf(" gwalk=%llu:%llu gsmom=%s,%s,%s,%s gsamp=", (unsigned long long)walk_evens,
               (unsigned long long)R->wn, m[0], m[1], m[2], m[3]);
        if (dense_empty(R->wh, GAP_BINS, 0)) printf("-"); else print_dense(R->wh, GAP_BINS, 0);
    }
    /* tail: by prime index from the first p >= tail_t */
    uint32_t t0 = 0;
    while (t0 < g_np && g_pr[t0] < TAIL_T) ++t0;
    printf(" tail=");
    if (dense_empty(R->tail_all, g_np, t0)) printf("-");
    else { printf("%u:", g_pr[t0]); print_dense(R->tail_all, g_np, t0); }
    printf("\n");
    /* state records of the p(n) histograms: sampled per class (all p),
     * tailcls per class (sampled segments, p >= tail_t), tailagg (the
     * other segments' tail, classes summed) */
    for (uint32_t cr = 0; cr < 15; ++cr) {
        printf("sampled %u ", cr);
        print_dense(&R->hist_c[cr * g_np], g_np, 0);
        printf("\n");
    }
    uint64_t agg[1300];
    memcpy(agg, R->tail_all, sizeof agg);
    for (uint32_t cr = 0; cr < 15; ++cr) {
        uint64_t v[1300] = {0};
        for (uint32_t i = 0; i < g_np; ++i)
            if (g_pr[i] >= TAIL_T) { v[i] = R->hist_c[cr * g_np + i]; agg[i] -= v[i]; }
        printf("tailcls %u ", cr);
        print_dense(v, g_np, 0);
        printf("\n");
    }
    printf("tailagg ");
    print_dense(agg, g_np, 0);
    printf("\n");
    if (R->sampled_evens != sampled_total) { fprintf(stderr, "sampled evens mismatch\n"); return 1; }
    return 0;
}

/* ---------------------------------------------------------------- ridges */

typedef struct {
    uint64_t start, evens, floor;
    unsigned nt, t;
    uint64_t* n;
    uint64_t* p;
    size_t cnt, cap;
} RidgeTask;

static void* ridge_worker(void* arg)
{
    RidgeTask* a = (RidgeTask*)arg;
    for (uint64_t i = a->t; i < a->evens; i += a->nt) {
        const uint64_t n = a->start + 2 * i;
        const uint64_t p = pn_scan(n);
        if (p < a->floor) continue;
        if (a->cnt == a->cap) {
            a->cap = a->cap ? 2 * a->cap : 256;
            a->n = realloc(a->n, a->cap * sizeof *a->n);
            a->p = realloc(a->p, a->cap * sizeof *a->p);
        }
        a->n[a->cnt] = n;
        a->p[a->cnt++] = p;
    }
    return NULL;
}

typedef struct { uint64_t n, p; } Ridge;
static int ridge_cmp(const void* x, const void* y)
{
    const Ridge* a = x;
    const Ridge* b = y;
    if (a->p != b->p) return a->p > b->p ? -1 : 1;
    return a->n > b->n ? -1 : a->n < b->n;
}

static int cmd_ridges(uint64_t start, uint64_t limit, uint64_t floor, size_t k, unsigned nt)
{
    RidgeTask* t = calloc(nt, sizeof *t);
    pthread_t* th = calloc(nt, sizeof *th);
    for (unsigned i = 0; i < nt; ++i) {
        t[i].start = start;
        t[i].evens = (limit - start) / 2 + 1;
        t[i].floor = floor;
        t[i].nt = nt;
        t[i].t = i;
        pthread_create(&th[i], NULL, ridge_worker, &t[i]);
    }
    size_t total = 0;
    for (unsigned i = 0; i < nt; ++i) { pthread_join(th[i], NULL); total += t[i].cnt; }
    Ridge* r = malloc((total + 1) * sizeof *r);
    size_t m = 0;
    for (unsigned i = 0; i < nt; ++i)
        for (size_t j = 0; j < t[i].cnt; ++j) { r[m].n = t[i].n[j]; r[m++].p = t[i].p[j]; }
    qsort(r, m, sizeof *r, ridge_cmp);
    for (size_t j = 0; j < m && j < k; ++j) printf("%llu %llu\n", (unsigned long long)r[j].n, (unsigned long long)r[j].p);
    return 0;
}

int main(int argc, char** argv)
{
    init_primes();
    if (argc >= 3 && !strcmp(argv[1], "pn")) {
        const uint64_t n = strtoull(argv[2], NULL, 10);
        printf("%llu\n", (unsigned long long)pn_exact(n));
        return 0;
    }
    if (argc >= 4 && !strcmp(argv[1], "ref")) {
        const uint64_t s = strtoull(argv[2], NULL, 10), l = strtoull(argv[3], NULL, 10);
        const unsigned nt = argc > 4 ? (unsigned)atoi(argv[4]) : 8;
        if (s % 2 || l % 2 || s < 1000000 || l < s) { fprintf(stderr, "bad range\n"); return 2; }
        return cmd_ref(s, l, nt ? nt : 1);
    }
    if (argc >= 7 && !strcmp(argv[1], "gbstats")) {
        const uint64_t s = strtoull(argv[2], NULL, 10), l = strtoull(argv[3], NULL, 10);
        const uint64_t seg = strtoull(argv[4], NULL, 10);
        const uint32_t sample = (uint32_t)strtoul(argv[5], NULL, 10);
        const unsigned nt = argc > 7 ? (unsigned)atoi(argv[7]) : 8;
        if (s < 1000000 || l < s || !sample) { fprintf(stderr, "bad range\n"); return 2; }
        return cmd_gbstats(s, l, seg, sample, argv[6], nt ? nt : 1);
    }
    if (argc >= 6 && !strcmp(argv[1], "ridges")) {
        const uint64_t s = strtoull(argv[2], NULL, 10), l = strtoull(argv[3], NULL, 10);
        const unsigned nt = argc > 6 ? (unsigned)atoi(argv[6]) : 8;
        if (s % 2 || l % 2 || s < 1000000 || l < s) { fprintf(stderr, "bad range\n"); return 2; }
        return cmd_ridges(s, l, strtoull(argv[4], NULL, 10), (size_t)strtoull(argv[5], NULL, 10), nt ? nt : 1);
    }
    if (argc >= 3 && !strcmp(argv[1], "isprime")) {
        /* one number per line on stdin -> 0/1 per line (the MR pin) */
        char line[64];
        while (fgets(line, sizeof line, stdin)) printf("%d\n", gb_is_prime_u64(strtoull(line, NULL, 10)));
        return 0;
    }
    fprintf(stderr, "usage: gb_oracle ref START LIMIT [THREADS] | gbstats START LIMIT SEGMENT SAMPLE GAPS [THREADS] | "
                    "ridges START LIMIT FLOOR K [THREADS] | pn N | isprime -\n");
    return 2;
}
