/*
 * workloads.c -- seeded synthetic inputs for the five benchmark workloads
 * (libbl_workloads.so, used by bench.py and the tests; no distance arithmetic).
 *
 * Every string is generated from its own counter-based stream (splitmix64 of
 * seed, set and index), so the output is identical for any thread count and
 * any item can be regenerated alone (a choice derived from query q regenerates
 * q). Two passes: blw_lengths() sizes a set, blw_fill() writes it at the
 * offsets the caller computed.
 *
 * Workloads (set 0 = a / queries, set 1 = b / choices):
 *   1  pairwise L32: a = 32 bytes over ACGT; b = a edited per position with
 *      p = 0.10 (substitution / insertion / deletion equally likely; the new
 *      byte is uniform over ACGT, so a substitution may leave the byte as is).
 *   2  cdist L64 search: queries and choices 64 bytes over 'a'-'z'; 30% of
 *      choices copy a uniformly chosen query and substitute e ~ U{0..8}
 *      distinct positions with a different letter (lengths stay 64).
 *   3  pairwise uneven: len(a) ~ U{1..256}, bytes U{0..255}; b = a edited per
 *      position at a per-pair rate r ~ U[0.05, 0.30] (sub/ins/del equal).
 *   4  long-pattern cdist: lengths ~ U{200..256} over ACGT; 10% of choices copy
 *      a random query, substitute each position with p = 0.01 (different
 *      base), then apply U{0..4} indels (insert or delete, equally likely).
 *   5  resident set for serving: lengths ~ U{16..128} over 'a'-'z'.
 *   6  serving pairs: len(a) ~ U{8..256} over 'a'-'z', b = a edited per
 *      position at r ~ U[0.05, 0.30].
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { uint64_t s; } rng_t;

static inline uint64_t sm64(uint64_t *x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline rng_t stream(uint64_t seed, int set, int64_t i)
{
    uint64_t x = seed ^ ((uint64_t)(set + 1) << 56) ^ (uint64_t)i * 0xD1B54A32D192ED03ull;
    rng_t r = {sm64(&x)};
    return r;
}

static inline uint64_t nx(rng_t *r) { return sm64(&r->s); }
/* uniform integer in [0, n) */
static inline uint32_t below(rng_t *r, uint32_t n) { return (uint32_t)(((nx(r) >> 32) * n) >> 32); }
/* uniform double in [0, 1) */
static inline double unit(rng_t *r) { return (double)(nx(r) >> 11) * (1.0 / 9007199254740992.0); }

static const uint8_t ACGT[4] = {'A', 'C', 'G', 'T'};
#define MAXLEN 1024

static int rand_string(rng_t *r, uint8_t *o, int len, const uint8_t *alpha, int na, int lo)
{
    for (int i = 0; i < len; i++) o[i] = alpha ? alpha[below(r, (uint32_t)na)] : (uint8_t)(lo + below(r, (uint32_t)na));
    return len;
}

/* Per-position edit process: with probability p one of sub / ins / del. */
static int edit_rate(rng_t *r, const uint8_t *a, int la, uint8_t *o, double p,
                     const uint8_t *alpha, int na, int lo)
{
    int n = 0;
    for (int i = 0; i < la; i++) {
        if (unit(r) < p) {
            const uint32_t op = below(r, 3);
            const uint8_t c = alpha ? alpha[below(r, (uint32_t)na)] : (uint8_t)(lo + below(r, (uint32_t)na));
            if (op == 0) o[n++] = c;                       /* substitution */
            else if (op == 1) { o[n++] = c; o[n++] = a[i]; } /* insertion before a[i] */
            /* op == 2: deletion */
        } else {
            o[n++] = a[i];
        }
    }
    return n;
}

static int gen(int cfg, uint64_t seed, int set, int64_t i, int64_t nq, uint8_t *o)
{
    uint8_t tmp[MAXLEN];
    rng_t r = stream(seed, set, i);
    switch (cfg) {
    case 1: {
        if (set == 0) return rand_string(&r, o, 32, ACGT, 4, 0);
        int la = gen(cfg, seed, 0, i, nq, tmp);
        return edit_rate(&r, tmp, la, o, 0.10, ACGT, 4, 0);
    }
    case 2: {
        if (set == 0 || unit(&r) >= 0.30) return rand_string(&r, o, 64, NULL, 26, 'a');
        const int64_t q = (int64_t)below(&r, (uint32_t)nq);
        gen(cfg, seed, 0, q, nq, o);
        const int e = (int)below(&r, 9);
        uint8_t pos[64];
        for (int j = 0; j < 64; j++) pos[j] = (uint8_t)j;
        for (int j = 0; j < e; j++) {                 /* e distinct positions */
            const int x = j + (int)below(&r, (uint32_t)(64 - j));
            const uint8_t t = pos[j]; pos[j] = pos[x]; pos[x] = t;
            const uint8_t old = o[pos[j]];
            uint8_t c = (uint8_t)('a' + below(&r, 25));
            if (c >= old) c++;                          /* a different letter */
            o[pos[j]] = c;
        }
        return 64;
    }
    case 3: {
        if (set == 0) return rand_string(&r, o, 1 + (int)below(&r, 256), NULL, 256, 0);
        const int la = gen(cfg, seed, 0, i, nq, tmp);
        const double rate = 0.05 + 0.25 * unit(&r);
        return edit_rate(&r, tmp, la, o, rate, NULL, 256, 0);
    }
    case 4: {
        if (set == 0 || unit(&r) >= 0.10) return rand_string(&r, o, 200 + (int)below(&r, 57), ACGT, 4, 0);
        const int64_t q = (int64_t)below(&r, (uint32_t)nq);
        int n = gen(cfg, seed, 0, q, nq, tmp);
        for (int j = 0; j < n; j++)
            if (unit(&r) < 0.01) {                     /* substitute with a different base */
                int old = 0;
                while (old < 3 && ACGT[old] != tmp[j]) old++;
                int c = (int)below(&r, 3);
                if (c >= old) c++;
                tmp[j] = ACGT[c];
            }
        const int ind = (int)below(&r, 5);
        for (int j = 0; j < ind; j++) {
            if (below(&r, 2) == 0) {                    /* insertion */
                const int at = (int)below(&r, (uint32_t)(n + 1));
                memmove(tmp + at + 1, tmp + at, (size_t)(n - at));
                tmp[at] = ACGT[below(&r, 4)];
                n++;
            } else if (n > 0) {                         /* deletion */
                const int at = (int)below(&r, (uint32_t)n);
                memmove(tmp + at, tmp + at + 1, (size_t)(n - at - 1));
                n--;
            }
        }
        memcpy(o, tmp, (size_t)n);
        return n;
    }
    case 5:
        return rand_string(&r, o, 16 + (int)below(&r, 113), NULL, 26, 'a');
    case 6: {
        if (set == 0) return rand_string(&r, o, 8 + (int)below(&r, 249), NULL, 26, 'a');
        const int la = gen(cfg, seed, 0, i, nq, tmp);
        const double rate = 0.05 + 0.25 * unit(&r);
        return edit_rate(&r, tmp, la, o, rate, NULL, 26, 'a');
    }
    default:
        return -1;
    }
}

typedef struct {
    int cfg, set;
    uint64_t seed;
    int64_t i0, i1, nq;
    int64_t *len;           /* pass 1 */
    const int64_t *off;     /* pass 2 */
    uint8_t *buf;
    int err;
} job_t;

static void *work(void *arg)
{
    job_t *J = (job_t *)arg;
    uint8_t tmp[MAXLEN];
    for (int64_t i = J->i0; i < J->i1; i++) {
        if (J->len) {
            const int n = gen(J->cfg, J->seed, J->set, i, J->nq, tmp);
            if (n < 0) { J->err = -2; return NULL; }
            J->len[i] = n;
        } else {
            const int n = gen(J->cfg, J->seed, J->set, i, J->nq, J->buf + J->off[i]);
            if (n != J->off[i + 1] - J->off[i]) { J->err = -3; return NULL; }
        }
    }
    return NULL;
}

static int run(int cfg, uint64_t seed, int set, int64_t n, int64_t nq, int64_t *len,
               const int64_t *off, uint8_t *buf, int threads)
{
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    job_t J[256];
    pthread_t th[256];
    int started = 0, err = 0;
    for (int t = 0; t < threads; t++) {
        J[t] = (job_t){cfg, set, seed, n * t / threads, n * (t + 1) / threads, nq, len, off, buf, 0};
        if (pthread_create(&th[t], NULL, work, &J[t]) == 0) started++;
        else { work(&J[t]); }
    }
    for (int t = 0; t < started; t++) pthread_join(th[t], NULL);
    for (int t = 0; t < threads; t++) if (J[t].err) err = J[t].err;
    return err;
}

/* Pass 1: lengths of strings [0, n) of a set. nq = number of queries (the
 * population derived choices draw from). Returns 0 or a negative error. */
int blw_lengths(int cfg, uint64_t seed, int set, int64_t n, int64_t nq, int64_t *lengths, int threads)
{
    return run(cfg, seed, set, n, nq, lengths, NULL, NULL, threads);
}

/* Pass 2: the strings themselves at buf[off[i] : off[i+1]]. */
int blw_fill(int cfg, uint64_t seed, int set, int64_t n, int64_t nq, const int64_t *off, uint8_t *buf,
             int threads)
{
    return run(cfg, seed, set, n, nq, NULL, off, buf, threads);
}
