/*
 * oracle/myers_oracle.c -- CPU restatement of bitlev's exact edit-distance core.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT. This is the checker the HIP kernels are
 * compared against. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it (through oracle/oracle.py). The shipped library
 * (libbitlev_hip.so) never links, calls or falls back to anything here.
 *
 * What it restates (paths relative to karims/editdist-bench):
 *   orc_step           csrc/myers.c:22-36        one Myers/Hyyro block step
 *   orc_peq_*          csrc/myers.c:39-55        256-row Peq table, cleared after use
 *   orc_scan           csrc/myers.c:58-77        run_text (exact)
 *                      csrc/myers.c:107-130      run_text_k (threshold, two exits)
 *   orc_trim           csrc/myers.c:80-87        exact common prefix/suffix trim
 *   orc_pair           csrc/myers.c:90-99, :132-143   pair_distance / pair_distance_k
 *   orc_pairs          csrc/edpy.c:59-88 (_edpy.pairs), csrc/myers.c:157-194
 *   orc_cross          csrc/edpy.c:91-127 (_edpy.cross), csrc/myers.c:198-225
 *   orc_dp             bench/reference.py:16-28  Wagner-Fischer
 *   orc_saturate       bitlev/_common.py:25-30, csrc/gpu_myers.cuh:149-152
 *
 * Semantics that must hold bit for bit:
 *   k < 0  -> exact distance; k >= 0 -> min(d, k + 1).
 *   Pattern = the shorter string (after trimming, pairs only); the text is the
 *   other one. Cross does not trim and does not swap: the query is the pattern.
 *   An empty cross query gives n, or k + 1 when k >= 0 and n > k.
 *
 * The band-column counters (orc_*_cols) report how many text columns the
 * reference's threshold exit rule evaluates; that count defines the useful
 * work of the thresholded benchmark configurations.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef uint64_t u64;

/* ORC_MUTANT (tests/test_oracle.py only): compile-time planted mistakes, each of
 * which the golden vectors must catch. 0 (default) is the real oracle. */
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

/* ---------------------------------------------------------------- column step */
/* One column of one 64-bit block. `hin` is the horizontal delta entering the
 * block's lowest row (-1, 0 or +1); the delta leaving row `bit` is returned.
 * Same recurrence as csrc/myers.c:22-36. */
static inline int orc_step(u64 *pv, u64 *mv, u64 eq, int hin, unsigned bit)
{
    const u64 P = *pv, M = *mv;
    const u64 neg = (u64)(hin < 0), pos = (u64)(hin > 0);
    const u64 xv = eq | M;
    const u64 e = eq | neg;
#if ORC_MUTANT == 1
    const u64 xh = (((e & P) + P) ^ P);                 /* dropped term */
#else
    const u64 xh = (((e & P) + P) ^ P) | e;
#endif
    u64 ph = M | ~(xh | P);
    u64 mh = P & xh;
#if ORC_MUTANT == 2
    bit = 63;                                           /* wrong output row */
#endif
    const int hout = (int)((ph >> bit) & 1u) - (int)((mh >> bit) & 1u);
#if ORC_MUTANT == 3
    ph = (ph << 1) | neg;                               /* carries swapped */
    mh = (mh << 1) | pos;
#else
    ph = (ph << 1) | pos;
    mh = (mh << 1) | neg;
#endif
    *pv = mh | ~(xv | ph);
    *mv = ph & xv;
    return hout;
}

/* ---------------------------------------------------------------- Peq table */
typedef struct {
    u64 *tab;    /* tab[c * words + w]: bit i of word w set <=> p[64w + i] == c */
    int words;
    u64 *pv, *mv;
} orc_peq;

static int orc_peq_init(orc_peq *q, int words)
{
    q->words = words < 1 ? 1 : words;
    q->tab = (u64 *)calloc((size_t)256 * q->words, sizeof(u64));
    q->pv = (u64 *)malloc(sizeof(u64) * q->words);
    q->mv = (u64 *)malloc(sizeof(u64) * q->words);
    if (!q->tab || !q->pv || !q->mv) {
        free(q->tab); free(q->pv); free(q->mv);
        q->tab = q->pv = q->mv = NULL;
        return -1;
    }
    return 0;
}

static void orc_peq_free(orc_peq *q)
{
    free(q->tab); free(q->pv); free(q->mv);
    q->tab = q->pv = q->mv = NULL;
}

static void orc_peq_load(orc_peq *q, const uint8_t *p, int m)
{
    for (int i = 0; i < m; i++)
        q->tab[(size_t)p[i] * q->words + (i >> 6)] |= (u64)1 << (i & 63);
}

/* Only the rows the pattern touched are non-zero; zero them again. */
static void orc_peq_unload(orc_peq *q, const uint8_t *p, int m)
{
    for (int i = 0; i < m; i++)
        q->tab[(size_t)p[i] * q->words + (i >> 6)] = 0;
}

/* ---------------------------------------------------------------- text scan */
/* Distance of the loaded pattern (length m >= 1) against text t[0..n).
 * k < 0: exact (run_text, csrc/myers.c:58-77).
 * k >= 0: run_text_k (csrc/myers.c:107-130): |n - m| > k answers k + 1 with no
 * column evaluated; after column j, score - (n - 1 - j) > k answers k + 1.
 * *cols (if given) receives the number of text columns evaluated. */
/* Lower bound on the final distance after jj columns: the DP cell on the
 * final diagonal, D[jj + m - n][jj] (n - m when that row is <= 0). Used only
 * to count the columns an exit test on this bound would evaluate. */
static int orc_diag_bound(const orc_peq *q, int m, int n, int jj)
{
    const int i0 = jj + m - n;
    if (i0 <= 0) return n - m;
    int s = jj;
    for (int w = 0; w * 64 < i0; w++) {
        const int r = i0 - 64 * w;
        const u64 mk = r >= 64 ? ~(u64)0 : (((u64)1 << r) - 1);
        s += __builtin_popcountll(q->pv[w] & mk) - __builtin_popcountll(q->mv[w] & mk);
    }
    return s;
}

/* What *cols counts: 0 = the columns the reference's bottom-row exit test
 * evaluates; 1 = the columns the final-diagonal bound needs (same results,
 * earlier exits); 2 = the DP cells the row-staged scan needs under the
 * final-diagonal bound (staged_cells below). */
static int g_count_rule = 0;
void orc_set_count_rule(int rule) { g_count_rule = rule; }

/* Cells of a scan in row stages (32 rows, then 64, then all m rows; each
 * stage restarts at column 0 and can test the bound at column j only while
 * its diagonal row j + m - n lies within the stage) whose bound first exceeds
 * k after column s (s = n: never). k < 0: one whole scan of m x n cells. */
static long long staged_cells(int m, int n, int k, int s)
{
    if (k < 0) return (long long)m * n;
    long long cells = 0;
    static const int R[2] = {32, 64};
    for (int t = 0; t < 2; t++) {
        if (R[t] >= m) break;
        const int J = R[t] - (m - n);          /* last column this stage can test */
        if (J < 1) continue;
        if (s <= J) return cells + (long long)s * R[t];
        cells += (long long)J * R[t];
    }
    return cells + (long long)s * m;
}

static int orc_scan(orc_peq *q, int m, const uint8_t *t, int n, int k, long long *cols)
{
    if (cols) *cols = 0;
    if (k >= 0 && (n > m ? n - m : m - n) > k) return k + 1;
    const int W = (m + 63) >> 6;
    const unsigned last = (unsigned)((m - 1) & 63);
    for (int w = 0; w < W; w++) { q->pv[w] = ~(u64)0; q->mv[w] = 0; }
    int score = m;
    for (int j = 0; j < n; j++) {
        const u64 *row = q->tab + (size_t)t[j] * q->words;
        int h = 1;                               /* row 0 of a global alignment: +1 per column */
        for (int w = 0; w + 1 < W; w++) h = orc_step(&q->pv[w], &q->mv[w], row[w], h, 63);
        score += orc_step(&q->pv[W - 1], &q->mv[W - 1], row[W - 1], h, last);
        if (cols && *cols >= 0) {
            *cols = j + 1;
            if (g_count_rule >= 1 && k >= 0 && orc_diag_bound(q, m, n, j + 1) > k)
                *cols = -(j + 1);               /* diagonal rule fired here; keep scanning */
        }
#if ORC_MUTANT == 4
        if (k >= 0 && score - (n - 2 - j) > k) return k + 1;   /* exit rule off by one */
#else
        if (k >= 0 && score - (n - 1 - j) > k) {
            if (cols && *cols < 0) *cols = -*cols;
            if (cols && g_count_rule == 2) *cols = staged_cells(m, n, k, (int)*cols);
            return k + 1;
        }
#endif
    }
    if (cols && *cols < 0) *cols = -*cols;
    if (cols && g_count_rule == 2) *cols = staged_cells(m, n, k, (int)*cols);
    return score;
}

/* ---------------------------------------------------------------- pairs */
/* d(xPy, xQy) == d(P, Q): drop the common prefix, then the common suffix of
 * what is left (csrc/myers.c:80-87). */
static void orc_trim(const uint8_t **a, int *la, const uint8_t **b, int *lb)
{
    const int n = *la < *lb ? *la : *lb;
    int p = 0;
    while (p < n && (*a)[p] == (*b)[p]) p++;
    int s = 0;
    while (s < n - p && (*a)[*la - 1 - s] == (*b)[*lb - 1 - s]) s++;
    *a += p; *b += p;
    *la -= p + s; *lb -= p + s;
}

/* One pair (pair_distance_k, csrc/myers.c:132-143; k < 0 is pair_distance,
 * :90-99). The table must hold >= ceil(min(la, lb) / 64) words. */
static int orc_pair(orc_peq *q, const uint8_t *a, int la, const uint8_t *b, int lb,
                    int trim, int k, long long *cols)
{
    if (cols) *cols = 0;
    if (trim) orc_trim(&a, &la, &b, &lb);
    if (la > lb) {
        const uint8_t *t = a; a = b; b = t;
        int x = la; la = lb; lb = x;
    }
#if ORC_MUTANT == 5
    if (k >= 0 && lb - la >= k) return k + 1;          /* gap test off by one */
#else
    if (k >= 0 && lb - la > k) return k + 1;
#endif
    if (la == 0) return lb;
    orc_peq_load(q, a, la);
    const int d = orc_scan(q, la, b, lb, k, cols);
    orc_peq_unload(q, a, la);
    return d;
}

static int words_for(int64_t len) { return (int)((len + 63) >> 6); }

/* ---------------------------------------------------------------- drivers */
/* Packed layout: string i of a set is buf[off[i] : off[i + 1]]. */
typedef struct {
    const uint8_t *ba; const int64_t *oa;
    const uint8_t *bb; const int64_t *ob;
    int64_t n; int32_t *out; int trim; int k; long long *cols;
    int words; int64_t chunk; int64_t next; pthread_mutex_t mu; int err;
} pairs_job;

static int64_t take(int64_t *next, pthread_mutex_t *mu, int64_t chunk, int64_t n)
{
    pthread_mutex_lock(mu);
    int64_t i = *next;
    *next = i + chunk < n ? i + chunk : n;
    pthread_mutex_unlock(mu);
    return i;
}

static void *pairs_worker(void *arg)
{
    pairs_job *J = (pairs_job *)arg;
    orc_peq q;
    if (orc_peq_init(&q, J->words)) { J->err = -1; return NULL; }
    for (;;) {
        int64_t i0 = take(&J->next, &J->mu, J->chunk, J->n);
        if (i0 >= J->n) break;
        int64_t i1 = i0 + J->chunk < J->n ? i0 + J->chunk : J->n;
        for (int64_t i = i0; i < i1; i++) {
            long long c = 0;
            J->out[i] = orc_pair(&q, J->ba + J->oa[i], (int)(J->oa[i + 1] - J->oa[i]),
                                 J->bb + J->ob[i], (int)(J->ob[i + 1] - J->ob[i]),
                                 J->trim, J->k, J->cols ? &c : NULL);
            if (J->cols) J->cols[i] = c;
        }
    }
    orc_peq_free(&q);
    return NULL;
}

/* out[i] = d(a_i, b_i) (min(d, k + 1) for k >= 0), as _edpy.pairs
 * (csrc/edpy.c:59-88) with trim; trim = 0 is ed_pairs (csrc/myers.c:190-194).
 * cols: optional per-pair count of evaluated text columns. threads >= 1.
 * Returns 0, -1 (allocation), -2 (bad argument). */
int orc_pairs(const uint8_t *ba, const int64_t *oa, const uint8_t *bb, const int64_t *ob,
              int64_t n, int32_t *out, int trim, int k, long long *cols, int threads)
{
    if (n < 0 || threads < 1 || (n > 0 && (!oa || !ob || !out))) return -2;
    if (n == 0) return 0;
    int64_t mx = 0;
    for (int64_t i = 0; i < n; i++) {
        int64_t la = oa[i + 1] - oa[i], lb = ob[i + 1] - ob[i];
        int64_t m = la < lb ? la : lb;
        if (m > mx) mx = m;
    }
    pairs_job J = {ba, oa, bb, ob, n, out, trim, k, cols, words_for(mx), 256, 0,
                   PTHREAD_MUTEX_INITIALIZER, 0};
    if (threads == 1) { pairs_worker(&J); return J.err; }
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    if (!th) return -1;
    int started = 0;
    for (; started < threads; started++)
        if (pthread_create(&th[started], NULL, pairs_worker, &J)) break;
    if (started == 0) pairs_worker(&J);
    for (int t = 0; t < started; t++) pthread_join(th[t], NULL);
    free(th);
    return J.err;
}

typedef struct {
    const uint8_t *bq; const int64_t *oq; int64_t nq;
    const uint8_t *bc; const int64_t *oc; int64_t nc;
    int32_t *out; int k; long long *cols;
    int words; int64_t tile; int64_t tiles_per_row; int64_t next; pthread_mutex_t mu; int err;
} cross_job;

static void *cross_worker(void *arg)
{
    cross_job *J = (cross_job *)arg;
    orc_peq q;
    if (orc_peq_init(&q, J->words)) { J->err = -1; return NULL; }
    const int64_t total = J->nq * J->tiles_per_row;
    for (;;) {
        int64_t t = take(&J->next, &J->mu, 1, total);
        if (t >= total) break;
        const int64_t qi = t / J->tiles_per_row;
        const int64_t c0 = (t % J->tiles_per_row) * J->tile;
        const int64_t c1 = c0 + J->tile < J->nc ? c0 + J->tile : J->nc;
        const uint8_t *p = J->bq + J->oq[qi];
        const int m = (int)(J->oq[qi + 1] - J->oq[qi]);
        int32_t *row = J->out + qi * J->nc;
        long long *crow = J->cols ? J->cols + qi * J->nc : NULL;
        if (m == 0) {                    /* csrc/edpy.c:111-115 */
            for (int64_t c = c0; c < c1; c++) {
                const int64_t n = J->oc[c + 1] - J->oc[c];
#if ORC_MUTANT == 6
                row[c] = (int32_t)n;                               /* k cap forgotten */
#else
                row[c] = (J->k >= 0 && n > J->k) ? J->k + 1 : (int32_t)n;
#endif
                if (crow) crow[c] = 0;
            }
            continue;
        }
        orc_peq_load(&q, p, m);
        for (int64_t c = c0; c < c1; c++) {
            long long cc = 0;
            row[c] = orc_scan(&q, m, J->bc + J->oc[c], (int)(J->oc[c + 1] - J->oc[c]), J->k,
                              crow ? &cc : NULL);
            if (crow) crow[c] = cc;
        }
        orc_peq_unload(&q, p, m);
    }
    orc_peq_free(&q);
    return NULL;
}

/* Row-major nq x nc matrix, as _edpy.cross (csrc/edpy.c:91-127): no trim, the
 * query is the pattern. Work is split into (query, column tile) units. */
int orc_cross(const uint8_t *bq, const int64_t *oq, int64_t nq,
              const uint8_t *bc, const int64_t *oc, int64_t nc,
              int32_t *out, int k, long long *cols, int threads)
{
    if (nq < 0 || nc < 0 || threads < 1) return -2;
    if (nq == 0 || nc == 0) return 0;
    if (!oq || !oc || !out) return -2;
    int64_t mx = 0;
    for (int64_t i = 0; i < nq; i++) {
        int64_t l = oq[i + 1] - oq[i];
        if (l > mx) mx = l;
    }
    const int64_t tile = 4096;
    cross_job J = {bq, oq, nq, bc, oc, nc, out, k, cols, words_for(mx), tile,
                   (nc + tile - 1) / tile, 0, PTHREAD_MUTEX_INITIALIZER, 0};
    if (threads == 1) { cross_worker(&J); return J.err; }
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    if (!th) return -1;
    int started = 0;
    for (; started < threads; started++)
        if (pthread_create(&th[started], NULL, cross_worker, &J)) break;
    if (started == 0) cross_worker(&J);
    for (int t = 0; t < started; t++) pthread_join(th[t], NULL);
    free(th);
    return J.err;
}

/* ---------------------------------------------------------------- Wagner-Fischer */
/* Textbook O(la * lb) DP with one row (bench/reference.py:16-28). Independent
 * of everything above; used to pin the bit-parallel restatement. */
int orc_dp(const uint8_t *a, int la, const uint8_t *b, int lb)
{
    if (la < lb) { const uint8_t *t = a; a = b; b = t; int x = la; la = lb; lb = x; }
    int *row = (int *)malloc(sizeof(int) * (size_t)(lb + 1));
    if (!row) return -1;
    for (int j = 0; j <= lb; j++) row[j] = j;
    for (int i = 1; i <= la; i++) {
        int diag = row[0];
        row[0] = i;
        for (int j = 1; j <= lb; j++) {
            const int up = row[j];
            int v = diag + (a[i - 1] != b[j - 1]);
            if (up + 1 < v) v = up + 1;
            if (row[j - 1] + 1 < v) v = row[j - 1] + 1;
            row[j] = v;
            diag = up;
        }
    }
    const int d = row[lb];
    free(row);
    return d;
}

/* DP over packed pairs: out[i] = d (min(d, k + 1) for k >= 0). */
int orc_dp_pairs(const uint8_t *ba, const int64_t *oa, const uint8_t *bb, const int64_t *ob,
                 int64_t n, int32_t *out, int k)
{
    for (int64_t i = 0; i < n; i++) {
        int d = orc_dp(ba + oa[i], (int)(oa[i + 1] - oa[i]), bb + ob[i], (int)(ob[i + 1] - ob[i]));
        if (d < 0) return -1;
        out[i] = (k >= 0 && d > k) ? k + 1 : d;
    }
    return 0;
}

/* ---------------------------------------------------------------- output types */
/* dtype codes: 0 int32, 1 uint16, 2 uint8. Narrow types saturate at their
 * maximum after the k cap (bitlev/_common.py:25-30; csrc/gpu_myers.cuh:149-152). */
int orc_saturate(const int32_t *in, int64_t n, int dtype, void *out)
{
    if (dtype == 0) { memcpy(out, in, sizeof(int32_t) * (size_t)n); return 0; }
    if (dtype == 1) {
        uint16_t *o = (uint16_t *)out;
        for (int64_t i = 0; i < n; i++) o[i] = (uint16_t)(in[i] > 65535 ? 65535 : in[i]);
        return 0;
    }
    if (dtype == 2) {
        uint8_t *o = (uint8_t *)out;
#if ORC_MUTANT == 7
        for (int64_t i = 0; i < n; i++) o[i] = (uint8_t)in[i];      /* wraps, no saturation */
#else
        for (int64_t i = 0; i < n; i++) o[i] = (uint8_t)(in[i] > 255 ? 255 : in[i]);
#endif
        return 0;
    }
    return -2;
}
