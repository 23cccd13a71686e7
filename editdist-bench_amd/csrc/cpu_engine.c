/*
 * cpu_engine.c -- the package's CPU engine, behind the reference's CPU
 * methods (bitlev.pairwise / bitlev.cdist with method="scalar" or "simd",
 * bitlev/__init__.py:53-101 there). The GPU API never calls it: it runs only
 * when a caller names one of those methods.
 *
 *   scalar  one pair at a time: exact prefix/suffix trim, the shorter string
 *           as the pattern, 64-row words with carries between them, and the
 *           two exact threshold exits (length gap; score minus the columns
 *           left). Results as the reference's C core (csrc/myers.c:22-225).
 *   simd    x86 AVX2 inter-pair kernels (the reference has NEON only,
 *           csrc/myers.c:227-633, and none on x86, README.md:418):
 *           cdist packs the choices as patterns, 16, 8 or 4 per 256-bit
 *           register (16-, 32- or 64-bit lanes, choices of at most 16, 32 or
 *           64 bytes), and runs the query as the text all lanes share, so
 *           the lanes advance in lockstep and a column's Eq words for every
 *           lane are one 32-byte load from the group's table; two groups are
 *           stepped together for instruction-level parallelism, and a group
 *           stops once every lane's lower bound exceeds k. pairwise packs 4
 *           pairs per register (patterns of at most 64 bytes after trimming),
 *           each lane's Eq gathered from its own table. Longer strings take
 *           the scalar path.
 *
 * Distances are symmetric, so swapping pattern and text never changes a
 * result. k < 0: exact; k >= 0: min(d, k + 1). Return 0, -1 (allocation),
 * -2 (invalid argument).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "cpu_engine.h"

typedef uint8_t u8;
typedef uint16_t u16;
typedef uint32_t u32;
typedef uint64_t u64;

/* ------------------------------------------------------------------ scalar */

/* Peq rows of a pattern, 256 x W words, kept all-zero between patterns. */
typedef struct {
    u64 *peq;
    int W;
    u64 *pv, *mv;
} scan_ws;

static int ws_reserve(scan_ws *ws, int W)
{
    if (W <= ws->W) return 0;
    free(ws->peq);
    free(ws->pv);
    ws->peq = calloc((size_t)256 * (size_t)W, sizeof(u64));
    ws->pv = malloc(sizeof(u64) * (size_t)W);
    ws->mv = malloc(sizeof(u64) * (size_t)W);
    if (!ws->peq || !ws->pv || !ws->mv) {
        free(ws->peq);
        free(ws->pv);
        free(ws->mv);
        ws->peq = ws->pv = ws->mv = NULL;
        ws->W = 0;
        return -1;
    }
    ws->W = W;
    return 0;
}

static void ws_release(scan_ws *ws)
{
    free(ws->peq);
    free(ws->pv);
    free(ws->mv);
}

static void peq_set(scan_ws *ws, const u8 *p, int m, int W)
{
    for (int i = 0; i < m; i++) ws->peq[(size_t)p[i] * (size_t)W + (size_t)(i >> 6)] |= (u64)1 << (i & 63);
}

static void peq_clear(scan_ws *ws, const u8 *p, int m, int W)
{
    for (int i = 0; i < m; i++) memset(&ws->peq[(size_t)p[i] * (size_t)W], 0, sizeof(u64) * (size_t)W);
}

/* Global distance of the loaded pattern (m >= 1, W = ceil(m / 64) words)
 * against text t (n >= 1). The horizontal delta entering word 0 of every
 * column is +1 (row 0 of a global alignment counts the columns); word w
 * passes its top row's delta to word w + 1. score follows D[m][j + 1]. With
 * k >= 0 the scan stops once score - (columns left) > k: every remaining
 * column lowers the bottom row by at most one. */
static int scan(scan_ws *ws, int m, int W, const u8 *t, int n, int k)
{
    u64 *Pv = ws->pv, *Mv = ws->mv;
    for (int w = 0; w < W; w++) {
        Pv[w] = ~(u64)0;
        Mv[w] = 0;
    }
    const int lastbit = (m - 1) & 63;
    int score = m;
    for (int j = 0; j < n; j++) {
        const u64 *row = &ws->peq[(size_t)t[j] * (size_t)W];
        u64 hp = 1, hn = 0;                       /* delta into the top row: +1 */
        for (int w = 0; w < W; w++) {
            u64 Eq = row[w];
            const u64 pv = Pv[w], mv = Mv[w];
            const u64 Xv = Eq | mv;
            Eq |= hn;
            const u64 Xh = (((Eq & pv) + pv) ^ pv) | Eq;
            u64 Ph = mv | ~(Xh | pv);
            u64 Mh = pv & Xh;
            const int top = (w == W - 1) ? lastbit : 63;
            const u64 po = (Ph >> top) & 1, mo = (Mh >> top) & 1;
            Ph = (Ph << 1) | hp;
            Mh = (Mh << 1) | hn;
            Pv[w] = Mh | ~(Xv | Ph);
            Mv[w] = Ph & Xv;
            hp = po;
            hn = mo;
        }
        score += (int)hp - (int)hn;
        if (k >= 0 && score - (n - 1 - j) > k) return k + 1;
    }
    return (k >= 0 && score > k) ? k + 1 : score;
}

/* Exact common prefix and suffix removal: d(xPy, xQy) = d(P, Q). */
static void trim(const u8 **a, int *la, const u8 **b, int *lb)
{
    int p = 0;
    const int mn = *la < *lb ? *la : *lb;
    while (p < mn && (*a)[p] == (*b)[p]) p++;
    *a += p;
    *b += p;
    *la -= p;
    *lb -= p;
    int s = 0;
    const int mn2 = *la < *lb ? *la : *lb;
    while (s < mn2 && (*a)[*la - 1 - s] == (*b)[*lb - 1 - s]) s++;
    *la -= s;
    *lb -= s;
}

static int pair_scalar(scan_ws *ws, const u8 *a, int la, const u8 *b, int lb, int do_trim, int k)
{
    if (do_trim) trim(&a, &la, &b, &lb);
    if (la > lb) {                                 /* the pattern is the shorter string */
        const u8 *tp = a; a = b; b = tp;
        const int tl = la; la = lb; lb = tl;
    }
    const int gap = lb - la;
    if (k >= 0 && gap > k) return k + 1;
    if (la == 0) return (k >= 0 && lb > k) ? k + 1 : lb;
    const int W = (la + 63) >> 6;
    if (ws_reserve(ws, W)) return -1;
    peq_set(ws, a, la, W);
    const int d = scan(ws, la, W, b, lb, k);
    peq_clear(ws, a, la, W);
    return d;
}

static int check_off(const int64_t *off, int64_t n)
{
    if (!off || n < 0) return -2;
    for (int64_t i = 0; i < n; i++)
        if (off[i + 1] < off[i] || off[i + 1] - off[i] > 0x7fffffff) return -2;
    return 0;
}

int cpu_pairs(const uint8_t *ba, const int64_t *oa, const uint8_t *bb, const int64_t *ob, int64_t n,
              int32_t *out, int do_trim, int k)
{
    if (check_off(oa, n) || check_off(ob, n) || (n > 0 && (!ba || !bb || !out))) return -2;
    scan_ws ws = {0};
    for (int64_t i = 0; i < n; i++) {
        const int d = pair_scalar(&ws, ba + oa[i], (int)(oa[i + 1] - oa[i]), bb + ob[i],
                                  (int)(ob[i + 1] - ob[i]), do_trim, k);
        if (d < 0) {
            ws_release(&ws);
            return -1;
        }
        out[i] = d;
    }
    ws_release(&ws);
    return 0;
}

int cpu_cross(const uint8_t *bq, const int64_t *oq, int64_t nq, const uint8_t *bc, const int64_t *oc,
              int64_t nc, int32_t *out, int k)
{
    if (check_off(oq, nq) || check_off(oc, nc) || (nq > 0 && nc > 0 && (!bq || !bc || !out))) return -2;
    scan_ws ws = {0};
    for (int64_t q = 0; q < nq; q++) {
        const u8 *p = bq + oq[q];
        const int m = (int)(oq[q + 1] - oq[q]);
        int32_t *row = out + q * nc;
        if (m == 0) {                              /* empty query: n, or k + 1 past k */
            for (int64_t c = 0; c < nc; c++) {
                const int n = (int)(oc[c + 1] - oc[c]);
                row[c] = (k >= 0 && n > k) ? k + 1 : n;
            }
            continue;
        }
        const int W = (m + 63) >> 6;
        if (ws_reserve(&ws, W)) {
            ws_release(&ws);
            return -1;
        }
        peq_set(&ws, p, m, W);
        for (int64_t c = 0; c < nc; c++) {
            const int n = (int)(oc[c + 1] - oc[c]);
            const int gap = n > m ? n - m : m - n;
            if (k >= 0 && gap > k) row[c] = k + 1;
            else if (n == 0) row[c] = (k >= 0 && m > k) ? k + 1 : m;
            else row[c] = scan(&ws, m, W, bc + oc[c], n, k);
        }
        peq_clear(&ws, p, m, W);
    }
    ws_release(&ws);
    return 0;
}

/* -------------------------------------------------------------------- AVX2 */
#if defined(__x86_64__) && (defined(__GNUC__) || defined(__clang__))
#include <immintrin.h>
#define AVX2 __attribute__((target("avx2")))

int cpu_simd_available(void)
{
    __builtin_cpu_init();
    return __builtin_cpu_supports("avx2") ? 1 : 0;
}

/* One lane type of the cdist kernel: LANES choices of at most BITS bytes per
 * 256-bit register, BITS-bit lanes. The lane operations come from the
 * element width: add, shift left by one, compare for equality. */
#define LANE_OPS(BITS, ST, ADD, SHL1, CMPEQ, CMPGT, SUB, SET1)                                  \
    static inline AVX2 __m256i add##BITS(__m256i a, __m256i b) { return ADD(a, b); }           \
    static inline AVX2 __m256i sub##BITS(__m256i a, __m256i b) { return SUB(a, b); }           \
    static inline AVX2 __m256i shl##BITS(__m256i a) { return SHL1(a, 1); }                     \
    static inline AVX2 __m256i eq##BITS(__m256i a, __m256i b) { return CMPEQ(a, b); }          \
    static inline AVX2 __m256i gt##BITS(__m256i a, __m256i b) { return CMPGT(a, b); }          \
    static inline AVX2 __m256i set##BITS(ST v) { return SET1(v); }
LANE_OPS(16, int16_t, _mm256_add_epi16, _mm256_slli_epi16, _mm256_cmpeq_epi16, _mm256_cmpgt_epi16,
         _mm256_sub_epi16, _mm256_set1_epi16)
LANE_OPS(32, int32_t, _mm256_add_epi32, _mm256_slli_epi32, _mm256_cmpeq_epi32, _mm256_cmpgt_epi32,
         _mm256_sub_epi32, _mm256_set1_epi32)
LANE_OPS(64, int64_t, _mm256_add_epi64, _mm256_slli_epi64, _mm256_cmpeq_epi64, _mm256_cmpgt_epi64,
         _mm256_sub_epi64, _mm256_set1_epi64x)

/* State of one group: Pv, Mv, the per-lane bottom-row bit, the score. */
typedef struct {
    __m256i pv, mv, last, score;
} lanes_t;

/* One column of one group, BITS-bit lanes: the Myers/Hyyro step with +1
 * entering the top row, and the bottom-row delta added to each lane's
 * score (compare masks are -1 where true, so score - maskP + maskM). */
#define STEP(BITS, S, EQ)                                                                       \
    do {                                                                                        \
        const __m256i st_eq_ = (EQ);                                                               \
        const __m256i st_xv_ = _mm256_or_si256(st_eq_, (S).mv);                                        \
        const __m256i st_xh_ = _mm256_or_si256(                                                     \
            _mm256_xor_si256(add##BITS(_mm256_and_si256(st_eq_, (S).pv), (S).pv), (S).pv), st_eq_);   \
        __m256i st_ph_ = _mm256_or_si256((S).mv, _mm256_xor_si256(_mm256_or_si256(st_xh_, (S).pv), ones)); \
        __m256i st_mh_ = _mm256_and_si256((S).pv, st_xh_);                                              \
        (S).score = sub##BITS((S).score, eq##BITS(_mm256_and_si256(st_ph_, (S).last), (S).last));    \
        (S).score = add##BITS((S).score, eq##BITS(_mm256_and_si256(st_mh_, (S).last), (S).last));    \
        st_ph_ = _mm256_or_si256(shl##BITS(st_ph_), one);                                               \
        st_mh_ = shl##BITS(st_mh_);                                                                     \
        (S).pv = _mm256_or_si256(st_mh_, _mm256_xor_si256(_mm256_or_si256(st_xv_, st_ph_), ones));          \
        (S).mv = _mm256_and_si256(st_ph_, st_xv_);                                                      \
    } while (0)

/* The cdist kernel for one lane type: `cnt` groups of LANES choices (idx,
 * all at most BITS bytes long and at least 1), each against every query,
 * two groups at a time. tables: scratch of 2 x 256 registers. */
#define DEFINE_CROSS(BITS, UT, ST, STMAX, LANES)                                                \
    static AVX2 void build##BITS(__m256i *table, const u8 *bc, const int64_t *oc,             \
                                 const int64_t *idx, int cnt, lanes_t *st)                      \
    {                                                                                           \
        memset(table, 0, 256 * sizeof(__m256i));                                               \
        UT last[LANES];                                                                         \
        ST len[LANES];                                                                          \
        for (int l = 0; l < LANES; l++) {                                                       \
            last[l] = 0;                                                                        \
            len[l] = (ST)(STMAX / 2);       /* unused lanes never hold the early exit back */   \
        }                                                                                       \
        for (int l = 0; l < cnt; l++) {                                                         \
            const u8 *p = bc + oc[idx[l]];                                                      \
            const int m = (int)(oc[idx[l] + 1] - oc[idx[l]]);                                   \
            for (int i = 0; i < m; i++) ((UT *)&table[p[i]])[l] |= (UT)((UT)1 << i);            \
            last[l] = (UT)((UT)1 << (m - 1));                                                   \
            len[l] = (ST)m;                                                                     \
        }                                                                                       \
        st->last = _mm256_loadu_si256((const __m256i *)last);                                  \
        st->score = _mm256_loadu_si256((const __m256i *)len);                                  \
    }                                                                                           \
    static AVX2 void cross##BITS(const u8 *bq, const int64_t *oq, int64_t nq, const u8 *bc,    \
                                 const int64_t *oc, int64_t nc, const int64_t *idx, int64_t n_idx, \
                                 int32_t *out, int k, __m256i *tables)                          \
    {                                                                                           \
        const __m256i ones = _mm256_set1_epi32(-1);                                            \
        const __m256i one = set##BITS(1);                                                       \
        for (int64_t g = 0; g < n_idx; g += 2 * LANES) {                                       \
            const int c0 = (int)(n_idx - g < LANES ? n_idx - g : LANES);                       \
            const int64_t rest = n_idx - g - c0;                                                \
            const int c1 = (int)(rest <= 0 ? 0 : rest < LANES ? rest : LANES);                 \
            lanes_t init0, init1;                                                               \
            build##BITS(tables, bc, oc, idx + g, c0, &init0);                                   \
            build##BITS(tables + 256, bc, oc, idx + g + c0, c1, &init1);                        \
            for (int64_t q = 0; q < nq; q++) {                                                  \
                const u8 *t = bq + oq[q];                                                       \
                const int n = (int)(oq[q + 1] - oq[q]);                                         \
                lanes_t s0 = init0, s1 = init1;                                                 \
                s0.pv = s1.pv = ones;                                                           \
                s0.mv = s1.mv = _mm256_setzero_si256();                                         \
                int stopped = 0;                                                                \
                for (int j = 0; j < n; j++) {                                                   \
                    const __m256i *row0 = tables + t[j];                                        \
                    STEP(BITS, s0, _mm256_load_si256(row0));                                    \
                    STEP(BITS, s1, _mm256_load_si256(row0 + 256));                              \
                    /* every 8 columns: stop once each lane's score minus the                   \
                     * columns left exceeds k (both groups); a bound past the lane's            \
                     * range can never be exceeded */                                           \
                    if (k >= 0 && (j & 7) == 7 && j + 1 < n) {                                  \
                        const long long bv = (long long)k + (long long)(n - 1 - j);             \
                        if (bv < (long long)(STMAX / 4)) {                                      \
                            const __m256i bound = set##BITS((ST)bv);                            \
                            const __m256i ok = _mm256_and_si256(gt##BITS(s0.score, bound),      \
                                                                gt##BITS(s1.score, bound));     \
                            if (_mm256_movemask_epi8(ok) == -1) {                               \
                                stopped = 1;                                                    \
                                break;                                                          \
                            }                                                                   \
                        }                                                                       \
                    }                                                                           \
                }                                                                               \
                ST sc[2][LANES];                                                                \
                _mm256_storeu_si256((__m256i *)sc[0], s0.score);                               \
                _mm256_storeu_si256((__m256i *)sc[1], s1.score);                               \
                int32_t *row = out + q * nc;                                                    \
                for (int h = 0; h < 2; h++) {                                                   \
                    const int cnt = h ? c1 : c0;                                                \
                    const int64_t *ix = idx + g + (h ? c0 : 0);                                 \
                    for (int l = 0; l < cnt; l++) {                                             \
                        const int d = stopped ? k + 1 : (int)sc[h][l];                          \
                        row[ix[l]] = (k >= 0 && d > k) ? k + 1 : d;                             \
                    }                                                                           \
                }                                                                               \
            }                                                                                   \
        }                                                                                       \
    }
DEFINE_CROSS(16, uint16_t, int16_t, 32767, 16)
DEFINE_CROSS(32, uint32_t, int32_t, 2147483647LL, 8)
DEFINE_CROSS(64, uint64_t, int64_t, 9223372036854775807LL, 4)

int cpu_cross_simd(const uint8_t *bq, const int64_t *oq, int64_t nq, const uint8_t *bc,
                   const int64_t *oc, int64_t nc, int32_t *out, int k)
{
    if (check_off(oq, nq) || check_off(oc, nc) || (nq > 0 && nc > 0 && (!bq || !bc || !out))) return -2;
    if (!cpu_simd_available()) return cpu_cross(bq, oq, nq, bc, oc, nc, out, k);
    if (nq == 0 || nc == 0) return 0;
    /* choices by lane class; queries the 16-bit lanes cannot score (scores
     * above 32767) and the longer choices go to the scalar path */
    int64_t *idx = malloc(sizeof(int64_t) * (size_t)nc);
    __m256i *tables = aligned_alloc(32, 512 * sizeof(__m256i));
    if (!idx || !tables) {
        free(idx);
        free(tables);
        return -1;
    }
    int64_t cnt[4] = {0, 0, 0, 0};
    for (int64_t c = 0; c < nc; c++) {
        const int64_t n = oc[c + 1] - oc[c];
        cnt[n == 0 || n > 64 ? 3 : n <= 16 ? 0 : n <= 32 ? 1 : 2]++;
    }
    int64_t base[4] = {0, cnt[0], cnt[0] + cnt[1], cnt[0] + cnt[1] + cnt[2]};
    for (int64_t c = 0; c < nc; c++) {
        const int64_t n = oc[c + 1] - oc[c];
        idx[base[n == 0 || n > 64 ? 3 : n <= 16 ? 0 : n <= 32 ? 1 : 2]++] = c;
    }
    const int64_t s1 = cnt[0], s2 = cnt[0] + cnt[1], s3 = cnt[0] + cnt[1] + cnt[2];
    int rc = 0;
    /* 16-bit lanes score queries of at most 32,000 bytes; longer ones take
     * the scalar path for the short choices */
    for (int64_t q0 = 0; q0 < nq;) {
        int64_t q1 = q0;
        while (q1 < nq && oq[q1 + 1] - oq[q1] <= 32000) q1++;
        if (q1 > q0) cross16(bq, oq + q0, q1 - q0, bc, oc, nc, idx, cnt[0], out + q0 * nc, k, tables);
        q0 = q1 + (q1 < nq ? 1 : 0);
    }
    cross32(bq, oq, nq, bc, oc, nc, idx + s1, cnt[1], out, k, tables);
    cross64(bq, oq, nq, bc, oc, nc, idx + s2, cnt[2], out, k, tables);
    /* scalar: empty and long choices against every query, and the long
     * queries against the 16-byte choices */
    scan_ws ws = {0};
    for (int64_t q = 0; q < nq && rc == 0; q++) {
        const u8 *p = bq + oq[q];
        const int m = (int)(oq[q + 1] - oq[q]);
        int32_t *row = out + q * nc;
        for (int64_t x = m > 32000 ? 0 : s3; x < nc; x++) {
            if (x == s1 && m > 32000) x = s3;               /* 32- and 64-bit lanes did those */
            if (x >= nc) break;
            const int64_t c = idx[x];
            const int d = pair_scalar(&ws, p, m, bc + oc[c], (int)(oc[c + 1] - oc[c]), 0, k);
            if (d < 0) {
                rc = -1;
                break;
            }
            row[c] = d;
        }
    }
    ws_release(&ws);
    free(idx);
    free(tables);
    return rc;
}

/* pairwise: 8 pairs per register (32-bit lanes, patterns of 1..32 bytes
 * after trimming) or 4 (64-bit lanes, 33..64 bytes). Lane l's Eq word comes
 * from its own 256-entry table (a gather); the group's texts are first laid
 * out column by column as table indices (l * 256 + byte), so a column is one
 * load and one gather; lanes whose text has ended keep their state and
 * score. The tables are cleared entry by entry after use. */
typedef struct {
    const u8 *pat[8], *txt[8];
    int m[8], n[8];
    int64_t who[8];
    int cnt;
} pair_group;

#define DEFINE_PAIRS(BITS, UT, LANES, GATHER, IDXT)                                            \
    static AVX2 void pairs##BITS(pair_group *G, int32_t *out, UT *tab, IDXT *cols)              \
    {                                                                                           \
        UT last[LANES];                                                                         \
        IDXT len[LANES], nl[LANES];                                                             \
        int nmax = 0;                                                                           \
        for (int l = 0; l < LANES; l++) {                                                       \
            last[l] = 0;                                                                        \
            len[l] = 0;                                                                         \
            nl[l] = 0;                                                                          \
        }                                                                                       \
        for (int l = 0; l < G->cnt; l++) {                                                      \
            for (int i = 0; i < G->m[l]; i++) tab[l * 256 + G->pat[l][i]] |= (UT)((UT)1 << i);  \
            last[l] = (UT)((UT)1 << (G->m[l] - 1));                                             \
            len[l] = (IDXT)G->m[l];                                                             \
            nl[l] = (IDXT)G->n[l];                                                              \
            if (G->n[l] > nmax) nmax = G->n[l];                                                 \
        }                                                                                       \
        for (int j = 0; j < nmax; j++)                                                          \
            for (int l = 0; l < LANES; l++)                                                     \
                cols[(size_t)j * LANES + l] =                                                   \
                    (IDXT)(l * 256 + ((l < G->cnt && j < G->n[l]) ? G->txt[l][j] : 0));         \
        const __m256i ones = _mm256_set1_epi32(-1);                                            \
        const __m256i one = set##BITS(1);                                                       \
        lanes_t s;                                                                              \
        s.pv = ones;                                                                            \
        s.mv = _mm256_setzero_si256();                                                          \
        s.last = _mm256_loadu_si256((const __m256i *)last);                                    \
        s.score = _mm256_loadu_si256((const __m256i *)len);                                    \
        const __m256i nlen = _mm256_loadu_si256((const __m256i *)nl);                          \
        for (int j = 0; j < nmax; j++) {                                                        \
            const __m256i vidx = _mm256_loadu_si256((const __m256i *)&cols[(size_t)j * LANES]); \
            const __m256i eqv = GATHER(tab, vidx);                                              \
            const __m256i live = gt##BITS(nlen, set##BITS(j));                                  \
            lanes_t t = s;                                                                      \
            STEP(BITS, t, eqv);                                                                 \
            s.pv = _mm256_blendv_epi8(s.pv, t.pv, live);                                        \
            s.mv = _mm256_blendv_epi8(s.mv, t.mv, live);                                        \
            s.score = _mm256_blendv_epi8(s.score, t.score, live);                               \
        }                                                                                       \
        IDXT sc[LANES];                                                                         \
        _mm256_storeu_si256((__m256i *)sc, s.score);                                           \
        for (int l = 0; l < G->cnt; l++) {                                                      \
            out[G->who[l]] = (int32_t)sc[l];                                                    \
            for (int i = 0; i < G->m[l]; i++) tab[l * 256 + G->pat[l][i]] = 0;                  \
        }                                                                                       \
        G->cnt = 0;                                                                             \
    }
#define GATHER32(tab, idx) _mm256_i32gather_epi32((const int *)(tab), (idx), 4)
#define GATHER64(tab, idx) _mm256_i64gather_epi64((const long long *)(tab), (idx), 8)
DEFINE_PAIRS(32, uint32_t, 8, GATHER32, int32_t)
DEFINE_PAIRS(64, uint64_t, 4, GATHER64, int64_t)

int cpu_pairs_simd(const uint8_t *ba, const int64_t *oa, const uint8_t *bb, const int64_t *ob,
                   int64_t n, int32_t *out, int do_trim)
{
    if (check_off(oa, n) || check_off(ob, n) || (n > 0 && (!ba || !bb || !out))) return -2;
    if (!cpu_simd_available()) return cpu_pairs(ba, oa, bb, ob, n, out, do_trim, -1);
    uint32_t *tab32 = calloc(8 * 256, sizeof(uint32_t));
    uint64_t *tab64 = calloc(4 * 256, sizeof(uint64_t));
    int64_t cap = 0;                 /* columns the index buffers hold */
    int32_t *cols32 = NULL;
    int64_t *cols64 = NULL;
    pair_group g32 = {0}, g64 = {0};
    scan_ws ws = {0};
    int rc = (tab32 && tab64) ? 0 : -1;
    int64_t nmax32 = 0, nmax64 = 0;
    for (int64_t i = 0; i < n && rc == 0; i++) {
        const u8 *a = ba + oa[i], *b = bb + ob[i];
        int la = (int)(oa[i + 1] - oa[i]), lb = (int)(ob[i + 1] - ob[i]);
        if (do_trim) trim(&a, &la, &b, &lb);
        if (la > lb) {
            const u8 *tp = a; a = b; b = tp;
            const int t = la; la = lb; lb = t;
        }
        if (la == 0) {
            out[i] = lb;
            continue;
        }
        if (la > 64) {
            const int d = pair_scalar(&ws, a, la, b, lb, 0, -1);
            if (d < 0) rc = -1;
            else out[i] = d;
            continue;
        }
        pair_group *G = la <= 32 ? &g32 : &g64;
        int64_t *nmax = la <= 32 ? &nmax32 : &nmax64;
        G->pat[G->cnt] = a; G->m[G->cnt] = la; G->txt[G->cnt] = b; G->n[G->cnt] = lb;
        G->who[G->cnt] = i;
        if (lb > *nmax) *nmax = lb;
        G->cnt++;
        if (G->cnt == (la <= 32 ? 8 : 4)) {
            if (*nmax > cap) {
                cap = *nmax * 2;
                free(cols32);
                free(cols64);
                cols32 = malloc(sizeof(int32_t) * 8 * (size_t)cap);
                cols64 = malloc(sizeof(int64_t) * 4 * (size_t)cap);
                if (!cols32 || !cols64) { rc = -1; break; }
            }
            if (la <= 32) pairs32(G, out, tab32, cols32);
            else pairs64(G, out, tab64, cols64);
            *nmax = 0;
        }
    }
    /* the partial groups */
    for (int h = 0; h < 2 && rc == 0; h++) {
        pair_group *G = h ? &g64 : &g32;
        const int64_t nm = h ? nmax64 : nmax32;
        if (!G->cnt) continue;
        if (nm > cap) {
            cap = nm;
            free(cols32);
            free(cols64);
            cols32 = malloc(sizeof(int32_t) * 8 * (size_t)cap);
            cols64 = malloc(sizeof(int64_t) * 4 * (size_t)cap);
            if (!cols32 || !cols64) { rc = -1; break; }
        }
        if (h) pairs64(G, out, tab64, cols64);
        else pairs32(G, out, tab32, cols32);
    }
    ws_release(&ws);
    free(tab32);
    free(tab64);
    free(cols32);
    free(cols64);
    return rc;
}

#else  /* no AVX2 on this target: the SIMD entry points run the scalar path */
int cpu_simd_available(void) { return 0; }
int cpu_cross_simd(const uint8_t *bq, const int64_t *oq, int64_t nq, const uint8_t *bc,
                   const int64_t *oc, int64_t nc, int32_t *out, int k)
{
    return cpu_cross(bq, oq, nq, bc, oc, nc, out, k);
}
int cpu_pairs_simd(const uint8_t *ba, const int64_t *oa, const uint8_t *bb, const int64_t *ob,
                   int64_t n, int32_t *out, int do_trim)
{
    return cpu_pairs(ba, oa, bb, ob, n, out, do_trim, -1);
}
#endif
