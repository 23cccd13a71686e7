/*
 * oracle/ref_driver.c -- a thin driver around the reference C core, built by
 * oracle/Makefile from the reference sources where they lie (it #includes
 * csrc/myers.c of karims/editdist-bench; nothing of it is copied here).
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT: it writes golden vectors (tests/golden/)
 * and is the "reference" CPU baseline of bench.py. The reference exports no
 * threshold cross entry point (its k-path for cross lives in the Python
 * extension, csrc/edpy.c:91-127), so these loops restate those extension loops
 * around the reference's own static pair_distance_k / run_text / run_text_k.
 */
#include <pthread.h>
#include "myers.c"

/* _edpy.pairs loop (csrc/edpy.c:71-83) on packed buffers. */
int ref_pairs_k(const uint8_t *ba, const int64_t *oa, const uint8_t *bb, const int64_t *ob,
                int64_t n, int32_t *out, int trim, int k)
{
    int W = max_words(oa, ob, n);
    ws_t ws;
    if (ws_init(&ws, W)) return -1;
    word *Pv = malloc(sizeof(word) * ws.W), *Mv = malloc(sizeof(word) * ws.W);
    if (!Pv || !Mv) { free(Pv); free(Mv); ws_free(&ws); return -1; }
    for (int64_t i = 0; i < n; i++)
        out[i] = pair_distance_k(&ws, Pv, Mv, ba + oa[i], (int)(oa[i + 1] - oa[i]),
                                 bb + ob[i], (int)(ob[i + 1] - ob[i]), trim, k);
    free(Pv); free(Mv); ws_free(&ws);
    return 0;
}

/* _edpy.cross loop (csrc/edpy.c:102-121) on packed buffers, for query rows
 * [q0, q1) and candidate columns [c0, c1) of the nq x nc output. */
static int cross_block(const uint8_t *bq, const int64_t *oq, int64_t q0, int64_t q1,
                       const uint8_t *bc, const int64_t *oc, int64_t c0, int64_t c1, int64_t nc,
                       int32_t *out, int k, int W)
{
    ws_t ws;
    if (ws_init(&ws, W)) return -1;
    word *Pv = malloc(sizeof(word) * ws.W), *Mv = malloc(sizeof(word) * ws.W);
    if (!Pv || !Mv) { free(Pv); free(Mv); ws_free(&ws); return -1; }
    for (int64_t i = q0; i < q1; i++) {
        const uint8_t *q = bq + oq[i];
        int lq = (int)(oq[i + 1] - oq[i]);
        int32_t *row = out + i * nc;
        if (lq == 0) {
            for (int64_t j = c0; j < c1; j++) {
                int64_t n = oc[j + 1] - oc[j];
                row[j] = (k >= 0 && n > k) ? k + 1 : (int32_t)n;
            }
            continue;
        }
        load_pattern(&ws, q, lq);
        for (int64_t j = c0; j < c1; j++) {
            const uint8_t *t = bc + oc[j];
            int n = (int)(oc[j + 1] - oc[j]);
            row[j] = k < 0 ? run_text(&ws, lq, t, n, Pv, Mv) : run_text_k(&ws, lq, t, n, Pv, Mv, k);
        }
        clear_pattern(&ws, q, lq);
    }
    free(Pv); free(Mv); ws_free(&ws);
    return 0;
}

static int query_words(const int64_t *oq, int64_t nq)
{
    int64_t mx = 0;
    for (int64_t i = 0; i < nq; i++)
        if (oq[i + 1] - oq[i] > mx) mx = oq[i + 1] - oq[i];
    return (int)((mx + 63) >> 6);
}

int ref_cross_k(const uint8_t *bq, const int64_t *oq, int64_t nq,
                const uint8_t *bc, const int64_t *oc, int64_t nc, int32_t *out, int k)
{
    if (nq == 0 || nc == 0) return 0;
    return cross_block(bq, oq, 0, nq, bc, oc, 0, nc, nc, out, k, query_words(oq, nq));
}

/* ---- all-cores variant: (query, column tile) units handed out dynamically */
typedef struct {
    const uint8_t *bq; const int64_t *oq; int64_t nq;
    const uint8_t *bc; const int64_t *oc; int64_t nc;
    int32_t *out; int k; int W; int64_t tile, tiles_per_row, next;
    pthread_mutex_t mu; int err;
} mt_job;

static void *mt_worker(void *arg)
{
    mt_job *J = (mt_job *)arg;
    const int64_t total = J->nq * J->tiles_per_row;
    for (;;) {
        pthread_mutex_lock(&J->mu);
        int64_t t = J->next++;
        pthread_mutex_unlock(&J->mu);
        if (t >= total) break;
        int64_t qi = t / J->tiles_per_row, c0 = (t % J->tiles_per_row) * J->tile;
        int64_t c1 = c0 + J->tile < J->nc ? c0 + J->tile : J->nc;
        if (cross_block(J->bq, J->oq, qi, qi + 1, J->bc, J->oc, c0, c1, J->nc, J->out, J->k, J->W))
            J->err = -1;
    }
    return NULL;
}

int ref_cross_k_mt(const uint8_t *bq, const int64_t *oq, int64_t nq,
                   const uint8_t *bc, const int64_t *oc, int64_t nc, int32_t *out, int k,
                   int threads)
{
    if (nq == 0 || nc == 0) return 0;
    if (threads < 1) return -2;
    mt_job J = {bq, oq, nq, bc, oc, nc, out, k, query_words(oq, nq), 8192, 0, 0,
                PTHREAD_MUTEX_INITIALIZER, 0};
    J.tiles_per_row = (nc + J.tile - 1) / J.tile;
    pthread_t *th = malloc(sizeof(pthread_t) * (size_t)threads);
    if (!th) return -1;
    int started = 0;
    for (; started < threads; started++)
        if (pthread_create(&th[started], NULL, mt_worker, &J)) break;
    if (started == 0) mt_worker(&J);
    for (int t = 0; t < started; t++) pthread_join(th[t], NULL);
    free(th);
    return J.err;
}

typedef struct {
    const uint8_t *ba; const int64_t *oa; const uint8_t *bb; const int64_t *ob;
    int64_t n; int32_t *out; int trim, k; int64_t chunk, next;
    pthread_mutex_t mu; int err;
} mtp_job;

static void *mtp_worker(void *arg)
{
    mtp_job *J = (mtp_job *)arg;
    for (;;) {
        pthread_mutex_lock(&J->mu);
        int64_t i0 = J->next;
        J->next += J->chunk;
        pthread_mutex_unlock(&J->mu);
        if (i0 >= J->n) break;
        int64_t i1 = i0 + J->chunk < J->n ? i0 + J->chunk : J->n;
        if (ref_pairs_k(J->ba, J->oa + i0, J->bb, J->ob + i0, i1 - i0, J->out + i0, J->trim, J->k))
            J->err = -1;
    }
    return NULL;
}

int ref_pairs_k_mt(const uint8_t *ba, const int64_t *oa, const uint8_t *bb, const int64_t *ob,
                   int64_t n, int32_t *out, int trim, int k, int threads)
{
    if (threads < 1) return -2;
    mtp_job J = {ba, oa, bb, ob, n, out, trim, k, 4096, 0, PTHREAD_MUTEX_INITIALIZER, 0};
    pthread_t *th = malloc(sizeof(pthread_t) * (size_t)threads);
    if (!th) return -1;
    int started = 0;
    for (; started < threads; started++)
        if (pthread_create(&th[started], NULL, mtp_worker, &J)) break;
    if (started == 0) mtp_worker(&J);
    for (int t = 0; t < started; t++) pthread_join(th[t], NULL);
    free(th);
    return J.err;
}
