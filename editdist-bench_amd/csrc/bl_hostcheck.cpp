/*
 * bl_hostcheck.cpp -- the kernels' per-lane code (bl_device.h) built for the
 * host, so tests/test_hostcheck.py checks it against the oracle on a CPU.
 * Counterpart of the reference's csrc/gpu_hostcheck.cpp:9-37.
 *
 * The long-string path is simulated lane by lane: the same long_lane_step()
 * runs for all 64 lanes of a wave in lockstep and the hand-off between lanes
 * is an array shift, which is what __shfl_up does on the device.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "bl_device.h"

using namespace bl;

static int long_sim(const u8 *p, int m, const u8 *t, int n, int is_pair, int trim, int k)
{
    const int r = long_prepare(p, m, t, n, is_pair, trim, k);
    if (r >= 0) return r;
    LongCtx C;
    C.p = p; C.m = m; C.t = t; C.n = n;
    C.W = (m + 63) >> 6;
    std::vector<int8_t> hbuf((size_t)n + 1);
    C.hbuf = hbuf.data();
    const int S = (C.W + 63) >> 6;
    static LongLane L[64];
    for (int s = 0; s < S; s++) {
        C.s = s;
        C.nl = C.W - 64 * s < 64 ? C.W - 64 * s : 64;
        for (int lane = 0; lane < 64; lane++) long_lane_init(C, L[lane], lane);
        u32 prev[64], cur[64];
        memset(prev, 0, sizeof prev);
        for (int step = 0; step < n + C.nl - 1; step++) {
            for (int lane = 0; lane < 64; lane++) {
                const u32 in = lane == 0 ? long_source(C, step) : prev[lane - 1];
                cur[lane] = long_lane_step(C, L[lane], lane, step, in);
            }
            memcpy(prev, cur, sizeof prev);
        }
    }
    const int d = m + L[(C.W - 1) - 64 * (S - 1)].delta;
    return (k >= 0 && d > k) ? k + 1 : d;
}

static int pair_host(const u8 *a, int la, const u8 *b, int lb, int trim, int k)
{
    const int m = la < lb ? la : lb;
    if (m <= 64) return pair_lane<1>(a, la, b, lb, trim, k);
    if (m <= 128) return pair_lane<2>(a, la, b, lb, trim, k);
    if (m <= 256) return pair_lane<4>(a, la, b, lb, trim, k);
    return long_sim(a, la, b, lb, 1, trim, k);
}

template <int NW, typename Text>
static int peq_host_nw(const u8 *q, int m, const Text &t, int n, int k)
{
    std::vector<u64> peq((size_t)256 * NW, 0);
    std::vector<u32> lo32(256, 0);
    for (int i = 0; i < m; i++) peq[(size_t)q[i] * NW + (i >> 6)] |= (u64)1 << (i & 63);
    for (int c = 0; c < 256; c++) lo32[(size_t)c] = (u32)peq[(size_t)c * NW];
    return peq_lane<NW, 0>(lo32.data(), peq.data(), m, t, n, k);
}

/* The search lane against a text read from choice stripes (the cdist kernels'
 * layout) or, for queries above 512 bytes, the wavefront on the packed bytes. */
static int peq_host(const u8 *q, int m, const StripeText &st, const u8 *t, int n, int k)
{
    if (m == 0) return empty_pattern(n, k);
    const int w = (m + 63) >> 6;
    if (w <= 1) return peq_host_nw<1>(q, m, st, n, k);
    if (w <= 2) return peq_host_nw<2>(q, m, st, n, k);
    if (w <= 4) return peq_host_nw<4>(q, m, st, n, k);
    if (w <= 8) return peq_host_nw<8>(q, m, st, n, k);
    return long_sim(q, m, t, n, 0, 0, k);
}

/* Stripes of a packed set, filled by the stripes kernel's own word rule. */
struct HostStripes {
    std::vector<int64_t> gbase;
    std::vector<u32> words;
    HostStripes(const u8 *buf, const int64_t *off, int64_t n) : gbase((size_t)((n + 63) / 64 + 1))
    {
        const int64_t total = stripes_layout(off, n, gbase.data());
        words.assign((size_t)(total > 0 ? total : 0), 0xDEADBEEFu);
        for (int64_t c = 0; c < n; c++) {
            const int64_t g = c / 64, gb = gbase[(size_t)g];
            const int64_t width = (gbase[(size_t)g + 1] - gb) / 64;
            for (int64_t j = 0; j < width; j++)
                words[(size_t)(gb + 64 * j + c % 64)] =
                    stripe_word(buf + off[c], (int)(off[c + 1] - off[c]), j);
        }
    }
    StripeText text(int64_t c) const { return StripeText{words.data() + gbase[(size_t)(c / 64)] + c % 64}; }
};

static int w32_pair(int w, const u8 *ba, const int64_t *oa, const u8 *bb, const int64_t *ob,
                    int64_t i, int trim, int k)
{
    switch (w) {
    case 1: return packed_pair_w32<1>(ba, oa, bb, ob, i, trim, k);
    case 2: return packed_pair_w32<2>(ba, oa, bb, ob, i, trim, k);
    case 3: return packed_pair_w32<3>(ba, oa, bb, ob, i, trim, k);
    case 4: return packed_pair_w32<4>(ba, oa, bb, ob, i, trim, k);
    case 5: return packed_pair_w32<5>(ba, oa, bb, ob, i, trim, k);
    case 6: return packed_pair_w32<6>(ba, oa, bb, ob, i, trim, k);
    case 7: return packed_pair_w32<7>(ba, oa, bb, ob, i, trim, k);
    default: return packed_pair_w32<8>(ba, oa, bb, ob, i, trim, k);
    }
}

/* The batched pairs pipeline emulated step by step with the kernels' shared
 * code and the same index arithmetic: keys from the offsets -> per-chunk
 * bucket scan (descending keys) -> chunk / tile prefix sums -> scatter, each
 * 1024-pair block ranking its pairs per key and reserving the key's range with
 * one add -> tiles taken in order, the chunk of each tile found by the same
 * bisection. Every index is checked, and so is the tile invariant (no pair in
 * a tile has a shorter string of more 32-row units than its first pair). chunk:
 * pairs per chunk (a multiple of 1024, the bucketing block, as on the GPU).
 * Returns 0, or a negative code naming the violated check. */
template <typename BandTile>
static int blh_pairs_batched_core(const u8 *ba, const int64_t *oa, const u8 *bb,
                                  const int64_t *ob, int64_t n, int32_t *out, int trim, int k,
                                  int bin_width, int64_t chunk, int band, BandTile band_tile)
{
    if (chunk < 1024 || chunk % 1024) return -9;
    const int64_t nchunks = n ? (n + chunk - 1) / chunk : 1;
    std::vector<int> key((size_t)n);
    std::vector<uint64_t> hist((size_t)(nchunks * PKEYS), 0), cursor(hist.size(), 0);
    std::vector<uint64_t> ccnt((size_t)nchunks), tcnt((size_t)nchunks);
    std::vector<uint64_t> cbase((size_t)nchunks + 1), tbase((size_t)nchunks + 1);
    auto minlen = [&](int64_t i) {
        const int64_t la = oa[i + 1] - oa[i], lb = ob[i + 1] - ob[i];
        return (int)(la < lb ? la : lb);
    };
    for (int64_t i = 0; i < n; i++) {
        int result = 0;
        const int kk = pair_key((int)(oa[i + 1] - oa[i]), (int)(ob[i + 1] - ob[i]), k, bin_width,
                                band, 8, false, result);
        if (kk < 0 || kk >= PKEYS) return -10;
        if (kk == 0) out[i] = result;
        key[(size_t)i] = kk;
        if (kk) hist[(size_t)((i / chunk) * PKEYS + kk)]++;
    }
    for (int64_t c = 0; c < nchunks; c++) {
        uint64_t acc = 0;
        for (int x = PKEYS - 1; x >= 1; x--) {
            cursor[(size_t)(c * PKEYS + x)] = acc;
            acc += hist[(size_t)(c * PKEYS + x)];
        }
        ccnt[(size_t)c] = acc;
        const uint64_t nl = hist[(size_t)(c * PKEYS + PKEY_LONG)];
        tcnt[(size_t)c] = nl + (acc - nl + 63) / 64;
    }
    cbase[0] = tbase[0] = 0;
    for (int64_t c = 0; c < nchunks; c++) {
        cbase[(size_t)c + 1] = cbase[(size_t)c] + ccnt[(size_t)c];
        tbase[(size_t)c + 1] = tbase[(size_t)c] + tcnt[(size_t)c];
    }
    const int64_t total = (int64_t)cbase[(size_t)nchunks];
    std::vector<int32_t> sorted((size_t)total, -1);
    for (int64_t b0 = 0; b0 < n; b0 += 1024) {            /* one scatter block */
        const int64_t b1 = b0 + 1024 < n ? b0 + 1024 : n;
        const int64_t c = b0 / chunk;
        std::vector<uint32_t> cnt(PKEYS, 0), rank((size_t)(b1 - b0), 0);
        for (int64_t i = b0; i < b1; i++)
            if (key[(size_t)i]) rank[(size_t)(i - b0)] = cnt[(size_t)key[(size_t)i]]++;
        std::vector<uint64_t> base(PKEYS, 0);
        for (int x = 1; x < PKEYS; x++) {
            base[(size_t)x] = cursor[(size_t)(c * PKEYS + x)];
            cursor[(size_t)(c * PKEYS + x)] += cnt[(size_t)x];
        }
        for (int64_t i = b0; i < b1; i++) {
            const int kk = key[(size_t)i];
            if (!kk) continue;
            const uint64_t pos = cbase[(size_t)c] + base[(size_t)kk] + rank[(size_t)(i - b0)];
            if (pos >= cbase[(size_t)c + 1]) return -11;
            if (sorted[(size_t)pos] != -1) return -18;
            sorted[(size_t)pos] = (int32_t)i;
        }
    }
    for (int64_t q = 0; q < total; q++)
        if (sorted[(size_t)q] < 0 || sorted[(size_t)q] >= n) return -12;
    const int64_t ntiles = (int64_t)tbase[(size_t)nchunks];
    int64_t c = 0;
    for (int64_t t = 0; t < ntiles; t++) {
        if ((int64_t)tbase[(size_t)c + 1] <= t || (int64_t)tbase[(size_t)c] > t) {
            int64_t lo = 0, hi = nchunks;
            while (hi - lo > 1) {
                const int64_t mid = (lo + hi) >> 1;
                if ((int64_t)tbase[(size_t)mid] <= t) lo = mid;
                else hi = mid;
            }
            c = lo;
        }
        if (!((int64_t)tbase[(size_t)c] <= t && t < (int64_t)tbase[(size_t)c + 1])) return -17;
        const int64_t lt = t - (int64_t)tbase[(size_t)c];
        const int64_t cb = (int64_t)cbase[(size_t)c];
        const int64_t nlong = (int64_t)hist[(size_t)(c * PKEYS + PKEY_LONG)];
        if (lt < nlong) {
            const int64_t i = sorted[(size_t)(cb + lt)];
            if (key[(size_t)i] != PKEY_LONG) return -13;
            out[i] = long_sim(ba + oa[i], (int)(oa[i + 1] - oa[i]), bb + ob[i],
                              (int)(ob[i + 1] - ob[i]), 1, trim, k);
            continue;
        }
        const int64_t s0 = cb + nlong + (lt - nlong) * 64;
        const int64_t ce = (int64_t)cbase[(size_t)c + 1];
        const int64_t s1 = s0 + 64 < ce ? s0 + 64 : ce;
        if (s0 >= ce) return -14;
        const int key0 = key[(size_t)sorted[(size_t)s0]];
        if (key_band(key0)) {
            if (band <= 0) return -19;
            band_tile(std::vector<int32_t>(sorted.begin() + s0, sorted.begin() + s1));
            continue;
        }
        const int w = key_units(key0);
        if (w < 1 || w > 8) return -15;
        for (int64_t q = s0; q < s1; q++) {
            if (key_band(key[(size_t)sorted[(size_t)q]])) return -20;
            if (((minlen(sorted[(size_t)q]) + 31) >> 5) > w) return -16;
        }
        for (int64_t q = s0; q < s1; q++) {
            const int64_t i = sorted[(size_t)q];
            out[i] = w32_pair(w, ba, oa, bb, ob, i, trim, k);
        }
    }
    return 0;
}

/* The exact-search lane (units_lane) over a word-major Peq table T (word w of
 * byte c at T[256 w + c]), as the exact kernel's LDS copy holds it. */
template <int U>
static int units_host(const u64 *T, int m, const StripeText &t, int n, int k)
{
    auto eq64 = [&](u32 ch, int w) { return T[w * 256 + ch]; };
    auto eq32 = [&](u32 ch, int u) { return ((const u32 *)T)[2 * ((u >> 1) * 256 + ch) + (u & 1)]; };
    return units_lane<U>(m, t, n, k, eq64, eq32);
}

extern "C" {

/* One pair through the lane code the pairs kernels run (class by the shorter
 * length: <= 64, <= 128, <= 256 in registers; longer via the wavefront). */
int blh_pair(const u8 *a, int la, const u8 *b, int lb, int trim, int k)
{
    return pair_host(a, la, b, lb, trim, k);
}

/* One pair through the 32-row-unit lane code of the batched tiles; w32 must
 * cover the shorter untrimmed string (1..8). */
int blh_pair_w32(const u8 *a, int la, const u8 *b, int lb, int trim, int k, int w32)
{
    switch (w32) {
    case 1: return pair_lane_w32<1>(a, la, b, lb, trim, k);
    case 2: return pair_lane_w32<2>(a, la, b, lb, trim, k);
    case 3: return pair_lane_w32<3>(a, la, b, lb, trim, k);
    case 4: return pair_lane_w32<4>(a, la, b, lb, trim, k);
    case 5: return pair_lane_w32<5>(a, la, b, lb, trim, k);
    case 6: return pair_lane_w32<6>(a, la, b, lb, trim, k);
    case 7: return pair_lane_w32<7>(a, la, b, lb, trim, k);
    case 8: return pair_lane_w32<8>(a, la, b, lb, trim, k);
    default: return -1;
    }
}

/* One pair through the band lane (win 32 or 64 rows at diagonal offset d0):
 * the pair's result, or -1 when the window cannot decide it. *reach receives
 * the threshold the window decides for this pair after trimming (-2 when the
 * gap / empty rules decided it without a scan). */
int blh_band(const u8 *a, int la, const u8 *b, int lb, int trim, int k, int d0, int win,
             int *reach)
{
    const int r = band_prepare(a, la, b, lb, trim, k);
    if (r >= 0) {
        *reach = -2;
        return r;
    }
    if (win == 32) {
        *reach = band_reach<32>(la - lb, d0);
        return band_pair<u32>(a, la, b, lb, k, d0);
    }
    *reach = band_reach<64>(la - lb, d0);
    return band_pair<u64>(a, la, b, lb, k, d0);
}

/* One batch of the count kernel's escalation list as cdist_list_kernel runs
 * it: the queries' bytes as patterns (no trim, no swap), the choices read
 * from stripes, one window offset for the batch from the lanes' m - n, 32
 * rows when that window reaches every lane's k, else 64 per lane where it
 * reaches, the whole pattern otherwise. n pairs (q[i], c[i]) in packed sets;
 * out[i] = min(d, k + 1); *banded = how many a band decided. */
int blh_list_batch(const u8 *bq, const int64_t *oq, const int32_t *qi, const u8 *bc, const int64_t *oc,
                   int64_t nc, const int32_t *ci, int n, int k, int32_t *out, int *banded)
{
    if (n < 1 || n > 64 || k < 0) return -1;
    const HostStripes hs(bc, oc, nc);
    int dmin = 0x7fffffff, dmax = -0x7fffffff;
    std::vector<int> m(n), len(n), dl(n);
    std::vector<bool> ok(n);
    for (int i = 0; i < n; i++) {
        m[i] = (int)(oq[qi[i] + 1] - oq[qi[i]]);
        len[i] = (int)(oc[ci[i] + 1] - oc[ci[i]]);
        dl[i] = m[i] - len[i];
        ok[i] = len[i] > 0 && m[i] > 0;
        if (ok[i]) {
            dmin = dl[i] < dmin ? dl[i] : dmin;
            dmax = dl[i] > dmax ? dl[i] : dmax;
        }
    }
    const int d32 = band_offset<32>(dmin, dmax), d64 = band_offset<64>(dmin, dmax);
    bool w32 = true;
    for (int i = 0; i < n; i++)
        if (ok[i] && k > band_reach<32>(dl[i], d32)) w32 = false;
    *banded = 0;
    for (int i = 0; i < n; i++) {
        const u8 *qb = bq + oq[qi[i]];
        const StripeText text = hs.text(ci[i]);
        int d = -1;
        if (ok[i]) {
            if (w32) d = band_scan<u32>(qb, m[i], text, len[i], d32, k, true);
            else if (k <= band_reach<64>(dl[i], d64)) d = band_scan<u64>(qb, m[i], text, len[i], d64, k, true);
        }
        if (d >= 0) (*banded)++;
        else d = peq_host(qb, m[i], text, bc + oc[ci[i]], len[i], k);
        out[i] = d;
    }
    return 0;
}

/* The window offset a tile uses for pairs with m - n in [dl_min, dl_max]. */
int blh_band_offset(int dl_min, int dl_max, int win)
{
    return win == 32 ? band_offset<32>(dl_min, dl_max) : band_offset<64>(dl_min, dl_max);
}

/* One (query, choice) through the Peq lane the cdist kernels run. */
int blh_peq(const u8 *q, int m, const u8 *t, int n, int k)
{
    const int64_t off[2] = {0, n};
    const HostStripes hs(t, off, 1);
    return peq_host(q, m, hs.text(0), t, n, k);
}

/* One (query, choice) through the exact-search lane in 32-row units; m <= 512. */
int blh_units(const u8 *q, int m, const u8 *t, int n, int k)
{
    if (m < 1 || m > 512) return -1;
    std::vector<u64> T((size_t)256 * 8, 0);
    for (int i = 0; i < m; i++) T[(size_t)(i >> 6) * 256 + q[i]] |= (u64)1 << (i & 63);
    const int64_t off[2] = {0, n};
    const HostStripes hs(t, off, 1);
    const StripeText st = hs.text(0);
    switch ((m + 31) >> 5) {
    case 1: return units_host<1>(T.data(), m, st, n, k);
    case 2: return units_host<2>(T.data(), m, st, n, k);
    case 3: return units_host<3>(T.data(), m, st, n, k);
    case 4: return units_host<4>(T.data(), m, st, n, k);
    case 5: return units_host<5>(T.data(), m, st, n, k);
    case 6: return units_host<6>(T.data(), m, st, n, k);
    case 7: return units_host<7>(T.data(), m, st, n, k);
    case 8: return units_host<8>(T.data(), m, st, n, k);
    case 9: return units_host<9>(T.data(), m, st, n, k);
    case 10: return units_host<10>(T.data(), m, st, n, k);
    case 11: return units_host<11>(T.data(), m, st, n, k);
    case 12: return units_host<12>(T.data(), m, st, n, k);
    case 13: return units_host<13>(T.data(), m, st, n, k);
    case 14: return units_host<14>(T.data(), m, st, n, k);
    case 15: return units_host<15>(T.data(), m, st, n, k);
    default: return units_host<16>(T.data(), m, st, n, k);
    }
}

/* The wavefront path for any lengths (is_pair: trim/swap as pairs do). */
int blh_long(const u8 *p, int m, const u8 *t, int n, int is_pair, int trim, int k)
{
    return long_sim(p, m, t, n, is_pair, trim, k);
}

/* Packed batches, int32 results. */
int blh_pairs(const u8 *ba, const int64_t *oa, const u8 *bb, const int64_t *ob, int64_t n,
              int32_t *out, int trim, int k, int force_long)
{
    for (int64_t i = 0; i < n; i++) {
        const u8 *a = ba + oa[i], *b = bb + ob[i];
        const int la = (int)(oa[i + 1] - oa[i]), lb = (int)(ob[i + 1] - ob[i]);
        out[i] = force_long ? long_sim(a, la, b, lb, 1, trim, k) : pair_host(a, la, b, lb, trim, k);
    }
    return 0;
}

int blh_cross(const u8 *bq, const int64_t *oq, int64_t nq, const u8 *bc, const int64_t *oc,
              int64_t nc, int32_t *out, int k, int force_long)
{
    const HostStripes hs(bc, oc, nc);
    for (int64_t i = 0; i < nq; i++)
        for (int64_t j = 0; j < nc; j++) {
            const u8 *q = bq + oq[i], *t = bc + oc[j];
            const int m = (int)(oq[i + 1] - oq[i]), n = (int)(oc[j + 1] - oc[j]);
            out[i * nc + j] = force_long ? long_sim(q, m, t, n, 0, 0, k)
                                         : peq_host(q, m, hs.text(j), t, n, k);
        }
    return 0;
}

/* band > 0: tiles of at least `band` 32-row units run as band tiles (one
 * window offset per tile from its lanes' length differences, 32 rows when
 * every threshold is <= 14, else 64, as pairs_tile_band does); undecided pairs
 * go through the whole scan 64 at a time in the units of the widest. (The
 * kernel's waves keep them on per-wave stacks; which pairs share a scan never
 * changes a result.) *escalated receives their number. */
int blh_pairs_batched_band(const u8 *ba, const int64_t *oa, const u8 *bb, const int64_t *ob,
                           int64_t n, int32_t *out, int trim, int k, int bin_width, int64_t chunk,
                           int band, int64_t *escalated)
{
    std::vector<int64_t> esc;
    auto band_tile = [&](const std::vector<int32_t> &idx) {
        struct L { int64_t i; const u8 *a, *b; int la, lb, res; };
        std::vector<L> ln;
        int dmin = 0x7fffffff, dmax = -0x7fffffff;
        bool narrow = true;
        for (int32_t i : idx) {
            L x{i, ba + oa[i], bb + ob[i], (int)(oa[i + 1] - oa[i]), (int)(ob[i + 1] - ob[i]), 0};
            x.res = band_prepare(x.a, x.la, x.b, x.lb, trim, k);
            if (x.res < 0) {
                const int dl = x.la - x.lb;
                dmin = dl < dmin ? dl : dmin;
                dmax = dl > dmax ? dl : dmax;
                if (!(k >= 0 && k <= 14)) narrow = false;
            }
            ln.push_back(x);
        }
        for (L &x : ln) {
            if (x.res < 0)
                x.res = narrow ? band_pair<u32>(x.a, x.la, x.b, x.lb, k, band_offset<32>(dmin, dmax))
                               : band_pair<u64>(x.a, x.la, x.b, x.lb, k, band_offset<64>(dmin, dmax));
            if (x.res >= 0) out[x.i] = x.res;
            else esc.push_back(x.i);
        }
    };
    const int rc = blh_pairs_batched_core(ba, oa, bb, ob, n, out, trim, k, bin_width, chunk, band,
                                          band_tile);
    for (size_t t0 = 0; t0 < esc.size(); t0 += 64) {
        const size_t t1 = t0 + 64 < esc.size() ? t0 + 64 : esc.size();
        int w = 0;
        for (size_t e = t0; e < t1; e++) {
            const int64_t i = esc[e];
            const int64_t la = oa[i + 1] - oa[i], lb = ob[i + 1] - ob[i];
            const int u = (int)(((la < lb ? la : lb) + 31) >> 5);
            w = u > w ? u : w;
        }
        for (size_t e = t0; e < t1; e++) out[esc[e]] = w32_pair(w, ba, oa, bb, ob, esc[e], trim, k);
    }
    if (escalated) *escalated = (int64_t)esc.size();
    return rc;
}

int blh_pairs_batched(const u8 *ba, const int64_t *oa, const u8 *bb, const int64_t *ob, int64_t n,
                      int32_t *out, int trim, int k, int bin_width, int64_t chunk)
{
    return blh_pairs_batched_band(ba, oa, bb, ob, n, out, trim, k, bin_width, chunk, 0, nullptr);
}

/* The same filter's SWAR test bound > k (k <= 3): bit 7 of byte s. */
/* Byte-class profile of s[0..n) into out[0..3], and the class-count test of
 * the search kernels for two strings at threshold k (1: d > k by the bound). */
void blh_class_profile(const u8 *s, int n, uint32_t *out)
{
    const Profile p = class_profile(s, n);
    out[0] = p.x; out[1] = p.y; out[2] = p.z; out[3] = p.w;
}

int blh_count_over(const u8 *a, int m, const u8 *b, int n, int k)
{
    return count_over(class_profile(a, m), class_profile(b, n), m, n, k) ? 1 : 0;
}

/* The class-count L2 test of the search kernels' matrix-core filter. */
int blh_count_l2_over(const u8 *a, int m, const u8 *b, int n, int k)
{
    return count_l2_over(class_profile(a, m), class_profile(b, n), k) ? 1 : 0;
}

void blh_class2_profile(const u8 *s, int n, uint32_t *out)
{
    const Profile p = class2_profile(s, n);
    out[0] = p.x; out[1] = p.y; out[2] = p.z; out[3] = p.w;
}

int blh_count2_over(const u8 *a, int m, const u8 *b, int n, int k)
{
    return count2_over(class2_profile(a, m), class2_profile(b, n), m, n, k) ? 1 : 0;
}

void blh_class3_profile(const u8 *s, int n, uint32_t *out)
{
    const Profile3 p = class3_profile(s, n);
    for (int i = 0; i < 4; i++) {
        out[4 * i] = p.p[i].x; out[4 * i + 1] = p.p[i].y; out[4 * i + 2] = p.p[i].z; out[4 * i + 3] = p.p[i].w;
    }
}

int blh_count3_over(const u8 *a, int m, const u8 *b, int n, int k)
{
    return count3_over(class3_profile(a, m), class3_profile(b, n), m, n, k) ? 1 : 0;
}

uint32_t blh_swar_over(const u8 *q, int m, const u8 *t, const int *n, int k)
{
    std::vector<u8> lo8(256, 0);
    for (int i = 0; i < 8 && i < m; i++) lo8[q[i]] |= (u8)(1u << i);
    u32 d[4];
    int nn[4];
    for (int s = 0; s < 4; s++) {
        memcpy(&d[s], t + 4 * s, 4);
        nn[s] = n[s];
    }
    u32 Pv, Mv;
    swar_steps4(lo8.data(), d, Pv, Mv, 1);
    return swar_over4(Pv, Mv, m, nn, k);
}

/* The 4-choice byte-segment filter: bounds of choices t[0..3] (first 4 bytes
 * each, lengths n[0..3]) against a pattern of m >= 8 bytes. */
void blh_swar_bounds(const u8 *q, int m, const u8 *t, const int *n, int *bound)
{
    std::vector<u32> lo32(256, 0);
    for (int i = 0; i < m && i < 32; i++) lo32[q[i]] |= 1u << i;
    u32 d[4];
    int nn[4], b[4];
    for (int s = 0; s < 4; s++) {
        memcpy(&d[s], t + 4 * s, 4);
        nn[s] = n[s];
    }
    swar_bounds4((const u8 *)lo32.data(), d, m, nn, b);
    for (int s = 0; s < 4; s++) bound[s] = b[s];
}

/* Output conversion exactly as the kernels store it. */
int blh_store(int d, int dtype)
{
    union { int32_t i; uint16_t h; u8 b; } o;
    o.i = 0;
    store_out(&o, 0, d, dtype);
    return dtype == DT_UINT8 ? o.b : dtype == DT_UINT16 ? o.h : o.i;
}

/* Pieces the tests pin individually. */
uint64_t blh_transpose8(uint64_t x) { return transpose8(x); }

/* The work-queue schedule of the persistent kernels. */
int blh_queue_of(int64_t wid, int64_t waves) { return queue_of(wid, waves); }
int64_t blh_queue_run(int64_t wid, int64_t waves, int first_static, int64_t taken)
{
    return queue_run(wid, waves, first_static != 0, taken);
}
uint32_t blh_bop3(uint32_t a, uint32_t b, uint32_t c, int which)
{
    switch (which) {
    case 0: return bop3<tt((TA ^ TB) | TC)>(a, b, c);
    case 1: return bop3<tt(TA | ~(TB | TC))>(a, b, c);
    case 2: return bop3<tt(TA & ~(TB ^ TC))>(a, b, c);
    default: return 0;
    }
}
void blh_trim(const u8 *a, int la, const u8 *b, int lb, int *res)
{
    const u8 *a0 = a, *b0 = b;
    trim_pair(a, la, b, lb);
    res[0] = (int)(a - a0);
    res[1] = la;
    res[2] = (int)(b - b0);
    res[3] = lb;
}

}  /* extern "C" */
