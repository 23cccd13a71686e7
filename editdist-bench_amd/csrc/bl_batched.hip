/*
 * bl_batched.hip -- the batched persistent form for gfx950.
 *
 * cdist (bl_cdist): one persistent launch. Waves pull (choice block, query)
 * tiles from an atomic work queue, block-major so that every query of a block
 * reads the same choices while they sit in L2 / MALL. Each wave builds its own
 * query's Peq table in its slice of LDS and keeps all 64 lanes busy: idle
 * lanes refill from a per-wave LDS queue of in-band candidates, which the wave
 * fills 64 choices at a time with one coalesced offsets load; choices the
 * length gap (|n - m| > k) or an empty string decide are answered right there
 * and never occupy a lane. Reference: ed_cross_w{NW} + the host class split,
 * csrc/gpu_kernels.cu:36-69, bitlev/gpu.py:106-138.
 *
 * pairwise (bl_pairs): on-device length bucketing, then one persistent launch.
 *   1. classify: trim, swap so the pattern is the shorter string, answer the
 *      trivial pairs (gap, empty) at once, and give every other pair a bucket
 *      key from (pattern words, text length), privatised histogram in LDS;
 *   2. scan: bucket bases, most expensive bucket first (longest-job-first);
 *   3. scatter: pair indices into bucket order;
 *   4. persistent kernel: waves pull 64-pair tiles; a tile's pattern word count
 *      is uniform, so its lanes run the same register-resident code with
 *      text lengths within one bucket width. Long pairs (pattern > 256 bytes)
 *      take the wave-cooperative wavefront inside the same launch.
 * Reference: ed_pairs_c{CAP}, csrc/gpu_kernels.cu:6-32, bitlev/gpu.py:71-103.
 *
 * Work-queue counters live in the caller's workspace (or a stream-ordered
 * allocation) and are cleared by a kernel, never by a memset, so the whole
 * pipeline can be captured into a HIP graph.
 */
#include <hip/hip_runtime.h>

#include "bl_device.h"
#include "bl_internal.h"

using namespace bl;

namespace {

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

__host__ __device__ __forceinline__ int clamp_k_dev(int k)
{
    return k < 0 ? -1 : (k > (1 << 30) ? (1 << 30) : k);
}

__device__ __forceinline__ int rank_below(u64 mask)
{
    return (int)__builtin_amdgcn_mbcnt_hi((u32)(mask >> 32), __builtin_amdgcn_mbcnt_lo((u32)mask, 0u));
}

__device__ __forceinline__ u64 wave_ballot(bool p) { return __ballot(p); }

/* Values every lane holds identically, made scalar so that branches on them
 * are uniform by construction. */
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t uni64(int64_t v)
{
    const u32 lo = __builtin_amdgcn_readfirstlane((u32)v);
    const u32 hi = __builtin_amdgcn_readfirstlane((u32)((u64)v >> 32));
    return (int64_t)(((u64)hi << 32) | lo);
}

/* Make this wave's LDS writes visible to its own later LDS reads by other lanes. */
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* Next work-queue index for the whole wave, as a scalar. Every lane executes
 * the atomic (the first active lane adds `step`, the others add 0) and the
 * wave takes that lane's ticket with readfirstlane. There must be no lane-specific branch here:
 * with `if (lane == 0) atomicAdd(...)` the loop that calls this is compiled
 * as divergent, and after a `continue` the structurizer sends only lane 0 back
 * through the atomic while the other lanes re-read a stale (zero) ticket in an
 * inner loop -- the same tile forever. */
__device__ __forceinline__ long long wave_take(unsigned long long *counter, long long step)
{
    /* the first active lane adds the step; readfirstlane reads that lane */
    const bool first = rank_below(__builtin_amdgcn_read_exec()) == 0;
    const unsigned long long mine = first ? (unsigned long long)step : 0ull;
    const unsigned long long t = atomicAdd(counter, mine);
    const u32 lo = __builtin_amdgcn_readfirstlane((u32)t);
    const u32 hi = __builtin_amdgcn_readfirstlane((u32)(t >> 32));
    return (long long)(((u64)hi << 32) | lo);
}

__global__ void clear_kernel(unsigned long long *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

/* ======================================================================= cdist */
struct QEnt {
    int64_t start;
    int32_t n;
    int32_t rel;
};

/* One (query, choice block) tile for one wave. peq: this wave's LDS table
 * (row c at peq[c * NW]); queue: this wave's LDS ring of QCAP entries. */
template <int NW, int QCAP>
__device__ void cdist_tile(const u64 *peq, QEnt *queue, int m, const u8 *__restrict__ bc,
                           const int64_t *__restrict__ oc, int64_t c0, int64_t c1, void *out,
                           int64_t orow, int dtype, int k)
{
    const int lane = lane_id();
    const int nw = NW == 1 ? 1 : (m + 63) >> 6;
    bool active = false;
    u64 Pv[NW], Mv[NW];
    ByteStream ts;
    u32 c4 = 0;
    int j = 0, n = 0, rel = 0, dmn = 0;
    int64_t wnext = c0;
    int qh = 0, qt = 0;
    for (;;) {
        /* ---- refill idle lanes from the queue; top the queue up from the window */
        for (;;) {
            const u64 need = wave_ballot(!active);
            const int nneed = __popcll(need);
            if (nneed == 0) break;
            const int avail = qt - qh;
            if (avail >= nneed || wnext >= c1) {
                const int r = rank_below(need);
                if (!active && r < avail) {
                    const QEnt e = queue[(qh + r) & (QCAP - 1)];
                    n = e.n;
                    rel = e.rel;
                    dmn = m - n;
                    ts.open(bc + e.start, n);
                    c4 = ts.next4();
                    j = 0;
#pragma unroll
                    for (int w = 0; w < NW; w++) { Pv[w] = ~(u64)0; Mv[w] = 0; }
                    active = true;
                }
                qh = uni(qh + (avail < nneed ? avail : nneed));
                break;
            }
            const int64_t c = wnext + lane;
            const bool valid = c < c1;
            int64_t o0 = 0, o1 = 0;
            if (valid) { o0 = oc[c]; o1 = oc[c + 1]; }
            const int nn = (int)(o1 - o0);
            const int gap = nn > m ? nn - m : m - nn;
            const bool decided = valid && ((k >= 0 && gap > k) || nn == 0);
            if (decided) store_out(out, orow + c, (k >= 0 && gap > k) ? k + 1 : m, dtype);
            const bool scan = valid && !decided;
            const u64 sm = wave_ballot(scan);
            if (scan) {
                QEnt e;
                e.start = o0;
                e.n = nn;
                e.rel = (int)(c - c0);
                queue[(qt + rank_below(sm)) & (QCAP - 1)] = e;
            }
            qt = uni(qt + __popcll(sm));
            wnext = uni64(wnext + 64);
            wave_lds_sync();
        }
        if (wave_ballot(active) == 0) break;
        /* ---- four columns for every busy lane */
        if (active) {
            const u32 c4n = ts.next4();
            const int cnt = n - j < 4 ? n - j : 4;
            if (cnt == 4) {
                u64 E[4][NW];
#pragma unroll
                for (int u = 0; u < 4; u++)
#pragma unroll
                    for (int w = 0; w < NW; w++) E[u][w] = peq[((c4 >> (8 * u)) & 0xFFu) * NW + w];
#pragma unroll
                for (int u = 0; u < 4; u++) column<NW>(Pv, Mv, nw, E[u]);
            } else {
#pragma unroll
                for (int u = 0; u < 3; u++) {
                    if (u < cnt) {
                        u64 E[NW];
#pragma unroll
                        for (int w = 0; w < NW; w++) E[w] = peq[((c4 >> (8 * u)) & 0xFFu) * NW + w];
                        column<NW>(Pv, Mv, nw, E);
                    }
                }
            }
            j += cnt;
            if (k >= 0 || j >= n) {
                const int bound = diag_bound<NW>(Pv, Mv, nw, j, m, m - dmn);
                const bool over = k >= 0 && bound > k;
                if (over || j >= n) {
                    store_out(out, orow + c0 + rel, over ? k + 1 : bound, dtype);
                    active = false;
                }
            }
            c4 = c4n;
        }
    }
}

/* Strided mode: lane l takes choices c0 + l, c0 + l + 64, ... and runs each to
 * its end (the plain kernel's lane loop inside the persistent tiles). Cheaper
 * than queue refill when every choice needs about the same number of columns. */
template <int NW>
__device__ void cdist_tile_strided(const u64 *peq, int m, const u8 *__restrict__ bc,
                                   const int64_t *__restrict__ oc, int64_t c0, int64_t c1,
                                   void *out, int64_t orow, int dtype, int k)
{
    for (int64_t c = c0 + lane_id(); c < c1; c += 64) {
        const int64_t o0 = oc[c];
        const int n = (int)(oc[c + 1] - o0);
        store_out(out, orow + c, peq_lane<NW>(peq, m, bc + o0, n, k), dtype);
    }
}

template <int NW>
__device__ void build_peq_wave(u64 *peq, const u8 *__restrict__ p, int m)
{
    const int lane = lane_id();
#pragma unroll
    for (int x = 0; x < 256 * NW / 64; x++) peq[x * 64 + lane] = 0;
    wave_lds_sync();
    for (int i = lane; i < m; i += 64)
        atomicOr((unsigned long long *)&peq[(int)p[i] * NW + (i >> 6)],
                 (unsigned long long)1 << (i & 63));
    wave_lds_sync();
}

constexpr int CD_QCAP = 128;

/* Persistent cdist over queries of at most 64*NWMAX bytes. LDS per wave:
 * 256*NWMAX*8 bytes of Peq + CD_QCAP queue entries. */
template <int NWMAX>
__global__ __launch_bounds__(256) void cdist_batched_kernel(
    const u8 *__restrict__ bq, const int64_t *__restrict__ oq, int64_t nq,
    const u8 *__restrict__ bc, const int64_t *__restrict__ oc, int64_t nc, void *out, int dtype,
    int k_all, const int32_t *__restrict__ kvec, int64_t tile, int strided,
    unsigned long long *counter)
{
    extern __shared__ u64 lds[];
    const int wave = threadIdx.x >> 6;
    const int per_wave = 256 * NWMAX + CD_QCAP * (int)(sizeof(QEnt) / sizeof(u64));
    u64 *peq = lds + wave * per_wave;
    QEnt *queue = (QEnt *)(peq + 256 * NWMAX);
    const int64_t nblk = (nc + tile - 1) / tile;
    const int64_t ntiles = nblk * nq;
    int64_t cur_q = -1;
    int cur_nw = 0;
    for (;;) {
        const long long t = wave_take(counter, 1);
        if (t >= ntiles) break;
        const int64_t blk = t / nq, q = t % nq;
        const int64_t c0 = blk * tile;
        const int64_t c1 = c0 + tile < nc ? c0 + tile : nc;
        const int64_t q0 = oq[q];
        const int m = (int)(oq[q + 1] - q0);
        const int64_t orow = q * nc;
        const int k = kvec ? clamp_k_dev(kvec[q]) : k_all;
        if (m == 0 || m > 64 * NWMAX) {
            /* empty query: n, or k + 1 (csrc/gpu_kernels.cu:53-54); too long for this
             * launch: caller error, marked -1 */
            for (int64_t c = c0 + lane_id(); c < c1; c += 64) {
                const int nn = (int)(oc[c + 1] - oc[c]);
                store_out(out, orow + c, m == 0 ? empty_pattern(nn, k) : -1, dtype);
            }
            continue;
        }
        const int nwq = (m + 63) >> 6;
        const int cls = nwq <= 1 ? 1 : nwq <= 2 ? 2 : nwq <= 4 ? 4 : 8;
        if (q != cur_q || cls != cur_nw) {
            if (cls == 1) build_peq_wave<1>(peq, bq + q0, m);
            else if (NWMAX >= 2 && cls == 2) build_peq_wave<NWMAX >= 2 ? 2 : 1>(peq, bq + q0, m);
            else if (NWMAX >= 4 && cls == 4) build_peq_wave<NWMAX >= 4 ? 4 : 1>(peq, bq + q0, m);
            else if (NWMAX >= 8) build_peq_wave<NWMAX >= 8 ? 8 : 1>(peq, bq + q0, m);
            cur_q = q;
            cur_nw = cls;
        }
        if (strided) {
            if (cls == 1) cdist_tile_strided<1>(peq, m, bc, oc, c0, c1, out, orow, dtype, k);
            else if (NWMAX >= 2 && cls == 2)
                cdist_tile_strided<NWMAX >= 2 ? 2 : 1>(peq, m, bc, oc, c0, c1, out, orow, dtype, k);
            else if (NWMAX >= 4 && cls == 4)
                cdist_tile_strided<NWMAX >= 4 ? 4 : 1>(peq, m, bc, oc, c0, c1, out, orow, dtype, k);
            else if (NWMAX >= 8)
                cdist_tile_strided<NWMAX >= 8 ? 8 : 1>(peq, m, bc, oc, c0, c1, out, orow, dtype, k);
            continue;
        }
        if (cls == 1) cdist_tile<1, CD_QCAP>(peq, queue, m, bc, oc, c0, c1, out, orow, dtype, k);
        else if (NWMAX >= 2 && cls == 2)
            cdist_tile<NWMAX >= 2 ? 2 : 1, CD_QCAP>(peq, queue, m, bc, oc, c0, c1, out, orow, dtype, k);
        else if (NWMAX >= 4 && cls == 4)
            cdist_tile<NWMAX >= 4 ? 4 : 1, CD_QCAP>(peq, queue, m, bc, oc, c0, c1, out, orow, dtype, k);
        else if (NWMAX >= 8)
            cdist_tile<NWMAX >= 8 ? 8 : 1, CD_QCAP>(peq, queue, m, bc, oc, c0, c1, out, orow, dtype, k);
    }
}

/* ======================================================================= pairs */
/* Pairs are bucketed within chunks of CH consecutive pairs (CH a multiple of
 * 256), chunk-major, so the 64 pairs of a tile lie close together in the
 * string buffers (a global sort would scatter every tile over gigabytes and
 * thrash the TLB). Per chunk c the order is: long pairs, then buckets by
 * descending key. Arrays: hist / cursor [nchunks * PKEYS], chunk_base and
 * tile_base [nchunks + 1]. */
__global__ __launch_bounds__(256) void pairs_classify_kernel(
    const u8 *__restrict__ ba, const int64_t *__restrict__ oa, const u8 *__restrict__ bb,
    const int64_t *__restrict__ ob, int64_t n, void *out, int dtype, int trim, int k_all,
    const int32_t *__restrict__ kvec, int bin_width, int64_t chunk, PDesc *desc,
    unsigned short *key, unsigned long long *hist)
{
    __shared__ unsigned int lh[PKEYS];
    for (int x = threadIdx.x; x < PKEYS; x += blockDim.x) lh[x] = 0;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const int k = kvec ? clamp_k_dev(kvec[i]) : k_all;
        PDesc d;
        int result = 0;
        const int kk = classify_pair(ba, oa, bb, ob, i, trim, k, bin_width, d, result);
        if (kk == 0) store_out(out, i, result, dtype);
        else desc[i] = d;
        key[i] = (unsigned short)kk;
        if (kk) atomicAdd(&lh[kk], 1u);
    }
    __syncthreads();
    /* a block never straddles chunks (chunk is a multiple of the block size) */
    const int64_t c = ((int64_t)blockIdx.x * blockDim.x) / chunk;
    unsigned long long *h = hist + c * PKEYS;
    for (int x = threadIdx.x; x < PKEYS; x += blockDim.x)
        if (lh[x]) atomicAdd(&h[x], (unsigned long long)lh[x]);
}

/* One block per chunk: within-chunk bases in descending key order (cursor),
 * the chunk's pair count and tile count (long pairs one tile each). */
__global__ __launch_bounds__(256) void pairs_chunk_scan_kernel(
    const unsigned long long *hist, unsigned long long *cursor, unsigned long long *chunk_cnt,
    unsigned long long *tile_cnt)
{
    __shared__ unsigned long long h[PKEYS];
    const int64_t c = blockIdx.x;
    for (int x = threadIdx.x; x < PKEYS; x += blockDim.x) h[x] = hist[c * PKEYS + x];
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long acc = 0;
        for (int x = PKEYS - 1; x >= 1; x--) {
            const unsigned long long v = h[x];
            h[x] = acc;
            acc += v;
        }
        chunk_cnt[c] = acc;
        const unsigned long long nlong = hist[c * PKEYS + PKEY_LONG];
        tile_cnt[c] = nlong + (acc - nlong + 63) / 64;
    }
    __syncthreads();
    for (int x = threadIdx.x; x < PKEYS; x += blockDim.x) cursor[c * PKEYS + x] = h[x];
}

/* Exclusive prefix sums of the per-chunk pair and tile counts (one block of
 * 1024 threads, any number of chunks): base[c] for c <= nchunks. */
__global__ __launch_bounds__(1024) void prefix_kernel(const unsigned long long *cnt_a,
                                                      unsigned long long *base_a,
                                                      const unsigned long long *cnt_b,
                                                      unsigned long long *base_b, int64_t nchunks)
{
    __shared__ unsigned long long sa[1024], sb[1024];
    const int t = threadIdx.x;
    const int64_t per = (nchunks + 1023) / 1024;
    const int64_t lo = t * per, hi = lo + per < nchunks ? lo + per : nchunks;
    unsigned long long a = 0, b = 0;
    for (int64_t c = lo; c < hi; c++) { a += cnt_a[c]; b += cnt_b[c]; }
    sa[t] = a;
    sb[t] = b;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {            /* inclusive Hillis-Steele scan */
        const unsigned long long xa = t >= off ? sa[t - off] : 0, xb = t >= off ? sb[t - off] : 0;
        __syncthreads();
        sa[t] += xa;
        sb[t] += xb;
        __syncthreads();
    }
    unsigned long long ra = sa[t] - a, rb = sb[t] - b;      /* exclusive start of this thread */
    for (int64_t c = lo; c < hi; c++) {
        base_a[c] = ra;
        base_b[c] = rb;
        ra += cnt_a[c];
        rb += cnt_b[c];
    }
    if (t == 1023) {
        base_a[nchunks] = sa[t];
        base_b[nchunks] = sb[t];
    }
}

__global__ __launch_bounds__(256) void pairs_scatter_kernel(
    const unsigned short *key, int64_t n, int64_t chunk, unsigned long long *cursor,
    const unsigned long long *chunk_base, int32_t *sorted)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int kk = key[i];
    if (!kk) return;
    const int64_t c = i / chunk;
    sorted[chunk_base[c] + atomicAdd(&cursor[c * PKEYS + kk], 1ull)] = (int32_t)i;
}

template <int NW>
__device__ void pairs_tile(const u8 *__restrict__ ba, const u8 *__restrict__ bb, const PDesc *desc,
                           const int32_t *sorted, int64_t s0, int64_t s1, void *out, int dtype)
{
    const int64_t s = s0 + lane_id();
    if (s >= s1) return;
    const int64_t i = sorted[s];
    store_out(out, i, desc_distance<NW>(ba, bb, desc[i]), dtype);
}

__device__ void pairs_long_item(const u8 *__restrict__ ba, const u8 *__restrict__ bb,
                                const PDesc &d, int64_t i, void *out, int dtype,
                                int8_t *hbuf, int64_t slot)
{
    const int k = d.k;
    const int lane = lane_id();
    const bool sw = uni(d.la) < 0;
    const int m = uni(d.la & 0x7fffffff), n = uni(d.lb);
    LongCtx C;
    C.p = (sw ? bb : ba) + uni64(d.a);
    C.m = m;
    C.t = (sw ? ba : bb) + uni64(d.b);
    C.n = n;
    C.W = (m + 63) >> 6;
    C.hbuf = hbuf;
    const int S = (C.W + 63) >> 6;
    int r;
    if (S > 1 && (hbuf == nullptr || n > slot)) {
        r = -1;                             /* scratch too small: caller error */
    } else {
        LongLane L;
        for (int s = 0; s < S; s++) {
            C.s = s;
            C.nl = C.W - 64 * s < 64 ? C.W - 64 * s : 64;
            long_lane_init(C, L, lane);
            u32 prev = 0;
            for (int step = 0; step < n + C.nl - 1; step++) {
                u32 in = (u32)__shfl_up((int)prev, 1);
                if (lane == 0) in = long_source(C, step);
                prev = long_lane_step(C, L, lane, step, in);
            }
            __threadfence();
        }
        const int dd = m + __shfl(L.delta, (C.W - 1) - 64 * (S - 1));
        r = (k >= 0 && dd > k) ? k + 1 : dd;
    }
    if (lane == 0) store_out(out, i, r, dtype);
}

__global__ __launch_bounds__(256) void pairs_batched_kernel(
    const u8 *__restrict__ ba, const u8 *__restrict__ bb, const PDesc *__restrict__ desc,
    const int32_t *__restrict__ sorted, const unsigned short *__restrict__ key_of,
    const unsigned long long *__restrict__ hist, const unsigned long long *__restrict__ chunk_base,
    const unsigned long long *__restrict__ tile_base, int64_t nchunks, void *out, int dtype,
    unsigned long long *counter, int8_t *scratch, int64_t slot)
{
    const int64_t ntiles = (int64_t)tile_base[nchunks];
    const int64_t wave_gid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    int8_t *hbuf = scratch ? scratch + wave_gid * slot : nullptr;
    int64_t c = 0;
    for (;;) {
        const long long t = wave_take(counter, 1);
        if (t >= ntiles) break;
        /* chunk of tile t: tile_base[c] <= t < tile_base[c + 1]; tiles are taken in
         * increasing order, so search upwards from the last chunk, then bisect */
        if ((int64_t)tile_base[c + 1] <= t || (int64_t)tile_base[c] > t) {
            int64_t lo = 0, hi = nchunks;
            while (hi - lo > 1) {
                const int64_t mid = (lo + hi) >> 1;
                if ((int64_t)tile_base[mid] <= t) lo = mid;
                else hi = mid;
            }
            c = uni64(lo);
        }
        const int64_t lt = t - (int64_t)tile_base[c];
        const int64_t cb = (int64_t)chunk_base[c];
        const int64_t nlong = (int64_t)hist[c * PKEYS + PKEY_LONG];
        if (lt < nlong) {                      /* long pairs lead their chunk, one per tile */
            const int64_t i = sorted[cb + lt];
            pairs_long_item(ba, bb, desc[i], i, out, dtype, hbuf, slot);
            continue;
        }
        const int64_t s0 = cb + nlong + (lt - nlong) * 64;
        const int64_t ce = (int64_t)chunk_base[c + 1];
        const int64_t s1 = s0 + 64 < ce ? s0 + 64 : ce;
        /* descending keys: the first pair of a tile has its largest word count */
        const int w = uni(key_words(key_of[sorted[s0]]));
        if (w <= 1) pairs_tile<1>(ba, bb, desc, sorted, s0, s1, out, dtype);
        else if (w == 2) pairs_tile<2>(ba, bb, desc, sorted, s0, s1, out, dtype);
        else if (w == 3) pairs_tile<3>(ba, bb, desc, sorted, s0, s1, out, dtype);
        else pairs_tile<4>(ba, bb, desc, sorted, s0, s1, out, dtype);
    }
}

int launch_clear(unsigned long long *p, int n, hipStream_t s)
{
    clear_kernel<<<(n + 255) / 256, 256, 0, s>>>(p, n);
    BL_LAUNCHED("clear_kernel");
    return BL_OK;
}

int device_cus()
{
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        if (hipGetDevice(&dev) == hipSuccess) {
            hipDeviceProp_t p;
            if (hipGetDeviceProperties(&p, dev) == hipSuccess) cus = p.multiProcessorCount;
        }
        if (!cus) cus = 256;
    }
    return cus;
}

}  // namespace

/* ======================================================================= C ABI */
extern "C" {

size_t bl_cdist_workspace_size(void) { return 256; }

int bl_cdist_tuned(const uint8_t *d_buf_q, const int64_t *d_off_q, int64_t nq,
                   const uint8_t *d_buf_c, const int64_t *d_off_c, int64_t nc, int max_qlen,
                   void *d_out, int dtype, int k, const int32_t *d_kvec, void *d_work,
                   size_t work_bytes, const bl_cdist_params *prm, void *stream)
{
    if (nq < 0 || nc < 0 || max_qlen < 0) return fail(BL_EINVAL, "bl_cdist: negative size");
    if (nq == 0 || nc == 0) return BL_OK;
    if (dtype < BL_INT32 || dtype > BL_UINT8) return fail(BL_EINVAL, "bl_cdist: bad dtype %d", dtype);
    if (!d_out || !d_buf_q || !d_off_q || !d_buf_c || !d_off_c)
        return fail(BL_EINVAL, "bl_cdist: null pointer");
    if (max_qlen > 512) return fail(BL_EINVAL, "bl_cdist: queries above 512 bytes go to bl_long");
    if (d_work && work_bytes < bl_cdist_workspace_size())
        return fail(BL_EINVAL, "bl_cdist: workspace below %zu bytes", bl_cdist_workspace_size());
    k = clamp_k_dev(k);
    hipStream_t s = (hipStream_t)stream;
    const int wpc = prm && prm->waves_per_cu > 0 ? prm->waves_per_cu : 16;
    const int64_t tile = prm && prm->tile > 0 ? prm->tile : 8192;
    const int strided = prm && prm->reserved[0] == 1 ? 1 : 0;
    const int words = (max_qlen + 63) >> 6;
    const int nwmax = words <= 1 ? 1 : words <= 2 ? 2 : words <= 4 ? 4 : 8;
    const size_t per_wave = (size_t)(256 * nwmax) * 8 + (size_t)CD_QCAP * sizeof(QEnt);
    const size_t lds = per_wave * 4;
    int64_t blocks = (int64_t)device_cus() * wpc / 4;
    const int64_t ntiles = ((nc + tile - 1) / tile) * nq;
    const int64_t need = (ntiles + 3) / 4;
    if (blocks > need) blocks = need;
    if (blocks < 1) blocks = 1;
    unsigned long long *ctr = (unsigned long long *)d_work;
    if (!ctr) BL_HIP(hipMallocAsync((void **)&ctr, 256, s));
    int rc = launch_clear(ctr, 1, s);
    if (rc == BL_OK) {
        switch (nwmax) {
        case 1:
            cdist_batched_kernel<1><<<(unsigned)blocks, 256, lds, s>>>(
                d_buf_q, d_off_q, nq, d_buf_c, d_off_c, nc, d_out, dtype, k, d_kvec, tile, strided, ctr);
            break;
        case 2:
            cdist_batched_kernel<2><<<(unsigned)blocks, 256, lds, s>>>(
                d_buf_q, d_off_q, nq, d_buf_c, d_off_c, nc, d_out, dtype, k, d_kvec, tile, strided, ctr);
            break;
        case 4:
            cdist_batched_kernel<4><<<(unsigned)blocks, 256, lds, s>>>(
                d_buf_q, d_off_q, nq, d_buf_c, d_off_c, nc, d_out, dtype, k, d_kvec, tile, strided, ctr);
            break;
        default:
            cdist_batched_kernel<8><<<(unsigned)blocks, 256, lds, s>>>(
                d_buf_q, d_off_q, nq, d_buf_c, d_off_c, nc, d_out, dtype, k, d_kvec, tile, strided, ctr);
            break;
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = fail(BL_EHIP, "cdist_batched_kernel launch: %s", hipGetErrorString(e));
    }
    if (!d_work) (void)hipFreeAsync(ctr, s);
    return rc;
}

int bl_cdist(const uint8_t *d_buf_q, const int64_t *d_off_q, int64_t nq, const uint8_t *d_buf_c,
             const int64_t *d_off_c, int64_t nc, int max_qlen, void *d_out, int dtype, int k,
             void *d_work, size_t work_bytes, void *stream)
{
    return bl_cdist_tuned(d_buf_q, d_off_q, nq, d_buf_c, d_off_c, nc, max_qlen, d_out, dtype, k,
                          nullptr, d_work, work_bytes, nullptr, stream);
}

static int64_t pairs_chunk(int64_t n, const bl_pairs_params *prm)
{
    int64_t ch = prm && prm->reserved[0] > 0 ? prm->reserved[0] : 65536;
    const int64_t floor_ch = (n + 4095) / 4096;           /* at most 4096 chunks */
    if (ch < floor_ch) ch = floor_ch;
    return (ch + 255) / 256 * 256;
}

struct PairsLayout {
    size_t head, desc, key, sorted, scratch, total;
    int64_t chunk, nchunks;
};

/* Workspace layout. header_chunks: chunks the header is sized for (at least
 * the chunks this call uses), so one size can serve every launch shape. */
static PairsLayout pairs_layout(int64_t n, int64_t max_len, const bl_pairs_params *prm,
                                int64_t header_chunks = 0)
{
    PairsLayout L;
    L.chunk = pairs_chunk(n, prm);
    L.nchunks = (n + L.chunk - 1) / L.chunk;
    if (L.nchunks < 1) L.nchunks = 1;
    const int64_t hc = header_chunks > L.nchunks ? header_chunks : L.nchunks;
    /* head: counter | hist | cursor | chunk_cnt | tile_cnt | chunk_base | tile_base */
    const size_t nk = (size_t)hc * PKEYS;
    L.head = (64 + (2 * nk + 4 * ((size_t)hc + 1)) * 8 + 255) & ~(size_t)255;
    L.desc = L.head;
    L.key = (L.desc + (size_t)n * sizeof(PDesc) + 255) & ~(size_t)255;
    L.sorted = (L.key + (size_t)n * sizeof(unsigned short) + 255) & ~(size_t)255;
    L.scratch = (L.sorted + (size_t)n * sizeof(int32_t) + 255) & ~(size_t)255;
    L.total = L.scratch + (max_len > 4096 ? (size_t)max_len * 2048 : 0) + 256;
    return L;
}

/* Enough for any launch shape: chunks are at least 256 pairs and at most 4096. */
static int64_t max_chunks(int64_t n)
{
    const int64_t c = (n + 255) / 256;
    return c < 4096 ? (c < 1 ? 1 : c) : 4096;
}

size_t bl_pairs_workspace_size(int64_t n, int64_t max_len)
{
    return pairs_layout(n, max_len, nullptr, max_chunks(n)).total;
}

int bl_pairs_tuned(const uint8_t *d_buf_a, const int64_t *d_off_a, const uint8_t *d_buf_b,
                   const int64_t *d_off_b, int64_t n, int64_t max_len, void *d_out, int dtype,
                   int trim, int k, const int32_t *d_kvec, void *d_work, size_t work_bytes,
                   const bl_pairs_params *prm, void *stream)
{
    if (n < 0 || max_len < 0) return fail(BL_EINVAL, "bl_pairs: negative size");
    if (n == 0) return BL_OK;
    if (dtype < BL_INT32 || dtype > BL_UINT8) return fail(BL_EINVAL, "bl_pairs: bad dtype %d", dtype);
    if (!d_out || !d_buf_a || !d_off_a || !d_buf_b || !d_off_b)
        return fail(BL_EINVAL, "bl_pairs: null pointer");
    if (n > 0x7fffffffLL) return fail(BL_EINVAL, "bl_pairs: at most 2^31 - 1 pairs per call");
    if (max_len > 0x7fffffffLL) return fail(BL_EINVAL, "bl_pairs: strings must be shorter than 2^31");
    const PairsLayout L = pairs_layout(n, max_len, prm, max_chunks(n));
    if (d_work && work_bytes < L.total)
        return fail(BL_EINVAL, "bl_pairs: workspace of %zu bytes below the %zu needed", work_bytes,
                    L.total);
    k = clamp_k_dev(k);
    hipStream_t s = (hipStream_t)stream;
    const int bw = prm && prm->bin_width > 0 ? prm->bin_width : 16;
    const int wpc = prm && prm->waves_per_cu > 0 ? prm->waves_per_cu : 16;
    const int wg = prm && (prm->block == 64 || prm->block == 128 || prm->block == 256) ? prm->block : 256;

    char *w = (char *)d_work;
    if (!w) BL_HIP(hipMallocAsync((void **)&w, L.total, s));
    const size_t nk = (size_t)L.nchunks * PKEYS;
    unsigned long long *ctr = (unsigned long long *)w;
    unsigned long long *hist = (unsigned long long *)(w + 64);
    unsigned long long *cursor = hist + nk;
    unsigned long long *chunk_cnt = cursor + nk;
    unsigned long long *tile_cnt = chunk_cnt + (L.nchunks + 1);
    unsigned long long *chunk_base = tile_cnt + (L.nchunks + 1);
    unsigned long long *tile_base = chunk_base + (L.nchunks + 1);
    PDesc *desc = (PDesc *)(w + L.desc);
    unsigned short *key = (unsigned short *)(w + L.key);
    int32_t *sorted = (int32_t *)(w + L.sorted);
    int8_t *scratch = max_len > 4096 ? (int8_t *)(w + L.scratch) : nullptr;

    /* the counter and the histograms start at zero */
    int rc = launch_clear(ctr, (int)(8 + nk), s);
    const unsigned cblocks = (unsigned)((n + 255) / 256);
    if (rc == BL_OK) {
        pairs_classify_kernel<<<cblocks, 256, 0, s>>>(d_buf_a, d_off_a, d_buf_b, d_off_b, n, d_out,
                                                      dtype, trim, k, d_kvec, bw, L.chunk, desc,
                                                      key, hist);
        pairs_chunk_scan_kernel<<<(unsigned)L.nchunks, 256, 0, s>>>(hist, cursor, chunk_cnt, tile_cnt);
        prefix_kernel<<<1, 1024, 0, s>>>(chunk_cnt, chunk_base, tile_cnt, tile_base, L.nchunks);
        pairs_scatter_kernel<<<cblocks, 256, 0, s>>>(key, n, L.chunk, cursor, chunk_base, sorted);
        int64_t blocks = (int64_t)device_cus() * wpc * 64 / wg;
        const int64_t maxb = (n + 63) / 64 * 64 / wg + 1;
        if (blocks > maxb) blocks = maxb;
        if (scratch && blocks * (wg / 64) > 2048) blocks = 2048 / (wg / 64);
        pairs_batched_kernel<<<(unsigned)blocks, wg, 0, s>>>(d_buf_a, d_buf_b, desc, sorted, key,
                                                             hist, chunk_base, tile_base,
                                                             L.nchunks, d_out, dtype, ctr, scratch,
                                                             scratch ? max_len : 0);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = fail(BL_EHIP, "pairs pipeline launch: %s", hipGetErrorString(e));
    }
    if (!d_work) (void)hipFreeAsync(w, s);
    return rc;
}

int bl_pairs(const uint8_t *d_buf_a, const int64_t *d_off_a, const uint8_t *d_buf_b,
             const int64_t *d_off_b, int64_t n, int64_t max_len, void *d_out, int dtype, int trim,
             int k, void *d_work, size_t work_bytes, void *stream)
{
    return bl_pairs_tuned(d_buf_a, d_off_a, d_buf_b, d_off_b, n, max_len, d_out, dtype, trim, k,
                          nullptr, d_work, work_bytes, nullptr, stream);
}

}  /* extern "C" */
