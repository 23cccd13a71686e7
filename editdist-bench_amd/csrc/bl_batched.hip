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
 *   4. persistent kernel: waves pull 64-pair tiles; a tile's pattern size in
 *      32-row units is uniform, so its lanes run the same register-resident code with
 *      text lengths within one bucket width. Long pairs (pattern > 256 bytes)
 *      take the wave-cooperative wavefront inside the same launch.
 * Reference: ed_pairs_c{CAP}, csrc/gpu_kernels.cu:6-32, bitlev/gpu.py:71-103.
 *
 * Work-queue counters live in the caller's workspace (or a stream-ordered
 * allocation) and are cleared by a kernel, never by a memset, so the whole
 * pipeline can be captured into a HIP graph.
 */
#include <hip/hip_runtime.h>

#include <mutex>
#include <type_traits>

#include "bl_device.h"
#include "bl_internal.h"

using namespace bl;

namespace {

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

typedef int v16i __attribute__((ext_vector_type(16)));   /* one 32x32 MFMA accumulator */
typedef u32 v4u __attribute__((ext_vector_type(4)));     /* one 16-byte store */
typedef int v4i __attribute__((ext_vector_type(4)));     /* one 16-byte MFMA operand */

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

__device__ __forceinline__ int wave_min(int v)
{
    for (int o = 32; o; o >>= 1) {
        const int x = __shfl_xor(v, o);
        v = x < v ? x : v;
    }
    return v;
}

__device__ __forceinline__ int wave_max(int v)
{
    for (int o = 32; o; o >>= 1) {
        const int x = __shfl_xor(v, o);
        v = x > v ? x : v;
    }
    return v;
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

/* Reserve n slots of an append counter for the active lanes, one atomic per
 * wave: only the first active lane adds, and readfirstlane hands its old
 * value to every lane. (wave_take has every lane add, the others zero, for
 * work-queue loops whose exit depends on the ticket; a same-address atomic
 * from 64 lanes is 64 operations in L2. Here the result only places stores,
 * and the branch has nothing after it that depends on which lane took it.) */
__device__ __forceinline__ unsigned long long wave_reserve(unsigned long long *counter, u32 n)
{
    unsigned long long t = 0;
    if (rank_below(__builtin_amdgcn_read_exec()) == 0) t = atomicAdd(counter, (unsigned long long)n);
    const u32 lo = __builtin_amdgcn_readfirstlane((u32)t);
    const u32 hi = __builtin_amdgcn_readfirstlane((u32)(t >> 32));
    return ((u64)hi << 32) | lo;
}

/* Work queues: QUEUES counters on separate 256-byte lines (QUEUE_BYTES of
 * workspace). Same-address device atomics are served one after another
 * (about 16 ns apart here), so one counter shared by 4,096 waves is a queue
 * in itself -- at the launch, at the end (every wave's last, failed
 * ticket) and for workloads of many small tiles. Wave w draws from queue
 * w % Q only (Q = min(QUEUES, waves)), and each queue owns an interleaved
 * share of the tiles, so every queue is balanced dynamically among its waves
 * while each address sees 1/Q of the traffic. */
constexpr int QUEUE_STRIDE = 32;                        /* u64 words: 256 bytes */
constexpr int QUEUE_BYTES = QUEUES * QUEUE_STRIDE * 8;

/* Tickets in runs of `grab` tiles; run r covers tiles r * grab .. + grab - 1,
 * scheduled by queue_run (bl_device.h). With `first_static` every wave's
 * first run is its own (run w), so a launch's waves do not all queue at once
 * (the 1,024 x 1e6 search at k = 2: 0.34 -> 0.30 ms with one counter); tiles
 * of uneven cost can prefer the dynamic start. One atomic per later run. A
 * wave takes its runs in increasing order. next / end are wave-uniform. */
struct Tickets {
    long long next = 0, end = 0;
    bool started = false;
    bool first_static = true;
    __device__ __forceinline__ long long get(unsigned long long *queues, int grab)
    {
        if (next >= end) {
            const long long waves = ((long long)gridDim.x * blockDim.x) >> 6;
            const long long wid = uni64(((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
            long long run;
            if (first_static && !started) {
                started = true;
                run = wid;
            } else {
                const int j = queue_of(wid, waves);
                run = queue_run(wid, waves, first_static, wave_take(queues + j * QUEUE_STRIDE, 1));
            }
            next = run * grab;
            end = next + grab;
        }
        return next++;
    }
};

__global__ void clear_kernel(unsigned long long *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

/* ======================================================================= cdist */
/* Where search results go: the dense row-major q x nc matrix in one of the
 * output dtypes (the reference's result), or a match list of (q, c, d) for
 * every d <= k, appended in no particular order. put() receives the result
 * min(d, k + 1) of output index q * nc + c (or -1 for a caller error). */
struct DenseSink {
    void *out;
    int dtype;
    static constexpr bool kDense = true;
    __device__ __forceinline__ void put(int64_t idx, int d, int) const { store_out(out, idx, d, dtype); }
};

__device__ __forceinline__ int dense_dtype(const DenseSink &s) { return s.dtype; }
__device__ __forceinline__ void *dense_out(const DenseSink &s) { return s.out; }

struct MatchSink {
    unsigned long long *count;  /* matches found (may exceed cap: the caller retries) */
    int32_t *mq;
    int64_t *mc;
    int32_t *md;
    int64_t cap, nc;
    static constexpr bool kDense = false;
    /* The active lanes' matches take one run of the list with one atomic
     * (the first active lane adds their number), each lane its rank in it. */
    __device__ __forceinline__ void put(int64_t idx, int d, int k) const
    {
        const bool hit = d >= 0 && d <= k;
        const u64 m = wave_ballot(hit);
        if (m == 0) return;
        const unsigned long long base = wave_reserve(count, (u32)__popcll(m));
        if (hit) {
            const unsigned long long pos = base + (unsigned long long)rank_below(m);
            if (pos < (unsigned long long)cap) {
                const int64_t q = idx / nc;
                mq[pos] = (int32_t)q;
                mc[pos] = idx - q * nc;
                md[pos] = d;
            }
        }
    }
};
__device__ __forceinline__ int dense_dtype(const MatchSink &) { return DT_INT32; }
__device__ __forceinline__ void *dense_out(const MatchSink &) { return nullptr; }

constexpr int CD_SCAP = 128;   /* pending-choice stack entries per wave (>= 64 + 64 - 1) */
constexpr int CD_ECAP = 128;   /* escalation stack entries per wave (>= 64 + 64 - 1) */
constexpr int CD_DIRECT = 32;  /* in-band choices a window needs to be scanned in place */
constexpr int CD_PARK = 16;    /* a pass parks its stragglers once at most this many remain */

/* A choice waiting for a pass of the 32-row stage, with its scan state after
 * j columns (j = 0: not started). */
struct Pend {
    int64_t start;             /* dword 0 of its stripe string */
    int32_t rel, n, j;
    u32 Pv, Mv, pad;
};

constexpr int CD_OCAP = 320;   /* open stack entries per wave (>= 64 + 256 - 1) */

/* One wave's slice of LDS. */
struct WaveLds {
    u32 lo32[256];             /* rows 0..31 of the current query's Peq table */
    u8 lo8[256];               /* rows 0..7 of it (the four-choice filter's table) */
    union {
        Pend pend[CD_SCAP];    /* compacted mode: choices waiting for a pass */
        int64_t open[CD_OCAP]; /* filter mode: (query, choice) pairs to rescan, as the
                                  output index q * nc + c, across tiles */
    };
    int32_t esc[CD_ECAP];      /* choices the 32-row stage cannot decide */
};

/* Peq tables of every query in global memory: row c of query q holds nwmax
 * words at peq[(q * 256 + c) * nwmax]; word w bit i = (query[64w + i] == c).
 * One workgroup per query, one thread per byte value. A query longer than
 * 64 * nwmax (a caller error, answered -1 by the search kernel) gets zeros. */
__global__ __launch_bounds__(256) void peq_build_kernel(const u8 *__restrict__ bq,
                                                        const int64_t *__restrict__ oq, int nwmax,
                                                        u64 *__restrict__ peq, Profile *__restrict__ qcnt,
                                                        Profile *__restrict__ qcnt2,
                                                        Profile3 *__restrict__ qcnt3,
                                                        unsigned long long *clr, int nclr,
                                                        unsigned long long *clr1,
                                                        unsigned long long *clr2)
{
    /* block 0 also zeroes the search's counters (the work queues' nclr and
     * up to two single ones), saving their own launches */
    if (blockIdx.x == 0) {
        for (int i = threadIdx.x; i < nclr; i += 256) clr[i] = 0;
        if (threadIdx.x == 0 && clr1) *clr1 = 0;
        if (threadIdx.x == 0 && clr2) *clr2 = 0;
    }
    __shared__ u8 qs[512];
    __shared__ u32 cls[16];
    __shared__ u32 cls2[16];
    __shared__ u32 cls3[64];
    const int64_t q = blockIdx.x;
    const int64_t q0 = oq[q];
    int m = (int)(oq[q + 1] - q0);
    if (m > 64 * nwmax) m = 0;
    for (int i = threadIdx.x; i < m; i += 256) qs[i] = bq[q0 + i];
    if (threadIdx.x < 16) {
        cls[threadIdx.x] = 0;
        cls2[threadIdx.x] = 0;
    }
    if (threadIdx.x < 64) cls3[threadIdx.x] = 0;
    __syncthreads();
    const u32 c = threadIdx.x;
    u64 *row = peq + ((size_t)q * 256 + c) * nwmax;
    u32 cnt = 0;
    for (int w = 0; w < nwmax; w++) {
        u64 word = 0;
        const int e = m < 64 * w + 64 ? m : 64 * w + 64;
        for (int i = 64 * w; i < e; i++) word |= (u64)(qs[i] == c) << (i - 64 * w);
        row[w] = word;
        cnt += (u32)popc64(word);
    }
    if (!qcnt) return;
    /* the query's byte-class profile (the class-count bound's query side):
     * the occurrences of byte c are the set bits of its Peq row */
    atomicAdd(&cls[c & 15u], cnt);
    if (qcnt2)                               /* the 2-gram profile, from the bytes */
        for (int i = threadIdx.x; i + 1 < m; i += 256)
            atomicAdd(&cls2[4u * class2(qs[i]) + class2(qs[i + 1])], 1u);
    if (qcnt3)                               /* and the 3-gram profile */
        for (int i = threadIdx.x; i + 2 < m; i += 256) atomicAdd(&cls3[class3(qs[i], qs[i + 1], qs[i + 2])], 1u);
    __syncthreads();
    if (qcnt3 && threadIdx.x == 0) qcnt3[q] = pack_profile3([&](int x) { return cls3[x]; });
    if (threadIdx.x < 4) {
        u32 v = 0, v2 = 0;
        for (int i = 0; i < 4; i++) {
            const u32 x = cls[4 * threadIdx.x + i], y = cls2[4 * threadIdx.x + i];
            v |= (x > 255u ? 255u : x) << (8 * i);
            v2 |= (y > 255u ? 255u : y) << (8 * i);
        }
        ((u32 *)(qcnt + q))[threadIdx.x] = v;
        if (qcnt2) ((u32 *)(qcnt2 + q))[threadIdx.x] = v2;
    }
}

/* Compacted mode, one (query, choice block) tile for one wave (c0 a multiple
 * of 64). The 32-row stage runs in lockstep passes of up to 64 choices, so the
 * lanes of a pass start together and read the same stripe dwords (coalesced),
 * and a pass ends as soon as few lanes still work -- their choices are parked
 * with their state and later continue in a pass of 64 parked choices:
 *   * a window is the next 64 choices of the tile (one stripe group): lengths
 *     and the first two text dwords come in with coalesced loads; the length
 *     gap (|n - m| > k) and empty strings are answered at once;
 *   * a window with at least CD_DIRECT in-band choices is scanned in place,
 *     a sparser one is parked whole;
 *   * a pass steps its lanes 4 columns at a time until at most CD_PARK are
 *     left (none once the tile's windows are used up), then parks those;
 *   * a choice whose next test would need rows past 32 goes on the escalation
 *     stack, drained 64 at a time through the wider stages (peq_lane from the
 *     64-row stage, Peq rows from the query's global table).
 * Needs m <= 32 or k >= 0 (otherwise every choice would escalate: the caller
 * runs strided). */
template <int NW, typename Sink>
__device__ void cdist_tile(WaveLds &L, const u64 *__restrict__ gpeq, int m,
                           const u32 *__restrict__ words, const int64_t *__restrict__ gbase,
                           const int32_t *__restrict__ clen, int64_t c0, int64_t c1,
                           const Sink &sink, int64_t orow, int k)
{
    const int lane = lane_id();
    const bool whole = m <= 32;
    int64_t w = c0;
    int ns = 0, ne = 0;
    for (;;) {
        /* ---- escalations: a full wave of them, or the rest at the very end */
        const bool tile_done = w >= c1 && ns == 0;
        if (ne >= 64 || (tile_done && ne > 0)) {
            const int cnt = ne < 64 ? ne : 64;
            const int base = ne - cnt;
            if (lane < cnt) {
                const int64_t c = c0 + L.esc[base + lane];
                const StripeText text{words + gbase[c >> 6] + (c & 63)};
                sink.put(orow + c, peq_lane<NW, 1>((const u32 *)nullptr, gpeq, m, text, clen[c], k), k);
            }
            ne = uni(base);
            wave_lds_sync();
            continue;
        }
        if (tile_done) break;
        /* ---- the pass's lanes: the next window, or up to 64 parked choices */
        bool act;
        int rel, n, j;
        const u32 *tp;
        u32 Pv, Mv, c4, c4n;
        if (ns < 64 && w < c1) {
            const int64_t c = w + lane;
            const bool valid = c < c1;
            const u32 *gp = words + gbase[w >> 6];          /* one group: a scalar load */
            n = valid ? clen[c] : 0;
            tp = gp + lane;
            rel = (int)(c - c0);
            j = 0;
            Pv = ~0u;
            Mv = 0u;
            w = uni64(w + 64);
            const int gap = n > m ? n - m : m - n;
            const bool decided = valid && ((k >= 0 && gap > k) || n == 0);
            if (decided) sink.put(orow + c, (k >= 0 && gap > k) ? k + 1 : m, k);
            /* the 32-row stage cannot test even its first group: escalate now */
            const bool direct = valid && !decided && !whole && 4 + m - n > 32;
            act = valid && !decided && !direct;
            const u64 dm = wave_ballot(direct);
            if (dm) {
                if (direct) L.esc[ne + rank_below(dm)] = rel;
                ne = uni(ne + __popcll(dm));
            }
            const u64 sm = wave_ballot(act);
            if (__popcll(sm) < CD_DIRECT) {
                /* sparse window: park its in-band choices */
                if (act) {
                    Pend e;
                    e.start = tp - words;
                    e.rel = rel;
                    e.n = n;
                    e.j = 0;
                    e.Pv = ~0u;
                    e.Mv = 0u;
                    e.pad = 0;
                    L.pend[ns + rank_below(sm)] = e;
                }
                ns = uni(ns + __popcll(sm));
                wave_lds_sync();
                continue;
            }
            c4 = act ? tp[0] : 0u;
            c4n = act && n > 4 ? tp[64] : 0u;
        } else {
            const int take = ns < 64 ? ns : 64;
            act = lane < take;
            rel = 0;
            n = 1;
            j = 0;
            Pv = ~0u;
            Mv = 0u;
            tp = words;
            if (act) {
                const Pend e = L.pend[ns - take + lane];
                tp = words + e.start;
                rel = e.rel;
                n = e.n;
                j = e.j;
                Pv = e.Pv;
                Mv = e.Mv;
            }
            ns = uni(ns - take);
            wave_lds_sync();                 /* entries read before their slots are reused */
            const int dj = j >> 2, last = (n - 1) >> 2;
            c4 = act ? tp[64 * dj] : 0u;
            c4n = act && dj < last ? tp[64 * (dj + 1)] : 0u;
        }
        /* ---- lockstep groups of 4 columns until few lanes are left */
        const int park = w < c1 ? CD_PARK : 0;
        bool escalate = false;
        for (;;) {
            if (act) {
                const int cnt = n - j < 4 ? n - j : 4;
                const u32 cur = c4;
                c4 = c4n;
                /* the dword after next, for the following group (clamped to the string) */
                const int nd = (j >> 2) + 2, last = (n - 1) >> 2;
                c4n = nd <= last ? tp[64 * nd] : 0u;
                u32 E[4];
#pragma unroll
                for (int u = 0; u < 4; u++) E[u] = L.lo32[(cur >> (8 * u)) & 0xFFu];
                u32 P1[1] = {Pv}, M1[1] = {Mv};
                if (cnt == 4) {
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const u32 e1[1] = {E[u]};
                        column<u32, 1>(P1, M1, 1, e1);
                    }
                } else {
#pragma unroll
                    for (int u = 0; u < 3; u++) {
                        if (u < cnt) {
                            const u32 e1[1] = {E[u]};
                            column<u32, 1>(P1, M1, 1, e1);
                        }
                    }
                }
                Pv = P1[0];
                Mv = M1[0];
                j += cnt;
                /* whole: exact at j = n, bound tests with k >= 0; otherwise (k >= 0)
                 * the bound row j + m - n is within the 32 rows by construction */
                if (!whole || k >= 0 || j >= n) {
                    const int bound = diag_bound<u32, 1>(P1, M1, 1, j, m, n);
                    const bool over = k >= 0 && bound > k;
                    if (over || (whole && j >= n)) {
                        sink.put(orow + c0 + rel, over ? k + 1 : bound, k);
                        act = false;
                    } else if (!whole && j + 4 + m - n > 32) {
                        escalate = true;
                        act = false;
                    }
                }
            }
            const int left = __popcll(wave_ballot(act));
            if (left <= park) break;
        }
        const u64 em = wave_ballot(escalate);
        if (em) {
            if (escalate) L.esc[ne + rank_below(em)] = rel;
            ne = uni(ne + __popcll(em));
        }
        const u64 am = wave_ballot(act);
        if (am) {
            if (act) {
                Pend e;
                e.start = tp - words;
                e.rel = rel;
                e.n = n;
                e.j = j;
                e.Pv = Pv;
                e.Mv = Mv;
                e.pad = 0;
                L.pend[ns + rank_below(am)] = e;
            }
            ns = uni(ns + __popcll(am));
        }
        wave_lds_sync();
    }
}

/* Strided mode: lane l takes choices c0 + l, c0 + l + 64, ... and runs each
 * through all its stages (the plain kernel's lane loop inside the persistent
 * tiles). Cheaper than the ring when every choice needs about the same work,
 * and the only mode for exact (k < 0) queries above 32 bytes. */
template <int NW, typename Sink>
__device__ void cdist_tile_strided(const u32 *lo32, const u64 *__restrict__ gpeq, int m,
                                   const u32 *__restrict__ words, const int64_t *__restrict__ gbase,
                                   const int32_t *__restrict__ clen, int64_t c0, int64_t c1,
                                   const Sink &sink, int64_t orow, int k)
{
    for (int64_t c = c0 + lane_id(); c < c1; c += 64) {
        const StripeText text{words + gbase[c >> 6] + (c & 63)};
        sink.put(orow + c, peq_lane<NW, 0>(lo32, gpeq, m, text, clen[c], k), k);
    }
}

/* k + 1 as stored in an output element of type OutT: saturated like every
 * other result (bitlev/gpu.py:64-68), so k = 300 stores 255 in uint8. */
template <typename OutT>
__device__ __forceinline__ u32 out_k1(int k)
{
    const u32 v = (u32)k + 1u;
    if (sizeof(OutT) == 1) return v > 255u ? 255u : v;
    if (sizeof(OutT) == 2) return v > 65535u ? 65535u : v;
    return v;
}

/* SWAR mode, for small thresholds (k <= 3) and queries of at least 8 bytes:
 * each lane carries four consecutive choices of a 256-choice window (w + 4l ..
 * w + 4l + 3, whose first text dwords are adjacent in the stripes, so a lane
 * reads them with one 16-byte load) in the four bytes of one u32, each byte
 * holding pattern rows 0..7 of its
 * choice. Four text columns of the Myers step run on all four at once with
 * byte-segmented arithmetic (carries and shifts kept inside each byte; row 0
 * of every segment gets the +1 of a global alignment), then the final-diagonal
 * bound is tested at column 4: an in-band choice has |n - m| <= k <= 3, so its
 * diagonal row 4 + m - n lies in 1..7, within the segment, and the low rows
 * evolve exactly as in the whole scan. Choices the bound decides get k + 1;
 * the rest (a few percent of unrelated strings) are stacked and rescanned 64
 * at a time by the staged lane code from column 0. */
/* One 256-choice window of the filter. FULL: all 256 choices lie in the tile
 * (no per-lane bounds checks). DENSE: results are stored through `ow` (the
 * output row at the window's first choice) as OutT. The choices the filter leaves
 * open are pushed on the wave's open stack as their output index orow + c
 * (the caller keeps fewer than 64 there between windows, so it never holds
 * more than 319). */
template <bool FULL, bool FILTER, bool DENSE, typename OutT>
__device__ __forceinline__ void swar_window(WaveLds &L, int m, const u32 *__restrict__ words,
                                            const int64_t *__restrict__ gbase,
                                            const int32_t *__restrict__ lw,
                                            const Profile *__restrict__ cw, const Profile &qc, OutT *ow,
                                            int64_t w, int64_t c1, int k, int64_t orow, int &ns)
{
    const int lane = lane_id();
    const u32 kk1 = out_k1<OutT>(k);
    /* lane l takes choices w + 4l .. w + 4l + 3: one stripe group ((w >> 6) +
     * l / 16, w a multiple of 256), whose first dwords sit side by side, so
     * the text, the lengths and (in full windows) the results move as one
     * 16-byte (results: 4-element) access per lane */
    const int64_t cl = w + 4 * lane;
    u32 d[4] = {0u, 0u, 0u, 0u};
    int n[4];
    bool valid[4];
    if (FULL) {
        const int4 l4 = *(const int4 *)(lw + 4 * lane);
        n[0] = l4.x; n[1] = l4.y; n[2] = l4.z; n[3] = l4.w;
#pragma unroll
        for (int s = 0; s < 4; s++) valid[s] = true;
    } else {
#pragma unroll
        for (int s = 0; s < 4; s++) {
            valid[s] = cl + s < c1;
            n[s] = valid[s] ? lw[4 * lane + s] : 0;
        }
    }
    /* out of band by the length gap or by the class-count bound (cw: the
     * window's choice profiles, qc the query's): the answer is k + 1 */
    bool decided[4];
    bool need = false;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const int gap = n[s] > m ? n[s] - m : m - n[s];
        decided[s] = gap > k;
        if (cw && !decided[s] && valid[s]) decided[s] = count_over(qc, cw[4 * lane + s], m, n[s], k);
        need = need || (valid[s] && !decided[s]);
    }
    if (FILTER && wave_ballot(need) != 0) {
        if (FULL) {
            const uint4 t = *(const uint4 *)(words + gbase[(w >> 6) + (lane >> 4)] + 4 * (lane & 15));
            d[0] = t.x; d[1] = t.y; d[2] = t.z; d[3] = t.w;
        } else {
            const u32 *gp = cl < c1 ? words + gbase[(w >> 6) + (lane >> 4)] + 4 * (lane & 15) : words;
#pragma unroll
            for (int s = 0; s < 4; s++) d[s] = valid[s] ? gp[s] : 0u;
        }
        u32 Pv, Mv;
        swar_steps4(L.lo8, d, Pv, Mv, 1);
        const u32 over = swar_over4(Pv, Mv, m, n, k);
        /* the bound at column 4 exceeds k (an empty choice has gap m >= 8 > k;
         * without the filter one with m <= k stays open and its scan returns m) */
#pragma unroll
        for (int s = 0; s < 4; s++) decided[s] = decided[s] || ((over >> (8 * s + 7)) & 1u);
    }
    bool open[4];
    bool any = false;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        open[s] = valid[s] && !decided[s];
        any = any || open[s];
        /* (k + 1 is never a match: a match list gets nothing) */
        if (DENSE && !FULL && valid[s] && decided[s]) ow[4 * lane + s] = (OutT)kk1;
    }
    if (DENSE && FULL) {
        /* all four get k + 1; an open one is overwritten by its rescan later
         * in this wave's program order */
        OutT *p = ow + 4 * lane;
        if (((uintptr_t)ow & (4 * sizeof(OutT) - 1)) == 0) {
            if (sizeof(OutT) == 1) *(u32 *)p = kk1 * 0x01010101u;
            else if (sizeof(OutT) == 2) *(uint2 *)p = make_uint2(kk1 * 0x00010001u, kk1 * 0x00010001u);
            else *(uint4 *)p = make_uint4(kk1, kk1, kk1, kk1);
        } else {
#pragma unroll
            for (int s = 0; s < 4; s++) p[s] = (OutT)kk1;
        }
    }
    if (wave_ballot(any) == 0) return;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const u64 sm = wave_ballot(open[s]);
        if (open[s]) L.open[ns + rank_below(sm)] = orow + cl + s;
        ns = uni(ns + __popcll(sm));
    }
    wave_lds_sync();
}

/* Gap mode (k >= 0): 256-choice windows settle the choices outside the band
 * (one length load and one result store per four choices) and stack the rest
 * (as output indices on the open stack); lanes scan the stacked choices'
 * 32-row stage in lockstep 4-column groups, and after every group the lanes
 * whose choice is decided take the next one from the stack, so lanes never
 * wait for a slow neighbour and no scan state is ever parked. Choices the
 * 32-row stage cannot decide go to the escalation stack and through the
 * 64-row and whole stages 64 at a time. */
template <int NW, typename Sink, typename Window>
__device__ void cdist_tile_refill(WaveLds &L, const u64 *__restrict__ gpeq, int m,
                                  const u32 *__restrict__ words, const int64_t *__restrict__ gbase,
                                  const int32_t *__restrict__ clen, int64_t c0, int64_t c1,
                                  const Sink &sink, int64_t orow, int k, Window window, int &ns)
{
    const int lane = lane_id();
    const bool whole = m <= 32;
    int64_t w = c0;
    int ne = 0;
    bool act = false;
    int64_t c = 0;
    int n = 0, j = 0;
    u32 Pv = ~0u, Mv = 0u, c4 = 0u, c4n = 0u;
    const u32 *tp = words;
    for (;;) {
        /* ---- feed the stack with windows while it is short */
        while (ns < 64 && w < c1) {
            window(w);
            w += 256;
        }
        /* ---- free lanes take stacked choices */
        const u64 fr = wave_ballot(!act);
        const int take = __popcll(fr) < ns ? __popcll(fr) : ns;
        if (take > 0) {
            const int r = rank_below(fr);
            if (!act && r < take) {
                c = L.open[ns - take + r] - orow;
                n = clen[c];
                j = 0;
                Pv = ~0u;
                Mv = 0u;
                tp = words + gbase[c >> 6] + (c & 63);
                act = true;
                if (n == 0) {                    /* (in band only when m <= k) */
                    sink.put(orow + c, m, k);
                    act = false;
                } else {
                    c4 = tp[0];
                    c4n = n > 4 ? tp[64] : 0u;
                }
            }
            ns = uni(ns - take);
            wave_lds_sync();                     /* entries read before their slots are reused */
        }
        /* ---- escalations: a full wave of them, or the rest at the very end */
        const u64 am0 = wave_ballot(act);
        const bool done = am0 == 0 && ns == 0 && w >= c1;
        if (ne >= 64 || (done && ne > 0)) {
            const int cnt = ne < 64 ? ne : 64;
            const int base = ne - cnt;
            if (lane < cnt) {
                const int64_t ce = c0 + L.esc[base + lane];
                const StripeText text{words + gbase[ce >> 6] + (ce & 63)};
                sink.put(orow + ce, peq_lane<NW, 1>((const u32 *)nullptr, gpeq, m, text, clen[ce], k), k);
            }
            ne = uni(base);
            wave_lds_sync();
            continue;
        }
        if (done) break;
        if (am0 == 0) continue;
        /* ---- one 4-column group of the 32-row stage */
        bool escalate = false;
        if (act) {
            const int cnt = n - j < 4 ? n - j : 4;
            const u32 cur = c4;
            c4 = c4n;
            const int nd = (j >> 2) + 2, last = (n - 1) >> 2;
            c4n = nd <= last ? tp[64 * nd] : 0u;
            u32 E[4];
#pragma unroll
            for (int u = 0; u < 4; u++) E[u] = L.lo32[(cur >> (8 * u)) & 0xFFu];
            u32 P1[1] = {Pv}, M1[1] = {Mv};
            if (cnt == 4) {
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const u32 e1[1] = {E[u]};
                    column<u32, 1>(P1, M1, 1, e1);
                }
            } else {
#pragma unroll
                for (int u = 0; u < 3; u++) {
                    if (u < cnt) {
                        const u32 e1[1] = {E[u]};
                        column<u32, 1>(P1, M1, 1, e1);
                    }
                }
            }
            Pv = P1[0];
            Mv = M1[0];
            j += cnt;
            /* the bound row j + m - n is within the 32 rows while the stage runs */
            const int bound = diag_bound<u32, 1>(P1, M1, 1, j, m, n);
            if (bound > k || (whole && j >= n)) {
                sink.put(orow + c, bound > k ? k + 1 : bound, k);
                act = false;
            } else if (!whole && j + 4 + m - n > 32) {
                escalate = true;
                act = false;
            }
        }
        const u64 em = wave_ballot(escalate);
        if (em) {
            if (escalate) L.esc[ne + rank_below(em)] = (int32_t)(c - c0);
            ne = uni(ne + __popcll(em));
            wave_lds_sync();
        }
    }
}

/* Count-mode rescans read a query's 32-row table from its global Peq table
 * (lane by lane, each lane's own query) instead of the wave's LDS copy. */
template <int NWMAX>
struct GlobalLo32 {
    const u64 *peq;
    __device__ __forceinline__ u32 operator[](u32 c) const { return (u32)peq[c * NWMAX]; }
};

/* One wave's LDS in count mode: the open stack, the escalation stack, and
 * the 32-row tables of up to GL queries of the current group (GL = 0: none). */
template <int GL>
struct CountLds {
    int64_t open[CD_OCAP];
    int64_t scan[CD_ECAP];      /* open entries the 2-gram bound leaves for a scan */
    int64_t esc[CD_ECAP];       /* entries the 32-row stage left undecided */
    u32 lo32[GL > 0 ? GL : 1][256];
};
constexpr int CD_GL = 4;        /* query groups up to this size keep their tables in LDS */

/* One wave's LDS for the matrix-core tiles: the group's int8 operands (two
 * 32-query row blocks) and one block's accumulator when it has open pairs,
 * so that nothing of the tile path stays in registers while rescans run. */
struct GramLds {
    v4i a[2][64];
    unsigned short m[16][64];   /* per 32 x 32 block of a window: lane l's open rows, bit r */
};

/* Persistent cdist over queries of at most 64*NWMAX bytes whose Peq tables
 * peq_build_kernel wrote to gpeq. Waves pull (choice block, query) tiles,
 * block-major, from the work-queue counter. */
template <int NWMAX, typename Sink>
__global__ __launch_bounds__(256) void cdist_batched_kernel(
    const int64_t *__restrict__ oq, int64_t nq, const u64 *__restrict__ gpeq,
    const u32 *__restrict__ words, const int64_t *__restrict__ gbase,
    const int32_t *__restrict__ clen, const Profile *__restrict__ ccnt,
    const Profile *__restrict__ qcnt, int64_t nc, Sink sink,
    int k_all, const int32_t *__restrict__ kvec, int64_t tile, int mode, int grab,
    unsigned long long *counter)
{
    extern __shared__ u64 lds[];
    WaveLds &L = ((WaveLds *)lds)[threadIdx.x >> 6];
    const int lane = lane_id();
    const int64_t nblk = (nc + tile - 1) / tile;
    const int64_t ntiles = nblk * nq;
    int64_t cur_q = -1;
    /* filter mode: choices the filter left open, as output indices, rescanned
     * 64 at a time with the current query's tables (the stack is emptied
     * before the wave moves to another tile) */
    int ns = 0;
    const u64 *qpeq = nullptr;
    int m = 0, k = 0;
    Profile qc{0u, 0u, 0u, 0u};   /* the current query's class profile */
    auto drain = [&](int cnt) {
        const int base = ns - cnt;
        if (lane < cnt) {
            const int64_t idx = L.open[base + lane];
            const int64_t c = idx % nc;
            const StripeText text{words + gbase[c >> 6] + (c & 63)};
            sink.put(idx, peq_lane<NWMAX, 0>(L.lo32, qpeq, m, text, clen[c], k), k);
        }
        ns = uni(base);
        wave_lds_sync();
    };
    Tickets tk;
    for (;;) {
        const long long t = tk.get(counter, grab);
        if (t >= ntiles) break;
        const int64_t blk = t / nq, q = t % nq;
        const int64_t c0 = blk * tile;
        const int64_t c1 = c0 + tile < nc ? c0 + tile : nc;
        m = (int)(oq[q + 1] - oq[q]);
        const int64_t orow = q * nc;
        k = kvec ? clamp_k_dev(kvec[q]) : k_all;
        if (m == 0 || m > 64 * NWMAX) {
            /* empty query: n, or k + 1 (csrc/gpu_kernels.cu:53-54); too long for this
             * launch: caller error, marked -1 */
            for (int64_t c = c0 + lane; c < c1; c += 64)
                sink.put(orow + c, m == 0 ? empty_pattern(clen[c], k) : -1, k);
            continue;
        }
        qpeq = gpeq + (size_t)q * 256 * NWMAX;
        if (q != cur_q) {
            wave_lds_sync();
#pragma unroll
            for (int x = 0; x < 4; x++) {
                const u32 v = (u32)qpeq[(x * 64 + lane) * NWMAX];
                L.lo32[x * 64 + lane] = v;
                L.lo8[x * 64 + lane] = (u8)v;
            }
            if (ccnt) qc = qcnt[q];
            wave_lds_sync();
            cur_q = q;
        }
        /* (the filter's rescan of 8-word queries would cost the whole launch its
         * occupancy: queries above 256 bytes use the other modes) */
        const bool filt = NWMAX <= 4 && mode == 2 && k >= 0 && k <= 3 && m >= 8;
        const bool gapw = mode == 3 && k >= 0;
        if ((filt || gapw) && (c0 & 255) == 0) {
            /* 256-choice windows through the four-choice filter (or the gap
             * test alone); the open stack is emptied to below 64 between windows */
            int64_t w = c0;
            auto window = [&](auto full, auto filter, bool drain_open = true) {
                constexpr bool F = decltype(full)::value;
                constexpr bool FI = decltype(filter)::value;
                int64_t w0 = w;
                const Profile *cw = ccnt ? ccnt + w0 : nullptr;
                if (!Sink::kDense)
                    swar_window<F, FI, false>(L, m, words, gbase, clen + w0, cw, qc, (u8 *)nullptr,
                                              w0, c1, k, orow, ns);
                else if (dense_dtype(sink) == DT_UINT8)
                    swar_window<F, FI, true>(L, m, words, gbase, clen + w0, cw, qc,
                                             (u8 *)dense_out(sink) + orow + w0, w0, c1, k, orow, ns);
                else if (dense_dtype(sink) == DT_UINT16)
                    swar_window<F, FI, true>(L, m, words, gbase, clen + w0, cw, qc,
                                             (uint16_t *)dense_out(sink) + orow + w0, w0, c1, k, orow,
                                             ns);
                else
                    swar_window<F, FI, true>(L, m, words, gbase, clen + w0, cw, qc,
                                             (int32_t *)dense_out(sink) + orow + w0, w0, c1, k, orow,
                                             ns);
                if (drain_open)
                    while (ns >= 64) drain(64);
            };
            if (filt) {
                for (; w + 256 <= c1; w += 256) window(std::true_type(), std::true_type());
                if (w < c1) window(std::false_type(), std::true_type());
                while (ns > 0) drain(ns < 64 ? ns : 64);
            } else {
                /* gap windows feed lanes that refill after every group */
                auto gap_window = [&](int64_t w0) {
                    w = w0;
                    if (w0 + 256 <= c1) window(std::true_type(), std::false_type(), false);
                    else window(std::false_type(), std::false_type(), false);
                };
                cdist_tile_refill<NWMAX>(L, qpeq, m, words, gbase, clen, c0, c1, sink, orow, k,
                                         gap_window, ns);
            }
        }
        else {
            if (mode == 1 || (m > 32 && k < 0))
                cdist_tile_strided<NWMAX>(L.lo32, qpeq, m, words, gbase, clen, c0, c1, sink, orow, k);
            else
                cdist_tile<NWMAX>(L, qpeq, m, words, gbase, clen, c0, c1, sink, orow, k);
        }
    }
}

/* Exact searches (lane mode 5; any threshold, lane by lane): one workgroup
 * per (query, block of `tile` slots), the grid walking the queries of a block
 * before the next block so that the block's stripes serve every query from
 * L2 (the query index rotated by the block, so every query visits every
 * XCD). The query's whole Peq table is copied into LDS word-major (word w of
 * byte c at T[256 w + c]: one 8-byte bank pair per byte value whatever the
 * word count) from the table peq_build_kernel wrote; each wave takes 64 slots
 * at a time and every lane scans its choice in the query's 32-row units
 * (units_lane: one u32 block, u64 blocks, or u64 blocks and a u32 tail), so a
 * 72-byte query costs 3 units, not 4. With `order` (stripes in length order)
 * slot c holds choice order[c]: the lanes of a wave scan strings of about the
 * same length and finish together, and each result goes to the caller's
 * column order[c]. */
constexpr int PEQ_STRIDE = 257;              /* u64 entries per word of an LDS Peq table */

template <int NWMAX, typename Sink>
__global__ __launch_bounds__(256) void cdist_exact_kernel(
    const int64_t *__restrict__ oq, int64_t nq, const u64 *__restrict__ gpeq,
    const u32 *__restrict__ words, const int64_t *__restrict__ gbase,
    const int32_t *__restrict__ clen, const int32_t *__restrict__ order, int64_t nc, Sink sink,
    int k_all, const int32_t *__restrict__ kvec, int64_t tile)
{
    /* word-major, word w of byte c at T[w * PEQ_STRIDE + c]: one u64 read per
     * word, lanes with bytes less than 32 apart on distinct banks; the stride
     * of 257 keeps the compiler from pairing two words of a byte in one
     * ds_read2 (their entries then sat on the same banks: 0.65 bank-conflict
     * cycles per LDS cycle on the serving-shaped search) */
    __shared__ u64 T[NWMAX * PEQ_STRIDE];
    /* workgroups go round-robin to the XCDs (b % 8), so with the query index
     * fastest query q would run on XCD q % 8 for every block and an XCD's
     * share would follow its queries' lengths; rotating by the block spreads
     * every query over all XCDs (64 queries of 16-128 bytes against 1e6
     * choices: 5.66 -> 4.00 ms) */
    const int64_t b = blockIdx.x;
    const int64_t blk = b / nq;
    const int64_t q = (b - blk * nq + blk) % nq;
    const int64_t c0 = blk * tile;
    const int64_t c1 = c0 + tile < nc ? c0 + tile : nc;
    const int m = (int)(oq[q + 1] - oq[q]);
    const int k = kvec ? clamp_k_dev(kvec[q]) : k_all;
    const int64_t orow = q * nc;
    if (m == 0 || m > 64 * NWMAX) {
        /* empty query: n, or k + 1 (csrc/gpu_kernels.cu:53-54); too long for this
         * launch: caller error, marked -1 */
        for (int64_t c = c0 + threadIdx.x; c < c1; c += 256)
            sink.put(orow + (order ? (int64_t)order[c] : c), m == 0 ? empty_pattern(clen[c], k) : -1, k);
        return;
    }
    const u64 *qp = gpeq + (size_t)q * 256 * NWMAX;
    for (int i = threadIdx.x; i < 256 * NWMAX; i += 256) T[(i % NWMAX) * PEQ_STRIDE + i / NWMAX] = qp[i];
    __syncthreads();
    const int lane = lane_id();
    const int64_t first = c0 + 64 * (threadIdx.x >> 6);
    auto run = [&](auto uc) {
        constexpr int U = decltype(uc)::value;
        auto eq64 = [&](u32 ch, int w) { return T[w * PEQ_STRIDE + ch]; };
        auto eq32 = [&](u32 ch, int u) { return ((const u32 *)T)[2 * ((u >> 1) * PEQ_STRIDE + ch) + (u & 1)]; };
        for (int64_t s = first; s < c1; s += 256) {
            const int64_t c = s + lane;
            if (c < c1) {
                const StripeText text{words + gbase[c >> 6] + (c & 63)};
                const int d = units_lane<U>(m, text, clen[c], k, eq64, eq32);
                sink.put(orow + (order ? (int64_t)order[c] : c), d, k);
            }
        }
    };
    using std::integral_constant;
    switch ((m + 31) >> 5) {
    case 1: run(integral_constant<int, 1>()); break;
    case 2: run(integral_constant<int, 2>()); break;
    case 3: if constexpr (NWMAX >= 2) run(integral_constant<int, 3>()); break;
    case 4: if constexpr (NWMAX >= 2) run(integral_constant<int, 4>()); break;
    case 5: if constexpr (NWMAX >= 4) run(integral_constant<int, 5>()); break;
    case 6: if constexpr (NWMAX >= 4) run(integral_constant<int, 6>()); break;
    case 7: if constexpr (NWMAX >= 4) run(integral_constant<int, 7>()); break;
    case 8: if constexpr (NWMAX >= 4) run(integral_constant<int, 8>()); break;
    case 9: if constexpr (NWMAX >= 8) run(integral_constant<int, 9>()); break;
    case 10: if constexpr (NWMAX >= 8) run(integral_constant<int, 10>()); break;
    case 11: if constexpr (NWMAX >= 8) run(integral_constant<int, 11>()); break;
    case 12: if constexpr (NWMAX >= 8) run(integral_constant<int, 12>()); break;
    case 13: if constexpr (NWMAX >= 8) run(integral_constant<int, 13>()); break;
    case 14: if constexpr (NWMAX >= 8) run(integral_constant<int, 14>()); break;
    case 15: if constexpr (NWMAX >= 8) run(integral_constant<int, 15>()); break;
    default: if constexpr (NWMAX >= 8) run(integral_constant<int, 16>()); break;
    }
}

/* Count mode, the windows of one tile: choices c0..c1 (c0 a multiple of 256)
 * against the queries q0 + j whose bit j is set in okmask, query j's length,
 * threshold and profile in lane j of lm, lk, lp. A lane holds four
 * consecutive choices of each 256-choice window (one 16-byte length load,
 * four profile loads, one packed result store), kept in registers for the
 * whole group. A choice is settled when its length gap exceeds k or its
 * class-count bound does (d >= (L1 + gap) / 2; L1 / 2 if a length exceeds
 * 255); when every length of the window and the query's fit a byte, the gap
 * joins the v_sad_u8 chain (|n - m| of the low bytes) and one compare decides
 * (L1 >= gap, so L1 + gap > 2k already covers gap > k). Settled choices get
 * k + 1 (DENSE: the row at out + q * nc); the rest are pushed on the open
 * stack as output indices for the per-lane rescans (drain). */
template <bool DENSE, bool PAIRS8, bool GRAM, typename OutT, typename Drain>
__device__ __forceinline__ void count_windows(int64_t *open_stack, const int32_t *__restrict__ clen,
                                              const Profile *__restrict__ ccnt, int64_t c0,
                                              int64_t c1, int64_t nc, int64_t q0, int gq, u64 okmask,
                                              int lm, int lk, bool uniform_k, const Profile &lp,
                                              OutT *out, bool rows_aligned, bool rows16, int &ns,
                                              Drain &drain, GramLds *gl)
{
    const int lane = lane_id();
    const int k_uni = __builtin_amdgcn_readfirstlane(lk);
    /* the common tile: every query of the group counted, one threshold, all
     * query lengths within a byte, aligned rows -- then full windows whose
     * choices fit a byte run a loop with nothing to decide per query */
    const u64 group_mask = gq >= 64 ? ~0ull : ((1ull << gq) - 1ull);
    /* lengths enter the byte sums as (length - lbase): lbase = the group's
     * shortest query less 128 (at least 0), so strings of 200-256 bytes fit
     * a byte too; a query or a window outside [lbase, lbase + 255] takes the
     * general loop */
    const int lbase = uni(max(0, wave_min(lane < gq && lm > 0 ? lm : 0x7fffffff) - 128));
    const int lmb = lm - lbase;
    const bool fast_tile = okmask == group_mask && uniform_k &&
                           wave_ballot(lane < gq && (u32)lmb > 255u) == 0 && (!DENSE || rows_aligned);
    const u32 k2_uni = 2u * (u32)k_uni;
    const u32 fill_uni = out_k1<OutT>(k_uni) * (sizeof(OutT) == 1 ? 0x01010101u
                                                : sizeof(OutT) == 2 ? 0x00010001u : 1u);
    /* Matrix-core tiles (GRAM): with every query of the group and every choice
     * of a window at most 64 bytes (class counts <= 64, doubled counts within
     * int8) and k <= 63, the whole window is tested with the class-count L2
     * bound (count_l2_over): per 32 queries x 32 choices one
     * v_mfma_i32_32x32x32_i8 gives D = ||a - b||^2 over the 16 class counts,
     * dims 0-15 carrying -2 a against b and dims 16-19 the norms in base-128
     * digits, [-h_a, l_a, -128, 1] against [-128, 1, -h_b, l_b] (|x|^2 = 128 h
     * + l, h <= 32); a pair stays open when D <= 2 k^2. Layout (checked on the device):
     * A and B lane l hold row / column l % 32, dims 16 (l / 32) .. + 15; D lane
     * l, register r holds column l % 32, row 8 (r / 4) + 4 (l / 32) + r % 4.
     * The window's k + 1 fills are plain streaming stores. */
    const int half = lane >> 5;
    bool gram = false;
    if (GRAM && fast_tile && k_uni <= 63 && wave_ballot(lane < gq && lm > 64) == 0) {
        gram = true;
        const u32 qn = lane < gq ? profile_dot(lp, lp) : 0u;
#pragma unroll
        for (int rb = 0; rb < 2; rb++) {
            const int j = 32 * rb + (lane & 31);
            const u32 sx = (u32)__shfl((int)lp.x, j), sy = (u32)__shfl((int)lp.y, j);
            const u32 sz = (u32)__shfl((int)lp.z, j), sw = (u32)__shfl((int)lp.w, j);
            const u32 n2 = (u32)__shfl((int)qn, j);
            if (half == 0) {
                const u32 v[4] = {sx, sy, sz, sw};
                u32 o[4];
#pragma unroll
                for (int d = 0; d < 4; d++) {
                    o[d] = 0;
#pragma unroll
                    for (int b = 0; b < 32; b += 8) o[d] |= ((0u - 2u * ((v[d] >> b) & 255u)) & 255u) << b;
                }
                gl->a[rb][lane] = j < gq ? v4i{(int)o[0], (int)o[1], (int)o[2], (int)o[3]} : v4i{0, 0, 0, 0};
            } else {
                /* rows past the group: 128 * 100 > 2 * 63^2, never open */
                const u32 h = j < gq ? n2 >> 7 : 100u, l = j < gq ? n2 & 127u : 0u;
                gl->a[rb][lane] = v4i{(int)(((0u - h) & 255u) | (l << 8) | (0x80u << 16) | (1u << 24)),
                                      0, 0, 0};
            }
        }
        wave_lds_sync();
    }
    const int two_k2 = 2 * k_uni * k_uni;
    auto gram_window = [&](int64_t w) {
        if (DENSE) {
            int64_t orow = q0 * nc;
            if (rows16) {
                /* 16 bytes per lane: 4 / sizeof(OutT) rows per store */
                constexpr int LPR = 16 * (int)sizeof(OutT);      /* lanes per row segment */
                constexpr int RPS = 64 / LPR;
                const int rsub = lane / LPR;
                const v4u f4 = {fill_uni, fill_uni, fill_uni, fill_uni};
                for (int j = 0; j < gq; j += RPS) {
                    if (j + rsub < gq) {
                        v4u *p = (v4u *)__builtin_assume_aligned(
                            out + orow + (int64_t)rsub * nc + w + (lane % LPR) * (16 / sizeof(OutT)), 16);
                        __builtin_nontemporal_store(f4, p);
                    }
                    orow += RPS * nc;
                }
            } else {
                for (int j = 0; j < gq; j++, orow += nc) {
                    u32 *p = (u32 *)__builtin_assume_aligned(out + orow + w + 4 * lane, 4 * sizeof(OutT));
#pragma unroll
                    for (int x = 0; x < (int)sizeof(OutT); x++) __builtin_nontemporal_store(fill_uni, p + x);
                }
            }
        }
        /* the window's profiles all at once (eight 16-byte loads a lane in
         * flight together), the group's operands from LDS */
        Profile cpv[8];
#pragma unroll
        for (int cb = 0; cb < 8; cb++) cpv[cb] = ccnt[w + 32 * cb + (lane & 31)];
        const v4i a0 = gl->a[0][lane], a1 = gl->a[1][lane];
        const bool two = gq > 32;
        /* blocks with open pairs: each lane's open rows as a 16-bit mask in
         * LDS; pushes and rescans run after the window, once D and the
         * profiles are dead */
        u32 any = 0;
#pragma unroll
        for (int cb = 0; cb < 8; cb++) {
            const Profile &cp = cpv[cb];
            /* lanes 0-31: the counts; lanes 32-63: [-128, 1, -h_b, l_b] */
            const u32 n2 = profile_dot(cp, cp);
            const u32 dig = 0x80u | (1u << 8) | (((0u - (n2 >> 7)) & 255u) << 16) | ((n2 & 127u) << 24);
            const v4i bop = half ? v4i{(int)dig, 0, 0, 0} : v4i{(int)cp.x, (int)cp.y, (int)cp.z, (int)cp.w};
#pragma unroll
            for (int rb = 0; rb < 2; rb++) {
                if (rb == 1 && !two) continue;
                const v16i D = __builtin_amdgcn_mfma_i32_32x32x32_i8(rb ? a1 : a0, bop, v16i{0}, 0, 0, 0);
                int mn = D[0];
#pragma unroll
                for (int r = 1; r < 16; r++) mn = D[r] < mn ? D[r] : mn;
                if (wave_ballot(mn <= two_k2) == 0) continue;
                u32 bits = 0;
#pragma unroll
                for (int r = 0; r < 16; r++) bits |= (D[r] <= two_k2 ? 1u : 0u) << r;
                gl->m[2 * cb + rb][lane] = (unsigned short)bits;
                any |= 1u << (2 * cb + rb);
            }
        }
        if (any == 0) return;
        wave_lds_sync();
        while (any) {
            const int b = __builtin_ctz(any);
            any &= any - 1;
            const int cb = b >> 1, rb = b & 1;
            const int64_t c = w + 32 * cb + (lane & 31);
#pragma unroll 1
            for (int r = 0; r < 16; r++) {
                const u64 m = wave_ballot((gl->m[b][lane] >> r) & 1u);
                if (m == 0) continue;
                const int64_t q = q0 + 32 * rb + 8 * (r >> 2) + 4 * half + (r & 3);
                if ((m >> lane) & 1u) open_stack[ns + rank_below(m)] = q * nc + c;
                ns = uni(ns + __popcll(m));
                if (ns >= 64) {
                    wave_lds_sync();
                    while (ns >= 64) drain(64);
                }
            }
        }
    };
    auto window = [&](int64_t w, auto fullc) {
        constexpr bool FULL = decltype(fullc)::value;
        const int64_t cl = w + 4 * lane;
        if (GRAM && FULL && gram) {
            const int4 l4 = *(const int4 *)(clen + cl);
            if (wave_ballot(l4.x > 64 || l4.y > 64 || l4.z > 64 || l4.w > 64) == 0) {
                gram_window(w);
                return;
            }
        }
        int n[4];
        bool valid[4];
        Profile cp[4];
        if (FULL) {
            const int4 l4 = *(const int4 *)(clen + cl);
            n[0] = l4.x; n[1] = l4.y; n[2] = l4.z; n[3] = l4.w;
#pragma unroll
            for (int s = 0; s < 4; s++) {
                valid[s] = true;
                cp[s] = ccnt[cl + s];
            }
        } else {
#pragma unroll
            for (int s = 0; s < 4; s++) {
                valid[s] = cl + s < c1;
                n[s] = valid[s] ? clen[cl + s] : 0;
                cp[s] = valid[s] ? ccnt[cl + s] : Profile{0u, 0u, 0u, 0u};
            }
        }
        bool big = false;
#pragma unroll
        for (int s = 0; s < 4; s++) big = big || (valid[s] && (u32)(n[s] - lbase) > 255u);
        const bool small = wave_ballot(big) == 0;
        int64_t orow = q0 * nc;
        if (FULL && fast_tile && small) {
            for (int j = 0; j < gq; j++, orow += nc) {
                const int mqb = __builtin_amdgcn_readlane(lmb, j);
                const Profile pq{(u32)__builtin_amdgcn_readlane((int)lp.x, j),
                                 (u32)__builtin_amdgcn_readlane((int)lp.y, j),
                                 (u32)__builtin_amdgcn_readlane((int)lp.z, j),
                                 (u32)__builtin_amdgcn_readlane((int)lp.w, j)};
                u32 t[4];
                bool half = false;
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    t[s] = sad_u8(pq.x, cp[s].x, 0u);
                    t[s] = sad_u8(pq.y, cp[s].y, t[s]);
                    t[s] = sad_u8((u32)(n[s] - lbase), (u32)mqb, t[s]);
                    half = half || t[s] <= k2_uni;
                }
                if (DENSE) {
                    u32 *p = (u32 *)__builtin_assume_aligned(out + orow + w + 4 * lane, 4 * sizeof(OutT));
#pragma unroll
                    for (int x = 0; x < (int)sizeof(OutT); x++) __builtin_nontemporal_store(fill_uni, p + x);
                }
                if (wave_ballot(half) == 0) continue;
                u64 om[4];
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    u32 u = sad_u8(pq.z, cp[s].z, t[s]);
                    u = sad_u8(pq.w, cp[s].w, u);
                    om[s] = wave_ballot(u <= k2_uni);
                }
                if ((om[0] | om[1] | om[2] | om[3]) == 0) continue;
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    if ((om[s] >> lane) & 1u) open_stack[ns + rank_below(om[s])] = orow + cl + s;
                    ns = uni(ns + __popcll(om[s]));
                }
                wave_lds_sync();
                while (ns >= 64) drain(64);
            }
            return;
        }
        for (int j = 0; j < gq; j++, orow += nc) {
            if (!((okmask >> j) & 1u)) continue;
            const int mq = __builtin_amdgcn_readlane(lm, j);
            const int kq = uniform_k ? k_uni : __builtin_amdgcn_readlane(lk, j);
            const Profile pq{(u32)__builtin_amdgcn_readlane((int)lp.x, j),
                             (u32)__builtin_amdgcn_readlane((int)lp.y, j),
                             (u32)__builtin_amdgcn_readlane((int)lp.z, j),
                             (u32)__builtin_amdgcn_readlane((int)lp.w, j)};
            const u32 k2 = 2u * (u32)kq;                 /* kq <= 2^30 */
            /* the open choices as wave masks, straight from the compares */
            u64 om[4];
            if (small && (u32)(mq - lbase) <= 255u) {
                /* classes 0-7 and the gap first: a partial L1 is still a lower
                 * bound, and it settles most choices; classes 8-15 only when a
                 * lane of the wave is left open */
                u32 t[4];
                bool half = false;
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    t[s] = sad_u8(pq.x, cp[s].x, 0u);
                    t[s] = sad_u8(pq.y, cp[s].y, t[s]);
                    t[s] = sad_u8((u32)(n[s] - lbase), (u32)(mq - lbase), t[s]);
                    half = half || (valid[s] && t[s] <= k2);
                }
                if (wave_ballot(half) == 0) {
#pragma unroll
                    for (int s = 0; s < 4; s++) om[s] = 0;
                } else {
#pragma unroll
                    for (int s = 0; s < 4; s++) {
                        u32 u = sad_u8(pq.z, cp[s].z, t[s]);
                        u = sad_u8(pq.w, cp[s].w, u);
                        om[s] = wave_ballot(valid[s] && u <= k2);
                    }
                }
            } else {
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    const u32 gap = (u32)(n[s] > mq ? n[s] - mq : mq - n[s]);
                    u32 t = sad_u8(pq.x, cp[s].x, 0u);
                    t = sad_u8(pq.y, cp[s].y, t);
                    t = sad_u8(pq.z, cp[s].z, t);
                    t = sad_u8(pq.w, cp[s].w, t);
                    t += gap;                        /* (saturation only shrinks L1) */
                    om[s] = wave_ballot(valid[s] && gap <= (u32)kq && t <= k2);
                }
            }
            if (DENSE) {
                const u32 kk1 = out_k1<OutT>(kq);
                OutT *ow = out + orow + w;
                if (FULL && (rows_aligned || ((uintptr_t)ow & (4 * sizeof(OutT) - 1)) == 0)) {
                    /* k + 1 for all four; an open one is overwritten by its
                     * rescan later in this wave's program order (same
                     * address, same L2 channel: the later store lands last).
                     * Non-temporal: the matrix is written once and streams
                     * past L2 instead of evicting the choice data (1,024 x
                     * 1e6 length-64 search at k = 2: 0.70 -> 0.56 ms) */
                    OutT *p = (OutT *)__builtin_assume_aligned(ow + 4 * lane, 4 * sizeof(OutT));
                    if (sizeof(OutT) == 1) {
                        __builtin_nontemporal_store(kk1 * 0x01010101u, (u32 *)p);
                    } else if (sizeof(OutT) == 2) {
                        __builtin_nontemporal_store(kk1 * 0x00010001u, (u32 *)p);
                        __builtin_nontemporal_store(kk1 * 0x00010001u, (u32 *)p + 1);
                    } else {
#pragma unroll
                        for (int s = 0; s < 4; s++) __builtin_nontemporal_store(kk1, (u32 *)p + s);
                    }
                } else {
#pragma unroll
                    for (int s = 0; s < 4; s++)
                        if (valid[s]) ow[4 * lane + s] = (OutT)kk1;
                }
            }
            if ((om[0] | om[1] | om[2] | om[3]) == 0) continue;
#pragma unroll
            for (int s = 0; s < 4; s++) {
                if ((om[s] >> lane) & 1u) open_stack[ns + rank_below(om[s])] = orow + cl + s;
                ns = uni(ns + __popcll(om[s]));
            }
            wave_lds_sync();
            while (ns >= 64) drain(64);
        }
    };
    int64_t w = c0;
    /* pairs of full windows through the fast loop: each lane holds eight
     * choices (four in each window), so a query's per-pass work (readlanes,
     * loop, store address, ballots) covers twice the pairs (PAIRS8: only
     * where the rescans' registers leave room -- queries of one word) */
    auto window2 = [&](int64_t w) {
        const int64_t cl = w + 4 * lane;
        int n[8];
        Profile cp[8];
        const int4 la = *(const int4 *)(clen + cl);
        const int4 lb = *(const int4 *)(clen + cl + 256);
        n[0] = la.x; n[1] = la.y; n[2] = la.z; n[3] = la.w;
        n[4] = lb.x; n[5] = lb.y; n[6] = lb.z; n[7] = lb.w;
#pragma unroll
        for (int s = 0; s < 4; s++) {
            cp[s] = ccnt[cl + s];
            cp[4 + s] = ccnt[cl + 256 + s];
        }
        bool big = false;
#pragma unroll
        for (int s = 0; s < 8; s++) big = big || (u32)(n[s] - lbase) > 255u;
        if (wave_ballot(big) != 0) {
            window(w, std::true_type());
            window(w + 256, std::true_type());
            return;
        }
        int64_t orow = q0 * nc;
        for (int j = 0; j < gq; j++, orow += nc) {
            const int mqb = __builtin_amdgcn_readlane(lmb, j);
            const Profile pq{(u32)__builtin_amdgcn_readlane((int)lp.x, j),
                             (u32)__builtin_amdgcn_readlane((int)lp.y, j),
                             (u32)__builtin_amdgcn_readlane((int)lp.z, j),
                             (u32)__builtin_amdgcn_readlane((int)lp.w, j)};
            u32 t[8];
            bool half = false;
#pragma unroll
            for (int s = 0; s < 8; s++) {
                t[s] = sad_u8(pq.x, cp[s].x, 0u);
                t[s] = sad_u8(pq.y, cp[s].y, t[s]);
                t[s] = sad_u8((u32)(n[s] - lbase), (u32)mqb, t[s]);
                half = half || t[s] <= k2_uni;
            }
            if (DENSE) {
                u32 *p0 = (u32 *)__builtin_assume_aligned(out + orow + w + 4 * lane, 4 * sizeof(OutT));
                u32 *p1 = (u32 *)__builtin_assume_aligned(out + orow + w + 256 + 4 * lane, 4 * sizeof(OutT));
#pragma unroll
                for (int x = 0; x < (int)sizeof(OutT); x++) {
                    __builtin_nontemporal_store(fill_uni, p0 + x);
                    __builtin_nontemporal_store(fill_uni, p1 + x);
                }
            }
            if (wave_ballot(half) == 0) continue;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                u64 om[4];
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    u32 u = sad_u8(pq.z, cp[4 * h + s].z, t[4 * h + s]);
                    u = sad_u8(pq.w, cp[4 * h + s].w, u);
                    om[s] = wave_ballot(u <= k2_uni);
                }
                if ((om[0] | om[1] | om[2] | om[3]) == 0) continue;
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    if ((om[s] >> lane) & 1u) open_stack[ns + rank_below(om[s])] = orow + cl + 256 * h + s;
                    ns = uni(ns + __popcll(om[s]));
                }
                wave_lds_sync();
                while (ns >= 64) drain(64);
            }
        }
    };
    if (PAIRS8 && fast_tile && !gram)
        for (; w + 512 <= c1; w += 512) window2(w);
    for (; w + 256 <= c1; w += 256) window(w, std::true_type());
    if (w < c1) window(w, std::false_type());
}

/* Count mode: persistent waves pull (choice block, query group) tiles. Each
 * 256-choice window of the block is tested against every query of the group
 * (qgroup queries) with the length and class-count bounds, so a window's
 * lengths and profiles are loaded once per group and stay in registers; the
 * choices neither bound settles are stacked as output indices and rescanned
 * 64 at a time through the 32-row stage, each lane with its own query (the
 * group's tables in LDS for groups of up to CD_GL queries, else the global
 * Peq tables); the few pairs that stage leaves open (near matches) go on an
 * escalation stack and through the wider stages 64 at a time. Empty,
 * too-long and exact queries run lane-strided. Needs the choices' and the
 * queries' class profiles. */
template <int NWMAX, int GL, typename Sink, bool GRAM = false>
__global__ __launch_bounds__(256)
/* the four-word instance sits at the 128-VGPR line: keep it at 4 waves/SIMD;
 * the others within 168 VGPRs, 3 waves/SIMD */
__attribute__((amdgpu_waves_per_eu(NWMAX == 4 ? 4 : 3))) void cdist_count_kernel(
    const int64_t *__restrict__ oq, int64_t nq, const u64 *__restrict__ gpeq,
    const u32 *__restrict__ words, const int64_t *__restrict__ gbase,
    const int32_t *__restrict__ clen, const Profile *__restrict__ ccnt,
    const Profile *__restrict__ qcnt, const Profile *__restrict__ ccnt2,
    const Profile *__restrict__ qcnt2, const Profile3 *__restrict__ ccnt3,
    const Profile3 *__restrict__ qcnt3, int64_t nc, Sink sink,
    int k_all, const int32_t *__restrict__ kvec, int64_t tile, int grab, int qgroup,
    unsigned long long *counter, int first_static, int64_t *__restrict__ elist,
    unsigned long long *elist_count, int64_t ecap)
{
    extern __shared__ u64 lds[];
    CountLds<GL> &L = ((CountLds<GL> *)lds)[threadIdx.x >> 6];
    GramLds *gl = GRAM ? (GramLds *)((CountLds<GL> *)lds + 4) + (threadIdx.x >> 6) : nullptr;
    const int lane = lane_id();
    const int64_t nblk = (nc + tile - 1) / tile;
    const int64_t ng = (nq + qgroup - 1) / qgroup;
    const int64_t ntiles = nblk * ng;
    int ns = 0, nsc = 0, ne = 0;
    int64_t q0 = 0;
    /* pairs the 32-row stage leaves undecided (near matches, which need the
     * whole pattern) go through the wider stages 64 at a time, so one hard
     * pair does not hold 63 easy lanes for its whole scan. With an
     * escalation list in the workspace (cap > 0) they are appended there
     * instead (one atomic per batch) for cdist_list_kernel, which runs a
     * diagonal band first; what does not fit the list is scanned here. */
    auto escalate = [&](int cnt) __attribute__((noinline)) {
        const int base = ne - cnt;
        bool here = lane < cnt;
        if (here && ecap > 0) {
            const u64 pos = wave_reserve(elist_count, (u32)cnt) + (u64)lane;
            if (pos < (u64)ecap) {
                elist[pos] = L.esc[base + lane];
                here = false;
            }
        }
        if (here) {
            const int64_t idx = L.esc[base + lane];
            const int64_t q = idx / nc;
            const int64_t c = idx - q * nc;
            const int mq = (int)(oq[q + 1] - oq[q]);
            const int kq = kvec ? clamp_k_dev(kvec[q]) : k_all;
            const u64 *qp = gpeq + (size_t)q * 256 * NWMAX;
            const StripeText text{words + gbase[c >> 6] + (c & 63)};
            sink.put(idx, peq_lane<NWMAX, 1, false>((const u32 *)nullptr, qp, mq, text, clen[c], kq), kq);
        }
        ne = uni(base);
        wave_lds_sync();
    };
    /* rescans: each lane takes an entry of its own query and runs the 32-row
     * stage (the whole scan when the query fits it); with GL > 0 the group's
     * 32-row tables sit in this wave's LDS, else they come from the global
     * Peq tables */
    auto scan_drain = [&](int cnt) {
        const int base = nsc - cnt;
        bool up = false;
        int64_t idx = 0;
        if (lane < cnt) {
            idx = L.scan[base + lane];
            const int64_t q = idx / nc;
            const int64_t c = idx - q * nc;
            const int mq = (int)(oq[q + 1] - oq[q]);
            const int kq = kvec ? clamp_k_dev(kvec[q]) : k_all;
            const int n = clen[c];
            const u64 *qp = gpeq + (size_t)q * 256 * NWMAX;
            const StripeText text{words + gbase[c >> 6] + (c & 63)};
            int d;
            if ((n > mq ? n - mq : mq - n) > kq) {
                d = kq + 1;                          /* (count windows settle these) */
            } else if (n == 0) {
                d = mq;
            } else if (GL > 0) {
                const u32 *lo = &L.lo32[(int)(q - q0)][0];
                d = scan_text<u32, 1, true>(mq, 1, 32, text, n, kq,
                                            [&](u32 ch, int) { return lo[ch]; });
            } else {
                const GlobalLo32<NWMAX> lo{qp};
                d = scan_text<u32, 1, true>(mq, 1, 32, text, n, kq,
                                            [&](u32 ch, int) { return lo[ch]; });
            }
            if (d >= 0) sink.put(idx, d, kq);
            else up = true;                          /* rows past 32 needed */
        }
        const u64 um = wave_ballot(up);
        if (up) L.esc[ne + rank_below(um)] = idx;
        ne = uni(ne + __popcll(um));
        nsc = uni(base);
        wave_lds_sync();
        while (ne >= 64) escalate(64);
    };
    /* open entries first meet the 2-gram bound (when the choices carry
     * 2-gram profiles); those it settles already hold k + 1, the rest are
     * compacted onto the scan stack, so scans run 64 lanes at a time */
    auto drain = [&](int cnt) {
        const int base = ns - cnt;
        bool keep = false;
        int64_t idx = 0;
        if (lane < cnt) {
            idx = L.open[base + lane];
            keep = true;
            if (ccnt2) {
                const int64_t q = idx / nc;
                const int64_t c = idx - q * nc;
                const int mq = (int)(oq[q + 1] - oq[q]);
                const int kq = kvec ? clamp_k_dev(kvec[q]) : k_all;
                const int n = clen[c];
                keep = !count2_over(qcnt2[q], ccnt2[c], mq, n, kq);
                if (keep && ccnt3) keep = !count3_over(qcnt3[q], ccnt3[c], mq, n, kq);
            }
        }
        const u64 km = wave_ballot(keep);
        if (GRAM || ccnt3) {
            /* after the L2 prefilter (or the 3-gram bound) the open pairs are
             * mostly near matches, which the 32-row stage cannot settle:
             * straight to the whole pattern */
            if (keep) L.esc[ne + rank_below(km)] = idx;
            ne = uni(ne + __popcll(km));
            ns = uni(base);
            wave_lds_sync();
            while (ne >= 64) escalate(64);
            return;
        }
        if (keep) L.scan[nsc + rank_below(km)] = idx;
        nsc = uni(nsc + __popcll(km));
        ns = uni(base);
        wave_lds_sync();
        while (nsc >= 64) scan_drain(64);
    };
    Tickets tk;
    tk.first_static = first_static != 0;
    for (;;) {
        const long long t = tk.get(counter, grab);
        if (t >= ntiles) break;
        const int64_t blk = t / ng;
        q0 = (t % ng) * qgroup;
        const int64_t q1 = q0 + qgroup < nq ? q0 + qgroup : nq;
        if (GL > 0) {
            /* the group's 32-row tables (qgroup <= GL), for the rescans */
            wave_lds_sync();
            for (int j = 0; j < (int)(q1 - q0); j++) {
                const u64 *qp = gpeq + (size_t)(q0 + j) * 256 * NWMAX;
#pragma unroll
                for (int x = 0; x < 4; x++) L.lo32[j][x * 64 + lane] = (u32)qp[(x * 64 + lane) * NWMAX];
            }
            wave_lds_sync();
        }
        const int64_t c0 = blk * tile;
        const int64_t c1 = c0 + tile < nc ? c0 + tile : nc;
        /* empty, too long and exact queries of the group (lane j looks at
         * query q0 + j, so the loads run in parallel), then only those */
        u64 special;
        {
            bool sp = false;
            if (lane < (int)(q1 - q0)) {
                const int64_t q = q0 + lane;
                const int mq = (int)(oq[q + 1] - oq[q]);
                const int kq = kvec ? clamp_k_dev(kvec[q]) : k_all;
                sp = mq == 0 || mq > 64 * NWMAX || kq < 0;
            }
            special = wave_ballot(sp);
        }
        while (special) {
            const int64_t q = q0 + __builtin_ctzll(special);
            special &= special - 1;
            const int mq = (int)(oq[q + 1] - oq[q]);
            const int kq = kvec ? clamp_k_dev(kvec[q]) : k_all;
            const int64_t orow = q * nc;
            if (mq == 0 || mq > 64 * NWMAX) {
                /* empty query: n, or k + 1 (csrc/gpu_kernels.cu:53-54); too long
                 * for this launch: caller error, marked -1 */
                for (int64_t c = c0 + lane; c < c1; c += 64)
                    sink.put(orow + c, mq == 0 ? empty_pattern(clen[c], kq) : -1, kq);
            } else if (kq < 0) {
                const u64 *qp = gpeq + (size_t)q * 256 * NWMAX;
                const GlobalLo32<NWMAX> lo{qp};
                for (int64_t c = c0 + lane; c < c1; c += 64) {
                    const StripeText text{words + gbase[c >> 6] + (c & 63)};
                    sink.put(orow + c, peq_lane<NWMAX, 0, false>(lo, qp, mq, text, clen[c], kq), kq);
                }
            }
        }
        /* the group's lengths, thresholds and profiles, query q0 + j in lane j
         * (qgroup <= 64), read back with readlane: no memory access per window */
        const int gq = (int)(q1 - q0);
        int lm = 0, lk = -1;
        Profile lp{0u, 0u, 0u, 0u};
        if (lane < gq) {
            const int64_t q = q0 + lane;
            lm = (int)(oq[q + 1] - oq[q]);
            lk = kvec ? clamp_k_dev(kvec[q]) : k_all;
            lp = qcnt[q];
        }
        const u64 okmask = wave_ballot(lane < gq && lm > 0 && lm <= 64 * NWMAX && lk >= 0);
        if (okmask) {
            if (!Sink::kDense) {
                count_windows<false, NWMAX == 1, GRAM>(L.open, clen, ccnt, c0, c1, nc, q0, gq, okmask, lm, lk,
                                                       !kvec, lp, (u8 *)nullptr, false, false, ns, drain, gl);
            } else {
                void *base = dense_out(sink);
                const int dt = dense_dtype(sink);
                const size_t esz = dt == DT_UINT8 ? 1 : dt == DT_UINT16 ? 2 : 4;
                const bool aligned = (nc & 3) == 0 && ((uintptr_t)base & (4 * esz - 1)) == 0;
                const bool a16 = ((nc * esz) & 15) == 0 && ((uintptr_t)base & 15) == 0;
                if (dt == DT_UINT8)
                    count_windows<true, NWMAX == 1, GRAM>(L.open, clen, ccnt, c0, c1, nc, q0, gq, okmask, lm, lk,
                                                          !kvec, lp, (u8 *)base, aligned, a16, ns, drain, gl);
                else if (dt == DT_UINT16)
                    count_windows<true, NWMAX == 1, GRAM>(L.open, clen, ccnt, c0, c1, nc, q0, gq, okmask, lm, lk,
                                                          !kvec, lp, (uint16_t *)base, aligned, a16, ns, drain, gl);
                else
                    count_windows<true, NWMAX == 1, GRAM>(L.open, clen, ccnt, c0, c1, nc, q0, gq, okmask, lm, lk,
                                                          !kvec, lp, (int32_t *)base, aligned, a16, ns, drain, gl);
            }
        }
        /* with the group's tables in LDS (GL > 0) the stacks empty before the
         * next group; with global tables they carry over to the next tile */
        if (GL > 0) {
            while (ns > 0) drain(ns < 64 ? ns : 64);
            while (nsc > 0) scan_drain(nsc < 64 ? nsc : 64);
            while (ne > 0) escalate(ne < 64 ? ne : 64);
        }
    }
    while (ns > 0) drain(ns < 64 ? ns : 64);
    while (nsc > 0) scan_drain(nsc < 64 ? nsc : 64);
    while (ne > 0) escalate(ne < 64 ? ne : 64);
}

/* The escalation list of a count-mode launch: every (query, choice) entry
 * (output index q * nc + c) the bounds and the 32-row stage left open --
 * near matches, mostly -- in batches of 64 per wave. A match within k lies on
 * at most k + 1 diagonals (Ukkonen), so a diagonal band decides most of them:
 * one window offset for the batch from its lanes' length differences, 32
 * rows when that window reaches every lane's k, else 64 (band_scan: bit
 * planes of the query bytes, the choice read from its stripe); a lane the
 * window cannot reach runs the whole pattern (peq_lane). A kernel of its own
 * so that the band's planes do not cost the count kernel its occupancy.
 * (The 256 x 4.2M DNA search at k = 8 spent 1.0 of its 1.8 ms in the whole
 * scans of its 420K near matches: four 64-row words a column, Eq from global
 * memory.) */
template <int NWMAX, typename Sink>
__global__ __launch_bounds__(256) void cdist_list_kernel(
    const u8 *__restrict__ bq, const int64_t *__restrict__ oq, const u64 *__restrict__ gpeq,
    const u32 *__restrict__ words, const int64_t *__restrict__ gbase, const int32_t *__restrict__ clen,
    int64_t nc, Sink sink, int k_all, const int32_t *__restrict__ kvec, const int64_t *__restrict__ elist,
    const unsigned long long *__restrict__ elist_count, int64_t ecap, const u8 *__restrict__ cbuf,
    const int64_t *__restrict__ coff)
{
    const int lane = lane_id();
    const unsigned long long cnt_all = *elist_count;
    const int64_t total = cnt_all < (unsigned long long)ecap ? (int64_t)cnt_all : ecap;
    const int64_t waves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t b = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; 64 * b < total; b += waves) {
        const int64_t e = 64 * b + lane;
        const bool live = e < total;
        int64_t idx = 0, q = 0, c = 0;
        int mq = 0, kq = -1, n = 0, dl = 0;
        if (live) {
            idx = elist[e];
            q = idx / nc;
            c = idx - q * nc;
            mq = (int)(oq[q + 1] - oq[q]);
            kq = kvec ? clamp_k_dev(kvec[q]) : k_all;
            n = clen[c];
            dl = mq - n;
        }
        const bool banded = live && kq >= 0 && n > 0 && mq > 0;
        const int dmin = wave_min(banded ? dl : 0x7fffffff), dmax = wave_max(banded ? dl : -0x7fffffff);
        const int d32 = band_offset<32>(dmin, dmax), d64 = band_offset<64>(dmin, dmax);
        const bool w32 = wave_ballot(banded && kq > band_reach<32>(dl, d32)) == 0;
        /* the text from the packed bytes when the set has them (16 aligned
         * bytes a load, the string in 2-3 cache lines), else from its stripe
         * (one line per dword: 58 lines for a 230-byte choice) */
        auto run = [&](const auto &text) {
            int d = -1;
            if (banded) {
                if (w32)
                    d = band_scan<u32>(bq + oq[q], mq, text, n, d32, kq, true);
                else if (kq <= band_reach<64>(dl, d64))
                    d = band_scan<u64>(bq + oq[q], mq, text, n, d64, kq, true);
            }
            if (live && d < 0) {
                const u64 *qp = gpeq + (size_t)q * 256 * NWMAX;
                d = peq_lane<NWMAX, 1, false>((const u32 *)nullptr, qp, mq, text, n, kq);
            }
            if (live) sink.put(idx, d, kq);
        };
        if (cbuf)
            run(BlockText{cbuf + coff[c]});
        else
            run(StripeText{words + gbase[c >> 6] + (c & 63)});
    }
}

/* ======================================================================= pairs */
/* Pairs are bucketed within chunks of CH consecutive pairs (CH a multiple of
 * 256), chunk-major, so the 64 pairs of a tile lie close together in the
 * string buffers (a global sort would scatter every tile over gigabytes and
 * thrash the TLB). Per chunk c the order is: long pairs, then buckets by
 * descending key. Arrays: hist / cursor [nchunks * PKEYS], chunk_base and
 * tile_base [nchunks + 1]. Keys come from the offsets alone (pair_key). */
__device__ __forceinline__ int pair_k(int k_all, const int32_t *kvec, int64_t i)
{
    return kvec ? clamp_k_dev(kvec[i]) : k_all;
}

__global__ __launch_bounds__(1024) void pairs_classify_kernel(
    const int64_t *__restrict__ oa, const int64_t *__restrict__ ob, int64_t n, void *out, int dtype,
    int k_all, const int32_t *__restrict__ kvec, int bin_width, int band, int long_units, int unit64,
    int64_t chunk, unsigned short *key, unsigned long long *hist)
{
    __shared__ unsigned int lh[PKEYS];
    for (int x = threadIdx.x; x < PKEYS; x += blockDim.x) lh[x] = 0;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        int result = 0;
        const int kk = pair_key((int)(oa[i + 1] - oa[i]), (int)(ob[i + 1] - ob[i]),
                                pair_k(k_all, kvec, i), bin_width, band, long_units, unit64 != 0, result);
        if (kk == 0) store_out(out, i, result, dtype);
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

/* Pair indices into bucket order. A block ranks its pairs per key in LDS and
 * reserves each key's range of the chunk with one global atomic, so pairs that
 * share a bucket (all of them, in a uniform batch) do not contend on one
 * global counter. The order inside a bucket is immaterial. */
__global__ __launch_bounds__(1024) void pairs_scatter_kernel(
    const unsigned short *key, int64_t n, int64_t chunk, unsigned long long *cursor,
    const unsigned long long *chunk_base, int32_t *sorted)
{
    __shared__ unsigned int cnt[PKEYS];
    __shared__ unsigned long long base[PKEYS];
    for (int x = threadIdx.x; x < PKEYS; x += blockDim.x) cnt[x] = 0;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int kk = i < n ? key[i] : 0;
    const unsigned int rank = kk ? atomicAdd(&cnt[kk], 1u) : 0u;
    __syncthreads();
    /* a block never straddles chunks (chunk is a multiple of the block size) */
    const int64_t c = ((int64_t)blockIdx.x * blockDim.x) / chunk;
    for (int x = threadIdx.x; x < PKEYS; x += blockDim.x)
        if (cnt[x]) base[x] = atomicAdd(&cursor[c * PKEYS + x], (unsigned long long)cnt[x]);
    __syncthreads();
    if (kk) sorted[chunk_base[c] + base[kk] + rank] = (int32_t)i;
}

/* One wave on one long pair (shorter string above 256 bytes before trimming):
 * trim, swap and the gap / empty rules (long_prepare), then the anti-diagonal
 * wavefront. The pair index is the same in every lane. */
__device__ void pairs_long_item(const u8 *__restrict__ ba, const int64_t *__restrict__ oa,
                                const u8 *__restrict__ bb, const int64_t *__restrict__ ob,
                                int64_t i, int trim, int k, void *out, int dtype, int8_t *hbuf,
                                int64_t slot)
{
    const int lane = lane_id();
    const int64_t a0 = oa[i], b0 = ob[i];
    const u8 *p = ba + a0, *t = bb + b0;
    int m = (int)(oa[i + 1] - a0), n = (int)(ob[i + 1] - b0);
    int r = long_prepare(p, m, t, n, 1, trim, k);
    if (r < 0) {
        m = uni(m);
        n = uni(n);
        LongCtx C;
        C.p = p;
        C.m = m;
        C.t = t;
        C.n = n;
        C.W = (m + 63) >> 6;
        C.hbuf = hbuf;
        const int S = (C.W + 63) >> 6;
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
    }
    if (lane == 0) store_out(out, i, r, dtype);
}

template <int W32>
__device__ void pairs_tile(const u8 *__restrict__ ba, const int64_t *__restrict__ oa,
                           const u8 *__restrict__ bb, const int64_t *__restrict__ ob,
                           const int32_t *sorted, int64_t s0, int64_t s1, int trim, int k_all,
                           const int32_t *__restrict__ kvec, void *out, int dtype)
{
    const int64_t s = s0 + lane_id();
    if (s >= s1) return;
    const int64_t i = sorted[s];
    store_out(out, i, packed_pair_w32<W32>(ba, oa, bb, ob, i, trim, pair_k(k_all, kvec, i)), dtype);
}

/* Pair indices a wave's band tiles could not decide, scanned in full 64 at a
 * time by the same wave (its LDS stack), so those scans overlap the other
 * waves' band tiles. */
constexpr int PB_ESC = 128;

/* A band tile: every lane scans its pair in a diagonal window at one offset
 * for the whole tile, chosen from the lanes' length differences (band_offset);
 * 32 rows when every threshold is at most 14, else 64. A lane whose distance
 * the window does not reach pushes its pair onto the wave's stack. */
__device__ void pairs_tile_band(const u8 *__restrict__ ba, const int64_t *__restrict__ oa,
                                const u8 *__restrict__ bb, const int64_t *__restrict__ ob,
                                const int32_t *sorted, int64_t s0, int64_t s1, int trim, int k_all,
                                const int32_t *__restrict__ kvec, void *out, int dtype, int32_t *esc,
                                int &nesc)
{
    const int64_t s = s0 + lane_id();
    const bool act = s < s1;
    int64_t i = 0;
    int k = 0, res = -1, la = 0, lb = 0;
    const u8 *a = ba, *b = bb;
    if (act) {
        i = sorted[s];
        k = pair_k(k_all, kvec, i);
        const int64_t a0 = oa[i], b0 = ob[i];
        a = ba + a0;
        b = bb + b0;
        la = (int)(oa[i + 1] - a0);
        lb = (int)(ob[i + 1] - b0);
        res = band_prepare(a, la, b, lb, trim, k);
    }
    const bool need = act && res < 0;
    const int dmin = uni(wave_min(need ? la - lb : 0x7fffffff));
    const int dmax = uni(wave_max(need ? la - lb : -0x7fffffff));
    const bool narrow = wave_ballot(need && !(k >= 0 && k <= 14)) == 0;
    if (need) {
        if (narrow) res = band_pair<u32>(a, la, b, lb, k, band_offset<32>(dmin, dmax));
        else res = band_pair<u64>(a, la, b, lb, k, band_offset<64>(dmin, dmax));
    }
    if (act && res >= 0) store_out(out, i, res, dtype);
    const u64 em = wave_ballot(act && res < 0);
    if (em) {
        if (act && res < 0) esc[nesc + rank_below(em)] = (int32_t)i;
        nesc += (int)__popcll(em);
        wave_lds_sync();
    }
}

/* The top cnt (<= 64) pairs of the wave's stack through the whole scan, in the
 * 32-row units of the widest of them. */
__device__ void pairs_esc_flush(const u8 *__restrict__ ba, const int64_t *__restrict__ oa,
                                const u8 *__restrict__ bb, const int64_t *__restrict__ ob,
                                int trim, int k_all, const int32_t *__restrict__ kvec, void *out,
                                int dtype, const int32_t *esc, int &nesc, int cnt)
{
    const int base = nesc - cnt;
    int w = 0;
    if (lane_id() < cnt) {
        const int64_t i = esc[base + lane_id()];
        const int la = (int)(oa[i + 1] - oa[i]), lb = (int)(ob[i + 1] - ob[i]);
        w = ((la < lb ? la : lb) + 31) >> 5;
    }
    w = uni(wave_max(w));
    switch (w) {
    case 1: pairs_tile<1>(ba, oa, bb, ob, esc, base, base + cnt, trim, k_all, kvec, out, dtype); break;
    case 2: pairs_tile<2>(ba, oa, bb, ob, esc, base, base + cnt, trim, k_all, kvec, out, dtype); break;
    case 3: pairs_tile<3>(ba, oa, bb, ob, esc, base, base + cnt, trim, k_all, kvec, out, dtype); break;
    case 4: pairs_tile<4>(ba, oa, bb, ob, esc, base, base + cnt, trim, k_all, kvec, out, dtype); break;
    case 5: pairs_tile<5>(ba, oa, bb, ob, esc, base, base + cnt, trim, k_all, kvec, out, dtype); break;
    case 6: pairs_tile<6>(ba, oa, bb, ob, esc, base, base + cnt, trim, k_all, kvec, out, dtype); break;
    case 7: pairs_tile<7>(ba, oa, bb, ob, esc, base, base + cnt, trim, k_all, kvec, out, dtype); break;
    default: pairs_tile<8>(ba, oa, bb, ob, esc, base, base + cnt, trim, k_all, kvec, out, dtype); break;
    }
    nesc = base;
    wave_lds_sync();                             /* reads done before the next pushes */
}

/* (at most 128 VGPRs: four waves per SIMD; without the cap the scheduler
 * interleaves the unrolled columns of the widest tiles into ~200 registers) */
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void pairs_batched_kernel(
    const u8 *__restrict__ ba, const int64_t *__restrict__ oa, const u8 *__restrict__ bb,
    const int64_t *__restrict__ ob, int trim, int k_all, const int32_t *__restrict__ kvec,
    const int32_t *__restrict__ sorted, const unsigned short *__restrict__ key_of,
    const unsigned long long *__restrict__ hist, const unsigned long long *__restrict__ chunk_base,
    const unsigned long long *__restrict__ tile_base, int64_t nchunks, void *out, int dtype,
    int grab, unsigned long long *counter, int8_t *scratch, int64_t slot)
{
    __shared__ int32_t esc_all[4][PB_ESC];
    int32_t *esc = esc_all[threadIdx.x >> 6];
    int nesc = 0;
    const int64_t ntiles = (int64_t)tile_base[nchunks];
    const int64_t wave_gid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    int8_t *hbuf = scratch ? scratch + wave_gid * slot : nullptr;
    int64_t c = 0;
    Tickets tk;
    for (;;) {
        const long long t = tk.get(counter, grab);
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
            const int64_t i = uni64(sorted[cb + lt]);
            pairs_long_item(ba, oa, bb, ob, i, trim, pair_k(k_all, kvec, i), out, dtype, hbuf, slot);
            continue;
        }
        const int64_t s0 = cb + nlong + (lt - nlong) * 64;
        const int64_t ce = (int64_t)chunk_base[c + 1];
        const int64_t s1 = s0 + 64 < ce ? s0 + 64 : ce;
        /* descending keys: the first pair of a tile has its largest key (a band
         * key, or the largest unit count of a whole-pattern tile) */
        const int key0 = uni(key_of[sorted[s0]]);
        if (key_band(key0)) {
            pairs_tile_band(ba, oa, bb, ob, sorted, s0, s1, trim, k_all, kvec, out, dtype, esc, nesc);
            if (nesc >= 64)
                pairs_esc_flush(ba, oa, bb, ob, trim, k_all, kvec, out, dtype, esc, nesc, 64);
            continue;
        }
        switch (key_units(key0)) {
        case 1: pairs_tile<1>(ba, oa, bb, ob, sorted, s0, s1, trim, k_all, kvec, out, dtype); break;
        case 2: pairs_tile<2>(ba, oa, bb, ob, sorted, s0, s1, trim, k_all, kvec, out, dtype); break;
        case 3: pairs_tile<3>(ba, oa, bb, ob, sorted, s0, s1, trim, k_all, kvec, out, dtype); break;
        case 4: pairs_tile<4>(ba, oa, bb, ob, sorted, s0, s1, trim, k_all, kvec, out, dtype); break;
        case 5: pairs_tile<5>(ba, oa, bb, ob, sorted, s0, s1, trim, k_all, kvec, out, dtype); break;
        case 6: pairs_tile<6>(ba, oa, bb, ob, sorted, s0, s1, trim, k_all, kvec, out, dtype); break;
        case 7: pairs_tile<7>(ba, oa, bb, ob, sorted, s0, s1, trim, k_all, kvec, out, dtype); break;
        default: pairs_tile<8>(ba, oa, bb, ob, sorted, s0, s1, trim, k_all, kvec, out, dtype); break;
        }
    }
    while (nesc > 0)
        pairs_esc_flush(ba, oa, bb, ob, trim, k_all, kvec, out, dtype, esc, nesc, nesc < 64 ? nesc : 64);
}

/* Direct mode: one launch, no bucketing. Waves take 64 consecutive pairs; each
 * lane runs its own pair through the lane code of its word count (up to
 * NWMAX words), and the wave then runs the long ones (shorter string above
 * 256 bytes, NWMAX = 4 only) one after another. Best for batches too small, or
 * too uniform, for bucketing to pay. */
template <int NWMAX>
__global__ __launch_bounds__(256) void pairs_direct_kernel(
    const u8 *__restrict__ ba, const int64_t *__restrict__ oa, const u8 *__restrict__ bb,
    const int64_t *__restrict__ ob, int64_t n, int trim, int k_all,
    const int32_t *__restrict__ kvec, void *out, int dtype, int8_t *scratch, int64_t slot)
{
    const int lane = lane_id();
    const int64_t wave_gid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    int8_t *hbuf = scratch ? scratch + wave_gid * slot : nullptr;
    for (int64_t b0 = wave_gid * 64; b0 < n; b0 += nwaves * 64) {
        const int64_t i = b0 + lane;
        bool is_long = false;
        if (i < n) {
            const int la = (int)(oa[i + 1] - oa[i]), lb = (int)(ob[i + 1] - ob[i]);
            const int mn = la < lb ? la : lb;
            const int k = pair_k(k_all, kvec, i);
            if (mn > 64 * NWMAX) {
                is_long = true;                  /* only when NWMAX = 4: above 256 bytes */
            } else {
                int d;
                if (NWMAX == 1 || mn <= 64) d = packed_pair<1>(ba, oa, bb, ob, i, trim, k);
                else if (NWMAX == 2 || mn <= 128) d = packed_pair<NWMAX >= 2 ? 2 : 1>(ba, oa, bb, ob, i, trim, k);
                else d = packed_pair<NWMAX>(ba, oa, bb, ob, i, trim, k);
                store_out(out, i, d, dtype);
            }
        }
        u64 lm = wave_ballot(is_long);
        while (lm) {
            const int l = __builtin_ctzll(lm);
            lm &= lm - 1;
            const int64_t il = b0 + l;
            pairs_long_item(ba, oa, bb, ob, il, trim, pair_k(k_all, kvec, il), out, dtype, hbuf, slot);
        }
    }
}

int launch_clear(unsigned long long *p, int n, hipStream_t s)
{
    clear_kernel<<<(n + 255) / 256, 256, 0, s>>>(p, n);
    BL_LAUNCHED("clear_kernel");
    return BL_OK;
}

int device_cus();

/* Blocks of `kernel` the device holds at once. A persistent launch of more
 * only adds waves that start when others have exited -- with the first run of
 * tickets given to every wave statically, their tiles would then run last. */
template <typename K>
int64_t resident_blocks(K kernel, int block, size_t lds)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, lds) != hipSuccess || per_cu < 1)
        per_cu = 1;
    return (int64_t)per_cu * device_cus();
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

static int cdist_nwmax(int max_qlen)
{
    const int words = (max_qlen + 63) >> 6;
    return words <= 1 ? 1 : words <= 2 ? 2 : words <= 4 ? 4 : 8;
}

/* Workspace: the work queues (QUEUE_BYTES), the queries' Peq tables,
 * then one 16-byte class profile and one 2-gram profile per query, then one
 * 64-byte 3-gram profile per query. */
static size_t cdist_peq_end(int64_t nq, int max_qlen)
{
    return QUEUE_BYTES + (size_t)nq * 256 * cdist_nwmax(max_qlen) * sizeof(u64);
}

/* The part of the workspace every launch needs: queues, Peq tables, the
 * queries' profiles. Then, 256-byte aligned, the escalation list of count
 * mode: its counter on a line of its own, then as many 8-byte entries as the
 * rest of the workspace holds (none: near matches are scanned in the count
 * kernel itself). */
static size_t cdist_ws_core(int64_t nq, int max_qlen)
{
    return cdist_peq_end(nq, max_qlen) + 2 * (size_t)nq * sizeof(Profile) + (size_t)nq * sizeof(Profile3);
}

static size_t cdist_list_start(int64_t nq, int max_qlen) { return (cdist_ws_core(nq, max_qlen) + 255) / 256 * 256; }

size_t bl_cdist_workspace_size(int64_t nq, int max_qlen)
{
    if (nq < 0 || max_qlen < 0) return 0;
    const int64_t entries = nq * 2048 < ((int64_t)1 << 20) ? nq * 2048 : ((int64_t)1 << 20);
    return cdist_list_start(nq, max_qlen) + 256 + (size_t)entries * 8;
}

}  /* extern "C" */

/* The search launch behind bl_cdist_tuned and bl_cdist_matches. */
template <typename Sink>
static int cdist_launch(const char *who, const uint8_t *d_buf_q, const int64_t *d_off_q, int64_t nq,
                        const bl_stripes *choices, int max_qlen, int k, const int32_t *d_kvec,
                        void *d_work, size_t work_bytes, const bl_cdist_params *prm, void *stream,
                        const Sink &sink, unsigned long long *extra_clear)
{
    const int64_t nc = choices->n;
    const size_t need_ws = cdist_ws_core(nq, max_qlen);
    if (d_work && work_bytes < need_ws)
        return fail(BL_EINVAL, "%s: workspace of %zu bytes below the %zu needed", who, work_bytes,
                    need_ws);
    const size_t ws_bytes = d_work ? work_bytes : bl_cdist_workspace_size(nq, max_qlen);
    const size_t lstart = cdist_list_start(nq, max_qlen);
    /* the escalation list: what the workspace holds past its core, unless
     * switched off (bit 2 of reserved[3]) */
    const bool use_list = !(prm && (prm->reserved[3] & 4));
    const int64_t ecap = use_list && ws_bytes >= lstart + 256 + 8 ? (int64_t)((ws_bytes - lstart - 256) / 8) : 0;
    hipStream_t s = (hipStream_t)stream;
    const int wpc = prm && prm->waves_per_cu > 0 ? prm->waves_per_cu : 16;
    /* tiles start on stripe groups (multiples of 64 choices; of 256 for the
     * four-choice filter's windows) */
    const int64_t tile = (prm && prm->tile > 0 ? (prm->tile + 255) / 256 * 256 : 8192);
    int mode = prm && prm->reserved[0] >= 1 && prm->reserved[0] <= 6 ? prm->reserved[0] : 0;
    if (mode == 0) mode = k < 0 && !d_kvec ? 5 : 6;  /* automatic: exact scans for exact searches */
    if (choices->order) mode = 5;                    /* a length-ordered set: exact scans only */
    if (mode == 4 && !choices->counts) mode = 6;   /* count mode needs the choices' profiles */
    const int grab = prm && prm->reserved[1] > 0 ? prm->reserved[1] : 1;
    const int qgroup = mode == 4 ? (prm->reserved[2] > 0 ? (prm->reserved[2] < 64 ? prm->reserved[2] : 64) : 16)
                                 : 1;
    const int nwmax = cdist_nwmax(max_qlen);
    /* count mode over queries of at most 64 bytes: matrix-core tiles unless
     * switched off (bit 0 of reserved[3]) */
    const bool gram = mode == 4 && max_qlen <= 64 && !(prm && (prm->reserved[3] & 1));
    /* count mode: each wave's first tile statically, unless switched off (bit 1) */
    const int first_static = prm && (prm->reserved[3] & 2) ? 0 : 1;
    const size_t lds = sizeof(WaveLds) * 4;
    /* exact scans: one workgroup per (query, block of xtile slots); by default
     * blocks sized for about 16 workgroups per CU over the batch, 256 to 4096
     * slots, and never more than 2^24 workgroups */
    int64_t xtile = prm && prm->tile > 0 ? (prm->tile + 63) / 64 * 64 : 0;
    if (!xtile) {
        xtile = nq * nc / ((int64_t)device_cus() * 16) / 256 * 256;
        xtile = xtile < 256 ? 256 : xtile > 4096 ? 4096 : xtile;
    }
    while (nq * ((nc + xtile - 1) / xtile) > ((int64_t)1 << 24)) xtile *= 2;
    const int64_t xblk = (nc + xtile - 1) / xtile;
    int64_t blocks = (int64_t)device_cus() * wpc / 4;
    const int64_t ntiles = ((nc + tile - 1) / tile) * ((nq + qgroup - 1) / qgroup);
    const int64_t need = (ntiles + 3) / 4;
    if (blocks > need) blocks = need;
    if (blocks < 1) blocks = 1;
    unsigned char *ws = (unsigned char *)d_work;
    if (!ws) BL_HIP(hipMallocAsync((void **)&ws, ws_bytes, s));
    unsigned long long *elist_count = ecap > 0 ? (unsigned long long *)(ws + lstart) : nullptr;
    int64_t *elist = ecap > 0 ? (int64_t *)(ws + lstart + 256) : nullptr;
    unsigned long long *ctr = (unsigned long long *)ws;
    u64 *gpeq = (u64 *)(ws + QUEUE_BYTES);
    /* query class profiles after the Peq tables, when the choices carry theirs */
    const Profile *ccnt = (const Profile *)choices->counts;
    Profile *qcnt = ccnt ? (Profile *)(ws + cdist_peq_end(nq, max_qlen)) : nullptr;
    const Profile *ccnt2 = ccnt ? (const Profile *)choices->counts2 : nullptr;
    Profile *qcnt2 = ccnt2 ? (Profile *)(ws + cdist_peq_end(nq, max_qlen)) + nq : nullptr;
    /* the 3-gram profiles after the 2-gram ones (64 bytes a query) */
    const Profile3 *ccnt3 = ccnt2 ? (const Profile3 *)choices->counts3 : nullptr;
    Profile3 *qcnt3 = ccnt3 ? (Profile3 *)((Profile *)(ws + cdist_peq_end(nq, max_qlen)) + 2 * nq) : nullptr;
    int rc = BL_OK;
    {
        /* the Peq tables and profiles; block 0 zeroes the counters */
        peq_build_kernel<<<(unsigned)nq, 256, 0, s>>>(d_buf_q, d_off_q, nwmax, gpeq, qcnt, qcnt2, qcnt3,
                                                     ctr, (int)(QUEUE_BYTES / 8), extra_clear,
                                                     mode == 4 ? elist_count : nullptr);
        auto capped = [&](auto kernel, size_t lds_bytes) {
            const int64_t cap = resident_blocks(kernel, 256, lds_bytes);
            return (unsigned)(blocks < cap ? blocks : cap);
        };
        auto go = [&](auto nwc) {
            constexpr int NW = decltype(nwc)::value;
            if (NW == 1 && mode == 4 && gram && qgroup > CD_GL) {
                auto kern = cdist_count_kernel<1, 0, Sink, true>;
                const size_t lb = (sizeof(CountLds<0>) + sizeof(GramLds)) * 4;
                kern<<<capped(kern, lb), 256, lb, s>>>(
                    d_off_q, nq, gpeq, choices->words, choices->gbase, choices->len, ccnt, qcnt, ccnt2,
                    qcnt2, ccnt3, qcnt3, nc, sink, k, d_kvec, tile, grab, qgroup, ctr, first_static, elist, elist_count, ecap);
            } else if (mode == 4 && qgroup <= CD_GL) {
                auto kern = cdist_count_kernel<NW, CD_GL, Sink>;
                const size_t lb = sizeof(CountLds<CD_GL>) * 4;
                kern<<<capped(kern, lb), 256, lb, s>>>(
                    d_off_q, nq, gpeq, choices->words, choices->gbase, choices->len, ccnt, qcnt, ccnt2,
                    qcnt2, ccnt3, qcnt3, nc, sink, k, d_kvec, tile, grab, qgroup, ctr, first_static, elist, elist_count, ecap);
            } else if (mode == 4) {
                auto kern = cdist_count_kernel<NW, 0, Sink>;
                const size_t lb = sizeof(CountLds<0>) * 4;
                kern<<<capped(kern, lb), 256, lb, s>>>(
                    d_off_q, nq, gpeq, choices->words, choices->gbase, choices->len, ccnt, qcnt, ccnt2,
                    qcnt2, ccnt3, qcnt3, nc, sink, k, d_kvec, tile, grab, qgroup, ctr, first_static, elist, elist_count, ecap);
            } else if (mode == 5) {
                cdist_exact_kernel<NW, Sink><<<(unsigned)(nq * xblk), 256, 0, s>>>(
                    d_off_q, nq, gpeq, choices->words, choices->gbase, choices->len, choices->order, nc,
                    sink, k, d_kvec, xtile);
            } else {
                auto kern = cdist_batched_kernel<NW, Sink>;
                kern<<<capped(kern, lds), 256, lds, s>>>(
                    d_off_q, nq, gpeq, choices->words, choices->gbase, choices->len, ccnt, qcnt, nc,
                    sink, k, d_kvec, tile, mode == 6 ? 0 : mode, grab, ctr);
            }
        };
        switch (nwmax) {
        case 1: go(std::integral_constant<int, 1>()); break;
        case 2: go(std::integral_constant<int, 2>()); break;
        case 4: go(std::integral_constant<int, 4>()); break;
        default: go(std::integral_constant<int, 8>()); break;
        }
        /* count mode: the near matches the count kernel listed */
        auto lgo = [&](auto nwc) {
            constexpr int NW = decltype(nwc)::value;
            cdist_list_kernel<NW, Sink><<<(unsigned)(device_cus() * 8), 256, 0, s>>>(
                d_buf_q, d_off_q, gpeq, choices->words, choices->gbase, choices->len, nc, sink, k, d_kvec,
                elist, elist_count, ecap, choices->order ? nullptr : choices->buf,
                choices->order ? nullptr : choices->off);
        };
        if (mode == 4 && ecap > 0) {
            switch (nwmax) {
            case 1: lgo(std::integral_constant<int, 1>()); break;
            case 2: lgo(std::integral_constant<int, 2>()); break;
            case 4: lgo(std::integral_constant<int, 4>()); break;
            default: lgo(std::integral_constant<int, 8>()); break;
            }
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = fail(BL_EHIP, "%s launch: %s", who, hipGetErrorString(e));
    }
    if (!d_work) (void)hipFreeAsync(ws, s);
    return rc;
}

static int cdist_check(const char *who, const uint8_t *d_buf_q, const int64_t *d_off_q, int64_t nq,
                       const bl_stripes *choices, int max_qlen)
{
    if (!choices) return fail(BL_EINVAL, "%s: null choices", who);
    if (nq < 0 || choices->n < 0 || max_qlen < 0) return fail(BL_EINVAL, "%s: negative size", who);
    if (nq > 0 && choices->n > 0 &&
        (!d_buf_q || !d_off_q || !choices->words || !choices->gbase || !choices->len))
        return fail(BL_EINVAL, "%s: null pointer", who);
    if (max_qlen > 512) return fail(BL_EINVAL, "%s: queries above 512 bytes go to bl_long", who);
    if (nq > 0x7fffffff) return fail(BL_EINVAL, "%s: too many queries", who);
    return BL_OK;
}

extern "C" {

int bl_cdist_tuned(const uint8_t *d_buf_q, const int64_t *d_off_q, int64_t nq,
                   const bl_stripes *choices, int max_qlen, void *d_out, int dtype, int k,
                   const int32_t *d_kvec, void *d_work, size_t work_bytes,
                   const bl_cdist_params *prm, void *stream)
{
    if (int rc = cdist_check("bl_cdist", d_buf_q, d_off_q, nq, choices, max_qlen)) return rc;
    if (nq == 0 || choices->n == 0) return BL_OK;
    if (dtype < BL_INT32 || dtype > BL_UINT8) return fail(BL_EINVAL, "bl_cdist: bad dtype %d", dtype);
    if (!d_out) return fail(BL_EINVAL, "bl_cdist: null pointer");
    DenseSink sink{d_out, dtype};
    return cdist_launch("bl_cdist", d_buf_q, d_off_q, nq, choices, max_qlen, clamp_k_dev(k), d_kvec,
                        d_work, work_bytes, prm, stream, sink, nullptr);
}

int bl_cdist_matches(const uint8_t *d_buf_q, const int64_t *d_off_q, int64_t nq,
                     const bl_stripes *choices, int max_qlen, int k, const int32_t *d_kvec,
                     int32_t *d_q, int64_t *d_c, int32_t *d_d, int64_t capacity,
                     unsigned long long *d_count, void *d_work, size_t work_bytes,
                     const bl_cdist_params *prm, void *stream)
{
    if (int rc = cdist_check("bl_cdist_matches", d_buf_q, d_off_q, nq, choices, max_qlen)) return rc;
    if (k < 0 && !d_kvec) return fail(BL_EINVAL, "bl_cdist_matches: needs max_dist >= 0");
    if (!d_count || capacity < 0 || (capacity > 0 && (!d_q || !d_c || !d_d)))
        return fail(BL_EINVAL, "bl_cdist_matches: null output or negative capacity");
    if (nq == 0 || choices->n == 0)
        return launch_clear(d_count, 1, (hipStream_t)stream);
    MatchSink sink{d_count, d_q, d_c, d_d, capacity, choices->n};
    return cdist_launch("bl_cdist_matches", d_buf_q, d_off_q, nq, choices, max_qlen, clamp_k_dev(k),
                        d_kvec, d_work, work_bytes, prm, stream, sink, d_count);
}

int bl_cdist(const uint8_t *d_buf_q, const int64_t *d_off_q, int64_t nq, const bl_stripes *choices,
             int max_qlen, void *d_out, int dtype, int k, void *d_work, size_t work_bytes,
             void *stream)
{
    return bl_cdist_tuned(d_buf_q, d_off_q, nq, choices, max_qlen, d_out, dtype, k, nullptr, d_work,
                          work_bytes, nullptr, stream);
}

}  /* extern "C" */

/* ------------------------------------------------ search without allocation */
/* The workspace of bl_cdist_noalloc: one array per device, held until the
 * process ends. In host code `scratch` only names the symbol; its device
 * address comes from hipGetSymbolAddress, once per device, with that device
 * current (noalloc_scratch). */
static __device__ __align__(256) unsigned char scratch[BL_NOALLOC_SCRATCH_BYTES];

static constexpr int NOALLOC_MAX_DEVICES = 64;
static void *noalloc_ptr[NOALLOC_MAX_DEVICES];
static std::mutex noalloc_lock;

/* The calling thread's current device's copy of `scratch` (its address
 * resolved on the first call on that device). */
static int noalloc_scratch(void **out)
{
    int dev = -1;
    BL_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return fail(BL_EINVAL, "bl_cdist_noalloc: device %d out of range", dev);
    std::lock_guard<std::mutex> g(noalloc_lock);
    if (!noalloc_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(BL_EHIP, "bl_cdist_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", dev,
                        hipGetErrorString(e));
        noalloc_ptr[dev] = p;
    }
    *out = noalloc_ptr[dev];
    return BL_OK;
}

extern "C" {

int bl_cdist_noalloc(const uint8_t *d_buf_q, const int64_t *d_off_q, int64_t nq,
                     const bl_stripes *choices, int max_qlen, void *d_out, int dtype, int k,
                     const int32_t *d_kvec, const bl_cdist_params *prm, void *stream)
{
    if (int rc = cdist_check("bl_cdist_noalloc", d_buf_q, d_off_q, nq, choices, max_qlen)) return rc;
    if (nq == 0 || choices->n == 0) return BL_OK;
    if (dtype < BL_INT32 || dtype > BL_UINT8)
        return fail(BL_EINVAL, "bl_cdist_noalloc: bad dtype %d", dtype);
    if (!d_out) return fail(BL_EINVAL, "bl_cdist_noalloc: null pointer");
    const size_t need = cdist_ws_core(nq, max_qlen);
    if (need > (size_t)BL_NOALLOC_SCRATCH_BYTES)
        return fail(BL_EINVAL,
                    "bl_cdist_noalloc: %lld queries of up to %d bytes need %zu bytes of scratch, above "
                    "the %d-byte static scratch (split the queries, or call bl_cdist_tuned)",
                    (long long)nq, max_qlen, need, (int)BL_NOALLOC_SCRATCH_BYTES);
    void *ws = nullptr;
    if (int rc = noalloc_scratch(&ws)) return rc;
    DenseSink sink{d_out, dtype};
    return cdist_launch("bl_cdist_noalloc", d_buf_q, d_off_q, nq, choices, max_qlen, clamp_k_dev(k),
                        d_kvec, ws, (size_t)BL_NOALLOC_SCRATCH_BYTES, prm, stream, sink, nullptr);
}

static int64_t pairs_chunk(int64_t n, const bl_pairs_params *prm)
{
    int64_t ch = prm && prm->reserved[0] > 0 ? prm->reserved[0] : 65536;
    const int64_t floor_ch = (n + 4095) / 4096;           /* at most 4096 chunks */
    if (ch < floor_ch) ch = floor_ch;
    return (ch + 1023) / 1024 * 1024;                     /* whole bucketing blocks */
}

struct PairsLayout {
    size_t head, key, sorted, scratch, total;
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
    L.head = (QUEUE_BYTES + (2 * nk + 4 * ((size_t)hc + 1)) * 8 + 255) & ~(size_t)255;
    L.key = L.head;
    L.sorted = (L.key + (size_t)n * sizeof(unsigned short) + 255) & ~(size_t)255;
    L.scratch = (L.sorted + (size_t)n * sizeof(int32_t) + 255) & ~(size_t)255;
    L.total = L.scratch + (max_len > 4096 ? (size_t)max_len * 2048 : 0) + 256;
    return L;
}

/* Enough for any launch shape: chunks are at least 1024 pairs and at most 4096. */
static int64_t max_chunks(int64_t n)
{
    const int64_t c = (n + 1023) / 1024;
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
    const bool direct = prm && prm->reserved[1] == 1;
    const int grab = prm && prm->reserved[2] > 0 ? prm->reserved[2] : 1;
    /* reserved[3]: bits 0-3 band tiles from this many units, bits 4-7 whole
     * waves for pairs of more units (0: 8), bit 8 64-row words */
    const int r3 = prm && prm->reserved[3] > 0 ? prm->reserved[3] : 0;
    const int band = r3 & 15;
    const int lu = (r3 >> 4) & 15;
    const int long_units = lu >= 1 && lu <= 8 ? lu : 8;
    const int unit64 = (r3 >> 8) & 1;

    char *w = (char *)d_work;
    if (!w) BL_HIP(hipMallocAsync((void **)&w, L.total, s));
    const size_t nk = (size_t)L.nchunks * PKEYS;
    unsigned long long *ctr = (unsigned long long *)w;
    unsigned long long *hist = (unsigned long long *)(w + QUEUE_BYTES);
    unsigned long long *cursor = hist + nk;
    unsigned long long *chunk_cnt = cursor + nk;
    unsigned long long *tile_cnt = chunk_cnt + (L.nchunks + 1);
    unsigned long long *chunk_base = tile_cnt + (L.nchunks + 1);
    unsigned long long *tile_base = chunk_base + (L.nchunks + 1);
    unsigned short *key = (unsigned short *)(w + L.key);
    int32_t *sorted = (int32_t *)(w + L.sorted);
    int8_t *scratch = max_len > 4096 ? (int8_t *)(w + L.scratch) : nullptr;
    const int64_t slot = scratch ? max_len : 0;

    int rc = BL_OK;
    int64_t blocks = (int64_t)device_cus() * wpc * 64 / wg;
    /* no more waves than tiles: 64 pairs a tile, but a whole wave for each
     * long pair when there can be any (so a handful of long pairs do not
     * queue on the few waves 64 pairs a tile would give them) */
    int64_t maxb = (n + 63) / 64 * 64 / wg + 1;
    if (!direct && (max_len > 256 || long_units < 8)) {
        const int64_t mb = (n + wg / 64 - 1) / (wg / 64) + 1;
        if (mb > maxb) maxb = mb;
    }
    if (blocks > maxb) blocks = maxb;
    if (scratch && blocks * (wg / 64) > 2048) blocks = 2048 / (wg / 64);
    if (direct) {
        /* the widest lane code the strings can need (shorter string <= max_len) */
        if (max_len <= 64)
            pairs_direct_kernel<1><<<(unsigned)blocks, wg, 0, s>>>(
                d_buf_a, d_off_a, d_buf_b, d_off_b, n, trim, k, d_kvec, d_out, dtype, scratch, slot);
        else if (max_len <= 128)
            pairs_direct_kernel<2><<<(unsigned)blocks, wg, 0, s>>>(
                d_buf_a, d_off_a, d_buf_b, d_off_b, n, trim, k, d_kvec, d_out, dtype, scratch, slot);
        else
            pairs_direct_kernel<4><<<(unsigned)blocks, wg, 0, s>>>(
                d_buf_a, d_off_a, d_buf_b, d_off_b, n, trim, k, d_kvec, d_out, dtype, scratch, slot);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = fail(BL_EHIP, "pairs_direct_kernel launch: %s", hipGetErrorString(e));
    } else {
        /* the counter and the histograms start at zero */
        rc = launch_clear(ctr, (int)(QUEUE_BYTES / 8 + nk), s);
        const unsigned cblocks = (unsigned)((n + 1023) / 1024);  /* 1024 pairs per block */
        if (rc == BL_OK) {
            pairs_classify_kernel<<<cblocks, 1024, 0, s>>>(d_off_a, d_off_b, n, d_out, dtype, k, d_kvec,
                                                          bw, band, long_units, unit64, L.chunk, key,
                                                          hist);
            pairs_chunk_scan_kernel<<<(unsigned)L.nchunks, 256, 0, s>>>(hist, cursor, chunk_cnt,
                                                                        tile_cnt);
            prefix_kernel<<<1, 1024, 0, s>>>(chunk_cnt, chunk_base, tile_cnt, tile_base, L.nchunks);
            pairs_scatter_kernel<<<cblocks, 1024, 0, s>>>(key, n, L.chunk, cursor, chunk_base, sorted);
            const int64_t cap = resident_blocks(pairs_batched_kernel, wg, 0);
            if (blocks > cap) blocks = cap;
            pairs_batched_kernel<<<(unsigned)blocks, wg, 0, s>>>(
                d_buf_a, d_off_a, d_buf_b, d_off_b, trim, k, d_kvec, sorted, key, hist, chunk_base,
                tile_base, L.nchunks, d_out, dtype, grab, ctr, scratch, slot);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) rc = fail(BL_EHIP, "pairs pipeline launch: %s", hipGetErrorString(e));
        }
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
