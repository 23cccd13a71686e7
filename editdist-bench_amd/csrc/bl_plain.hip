/*
 * bl_plain.hip -- the plain (drop-in) kernels for gfx950 and their C ABI.
 *
 *   pairs_class_kernel<NW>   one lane per pair, pattern bit-planes in VGPRs,
 *                            exact trim, up to NW 64-bit words (shorter <= 64*NW).
 *                            Reference: ed_pairs_c{CAP}, csrc/gpu_kernels.cu:6-32.
 *   cdist_class_kernel<NW>   one workgroup per (query, choice block); the query's
 *                            256 x NW Peq table (and its low-32-row copy) in LDS;
 *                            lanes stride over choices, read from the packed
 *                            buffer (bl_cdist_class, the reference's arguments)
 *                            or from a resident stripes layout
 *                            (bl_cdist_class_stripes).
 *                            Reference: ed_cross_w{NW}, csrc/gpu_kernels.cu:36-69.
 *   long_kernel              one wave per long (pattern, text), anti-diagonal
 *                            wavefront over 64-row words; replaces the reference's
 *                            CPU fallbacks (bitlev/gpu.py:97-101, :131-136).
 */
#include <hip/hip_runtime.h>

#include "bl_device.h"
#include "bl_internal.h"

using namespace bl;

namespace {

constexpr int THREADS = 256;
static_assert(THREADS == 256, "cdist_class_kernel fills its 256-entry tables one entry per thread");

static inline int clamp_k(int k) { return k < 0 ? -1 : (k > (1 << 30) ? (1 << 30) : k); }

template <int NW>
__global__ __launch_bounds__(THREADS) void pairs_class_kernel(
    const u8 *__restrict__ ba, const int64_t *__restrict__ oa, const u8 *__restrict__ bb,
    const int64_t *__restrict__ ob, const int32_t *__restrict__ idx, int64_t cnt, void *out,
    int dtype, int trim, int k)
{
    const int64_t t = (int64_t)blockIdx.x * THREADS + threadIdx.x;
    if (t >= cnt) return;
    const int64_t i = idx[t];
    const int64_t a0 = oa[i], b0 = ob[i];
    const int la = (int)(oa[i + 1] - a0), lb = (int)(ob[i + 1] - b0);
    /* A pair outside the class (caller error) gets -1 (int32) or the narrow
     * dtype's maximum, never a partial scan. */
    const int d = (la < lb ? la : lb) > 64 * NW
                      ? -1
                      : pair_lane<NW>(ba + a0, la, bb + b0, lb, trim, k);
    store_out(out, i, d, dtype);
}

/* Choices straight from the packed buffer (the reference's layout): each lane
 * reads its own string with aligned dword loads. */
struct PackedChoices {
    const u8 *buf;
    const int64_t *off;
    BL_HDM int len(int64_t c) const { return (int)(off[c + 1] - off[c]); }
    BL_HDM ByteText text(int64_t c) const { return ByteText{buf + off[c]}; }
};

/* Choices from a resident stripes layout: the lanes of a wave read 256
 * contiguous bytes per dword step. */
struct StripeChoices {
    const u32 *words;
    const int64_t *gbase;
    const int32_t *clen;
    BL_HDM int len(int64_t c) const { return clen[c]; }
    BL_HDM StripeText text(int64_t c) const { return StripeText{words + gbase[c >> 6] + (c & 63)}; }
};

/* The query's Peq table in LDS, word-major: word w of byte c at
 * t[w * STRIDE + c], so that one u64 read per word puts lanes whose bytes
 * are less than 32 apart on distinct banks (byte-major rows of NW words gave
 * bytes 64 / NW apart the same banks: 0.43 bank-conflict cycles per LDS
 * cycle for two-word queries), and the stride of 257 keeps two words of a
 * byte out of one ds_read2. Read through peq_lane's row-major indexing. */
template <int NW>
struct WordMajorPeq {
    static constexpr int STRIDE = 257;
    const u64 *t;
    BL_HDM u64 operator[](int i) const { return t[(i % NW) * STRIDE + i / NW]; }
};

template <int NW, typename Choices>
__global__ __launch_bounds__(THREADS) void cdist_class_kernel(
    const u8 *__restrict__ bq, const int64_t *__restrict__ oq, const int32_t *__restrict__ qidx,
    const Choices ch, int64_t nc, void *out, int dtype, int k)
{
    constexpr int STRIDE = WordMajorPeq<NW>::STRIDE;
    __shared__ u64 peq[STRIDE * NW];
    __shared__ u32 lo32[256];               /* rows 0..31: the 32-row stage's table */
    const int64_t q = qidx[blockIdx.x];
    const int64_t q0 = oq[q];
    const int m = (int)(oq[q + 1] - q0);
    const bool fits = m <= 64 * NW;         /* a longer query is a caller error */
    for (int x = threadIdx.x; x < STRIDE * NW; x += THREADS) peq[x] = 0;
    __syncthreads();
    for (int i = threadIdx.x; fits && i < m; i += THREADS)
        atomicOr((unsigned long long *)&peq[(i >> 6) * STRIDE + (int)bq[q0 + i]],
                 (unsigned long long)1 << (i & 63));
    __syncthreads();
    lo32[threadIdx.x] = (u32)peq[threadIdx.x];
    __syncthreads();
    const WordMajorPeq<NW> tab{peq};
    const int64_t stride = (int64_t)gridDim.y * THREADS;
    for (int64_t c = (int64_t)blockIdx.y * THREADS + threadIdx.x; c < nc; c += stride) {
        const int n = ch.len(c);
        const int d = !fits ? -1 : m == 0 ? empty_pattern(n, k) : peq_lane<NW, 0>(lo32, tab, m, ch.text(c), n, k);
        store_out(out, q * nc + c, d, dtype);
    }
}

__global__ __launch_bounds__(64) void long_kernel(
    int is_pair, const u8 *__restrict__ ba, const int64_t *__restrict__ oa,
    const u8 *__restrict__ bb, const int64_t *__restrict__ ob, const int32_t *__restrict__ idx,
    int64_t cnt, int64_t nc, void *out, int dtype, int trim, int k, int8_t *scratch,
    int64_t slot)
{
    const int lane = threadIdx.x;
    const int64_t items = is_pair ? cnt : cnt * nc;
    for (int64_t it = blockIdx.x; it < items; it += gridDim.x) {
        const u8 *p, *t;
        int m, n;
        int64_t o;
        if (is_pair) {
            const int64_t i = idx[it];
            p = ba + oa[i]; m = (int)(oa[i + 1] - oa[i]);
            t = bb + ob[i]; n = (int)(ob[i + 1] - ob[i]);
            o = i;
        } else {
            const int64_t q = idx[it / nc], c = it % nc;
            p = ba + oa[q]; m = (int)(oa[q + 1] - oa[q]);
            t = bb + ob[c]; n = (int)(ob[c + 1] - ob[c]);
            o = q * nc + c;
        }
        int r = long_prepare(p, m, t, n, is_pair, trim, k);
        /* Stripes need a scratch slot of n bytes; a max_len below n is a caller error. */
        if (r < 0 && m > 4096 && (scratch == nullptr || n > slot)) r = -1;
        else if (r < 0) {
            LongCtx C;
            C.p = p; C.m = m; C.t = t; C.n = n;
            C.W = (m + 63) >> 6;
            C.hbuf = scratch ? scratch + (int64_t)blockIdx.x * slot : nullptr;
            const int S = (C.W + 63) >> 6;
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
                __threadfence();            /* this stripe's deltas before the next reads them */
            }
            const int d = m + __shfl(L.delta, (C.W - 1) - 64 * (S - 1));
            r = (k >= 0 && d > k) ? k + 1 : d;
        }
        if (lane == 0) store_out(out, o, r, dtype);     /* -1: caller error marker */
    }
}

int check_common(const void *out, int dtype, const char *what)
{
    if (dtype < BL_INT32 || dtype > BL_UINT8) return fail(BL_EINVAL, "%s: bad dtype %d", what, dtype);
    if (!out) return fail(BL_EINVAL, "%s: null output", what);
    return BL_OK;
}

}  // namespace

extern "C" {

int bl_pairs_class(int cap, const uint8_t *d_buf_a, const int64_t *d_off_a,
                   const uint8_t *d_buf_b, const int64_t *d_off_b, const int32_t *d_idx,
                   int64_t cnt, void *d_out, int dtype, int trim, int k, void *stream)
{
    if (cnt < 0) return fail(BL_EINVAL, "bl_pairs_class: negative count");
    if (cnt == 0) return BL_OK;
    if (int rc = check_common(d_out, dtype, "bl_pairs_class")) return rc;
    if (!d_buf_a || !d_off_a || !d_buf_b || !d_off_b || !d_idx)
        return fail(BL_EINVAL, "bl_pairs_class: null input");
    k = clamp_k(k);
    const dim3 grid((unsigned)((cnt + THREADS - 1) / THREADS));
    hipStream_t s = (hipStream_t)stream;
    switch (cap) {
    case 16: case 32: case 64:
        pairs_class_kernel<1><<<grid, THREADS, 0, s>>>(d_buf_a, d_off_a, d_buf_b, d_off_b, d_idx,
                                                       cnt, d_out, dtype, trim, k);
        break;
    case 128:
        pairs_class_kernel<2><<<grid, THREADS, 0, s>>>(d_buf_a, d_off_a, d_buf_b, d_off_b, d_idx,
                                                       cnt, d_out, dtype, trim, k);
        break;
    case 256:
        pairs_class_kernel<4><<<grid, THREADS, 0, s>>>(d_buf_a, d_off_a, d_buf_b, d_off_b, d_idx,
                                                       cnt, d_out, dtype, trim, k);
        break;
    default:
        return fail(BL_EINVAL, "bl_pairs_class: cap must be 16, 32, 64, 128 or 256 (got %d)", cap);
    }
    BL_LAUNCHED("pairs_class_kernel");
    return BL_OK;
}

}  /* extern "C" */

namespace {

template <typename Choices>
int cdist_class_launch(const char *who, int nw, const uint8_t *d_buf_q, const int64_t *d_off_q,
                       const int32_t *d_qidx, int64_t nqc, const Choices &ch, int64_t nc,
                       void *d_out, int dtype, int k, void *stream)
{
    k = clamp_k(k);
    int64_t gy = (nc + THREADS - 1) / THREADS;
    if (gy > 64) gy = 64;
    const dim3 grid((unsigned)nqc, (unsigned)gy);
    hipStream_t s = (hipStream_t)stream;
    switch (nw) {
    case 1: cdist_class_kernel<1><<<grid, THREADS, 0, s>>>(d_buf_q, d_off_q, d_qidx, ch, nc, d_out, dtype, k); break;
    case 2: cdist_class_kernel<2><<<grid, THREADS, 0, s>>>(d_buf_q, d_off_q, d_qidx, ch, nc, d_out, dtype, k); break;
    case 4: cdist_class_kernel<4><<<grid, THREADS, 0, s>>>(d_buf_q, d_off_q, d_qidx, ch, nc, d_out, dtype, k); break;
    case 8: cdist_class_kernel<8><<<grid, THREADS, 0, s>>>(d_buf_q, d_off_q, d_qidx, ch, nc, d_out, dtype, k); break;
    default:
        return fail(BL_EINVAL, "%s: nw must be 1, 2, 4 or 8 (got %d)", who, nw);
    }
    BL_LAUNCHED("cdist_class_kernel");
    return BL_OK;
}

}  // namespace

extern "C" {

int bl_cdist_class(int nw, const uint8_t *d_buf_q, const int64_t *d_off_q,
                   const int32_t *d_qidx, int64_t nqc, const uint8_t *d_buf_c,
                   const int64_t *d_off_c, int64_t nc, void *d_out, int dtype, int k, void *stream)
{
    if (nqc < 0 || nc < 0) return fail(BL_EINVAL, "bl_cdist_class: negative count");
    if (nqc == 0 || nc == 0) return BL_OK;
    if (int rc = check_common(d_out, dtype, "bl_cdist_class")) return rc;
    if (!d_buf_q || !d_off_q || !d_qidx || !d_buf_c || !d_off_c)
        return fail(BL_EINVAL, "bl_cdist_class: null input");
    if (nqc > 0x7fffffff) return fail(BL_EINVAL, "bl_cdist_class: too many queries");
    return cdist_class_launch("bl_cdist_class", nw, d_buf_q, d_off_q, d_qidx, nqc,
                              PackedChoices{d_buf_c, d_off_c}, nc, d_out, dtype, k, stream);
}

int bl_cdist_class_stripes(int nw, const uint8_t *d_buf_q, const int64_t *d_off_q,
                           const int32_t *d_qidx, int64_t nqc, const bl_stripes *choices,
                           void *d_out, int dtype, int k, void *stream)
{
    if (!choices) return fail(BL_EINVAL, "bl_cdist_class_stripes: null choices");
    const int64_t nc = choices->n;
    if (nqc < 0 || nc < 0) return fail(BL_EINVAL, "bl_cdist_class_stripes: negative count");
    if (nqc == 0 || nc == 0) return BL_OK;
    if (int rc = check_common(d_out, dtype, "bl_cdist_class_stripes")) return rc;
    if (!d_buf_q || !d_off_q || !d_qidx || !choices->words || !choices->gbase || !choices->len)
        return fail(BL_EINVAL, "bl_cdist_class_stripes: null input");
    if (nqc > 0x7fffffff) return fail(BL_EINVAL, "bl_cdist_class_stripes: too many queries");
    if (choices->order)
        return fail(BL_EINVAL, "bl_cdist_class_stripes: choices in length order (use bl_cdist)");
    return cdist_class_launch("bl_cdist_class_stripes", nw, d_buf_q, d_off_q, d_qidx, nqc,
                              StripeChoices{choices->words, choices->gbase, choices->len}, nc,
                              d_out, dtype, k, stream);
}

int bl_long(int is_pair, const uint8_t *d_buf_a, const int64_t *d_off_a,
            const uint8_t *d_buf_b, const int64_t *d_off_b, const int32_t *d_idx, int64_t cnt,
            int64_t nc, void *d_out, int dtype, int trim, int k, int64_t max_len, void *stream)
{
    if (cnt < 0 || nc < 0 || max_len < 0) return fail(BL_EINVAL, "bl_long: negative size");
    const int64_t items = is_pair ? cnt : cnt * nc;
    if (items == 0) return BL_OK;
    if (int rc = check_common(d_out, dtype, "bl_long")) return rc;
    if (!d_buf_a || !d_off_a || !d_buf_b || !d_off_b || !d_idx)
        return fail(BL_EINVAL, "bl_long: null input");
    if (max_len > 0x7fffffff) return fail(BL_EINVAL, "bl_long: strings must be shorter than 2^31");
    k = clamp_k(k);
    hipStream_t s = (hipStream_t)stream;
    /* Patterns above 4096 rows need the bottom-row scratch: one slot per wave. */
    const bool striped = max_len > 4096;
    int64_t grid = items < 4096 ? items : 4096;
    if (striped && grid > 512) grid = 512;
    int8_t *scratch = nullptr;
    if (striped) BL_HIP(hipMallocAsync((void **)&scratch, (size_t)(grid * max_len), s));
    long_kernel<<<dim3((unsigned)grid), 64, 0, s>>>(is_pair, d_buf_a, d_off_a, d_buf_b, d_off_b,
                                                    d_idx, cnt, nc, d_out, dtype, trim, k,
                                                    scratch, striped ? max_len : 0);
    hipError_t e = hipGetLastError();
    if (scratch) (void)hipFreeAsync(scratch, s);
    if (e != hipSuccess) return fail(BL_EHIP, "long_kernel launch: %s", hipGetErrorString(e));
    return BL_OK;
}

}  /* extern "C" */
