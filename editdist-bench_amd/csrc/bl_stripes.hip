/*
 * bl_stripes.hip -- choice sets laid out for the search kernels (gfx950).
 *
 * The search kernels give consecutive choices to consecutive lanes. In the
 * packed layout (csrc/edpy.c:200-224: one byte buffer + int64 offsets) the
 * lanes of a wave then read 64 different cache lines for every 4 text bytes,
 * and the vector memory pipeline, not the arithmetic, sets the pace. Stripes
 * interleave each group of 64 choices by dword:
 *
 *     dword j of choice c  at  words[gbase[c / 64] + 64 * j + c % 64]
 *
 * so a wave stepping its lanes' texts by one dword reads 256 contiguous bytes.
 * Every string of a group is zero-padded to the group's widest (at least one
 * dword), and len[c] keeps its length. Built once per resident choice set.
 *
 * A set may also be laid out in length order: within each block of `block`
 * slots the choices are sorted by length (stable), slot c holding choice
 * order[c]. The lanes of a wave then scan strings of about the same length
 * and finish together (an exact search costs m x n cells per pair, so a wave
 * of unsorted lengths 16..128 runs as long as its longest string), the
 * groups pad less, and the results still land in the caller's columns
 * (order[c]) within the block's range of each row.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "bl_device.h"
#include "bl_internal.h"

using namespace bl;

namespace {

/* One wave per group of 64 slots; slot c holds choice order[c] (c when order
 * is null). The 3-gram counts (counts3) go through a per-lane LDS histogram
 * of 64 u16 counters, class-major (counter x of thread t at h3[x][t]), fed
 * from the dwords as they are written. */
__global__ __launch_bounds__(256) void stripes_build_kernel(const u8 *__restrict__ buf,
                                                            const int64_t *__restrict__ off,
                                                            int64_t n,
                                                            const int32_t *__restrict__ order,
                                                            const int64_t *__restrict__ gbase,
                                                            u32 *__restrict__ words,
                                                            int32_t *__restrict__ len,
                                                            Profile *__restrict__ counts,
                                                            Profile *__restrict__ counts2,
                                                            Profile3 *__restrict__ counts3)
{
    extern __shared__ uint16_t h3s[];        /* 64 x 256 counters, only with counts3 */
    uint16_t(*h3)[256] = (uint16_t(*)[256])h3s;
    const int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = (int)(threadIdx.x & 63);
    const int t = (int)threadIdx.x;
    if (counts3)
        for (int x = 0; x < 64; x++) h3[x][t] = 0;
    if (g >= (n + 63) >> 6) return;
    const int64_t c = g * 64 + lane;
    const int64_t gb = gbase[g];
    const int64_t width = (gbase[g + 1] - gb) >> 6;
    int ln = 0;
    const u8 *s = buf;
    if (c < n) {
        const int64_t src = order ? (int64_t)order[c] : c;
        const int64_t o = off[src];
        ln = (int)(off[src + 1] - o);
        s = buf + o;
        len[c] = ln;
        if (counts) counts[c] = class_profile(s, ln);
        if (counts2) counts2[c] = class2_profile(s, ln);
    }
    if (!counts3) {
        for (int64_t j = 0; j < width; j++) words[gb + 64 * j + lane] = stripe_word(s, ln, j);
        return;
    }
    u32 b0 = 0, b1 = 0;                   /* the two bytes before the current one */
    for (int64_t j = 0; j < width; j++) {
        const u32 w = stripe_word(s, ln, j);
        words[gb + 64 * j + lane] = w;
        for (int i = 0; i < 4; i++) {
            const int64_t at = 4 * j + i;
            const u32 b2 = (w >> (8 * i)) & 0xFFu;
            if (at >= 2 && at < ln) {
                uint16_t &h = h3[class3(b0, b1, b2)][t];
                h = (uint16_t)(h + (h < 0xFFFFu ? 1u : 0u));
            }
            b0 = b1;
            b1 = b2;
        }
    }
    if (c < n) counts3[c] = pack_profile3([&](int x) { return h3[x][t]; });
}

/* dst[r][i] = src[r][inv[i]]: one thread per column, rows in turn, so every
 * row being gathered from stays in L2 and the writes are coalesced. */
template <typename T>
__global__ __launch_bounds__(256) void gather_cols_kernel(const T *__restrict__ src,
                                                          const int64_t *__restrict__ inv,
                                                          T *__restrict__ dst, int64_t rows,
                                                          int64_t cols)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cols) return;
    const int64_t s = inv[i];
    for (int64_t r = 0; r < rows; r++) dst[r * cols + i] = src[r * cols + s];
}

}  // namespace

extern "C" {

int bl_gather_cols(const void *d_src, const int64_t *d_inv, void *d_dst, int64_t rows, int64_t cols,
                   int dtype, void *stream)
{
    if (rows < 0 || cols < 0) return fail(BL_EINVAL, "bl_gather_cols: negative size");
    if (rows == 0 || cols == 0) return BL_OK;
    if (!d_src || !d_inv || !d_dst) return fail(BL_EINVAL, "bl_gather_cols: null pointer");
    if (d_src == d_dst) return fail(BL_EINVAL, "bl_gather_cols: source and destination overlap");
    const int64_t blocks = (cols + 255) / 256;
    if (blocks > 0x7fffffff) return fail(BL_EINVAL, "bl_gather_cols: too many columns");
    hipStream_t s = (hipStream_t)stream;
    switch (dtype) {
    case BL_INT32:
        gather_cols_kernel<int32_t><<<(unsigned)blocks, 256, 0, s>>>(
            (const int32_t *)d_src, d_inv, (int32_t *)d_dst, rows, cols);
        break;
    case BL_UINT16:
        gather_cols_kernel<uint16_t><<<(unsigned)blocks, 256, 0, s>>>(
            (const uint16_t *)d_src, d_inv, (uint16_t *)d_dst, rows, cols);
        break;
    case BL_UINT8:
        gather_cols_kernel<uint8_t><<<(unsigned)blocks, 256, 0, s>>>(
            (const uint8_t *)d_src, d_inv, (uint8_t *)d_dst, rows, cols);
        break;
    default:
        return fail(BL_EINVAL, "bl_gather_cols: bad dtype %d", dtype);
    }
    BL_LAUNCHED("gather_cols_kernel");
    return BL_OK;
}

int64_t bl_stripes_layout(const int64_t *h_off, int64_t n, int64_t *h_gbase)
{
    if (n < 0) return fail(BL_EINVAL, "bl_stripes_layout: negative count");
    if (!h_off || !h_gbase) return fail(BL_EINVAL, "bl_stripes_layout: null pointer");
    const int64_t total = stripes_layout(h_off, n, h_gbase);
    if (total < 0) return fail(BL_EINVAL, "bl_stripes_layout: offsets decrease");
    return total;
}

int64_t bl_stripes_layout_sorted(const int64_t *h_off, int64_t n, int64_t block, int32_t *h_order,
                                 int64_t *h_gbase)
{
    if (n < 0) return fail(BL_EINVAL, "bl_stripes_layout_sorted: negative count");
    if (!h_off || !h_order || !h_gbase) return fail(BL_EINVAL, "bl_stripes_layout_sorted: null pointer");
    if (block < 64 || block % 64 != 0)
        return fail(BL_EINVAL, "bl_stripes_layout_sorted: block %lld is not a positive multiple of 64",
                    (long long)block);
    if (n > 0x7fffffff) return fail(BL_EINVAL, "bl_stripes_layout_sorted: too many strings");
    for (int64_t c = 0; c < n; c++)
        if (h_off[c + 1] < h_off[c]) return fail(BL_EINVAL, "bl_stripes_layout_sorted: offsets decrease");
    auto len = [&](int64_t c) { return h_off[c + 1] - h_off[c]; };
    std::vector<int64_t> cnt;
    for (int64_t b0 = 0; b0 < n; b0 += block) {
        const int64_t e = b0 + block < n ? b0 + block : n;
        int64_t lo = len(b0), hi = lo;
        for (int64_t c = b0; c < e; c++) {
            const int64_t l = len(c);
            lo = l < lo ? l : lo;
            hi = l > hi ? l : hi;
        }
        if (hi - lo < 65536) {
            /* counting sort, stable: slots taken in string order within each length */
            cnt.assign((size_t)(hi - lo + 2), 0);
            for (int64_t c = b0; c < e; c++) cnt[(size_t)(len(c) - lo + 1)]++;
            for (size_t i = 1; i < cnt.size(); i++) cnt[i] += cnt[i - 1];
            for (int64_t c = b0; c < e; c++) h_order[b0 + cnt[(size_t)(len(c) - lo)]++] = (int32_t)c;
        } else {
            for (int64_t c = b0; c < e; c++) h_order[c] = (int32_t)c;
            std::stable_sort(h_order + b0, h_order + e,
                             [&](int32_t x, int32_t y) { return len(x) < len(y); });
        }
    }
    /* the group bases of the strings in slot order */
    std::vector<int64_t> soff((size_t)n + 1);
    soff[0] = 0;
    for (int64_t c = 0; c < n; c++) soff[(size_t)c + 1] = soff[(size_t)c] + len(h_order[c]);
    return stripes_layout(soff.data(), n, h_gbase);
}

static int stripes_build(const char *who, const uint8_t *d_buf, const int64_t *d_off, int64_t n,
                         const int32_t *d_order, const int64_t *d_gbase, uint32_t *d_words,
                         int32_t *d_len, uint32_t *d_counts, uint32_t *d_counts2, uint32_t *d_counts3,
                         void *stream)
{
    if (n < 0) return fail(BL_EINVAL, "%s: negative count", who);
    if (n == 0) return BL_OK;
    if (!d_buf || !d_off || !d_gbase || !d_words || !d_len) return fail(BL_EINVAL, "%s: null pointer", who);
    const int64_t groups = (n + 63) / 64;
    const int64_t blocks = (groups + 3) / 4;
    if (blocks > 0x7fffffff) return fail(BL_EINVAL, "%s: too many strings", who);
    if (((uintptr_t)d_counts & 15) != 0 || ((uintptr_t)d_counts2 & 15) != 0 ||
        ((uintptr_t)d_counts3 & 15) != 0)
        return fail(BL_EINVAL, "%s: profiles not 16-byte aligned", who);
    /* 32 KB of LDS per block whether or not the 3-gram histogram needs it:
     * every lane reads its own string with scattered dword loads, and more
     * resident waves only thrash L1 and L2 (4.2M strings of 200-256 bytes:
     * 35 ms without the cap, 13.6 ms with it) */
    const size_t lds = 64 * 256 * sizeof(uint16_t);
    stripes_build_kernel<<<(unsigned)blocks, 256, lds, (hipStream_t)stream>>>(
        d_buf, d_off, n, d_order, d_gbase, d_words, d_len, (Profile *)d_counts, (Profile *)d_counts2,
        (Profile3 *)d_counts3);
    BL_LAUNCHED("stripes_build_kernel");
    return BL_OK;
}

int bl_stripes_build(const uint8_t *d_buf, const int64_t *d_off, int64_t n, const int64_t *d_gbase,
                     uint32_t *d_words, int32_t *d_len, uint32_t *d_counts, uint32_t *d_counts2,
                     void *stream)
{
    return stripes_build("bl_stripes_build", d_buf, d_off, n, nullptr, d_gbase, d_words, d_len, d_counts,
                         d_counts2, nullptr, stream);
}

int bl_stripes_build_order(const uint8_t *d_buf, const int64_t *d_off, int64_t n, const int32_t *d_order,
                           const int64_t *d_gbase, uint32_t *d_words, int32_t *d_len, uint32_t *d_counts,
                           uint32_t *d_counts2, uint32_t *d_counts3, void *stream)
{
    return stripes_build("bl_stripes_build_order", d_buf, d_off, n, d_order, d_gbase, d_words, d_len,
                         d_counts, d_counts2, d_counts3, stream);
}

}  /* extern "C" */
