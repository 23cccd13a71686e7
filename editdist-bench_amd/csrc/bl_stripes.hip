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

/* One wave (a block of 64) per group of 64 slots; slot c holds choice
 * order[c] (c when order is null). When the group's strings lie together in
 * the packed buffer (no order) and span at most STAGE bytes, the wave first
 * copies that span into LDS with aligned 16-byte loads (a wave load reads
 * 1 KB of consecutive bytes); every lane then builds its stripe words from
 * LDS. Otherwise each lane reads its own string's dwords from global memory.
 * Either way the profiles are counted from the bytes the words are built
 * from, 32 bytes a step. The byte-class and 2-gram counts are bit-sliced in
 * registers: an 8 x 8 bit transpose gives the planes of bits 0-3 of the 32
 * bytes, and each of the 16 classes adds the popcount of the AND of its four
 * planes (or their complements) under the mask of valid positions; a 2-gram
 * class takes bits 1-2 of the byte and of the one before it (the planes
 * shifted by one position, the previous step's last byte shifted in). With
 * counts3 all three profiles go through per-lane u32 histograms in LDS
 * instead, counter x of lane l at h[x][l] (bank l whatever x), by adds that
 * return nothing. (The LDS histograms for the first two cost their LDS in
 * occupancy: 1.44 ms for 4.2M strings of 200-256 bytes against 0.56 ms
 * without profiles; reading every string three times with scattered loads
 * took 13.9 ms.) */
typedef u32 v4u __attribute__((ext_vector_type(4)));
constexpr int STAGE = 16384;                 /* bytes of a group staged in LDS */
constexpr int H1 = 0, H2 = 16, H3 = 32;      /* histogram rows: classes, 2-grams, 3-grams */

constexpr int STAGE_SPARE = 48;              /* read past the span, masked off */

static size_t stripes_build_lds(bool counts3)
{
    return (size_t)STAGE + STAGE_SPARE + (counts3 ? (size_t)96 * 64 * sizeof(u32) : 0);
}

__global__ __launch_bounds__(64) void stripes_build_kernel(const u8 *__restrict__ buf,
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
    extern __shared__ __align__(16) u32 lds[];
    u8 *stage = (u8 *)lds;                                      /* STAGE + STAGE_SPARE */
    u32(*h)[64] = (u32(*)[64])(stage + STAGE + STAGE_SPARE);     /* with counts3: 96 rows */
    const int64_t g = blockIdx.x;
    const int lane = (int)threadIdx.x;
    const int64_t c0 = g * 64, c1 = c0 + 64 < n ? c0 + 64 : n;
    const int64_t c = c0 + lane;
    const int rows = counts3 ? 96 : 0;
    const bool bits = !counts3 && (counts || counts2);
    for (int x = 0; x < rows; x++) h[x][lane] = 0;
    int ln = 0;
    int64_t o = 0;
    if (c < n) {
        const int64_t src = order ? (int64_t)order[c] : c;
        o = off[src];
        ln = (int)(off[src + 1] - o);
        len[c] = ln;
    }
    const u8 *s = buf + o;
    /* stage the group's bytes when they lie together: the aligned 16-byte
     * blocks from the one holding byte off[c0] up to the one holding byte
     * off[c1] - 1 (an aligned block never crosses a page, and each holds a
     * byte of the buffer); only when every lane's string lies in that span
     * (offsets that decrease, a caller error, take the direct reads) */
    bool staged = false;
    const u8 *pa = buf;
    if (!order) {
        const int64_t s0 = off[c0], s1 = off[c1];
        const u8 *p0 = buf + s0, *p1 = buf + s1;
        pa = (const u8 *)((uintptr_t)p0 & ~(uintptr_t)15);
        const bool inside = c >= n || (ln >= 0 && o >= s0 && o + ln <= s1);
        staged = p1 > p0 && p1 - pa <= STAGE && __all(inside ? 1 : 0);
        if (staged) {
            const int64_t span = p1 - pa;
#pragma unroll 16
            for (int64_t a = 16 * lane; a < span; a += 64 * 16)
                *(v4u *)(stage + a) = *(const v4u *)(pa + a);
        }
    }
    __syncthreads();
    const int64_t gb = gbase[g];
    const int64_t width = (gbase[g + 1] - gb) >> 6;
    const u8 *ls = stage + (s - pa);          /* (staged) the string in LDS */
    u32 b0 = 0, b1 = 0;                       /* the two bytes before the current one */
    u32 n1[16], n2[16];                       /* bit-sliced class and 2-gram counts */
#pragma unroll
    for (int x = 0; x < 16; x++) n1[x] = n2[x] = 0;
    u32 prev1 = 0, prev2 = 0;                 /* bit planes 1 and 2 of the previous step */
    /* eight dwords a step: their LDS reads issue together, one wait a step */
    for (int64_t j0 = 0; j0 < width; j0 += 8) {
        u32 w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (4 * j0 < ln) {
            if (staged) {
                /* nine aligned LDS dwords from the one holding byte 4 j0 (the
                 * stage starts 16-byte aligned; the reads end at most 35
                 * bytes past the staged span, inside the stage's spare
                 * bytes, and what passes the string is masked off) */
                const u8 *p = ls + 4 * j0;
                const u32 sh = (u32)((uintptr_t)p & 3);
                const u32 *al = (const u32 *)(p - sh);
                u32 d[9];
#pragma unroll
                for (int k = 0; k < 9; k++) d[k] = al[k];
#pragma unroll
                for (int k = 0; k < 8; k++) w[k] = align_bytes(d[k + 1], d[k], sh);
            } else {
#pragma unroll
                for (int k = 0; k < 8; k++) w[k] = j0 + k < width ? stripe_word(s, ln, j0 + k) : 0u;
            }
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int64_t r = ln - 4 * (j0 + k);
                w[k] = r <= 0 ? 0u : (r < 4 ? w[k] & ((1u << (8 * r)) - 1u) : w[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (j0 + k < width) words[gb + 64 * (j0 + k) + lane] = w[k];
        if (bits && 4 * j0 < ln) {
            /* planes of bits 0-3 of the 32 bytes: bit r of P[b] = bit b of
             * byte r (transpose8: byte b of y holds bit b of each byte of x) */
            u32 P[4] = {0, 0, 0, 0};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const u64 y = transpose8(((u64)w[2 * q + 1] << 32) | w[2 * q]);
#pragma unroll
                for (int b = 0; b < 4; b++) P[b] |= ((u32)(y >> (8 * b)) & 0xFFu) << (8 * q);
            }
            const int64_t rem = ln - 4 * j0;
            const u32 V = rem >= 32 ? ~0u : (1u << rem) - 1u;      /* bytes of the string */
            const u32 V2 = j0 == 0 ? V & ~1u : V;                  /* with a byte before */
            /* class x = b & 15: bits 0-1 of x from A, bits 2-3 from B */
            const u32 A[4] = {V & ~P[0] & ~P[1], V & P[0] & ~P[1], V & ~P[0] & P[1], V & P[0] & P[1]};
            const u32 B[4] = {~P[2] & ~P[3], P[2] & ~P[3], ~P[2] & P[3], P[2] & P[3]};
#pragma unroll
            for (int x = 0; x < 16; x++) n1[x] += (u32)popc(A[x & 3] & B[x >> 2]);
            /* 2-gram class 4 c2(previous) + c2(this), c2 = bits 1-2 */
            const u32 S1 = funnel32(P[1], prev1, 31), S2 = funnel32(P[2], prev2, 31);
            const u32 C[4] = {V2 & ~P[1] & ~P[2], V2 & P[1] & ~P[2], V2 & ~P[1] & P[2], V2 & P[1] & P[2]};
            const u32 D[4] = {~S1 & ~S2, S1 & ~S2, ~S1 & S2, S1 & S2};
#pragma unroll
            for (int x = 0; x < 16; x++) n2[x] += (u32)popc(C[x & 3] & D[x >> 2]);
            prev1 = P[1];
            prev2 = P[2];
        }
        if (rows && 4 * j0 < ln) {
#pragma unroll
            for (int k = 0; k < 8; k++) {
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int64_t at = 4 * (j0 + k) + i;
                    const u32 bb = (w[k] >> (8 * i)) & 0xFFu;
                    if (at < ln) {
                        atomicAdd(&h[H1 + (bb & 15u)][lane], 1u);
                        if (at >= 1) atomicAdd(&h[H2 + 4u * class2(b1) + class2(bb)][lane], 1u);
                        atomicAdd(&h[H3 + class3(b0, b1, bb)][lane], at >= 2 ? 1u : 0u);
                    }
                    b0 = b1;
                    b1 = bb;
                }
            }
        }
    }
    if (c < n && (bits || rows)) {
        auto pack16 = [&](auto get) {
            u32 v[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                u32 x = 0;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const u32 cnt = get(4 * q + i);
                    x |= (cnt > 255u ? 255u : cnt) << (8 * i);
                }
                v[q] = x;
            }
            return Profile{v[0], v[1], v[2], v[3]};
        };
        if (bits) {
            if (counts) counts[c] = pack16([&](int x) { return n1[x]; });
            if (counts2) counts2[c] = pack16([&](int x) { return n2[x]; });
        } else {
            if (counts) counts[c] = pack16([&](int x) { return h[H1 + x][lane]; });
            if (counts2) counts2[c] = pack16([&](int x) { return h[H2 + x][lane]; });
            counts3[c] = pack_profile3([&](int x) { return h[H3 + x][lane]; });
        }
    }
}

/* The stripes layout on the device, for offsets already in HBM (a host pass
 * over 4.2M offsets takes about 4 ms, longer than the search it prepares).
 * Three launches over blocks of LAYOUT_GROUPS groups: the widths and each
 * block's own inclusive prefix of them (part[b]: the block's total, or -1
 * for a decreasing offset), one block scanning the block totals in place
 * into exclusive prefixes (and part[nb] = the grand total, or -1), and the
 * prefixes added. gbase[0] = 0, gbase[g + 1] = sum of widths 0..g, a width
 * 64 x the group's widest string in dwords, at least 1: the host layout's
 * arithmetic. */
constexpr int LAYOUT_GROUPS = 256;

template <typename T>
__device__ T wave_scan_incl(T v, int lane)
{
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const T u = __shfl_up(v, d, 64);
        if (lane >= d) v += u;
    }
    return v;
}

__global__ __launch_bounds__(256) void layout_widths_kernel(const int64_t *__restrict__ off, int64_t n,
                                                            int64_t groups, int64_t *__restrict__ gbase,
                                                            int64_t *__restrict__ part)
{
    __shared__ int64_t wsum[4];
    __shared__ int wbad[4];
    const int lane = (int)(threadIdx.x & 63), wv = (int)(threadIdx.x >> 6);
    const int64_t g0 = (int64_t)blockIdx.x * LAYOUT_GROUPS + 64 * wv;
    int64_t widest = 0;                 /* lane i: group g0 + i */
    bool bad = false;
    /* eight groups a step: their sixteen offset loads issue together */
    for (int i0 = 0; i0 < 64 && g0 + i0 < groups; i0 += 8) {
        int64_t l[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int64_t c = 64 * (g0 + i0 + u) + lane;
            l[u] = g0 + i0 + u < groups && c < n ? off[c + 1] - off[c] : 0;
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            bad |= l[u] < 0;
            int64_t mx = l[u];
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) {
                const int64_t o = __shfl_xor(mx, d, 64);
                mx = o > mx ? o : mx;
            }
            if (lane == i0 + u) widest = mx;
        }
    }
    const int64_t g = g0 + lane;
    const int64_t w = g < groups ? 64 * (widest > 0 ? (widest + 3) / 4 : 1) : 0;
    const int64_t incl = wave_scan_incl(w, lane);
    if (lane == 63) wsum[wv] = incl;
    const bool any_bad = __any(bad);
    if (lane == 0) wbad[wv] = any_bad ? 1 : 0;
    __syncthreads();
    int64_t before = 0;
    for (int x = 0; x < wv; x++) before += wsum[x];
    if (g < groups) gbase[g + 1] = before + incl;
    if (threadIdx.x == 0) {
        const bool b = wbad[0] | wbad[1] | wbad[2] | wbad[3];
        part[blockIdx.x] = b ? -1 : wsum[0] + wsum[1] + wsum[2] + wsum[3];
    }
}

/* One block of 1024: thread t scans entries [t per, (t + 1) per) of part in
 * turn; a block scan of the thread sums gives each run its start. */
__global__ __launch_bounds__(1024) void layout_parts_kernel(int64_t *__restrict__ part, int64_t nb,
                                                            int64_t *__restrict__ gbase)
{
    __shared__ int64_t wsum[16];
    __shared__ int wbad[16];
    const int t = (int)threadIdx.x, lane = t & 63, wv = t >> 6;
    const int64_t per = (nb + 1023) / 1024;
    const int64_t a = t * per, e = a + per < nb ? a + per : nb;
    int64_t sum = 0;
    bool bad = false;
    for (int64_t i = a; i < e; i++) {
        const int64_t v = part[i];
        bad |= v < 0;
        sum += v;
    }
    const int64_t incl = wave_scan_incl(sum, lane);
    if (lane == 63) wsum[wv] = incl;
    const bool any_bad = __any(bad);
    if (lane == 0) wbad[wv] = any_bad ? 1 : 0;
    __syncthreads();
    int64_t run = incl - sum;
    for (int x = 0; x < wv; x++) run += wsum[x];
    for (int64_t i = a; i < e; i++) {
        const int64_t v = part[i];
        part[i] = run;
        run += v;
    }
    if (t == 0) {
        int64_t total = 0;
        bool b = false;
        for (int x = 0; x < 16; x++) {
            total += wsum[x];
            b |= wbad[x] != 0;
        }
        part[nb] = b ? -1 : total;
        gbase[0] = 0;
    }
}

__global__ __launch_bounds__(256) void layout_add_kernel(const int64_t *__restrict__ part, int64_t groups,
                                                         int64_t *__restrict__ gbase)
{
    const int64_t g = (int64_t)blockIdx.x * LAYOUT_GROUPS + threadIdx.x;
    if (g < groups) gbase[g + 1] += part[blockIdx.x];
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

size_t bl_stripes_layout_workspace_size(int64_t n)
{
    if (n < 0) return 0;
    const int64_t groups = (n + 63) / 64;
    const int64_t nb = (groups + LAYOUT_GROUPS - 1) / LAYOUT_GROUPS;
    return ((size_t)(nb + 1) * sizeof(int64_t) + 255) & ~(size_t)255;
}

int64_t bl_stripes_layout_device(const int64_t *d_off, int64_t n, int64_t *d_gbase, void *d_work,
                                 size_t work_bytes, void *stream)
{
    const char *who = "bl_stripes_layout_device";
    if (n < 0) return fail(BL_EINVAL, "%s: negative count", who);
    if (!d_off || !d_gbase || !d_work) return fail(BL_EINVAL, "%s: null pointer", who);
    if (work_bytes < bl_stripes_layout_workspace_size(n))
        return fail(BL_EINVAL, "%s: workspace of %zu bytes, %zu needed", who, work_bytes,
                    bl_stripes_layout_workspace_size(n));
    const int64_t groups = (n + 63) / 64;
    const int64_t nb = (groups + LAYOUT_GROUPS - 1) / LAYOUT_GROUPS;
    if (nb > 0x7fffffff) return fail(BL_EINVAL, "%s: too many strings", who);
    hipStream_t s = (hipStream_t)stream;
    int64_t *part = (int64_t *)d_work;
    if (nb > 0) {
        layout_widths_kernel<<<(unsigned)nb, 256, 0, s>>>(d_off, n, groups, d_gbase, part);
        BL_LAUNCHED("layout_widths_kernel");
    }
    layout_parts_kernel<<<1, 1024, 0, s>>>(part, nb, d_gbase);
    BL_LAUNCHED("layout_parts_kernel");
    if (nb > 0) {
        layout_add_kernel<<<(unsigned)nb, LAYOUT_GROUPS, 0, s>>>(part, groups, d_gbase);
        BL_LAUNCHED("layout_add_kernel");
    }
    int64_t total = 0;
    BL_HIP(hipMemcpyAsync(&total, part + nb, sizeof total, hipMemcpyDeviceToHost, s));
    BL_HIP(hipStreamSynchronize(s));
    if (total < 0) return fail(BL_EINVAL, "%s: offsets decrease", who);
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
    if (groups > 0x7fffffff) return fail(BL_EINVAL, "%s: too many strings", who);
    if (((uintptr_t)d_counts & 15) != 0 || ((uintptr_t)d_counts2 & 15) != 0 ||
        ((uintptr_t)d_counts3 & 15) != 0)
        return fail(BL_EINVAL, "%s: profiles not 16-byte aligned", who);
    const size_t lds = stripes_build_lds(d_counts3 != nullptr);
    stripes_build_kernel<<<(unsigned)groups, 64, lds, (hipStream_t)stream>>>(
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
