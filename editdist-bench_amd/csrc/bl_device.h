/*
 * bl_device.h -- per-lane building blocks of the gfx950 edit-distance kernels.
 *
 * Everything here compiles both as HIP device code (hipcc, gfx950) and as plain
 * C++ (g++), so bl_hostcheck.cpp runs the very same per-lane algorithms on the
 * CPU against the oracle before any of it reaches a GPU.
 *
 * Reference behaviour followed (karims/editdist-bench):
 *   word_step          csrc/myers.c:22-36, csrc/gpu_myers.cuh:45-59 (same recurrence)
 *   trim_pair          csrc/myers.c:80-87, csrc/gpu_myers.cuh:62-68 (exact affix trim)
 *   pair_lane<NW>      csrc/gpu_myers.cuh:73-119  (pattern in registers, one lane per pair)
 *   peq_lane<NW>       csrc/gpu_myers.cuh:124-145 (pattern via a Peq table)
 *   store_out          csrc/gpu_myers.cuh:149-152, bitlev/gpu.py:64-68 (saturating dtypes)
 *
 * Design points that differ from the reference (results are identical):
 *   * Eq comes from eight bit-planes of the pattern, not byte-wise SWAR compares:
 *     Eq = AND_b ~(plane_b ^ broadcast(bit b of the text byte)), one v_bitop3 per
 *     plane per 32 pattern rows. All 64 rows of a word compare in 16 ops.
 *   * The score is never tracked per column. Row 0 of a global alignment is j, so
 *     D[m][j] = j + popcount(Pv & mask) - popcount(Mv & mask); the threshold exit is
 *     tested every 4 columns (a later exit never changes min(d, k + 1)).
 *   * Text bytes are read 4 at a time with aligned dword loads only, and only
 *     dwords that hold at least one byte of the string are ever touched, so no
 *     load can leave the caller's allocation whatever the string alignment.
 */
#pragma once
#include <stdint.h>
#include <string.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define BL_HD __host__ __device__ __forceinline__
#define BL_HDM __host__ __device__ __forceinline__
#else
#define BL_HD static inline
#define BL_HDM inline
#endif

namespace bl {

typedef uint8_t u8;
typedef uint32_t u32;
typedef uint64_t u64;

/* Output dtype codes of the C ABI. */
enum { DT_INT32 = 0, DT_UINT16 = 1, DT_UINT8 = 2 };

/* ------------------------------------------------------------ bit utilities */
/* v_bitop3_b32 truth tables: operand S0, S1, S2 stand for 0xF0, 0xCC, 0xAA. */
constexpr u32 TA = 0xF0, TB = 0xCC, TC = 0xAA;
constexpr u32 tt(u32 v) { return v & 0xFFu; }

template <u32 TT>
BL_HD u32 bop3(u32 a, u32 b, u32 c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, c, TT);
#else
    u32 r = 0;
    for (int i = 0; i < 8; i++)
        if ((TT >> i) & 1u)
            r |= ((i & 4) ? a : ~a) & ((i & 2) ? b : ~b) & ((i & 1) ? c : ~c);
    return r;
#endif
}

template <u32 TT>
BL_HD u64 bop3(u64 a, u64 b, u64 c)
{
    const u32 lo = bop3<TT>((u32)a, (u32)b, (u32)c);
    const u32 hi = bop3<TT>((u32)(a >> 32), (u32)(b >> 32), (u32)(c >> 32));
    return ((u64)hi << 32) | lo;
}

BL_HD int popc64(u64 x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __popcll(x);
#else
    return __builtin_popcountll(x);
#endif
}

BL_HD int ctz32(u32 x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_ctz(x);
#else
    return __builtin_ctz(x);
#endif
}

BL_HD int clz32(u32 x) { return __builtin_clz(x); }

/* v_alignbit_b32: the low 32 bits of (hi:lo) >> (sh & 31). */
BL_HD u32 funnel32(u32 hi, u32 lo, u32 sh)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_alignbit(hi, lo, sh);
#else
    return (u32)((((u64)hi << 32) | lo) >> (sh & 31));
#endif
}

/* bytes [sh, sh+4) of the little-endian 8-byte value hi:lo (v_alignbyte_b32). */
BL_HD u32 align_bytes(u32 hi, u32 lo, u32 sh)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_alignbyte(hi, lo, sh);
#else
    return (u32)((((u64)hi << 32) | lo) >> (8 * (sh & 3)));
#endif
}

/* 0 or ~0: bit b of c broadcast (v_bfe_i32 c, b, 1). */
BL_HD u32 bit_mask(u32 c, int b) { return (u32)((int32_t)(c << (31 - b)) >> 31); }

/* Aligned 4-byte load. */
BL_HD u32 ld32(const u8 *p)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return *(const u32 *)p;
#else
    u32 v;
    memcpy(&v, p, 4);
    return v;
#endif
}

/* s[i .. i+4) as a little-endian u32, 0 <= i < n. Bytes at or past n are
 * unspecified. Only aligned dwords holding a byte of s[0..n) are read.
 * (Pointer arithmetic, not integer casts, so the compiler keeps the global
 * address space and emits global loads.) */
BL_HD u32 load4(const u8 *s, int i, int n)
{
    const u8 *p = s + i;
    const u32 sh = (u32)((uintptr_t)p & 3);
    const u8 *al = p - sh;
    const u8 *e = s + (n - 1);
    const u8 *last = e - ((uintptr_t)e & 3);
    const u32 lo = ld32(al);
    const u32 hi = ld32(al + 4 <= last ? al + 4 : last);
    return align_bytes(hi, lo, sh);
}

/* Sequential reader of a string, 4 bytes per call, one aligned load each. */
struct ByteStream {
    const u8 *al, *last;
    u32 cur, sh;
    BL_HDM void open(const u8 *s, int n)
    {
        sh = (u32)((uintptr_t)s & 3);
        al = s - sh;
        const u8 *e = s + (n > 0 ? n - 1 : 0);
        last = e - ((uintptr_t)e & 3);
        cur = n > 0 ? ld32(al) : 0;
    }
    /* The first call returns bytes 0..3, the next 4..7, ... (the caller stops
     * at n; bytes past n are unspecified). */
    BL_HDM u32 next4()
    {
        const u8 *nxa = al + 4;
        const u32 nx = ld32(nxa <= last ? nxa : last);
        const u32 v = align_bytes(nx, cur, sh);
        cur = nx;
        al = nxa;
        return v;
    }
};

/* The same reader fetching 16 aligned bytes at a time (one dwordx4 instead of
 * four dword loads, so a lane asks its cache for each line a quarter as
 * often): the block's dwords are handed out one per call. Only aligned
 * 16-byte blocks that hold a byte of the string are read (such a block never
 * crosses a page, and the caller's buffer covers the byte). The band lanes
 * use it: their columns are cheap, so their text requests would otherwise
 * miss the L2 about four times as often; lanes of short or whole-pattern scans
 * gain nothing from it and pay its selects. */
struct BlockStream {
    const u8 *blk, *last;          /* next block to load; the block of the last byte */
    u32 b0, b1, b2, b3;            /* the current block */
    u32 cur, sh;
    int i;                         /* next dword of the block to hand out (4: load) */
    BL_HDM u32 pick() const { return (i & 2) ? ((i & 1) ? b3 : b2) : ((i & 1) ? b1 : b0); }
    BL_HDM void load()
    {
        const u8 *p = blk <= last ? blk : last;
#if defined(__HIP_DEVICE_COMPILE__)
        const uint4 v = *(const uint4 *)__builtin_assume_aligned(p, 16);
        b0 = v.x; b1 = v.y; b2 = v.z; b3 = v.w;
#else
        b0 = ld32(p); b1 = ld32(p + 4); b2 = ld32(p + 8); b3 = ld32(p + 12);
#endif
        blk += 16;
        i = 0;
    }
    BL_HDM void open(const u8 *s, int n)
    {
        sh = (u32)((uintptr_t)s & 3);
        blk = s - ((uintptr_t)s & 15);
        const u8 *e = s + (n > 0 ? n - 1 : 0);
        last = e - ((uintptr_t)e & 15);
        load();
        i = (int)(((uintptr_t)s >> 2) & 3);
        cur = pick();
        i += 1;
    }
    BL_HDM u32 next4()
    {
        if (i == 4) load();
        const u32 nx = pick();
        i += 1;
        const u32 v = align_bytes(nx, cur, sh);
        cur = nx;
        return v;
    }
};

struct BlockText {                 /* a string read 16 aligned bytes at a time */
    const u8 *t;
    BL_HDM BlockStream open(int n) const
    {
        BlockStream s;
        s.open(t, n);
        return s;
    }
};

/* Text sources. A scan may restart from column 0 (row stages), so it takes a
 * source and opens a fresh stream per stage. */
struct ByteText {                  /* a string in a packed byte buffer */
    const u8 *t;
    BL_HDM ByteStream open(int n) const
    {
        ByteStream s;
        s.open(t, n);
        return s;
    }
};

/* Choice-set stripes: the strings of group g (choices 64g .. 64g+63) are
 * interleaved by dword, dword j of choice c at words[gbase[g] + 64 * j + c % 64],
 * each string zero-padded to the widest of its group (gbase[g + 1] - gbase[g]
 * = 64 * that width in dwords). A wave whose lanes hold consecutive choices
 * reads 256 contiguous bytes per dword step. */
struct StripeStream {
    const u32 *p, *last;           /* next dword; the string's last dword */
    BL_HDM u32 next4()
    {
        const u32 v = *(p <= last ? p : last);
        p += 64;
        return v;
    }
};

/* Dword j of a string of ln bytes as the stripes hold it (zero past ln). */
BL_HD u32 stripe_word(const u8 *s, int ln, int64_t j)
{
    const int64_t b0 = 4 * j;
    if (b0 >= ln) return 0;
    u32 v = load4(s, (int)b0, ln);
    const int64_t r = ln - b0;
    if (r < 4) v &= (1u << (8 * r)) - 1u;
    return v;
}

/* Group bases of the stripes of n strings (host offsets off[0..n]) into
 * gbase[0..ceil(n/64)]: each group is 64 x its widest string in dwords (at
 * least 1). Returns the total dwords, or -1 for decreasing offsets. */
static inline int64_t stripes_layout(const int64_t *off, int64_t n, int64_t *gbase)
{
    const int64_t groups = (n + 63) / 64;
    int64_t acc = 0;
    for (int64_t g = 0; g < groups; g++) {
        int64_t widest = 0;
        const int64_t e = (g + 1) * 64 < n ? (g + 1) * 64 : n;
        for (int64_t c = g * 64; c < e; c++) {
            const int64_t l = off[c + 1] - off[c];
            if (l < 0) return -1;
            if (l > widest) widest = l;
        }
        gbase[g] = acc;
        acc += 64 * (widest > 0 ? (widest + 3) / 4 : 1);
    }
    gbase[groups] = acc;
    return acc;
}

struct StripeText {
    const u32 *first;              /* dword 0 of the string */
    BL_HDM StripeStream open(int n) const
    {
        StripeStream s;
        s.p = first;
        s.last = first + 64 * (n > 0 ? (n - 1) >> 2 : 0);
        return s;
    }
};

/* The same string read with four dwords (16 columns) in flight: for lanes
 * that each scan a different, scattered string (rescans of single choices),
 * where one dword ahead leaves every fourth column waiting on memory. The
 * window shifts through registers (no indexed array). */
struct DeepStripeStream {
    const u32 *p, *last;           /* next dword to load; the string's last dword */
    u32 b0, b1, b2, b3;
    BL_HDM u32 load()
    {
        const u32 v = *(p <= last ? p : last);
        p += 64;
        return v;
    }
    BL_HDM u32 next4()
    {
        const u32 v = b0;
        b0 = b1; b1 = b2; b2 = b3;
        b3 = load();
        return v;
    }
};

struct DeepStripeText {
    const u32 *first;
    BL_HDM DeepStripeStream open(int n) const
    {
        DeepStripeStream s;
        s.p = first;
        s.last = first + 64 * (n > 0 ? (n - 1) >> 2 : 0);
        s.b0 = s.load(); s.b1 = s.load(); s.b2 = s.load(); s.b3 = s.load();
        return s;
    }
};

/* ------------------------------------------------------------ output */
BL_HD void store_out(void *out, int64_t i, int d, int dtype)
{
    if (dtype == DT_UINT8) ((u8 *)out)[i] = (u8)(d > 255 ? 255 : d);
    else if (dtype == DT_UINT16) ((uint16_t *)out)[i] = (uint16_t)(d > 65535 ? 65535 : d);
    else ((int32_t *)out)[i] = d;
}

/* ------------------------------------------------------------ Myers column step */
/* Pattern rows live in blocks of 32 (u32) or 64 (u64) rows. */
template <typename T> struct Rows;
template <> struct Rows<u32> { static constexpr int n = 32; };
template <> struct Rows<u64> { static constexpr int n = 64; };

BL_HD int popc(u32 x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __popc(x);
#else
    return __builtin_popcount(x);
#endif
}
BL_HD int popc(u64 x) { return popc64(x); }

/* One text column on one block of pattern rows. hp / hn (0 or 1) are the +1 /
 * -1 horizontal delta entering the lowest row. Ph / Mh receive the deltas
 * leaving every row (bit i = row i) before they are shifted down; a following
 * block takes the top bit of them as its own hp / hn. Carries and shifts only
 * move towards higher rows, so the low rows of a block never depend on the
 * rows above them. */
template <typename T>
BL_HD void word_step(T &Pv, T &Mv, T Eq, u32 hp, u32 hn, T &Ph, T &Mh)
{
    const T Xv = Eq | Mv;
    const T E = Eq | (T)hn;
    const T S = (E & Pv) + Pv;
    const T Xh = bop3<tt((TA ^ TB) | TC)>(S, Pv, E);            /* (S ^ Pv) | E   */
    Ph = bop3<tt(TA | ~(TB | TC))>(Mv, Xh, Pv);                 /* Mv | ~(Xh | Pv) */
    Mh = Pv & Xh;
    /* (the shifted-in bit 0 is free, so "+" equals "|"; on 64-bit blocks it
     * maps to one v_lshl_add_u64) */
    const T Phs = (T)(Ph << 1) + (T)hp;
    const T Mhs = (T)(Mh << 1) + (T)hn;
    Pv = bop3<tt(TA | ~(TB | TC))>(Mhs, Xv, Phs);               /* Mhs | ~(Xv | Phs) */
    Mv = Phs & Xv;
}

/* word_step with the carries taken from the block below as the high dwords
 * of its Ph / Mh (bit 31: the carry; the lowest block gets 0x80000000 and 0,
 * row 0's +1), so that every shift is one v_alignbit_b32 per dword: (Ph << 1)
 * | carry without extracting the carry or shifting a 64-bit value. */
BL_HD u32 hi32(u32 x) { return x; }
BL_HD u32 hi32(u64 x) { return (u32)(x >> 32); }

template <typename T>
BL_HD T shift_in(T x, u32 below)
{
    if constexpr (sizeof(T) == 4) {
        return funnel32(x, below, 31);
    } else {
        const u32 lo = (u32)x, hi = (u32)(x >> 32);
        return ((u64)funnel32(hi, lo, 31) << 32) | funnel32(lo, below, 31);
    }
}

template <typename T>
BL_HD void word_step_f(T &Pv, T &Mv, T Eq, u32 ph_below, u32 mh_below, T &Ph, T &Mh)
{
    const T Xv = Eq | Mv;
    const T E = Eq | (T)(mh_below >> 31);
    const T S = (E & Pv) + Pv;
    const T Xh = bop3<tt((TA ^ TB) | TC)>(S, Pv, E);            /* (S ^ Pv) | E   */
    Ph = bop3<tt(TA | ~(TB | TC))>(Mv, Xh, Pv);                 /* Mv | ~(Xh | Pv) */
    Mh = Pv & Xh;
    const T Phs = shift_in<T>(Ph, ph_below);
    const T Mhs = shift_in<T>(Mh, mh_below);
    Pv = bop3<tt(TA | ~(TB | TC))>(Mhs, Xv, Phs);               /* Mhs | ~(Xv | Phs) */
    Mv = Phs & Xv;
}

/* Rows [0, r) of a block (r clamped to [0, rows of T]). */
template <typename T>
BL_HD T rows_below(int r)
{
    return r >= Rows<T>::n ? ~(T)0 : r <= 0 ? (T)0 : (T)(((T)1 << r) - 1);
}

/* ------------------------------------------------------------ bit-plane pattern */
/* 8x8 bit transpose: byte c of the result holds, in bit r, bit c of byte r. */
BL_HD u64 transpose8(u64 x)
{
    u64 t;
    t = (x ^ (x >> 7)) & 0x00AA00AA00AA00AAull;  x ^= t ^ (t << 7);
    t = (x ^ (x >> 14)) & 0x0000CCCC0000CCCCull; x ^= t ^ (t << 14);
    t = (x ^ (x >> 28)) & 0x00000000F0F0F0F0ull; x ^= t ^ (t << 28);
    return x;
}

/* Planes of one 64-row block: pl[b] bit i = bit b of p[i] (i < n; rows at or
 * past n hold unspecified bits, which never reach a row below them). */
BL_HD void build_planes(const u8 *p, int n, u64 (&pl)[8])
{
    u32 half[2][8];
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int b = 0; b < 8; b++) half[h][b] = 0;
#pragma unroll
    for (int g = 0; g < 8; g++) {                       /* 8 rows per group */
        if (8 * g < n) {
            const u32 lo = load4(p, 8 * g, n);
            const u32 hi = 8 * g + 4 < n ? load4(p, 8 * g + 4, n) : 0u;
            const u64 y = transpose8(((u64)hi << 32) | lo);
#pragma unroll
            for (int b = 0; b < 8; b++)
                half[g >> 2][b] |= (u32)((y >> (8 * b)) & 0xFFu) << (8 * (g & 3));
        }
    }
#pragma unroll
    for (int b = 0; b < 8; b++) pl[b] = ((u64)half[1][b] << 32) | half[0][b];
}

/* Per text byte: m[b] = bit b of c broadcast; m[0] is stored inverted. */
BL_HD void text_masks(u32 c, u32 (&m)[8])
{
#pragma unroll
    for (int b = 0; b < 8; b++) m[b] = bit_mask(c, b);
    m[0] = ~m[0];
}

/* Eq of one block: AND over b of ~(plane_b ^ m_b). */
BL_HD u64 eq_planes(const u64 (&pl)[8], const u32 (&m)[8])
{
    u32 lo = (u32)pl[0] ^ m[0], hi = (u32)(pl[0] >> 32) ^ m[0];
#pragma unroll
    for (int b = 1; b < 8; b++) {
        lo = bop3<tt(TA & ~(TB ^ TC))>(lo, (u32)pl[b], m[b]);
        hi = bop3<tt(TA & ~(TB ^ TC))>(hi, (u32)(pl[b] >> 32), m[b]);
    }
    return ((u64)hi << 32) | lo;
}

/* Eq of the low 32 rows of a block only. */
BL_HD u32 eq_planes_lo(const u64 (&pl)[8], const u32 (&m)[8])
{
    u32 lo = (u32)pl[0] ^ m[0];
#pragma unroll
    for (int b = 1; b < 8; b++) lo = bop3<tt(TA & ~(TB ^ TC))>(lo, (u32)pl[b], m[b]);
    return lo;
}

/* ------------------------------------------------------------ trimming */
/* d(xPy, xQy) == d(P, Q): drop the common prefix, then the common suffix of
 * what remains (never overlapping the prefix). 4 bytes per compare. */
BL_HD void trim_pair(const u8 *&a, int &la, const u8 *&b, int &lb)
{
    const int n = la < lb ? la : lb;
    int p = 0;
    while (p < n) {
        u32 x = load4(a, p, la) ^ load4(b, p, lb);
        const int r = n - p;
        if (r < 4) x &= (1u << (8 * r)) - 1u;
        if (x) { p += ctz32(x) >> 3; break; }
        p += 4;
    }
    if (p > n) p = n;
    const int lim = n - p;
    int s = 0;
    while (s < lim) {
        /* bytes a[la-4-s .. la-1-s] (top byte = a[la-1-s]) */
        const int ea = la - s, eb = lb - s;
        u32 xa = ea >= 4 ? load4(a, ea - 4, la) : load4(a, 0, la) << (8 * (4 - ea));
        u32 xb = eb >= 4 ? load4(b, eb - 4, lb) : load4(b, 0, lb) << (8 * (4 - eb));
        u32 x = xa ^ xb;
        const int r = lim - s;
        if (r < 4) x &= ~((1u << (8 * (4 - r))) - 1u);
        if (x) { s += clz32(x) >> 3; break; }
        s += 4;
    }
    if (s > lim) s = lim;
    a += p; b += p;
    la -= p + s; lb -= p + s;
}

/* ------------------------------------------------------------ exit bound */
/* After jj text columns, every alignment still has to cross column jj, so
 *   d >= min_i D[i][jj] + |(m - i) - (n - jj)|,
 * and because D changes by at most 1 per row the minimum sits on the final
 * diagonal i0 = jj + m - n: d >= D[i0][jj] (d >= n - m when i0 <= 0). With
 * D[0][jj] = jj, D[i0][jj] = jj + popcount(Pv & rows<i0) - popcount(Mv & rows<i0).
 * This dominates the reference's bottom-row test (score - (n-1-j) > k,
 * csrc/myers.c:117, :127), so exits come earlier and results are identical:
 * when the bound exceeds k the answer is k + 1 either way. At jj = n the
 * bound is the exact distance. Only rows below i0 are read, so the bound is
 * exact from any block state that holds those rows. */
template <typename T, int NW>
BL_HD int diag_bound(const T (&Pv)[NW], const T (&Mv)[NW], int nw, int jj, int m, int n)
{
    const int i0 = jj + m - n;
    if (i0 <= 0) return n - m;
    int s = jj;
#pragma unroll
    for (int w = 0; w < NW; w++) {
        if (w < nw) {
            const T mk = rows_below<T>(i0 - Rows<T>::n * w);
            s += popc((T)(Pv[w] & mk)) - popc((T)(Mv[w] & mk));
        }
    }
    return s;
}

/* One text column over the first nw blocks, Eq blocks given. */
template <typename T, int NW>
BL_HD void column(T (&Pv)[NW], T (&Mv)[NW], int nw, const T (&E)[NW])
{
    u32 ph = 0x80000000u, mh = 0;          /* row 0 of a global alignment: +1 per column */
#pragma unroll
    for (int w = 0; w < NW; w++) {
        if (NW == 1 || w < nw) {
            T Ph, Mh;
            word_step_f<T>(Pv[w], Mv[w], E[w], ph, mh, Ph, Mh);
            ph = hi32(Ph);
            mh = hi32(Mh);
        }
    }
}

/* Scan a text of n >= 1 bytes (from `text`) against the low R rows of a pattern of m rows,
 * held in nw <= NW blocks of type T (R = the rows those blocks cover); eq(c, w)
 * yields Eq of block w for text byte c. Four columns per aligned text load (the
 * next load is issued before the current columns run), the exit bound tested
 * after each group of four.
 *   R >= m: the whole scan. k < 0: exact distance; k >= 0: min(d, k + 1).
 *   R <  m (k >= 0 only): the low R rows evolve exactly as in the whole scan,
 *          so the bound is exact while its diagonal row jj + m - n stays within
 *          R. Returns k + 1 when it exceeds k by then, else -1 (undecided: the
 *          caller scans again with more rows).
 * EAGER: fetch the Eq blocks of all four columns before stepping (table
 * lookups: the four loads overlap); otherwise compute each column's Eq just
 * before its step (register bit-planes: nothing to overlap, fewer live VGPRs). */
template <typename T, int NW, bool EAGER, typename Text, typename EqFn>
BL_HD int scan_text(int m, int nw, int R, const Text &text, int n, int k, EqFn eq)
{
    T Pv[NW], Mv[NW];
#pragma unroll
    for (int w = 0; w < NW; w++) { Pv[w] = ~(T)0; Mv[w] = 0; }
    const bool whole = R >= m;
    auto ts = text.open(n);
    u32 c4 = ts.next4();
    int j = 0;
    for (;;) {
        /* a partial scan stops before a test it could not evaluate (it never
         * reaches j = n: that test would need all m rows) */
        if (!whole && j + 4 + m - n > R) return -1;
        const u32 c4n = ts.next4();
        const int cnt = n - j < 4 ? n - j : 4;
        if (cnt == 4) {
            if (EAGER) {
                T E[4][NW];
#pragma unroll
                for (int u = 0; u < 4; u++)
#pragma unroll
                    for (int w = 0; w < NW; w++) E[u][w] = eq((c4 >> (8 * u)) & 0xFFu, w);
#pragma unroll
                for (int u = 0; u < 4; u++) column<T, NW>(Pv, Mv, nw, E[u]);
            } else {
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    T E[NW];
#pragma unroll
                    for (int w = 0; w < NW; w++) E[w] = eq((c4 >> (8 * u)) & 0xFFu, w);
                    column<T, NW>(Pv, Mv, nw, E);
                }
            }
        } else {
#pragma unroll
            for (int u = 0; u < 3; u++) {
                if (u < cnt) {
                    T E[NW];
#pragma unroll
                    for (int w = 0; w < NW; w++) E[w] = eq((c4 >> (8 * u)) & 0xFFu, w);
                    column<T, NW>(Pv, Mv, nw, E);
                }
            }
        }
        j += cnt;
        if (j >= n) {
            const int d = diag_bound<T, NW>(Pv, Mv, nw, j, m, n);  /* exact at j = n */
            return (k >= 0 && d > k) ? k + 1 : d;
        }
        if (k >= 0 && diag_bound<T, NW>(Pv, Mv, nw, j, m, n) > k) return k + 1;
        c4 = c4n;
    }
}

/* ------------------------------------------------------------ row stages */
/* With a threshold most pairs are decided within a few columns, when the exit
 * bound needs only the rows near the top-left corner. The lanes therefore scan
 * in stages of growing height and stop at the first stage that decides:
 *   32 rows (one u32 block) -> 64 rows (one u64 block) -> all rows.
 * Each stage starts again from column 0; a stage that decides returns exactly
 * what the whole scan returns (k + 1), and the last stage is the whole scan.
 * Without a threshold only the last stage runs; a pattern of at most 32 rows
 * runs entirely on u32 blocks. */

/* ------------------------------------------------------------ pair lane */
/* Distance of one pair whose shorter string (after the optional trim) has at
 * most 64*NW bytes: the pattern lives in registers as bit-planes.
 * k < 0: exact; k >= 0: min(d, k + 1). Semantics of pair_distance_swar,
 * csrc/gpu_myers.cuh:73-119, and pair_distance_k, csrc/myers.c:132-143. */
template <int NW>
BL_HD int pair_lane(const u8 *a, int la, const u8 *b, int lb, int trim, int k)
{
    if (trim) trim_pair(a, la, b, lb);
    if (la > lb) {
        const u8 *t = a; a = b; b = t;
        const int x = la; la = lb; lb = x;
    }
    if (k >= 0 && lb - la > k) return k + 1;
    if (la == 0) return lb;

    const ByteText text{b};
    u64 pl[NW][8];
    build_planes(a, la, pl[0]);
    auto eq32 = [&](u32 c, int) {
        u32 m8[8];
        text_masks(c, m8);
        return eq_planes_lo(pl[0], m8);
    };
    if (la <= 32) return scan_text<u32, 1, false>(la, 1, 32, text, lb, k, eq32);
    if (k >= 0) {
        const int r = scan_text<u32, 1, false>(la, 1, 32, text, lb, k, eq32);
        if (r >= 0) return r;
    }
    auto eq64 = [&](u32 c, int w) {
        u32 m8[8];
        text_masks(c, m8);
        return eq_planes(pl[w], m8);
    };
    if (NW == 1 || la <= 64) return scan_text<u64, 1, false>(la, 1, 64, text, lb, k, eq64);
    if (k >= 0) {
        const int r = scan_text<u64, 1, false>(la, 1, 64, text, lb, k, eq64);
        if (r >= 0) return r;
    }
#pragma unroll
    for (int w = 1; w < NW; w++) {
#pragma unroll
        for (int b8 = 0; b8 < 8; b8++) pl[w][b8] = 0;
        if (64 * w < la) build_planes(a + 64 * w, la - 64 * w, pl[w]);
    }
    const int nw = (la + 63) >> 6;
    return scan_text<u64, NW, false>(la, nw, 64 * nw, text, lb, k, eq64);
}

/* ------------------------------------------------------------ 32-row units */
/* The whole scan over NW64 64-row blocks plus, with TAIL, one 32-row block on
 * top of them (rows 64*NW64 .. 64*NW64 + 31), so a pattern of 32*(2*NW64 + 1)
 * rows costs no more than its 32-row units. eq64(c, w) and eq32(c) give the
 * Eq blocks of text byte c. Results as scan_text with R >= m. */
template <int NW64, bool TAIL, typename Text, typename Eq64, typename Eq32>
BL_HD int scan_mixed(int m, const Text &text, int n, int k, Eq64 eq64, Eq32 eq32)
{
    static_assert(NW64 >= 1, "at least one 64-row block");
    u64 Pv[NW64], Mv[NW64];
#pragma unroll
    for (int w = 0; w < NW64; w++) { Pv[w] = ~(u64)0; Mv[w] = 0; }
    u32 Pt = ~0u, Mt = 0u;
    auto col = [&](u32 c) {
        u32 ph = 0x80000000u, mh = 0;
#pragma unroll
        for (int w = 0; w < NW64; w++) {
            u64 Ph, Mh;
            word_step_f<u64>(Pv[w], Mv[w], eq64(c, w), ph, mh, Ph, Mh);
            ph = hi32(Ph);
            mh = hi32(Mh);
        }
        if (TAIL) {
            u32 Ph, Mh;
            word_step_f<u32>(Pt, Mt, eq32(c), ph, mh, Ph, Mh);
        }
    };
    auto bound = [&](int jj) {
        const int i0 = jj + m - n;
        if (i0 <= 0) return n - m;
        int s = jj;
#pragma unroll
        for (int w = 0; w < NW64; w++) {
            const u64 mk = rows_below<u64>(i0 - 64 * w);
            s += popc((u64)(Pv[w] & mk)) - popc((u64)(Mv[w] & mk));
        }
        if (TAIL) {
            const u32 mk = rows_below<u32>(i0 - 64 * NW64);
            s += popc((u32)(Pt & mk)) - popc((u32)(Mt & mk));
        }
        return s;
    };
    auto ts = text.open(n);
    u32 c4 = ts.next4();
    int j = 0;
    for (;;) {
        const u32 c4n = ts.next4();
        const int cnt = n - j < 4 ? n - j : 4;
        if (cnt == 4) {
#pragma unroll
            for (int u = 0; u < 4; u++) col((c4 >> (8 * u)) & 0xFFu);
        } else {
#pragma unroll
            for (int u = 0; u < 3; u++)
                if (u < cnt) col((c4 >> (8 * u)) & 0xFFu);
        }
        j += cnt;
        if (j >= n) {
            const int d = bound(j);                     /* exact at j = n */
            return (k >= 0 && d > k) ? k + 1 : d;
        }
        if (k >= 0 && bound(j) > k) return k + 1;
        c4 = c4n;
    }
}

/* pair_lane for a shorter string (before trimming) of at most 32*W32 bytes,
 * scanned in 32-row units: the same stages, then the whole scan over W32/2
 * 64-row blocks and, for odd W32, a 32-row tail block. */
template <int W32>
BL_HD int pair_lane_w32(const u8 *a, int la, const u8 *b, int lb, int trim, int k)
{
    constexpr int NW64 = W32 / 2 > 0 ? W32 / 2 : 1;
    constexpr bool TAIL = (W32 & 1) && W32 > 1;
    constexpr int NPL = NW64 + (TAIL ? 1 : 0);
    if (trim) trim_pair(a, la, b, lb);
    if (la > lb) {
        const u8 *t = a; a = b; b = t;
        const int x = la; la = lb; lb = x;
    }
    if (k >= 0 && lb - la > k) return k + 1;
    if (la == 0) return lb;

    const ByteText text{b};
    u64 pl[NPL][8];
    build_planes(a, la, pl[0]);
    auto eq32 = [&](u32 c, int) {
        u32 m8[8];
        text_masks(c, m8);
        return eq_planes_lo(pl[0], m8);
    };
    if (W32 == 1 || la <= 32) return scan_text<u32, 1, false>(la, 1, 32, text, lb, k, eq32);
    if (k >= 0) {
        const int r = scan_text<u32, 1, false>(la, 1, 32, text, lb, k, eq32);
        if (r >= 0) return r;
    }
    auto eq64 = [&](u32 c, int w) {
        u32 m8[8];
        text_masks(c, m8);
        return eq_planes(pl[w], m8);
    };
    if (W32 <= 2 || la <= 64) return scan_text<u64, 1, false>(la, 1, 64, text, lb, k, eq64);
    if (k >= 0) {
        const int r = scan_text<u64, 1, false>(la, 1, 64, text, lb, k, eq64);
        if (r >= 0) return r;
    }
#pragma unroll
    for (int w = 1; w < NPL; w++) {
#pragma unroll
        for (int b8 = 0; b8 < 8; b8++) pl[w][b8] = 0;
        if (64 * w < la) build_planes(a + 64 * w, la - 64 * w, pl[w]);
    }
    auto eqt = [&](u32 c) {
        u32 m8[8];
        text_masks(c, m8);
        return eq_planes_lo(pl[NPL - 1], m8);
    };
    return scan_mixed<NW64, TAIL>(la, text, lb, k, eq64, eqt);
}

/* ------------------------------------------------------------ diagonal band */
/* An alignment of cost <= k passes only through cells (i, j) with
 *     |i - j| + |(m - n) - (i - j)| <= k
 * (it needs |i - j| edits to reach the cell and |(m - n) - (i - j)| more to
 * leave it), i.e. through the diagonals i - j in [min(0, dl) - e, max(0, dl) + e],
 * dl = m - n, e = floor((k - |dl|) / 2): at most k + 1 diagonals (Ukkonen's band).
 *
 * A band scan keeps WIN pattern rows (one T word) in a window that slides down
 * one row per text column, covering the diagonals [d0, d0 + WIN - 2] of every
 * column plus one row below them. Values outside the window are replaced by
 * upper bounds (+1 horizontal delta above it, +1 vertical delta below it, the
 * same conventions as the DP's own first row and column), so every value the
 * scan holds is >= the true one, and a cell whose optimal path stays inside
 * the window is exact. With the band of k inside the window that covers an
 * optimal alignment whenever d <= k: the scan returns d exactly then, and a
 * value > k otherwise; the final-diagonal exit bound stays valid as well (its
 * cell is on the final diagonal, inside the band). Cost per column: one T
 * step plus Eq of WIN / 32 + 1 plane words, whatever m is. */


/* The largest threshold a window of WIN rows at diagonal offset d0 decides for
 * a pair with m - n = dl (-1: none). */
template <int WIN>
BL_HD int band_reach(int dl, int d0)
{
    const int lo = dl < 0 ? dl : 0, hi = dl > 0 ? dl : 0;
    const int e1 = lo - d0, e2 = d0 + WIN - 2 - hi;
    const int e = e1 < e2 ? e1 : e2;
    return e < 0 ? -1 : (dl < 0 ? -dl : dl) + 2 * e + 1;
}

/* The offset that maximises the least reach over dl in [dl_min, dl_max]. */
template <int WIN>
BL_HD int band_offset(int dl_min, int dl_max)
{
    const int t = (dl_min < 0 ? dl_min : 0) + (dl_max > 0 ? dl_max : 0) - (WIN - 2);
    return t >= 0 ? t / 2 : -((1 - t) / 2);          /* floor(t / 2) */
}

/* Raw bytes of pattern rows [32q, 32q + 32) as 8 dwords (bytes at or past m
 * unspecified; none at all when 32q >= m): 9 aligned dword loads, all within
 * the dwords that hold a byte of p[0 .. m). (Aligned 16-byte loads need
 * selects to undo each lane's own misalignment, which cost more than the
 * loads they save here.) */
BL_HD void word_bytes(const u8 *p, int m, int q, u32 (&raw)[8])
{
    if (32 * q >= m) {
#pragma unroll
        for (int g = 0; g < 8; g++) raw[g] = 0;
        return;
    }
    const u8 *s0 = p + 32 * q;
    const u32 sh = (u32)((uintptr_t)s0 & 3);
    const u8 *al = s0 - sh;
    const u8 *e = p + (m - 1);
    const u8 *last = e - ((uintptr_t)e & 3);
    u32 d[9];
#pragma unroll
    for (int g = 0; g < 9; g++) d[g] = al + 4 * g <= last ? ld32(al + 4 * g) : 0u;
#pragma unroll
    for (int g = 0; g < 8; g++) raw[g] = align_bytes(d[g + 1], d[g], sh);
}

/* Bit-planes of those rows: bit r of pw[b] = bit b of row 32q + r (rows at or
 * past m hold unspecified bits, which never reach the rows below them). */
BL_HD void word_planes(const u32 (&raw)[8], u32 (&pw)[8])
{
#pragma unroll
    for (int b = 0; b < 8; b++) pw[b] = 0;
#pragma unroll
    for (int g = 0; g < 4; g++) {
        const u64 y = transpose8(((u64)raw[2 * g + 1] << 32) | raw[2 * g]);
#pragma unroll
        for (int b = 0; b < 8; b++) pw[b] |= (u32)((y >> (8 * b)) & 0xFFu) << (8 * g);
    }
}

/* Eq of one plane word: AND over b of ~(plane_b ^ m_b) (m[0] inverted). */
BL_HD u32 eq_word(const u32 (&pw)[8], const u32 (&m)[8])
{
    u32 e = pw[0] ^ m[0];
#pragma unroll
    for (int b = 1; b < 8; b++) e = bop3<tt(TA & ~(TB ^ TC))>(e, pw[b], m[b]);
    return e;
}

/* Band scan of pattern p (m >= 1 rows, bytes in memory) against a text of
 * n >= 1 bytes, window T (32 or 64 rows) at diagonal offset d0, for a threshold
 * kq <= band_reach<WIN>(m - n, d0): returns d if d <= kq, else kq + 1. early:
 * test the final-diagonal bound every 4 columns and stop once it exceeds kq.
 * The window's plane words are built from the pattern bytes as it slides (the
 * bytes of the next word are loaded one word ahead). Groups of four columns
 * that neither start sliding nor enter a new plane word -- all but about one
 * in eight -- run without those checks; the rest go column by column. */
template <typename T, typename Text>
BL_HD int band_scan(const u8 *p, int m, const Text &text, int n, int d0, int kq, bool early)
{
    constexpr int WIN = Rows<T>::n;
    constexpr int NE = WIN / 32 + 1;             /* plane words under one window */
    u32 cur[NE][8], raw[8];
#pragma unroll
    for (int w = 0; w < NE; w++) {
        word_bytes(p, m, w, raw);
        word_planes(raw, cur[w]);
    }
    word_bytes(p, m, NE, raw);
    T Pv = ~(T)0, Mv = 0;
    int V = 0;                /* value of the window's base row s at column j */
    int s = 0, q = 0;         /* window = rows s+1 .. s+WIN; cur[0] = plane word q */
    int j = 0;
    /* Eq of text byte c on the current window */
    auto eqwin = [&](u32 c) {
        u32 m8[8];
        text_masks(c, m8);
        u32 E[NE];
#pragma unroll
        for (int w = 0; w < NE; w++) E[w] = eq_word(cur[w], m8);
        const u32 sh = (u32)s;
        if (WIN == 32) return (T)funnel32(E[1], E[0], sh);
        return (T)(((u64)funnel32(E[NE - 1], E[1], sh) << 32) | funnel32(E[1], E[0], sh));
    };
    /* one column (byte c), the window staying put */
    auto step = [&](u32 c) {
        T Ph, Mh;
        word_step_f<T>(Pv, Mv, eqwin(c), 0x80000000u, 0u, Ph, Mh);
        V += 1;
    };
    /* one column, then the window one row down. The new deltas shifted down a
     * row undo the step's own up-shift of Ph / Mh, and the row that leaves the
     * window always has Pv = 0, Mv = Xv bit 0 (hp = 1 enters below it), so:
     *   Pv' = Mh | ~((Xv >> 1) | Ph) | top,  Mv' = Ph & (Xv >> 1),
     *   V' = V + 1 - (Xv & 1). */
    auto step_slide = [&](u32 c) {
        const T Eq = eqwin(c);
        const T Xv = Eq | Mv;
        const T S = (T)((Eq & Pv) + Pv);
        const T Xh = bop3<tt((TA ^ TB) | TC)>(S, Pv, Eq);          /* (S ^ Pv) | Eq */
        const T Ph = bop3<tt(TA | ~(TB | TC))>(Mv, Xh, Pv);        /* Mv | ~(Xh | Pv) */
        const T Mh = Pv & Xh;
        const T X1 = (T)(Xv >> 1);
        V += 1 - (int)(u32)(Xv & 1u);
        Pv = bop3<tt(TA | ~(TB | TC))>(Mh, X1, Ph) | (T)((T)1 << (WIN - 1));
        Mv = Ph & X1;
        s += 1;
    };
    /* the window entered the next plane word */
    auto refresh = [&]() {
#pragma unroll
        for (int w = 0; w + 1 < NE; w++)
#pragma unroll
            for (int b = 0; b < 8; b++) cur[w][b] = cur[w + 1][b];
        word_planes(raw, cur[NE - 1]);
        q += 1;
        word_bytes(p, m, q + NE, raw);
    };
    auto value = [&](int i) {                    /* s <= i <= s + WIN */
        const T mk = rows_below<T>(i - s);
        return V + popc((T)(Pv & mk)) - popc((T)(Mv & mk));
    };
    auto over = [&]() {                          /* early exit: bound above kq */
        const int i0 = j + m - n;
        return (i0 <= 0 ? n - m : value(i0)) > kq;
    };
    /* text dwords two groups ahead: a band column costs few instructions, so
     * one group does not cover a memory round trip */
    auto ts = text.open(n);
    u32 c4 = ts.next4(), c4n = ts.next4();
    for (;;) {
        /* The two inner loops leave the plane words alone, so nothing but the
         * scan state is carried around them. */
        while (n - j >= 4 && j + 4 + d0 - 1 <= 0) {          /* no slide in these 4 */
            const u32 c4nn = ts.next4();
#pragma unroll
            for (int u = 0; u < 4; u++) step((c4 >> (8 * u)) & 0xFFu);
            j += 4;
            if (early && over()) return kq + 1;
            c4 = c4n;
            c4n = c4nn;
        }
        while (n - j >= 4 && j + d0 - 1 >= 0 && (s & 31) <= 27) {  /* 4 slides, one word */
            const u32 c4nn = ts.next4();
#pragma unroll
            for (int u = 0; u < 4; u++) step_slide((c4 >> (8 * u)) & 0xFFu);
            j += 4;
            if (early && over()) return kq + 1;
            c4 = c4n;
            c4n = c4nn;
        }
        /* one group column by column: the slide starts, the window enters the
         * next plane word, or fewer than 4 columns are left */
        const u32 c4nn = ts.next4();
        const int cnt = n - j < 4 ? n - j : 4;
        for (int u = 0; u < cnt; u++) {
            j += 1;
            if (j + d0 - 1 > s) {
                step_slide((c4 >> (8 * u)) & 0xFFu);
                if ((s & 31) == 0) refresh();
            } else {
                step((c4 >> (8 * u)) & 0xFFu);
            }
        }
        if (j >= n) {
            const int d = value(m);
            return d <= kq ? d : kq + 1;
        }
        if (early && over()) return kq + 1;
        c4 = c4n;
        c4n = c4nn;
    }
}

/* A pair before its band scan: trim, swap (pattern a = the shorter string),
 * then the gap / empty rules. Returns the result when these decide it, else
 * -1 (a scan is needed; m - n = la - lb <= 0). */
BL_HD int band_prepare(const u8 *&a, int &la, const u8 *&b, int &lb, int trim, int k)
{
    if (trim) trim_pair(a, la, b, lb);
    if (la > lb) {
        const u8 *t = a; a = b; b = t;
        const int x = la; la = lb; lb = x;
    }
    if (k >= 0 && lb - la > k) return k + 1;
    if (la == 0) return lb;
    return -1;
}

/* A prepared pair through band_scan at the caller's offset d0: the pair's
 * result (min(d, k + 1), or d for k < 0), or -1 when d is above what the
 * window reaches (the caller then scans the whole pattern). */
template <typename T>
BL_HD int band_pair(const u8 *a, int la, const u8 *b, int lb, int k, int d0)
{
    constexpr int WIN = Rows<T>::n;
    const int reach = band_reach<WIN>(la - lb, d0);
    if (reach < 0) return -1;
    const int kq = k >= 0 && k < reach ? k : reach;
    const BlockText text{b};
    const int r = band_scan<T>(a, la, text, lb, d0, kq, k >= 0);
    if (r <= kq) return r;
    return k >= 0 && kq == k ? k + 1 : -1;
}

/* ------------------------------------------------------------ Peq lane */
/* Distance of a pattern given by its Peq table (row c: NW words, peq[c*NW + w])
 * of length 1 <= m <= 64*NW against a text of n bytes (a ByteText or a
 * StripeText). No trim: the pattern is the query.
 * lo32[c] = the low 32 bits of peq[c*NW] (a compact table for the 32-row
 * stage). Semantics of text_distance_peq, csrc/gpu_myers.cuh:124-145, and
 * run_text / run_text_k, csrc/myers.c:58-77, :107-130.
 * FROM: the first stage to run (0: 32 rows, 1: 64 rows); a caller that already
 * ran the 32-row stage passes 1 (lo32 is then not read). EAGER: the whole
 * stage fetches its four columns' Eq words before stepping (false: column by
 * column, fewer live registers). */
template <int NW, int FROM, bool EAGER = true, typename Lo32Ptr, typename PeqPtr, typename Text>
BL_HD int peq_lane(Lo32Ptr lo32, PeqPtr peq, int m, const Text &t, int n, int k)
{
    if (k >= 0 && (n > m ? n - m : m - n) > k) return k + 1;
    if (n == 0) return m;
    if (FROM == 0) {
        auto eq32 = [&](u32 c, int) { return (u32)lo32[c]; };
        if (m <= 32) return scan_text<u32, 1, true>(m, 1, 32, t, n, k, eq32);
        if (k >= 0) {
            const int r = scan_text<u32, 1, true>(m, 1, 32, t, n, k, eq32);
            if (r >= 0) return r;
        }
    }
    auto eq64 = [&](u32 c, int w) { return (u64)peq[c * NW + w]; };
    if (NW == 1 || m <= 64) return scan_text<u64, 1, true>(m, 1, 64, t, n, k, eq64);
    if (k >= 0) {
        const int r = scan_text<u64, 1, true>(m, 1, 64, t, n, k, eq64);
        if (r >= 0) return r;
    }
    const int nw = (m + 63) >> 6;
    return scan_text<u64, NW, EAGER>(m, nw, 64 * nw, t, n, k, eq64);
}

/* The search lane in 32-row units: a query of m rows, 32 (U - 1) < m <= 32 U,
 * given by its Peq table -- eq64(c, w) = rows 64w .. 64w + 63 of the row of
 * byte c, eq32(c, u) = rows 32u .. 32u + 31 -- against a text of n bytes.
 * U = 1 scans one u32 block, an even U U/2 u64 blocks, an odd U > 1 (U - 1)/2
 * u64 blocks and a u32 tail, so a pattern costs its 32-row units and no more
 * (peq_lane rounds up to 64-row words). k < 0: the exact distance; k >= 0:
 * min(d, k + 1), the 32-row stage first as in peq_lane. Same results as
 * peq_lane (csrc/gpu_myers.cuh:124-145, csrc/myers.c:58-77, :107-130). */
template <int U, typename Text, typename Eq64, typename Eq32>
BL_HD int units_lane(int m, const Text &t, int n, int k, Eq64 eq64, Eq32 eq32)
{
    static_assert(U >= 1 && U <= 16, "1 to 16 units of 32 rows");
    if (k >= 0 && (n > m ? n - m : m - n) > k) return k + 1;
    if (n == 0) return m;
    auto lo = [&](u32 c, int) { return eq32(c, 0); };
    if constexpr (U == 1) {
        return scan_text<u32, 1, true>(m, 1, 32, t, n, k, lo);
    } else {
        if (k >= 0) {
            const int r = scan_text<u32, 1, true>(m, 1, 32, t, n, k, lo);
            if (r >= 0) return r;
        }
        constexpr int W = U / 2;
        if constexpr (U % 2 == 0) {
            return scan_text<u64, W, (W <= 4)>(m, W, 64 * W, t, n, k, eq64);
        } else {
            return scan_mixed<W, true>(m, t, n, k, eq64, [&](u32 c) { return eq32(c, U - 1); });
        }
    }
}

/* Empty cross query: d = n, or k + 1 when k >= 0 and n > k
 * (csrc/gpu_kernels.cu:53-54, csrc/edpy.c:111-115). */
BL_HD int empty_pattern(int n, int k) { return (k >= 0 && n > k) ? k + 1 : n; }

/* ------------------------------------------------------------ 4-choice filter */
/* Bytewise x + y with no carry between bytes (the top bit of each byte loses
 * its carry-out, which belongs to rows beyond the byte's segment). */
BL_HD u32 swar_add(u32 x, u32 y)
{
    const u32 lo = (x & 0x7F7F7F7Fu) + (y & 0x7F7F7F7Fu);
    return lo ^ ((x ^ y) & 0x80808080u);
}

/* Four choices in the four bytes of u32 blocks, byte s holding pattern rows
 * 0..7 of choice s (a pattern of m >= 8 rows), stepped through their first 4
 * text columns (d[s]: their first 4 bytes, n[s] >= 4) at once: carries and
 * shifts stay inside each byte, row 0 of every byte gets the +1 of a global
 * alignment. lo8[stride * c] = rows 0..7 of the Peq row of byte value c.
 * bound[s] = the final-diagonal bound D[4 + m - n[s]][4] when that row lies
 * in 1..8 (exact: rows 0..7 evolve as in the whole scan), else -1. */
template <typename Lo8Ptr>
BL_HD void swar_steps4(Lo8Ptr lo8, const u32 (&d)[4], u32 &Pv, u32 &Mv, int stride)
{
    Pv = ~0u;
    Mv = 0u;
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const u32 E = (u32)lo8[stride * ((d[0] >> (8 * u)) & 0xFFu)] |
                      ((u32)lo8[stride * ((d[1] >> (8 * u)) & 0xFFu)] << 8) |
                      ((u32)lo8[stride * ((d[2] >> (8 * u)) & 0xFFu)] << 16) |
                      ((u32)lo8[stride * ((d[3] >> (8 * u)) & 0xFFu)] << 24);
        const u32 Xv = E | Mv;
        const u32 S = swar_add(E & Pv, Pv);
        const u32 Xh = bop3<tt((TA ^ TB) | TC)>(S, Pv, E);          /* (S ^ Pv) | E    */
        const u32 Ph = bop3<tt(TA | ~(TB | TC))>(Mv, Xh, Pv);       /* Mv | ~(Xh | Pv) */
        const u32 Mh = Pv & Xh;
        const u32 Phs = (Ph << 1) | 0x01010101u;                    /* +1 into row 0 */
        const u32 Mhs = (Mh << 1) & 0xFEFEFEFEu;
        Pv = bop3<tt(TA | ~(TB | TC))>(Mhs, Xv, Phs);               /* Mhs | ~(Xv | Phs) */
        Mv = Phs & Xv;
    }
}

template <typename Lo8Ptr>
BL_HD void swar_bounds4(Lo8Ptr lo8, const u32 (&d)[4], int m, const int (&n)[4], int (&bound)[4],
                        int stride = 4)
{
    u32 Pv, Mv;
    swar_steps4(lo8, d, Pv, Mv, stride);
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const int i0 = 4 + m - n[s];
        const u32 mk = i0 >= 8 ? 0xFFu : ((1u << (i0 > 0 ? i0 : 0)) - 1u);
        const int b = 4 + popc((u32)((Pv >> (8 * s)) & mk)) - popc((u32)((Mv >> (8 * s)) & mk));
        bound[s] = (i0 >= 1 && i0 <= 8) ? b : -1;
    }
}

/* The four bounds of swar_bounds4 tested against k <= 3 at once: bit 7 of
 * byte s is set when bound[s] > k, for every choice whose row 4 + m - n[s]
 * lies in 1..8 (the others have |m - n| >= 4 > k, decided by the gap). Per
 * byte: a mask of the rows below that one, popcounts p of Pv and q of Mv
 * under it (<= 8 each), then 128 + p - q compared with 124 + k by one add
 * (no borrow or carry crosses a byte). */
BL_HD u32 swar_over4(u32 Pv, u32 Mv, int m, const int (&n)[4], int k)
{
    u32 mk = 0;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        int r = 4 + m - n[s];
        r = r < 0 ? 0 : r > 8 ? 8 : r;
        mk |= (u32)((1u << r) - 1u) << (8 * s);
    }
    auto pc8 = [](u32 x) {                       /* popcount of every byte */
        x = x - ((x >> 1) & 0x55555555u);
        x = (x & 0x33333333u) + ((x >> 2) & 0x33333333u);
        return (x + (x >> 4)) & 0x0F0F0F0Fu;
    };
    const u32 t = (pc8(Pv & mk) | 0x80808080u) - pc8(Mv & mk);
    return (t + (u32)(3 - k) * 0x01010101u) & 0x80808080u;
}

/* ------------------------------------------------------------ class counts */
/* A lower bound on the distance from byte-class counts, used by the search
 * kernels before any scan. Bytes fall into 16 classes, class(b) = b & 15; a
 * string's profile holds, in byte i of dword j, the number of its bytes of
 * class 4j + i, saturated at 255.
 *
 * Let a_c, b_c be the true class counts of strings a, b and P = sum of
 * max(a_c - b_c, 0), N = sum of max(b_c - a_c, 0). An insertion or deletion
 * changes one count by one, a substitution lowers one count and raises
 * another, so every edit lowers P by at most one and N by at most one; hence
 * d(a, b) >= max(P, N) = (L1 + |len a - len b|) / 2 with L1 = P + N (P - N
 * = len a - len b). With saturated counts L1' <= L1 (clamping moves no two
 * counts apart) while the length gap comes from the lengths themselves, so
 * d >= (L1' + |len a - len b|) / 2 at any length. The bound is exact in the
 * sense the results need: a pair it puts above k has d > k, and its stored
 * result min(d, k + 1) = k + 1 is what the scan would give. */
struct alignas(16) Profile {
    u32 x, y, z, w;
};

BL_HD u32 sad_u8(u32 a, u32 b, u32 acc)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sad_u8(a, b, acc);
#else
    for (int i = 0; i < 32; i += 8) {
        const int x = (int)((a >> i) & 255u), y = (int)((b >> i) & 255u);
        acc += (u32)(x > y ? x - y : y - x);
    }
    return acc;
#endif
}

/* Does the class-count bound put d(query, choice) above k? qc, cc: the
 * profiles; m, n: the lengths. */
BL_HD bool count_over(const Profile &qc, const Profile &cc, int m, int n, int k)
{
    u32 l1 = sad_u8(qc.x, cc.x, 0u);
    l1 = sad_u8(qc.y, cc.y, l1);
    l1 = sad_u8(qc.z, cc.z, l1);
    l1 = sad_u8(qc.w, cc.w, l1);
    const u32 gap = (u32)(n > m ? n - m : m - n);
    return (u64)l1 + gap > 2ull * (u64)k;
}

/* A second bound from 2-grams. Byte b falls into class c2(b) = (b >> 1) & 3
 * (A, C, G, T apart), a 2-gram (x, y) into class 4 c2(x) + c2(y); the
 * profile holds the saturated counts of the 16 classes over the string's
 * max(n - 1, 0) 2-grams. With P and N the excess and the deficit of a's
 * counts against b's: a substitution removes two 2-grams and adds two, an
 * insertion removes one and adds two, a deletion removes two and adds one,
 * so each edit lowers P by at most two and N by at most two, and d >=
 * max(P, N) / 2 = (L1 + |t_a - t_b|) / 4 with t the 2-gram totals; as for
 * the class counts, saturation only shrinks L1, so this holds at any length. */
BL_HD u32 class2(u32 b) { return (b >> 1) & 3u; }

BL_HD bool count2_over(const Profile &qc, const Profile &cc, int m, int n, int k)
{
    u32 l1 = sad_u8(qc.x, cc.x, 0u);
    l1 = sad_u8(qc.y, cc.y, l1);
    l1 = sad_u8(qc.z, cc.z, l1);
    l1 = sad_u8(qc.w, cc.w, l1);
    const int ta = m > 0 ? m - 1 : 0, tb = n > 0 ? n - 1 : 0;
    const u32 gap = (u32)(ta > tb ? ta - tb : tb - ta);
    return (u64)l1 + gap > 4ull * (u64)k;   /* 4k overflows 32 bits at k = 2^30 */
}

/* A fourth bound from 3-grams, for small alphabets (DNA), where byte classes
 * and 2-grams leave many unrelated pairs open. A 3-gram (x, y, z) falls into
 * class 16 c2(x) + 4 c2(y) + c2(z) of 64; the profile holds the saturated
 * counts over the string's max(n - 2, 0) 3-grams, byte i of dword j (of 16)
 * counting class 4j + i. A substitution removes three 3-grams and adds three,
 * an insertion removes two and adds three, a deletion removes three and adds
 * two, so each edit lowers the excess P and the deficit N of a's counts by
 * at most three each: d >= max(P, N) / 3 = (L1 + |t_a - t_b|) / 6, t the
 * 3-gram totals (at any length: saturation only shrinks L1).
 * Two random DNA strings of 230 bytes differ by about 137 in L1 over the 64
 * classes (bound about 23), so at k = 8 almost only true near matches stay
 * open. */
struct alignas(16) Profile3 {
    Profile p[4];
};

BL_HD bool count3_over(const Profile3 &qc, const Profile3 &cc, int m, int n, int k)
{
    u32 l1 = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        l1 = sad_u8(qc.p[i].x, cc.p[i].x, l1);
        l1 = sad_u8(qc.p[i].y, cc.p[i].y, l1);
        l1 = sad_u8(qc.p[i].z, cc.p[i].z, l1);
        l1 = sad_u8(qc.p[i].w, cc.p[i].w, l1);
    }
    const int ta = m > 1 ? m - 2 : 0, tb = n > 1 ? n - 2 : 0;
    const u32 gap = (u32)(ta > tb ? ta - tb : tb - ta);
    return (u64)l1 + gap > 6ull * (u64)k;
}

BL_HD u32 class3(u32 x, u32 y, u32 z) { return 16u * class2(x) + 4u * class2(y) + class2(z); }

/* Pack 64 counts, count(x) for class x, into a Profile3, saturated at 255. */
template <typename Get>
BL_HD Profile3 pack_profile3(Get count)
{
    Profile3 r;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        u32 w[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            u32 v = 0;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const u32 c = (u32)count(16 * q + 4 * j + i);
                v |= (c > 255u ? 255u : c) << (8 * i);
            }
            w[j] = v;
        }
        r.p[q] = Profile{w[0], w[1], w[2], w[3]};
    }
    return r;
}

#ifndef __HIP_DEVICE_COMPILE__
/* The 3-gram profile of s[0..n) (host; the kernels count into LDS). */
static inline Profile3 class3_profile(const u8 *s, int n)
{
    u32 cnt[64] = {0};
    for (int i = 0; i + 2 < n; i++) cnt[class3(s[i], s[i + 1], s[i + 2])]++;
    return pack_profile3([&](int x) { return cnt[x]; });
}
#endif

/* A third bound over the byte-class counts, in the form the matrix cores
 * evaluate. An insertion or a deletion moves the count vector by a unit
 * vector, a substitution by the difference of two (L2 norm sqrt 2), so after
 * d edits ||count(a) - count(b)||_2 <= sqrt(2) d, and d > k whenever
 * ||count(a) - count(b)||^2 > 2 k^2. Weaker than the L1 bound (an L2 norm is
 * at most the L1 norm), but ||a||^2 + ||b||^2 - 2 a.b needs only one int8
 * matrix product for a block of query and choice profiles. Saturated counts
 * only shrink the norm (clamping is 1-Lipschitz), so it holds at any length;
 * the kernels use it for strings of at most 64 bytes, whose counts and
 * doubled counts fit int8. */
BL_HD u32 profile_dot(const Profile &a, const Profile &b)
{
    u32 s = 0;
    const u32 A[4] = {a.x, a.y, a.z, a.w}, B[4] = {b.x, b.y, b.z, b.w};
    for (int j = 0; j < 4; j++) {
#if defined(__HIP_DEVICE_COMPILE__)
        s = __builtin_amdgcn_udot4(A[j], B[j], s, false);
#else
        for (int i = 0; i < 32; i += 8) s += ((A[j] >> i) & 255u) * ((B[j] >> i) & 255u);
#endif
    }
    return s;
}

BL_HD bool count_l2_over(const Profile &qc, const Profile &cc, int k)
{
    const int64_t d2 = (int64_t)profile_dot(qc, qc) + (int64_t)profile_dot(cc, cc) -
                       2 * (int64_t)profile_dot(qc, cc);
    return d2 > 2 * (int64_t)k * (int64_t)k;    /* k <= 2^30: fits 64 bits */
}

/* Pack 16 counts, count(x) for class x, into a Profile, saturated at 255. */
template <typename Get>
BL_HD Profile pack_profile(Get count)
{
    u32 w[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        u32 v = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const u32 c = (u32)count(4 * j + i);
            v |= (c > 255u ? 255u : c) << (8 * i);
        }
        w[j] = v;
    }
    return Profile{w[0], w[1], w[2], w[3]};
}

#ifndef __HIP_DEVICE_COMPILE__
/* The byte-class profile of s[0..n) (class b & 15; host, the kernels count
 * into LDS). */
static inline Profile class_profile(const u8 *s, int n)
{
    u32 cnt[16] = {0};
    for (int i = 0; i < n; i++) cnt[s[i] & 15u]++;
    return pack_profile([&](int x) { return cnt[x]; });
}

/* The 2-gram profile of s[0..n) (classes as for count2_over). */
static inline Profile class2_profile(const u8 *s, int n)
{
    u32 cnt[16] = {0};
    for (int i = 0; i + 1 < n; i++) cnt[4u * class2(s[i]) + class2(s[i + 1])]++;
    return pack_profile([&](int x) { return cnt[x]; });
}
#endif

/* ------------------------------------------------------------ long strings */
/* Pairs whose shorter string exceeds 256 bytes, and cross queries longer than
 * 512 bytes, run here instead of on the CPU (the reference falls back to its C
 * core for them, bitlev/gpu.py:97-101, :131-136). One wave per (pattern, text):
 * lane l owns pattern word 64*s + l of stripe s (4096 rows per stripe). The
 * carry between words moves one lane per step, so lane l works on text column
 * step - l (an anti-diagonal wavefront); each step hands (text byte, hp, hn) to
 * the next lane. A stripe's bottom-row deltas go to hbuf[j] for the next
 * stripe, so patterns of any length are covered. No threshold early exit here:
 * with k >= 0 the result is min(d, k + 1) after the gap test.
 *
 * The lane logic below is shared by the device kernel and the host simulation
 * in bl_hostcheck.cpp; only the lane-to-lane hand-off differs. */
struct LongCtx {
    const u8 *p;        /* pattern (the shorter string for pairs; the query for cross) */
    int m;
    const u8 *t;        /* text */
    int n;
    int W;              /* pattern words, ceil(m / 64) */
    int s;              /* stripe */
    int nl;             /* lanes with a word in this stripe */
    int8_t *hbuf;       /* n bottom-row deltas between stripes (W > 64 only) */
};

struct LongLane {
    u64 pl[8];
    u64 Pv, Mv;
    int w, bit, delta;
};

BL_HD void long_lane_init(const LongCtx &C, LongLane &L, int lane)
{
    L.w = 64 * C.s + lane;
    L.Pv = ~(u64)0;
    L.Mv = 0;
    L.bit = L.w == C.W - 1 ? ((C.m - 1) & 63) : 63;
    L.delta = 0;
    if (lane < C.nl) build_planes(C.p + 64 * L.w, C.m - 64 * L.w, L.pl);
}

/* Input of lane 0 at `step`: bits 0..7 text byte, bit 8 hp, bit 9 hn. */
BL_HD u32 long_source(const LongCtx &C, int step)
{
    if (step >= C.n) return 0;
    const u32 c = C.t[step];
    if (C.s == 0) return c | (1u << 8);                /* row 0: +1 per column */
    const int h = ((volatile int8_t *)C.hbuf)[step];
    return c | ((u32)(h > 0) << 8) | ((u32)(h < 0) << 9);
}

/* One step of one lane; returns what the next lane receives on the next step. */
BL_HD u32 long_lane_step(const LongCtx &C, LongLane &L, int lane, int step, u32 in)
{
    const int j = step - lane;
    if (lane >= C.nl || j < 0 || j >= C.n) return 0;
    u32 m8[8];
    text_masks(in & 0xFFu, m8);
    u64 Ph, Mh;
    word_step(L.Pv, L.Mv, eq_planes(L.pl, m8), (in >> 8) & 1u, (in >> 9) & 1u, Ph, Mh);
    const u32 hp = (u32)(Ph >> L.bit) & 1u, hn = (u32)(Mh >> L.bit) & 1u;
    if (L.w == C.W - 1) L.delta += (int)hp - (int)hn;
    else if (lane == C.nl - 1) ((volatile int8_t *)C.hbuf)[j] = (int8_t)((int)hp - (int)hn);
    return (in & 0xFFu) | (hp << 8) | (hn << 9);
}

/* Resolve one long item to (pattern, text). Pairs trim (optionally) as the
 * reference does; an empty cross query keeps the reference's rule. Then the
 * shorter string becomes the pattern (distance is symmetric, so this only
 * shortens the wavefront) and the gap test applies. Returns the final result
 * (>= 0) when no scan is needed, else -1. */
BL_HD int long_prepare(const u8 *&p, int &m, const u8 *&t, int &n, int is_pair, int trim, int k)
{
    if (is_pair && trim) trim_pair(p, m, t, n);
    if (!is_pair && m == 0) return empty_pattern(n, k);
    if (m > n) {
        const u8 *x = p; p = t; t = x;
        const int y = m; m = n; n = y;
    }
    if (k >= 0 && n - m > k) return k + 1;
    if (m == 0) return n;
    return -1;
}

/* ------------------------------------------------------------ batched pairs */
/* Bucket keys of the batched pairs pipeline, from the untrimmed lengths (so
 * classifying reads no string bytes; trimming only shortens both strings by
 * the same amount, so the pattern words of the untrimmed shorter string bound
 * the trimmed ones and the length gap is unchanged):
 *   1 + (w - 1) * LB_BINS + min(max(la, lb) / bin_width, LB_BINS - 1)
 *       for a shorter string of w <= long_units units of 32 bytes
 *       (whole-pattern tiles, one lane per pair); with unit64 w counts
 *       64-byte words twice (w even: the lanes hold u64 blocks only, in fewer
 *       and fuller buckets);
 *   PKEY_BAND0 + that length bin * GAP_BINS + |la - lb| / 4
 *       instead when band > 0, w >= band and the gap is at most 63 (band tiles:
 *       one window offset serves a tile well when its gaps are close);
 *   PKEY_LONG for a shorter string of more units (above 256 bytes by
 *       default; long_units < 8 sends shorter ones there too): one whole wave
 *       per pair on the anti-diagonal wavefront, which for a handful of long
 *       pairs finishes sooner than lanes of their own;
 *   0 for a pair the gap or an empty string decides.
 * Larger keys mean more work; the pipeline schedules large keys first. */
constexpr int LB_BINS = 64;
constexpr int GAP_BINS = 16;
constexpr int PKEY_BAND0 = 1 + 8 * LB_BINS;
constexpr int PKEY_LONG = PKEY_BAND0 + LB_BINS * GAP_BINS;
constexpr int PKEYS = PKEY_LONG + 1;

/* Key of a pair of lengths la, lb; for key 0, *result holds the answer
 * (the gap rule, csrc/myers.c:137, and the empty rule, csrc/myers.c:90-99). */
BL_HD int pair_key(int la, int lb, int k, int bin_width, int band, int long_units, bool unit64,
                  int &result)
{
    const int mn = la < lb ? la : lb, mx = la < lb ? lb : la;
    if (k >= 0 && mx - mn > k) { result = k + 1; return 0; }
    if (mn == 0) { result = mx; return 0; }
    if (mn > 256) return PKEY_LONG;
    const int w = unit64 ? 2 * ((mn + 63) >> 6) : (mn + 31) >> 5;   /* 32-row units */
    if (w > long_units) return PKEY_LONG;
    int bin = mx / bin_width;
    bin = bin < LB_BINS - 1 ? bin : LB_BINS - 1;
    if (band > 0 && w >= band && mx - mn <= 63) return PKEY_BAND0 + bin * GAP_BINS + (mx - mn) / 4;
    return 1 + (w - 1) * LB_BINS + bin;
}

/* Band tiles: those whose largest key is a band key. */
BL_HD bool key_band(int key) { return key >= PKEY_BAND0 && key < PKEY_LONG; }

/* 32-row units of a whole-pattern tile whose largest key is `key`. */
BL_HD int key_units(int key) { return (key - 1) / LB_BINS + 1; }

/* Pair i of packed sets through the register-resident lane code (shorter
 * string at most 64*NW bytes before trimming). */
template <int NW>
BL_HD int packed_pair(const u8 *ba, const int64_t *oa, const u8 *bb, const int64_t *ob, int64_t i,
                      int trim, int k)
{
    const int64_t a0 = oa[i], b0 = ob[i];
    return pair_lane<NW>(ba + a0, (int)(oa[i + 1] - a0), bb + b0, (int)(ob[i + 1] - b0), trim, k);
}

/* The same in 32-row units (shorter string at most 32*W32 bytes before trimming). */
template <int W32>
BL_HD int packed_pair_w32(const u8 *ba, const int64_t *oa, const u8 *bb, const int64_t *ob,
                          int64_t i, int trim, int k)
{
    const int64_t a0 = oa[i], b0 = ob[i];
    return pair_lane_w32<W32>(ba + a0, (int)(oa[i + 1] - a0), bb + b0, (int)(ob[i + 1] - b0), trim,
                              k);
}

/* ------------------------------------------------------------ work queues */
/* The persistent kernels' tile schedule. Tiles go out in runs; with
 * first_static, wave w's first run is run w. After that wave w draws from
 * queue queue_of(w, waves) only, and its draw that finds the queue's counter
 * at `taken` gets run queue_run(w, waves, first_static, taken): queue j hands
 * out S + j, S + j + Q, S + j + 2Q, ... (S the static round, Q the queues in
 * use). Q = min(QUEUES, waves), so every queue in use has a wave to drain it. */
constexpr int QUEUES = 16;
BL_HD int64_t queue_count(int64_t waves) { return waves < QUEUES ? waves : QUEUES; }
BL_HD int queue_of(int64_t wid, int64_t waves) { return (int)(wid % queue_count(waves)); }
BL_HD int64_t queue_run(int64_t wid, int64_t waves, bool first_static, int64_t taken)
{
    return (first_static ? waves : 0) + queue_of(wid, waves) + queue_count(waves) * taken;
}

}  // namespace bl
