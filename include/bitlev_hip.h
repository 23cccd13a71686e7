/*
 * bitlev_hip.h -- C ABI of libbitlev_hip.so, the MI355X (gfx950) backend of
 * bitlev's exact unit-cost Levenshtein distance.
 *
 * Conventions (all entry points):
 *   * Strings are packed: string i of a set is buf[off[i] : off[i+1]] with
 *     int64 offsets (the layout of the reference's _edpy.pack, csrc/edpy.c:200-224).
 *     Byte buffers need no padding or alignment.
 *   * Pointers named d_* are device pointers (from bl_malloc or any HIP allocation
 *     on the current device); h_* are host pointers.
 *   * k < 0 asks for the exact distance; k >= 0 returns min(d, k + 1)
 *     (bitlev/__init__.py:14-18). Values above 2^30 are clamped like the
 *     reference's check_k (bitlev/_common.py:9-15).
 *   * dtype: BL_INT32, BL_UINT16, BL_UINT8; narrow types saturate at their
 *     maximum after the k cap (bitlev/_common.py:25-30).
 *   * stream: a hipStream_t cast to void* (NULL = the default stream). Calls are
 *     asynchronous on that stream unless documented otherwise.
 *   * Return value: BL_OK (0) or a negative BL_E* code; bl_last_error() gives
 *     the message of the calling thread's last failure.
 *   * No entry point calls back into its caller; callers may release the GIL.
 */
#ifndef BITLEV_HIP_H
#define BITLEV_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    BL_OK = 0,
    BL_EALLOC = -1,    /* allocation failed (MemoryError)      -- csrc/myers.c:155-156 */
    BL_EINVAL = -2,    /* invalid argument (ValueError)        -- csrc/myers.c:372-373 */
    BL_EHIP = -3,      /* HIP runtime error                                        */
    BL_ENODEV = -4,    /* no usable GPU                                            */
};

enum { BL_INT32 = 0, BL_UINT16 = 1, BL_UINT8 = 2 };

/* ---------------------------------------------------------------- runtime */
int bl_version(void);
const char *bl_last_error(void);
/* Number of visible HIP devices (0 when none); never fails. Replaces
 * bitlev/gpu.py:36-41 (available()). */
int bl_device_count(void);
int bl_set_device(int device);
/* Name and compute-unit count of the current device. */
int bl_device_info(char *name, size_t name_len, int *cu_count, int *arch_major, int *arch_minor);

int bl_malloc(void **d_ptr, size_t bytes);
int bl_free(void *d_ptr);
int bl_host_malloc(void **h_ptr, size_t bytes);        /* pinned host memory */
int bl_host_free(void *h_ptr);
int bl_memcpy_htod(void *d_dst, const void *h_src, size_t bytes, void *stream);
int bl_memcpy_dtoh(void *h_dst, const void *d_src, size_t bytes, void *stream);
int bl_memcpy_dtod(void *d_dst, const void *d_src, size_t bytes, void *stream);
int bl_memset(void *d_dst, int value, size_t bytes, void *stream);
int bl_stream_create(void **stream);
int bl_stream_destroy(void *stream);
int bl_stream_sync(void *stream);
int bl_device_sync(void);
int bl_event_create(void **event);
int bl_event_destroy(void *event);
int bl_event_record(void *event, void *stream);
/* Work issued to `stream` after this call waits for `event` (no host wait). */
int bl_stream_wait_event(void *stream, void *event);
int bl_event_sync(void *event);
int bl_event_query(void *event);                       /* 1 done, 0 pending, <0 error */
int bl_event_elapsed_ms(void *start, void *stop, float *ms);

/* HIP graphs: capture the work issued on `stream` (not the default stream)
 * between bl_graph_begin and bl_graph_end into an executable graph; replay it
 * with bl_graph_launch. The batched pipelines (bl_pairs, bl_cdist) with a
 * caller-provided workspace are capture-safe: they clear their counters with a
 * kernel, never a memset, and allocate nothing. */
int bl_graph_begin(void *stream);
int bl_graph_end(void *stream, void **graph_exec);
int bl_graph_launch(void *graph_exec, void *stream);
int bl_graph_destroy(void *graph_exec);

/* ---------------------------------------------------------------- plain form */
/* One pairs class, one lane per pair: pair i = d_idx[t] for t < cnt, result
 * written to out[i]. cap in {16, 32, 64, 128, 256}: every listed pair has
 * min(la, lb) <= cap. Replaces the kernels ed_pairs_c{16,32,64,128,256}
 * (csrc/gpu_kernels.cu:21-32) and their launch in bitlev/gpu.py:89-96. */
int bl_pairs_class(int cap, const uint8_t *d_buf_a, const int64_t *d_off_a,
                   const uint8_t *d_buf_b, const int64_t *d_off_b, const int32_t *d_idx,
                   int64_t cnt, void *d_out, int dtype, int trim, int k, void *stream);

/* ---------------------------------------------------------------- choice stripes */
/* A choice set laid out for the search kernels: each group of 64 choices is
 * interleaved by dword, dword j of choice c at words[gbase[c / 64] + 64 * j +
 * c % 64] (zero-padded to the group's widest string, at least one dword), so
 * the lanes of a wave, holding consecutive choices, read 256 contiguous bytes
 * per step. len[c] is the length of choice c. counts (optional, 16-byte
 * aligned) holds each choice's byte-class profile in dwords 4c .. 4c + 3: byte
 * i of dword 4c + j counts the bytes b of choice c with b & 15 == 4j + i,
 * saturated at 255. The batched search kernels use it for an exact lower
 * bound on the distance, d >= (L1 + |m - n|) / 2 over the two profiles (d >=
 * L1 / 2 if either string exceeds 255 bytes), before any scan; NULL turns
 * that test off. All pointers are device memory. The reference searches the
 * packed buffer directly (csrc/gpu_kernels.cu:62-68, strided byte loads); the
 * stripes are this library's own resident format. */
typedef struct {
    const uint32_t *words;   /* gbase[ceil(n / 64)] dwords */
    const int64_t *gbase;    /* ceil(n / 64) + 1 group bases, in dwords */
    const int32_t *len;      /* n lengths */
    int64_t n;               /* number of choices */
    const uint32_t *counts;  /* 4 * n dwords of class profiles, or NULL */
    const uint32_t *counts2; /* 4 * n dwords of 2-gram profiles, or NULL: byte i of
                                dword 4c + j counts the 2-grams (x, y) of choice c
                                with 4 ((x >> 1) & 3) + ((y >> 1) & 3) == 4j + i,
                                saturated at 255; a second lower bound, d >=
                                (L1 + |t_q - t_c|) / 4 over the two (t: the
                                2-gram totals; L1 / 4 past 256 bytes), applied to
                                what the first leaves open */
    const int32_t *order;    /* NULL: slot c holds choice c. Otherwise the set is
                                in length order (bl_stripes_layout_sorted): slot c
                                holds choice order[c], and words, len and the
                                profiles are indexed by slot; results still go to
                                the caller's column order[c]. Only the exact-search
                                kernel (lane mode 5) reads such a set: bl_cdist*
                                run it whatever the mode asked for, and
                                bl_cdist_class_stripes rejects it */
    const uint32_t *counts3; /* 16 * n dwords of 3-gram profiles, or NULL: byte i
                                of dword 16c + j counts the 3-grams (x, y, z) of
                                choice c with 16 ((x >> 1) & 3) + 4 ((y >> 1) & 3)
                                + ((z >> 1) & 3) == 4j + i, saturated at 255; a
                                third bound, d >= (L1 + |t_q - t_c|) / 6 (L1 / 6
                                past 257 bytes), applied by the count windows to
                                what the 2-gram bound leaves open (16-byte
                                aligned) */
    const uint8_t *buf;      /* the packed set the stripes were built from (buf,
                                off: the choices in the caller's order), or NULL:
                                with it the near matches of count mode read their
                                text from the packed bytes (a string in 2-3 cache
                                lines) instead of the stripes (a line per dword) */
    const int64_t *off;
} bl_stripes;

/* Host: group bases of the stripes of n strings with host offsets
 * h_off[0..n] into h_gbase[0..ceil(n/64)]. Returns the number of dwords the
 * words array needs (h_gbase[ceil(n/64)]), or a negative status. */
int64_t bl_stripes_layout(const int64_t *h_off, int64_t n, int64_t *h_gbase);

/* Device: bl_stripes_layout for offsets already in HBM: the group bases of
 * the n strings with device offsets d_off[0..n] into device d_gbase[0..
 * ceil(n/64)], computed on `stream` with d_work (at least
 * bl_stripes_layout_workspace_size(n) bytes) as scratch. Synchronises the
 * stream and returns the number of dwords the words array needs, or a
 * negative status (decreasing offsets: BL_EINVAL). No counterpart in the
 * reference, whose kernels read the packed bytes: the host pass over 4.2M
 * offsets takes about 4 ms, several times the device build it prepares. */
size_t bl_stripes_layout_workspace_size(int64_t n);
int64_t bl_stripes_layout_device(const int64_t *d_off, int64_t n, int64_t *d_gbase, void *d_work,
                                 size_t work_bytes, void *stream);

/* Device: fill words, len and (unless NULL) the class and 2-gram profiles
 * from packed strings (d_buf, d_off) using the group bases d_gbase that
 * bl_stripes_layout or bl_stripes_layout_device computed for the same
 * offsets. */
int bl_stripes_build(const uint8_t *d_buf, const int64_t *d_off, int64_t n,
                     const int64_t *d_gbase, uint32_t *d_words, int32_t *d_len,
                     uint32_t *d_counts, uint32_t *d_counts2, void *stream);

/* Host: the length-ordered layout of n strings with host offsets
 * h_off[0..n]: within each block of `block` slots (a positive multiple of
 * 64; the last block may be short) the strings are sorted by length, stably,
 * into h_order[0..n) (slot c holds string h_order[c], a string of the same
 * block), and h_gbase[0..ceil(n/64)] receives the group bases of the strings
 * in that order. Returns the number of dwords of the words array, or a
 * negative status. No counterpart in the reference: a wave of an exact search
 * runs as long as its longest string, so lanes of equal lengths waste
 * nothing. */
int64_t bl_stripes_layout_sorted(const int64_t *h_off, int64_t n, int64_t block,
                                 int32_t *h_order, int64_t *h_gbase);

/* Device: bl_stripes_build for the slot order d_order (from
 * bl_stripes_layout_sorted, uploaded; NULL: slot c gets string c, with the
 * group bases of bl_stripes_layout): slot c gets string d_order[c]. Also
 * fills the 3-gram profiles unless d_counts3 is NULL. */
int bl_stripes_build_order(const uint8_t *d_buf, const int64_t *d_off, int64_t n,
                           const int32_t *d_order, const int64_t *d_gbase, uint32_t *d_words,
                           int32_t *d_len, uint32_t *d_counts, uint32_t *d_counts2,
                           uint32_t *d_counts3, void *stream);

/* dst[r][i] = src[r][inv[i]] for r < rows, i < cols (dtype BL_INT32 /
 * BL_UINT16 / BL_UINT8; src and dst distinct). No counterpart in the
 * reference: a resident choice set kept sorted by length (so a wave's
 * choices have similar lengths) gets its search results back in the
 * caller's choice order with it (inv[i] = sorted position of choice i). */
int bl_gather_cols(const void *d_src, const int64_t *d_inv, void *d_dst, int64_t rows, int64_t cols,
                   int dtype, void *stream);

/* One cdist class: queries d_qidx[0..nqc) (each at most 64*nw bytes, nw in
 * {1, 2, 4, 8}) against the nc packed choices (d_buf_c, d_off_c); out is the
 * row-major nq x nc matrix of the call (out[q * nc + c], q the query's index
 * in d_off_q). One workgroup per (query, choice block), the query's Peq table
 * in LDS, lanes over consecutive choices, each reading its own string from
 * the packed buffer. Same arguments as ed_cross_w{1,2,4,8}(buf_q, off_q, qidx,
 * buf_c, off_c, nc, out, k, out8) (csrc/gpu_kernels.cu:59-65), with nw for
 * the kernel name's suffix, nqc for the grid size and dtype for out8;
 * replaces those kernels and their launch in bitlev/gpu.py:121-130. */
int bl_cdist_class(int nw, const uint8_t *d_buf_q, const int64_t *d_off_q,
                   const int32_t *d_qidx, int64_t nqc, const uint8_t *d_buf_c,
                   const int64_t *d_off_c, int64_t nc, void *d_out, int dtype, int k, void *stream);

/* bl_cdist_class over a resident choice set in the stripes layout
 * (bl_stripes_build; nc = choices->n): the lanes of a wave read 256
 * contiguous bytes per step instead of 64 scattered strings. No counterpart
 * in the reference, which searches the packed buffer on every call. */
int bl_cdist_class_stripes(int nw, const uint8_t *d_buf_q, const int64_t *d_off_q,
                           const int32_t *d_qidx, int64_t nqc, const bl_stripes *choices,
                           void *d_out, int dtype, int k, void *stream);

/* Long strings on the GPU: pairs whose shorter string exceeds 256 bytes
 * (is_pair = 1: d_idx lists pair indices; trim applies) or cdist queries
 * longer than 512 bytes (is_pair = 0: d_idx lists query indices, each against
 * all nc choices). One wave per (pattern, text), any length. max_len bounds
 * every string involved; above 4096 it sizes a stream-ordered scratch of
 * bottom-row deltas. Replaces the CPU fallbacks of bitlev/gpu.py:97-101 and
 * :131-136. */
int bl_long(int is_pair, const uint8_t *d_buf_a, const int64_t *d_off_a,
            const uint8_t *d_buf_b, const int64_t *d_off_b, const int32_t *d_idx, int64_t cnt,
            int64_t nc, void *d_out, int dtype, int trim, int k, int64_t max_len,
            void *stream);

/* ---------------------------------------------------------------- batched form */
/* Launch shape of the persistent kernels; zero fields take the defaults.
 * Chosen per workload by bitlev.autotune. */
typedef struct {
    int waves_per_cu;      /* persistent waves per compute unit (default 16) */
    int64_t tile;          /* choices per (query, block) tile (default 8192) */
    int reserved[4];       /* [0]: lane mode -- 0 automatic (5 for an exact search
                              with no per-query thresholds, else 6), 1 strided
                              lanes, 2 four-choice byte filter for k <= 3 (queries of
                              8 to 256 bytes; other queries run compacted), 3 gap
                              windows for k >= 0 (the length test on 256 choices at
                              a time, the in-band ones scanned 64 at a time), 4 count
                              windows for k >= 0 (needs the choices' class profiles,
                              else mode 6): a tile is a choice block against a group
                              of queries, the length and class-count bounds settle
                              most choices, the rest are rescanned with per-lane
                              queries; 5 exact scans: one workgroup per (query,
                              block of `tile` choices; default sized from the batch),
                              the query's whole Peq table in LDS, each lane scanning
                              its choice in 32-row units (any threshold; the only
                              mode for choices in length order); 6 compacted passes;
                              [1]: tiles per work-queue ticket (default 1);
                              [2]: count mode: queries per tile (default 16, at most 64);
                              [3]: count mode options: bit 0 switches off the
                              matrix-core tiles (queries of at most 64 bytes and
                              windows of choices of at most 64 bytes are otherwise
                              tested 32 x 32 pairs at a time with an int8 MFMA over
                              the class counts, a squared-L2 bound); bit 1 makes
                              every tile come from the work queue (by default each
                              wave's first tile is assigned statically); bit 2
                              leaves the escalation list unused, so the count
                              kernel scans its near matches itself (for short
                              queries that costs less than the list kernel's
                              launch: 1,024 x 1e6 at k = 2, 0.316 -> 0.311 ms) */
} bl_cdist_params;

typedef struct {
    int waves_per_cu;      /* persistent waves per compute unit (default 16) */
    int block;             /* workgroup size: 64, 128 or 256 (default 256) */
    int bin_width;         /* text-length bucket width in bytes (default 16) */
    int reserved[4];       /* [0]: pairs per bucketing chunk (default 65536);
                              [1]: 1 = direct mode (one launch, no bucketing);
                              [2]: tiles per work-queue ticket (default 1);
                              [3]: bits 0-3: tiles of at least this many 32-row
                              units (1..8) scan a diagonal band first (Ukkonen: an
                              alignment of cost <= k stays within k + 1 diagonals),
                              pairs the band cannot decide then the whole pattern
                              (default 0: no band tiles); bits 4-7 (lanes per
                              pair): pairs whose shorter string needs more than
                              this many 32-row units (1..8; default 0 = 8, i.e.
                              above 256 bytes) run one whole wave per pair on the
                              anti-diagonal wavefront instead of a lane each;
                              bit 8 (words per lane): lanes hold their pattern in
                              64-row words only, no 32-row tail unit */
} bl_pairs_params;

/* Bytes of device workspace bl_cdist needs for nq queries of at most
 * max_qlen bytes: the work-queue counters (4 KiB), one Peq table per query
 * (256 x ceil(max_qlen / 64) u64 words, the word count rounded up to 1, 2, 4
 * or 8), then two 16-byte profiles per query (its byte-class and 2-gram
 * profiles) and one 64-byte 3-gram profile; then the escalation list of the
 * count windows (lane mode 4): a counter and min(2048 nq, 2^20) 8-byte
 * entries. A workspace may be smaller, down to the part before the list
 * (the near matches the list would hold are then scanned in the count
 * kernel: same results, slower), or larger (a longer list). Size workspaces
 * from the return value, not from this list. 0 for a negative argument. */
size_t bl_cdist_workspace_size(int64_t nq, int max_qlen);

/* All queries (each at most max_qlen <= 512 bytes) against all nc choices:
 * a pre-pass writes every query's Peq table to the workspace, then one
 * persistent launch in which waves pull (choice block, query) tiles from a work
 * queue. Lanes scan the low 32 pattern rows first (the table's low dwords,
 * copied into the wave's LDS slice) and refill from a per-wave ring of in-band
 * choices as they finish (length-gap and empty choices are answered without
 * occupying a lane); the few choices that stage cannot decide are rescanned
 * with all rows. Row-major out[q * nc + c], nc = choices->n. d_work: >=
 * bl_cdist_workspace_size(nq, max_qlen) bytes of device memory the call may
 * overwrite, or NULL for a stream-ordered allocation. Queries above 512 bytes
 * go to bl_long. Replaces the per-class ed_cross_w{NW} launches and host split
 * of bitlev/gpu.py:116-130. */
int bl_cdist(const uint8_t *d_buf_q, const int64_t *d_off_q, int64_t nq,
             const bl_stripes *choices, int max_qlen, void *d_out, int dtype, int k,
             void *d_work, size_t work_bytes, void *stream);
/* bl_cdist with an optional per-query threshold d_kvec[q] (NULL: k for all)
 * and an explicit launch shape (NULL: defaults). Lets one launch serve
 * requests with different max_dist. */
int bl_cdist_tuned(const uint8_t *d_buf_q, const int64_t *d_off_q, int64_t nq,
                   const bl_stripes *choices, int max_qlen, void *d_out, int dtype, int k,
                   const int32_t *d_kvec, void *d_work, size_t work_bytes,
                   const bl_cdist_params *params, void *stream);

/* Device scratch held by bl_cdist_noalloc on each GPU it runs on (until the
 * process ends): 8 MiB. */
#define BL_NOALLOC_SCRATCH_BYTES (8 << 20)

/* bl_cdist_tuned without any allocation: the workspace (work-queue counter,
 * Peq tables, query profiles, and an escalation list in what remains) is a
 * static device array of BL_NOALLOC_SCRATCH_BYTES per GPU instead of a
 * caller buffer or a stream-ordered allocation; the call makes no hipMalloc,
 * hipFree or other allocating call. Returns BL_EINVAL when the queries need
 * more than that (the part of bl_cdist_workspace_size(nq, max_qlen) before
 * the list above BL_NOALLOC_SCRATCH_BYTES: about 3,900 queries of up to 64
 * bytes, 500 of up to 512). Every call on one GPU
 * shares that GPU's array, so calls that overlap on one GPU (on different
 * streams or from different threads) race; calls ordered on one stream do
 * not. No counterpart in the reference: for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). */
int bl_cdist_noalloc(const uint8_t *d_buf_q, const int64_t *d_off_q, int64_t nq,
                     const bl_stripes *choices, int max_qlen, void *d_out, int dtype, int k,
                     const int32_t *d_kvec, const bl_cdist_params *params, void *stream);

/* Threshold match list instead of the dense matrix: every (query q, choice c)
 * with d(q, c) <= k is appended as d_q[i] = q, d_c[i] = c, d_d[i] = d, in no
 * particular order, for i < min(*d_count, capacity); *d_count receives the
 * number of matches (more than capacity: call again with a larger buffer).
 * k >= 0 (or a per-query d_kvec; queries with a negative threshold add
 * nothing). Same launch as bl_cdist_tuned, with no nq x nc write: the four-
 * choice filter and the band exits decide almost every cell without any
 * store. d_count is device memory, cleared by the call. No counterpart in the
 * reference; it names this output as its next step (README.md:319-322). */
int bl_cdist_matches(const uint8_t *d_buf_q, const int64_t *d_off_q, int64_t nq,
                     const bl_stripes *choices, int max_qlen, int k, const int32_t *d_kvec,
                     int32_t *d_q, int64_t *d_c, int32_t *d_d, int64_t capacity,
                     unsigned long long *d_count, void *d_work, size_t work_bytes,
                     const bl_cdist_params *params, void *stream);

/* Bytes of device workspace bl_pairs needs for n pairs whose strings are at
 * most max_len bytes (descriptors, bucket keys, sorted order, counters, and
 * the wavefront scratch when max_len > 4096). */
size_t bl_pairs_workspace_size(int64_t n, int64_t max_len);

/* All n pairs in one pipeline with no host round trip: classify (trim, swap,
 * trivial answers, bucket key) -> bucket scan -> scatter -> one persistent
 * launch over 64-pair tiles, most expensive bucket first; long pairs (shorter
 * string > 256 bytes) run the wavefront in the same launch. out[i] in input
 * order. Replaces the host class split and ed_pairs_c{CAP} launches and the CPU
 * fallback of bitlev/gpu.py:85-101. */
int bl_pairs(const uint8_t *d_buf_a, const int64_t *d_off_a, const uint8_t *d_buf_b,
             const int64_t *d_off_b, int64_t n, int64_t max_len, void *d_out, int dtype,
             int trim, int k, void *d_work, size_t work_bytes, void *stream);
/* bl_pairs with an optional per-pair threshold d_kvec[i] (NULL: k for all)
 * and an explicit launch shape (NULL: defaults). */
int bl_pairs_tuned(const uint8_t *d_buf_a, const int64_t *d_off_a, const uint8_t *d_buf_b,
                   const int64_t *d_off_b, int64_t n, int64_t max_len, void *d_out, int dtype,
                   int trim, int k, const int32_t *d_kvec, void *d_work, size_t work_bytes,
                   const bl_pairs_params *params, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* BITLEV_HIP_H */
