/* cpu_engine.h -- the CPU engine behind the reference's CPU methods
 * (method="scalar" / "simd"): packed buffers + int64 offsets, int32 results,
 * k < 0 exact else min(d, k + 1); 0, -1 (allocation) or -2 (invalid
 * argument). The SIMD entry points run the scalar path where AVX2 is
 * missing. */
#pragma once
#include <stdint.h>

int cpu_simd_available(void);
int cpu_pairs(const uint8_t *ba, const int64_t *oa, const uint8_t *bb, const int64_t *ob, int64_t n,
              int32_t *out, int trim, int k);
int cpu_cross(const uint8_t *bq, const int64_t *oq, int64_t nq, const uint8_t *bc, const int64_t *oc,
              int64_t nc, int32_t *out, int k);
/* exact distances (the reference's pairs_simd caps with k afterwards) */
int cpu_pairs_simd(const uint8_t *ba, const int64_t *oa, const uint8_t *bb, const int64_t *ob,
                   int64_t n, int32_t *out, int trim);
int cpu_cross_simd(const uint8_t *bq, const int64_t *oq, int64_t nq, const uint8_t *bc,
                   const int64_t *oc, int64_t nc, int32_t *out, int k);
