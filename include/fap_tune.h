/*
 * fap_tune.h -- autotuner and its cache.
 *
 * No counterpart in the reference (its HIP kernels use fixed shapes:
 * 32 x 32 FW tiles, fap/kernel/hip/floyd.cpp:253-295; 32 x 32 min-plus
 * tiles, fap/kernel/hip/minplus.cpp:91-128; one SSSP source at a time,
 * fap/kernel/hip/batch_sssp.cpp:91-158).
 *
 * Every tunable changes speed only, never a result: the kernels compute the
 * same exact distances under every setting.  A context consults its cache
 * for each option left at "auto" (fap_ctx_set_option value 0, or -1 for
 * fw_lookahead):
 *
 *   kind     key                                   params
 *   "fw"     fap_tune_key_fw(n)                    {tile 64|128, lookahead 0|1}
 *            (quarter octaves of n; a dense closure's cost does not depend on
 *            density, the SSSP key carries the graph's mean degree)
 *   "sssp"   fap_tune_key_sssp(V, E, sources)      {threads 256|512|1024, delta / mean weight}
 *   "batch"  fap_tune_key_batch(count, n)          {workgroups per CU per launch}
 *
 * With no entry the defaults are tile 128 when ceil(n / 128) - 1 squared is
 * at least 8 per CU (64 below), no look-ahead, SSSP threads 1024
 * while the sources fit on the CUs (512 up to twice that, 256 beyond),
 * delta 8 x mean weight while the sources fit on the CUs (6 x up to twice
 * that, 4 x beyond), 64 workgroups per CU.
 *
 * The cache is a text file, one entry per line: `kind key p0 [p1] ms`.
 * A process whose environment names a file in FAP_TUNE_CACHE loads it once
 * and attaches it to every context it creates.
 */
#ifndef FAP_TUNE_H_
#define FAP_TUNE_H_

#include <stdint.h>

#include "fap_dev.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fap_tune_cache fap_tune_cache;

int fap_tune_cache_create(fap_tune_cache **out);
int fap_tune_cache_destroy(fap_tune_cache *c);
/* merges the file's entries (later lines win); FAP_E_IO if unreadable or malformed */
int fap_tune_cache_load(fap_tune_cache *c, const char *path);
int fap_tune_cache_save(fap_tune_cache *c, const char *path);
int fap_tune_cache_size(fap_tune_cache *c, int32_t *entries);
/* params: 2 for "fw" and "sssp", 1 for "batch"; values are checked */
int fap_tune_cache_set(fap_tune_cache *c, const char *kind, int64_t key, const int32_t *params,
                       int32_t nparams, double ms);
/* FAP_E_STATE when there is no entry */
int fap_tune_cache_get(fap_tune_cache *c, const char *kind, int64_t key, int32_t *params,
                       int32_t nparams, double *ms);

/* buckets: ceil(log2 n^4) (quarter octaves); (ceil(log2 V) * 64 + round(2 log2(1 + E / V))) * 32
 * + ceil(log2 sources); ceil(log2 of the batch's 64-tile triple count) */
int64_t fap_tune_key_fw(int32_t n);
int64_t fap_tune_key_sssp(int32_t V, int64_t E, int32_t n_src);
int64_t fap_tune_key_batch(int32_t count, const int32_t *n);

/* the cache a context consults (not owned; NULL detaches) */
int fap_ctx_set_tune_cache(fap_ctx *ctx, fap_tune_cache *c);

/* Measure every candidate on the device and record the fastest in the
 * context's cache (FAP_E_STATE if it has none).  fw: an n x n matrix of
 * the dense benchmark's kind (arcs with probability 1/64, weights 1..1000);
 * sssp: the given graph from the given host source ids; batch: matrices of
 * the given sizes.  reps timed runs per candidate after one warm-up. */
int fap_tune_fw(fap_ctx *ctx, int32_t n, int32_t reps);
int fap_tune_sssp(fap_ctx *ctx, const fap_graph_dev *g, const int32_t *h_sources, int32_t n_src,
                  int32_t reps);
int fap_tune_batch(fap_ctx *ctx, int32_t count, const int32_t *n, int32_t reps);

/* Measured ceiling of the min-plus instruction mix on this device (no
 * counterpart in the reference): one v_add_u32 per update and one
 * v_min3_i32 per two updates, eight independent chains per lane, every CU
 * busy, no memory traffic.  Reported in TOP/s at 2 ops per update, the
 * unit of the min-plus and FW rates; on gfx950 v_add_u32 issues at twice
 * the v_min3_i32 rate, so this exceeds 1 op per lane per clock. */
int fap_probe_minplus_ceiling(fap_ctx *ctx, double *tops);

#ifdef __cplusplus
}
#endif

#endif /* FAP_TUNE_H_ */
