/*
 * fap_serve.h -- request loop on one device: whole-graph APSP solves and
 * point-query batches, admitted by device-memory footprint and batched
 * continuously.
 *
 * No counterpart in the reference.  It takes the place of the multi-task
 * driver (fap/fap_more_task_mpi.cpp:34-91: part ids shuffled over MPI
 * processes, each solving its share in a host loop) for one device, and
 * answers the point queries the reference leaves as stubs
 * (getDistanceP2P, fap/fap.cpp:263-265) against a resident graph.
 *
 * Model.  A server owns an arena of device memory (the HBM budget), a few
 * lanes (a stream each, with pinned and device staging) and one host thread.
 * Requests are queued on submit.  Whenever a lane is free the thread takes
 * a wave from the queue -- requests of the head's kind in arrival order, as
 * many as fit in the arena, the staging and the wave limits -- and launches
 * it: for graphs, one dense build and one batched closure of every graph in
 * the wave; for queries, one lookup launch over the query index (or, without
 * one, one SSSP launch over the distinct sources of every request in the
 * wave and one gather of the answers).  Waves on different
 * lanes run concurrently; a request admitted while others run overlaps
 * with them on the device.  A graph's n x n int32 result stays in the arena
 * until released; query answers are copied to the host on completion.
 *
 * Results do not depend on the wave a request lands in: each request's
 * matrix or rows are computed by the same kernels on its own memory.
 *
 * Thread safety: every function may be called from any host thread.
 */
#ifndef FAP_SERVE_H_
#define FAP_SERVE_H_

#include <stdint.h>

#include "fap_dev.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fap_server fap_server;

typedef struct fap_serve_opts {
  int64_t hbm_budget;  /* arena bytes: graph result blocks (n * ld * 4) + query rows */
  int64_t stage_bytes; /* pinned + device staging per lane (CSR, descriptors, answers) */
  int32_t lanes;       /* waves in flight at once, 1..8 */
  int32_t max_batch;   /* requests per wave */
  int32_t max_sources; /* distinct query sources per wave */
  int32_t poll_us;     /* completion polling period of the server thread */
  int32_t query_parts; /* > 0: point queries use an index over this many parts of the
                        * resident graph (built by fap_serve_attach_graph); 0: each
                        * query batch runs SSSP from its sources */
} fap_serve_opts;

enum { FAP_REQ_GRAPH = 0, FAP_REQ_QUERY = 1 };
enum { FAP_REQ_PENDING = 0, FAP_REQ_RUNNING = 1, FAP_REQ_DONE = 2, FAP_REQ_FAILED = 3 };

typedef struct fap_serve_info {
  int32_t state;     /* FAP_REQ_* */
  int32_t kind;      /* FAP_REQ_GRAPH or FAP_REQ_QUERY */
  int32_t n;         /* graph: vertices; queries: pairs */
  int32_t wave_size; /* requests in the wave it ran in */
  int64_t wave;      /* wave number (0, 1, ...) */
  int64_t ld;        /* graph: row stride of the result block, in elements */
  double submit_us;  /* server clock, microseconds */
  double admit_us;
  double done_us;
} fap_serve_info;

typedef struct fap_serve_stats {
  int64_t submitted, completed, failed, pending, running;
  int64_t waves;
  int64_t arena_bytes, arena_used, arena_peak;
  int64_t fw_ops;       /* 2 n^3 over completed graph requests */
  int64_t sssp_sources; /* distinct sources over completed query requests */
  int32_t max_wave;     /* largest wave so far */
  int32_t lanes;
  int64_t index_bytes;  /* device memory of the query index (0 without one) */
  double index_build_ms;
} fap_serve_stats;

/* defaults: 8 GiB arena, 64 MiB staging, 3 lanes, 64 requests and 256
 * sources per wave, 20 us polling */
void fap_serve_default_opts(fap_serve_opts *opts);
int fap_serve_create(int device, const fap_serve_opts *opts, fap_server **out);
/* waits for waves in flight, fails requests still queued, frees everything */
int fap_serve_destroy(fap_server *s);

/* the resident graph point queries run against (CSR, int32 weights in
 * [0, FAP_INF)); FAP_E_BUSY while query requests are queued or running.
 * With opts.query_parts = K > 0 it also builds the query index -- the
 * reference's per-part pipeline without its last product: the graph is
 * cut into K parts (preCondition), SSSP rows from every boundary vertex
 * (sum of boundary counts x V int32) and every part's closed matrix stay
 * resident, and a query (s, t) is one cell of s's part block: the part
 * matrix entry when t is in s's part, the SSSP row when s is a boundary
 * vertex, else the min over s's part boundary k of M[s][k] + row_k[t]
 * (fap/fap.cpp:92-181).  Exact, like the full block. */
int fap_serve_attach_graph(fap_server *s, int32_t V, int64_t E, const int32_t *row_offset,
                           const int32_t *col, const int32_t *w);

/* whole-graph APSP (the K = 1 pipeline: dense matrix, closure).  Inputs are
 * copied; the result is an n x n int32 block (row stride info.ld). */
int fap_serve_submit_graph(fap_server *s, int32_t V, int64_t E, const int32_t *row_offset,
                           const int32_t *col, const int32_t *w, int64_t *id);
/* count (src, dst) pairs against the resident graph */
int fap_serve_submit_queries(fap_server *s, int32_t count, const int32_t *src, const int32_t *dst,
                             int64_t *id);

/* FAP_OK once done (info filled); FAP_E_BUSY if not done within timeout_ms
 * (0 = just look); the request's own error code and message if it failed */
int fap_serve_wait(fap_server *s, int64_t id, int32_t timeout_ms, fap_serve_info *info);
/* queries: `count` int32 distances (FAP_INF when unreachable) */
int fap_serve_answers(fap_server *s, int64_t id, int32_t *out);
/* graphs: rows [row0, row0 + nrows) of the n x n result, packed (stride n) */
int fap_serve_rows(fap_server *s, int64_t id, int32_t row0, int32_t nrows, int32_t *out);
/* graphs: fap_checksum_i32 of the n x n result as if packed (stride n) */
int fap_serve_checksum(fap_server *s, int64_t id, uint64_t *sum, uint64_t *weighted);
/* graphs: the device block itself (valid until release) */
int fap_serve_block(fap_server *s, int64_t id, int32_t **d_block, int64_t *ld, int32_t *n);
/* frees the request and its block; FAP_E_BUSY while queued or running */
int fap_serve_release(fap_server *s, int64_t id);
int fap_serve_get_stats(fap_server *s, fap_serve_stats *out);

#ifdef __cplusplus
}
#endif

#endif /* FAP_SERVE_H_ */
