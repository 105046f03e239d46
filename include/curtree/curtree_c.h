/*
 * curtree_c.h -- C ABI of the MI355X (gfx950) R-tree library libcurtree_amd.so.
 *
 * The structs are layout-identical to the reference's data model
 * (src/rtree.h:3-29): int sizes followed by raw device pointers.  Every entry
 * point returns a status code (CURTREE_OK == 0); curtree_last_error() gives the
 * message of the calling thread's last failure.  Device pointers are HIP device
 * memory; `stream` is a hipStream_t (NULL = the null stream).  Entry points that
 * mirror a synchronous reference call synchronise `stream` before returning.
 */
#ifndef CURTREE_C_H
#define CURTREE_C_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CURTREE_API __attribute__((visibility("default")))

/* status codes */
#define CURTREE_OK 0
#define CURTREE_EINVAL 1   /* bad argument */
#define CURTREE_ENOMEM 2   /* device allocation failed */
#define CURTREE_EHIP 3     /* HIP runtime error */
#define CURTREE_ERANGE 4   /* result larger than the caller's buffer / int32 limit */
#define CURTREE_EBUSY 5    /* serving queue full / budget exhausted */

/* key modes for curtree_build_f32 */
#define CURTREE_KEY_REF32 0    /* the reference's 16-bit-per-axis Z-order key, int32 order */
#define CURTREE_KEY_MORTON64 1 /* 32 bits per axis over the centroid extent, uint64 order */

/* query algorithms */
/* level-synchronous BFS frontier in HBM; each level by the launches that suit
 * the run's size (chained or count / scan / write) */
#define CURTREE_ALGO_FRONTIER 0
/* the same frontier, every level (leaf stage included) one launch of
 * persistent blocks that take tiles from a ticket counter and chain them by a
 * decoupled look-back (fewest launches; the same pairs) */
#define CURTREE_ALGO_PERSISTENT 1

/* == BBOX<float>, src/rtree.h:3-11 */
typedef struct curtree_bbox_f32 {
  int sz;
  float* xmin;
  float* ymin;
  float* xmax;
  float* ymax;
  int* id;
} curtree_bbox_f32;

/* == RTREE<float>, src/rtree.h:13-23.  Nodes root-level first; the root is
 * node 0 and the leaf-parent level is the last block.  pos = first child (for
 * leaf-parents: R_sz + first sorted-box position), len = child count. */
typedef struct curtree_rtree_f32 {
  int sz;
  int fanout;
  float* xmin;
  float* ymin;
  float* xmax;
  float* ymax;
  int* pos;
  int* len;
} curtree_rtree_f32;

/* == IDPAIR, src/rtree.h:25-29: fid = data box id, tid = query index */
typedef struct curtree_idpair {
  int sz;
  int* fid;
  int* tid;
} curtree_idpair;

/* IDPAIR with a 64-bit count, for result sets beyond 2^31-1 pairs */
typedef struct curtree_idpair64 {
  int64_t sz;
  int* fid;
  int* tid;
} curtree_idpair64;

/* ---- bulk load ---------------------------------------------------------- */

/* Replaces build_rtree<float> (src/rtree_bulkload_thrust.cuh:62-213).
 * Sorts the caller's box arrays IN PLACE into key order (the in-place
 * sort_by_key of :83-86), allocates the six node arrays (as rt_d_alloc,
 * src/utility.cuh:114-135; free with curtree_rt_d_free) and fills them level
 * by level.  rt->fanout must be set by the caller (2..any); rt->sz receives
 * R_sz.  Synchronises `stream`. */
CURTREE_API int curtree_build_f32(curtree_bbox_f32* box, curtree_rtree_f32* rt, int key_mode,
                                  void* stream);

/* Same as curtree_build_f32 but asynchronous: returns after enqueueing all
 * kernels on `stream` (the node arrays are allocated before it returns). */
CURTREE_API int curtree_build_f32_async(curtree_bbox_f32* box, curtree_rtree_f32* rt,
                                        int key_mode, void* stream);

/* Same as curtree_build_f32_async, into node arrays the caller allocated:
 * rt->sz must equal the node count for box->sz boxes at rt->fanout (see
 * curtree_tree_size).  Allocates no node arrays (work buffers are cached). */
CURTREE_API int curtree_build_into_f32(curtree_bbox_f32* box, curtree_rtree_f32* rt,
                                       int key_mode, void* stream);

/* Node count R_sz and number of levels of a bulk-loaded tree of n boxes. */
CURTREE_API int curtree_tree_size(int64_t n, int fanout, int64_t* r_sz, int* levels);

/* ---- window query ------------------------------------------------------- */

/* Replaces bfs_frontier<float> (src/rtree_bfs_frontier.cuh:292-529) and
 * bfs_thrust<float> (src/rtree_bfs_thrust.cuh:58-256): all (box id, query
 * index) pairs whose boxes intersect the query windows (closed intervals),
 * in (query index, sorted box position) order -- the order the reference
 * produces before its host sort.  `rbox` is the box set as sorted by the
 * build.  Allocates out->fid/out->tid (free with curtree_idpair64_d_free).
 * Never aborts on large frontiers.  Synchronises `stream`. */
CURTREE_API int curtree_query_f32(const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox,
                                  const curtree_bbox_f32* qbox, curtree_idpair64* out, int algo,
                                  void* stream);

/* Same result written into caller buffers of `cap` pairs, in one pass over the
 * leaf level.  *n_out receives the pair count; if it exceeds `cap`,
 * CURTREE_ERANGE is returned and the buffer contents are unspecified.
 * Synchronises `stream`. */
CURTREE_API int curtree_query_into_f32(const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox,
                                       const curtree_bbox_f32* qbox, int* fid, int* tid,
                                       int64_t cap, int64_t* n_out, int algo, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc): the result of
 * curtree_query_into_f32, computed with no allocation at all.  The work
 * buffers are one file-scope device array of 128 MiB that every GPU the
 * library runs on holds until the process ends.  Limits: at most 1,572,864
 * windows per call (else CURTREE_EINVAL) and at most 3,145,728 (window, node)
 * entries on any tree level (else CURTREE_ERANGE, the pairs are then not
 * written); the containment shortcut is off, which never changes the pairs.
 * Calls in one process run one at a time (a lock); the array keeps the last
 * call's contents and is shared by every call on the GPU, so calls that
 * overlap on one GPU from several processes or libraries race.  Synchronises
 * `stream`. */
CURTREE_API int curtree_query_into_noalloc_f32(const curtree_rtree_f32* rt,
                                               const curtree_bbox_f32* rbox,
                                               const curtree_bbox_f32* qbox, int* fid, int* tid,
                                               int64_t cap, int64_t* n_out, void* stream);

/* Count pass only: number of pairs per query (`counts`, Q int64 in device
 * memory, may be NULL) and in total (*n_out). */
CURTREE_API int curtree_query_count_f32(const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox,
                                        const curtree_bbox_f32* qbox, int64_t* counts,
                                        int64_t* n_out, void* stream);

/* Batched query of a window set whose pairs may exceed device memory (no
 * reference counterpart: the reference runs one batch and aborts on large
 * frontiers, src/rtree_bfs_frontier.cuh:385-403).  Pass 1 counts every
 * window's pairs; windows are then cut into chunks of consecutive windows
 * whose pairs fit `budget_pairs` (a window with more pairs than the budget
 * gets a chunk of its own); pass 2 replays one captured hipGraph per chunk.
 * `fn` receives each chunk's pairs in device memory, in (query, sorted box
 * position) order, valid until it returns; a nonzero return stops the batch. */
typedef struct curtree_batch_stats {
  int64_t pairs;           /* pairs delivered */
  int64_t chunks;          /* graph replays */
  int64_t max_chunk_pairs; /* largest chunk */
  double count_ms;         /* pass 1, device time */
  double write_ms;         /* pass 2, device time summed over chunks */
} curtree_batch_stats;
typedef int (*curtree_chunk_fn)(void* user, const int* fid, const int* tid, int64_t n,
                                int64_t q_first, int64_t q_end);
CURTREE_API int curtree_query_batched_f32(const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox,
                                          const curtree_bbox_f32* qbox, int64_t budget_pairs,
                                          curtree_chunk_fn fn, void* user,
                                          curtree_batch_stats* stats);

/* ---- serving (no reference counterpart) ----------------------------------
 * Several bulk-loaded trees stay resident; window requests of any size are
 * submitted concurrently and served by continuous batching: a dispatcher
 * thread merges queued requests on the same tree into one query batch per
 * free slot (stream), admits a batch when its pairs fit the result budget,
 * and overlaps the slots' work on the device.  A request's result (device
 * pairs, query indices relative to the request, (query, sorted box position)
 * order) is identical alone and in any mix, and stays valid until released.
 * Window arrays passed to submit must stay valid until the request is done. */
#define CURTREE_REQ_QUEUED 0
#define CURTREE_REQ_RUNNING 1
#define CURTREE_REQ_DONE 2
#define CURTREE_REQ_FAILED 3
typedef struct curtree_server curtree_server;
typedef struct curtree_server_opts {
  int slots;                 /* concurrent batches (streams), 1..16 */
  int64_t budget_bytes;      /* result memory (8 bytes per pair) */
  int64_t max_batch_windows; /* window cap of one batch (and of one request) */
  int batching;              /* 0: each request alone (plain per-request launches) */
} curtree_server_opts;
typedef struct curtree_server_stats {
  int64_t requests_done, pairs_done, batches, queued;
  double p50_us, p99_us; /* submit -> completion latency */
  double span_us;        /* first submit -> last completion */
  double launch_us;      /* time spent enqueuing device work, summed over slots */
} curtree_server_stats;
CURTREE_API int curtree_server_create(const curtree_server_opts* opts, curtree_server** out);
CURTREE_API int curtree_server_add_tree(curtree_server* s, const curtree_rtree_f32* rt,
                                        const curtree_bbox_f32* rbox, int* tree_id);
CURTREE_API int curtree_server_submit(curtree_server* s, int tree_id,
                                      const curtree_bbox_f32* windows, int64_t* request_id);
/* state: CURTREE_REQ_*; result/latency filled once done */
CURTREE_API int curtree_server_poll(curtree_server* s, int64_t request_id, int* state,
                                    curtree_idpair64* result, double* latency_us);
/* Completion queue: ids of requests that finished (done or failed) since the
 * last call, oldest first, at most cap of them; *n = how many.  Each id is
 * reported once; poll it for the result. */
CURTREE_API int curtree_server_completed(curtree_server* s, int64_t* ids, int cap, int* n);
CURTREE_API int curtree_server_release(curtree_server* s, int64_t request_id);
CURTREE_API int curtree_server_get_stats(curtree_server* s, curtree_server_stats* stats);
CURTREE_API int curtree_server_reset_stats(curtree_server* s);
CURTREE_API int curtree_server_destroy(curtree_server* s);

/* Canonical result order: sorts pairs on the device into (query id, box id)
 * lexicographic order -- the order the reference's two host stable sorts give
 * (src/rtree_test_frontier.cu:102-105).  Pairs must arrive grouped by query
 * (as every query entry point returns them). */
CURTREE_API int curtree_pairs_canonical_sort(int* fid, int* tid, int64_t n, void* stream);

/* Distance self-join: all (box id, querying box id) pairs whose boxes are
 * within r of each other along both axes -- the tree queried with every data
 * box expanded by r, as the reference's drivers do (query file = data file
 * read with r = 10, src/rtree_test_frontier.cu:84-85; Boost check
 * src/verify.hpp:75-104).  `rbox` is the sorted box set of the build; tid
 * holds the querying box's id.  Order: (querying box's sorted position, box
 * sorted position).  Allocates out->fid/tid.  Synchronises `stream`. */
CURTREE_API int curtree_self_join_f32(const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox,
                                      float r, curtree_idpair64* out, void* stream);

/* ---- ingest and persistence --------------------------------------------- */

/* Replaces readMBR (src/utility.cpp:5-40): reads `size` lines
 * "x1,y1,x2,y2,id,pid,nume" into host arrays the caller allocated; box =
 * [min(x1,x2)-r, max(x1,x2)+r] x [min(y1,y2)-r, max(y1,y2)+r].  Open or parse
 * failures return CURTREE_EINVAL (the reference exits / throws).  Host only. */
CURTREE_API int curtree_read_mbr_csv(const char* path, int size, float r, float* xmin, float* ymin,
                                     float* xmax, float* ymax, int* id);

/* Write the sorted boxes and node arrays of one tree to a binary file / read
 * them back into newly allocated device arrays (no reference counterpart: the
 * reference never persists a tree; used to keep serving trees on disk). */
CURTREE_API int curtree_tree_save(const char* path, const curtree_rtree_f32* rt,
                                  const curtree_bbox_f32* rbox);
CURTREE_API int curtree_tree_load(const char* path, curtree_rtree_f32* rt,
                                  curtree_bbox_f32* rbox);

/* ---- memory helpers (src/utility.cuh:16-205, over hipMalloc/hipFree) ----- */

/* box_h2d: allocate device arrays for h->sz boxes and copy them in */
CURTREE_API int curtree_box_h2d(curtree_bbox_f32* d, const curtree_bbox_f32* h);
/* box_d2h: copy into host arrays the caller allocated (h->xmin.. of d->sz) */
CURTREE_API int curtree_box_d2h(curtree_bbox_f32* h, const curtree_bbox_f32* d);
CURTREE_API int curtree_box_d_alloc(curtree_bbox_f32* d, int sz);
CURTREE_API int curtree_box_d_free(curtree_bbox_f32* d);
CURTREE_API int curtree_rt_d_alloc(curtree_rtree_f32* rt, int sz);
CURTREE_API int curtree_rt_d_free(curtree_rtree_f32* rt);
CURTREE_API int curtree_idpair_d_alloc(curtree_idpair* p, int sz);
CURTREE_API int curtree_idpair_d_free(curtree_idpair* p);
CURTREE_API int curtree_idpair64_d_free(curtree_idpair64* p);
/* idpair_d2h: copy into host arrays of d->sz entries the caller allocated */
CURTREE_API int curtree_idpair64_d2h(int* h_fid, int* h_tid, const curtree_idpair64* d);

/* raw device memory (used by non-C hosts: Python, Go, ...) */
CURTREE_API int curtree_malloc(void** p, int64_t bytes);
CURTREE_API int curtree_free(void* p);
CURTREE_API int curtree_memcpy_h2d(void* dst, const void* src, int64_t bytes);
CURTREE_API int curtree_memcpy_d2h(void* dst, const void* src, int64_t bytes);
CURTREE_API int curtree_memcpy_d2d(void* dst, const void* src, int64_t bytes, void* stream);
CURTREE_API int curtree_memset(void* dst, int value, int64_t bytes);
CURTREE_API int curtree_synchronize(void* stream);
CURTREE_API int curtree_stream_create(void** stream);
CURTREE_API int curtree_stream_destroy(void* stream);

/* ---- diagnostics -------------------------------------------------------- */

CURTREE_API const char* curtree_last_error(void);
CURTREE_API const char* curtree_version(void);
/* device count, name of device 0, CU count, HBM bytes */
CURTREE_API int curtree_device_info(int* count, char* name, int name_len, int* cus,
                                    int64_t* hbm_bytes);
/* Frontier sizes of the last plain query run in this process (diagnostics
 * for measurement; no counterpart in the reference): entries[l] = (window,
 * node) entries processed at depth l (l = 0: one per window at the root; the
 * last depth: leaf-parent entries plus contained-subtree entries), *levels =
 * number of depths, *pairs = pairs of the run, *pieces = long box ranges of
 * contained subtrees written by the piece writer.  Zero levels before any
 * query. */
CURTREE_API int curtree_query_last_stats(int64_t* entries, int cap, int* levels, int64_t* pairs,
                                         int64_t* pieces);

/* Tuning knobs (process-wide):
 *   "containment"  -1 auto (default: on for >= 256 windows), 0 off, 1 on.  A node
 *                  whose MBR lies inside a window and whose boxes are all
 *                  well-formed is expanded without reading child MBRs/boxes;
 *                  results are identical either way.
 *   "query_tile_steps"  4, 8 (default) or 16: wave-steps per frontier tile
 *                  (tile = 4 waves x steps x 64/next_pow2(F) entries).
 *   "query_grid"   blocks of each chained query launch (they loop over
 *                  tiles), i.e. waves per CU, for batches of at least 8 x that
 *                  many windows (smaller batches: windows/8, at least 512);
 *                  0 (default): windows/8 clamped to [512, 2048].
 *   "box_staging"  -1 or 0 (default: off), 1 on: a query copies boxes,
 *                  node MBRs and windows into AoS float4 arrays first (one
 *                  extra pass over the boxes; a server made while it is on
 *                  keeps such copies of its trees).  Results are identical
 *                  either way; the SoA arrays measured faster.
 *   "expand_split" -1 (default: by run size), 0, 1: internal tree levels are
 *                  expanded by one chained-scan kernel (0) or by count, tile
 *                  scan and write kernels with no chain between tiles (1;
 *                  auto for runs of at least 32768 windows and for batched
 *                  chunks).  Results are identical either way.
 *   "leaf_fused"   0 (default) never, 1 always, -1 for runs below 32768
 *                  windows: a writing query tests the leaf entries' boxes and
 *                  stores their pairs in one chained pass (as
 *                  CURTREE_ALGO_PERSISTENT does), instead of a count pass, a
 *                  tile scan and a write pass.  Results are identical.
 *   "window_order" -1 (default: for batches of at least 4194304 windows), 0
 *                  off, 1 on: the plain entry points process a batch's windows
 *                  in Morton order of their centres (windows that meet the
 *                  same leaf-parents are tested close together, so repeated
 *                  box reads hit in cache) and write every window's pairs
 *                  where the input order puts them.  Results are identical.
 *   "tile_xcd"     0 (default) or 1: the count and write kernels' blocks on
 *                  one XCD walk a contiguous eighth of the tiles.  Results
 *                  are identical.
 * Returns CURTREE_EINVAL for an unknown name or value. */
CURTREE_API int curtree_set_option(const char* name, int64_t value);

/* Release the library's cached work buffers (they are reused across calls). */
CURTREE_API int curtree_release_workspace(void);

/* Per-kernel timing with HIP events on the launch stream.  While enabled, every
 * kernel the library launches is bracketed by events; curtree_prof_get reports
 * the summed milliseconds and launch count of kernel `name` since the last
 * reset (events are resolved at the call, which synchronises). */
CURTREE_API int curtree_prof_enable(int on);
/* Bracket only the launches of kernel `name` (NULL or "": every kernel), so
 * that timing one kernel adds two events per launch of it and no others. */
CURTREE_API int curtree_prof_only(const char* name);
CURTREE_API int curtree_prof_reset(void);
CURTREE_API int curtree_prof_get(const char* name, double* total_ms, int64_t* launches);
/* Elapsed milliseconds between two events recorded with curtree_event_record. */
CURTREE_API int curtree_event_create(void** ev);
CURTREE_API int curtree_event_destroy(void* ev);
CURTREE_API int curtree_event_record(void* ev, void* stream);
CURTREE_API int curtree_event_elapsed(void* ev0, void* ev1, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* CURTREE_C_H */
