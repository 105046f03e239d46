// ct_internal.hpp -- host-side internals shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>

#include <algorithm>
#include <map>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/curtree/curtree_c.h"
#include "ct_common.hpp"

namespace ct {

constexpr int kKeyRef32 = CURTREE_KEY_REF32;
constexpr int kKeyMorton64 = CURTREE_KEY_MORTON64;
constexpr int kMaxDigits = 8;

// Set the calling thread's last error; returns `code`.
int set_error(int code, const char* fmt, ...);

// Non-owning view of the six node arrays (device pointers).
struct TreeView {
  float* xmin;
  float* ymin;
  float* xmax;
  float* ymax;
  int* pos;
  int* len;
};

// Level plan: n[0] = leaf-parent level ... n[nl-1] = 1 (root).  Integer
// ceil-division; N == 1 gives a single root level (see level_plan in ct_capi).
struct LevelPlan {
  int nl = 0;
  int64_t r_sz = 0;
  int64_t n[64];
};
void level_plan(int64_t n, int fanout, LevelPlan* lp);

int rt_alloc(curtree_rtree_f32* rt, int64_t sz);
void rt_free(curtree_rtree_f32* rt);

// State of the trees this library built, kept with the tree (ct_capi.cpp):
// one device word per tree, written by its build, nonzero when every box is
// well-formed (xmin <= xmax and ymin <= ymax, no NaN).  A query on that tree
// with the same sorted boxes reads it instead of computing per-node flags
// (the containment shortcut needs them).  An entry is keyed by the six node
// arrays, the five box arrays and the sizes.  Building into, freeing, loading
// over or writing (through the library's memory helpers) any of those arrays
// forgets it; boxes changed by other means after the build need a rebuild.
// register: the word of the tree about to be built (nullptr: not kept, e.g.
// under graph capture); find: the word of a kept tree or nullptr.
uint32_t* tree_state_register(const curtree_rtree_f32* rt, const curtree_bbox_f32* box,
                              bool may_alloc);
const uint32_t* tree_state_find(const curtree_rtree_f32* rt, const curtree_bbox_f32* box);
// Leaf codes kept with a tree (`bytes` of them): the ready buffer, or null.
// With `build` and none ready, a buffer is found (a spare one, after a device
// synchronisation, since its last tree's queries may still read it, or a new
// one) and build(buffer) fills it; build must finish its device work before
// it returns.  Null when the tree is not kept or on any failure (the leaf
// count then tests the boxes directly).
const void* tree_state_codes(const curtree_rtree_f32* rt, const curtree_bbox_f32* box,
                             int64_t bytes, const std::function<int(void*)>* build);
// Frees the leaf-code buffers of forgotten trees (the device must be idle).
void tree_state_free_spare_codes();
// Forget every entry that names one of these arrays (null entries ignored).
void tree_state_forget(const void* const* arrays, int n);
// Forget every entry with an array that overlaps [p, p + bytes).
void tree_state_forget_range(const void* p, int64_t bytes);

// Cached device work buffers, grown on demand and reused across calls.
enum WsSlot {
  WsBuildKeys0,
  WsBuildKeys1,
  WsBuildVals0,
  WsBuildVals1,
  WsBuildAos,
  WsBuildStatus,
  WsBuildCtrl,
  WsBuildRuns,
  WsQFrontA,
  WsQFrontB,
  WsQStatus,
  WsQCtrl,
  WsQMask,
  WsQOff,
  WsQSafe,
  WsJoinWin,
  WsBatchCounts,
  WsBatchOut,
  WsSortKeys0,
  WsSortKeys1,
  WsSortVals0,
  WsSortVals1,
  WsSortStatus,
  WsSortCtrl,
  WsSlots
};

struct Workspace {
  void* p[WsSlots] = {};
  size_t cap[WsSlots] = {};
  int reserve(WsSlot s, size_t bytes);  // keeps contents only if no growth
  bool fits(WsSlot s, size_t bytes) const { return cap[s] >= (bytes ? bytes : 256); }
  void* ptr(WsSlot s) const { return p[s]; }
  void release();
};

// Optional per-kernel event timing.
struct Prof {
  bool on = false;
  std::string only;  // non-empty: bracket only the launches of this kernel
  struct Rec {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Rec> pending;
  std::map<std::string, std::pair<double, int64_t>> acc;
  hipStream_t open_stream = nullptr;
  const char* open_name = nullptr;
  hipEvent_t open_ev = nullptr;
  std::vector<hipEvent_t> pool;
  // Close the open interval (if any) and open one named `name` (nullptr: just close).
  void mark(hipStream_t s, const char* name);
  void resolve();
  hipEvent_t get_event();
};

// Device-side state of one query pipeline run (one chain of kernels: run
// begin, one expansion per internal level, leaf stage).  Every size lives
// here, so a run needs no host round trip and can be captured in a graph.
constexpr int kMaxRunSlots = 72;
// Level table of a bulk-loaded tree, leaf-parent level first: node i of level
// l (global index lstart[l] + i) covers sorted boxes [i*span[l], (i+1)*span[l]).
constexpr int kMaxLevels = 32;
struct QLevels {
  int nl;
  int64_t lstart[kMaxLevels];
  int64_t lcount[kMaxLevels];
  int64_t span[kMaxLevels];
};
struct QCtrl {
  uint64_t count[kMaxRunSlots];  // count[l]: entries of frontier l (count[0]: queries of the run)
  uint64_t total;                // pairs of the run
  unsigned long long npieces;    // long box ranges queued for the piece writer
  uint32_t overflow;             // nonzero: a frontier exceeded the buffer capacity
  uint32_t zero_needed;          // status words must be cleared before this run
  uint32_t epoch_base;           // chained-scan epoch of slot 0 of this run
  uint32_t epoch_next;           // first free epoch (kEpochs: clear before the next run)
  uint32_t chunk;                // next entry of the chunk plan (graph replays)
  uint32_t pad;
  int64_t q0, q1;                // query range of the run
  uint32_t ticket[kMaxRunSlots];
  QLevels lv;                    // levels of the run's tree (written by the run's begin)
};

// Work buffers of one query pipeline; one engine serves one stream at a time.
// One leaf entry between the leaf kernels: hit mask (partial leaf-parent) or
// range length (contained subtree), first sorted box, query | range flag.
struct alignas(16) QLeaf {
  uint64_t val;
  uint32_t first;
  uint32_t qf;
};
// One internal entry between the expansion kernels: hit mask and the mask of
// hits that are contained, well-formed subtrees, query, first child.
struct alignas(8) QExp {
  uint64_t hit;
  uint64_t full;
  uint32_t q;
  uint32_t first;
};
constexpr size_t kEntryBytes = sizeof(QExp) > sizeof(QLeaf) ? sizeof(QExp) : sizeof(QLeaf);
struct QEngine {
  void* front[2] = {nullptr, nullptr};  // (query, node) frontiers, uint2 entries
  int64_t cap_front = 0;                // entries per frontier buffer (and leaf entries)
  void* est = nullptr;                  // per-entry state between count and write kernels
  uint64_t* ttot = nullptr;             // pairs per leaf tile, then their exclusive scan
  void* pieces = nullptr;               // long box ranges of contained subtrees
  int64_t cap_pieces = 0;
  uint64_t* status = nullptr;           // chained-scan status words
  int64_t status_words = 0;
  QCtrl* ctrl = nullptr;
  QCtrl* host_ctrl = nullptr;           // pinned copy for read-back
  uint8_t* safe = nullptr;              // per-node well-formedness flags
  int64_t cap_safe = 0;
  int64_t* plan = nullptr;              // chunk plan (q0, q1) pairs for graph replays
  int64_t cap_plan = 0;
  float4* box4 = nullptr;               // AoS copies staged for one query: boxes,
  int64_t cap_box4 = 0;                 // node MBRs and windows
  float4* node4 = nullptr;
  int64_t cap_node4 = 0;
  float4* win4 = nullptr;
  int64_t cap_win4 = 0;
  uint32_t* canon = nullptr;            // the tree has the bulk-load layout (device word)
  // windows in a run order (large writing runs): the order, per-window pair
  // counts, each window's first pair in input and in run order, block sums
  void* order = nullptr;
  int64_t cap_order = 0;
  // fixed: every buffer above is a slice of memory the engine does not own
  // (the allocation-free entry point's scratch); a request beyond a buffer's
  // capacity fails with CURTREE_ERANGE instead of allocating
  bool fixed = false;
  bool wide = false;                    // tile counts hold one word per entry (fanout > 64)
  bool needs_clear = true;              // status/ctrl are fresh: clear on the next run
  int last_nl = 0;                      // depths of the last run read back (host_ctrl)
  int reserve_front(int64_t entries);
  int reserve_pieces(int64_t n);
  int reserve_status(int64_t words);
  int reserve_safe(int64_t nodes);
  int reserve_plan(int64_t chunks);
  int reserve_order(int64_t windows);
  int ensure_ctrl();
  void release();
  ~QEngine() { release(); }
};

// Resources of the batched entry point, owned by it alone: its chunk loop runs
// without the context lock so that chunk callbacks may call the library.
struct BatchState {
  std::mutex mu;  // one batched call at a time
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  QEngine qe;
  Workspace ws;   // WsBatchCounts, WsBatchOut
};

struct Context {
  std::mutex mu;
  Workspace ws;
  Prof prof;
  QEngine qe;            // engine of the plain (synchronous) query entry points
  BatchState batch;
  int opt_contain = -1;  // containment shortcut: -1 auto, 0 off, 1 on
  // recorded after every build's kernels: the next build, on any stream,
  // waits for it before it touches the shared build buffers
  hipEvent_t build_done = nullptr;
};

Context& context();

// alloc: allocate the node arrays (else use the caller's, rt->sz must match)
int build_f32(Context& ctx, curtree_bbox_f32* box, curtree_rtree_f32* rt, int key_mode,
              hipStream_t s, bool alloc = true);

// Query: run count (and, if fid != nullptr, write) passes.
struct QueryOut {
  int* fid = nullptr;         // caller buffers, or nullptr to allocate
  int* tid = nullptr;
  int64_t cap = 0;            // capacity of caller buffers
  bool alloc = false;         // allocate exact-size output
  int64_t* counts = nullptr;  // optional per-query counts (device)
  bool count_only = false;
  bool contain = true;        // use the containment shortcut (set by the entry point)
  int algo = CURTREE_ALGO_FRONTIER;  // CURTREE_ALGO_PERSISTENT: chained launches only
  // results
  int64_t n = 0;
  int* out_fid = nullptr;
  int* out_tid = nullptr;
};
int query_frontier_f32(Context& ctx, const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox,
                       const curtree_bbox_f32* qbox, QueryOut* qo, hipStream_t s);

// Same, on an explicit engine and stream, for queries [q0, q1) of qbox (pair
// tids stay global query indices).  contain: -1 auto, 0 off, 1 on.
// ctx (or null): the library context, whose sort buffers order the windows
// of large runs in Morton order of their centres (see ct_query.hip).
int query_run_f32(QEngine& qe, Prof& pf, const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox,
                  const curtree_bbox_f32* qbox, int64_t q0, int64_t q1, int contain, QueryOut* qo,
                  hipStream_t s, Context* ctx = nullptr);

int pairs_canonical_sort(Context& ctx, int* fid, int* tid, int64_t n, hipStream_t s);
// (query id, box id) order of any n < 2^30 pairs by a full radix sort of the
// 64-bit (tid, fid) keys (signed order of both); no host synchronisation
int pairs_full_sort(Context& ctx, int* fid, int* tid, int64_t n, hipStream_t s,
                    bool tid_only = false);

// query_into on the fixed scratch engine (no allocation; see ct_query.hip).
int query_into_noalloc_f32(const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox,
                           const curtree_bbox_f32* qbox, int* fid, int* tid, int64_t cap,
                           int64_t* n_out, hipStream_t s);

int query_batched_f32(Context& ctx, const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox,
                      const curtree_bbox_f32* qbox, int64_t budget, curtree_chunk_fn fn,
                      void* user, curtree_batch_stats* stats);
// Launch the control-block reset now (outside any graph capture).
void enqueue_ctrl_init(QEngine& qe, hipStream_t s);

int self_join_f32(Context& ctx, const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox, float r,
                  curtree_idpair64* out, hipStream_t s);

#define CT_HIP(call)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return ::ct::set_error(e_ == hipErrorOutOfMemory ? CURTREE_ENOMEM : CURTREE_EHIP,     \
                             "%s:%d %s: %s", __FILE__, __LINE__, #call,                     \
                             hipGetErrorString(e_));                                         \
  } while (0)

}  // namespace ct
