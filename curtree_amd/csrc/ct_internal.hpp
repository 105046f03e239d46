// ct_internal.hpp -- host-side internals shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>

#include <algorithm>
#include <map>
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

// Cached device work buffers, grown on demand and reused across calls.
enum WsSlot {
  WsBuildKeys0,
  WsBuildKeys1,
  WsBuildVals0,
  WsBuildVals1,
  WsBuildAos,
  WsBuildStatus,
  WsBuildCtrl,
  WsQFrontA,
  WsQFrontB,
  WsQStatus,
  WsQCtrl,
  WsQMask,
  WsQOff,
  WsQSafe,
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
  void* ptr(WsSlot s) const { return p[s]; }
  void release();
};

// Optional per-kernel event timing.
struct Prof {
  bool on = false;
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

struct Context {
  std::mutex mu;
  Workspace ws;
  Prof prof;
  uint32_t qepoch = 0;  // query chained-scan epoch, cycles through [0, kEpochs)
  size_t qstatus_words = 0;
  int qfanout = 0;       // fanout of the tree being queried
  int opt_contain = -1;  // containment shortcut: -1 auto, 0 off, 1 on
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
  // results
  int64_t n = 0;
  int* out_fid = nullptr;
  int* out_tid = nullptr;
};
int query_frontier_f32(Context& ctx, const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox,
                       const curtree_bbox_f32* qbox, QueryOut* qo, hipStream_t s);

int pairs_canonical_sort(Context& ctx, int* fid, int* tid, int64_t n, hipStream_t s);

#define CT_HIP(call)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return ::ct::set_error(e_ == hipErrorOutOfMemory ? CURTREE_ENOMEM : CURTREE_EHIP,     \
                             "%s:%d %s: %s", __FILE__, __LINE__, #call,                     \
                             hipGetErrorString(e_));                                         \
  } while (0)

}  // namespace ct
