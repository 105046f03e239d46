// ct_capi.cpp -- the extern "C" entry points of libcurtree_amd.so and the host
// runtime pieces behind them (errors, level plan, work buffers, timing).
//
// Ownership follows the reference's helpers (src/utility.cuh:16-205): the
// caller owns box arrays, build allocates the node arrays, queries allocate the
// pair arrays; everything is hipMalloc'd and released with hipFree.  Errors are
// returned as status codes instead of the reference's exit() (utility.cuh:8-14).
#include <stdio.h>
#include <string.h>

#include "ct_query.hpp"

namespace ct {

static thread_local std::string g_err;

int set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

// Level plan of src/rtree_bulkload_thrust.cuh:101-115 with integer ceil-div.
// The reference's float form undercounts for some N > 2^24 and never ends for
// N == 0; here N == 0 is an empty tree and N == 1 a single root level.
void level_plan(int64_t n, int fanout, LevelPlan* lp) {
  lp->nl = 0;
  lp->r_sz = 0;
  if (n <= 0 || fanout < 2) return;
  int64_t len = n;
  do {
    int64_t nodes = (len + fanout - 1) / fanout;
    lp->n[lp->nl++] = nodes;
    lp->r_sz += nodes;
    len = nodes;
  } while (len != 1 && lp->nl < 64);
}

int rt_alloc(curtree_rtree_f32* rt, int64_t sz) {
  rt->sz = (int)sz;
  float** f[4] = {&rt->xmin, &rt->ymin, &rt->xmax, &rt->ymax};
  for (auto p : f) *p = nullptr;
  rt->pos = rt->len = nullptr;
  if (sz <= 0) return CURTREE_OK;
  for (auto p : f) {
    if (hipMalloc((void**)p, sz * sizeof(float)) != hipSuccess) {
      rt_free(rt);
      return set_error(CURTREE_ENOMEM, "cannot allocate %lld tree nodes", (long long)sz);
    }
  }
  if (hipMalloc((void**)&rt->pos, sz * sizeof(int)) != hipSuccess ||
      hipMalloc((void**)&rt->len, sz * sizeof(int)) != hipSuccess) {
    rt_free(rt);
    return set_error(CURTREE_ENOMEM, "cannot allocate %lld tree nodes", (long long)sz);
  }
  // (fresh arrays: an entry naming their addresses outlived arrays freed elsewhere)
  const void* ps[6] = {rt->xmin, rt->ymin, rt->xmax, rt->ymax, rt->pos, rt->len};
  tree_state_forget(ps, 6);
  return CURTREE_OK;
}

// ---- state of the trees built here ------------------------------------------
namespace {
struct TreeState {
  const void* arr[11];  // node arrays (6), then the sorted boxes (5)
  int64_t sz = 0, nbox = 0;
  int fanout = 0;
  uint32_t* words = nullptr;  // device: [0] every box is well-formed
  void* codes = nullptr;      // device: leaf codes (leaf_codes_bytes), or null
  int64_t codes_bytes = 0;
  bool codes_ready = false;
  bool overlaps(const char* p, int64_t bytes) const {
    for (int j = 0; j < 11; j++) {
      const char* a = (const char*)arr[j];
      const int64_t len = 4 * (j < 6 ? sz : nbox);
      if (a && p < a + len && a < p + bytes) return true;
    }
    return false;
  }
};
std::mutex g_state_mu;
std::vector<TreeState> g_states;
// words of forgotten entries, reused (never freed: 256 B each, and a free
// could stall an asynchronous build or break a graph capture)
std::vector<uint32_t*> g_spare_words;
// leaf-code buffers of forgotten entries (bytes), reused by later trees
std::vector<std::pair<void*, int64_t>> g_spare_codes;

// Moves an entry's words and codes to the spare lists.
void retire_locked(TreeState& e) {
  g_spare_words.push_back(e.words);
  if (e.codes) g_spare_codes.push_back({e.codes, e.codes_bytes});
  e.codes = nullptr;
  e.codes_bytes = 0;
  e.codes_ready = false;
}

TreeState state_key(const curtree_rtree_f32* rt, const curtree_bbox_f32* box) {
  TreeState k;
  const void* a[11] = {rt->xmin, rt->ymin, rt->xmax, rt->ymax, rt->pos, rt->len,
                       box->xmin, box->ymin, box->xmax, box->ymax, box->id};
  for (int i = 0; i < 11; i++) k.arr[i] = a[i];
  k.sz = rt->sz;
  k.nbox = box->sz;
  k.fanout = rt->fanout;
  return k;
}

// Drops (under g_state_mu) every entry that shares an array with `a`.
void forget_locked(const void* const* a, int n) {
  for (size_t i = 0; i < g_states.size();) {
    bool hit = false;
    for (int j = 0; j < 11 && !hit; j++)
      for (int k = 0; k < n && !hit; k++) hit = a[k] && g_states[i].arr[j] == a[k];
    if (hit) {
      retire_locked(g_states[i]);
      g_states[i] = g_states.back();
      g_states.pop_back();
    } else {
      i++;
    }
  }
}
}  // namespace

uint32_t* tree_state_register(const curtree_rtree_f32* rt, const curtree_bbox_f32* box,
                              bool may_alloc) {
  std::lock_guard<std::mutex> g(g_state_mu);
  TreeState k = state_key(rt, box);
  forget_locked(k.arr, 11);  // trees over these arrays are rebuilt or re-sorted
  if (g_spare_words.empty()) {
    if (!may_alloc) return nullptr;
    uint32_t* w = nullptr;
    if (hipMalloc((void**)&w, 256) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;  // not kept: queries compute the flags themselves
    }
    g_spare_words.push_back(w);
  }
  k.words = g_spare_words.back();
  g_spare_words.pop_back();
  g_states.push_back(k);
  return k.words;
}

namespace {
TreeState* find_locked(const curtree_rtree_f32* rt, const curtree_bbox_f32* box) {
  const TreeState k = state_key(rt, box);
  for (TreeState& e : g_states) {
    if (e.sz != k.sz || e.nbox != k.nbox || e.fanout != k.fanout) continue;
    bool same = true;
    for (int j = 0; j < 11 && same; j++) same = e.arr[j] == k.arr[j];
    if (same) return &e;
  }
  return nullptr;
}
}  // namespace

const uint32_t* tree_state_find(const curtree_rtree_f32* rt, const curtree_bbox_f32* box) {
  std::lock_guard<std::mutex> g(g_state_mu);
  const TreeState* e = find_locked(rt, box);
  return e ? e->words : nullptr;
}

const void* tree_state_codes(const curtree_rtree_f32* rt, const curtree_bbox_f32* box,
                             int64_t bytes, const std::function<int(void*)>* build) {
  std::lock_guard<std::mutex> g(g_state_mu);
  TreeState* e = find_locked(rt, box);
  if (!e) return nullptr;
  if (e->codes_ready) return e->codes;
  if (!build) return nullptr;
  if (!e->codes) {
    // the smallest spare buffer that holds them, else a new one
    size_t best = g_spare_codes.size();
    for (size_t i = 0; i < g_spare_codes.size(); i++)
      if (g_spare_codes[i].second >= bytes &&
          (best == g_spare_codes.size() || g_spare_codes[i].second < g_spare_codes[best].second))
        best = i;
    if (best < g_spare_codes.size()) {
      // a query of the tree that had it may still be reading it
      if (hipDeviceSynchronize() != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
      }
      e->codes = g_spare_codes[best].first;
      e->codes_bytes = g_spare_codes[best].second;
      g_spare_codes[best] = g_spare_codes.back();
      g_spare_codes.pop_back();
    } else {
      if (hipMalloc(&e->codes, (size_t)bytes) != hipSuccess) {
        (void)hipGetLastError();
        e->codes = nullptr;
        return nullptr;  // no codes: the boxes are tested directly
      }
      e->codes_bytes = bytes;
    }
  }
  if ((*build)(e->codes) != CURTREE_OK) return nullptr;
  e->codes_ready = true;
  return e->codes;
}

void tree_state_free_spare_codes() {
  std::lock_guard<std::mutex> g(g_state_mu);
  for (auto& sp : g_spare_codes) (void)hipFree(sp.first);
  g_spare_codes.clear();
}

void tree_state_forget(const void* const* arrays, int n) {
  std::lock_guard<std::mutex> g(g_state_mu);
  forget_locked(arrays, n);
}

void tree_state_forget_range(const void* p, int64_t bytes) {
  if (!p || bytes <= 0) return;
  std::lock_guard<std::mutex> g(g_state_mu);
  for (size_t i = 0; i < g_states.size();) {
    if (g_states[i].overlaps((const char*)p, bytes)) {
      retire_locked(g_states[i]);
      g_states[i] = g_states.back();
      g_states.pop_back();
    } else {
      i++;
    }
  }
}

void rt_free(curtree_rtree_f32* rt) {
  void* ps[6] = {rt->xmin, rt->ymin, rt->xmax, rt->ymax, rt->pos, rt->len};
  tree_state_forget(ps, 6);
  for (void* p : ps)
    if (p) (void)hipFree(p);
  rt->xmin = rt->ymin = rt->xmax = rt->ymax = nullptr;
  rt->pos = rt->len = nullptr;
  rt->sz = 0;
}

int Workspace::reserve(WsSlot s, size_t bytes) {
  if (bytes == 0) bytes = 256;
  if (cap[s] >= bytes) return CURTREE_OK;
  if (p[s]) (void)hipFree(p[s]);
  p[s] = nullptr;
  cap[s] = 0;
  size_t want = (bytes + 4095) & ~(size_t)4095;
  if (hipMalloc(&p[s], want) != hipSuccess) {
    (void)hipGetLastError();
    p[s] = nullptr;
    return set_error(CURTREE_ENOMEM, "cannot allocate %zu bytes of work memory", want);
  }
  cap[s] = want;
  return CURTREE_OK;
}

void Workspace::release() {
  for (int i = 0; i < WsSlots; i++) {
    if (p[i]) (void)hipFree(p[i]);
    p[i] = nullptr;
    cap[i] = 0;
  }
}

hipEvent_t Prof::get_event() {
  if (!pool.empty()) {
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

void Prof::mark(hipStream_t s, const char* name) {
  if (!on) return;
  if (open_name) {
    hipEvent_t b = get_event();
    (void)hipEventRecord(b, open_stream);
    pending.push_back({open_name, open_ev, b});
    open_name = nullptr;
    open_ev = nullptr;
  }
  if (name && (only.empty() || only == name)) {
    open_ev = get_event();
    (void)hipEventRecord(open_ev, s);
    open_stream = s;
    open_name = name;
  }
}

void Prof::resolve() {
  for (auto& r : pending) {
    (void)hipEventSynchronize(r.b);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, r.a, r.b);
    auto& a = acc[r.name];
    a.first += ms;
    a.second += 1;
    pool.push_back(r.a);
    pool.push_back(r.b);
  }
  pending.clear();
}

Context& context() {
  static Context* c = new Context();  // never destroyed: outlives static teardown
  return *c;
}

}  // namespace ct

using namespace ct;

static hipStream_t as_stream(void* s) { return (hipStream_t)s; }

extern "C" {

CURTREE_API const char* curtree_last_error(void) { return g_err.c_str(); }
CURTREE_API const char* curtree_version(void) { return "curtree_amd 0.1 (gfx950)"; }

CURTREE_API int curtree_build_f32_async(curtree_bbox_f32* box, curtree_rtree_f32* rt,
                                        int key_mode, void* stream) {
  Context& c = context();
  std::lock_guard<std::mutex> g(c.mu);
  return build_f32(c, box, rt, key_mode, as_stream(stream));
}

CURTREE_API int curtree_build_f32(curtree_bbox_f32* box, curtree_rtree_f32* rt, int key_mode,
                                  void* stream) {
  Context& c = context();
  std::lock_guard<std::mutex> g(c.mu);
  int rc = build_f32(c, box, rt, key_mode, as_stream(stream));
  if (rc) return rc;
  CT_HIP(hipStreamSynchronize(as_stream(stream)));
  return CURTREE_OK;
}

CURTREE_API int curtree_build_into_f32(curtree_bbox_f32* box, curtree_rtree_f32* rt,
                                       int key_mode, void* stream) {
  Context& c = context();
  std::lock_guard<std::mutex> g(c.mu);
  return build_f32(c, box, rt, key_mode, as_stream(stream), false);
}

CURTREE_API int curtree_tree_size(int64_t n, int fanout, int64_t* r_sz, int* levels) {
  if (n < 0 || fanout < 2) return set_error(CURTREE_EINVAL, "tree_size: bad argument");
  LevelPlan lp;
  level_plan(n, fanout, &lp);
  if (r_sz) *r_sz = lp.r_sz;
  if (levels) *levels = lp.nl;
  return CURTREE_OK;
}

static int check_algo(int algo) {
  if (algo != CURTREE_ALGO_FRONTIER && algo != CURTREE_ALGO_PERSISTENT)
    return set_error(CURTREE_EINVAL, "query: unknown algorithm %d", algo);
  return CURTREE_OK;
}

CURTREE_API int curtree_query_f32(const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox,
                                  const curtree_bbox_f32* qbox, curtree_idpair64* out, int algo,
                                  void* stream) {
  if (!out) return set_error(CURTREE_EINVAL, "query: null output");
  if (int rc = check_algo(algo)) return rc;
  out->sz = 0;
  out->fid = out->tid = nullptr;
  Context& c = context();
  std::lock_guard<std::mutex> g(c.mu);
  QueryOut qo;
  qo.alloc = true;
  qo.algo = algo;
  int rc = query_frontier_f32(c, rt, rbox, qbox, &qo, as_stream(stream));
  if (rc) return rc;
  CT_HIP(hipStreamSynchronize(as_stream(stream)));
  out->sz = qo.n;
  out->fid = qo.out_fid;
  out->tid = qo.out_tid;
  return CURTREE_OK;
}

CURTREE_API int curtree_query_into_f32(const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox,
                                       const curtree_bbox_f32* qbox, int* fid, int* tid,
                                       int64_t cap, int64_t* n_out, int algo, void* stream) {
  if (int rc = check_algo(algo)) return rc;
  if (cap > 0 && (!fid || !tid)) return set_error(CURTREE_EINVAL, "query: null output buffer");
  Context& c = context();
  std::lock_guard<std::mutex> g(c.mu);
  QueryOut qo;
  qo.fid = fid;
  qo.tid = tid;
  qo.cap = cap;
  qo.algo = algo;
  int rc = query_frontier_f32(c, rt, rbox, qbox, &qo, as_stream(stream));
  if (n_out) *n_out = qo.n;
  if (rc) return rc;
  CT_HIP(hipStreamSynchronize(as_stream(stream)));
  return CURTREE_OK;
}

CURTREE_API int curtree_query_into_noalloc_f32(const curtree_rtree_f32* rt,
                                               const curtree_bbox_f32* rbox,
                                               const curtree_bbox_f32* qbox, int* fid, int* tid,
                                               int64_t cap, int64_t* n_out, void* stream) {
  return query_into_noalloc_f32(rt, rbox, qbox, fid, tid, cap, n_out, as_stream(stream));
}

CURTREE_API int curtree_query_count_f32(const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox,
                                        const curtree_bbox_f32* qbox, int64_t* counts,
                                        int64_t* n_out, void* stream) {
  Context& c = context();
  std::lock_guard<std::mutex> g(c.mu);
  QueryOut qo;
  qo.count_only = true;
  qo.counts = counts;
  int rc = query_frontier_f32(c, rt, rbox, qbox, &qo, as_stream(stream));
  if (n_out) *n_out = qo.n;
  if (rc) return rc;
  CT_HIP(hipStreamSynchronize(as_stream(stream)));
  return CURTREE_OK;
}

CURTREE_API int curtree_query_batched_f32(const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox,
                                          const curtree_bbox_f32* qbox, int64_t budget_pairs,
                                          curtree_chunk_fn fn, void* user,
                                          curtree_batch_stats* stats) {
  Context& c = context();
  std::unique_lock<std::mutex> b(c.batch.mu, std::try_to_lock);
  if (!b.owns_lock())
    return set_error(CURTREE_EBUSY, "batched query already running (nested in a chunk callback?)");
  return query_batched_f32(c, rt, rbox, qbox, budget_pairs, fn, user, stats);
}

CURTREE_API int curtree_self_join_f32(const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox,
                                      float r, curtree_idpair64* out, void* stream) {
  Context& c = context();
  std::lock_guard<std::mutex> g(c.mu);
  int rc = self_join_f32(c, rt, rbox, r, out, as_stream(stream));
  if (rc) return rc;
  CT_HIP(hipStreamSynchronize(as_stream(stream)));
  return CURTREE_OK;
}

CURTREE_API int curtree_pairs_canonical_sort(int* fid, int* tid, int64_t n, void* stream) {
  Context& c = context();
  std::lock_guard<std::mutex> g(c.mu);
  int rc = pairs_canonical_sort(c, fid, tid, n, as_stream(stream));
  if (rc) return rc;
  CT_HIP(hipStreamSynchronize(as_stream(stream)));
  return CURTREE_OK;
}

// ---- memory helpers -------------------------------------------------------

CURTREE_API int curtree_box_d_alloc(curtree_bbox_f32* d, int sz) {
  if (!d || sz < 0) return set_error(CURTREE_EINVAL, "box_d_alloc: bad argument");
  d->sz = sz;
  d->xmin = d->ymin = d->xmax = d->ymax = nullptr;
  d->id = nullptr;
  if (sz == 0) return CURTREE_OK;
  float** f[4] = {&d->xmin, &d->ymin, &d->xmax, &d->ymax};
  for (auto p : f)
    if (hipMalloc((void**)p, (size_t)sz * sizeof(float)) != hipSuccess) {
      curtree_box_d_free(d);
      return set_error(CURTREE_ENOMEM, "box_d_alloc: cannot allocate %d boxes", sz);
    }
  if (hipMalloc((void**)&d->id, (size_t)sz * sizeof(int)) != hipSuccess) {
    curtree_box_d_free(d);
    return set_error(CURTREE_ENOMEM, "box_d_alloc: cannot allocate %d boxes", sz);
  }
  const void* ps[5] = {d->xmin, d->ymin, d->xmax, d->ymax, d->id};
  tree_state_forget(ps, 5);
  return CURTREE_OK;
}

CURTREE_API int curtree_box_d_free(curtree_bbox_f32* d) {
  if (!d) return CURTREE_OK;
  void* ps[5] = {d->xmin, d->ymin, d->xmax, d->ymax, d->id};
  tree_state_forget(ps, 5);
  for (void* p : ps)
    if (p) (void)hipFree(p);
  d->xmin = d->ymin = d->xmax = d->ymax = nullptr;
  d->id = nullptr;
  return CURTREE_OK;
}

CURTREE_API int curtree_box_h2d(curtree_bbox_f32* d, const curtree_bbox_f32* h) {
  if (!d || !h || h->sz < 0) return set_error(CURTREE_EINVAL, "box_h2d: bad argument");
  if (int rc = curtree_box_d_alloc(d, h->sz)) return rc;
  const size_t b = (size_t)h->sz * 4;
  if (!b) return CURTREE_OK;
  CT_HIP(hipMemcpy(d->xmin, h->xmin, b, hipMemcpyHostToDevice));
  CT_HIP(hipMemcpy(d->ymin, h->ymin, b, hipMemcpyHostToDevice));
  CT_HIP(hipMemcpy(d->xmax, h->xmax, b, hipMemcpyHostToDevice));
  CT_HIP(hipMemcpy(d->ymax, h->ymax, b, hipMemcpyHostToDevice));
  CT_HIP(hipMemcpy(d->id, h->id, b, hipMemcpyHostToDevice));
  return CURTREE_OK;
}

CURTREE_API int curtree_box_d2h(curtree_bbox_f32* h, const curtree_bbox_f32* d) {
  if (!d || !h || d->sz < 0) return set_error(CURTREE_EINVAL, "box_d2h: bad argument");
  h->sz = d->sz;
  const size_t b = (size_t)d->sz * 4;
  if (!b) return CURTREE_OK;
  CT_HIP(hipMemcpy(h->xmin, d->xmin, b, hipMemcpyDeviceToHost));
  CT_HIP(hipMemcpy(h->ymin, d->ymin, b, hipMemcpyDeviceToHost));
  CT_HIP(hipMemcpy(h->xmax, d->xmax, b, hipMemcpyDeviceToHost));
  CT_HIP(hipMemcpy(h->ymax, d->ymax, b, hipMemcpyDeviceToHost));
  CT_HIP(hipMemcpy(h->id, d->id, b, hipMemcpyDeviceToHost));
  return CURTREE_OK;
}

CURTREE_API int curtree_rt_d_alloc(curtree_rtree_f32* rt, int sz) {
  if (!rt || sz < 0) return set_error(CURTREE_EINVAL, "rt_d_alloc: bad argument");
  return rt_alloc(rt, sz);
}

CURTREE_API int curtree_rt_d_free(curtree_rtree_f32* rt) {
  if (rt) rt_free(rt);
  return CURTREE_OK;
}

CURTREE_API int curtree_idpair_d_alloc(curtree_idpair* p, int sz) {
  if (!p || sz < 0) return set_error(CURTREE_EINVAL, "idpair_d_alloc: bad argument");
  p->sz = sz;
  p->fid = p->tid = nullptr;
  if (!sz) return CURTREE_OK;
  CT_HIP(hipMalloc((void**)&p->fid, (size_t)sz * sizeof(int)));
  CT_HIP(hipMalloc((void**)&p->tid, (size_t)sz * sizeof(int)));
  return CURTREE_OK;
}

CURTREE_API int curtree_idpair_d_free(curtree_idpair* p) {
  if (!p) return CURTREE_OK;
  if (p->fid) (void)hipFree(p->fid);
  if (p->tid) (void)hipFree(p->tid);
  p->fid = p->tid = nullptr;
  return CURTREE_OK;
}

CURTREE_API int curtree_idpair64_d_free(curtree_idpair64* p) {
  if (!p) return CURTREE_OK;
  if (p->fid) (void)hipFree(p->fid);
  if (p->tid) (void)hipFree(p->tid);
  p->fid = p->tid = nullptr;
  return CURTREE_OK;
}

CURTREE_API int curtree_idpair64_d2h(int* h_fid, int* h_tid, const curtree_idpair64* d) {
  if (!d || d->sz < 0) return set_error(CURTREE_EINVAL, "idpair_d2h: bad argument");
  if (!d->sz) return CURTREE_OK;
  CT_HIP(hipMemcpy(h_fid, d->fid, (size_t)d->sz * sizeof(int), hipMemcpyDeviceToHost));
  CT_HIP(hipMemcpy(h_tid, d->tid, (size_t)d->sz * sizeof(int), hipMemcpyDeviceToHost));
  return CURTREE_OK;
}

CURTREE_API int curtree_malloc(void** p, int64_t bytes) {
  if (!p || bytes < 0) return set_error(CURTREE_EINVAL, "malloc: bad argument");
  *p = nullptr;
  if (!bytes) return CURTREE_OK;
  CT_HIP(hipMalloc(p, (size_t)bytes));
  const void* fresh[1] = {*p};
  tree_state_forget(fresh, 1);
  return CURTREE_OK;
}

CURTREE_API int curtree_free(void* p) {
  const void* gone[1] = {p};
  tree_state_forget(gone, 1);
  if (p) CT_HIP(hipFree(p));
  return CURTREE_OK;
}

CURTREE_API int curtree_memcpy_h2d(void* dst, const void* src, int64_t bytes) {
  tree_state_forget_range(dst, bytes);  // a kept tree's boxes may change
  if (bytes > 0) CT_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
  return CURTREE_OK;
}

CURTREE_API int curtree_memcpy_d2h(void* dst, const void* src, int64_t bytes) {
  if (bytes > 0) CT_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
  return CURTREE_OK;
}

CURTREE_API int curtree_memcpy_d2d(void* dst, const void* src, int64_t bytes, void* stream) {
  tree_state_forget_range(dst, bytes);
  if (bytes > 0)
    CT_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
  return CURTREE_OK;
}

CURTREE_API int curtree_memset(void* dst, int value, int64_t bytes) {
  tree_state_forget_range(dst, bytes);
  if (bytes > 0) CT_HIP(hipMemset(dst, value, (size_t)bytes));
  return CURTREE_OK;
}

CURTREE_API int curtree_synchronize(void* stream) {
  if (stream) CT_HIP(hipStreamSynchronize(as_stream(stream)));
  else CT_HIP(hipDeviceSynchronize());
  return CURTREE_OK;
}

CURTREE_API int curtree_stream_create(void** stream) {
  hipStream_t s = nullptr;
  CT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = (void*)s;
  return CURTREE_OK;
}

CURTREE_API int curtree_stream_destroy(void* stream) {
  if (stream) CT_HIP(hipStreamDestroy(as_stream(stream)));
  return CURTREE_OK;
}

CURTREE_API int curtree_device_info(int* count, char* name, int name_len, int* cus,
                                    int64_t* hbm_bytes) {
  int n = 0;
  CT_HIP(hipGetDeviceCount(&n));
  if (count) *count = n;
  if (n <= 0) return CURTREE_OK;
  hipDeviceProp_t p;
  CT_HIP(hipGetDeviceProperties(&p, 0));
  if (name && name_len > 0) {
    snprintf(name, (size_t)name_len, "%s", p.gcnArchName);
  }
  if (cus) *cus = p.multiProcessorCount;
  if (hbm_bytes) *hbm_bytes = (int64_t)p.totalGlobalMem;
  return CURTREE_OK;
}

CURTREE_API int curtree_release_workspace(void) {
  Context& c = context();
  std::lock_guard<std::mutex> g(c.mu);
  std::unique_lock<std::mutex> b(c.batch.mu, std::try_to_lock);
  if (!b.owns_lock()) return set_error(CURTREE_EBUSY, "release_workspace: batched query running");
  CT_HIP(hipDeviceSynchronize());
  c.ws.release();
  c.qe.release();
  c.batch.ws.release();
  c.batch.qe.release();
  tree_state_free_spare_codes();
  return CURTREE_OK;
}

CURTREE_API int curtree_query_last_stats(int64_t* entries, int cap, int* levels, int64_t* pairs,
                                         int64_t* pieces) {
  Context& c = context();
  std::lock_guard<std::mutex> g(c.mu);
  const int nl = c.qe.host_ctrl ? c.qe.last_nl : 0;
  if (levels) *levels = nl;
  for (int l = 0; l < nl && l < cap && entries; l++) entries[l] = (int64_t)c.qe.host_ctrl->count[l];
  if (pairs) *pairs = nl ? (int64_t)c.qe.host_ctrl->total : 0;
  if (pieces) *pieces = nl ? (int64_t)c.qe.host_ctrl->npieces : 0;
  return CURTREE_OK;
}

CURTREE_API int curtree_set_option(const char* name, int64_t value) {
  Context& c = context();
  std::lock_guard<std::mutex> g(c.mu);
  if (!name) return set_error(CURTREE_EINVAL, "set_option: null name");
  if (!strcmp(name, "containment")) {
    if (value < -1 || value > 1) return set_error(CURTREE_EINVAL, "containment: -1, 0 or 1");
    c.opt_contain = (int)value;
    return CURTREE_OK;
  }
  if (!strcmp(name, "query_tile_steps")) {
    if (value != 4 && value != 8 && value != 16)
      return set_error(CURTREE_EINVAL, "query_tile_steps: 4, 8 or 16");
    qtuning().steps = (int)value;
    return CURTREE_OK;
  }
  if (!strcmp(name, "box_staging")) {
    if (value < -1 || value > 1) return set_error(CURTREE_EINVAL, "box_staging: -1, 0 or 1");
    qtuning().staging = (int)value;
    return CURTREE_OK;
  }
  if (!strcmp(name, "expand_split")) {
    if (value < -1 || value > 1) return set_error(CURTREE_EINVAL, "expand_split: -1, 0 or 1");
    qtuning().expand = (int)value;
    return CURTREE_OK;
  }
  if (!strcmp(name, "leaf_fused")) {
    if (value < -1 || value > 1) return set_error(CURTREE_EINVAL, "leaf_fused: -1, 0 or 1");
    qtuning().fused = (int)value;
    return CURTREE_OK;
  }
  if (!strcmp(name, "window_order")) {
    if (value < -1 || value > 1) return set_error(CURTREE_EINVAL, "window_order: -1, 0 or 1");
    qtuning().order = (int)value;
    return CURTREE_OK;
  }
  if (!strcmp(name, "leaf_codes")) {
    if (value < -1 || value > 1) return set_error(CURTREE_EINVAL, "leaf_codes: -1, 0 or 1");
    qtuning().codes = (int)value;
    return CURTREE_OK;
  }
  if (!strcmp(name, "tile_xcd")) {
    if (value < 0 || value > 1) return set_error(CURTREE_EINVAL, "tile_xcd: 0 or 1");
    qtuning().xcd = (int)value;
    return CURTREE_OK;
  }
  if (!strcmp(name, "query_grid")) {
    if (value < 0 || value > 65535) return set_error(CURTREE_EINVAL, "query_grid: 0..65535");
    qtuning().grid = (int)value;
    return CURTREE_OK;
  }
  return set_error(CURTREE_EINVAL, "set_option: unknown option %s", name);
}

CURTREE_API int curtree_prof_enable(int on) {
  Context& c = context();
  std::lock_guard<std::mutex> g(c.mu);
  c.prof.on = on != 0;
  return CURTREE_OK;
}

CURTREE_API int curtree_prof_only(const char* name) {
  Context& c = context();
  std::lock_guard<std::mutex> g(c.mu);
  c.prof.only = name ? name : "";
  return CURTREE_OK;
}

CURTREE_API int curtree_prof_reset(void) {
  Context& c = context();
  std::lock_guard<std::mutex> g(c.mu);
  c.prof.resolve();
  c.prof.acc.clear();
  return CURTREE_OK;
}

CURTREE_API int curtree_prof_get(const char* name, double* total_ms, int64_t* launches) {
  Context& c = context();
  std::lock_guard<std::mutex> g(c.mu);
  c.prof.resolve();
  auto it = c.prof.acc.find(name ? name : "");
  if (total_ms) *total_ms = it == c.prof.acc.end() ? 0.0 : it->second.first;
  if (launches) *launches = it == c.prof.acc.end() ? 0 : it->second.second;
  return CURTREE_OK;
}

CURTREE_API int curtree_event_create(void** ev) {
  hipEvent_t e = nullptr;
  CT_HIP(hipEventCreate(&e));
  *ev = (void*)e;
  return CURTREE_OK;
}

CURTREE_API int curtree_event_destroy(void* ev) {
  if (ev) CT_HIP(hipEventDestroy((hipEvent_t)ev));
  return CURTREE_OK;
}

CURTREE_API int curtree_event_record(void* ev, void* stream) {
  CT_HIP(hipEventRecord((hipEvent_t)ev, as_stream(stream)));
  return CURTREE_OK;
}

CURTREE_API int curtree_event_elapsed(void* ev0, void* ev1, float* ms) {
  CT_HIP(hipEventSynchronize((hipEvent_t)ev1));
  CT_HIP(hipEventElapsedTime(ms, (hipEvent_t)ev0, (hipEvent_t)ev1));
  return CURTREE_OK;
}

}  // extern "C"
