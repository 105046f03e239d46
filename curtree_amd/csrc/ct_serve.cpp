// ct_serve.cpp -- serving: several bulk-loaded trees resident in HBM, window
// requests of any size submitted concurrently, continuous batching.
//
// A dispatcher thread owns S slots (stream + query engine + window staging).
// Whenever a slot is free it takes the oldest queued request, adds every other
// queued request on the same tree (up to a window cap), counts the batch's
// pairs on the slot's stream, reserves that many pairs in the result pool
// (the memory budget; a batch waits while the pool is full), and enqueues the
// writing run.  Slots run concurrently on their own streams, so requests that
// arrive while others are on the device are admitted as slots and memory free
// up and their work overlaps.  A request's pairs are a contiguous slice of its
// batch's output; their query indices are rebased to the request, so the
// result is the same alone and in any mix.  Results stay in the pool until the
// client releases them.  With batching off each request runs alone on one
// slot: the plain per-request launches the serving mode is compared with.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <unordered_map>

#include "ct_query.hpp"

namespace ct {

// One staging / rebasing job per request of a batch.
struct ServeJob {
  const float* src[4];  // the request's windows
  int64_t n;            // windows (stage) or pairs (rebase)
  int64_t dst;          // first staged window / first pool pair
  int64_t base;         // window offset of the request in its batch
};

// Copy every request's windows into the slot's staging arrays (one block per
// request, grid-stride over its windows).
__global__ void k_stage(const ServeJob* __restrict__ jobs, int njobs, float* __restrict__ win,
                        int64_t cap) {
  for (int j = blockIdx.x; j < njobs; j += gridDim.x) {
    const ServeJob jb = jobs[j];
    for (int64_t i = threadIdx.x; i < jb.n; i += 256) {
#pragma unroll
      for (int k = 0; k < 4; k++) win[k * cap + jb.dst + i] = jb.src[k][i];
    }
  }
}

// Query indices of each request's pairs relative to the request.
__global__ void k_rebase(const ServeJob* __restrict__ jobs, int njobs, int* __restrict__ tid) {
  for (int j = blockIdx.x; j < njobs; j += gridDim.x) {
    const ServeJob jb = jobs[j];
    for (int64_t i = threadIdx.x; i < jb.n; i += 256) tid[jb.dst + i] -= (int)jb.base;
  }
}

namespace {

double now_us() {
  using namespace std::chrono;
  return duration_cast<duration<double, std::micro>>(steady_clock::now().time_since_epoch())
      .count();
}

// First-fit allocator over the result pool (units: pairs).
struct Pool {
  int64_t cap = 0;
  std::map<int64_t, int64_t> free_;  // offset -> length
  void init(int64_t c) {
    cap = c;
    free_.clear();
    if (c > 0) free_[0] = c;
  }
  int64_t alloc(int64_t n) {
    if (n == 0) return 0;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        int64_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > n) free_[off + n] = len - n;
        return off;
      }
    }
    return -1;
  }
  void release(int64_t off, int64_t n) {
    if (n == 0) return;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

struct Request {
  int64_t id = 0;
  int tree = 0;
  curtree_bbox_f32 win{};
  double t_submit = 0, t_done = 0;
  int state = CURTREE_REQ_QUEUED;
  int err = 0;
  int64_t pool_off = 0, npairs = 0;
};

struct Tree {
  curtree_rtree_f32 rt{};
  curtree_bbox_f32 rbox{};
  uint8_t* safe = nullptr;  // owned: containment flags computed once
  float4* box4 = nullptr;   // owned: AoS copies of the boxes and node MBRs
  float4* node4 = nullptr;
  uint32_t* canon = nullptr;  // owned: bulk-load layout flag (device word)
};

struct Slot {
  hipStream_t s = nullptr;
  hipEvent_t done = nullptr;
  QEngine qe;
  float* win = nullptr;  // staging: 4 arrays of win_cap windows
  int64_t win_cap = 0;
  int64_t* counts = nullptr;  // per-window pair counts
  std::vector<int64_t> hcounts;
  ServeJob* jobs = nullptr;     // device descriptors (stage, then rebase)
  ServeJob* hjobs = nullptr;    // pinned host copy
  int jobs_cap = 0;
  bool busy = false;
  std::vector<int64_t> reqs;
  int64_t pool_off = 0, pool_len = 0;
};

}  // namespace

struct Server {
  curtree_server_opts opt{};
  std::vector<Tree> trees;
  std::vector<Slot> slots;
  std::unordered_map<int64_t, Request> reqs;
  std::deque<int64_t> queue;
  Pool pool;
  int* pool_fid = nullptr;
  int* pool_tid = nullptr;
  Prof prof;  // never enabled
  std::mutex mu;
  std::condition_variable cv;
  std::thread th;
  bool stop = false;
  int64_t next_id = 1;
  std::vector<double> lat_us;
  int64_t pairs_done = 0, batches = 0, requests_done = 0;
  uint64_t pool_gen = 0, busy_gen = ~0ull;  // releases so far / at the last full-pool stall
  double t_first = 0, t_last = 0;
  std::string fatal;

  int start();
  void loop();
  bool harvest();
  bool launch_one();
  int run_batch(Slot& sl, std::vector<int64_t>& batch, int64_t* pool_off, int64_t* pool_len);
};

int Server::start() {
  int rc;
  slots.resize((size_t)opt.slots);
  const int64_t cap = std::max<int64_t>(opt.max_batch_windows, 1);
  for (auto& sl : slots) {
    CT_HIP(hipStreamCreateWithFlags(&sl.s, hipStreamNonBlocking));
    CT_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    CT_HIP(hipMalloc((void**)&sl.win, (size_t)cap * 16));
    CT_HIP(hipMalloc((void**)&sl.counts, (size_t)cap * 8));
    sl.win_cap = cap;
    sl.hcounts.resize((size_t)cap);
    sl.jobs_cap = 4096;
    CT_HIP(hipMalloc((void**)&sl.jobs, (size_t)sl.jobs_cap * 2 * sizeof(ServeJob)));
    CT_HIP(hipHostMalloc((void**)&sl.hjobs, (size_t)sl.jobs_cap * 2 * sizeof(ServeJob),
                         hipHostMallocDefault));
    if ((rc = sl.qe.ensure_ctrl())) return rc;
  }
  const int64_t pairs = std::max<int64_t>(opt.budget_bytes / 8, 1);
  CT_HIP(hipMalloc((void**)&pool_fid, (size_t)pairs * 4));
  CT_HIP(hipMalloc((void**)&pool_tid, (size_t)pairs * 4));
  pool.init(pairs);
  th = std::thread([this] { loop(); });
  return CURTREE_OK;
}

// Completed slots -> their requests are done, their pool space stays with the
// requests until released.
bool Server::harvest() {
  bool any = false;
  for (auto& sl : slots) {
    if (!sl.busy) continue;
    hipError_t e = hipEventQuery(sl.done);
    if (e == hipErrorNotReady) continue;
    const double t = now_us();
    for (int64_t id : sl.reqs) {
      auto it = reqs.find(id);
      if (it == reqs.end()) continue;
      Request& r = it->second;
      if (e != hipSuccess) {
        r.state = CURTREE_REQ_FAILED;
        r.err = CURTREE_EHIP;
      } else {
        r.state = CURTREE_REQ_DONE;
      }
      r.t_done = t;
      lat_us.push_back(t - r.t_submit);
      pairs_done += r.npairs;
      requests_done++;
      t_last = t;
    }
    sl.reqs.clear();
    sl.busy = false;
    any = true;
  }
  return any;
}

// Count, reserve, write: the batch's requests (same tree) on slot `sl`.  The
// host waits for the count only; the write runs on while other slots start.
int Server::run_batch(Slot& sl, std::vector<int64_t>& batch, int64_t* pool_off,
                      int64_t* pool_len) {
  Tree& tr = trees[(size_t)reqs[batch[0]].tree];
  const int nj = (int)batch.size();
  int64_t nwin = 0;
  for (int j = 0; j < nj; j++) {
    const Request& r = reqs[batch[(size_t)j]];
    ServeJob& jb = sl.hjobs[j];
    jb.src[0] = r.win.xmin;
    jb.src[1] = r.win.ymin;
    jb.src[2] = r.win.xmax;
    jb.src[3] = r.win.ymax;
    jb.n = r.win.sz;
    jb.dst = nwin;
    jb.base = nwin;
    nwin += r.win.sz;
  }
  CT_HIP(hipMemcpyAsync(sl.jobs, sl.hjobs, (size_t)nj * sizeof(ServeJob), hipMemcpyHostToDevice,
                        sl.s));
  hipLaunchKernelGGL(k_stage, dim3((unsigned)std::min(nj, 1024)), dim3(256), 0, sl.s,
                     (const ServeJob*)sl.jobs, nj, sl.win, sl.win_cap);
  curtree_bbox_f32 q{(int)nwin, sl.win, sl.win + sl.win_cap, sl.win + 2 * sl.win_cap,
                     sl.win + 3 * sl.win_cap, nullptr};
  // count pass with per-window counts (the tree's cached containment flags)
  QRunArgs a;
  int rc;
  if ((rc = prepare_query(sl.qe, prof, &tr.rt, &tr.rbox, &q, 0, nwin, &a, sl.s))) return rc;
  a.tr.safe = tr.safe;
  a.rb.b4 = tr.box4;
  a.tr.mbr.b4 = tr.node4;
  a.tr.canon = tr.canon;
  a.q0 = 0;
  a.q1 = nwin;
  a.write = false;
  a.counts = (unsigned long long*)sl.counts;
  if ((rc = sl.qe.reserve_front(std::max<int64_t>(2 * nwin, 1 << 16)))) return rc;
  for (int attempt = 0;; attempt++) {
    enqueue_fill_counts(sl.counts, nwin, sl.s);
    if ((rc = sl.qe.reserve_status(std::max<int64_t>(sl.qe.cap_front, nwin) / per_tile(a.G, a.steps) + 2)))
      return rc;
    enqueue_run(sl.qe, prof, a, sl.s);
    CT_HIP(hipGetLastError());
    if ((rc = read_ctrl(sl.qe, sl.s))) return rc;
    if (!sl.qe.host_ctrl->overflow) break;
    if (attempt >= 3) return set_error(CURTREE_ENOMEM, "serve: frontier kept overflowing");
    uint64_t need = 0;
    for (int l = 1; l < a.nl; l++) need = std::max<uint64_t>(need, sl.qe.host_ctrl->count[l]);
    if ((rc = sl.qe.reserve_front((int64_t)(need + need / 4 + 1024)))) return rc;
  }
  const int64_t H = (int64_t)sl.qe.host_ctrl->total;
  if (H > pool.cap)
    return set_error(CURTREE_ERANGE, "serve: %lld pairs exceed the result budget", (long long)H);
  int64_t off = 0;
  if (H > 0) {
    off = pool.alloc(H);
    if (off < 0) return CURTREE_EBUSY;  // wait for released results
  }
  CT_HIP(hipMemcpyAsync(sl.hcounts.data(), sl.counts, (size_t)nwin * 8, hipMemcpyDeviceToHost,
                        sl.s));
  CT_HIP(hipStreamSynchronize(sl.s));
  if (H > 0) {
    // write pass into the reserved pool slice
    a.write = true;
    a.counts = nullptr;
    a.out_fid = pool_fid + off;
    a.out_tid = pool_tid + off;
    a.out_cap = H;
    if ((rc = sl.qe.reserve_pieces(pieces_needed(sl.qe, a)))) {
      pool.release(off, H);
      return rc;
    }
    enqueue_run(sl.qe, prof, a, sl.s);
  }
  // per-request slices of the batch output; query indices rebased per request
  int64_t p0 = 0;
  int nr = 0;
  ServeJob* rj = sl.hjobs + sl.jobs_cap;
  for (int j = 0; j < nj; j++) {
    Request& r = reqs[batch[(size_t)j]];
    const int64_t w0 = sl.hjobs[j].dst;
    int64_t np = 0;
    for (int64_t i = 0; i < r.win.sz; i++) np += sl.hcounts[(size_t)(w0 + i)];
    r.pool_off = off + p0;
    r.npairs = np;
    r.state = CURTREE_REQ_RUNNING;
    if (np > 0 && w0 > 0) {
      ServeJob& jb = rj[nr++];
      jb.n = np;
      jb.dst = off + p0;
      jb.base = w0;
    }
    p0 += np;
  }
  if (nr > 0) {
    CT_HIP(hipMemcpyAsync(sl.jobs + sl.jobs_cap, rj, (size_t)nr * sizeof(ServeJob),
                          hipMemcpyHostToDevice, sl.s));
    hipLaunchKernelGGL(k_rebase, dim3((unsigned)std::min(nr, 1024)), dim3(256), 0, sl.s,
                       (const ServeJob*)(sl.jobs + sl.jobs_cap), nr, pool_tid);
  }
  CT_HIP(hipEventRecord(sl.done, sl.s));
  *pool_off = off;
  *pool_len = H;
  return CURTREE_OK;
}

// Admit the next batch onto a free slot.  Returns true if something launched.
bool Server::launch_one() {
  if (queue.empty()) return false;
  if (busy_gen == pool_gen) return false;  // the pool was full; nothing released since
  Slot* free_slot = nullptr;
  for (auto& sl : slots)
    if (!sl.busy) {
      free_slot = &sl;
      break;
    }
  if (!free_slot) return false;
  // the oldest request decides the tree; same-tree requests join it
  std::vector<int64_t> batch;
  const int tree = reqs[queue.front()].tree;
  int64_t nwin = 0;
  for (auto it = queue.begin(); it != queue.end();) {
    const Request& r = reqs[*it];
    const bool first = batch.empty();
    if (r.tree == tree && (first || (opt.batching && nwin + r.win.sz <= free_slot->win_cap &&
                                     (int)batch.size() < free_slot->jobs_cap))) {
      batch.push_back(*it);
      nwin += r.win.sz;
      it = queue.erase(it);
      if (!opt.batching) break;
    } else {
      ++it;
    }
  }
  int64_t off = 0, len = 0;
  int rc = run_batch(*free_slot, batch, &off, &len);
  if (rc == CURTREE_EBUSY) {
    // not enough result memory now: requeue in front, keep their order
    for (auto it = batch.rbegin(); it != batch.rend(); ++it) queue.push_front(*it);
    busy_gen = pool_gen;
    return false;
  }
  if (rc) {
    for (int64_t id : batch) {
      Request& r = reqs[id];
      r.state = CURTREE_REQ_FAILED;
      r.err = rc;
      r.t_done = now_us();
    }
    return true;
  }
  free_slot->busy = true;
  free_slot->reqs = batch;
  batches++;
  return true;
}

void Server::loop() {
  std::unique_lock<std::mutex> lk(mu);
  while (!stop) {
    bool progressed = harvest();
    while (launch_one()) progressed = true;
    bool busy = false;
    for (auto& sl : slots) busy |= sl.busy;
    if (!progressed) {
      if (busy) {
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(20));
        lk.lock();
      } else {
        cv.wait_for(lk, std::chrono::milliseconds(5));
      }
    }
  }
}

}  // namespace ct

using namespace ct;

extern "C" {

CURTREE_API int curtree_server_create(const curtree_server_opts* opts, curtree_server** out) {
  if (!out) return set_error(CURTREE_EINVAL, "server_create: null output");
  curtree_server_opts o{4, (int64_t)16 << 30, 1 << 20, 1};
  if (opts) o = *opts;
  if (o.slots < 1 || o.slots > 16 || o.budget_bytes < 8 || o.max_batch_windows < 1)
    return set_error(CURTREE_EINVAL, "server_create: bad options");
  Server* sv = new Server();
  sv->opt = o;
  int rc = sv->start();
  if (rc) {
    curtree_server_destroy((curtree_server*)sv);
    return rc;
  }
  *out = (curtree_server*)sv;
  return CURTREE_OK;
}

CURTREE_API int curtree_server_add_tree(curtree_server* s, const curtree_rtree_f32* rt,
                                        const curtree_bbox_f32* rbox, int* tree_id) {
  Server* sv = (Server*)s;
  if (!sv || !rt || !rbox || !tree_id) return set_error(CURTREE_EINVAL, "server_add_tree: null");
  LevelPlan lp;
  level_plan(rbox->sz, rt->fanout, &lp);
  if (rt->fanout < 2 || rt->fanout > 64 || lp.r_sz != rt->sz)
    return set_error(CURTREE_EINVAL, "server_add_tree: tree does not match its boxes");
  Tree t;
  t.rt = *rt;
  t.rbox = *rbox;
  if (rbox->sz > 0) {
    // containment flags once per resident tree, on a private engine
    QEngine qe;
    Prof pf;
    QBoxes rb{rbox->xmin, rbox->ymin, rbox->xmax, rbox->ymax, nullptr};
    QBoxes nb{rt->xmin, rt->ymin, rt->xmax, rt->ymax, nullptr};
    int64_t cb = 0, cn = 0;
    int rc = compute_safe(qe, pf, rb, rt->fanout, rt->sz, rbox->sz, nullptr);
    if (!rc) rc = stage_boxes(&t.box4, &cb, rb, rbox->sz, nullptr);
    if (!rc) rc = stage_boxes(&t.node4, &cn, nb, rt->sz, nullptr);
    if (!rc) rc = check_layout(&t.canon, rt, rbox->sz, nullptr);
    if (!rc && hipDeviceSynchronize() != hipSuccess) rc = set_error(CURTREE_EHIP, "server_add_tree: staging failed");
    t.safe = qe.safe;
    qe.safe = nullptr;  // ownership moves to the tree
    if (rc) {
      void* ps[] = {t.safe, t.box4, t.node4, t.canon};
      for (void* p : ps)
        if (p) (void)hipFree(p);
      return rc;
    }
  }
  std::lock_guard<std::mutex> g(sv->mu);
  sv->trees.push_back(t);
  *tree_id = (int)sv->trees.size() - 1;
  return CURTREE_OK;
}

CURTREE_API int curtree_server_submit(curtree_server* s, int tree_id,
                                      const curtree_bbox_f32* windows, int64_t* request_id) {
  Server* sv = (Server*)s;
  if (!sv || !windows || !request_id || windows->sz < 0)
    return set_error(CURTREE_EINVAL, "server_submit: bad argument");
  std::lock_guard<std::mutex> g(sv->mu);
  if (tree_id < 0 || tree_id >= (int)sv->trees.size())
    return set_error(CURTREE_EINVAL, "server_submit: unknown tree %d", tree_id);
  if (windows->sz > sv->opt.max_batch_windows)
    return set_error(CURTREE_EINVAL, "server_submit: %d windows exceed the batch cap %lld",
                     windows->sz, (long long)sv->opt.max_batch_windows);
  Request r;
  r.id = sv->next_id++;
  r.tree = tree_id;
  r.win = *windows;
  r.t_submit = now_us();
  if (sv->t_first == 0) sv->t_first = r.t_submit;
  sv->reqs[r.id] = r;
  sv->queue.push_back(r.id);
  *request_id = r.id;
  sv->cv.notify_one();
  return CURTREE_OK;
}

CURTREE_API int curtree_server_poll(curtree_server* s, int64_t request_id, int* state,
                                    curtree_idpair64* result, double* latency_us) {
  Server* sv = (Server*)s;
  if (!sv || !state) return set_error(CURTREE_EINVAL, "server_poll: bad argument");
  std::lock_guard<std::mutex> g(sv->mu);
  auto it = sv->reqs.find(request_id);
  if (it == sv->reqs.end()) return set_error(CURTREE_EINVAL, "server_poll: unknown request");
  const Request& r = it->second;
  *state = r.state;
  if (r.state == CURTREE_REQ_DONE && result) {
    result->sz = r.npairs;
    result->fid = sv->pool_fid + r.pool_off;
    result->tid = sv->pool_tid + r.pool_off;
  }
  if (latency_us) *latency_us = r.state >= CURTREE_REQ_DONE ? r.t_done - r.t_submit : 0.0;
  if (r.state == CURTREE_REQ_FAILED) return set_error(r.err, "request failed");
  return CURTREE_OK;
}

CURTREE_API int curtree_server_release(curtree_server* s, int64_t request_id) {
  Server* sv = (Server*)s;
  if (!sv) return set_error(CURTREE_EINVAL, "server_release: null");
  std::lock_guard<std::mutex> g(sv->mu);
  auto it = sv->reqs.find(request_id);
  if (it == sv->reqs.end()) return set_error(CURTREE_EINVAL, "server_release: unknown request");
  if (it->second.state < CURTREE_REQ_DONE)
    return set_error(CURTREE_EINVAL, "server_release: request still in flight");
  if (it->second.state == CURTREE_REQ_DONE) {
    sv->pool.release(it->second.pool_off, it->second.npairs);
    sv->pool_gen++;
  }
  sv->reqs.erase(it);
  sv->cv.notify_one();
  return CURTREE_OK;
}

CURTREE_API int curtree_server_get_stats(curtree_server* s, curtree_server_stats* st) {
  Server* sv = (Server*)s;
  if (!sv || !st) return set_error(CURTREE_EINVAL, "server_stats: null");
  std::lock_guard<std::mutex> g(sv->mu);
  st->requests_done = sv->requests_done;
  st->pairs_done = sv->pairs_done;
  st->batches = sv->batches;
  st->queued = (int64_t)sv->queue.size();
  std::vector<double> l = sv->lat_us;
  std::sort(l.begin(), l.end());
  auto pct = [&](double p) {
    if (l.empty()) return 0.0;
    size_t i = (size_t)std::min<double>(l.size() - 1, std::ceil(p * l.size()) - 1);
    return l[i];
  };
  st->p50_us = pct(0.50);
  st->p99_us = pct(0.99);
  st->span_us = sv->t_last > sv->t_first ? sv->t_last - sv->t_first : 0.0;
  return CURTREE_OK;
}

CURTREE_API int curtree_server_reset_stats(curtree_server* s) {
  Server* sv = (Server*)s;
  if (!sv) return set_error(CURTREE_EINVAL, "server_reset_stats: null");
  std::lock_guard<std::mutex> g(sv->mu);
  sv->lat_us.clear();
  sv->pairs_done = sv->batches = sv->requests_done = 0;
  sv->t_first = sv->t_last = 0;
  return CURTREE_OK;
}

CURTREE_API int curtree_server_destroy(curtree_server* s) {
  Server* sv = (Server*)s;
  if (!sv) return CURTREE_OK;
  {
    std::lock_guard<std::mutex> g(sv->mu);
    sv->stop = true;
  }
  sv->cv.notify_one();
  if (sv->th.joinable()) sv->th.join();
  (void)hipDeviceSynchronize();
  for (auto& sl : sv->slots) {
    if (sl.s) (void)hipStreamDestroy(sl.s);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.win) (void)hipFree(sl.win);
    if (sl.counts) (void)hipFree(sl.counts);
    if (sl.jobs) (void)hipFree(sl.jobs);
    if (sl.hjobs) (void)hipHostFree(sl.hjobs);
    sl.qe.release();
  }
  for (auto& t : sv->trees) {
    if (t.safe) (void)hipFree(t.safe);
    if (t.box4) (void)hipFree(t.box4);
    if (t.node4) (void)hipFree(t.node4);
    if (t.canon) (void)hipFree(t.canon);
  }
  if (sv->pool_fid) (void)hipFree(sv->pool_fid);
  if (sv->pool_tid) (void)hipFree(sv->pool_tid);
  delete sv;
  return CURTREE_OK;
}

}  // extern "C"
