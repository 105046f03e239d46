// ct_serve.cpp -- serving: several bulk-loaded trees resident in HBM, window
// requests of any size submitted concurrently, continuous batching.
//
// The server owns S slots (stream + query engine + window staging), each run
// by a worker thread.  Whenever its slot is free, a worker takes the oldest
// queued request, adds every other queued request on the same tree (up to a
// window cap), reserves result space in the pool (the memory budget; a batch
// waits while the pool is full) and runs the batch's query writing straight
// into it, with per-window counts.  When the pairs fit the reservation shrinks
// to them: one pass, one host
// round trip per batch; otherwise a second pass writes into an exact
// reservation.  Slots run concurrently on their own streams, and a worker
// enqueues and waits without the server lock, so requests that arrive while
// others are on the device are admitted as slots and memory free up, and
// their work overlaps.  A request's pairs are a contiguous slice of its batch's output;
// their query indices are rebased to the request on the device, so the result
// is the same alone and in any mix.  Results stay in the pool until the
// client releases them.  With batching off each request runs alone: the plain
// per-request launches the serving mode is compared with.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <unordered_map>

#include "ct_pool.hpp"
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

// Query indices of a batch's pairs relative to their request: the request of
// a pair is the last one whose first batch window is <= the pair's window
// (binary search over the requests' first windows).  The pair count is read
// from the run's control block, so no host round trip is needed.
__global__ void k_rebase_search(const QCtrl* __restrict__ c, int64_t cap,
                                const int64_t* __restrict__ first_win, int nreq,
                                int* __restrict__ tid) {
  const int64_t n = (int64_t)c->total < cap ? (int64_t)c->total : cap;
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const int64_t t = tid[i];
    int lo = 0, hi = nreq - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (first_win[mid] <= t) lo = mid;
      else hi = mid - 1;
    }
    tid[i] = (int)(t - first_win[lo]);
  }
}

namespace {

double now_us() {
  using namespace std::chrono;
  return duration_cast<duration<double, std::micro>>(steady_clock::now().time_since_epoch())
      .count();
}

// Slot streams with a hardware queue of their own.  The runtime shares a
// small pool of hardware queues among ordinary streams (4 per process by
// default, one of them the null stream's), so two slots on one queue would
// run their kernels one after the other.  A stream created with a CU mask (the
// mask enables every CU) is given a queue outside that pool.  The library
// creates at most kDedicatedQueues of them per device, once, and lends them
// to servers (a slot returns its stream when the server is destroyed); slots
// beyond them take ordinary streams.
constexpr int kDedicatedQueues = 4;
std::mutex g_qmu;
std::unordered_map<int, std::vector<hipStream_t>> g_qfree;  // per device: idle streams
std::unordered_map<int, int> g_qmade;                       // per device: streams created

hipStream_t acquire_queue(int dev) {
  std::lock_guard<std::mutex> g(g_qmu);
  auto& fr = g_qfree[dev];
  if (!fr.empty()) {
    hipStream_t s = fr.back();
    fr.pop_back();
    return s;
  }
  if (g_qmade[dev] >= kDedicatedQueues) return nullptr;
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  std::vector<uint32_t> mask((size_t)std::max(1, (ncu + 31) / 32), 0xFFFFFFFFu);
  hipStream_t s = nullptr;
  if (hipExtStreamCreateWithCUMask(&s, (uint32_t)mask.size(), mask.data()) != hipSuccess) {
    (void)hipGetLastError();
    g_qmade[dev] = kDedicatedQueues;  // not available: ordinary streams from now on
    return nullptr;
  }
  g_qmade[dev]++;
  return s;
}

void release_queue(int dev, hipStream_t s) {
  std::lock_guard<std::mutex> g(g_qmu);
  g_qfree[dev].push_back(s);
}

// A batch's completion: polled for up to kSpinUs (a batch usually ends within
// a few hundred microseconds, and a poll sees it at once), then a blocking
// wait.
constexpr double kSpinUs = 2000.0;
hipError_t wait_event(hipEvent_t ev) {
  const double t0 = now_us();
  for (;;) {
    const hipError_t e = hipEventQuery(ev);
    if (e != hipErrorNotReady) return e;
    if (now_us() - t0 > kSpinUs) break;
    std::this_thread::yield();
  }
  return hipEventSynchronize(ev);
}

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

// A slot runs one batch at a time.  Pass 1 writes the batch's pairs straight
// into a pool reservation of `reserve` pairs (with per-window counts), rebases
// query indices on the device and copies the counts back; if the pairs fit,
// the reservation shrinks to them and the batch is done after ONE pass and
// one host round trip.  Otherwise the counts give the exact size: a pass 2
// writes into an exact reservation.  Every step is enqueued on the slot's
// stream and finishes on an event, so the dispatcher never waits on the
// device and all slots' work overlaps.
enum SlotState { kIdle, kWaitPool1, kPass1, kWaitPool2, kPass2 };

struct Slot {
  hipStream_t s = nullptr;
  bool dedicated = false;  // s is a stream with a hardware queue of its own (acquire_queue)
  int dev = 0;
  hipEvent_t ev = nullptr;
  QEngine qe;
  float* win = nullptr;  // staging: 4 arrays of win_cap windows
  int64_t win_cap = 0;
  int64_t* counts = nullptr;   // per-window pair counts (device)
  int64_t* hcounts = nullptr;  // pinned copy
  ServeJob* jobs = nullptr;    // device descriptors (window staging)
  ServeJob* hjobs = nullptr;   // pinned host copy
  int64_t* first = nullptr;    // device: first batch window of each request
  int64_t* hfirst = nullptr;   // pinned host copy
  int jobs_cap = 0;
  SlotState state = kIdle;
  // device work decided under the server lock and enqueued outside it
  enum Todo { kNone, kBegin, kPass, kGrowPass } todo = kNone;
  int64_t grow_to = 0;         // kGrowPass: frontier entries
  Tree tree;                   // the batch's tree (a copy: trees may grow)
  int attempts = 0;
  QRunArgs a;                  // the batch's run
  int64_t nwin = 0;
  std::vector<int64_t> reqs;
  int64_t pool_off = 0, pool_len = 0;  // current reservation
};

}  // namespace

struct Server {
  curtree_server_opts opt{};
  std::vector<Tree> trees;
  std::vector<Slot> slots;
  std::unordered_map<int64_t, Request> reqs;
  std::deque<int64_t> queue;
  std::deque<int64_t> completed;  // finished request ids not yet reported
  std::atomic<int64_t> ncompleted{0};  // completed.size(), readable without the lock
  Pool pool;
  int* pool_fid = nullptr;
  int* pool_tid = nullptr;
  int64_t reserve = 0;  // pass-1 reservation (pairs)
  Prof prof;  // never enabled
  std::mutex mu;
  std::condition_variable cv;
  std::vector<std::thread> workers;  // one per slot
  bool stop = false;
  int64_t next_id = 1;
  std::vector<double> lat_us;
  int64_t pairs_done = 0, batches = 0, requests_done = 0;
  double launch_us = 0;  // worker time spent enqueuing device work (summed over slots)
  uint64_t pool_gen = 0;
  double t_first = 0, t_last = 0;

  int start();
  void worker(Slot& sl);
  bool take_batch(Slot& sl);
  int begin_batch(Slot& sl);
  int try_pass(Slot& sl, bool exact);
  int enqueue_pass(Slot& sl);
  int after_pass(Slot& sl);
  int run_todo(Slot& sl);
  void finish(Slot& sl, int err);
};

int Server::start() {
  int rc;
  slots.resize((size_t)opt.slots);
  const int64_t cap = std::max<int64_t>(opt.max_batch_windows, 1);
  int dev = 0;
  CT_HIP(hipGetDevice(&dev));
  for (auto& sl : slots) {
    sl.dev = dev;
    sl.s = acquire_queue(dev);
    if (sl.s) sl.dedicated = true;
    else CT_HIP(hipStreamCreateWithFlags(&sl.s, hipStreamNonBlocking));
    CT_HIP(hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming));
    CT_HIP(hipMalloc((void**)&sl.win, (size_t)cap * 16));
    CT_HIP(hipMalloc((void**)&sl.counts, (size_t)cap * 8));
    CT_HIP(hipHostMalloc((void**)&sl.hcounts, (size_t)cap * 8, hipHostMallocDefault));
    sl.win_cap = cap;
    sl.jobs_cap = 4096;
    CT_HIP(hipMalloc((void**)&sl.jobs, (size_t)sl.jobs_cap * sizeof(ServeJob)));
    CT_HIP(hipHostMalloc((void**)&sl.hjobs, (size_t)sl.jobs_cap * sizeof(ServeJob),
                         hipHostMallocDefault));
    CT_HIP(hipMalloc((void**)&sl.first, (size_t)sl.jobs_cap * 8));
    CT_HIP(hipHostMalloc((void**)&sl.hfirst, (size_t)sl.jobs_cap * 8, hipHostMallocDefault));
    if ((rc = sl.qe.ensure_ctrl())) return rc;
  }
  const int64_t pairs = std::max<int64_t>(opt.budget_bytes / 8, 1);
  CT_HIP(hipMalloc((void**)&pool_fid, (size_t)pairs * 4));
  CT_HIP(hipMalloc((void**)&pool_tid, (size_t)pairs * 4));
  pool.init(pairs);
  reserve = std::max<int64_t>(pairs / (2 * (int64_t)opt.slots), 1);
  for (auto& sl : slots) workers.emplace_back([this, &sl] { worker(sl); });
  return CURTREE_OK;
}

// The batch's requests are finished (err == 0: done, their pairs stay in the
// pool until released; else failed, the reservation is returned).
void Server::finish(Slot& sl, int err) {
  const double t = now_us();
  if (err && sl.pool_len > 0) {
    pool.release(sl.pool_off, sl.pool_len);
    pool_gen++;
  }
  for (int64_t id : sl.reqs) {
    auto it = reqs.find(id);
    if (it == reqs.end()) continue;
    Request& r = it->second;
    r.state = err ? CURTREE_REQ_FAILED : CURTREE_REQ_DONE;
    r.err = err;
    r.t_done = t;
    if (!err) {
      lat_us.push_back(t - r.t_submit);
      pairs_done += r.npairs;
      requests_done++;
    }
    completed.push_back(id);
    t_last = t;
  }
  ncompleted.store((int64_t)completed.size(), std::memory_order_release);
  sl.reqs.clear();
  sl.pool_off = sl.pool_len = 0;
  sl.state = kIdle;
}

// Under the lock: the batch's window descriptors and request offsets.
static int64_t fill_jobs(Slot& sl, std::unordered_map<int64_t, Request>& reqs) {
  int64_t nwin = 0;
  for (size_t j = 0; j < sl.reqs.size(); j++) {
    const Request& r = reqs[sl.reqs[j]];
    ServeJob& jb = sl.hjobs[j];
    jb.src[0] = r.win.xmin;
    jb.src[1] = r.win.ymin;
    jb.src[2] = r.win.xmax;
    jb.src[3] = r.win.ymax;
    jb.n = r.win.sz;
    jb.dst = nwin;
    jb.base = nwin;
    sl.hfirst[j] = nwin;
    nwin += r.win.sz;
  }
  return nwin;
}

// Outside the lock: stage the batch's windows and enqueue pass 1 (the pool
// reservation is already held).
int Server::begin_batch(Slot& sl) {
  const Tree& tr = sl.tree;
  const int nj = (int)sl.reqs.size();
  const int64_t nwin = sl.nwin;
  curtree_bbox_f32 q{(int)nwin, sl.win, sl.win + sl.win_cap, sl.win + 2 * sl.win_cap,
                     sl.win + 3 * sl.win_cap, nullptr};
  if (nj == 1) {
    // one request: the run reads its windows where the client put them
    const ServeJob& jb = sl.hjobs[0];
    q = curtree_bbox_f32{(int)nwin, (float*)jb.src[0], (float*)jb.src[1], (float*)jb.src[2],
                         (float*)jb.src[3], nullptr};
  } else {
    CT_HIP(hipMemcpyAsync(sl.jobs, sl.hjobs, (size_t)nj * sizeof(ServeJob),
                          hipMemcpyHostToDevice, sl.s));
    CT_HIP(hipMemcpyAsync(sl.first, sl.hfirst, (size_t)nj * 8, hipMemcpyHostToDevice, sl.s));
    hipLaunchKernelGGL(k_stage, dim3((unsigned)std::min(nj, 1024)), dim3(256), 0, sl.s,
                       (const ServeJob*)sl.jobs, nj, sl.win, sl.win_cap);
  }
  // the tree's cached containment flags, AoS copies (if any) and layout check
  int rc;
  if ((rc = prepare_query(sl.qe, prof, &tr.rt, &tr.rbox, &q, 0, nwin, &sl.a, sl.s, tr.canon)))
    return rc;
  sl.a.tr.safe = tr.safe;
  sl.a.rb.b4 = tr.box4;
  sl.a.tr.mbr.b4 = tr.node4;
  sl.a.q0 = 0;
  sl.a.q1 = nwin;
  if ((rc = sl.qe.reserve_front(std::max<int64_t>(2 * nwin, 1 << 16)))) return rc;
  return enqueue_pass(sl);
}

// Outside the lock: the slot's pending device work.
int Server::run_todo(Slot& sl) {
  const Slot::Todo t = sl.todo;
  sl.todo = Slot::kNone;
  int rc;
  switch (t) {
    case Slot::kBegin:
      return begin_batch(sl);
    case Slot::kGrowPass:
      if ((rc = sl.qe.reserve_front(sl.grow_to))) return rc;
      return enqueue_pass(sl);
    case Slot::kPass:
      return enqueue_pass(sl);
    default:
      return CURTREE_OK;
  }
}

// Reserve pool space (pass 1: the standing reservation or the largest free
// block; pass 2: exactly the counted pairs) and enqueue the pass, or wait.
int Server::try_pass(Slot& sl, bool exact) {
  int64_t want = 0;
  if (exact) {
    want = sl.qe.host_ctrl->total;
    if (want > pool.cap)
      return set_error(CURTREE_ERANGE, "serve: %lld pairs exceed the result budget",
                       (long long)want);
  } else {
    want = std::min<int64_t>(reserve, std::max<int64_t>(pool.largest(), 0));
    if (want < std::min<int64_t>(reserve, 1 << 16)) want = -1;  // too little free: wait
  }
  const int64_t off = want > 0 ? pool.alloc(want) : (want == 0 ? 0 : -1);
  if (off < 0) {
    sl.state = exact ? kWaitPool2 : kWaitPool1;
    return CURTREE_OK;
  }
  sl.pool_off = off;
  sl.pool_len = want;
  const bool first = sl.state == kIdle || sl.state == kWaitPool1;
  sl.state = exact ? kPass2 : kPass1;
  sl.todo = first && !exact ? Slot::kBegin : Slot::kPass;
  return CURTREE_OK;
}

int Server::enqueue_pass(Slot& sl) {
  int rc;
  QRunArgs& a = sl.a;
  const bool pass1 = sl.state == kPass1;
  // per-window counts split a batch's pairs among its requests; a batch of
  // one request takes them all (the run's total)
  const bool split = pass1 && sl.reqs.size() > 1;
  if (a.nbox == 0 || sl.nwin == 0) {
    if (split) enqueue_fill_counts(sl.counts, sl.nwin, sl.s);
    sl.qe.host_ctrl->total = 0;
    sl.qe.host_ctrl->overflow = 0;
  } else {
    a.write = true;
    a.counts = split ? (unsigned long long*)sl.counts : nullptr;
    a.out_fid = pool_fid + sl.pool_off;
    a.out_tid = pool_tid + sl.pool_off;
    a.out_cap = sl.pool_len;
    if ((rc = sl.qe.reserve_status(std::max<int64_t>(sl.qe.cap_front, sl.nwin) /
                                       per_tile(a.G, a.steps) + 2)) ||
        (rc = sl.qe.reserve_pieces(pieces_needed(sl.qe, a))))
      return rc;
    if (split) enqueue_fill_counts(sl.counts, sl.nwin, sl.s);
    enqueue_run(sl.qe, prof, a, sl.s);
    if (sl.reqs.size() > 1 && sl.pool_len > 0)
      hipLaunchKernelGGL(k_rebase_search, dim3(1024), dim3(256), 0, sl.s,
                         (const QCtrl*)sl.qe.ctrl, sl.pool_len, (const int64_t*)sl.first,
                         (int)sl.reqs.size(), pool_tid + sl.pool_off);
    CT_HIP(hipMemcpyAsync(sl.qe.host_ctrl, sl.qe.ctrl, sizeof(QCtrl), hipMemcpyDeviceToHost,
                          sl.s));
  }
  if (split)
    CT_HIP(hipMemcpyAsync(sl.hcounts, sl.counts, (size_t)sl.nwin * 8, hipMemcpyDeviceToHost,
                          sl.s));
  CT_HIP(hipEventRecord(sl.ev, sl.s));
  return CURTREE_OK;
}

// A pass finished: rerun after a frontier overflow; after pass 1 shrink the
// reservation to the pairs (done) or go for an exact pass 2.
int Server::after_pass(Slot& sl) {
  QRunArgs& a = sl.a;
  const QCtrl& hc = *sl.qe.host_ctrl;
  const bool ran = a.nbox > 0 && sl.nwin > 0;
  if (ran && hc.overflow) {
    if (++sl.attempts > a.nl + 1)
      return set_error(CURTREE_ENOMEM, "serve: frontier kept overflowing");
    uint64_t need = 0;
    for (int l = 1; l < a.nl; l++) need = std::max<uint64_t>(need, hc.count[l]);
    sl.grow_to = (int64_t)(need + need / 4 + 1024);
    sl.todo = Slot::kGrowPass;
    return CURTREE_OK;
  }
  const int64_t H = ran ? (int64_t)hc.total : 0;
  if (sl.state == kPass1) {
    // per-request slices from the per-window counts (one request: all pairs)
    int64_t p0 = 0;
    for (size_t j = 0; j < sl.reqs.size(); j++) {
      Request& r = reqs[sl.reqs[j]];
      const int64_t w0 = sl.hfirst[j];
      int64_t np = 0;
      if (sl.reqs.size() == 1) np = H;
      else
        for (int64_t i = 0; i < r.win.sz; i++) np += sl.hcounts[w0 + i];
      r.pool_off = p0;  // relative until the final reservation is known
      r.npairs = np;
      r.state = CURTREE_REQ_RUNNING;
      p0 += np;
    }
    if (H > sl.pool_len) {
      // did not fit: return the reservation, write again into an exact one
      if (sl.pool_len > 0) {
        pool.release(sl.pool_off, sl.pool_len);
        pool_gen++;
      }
      sl.pool_off = sl.pool_len = 0;
      return try_pass(sl, true);
    }
    if (H < sl.pool_len) {
      pool.release(sl.pool_off + H, sl.pool_len - H);
      pool_gen++;
      sl.pool_len = H;
    }
  }
  for (int64_t id : sl.reqs) reqs[id].pool_off += sl.pool_off;
  finish(sl, CURTREE_OK);
  return CURTREE_OK;
}

// Admit the oldest request (and, with batching, every queued request on the
// same tree that fits) onto free slot `sl`.  Returns true if a batch started.
bool Server::take_batch(Slot& sl) {
  if (queue.empty()) return false;
  // a batch waiting for result memory goes first: no new batch may take the
  // space that releases free up
  for (auto& o : slots)
    if (o.state == kWaitPool1 || o.state == kWaitPool2) return false;
  const int tree = reqs[queue.front()].tree;
  int64_t nwin = 0;
  sl.reqs.clear();
  for (auto it = queue.begin(); it != queue.end();) {
    const Request& r = reqs[*it];
    const bool first = sl.reqs.empty();
    if (r.tree == tree && (first || (opt.batching && nwin + r.win.sz <= sl.win_cap &&
                                     (int)sl.reqs.size() < sl.jobs_cap))) {
      sl.reqs.push_back(*it);
      nwin += r.win.sz;
      it = queue.erase(it);
      if (!opt.batching) break;
    } else {
      ++it;
    }
  }
  batches++;
  sl.tree = trees[(size_t)tree];
  sl.nwin = fill_jobs(sl, reqs);
  sl.attempts = 0;
  const int rc = try_pass(sl, false);  // reserve now, stage outside the lock
  if (rc) finish(sl, rc);
  return true;
}

// One worker thread per slot: it takes a batch when its slot is free, does
// the slot's device work outside the server lock and waits on the slot's own
// event, so the slots' launch work and waits proceed in parallel.
void Server::worker(Slot& sl) {
  std::unique_lock<std::mutex> lk(mu);
  while (!stop) {
    if (sl.state == kIdle) {
      if (!take_batch(sl)) cv.wait_for(lk, std::chrono::milliseconds(5));
      continue;
    }
    int rc = CURTREE_OK;
    if (sl.todo != Slot::kNone) {
      lk.unlock();
      const double t0 = now_us();
      rc = run_todo(sl);
      const double dt = now_us() - t0;
      lk.lock();
      launch_us += dt;
    } else if (sl.state == kPass1 || sl.state == kPass2) {
      lk.unlock();
      const hipError_t e = wait_event(sl.ev);
      lk.lock();
      rc = e != hipSuccess ? set_error(CURTREE_EHIP, "serve: %s", hipGetErrorString(e))
                           : after_pass(sl);
    } else {  // waiting for result memory
      const SlotState was = sl.state;
      rc = try_pass(sl, was == kWaitPool2);
      if (!rc && sl.state == was) cv.wait_for(lk, std::chrono::milliseconds(1));
    }
    if (rc) finish(sl, rc);
    if (sl.state == kIdle) cv.notify_all();  // released pool space, a free slot
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
  if (rt->fanout < 2 || lp.r_sz != rt->sz)
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
    // AoS copies of the boxes and node MBRs only when asked for ("box_staging"
    // 1): the query reads the caller's SoA arrays faster
    if (!rc && qtuning().staging == 1) rc = stage_boxes(&t.box4, &cb, rb, rbox->sz, nullptr);
    if (!rc && qtuning().staging == 1) rc = stage_boxes(&t.node4, &cn, nb, rt->sz, nullptr);
    if (!rc) rc = check_layout(&t.canon, rt, rbox->sz, nullptr);
    // the leaf codes of a tree built by this library, now rather than in
    // the first large batch (a no-op for other trees)
    if (!rc) rc = ensure_leaf_codes(rt, rbox, nullptr);
    if (!rc && hipDeviceSynchronize() != hipSuccess)
      rc = set_error(CURTREE_EHIP, "server_add_tree: staging failed");
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
  sv->cv.notify_all();
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

CURTREE_API int curtree_server_completed(curtree_server* s, int64_t* ids, int cap, int* n) {
  Server* sv = (Server*)s;
  if (!sv || !n || cap < 0 || (cap > 0 && !ids))
    return set_error(CURTREE_EINVAL, "server_completed: bad argument");
  // an empty queue is seen without the lock, so a client may poll often
  // without holding up the workers
  if (sv->ncompleted.load(std::memory_order_acquire) == 0) {
    *n = 0;
    return CURTREE_OK;
  }
  std::lock_guard<std::mutex> g(sv->mu);
  int k = 0;
  while (k < cap && !sv->completed.empty()) {
    ids[k++] = sv->completed.front();
    sv->completed.pop_front();
  }
  sv->ncompleted.store((int64_t)sv->completed.size(), std::memory_order_release);
  *n = k;
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
  sv->cv.notify_all();
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
  st->launch_us = sv->launch_us;
  return CURTREE_OK;
}

CURTREE_API int curtree_server_reset_stats(curtree_server* s) {
  Server* sv = (Server*)s;
  if (!sv) return set_error(CURTREE_EINVAL, "server_reset_stats: null");
  std::lock_guard<std::mutex> g(sv->mu);
  sv->lat_us.clear();
  sv->pairs_done = sv->batches = sv->requests_done = 0;
  sv->launch_us = 0;
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
  sv->cv.notify_all();
  for (auto& w : sv->workers)
    if (w.joinable()) w.join();
  (void)hipDeviceSynchronize();
  for (auto& sl : sv->slots) {
    if (sl.dedicated) release_queue(sl.dev, sl.s);  // kept for the next server
    else if (sl.s) (void)hipStreamDestroy(sl.s);
    if (sl.ev) (void)hipEventDestroy(sl.ev);
    if (sl.win) (void)hipFree(sl.win);
    if (sl.counts) (void)hipFree(sl.counts);
    if (sl.hcounts) (void)hipHostFree(sl.hcounts);
    if (sl.jobs) (void)hipFree(sl.jobs);
    if (sl.hjobs) (void)hipHostFree(sl.hjobs);
    if (sl.first) (void)hipFree(sl.first);
    if (sl.hfirst) (void)hipHostFree(sl.hfirst);
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
