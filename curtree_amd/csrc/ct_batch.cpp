// ct_batch.cpp -- batched queries whose results exceed device memory: a count
// pass gives every window's pair count; windows are cut into chunks whose
// pairs fit a result budget; each chunk is one replay of a hipGraph holding a
// whole query run (begin, one expansion per level, leaf stage, piece writer),
// so chunks need no kernel-launch or host round trips beyond the replay.
//
// The reference runs one query batch per call and aborts when a block's
// frontier exceeds its LDS queue (src/rtree_bfs_frontier.cuh:385-403); here a
// window set of any size and any result volume streams through a fixed budget.
#include <stdio.h>

#include "ct_query.hpp"

namespace ct {

// Chunks of consecutive windows with at most `budget` pairs each (a window
// with more pairs than the budget gets a chunk of its own).
static void plan_chunks(const int64_t* counts, int64_t q, int64_t budget,
                        std::vector<int64_t>* plan, std::vector<int64_t>* sums,
                        int64_t* max_chunk) {
  plan->clear();
  sums->clear();
  *max_chunk = 0;
  int64_t start = 0, acc = 0;
  for (int64_t i = 0; i < q; i++) {
    if (acc > 0 && acc + counts[i] > budget) {
      plan->push_back(start);
      plan->push_back(i);
      sums->push_back(acc);
      *max_chunk = std::max(*max_chunk, acc);
      start = i;
      acc = 0;
    }
    acc += counts[i];
  }
  if (q > start) {
    plan->push_back(start);
    plan->push_back(q);
    sums->push_back(acc);
    *max_chunk = std::max(*max_chunk, acc);
  }
}

int query_batched_f32(Context& ctx, const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox,
                      const curtree_bbox_f32* qbox, int64_t budget, curtree_chunk_fn fn,
                      void* user, curtree_batch_stats* stats) {
  if (!rt || !rbox || !qbox || budget <= 0)
    return set_error(CURTREE_EINVAL, "batched: bad argument");
  BatchState& bs = ctx.batch;  // caller holds bs.mu
  // the context lock covers the setup (profiler, options); the chunk loop
  // below touches batch-owned resources only and runs without it, so chunk
  // callbacks may call other entry points
  std::unique_lock<std::mutex> lk(ctx.mu);
  int rc;
  if (!bs.stream) CT_HIP(hipStreamCreateWithFlags(&bs.stream, hipStreamNonBlocking));
  if (!bs.ev0) {
    CT_HIP(hipEventCreate(&bs.ev0));
    CT_HIP(hipEventCreate(&bs.ev1));
  }
  hipStream_t s = bs.stream;
  QEngine& qe = bs.qe;
  curtree_batch_stats st{};
  const int64_t Q = qbox->sz;
  if (Q == 0 || rbox->sz == 0) {
    if (stats) *stats = st;
    return CURTREE_OK;
  }
  // ---- pass 1: per-window pair counts
  if ((rc = bs.ws.reserve(WsBatchCounts, (size_t)Q * 8))) return rc;
  int64_t* counts = (int64_t*)bs.ws.ptr(WsBatchCounts);
  CT_HIP(hipEventRecord(bs.ev0, s));
  QueryOut qo;
  qo.count_only = true;
  qo.counts = counts;
  if ((rc = query_run_f32(qe, ctx.prof, rt, rbox, qbox, 0, Q, ctx.opt_contain, &qo, s))) return rc;
  CT_HIP(hipEventRecord(bs.ev1, s));
  CT_HIP(hipEventSynchronize(bs.ev1));
  float ms = 0.f;
  CT_HIP(hipEventElapsedTime(&ms, bs.ev0, bs.ev1));
  st.count_ms = ms;
  uint64_t front_need = 0;
  for (int l = 1; l < kMaxRunSlots; l++) front_need = std::max(front_need, qe.host_ctrl->count[l]);
  std::vector<int64_t> hc((size_t)Q);
  CT_HIP(hipMemcpy(hc.data(), counts, (size_t)Q * 8, hipMemcpyDeviceToHost));
  std::vector<int64_t> plan, sums;
  int64_t max_chunk = 0;
  plan_chunks(hc.data(), Q, budget, &plan, &sums, &max_chunk);
  const int64_t nchunks = (int64_t)sums.size();
  const int64_t out_cap = std::max<int64_t>(max_chunk, 1);

  // ---- buffers sized once for every chunk (the graph holds their addresses)
  QRunArgs a;
  if ((rc = prepare_query(qe, ctx.prof, rt, rbox, qbox, ctx.opt_contain, Q, &a, s))) return rc;
  int64_t max_q = 0;
  for (int64_t k = 0; k < nchunks; k++) max_q = std::max(max_q, plan[2 * k + 1] - plan[2 * k]);
  // frontiers and leaf entries (a one-level tree has one leaf entry per window)
  if ((rc = qe.reserve_front(
           std::max<int64_t>(std::max<int64_t>((int64_t)front_need + 1024, max_q), 1 << 16))))
    return rc;
  if ((rc = qe.reserve_status(std::max<int64_t>(qe.cap_front, max_q) / per_tile(a.G, a.steps) + 2)))
    return rc;
  if ((rc = bs.ws.reserve(WsBatchOut, (size_t)out_cap * 8))) return rc;
  int* fid = (int*)bs.ws.ptr(WsBatchOut);
  int* tid = fid + out_cap;
  a.write = true;
  a.out_fid = fid;
  a.out_tid = tid;
  a.out_cap = out_cap;
  a.plan = nullptr;
  if ((rc = qe.reserve_pieces(pieces_needed(qe, a)))) return rc;
  if ((rc = qe.reserve_plan(nchunks))) return rc;
  a.plan = qe.plan;
  CT_HIP(hipMemcpyAsync(qe.plan, plan.data(), (size_t)nchunks * 16, hipMemcpyHostToDevice, s));
  // fresh control block: chunk 0 next, status words cleared by the first run
  qe.needs_clear = true;
  // ---- capture one run, replay it per chunk
  const bool prof_on = ctx.prof.on;
  ctx.prof.on = false;  // no event nodes inside the graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // k_ctrl_init runs eagerly, before the capture
  enqueue_ctrl_init(qe, s);
  CT_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  enqueue_run(qe, ctx.prof, a, s);
  hipError_t ce = hipStreamEndCapture(s, &graph);
  ctx.prof.on = prof_on;
  if (ce != hipSuccess)
    return set_error(CURTREE_EHIP, "batched: graph capture failed: %s", hipGetErrorString(ce));
  ce = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (ce != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return set_error(CURTREE_EHIP, "batched: graph instantiation failed: %s",
                     hipGetErrorString(ce));
  }
  lk.unlock();
  rc = CURTREE_OK;
  double write_ms = 0.0;
  for (int64_t k = 0; k < nchunks && !rc; k++) {
    hipError_t e = hipEventRecord(bs.ev0, s);
    if (e == hipSuccess) e = hipGraphLaunch(exec, s);
    if (e == hipSuccess) e = hipEventRecord(bs.ev1, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(qe.host_ctrl, qe.ctrl, sizeof(QCtrl), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      rc = set_error(CURTREE_EHIP, "batched: chunk %lld: %s", (long long)k, hipGetErrorString(e));
      break;
    }
    float cm = 0.f;
    (void)hipEventElapsedTime(&cm, bs.ev0, bs.ev1);
    write_ms += cm;
    const QCtrl& hc2 = *qe.host_ctrl;
    if (hc2.overflow || (int64_t)hc2.total != sums[(size_t)k] || hc2.q0 != plan[2 * k] ||
        hc2.q1 != plan[2 * k + 1]) {
      rc = set_error(CURTREE_EHIP,
                     "batched: chunk %lld wrote %llu pairs for windows [%lld,%lld), planned %lld",
                     (long long)k, (unsigned long long)hc2.total, (long long)hc2.q0,
                     (long long)hc2.q1, (long long)sums[(size_t)k]);
      break;
    }
    st.pairs += sums[(size_t)k];
    st.chunks++;
    if (fn && fn(user, fid, tid, sums[(size_t)k], plan[2 * k], plan[2 * k + 1]) != 0) {
      rc = set_error(CURTREE_EINVAL, "batched: stopped by the chunk callback");
      break;
    }
  }
  (void)hipGraphExecDestroy(exec);
  (void)hipGraphDestroy(graph);
  st.write_ms = write_ms;
  st.max_chunk_pairs = max_chunk;
  if (stats) *stats = st;
  return rc;
}

}  // namespace ct
