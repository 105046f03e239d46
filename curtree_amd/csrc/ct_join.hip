// ct_join.hip -- distance self-join ("neighbor edges"): every pair of boxes
// whose distance along both axes is at most r, found by querying the tree
// with each data box expanded by r.
//
// This is the workload the reference's drivers run: the query file is the
// data file read again with the boxes expanded by r = 10
// (src/rtree_test_frontier.cu:84-85, readMBR's +/- r at src/utility.cpp:34-37),
// and the Boost.Geometry check does the same self-join (src/verify.hpp:75-104).
// Query ids are the data boxes' ids (the reference's query index equals the
// file position, which is the id when ids are 0..N-1).
#include "ct_internal.hpp"

namespace ct {

__global__ __launch_bounds__(256) void k_expand_windows(int64_t n, const float* __restrict__ x0,
                                                        const float* __restrict__ y0,
                                                        const float* __restrict__ x1,
                                                        const float* __restrict__ y1, float r,
                                                        float* __restrict__ qx0,
                                                        float* __restrict__ qy0,
                                                        float* __restrict__ qx1,
                                                        float* __restrict__ qy1) {
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    qx0[i] = x0[i] - r;
    qy0[i] = y0[i] - r;
    qx1[i] = x1[i] + r;
    qy1[i] = y1[i] + r;
  }
}

// tid (a sorted position of the querying box) -> that box's id
__global__ __launch_bounds__(256) void k_map_query_ids(int64_t h, int* __restrict__ tid,
                                                       const int* __restrict__ ids) {
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < h; i += (int64_t)gridDim.x * 256)
    tid[i] = ids[tid[i]];
}

int self_join_f32(Context& ctx, const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox, float r,
                  curtree_idpair64* out, hipStream_t s) {
  if (!rt || !rbox || !out) return set_error(CURTREE_EINVAL, "self_join: null argument");
  out->sz = 0;
  out->fid = out->tid = nullptr;
  const int64_t n = rbox->sz;
  if (n <= 0) return CURTREE_OK;
  if (!(r >= 0.0f)) return set_error(CURTREE_EINVAL, "self_join: r must be >= 0");
  int rc;
  if ((rc = ctx.ws.reserve(WsJoinWin, (size_t)n * 16))) return rc;
  float* w = (float*)ctx.ws.ptr(WsJoinWin);
  const unsigned grid = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_expand_windows, dim3(grid), dim3(256), 0, s, n, rbox->xmin, rbox->ymin,
                     rbox->xmax, rbox->ymax, r, w, w + n, w + 2 * n, w + 3 * n);
  curtree_bbox_f32 q{(int)n, w, w + n, w + 2 * n, w + 3 * n, nullptr};
  QueryOut qo;
  qo.alloc = true;
  if ((rc = query_frontier_f32(ctx, rt, rbox, &q, &qo, s))) return rc;
  if (qo.n > 0) {
    const unsigned g2 = (unsigned)std::min<int64_t>((qo.n + 255) / 256, 1 << 16);
    hipLaunchKernelGGL(k_map_query_ids, dim3(g2), dim3(256), 0, s, qo.n, qo.out_tid, rbox->id);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    if (qo.out_fid) (void)hipFree(qo.out_fid);
    if (qo.out_tid) (void)hipFree(qo.out_tid);
    return set_error(CURTREE_EHIP, "self_join: launch failed: %s", hipGetErrorString(e));
  }
  out->sz = qo.n;
  out->fid = qo.out_fid;
  out->tid = qo.out_tid;
  return CURTREE_OK;
}

}  // namespace ct
