// dropin_frontier.cpp -- a caller written against the reference's template API
// (src/rtree_test_frontier.cu:37-137 is the program it is shaped after),
// compiled against include/curtree/curtree.hpp and linked with
// libcurtree_amd.so.  It reads an edge file and a query file in the readMBR
// format, bulk-loads a tree, queries it with bfs_frontier and bfs_thrust,
// sorts the pairs on the host into (query id, box id) order with two stable
// sorts, and compares them with a host nested loop over the input boxes.
//
//   dropin_frontier input_mbr input_num query_mbr query_num fanout [key_mode]
//
// Prints "verified successfully" and exits 0 when every check holds.
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "curtree/curtree.hpp"

// closed intervals, as the reference's intersect_cpu (src/utility.h:16-20)
static bool overlap(float x1, float y1, float x2, float y2, float m1, float n1, float m2,
                    float n2) {
  return !(x1 > m2 || x2 < m1 || y1 > n2 || y2 < n1);
}

// Host order of the reference driver: stable by box id, then stable by query.
static void host_sort(std::vector<int>& fid, std::vector<int>& tid) {
  std::vector<size_t> o(fid.size());
  std::iota(o.begin(), o.end(), size_t(0));
  std::stable_sort(o.begin(), o.end(), [&](size_t a, size_t b) { return fid[a] < fid[b]; });
  std::stable_sort(o.begin(), o.end(), [&](size_t a, size_t b) { return tid[a] < tid[b]; });
  std::vector<int> f(fid.size()), t(tid.size());
  for (size_t i = 0; i < o.size(); i++) {
    f[i] = fid[o[i]];
    t[i] = tid[o[i]];
  }
  fid.swap(f);
  tid.swap(t);
}

static bool run_query(const char* what, bool thrust_form, RTREE<float>& d_rt, BBOX<float>& d_r_box,
                      BBOX<float>& d_q_box, const std::vector<int>& c_did,
                      const std::vector<int>& c_qid) {
  IDPAIR d_dq, h_dq;
  if (thrust_form)
    bfs_thrust(d_rt, d_r_box, d_q_box, d_dq);
  else
    bfs_frontier(d_rt, d_r_box, d_q_box, d_dq);
  printf("%s: num d_dq pairs=%d\n", what, d_dq.sz);
  idpair_d2h(h_dq, d_dq);
  std::vector<int> g_did(h_dq.fid, h_dq.fid + h_dq.sz), g_qid(h_dq.tid, h_dq.tid + h_dq.sz);
  host_sort(g_did, g_qid);
  const bool same = g_did == c_did && g_qid == c_qid;
  if (!same) printf("%s: %zu pairs, expected %zu\n", what, g_did.size(), c_did.size());
  idpair_d_free(d_dq);
  idpair_h_free(h_dq);
  return same;
}

int main(int argc, char* argv[]) {
  if (argc != 6 && argc != 7) {
    printf("EXEC input_mbr input_num query_mbr query_num rtree_fanout [key_mode]\n");
    return 2;
  }
  BBOX<float> h_d_box, d_r_box, h_q_box, d_q_box;
  RTREE<float> d_rt;
  h_d_box.sz = atoi(argv[2]);
  h_q_box.sz = atoi(argv[4]);
  d_rt.fanout = atoi(argv[5]);
  const int key_mode = argc == 7 ? atoi(argv[6]) : CURTREE_KEY_REF32;

  readMBR(argv[1], h_d_box.sz, 0, h_d_box.xmin, h_d_box.ymin, h_d_box.xmax, h_d_box.ymax,
          h_d_box.id);
  box_h2d(d_r_box, h_d_box);
  const int r_sz = build_rtree(d_r_box, d_rt, key_mode);
  printf("build_rtree: %d nodes at fanout %d\n", r_sz, d_rt.fanout);

  // the build sorted d_r_box in place: the same box multiset, reordered
  BBOX<float> h_sorted;
  box_d2h(h_sorted, d_r_box);
  std::vector<int> ids_in(h_d_box.id, h_d_box.id + h_d_box.sz),
      ids_out(h_sorted.id, h_sorted.id + h_sorted.sz);
  std::sort(ids_in.begin(), ids_in.end());
  std::sort(ids_out.begin(), ids_out.end());
  bool ok = ids_in == ids_out && r_sz == d_rt.sz;
  box_h_free(h_sorted);

  // queries: the query file read with r = 10, as the reference driver does
  readMBR(argv[3], h_q_box.sz, 10, h_q_box.xmin, h_q_box.ymin, h_q_box.xmax, h_q_box.ymax,
          h_q_box.id);
  box_h2d(d_q_box, h_q_box);

  // verification: nested loop over the unsorted input, query-major
  std::vector<int> c_did, c_qid;
  for (int i = 0; i < h_q_box.sz; i++)
    for (int j = 0; j < h_d_box.sz; j++)
      if (overlap(h_q_box.xmin[i], h_q_box.ymin[i], h_q_box.xmax[i], h_q_box.ymax[i],
                  h_d_box.xmin[j], h_d_box.ymin[j], h_d_box.xmax[j], h_d_box.ymax[j])) {
        c_did.push_back(h_d_box.id[j]);
        c_qid.push_back(i);  // the query's position (the GPU reports positions)
      }
  host_sort(c_did, c_qid);
  printf("num_h=%zu\n", c_did.size());
  ok = run_query("bfs_frontier", false, d_rt, d_r_box, d_q_box, c_did, c_qid) && ok;
  ok = run_query("bfs_thrust", true, d_rt, d_r_box, d_q_box, c_did, c_qid) && ok;
  printf(ok ? "verified successfully\n" : "verification failed\n");

  box_d_free(d_r_box);
  box_h_free(h_d_box);
  box_d_free(d_q_box);
  box_h_free(h_q_box);
  rt_d_free(d_rt);
  return ok ? 0 : 1;
}
