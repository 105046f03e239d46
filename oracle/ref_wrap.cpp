// ref_wrap.cpp -- C entry points over the reference's own CPU code, compiled
// from the curtree sources where they lie (see ref.mk).  Test infrastructure:
// the oracle's brute force and CSV reader are checked against these.
//
//   ref_nested_loop_query  -> nested_loop_query<float>, src/utility.h:22-45
//   ref_read_mbr           -> readMBR, src/utility.cpp:5-40
//
// rocThrust supplies thrust::host_vector; the headers below are the ones
// utility.h relies on transitively under CUDA's Thrust.
#include <stdint.h>
#include <string.h>
#include <thrust/equal.h>
#include <thrust/execution_policy.h>
#include <thrust/iterator/constant_iterator.h>
#include <thrust/reduce.h>

#include <iterator>

#include "utility.h"

extern "C" int64_t ref_nested_loop_query(int nq, float* qx0, float* qy0, float* qx1, float* qy1,
                                         int* qid, int n, float* x0, float* y0, float* x1,
                                         float* y1, int* id, int* out_fid, int* out_tid,
                                         int64_t cap) {
  BBOX<float> q{nq, qx0, qy0, qx1, qy1, qid};
  BBOX<float> d{n, x0, y0, x1, y1, id};
  thrust::host_vector<int> did, tid;
  int h = nested_loop_query<float>(q, d, did, tid);
  for (int64_t i = 0; i < h && i < cap; i++) {
    out_fid[i] = did[i];
    out_tid[i] = tid[i];
  }
  return h;
}

// Returns 0, or -1 when readMBR throws on a malformed line.
extern "C" int ref_read_mbr(char* file, int size, float r, float* xmin, float* ymin, float* xmax,
                            float* ymax, int* id) {
  float *a = nullptr, *b = nullptr, *c = nullptr, *d = nullptr;
  int* e = nullptr;
  try {
    readMBR(file, size, r, a, b, c, d, e);
  } catch (const std::exception&) {
    return -1;
  }
  memcpy(xmin, a, sizeof(float) * size);
  memcpy(ymin, b, sizeof(float) * size);
  memcpy(xmax, c, sizeof(float) * size);
  memcpy(ymax, d, sizeof(float) * size);
  memcpy(id, e, sizeof(int) * size);
  delete[] a;
  delete[] b;
  delete[] c;
  delete[] d;
  delete[] e;
  return 0;
}
