/*
 * curtree.hpp -- drop-in C++ API over libcurtree_amd.so for code written
 * against the reference's header templates:
 *
 *   build_rtree<float>   src/rtree_bulkload_thrust.cuh:62-213
 *   bfs_frontier<float>  src/rtree_bfs_frontier.cuh:292-529
 *   bfs_thrust<float>    src/rtree_bfs_thrust.cuh:58-256
 *   box_h2d, box_d2h, box_d_free, box_h_free, rt_d_alloc, rt_d_free,
 *   idpair_d_alloc, idpair_d2h, idpair_d_free, idpair_h_free
 *                        src/utility.cuh:16-205
 *   readMBR              src/utility.h:7-14, src/utility.cpp:5-40
 *
 * Same signatures, ownership and synchronous semantics (null stream, complete
 * on return).  Errors print the library's message and exit(EXIT_FAILURE), as
 * the reference's HANDLE_ERROR does (src/utility.cuh:8-14), unless
 * CURTREE_NO_EXIT is defined, in which case they throw std::runtime_error.
 * Unlike the reference, the queries never abort on large frontiers.
 * Only T = float is provided (the only instantiation the reference uses).
 */
#pragma once

#include <stdio.h>
#include <stdlib.h>

#include <cstddef>
#include <stdexcept>
#include <string>

#include "curtree_c.h"
#include "rtree.h"

static_assert(sizeof(BBOX<float>) == sizeof(curtree_bbox_f32), "BBOX layout");
static_assert(offsetof(BBOX<float>, id) == offsetof(curtree_bbox_f32, id), "BBOX layout");
static_assert(sizeof(RTREE<float>) == sizeof(curtree_rtree_f32), "RTREE layout");
static_assert(offsetof(RTREE<float>, len) == offsetof(curtree_rtree_f32, len), "RTREE layout");
static_assert(sizeof(IDPAIR) == sizeof(curtree_idpair), "IDPAIR layout");

namespace curtree_detail {
inline void check(int rc, const char* what) {
  if (rc == CURTREE_OK) return;
#ifdef CURTREE_NO_EXIT
  throw std::runtime_error(std::string(what) + ": " + curtree_last_error());
#else
  printf("%s failed: %s\n", what, curtree_last_error());
  exit(EXIT_FAILURE);
#endif
}
inline curtree_bbox_f32* cb(BBOX<float>& b) { return reinterpret_cast<curtree_bbox_f32*>(&b); }
inline const curtree_bbox_f32* cb(const BBOX<float>& b) {
  return reinterpret_cast<const curtree_bbox_f32*>(&b);
}
inline curtree_rtree_f32* cr(RTREE<float>& r) { return reinterpret_cast<curtree_rtree_f32*>(&r); }
inline const curtree_rtree_f32* cr(const RTREE<float>& r) {
  return reinterpret_cast<const curtree_rtree_f32*>(&r);
}
inline void query(const RTREE<float>& rt, const BBOX<float>& rbox, const BBOX<float>& qbox,
                  IDPAIR& out, int algo) {
  curtree_idpair64 p;
  check(curtree_query_f32(cr(rt), cb(rbox), cb(qbox), &p, algo, nullptr), "bfs query");
  if (p.sz > 0x7FFFFFFF) {
    curtree_idpair64_d_free(&p);
    check(CURTREE_ERANGE, "bfs query: more than 2^31-1 pairs for IDPAIR (use curtree_query_f32)");
  }
  out.sz = (int)p.sz;
  out.fid = p.fid;
  out.tid = p.tid;
}
}  // namespace curtree_detail

/* Bulk load: sorts d_box in place, allocates and fills d_rt (d_rt.fanout set
 * by the caller), returns R_sz.  key_mode defaults to the reference's key. */
template <class T>
int build_rtree(BBOX<T>& d_box, RTREE<T>& d_rt, int key_mode = CURTREE_KEY_REF32) {
  static_assert(sizeof(T) == sizeof(float), "only float boxes are supported");
  curtree_detail::check(curtree_build_f32(curtree_detail::cb(d_box), curtree_detail::cr(d_rt),
                                          key_mode, nullptr),
                        "build_rtree");
  return d_rt.sz;
}

/* d_dq: (box id, query index) pairs, allocated here (free with idpair_d_free). */
template <class T>
void bfs_frontier(const RTREE<T>& d_rt, const BBOX<T>& d_rbox, const BBOX<T>& d_qbox,
                  IDPAIR& d_dq) {
  curtree_detail::query(d_rt, d_rbox, d_qbox, d_dq, CURTREE_ALGO_FRONTIER);
}

template <class T>
void bfs_thrust(const RTREE<T>& d_rt, const BBOX<T>& d_rbox, const BBOX<T>& d_qbox,
                IDPAIR& d_dq) {
  curtree_detail::query(d_rt, d_rbox, d_qbox, d_dq, CURTREE_ALGO_FRONTIER);
}

/* ---- src/utility.cuh helpers ------------------------------------------- */

template <typename T>
void box_h2d(BBOX<T>& d_box, const BBOX<T>& h_box) {
  curtree_detail::check(curtree_box_h2d(curtree_detail::cb(d_box), curtree_detail::cb(h_box)),
                        "box_h2d");
}

template <typename T>
void box_d2h(BBOX<T>& h_box, const BBOX<T>& d_box) {
  int sz = d_box.sz;
  h_box.sz = sz;
  h_box.xmin = new T[sz];
  h_box.ymin = new T[sz];
  h_box.xmax = new T[sz];
  h_box.ymax = new T[sz];
  h_box.id = new int[sz];
  curtree_detail::check(curtree_box_d2h(curtree_detail::cb(h_box), curtree_detail::cb(d_box)),
                        "box_d2h");
}

template <typename T>
void box_d_free(BBOX<T>& box) {
  curtree_box_d_free(curtree_detail::cb(box));
}

/* host arrays from box_d2h / the caller's new[] */
template <typename T>
void box_h_free(BBOX<T>& box) {
  delete[] box.xmin;
  delete[] box.ymin;
  delete[] box.xmax;
  delete[] box.ymax;
  delete[] box.id;
  box.xmin = box.ymin = box.xmax = box.ymax = nullptr;
  box.id = nullptr;
}

template <typename T>
void rt_d_alloc(unsigned sz, RTREE<T>& rt) {
  curtree_detail::check(curtree_rt_d_alloc(curtree_detail::cr(rt), (int)sz), "rt_d_alloc");
}

template <typename T>
void rt_d_free(RTREE<T>& rt) {
  curtree_rt_d_free(curtree_detail::cr(rt));
}

inline void idpair_d_alloc(unsigned sz, IDPAIR& ft) {
  curtree_detail::check(curtree_idpair_d_alloc(reinterpret_cast<curtree_idpair*>(&ft), (int)sz),
                        "idpair_d_alloc");
}

inline void idpair_d2h(IDPAIR& h_ft, const IDPAIR& d_ft) {
  h_ft.sz = d_ft.sz;
  h_ft.fid = new int[d_ft.sz > 0 ? d_ft.sz : 1];
  h_ft.tid = new int[d_ft.sz > 0 ? d_ft.sz : 1];
  curtree_idpair64 d{d_ft.sz, d_ft.fid, d_ft.tid};
  curtree_detail::check(curtree_idpair64_d2h(h_ft.fid, h_ft.tid, &d), "idpair_d2h");
}

inline void idpair_d_free(IDPAIR& ft) {
  curtree_idpair_d_free(reinterpret_cast<curtree_idpair*>(&ft));
}

inline void idpair_h_free(IDPAIR& ft) {
  delete[] ft.fid;
  delete[] ft.tid;
  ft.fid = ft.tid = nullptr;
}

/* ---- src/utility.h: readMBR ---------------------------------------------- */

/* Reads `size` lines "x1,y1,x2,y2,id,pid,nume" into host arrays allocated
 * here with new[] (free with box_h_free); box = [min(x1,x2)-r, max(x1,x2)+r] x
 * [min(y1,y2)-r, max(y1,y2)+r].  A missing file or a bad line is an error. */
inline void readMBR(const char* file, const int size, const float r, float*& xmin, float*& ymin,
                    float*& xmax, float*& ymax, int*& id) {
  const int n = size > 0 ? size : 0;
  xmin = new float[n > 0 ? n : 1];
  ymin = new float[n > 0 ? n : 1];
  xmax = new float[n > 0 ? n : 1];
  ymax = new float[n > 0 ? n : 1];
  id = new int[n > 0 ? n : 1];
  curtree_detail::check(curtree_read_mbr_csv(file, n, r, xmin, ymin, xmax, ymax, id), "readMBR");
}
