/*
 * rtree.h -- the reference's data model (src/rtree.h:3-29), declared here so
 * that callers of the reference's template API compile unchanged against this
 * library.  The interface has to be identical for a drop-in: the same three
 * structs, the same member order and types (int sizes, raw device pointers).
 */
#pragma once

template <class T>
struct BBOX {
  int sz;
  T* xmin;
  T* ymin;
  T* xmax;
  T* ymax;
  int* id;
};

template <class T>
struct RTREE {
  int sz;
  int fanout;
  T* xmin;
  T* ymin;
  T* xmax;
  T* ymax;
  int* pos;
  int* len;
};

struct IDPAIR {
  int sz;
  int* fid;
  int* tid;
};
