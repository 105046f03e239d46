/*
 * oracle.c -- CPU restatement of curtree's bulk-load + window-query path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library in
 * curtree_amd/ is compared against.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product never links or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the curtree source tree).  Built with -ffp-contract=off so
 * that every FP32 operation is one IEEE-754 round-to-nearest-even operation.
 *
 * Pins (tests/test_oracle.py):
 *   or_dilate16       exhaustive vs the 256-entry dilate table parsed from
 *                     src/z_order.cuh:20-42 (tests/golden/dilate_table.json)
 *   or_keys_ref32     closed forms: integer-coordinate data -> textbook Morton,
 *                     [0,1)^2 data -> key 0, saturation/NaN choices
 *   or_keys_morton64  closed form on exactly representable grids; extent pinned
 *                     by numpy min/max
 *   or_sort_perm      numpy stable argsort (mergesort) of the same keys
 *   or_level_plan     closed form and a brute-force loop (float plan for N<=2^24)
 *   or_build_levels   invariants (parent MBR = min/max of children, pos/len
 *                     tile each level) + numpy reduceat over groups
 *   or_bfs            equal, after canonical ordering, to or_brute_force; its
 *                     work counters (X, T) and per-depth frontier sizes equal
 *                     a numpy closed form (the depth-d frontier of a tree whose
 *                     parent MBRs contain their children is every (window, node
 *                     at depth d) pair that intersects)
 *   or_canonical_sort numpy lexsort by (query id, box id)
 *   or_brute_force    src/utility.h:22-45 semantics; checked against a numpy
 *                     vectorised brute force and, where it builds, against the
 *                     reference's own nested_loop_query (oracle/_ref)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* intersect: closed-interval overlap, src/utility.h:16-20 (== the device
 * versions at src/rtree_bfs_frontier.cuh:40-44 and src/rtree_bfs_thrust.cuh:44-47).
 * (x1,y1,x2,y2) is the query window, (m1,n1,m2,n2) the tested box. */
static inline int or_intersect(float x1, float y1, float x2, float y2,
                               float m1, float n1, float m2, float n2) {
  if (x1 > m2 || x2 < m1 || y1 > n2 || y2 < n1) return 0;
  return 1;
}

OR_API int or_intersect_one(float x1, float y1, float x2, float y2,
                            float m1, float n1, float m2, float n2) {
  return or_intersect(x1, y1, x2, y2, m1, n1, m2, n2);
}

/* ------------------------------------------------------------------------ */
/* Morton dilation.  The reference uses the table-driven method of Raman & Wise
 * (IEEE TC 2007), src/z_order.cuh:20-42,63-65.  The oracle uses the closed-form
 * binary-magic-number bit spread; the test checks all 65,536 inputs against the
 * reference's table. */
OR_API uint32_t or_dilate16(uint32_t x) {
  x &= 0xFFFFu;
  x = (x | (x << 8)) & 0x00FF00FFu;
  x = (x | (x << 4)) & 0x0F0F0F0Fu;
  x = (x | (x << 2)) & 0x33333333u;
  x = (x | (x << 1)) & 0x55555555u;
  return x;
}

OR_API uint64_t or_dilate32(uint64_t x) {
  x &= 0xFFFFFFFFull;
  x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
  x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
  x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return x;
}

/* z_order(x, y) = (dilate_2(y) << 1) | dilate_2(x): src/z_order.cuh:67-69 */
OR_API uint32_t or_zorder16(uint32_t x, uint32_t y) {
  return (or_dilate16(y) << 1) | or_dilate16(x);
}

/* (ushort)f as the device conversion does it: truncation toward zero, saturated
 * to [0, 65535], NaN -> 0.  src/rtree_bulkload_thrust.cuh:55-56 casts with
 * (ushort); out-of-range and NaN are undefined in C++, so this saturating choice
 * is this project's definition (documented, not pinned by the reference). */
static inline uint32_t or_f2u16(float v) {
  if (!(v > 0.0f)) return 0u;           /* NaN, negative, zero */
  if (v >= 65535.0f) return 65535u;
  return (uint32_t)v;                   /* truncation toward zero */
}

/* mbr2key: src/rtree_bulkload_thrust.cuh:52-60, applied by the transform at
 * :76-81 and stored into a device_vector<int> (:72), i.e. as a signed int32. */
OR_API void or_keys_ref32(int64_t n, const float* xmin, const float* ymin,
                          const float* xmax, const float* ymax, int32_t* key) {
  for (int64_t i = 0; i < n; i++) {
    float cx = (xmin[i] + xmax[i]) / 2.0f;
    float cy = (ymin[i] + ymax[i]) / 2.0f;
    uint32_t z = or_zorder16(or_f2u16(cx), or_f2u16(cy));
    key[i] = (int32_t)z;
  }
}

/* ------------------------------------------------------------------------ */
/* Min / max used for every MBR reduction.  The reference calls CUDA's float
 * min/max (fminf/fmaxf, src/rtree_bulkload_thrust.cuh:42-45): NaN operands are
 * ignored.  The choice between -0 and +0 and the NaN payload when both are NaN
 * are not fixed by C; this project defines -0 < +0 and a canonical quiet NaN,
 * which makes both functions commutative and associative on bit patterns. */
static inline float or_canon_nan(void) {
  uint32_t u = 0x7FC00000u;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

static inline float or_fmin(float a, float b) {
  if (a != a) return (b != b) ? or_canon_nan() : b;
  if (b != b) return a;
  if (a < b) return a;
  if (b < a) return b;
  return signbit(a) ? a : b;
}

static inline float or_fmax(float a, float b) {
  if (a != a) return (b != b) ? or_canon_nan() : b;
  if (b != b) return a;
  if (a > b) return a;
  if (b > a) return b;
  return signbit(a) ? b : a;
}

OR_API float or_fmin_one(float a, float b) { return or_fmin(a, b); }
OR_API float or_fmax_one(float a, float b) { return or_fmax(a, b); }

/* ------------------------------------------------------------------------ */
/* 64-bit Morton key (this project's MORTON64 mode; the reference has only the
 * 16-bit-per-axis key).  Centroids are quantised to 32 bits per axis over the
 * extent of all centroids, then interleaved with x in the even bits, y in the
 * odd bits (the reference's z_order convention, src/z_order.cuh:67-69). */
OR_API void or_centroid_extent(int64_t n, const float* xmin, const float* ymin,
                               const float* xmax, const float* ymax, float* ext) {
  float lox = INFINITY, loy = INFINITY, hix = -INFINITY, hiy = -INFINITY;
  for (int64_t i = 0; i < n; i++) {
    float cx = (xmin[i] + xmax[i]) * 0.5f;
    float cy = (ymin[i] + ymax[i]) * 0.5f;
    lox = or_fmin(lox, cx);
    hix = or_fmax(hix, cx);
    loy = or_fmin(loy, cy);
    hiy = or_fmax(hiy, cy);
  }
  if (!(lox <= hix)) { lox = 0.0f; hix = 0.0f; }
  if (!(loy <= hiy)) { loy = 0.0f; hiy = 0.0f; }
  ext[0] = lox; ext[1] = loy; ext[2] = hix; ext[3] = hiy;
}

OR_API float or_m64_scale(float lo, float hi) {
  float w = hi - lo;
  return (w > 0.0f) ? 4294967296.0f / w : 0.0f;
}

static inline uint32_t or_q32(float c, float lo, float s) {
  float t = (c - lo) * s;
  if (!(t > 0.0f)) return 0u;
  if (t >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)t;
}

/* ext == NULL: use the centroid extent of the input. */
OR_API void or_keys_morton64(int64_t n, const float* xmin, const float* ymin,
                             const float* xmax, const float* ymax, const float* ext_in,
                             uint64_t* key) {
  float ext[4];
  if (ext_in) memcpy(ext, ext_in, sizeof ext);
  else or_centroid_extent(n, xmin, ymin, xmax, ymax, ext);
  float sx = or_m64_scale(ext[0], ext[2]);
  float sy = or_m64_scale(ext[1], ext[3]);
  for (int64_t i = 0; i < n; i++) {
    float cx = (xmin[i] + xmax[i]) * 0.5f;
    float cy = (ymin[i] + ymax[i]) * 0.5f;
    uint64_t qx = or_q32(cx, ext[0], sx);
    uint64_t qy = or_q32(cy, ext[1], sy);
    key[i] = (or_dilate32(qy) << 1) | or_dilate32(qx);
  }
}

/* ------------------------------------------------------------------------ */
/* Stable sort permutation.  thrust::sort_by_key at
 * src/rtree_bulkload_thrust.cuh:83-86 sorts int32 keys (signed order) with
 * Thrust's stable radix path; perm[i] = input position of the i-th sorted box.
 * LSD radix sort, 8-bit digits, which is stable by construction. */
static void or_lsd(int64_t n, const uint64_t* k, int nbytes, uint32_t* perm) {
  uint32_t* tmp = (uint32_t*)malloc((size_t)(n ? n : 1) * sizeof(uint32_t));
  for (int64_t i = 0; i < n; i++) perm[i] = (uint32_t)i;
  for (int d = 0; d < nbytes; d++) {
    int64_t cnt[257];
    memset(cnt, 0, sizeof cnt);
    int sh = 8 * d;
    for (int64_t i = 0; i < n; i++) cnt[((k[perm[i]] >> sh) & 0xFF) + 1]++;
    for (int b = 0; b < 256; b++) cnt[b + 1] += cnt[b];
    for (int64_t i = 0; i < n; i++) tmp[cnt[(k[perm[i]] >> sh) & 0xFF]++] = perm[i];
    memcpy(perm, tmp, (size_t)n * sizeof(uint32_t));
  }
  free(tmp);
}

OR_API void or_sort_perm_i32(int64_t n, const int32_t* key, uint32_t* perm) {
  uint64_t* k = (uint64_t*)malloc((size_t)(n ? n : 1) * sizeof(uint64_t));
  for (int64_t i = 0; i < n; i++) k[i] = (uint32_t)key[i] ^ 0x80000000u; /* signed order */
  or_lsd(n, k, 4, perm);
  free(k);
}

OR_API void or_sort_perm_u64(int64_t n, const uint64_t* key, uint32_t* perm) {
  or_lsd(n, key, 8, perm);
}

/* Apply the permutation to the caller's BBOX arrays (the in-place reorder of
 * src/rtree_bulkload_thrust.cuh:83-86 payload zip (xmin,ymin,xmax,ymax,id)). */
OR_API void or_permute_boxes(int64_t n, const uint32_t* perm, float* xmin, float* ymin,
                             float* xmax, float* ymax, int32_t* id) {
  float* t = (float*)malloc((size_t)(n ? n : 1) * sizeof(float));
  float* arr[4] = {xmin, ymin, xmax, ymax};
  for (int a = 0; a < 4; a++) {
    for (int64_t i = 0; i < n; i++) t[i] = arr[a][perm[i]];
    memcpy(arr[a], t, (size_t)n * sizeof(float));
  }
  int32_t* ti = (int32_t*)t;
  for (int64_t i = 0; i < n; i++) ti[i] = id[perm[i]];
  memcpy(id, ti, (size_t)n * sizeof(int32_t));
  free(t);
}

/* ------------------------------------------------------------------------ */
/* Level plan: src/rtree_bulkload_thrust.cuh:101-115 repeats
 * num_nodes = ceil((float)length / d) until one node remains.  The oracle uses
 * integer ceil-division (equal to the float form for N <= 2^24 and for the
 * configured round N) and defines N == 1 as one root level holding the box
 * (the reference yields an empty tree there) and N == 0 as an empty tree.
 * levels[0] is the leaf-parent level.  Returns the number of levels, or -1 if
 * `cap` is too small. */
OR_API int or_level_plan(int64_t n, int fanout, int64_t* levels, int cap, int64_t* r_sz) {
  int nl = 0;
  int64_t len = n, tot = 0;
  if (n <= 0 || fanout < 2) { *r_sz = 0; return 0; }
  do {
    int64_t nodes = (len + fanout - 1) / fanout;
    if (nl >= cap) return -1;
    levels[nl++] = nodes;
    tot += nodes;
    len = nodes;
  } while (len != 1);
  *r_sz = tot;
  return nl;
}

/* The reference's own float-based count, for documenting where it differs. */
OR_API int or_level_plan_float(int64_t n, int fanout, int64_t* levels, int cap, int64_t* r_sz) {
  int nl = 0;
  int length = (int)n;
  int64_t tot = 0;
  if (n <= 1 || fanout < 2) { *r_sz = 0; return 0; }
  while (length != 1) {
    int num_nodes = (int)ceilf((float)length / (float)fanout);
    if (nl >= cap) return -1;
    levels[nl++] = num_nodes;
    tot += num_nodes;
    length = num_nodes;
  }
  *r_sz = tot;
  return nl;
}

/* Level-by-level reduce: reduce_trans (i / d) + reduce_op (min,min,max,max,
 * min-pos, count) driven by src/rtree_bulkload_thrust.cuh:32-50,126-171.
 * The node arrays are laid out root level first; the leaf-parent level is the
 * last block, [R_sz - levels[0], R_sz).  Leaf-parent g covers sorted boxes
 * [g*F, min(g*F+F, N)) with pos = R_sz + g*F (:140); a node of level k covers
 * F consecutive nodes of level k-1 with pos = absolute index of its first child
 * (:158).  Each group is reduced left to right. */
OR_API int or_build_levels(int64_t n, int fanout, const float* bxmin, const float* bymin,
                           const float* bxmax, const float* bymax, float* rxmin, float* rymin,
                           float* rxmax, float* rymax, int32_t* rpos, int32_t* rlen) {
  int64_t levels[64], r_sz;
  int nl = or_level_plan(n, fanout, levels, 64, &r_sz);
  if (nl <= 0) return nl;
  int64_t start = r_sz;
  int64_t child_start = 0; /* absolute index of the child level (boxes: 0) */
  int64_t child_n = n;
  for (int l = 0; l < nl; l++) {
    start -= levels[l];
    const float *cx0, *cy0, *cx1, *cy1;
    int64_t pos_base;
    if (l == 0) {
      cx0 = bxmin; cy0 = bymin; cx1 = bxmax; cy1 = bymax;
      pos_base = r_sz;
    } else {
      cx0 = rxmin + child_start; cy0 = rymin + child_start;
      cx1 = rxmax + child_start; cy1 = rymax + child_start;
      pos_base = child_start;
    }
    for (int64_t g = 0; g < levels[l]; g++) {
      int64_t b = g * fanout, e = b + fanout;
      if (e > child_n) e = child_n;
      float a0 = cx0[b], a1 = cy0[b], a2 = cx1[b], a3 = cy1[b];
      for (int64_t i = b + 1; i < e; i++) {
        a0 = or_fmin(a0, cx0[i]);
        a1 = or_fmin(a1, cy0[i]);
        a2 = or_fmax(a2, cx1[i]);
        a3 = or_fmax(a3, cy1[i]);
      }
      rxmin[start + g] = a0; rymin[start + g] = a1;
      rxmax[start + g] = a2; rymax[start + g] = a3;
      rpos[start + g] = (int32_t)(pos_base + b);
      rlen[start + g] = (int32_t)(e - b);
    }
    child_start = start;
    child_n = levels[l];
  }
  return nl;
}

/* ------------------------------------------------------------------------ */
/* Window query, level-synchronous BFS with an unbounded frontier: the
 * semantics of bfs_thrust (src/rtree_bfs_thrust.cuh:58-250) and of
 * bfs_frontier when it does not overflow (src/rtree_bfs_frontier.cuh:60-182,
 * 184-289, 385-512):
 *   - every query starts at the root, node 0, and the root MBR is not tested
 *     (:96-99 / :70);
 *   - a (query, node) entry whose children are nodes (pos < R_sz, :110) tests
 *     each child's MBR and keeps the intersecting ones, in child order;
 *   - a leaf-parent entry tests its boxes (sorted positions pos-R_sz ..
 *     pos-R_sz+len-1; :221-222 walks F boxes capped at N, the same boxes for a
 *     bulk-loaded tree) and emits (box id = rbox.id[position], query index).
 * Output order is (query, sorted box position) -- the order the reference
 * produces before its host sort.  Counters: X = (query, node) entries below
 * the root that passed their MBR test (every materialised frontier entry,
 * leaf-parents included), T = MBR + box tests, H = pairs.
 * Returns H, or -1 if out_cap is too small (then only the counters are valid). */
typedef struct { uint32_t q, node; } or_entry;

OR_API int64_t or_bfs(int64_t r_sz, const float* rxmin, const float* rymin, const float* rxmax,
                      const float* rymax, const int32_t* rpos, const int32_t* rlen,
                      const float* bxmin, const float* bymin, const float* bxmax,
                      const float* bymax, const int32_t* bid, int64_t nq, const float* qxmin,
                      const float* qymin, const float* qxmax, const float* qymax,
                      int32_t* out_fid, int32_t* out_tid, int64_t out_cap, int64_t* counters,
                      int64_t* depth_sizes, int depth_cap) {
  int64_t X = 0, T = 0, H = 0;
  int depth = 0;
  for (int d = 0; depth_sizes && d < depth_cap; d++) depth_sizes[d] = 0;
  if (r_sz <= 0 || nq <= 0) {
    if (counters) { counters[0] = 0; counters[1] = 0; counters[2] = 0; }
    return 0;
  }
  int64_t cap = nq, n = nq;
  or_entry* cur = (or_entry*)malloc((size_t)cap * sizeof(or_entry));
  for (int64_t q = 0; q < nq; q++) { cur[q].q = (uint32_t)q; cur[q].node = 0; }
  int overflow = 0, root = 1;
  while (n > 0) {
    /* all entries of one frontier sit on the same level of a balanced tree */
    int leaf = rpos[cur[0].node] >= r_sz;
    if (depth_sizes && depth < depth_cap) depth_sizes[depth] = n;
    depth++;
    if (!root) X += n;
    root = 0;
    if (leaf) {
      for (int64_t e = 0; e < n; e++) {
        uint32_t q = cur[e].q, nd = cur[e].node;
        float x1 = qxmin[q], y1 = qymin[q], x2 = qxmax[q], y2 = qymax[q];
        int64_t b0 = (int64_t)rpos[nd] - r_sz;
        for (int64_t b = b0; b < b0 + rlen[nd]; b++) {
          T++;
          if (or_intersect(x1, y1, x2, y2, bxmin[b], bymin[b], bxmax[b], bymax[b])) {
            if (H < out_cap) { out_fid[H] = bid[b]; out_tid[H] = (int32_t)q; }
            else overflow = 1;
            H++;
          }
        }
      }
      break;
    }
    int64_t ncap = 16, nn = 0;
    or_entry* nxt = (or_entry*)malloc((size_t)ncap * sizeof(or_entry));
    for (int64_t e = 0; e < n; e++) {
      uint32_t q = cur[e].q, nd = cur[e].node;
      float x1 = qxmin[q], y1 = qymin[q], x2 = qxmax[q], y2 = qymax[q];
      for (int32_t c = rpos[nd]; c < rpos[nd] + rlen[nd]; c++) {
        T++;
        if (or_intersect(x1, y1, x2, y2, rxmin[c], rymin[c], rxmax[c], rymax[c])) {
          if (nn == ncap) { ncap *= 2; nxt = (or_entry*)realloc(nxt, (size_t)ncap * sizeof(or_entry)); }
          nxt[nn].q = q; nxt[nn].node = (uint32_t)c; nn++;
        }
      }
    }
    free(cur);
    cur = nxt;
    n = nn;
  }
  free(cur);
  if (counters) { counters[0] = X; counters[1] = T; counters[2] = H; }
  return overflow ? -1 : H;
}

/* ------------------------------------------------------------------------ */
/* Brute force: nested_loop_query, src/utility.h:22-45.  Query-major, data
 * order; emits the data box's id and the query's position (the reference
 * emits q_box.id[i], which is the position when query ids are 0..Q-1).
 * Returns H, or -1 if out_cap is too small (H still counted). */
OR_API int64_t or_brute_force(int64_t nq, const float* qxmin, const float* qymin,
                              const float* qxmax, const float* qymax, int64_t n,
                              const float* bxmin, const float* bymin, const float* bxmax,
                              const float* bymax, const int32_t* bid, int32_t* out_fid,
                              int32_t* out_tid, int64_t out_cap) {
  int64_t H = 0;
  int overflow = 0;
  for (int64_t i = 0; i < nq; i++) {
    float x1 = qxmin[i], y1 = qymin[i], x2 = qxmax[i], y2 = qymax[i];
    for (int64_t j = 0; j < n; j++) {
      if (or_intersect(x1, y1, x2, y2, bxmin[j], bymin[j], bxmax[j], bymax[j])) {
        if (H < out_cap) { out_fid[H] = bid[j]; out_tid[H] = (int32_t)i; }
        else overflow = 1;
        H++;
      }
    }
  }
  return overflow ? -1 : H;
}

/* Count-only brute force over a query range, parallel over queries with
 * OpenMP when built with -fopenmp (the CPU baseline's all-cores leg). */
OR_API int64_t or_brute_count(int64_t q0, int64_t q1, const float* qxmin, const float* qymin,
                              const float* qxmax, const float* qymax, int64_t n,
                              const float* bxmin, const float* bymin, const float* bxmax,
                              const float* bymax, int threads) {
  int64_t H = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : H) num_threads(threads > 0 ? threads : 1)
#endif
  for (int64_t i = q0; i < q1; i++) {
    float x1 = qxmin[i], y1 = qymin[i], x2 = qxmax[i], y2 = qymax[i];
    int64_t h = 0;
    for (int64_t j = 0; j < n; j++)
      h += or_intersect(x1, y1, x2, y2, bxmin[j], bymin[j], bxmax[j], bymax[j]);
    H += h;
  }
  (void)threads;
  return H;
}

/* ------------------------------------------------------------------------ */
/* Canonical result order: the two host stable sorts at
 * src/rtree_test_frontier.cu:102-105 (by box id, then by query id) leave the
 * pairs in lexicographic (query id, box id) order. */
typedef struct { int32_t t, f; } or_pair;
static int or_pair_cmp(const void* a, const void* b) {
  const or_pair* x = (const or_pair*)a;
  const or_pair* y = (const or_pair*)b;
  if (x->t != y->t) return x->t < y->t ? -1 : 1;
  if (x->f != y->f) return x->f < y->f ? -1 : 1;
  return 0;
}

OR_API void or_canonical_sort(int64_t h, int32_t* fid, int32_t* tid) {
  or_pair* p = (or_pair*)malloc((size_t)(h ? h : 1) * sizeof(or_pair));
  for (int64_t i = 0; i < h; i++) { p[i].t = tid[i]; p[i].f = fid[i]; }
  qsort(p, (size_t)h, sizeof(or_pair), or_pair_cmp);
  for (int64_t i = 0; i < h; i++) { tid[i] = p[i].t; fid[i] = p[i].f; }
  free(p);
}

/* ------------------------------------------------------------------------ */
/* Whole bulk load in one call (the CPU baseline's host R-tree build):
 * key -> stable sort -> permute the caller's arrays -> level reduce.
 * key_mode 0 = REF32 (reference key), 1 = MORTON64.  Returns R_sz or -1. */
OR_API int64_t or_build_rtree(int64_t n, int fanout, int key_mode, float* xmin, float* ymin,
                              float* xmax, float* ymax, int32_t* id, float* rxmin, float* rymin,
                              float* rxmax, float* rymax, int32_t* rpos, int32_t* rlen) {
  int64_t levels[64], r_sz;
  if (or_level_plan(n, fanout, levels, 64, &r_sz) < 0) return -1;
  if (n <= 0) return 0;
  uint32_t* perm = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
  if (key_mode == 0) {
    int32_t* k = (int32_t*)malloc((size_t)n * sizeof(int32_t));
    or_keys_ref32(n, xmin, ymin, xmax, ymax, k);
    or_sort_perm_i32(n, k, perm);
    free(k);
  } else {
    uint64_t* k = (uint64_t*)malloc((size_t)n * sizeof(uint64_t));
    or_keys_morton64(n, xmin, ymin, xmax, ymax, NULL, k);
    or_sort_perm_u64(n, k, perm);
    free(k);
  }
  or_permute_boxes(n, perm, xmin, ymin, xmax, ymax, id);
  free(perm);
  or_build_levels(n, fanout, xmin, ymin, xmax, ymax, rxmin, rymin, rxmax, rymax, rpos, rlen);
  return r_sz;
}
