// ct_query.hpp -- query pipeline internals shared by the plain, batched and
// serving paths.
#pragma once

#include "ct_internal.hpp"

namespace ct {

constexpr int kQThreads = 256;
constexpr int kQSteps = 8;       // wave-steps per chained tile
constexpr int kChainGrid = 8192;  // blocks of a chained launch (they loop over tiles)

struct QTree {
  const float* x0;
  const float* y0;
  const float* x1;
  const float* y1;
  const int* pos;
  const int* len;
  const uint8_t* safe;  // per node: every box below is well-formed (or null: no shortcut)
};
struct QBoxes {
  const float* x0;
  const float* y0;
  const float* x1;
  const float* y1;
};

// Frontier entries are (query, node | kFull).  kFull marks a node whose MBR
// lies inside the query window and whose boxes are all well-formed
// (xmin <= xmax, ymin <= ymax): every box below it then intersects the window
// (box inside its leaf-parent MBR inside the window), so its subtree is
// expanded without reading child MBRs or boxes.  Results are identical to
// testing every child, as the reference does.
constexpr uint32_t kFull = 0x80000000u;

__device__ __forceinline__ bool contains(float x1, float y1, float x2, float y2, float m1,
                                         float n1, float m2, float n2) {
  return x1 <= m1 && m2 <= x2 && y1 <= n1 && n2 <= y2;
}

template <int G>
struct QShape {
  static constexpr int E = kWave / G;  // entries per wave-step
  static constexpr int kPerWave = kQSteps * E;
  static constexpr int kPerTile = (kQThreads / kWave) * kPerWave;
  static constexpr uint64_t kGroupMask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
};

// Level table of a bulk-loaded tree, leaf-parent level first: node i of level
// l (global index lstart[l] + i) covers sorted boxes [i*span[l], (i+1)*span[l]).
constexpr int kMaxLevels = 32;
struct QLevels {
  int nl;
  int64_t lstart[kMaxLevels];
  int64_t lcount[kMaxLevels];
  int64_t span[kMaxLevels];
};

// A piece of a contained subtree's box range, written by one block.
constexpr uint64_t kPiece = 4096;
struct QPiece {
  uint64_t off;    // first output slot
  int64_t first;   // first sorted box
  uint32_t len;    // boxes (<= kPiece)
  uint32_t q;      // query index
};

// Everything one run needs (device pointers and sizes).
struct QRunArgs {
  QLevels lv{};
  QTree tr{};
  QBoxes qb{};
  QBoxes rb{};
  const int* rid = nullptr;
  int64_t r_sz = 0, nbox = 0;
  int nl = 0, F = 0, G = 0;
  int64_t q0 = 0, q1 = 0;
  const int64_t* plan = nullptr;  // take (q0, q1) from the next plan chunk instead
  bool write = false;             // leaf stage writes pairs (else counts only)
  int* out_fid = nullptr;
  int* out_tid = nullptr;
  int64_t out_cap = 0;
  unsigned long long* counts = nullptr;  // per-query totals (count pass)
};

int next_pow2_group(int F);
int per_tile(int G);
// Validate, fill `a` (tree/boxes/levels) and compute the containment flags.
int prepare_query(QEngine& qe, Prof& pf, const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox,
                  const curtree_bbox_f32* qbox, int contain, int64_t nq_hint, QRunArgs* a,
                  hipStream_t s);
int compute_safe(QEngine& qe, Prof& pf, const QBoxes& rb, int F, int64_t r_sz, int64_t nbox,
                 hipStream_t s);
// Enqueue one run (begin, expansions, leaf stage, piece writer); no host
// synchronisation.  Needs qe.cap_pieces >= pieces_needed(qe, a).
void enqueue_run(QEngine& qe, Prof& pf, const QRunArgs& a, hipStream_t s);
int64_t pieces_needed(const QEngine& qe, const QRunArgs& a);
void enqueue_fill_counts(int64_t* counts, int64_t n, hipStream_t s);
int read_ctrl(QEngine& qe, hipStream_t s);

}  // namespace ct
