// ct_query.hpp -- query pipeline internals shared by the plain, batched and
// serving paths.
#pragma once

#include <type_traits>

#include "ct_internal.hpp"

namespace ct {

#ifndef CT_QTHREADS
#define CT_QTHREADS 256
#endif
constexpr int kQThreads = CT_QTHREADS;  // threads per block of the chained query kernels
constexpr int kQWaves = kQThreads / 64;
// Tile of a chained launch: 4 waves x STEPS wave-steps x (64/G) entries.
// STEPS (4, 8 or 16) and the grid (blocks looping over tiles, i.e. waves per
// CU) are process-wide tuning knobs (curtree_set_option), picked by the
// autotuner.
constexpr int kDefaultSteps = 8;
constexpr int kDefaultChainGrid = 2048;  // resident blocks: 8 per CU on 256 CUs
struct QTuning {
  int steps = kDefaultSteps;
  int grid = 0;      // blocks per chained launch; 0: by batch size, up to kDefaultChainGrid
  int staging = -1;  // AoS box staging: -1 auto, 0 off, 1 on
  int expand = -1;   // internal levels: -1 by run size, 0 chained scan, 1 count/scan/write
  int fused = 0;     // writing leaf stage: 1 one chained pass, 0 count / tile scan / write,
                     // -1 the chained pass for runs below kSplitMinQueries windows
  int xcd = 0;       // chain-free kernels: 1 each XCD walks a contiguous eighth of the tiles
  int order = -1;    // windows in Morton order of their centres: -1 from kOrderMinWindows, 0, 1
  int codes = -1;    // leaf codes of kept trees: -1 built from kCodesMinWindows, 0 off, 1 on
};
QTuning& qtuning();

// Boxes are read from the caller's SoA arrays, or from an AoS copy (one
// float4 per box: xmin, ymin, xmax, ymax) when the query stages one: a group's
// F consecutive boxes are then F*16 contiguous bytes read by one dwordx4 load
// per lane, where SoA needs four loads touching four half-used cache lines.
struct QBoxes {
  const float* x0 = nullptr;
  const float* y0 = nullptr;
  const float* x1 = nullptr;
  const float* y1 = nullptr;
  const float4* b4 = nullptr;  // AoS copy (or null)
};
struct QTree {
  QBoxes mbr;                   // node MBRs
  const int* pos = nullptr;
  const int* len = nullptr;
  const uint8_t* safe = nullptr;  // per node: every box below is well-formed
  // or, for a tree built by this library, one device word: every box of the
  // tree is well-formed (both null: no containment shortcut)
  const uint32_t* allwf = nullptr;
  // Device word, nonzero when every node's pos/len is the bulk-load layout
  // (children of node i of a level are i*F .. i*F+F-1 of the level below):
  // child ranges are then computed, not loaded.  Null: always load.
  const uint32_t* canon = nullptr;
  // Leaf codes (or both null): per leaf-parent a frame and G slots of four
  // 8-bit cells of that frame per box (see box_cells); bulk-load layout only.
  const float4* lframe = nullptr;
  const uint32_t* lcode = nullptr;
  // Node codes (with the leaf codes): per internal node (0 .. lstart[0]-1) a
  // frame and G slots of its children's cells.
  const float4* nframe = nullptr;
  const uint32_t* ncode = nullptr;
  int F = 0;
};

// Leaf codes.  A leaf-parent's frame is (x0, y0, sx, sy): the low corner of
// its MBR and kCells / extent per axis (sx = 0: no codes for this node).  A
// coordinate v maps to a cell of that frame by one function, the same for
// boxes and windows:
//   t = (v - x0) * sx + 1, rounded after each step (no contraction);
//   cell = t clamped to [0, 255] (v_med3), truncated.
// The map is monotone (v <= w gives cell(v) <= cell(w)); nothing else about it
// matters.  The box test is "not separated": !(lo > hi) for the four pairs
// (window low side, box high side) and (box low side, window high side).
// cell(lo) > cell(hi) proves a separation, cell(lo) < cell(hi) proves there is
// none, equal cells leave the pair open, and only then is the box itself
// read.  NaN takes cell 0 on the low side (xmin, ymin) and cell 255 on the
// high side (xmax, ymax), so a pair with a NaN is never proven separated, as
// the float compare (false) says.
constexpr float kCells = 254.f;
__device__ __forceinline__ uint32_t cell_of(float v, float x0, float sx, bool low_side) {
  if (v != v) return low_side ? 0u : 255u;
  const float t = __fadd_rn(__fmul_rn(__fsub_rn(v, x0), sx), 1.f);
  return (uint32_t)__builtin_amdgcn_fmed3f(t, 0.f, 255.f);
}
// Cells of a box (or window): xmin | ymin << 8 | xmax << 16 | ymax << 24.
__device__ __forceinline__ uint32_t box_cells(float x1, float y1, float x2, float y2,
                                              const float4& fr) {
  return cell_of(x1, fr.x, fr.z, true) | (cell_of(y1, fr.y, fr.w, true) << 8) |
         (cell_of(x2, fr.x, fr.z, false) << 16) | (cell_of(y2, fr.y, fr.w, false) << 24);
}
// The low (xmin, ymin) and high (xmax, ymax) cells of c as two 16-bit lanes
// (x in the low half, y in the high half).
typedef short ct_s16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ ct_s16x2 cells_low(uint32_t c) {
  return __builtin_bit_cast(ct_s16x2, __builtin_amdgcn_perm(0u, c, 0x0c010c00u));
}
__device__ __forceinline__ ct_s16x2 cells_high(uint32_t c) {
  return __builtin_bit_cast(ct_s16x2, __builtin_amdgcn_perm(0u, c, 0x0c030c02u));
}
// 1: both lanes of m positive; 0: a lane negative; -1: otherwise (a zero lane).
__device__ __forceinline__ int lanes_sign(ct_s16x2 m) {
  const uint32_t u = __builtin_bit_cast(uint32_t, m);
  if (u & 0x80008000u) return 0;
  const ct_s16x2 one = {1, 1};
  return (__builtin_bit_cast(uint32_t, m - one) & 0x80008000u) ? -1 : 1;
}
// Intersection (no separation) of window cells (low wl, high wh) and box
// cells b: 1 proven, 0 disproven (a separation is proven), -1 open.  Per axis
// box high - window low and window high - box low: a negative difference
// proves a separation, all positive prove none.
__device__ __forceinline__ int cells_intersect(ct_s16x2 wl, ct_s16x2 wh, uint32_t b) {
  const ct_s16x2 d1 = cells_high(b) - wl, d2 = wh - cells_low(b);
  return lanes_sign(__builtin_elementwise_min(d1, d2));
}
__device__ __forceinline__ int cells_intersect(uint32_t w, uint32_t b) {
  return cells_intersect(cells_low(w), cells_high(w), b);
}

// Tiles of block blockIdx.x in a chain-free launch: first, first + step, ...
// below end.  xcd: blocks are dispatched to the 8 XCDs round-robin (block b
// on XCD b % 8), and XCD x walks the x-th contiguous eighth of the tiles, so
// the tiles in flight on one XCD are neighbours in the frontier (nearby
// windows when the windows are in spatial order) and share that XCD's L2.
struct TileWalk {
  int64_t first, end, step;
};
__device__ __forceinline__ TileWalk tile_walk(int64_t ntiles, bool xcd) {
  if (!xcd || (gridDim.x & 7u) != 0) return TileWalk{(int64_t)blockIdx.x, ntiles, gridDim.x};
  const int64_t slab = (ntiles + 7) / 8, x = blockIdx.x & 7u;
  const int64_t end = (x + 1) * slab < ntiles ? (x + 1) * slab : ntiles;
  return TileWalk{x * slab + (blockIdx.x >> 3), end, (int64_t)(gridDim.x >> 3)};
}

__device__ __forceinline__ float4 ld_box(const QBoxes& b, int64_t i) {
  if (b.b4) return b.b4[i];
  return make_float4(b.x0[i], b.y0[i], b.x1[i], b.y1[i]);
}

// Row broadcast: every lane receives the value of the first lane of its row of
// 16 lanes (DPP row_newbcast:0).  Called with every lane of the wave active.
__device__ __forceinline__ uint32_t row_first(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x150, 0xF, 0xF, false);
}
__device__ __forceinline__ float row_first(float v) {
  return __uint_as_float(row_first(__float_as_uint(v)));
}

// Frontier entries are (query, node | kFull).  kFull marks a node whose MBR
// lies inside the query window and whose boxes are all well-formed
// (xmin <= xmax, ymin <= ymax): every box below it then intersects the window
// (box inside its leaf-parent MBR inside the window), so its subtree is
// expanded without reading child MBRs or boxes.  Results are identical to
// testing every child, as the reference does.
constexpr uint32_t kFull = 0x80000000u;

// The containment shortcut is allowed on this tree (call once per kernel),
// and the subtree of `node` holds only well-formed boxes.
__device__ __forceinline__ bool shortcut_on(const QTree& tr, bool canon) {
  return canon && (tr.safe != nullptr || (tr.allwf != nullptr && *tr.allwf != 0u));
}
__device__ __forceinline__ bool node_safe(const QTree& tr, uint32_t node) {
  return tr.safe == nullptr || tr.safe[node] != 0;
}

__device__ __forceinline__ bool contains(float x1, float y1, float x2, float y2, float m1,
                                         float n1, float m2, float n2) {
  return x1 <= m1 && m2 <= x2 && y1 <= n1 && n2 <= y2;
}

template <int G, int STEPS>
struct QShape {
  static constexpr int E = kWave / G;  // entries per wave-step
  static constexpr int kPerWave = STEPS * E;
  static constexpr int kPerTile = (kQThreads / kWave) * kPerWave;
  static constexpr uint64_t kGroupMask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
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
  int steps = kDefaultSteps;      // wave-steps per tile: 4, 8 or 16
  int grid = kDefaultChainGrid;   // blocks of each chained launch
  int64_t q0 = 0, q1 = 0;
  const int64_t* plan = nullptr;  // take (q0, q1) from the next plan chunk instead
  bool write = false;             // leaf stage writes pairs (else counts only)
  bool split = false;             // internal levels as count/scan/write (no tile chain)
  bool chained = false;           // every level one chained launch (CURTREE_ALGO_PERSISTENT)
  bool generic = false;           // fanout above 64: one wave per entry (k_gen_*)
  int* out_fid = nullptr;
  int* out_tid = nullptr;
  int64_t out_cap = 0;
  unsigned long long* counts = nullptr;  // per-query totals (count pass)
  // windows in a run order (tmap: the input window at each rank; the run's
  // windows and query indices are ranks): a window's first pair in input
  // order (ooff) and in run order (poff), both by rank; the leaf write moves
  // each entry by their difference and writes input indices
  const uint32_t* tmap = nullptr;
  int64_t* ooff = nullptr;
  int64_t* poff = nullptr;
  int64_t* osum = nullptr;  // two block sums per kOrdBlock windows
};

int next_pow2_group(int F);
int per_tile(int G, int steps);
// Validate, fill `a` (tree/boxes/levels) and compute the containment flags.
// canon: the device word of a layout check the caller made once for this
// tree (check_layout), used instead of checking again; null: checked here
// where the run needs it.
int prepare_query(QEngine& qe, Prof& pf, const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox,
                  const curtree_bbox_f32* qbox, int contain, int64_t nq_hint, QRunArgs* a,
                  hipStream_t s, const uint32_t* canon = nullptr);
// Well-formedness flags of every node; with box4 (power-of-two F) the same
// pass writes the AoS copy of the boxes and sets *staged.
int compute_safe(QEngine& qe, Prof& pf, const QBoxes& rb, int F, int64_t r_sz, int64_t nbox,
                 hipStream_t s, float4* box4 = nullptr, bool* staged = nullptr);
int reserve_f4(float4** dst, int64_t* cap, int64_t n);
// Checks that rt's pos/len arrays hold the bulk-load layout; the result lands
// in the device word *flag (allocated on first use), no host synchronisation.
int check_layout(uint32_t** flag, const curtree_rtree_f32* rt, int64_t nbox, hipStream_t s);
// AoS copy of n boxes into *dst (grown as needed; *cap in boxes).
int stage_boxes(float4** dst, int64_t* cap, const QBoxes& src, int64_t n, hipStream_t s);
// Enqueue one run (begin, expansions, leaf stage, piece writer); no host
// synchronisation.  Needs qe.cap_pieces >= pieces_needed(qe, a).
void enqueue_run(QEngine& qe, Prof& pf, const QRunArgs& a, hipStream_t s);
int64_t pieces_needed(const QEngine& qe, const QRunArgs& a);
void enqueue_fill_counts(int64_t* counts, int64_t n, hipStream_t s);
int read_ctrl(QEngine& qe, hipStream_t s);
// Codes of a bulk-load tree (levels lv, G >= F slots per parent): the frames
// of the leaf-parents, G u32 cells per leaf-parent, then the frames of the
// internal nodes (0 .. lstart[0]-1) and G cells of their children each, in
// one buffer of leaf_codes_bytes(leaf-parents, internal nodes, G).
int64_t leaf_codes_bytes(int64_t nleafparents, int64_t nparents, int G);
void enqueue_leaf_codes(const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox,
                        const QLevels& lv, int G, void* buf, hipStream_t s);
// Builds the leaf codes of a tree built by this library (kept), if it has
// none yet and they apply (fanout 3..64, option leaf_codes not 0); a tree
// that is not kept is left as it is.  Synchronises s.
int ensure_leaf_codes(const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox, hipStream_t s);

}  // namespace ct
