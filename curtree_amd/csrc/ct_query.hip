// ct_query.hip -- batched window query on gfx950: level-synchronous BFS whose
// frontier of (query, node) entries lives in HBM.
//
// Reference semantics: query_tree_bfs_frontier + count_hit/write_hit
// (src/rtree_bfs_frontier.cuh:60-182, 191-289, 385-512) and bfs_thrust
// (src/rtree_bfs_thrust.cuh:58-250): every query starts at the root (node 0,
// whose MBR is not tested), intersecting children are kept level by level,
// and leaf-parent entries test their boxes and emit (box id, query index).
//
// Each level is ONE kernel: G = next_pow2(F) lanes test the children of one
// frontier entry, survivors are ranked by wave ballot, wave totals by a block
// scan, and tiles are chained by a decoupled look-back, so the next frontier is
// written compacted AND in (query, node) order with one read of the input.  The
// leaf level runs count (test boxes, keep a hit mask per entry, chained 64-bit
// offsets) and write (emit pairs from the masks, no re-test) passes, so the
// output is allocated exactly and lands in (query, sorted box position) order.
#include "ct_internal.hpp"

namespace ct {

constexpr int kQThreads = 256;
constexpr int kQSteps = 8;  // wave-steps per tile

struct QCtrl {
  uint64_t total;
  uint64_t pad[7];
  uint32_t ticket[kEpochs + 1];
};

struct QTree {
  const float* x0;
  const float* y0;
  const float* x1;
  const float* y1;
  const int* pos;
  const int* len;
  const uint8_t* safe;  // per node: every box below is well-formed (or null: no shortcut)
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
struct QBoxes {
  const float* x0;
  const float* y0;
  const float* x1;
  const float* y1;
};

template <int G>
struct QShape {
  static constexpr int E = kWave / G;                   // entries per wave-step
  static constexpr int kPerWave = kQSteps * E;
  static constexpr int kPerTile = (kQThreads / kWave) * kPerWave;
  static constexpr uint64_t kGroupMask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
};

__global__ void k_fill_u64(uint64_t* p, int64_t n, uint64_t v) {
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = v;
}
__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v) {
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = v;
}

// Tile prologue/epilogue shared by the chained kernels: wave totals -> tile
// total -> look-back -> exclusive base of this wave.
__device__ __forceinline__ uint64_t chain_tile(uint64_t wcount, int64_t tile, int64_t ntiles,
                                               uint64_t* status, uint32_t epoch,
                                               uint64_t* total_out, uint64_t* s_wtot,
                                               uint64_t* s_base) {
  const int w = threadIdx.x >> 6, lane = lane_id();
  if (lane == 0) s_wtot[w] = wcount;
  __syncthreads();
  if (w == 0) {
    uint64_t tt = s_wtot[0] + s_wtot[1] + s_wtot[2] + s_wtot[3];
    uint64_t excl = 0;
    if (tile == 0) {
      if (lane == 0) st_agent(status, st_pack(2u * epoch + 2u, tt));
    } else {
      if (lane == 0) st_agent(status + tile, st_pack(2u * epoch + 1u, tt));
      excl = lookback64(status, tile, epoch);
      if (lane == 0) st_agent(status + tile, st_pack(2u * epoch + 2u, excl + tt));
    }
    if (lane == 0) {
      *s_base = excl;
      if (tile == ntiles - 1) *total_out = excl + tt;
    }
  }
  __syncthreads();
  uint64_t b = *s_base;
  for (int i = 0; i < w; i++) b += s_wtot[i];
  return b;
}

// Internal level: entries (query, node) whose children are nodes.  fin == null
// means the implicit first frontier (query e, root).
template <int G>
__global__ __launch_bounds__(kQThreads) void k_expand(int64_t n_in, const uint2* __restrict__ fin,
                                                      QTree tr, QBoxes qb,
                                                      uint2* __restrict__ fout, int64_t cap,
                                                      uint64_t* status, uint32_t epoch,
                                                      QCtrl* ctl, int64_t ntiles, int64_t r_sz) {
  using S = QShape<G>;
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_wtot[4];
  __shared__ uint64_t s_base;
  const int t = threadIdx.x, w = t >> 6, lane = lane_id();
  const int sub = lane / G, c = lane % G;
  for (;;) {  // tiles by ticket until all are taken (grid may be smaller)
  __syncthreads();  // the previous tile is done with s_tile, s_wtot, s_base
  if (t == 0) s_tile = atomicAdd(&ctl->ticket[epoch], 1u);
  __syncthreads();
  const int64_t tile = s_tile;
  if (tile >= ntiles) return;
  uint64_t masks[kQSteps];
  uint32_t qk[kQSteps], ck[kQSteps];
  uint64_t wcount = 0;
  const int64_t ebase = tile * S::kPerTile + (int64_t)w * S::kPerWave + sub;
#pragma unroll
  for (int k = 0; k < kQSteps; k++) {
    const int64_t e = ebase + k * S::E;
    bool hit = false;
    uint32_t q = 0, ci = 0;
    if (e < n_in) {
      uint32_t node = 0, full = 0;
      if (fin) {
        uint2 en = fin[e];
        q = en.x;
        node = en.y & ~kFull;
        full = en.y & kFull;
      } else {
        q = (uint32_t)e;
      }
      const int p = tr.pos[node], L = tr.len[node];
      if (c < L && (int64_t)p + L <= r_sz) {  // children must be nodes
        const uint32_t child = (uint32_t)(p + c);
        if (full) {
          hit = true;
          ci = child | kFull;
        } else {
          const float wx1 = qb.x0[q], wy1 = qb.y0[q], wx2 = qb.x1[q], wy2 = qb.y1[q];
          const float m1 = tr.x0[child], n1 = tr.y0[child], m2 = tr.x1[child], n2 = tr.y1[child];
          hit = intersect(wx1, wy1, wx2, wy2, m1, n1, m2, n2);
          ci = child;
          if (hit && tr.safe && contains(wx1, wy1, wx2, wy2, m1, n1, m2, n2) && tr.safe[child])
            ci = child | kFull;
        }
      }
    }
    const uint64_t m = ballot(hit);
    masks[k] = m;
    qk[k] = q;
    ck[k] = ci;
    wcount += (uint64_t)popc64(m);
  }
  uint64_t b = chain_tile(wcount, tile, ntiles, status, epoch, &ctl->total, s_wtot, &s_base);
  const uint64_t lt = lanemask_lt();
#pragma unroll
  for (int k = 0; k < kQSteps; k++) {
    const uint64_t m = masks[k];
    if ((m >> lane) & 1ull) {
      const uint64_t o = b + (uint64_t)popc64(m & lt);
      if (o < (uint64_t)cap) fout[o] = make_uint2(qk[k], ck[k]);
    }
    b += (uint64_t)popc64(m);
  }
  }
}

// Box c of leaf-parent entry e: does it intersect the entry's window?
__device__ __forceinline__ bool leaf_hit(int64_t e, int c, const uint2* __restrict__ fin,
                                         const QTree& tr, const QBoxes& qb, const QBoxes& rb,
                                         int64_t r_sz, int64_t nbox, uint32_t* qout,
                                         int64_t* biout) {
  uint32_t node = 0, full = 0, q;
  if (fin) {
    uint2 en = fin[e];
    q = en.x;
    node = en.y & ~kFull;
    full = en.y & kFull;
  } else {
    q = (uint32_t)e;
  }
  *qout = q;
  const int p = tr.pos[node], L = tr.len[node];
  const int64_t bi = (int64_t)p - r_sz + c;
  *biout = bi;
  if (c >= L || bi < 0 || bi >= nbox) return false;  // children must be boxes
  if (full) return true;
  return intersect(qb.x0[q], qb.y0[q], qb.x1[q], qb.y1[q], rb.x0[bi], rb.y0[bi], rb.x1[bi],
                   rb.y1[bi]);
}

// Leaf level, count pass: test the boxes of each leaf-parent entry, store the
// entry's hit mask and its exclusive pair offset (chained across tiles).
template <int G>
__global__ __launch_bounds__(kQThreads) void k_leaf_count(
    int64_t n_in, const uint2* __restrict__ fin, QTree tr, QBoxes qb, QBoxes rb, int64_t r_sz,
    int64_t nbox, uint64_t* __restrict__ emask, uint64_t* __restrict__ eoff, unsigned long long* counts,
    uint64_t* status, uint32_t epoch, QCtrl* ctl, int64_t ntiles) {
  using S = QShape<G>;
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_wtot[4];
  __shared__ uint64_t s_base;
  const int t = threadIdx.x, w = t >> 6, lane = lane_id();
  const int sub = lane / G, c = lane % G;
  for (;;) {  // tiles by ticket until all are taken (grid may be smaller)
  __syncthreads();  // the previous tile is done with s_tile, s_wtot, s_base
  if (t == 0) s_tile = atomicAdd(&ctl->ticket[epoch], 1u);
  __syncthreads();
  const int64_t tile = s_tile;
  if (tile >= ntiles) return;
  uint64_t masks[kQSteps];
  uint64_t wcount = 0;
  const int64_t ebase = tile * S::kPerTile + (int64_t)w * S::kPerWave + sub;
#pragma unroll
  for (int k = 0; k < kQSteps; k++) {
    const int64_t e = ebase + k * S::E;
    bool hit = false;
    uint32_t q = 0;
    if (e < n_in) {
      int64_t bi;
      hit = leaf_hit(e, c, fin, tr, qb, rb, r_sz, nbox, &q, &bi);
    }
    const uint64_t m = ballot(hit);
    masks[k] = m;
    wcount += (uint64_t)popc64(m);
    if (counts && c == 0 && e < n_in) {
      // per-query totals: one atomic per entry with hits
      const uint64_t me = (m >> (sub * G)) & S::kGroupMask;
      if (me) atomicAdd(counts + q, (unsigned long long)popc64(me));
    }
  }
  uint64_t b = chain_tile(wcount, tile, ntiles, status, epoch, &ctl->total, s_wtot, &s_base);
#pragma unroll
  for (int k = 0; k < kQSteps; k++) {
    const int64_t e = ebase + k * S::E;
    const uint64_t m = masks[k];
    const uint64_t me = (m >> (sub * G)) & S::kGroupMask;
    const uint64_t before = sub == 0 ? 0ull : (m & ((1ull << (sub * G)) - 1ull));
    if (c == 0 && e < n_in) {
      emask[e] = me;
      eoff[e] = b + (uint64_t)popc64(before);
    }
    b += (uint64_t)popc64(m);
  }
  }
}

// Leaf level, write pass: G lanes per entry; lane c writes pair for box c if
// bit c of the entry's mask is set.
template <int G>
__global__ __launch_bounds__(kQThreads) void k_leaf_write(int64_t n_in, const uint2* __restrict__ fin,
                                                          QTree tr, const int* __restrict__ rid,
                                                          int64_t r_sz,
                                                          const uint64_t* __restrict__ emask,
                                                          const uint64_t* __restrict__ eoff,
                                                          int* __restrict__ out_fid,
                                                          int* __restrict__ out_tid) {
  const int64_t total = n_in * G;
  for (int64_t gt = blockIdx.x * (int64_t)kQThreads + threadIdx.x; gt < total;
       gt += (int64_t)gridDim.x * kQThreads) {
    const int64_t e = gt / G;
    const int c = (int)(gt % G);
    const uint64_t m = emask[e];
    if (!((m >> c) & 1ull)) continue;
    uint32_t node = 0, q;
    if (fin) {
      uint2 en = fin[e];
      q = en.x;
      node = en.y & ~kFull;
    } else {
      q = (uint32_t)e;
    }
    const uint64_t o = eoff[e] + (uint64_t)popc64(m & ((1ull << c) - 1ull));
    const int64_t bi = (int64_t)tr.pos[node] - r_sz + c;
    out_fid[o] = rid[bi];
    out_tid[o] = (int)q;
  }
}

// Leaf level, single pass into a caller buffer of `cap` pairs: test, chain
// the tile offsets, write pairs in (query, sorted position) order.  Pairs
// beyond `cap` are dropped; the total is always counted.
template <int G>
__global__ __launch_bounds__(kQThreads) void k_leaf_fused(
    int64_t n_in, const uint2* __restrict__ fin, QTree tr, QBoxes qb, QBoxes rb,
    const int* __restrict__ rid, int64_t r_sz, int64_t nbox, int* __restrict__ out_fid,
    int* __restrict__ out_tid, int64_t cap, uint64_t* status, uint32_t epoch, QCtrl* ctl,
    int64_t ntiles) {
  using S = QShape<G>;
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_wtot[4];
  __shared__ uint64_t s_base;
  const int t = threadIdx.x, w = t >> 6, lane = lane_id();
  const int sub = lane / G, c = lane % G;
  for (;;) {
  __syncthreads();
  if (t == 0) s_tile = atomicAdd(&ctl->ticket[epoch], 1u);
  __syncthreads();
  const int64_t tile = s_tile;
  if (tile >= ntiles) return;
  uint64_t masks[kQSteps];
  uint32_t qk[kQSteps];
  int64_t bk[kQSteps];
  uint64_t wcount = 0;
  const int64_t ebase = tile * S::kPerTile + (int64_t)w * S::kPerWave + sub;
#pragma unroll
  for (int k = 0; k < kQSteps; k++) {
    const int64_t e = ebase + k * S::E;
    bool hit = false;
    uint32_t q = 0;
    int64_t bi = 0;
    if (e < n_in) hit = leaf_hit(e, c, fin, tr, qb, rb, r_sz, nbox, &q, &bi);
    const uint64_t m = ballot(hit);
    masks[k] = m;
    qk[k] = q;
    bk[k] = bi;
    wcount += (uint64_t)popc64(m);
  }
  uint64_t b = chain_tile(wcount, tile, ntiles, status, epoch, &ctl->total, s_wtot, &s_base);
  const uint64_t lt = lanemask_lt();
#pragma unroll
  for (int k = 0; k < kQSteps; k++) {
    const uint64_t m = masks[k];
    if ((m >> lane) & 1ull) {
      const uint64_t o = b + (uint64_t)popc64(m & lt);
      if (o < (uint64_t)cap) {
        out_fid[o] = rid[bk[k]];
        out_tid[o] = (int)qk[k];
      }
    }
    b += (uint64_t)popc64(m);
  }
  }
}

// Well-formedness flags (xmin <= xmax and ymin <= ymax for every box below a
// node; NaN counts as not well-formed).  Leaf-parents from their boxes,
// upper levels as the AND of their children.
template <int G>
__global__ __launch_bounds__(256) void k_safe_leaf_pow2(int64_t nbox, QBoxes rb, uint8_t* safe0) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  const bool wf = i >= nbox || (rb.x0[i] <= rb.x1[i] && rb.y0[i] <= rb.y1[i]);
  const uint64_t bad = ballot(!wf);
  const int lane = lane_id();
  const uint64_t gm = G == 64 ? ~0ull : ((1ull << G) - 1ull);
  if ((lane % G) == 0 && i < nbox) safe0[i / G] = ((bad >> (lane - lane % G)) & gm) == 0 ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_safe_leaf_generic(int64_t nbox, int F, QBoxes rb,
                                                           uint8_t* safe0, int64_t n0) {
  const int64_t g = blockIdx.x * 256ll + threadIdx.x;
  if (g >= n0) return;
  int64_t b = g * F, e = b + F < nbox ? b + F : nbox;
  uint8_t ok = 1;
  for (int64_t i = b; i < e; i++) ok &= (rb.x0[i] <= rb.x1[i] && rb.y0[i] <= rb.y1[i]) ? 1 : 0;
  safe0[g] = ok;
}

__global__ __launch_bounds__(256) void k_safe_up(int64_t cn, int F, const uint8_t* child,
                                                 uint8_t* parent, int64_t pn) {
  const int64_t g = blockIdx.x * 256ll + threadIdx.x;
  if (g >= pn) return;
  int64_t b = g * F, e = b + F < cn ? b + F : cn;
  uint8_t ok = 1;
  for (int64_t i = b; i < e; i++) ok &= child[i];
  parent[g] = ok;
}

// ---------------------------------------------------------------------------
// Host side.
// Blocks for a ticketed chained launch: enough to fill the chip; each block
// loops over tiles, so the grid stays far below the 2^32 work-item limit.
static unsigned chain_grid(int64_t ntiles) {
  return (unsigned)std::min<int64_t>(ntiles, 8192);
}

static int next_pow2_group(int F) {
  int g = 2;
  while (g < F) g <<= 1;
  return g;
}

// Fresh epoch for one chained launch of `ntiles` tiles; zeroes the status
// words and tickets when the array grows or the epoch counter wraps.
static int chain_epoch(Context& ctx, hipStream_t s, int64_t ntiles, uint32_t* epoch) {
  Workspace& ws = ctx.ws;
  int rc;
  bool zero = false;
  if ((size_t)ntiles > ctx.qstatus_words) {
    size_t words = std::max<size_t>((size_t)ntiles + ntiles / 2, 1 << 16);
    if ((rc = ws.reserve(WsQStatus, words * 8))) return rc;
    ctx.qstatus_words = words;
    zero = true;
  }
  if (!ws.ptr(WsQCtrl)) {
    if ((rc = ws.reserve(WsQCtrl, sizeof(QCtrl)))) return rc;
    zero = true;
  }
  if (zero || ctx.qepoch + 1 >= (uint32_t)kEpochs) {
    hipLaunchKernelGGL(k_fill_u64, dim3(1024), dim3(256), 0, s, (uint64_t*)ws.ptr(WsQStatus),
                       (int64_t)ctx.qstatus_words, 0ull);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, s,
                       (uint32_t*)ws.ptr(WsQCtrl), (int64_t)(sizeof(QCtrl) / 4), 0u);
    ctx.qepoch = 0;
  } else {
    ctx.qepoch++;
  }
  *epoch = ctx.qepoch;
  return CURTREE_OK;
}

static int read_total(Context& ctx, hipStream_t s, uint64_t* v) {
  static uint64_t* host = nullptr;
  if (!host) CT_HIP(hipHostMalloc((void**)&host, 64, hipHostMallocDefault));
  CT_HIP(hipMemcpyAsync(host, &((QCtrl*)ctx.ws.ptr(WsQCtrl))->total, 8, hipMemcpyDeviceToHost, s));
  CT_HIP(hipStreamSynchronize(s));
  *v = *host;
  return CURTREE_OK;
}

// Per-node well-formedness flags into WsQSafe (R_sz bytes).
static int safe_flags(Context& ctx, const QTree& tr, const QBoxes& rb, int64_t r_sz, int64_t nbox,
                      hipStream_t s) {
  (void)tr;
  Workspace& ws = ctx.ws;
  int rc;
  if ((rc = ws.reserve(WsQSafe, (size_t)r_sz))) return rc;
  uint8_t* safe = (uint8_t*)ws.ptr(WsQSafe);
  LevelPlan lp;
  level_plan(nbox, ctx.qfanout, &lp);
  const int F = ctx.qfanout;
  int64_t start = r_sz - lp.n[0];
  ctx.prof.mark(s, "k_safe");
  switch (F) {
#define CT_SAFE(g)                                                                              \
  case g:                                                                                       \
    hipLaunchKernelGGL(k_safe_leaf_pow2<g>, dim3((unsigned)((nbox + 255) / 256)), dim3(256), 0, s, \
                       nbox, rb, safe + start);                                                 \
    break;
    CT_SAFE(2) CT_SAFE(4) CT_SAFE(8) CT_SAFE(16) CT_SAFE(32) CT_SAFE(64)
#undef CT_SAFE
    default:
      hipLaunchKernelGGL(k_safe_leaf_generic, dim3((unsigned)((lp.n[0] + 255) / 256)), dim3(256),
                         0, s, nbox, F, rb, safe + start, lp.n[0]);
  }
  for (int l = 1; l < lp.nl; l++) {
    int64_t pstart = start - lp.n[l];
    hipLaunchKernelGGL(k_safe_up, dim3((unsigned)((lp.n[l] + 255) / 256)), dim3(256), 0, s,
                       lp.n[l - 1], F, (const uint8_t*)(safe + start), safe + pstart, lp.n[l]);
    start = pstart;
  }
  ctx.prof.mark(s, nullptr);
  CT_HIP(hipGetLastError());
  return CURTREE_OK;
}

template <int G>
static int run_query(Context& ctx, const QTree& tr_in, const QBoxes& qb, const QBoxes& rb,
                     const int* rid, int64_t r_sz, int64_t nbox, int nl, int64_t Q,
                     QueryOut* qo, hipStream_t s) {
  using S = QShape<G>;
  Workspace& ws = ctx.ws;
  Prof& pf = ctx.prof;
  int rc;
  QTree tr = tr_in;
  tr.safe = nullptr;
  if (qo->contain) {
    if ((rc = safe_flags(ctx, tr, rb, r_sz, nbox, s))) return rc;
    tr.safe = (const uint8_t*)ws.ptr(WsQSafe);
  }
  const uint2* cur = nullptr;  // implicit (query, root) frontier
  int64_t ncur = Q;
  WsSlot cur_slot = WsQFrontB;
  for (int l = 0; l + 1 < nl; l++) {
    const WsSlot out_slot = cur_slot == WsQFrontA ? WsQFrontB : WsQFrontA;
    const int64_t ntiles = (ncur + S::kPerTile - 1) / S::kPerTile;
    int64_t cap = (int64_t)(ws.cap[out_slot] / sizeof(uint2));
    if (cap < std::max<int64_t>(2 * ncur, 1 << 20)) {
      cap = std::max<int64_t>(2 * ncur, 1 << 20);
      if ((rc = ws.reserve(out_slot, (size_t)cap * sizeof(uint2)))) return rc;
    }
    uint64_t total = 0;
    for (int attempt = 0; attempt < 2; attempt++) {
      uint32_t epoch;
      if ((rc = chain_epoch(ctx, s, ntiles, &epoch))) return rc;
      pf.mark(s, "k_expand");
      hipLaunchKernelGGL(k_expand<G>, dim3(chain_grid(ntiles)), dim3(kQThreads), 0, s, ncur, cur, tr,
                         qb, (uint2*)ws.ptr(out_slot), cap, (uint64_t*)ws.ptr(WsQStatus), epoch,
                         (QCtrl*)ws.ptr(WsQCtrl), ntiles, r_sz);
      pf.mark(s, nullptr);
      CT_HIP(hipGetLastError());
      if ((rc = read_total(ctx, s, &total))) return rc;
      if (total <= (uint64_t)cap) break;
      // frontier larger than the buffer: grow and run the level again
      cap = (int64_t)(total + total / 4);
      if ((rc = ws.reserve(out_slot, (size_t)cap * sizeof(uint2)))) return rc;
    }
    if (total > 0xFFFFFFFFull)
      return set_error(CURTREE_ERANGE, "query: frontier of %llu entries exceeds 2^32",
                       (unsigned long long)total);
    ncur = (int64_t)total;
    cur = (const uint2*)ws.ptr(out_slot);
    cur_slot = out_slot;
    if (ncur == 0) break;
  }
  qo->n = 0;
  if (ncur == 0) {
    if (qo->counts)
      hipLaunchKernelGGL(k_fill_u64, dim3(256), dim3(256), 0, s, (uint64_t*)qo->counts, Q, 0ull);
    return CURTREE_OK;
  }
  const int64_t ntiles = (ncur + S::kPerTile - 1) / S::kPerTile;
  uint32_t epoch;
  if (!qo->alloc && !qo->count_only && !qo->counts && qo->fid) {
    // leaf level, single pass into the caller's buffer
    if ((rc = chain_epoch(ctx, s, ntiles, &epoch))) return rc;
    pf.mark(s, "k_leaf_fused");
    hipLaunchKernelGGL(k_leaf_fused<G>, dim3(chain_grid(ntiles)), dim3(kQThreads), 0, s, ncur, cur,
                       tr, qb, rb, rid, r_sz, nbox, qo->fid, qo->tid, qo->cap,
                       (uint64_t*)ws.ptr(WsQStatus), epoch, (QCtrl*)ws.ptr(WsQCtrl), ntiles);
    pf.mark(s, nullptr);
    CT_HIP(hipGetLastError());
    uint64_t H = 0;
    if ((rc = read_total(ctx, s, &H))) return rc;
    qo->n = (int64_t)H;
    qo->out_fid = qo->fid;
    qo->out_tid = qo->tid;
    if ((int64_t)H > qo->cap)
      return set_error(CURTREE_ERANGE, "query: %llu pairs exceed the buffer of %lld",
                       (unsigned long long)H, (long long)qo->cap);
    return CURTREE_OK;
  }
  // leaf level: count pass
  if ((rc = ws.reserve(WsQMask, (size_t)ncur * 8)) || (rc = ws.reserve(WsQOff, (size_t)ncur * 8)))
    return rc;
  if ((rc = chain_epoch(ctx, s, ntiles, &epoch))) return rc;
  if (qo->counts)
    hipLaunchKernelGGL(k_fill_u64, dim3(256), dim3(256), 0, s, (uint64_t*)qo->counts, Q, 0ull);
  pf.mark(s, "k_leaf_count");
  hipLaunchKernelGGL(k_leaf_count<G>, dim3(chain_grid(ntiles)), dim3(kQThreads), 0, s, ncur, cur, tr,
                     qb, rb, r_sz, nbox, (uint64_t*)ws.ptr(WsQMask), (uint64_t*)ws.ptr(WsQOff),
                     (unsigned long long*)qo->counts, (uint64_t*)ws.ptr(WsQStatus), epoch,
                     (QCtrl*)ws.ptr(WsQCtrl), ntiles);
  pf.mark(s, nullptr);
  CT_HIP(hipGetLastError());
  uint64_t H = 0;
  if ((rc = read_total(ctx, s, &H))) return rc;
  qo->n = (int64_t)H;
  if (qo->count_only || H == 0) return CURTREE_OK;
  int* fid = qo->fid;
  int* tid = qo->tid;
  if (qo->alloc) {
    CT_HIP(hipMalloc((void**)&fid, H * sizeof(int)));
    hipError_t e2 = hipMalloc((void**)&tid, H * sizeof(int));
    if (e2 != hipSuccess) {
      (void)hipFree(fid);
      return set_error(CURTREE_ENOMEM, "query: cannot allocate %llu result pairs",
                       (unsigned long long)H);
    }
  } else if ((int64_t)H > qo->cap) {
    return set_error(CURTREE_ERANGE, "query: %llu pairs exceed the buffer of %lld",
                     (unsigned long long)H, (long long)qo->cap);
  }
  qo->out_fid = fid;
  qo->out_tid = tid;
  const int64_t threads = ncur * G;
  const unsigned wgrid = (unsigned)std::min<int64_t>((threads + kQThreads - 1) / kQThreads, 1 << 16);
  pf.mark(s, "k_leaf_write");
  hipLaunchKernelGGL(k_leaf_write<G>, dim3(wgrid),
                     dim3(kQThreads), 0, s, ncur, cur, tr, rid, r_sz, (const uint64_t*)ws.ptr(WsQMask),
                     (const uint64_t*)ws.ptr(WsQOff), fid, tid);
  pf.mark(s, nullptr);
  CT_HIP(hipGetLastError());
  return CURTREE_OK;
}

int query_frontier_f32(Context& ctx, const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox,
                       const curtree_bbox_f32* qbox, QueryOut* qo, hipStream_t s) {
  if (!rt || !rbox || !qbox || !qo) return set_error(CURTREE_EINVAL, "query: null argument");
  const int64_t N = rbox->sz, Q = qbox->sz;
  const int F = rt->fanout;
  qo->n = 0;
  if (N < 0 || Q < 0) return set_error(CURTREE_EINVAL, "query: negative size");
  if (F < 2) return set_error(CURTREE_EINVAL, "query: fanout must be >= 2");
  if (F > 64)
    return set_error(CURTREE_EINVAL, "query: fanout %d > 64 is not supported by the GPU query", F);
  if (N == 0 || Q == 0) {
    if (qo->counts && Q > 0)
      hipLaunchKernelGGL(k_fill_u64, dim3(256), dim3(256), 0, s, (uint64_t*)qo->counts, Q, 0ull);
    return CURTREE_OK;
  }
  LevelPlan lp;
  level_plan(N, F, &lp);
  if (lp.r_sz != rt->sz)
    return set_error(CURTREE_EINVAL,
                     "query: tree size %d does not match %lld boxes at fanout %d (expected %lld)",
                     rt->sz, (long long)N, F, (long long)lp.r_sz);
  if (!rt->xmin || !rt->ymin || !rt->xmax || !rt->ymax || !rt->pos || !rt->len || !rbox->xmin ||
      !rbox->ymin || !rbox->xmax || !rbox->ymax || !rbox->id || !qbox->xmin || !qbox->ymin ||
      !qbox->xmax || !qbox->ymax)
    return set_error(CURTREE_EINVAL, "query: null device array");
  QTree tr{rt->xmin, rt->ymin, rt->xmax, rt->ymax, rt->pos, rt->len, nullptr};
  ctx.qfanout = F;
  if (ctx.opt_contain == 0) qo->contain = false;
  else if (ctx.opt_contain == 1) qo->contain = true;
  else qo->contain = Q >= 256;  // the flags cost one pass over the boxes
  QBoxes qb{qbox->xmin, qbox->ymin, qbox->xmax, qbox->ymax};
  QBoxes rb{rbox->xmin, rbox->ymin, rbox->xmax, rbox->ymax};
  switch (next_pow2_group(F)) {
    case 2: return run_query<2>(ctx, tr, qb, rb, rbox->id, lp.r_sz, N, lp.nl, Q, qo, s);
    case 4: return run_query<4>(ctx, tr, qb, rb, rbox->id, lp.r_sz, N, lp.nl, Q, qo, s);
    case 8: return run_query<8>(ctx, tr, qb, rb, rbox->id, lp.r_sz, N, lp.nl, Q, qo, s);
    case 16: return run_query<16>(ctx, tr, qb, rb, rbox->id, lp.r_sz, N, lp.nl, Q, qo, s);
    case 32: return run_query<32>(ctx, tr, qb, rb, rbox->id, lp.r_sz, N, lp.nl, Q, qo, s);
    default: return run_query<64>(ctx, tr, qb, rb, rbox->id, lp.r_sz, N, lp.nl, Q, qo, s);
  }
}

}  // namespace ct
