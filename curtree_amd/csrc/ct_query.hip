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
// written compacted AND in (query, node) order with one read of the input.
// Tiles are taken from a ticket counter by a capped grid of blocks that loop
// until the level's tiles run out (work is balanced by (query, node) entry, not
// by query, so one window with 1e5 hits spreads over many blocks).
//
// All sizes of a run live in device memory (QCtrl): a run is begin -> one
// expansion per internal level -> leaf stage, with no host round trip, so the
// batched path can capture a run in a hipGraph and replay it per chunk.  The
// leaf stage either writes pairs in one pass into a buffer of known capacity,
// or counts (hit mask + 64-bit offset per entry) and writes in a second pass
// after the exact output was allocated.  Output order is (query, sorted box
// position), the reference's pre-sort order.
#include "ct_query.hpp"

namespace ct {

__global__ void k_fill_u64(uint64_t* p, int64_t n, uint64_t v) {
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = v;
}

// Fresh control block: the first run clears the status words.
__global__ void k_ctrl_init(QCtrl* c) {
  if (threadIdx.x == 0) {
    c->epoch_next = kEpochs;
    c->chunk = 0;
    c->overflow = 0;
    c->total = 0;
  }
}

// Start of a run: query range (from the arguments, or the next chunk of
// `plan`), epochs for its `nslots` chained launches, cleared tickets/counts.
__global__ void k_run_begin(QCtrl* c, int64_t q0, int64_t q1, const int64_t* plan,
                            uint32_t nslots) {
  if (threadIdx.x == 0) {
    if (plan) {
      const uint32_t k = c->chunk++;
      q0 = plan[2 * k];
      q1 = plan[2 * k + 1];
    }
    uint32_t eb = c->epoch_next;
    if (eb + nslots > (uint32_t)kEpochs) {
      eb = 0;
      c->zero_needed = 1;
    } else {
      c->zero_needed = 0;
    }
    c->epoch_base = eb;
    c->epoch_next = eb + nslots;
    c->q0 = q0;
    c->q1 = q1;
    c->count[0] = (uint64_t)(q1 - q0);
    c->overflow = 0;
    c->total = 0;
    c->npieces = 0;
  }
  for (int i = threadIdx.x; i < kMaxRunSlots; i += blockDim.x) {
    c->ticket[i] = 0;
    if (i > 0) c->count[i] = 0;
  }
}

__global__ void k_status_zero(const QCtrl* c, uint64_t* status, int64_t words) {
  if (!c->zero_needed) return;
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < words; i += (int64_t)gridDim.x * 256)
    status[i] = 0;
}

// Tile epilogue shared by the chained kernels: wave totals -> tile total ->
// look-back -> exclusive base of this wave.  The last tile leaves the grand
// total in *s_total.
__device__ __forceinline__ uint64_t chain_tile(uint64_t wcount, int64_t tile, int64_t ntiles,
                                               uint64_t* status, uint32_t epoch, uint64_t* s_wtot,
                                               uint64_t* s_base, uint64_t* s_total) {
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
      if (tile == ntiles - 1) *s_total = excl + tt;
    }
  }
  __syncthreads();
  uint64_t b = *s_base;
  for (int i = 0; i < w; i++) b += s_wtot[i];
  return b;
}

// Take the next tile; returns -1 when the level's tiles are exhausted.
__device__ __forceinline__ int64_t next_tile(uint32_t* ticket, int64_t ntiles, uint32_t* s_tile) {
  __syncthreads();  // the previous tile is done with the shared words
  if (threadIdx.x == 0) *s_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  const int64_t t = *s_tile;
  return t < ntiles ? t : -1;
}

// Internal level `slot`: entries (query, node) whose children are nodes.
// fin == null: the implicit first frontier (query q0 + e, root).
template <int G>
__global__ __launch_bounds__(kQThreads) void k_expand(int slot, QTree tr, QBoxes qb, QCtrl* c,
                                                      const uint2* __restrict__ fin,
                                                      uint2* __restrict__ fout, int64_t cap,
                                                      uint64_t* status, int64_t r_sz) {
  using S = QShape<G>;
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_wtot[4];
  __shared__ uint64_t s_base, s_total;
  if (c->overflow) return;
  const int64_t n_in = (int64_t)c->count[slot];
  const uint32_t epoch = c->epoch_base + slot;
  const int64_t q0 = c->q0;
  const int64_t ntiles = (n_in + S::kPerTile - 1) / S::kPerTile;
  const int t = threadIdx.x, w = t >> 6, lane = lane_id();
  if (ntiles == 0) {
    if (blockIdx.x == 0 && t == 0) c->count[slot + 1] = 0;
    return;
  }
  const int sub = lane / G, c_ = lane % G;
  int64_t tile;
  while ((tile = next_tile(&c->ticket[slot], ntiles, &s_tile)) >= 0) {
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
          q = (uint32_t)(q0 + e);
        }
        if (full) {
          // a contained subtree travels down as one entry (its box range)
          hit = c_ == 0;
          ci = node | kFull;
        } else {
          const int p = tr.pos[node], L = tr.len[node];
          if (c_ < L && (int64_t)p + L <= r_sz) {  // children must be nodes
            const uint32_t child = (uint32_t)(p + c_);
            const float wx1 = qb.x0[q], wy1 = qb.y0[q], wx2 = qb.x1[q], wy2 = qb.y1[q];
            const float m1 = tr.x0[child], n1 = tr.y0[child], m2 = tr.x1[child],
                        n2 = tr.y1[child];
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
    uint64_t b = chain_tile(wcount, tile, ntiles, status, epoch, s_wtot, &s_base, &s_total);
    if (tile == ntiles - 1 && t == 0) {
      c->count[slot + 1] = s_total;
      if (s_total > (uint64_t)cap) c->overflow = 1;
    }
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

// Box range [first, first + cnt) below node v (a complete F-ary tree over the
// sorted boxes: node i of level l covers boxes [i*F^(l+1), (i+1)*F^(l+1))).
__device__ __forceinline__ void node_range(uint32_t v, const QLevels& L, int64_t nbox,
                                           int64_t* first, int64_t* cnt) {
  int lv = 0;
#pragma unroll
  for (int k = 0; k < kMaxLevels; k++)
    if (k < L.nl && (int64_t)v >= L.lstart[k] && (int64_t)v < L.lstart[k] + L.lcount[k]) lv = k;
  const int64_t f = ((int64_t)v - L.lstart[lv]) * L.span[lv];
  const int64_t rem = nbox - f;
  *first = f;
  *cnt = rem < L.span[lv] ? rem : L.span[lv];
}

// Leaf stage.  Each entry is either a partial leaf-parent (its boxes are
// tested: G lanes, one box each) or a contained subtree (all of its box range
// hits, no reads).  Entry counts are prefix-summed across the wave, the waves
// of the tile and the tiles (chained), giving every pair its output slot.
//   COUNT: totals (and per-query totals) only.
//   WRITE: partial entries and short ranges are written by their lanes; longer
//          ranges are cut into pieces for k_write_pieces.  Slots beyond
//          out_cap are dropped; the total is always counted.
template <int G, bool WRITE>
__global__ __launch_bounds__(kQThreads) void k_leaf_pass(
    int slot, QTree tr, QBoxes qb, QBoxes rb, const int* __restrict__ rid, QLevels lv, QCtrl* c,
    const uint2* __restrict__ fin, int64_t r_sz, int64_t nbox, int* __restrict__ out_fid,
    int* __restrict__ out_tid, int64_t out_cap, QPiece* __restrict__ pieces, int64_t piece_cap,
    unsigned long long* counts, uint64_t* status) {
  using S = QShape<G>;
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_wtot[4];
  __shared__ uint64_t s_base, s_total;
  if (c->overflow) return;
  const int64_t n_in = (int64_t)c->count[slot];
  const uint32_t epoch = c->epoch_base + slot;
  const int64_t q0 = c->q0;
  const int64_t ntiles = (n_in + S::kPerTile - 1) / S::kPerTile;
  const int t = threadIdx.x, w = t >> 6, lane = lane_id();
  if (ntiles == 0) {
    if (blockIdx.x == 0 && t == 0) c->total = 0;
    return;
  }
  const int sub = lane / G, c_ = lane % G;
  const int lead = sub * G;  // first lane of this lane's group
  int64_t tile;
  while ((tile = next_tile(&c->ticket[slot], ntiles, &s_tile)) >= 0) {
    uint64_t gmask[kQSteps];   // the group's hit bits (partial entries)
    uint64_t ecnt[kQSteps];    // pairs of this lane's entry
    uint64_t epre[kQSteps];    // pairs of earlier entries in the same step
    uint64_t stot[kQSteps];    // pairs of the whole step
    int64_t efirst[kQSteps];   // first sorted box of the entry
    uint32_t qk[kQSteps];
    bool efull[kQSteps];
    uint64_t wcount = 0;
    const int64_t ebase = tile * S::kPerTile + (int64_t)w * S::kPerWave + sub;
#pragma unroll
    for (int k = 0; k < kQSteps; k++) {
      const int64_t e = ebase + k * S::E;
      bool hit = false, full = false;
      uint32_t q = 0;
      int64_t first = 0, rcnt = 0;
      if (e < n_in) {
        uint32_t node = 0;
        if (fin) {
          const uint2 en = fin[e];
          q = en.x;
          node = en.y & ~kFull;
          full = (en.y & kFull) != 0;
        } else {
          q = (uint32_t)(q0 + e);
        }
        if (full) {
          node_range(node, lv, nbox, &first, &rcnt);
        } else {
          const int p = tr.pos[node], L = tr.len[node];
          first = (int64_t)p - r_sz;
          const int64_t bi = first + c_;
          if (c_ < L && bi >= 0 && bi < nbox)  // children must be boxes
            hit = intersect(qb.x0[q], qb.y0[q], qb.x1[q], qb.y1[q], rb.x0[bi], rb.y0[bi],
                            rb.x1[bi], rb.y1[bi]);
        }
      }
      const uint64_t m = ballot(hit);
      const uint64_t gm = (m >> lead) & S::kGroupMask;
      const uint64_t cnt = full ? (uint64_t)rcnt : (uint64_t)popc64(gm);
      // exclusive prefix of entry counts over the step's groups
      uint64_t v = c_ == 0 ? cnt : 0ull, inc = v;
#pragma unroll
      for (int o = 1; o < kWave; o <<= 1) {
        const uint64_t y = __shfl_up(inc, o);
        if (lane >= o) inc += y;
      }
      const uint64_t pre = __shfl(inc - v, lead);
      const uint64_t tot = __shfl(inc, kWave - 1);
      gmask[k] = gm;
      ecnt[k] = cnt;
      epre[k] = pre;
      stot[k] = tot;
      efirst[k] = first;
      qk[k] = q;
      efull[k] = full;
      wcount += tot;
      if (counts && c_ == 0 && e < n_in && cnt) atomicAdd(counts + q, (unsigned long long)cnt);
    }
    uint64_t b = chain_tile(wcount, tile, ntiles, status, epoch, s_wtot, &s_base, &s_total);
    if (tile == ntiles - 1 && t == 0) c->total = s_total;
    if (WRITE) {
#pragma unroll
      for (int k = 0; k < kQSteps; k++) {
        const uint64_t off = b + epre[k];  // the entry's first output slot
        if (!efull[k]) {
          const uint64_t gm = gmask[k];
          if ((gm >> c_) & 1ull) {
            const uint64_t o = off + (uint64_t)popc64(gm & ((1ull << c_) - 1ull));
            if (o < (uint64_t)out_cap) {
              out_fid[o] = rid[efirst[k] + c_];
              out_tid[o] = (int)qk[k];
            }
          }
        } else if (ecnt[k] <= (uint64_t)(4 * G)) {
          for (uint64_t j = c_; j < ecnt[k]; j += G) {
            if (off + j < (uint64_t)out_cap) {
              out_fid[off + j] = rid[efirst[k] + (int64_t)j];
              out_tid[off + j] = (int)qk[k];
            }
          }
        } else if (c_ == 0) {
          const uint64_t np = (ecnt[k] + kPiece - 1) / kPiece;
          const uint64_t at = atomicAdd(&c->npieces, (unsigned long long)np);
          for (uint64_t p = 0; p < np; p++) {
            if (at + p >= (uint64_t)piece_cap) break;  // only when the total exceeds out_cap
            const uint64_t rest = ecnt[k] - p * kPiece;
            QPiece pc;
            pc.off = off + p * kPiece;
            pc.first = efirst[k] + (int64_t)(p * kPiece);
            pc.len = (uint32_t)(rest < kPiece ? rest : kPiece);
            pc.q = qk[k];
            pieces[at + p] = pc;
          }
        }
        b += stot[k];
      }
    }
  }
}

// Long box ranges of contained subtrees: one block per piece, coalesced id
// reads and pair writes.
__global__ __launch_bounds__(256) void k_write_pieces(const QCtrl* c, const QPiece* __restrict__ pieces,
                                                      int64_t piece_cap, const int* __restrict__ rid,
                                                      int* __restrict__ out_fid,
                                                      int* __restrict__ out_tid, int64_t out_cap) {
  if (c->overflow) return;
  const int64_t np = (int64_t)c->npieces < piece_cap ? (int64_t)c->npieces : piece_cap;
  for (int64_t p = blockIdx.x; p < np; p += gridDim.x) {
    const QPiece pc = pieces[p];
    for (uint32_t j = threadIdx.x; j < pc.len; j += 256) {
      const uint64_t o = pc.off + j;
      if (o < (uint64_t)out_cap) {
        out_fid[o] = rid[pc.first + j];
        out_tid[o] = (int)pc.q;
      }
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
// Host side: engine buffers.
static int dev_alloc(void** p, size_t bytes) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  if (hipMalloc(p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return set_error(CURTREE_ENOMEM, "query: cannot allocate %zu bytes of work memory", bytes);
  }
  return CURTREE_OK;
}

int QEngine::reserve_front(int64_t entries) {
  if (entries <= cap_front) return CURTREE_OK;
  int rc;
  if ((rc = dev_alloc(&front[0], (size_t)entries * 8)) ||
      (rc = dev_alloc(&front[1], (size_t)entries * 8))) {
    cap_front = 0;
    return rc;
  }
  cap_front = entries;
  return CURTREE_OK;
}

int QEngine::reserve_pieces(int64_t n) {
  if (n <= cap_pieces) return CURTREE_OK;
  n = std::max<int64_t>(n + n / 4, 4096);
  int rc = dev_alloc(&pieces, (size_t)n * sizeof(QPiece));
  cap_pieces = rc ? 0 : n;
  return rc;
}

int QEngine::reserve_status(int64_t words) {
  if (words <= status_words) return CURTREE_OK;
  words = std::max<int64_t>(words + words / 2, 1 << 16);
  int rc = dev_alloc((void**)&status, (size_t)words * 8);
  status_words = rc ? 0 : words;
  needs_clear = true;
  return rc;
}

int QEngine::reserve_safe(int64_t nodes) {
  if (nodes <= cap_safe) return CURTREE_OK;
  int rc = dev_alloc((void**)&safe, (size_t)std::max<int64_t>(nodes, 256));
  cap_safe = rc ? 0 : nodes;
  return rc;
}

int QEngine::reserve_plan(int64_t chunks) {
  if (chunks <= cap_plan) return CURTREE_OK;
  int rc = dev_alloc((void**)&plan, (size_t)chunks * 16);
  cap_plan = rc ? 0 : chunks;
  return rc;
}

int QEngine::ensure_ctrl() {
  if (ctrl) return CURTREE_OK;
  int rc = dev_alloc((void**)&ctrl, sizeof(QCtrl));
  if (rc) return rc;
  if (hipHostMalloc((void**)&host_ctrl, sizeof(QCtrl), hipHostMallocDefault) != hipSuccess) {
    host_ctrl = nullptr;
    return set_error(CURTREE_ENOMEM, "query: cannot allocate pinned control block");
  }
  needs_clear = true;
  return CURTREE_OK;
}

void QEngine::release() {
  void* ps[] = {front[0], front[1], pieces, status, ctrl, safe, plan};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  if (host_ctrl) (void)hipHostFree(host_ctrl);
  front[0] = front[1] = pieces = nullptr;
  status = nullptr;
  ctrl = nullptr;
  host_ctrl = nullptr;
  safe = nullptr;
  plan = nullptr;
  cap_front = cap_pieces = status_words = cap_safe = cap_plan = 0;
  needs_clear = true;
}

int next_pow2_group(int F) {
  int g = 2;
  while (g < F) g <<= 1;
  return g;
}

int per_tile(int G) { return (kQThreads / kWave) * kQSteps * (kWave / G); }

int compute_safe(QEngine& qe, Prof& pf, const QBoxes& rb, int F, int64_t r_sz, int64_t nbox,
                 hipStream_t s) {
  int rc;
  if ((rc = qe.reserve_safe(r_sz))) return rc;
  uint8_t* safe = qe.safe;
  LevelPlan lp;
  level_plan(nbox, F, &lp);
  int64_t start = r_sz - lp.n[0];
  pf.mark(s, "k_safe");
  switch (F) {
#define CT_SAFE(g)                                                                        \
  case g:                                                                                 \
    hipLaunchKernelGGL(k_safe_leaf_pow2<g>, dim3((unsigned)((nbox + 255) / 256)), dim3(256), \
                       0, s, nbox, rb, safe + start);                                     \
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
  pf.mark(s, nullptr);
  CT_HIP(hipGetLastError());
  return CURTREE_OK;
}

template <int G>
void enqueue_run_g(QEngine& qe, Prof& pf, const QRunArgs& a, hipStream_t s) {
  const unsigned nslots = (unsigned)a.nl;  // nl-1 expansions + 1 leaf stage
  if (qe.needs_clear) {
    hipLaunchKernelGGL(k_ctrl_init, dim3(1), dim3(64), 0, s, qe.ctrl);
    qe.needs_clear = false;
  }
  pf.mark(s, "k_run_begin");
  hipLaunchKernelGGL(k_run_begin, dim3(1), dim3(64), 0, s, qe.ctrl, a.q0, a.q1, a.plan, nslots);
  hipLaunchKernelGGL(k_status_zero, dim3(1024), dim3(256), 0, s, (const QCtrl*)qe.ctrl, qe.status,
                     qe.status_words);
  const uint2* cur = nullptr;
  int which = 0;
  for (int l = 0; l + 1 < a.nl; l++) {
    uint2* out = (uint2*)qe.front[which];
    pf.mark(s, "k_expand");
    hipLaunchKernelGGL(k_expand<G>, dim3(kChainGrid), dim3(kQThreads), 0, s, l, a.tr, a.qb,
                       qe.ctrl, cur, out, qe.cap_front, qe.status, a.r_sz);
    cur = out;
    which ^= 1;
  }
  const int leaf = a.nl - 1;
  if (a.write) {
    pf.mark(s, "k_leaf_pass");
    hipLaunchKernelGGL((k_leaf_pass<G, true>), dim3(kChainGrid), dim3(kQThreads), 0, s, leaf, a.tr,
                       a.qb, a.rb, a.rid, a.lv, qe.ctrl, cur, a.r_sz, a.nbox, a.out_fid,
                       a.out_tid, a.out_cap, (QPiece*)qe.pieces, qe.cap_pieces, a.counts,
                       qe.status);
    pf.mark(s, "k_write_pieces");
    hipLaunchKernelGGL(k_write_pieces, dim3(4096), dim3(256), 0, s, (const QCtrl*)qe.ctrl,
                       (const QPiece*)qe.pieces, qe.cap_pieces, a.rid, a.out_fid, a.out_tid,
                       a.out_cap);
  } else {
    pf.mark(s, "k_leaf_count");
    hipLaunchKernelGGL((k_leaf_pass<G, false>), dim3(kChainGrid), dim3(kQThreads), 0, s, leaf, a.tr,
                       a.qb, a.rb, a.rid, a.lv, qe.ctrl, cur, a.r_sz, a.nbox, (int*)nullptr,
                       (int*)nullptr, (int64_t)0, (QPiece*)nullptr, (int64_t)0, a.counts,
                       qe.status);
  }
  pf.mark(s, nullptr);
}

void enqueue_run(QEngine& qe, Prof& pf, const QRunArgs& a, hipStream_t s) {
  switch (a.G) {
    case 2: enqueue_run_g<2>(qe, pf, a, s); break;
    case 4: enqueue_run_g<4>(qe, pf, a, s); break;
    case 8: enqueue_run_g<8>(qe, pf, a, s); break;
    case 16: enqueue_run_g<16>(qe, pf, a, s); break;
    case 32: enqueue_run_g<32>(qe, pf, a, s); break;
    default: enqueue_run_g<64>(qe, pf, a, s); break;
  }
}

int64_t pieces_needed(const QEngine& qe, const QRunArgs& a) {
  // every long range is at most one leaf entry and cut into kPiece slots
  return qe.cap_front + a.out_cap / (int64_t)kPiece + 1;
}

void enqueue_ctrl_init(QEngine& qe, hipStream_t s) {
  hipLaunchKernelGGL(k_ctrl_init, dim3(1), dim3(64), 0, s, qe.ctrl);
  qe.needs_clear = false;
}

void enqueue_fill_counts(int64_t* counts, int64_t n, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_fill_u64, dim3(256), dim3(256), 0, s, (uint64_t*)counts, n, 0ull);
}

int read_ctrl(QEngine& qe, hipStream_t s) {
  CT_HIP(hipMemcpyAsync(qe.host_ctrl, qe.ctrl, sizeof(QCtrl), hipMemcpyDeviceToHost, s));
  CT_HIP(hipStreamSynchronize(s));
  return CURTREE_OK;
}

int prepare_query(QEngine& qe, Prof& pf, const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox,
                  const curtree_bbox_f32* qbox, int contain, int64_t nq_hint, QRunArgs* a,
                  hipStream_t s) {
  if (!rt || !rbox || !qbox) return set_error(CURTREE_EINVAL, "query: null argument");
  const int64_t N = rbox->sz, Q = qbox->sz;
  const int F = rt->fanout;
  if (N < 0 || Q < 0) return set_error(CURTREE_EINVAL, "query: negative size");
  if (F < 2) return set_error(CURTREE_EINVAL, "query: fanout must be >= 2");
  if (F > 64)
    return set_error(CURTREE_EINVAL, "query: fanout %d > 64 is not supported by the GPU query", F);
  if (Q > (int64_t)0xFFFFFFFF) return set_error(CURTREE_EINVAL, "query: too many windows");
  *a = QRunArgs();
  a->nbox = N;
  a->F = F;
  a->G = next_pow2_group(F);
  if (N == 0) return CURTREE_OK;
  LevelPlan lp;
  level_plan(N, F, &lp);
  if (lp.r_sz != rt->sz)
    return set_error(CURTREE_EINVAL,
                     "query: tree size %d does not match %lld boxes at fanout %d (expected %lld)",
                     rt->sz, (long long)N, F, (long long)lp.r_sz);
  if (lp.nl + 1 > kMaxRunSlots) return set_error(CURTREE_EINVAL, "query: tree too deep");
  if (!rt->xmin || !rt->ymin || !rt->xmax || !rt->ymax || !rt->pos || !rt->len || !rbox->xmin ||
      !rbox->ymin || !rbox->xmax || !rbox->ymax || !rbox->id ||
      (Q > 0 && (!qbox->xmin || !qbox->ymin || !qbox->xmax || !qbox->ymax)))
    return set_error(CURTREE_EINVAL, "query: null device array");
  int rc;
  if ((rc = qe.ensure_ctrl())) return rc;
  a->tr = QTree{rt->xmin, rt->ymin, rt->xmax, rt->ymax, rt->pos, rt->len, nullptr};
  a->qb = QBoxes{qbox->xmin, qbox->ymin, qbox->xmax, qbox->ymax};
  a->rb = QBoxes{rbox->xmin, rbox->ymin, rbox->xmax, rbox->ymax};
  a->rid = rbox->id;
  a->r_sz = lp.r_sz;
  a->nl = lp.nl;
  a->lv.nl = lp.nl;
  {
    int64_t start = lp.r_sz, span = 1;
    for (int l = 0; l < lp.nl; l++) {
      start -= lp.n[l];
      span = span > ((int64_t)1 << 40) / F ? ((int64_t)1 << 40) : span * F;
      a->lv.lstart[l] = start;
      a->lv.lcount[l] = lp.n[l];
      a->lv.span[l] = span;
    }
  }
  // the flags cost one pass over the boxes; worth it for larger batches
  const bool use = contain == 1 || (contain < 0 && nq_hint >= 256);
  if (use) {
    if ((rc = compute_safe(qe, pf, a->rb, F, lp.r_sz, N, s))) return rc;
    a->tr.safe = qe.safe;
  }
  return CURTREE_OK;
}

int query_run_f32(QEngine& qe, Prof& pf, const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox,
                  const curtree_bbox_f32* qbox, int64_t q0, int64_t q1, int contain, QueryOut* qo,
                  hipStream_t s) {
  if (!qo || !qbox) return set_error(CURTREE_EINVAL, "query: null argument");
  qo->n = 0;
  if (q0 < 0 || q1 > qbox->sz || q0 > q1) return set_error(CURTREE_EINVAL, "query: bad range");
  QRunArgs a;
  int rc;
  if ((rc = prepare_query(qe, pf, rt, rbox, qbox, contain, q1 - q0, &a, s))) return rc;
  const int64_t nq = q1 - q0;
  if (a.nbox == 0 || nq == 0) {
    if (qo->counts) enqueue_fill_counts(qo->counts + q0, nq, s);
    return CURTREE_OK;
  }
  a.q0 = q0;
  a.q1 = q1;
  a.counts = (unsigned long long*)qo->counts;
  // caller buffer: one writing run; otherwise count first
  const bool direct = !qo->alloc && !qo->count_only && !qo->counts && qo->fid;
  a.write = direct;
  a.out_fid = qo->fid;
  a.out_tid = qo->tid;
  a.out_cap = direct ? qo->cap : 0;
  if ((rc = qe.reserve_front(std::max<int64_t>(2 * nq, 1 << 20)))) return rc;
  for (int pass = 0; pass < 2; pass++) {
    for (int attempt = 0;; attempt++) {
      if (a.counts) enqueue_fill_counts(qo->counts + q0, nq, s);
      if ((rc = qe.reserve_status(std::max<int64_t>(qe.cap_front, nq) / per_tile(a.G) + 2)))
        return rc;
      if (a.write && (rc = qe.reserve_pieces(pieces_needed(qe, a)))) return rc;
      enqueue_run(qe, pf, a, s);
      CT_HIP(hipGetLastError());
      if ((rc = read_ctrl(qe, s))) return rc;
      if (!qe.host_ctrl->overflow) break;
      if (attempt >= 3) return set_error(CURTREE_ENOMEM, "query: frontier kept overflowing");
      // grow to the largest frontier the run reported and run again
      uint64_t need = 0;
      for (int l = 1; l < a.nl; l++) need = std::max<uint64_t>(need, qe.host_ctrl->count[l]);
      if (need > 0xFFFFFFFFull)
        return set_error(CURTREE_ERANGE, "query: frontier of %llu entries exceeds 2^32",
                         (unsigned long long)need);
      if ((rc = qe.reserve_front((int64_t)(need + need / 4 + 1024)))) return rc;
    }
    const uint64_t H = qe.host_ctrl->total;
    qo->n = (int64_t)H;
    if (a.write) {
      qo->out_fid = a.out_fid;
      qo->out_tid = a.out_tid;
      if ((int64_t)H > a.out_cap)
        return set_error(CURTREE_ERANGE, "query: %llu pairs exceed the buffer of %lld",
                         (unsigned long long)H, (long long)a.out_cap);
      return CURTREE_OK;
    }
    if (qo->count_only || H == 0) return CURTREE_OK;
    // second pass: write into an exact allocation (or the caller's buffer)
    if (qo->alloc) {
      int* fid = nullptr;
      int* tid = nullptr;
      CT_HIP(hipMalloc((void**)&fid, H * sizeof(int)));
      if (hipMalloc((void**)&tid, H * sizeof(int)) != hipSuccess) {
        (void)hipFree(fid);
        return set_error(CURTREE_ENOMEM, "query: cannot allocate %llu result pairs",
                         (unsigned long long)H);
      }
      a.out_fid = fid;
      a.out_tid = tid;
      a.out_cap = (int64_t)H;
    } else if ((int64_t)H > qo->cap) {
      return set_error(CURTREE_ERANGE, "query: %llu pairs exceed the buffer of %lld",
                       (unsigned long long)H, (long long)qo->cap);
    } else {
      a.out_cap = qo->cap;
    }
    a.write = true;
    a.counts = nullptr;
  }
  return CURTREE_OK;
}

int query_frontier_f32(Context& ctx, const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox,
                       const curtree_bbox_f32* qbox, QueryOut* qo, hipStream_t s) {
  if (!qbox) return set_error(CURTREE_EINVAL, "query: null argument");
  return query_run_f32(ctx.qe, ctx.prof, rt, rbox, qbox, 0, qbox->sz, ctx.opt_contain, qo, s);
}

}  // namespace ct
