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
                            uint32_t nslots, QLevels lv) {
  if (threadIdx.x == 0) {
    c->lv = lv;
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
    uint64_t tt = 0;
#pragma unroll
    for (int i = 0; i < kQWaves; i++) tt += s_wtot[i];
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

// Take the next tile; returns -1 when the level's tiles are exhausted.  The
// ticket is taken when the block is ready to work on it: a tile taken early
// would hold up the look-back of every later tile until its block gets to it.
__device__ __forceinline__ int64_t next_tile(uint32_t* ticket, int64_t ntiles, uint32_t* s_tile) {
  __syncthreads();  // the previous tile is done with the shared words
  if (threadIdx.x == 0) *s_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  const int64_t t = *s_tile;
  return t < ntiles ? t : -1;
}

// Internal level `slot`: entries (query, node) whose children are nodes.
// fin == null: the implicit first frontier (query q0 + e, root).
template <int G, int STEPS>
__global__ __launch_bounds__(kQThreads) void k_expand(int slot, QTree tr, QBoxes qb, QCtrl* c,
                                                      const uint2* __restrict__ fin,
                                                      uint2* __restrict__ fout, int64_t cap,
                                                      uint64_t* status, int64_t r_sz) {
  using S = QShape<G, STEPS>;
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_wtot[kQWaves];
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
  // bulk-load layout: the entries of this launch sit on level nl-1-slot and
  // node i's children are nodes cs + i*F .. of the level below
  const bool canon = tr.canon && *tr.canon;
  const bool shortcut = shortcut_on(tr, canon);
  const int lvl = c->lv.nl - 1 - slot;
  const int64_t ls = c->lv.lstart[lvl];
  const int64_t cs = lvl > 0 ? c->lv.lstart[lvl - 1] : 0;
  const int64_t cc = lvl > 0 ? c->lv.lcount[lvl - 1] : 0;
  int64_t tile;
  while ((tile = next_tile(&c->ticket[slot], ntiles, &s_tile)) >= 0) {
    uint64_t masks[STEPS];
    uint32_t qk[STEPS], ck[STEPS];
    uint64_t wcount = 0;
    const int64_t ebase = tile * S::kPerTile + (int64_t)w * S::kPerWave + sub;
#pragma unroll
    for (int k = 0; k < STEPS; k++) {
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
          int64_t p, L;
          if (canon) {
            const int64_t f = ((int64_t)node - ls) * tr.F;
            p = cs + f;
            L = cc - f < tr.F ? cc - f : tr.F;
          } else {
            p = tr.pos[node];
            L = tr.len[node];
          }
          if (p >= 0 && c_ < L && p + L <= r_sz) {  // children must be nodes
            const uint32_t child = (uint32_t)(p + c_);
            const float4 wq = ld_box(qb, q);
            const float4 mb = ld_box(tr.mbr, child);
            hit = intersect(wq.x, wq.y, wq.z, wq.w, mb.x, mb.y, mb.z, mb.w);
            ci = child;
            // (a contained subtree's box range comes from the level table,
            // so the shortcut needs the bulk-load layout)
            if (hit && shortcut &&
                contains(wq.x, wq.y, wq.z, wq.w, mb.x, mb.y, mb.z, mb.w) && node_safe(tr, child))
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
    for (int k = 0; k < STEPS; k++) {
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
// Levels are stored root level first, so lstart falls with l and v's level is
// the first l with v >= lstart[l]; contained subtrees sit low in the tree, so
// the loop usually ends after one or two steps.
__device__ __forceinline__ void node_range(uint32_t v, const QLevels* L, int64_t nbox,
                                           int64_t* first, int64_t* cnt) {
  const int nl = L->nl;
  int lv = nl - 1;
  for (int k = 0; k < nl; k++) {
    if ((int64_t)v >= L->lstart[k]) {
      lv = k;
      break;
    }
  }
  const int64_t span = L->span[lv];
  const int64_t f = ((int64_t)v - L->lstart[lv]) * span;
  const int64_t rem = nbox - f;
  *first = f;
  *cnt = rem < span ? rem : span;
}

// Leaf stage.  Each entry is either a partial leaf-parent (its boxes are
// tested: G lanes, one box each) or a contained subtree (all of its box range
// hits, no reads).  Three launches, none of which waits on another tile:
//   k_leaf_count  tests the boxes; per entry it stores the hit mask (or range
//                 length), first box and query (16 B), per tile the pair
//                 count; totals and per-query counts by atomics.  Count-only
//                 runs stop here and store nothing.
//   k_tile_scan   exclusive scan of the tile counts (a short chain: one link
//                 per kQThreads * 8 tiles).
//   k_leaf_write  per tile: loads the entries, scans them locally, lists the
//                 entries with pairs to write, adds the tile's base and
//                 writes the tile's pairs as one flat, coalesced sequence
//                 (each thread walks a run of consecutive slots through the
//                 list and stages their box ids in LDS, the block stores them
//                 linearly); ranges longer than kFlatMax are cut into pieces
//                 for k_write_pieces.  Slots beyond out_cap are dropped.
// Output order is the entries' order: (query, node), then box position.
constexpr uint32_t kQRange = 0x80000000u;
constexpr uint32_t kQNone = 0xFFFFFFFFu;
constexpr uint64_t kFlatMax = 2048;  // longer ranges go to the piece writer
constexpr int kScanPer = 8;          // tile counts per thread in k_tile_scan

// Position of the r-th (0-based) set bit of m (r < popcount(m)); m has no
// bit at W or above.
template <int W = 64>
__device__ __forceinline__ uint32_t nth_set_bit(uint64_t m, uint32_t r) {
  uint32_t pos = 0;
#pragma unroll
  for (int w = W / 2; w > 0; w >>= 1) {
    const uint64_t low = m & ((1ull << w) - 1ull);
    const uint32_t c = (uint32_t)__popcll(low);
    if (r >= c) {
      r -= c;
      m >>= w;
      pos += w;
    } else {
      m = low;
    }
  }
  return pos;
}

// Exclusive scan of one u64 per thread over the block.
__device__ __forceinline__ uint64_t block_excl_scan64(uint64_t v, uint64_t* s_tmp,
                                                      uint64_t* total) {
  const int lane = lane_id(), w = threadIdx.x >> 6;
  uint64_t x = v;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint64_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == kWave - 1) s_tmp[w] = x;
  __syncthreads();
  uint64_t off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kQWaves; i++) {
    const uint64_t t = s_tmp[i];
    if (i < w) off += t;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return x - v + off;
}

// The entries of one lane group for STEPS wave-steps (entry e0 + k*E at step
// k; past the frontier: its last entry, unused; fin == null: the implicit
// first frontier, window q0 + e at the root).  Groups of 16 lanes or more:
// the first lane of each row of 16 loads and DPP broadcasts to the row (every
// row of a group loads the same entry), so a wave issues a quarter-row of
// loads instead of 64 identical ones per group.
template <int G, int STEPS, int PERWAVE, int E>
__device__ __forceinline__ void load_entries(uint2 (&ent)[STEPS], const uint2* __restrict__ fin,
                                             int64_t e0, int64_t n_in, int64_t q0) {
  if (!fin) {
#pragma unroll
    for (int k = 0; k < STEPS; k++) {
      const int64_t e = e0 + k * E;
      ent[k] = make_uint2((uint32_t)(q0 + (e < n_in ? e : n_in - 1)), 0u);
    }
    return;
  }
  if constexpr (G >= 16) {
    uint2 raw[STEPS];
#pragma unroll
    for (int k = 0; k < STEPS; k++) raw[k] = make_uint2(0u, 0u);
    if ((lane_id() & 15) == 0) {
#pragma unroll
      for (int k = 0; k < STEPS; k++) {
        const int64_t e = e0 + k * E;
        raw[k] = fin[e < n_in ? e : n_in - 1];
      }
    }
#pragma unroll
    for (int k = 0; k < STEPS; k++) ent[k] = make_uint2(row_first(raw[k].x), row_first(raw[k].y));
  } else {
#pragma unroll
    for (int k = 0; k < STEPS; k++) {
      const int64_t e = e0 + k * E;
      ent[k] = fin[e < n_in ? e : n_in - 1];
    }
  }
}

// The windows of H entries (queries q[k], the same in every lane of a group):
// load_windows issues the loads (groups of 16 lanes or more: the first lane of
// each row only), bcast_windows completes them (the row broadcast), so other
// loads can be issued between the two.
template <int G, int H>
__device__ __forceinline__ void load_windows(float4 (&wq)[H], const QBoxes& qb,
                                             const uint32_t (&q)[H]) {
  if constexpr (G >= 16) {
#pragma unroll
    for (int k = 0; k < H; k++) wq[k] = make_float4(0.f, 0.f, 0.f, 0.f);
    if ((lane_id() & 15) == 0) {
#pragma unroll
      for (int k = 0; k < H; k++) wq[k] = ld_box(qb, q[k]);
    }
  } else {
#pragma unroll
    for (int k = 0; k < H; k++) wq[k] = ld_box(qb, q[k]);
  }
}
template <int G, int H>
__device__ __forceinline__ void bcast_windows(float4 (&wq)[H]) {
  if constexpr (G >= 16) {
#pragma unroll
    for (int k = 0; k < H; k++)
      wq[k] = make_float4(row_first(wq[k].x), row_first(wq[k].y), row_first(wq[k].z),
                          row_first(wq[k].w));
  }
}

#ifndef CT_LEAF_H
#define CT_LEAF_H 2
#endif
// QCOUNTS: per-query counts into `counts` as well
template <int G, int STEPS, bool STORE, bool QCOUNTS>
__global__ __launch_bounds__(kQThreads) void k_leaf_count(
    int slot, QTree tr, QBoxes qb, QBoxes rb, QCtrl* c, const uint2* __restrict__ fin,
    int64_t r_sz, int64_t nbox, QLeaf* __restrict__ est, uint64_t* __restrict__ ttot,
    unsigned long long* counts, bool xcd) {
  using S = QShape<G, STEPS>;
  constexpr int P = S::kPerTile;
  // memory round trips per tile: the entries of all steps, then the windows
  // and boxes of H steps at a time
  constexpr int H = STEPS < CT_LEAF_H ? STEPS : CT_LEAF_H;
  static_assert(STEPS % H == 0, "steps in whole groups");
  __shared__ uint64_t s_red[kQWaves];
  __shared__ QLevels s_lv;  // the level table (contained ranges)
  // the tile's entries, written out once at the end of the tile: a store
  // between two rounds of loads would make the second wait for it (stores
  // and loads share one counter)
  __shared__ QLeaf s_est[STORE ? P : 1];
  // per-query counts: the query and pair count of every entry of the tile,
  // summed per run of one query (the frontier keeps a query's entries
  // together) and added with one atomic at each end of the run
  __shared__ uint32_t s_cq[QCOUNTS ? P : 1];
  __shared__ uint32_t s_cn[QCOUNTS ? P : 1];
  __shared__ uint64_t s_scan[kQWaves];
  if (c->overflow) return;
  const int64_t n_in = (int64_t)c->count[slot];
  const int64_t q0 = c->q0;
  const int64_t ntiles = (n_in + P - 1) / P;
  const int t = threadIdx.x, w = t >> 6, lane = lane_id();
  if (ntiles == 0) return;  // the run's total stays 0
  for (int i = t; i < (int)(sizeof(QLevels) / 4); i += kQThreads)
    ((uint32_t*)&s_lv)[i] = ((const uint32_t*)&c->lv)[i];
  const int sub = lane / G, c_ = lane % G;
  const int lead = sub * G;  // first lane of this lane's group
  // bulk-load layout: leaf-parent i holds sorted boxes i*F .. i*F+F-1
  const bool canon = tr.canon && *tr.canon;
  const int64_t ls = c->lv.lstart[0];
  uint64_t block_total = 0;  // thread 0: one atomic per block, not per tile
  // tiles are independent (no chain): a fixed stride over them, no ticket
  const TileWalk tw = tile_walk(ntiles, xcd);
  for (int64_t tile = tw.first; tile < tw.end; tile += tw.step) {
    __syncthreads();  // the previous tile is done with the shared words (and s_lv is in)
    uint64_t wsum = 0;
    const int64_t tbase = tile * P;
    // the tile's entries (a position past the frontier reads its last entry,
    // unused)
    uint2 ent[STEPS];
    if (fin) {
#pragma unroll
      for (int k = 0; k < STEPS; k++) {
        const int64_t e = tbase + w * S::kPerWave + k * S::E + sub;
        ent[k] = fin[e < n_in ? e : n_in - 1];
      }
    } else {
#pragma unroll
      for (int k = 0; k < STEPS; k++) {
        const int64_t e = tbase + w * S::kPerWave + k * S::E + sub;
        ent[k] = make_uint2((uint32_t)(q0 + (e < n_in ? e : n_in - 1)), 0u);
      }
    }
#pragma unroll
    for (int k0 = 0; k0 < STEPS; k0 += H) {
      // the box range of each entry
      // (box positions and counts fit 32 bits: a tree holds < 2^30 boxes)
      uint32_t q[H];
      int first[H], cnt[H];  // contained: range length; partial: children
      bool full[H];
      const uint2* en = ent + k0;
#pragma unroll
      for (int k = 0; k < H; k++) {
        q[k] = en[k].x;
        const uint32_t node = en[k].y & ~kFull;
        full[k] = (en[k].y & kFull) != 0;
        if (full[k]) {
          int64_t f, n;
          node_range(node, &s_lv, nbox, &f, &n);
          first[k] = (int)f;
          cnt[k] = (int)n;
        } else {
          const int64_t f = ((int64_t)node - ls) * tr.F;
          first[k] = (int)f;
          cnt[k] = (int)(nbox - f < tr.F ? nbox - f : tr.F);
        }
      }
      if (!canon) {  // any other layout: the node's pos/len
        int pos[H], len[H];
#pragma unroll
        for (int k = 0; k < H; k++) {
          const uint32_t node = en[k].y & ~kFull;
          pos[k] = tr.pos[node];
          len[k] = tr.len[node];
        }
#pragma unroll
        for (int k = 0; k < H; k++)
          if (!full[k]) {
            first[k] = (int)((int64_t)pos[k] - r_sz);
            cnt[k] = len[k];
          }
      }
      // windows and boxes (a lane with no child reads box 0, unused)
      float4 wq[H], bx[H];
      bool test[H];
#pragma unroll
      for (int k = 0; k < H; k++) {
        const int bi = first[k] + c_;
        test[k] = !full[k] && c_ < cnt[k] && bi >= 0 && bi < nbox;  // children must be boxes
        wq[k] = ld_box(qb, q[k]);
        bx[k] = ld_box(rb, test[k] ? bi : 0);
      }
#pragma unroll
      for (int k = 0; k < H; k++) {
        const int64_t e = tbase + w * S::kPerWave + (k0 + k) * S::E + sub;
        const bool valid = e < n_in;
        const bool hit = valid && test[k] &&
                         intersect(wq[k].x, wq[k].y, wq[k].z, wq[k].w, bx[k].x, bx[k].y, bx[k].z,
                                   bx[k].w);
        const uint64_t m = ballot(hit);
        const uint64_t gm = (m >> lead) & S::kGroupMask;
        const uint64_t nh = full[k] ? (uint64_t)cnt[k] : (uint64_t)popc64(gm);
        if (c_ == 0 && valid) {
          if (STORE) {
            QLeaf x;
            x.val = full[k] ? nh : gm;
            x.first = (uint32_t)first[k];
            x.qf = q[k] | (full[k] ? kQRange : 0u);
            s_est[e - tbase] = x;
          }
          if (QCOUNTS) {
            s_cq[e - tbase] = q[k];
            s_cn[e - tbase] = (uint32_t)nh;  // (at most F, or a range of < 2^30 boxes)
          }
        }
        if (c_ == 0 && valid) wsum += nh;  // (this lane's share; summed over the wave below)
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) wsum += __shfl_xor(wsum, o);
    if (lane == 0) s_red[w] = wsum;
    __syncthreads();
    if (STORE)
      for (int i = t; i < P; i += kQThreads)
        if (tbase + i < n_in) est[tbase + i] = s_est[i];
    if (t == 0) {
      uint64_t tot = 0;
#pragma unroll
      for (int i = 0; i < kQWaves; i++) tot += s_red[i];
      if (STORE) ttot[tile] = tot;
      block_total += tot;
    }
    if (QCOUNTS) {
      // thread t holds entries t*R .. t*R+R-1: their sum, a block scan, then
      // count[q] -= (pairs before the run) at a run's first entry and
      // += (pairs up to its end) at its last (unsigned, so the two cancel)
      constexpr int R = (P + kQThreads - 1) / kQThreads;
      const int nv = n_in - tbase < P ? (int)(n_in - tbase) : P;
      uint64_t sum = 0;
#pragma unroll
      for (int r = 0; r < R; r++)
        if (t * R + r < nv) sum += s_cn[t * R + r];
      uint64_t tile_sum;
      uint64_t before = block_excl_scan64(sum, s_scan, &tile_sum);
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int i = t * R + r;
        if (i >= nv) break;
        const uint32_t qi = s_cq[i];
        if (i == 0 || s_cq[i - 1] != qi) atomicAdd(counts + qi, 0ull - before);
        before += s_cn[i];
        if (i + 1 == nv || s_cq[i + 1] != qi) atomicAdd(counts + qi, before);
      }
    }
  }
  if (threadIdx.x == 0 && block_total)
    atomicAdd((unsigned long long*)&c->total, (unsigned long long)block_total);
}

// The leaf count over leaf codes (tr.lframe / tr.lcode, G >= 4): the same
// tiles of P entries and the same outputs as k_leaf_count, but G/4 lanes per
// entry, lane g holding the cells of child slots g, g + G/4, g + 2G/4,
// g + 3G/4 (one 16-byte load, see cell_slot), so one wave-step covers four
// times the entries and a whole tile is one round of loads (entries, then
// windows, frames and cells); an entry's hit mask is then four shifted
// nibbles of four ballots.  A
// test the cells leave open is settled at the end of the tile, in one more
// round: the box and the window read, the float test.  A layout other than
// the bulk-load one has no codes: every test is open (pos/len give the
// children).
#ifndef CT_CODED_WAVES
#define CT_CODED_WAVES 6  // waves per SIMD asked of the compiler (0: its choice)
#endif
template <int G, int STEPS, bool STORE, bool QCOUNTS>
__global__ __launch_bounds__(kQThreads)
#if CT_CODED_WAVES > 0
__attribute__((amdgpu_waves_per_eu(CT_CODED_WAVES)))
#endif
void k_leaf_count_coded(
    int slot, QTree tr, QBoxes qb, QBoxes rb, QCtrl* c, const uint2* __restrict__ fin,
    int64_t r_sz, int64_t nbox, QLeaf* __restrict__ est, uint64_t* __restrict__ ttot,
    unsigned long long* counts, bool xcd) {
  using S = QShape<G, STEPS>;
  constexpr int P = S::kPerTile;
  constexpr int GC = G / 4;           // lanes per entry
  constexpr int EC = kWave / GC;      // entries per wave-step
  constexpr int SC = S::kPerWave / EC;  // wave-steps per tile
  static_assert(G >= 4 && SC * EC == S::kPerWave, "whole wave-steps");
  constexpr uint32_t kNib = GC >= 32 ? 0xFFFFFFFFu : ((1u << GC) - 1u);
  __shared__ uint64_t s_red[kQWaves];
  __shared__ QLevels s_lv;
  __shared__ QLeaf s_est[STORE ? P : 1];
  __shared__ uint32_t s_cq[QCOUNTS ? P : 1];
  __shared__ uint32_t s_cn[QCOUNTS ? P : 1];
  __shared__ uint64_t s_scan[kQWaves];
  if (c->overflow) return;
  const int64_t n_in = (int64_t)c->count[slot];
  const int64_t q0 = c->q0;
  const int64_t ntiles = (n_in + P - 1) / P;
  const int t = threadIdx.x, w = t >> 6, lane = lane_id();
  if (ntiles == 0) return;
  for (int i = t; i < (int)(sizeof(QLevels) / 4); i += kQThreads)
    ((uint32_t*)&s_lv)[i] = ((const uint32_t*)&c->lv)[i];
  const int sub = lane / GC, cg = lane % GC;  // entry of the wave-step, slot group
  const int lead = sub * GC;
  const bool canon = tr.canon && *tr.canon;
  const int64_t ls = c->lv.lstart[0];
  uint64_t block_total = 0;
  const TileWalk tw = tile_walk(ntiles, xcd);
  for (int64_t tile = tw.first; tile < tw.end; tile += tw.step) {
    __syncthreads();
    uint64_t wsum = 0;
    const int64_t tbase = tile * P;
    uint2 ent[SC];
#pragma unroll
    for (int k = 0; k < SC; k++) {
      const int64_t e = tbase + w * S::kPerWave + k * EC + sub;
      ent[k] = fin ? fin[e < n_in ? e : n_in - 1]
                   : make_uint2((uint32_t)(q0 + (e < n_in ? e : n_in - 1)), 0u);
    }
    uint32_t q[SC];
    int first[SC], cnt[SC], lpn[SC];
    bool full[SC];
#pragma unroll
    for (int k = 0; k < SC; k++) {
      q[k] = ent[k].x;
      const uint32_t node = ent[k].y & ~kFull;
      full[k] = (ent[k].y & kFull) != 0;
      lpn[k] = -1;
      if (full[k]) {
        int64_t f, n;
        node_range(node, &s_lv, nbox, &f, &n);
        first[k] = (int)f;
        cnt[k] = (int)n;
      } else if (canon) {
        const int64_t f = ((int64_t)node - ls) * tr.F;
        first[k] = (int)f;
        cnt[k] = (int)(nbox - f < tr.F ? nbox - f : tr.F);
        lpn[k] = (int)((int64_t)node - ls);
      } else {
        first[k] = (int)((int64_t)tr.pos[node] - r_sz);
        cnt[k] = tr.len[node];
      }
    }
    // windows, frames and the cells of this lane's four slots
    float4 wq[SC], fr[SC];
    uint4 cd[SC];
    uint32_t test[SC];  // bit j: slot j * GC + cg is a child box to test
#pragma unroll
    for (int k = 0; k < SC; k++) {
      test[k] = 0;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int sl = j * GC + cg, bi = first[k] + sl;
        if (!full[k] && sl < cnt[k] && bi >= 0 && bi < nbox) test[k] |= 1u << j;
      }
      wq[k] = ld_box(qb, q[k]);
      const bool coded = lpn[k] >= 0 && test[k] != 0;
      fr[k] = coded ? tr.lframe[lpn[k]] : make_float4(0.f, 0.f, 0.f, 0.f);
      cd[k] = coded ? ((const uint4*)tr.lcode)[(int64_t)lpn[k] * GC + cg] : make_uint4(0, 0, 0, 0);
    }
    // decisions from the cells; open tests settled by the boxes
    uint32_t hit[SC], open[SC];
#pragma unroll
    for (int k = 0; k < SC; k++) {
      const int64_t e = tbase + w * S::kPerWave + k * EC + sub;
      hit[k] = open[k] = 0;
      if (e < n_in && test[k]) {
        if (fr[k].z > 0.f) {
          const uint32_t wc = box_cells(wq[k].x, wq[k].y, wq[k].z, wq[k].w, fr[k]);
          const ct_s16x2 wl = cells_low(wc), wh = cells_high(wc);
          const uint32_t b4[4] = {cd[k].x, cd[k].y, cd[k].z, cd[k].w};
#pragma unroll
          for (int j = 0; j < 4; j++) {
            if (!((test[k] >> j) & 1u)) continue;
            const int r = cells_intersect(wl, wh, b4[j]);
            if (r > 0) hit[k] |= 1u << j;
            else if (r < 0) open[k] |= 1u << j;
          }
        } else {
          open[k] = test[k];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < SC; k++) {
      if (open[k]) {
        float4 bx[4];
#pragma unroll
        for (int j = 0; j < 4; j++)
          if ((open[k] >> j) & 1u) bx[j] = ld_box(rb, first[k] + j * GC + cg);
#pragma unroll
        for (int j = 0; j < 4; j++)
          if (((open[k] >> j) & 1u) &&
              intersect(wq[k].x, wq[k].y, wq[k].z, wq[k].w, bx[j].x, bx[j].y, bx[j].z, bx[j].w))
            hit[k] |= 1u << j;
      }
    }
    // per entry: the hit mask over its slots (slot j * GC + cg from lane cg's bit j)
#pragma unroll
    for (int k = 0; k < SC; k++) {
      const int64_t e = tbase + w * S::kPerWave + k * EC + sub;
      const bool valid = e < n_in;
      uint64_t gm = 0;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const uint64_t m = ballot((hit[k] >> j) & 1u);
        gm |= (uint64_t)((uint32_t)(m >> lead) & kNib) << (j * GC);
      }
      const uint64_t nh = full[k] ? (uint64_t)cnt[k] : (uint64_t)popc64(gm);
      if (cg == 0 && valid) {
        if (STORE) {
          QLeaf x;
          x.val = full[k] ? nh : gm;
          x.first = (uint32_t)first[k];
          x.qf = q[k] | (full[k] ? kQRange : 0u);
          s_est[e - tbase] = x;
        }
        if (QCOUNTS) {
          s_cq[e - tbase] = q[k];
          s_cn[e - tbase] = (uint32_t)nh;
        }
        wsum += nh;
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) wsum += __shfl_xor(wsum, o);
    if (lane == 0) s_red[w] = wsum;
    __syncthreads();
    if (STORE)
      for (int i = t; i < P; i += kQThreads)
        if (tbase + i < n_in) est[tbase + i] = s_est[i];
    if (t == 0) {
      uint64_t tot = 0;
#pragma unroll
      for (int i = 0; i < kQWaves; i++) tot += s_red[i];
      if (STORE) ttot[tile] = tot;
      block_total += tot;
    }
    if (QCOUNTS) {
      constexpr int R = (P + kQThreads - 1) / kQThreads;
      const int nv = n_in - tbase < P ? (int)(n_in - tbase) : P;
      uint64_t sum = 0;
#pragma unroll
      for (int r = 0; r < R; r++)
        if (t * R + r < nv) sum += s_cn[t * R + r];
      uint64_t tile_sum;
      uint64_t before = block_excl_scan64(sum, s_scan, &tile_sum);
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int i = t * R + r;
        if (i >= nv) break;
        const uint32_t qi = s_cq[i];
        if (i == 0 || s_cq[i - 1] != qi) atomicAdd(counts + qi, 0ull - before);
        before += s_cn[i];
        if (i + 1 == nv || s_cq[i + 1] != qi) atomicAdd(counts + qi, before);
      }
    }
  }
  if (threadIdx.x == 0 && block_total)
    atomicAdd((unsigned long long*)&c->total, (unsigned long long)block_total);
}


// Exclusive scan in place of the tile counts of frontier `cslot` (tiles of P
// entries); `slot` names the launch's ticket and epoch.
__global__ __launch_bounds__(kQThreads) void k_tile_scan(int slot, int cslot, int P, QCtrl* c,
                                                         uint64_t* __restrict__ ttot,
                                                         uint64_t* status) {
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_wtot[kQWaves];
  __shared__ uint64_t s_base, s_total;
  if (c->overflow) return;
  const int64_t n = ((int64_t)c->count[cslot] + P - 1) / P;
  constexpr int T = kQThreads * kScanPer;
  const int64_t ntiles = (n + T - 1) / T;
  const uint32_t epoch = c->epoch_base + slot;
  const int t = threadIdx.x, lane = lane_id();
  int64_t tile;
  while ((tile = next_tile(&c->ticket[slot], ntiles, &s_tile)) >= 0) {
    const int64_t b0 = tile * T + (int64_t)t * kScanPer;
    uint64_t v[kScanPer];
    uint64_t sum = 0;
#pragma unroll
    for (int j = 0; j < kScanPer; j++) {
      v[j] = b0 + j < n ? ttot[b0 + j] : 0ull;
      sum += v[j];
    }
    uint64_t x = sum;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const uint64_t y = __shfl_up(x, o);
      if (lane >= o) x += y;
    }
    const uint64_t wtot = __shfl(x, kWave - 1);
    uint64_t b = chain_tile(wtot, tile, ntiles, status, epoch, s_wtot, &s_base, &s_total);
    b += x - sum;
#pragma unroll
    for (int j = 0; j < kScanPer; j++) {
      if (b0 + j < n) ttot[b0 + j] = b;
      b += v[j];
    }
  }
}

// Internal level `slot` without a tile chain (large runs): entries (query,
// node) whose children are nodes; fin == null: the implicit first frontier
// (query q0 + e, root).  Three launches, none of which waits on another tile:
//   k_expand_count  tests the children of each entry; per entry it stores the
//                   hit mask, the contained-and-safe mask, the query and the
//                   first child (24 B), per tile the survivor count; the
//                   level's total by one atomic per block.
//   k_tile_scan     exclusive scan of the tile counts.
//   k_expand_write  per tile: loads the entries into LDS, scans them locally,
//                   adds the tile's base and writes the survivors (query,
//                   child | FULL) as one flat, coalesced sequence in
//                   (query, node) order.
// A contained subtree travels down as one FULL entry (its box range is taken
// at the leaf stage); it survives every level as its own single child.
#ifndef CT_EXPAND_H
#define CT_EXPAND_H 2
#endif
template <int G, int STEPS>
__global__ __launch_bounds__(kQThreads) void k_expand_count(int slot, QTree tr, QBoxes qb,
                                                            QCtrl* c, const uint2* __restrict__ fin,
                                                            QExp* __restrict__ est,
                                                            uint64_t* __restrict__ ttot,
                                                            int64_t r_sz, bool xcd) {
  using S = QShape<G, STEPS>;
  constexpr int P = S::kPerTile;
  // memory round trips per tile: the entries of all steps, then the windows,
  // child MBRs and flags of H steps at a time
  constexpr int H = STEPS < CT_EXPAND_H ? STEPS : CT_EXPAND_H;
  static_assert(STEPS % H == 0, "steps in whole groups");
  __shared__ uint64_t s_red[kQWaves];
  // the tile's entries, written out once at the end of the tile (a store
  // between two rounds of loads would make the second wait for it)
  __shared__ QExp s_est[P];
  if (c->overflow) return;
  const int64_t n_in = (int64_t)c->count[slot];
  const int64_t q0 = c->q0;
  const int64_t ntiles = (n_in + P - 1) / P;
  const int t = threadIdx.x, w = t >> 6, lane = lane_id();
  if (ntiles == 0) return;  // the next level's count stays 0
  const int sub = lane / G, c_ = lane % G;
  const int lead = sub * G;
  // bulk-load layout: the entries of this launch sit on level nl-1-slot and
  // node i's children are nodes cs + i*F .. of the level below
  const bool canon = tr.canon && *tr.canon;
  const bool shortcut = shortcut_on(tr, canon);
  const int lvl = c->lv.nl - 1 - slot;
  const int64_t ls = c->lv.lstart[lvl];
  const int64_t cs = lvl > 0 ? c->lv.lstart[lvl - 1] : 0;
  const int64_t cc = lvl > 0 ? c->lv.lcount[lvl - 1] : 0;
  uint64_t block_total = 0;
  // tiles are independent (no chain): a fixed stride over them, no ticket
  const TileWalk tw = tile_walk(ntiles, xcd);
  for (int64_t tile = tw.first; tile < tw.end; tile += tw.step) {
    __syncthreads();  // the previous tile is done with the shared words
    uint64_t wsum = 0;
    const int64_t tbase = tile * P;
    // the tile's entries (a position past the frontier reads its last entry,
    // unused)
    uint2 ent[STEPS];
    load_entries<G, STEPS, S::kPerWave, S::E>(ent, fin, tbase + w * S::kPerWave + sub, n_in, q0);
#pragma unroll
    for (int k0 = 0; k0 < STEPS; k0 += H) {
      const uint2* en = ent + k0;
      // children of each entry (a contained entry is its own single child)
      // (node indices fit 32 bits)
      int p[H], L[H];
#pragma unroll
      for (int k = 0; k < H; k++) {
        const int64_t f = ((int64_t)(en[k].y & ~kFull) - ls) * tr.F;
        p[k] = (int)(cs + f);
        L[k] = (int)(cc - f < tr.F ? cc - f : tr.F);
      }
      if (!canon) {  // any other layout: the node's pos/len
        int pos[H], len[H];
#pragma unroll
        for (int k = 0; k < H; k++) {
          const uint32_t node = en[k].y & ~kFull;
          pos[k] = tr.pos[node];
          len[k] = tr.len[node];
        }
#pragma unroll
        for (int k = 0; k < H; k++) {
          p[k] = pos[k];
          L[k] = len[k];
        }
      }
      // windows, child MBRs and well-formed flags (a lane with no child
      // reads node 0, unused)
      float4 wq[H], mb[H];
      uint8_t sf[H];
      bool test[H];
      uint32_t qv[H];
#pragma unroll
      for (int k = 0; k < H; k++) qv[k] = en[k].x;
      load_windows<G, H>(wq, qb, qv);
#pragma unroll
      for (int k = 0; k < H; k++) {
        const bool full = (en[k].y & kFull) != 0;
        // children must be nodes (a corrupt pos/len is never followed)
        test[k] = !full && p[k] >= 0 && c_ < L[k] && (int64_t)p[k] + L[k] <= r_sz;
        const uint32_t child = test[k] ? (uint32_t)(p[k] + c_) : 0u;
        mb[k] = ld_box(tr.mbr, child);
        sf[k] = shortcut && node_safe(tr, child) ? 1 : 0;
      }
      bcast_windows<G, H>(wq);
#pragma unroll
      for (int k = 0; k < H; k++) {
        const int64_t e = tbase + w * S::kPerWave + (k0 + k) * S::E + sub;
        const bool valid = e < n_in;
        const bool full = (en[k].y & kFull) != 0;
        bool hit, cont;
        if (full) {
          hit = valid && c_ == 0;
          cont = hit;
        } else {
          hit = valid && test[k] &&
                intersect(wq[k].x, wq[k].y, wq[k].z, wq[k].w, mb[k].x, mb[k].y, mb[k].z, mb[k].w);
          // (a contained subtree's box range comes from the level table, so
          // the shortcut needs the bulk-load layout)
          cont = hit && shortcut &&
                 contains(wq[k].x, wq[k].y, wq[k].z, wq[k].w, mb[k].x, mb[k].y, mb[k].z,
                          mb[k].w) &&
                 sf[k];
        }
        const uint64_t gm = (ballot(hit) >> lead) & S::kGroupMask;
        const uint64_t fm = (ballot(cont) >> lead) & S::kGroupMask;
        if (c_ == 0 && valid) {
          QExp x;
          x.hit = gm;
          x.full = fm;
          x.q = en[k].x;
          x.first = full ? (en[k].y & ~kFull) : (uint32_t)p[k];
          s_est[e - tbase] = x;
        }
        if (c_ == 0 && valid) wsum += (uint64_t)popc64(gm);  // (summed over the wave below)
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) wsum += __shfl_xor(wsum, o);
    if (lane == 0) s_red[w] = wsum;
    __syncthreads();
    for (int i = t; i < P; i += kQThreads)
      if (tbase + i < n_in) est[tbase + i] = s_est[i];
    if (t == 0) {
      uint64_t tot = 0;
#pragma unroll
      for (int i = 0; i < kQWaves; i++) tot += s_red[i];
      ttot[tile] = tot;
      block_total += tot;
    }
  }
  if (t == 0 && block_total)
    atomicAdd((unsigned long long*)&c->count[slot + 1], (unsigned long long)block_total);
}

// Containment of box cells b in window cells (low wl, high wh): window low
// side <= box low side and box high side <= window high side; 1 proven, 0
// disproven, -1 open.  Only for a window without NaN (a NaN window contains
// nothing; its cells could prove otherwise); a NaN in the box takes cell 0
// (low side) or 255 (high side) and is never proven contained.
__device__ __forceinline__ int cells_contain(ct_s16x2 wl, ct_s16x2 wh, uint32_t b) {
  const ct_s16x2 d1 = cells_low(b) - wl, d2 = wh - cells_high(b);
  return lanes_sign(__builtin_elementwise_min(d1, d2));
}

// Internal level `slot` over the node codes (tr.nframe / tr.ncode, G >= 4):
// the same tiles and QExp records as k_expand_count, but G/4 lanes per entry,
// each lane the cells of four children (slots g + j * G/4, one 16-byte load),
// so a tile is one round of loads.  A child whose cells leave the intersection open reads its
// MBR at the end of the tile, and so does a child that intersects but whose
// containment (the shortcut) the cells leave open.
template <int G, int STEPS>
__global__ __launch_bounds__(kQThreads)
#if CT_CODED_WAVES > 0
__attribute__((amdgpu_waves_per_eu(CT_CODED_WAVES)))
#endif
void k_expand_count_coded(int slot, QTree tr, QBoxes qb, QCtrl* c,
                          const uint2* __restrict__ fin, QExp* __restrict__ est,
                          uint64_t* __restrict__ ttot, int64_t r_sz, bool xcd) {
  using S = QShape<G, STEPS>;
  constexpr int P = S::kPerTile;
  constexpr int GC = G / 4;             // lanes per entry
  constexpr int EC = kWave / GC;        // entries per wave-step
  constexpr int SC = S::kPerWave / EC;  // wave-steps per tile
  static_assert(G >= 4 && SC * EC == S::kPerWave, "whole wave-steps");
  constexpr uint32_t kNib = GC >= 32 ? 0xFFFFFFFFu : ((1u << GC) - 1u);
  __shared__ uint64_t s_red[kQWaves];
  __shared__ QExp s_est[P];
  if (c->overflow) return;
  const int64_t n_in = (int64_t)c->count[slot];
  const int64_t q0 = c->q0;
  const int64_t ntiles = (n_in + P - 1) / P;
  const int t = threadIdx.x, w = t >> 6, lane = lane_id();
  if (ntiles == 0) return;
  const int sub = lane / GC, cg = lane % GC;
  const int lead = sub * GC;
  const bool canon = tr.canon && *tr.canon;
  // (a tree with node codes is a kept tree: its shortcut needs no per-node flags)
  const bool shortcut = shortcut_on(tr, canon) && tr.safe == nullptr;
  const int lvl = c->lv.nl - 1 - slot;
  const int64_t ls = c->lv.lstart[lvl];
  const int64_t cs = lvl > 0 ? c->lv.lstart[lvl - 1] : 0;
  const int64_t cc = lvl > 0 ? c->lv.lcount[lvl - 1] : 0;
  uint64_t block_total = 0;
  const TileWalk tw = tile_walk(ntiles, xcd);
  for (int64_t tile = tw.first; tile < tw.end; tile += tw.step) {
    __syncthreads();
    uint64_t wsum = 0;
    const int64_t tbase = tile * P;
    uint2 ent[SC];
#pragma unroll
    for (int k = 0; k < SC; k++) {
      const int64_t e = tbase + w * S::kPerWave + k * EC + sub;
      ent[k] = fin ? fin[e < n_in ? e : n_in - 1]
                   : make_uint2((uint32_t)(q0 + (e < n_in ? e : n_in - 1)), 0u);
    }
    int p[SC], L[SC];
    bool coded[SC];
#pragma unroll
    for (int k = 0; k < SC; k++) {
      const uint32_t node = ent[k].y & ~kFull;
      if (canon) {
        const int64_t f = ((int64_t)node - ls) * tr.F;
        p[k] = (int)(cs + f);
        L[k] = (int)(cc - f < tr.F ? cc - f : tr.F);
      } else {
        p[k] = tr.pos[node];
        L[k] = tr.len[node];
      }
      coded[k] = canon && (ent[k].y & kFull) == 0;
    }
    float4 wq[SC], fr[SC];
    uint4 cd[SC];
    uint32_t test[SC];  // bit j: child slot j * GC + cg exists (children must be nodes)
#pragma unroll
    for (int k = 0; k < SC; k++) {
      test[k] = 0;
      const bool full = (ent[k].y & kFull) != 0;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int sl = j * GC + cg;
        if (!full && p[k] >= 0 && sl < L[k] && (int64_t)p[k] + L[k] <= r_sz) test[k] |= 1u << j;
      }
      wq[k] = ld_box(qb, ent[k].x);
      const uint32_t node = ent[k].y & ~kFull;
      const bool ld = coded[k] && test[k] != 0;
      fr[k] = ld ? tr.nframe[node] : make_float4(0.f, 0.f, 0.f, 0.f);
      cd[k] = ld ? ((const uint4*)tr.ncode)[(int64_t)node * GC + cg] : make_uint4(0, 0, 0, 0);
    }
    uint32_t hit[SC], cont[SC], open[SC];
#pragma unroll
    for (int k = 0; k < SC; k++) {
      const int64_t e = tbase + w * S::kPerWave + k * EC + sub;
      hit[k] = cont[k] = open[k] = 0;
      if (e < n_in && test[k]) {
        if (fr[k].z > 0.f) {
          const uint32_t wc = box_cells(wq[k].x, wq[k].y, wq[k].z, wq[k].w, fr[k]);
          const ct_s16x2 wl = cells_low(wc), wh = cells_high(wc);
          const bool wnan = wq[k].x != wq[k].x || wq[k].y != wq[k].y || wq[k].z != wq[k].z ||
                            wq[k].w != wq[k].w;
          const uint32_t b4[4] = {cd[k].x, cd[k].y, cd[k].z, cd[k].w};
#pragma unroll
          for (int j = 0; j < 4; j++) {
            if (!((test[k] >> j) & 1u)) continue;
            const int r = cells_intersect(wl, wh, b4[j]);
            if (r > 0) {
              hit[k] |= 1u << j;
              const int rc = shortcut && !wnan ? cells_contain(wl, wh, b4[j]) : 0;
              if (rc > 0) cont[k] |= 1u << j;
              else if (rc < 0) open[k] |= 1u << j;  // (the MBR settles the containment)
            } else if (r < 0) {
              open[k] |= 1u << j;
            }
          }
        } else {
          open[k] = test[k];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < SC; k++) {
      if (open[k]) {
        float4 mb[4];
#pragma unroll
        for (int j = 0; j < 4; j++)
          if ((open[k] >> j) & 1u) mb[j] = ld_box(tr.mbr, (uint32_t)(p[k] + j * GC + cg));
#pragma unroll
        for (int j = 0; j < 4; j++)
          if (((open[k] >> j) & 1u) &&
              intersect(wq[k].x, wq[k].y, wq[k].z, wq[k].w, mb[j].x, mb[j].y, mb[j].z, mb[j].w)) {
            hit[k] |= 1u << j;
            if (shortcut &&
                contains(wq[k].x, wq[k].y, wq[k].z, wq[k].w, mb[j].x, mb[j].y, mb[j].z, mb[j].w))
              cont[k] |= 1u << j;
          }
      }
    }
#pragma unroll
    for (int k = 0; k < SC; k++) {
      const int64_t e = tbase + w * S::kPerWave + k * EC + sub;
      const bool valid = e < n_in;
      const bool full = (ent[k].y & kFull) != 0;
      uint64_t gm = 0, fm = 0;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        gm |= (uint64_t)((uint32_t)(ballot((hit[k] >> j) & 1u) >> lead) & kNib) << (j * GC);
        fm |= (uint64_t)((uint32_t)(ballot((cont[k] >> j) & 1u) >> lead) & kNib) << (j * GC);
      }
      if (full) gm = fm = 1;  // a contained entry is its own single child
      if (cg == 0 && valid) {
        QExp x;
        x.hit = gm;
        x.full = fm;
        x.q = ent[k].x;
        x.first = full ? (ent[k].y & ~kFull) : (uint32_t)p[k];
        s_est[e - tbase] = x;
        wsum += (uint64_t)popc64(gm);
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) wsum += __shfl_xor(wsum, o);
    if (lane == 0) s_red[w] = wsum;
    __syncthreads();
    for (int i = t; i < P; i += kQThreads)
      if (tbase + i < n_in) est[tbase + i] = s_est[i];
    if (t == 0) {
      uint64_t tot = 0;
#pragma unroll
      for (int i = 0; i < kQWaves; i++) tot += s_red[i];
      ttot[tile] = tot;
      block_total += tot;
    }
  }
  if (t == 0 && block_total)
    atomicAdd((unsigned long long*)&c->count[slot + 1], (unsigned long long)block_total);
}


template <int G, int STEPS>
__global__ __launch_bounds__(kQThreads) void k_expand_write(int slot, QCtrl* c,
                                                            const QExp* __restrict__ est,
                                                            const uint64_t* __restrict__ tbase_arr,
                                                            uint2* __restrict__ fout, int64_t cap,
                                                            bool xcd) {
  using S = QShape<G, STEPS>;
  // a write tile is M count tiles (at least one entry per thread): its base is
  // the first count tile's scanned total
  constexpr int PC = S::kPerTile;
  constexpr int M = PC < kQThreads ? kQThreads / PC : 1;
  constexpr int P = PC * M;                           // entries per write tile
  constexpr int R = (P + kQThreads - 1) / kQThreads;
  __shared__ uint64_t s_hit[P];
  __shared__ uint64_t s_full[P];
  __shared__ uint32_t s_q[P];
  __shared__ uint32_t s_first[P];
  __shared__ uint32_t s_off[P];  // tile-relative first output slot
  __shared__ uint32_t s_tot;
  __shared__ uint64_t s_scan[kQWaves];
  if (c->overflow) return;
  if ((int64_t)c->count[slot + 1] > cap) {  // the next frontier does not fit
    if (blockIdx.x == 0 && threadIdx.x == 0) c->overflow = 1;
    return;
  }
  const int64_t n_in = (int64_t)c->count[slot];
  const int64_t ntiles = (n_in + P - 1) / P;
  const int t = threadIdx.x;
  const TileWalk tw = tile_walk(ntiles, xcd);
  for (int64_t tile = tw.first; tile < tw.end; tile += tw.step) {
    __syncthreads();  // the previous tile is done with the shared arrays
    uint64_t loc[R];
    uint64_t sum = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int li = t * R + r;
      loc[r] = 0;
      if (li < P) {
        const int64_t e = tile * P + li;
        uint64_t h = 0, f = 0;
        uint32_t q = 0, fc = 0;
        if (e < n_in) {
          const QExp x = est[e];
          h = x.hit;
          f = x.full;
          q = x.q;
          fc = x.first;
        }
        s_hit[li] = h;
        s_full[li] = f;
        s_q[li] = q;
        s_first[li] = fc;
        loc[r] = (uint64_t)popc64(h);
        sum += loc[r];
      }
    }
    uint64_t ttot;
    uint64_t ex = block_excl_scan64(sum, s_scan, &ttot);
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int li = t * R + r;
      if (li < P) s_off[li] = (uint32_t)ex;
      ex += loc[r];
    }
    if (t == 0) s_tot = (uint32_t)ttot;
    __syncthreads();
    const uint64_t base = tbase_arr[tile * M];
    const uint32_t tn = s_tot;
    // four slots per thread at a time, their searches stepping together
    constexpr int U = 4;
    for (uint32_t j0 = t; j0 < tn; j0 += U * kQThreads) {
      uint32_t jc[U];
      int lo[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const uint32_t j = j0 + u * kQThreads;
        jc[u] = j < tn ? j : tn - 1;
        lo[u] = 0;
      }
#pragma unroll
      for (int step = P / 2; step > 0; step >>= 1) {
        uint32_t f[U];
#pragma unroll
        for (int u = 0; u < U; u++) f[u] = s_off[lo[u] + step];
#pragma unroll
        for (int u = 0; u < U; u++)
          if (f[u] <= jc[u]) lo[u] += step;
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const uint32_t j = j0 + u * kQThreads;
        const uint32_t r = jc[u] - s_off[lo[u]];
        const uint32_t bit = nth_set_bit<G>(s_hit[lo[u]], r);
        const uint32_t child = s_first[lo[u]] + bit;
        const uint32_t fl = ((s_full[lo[u]] >> bit) & 1ull) ? kFull : 0u;
        const uint64_t o = base + j;
        if (j < tn && o < (uint64_t)cap) fout[o] = make_uint2(s_q[lo[u]], child | fl);
      }
    }
  }
}

// Exclusive scans of two u64 per thread over the block (one pair of barriers).
__device__ __forceinline__ void block_excl_scan64x2(uint64_t& a, uint64_t& b, uint64_t* s_tmp,
                                                    uint64_t* ta, uint64_t* tb) {
  const int lane = lane_id(), w = threadIdx.x >> 6;
  uint64_t x = a, y = b;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint64_t xu = __shfl_up(x, o), yu = __shfl_up(y, o);
    if (lane >= o) {
      x += xu;
      y += yu;
    }
  }
  if (lane == kWave - 1) {
    s_tmp[w] = x;
    s_tmp[kQWaves + w] = y;
  }
  __syncthreads();
  uint64_t ox = 0, oy = 0, tx = 0, ty = 0;
#pragma unroll
  for (int i = 0; i < kQWaves; i++) {
    const uint64_t u = s_tmp[i], v = s_tmp[kQWaves + i];
    if (i < w) {
      ox += u;
      oy += v;
    }
    tx += u;
    ty += v;
  }
  __syncthreads();
  *ta = tx;
  *tb = ty;
  a = x - a + ox;
  b = y - b + oy;
}

// The leaf write's staging: K consecutive flat slots per thread, one uint2
// per slot (sorted box; list index << 16 | slot within the entry), padded by one
// uint2 every 16 so that neither the per-thread writes (stride K) nor the
// linear reads of the copy-out meet more than two-way bank conflicts.
__device__ __forceinline__ int stage_at(int p) { return p + (p >> 4); }

template <int G, int STEPS>
__global__ __launch_bounds__(kQThreads) void k_leaf_write(
    int slot, QCtrl* c, const QLeaf* __restrict__ est, const uint64_t* __restrict__ tbase_arr,
    const int* __restrict__ rid, int* __restrict__ out_fid, int* __restrict__ out_tid,
    int64_t out_cap, QPiece* __restrict__ pieces, int64_t piece_cap, bool xcd,
    const uint32_t* __restrict__ tmap, const int64_t* __restrict__ ooff,
    const int64_t* __restrict__ poff) {
  using S = QShape<G, STEPS>;
  // a write tile is M count tiles (at least one entry per thread): its base is
  // the first count tile's scanned total
  constexpr int PC = S::kPerTile;
  constexpr int M = PC < kQThreads ? kQThreads / PC : 1;
  constexpr int P = PC * M;                           // entries per write tile
  constexpr int R = (P + kQThreads - 1) / kQThreads;  // entries per thread
  constexpr int K = P >= 2048 ? 4 : 8;                // flat slots per thread and chunk
  constexpr int C = K * kQThreads;                    // flat slots per chunk
  static_assert(P <= 65536 && kFlatMax < 65536, "list index and slot share a word");
  // the flat entries (partial entries with hits, ranges of 1..kFlatMax boxes)
  // in tile order: first flat slot (~0u past the list), hit mask or length,
  // first box, query | kQRange, tile-relative first output slot
  __shared__ uint32_t s_lfoff[P];
  __shared__ uint64_t s_lval[P];
  __shared__ uint32_t s_lfirst[P];
  __shared__ uint32_t s_lq[P];
  __shared__ uint64_t s_loff[P];
  __shared__ uint2 s_stage[C + C / 16];
  __shared__ uint64_t s_scan[2 * kQWaves];
  if (c->overflow) return;
  const int64_t n_in = (int64_t)c->count[slot];
  const int64_t ntiles = (n_in + P - 1) / P;
  const int t = threadIdx.x;
  const TileWalk tw = tile_walk(ntiles, xcd);
  for (int64_t tile = tw.first; tile < tw.end; tile += tw.step) {
    __syncthreads();  // the previous tile is done with the shared arrays
    // the tile's entries (in registers), then scans of all pairs and of the
    // flat slots and entries (slots in the low word, entries in the high)
    uint64_t loc[R], shift[R], val[R];
    uint32_t qv[R], fv[R];
    uint64_t sum = 0, fsum = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int li = t * R + r;
      loc[r] = 0;
      shift[r] = 0;
      val[r] = 0;
      fv[r] = 0;
      uint32_t qf = kQNone;
      if (li < P) {
        const int64_t e = tile * P + li;
        if (e < n_in) {
          const QLeaf x = est[e];
          qf = x.qf;
          val[r] = x.val;
          fv[r] = x.first;
          loc[r] = (qf & kQRange) ? x.val : (uint64_t)popc64(x.val);
          // windows in a run order: the entry's query is a rank; its pairs
          // move from the window's first pair in run order to its first pair
          // in input order (modulo 2^64), and carry the input index
          if (tmap) {
            const uint32_t rk = qf & ~kQRange, qi = tmap[rk];
            shift[r] = (uint64_t)ooff[rk] - (uint64_t)poff[rk];
            qf = qi | (qf & kQRange);
          }
        }
        sum += loc[r];
        const bool piece = (qf & kQRange) && qf != kQNone && loc[r] > kFlatMax;
        if (!piece && loc[r] > 0) fsum += loc[r] | (1ull << 32);
      }
      qv[r] = qf;
    }
    uint64_t ttot, ftot;
    uint64_t ex = sum, fex = fsum;
    block_excl_scan64x2(ex, fex, s_scan, &ttot, &ftot);
    const uint32_t fn = (uint32_t)ftot;          // flat slots of the tile
    const int nl = (int)(ftot >> 32);            // flat entries
    const uint64_t base = tbase_arr[tile * M];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int li = t * R + r;
      if (li < P) {
        const uint32_t qf = qv[r];
        const bool piece = (qf & kQRange) && qf != kQNone && loc[r] > kFlatMax;
        const uint64_t off = ex + shift[r];
        if (!piece && loc[r] > 0) {
          const int k = (int)(fex >> 32);
          s_lfoff[k] = (uint32_t)fex;
          s_lval[k] = val[r];
          s_lfirst[k] = fv[r];
          s_lq[k] = qf;
          s_loff[k] = off;
          fex += loc[r] | (1ull << 32);
        }
        if (li >= nl) s_lfoff[li] = ~0u;  // (list entries are below nl: no overlap)
        // long ranges: pieces for k_write_pieces
        if (piece) {
          const uint64_t ev = val[r];
          const uint64_t np = (ev + kPiece - 1) / kPiece;
          const uint64_t at = atomicAdd(&c->npieces, (unsigned long long)np);
          for (uint64_t p = 0; p < np; p++) {
            if (at + p >= (uint64_t)piece_cap) break;  // only when the total exceeds out_cap
            const uint64_t rest = ev - p * kPiece;
            QPiece pc;
            pc.off = base + off + p * kPiece;
            pc.first = (int64_t)fv[r] + (int64_t)(p * kPiece);
            pc.len = (uint32_t)(rest < kPiece ? rest : kPiece);
            pc.q = qf & ~kQRange;
            pieces[at + p] = pc;
          }
        }
      }
      ex += loc[r];
    }
    __syncthreads();
    // flat slots in chunks of C: thread t walks slots j0 .. j0+K-1 of the
    // chunk (one binary search for the first, then entry by entry and hit by
    // hit) and stages their boxes; the block then loads the box ids and
    // stores the chunk linearly (coalesced)
    for (uint32_t c0 = 0; c0 < fn; c0 += C) {
      const uint32_t j0 = c0 + (uint32_t)t * K;
      if (j0 < fn) {
        int lo = 0;
#pragma unroll
        for (int step = P / 2; step > 0; step >>= 1)
          if (s_lfoff[lo + step] <= j0) lo += step;
        uint32_t r = j0 - s_lfoff[lo];
        uint64_t v = s_lval[lo];
        uint32_t first = s_lfirst[lo];
        bool rng = (s_lq[lo] & kQRange) != 0;
        uint32_t n_e = rng ? (uint32_t)v : (uint32_t)popc64(v);
        // partial entries: the hits still to write (the lowest is this slot's)
        uint64_t m = rng ? 0ull : v & (~0ull << nth_set_bit<G>(v, r));
        uint32_t bi[K], pk[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
          bi[k] = first + (rng ? r : (uint32_t)__builtin_ctzll(m | (1ull << 63)));
          pk[k] = ((uint32_t)lo << 16) | r;
          if (k + 1 < K) {
            r++;
            m &= m - 1ull;
            if (r == n_e && lo + 1 < nl) {  // the next flat entry
              lo++;
              r = 0;
              v = s_lval[lo];
              first = s_lfirst[lo];
              rng = (s_lq[lo] & kQRange) != 0;
              n_e = rng ? (uint32_t)v : (uint32_t)popc64(v);
              m = rng ? 0ull : v;
            }
          }
        }
#pragma unroll
        for (int k = 0; k < K; k++) s_stage[stage_at(t * K + k)] = make_uint2(bi[k], pk[k]);
      }
      __syncthreads();
      // the box ids are loaded here, in slot order (a wave's loads are the
      // consecutive boxes of a few entries)
      const uint32_t cn = fn - c0 < (uint32_t)C ? fn - c0 : (uint32_t)C;
      uint64_t o[K];
      uint32_t qo[K];
      int idv[K];
#pragma unroll
      for (int k = 0; k < K; k++) {
        const uint32_t cj = (uint32_t)(k * kQThreads + t);
        o[k] = ~0ull;
        qo[k] = 0;
        uint32_t b = 0;
        if (cj < cn) {
          const uint2 x = s_stage[stage_at((int)cj)];
          const uint32_t li = x.y >> 16;
          o[k] = base + s_loff[li] + (x.y & 0xFFFFu);
          qo[k] = s_lq[li] & ~kQRange;
          b = x.x;
        }
        idv[k] = o[k] < (uint64_t)out_cap ? rid[b] : 0;
      }
#pragma unroll
      for (int k = 0; k < K; k++)
        if (o[k] < (uint64_t)out_cap) {
          out_fid[o[k]] = idv[k];
          out_tid[o[k]] = (int)qo[k];
        }
      __syncthreads();  // the stage is free again
    }
  }
}

// Leaf stage in one pass (writing runs): the boxes of every entry are tested
// once, the tile's pair total is published through a decoupled look-back
// (tiles taken in order from a ticket), and every hit is stored straight from
// the registers that tested it -- the box id is loaded beside the box, and the
// hits of one wave-step are consecutive output slots, so the stores of a step
// are one contiguous run.  Contained ranges are written by the block after the
// partial entries (short ones flat and load-balanced, long ones cut into
// pieces for k_write_pieces).  QCOUNTS: per-query counts as k_leaf_count.
// The last tile leaves the run's pair total in c->total.
template <int G, int STEPS, bool QCOUNTS>
__global__ __launch_bounds__(kQThreads) void k_leaf_fused(
    int slot, QTree tr, QBoxes qb, QBoxes rb, QCtrl* c, const uint2* __restrict__ fin,
    int64_t r_sz, int64_t nbox, const int* __restrict__ rid, int* __restrict__ out_fid,
    int* __restrict__ out_tid, int64_t out_cap, uint64_t* status, QPiece* __restrict__ pieces,
    int64_t piece_cap, unsigned long long* counts) {
  using S = QShape<G, STEPS>;
  constexpr int P = S::kPerTile;
  constexpr int H = STEPS < CT_LEAF_H ? STEPS : CT_LEAF_H;
  static_assert(STEPS % H == 0, "steps in whole groups");
  constexpr int R = (P + kQThreads - 1) / kQThreads;  // entries per thread in the scans
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_base;
  __shared__ uint64_t s_scan[kQWaves];
  __shared__ uint64_t s_off[P];    // tile-relative first output slot of each entry
  __shared__ uint32_t s_cnt[P];    // pairs of each entry (ranges: their length)
  __shared__ uint32_t s_q[P];      // query | kQRange for ranges, kQNone past the end
  __shared__ uint32_t s_first[P];  // ranges: first sorted box
  __shared__ uint32_t s_foff[P];   // ranges up to kFlatMax: tile-relative flat offset
  __shared__ uint32_t s_ftot;
  __shared__ QLevels s_lv;
  if (c->overflow) return;
  const int64_t n_in = (int64_t)c->count[slot];
  const uint32_t epoch = c->epoch_base + slot;
  const int64_t q0 = c->q0;
  const int64_t ntiles = (n_in + P - 1) / P;
  const int t = threadIdx.x, w = t >> 6, lane = lane_id();
  if (ntiles == 0) return;  // the run's total stays 0
  for (int i = t; i < (int)(sizeof(QLevels) / 4); i += kQThreads)
    ((uint32_t*)&s_lv)[i] = ((const uint32_t*)&c->lv)[i];
  const int sub = lane / G, c_ = lane % G;
  const int lead = sub * G;
  const uint64_t below = (1ull << c_) - 1ull;  // group lanes before this one
  const bool canon = tr.canon && *tr.canon;
  const int64_t ls = c->lv.lstart[0];
  int64_t tile;
  while ((tile = next_tile(&c->ticket[slot], ntiles, &s_tile)) >= 0) {
    const int64_t tbase = tile * P;
    uint2 ent[STEPS];
    load_entries<G, STEPS, S::kPerWave, S::E>(ent, fin, tbase + w * S::kPerWave + sub, n_in, q0);
    // kept for the stores: this lane's group mask and box id per step
    uint64_t gmk[STEPS];
    int idk[STEPS];
#pragma unroll
    for (int k0 = 0; k0 < STEPS; k0 += H) {
      uint32_t q[H];
      int first[H], cnt[H];  // contained: range length; partial: children
      bool full[H];
      const uint2* en = ent + k0;
#pragma unroll
      for (int k = 0; k < H; k++) {
        q[k] = en[k].x;
        const uint32_t node = en[k].y & ~kFull;
        full[k] = (en[k].y & kFull) != 0;
        if (full[k]) {
          int64_t f, n;
          node_range(node, &s_lv, nbox, &f, &n);
          first[k] = (int)f;
          cnt[k] = (int)n;
        } else {
          const int64_t f = ((int64_t)node - ls) * tr.F;
          first[k] = (int)f;
          cnt[k] = (int)(nbox - f < tr.F ? nbox - f : tr.F);
        }
      }
      if (!canon) {  // any other layout: the node's pos/len
        int pos[H], len[H];
#pragma unroll
        for (int k = 0; k < H; k++) {
          const uint32_t node = en[k].y & ~kFull;
          pos[k] = tr.pos[node];
          len[k] = tr.len[node];
        }
#pragma unroll
        for (int k = 0; k < H; k++)
          if (!full[k]) {
            first[k] = (int)((int64_t)pos[k] - r_sz);
            cnt[k] = len[k];
          }
      }
      float4 wq[H], bx[H];
      int id[H];
      bool test[H];
      load_windows<G, H>(wq, qb, q);
#pragma unroll
      for (int k = 0; k < H; k++) {
        const int bi = first[k] + c_;
        test[k] = !full[k] && c_ < cnt[k] && bi >= 0 && bi < nbox;  // children must be boxes
        bx[k] = ld_box(rb, test[k] ? bi : 0);
        id[k] = rid[test[k] ? bi : 0];
      }
      bcast_windows<G, H>(wq);
#pragma unroll
      for (int k = 0; k < H; k++) {
        const int el = w * S::kPerWave + (k0 + k) * S::E + sub;  // entry in the tile
        const bool valid = tbase + el < n_in;
        const bool hit = valid && test[k] &&
                         intersect(wq[k].x, wq[k].y, wq[k].z, wq[k].w, bx[k].x, bx[k].y, bx[k].z,
                                   bx[k].w);
        const uint64_t gm = (ballot(hit) >> lead) & S::kGroupMask;
        gmk[k0 + k] = gm;
        idk[k0 + k] = id[k];
        if (c_ == 0) {
          s_cnt[el] = !valid ? 0u : full[k] ? (uint32_t)cnt[k] : (uint32_t)popc64(gm);
          s_q[el] = !valid ? kQNone : (q[k] | (full[k] ? kQRange : 0u));
          s_first[el] = (uint32_t)first[k];
        }
      }
    }
    __syncthreads();
    // tile-relative offsets of every entry and flat offsets of short ranges
    uint64_t loc[R];
    uint64_t sum = 0, fsum = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int li = t * R + r;
      loc[r] = li < P ? s_cnt[li] : 0;
      sum += loc[r];
      if (li < P && (s_q[li] & kQRange) && s_q[li] != kQNone && loc[r] <= kFlatMax)
        fsum += loc[r];
    }
    uint64_t ttot, ftot;
    uint64_t ex = block_excl_scan64(sum, s_scan, &ttot);
    uint64_t fex = block_excl_scan64(fsum, s_scan, &ftot);
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int li = t * R + r;
      if (li < P) {
        s_off[li] = ex;
        s_foff[li] = (uint32_t)fex;
        if ((s_q[li] & kQRange) && s_q[li] != kQNone && loc[r] <= kFlatMax) fex += loc[r];
      }
      ex += loc[r];
    }
    // the tile's base: decoupled look-back over the tiles before it
    if (w == 0) {
      uint64_t excl = 0;
      if (tile == 0) {
        if (lane == 0) st_agent(status, st_pack(2u * epoch + 2u, ttot));
      } else {
        if (lane == 0) st_agent(status + tile, st_pack(2u * epoch + 1u, ttot));
        excl = lookback64(status, tile, epoch);
        if (lane == 0) st_agent(status + tile, st_pack(2u * epoch + 2u, excl + ttot));
      }
      if (lane == 0) {
        s_base = excl;
        s_ftot = (uint32_t)ftot;
        if (tile == ntiles - 1) c->total = excl + ttot;
      }
    }
    __syncthreads();
    const uint64_t base = s_base;
    // partial entries: each hit from its own lane (ranges have no hit bits)
#pragma unroll
    for (int k = 0; k < STEPS; k++) {
      const uint64_t gm = gmk[k];
      if ((gm >> c_) & 1ull) {
        const int el = w * S::kPerWave + k * S::E + sub;
        const uint64_t o = base + s_off[el] + (uint64_t)popc64(gm & below);
        if (o < (uint64_t)out_cap) {
          out_fid[o] = idk[k];
          out_tid[o] = (int)(s_q[el]);
        }
      }
    }
    // contained ranges: long ones as pieces, short ones flat over the block
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int li = t * R + r;
      if (li >= P) break;
      const uint32_t qf = s_q[li];
      if (qf == kQNone || !(qf & kQRange) || s_cnt[li] <= kFlatMax) continue;
      const uint64_t ev = s_cnt[li];
      const uint64_t off = base + s_off[li];
      const uint64_t np = (ev + kPiece - 1) / kPiece;
      const uint64_t at = atomicAdd(&c->npieces, (unsigned long long)np);
      for (uint64_t p = 0; p < np; p++) {
        if (at + p >= (uint64_t)piece_cap) break;  // only when the total exceeds out_cap
        const uint64_t rest = ev - p * kPiece;
        QPiece pc;
        pc.off = off + p * kPiece;
        pc.first = (int64_t)s_first[li] + (int64_t)(p * kPiece);
        pc.len = (uint32_t)(rest < kPiece ? rest : kPiece);
        pc.q = qf & ~kQRange;
        pieces[at + p] = pc;
      }
    }
    const uint32_t fn = s_ftot;
    for (uint32_t j = t; j < fn; j += kQThreads) {
      int lo = 0;
#pragma unroll
      for (int step = P / 2; step > 0; step >>= 1)
        if (s_foff[lo + step] <= j) lo += step;
      // (entries that are not short ranges hold no flat slots: the search
      // lands on the short range that holds slot j)
      const uint32_t r = j - s_foff[lo];
      const uint64_t o = base + s_off[lo] + r;
      if (o < (uint64_t)out_cap) {
        out_fid[o] = rid[(int64_t)s_first[lo] + r];
        out_tid[o] = (int)(s_q[lo] & ~kQRange);
      }
    }
    if (QCOUNTS) {
      // per-query counts: sums over runs of one query's entries in the tile
      const int nv = n_in - tbase < P ? (int)(n_in - tbase) : P;
      uint64_t csum = 0;
#pragma unroll
      for (int r = 0; r < R; r++)
        if (t * R + r < nv) csum += s_cnt[t * R + r];
      uint64_t tile_sum;
      uint64_t before = block_excl_scan64(csum, s_scan, &tile_sum);
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int i = t * R + r;
        if (i >= nv) break;
        const uint32_t qi = s_q[i] & ~kQRange;
        if (i == 0 || (s_q[i - 1] & ~kQRange) != qi) atomicAdd(counts + qi, 0ull - before);
        before += s_cnt[i];
        if (i + 1 == nv || (s_q[i + 1] & ~kQRange) != qi) atomicAdd(counts + qi, before);
      }
    }
  }
}

// Long box ranges of contained subtrees: one block per piece, coalesced id
// reads and pair writes.
__global__ __launch_bounds__(256) void k_write_pieces(const QCtrl* c,
                                                      const QPiece* __restrict__ pieces,
                                                      int64_t piece_cap,
                                                      const int* __restrict__ rid,
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
// (box4 non-null: the AoS copy of the boxes is written in the same pass)
__global__ __launch_bounds__(256) void k_safe_leaf_pow2(int64_t nbox, QBoxes rb, uint8_t* safe0,
                                                        float4* __restrict__ box4) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  float4 b = make_float4(0.f, 0.f, 0.f, 0.f);
  if (i < nbox) {
    b = make_float4(rb.x0[i], rb.y0[i], rb.x1[i], rb.y1[i]);
    if (box4) box4[i] = b;
  }
  const bool wf = i >= nbox || (b.x <= b.z && b.y <= b.w);
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

static int fixed_full(const char* what, int64_t want, int64_t cap) {
  return set_error(CURTREE_ERANGE,
                   "query: %s needs %lld entries, the fixed scratch holds %lld", what,
                   (long long)want, (long long)cap);
}

int QEngine::reserve_front(int64_t entries) {
  if (entries <= cap_front) return CURTREE_OK;
  if (fixed) return fixed_full("the frontier", entries, cap_front);
  int rc;
  // per-entry state (a QLeaf or QExp each) and tile counts (tiles hold >= 16
  // entries)
  if ((rc = dev_alloc(&front[0], (size_t)entries * 8)) ||
      (rc = dev_alloc(&front[1], (size_t)entries * 8)) ||
      (rc = dev_alloc(&est, (size_t)entries * kEntryBytes)) ||
      (rc = dev_alloc((void**)&ttot, (size_t)((wide ? entries : entries / 16) + 2) * 8))) {
    cap_front = 0;
    return rc;
  }
  cap_front = entries;
  return CURTREE_OK;
}

int QEngine::reserve_pieces(int64_t n) {
  if (n <= cap_pieces) return CURTREE_OK;
  if (fixed) return fixed_full("the piece list", n, cap_pieces);
  n = std::max<int64_t>(n + n / 4, 4096);
  int rc = dev_alloc(&pieces, (size_t)n * sizeof(QPiece));
  cap_pieces = rc ? 0 : n;
  return rc;
}

int QEngine::reserve_status(int64_t words) {
  if (words <= status_words) return CURTREE_OK;
  if (fixed) return fixed_full("the scan status", words, status_words);
  words = std::max<int64_t>(words + words / 2, 1 << 16);
  int rc = dev_alloc((void**)&status, (size_t)words * 8);
  status_words = rc ? 0 : words;
  needs_clear = true;
  return rc;
}

int QEngine::reserve_safe(int64_t nodes) {
  if (nodes <= cap_safe) return CURTREE_OK;
  if (fixed) return fixed_full("the node flags", nodes, cap_safe);
  int rc = dev_alloc((void**)&safe, (size_t)std::max<int64_t>(nodes, 256));
  cap_safe = rc ? 0 : nodes;
  return rc;
}

int QEngine::reserve_plan(int64_t chunks) {
  if (chunks <= cap_plan) return CURTREE_OK;
  if (fixed) return fixed_full("the chunk plan", chunks, cap_plan);
  int rc = dev_alloc((void**)&plan, (size_t)chunks * 16);
  cap_plan = rc ? 0 : chunks;
  return rc;
}

// Window order buffers for `windows` windows: the order and the ranks (4 B
// each), the windows in run order (16 B), counts (8 B), first pair in input
// and in run order (8 B each), and two block sums per kOrdBlock windows.
constexpr int64_t kOrdBlock = 4096;
int QEngine::reserve_order(int64_t windows) {
  if (windows <= cap_order) return CURTREE_OK;
  if (fixed) return fixed_full("the window order", windows, cap_order);
  const int64_t nb = (windows + kOrdBlock - 1) / kOrdBlock;
  int rc = dev_alloc(&order, (size_t)windows * 48 + (size_t)nb * 16 + 1024);
  cap_order = rc ? 0 : windows;
  return rc;
}

int QEngine::ensure_ctrl() {
  if (ctrl) return CURTREE_OK;
  if (fixed) return set_error(CURTREE_EINVAL, "query: fixed engine without a control block");
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
  if (fixed) return;  // slices of memory the engine does not own
  void* ps[] = {front[0], front[1], pieces, status, ctrl, safe, plan,
                box4,     node4,    win4,   canon,  est,  ttot, order};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  box4 = node4 = win4 = nullptr;
  est = nullptr;
  ttot = nullptr;
  canon = nullptr;
  cap_box4 = cap_node4 = cap_win4 = 0;
  if (host_ctrl) (void)hipHostFree(host_ctrl);
  front[0] = front[1] = pieces = nullptr;
  status = nullptr;
  ctrl = nullptr;
  host_ctrl = nullptr;
  safe = nullptr;
  plan = nullptr;
  order = nullptr;
  cap_front = cap_pieces = status_words = cap_safe = cap_plan = cap_order = 0;
  needs_clear = true;
}

QTuning& qtuning() {
  static QTuning t;
  return t;
}

int next_pow2_group(int F) {
  int g = 2;
  while (g < F) g <<= 1;
  return g;
}

int per_tile(int G, int steps) { return (kQThreads / kWave) * steps * (kWave / G); }

__global__ void k_set_u32(uint32_t* p, uint32_t v) {
  if (threadIdx.x == 0) *p = v;
}

// Nodes [start, start+count) of one level must point at F consecutive
// children starting at child0 + i*F (the last one fewer: nchild in all).
__global__ __launch_bounds__(256) void k_check_level(int64_t start, int64_t count, int64_t child0,
                                                     int64_t nchild, int F, const int* pos,
                                                     const int* len, uint32_t* flag) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= count) return;
  const int64_t f = i * F;
  const int64_t want_len = nchild - f < F ? nchild - f : F;
  if ((int64_t)pos[start + i] != child0 + f || (int64_t)len[start + i] != want_len) *flag = 0u;
}

int check_layout(uint32_t** flag, const curtree_rtree_f32* rt, int64_t nbox, hipStream_t s) {
  if (!*flag && hipMalloc((void**)flag, 256) != hipSuccess) {
    (void)hipGetLastError();
    *flag = nullptr;
    return set_error(CURTREE_ENOMEM, "query: cannot allocate a flag");
  }
  LevelPlan lp;
  level_plan(nbox, rt->fanout, &lp);
  hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, s, *flag, lp.nl > 0 ? 1u : 0u);
  int64_t start = lp.r_sz;
  for (int l = 0; l < lp.nl; l++) {
    start -= lp.n[l];
    const int64_t child0 = l == 0 ? lp.r_sz : start + lp.n[l];
    const int64_t nchild = l == 0 ? nbox : lp.n[l - 1];
    hipLaunchKernelGGL(k_check_level, dim3((unsigned)((lp.n[l] + 255) / 256)), dim3(256), 0, s,
                       start, lp.n[l], child0, nchild, rt->fanout, (const int*)rt->pos,
                       (const int*)rt->len, *flag);
  }
  CT_HIP(hipGetLastError());
  return CURTREE_OK;
}

__global__ __launch_bounds__(256) void k_stage4(int64_t n, QBoxes src, float4* __restrict__ dst) {
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    dst[i] = make_float4(src.x0[i], src.y0[i], src.x1[i], src.y1[i]);
}

int reserve_f4(float4** dst, int64_t* cap, int64_t n) {
  if (n > *cap) {
    if (*dst) (void)hipFree(*dst);
    *dst = nullptr;
    *cap = 0;
    const int64_t want = n + n / 8 + 1024;
    if (hipMalloc((void**)dst, (size_t)want * sizeof(float4)) != hipSuccess) {
      (void)hipGetLastError();
      *dst = nullptr;
      return set_error(CURTREE_ENOMEM, "query: cannot stage %lld boxes", (long long)n);
    }
    *cap = want;
  }
  return CURTREE_OK;
}

// Where the cells of child slot sl (of G slots per parent) are stored: slots
// are transposed in groups of four, so that the 16 bytes at 16 * g hold slots
// g, g + G/4, g + 2G/4, g + 3G/4 (one load per lane of a coded kernel, whose
// lane g then finds slot j * G/4 + g in component j).
__device__ __forceinline__ int cell_slot(int sl, int G) {
  const int gc = G >> 2;
  return (sl % gc) * 4 + sl / gc;
}

// Frame of a leaf-parent with MBR (a1, b1, a2, b2).  A frame needs a finite
// low corner and a finite, positive extent on both axes with finite scales;
// otherwise its scales are 0 and the leaf count tests that node's boxes
// directly.
__device__ __forceinline__ float4 leaf_frame(float a1, float b1, float a2, float b2) {
  const float ex = a2 - a1, ey = b2 - b1;
  if (isfinite(a1) && isfinite(b1) && isfinite(ex) && isfinite(ey) && ex > 0.f && ey > 0.f) {
    const float sx = kCells / ex, sy = kCells / ey;
    if (isfinite(sx) && isfinite(sy)) return make_float4(a1, b1, sx, sy);
  }
  return make_float4(a1, b1, 0.f, 0.f);
}

// Leaf codes of a bulk-load tree: box i is coded in the frame of its
// leaf-parent (node ls + i / F) into slot i % F of its G >= F slots (stored at
// cell_slot); the first box of each leaf-parent also writes the frame.  vec
// (F == G, a multiple of 4, box arrays 16-byte aligned): a thread codes four
// consecutive boxes, read with 16-byte loads; otherwise one box per thread.
__global__ __launch_bounds__(256) void k_leaf_codes(int64_t nbox, int F, int G, bool vec,
                                                    int64_t ls,
                                                    const float* __restrict__ nx1,
                                                    const float* __restrict__ ny1,
                                                    const float* __restrict__ nx2,
                                                    const float* __restrict__ ny2, QBoxes rb,
                                                    float4* __restrict__ frame,
                                                    uint32_t* __restrict__ code) {
  const int64_t step = (int64_t)gridDim.x * 256;
  if (vec) {
    // whole groups of four (a leaf-parent holds F = G boxes, so a group never
    // straddles two); the last nbox % 4 boxes below
    const int64_t n4 = nbox >> 2;
    for (int64_t j = blockIdx.x * 256ll + threadIdx.x; j < n4; j += step) {
      const int64_t i = 4 * j, lp = i / F;
      const float4 fr = leaf_frame(nx1[ls + lp], ny1[ls + lp], nx2[ls + lp], ny2[ls + lp]);
      if (i == lp * F) frame[lp] = fr;
      const float4 x1 = ((const float4*)rb.x0)[j], y1 = ((const float4*)rb.y0)[j];
      const float4 x2 = ((const float4*)rb.x1)[j], y2 = ((const float4*)rb.y1)[j];
      uint32_t* cp = code + lp * G;
      const int s0 = (int)(i - lp * F);
      cp[cell_slot(s0, G)] = box_cells(x1.x, y1.x, x2.x, y2.x, fr);
      cp[cell_slot(s0 + 1, G)] = box_cells(x1.y, y1.y, x2.y, y2.y, fr);
      cp[cell_slot(s0 + 2, G)] = box_cells(x1.z, y1.z, x2.z, y2.z, fr);
      cp[cell_slot(s0 + 3, G)] = box_cells(x1.w, y1.w, x2.w, y2.w, fr);
    }
    for (int64_t i = 4 * n4 + blockIdx.x * 256ll + threadIdx.x; i < nbox; i += step) {
      const int64_t lp = i / F;
      const float4 fr = leaf_frame(nx1[ls + lp], ny1[ls + lp], nx2[ls + lp], ny2[ls + lp]);
      if (i == lp * F) frame[lp] = fr;
      code[lp * G + cell_slot((int)(i - lp * F), G)] =
          box_cells(rb.x0[i], rb.y0[i], rb.x1[i], rb.y1[i], fr);
    }
    return;
  }
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < nbox; i += step) {
    const int64_t lp = i / F;
    const float4 fr = leaf_frame(nx1[ls + lp], ny1[ls + lp], nx2[ls + lp], ny2[ls + lp]);
    if (i == lp * F) frame[lp] = fr;
    code[lp * G + cell_slot((int)(i - lp * F), G)] =
        box_cells(rb.x0[i], rb.y0[i], rb.x1[i], rb.y1[i], fr);
  }
}

// Node codes of one internal level: child j (node cs + j) is coded in the
// frame of its parent (node ls + j / F, its MBR) into slot j % F of the
// parent's G slots (stored at cell_slot); the first child of each parent
// writes the parent's frame at the parent's node index.
__global__ __launch_bounds__(256) void k_node_codes(int64_t cc, int F, int G, int64_t ls,
                                                    int64_t cs, QBoxes nb,
                                                    float4* __restrict__ frame,
                                                    uint32_t* __restrict__ code) {
  for (int64_t j = blockIdx.x * 256ll + threadIdx.x; j < cc; j += (int64_t)gridDim.x * 256) {
    const int64_t par = ls + j / F, sl = j % F;
    const float4 fr = leaf_frame(nb.x0[par], nb.y0[par], nb.x1[par], nb.y1[par]);
    if (sl == 0) frame[par] = fr;
    const int64_t ch = cs + j;
    code[par * G + cell_slot((int)sl, G)] = box_cells(nb.x0[ch], nb.y0[ch], nb.x1[ch], nb.y1[ch], fr);
  }
}

int64_t leaf_codes_bytes(int64_t nleafparents, int64_t nparents, int G) {
  return (nleafparents + nparents) * ((int64_t)sizeof(float4) + 4 * (int64_t)G);
}

void enqueue_leaf_codes(const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox,
                        const QLevels& lv, int G, void* buf, hipStream_t s) {
  const int64_t n = rbox->sz;
  if (n <= 0) return;
  const int64_t nlp = lv.lcount[0], ls = lv.lstart[0];
  // the internal levels' node codes, after the leaf codes
  {
    const int64_t np = ls;  // nodes above the leaf-parent level: 0 .. ls-1
    float4* nf = (float4*)((char*)buf + nlp * ((int64_t)sizeof(float4) + 4 * (int64_t)G));
    uint32_t* nc = (uint32_t*)(nf + np);
    const QBoxes nb{rt->xmin, rt->ymin, rt->xmax, rt->ymax, nullptr};
    for (int l = lv.nl - 1; l >= 1; l--) {
      const int64_t cc = lv.lcount[l - 1];
      const unsigned grid = (unsigned)std::min<int64_t>((cc + 255) / 256, 8192);
      hipLaunchKernelGGL(k_node_codes, dim3(grid), dim3(256), 0, s, cc, rt->fanout, G,
                         lv.lstart[l], lv.lstart[l - 1], nb, nf, nc);
    }
  }
  // (16-byte loads of the boxes need 16-byte aligned arrays: else one box per thread)
  const bool al = ((uintptr_t)rbox->xmin | (uintptr_t)rbox->ymin | (uintptr_t)rbox->xmax |
                   (uintptr_t)rbox->ymax) % 16 == 0;
  const bool vec = al && rt->fanout == G && (G & 3) == 0;
  const unsigned grid = (unsigned)std::min<int64_t>((n / (vec ? 4 : 1) + 255) / 256 + 1, 16384);
  hipLaunchKernelGGL(k_leaf_codes, dim3(grid), dim3(256), 0, s, n, rt->fanout, G, vec, ls,
                     rt->xmin,
                     rt->ymin, rt->xmax, rt->ymax,
                     QBoxes{rbox->xmin, rbox->ymin, rbox->xmax, rbox->ymax, nullptr},
                     (float4*)buf, (uint32_t*)((float4*)buf + nlp));
}

int stage_boxes(float4** dst, int64_t* cap, const QBoxes& src, int64_t n, hipStream_t s) {
  int rc;
  if ((rc = reserve_f4(dst, cap, n))) return rc;
  if (n > 0) {
    const unsigned grid = (unsigned)std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_stage4, dim3(grid), dim3(256), 0, s, n, src, *dst);
  }
  CT_HIP(hipGetLastError());
  return CURTREE_OK;
}

int compute_safe(QEngine& qe, Prof& pf, const QBoxes& rb, int F, int64_t r_sz, int64_t nbox,
                 hipStream_t s, float4* box4, bool* staged) {
  int rc;
  if (staged) *staged = false;
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
                       0, s, nbox, rb, safe + start, box4);                               \
    if (staged) *staged = box4 != nullptr;                                                \
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

// Runs of at least this many queries expand without a tile chain.
constexpr int64_t kSplitMinQueries = 32768;
// a query of at least this many windows builds the leaf codes of a kept tree
constexpr int64_t kCodesMinWindows = 65536;

// Blocks of `kernel` (kQThreads threads) that the device holds at once.  The
// chain-free count and write kernels stride over their tiles with a fixed
// step, so they are launched with no more blocks than can run together (a
// block that starts late would add its whole share of tiles to the end of
// the launch).
static unsigned resident_grid(const void* kernel) {
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, kQThreads, 0) != hipSuccess ||
      cus <= 0 || per <= 0) {
    (void)hipGetLastError();
    return 1024;
  }
  return (unsigned)(cus * per);
}

template <int G, int STEPS, bool STORE, bool QCOUNTS>
static unsigned leaf_count_grid() {
  static const unsigned g = resident_grid((const void*)k_leaf_count<G, STEPS, STORE, QCOUNTS>);
  return g;
}

template <int G, int STEPS, bool STORE, bool QCOUNTS>
static unsigned leaf_count_coded_grid() {
  static const unsigned g =
      resident_grid((const void*)k_leaf_count_coded<G, STEPS, STORE, QCOUNTS>);
  return g;
}

// The leaf count: over the leaf codes when the run has them (G >= 4; tiles
// of up to 8 wave-steps, two rounds of coded loads per tile), else the box
// tests.
template <int G, int STEPS, bool STORE, bool QCOUNTS>
static void launch_leaf_count(const QRunArgs& a, int leaf, const uint2* cur, QCtrl* ctrl,
                              QLeaf* est, uint64_t* ttot, unsigned long long* counts, bool xcd,
                              hipStream_t s) {
  if constexpr (G >= 4 && STEPS <= 8) {
    if (a.tr.lcode) {
      hipLaunchKernelGGL((k_leaf_count_coded<G, STEPS, STORE, QCOUNTS>),
                         dim3(std::min<unsigned>(a.grid,
                                                 leaf_count_coded_grid<G, STEPS, STORE, QCOUNTS>())),
                         dim3(kQThreads), 0, s, leaf, a.tr, a.qb, a.rb, ctrl, cur, a.r_sz, a.nbox,
                         est, ttot, counts, xcd);
      return;
    }
  }
  hipLaunchKernelGGL((k_leaf_count<G, STEPS, STORE, QCOUNTS>),
                     dim3(std::min<unsigned>(a.grid, leaf_count_grid<G, STEPS, STORE, QCOUNTS>())),
                     dim3(kQThreads), 0, s, leaf, a.tr, a.qb, a.rb, ctrl, cur, a.r_sz, a.nbox, est,
                     ttot, counts, xcd);
}

template <int G, int STEPS>
static unsigned expand_count_coded_grid() {
  static const unsigned g = resident_grid((const void*)k_expand_count_coded<G, STEPS>);
  return g;
}

template <int G, int STEPS>
static unsigned expand_count_grid();

// An internal level's count: over the node codes when the run has them (G >= 4,
// tiles of up to 8 wave-steps), else the MBR tests.
template <int G, int STEPS>
static void launch_expand_count(const QRunArgs& a, int l, const uint2* cur, QEngine& qe, bool xcd,
                                hipStream_t s) {
  if constexpr (G >= 4 && STEPS <= 8) {
    if (a.tr.ncode) {
      hipLaunchKernelGGL((k_expand_count_coded<G, STEPS>),
                         dim3(std::min<unsigned>(a.grid, expand_count_coded_grid<G, STEPS>())),
                         dim3(kQThreads), 0, s, l, a.tr, a.qb, qe.ctrl, cur, (QExp*)qe.est,
                         qe.ttot, a.r_sz, xcd);
      return;
    }
  }
  hipLaunchKernelGGL((k_expand_count<G, STEPS>),
                     dim3(std::min<unsigned>(a.grid, expand_count_grid<G, STEPS>())),
                     dim3(kQThreads), 0, s, l, a.tr, a.qb, qe.ctrl, cur, (QExp*)qe.est, qe.ttot,
                     a.r_sz, xcd);
}

template <int G, int STEPS>
static unsigned expand_count_grid() {
  static const unsigned g = resident_grid((const void*)k_expand_count<G, STEPS>);
  return g;
}

template <int G, int STEPS>
static unsigned expand_write_grid() {
  static const unsigned g = resident_grid((const void*)k_expand_write<G, STEPS>);
  return g;
}

template <int G, int STEPS>
static unsigned leaf_write_grid() {
  static const unsigned g = resident_grid((const void*)k_leaf_write<G, STEPS>);
  return g;
}

template <int G, int STEPS>
static void warm_coded() {
  if constexpr (G >= 4 && STEPS <= 8) {
    (void)expand_count_coded_grid<G, STEPS>();
    (void)leaf_count_coded_grid<G, STEPS, true, false>();
    (void)leaf_count_coded_grid<G, STEPS, true, true>();
    (void)leaf_count_coded_grid<G, STEPS, false, false>();
    (void)leaf_count_coded_grid<G, STEPS, false, true>();
  }
}

// Queries the occupancy of the count kernels for (G, steps) ahead of any
// graph capture (the first query of each is cached).
template <int G>
static void warm_grids_g(int steps) {
  switch (steps) {
#define CT_WARM(st)                               \
  case st:                                        \
    (void)expand_count_grid<G, st>();             \
    (void)expand_write_grid<G, st>();             \
    (void)leaf_count_grid<G, st, true, false>();  \
    (void)leaf_count_grid<G, st, true, true>();   \
    (void)leaf_count_grid<G, st, false, false>(); \
    (void)leaf_count_grid<G, st, false, true>();  \
    warm_coded<G, st>();                          \
    (void)leaf_write_grid<G, st>();               \
    break;
    CT_WARM(4) CT_WARM(16)
    default: CT_WARM(8)
#undef CT_WARM
  }
}

static void warm_grids(int G, int steps) {
  switch (G) {
    case 2: warm_grids_g<2>(steps); break;
    case 4: warm_grids_g<4>(steps); break;
    case 8: warm_grids_g<8>(steps); break;
    case 16: warm_grids_g<16>(steps); break;
    case 32: warm_grids_g<32>(steps); break;
    default: warm_grids_g<64>(steps); break;
  }
}

// (window order kernels, below)
__global__ void k_ord_sums(int64_t nq, const unsigned long long* __restrict__ cnt,
                           const uint32_t* __restrict__ inv, int64_t* __restrict__ sums);
__global__ void k_ord_top(int64_t nb, int64_t* __restrict__ sums);
__global__ void k_ord_apply(int64_t nq, const unsigned long long* __restrict__ cnt,
                            const uint32_t* __restrict__ inv, const int64_t* __restrict__ sums,
                            int64_t* __restrict__ ooff, int64_t* __restrict__ poff);

template <int G, int STEPS>
void enqueue_run_g(QEngine& qe, Prof& pf, const QRunArgs& a, hipStream_t s) {
  using S = QShape<G, STEPS>;
  // slots: expansions 0..nl-2, the leaf stage nl-1 and its tile scan nl, the
  // expansions' tile scans nl+1..2nl-1
  const unsigned nslots = 2u * (unsigned)a.nl;
  const int ex = qtuning().expand;
  const bool small = a.plan == nullptr && a.q1 - a.q0 < kSplitMinQueries;
  // the persistent algorithm: one chained launch per level, leaf stage included
  const bool split = !a.chained && (ex == 1 || (ex < 0 && !small));
  const int fz = qtuning().fused;
  const bool fused = a.chained || fz == 1 || (fz < 0 && small);
  const bool xcd = qtuning().xcd == 1;
  if (qe.needs_clear) {
    hipLaunchKernelGGL(k_ctrl_init, dim3(1), dim3(64), 0, s, qe.ctrl);
    qe.needs_clear = false;
  }
  pf.mark(s, "k_run_begin");
  hipLaunchKernelGGL(k_run_begin, dim3(1), dim3(64), 0, s, qe.ctrl, a.q0, a.q1, a.plan, nslots,
                     a.lv);
  hipLaunchKernelGGL(k_status_zero, dim3(256), dim3(256), 0, s, (const QCtrl*)qe.ctrl, qe.status,
                     qe.status_words);
  const uint2* cur = nullptr;
  int which = 0;
  for (int l = 0; l + 1 < a.nl; l++) {
    uint2* out = (uint2*)qe.front[which];
    if (split) {
      pf.mark(s, "k_expand_count");
      launch_expand_count<G, STEPS>(a, l, cur, qe, xcd, s);
      pf.mark(s, "k_tile_scan");
      hipLaunchKernelGGL(k_tile_scan, dim3(64), dim3(kQThreads), 0, s, a.nl + 1 + l, l,
                         S::kPerTile, qe.ctrl, qe.ttot, qe.status);
      pf.mark(s, "k_expand_write");
      hipLaunchKernelGGL((k_expand_write<G, STEPS>),
                         dim3(std::min<unsigned>(a.grid, expand_write_grid<G, STEPS>())),
                         dim3(kQThreads), 0, s, l,
                         qe.ctrl, (const QExp*)qe.est, (const uint64_t*)qe.ttot, out,
                         qe.cap_front, xcd);
    } else {
      pf.mark(s, "k_expand");
      hipLaunchKernelGGL((k_expand<G, STEPS>), dim3(a.grid), dim3(kQThreads), 0, s, l, a.tr,
                         a.qb, qe.ctrl, cur, out, qe.cap_front, qe.status, a.r_sz);
    }
    cur = out;
    which ^= 1;
  }
  const int leaf = a.nl - 1;
  if (a.write && fused) {
    pf.mark(s, "k_leaf_fused");
    if (a.counts)
      hipLaunchKernelGGL((k_leaf_fused<G, STEPS, true>), dim3(a.grid), dim3(kQThreads), 0, s,
                         leaf, a.tr, a.qb, a.rb, qe.ctrl, cur, a.r_sz, a.nbox, a.rid, a.out_fid,
                         a.out_tid, a.out_cap, qe.status, (QPiece*)qe.pieces, qe.cap_pieces,
                         a.counts);
    else
      hipLaunchKernelGGL((k_leaf_fused<G, STEPS, false>), dim3(a.grid), dim3(kQThreads), 0, s,
                         leaf, a.tr, a.qb, a.rb, qe.ctrl, cur, a.r_sz, a.nbox, a.rid, a.out_fid,
                         a.out_tid, a.out_cap, qe.status, (QPiece*)qe.pieces, qe.cap_pieces,
                         (unsigned long long*)nullptr);
    pf.mark(s, "k_write_pieces");
    hipLaunchKernelGGL(k_write_pieces, dim3(4096), dim3(256), 0, s, (const QCtrl*)qe.ctrl,
                       (const QPiece*)qe.pieces, qe.cap_pieces, a.rid, a.out_fid, a.out_tid,
                       a.out_cap);
  } else if (a.write) {
    pf.mark(s, "k_leaf_count");
    if (a.counts)
      launch_leaf_count<G, STEPS, true, true>(a, leaf, cur, qe.ctrl, (QLeaf*)qe.est, qe.ttot,
                                              a.counts, xcd, s);
    else
      launch_leaf_count<G, STEPS, true, false>(a, leaf, cur, qe.ctrl, (QLeaf*)qe.est, qe.ttot,
                                               (unsigned long long*)nullptr, xcd, s);
    pf.mark(s, "k_tile_scan");
    hipLaunchKernelGGL(k_tile_scan, dim3(64), dim3(kQThreads), 0, s, leaf + 1, leaf, S::kPerTile,
                       qe.ctrl, qe.ttot, qe.status);
    if (a.tmap) {  // windows in a run order: their first pairs in both orders
      const int64_t nq = a.q1 - a.q0, nb = (nq + kOrdBlock - 1) / kOrdBlock;
      const uint32_t* inv = a.tmap + nq;  // the rank of each input window
      pf.mark(s, "k_win_offsets");
      hipLaunchKernelGGL(k_ord_sums, dim3((unsigned)nb), dim3(256), 0, s, nq,
                         (const unsigned long long*)a.counts, inv, a.osum);
      hipLaunchKernelGGL(k_ord_top, dim3(1), dim3(256), 0, s, nb, a.osum);
      hipLaunchKernelGGL(k_ord_apply, dim3((unsigned)nb), dim3(256), 0, s, nq,
                         (const unsigned long long*)a.counts, inv, (const int64_t*)a.osum,
                         a.ooff, a.poff);
    }
    pf.mark(s, "k_leaf_write");
    hipLaunchKernelGGL((k_leaf_write<G, STEPS>),
                       dim3(std::min<unsigned>(a.grid, leaf_write_grid<G, STEPS>())),
                       dim3(kQThreads), 0, s, leaf,
                       qe.ctrl, (const QLeaf*)qe.est, (const uint64_t*)qe.ttot, a.rid, a.out_fid,
                       a.out_tid, a.out_cap, (QPiece*)qe.pieces, qe.cap_pieces, xcd, a.tmap,
                       a.ooff, a.poff);
    pf.mark(s, "k_write_pieces");
    hipLaunchKernelGGL(k_write_pieces, dim3(4096), dim3(256), 0, s, (const QCtrl*)qe.ctrl,
                       (const QPiece*)qe.pieces, qe.cap_pieces, a.rid, a.out_fid, a.out_tid,
                       a.out_cap);
  } else {
    pf.mark(s, "k_leaf_count");
    if (a.counts)
      launch_leaf_count<G, STEPS, false, true>(a, leaf, cur, qe.ctrl, (QLeaf*)nullptr,
                                               (uint64_t*)nullptr, a.counts, xcd, s);
    else
      launch_leaf_count<G, STEPS, false, false>(a, leaf, cur, qe.ctrl, (QLeaf*)nullptr,
                                                (uint64_t*)nullptr, (unsigned long long*)nullptr,
                                                xcd, s);
  }
  pf.mark(s, nullptr);
}

template <int G>
static void enqueue_run_steps(QEngine& qe, Prof& pf, const QRunArgs& a, hipStream_t s) {
  switch (a.steps) {
    case 4: enqueue_run_g<G, 4>(qe, pf, a, s); break;
    case 16: enqueue_run_g<G, 16>(qe, pf, a, s); break;
    default: enqueue_run_g<G, 8>(qe, pf, a, s); break;
  }
}

// ---------------------------------------------------------------------------
// Trees of fanout above 64 (one lane per child does not fit a wave): one wave
// per frontier entry visits the entry's children 64 at a time.  Per level:
//   k_gen_count  the children that intersect the entry's window (internal
//                levels: child nodes, the leaf level: boxes), counted into
//                cnt[entry]; the level's total by one atomic per wave, and
//                per-query counts at the leaf level when asked;
//   k_tile_scan  the counts, scanned in place (one count per "tile");
//   k_gen_write  the same children again, stored compacted at the entry's
//                offset: (query, child) entries of the next frontier, or
//                (box id, query) pairs.
// Output order is the entries' order, then child order, as on the other
// paths; there is no containment shortcut here.
template <bool LEAF>
__device__ __forceinline__ void gen_children(const QTree& tr, const QCtrl* c, int slot,
                                             uint32_t node, int64_t r_sz, int64_t nbox,
                                             bool canon, int64_t* first, int64_t* len) {
  const int lvl = LEAF ? 0 : c->lv.nl - 1 - slot;
  const int64_t ls = c->lv.lstart[lvl];
  if (canon) {
    const int64_t f = ((int64_t)node - ls) * tr.F;
    const int64_t cnt = LEAF ? nbox : c->lv.lcount[lvl - 1];
    *first = (LEAF ? 0 : c->lv.lstart[lvl - 1]) + f;
    *len = cnt - f < tr.F ? cnt - f : tr.F;
  } else {
    *first = LEAF ? (int64_t)tr.pos[node] - r_sz : (int64_t)tr.pos[node];
    *len = tr.len[node];
  }
  // children must be boxes (leaf level) or nodes (a corrupt pos/len is never
  // followed)
  const int64_t lim = LEAF ? nbox : r_sz;
  if (*first < 0 || *len < 0 || *first + *len > lim) *len = 0;
}

template <bool LEAF>
__global__ __launch_bounds__(256) void k_gen_count(int slot, QTree tr, QBoxes qb, QBoxes rb,
                                                   QCtrl* c, const uint2* __restrict__ fin,
                                                   int64_t r_sz, int64_t nbox,
                                                   uint64_t* __restrict__ cnt,
                                                   unsigned long long* counts) {
  if (c->overflow) return;
  const int64_t n_in = (int64_t)c->count[slot];
  const int64_t q0 = c->q0;
  const int lane = lane_id();
  const bool canon = tr.canon && *tr.canon;
  const QBoxes& kids = LEAF ? rb : tr.mbr;
  uint64_t wsum = 0;
  for (int64_t e = blockIdx.x * 4ll + (threadIdx.x >> 6); e < n_in;
       e += (int64_t)gridDim.x * 4) {
    const uint2 en = fin ? fin[e] : make_uint2((uint32_t)(q0 + e), 0u);
    int64_t first, L;
    gen_children<LEAF>(tr, c, slot, en.y & ~kFull, r_sz, nbox, canon, &first, &L);
    const float4 wq = ld_box(qb, en.x);
    uint64_t hits = 0;
    for (int64_t j = 0; j < L; j += kWave) {
      const bool ok = j + lane < L;
      const float4 b = ld_box(kids, ok ? first + j + lane : 0);
      hits += (uint64_t)popc64(ballot(ok && intersect(wq.x, wq.y, wq.z, wq.w, b.x, b.y, b.z, b.w)));
    }
    if (lane == 0) {
      cnt[e] = hits;
      wsum += hits;
      if (LEAF && counts && hits) atomicAdd(counts + en.x, (unsigned long long)hits);
    }
  }
  if (lane == 0 && wsum)
    atomicAdd((unsigned long long*)(LEAF ? &c->total : &c->count[slot + 1]),
              (unsigned long long)wsum);
}

template <bool LEAF>
__global__ __launch_bounds__(256) void k_gen_write(int slot, QTree tr, QBoxes qb, QBoxes rb,
                                                   QCtrl* c, const uint2* __restrict__ fin,
                                                   int64_t r_sz, int64_t nbox,
                                                   const uint64_t* __restrict__ off,
                                                   uint2* __restrict__ fout, int64_t cap,
                                                   const int* __restrict__ rid,
                                                   int* __restrict__ out_fid,
                                                   int* __restrict__ out_tid, int64_t out_cap) {
  if (c->overflow) return;
  if (!LEAF && (int64_t)c->count[slot + 1] > cap) {  // the next frontier does not fit
    if (blockIdx.x == 0 && threadIdx.x == 0) c->overflow = 1;
    return;
  }
  const int64_t n_in = (int64_t)c->count[slot];
  const int64_t q0 = c->q0;
  const int lane = lane_id();
  const uint64_t lt = lanemask_lt();
  const bool canon = tr.canon && *tr.canon;
  const QBoxes& kids = LEAF ? rb : tr.mbr;
  for (int64_t e = blockIdx.x * 4ll + (threadIdx.x >> 6); e < n_in;
       e += (int64_t)gridDim.x * 4) {
    const uint2 en = fin ? fin[e] : make_uint2((uint32_t)(q0 + e), 0u);
    int64_t first, L;
    gen_children<LEAF>(tr, c, slot, en.y & ~kFull, r_sz, nbox, canon, &first, &L);
    const float4 wq = ld_box(qb, en.x);
    uint64_t o = off[e];
    for (int64_t j = 0; j < L; j += kWave) {
      const bool ok = j + lane < L;
      const int64_t ci = ok ? first + j + lane : 0;
      const float4 b = ld_box(kids, ci);
      const bool hit = ok && intersect(wq.x, wq.y, wq.z, wq.w, b.x, b.y, b.z, b.w);
      const uint64_t m = ballot(hit);
      if (hit) {
        const uint64_t slot_o = o + (uint64_t)popc64(m & lt);
        if (LEAF) {
          if (slot_o < (uint64_t)out_cap) {
            out_fid[slot_o] = rid[ci];
            out_tid[slot_o] = (int)en.x;
          }
        } else if (slot_o < (uint64_t)cap) {
          fout[slot_o] = make_uint2(en.x, (uint32_t)ci);
        }
      }
      o += (uint64_t)popc64(m);
    }
  }
}

static void enqueue_run_generic(QEngine& qe, Prof& pf, const QRunArgs& a, hipStream_t s) {
  const unsigned nslots = 2u * (unsigned)a.nl;
  if (qe.needs_clear) {
    hipLaunchKernelGGL(k_ctrl_init, dim3(1), dim3(64), 0, s, qe.ctrl);
    qe.needs_clear = false;
  }
  pf.mark(s, "k_run_begin");
  hipLaunchKernelGGL(k_run_begin, dim3(1), dim3(64), 0, s, qe.ctrl, a.q0, a.q1, a.plan, nslots,
                     a.lv);
  hipLaunchKernelGGL(k_status_zero, dim3(256), dim3(256), 0, s, (const QCtrl*)qe.ctrl, qe.status,
                     qe.status_words);
  const dim3 grid((unsigned)std::max(a.grid, 1024));
  const uint2* cur = nullptr;
  int which = 0;
  for (int l = 0; l + 1 < a.nl; l++) {
    uint2* out = (uint2*)qe.front[which];
    pf.mark(s, "k_gen_count");
    hipLaunchKernelGGL(k_gen_count<false>, grid, dim3(256), 0, s, l, a.tr, a.qb, a.rb, qe.ctrl,
                       cur, a.r_sz, a.nbox, qe.ttot, (unsigned long long*)nullptr);
    pf.mark(s, "k_tile_scan");
    hipLaunchKernelGGL(k_tile_scan, dim3(64), dim3(kQThreads), 0, s, a.nl + 1 + l, l, 1, qe.ctrl,
                       qe.ttot, qe.status);
    pf.mark(s, "k_gen_write");
    hipLaunchKernelGGL(k_gen_write<false>, grid, dim3(256), 0, s, l, a.tr, a.qb, a.rb, qe.ctrl,
                       cur, a.r_sz, a.nbox, (const uint64_t*)qe.ttot, out, qe.cap_front,
                       (const int*)nullptr, (int*)nullptr, (int*)nullptr, (int64_t)0);
    cur = out;
    which ^= 1;
  }
  const int leaf = a.nl - 1;
  pf.mark(s, "k_gen_count");
  hipLaunchKernelGGL(k_gen_count<true>, grid, dim3(256), 0, s, leaf, a.tr, a.qb, a.rb, qe.ctrl,
                     cur, a.r_sz, a.nbox, qe.ttot, a.counts);
  if (a.write) {
    pf.mark(s, "k_tile_scan");
    hipLaunchKernelGGL(k_tile_scan, dim3(64), dim3(kQThreads), 0, s, leaf + 1, leaf, 1, qe.ctrl,
                       qe.ttot, qe.status);
    pf.mark(s, "k_gen_write");
    hipLaunchKernelGGL(k_gen_write<true>, grid, dim3(256), 0, s, leaf, a.tr, a.qb, a.rb, qe.ctrl,
                       cur, a.r_sz, a.nbox, (const uint64_t*)qe.ttot, (uint2*)nullptr,
                       (int64_t)0, a.rid, a.out_fid, a.out_tid, a.out_cap);
  }
  pf.mark(s, nullptr);
}

void enqueue_run(QEngine& qe, Prof& pf, const QRunArgs& a, hipStream_t s) {
  if (a.generic) {
    enqueue_run_generic(qe, pf, a, s);
    return;
  }
  switch (a.G) {
    case 2: enqueue_run_steps<2>(qe, pf, a, s); break;
    case 4: enqueue_run_steps<4>(qe, pf, a, s); break;
    case 8: enqueue_run_steps<8>(qe, pf, a, s); break;
    case 16: enqueue_run_steps<16>(qe, pf, a, s); break;
    case 32: enqueue_run_steps<32>(qe, pf, a, s); break;
    default: enqueue_run_steps<64>(qe, pf, a, s); break;
  }
}

int64_t pieces_needed(const QEngine& qe, const QRunArgs& a) {
  // contained ranges exist only with the containment shortcut (node flags or
  // the tree's well-formed word)
  if (!a.tr.safe && !a.tr.allwf) return 0;
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

// The leaf codes of a kept tree (nlp leaf-parents from node ls, G slots
// each): the ready buffer, or null when the tree is not kept, or none are
// ready and make is false.  make: built now on stream s (a synchronisation).
static const void* leaf_codes_for(Prof& pf, const curtree_rtree_f32* rt,
                                  const curtree_bbox_f32* rbox, int G, const QLevels& lv,
                                  bool make, hipStream_t s) {
  const std::function<int(void*)> build = [&](void* buf) {
    pf.mark(s, "k_leaf_codes");
    enqueue_leaf_codes(rt, rbox, lv, G, buf, s);
    pf.mark(s, nullptr);
    CT_HIP(hipGetLastError());
    CT_HIP(hipStreamSynchronize(s));
    return CURTREE_OK;
  };
  return tree_state_codes(rt, rbox, leaf_codes_bytes(lv.lcount[0], lv.lstart[0], G),
                          make ? &build : nullptr);
}

// The level table of a tree over nbox boxes (root level last in lstart order:
// level l starts at lstart[l], level 0 being the leaf-parents).
static void level_table(int F, const LevelPlan& lp, QLevels* lv) {
  lv->nl = lp.nl;
  int64_t start = lp.r_sz, span = 1;
  for (int l = 0; l < lp.nl; l++) {
    start -= lp.n[l];
    span = span > ((int64_t)1 << 40) / F ? ((int64_t)1 << 40) : span * F;
    lv->lstart[l] = start;
    lv->lcount[l] = lp.n[l];
    lv->span[l] = span;
  }
}

int ensure_leaf_codes(const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox, hipStream_t s) {
  const int64_t N = rbox->sz;
  const int F = rt->fanout;
  if (N <= 0 || F < 2 || F > 64 || next_pow2_group(F) < 4 || qtuning().codes == 0)
    return CURTREE_OK;
  LevelPlan lp;
  level_plan(N, F, &lp);
  if (lp.r_sz != rt->sz) return CURTREE_OK;
  QLevels lv{};
  level_table(F, lp, &lv);
  Prof pf;  // (never enabled)
  (void)leaf_codes_for(pf, rt, rbox, next_pow2_group(F), lv, true, s);
  return CURTREE_OK;
}

int prepare_query(QEngine& qe, Prof& pf, const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox,
                  const curtree_bbox_f32* qbox, int contain, int64_t nq_hint, QRunArgs* a,
                  hipStream_t s, const uint32_t* canon) {
  if (!rt || !rbox || !qbox) return set_error(CURTREE_EINVAL, "query: null argument");
  const int64_t N = rbox->sz, Q = qbox->sz;
  const int F = rt->fanout;
  if (N < 0 || Q < 0) return set_error(CURTREE_EINVAL, "query: negative size");
  if (F < 2) return set_error(CURTREE_EINVAL, "query: fanout must be >= 2");
  if (F > 64 && qe.fixed)
    return set_error(CURTREE_EINVAL,
                     "query_noalloc: fanout %d above 64 needs per-entry counts beyond the scratch",
                     F);
  if (Q > (int64_t)0xFFFFFFFF) return set_error(CURTREE_EINVAL, "query: too many windows");
  *a = QRunArgs();
  a->steps = qtuning().steps;
  // a block loops over tiles, so any grid is correct; small batches get a
  // small grid (cheaper launches that leave room for concurrent runs), also
  // under a tuned grid (tuned on large batches)
  // (at least 128 blocks: for a request of a few hundred windows most blocks
  // of a chained launch only take a ticket and leave; with 128 rather than
  // 512 the serving workload's p50 at 16k requests/s fell from 0.48 to 0.40 ms
  // with the server and from 0.31 to 0.29 ms at 4.6k with the plain entry point)
  const int64_t by_size = std::max<int64_t>(128, nq_hint / 8);
  a->grid = (int)std::min<int64_t>(qtuning().grid > 0 ? qtuning().grid : kDefaultChainGrid,
                                   by_size);
  a->nbox = N;
  a->F = F;
  a->generic = F > 64;  // one wave per entry, children 64 at a time
  a->G = a->generic ? 64 : next_pow2_group(F);
  warm_grids(a->G, a->steps);
  if (a->generic && !qe.wide) {  // per-entry counts: the buffers are sized again
    qe.wide = true;
    qe.cap_front = 0;
  }
  if (N == 0) return CURTREE_OK;
  LevelPlan lp;
  level_plan(N, F, &lp);
  if (lp.r_sz != rt->sz)
    return set_error(CURTREE_EINVAL,
                     "query: tree size %d does not match %lld boxes at fanout %d (expected %lld)",
                     rt->sz, (long long)N, F, (long long)lp.r_sz);
  if (2 * lp.nl > kMaxRunSlots) return set_error(CURTREE_EINVAL, "query: tree too deep");
  if (!rt->xmin || !rt->ymin || !rt->xmax || !rt->ymax || !rt->pos || !rt->len || !rbox->xmin ||
      !rbox->ymin || !rbox->xmax || !rbox->ymax || !rbox->id ||
      (Q > 0 && (!qbox->xmin || !qbox->ymin || !qbox->xmax || !qbox->ymax)))
    return set_error(CURTREE_EINVAL, "query: null device array");
  int rc;
  if ((rc = qe.ensure_ctrl())) return rc;
  a->tr.mbr = QBoxes{rt->xmin, rt->ymin, rt->xmax, rt->ymax, nullptr};
  a->tr.pos = rt->pos;
  a->tr.len = rt->len;
  a->tr.F = F;
  a->qb = QBoxes{qbox->xmin, qbox->ymin, qbox->xmax, qbox->ymax, nullptr};
  a->rb = QBoxes{rbox->xmin, rbox->ymin, rbox->xmax, rbox->ymax, nullptr};
  a->rid = rbox->id;
  a->r_sz = lp.r_sz;
  a->nl = lp.nl;
  level_table(F, lp, &a->lv);
  // the flags cost one pass over the boxes; worth it for larger batches
  // (a fixed engine holds no tree-sized buffers: no flags, no staging)
  const bool use = !qe.fixed && !a->generic && (contain == 1 || (contain < 0 && nq_hint >= 256));
  // AoS staging (a read and a write of every box) only when asked for: the
  // leaf and expansion kernels read the caller's SoA arrays faster (on the
  // 100M-box workload 4.73 ms against 5.71 ms for the leaf count)
  const int st = qtuning().staging;
  const bool stage = !qe.fixed && st == 1;
  bool staged = false;
  if (stage && (rc = reserve_f4(&qe.box4, &qe.cap_box4, N))) return rc;
  // A tree built by this library over these boxes keeps state with it: its
  // build recorded whether every box is well-formed, so no per-node flags are
  // computed (the layout is still checked: two small launches over pos/len),
  // and its leaf codes, built once by the first large enough query (not under
  // graph capture: the build synchronises), serve every later query.
  const int cz = qtuning().codes;
  const bool codes = !qe.fixed && !a->generic && a->G >= 4 && cz != 0 &&
                     (cz == 1 || nq_hint >= 256);
  const uint32_t* kept = use || codes ? tree_state_find(rt, rbox) : nullptr;
  if (kept && codes) {
    {
      hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
      if (hipStreamIsCapturing(s, &cs) != hipSuccess) {
        (void)hipGetLastError();
        cs = hipStreamCaptureStatusActive;
      }
      const bool make = cs == hipStreamCaptureStatusNone && (cz == 1 || nq_hint >= kCodesMinWindows);
      const int64_t nlp = lp.n[0], np = a->lv.lstart[0];
      const void* cb = leaf_codes_for(pf, rt, rbox, a->G, a->lv, make, s);
      if (cb) {
        a->tr.lframe = (const float4*)cb;
        a->tr.lcode = (const uint32_t*)((const float4*)cb + nlp);
#ifndef CT_NODE_CODES
#define CT_NODE_CODES 1  // internal levels over the node codes
#endif
        if (CT_NODE_CODES) {
          a->tr.nframe = (const float4*)((const char*)cb +
                                         nlp * ((int64_t)sizeof(float4) + 4 * (int64_t)a->G));
          a->tr.ncode = (const uint32_t*)(a->tr.nframe + np);
        }
      }
    }
  }
  if (kept && use) {
    a->tr.allwf = kept;
  } else if (use) {  // the flag pass also writes the AoS boxes when both are wanted
    if ((rc = compute_safe(qe, pf, a->rb, F, lp.r_sz, N, s, stage ? qe.box4 : nullptr, &staged)))
      return rc;
    a->tr.safe = qe.safe;
  }
  if (stage) {
    pf.mark(s, "k_stage4");
    if ((!staged && (rc = stage_boxes(&qe.box4, &qe.cap_box4, a->rb, N, s))) ||
        (rc = stage_boxes(&qe.node4, &qe.cap_node4, a->tr.mbr, lp.r_sz, s)) ||
        (rc = stage_boxes(&qe.win4, &qe.cap_win4, a->qb, Q, s)))
      return rc;
    pf.mark(s, nullptr);
    a->rb.b4 = qe.box4;
    a->tr.mbr.b4 = qe.node4;
    a->qb.b4 = qe.win4;
  }
  if (canon) {  // the caller checked this tree's layout once
    a->tr.canon = canon;
  } else if (use || stage || qe.fixed || a->tr.lcode) {  // (the codes need the bulk-load layout)
    if ((rc = check_layout(&qe.canon, rt, N, s))) return rc;
    a->tr.canon = qe.canon;
  }
  return CURTREE_OK;
}

// ---------------------------------------------------------------------------
// Windows in Morton order of their centres.  In a large batch each leaf-parent
// is tested by many windows (about 14 at 100M boxes and 1M windows); with the
// windows in input order those tests are far apart in time and every one reads
// the leaf-parent's boxes from HBM again.  A run may take its windows in any
// order as long as every window's pairs land where the input order puts them:
// the windows are sorted by the Morton code of their centres (quantised over
// the root MBR) and copied in that order, the run works on ranks (its frontier,
// window loads and per-window counts), and the leaf write moves every entry
// from its window's first pair in run order to the window's first pair in
// input order and writes the input index.  Within a window the pairs keep
// their (sorted box position) order, so the output is the same.
// The price is the write: the pairs of the tiles in flight land far apart in
// the output instead of in one advancing run.  On 100M boxes and 1M windows
// (1.17e9 pairs) the leaf count saves 0.47 ms and the leaf write loses 0.44 ms
// (0.26 ms more for the sort and offsets); on 1B skewed boxes and 10M windows
// (7.9e8 pairs) the query is 0.9 ms faster.  Default: batches of 4M windows
// or more.
constexpr int64_t kOrderMinWindows = int64_t(1) << 22;

// The sort keys (code ^ 0x80000000, so that the signed order of the pair sort
// is the unsigned order of the code) and the window indices.
__global__ __launch_bounds__(256) void k_win_keys(int64_t nq, QBoxes qb, QBoxes root,
                                                  int* __restrict__ key, int* __restrict__ idx) {
  const float rx0 = root.x0[0], ry0 = root.y0[0], rx1 = root.x1[0], ry1 = root.y1[0];
  const float sx = rx1 > rx0 ? 65535.0f / (rx1 - rx0) : 0.0f;
  const float sy = ry1 > ry0 ? 65535.0f / (ry1 - ry0) : 0.0f;
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < nq; i += (int64_t)gridDim.x * 256) {
    const float cx = (qb.x0[i] + qb.x1[i]) * 0.5f, cy = (qb.y0[i] + qb.y1[i]) * 0.5f;
    float fx = (cx - rx0) * sx, fy = (cy - ry0) * sy;
    fx = fx > 0.0f ? (fx < 65535.0f ? fx : 65535.0f) : 0.0f;  // (NaN: 0)
    fy = fy > 0.0f ? (fy < 65535.0f ? fy : 65535.0f) : 0.0f;
    const uint32_t code = (spread16((uint32_t)fy) << 1) | spread16((uint32_t)fx);
    key[i] = (int)(code ^ 0x80000000u);
    idx[i] = (int)i;
  }
}

// The windows copied in run order (win4[r] = window perm[r]) and the rank of
// each input window (inv[perm[r]] = r).
__global__ __launch_bounds__(256) void k_win_stage(int64_t nq, QBoxes qb,
                                                   const uint32_t* __restrict__ perm,
                                                   float4* __restrict__ win4,
                                                   uint32_t* __restrict__ inv) {
  for (int64_t r = blockIdx.x * 256ll + threadIdx.x; r < nq; r += (int64_t)gridDim.x * 256) {
    const uint32_t i = perm[r];
    win4[r] = make_float4(qb.x0[i], qb.y0[i], qb.x1[i], qb.y1[i]);
    inv[i] = (uint32_t)r;
  }
}

// Per block of kOrdBlock windows: their pair counts (by rank) summed in input
// order (sums[b], windows b*kOrdBlock.. of the input) and in run order
// (sums[nb + b], ranks b*kOrdBlock..).
__global__ __launch_bounds__(256) void k_ord_sums(int64_t nq, const unsigned long long* __restrict__ cnt,
                                                  const uint32_t* __restrict__ inv,
                                                  int64_t* __restrict__ sums) {
  __shared__ uint64_t s_a[kQWaves], s_b[kQWaves];
  const int64_t nb = (nq + kOrdBlock - 1) / kOrdBlock;
  const int64_t b = blockIdx.x;
  uint64_t so = 0, sp = 0;
  for (int64_t i = b * kOrdBlock + threadIdx.x; i < nq && i < (b + 1) * kOrdBlock; i += 256) {
    so += cnt[inv[i]];
    sp += cnt[i];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    so += __shfl_xor(so, o);
    sp += __shfl_xor(sp, o);
  }
  if (lane_id() == 0) {
    s_a[threadIdx.x >> 6] = so;
    s_b[threadIdx.x >> 6] = sp;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t x = 0, z = 0;
#pragma unroll
    for (int w = 0; w < kQWaves; w++) {
      x += s_a[w];
      z += s_b[w];
    }
    sums[b] = (int64_t)x;
    sums[nb + b] = (int64_t)z;
  }
}

// Exclusive scans of both rows of block sums in place (one block).
__global__ __launch_bounds__(256) void k_ord_top(int64_t nb, int64_t* __restrict__ sums) {
  __shared__ uint64_t s_scan[kQWaves];
  for (int row = 0; row < 2; row++) {
    int64_t* v = sums + row * nb;
    uint64_t carry = 0;
    for (int64_t r0 = 0; r0 < nb; r0 += 256) {
      const int64_t i = r0 + threadIdx.x;
      const uint64_t x = i < nb ? (uint64_t)v[i] : 0ull;
      uint64_t tot;
      const uint64_t ex = block_excl_scan64(x, s_scan, &tot);
      if (i < nb) v[i] = (int64_t)(carry + ex);
      carry += tot;
    }
  }
}

// Each window's first pair in input order and in run order, both stored by
// rank (ooff[rank], poff[rank]: the leaf write reads them beside the rank's
// input index, all in run order): thread t of block b takes 16 consecutive
// windows of the block's kOrdBlock, in input order for ooff and in run order
// for poff.
__global__ __launch_bounds__(256) void k_ord_apply(int64_t nq, const unsigned long long* __restrict__ cnt,
                                                   const uint32_t* __restrict__ inv,
                                                   const int64_t* __restrict__ sums,
                                                   int64_t* __restrict__ ooff,
                                                   int64_t* __restrict__ poff) {
  __shared__ uint64_t s_scan[kQWaves];
  constexpr int kPer = kOrdBlock / 256;
  const int64_t nb = (nq + kOrdBlock - 1) / kOrdBlock;
  const int64_t b = blockIdx.x;
  const int64_t i0 = b * kOrdBlock + (int64_t)threadIdx.x * kPer;
  uint64_t co[kPer], cp[kPer];
  uint32_t rk[kPer];
  uint64_t so = 0, sp = 0;
#pragma unroll
  for (int k = 0; k < kPer; k++) {
    const int64_t i = i0 + k;
    rk[k] = i < nq ? inv[i] : 0u;
    co[k] = i < nq ? cnt[rk[k]] : 0ull;
    cp[k] = i < nq ? cnt[i] : 0ull;
    so += co[k];
    sp += cp[k];
  }
  uint64_t tot;
  uint64_t eo = (uint64_t)sums[b] + block_excl_scan64(so, s_scan, &tot);
  uint64_t ep = (uint64_t)sums[nb + b] + block_excl_scan64(sp, s_scan, &tot);
#pragma unroll
  for (int k = 0; k < kPer; k++) {
    const int64_t i = i0 + k;
    if (i < nq) {
      ooff[rk[k]] = (int64_t)eo;
      poff[i] = (int64_t)ep;
    }
    eo += co[k];
    ep += cp[k];
  }
}

int query_run_f32(QEngine& qe, Prof& pf, const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox,
                  const curtree_bbox_f32* qbox, int64_t q0, int64_t q1, int contain, QueryOut* qo,
                  hipStream_t s, Context* ctx) {
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
  a.chained = qo->algo == CURTREE_ALGO_PERSISTENT;
  // caller buffer: one writing run; otherwise count first
  const bool direct = !qo->alloc && !qo->count_only && !qo->counts && qo->fid;
  a.write = direct;
  a.out_fid = qo->fid;
  a.out_tid = qo->tid;
  a.out_cap = direct ? qo->cap : 0;
  if ((rc = qe.reserve_front(std::max<int64_t>(2 * nq, 1 << 20)))) return rc;
  // the exact output of an allocating query: freed on every error return
  struct Owned {
    int* fid = nullptr;
    int* tid = nullptr;
    ~Owned() {
      if (fid) (void)hipFree(fid);
      if (tid) (void)hipFree(tid);
    }
  } owned;
  // windows in Morton order of their centres (large batches of the plain entry
  // points; sorted and copied once, before the passes).  Per-window counts
  // for the caller are by input index: such runs keep the input order.
  const int ow = qtuning().order;
  const bool order = ctx && q0 == 0 && !a.generic && !a.chained && !qe.fixed && !qo->counts &&
                     qtuning().fused != 1 && nq >= kSplitMinQueries &&
                     nq < ((int64_t)1 << 30) &&
                     (ow == 1 || (ow < 0 && nq >= kOrderMinWindows));
  uint32_t* perm = nullptr;
  unsigned long long* ocnt = nullptr;
  int64_t *ooff = nullptr, *poff = nullptr, *osum = nullptr;
  if (order) {
    if ((rc = qe.reserve_order(nq))) return rc;
    char* base = (char*)qe.order;
    const size_t at = ((size_t)nq * 8 + 255) & ~(size_t)255;
    perm = (uint32_t*)base;  // then the rank of each input window (nq more)
    float4* win4 = (float4*)(base + at);
    ocnt = (unsigned long long*)(win4 + nq);
    ooff = (int64_t*)(ocnt + nq);
    poff = ooff + nq;
    osum = poff + nq;
    // sort keys in the counts' space (cleared before every writing run)
    int* key = (int*)ocnt;
    pf.mark(s, "k_win_keys");
    const unsigned g = (unsigned)std::min<int64_t>((nq + 255) / 256, 2048);
    hipLaunchKernelGGL(k_win_keys, dim3(g), dim3(256), 0, s, nq, a.qb, a.tr.mbr, key, (int*)perm);
    CT_HIP(hipGetLastError());
    if ((rc = pairs_full_sort(*ctx, (int*)perm, key, nq, s, true))) return rc;
    pf.mark(s, "k_win_stage");
    hipLaunchKernelGGL(k_win_stage, dim3(g), dim3(256), 0, s, nq, a.qb, (const uint32_t*)perm,
                       win4, perm + nq);
    CT_HIP(hipGetLastError());
    pf.mark(s, nullptr);
    a.qb.b4 = win4;  // the run's windows are ranks
    a.tmap = perm;
  }
  for (int pass = 0; pass < 2; pass++) {
    if (order && a.write) {  // counts per rank, and the offsets from them
      a.counts = ocnt;
      a.ooff = ooff;
      a.poff = poff;
      a.osum = osum;
    }
    for (int attempt = 0;; attempt++) {
      if (a.counts) enqueue_fill_counts((int64_t*)a.counts + (a.counts == ocnt ? 0 : q0), nq, s);
      if ((rc = qe.reserve_status(std::max<int64_t>(qe.cap_front, nq) / per_tile(a.G, a.steps) +
                                  2)))
        return rc;
      if (a.write && (rc = qe.reserve_pieces(pieces_needed(qe, a)))) return rc;
      enqueue_run(qe, pf, a, s);
      CT_HIP(hipGetLastError());
      if ((rc = read_ctrl(qe, s))) return rc;
      qe.last_nl = a.nl;
      if (!qe.host_ctrl->overflow) break;
      // an overflow stops the run at the level that did not fit, so every
      // attempt learns at least one more level: nl attempts always suffice
      if (attempt > a.nl) return set_error(CURTREE_ENOMEM, "query: frontier kept overflowing");
      const uint64_t* cnt = qe.host_ctrl->count;
      uint64_t need = 0;
      int deepest = 1;
      for (int l = 1; l < a.nl; l++) {
        if (cnt[l] > need) need = cnt[l];
        if (cnt[l]) deepest = l;
      }
      if (need > 0xFFFFFFFFull)
        return set_error(CURTREE_ERANGE, "query: frontier of %llu entries exceeds 2^32",
                         (unsigned long long)need);
      // the levels below the cut are unknown: extrapolate the last growth
      // factor (at most F) to save reruns, falling back to the known size
      uint64_t guess = need + need / 4 + 1024;
      if (deepest + 1 < a.nl && deepest >= 2 && cnt[deepest - 1] > 0) {
        double r = (double)cnt[deepest] / (double)cnt[deepest - 1];
        r = std::min<double>(std::max(r, 1.0), (double)a.F);
        guess = std::min<uint64_t>((uint64_t)((double)need * r * 1.25) + 1024, 0xFFFFFFFFull);
      }
      if (qe.reserve_front((int64_t)guess) != CURTREE_OK &&
          (rc = qe.reserve_front((int64_t)(need + need / 4 + 1024))))
        return rc;
    }
    const uint64_t H = qe.host_ctrl->total;
    qo->n = (int64_t)H;
    if (a.write) {
      if ((int64_t)H > a.out_cap)
        return set_error(CURTREE_ERANGE, "query: %llu pairs exceed the buffer of %lld",
                         (unsigned long long)H, (long long)a.out_cap);
      qo->out_fid = a.out_fid;
      qo->out_tid = a.out_tid;
      owned.fid = owned.tid = nullptr;  // handed to the caller
      return CURTREE_OK;
    }
    if (qo->count_only || H == 0) return CURTREE_OK;
    // second pass: write into an exact allocation (or the caller's buffer)
    if (qo->alloc) {
      if (hipMalloc((void**)&owned.fid, H * sizeof(int)) != hipSuccess ||
          hipMalloc((void**)&owned.tid, H * sizeof(int)) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(CURTREE_ENOMEM, "query: cannot allocate %llu result pairs",
                         (unsigned long long)H);
      }
      a.out_fid = owned.fid;
      a.out_tid = owned.tid;
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
  return query_run_f32(ctx.qe, ctx.prof, rt, rbox, qbox, 0, qbox->sz, ctx.opt_contain, qo, s,
                       &ctx);
}

// ---------------------------------------------------------------------------
// The allocation-free query (curtree_query_into_noalloc_f32): the plain
// query's kernels on an engine whose buffers are slices of `scratch`, a
// file-scope device array that every GPU holds until the process ends.  The
// containment shortcut and the AoS staging are off (they need buffers the size
// of the tree), so the run needs only the control block, the layout flag, the
// scan status words, two frontier buffers, the per-entry state and the tile
// counts, all sized for kScratchEntries frontier entries.
constexpr int64_t kScratchEntries = int64_t(3) << 20;
constexpr int64_t kScratchMaxWindows = kScratchEntries / 2;  // reserve_front asks for 2 per window
constexpr int64_t kScratchTileWords = kScratchEntries / 16 + 2;  // tiles hold >= 16 entries
constexpr size_t scr_up(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kScrCtrl = 0;
constexpr size_t kScrCanon = 4096;
constexpr size_t kScrStatus = 8192;
constexpr size_t kScrFront0 = kScrStatus + scr_up((size_t)kScratchTileWords * 8);
constexpr size_t kScrFront1 = kScrFront0 + scr_up((size_t)kScratchEntries * 8);
constexpr size_t kScrEst = kScrFront1 + scr_up((size_t)kScratchEntries * 8);
constexpr size_t kScrTtot = kScrEst + scr_up((size_t)kScratchEntries * kEntryBytes);
constexpr size_t kScrEnd = kScrTtot + scr_up((size_t)kScratchTileWords * 8);
constexpr size_t kScratchBytes = size_t(128) << 20;
static_assert(sizeof(QCtrl) <= kScrCanon, "control block slice");
static_assert(kScrEnd <= kScratchBytes, "scratch layout");
static_assert(kScrFront0 % 256 == 0 && kScrEst % 256 == 0 && kScrTtot % 256 == 0,
              "slices stay aligned");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kScratchDevices = 64;
struct ScratchSlot {
  bool ready = false;
  QEngine qe;
  QCtrl host_ctrl;  // read-back copy (plain host memory: nothing is allocated)
};
std::mutex g_scratch_mu;  // one allocation-free query at a time in the process
ScratchSlot g_scratch[kScratchDevices];

// The engine of the current device over that device's copy of `scratch`; its
// address is looked up once per device, with the device current.
int scratch_engine(QEngine** out) {
  int dev = 0;
  CT_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kScratchDevices)
    return set_error(CURTREE_EINVAL, "query_noalloc: device %d beyond %d", dev, kScratchDevices);
  ScratchSlot& sl = g_scratch[dev];
  if (!sl.ready) {
    unsigned char* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return set_error(CURTREE_EHIP, "query_noalloc: no address for the scratch array: %s",
                       hipGetErrorString(e));
    QEngine& q = sl.qe;
    q.fixed = true;
    q.ctrl = (QCtrl*)(ptr + kScrCtrl);
    q.host_ctrl = &sl.host_ctrl;
    q.canon = (uint32_t*)(ptr + kScrCanon);
    q.status = (uint64_t*)(ptr + kScrStatus);
    q.status_words = kScratchTileWords;
    q.front[0] = ptr + kScrFront0;
    q.front[1] = ptr + kScrFront1;
    q.est = ptr + kScrEst;
    q.ttot = (uint64_t*)(ptr + kScrTtot);
    q.cap_front = kScratchEntries;
    q.needs_clear = true;  // the status words start at zero on the first run
    sl.ready = true;
  }
  *out = &sl.qe;
  return CURTREE_OK;
}
}  // namespace

int query_into_noalloc_f32(const curtree_rtree_f32* rt, const curtree_bbox_f32* rbox,
                           const curtree_bbox_f32* qbox, int* fid, int* tid, int64_t cap,
                           int64_t* n_out, hipStream_t s) {
  if (n_out) *n_out = 0;
  if (!qbox) return set_error(CURTREE_EINVAL, "query_noalloc: null argument");
  if (cap > 0 && (!fid || !tid)) return set_error(CURTREE_EINVAL, "query_noalloc: null output");
  if (qbox->sz > kScratchMaxWindows)
    return set_error(CURTREE_EINVAL, "query_noalloc: %d windows, at most %lld per call",
                     qbox->sz, (long long)kScratchMaxWindows);
  std::lock_guard<std::mutex> g(g_scratch_mu);
  QEngine* qe = nullptr;
  int rc;
  if ((rc = scratch_engine(&qe))) return rc;
  Prof pf;  // no per-kernel events (they would be allocated)
  QueryOut qo;
  qo.fid = fid;
  qo.tid = tid;
  qo.cap = cap;
  rc = query_run_f32(*qe, pf, rt, rbox, qbox, 0, qbox->sz, 0, &qo, s);
  if (n_out) *n_out = qo.n;
  if (rc) return rc;
  CT_HIP(hipStreamSynchronize(s));
  return CURTREE_OK;
}

}  // namespace ct
