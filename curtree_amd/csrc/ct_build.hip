// ct_build.hip -- bulk load on gfx950: Morton keys -> onesweep LSD radix sort
// of 32-bit keys that carries the boxes and ids (no gather; its last pass
// writes the caller's arrays) -> final pass, in place: ties of the 64-bit keys
// settled and the leaf-parent level reduced -> upper levels.
//
// Reference algorithm: src/rtree_bulkload_thrust.cuh:62-213 (key transform
// :76-81, sort_by_key :83-86, level plan :101-115, reduce_by_key :126-171).
#include "ct_internal.hpp"

namespace ct {

// ---------------------------------------------------------------------------
// Block-wide exclusive scan over NW waves.  s_tmp holds NW words.
template <int NW>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_tmp,
                                                    uint32_t* total) {
  const int lane = lane_id(), w = threadIdx.x >> 6;
  uint32_t x = v;
  for (int o = 1; o < kWave; o <<= 1) {
    uint32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == kWave - 1) s_tmp[w] = x;
  __syncthreads();
  uint32_t off = 0, tot = 0;
  for (int i = 0; i < NW; i++) {
    uint32_t t = s_tmp[i];
    if (i < w) off += t;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return x - v + off;
}

// ---------------------------------------------------------------------------
// Control block layout (device memory, zeroed by k_build_init).
constexpr int kExtSlots = 64;  // extent partials are spread over slots to avoid atomic contention

// Sort steps: one onesweep pass each, over digit digit[step].  The payload
// sort runs steps 0..3 over the four digits of its 32-bit key (REF32's key or
// the high half of the Morton key); runs of equal high halves are settled by
// k_final / k_fixup_long, and only if a run is too long for that does the
// full 64-bit LSD sort run (steps 4..11).  The pair sort runs steps 0..7.
constexpr int kMaxSteps = 12;
struct BuildCtrl {
  uint32_t ext[kExtSlots][4];      // ordered-uint lox, loy, hix, hiy of centroids, spread
  uint32_t hist[kMaxDigits * 256];  // global digit histograms
  uint32_t base[kMaxDigits * 256];  // exclusive digit offsets
  uint32_t ticket[kMaxSteps];
  int32_t digit[kMaxSteps];        // digit of each step
  int32_t active[kMaxSteps];       // step s runs
  int32_t src[kMaxSteps];          // step s reads buffer src[s]
  int32_t write_keys[kMaxSteps];   // step s writes keys (not the last step: values only)
  int32_t npass;                   // steps that ran so far (planned): parity of the data
  int32_t first;                   // first step that runs (-1: none)
  uint32_t fallback;               // a run is too long to settle in place: full 64-bit sort
  uint32_t nruns;                  // runs listed by k_final
  uint32_t nmid;                   // of them 65..kWaveSortMax long (k_fixup_long, a wave each)
  uint32_t nlong;                  // of them up to kFixLong long (k_fixup_long, a block each)
  uint32_t nhuge;                  // of them longer than kFixLong (k_fixup_huge)
  uint32_t huge_next;              // k_fixup_huge: next listed run to claim
  uint32_t trivial;                // bit d: digit d is the same for every key
};

// wf: the tree's well-formed word (or null), see tree_state_register: set
// here, cleared by k_hist if a box is not well-formed.
__global__ __launch_bounds__(256) void k_build_init(BuildCtrl* c, uint32_t* wf) {
  if (wf && threadIdx.x == 0) *wf = 1u;
  for (int i = threadIdx.x; i < kMaxDigits * 256; i += 256) {
    c->hist[i] = 0;
    c->base[i] = 0;
  }
  if (threadIdx.x < kMaxSteps) {
    c->ticket[threadIdx.x] = 0;
    c->digit[threadIdx.x] = 0;
    c->active[threadIdx.x] = 0;
    c->src[threadIdx.x] = 0;
    c->write_keys[threadIdx.x] = 0;
  }
  if (threadIdx.x < kExtSlots) {
    c->ext[threadIdx.x][0] = 0xFFFFFFFFu;
    c->ext[threadIdx.x][1] = 0xFFFFFFFFu;
    c->ext[threadIdx.x][2] = 0u;
    c->ext[threadIdx.x][3] = 0u;
  }
  if (threadIdx.x == 0) {
    c->npass = 0;
    c->first = -1;
    c->fallback = 0;
    c->nruns = 0;
    c->nmid = 0;
    c->nlong = 0;
    c->nhuge = 0;
    c->huge_next = 0;
    c->trivial = 0;
  }
}

// Centroid extent (MORTON64 mode).  NaN centroids are skipped.
__global__ __launch_bounds__(256) void k_extent(int64_t n, const float* __restrict__ xmin,
                                                const float* __restrict__ ymin,
                                                const float* __restrict__ xmax,
                                                const float* __restrict__ ymax, BuildCtrl* c) {
  uint32_t lx = 0xFFFFFFFFu, ly = 0xFFFFFFFFu, hx = 0u, hy = 0u;
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    float cx = (xmin[i] + xmax[i]) * 0.5f;
    float cy = (ymin[i] + ymax[i]) * 0.5f;
    if (cx == cx) {
      uint32_t o = f2ord(cx);
      lx = min(lx, o);
      hx = max(hx, o);
    }
    if (cy == cy) {
      uint32_t o = f2ord(cy);
      ly = min(ly, o);
      hy = max(hy, o);
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    lx = min(lx, (uint32_t)__shfl_xor(lx, o));
    ly = min(ly, (uint32_t)__shfl_xor(ly, o));
    hx = max(hx, (uint32_t)__shfl_xor(hx, o));
    hy = max(hy, (uint32_t)__shfl_xor(hy, o));
  }
  __shared__ uint32_t s_e[4][4];
  const int w = threadIdx.x >> 6;
  if (lane_id() == 0) {
    s_e[w][0] = lx;
    s_e[w][1] = ly;
    s_e[w][2] = hx;
    s_e[w][3] = hy;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int k = threadIdx.x;
    uint32_t v = s_e[0][k];
    for (int i = 1; i < 4; i++) v = k < 2 ? min(v, s_e[i][k]) : max(v, s_e[i][k]);
    uint32_t* slot = &c->ext[blockIdx.x % kExtSlots][k];
    if (k < 2) atomicMin(slot, v);
    else atomicMax(slot, v);
  }
}

// Final extent from the spread slots (every thread of the block gets it).
__device__ __forceinline__ void read_extent(const BuildCtrl* c, uint32_t* out4) {
  __shared__ uint32_t s_x[4];
  if (threadIdx.x < 64) {
    uint32_t lx = c->ext[threadIdx.x][0], ly = c->ext[threadIdx.x][1];
    uint32_t hx = c->ext[threadIdx.x][2], hy = c->ext[threadIdx.x][3];
    for (int o = 32; o > 0; o >>= 1) {
      lx = min(lx, (uint32_t)__shfl_xor(lx, o));
      ly = min(ly, (uint32_t)__shfl_xor(ly, o));
      hx = max(hx, (uint32_t)__shfl_xor(hx, o));
      hy = max(hy, (uint32_t)__shfl_xor(hy, o));
    }
    if (threadIdx.x == 0) {
      s_x[0] = lx;
      s_x[1] = ly;
      s_x[2] = hx;
      s_x[3] = hy;
    }
  }
  __syncthreads();
  for (int k = 0; k < 4; k++) out4[k] = s_x[k];
}

// Extent -> (lo, scale) per axis, the same FP32 operations as the oracle.
__device__ __forceinline__ void axis_params(uint32_t lo_o, uint32_t hi_o, float* lo, float* s) {
  if (lo_o > hi_o) {  // no non-NaN centroid
    *lo = 0.0f;
    *s = 0.0f;
    return;
  }
  float l = ord2f(lo_o), h = ord2f(hi_o);
  float w = h - l;
  *lo = l;
  *s = (w > 0.0f) ? 4294967296.0f / w : 0.0f;
}

// ---------------------------------------------------------------------------
// Keys.  REF32: key = z_order((ushort)cx, (ushort)cy) as in mbr2key
// (src/rtree_bulkload_thrust.cuh:52-60), stored with the sign bit flipped so
// that an unsigned radix sort yields the reference's signed int32 order.
// MORTON64: 32-bit quantised centroids over the centroid extent, interleaved.
// The payload sort runs on a 32-bit key: REF32's key, or the high half of the
// Morton key (ties of the high half are settled on the low half afterwards).
struct KeyParams {
  float lox, sx, loy, sy;
};

// Block-wide: every thread of the block calls it.
template <int MODE>
__device__ __forceinline__ KeyParams key_params(const BuildCtrl* c) {
  KeyParams p{0.f, 0.f, 0.f, 0.f};
  if (MODE == kKeyMorton64) {
    uint32_t e4[4];
    read_extent(c, e4);
    axis_params(e4[0], e4[2], &p.lox, &p.sx);
    axis_params(e4[1], e4[3], &p.loy, &p.sy);
  }
  return p;
}

__device__ __forceinline__ uint64_t morton64(float a, float b, float cc, float d,
                                             const KeyParams& p) {
  const float cx = (a + cc) * 0.5f;
  const float cy = (b + d) * 0.5f;
  return (spread32(q32(cy, p.loy, p.sy)) << 1) | spread32(q32(cx, p.lox, p.sx));
}

// The two halves of morton64 in comparable form without interleaving: the
// cell (qy_hi16 << 16) | qx_hi16, equal exactly when the high halves are, and
// *lo = (qy_lo16 << 16) | qx_lo16 (order: morton_less16).
__device__ __forceinline__ uint32_t morton_cell(float a, float b, float cc, float d,
                                                const KeyParams& p, uint32_t* lo) {
  const float cx = (a + cc) * 0.5f;
  const float cy = (b + d) * 0.5f;
  const uint32_t qx = q32(cx, p.lox, p.sx), qy = q32(cy, p.loy, p.sy);
  *lo = (qy << 16) | (qx & 0xFFFFu);
  return (qy & 0xFFFF0000u) | (qx >> 16);
}

// Order of the low halves of two Morton keys given as morton_cell's *lo: the
// key interleaves the coordinates with y above x at every bit, so the
// coordinate whose difference has the higher top bit decides (y on a tie).
__device__ __forceinline__ bool morton_less16(uint32_t a, uint32_t b) {
  const uint32_t d = a ^ b;
  const uint32_t dx = d & 0xFFFFu, dy = d >> 16;
  const bool xdec = dy < dx && dy < (dy ^ dx);  // top bit of dy below that of dx
  return xdec ? (a & 0xFFFFu) < (b & 0xFFFFu) : (a >> 16) < (b >> 16);
}

__device__ __forceinline__ uint32_t ref32(float a, float b, float cc, float d) {
  const float cx = (a + cc) / 2.0f;
  const float cy = (b + d) / 2.0f;
  return ((spread16(f2u16_sat(cy)) << 1) | spread16(f2u16_sat(cx))) ^ 0x80000000u;
}

// The high half of morton64 from the high 16 bits of each quantised
// coordinate (bit 16 + j of x lands on bit 2j, of y on bit 2j + 1): the same
// bits as morton64(...) >> 32 with 32-bit spreads only.
template <int MODE>
__device__ __forceinline__ uint32_t sort_key32(float a, float b, float cc, float d,
                                               const KeyParams& p) {
  if (MODE == kKeyRef32) return ref32(a, b, cc, d);
  const float cx = (a + cc) * 0.5f;
  const float cy = (b + d) * 0.5f;
  return (spread16(q32(cy, p.loy, p.sy) >> 16) << 1) | spread16(q32(cx, p.lox, p.sx) >> 16);
}

// The histograms of the four digits of the 32-bit sort keys (the first sort
// pass computes the keys again from the boxes it reads), and the status words
// of the sort steps cleared.  wf (or null): cleared when a box is not
// well-formed (xmin <= xmax and ymin <= ymax fails, NaN included); the sort
// only permutes the boxes, so this holds for the sorted boxes as well.
template <int MODE>
__global__ __launch_bounds__(256) void k_hist(int64_t n, const float* __restrict__ xmin,
                                              const float* __restrict__ ymin,
                                              const float* __restrict__ xmax,
                                              const float* __restrict__ ymax, BuildCtrl* c,
                                              uint32_t* __restrict__ zero_ptr,
                                              int64_t zero_words, uint32_t* wf) {
  __shared__ uint32_t sh[4 * 256];
  for (int t = threadIdx.x; t < 4 * 256; t += 256) sh[t] = 0;
  const KeyParams kp = key_params<MODE>(c);
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * 256;
  constexpr int U = 4;  // boxes per thread and step, all loaded before any is used
  for (int64_t i0 = blockIdx.x * (256ll * U) + threadIdx.x; i0 < n; i0 += stride * U) {
    float4 b[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t i = i0 + 256 * u;
      b[u] = i < n ? make_float4(xmin[i], ymin[i], xmax[i], ymax[i])
                   : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    if (wf) {
      bool bad = false;
#pragma unroll
      for (int u = 0; u < U; u++)
        bad |= i0 + 256 * u < n && !(b[u].x <= b[u].z && b[u].y <= b[u].w);
      if (ballot(bad) && lane_id() == 0) *wf = 0u;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const bool ok = i0 + 256 * u < n;
      const uint32_t k = sort_key32<MODE>(b[u].x, b[u].y, b[u].z, b[u].w, kp);
      // a digit that the whole wave shares is added once (degenerate keys,
      // such as REF32 on unit-square data, would otherwise serialise on one
      // address)
      const uint64_t act = ballot(ok);
      if (!act) break;
#pragma unroll
      for (int p = 0; p < 4; p++) {
        const uint32_t d = (k >> (8 * p)) & 255u;
        const uint32_t d0 = __builtin_amdgcn_readlane(d, __ffsll((long long)act) - 1);
        const uint64_t same = ballot(ok && d == d0);
        if (same == act) {
          if (lane_id() == __ffsll((long long)act) - 1)
            atomicAdd(&sh[p * 256 + d0], (uint32_t)popc64(act));
        } else if (ok) {
          atomicAdd(&sh[p * 256 + d], 1u);
        }
      }
    }
  }
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < zero_words; i += stride) zero_ptr[i] = 0u;
  __syncthreads();
  for (int t = threadIdx.x; t < 4 * 256; t += 256)
    if (sh[t]) atomicAdd(&c->hist[t], sh[t]);
}

// 64-bit keys + 32-B staging records + all-digit histograms: the input of the
// full 8-pass LSD sort, which runs only when the high-half sort leaves a run
// of equal high halves too long to settle in place (`gate`: skip unless
// c->fallback; the keys go to the buffer the next step reads).
__global__ __launch_bounds__(256) void k_keygen64(int64_t n, const float* __restrict__ xmin,
                                                  const float* __restrict__ ymin,
                                                  const float* __restrict__ xmax,
                                                  const float* __restrict__ ymax,
                                                  const int* __restrict__ id, BuildCtrl* c,
                                                  uint64_t* __restrict__ k0,
                                                  uint64_t* __restrict__ k1,
                                                  uint4* __restrict__ aos, int gate) {
  __shared__ uint32_t sh[8 * 256];
  if (gate && !c->fallback) return;
  for (int t = threadIdx.x; t < 8 * 256; t += 256) sh[t] = 0;
  const KeyParams kp = key_params<kKeyMorton64>(c);
  uint64_t* keys_out = (c->npass & 1) ? k1 : k0;
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += stride) {
    const float a = xmin[i], b = ymin[i], cc = xmax[i], d = ymax[i];
    aos[2 * i] = make_uint4(__float_as_uint(a), __float_as_uint(b), __float_as_uint(cc),
                            __float_as_uint(d));
    aos[2 * i + 1] = make_uint4((uint32_t)id[i], 0u, 0u, 0u);
    const uint64_t k = morton64(a, b, cc, d, kp);
    keys_out[i] = k;
#pragma unroll
    for (int p = 0; p < 8; p++) atomicAdd(&sh[p * 256 + (uint32_t)((k >> (8 * p)) & 255u)], 1u);
  }
  __syncthreads();
  for (int t = threadIdx.x; t < 8 * 256; t += 256)
    if (sh[t]) atomicAdd(&c->hist[t], sh[t]);
}

// Step plan: exclusive digit offsets for digits 0..D-1 and, for steps
// [s0, s0+ns) over the digits packed 4 bits each in `digits` (lowest first),
// which steps run and from which buffer.  A step whose digit is the same for
// every key is skipped (it would be the identity permutation), except that at
// least `min_steps` steps run: the first trivial steps are kept when fewer
// digits vary (a later stage reads the buffers a step writes; the payload
// sort's last step must not be its first, which reads the arrays it writes),
// and that `keep_first` keeps the range's first step (the payload sort's
// first launch is the one that computes the keys).
__device__ void plan_steps(BuildCtrl* c, int s0, int ns, uint64_t digits, int min_steps,
                           bool keep_first) {
  int varying = 0;
  for (int k = 0; k < ns; k++)
    if (!((c->trivial >> (int)((digits >> (4 * k)) & 15u)) & 1u) || (keep_first && k == 0))
      varying++;
  int force = min_steps - varying;  // identity steps that still run
  int J = c->npass, first = c->first;
  for (int k = 0; k < ns; k++) {
    const int st = s0 + k, d = (int)((digits >> (4 * k)) & 15u);
    c->digit[st] = d;
    c->ticket[st] = 0;
    bool run = !((c->trivial >> d) & 1u) || (keep_first && k == 0);
    if (!run && force > 0) {
      run = true;
      force--;
    }
    if (!run) {
      c->active[st] = 0;
    } else {
      c->active[st] = 1;
      c->src[st] = J & 1;
      c->write_keys[st] = 1;
      if (first < 0) first = st;
      J++;
    }
  }
  c->npass = J;
  c->first = first;
}

// Digit offsets for digits 0..D-1 from the histograms and the plan of steps
// [s0, s0+ns) (at least min_steps of them run; keep_first: the first one
// runs).  `gate`: plan only if the fallback sort was asked for (its first step
// then reads input positions as values again).
__global__ __launch_bounds__(256) void k_sort_plan(int64_t n, int D, BuildCtrl* c, int s0, int ns,
                                                   uint64_t digits, int keep_last_keys, int gate,
                                                   int min_steps, int keep_first) {
  __shared__ uint32_t s_tmp[4];
  __shared__ uint32_t s_triv;
  if (gate && !c->fallback) return;
  if (threadIdx.x == 0) s_triv = 0;
  __syncthreads();
  for (int p = 0; p < D; p++) {
    uint32_t v = c->hist[p * 256 + threadIdx.x];
    int triv = __syncthreads_or(v == (uint32_t)n);
    uint32_t tot;
    uint32_t ex = block_excl_scan<4>(v, s_tmp, &tot);
    c->base[p * 256 + threadIdx.x] = ex;
    if (threadIdx.x == 0 && triv) s_triv |= 1u << p;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    c->trivial = s_triv;
    if (gate) c->first = -1;
    plan_steps(c, s0, ns, digits, min_steps, keep_first != 0);
    // the last step writes values only, unless a later stage reads the keys
    if (!keep_last_keys)
      for (int st = s0 + ns - 1; st >= s0; st--)
        if (c->active[st]) {
          c->write_keys[st] = 0;
          break;
        }
  }
}

// ---------------------------------------------------------------------------
// Onesweep LSD passes over 8 bits.  Tile = kSortThreads x kSortItems items,
// warp-striped (wave w owns items [w*64*I, (w+1)*64*I) of the tile, item i of
// lane l is item w*64*I + i*64 + l), so ballot ranking in (item, lane) order is
// the input order and the pass is stable.  Per-digit tile counts are chained
// across tiles by decoupled look-back; tiles are numbered by an atomic ticket
// so every tile waits only on tiles that have already started.
#ifndef CT_SORT_THREADS
#define CT_SORT_THREADS 1024
#endif
#ifndef CT_SORT_ITEMS
#define CT_SORT_ITEMS 10
#endif
#ifndef CT_SWEEP_ITEMS
#define CT_SWEEP_ITEMS 12
#endif
#ifndef CT_LOOKBACK
#define CT_LOOKBACK 4
#endif
// One block per CU: large tiles make long runs of equal digits (fewer written
// cache lines shared with the neighbouring tile) and a short look-back chain.
constexpr int kSortThreads = CT_SORT_THREADS;  // 256 .. 1024
constexpr int kSortWaves = kSortThreads / kWave;
constexpr int kSortItems = CT_SORT_ITEMS;      // 64-bit LSD passes
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kSweepItems = CT_SWEEP_ITEMS;    // payload passes
constexpr int kSweepTile = kSortThreads * kSweepItems;
constexpr uint32_t kFlagA = 1u << 30, kFlagP = 2u << 30, kValMask = (1u << 30) - 1u;
// grid of the 64-bit LSD passes (fallback sort, pair order): one block per CU
// (their LDS admits no second) looping over tickets, so a launch that has
// nothing to do costs one round of blocks
constexpr int64_t kOnesweepGrid = 256;
static_assert(kSortThreads >= 256, "one thread per digit");

// Lanes of the wave whose 8-bit digit d equals this lane's (valid lanes
// only): every valid lane ORs its bit into the digit's entry of a per-wave
// table of 256 lane masks in LDS and reads the entry back; the lowest lane of
// each digit then clears it, so the table is all zero again on return (it
// must be on entry).  Three LDS operations per item instead of eight ballots
// and the mask arithmetic around them.  The wave's LDS operations execute in
// order; the fences keep the compiler from moving them.
__device__ __forceinline__ uint64_t peer_mask(uint32_t d, bool valid, uint64_t* tab) {
  if (valid) __hip_atomic_fetch_or(&tab[d], 1ull << lane_id(), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WAVEFRONT);
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  const uint64_t m =
      valid ? __hip_atomic_load(&tab[d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT) : 0ull;
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (valid && (m & lanemask_lt()) == 0)
    __hip_atomic_store(&tab[d], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  return m;
}

struct SweepShared {
  uint64_t match[kSortWaves][256];  // peer_mask tables (zero between items)
  uint32_t wh[kSortWaves][256];  // per-wave digit counts, then wave offsets
  uint32_t start[256];           // first tile-local position of each digit
  uint32_t adj[256];             // output position of tile-local position j: adj[d] + j
  uint32_t tmp[kSortWaves];
  uint32_t tile;
};

// Stable rank of the tile's items by their digits dig[i] and the per-digit
// look-back.  On return pos[i] is item i's position in the tile reordered by
// digit, and the item at reordered position j with digit d belongs at output
// position sh.adj[d] + j.  sh.wh must be zero on entry (and synchronised).
template <int I>
__device__ __forceinline__ void sweep_rank(const uint32_t (&dig)[I], uint32_t (&pos)[I],
                                           int64_t wbase,
                                           int64_t n, int64_t tile, const uint32_t* gbase,
                                           uint32_t* status, SweepShared& sh) {
  const int t = threadIdx.x, w = t >> 6, lane = lane_id();
  const uint64_t lt = lanemask_lt();
#pragma unroll
  for (int i = 0; i < I; i++) {
    const int64_t idx = wbase + i * kWave + lane;
    const bool valid = idx < n;
    const uint32_t d = dig[i];
    const uint64_t m = peer_mask(d, valid, sh.match[w]);
    const uint32_t old = sh.wh[w][d];
    pos[i] = old + (uint32_t)popc64(m & lt);
    if (valid && (m & lt) == 0) sh.wh[w][d] = old + (uint32_t)popc64(m);
  }
  __syncthreads();

  // threads 0..255 own one digit each: wave offsets, tile count, look-back
  const int dg = t;
  uint32_t tot = 0, excl = 0;
  if (t < 256) {
#pragma unroll
    for (int i = 0; i < kSortWaves; i++) {
      const uint32_t ci = sh.wh[i][dg];
      sh.wh[i][dg] = tot;
      tot += ci;
    }
    uint32_t* st = status + (size_t)tile * 256;
    if (tile == 0) {
      st_agent(st + dg, kFlagP | tot);
    } else {
      st_agent(st + dg, kFlagA | tot);
      // look back kLB tiles per step (kLB loads in flight); stop at the
      // nearest inclusive prefix; tiles before the first count as prefix 0
      constexpr int kLB = CT_LOOKBACK;
      int64_t j = tile - 1;
      for (;;) {
        uint32_t v[kLB];
#pragma unroll
        for (int k = 0; k < kLB; k++)
          v[k] = (j - k >= 0) ? ld_agent(status + (size_t)(j - k) * 256 + dg) : kFlagP;
        int stop = kLB;
        bool ready = true;
#pragma unroll
        for (int k = 0; k < kLB; k++) {
          if (stop < kLB) break;
          const uint32_t f = v[k] & ~kValMask;
          if (f == 0u) {
            ready = false;
            break;
          }
          if (f == kFlagP) stop = k;
        }
        if (!ready) continue;
#pragma unroll
        for (int k = 0; k < kLB; k++)
          if (k <= stop) excl += v[k] & kValMask;
        if (stop < kLB) break;
        j -= kLB;
      }
      st_agent(st + dg, kFlagP | (excl + tot));
    }
  }
  uint32_t ttot;
  const uint32_t start = block_excl_scan<kSortWaves>(tot, sh.tmp, &ttot);
  if (t < 256) {
    sh.start[dg] = start;
    sh.adj[dg] = gbase[dg] + excl - start;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < I; i++) pos[i] += sh.start[dig[i]] + sh.wh[w][dig[i]];
}

// One 8-bit pass of (key, u32 value) or keys alone: the full 8-pass LSD sort
// of 64-bit keys (fallback of the Morton sort, canonical pair order).
template <class KT>
__global__ __launch_bounds__(kSortThreads) void k_onesweep(
    int64_t n, int step, KT* __restrict__ k0, KT* __restrict__ k1, uint32_t* __restrict__ v0,
    uint32_t* __restrict__ v1, BuildCtrl* c, uint32_t* __restrict__ status, int keys_only) {
  __shared__ SweepShared sh;
  __shared__ KT s_buf[kSortTile];
  __shared__ uint8_t s_dig[kSortTile];

  if (!c->active[step]) return;
  const int t = threadIdx.x, w = t >> 6, lane = lane_id();
  const int pass = c->digit[step];
  const int shift = 8 * pass;
  const int src = c->src[step];
  const bool ident = step == c->first;  // first step: values are input positions
  const KT* kin = src ? k1 : k0;
  KT* kout = src ? k0 : k1;
  const uint32_t* vin = src ? v1 : v0;
  uint32_t* vout = src ? v0 : v1;
  const bool wk = keys_only || c->write_keys[step] != 0;
  const int64_t ntiles = (n + kSortTile - 1) / kSortTile;
  // blocks loop over tickets when the grid is smaller than the tile count
  for (;;) {
    if (t == 0) sh.tile = atomicAdd(&c->ticket[step], 1u);
    for (int i = t; i < kSortWaves * 256; i += kSortThreads) {
      (&sh.wh[0][0])[i] = 0;
      (&sh.match[0][0])[i] = 0;
    }
    __syncthreads();
    const int64_t tile = sh.tile;
    if (tile >= ntiles) break;

    const int64_t tbase = tile * kSortTile;
    const int64_t wbase = tbase + (int64_t)w * (kWave * kSortItems);
    KT key[kSortItems];
    uint32_t dig[kSortItems], pos[kSortItems];
#pragma unroll
    for (int i = 0; i < kSortItems; i++) {
      const int64_t idx = wbase + i * kWave + lane;
      key[i] = idx < n ? kin[idx] : (KT)0;
      dig[i] = (uint32_t)(key[i] >> shift) & 255u;
    }
    sweep_rank(dig, pos, wbase, n, tile, c->base + pass * 256, status, sh);

    // local reorder of keys through LDS, then scatter runs of equal digits
#pragma unroll
    for (int i = 0; i < kSortItems; i++) {
      const int64_t idx = wbase + i * kWave + lane;
      if (idx < n) s_buf[pos[i]] = key[i];
    }
    __syncthreads();
    const int64_t rem = n - tbase;
    const int tn = rem < kSortTile ? (int)rem : kSortTile;
#pragma unroll
    for (int i = 0; i < kSortItems; i++) {
      const int j = t + i * kSortThreads;
      if (j < tn) {
        const KT kk = s_buf[j];
        const uint32_t d = (uint32_t)(kk >> shift) & 255u;
        s_dig[j] = (uint8_t)d;
        if (wk) kout[sh.adj[d] + (uint32_t)j] = kk;
      }
    }
    if (!keys_only) {
      // values: reorder through LDS, scatter
      uint32_t val[kSortItems];
#pragma unroll
      for (int i = 0; i < kSortItems; i++) {
        const int64_t idx = wbase + i * kWave + lane;
        val[i] = idx < n ? (ident ? (uint32_t)idx : vin[idx]) : 0u;
      }
      __syncthreads();
      uint32_t* s_val = (uint32_t*)s_buf;
#pragma unroll
      for (int i = 0; i < kSortItems; i++) {
        const int64_t idx = wbase + i * kWave + lane;
        if (idx < n) s_val[pos[i]] = val[i];
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < kSortItems; i++) {
        const int j = t + i * kSortThreads;
        if (j < tn) vout[sh.adj[s_dig[j]] + (uint32_t)j] = s_val[j];
      }
    }
    if ((int64_t)gridDim.x >= ntiles) break;  // one tile per block
    __syncthreads();  // the tile is done with the shared arrays
  }
}

// Payload sort passes: the items are (key32 << 32 | id) plus the box as two
// float2 halves (xmin, ymin) and (xmax, ymax), all moved by every pass, so
// that the sorted boxes need no gather.  The first pass that runs reads the
// keys k_hist stored and the caller's SoA arrays; later passes read the
// staging buffers the previous pass wrote.  The three 8-B fields are
// reordered through LDS one after the other, so the tile's LDS buffer stays
// 8 B per item.
// ---------------------------------------------------------------------------
// Runs of equal keys inside one wave, whose lanes hold 64 consecutive items.
// dpp_perm moves a value by a DPP lane pattern; a lane whose source lies
// outside the wave, or outside the active lanes, reads 0 (so every lane of
// the wave runs these helpers; with no old value to keep, a move folds into
// the min/max that uses it).
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_perm(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, true);
}
// whole-wave shifts by one lane: lane i takes lane i - 1 (shr) or i + 1 (shl)
constexpr int kDppWaveShr1 = 0x138;
constexpr int kDppWaveShl1 = 0x130;

// First and last lane of this lane's run of equal keys (eqp / eqn: the same
// key as the previous / next item); -1 / kWave when the run goes on beyond
// the wave.
__device__ __forceinline__ void run_bounds(bool eqp, bool eqn, int* st, int* en) {
  const int lane = lane_id();
  const uint64_t S = ballot(!eqp && eqn), E = ballot(eqp && !eqn);
  const uint64_t le = S & (lanemask_lt() | (1ull << lane)), ge = E & ~lanemask_lt();
  *st = le ? 63 - __clzll((long long)le) : -1;
  *en = ge ? __ffsll((long long)ge) - 1 : kWave;
}

// Destination lane of this lane's item when every run marked `local` (lanes
// st..en) is sorted by (low half, lane), low halves in morton_cell form;
// other lanes stay.  Runs of two (nearly all runs) swap when the second key
// is strictly less; longer runs rank each member against the others, shifted
// in one lane per step (a lower lane with an equal key comes first).
__device__ __forceinline__ int local_run_dest(bool local, int st, int en, uint32_t lo) {
  const int lane = lane_id();
  const uint32_t lo_next = dpp_perm<kDppWaveShl1>(lo), lo_prev = dpp_perm<kDppWaveShr1>(lo);
  int dest = lane;
  if (local && en - st == 1) {
    const bool head = lane == st;
    if (head ? morton_less16(lo_next, lo) : morton_less16(lo, lo_prev))
      dest = head ? lane + 1 : lane - 1;
  }
  const bool longer = local && en - st > 1;
  if (ballot(longer)) {
    uint32_t vl = lo, vr = lo;
    int rank = 0;
    for (int d = 1; ballot(longer && d <= en - st); d++) {
      vl = dpp_perm<kDppWaveShr1>(vl);  // key of lane - d
      vr = dpp_perm<kDppWaveShl1>(vr);  // key of lane + d
      if (longer && lane - d >= st && !morton_less16(lo, vl)) rank++;
      if (longer && lane + d <= en && morton_less16(vr, lo)) rank++;
    }
    if (longer) dest = st + rank;
  }
  return dest;
}

// Forward permutation over the wave: this lane's box and id go to lane `dest`
// (dest is a permutation of the lanes).
__device__ __forceinline__ void wave_permute(int dest, float4& b, int& v) {
  const int a = dest << 2;
  b.x = __int_as_float(__builtin_amdgcn_ds_permute(a, __float_as_int(b.x)));
  b.y = __int_as_float(__builtin_amdgcn_ds_permute(a, __float_as_int(b.y)));
  b.z = __int_as_float(__builtin_amdgcn_ds_permute(a, __float_as_int(b.z)));
  b.w = __int_as_float(__builtin_amdgcn_ds_permute(a, __float_as_int(b.w)));
  v = __builtin_amdgcn_ds_permute(a, v);
}

// Moves the ranked tile of the payload sort: ids, then the two box halves,
// each reordered through LDS and scattered in runs of equal digits.  On entry
// the low halves wait in LDS at each item's tile index (s_buf as float2) and
// the high halves in vhi.  The ids are loaded here, after the ranking (FIRST:
// the caller's array, later passes the work array), so that their loads do not
// hold registers through the ranking.  FULL: every item of the tile exists (no
// predicated memory operations, so nothing waits on earlier stores at block
// joins).
template <bool FULL, bool FIRST>
__device__ __forceinline__ void sweep_move(
    const uint32_t (&pos)[kSweepItems], const uint32_t (&dig)[kSweepItems],
    float2 (&vhi)[kSweepItems], int64_t wbase, int64_t n, int tn, bool last,
    float* __restrict__ xmin, float* __restrict__ ymin, float* __restrict__ xmax,
    float* __restrict__ ymax, int* __restrict__ id, const uint32_t* __restrict__ idin,
    uint32_t* __restrict__ idout, float2* __restrict__ lout, float2* __restrict__ hout,
    SweepShared& sh, uint64_t* s_buf, uint8_t* s_dig) {
  const int t = threadIdx.x, lane = lane_id();
  const int64_t tbase = wbase - (threadIdx.x >> 6) * (kWave * kSweepItems);
  float2 vlo[kSweepItems];
  uint32_t vid[kSweepItems];
  {
    const int64_t b0 = wbase + lane;
#pragma unroll
    for (int i = 0; i < kSweepItems; i++) {
      const bool ok = FULL || b0 + i * kWave < n;
      vid[i] = ok ? (FIRST ? (uint32_t)id[b0 + i * kWave] : idin[b0 + i * kWave]) : 0u;
      vlo[i] = ((const float2*)s_buf)[(wbase - tbase) + i * kWave + lane];
    }
  }
  __syncthreads();  // every stashed half is read before the ids overwrite LDS
  uint32_t* s_id = (uint32_t*)s_buf;
#pragma unroll
  for (int i = 0; i < kSweepItems; i++)
    if (FULL || wbase + i * kWave + lane < n) {
      s_id[pos[i]] = vid[i];
      s_dig[pos[i]] = (uint8_t)dig[i];
    }
  __syncthreads();
  if (last) {  // the ids, into the caller's array
#pragma unroll
    for (int i = 0; i < kSweepItems; i++) {
      const int j = t + i * kSortThreads;
      if (FULL || j < tn) id[sh.adj[s_dig[j]] + (uint32_t)j] = (int)s_id[j];
    }
  } else {
#pragma unroll
    for (int i = 0; i < kSweepItems; i++) {
      const int j = t + i * kSortThreads;
      if (FULL || j < tn) idout[sh.adj[s_dig[j]] + (uint32_t)j] = s_id[j];
    }
  }
  float2* s_f2 = (float2*)s_buf;
#pragma unroll
  for (int half = 0; half < 2; half++) {
    __syncthreads();  // the previous field's LDS reads are done
#pragma unroll
    for (int i = 0; i < kSweepItems; i++)
      if (FULL || wbase + i * kWave + lane < n) s_f2[pos[i]] = half ? vhi[i] : vlo[i];
    __syncthreads();
    if (last) {  // one coordinate array each
      float* dx = half ? xmax : xmin;
      float* dy = half ? ymax : ymin;
#pragma unroll
      for (int i = 0; i < kSweepItems; i++) {
        const int j = t + i * kSortThreads;
        if (FULL || j < tn) {
          const float2 v = s_f2[j];
          const uint32_t o = sh.adj[s_dig[j]] + (uint32_t)j;
          dx[o] = v.x;
          dy[o] = v.y;
        }
      }
    } else {
      float2* hdst = half ? hout : lout;
#pragma unroll
      for (int i = 0; i < kSweepItems; i++) {
        const int j = t + i * kSortThreads;
        if (FULL || j < tn) hdst[sh.adj[s_dig[j]] + (uint32_t)j] = s_f2[j];
      }
    }
  }
}

// One 8-bit pass of the payload sort over the 32-bit sort key of MODE.  An
// item is the box and its id (40 B moved per box); the key is computed again
// from the box in every pass instead of being carried.  FIRST (step 0, which
// the plan always runs) reads the caller's arrays; the last pass writes the
// ids and boxes into the caller's arrays in the order of the key (the plan
// runs at least two passes, so the last never reads what it writes).
template <bool FIRST, int MODE>
__global__ __launch_bounds__(kSortThreads) void k_sweep_kv(
    int64_t n, int step, float* __restrict__ xmin, float* __restrict__ ymin,
    float* __restrict__ xmax, float* __restrict__ ymax, int* __restrict__ id,
    uint32_t* __restrict__ id0, uint32_t* __restrict__ id1, float2* __restrict__ lo0,
    float2* __restrict__ lo1, float2* __restrict__ hi0, float2* __restrict__ hi1, BuildCtrl* c,
    uint32_t* __restrict__ status) {
  __shared__ SweepShared sh;
  __shared__ uint64_t s_buf[kSweepTile];
  __shared__ uint8_t s_dig[kSweepTile];

  if (!c->active[step]) return;
  const int t = threadIdx.x, w = t >> 6, lane = lane_id();
  const int pass = c->digit[step];
  const int shift = 8 * pass;
  const bool last = c->write_keys[step] == 0;
  const KeyParams kp = key_params<MODE>(c);
  if (t == 0) sh.tile = atomicAdd(&c->ticket[step], 1u);
  for (int i = t; i < kSortWaves * 256; i += kSortThreads) {
    (&sh.wh[0][0])[i] = 0;
    (&sh.match[0][0])[i] = 0;
  }
  __syncthreads();
  const int64_t tile = sh.tile;
  const int src = c->src[step];
  const uint32_t* idin = src ? id1 : id0;
  uint32_t* idout = src ? id0 : id1;
  const float2* lin = src ? lo1 : lo0;
  float2* lout = src ? lo0 : lo1;
  const float2* hin = src ? hi1 : hi0;
  float2* hout = src ? hi0 : hi1;

  const int64_t tbase = tile * kSweepTile;
  const int64_t wbase = tbase + (int64_t)w * (kWave * kSweepItems);
  uint32_t dig[kSweepItems], pos[kSweepItems];
  float2 vlo[kSweepItems], vhi[kSweepItems];
  // the boxes (each load loop issues all of its loads before the first use;
  // one base per array and constant offsets: few address registers)
  const int64_t b0 = wbase + lane;
  if (FIRST) {
    const float *px0 = xmin + b0, *py0 = ymin + b0, *px1 = xmax + b0, *py1 = ymax + b0;
#pragma unroll
    for (int i = 0; i < kSweepItems; i++) {
      const bool ok = b0 + i * kWave < n;
      vlo[i] = ok ? make_float2(px0[i * kWave], py0[i * kWave]) : make_float2(0.f, 0.f);
      vhi[i] = ok ? make_float2(px1[i * kWave], py1[i * kWave]) : make_float2(0.f, 0.f);
    }
  } else {
    const float2 *pl = lin + b0, *ph = hin + b0;
#pragma unroll
    for (int i = 0; i < kSweepItems; i++) {
      const bool ok = b0 + i * kWave < n;
      vlo[i] = ok ? pl[i * kWave] : make_float2(0.f, 0.f);
      vhi[i] = ok ? ph[i * kWave] : make_float2(0.f, 0.f);
    }
  }
#pragma unroll
  for (int i = 0; i < kSweepItems; i++) {
    dig[i] = (sort_key32<MODE>(vlo[i].x, vlo[i].y, vhi[i].x, vhi[i].y, kp) >> shift) & 255u;
    // the low halves wait in LDS (unused until the moves) while the tile is
    // ranked: registers for all of the box would not fit
    ((float2*)s_buf)[(wbase - tbase) + i * kWave + lane] = vlo[i];
  }
  sweep_rank(dig, pos, wbase, n, tile, c->base + pass * 256, status, sh);

  const int64_t rem = n - tbase;
  if (rem >= kSweepTile)  // full tile: no predicated loads or stores
    sweep_move<true, FIRST>(pos, dig, vhi, wbase, n, kSweepTile, last, xmin, ymin, xmax, ymax,
                            id, idin, idout, lout, hout, sh, s_buf, s_dig);
  else
    sweep_move<false, FIRST>(pos, dig, vhi, wbase, n, (int)rem, last, xmin, ymin, xmax, ymax, id,
                             idin, idout, lout, hout, sh, s_buf, s_dig);
}

// ---------------------------------------------------------------------------
// Ties of the high half (64-bit Morton keys).  After the high-half passes the
// caller's arrays are sorted by the high 32 bits, and a stable sort leaves each
// run of equal high halves in input order.  Sorting every run by (low half,
// position in the run) then gives exactly the order of a stable sort on the
// full key, so the four low-digit passes are not needed.  With n boxes spread
// over 2^32 high-half cells runs are short.  The runs are sorted in place:
//   k_final sorts, within one wave, every run that lies inside one chunk of 64
// positions (a wave's lanes), unless it holds the first or last position of
// the block's tile: the neighbouring blocks read those two positions, so they
// never change during the launch.  Every other run is listed by the chunk
// that holds its start; k_run_len measures the listed runs (reading only),
// k_fixup_short sorts those of up to 64 boxes (one wave each), k_fixup_long
// those of up to kWaveSortMax (one wave each) and kFixLong (one block each),
// with an LSD sort in registers and LDS, k_fixup_huge the longer ones still
// (one block each, in LDS after a split by the top digit), and
// k_fixup_groups recomputes the leaf-parent nodes over them.  A run too long
// for one block, or a full list, sets `fallback`: the full 8-pass LSD sort of
// 64-bit keys then runs from the order reached so far (in which equal full
// keys are still in input order).
constexpr int kFixLong = 4096;
constexpr int kWaveSortMax = 1024;   // runs sorted by one wave (16 items per lane)
constexpr int kFixSlots = 4096;      // blocks of the fixup launches (they loop over the list)
constexpr int kFixLongSlots = 1024;  // blocks of k_fixup_long (four per CU)
constexpr int kHugeThreads = 512;    // k_fixup_huge: threads per block
constexpr int kHugeWaves = kHugeThreads / 64;
constexpr int kFixHugeSlots = 512;   // blocks of k_fixup_huge (two per CU)
constexpr int kHugeK = 8;            // k_fixup_huge: positions per thread in one tile
constexpr int kHugeLds = kHugeWaves * 64 * 16;  // k_fixup_huge: longest run sorted whole in LDS
#ifndef CT_FINAL_GRID
#define CT_FINAL_GRID 4096
#endif
constexpr int kFinTile = 2048;       // positions owned by one block of k_final (multiple of 64)

// Length of every listed run and the lists of the long ones.  Each wave takes
// 64 listed runs (one per lane); the open ones (length 0: they reached the end
// of a k_final tile) are measured one at a time by the whole wave, which reads
// the boxes only, so every wave reads stable data.  Runs of 65..kWaveSortMax
// boxes are listed for the waves of k_fixup_long, runs of up to kFixLong for
// its blocks, longer ones (counted only up to kFixLong + 1) for k_fixup_huge
// from the far end of the first list, with one atomic per wave and list.
__global__ __launch_bounds__(256) void k_run_len(int64_t n, BuildCtrl* c,
                                                 const float* __restrict__ xmin,
                                                 const float* __restrict__ ymin,
                                                 const float* __restrict__ xmax,
                                                 const float* __restrict__ ymax, uint32_t* runs,
                                                 uint32_t runs_cap) {
  const uint32_t nr = c->nruns < runs_cap ? c->nruns : runs_cap;
  if (blockIdx.x * 256u >= nr) return;
  const KeyParams kp = key_params<kKeyMorton64>(c);
  const int lane = lane_id();
  const uint64_t lt = lanemask_lt();
  const uint32_t wave = (blockIdx.x * 256u + threadIdx.x) >> 6, nw = gridDim.x * 4u;
  for (uint32_t q0 = wave * 64u; q0 < nr; q0 += nw * 64u) {
    const uint32_t q = q0 + lane;
    const bool has = q < nr;
    int L = has ? (int)runs[runs_cap + q] : 1;
    const bool was_open = has && L == 0;
    for (uint64_t open = ballot(was_open); open; open &= open - 1) {
      const int j = __ffsll((long long)open) - 1;
      const int64_t i = runs[q0 + j];
      const uint32_t h0 = (uint32_t)(morton64(xmin[i], ymin[i], xmax[i], ymax[i], kp) >> 32);
      int Lj = kFixLong + 1;
      // four chunks of 64 positions per round (independent loads in flight)
      for (int64_t b = i; Lj > kFixLong && b - i <= kFixLong; b += 4 * kWave) {
        bool eq[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int64_t p = b + kWave * u + lane;
          eq[u] = p < n &&
                  (uint32_t)(morton64(xmin[p], ymin[p], xmax[p], ymax[p], kp) >> 32) == h0;
        }
#pragma unroll
        for (int u = 3; u >= 0; u--) {
          const uint64_t ne = ballot(!eq[u]);
          if (ne) Lj = (int)(b - i) + kWave * u + __ffsll((long long)ne) - 1;
        }
      }
      if (lane == j) L = Lj;
    }
    if (was_open) runs[runs_cap + q] = (uint32_t)L;
    const bool mid = has && L > kWave && L <= kWaveSortMax;
    const bool lng = has && L > kWaveSortMax && L <= kFixLong, hug = has && L > kFixLong;
    const uint64_t mm = ballot(mid), ml = ballot(lng), mh = ballot(hug);
    uint32_t bm = 0, bl = 0, bh = 0;
    if (lane == 0) {
      if (mm) bm = atomicAdd(&c->nmid, (uint32_t)popc64(mm));
      if (ml) bl = atomicAdd(&c->nlong, (uint32_t)popc64(ml));
      if (mh) bh = atomicAdd(&c->nhuge, (uint32_t)popc64(mh));
    }
    bm = __shfl(bm, 0);
    bl = __shfl(bl, 0);
    bh = __shfl(bh, 0);
    if (mid) runs[2 * (size_t)runs_cap + bm + popc64(mm & lt)] = q;
    if (lng) runs[3 * (size_t)runs_cap + bl + popc64(ml & lt)] = q;
    if (hug) runs[3 * (size_t)runs_cap - 1 - (bh + popc64(mh & lt))] = q;
  }
}

// Listed runs of up to 64 boxes: eight runs of up to 8 per wave, longer ones
// one per wave (lane l holds position i + l of its run).
__global__ __launch_bounds__(256) void k_fixup_short(int64_t n, BuildCtrl* c,
                                                     float* __restrict__ xmin,
                                                     float* __restrict__ ymin,
                                                     float* __restrict__ xmax,
                                                     float* __restrict__ ymax,
                                                     int* __restrict__ id,
                                                     const uint32_t* __restrict__ runs,
                                                     uint32_t runs_cap) {
  const uint32_t nr = c->nruns < runs_cap ? c->nruns : runs_cap;
  if (blockIdx.x * 4u >= nr || c->fallback) return;
  const KeyParams kp = key_params<kKeyMorton64>(c);
  const int lane = lane_id();
  const uint32_t wave = (blockIdx.x * 256u + threadIdx.x) >> 6, nw = gridDim.x * 4u;
  // lane `member` of a run at position p: sort the wave's runs, move what moved
  auto settle = [&](bool member, int st, int en, int64_t p) {
    float4 b = make_float4(0.f, 0.f, 0.f, 0.f);
    if (member) b = make_float4(xmin[p], ymin[p], xmax[p], ymax[p]);
    uint32_t lo;
    (void)morton_cell(b.x, b.y, b.z, b.w, kp, &lo);
    const int dest = local_run_dest(member, st, en, lo);
    int v = dest != lane ? id[p] : 0;
    wave_permute(dest, b, v);
    if (dest != lane) {
      xmin[p] = b.x;
      ymin[p] = b.y;
      xmax[p] = b.z;
      ymax[p] = b.w;
      id[p] = v;
    }
  };
  // runs of up to kSeg boxes (nearly all): kWave / kSeg of them per wave,
  // lanes kSeg*j .. kSeg*j+kSeg-1 holding run j
  constexpr int kSeg = 8, kPerWave = kWave / kSeg;
  for (uint32_t q0 = wave * kPerWave; q0 < nr; q0 += nw * kPerWave) {
    const int j = lane / kSeg, i = lane % kSeg;
    const uint32_t q = q0 + j;
    const int L = q < nr ? (int)runs[runs_cap + q] : 0;
    const bool member = L <= kSeg && i < L;
    settle(member, j * kSeg, j * kSeg + L - 1, member ? (int64_t)runs[q] + i : 0);
  }
  // runs of kSeg+1 .. kWave boxes: one per wave
  for (uint32_t q = wave; q < nr; q += nw) {
    const int L = (int)runs[runs_cap + q];
    if (L <= kSeg || L > kWave) continue;
    settle(lane < L, 0, L - 1, (int64_t)runs[q] + lane);
  }
}

// Four stable LSD passes over 8-bit digits of the high words of x (the low
// half of the 64-bit key above the index in the run), with a block of NW
// waves: wave w, item k, lane l holds index wbase + 64 * k with
// wbase = w * 64 * kUse + l; slots past the run hold ~0, which sorts after
// every real key.  Ranks from ballots and per-wave running counts, offsets
// from a scan of the counts, the scatter and the reload through LDS.  Every
// thread of the block calls it.
constexpr int kRunPer = 16;  // items per lane of the LDS run sorts
// One stable pass of lds_lsd over the 8-bit digit at bit `sh` of the items
// (T: 64-bit items, a key above an index, or 32-bit keys alone).  On return
// s_cnt[0][d] is the first position of digit d in the reordered items, and
// x[k] the item at position wbase + 64 * k.  s_match (or null: ballots):
// the waves' peer_mask tables, zero on entry and on return.
template <int NW, typename T>
__device__ __forceinline__ void lds_pass(T (&x)[kRunPer], int kUse, int wbase, int sh, T* s_net,
                                         uint32_t (*s_cnt)[256], uint32_t* s_wsum,
                                         uint64_t (*s_match)[256] = nullptr) {
  const int t = threadIdx.x, lane = lane_id(), w = t >> 6;
  const uint64_t lt = lanemask_lt();
#pragma unroll
  for (int k = 0; k < 4; k++) s_cnt[w][lane + 64 * k] = 0;
  uint32_t rank[kRunPer];
#pragma unroll
  for (int k = 0; k < kRunPer; k++) {
    if (k < kUse) {
      const uint32_t d = (uint32_t)(x[k] >> sh) & 255u;
      uint64_t m = ~0ull;
      if (s_match) {
        m = peer_mask(d, true, s_match[w]);
      } else {
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
          const uint64_t bb = ballot((d >> bit) & 1u);
          m &= ((d >> bit) & 1u) ? bb : ~bb;
        }
      }
      const uint32_t before = s_cnt[w][d];
      rank[k] = before + (uint32_t)popc64(m & lt);
      if ((m & lt) == 0) s_cnt[w][d] = before + (uint32_t)popc64(m);
    }
  }
  __syncthreads();
  // digit t (threads 0..255): offsets of the waves' shares, after all
  // smaller digits
  uint32_t cw[NW], tot = 0, inc = 0;
  if (t < 256) {
#pragma unroll
    for (int k = 0; k < NW; k++) {
      cw[k] = s_cnt[k][t];
      tot += cw[k];
    }
    inc = tot;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t u = __shfl_up(inc, d);
      if (lane >= d) inc += u;
    }
    if (lane == 63) s_wsum[w] = inc;
  }
  __syncthreads();
  if (t < 256) {
    uint32_t off = inc - tot;
    for (int k = 0; k < w; k++) off += s_wsum[k];
#pragma unroll
    for (int k = 0; k < NW; k++) {
      s_cnt[k][t] = off;
      off += cw[k];
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kRunPer; k++)
    if (k < kUse) s_net[s_cnt[w][(uint32_t)(x[k] >> sh) & 255u] + rank[k]] = x[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kRunPer; k++)
    if (k < kUse) x[k] = s_net[wbase + 64 * k];
}

// (64-bit items: the four digits of the high word; 32-bit keys: of the key)
template <int NW, typename T>
__device__ __forceinline__ void lds_lsd(T (&x)[kRunPer], int kUse, int wbase, T* s_net,
                                        uint32_t (*s_cnt)[256], uint32_t* s_wsum,
                                        uint64_t (*s_match)[256] = nullptr) {
  // (a pass reuses only its own wave's row of s_cnt before its first barrier,
  // and writes s_net only after two, so passes need no barrier between them)
  constexpr int sh0 = sizeof(T) == 8 ? 32 : 0;
  for (int pass = 0; pass < 4; pass++)
    lds_pass<NW, T>(x, kUse, wbase, sh0 + 8 * pass, s_net, s_cnt, s_wsum, s_match);
}

// Sorts the run of L <= NW * 64 * kRunPer boxes at position i by (low half,
// index in the run), in place, with a block of NW waves (lds_lsd over the
// keys computed from the boxes); the moved boxes come from a copy of the run
// in its span of the staging records.  Every thread of the block calls it; it ends with a
// barrier.
template <int NW>
__device__ __forceinline__ void lds_sort_run(int64_t i, int L, const KeyParams& kp,
                                             float* __restrict__ xmin, float* __restrict__ ymin,
                                             float* __restrict__ xmax, float* __restrict__ ymax,
                                             int* __restrict__ id, float4* __restrict__ stage,
                                             uint32_t* __restrict__ stage_id, uint64_t* s_net,
                                             uint32_t (*s_cnt)[256], uint32_t* s_wsum,
                                             uint64_t (*s_match)[256] = nullptr) {
  const int kUse = (L + NW * 64 - 1) / (NW * 64);  // items per lane
  const int wbase = (threadIdx.x >> 6) * 64 * kUse + lane_id();
  uint64_t x[kRunPer];
  // keys from the boxes, and the run copied (coalesced) into its span of the
  // staging records, from which the moved positions read their source
#pragma unroll
  for (int k = 0; k < kRunPer; k++) {
    x[k] = ~0ull;
    const int e = wbase + 64 * k;
    if (k < kUse && e < L) {
      const int64_t p = i + e;
      const float4 b = make_float4(xmin[p], ymin[p], xmax[p], ymax[p]);
      stage[p] = b;
      stage_id[p] = (uint32_t)id[p];
      x[k] = ((uint64_t)(uint32_t)morton64(b.x, b.y, b.z, b.w, kp) << 32) | (uint32_t)e;
    }
  }
  lds_lsd<NW, uint64_t>(x, kUse, wbase, s_net, s_cnt, s_wsum, s_match);  // (its barriers order the copy)
#pragma unroll
  for (int k = 0; k < kRunPer; k++) {
    const int e = wbase + 64 * k;
    if (k < kUse && e < L && (uint32_t)x[k] != (uint32_t)e) {
      const int64_t p = i + e, sp = i + (uint32_t)x[k];
      const float4 b = stage[sp];
      xmin[p] = b.x;
      ymin[p] = b.y;
      xmax[p] = b.z;
      ymax[p] = b.w;
      id[p] = (int)stage_id[sp];
    }
  }
  __syncthreads();  // s_net and s_cnt are reused by the next run
}

// Sorts L <= NW * 64 * kRunPer keys (low half << 32 | index) from `src` into
// `dst` with a block of NW waves (lds_lsd).  Every thread of the block calls
// it; it ends with a barrier.
template <int NW>
__device__ __forceinline__ void lds_sort_keys(const uint64_t* __restrict__ src,
                                              uint64_t* __restrict__ dst, int L,
                                              uint64_t* s_net, uint32_t (*s_cnt)[256],
                                              uint32_t* s_wsum) {
  const int kUse = (L + NW * 64 - 1) / (NW * 64);
  const int wbase = (threadIdx.x >> 6) * 64 * kUse + lane_id();
  uint64_t x[kRunPer];
#pragma unroll
  for (int k = 0; k < kRunPer; k++) {
    const int e = wbase + 64 * k;
    x[k] = k < kUse && e < L ? src[e] : ~0ull;
  }
  lds_lsd<NW, uint64_t>(x, kUse, wbase, s_net, s_cnt, s_wsum);
#pragma unroll
  for (int k = 0; k < kRunPer; k++) {
    const int e = wbase + 64 * k;
    if (k < kUse && e < L) dst[e] = x[k];
  }
  __syncthreads();  // s_net and s_cnt are reused
}

// Sorts the run of 65..kWaveSortMax boxes at position i like lds_sort_run,
// with one wave and no barriers: lane l, item k holds index 64 * k + l; the
// wave's own LDS rows hold the counts (lane l scans digits 4l..4l+3) and the
// scatter.  LDS written by one lane and read by another is ordered by the
// wave's program order (LDS operations of one wave complete in order).
__device__ __forceinline__ void wave_sort_run(int64_t i, int L, const KeyParams& kp,
                                              float* __restrict__ xmin, float* __restrict__ ymin,
                                              float* __restrict__ xmax, float* __restrict__ ymax,
                                              int* __restrict__ id, float4* __restrict__ stage,
                                              uint32_t* __restrict__ stage_id, uint64_t* s_w,
                                              uint32_t* s_c) {
  const int lane = lane_id();
  const uint64_t lt = lanemask_lt();
  const int kUse = (L + 63) >> 6;
  uint64_t x[kRunPer];
  // keys from the boxes, and the run copied (coalesced) into its span of the
  // staging records, from which the moved positions read their source
#pragma unroll
  for (int k = 0; k < kRunPer; k++) {
    x[k] = ~0ull;
    const int e = 64 * k + lane;
    if (k < kUse && e < L) {
      const int64_t p = i + e;
      const float4 b = make_float4(xmin[p], ymin[p], xmax[p], ymax[p]);
      stage[p] = b;
      stage_id[p] = (uint32_t)id[p];
      x[k] = ((uint64_t)(uint32_t)morton64(b.x, b.y, b.z, b.w, kp) << 32) | (uint32_t)e;
    }
  }
  for (int pass = 0; pass < 4; pass++) {
    const int sh = 32 + 8 * pass;
#pragma unroll
    for (int k = 0; k < 4; k++) s_c[lane + 64 * k] = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    uint32_t rank[kRunPer];
#pragma unroll
    for (int k = 0; k < kRunPer; k++) {
      if (k < kUse) {
        const uint32_t d = (uint32_t)(x[k] >> sh) & 255u;
        // (ballots, not peer_mask: a run's keys share digits, and lanes that
        // OR into one LDS entry are served one after another)
        uint64_t m = ~0ull;
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
          const uint64_t bb = ballot((d >> bit) & 1u);
          m &= ((d >> bit) & 1u) ? bb : ~bb;
        }
        const uint32_t before = s_c[d];
        rank[k] = before + (uint32_t)popc64(m & lt);
        if ((m & lt) == 0) s_c[d] = before + (uint32_t)popc64(m);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
    }
    const uint4 cnt = *(const uint4*)&s_c[4 * lane];
    const uint32_t tot = cnt.x + cnt.y + cnt.z + cnt.w;
    uint32_t inc = tot;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t u = __shfl_up(inc, d);
      if (lane >= d) inc += u;
    }
    const uint32_t o = inc - tot;
    *(uint4*)&s_c[4 * lane] = make_uint4(o, o + cnt.x, o + cnt.x + cnt.y,
                                         o + cnt.x + cnt.y + cnt.z);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int k = 0; k < kRunPer; k++)
      if (k < kUse) s_w[s_c[(uint32_t)(x[k] >> sh) & 255u] + rank[k]] = x[k];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int k = 0; k < kRunPer; k++)
      if (k < kUse) x[k] = s_w[64 * k + lane];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  // the copy's stores complete before the moved positions read other lanes'
  // records (a fence at work-group scope waits for them)
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int k = 0; k < kRunPer; k++) {
    const int e = 64 * k + lane;
    if (k < kUse && e < L && (uint32_t)x[k] != (uint32_t)e) {
      const int64_t p = i + e, sp = i + (uint32_t)x[k];
      const float4 b = stage[sp];
      xmin[p] = b.x;
      ymin[p] = b.y;
      xmax[p] = b.z;
      ymax[p] = b.w;
      id[p] = (int)stage_id[sp];
    }
  }
}

// Listed runs of 1025..kFixLong boxes, one block of four waves each, then the
// listed runs of 65..kWaveSortMax boxes, one wave each.
static_assert(kFixLong == 4 * 64 * kRunPer, "k_fixup_long: one item slot per box");
static_assert(kWaveSortMax == 64 * kRunPer, "wave_sort_run: one item slot per box");
static_assert(kHugeLds == kHugeWaves * 64 * kRunPer, "k_fixup_huge: one item slot per box");
__global__ __launch_bounds__(256) void k_fixup_long(int64_t n, BuildCtrl* c,
                                                    float* __restrict__ xmin,
                                                    float* __restrict__ ymin,
                                                    float* __restrict__ xmax,
                                                    float* __restrict__ ymax,
                                                    int* __restrict__ id,
                                                    const uint32_t* __restrict__ runs,
                                                    uint32_t runs_cap,
                                                    float4* __restrict__ stage,
                                                    uint32_t* __restrict__ stage_id) {
  __shared__ uint64_t s_net[kFixLong];
  __shared__ __align__(16) uint32_t s_cnt[4][256];  // per wave, per digit: count -> offset
  __shared__ uint32_t s_wsum[4];
  const uint32_t nl = c->nlong, nm = c->nmid;
  if ((blockIdx.x >= nl && blockIdx.x * 4u >= nm) || c->fallback) return;
  const KeyParams kp = key_params<kKeyMorton64>(c);
  // (ballot ranking: the keys of one run share digits, which would serialise
  // peer_mask's LDS ORs)
  for (uint32_t r = blockIdx.x; r < nl; r += gridDim.x) {
    const uint32_t q = runs[3 * (size_t)runs_cap + r];
    lds_sort_run<4>(runs[q], (int)runs[runs_cap + q], kp, xmin, ymin, xmax, ymax, id, stage,
                    stage_id, s_net, s_cnt, s_wsum);
  }
  const int w = threadIdx.x >> 6;
  for (uint32_t r = blockIdx.x * 4u + w; r < nm; r += gridDim.x * 4u) {
    const uint32_t q = runs[2 * (size_t)runs_cap + r];
    wave_sort_run(runs[q], (int)runs[runs_cap + q], kp, xmin, ymin, xmax, ymax, id, stage,
                  stage_id, s_net + kWaveSortMax * w, s_cnt[w]);
  }
}

// Listed runs longer than kFixLong, one block each, claimed by ticket.  The
// block measures the run (k_run_len stopped at kFixLong + 1) and records its
// length for k_fixup_groups.  A run of up to kHugeLds boxes is sorted in LDS
// (lds_sort_run).  A longer one is split stably by the top 8 bits of the low
// half (one LSD-style scatter of (low half << 32 | index in the run) over the
// run's span of the free key buffers), groups of whole digits of up to
// kHugeLds keys are sorted in LDS (lds_sort_keys), and the boxes are moved
// once, through staging records; if one digit alone holds more, four LSD
// passes over the whole run in the key buffers sort it.  A run longer than `hmax` (one block would take
// longer than the full 64-bit sort) sets `fallback` and stays as it is; every
// other run is settled completely, so the fallback may start from any mix of
// the two.
// (two blocks per CU: at most 128 VGPRs)
__global__ __launch_bounds__(kHugeThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_fixup_huge(
    int64_t n, BuildCtrl* c, float* __restrict__ xmin, float* __restrict__ ymin,
    float* __restrict__ xmax, float* __restrict__ ymax, int* __restrict__ id,
    uint32_t* __restrict__ runs, uint32_t runs_cap, uint64_t* __restrict__ ka,
    uint64_t* __restrict__ kb, float4* __restrict__ stage, uint32_t* __restrict__ stage_id,
    int64_t hmax) {
  __shared__ uint32_t s_hist[2][256];  // digit counts of this pass / the next
  __shared__ uint32_t s_base[256];     // running offset of each digit
  __shared__ uint32_t s_bstart[256];   // first offset of each digit (the split pass)
  __shared__ uint32_t s_wcnt[kHugeWaves][256];  // a tile's count per (wave, digit) -> offset
  __shared__ uint32_t s_wsum[4];
  __shared__ uint32_t s_claim, s_len, s_nchunk, s_big;
  __shared__ uint32_t s_chunk[258];
  __shared__ uint64_t s_net[kHugeLds];
  const uint32_t nh = c->nhuge;
  if (blockIdx.x >= nh || c->fallback) return;
  const KeyParams kp = key_params<kKeyMorton64>(c);
  const int t = threadIdx.x, lane = lane_id(), w = t >> 6;
  const uint64_t lt = lanemask_lt();

  // (low half << 32 | index) of the run's boxes into ka, and the counts of
  // the digit at `hsh` of the low half into s_hist[0]
  auto load_keys = [&](int64_t i, uint32_t L, int hsh) {
    if (t < 256) {
      s_hist[0][t] = 0;
      s_hist[1][t] = 0;
    }
    __syncthreads();
    // four positions per thread and round: the loads of a round in flight together
    for (uint32_t e0 = t; e0 < L; e0 += 4 * kHugeThreads) {
      uint32_t lo[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const uint32_t e = e0 + kHugeThreads * u;
        const int64_t p = i + e;
        lo[u] = e < L ? (uint32_t)morton64(xmin[p], ymin[p], xmax[p], ymax[p], kp) : 0u;
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const uint32_t e = e0 + kHugeThreads * u;
        if (e < L) {
          ka[i + e] = (uint64_t)lo[u] << 32 | e;
          atomicAdd(&s_hist[0][(lo[u] >> hsh) & 255u], 1u);
        }
      }
    }
  };
  // one stable scatter of the run's keys by the digit at bit `sh`, src -> dst,
  // from the counts in s_hist[h]; with `next`, the counts of the following
  // digit go to s_hist[h ^ 1]; with `starts`, the digits' first offsets to
  // s_bstart
  auto lsd_pass = [&](int64_t i, uint32_t L, const uint64_t* src, uint64_t* dst, int sh, int h,
                      bool next, bool starts) {
    __syncthreads();  // s_hist[h] is complete
    uint32_t v = 0, inc = 0;
    if (t < 256) {
      v = s_hist[h][t];
      inc = v;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t u = __shfl_up(inc, d);
        if (lane >= d) inc += u;
      }
      if (lane == 63) s_wsum[w] = inc;
      s_hist[h ^ 1][t] = 0;
    }
    __syncthreads();
    if (t < 256) {
      uint32_t off = 0;
      for (int k = 0; k < w; k++) off += s_wsum[k];
      s_base[t] = off + inc - v;
      if (starts) s_bstart[t] = off + inc - v;
    }
    for (uint32_t e0 = 0; e0 < L; e0 += kHugeThreads * kHugeK) {
      // wave w, item k, lane l: index e0 + w * 64 * kHugeK + 64 * k + l
      const uint32_t eb = e0 + (uint32_t)(w * 64 * kHugeK + lane);
      uint64_t x[kHugeK];
      uint32_t rank[kHugeK];
#pragma unroll
      for (int k = 0; k < kHugeK; k++) {
        const uint32_t e = eb + 64u * k;
        x[k] = e < L ? src[i + e] : 0ull;
      }
#pragma unroll
      for (int k = 0; k < 4; k++) s_wcnt[w][lane + 64 * k] = 0;
#pragma unroll
      for (int k = 0; k < kHugeK; k++) {
        const bool valid = eb + 64u * k < L;
        const uint32_t d = (uint32_t)(x[k] >> sh) & 255u;
        uint64_t m = ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
          const uint64_t bb = ballot((d >> bit) & 1u);
          m &= ((d >> bit) & 1u) ? bb : ~bb;
        }
        const uint32_t before = s_wcnt[w][d];
        rank[k] = before + (uint32_t)popc64(m & lt);
        if (valid && (m & lt) == 0) s_wcnt[w][d] = before + (uint32_t)popc64(m);
      }
      __syncthreads();
      if (t < 256) {
        uint32_t run = s_base[t];
#pragma unroll
        for (int k = 0; k < kHugeWaves; k++) {
          const uint32_t cnt = s_wcnt[k][t];
          s_wcnt[k][t] = run;
          run += cnt;
        }
        s_base[t] = run;
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kHugeK; k++) {
        if (eb + 64u * k < L) {
          dst[i + s_wcnt[w][(uint32_t)(x[k] >> sh) & 255u] + rank[k]] = x[k];
          if (next)
            atomicAdd(&s_hist[h ^ 1][(uint32_t)(x[k] >> (sh + 8)) & 255u], 1u);
        }
      }
      __syncthreads();  // s_wcnt is cleared by the next tile
    }
  };
  // the run's boxes put in the order of `k` (the index halves): a coalesced
  // copy of the run into its span of the staging records, then each position
  // reads its source record (two random reads instead of one per array)
  auto gather = [&](int64_t i, uint32_t L, const uint64_t* k) {
    for (uint32_t e0 = t; e0 < L; e0 += 4 * kHugeThreads) {
      float4 b[4];
      int v[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const uint32_t e = e0 + kHugeThreads * u;
        if (e < L) {
          const int64_t p = i + e;
          b[u] = make_float4(xmin[p], ymin[p], xmax[p], ymax[p]);
          v[u] = id[p];
        }
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const uint32_t e = e0 + kHugeThreads * u;
        if (e < L) {
          stage[i + e] = b[u];
          stage_id[i + e] = (uint32_t)v[u];
        }
      }
    }
    __syncthreads();
    for (uint32_t e0 = t; e0 < L; e0 += 4 * kHugeThreads) {
      uint32_t src[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const uint32_t e = e0 + kHugeThreads * u;
        src[u] = e < L ? (uint32_t)k[i + e] : e;
      }
      float4 b[4];
      uint32_t v[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        if (src[u] != e0 + kHugeThreads * u) {
          b[u] = stage[i + src[u]];
          v[u] = stage_id[i + src[u]];
        }
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const uint32_t e = e0 + kHugeThreads * u;
        if (src[u] != e) {
          const int64_t p = i + e;
          xmin[p] = b[u].x;
          ymin[p] = b[u].y;
          xmax[p] = b[u].z;
          ymax[p] = b[u].w;
          id[p] = (int)v[u];
        }
      }
    }
    __syncthreads();
  };

  for (;;) {
    if (t == 0) {
      s_claim = atomicAdd(&c->huge_next, 1u);
      s_len = 0xFFFFFFFFu;
    }
    __syncthreads();
    const uint32_t r = s_claim;
    if (r >= nh) break;
    const uint32_t q = runs[3 * (size_t)runs_cap - 1 - r];
    const int64_t i = runs[q];
    const uint32_t h0 = (uint32_t)(morton64(xmin[i], ymin[i], xmax[i], ymax[i], kp) >> 32);
    // the run's end: the first later position with another high half
    const int64_t lim = n - i < hmax + 1 ? n - i : hmax + 1;
    for (int64_t b = kFixLong; b < lim; b += 4 * kHugeThreads) {
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int64_t rel = b + kHugeThreads * u + t, p = i + rel;
        if (rel < lim &&
            (uint32_t)(morton64(xmin[p], ymin[p], xmax[p], ymax[p], kp) >> 32) != h0)
          atomicMin(&s_len, (uint32_t)rel);
      }
      __syncthreads();
      const bool found = s_len != 0xFFFFFFFFu;
      __syncthreads();
      if (found) break;
    }
    const uint32_t L = s_len != 0xFFFFFFFFu ? s_len : (uint32_t)lim;
    if ((int64_t)L > hmax) {
      if (t == 0) atomicOr(&c->fallback, 1u);
      __syncthreads();  // s_claim / s_len are rewritten for the next run
      continue;
    }
    if (t == 0) runs[runs_cap + q] = L;
    if ((int)L <= kHugeLds) {
      lds_sort_run<kHugeWaves>(i, (int)L, kp, xmin, ymin, xmax, ymax, id, stage, stage_id, s_net,
                               s_wcnt, s_wsum);
      continue;
    }
    // a stable split by the top 8 bits of the low half, then groups of whole
    // digits of up to kHugeLds boxes sorted in LDS (within a group the top
    // digits are already in order, and equal keys keep their input order)
    load_keys(i, L, 24);
    lsd_pass(i, L, ka, kb, 56, 0, false, true);
    if (t == 0) {
      uint32_t nc = 0, cs = 0, big = 0;
      for (int d = 0; d < 256; d++) {
        const uint32_t bs = s_bstart[d], be = d < 255 ? s_bstart[d + 1] : L;
        if (be - bs > (uint32_t)kHugeLds) big = 1;
        if (be - cs > (uint32_t)kHugeLds) {
          s_chunk[nc++] = cs;
          cs = bs;
        }
      }
      s_chunk[nc++] = cs;
      s_chunk[nc] = L;
      s_nchunk = nc;
      s_big = big;
    }
    __syncthreads();
    if (!s_big) {
      for (uint32_t g = 0; g < s_nchunk; g++) {
        const uint32_t cs = s_chunk[g], ce = s_chunk[g + 1];
        lds_sort_keys<kHugeWaves>(kb + i + cs, ka + i + cs, (int)(ce - cs), s_net, s_wcnt,
                                  s_wsum);
      }
      gather(i, L, ka);
    } else {
      // one digit holds more than kHugeLds boxes: four LSD passes over the
      // whole run in the key buffers
      load_keys(i, L, 0);
      for (int pass = 0; pass < 4; pass++)
        lsd_pass(i, L, pass & 1 ? kb : ka, pass & 1 ? ka : kb, 32 + 8 * pass, pass & 1, pass < 3,
                 false);
      gather(i, L, ka);
    }
    __syncthreads();  // s_claim / s_len / s_chunk are rewritten for the next run
  }
}

// Status words of the fallback steps and the histograms (cleared only when
// the fallback sort will run).
__global__ __launch_bounds__(256) void k_fallback_zero(BuildCtrl* c, uint32_t* p, int64_t words) {
  if (!c->fallback) return;
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < kMaxDigits * 256; i += 256) c->hist[i] = 0;
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < words; i += (int64_t)gridDim.x * 256)
    p[i] = 0u;
}

// ---------------------------------------------------------------------------
// Group reduction of an MBR over G consecutive lanes (G a power of two <= 64;
// every lane of the group ends with the result).  `has` marks lanes that hold
// a value; invalid lanes are skipped.
//   The reduction runs on ordered integer keys: f2ord maps floats to uint32 in
// float order with -0 below +0, and NaN takes the key that loses every
// comparison (all ones for the min, zero for the max), which an invalid lane
// takes as well.  A u32 min/max over the keys then gives exactly the bits of
// the mbr_min/mbr_max fold: the least (greatest) non-NaN value, and the
// canonical NaN when every value is NaN.  Partners within 16 lanes come from
// DPP lane permutations (quad xor 1, quad xor 2, half-row mirror, row mirror:
// each pairs the lane with one in the other half of its 2-, 4-, 8- or
// 16-lane block), wider ones from ds_bpermute.
__device__ __forceinline__ uint32_t min_key(float x, bool has) {
  return (has && x == x) ? f2ord(x) : 0xFFFFFFFFu;
}
__device__ __forceinline__ uint32_t max_key(float x, bool has) {
  return (has && x == x) ? f2ord(x) : 0u;
}

template <int CTRL>
__device__ __forceinline__ void key_step_dpp(uint32_t& a, uint32_t& b, uint32_t& cc, uint32_t& d) {
  a = min(a, dpp_perm<CTRL>(a));
  b = min(b, dpp_perm<CTRL>(b));
  cc = max(cc, dpp_perm<CTRL>(cc));
  d = max(d, dpp_perm<CTRL>(d));
}

template <int G>
__device__ __forceinline__ void group_mbr(float& a, float& b, float& cc, float& d, bool has) {
  uint32_t ka = min_key(a, has), kb = min_key(b, has);
  uint32_t kc = max_key(cc, has), kd = max_key(d, has);
  if (G >= 2) key_step_dpp<0xB1>(ka, kb, kc, kd);   // quad_perm [1,0,3,2]
  if (G >= 4) key_step_dpp<0x4E>(ka, kb, kc, kd);   // quad_perm [2,3,0,1]
  if (G >= 8) key_step_dpp<0x141>(ka, kb, kc, kd);  // row_half_mirror
  if (G >= 16) key_step_dpp<0x140>(ka, kb, kc, kd); // row_mirror
#pragma unroll
  for (int o = 16; o < G; o <<= 1) {
    ka = min(ka, (uint32_t)__shfl_xor((int)ka, o));
    kb = min(kb, (uint32_t)__shfl_xor((int)kb, o));
    kc = max(kc, (uint32_t)__shfl_xor((int)kc, o));
    kd = max(kd, (uint32_t)__shfl_xor((int)kd, o));
  }
  a = ka == 0xFFFFFFFFu ? canon_nan() : ord2f(ka);
  b = kb == 0xFFFFFFFFu ? canon_nan() : ord2f(kb);
  cc = kc == 0u ? canon_nan() : ord2f(kc);
  d = kd == 0u ? canon_nan() : ord2f(kd);
}

// Leaf-parent node g over sorted boxes [g*G, min(g*G+G, n)) from a group
// reduction over G lanes (lanes of a group hold consecutive positions, the
// group's first lane writes): pos = R_sz + g*G, len = count
// (src/rtree_bulkload_thrust.cuh:136-149).
// (positions, node indices and pos values fit 32 bits: n < 2^30 and
// r_sz + n < 2^31)
template <int G>
__device__ __forceinline__ void leaf_parent(uint32_t i, uint32_t n, float a, float b, float cc,
                                            float d, bool has, const TreeView& tr,
                                            uint32_t l0start, uint32_t r_sz, bool mbr_only) {
  group_mbr<G>(a, b, cc, d, has);
  if (has && (i % G) == 0) {
    const uint32_t node = l0start + i / G;
    tr.xmin[node] = a;
    tr.ymin[node] = b;
    tr.xmax[node] = cc;
    tr.ymax[node] = d;
    if (!mbr_only) {
      tr.pos[node] = (int)(r_sz + i);
      const uint32_t len = n - i;
      tr.len[node] = (int)(len < (uint32_t)G ? len : (uint32_t)G);
    }
  }
}

// Last stage of the payload sort, in place on the caller's arrays (sorted by
// the 32-bit key): for 64-bit keys the runs of equal high halves that lie in
// one chunk of 64 positions sorted by (low half, position) within the wave,
// the others listed; for power-of-two F <= 64 the leaf-parent level.  Thread
// t holds positions tb + t + 256k, so each wave holds chunks of 64 consecutive
// positions, and a group of G <= 64 never leaves its chunk.
constexpr int kFinWaves = 256 / kWave;
constexpr int kFinChunks = kFinTile / kWave;
constexpr int kListCap = 512;  // listed runs buffered per block (at most kFinChunks + 1 per tile)

template <int MODE, int G>
__global__ __launch_bounds__(256) void k_final(int64_t n, BuildCtrl* c, float* __restrict__ xmin,
                                               float* __restrict__ ymin,
                                               float* __restrict__ xmax,
                                               float* __restrict__ ymax, int* __restrict__ id,
                                               TreeView tr, int64_t l0start, int64_t r_sz,
                                               uint32_t* runs, uint32_t runs_cap) {
  constexpr bool M64 = MODE == kKeyMorton64;
  constexpr int kPer = kFinTile / 256;  // positions per thread
  // high halves at each chunk's first and last position, and at the
  // positions just before and after the tile
  __shared__ uint32_t s_eh[2][M64 ? kFinChunks : 1];
  __shared__ uint32_t s_nb[2];
  // listed runs: start position and length (kOpen while the tile is open: a
  // run that goes on beyond its chunk)
  __shared__ uint32_t s_list[M64 ? kListCap : 1];
  __shared__ uint32_t s_len[M64 ? kListCap : 1];
  // per chunk: how many of its first positions continue the previous chunk's
  // run (kWave + 1: all of them, and the run goes on)
  __shared__ uint8_t s_cont[M64 ? kFinChunks : 1];
  __shared__ uint32_t s_nlist, s_base;
  constexpr uint32_t kOpen = 0x80000000u;
  const int t = threadIdx.x, w = t >> 6, lane = lane_id();
  const KeyParams kp = key_params<MODE>(c);
  // lengths of the tile's open runs from the chunks after them (0: the run
  // goes on into the next tile, k_run_len measures it)
  auto close_runs = [&](uint32_t from, uint32_t tb) {
    for (uint32_t i = from + t; i < s_nlist; i += 256) {
      if (!(s_len[i] & kOpen)) continue;
      const uint32_t off = s_list[i] - tb;  // the start's chunk and lane
      uint32_t L = kWave - (off % kWave);
      int ci = (int)(off / kWave) + 1;
      for (; ci < kFinChunks; ci++) {
        const uint32_t v = s_cont[ci];
        if (v <= (uint32_t)kWave) {
          L += v;
          break;
        }
        L += kWave;
      }
      s_len[i] = ci < kFinChunks ? L : 0u;
    }
  };
  // listed runs go to the global list in one atomic per block and flush
  auto flush = [&]() {
    const uint32_t cnt = s_nlist;  // (the same for every thread: read after a barrier)
    if (cnt == 0) return;
    if (t == 0) s_base = atomicAdd(&c->nruns, cnt);
    __syncthreads();
    const uint32_t base = s_base;
    for (uint32_t i = t; i < cnt; i += 256) {
      if (base + i < runs_cap) {
        runs[base + i] = s_list[i];
        runs[runs_cap + base + i] = s_len[i];
      } else {
        atomicOr(&c->fallback, 1u);
      }
    }
    __syncthreads();
    if (t == 0) s_nlist = 0;
  };
  if (M64 && t == 0) s_nlist = 0;
  const uint32_t un = (uint32_t)n;  // (n < 2^30: positions fit 32 bits)
  // one tile; FULL: every position of the tile exists (no bounds checks)
  auto tile_body = [&](auto full_tag, uint32_t tb) {
    constexpr bool FULL = decltype(full_tag)::value;
    float4 b8[kPer];
    int v8[kPer];  // ids (64-bit keys: they move with the boxes)
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const uint32_t gp = tb + t + 256u * k;
      const bool ok = FULL || gp < un;
      b8[k] = ok ? make_float4(xmin[gp], ymin[gp], xmax[gp], ymax[gp])
                 : make_float4(0.f, 0.f, 0.f, 0.f);
      if (M64) v8[k] = ok ? id[gp] : 0;
    }
    if constexpr (M64) {
      const uint32_t te = FULL ? tb + kFinTile : un;
      const bool before = tb > 0, after = te < un;
      if (t < 2 && (t == 0 ? before : after)) {
        const uint32_t p = t == 0 ? tb - 1 : te;
        uint32_t lo;
        s_nb[t] = morton_cell(xmin[p], ymin[p], xmax[p], ymax[p], kp, &lo);
      }
      // high halves compared as cells, low halves as coordinates
      uint32_t h8[kPer], l8[kPer];
#pragma unroll
      for (int k = 0; k < kPer; k++) {
        h8[k] = morton_cell(b8[k].x, b8[k].y, b8[k].z, b8[k].w, kp, &l8[k]);
        if (lane == 0) s_eh[0][k * kFinWaves + w] = h8[k];
        if (lane == kWave - 1) s_eh[1][k * kFinWaves + w] = h8[k];
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kPer; k++) {
        const uint32_t gp = tb + t + 256u * k;
        const int ci = k * kFinWaves + w;  // chunk index in the tile
        const uint32_t h = h8[k];
        // same cell as the previous / next position (lanes 0 and 63: the
        // neighbouring chunk's edge from LDS)
        bool eqp = dpp_perm<kDppWaveShr1>(h) == h, eqn = dpp_perm<kDppWaveShl1>(h) == h;
        if (lane == 0) eqp = (ci > 0 || before) && (ci > 0 ? s_eh[1][ci - 1] : s_nb[0]) == h;
        if (lane == kWave - 1)
          eqn = (ci + 1 < kFinChunks ? s_eh[0][ci + 1] : s_nb[1]) == h;
        eqp = eqp && (FULL || gp < un);
        eqn = eqn && gp + 1 < un;
        int st, en;  // this lane's run in the chunk (-1 / kWave: beyond it)
        run_bounds(eqp, eqn, &st, &en);
        const bool inrun = eqp || eqn;
        const bool edge = (ci == 0 && st == 0 && before) ||
                          (ci == kFinChunks - 1 && en == kWave - 1 && after);
        const bool local = inrun && st >= 0 && en < kWave && !edge;
        if (lane == 0) s_cont[ci] = (uint8_t)(eqp ? (en < kWave ? en + 1 : kWave + 1) : 0);
        if (inrun && st == lane && !local) {
          const uint32_t at = atomicAdd(&s_nlist, 1u);
          s_list[at] = gp;
          s_len[at] = en < kWave ? (uint32_t)(en - st + 1) : kOpen;
        }
        if (!ballot(local)) continue;
        // (runs that the last sort pass could not settle: split between its
        // chunks or tiles)
        const int dest = local_run_dest(local, st, en, l8[k]);
        const uint64_t moved = ballot(dest != lane);
        if (!moved) continue;
        wave_permute(dest, b8[k], v8[k]);
        // whole 32-byte sectors (8 positions) around the moved boxes are
        // written: a partly written sector costs the memory a read first
        if (((moved >> (lane & ~7)) & 0xFFull) && (FULL || gp < un)) {
          xmin[gp] = b8[k].x;
          ymin[gp] = b8[k].y;
          xmax[gp] = b8[k].z;
          ymax[gp] = b8[k].w;
          id[gp] = v8[k];
        }
      }
    }
    if constexpr (G > 0) {
#pragma unroll
      for (int k = 0; k < kPer; k++) {
        const uint32_t gp = tb + t + 256u * k;
        leaf_parent<G>(gp, un, b8[k].x, b8[k].y, b8[k].z, b8[k].w, FULL || gp < un, tr,
                       (uint32_t)l0start, (uint32_t)r_sz, false);
      }
    }
  };
  const uint32_t ntiles = (un + kFinTile - 1) / kFinTile;
  // the previous tile's first entry and first position (the same in every thread)
  uint32_t opened = 0, prev_tb = 0;
  for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    __syncthreads();  // the previous tile is done with the shared words
    if (M64) {
      // (before this tile's chunks overwrite s_cont)
      close_runs(opened, prev_tb);
      if (s_nlist > (uint32_t)(kListCap - 2 * kFinChunks)) {
        __syncthreads();
        flush();
        opened = 0;
      } else {
        opened = s_nlist;
      }
    }
    const uint32_t tb = tile * kFinTile;
    if (tb + kFinTile <= un) tile_body(std::true_type{}, tb);
    else tile_body(std::false_type{}, tb);
    prev_tb = tb;
  }
  if constexpr (M64) {
    __syncthreads();
    close_runs(opened, prev_tb);
    __syncthreads();
    flush();
  }
}

// Leaf-parent MBRs over the listed runs, one wave each (nodes shared by two
// runs are computed twice, identically).
template <int G>
__global__ __launch_bounds__(256) void k_fixup_groups(
    int64_t n, const BuildCtrl* c, const uint32_t* __restrict__ runs, uint32_t runs_cap,
    const float* __restrict__ xmin, const float* __restrict__ ymin,
    const float* __restrict__ xmax, const float* __restrict__ ymax, TreeView tr,
    int64_t l0start, int64_t r_sz) {
  const uint32_t nr = c->nruns < runs_cap ? c->nruns : runs_cap;
  if (blockIdx.x * 4u >= nr || c->fallback) return;
  const uint32_t nw = gridDim.x * 4u;
  for (uint32_t q = (blockIdx.x * 256u + threadIdx.x) >> 6; q < nr; q += nw) {
    const int64_t i = runs[q], L = runs[runs_cap + q];
    const int64_t p0 = (i / G) * G;
    int64_t p1 = ((i + L - 1) / G + 1) * G;
    if (p1 > n) p1 = n;
    // four chunks of 64 positions per round (the loads in flight together)
    for (int64_t base = p0; base < p1; base += 4 * kWave) {
      float a[4], b[4], cc[4], d[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int64_t p = base + kWave * u + lane_id();
        a[u] = b[u] = cc[u] = d[u] = 0.f;
        if (p < p1) {
          a[u] = xmin[p];
          b[u] = ymin[p];
          cc[u] = xmax[p];
          d[u] = ymax[p];
        }
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int64_t p = base + kWave * u + lane_id();
        if (base + kWave * u < p1)
          leaf_parent<G>(p, n, a[u], b[u], cc[u], d[u], p < p1, tr, l0start, r_sz, true);
      }
    }
  }
}

// Fallback sort's last stage: gather the boxes into sorted order (in place on
// the caller's arrays, reading the 32-B staging records) and, for power-of-two
// F <= 64, the leaf-parent level.  Runs only if c->fallback is set.
template <int G>
__global__ __launch_bounds__(256) void k_gather_l0(int64_t n, const BuildCtrl* c,
                                                   const uint32_t* __restrict__ v0,
                                                   const uint32_t* __restrict__ v1,
                                                   const uint4* __restrict__ aos, float* xmin,
                                                   float* ymin, float* xmax, float* ymax,
                                                   int* id, TreeView tr, int64_t l0start,
                                                   int64_t r_sz) {
  if (!c->fallback) return;
  const uint32_t* vf = (c->npass & 1) ? v1 : v0;
  for (int64_t base = blockIdx.x * 256ll; base < n; base += (int64_t)gridDim.x * 256) {
    const int64_t i = base + threadIdx.x;
    float a = 0.f, b = 0.f, cc = 0.f, d = 0.f;
    const bool has = i < n;
    if (has) {
      const uint32_t s = vf[i];
      const uint4 r0 = aos[2 * (size_t)s];
      const uint4 r1 = aos[2 * (size_t)s + 1];
      a = __uint_as_float(r0.x);
      b = __uint_as_float(r0.y);
      cc = __uint_as_float(r0.z);
      d = __uint_as_float(r0.w);
      xmin[i] = a;
      ymin[i] = b;
      xmax[i] = cc;
      ymax[i] = d;
      id[i] = (int)r1.x;
    }
    if constexpr (G > 0) leaf_parent<G>(i, n, a, b, cc, d, has, tr, l0start, r_sz, false);
  }
}

// Upper level (or a level 0 that is not fused): child i of [0, cn) reads the
// child arrays at cbase, parent i/G is written at pbase.  pos = absolute child
// index (src/rtree_bulkload_thrust.cuh:151-168).
template <int G>
__global__ __launch_bounds__(256) void k_level_pow2(int64_t cn, const float* cx0, const float* cy0,
                                                    const float* cx1, const float* cy1,
                                                    int64_t pos_base, TreeView tr,
                                                    int64_t pbase) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  float a = 0.f, b = 0.f, cc = 0.f, d = 0.f;
  const bool has = i < cn;
  if (has) {
    a = cx0[i];
    b = cy0[i];
    cc = cx1[i];
    d = cy1[i];
  }
  group_mbr<G>(a, b, cc, d, has);
  if (has && (i % G) == 0) {
    int64_t node = pbase + i / G;
    tr.xmin[node] = a;
    tr.ymin[node] = b;
    tr.xmax[node] = cc;
    tr.ymax[node] = d;
    tr.pos[node] = (int)(pos_base + i);
    int64_t len = cn - i;
    tr.len[node] = (int)(len < G ? len : G);
  }
}

// Any fanout: one thread per parent, left-to-right loop over its children.
__global__ __launch_bounds__(256) void k_level_generic(int64_t cn, int F, const float* cx0,
                                                       const float* cy0, const float* cx1,
                                                       const float* cy1, int64_t pos_base,
                                                       TreeView tr, int64_t pbase, int64_t pn) {
  const int64_t g = blockIdx.x * 256ll + threadIdx.x;
  if (g >= pn) return;
  int64_t b0 = g * F, e = b0 + F;
  if (e > cn) e = cn;
  float a = cx0[b0], b = cy0[b0], cc = cx1[b0], d = cy1[b0];
  for (int64_t i = b0 + 1; i < e; i++) {
    a = mbr_min(a, cx0[i]);
    b = mbr_min(b, cy0[i]);
    cc = mbr_max(cc, cx1[i]);
    d = mbr_max(d, cy1[i]);
  }
  int64_t node = pbase + g;
  tr.xmin[node] = a;
  tr.ymin[node] = b;
  tr.xmax[node] = cc;
  tr.ymax[node] = d;
  tr.pos[node] = (int)(pos_base + b0);
  tr.len[node] = (int)(e - b0);
}

// ---------------------------------------------------------------------------
// Host side.
static int pow2_group(int F) {
  switch (F) {
    case 2: case 4: case 8: case 16: case 32: case 64: return F;
    default: return 0;
  }
}

struct SortBufs {
  uint64_t* kv[2];
  uint32_t* keys;    // sort keys of the input order (k_hist -> first pass)
  float2* lo[2];
  float2* hi[2];
  uint64_t* k64[2];  // fallback: 64-bit keys (the kv buffers)
  uint32_t* v[2];    // fallback: positions
  uint4* aos;        // fallback: 32-B staging records (the lo/hi buffers)
  uint32_t* status;
  size_t step_words;
  uint32_t* runs;
  uint32_t runs_cap;
};

// Ties of the 64-bit keys and the leaf-parent level, in place on the caller's
// arrays (nothing to do for 32-bit keys and a fanout without the fused level).
template <int MODE, int G>
static void launch_tail(Prof& pf, hipStream_t s, int64_t n, BuildCtrl* c, const SortBufs& b,
                        curtree_bbox_f32* box, TreeView tr, int64_t l0start, int64_t r_sz) {
  if (MODE != kKeyMorton64 && G == 0) return;
  pf.mark(s, "k_final");
  hipLaunchKernelGGL((k_final<MODE, G>),
                     dim3((unsigned)std::min<int64_t>((n + kFinTile - 1) / kFinTile,
                                                      CT_FINAL_GRID)),
                     dim3(256), 0, s, n, c, box->xmin, box->ymin, box->xmax, box->ymax, box->id,
                     tr, l0start, r_sz, b.runs, b.runs_cap);
  if (MODE != kKeyMorton64) return;
  pf.mark(s, "k_run_len");
  hipLaunchKernelGGL(k_run_len, dim3(kFixSlots), dim3(256), 0, s, n, c, (const float*)box->xmin,
                     (const float*)box->ymin, (const float*)box->xmax, (const float*)box->ymax,
                     b.runs, b.runs_cap);
  pf.mark(s, "k_fixup_short");
  hipLaunchKernelGGL(k_fixup_short, dim3(kFixSlots), dim3(256), 0, s, n, c, box->xmin,
                     box->ymin, box->xmax, box->ymax, box->id, (const uint32_t*)b.runs,
                     b.runs_cap);
  pf.mark(s, "k_fixup_long");
  hipLaunchKernelGGL(k_fixup_long, dim3(kFixLongSlots), dim3(256), 0, s, n, c, box->xmin,
                     box->ymin, box->xmax, box->ymax, box->id, (const uint32_t*)b.runs,
                     b.runs_cap, (float4*)b.aos, b.v[0]);
  // a run longer than n / 64 (at least 65536) takes the full 64-bit sort
  pf.mark(s, "k_fixup_huge");
  hipLaunchKernelGGL(k_fixup_huge, dim3(kFixHugeSlots), dim3(kHugeThreads), 0, s, n, c, box->xmin,
                     box->ymin, box->xmax, box->ymax, box->id, b.runs, b.runs_cap, b.k64[0],
                     b.k64[1], (float4*)b.aos, b.v[0],
                     std::max<int64_t>(int64_t(1) << 16, n / 64));
  pf.mark(s, "k_fixup_groups");
  if constexpr (G > 0)
    hipLaunchKernelGGL(k_fixup_groups<G>, dim3(kFixSlots), dim3(256), 0, s, n,
                       (const BuildCtrl*)c, (const uint32_t*)b.runs, b.runs_cap, box->xmin,
                       box->ymin, box->xmax, box->ymax, tr, l0start, r_sz);
  // the full LSD sort of 64-bit keys: every launch exits at once unless a
  // run of equal high halves was too long to settle in place
  pf.mark(s, "k_fallback");
  hipLaunchKernelGGL(k_fallback_zero, dim3(256), dim3(256), 0, s, c, b.status + 4 * b.step_words,
                     (int64_t)(8 * b.step_words));
  const unsigned sgrid = (unsigned)std::min<int64_t>((n + 1023) / 1024, 2048);
  hipLaunchKernelGGL(k_keygen64, dim3(sgrid), dim3(256), 0, s, n, (const float*)box->xmin,
                     (const float*)box->ymin, (const float*)box->xmax, (const float*)box->ymax,
                     (const int*)box->id, c, b.k64[0], b.k64[1], b.aos, 1);
  hipLaunchKernelGGL(k_sort_plan, dim3(1), dim3(256), 0, s, n, 8, c, 4, 8, 0x76543210ull, 0, 1,
                     1, 0);
  const int64_t ntiles = (n + kSortTile - 1) / kSortTile;
  for (int st = 4; st < kMaxSteps; st++)
    hipLaunchKernelGGL(k_onesweep<uint64_t>, dim3((unsigned)std::min(ntiles, kOnesweepGrid)),
                       dim3(kSortThreads), 0, s, n, st, b.k64[0], b.k64[1], b.v[0], b.v[1], c,
                       b.status + (size_t)st * b.step_words, 0);
  const unsigned ggrid = (unsigned)std::min<int64_t>((n + 255) / 256, 2048);
  hipLaunchKernelGGL(k_gather_l0<G>, dim3(ggrid), dim3(256), 0, s, n, (const BuildCtrl*)c,
                     (const uint32_t*)b.v[0], (const uint32_t*)b.v[1], (const uint4*)b.aos,
                     box->xmin, box->ymin, box->xmax, box->ymax, box->id, tr, l0start, r_sz);
}

template <int MODE>
static void launch_tail_g(Prof& pf, hipStream_t s, int G, int64_t n, BuildCtrl* c,
                          const SortBufs& b, curtree_bbox_f32* box, TreeView tr, int64_t l0start,
                          int64_t r_sz) {
  switch (G) {
    case 2: launch_tail<MODE, 2>(pf, s, n, c, b, box, tr, l0start, r_sz); break;
    case 4: launch_tail<MODE, 4>(pf, s, n, c, b, box, tr, l0start, r_sz); break;
    case 8: launch_tail<MODE, 8>(pf, s, n, c, b, box, tr, l0start, r_sz); break;
    case 16: launch_tail<MODE, 16>(pf, s, n, c, b, box, tr, l0start, r_sz); break;
    case 32: launch_tail<MODE, 32>(pf, s, n, c, b, box, tr, l0start, r_sz); break;
    case 64: launch_tail<MODE, 64>(pf, s, n, c, b, box, tr, l0start, r_sz); break;
    default: launch_tail<MODE, 0>(pf, s, n, c, b, box, tr, l0start, r_sz);
  }
}

static void launch_level(hipStream_t s, int F, int64_t cn, const float* cx0, const float* cy0,
                         const float* cx1, const float* cy1, int64_t pos_base, TreeView tr,
                         int64_t pbase, int64_t pn) {
  int G = pow2_group(F);
  unsigned grid = (unsigned)((cn + 255) / 256);
  switch (G) {
#define CT_LVL(g)                                                                              \
  case g:                                                                                      \
    hipLaunchKernelGGL(k_level_pow2<g>, dim3(grid), dim3(256), 0, s, cn, cx0, cy0, cx1, cy1,   \
                       pos_base, tr, pbase);                                                   \
    return;
    CT_LVL(2) CT_LVL(4) CT_LVL(8) CT_LVL(16) CT_LVL(32) CT_LVL(64)
#undef CT_LVL
    default: {
      unsigned pg = (unsigned)((pn + 255) / 256);
      hipLaunchKernelGGL(k_level_generic, dim3(pg), dim3(256), 0, s, cn, F, cx0, cy0, cx1, cy1,
                         pos_base, tr, pbase, pn);
    }
  }
}

int build_f32(Context& ctx, curtree_bbox_f32* box, curtree_rtree_f32* rt, int key_mode,
              hipStream_t s, bool alloc) {
  if (!box || !rt) return set_error(CURTREE_EINVAL, "build: null argument");
  const int F = rt->fanout;
  const int64_t n = box->sz;
  if (F < 2) return set_error(CURTREE_EINVAL, "build: fanout must be >= 2 (got %d)", F);
  if (n < 0) return set_error(CURTREE_EINVAL, "build: negative box count");
  if (n >= (int64_t)1 << 30)
    return set_error(CURTREE_EINVAL, "build: at most 2^30-1 boxes per tree (got %lld)",
                     (long long)n);
  if (key_mode != kKeyRef32 && key_mode != kKeyMorton64)
    return set_error(CURTREE_EINVAL, "build: unknown key mode %d", key_mode);
  if (n > 0 && (!box->xmin || !box->ymin || !box->xmax || !box->ymax || !box->id))
    return set_error(CURTREE_EINVAL, "build: null box array");

  LevelPlan lp;
  level_plan(n, F, &lp);
  if (lp.r_sz > INT32_MAX - n)
    return set_error(CURTREE_EINVAL, "build: node count overflows int32 positions");
  int rc;
  if (alloc) {
    rt->sz = 0;
    rt->xmin = rt->ymin = rt->xmax = rt->ymax = nullptr;
    rt->pos = rt->len = nullptr;
    if (n == 0) return CURTREE_OK;
    // node arrays: owned by the caller afterwards (freed with curtree_rt_d_free)
    if ((rc = rt_alloc(rt, lp.r_sz))) return rc;
  } else {
    if (rt->sz != lp.r_sz)
      return set_error(CURTREE_EINVAL, "build: tree holds %d nodes, %lld boxes need %lld", rt->sz,
                       (long long)n, (long long)lp.r_sz);
    if (n == 0) return CURTREE_OK;
    if (!rt->xmin || !rt->ymin || !rt->xmax || !rt->ymax || !rt->pos || !rt->len)
      return set_error(CURTREE_EINVAL, "build: null node array");
  }
  auto fail = [&](int code) {
    const void* ps[6] = {rt->xmin, rt->ymin, rt->xmax, rt->ymax, rt->pos, rt->len};
    tree_state_forget(ps, 6);
    if (alloc) rt_free(rt);
    return code;
  };

  const int64_t ntiles = (n + kSweepTile - 1) / kSweepTile;      // payload passes
  const int64_t ntiles_lsd = (n + kSortTile - 1) / kSortTile;    // 64-bit LSD passes
  const bool m64 = key_mode == kKeyMorton64;
  const int nsteps = m64 ? kMaxSteps : 4;  // Morton: 4 high-half steps + 8 fallback steps
  const size_t step_words = (size_t)std::max(ntiles, ntiles_lsd) * 256;
  const size_t status_words = (size_t)nsteps * step_words;
  Workspace& ws = ctx.ws;
  const uint32_t runs_cap = m64 ? (uint32_t)(2 * ((n + kWave - 1) / kWave) + 2) : 1;
  // a buffer that grows is freed first: the previous build (possibly on
  // another stream) must be done with it
  if (ctx.build_done &&
      (!ws.fits(WsBuildKeys0, n * 8) || !ws.fits(WsBuildKeys1, n * 8) ||
       !ws.fits(WsBuildAos, n * 32) || !ws.fits(WsBuildVals0, n * 4) ||
       !ws.fits(WsBuildVals1, n * 4) || !ws.fits(WsBuildStatus, status_words * 4) ||
       !ws.fits(WsBuildRuns, (size_t)runs_cap * 16)) &&
      hipEventSynchronize(ctx.build_done) != hipSuccess)
    return fail(set_error(CURTREE_EHIP, "build: cannot wait for the previous build"));
  // payload sort: the ids and the box halves, double-buffered (40 B/box); the
  // fallback uses the key buffers for 64-bit keys and the box halves for 32-B
  // staging records
  if ((rc = ws.reserve(WsBuildKeys0, n * 8)) || (rc = ws.reserve(WsBuildKeys1, n * 8)) ||
      (rc = ws.reserve(WsBuildAos, n * 32)) ||
      (rc = ws.reserve(WsBuildVals0, n * 4)) || (rc = ws.reserve(WsBuildVals1, n * 4)) ||
      (rc = ws.reserve(WsBuildStatus, status_words * 4)) ||
      (rc = ws.reserve(WsBuildCtrl, sizeof(BuildCtrl))))
    return fail(rc);
  // listed runs of equal high halves (64-bit keys): start, length and the
  // lists of the long ones; at most two per chunk of 64 positions
  if ((rc = ws.reserve(WsBuildRuns, (size_t)runs_cap * 16))) return fail(rc);
  BuildCtrl* c = (BuildCtrl*)ws.ptr(WsBuildCtrl);
  SortBufs sb;
  sb.kv[0] = (uint64_t*)ws.ptr(WsBuildKeys0);
  sb.kv[1] = (uint64_t*)ws.ptr(WsBuildKeys1);
  float2* f2 = (float2*)ws.ptr(WsBuildAos);
  sb.lo[0] = f2;
  sb.hi[0] = f2 + n;
  sb.lo[1] = f2 + 2 * n;
  sb.hi[1] = f2 + 3 * n;
  sb.k64[0] = sb.kv[0];
  sb.k64[1] = sb.kv[1];
  sb.v[0] = (uint32_t*)ws.ptr(WsBuildVals0);
  sb.v[1] = (uint32_t*)ws.ptr(WsBuildVals1);
  sb.keys = sb.v[1];
  sb.aos = (uint4*)ws.ptr(WsBuildAos);
  sb.status = (uint32_t*)ws.ptr(WsBuildStatus);
  sb.step_words = step_words;
  sb.runs = (uint32_t*)ws.ptr(WsBuildRuns);
  sb.runs_cap = runs_cap;

  // every build shares the work buffers: a build on another stream waits for
  // the previous build's kernels (asynchronous builds may overlap on the host)
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cap) != hipSuccess) {
    (void)hipGetLastError();
    cap = hipStreamCaptureStatusNone;
  }
  const bool order = cap == hipStreamCaptureStatusNone;
  if (order && !ctx.build_done &&
      hipEventCreateWithFlags(&ctx.build_done, hipEventDisableTiming) != hipSuccess) {
    ctx.build_done = nullptr;
    return fail(set_error(CURTREE_EHIP, "build: cannot create an event"));
  }
  if (order && hipStreamWaitEvent(s, ctx.build_done, 0) != hipSuccess)
    return fail(set_error(CURTREE_EHIP, "build: cannot order after the previous build"));
  // the tree's well-formed word, kept for its queries (no allocation under
  // capture)
  uint32_t* wf = tree_state_register(rt, box, order);
  Prof& pf = ctx.prof;
  pf.mark(s, "k_build_init");
  hipLaunchKernelGGL(k_build_init, dim3(1), dim3(256), 0, s, c, wf);
  const unsigned sgrid = (unsigned)std::min<int64_t>((n + 1023) / 1024, 2048);
  if (m64) {
    pf.mark(s, "k_extent");
    hipLaunchKernelGGL(k_extent, dim3(sgrid), dim3(256), 0, s, n, box->xmin, box->ymin,
                       box->xmax, box->ymax, c);
  }
  // histograms of the sort keys' digits; clears the status words of steps 0..3
  pf.mark(s, "k_hist");
  if (m64)
    hipLaunchKernelGGL(k_hist<kKeyMorton64>, dim3(sgrid), dim3(256), 0, s, n, box->xmin,
                       box->ymin, box->xmax, box->ymax, c, sb.status, (int64_t)(4 * step_words), wf);
  else
    hipLaunchKernelGGL(k_hist<kKeyRef32>, dim3(sgrid), dim3(256), 0, s, n, box->xmin, box->ymin,
                       box->xmax, box->ymax, c, sb.status, (int64_t)(4 * step_words), wf);
  // (the last step that runs writes the caller's arrays)
  pf.mark(s, "k_sort_plan");
  hipLaunchKernelGGL(k_sort_plan, dim3(1), dim3(256), 0, s, n, 4, c, 0, 4, 0x3210ull, 0, 0, 2,
                     1);
  for (int st = 0; st < 4; st++) {
    pf.mark(s, "k_sweep_kv");
    auto kern = m64 ? (st > 0 ? k_sweep_kv<false, kKeyMorton64> : k_sweep_kv<true, kKeyMorton64>)
                    : (st > 0 ? k_sweep_kv<false, kKeyRef32> : k_sweep_kv<true, kKeyRef32>);
    hipLaunchKernelGGL(kern, dim3((unsigned)ntiles), dim3(kSortThreads), 0, s, n, st, box->xmin,
                       box->ymin, box->xmax, box->ymax, box->id, sb.v[0], sb.v[1], sb.lo[0],
                       sb.lo[1], sb.hi[0], sb.hi[1], c, sb.status + (size_t)st * step_words);
  }

  TreeView tr{rt->xmin, rt->ymin, rt->xmax, rt->ymax, rt->pos, rt->len};
  const int64_t l0start = lp.r_sz - lp.n[0];
  const int G = pow2_group(F);
  if (m64)
    launch_tail_g<kKeyMorton64>(pf, s, G, n, c, sb, box, tr, l0start, lp.r_sz);
  else
    launch_tail_g<kKeyRef32>(pf, s, G, n, c, sb, box, tr, l0start, lp.r_sz);
  if (G == 0) {
    pf.mark(s, "k_level");
    launch_level(s, F, n, box->xmin, box->ymin, box->xmax, box->ymax, lp.r_sz, tr, l0start,
                 lp.n[0]);
  }
  // upper levels: level l (l >= 1) reads level l-1
  int64_t child = l0start;
  for (int l = 1; l < lp.nl; l++) {
    int64_t pstart = child - lp.n[l];
    pf.mark(s, "k_level");
    launch_level(s, F, lp.n[l - 1], rt->xmin + child, rt->ymin + child, rt->xmax + child,
                 rt->ymax + child, child, tr, pstart, lp.n[l]);
    child = pstart;
  }
  pf.mark(s, nullptr);
  if (order) (void)hipEventRecord(ctx.build_done, s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return fail(set_error(CURTREE_EHIP, "build: launch failed: %s", hipGetErrorString(e)));
  return CURTREE_OK;
}

// ---------------------------------------------------------------------------
// Canonical pair order by segments.  Query results arrive in (query, sorted
// box position) order: query ids never decrease, so the pairs of one query
// are one contiguous segment and the canonical (query id, box id) order is
// each segment's box ids sorted in place (the query ids stay where they are).
//   k_seg_count   per chunk of kSegChunk pairs: its segment heads (positions
//                 whose query id differs from the one before), and a flag if
//                 a query id ever decreases (wide loads);
//   k_seg_scan    exclusive scan of the chunks' head counts (one block);
//   k_seg_heads   the head positions, written in order (the head list);
//   k_seg_classify  the segments of 2 or more pairs listed by length class
//                 (a segment ends at the next head);
//   k_seg_wave    segments of 2..1024 pairs, one wave each (up to 64 a bitonic
//                 network over the lanes, longer ones four LSD passes in
//                 registers and the wave's LDS);
//   k_seg_block   1025..4096 pairs, one block of four waves each, and
//                 4097..16384, one block of sixteen waves each (lds_lsd);
//   k_seg_huge    longer segments, one block each: the four digit histograms
//                 in one read, then four stable LSD passes between the box ids
//                 and a work array, in chunks of 4096 ranked in LDS.
// Box ids are sorted as unsigned keys (id ^ 0x80000000), i.e. in signed order.
constexpr int kSegWaveMax = 1024;
constexpr int kSegBlockMax = 4096;
constexpr int kSegBigMax = 16384;
constexpr int kSegPer = 16;               // query ids per thread of a chunk
constexpr int kSegChunk = 256 * kSegPer;  // pairs per chunk (one block step)
struct SegCtrl {
  // listed segments: wave, block, big-block and long class, and the buckets of
  // long segments too long for a block
  unsigned long long nseg[5];
  unsigned long long nheads;   // length of the head list
  uint32_t unordered;          // a query id is below the one before it
  uint32_t too_long;           // a segment holds 2^32 pairs or more (not sorted)
};
// A listed segment's ids are in the box id array, or (the buckets of a long
// segment, kSegFromTmp) at the same positions of the work array.
struct SegBufs {
  uint64_t* small;  // wave class: start << 12 | from_tmp << 11 | length (<= 1024)
  uint64_t* big;    // the other classes: (start, length | from_tmp << 63), a region each
  uint64_t off[5];  // first entry of each class's region in `big` (classes 1..4)
  uint64_t cap[5];  // entries each class's list holds
};
constexpr uint64_t kSegFromTmp = 1ull << 63;

__global__ void k_seg_init(SegCtrl* c) {
  if (threadIdx.x < 5) c->nseg[threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    c->nheads = 0;
    c->unordered = 0;
    c->too_long = 0;
  }
}

// Head bits of this thread's kSegPer positions p0 .. p0 + kSegPer - 1 of the
// chunk at c0 (bit k: position p0 + k starts a segment); *down: a query id
// below the one before it.  s_last: kSegChunk / kSegPer words of LDS.  Every
// thread of the block calls it.
__device__ __forceinline__ uint32_t seg_head_bits(int64_t n, const int* __restrict__ tid,
                                                  int64_t c0, uint32_t* s_last, bool* down) {
  const int t = threadIdx.x;
  const int64_t p0 = c0 + (int64_t)t * kSegPer;
  int v[kSegPer];
  // (wide loads only from a 16-byte aligned array: p0 is a multiple of 16)
  if (p0 + kSegPer <= n && ((uintptr_t)tid & 15u) == 0) {
#pragma unroll
    for (int q = 0; q < kSegPer / 4; q++) {
      const int4 a = *(const int4*)(tid + p0 + 4 * q);
      v[4 * q] = a.x;
      v[4 * q + 1] = a.y;
      v[4 * q + 2] = a.z;
      v[4 * q + 3] = a.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < kSegPer; k++) v[k] = p0 + k < n ? tid[p0 + k] : 0;
  }
  __syncthreads();  // the previous chunk is done with s_last
  s_last[t] = (uint32_t)v[kSegPer - 1];
  __syncthreads();
  // the id before this thread's first: the previous thread's last (thread 0:
  // a load; position 0 always starts a segment)
  int prev = t ? (int)s_last[t - 1] : (c0 > 0 ? tid[c0 - 1] : ~v[0]);
  uint32_t hb = 0;
  bool dn = false;
#pragma unroll
  for (int k = 0; k < kSegPer; k++) {
    const bool valid = p0 + k < n;
    if (valid && v[k] != prev) hb |= 1u << k;
    dn |= valid && p0 + k > 0 && v[k] < prev;
    prev = v[k];
  }
  *down = dn;
  return hb;
}

__global__ __launch_bounds__(256) void k_seg_count(int64_t n, const int* __restrict__ tid,
                                                   uint32_t* __restrict__ cnt, SegCtrl* c) {
  __shared__ uint32_t s_last[256];
  __shared__ uint32_t s_tmp[4];
  const int64_t nchunks = (n + kSegChunk - 1) / kSegChunk;
  for (int64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    bool dn;
    const uint32_t hb = seg_head_bits(n, tid, ch * kSegChunk, s_last, &dn);
    if (ballot(dn) && lane_id() == 0) c->unordered = 1u;
    uint32_t tot;
    (void)block_excl_scan<4>((uint32_t)__popc(hb), s_tmp, &tot);
    if (threadIdx.x == 0) cnt[ch] = tot;
  }
}

// Exclusive scan of the chunk head counts into 64-bit offsets, by one block
// of 1024 threads in rounds of 8192 counts (thread t loads counts 8t .. 8t+7
// of the round as two 16-byte words), carrying the running total; the total
// is the head list's length.
__global__ __launch_bounds__(1024) void k_seg_scan(int64_t nchunks,
                                                   const uint32_t* __restrict__ cnt,
                                                   uint64_t* __restrict__ off, SegCtrl* c) {
  __shared__ uint64_t s_w[16];
  const int t = threadIdx.x, lane = lane_id(), w = t >> 6;
  uint64_t carry = 0;
  for (int64_t r0 = 0; r0 < nchunks; r0 += 8192) {
    const int64_t i0 = r0 + 8 * (int64_t)t;
    uint32_t v[8];
    if (i0 + 8 <= nchunks) {
      const uint4 a = *(const uint4*)(cnt + i0);
      const uint4 z = *(const uint4*)(cnt + i0 + 4);
      v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
      v[4] = z.x; v[5] = z.y; v[6] = z.z; v[7] = z.w;
    } else {
#pragma unroll
      for (int k = 0; k < 8; k++) v[k] = i0 + k < nchunks ? cnt[i0 + k] : 0u;
    }
    uint64_t sum = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) sum += v[k];
    uint64_t x = sum;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const uint64_t y = __shfl_up(x, o);
      if (lane >= o) x += y;
    }
    if (lane == kWave - 1) s_w[w] = x;
    __syncthreads();
    uint64_t before = carry + x - sum, total = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      if (i < w) before += s_w[i];
      total += s_w[i];
    }
    __syncthreads();  // s_w is rewritten by the next round
#pragma unroll
    for (int k = 0; k < 8; k++) {
      if (i0 + k < nchunks) off[i0 + k] = before;
      before += v[k];
    }
    carry += total;
  }
  if (t == 0) c->nheads = carry;
}

__global__ __launch_bounds__(256) void k_seg_heads(int64_t n, const int* __restrict__ tid,
                                                   const uint64_t* __restrict__ off,
                                                   uint64_t* __restrict__ heads) {
  __shared__ uint32_t s_last[256];
  __shared__ uint32_t s_tmp[4];
  const int64_t nchunks = (n + kSegChunk - 1) / kSegChunk;
  for (int64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    bool dn;
    const int64_t c0 = ch * kSegChunk;
    uint32_t hb = seg_head_bits(n, tid, c0, s_last, &dn);
    uint32_t tot;
    uint64_t at = off[ch] + block_excl_scan<4>((uint32_t)__popc(hb), s_tmp, &tot);
    const int64_t p0 = c0 + (int64_t)threadIdx.x * kSegPer;
    while (hb) {
      const int k = __ffs(hb) - 1;
      hb &= hb - 1;
      heads[at++] = (uint64_t)(p0 + k);
    }
  }
}

// Sorts the head list into the class lists: segment lengths 2..1024 (wave
// class, start << 11 | length), 1025..4096, 4097..16384 and longer ((start,
// length) pairs); one-pair segments need nothing.  One thread per head; each
// block reserves its entries of a class with one atomic.
__global__ __launch_bounds__(256) void k_seg_classify(int64_t n, const uint64_t* __restrict__ heads,
                                                      SegCtrl* c, SegBufs b) {
  __shared__ uint32_t s_n[4];
  __shared__ unsigned long long s_at[4];
  const uint64_t nh = c->nheads;
  for (uint64_t i0 = (uint64_t)blockIdx.x * 256; i0 < nh; i0 += (uint64_t)gridDim.x * 256) {
    __syncthreads();  // the previous batch is done with the shared words
    if (threadIdx.x < 4) s_n[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t i = i0 + threadIdx.x;
    int64_t st = 0, len = 0;
    int q = -1;
    uint32_t r = 0;
    if (i < nh) {
      st = (int64_t)heads[i];
      len = (i + 1 < nh ? (int64_t)heads[i + 1] : n) - st;
      q = len < 2 ? -1 : len <= kSegWaveMax ? 0 : len <= kSegBlockMax ? 1 : len <= kSegBigMax ? 2 : 3;
      if (len > (int64_t)0xFFFFFFFFll) {  // (positions in a segment are 32-bit)
        q = -1;
        c->too_long = 1u;
      }
      if (q >= 0) r = atomicAdd(&s_n[q], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 4)
      s_at[threadIdx.x] = s_n[threadIdx.x] ? atomicAdd(&c->nseg[threadIdx.x],
                                                       (unsigned long long)s_n[threadIdx.x])
                                           : 0ull;
    __syncthreads();
    if (q >= 0) {
      const uint64_t at = s_at[q] + r;
      if (at < b.cap[q]) {
        if (q == 0) {
          b.small[at] = ((uint64_t)st << 12) | (uint64_t)len;
        } else {
          b.big[2 * (b.off[q] + at)] = (uint64_t)st;
          b.big[2 * (b.off[q] + at) + 1] = (uint64_t)len;
        }
      }
    }
  }
}

__device__ __forceinline__ uint32_t id_key(int v) { return (uint32_t)v ^ 0x80000000u; }
__device__ __forceinline__ int key_id(uint32_t k) { return (int)(k ^ 0x80000000u); }

// Segments of 2..1024 box ids, one wave each.
__global__ __launch_bounds__(256) void k_seg_wave(const SegCtrl* c, SegBufs b,
                                                  int* __restrict__ fid,
                                                  const int* __restrict__ tmp) {
  __shared__ uint32_t s_w[4][kSegWaveMax];
  __shared__ __align__(16) uint32_t s_c[4][256];
  __shared__ uint64_t s_m[4][256];  // peer_mask tables
  const uint64_t nseg = c->nseg[0] < b.cap[0] ? c->nseg[0] : b.cap[0];
  const int w = threadIdx.x >> 6, lane = lane_id();
  const uint64_t lt = lanemask_lt();
#pragma unroll
  for (int k = 0; k < 4; k++) s_m[w][lane + 64 * k] = 0;  // (each wave uses its own row)
  for (uint64_t sg = (uint64_t)blockIdx.x * 4 + w; sg < nseg; sg += (uint64_t)gridDim.x * 4) {
    const uint64_t e = b.small[sg];
    const int64_t st = (int64_t)(e >> 12);
    const int L = (int)(e & 2047u);
    const int* src = ((e >> 11) & 1u) ? tmp : fid;
    if (L <= 64) {
      // bitonic network over the lanes (lanes past the segment hold the
      // largest key and stay at the end)
      uint32_t x = lane < L ? id_key(src[st + lane]) : 0xFFFFFFFFu;
#pragma unroll
      for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
          const uint32_t y = (uint32_t)__shfl_xor((int)x, j);
          const bool up = (lane & k) == 0;   // this block of k sorts ascending
          const bool low = (lane & j) == 0;  // this lane keeps the smaller one
          const uint32_t mn = x < y ? x : y, mx = x < y ? y : x;
          x = (up == low) ? mn : mx;
        }
      }
      if (lane < L) fid[st + lane] = key_id(x);
      continue;
    }
    // four LSD passes: item k of this lane is position 64 * k + lane
    const int kUse = (L + 63) >> 6;
    uint32_t x[16];
#pragma unroll
    for (int k = 0; k < 16; k++) {
      const int p = 64 * k + lane;
      x[k] = (k < kUse && p < L) ? id_key(src[st + p]) : 0xFFFFFFFFu;
    }
    // digits above the highest bit where the keys differ are shared by all
    // of them (keys between the smallest and the largest share their common
    // prefix): passes over them change nothing.  (The padding keys hold the
    // largest digit in every pass, so they stay after the segment.)
    uint32_t mn = 0xFFFFFFFFu, mx = 0u;
#pragma unroll
    for (int k = 0; k < 16; k++)
      if (k < kUse && 64 * k + lane < L) {
        mn = x[k] < mn ? x[k] : mn;
        mx = x[k] > mx ? x[k] : mx;
      }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const uint32_t a = (uint32_t)__shfl_xor((int)mn, o), z = (uint32_t)__shfl_xor((int)mx, o);
      mn = a < mn ? a : mn;
      mx = z > mx ? z : mx;
    }
    const int npass = mn != mx ? ((31 - __clz((int)(mn ^ mx))) >> 3) + 1 : 0;
    for (int pass = 0; pass < npass; pass++) {
      const int sh = 8 * pass;
#pragma unroll
      for (int k = 0; k < 4; k++) s_c[w][lane + 64 * k] = 0;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      uint32_t rank[16];
#pragma unroll
      for (int k = 0; k < 16; k++) {
        if (k < kUse) {
          const uint32_t d = (x[k] >> sh) & 255u;
          const uint64_t m = peer_mask(d, true, s_m[w]);
          const uint32_t before = s_c[w][d];
          rank[k] = before + (uint32_t)popc64(m & lt);
          if ((m & lt) == 0) s_c[w][d] = before + (uint32_t)popc64(m);
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
        }
      }
      const uint4 cnt = *(const uint4*)&s_c[w][4 * lane];
      const uint32_t tot = cnt.x + cnt.y + cnt.z + cnt.w;
      uint32_t inc = tot;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t u = __shfl_up(inc, d);
        if (lane >= d) inc += u;
      }
      const uint32_t o = inc - tot;
      *(uint4*)&s_c[w][4 * lane] =
          make_uint4(o, o + cnt.x, o + cnt.x + cnt.y, o + cnt.x + cnt.y + cnt.z);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int k = 0; k < 16; k++)
        if (k < kUse) s_w[w][s_c[w][(x[k] >> sh) & 255u] + rank[k]] = x[k];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int k = 0; k < 16; k++)
        if (k < kUse) x[k] = s_w[w][64 * k + lane];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
#pragma unroll
    for (int k = 0; k < 16; k++) {
      const int p = 64 * k + lane;
      if (k < kUse && p < L) fid[st + p] = key_id(x[k]);
    }
  }
}

// Segments of 1025..4096 box ids (NW = 4, class 1) or 4097..16384 (NW = 16,
// class 2), one block of NW waves each, sorted whole in LDS.
template <int NW, int CLS>
__global__ __launch_bounds__(NW * 64) void k_seg_block(const SegCtrl* c, SegBufs b,
                                                       int* __restrict__ fid,
                                                       const int* __restrict__ tmp) {
  __shared__ uint32_t s_net[NW * 64 * kRunPer];
  __shared__ __align__(16) uint32_t s_cnt[NW][256];
  __shared__ uint32_t s_wsum[NW];
  __shared__ uint64_t s_match[NW][256];  // peer_mask tables
  __shared__ uint32_t s_mm[2][NW];
  const uint64_t nseg = c->nseg[CLS] < b.cap[CLS] ? c->nseg[CLS] : b.cap[CLS];
  for (int i = threadIdx.x; i < NW * 256; i += NW * 64) (&s_match[0][0])[i] = 0;
  __syncthreads();
  for (uint64_t sg = blockIdx.x; sg < nseg; sg += gridDim.x) {
    const int64_t st = (int64_t)b.big[2 * (b.off[CLS] + sg)];
    const uint64_t lf = b.big[2 * (b.off[CLS] + sg) + 1];
    const int L = (int)(lf & ~kSegFromTmp);
    const int* src = (lf & kSegFromTmp) ? tmp : fid;
    const int kUse = (L + NW * 64 - 1) / (NW * 64);
    const int wbase = (threadIdx.x >> 6) * 64 * kUse + lane_id();
    // (box ids alone: equal ids are indistinguishable, so no index rides
    // along; slots past the segment hold the largest key and stay after it)
    uint32_t x[kRunPer];
#pragma unroll
    for (int k = 0; k < kRunPer; k++) {
      const int p = wbase + 64 * k;
      x[k] = (k < kUse && p < L) ? id_key(src[st + p]) : 0xFFFFFFFFu;
    }
    // buckets of long segments (their keys share the top bits): only the
    // digits up to the highest bit where the keys differ (as in k_seg_wave);
    // other segments take all four passes
    int npass = 4;
    if (lf & kSegFromTmp) {
      uint32_t mn = 0xFFFFFFFFu, mx = 0u;
#pragma unroll
      for (int k = 0; k < kRunPer; k++)
        if (k < kUse && wbase + 64 * k < L) {
          mn = x[k] < mn ? x[k] : mn;
          mx = x[k] > mx ? x[k] : mx;
        }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const uint32_t a = (uint32_t)__shfl_xor((int)mn, o), z = (uint32_t)__shfl_xor((int)mx, o);
        mn = a < mn ? a : mn;
        mx = z > mx ? z : mx;
    }
    if (lane_id() == 0) {
      s_mm[0][threadIdx.x >> 6] = mn;
      s_mm[1][threadIdx.x >> 6] = mx;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NW; i++) {
      mn = s_mm[0][i] < mn ? s_mm[0][i] : mn;
      mx = s_mm[1][i] > mx ? s_mm[1][i] : mx;
    }
    npass = mn != mx ? ((31 - __clz((int)(mn ^ mx))) >> 3) + 1 : 0;
    }
    // (a pass reuses only its own wave's row of s_cnt before its first
    // barrier, and writes s_net only after two, so passes need no barrier
    // between them)
    for (int pass = 0; pass < npass; pass++)
      lds_pass<NW, uint32_t>(x, kUse, wbase, 8 * pass, s_net, s_cnt, s_wsum, s_match);
#pragma unroll
    for (int k = 0; k < kRunPer; k++) {
      const int p = wbase + 64 * k;
      if (k < kUse && p < L) fid[st + p] = key_id(x[k]);
    }
    __syncthreads();  // s_net and s_cnt are reused by the next segment
  }
}

// Segments longer than 16384 box ids, one block of 1024 threads each: the
// ids' key range (a first read), a histogram of the top bits of (key - min)
// over nb <= kSplitBins buckets chosen so that a bucket averages kSplitTarget
// ids (a second read), and the ids scattered into their buckets in the work
// array at the segment's positions (a third read).  Buckets hold consecutive
// key ranges, so the segment is sorted once every bucket is: a bucket of one
// id is copied back here, the others are listed in the class of their length,
// marked as read from the work array, and sorted by the class kernels (a
// bucket longer than kSegBigMax goes to k_seg_huge).
constexpr int kSplitBins = 4096;
constexpr int kSplitTarget = 2048;
__global__ __launch_bounds__(1024) void k_seg_split(SegCtrl* c, SegBufs b, int* __restrict__ fid,
                                                    int* __restrict__ tmp) {
  __shared__ uint32_t s_cnt[kSplitBins];      // counts, cursors, then list slots
  __shared__ uint32_t s_off[kSplitBins + 1];  // first position of each bucket
  __shared__ uint32_t s_r[2][16];
  __shared__ uint32_t s_n[5];
  __shared__ unsigned long long s_at[5];
  const int t = threadIdx.x, lane = lane_id(), w = t >> 6;
  const uint64_t nseg = c->nseg[3] < b.cap[3] ? c->nseg[3] : b.cap[3];
  for (uint64_t sg = blockIdx.x; sg < nseg; sg += gridDim.x) {
    const uint64_t slot = b.off[3] + sg;
    const int64_t st = (int64_t)b.big[2 * slot];
    const int64_t L = (int64_t)b.big[2 * slot + 1];  // (< 2^32: checked when listed)
    // the key range
    uint32_t mn = 0xFFFFFFFFu, mx = 0u;
    for (int64_t i = t; i < L; i += 1024) {
      const uint32_t k = id_key(fid[st + i]);
      mn = k < mn ? k : mn;
      mx = k > mx ? k : mx;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const uint32_t a = (uint32_t)__shfl_xor((int)mn, o), z = (uint32_t)__shfl_xor((int)mx, o);
      mn = a < mn ? a : mn;
      mx = z > mx ? z : mx;
    }
    __syncthreads();  // the previous segment is done with the shared words
    if (lane == 0) {
      s_r[0][w] = mn;
      s_r[1][w] = mx;
    }
    if (t < 5) s_n[t] = 0;
    __syncthreads();
    mn = s_r[0][0];
    mx = s_r[1][0];
#pragma unroll
    for (int i = 1; i < 16; i++) {
      mn = s_r[0][i] < mn ? s_r[0][i] : mn;
      mx = s_r[1][i] > mx ? s_r[1][i] : mx;
    }
    // nb = 2^lb buckets; bucket of key k: (k - mn) >> sh < nb
    int lb = 1;
    while (lb < 12 && ((int64_t)1 << lb) * kSplitTarget < L) lb++;
    const int nb = 1 << lb;
    const uint32_t span = mx - mn;
    const int bits = span ? 32 - __clz((int)span) : 0;
    const int sh = bits > lb ? bits - lb : 0;
    for (int j = t; j < nb; j += 1024) s_cnt[j] = 0;
    __syncthreads();
    for (int64_t i = t; i < L; i += 1024)
      atomicAdd(&s_cnt[(id_key(fid[st + i]) - mn) >> sh], 1u);
    __syncthreads();
    // exclusive scan of the counts (thread t: buckets 4t .. 4t+3), cursors 0
    {
      uint32_t v[4], sum = 0;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int j = 4 * t + q;
        v[q] = j < nb ? s_cnt[j] : 0u;
        sum += v[q];
      }
      uint32_t tot;
      uint32_t ex = block_excl_scan<16>(sum, &s_r[0][0], &tot);
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int j = 4 * t + q;
        if (j < nb) {
          s_off[j] = ex;
          s_cnt[j] = 0;
        }
        ex += v[q];
      }
      if (t == 0) s_off[nb] = tot;
    }
    __syncthreads();
    for (int64_t i = t; i < L; i += 1024) {
      const int v = fid[st + i];
      const uint32_t d = (id_key(v) - mn) >> sh;
      tmp[st + s_off[d] + atomicAdd(&s_cnt[d], 1u)] = v;
    }
    __syncthreads();  // (the block's stores to tmp are visible to it after the barrier)
    // buckets: one id copied back; the others listed by the class of their
    // length (local slot in s_cnt, then one reservation per class)
    for (int j = t; j < nb; j += 1024) {
      const uint32_t len = s_off[j + 1] - s_off[j];
      if (len == 1) fid[st + s_off[j]] = tmp[st + s_off[j]];
      const int q = len < 2 ? -1 : len <= (uint32_t)kSegWaveMax ? 0 : len <= (uint32_t)kSegBlockMax ? 1
                  : len <= (uint32_t)kSegBigMax ? 2 : 4;
      if (q >= 0) s_cnt[j] = atomicAdd(&s_n[q], 1u);
    }
    __syncthreads();
    if (t < 5) s_at[t] = s_n[t] ? atomicAdd(&c->nseg[t], (unsigned long long)s_n[t]) : 0ull;
    __syncthreads();
    for (int j = t; j < nb; j += 1024) {
      const uint32_t len = s_off[j + 1] - s_off[j];
      const int q = len < 2 ? -1 : len <= (uint32_t)kSegWaveMax ? 0 : len <= (uint32_t)kSegBlockMax ? 1
                  : len <= (uint32_t)kSegBigMax ? 2 : 4;
      if (q < 0) continue;
      const uint64_t at = s_at[q] + s_cnt[j];
      if (at >= b.cap[q]) continue;  // (capacities cover every bucket)
      const uint64_t bst = (uint64_t)(st + s_off[j]);
      if (q == 0) {
        b.small[at] = (bst << 12) | (1ull << 11) | (uint64_t)len;
      } else {
        b.big[2 * (b.off[q] + at)] = bst;
        b.big[2 * (b.off[q] + at) + 1] = (uint64_t)len | kSegFromTmp;
      }
    }
  }
}

// Buckets of long segments that are still longer than 16384 box ids (ids
// crowded in a small part of their range), one block each: copied back from
// the work array, then stable LSD passes between fid and the work array `tmp`
// (same positions).  A digit that every id of the bucket shares is skipped;
// an odd number of passes ends with a copy back.
__global__ __launch_bounds__(256) void k_seg_huge(const SegCtrl* c, SegBufs b,
                                                  int* __restrict__ fid,
                                                  uint32_t* __restrict__ tmp) {
  __shared__ uint64_t s_net[kSegBlockMax];
  __shared__ __align__(16) uint32_t s_cnt[4][256];
  __shared__ uint32_t s_wsum[4];
  __shared__ uint32_t s_hist[4][256];
  __shared__ uint32_t s_base[256];
  __shared__ int s_skip[4];
  __shared__ uint64_t s_match[4][256];  // peer_mask tables
  const int t = threadIdx.x;
  for (int i = t; i < 4 * 256; i += 256) (&s_match[0][0])[i] = 0;
  const uint64_t nseg = c->nseg[4] < b.cap[4] ? c->nseg[4] : b.cap[4];
  for (uint64_t sg = blockIdx.x; sg < nseg; sg += gridDim.x) {
    const uint64_t slot = b.off[4] + sg;
    const int64_t st = (int64_t)b.big[2 * slot];
    const uint64_t lf = b.big[2 * slot + 1];
    const int64_t L = (int64_t)(lf & ~kSegFromTmp);
    if (lf & kSegFromTmp)
      for (int64_t i = t; i < L; i += 256) fid[st + i] = (int)tmp[st + i];
#pragma unroll
    for (int p = 0; p < 4; p++) s_hist[p][t] = 0;
    __syncthreads();
    for (int64_t i = t; i < L; i += 256) {
      const uint32_t k = id_key(fid[st + i]);
#pragma unroll
      for (int p = 0; p < 4; p++) atomicAdd(&s_hist[p][(k >> (8 * p)) & 255u], 1u);
    }
    __syncthreads();
    if (t < 4) s_skip[t] = 0;
    __syncthreads();
#pragma unroll
    for (int p = 0; p < 4; p++)
      if (s_hist[p][t] == (uint32_t)L) s_skip[p] = 1;  // digit p is the same for all
    __syncthreads();
    // the keys: in fid as raw ids before the first pass that runs, then as
    // keys in the buffer the last pass wrote
    bool in_tmp = false, fid_raw = true;
    for (int p = 0; p < 4; p++) {
      if (s_skip[p]) continue;
      // exclusive offsets of digit p within the segment
      {
        const uint32_t v = s_hist[p][t];
        uint32_t tot;
        const uint32_t ex = block_excl_scan<4>(v, s_wsum, &tot);
        s_base[t] = ex;
      }
      __syncthreads();
      for (int64_t c0 = 0; c0 < L; c0 += kSegBlockMax) {
        const int Lc = (int)(L - c0 < kSegBlockMax ? L - c0 : kSegBlockMax);
        const int kUse = (Lc + 255) / 256;
        const int wbase = (t >> 6) * 64 * kUse + lane_id();
        uint64_t x[kRunPer];
#pragma unroll
        for (int k = 0; k < kRunPer; k++) {
          const int q = wbase + 64 * k;
          uint32_t key = 0;
          if (k < kUse && q < Lc)
            key = in_tmp    ? tmp[st + c0 + q]
                  : fid_raw ? id_key(fid[st + c0 + q])
                            : (uint32_t)fid[st + c0 + q];
          x[k] = (k < kUse && q < Lc) ? ((uint64_t)key << 32) | (uint32_t)q : ~0ull;
        }
        lds_pass<4, uint64_t>(x, kUse, wbase, 32 + 8 * p, s_net, s_cnt, s_wsum, s_match);
        // item at chunk position q: digit d, output base[d] + q - start of d
        // in the chunk (the padding sorts after every item of the chunk)
#pragma unroll
        for (int k = 0; k < kRunPer; k++) {
          const int q = wbase + 64 * k;
          if (k < kUse && q < Lc) {
            const uint32_t key = (uint32_t)(x[k] >> 32);
            const uint32_t d = (key >> (8 * p)) & 255u;
            const int64_t o = st + (int64_t)s_base[d] + (q - (int)s_cnt[0][d]);
            if (in_tmp) fid[o] = (int)key;  // (a key; converted after the last pass)
            else tmp[o] = key;
          }
        }
        __syncthreads();
        // the digits' next offsets: this chunk's items of digit t
        {
          const uint32_t s0 = s_cnt[0][t];
          const uint32_t s1 = t < 255 ? s_cnt[0][t + 1] : (uint32_t)Lc;
          // (digit 255 of the padding sorts after the real ones)
          const uint32_t cnt = (t < 255 ? s1 : (s1 < s0 ? s0 : s1)) - s0;
          s_base[t] += cnt < (uint32_t)Lc ? cnt : (uint32_t)Lc;
        }
        __syncthreads();
      }
      if (in_tmp) fid_raw = false;  // this pass wrote keys into fid
      in_tmp = !in_tmp;
    }
    // keys -> ids: the last pass left keys in tmp (copy back) or in fid
    if (in_tmp || !fid_raw) {
      for (int64_t i = t; i < L; i += 256) {
        const uint32_t key = in_tmp ? tmp[st + i] : (uint32_t)fid[st + i];
        fid[st + i] = key_id(key);
      }
    }
    __syncthreads();  // the shared arrays are reused by the next segment
  }
}

// ---------------------------------------------------------------------------
// Canonical pair order: (query id, box id) lexicographic, as the reference's
// two host stable sorts leave it (src/rtree_test_frontier.cu:102-105).  Pairs
// are packed into one 64-bit key (query id high, box id with its sign bit
// flipped low, so that unsigned order is signed int order) and radix sorted.
__global__ __launch_bounds__(256) void k_pair_pack(int64_t n, const int* __restrict__ fid,
                                                   const int* __restrict__ tid, BuildCtrl* c,
                                                   uint64_t* __restrict__ keys,
                                                   uint32_t* __restrict__ zero_ptr,
                                                   int64_t zero_words) {
  __shared__ uint32_t sh[8 * 256];
  for (int t = threadIdx.x; t < 8 * 256; t += 256) sh[t] = 0;
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += stride) {
    uint64_t k = ((uint64_t)((uint32_t)tid[i] ^ 0x80000000u) << 32) |
                 (uint64_t)((uint32_t)fid[i] ^ 0x80000000u);
    keys[i] = k;
#pragma unroll
    for (int p = 0; p < 8; p++) atomicAdd(&sh[p * 256 + (uint32_t)((k >> (8 * p)) & 255u)], 1u);
  }
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < zero_words; i += stride) zero_ptr[i] = 0u;
  __syncthreads();
  for (int t = threadIdx.x; t < 8 * 256; t += 256)
    if (sh[t]) atomicAdd(&c->hist[t], sh[t]);
}

__global__ __launch_bounds__(256) void k_pair_unpack(int64_t n, const BuildCtrl* c,
                                                     const uint64_t* __restrict__ k0,
                                                     const uint64_t* __restrict__ k1,
                                                     int* __restrict__ fid,
                                                     int* __restrict__ tid) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = (c->npass & 1) ? k1[i] : k0[i];
  tid[i] = (int)((uint32_t)(k >> 32) ^ 0x80000000u);
  fid[i] = (int)((uint32_t)k ^ 0x80000000u);
}

// Any order of pairs: a full radix sort of (query id, box id) keys.
// tid_only: the pairs are already in fid order, so the stable passes over the
// tid digits alone sort them.
int pairs_full_sort(Context& ctx, int* fid, int* tid, int64_t n, hipStream_t s, bool tid_only) {
  if (n >= (int64_t)1 << 30)
    return set_error(CURTREE_ERANGE,
                     "canonical sort: pairs not in query order take at most 2^30-1 per call");
  const int64_t ntiles = (n + kSortTile - 1) / kSortTile;
  const size_t status_words = (size_t)8 * ntiles * 256;
  Workspace& ws = ctx.ws;
  int rc;
  if ((rc = ws.reserve(WsSortKeys0, n * 8)) || (rc = ws.reserve(WsSortKeys1, n * 8)) ||
      (rc = ws.reserve(WsSortStatus, status_words * 4)) ||
      (rc = ws.reserve(WsSortCtrl, sizeof(BuildCtrl))))
    return rc;
  BuildCtrl* c = (BuildCtrl*)ws.ptr(WsSortCtrl);
  uint64_t* k0 = (uint64_t*)ws.ptr(WsSortKeys0);
  uint64_t* k1 = (uint64_t*)ws.ptr(WsSortKeys1);
  uint32_t* status = (uint32_t*)ws.ptr(WsSortStatus);
  Prof& pf = ctx.prof;
  pf.mark(s, "k_build_init");
  hipLaunchKernelGGL(k_build_init, dim3(1), dim3(256), 0, s, c, (uint32_t*)nullptr);
  const unsigned sgrid = (unsigned)std::min<int64_t>((n + 1023) / 1024, 2048);
  pf.mark(s, "k_pair_pack");
  hipLaunchKernelGGL(k_pair_pack, dim3(sgrid), dim3(256), 0, s, n, fid, tid, c, k0, status,
                     (int64_t)status_words);
  pf.mark(s, "k_sort_plan");
  const int steps = tid_only ? 4 : 8;
  hipLaunchKernelGGL(k_sort_plan, dim3(1), dim3(256), 0, s, n, 8, c, 0, steps,
                     tid_only ? 0x7654ull : 0x76543210ull, 1, 0, 1, 0);
  for (int p = 0; p < steps; p++) {
    pf.mark(s, "k_onesweep_pairs");
    hipLaunchKernelGGL(k_onesweep<uint64_t>, dim3((unsigned)std::min(ntiles, kOnesweepGrid)),
                       dim3(kSortThreads), 0, s, n, p, k0, k1, (uint32_t*)nullptr,
                       (uint32_t*)nullptr, c, status + (size_t)p * ntiles * 256, 1);
  }
  pf.mark(s, "k_pair_unpack");
  hipLaunchKernelGGL(k_pair_unpack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, c, k0,
                     k1, fid, tid);
  pf.mark(s, nullptr);
  CT_HIP(hipGetLastError());
  return CURTREE_OK;
}

int pairs_canonical_sort(Context& ctx, int* fid, int* tid, int64_t n, hipStream_t s) {
  if (n < 0) return set_error(CURTREE_EINVAL, "canonical sort: negative count");
  if (n <= 1) return CURTREE_OK;
  if (!fid || !tid) return set_error(CURTREE_EINVAL, "canonical sort: null array");
  Workspace& ws = ctx.ws;
  int rc;
  // lists: wave-class segments (8 B, at most n/2 of them), longer ones (16 B,
  // at most n/1025); the huge class's work array (4 B per pair)
  // (a class of segments longer than m holds at most n / (m + 1) of them)
  // chunk head counts and their offsets; then the head list and the class
  // lists (sized once the number of heads is known); the huge class's work
  // array (4 B per pair)
  const int64_t nchunks = (n + kSegChunk - 1) / kSegChunk;
  if ((rc = ws.reserve(WsSortKeys0, (size_t)n * 4)) ||
      (rc = ws.reserve(WsSortVals0, (size_t)nchunks * 4)) ||
      (rc = ws.reserve(WsSortVals1, (size_t)nchunks * 8)) ||
      (rc = ws.reserve(WsSortCtrl, sizeof(SegCtrl) > sizeof(BuildCtrl) ? sizeof(SegCtrl)
                                                                         : sizeof(BuildCtrl))))
    return rc;
  SegCtrl* c = (SegCtrl*)ws.ptr(WsSortCtrl);
  uint32_t* cnt = (uint32_t*)ws.ptr(WsSortVals0);
  uint64_t* off = (uint64_t*)ws.ptr(WsSortVals1);
  const unsigned cgrid = (unsigned)std::min<int64_t>(nchunks, 4096);
  Prof& pf = ctx.prof;
  pf.mark(s, "k_seg_find");
  hipLaunchKernelGGL(k_seg_init, dim3(1), dim3(64), 0, s, c);
  hipLaunchKernelGGL(k_seg_count, dim3(cgrid), dim3(256), 0, s, n, (const int*)tid, cnt, c);
  hipLaunchKernelGGL(k_seg_scan, dim3(1), dim3(1024), 0, s, nchunks, (const uint32_t*)cnt, off,
                     c);
  pf.mark(s, nullptr);
  CT_HIP(hipGetLastError());
  SegCtrl hc;
  CT_HIP(hipMemcpyAsync(&hc, c, sizeof hc, hipMemcpyDeviceToHost, s));
  CT_HIP(hipStreamSynchronize(s));
  // query ids that ever decrease (e.g. the self-join's box ids): the full sort
  if (hc.unordered) return pairs_full_sort(ctx, fid, tid, n, s);
  const uint64_t nh = hc.nheads;
  // (a class of segments longer than m holds at most n / (m + 1) of them; the
  // buckets of long segments come on top: at most 2 + (segment length) / 1024
  // per segment, so fewer than n / 1024 + n / 8192 in all)
  SegBufs b;
  const uint64_t extra = (uint64_t)(n / 1024 + n / 8192 + 16);
  b.cap[0] = std::min<uint64_t>(nh, (uint64_t)(n / 2 + 1)) + extra;
  b.cap[1] = std::min<uint64_t>(nh, (uint64_t)(n / (kSegWaveMax + 1) + 1)) + extra;
  b.cap[2] = std::min<uint64_t>(nh, (uint64_t)(n / (kSegBlockMax + 1) + 1)) + extra;
  b.cap[3] = std::min<uint64_t>(nh, (uint64_t)(n / (kSegBigMax + 1) + 1));
  b.cap[4] = (uint64_t)(n / (kSegBigMax + 1) + 1);
  b.off[0] = 0;
  b.off[1] = 0;
  b.off[2] = b.cap[1];
  b.off[3] = b.off[2] + b.cap[2];
  b.off[4] = b.off[3] + b.cap[3];
  const uint64_t nbig = b.off[4] + b.cap[4];
  if ((rc = ws.reserve(WsSortKeys1, (size_t)(nh * 8 + b.cap[0] * 8 + nbig * 16)))) return rc;
  uint64_t* heads = (uint64_t*)ws.ptr(WsSortKeys1);
  b.small = heads + nh;
  b.big = b.small + b.cap[0];
  int* tmp = (int*)ws.ptr(WsSortKeys0);
  pf.mark(s, "k_seg_find");
  hipLaunchKernelGGL(k_seg_heads, dim3(cgrid), dim3(256), 0, s, n, (const int*)tid,
                     (const uint64_t*)off, heads);
  hipLaunchKernelGGL(k_seg_classify, dim3((unsigned)std::min<uint64_t>((nh + 255) / 256, 4096)),
                     dim3(256), 0, s, n, (const uint64_t*)heads, c, b);
  // long segments: bucketed into the work array and listed in the classes
  // below (every class launch reads its count on the device; an empty class
  // exits)
  auto grid = [](uint64_t want, uint64_t cap) {
    return dim3((unsigned)std::max<uint64_t>(1, std::min<uint64_t>(want, cap)));
  };
  pf.mark(s, "k_seg_split");
  hipLaunchKernelGGL(k_seg_split, grid(b.cap[3], 512), dim3(1024), 0, s, c, b, fid, tmp);
  pf.mark(s, "k_seg_wave");
  hipLaunchKernelGGL(k_seg_wave, grid((b.cap[0] + 3) / 4, 2048), dim3(256), 0, s,
                     (const SegCtrl*)c, b, fid, (const int*)tmp);
  pf.mark(s, "k_seg_block");
  hipLaunchKernelGGL((k_seg_block<4, 1>), grid(b.cap[1], 1024), dim3(256), 0, s,
                     (const SegCtrl*)c, b, fid, (const int*)tmp);
  pf.mark(s, "k_seg_block16");
  hipLaunchKernelGGL((k_seg_block<16, 2>), grid(b.cap[2], 256), dim3(1024), 0, s,
                     (const SegCtrl*)c, b, fid, (const int*)tmp);
  pf.mark(s, "k_seg_huge");
  hipLaunchKernelGGL(k_seg_huge, grid(b.cap[4], 768), dim3(256), 0, s, (const SegCtrl*)c, b,
                     fid, (uint32_t*)tmp);
  pf.mark(s, nullptr);
  CT_HIP(hipGetLastError());
  CT_HIP(hipMemcpyAsync(&hc, c, sizeof hc, hipMemcpyDeviceToHost, s));
  CT_HIP(hipStreamSynchronize(s));
  if (hc.too_long)
    return set_error(CURTREE_ERANGE,
                     "canonical sort: a query holds 2^32 or more pairs (not sorted)");
  return CURTREE_OK;
}

}  // namespace ct

