// ct_build.hip -- bulk load on gfx950: Morton keys -> onesweep LSD radix sort of
// (key, input position) -> gather of the caller's box arrays into sorted order
// fused with the leaf-parent MBR reduce -> upper levels.
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

// Sort steps: one onesweep pass each, over digit digit[step].  A plain LSD
// sort runs digits 0..D-1; the 64-bit Morton sort runs the high half first
// (digits 4..7), fixes runs of equal high halves in place (k_fixup) and only
// if a run is too long for that runs the full LSD sort (steps 4..11).
constexpr int kMaxSteps = 12;
struct BuildCtrl {
  uint32_t ext[kExtSlots][4];      // ordered-uint lox, loy, hix, hiy of centroids, spread
  uint32_t hist[kMaxDigits * 256];  // global digit histograms
  uint32_t base[kMaxDigits * 256];  // exclusive digit offsets
  uint32_t ticket[kMaxSteps];
  int32_t digit[kMaxSteps];        // digit of each step
  int32_t active[kMaxSteps];       // step s runs
  int32_t src[kMaxSteps];          // step s reads buffer src[s]
  int32_t write_keys[kMaxSteps];   // step s writes keys (not the last step)
  int32_t npass;                   // steps that ran so far (planned): parity of the data
  int32_t first;                   // first step that runs (-1: none)
  uint32_t fallback;               // k_fixup found a run too long to fix in place
  uint32_t nruns;                  // long runs listed for k_fixup_long
  uint32_t trivial;                // bit d: digit d is the same for every key
};

__global__ __launch_bounds__(256) void k_build_init(BuildCtrl* c) {
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

// Key generation + AoS staging copy + all-digit histograms + status zeroing.
// REF32: key = z_order((ushort)cx, (ushort)cy) as in mbr2key
// (src/rtree_bulkload_thrust.cuh:52-60), stored with the sign bit flipped so
// that an unsigned radix sort yields the reference's signed int32 order.
template <int MODE>
__global__ __launch_bounds__(256) void k_keygen(int64_t n, const float* __restrict__ xmin,
                                                const float* __restrict__ ymin,
                                                const float* __restrict__ xmax,
                                                const float* __restrict__ ymax,
                                                const int* __restrict__ id, BuildCtrl* c,
                                                void* __restrict__ keys_out,
                                                uint4* __restrict__ aos,
                                                uint32_t* __restrict__ zero_ptr,
                                                int64_t zero_words) {
  constexpr int D = MODE == kKeyRef32 ? 4 : 8;
  __shared__ uint32_t sh[D * 256];
  for (int t = threadIdx.x; t < D * 256; t += 256) sh[t] = 0;
  float lox = 0.f, loy = 0.f, sx = 0.f, sy = 0.f;
  if (MODE == kKeyMorton64) {
    uint32_t e4[4];
    read_extent(c, e4);
    axis_params(e4[0], e4[2], &lox, &sx);
    axis_params(e4[1], e4[3], &loy, &sy);
  }
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += stride) {
    float a = xmin[i], b = ymin[i], cc = xmax[i], d = ymax[i];
    int ident = id[i];
    aos[2 * i] = make_uint4(__float_as_uint(a), __float_as_uint(b), __float_as_uint(cc),
                            __float_as_uint(d));
    aos[2 * i + 1] = make_uint4((uint32_t)ident, 0u, 0u, 0u);
    if (MODE == kKeyRef32) {
      float cx = (a + cc) / 2.0f;
      float cy = (b + d) / 2.0f;
      uint32_t z = (spread16(f2u16_sat(cy)) << 1) | spread16(f2u16_sat(cx));
      uint32_t k = z ^ 0x80000000u;
      ((uint32_t*)keys_out)[i] = k;
#pragma unroll
      for (int p = 0; p < 4; p++) atomicAdd(&sh[p * 256 + ((k >> (8 * p)) & 255u)], 1u);
    } else {
      float cx = (a + cc) * 0.5f;
      float cy = (b + d) * 0.5f;
      uint64_t k = (spread32(q32(cy, loy, sy)) << 1) | spread32(q32(cx, lox, sx));
      ((uint64_t*)keys_out)[i] = k;
#pragma unroll
      for (int p = 0; p < 8; p++)
        atomicAdd(&sh[p * 256 + (uint32_t)((k >> (8 * p)) & 255u)], 1u);
    }
  }
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < zero_words; i += stride) zero_ptr[i] = 0u;
  __syncthreads();
  for (int t = threadIdx.x; t < D * 256; t += 256)
    if (sh[t]) atomicAdd(&c->hist[t], sh[t]);
}

// Step plan: exclusive digit offsets for digits 0..D-1 and, for steps
// [s0, s0+ns) over the digits packed 4 bits each in `digits` (lowest first),
// which steps run and from which buffer.  A step whose digit is the same for
// every key is skipped (it would be the identity permutation).  `need_one`:
// at least one step must run (a later stage reads the sorted buffers).
__device__ void plan_steps(BuildCtrl* c, int s0, int ns, uint64_t digits, bool need_one) {
  int J = c->npass, last = -1, first = c->first;
  for (int k = 0; k < ns; k++) {
    const int st = s0 + k, d = (int)((digits >> (4 * k)) & 15u);
    c->digit[st] = d;
    c->ticket[st] = 0;
    const bool triv = (c->trivial >> d) & 1u;
    if (triv && !(need_one && k == ns - 1 && J == c->npass && last < 0)) {
      c->active[st] = 0;
    } else {
      c->active[st] = 1;
      c->src[st] = J & 1;
      c->write_keys[st] = 1;
      if (first < 0) first = st;
      last = st;
      J++;
    }
  }
  c->npass = J;
  c->first = first;
}

__global__ __launch_bounds__(256) void k_sort_plan(int64_t n, int D, BuildCtrl* c, int ns,
                                                   uint64_t digits, int keep_last_keys) {
  __shared__ uint32_t s_tmp[4];
  __shared__ uint32_t s_triv;
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
    plan_steps(c, 0, ns, digits, keep_last_keys != 0);
    // the last step writes values only, unless a later stage reads the keys
    if (!keep_last_keys)
      for (int st = ns - 1; st >= 0; st--)
        if (c->active[st]) {
          c->write_keys[st] = 0;
          break;
        }
  }
}

// ---------------------------------------------------------------------------
// One onesweep LSD pass over 8 bits.  Tile = kSortThreads x kSortItems keys,
// warp-striped (wave w owns keys [w*64*I, (w+1)*64*I) of the tile, item i of
// lane l is key w*64*I + i*64 + l), so ballot ranking in (item, lane) order is
// the input order and the pass is stable.  Per-digit tile counts are chained
// across tiles by decoupled look-back; tiles are numbered by an atomic ticket
// so every tile waits only on tiles that have already started.
#ifndef CT_SORT_THREADS
#define CT_SORT_THREADS 512
#endif
#ifndef CT_SORT_ITEMS
#define CT_SORT_ITEMS 12
#endif
#ifndef CT_LOOKBACK
#define CT_LOOKBACK 4
#endif
#ifndef CT_VALS_EARLY
#define CT_VALS_EARLY 0
#endif
constexpr int kSortThreads = CT_SORT_THREADS;  // 256 or 512
constexpr int kSortWaves = kSortThreads / kWave;
constexpr int kSortItems = CT_SORT_ITEMS;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr uint32_t kFlagA = 1u << 30, kFlagP = 2u << 30, kValMask = (1u << 30) - 1u;
static_assert(kSortThreads >= 256, "one thread per digit");

template <class KT>
__global__ __launch_bounds__(kSortThreads) void k_onesweep(
    int64_t n, int step, KT* __restrict__ k0, KT* __restrict__ k1, uint32_t* __restrict__ v0,
    uint32_t* __restrict__ v1, BuildCtrl* c, uint32_t* __restrict__ status, int keys_only) {
  __shared__ uint32_t s_wh[kSortWaves][256];
  __shared__ uint32_t s_start[256];
  __shared__ uint32_t s_adj[256];
  __shared__ uint32_t s_tmp[kSortWaves];
  __shared__ uint32_t s_tile;
  __shared__ KT s_buf[kSortTile];
  __shared__ uint8_t s_dig[kSortTile];

  if (!c->active[step]) return;
  const int t = threadIdx.x, w = t >> 6, lane = lane_id();
  const int pass = c->digit[step];
  const int shift = 8 * pass;
  if (t == 0) s_tile = atomicAdd(&c->ticket[step], 1u);
  for (int i = t; i < kSortWaves * 256; i += kSortThreads) (&s_wh[0][0])[i] = 0;
  __syncthreads();
  const int64_t tile = s_tile;
  const int src = c->src[step];
  const bool ident = step == c->first;  // first step: values are input positions
  const KT* kin = src ? k1 : k0;
  KT* kout = src ? k0 : k1;
  const uint32_t* vin = src ? v1 : v0;
  uint32_t* vout = src ? v0 : v1;
  const bool wk = keys_only || c->write_keys[step] != 0;

  const int64_t tbase = tile * kSortTile;
  const int64_t wbase = tbase + (int64_t)w * (kWave * kSortItems);
  KT key[kSortItems];
  uint32_t pos[kSortItems];  // rank within (wave, digit), later the tile-local position
#pragma unroll
  for (int i = 0; i < kSortItems; i++) {
    int64_t idx = wbase + i * kWave + lane;
    key[i] = idx < n ? kin[idx] : (KT)0;
  }
#if CT_VALS_EARLY
  // values in flight with the keys (one memory round trip per tile)
  uint32_t val[kSortItems];
#pragma unroll
  for (int i = 0; i < kSortItems; i++) {
    int64_t idx = wbase + i * kWave + lane;
    val[i] = (idx < n && !keys_only) ? (ident ? (uint32_t)idx : vin[idx]) : 0u;
  }
#endif
  const uint64_t lt = lanemask_lt();
#pragma unroll
  for (int i = 0; i < kSortItems; i++) {
    int64_t idx = wbase + i * kWave + lane;
    bool valid = idx < n;
    uint32_t d = (uint32_t)(key[i] >> shift) & 255u;
    uint64_t m = ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      uint64_t bb = ballot((d >> b) & 1u);
      m &= ((d >> b) & 1u) ? bb : ~bb;
    }
    uint32_t old = s_wh[w][d];
    pos[i] = old + (uint32_t)popc64(m & lt);
    if (valid && (m & lt) == 0) s_wh[w][d] = old + (uint32_t)popc64(m);
  }
  __syncthreads();

  // threads 0..255 own one digit each: wave offsets, tile count, look-back
  const int dg = t;
  uint32_t tot = 0, excl = 0;
  if (t < 256) {
#pragma unroll
    for (int i = 0; i < kSortWaves; i++) {
      const uint32_t ci = s_wh[i][dg];
      s_wh[i][dg] = tot;
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
          uint32_t f = v[k] & ~kValMask;
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
  const uint32_t start = block_excl_scan<kSortWaves>(tot, s_tmp, &ttot);
  if (t < 256) {
    s_start[dg] = start;
    s_adj[dg] = c->base[pass * 256 + dg] + excl - start;
  }
  __syncthreads();

  // local reorder of keys through LDS, then scatter runs of equal digits
#pragma unroll
  for (int i = 0; i < kSortItems; i++) {
    int64_t idx = wbase + i * kWave + lane;
    uint32_t d = (uint32_t)(key[i] >> shift) & 255u;
    pos[i] += s_start[d] + s_wh[w][d];
    if (idx < n) s_buf[pos[i]] = key[i];
  }
  __syncthreads();
  const int64_t rem = n - tbase;
  const int tn = rem < kSortTile ? (int)rem : kSortTile;
#pragma unroll
  for (int i = 0; i < kSortItems; i++) {
    int j = t + i * kSortThreads;
    if (j < tn) {
      KT kk = s_buf[j];
      uint32_t d = (uint32_t)(kk >> shift) & 255u;
      s_dig[j] = (uint8_t)d;
      if (wk) kout[s_adj[d] + (uint32_t)j] = kk;
    }
  }
  if (keys_only) return;
  // values: reorder through LDS, scatter
#if !CT_VALS_EARLY
  uint32_t val[kSortItems];
#pragma unroll
  for (int i = 0; i < kSortItems; i++) {
    int64_t idx = wbase + i * kWave + lane;
    val[i] = idx < n ? (ident ? (uint32_t)idx : vin[idx]) : 0u;
  }
#endif
  __syncthreads();
  uint32_t* s_val = (uint32_t*)s_buf;
#pragma unroll
  for (int i = 0; i < kSortItems; i++) {
    int64_t idx = wbase + i * kWave + lane;
    if (idx < n) s_val[pos[i]] = val[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kSortItems; i++) {
    int j = t + i * kSortThreads;
    if (j < tn) vout[s_adj[s_dig[j]] + (uint32_t)j] = s_val[j];
  }
}

// ---------------------------------------------------------------------------
// Runs of keys with equal high halves, after the high-half passes: each run is
// sorted in place by (low half, input position), which makes the whole array
// sorted by (key, input position) -- the stable full-key order -- without the
// four low-digit passes.  With n boxes spread over 2^32 high-half cells runs
// are short.
//   k_fixup: a block loads a tile (plus a halo of kFixShort) into LDS; the
// owner thread of each run of 2..kFixShort starting in the tile sorts it by
// insertion, and the block writes the rewritten positions back (keys and
// values, so that fallback passes see consistent data).  Longer runs are
// listed for k_fixup_long (one block per run, bitonic in LDS, up to
// kFixLong); a longer run, or a full list, sets `fallback` and the full LSD
// sort runs instead.
constexpr int kFixShort = 16;
constexpr int kFixLong = 4096;
constexpr int kFixLongSlots = 4096;  // blocks of k_fixup_long (they loop over the list)

__device__ __forceinline__ uint32_t key_hi(const uint64_t* k, int64_t i) {
  return (uint32_t)(k[i] >> 32);
}

constexpr int kFixTile = 2048;  // run starts owned by one block of k_fixup
constexpr int kFixSpan = kFixTile + kFixShort;

__global__ __launch_bounds__(256) void k_fixup(int64_t n, BuildCtrl* c, uint64_t* k0,
                                               uint64_t* k1, uint32_t* v0, uint32_t* v1,
                                               uint32_t* runs, uint32_t runs_cap) {
  __shared__ uint32_t s_hi[kFixSpan + 1];  // s_hi[j + 1]: high half of position tb + j
  __shared__ uint64_t s_ck[kFixSpan];      // (low half << 32) | input position
  __shared__ uint8_t s_dirty[kFixSpan];    // position rewritten by its run's owner
  const int cur = c->npass & 1;
  uint64_t* keys = cur ? k1 : k0;
  uint32_t* vals = cur ? v1 : v0;
  const int64_t tb = (int64_t)blockIdx.x * kFixTile;
  const int64_t te = tb + kFixTile < n ? tb + kFixTile : n;
  const int64_t le = te + kFixShort < n ? te + kFixShort : n;
  const int nl = (int)(le - tb), nown = (int)(te - tb);
  const int t = threadIdx.x;
  for (int j = t; j < nl; j += 256) {
    const uint64_t k = keys[tb + j];
    s_hi[j + 1] = (uint32_t)(k >> 32);
    s_ck[j] = ((k & 0xFFFFFFFFull) << 32) | vals[tb + j];
    s_dirty[j] = 0;
  }
  // the position before the tile (a value unequal to the first for tile 0)
  if (t == 0) s_hi[0] = tb > 0 ? (uint32_t)(keys[tb - 1] >> 32) : ~(uint32_t)(keys[0] >> 32);
  __syncthreads();
  // owner threads: runs of 2..kFixShort starting in the tile, insertion sort
  for (int j = t; j < nown; j += 256) {
    const uint32_t h = s_hi[j + 1];
    if (s_hi[j] == h || j + 1 >= nl || s_hi[j + 2] != h) continue;  // no start of a run >= 2
    int L = 2;
    while (L <= kFixShort && j + L < nl && s_hi[j + L + 1] == h) L++;
    if (L > kFixShort || (j + L == nl && le < n)) {  // long (or reaching past the halo)
      const uint32_t at = atomicAdd(&c->nruns, 1u);
      if (at < runs_cap) runs[at] = (uint32_t)(tb + j);
      else atomicOr(&c->fallback, 1u);
      continue;
    }
    for (int a2 = j + 1; a2 < j + L; a2++) {
      const uint64_t v = s_ck[a2];
      int b2 = a2 - 1;
      while (b2 >= j && s_ck[b2] > v) {
        s_ck[b2 + 1] = s_ck[b2];
        b2--;
      }
      s_ck[b2 + 1] = v;
    }
    for (int q = j; q < j + L; q++) s_dirty[q] = 1;
  }
  __syncthreads();
  // coalesced write-back of the rewritten positions
  for (int j = t; j < nl; j += 256) {
    if (!s_dirty[j]) continue;
    const uint64_t v = s_ck[j];
    keys[tb + j] = ((uint64_t)s_hi[j + 1] << 32) | (v >> 32);
    vals[tb + j] = (uint32_t)v;
  }
}

__global__ __launch_bounds__(256) void k_fixup_long(int64_t n, BuildCtrl* c, uint64_t* k0,
                                                    uint64_t* k1, uint32_t* v0, uint32_t* v1,
                                                    const uint32_t* runs, uint32_t runs_cap) {
  __shared__ uint64_t s_net[kFixLong];
  __shared__ int s_len;
  const uint32_t nr = c->nruns < runs_cap ? c->nruns : runs_cap;
  if (blockIdx.x >= nr || c->fallback) return;
  const int cur = c->npass & 1;
  uint64_t* keys = cur ? k1 : k0;
  uint32_t* vals = cur ? v1 : v0;
  const int t = threadIdx.x;
  for (uint32_t q = blockIdx.x; q < nr; q += gridDim.x) {
    const int64_t i = runs[q];
    const uint32_t h = key_hi(keys, i);
    if (t == 0) {
      int64_t L = kFixShort;
      while (L <= kFixLong && i + L < n && key_hi(keys, i + L) == h) L++;
      s_len = (int)L;
      if (L > kFixLong) atomicOr(&c->fallback, 1u);
    }
    __syncthreads();
    const int L = s_len;
    if (L > kFixLong) return;  // the full LSD sort takes over
    int m = 1;
    while (m < L) m <<= 1;
    for (int j = t; j < m; j += 256)
      s_net[j] = j < L ? ((keys[i + j] & 0xFFFFFFFFull) << 32) | vals[i + j] : ~0ull;
    __syncthreads();
    for (int k = 2; k <= m; k <<= 1) {
      for (int d = k >> 1; d > 0; d >>= 1) {
        for (int j = t; j < m; j += 256) {
          const int p = j ^ d;
          if (p > j) {
            const uint64_t x = s_net[j], y = s_net[p];
            if ((x > y) == ((j & k) == 0)) {
              s_net[j] = y;
              s_net[p] = x;
            }
          }
        }
        __syncthreads();
      }
    }
    const uint64_t hi = (uint64_t)h << 32;
    for (int j = t; j < L; j += 256) {
      keys[i + j] = hi | (s_net[j] >> 32);
      vals[i + j] = (uint32_t)s_net[j];
    }
    __syncthreads();
  }
}

// After k_fixup: if a run was too long, plan the full LSD sort (steps 4..11
// over digits 0..7) from the current order; else nothing more runs.
__global__ void k_fallback_plan(BuildCtrl* c) {
  if (threadIdx.x != 0 || !c->fallback) return;
  plan_steps(c, 4, 8, 0x76543210ull, false);
  for (int st = kMaxSteps - 1; st >= 4; st--)
    if (c->active[st]) {
      c->write_keys[st] = 0;
      break;
    }
}

// Status words of the fallback steps (cleared only when they will run).
__global__ __launch_bounds__(256) void k_fallback_zero(const BuildCtrl* c, uint32_t* p,
                                                       int64_t words) {
  if (!c->fallback) return;
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < words; i += (int64_t)gridDim.x * 256)
    p[i] = 0u;
}

// ---------------------------------------------------------------------------
// Group reduction of an MBR over G consecutive lanes (G a power of two <= 64).
// `has` marks lanes that hold a value; invalid lanes are skipped.
template <int G>
__device__ __forceinline__ void group_mbr(float& a, float& b, float& cc, float& d, bool has) {
#pragma unroll
  for (int o = 1; o < G; o <<= 1) {
    float pa = __shfl_xor(a, o), pb = __shfl_xor(b, o);
    float pc = __shfl_xor(cc, o), pd = __shfl_xor(d, o);
    bool ph = __shfl_xor((int)has, o) != 0;
    if (ph) {
      if (has) {
        a = mbr_min(a, pa);
        b = mbr_min(b, pb);
        cc = mbr_max(cc, pc);
        d = mbr_max(d, pd);
      } else {
        a = pa;
        b = pb;
        cc = pc;
        d = pd;
        has = true;
      }
    }
  }
}

// Gather the boxes into sorted order (in place on the caller's arrays, reading
// the staging copy) and, for power-of-two F <= 64, build the leaf-parent level
// in the same pass: node g covers sorted boxes [g*F, min(g*F+F, N)) with
// pos = R_sz + g*F and len = count (src/rtree_bulkload_thrust.cuh:136-149).
template <int G>
__global__ __launch_bounds__(256) void k_gather_l0(int64_t n, const BuildCtrl* c,
                                                   const uint32_t* __restrict__ v0,
                                                   const uint32_t* __restrict__ v1,
                                                   const uint4* __restrict__ aos, float* xmin,
                                                   float* ymin, float* xmax, float* ymax,
                                                   int* id, TreeView tr, int64_t l0start,
                                                   int64_t r_sz) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  const int J = c->npass;
  const uint32_t* vf = (J & 1) ? v1 : v0;
  float a = 0.f, b = 0.f, cc = 0.f, d = 0.f;
  const bool has = i < n;
  if (has) {
    uint32_t s = J ? vf[i] : (uint32_t)i;
    uint4 r0 = aos[2 * (size_t)s];
    uint4 r1 = aos[2 * (size_t)s + 1];
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
  if constexpr (G > 0) {
    group_mbr<G>(a, b, cc, d, has);
    if (has && (i % G) == 0) {
      int64_t g = i / G, node = l0start + g;
      tr.xmin[node] = a;
      tr.ymin[node] = b;
      tr.xmax[node] = cc;
      tr.ymax[node] = d;
      tr.pos[node] = (int)(r_sz + i);
      int64_t len = n - i;
      tr.len[node] = (int)(len < G ? len : G);
    }
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

template <int G>
static void launch_gather(hipStream_t s, int64_t n, const BuildCtrl* c, const uint32_t* v0,
                          const uint32_t* v1, const uint4* aos, curtree_bbox_f32* box,
                          TreeView tr, int64_t l0start, int64_t r_sz) {
  unsigned grid = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_gather_l0<G>, dim3(grid), dim3(256), 0, s, n, c, v0, v1, aos, box->xmin,
                     box->ymin, box->xmax, box->ymax, box->id, tr, l0start, r_sz);
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
    if (alloc) rt_free(rt);
    return code;
  };

  const int D = key_mode == kKeyRef32 ? 4 : 8;
  const size_t kb = key_mode == kKeyRef32 ? 4 : 8;
  const int64_t ntiles = (n + kSortTile - 1) / kSortTile;
  const bool m64 = key_mode == kKeyMorton64;
  const int nsteps = m64 ? kMaxSteps : 4;  // Morton: 4 high-half steps + 8 fallback steps
  const size_t step_words = (size_t)ntiles * 256;
  const size_t status_words = (size_t)nsteps * step_words;
  Workspace& ws = ctx.ws;
  if ((rc = ws.reserve(WsBuildKeys0, n * kb)) || (rc = ws.reserve(WsBuildKeys1, n * kb)) ||
      (rc = ws.reserve(WsBuildVals0, n * 4)) || (rc = ws.reserve(WsBuildVals1, n * 4)) ||
      (rc = ws.reserve(WsBuildAos, n * 32)) ||
      (rc = ws.reserve(WsBuildStatus, status_words * 4)) ||
      (rc = ws.reserve(WsBuildCtrl, sizeof(BuildCtrl))))
    return fail(rc);
  // long runs of equal high halves (64-bit keys): at most n / (kFixShort + 1)
  const uint32_t runs_cap = m64 ? (uint32_t)std::min<int64_t>(n / (kFixShort + 1) + 1, 1 << 22) : 1;
  if ((rc = ws.reserve(WsBuildRuns, (size_t)runs_cap * 4))) return fail(rc);
  uint32_t* runs = (uint32_t*)ws.ptr(WsBuildRuns);
  BuildCtrl* c = (BuildCtrl*)ws.ptr(WsBuildCtrl);
  void* k0 = ws.ptr(WsBuildKeys0);
  void* k1 = ws.ptr(WsBuildKeys1);
  uint32_t* v0 = (uint32_t*)ws.ptr(WsBuildVals0);
  uint32_t* v1 = (uint32_t*)ws.ptr(WsBuildVals1);
  uint4* aos = (uint4*)ws.ptr(WsBuildAos);
  uint32_t* status = (uint32_t*)ws.ptr(WsBuildStatus);

  Prof& pf = ctx.prof;
  pf.mark(s, "k_build_init");
  hipLaunchKernelGGL(k_build_init, dim3(1), dim3(256), 0, s, c);
  const unsigned sgrid = (unsigned)std::min<int64_t>((n + 1023) / 1024, 2048);
  if (key_mode == kKeyMorton64) {
    pf.mark(s, "k_extent");
    hipLaunchKernelGGL(k_extent, dim3(sgrid), dim3(256), 0, s, n, box->xmin, box->ymin,
                       box->xmax, box->ymax, c);
  }
  // keygen clears the status words of the first four steps
  pf.mark(s, "k_keygen");
  if (key_mode == kKeyRef32)
    hipLaunchKernelGGL(k_keygen<kKeyRef32>, dim3(sgrid), dim3(256), 0, s, n, box->xmin,
                       box->ymin, box->xmax, box->ymax, box->id, c, k0, aos, status,
                       (int64_t)(4 * step_words));
  else
    hipLaunchKernelGGL(k_keygen<kKeyMorton64>, dim3(sgrid), dim3(256), 0, s, n, box->xmin,
                       box->ymin, box->xmax, box->ymax, box->id, c, k0, aos, status,
                       (int64_t)(4 * step_words));
  pf.mark(s, "k_sort_plan");
  // REF32: digits 0..3; Morton: the high half (digits 4..7) first
  hipLaunchKernelGGL(k_sort_plan, dim3(1), dim3(256), 0, s, n, D, c, 4,
                     m64 ? 0x7654ull : 0x3210ull, m64 ? 1 : 0);
  for (int st = 0; st < nsteps; st++) {
    if (st == 4) {
      pf.mark(s, "k_fixup");
      const unsigned fgrid = (unsigned)((n + kFixTile - 1) / kFixTile);
      hipLaunchKernelGGL(k_fixup, dim3(fgrid), dim3(256), 0, s, n, c, (uint64_t*)k0,
                         (uint64_t*)k1, v0, v1, runs, runs_cap);
      hipLaunchKernelGGL(k_fixup_long, dim3(kFixLongSlots), dim3(256), 0, s, n, c, (uint64_t*)k0,
                         (uint64_t*)k1, v0, v1, (const uint32_t*)runs, runs_cap);
      hipLaunchKernelGGL(k_fallback_plan, dim3(1), dim3(64), 0, s, c);
      hipLaunchKernelGGL(k_fallback_zero, dim3(1024), dim3(256), 0, s, (const BuildCtrl*)c,
                         status + 4 * step_words, (int64_t)(8 * step_words));
    }
    // fallback steps exit at once unless k_fixup asked for them
    pf.mark(s, st < 4 ? "k_onesweep" : "k_onesweep_fallback");
    uint32_t* stw = status + (size_t)st * step_words;
    if (key_mode == kKeyRef32)
      hipLaunchKernelGGL(k_onesweep<uint32_t>, dim3((unsigned)ntiles), dim3(kSortThreads), 0, s,
                         n, st, (uint32_t*)k0, (uint32_t*)k1, v0, v1, c, stw, 0);
    else
      hipLaunchKernelGGL(k_onesweep<uint64_t>, dim3((unsigned)ntiles), dim3(kSortThreads), 0, s,
                         n, st, (uint64_t*)k0, (uint64_t*)k1, v0, v1, c, stw, 0);
  }

  TreeView tr{rt->xmin, rt->ymin, rt->xmax, rt->ymax, rt->pos, rt->len};
  const int64_t l0start = lp.r_sz - lp.n[0];
  const int G = pow2_group(F);
  pf.mark(s, "k_gather_l0");
  switch (G) {
    case 2: launch_gather<2>(s, n, c, v0, v1, aos, box, tr, l0start, lp.r_sz); break;
    case 4: launch_gather<4>(s, n, c, v0, v1, aos, box, tr, l0start, lp.r_sz); break;
    case 8: launch_gather<8>(s, n, c, v0, v1, aos, box, tr, l0start, lp.r_sz); break;
    case 16: launch_gather<16>(s, n, c, v0, v1, aos, box, tr, l0start, lp.r_sz); break;
    case 32: launch_gather<32>(s, n, c, v0, v1, aos, box, tr, l0start, lp.r_sz); break;
    case 64: launch_gather<64>(s, n, c, v0, v1, aos, box, tr, l0start, lp.r_sz); break;
    default:
      launch_gather<0>(s, n, c, v0, v1, aos, box, tr, l0start, lp.r_sz);
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
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return fail(set_error(CURTREE_EHIP, "build: launch failed: %s", hipGetErrorString(e)));
  return CURTREE_OK;
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

int pairs_canonical_sort(Context& ctx, int* fid, int* tid, int64_t n, hipStream_t s) {
  if (n < 0) return set_error(CURTREE_EINVAL, "canonical sort: negative count");
  if (n <= 1) return CURTREE_OK;
  if (!fid || !tid) return set_error(CURTREE_EINVAL, "canonical sort: null array");
  if (n >= (int64_t)1 << 30)
    return set_error(CURTREE_ERANGE, "canonical sort: at most 2^30-1 pairs per call");
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
  hipLaunchKernelGGL(k_build_init, dim3(1), dim3(256), 0, s, c);
  const unsigned sgrid = (unsigned)std::min<int64_t>((n + 1023) / 1024, 2048);
  pf.mark(s, "k_pair_pack");
  hipLaunchKernelGGL(k_pair_pack, dim3(sgrid), dim3(256), 0, s, n, fid, tid, c, k0, status,
                     (int64_t)status_words);
  pf.mark(s, "k_sort_plan");
  hipLaunchKernelGGL(k_sort_plan, dim3(1), dim3(256), 0, s, n, 8, c, 8, 0x76543210ull, 1);
  for (int p = 0; p < 8; p++) {
    pf.mark(s, "k_onesweep_pairs");
    hipLaunchKernelGGL(k_onesweep<uint64_t>, dim3((unsigned)ntiles), dim3(kSortThreads), 0, s, n,
                       p, k0, k1, (uint32_t*)nullptr, (uint32_t*)nullptr, c,
                       status + (size_t)p * ntiles * 256, 1);
  }
  pf.mark(s, "k_pair_unpack");
  hipLaunchKernelGGL(k_pair_unpack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, c, k0,
                     k1, fid, tid);
  pf.mark(s, nullptr);
  CT_HIP(hipGetLastError());
  return CURTREE_OK;
}

}  // namespace ct
