// engine.hip -- many light-map frames in shared launches: batching and continuous serving.
//
// Frames live in slots. A slot at stage s (bounces done) takes part in the next step as:
//   s == 0      : clear its light map, generate its rays, then bounce 1
//   0 <= s < d  : hit (live rays) -> count -> per-frame scan -> fused bounce pass (evolve, splat of
//                 the rays that become DIFFUSE, stable partition; bounce.hpp)     (bounce s+1)
//   s == d      : realigning hit -> copy of the live rays -> free
// A step is one fixed sequence of launches for every slot at once: a planning kernel turns the
// device-side live counts into per-kind work-item prefixes (and the hitter's split factor); the
// hitter runs one wave per (64-ray group, split) item of any frame, so frames of any size and
// scene cost share the chip; the bounce pass takes one tile per workgroup, the other stages walk
// their items grid-stride. The grids are powers of two sized by the host from the frames' ray
// counts and stages, so each set of them is captured into a HIP graph once and replayed. The host tracks stages itself (every admitted frame advances exactly one stage per
// step), so it never reads device state to learn that a frame has finished: an event recorded
// after each step tells when.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <deque>
#include <array>
#include <map>
#include <vector>

#include "hitcore.hpp"
#include "path.hpp"
#include "bounce.hpp"

namespace spl {

constexpr int MAX_SLOTS = 1024;
constexpr int MAXD = SPLAT_ENGINE_MAX_DEPTH;
// K_HIT: 64-ray groups; K_WAVE: hitter waves; K_KEYS: the 64-ray groups whose miss keys are filled
// before the first hit (new frames hit with splits, which merge into the keys)
enum Kind { K_RAYGEN = 0, K_CLEAR, K_HIT, K_TILE, K_FINAL, K_WAVE, K_KEYS, N_KINDS };
constexpr int HIT_ITEM = 64;     // rays per hitter work item (one wave)
constexpr int CLEAR_ITEM = 4096;  // light-map floats per clear item
constexpr int QUEUE_BLOCKS_PER_CU = 8;  // the queue hitter's grid: 8 workgroups of 4 waves per CU
constexpr int QUEUES = 8;               // its item ranges: one per XCD
constexpr int QUEUE_STRIDE = 32;        // int32 counters a cache line apart

struct SlotDev {
  splat_adray_t* F;        // final rays (the frame's output order); null: not kept
  splat_adray_t* S[2];     // live rays, ping-pong by stage parity
  int32_t* hit_final;      // their hit indices; null: not kept
  uint64_t* keys[2];       // packed (t, index) nearest hit of each live ray, by stage parity
  int32_t* tile_counts;
  int32_t* tile_offsets;
  float* tex;
  const splat_ftri_t* tris;
  BVScene scene;
  int64_t R;
  int64_t tex_n;
  int32_t n_tris, first_diffuse, depth;
  int32_t last;            // the frame's last stage: depth, or depth - 1 when it keeps only its light map
  float force_rand;
  uint64_t seed;
  splat_optics_t optics;
  splat_pinhole_t cam;
};

struct EngineDev {
  SlotDev* slots;
  splat_lightmap_params_t* lm;
  int32_t* stage;      // -1 = empty
  int64_t* counters;   // [slot][MAXD + 2]: live rays after bounce k (k = 0: R)
  int32_t* pre;        // [kind][MAX_SLOTS + 1] exclusive prefixes of work items
  int32_t* hit_split;  // [slot] splits per 64-ray group this step; negative: flat-scan tail
  int32_t* admit;      // [MAX_SLOTS + 1]: count, then the slots admitted since the last step
  int32_t* queue;      // the step's work queue of hitter items: items taken from each range, zeroed by k_plan
  int32_t n_slots;
};

// The slot of work item i: the largest s with pre[s] <= i (slots with no items share their
// successor's prefix; pre[0] = 0 <= i < pre[n]). The prefixes do not decrease, so that s is the
// number of s' in [1, n) with pre[s'] <= i: the wave compares 64 prefixes per (independent) load
// and counts with a ballot, instead of a binary search's chain of dependent loads. Called by whole
// waves with i the same on every lane.
__device__ __forceinline__ int32_t item_slot(const int32_t* __restrict__ pre, int32_t n, int32_t i) {
  const int32_t lane = threadIdx.x & 63;
  int32_t s = 0;
  for (int32_t b = 1; b < n; b += 64) {
    const int32_t k = b + lane;
    s += __builtin_popcountll(__builtin_amdgcn_ballot_w64(k < n && pre[k] <= i));
  }
  return s;
}
__device__ __forceinline__ int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// one workgroup of MAX_SLOTS threads: per-slot item counts -> exclusive prefixes
__global__ __launch_bounds__(MAX_SLOTS) void k_plan(EngineDev E) {
  const int t = threadIdx.x;
  // frames admitted since the last step start at stage 0 (their slot records were uploaded
  // with the step); the list is consumed
  const int32_t n_admit = E.admit[0];
  if (t < n_admit) E.stage[E.admit[1 + t]] = 0;
  __syncthreads();
  if (t == 0) E.admit[0] = 0;
  if (t < QUEUES) E.queue[t * QUEUE_STRIDE] = 0;
  const bool in = t < E.n_slots;
  const int32_t s = in ? E.stage[t] : -1;
  int32_t cnt[N_KINDS] = {0, 0, 0, 0, 0, 0, 0};
  bool tail = false;
  int32_t leaves = 1;
  if (s >= 0) {
    const SlotDev& D = E.slots[t];
    int64_t* C = E.counters + (int64_t)t * (MAXD + 2);
    if (s == 0) C[0] = D.R;
    const int64_t A = s == 0 ? D.R : C[s];
    if (s == 0) {
      // the first bounce generates its rays where it needs them; a frame without bounces stores them
      cnt[K_RAYGEN] = D.depth == 0 ? (int32_t)cdiv(D.R, TILE) : 0;
      cnt[K_CLEAR] = (int32_t)cdiv(D.tex_n, CLEAR_ITEM);
    }
    cnt[K_HIT] = (int32_t)cdiv(A, HIT_ITEM);
    if (s < D.depth) cnt[K_TILE] = (int32_t)cdiv(A, TILE);
    if (s == D.depth) cnt[K_FINAL] = (int32_t)cdiv(A, TILE);
    tail = bv_mode(D.scene) == BV_TREE && cnt[K_HIT] <= TAIL_GROUPS && !D.scene.no_tail;
    leaves = D.scene.n_bv;
  }
  // the hitter's splits: one factor for the tree walks of all frames, a flat-scan tail split for
  // frames with few live rays (both from the totals, so the items fit the host's grid)
  // (sums and prefix sums over the slots: within each wave by shuffles, then across the waves
  // through LDS -- a few barriers in all, not a tree of them per kind)
  constexpr int NW = MAX_SLOTS / 64;
  __shared__ int32_t wtot[N_KINDS + 2][NW];
  const int lane = t & 63, wv = t >> 6;
  auto wave_incl = [&](int32_t v) {  // inclusive scan within the wave
    for (int off = 1; off < 64; off <<= 1) {
      const int32_t u = __shfl_up(v, off);
      if (lane >= off) v += u;
    }
    return v;
  };
  const int32_t iw = wave_incl(tail ? 0 : cnt[K_HIT]), it = wave_incl(tail ? cnt[K_HIT] : 0);
  if (lane == 63) {
    wtot[N_KINDS][wv] = iw;
    wtot[N_KINDS + 1][wv] = it;
  }
  __syncthreads();
  int32_t g_walk = 0, g_tail = 0;
  for (int j = 0; j < NW; ++j) {
    g_walk += wtot[N_KINDS][j];
    g_tail += wtot[N_KINDS + 1][j];
  }
  const int32_t S_walk = split_factor(g_walk, MAX_SPLIT);
  const int32_t S_slot = tail ? tail_split_factor(g_tail, leaves) : S_walk;
  if (in) E.hit_split[t] = tail ? -S_slot : S_slot;
  cnt[K_WAVE] = cnt[K_HIT] * S_slot;
  cnt[K_KEYS] = s == 0 && S_slot > 1 ? cnt[K_HIT] : 0;  // without splits the hitter stores every key
  // exclusive prefixes of every kind (slots past n_slots count nothing, so the total is theirs)
  int32_t inc[N_KINDS];
#pragma unroll
  for (int k = 0; k < N_KINDS; ++k) inc[k] = wave_incl(cnt[k]);
  if (lane == 63)
#pragma unroll
    for (int k = 0; k < N_KINDS; ++k) wtot[k][wv] = inc[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N_KINDS; ++k) {
    int32_t before = 0, total = 0;
    for (int j = 0; j < NW; ++j) {
      const int32_t v = wtot[k][j];
      before += j < wv ? v : 0;
      total += v;
    }
    int32_t* P = E.pre + k * (MAX_SLOTS + 1);
    if (t < E.n_slots) P[t] = before + inc[k] - cnt[k];
    if (t == 0) P[E.n_slots] = total;
  }
}

__global__ __launch_bounds__(256) void k_b_raygen(EngineDev E) {
  const int32_t n = E.n_slots;
  const int32_t* P = E.pre + K_RAYGEN * (MAX_SLOTS + 1);
  for (int32_t i = blockIdx.x; i < P[n]; i += gridDim.x) {
    const int32_t s = item_slot(P, n, i);
    const SlotDev& D = E.slots[s];
    const int64_t q = (int64_t)(i - P[s]) * TILE + threadIdx.x;
    if (q < D.R) raygen_one(D.cam, q, D.S[0] + q);
  }
  const int32_t* Q = E.pre + K_CLEAR * (MAX_SLOTS + 1);
  for (int32_t i = blockIdx.x; i < Q[n]; i += gridDim.x) {
    const int32_t s = item_slot(Q, n, i);
    // the slot's fields in registers: read through the (unwritten) record each iteration, the
    // float stores below could alias them and force a reload per store
    const cptr<SlotDev> D = as_const(E.slots + s);
    float* const tex = D->tex;
    const int64_t tex_n = D->tex_n;
    const int64_t b = (int64_t)(i - Q[s]) * CLEAR_ITEM;
    if (b + CLEAR_ITEM <= tex_n && (reinterpret_cast<uintptr_t>(tex) & 15) == 0) {
      float4* const t4 = reinterpret_cast<float4*>(tex + b);
      for (int k = threadIdx.x; k < CLEAR_ITEM / 4; k += blockDim.x) t4[k] = make_float4(0.f, 0.f, 0.f, 0.f);
    } else {
      for (int64_t k = b + threadIdx.x; k < b + CLEAR_ITEM && k < tex_n; k += blockDim.x) tex[k] = 0.0f;
    }
  }
  // keys of a new frame's rays <- (Inf, 1) before the hitter's splits merge into them (later
  // stages get theirs from the bounce pass)
  const int32_t* H = E.pre + K_KEYS * (MAX_SLOTS + 1);
  constexpr int PER_BLOCK = 256 / HIT_ITEM;
  for (int32_t i = blockIdx.x * PER_BLOCK + threadIdx.x / HIT_ITEM; i < H[n]; i += gridDim.x * PER_BLOCK) {
    const int32_t s = item_slot(H, n, i);
    const cptr<SlotDev> D = as_const(E.slots + s);
    const int64_t r = (int64_t)(i - H[s]) * HIT_ITEM + (threadIdx.x & (HIT_ITEM - 1));
    if (r < D->R) D->keys[0][r] = KEY_MISS;
  }
}

// one hitter work item: (64-ray group, split) `i` of the step's items over every frame's live rays
__device__ __forceinline__ void b_hit_item(const EngineDev& E, const int32_t* __restrict__ P, int32_t n, int32_t i,
                                           HitStats& hs) {
  const int32_t s = item_slot(P, n, i);
  const int32_t sv = E.hit_split[s];
  const bool tail = sv < 0;
  const int32_t S = tail ? -sv : sv;
  const int32_t local = i - P[s], group = local / S, sp = local - group * S;
  const SlotDev& D = E.slots[s];
  const int32_t st = E.stage[s];
  const int64_t A = E.counters[(int64_t)s * (MAXD + 2) + st];
  const int64_t base = (int64_t)group * HIT_ITEM;
  uint64_t* keys = D.keys[st & 1];
  if (tail) {  // few live rays: flat over the leaves
    if (st == 0 && D.depth > 0) bv_item<false, BV_FLAT, true>(nullptr, A, base, sp, S, D.scene, keys, hs, &D.cam);
    else bv_item<false, BV_FLAT>(D.S[st & 1], A, base, sp, S, D.scene, keys, hs);
  } else if (st == 0 && D.depth > 0) {  // the first bounce: the pinhole's rays, generated here
    switch (bv_mode(D.scene)) {
      case BV_TREE: bv_item<false, BV_TREE, true>(nullptr, A, base, sp, S, D.scene, keys, hs, &D.cam); break;
      case BV_GROUPS: bv_item<false, BV_GROUPS, true>(nullptr, A, base, sp, S, D.scene, keys, hs, &D.cam); break;
      default: bv_item<false, BV_FLAT, true>(nullptr, A, base, sp, S, D.scene, keys, hs, &D.cam); break;
    }
  } else {
    switch (bv_mode(D.scene)) {
      case BV_TREE: bv_item<false, BV_TREE>(D.S[st & 1], A, base, sp, S, D.scene, keys, hs); break;
      case BV_GROUPS: bv_item<false, BV_GROUPS>(D.S[st & 1], A, base, sp, S, D.scene, keys, hs); break;
      default: bv_item<false, BV_FLAT>(D.S[st & 1], A, base, sp, S, D.scene, keys, hs); break;
    }
  }
}

// Hitter: one wave per (64-ray group, split) item over every frame's live rays; the splits of a
// group merge into its keys with a 64-bit atomic minimum. The grid (chosen by the host from the
// frames' ray counts) covers the most items the live counts can need; waves past them leave.
__global__ __launch_bounds__(256) void k_b_hit(EngineDev E) {
  const int32_t n = E.n_slots;
  const int32_t* P = E.pre + K_WAVE * (MAX_SLOTS + 1);
  HitStats hs;
  const int32_t w = (int32_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (w < P[n]) b_hit_item(E, P, n, w, hs);
}

// The same items from one persistent launch (splat_engine_use_queue): one round of workgroups on
// every CU. Each wave first takes the item of its own index; the items past the grid (a step with
// more items than waves) come from the step's work queue: they split into QUEUES contiguous
// ranges, each with a device counter of the items taken (zeroed by the planning kernel, on a
// cache line of its own); a wave takes chunks from the range of its workgroup's XCD (workgroups
// go to the XCDs in turn), then from the others', each chunk an eighth of its range's remaining
// share of a wave (single items at the end), and asks for its next chunk before it works on the
// current one (the atomic's round trip overlaps the work); it leaves when every range is taken.
// A step with no more items than waves never touches the queue: every wave probing all the
// ranges cost serving steps 16% of their throughput. With one counter for all items every take
// serialises on one address (about 10 ns each: the 64-frame batch 2.2x slower with single items);
// larger chunks leave long late-bounce items queued behind each other in one wave.
__global__ __launch_bounds__(256, 8) void k_b_hit_queue(EngineDev E) {
  const int32_t n = E.n_slots;
  const int32_t* P = E.pre + K_WAVE * (MAX_SLOTS + 1);
  const int32_t total = P[n];
  HitStats hs;
  const int32_t w = (int32_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (w < total) b_hit_item(E, P, n, w, hs);  // a wave's first item is its own index
  const int32_t W = (int32_t)gridDim.x * 4;
  if (total <= W) return;  // no item past the grid: the queue is never touched
  // the items past the grid, [W, total), from the work queue
  const int32_t rem = total - W, Wq = W / QUEUES > 0 ? W / QUEUES : 1;
  int32_t q = (int32_t)(blockIdx.x % QUEUES), tried = 0, taken = 0;  // taken: range q's count as last seen
  auto lo_of = [&](int32_t r) { return W + (int32_t)((int64_t)rem * r / QUEUES); };
  // a take from range q: its chunk size and the count before it (lane 0's atomic; read later, so
  // the take is in flight while the wave works on the chunk before)
  int32_t chunk = 0, at_raw = 0;
  auto take = [&]() {
    const int32_t rest = lo_of(q + 1) - lo_of(q) - taken;
    chunk = rest > 8 * Wq ? rest / (8 * Wq) : 1;
    if ((threadIdx.x & 63) == 0) at_raw = atomicAdd(E.queue + q * QUEUE_STRIDE, chunk);
  };
  take();
  while (true) {
    const int32_t at = __builtin_amdgcn_readfirstlane(at_raw), c = chunk;
    const int32_t lo = lo_of(q), hi = lo_of(q + 1);
    if (at >= hi - lo) {  // range q is taken: the next one
      if (++tried == QUEUES) break;
      q = q + 1 < QUEUES ? q + 1 : 0;
      taken = 0;
      take();
      continue;
    }
    taken = at + c;
    take();  // the next chunk, requested before this one's work
    const int32_t end = lo + at + c < hi ? lo + at + c : hi;
    for (int32_t i = lo + at; i < end; ++i) b_hit_item(E, P, n, i, hs);
  }
}

// count pass of the fused bounce (bounce.hpp): ACTIVE rays per tile from the hit keys, one tile
// per wave (its grid holds a wave per tile of the host's bound)
__global__ __launch_bounds__(TILE) void k_b_count(EngineDev E) {
  const int32_t n = E.n_slots;
  const int32_t* P = E.pre + K_TILE * (MAX_SLOTS + 1);
  const int32_t i = (int32_t)blockIdx.x * (TILE / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (i >= P[n]) return;  // whole waves
  const int32_t s = item_slot(P, n, i);
  const cptr<SlotDev> D = as_const(E.slots + s);
  const int32_t st = *as_const(E.stage + s);
  const int64_t A = *as_const(E.counters + (int64_t)s * (MAXD + 2) + st);
  const int32_t tile = i - P[s];
  const int64_t base = (int64_t)tile * TILE;
  const int nvalid = (int)(A - base < TILE ? A - base : TILE);
  BounceParams BP;
  BP.tris = D->tris;
  BP.n_tris = D->n_tris;
  BP.first_diffuse = D->first_diffuse;
  BP.bounce = st + 1;
  BP.force_rand = D->force_rand;
  BP.seed = D->seed;
  BP.optics = D->optics;
  BP.tex = D->tex;
  const bool odd = (st & 1) != 0;
  const uint64_t* keys = (odd ? D->keys[1] : D->keys[0]) + base;
  int c;
  if (st == 0) {
    const splat_pinhole_t cam = D->cam;
    c = bounce_count<true>(nullptr, keys, nvalid, BP, true, cam, base);
  } else {
    const splat_pinhole_t cam{};
    c = bounce_count<false>((odd ? D->S[1] : D->S[0]) + base, keys, nvalid, BP, true, cam, base);
  }
  if ((threadIdx.x & 63) == 0) D->tile_counts[tile] = c;
}

// one workgroup per slot: exclusive scan of its live tiles' ACTIVE counts
__global__ __launch_bounds__(1024) void k_b_scan(EngineDev E) {
  __shared__ int64_t sums[1024];
  const int32_t s = blockIdx.x;
  const int32_t st = E.stage[s];
  if (st < 0) return;
  const SlotDev& D = E.slots[s];
  if (st >= D.depth) return;
  int64_t* C = E.counters + (int64_t)s * (MAXD + 2);
  const int64_t n_tiles = cdiv(C[st], TILE);
  const int t = threadIdx.x;
  const int64_t chunk = (n_tiles + 1023) / 1024;
  const int64_t b = t * chunk, e = b + chunk < n_tiles ? b + chunk : n_tiles;
  int64_t sum = 0;
  for (int64_t i = b; i < e; ++i) sum += D.tile_counts[i];
  sums[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int64_t v = t >= off ? sums[t - off] : 0;
    __syncthreads();
    sums[t] += v;
    __syncthreads();
  }
  int64_t run = sums[t] - sum;
  for (int64_t i = b; i < e; ++i) {
    D.tile_offsets[i] = (int32_t)run;
    run += D.tile_counts[i];
  }
  if (t == 1023) C[st + 1] = sums[1023];
}

// bounce pass (bounce.hpp): evolve, splat the rays that become DIFFUSE, stable partition
// One tile per workgroup (no loop: a grid-stride loop over the tiles made the compiler keep the
// slot's parameters live across iterations, 188 VGPRs and 2 waves per SIMD). Slot fields read one
// by one and the light-map parameters in LDS bring it to 128 VGPRs (4 waves; a few spilled
// registers cost less than the lost wave). The grid holds a workgroup per tile of the host's bound
// (the ray counts of the frames that bounce in this step), which covers every live tile.
__global__ __launch_bounds__(TILE, 4) void k_b_bounce(EngineDev E) {
  const int32_t n = E.n_slots;
  const int32_t* P = E.pre + K_TILE * (MAX_SLOTS + 1);
  __shared__ splat_lightmap_params_t s_lm;  // the slot's light-map parameters
  {
    const int32_t i = blockIdx.x;
    if (i >= P[n]) return;
    const int32_t s = item_slot(P, n, i);
    {
      constexpr int NW = sizeof(splat_lightmap_params_t) / 4;
      const uint32_t* src = reinterpret_cast<const uint32_t*>(E.lm + s);
      for (int w = threadIdx.x; w < NW; w += TILE) reinterpret_cast<uint32_t*>(&s_lm)[w] = src[w];
      __syncthreads();
    }
    // the slot record and counters are read-only in this kernel: read field by field through the
    // constant address space they are scalar loads into SGPRs rather than vector loads into VGPRs
    const cptr<SlotDev> D = as_const(E.slots + s);
    const int32_t st = *as_const(E.stage + s);
    const cptr<int64_t> C = as_const(E.counters + (int64_t)s * (MAXD + 2));
    const int64_t A = C[st];
    const int32_t tile = i - P[s];
    const int64_t base = (int64_t)tile * TILE;
    const int nvalid = (int)(A - base < TILE ? A - base : TILE);
    const int64_t off_a = D->tile_offsets[tile];
    const int64_t off_r = C[st + 1] + (base - off_a);
    BounceParams BP;
    BP.tris = D->tris;
    BP.n_tris = D->n_tris;
    BP.first_diffuse = D->first_diffuse;
    BP.bounce = st + 1;
    BP.force_rand = D->force_rand;
    BP.seed = D->seed;
    BP.optics = D->optics;
    BP.tex = D->tex;
    const bool odd = (st & 1) != 0;
    splat_adray_t* const S0 = D->S[0];
    splat_adray_t* const S1 = D->S[1];
    uint64_t* const K0 = D->keys[0];
    uint64_t* const K1 = D->keys[1];
    const uint64_t* keys = (odd ? K1 : K0) + base;
    uint64_t* keys_next = odd ? K0 : K1;
    if (st == 0) {
      const splat_pinhole_t cam = D->cam;
      bounce_tile<true>(nullptr, keys, keys_next, nvalid, BP, s_lm, cam, base, off_a, off_r, S1, D->F,
                        D->hit_final);
    } else {
      const splat_pinhole_t cam{};
      bounce_tile<false>((odd ? S1 : S0) + base, keys, keys_next, nvalid, BP, s_lm, cam, base, off_a, off_r,
                         odd ? S0 : S1, D->F, D->hit_final);
    }
  }
}

// final stage: live rays to their final slots, then final evolution + splat over the frame
__global__ __launch_bounds__(TILE) void k_b_final(EngineDev E) {
  const int32_t n = E.n_slots;
  const int32_t* P = E.pre + K_FINAL * (MAX_SLOTS + 1);
  constexpr int W = sizeof(splat_adray_t) / 8;
  for (int32_t i = blockIdx.x; i < P[n]; i += gridDim.x) {
    const int32_t s = item_slot(P, n, i);
    const SlotDev& D = E.slots[s];
    const int32_t st = E.stage[s];
    const int64_t A = E.counters[(int64_t)s * (MAXD + 2) + st];
    const int64_t base = (int64_t)(i - P[s]) * TILE;
    const int64_t nw = (A - base < TILE ? A - base : TILE) * W;
    if (D.F) {
      const uint2* src = reinterpret_cast<const uint2*>(D.S[st & 1] + base);
      uint2* dst = reinterpret_cast<uint2*>(D.F + base);
      for (int64_t k = threadIdx.x; k < nw; k += TILE) dst[k] = src[k];
    }
    if (D.hit_final && base + threadIdx.x < A)
      D.hit_final[base + threadIdx.x] = key_index(D.keys[st & 1][base + threadIdx.x]);
  }
}

__global__ void k_b_finish(EngineDev E) {
  for (int s = threadIdx.x; s < E.n_slots; s += blockDim.x) {
    const int32_t st = E.stage[s];
    if (st < 0) continue;
    E.stage[s] = st >= E.slots[s].last ? -1 : st + 1;
  }
}

// ---- host side --------------------------------------------------------------------------------
constexpr size_t ALIGNB = 256;
inline size_t al(size_t x) { return (x + ALIGNB - 1) & ~(ALIGNB - 1); }

struct Block {
  int64_t off, size;
};

// first-fit allocator over the arena, with neighbour merging on free
struct Arena {
  int64_t size = 0;
  std::map<int64_t, int64_t> free_;  // offset -> size
  void init(int64_t n) { size = n; free_.clear(); free_[0] = n; }
  bool alloc(int64_t n, int64_t* off) {
    n = (int64_t)al((size_t)n);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        *off = it->first;
        int64_t rest = it->second - n, o = it->first + n;
        free_.erase(it);
        if (rest > 0) free_[o] = rest;
        return true;
      }
    }
    return false;
  }
  void release(int64_t off, int64_t n) {
    n = (int64_t)al((size_t)n);
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) { it->second += nx->second; free_.erase(nx); }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) { pv->second += it->second; free_.erase(it); }
    }
  }
  int64_t free_bytes() const {
    int64_t s = 0;
    for (auto& kv : free_) s += kv.second;
    return s;
  }
};

struct HostSlot {
  bool used = false;
  int32_t stage = -1;  // host mirror: stage the slot runs at in the next step
  int32_t depth = 0;
  int32_t last = 0;  // last stage (SlotDev::last)
  int64_t R = 0;
  int64_t tag = 0;
  std::vector<Block> blocks;
};

struct Pending {
  hipEvent_t ev;
  std::vector<int32_t> finishing;  // slots whose last stage ran in this step
};

// arena bytes of a frame of R rays (two live-ray buffers, two key buffers, tile counts and
// offsets); with_outputs adds its output arrays (final rays and hit indices), which the caller
// allocates, for a budget that covers both
size_t frame_bytes(int64_t R, bool with_outputs) {
  int64_t tiles = (R + TILE - 1) / TILE;
  size_t b = 2 * al((size_t)R * sizeof(splat_adray_t)) + 2 * al((size_t)R * 8) + 2 * al((size_t)tiles * 4);
  if (with_outputs) b += al((size_t)R * sizeof(splat_adray_t)) + al((size_t)R * 4);
  return b;
}

}  // namespace spl

using namespace spl;

struct splat_engine {
  int32_t max_slots = 0;
  char* arena_base = nullptr;
  Arena arena;
  EngineDev dev{};
  char* dev_mem = nullptr;
  SlotDev* h_slots = nullptr;  // pinned staging of the slot table
  splat_lightmap_params_t* h_lm = nullptr;
  int32_t* h_stage0 = nullptr;
  // admissions are uploaded by the next step: the slot records [up_lo, up_hi] and the list of
  // admitted slots (two pinned buffers, each reused once the copy that read it has completed)
  std::vector<int32_t> admits;
  int32_t up_lo = -1, up_hi = -1;
  int32_t* h_admit[2] = {nullptr, nullptr};
  hipEvent_t admit_ev[2] = {nullptr, nullptr};
  int admit_cur = 0;
  std::vector<HostSlot> slots;
  std::deque<Pending> pending;
  std::vector<hipEvent_t> free_events;
  std::map<std::array<int64_t, 4>, hipGraphExec_t> graphs;  // one captured step per set of grid sizes
  hipStream_t cap_stream = nullptr;  // graphs are captured here (the null stream cannot capture)
  int use_graph = 1;
  int use_queue = 0;                 // the hitter: a persistent grid taking items from a work queue
  int64_t queue_hit_blocks = 256 * 8;  // its grid: one round of workgroups on every CU
};

namespace {

// Grids of one step, sized by the host from the frames' ray counts and stages: the hitter's covers
// every item the live counts can need; tile_blocks >= the tile count of the frames that bounce in
// this step, so the bounce pass takes one tile per workgroup over it and the count pass one tile
// per wave over a quarter of it; the first-stage and last-stage kernels stride over grids for the
// frames at those stages (1 when none is).
struct StepGrids {
  int64_t hit_blocks, tile_blocks, first_blocks, final_blocks;  // grid classes (grid_class)
  std::array<int64_t, 4> key() const { return {hit_blocks, tile_blocks, first_blocks, final_blocks}; }
};

// The smallest grid class >= n: 1, 2, 3, then m * 2^k with m in {4, 5, 6, 7}. Four classes per
// power of two bound the dispatch of workgroups past the items to 25% (a power-of-two grid
// dispatched up to twice the items), with a few graph variants per octave.
int64_t grid_class(int64_t n) {
  if (n <= 3) return n < 1 ? 1 : n;
  int64_t k = 0;
  while ((7LL << k) < n) ++k;  // now 7 * 2^k >= n
  for (int64_t m = 4; m <= 7; ++m)
    if ((m << k) >= n) return m << k;
  return 7LL << k;
}

int32_t launch_step_kernels(splat_engine* e, const StepGrids& g, hipStream_t s) {
  const EngineDev& E = e->dev;
  const int64_t hit_blocks = g.hit_blocks;
  hipLaunchKernelGGL(k_plan, dim3(1), dim3(MAX_SLOTS), 0, s, E);
  hipLaunchKernelGGL(k_b_raygen, dim3((unsigned)g.first_blocks), dim3(256), 0, s, E);
  if (e->use_queue) hipLaunchKernelGGL(k_b_hit_queue, dim3((unsigned)hit_blocks), dim3(256), 0, s, E);
  else hipLaunchKernelGGL(k_b_hit, dim3((unsigned)hit_blocks), dim3(256), 0, s, E);
  hipLaunchKernelGGL(k_b_count, dim3((unsigned)((g.tile_blocks + 3) / 4)), dim3(TILE), 0, s, E);
  hipLaunchKernelGGL(k_b_scan, dim3(e->max_slots), dim3(1024), 0, s, E);
  hipLaunchKernelGGL(k_b_bounce, dim3((unsigned)g.tile_blocks), dim3(TILE), 0, s, E);
  hipLaunchKernelGGL(k_b_final, dim3((unsigned)g.final_blocks), dim3(TILE), 0, s, E);
  hipLaunchKernelGGL(k_b_finish, dim3(1), dim3(1024), 0, s, E);
  return hipGetLastError() == hipSuccess ? SPLAT_OK : SPLAT_ERR_HIP;
}

void release_slot(splat_engine* e, int32_t s) {
  for (const Block& b : e->slots[s].blocks) e->arena.release(b.off, b.size);
  e->slots[s].blocks.clear();
  e->slots[s].used = false;
  e->slots[s].stage = -1;
}

}  // namespace

extern "C" {

size_t splat_engine_frame_bytes(const splat_pinhole_t* cam, int32_t with_outputs) {
  if (!cam || cam->height < 1 || cam->width < 1 || cam->samples_per_pixel < 1) return 0;
  return frame_bytes((int64_t)cam->height * cam->width * cam->samples_per_pixel, with_outputs != 0);
}

int32_t splat_engine_create(int64_t arena_bytes, int32_t max_slots, splat_engine_t** out) {
  if (!out || arena_bytes <= 0 || max_slots < 1 || max_slots > MAX_SLOTS) return SPLAT_ERR_INVALID_ARG;
  *out = nullptr;
  splat_engine* e = new splat_engine();
  e->max_slots = max_slots;
  size_t dev_bytes = al(sizeof(SlotDev) * max_slots) + al(sizeof(splat_lightmap_params_t) * max_slots) +
                     al(sizeof(int32_t) * max_slots) + al(sizeof(int64_t) * (size_t)max_slots * (MAXD + 2)) +
                     al(sizeof(int32_t) * N_KINDS * (MAX_SLOTS + 1)) + al(sizeof(int32_t) * MAX_SLOTS) +
                     al(sizeof(int32_t) * (MAX_SLOTS + 1)) + al(sizeof(int32_t) * QUEUES * QUEUE_STRIDE);
  if (hipMalloc(&e->arena_base, (size_t)arena_bytes) != hipSuccess) { delete e; return SPLAT_ERR_BUDGET; }
  if (hipMalloc(&e->dev_mem, dev_bytes) != hipSuccess) { (void)hipFree(e->arena_base); delete e; return SPLAT_ERR_HIP; }
  if (hipHostMalloc(reinterpret_cast<void**>(&e->h_slots), sizeof(SlotDev) * max_slots) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&e->h_lm), sizeof(splat_lightmap_params_t) * max_slots) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&e->h_stage0), sizeof(int32_t) * max_slots) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&e->h_admit[0]), sizeof(int32_t) * (MAX_SLOTS + 1)) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&e->h_admit[1]), sizeof(int32_t) * (MAX_SLOTS + 1)) != hipSuccess ||
      hipEventCreateWithFlags(&e->admit_ev[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&e->admit_ev[1], hipEventDisableTiming) != hipSuccess) {
    splat_engine_destroy(e);
    return SPLAT_ERR_HIP;
  }
  e->arena.init(arena_bytes);
  char* p = e->dev_mem;
  e->dev.slots = reinterpret_cast<SlotDev*>(p); p += al(sizeof(SlotDev) * max_slots);
  e->dev.lm = reinterpret_cast<splat_lightmap_params_t*>(p); p += al(sizeof(splat_lightmap_params_t) * max_slots);
  e->dev.stage = reinterpret_cast<int32_t*>(p); p += al(sizeof(int32_t) * max_slots);
  e->dev.counters = reinterpret_cast<int64_t*>(p); p += al(sizeof(int64_t) * (size_t)max_slots * (MAXD + 2));
  e->dev.pre = reinterpret_cast<int32_t*>(p); p += al(sizeof(int32_t) * N_KINDS * (MAX_SLOTS + 1));
  e->dev.hit_split = reinterpret_cast<int32_t*>(p); p += al(sizeof(int32_t) * MAX_SLOTS);
  e->dev.admit = reinterpret_cast<int32_t*>(p); p += al(sizeof(int32_t) * (MAX_SLOTS + 1));
  e->dev.queue = reinterpret_cast<int32_t*>(p);
  e->dev.n_slots = max_slots;
  {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
      e->queue_hit_blocks = (int64_t)cus * QUEUE_BLOCKS_PER_CU;
  }
  for (int i = 0; i < max_slots; ++i) e->h_stage0[i] = -1;
  e->h_admit[0][0] = 0;
  if (hipMemcpy(e->dev.stage, e->h_stage0, sizeof(int32_t) * max_slots, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->dev.admit, e->h_admit[0], sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess) {
    splat_engine_destroy(e);
    return SPLAT_ERR_HIP;
  }
  e->h_stage0[0] = 0;
  e->slots.resize(max_slots);
  *out = e;
  return SPLAT_OK;
}

int32_t splat_engine_destroy(splat_engine_t* e) {
  if (!e) return SPLAT_ERR_INVALID_ARG;
  (void)hipDeviceSynchronize();
  for (auto& p : e->pending) (void)hipEventDestroy(p.ev);
  for (auto ev : e->free_events) (void)hipEventDestroy(ev);
  for (auto& g : e->graphs)
    if (g.second) (void)hipGraphExecDestroy(g.second);
  if (e->cap_stream) (void)hipStreamDestroy(e->cap_stream);
  if (e->h_slots) (void)hipHostFree(e->h_slots);
  if (e->h_lm) (void)hipHostFree(e->h_lm);
  if (e->h_stage0) (void)hipHostFree(e->h_stage0);
  for (int i = 0; i < 2; ++i) {
    if (e->h_admit[i]) (void)hipHostFree(e->h_admit[i]);
    if (e->admit_ev[i]) (void)hipEventDestroy(e->admit_ev[i]);
  }
  if (e->dev_mem) (void)hipFree(e->dev_mem);
  if (e->arena_base) (void)hipFree(e->arena_base);
  delete e;
  return SPLAT_OK;
}

int32_t splat_engine_admit(splat_engine_t* e, const splat_frame_request_t* rq, void* stream) {
  if (!e || !rq || !rq->tex || !rq->tris || rq->n_tris < 1) return SPLAT_ERR_INVALID_ARG;
  const splat_pinhole_t& cam = rq->cam;
  if (cam.height < 1 || cam.width < 1 || cam.samples_per_pixel < 1 || !(cam.fov_half_sin > 0.f)) return SPLAT_ERR_INVALID_ARG;
  if (rq->hs.kind != SPLAT_HITTER_BV) return SPLAT_ERR_INVALID_ARG;
  if (int32_t err = check_hitset(&rq->hs)) return err;
  if (rq->depth < 0 || rq->depth > MAXD) return SPLAT_ERR_UNSUPPORTED;
  const splat_lightmap_params_t& lm = rq->lm;
  if (int32_t err = check_lightmap(&lm)) return err;
  if (rq->optics.ior_model != SPLAT_IOR_REFERENCE_LINEAR && rq->optics.ior_model != SPLAT_IOR_CAUCHY) return SPLAT_ERR_INVALID_ARG;
  const int64_t R = (int64_t)cam.height * cam.width * cam.samples_per_pixel;
  if (R > (int64_t)INT32_MAX / 2) return SPLAT_ERR_UNSUPPORTED;
  int32_t slot = -1;
  for (int32_t i = 0; i < e->max_slots; ++i)
    if (!e->slots[i].used) { slot = i; break; }
  if (slot < 0) return SPLAT_ERR_BUDGET;
  HostSlot& hsl = e->slots[slot];
  const int64_t tiles = (R + TILE - 1) / TILE;
  const std::vector<int64_t> sizes = {R * (int64_t)sizeof(splat_adray_t), R * (int64_t)sizeof(splat_adray_t), R * 8,
                                      tiles * 4, tiles * 4, R * 8};
  std::vector<Block> got;
  for (int64_t n : sizes) {
    int64_t off;
    if (!e->arena.alloc(n, &off)) {
      for (const Block& b : got) e->arena.release(b.off, b.size);
      return SPLAT_ERR_BUDGET;
    }
    got.push_back(Block{off, n});
  }
  auto at = [&](size_t k) { return e->arena_base + got[k].off; };
  SlotDev& D = e->h_slots[slot];
  std::memset(&D, 0, sizeof(D));
  D.S[0] = reinterpret_cast<splat_adray_t*>(at(0));
  D.S[1] = reinterpret_cast<splat_adray_t*>(at(1));
  D.keys[0] = reinterpret_cast<uint64_t*>(at(2));
  D.tile_counts = reinterpret_cast<int32_t*>(at(3));
  D.tile_offsets = reinterpret_cast<int32_t*>(at(4));
  D.keys[1] = reinterpret_cast<uint64_t*>(at(5));
  D.F = rq->rays_out;
  D.hit_final = rq->hit_out;
  // a frame that keeps only its light map ends with its last bounce: the final stage (realigning
  // hit, copy of the live rays) only feeds the outputs it does not keep
  const bool map_only = !rq->rays_out && !rq->hit_out;
  D.last = map_only && rq->depth > 0 ? rq->depth - 1 : rq->depth;
  D.tex = rq->tex;
  D.tris = rq->tris;
  D.scene = bv_scene(&rq->hs);
  D.R = R;
  D.tex_n = (int64_t)lm.width * lm.height * lm.channels;
  D.n_tris = rq->n_tris;
  D.first_diffuse = rq->first_diffuse;
  D.depth = rq->depth;
  D.force_rand = rq->force_rand;
  D.seed = rq->seed;
  D.optics = rq->optics;
  D.cam = cam;
  e->h_lm[slot] = lm;
  (void)stream;  // nothing is enqueued here: the next step uploads the admissions
  e->admits.push_back(slot);
  e->up_lo = e->up_lo < 0 || slot < e->up_lo ? slot : e->up_lo;
  e->up_hi = slot > e->up_hi ? slot : e->up_hi;
  hsl.used = true;
  hsl.stage = 0;
  hsl.depth = rq->depth;
  hsl.last = D.last;
  hsl.R = R;
  hsl.tag = rq->tag;
  hsl.blocks = got;
  return SPLAT_OK;
}

int32_t splat_engine_step(splat_engine_t* e, void* stream) {
  if (!e) return SPLAT_ERR_INVALID_ARG;
  hipStream_t s = static_cast<hipStream_t>(stream);
  bool any = false;
  int64_t groups = 0;       // upper bound of the live 64-ray groups in this step
  int64_t tail_groups = 0;  // upper bound of those in frames that may scan flat (<= TAIL_GROUPS each)
  int64_t first = 0, last = 0;  // work items of the first-stage and last-stage kernels
  int64_t tiles = 0;            // bound of the bounce tiles: frames that bounce in this step
  int64_t first_rg = 0, first_clear = 0, first_keys = 0;
  for (size_t i = 0; i < e->slots.size(); ++i) {
    const HostSlot& h = e->slots[i];
    if (!h.used || h.stage < 0) continue;
    any = true;
    const int64_t g = (h.R + HIT_ITEM - 1) / HIT_ITEM;
    groups += g;
    tail_groups += g < TAIL_GROUPS ? g : TAIL_GROUPS;
    if (h.stage == 0) {
      if (h.depth == 0) first_rg += (h.R + TILE - 1) / TILE;
      first_clear += (e->h_slots[i].tex_n + CLEAR_ITEM - 1) / CLEAR_ITEM;
      first_keys += (h.R + 4 * HIT_ITEM - 1) / (4 * HIT_ITEM);
    }
    if (h.stage == h.depth) last += (h.R + TILE - 1) / TILE;
    else tiles += (h.R + TILE - 1) / TILE;
  }
  if (!any) return SPLAT_OK;
  first = first_rg > first_clear ? first_rg : first_clear;
  first = first > first_keys ? first : first_keys;
  if (!e->admits.empty()) {
    // the slot records of the frames admitted since the last step (one copy per table over the
    // range they span: the other records in it are unchanged copies of their device records),
    // then the list of those slots; k_plan starts them at stage 0
    const int32_t lo = e->up_lo, n = e->up_hi - e->up_lo + 1;
    int32_t* buf = e->h_admit[e->admit_cur];
    if (hipEventSynchronize(e->admit_ev[e->admit_cur]) != hipSuccess) return SPLAT_ERR_HIP;  // its last copy is done
    buf[0] = (int32_t)e->admits.size();
    for (size_t i = 0; i < e->admits.size(); ++i) buf[1 + i] = e->admits[i];
    if (hipMemcpyAsync(e->dev.slots + lo, e->h_slots + lo, sizeof(SlotDev) * n, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(e->dev.lm + lo, e->h_lm + lo, sizeof(splat_lightmap_params_t) * n, hipMemcpyHostToDevice, s) !=
            hipSuccess ||
        hipMemcpyAsync(e->dev.admit, buf, sizeof(int32_t) * (1 + e->admits.size()), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipEventRecord(e->admit_ev[e->admit_cur], s) != hipSuccess)
      return SPLAT_ERR_HIP;
    e->admit_cur ^= 1;
    e->admits.clear();
    e->up_lo = e->up_hi = -1;
  }
  // hitter grid: a grid class of workgroups (4 waves each) that covers every item the live
  // counts can need, so a few graphs serve any mix of frames
  const int64_t need = (max_hit_items(groups, tail_groups) + 3) / 4;
  StepGrids grids{grid_class(need), 1, 1, 1};
  if (e->use_queue && grids.hit_blocks > e->queue_hit_blocks) grids.hit_blocks = e->queue_hit_blocks;
  if (grids.hit_blocks > (int64_t)1 << 24) return SPLAT_ERR_UNSUPPORTED;  // grid of at most 2^32 threads
  // the striding kernels: powers of two up to 65536 workgroups (a few graph variants in all)
  while (grids.first_blocks < first && grids.first_blocks < 65536) grids.first_blocks <<= 1;
  while (grids.final_blocks < last && grids.final_blocks < 65536) grids.final_blocks <<= 1;
  // one bounce workgroup (and one count wave) per tile of the frames that bounce: tiles <=
  // groups / 4 + slots, so this never passes hit_blocks
  grids.tile_blocks = grid_class(tiles);
  if (e->use_graph) {
    // one graph per set of grids and hitter kernel (the mode can change between steps)
    std::array<int64_t, 4> key = grids.key();
    if (e->use_queue) key[0] = -key[0];
    hipGraphExec_t& ge = e->graphs[key];
    if (!ge) {
      hipGraph_t g;
      if (!e->cap_stream && hipStreamCreateWithFlags(&e->cap_stream, hipStreamNonBlocking) != hipSuccess) return SPLAT_ERR_HIP;
      if (hipStreamBeginCapture(e->cap_stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return SPLAT_ERR_HIP;
      int32_t st = launch_step_kernels(e, grids, e->cap_stream);
      if (hipStreamEndCapture(e->cap_stream, &g) != hipSuccess || st != SPLAT_OK) return SPLAT_ERR_HIP;
      hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) {
        ge = nullptr;
        return SPLAT_ERR_HIP;
      }
    }
    if (hipGraphLaunch(ge, s) != hipSuccess) return SPLAT_ERR_HIP;
  } else if (int32_t st = launch_step_kernels(e, grids, s)) {
    return st;
  }
  Pending p;
  if (e->free_events.empty()) {
    if (hipEventCreateWithFlags(&p.ev, hipEventDisableTiming) != hipSuccess) return SPLAT_ERR_HIP;
  } else {
    p.ev = e->free_events.back();
    e->free_events.pop_back();
  }
  for (int32_t i = 0; i < e->max_slots; ++i) {
    HostSlot& h = e->slots[i];
    if (!h.used || h.stage < 0) continue;
    if (h.stage >= h.last) {
      p.finishing.push_back(i);
      h.stage = -2;  // finished on the device once this step's event completes
    } else {
      ++h.stage;
    }
  }
  if (hipEventRecord(p.ev, s) != hipSuccess) return SPLAT_ERR_HIP;
  e->pending.push_back(std::move(p));
  return SPLAT_OK;
}

int32_t splat_engine_poll(splat_engine_t* e, int32_t wait, int64_t* tags, int32_t max_tags, int32_t* n_out) {
  if (!e || !n_out || (max_tags > 0 && !tags)) return SPLAT_ERR_INVALID_ARG;
  *n_out = 0;
  while (!e->pending.empty()) {
    Pending& p = e->pending.front();
    if ((int32_t)p.finishing.size() > max_tags - *n_out) break;  // caller's buffer is full
    if (wait) {
      if (hipEventSynchronize(p.ev) != hipSuccess) return SPLAT_ERR_HIP;
    } else {
      hipError_t q = hipEventQuery(p.ev);
      if (q == hipErrorNotReady) break;
      if (q != hipSuccess) return SPLAT_ERR_HIP;
    }
    for (int32_t sl : p.finishing) {
      tags[(*n_out)++] = e->slots[sl].tag;
      release_slot(e, sl);
    }
    e->free_events.push_back(p.ev);
    e->pending.pop_front();
  }
  return SPLAT_OK;
}

int32_t splat_engine_sync_steps(splat_engine_t* e, int32_t keep) {
  if (!e || keep < 0) return SPLAT_ERR_INVALID_ARG;
  const int32_t n = (int32_t)e->pending.size();
  if (n > keep && hipEventSynchronize(e->pending[(size_t)(n - keep - 1)].ev) != hipSuccess) return SPLAT_ERR_HIP;
  return SPLAT_OK;
}

int32_t splat_engine_in_flight(const splat_engine_t* e) {
  if (!e) return 0;
  int32_t n = 0;
  for (auto& h : e->slots) n += h.used ? 1 : 0;
  return n;
}

int32_t splat_engine_runnable(const splat_engine_t* e) {
  if (!e) return 0;
  int32_t n = 0;
  for (auto& h : e->slots) n += h.used && h.stage >= 0 ? 1 : 0;
  return n;
}

int64_t splat_engine_arena_free(const splat_engine_t* e) { return e ? e->arena.free_bytes() : 0; }

int32_t splat_param_sizes(int64_t* out6) {
  if (!out6) return SPLAT_ERR_INVALID_ARG;
  out6[0] = sizeof(splat_optics_t);
  out6[1] = sizeof(splat_hitset_t);
  out6[2] = sizeof(splat_lightmap_params_t);
  out6[3] = sizeof(splat_pinhole_t);
  out6[4] = sizeof(splat_run_params_t);
  out6[5] = sizeof(splat_frame_request_t);
  return SPLAT_OK;
}

int32_t splat_engine_use_graph(splat_engine_t* e, int32_t on) {
  if (!e) return SPLAT_ERR_INVALID_ARG;
  e->use_graph = on != 0;
  return SPLAT_OK;
}

int32_t splat_engine_use_queue(splat_engine_t* e, int32_t on) {
  if (!e) return SPLAT_ERR_INVALID_ARG;
  e->use_queue = on != 0;
  return SPLAT_OK;
}

}  // extern "C"
