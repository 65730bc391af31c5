// hitcore.hpp -- the per-wave ray/triangle traversal shared by the single-frame hitter kernels
// (hit.hip) and the multi-frame engine (engine.hip).
#pragma once

#include "path.hpp"

namespace spl {

// ---- ray loading ----------------------------------------------------------------------------
struct RayLanes {
  float ox, oy, oz, dx, dy, dz;
  int64_t ig;
};
__device__ __forceinline__ RayLanes load_ray_lanes(const splat_adray_t* __restrict__ rays, int64_t r, bool valid) {
  RayLanes L{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0};
  if (valid) {
    const splat_adray_t* p = rays + r;
    L.ox = p->pos[0]; L.oy = p->pos[1]; L.oz = p->pos[2];
    L.dx = p->dir[0]; L.dy = p->dir[1]; L.dz = p->dir[2];
    L.ig = p->ignore_tri;
  }
  return L;
}

// the lanes of pinhole ray r, generated on the fly (the value parts of camera_ray)
__device__ __forceinline__ RayLanes cam_ray_lanes(const splat_pinhole_t& cam, int64_t r, bool valid) {
  RayLanes L{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0};
  if (valid) {
    const splat_adray_t g = raygen_ray(cam, r);
    L.ox = g.pos[0]; L.oy = g.pos[1]; L.oz = g.pos[2];
    L.dx = g.dir[0]; L.dy = g.dir[1]; L.dz = g.dir[2];
    L.ig = g.ignore_tri;
  }
  return L;
}
template <bool CAM>
__device__ __forceinline__ RayLanes ray_lanes(const splat_adray_t* __restrict__ rays, const splat_pinhole_t& cam,
                                              int64_t r, bool valid) {
  if constexpr (CAM) return cam_ray_lanes(cam, r, valid);
  else return load_ray_lanes(rays, r, valid);
}

__device__ __forceinline__ int64_t live_count(int64_t R, const int64_t* __restrict__ d_count) {
  if (!d_count) return R;
  int64_t c = *d_count;
  return c < R ? c : R;
}

// ---- bounding volumes -------------------------------------------------------------------------
// The per-lane slab test against a conservative box (padded member bounds). Reciprocals are
// precomputed per ray; |d| < 1e-30 is clamped so 0 * inf never appears.
// slab_entry also gives the entry distance tmin (>= 0).
__device__ __forceinline__ bool slab_entry(const float4& lo, const float4& hi, const RayLanes& L, float ix, float iy,
                                           float iz, float best, float& tmin_out) {
  float t1 = (lo.x - L.ox) * ix, t2 = (hi.x - L.ox) * ix;
  float tmin = fmaxf(0.0f, fminf(t1, t2)), tmax = fmaxf(t1, t2);
  t1 = (lo.y - L.oy) * iy;
  t2 = (hi.y - L.oy) * iy;
  tmin = fmaxf(tmin, fminf(t1, t2));
  tmax = fminf(tmax, fmaxf(t1, t2));
  t1 = (lo.z - L.oz) * iz;
  t2 = (hi.z - L.oz) * iz;
  tmin = fmaxf(tmin, fminf(t1, t2));
  tmax = fminf(tmax, fmaxf(t1, t2));
  tmin_out = tmin;
  return (tmin <= tmax) && (tmin <= best);
}
__device__ __forceinline__ bool slab_enter(const float4& lo, const float4& hi, const RayLanes& L, float ix, float iy,
                                           float iz, float best) {
  float tmin;
  return slab_entry(lo, hi, L, ix, iy, iz, best, tmin);
}

// keep (t, idx) when it is a hit lexicographically before (best, bi); selects, not branches, so
// the record loads of consecutive members can be issued ahead of their tests
__device__ __forceinline__ void take_nearer(bool h, float t, int32_t idx, float& best, int32_t& bi) {
  const bool take = h && (t < best || (t == best && idx < bi));
  best = take ? t : best;
  bi = take ? idx : bi;
}

// Up to 16 consecutive records recs[0, cnt) tested by the whole wave: one coalesced vector load
// brings them in (lane l holds quarter l % 4 of record l / 4) and writes them to the wave's own
// 1 KB of LDS, from which every lane reads each record at the same address (a broadcast: four
// ds_read_b128 per record, no bank conflict). Scalar loads return out of order, so a loop of them
// waits for each record in turn; this waits once per 16. (Broadcasting with v_readlane instead,
// 16 per record, spent a quarter of a leaf scan's VALU issue: the LDS read is 4.5% faster per
// frame.) Called by whole waves of a workgroup of at most 4 waves (every hitter launch is 256
// threads): 4 KB of LDS per workgroup, so a CU's eight hitter workgroups leave room for other
// kernels' workgroups beside them. LDS operations of one wave complete in order, so the reads see
// the wave's own writes without a barrier.
__device__ __forceinline__ void scan_records16(const TriRec* __restrict__ recs, int32_t cnt, const RayLanes& L,
                                               float& best, int32_t& bi) {
  __shared__ float4 rec_stage[4][64];
  float4* const buf = rec_stage[threadIdx.x >> 6];
  const int lane = threadIdx.x & 63;
  float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
  if ((lane >> 2) < cnt) v = reinterpret_cast<const float4*>(recs + (lane >> 2))[lane & 3];
  buf[lane] = v;
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    if (j < cnt) {
      const float4 q0 = buf[4 * j], q1 = buf[4 * j + 1], q2 = buf[4 * j + 2], q3 = buf[4 * j + 3];
      float t;
      const bool h = hit_rec(q0, q1, q2, q3, L.ox, L.oy, L.oz, L.dx, L.dy, L.dz, L.ig, t);
      take_nearer(h, t, __float_as_int(q1.w), best, bi);
    }
  }
  __builtin_amdgcn_wave_barrier();  // the next block's writes stay after these reads
}

// Counters of a STATS build, per lane: n_box -- ray-box tests on the ray's own path (the boxes
// a per-ray traversal of the same structure tests: the root, then the children of every node the
// ray entered; every box of a flat scan); n_enter -- leaves the ray entered; n_alg -- triangle
// tests of those leaves (and of the large-triangle list). Per wave (the same in every lane):
// n_exec -- triangle tests and n_box_exec -- box tests the wave executed (each for all 64 lanes).
struct HitStats {
  unsigned long long n_box = 0, n_enter = 0, n_alg = 0, n_exec = 0, n_box_exec = 0;
};

// (t, index) key of a miss: (Inf, 1)
constexpr uint64_t KEY_MISS = 0xFF80000000000001ull;

__device__ __forceinline__ float key_t(uint64_t k) {
  const uint32_t m = (uint32_t)(k >> 32);
  return __uint_as_float((m >> 31) == 0 ? ~m : (m ^ 0x80000000u));
}

// One wave, one ray per lane: the lexicographic (t, index) minimum over the scene's BV leaves.
// The wave walks the top structure together: a box no lane enters (one ballot) is skipped with
// everything under it; a leaf some lane enters has all its member records tested by every lane
// (a triangle held by two leaves ties with itself). MODE (BVMode) picks the top structure:
//   BV_FLAT   -- every leaf box in order (the reference's one-level DPBVHitter);
//   BV_GROUPS -- group boxes of gsize consecutive leaves, then their leaves;
//   BV_TREE   -- depth-first over the binary tree two levels at a time (a node's four
//                grandchildren at once), nearer first along the split axes for the wave's first
//                live ray; a wave whose rays diverge in direction may walk per lane
//                (tree_walk_lanes).
// Work split `split` of `n_split` takes groups/leaves split, split + n_split, ... (flat, groups)
// or the subtree under node n_split + split (tree; n_split a power of two). Must be called by
// whole waves.
// Scene data is read-only for the whole launch: reading it through the constant address space
// lets every wave-uniform access become a scalar load, whatever the kernel writes elsewhere.
// (The host pass only parses these functions; it has no such address space.)
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
using cptr = const __attribute__((address_space(4))) T*;
#else
template <class T>
using cptr = const T*;
#endif
template <class T>
__device__ __forceinline__ cptr<T> as_const(const T* p) {
  return (cptr<T>)(p);
}

// Per-lane variant of the tree walk, for waves whose rays diverge: every lane walks its own
// depth-first path through the subtree under `root` (nearer child first along its own
// direction), stackless -- a per-lane bit per level records whether the node on the path is its
// parent's first child, so climbing needs no memory. Node boxes and member records come as
// per-lane vector loads (cached: the tree and records are small and shared). Lanes first walk
// until they reach a leaf they enter (or finish); then the lanes holding a leaf test its members
// together; repeat until every lane is done. A lane's work is its own path, not the union of the
// wave's paths. best/bi are in-out; shared_key as in bv_traverse.
template <bool STATS>
__device__ __forceinline__ void tree_walk_lanes(const RayLanes& L, bool valid, float ix, float iy, float iz,
                                                const BVScene& S, int32_t root, HitStats& st, float& best,
                                                int32_t& bi, const uint64_t* shared_key) {
  const float4* __restrict__ nodes = S.nodes;
  const int2* __restrict__ ranges = S.ranges;
  const TriRec* __restrict__ recs = S.recs;
  const int32_t leaf0 = 1 << S.levels;
  const int32_t neg = (L.dx < 0.f ? 1 : 0) | (L.dy < 0.f ? 2 : 0) | (L.dz < 0.f ? 4 : 0);
  int32_t n = root;
  uint32_t firsts = 0;  // bit l: the node at depth l below root is its parent's first child
  int32_t depth = 0;
  bool done = !valid;
  // advance past node n (its subtree finished or skipped)
  auto next = [&]() {
    while (depth > 0 && !((firsts >> depth) & 1u)) {
      n >>= 1;
      --depth;
    }
    if (depth == 0) {
      done = true;
    } else {
      firsts &= ~(1u << depth);
      n ^= 1;
    }
  };
  while (__builtin_amdgcn_ballot_w64(!done) != 0) {
    // walk until this lane holds a leaf it enters, or is done
    int32_t leaf = -1;
    while (__builtin_amdgcn_ballot_w64(!done && leaf < 0) != 0) {
      if constexpr (STATS) st.n_box_exec += 1ull;
      if (!done && leaf < 0) {
        const float4 lo = nodes[2 * n], hi = nodes[2 * n + 1];
        const int32_t code = __float_as_int(lo.w);
        const bool enter = !(code & TREE_EMPTY) && slab_enter(lo, hi, L, ix, iy, iz, best);
        if constexpr (STATS) st.n_box += 1ull;
        if (enter && n < leaf0) {
          n = 2 * n + (((neg >> (code & TREE_AXIS)) & 1) ^ ((code & TREE_SWAP) ? 1 : 0));
          ++depth;
          firsts |= 1u << depth;
        } else if (enter) {
          leaf = n - leaf0;
        } else {
          next();
        }
      }
    }
    // the lanes holding a leaf test its members
    int2 rg = make_int2(0, 0);
    if (leaf >= 0) {
      rg = ranges[leaf];
      if (shared_key) {
        const uint64_t k = __hip_atomic_load(shared_key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (k < pack_key(best, bi)) {
          best = key_t(k);
          bi = key_index(k);
        }
      }
      if constexpr (STATS) {
        st.n_enter += 1ull;
        st.n_alg += (unsigned long long)rg.y;
      }
    }
    int32_t m_max = rg.y;
    for (int off = 32; off > 0; off >>= 1) m_max = max(m_max, __shfl_xor(m_max, off));
    if constexpr (STATS) st.n_exec += (threadIdx.x & 63) == 0 ? (unsigned long long)m_max : 0ull;
    for (int32_t m = 0; m < m_max; ++m) {
      if (m < rg.y) {
        const TriRec& T = recs[rg.x + m];
        const float4 q0 = T.q0, q1 = T.q1, q2 = T.q2, q3 = T.q3;
        float t;
        const bool h = hit_rec(q0, q1, q2, q3, L.ox, L.oy, L.oz, L.dx, L.dy, L.dz, L.ig, t);
        const int32_t idx = __float_as_int(q1.w);
        take_nearer(h, t, idx, best, bi);
      }
    }
    if (leaf >= 0) next();
  }
}

template <bool STATS, int MODE>
__device__ __forceinline__ int32_t bv_traverse(const RayLanes& L, bool valid, const BVScene& S, HitStats& st,
                                               float* best_io = nullptr, int32_t split = 0, int32_t n_split = 1,
                                               const uint64_t* shared_key = nullptr) {
  const cptr<float4> boxes = as_const(S.boxes);
  const cptr<int2> ranges = as_const(S.ranges);
  unsigned long long& n_box = st.n_box;
  unsigned long long& n_enter = st.n_enter;
  unsigned long long& n_alg = st.n_alg;
  unsigned long long& n_exec = st.n_exec;
  unsigned long long& n_box_exec = st.n_box_exec;
  auto safe_inv = [](float d) {
    float a = fabsf(d) < 1e-30f ? copysignf(1e-30f, d) : d;
    return 1.0f / a;
  };
  const float ix = safe_inv(L.dx), iy = safe_inv(L.dy), iz = safe_inv(L.dz);
  float best = INFINITY;
  int32_t bi = 1;
  // another split of the same rays may already hold a nearer hit: prune with it (it is a real
  // hit of this ray, so the lexicographic minimum is unchanged). The load is issued before a
  // leaf's members are scanned and merged after, so its latency hides behind the scan.
  auto fetch = [&]() -> uint64_t {
    return shared_key && valid ? __hip_atomic_load(shared_key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : ~0ull;
  };
  auto merge = [&](uint64_t k) {
    if (k < pack_key(best, bi)) {
      best = key_t(k);
      bi = key_index(k);
    }
  };
  // every lane tests the member records rg = (first, count) of a leaf (some lane entered it)
  auto scan_range = [&](int2 rg, bool entered) {
    if constexpr (STATS) {
      n_enter += entered ? 1ull : 0ull;
      n_alg += entered ? (unsigned long long)rg.y : 0ull;
      n_exec += (unsigned long long)rg.y;
    }
    for (int32_t m = 0; m < rg.y; m += 16)
      scan_records16(S.recs + rg.x + m, rg.y - m < 16 ? rg.y - m : 16, L, best, bi);
  };
  auto scan_members = [&](int32_t b, bool entered) { scan_range(ranges[b], entered); };

  // the records every ray tests before the structure (large triangles): in one split only, the
  // splits' keys merge into the minimum over all of them
  if (S.n_top > 0 && split == 0) {
    const int32_t nt = S.n_top;
    if constexpr (STATS) {
      n_alg += valid ? (unsigned long long)nt : 0ull;
      n_exec += (unsigned long long)nt;
    }
    for (int32_t m = 0; m < nt; m += 16) scan_records16(S.recs + m, nt - m < 16 ? nt - m : 16, L, best, bi);
  }
  auto scan_leaf = [&](int32_t b, bool in_group) {
    const float4 blo = boxes[2 * b], bhi = boxes[2 * b + 1];
    const bool enter = valid && slab_enter(blo, bhi, L, ix, iy, iz, best);
    if constexpr (STATS) {
      n_box += in_group ? 1ull : 0ull;
      n_box_exec += 1ull;
    }
    if (__builtin_amdgcn_ballot_w64(enter) == 0) return;
    scan_members(b, enter);
  };

  if constexpr (MODE == BV_GROUPS) {
    const cptr<float4> groups = as_const(S.groups);
    const int32_t n_bv = S.n_bv, gsize = S.gsize, n_groups = S.n_groups;
    for (int32_t g = split; g < n_groups; g += n_split) {
      merge(fetch());
      const float4 glo = groups[2 * g], ghi = groups[2 * g + 1];
      const bool enter = valid && slab_enter(glo, ghi, L, ix, iy, iz, best);
      if constexpr (STATS) {
        n_box += valid ? 1ull : 0ull;
        n_box_exec += 1ull;
      }
      if (__builtin_amdgcn_ballot_w64(enter) == 0) continue;
      const int32_t b1 = (g + 1) * gsize < n_bv ? (g + 1) * gsize : n_bv;
      for (int32_t b = g * gsize; b < b1; ++b) scan_leaf(b, enter);
    }
  } else if constexpr (MODE == BV_TREE) {
    const cptr<float4> nodes = as_const(S.nodes);
    const int32_t leaf0 = 1 << S.levels;
    // splits beyond the leaf count have no subtree of their own
    const int32_t n_sub = n_split < leaf0 ? n_split : leaf0;
    const uint64_t live = __builtin_amdgcn_ballot_w64(valid);
    // direction signs of the wave's first live ray order the children
    const int32_t fl = live ? (int32_t)__builtin_ctzll(live) : 0;
    const float rdx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(L.dx), fl));
    const float rdy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(L.dy), fl));
    const float rdz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(L.dz), fl));
    const int32_t neg = (rdx < 0.f ? 1 : 0) | (rdy < 0.f ? 2 : 0) | (rdz < 0.f ? 4 : 0);
    // a wave whose live rays do not all point into the leader's octant walks per lane
    const int32_t own = (L.dx < 0.f ? 1 : 0) | (L.dy < 0.f ? 2 : 0) | (L.dz < 0.f ? 4 : 0);
    const uint64_t same = __builtin_amdgcn_ballot_w64(valid && own == neg);
    const bool per_lane = S.walk == BV_WALK_LANES || (S.walk == BV_WALK_AUTO && same != live);
    if (live != 0 && split < n_sub && per_lane) {
      tree_walk_lanes<STATS>(L, valid, ix, iy, iz, S, n_sub + split, st, best, bi, shared_key);
    } else if (live != 0 && split < n_sub) {
      // The wave walks the binary tree two levels at a time. A node's four grandchildren are
      // contiguous in the heap (4n .. 4n + 3: one 128-byte scalar load) and are tested at once: a
      // parent's box is the union of its children's, so a ray entering a grandchild enters its
      // parent too and the leaves reached are those of a level-by-level walk, with half the chain
      // of dependent loads. Entered grandchildren go on a stack, farthest first for the wave's
      // first live ray along the split axes (nearer child, then its nearer half: measured faster
      // than ordering by that ray's entry distances), and are popped without a re-test. The stack
      // holds wave-uniform node indices, one per lane of a VGPR: a step pops one entry and pushes
      // at most four, over at most 12 steps (24 levels), so it never holds more than 37 of the 64
      // lanes. Leaves are scanned as soon as they are reached, nearer first, each re-tested with
      // the hits found meanwhile.
      const int32_t root = n_sub + split;
      int32_t stack = 0, sp = 0;  // lane k of `stack`: entry k
      // STATS: lane k of mlo/mhi holds the mask of the lanes that entered entry k
      uint32_t mlo = 0, mhi = 0;
      const int32_t my_lane = threadIdx.x & 63;
      auto push = [&](int32_t v, uint64_t m) {
        stack = my_lane == sp ? v : stack;
        if constexpr (STATS) {
          mlo = my_lane == sp ? (uint32_t)m : mlo;
          mhi = my_lane == sp ? (uint32_t)(m >> 32) : mhi;
        }
        ++sp;
      };
      // entry distance of this lane's ray into a node box: tmin where the slab intervals overlap,
      // NaN where they do not (or the lane is idle, or the subtree is empty), so `t <= best` is
      // the slab test, and stays exact later for a nearer best (tmin and tmax do not depend on it)
      auto lane_t = [&](const float4& lo, const float4& hi) -> float {
        if (!valid || (__float_as_int(lo.w) & TREE_EMPTY)) return __builtin_nanf("");
        float tmin;
        return slab_entry(lo, hi, L, ix, iy, iz, INFINITY, tmin) ? tmin : __builtin_nanf("");
      };
      // the child of a node with split code c that the wave's first live ray meets first (0 or 1):
      // the lower one along the split axis when the ray points up it
      auto near_half = [&](int32_t c) { return ((neg >> (c & TREE_AXIS)) & 1) ^ ((c & TREE_SWAP) ? 1 : 0); };
      // a leaf node's member range, stored with its box
      auto leaf_range = [](const float4& lo, const float4& hi) {
        return make_int2(__float_as_int(hi.w), __float_as_int(lo.w) >> TREE_COUNT_SHIFT);
      };
      auto leaf = [&](int2 rg, bool e) {
        const uint64_t k = fetch();
        scan_range(rg, e);
        merge(k);
      };
      // the leaf of range rg whose box this lane entered at distance t, unless the hits found
      // since rule it out
      auto visit = [&](int2 rg, float t) {
        const bool e = t <= best;
        if (__builtin_amdgcn_ballot_w64(e) != 0) leaf(rg, e);
      };
      {
        const float t = lane_t(nodes[2 * root], nodes[2 * root + 1]);
        if constexpr (STATS) {
          n_box += valid ? 1ull : 0ull;
          n_box_exec += 1ull;
        }
        const uint64_t m = __builtin_amdgcn_ballot_w64(t <= best);
        if (m != 0) {
          if (root >= leaf0) visit(leaf_range(nodes[2 * root], nodes[2 * root + 1]), t);
          else push(root, m);
        }
      }
      while (sp > 0) {
        const int32_t n = __builtin_amdgcn_readlane(stack, --sp);
        // STATS: did this lane's ray enter n (its own path continues below n)?
        bool in_n = false;
        if constexpr (STATS) {
          const uint64_t m = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int32_t)mhi, sp) << 32) |
                             (uint32_t)__builtin_amdgcn_readlane((int32_t)mlo, sp);
          in_n = ((m >> my_lane) & 1ull) != 0;
        }
        if (2 * n >= leaf0) {  // the children are leaves: nearer first
          const float ta = lane_t(nodes[4 * n], nodes[4 * n + 1]);
          const float tb = lane_t(nodes[4 * n + 2], nodes[4 * n + 3]);
          if constexpr (STATS) {
            n_box += in_n ? 2ull : 0ull;
            n_box_exec += 2ull;
          }
          const bool wa = __builtin_amdgcn_ballot_w64(ta <= best) != 0, wb = __builtin_amdgcn_ballot_w64(tb <= best) != 0;
          const bool b_first = wb && (!wa || near_half(__float_as_int(nodes[2 * n].w)) == 1);
          const int2 ra = leaf_range(nodes[4 * n], nodes[4 * n + 1]), rb = leaf_range(nodes[4 * n + 2], nodes[4 * n + 3]);
          visit(b_first ? rb : ra, b_first ? tb : ta);
          if (wa && wb) visit(b_first ? ra : rb, b_first ? ta : tb);
          continue;
        }
        // named values, no arrays: a select between array elements puts the array in scratch
        const int32_t g = 8 * n;
        const float t0 = lane_t(nodes[g], nodes[g + 1]), t1 = lane_t(nodes[g + 2], nodes[g + 3]),
                    t2 = lane_t(nodes[g + 4], nodes[g + 5]), t3 = lane_t(nodes[g + 6], nodes[g + 7]);
        const uint64_t m0 = __builtin_amdgcn_ballot_w64(t0 <= best), m1 = __builtin_amdgcn_ballot_w64(t1 <= best),
                       m2 = __builtin_amdgcn_ballot_w64(t2 <= best), m3 = __builtin_amdgcn_ballot_w64(t3 <= best);
        const bool w0 = m0 != 0, w1 = m1 != 0, w2 = m2 != 0, w3 = m3 != 0;
        if constexpr (STATS) {
          // a per-ray walk tests n's two children, then the two children of each it enters
          const bool e0 = in_n && lane_t(nodes[4 * n], nodes[4 * n + 1]) <= best;
          const bool e1 = in_n && lane_t(nodes[4 * n + 2], nodes[4 * n + 3]) <= best;
          n_box += (in_n ? 2ull : 0ull) + (e0 ? 2ull : 0ull) + (e1 ? 2ull : 0ull);
          n_box_exec += 4ull;
        }
        const uint32_t wmask = (w0 ? 1u : 0u) | (w1 ? 2u : 0u) | (w2 ? 4u : 0u) | (w3 ? 8u : 0u);
        if (wmask == 0) continue;
        // visiting rank of grandchild j = 2c + h (child c, half h): nearer child first, then its
        // nearer half, by the split codes of n and of its two children (loaded with the boxes)
        const int32_t cn = __float_as_int(nodes[2 * n].w), ca = __float_as_int(nodes[4 * n].w),
                      cb = __float_as_int(nodes[4 * n + 2].w);
        const int32_t nn = near_half(cn), na = near_half(ca), nb = near_half(cb);
        auto rank = [&](int32_t j) {
          const int32_t c = j >> 1, h = j & 1;
          return (c == nn ? 0 : 2) + (h == (c ? nb : na) ? 0 : 1);
        };
        int32_t k0 = w0 ? rank(0) : INT32_MAX, k1 = w1 ? rank(1) : INT32_MAX,
                k2 = w2 ? rank(2) : INT32_MAX, k3 = w3 ? rank(3) : INT32_MAX;
        int32_t i0 = 0, i1 = 1, i2 = 2, i3 = 3;
        // sort (key, index) ascending: a network of five compare-exchanges
        auto cx = [](int32_t& ka, int32_t& ia, int32_t& kb, int32_t& ib) {
          const bool sw = kb < ka;
          const int32_t k = ka, i = ia;
          ka = sw ? kb : ka; ia = sw ? ib : ia;
          kb = sw ? k : kb; ib = sw ? i : ib;
        };
        cx(k0, i0, k1, i1); cx(k2, i2, k3, i3); cx(k0, i0, k2, i2); cx(k1, i1, k3, i3); cx(k1, i1, k2, i2);
        if (4 * n >= leaf0) {  // the grandchildren are leaves: scan them nearest first
          auto pick = [](int32_t j, float a0, float a1, float a2, float a3) {
            return j == 0 ? a0 : j == 1 ? a1 : j == 2 ? a2 : a3;
          };
          const int2 r0 = leaf_range(nodes[g], nodes[g + 1]), r1 = leaf_range(nodes[g + 2], nodes[g + 3]),
                     r2 = leaf_range(nodes[g + 4], nodes[g + 5]), r3 = leaf_range(nodes[g + 6], nodes[g + 7]);
          auto pick_r = [&](int32_t j) { return j == 0 ? r0 : j == 1 ? r1 : j == 2 ? r2 : r3; };
          if ((wmask >> i0) & 1u) visit(pick_r(i0), pick(i0, t0, t1, t2, t3));
          if ((wmask >> i1) & 1u) visit(pick_r(i1), pick(i1, t0, t1, t2, t3));
          if ((wmask >> i2) & 1u) visit(pick_r(i2), pick(i2, t0, t1, t2, t3));
          if ((wmask >> i3) & 1u) visit(pick_r(i3), pick(i3, t0, t1, t2, t3));
        } else {
          auto mask_of = [&](int32_t j) { return j == 0 ? m0 : j == 1 ? m1 : j == 2 ? m2 : m3; };
          if ((wmask >> i3) & 1u) push(4 * n + i3, mask_of(i3));
          if ((wmask >> i2) & 1u) push(4 * n + i2, mask_of(i2));
          if ((wmask >> i1) & 1u) push(4 * n + i1, mask_of(i1));
          if ((wmask >> i0) & 1u) push(4 * n + i0, mask_of(i0));
        }
      }
    }
  } else {
    // leaves split, split + n_split, ...: eight box loads and tests at a time (independent
    // loads, no chain), then the members of the leaves some lane entered
    const int32_t n_bv = S.n_bv;
    for (int32_t b0 = split; b0 < n_bv; b0 += 8 * n_split) {
      const uint64_t k = fetch();
      uint32_t wave_mask = 0, lane_mask = 0;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int32_t b = b0 + j * n_split;
        if (b < n_bv) {
          const float4 lo = boxes[2 * b], hi = boxes[2 * b + 1];
          const bool e = valid && slab_enter(lo, hi, L, ix, iy, iz, best);
          if constexpr (STATS) {
            n_box += valid ? 1ull : 0ull;
            n_box_exec += 1ull;
          }
          if (__builtin_amdgcn_ballot_w64(e) != 0) wave_mask |= 1u << j;
          lane_mask |= (e ? 1u : 0u) << j;
        }
      }
      while (wave_mask != 0) {
        const int j = __builtin_ctz(wave_mask);
        wave_mask &= wave_mask - 1;
        scan_members(b0 + j * n_split, ((lane_mask >> j) & 1u) != 0);
      }
      merge(k);
    }
  }
  if (best_io) *best_io = best;
  return bi;
}

// Split factor for `groups64` live 64-ray groups over a scene with `units` top-level boxes (or
// tree leaves): enough (group, split) items to fill the chip, never more splits than units.
// (12288: 1.2% faster per 1 MP frame than 16384, mostly from the smaller grids of the late
// launches; 24576 splits the first bounce's coherent rays, 8192 too few for bounces 2-3.)
constexpr int64_t SPLIT_TARGET_ITEMS = 12288;
constexpr int32_t MAX_SPLIT = 32;
inline __host__ __device__ int32_t split_factor(int64_t groups64, int32_t units,
                                                int64_t target = SPLIT_TARGET_ITEMS) {
  int32_t S = 1;
  while (S < MAX_SPLIT && S < units && groups64 * S < target) S *= 2;
  return S;
}
// A launch with at most TAIL_GROUPS live 64-ray groups (the last bounces: a few thousand rays
// trapped in glass, pointing every way) scans the tree's leaf boxes flat, split over many items:
// the tree walk of such a wave is one long chain of dependent node loads, the flat scan is
// independent loads. Split: up to TAIL_MAX_SPLIT items per group, at least 8 leaves each.
constexpr int64_t TAIL_GROUPS = 64;
constexpr int32_t TAIL_MAX_SPLIT = 128;
constexpr int64_t TAIL_ITEMS = 8192;  // at most this many flat-scan items per launch
// split of `groups64` tail groups (in all) over scenes of `leaves` leaves: S * 8 <= leaves,
// groups64 * S <= TAIL_ITEMS
inline __host__ __device__ int32_t tail_split_factor(int64_t groups64, int32_t leaves) {
  int32_t S = 1;
  while (S < TAIL_MAX_SPLIT && S * 16 <= leaves && groups64 * (S * 2) <= TAIL_ITEMS) S *= 2;
  return S;
}
// Work items (waves) the hitter needs for at most `groups64` live groups, of which at most
// `tail_groups64` can belong to flat-scan tails (a frame scans flat only with at most TAIL_GROUPS
// live groups): the walk split only grows while groups * S < SPLIT_TARGET_ITEMS, so the walks
// never take more than max(groups, 2 SPLIT_TARGET_ITEMS); the tail split only grows while
// tail groups * S <= TAIL_ITEMS and is 1 above that, so the tails take at most
// max(tail groups, TAIL_ITEMS). A step of many frames can hold both kinds.
inline __host__ __device__ int64_t max_hit_items(int64_t groups64, int64_t tail_groups64) {
  return (groups64 > 2 * SPLIT_TARGET_ITEMS ? groups64 : 2 * SPLIT_TARGET_ITEMS) +
         (tail_groups64 > TAIL_ITEMS ? tail_groups64 : TAIL_ITEMS);
}
// The same for a launch over one frame, whose items are all tree walks or all a flat-scan tail:
// the larger bound, not the sum (every wave past the items still has to be dispatched).
inline __host__ __device__ int64_t max_hit_items_one(int64_t groups64, int64_t target = SPLIT_TARGET_ITEMS) {
  const int64_t walk = groups64 > 2 * target ? groups64 : 2 * target;
  return walk > TAIL_ITEMS ? walk : TAIL_ITEMS;
}

// One (64-ray group, split) work item: rays [base, base + 64) of `rays` (CAM: of the pinhole) with
// A live, scanning split `split` of `n_split`; the result is merged into keys[r] with a 64-bit
// atomic minimum of the packed (t, index) key. Called by whole waves.
template <bool STATS, int MODE, bool CAM = false>
__device__ __forceinline__ void bv_item(const splat_adray_t* __restrict__ rays, int64_t A, int64_t base, int32_t split,
                                        int32_t n_split, const BVScene& scene, uint64_t* __restrict__ keys,
                                        HitStats& st, const splat_pinhole_t* cam = nullptr) {
  const int64_t r = base + (threadIdx.x & 63);
  const bool valid = r < A;
  const RayLanes L = CAM ? cam_ray_lanes(*cam, r, valid) : load_ray_lanes(rays, r, valid);
  float best;
  const int32_t bi = bv_traverse<STATS, MODE>(L, valid, scene, st, &best, split, n_split,
                                                n_split > 1 && valid ? keys + r : nullptr);
  if (valid) {
    const uint64_t k = pack_key(best, bi);
    if (n_split == 1) keys[r] = k;
    else if (k < KEY_MISS) atomicMin(reinterpret_cast<unsigned long long*>(keys + r), (unsigned long long)k);
  }
}

}  // namespace spl
