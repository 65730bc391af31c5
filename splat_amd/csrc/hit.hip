// hit.hip -- ray/triangle hitters for gfx950.
//
// Design (replaces EH1-EH6 and DPBVHitter, src/hitters.jl:34-648):
//  * one lane = one ray, kept in VGPRs for the whole scan; a wave is 64 rays;
//  * the triangle stream is wave-uniform, so each 64-byte prepared record arrives through the
//    scalar unit (s_load_dwordx16) into SGPRs and feeds all 64 lanes as an instruction operand:
//    no LDS traffic and no cross-lane shuffles per test (the reference rotates rays through
//    32-lane shuffles, src/hitters.jl:111-130);
//  * brute force splits the triangle list across W waves of a workgroup that share the same 64
//    rays, then reduces the W packed (t, index) keys per ray through LDS -- deterministic, no
//    global atomics, ties to the lowest index;
//  * the BV hitter walks the conservative member boxes per wave; a box whose slab test fails for
//    all 64 lanes (one ballot) is skipped, otherwise its member records are scanned.
#include <cstring>

#include "hitcore.hpp"

namespace spl {

// ---- preparation --------------------------------------------------------------------------
__global__ void k_prep_ntris(const splat_ntri_t* __restrict__ tris, int32_t N, TriRec* __restrict__ out) {
  int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const splat_ntri_t& t = tris[j];
  out[j] = tri_prep(ld3(t.v[0]), ld3(t.v[1]), ld3(t.v[2]), ld3(t.v[3]), t.idx);
}

__global__ void k_prep_ftris(const splat_ftri_t* __restrict__ tris, int32_t N, TriRec* __restrict__ out) {
  int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const splat_ftri_t& t = tris[j];
  out[j] = tri_prep(ld3(t.v[0]), ld3(t.v[1]), ld3(t.v[2]), ld3(t.v[3]), j + 1);
}

// members: 1-based triangle indices, concatenated per BV
__global__ void k_prep_members(const splat_ftri_t* __restrict__ tris, const int32_t* __restrict__ members,
                               int64_t total, TriRec* __restrict__ out) {
  int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= total) return;
  int32_t idx = members[m];
  const splat_ftri_t& t = tris[idx - 1];
  out[m] = tri_prep(ld3(t.v[0]), ld3(t.v[1]), ld3(t.v[2]), ld3(t.v[3]), idx);
}

// one workgroup per BV: member vertex bounds, padded (relative 1e-3 of the extent, 1e-5 of the
// magnitude, 1e-6 absolute) so that rounding in the slab test can never cull a hit.
__global__ __launch_bounds__(256) void k_bv_boxes(const splat_ftri_t* __restrict__ tris, const int32_t* __restrict__ members,
                                                  const int2* __restrict__ ranges, float4* __restrict__ boxes) {
  __shared__ float lo_s[3][256], hi_s[3][256];
  const int b = blockIdx.x;
  const int2 rg = ranges[b];
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int m = threadIdx.x; m < rg.y; m += blockDim.x) {
    const splat_ftri_t& t = tris[members[rg.x + m] - 1];
    for (int v = 1; v <= 3; ++v)
      for (int k = 0; k < 3; ++k) {
        lo[k] = fminf(lo[k], t.v[v][k]);
        hi[k] = fmaxf(hi[k], t.v[v][k]);
      }
  }
  for (int k = 0; k < 3; ++k) { lo_s[k][threadIdx.x] = lo[k]; hi_s[k][threadIdx.x] = hi[k]; }
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int k = 0; k < 3; ++k) {
        lo_s[k][threadIdx.x] = fminf(lo_s[k][threadIdx.x], lo_s[k][threadIdx.x + s]);
        hi_s[k][threadIdx.x] = fmaxf(hi_s[k][threadIdx.x], hi_s[k][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    float ext = 0.f, mag = 0.f;
    for (int k = 0; k < 3; ++k) {
      ext = fmaxf(ext, hi_s[k][0] - lo_s[k][0]);
      mag = fmaxf(mag, fmaxf(fabsf(lo_s[k][0]), fabsf(hi_s[k][0])));
    }
    float pad = 1e-3f * ext + 1e-5f * mag + 1e-6f;
    if (rg.y == 0) {  // empty BV: a box no ray can enter
      boxes[2 * b] = make_float4(INFINITY, INFINITY, INFINITY, 0.f);
      boxes[2 * b + 1] = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.f);
    } else {
      boxes[2 * b] = make_float4(lo_s[0][0] - pad, lo_s[1][0] - pad, lo_s[2][0] - pad, 0.f);
      boxes[2 * b + 1] = make_float4(hi_s[0][0] + pad, hi_s[1][0] + pad, hi_s[2][0] + pad, 0.f);
    }
  }
}

// ---- brute force ------------------------------------------------------------------------------
// Workgroup = W waves sharing 64 rays; wave w scans record slots 1+w, 1+w+W, ... (slot 0 is the
// degenerate sentinel). Within a wave indices increase, so a strict t < best keeps the lowest
// index on ties; across waves the packed-key minimum does.
template <int W, bool STATS>
__global__ __launch_bounds__(64 * W) void k_hit_bf(const splat_adray_t* __restrict__ rays, int64_t R,
                                                   const int64_t* __restrict__ d_count,
                                                   const TriRec* __restrict__ recs, int32_t N,
                                                   int32_t* __restrict__ hit_idx, uint64_t* __restrict__ keys,
                                                   unsigned long long* stats, const splat_pinhole_t cam,
                                                   int32_t from_cam) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t Rl = live_count(R, d_count);
  const int64_t base = (int64_t)blockIdx.x * 64;
  if (base >= Rl) return;  // uniform over the workgroup
  const int64_t r = base + lane;
  const bool valid = r < Rl;
  RayLanes L = from_cam ? cam_ray_lanes(cam, r, valid) : load_ray_lanes(rays, r, valid);
  float best = INFINITY;
  int32_t bi = 1;
#pragma unroll 4
  for (int32_t j = 1 + wave; j < N; j += W) {
    const TriRec& T = recs[j];
    float t;
    bool h = hit_rec(T.q0, T.q1, T.q2, T.q3, L.ox, L.oy, L.oz, L.dx, L.dy, L.dz, L.ig, t);
    if (h && t < best) {
      best = t;
      bi = __float_as_int(T.q1.w);
    }
  }
  if constexpr (STATS) {
    const unsigned long long nv = (unsigned long long)__builtin_popcountll(__builtin_amdgcn_ballot_w64(valid));
    if (lane == 0) {
      const unsigned long long scanned = N > 1 + wave ? (unsigned long long)((N - 2 - wave) / W + 1) : 0ull;
      atomicAdd(stats + 2, nv * scanned);
      atomicAdd(stats + 3, 64ull * scanned);
    }
  }
  if constexpr (W == 1) {
    if (valid) {
      if (keys) keys[r] = pack_key(best, bi);
      else hit_idx[r] = bi;
    }
  } else {
    __shared__ uint64_t part[W][64];
    part[wave][lane] = pack_key(best, bi);
    __syncthreads();
    if (wave == 0) {
      uint64_t k = part[0][lane];
#pragma unroll
      for (int w = 1; w < W; ++w) {
        uint64_t o = part[w][lane];
        k = o < k ? o : k;
      }
      if (valid) {
        if (keys) keys[r] = k;
        else hit_idx[r] = key_index(k);
      }
    }
  }
}

// BV hitter over (64-ray group, split) work items, one wave each. The split factor follows the
// live count on the device, so bounces with few, incoherent survivors still spread over the whole
// chip; the splits of a group merge their (t, index) keys with a 64-bit atomic minimum. The grid
// covers the most items any live count can need (max_hit_items_one); waves past the count leave.
template <bool STATS, int MODE, bool STRIDE>
__global__ __launch_bounds__(256) void k_hit_bv(const splat_adray_t* __restrict__ rays, int64_t R,
                                                const int64_t* __restrict__ d_count, BVScene scene,
                                                uint64_t* __restrict__ keys, unsigned long long* stats,
                                                const splat_pinhole_t cam, int32_t from_cam) {
  const int64_t A = live_count(R, d_count);
  const int64_t groups64 = (A + 63) / 64;
  const int32_t units = MODE == BV_TREE ? (1 << scene.levels) : MODE == BV_GROUPS ? scene.n_groups : scene.n_bv;
  const bool tail = MODE == BV_TREE && groups64 <= TAIL_GROUPS && !scene.no_tail;  // few rays: flat over the leaves
  const int32_t S = tail ? tail_split_factor(groups64, scene.n_bv) : split_factor(groups64, units, scene.split_target);
  const int64_t items = groups64 * S;
  HitStats st;
  const int64_t first = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if constexpr (STRIDE) {
    // a capped grid (HIT_GRID_CAP) after the first bounce of a large frame: the items stride over it
    // (no camera rays here: the first bounce's grid holds all its items)
    for (int64_t i = first; i < items; i += (int64_t)gridDim.x * 4) {  // whole waves
      const int64_t g = i / S;
      const int32_t sp = (int32_t)(i - g * S);
      if (tail) bv_item<STATS, BV_FLAT, false>(rays, A, g * 64, sp, S, scene, keys, st);
      else bv_item<STATS, MODE, false>(rays, A, g * 64, sp, S, scene, keys, st);
    }
  } else {
    if (first >= items) return;  // whole waves
    const int64_t g = first / S;
    const int32_t sp = (int32_t)(first - g * S);
    if (tail) {
      if (from_cam) bv_item<STATS, BV_FLAT, true>(rays, A, g * 64, sp, S, scene, keys, st, &cam);
      else bv_item<STATS, BV_FLAT, false>(rays, A, g * 64, sp, S, scene, keys, st);
    } else {
      if (from_cam) bv_item<STATS, MODE, true>(rays, A, g * 64, sp, S, scene, keys, st, &cam);
      else bv_item<STATS, MODE, false>(rays, A, g * 64, sp, S, scene, keys, st);
    }
  }
  if constexpr (STATS) {
    atomicAdd(stats + 0, st.n_box);
    atomicAdd(stats + 1, st.n_enter);
    atomicAdd(stats + 2, st.n_alg);
    if ((threadIdx.x & 63) == 0) {
      atomicAdd(stats + 3, 64ull * st.n_exec);
      atomicAdd(stats + 4, 64ull * st.n_box_exec);
    }
  }
}

// keys of the live rays <- (Inf, 1)
__global__ void k_key_init(uint64_t* __restrict__ keys, int64_t R, const int64_t* __restrict__ d_count) {
  const int64_t A = live_count(R, d_count);
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < A; r += (int64_t)gridDim.x * blockDim.x)
    keys[r] = KEY_MISS;
}

// hit_idx[r] <- index part of keys[r] for the live rays (retreive_arg, src/atomic_argmin.jl:17-20)
__global__ void k_key_final(const uint64_t* __restrict__ keys, int64_t R, const int64_t* __restrict__ d_count,
                            int32_t* __restrict__ hit_idx) {
  const int64_t A = live_count(R, d_count);
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < A; r += (int64_t)gridDim.x * blockDim.x)
    hit_idx[r] = key_index(keys[r]);
}

// group boxes: union of each group's (already padded) leaf boxes
__global__ void k_group_boxes(const float4* __restrict__ boxes, int32_t n_bv, int32_t gsize, float4* __restrict__ groups) {
  const int32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t n_groups = (n_bv + gsize - 1) / gsize;
  if (g >= n_groups) return;
  float4 lo = make_float4(INFINITY, INFINITY, INFINITY, 0.f), hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.f);
  const int32_t b1 = (g + 1) * gsize < n_bv ? (g + 1) * gsize : n_bv;
  for (int32_t b = g * gsize; b < b1; ++b) {
    const float4 l = boxes[2 * b], h = boxes[2 * b + 1];
    lo.x = fminf(lo.x, l.x); lo.y = fminf(lo.y, l.y); lo.z = fminf(lo.z, l.z);
    hi.x = fmaxf(hi.x, h.x); hi.y = fmaxf(hi.y, h.y); hi.z = fmaxf(hi.z, h.z);
  }
  groups[2 * g] = lo;
  groups[2 * g + 1] = hi;
}

// tree leaves: node 2^L + b <- box of BV b (empty when b is padding or has no members), with its
// member range: count in the code's upper bits, first record in hi.w (a walk that reaches the
// leaf has its range with the box, no dependent load)
__global__ void k_tree_leaves(const float4* __restrict__ boxes, const int2* __restrict__ ranges, int32_t n_bv,
                              int32_t levels, float4* __restrict__ nodes) {
  const int32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t n_leaves = 1 << levels;
  if (b >= n_leaves) return;
  const int32_t n = n_leaves + b;
  float4 lo = make_float4(INFINITY, INFINITY, INFINITY, 0.f), hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.f);
  int32_t code = TREE_EMPTY, first = 0;
  if (b < n_bv && ranges[b].y > 0) {
    lo = boxes[2 * b];
    hi = boxes[2 * b + 1];
    code = ranges[b].y << TREE_COUNT_SHIFT;
    first = ranges[b].x;
  }
  lo.w = __int_as_float(code);
  hi.w = __int_as_float(first);
  nodes[2 * n] = lo;
  nodes[2 * n + 1] = hi;
}

// one tree level: node n <- union of its non-empty children, with the split code
__global__ void k_tree_level(int32_t level, float4* __restrict__ nodes) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (1 << level)) return;
  const int32_t n = (1 << level) + i, a = 2 * n, b = 2 * n + 1;
  const float4 alo = nodes[2 * a], ahi = nodes[2 * a + 1], blo = nodes[2 * b], bhi = nodes[2 * b + 1];
  const bool ea = (__float_as_int(alo.w) & TREE_EMPTY) != 0, eb = (__float_as_int(blo.w) & TREE_EMPTY) != 0;
  float4 lo, hi;
  int32_t code;
  if (ea && eb) {
    lo = alo; hi = ahi; code = TREE_EMPTY;
  } else if (ea || eb) {
    lo = ea ? blo : alo; hi = ea ? bhi : ahi; code = 0;
  } else {
    lo = make_float4(fminf(alo.x, blo.x), fminf(alo.y, blo.y), fminf(alo.z, blo.z), 0.f);
    hi = make_float4(fmaxf(ahi.x, bhi.x), fmaxf(ahi.y, bhi.y), fmaxf(ahi.z, bhi.z), 0.f);
    // centres (doubled) of the children along each axis
    const float dx = (blo.x + bhi.x) - (alo.x + ahi.x), dy = (blo.y + bhi.y) - (alo.y + ahi.y),
                dz = (blo.z + bhi.z) - (alo.z + ahi.z);
    int32_t axis = 0;
    float d = dx;
    if (fabsf(dy) > fabsf(d)) { axis = 1; d = dy; }
    if (fabsf(dz) > fabsf(d)) { axis = 2; d = dz; }
    code = axis | (d < 0.f ? TREE_SWAP : 0);
  }
  lo.w = __int_as_float(code);
  hi.w = 0.f;
  nodes[2 * n] = lo;
  nodes[2 * n + 1] = hi;
}

// ---- launchers -------------------------------------------------------------------------------
void launch_prep_ntris(const splat_ntri_t* tris, int32_t N, TriRec* out, hipStream_t s) {
  hipLaunchKernelGGL(k_prep_ntris, dim3((N + 255) / 256), dim3(256), 0, s, tris, N, out);
}
void launch_prep_ftris(const splat_ftri_t* tris, int32_t N, TriRec* out, hipStream_t s) {
  hipLaunchKernelGGL(k_prep_ftris, dim3((N + 255) / 256), dim3(256), 0, s, tris, N, out);
}
void launch_prep_members(const splat_ftri_t* tris, const int32_t* members, int64_t total, TriRec* out, hipStream_t s) {
  if (total <= 0) return;
  hipLaunchKernelGGL(k_prep_members, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, tris, members, total, out);
}
void launch_bv_boxes(const splat_ftri_t* tris, const int32_t* members, const int2* ranges, int32_t n_bv, float4* boxes,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_bv_boxes, dim3(n_bv), dim3(256), 0, s, tris, members, ranges, boxes);
}

int choose_bf_waves(int64_t R, int32_t N) {
  // Measured on MI355X over rays 256^2..1024^2 x triangles 256..65536: eight waves per group of 64
  // rays beat one, two and four at every size (55 vs 47 TFLOP/s at 1024^2 rays, 55 vs 38 at
  // 768^2, where one wave per group leaves a 1.125-round tail on the 8192 wave slots), sixteen
  // win only when the grid would otherwise hold fewer than two rounds of waves. Each wave keeps
  // at least 32 triangles.
  const int64_t groups = (R + 63) / 64;
  int w = groups * 8 < 2 * 8192 && N >= 1024 ? 16 : 8;
  while (w > 1 && N / w < 32) w /= 2;
  return w;
}

template <bool ST>
static void launch_hit_bf_t(const splat_adray_t* rays, int64_t R, const int64_t* d_count, const TriRec* recs, int32_t N,
                            int32_t* hit_idx, uint64_t* keys, int waves, unsigned long long* st,
                            const splat_pinhole_t& c, int32_t fc, hipStream_t s) {
  dim3 grid((unsigned)((R + 63) / 64));
#define SPL_HIT_BF(W) \
  hipLaunchKernelGGL((k_hit_bf<W, ST>), grid, dim3(64 * W), 0, s, rays, R, d_count, recs, N, hit_idx, keys, st, c, fc)
  switch (waves) {
    case 1: SPL_HIT_BF(1); break;
    case 2: SPL_HIT_BF(2); break;
    case 4: SPL_HIT_BF(4); break;
    case 8: SPL_HIT_BF(8); break;
    default: SPL_HIT_BF(16); break;
  }
#undef SPL_HIT_BF
}

void launch_hit_bf(const splat_adray_t* rays, int64_t R, const int64_t* d_count, const TriRec* recs, int32_t N,
                   int32_t* hit_idx, uint64_t* keys, int waves, uint64_t* stats, hipStream_t s,
                   const splat_pinhole_t* cam) {
  if (R <= 0) return;
  auto* st = reinterpret_cast<unsigned long long*>(stats);
  splat_pinhole_t c;
  if (cam) c = *cam;
  else std::memset(&c, 0, sizeof(c));
  if (st) launch_hit_bf_t<true>(rays, R, d_count, recs, N, hit_idx, keys, waves, st, c, cam ? 1 : 0, s);
  else launch_hit_bf_t<false>(rays, R, d_count, recs, N, hit_idx, keys, waves, st, c, cam ? 1 : 0, s);
}

void launch_tree_build(const float4* boxes, const int2* ranges, int32_t n_bv, int32_t levels, float4* nodes,
                       hipStream_t s) {
  const int32_t n_leaves = 1 << levels;
  hipLaunchKernelGGL(k_tree_leaves, dim3((n_leaves + 255) / 256), dim3(256), 0, s, boxes, ranges, n_bv, levels, nodes);
  for (int32_t l = levels - 1; l >= 0; --l)
    hipLaunchKernelGGL(k_tree_level, dim3(((1 << l) + 255) / 256), dim3(256), 0, s, l, nodes);
}

void launch_group_boxes(const float4* boxes, int32_t n_bv, int32_t gsize, float4* groups, hipStream_t s) {
  const int32_t n_groups = (n_bv + gsize - 1) / gsize;
  hipLaunchKernelGGL(k_group_boxes, dim3((n_groups + 63) / 64), dim3(64), 0, s, boxes, n_bv, gsize, groups);
}

static unsigned stream_blocks(int64_t n) {
  int64_t g = (n + 255) / 256;
  return (unsigned)(g < 4096 ? (g < 1 ? 1 : g) : 4096);
}

void launch_key_init(uint64_t* keys, int64_t R, const int64_t* d_count, hipStream_t s) {
  if (R <= 0) return;
  hipLaunchKernelGGL(k_key_init, dim3(stream_blocks(R)), dim3(256), 0, s, keys, R, d_count);
}
void launch_key_final(const uint64_t* keys, int64_t R, const int64_t* d_count, int32_t* hit_idx, hipStream_t s) {
  if (R <= 0) return;
  hipLaunchKernelGGL(k_key_final, dim3(stream_blocks(R)), dim3(256), 0, s, keys, R, d_count, hit_idx);
}

// keys[] must have been initialised (launch_key_init) on the same stream
void launch_hit_bv(const splat_adray_t* rays, int64_t R, const int64_t* d_count, const BVScene& scene, uint64_t* keys,
                   uint64_t* stats, hipStream_t s, const splat_pinhole_t* cam) {
  if (R <= 0) return;
  auto* st = reinterpret_cast<unsigned long long*>(stats);
  // a workgroup (4 waves) per item any live count can need; past the first bounce (whose rays
  // from the pinhole are all live) at most HIT_GRID_CAP, the items striding over it, so that a
  // launch over few live rays of a large frame dispatches no more (a separate kernel: the loop
  // costs the one-item kernel 2% of a 1 MP frame)
  int64_t wgs = (max_hit_items_one((R + 63) / 64, scene.split_target) + 3) / 4;
  if (cam) {  // every ray live: exactly the first bounce's items (the kernel's own split choice)
    const int64_t groups64 = (R + 63) / 64;
    const int mode = bv_mode(scene);
    const int32_t units = mode == BV_TREE ? (1 << scene.levels) : mode == BV_GROUPS ? scene.n_groups : scene.n_bv;
    const bool tail = mode == BV_TREE && groups64 <= TAIL_GROUPS && !scene.no_tail;
    const int32_t S = tail ? tail_split_factor(groups64, scene.n_bv) : split_factor(groups64, units, scene.split_target);
    wgs = (groups64 * S + 3) / 4;
  }
  const bool stride = !cam && !st && wgs > HIT_GRID_CAP;
  dim3 grid((unsigned)(stride ? HIT_GRID_CAP : wgs)), block(256);
  splat_pinhole_t c;
  if (cam) c = *cam;
  else std::memset(&c, 0, sizeof(c));
  const int32_t fc = cam ? 1 : 0;
#define SPL_HIT_BV(ST, M, SD) \
  hipLaunchKernelGGL((k_hit_bv<ST, M, SD>), grid, block, 0, s, rays, R, d_count, scene, keys, st, c, fc)
#define SPL_HIT_BV_MODE(M) \
  if (st) SPL_HIT_BV(true, M, false); else if (stride) SPL_HIT_BV(false, M, true); else SPL_HIT_BV(false, M, false)
  switch (bv_mode(scene)) {
    case BV_TREE: SPL_HIT_BV_MODE(BV_TREE); break;
    case BV_GROUPS: SPL_HIT_BV_MODE(BV_GROUPS); break;
    default: SPL_HIT_BV_MODE(BV_FLAT); break;
  }
#undef SPL_HIT_BV_MODE
#undef SPL_HIT_BV
}

}  // namespace spl
