// trace.hip -- dual-number evolve, final evolution + light-map splat, pinhole ray source, and the
// fused count / bounce passes of the bounce loop.
#include <cstring>

#include "bounce.hpp"

namespace spl {

// ---- fused bounce loop passes (bounce.hpp) ----------------------------------------------------
// from_cam: the live rays are the pinhole's rays 0..R-1, not yet generated (the first bounce)
// Count pass: ACTIVE rays per live tile after this bounce (a separate one-workgroup scan follows:
// a last-workgroup scan here would need one same-address atomic per tile, which serialises)
__global__ __launch_bounds__(TILE) void k_bounce_count(const splat_adray_t* __restrict__ rays, int64_t R,
                                                       const int64_t* __restrict__ d_count,
                                                       const uint64_t* __restrict__ keys, BounceParams P,
                                                       const splat_pinhole_t cam, int32_t from_cam,
                                                       int32_t* __restrict__ tile_counts) {
  const int64_t A = live_count_t(R, d_count);
  // one tile per wave; a capped grid strides over the live tiles (the work per tile is even: its keys)
  const int64_t wave = (int64_t)blockIdx.x * (TILE / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (int64_t tile = wave; tile * TILE < A; tile += (int64_t)gridDim.x * (TILE / 64)) {  // uniform over the wave
    const int64_t base = tile * TILE;
    const int nvalid = (int)(A - base < TILE ? A - base : TILE);
    const int c = from_cam ? bounce_count<true>(rays + base, keys + base, nvalid, P, true, cam, base)
                           : bounce_count<false>(rays + base, keys + base, nvalid, P, true, cam, base);
    if ((threadIdx.x & 63) == 0) tile_counts[tile] = c;
  }
}

// Start of a frame's loop: light map (when given) cleared, live count R, the first bounce's miss
// keys (when given).
__global__ void k_frame_start(float* __restrict__ tex, int64_t tex_n, int64_t* __restrict__ count0, int64_t R,
                              uint64_t* __restrict__ keys) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tex)
    for (int64_t i = g; i < tex_n; i += stride) tex[i] = 0.0f;
  if (keys)
    for (int64_t i = g; i < R; i += stride) keys[i] = KEY_MISS;
  if (g == 0) *count0 = R;
}

// End of the loop: the live rays copied to their final slots (8-byte words, coalesced) and, when
// keys are given, their realigning hit indices taken from them.
__global__ void k_frame_end(const uint2* __restrict__ src, uint2* __restrict__ dst, int64_t R,
                            const int64_t* __restrict__ d_count, const uint64_t* __restrict__ keys,
                            int32_t* __restrict__ hit_idx) {
  const int64_t A = live_count_t(R, d_count);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = A * (int64_t)(sizeof(splat_adray_t) / 8);
  for (int64_t i = g; i < n; i += stride) dst[i] = src[i];
  if (keys)
    for (int64_t i = g; i < A; i += stride) hit_idx[i] = key_index(keys[i]);
}

__global__ __launch_bounds__(TILE) void k_bounce(const splat_adray_t* __restrict__ rays, int64_t R,
                                                 const int64_t* __restrict__ d_count,
                                                 const uint64_t* __restrict__ keys, uint64_t* __restrict__ keys_next,
                                                 BounceParams P, const splat_lightmap_params_t lm,
                                                 const splat_pinhole_t cam, int32_t from_cam,
                                                 const int32_t* __restrict__ tile_sums, int32_t local_scan,
                                                 int64_t* __restrict__ d_total,
                                                 splat_adray_t* __restrict__ out_active,
                                                 splat_adray_t* __restrict__ out_rest, int32_t* __restrict__ hit_rest) {
  const int64_t A = live_count_t(R, d_count);
  const int64_t base = (int64_t)blockIdx.x * TILE;
  if (base >= A) {
    if (local_scan && blockIdx.x == 0 && threadIdx.x == 0) *d_total = 0;  // no live ray: none stays ACTIVE
    return;
  }
  const int nvalid = (int)(A - base < TILE ? A - base : TILE);
  int64_t off_a, total;
  if (local_scan) {
    // tile_sums are the count pass's tile counts (at most SCAN_LOCAL_TILES): the workgroup sums
    // those before its tile and all of them itself, in place of a scan launch; workgroup 0
    // publishes the total (the next bounce's live count)
    const int32_t n = (int32_t)((A + TILE - 1) / TILE), tile = (int32_t)blockIdx.x;
    int32_t before = 0, all = 0;
    for (int32_t i = threadIdx.x; i < n; i += TILE) {
      const int32_t c = tile_sums[i];
      all += c;
      before += i < tile ? c : 0;
    }
    for (int off = 32; off > 0; off >>= 1) {
      before += __shfl_xor(before, off);
      all += __shfl_xor(all, off);
    }
    __shared__ int32_t red[2][TILE / 64];
    if ((threadIdx.x & 63) == 0) {
      red[0][threadIdx.x >> 6] = before;
      red[1][threadIdx.x >> 6] = all;
    }
    __syncthreads();
    before = 0;
    all = 0;
#pragma unroll
    for (int w = 0; w < TILE / 64; ++w) {
      before += red[0][w];
      all += red[1][w];
    }
    off_a = before;
    total = all;
    if (tile == 0 && threadIdx.x == 0) *d_total = total;
  } else {  // tile_sums are the scan's tile offsets, d_total its total
    off_a = tile_sums[blockIdx.x];
    total = *d_total;
  }
  const int64_t off_r = total + (base - off_a);
  if (from_cam)
    bounce_tile<true>(rays + base, keys + base, keys_next, nvalid, P, lm, cam, base, off_a, off_r, out_active, out_rest,
                      hit_rest);
  else
    bounce_tile<false>(rays + base, keys + base, keys_next, nvalid, P, lm, cam, base, off_a, off_r, out_active, out_rest,
                       hit_rest);
}

__global__ __launch_bounds__(TILE) void k_evolve(splat_adray_t* __restrict__ rays, int64_t R,
                                                 const int64_t* __restrict__ d_count,
                                                 const int32_t* __restrict__ hit_idx,
                                                 const uint64_t* __restrict__ keys,
                                                 const splat_ftri_t* __restrict__ tris, int32_t N,
                                                 const float* __restrict__ rndm, float force_rand, uint64_t seed,
                                                 int32_t bounce, int32_t first_diffuse, splat_optics_t opt,
                                                 int32_t* __restrict__ tile_counts) {
  const int64_t Rl = live_count_t(R, d_count);
  const int64_t r = (int64_t)blockIdx.x * TILE + threadIdx.x;
  bool active_after = false;
  if (r < Rl && rays[r].status == SPLAT_RAY_ACTIVE) {
    RayF ray = load_rayf(rays + r);
    int32_t i = keys ? key_index(keys[r]) : hit_idx[r];
    if (i < 1 || i > N) i = 1;  // an index outside tris reads as the miss sentinel
    FTriF F = load_ftri(tris + (i - 1));
    float u = bounce_random(force_rand, rndm, r, seed, bounce, ray.dest);
    RayF o = evolve_one(ray, i, F, u, first_diffuse, opt);
    store_rayf(rays + r, o);
    active_after = o.status == SPLAT_RAY_ACTIVE;
  }
  if (tile_counts) {
    int c = __syncthreads_count(active_after);
    if (threadIdx.x == 0) tile_counts[blockIdx.x] = c;
  }
}

__global__ __launch_bounds__(256) void k_light_map(splat_adray_t* __restrict__ rays, int64_t R,
                                                   const int64_t* __restrict__ d_count,
                                                   const int32_t* __restrict__ hit_idx,
                                                   const splat_ftri_t* __restrict__ tris, int32_t N,
                                                   const splat_lightmap_params_t lm, float* __restrict__ tex) {
  const int64_t Rl = live_count_t(R, d_count);
  const int64_t rr = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (rr >= Rl || rays[rr].status != SPLAT_RAY_DIFFUSE) return;
  RayF r = load_rayf(rays + rr);
  int32_t i = hit_idx[rr];
  if (i < 1 || i > N) i = 1;  // an index outside tris reads as the miss sentinel
  final_evolve_and_splat(r, i, load_ftri(tris + (i - 1)), lm, tex, rays + rr);
}

// ---- pinhole ray source: camera_ray (src/ray_generators.jl:18-50), evaluated in Float64 and
// rounded to the Float32 ADRay fields as the reference does ----------------------------------
__global__ __launch_bounds__(256) void k_raygen(const splat_pinhole_t cam, splat_adray_t* __restrict__ rays, int64_t R) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= R) return;
  raygen_one(cam, q, rays + q);
}

// ---- rectangular light sources: rays_from_lights (src/ray_generators.jl:64-113) -------------
__global__ __launch_bounds__(256) void k_raygen_lights(const LightSet A, splat_adray_t* __restrict__ rays, int64_t R) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= R) return;
  rays[q] = light_ray(A, q);
}

// ---- backward shading (continuum_shade!, src/ray_imagers.jl:440-503) ----------------------------
// one thread per ray (rays in pixel order on an h-row grid): final_evolution in place, then its
// n_delta^2 samples into the (h * n_delta)-row interleaved RGB image
__global__ __launch_bounds__(256) void k_shade(splat_adray_t* __restrict__ rays, int64_t R,
                                               const int32_t* __restrict__ hit_idx,
                                               const splat_ftri_t* __restrict__ tris, int32_t N,
                                               const splat_lightmap_params_t lm, const float* __restrict__ tex,
                                               const float* __restrict__ sky, int32_t SW, int32_t SH, int64_t h,
                                               float* __restrict__ rgb) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= R) return;
  RayF r = load_rayf(rays + q);
  int32_t i = hit_idx[q];
  if (i < 1 || i > N) i = 1;
  const FTriF F = load_ftri(tris + (i - 1));
  if (r.status == SPLAT_RAY_DIFFUSE) {
    final_evolve(r, i, F);
    store_rayf(rays + q, r);
  }
  const int nd = lm.n_delta;
  const int64_t img_h = h * nd, x = q % h, y = q / h;
  for (int b = 0; b < nd; ++b)
    for (int a = 0; a < nd; ++a) {
      float c3[3];
      shade_sample(r, F, lm, tex, sky, SW, SH, a, b, c3);
      float* o = rgb + ((a + nd * x) + img_h * (b + nd * y)) * 3;
      o[0] = c3[0];
      o[1] = c3[1];
      o[2] = c3[2];
    }
}

// run_evolution!'s reorder (src/tracer.jl:270-278): record and hit index of ray r to slot dest - 1
// (a dest outside 1..R is skipped)
__global__ __launch_bounds__(256) void k_reorder(const splat_adray_t* __restrict__ src, const int32_t* __restrict__ hsrc,
                                                 int64_t R, splat_adray_t* __restrict__ dst, int32_t* __restrict__ hdst) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  const int64_t d = src[r].dest;
  if (d < 1 || d > R) return;
  const uint2* s = reinterpret_cast<const uint2*>(src + r);
  uint2* t = reinterpret_cast<uint2*>(dst + (d - 1));
#pragma unroll
  for (int w = 0; w < (int)(sizeof(splat_adray_t) / 8); ++w) t[w] = s[w];
  hdst[d - 1] = hsrc[r];
}

// ---- small utilities (graph-capturable: plain kernels, no memset nodes) ----------------------
__global__ void k_fill_f32(float* __restrict__ dst, int64_t n, float v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) dst[i] = v;
}
__global__ void k_set_i64(int64_t* dst, int64_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = v;
}
// copies the first min(*d_count, R) records as 8-byte words, coalesced
__global__ void k_copy_rays(const uint2* __restrict__ src, uint2* __restrict__ dst, int64_t R, const int64_t* __restrict__ d_count) {
  const int64_t n = live_count_t(R, d_count) * (int64_t)(sizeof(splat_adray_t) / 8);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) dst[i] = src[i];
}

constexpr int64_t COUNT_GRID = 1024;  // workgroups of the count pass (about one resident set)
void launch_bounce_count(const splat_adray_t* rays, int64_t R, const int64_t* d_count, const uint64_t* keys,
                         const BounceParams& P, const splat_pinhole_t* cam, int32_t* tile_counts, hipStream_t s) {
  if (R <= 0) return;
  splat_pinhole_t c;
  if (cam) c = *cam;
  else std::memset(&c, 0, sizeof(c));
  const int64_t groups = (R + 4 * TILE - 1) / (4 * TILE);  // four tiles per workgroup
  hipLaunchKernelGGL(k_bounce_count, dim3((unsigned)(groups < COUNT_GRID ? groups : COUNT_GRID)), dim3(TILE), 0, s, rays,
                     R, d_count, keys, P, c, cam ? 1 : 0, tile_counts);
}
static unsigned frame_grid(int64_t n) {
  int64_t g = (n + 255) / 256;
  return (unsigned)(g < 4096 ? (g < 1 ? 1 : g) : 4096);
}
void launch_frame_start(float* tex, int64_t tex_n, int64_t* count0, int64_t R, uint64_t* keys, hipStream_t s) {
  const int64_t n = (tex ? tex_n : 0) > (keys ? R : 0) ? (tex ? tex_n : 0) : (keys ? R : 0);
  hipLaunchKernelGGL(k_frame_start, dim3(frame_grid(n)), dim3(256), 0, s, tex, tex_n, count0, R, keys);
}
void launch_frame_end(const splat_adray_t* src, splat_adray_t* dst, int64_t R, const int64_t* d_count,
                      const uint64_t* keys, int32_t* hit_idx, hipStream_t s) {
  if (R <= 0) return;
  hipLaunchKernelGGL(k_frame_end, dim3(frame_grid(R * 17)), dim3(256), 0, s, (const uint2*)src, (uint2*)dst, R,
                     d_count, keys, hit_idx);
}
void launch_bounce(const splat_adray_t* rays, int64_t R, const int64_t* d_count, const uint64_t* keys,
                   uint64_t* keys_next, const BounceParams& P, const splat_lightmap_params_t* lm,
                   const splat_pinhole_t* cam, const int32_t* tile_sums, bool local_scan, int64_t* d_total,
                   splat_adray_t* out_active, splat_adray_t* out_rest, int32_t* hit_rest, hipStream_t s) {
  if (R <= 0) return;
  splat_lightmap_params_t l;
  if (lm) l = *lm;
  else std::memset(&l, 0, sizeof(l));
  splat_pinhole_t c;
  if (cam) c = *cam;
  else std::memset(&c, 0, sizeof(c));
  BounceParams Q = P;
  if (!lm) Q.tex = nullptr;
  hipLaunchKernelGGL(k_bounce, dim3((unsigned)((R + TILE - 1) / TILE)), dim3(TILE), 0, s, rays, R, d_count, keys,
                     keys_next, Q, l, c, cam ? 1 : 0, tile_sums, local_scan ? 1 : 0, d_total, out_active, out_rest,
                     hit_rest);
}

void launch_evolve(splat_adray_t* rays, int64_t R, const int64_t* d_count, const int32_t* hit_idx, const uint64_t* keys,
                   const splat_ftri_t* tris,
                   int32_t N, const float* rndm, float force_rand, uint64_t seed, int32_t bounce, int32_t first_diffuse,
                   const splat_optics_t& opt, int32_t* tile_counts, hipStream_t s) {
  if (R <= 0) return;
  hipLaunchKernelGGL(k_evolve, dim3((unsigned)((R + TILE - 1) / TILE)), dim3(TILE), 0, s, rays, R, d_count, hit_idx, keys, tris, N,
                     rndm, force_rand, seed, bounce, first_diffuse, opt, tile_counts);
}
void launch_light_map(splat_adray_t* rays, int64_t R, const int64_t* d_count, const int32_t* hit_idx, const splat_ftri_t* tris,
                      int32_t N, const splat_lightmap_params_t& lm, float* tex, hipStream_t s) {
  if (R <= 0) return;
  hipLaunchKernelGGL(k_light_map, dim3((unsigned)((R + 255) / 256)), dim3(256), 0, s, rays, R, d_count, hit_idx, tris, N, lm, tex);
}
void launch_shade(splat_adray_t* rays, int64_t R, const int32_t* hit_idx, const splat_ftri_t* tris, int32_t N,
                  const splat_lightmap_params_t& lm, const float* tex, const float* sky, int32_t SW, int32_t SH,
                  int64_t h, float* rgb, hipStream_t s) {
  if (R <= 0) return;
  hipLaunchKernelGGL(k_shade, dim3((unsigned)((R + 255) / 256)), dim3(256), 0, s, rays, R, hit_idx, tris, N, lm, tex, sky,
                     SW, SH, h, rgb);
}
void launch_reorder(const splat_adray_t* src, const int32_t* hsrc, int64_t R, splat_adray_t* dst, int32_t* hdst,
                    hipStream_t s) {
  if (R <= 0) return;
  hipLaunchKernelGGL(k_reorder, dim3((unsigned)((R + 255) / 256)), dim3(256), 0, s, src, hsrc, R, dst, hdst);
}
void launch_raygen(const splat_pinhole_t& cam, splat_adray_t* rays, int64_t R, hipStream_t s) {
  if (R <= 0) return;
  hipLaunchKernelGGL(k_raygen, dim3((unsigned)((R + 255) / 256)), dim3(256), 0, s, cam, rays, R);
}
void launch_raygen_lights(const LightSet& A, splat_adray_t* rays, int64_t R, hipStream_t s) {
  if (R <= 0) return;
  hipLaunchKernelGGL(k_raygen_lights, dim3((unsigned)((R + 255) / 256)), dim3(256), 0, s, A, rays, R);
}
static unsigned stream_grid(int64_t n) {
  int64_t g = (n + 255) / 256;
  return (unsigned)(g < 8192 ? (g < 1 ? 1 : g) : 8192);
}
void launch_fill_f32(float* dst, int64_t n, float v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_f32, dim3(stream_grid(n)), dim3(256), 0, s, dst, n, v);
}
void launch_set_i64(int64_t* dst, int64_t v, hipStream_t s) { hipLaunchKernelGGL(k_set_i64, dim3(1), dim3(64), 0, s, dst, v); }
void launch_copy_rays(const splat_adray_t* src, splat_adray_t* dst, int64_t R, const int64_t* d_count, hipStream_t s) {
  if (R <= 0) return;
  hipLaunchKernelGGL(k_copy_rays, dim3(stream_grid(R * 17)), dim3(256), 0, s, (const uint2*)src, (uint2*)dst, R, d_count);
}

}  // namespace spl
