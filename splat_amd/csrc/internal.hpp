// internal.hpp -- launchers shared between the kernel files and the C ABI (not installed).
#pragma once

#include "spec.hpp"

namespace spl {

constexpr int TILE = 256;  // rays per partition tile / evolve workgroup

// hit.hip
void launch_prep_ntris(const splat_ntri_t* tris, int32_t N, TriRec* out, hipStream_t s);
void launch_prep_ftris(const splat_ftri_t* tris, int32_t N, TriRec* out, hipStream_t s);
void launch_prep_members(const splat_ftri_t* tris, const int32_t* members, int64_t total, TriRec* out, hipStream_t s);
void launch_bv_boxes(const splat_ftri_t* tris, const int32_t* members, const int2* ranges, int32_t n_bv, float4* boxes,
                     hipStream_t s);
int choose_bf_waves(int64_t R, int32_t N);
// brute force: writes hit_idx, or the packed (t, index) keys when keys != nullptr
void launch_hit_bf(const splat_adray_t* rays, int64_t R, const int64_t* d_count, const TriRec* recs, int32_t N,
                   int32_t* hit_idx, uint64_t* keys, int waves, uint64_t* stats, hipStream_t s);
void launch_group_boxes(const float4* boxes, int32_t n_bv, int32_t group_size, float4* groups, hipStream_t s);

// Prepared scene of the BV hitter as the kernels see it.
struct BVScene {
  const float4* boxes;
  const int2* ranges;
  const TriRec* recs;
  const float4* groups;
  int32_t n_bv, gsize, n_groups;
};
// BV hitter into packed keys (initialised by launch_key_init on the same stream)
void launch_key_init(uint64_t* keys, int64_t R, const int64_t* d_count, hipStream_t s);
void launch_hit_bv(const splat_adray_t* rays, int64_t R, const int64_t* d_count, const BVScene& scene, uint64_t* keys,
                   uint64_t* stats, hipStream_t s);
void launch_key_final(const uint64_t* keys, int64_t R, const int64_t* d_count, int32_t* hit_idx, hipStream_t s);

// trace.hip
// hit index per ray from hit_idx, or from the packed keys when keys != nullptr
void launch_evolve(splat_adray_t* rays, int64_t R, const int64_t* d_count, const int32_t* hit_idx, const uint64_t* keys,
                   const splat_ftri_t* tris, int32_t N, const float* rndm, float force_rand, uint64_t seed, int32_t bounce,
                   int32_t first_diffuse, const splat_optics_t& opt, int32_t* tile_counts, hipStream_t s);
void launch_light_map(splat_adray_t* rays, int64_t R, const int64_t* d_count, const int32_t* hit_idx,
                      const splat_ftri_t* tris, int32_t N, const splat_lightmap_params_t& lm, float* tex,
                      hipStream_t s);
void launch_raygen(const splat_pinhole_t& cam, splat_adray_t* rays, int64_t R, hipStream_t s);
void launch_fill_f32(float* dst, int64_t n, float v, hipStream_t s);
void launch_set_i64(int64_t* dst, int64_t v, hipStream_t s);
void launch_copy_rays(const splat_adray_t* src, splat_adray_t* dst, int64_t R, const int64_t* d_count, hipStream_t s);

// partition.hip
void launch_count_active(const splat_adray_t* rays, int64_t R, const int64_t* d_count, int32_t* tile_counts,
                         hipStream_t s);
// exclusive scan of tile counts -> tile offsets; total -> *d_total
void launch_scan_tiles(const int32_t* tile_counts, int64_t n_tiles, const int64_t* d_count, int64_t R,
                       int32_t* tile_offsets, int64_t* d_total, hipStream_t s);
// stable scatter: ACTIVE -> out_active[offset + rank], others -> out_rest[*d_total + rest_rank];
// the optional carry array moves with the rest part (carry_rest aligned with out_rest).
void launch_scatter(const splat_adray_t* in, int64_t R, const int64_t* d_count, const int32_t* tile_offsets,
                    const int64_t* d_total, splat_adray_t* out_active, splat_adray_t* out_rest, const int32_t* carry_in,
                    const uint64_t* carry_keys, int32_t* carry_rest, hipStream_t s);

}  // namespace spl
