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
// cam != nullptr (hitters): the rays are that pinhole's rays 0..R-1, generated on the fly
void launch_hit_bf(const splat_adray_t* rays, int64_t R, const int64_t* d_count, const TriRec* recs, int32_t N,
                   int32_t* hit_idx, uint64_t* keys, int waves, uint64_t* stats, hipStream_t s,
                   const splat_pinhole_t* cam = nullptr);
void launch_group_boxes(const float4* boxes, int32_t n_bv, int32_t group_size, float4* groups, hipStream_t s);
// Binary tree over the BVs in their order, heap-indexed: node n has children 2n and 2n + 1, the
// 2^levels leaves are nodes 2^levels + b (BV b, or empty padding). nodes[2n] / nodes[2n + 1] are
// the node's box; nodes[2n].w holds a code: bits 0-1 = the axis that separates the children's
// centres most, bit 2 = the right child lies lower on it, bit 3 = empty subtree.
constexpr int32_t TREE_AXIS = 3, TREE_SWAP = 4, TREE_EMPTY = 8;
// a non-empty leaf node carries its member range: count = code >> TREE_COUNT_SHIFT, first record
// index = the bits of nodes[2n + 1].w
constexpr int32_t TREE_COUNT_SHIFT = 4;
constexpr int32_t MAX_TREE_LEVELS = 24;
// light-map parameters a kernel can use (sizes in range; curve abscissae ascending)
int32_t check_lightmap(const splat_lightmap_params_t* lm);
// a prepared hit set the kernels can use (pointers present, tree depth within MAX_TREE_LEVELS)
int32_t check_hitset(const splat_hitset_t* hs);
// the device BV build (bvbuild.hip): member list and ranges of 1 + 2^depth BVs (the large
// triangles, then the object-median leaves) of triangles 2..n+1; h_summary <- {n_large, n - n_large}
// once the stream reaches it
size_t bv_build_scratch_bytes(int32_t n, int32_t depth);
int32_t bv_build_members(const splat_ftri_t* tris, int32_t n, int32_t depth, double big_factor, int32_t max_big,
                         void* scratch, int32_t* members, int2* ranges, int32_t* h_summary, hipStream_t s);
void launch_tree_build(const float4* boxes, const int2* ranges, int32_t n_bv, int32_t levels, float4* nodes,
                       hipStream_t s);

// Prepared scene of the BV hitter as the kernels see it: the leaves (boxes, member ranges,
// records) under one of three top structures -- none (every leaf box tested), a coarse level of
// groups of gsize consecutive leaves, or a binary tree over the leaves (nodes != nullptr).
struct BVScene {
  const float4* boxes;
  const int2* ranges;
  const TriRec* recs;
  const float4* groups;
  int32_t n_bv, gsize, n_groups;
  const float4* nodes;
  int32_t levels;
  int32_t walk;     // tree only: BV_WALK_*
  int32_t n_top;    // records recs[0, n_top) that every ray tests before the structure
  int32_t no_tail;  // tree only: walk the tree even for launches with few live rays
  int32_t split_target;  // single-frame launches: items the split factor aims for (SPLIT_TARGETS)
};
// How a wave walks the tree: together (one path, the union of its rays'; scalar node loads), or
// each lane its own path; AUTO: together when all its rays point into one octant.
enum BVWalk { BV_WALK_WAVE = 0, BV_WALK_AUTO = 1, BV_WALK_LANES = 2 };
enum BVMode { BV_FLAT = 0, BV_GROUPS = 1, BV_TREE = 2 };
// split targets of the single-frame BV hitter (splat_hitset_t.tune_waves bits 8-10; 0: the default)
constexpr int32_t SPLIT_TARGETS[8] = {12288, 3072, 6144, 24576, 49152, 12288, 12288, 12288};
// the kernels' view of a prepared BV hit set
inline BVScene bv_scene(const splat_hitset_t* hs) {
  BVScene b;
  const int32_t top = hs->top_bvs > 0 && hs->top_bvs < hs->n_bv ? hs->top_bvs : 0;
  b.boxes = static_cast<const float4*>(hs->d_boxes) + 2 * top;
  b.ranges = static_cast<const int2*>(hs->d_ranges) + top;
  b.recs = static_cast<const TriRec*>(hs->d_recs);
  b.n_bv = hs->n_bv - top;
  b.n_top = top > 0 ? hs->top_recs : 0;
  const bool tree = hs->d_groups && hs->tree_levels > 0;
  const bool g = !tree && hs->d_groups && hs->group_size > 0 && hs->n_groups > 0;
  b.groups = g ? static_cast<const float4*>(hs->d_groups) : nullptr;
  b.gsize = g ? hs->group_size : 0;
  b.n_groups = g ? hs->n_groups : 0;
  b.nodes = tree ? static_cast<const float4*>(hs->d_groups) : nullptr;
  b.levels = tree ? hs->tree_levels - 1 : 0;
  const int32_t w = hs->tune_waves & 3;
  b.walk = tree && w <= BV_WALK_LANES ? w : BV_WALK_WAVE;
  b.no_tail = tree && (hs->tune_waves & 4) ? 1 : 0;
  b.split_target = SPLIT_TARGETS[(hs->tune_waves >> 8) & 7];
  return b;
}

inline __host__ __device__ int bv_mode(const BVScene& s) {
  return s.nodes ? BV_TREE : (s.groups && s.gsize > 0 && s.n_groups > 0 ? BV_GROUPS : BV_FLAT);
}
// Largest grid (workgroups) of the single-frame hitter after the first bounce: four rounds of the
// chip's wave slots; the items of a larger frame stride over it. A grid sized for every ray of a
// 33.5M-ray frame dispatched 131,072 workgroups at each late bounce for a few thousand live rays.
constexpr int64_t HIT_GRID_CAP = 8192;
// BV hitter into packed keys (initialised by launch_key_init on the same stream)
void launch_key_init(uint64_t* keys, int64_t R, const int64_t* d_count, hipStream_t s);
void launch_hit_bv(const splat_adray_t* rays, int64_t R, const int64_t* d_count, const BVScene& scene, uint64_t* keys,
                   uint64_t* stats, hipStream_t s, const splat_pinhole_t* cam = nullptr);
void launch_key_final(const uint64_t* keys, int64_t R, const int64_t* d_count, int32_t* hit_idx, hipStream_t s);

// trace.hip
struct BounceParams;  // bounce.hpp
// cam != nullptr: the live rays are that pinhole's rays, generated on the fly (first bounce)
void launch_bounce_count(const splat_adray_t* rays, int64_t R, const int64_t* d_count, const uint64_t* keys,
                         const BounceParams& P, const splat_pinhole_t* cam, int32_t* tile_counts, hipStream_t s);
// tex (nullable) cleared, *count0 = R, keys (nullable) set to the miss key
void launch_frame_start(float* tex, int64_t tex_n, int64_t* count0, int64_t R, uint64_t* keys, hipStream_t s);
// live rays src -> dst; hit_idx from keys when keys != nullptr
void launch_frame_end(const splat_adray_t* src, splat_adray_t* dst, int64_t R, const int64_t* d_count,
                      const uint64_t* keys, int32_t* hit_idx, hipStream_t s);
// lm == nullptr: no splat (P.tex ignored); keys_next (nullable): miss keys at the next positions.
// local_scan: tile_sums are the count pass's tile counts (R within SCAN_LOCAL_TILES tiles) and the
// bounce pass sums them itself, writing the ACTIVE total to d_total; otherwise tile_sums and
// d_total are launch_scan_tiles' offsets and total.
constexpr int64_t SCAN_LOCAL_TILES = 4096;
void launch_bounce(const splat_adray_t* rays, int64_t R, const int64_t* d_count, const uint64_t* keys,
                   uint64_t* keys_next, const BounceParams& P, const splat_lightmap_params_t* lm,
                   const splat_pinhole_t* cam, const int32_t* tile_sums, bool local_scan, int64_t* d_total,
                   splat_adray_t* out_active, splat_adray_t* out_rest, int32_t* hit_rest, hipStream_t s);
// hit index per ray from hit_idx, or from the packed keys when keys != nullptr
void launch_evolve(splat_adray_t* rays, int64_t R, const int64_t* d_count, const int32_t* hit_idx, const uint64_t* keys,
                   const splat_ftri_t* tris, int32_t N, const float* rndm, float force_rand, uint64_t seed, int32_t bounce,
                   int32_t first_diffuse, const splat_optics_t& opt, int32_t* tile_counts, hipStream_t s);
void launch_light_map(splat_adray_t* rays, int64_t R, const int64_t* d_count, const int32_t* hit_idx,
                      const splat_ftri_t* tris, int32_t N, const splat_lightmap_params_t& lm, float* tex,
                      hipStream_t s);
void launch_shade(splat_adray_t* rays, int64_t R, const int32_t* hit_idx, const splat_ftri_t* tris, int32_t N,
                  const splat_lightmap_params_t& lm, const float* tex, const float* sky, int32_t SW, int32_t SH,
                  int64_t h, float* rgb, hipStream_t s);
void launch_reorder(const splat_adray_t* src, const int32_t* hsrc, int64_t R, splat_adray_t* dst, int32_t* hdst,
                    hipStream_t s);
void launch_raygen(const splat_pinhole_t& cam, splat_adray_t* rays, int64_t R, hipStream_t s);
struct LightSet;
void launch_raygen_lights(const LightSet& A, splat_adray_t* rays, int64_t R, hipStream_t s);
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
