/*
 * splat_hip.h -- C ABI of the MI355X (gfx950) per-frame tracing hot path of SPLAT.
 *
 * Every entry point is a plain C function over device pointers and sizes. It enqueues its work
 * on the caller's HIP stream (passed as an opaque `void*`, NULL = the null stream) and returns
 * an int32 status: 0 on success, a negative SPLAT_ERR_* code otherwise (splat_strerror() gives
 * the message). Nothing throws across the ABI. Apart from the serving loop, no entry point
 * allocates device memory: workspaces are sized by the *_bytes() helpers and owned by the
 * caller. All layouts below are byte-for-byte the Julia isbits layouts of the reference, so a
 * Julia caller passes `pointer(::CuArray)`-style device pointers with zero copies.
 *
 * Index conventions follow the reference: triangle indices are 1-based Julia indices, index 1 is
 * the degenerate all-zero triangle that doubles as the "miss" sentinel (src/geo.jl:312,
 * src/scenes/sphere.jl:36), triangles >= first_diffuse are diffuse (src/tracer.jl:186).
 */
#ifndef SPLAT_HIP_H
#define SPLAT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPLAT_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------------------------ */
#define SPLAT_OK 0
#define SPLAT_ERR_INVALID_ARG (-1) /* null pointer, negative size, bad enum               */
#define SPLAT_ERR_WORKSPACE (-2)   /* caller workspace too small                          */
#define SPLAT_ERR_HIP (-3)         /* a HIP runtime call failed                           */
#define SPLAT_ERR_UNSUPPORTED (-4) /* parameter outside the supported range               */
#define SPLAT_ERR_CAPACITY (-5)    /* a fixed-capacity table would overflow               */
#define SPLAT_ERR_BUDGET (-6)      /* serving: request does not fit the HBM budget         */

/* ---- ray status (src/geo.jl:78-80) ------------------------------------------------------- */
#define SPLAT_RAY_DIFFUSE 0
#define SPLAT_RAY_ACTIVE 1
#define SPLAT_RAY_INFINITY 2

/* ---- data layouts (Julia isbits, src/geo.jl:12-113, src/tracer.jl:313) ------------------- */

/* ADRay, src/geo.jl:60-76. 136 bytes: 8 x vec3 (96 B), Bool @96, Int64 @104/@112,
 * Float32 lambda/x/y @120/124/128, UInt8 status @132. */
typedef struct splat_adray {
  float pos[3];
  float dpdl[3]; /* d pos / d lambda */
  float dpdx[3]; /* d pos / d x      */
  float dpdy[3]; /* d pos / d y      */
  float dir[3];
  float dddl[3]; /* d dir / d lambda */
  float dddx[3];
  float dddy[3];
  uint8_t in_medium;
  uint8_t _pad0[7];
  int64_t ignore_tri; /* 1-based triangle index the ray just left (1 = none) */
  int64_t dest;       /* sample id (camera_ray: pixel index, src/ray_generators.jl:32) */
  float lambda;       /* wavelength, nm */
  float x, y;         /* expansion point of the pixel-offset differentials */
  uint8_t status;     /* SPLAT_RAY_* */
  uint8_t _pad1[3];
} splat_adray_t;

/* Tri = SVector{4,R3}: [face normal, a, b, c] (src/geo.jl:16). 48 bytes. */
typedef struct splat_tri {
  float v[4][3];
} splat_tri_t;

/* n_tris element = Tuple{Int32,Tri} (src/tracer.jl:313). 52 bytes. */
typedef struct splat_ntri {
  int32_t idx;
  float v[4][3];
} splat_ntri_t;

/* FTri = SVector{10,R3}: [n, a, b, c, n_a, n_b, n_c, uv_a, uv_b, uv_c] (src/geo.jl:18,
 * src/geo.jl:337). 120 bytes. */
typedef struct splat_ftri {
  float v[10][3];
} splat_ftri_t;

/* BVBox (src/geo.jl:23-26), 24 bytes; Sphere (src/geo.jl:40-43), 16 bytes. */
typedef struct splat_bvbox {
  float min[3];
  float max[3];
} splat_bvbox_t;

typedef struct splat_sphere {
  float origin[3];
  float radius;
} splat_sphere_t;

/* ---- optics ------------------------------------------------------------------------------ */
#define SPLAT_IOR_REFERENCE_LINEAR 0 /* glass(lambda): 1.3 @400nm -> 1.4 @700nm (src/material.jl:1-9) */
#define SPLAT_IOR_CAUCHY 1           /* n = A + B / lambda_um^2                                      */

/* Reference-compatibility switches (both on reproduces src/tracer.jl:118-179 exactly). */
#define SPLAT_COMPAT_DY_AT_X (1 << 0)       /* d dir / d y taken at y = r.x (src/tracer.jl:138-141, :166-169) */
#define SPLAT_COMPAT_REFRACT_NORMAL_LAMBDA (1 << 1) /* refract's lambda tangent holds the normal fixed (:130-133) */
#define SPLAT_COMPAT_REFERENCE (SPLAT_COMPAT_DY_AT_X | SPLAT_COMPAT_REFRACT_NORMAL_LAMBDA)

typedef struct splat_optics {
  int32_t ior_model; /* SPLAT_IOR_* */
  float cauchy_a;    /* Cauchy A (dimensionless)          */
  float cauchy_b;    /* Cauchy B (um^2)                   */
  int32_t compat;    /* SPLAT_COMPAT_* bits               */
} splat_optics_t;

/* ---- prepared hit sets (the hitter's per-scene state, cf. src/alg_types.jl:10-90) ------- */
#define SPLAT_HITTER_BF 0 /* brute force: StableHitter, EH1..EH6 (src/hitters.jl:24-310)         */
#define SPLAT_HITTER_BV 1 /* one-level bounding volumes: DPBVHitter (src/hitters.jl:486-648)      */

/* Host-side descriptor filled by splat_hitset_prepare_*(); the device arrays it points to live
 * inside the caller's workspace. Plain data: copy it freely, it owns nothing. */
typedef struct splat_hitset {
  int32_t kind;          /* SPLAT_HITTER_*                                              */
  int32_t n_tris;        /* N, including the degenerate index 1                          */
  int32_t n_bv;          /* number of bounding volumes (BV only)                         */
  int32_t max_members;   /* largest BV membership (BV only)                              */
  int64_t total_members; /* sum of BV memberships (BV only)                              */
  void* d_recs;          /* device: prepared triangle records, 64 B each                 */
  void* d_boxes;         /* device: conservative member boxes, 32 B each (BV only)       */
  void* d_ranges;        /* device: int32 (first record, count) per BV (BV only)          */
  void* d_groups;        /* device: top-structure boxes, 32 B each (BV only): group boxes,  */
                         /* or the binary tree's node boxes when tree_levels > 0             */
  int32_t group_size;    /* BVs per group; 0 = no coarse level (BV only)                  */
  int32_t n_groups;      /* ceil(n_bv / group_size)                                        */
  int32_t tune_waves;    /* BF: triangle-split waves per workgroup (0 = autotuned default); */
                         /* BV tree: 0 = waves walk together, 1 = per wave, together when */
                         /* its rays share a direction octant, else per lane; 2 = per lane; */
                         /* + 4: walk the tree also when few rays are live (by default a   */
                         /* launch with <= 4096 live rays scans the leaf boxes flat);        */
                         /* BV, single-frame launches: + 256 x c, c = 1..4: the split factor */
                         /* aims for 3072, 6144, 24576 or 49152 (64-ray group, split) items  */
                         /* instead of 12288 (c = 0)                                         */
  int32_t tree_levels;   /* 1 + levels of the tree over the BVs; 0 = no tree (BV only)    */
  int32_t top_bvs;       /* leading BVs whose members every ray tests directly (BV only) */
  int32_t top_recs;      /* their records: d_recs[0, top_recs)                            */
  /* Optional device uint64[5] counters, accumulated by every hitter launch when non-NULL:
   * [0] ray-box slab tests on each ray's own path (the root of its split, then the children of
   * every node that ray entered; every box of a flat scan), [1] (ray, leaf) pairs entered,
   * [2] ray-triangle tests of the algorithm (brute force: every live ray x every triangle; BV:
   * the members of the leaves each ray entered, and the large-triangle list), [3] ray-triangle
   * and [4] ray-box tests the wavefronts executed (64 lanes per wave test, live or not). NULL
   * costs nothing. */
  uint64_t* d_stats;
} splat_hitset_t;

/* ---- light-map (splat) parameters (src/ray_imagers.jl:78-165) ---------------------------- */
#define SPLAT_MAX_LAMBDA 64
#define SPLAT_MAX_DELTA 8
#define SPLAT_MAX_CURVE 24
#define SPLAT_WEIGHTS_IDENTITY 0 /* channel k <- spectrum[k], weight 1 (the reference's tex[i,j,k]) */
#define SPLAT_WEIGHTS_TABLE 1    /* channel c <- weights[k*3+c] * cos (c < channels <= 3)          */
#define SPLAT_WEIGHTS_CURVES 2   /* channel c <- curve_c(lambda) * cos, piecewise linear (the retina
                                    response of src/rgb_spectrum.jl:116-148), lambda = splat lambda */

typedef struct splat_lightmap_params {
  int32_t n_lambda;  /* grid size; 0 = own-lambda mode (one splat per (dx,dy) at the ray's lambda) */
  int32_t n_delta;   /* sub-pixel offsets per axis (upres(N), src/alg_types.jl:108)              */
  int32_t channels;  /* light-map planes                                                         */
  int32_t weight_mode; /* SPLAT_WEIGHTS_*                                                         */
  int32_t width, height; /* light-map texels (tex[i,j,c], i fastest, Julia column-major)          */
  float spectrum[SPLAT_MAX_LAMBDA];
  float delta[SPLAT_MAX_DELTA];
  float weights[SPLAT_MAX_LAMBDA * 3]; /* SPLAT_WEIGHTS_TABLE: [k][c] */
  int32_t curve_n[3];                  /* SPLAT_WEIGHTS_CURVES: points per channel (>= 2) */
  float curve_x[3][SPLAT_MAX_CURVE];   /* strictly ascending abscissae (nm; else SPLAT_ERR_INVALID_ARG) */
  float curve_y[3][SPLAT_MAX_CURVE];   /* ordinates (already normalised)                  */
} splat_lightmap_params_t;

/* ---- pinhole ray source (camera_ray, src/ray_generators.jl:18-62) ------------------------ */
#define SPLAT_ORDER_COLUMN_MAJOR 0 /* ray k <-> pixel k, rows fastest (the reference's order) */
#define SPLAT_ORDER_TILED8 1       /* 8x8 pixel tiles, tiles column-major (coherent waves)     */
#define SPLAT_LAMBDA_FIXED 0       /* every ray at lambda_min (the reference uses 550 nm)       */
#define SPLAT_LAMBDA_UNIFORM 1     /* per-sample uniform in [lambda_min, lambda_max) from seed  */
#define SPLAT_LAMBDA_STRATIFIED 2  /* sample l of L: stratum l of [lambda_min, lambda_max)      */

typedef struct splat_pinhole {
  float pos[3];   /* pinhole (the point light)             */
  float dir[3];   /* unit view direction                    */
  float up[3];    /* unit up                                */
  float right[3]; /* unit right                             */
  float fov_half_sin;
  int32_t height; /* rows (x in 1..height)                  */
  int32_t width;  /* columns (y in 1..width)                */
  int32_t samples_per_pixel;
  int32_t order;       /* SPLAT_ORDER_*  */
  int32_t lambda_mode; /* SPLAT_LAMBDA_* */
  float lambda_min, lambda_max;
  float jitter; /* dv scale numerator: dir += dv * jitter / max(h, w) (reference: 0.025) */
  int32_t _pad;
  uint64_t seed;
} splat_pinhole_t;

/* ---- rectangular light sources (RectLight, rays_from_light{,s}, src/ray_generators.jl:4-13,
 * :64-113) ---------------------------------------------------------------------------------- */
#define SPLAT_MAX_LIGHTS 16

typedef struct splat_rect_light {
  float center[3]; /* centre of the emitting rectangle                                  */
  float dir[3];    /* emission direction, used as given                                 */
  float dim1[3];   /* half extent along rows (x)                                        */
  float dim2[3];   /* half extent along columns (y)                                     */
  int64_t res1;    /* rows before upscale                                               */
  int64_t res2;    /* columns before upscale                                            */
} splat_rect_light_t; /* 64 bytes: Julia's isbits RectLight */

typedef struct splat_light_rays {
  int32_t upscale;     /* forward_upscale: a light emits (res1/upscale) x (res2/upscale) rays */
  int32_t lambda_mode; /* SPLAT_LAMBDA_FIXED (the reference: 550 nm) or SPLAT_LAMBDA_UNIFORM  */
  float lambda_min, lambda_max;
  float jitter;        /* cell offsets dx, dy = jitter * U[0,1) (the reference: rand(Float32)) */
  int32_t _pad;
  uint64_t seed;
} splat_light_rays_t;

/* ---- fused bounce loop (run_evolution!, src/tracer.jl:233-282) --------------------------- */
typedef struct splat_run_params {
  int32_t depth;         /* bounces                                                    */
  int32_t first_diffuse; /* 1-based index of the first diffuse triangle                */
  float force_rand;      /* NaN = seeded per-ray random; 0 = always reflect; 1 = refract */
  int32_t _pad;
  uint64_t seed;         /* bounce random numbers: f(seed, bounce, ray.dest)           */
  splat_optics_t optics;
  /* Optional instrumentation: hipEvent_t handles (as void*) recorded on the stream immediately
   * before and after launches: [0, 2*(depth+1)) each hitter launch (bounces 1..depth, then the
   * realigning pass); with 4*depth+2 handles also [2*(depth+1), 4*depth+2) each bounce's fused
   * pass (evolve, splat of the rays that become DIFFUSE, stable partition). NULL or fewer
   * handles record nothing (resp. only the hitter launches). */
  void** hit_events;
  int32_t n_hit_events;
  int32_t _pad2;
} splat_run_params_t;

/* ===================================== entry points ====================================== */

const char* splat_strerror(int32_t code);
int32_t splat_abi_version(void);
/* Writes sizeof of (adray, tri, ntri, ftri, bvbox, sphere) into out[0..5]. */
int32_t splat_layout_sizes(int64_t* out6);
/* Writes sizeof of (optics, hitset, lightmap_params, pinhole, run_params, frame_request) into
 * out[0..5], for bindings that mirror these structs. */
int32_t splat_param_sizes(int64_t* out6);
/* Device count visible to HIP (0 when none; never fails). */
int32_t splat_device_count(void);

/* Triangle preparation: hoists the per-triangle terms of distance_to_plane and in_triangle
 * (src/geo.jl:433-470) into 64-byte records. Replaces the per-call n_tris tuple build of
 * trace! (src/tracer.jl:313). */
size_t splat_hitset_bf_bytes(int32_t n_tris);
int32_t splat_hitset_prepare_bf(const splat_ntri_t* n_tris, int32_t n_tris_len, void* ws,
                                size_t ws_bytes, splat_hitset_t* out, void* stream);
int32_t splat_hitset_prepare_bf_ftri(const splat_ftri_t* tris, int32_t n_tris_len, void* ws,
                                     size_t ws_bytes, splat_hitset_t* out, void* stream);

/* Bounding-volume preparation from the reference's BV memberships (DPBVHitter constructor,
 * src/alg_types.jl:81-90, pack_bv_tris src/bv.jl:126-136). `bv_tri_count` (n_bv) and
 * `bv_tris` (Julia column-major n_bv x bv_stride matrix of 1-based indices) are HOST arrays.
 * The boxes are rebuilt from the members and padded so the test is conservative. Over them:
 * group_size == 0 -- a binary tree over the BVs in their order (children of node n: 2n, 2n+1),
 *                    walked depth-first, nearer child first, by each wavefront (default);
 * group_size > 0  -- one coarse level of boxes over `group_size` consecutive BVs;
 * group_size < 0  -- nothing: every BV box is tested (the reference's one-level structure).
 * The first n_top BVs are excluded from that structure: every ray tests their members directly,
 * first (meant for a few large triangles -- a ground plane -- that would otherwise inflate every
 * box above them). Any memberships that cover the triangles give the brute-force result. */
size_t splat_hitset_bv_bytes(int32_t n_bv, int64_t total_members);

/* The BV structure built on the device from the triangles alone (bv_partition + pack_bv_tris +
 * the DPBVHitter constructor, src/bv.jl:68-136, src/alg_types.jl:81-90, for the tree hitter):
 * 1 + 2^depth BVs -- BV 0 holds the large triangles (bounding-box diagonal > big_factor x the
 * median, at most max_big, largest first), tested by every ray; BVs 1.. are the object-median
 * leaves of the rest (split at the median centroid of each node's widest centroid axis, ties
 * to the lower index), some possibly empty -- with the binary tree over the leaves. The same
 * memberships as splat_amd.bv.tree_partition on the host. One stream synchronisation at the
 * end. `ws` (splat_hitset_build_bv_bytes) holds the hitset's arrays first; the rest of it is
 * build scratch, free again on return. */
size_t splat_hitset_build_bv_bytes(int32_t n_tris_len, int32_t depth);
int32_t splat_hitset_build_bv(const splat_ftri_t* tris, int32_t n_tris_len, int32_t depth,
                              float big_factor, int32_t max_big, void* ws, size_t ws_bytes,
                              splat_hitset_t* out, void* stream);
int32_t splat_hitset_prepare_bv(const splat_ftri_t* tris, int32_t n_tris_len, int32_t n_bv,
                                const int64_t* bv_tri_count, const int64_t* bv_tris,
                                int64_t bv_stride, int32_t group_size, int32_t n_top, void* ws,
                                size_t ws_bytes, splat_hitset_t* out, void* stream);

/* next_hit!(tracer, hitter, rays, n_tris) (src/hitters.jl:24-30, :34-82, :168-226, :615-648):
 * hit_idx[r] = the lexicographic minimum over triangles j of (t_rj, j), ties to the lowest
 * index, 1 on a miss. `d_count` (device int64, nullable) limits the rays to the first
 * min(*d_count, R) -- the live count kept on the device between bounces. A BV hit set needs a
 * device workspace of splat_next_hit_ws_bytes(R) bytes (the packed (t, index) key per ray
 * that the reference encodes in src/atomic_argmin.jl:5-20);
 * brute force takes ws = NULL. */
size_t splat_next_hit_ws_bytes(int64_t R);
int32_t splat_next_hit(const splat_adray_t* rays, int64_t R, const int64_t* d_count,
                       const splat_hitset_t* hs, int32_t* hit_idx, void* ws, size_t ws_bytes,
                       void* stream);
/* One-call brute force over an n_tris array (prepares into ws, then hits). */
int32_t splat_next_hit_bf(const splat_adray_t* rays, int64_t R, const splat_ntri_t* n_tris,
                          int32_t n_tris_len, int32_t* hit_idx, void* ws, size_t ws_bytes,
                          void* stream);

/* splat_next_hit without a workspace. No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). Same result as splat_next_hit; a BV hit set's per-ray keys live in a
 * library-owned device array of SPLAT_NOALLOC_MAX_RAYS keys (8 B each, 16 MiB), one copy per
 * device, held until the process ends. R above SPLAT_NOALLOC_MAX_RAYS returns
 * SPLAT_ERR_CAPACITY. Every call on one device shares that device's array (each call
 * re-initialises the keys it uses), so calls that overlap on one device -- on different streams
 * or host threads -- race: serialise them. Allocates nothing (no hipMalloc/hipFree). */
#define SPLAT_NOALLOC_MAX_RAYS (1 << 21)
int32_t splat_next_hit_noalloc(const splat_adray_t* rays, int64_t R, const int64_t* d_count,
                               const splat_hitset_t* hs, int32_t* hit_idx, void* stream);

/* rays .= evolve_ray.(rays, hit_idx, tris[hit_idx], rndm, first_diffuse)
 * (src/tracer.jl:181-201, :118-179). Random number per ray: force_rand when not NaN
 * (src/tracer.jl:256-258), else rndm[r] when rndm != NULL, else f(seed, bounce, ray.dest). */
int32_t splat_evolve(splat_adray_t* rays, int64_t R, const int64_t* d_count,
                     const int32_t* hit_idx, const splat_ftri_t* tris, int32_t n_tris_len,
                     const float* rndm, float force_rand, uint64_t seed, int32_t bounce,
                     int32_t first_diffuse, const splat_optics_t* optics, void* stream);

/* partition!(rays, swap; by = r.status != ACTIVE) (src/partition.jl:122-132): stable -- ACTIVE
 * rays first in their original order, then the rest in their original order -- and in place
 * (the result is copied back into `rays`). The ACTIVE count goes to d_count_out (device) and,
 * if h_count_out != NULL, also to the host (this synchronises the stream). */
size_t splat_partition_ws_bytes(int64_t R);
int32_t splat_partition_active(splat_adray_t* rays, splat_adray_t* swap, int64_t R,
                               int64_t* d_count_out, int64_t* h_count_out, void* ws,
                               size_t ws_bytes, void* stream);

/* continuum_light_map2! (src/ray_imagers.jl:142-165): final_evolution of DIFFUSE rays
 * (:4-37, mutates rays) then the atomic splat (:78-113) into tex (float, channels x height x
 * width planes, texel (i,j,c) at i + width*(j + height*c)). tex is accumulated into, not
 * cleared. */
int32_t splat_light_map(splat_adray_t* rays, int64_t R, const int64_t* d_count,
                        const int32_t* hit_idx, const splat_ftri_t* tris, int32_t n_tris_len,
                        const splat_lightmap_params_t* lm, float* tex, void* stream);

/* Seeded pinhole ray source: R = height * width * samples_per_pixel rays. */
int32_t splat_raygen_pinhole(const splat_pinhole_t* cam, splat_adray_t* rays, int64_t R,
                             void* stream);

/* rays_from_lights(lights, upscale) (src/ray_generators.jl:64-113). `lights` is host memory
 * (a Julia Vector{RectLight}). Light i's rays follow light i-1's; ray j of a light is cell
 * (j % h + 1, j / h + 1) of its h x w = (res1/upscale) x (res2/upscale) grid, column-major, at
 * offset (x + dx, y + dy) with origin center + dim1 (x+dx - h/2)/(h/2) + dim2 (y+dy - w/2)/(w/2)
 * (integer halves), dest 0. With several lights each must emit res1*res2/upscale^2 rays, the
 * slice the reference sizes for it (src/ray_generators.jl:107). Returns the total ray count, or
 * a negative error code. */
int64_t splat_lights_ray_count(const splat_rect_light_t* lights, int32_t n_lights, int32_t upscale);
int32_t splat_raygen_lights(const splat_rect_light_t* lights, int32_t n_lights,
                            const splat_light_rays_t* params, splat_adray_t* rays, int64_t R,
                            void* stream);

/* Fill helper (graph-safe clear, no hipMemsetAsync): n floats <- value. */
int32_t splat_fill_f32(float* dst, int64_t n, float value, void* stream);

/* run_evolution!(hitter, ::ExperimentalTracer; depth, first_diffuse, ...) fused: every count
 * stays on the device (no host sync per bounce), so the whole loop can be captured in a HIP
 * graph. On return `rays` and `hit_idx` hold exactly what the reference's loop plus its final
 * realigning next_hit! leave (src/tracer.jl:247-281, reorder = false): live rays first, then
 * the rays retired at bounce depth, ..., then those retired at bounce 1, each group in
 * order. `d_live_out` (device int64) receives the final live count. */
size_t splat_run_ws_bytes(int64_t R);
int32_t splat_run_evolution(splat_adray_t* rays, int64_t R, const splat_hitset_t* hs,
                            const splat_ftri_t* tris, int32_t n_tris_len, int32_t* hit_idx,
                            const splat_run_params_t* params, void* ws, size_t ws_bytes,
                            int64_t* d_live_out, void* stream);

/* One full light-map frame: pinhole rays -> run_evolution -> light map, all on the device.
 * `tex` is cleared first. `rays`/`hit_idx` receive cam's R final rays and hit indices, as
 * splat_run_evolution leaves them; both NULL: only the light map is kept (the reference's forward
 * pass discards its rays after the splat), and the writes of retired rays and the realigning hit
 * that only feed those outputs are skipped. ws from splat_run_ws_bytes. */
int32_t splat_trace_frame(const splat_pinhole_t* cam, splat_adray_t* rays, int64_t R,
                          const splat_hitset_t* hs, const splat_ftri_t* tris,
                          int32_t n_tris_len, int32_t* hit_idx, const splat_run_params_t* params,
                          const splat_lightmap_params_t* lm, float* tex, void* ws,
                          size_t ws_bytes, void* stream);

/* The forward pass of trace! (src/tracer.jl:316-341): rays_from_lights(lights, upscale) ->
 * run_evolution! -> continuum_light_map2!, all on the device, the splat inside the bounce loop.
 * R = splat_lights_ray_count(lights, n_lights, lp->upscale); the rest as splat_trace_frame. */
int32_t splat_trace_lights(const splat_rect_light_t* lights, int32_t n_lights,
                           const splat_light_rays_t* lp, splat_adray_t* rays, int64_t R,
                           const splat_hitset_t* hs, const splat_ftri_t* tris, int32_t n_tris_len,
                           int32_t* hit_idx, const splat_run_params_t* params,
                           const splat_lightmap_params_t* lm, float* tex, void* ws, size_t ws_bytes,
                           void* stream);

/* ---- backward pass: camera rays -> image (the rest of trace!, src/tracer.jl:346-385) ------ */
/* run_evolution!(...; reorder = true)'s reorder (src/tracer.jl:270-278): ray r and its hit index
 * move to slot r.dest - 1, i.e. back to pixel order (dest must be a permutation of 1..R; a dest
 * outside 1..R is dropped). Workspace: splat_reorder_ws_bytes(R). */
size_t splat_reorder_ws_bytes(int64_t R);
int32_t splat_reorder_by_dest(splat_adray_t* rays, int64_t R, int32_t* hit_idx, void* ws, size_t ws_bytes,
                              void* stream);
/* continuum_shade!(::ExperimentalImager2; tracer, RGB, tris, rays, spectrum, retina_factor, tex,
 * skydome) (src/ray_imagers.jl:440-503, shade_tex4_kernel :381-436) for rays in pixel order on a
 * grid of grid_h rows (ray q is pixel (q % grid_h, q / grid_h)): DIFFUSE rays are final-evolved in
 * place; then per ray and sub-pixel offset (a, b) of lm->delta: an INFINITY ray reads the skydome
 * (3 float planes sky_w x sky_h, u fastest) at u = acos(d_z)/pi * sky_w, v = (atan2(d_y, d_x)/pi
 * + 1)/2 * sky_h; a DIFFUSE ray sums, over the lm->n_lambda planes of the light map `tex` at its
 * texture coordinate, lm->weights[k*3 + c] x the plane's value; channels clamped to [0, 1].
 * Textures are read bilinearly at 1-based coordinates, clamped at the edges (CUDA.jl's linear
 * texture reads, with exact FP32 weights). rgb: (grid_h * n_delta) x (R / grid_h * n_delta)
 * pixels, 3 interleaved floats each, row index fastest (the reference's RGB layout). */
int32_t splat_shade(splat_adray_t* rays, int64_t R, const int32_t* hit_idx, const splat_ftri_t* tris,
                    int32_t n_tris_len, const splat_lightmap_params_t* lm, const float* tex, const float* sky,
                    int32_t sky_w, int32_t sky_h, int64_t grid_h, float* rgb, void* stream);

/* ---- multi-frame engine: batching and continuous serving --------------------------------
 * Frames (each with its own resident scene, camera, depth and light map) occupy slots of an
 * engine. One step advances every admitted frame by one stage -- raygen + bounce 1, bounce k,
 * ..., final hit + splat -- in ONE fixed sequence of launches shared by all frames: a planning
 * kernel turns the frames' device-side live counts into work-item prefixes, and every kernel
 * maps its hardware-dispatched workgroups onto the items of all frames (the hitter's items are
 * (64-ray group, split) pairs). A step has fixed launch shapes per hitter grid size, so it is
 * captured once into a HIP graph and replayed. Frames
 * can be admitted between any two steps (continuous batching); each frame's result is the same
 * as splat_trace_frame's for the same request. The engine owns one device arena (allocated at
 * creation, the only allocation) from which each admitted frame takes its ray buffers;
 * admission fails with SPLAT_ERR_BUDGET when the arena or the slot table is full. */
typedef struct splat_engine splat_engine_t;

typedef struct splat_frame_request {
  splat_pinhole_t cam;       /* the frame's rays                                      */
  splat_hitset_t hs;         /* prepared BV hit set of a resident scene (SPLAT_HITTER_BV) */
  const splat_ftri_t* tris;  /* device: the scene's FTri array                         */
  int32_t n_tris;
  int32_t depth;             /* bounces, <= SPLAT_ENGINE_MAX_DEPTH                      */
  int32_t first_diffuse;
  float force_rand;          /* NaN = seeded per-ray random numbers                    */
  uint64_t seed;
  splat_optics_t optics;
  float* tex;                /* device light map (channels x height x width), cleared by the engine */
  splat_adray_t* rays_out;   /* device, nullable: final rays (else kept in the arena)   */
  int32_t* hit_out;          /* device, nullable: final hit indices                     */
  int64_t tag;               /* caller's id, reported back by splat_engine_poll         */
  splat_lightmap_params_t lm;
} splat_frame_request_t;

#define SPLAT_ENGINE_MAX_DEPTH 16
size_t splat_engine_frame_bytes(const splat_pinhole_t* cam, int32_t with_outputs);
int32_t splat_engine_create(int64_t arena_bytes, int32_t max_slots, splat_engine_t** out);
int32_t splat_engine_destroy(splat_engine_t* e);
/* Takes a slot and arena space for the request; enqueues nothing: the next splat_engine_step
 * uploads every request admitted since the previous step (a few copies per step, not per frame)
 * and starts them. `stream` is not used. Steps of one engine go to one stream. */
int32_t splat_engine_admit(splat_engine_t* e, const splat_frame_request_t* req, void* stream);
int32_t splat_engine_step(splat_engine_t* e, void* stream);
/* Reports the tags of frames whose final step has completed (wait != 0: block until every
 * launched step is done). Their slots and arena space are released. */
int32_t splat_engine_poll(splat_engine_t* e, int32_t wait, int64_t* tags, int32_t max_tags, int32_t* n_out);
/* Admitted frames not yet reported by splat_engine_poll. */
int32_t splat_engine_in_flight(const splat_engine_t* e);
/* Frames the next splat_engine_step advances (in flight minus those whose last step has been
 * launched): 0 means a step would launch nothing and the caller should poll with wait != 0. */
int32_t splat_engine_runnable(const splat_engine_t* e);
/* Blocks until at most `keep` launched steps are still running on the device. */
int32_t splat_engine_sync_steps(splat_engine_t* e, int32_t keep);
int64_t splat_engine_arena_free(const splat_engine_t* e);
/* on != 0 (default): each step replays one captured HIP graph; 0: plain stream launches. */
int32_t splat_engine_use_graph(splat_engine_t* e, int32_t on);
/* on != 0: a step's hitter is one persistent launch (a round of workgroups on every CU) whose
 * waves take their (64-ray group, split) items of every frame from device work queues; 0
 * (default, faster on the 64-frame batch and in serving): one wave per item over a grid the host
 * sizes from the frames' ray counts. The results are the same. No reference counterpart (the
 * reference traces one frame per launch sequence; the nearest is the dynamic-parallelism
 * launcher, src/hitters.jl:592-612). */
int32_t splat_engine_use_queue(splat_engine_t* e, int32_t on);

#ifdef __cplusplus
}
#endif

#endif /* SPLAT_HIP_H */
