// capi.hip -- the C ABI (include/splat_hip.h): argument checks, workspace carving, launches and
// the device-resident bounce loop.
#include <cmath>
#include <cstring>
#include <mutex>
#include <utility>
#include <vector>

#include "bounce.hpp"

using namespace spl;

namespace {

constexpr size_t ALIGN = 256;
inline size_t align_up(size_t x) { return (x + ALIGN - 1) & ~(ALIGN - 1); }

struct Carver {
  char* base;
  size_t used = 0;
  explicit Carver(void* p) : base(static_cast<char*>(p)) {}
  template <class T> T* take(size_t count) {
    T* p = reinterpret_cast<T*>(base ? base + used : nullptr);
    used += align_up(count * sizeof(T));
    return p;
  }
};

inline hipStream_t S(void* s) { return static_cast<hipStream_t>(s); }

inline int32_t launch_status() {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? SPLAT_OK : SPLAT_ERR_HIP;
}

constexpr int MAX_DEPTH = 64;

struct RunWs {
  splat_adray_t* s1;
  splat_adray_t* s2;
  uint64_t* keys;  // packed (t, index) nearest hit of each live ray, ping-pong with keys2
  uint64_t* keys2;
  int32_t* tile_counts;
  int32_t* tile_offsets;
  int64_t* counters;  // counters[k] = live rays after bounce k (counters[0] = R)
  size_t bytes;
};

RunWs carve_run(void* ws, int64_t R) {
  Carver c(ws);
  int64_t n_tiles = (R + TILE - 1) / TILE;
  RunWs w;
  w.s1 = c.take<splat_adray_t>((size_t)R);
  w.s2 = c.take<splat_adray_t>((size_t)R);
  w.keys = c.take<uint64_t>((size_t)R);
  w.keys2 = c.take<uint64_t>((size_t)R);
  w.tile_counts = c.take<int32_t>((size_t)n_tiles);
  w.tile_offsets = c.take<int32_t>((size_t)n_tiles);
  w.counters = c.take<int64_t>(MAX_DEPTH + 2);
  w.bytes = c.used;
  return w;
}

struct PartWs {
  int32_t* tile_counts;
  int32_t* tile_offsets;
  int64_t* total;
  size_t bytes;
};
PartWs carve_part(void* ws, int64_t R) {
  Carver c(ws);
  int64_t n_tiles = (R + TILE - 1) / TILE;
  PartWs w;
  w.tile_counts = c.take<int32_t>((size_t)n_tiles);
  w.tile_offsets = c.take<int32_t>((size_t)n_tiles);
  w.total = c.take<int64_t>(1);
  w.bytes = c.used;
  return w;
}


int bf_waves(const splat_hitset_t* hs, int64_t R) {
  return hs->tune_waves > 0 ? hs->tune_waves : choose_bf_waves(R, hs->n_tris);
}

// nearest hit of the live rays as packed (t, index) keys: hit_keys_init, then hit_keys_launch
void hit_keys_init(int64_t R, const int64_t* d_count, const splat_hitset_t* hs, uint64_t* keys, hipStream_t s) {
  if (hs->kind == SPLAT_HITTER_BV) launch_key_init(keys, R, d_count, s);  // brute force writes every key
}
void hit_keys_launch(const splat_adray_t* rays, int64_t R, const int64_t* d_count, const splat_hitset_t* hs,
                     uint64_t* keys, hipStream_t s, const splat_pinhole_t* cam = nullptr) {
  if (hs->kind == SPLAT_HITTER_BF)
    launch_hit_bf(rays, R, d_count, static_cast<const TriRec*>(hs->d_recs), hs->n_tris, nullptr, keys, bf_waves(hs, R),
                  hs->d_stats, s, cam);
  else
    launch_hit_bv(rays, R, d_count, bv_scene(hs), keys, hs->d_stats, s, cam);
}

// nearest hit index of the live rays into hit_idx (keys: scratch of the BV hitter)
void hit_index(const splat_adray_t* rays, int64_t R, const int64_t* d_count, const splat_hitset_t* hs, int32_t* hit_idx,
               uint64_t* keys, hipStream_t s) {
  if (hs->kind == SPLAT_HITTER_BF) {
    launch_hit_bf(rays, R, d_count, static_cast<const TriRec*>(hs->d_recs), hs->n_tris, hit_idx, nullptr, bf_waves(hs, R),
                  hs->d_stats, s);
  } else {
    hit_keys_init(R, d_count, hs, keys, s);
    hit_keys_launch(rays, R, d_count, hs, keys, s);
    launch_key_final(keys, R, d_count, hit_idx, s);
  }
}

struct HitWs {
  uint64_t* keys;
  size_t bytes;
};
HitWs carve_hit(void* ws, int64_t R) {
  Carver c(ws);
  HitWs w;
  w.keys = c.take<uint64_t>((size_t)R);
  w.bytes = c.used;
  return w;
}

int32_t check_optics(const splat_optics_t* o) {
  if (!o) return SPLAT_ERR_INVALID_ARG;
  if (o->ior_model != SPLAT_IOR_REFERENCE_LINEAR && o->ior_model != SPLAT_IOR_CAUCHY) return SPLAT_ERR_INVALID_ARG;
  return SPLAT_OK;
}


// the device-resident bounce loop; results land in `final_rays`. The live rays start in w.s1, or,
// with cam != nullptr (and depth >= 1), are that pinhole's rays, generated where the first
// bounce needs them. With lm != nullptr, rays that become DIFFUSE are final-evolved and splatted
// into tex in the loop (the light map of the frame), otherwise they are left as run_evolution!
// leaves them. A light-map frame with neither final_rays nor hit_final keeps no rays: retired
// rays are not written, and the realigning hit and final copy (which only feed those outputs)
// are skipped.
int32_t run_core(splat_adray_t* final_rays, int64_t R, const splat_hitset_t* hs, const splat_ftri_t* tris,
                 int32_t* hit_final, const splat_run_params_t* p, const RunWs& w, hipStream_t s,
                 const splat_lightmap_params_t* lm = nullptr, float* tex = nullptr,
                 const splat_pinhole_t* cam = nullptr) {
  const int64_t n_tiles = (R + TILE - 1) / TILE;
  const bool ev = p->hit_events && p->n_hit_events >= 2 * (p->depth + 1);
  const bool ev_all = p->hit_events && p->n_hit_events >= 4 * p->depth + 2;
  auto mark = [&](int i) { if (ev) (void)hipEventRecord(static_cast<hipEvent_t>(p->hit_events[i]), s); };
  auto mark_part = [&](int i) {
    if (ev_all) (void)hipEventRecord(static_cast<hipEvent_t>(p->hit_events[2 * (p->depth + 1) + i]), s);
  };
  const bool bv = hs->kind == SPLAT_HITTER_BV;
  BounceParams P;
  P.tris = tris;
  P.n_tris = hs->n_tris;
  P.first_diffuse = p->first_diffuse;
  P.force_rand = p->force_rand;
  P.seed = p->seed;
  P.optics = p->optics;
  P.tex = lm ? tex : nullptr;
  splat_adray_t* cur = w.s1;
  splat_adray_t* nxt = w.s2;
  uint64_t* kc = w.keys;   // this bounce's keys
  uint64_t* kn = w.keys2;  // the next bounce's, set to the miss key by the bounce pass
  // the first bounce's keys need the miss key first unless every one is stored by the hitter
  // itself: a BV launch over at least its split target's 64-ray groups runs one item per group
  // (split factor 1), which stores its rays' keys instead of merging them
  const bool keys_stored = bv && (R + 63) / 64 >= bv_scene(hs).split_target;
  launch_frame_start(lm ? tex : nullptr, lm ? (int64_t)lm->width * lm->height * lm->channels : 0, &w.counters[0], R,
                     bv && !keys_stored ? kc : nullptr, s);
  for (int k = 1; k <= p->depth; ++k) {
    const int64_t* live = &w.counters[k - 1];
    const splat_pinhole_t* gen = k == 1 ? cam : nullptr;
    mark(2 * (k - 1));
    hit_keys_launch(cur, R, live, hs, kc, s, gen);
    mark(2 * (k - 1) + 1);
    P.bounce = k;
    launch_bounce_count(cur, R, live, kc, P, gen, w.tile_counts, s);
    // after the first bounce (whose tiles are all live: a scan launch costs less than every
    // workgroup summing them all), frames of up to SCAN_LOCAL_TILES tiles let the bounce pass sum
    // the live tiles' counts itself, one launch fewer
    const bool local_scan = k >= 2 && n_tiles <= SCAN_LOCAL_TILES;
    if (!local_scan) launch_scan_tiles(w.tile_counts, n_tiles, live, R, w.tile_offsets, &w.counters[k], s);
    mark_part(2 * (k - 1));
    launch_bounce(cur, R, live, kc, bv ? kn : nullptr, P, lm, gen, local_scan ? w.tile_counts : w.tile_offsets,
                  local_scan, &w.counters[k], nxt, final_rays, hit_final, s);
    mark_part(2 * (k - 1) + 1);
    std::swap(cur, nxt);
    std::swap(kc, kn);
  }
  const int64_t* live = &w.counters[p->depth];
  if (!final_rays && !hit_final) return launch_status();  // a light map only: nothing to realign
  // the reference's realigning next_hit! (src/tracer.jl:280)
  if (!bv) {
    mark(2 * p->depth);
    hit_index(cur, R, live, hs, hit_final, nullptr, s);
    mark(2 * p->depth + 1);
  } else {
    mark(2 * p->depth);
    hit_keys_launch(cur, R, live, hs, kc, s);  // kc holds the miss keys of the live rays
    mark(2 * p->depth + 1);
  }
  launch_frame_end(cur, final_rays, R, live, bv ? kc : nullptr, hit_final, s);
  return launch_status();
}

// (outputs_optional: a light-map frame may keep neither its rays nor their hit indices, both NULL)
int32_t check_run(const splat_adray_t* rays, int64_t R, const splat_hitset_t* hs, const splat_ftri_t* tris, int32_t N,
                  const int32_t* hit_idx, const splat_run_params_t* p, void* ws, size_t ws_bytes,
                  bool outputs_optional = false) {
  const bool outputs_ok = (rays && hit_idx) || (outputs_optional && !rays && !hit_idx);
  if (!outputs_ok || R < 0 || !tris || N < 1 || !p || !ws) return SPLAT_ERR_INVALID_ARG;
  if (int32_t e = check_hitset(hs)) return e;
  if (int32_t e = check_optics(&p->optics)) return e;
  if (p->depth < 0 || p->depth > MAX_DEPTH) return SPLAT_ERR_UNSUPPORTED;
  if (R > (int64_t)INT32_MAX) return SPLAT_ERR_UNSUPPORTED;
  if (ws_bytes < carve_run(nullptr, R).bytes) return SPLAT_ERR_WORKSPACE;
  return SPLAT_OK;
}

}  // namespace

// (shared with the engine's admission) the prepared hit set's sizes within what the kernels
// assume: the tree walk's stack holds the pending nodes of at most MAX_TREE_LEVELS levels
int32_t spl::check_hitset(const splat_hitset_t* hs) {
  if (!hs || !hs->d_recs || hs->n_tris < 1) return SPLAT_ERR_INVALID_ARG;
  if (hs->kind == SPLAT_HITTER_BV) {
    if (!hs->d_boxes || !hs->d_ranges || hs->n_bv < 0) return SPLAT_ERR_INVALID_ARG;
    if (hs->tree_levels < 0 || hs->tree_levels > MAX_TREE_LEVELS + 1 || hs->top_bvs < 0 || hs->top_recs < 0 ||
        hs->group_size < 0 || hs->n_groups < 0)
      return SPLAT_ERR_INVALID_ARG;
    // the structure must reach every BV below the top list (the kernels visit nothing else)
    const int32_t top = hs->top_bvs > 0 && hs->top_bvs < hs->n_bv ? hs->top_bvs : 0;
    const int64_t below = (int64_t)hs->n_bv - top;
    if (top > 0 && hs->top_recs > hs->total_members) return SPLAT_ERR_INVALID_ARG;
    if (hs->d_groups && hs->tree_levels > 0) {
      if (((int64_t)1 << (hs->tree_levels - 1)) < below) return SPLAT_ERR_INVALID_ARG;
    } else if (hs->d_groups && hs->group_size > 0 && hs->n_groups > 0) {
      if ((int64_t)hs->n_groups * hs->group_size < below) return SPLAT_ERR_INVALID_ARG;
    }
  } else if (hs->kind != SPLAT_HITTER_BF) {
    return SPLAT_ERR_INVALID_ARG;
  }
  return SPLAT_OK;
}

// (shared with the engine's admission)
int32_t spl::check_lightmap(const splat_lightmap_params_t* lm) {
  if (!lm || lm->width < 1 || lm->height < 1 || lm->channels < 1) return SPLAT_ERR_INVALID_ARG;
  if (lm->n_lambda < 0 || lm->n_lambda > SPLAT_MAX_LAMBDA) return SPLAT_ERR_UNSUPPORTED;
  if (lm->n_delta < 1 || lm->n_delta > SPLAT_MAX_DELTA) return SPLAT_ERR_UNSUPPORTED;
  switch (lm->weight_mode) {
    case SPLAT_WEIGHTS_IDENTITY:
      if (lm->channels != (lm->n_lambda > 0 ? lm->n_lambda : 1)) return SPLAT_ERR_INVALID_ARG;
      break;
    case SPLAT_WEIGHTS_TABLE:
      if (lm->channels > 3 || lm->n_lambda < 1) return SPLAT_ERR_INVALID_ARG;
      break;
    case SPLAT_WEIGHTS_CURVES:
      if (lm->channels > 3) return SPLAT_ERR_INVALID_ARG;
      for (int c = 0; c < lm->channels; ++c) {
        if (lm->curve_n[c] < 2 || lm->curve_n[c] > SPLAT_MAX_CURVE) return SPLAT_ERR_INVALID_ARG;
        // the abscissae ascend strictly (not NaN, no repeat: a zero-width segment divides by zero)
        for (int j = 0; j + 1 < lm->curve_n[c]; ++j)
          if (!(lm->curve_x[c][j] < lm->curve_x[c][j + 1])) return SPLAT_ERR_INVALID_ARG;
      }
      break;
    default:
      return SPLAT_ERR_INVALID_ARG;
  }
  return SPLAT_OK;
}

extern "C" {

const char* splat_strerror(int32_t code) {
  switch (code) {
    case SPLAT_OK: return "ok";
    case SPLAT_ERR_INVALID_ARG: return "invalid argument";
    case SPLAT_ERR_WORKSPACE: return "workspace too small";
    case SPLAT_ERR_HIP: return "HIP runtime error";
    case SPLAT_ERR_UNSUPPORTED: return "unsupported parameter";
    case SPLAT_ERR_CAPACITY: return "capacity exceeded";
    case SPLAT_ERR_BUDGET: return "request exceeds the memory budget";
    default: return "unknown error";
  }
}

int32_t splat_abi_version(void) { return SPLAT_ABI_VERSION; }

int32_t splat_layout_sizes(int64_t* out6) {
  if (!out6) return SPLAT_ERR_INVALID_ARG;
  out6[0] = sizeof(splat_adray_t);
  out6[1] = sizeof(splat_tri_t);
  out6[2] = sizeof(splat_ntri_t);
  out6[3] = sizeof(splat_ftri_t);
  out6[4] = sizeof(splat_bvbox_t);
  out6[5] = sizeof(splat_sphere_t);
  return SPLAT_OK;
}

int32_t splat_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

size_t splat_hitset_bf_bytes(int32_t n_tris) { return n_tris > 0 ? align_up((size_t)n_tris * sizeof(TriRec)) : 0; }

int32_t splat_hitset_prepare_bf(const splat_ntri_t* n_tris, int32_t N, void* ws, size_t ws_bytes, splat_hitset_t* out,
                                void* stream) {
  if (!n_tris || N < 1 || !ws || !out) return SPLAT_ERR_INVALID_ARG;
  if (ws_bytes < splat_hitset_bf_bytes(N)) return SPLAT_ERR_WORKSPACE;
  launch_prep_ntris(n_tris, N, static_cast<TriRec*>(ws), S(stream));
  std::memset(out, 0, sizeof(*out));
  out->kind = SPLAT_HITTER_BF;
  out->n_tris = N;
  out->d_recs = ws;
  return launch_status();
}

int32_t splat_hitset_prepare_bf_ftri(const splat_ftri_t* tris, int32_t N, void* ws, size_t ws_bytes, splat_hitset_t* out,
                                     void* stream) {
  if (!tris || N < 1 || !ws || !out) return SPLAT_ERR_INVALID_ARG;
  if (ws_bytes < splat_hitset_bf_bytes(N)) return SPLAT_ERR_WORKSPACE;
  launch_prep_ftris(tris, N, static_cast<TriRec*>(ws), S(stream));
  std::memset(out, 0, sizeof(*out));
  out->kind = SPLAT_HITTER_BF;
  out->n_tris = N;
  out->d_recs = ws;
  return launch_status();
}

// levels of the binary tree over n_bv leaves (2^levels >= n_bv)
static int32_t tree_levels_for(int32_t n_bv) {
  int32_t l = 0;
  while ((1 << l) < n_bv) ++l;
  return l;
}

// the top structure's boxes: group boxes (at most n_bv groups) or tree nodes (2^(levels+1))
static size_t top_boxes(int32_t n_bv) {
  size_t tree_nodes = (size_t)2 << tree_levels_for(n_bv);
  return tree_nodes > (size_t)n_bv ? tree_nodes : (size_t)n_bv;
}

size_t splat_hitset_bv_bytes(int32_t n_bv, int64_t total_members) {
  if (n_bv < 1 || total_members < 0 || tree_levels_for(n_bv) > MAX_TREE_LEVELS) return 0;
  return align_up((size_t)total_members * sizeof(TriRec)) + align_up((size_t)n_bv * 2 * sizeof(float4)) +
         align_up(top_boxes(n_bv) * 2 * sizeof(float4)) + align_up((size_t)n_bv * sizeof(int2)) +
         align_up((size_t)(total_members > 0 ? total_members : 1) * sizeof(int32_t));
}

int32_t splat_hitset_prepare_bv(const splat_ftri_t* tris, int32_t N, int32_t n_bv, const int64_t* bv_tri_count,
                                const int64_t* bv_tris, int64_t bv_stride, int32_t group_size, int32_t n_top, void* ws,
                                size_t ws_bytes, splat_hitset_t* out, void* stream) {
  if (!tris || N < 1 || n_bv < 1 || !bv_tri_count || !bv_tris || bv_stride < 0 || !ws || !out) return SPLAT_ERR_INVALID_ARG;
  if (n_top < 0 || n_top >= n_bv) return SPLAT_ERR_INVALID_ARG;  // at least one BV under the structure
  if (tree_levels_for(n_bv) > MAX_TREE_LEVELS) return SPLAT_ERR_UNSUPPORTED;
  // flatten the Julia column-major (n_bv x stride) membership matrix; drop the sentinel index 1
  std::vector<int32_t> members;
  std::vector<int2> ranges((size_t)n_bv);
  int32_t max_members = 0;
  for (int32_t b = 0; b < n_bv; ++b) {
    int64_t cnt = bv_tri_count[b];
    if (cnt < 0 || cnt > bv_stride) return SPLAT_ERR_INVALID_ARG;
    int32_t start = (int32_t)members.size();
    for (int64_t m = 0; m < cnt; ++m) {
      int64_t idx = bv_tris[b + m * (int64_t)n_bv];
      if (idx < 1 || idx > N) return SPLAT_ERR_INVALID_ARG;
      if (idx == 1) continue;
      members.push_back((int32_t)idx);
    }
    int32_t c = (int32_t)members.size() - start;
    ranges[(size_t)b] = make_int2(start, c);
    max_members = c > max_members ? c : max_members;
  }
  const int64_t total = (int64_t)members.size();
  if (ws_bytes < splat_hitset_bv_bytes(n_bv, total)) return SPLAT_ERR_WORKSPACE;
  // the structure covers BVs [n_top, n_bv); the first n_top BVs' records lead d_recs
  const int32_t m = n_bv - n_top;
  const bool tree = group_size == 0;
  const int32_t levels = tree_levels_for(m);
  if (group_size < 0 || m < 4) group_size = 0;
  const int32_t n_groups = group_size > 0 ? (m + group_size - 1) / group_size : 0;
  Carver c(ws);
  TriRec* recs = c.take<TriRec>((size_t)total);
  float4* boxes = c.take<float4>((size_t)n_bv * 2);
  float4* groups = c.take<float4>(top_boxes(n_bv) * 2);
  int2* d_ranges = c.take<int2>((size_t)n_bv);
  int32_t* d_members = c.take<int32_t>((size_t)(total > 0 ? total : 1));
  hipStream_t s = S(stream);
  if (total > 0 && hipMemcpyAsync(d_members, members.data(), (size_t)total * sizeof(int32_t), hipMemcpyHostToDevice, s) != hipSuccess)
    return SPLAT_ERR_HIP;
  if (hipMemcpyAsync(d_ranges, ranges.data(), ranges.size() * sizeof(int2), hipMemcpyHostToDevice, s) != hipSuccess)
    return SPLAT_ERR_HIP;
  launch_prep_members(tris, d_members, total, recs, s);
  launch_bv_boxes(tris, d_members, d_ranges, n_bv, boxes, s);
  if (tree) launch_tree_build(boxes + 2 * n_top, d_ranges + n_top, m, levels, groups, s);
  else if (n_groups > 0) launch_group_boxes(boxes + 2 * n_top, m, group_size, groups, s);
  if (int32_t e = launch_status()) return e;
  // the host vectors must outlive the copies
  if (hipStreamSynchronize(s) != hipSuccess) return SPLAT_ERR_HIP;
  std::memset(out, 0, sizeof(*out));
  out->kind = SPLAT_HITTER_BV;
  out->n_tris = N;
  out->n_bv = n_bv;
  out->max_members = max_members;
  out->total_members = total;
  out->d_recs = recs;
  out->d_boxes = boxes;
  out->d_ranges = d_ranges;
  out->d_groups = tree || n_groups > 0 ? groups : nullptr;
  out->group_size = tree ? 0 : group_size;
  out->n_groups = tree ? 0 : n_groups;
  out->tree_levels = tree ? levels + 1 : 0;
  out->top_bvs = n_top;
  out->top_recs = n_top > 0 ? ranges[(size_t)n_top].x : 0;
  return SPLAT_OK;
}

size_t splat_hitset_build_bv_bytes(int32_t N, int32_t depth) {
  if (N < 2 || N > (1 << 26) || depth < 1 || depth > 20) return 0;
  const int32_t n_bv = 1 + (1 << depth);
  return splat_hitset_bv_bytes(n_bv, N - 1) + bv_build_scratch_bytes(N - 1, depth);
}

int32_t splat_hitset_build_bv(const splat_ftri_t* tris, int32_t N, int32_t depth, float big_factor, int32_t max_big,
                              void* ws, size_t ws_bytes, splat_hitset_t* out, void* stream) {
  if (!tris || !ws || !out || max_big < 0 || !(big_factor > 0.f)) return SPLAT_ERR_INVALID_ARG;
  const size_t need = splat_hitset_build_bv_bytes(N, depth);
  if (need == 0) return SPLAT_ERR_UNSUPPORTED;
  if (ws_bytes < need) return SPLAT_ERR_WORKSPACE;
  const int32_t n = N - 1, n_bv = 1 + (1 << depth), m = n_bv - 1;
  // the hitset's arrays, laid out as splat_hitset_prepare_bv lays them; the build's scratch after
  Carver c(ws);
  TriRec* recs = c.take<TriRec>((size_t)n);
  float4* boxes = c.take<float4>((size_t)n_bv * 2);
  float4* groups = c.take<float4>(top_boxes(n_bv) * 2);
  int2* d_ranges = c.take<int2>((size_t)n_bv);
  int32_t* d_members = c.take<int32_t>((size_t)n);
  void* scratch = static_cast<char*>(ws) + splat_hitset_bv_bytes(n_bv, n);
  hipStream_t s = S(stream);
  int32_t summary[2] = {0, 0};
  if (int32_t e = bv_build_members(tris, n, depth, (double)big_factor, max_big, scratch, d_members, d_ranges, summary, s))
    return e;
  const int32_t levels = tree_levels_for(m);
  launch_prep_members(tris, d_members, n, recs, s);
  launch_bv_boxes(tris, d_members, d_ranges, n_bv, boxes, s);
  launch_tree_build(boxes + 2, d_ranges + 1, m, levels, groups, s);
  if (int32_t e = launch_status()) return e;
  if (hipStreamSynchronize(s) != hipSuccess) return SPLAT_ERR_HIP;  // the summary is on the host now
  const int32_t n_big = summary[0], rest = summary[1];
  std::memset(out, 0, sizeof(*out));
  out->kind = SPLAT_HITTER_BV;
  out->n_tris = N;
  out->n_bv = n_bv;
  const int32_t leaf_max = (rest + m - 1) / m;
  out->max_members = n_big > leaf_max ? n_big : leaf_max;
  out->total_members = n;
  out->d_recs = recs;
  out->d_boxes = boxes;
  out->d_ranges = d_ranges;
  out->d_groups = groups;
  out->tree_levels = levels + 1;
  out->top_bvs = 1;
  out->top_recs = n_big;
  return SPLAT_OK;
}

size_t splat_next_hit_ws_bytes(int64_t R) { return R >= 0 ? carve_hit(nullptr, R).bytes : 0; }

int32_t splat_next_hit(const splat_adray_t* rays, int64_t R, const int64_t* d_count, const splat_hitset_t* hs,
                       int32_t* hit_idx, void* ws, size_t ws_bytes, void* stream) {
  if (R < 0 || (R > 0 && (!rays || !hit_idx))) return SPLAT_ERR_INVALID_ARG;
  if (int32_t e = check_hitset(hs)) return e;
  HitWs w{nullptr, 0};
  if (hs->kind == SPLAT_HITTER_BV) {
    if (!ws) return SPLAT_ERR_INVALID_ARG;
    if (ws_bytes < splat_next_hit_ws_bytes(R)) return SPLAT_ERR_WORKSPACE;
    w = carve_hit(ws, R);
  }
  hit_index(rays, R, d_count, hs, hit_idx, w.keys, S(stream));
  return launch_status();
}

int32_t splat_next_hit_bf(const splat_adray_t* rays, int64_t R, const splat_ntri_t* n_tris, int32_t N, int32_t* hit_idx,
                          void* ws, size_t ws_bytes, void* stream) {
  splat_hitset_t hs;
  if (int32_t e = splat_hitset_prepare_bf(n_tris, N, ws, ws_bytes, &hs, stream)) return e;
  return splat_next_hit(rays, R, nullptr, &hs, hit_idx, nullptr, 0, stream);
}

// ---- allocation-free next_hit: the BV hitter's key scratch in a library-owned device array ----
// Every call on one device shares that device's copy of `scratch` (the packed key of each ray,
// re-initialised by every call before the hitter reads it), so calls that overlap on one device
// race. Its address is taken once per device, with that device current; host code never uses
// the array's name itself, which on the host is a stand-in and not a device address.
namespace {
static __device__ __align__(256) uint64_t scratch[SPLAT_NOALLOC_MAX_RAYS];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
uint64_t* g_scratch_ptr[MAX_DEVICES] = {};

int32_t device_scratch(uint64_t** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) return SPLAT_ERR_HIP;
  if (dev >= MAX_DEVICES) return SPLAT_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return SPLAT_ERR_HIP;
    g_scratch_ptr[dev] = static_cast<uint64_t*>(p);
  }
  *out = g_scratch_ptr[dev];
  return SPLAT_OK;
}
}  // namespace

int32_t splat_next_hit_noalloc(const splat_adray_t* rays, int64_t R, const int64_t* d_count, const splat_hitset_t* hs,
                               int32_t* hit_idx, void* stream) {
  if (R < 0 || (R > 0 && (!rays || !hit_idx))) return SPLAT_ERR_INVALID_ARG;
  if (R > SPLAT_NOALLOC_MAX_RAYS) return SPLAT_ERR_CAPACITY;
  if (int32_t e = check_hitset(hs)) return e;
  uint64_t* keys = nullptr;
  if (hs->kind == SPLAT_HITTER_BV) {
    if (int32_t e = device_scratch(&keys)) return e;
  }
  hit_index(rays, R, d_count, hs, hit_idx, keys, S(stream));
  return launch_status();
}

int32_t splat_evolve(splat_adray_t* rays, int64_t R, const int64_t* d_count, const int32_t* hit_idx,
                     const splat_ftri_t* tris, int32_t N, const float* rndm, float force_rand, uint64_t seed,
                     int32_t bounce, int32_t first_diffuse, const splat_optics_t* optics, void* stream) {
  if (R < 0 || (R > 0 && (!rays || !hit_idx)) || !tris || N < 1) return SPLAT_ERR_INVALID_ARG;
  if (int32_t e = check_optics(optics)) return e;
  launch_evolve(rays, R, d_count, hit_idx, nullptr, tris, N, rndm, force_rand, seed, bounce, first_diffuse, *optics, nullptr, S(stream));
  return launch_status();
}

size_t splat_partition_ws_bytes(int64_t R) { return R >= 0 ? carve_part(nullptr, R).bytes : 0; }

int32_t splat_partition_active(splat_adray_t* rays, splat_adray_t* swap, int64_t R, int64_t* d_count_out,
                               int64_t* h_count_out, void* ws, size_t ws_bytes, void* stream) {
  if (R < 0 || (R > 0 && (!rays || !swap)) || !ws) return SPLAT_ERR_INVALID_ARG;
  if (R > (int64_t)INT32_MAX) return SPLAT_ERR_UNSUPPORTED;
  if (ws_bytes < splat_partition_ws_bytes(R)) return SPLAT_ERR_WORKSPACE;
  PartWs w = carve_part(ws, R);
  hipStream_t s = S(stream);
  int64_t* total = d_count_out ? d_count_out : w.total;
  const int64_t n_tiles = (R + TILE - 1) / TILE;
  if (R == 0) {
    launch_set_i64(total, 0, s);
  } else {
    launch_count_active(rays, R, nullptr, w.tile_counts, s);
    launch_scan_tiles(w.tile_counts, n_tiles, nullptr, R, w.tile_offsets, total, s);
    launch_scatter(rays, R, nullptr, w.tile_offsets, total, swap, swap, nullptr, nullptr, nullptr, s);
    if (hipMemcpyAsync(rays, swap, (size_t)R * sizeof(splat_adray_t), hipMemcpyDeviceToDevice, s) != hipSuccess)
      return SPLAT_ERR_HIP;
  }
  if (int32_t e = launch_status()) return e;
  if (h_count_out) {
    if (hipMemcpyAsync(h_count_out, total, sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess) return SPLAT_ERR_HIP;
    if (hipStreamSynchronize(s) != hipSuccess) return SPLAT_ERR_HIP;
  }
  return SPLAT_OK;
}

int32_t splat_light_map(splat_adray_t* rays, int64_t R, const int64_t* d_count, const int32_t* hit_idx,
                        const splat_ftri_t* tris, int32_t N, const splat_lightmap_params_t* lm, float* tex, void* stream) {
  if (R < 0 || (R > 0 && (!rays || !hit_idx)) || !tris || N < 1 || !tex) return SPLAT_ERR_INVALID_ARG;
  if (int32_t e = check_lightmap(lm)) return e;
  launch_light_map(rays, R, d_count, hit_idx, tris, N, *lm, tex, S(stream));
  return launch_status();
}

int32_t splat_raygen_pinhole(const splat_pinhole_t* cam, splat_adray_t* rays, int64_t R, void* stream) {
  if (!cam || !rays || cam->height < 1 || cam->width < 1 || cam->samples_per_pixel < 1) return SPLAT_ERR_INVALID_ARG;
  if ((int64_t)cam->height * cam->width * cam->samples_per_pixel != R) return SPLAT_ERR_INVALID_ARG;
  if (!(cam->fov_half_sin > 0.f)) return SPLAT_ERR_INVALID_ARG;
  launch_raygen(*cam, rays, R, S(stream));
  return launch_status();
}

static_assert(sizeof(splat_rect_light_t) == 64, "RectLight is 64 bytes");
static_assert(sizeof(splat_light_rays_t) == 32, "light ray parameters are 32 bytes");

int64_t splat_lights_ray_count(const splat_rect_light_t* lights, int32_t n, int32_t upscale) {
  if (n < 0 || n > SPLAT_MAX_LIGHTS || (n > 0 && !lights) || upscale < 1) return SPLAT_ERR_INVALID_ARG;
  int64_t R = 0;
  for (int32_t i = 0; i < n; ++i) {
    const int64_t r1 = lights[i].res1, r2 = lights[i].res2;
    if (r1 < 0 || r2 < 0 || r1 > (1 << 20) || r2 > (1 << 20)) return SPLAT_ERR_INVALID_ARG;
    const int64_t c = (r1 / upscale) * (r2 / upscale);
    // several lights: light i fills a slice of res1 * res2 / upscale^2 rays (src/ray_generators.jl:107)
    if (n > 1 && c != (r1 * r2) / ((int64_t)upscale * upscale)) return SPLAT_ERR_INVALID_ARG;
    R += c;
  }
  return R;
}

int32_t splat_raygen_lights(const splat_rect_light_t* lights, int32_t n, const splat_light_rays_t* p,
                            splat_adray_t* rays, int64_t R, void* stream) {
  if (!p) return SPLAT_ERR_INVALID_ARG;
  const int64_t total = splat_lights_ray_count(lights, n, p->upscale);
  if (total < 0) return (int32_t)total;
  if (total != R || (R > 0 && !rays)) return SPLAT_ERR_INVALID_ARG;
  if (p->lambda_mode != SPLAT_LAMBDA_FIXED && p->lambda_mode != SPLAT_LAMBDA_UNIFORM) return SPLAT_ERR_INVALID_ARG;
  LightSet A{};
  A.n = n;
  A.p = *p;
  A.base[0] = 0;
  for (int32_t i = 0; i < n; ++i) {
    A.light[i] = lights[i];
    A.base[i + 1] = A.base[i] + (lights[i].res1 / p->upscale) * (lights[i].res2 / p->upscale);
  }
  launch_raygen_lights(A, rays, R, S(stream));
  return launch_status();
}

int32_t splat_fill_f32(float* dst, int64_t n, float value, void* stream) {
  if (n < 0 || (n > 0 && !dst)) return SPLAT_ERR_INVALID_ARG;
  launch_fill_f32(dst, n, value, S(stream));
  return launch_status();
}

size_t splat_run_ws_bytes(int64_t R) { return R >= 0 ? carve_run(nullptr, R).bytes : 0; }

int32_t splat_run_evolution(splat_adray_t* rays, int64_t R, const splat_hitset_t* hs, const splat_ftri_t* tris, int32_t N,
                            int32_t* hit_idx, const splat_run_params_t* p, void* ws, size_t ws_bytes, int64_t* d_live_out,
                            void* stream) {
  if (int32_t e = check_run(rays, R, hs, tris, N, hit_idx, p, ws, ws_bytes)) return e;
  RunWs w = carve_run(ws, R);
  hipStream_t s = S(stream);
  if (R > 0 && hipMemcpyAsync(w.s1, rays, (size_t)R * sizeof(splat_adray_t), hipMemcpyDeviceToDevice, s) != hipSuccess)
    return SPLAT_ERR_HIP;
  if (int32_t e = run_core(rays, R, hs, tris, hit_idx, p, w, s)) return e;
  if (d_live_out &&
      hipMemcpyAsync(d_live_out, &w.counters[p->depth], sizeof(int64_t), hipMemcpyDeviceToDevice, s) != hipSuccess)
    return SPLAT_ERR_HIP;
  return SPLAT_OK;
}

int32_t splat_trace_frame(const splat_pinhole_t* cam, splat_adray_t* rays, int64_t R, const splat_hitset_t* hs,
                          const splat_ftri_t* tris, int32_t N, int32_t* hit_idx, const splat_run_params_t* p,
                          const splat_lightmap_params_t* lm, float* tex, void* ws, size_t ws_bytes, void* stream) {
  if (!cam || !tex) return SPLAT_ERR_INVALID_ARG;
  if (int32_t e = check_run(rays, R, hs, tris, N, hit_idx, p, ws, ws_bytes, true)) return e;
  if (int32_t e = check_lightmap(lm)) return e;
  if ((int64_t)cam->height * cam->width * cam->samples_per_pixel != R || !(cam->fov_half_sin > 0.f))
    return SPLAT_ERR_INVALID_ARG;
  RunWs w = carve_run(ws, R);
  hipStream_t s = S(stream);
  // (the light map is cleared by run_core) the first bounce generates its rays itself; only a
  // frame without bounces needs them stored
  if (p->depth == 0) launch_raygen(*cam, w.s1, R, s);
  // the light map is splatted inside the loop, as rays become DIFFUSE
  if (int32_t e = run_core(rays, R, hs, tris, hit_idx, p, w, s, lm, tex, p->depth > 0 ? cam : nullptr)) return e;
  return launch_status();
}

int32_t splat_trace_lights(const splat_rect_light_t* lights, int32_t n, const splat_light_rays_t* lp,
                           splat_adray_t* rays, int64_t R, const splat_hitset_t* hs, const splat_ftri_t* tris,
                           int32_t N, int32_t* hit_idx, const splat_run_params_t* p,
                           const splat_lightmap_params_t* lm, float* tex, void* ws, size_t ws_bytes, void* stream) {
  if (!lp || !tex) return SPLAT_ERR_INVALID_ARG;
  if (int32_t e = check_run(rays, R, hs, tris, N, hit_idx, p, ws, ws_bytes, true)) return e;
  if (int32_t e = check_lightmap(lm)) return e;
  const int64_t total = splat_lights_ray_count(lights, n, lp->upscale);
  if (total < 0) return (int32_t)total;
  if (total != R) return SPLAT_ERR_INVALID_ARG;
  RunWs w = carve_run(ws, R);
  hipStream_t s = S(stream);
  // the rays go where the bounce loop starts; the light map is cleared by run_core and splatted
  // inside the loop, as rays become DIFFUSE
  if (int32_t e = splat_raygen_lights(lights, n, lp, w.s1, R, stream)) return e;
  if (int32_t e = run_core(rays, R, hs, tris, hit_idx, p, w, s, lm, tex)) return e;
  return launch_status();
}

size_t splat_reorder_ws_bytes(int64_t R) {
  if (R < 0) return 0;
  return align_up((size_t)R * sizeof(splat_adray_t)) + align_up((size_t)(R > 0 ? R : 1) * sizeof(int32_t));
}

int32_t splat_reorder_by_dest(splat_adray_t* rays, int64_t R, int32_t* hit_idx, void* ws, size_t ws_bytes,
                              void* stream) {
  if (R < 0 || (R > 0 && (!rays || !hit_idx || !ws))) return SPLAT_ERR_INVALID_ARG;
  if (R == 0) return SPLAT_OK;
  if (ws_bytes < splat_reorder_ws_bytes(R)) return SPLAT_ERR_WORKSPACE;
  Carver c(ws);
  splat_adray_t* tr = c.take<splat_adray_t>((size_t)R);
  int32_t* th = c.take<int32_t>((size_t)R);
  hipStream_t s = S(stream);
  if (hipMemcpyAsync(tr, rays, (size_t)R * sizeof(splat_adray_t), hipMemcpyDeviceToDevice, s) != hipSuccess ||
      hipMemcpyAsync(th, hit_idx, (size_t)R * sizeof(int32_t), hipMemcpyDeviceToDevice, s) != hipSuccess)
    return SPLAT_ERR_HIP;
  launch_reorder(tr, th, R, rays, hit_idx, s);
  return launch_status();
}

int32_t splat_shade(splat_adray_t* rays, int64_t R, const int32_t* hit_idx, const splat_ftri_t* tris, int32_t N,
                    const splat_lightmap_params_t* lm, const float* tex, const float* sky, int32_t sky_w, int32_t sky_h,
                    int64_t grid_h, float* rgb, void* stream) {
  if (R < 0 || !tris || N < 1 || !lm || grid_h < 1) return SPLAT_ERR_INVALID_ARG;
  if (R == 0) return SPLAT_OK;
  if (!rays || !hit_idx || !rgb || R % grid_h != 0) return SPLAT_ERR_INVALID_ARG;
  if (lm->width < 1 || lm->height < 1 || lm->n_delta < 1 || lm->n_delta > SPLAT_MAX_DELTA || lm->n_lambda < 0 ||
      lm->n_lambda > SPLAT_MAX_LAMBDA)
    return SPLAT_ERR_INVALID_ARG;
  if (lm->n_lambda > 0 && !tex) return SPLAT_ERR_INVALID_ARG;
  if (sky && (sky_w < 1 || sky_h < 1)) return SPLAT_ERR_INVALID_ARG;
  launch_shade(rays, R, hit_idx, tris, N, *lm, tex, sky, sky ? sky_w : 0, sky ? sky_h : 0, grid_h, rgb, S(stream));
  return launch_status();
}

}  // extern "C"
