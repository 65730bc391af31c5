// bvbuild.hip -- the BV hitter's scene structure built on the device (the build step of
// bv_partition / pack_bv_tris, src/bv.jl:68-136, for the tree hitter's object-median leaves).
//
// Triangles t = 0..n-1 are the 1-based indices t + 2 (index 1, the degenerate sentinel, is never
// a member). The build, in the order the host builder (splat_amd/bv.py tree_partition) defines:
//  1. centroid ((a + b) + c) / 3 and bounding-box diagonal sqrt((dx^2 + dy^2) + dz^2) per
//     triangle, in FP64;
//  2. the large triangles: diagonal > big_factor x the median diagonal (median of the sorted
//     diagonals, the mean of the middle two for an even count), at most max_big of them, the
//     largest first, ties to the lower index; they form the leading member list;
//  3. the rest is split object-median `depth` times: a node's triangles are ordered by
//     (centroid on its split axis, index) and the first half goes left. The split axis of a node
//     of 2..SAH_MAX triangles has the least surface-area cost (area of each half's triangle
//     bounds x its count, summed; the lower axis on ties), of a larger node the widest centroid
//     extent. Each of three
//     lists (one per axis) stays sorted by (centroid on that axis, index) within every node, so
//     a node's widest axis is read off the ends of its segments, and its split is a stable
//     partition of every list, done for all nodes of a level at once with one exclusive scan.
//     A fourth list in index order gives each leaf's members sorted by index.
// Each step is a handful of launches over all triangles; the sorts and scans are rocPRIM's
// (hipCUB) device radix sort and scan.
#include <hipcub/hipcub.hpp>

#include "internal.hpp"

namespace spl {
namespace {

constexpr int BLK = 256;
constexpr int N_LISTS = 4;  // x, y, z centroid order; index order
constexpr int32_t SAH_MAX = 4096;

unsigned blocks(int64_t n) { return (unsigned)((n + BLK - 1) / BLK > 0 ? (n + BLK - 1) / BLK : 1); }

// 1. per-triangle centroid (+0.0 makes -0 and +0 one key) and diagonal
__global__ __launch_bounds__(BLK) void k_info(const splat_ftri_t* __restrict__ tris, int32_t n,
                                              double* __restrict__ cent, double* __restrict__ diag,
                                              float* __restrict__ bounds, int32_t* __restrict__ ids,
                                              uint8_t* __restrict__ big) {
  const int32_t t = blockIdx.x * BLK + threadIdx.x;
  if (t >= n) return;
  const splat_ftri_t& T = tris[t + 1];
  double sq = 0.0;
  for (int a = 0; a < 3; ++a) {
    const double p = T.v[1][a], q = T.v[2][a], r = T.v[3][a];
    cent[(int64_t)a * n + t] = ((p + q) + r) / 3.0 + 0.0;
    const double hi = fmax(fmax(p, q), r), lo = fmin(fmin(p, q), r);
    bounds[(int64_t)a * n + t] = (float)lo;  // the triangle's bounds: lo x, y, z then hi x, y, z
    bounds[(int64_t)(3 + a) * n + t] = (float)hi;
    const double d = hi - lo;
    sq = a == 0 ? d * d : sq + d * d;
  }
  diag[t] = __builtin_sqrt(sq);
  ids[t] = t;
  big[t] = 0;
}

// 2. the large triangles from the diagonals sorted in descending order (ties: lower index first)
__global__ __launch_bounds__(BLK) void k_big(const double* __restrict__ d_desc, const int32_t* __restrict__ id_desc,
                                             int32_t n, double big_factor, int32_t max_big, uint8_t* __restrict__ big,
                                             int32_t* __restrict__ summary, int32_t* __restrict__ seg0) {
  __shared__ int32_t s_nbig;
  if (threadIdx.x == 0) {
    int32_t nb = 0;
    if (n > 0) {
      const double med = (n & 1) ? d_desc[n / 2] : (d_desc[n / 2 - 1] + d_desc[n / 2]) / 2.0;
      if (med > 0.0) {
        const double thr = big_factor * med;
        int32_t lo = 0, hi = n;  // first k with d_desc[k] <= thr
        while (lo < hi) {
          const int32_t mid = (lo + hi) / 2;
          if (d_desc[mid] > thr) lo = mid + 1;
          else hi = mid;
        }
        nb = lo < max_big ? lo : max_big;
      }
    }
    s_nbig = nb;
    summary[0] = nb;
    summary[1] = n - nb;
    // the root segment of the split: [0, n - n_big)
    seg0[0] = 0;
    seg0[1] = n - nb;
  }
  __syncthreads();
  for (int32_t k = threadIdx.x; k < s_nbig; k += BLK) big[id_desc[k]] = 1;
}

__global__ __launch_bounds__(BLK) void k_axis_keys(const double* __restrict__ cent_a, int32_t n, double* __restrict__ keys,
                                                   int32_t* __restrict__ ids) {
  const int32_t t = blockIdx.x * BLK + threadIdx.x;
  if (t >= n) return;
  keys[t] = cent_a[t];
  ids[t] = t;
}

__global__ __launch_bounds__(BLK) void k_iota(int32_t* __restrict__ ids, int32_t n) {
  const int32_t t = blockIdx.x * BLK + threadIdx.x;
  if (t < n) ids[t] = t;
}

// compaction of the four sorted lists and the index list: keep[q] over the concatenation of
// lists 0..3 (non-large triangles) and list 4 (large triangles)
__global__ __launch_bounds__(BLK) void k_keep(const int32_t* __restrict__ raw, int32_t n, const uint8_t* __restrict__ big,
                                              int32_t* __restrict__ keep) {
  const int64_t q = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (q >= (int64_t)(N_LISTS + 1) * n) return;
  const int32_t b = (int32_t)(q / n);
  const uint8_t g = big[raw[q]];
  keep[q] = b < N_LISTS ? (g == 0) : (g != 0);
}

__global__ __launch_bounds__(BLK) void k_compact(const int32_t* __restrict__ raw, int32_t n, const int32_t* __restrict__ keep,
                                                 const int32_t* __restrict__ excl, int32_t* __restrict__ lists,
                                                 int32_t* __restrict__ big_list, int32_t* __restrict__ node_of) {
  const int64_t q = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (q >= (int64_t)(N_LISTS + 1) * n || !keep[q]) return;
  const int32_t b = (int32_t)(q / n);
  const int32_t pos = excl[q] - excl[(int64_t)b * n];
  if (b < N_LISTS) {
    lists[(int64_t)b * n + pos] = raw[q];
    if (b == 0) node_of[pos] = 0;
  } else {
    big_list[pos] = raw[q];
  }
}

// 3. one level of the split, one workgroup per node: its split axis and half size; the
// children's segments.
__device__ __forceinline__ double box_area(const float (&b)[6]) {
  double d[3];
  for (int c = 0; c < 3; ++c) {
    const double e = (double)b[3 + c] - (double)b[c];
    d[c] = e > 0.0 ? e : 0.0;
  }
  return 2.0 * ((d[0] * d[1] + d[1] * d[2]) + d[2] * d[0]);
}
__global__ __launch_bounds__(BLK) void k_nodes(const int32_t* __restrict__ seg, int32_t nodes, const int32_t* __restrict__ lists,
                                               int32_t n, const double* __restrict__ cent, const float* __restrict__ bounds,
                                               int32_t* __restrict__ axis, int32_t* __restrict__ half,
                                               int32_t* __restrict__ seg_next) {
  __shared__ float part[BLK / 64][12];
  const int32_t k = blockIdx.x;
  const int32_t s = seg[k], e = seg[k + 1], cnt = e - s, h = cnt / 2;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int32_t ax = 0;
  if (cnt > SAH_MAX) {  // widest centroid extent (each list is sorted within the node)
    double best = -1.0;
    for (int a = 0; a < 3; ++a) {
      const int32_t* L = lists + (int64_t)a * n;
      const double ext = cent[(int64_t)a * n + L[e - 1]] - cent[(int64_t)a * n + L[s]];
      if (a == 0 || ext > best) {
        best = ext;
        ax = a;
      }
    }
  } else if (cnt >= 2) {  // least surface-area cost of the median split
    double best = 0.0;
    for (int a = 0; a < 3; ++a) {
      const int32_t* L = lists + (int64_t)a * n;
      // bounds of the left half [s, s + h) and the right half [s + h, e) in this axis' order
      float b[12];
      for (int c = 0; c < 3; ++c) {
        b[c] = b[6 + c] = INFINITY;
        b[3 + c] = b[9 + c] = -INFINITY;
      }
      for (int32_t p = s + threadIdx.x; p < e; p += BLK) {
        const int32_t t = L[p];
        const int o = p < s + h ? 0 : 6;
        for (int c = 0; c < 3; ++c) {
          b[o + c] = fminf(b[o + c], bounds[(int64_t)c * n + t]);
          b[o + 3 + c] = fmaxf(b[o + 3 + c], bounds[(int64_t)(3 + c) * n + t]);
        }
      }
      for (int off = 32; off > 0; off >>= 1)
        for (int j = 0; j < 12; ++j) {
          const float o = __shfl_xor(b[j], off);
          b[j] = (j % 6) < 3 ? fminf(b[j], o) : fmaxf(b[j], o);
        }
      if (lane == 0)
        for (int j = 0; j < 12; ++j) part[wave][j] = b[j];
      __syncthreads();
      if (threadIdx.x == 0) {
        for (int w = 1; w < BLK / 64; ++w)
          for (int j = 0; j < 12; ++j) b[j] = (j % 6) < 3 ? fminf(b[j], part[w][j]) : fmaxf(b[j], part[w][j]);
        float bl[6], br[6];
        for (int j = 0; j < 6; ++j) {
          bl[j] = b[j];
          br[j] = b[6 + j];
        }
        const double cost = box_area(bl) * (double)h + box_area(br) * (double)(cnt - h);
        if (a == 0 || cost < best) {
          best = cost;
          ax = a;
        }
      }
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) {
    axis[k] = ax;
    half[k] = h;
    seg_next[2 * k] = s;
    seg_next[2 * k + 1] = s + h;
    if (k == nodes - 1) seg_next[2 * nodes] = e;
  }
}

// side of every triangle at this level: right (1) when it is past the half of its node's order
// on the node's axis
__global__ __launch_bounds__(BLK) void k_side(const int32_t* __restrict__ lists, int32_t n, const int32_t* __restrict__ summary,
                                              const int32_t* __restrict__ node_of, const int32_t* __restrict__ seg,
                                              const int32_t* __restrict__ axis, const int32_t* __restrict__ half,
                                              uint8_t* __restrict__ right) {
  const int32_t p = blockIdx.x * BLK + threadIdx.x;
  if (p >= summary[1]) return;
  const int32_t k = node_of[p];
  right[lists[(int64_t)axis[k] * n + p]] = (p - seg[k]) >= half[k];
}

__global__ __launch_bounds__(BLK) void k_left(const int32_t* __restrict__ lists, int32_t n, const int32_t* __restrict__ summary,
                                              const uint8_t* __restrict__ right, int32_t* __restrict__ is_left) {
  const int64_t q = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (q >= (int64_t)N_LISTS * n) return;
  const int32_t p = (int32_t)(q % n);
  is_left[q] = p < summary[1] ? (right[lists[q]] == 0) : 0;
}

// stable partition of every node's segment in every list: lefts first, then rights, each in order
__global__ __launch_bounds__(BLK) void k_split(const int32_t* __restrict__ lists, int32_t n, const int32_t* __restrict__ summary,
                                               const int32_t* __restrict__ node_of, const int32_t* __restrict__ seg,
                                               const int32_t* __restrict__ half, const int32_t* __restrict__ is_left,
                                               const int32_t* __restrict__ excl, int32_t* __restrict__ lists_next,
                                               int32_t* __restrict__ node_next) {
  const int64_t q = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (q >= (int64_t)N_LISTS * n) return;
  const int32_t b = (int32_t)(q / n), p = (int32_t)(q % n);
  if (p >= summary[1]) return;
  const int32_t k = node_of[p], s = seg[k];
  const int32_t lb = excl[q] - excl[(int64_t)b * n + s];  // lefts before p in the segment
  const bool left = is_left[q] != 0;
  const int32_t np = left ? s + lb : s + half[k] + ((p - s) - lb);
  lists_next[(int64_t)b * n + np] = lists[q];
  if (b == 0) node_next[np] = 2 * k + (left ? 0 : 1);
}

// 4. the hitset's member list (1-based indices) and per-BV ranges: BV 0 = the large triangles,
// BV 1 + k = leaf k
__global__ __launch_bounds__(BLK) void k_members(const int32_t* __restrict__ big_list, const int32_t* __restrict__ idx_list,
                                                 int32_t n, const int32_t* __restrict__ summary, int32_t* __restrict__ members) {
  const int32_t t = blockIdx.x * BLK + threadIdx.x;
  if (t >= n) return;
  const int32_t nb = summary[0];
  members[t] = (t < nb ? big_list[t] : idx_list[t - nb]) + 2;
}

__global__ __launch_bounds__(BLK) void k_ranges(const int32_t* __restrict__ seg, int32_t leaves, const int32_t* __restrict__ summary,
                                                int2* __restrict__ ranges) {
  const int32_t k = blockIdx.x * BLK + threadIdx.x;
  if (k > leaves) return;
  const int32_t nb = summary[0];
  if (k == 0) ranges[0] = make_int2(0, nb);
  else ranges[k] = make_int2(nb + seg[k - 1], seg[k] - seg[k - 1]);
}

struct Scratch {
  double *cent, *keys_a, *keys_b;
  float* bounds;
  int32_t *ids_a, *ids_b, *raw, *lists, *lists2, *big_list, *keep, *excl, *node_of, *node2, *seg, *seg2, *axis, *half,
      *summary;
  uint8_t *big, *right;
  void* cub;
  size_t cub_bytes;
  size_t bytes;
};

size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

Scratch carve(void* base, int32_t n, int32_t depth) {
  Scratch S{};
  const size_t nn = (size_t)(n > 0 ? n : 1);
  const size_t leaves = (size_t)1 << depth;
  size_t sort_bytes = 0, scan_bytes = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (const double*)nullptr, (double*)nullptr,
                                           (const int32_t*)nullptr, (int32_t*)nullptr, (int)nn);
  size_t desc_bytes = 0;
  (void)hipcub::DeviceRadixSort::SortPairsDescending(nullptr, desc_bytes, (const double*)nullptr, (double*)nullptr,
                                                     (const int32_t*)nullptr, (int32_t*)nullptr, (int)nn);
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (const int32_t*)nullptr, (int32_t*)nullptr,
                                         (int)((N_LISTS + 1) * nn));
  S.cub_bytes = std::max(std::max(sort_bytes, desc_bytes), scan_bytes);
  char* p = static_cast<char*>(base);
  size_t used = 0;
  auto take = [&](size_t bytes) -> void* {
    void* r = p ? p + used : nullptr;
    used += up256(bytes);
    return r;
  };
  S.cent = (double*)take(3 * nn * sizeof(double));
  S.bounds = (float*)take(6 * nn * sizeof(float));
  S.keys_a = (double*)take(nn * sizeof(double));
  S.keys_b = (double*)take(nn * sizeof(double));
  S.ids_a = (int32_t*)take(nn * sizeof(int32_t));
  S.ids_b = (int32_t*)take(nn * sizeof(int32_t));
  S.raw = (int32_t*)take((N_LISTS + 1) * nn * sizeof(int32_t));
  S.lists = (int32_t*)take(N_LISTS * nn * sizeof(int32_t));
  S.lists2 = (int32_t*)take(N_LISTS * nn * sizeof(int32_t));
  S.big_list = (int32_t*)take(nn * sizeof(int32_t));
  S.keep = (int32_t*)take((N_LISTS + 1) * nn * sizeof(int32_t));
  S.excl = (int32_t*)take((N_LISTS + 1) * nn * sizeof(int32_t));
  S.node_of = (int32_t*)take(nn * sizeof(int32_t));
  S.node2 = (int32_t*)take(nn * sizeof(int32_t));
  S.seg = (int32_t*)take((leaves + 1) * sizeof(int32_t));
  S.seg2 = (int32_t*)take((leaves + 1) * sizeof(int32_t));
  S.axis = (int32_t*)take(leaves * sizeof(int32_t));
  S.half = (int32_t*)take(leaves * sizeof(int32_t));
  S.summary = (int32_t*)take(4 * sizeof(int32_t));
  S.big = (uint8_t*)take(nn);
  S.right = (uint8_t*)take(nn);
  S.cub = take(S.cub_bytes);
  S.bytes = used;
  return S;
}

}  // namespace

size_t bv_build_scratch_bytes(int32_t n, int32_t depth) { return carve(nullptr, n, depth).bytes; }

int32_t bv_build_members(const splat_ftri_t* tris, int32_t n, int32_t depth, double big_factor, int32_t max_big,
                         void* scratch, int32_t* members, int2* ranges, int32_t* h_summary, hipStream_t s) {
  Scratch S = carve(scratch, n, depth);
  size_t cb = S.cub_bytes;
  hipLaunchKernelGGL(k_info, dim3(blocks(n)), dim3(BLK), 0, s, tris, n, S.cent, S.keys_a, S.bounds, S.ids_a, S.big);
  if (hipcub::DeviceRadixSort::SortPairsDescending(S.cub, cb, S.keys_a, S.keys_b, S.ids_a, S.ids_b, n, 0, 64, s) !=
      hipSuccess)
    return SPLAT_ERR_HIP;
  hipLaunchKernelGGL(k_big, dim3(1), dim3(BLK), 0, s, S.keys_b, S.ids_b, n, big_factor, max_big, S.big, S.summary,
                     S.seg);
  for (int a = 0; a < 3; ++a) {
    hipLaunchKernelGGL(k_axis_keys, dim3(blocks(n)), dim3(BLK), 0, s, S.cent + (int64_t)a * n, n, S.keys_a, S.ids_a);
    cb = S.cub_bytes;
    if (hipcub::DeviceRadixSort::SortPairs(S.cub, cb, S.keys_a, S.keys_b, S.ids_a, S.raw + (int64_t)a * n, n, 0, 64,
                                           s) != hipSuccess)
      return SPLAT_ERR_HIP;
  }
  hipLaunchKernelGGL(k_iota, dim3(blocks(n)), dim3(BLK), 0, s, S.raw + (int64_t)3 * n, n);
  hipLaunchKernelGGL(k_iota, dim3(blocks(n)), dim3(BLK), 0, s, S.raw + (int64_t)4 * n, n);
  const int64_t n5 = (int64_t)(N_LISTS + 1) * n, n4 = (int64_t)N_LISTS * n;
  hipLaunchKernelGGL(k_keep, dim3(blocks(n5)), dim3(BLK), 0, s, S.raw, n, S.big, S.keep);
  cb = S.cub_bytes;
  if (hipcub::DeviceScan::ExclusiveSum(S.cub, cb, S.keep, S.excl, (int)n5, s) != hipSuccess) return SPLAT_ERR_HIP;
  hipLaunchKernelGGL(k_compact, dim3(blocks(n5)), dim3(BLK), 0, s, S.raw, n, S.keep, S.excl, S.lists, S.big_list,
                     S.node_of);
  int32_t *lists = S.lists, *lists2 = S.lists2, *node_of = S.node_of, *node2 = S.node2, *seg = S.seg, *seg2 = S.seg2;
  for (int l = 0; l < depth; ++l) {
    const int32_t nodes = 1 << l;
    hipLaunchKernelGGL(k_nodes, dim3((unsigned)nodes), dim3(BLK), 0, s, seg, nodes, lists, n, S.cent, S.bounds, S.axis,
                       S.half, seg2);
    hipLaunchKernelGGL(k_side, dim3(blocks(n)), dim3(BLK), 0, s, lists, n, S.summary, node_of, seg, S.axis, S.half,
                       S.right);
    hipLaunchKernelGGL(k_left, dim3(blocks(n4)), dim3(BLK), 0, s, lists, n, S.summary, S.right, S.keep);
    cb = S.cub_bytes;
    if (hipcub::DeviceScan::ExclusiveSum(S.cub, cb, S.keep, S.excl, (int)n4, s) != hipSuccess) return SPLAT_ERR_HIP;
    hipLaunchKernelGGL(k_split, dim3(blocks(n4)), dim3(BLK), 0, s, lists, n, S.summary, node_of, seg, S.half, S.keep,
                       S.excl, lists2, node2);
    std::swap(lists, lists2);
    std::swap(node_of, node2);
    std::swap(seg, seg2);
  }
  const int32_t leaves = 1 << depth;
  hipLaunchKernelGGL(k_members, dim3(blocks(n)), dim3(BLK), 0, s, S.big_list, lists + (int64_t)3 * n, n, S.summary,
                     members);
  hipLaunchKernelGGL(k_ranges, dim3(blocks(leaves + 1)), dim3(BLK), 0, s, seg, leaves, S.summary, ranges);
  if (hipMemcpyAsync(h_summary, S.summary, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess)
    return SPLAT_ERR_HIP;
  return hipGetLastError() == hipSuccess ? SPLAT_OK : SPLAT_ERR_HIP;
}

}  // namespace spl
