// adsr_device.hip -- gfx950 kernels and launches for
// PeakGroupSelection.search (src/peak_group_selection/mod.rs:87-195).
//
// Pipeline (one stream; captured in a hipGraph for repeated searches):
//   plan    one thread per precursor: RT cycle window
//           (src/rt_index/mod.rs:89-138), valid quadrupole windows
//           (src/dia_data/mod.rs:79-87) and the fragment filter
//           (src/speclib_flat/mod.rs:240-345).
//   bins    one thread per (precursor, selected fragment): the 1-ppm bins of
//           [mz - d, mz + d] (src/mz_index/mod.rs:95-127) via the f32-bit
//           bucket table, then the CSR entry range of every valid window.
//   select  one wavefront per precursor, persistent over an atomic queue:
//           coalesced scan of the CSR entries into an LDS-resident XIC tile
//           (src/quadrupole_observation/mod.rs:70-107), sparse 21-tap
//           convolution on the compacted active cycles
//           (src/convolution/scalar.rs:5-56), exact ln and ordered weighted
//           accumulation (src/score/scalar.rs:4-25), strict 5-point maxima
//           and top-k (src/peak_group_selection/mod.rs:23-64,167-192).
//   compact exclusive scan of per-precursor counts, order-preserving
//           scatter of the candidate rows (library order, as the rayon
//           collect at mod.rs:101-119).
//
// Every f32 operation follows the reference's evaluation order with one
// rounding per operation: the library is built with -ffp-contract=off.
// Skipping work is only done where the skipped terms are exact zeros.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "adsr_internal.hpp"
#include "adsr_logf.hpp"

namespace adsr {

static const LogfEntry kHostLogfTab[16] = ADSR_LOGF_TABLE;
__constant__ LogfEntry kDevLogfTab[16] = ADSR_LOGF_TABLE;

constexpr int kWave = 64;
constexpr int kMaxSel = 64;      // selected fragments per precursor
constexpr int kScoreMaxSel = 256;  // the same for scoring and quantification
constexpr int kObsMax = 4;       // valid windows per precursor
constexpr int kHalfMax = 15;     // kernel half width (kernel_size <= 30)
constexpr int kLCapMax = 4096;   // cycles per precursor window
constexpr int kStashFrag = 48;   // per-thread LDS stash in the plan kernel
constexpr size_t kMaxDynLds = 160 * 1024 - 512;  // bytes per block
constexpr uint32_t kLdsBudget = 10 * 1024;       // preferred bytes per wave

enum StatusBits : uint32_t {
  kStTooManyFragments = 1u,
  kStTooManyWindows = 2u,
  kStWindowTooLong = 4u,
  kStSelectFallback = 8u,  // top-k selection needs median of medians
};

#define HIP_TRY(expr)                                                   \
  do {                                                                  \
    hipError_t _e = (expr);                                             \
    if (_e != hipSuccess)                                               \
      return set_error(_e == hipErrorOutOfMemory ? ADSR_ERR_OUT_OF_MEMORY \
                                                 : ADSR_ERR_DEVICE,     \
                       "%s failed: %s", #expr, hipGetErrorString(_e));  \
  } while (0)

// Every device allocation of the library: hipMalloc, or out of memory for
// the allocation ADSR_FAIL_ALLOC names.
template <class T>
static hipError_t dev_malloc(T **p, size_t bytes) {
  if (alloc_should_fail()) {
    *p = nullptr;
    return hipErrorOutOfMemory;
  }
  return hipMalloc(reinterpret_cast<void **>(p), bytes);
}

static void *host_malloc(size_t bytes) {
  return alloc_should_fail() ? nullptr : std::malloc(bytes);
}

// ---------------------------------------------------------------------------
// kernel argument blocks
// ---------------------------------------------------------------------------
struct KDia {
  const float *grid;
  const uint32_t *bucket;
  uint32_t n_grid, base_key;
  const float *rt;
  uint32_t n_rt;
  float delta_t;
  uint32_t rt_strict;  // RT axis strictly increasing
  float rt_inv_dt;     // 1 / mean spacing when rt_strict (else 0)
  uint32_t n_obs;
  const float *iso_lo, *iso_hi;
  const uint32_t *ss;  // n_obs * (n_grid + 1)
  const uint64_t *entry_base;
  const uint16_t *cycles;
  const float *intens;
  const uint4 *bucket2;  // (count below, grid[count], grid[count + 1], -)
  // per bucket: count below (22 bits) | grid points in it (2) << 22 |
  // ulp offset of the first << 24 | of the second << 28
  const uint32_t *bucket1;
  float grid_first, grid_last;
};

struct KLib {
  uint32_t n_prec;
  const float *p_mz, *p_rt;
  const uint64_t *f_start, *f_stop;
  const float *f_mz, *f_int;
  const uint8_t *f_type, *f_num;
};

struct KParams {
  float mass_tol;
  float rt_tol;
  const float *rt_tol_pp;  // nullable
  float pad_cycles;        // (kernel_size as f32)
  uint32_t top_k;
  uint32_t cand_count;
  uint32_t half;
  uint32_t klen;
  uint32_t per_window;
  const float *kernels;  // (1 + n_obs*per_window) * klen
};

struct KPlan {
  uint4 *rec;         // {c0, c1, nsel | nobs << 16, kernel index}
  uint16_t *obs;      // [row][kObsMax]
  uint16_t *sel;      // [row][KS] local fragment index, output order
  float *sel_mz;      // [row][KS]
  float *sel_w;       // [row][KS]
  uint2 *rng;         // [row][kObsMax][KS] CSR entry range, per window
  uint32_t *b0;       // [row][kObsMax][KS] first bin
  uint32_t ks;
};

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
// Rust 1.88 slice::binary_search_by over f32 (src/rt_index/mod.rs:254-265):
// base moves to mid whenever v[mid] <= key; Ok(base) on equality, else the
// insertion point.
__device__ __forceinline__ uint32_t rust_bsearch_f32(const float *v,
                                                     uint32_t n, float key) {
  if (n == 0) return 0;
  uint32_t size = n, base = 0;
  while (size > 1) {
    uint32_t half = size >> 1, mid = base + half;
    if (!(v[mid] > key)) base = mid;
    size -= half;
  }
  const float b = v[base];
  if (b == key) return base;
  return base + (b < key ? 1u : 0u);
}

// rust_bsearch_f32 over the RT axis.  On a strictly increasing axis every
// value is unique, so the binary search's result is the lower bound of the
// key, found here by a walk from the index g the mean spacing predicts.
// rt[g - 1], rt[g] and rt[g + 1] are read at once, which settles an evenly
// spaced axis in one round trip; the walk goes on only past them.
__device__ __forceinline__ uint32_t rt_search(const float *rt, uint32_t n, uint32_t strict,
                                              float inv_dt, float key) {
  if (!strict || n == 0 || key != key) return rust_bsearch_f32(rt, n, key);
  const float gf = fminf(fmaxf((key - rt[0]) * inv_dt, 0.0f), (float)(n - 1));
  uint32_t g = (uint32_t)gf;  // g < n
  const float vm = g > 0 ? rt[g - 1] : 0.0f;
  const float v0 = rt[g];
  const float vp = g + 1 < n ? rt[g + 1] : 0.0f;
  if (g > 0 && vm >= key) {
    // rt[g - 1] >= key: walk down; rt[g] >= key holds at every stop
    --g;
    while (g > 0 && rt[g - 1] >= key) --g;
    return g;
  }
  if (!(v0 < key)) return g;
  ++g;
  if (!(g < n && vp < key)) return g;
  ++g;
  while (g < n && rt[g] < key) ++g;
  return g;
}

__device__ __forceinline__ uint32_t grid_lb(const KDia &d, float x) {
  const float *m = d.grid;
  const uint32_t n = d.n_grid;
  if (!(x > m[0])) return 0;
  if (x > m[n - 1]) return n;
  uint32_t b = d.bucket[(__float_as_uint(x) >> kBucketShift) - d.base_key];
  while (b < n && m[b] < x) ++b;
  return b;
}

__device__ __forceinline__ uint32_t grid_ub(const KDia &d, float x) {
  const float *m = d.grid;
  const uint32_t n = d.n_grid;
  if (!(x >= m[0])) return 0;
  if (x >= m[n - 1]) return n;
  uint32_t b = d.bucket[(__float_as_uint(x) >> kBucketShift) - d.base_key];
  while (b < n && m[b] <= x) ++b;
  return b;
}

// Bin bounds from one bucket2 entry.  A bucket spans 16 f32 ulps and a
// 1-ppm bin at least 8.39, so at most two grid points lie in a bucket and
// the entry's two values settle the bound exactly; a value past the bucket
// compares false.  Valid for grid_first < x <= grid_last (lower) and
// grid_first <= x < grid_last (upper).
__device__ __forceinline__ uint32_t bucket_key(const KDia &d, float x) {
  return (__float_as_uint(x) >> kBucketShift) - d.base_key;
}
__device__ __forceinline__ uint32_t lb_from(const uint4 &e, float x) {
  return e.x + (__uint_as_float(e.y) < x ? 1u : 0u) +
         (__uint_as_float(e.z) < x ? 1u : 0u);
}
__device__ __forceinline__ uint32_t ub_from(const uint4 &e, float x) {
  return e.x + (__uint_as_float(e.y) <= x ? 1u : 0u) +
         (__uint_as_float(e.z) <= x ? 1u : 0u);
}

// src/mz_index/mod.rs:95-127 -> [start, end)
__device__ __forceinline__ void mz_range(const KDia &d, float lo, float hi,
                                         uint32_t &start, uint32_t &end) {
  start = grid_lb(d, lo);
  end = start;
  if (start < d.n_grid) end = max(start, grid_ub(d, hi));
}

__device__ __forceinline__ uint64_t ballot64(bool p) {
  return __ballot(p);
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

// Two u16 lanes of one VGPR, for packed integer arithmetic.
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
static_assert(kLCapMax < 65536, "window lengths fit 16 bits");

// The lane index as a value the compiler cannot move.  A per-lane address
// built from it stays at its use, as a uniform 64-bit base in SGPRs plus a
// 32-bit lane offset, instead of being hoisted out of a row loop as a 64-bit
// per-lane pointer that the register allocator then spills to scratch and
// reloads once per row.
__device__ __forceinline__ uint32_t opaque_lane() {
  uint32_t v = threadIdx.x & (kWave - 1);
  asm volatile("" : "+v"(v));
  return v;
}

// inclusive wave scan of u32
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    uint32_t t = __shfl_up(v, o, kWave);
    if (lane >= o) v += t;
  }
  return v;
}

// ---------------------------------------------------------------------------
// plan: one thread per precursor row
// ---------------------------------------------------------------------------
// Rust 1.88 select_nth_unstable_by(k, |a, b| b.I.partial_cmp(&a.I)) for
// 0 < k < len - 1 (src/speclib_flat/mod.rs:286-292): core::slice::select's
// partition_at_index_loop, sequential, over a view V of fragment indices
// (get/set by position, less(a, b) = I[b] < I[a]).  Slices of at most 16
// elements are insertion-sorted; longer ones are partitioned around a
// median-of-3 pivot (recursive for 64 and more elements) by the branchless
// cyclic Lomuto partition of the unstable quicksort.  Returns false when the
// 16 partition rounds run out (the reference then switches to median of
// medians, which is not restated here) or for 4096 or more elements; the
// caller reports it.
template <class V>
__device__ void rs_insertion_sort(const V &v, uint32_t o, uint32_t n) {
  for (uint32_t i = 1; i < n; ++i) {
    const uint32_t tmp = v.get(o + i);
    uint32_t j = i;
    while (j > 0 && v.less(tmp, v.get(o + j - 1))) {
      v.set(o + j, v.get(o + j - 1));
      --j;
    }
    v.set(o + j, tmp);
  }
}

template <class V>
__device__ uint32_t rs_median3(const V &v, uint32_t a, uint32_t b, uint32_t c) {
  const bool x = v.less(v.get(a), v.get(b));
  const bool y = v.less(v.get(a), v.get(c));
  if (x == y) return (v.less(v.get(b), v.get(c)) != x) ? c : b;
  return a;
}

// pivot::choose_pivot for 16 < len < 4096: median3, or median3_rec with its
// (at most two) levels of recursion written out; len = n * 8 + r
template <class V>
__device__ uint32_t rs_choose_pivot(const V &v, uint32_t o, uint32_t len) {
  const uint32_t n = len / 8;
  const uint32_t a = o, b = o + n * 4, c = o + n * 7;
  if (len < 64) return rs_median3(v, a, b, c);
  const uint32_t n8 = n / 8, n64 = n8 / 8;
  uint32_t m[3];
  const uint32_t x3[3] = {a, b, c};
  for (int i = 0; i < 3; ++i) {
    uint32_t p = x3[i], q = x3[i] + n8 * 4, r = x3[i] + n8 * 7;
    if (n8 * 8 >= 64) {
      p = rs_median3(v, p, p + n64 * 4, p + n64 * 7);
      q = rs_median3(v, q, q + n64 * 4, q + n64 * 7);
      r = rs_median3(v, r, r + n64 * 4, r + n64 * 7);
    }
    m[i] = rs_median3(v, p, q, r);
  }
  return rs_median3(v, m[0], m[1], m[2]);
}

// partition of v[o, o + len) around the element at o + pivot; `le` selects
// the comparison !less(pivot, e) used for elements equal to an ancestor
// pivot.  Returns the number of elements placed before the pivot.
template <class V>
__device__ uint32_t rs_partition(const V &v, uint32_t o, uint32_t len, uint32_t pivot,
                                 bool le) {
  uint32_t t = v.get(o);
  v.set(o, v.get(o + pivot));
  v.set(o + pivot, t);
  const uint32_t pv = v.get(o);
  const uint32_t r = o + 1, n = len - 1;
  uint32_t num_lt = 0;
  if (n > 0) {
    const uint32_t gap_value = v.get(r);
    uint32_t gap = 0;
    for (uint32_t right = 1; right <= n; ++right) {
      const uint32_t e = right == n ? gap_value : v.get(r + right);
      const bool lt = le ? !v.less(pv, e) : v.less(e, pv);
      v.set(r + gap, v.get(r + num_lt));
      v.set(r + num_lt, e);
      gap = right;
      num_lt += lt ? 1u : 0u;
    }
  }
  t = v.get(o);
  v.set(o, v.get(o + num_lt));
  v.set(o + num_lt, t);
  return num_lt;
}

template <class V>
__device__ bool rs_select(const V &v, uint32_t len, uint32_t index) {
  if (len >= 4096) return false;  // median3_rec deeper than two levels
  uint32_t o = 0;
  int limit = 16;
  bool have_anc = false;
  uint32_t anc = 0;
  for (;;) {
    if (len <= 16) {
      if (len >= 2) rs_insertion_sort(v, o, len);
      return true;
    }
    if (limit == 0) return false;
    --limit;
    const uint32_t pivot = rs_choose_pivot(v, o, len);
    if (have_anc && !v.less(anc, v.get(pivot))) {
      const uint32_t mid = rs_partition(v, o, len, pivot - o, true) + 1;
      if (mid > index) return true;
      o += mid;
      len -= mid;
      index -= mid;
      have_anc = false;
      continue;
    }
    const uint32_t mid = rs_partition(v, o, len, pivot - o, false);
    if (mid == index) return true;
    if (mid < index) {
      anc = v.get(o + mid);
      have_anc = true;
      o += mid + 1;
      len -= mid + 1;
      index -= mid + 1;
    } else {
      len = mid;
    }
  }
}

// Fragment filter and order of src/speclib_flat/mod.rs:240-345 with the
// Rust 1.88 select_nth_unstable_by behaviour:
//   k >= kept        all kept fragments, kept order
//   k == 0           none
//   k == kept - 1    drop the first fragment of minimum intensity; the last
//                    kept fragment takes its place (max_index + swap)
//   otherwise        rs_select over the kept fragments in order, first k
//                    (for at most 16 kept: the stable order by intensity)
// followed by the stable sort by fragment m/z.  Accessor A gives the
// intensity of fragment j (-inf when the fragment is filtered out) and mz.
template <class A>
struct KeyView {  // positions in A's key storage
  const A &acc;
  __device__ uint32_t get(uint32_t p) const { return acc.key(p); }
  __device__ void set(uint32_t p, uint32_t j) const { acc.set_key(p, j); }
  __device__ bool less(uint32_t a, uint32_t b) const { return acc.I(b) < acc.I(a); }
};

template <class A>
__device__ uint32_t select_fragments(const A &acc, uint32_t n, uint32_t top_k,
                                     uint16_t *sel, float *sel_mz,
                                     float *sel_w, const float *f_int,
                                     uint32_t ks, uint32_t &status) {
  uint32_t kept = 0;
  for (uint32_t j = 0; j < n; ++j) kept += acc.I(j) > 0.0f ? 1u : 0u;
  const uint32_t k = min(top_k, kept);
  if (k == 0) return 0;
  if (k > ks) {  // ks <= kMaxSel for a search, kScoreMaxSel for scoring
    status |= kStTooManyFragments;
    return 0;
  }
  const int mode = (k >= kept) ? 0 : (k == kept - 1 ? 1 : 2);
  if (mode == 2) {
    // the key storage holds the kept fragments' indices by position
    uint32_t q = 0;
    for (uint32_t j = 0; j < n; ++j)
      if (acc.I(j) > 0.0f) acc.set_key(q++, j);
    if (!rs_select(KeyView<A>{acc}, kept, k)) {
      status |= kStSelectFallback;
      return 0;
    }
    for (uint32_t p = 0; p < k; ++p) {
      const uint32_t j = acc.key(p);
      const float mj = acc.mz(j);
      uint32_t pos = 0;
      for (uint32_t p2 = 0; p2 < k; ++p2) {
        if (p2 == p) continue;
        const float mi = acc.mz(acc.key(p2));
        if (mi < mj || (!(mi > mj) && !(mi < mj) && p2 < p)) ++pos;
      }
      sel[pos] = (uint16_t)j;
      sel_mz[pos] = mj;
      sel_w[pos] = f_int[j];
    }
    return k;
  }
  uint32_t drop = 0xffffffffu, last = 0, drop_pos = 0;
  if (mode == 1) {
    float mn = 0.0f;
    uint32_t pos = 0;
    for (uint32_t j = 0; j < n; ++j) {
      const float v = acc.I(j);
      if (!(v > 0.0f)) continue;
      if (drop == 0xffffffffu || v < mn) {
        mn = v;
        drop = j;
        drop_pos = pos;
      }
      last = j;
      ++pos;
    }
  }
  // pre-sort key of every fragment (its position before the m/z sort);
  // 0xffffffff marks fragments that are not selected
  uint32_t pos_kept = 0;
  for (uint32_t j = 0; j < n; ++j) {
    const float v = acc.I(j);
    uint32_t key = 0xffffffffu;
    if (v > 0.0f) {
      if (mode == 1) {
        key = (j == drop) ? 0xffffffffu
                          : ((j == last && last != drop) ? drop_pos : pos_kept);
      } else {
        key = pos_kept;
      }
      ++pos_kept;
    }
    acc.set_key(j, key);
  }
  for (uint32_t j = 0; j < n; ++j) {
    const uint32_t kj = acc.key(j);
    if (kj == 0xffffffffu) continue;
    const float mj = acc.mz(j);
    uint32_t pos = 0;
    for (uint32_t i = 0; i < n; ++i) {
      const uint32_t ki = acc.key(i);
      if (i == j || ki == 0xffffffffu) continue;
      const float mi = acc.mz(i);
      if (mi < mj || (!(mi > mj) && !(mi < mj) && ki < kj)) ++pos;
    }
    sel[pos] = (uint16_t)j;
    sel_mz[pos] = mj;
    sel_w[pos] = f_int[j];
  }
  return k;
}

// Fragments beyond the LDS stash: the same filter evaluated from global
// memory, with the pre-sort keys in a per-row global scratch.
struct GlobalFragAcc {
  const float *mz_, *int_;
  const uint8_t *type_, *num_;
  uint32_t *key_;
  __device__ float I(uint32_t j) const {
    const float v = int_[j];
    const bool keep = v > 0.0f && !(type_[j] == 121 && num_[j] == 1);
    return keep ? v : -__builtin_inff();
  }
  __device__ float mz(uint32_t j) const { return mz_[j]; }
  __device__ uint32_t key(uint32_t j) const { return key_[j]; }
  __device__ void set_key(uint32_t j, uint32_t k) const { key_[j] = k; }
};

struct StashFragAcc {
  const float *mz_, *I_;  // LDS, stride kWave (one column per lane)
  uint32_t *key_;
  __device__ float I(uint32_t j) const { return I_[j * kWave]; }
  __device__ float mz(uint32_t j) const { return mz_[j * kWave]; }
  __device__ uint32_t key(uint32_t j) const { return key_[j * kWave]; }
  __device__ void set_key(uint32_t j, uint32_t k) const { key_[j * kWave] = k; }
};

// STASH: fragments per thread kept in LDS (16, 32 or kStashFrag), chosen
// from the library's longest fragment list so that short lists do not pay
// for LDS they never use (occupancy).
template <int STASH>
__global__ __launch_bounds__(kWave) void plan_kernel(KDia dia, KLib lib,
                                                     KParams prm, KPlan plan,
                                                     uint32_t *key_scratch,
                                                     uint32_t *status) {
  __shared__ float st_mz[STASH * kWave];
  __shared__ float st_I[STASH * kWave];
  __shared__ uint32_t st_key[STASH * kWave];
  const uint32_t row = blockIdx.x * kWave + threadIdx.x;
  if (row >= lib.n_prec) return;
  uint32_t st = 0;
  // src/peak_group_selection/mod.rs:139-148 -> rt_index/mod.rs:89-138
  const float rt_tol = prm.rt_tol_pp ? prm.rt_tol_pp[row] : prm.rt_tol;
  const float prt = lib.p_rt[row];
  uint32_t c0 = 0, c1 = 0;
  const uint32_t n = dia.n_rt;
  if (n > 0) {
    const float min_tol = (10.0f + prm.pad_cycles) * dia.delta_t;
    const float eff = fmaxf(rt_tol, min_tol);
    const float lo = prt - eff, hi = prt + eff;
    if (hi < dia.rt[0]) {
      c0 = c1 = 0;
    } else if (lo > dia.rt[n - 1]) {
      c0 = c1 = n;
    } else {
      c0 = (lo <= dia.rt[0]) ? 0u : rt_search(dia.rt, n, dia.rt_strict, dia.rt_inv_dt, lo);
      c1 = (hi >= dia.rt[n - 1]) ? n : rt_search(dia.rt, n, dia.rt_strict, dia.rt_inv_dt, hi);
    }
  }
  // src/dia_data/mod.rs:79-87 valid observations (inclusive, ascending)
  const float pmz = lib.p_mz[row];
  uint32_t nobs = 0, first = 0xffffffffu;
  for (uint32_t o = 0; o < dia.n_obs; ++o) {
    if (dia.iso_lo[o] <= pmz && dia.iso_hi[o] >= pmz) {
      if (nobs < (uint32_t)kObsMax) plan.obs[(size_t)row * kObsMax + nobs] = (uint16_t)o;
      if (first == 0xffffffffu) first = o;
      ++nobs;
    }
  }
  if (nobs > (uint32_t)kObsMax) st |= kStTooManyWindows;
  // src/speclib_flat/mod.rs:388-456 get_precursor_filtered(i, true, true, k)
  const uint64_t fs = lib.f_start[row];
  const uint32_t nraw = (uint32_t)(lib.f_stop[row] - fs);
  uint16_t *sel = plan.sel + (size_t)row * plan.ks;
  float *smz = plan.sel_mz + (size_t)row * plan.ks;
  float *sw = plan.sel_w + (size_t)row * plan.ks;
  uint32_t nsel;
  if (nraw <= (uint32_t)STASH) {
    for (uint32_t j = 0; j < nraw; ++j) {
      const float v = lib.f_int[fs + j];
      const bool keep =
          v > 0.0f && !(lib.f_type[fs + j] == 121 && lib.f_num[fs + j] == 1);
      st_I[j * kWave + threadIdx.x] = keep ? v : -__builtin_inff();
      st_mz[j * kWave + threadIdx.x] = lib.f_mz[fs + j];
    }
    StashFragAcc acc{st_mz + threadIdx.x, st_I + threadIdx.x,
                     st_key + threadIdx.x};
    nsel = select_fragments(acc, nraw, prm.top_k, sel, smz, sw, lib.f_int + fs,
                            plan.ks, st);
  } else {
    GlobalFragAcc acc{lib.f_mz + fs, lib.f_int + fs, lib.f_type + fs,
                      lib.f_num + fs, key_scratch + fs};
    nsel = select_fragments(acc, nraw, prm.top_k, sel, smz, sw, lib.f_int + fs,
                            plan.ks, st);
  }
  const uint32_t kidx =
      (prm.per_window && first != 0xffffffffu) ? 1u + first : 0u;
  plan.rec[row] = make_uint4(c0, c1, nsel | (min(nobs, (uint32_t)kObsMax) << 16),
                             kidx);
  if (c1 > c0 && c1 - c0 > (uint32_t)kLCapMax) st |= kStWindowTooLong;
  if (st) atomicOr(status, st);
}

// ---------------------------------------------------------------------------
// plan with LP lanes per precursor row (LP = 16, 32 or 64; every raw
// fragment list of the library has at most LP entries): the same RT
// window, valid windows and fragment filter/order as plan_kernel, one
// fragment per lane, so the O(n^2) orderings cost O(n) per lane.
// ---------------------------------------------------------------------------
constexpr int kPlanThreads = 256;  // 4 waves per block: more waves per CU than 1-wave blocks

// rs_select's view of a group's lanes: positions in g_key, intensities g_I
struct GroupView {
  uint32_t *key;
  const float *I;
  __device__ uint32_t get(uint32_t p) const { return key[p]; }
  __device__ void set(uint32_t p, uint32_t j) const { key[p] = j; }
  __device__ bool less(uint32_t a, uint32_t b) const { return I[b] < I[a]; }
};

template <int LP>
__global__ __launch_bounds__(kPlanThreads) void plan_group_kernel(KDia dia, KLib lib,
                                                                  KParams prm, KPlan plan,
                                                                  uint32_t *status) {
  __shared__ float g_mz_b[kPlanThreads], g_I_b[kPlanThreads];
  __shared__ uint32_t g_key_b[kPlanThreads];
  const int lane = lane_id();
  const int wv = (int)threadIdx.x / kWave;
  // this wave's slice of the block's LDS, indexed by lane
  float *const g_mz = g_mz_b + wv * kWave, *const g_I = g_I_b + wv * kWave;
  uint32_t *const g_key = g_key_b + wv * kWave;
  const int gl = lane % LP;           // lane within the row's group
  const int gb = lane - gl;           // first lane of the group
  const uint64_t gmask = (LP == 64) ? ~0ull : (((1ull << LP) - 1) << gb);
  const uint32_t row = (blockIdx.x * (kPlanThreads / kWave) + (uint32_t)wv) * (kWave / LP) +
                       (uint32_t)(lane / LP);
  const bool live = row < lib.n_prec;
  uint32_t st = 0;
  // The row's independent loads are issued first, so that their latencies
  // overlap: RT, tolerance, m/z, fragment range, the first isolation windows
  // of the group's lanes, then the fragments.
  const float rt_tol = live && prm.rt_tol_pp ? prm.rt_tol_pp[row] : prm.rt_tol;
  const float prt = live ? lib.p_rt[row] : 0.0f;
  const float pmz = live ? lib.p_mz[row] : 0.0f;
  const uint64_t fs = live ? lib.f_start[row] : 0;
  const uint64_t fe = live ? lib.f_stop[row] : 0;
  constexpr int kIsoBatch = 4;  // LP-window batches per round trip
  float iso_lo[kIsoBatch], iso_hi[kIsoBatch];
  auto load_iso = [&](uint32_t w0) {
#pragma unroll
    for (int u = 0; u < kIsoBatch; ++u) {
      const uint32_t w = w0 + (uint32_t)(u * LP + gl);
      const bool in = w < dia.n_obs;
      iso_lo[u] = in ? dia.iso_lo[w] : 0.0f;
      iso_hi[u] = in ? dia.iso_hi[w] : 0.0f;
    }
  };
  load_iso(0);
  const uint32_t nraw = (uint32_t)(fe - fs);
  const bool mine = gl < (int)nraw;
  float v = 0.0f, mz = 0.0f;
  uint32_t ftype = 0, fnum = 0;
  if (mine) {
    v = lib.f_int[fs + gl];
    mz = lib.f_mz[fs + gl];
    ftype = lib.f_type[fs + gl];
    fnum = lib.f_num[fs + gl];
  }
  // src/peak_group_selection/mod.rs:139-148 -> rt_index/mod.rs:89-138
  uint32_t c0 = 0, c1 = 0;
  if (live) {
    const uint32_t n = dia.n_rt;
    if (n > 0) {
      const float min_tol = (10.0f + prm.pad_cycles) * dia.delta_t;
      const float eff = fmaxf(rt_tol, min_tol);
      const float lo = prt - eff, hi = prt + eff;
      if (hi < dia.rt[0]) {
        c0 = c1 = 0;
      } else if (lo > dia.rt[n - 1]) {
        c0 = c1 = n;
      } else {
        c0 = (lo <= dia.rt[0]) ? 0u : rt_search(dia.rt, n, dia.rt_strict, dia.rt_inv_dt, lo);
        c1 = (hi >= dia.rt[n - 1]) ? n : rt_search(dia.rt, n, dia.rt_strict, dia.rt_inv_dt, hi);
      }
    }
  }
  // src/dia_data/mod.rs:79-87 valid observations (inclusive, ascending)
  uint32_t nobs = 0, first = 0xffffffffu;
  for (uint32_t wb = 0; wb < dia.n_obs; wb += kIsoBatch * LP) {
    if (wb > 0) load_iso(wb);
#pragma unroll
    for (int u = 0; u < kIsoBatch; ++u) {
      const uint32_t w0 = wb + (uint32_t)(u * LP);
      if (w0 >= dia.n_obs) break;
      const uint32_t w = w0 + (uint32_t)gl;
      const bool ok = live && w < dia.n_obs && iso_lo[u] <= pmz && iso_hi[u] >= pmz;
      const uint64_t m = ballot64(ok) & gmask;
      if (ok) {
        const uint32_t rank = nobs + (uint32_t)__builtin_popcountll(m & ((1ull << lane) - 1));
        if (rank < (uint32_t)kObsMax) plan.obs[(size_t)row * kObsMax + rank] = (uint16_t)w;
      }
      if (m && first == 0xffffffffu) first = w0 + (uint32_t)(__builtin_ctzll(m) - gb);
      nobs += (uint32_t)__builtin_popcountll(m);
    }
  }
  if (nobs > (uint32_t)kObsMax) st |= kStTooManyWindows;
  // src/speclib_flat/mod.rs:388-456 get_precursor_filtered(i, true, true, k)
  float I = -__builtin_inff();
  if (mine) {
    const bool keep = v > 0.0f && !(ftype == 121u && fnum == 1u);
    I = keep ? v : -__builtin_inff();
  }
  g_I[lane] = I;
  g_mz[lane] = mz;
  const bool kept = I > 0.0f;
  const uint64_t km = ballot64(kept) & gmask;
  const uint32_t nkept = (uint32_t)__builtin_popcountll(km);
  const uint32_t pos_kept = (uint32_t)__builtin_popcountll(km & ((1ull << lane) - 1));
  const uint32_t k = min(prm.top_k, nkept);
  uint32_t nsel = 0;
  if (k > 0 && k > plan.ks) {  // ks <= kMaxSel for a search
    st |= kStTooManyFragments;
  } else if (k > 0) {
    nsel = k;
  }
  const int mode = (k >= nkept) ? 0 : (k == nkept - 1 ? 1 : 2);
  // mode 1: the first fragment of minimum intensity is dropped and the
  // last kept one takes its place
  uint32_t drop = 0xffffffffu, drop_pos = 0;
  const uint32_t last = km ? (uint32_t)(63 - __builtin_clzll(km) - gb) : 0u;
  {
    float mv = kept ? I : __builtin_inff();
    uint32_t mi = kept ? (uint32_t)gl : 0xffffffffu;
#pragma unroll
    for (int o = LP / 2; o >= 1; o >>= 1) {
      const float ov = __shfl_xor(mv, o, kWave);
      const uint32_t oi = __shfl_xor(mi, o, kWave);
      if (oi != 0xffffffffu && (mi == 0xffffffffu || ov < mv || (ov == mv && oi < mi))) {
        mv = ov;
        mi = oi;
      }
    }
    drop = mi;
    drop_pos = (uint32_t)__shfl((int)pos_kept, (int)(gb + (mi & (LP - 1))), kWave);
  }
  __builtin_amdgcn_wave_barrier();
  // more than 16 kept fragments (mode 2): the group's first lane runs the
  // selection over the kept fragments' lanes, in order, in g_key; each
  // lane then reads its position from the inverse permutation
  const bool big = live && nsel > 0 && mode == 2 && nkept > 16;
  if (big && gl == 0) {
    uint32_t q = 0;
    for (uint32_t i = 0; i < nraw; ++i)
      if (g_I[gb + i] > 0.0f) g_key[gb + q++] = i;
    const GroupView gv{g_key + gb, g_I + gb};
    if (!rs_select(gv, nkept, k)) st |= kStSelectFallback;
  }
  __builtin_amdgcn_wave_barrier();
  const uint32_t at = big && (uint32_t)gl < nkept ? g_key[lane] : 0xffffffffu;
  __builtin_amdgcn_wave_barrier();
  if (big) g_key[lane] = 0xffffffffu;
  __builtin_amdgcn_wave_barrier();
  if (at != 0xffffffffu && (uint32_t)gl < k && !(st & kStSelectFallback))
    g_key[gb + at] = (uint32_t)gl;
  __builtin_amdgcn_wave_barrier();
  uint32_t key = 0xffffffffu;
  if (kept && nsel > 0) {
    if (mode == 2 && big) {
      key = g_key[lane];
    } else if (mode == 2) {
      uint32_t r = 0;
      for (uint32_t i = 0; i < nraw; ++i) {
        const float u = g_I[gb + i];
        if (u > I || (u == I && i < (uint32_t)gl)) ++r;
      }
      key = r < k ? r : 0xffffffffu;
    } else if (mode == 1) {
      key = ((uint32_t)gl == drop) ? 0xffffffffu
            : (((uint32_t)gl == last && last != drop) ? drop_pos : pos_kept);
    } else {
      key = pos_kept;
    }
  }
  g_key[lane] = key;
  __builtin_amdgcn_wave_barrier();
  if (key != 0xffffffffu) {
    uint32_t pos = 0;
    for (uint32_t i = 0; i < nraw; ++i) {
      const uint32_t ki = g_key[gb + i];
      if (i == (uint32_t)gl || ki == 0xffffffffu) continue;
      const float mi = g_mz[gb + i];
      if (mi < mz || (!(mi > mz) && !(mi < mz) && ki < key)) ++pos;
    }
    plan.sel[(size_t)row * plan.ks + pos] = (uint16_t)gl;
    plan.sel_mz[(size_t)row * plan.ks + pos] = mz;
    plan.sel_w[(size_t)row * plan.ks + pos] = v;
  }
  if (live && gl == 0) {
    const uint32_t kidx = (prm.per_window && first != 0xffffffffu) ? 1u + first : 0u;
    plan.rec[row] = make_uint4(c0, c1, nsel | (min(nobs, (uint32_t)kObsMax) << 16), kidx);
    if (c1 > c0 && c1 - c0 > (uint32_t)kLCapMax) st |= kStWindowTooLong;
    if (st) atomicOr(status, st);
  }
}

// ---------------------------------------------------------------------------
// bins: one thread per (row, selected-fragment slot)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void bins_kernel(KDia dia, KLib lib,
                                                   KParams prm, KPlan plan) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t row = t / plan.ks;
  const uint32_t slot = (uint32_t)(t - row * plan.ks);
  if (row >= lib.n_prec) return;
  const uint4 rec = plan.rec[row];
  const uint32_t nsel = rec.z & 0xffffu, nobs = rec.z >> 16;
  if (slot >= nsel || rec.y <= rec.x) return;
  // src/quadrupole_observation/mod.rs:79-83
  const float mz = plan.sel_mz[row * plan.ks + slot];
  const float delta = mz * prm.mass_tol * 1e-6f;
  const float lo = mz - delta, hi = mz + delta;
  uint32_t start, end;
  mz_range(dia, lo, hi, start, end);
  for (uint32_t j = 0; j < nobs; ++j) {
    const uint32_t o = plan.obs[row * kObsMax + j];
    const uint32_t *ss = dia.ss + (size_t)o * (dia.n_grid + 1);
    const size_t idx = (row * kObsMax + j) * plan.ks + slot;
    plan.rng[idx] = make_uint2(ss[start], ss[end]);
    plan.b0[idx] = start;
  }
}

// ---------------------------------------------------------------------------
// select: one wavefront per precursor
// ---------------------------------------------------------------------------
struct SelArgs {
  KDia dia;
  KLib lib;
  KParams prm;
  KPlan plan;
  uint32_t n_rows;
  uint32_t *queue;
  uint32_t chunk;
  uint32_t lds_per_wave;  // bytes
  uint32_t lcap;          // max cycles per window (multiple of 64)
  uint32_t xcap;          // floats in the XIC tile
  uint32_t gcap;          // fragment rows of the longest row (host side)
  float *cand_score;      // [row][cand_count]
  uint32_t *cand_center;  // [row][cand_count]
  uint32_t *cand_count;   // [row]
  uint32_t *status;
  float *dbg_scores;
  uint32_t dbg_stride, dbg_row0, dbg_rows;
  // Rows longer than split_l cycles (<= lcap) are not searched but appended
  // to defer_rows; a second launch searches the rows of such a list:
  // rows[q] for q < *n_rows_at (rows == nullptr: row q, q < n_rows).
  uint32_t split_l;
  uint32_t *defer_rows;
  uint32_t *defer_count;
  const uint32_t *rows;
  const uint32_t *n_rows_at;
  uint32_t row0;  // library row of row 0 (chunked searches): defer list entries
  // differentiable forward: per row, the XIC around the rank-0 apex
  // (xwin_stride floats: fragment-major, kWin cycles from apex - half) and
  // whether it was written (rows of the dense tile are not)
  float *xwin;
  uint32_t *xwin_ok;
  uint32_t xwin_stride;
};

// Per-wave LDS.  XIC rows have kXPad zero floats on both sides so that a
// block of 4 outputs reads its 21-tap neighbourhood with aligned 16-byte
// loads.  BD holds one bit per 4-cycle output block that can be non-zero.
// hits is the extraction's hit list (entry or intensity, key), items the
// (fragment, block) work list of the convolution, own the duplicate
// detector of the hit accumulation (one byte per fragment and cycle).
constexpr int kXPad = 16;
constexpr int kItemCap = 512;   // u16 (fragment, block) items
constexpr int kHitCap = 256;    // (u32 entry or intensity, u32 key) hits
// cycles of XIC kept around a row's rank-0 apex for the backward pass
// (2 * half + 1 <= 31)
constexpr int kWin = 32;
constexpr int kTileHits = 4 * 64;  // hits one extraction tile can add
// extraction tiles in flight (8 entries per lane each); one measured
// fastest: a second tile's registers cost more than its latency hiding
constexpr int kGroupTiles = 1;
// a batch's last items up to this count run one output per lane (one
// pass of 64 outputs; 16 measured best, 32 and 48 slower)
constexpr int kTailItems = 16;

struct LdsShape {
  uint32_t lp;     // XIC row stride (floats)
  uint32_t wb;     // BD row stride (u64 words)
};

__host__ __device__ inline LdsShape lds_shape(uint32_t lcap) {
  LdsShape s;
  s.lp = lcap + 2 * kXPad;
  s.wb = (lcap / 4 + 63) / 64;
  return s;
}

// Bytes of a wave's LDS that do not depend on the number of fragment rows
// (R with 4 + 8 pad floats, hits, items, kernel taps, weights), and the
// bytes of one fragment row (X row, BD row, own row), for rows of up to
// lrow cycles.
__host__ __device__ inline uint32_t lds_fixed(uint32_t lrow) {
  return (lrow + 12) * 4 + kHitCap * 8 + kItemCap * 2 + 32 * 4 + kMaxSel * 4;
}

__host__ __device__ inline uint32_t lds_per_frag(uint32_t lrow) {
  const LdsShape s = lds_shape(lrow);
  return s.lp * 4 + s.wb * 8 + lrow;
}

__host__ __device__ inline uint32_t lds_bytes(uint32_t lcap, uint32_t gcap) {
  const uint32_t b = lds_fixed(lcap) + gcap * lds_per_frag(lcap);
  return (b + 15) / 16 * 16;
}

struct WaveLds {
  float *R;
  float *X;
  uint64_t *BD;
  uint16_t *items;
  uint2 *hits;
  uint8_t *own;
  float *w;
  float *kern;
  uint32_t lrow;  // row capacity in cycles (multiple of 64, >= L)
  uint32_t lp;    // XIC row stride (floats)
  uint32_t wb;    // BD row stride (u64 words)
  uint32_t gcap;  // fragment rows in the XIC tile
  uint32_t tile_bytes;  // bytes from X to the end of the wave's share
  uint32_t hcap;  // hit list capacity (entries)
  uint32_t room;  // bytes from the hit list to the end of the wave's share
};

// A wave's LDS for one row of L <= lcap cycles and nsel fragments, carved
// from the row's own length: the score row, the item list, kernel taps and
// weights, then the hit list.  The hit list holds kHitCap hits, plus what
// the share holds beyond the fixed parts and two dense fragment rows.  The
// XIC tile (carve_tile) follows the part of the hit list in use, so that
// rows with few hits or short RT windows take more fragments per group.
// `bytes` >= lds_bytes(lcap, 1) guarantees at least one.
__device__ __forceinline__ WaveLds carve(char *base, uint32_t L, uint32_t bytes) {
  WaveLds l;
  l.lrow = max(64u, (L + 63) / 64 * 64);
  const LdsShape sh = lds_shape(l.lrow);
  l.lp = sh.lp;
  l.wb = sh.wb;
  const uint32_t need = lds_fixed(l.lrow) + 2u * lds_per_frag(l.lrow);
  l.hcap = (uint32_t)kHitCap + (bytes > need ? (bytes - need) / 128u * 16u : 0u);
  char *p = base;
  l.R = (float *)p + 4;
  p += (l.lrow + 12) * 4;
  l.items = (uint16_t *)p;
  p += kItemCap * 2;
  l.kern = (float *)p;
  p += 32 * 4;
  l.w = (float *)p;
  p += kMaxSel * 4;
  l.hits = (uint2 *)p;
  l.room = bytes - (uint32_t)(p - base);
  return l;
}

// The dense XIC tile after the first `used` entries of the hit list.
__device__ __forceinline__ void carve_tile(WaveLds &l, uint32_t used, uint32_t nsel) {
  const uint32_t off = (used * 8u + 15u) / 16u * 16u;
  char *p = reinterpret_cast<char *>(l.hits) + off;
  l.X = (float *)p;
  l.tile_bytes = l.room - off;
  l.gcap = max(1u, min(nsel, l.tile_bytes / lds_per_frag(l.lrow)));
  p += l.gcap * l.lp * 4;
  l.BD = (uint64_t *)p;
  p += l.gcap * l.wb * 8;
  l.own = (uint8_t *)p;
}

// The block-sparse XIC tile of a row: every selected fragment at once, with
// only the 4-cycle cell blocks that hold a hit stored.  Per fragment: NZB, a
// bit per cell block that holds a hit; BD, a bit per 4-output block that a
// hit reaches (the convolution's work items); ST, the storage slot of each
// cell block (byte b + 4 for block b, -4 <= b < nb + 4; 0 = the zero block).
// CB holds the slots' cells (slot 0 all zero) and own a claim byte per cell.
struct SparseTile {
  uint64_t *NZB;
  uint64_t *BD;
  uint8_t *ST;
  float4 *CB;
  uint8_t *own;
  uint32_t sts;  // ST bytes per fragment
  uint32_t cap;  // slots (including the zero block), at most 255
};

__device__ __forceinline__ SparseTile carve_sparse(const WaveLds &l, uint32_t nsel) {
  SparseTile s;
  char *p = reinterpret_cast<char *>(l.X);
  const uint32_t words = nsel * l.wb;
  s.sts = l.lrow / 4 + 8;
  s.NZB = reinterpret_cast<uint64_t *>(p);
  s.BD = s.NZB + words;
  s.ST = reinterpret_cast<uint8_t *>(s.BD + words);
  const uint32_t head = (16u * words + nsel * s.sts + 15u) & ~15u;
  s.CB = reinterpret_cast<float4 *>(p + head);
  s.cap = head + 40u <= l.tile_bytes ? min(255u, (l.tile_bytes - head) / 20u) : 0u;
  s.own = reinterpret_cast<uint8_t *>(s.CB + s.cap);
  return s;
}

// small[0]: queue, [1]: status, [3]: queue of the long-row launch,
// [4]: deferred row count, [8 + k]: queue of chunk k of a chunked search
// ([2], the candidate total, is written by the scan)
constexpr int kSmallWords = 32;
constexpr int kMaxChunks = kSmallWords - 8;
// pipelined chunks: chunk k's queue is small[8 + k], its candidate count
// small[16 + k]
constexpr int kPipeMax = 8;
__global__ void reset_queue_kernel(uint32_t *small) {
  if (threadIdx.x < kSmallWords && threadIdx.x != 2) small[threadIdx.x] = 0;
}

// The small words into page-locked host memory (`host` is its device
// address): stores from a kernel, where a copy would be a blit kernel that
// waits for compute units.
__global__ void publish_kernel(const uint32_t *small, uint32_t *host) {
  if (threadIdx.x < kSmallWords) host[threadIdx.x] = small[threadIdx.x];
}

__device__ __forceinline__ uint32_t rdlane(uint32_t v, uint32_t lane) {
  return __builtin_amdgcn_readlane(v, lane);
}

__device__ __forceinline__ uint32_t uni(uint32_t v) {
  return __builtin_amdgcn_readfirstlane(v);
}

// True when entry e (relative to the window) of a hit with cycle == c0 is
// followed, within its bin, by another entry with cycle == c0: the
// reference's binary_search(c0) starts at the last such entry
// (src/quadrupole_observation/mod.rs:92-94).  `next` is the cycle of the
// window's entry e + 1; only when it equals c0 is e's bin found (a walk over
// the fragment's bins from b0, in global memory).
__device__ __forceinline__ bool skipped_duplicate(const KDia &dia, uint32_t o,
                                                  uint32_t b0, uint32_t e,
                                                  uint32_t c0, uint32_t next) {
  if (next != c0) return false;
  const uint32_t *ss = dia.ss + (size_t)o * (dia.n_grid + 1);
  uint32_t b = b0;
  while (ss[b + 1] <= e) ++b;
  return e + 1 < ss[b + 1];
}

// Four consecutive outputs i0..i0+3 of src/convolution/scalar.rs:14-36
// (s = 0 + x[i]*k[h], then s += (x[i-d] + x[i+d]) * k[h-d] for d = h..1)
// for h = 10: xb holds the cells i0 - 12 .. i0 + 15 of the zero-padded row
// as seven 4-cell blocks.  The leading 0 + is dropped: it only turns a -0
// product into +0, and an output that ends up +-0 scores ln(+-0 + 1) = 0
// either way.
__device__ __forceinline__ void conv4_h10(const float4 (&xb)[7], const float (&k)[11],
                                          float (&y)[4]) {
  float x[28];
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    const float4 v = xb[q];
    x[4 * q + 0] = v.x;
    x[4 * q + 1] = v.y;
    x[4 * q + 2] = v.z;
    x[4 * q + 3] = v.w;
  }
#pragma unroll
  for (int o = 0; o < 4; ++o) {
    const int c = 12 + o;
    float s = x[c] * k[10];
#pragma unroll
    for (int j = 0; j < 10; ++j) {
      const int d = 10 - j;
      s = s + (x[c - d] + x[c + d]) * k[j];
    }
    y[o] = s;
  }
}

// Same for any half width <= 15, taps read from LDS.
__device__ __forceinline__ void conv4_any(const float *row, uint32_t i0,
                                          uint32_t half, const float *k,
                                          float (&y)[4]) {
#pragma unroll
  for (int o = 0; o < 4; ++o) {
    const float *xc = row + i0 + o;
    float s = xc[0] * k[half];
    for (uint32_t j = 0; j < half; ++j) {
      const uint32_t d = half - j;
      s = s + (xc[-(int)d] + xc[d]) * k[j];
    }
    y[o] = s;
  }
}

// Slots of each lane's hits in a list ordered by lane, then entry: `pre` is
// the number of hits (set bits of the 8-bit mask hm) of the active lanes
// below this one, `tot` the active lanes' total.  The per-lane count (at
// most 8) is summed bit by bit with one ballot per bit.
__device__ __forceinline__ void list_slots(uint32_t hm, bool active, uint32_t &pre,
                                           uint32_t &tot) {
  const uint32_t h = active ? (uint32_t)__builtin_popcount(hm) : 0u;
  pre = 0;
  tot = 0;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const uint64_t bb = ballot64((h >> b) & 1u);
    pre += __builtin_amdgcn_mbcnt_hi((uint32_t)(bb >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)bb, 0u))
           << b;
    tot += (uint32_t)__builtin_popcountll(bb) << b;
  }
}

// Inclusive prefix maximum over the wave's lanes through DPP: row_shr 1, 2,
// 4, 8 within each row of 16 lanes, then row_bcast 15 and 31 carry rows
// 0..2 forward.  Lanes without a source keep their own value (max is
// idempotent).
__device__ __forceinline__ uint32_t wave_incl_max_dpp(uint32_t x) {
  x = max(x, (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0x111, 0xf, 0xf, false));
  x = max(x, (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0x112, 0xf, 0xf, false));
  x = max(x, (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0x114, 0xf, 0xf, false));
  x = max(x, (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0x118, 0xf, 0xf, false));
  x = max(x, (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0x142, 0xa, 0xf, false));
  x = max(x, (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0x143, 0xc, 0xf, false));
  return x;
}

// Order-preserving u32 image of a float that is not NaN: a larger float
// gives a larger key, -0 and +0 give the same key, and 0 lies below the key
// of every float.
__device__ __forceinline__ uint32_t f32_order_key(float v) {
  const uint32_t u = v == 0.0f ? 0u : __builtin_bit_cast(uint32_t, v);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Wave-wide arg-max over one (key, position) pair per lane: the largest
// key, ties to the lowest position; key 0 marks a lane without a pair.
// Returns the winning position in every lane, ~0u when no lane has a pair.
__device__ __forceinline__ uint32_t wave_argmax_key(uint32_t key, uint32_t pos) {
  const uint32_t m = __builtin_amdgcn_readlane((int)wave_incl_max_dpp(key), 63);
  if (m == 0u) return 0xffffffffu;
  const uint64_t tie = ballot64(key == m);
  if ((tie & (tie - 1)) == 0ull)
    return (uint32_t)__builtin_amdgcn_readlane((int)pos, (int)__builtin_ctzll(tie));
  // several lanes hold the largest key: the lowest position among them
  const uint32_t c = key == m ? ~pos : 0u;
  return ~(uint32_t)__builtin_amdgcn_readlane((int)wave_incl_max_dpp(c), 63);
}

// Occupancy: at most 128 VGPRs (4 waves per SIMD) and, with the default
// 10 KB of LDS per wave, 16 waves per CU; the kernel is latency-bound, and
// measured faster there than with more registers or more LDS per wave.
// static LDS of select_kernel per block (the logf table)
constexpr uint32_t kSelectStaticLds = 16 * sizeof(LogfEntry);

// SAVE (h = 10 only): the differentiable forward, which leaves each row's
// apex XIC window for the backward pass (SelArgs::xwin)
template <int WAVES, int HALF, bool SAVE = false>
__global__ __launch_bounds__(WAVES *kWave) __attribute__((amdgpu_waves_per_eu(4)))
void select_kernel(SelArgs a) {
  extern __shared__ __align__(16) char smem[];
  __shared__ LogfEntry ltab[16];
  if (threadIdx.x < 16) ltab[threadIdx.x] = kDevLogfTab[threadIdx.x];
  __syncthreads();
  const int lane = lane_id();
  const int wv = threadIdx.x / kWave;
  char *const wave_lds = smem + (size_t)wv * a.lds_per_wave;
  const uint32_t half = HALF ? (uint32_t)HALF : a.prm.half;
  const uint32_t cc = a.prm.cand_count;
  const uint32_t ks = a.plan.ks;

  // Row schedule: the queue hands out `chunk` rows at a time.  The next
  // row's plan data is loaded before the current row is processed (stage
  // 1: record, weights, windows, first-window CSR ranges) and its dependent
  // values half-way through (stage 2: kernel taps, window entry bases), so
  // that their latency overlaps the current row's work.
  // The queue is grabbed one chunk ahead: `ahead` is the start of the
  // chunk after the current one, requested when the current one began.
  const uint32_t n_rows = a.n_rows_at ? *a.n_rows_at : a.n_rows;
  uint32_t nxt = 0, cend = 0, ahead = 0;
  auto request = [&]() {
    uint32_t r0 = 0;
    if (lane == 0) r0 = atomicAdd(a.queue, a.chunk);
    return r0;
  };
  auto grab = [&]() {
    nxt = uni(__shfl(ahead, 0, kWave));
    cend = min(nxt + a.chunk, n_rows);
    ahead = nxt < n_rows ? request() : 0xffffffffu;
  };
  ahead = request();
  // Outputs of the last finished row are stored after the next row's first
  // loads are issued, so that those loads do not wait for the stores.
  uint32_t out_row = 0xffffffffu, out_n = 0, out_c = 0;
  float out_s = 0.0f;
  auto flush_out = [&]() {
    if (out_row != 0xffffffffu) {
      if ((uint32_t)lane < min(out_n, (uint32_t)kWave)) {
        const uint32_t ol = opaque_lane();
        (a.cand_score + (size_t)out_row * cc)[ol] = out_s;
        (a.cand_center + (size_t)out_row * cc)[ol] = out_c;
      }
      a.cand_count[out_row] = out_n;
      out_row = 0xffffffffu;
    }
  };
  uint4 p_rec = make_uint4(0u, 0u, 0u, 0u);
  float p_sw = 0.0f, p_tap = 0.0f, p_smz = 0.0f, p_lo = 0.0f, p_hi = 0.0f;
  uint32_t p_obs = 0, p_bs = 0, p_be = 0;
  uint4 p_elo = make_uint4(0u, 0u, 0u, 0u), p_ehi = make_uint4(0u, 0u, 0u, 0u);
  uint2 p_rng = make_uint2(0u, 0u);
  uint64_t p_eb = 0;
  uint32_t p_row = 0;
  const float mtol = a.prm.mass_tol;
  auto stage1 = [&](uint32_t q) {
    const uint32_t r = a.rows ? a.rows[q] : q;
    p_row = r;
    p_rec = a.plan.rec[r];
    const uint32_t ol = opaque_lane();
    p_sw = lane < (int)ks ? (a.plan.sel_w + (size_t)r * ks)[ol] : 0.0f;
    p_smz = lane < (int)ks ? (a.plan.sel_mz + (size_t)r * ks)[ol] : 0.0f;
    p_obs = lane < kObsMax ? (a.plan.obs + (size_t)r * kObsMax)[ol] : 0u;
  };
  // src/quadrupole_observation/mod.rs:79-83 and src/mz_index/mod.rs:95-127:
  // the fragment's ppm window and the bucket entries of its two bounds
  auto stage2 = [&]() {
    const uint32_t kidx = uni(p_rec.w), pobs = uni(p_rec.z >> 16);
    const uint32_t pnsel = uni(p_rec.z & 0xffffu);
    p_tap = lane < (int)a.prm.klen
                ? (a.prm.kernels + (size_t)kidx * a.prm.klen)[opaque_lane()]
                : 0.0f;
    p_eb = lane < (int)min(pobs, (uint32_t)kObsMax) ? a.dia.entry_base[p_obs] : 0ull;
    const float delta = p_smz * mtol * 1e-6f;
    p_lo = p_smz - delta;
    p_hi = p_smz + delta;
    const bool on = lane < (int)pnsel;
    if (on && p_lo > a.dia.grid_first && !(p_lo > a.dia.grid_last))
      p_elo = a.dia.bucket2[bucket_key(a.dia, p_lo)];
    if (on && p_hi >= a.dia.grid_first && p_hi < a.dia.grid_last)
      p_ehi = a.dia.bucket2[bucket_key(a.dia, p_hi)];
  };
  // bin range [p_bs, p_be) of each fragment, CSR entry range in window 0
  auto stage3 = [&]() {
    const uint32_t pobs = uni(p_rec.z >> 16), pnsel = uni(p_rec.z & 0xffffu);
    const uint32_t n = a.dia.n_grid;
    uint32_t st = 0, en = 0;
    if (lane < (int)pnsel) {
      st = !(p_lo > a.dia.grid_first) ? 0u
           : (p_lo > a.dia.grid_last) ? n
                                      : lb_from(p_elo, p_lo);
      const uint32_t ub = !(p_hi >= a.dia.grid_first) ? 0u
                          : (p_hi >= a.dia.grid_last) ? n
                                                      : ub_from(p_ehi, p_hi);
      en = st < n ? max(st, ub) : st;
    }
    p_bs = st;
    p_be = en;
    p_rng = make_uint2(0u, 0u);
    if (pobs > 0 && lane < (int)pnsel) {
      const uint32_t o0 = (uint32_t)__builtin_amdgcn_readlane((int)p_obs, 0);
      const uint32_t *ssw = a.dia.ss + (size_t)o0 * (n + 1);
      p_rng = make_uint2(ssw[st], ssw[en]);
    }
  };
  grab();
  if (nxt < n_rows) {
    stage1(nxt);
    stage2();
    stage3();
  }
  while (nxt < n_rows) {
    {
      const uint32_t row = p_row;
      const uint4 rec = p_rec;
      const float sw = p_sw, tap = p_tap;
      const uint32_t obs_l = p_obs, bs_l = p_bs, be_l = p_be;
      const uint2 rng0 = p_rng;
      const uint64_t eb_l = p_eb;
      if (++nxt >= cend) grab();
      const bool more = nxt < n_rows;
      if (more) stage1(nxt);
      const uint32_t c0 = uni(rec.x), c1 = uni(rec.y);
      const uint32_t nsel = uni(rec.z & 0xffffu), nobs = uni(rec.z >> 16);
      // Stores next to a `continue` or the loop's back edge are made by all
      // lanes (same value, same address): a lane-0-only region there lets
      // the compiler thread it into the lane-0-only queue grab, after which
      // lanes 1..63 no longer follow lane 0 through the loop.
      if (c1 <= c0) {  // mod.rs:146-148 -> None
        a.cand_count[row] = 0;
        if (more) {
          stage2();
          stage3();
        }
        continue;
      }
      const uint32_t L = c1 - c0;
      if (L > a.split_l) {
        if (L > a.lcap) {  // excluded by the host bound; flagged if it happens
          a.cand_count[row] = 0;
          atomicOr(a.status, (uint32_t)kStWindowTooLong);
        } else {
          // deferred to the launch over long rows.  Every lane adds (lane 0
          // one, the others zero), so lane 0's old value is the row's slot.
          const uint32_t slot = uni(atomicAdd(a.defer_count, lane == 0 ? 1u : 0u));
          a.defer_rows[slot] = a.row0 + row;
        }
        if (more) {
          stage2();
          stage3();
        }
        continue;
      }
      WaveLds l = carve(wave_lds, L, a.lds_per_wave);
      for (uint32_t i = 4 * opaque_lane(); i < L; i += 4 * kWave)
        *reinterpret_cast<float4 *>(l.R + i) = make_float4(0.f, 0.f, 0.f, 0.f);
      float k10[11];
      if (HALF == 10) {
#pragma unroll
        for (int j = 0; j < 11; ++j)
          k10[j] = __builtin_bit_cast(
              float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, tap), j));
      } else {
        if (lane <= (int)half) l.kern[lane] = tap;
      }
      if (lane < (int)nsel) l.w[lane] = sw;
      __builtin_amdgcn_wave_barrier();
      // window j of this row: obs index and entry base
      auto win_obs = [&](uint32_t j) { return (uint32_t)__builtin_amdgcn_readlane((int)obs_l, (int)j); };
      auto win_base = [&](uint32_t j) {
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)eb_l, (int)j);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(eb_l >> 32), (int)j);
        return ((uint64_t)hi << 32) | lo;
      };

      // ---- extraction: src/dense_xic_observation/mod.rs:42-55 ------------
      // Windows in order; within a window the fragments' CSR ranges are cut
      // into 8-entry groups (one 16-byte load of u16 cycles per lane), 64
      // groups per tile across fragment boundaries, kGroupTiles tiles in
      // flight.  In-window hits go to an LDS list in entry order (window,
      // fragment, bin, entry), which is the reference's summation order.
      // Hit key: t (12 bits) | fragment (6) << 12 | window (2) << 18.
      uint32_t nh = 0;
      uint32_t gA = 0, gG = 0;  // fragment group being accumulated
      // spm: the row's XIC is the block-sparse tile st (all fragments)
      bool spm = false;
      SparseTile st;
      // Adds the hits in list slots [h0, h1) of fragments [gA, gA + gG)
      // into X (or the sparse tile); the list holds intensities (loaded) or
      // window-relative entries of window `base`.
      auto accumulate = [&](bool loaded, uint64_t base, uint32_t h0, uint32_t h1) {
        for (uint32_t c = h0; c < h1; c += kWave) {
          const uint32_t idx = c + lane;
          uint2 h = make_uint2(0u, 0u);
          if (idx < h1) h = l.hits[idx];
          const uint32_t f = (h.y >> 12) & 63u, t = h.y & 0xfffu;
          const bool v = idx < h1 && f >= gA && f < gA + gG;
          const uint32_t fl = v ? f - gA : 0u;
          float inten = 0.0f;
          if (v) inten = loaded ? __builtin_bit_cast(float, h.x) : a.dia.intens[base + h.x];
          uint8_t *ow;
          float *xc;  // the hit's XIC cell
          uint64_t *br;
          if (spm) {
            const uint32_t slot = v ? st.ST[fl * st.sts + (t >> 2) + 4u] : 0u;
            ow = st.own + slot * 4u + (t & 3u);
            xc = reinterpret_cast<float *>(st.CB + slot) + (t & 3u);
            br = st.BD + fl * l.wb;
          } else {
            ow = l.own + fl * l.lrow + t;
            xc = l.X + fl * l.lp + kXPad + t;
            br = l.BD + fl * l.wb;
          }
          if (v) *ow = (uint8_t)lane;
          __builtin_amdgcn_wave_barrier();
          // lanes whose cell another lane overwrote share it with that lane
          const bool lost = v && *ow != (uint8_t)lane;
          if (ballot64(lost) == 0) {
            if (v) *xc = *xc + inten;
          } else {
            // The losers write again, so that the lane that kept a shared
            // cell reads another id: every lane of a shared cell is then
            // marked.  Unshared cells add at once, shared ones one lane at a
            // time in lane (= entry) order.
            if (lost) *ow = (uint8_t)lane;
            __builtin_amdgcn_wave_barrier();
            const bool shared = lost || (v && *ow != (uint8_t)lane);
            if (v && !shared) *xc = *xc + inten;
            __builtin_amdgcn_wave_barrier();
            uint64_t it2 = ballot64(shared);
            while (it2) {
              const int j2 = __builtin_ctzll(it2);
              it2 &= it2 - 1;
              if (lane == j2) *xc = *xc + inten;
              __builtin_amdgcn_wave_barrier();
            }
          }
          // blocks of the outputs this hit reaches, within the computed
          // range [half, L - half) of the convolution
          if (v) {
            const int ilo = max((int)t - (int)half, (int)half);
            const int ihi = min((int)(t + half), (int)L - (int)half - 1);
            if (ilo <= ihi) {
              const uint32_t blo = (uint32_t)ilo >> 2, bhi = (uint32_t)ihi >> 2;
              const uint32_t wlo = blo >> 6, whi = bhi >> 6;
              const uint64_t ones = ~0ull;
              if (wlo == whi) {
                const uint64_t m = (ones >> (63 - (bhi - blo))) << (blo & 63);
                atomicOr((unsigned long long *)&br[wlo], m);
              } else {
                atomicOr((unsigned long long *)&br[wlo], ones << (blo & 63));
                atomicOr((unsigned long long *)&br[whi], ones >> (63 - (bhi & 63)));
              }
            }
          }
          __builtin_amdgcn_wave_barrier();
        }
      };
      // Lists the hits of fragments [fa, fb) of every window.  Deferred: the
      // list holds whole fragments; the first fragment not listed is
      // returned (fb when all are).  When the list would overflow, a row
      // with one window keeps the fragments before the one being scanned
      // (the list is in fragment order); otherwise nothing is kept and fa is
      // returned.  Direct: the list is accumulated into X whenever it fills
      // and at the end of each window.
      auto extract = [&](uint32_t fa, uint32_t fb, bool deferred) -> uint32_t {
        const uint32_t FN = fb - fa;
        for (uint32_t j = 0; j < nobs; ++j) {
          const uint32_t o = win_obs(j);
          const uint64_t base = win_base(j);
          uint32_t rx = 0, ry = 0, rb0 = 0;
          if (j == 0 && fa == 0) {
            if (lane < (int)FN) {
              rx = rng0.x;
              ry = rng0.y;
              rb0 = bs_l;
            }
          } else {
            // bins of fragment fa + lane, entries of window j
            const uint32_t bs = (uint32_t)__shfl((int)bs_l, (int)min(fa + (uint32_t)lane, 63u), kWave);
            const uint32_t be = (uint32_t)__shfl((int)be_l, (int)min(fa + (uint32_t)lane, 63u), kWave);
            if (lane < (int)FN) {
              const uint32_t *ssw = a.dia.ss + (size_t)o * (a.dia.n_grid + 1);
              rx = ssw[bs];
              ry = ssw[be];
              rb0 = bs;
            }
          }
          // Groups of 8 entries are counted from the window's 8-aligned
          // start (group g covers entries 8g - bo .. 8g - bo + 7 of the
          // window), so one 16-byte load at cw + 8g reads a group.
          // Fragment fa + lane spans groups gsr .. gsr + ng - 1; its first
          // and last groups are partial.
          const uint32_t bo = (uint32_t)base & 7u;
          const uint16_t *cw = a.dia.cycles + (base - bo);
          const uint32_t gsr = (bo + rx) >> 3;
          const uint32_t ng = ry > rx ? ((bo + ry - 1) >> 3) - gsr + 1 : 0u;
          // packed: first entry in the first group, last in the last, count
          const uint32_t meta = ((bo + rx) & 7u) | (((bo + ry - 1) & 7u) << 3) | (ng << 6);
          const uint32_t gincl = wave_incl_scan(ng);
          const uint32_t gp = gincl - ng;  // first virtual group
          const uint32_t NG = rdlane(gincl, FN - 1);
          // The item list's LDS is unused during extraction: slots 0-63
          // hold the tile's fragment-start markers, then each fragment's
          // first group, packed bounds and first bin, indexed by fragment.
          static_assert(kItemCap * 2 >= 4 * kWave * 4,
                        "the item list's LDS holds the tile markers and the "
                        "fragment table during extraction");
          uint32_t *const mk = reinterpret_cast<uint32_t *>(l.items);
          uint32_t *const ft = mk + kWave;
          if (lane < (int)FN) {
            ft[lane] = gsr;
            ft[kWave + lane] = meta;
            ft[2 * kWave + lane] = rb0;
          }
          for (uint32_t vb = 0; vb < NG; vb += kWave * kGroupTiles) {
            uint4 raw[kGroupTiles];
            uint32_t tg[kGroupTiles], tfr[kGroupTiles], tvm[kGroupTiles],
                tb0[kGroupTiles];
#pragma unroll
            for (int u = 0; u < kGroupTiles; ++u) {
              const uint32_t v0 = vb + (uint32_t)(u * kWave);
              const uint32_t v = v0 + lane;
              raw[u] = make_uint4(0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu);
              tg[u] = 0;
              tfr[u] = 0;
              tvm[u] = 0;  // entries of this group that belong to the fragment
              tb0[u] = 0;
              if (v0 < NG) {
                const uint32_t vmax = min(v0 + (uint32_t)kWave - 1, NG - 1);
                // v0 < NG: the first non-empty fragment of [fa, fb) has
                // gp = 0 <= v0 (the ones before it are empty), so lo_m != 0
                const uint64_t lo_m = ballot64(lane < (int)FN && ng > 0 && gp <= v0);
                const uint32_t flo = 63u - (uint32_t)__builtin_clzll(lo_m);
                // Fragment of this lane's group: the last non-empty one
                // starting at or before it (empty fragments share the next
                // one's start).  Non-empty fragments start at distinct
                // groups; those starting inside the tile mark 1 + their
                // index at their start's slot, and each lane takes the
                // marked slot at or below it (a prefix maximum; lane 0's
                // slot, never marked, holds 1 + flo).
                mk[lane] = lane == 0 ? flo + 1u : 0u;
                __builtin_amdgcn_wave_barrier();
                if (lane < (int)FN && ng > 0 && gp > v0 && gp <= vmax) mk[gp - v0] = (uint32_t)lane + 1u;
                __builtin_amdgcn_wave_barrier();
                const uint32_t f = wave_incl_max_dpp(mk[lane]) - 1u;
                const uint32_t fgp = (uint32_t)__shfl((int)gp, (int)f, kWave);
                const uint32_t fgs = ft[f];
                const uint32_t fmeta = ft[kWave + f];
                const uint32_t fb0 = ft[2 * kWave + f];
                __builtin_amdgcn_wave_barrier();
                if (v < NG) {
                  const uint32_t q = v - fgp;  // group within the fragment
                  tg[u] = fgs + q;
                  raw[u] = *reinterpret_cast<const uint4 *>(cw + 8 * (size_t)tg[u]);
                  tfr[u] = fa + f;
                  const uint32_t klo = (q == 0) ? (fmeta & 7u) : 0u;
                  const uint32_t khi = (q + 1 == (fmeta >> 6)) ? ((fmeta >> 3) & 7u) : 7u;
                  tvm[u] = (0xffu << klo) & (0xffu >> (7u - khi));
                  tb0[u] = fb0;
                }
              }
            }
            flush_out();
#pragma unroll
            for (int u = 0; u < kGroupTiles; ++u) {
              if (vb + (uint32_t)(u * kWave) >= NG) break;
              const uint32_t cyc[8] = {raw[u].x & 0xffffu, raw[u].x >> 16,
                                       raw[u].y & 0xffffu, raw[u].y >> 16,
                                       raw[u].z & 0xffffu, raw[u].z >> 16,
                                       raw[u].w & 0xffffu, raw[u].w >> 16};
              // Entry k is in the window when (cycle - c0) mod 2^16 < L:
              // cycles, c0 and c0 + L all fit 16 bits, so the wrapped
              // difference of a cycle below c0 is at least L.  Two entries
              // per packed u16 operation: L - d saturated at 0, then
              // min(., 1), leaves 1 in each half whose entry is in the
              // window; d == 0 marks an entry at c0.
              uint32_t hm, tmin;
              {
                const u16x2 c0v = {(uint16_t)c0, (uint16_t)c0};
                const u16x2 lv = {(uint16_t)L, (uint16_t)L};
                const u16x2 one = {1, 1};
                const uint32_t rw[4] = {raw[u].x, raw[u].y, raw[u].z, raw[u].w};
                u16x2 dmin = {0xffff, 0xffff};
                uint32_t t = 0;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                  const u16x2 d = __builtin_bit_cast(u16x2, rw[q]) - c0v;
                  dmin = __builtin_elementwise_min(dmin, d);
                  const u16x2 r = __builtin_elementwise_min(__builtin_elementwise_sub_sat(lv, d), one);
                  t |= __builtin_bit_cast(uint32_t, r) << (2 * q);
                }
                hm = (t & 0x55u) | ((t >> 15) & 0xaau);
                const uint32_t dm = __builtin_bit_cast(uint32_t, dmin);
                tmin = min(dm & 0xffffu, dm >> 16);
              }
              hm &= tvm[u];
              // the reference's binary_search(c0) starts at the last entry
              // equal to c0 in a bin: earlier ones are skipped.  tmin == 0
              // says that one of the eight entries (listed or not) is at c0;
              // the exact test per entry follows.
              if (ballot64(tmin == 0u && hm != 0u) != 0) {
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                  if (((hm >> k) & 1u) && cyc[k] == c0) {
                    const uint32_t e = 8 * tg[u] + k - bo;
                    const uint32_t next = k < 7 ? cyc[k < 7 ? k + 1 : 7]
                                                : (uint32_t)a.dia.cycles[base + e + 1];
                    if (skipped_duplicate(a.dia, o, tb0[u], e, c0, next))
                      hm &= ~(1u << k);
                  }
                }
              }
              // list slots in (lane, k) order = entry order.  When the
              // tile's hits fit the list, each lane writes its own hits
              // (one iteration per hit of the busiest lane); otherwise lanes
              // 0-31 first, then 32-63 (at most 32 x 8 hits each), so that
              // an overflow is met half a tile at a time.
              {
                uint32_t pre, tot;
                list_slots(hm, true, pre, tot);
                if (tot == 0) continue;
                if (nh + tot <= l.hcap) {
                  uint32_t pos = nh + pre, mm = hm;
                  const uint32_t key = (j << 18) | (tfr[u] << 12);
                  while (mm) {
                    const uint32_t k = (uint32_t)__builtin_ctz(mm);
                    mm &= mm - 1u;
                    const uint32_t w = k < 4u ? (k < 2u ? raw[u].x : raw[u].y)
                                              : (k < 6u ? raw[u].z : raw[u].w);
                    const uint32_t cy = (w >> ((k & 1u) * 16u)) & 0xffffu;
                    l.hits[pos] = make_uint2(8 * tg[u] + k - bo, key | (cy - c0));
                    ++pos;
                  }
                  nh += tot;
                  __builtin_amdgcn_wave_barrier();
                  if (!deferred) {
                    accumulate(false, base, 0, nh);
                    nh = 0;
                  }
                  continue;
                }
              }
#pragma unroll
              for (int hh = 0; hh < 2; ++hh) {
                uint32_t pre, tot;
                list_slots(hm, (lane >> 5) == hh, pre, tot);
                if (tot == 0) continue;
                if (deferred && nh + tot > l.hcap) {
                  if (nobs != 1) return fa;
                  // lane 32 hh scans this half-tile's first fragment
                  const uint32_t fcut = rdlane(tfr[u], 32 * hh);
                  uint32_t keep = 0;
                  for (uint32_t c = 0; c < nh; c += kWave) {
                    const uint32_t idx = c + lane;
                    const bool kp = idx < nh && ((l.hits[idx].y >> 12) & 63u) < fcut;
                    keep += (uint32_t)__builtin_popcountll(ballot64(kp));
                  }
                  nh = keep;
                  return fcut;
                }
                if ((lane >> 5) == hh) {
                  uint32_t pos = nh + pre;
#pragma unroll
                  for (int k = 0; k < 8; ++k) {
                    if ((hm >> k) & 1u) {
                      const uint32_t e = 8 * tg[u] + k - bo;
                      l.hits[pos] = make_uint2(e, (j << 18) | (tfr[u] << 12) | (cyc[k] - c0));
                      ++pos;
                    }
                  }
                }
                nh += tot;
                __builtin_amdgcn_wave_barrier();
                if (!deferred && nh > l.hcap - (uint32_t)kTileHits) {
                  accumulate(false, base, 0, nh);
                  nh = 0;
                }
              }
            }
          }
          if (!deferred && nh) {
            accumulate(false, base, 0, nh);
            nh = 0;
          }
          __builtin_amdgcn_wave_barrier();
        }
        return fb;
      };
      auto zero_group = [&](uint32_t G) {
        float4 *xz = reinterpret_cast<float4 *>(l.X);
        const uint32_t n4 = G * l.lp / 4;
        for (uint32_t i = lane; i < n4; i += kWave)
          xz[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        for (uint32_t i = lane; i < G * l.wb; i += kWave) l.BD[i] = 0ull;
        __builtin_amdgcn_wave_barrier();
      };
      // The hits of fragments [0, fstop) in one pass, their intensities in
      // one batch; fragments [fstop, nsel), which did not fit the list, are
      // extracted directly, group by group.
      uint32_t fstop = 0;
      if (nobs > 0 && nsel > 0) {
        nh = 0;
        fstop = extract(0, nsel, true);
        uint64_t wb[kObsMax];
#pragma unroll
        for (int j = 0; j < kObsMax; ++j) wb[j] = win_base((uint32_t)j);
        for (uint32_t c = 0; c < nh; c += kWave * 4) {
          float iv[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const uint32_t idx = c + (uint32_t)(u * kWave) + lane;
            iv[u] = 0.0f;
            if (idx < nh) {
              const uint2 h = l.hits[idx];
              const uint32_t j = h.y >> 18;
              const uint64_t base = j == 0 ? wb[0] : j == 1 ? wb[1] : j == 2 ? wb[2] : wb[3];
              iv[u] = a.dia.intens[base + h.x];
            }
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const uint32_t idx = c + (uint32_t)(u * kWave) + lane;
            if (idx < nh) l.hits[idx].x = __builtin_bit_cast(uint32_t, iv[u]);
          }
        }
        __builtin_amdgcn_wave_barrier();
      }

      if (more) stage2();
      // the list is kept whole while the tile is in use; fragments that did
      // not fit it are extracted through the whole list capacity
      carve_tile(l, fstop < nsel ? l.hcap : nh, nsel);
      // A row whose hits are all listed takes the block-sparse tile when its
      // cell blocks fit: every fragment in one group.  Each hit marks its
      // cell block; the first hit of a block (an LDS atomic's old value)
      // gives it the next storage slot, in list order.  Otherwise the dense
      // tile, gcap fragments at a time.
      if (HALF == 10 && nobs > 0 && nsel > 0 && fstop == nsel) {
        st = carve_sparse(l, nsel);
        if (st.cap >= 2u) {
          const uint32_t w64 = 2u * nsel * l.wb;  // NZB and BD
          for (uint32_t i = lane; i < w64; i += kWave) st.NZB[i] = 0ull;
          uint2 *stz = reinterpret_cast<uint2 *>(st.ST);
          for (uint32_t i = lane; i < nsel * st.sts / 8u; i += kWave) stz[i] = make_uint2(0u, 0u);
          if (lane == 0) st.CB[0] = make_float4(0.f, 0.f, 0.f, 0.f);
          __builtin_amdgcn_wave_barrier();
          uint32_t ns = 0;  // slots given out (slot 0 excluded)
          for (uint32_t c = 0; c < nh; c += kWave) {
            const uint32_t idx = c + (uint32_t)lane;
            const bool v = idx < nh;
            const uint32_t key = v ? l.hits[idx].y : 0u;
            const uint32_t f = (key >> 12) & 63u, be = (key & 0xfffu) >> 2;
            uint64_t old = ~0ull;
            if (v)
              old = atomicOr((unsigned long long *)&st.NZB[f * l.wb + (be >> 6)],
                             1ull << (be & 63u));
            const bool first = v && ((old >> (be & 63u)) & 1ull) == 0ull;
            const uint64_t fb = ballot64(first);
            const uint32_t slot = ns + 1u +
                __builtin_amdgcn_mbcnt_hi((uint32_t)(fb >> 32),
                                          __builtin_amdgcn_mbcnt_lo((uint32_t)fb, 0u));
            if (first && slot < st.cap) {
              st.ST[f * st.sts + be + 4u] = (uint8_t)slot;
              st.CB[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
            ns += (uint32_t)__builtin_popcountll(fb);
          }
          __builtin_amdgcn_wave_barrier();
          spm = ns < st.cap;
        }
      }
      // With one window the list is in fragment order, so each group's
      // hits follow the previous group's: [hcur, hend)
      uint32_t hcur = 0;
      for (uint32_t g0 = 0, G = 0; g0 < nsel && nobs > 0; g0 += G) {
        const bool listed = g0 < fstop;
        G = spm ? nsel : min(l.gcap, (listed ? fstop : nsel) - g0);
        if (!spm) zero_group(G);
        gA = g0;
        gG = G;
        if (listed && nobs == 1 && !spm) {
          uint32_t hend = hcur;
          for (;;) {
            const uint32_t idx = hend + (uint32_t)lane;
            const bool in = idx < nh && ((l.hits[idx].y >> 12) & 63u) < g0 + G;
            const uint32_t cnt = (uint32_t)__builtin_popcountll(ballot64(in));
            hend += cnt;
            if (cnt < (uint32_t)kWave) break;
          }
          accumulate(true, 0, hcur, hend);
          hcur = hend;
        } else if (listed) {
          accumulate(true, 0, 0, nh);
        } else {
          nh = 0;
          extract(g0, g0 + G, false);
        }
        __builtin_amdgcn_wave_barrier();
        // ---- compaction of (fragment, block) items, fragment-major -------
        uint32_t nitems = 0;
        auto flush = [&](uint32_t n) {
          // conv + ln + weighted score, 1 block per lane per pass; score
          // updates are applied fragment by fragment (reference row order)
          for (uint32_t c = 0; c < n; c += kWave) {
            if (HALF == 10 && !spm && n - c <= (uint32_t)kTailItems) {
              // The last items, one output per lane (4 lanes per item): a
              // pass of 64 blocks for them would leave most lanes idle.
              // Each output is the block pass's chain for that output.
              const uint32_t no = 4 * (n - c);
              for (uint32_t c2 = 0; c2 < no; c2 += kWave) {
                const uint32_t oi = c2 + lane;
                const bool v = oi < no;
                const uint32_t item = l.items[c + (v ? oi >> 2 : 0u)];
                const uint32_t f = item >> 10;
                const uint32_t pos = (item & 0x3ffu) * 4 + (oi & 3u);
                const float *xc = l.X + f * l.lp + kXPad + pos;
                float y = xc[0] * k10[10];
#pragma unroll
                for (int j = 0; j < 10; ++j) {
                  const int d = 10 - j;
                  y = y + (xc[-d] + xc[d]) * k10[j];
                }
                const bool act = v && pos >= half && pos + half < L && y != 0.0f;
                const float wf = l.w[g0 + f];
                float t = 0.0f;
                if (ballot64(act && !logf_is_normal(y + 1.0f)) == 0) {
                  const float lq = logf_normal(y + 1.0f, ltab);
                  if (act) t = lq * wf;
                } else {
                  if (act) t = logf_exact(y + 1.0f, ltab) * wf;
                }
                const uint32_t flo = rdlane(item >> 10, 0);
                const uint32_t fhi = rdlane(item >> 10, min(no - c2, (uint32_t)kWave) - 1);
                for (uint32_t ff = flo; ff <= fhi; ++ff) {
                  if (act && f == ff) l.R[pos] = l.R[pos] + t;
                  __builtin_amdgcn_wave_barrier();
                }
              }
              break;
            }
            const uint32_t idx = c + lane;
            const bool v = idx < n;
            const uint32_t item = v ? l.items[idx] : l.items[c];
            const uint32_t f = item >> 10, i0 = (item & 0x3ffu) * 4;
            const float *row = l.X + f * l.lp + kXPad;
            float y[4];
            if (HALF == 10) {
              float4 xb[7];  // cells i0 - 12 .. i0 + 15
              if (spm) {
                // slots of cell blocks b - 3 .. b + 3: ST bytes sa .. sa + 6
                const uint32_t sa = f * st.sts + (item & 0x3ffu) + 1u;
                const uint32_t *stw = reinterpret_cast<const uint32_t *>(st.ST) + (sa >> 2);
                const uint32_t w0 = stw[0], w1 = stw[1], w2 = stw[2];
                const uint32_t u0 = __builtin_amdgcn_alignbyte(w1, w0, sa & 3u);
                const uint32_t u1 = __builtin_amdgcn_alignbyte(w2, w1, sa & 3u);
#pragma unroll
                for (int j = 0; j < 7; ++j)
                  xb[j] = st.CB[((j < 4 ? u0 : u1) >> (8 * (j & 3))) & 0xffu];
              } else {
                const float4 *xs = reinterpret_cast<const float4 *>(row + i0 - 12);
#pragma unroll
                for (int j = 0; j < 7; ++j) xb[j] = xs[j];
              }
              conv4_h10(xb, k10, y);
            } else {
              conv4_any(row, i0, half, l.kern, y);
            }
            const float wf = l.w[g0 + f];
            // outputs outside [half, L - half) stay zero (scalar.rs:30-38)
            float t[4], u[4];
            bool act[4];
            // all four arguments of ln positive normal floats: their bit
            // patterns, as signed integers, lie in [0x00800000, 0x7f800000)
            // (negative values, zeros, subnormals, infinities and NaN fall
            // outside); otherwise the exact ln handles every case
            int blo = 0x7fffffff, bhi = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const uint32_t pos = i0 + q;
              const bool inside = pos >= half && pos + half < L;
              act[q] = v && inside && y[q] != 0.0f;
              u[q] = y[q] + 1.0f;
              const int b = __builtin_bit_cast(int, u[q]);
              blo = min(blo, b);
              bhi = max(bhi, b);
            }
            const bool odd = v && (blo < 0x00800000 || bhi >= 0x7f800000);
            if (ballot64(odd) == 0) {
#pragma unroll
              for (int q = 0; q < 4; ++q) {
                const float lq = logf_normal(u[q], ltab);
                t[q] = act[q] ? lq * wf : 0.0f;
              }
            } else {
#pragma unroll
              for (int q = 0; q < 4; ++q)
                t[q] = act[q] ? logf_exact(u[q], ltab) * wf : 0.0f;
            }
            const uint32_t flo = rdlane(item >> 10, 0);
            const uint32_t last = min(n - c, (uint32_t)kWave) - 1;
            const uint32_t fhi = rdlane(item >> 10, last);
            for (uint32_t ff = flo; ff <= fhi; ++ff) {
              if (v && f == ff) {
                float4 *rp = reinterpret_cast<float4 *>(l.R + i0);
                float4 r = *rp;
                r.x = r.x + t[0];
                r.y = r.y + t[1];
                r.z = r.z + t[2];
                r.w = r.w + t[3];
                *rp = r;
              }
              __builtin_amdgcn_wave_barrier();
            }
          }
          __builtin_amdgcn_wave_barrier();
        };
        // The group's block words (G x wb, fragment-major; a row of L cycles
        // has ceil(L / 256) = wb words) are read 64 at a time, one per lane,
        // and then taken from their lanes two at a time: lane l writes the
        // items of bit l of both words, with no branch.  A lane whose bit is
        // clear writes to the first slot after the pair's items, which a
        // later pair overwrites or no pass reads (nitems stays below
        // kItemCap - 2 x 64 between pairs, so that slot is inside the list).
        const uint32_t nwords = G * l.wb;
        const uint64_t *const bdw = spm ? st.BD : l.BD;
        uint32_t wlo = 0, whi = 0;
        uint32_t fa = 0, qa = 0;  // fragment and word of word wq
        for (uint32_t wq = 0; wq < nwords; wq += 2) {
          const uint32_t j = wq & (uint32_t)(kWave - 1);
          if (j == 0u) {
            const uint32_t wi = wq + (uint32_t)lane;
            const uint64_t bw = wi < nwords ? bdw[wi] : 0ull;
            wlo = (uint32_t)bw;
            whi = (uint32_t)(bw >> 32);
          }
          // word wq + 1 (j + 1 <= 63: j is even); zero past the last word
          const bool two = wq + 1u < nwords;
          const uint32_t fb = qa + 1u == l.wb ? fa + 1u : fa;
          const uint32_t qb = qa + 1u == l.wb ? 0u : qa + 1u;
          const uint32_t alo = rdlane(wlo, j), ahi = rdlane(whi, j);
          const uint32_t blo = two ? rdlane(wlo, j + 1u) : 0u;
          const uint32_t bhi = two ? rdlane(whi, j + 1u) : 0u;
          if ((alo | ahi | blo | bhi) != 0u) {
            const uint64_t A = ((uint64_t)ahi << 32) | alo;
            const uint64_t B = ((uint64_t)bhi << 32) | blo;
            const uint32_t na = (uint32_t)__builtin_popcountll(A);
            const uint32_t nb = (uint32_t)__builtin_popcountll(B);
            const uint32_t spare = nitems + na + nb;
            const uint32_t pa = nitems + __builtin_amdgcn_mbcnt_hi(ahi, __builtin_amdgcn_mbcnt_lo(alo, 0u));
            const uint32_t pb =
                nitems + na + __builtin_amdgcn_mbcnt_hi(bhi, __builtin_amdgcn_mbcnt_lo(blo, 0u));
            l.items[((A >> lane) & 1ull) ? pa : spare] = (uint16_t)((fa << 10) | (qa * 64 + lane));
            l.items[((B >> lane) & 1ull) ? pb : spare] = (uint16_t)((fb << 10) | (qb * 64 + lane));
            nitems = spare;
            if (nitems > (uint32_t)(kItemCap - 2 * kWave - 1)) {
              __builtin_amdgcn_wave_barrier();
              flush(nitems);
              nitems = 0;
            }
          }
          // advance two words
          fa = qb + 1u == l.wb ? fb + 1u : fb;
          qa = qb + 1u == l.wb ? 0u : qb + 1u;
        }
        __builtin_amdgcn_wave_barrier();
        if (nitems) flush(nitems);
      }
      __builtin_amdgcn_wave_barrier();
      if (a.dbg_scores && row >= a.dbg_row0 && row < a.dbg_row0 + a.dbg_rows &&
          L <= a.dbg_stride) {
        float *out = a.dbg_scores + (size_t)(row - a.dbg_row0) * a.dbg_stride;
        for (uint32_t i = opaque_lane(); i < L; i += kWave) out[i] = l.R[i];
      }
      // ---- local maxima and top-k: mod.rs:23-64, 167-192 -----------------
      // lane l tests positions 256m + 4l + k (k < 4) from three float4
      // reads; bit 4m + k of ismax marks a strict 5-point maximum.  Each
      // lane keeps its best maximum (lbv, lbi): the largest score, ties to
      // the lower position (positions are visited in increasing order).
      uint64_t ismax = 0;
      float lbv = 0.0f;
      uint32_t lbi = 0xffffffffu;
      for (uint32_t m = 0; 256 * m < L; ++m) {
        const uint32_t i0 = 256 * m + 4 * lane;
        if (i0 < L) {
          const float4 pa = *reinterpret_cast<const float4 *>(l.R + i0 - 4);
          const float4 pb = *reinterpret_cast<const float4 *>(l.R + i0);
          const float4 pc = *reinterpret_cast<const float4 *>(l.R + i0 + 4);
          const float w[12] = {pa.x, pa.y, pa.z, pa.w, pb.x, pb.y,
                               pb.z, pb.w, pc.x, pc.y, pc.z, pc.w};
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const uint32_t i = i0 + k;
            const float v = w[4 + k];
            if (i >= 2 && i + 2 < L && w[2 + k] < w[3 + k] && w[3 + k] < v &&
                v > w[5 + k] && w[5 + k] > w[6 + k]) {
              ismax |= 1ull << (4 * m + k);
              if (lbi == 0xffffffffu || v > lbv) {
                lbv = v;
                lbi = i;
              }
            }
          }
        }
      }
      if (more) stage3();
      // ranks: descending score, ties to the lower cycle (stable sort).  A
      // maximum is never NaN (it compares greater than its neighbours), so
      // the order key ranks the lanes' best maxima exactly; only the lane
      // that supplied a rank looks for its next best.
      flush_out();  // rows without extraction still own pending outputs
      uint32_t ncand = 0, apex = 0;
      for (uint32_t r = 0; r < cc; ++r) {
        const uint32_t bi =
            wave_argmax_key(lbi == 0xffffffffu ? 0u : f32_order_key(lbv), lbi);
        if (bi == 0xffffffffu) break;
        if (r == 0) apex = bi;
        const uint32_t wl = (bi & 255u) >> 2;
        const float bv = __builtin_bit_cast(
            float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, lbv), (int)wl));
        if ((uint32_t)lane == wl) {
          ismax &= ~(1ull << (((bi >> 8) << 2) | (bi & 3u)));
          lbi = 0xffffffffu;
          uint64_t it = ismax;
          while (it) {
            const uint32_t b = __builtin_ctzll(it);
            it &= it - 1;
            const uint32_t i = 256 * (b >> 2) + 4 * lane + (b & 3);
            const float v = l.R[i];
            if (lbi == 0xffffffffu || v > lbv) {
              lbv = v;
              lbi = i;
            }
          }
        }
        if (r < (uint32_t)kWave) {
          if ((uint32_t)lane == r) {
            out_s = bv;
            out_c = c0 + bi;
          }
        } else if (lane == 0) {  // ranks beyond one per lane: stored now
          a.cand_score[(size_t)row * cc + r] = bv;
          a.cand_center[(size_t)row * cc + r] = c0 + bi;
        }
        ++ncand;
      }
      if (HALF == 10 && SAVE) {
        // differentiable forward: a row in the block-sparse tile leaves the
        // XIC cells around its rank-0 apex for the backward pass, which
        // re-extracts the other rows' windows itself
        const bool saved = spm && ncand > 0;
        if (saved) {
          // fragment-major, 2 * half + 1 cells from apex - half; two
          // fragments per pass, one per half-wave
          float *xo = a.xwin + (size_t)row * a.xwin_stride;
          const int t0 = (int)apex - (int)half;
          const uint32_t wlen = 2 * half + 1;
          for (uint32_t f0 = 0; f0 < nsel; f0 += 2) {
            const uint32_t f = f0 + ((uint32_t)lane >> 5), t = (uint32_t)lane & 31u;
            if (f < nsel && t < wlen) {
              const int cell = t0 + (int)t;
              float v = 0.0f;
              if (cell >= 0 && cell < (int)L) {
                const uint32_t slot = st.ST[f * st.sts + ((uint32_t)cell >> 2) + 4u];
                v = reinterpret_cast<const float *>(st.CB + slot)[cell & 3];
              }
              xo[f * wlen + t] = v;
            }
          }
        }
        a.xwin_ok[row] = saved ? 1u : 0u;  // every lane, the same value
      }
      out_row = row;
      out_n = ncand;
      __builtin_amdgcn_wave_barrier();
    }
  }
  flush_out();
}

// ---------------------------------------------------------------------------
// order-preserving compaction
// ---------------------------------------------------------------------------
constexpr int kScanBlock = 1024;  // rows per scan block (256 threads x 4)

__global__ __launch_bounds__(256) void scan_local_kernel(
    const uint32_t *counts, uint32_t n, uint32_t *local, uint32_t *bsum) {
  __shared__ uint32_t wsum[4];
  const uint32_t base = blockIdx.x * kScanBlock + threadIdx.x * 4;
  uint32_t v[4], s = 0;
  for (int k = 0; k < 4; ++k) {
    v[k] = (base + k < n) ? counts[base + k] : 0u;
    s += v[k];
  }
  const uint32_t incl = wave_incl_scan(s);
  const int wv = threadIdx.x / kWave;
  if (lane_id() == kWave - 1) wsum[wv] = incl;
  __syncthreads();
  uint32_t off = 0;
  for (int k = 0; k < wv; ++k) off += wsum[k];
  uint32_t run = off + incl - s;
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) local[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 255) bsum[blockIdx.x] = off + incl;
}

__global__ __launch_bounds__(1024) void scan_blocks_kernel(uint32_t *bsum,
                                                           uint32_t nb,
                                                           uint32_t *total) {
  __shared__ uint32_t wsum[16];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t b0 = 0; b0 < nb; b0 += 1024) {
    const uint32_t i = b0 + threadIdx.x;
    const uint32_t v = i < nb ? bsum[i] : 0u;
    const uint32_t incl = wave_incl_scan(v);
    const int wv = threadIdx.x / kWave;
    if (lane_id() == kWave - 1) wsum[wv] = incl;
    __syncthreads();
    uint32_t off = carry;
    for (int k = 0; k < wv; ++k) off += wsum[k];
    if (i < nb) bsum[i] = off + incl - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry = off + incl;
    __syncthreads();
  }
  if (threadIdx.x == 0) total[0] = carry;
}

struct CandRow {
  uint32_t row, rank;
  float score;
  uint32_t center;
};

__global__ __launch_bounds__(256) void scatter_kernel(
    const uint32_t *counts, const uint32_t *local, const uint32_t *bsum,
    uint32_t n, uint32_t cc, const float *cs, const uint32_t *cctr,
    CandRow *out) {
  const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  const uint32_t k = counts[row];
  const uint32_t off = local[row] + bsum[row / kScanBlock];
  for (uint32_t r = 0; r < k; ++r) {
    CandRow c;
    c.row = row;
    c.rank = r;
    c.score = cs[(size_t)row * cc + r];
    c.center = cctr[(size_t)row * cc + r];
    out[off + r] = c;
  }
}

// Chunk k of a pipelined search (rows [r0, r0 + n), r0 a multiple of
// kScanBlock): the chunk's candidate rows go after those of chunks 0..k-1,
// whose counts tot[0..k) the scans of those chunks wrote.
__global__ __launch_bounds__(256) void scatter_chunk_kernel(
    const uint32_t *counts, const uint32_t *local, const uint32_t *bsum_k,
    uint32_t r0, uint32_t n, uint32_t cc, const float *cs, const uint32_t *cctr,
    const uint32_t *tot, uint32_t k, CandRow *out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t base = 0;
  for (uint32_t j = 0; j < k; ++j) base += tot[j];
  const uint32_t row = r0 + i;
  const uint32_t cnt = counts[row];
  const uint32_t off = base + local[row] + bsum_k[i / kScanBlock];
  for (uint32_t r = 0; r < cnt; ++r) {
    CandRow c;
    c.row = row;
    c.rank = r;
    c.score = cs[(size_t)row * cc + r];
    c.center = cctr[(size_t)row * cc + r];
    out[off + r] = c;
  }
}

// A candidate row kept in HBM: the library's precursor_idx resolved, so the
// table outlives the library handle.
struct DevCand {
  uint64_t pidx;
  uint32_t rank;
  float score;
  uint32_t center;
  uint32_t pad;
};

__global__ __launch_bounds__(256) void resolve_kernel(const CandRow *rows,
                                                      uint32_t n,
                                                      const uint64_t *p_idx,
                                                      DevCand *out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const CandRow r = rows[i];
  DevCand c;
  c.pidx = p_idx[r.row];
  c.rank = r.rank;
  c.score = r.score;
  c.center = r.center;
  c.pad = 0;
  out[i] = c;
}

// CandidateCollection::to_arrays (src/candidate/entry.rs:124-173) into
// device columns; scan_* are 0 from selection (entry.rs:31-33) and the
// cycle bounds follow src/peak_group_selection/mod.rs:174-180.
__global__ __launch_bounds__(256) void export_kernel(
    const DevCand *rows, uint64_t n, uint64_t pl, uint64_t nrt, uint64_t *pidx,
    uint64_t *rank, float *score, uint64_t *scan_center, uint64_t *scan_start,
    uint64_t *scan_stop, uint64_t *cycle_center, uint64_t *cycle_start,
    uint64_t *cycle_stop) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const DevCand c = rows[i];
  const uint64_t center = c.center;
  if (pidx) pidx[i] = c.pidx;
  if (rank) rank[i] = c.rank;
  if (score) score[i] = c.score;
  if (scan_center) scan_center[i] = 0;
  if (scan_start) scan_start[i] = 0;
  if (scan_stop) scan_stop[i] = 0;
  if (cycle_center) cycle_center[i] = center;
  if (cycle_start) cycle_start[i] = center - pl;  // usize arithmetic, mod.rs:176
  if (cycle_stop) cycle_stop[i] = min(center + pl + 1, nrt);
}

// The host table's block (eight u64 columns, then the f32 score; see
// adsr_candidates) for the compacted rows, n = *total read on the device so
// that the kernel is queued before the host knows the count.  Columns as in
// export_kernel.
__global__ __launch_bounds__(256) void table_kernel(const CandRow *rows,
                                                    const uint32_t *total,
                                                    const uint64_t *p_idx,
                                                    uint64_t pl, uint64_t nrt,
                                                    uint64_t cap, uint64_t *blk) {
  const uint64_t n = min((uint64_t)*total, cap);
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const CandRow r = rows[i];
  const uint64_t center = r.center;
  blk[i] = p_idx[r.row];
  blk[n + i] = r.rank;
  blk[2 * n + i] = 0;
  blk[3 * n + i] = 0;
  blk[4 * n + i] = 0;
  blk[5 * n + i] = center;
  blk[6 * n + i] = center - pl;  // usize arithmetic, mod.rs:176
  blk[7 * n + i] = min(center + pl + 1, nrt);
  reinterpret_cast<float *>(blk + 8 * n)[i] = r.score;
}

// Chunk k's rows of the host table, in a block whose columns are `stride`
// rows apart (the rows of chunks 0..k-1 come first), so that each chunk's
// column segments can be copied as soon as the chunk is done.
__global__ __launch_bounds__(256) void table_chunk_kernel(const CandRow *rows,
                                                          const uint32_t *tot, uint32_t k,
                                                          const uint64_t *p_idx, uint64_t pl,
                                                          uint64_t nrt, uint64_t stride,
                                                          uint64_t *blk) {
  uint64_t base = 0;
  for (uint32_t j = 0; j < k; ++j) base += tot[j];
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= tot[k]) return;
  const uint64_t q = base + i;
  const CandRow r = rows[q];
  const uint64_t center = r.center;
  blk[q] = p_idx[r.row];
  blk[stride + q] = r.rank;
  blk[2 * stride + q] = 0;
  blk[3 * stride + q] = 0;
  blk[4 * stride + q] = 0;
  blk[5 * stride + q] = center;
  blk[6 * stride + q] = center - pl;  // usize arithmetic, mod.rs:176
  blk[7 * stride + q] = min(center + pl + 1, nrt);
  reinterpret_cast<float *>(blk + 8 * stride)[q] = r.score;
}

// The nine to_arrays columns of the compacted rows straight into caller
// device buffers of `cap` rows (the no-allocation search): n = *total read
// on the device; rows past cap are not written (the host reports the
// shortfall).  p_idx starts at the searched rows' first library row.
__global__ __launch_bounds__(256) void columns_kernel(
    const CandRow *rows, const uint32_t *total, const uint64_t *p_idx, uint64_t pl,
    uint64_t nrt, uint64_t cap, uint64_t *pidx, uint64_t *rank, float *score,
    uint64_t *scan_center, uint64_t *scan_start, uint64_t *scan_stop,
    uint64_t *cycle_center, uint64_t *cycle_start, uint64_t *cycle_stop) {
  const uint64_t n = min((uint64_t)*total, cap);
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const CandRow r = rows[i];
  const uint64_t center = r.center;
  pidx[i] = p_idx[r.row];
  rank[i] = r.rank;
  score[i] = r.score;
  scan_center[i] = 0;
  scan_start[i] = 0;
  scan_stop[i] = 0;
  cycle_center[i] = center;
  cycle_start[i] = center - pl;  // usize arithmetic, mod.rs:176
  cycle_stop[i] = min(center + pl + 1, nrt);
}

// Algorithmic bytes per (row, selected fragment): slice_starts of the bin
// range, every CSR entry's cycle index, the intensity of entries in the RT
// window, per valid window.  Reduced with one atomic per block.
__global__ __launch_bounds__(256) void traffic_kernel(KDia dia, KLib lib,
                                                      KParams prm, KPlan plan,
                                                      unsigned long long *acc) {
  __shared__ unsigned long long part[3];
  if (threadIdx.x < 3) part[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t row = t / plan.ks;
  const uint32_t slot = (uint32_t)(t - row * plan.ks);
  unsigned long long bytes = 0, ents = 0, hits = 0;
  if (row < lib.n_prec) {
    const uint4 rec = plan.rec[row];
    const uint32_t nsel = rec.z & 0xffffu, nobs = rec.z >> 16;
    if (slot == 0)
      bytes += 24 + 10ull * (lib.f_stop[row] - lib.f_start[row]);
    if (slot < nsel && rec.y > rec.x) {
      const float mz = plan.sel_mz[row * plan.ks + slot];
      const float delta = mz * prm.mass_tol * 1e-6f;
      uint32_t start, end;
      mz_range(dia, mz - delta, mz + delta, start, end);
      for (uint32_t j = 0; j < nobs; ++j) {
        const size_t idx = (row * kObsMax + j) * plan.ks + slot;
        const uint2 r = plan.rng[idx];
        const uint32_t o = plan.obs[row * kObsMax + j];
        bytes += 4ull * (end - start + 1);
        const uint16_t *cy = dia.cycles + dia.entry_base[o];
        for (uint32_t e = r.x; e < r.y; ++e) {
          ++ents;
          const uint32_t c = cy[e];
          if (c >= rec.x && c < rec.y) ++hits;
        }
      }
    }
  }
  bytes += 2 * ents + 4 * hits;
  atomicAdd(&part[0], bytes);
  atomicAdd(&part[1], ents);
  atomicAdd(&part[2], hits);
  __syncthreads();
  if (threadIdx.x < 3) atomicAdd(&acc[threadIdx.x], part[threadIdx.x]);
}

// ---------------------------------------------------------------------------
// differentiable selection: backward through the convolution and the
// weighted ln(1+x) score at each precursor's rank-0 apex
// ---------------------------------------------------------------------------
struct BwdArgs {
  KDia dia;
  KLib lib;
  KParams prm;
  KPlan plan;
  uint32_t n_rows;
  uint32_t cc;
  const uint32_t *cand_center;
  const uint32_t *cand_count;
  const float *grad_s;  // dLoss/ds_star per row
  const double *dk;     // d(kernel tap)/d(sigma), (half+1) per kernel index
  float *grad_w;        // per fragment
  double *contrib;      // per row: dLoss/ds_star * ds_star/dsigma
  uint32_t *queue;
  const float *xwin;    // apex XIC windows left by the forward (or nullptr)
  const uint32_t *xwin_ok;
  uint32_t xwin_stride;
};


__global__ void fill_f32_kernel(float *p, uint64_t n, float v) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

// Hinge loss over (target, decoy) pairs of rank-0 scores: terms[p] =
// max(0, margin - (s[t] - s[d])), NaN kept (torch.clamp), and the loss's
// gradient added to g: -1 to the target and +1 to the decoy of every pair
// with margin - (s[t] - s[d]) >= 0 (torch.clamp passes the gradient where
// its input >= min).  The +-1 float atomics are exact in any order while a
// row's count stays below 2^24.  Pairs with an index outside [0, n) get a
// zero term and no gradient.
__global__ __launch_bounds__(256) void hinge_kernel(const float *s, uint64_t n,
                                                    const int64_t *t, const int64_t *d,
                                                    uint64_t np, float margin,
                                                    float *terms, float *g) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += stride) {
    const int64_t ti = t[p], di = d[p];
    if (ti < 0 || di < 0 || (uint64_t)ti >= n || (uint64_t)di >= n) {
      terms[p] = 0.0f;
      continue;
    }
    const float m = margin - (s[ti] - s[di]);
    terms[p] = m != m ? m : (m > 0.0f ? m : 0.0f);
    if (m >= 0.0f) {
      atomicAdd(g + ti, -1.0f);
      atomicAdd(g + di, 1.0f);
    }
  }
}

__global__ void sstar_kernel(const float *cs, const uint32_t *cnt, uint32_t cc,
                             uint32_t n, float *s_star, uint8_t *has) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const bool h = cnt[r] > 0;
  s_star[r] = h ? cs[(size_t)r * cc] : 0.0f;
  if (has) has[r] = h ? 1 : 0;
}

constexpr int kBwdHitCap = 256;
constexpr int kBwdPreFrag = 24;  // saved-window fragments loaded with the row's first loads

// Bin range [st, en) of one fragment m/z (src/quadrupole_observation/mod.rs:
// 79-83, src/mz_index/mod.rs:95-127) from the bucket entries of its bounds.
__device__ __forceinline__ void bin_range(const KDia &d, float mz, float tol,
                                          uint32_t &st, uint32_t &en) {
  const float delta = mz * tol * 1e-6f;
  const float lo = mz - delta, hi = mz + delta;
  const uint32_t n = d.n_grid;
  st = !(lo > d.grid_first) ? 0u
       : (lo > d.grid_last) ? n
                            : lb_from(d.bucket2[bucket_key(d, lo)], lo);
  const uint32_t ub = !(hi >= d.grid_first) ? 0u
                      : (hi >= d.grid_last) ? n
                                            : ub_from(d.bucket2[bucket_key(d, hi)], hi);
  en = st < n ? max(st, ub) : st;
}

// Per-wave LDS of the backward: XIC window per selected fragment, its
// duplicate detector, the hit list and the per-fragment sigma terms.
__host__ __device__ inline uint32_t bwd_lds_bytes(uint32_t ks) {
  return (ks * kWin * 4 + ks * 8 + kBwdHitCap * 8 + kWave * 4 + ks * kWin + 15) / 16 * 16;
}

// One wave per precursor (launched over all rows; rows without a candidate
// or with a zero loss gradient leave at once).
__global__ __launch_bounds__(256) void diff_backward_kernel(BwdArgs a) {
  extern __shared__ __align__(16) char bsm[];
  const int lane = lane_id();
  const uint32_t ks = a.plan.ks;
  const uint32_t row = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  char *base_lds = bsm + (size_t)(threadIdx.x / kWave) * bwd_lds_bytes(ks);
  float *xw = reinterpret_cast<float *>(base_lds);
  double *part = reinterpret_cast<double *>(base_lds + ks * kWin * 4);
  uint2 *hits = reinterpret_cast<uint2 *>(base_lds + ks * kWin * 4 + ks * 8);
  // fragment-start markers of a tile (see the select kernel's extraction)
  uint32_t *mk = reinterpret_cast<uint32_t *>(base_lds + ks * kWin * 4 + ks * 8 +
                                              kBwdHitCap * 8);
  uint8_t *own = reinterpret_cast<uint8_t *>(base_lds + ks * kWin * 4 + ks * 8 +
                                             kBwdHitCap * 8 + kWave * 4);
  if (row >= a.n_rows) return;  // whole wave
  const uint32_t half = a.prm.half;
  // The row's loads are issued in three rounds of independent loads: what
  // the row index gives, then what those values give (bucket entries,
  // window bases, kernel taps), then the first window's CSR ranges.
  const uint4 rec = a.plan.rec[row];
  const uint32_t g_raw = __builtin_bit_cast(uint32_t, a.grad_s[row]);
  const uint32_t ncand_raw = a.cand_count[row];
  const uint32_t cen_raw = a.cand_center[(size_t)row * a.cc];
  const uint32_t ok_raw = a.xwin_ok ? a.xwin_ok[row] : 0u;
  // The apex window the forward may have left, loaded with the row's first
  // loads rather than after its flag: fragments 2p + (lane >> 5), cells
  // lane & 31, for the first kBwdPreFrag fragments of the stride.
  const uint32_t wlen = 2 * a.prm.half + 1;
  const uint32_t hl = (uint32_t)lane >> 5, tl = (uint32_t)lane & 31u;
  float pre[kBwdPreFrag / 2];
#pragma unroll
  for (int q = 0; q < kBwdPreFrag / 2; ++q) {
    const uint32_t f = 2u * q + hl;
    pre[q] = a.xwin && f < ks && tl < wlen ? a.xwin[(size_t)row * a.xwin_stride + f * wlen + tl]
                                           : 0.0f;
  }
  const bool fl = lane < (int)ks;
  const float smz_l = fl ? a.plan.sel_mz[(size_t)row * ks + lane] : 0.0f;
  const uint32_t sel_l = fl ? (uint32_t)a.plan.sel[(size_t)row * ks + lane] : 0u;
  const float selw_l = fl ? a.plan.sel_w[(size_t)row * ks + lane] : 0.0f;
  const uint32_t obs_l = lane < kObsMax ? (uint32_t)a.plan.obs[(size_t)row * kObsMax + lane] : 0u;
  const uint64_t fs = a.lib.f_start[row];
  const uint32_t c0 = uni(rec.x), c1 = uni(rec.y);
  const uint32_t nsel = uni(rec.z & 0xffffu), nobs = uni(rec.z >> 16);
  const float g = __builtin_bit_cast(float, uni(g_raw));
  const uint32_t ncand = uni(ncand_raw);
  if (ncand == 0 || g == 0.0f || c1 <= c0 || nobs == 0 || nsel == 0) {
    if (lane == 0) a.contrib[row] = 0.0;
    return;
  }
  const uint32_t istar = uni(cen_raw) - c0;
  // cycles [hc0, hc1) of the apex window; x[t] = X[hc0 + t]
  const uint32_t hc0 = c0 + istar - half;
  const uint32_t wl = max(hc0, c0), wh = min(hc0 + 2 * half + 1, c1);
  // the forward left this row's XIC window (block-sparse tile rows)
  const bool saved = uni(ok_raw) != 0u;
  // bin range of fragment `lane`; window bases; kernel taps and their
  // sigma derivatives (lane j holds tap j)
  uint32_t bst = 0, ben = 0;
  if (!saved && lane < (int)nsel) bin_range(a.dia, smz_l, a.prm.mass_tol, bst, ben);
  const uint32_t nob = min(nobs, (uint32_t)kObsMax);
  const uint64_t eb_l = !saved && lane < (int)nob ? a.dia.entry_base[obs_l] : 0ull;
  const float kk_l = lane <= (int)half ? a.prm.kernels[(size_t)rec.w * a.prm.klen + lane] : 0.0f;
  const double dk_l = lane <= (int)half ? a.dk[(size_t)rec.w * (half + 1) + lane] : 0.0;
  // CSR entry range of fragment `lane` in window j
  auto csr_range = [&](uint32_t j, uint32_t &rx, uint32_t &ry) {
    const uint32_t o = rdlane(obs_l, j);
    rx = 0;
    ry = 0;
    if (lane < (int)nsel) {
      const uint32_t *ssw = a.dia.ss + (size_t)o * (a.dia.n_grid + 1);
      rx = ssw[bst];
      ry = ssw[ben];
    }
  };
  if (saved) {
    // fragment-major, 2 * half + 1 cells per fragment
#pragma unroll
    for (int q = 0; q < kBwdPreFrag / 2; ++q) {
      const uint32_t f = 2u * q + hl;
      if (f < nsel && tl < wlen) xw[f * kWin + tl] = pre[q];
    }
    const float *xi = a.xwin + (size_t)row * a.xwin_stride;
    for (uint32_t f0 = kBwdPreFrag; f0 < nsel; f0 += 2) {
      const uint32_t f = f0 + hl;
      if (f < nsel && tl < wlen) xw[f * kWin + tl] = xi[f * wlen + tl];
    }
    __builtin_amdgcn_wave_barrier();
  }
  uint32_t rx0 = 0, ry0 = 0;
  if (!saved) csr_range(0, rx0, ry0);
  if (!saved)
    for (uint32_t i = lane; i < nsel * kWin; i += kWave) xw[i] = 0.0f;
  __builtin_amdgcn_wave_barrier();
  // the XIC window, summed in the forward pass's order: windows, then the
  // fragments' CSR entries in order (8-entry groups, 64 per tile)
  for (uint32_t j = 0; j < (saved ? 0u : nob); ++j) {
    const uint32_t o = rdlane(obs_l, j);
    const uint64_t base = ((uint64_t)rdlane((uint32_t)(eb_l >> 32), j) << 32) |
                          rdlane((uint32_t)eb_l, j);
    uint32_t rx = rx0, ry = ry0;
    if (j > 0) csr_range(j, rx, ry);
    const uint32_t bo = (uint32_t)base & 7u;
    const uint16_t *cw = a.dia.cycles + (base - bo);
    const uint32_t gsr = (bo + rx) >> 3;
    const uint32_t ng = ry > rx ? ((bo + ry - 1) >> 3) - gsr + 1 : 0u;
    const uint32_t meta = ((bo + rx) & 7u) | (((bo + ry - 1) & 7u) << 3) | (ng << 6);
    const uint32_t gincl = wave_incl_scan(ng);
    const uint32_t gp = gincl - ng;
    const uint32_t NG = rdlane(gincl, nsel - 1);
    uint32_t nh = 0;
    auto flush = [&]() {
      for (uint32_t c = 0; c < nh; c += kWave) {
        const uint32_t idx = c + lane;
        const bool v = idx < nh;
        uint2 h = make_uint2(0u, 0u);
        if (v) h = hits[idx];
        const float inten = v ? a.dia.intens[base + h.x] : 0.0f;
        const uint32_t slot = h.y;  // f * kWin + t
        if (v) own[slot] = (uint8_t)lane;
        __builtin_amdgcn_wave_barrier();
        const bool dup = v && own[slot] != (uint8_t)lane;
        if (ballot64(dup) == 0) {
          if (v) xw[slot] = xw[slot] + inten;
        } else {
          uint64_t it2 = ballot64(v);
          while (it2) {
            const int j2 = __builtin_ctzll(it2);
            it2 &= it2 - 1;
            if (lane == j2) xw[slot] = xw[slot] + inten;
            __builtin_amdgcn_wave_barrier();
          }
        }
        __builtin_amdgcn_wave_barrier();
      }
      nh = 0;
    };
    for (uint32_t v0 = 0; v0 < NG; v0 += kWave) {
      const uint32_t v = v0 + lane;
      const uint32_t vmax = min(v0 + (uint32_t)kWave - 1, NG - 1);
      const uint64_t lo_m = ballot64(lane < (int)nsel && ng > 0 && gp <= v0);
      const uint32_t flo = 63u - (uint32_t)__builtin_clzll(lo_m);
      // fragment of this lane's group: the last non-empty one starting at
      // or before it, by start markers and a prefix maximum
      mk[lane] = lane == 0 ? flo + 1u : 0u;
      __builtin_amdgcn_wave_barrier();
      if (lane < (int)nsel && ng > 0 && gp > v0 && gp <= vmax) mk[gp - v0] = (uint32_t)lane + 1u;
      __builtin_amdgcn_wave_barrier();
      const uint32_t f = wave_incl_max_dpp(mk[lane]) - 1u;
      const uint32_t fgp = (uint32_t)__shfl((int)gp, (int)f, kWave);
      const uint32_t fgs = (uint32_t)__shfl((int)gsr, (int)f, kWave);
      const uint32_t fmeta = (uint32_t)__shfl((int)meta, (int)f, kWave);
      const uint32_t fb0 = (uint32_t)__shfl((int)bst, (int)f, kWave);
      __builtin_amdgcn_wave_barrier();
      uint4 raw = make_uint4(0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu);
      uint32_t tg = 0, vm = 0;
      if (v < NG) {
        const uint32_t q = v - fgp;
        tg = fgs + q;
        raw = *reinterpret_cast<const uint4 *>(cw + 8 * (size_t)tg);
        const uint32_t klo = (q == 0) ? (fmeta & 7u) : 0u;
        const uint32_t khi = (q + 1 == (fmeta >> 6)) ? ((fmeta >> 3) & 7u) : 7u;
        vm = (0xffu << klo) & (0xffu >> (7u - khi));
      }
      const uint32_t cyc[8] = {raw.x & 0xffffu, raw.x >> 16, raw.y & 0xffffu,
                               raw.y >> 16,     raw.z & 0xffffu, raw.z >> 16,
                               raw.w & 0xffffu, raw.w >> 16};
      uint32_t hm = 0;
#pragma unroll
      for (int k = 0; k < 8; ++k) hm |= (cyc[k] >= wl && cyc[k] < wh ? 1u : 0u) << k;
      hm &= vm;
      bool at_c0 = false;
#pragma unroll
      for (int k = 0; k < 8; ++k) at_c0 = at_c0 || (((hm >> k) & 1u) && cyc[k] == c0);
      if (ballot64(at_c0) != 0) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          if (((hm >> k) & 1u) && cyc[k] == c0 &&
              skipped_duplicate(a.dia, o, fb0, 8 * tg + k - bo, c0,
                                k < 7 ? cyc[k < 7 ? k + 1 : 7]
                                      : (uint32_t)a.dia.cycles[base + 8 * tg + k - bo + 1]))
            hm &= ~(1u << k);
        }
      }
      // list slots in (lane, k) order = entry order; lanes 0-31 first,
      // then 32-63 (at most 32 x 8 hits each)
#pragma unroll
      for (int hh = 0; hh < 2; ++hh) {
        uint32_t pre, tot;
        list_slots(hm, (lane >> 5) == hh, pre, tot);
        if (tot == 0) continue;
        if ((lane >> 5) == hh) {
          uint32_t pos = nh + pre;
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            if ((hm >> k) & 1u) {
              hits[pos] = make_uint2(8 * tg + k - bo, f * kWin + (cyc[k] - hc0));
              ++pos;
            }
          }
        }
        nh += tot;
        __builtin_amdgcn_wave_barrier();
        if (nh > (uint32_t)(kBwdHitCap - 4 * kWave)) flush();
      }
    }
    if (nh) flush();
    __builtin_amdgcn_wave_barrier();
  }
  __builtin_amdgcn_wave_barrier();
  // per fragment: Y at the apex, dY/dsigma, ln(Y+1)
  auto tap = [&](uint32_t j) {
    return __builtin_bit_cast(float, rdlane(__builtin_bit_cast(uint32_t, kk_l), j));
  };
  auto dtap = [&](uint32_t j) {
    const uint64_t b = __builtin_bit_cast(uint64_t, dk_l);
    return __builtin_bit_cast(double, ((uint64_t)rdlane((uint32_t)(b >> 32), j) << 32) |
                                          rdlane((uint32_t)b, j));
  };
  const float *x = xw + min((uint32_t)lane, nsel - 1) * kWin + half;  // x[0] = apex cycle
  float y = 0.0f;
  y = y + x[0] * tap(half);
  double dy = (double)x[0] * dtap(half);
  for (uint32_t jj = 0; jj < half; ++jj) {
    const uint32_t d = half - jj;
    y = y + (x[-(int)d] + x[d]) * tap(jj);
    dy += ((double)x[-(int)d] + (double)x[d]) * dtap(jj);
  }
  if (lane < (int)nsel) {
    // d s*/d w_f = ln(1 + Y_f), evaluated as log1p in f64: the f32
    // logf(Y + 1) of the forward loses the low bits of Y when Y << 1
    const double lnv = log1p((double)y);
    a.grad_w[fs + sel_l] = (float)((double)g * lnv);
    part[lane] = (double)selw_l * dy / ((double)y + 1.0);
  }
  __builtin_amdgcn_wave_barrier();
  if (lane == 0) {
    double acc = 0.0;  // fixed fragment order
    for (uint32_t f = 0; f < nsel; ++f) acc += part[f];
    a.contrib[row] = (double)g * acc;
  }
}

// grad_sigma[w] = sum of contrib over the rows whose kernel is window w's,
// in a fixed order: block (c, w) sums rows [c * kSigmaChunk, ...) with
// per-thread strided sums and a fixed tree; the chunks are then added in
// chunk order.  Deterministic for a given row count.
constexpr uint32_t kSigmaChunk = 8192;

__global__ __launch_bounds__(256) void sigma_partial_kernel(const uint4 *rec,
                                                            const double *contrib,
                                                            uint32_t n,
                                                            double *part) {
  __shared__ double red[256];
  const uint32_t c = blockIdx.x, w = blockIdx.y;
  const uint32_t r0 = c * kSigmaChunk, r1 = min(r0 + kSigmaChunk, n);
  double acc = 0.0;
  for (uint32_t r = r0 + threadIdx.x; r < r1; r += 256)
    if (rec[r].w == w + 1) acc += contrib[r];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s2 = 128; s2 > 0; s2 >>= 1) {
    if ((int)threadIdx.x < s2) red[threadIdx.x] += red[threadIdx.x + s2];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(size_t)w * gridDim.x + c] = red[0];
}

__global__ __launch_bounds__(64) void sigma_final_kernel(const double *part,
                                                         uint32_t nch,
                                                         double *out) {
  if (threadIdx.x != 0) return;
  const uint32_t w = blockIdx.x;
  double acc = 0.0;
  for (uint32_t c = 0; c < nch; ++c) acc += part[(size_t)w * nch + c];
  out[w] = acc;
}

// ---------------------------------------------------------------------------
// per-stage kernels (SimdBackend operations)
// ---------------------------------------------------------------------------
// src/convolution/scalar.rs:5-56, one thread per output element
__global__ void conv_dense_kernel(const float *k, uint32_t klen, const float *x,
                                  uint32_t rows, uint32_t cols, float *y) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)rows * cols) return;
  const uint32_t f = (uint32_t)(t / cols), i = (uint32_t)(t % cols);
  const uint32_t half = klen / 2;
  float out = 0.0f;
  if (cols >= klen && i >= half && i + half < cols) {
    const float *xr = x + (size_t)f * cols;
    float s = 0.0f;
    s = s + xr[i] * k[half];
    for (uint32_t kk = 0; kk < half; ++kk) {
      const float kv = k[kk];
      s = s + (xr[i - (half - kk)] + xr[i + (half - kk)]) * kv;
    }
    out = s;
  }
  y[t] = out;
}

// src/score/scalar.rs:4-25, one thread per column, rows in order
__global__ void logdot_kernel(const float *a, uint32_t rows, uint32_t cols,
                              const float *w, float *out) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  float r = 0.0f;
  for (uint32_t i = 0; i < rows; ++i) {
    const float v = logf_exact(a[(size_t)i * cols + j] + 1.0f, kDevLogfTab);
    r = r + v * w[i];
  }
  out[j] = r;
}

// logf_exact of every input.  For a positive normal input the fused kernels
// may take the shorter logf_normal instead; where the two differ the output
// is the NaN 0x7fc0dead, a value glibc never returns for such an input, so a
// comparison with glibc checks both paths at once.
__global__ void logf_kernel(const float *x, uint64_t n, float *y) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) {
    const float v = x[t];
    float r = logf_exact(v, kDevLogfTab);
    if (logf_is_normal(v) &&
        __float_as_uint(logf_normal(v, kDevLogfTab)) != __float_as_uint(r))
      r = __uint_as_float(0x7fc0deadu);
    y[t] = r;
  }
}

// src/peak_group_selection/mod.rs:23-64 on one wavefront: flags in parallel,
// then a stable descending insertion order by lane 0.
__global__ void maxima_kernel(const float *a, uint32_t n, uint64_t offset,
                              uint64_t *idx, float *val, uint32_t *count) {
  if (threadIdx.x != 0) return;
  uint32_t m = 0;
  if (n >= 5) {
    for (uint32_t i = 2; i < n - 2; ++i) {
      const float v = a[i];
      if (a[i - 2] < a[i - 1] && a[i - 1] < v && v > a[i + 1] &&
          a[i + 1] > a[i + 2]) {
        uint32_t j = m;
        while (j > 0 && v > val[j - 1]) {
          val[j] = val[j - 1];
          idx[j] = idx[j - 1];
          --j;
        }
        val[j] = v;
        idx[j] = i + offset;
        ++m;
      }
    }
  }
  *count = m;
}

// src/dense_xic_observation/mod.rs:30-64: one thread per fragment row, valid
// windows in ascending order, bins ascending, entries from the last entry
// equal to cycle_start (Rust binary_search).
__global__ void dense_xic_kernel(KDia dia, float pmz, uint32_t c0, uint32_t c1,
                                 float tol, const float *fmz, uint32_t nf,
                                 float *out) {
  const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const uint32_t L = c1 - c0;
  float *row = out + (size_t)f * L;
  for (uint32_t i = 0; i < L; ++i) row[i] = 0.0f;
  const float mz = fmz[f];
  const float delta = mz * tol * 1e-6f;
  uint32_t start, end;
  mz_range(dia, mz - delta, mz + delta, start, end);
  for (uint32_t o = 0; o < dia.n_obs; ++o) {
    if (!(dia.iso_lo[o] <= pmz && dia.iso_hi[o] >= pmz)) continue;
    const uint32_t *ss = dia.ss + (size_t)o * (dia.n_grid + 1);
    const uint16_t *cy = dia.cycles + dia.entry_base[o];
    const float *in = dia.intens + dia.entry_base[o];
    for (uint32_t b = start; b < end; ++b) {
      const uint32_t s0 = ss[b], n = ss[b + 1] - s0;
      // Rust binary_search on u16
      uint32_t pos = 0;
      if (n > 0) {
        const uint16_t key = (uint16_t)c0;
        uint32_t size = n, base = 0;
        while (size > 1) {
          const uint32_t h = size >> 1, mid = base + h;
          if (!(cy[s0 + mid] > key)) base = mid;
          size -= h;
        }
        pos = (cy[s0 + base] == key) ? base : base + (cy[s0 + base] < key ? 1 : 0);
      }
      for (uint32_t i = pos; i < n; ++i) {
        const uint32_t c = cy[s0 + i];
        if (c >= c1) break;
        row[c - c0] = row[c - c0] + in[s0 + i];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// host-side device management
// ---------------------------------------------------------------------------
static int g_device_state = -1;  // -1 unknown, 0 none, 1 gfx950 present
static std::mutex g_dev_mu;
static std::atomic<int> g_backend_override{-1};

static bool device_present() {
  std::lock_guard<std::mutex> lk(g_dev_mu);
  if (g_device_state >= 0) return g_device_state == 1;
  int n = 0;
  g_device_state = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return false;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return false;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return false;
  g_device_state = 1;
  return true;
}

static adsr_status require_device() {
  if (!device_present())
    return set_error(ADSR_ERR_DEVICE,
                     "no gfx950 (MI355X) device is visible; the hip backend "
                     "is the only selection backend");
  return ADSR_OK;
}

template <class T>
static adsr_status dmalloc_copy(T **dst, const T *src, size_t n, size_t &acc,
                                size_t pad = 0) {
  *dst = nullptr;
  const size_t bytes = std::max<size_t>(n + pad, 1) * sizeof(T);
  HIP_TRY(dev_malloc((void **)dst, bytes));
  if (n) HIP_TRY(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
  if (pad) HIP_TRY(hipMemset(*dst + n, 0xff, pad * sizeof(T)));
  acc += bytes;
  return ADSR_OK;
}

void device_dia_free(DeviceDia &d) {
  if (!d.ready && !d.grid) return;
  (void)hipFree(d.grid);
  (void)hipFree(d.bucket);
  (void)hipFree(d.bucket2);
  (void)hipFree(d.bucket1);
  (void)hipFree(d.rt);
  (void)hipFree(d.iso_lo);
  (void)hipFree(d.iso_hi);
  (void)hipFree(d.slice_starts);
  (void)hipFree(d.entry_base);
  (void)hipFree(d.cycles);
  (void)hipFree(d.intens);
  d = DeviceDia();
}

void device_lib_free(DeviceLib &l) {
  if (!l.ready && !l.p_mz) return;
  (void)hipFree(l.p_idx);
  (void)hipFree(l.p_mz);
  (void)hipFree(l.p_rt);
  (void)hipFree(l.f_start);
  (void)hipFree(l.f_stop);
  (void)hipFree(l.f_mz);
  (void)hipFree(l.f_int);
  (void)hipFree(l.f_type);
  (void)hipFree(l.f_num);
  (void)hipFree(l.p_naa);
  (void)hipFree(l.f_mzl);
  (void)hipFree(l.idf_counts);
  l = DeviceLib();
}

#define UP(dst, src, n, ...)                                              \
  do {                                                                    \
    adsr_status s_ = dmalloc_copy(&(dst), (src), (n), acc, ##__VA_ARGS__); \
    if (s_) {                                                             \
      device_dia_free(v);                                                 \
      return s_;                                                          \
    }                                                                     \
  } while (0)

// Grid tables, RT axis, window bounds and entry bases: everything of a
// DIAData on the device except the CSR arrays.
static adsr_status upload_dia_meta(adsr_dia *d, size_t &acc) {
  const Grid &g = global_grid();
  DeviceDia &v = d->dev;
  UP(v.grid, g.mz.data(), g.mz.size());
  UP(v.bucket, g.bucket.data(), g.bucket.size());
  {
    std::vector<uint32_t> b2(4 * g.bucket.size());
    const size_t n = g.mz.size();
    const uint32_t inf = 0x7f800000u;
    for (size_t b = 0; b < g.bucket.size(); ++b) {
      const uint32_t c = g.bucket[b];
      b2[4 * b + 0] = c;
      b2[4 * b + 1] = c < n ? __builtin_bit_cast(uint32_t, g.mz[c]) : inf;
      b2[4 * b + 2] = c + 1 < n ? __builtin_bit_cast(uint32_t, g.mz[c + 1]) : inf;
      b2[4 * b + 3] = 0;
    }
    UP(v.bucket2, b2.data(), b2.size());
    // the same buckets in 4 bytes: the grid values themselves are the
    // bucket's bit pattern with the stored ulp offsets
    std::vector<uint32_t> b1(g.bucket.size());
    for (size_t b = 0; b < g.bucket.size(); ++b) {
      const uint32_t c = g.bucket[b];
      uint32_t nin = 0, off[2] = {0, 0};
      for (uint32_t k = c; k < n && nin < 3; ++k) {
        const uint32_t bits = __builtin_bit_cast(uint32_t, g.mz[k]);
        if ((bits >> kBucketShift) - g.base_key != b) break;
        if (nin < 2) off[nin] = bits & 15u;
        ++nin;
      }
      if (nin > 2 || c >= (1u << 22))
        return set_error(ADSR_ERR_INTERNAL, "m/z grid bucket table invariant");
      b1[b] = c | (nin << 22) | (off[0] << 24) | (off[1] << 28);
    }
    UP(v.bucket1, b1.data(), b1.size());
  }
  UP(v.rt, d->rt_index.data(), d->rt_index.size());
  UP(v.iso_lo, d->iso_lo.data(), d->iso_lo.size());
  UP(v.iso_hi, d->iso_hi.data(), d->iso_hi.size());
  UP(v.entry_base, d->entry_base.data(), d->entry_base.size());
  return ADSR_OK;
}

static adsr_status upload_dia(adsr_dia *d) {
  std::lock_guard<std::mutex> lk(d->mu);
  if (d->dev.ready) return ADSR_OK;
  adsr_status st = require_device();
  if (st) return st;
  if (!d->host_csr)
    return set_error(ADSR_ERR_INVALID_ARGUMENT,
                     "device-built DIAData has no device copy");
  DeviceDia &v = d->dev;
  size_t acc = 0;
  st = upload_dia_meta(d, acc);
  if (st) return st;
  UP(v.slice_starts, d->slice_starts.data(), d->slice_starts.size());
  // 8 spare entries: the select kernel reads cycles in aligned groups of 8
  UP(v.cycles, d->cycles.data(), d->cycles.size(), 8);
  UP(v.intens, d->intens.data(), d->intens.size());
  v.bytes = acc;
  v.ready = true;
  (void)hipGetDevice(&v.device);
  return ADSR_OK;
}
#undef UP

static KDia make_kdia(const adsr_dia *d) {
  const Grid &g = global_grid();
  const DeviceDia &dv = d->dev;
  KDia k;
  k.grid = dv.grid;
  k.bucket = dv.bucket;
  k.n_grid = (uint32_t)g.mz.size();
  k.base_key = g.base_key;
  k.rt = dv.rt;
  k.n_rt = (uint32_t)d->rt_index.size();
  k.delta_t = d->delta_t;
  k.rt_strict = d->rt_strict ? 1u : 0u;
  k.rt_inv_dt = (d->rt_strict && std::isfinite(d->delta_t) && d->delta_t > 0.0f)
                    ? 1.0f / d->delta_t : 0.0f;
  k.n_obs = (uint32_t)d->n_obs;
  k.iso_lo = dv.iso_lo;
  k.iso_hi = dv.iso_hi;
  k.ss = dv.slice_starts;
  k.entry_base = dv.entry_base;
  k.cycles = dv.cycles;
  k.intens = dv.intens;
  k.bucket2 = reinterpret_cast<const uint4 *>(dv.bucket2);
  k.bucket1 = dv.bucket1;
  k.grid_first = g.mz.front();
  k.grid_last = g.mz.back();
  return k;
}

static adsr_status upload_lib(adsr_lib *l) {
  std::lock_guard<std::mutex> lk(l->mu);
  if (l->dev.ready) return ADSR_OK;
  adsr_status st = require_device();
  if (st) return st;
  DeviceLib &v = l->dev;
  size_t acc = 0;
#define UP(dst, src, n, ...)                             \
  do {                                                   \
    adsr_status s_ = dmalloc_copy(&(dst), (src), (n), acc, ##__VA_ARGS__); \
    if (s_) {                                            \
      device_lib_free(v);                                \
      return s_;                                         \
    }                                                    \
  } while (0)
  UP(v.p_idx, l->precursor_idx.data(), l->n_prec);
  UP(v.p_mz, l->precursor_mz.data(), l->n_prec);
  UP(v.p_rt, l->precursor_rt.data(), l->n_prec);
  UP(v.f_start, l->frag_start.data(), l->n_prec);
  UP(v.f_stop, l->frag_stop.data(), l->n_prec);
  UP(v.f_mz, l->fragment_mz.data(), l->n_frag);
  UP(v.f_int, l->fragment_intensity.data(), l->n_frag);
  UP(v.f_type, l->fragment_type.data(), l->n_frag);
  UP(v.f_num, l->fragment_number.data(), l->n_frag);
#undef UP
  v.bytes = acc;
  v.ready = true;
  (void)hipGetDevice(&v.device);
  return ADSR_OK;
}

// ---------------------------------------------------------------------------
// search workspace and launch
// ---------------------------------------------------------------------------
struct LaunchShape {
  int waves = 1;  // waves per block: 1, 2 or 4 (64/128/256 threads)
  int chunk = 1;  // precursors per queue grab: 1, 2, 4, 8 (16 on request)
};

struct Workspace {
  size_t n_rows = 0, ks = 0, cc = 0, nker = 0;
  uint4 *rec = nullptr;
  uint16_t *obs = nullptr, *sel = nullptr;
  float *sel_mz = nullptr, *sel_w = nullptr;
  uint2 *rng = nullptr;
  uint32_t *b0 = nullptr;
  float *cand_score = nullptr;
  uint32_t *cand_center = nullptr, *cand_count = nullptr;
  uint32_t *local = nullptr, *bsum = nullptr, *small = nullptr;  // small: queue, status, total
  CandRow *out = nullptr;
  float *kernels = nullptr;
  float *rt_tol_pp = nullptr;
  uint32_t *key_scratch = nullptr;
  uint64_t *table = nullptr;  // host table block on the device (68 B per row)
  bool pp = false;            // per-precursor RT tolerances
  size_t nfs = 0;             // key_scratch entries
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  // pipelined host tables: the column copies' stream
  hipStream_t stream2 = nullptr;
  // pipelined host tables: chunk k done (ev_pipe[k]); page-locked copies of
  // the small words after each chunk (hsmall + kSmallWords * k)
  hipEvent_t ev_pipe[kPipeMax] = {};
  uint32_t *hsmall = nullptr, *hsmall_dev = nullptr;
  hipGraphExec_t graph = nullptr;
  hipGraph_t graph_src = nullptr;
};

static void ws_free(Workspace &w) {
  // work of a failed search may still be queued: it ends before its
  // buffers are freed
  if (w.stream2) (void)hipStreamSynchronize(w.stream2);
  if (w.stream) (void)hipStreamSynchronize(w.stream);
  if (w.graph) (void)hipGraphExecDestroy(w.graph);
  if (w.graph_src) (void)hipGraphDestroy(w.graph_src);
  (void)hipFree(w.rec);
  (void)hipFree(w.obs);
  (void)hipFree(w.sel);
  (void)hipFree(w.sel_mz);
  (void)hipFree(w.sel_w);
  (void)hipFree(w.rng);
  (void)hipFree(w.b0);
  (void)hipFree(w.cand_score);
  (void)hipFree(w.cand_center);
  (void)hipFree(w.cand_count);
  (void)hipFree(w.local);
  (void)hipFree(w.bsum);
  (void)hipFree(w.small);
  (void)hipFree(w.out);
  (void)hipFree(w.kernels);
  (void)hipFree(w.rt_tol_pp);
  (void)hipFree(w.key_scratch);
  (void)hipFree(w.table);
  for (auto &e : w.ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : w.ev_pipe)
    if (e) (void)hipEventDestroy(e);
  if (w.hsmall) (void)hipHostFree(w.hsmall);
  if (w.stream2) (void)hipStreamDestroy(w.stream2);
  if (w.stream) (void)hipStreamDestroy(w.stream);
  w = Workspace();
}

// Libraries of more than kChunkAbove rows are searched in chunks of at most
// kChunkRows rows (ADSR_CHUNK_ROWS overrides the chunk size; 0 turns
// chunking off): each chunk is planned and selected in turn, and one
// compaction over all rows keeps library order.
constexpr uint32_t kChunkAbove = 1u << 20;
constexpr uint32_t kChunkRows = 1u << 20;  // 2 chunks at 2M rows: 1.7% over one launch, 4 chunks 4%

static uint32_t search_chunks(uint32_t n_rows) {
  static const long env = [] {
    const char *e = std::getenv("ADSR_CHUNK_ROWS");
    return e ? std::strtol(e, nullptr, 10) : -1L;
  }();
  uint32_t rows = kChunkRows;
  if (env == 0) return 1;
  if (env > 0) rows = (uint32_t)std::min<long>(env, 0x7fffffffL);
  else if (n_rows <= kChunkAbove) return 1;
  const uint64_t k = ((uint64_t)n_rows + rows - 1) / rows;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(k, kMaxChunks));
}

// A search whose table goes to host memory runs in chunks of library rows
// (multiples of kScanBlock), each compacted and tabled on its own, so that
// chunk k's columns are copied to the host while chunk k + 1 is selected.
// ADSR_PIPE_CHUNKS sets the count (0 or 1: off); by default 3 for
// libraries of kPipeAbove rows or more, below kChunkAbove rows (larger
// libraries are planned in chunks and compacted once).  The copies to host
// memory run as blit kernels on the compute units, slowing the select they
// overlap, and each select launch pays its start and tail: at 400k rows,
// 3 chunks measured best (about 10% less than one launch and one copy).
constexpr uint32_t kPipeAbove = 1u << 17;

static uint32_t pipe_chunks(uint32_t n_rows) {
  static const long env = [] {
    const char *e = std::getenv("ADSR_PIPE_CHUNKS");
    return e ? std::strtol(e, nullptr, 10) : -1L;
  }();
  if (n_rows <= kScanBlock || search_chunks(n_rows) > 1) return 1;
  uint32_t k = env >= 0 ? (uint32_t)std::min<long>(env, kPipeMax)
                        : (n_rows >= kPipeAbove ? 3u : 1u);
  // every chunk holds at least one scan block
  k = std::min<uint32_t>(k, (n_rows + kScanBlock - 1) / kScanBlock);
  return std::max<uint32_t>(k, 1u);
}

// Idle workspaces of a library (at most kIdleWorkspaces) are taken by the
// next search with the same shape; a workspace is used by one search at a
// time, so concurrent searches each get their own.
constexpr size_t kIdleWorkspaces = 2;

static adsr_status ws_alloc(Workspace &w, size_t n_rows, size_t ks, size_t cc,
                            size_t nker, bool pp, size_t n_frag_scratch,
                            adsr_lib *pool = nullptr) {
  ws_free(w);
  if (pool) {
    std::lock_guard<std::mutex> lk(pool->mu);
    auto &v = pool->idle_workspaces;
    for (size_t i = 0; i < v.size(); ++i) {
      Workspace *c = static_cast<Workspace *>(v[i]);
      if (c->n_rows == n_rows && c->ks == ks && c->cc == cc && c->nker == nker &&
          c->pp == pp && c->nfs == n_frag_scratch) {
        w = *c;
        delete c;
        v.erase(v.begin() + (long)i);
        return ADSR_OK;
      }
    }
  }
  w.pp = pp;
  w.nfs = n_frag_scratch;
  w.n_rows = n_rows;
  w.ks = ks;
  w.cc = cc;
  w.nker = nker;
  const size_t R = std::max<size_t>(n_rows, 1);
  const size_t nb = (R + kScanBlock - 1) / kScanBlock;
  HIP_TRY(hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking));
  for (auto &e : w.ev) HIP_TRY(hipEventCreate(&e));
  if (pipe_chunks((uint32_t)n_rows) > 1) {
    HIP_TRY(hipStreamCreateWithFlags(&w.stream2, hipStreamNonBlocking));
    for (auto &e : w.ev_pipe) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    if (alloc_should_fail() ||
        hipHostMalloc((void **)&w.hsmall, kSmallWords * 4 * kPipeMax,
                      hipHostMallocMapped) != hipSuccess) {
      w.hsmall = nullptr;
      return set_error(ADSR_ERR_OUT_OF_MEMORY, "page-locked status words");
    }
    HIP_TRY(hipHostGetDevicePointer((void **)&w.hsmall_dev, w.hsmall, 0));
  }
  HIP_TRY(dev_malloc(&w.rec, R * sizeof(uint4)));
  HIP_TRY(dev_malloc(&w.obs, R * kObsMax * 2));
  HIP_TRY(dev_malloc(&w.sel, R * ks * 2));
  HIP_TRY(dev_malloc(&w.sel_mz, R * ks * 4));
  HIP_TRY(dev_malloc(&w.sel_w, R * ks * 4));
  HIP_TRY(dev_malloc(&w.rng, R * kObsMax * ks * sizeof(uint2)));
  HIP_TRY(dev_malloc(&w.b0, R * kObsMax * ks * 4));
  HIP_TRY(dev_malloc(&w.cand_score, R * std::max<size_t>(cc, 1) * 4));
  HIP_TRY(dev_malloc(&w.cand_center, R * std::max<size_t>(cc, 1) * 4));
  HIP_TRY(dev_malloc(&w.cand_count, R * 4));
  HIP_TRY(dev_malloc(&w.local, R * 4));
  HIP_TRY(dev_malloc(&w.bsum, nb * 4));
  HIP_TRY(dev_malloc(&w.small, kSmallWords * 4));
  HIP_TRY(dev_malloc(&w.out, R * std::max<size_t>(cc, 1) * sizeof(CandRow)));
  HIP_TRY(dev_malloc(&w.kernels, nker * 4));
  HIP_TRY(dev_malloc(&w.rt_tol_pp, (pp ? R : 1) * 4));
  HIP_TRY(dev_malloc(&w.key_scratch, std::max<size_t>(n_frag_scratch, 1) * 4));
  return ADSR_OK;
}

// Returns a finished search's workspace to its library's idle list (or
// frees it when the list is full).
static void ws_recycle(adsr_lib *l, Workspace &w) {
  if (w.graph) (void)hipGraphExecDestroy(w.graph);
  if (w.graph_src) (void)hipGraphDestroy(w.graph_src);
  w.graph = nullptr;
  w.graph_src = nullptr;
  if (w.stream) {
    std::lock_guard<std::mutex> lk(l->mu);
    if (l->idle_workspaces.size() < kIdleWorkspaces) {
      l->idle_workspaces.push_back(new Workspace(w));
      w = Workspace();
      return;
    }
  }
  ws_free(w);
}

void release_lib_workspaces(adsr_lib *l) {
  std::vector<void *> v;
  {
    std::lock_guard<std::mutex> lk(l->mu);
    v.swap(l->idle_workspaces);
  }
  for (void *p : v) {
    Workspace *w = static_cast<Workspace *>(p);
    ws_free(*w);
    delete w;
  }
}

// ---- page-locked blocks for host candidate tables ------------------------
// Blocks are kept after release up to kPinnedPoolBytes in total and handed
// to the next request they fit (capacity at most twice the request).  The
// pool is never destroyed, so tables freed during interpreter shutdown can
// still return their blocks.
constexpr size_t kPinnedPoolBytes = size_t(1) << 30;

struct PinnedPool {
  std::mutex mu;
  std::multimap<size_t, void *> idle;
  size_t idle_bytes = 0;
};

static PinnedPool &pinned_pool() {
  static PinnedPool *p = new PinnedPool();
  return *p;
}

void *pinned_acquire(size_t bytes, size_t *cap) {
  const size_t want = (std::max<size_t>(bytes, 1) + (size_t(2) << 20) - 1) &
                      ~((size_t(2) << 20) - 1);
  PinnedPool &pp = pinned_pool();
  {
    std::lock_guard<std::mutex> lk(pp.mu);
    auto it = pp.idle.lower_bound(want);
    if (it != pp.idle.end() && it->first <= 2 * want) {
      void *p = it->second;
      *cap = it->first;
      pp.idle_bytes -= it->first;
      pp.idle.erase(it);
      return p;
    }
  }
  void *p = nullptr;
  if (alloc_should_fail() ||
      hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess)
    return nullptr;
  *cap = want;
  return p;
}

void pinned_release(void *p, size_t cap) {
  if (!p) return;
  PinnedPool &pp = pinned_pool();
  {
    std::lock_guard<std::mutex> lk(pp.mu);
    if (pp.idle_bytes + cap <= kPinnedPoolBytes) {
      pp.idle.emplace(cap, p);
      pp.idle_bytes += cap;
      return;
    }
  }
  (void)hipHostFree(p);
}

struct SearchPlan {
  KDia dia;
  KLib lib;
  KParams prm;
  KPlan plan;
  uint32_t lcap = 0, xcap = 0, gcap = 0, lds_per_wave = 0;
  uint32_t lds_budget = 0;  // per-wave LDS share (incl. static) to fill
  bool lds_grow = false;    // rows shorter than lcap can use more LDS
  int cus = 256;
  uint64_t max_raw = 0;  // longest raw fragment list of the library
  std::vector<float> kernels;
  size_t n_taps = 0;
  // differentiable forward: apex XIC windows (SelArgs::xwin)
  float *xwin = nullptr;
  uint32_t *xwin_ok = nullptr;
  uint32_t xwin_stride = 0;
};

// Largest number of RT-index points inside any closed window of width
// 2*eff (+ rounding slack).  Bounds c1 - c0 for every precursor when the
// RT axis is sorted; the full axis otherwise.
static uint32_t window_bound(const adsr_dia *d, float eff) {
  const std::vector<float> &rt = d->rt_index;
  const size_t n = rt.size();
  if (n == 0) return 0;
  if (!d->rt_sorted || !std::isfinite(eff)) return (uint32_t)n;
  const double width = 2.0 * (double)eff * (1.0 + 1e-5) + 1e-3;
  size_t best = 0, j = 0;
  for (size_t i = 0; i < n; ++i) {
    if (j < i) j = i;
    while (j < n && (double)rt[j] - (double)rt[i] <= width) ++j;
    best = std::max(best, j - i);
  }
  return (uint32_t)std::min<size_t>(n, best + 2);
}

// Taps go to sp.kernels, or to taps[0, taps_cap) when `taps` is given (the
// no-allocation search), with their count in sp.n_taps.
static adsr_status prepare(const adsr_dia *d, const adsr_lib *l,
                           const adsr_selection_params *p, SearchPlan &sp,
                           size_t &ks_out, float *taps = nullptr,
                           size_t taps_cap = 0, size_t kmax = kMaxSel) {
  if (p->kernel_size == 0)
    return set_error(ADSR_ERR_INVALID_ARGUMENT,
                     "kernel_size must be >= 1 (the reference indexes "
                     "kernel_array[kernel_size/2])");
  const size_t klen = (size_t)std::ceil((float)p->kernel_size / 2.0f) * 2;
  if (klen / 2 > (size_t)kHalfMax)
    return set_error(ADSR_ERR_UNSUPPORTED,
                     "kernel_size %llu exceeds the device limit of %d",
                     (unsigned long long)p->kernel_size, 2 * kHalfMax);
  if (p->sigma_per_window && p->n_sigma_per_window != d->n_obs)
    return set_error(ADSR_ERR_INVALID_ARGUMENT,
                     "sigma_per_window has %llu values for %zu observations",
                     (unsigned long long)p->n_sigma_per_window, d->n_obs);
  sp.n_taps = klen * (p->sigma_per_window ? 1 + d->n_obs : 1);
  float *kt = taps;
  if (!kt) {
    sp.kernels.assign(sp.n_taps, 0.0f);
    kt = sp.kernels.data();
  } else if (sp.n_taps > taps_cap) {
    return set_error(ADSR_ERR_UNSUPPORTED,
                     "%zu kernel taps (one kernel per window) exceed the %zu of "
                     "the no-allocation search", sp.n_taps, taps_cap);
  }
  adsr_gaussian_kernel(p->fwhm_rt, 1.0f, p->kernel_size, 1.0f, kt, klen);
  if (p->sigma_per_window)
    for (size_t o = 0; o < d->n_obs; ++o)
      adsr_gaussian_kernel_sigma(p->sigma_per_window[o], klen, kt + klen * (1 + o),
                                 klen);
  // selected-fragment stride
  uint64_t ks = std::min<uint64_t>(p->top_k_fragments, l->max_raw_fragments);
  ks = std::min<uint64_t>(ks, kmax);
  if (ks == 0) ks = 1;
  ks_out = ks;
  // window bound
  float tol = p->rt_tolerance;
  if (p->rt_tolerance_per_precursor) {
    tol = -__builtin_inff();
    for (size_t i = 0; i < l->n_prec; ++i)
      tol = std::max(tol, p->rt_tolerance_per_precursor[i]);
    if (l->n_prec == 0) tol = 0.0f;
  }
  const float min_tol = (10.0f + (float)p->kernel_size) * d->delta_t;
  const float eff = std::max(tol, min_tol);
  uint32_t lb = window_bound(d, eff);
  if (lb > (uint32_t)kLCapMax) lb = kLCapMax;  // longer windows are flagged
  sp.lcap = std::max<uint32_t>(64, (lb + 63) / 64 * 64);
  // XIC tile: all selected fragments when the wave's LDS stays within
  // kLdsBudget, else groups of fragments processed one after another
  static const uint32_t budget = [] {
    const char *e = std::getenv("ADSR_LDS_BUDGET");
    const long v = e ? std::strtol(e, nullptr, 10) : 0;
    return v > 0 ? (uint32_t)v : kLdsBudget;
  }();
  uint32_t g = (uint32_t)ks;
  while (g > 1 && lds_bytes(sp.lcap, g) > budget) --g;
  sp.gcap = g;
  sp.xcap = g * lds_shape(sp.lcap).lp;
  sp.lds_per_wave = lds_bytes(sp.lcap, g);
  // rows shorter than lcap carve their own tile from the wave's share: when
  // the longest row cannot take every fragment at once, each wave gets the
  // whole budget at launch so that shorter rows take more fragments per group
  sp.lds_budget = budget;
  sp.lds_grow = g < ks;
  if (sp.lds_per_wave > kMaxDynLds)
    return set_error(ADSR_ERR_UNSUPPORTED,
                     "per-precursor LDS tile of %u bytes exceeds the device",
                     sp.lds_per_wave);
  static const int cus = [] {
    int n = 0;
    return hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, 0) ==
                   hipSuccess && n > 0
               ? n
               : 256;
  }();
  sp.cus = cus;
  return ADSR_OK;
}

static void set_lds_limits() {
  static std::once_flag once;
  std::call_once(once, [] {
    const void *fns[] = {
        (const void *)select_kernel<1, 10>, (const void *)select_kernel<2, 10>,
        (const void *)select_kernel<4, 10>, (const void *)select_kernel<1, 0>,
        (const void *)select_kernel<2, 0>,  (const void *)select_kernel<4, 0>,
        (const void *)select_kernel<1, 10, true>, (const void *)select_kernel<2, 10, true>,
        (const void *)select_kernel<4, 10, true>};
    for (const void *fn : fns)
      (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)kMaxDynLds);
  });
}

static int occupancy_blocks(const void *fn, int threads, size_t lds) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, threads, lds) !=
      hipSuccess)
    return 1;
  return std::max(nb, 1);
}

// A pipelined search's host table (see pipe_chunks): `chunks` chunks, the
// table's precursor_idx source and cycle bounds, its row stride; `used` is
// set when enqueue ran the search that way.
struct PipeSpec {
  uint32_t chunks = 0;
  const uint64_t *p_idx = nullptr;
  uint64_t pl = 0, nrt = 0, stride = 0;
  bool used = false;
};

static adsr_status enqueue(Workspace &w, const SearchPlan &sp, uint32_t n_rows,
                           LaunchShape shape, float *dbg,
                           uint32_t dbg_stride, uint32_t dbg_row0,
                           uint32_t dbg_rows, bool timed_select,
                           bool plan_only = false, bool with_bins = false,
                           PipeSpec *pipe = nullptr) {
  hipStream_t s = w.stream;
  uint32_t *queue = w.small, *status = w.small + 1, *total = w.small + 2;
  reset_queue_kernel<<<1, 64, 0, s>>>(w.small);
  // fragment lists of at most 64 entries: a lane group per row; longer
  // ones: one thread per row with an LDS stash / global scratch
  auto plan_rows = [&](const KLib &lb, const KParams &pm, const KPlan &pl, uint32_t n,
                       hipStream_t st) {
    constexpr uint32_t wpb = kPlanThreads / kWave;  // rows per block: wpb * 64 / LP
    if (sp.max_raw <= 16)
      plan_group_kernel<16><<<(n + 4 * wpb - 1) / (4 * wpb), kPlanThreads, 0, st>>>(
          sp.dia, lb, pm, pl, status);
    else if (sp.max_raw <= 32)
      plan_group_kernel<32><<<(n + 2 * wpb - 1) / (2 * wpb), kPlanThreads, 0, st>>>(
          sp.dia, lb, pm, pl, status);
    else if (sp.max_raw <= 64)
      plan_group_kernel<64><<<(n + wpb - 1) / wpb, kPlanThreads, 0, st>>>(
          sp.dia, lb, pm, pl, status);
    else
      plan_kernel<kStashFrag><<<(n + kWave - 1) / kWave, kWave, 0, st>>>(
          sp.dia, lb, pm, pl, w.key_scratch, status);
  };
  auto launch_plan = [&]() { plan_rows(sp.lib, sp.prm, sp.plan, n_rows, s); };
  // rows [r0, r0 + n) as a search of their own (library arrays, tolerances
  // and plan slots offset by r0; fragment ranges stay absolute)
  auto chunk_args = [&](uint32_t r0, uint32_t n, KLib &lb, KParams &pm, KPlan &pl) {
    lb = sp.lib;
    lb.n_prec = n;
    lb.p_mz += r0;
    lb.p_rt += r0;
    lb.f_start += r0;
    lb.f_stop += r0;
    pm = sp.prm;
    if (pm.rt_tol_pp) pm.rt_tol_pp += r0;
    pl = sp.plan;
    pl.rec += r0;
    pl.obs += (size_t)r0 * kObsMax;
    pl.sel += (size_t)r0 * pl.ks;
    pl.sel_mz += (size_t)r0 * pl.ks;
    pl.sel_w += (size_t)r0 * pl.ks;
    if (pl.rng) pl.rng += (size_t)r0 * kObsMax * pl.ks;
    if (pl.b0) pl.b0 += (size_t)r0 * kObsMax * pl.ks;
  };
  const uint32_t nchunk =
      (!dbg && !with_bins && !plan_only) ? search_chunks(n_rows) : 1u;
  const uint32_t crows = (n_rows + nchunk - 1) / nchunk;
  if (n_rows && plan_only) {
    launch_plan();
  } else if (n_rows) {
    // a chunked search plans each chunk right before its select (below):
    // a plan running beside a select slows both about as much as it gains
    if (nchunk == 1) launch_plan();
    // the select kernel finds bin and CSR ranges itself; the per-window
    // range table is only built for the byte count and the backward pass
    if (with_bins) {
      const uint64_t nt = (uint64_t)n_rows * sp.plan.ks;
      bins_kernel<<<(uint32_t)((nt + 255) / 256), 256, 0, s>>>(sp.dia, sp.lib,
                                                                sp.prm, sp.plan);
    }
    SelArgs a;
    a.dia = sp.dia;
    a.lib = sp.lib;
    a.prm = sp.prm;
    a.plan = sp.plan;
    a.n_rows = n_rows;
    a.queue = queue;
    a.chunk = (uint32_t)shape.chunk;
    a.lcap = sp.lcap;
    a.xcap = sp.xcap;
    a.gcap = sp.gcap;
    a.cand_score = w.cand_score;
    a.cand_center = w.cand_center;
    a.cand_count = w.cand_count;
    a.status = status;
    a.dbg_scores = dbg;
    a.dbg_stride = dbg_stride;
    a.dbg_row0 = dbg_row0;
    a.dbg_rows = dbg_rows;
    a.row0 = 0;
    a.xwin = sp.xwin;
    a.xwin_ok = sp.xwin_ok;
    a.xwin_stride = sp.xwin_stride;
    while (shape.waves > 1 &&
           (size_t)sp.lds_per_wave * shape.waves > kMaxDynLds)
      shape.waves >>= 1;
    // the wave's share of the budget, less its part of the block's static
    // LDS (logf table), so that the budget's occupancy is kept
    const uint32_t wv = (uint32_t)shape.waves;
    const uint32_t fit = (sp.lds_budget * wv - kSelectStaticLds) / wv / 16 * 16;
    a.lds_per_wave = sp.lds_per_wave;
    if (sp.lds_grow && (size_t)fit * wv <= kMaxDynLds)
      a.lds_per_wave = std::max(a.lds_per_wave, fit);
    // Rows whose one-fragment tile does not fit the budget's share would set
    // the share, and so the occupancy, of the whole launch: when there are
    // such rows, a first launch at the budget's share defers them (rows of
    // up to split_l cycles are searched) and a second one searches them.
    a.split_l = sp.lcap;
    a.defer_rows = w.local;  // free until the scan after the select launches
    a.defer_count = w.small + 4;
    a.rows = nullptr;
    a.n_rows_at = nullptr;
    if (lds_bytes(sp.lcap, 1) > fit) {
      uint32_t sl = 0;
      for (uint32_t L = 64; L < sp.lcap && lds_bytes(L, 1) <= fit; L += 64) sl = L;
      if (sl > 0) a.split_l = sl;
    }
    const bool split = a.split_l < sp.lcap;
    const bool h10 = sp.prm.half == 10;
    const bool save = h10 && sp.xwin != nullptr;
    const void *fn =
        save ? (shape.waves == 1   ? (const void *)select_kernel<1, 10, true>
                : shape.waves == 2 ? (const void *)select_kernel<2, 10, true>
                                   : (const void *)select_kernel<4, 10, true>)
        : h10 ? (shape.waves == 1   ? (const void *)select_kernel<1, 10>
                 : shape.waves == 2 ? (const void *)select_kernel<2, 10>
                                    : (const void *)select_kernel<4, 10>)
              : (shape.waves == 1   ? (const void *)select_kernel<1, 0>
                 : shape.waves == 2 ? (const void *)select_kernel<2, 0>
                                    : (const void *)select_kernel<4, 0>);
    auto launch_select_on = [&](const SelArgs &x, hipStream_t s) {
      const size_t lds = (size_t)x.lds_per_wave * shape.waves;
      const int per_cu = occupancy_blocks(fn, shape.waves * kWave, lds);
      const uint64_t want = ((uint64_t)x.n_rows + shape.chunk - 1) / shape.chunk;
      const uint32_t blocks = (uint32_t)std::max<uint64_t>(
          1, std::min<uint64_t>((uint64_t)per_cu * sp.cus,
                                (want + shape.waves - 1) / shape.waves));
      if (save) {
        if (shape.waves == 1) select_kernel<1, 10, true><<<blocks, 64, lds, s>>>(x);
        else if (shape.waves == 2) select_kernel<2, 10, true><<<blocks, 128, lds, s>>>(x);
        else select_kernel<4, 10, true><<<blocks, 256, lds, s>>>(x);
      } else if (h10) {
        if (shape.waves == 1) select_kernel<1, 10><<<blocks, 64, lds, s>>>(x);
        else if (shape.waves == 2) select_kernel<2, 10><<<blocks, 128, lds, s>>>(x);
        else select_kernel<4, 10><<<blocks, 256, lds, s>>>(x);
      } else {
        if (shape.waves == 1) select_kernel<1, 0><<<blocks, 64, lds, s>>>(x);
        else if (shape.waves == 2) select_kernel<2, 0><<<blocks, 128, lds, s>>>(x);
        else select_kernel<4, 0><<<blocks, 256, lds, s>>>(x);
      }
    };
    auto launch_select = [&](const SelArgs &x) { launch_select_on(x, s); };
    if (timed_select) HIP_TRY(hipEventRecord(w.ev[2], s));
    if (pipe && pipe->chunks > 1 && nchunk == 1 && !split && !dbg && !with_bins &&
        w.hsmall_dev && w.table) {
      // chunk k: select, compaction after chunks 0..k-1, table rows, the
      // small words to page-locked memory, then ev_pipe[k]
      pipe->used = true;
      const uint32_t nbl = (n_rows + kScanBlock - 1) / kScanBlock;
      const uint32_t pk = pipe->chunks;
      const uint32_t cc = (uint32_t)sp.prm.cand_count;
      for (uint32_t k = 0; k < pk; ++k) {
        const uint32_t b0 = (uint32_t)((uint64_t)nbl * k / pk);
        const uint32_t b1 = (uint32_t)((uint64_t)nbl * (k + 1) / pk);
        const uint32_t r0 = b0 * kScanBlock;
        const uint32_t n = std::min<uint32_t>(b1 * kScanBlock, n_rows) - r0;
        SelArgs x = a;
        chunk_args(r0, n, x.lib, x.prm, x.plan);
        x.n_rows = n;
        x.queue = w.small + 8 + k;
        x.cand_score += (size_t)r0 * cc;
        x.cand_center += (size_t)r0 * cc;
        x.cand_count += r0;
        x.row0 = r0;
        if (x.xwin) {
          x.xwin += (size_t)r0 * x.xwin_stride;
          x.xwin_ok += r0;
        }
        launch_select(x);
        scan_local_kernel<<<b1 - b0, 256, 0, s>>>(w.cand_count + r0, n, w.local + r0,
                                                  w.bsum + b0);
        scan_blocks_kernel<<<1, 1024, 0, s>>>(w.bsum + b0, b1 - b0, w.small + 16 + k);
        scatter_chunk_kernel<<<(n + 255) / 256, 256, 0, s>>>(
            w.cand_count, w.local, w.bsum + b0, r0, n, cc, w.cand_score, w.cand_center,
            w.small + 16, k, w.out);
        const uint64_t most = (uint64_t)n * std::max<uint32_t>(cc, 1u);
        table_chunk_kernel<<<(uint32_t)((most + 255) / 256), 256, 0, s>>>(
            w.out, w.small + 16, k, pipe->p_idx, pipe->pl, pipe->nrt, pipe->stride,
            w.table);
        publish_kernel<<<1, 64, 0, s>>>(w.small, w.hsmall_dev + (size_t)kSmallWords * k);
        HIP_TRY(hipEventRecord(w.ev_pipe[k], s));
      }
      if (timed_select) HIP_TRY(hipEventRecord(w.ev[3], s));
      HIP_TRY(hipGetLastError());
      return ADSR_OK;
    }
    if (nchunk > 1) {
      // chunk k's plan, then its rows through the first select launch;
      // long rows of every chunk go to the one long-row launch
      const SelArgs b = a;
      for (uint32_t k = 0; k < nchunk; ++k) {
        const uint32_t r0 = k * crows, n = std::min(crows, n_rows - r0);
        SelArgs x = a;
        chunk_args(r0, n, x.lib, x.prm, x.plan);
        x.n_rows = n;
        x.queue = w.small + 8 + k;
        x.cand_score += (size_t)r0 * sp.prm.cand_count;
        x.cand_center += (size_t)r0 * sp.prm.cand_count;
        x.cand_count += r0;
        x.row0 = r0;
        if (x.xwin) {
          x.xwin += (size_t)r0 * x.xwin_stride;
          x.xwin_ok += r0;
        }
        if (split) x.lds_per_wave = fit;
        plan_rows(x.lib, x.prm, x.plan, n, s);
        launch_select(x);
      }
      if (split) {
        SelArgs y = b;
        y.split_l = sp.lcap;
        y.queue = w.small + 3;
        y.rows = w.local;
        y.n_rows_at = w.small + 4;
        launch_select(y);
      }
    } else if (split) {
      SelArgs b = a;
      a.lds_per_wave = fit;
      launch_select(a);
      b.split_l = sp.lcap;
      b.queue = w.small + 3;
      b.rows = w.local;
      b.n_rows_at = w.small + 4;
      launch_select(b);
    } else {
      launch_select(a);
    }
    if (timed_select) HIP_TRY(hipEventRecord(w.ev[3], s));
    const uint32_t nb = (n_rows + kScanBlock - 1) / kScanBlock;
    scan_local_kernel<<<nb, 256, 0, s>>>(w.cand_count, n_rows, w.local, w.bsum);
    scan_blocks_kernel<<<1, 1024, 0, s>>>(w.bsum, nb, total);
    scatter_kernel<<<(n_rows + 255) / 256, 256, 0, s>>>(
        w.cand_count, w.local, w.bsum, n_rows, (uint32_t)sp.prm.cand_count,
        w.cand_score, w.cand_center, w.out);
  }
  HIP_TRY(hipGetLastError());
  return ADSR_OK;
}

static void fill_plan(const adsr_dia *d, const adsr_lib *l,
                      const adsr_selection_params *p, Workspace &w,
                      SearchPlan &sp, size_t ks, size_t cc, size_t row0,
                      size_t n);

static adsr_status build_search(const adsr_dia *d, const adsr_lib *l,
                                const adsr_selection_params *p, Workspace &w,
                                SearchPlan &sp, bool pooled = false,
                                size_t kmax = kMaxSel) {
  set_lds_limits();
  size_t ks = 1;
  adsr_status st = prepare(d, l, p, sp, ks, nullptr, 0, kmax);
  if (st) return st;
  const size_t cc = (size_t)std::min<uint64_t>(p->candidate_count, 1u << 20);
  st = ws_alloc(w, l->n_prec, ks, cc, sp.kernels.size(),
                p->rt_tolerance_per_precursor != nullptr,
                l->max_raw_fragments > (uint64_t)kStashFrag ? l->n_frag : 0,
                pooled ? const_cast<adsr_lib *>(l) : nullptr);
  if (st) return st;
  HIP_TRY(hipMemcpy(w.kernels, sp.kernels.data(), sp.kernels.size() * 4,
                    hipMemcpyHostToDevice));
  if (p->rt_tolerance_per_precursor && l->n_prec)
    HIP_TRY(hipMemcpy(w.rt_tol_pp, p->rt_tolerance_per_precursor,
                      l->n_prec * 4, hipMemcpyHostToDevice));
  fill_plan(d, l, p, w, sp, ks, cc, 0, l->n_prec);
  return ADSR_OK;
}

// The kernel arguments of a search over library rows [row0, row0 + n) in
// workspace w: the library arrays indexed by row start at row0, so the
// plan and outputs are indexed by row - row0 (fragment ranges stay
// absolute).
static void fill_plan(const adsr_dia *d, const adsr_lib *l,
                      const adsr_selection_params *p, Workspace &w,
                      SearchPlan &sp, size_t ks, size_t cc, size_t row0,
                      size_t n) {
  sp.dia = make_kdia(d);
  const DeviceLib &lv = l->dev;
  sp.lib = KLib{(uint32_t)n,        lv.p_mz + row0, lv.p_rt + row0, lv.f_start + row0,
                lv.f_stop + row0,   lv.f_mz,        lv.f_int,       lv.f_type,
                lv.f_num};
  sp.prm.mass_tol = p->mass_tolerance;
  sp.prm.rt_tol = p->rt_tolerance;
  sp.prm.rt_tol_pp = p->rt_tolerance_per_precursor ? w.rt_tol_pp : nullptr;
  sp.prm.pad_cycles = (float)p->kernel_size;
  sp.prm.top_k = (uint32_t)std::min<uint64_t>(p->top_k_fragments, 0xffffffffu);
  sp.prm.cand_count = (uint32_t)cc;
  const size_t klen = (size_t)std::ceil((float)p->kernel_size / 2.0f) * 2;
  sp.prm.half = (uint32_t)(klen / 2);
  sp.prm.klen = (uint32_t)klen;
  sp.prm.per_window = p->sigma_per_window ? 1u : 0u;
  sp.prm.kernels = w.kernels;
  sp.max_raw = l->max_raw_fragments;
  sp.plan = KPlan{w.rec, w.obs, w.sel, w.sel_mz, w.sel_w, w.rng, w.b0,
                  (uint32_t)ks};
}

static adsr_status status_error(uint32_t st) {
  if (st & kStTooManyFragments)
    return set_error(ADSR_ERR_UNSUPPORTED,
                     "a precursor selects more than %d fragments", kMaxSel);
  if (st & kStTooManyWindows)
    return set_error(ADSR_ERR_UNSUPPORTED,
                     "a precursor lies in more than %d isolation windows",
                     kObsMax);
  if (st & kStWindowTooLong)
    return set_error(ADSR_ERR_UNSUPPORTED,
                     "an RT window exceeds %d cycles", kLCapMax);
  if (st & kStSelectFallback)
    return set_error(ADSR_ERR_UNSUPPORTED,
                     "a precursor's top-k fragment selection needs the median-of-"
                     "medians fallback of select_nth_unstable (more than 16 "
                     "partition rounds) or has 4096 or more kept fragments");
  return ADSR_OK;
}

static adsr_status check_status(Workspace &w, uint32_t *total = nullptr) {
  uint32_t host[3] = {0, 0, 0};
  HIP_TRY(hipMemcpyAsync(host, w.small, 12, hipMemcpyDeviceToHost, w.stream));
  HIP_TRY(hipStreamSynchronize(w.stream));
  if (total) *total = host[2];
  return status_error(host[1]);
}

// Queues the host table's columns (table_kernel) behind the search, so
// that one device-to-host copy follows the status read.
static adsr_status queue_table(Workspace &w, const adsr_dia *d, const adsr_lib *l,
                               const adsr_selection_params *p) {
  const size_t cap = std::max<size_t>(w.n_rows, 1) * std::max<size_t>(w.cc, 1);
  if (!w.table)
    HIP_TRY(dev_malloc(&w.table, cap * adsr_candidates::kRowBytes));
  if (w.n_rows)
    table_kernel<<<(uint32_t)((cap + 255) / 256), 256, 0, w.stream>>>(
        w.out, w.small + 2, l->dev.p_idx, p->peak_length, d->rt_index.size(), cap,
        w.table);
  HIP_TRY(hipGetLastError());
  return ADSR_OK;
}

// The table queued by queue_table, `total` rows, into one page-locked block
// (pageable memory when page-locked memory cannot be had).
static adsr_status fetch(Workspace &w, const adsr_lib *l, uint32_t total,
                         adsr_candidates **out) {
  if (l->n_prec == 0) total = 0;
  const size_t bytes = (size_t)total * adsr_candidates::kRowBytes;
  size_t cap = 0;
  bool pinned = true;
  void *blk = pinned_acquire(bytes, &cap);
  if (!blk) {
    pinned = false;
    cap = std::max<size_t>(bytes, 16);
    blk = host_malloc(cap);
    if (!blk) return set_error(ADSR_ERR_OUT_OF_MEMORY, "candidate table");
  }
  auto *c = new adsr_candidates();
  c->pinned = pinned;
  c->block_cap = cap;
  c->bind(blk, total);
  if (bytes) {
    if (hipMemcpyAsync(blk, w.table, bytes, hipMemcpyDeviceToHost, w.stream) !=
            hipSuccess ||
        hipStreamSynchronize(w.stream) != hipSuccess) {
      delete c;
      return set_error(ADSR_ERR_DEVICE, "candidate table copy failed");
    }
  }
  *out = c;
  return ADSR_OK;
}

// A search whose table ends in host memory, pipelined (pipe_chunks): the
// host block is laid out with the device table's stride (one row slot per
// possible candidate), and chunk k's nine column segments are copied on the
// second stream once ev_pipe[k] says that the chunk is tabled, while later
// chunks are still being selected.  When enqueue cannot pipeline the search
// (long rows in a second launch), the table is fetched as usual.
static adsr_status search_to_host_pipelined(Workspace &w, const SearchPlan &sp,
                                            uint32_t n, LaunchShape shape,
                                            const adsr_dia *d, const adsr_lib *l,
                                            const adsr_selection_params *p,
                                            uint32_t chunks, adsr_candidates **out) {
  const size_t cap = std::max<size_t>(w.n_rows, 1) * std::max<size_t>(w.cc, 1);
  if (!w.table)
    HIP_TRY(dev_malloc(&w.table, cap * adsr_candidates::kRowBytes));
  PipeSpec ps;
  ps.chunks = chunks;
  ps.p_idx = l->dev.p_idx;
  ps.pl = p->peak_length;
  ps.nrt = d->rt_index.size();
  ps.stride = cap;
  adsr_status st = enqueue(w, sp, n, shape, nullptr, 0, 0, 0, false, false, false, &ps);
  if (st) return st;
  if (!ps.used) {
    st = queue_table(w, d, l, p);
    uint32_t total = 0;
    if (!st) st = check_status(w, &total);
    if (!st) st = fetch(w, l, total, out);
    return st;
  }
  const size_t bytes = cap * adsr_candidates::kRowBytes;
  size_t bcap = 0;
  bool pinned = true;
  void *blk = pinned_acquire(bytes, &bcap);
  if (!blk) {
    pinned = false;
    bcap = bytes;
    blk = host_malloc(bytes);
    if (!blk) {
      (void)hipStreamSynchronize(w.stream);
      return set_error(ADSR_ERR_OUT_OF_MEMORY, "candidate table");
    }
  }
  auto *c = new adsr_candidates();
  c->pinned = pinned;
  c->block_cap = bcap;
  c->block = blk;
  auto fail = [&](adsr_status e) {
    (void)hipStreamSynchronize(w.stream);
    (void)hipStreamSynchronize(w.stream2);
    delete c;
    return e;
  };
  uint64_t *hb = static_cast<uint64_t *>(blk);
  const uint64_t *db = w.table;
  uint64_t off = 0;
  uint32_t status = 0;
  for (uint32_t k = 0; k < chunks; ++k) {
    if (hipEventSynchronize(w.ev_pipe[k]) != hipSuccess)
      return fail(set_error(ADSR_ERR_DEVICE, "search failed: %s",
                            hipGetErrorString(hipGetLastError())));
    const uint32_t *hs = w.hsmall + (size_t)kSmallWords * k;
    status |= hs[1];
    const uint64_t nk = hs[16 + k];
    if (status || nk == 0) continue;
    if (off + nk > cap) return fail(set_error(ADSR_ERR_INTERNAL, "candidate count"));
    for (int col = 0; col < 8; ++col) {
      if (hipMemcpyAsync(hb + (size_t)col * cap + off, db + (size_t)col * cap + off,
                         nk * 8, hipMemcpyDeviceToHost, w.stream2) != hipSuccess)
        return fail(set_error(ADSR_ERR_DEVICE, "candidate table copy failed"));
    }
    if (hipMemcpyAsync(reinterpret_cast<float *>(hb + 8 * cap) + off,
                       reinterpret_cast<const float *>(db + 8 * cap) + off, nk * 4,
                       hipMemcpyDeviceToHost, w.stream2) != hipSuccess)
      return fail(set_error(ADSR_ERR_DEVICE, "candidate table copy failed"));
    off += nk;
  }
  if (hipStreamSynchronize(w.stream2) != hipSuccess)
    return fail(set_error(ADSR_ERR_DEVICE, "candidate table copy failed"));
  st = status_error(status);
  if (st) return fail(st);
  c->bind(blk, l->n_prec ? off : 0, cap);
  *out = c;
  return ADSR_OK;
}

// The compacted table stays in HBM: rows resolved to precursor_idx.
static adsr_status fetch_device(Workspace &w, const adsr_dia *d,
                                const adsr_lib *l, const adsr_selection_params *p,
                                adsr_device_candidates **out) {
  uint32_t total = 0;
  HIP_TRY(hipMemcpyAsync(&total, w.small + 2, 4, hipMemcpyDeviceToHost,
                         w.stream));
  HIP_TRY(hipStreamSynchronize(w.stream));
  if (l->n_prec == 0) total = 0;
  auto *c = new adsr_device_candidates();
  c->n = total;
  c->pl = p->peak_length;
  c->nrt = d->rt_index.size();
  if (dev_malloc((void **)&c->rows, std::max<size_t>(total, 1) * sizeof(DevCand)) !=
      hipSuccess) {
    delete c;
    return set_error(ADSR_ERR_OUT_OF_MEMORY, "device candidate table");
  }
  if (total)
    resolve_kernel<<<(total + 255) / 256, 256, 0, w.stream>>>(
        w.out, total, l->dev.p_idx, (DevCand *)c->rows);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(w.stream) != hipSuccess) {
    (void)hipFree(c->rows);
    delete c;
    return set_error(ADSR_ERR_DEVICE, "candidate table kernel failed");
  }
  *out = c;
  return ADSR_OK;
}

static std::map<std::tuple<size_t, uint32_t, size_t>, LaunchShape> g_tuned;
static std::mutex g_tuned_mu;

}  // namespace adsr

using namespace adsr;

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" int adsr_device_available(void) { return device_present() ? 1 : 0; }

// The registry knows the reference's names (src/simd/mod.rs:53-58: "scalar",
// and "neon" on aarch64) and "hip".  Only "hip" has an implementation here,
// so the CPU names answer as an unavailable backend does in the reference
// (src/simd/mod.rs:105-111); unknown names list the known ones (:113-118).
extern "C" adsr_status adsr_backend_set(const char *name) {
  clear_error();
  static const char *const kKnown[] = {"scalar", "neon", "hip"};
  bool known = false;
  for (const char *k : kKnown) known = known || (name && std::strcmp(name, k) == 0);
  if (!known)
    return set_error(ADSR_ERR_INVALID_ARGUMENT,
                     "Unknown backend '%s'. Available backends: [\"scalar\", \"neon\", \"hip\"]",
                     name ? name : "(null)");
  if (std::strcmp(name, "hip") != 0 || !device_present())
    return set_error(ADSR_ERR_INVALID_ARGUMENT,
                     "Backend '%s' is not available on this system", name);
  g_backend_override.store(0);
  return ADSR_OK;
}

extern "C" void adsr_backend_clear(void) { g_backend_override.store(-1); }

static adsr_status write_name(char *out, size_t cap, const char *name) {
  if (!out || cap == 0) return set_error(ADSR_ERR_INVALID_ARGUMENT, "no buffer");
  std::snprintf(out, cap, "%s", name);
  return ADSR_OK;
}

extern "C" adsr_status adsr_backend_current(char *out, size_t cap) {
  clear_error();
  return write_name(out, cap, device_present() ? "hip" : "none");
}

extern "C" adsr_status adsr_backend_optimal(char *out, size_t cap) {
  clear_error();
  return write_name(out, cap, device_present() ? "hip" : "none");
}

extern "C" adsr_status adsr_dia_upload(adsr_dia *d) {
  clear_error();
  if (!d) return set_error(ADSR_ERR_INVALID_ARGUMENT, "NULL handle");
  return upload_dia(d);
}

extern "C" adsr_status adsr_lib_upload(adsr_lib *l) {
  clear_error();
  if (!l) return set_error(ADSR_ERR_INVALID_ARGUMENT, "NULL handle");
  return upload_lib(l);
}

// Shared body of adsr_select / adsr_select_timed.
static adsr_status select_common(const adsr_dia *dc, const adsr_lib *lc,
                                 const adsr_selection_params *p, int warmup,
                                 int steps, int *tpb, int *ppw, double *step_ms,
                                 double *kernel_ms, double *graph_ms,
                                 adsr_candidates **out,
                                 adsr_device_candidates **dout = nullptr) {
  clear_error();
  if (!dc || !lc || !p || (!out && !dout))
    return set_error(ADSR_ERR_INVALID_ARGUMENT, "NULL argument");
  if (out) *out = nullptr;
  if (dout) *dout = nullptr;
  adsr_dia *d = const_cast<adsr_dia *>(dc);
  adsr_lib *l = const_cast<adsr_lib *>(lc);
  adsr_status st = upload_dia(d);
  if (st) return st;
  st = upload_lib(l);
  if (st) return st;
  Workspace w;
  SearchPlan sp;
  st = build_search(d, l, p, w, sp, true);
  if (st) {
    ws_free(w);
    return st;
  }
  const uint32_t n = (uint32_t)l->n_prec;
  auto fail = [&](adsr_status s2) {
    ws_free(w);
    return s2;
  };
  auto sync = [&]() -> adsr_status {
    if (hipStreamSynchronize(w.stream) != hipSuccess)
      return set_error(ADSR_ERR_DEVICE, "stream synchronize failed: %s",
                       hipGetErrorString(hipGetLastError()));
    return ADSR_OK;
  };
  // untimed searches use the shape tuned for this (rows, window, fragments)
  // key when there is one, else 256 threads x 2 precursors per grab
  LaunchShape shape{4, 2};
  {
    std::lock_guard<std::mutex> lk(g_tuned_mu);
    auto it = g_tuned.find(std::make_tuple((size_t)n, sp.lcap, w.ks));
    if (it != g_tuned.end()) shape = it->second;
  }
  const int want_tpb = tpb ? *tpb : 0, want_ppw = ppw ? *ppw : 0;
  if (want_tpb == 64 || want_tpb == 128 || want_tpb == 256) {
    shape.waves = want_tpb / 64;
    shape.chunk = (want_ppw == 1 || want_ppw == 2 || want_ppw == 4 ||
                   want_ppw == 8 || want_ppw == 16)
                      ? want_ppw
                      : 1;
  } else if (steps > 0) {
    auto key = std::make_tuple((size_t)n, sp.lcap, w.ks);
    bool cached = false;
    {
      std::lock_guard<std::mutex> lk(g_tuned_mu);
      auto it = g_tuned.find(key);
      if (it != g_tuned.end()) {
        shape = it->second;
        cached = true;
      }
    }
    if (!cached) {
      // autotuner: {64,128,256} threads x {1,2,4,8} precursors per queue
      // grab, each timed on its second launch
      float best = 1e30f;
      for (int wv : {1, 2, 4})
        for (int ch : {1, 2, 4, 8}) {
          LaunchShape sh{wv, ch};
          for (int rep = 0; rep < 2; ++rep) {
            st = enqueue(w, sp, n, sh, nullptr, 0, 0, 0, true);
            if (st) return fail(st);
          }
          st = sync();
          if (st) return fail(st);
          float ms = 0;
          (void)hipEventElapsedTime(&ms, w.ev[2], w.ev[3]);
          if (ms < best) {
            best = ms;
            shape = sh;
          }
        }
      std::lock_guard<std::mutex> lk(g_tuned_mu);
      g_tuned[key] = shape;
    }
  }
  if (tpb) *tpb = shape.waves * 64;
  if (ppw) *ppw = shape.chunk;
  if (steps <= 0 && warmup <= 0 && out && !dout && pipe_chunks(n) > 1) {
    st = search_to_host_pipelined(w, sp, n, shape, d, l, p, pipe_chunks(n), out);
    if (st) return fail(st);
    ws_recycle(l, w);
    return ADSR_OK;
  }
  // first search: checks the device-side limits before anything is timed
  st = enqueue(w, sp, n, shape, nullptr, 0, 0, 0, false);
  if (!st) st = check_status(w);
  if (st) return fail(st);
  for (int i = 0; i < warmup; ++i) {
    st = enqueue(w, sp, n, shape, nullptr, 0, 0, 0, false);
    if (st) return fail(st);
  }
  if (steps > 0) {
    std::vector<hipEvent_t> ev(2 * (size_t)steps, nullptr);
    for (auto &e : ev)
      if (hipEventCreate(&e) != hipSuccess) {
        for (auto &e2 : ev)
          if (e2) (void)hipEventDestroy(e2);
        return fail(set_error(ADSR_ERR_DEVICE, "event creation failed"));
      }
    st = sync();
    if (!st && hipEventRecord(w.ev[0], w.stream) != hipSuccess)
      st = set_error(ADSR_ERR_DEVICE, "event record failed");
    for (int i = 0; !st && i < steps; ++i) {
      w.ev[2] = ev[2 * i];
      w.ev[3] = ev[2 * i + 1];
      st = enqueue(w, sp, n, shape, nullptr, 0, 0, 0, true);
    }
    if (!st && hipEventRecord(w.ev[1], w.stream) != hipSuccess)
      st = set_error(ADSR_ERR_DEVICE, "event record failed");
    if (!st) st = sync();
    double ksum = 0.0;
    float total = 0.0f;
    if (!st) {
      (void)hipEventElapsedTime(&total, w.ev[0], w.ev[1]);
      for (int i = 0; i < steps; ++i) {
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]);
        ksum += ms;
      }
    }
    // the per-step events belong to `ev`; restore the workspace's own pair
    w.ev[2] = nullptr;
    w.ev[3] = nullptr;
    (void)hipEventCreate(&w.ev[2]);
    (void)hipEventCreate(&w.ev[3]);
    for (auto &e : ev) (void)hipEventDestroy(e);
    if (st) return fail(st);
    if (step_ms) *step_ms = total / steps;
    if (kernel_ms) *kernel_ms = ksum / steps;
    if (graph_ms) {
      if (hipStreamBeginCapture(w.stream, hipStreamCaptureModeThreadLocal) !=
          hipSuccess)
        return fail(set_error(ADSR_ERR_DEVICE, "graph capture failed to start"));
      st = enqueue(w, sp, n, shape, nullptr, 0, 0, 0, false);
      hipError_t e = hipStreamEndCapture(w.stream, &w.graph_src);
      if (st) return fail(st);
      if (e != hipSuccess ||
          hipGraphInstantiate(&w.graph, w.graph_src, nullptr, nullptr, 0) !=
              hipSuccess)
        return fail(set_error(ADSR_ERR_DEVICE, "graph capture failed"));
      if (hipGraphLaunch(w.graph, w.stream) != hipSuccess)
        return fail(set_error(ADSR_ERR_DEVICE, "graph launch failed"));
      st = sync();
      if (st) return fail(st);
      (void)hipEventRecord(w.ev[0], w.stream);
      for (int i = 0; i < steps; ++i)
        if (hipGraphLaunch(w.graph, w.stream) != hipSuccess)
          return fail(set_error(ADSR_ERR_DEVICE, "graph launch failed"));
      (void)hipEventRecord(w.ev[1], w.stream);
      st = sync();
      if (st) return fail(st);
      float gms = 0.0f;
      (void)hipEventElapsedTime(&gms, w.ev[0], w.ev[1]);
      *graph_ms = gms / steps;
    }
  }
  if (!dout) {
    st = queue_table(w, d, l, p);
    if (st) return fail(st);
  }
  uint32_t total = 0;
  st = check_status(w, &total);
  if (!st) st = dout ? fetch_device(w, d, l, p, dout) : fetch(w, l, total, out);
  if (st) return fail(st);
  ws_recycle(l, w);
  return ADSR_OK;
}

extern "C" adsr_status adsr_select_device(const adsr_dia *d, const adsr_lib *l,
                                          const adsr_selection_params *p,
                                          adsr_device_candidates **out) {
  if (!out) {
    clear_error();
    return set_error(ADSR_ERR_INVALID_ARGUMENT, "NULL argument");
  }
  return select_common(d, l, p, 0, 0, nullptr, nullptr, nullptr, nullptr,
                       nullptr, nullptr, out);
}

// ---------------------------------------------------------------------------
// search without allocation: every buffer in one file-scope device array
// ---------------------------------------------------------------------------
// Callers that cannot allocate device memory (code under HIP graph capture,
// serving loops, harnesses that forbid hipMalloc) search library rows
// [row_begin, row_end) with adsr_select_rows_noalloc.  The plan, candidate
// slots, compaction and queue live in `scratch`, each GPU's own copy of
// this array; its address on the current device is looked up once per
// device with hipGetSymbolAddress, and only that address is used on the
// host.  Every call on a GPU shares the array: calls that overlap on one
// GPU race.  The queue and status words are cleared by a kernel at the
// start of every call; every other buffer is written before it is read.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_at[kScratchDevices];
constexpr size_t kNoAllocTaps = 4096;  // kernel taps (one kernel per window)

static adsr_status scratch_base(int dev, unsigned char **out) {
  if (dev < 0 || dev >= kScratchDevices)
    return set_error(ADSR_ERR_UNSUPPORTED, "device %d beyond the scratch table", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_at[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(ADSR_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed: %s",
                       hipGetErrorString(e));
    g_scratch_at[dev] = p;
  }
  *out = static_cast<unsigned char *>(g_scratch_at[dev]);
  return ADSR_OK;
}

extern "C" adsr_status adsr_select_rows_noalloc(
    const adsr_dia *d, const adsr_lib *l, const adsr_selection_params *p,
    size_t row_begin, size_t row_end, uint64_t *precursor_idx, uint64_t *rank,
    float *score, uint64_t *scan_center, uint64_t *scan_start, uint64_t *scan_stop,
    uint64_t *cycle_center, uint64_t *cycle_start, uint64_t *cycle_stop,
    size_t capacity, size_t *n_out, void *stream) {
  clear_error();
  if (!d || !l || !p || !n_out)
    return set_error(ADSR_ERR_INVALID_ARGUMENT, "NULL argument");
  *n_out = 0;
  if (capacity && (!precursor_idx || !rank || !score || !scan_center || !scan_start ||
                   !scan_stop || !cycle_center || !cycle_start || !cycle_stop))
    return set_error(ADSR_ERR_INVALID_ARGUMENT, "column buffers are NULL");
  if (row_begin > row_end || row_end > l->n_prec)
    return set_error(ADSR_ERR_INVALID_ARGUMENT, "rows [%zu, %zu) outside the library of %zu",
                     row_begin, row_end, l->n_prec);
  if (!d->dev.ready || !l->dev.ready)
    return set_error(ADSR_ERR_INVALID_ARGUMENT,
                     "the DIAData and the library must be on the device first "
                     "(adsr_dia_upload, adsr_lib_upload)");
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (d->dev.device != dev || l->dev.device != dev)
    return set_error(ADSR_ERR_INVALID_ARGUMENT,
                     "the handles live on another device than the current one");
  if (l->max_raw_fragments > (uint64_t)kStashFrag)
    return set_error(ADSR_ERR_UNSUPPORTED,
                     "fragment lists longer than %d are not searched without "
                     "allocation", kStashFrag);
  unsigned char *base = nullptr;
  adsr_status st = scratch_base(dev, &base);
  if (st) return st;
  set_lds_limits();
  SearchPlan sp;
  size_t ks = 1;
  float taps[kNoAllocTaps];
  st = prepare(d, l, p, sp, ks, taps, kNoAllocTaps);
  if (st) return st;
  const size_t n = row_end - row_begin;
  if (n == 0) return ADSR_OK;
  const size_t cc = (size_t)std::min<uint64_t>(p->candidate_count, 1u << 20);
  const size_t ccs = std::max<size_t>(cc, 1);  // slots per row
  const bool pp = p->rt_tolerance_per_precursor != nullptr;
  // layout: each buffer 256-byte aligned
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += (bytes + 255) / 256 * 256;
    return at;
  };
  const size_t nb = (n + kScanBlock - 1) / kScanBlock;
  const size_t o_small = take(kSmallWords * 4), o_rec = take(n * sizeof(uint4)),
               o_obs = take(n * kObsMax * 2), o_sel = take(n * ks * 2),
               o_smz = take(n * ks * 4), o_sw = take(n * ks * 4),
               o_cs = take(n * ccs * 4), o_cc = take(n * ccs * 4), o_cnt = take(n * 4),
               o_loc = take(n * 4), o_bsum = take(nb * 4),
               o_out = take(n * ccs * sizeof(CandRow)), o_ker = take(sp.n_taps * 4),
               o_rt = take(pp ? n * 4 : 0);
  if (off > kScratchBytes) {
    const size_t per_row = off / n + 1;
    return set_error(ADSR_ERR_INVALID_ARGUMENT,
                     "rows [%zu, %zu) need %zu bytes of scratch, the limit is %zu: "
                     "search at most %zu rows per call", row_begin, row_end, off,
                     kScratchBytes, kScratchBytes / per_row);
  }
  Workspace w;
  w.n_rows = n;
  w.ks = ks;
  w.cc = cc;
  w.nker = sp.n_taps;
  w.small = reinterpret_cast<uint32_t *>(base + o_small);
  w.rec = reinterpret_cast<uint4 *>(base + o_rec);
  w.obs = reinterpret_cast<uint16_t *>(base + o_obs);
  w.sel = reinterpret_cast<uint16_t *>(base + o_sel);
  w.sel_mz = reinterpret_cast<float *>(base + o_smz);
  w.sel_w = reinterpret_cast<float *>(base + o_sw);
  w.cand_score = reinterpret_cast<float *>(base + o_cs);
  w.cand_center = reinterpret_cast<uint32_t *>(base + o_cc);
  w.cand_count = reinterpret_cast<uint32_t *>(base + o_cnt);
  w.local = reinterpret_cast<uint32_t *>(base + o_loc);
  w.bsum = reinterpret_cast<uint32_t *>(base + o_bsum);
  w.out = reinterpret_cast<CandRow *>(base + o_out);
  w.kernels = reinterpret_cast<float *>(base + o_ker);
  w.rt_tol_pp = pp ? reinterpret_cast<float *>(base + o_rt) : nullptr;
  w.stream = (hipStream_t)stream;
  // scratch memory that is copied, not allocated: none of it is freed
  auto release = [&](adsr_status s2) {
    w = Workspace();
    return s2;
  };
  if (hipMemcpyAsync(w.kernels, taps, sp.n_taps * 4, hipMemcpyHostToDevice, w.stream) !=
          hipSuccess ||
      (pp && hipMemcpyAsync(w.rt_tol_pp, p->rt_tolerance_per_precursor + row_begin, n * 4,
                            hipMemcpyHostToDevice, w.stream) != hipSuccess))
    return release(set_error(ADSR_ERR_DEVICE, "copy of taps or tolerances failed"));
  fill_plan(d, l, p, w, sp, ks, cc, row_begin, n);
  LaunchShape shape{4, 2};
  {
    std::lock_guard<std::mutex> lk(g_tuned_mu);
    auto it = g_tuned.find(std::make_tuple(n, sp.lcap, ks));
    if (it != g_tuned.end()) shape = it->second;
  }
  st = enqueue(w, sp, (uint32_t)n, shape, nullptr, 0, 0, 0, false);
  if (st) return release(st);
  if (capacity) {
    const uint64_t cap_rows = (uint64_t)n * ccs;
    columns_kernel<<<(uint32_t)((cap_rows + 255) / 256), 256, 0, w.stream>>>(
        w.out, w.small + 2, l->dev.p_idx + row_begin, p->peak_length,
        d->rt_index.size(), capacity, precursor_idx, rank, score, scan_center,
        scan_start, scan_stop, cycle_center, cycle_start, cycle_stop);
    if (hipGetLastError() != hipSuccess)
      return release(set_error(ADSR_ERR_DEVICE, "column kernel launch failed"));
  }
  uint32_t total = 0;
  st = check_status(w, &total);
  if (st) return release(st);
  *n_out = total;
  if (total > capacity)
    return release(set_error(ADSR_ERR_INVALID_ARGUMENT,
                             "%u candidates found, the columns hold %zu: the first %zu "
                             "are written", total, capacity, capacity));
  return release(ADSR_OK);
}

extern "C" size_t adsr_device_candidates_len(const adsr_device_candidates *c) {
  return c ? c->n : 0;
}

extern "C" adsr_status adsr_device_candidates_export(
    const adsr_device_candidates *c, uint64_t *precursor_idx, uint64_t *rank,
    float *score, uint64_t *scan_center, uint64_t *scan_start,
    uint64_t *scan_stop, uint64_t *cycle_center, uint64_t *cycle_start,
    uint64_t *cycle_stop, void *stream) {
  clear_error();
  if (!c) return set_error(ADSR_ERR_INVALID_ARGUMENT, "NULL handle");
  if (c->n == 0) return ADSR_OK;
  export_kernel<<<(uint32_t)((c->n + 255) / 256), 256, 0, (hipStream_t)stream>>>(
      (const DevCand *)c->rows, c->n, c->pl, c->nrt, precursor_idx, rank, score,
      scan_center, scan_start, scan_stop, cycle_center, cycle_start, cycle_stop);
  HIP_TRY(hipGetLastError());
  return ADSR_OK;
}

extern "C" adsr_status adsr_device_candidates_to_host(
    const adsr_device_candidates *c, adsr_candidates **out) {
  clear_error();
  if (!c || !out) return set_error(ADSR_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  const size_t n = c->n;
  const size_t bytes = n * adsr_candidates::kRowBytes;
  size_t cap = 0;
  bool pinned = true;
  void *blk = pinned_acquire(bytes, &cap);
  if (!blk) {
    pinned = false;
    cap = std::max<size_t>(bytes, 16);
    blk = host_malloc(cap);
    if (!blk) return set_error(ADSR_ERR_OUT_OF_MEMORY, "candidate table");
  }
  auto *h = new adsr_candidates();
  h->pinned = pinned;
  h->block_cap = cap;
  h->bind(blk, n);
  if (n) {
    // the columns expanded on the device, then one copy
    uint64_t *dblk = nullptr;
    if (dev_malloc(&dblk, bytes) != hipSuccess) {
      delete h;
      return set_error(ADSR_ERR_OUT_OF_MEMORY, "device table block");
    }
    export_kernel<<<(uint32_t)((n + 255) / 256), 256>>>(
        (const DevCand *)c->rows, n, c->pl, c->nrt, dblk, dblk + n,
        reinterpret_cast<float *>(dblk + 8 * n), dblk + 2 * n, dblk + 3 * n,
        dblk + 4 * n, dblk + 5 * n, dblk + 6 * n, dblk + 7 * n);
    const bool ok = hipGetLastError() == hipSuccess &&
                    hipMemcpy(blk, dblk, bytes, hipMemcpyDeviceToHost) == hipSuccess;
    (void)hipFree(dblk);
    if (!ok) {
      delete h;
      return set_error(ADSR_ERR_DEVICE, "candidate table copy failed");
    }
  }
  *out = h;
  return ADSR_OK;
}

extern "C" void adsr_device_candidates_free(adsr_device_candidates *c) {
  if (!c) return;
  (void)hipFree(c->rows);
  delete c;
}

extern "C" adsr_status adsr_select(const adsr_dia *d, const adsr_lib *l,
                                   const adsr_selection_params *p,
                                   adsr_candidates **out) {
  return select_common(d, l, p, 0, 0, nullptr, nullptr, nullptr, nullptr,
                       nullptr, out);
}

extern "C" adsr_status adsr_select_timed(
    const adsr_dia *d, const adsr_lib *l, const adsr_selection_params *p,
    int warmup, int steps, int *tpb, int *ppw, double *step_ms,
    double *kernel_ms, double *graph_ms, adsr_candidates **out) {
  return select_common(d, l, p, std::max(warmup, 0), std::max(steps, 0), tpb,
                       ppw, step_ms, kernel_ms, graph_ms, out);
}

extern "C" adsr_status adsr_select_algorithmic_bytes(
    const adsr_dia *dc, const adsr_lib *lc, const adsr_selection_params *p,
    double *bytes, double *entries, double *hits) {
  clear_error();
  if (!dc || !lc || !p || !bytes)
    return set_error(ADSR_ERR_INVALID_ARGUMENT, "NULL argument");
  adsr_dia *d = const_cast<adsr_dia *>(dc);
  adsr_lib *l = const_cast<adsr_lib *>(lc);
  adsr_status st = upload_dia(d);
  if (!st) st = upload_lib(l);
  if (st) return st;
  Workspace w;
  SearchPlan sp;
  st = build_search(d, l, p, w, sp);
  const uint32_t n = (uint32_t)l->n_prec;
  if (!st) st = enqueue(w, sp, n, LaunchShape(), nullptr, 0, 0, 0, false, false, true);
  if (!st) st = check_status(w);
  unsigned long long *acc = nullptr;
  if (!st && dev_malloc(&acc, 32) != hipSuccess)
    st = set_error(ADSR_ERR_OUT_OF_MEMORY, "accumulator");
  if (!st) {
    unsigned long long zero[4] = {0, 0, 0, 0};
    (void)hipMemcpy(acc, zero, 32, hipMemcpyHostToDevice);
    const uint64_t nt = (uint64_t)n * sp.plan.ks;
    if (nt)
      traffic_kernel<<<(uint32_t)((nt + 255) / 256), 256, 0, w.stream>>>(
          sp.dia, sp.lib, sp.prm, sp.plan, acc);
    unsigned long long host[3] = {0, 0, 0};
    uint32_t total = 0;
    if (hipStreamSynchronize(w.stream) != hipSuccess ||
        hipMemcpy(host, acc, 24, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(&total, w.small + 2, 4, hipMemcpyDeviceToHost) != hipSuccess)
      st = set_error(ADSR_ERR_DEVICE, "traffic count failed");
    if (n == 0) total = 0;
    *bytes = (double)host[0] + 68.0 * total;
    if (entries) *entries = (double)host[1];
    if (hits) *hits = (double)host[2];
  }
  (void)hipFree(acc);
  ws_free(w);
  return st;
}

extern "C" adsr_status adsr_select_scores(const adsr_dia *dc,
                                          const adsr_lib *lc,
                                          const adsr_selection_params *p,
                                          size_t row_begin, size_t row_end,
                                          size_t max_cycles, float *scores,
                                          uint64_t *cstart, uint64_t *cstop) {
  clear_error();
  if (!dc || !lc || !p || !scores || !cstart || !cstop)
    return set_error(ADSR_ERR_INVALID_ARGUMENT, "NULL argument");
  adsr_dia *d = const_cast<adsr_dia *>(dc);
  adsr_lib *l = const_cast<adsr_lib *>(lc);
  if (row_end > l->n_prec || row_begin > row_end)
    return set_error(ADSR_ERR_INVALID_ARGUMENT, "row range out of bounds");
  adsr_status st = upload_dia(d);
  if (!st) st = upload_lib(l);
  if (st) return st;
  Workspace w;
  SearchPlan sp;
  st = build_search(d, l, p, w, sp);
  if (st) {
    ws_free(w);
    return st;
  }
  const size_t nr = row_end - row_begin;
  float *dbg = nullptr;
  const size_t bytes = std::max<size_t>(nr * max_cycles, 1) * 4;
  if (dev_malloc(&dbg, bytes) != hipSuccess) {
    ws_free(w);
    return set_error(ADSR_ERR_OUT_OF_MEMORY, "debug buffer");
  }
  std::vector<float> nanv(nr * max_cycles, __builtin_nanf(""));
  (void)hipMemcpy(dbg, nanv.data(), nanv.size() * 4, hipMemcpyHostToDevice);
  LaunchShape shape;
  st = enqueue(w, sp, (uint32_t)l->n_prec, shape, dbg, (uint32_t)max_cycles,
               (uint32_t)row_begin, (uint32_t)nr, false);
  if (!st) st = check_status(w);
  if (!st && nr) {
    if (hipMemcpy(scores, dbg, nr * max_cycles * 4, hipMemcpyDeviceToHost) !=
        hipSuccess)
      st = set_error(ADSR_ERR_DEVICE, "copy of scores failed");
    std::vector<uint4> rec(nr);
    if (!st && hipMemcpy(rec.data(), w.rec + row_begin, nr * sizeof(uint4),
                         hipMemcpyDeviceToHost) != hipSuccess)
      st = set_error(ADSR_ERR_DEVICE, "copy of plan failed");
    for (size_t i = 0; !st && i < nr; ++i) {
      cstart[i] = rec[i].x;
      cstop[i] = rec[i].y;
    }
  }
  (void)hipFree(dbg);
  ws_free(w);
  return st;
}

extern "C" adsr_status adsr_select_plan(const adsr_dia *dc, const adsr_lib *lc,
                                        const adsr_selection_params *p,
                                        size_t row_begin, size_t row_end,
                                        uint64_t *cstart, uint64_t *cstop,
                                        uint32_t *n_selected,
                                        uint64_t *selected, size_t stride) {
  clear_error();
  if (!dc || !lc || !p || !cstart || !cstop || !n_selected || !selected)
    return set_error(ADSR_ERR_INVALID_ARGUMENT, "NULL argument");
  adsr_dia *d = const_cast<adsr_dia *>(dc);
  adsr_lib *l = const_cast<adsr_lib *>(lc);
  if (row_end > l->n_prec || row_begin > row_end)
    return set_error(ADSR_ERR_INVALID_ARGUMENT, "row range out of bounds");
  adsr_status st = upload_dia(d);
  if (!st) st = upload_lib(l);
  if (st) return st;
  Workspace w;
  SearchPlan sp;
  st = build_search(d, l, p, w, sp);
  if (!st)
    st = enqueue(w, sp, (uint32_t)l->n_prec, LaunchShape(), nullptr, 0, 0, 0,
                 false, true);
  if (!st) st = check_status(w);
  const size_t nr = row_end - row_begin;
  if (!st && nr) {
    std::vector<uint4> rec(nr);
    std::vector<uint16_t> sel(nr * w.ks);
    if (hipMemcpy(rec.data(), w.rec + row_begin, nr * sizeof(uint4),
                  hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(sel.data(), w.sel + row_begin * w.ks, nr * w.ks * 2,
                  hipMemcpyDeviceToHost) != hipSuccess)
      st = set_error(ADSR_ERR_DEVICE, "copy of plan failed");
    for (size_t i = 0; !st && i < nr; ++i) {
      cstart[i] = rec[i].x;
      cstop[i] = rec[i].y;
      const uint32_t ns = rec[i].z & 0xffffu;
      n_selected[i] = ns;
      const uint64_t fs = l->frag_start[row_begin + i];
      for (uint32_t j = 0; j < ns && j < stride; ++j)
        selected[i * stride + j] = fs + sel[i * w.ks + j];
    }
  }
  ws_free(w);
  return st;
}

// ---- differentiable selection ---------------------------------------------
namespace adsr {
struct DiffState {
  adsr_dia *d = nullptr;
  adsr_lib *l = nullptr;
  adsr_selection_params p{};
  std::vector<float> sig;
  Workspace w;
  SearchPlan sp;
  double *dk = nullptr;
  double *contrib = nullptr;
  double *gsig = nullptr;
  double *gpart = nullptr;  // per-window, per-chunk sigma partial sums
  uint32_t *bq = nullptr;
  // the forward's apex XIC windows (sp.xwin, sp.xwin_ok), when allocated
  float *xwin = nullptr;
  uint32_t *xwin_ok = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  float fwd_ms = 0.0f, bwd_ms = 0.0f;
  bool have_forward = false;
  // host copies of the taps and their sigma derivatives, uploaded on the
  // forward's stream; alive until the next forward, which starts after the
  // previous one has been waited for
  std::vector<float> taps_h;
  std::vector<double> dk_h;
  // a forward enqueued by adsr_diff_forward_async whose device status has
  // not been read yet, and the caller's stream it runs on
  bool pending = false;
  hipStream_t pending_stream = nullptr;
};

// Waits for a pending asynchronous forward and reports its device status.
static adsr_status diff_settle(DiffState *s) {
  if (!s->pending) return ADSR_OK;
  s->pending = false;
  HIP_TRY(hipStreamSynchronize(s->pending_stream));
  const adsr_status st = check_status(s->w);
  if (st) {
    s->have_forward = false;
    return st;
  }
  (void)hipEventElapsedTime(&s->fwd_ms, s->ev[0], s->ev[1]);
  return ADSR_OK;
}

// d k[j] / d sigma for the taps j = 0..half of gaussian_kernel_1d
// (src/kernel/mod.rs:32-56): k = e / sum(e), e_j = exp(-x_j^2 / (2 sigma^2)),
// x_j = j - half over the kernel_width taps, so
// dk_j/dsigma = k_j (x_j^2 - sum_m k_m x_m^2) / sigma^3.
static void kernel_dsigma(double sigma, size_t width, double *out) {
  const int half = (int)(width / 2);
  std::vector<double> e(width);
  double S = 0.0;
  for (int j = 0; j < (int)width; ++j) {
    const double x = j - half;
    e[j] = std::exp(-x * x / (2.0 * sigma * sigma));
    S += e[j];
  }
  double m2 = 0.0;
  for (int j = 0; j < (int)width; ++j) {
    const double x = j - half;
    m2 += e[j] / S * x * x;
  }
  for (int j = 0; j <= half; ++j) {
    const double x = j - half;
    out[j] = (e[j] / S) * (x * x - m2) / (sigma * sigma * sigma);
  }
}
}  // namespace adsr
extern "C" {

adsr_status adsr_diff_create(const adsr_dia *dc, const adsr_lib *lc,
                             const adsr_selection_params *p, adsr_diff **out) {
  clear_error();
  if (!dc || !lc || !p || !out)
    return set_error(ADSR_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  adsr_dia *d = const_cast<adsr_dia *>(dc);
  adsr_lib *l = const_cast<adsr_lib *>(lc);
  if (!l->frag_ranges_disjoint)
    return set_error(ADSR_ERR_UNSUPPORTED,
                     "precursors share fragment indices; per-fragment "
                     "gradients need disjoint fragment ranges");
  adsr_status st = upload_dia(d);
  if (!st) st = upload_lib(l);
  if (st) return st;
  auto *s = new DiffState();
  s->d = d;
  s->l = l;
  s->p = *p;
  s->sig.assign(d->n_obs, 2.0f);
  s->p.sigma_per_window = s->sig.data();
  s->p.n_sigma_per_window = d->n_obs;
  s->p.rt_tolerance_per_precursor = nullptr;
  // the per-precursor tolerances are uploaded into the workspace here; the
  // caller's array only has to live during this call
  s->p.rt_tolerance_per_precursor = p->rt_tolerance_per_precursor;
  st = build_search(d, l, &s->p, s->w, s->sp);
  s->p.rt_tolerance_per_precursor = nullptr;
  const size_t nk = 1 + d->n_obs;
  const size_t half = s->sp.prm.half;
  auto fail = [&](adsr_status e) {
    ws_free(s->w);
    (void)hipFree(s->dk);
    (void)hipFree(s->contrib);
    (void)hipFree(s->gsig);
    (void)hipFree(s->gpart);
    (void)hipFree(s->bq);
    (void)hipFree(s->xwin);
    (void)hipFree(s->xwin_ok);
    delete s;
    return e;
  };
  if (st) return fail(st);
  if (dev_malloc(&s->dk, nk * (half + 1) * 8) != hipSuccess ||
      dev_malloc(&s->contrib, std::max<size_t>(l->n_prec, 1) * 8) != hipSuccess ||
      dev_malloc(&s->gsig, std::max<size_t>(d->n_obs, 1) * 8) != hipSuccess ||
      dev_malloc(&s->gpart, std::max<size_t>(d->n_obs, 1) *
                               ((l->n_prec + kSigmaChunk - 1) / kSigmaChunk + 1) * 8) !=
          hipSuccess ||
      dev_malloc(&s->bq, 64) != hipSuccess)
    return fail(set_error(ADSR_ERR_OUT_OF_MEMORY, "diff buffers"));
  // Apex XIC windows: the backward reads the forward's instead of
  // re-extracting them (ks x (2 half + 1) floats per row; the forward
  // writes them when the half width is 10).  Without them (another width,
  // ADSR_DIFF_REEXTRACT=1 at creation, or the allocation failed) every row
  // is re-extracted, with the same results.
  {
    const char *rx = std::getenv("ADSR_DIFF_REEXTRACT");
    const bool reextract = rx && rx[0] && rx[0] != '0';
    const size_t stride = (size_t)s->sp.plan.ks * (2 * half + 1);
    if (!reextract && half == 10 && stride <= 0xffffffffu && l->n_prec &&
        dev_malloc(&s->xwin, (size_t)l->n_prec * stride * 4) == hipSuccess &&
        dev_malloc(&s->xwin_ok, (size_t)l->n_prec * 4) == hipSuccess) {
      s->sp.xwin = s->xwin;
      s->sp.xwin_ok = s->xwin_ok;
      s->sp.xwin_stride = (uint32_t)stride;
    } else {
      (void)hipGetLastError();
      (void)hipFree(s->xwin);
      s->xwin = nullptr;
    }
  }
  for (auto &e : s->ev)
    if (hipEventCreate(&e) != hipSuccess)
      return fail(set_error(ADSR_ERR_DEVICE, "event creation failed"));
  *out = reinterpret_cast<adsr_diff *>(s);
  return ADSR_OK;
}

// The forward on the state's own stream (async false: waits for the device
// and reads the status) or enqueued on the caller's stream (async true).
static adsr_status diff_forward(DiffState *s, const float *sigma,
                                const float *weights, float *s_star,
                                uint8_t *has, bool async, hipStream_t stream) {
  clear_error();
  if (!s || !sigma || !weights || !s_star)
    return set_error(ADSR_ERR_INVALID_ARGUMENT, "NULL argument");
  adsr_status pst = diff_settle(s);
  if (pst) return pst;
  const adsr_dia *d = s->d;
  const adsr_lib *l = s->l;
  const size_t klen = s->sp.prm.klen, half = s->sp.prm.half;
  const size_t nk = 1 + d->n_obs;
  s->sig.assign(sigma, sigma + d->n_obs);
  for (size_t o = 0; o < d->n_obs; ++o)
    if (!(s->sig[o] > 0.0f) || !std::isfinite(s->sig[o]))
      return set_error(ADSR_ERR_INVALID_ARGUMENT, "sigma[%zu] must be > 0", o);
  // taps: index 0 from fwhm_rt, 1 + o from sigma[o]
  std::vector<float> &taps = s->taps_h;
  taps.assign(nk * klen, 0.0f);
  adsr_gaussian_kernel(s->p.fwhm_rt, 1.0f, s->p.kernel_size, 1.0f, taps.data(), klen);
  for (size_t o = 0; o < d->n_obs; ++o)
    adsr_gaussian_kernel_sigma(s->sig[o], klen, taps.data() + (1 + o) * klen, klen);
  std::vector<double> &dk = s->dk_h;
  dk.assign(nk * (half + 1), 0.0);
  kernel_dsigma((double)(s->p.fwhm_rt / 2.3548f), klen, dk.data());
  for (size_t o = 0; o < d->n_obs; ++o)
    kernel_dsigma((double)s->sig[o], klen, dk.data() + (1 + o) * (half + 1));
  // the workspace's launches go to the caller's stream for an asynchronous
  // forward; its own stream is put back before returning
  const hipStream_t own = s->w.stream;
  const hipStream_t fs = async ? stream : own;
  s->w.stream = fs;
  adsr_status st = [&]() -> adsr_status {
    HIP_TRY(hipMemcpyAsync(s->w.kernels, taps.data(), taps.size() * 4,
                           hipMemcpyHostToDevice, fs));
    HIP_TRY(hipMemcpyAsync(s->dk, dk.data(), dk.size() * 8, hipMemcpyHostToDevice, fs));
    s->sp.lib.f_int = weights;
    LaunchShape shape{4, 2};
    const uint32_t n = (uint32_t)l->n_prec;
    HIP_TRY(hipEventRecord(s->ev[0], fs));
    adsr_status e = enqueue(s->w, s->sp, n, shape, nullptr, 0, 0, 0, false);
    if (e) return e;
    if (n)
      sstar_kernel<<<(n + 255) / 256, 256, 0, fs>>>(
          s->w.cand_score, s->w.cand_count, (uint32_t)s->sp.prm.cand_count, n,
          s_star, has);
    HIP_TRY(hipEventRecord(s->ev[1], fs));
    return ADSR_OK;
  }();
  s->w.stream = own;
  if (st) {
    // work may have been queued: wait for it before the state is reused
    (void)hipStreamSynchronize(fs);
    s->have_forward = false;
    return st;
  }
  s->have_forward = true;
  if (async) {
    s->pending = true;
    s->pending_stream = fs;
    return ADSR_OK;
  }
  st = check_status(s->w);
  if (st) {
    s->have_forward = false;
    return st;
  }
  (void)hipEventElapsedTime(&s->fwd_ms, s->ev[0], s->ev[1]);
  return ADSR_OK;
}

adsr_status adsr_diff_forward(adsr_diff *h, const float *sigma,
                              const float *weights, float *s_star,
                              uint8_t *has) {
  return diff_forward(reinterpret_cast<DiffState *>(h), sigma, weights, s_star,
                      has, false, nullptr);
}

adsr_status adsr_diff_forward_async(adsr_diff *h, const float *sigma,
                                    const float *weights, float *s_star,
                                    uint8_t *has, void *stream) {
  return diff_forward(reinterpret_cast<DiffState *>(h), sigma, weights, s_star,
                      has, true, (hipStream_t)stream);
}

adsr_status adsr_diff_backward(adsr_diff *h, const float *grad_s,
                               float *grad_w, double *grad_sigma) {
  clear_error();
  auto *s = reinterpret_cast<DiffState *>(h);
  if (!s || !grad_s || !grad_w || !grad_sigma)
    return set_error(ADSR_ERR_INVALID_ARGUMENT, "NULL argument");
  adsr_status pst = diff_settle(s);
  if (pst) return pst;
  if (!s->have_forward)
    return set_error(ADSR_ERR_INVALID_ARGUMENT, "backward before forward");
  const adsr_dia *d = s->d;
  const adsr_lib *l = s->l;
  hipStream_t st_ = s->w.stream;
  const uint32_t n = (uint32_t)l->n_prec;
  HIP_TRY(hipEventRecord(s->ev[2], st_));
  if (l->n_frag)
    fill_f32_kernel<<<(uint32_t)((l->n_frag + 255) / 256), 256, 0, st_>>>(
        grad_w, l->n_frag, 0.0f);
  if (n) {
    BwdArgs a;
    a.dia = s->sp.dia;
    a.lib = s->sp.lib;
    a.prm = s->sp.prm;
    a.plan = s->sp.plan;
    a.n_rows = n;
    a.cc = (uint32_t)s->sp.prm.cand_count;
    a.cand_center = s->w.cand_center;
    a.cand_count = s->w.cand_count;
    a.grad_s = grad_s;
    a.dk = s->dk;
    a.grad_w = grad_w;
    a.contrib = s->contrib;
    a.queue = nullptr;
    a.xwin = s->sp.xwin;
    a.xwin_ok = s->sp.xwin_ok;
    a.xwin_stride = s->sp.xwin_stride;
    // four rows (waves) per block
    const size_t lds = 4 * (size_t)bwd_lds_bytes(s->sp.plan.ks);
    diff_backward_kernel<<<(n + 3) / 4, 256, lds, st_>>>(a);
    if (d->n_obs) {
      const uint32_t nch = (n + kSigmaChunk - 1) / kSigmaChunk;
      sigma_partial_kernel<<<dim3(nch, (uint32_t)d->n_obs), 256, 0, st_>>>(
          s->sp.plan.rec, s->contrib, n, s->gpart);
      sigma_final_kernel<<<(uint32_t)d->n_obs, 64, 0, st_>>>(s->gpart, nch, s->gsig);
    }
  }
  HIP_TRY(hipEventRecord(s->ev[3], st_));
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st_));
  std::vector<double> g(d->n_obs, 0.0);
  if (n && d->n_obs)
    HIP_TRY(hipMemcpy(g.data(), s->gsig, d->n_obs * 8, hipMemcpyDeviceToHost));
  std::memcpy(grad_sigma, g.data(), d->n_obs * 8);
  (void)hipEventElapsedTime(&s->bwd_ms, s->ev[2], s->ev[3]);
  return ADSR_OK;
}

adsr_status adsr_hinge_loss(const float *s_star, size_t n_rows,
                            const int64_t *targets, const int64_t *decoys,
                            size_t n_pairs, float margin, float *terms,
                            float *grad_s, void *stream) {
  clear_error();
  adsr_status st = require_device();
  if (st) return st;
  if (n_pairs == 0) return ADSR_OK;
  if (!s_star || !targets || !decoys || !terms || !grad_s)
    return set_error(ADSR_ERR_INVALID_ARGUMENT, "NULL argument");
  const uint32_t blocks =
      (uint32_t)std::min<uint64_t>((n_pairs + 255) / 256, 1u << 16);
  hinge_kernel<<<blocks, 256, 0, (hipStream_t)stream>>>(
      s_star, n_rows, targets, decoys, n_pairs, margin, terms, grad_s);
  HIP_TRY(hipGetLastError());
  return ADSR_OK;
}

adsr_status adsr_diff_timings(const adsr_diff *h, double *fwd, double *bwd) {
  // a pending forward is waited for (the state itself is not const)
  auto *s = reinterpret_cast<DiffState *>(const_cast<adsr_diff *>(h));
  if (!s) return set_error(ADSR_ERR_INVALID_ARGUMENT, "NULL handle");
  adsr_status pst = diff_settle(s);
  if (pst) return pst;
  if (fwd) *fwd = s->fwd_ms;
  if (bwd) *bwd = s->bwd_ms;
  return ADSR_OK;
}

void adsr_diff_free(adsr_diff *h) {
  auto *s = reinterpret_cast<DiffState *>(h);
  if (!s) return;
  // kernels of an asynchronous forward may still use the workspace
  if (s->pending) (void)hipStreamSynchronize(s->pending_stream);
  ws_free(s->w);
  (void)hipFree(s->dk);
  (void)hipFree(s->contrib);
  (void)hipFree(s->gsig);
  (void)hipFree(s->gpart);
  (void)hipFree(s->bq);
  (void)hipFree(s->xwin);
  (void)hipFree(s->xwin_ok);
  for (auto &e : s->ev)
    if (e) (void)hipEventDestroy(e);
  delete s;
}

}  // extern "C"

// ---- per-stage operations --------------------------------------------------
template <class F>
static adsr_status with_buffers(size_t in_bytes, size_t out_bytes, F &&fn) {
  void *din = nullptr, *dout = nullptr;
  HIP_TRY(dev_malloc(&din, std::max<size_t>(in_bytes, 16)));
  if (dev_malloc(&dout, std::max<size_t>(out_bytes, 16)) != hipSuccess) {
    (void)hipFree(din);
    return set_error(ADSR_ERR_OUT_OF_MEMORY, "device buffer");
  }
  adsr_status st = fn(din, dout);
  (void)hipFree(din);
  (void)hipFree(dout);
  return st;
}

extern "C" adsr_status adsr_convolution(const float *kernel, size_t klen,
                                        const float *xic, size_t rows,
                                        size_t cols, float *out) {
  clear_error();
  adsr_status st = require_device();
  if (st) return st;
  if (klen == 0)
    return set_error(ADSR_ERR_INVALID_ARGUMENT, "empty kernel");
  const size_t n = rows * cols;
  if (n == 0) return ADSR_OK;
  return with_buffers((n + klen) * 4, n * 4, [&](void *din, void *dout) {
    float *dk = (float *)din, *dx = dk + klen;
    HIP_TRY(hipMemcpy(dk, kernel, klen * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dx, xic, n * 4, hipMemcpyHostToDevice));
    conv_dense_kernel<<<(uint32_t)((n + 255) / 256), 256>>>(
        dk, (uint32_t)klen, dx, (uint32_t)rows, (uint32_t)cols, (float *)dout);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(out, dout, n * 4, hipMemcpyDeviceToHost));
    return ADSR_OK;
  });
}

extern "C" adsr_status adsr_axis_log_dot_product(const float *array,
                                                 size_t rows, size_t cols,
                                                 const float *weights,
                                                 float *out) {
  clear_error();
  adsr_status st = require_device();
  if (st) return st;
  if (cols == 0) return ADSR_OK;
  const size_t n = rows * cols;
  return with_buffers((n + rows) * 4, cols * 4, [&](void *din, void *dout) {
    float *da = (float *)din, *dw = da + n;
    if (n) HIP_TRY(hipMemcpy(da, array, n * 4, hipMemcpyHostToDevice));
    if (rows) HIP_TRY(hipMemcpy(dw, weights, rows * 4, hipMemcpyHostToDevice));
    logdot_kernel<<<(uint32_t)((cols + 255) / 256), 256>>>(
        da, (uint32_t)rows, (uint32_t)cols, dw, (float *)dout);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(out, dout, cols * 4, hipMemcpyDeviceToHost));
    return ADSR_OK;
  });
}

extern "C" adsr_status adsr_device_logf(const float *x, size_t n, float *out) {
  clear_error();
  adsr_status st = require_device();
  if (st) return st;
  if (n == 0) return ADSR_OK;
  return with_buffers(n * 4, n * 4, [&](void *din, void *dout) {
    HIP_TRY(hipMemcpy(din, x, n * 4, hipMemcpyHostToDevice));
    logf_kernel<<<(uint32_t)((n + 255) / 256), 256>>>((const float *)din, n,
                                                      (float *)dout);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(out, dout, n * 4, hipMemcpyDeviceToHost));
    return ADSR_OK;
  });
}

extern "C" adsr_status adsr_find_local_maxima(const float *array, size_t n,
                                              uint64_t offset,
                                              uint64_t *indices, float *values,
                                              size_t *n_out) {
  clear_error();
  adsr_status st = require_device();
  if (st) return st;
  if (!n_out) return set_error(ADSR_ERR_INVALID_ARGUMENT, "n_out is NULL");
  const size_t cap = n;  // at most n/2 maxima
  return with_buffers(n * 4, cap * 12 + 16, [&](void *din, void *dout) {
    if (n) HIP_TRY(hipMemcpy(din, array, n * 4, hipMemcpyHostToDevice));
    uint64_t *di = (uint64_t *)dout;
    float *dv = (float *)(di + cap);
    uint32_t *dc = (uint32_t *)(dv + cap);
    maxima_kernel<<<1, 64>>>((const float *)din, (uint32_t)n, offset, di, dv,
                             dc);
    HIP_TRY(hipGetLastError());
    uint32_t m = 0;
    HIP_TRY(hipMemcpy(&m, dc, 4, hipMemcpyDeviceToHost));
    if (m && indices) HIP_TRY(hipMemcpy(indices, di, m * 8, hipMemcpyDeviceToHost));
    if (m && values) HIP_TRY(hipMemcpy(values, dv, m * 4, hipMemcpyDeviceToHost));
    *n_out = m;
    return ADSR_OK;
  });
}

extern "C" adsr_status adsr_dense_xic(const adsr_dia *dc, float pmz,
                                      uint64_t c0, uint64_t c1, float tol,
                                      const float *fmz, size_t nf,
                                      float *out) {
  clear_error();
  if (!dc) return set_error(ADSR_ERR_INVALID_ARGUMENT, "NULL handle");
  if (c1 < c0) return set_error(ADSR_ERR_INVALID_ARGUMENT, "cycle_stop < cycle_start");
  adsr_dia *d = const_cast<adsr_dia *>(dc);
  adsr_status st = upload_dia(d);
  if (st) return st;
  const size_t L = c1 - c0;
  if (nf == 0 || L == 0) return ADSR_OK;
  const Grid &g = global_grid();
  const DeviceDia &dv = d->dev;
  KDia kd = make_kdia(d);
  return with_buffers(nf * 4, nf * L * 4, [&](void *din, void *dout) {
    HIP_TRY(hipMemcpy(din, fmz, nf * 4, hipMemcpyHostToDevice));
    dense_xic_kernel<<<(uint32_t)((nf + 63) / 64), 64>>>(
        kd, pmz, (uint32_t)c0, (uint32_t)c1, tol, (const float *)din,
        (uint32_t)nf, (float *)dout);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(out, dout, nf * L * 4, hipMemcpyDeviceToHost));
    return ADSR_OK;
  });
}

// Host instance of the device ln; used to check the device code path
// against the platform libm on machines without a GPU.
extern "C" float adsr_logf_host(float x) { return logf_exact(x, kHostLogfTab); }

// ===========================================================================
// DIAData CSR built in HBM: src/dia_data_builder/mod.rs:52-139
// ===========================================================================
// The reference pushes, per observation, one (bin, cycle, intensity) triple
// per peak in spectrum order, stable-sorts the triples by (bin, cycle) and
// turns the bins into slice_starts.  Here the host orders each
// observation's spectra by (u16 cycle, spectrum index) -- the order a stable
// sort by cycle gives -- and the device
//   1. snaps every peak to its 1-ppm bin (find_closest_index) and writes
//      (bin, cycle << 32 | intensity bits) in that order,
//   2. stable-sorts each observation's entries by bin: three 8-bit LSD
//      radix passes over 4096-entry tiles that never straddle two
//      observations (tile histograms, one exclusive scan, stable ranking
//      with wave ballots, scatter through LDS),
//   3. marks the end of every bin's run and takes a per-observation prefix
//      maximum, which is slice_starts.
// The result equals the host builder's CSR bit for bit.

struct SnapItem {
  uint64_t src;    // first peak
  uint64_t dst;    // first entry (global, observation-major)
  uint32_t n;      // peaks
  uint32_t cycle;  // u16 cycle index of the spectrum
};
constexpr uint32_t kSnapPeaks = 2048;  // peaks per snap item

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;  // entries per tile
constexpr int kSortWaves = kSortThreads / kWave;
constexpr int kSortRun = kSortTile / kSortWaves;  // entries per wave
constexpr int kMaxScanChunk = 8192;               // prefix-max chunk

// src/mz_index/mod.rs:55-90: nearest grid point, ties to the lower index,
// exact grid values returned as themselves, NaN and values at or below the
// first point to 0, values beyond the last point to n - 1.  From x's
// 4-byte bucket entry: its count gives the lower bound, and the grid points
// next to x are rebuilt from the bucket bits and the stored ulp offsets --
// in x's bucket, else in the next (previous) non-empty one, which is at most
// two buckets away because 1-ppm steps span 8.4 to 16.8 ulps.
__device__ __forceinline__ uint32_t closest_from(const KDia &dia, float x,
                                                 uint32_t e) {
  const uint32_t bx = __float_as_uint(x), xo = bx & 15u;
  const uint32_t key = (bx >> kBucketShift) - dia.base_key;
  const uint32_t c = e & 0x3fffffu, nin = (e >> 22) & 3u;
  const uint32_t o0 = (e >> 24) & 15u, o1 = e >> 28;
  const bool p0 = nin >= 1, p1 = nin == 2;
  if (p0 && o0 == xo) return c;
  if (p1 && o1 == xo) return c + 1;
  const uint32_t lb = c + (p0 && o0 < xo ? 1u : 0u) + (p1 && o1 < xo ? 1u : 0u);
  const uint32_t kb = bx & ~15u;
  uint32_t above, below;
  if (p0 && o0 > xo) {
    above = kb | o0;
  } else if (p1 && o1 > xo) {
    above = kb | o1;
  } else {
    uint32_t k2 = key + 1, e2 = dia.bucket1[k2];
    if (((e2 >> 22) & 3u) == 0) e2 = dia.bucket1[++k2];
    above = ((k2 + dia.base_key) << kBucketShift) | ((e2 >> 24) & 15u);
  }
  if (p1 && o1 < xo) {
    below = kb | o1;
  } else if (p0 && o0 < xo) {
    below = kb | o0;
  } else {
    uint32_t k2 = key - 1, e2 = dia.bucket1[k2];
    if (((e2 >> 22) & 3u) == 0) e2 = dia.bucket1[--k2];
    below = ((k2 + dia.base_key) << kBucketShift) |
            (((e2 >> 22) & 3u) == 2 ? (e2 >> 28) : ((e2 >> 24) & 15u));
  }
  const float left = fabsf(__uint_as_float(above) - x);
  const float right = fabsf(__uint_as_float(below) - x);
  return left < right ? lb : lb - 1;
}

// Each thread snaps kSnapPeaks / 256 peaks of an item, with the m/z loads,
// then the bucket loads, of all of them in flight together.  (Loading the
// two neighbouring buckets up front as well measured no faster: the extra
// divergent gathers cost what the shorter dependency chain saves.)
__global__ __launch_bounds__(256) void snap_kernel(KDia dia, const SnapItem *items,
                                                   const float *mz, const float *inten,
                                                   uint32_t *keys, uint64_t *vals) {
  constexpr int kPer = kSnapPeaks / 256;
  const SnapItem it = items[blockIdx.x];
  float x[kPer];
  uint32_t ib[kPer], e[kPer];
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const uint32_t i = threadIdx.x + 256u * j;
    x[j] = 0.0f;
    ib[j] = 0;
    if (i < it.n) {
      x[j] = mz[it.src + i];
      ib[j] = __float_as_uint(inten[it.src + i]);
    }
  }
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    e[j] = 0;
    const bool inside = x[j] > dia.grid_first && !(x[j] > dia.grid_last);
    if (threadIdx.x + 256u * j < it.n && inside)
      e[j] = dia.bucket1[(__float_as_uint(x[j]) >> kBucketShift) - dia.base_key];
  }
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const uint32_t i = threadIdx.x + 256u * j;
    if (i < it.n) {
      uint32_t b;
      if (!(x[j] > dia.grid_first)) b = 0;  // also NaN
      else if (x[j] > dia.grid_last) b = dia.n_grid - 1;
      else b = closest_from(dia, x[j], e[j]);
      keys[it.dst + i] = b;
      vals[it.dst + i] = ((uint64_t)it.cycle << 32) | ib[j];
    }
  }
}

struct SortGeo {
  const uint64_t *ebase;   // n_obs + 1 entry bases
  const uint32_t *tfirst;  // n_obs + 1 first tiles
  uint32_t n_obs;
};

struct SortTile {
  uint32_t o;      // observation
  uint32_t tl;     // tile within the observation
  uint32_t nt;     // tiles of the observation
  uint32_t cnt;    // entries in this tile
  uint32_t pos0;   // first entry of the tile, relative to the observation
  uint64_t eb;     // entry base of the observation
  uint64_t n_o;    // entries of the observation
};

// Observation o with tfirst[o] <= tile < tfirst[o + 1] (never an empty one).
__device__ __forceinline__ SortTile sort_tile(const SortGeo &g, uint32_t tile) {
  uint32_t lo = 0, hi = g.n_obs;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (g.tfirst[mid] <= tile) lo = mid;
    else hi = mid;
  }
  SortTile t;
  t.o = lo;
  t.tl = tile - g.tfirst[lo];
  t.nt = g.tfirst[lo + 1] - g.tfirst[lo];
  t.eb = g.ebase[lo];
  t.n_o = g.ebase[lo + 1] - t.eb;
  t.pos0 = t.tl * (uint32_t)kSortTile;
  t.cnt = (uint32_t)min((uint64_t)kSortTile, t.n_o - t.pos0);
  return t;
}

// Digit histogram of one tile, stored digit-major within the observation:
// H[tfirst[o] * 256 + d * nt + tl], so that one exclusive scan over H gives
// every (observation, digit, tile) its output offset.
__global__ __launch_bounds__(kSortThreads) void radix_upsweep_kernel(
    SortGeo g, const uint32_t *keys, uint32_t shift, uint32_t *H) {
  __shared__ uint32_t hist[kSortWaves][256];
  for (int i = threadIdx.x; i < kSortWaves * 256; i += kSortThreads)
    (&hist[0][0])[i] = 0;
  __syncthreads();
  const SortTile t = sort_tile(g, blockIdx.x);
  const int lane = lane_id(), w = threadIdx.x / kWave;
  uint32_t k[kSortItems];
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const uint32_t p = (uint32_t)(w * kSortRun + j * kWave + lane);
    k[j] = p < t.cnt ? keys[t.eb + t.pos0 + p] : 0u;
  }
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const uint32_t p = (uint32_t)(w * kSortRun + j * kWave + lane);
    if (p < t.cnt) atomicAdd(&hist[w][(k[j] >> shift) & 255u], 1u);
  }
  __syncthreads();
  const uint32_t d = threadIdx.x;
  uint32_t c = 0;
#pragma unroll
  for (int q = 0; q < kSortWaves; ++q) c += hist[q][d];
  H[(size_t)g.tfirst[t.o] * 256 + (size_t)d * t.nt + t.tl] = c;
}

// Stable scatter of one tile by digit.  Wave w ranks entries
// [w * kSortRun, (w + 1) * kSortRun) of the tile 64 at a time: lanes with
// equal digits find each other with 8 ballots, and a per-wave running count
// per digit in LDS gives each entry its rank among the earlier entries of
// its digit.  Entries are then placed in LDS in digit order and written out
// in runs per digit.  The scan offsets are taken modulo 2^32 and made
// relative to the observation, which holds fewer than 2^32 entries.
template <bool FINAL>
__global__ __launch_bounds__(kSortThreads) void radix_downsweep_kernel(
    SortGeo g, const uint32_t *kin, const uint64_t *vin, uint32_t shift,
    const uint32_t *hscan, const uint32_t *hbsum, uint32_t *kout,
    uint64_t *vout, uint16_t *cout, float *iout) {
  __shared__ uint32_t whist[kSortWaves][256];
  __shared__ uint32_t goff[256];
  __shared__ uint32_t wsum[kSortWaves];
  __shared__ uint64_t stage[kSortTile];
  for (int i = threadIdx.x; i < kSortWaves * 256; i += kSortThreads)
    (&whist[0][0])[i] = 0;
  __syncthreads();
  const SortTile t = sort_tile(g, blockIdx.x);
  const int lane = lane_id(), w = threadIdx.x / kWave;
  const uint64_t src0 = t.eb + t.pos0;
  uint32_t key[kSortItems], rk[kSortItems];
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const uint32_t p = (uint32_t)(w * kSortRun + j * kWave + lane);
    key[j] = p < t.cnt ? kin[src0 + p] : 0u;
  }
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const uint32_t p = (uint32_t)(w * kSortRun + j * kWave + lane);
    const bool valid = p < t.cnt;
    const uint32_t d = (key[j] >> shift) & 255u;
    uint64_t m = ballot64(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = ballot64(bit);
      m &= bit ? bb : ~bb;
    }
    const uint32_t below = __builtin_amdgcn_mbcnt_hi(
        (uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    uint32_t pre = 0;
    if (valid) pre = whist[w][d];
    if (valid && below == 0) whist[w][d] = pre + (uint32_t)__builtin_popcountll(m);
    rk[j] = pre + below;
  }
  __syncthreads();
  // thread d owns digit d: tile-local start of the digit, start of each
  // wave's entries within it, and the digit's output offset
  {
    const uint32_t d = threadIdx.x;
    uint32_t c[kSortWaves], tot = 0;
#pragma unroll
    for (int q = 0; q < kSortWaves; ++q) {
      c[q] = whist[q][d];
      tot += c[q];
    }
    const uint32_t incl = wave_incl_scan(tot);
    if (lane == kWave - 1) wsum[w] = incl;
    __syncthreads();
    uint32_t dstart = incl - tot;
    for (int q = 0; q < w; ++q) dstart += wsum[q];
    uint32_t run = dstart;
#pragma unroll
    for (int q = 0; q < kSortWaves; ++q) {
      whist[q][d] = run;
      run += c[q];
    }
    const size_t hidx = (size_t)g.tfirst[t.o] * 256 + (size_t)d * t.nt + t.tl;
    const uint32_t rel = hscan[hidx] + hbsum[hidx / kScanBlock] - (uint32_t)t.eb;
    goff[d] = rel - dstart;
  }
  __syncthreads();
  uint32_t *skey = reinterpret_cast<uint32_t *>(stage);
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const uint32_t p = (uint32_t)(w * kSortRun + j * kWave + lane);
    if (p < t.cnt) {
      const uint32_t d = (key[j] >> shift) & 255u;
      rk[j] += whist[w][d];  // slot in the digit-ordered tile
      skey[rk[j]] = key[j];
    }
  }
  __syncthreads();
  uint32_t dst[kSortItems];  // observation-relative destination of slot q
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const uint32_t q = (uint32_t)(j * kSortThreads) + threadIdx.x;
    dst[j] = 0;
    if (q < t.cnt) {
      const uint32_t k = skey[q];
      dst[j] = goff[(k >> shift) & 255u] + q;
      kout[t.eb + dst[j]] = k;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const uint32_t p = (uint32_t)(w * kSortRun + j * kWave + lane);
    if (p < t.cnt) stage[rk[j]] = vin[src0 + p];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const uint32_t q = (uint32_t)(j * kSortThreads) + threadIdx.x;
    if (q < t.cnt) {
      const uint64_t v = stage[q];
      if (FINAL) {
        cout[t.eb + dst[j]] = (uint16_t)(v >> 32);
        iout[t.eb + dst[j]] = __uint_as_float((uint32_t)v);
      } else {
        vout[t.eb + dst[j]] = v;
      }
    }
  }
}

// slice_starts[o][b + 1] = end of bin b's run in the sorted observation;
// the other entries stay 0 until the prefix maximum fills them.
__global__ __launch_bounds__(kSortThreads) void run_end_kernel(
    SortGeo g, const uint32_t *keys, uint32_t n_grid, uint32_t *ss) {
  const SortTile t = sort_tile(g, blockIdx.x);
  uint32_t *sso = ss + (size_t)t.o * (n_grid + 1);
  for (uint32_t p = threadIdx.x; p < t.cnt; p += kSortThreads) {
    const uint64_t r = (uint64_t)t.pos0 + p;  // position in the observation
    const uint32_t k = keys[t.eb + r];
    const bool last = r + 1 == t.n_o || keys[t.eb + r + 1] != k;
    if (last) sso[k + 1] = (uint32_t)(r + 1);
  }
}

__device__ __forceinline__ uint32_t wave_incl_max(uint32_t v) {
  const int lane = lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t t = __shfl_up(v, o, kWave);
    if (lane >= o) v = max(v, t);
  }
  return v;
}

// Per (observation, chunk of kMaxScanChunk entries of slice_starts): max.
__global__ __launch_bounds__(256) void chunk_max_kernel(const uint32_t *ss,
                                                        uint32_t len,
                                                        uint32_t *cm) {
  __shared__ uint32_t wm[4];
  const uint32_t o = blockIdx.y, c = blockIdx.x;
  const uint32_t *row = ss + (size_t)o * len;
  uint32_t m = 0;
  for (uint32_t i = c * kMaxScanChunk + threadIdx.x;
       i < min(len, (c + 1) * (uint32_t)kMaxScanChunk); i += 256)
    m = max(m, row[i]);
  m = wave_incl_max(m);
  if (lane_id() == kWave - 1) wm[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0)
    cm[(size_t)o * gridDim.x + c] = max(max(wm[0], wm[1]), max(wm[2], wm[3]));
}

// Exclusive prefix maximum over the chunk maxima of each observation.
__global__ __launch_bounds__(256) void chunk_scan_kernel(uint32_t *cm,
                                                         uint32_t nc) {
  __shared__ uint32_t wm[4];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  uint32_t *row = cm + (size_t)blockIdx.x * nc;
  for (uint32_t c0 = 0; c0 < nc; c0 += 256) {
    const uint32_t i = c0 + threadIdx.x;
    const uint32_t v = i < nc ? row[i] : 0u;
    const uint32_t incl = wave_incl_max(v);
    const int wv = threadIdx.x / kWave;
    if (lane_id() == kWave - 1) wm[wv] = incl;
    __syncthreads();
    uint32_t before = carry;
    for (int q = 0; q < wv; ++q) before = max(before, wm[q]);
    uint32_t excl = __shfl_up(incl, 1, kWave);
    if (lane_id() == 0) excl = 0;
    if (i < nc) row[i] = max(before, excl);
    __syncthreads();
    if (threadIdx.x == 255) carry = max(before, incl);
    __syncthreads();
  }
}

// In-place inclusive prefix maximum of each chunk, seeded with the maximum
// of the chunks before it: slice_starts.  The chunk is loaded and stored
// coalesced through LDS; thread t scans entries 32t .. 32t + 31, stored at
// a stride of 33 words so that the 32 threads of a bank cycle hit 32 banks.
__global__ __launch_bounds__(256) void chunk_fill_kernel(uint32_t *ss,
                                                         uint32_t len,
                                                         const uint32_t *cmx) {
  constexpr int kPer = kMaxScanChunk / 256;
  __shared__ uint32_t buf[kMaxScanChunk + kMaxScanChunk / 32];
  __shared__ uint32_t wm[4];
  const uint32_t o = blockIdx.y, c = blockIdx.x;
  uint32_t *row = ss + (size_t)o * len + (size_t)c * kMaxScanChunk;
  const uint32_t n = min((uint32_t)kMaxScanChunk, len - c * kMaxScanChunk);
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const uint32_t i = (uint32_t)j * 256 + threadIdx.x;
    buf[i + i / 32] = i < n ? row[i] : 0u;
  }
  __syncthreads();
  uint32_t *mine = buf + threadIdx.x * (kPer + 1);
  uint32_t m = 0;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    m = max(m, mine[j]);
    mine[j] = m;
  }
  const uint32_t incl = wave_incl_max(m);
  const int wv = threadIdx.x / kWave;
  if (lane_id() == kWave - 1) wm[wv] = incl;
  __syncthreads();
  uint32_t before = cmx[(size_t)o * gridDim.x + c];
  for (int q = 0; q < wv; ++q) before = max(before, wm[q]);
  uint32_t excl = __shfl_up(incl, 1, kWave);
  if (lane_id() == 0) excl = 0;
  before = max(before, excl);
#pragma unroll
  for (int j = 0; j < kPer; ++j) mine[j] = max(before, mine[j]);
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const uint32_t i = (uint32_t)j * 256 + threadIdx.x;
    if (i < n) row[i] = buf[i + i / 32];
  }
}

namespace {
// Device buffers of one build, freed together.
struct BuildBufs {
  std::vector<void *> p;
  ~BuildBufs() {
    // a build that failed part-way may have kernels queued on them
    (void)hipDeviceSynchronize();
    for (void *q : p) (void)hipFree(q);
  }
  template <class T>
  adsr_status alloc(T **out, size_t n) {
    *out = nullptr;
    HIP_TRY(dev_malloc((void **)out, std::max<size_t>(n, 1) * sizeof(T)));
    p.push_back(*out);
    return ADSR_OK;
  }
  void release(void *q) {
    for (auto &x : p)
      if (x == q) {
        (void)hipFree(x);
        x = nullptr;
      }
  }
};
}  // namespace

#define BTRY(expr)                   \
  do {                               \
    adsr_status s_ = (expr);         \
    if (s_) return s_;               \
  } while (0)

// Builds d's CSR on the device from the peaks; d->dev holds the metadata.
static adsr_status build_csr_device(adsr_dia *d,
                                    const std::vector<std::vector<size_t>> &spectra,
                                    const int64_t *pstart, const int64_t *pstop,
                                    const int64_t *cyc, const float *peak_mz,
                                    const float *peak_int, size_t n_peaks) {
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  const Grid &g = global_grid();
  const uint32_t B = (uint32_t)g.mz.size();
  const size_t n_obs = d->n_obs;
  const std::vector<uint64_t> &eb = d->entry_base;
  const uint64_t total = eb[n_obs];
  // snap items: each observation's spectra by (u16 cycle, spectrum index)
  std::vector<SnapItem> items;
  std::vector<uint32_t> tfirst(n_obs + 1, 0);
  uint64_t min_src = ~0ull, max_src = 0;
  for (size_t o = 0; o < n_obs; ++o) {
    std::vector<size_t> sp = spectra[o];
    auto by_cycle = [&](size_t a, size_t b) {
      return (uint16_t)cyc[a] < (uint16_t)cyc[b];
    };
    if (!std::is_sorted(sp.begin(), sp.end(), by_cycle))
      std::stable_sort(sp.begin(), sp.end(), by_cycle);
    uint64_t dst = eb[o];
    for (size_t s : sp) {
      const uint64_t a = (uint64_t)pstart[s], b = (uint64_t)pstop[s];
      if (a >= b) continue;
      min_src = std::min(min_src, a);
      max_src = std::max(max_src, b);
      for (uint64_t q = a; q < b; q += kSnapPeaks) {
        SnapItem it;
        it.src = q;
        it.dst = dst;
        it.n = (uint32_t)std::min<uint64_t>(kSnapPeaks, b - q);
        it.cycle = (uint16_t)cyc[s];
        items.push_back(it);
        dst += it.n;
      }
    }
    const uint64_t n_o = eb[o + 1] - eb[o];
    tfirst[o + 1] = tfirst[o] + (uint32_t)((n_o + kSortTile - 1) / kSortTile);
  }
  const uint32_t n_tiles = tfirst[n_obs];
  if ((uint64_t)n_tiles * 256 >= (1ull << 32) || items.size() >= (1ull << 31))
    return set_error(ADSR_ERR_UNSUPPORTED, "run too large for the device builder");
  const auto t1 = clk::now();

  BuildBufs tmp;
  DeviceDia &v = d->dev;
  size_t acc = v.bytes;
  // peaks referenced by some observation, copied as one range
  const uint64_t lo = total ? min_src : 0, np = total ? max_src - min_src : 0;
  float *dmz = nullptr, *dint = nullptr;
  BTRY(tmp.alloc(&dmz, np));
  BTRY(tmp.alloc(&dint, np));
  if (np) {
    HIP_TRY(hipMemcpy(dmz, peak_mz + lo, np * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dint, peak_int + lo, np * 4, hipMemcpyHostToDevice));
  }
  for (auto &it : items) it.src -= lo;
  SnapItem *ditems = nullptr;
  uint32_t *dtf = nullptr;
  BTRY(tmp.alloc(&ditems, items.size()));
  BTRY(tmp.alloc(&dtf, tfirst.size()));
  HIP_TRY(hipMemcpy(ditems, items.data(), items.size() * sizeof(SnapItem),
                    hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dtf, tfirst.data(), tfirst.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipDeviceSynchronize());
  const auto t2 = clk::now();

  uint32_t *keys[2];
  uint64_t *vals[2];
  BTRY(tmp.alloc(&keys[0], total));
  BTRY(tmp.alloc(&keys[1], total));
  BTRY(tmp.alloc(&vals[0], total));
  BTRY(tmp.alloc(&vals[1], total));
  const size_t nh = (size_t)n_tiles * 256;
  const size_t nhb = (nh + kScanBlock - 1) / kScanBlock;
  uint32_t *H, *hscan, *hbsum, *htot;
  BTRY(tmp.alloc(&H, nh));
  BTRY(tmp.alloc(&hscan, nh));
  BTRY(tmp.alloc(&hbsum, nhb));
  BTRY(tmp.alloc(&htot, 1));
  // CSR arrays of the handle
  const size_t len = (size_t)B + 1;
  HIP_TRY(dev_malloc((void **)&v.slice_starts, std::max<size_t>(n_obs * len, 1) * 4));
  HIP_TRY(dev_malloc((void **)&v.cycles, (total + 8) * 2));
  HIP_TRY(dev_malloc((void **)&v.intens, std::max<uint64_t>(total, 1) * 4));
  acc += n_obs * len * 4 + (total + 8) * 2 + total * 4;
  HIP_TRY(hipMemset(v.cycles + total, 0xff, 8 * 2));
  if (n_obs) HIP_TRY(hipMemset(v.slice_starts, 0, n_obs * len * 4));

  const KDia kd = make_kdia(d);
  SortGeo geo;
  geo.ebase = v.entry_base;
  geo.tfirst = dtf;
  geo.n_obs = (uint32_t)n_obs;
  if (!items.empty()) {
    snap_kernel<<<(uint32_t)items.size(), 256>>>(kd, ditems, dmz, dint, keys[0], vals[0]);
    HIP_TRY(hipGetLastError());
  }
  if (n_tiles) {
    int cur = 0;
    for (int pass = 0; pass < 3; ++pass) {
      const uint32_t shift = 8u * (uint32_t)pass;
      radix_upsweep_kernel<<<n_tiles, kSortThreads>>>(geo, keys[cur], shift, H);
      HIP_TRY(hipGetLastError());
      scan_local_kernel<<<(uint32_t)nhb, 256>>>(H, (uint32_t)nh, hscan, hbsum);
      HIP_TRY(hipGetLastError());
      scan_blocks_kernel<<<1, 1024>>>(hbsum, (uint32_t)nhb, htot);
      HIP_TRY(hipGetLastError());
      if (pass < 2) {
        radix_downsweep_kernel<false><<<n_tiles, kSortThreads>>>(
            geo, keys[cur], vals[cur], shift, hscan, hbsum, keys[cur ^ 1],
            vals[cur ^ 1], nullptr, nullptr);
      } else {
        radix_downsweep_kernel<true><<<n_tiles, kSortThreads>>>(
            geo, keys[cur], vals[cur], shift, hscan, hbsum, keys[cur ^ 1],
            nullptr, v.cycles, v.intens);
      }
      HIP_TRY(hipGetLastError());
      cur ^= 1;
    }
    run_end_kernel<<<n_tiles, kSortThreads>>>(geo, keys[cur], B, v.slice_starts);
    HIP_TRY(hipGetLastError());
  }
  if (n_obs) {
    const uint32_t nc = (uint32_t)((len + kMaxScanChunk - 1) / kMaxScanChunk);
    uint32_t *cm;
    BTRY(tmp.alloc(&cm, n_obs * nc));
    chunk_max_kernel<<<dim3(nc, (uint32_t)n_obs), 256>>>(v.slice_starts, (uint32_t)len, cm);
    HIP_TRY(hipGetLastError());
    chunk_scan_kernel<<<(uint32_t)n_obs, 256>>>(cm, nc);
    HIP_TRY(hipGetLastError());
    chunk_fill_kernel<<<dim3(nc, (uint32_t)n_obs), 256>>>(v.slice_starts, (uint32_t)len, cm);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipDeviceSynchronize());
  const auto t3 = clk::now();
  v.bytes = acc;
  auto ms = [](clk::duration x) { return std::chrono::duration<double, std::milli>(x).count(); };
  d->build_ms[0] = ms(t1 - t0);
  d->build_ms[1] = ms(t2 - t1);
  d->build_ms[2] = ms(t3 - t2);
  return ADSR_OK;
}
#undef BTRY

extern "C" adsr_status adsr_dia_create_device(
    const int64_t *delta, const float *iso_lo, const float *iso_hi,
    const int64_t *pstart, const int64_t *pstop, const int64_t *cyc,
    const float *srt, size_t n_spectra, const float *peak_mz,
    const float *peak_int, size_t n_peaks, const float *cycle,
    const size_t cycle_shape[4], adsr_dia **out) {
  clear_error();
  if (!out) return set_error(ADSR_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  adsr_status st = require_device();
  if (st) return st;
  std::vector<std::vector<size_t>> spectra;
  adsr_dia *d = nullptr;
  st = dia_prepare(delta, iso_lo, iso_hi, pstart, pstop, cyc, srt, n_spectra,
                   peak_mz, peak_int, n_peaks, cycle, cycle_shape, false,
                   spectra, &d);
  if (st) return st;
  {
    std::lock_guard<std::mutex> lk(d->mu);
    size_t acc = 0;
    st = upload_dia_meta(d, acc);
    if (!st) {
      d->dev.bytes = acc;
      st = build_csr_device(d, spectra, pstart, pstop, cyc, peak_mz, peak_int,
                            n_peaks);
    }
    if (st) {
      device_dia_free(d->dev);
    } else {
      d->dev.ready = true;
      (void)hipGetDevice(&d->dev.device);
    }
  }
  if (st) {
    delete d;
    return st;
  }
  *out = d;
  return ADSR_OK;
}

extern "C" adsr_status adsr_dia_build_timings(const adsr_dia *d, double *prep_ms,
                                              double *copy_ms, double *build_ms) {
  clear_error();
  if (!d) return set_error(ADSR_ERR_INVALID_ARGUMENT, "NULL handle");
  if (prep_ms) *prep_ms = d->build_ms[0];
  if (copy_ms) *copy_ms = d->build_ms[1];
  if (build_ms) *build_ms = d->build_ms[2];
  return ADSR_OK;
}

namespace adsr {
adsr_status device_dia_window(const adsr_dia *d, size_t o, uint32_t *ss,
                              uint16_t *ci, float *in) {
  const DeviceDia &v = d->dev;
  if (!v.ready) return set_error(ADSR_ERR_DEVICE, "DIAData has no device CSR");
  const size_t len = d->n_grid + 1;
  const uint64_t a = d->entry_base[o], b = d->entry_base[o + 1];
  if (ss) HIP_TRY(hipMemcpy(ss, v.slice_starts + o * len, len * 4, hipMemcpyDeviceToHost));
  if (ci && b > a) HIP_TRY(hipMemcpy(ci, v.cycles + a, (b - a) * 2, hipMemcpyDeviceToHost));
  if (in && b > a) HIP_TRY(hipMemcpy(in, v.intens + a, (b - a) * 4, hipMemcpyDeviceToHost));
  return ADSR_OK;
}
}  // namespace adsr

// ===========================================================================
// PeakGroupScoring.score: src/peak_group_scoring/mod.rs:43-341
// ===========================================================================
// One wave per candidate, lane f owning selected fragment f (the m/z-sorted
// output of get_precursor_filtered with the scoring top-k, from the same
// plan pass as selection).  Per-fragment work (extraction, normalisation,
// correlation, m/z) runs per lane; reductions over fragments run as loops
// that every lane executes alike over LDS copies, in fragment order, so the
// f32 sums keep the reference's order.  Features are written column-major.

constexpr int kScoreLMax = 32;  // cycles per candidate (2 * peak_length + 1)
constexpr int kFragB = 98, kFragY = 121;  // src/constants.rs FragmentType

struct ScoreArgs {
  KDia dia;
  KLib lib;
  KPlan plan;
  const uint8_t *p_naa;
  const float *f_mzl;
  const uint32_t *idf_counts;
  float idf_total;
  float mass_tol;
  uint32_t n;
  const uint64_t *c_pidx, *c_start, *c_stop, *c_center;
  const uint32_t *c_row;  // library row of each candidate, ~0u: not in the library
  const float *c_score;
  float *feat;  // [ADSR_N_FEATURES][n]
  uint8_t *found;
  uint32_t *status;
  uint32_t lds_per_wave;
  // quantification outputs: per candidate and fragment slot [n][ks]
  float *q_int, *q_mzo, *q_corr, *q_err;
  uint32_t *q_raw;           // library fragment index of the slot
  uint32_t *q_nf, *q_row;    // per candidate: fragments, library row
  float *q_rt;               // per candidate: RT at cycle_center
};

constexpr uint32_t kStScoreWindowTooLong = 1u << 8;
constexpr uint32_t kStScoreRtIndex = 1u << 9;  // the reference panics there
constexpr uint32_t kStScoreTooManyWindows = 1u << 10;
constexpr uint32_t kScoreObsMax = 16;  // isolation windows holding one precursor
// in-window entries staged per pass of 64 bins; 136 so that the staging
// area also holds the column sorts of 24 fragments x 11 cycles
constexpr uint32_t kScorePool = 136;

// Rust 1.88 slice::binary_search over u16: the last equal element or the
// insertion point (src/quadrupole_observation/mod.rs:175-178).
__device__ __forceinline__ uint32_t rust_bsearch_u16(const uint16_t *v, uint32_t n,
                                                     uint16_t key) {
  if (n == 0) return 0;
  uint32_t size = n, base = 0;
  while (size > 1) {
    const uint32_t half = size >> 1, mid = base + half;
    if (!(v[mid] > key)) base = mid;
    size -= half;
  }
  const uint16_t b = v[base];
  if (b == key) return base;
  return base + (b < key ? 1u : 0u);
}

// ndarray 0.16 sum() of a contiguous lane: eight partial sums, combined
// pairwise, then the tail in order (sum_axis(Axis(1)), mod.rs:154)
__device__ __forceinline__ float ndarray_sum(const float *x, uint32_t n) {
  float p[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  uint32_t i = 0;
  for (; i + 8 <= n; i += 8)
    for (int k = 0; k < 8; ++k) p[k] = p[k] + x[i + k];
  float acc = 0.0f;
  acc = acc + (p[0] + p[4]);
  acc = acc + (p[1] + p[5]);
  acc = acc + (p[2] + p[6]);
  acc = acc + (p[3] + p[7]);
  for (; i < n; ++i) acc = acc + x[i];
  return acc;
}

// src/peak_group_scoring/utils.rs:126-172 over x[i * sx], y[i * sy]
__device__ float corr_safe(const float *x, const float *y, uint32_t n) {
  if (n == 0) return 0.0f;
  float xs = -0.0f, ys = -0.0f;  // Iterator::sum folds from -0.0
  for (uint32_t i = 0; i < n; ++i) {
    xs = xs + x[i];
    ys = ys + y[i];
  }
  if (xs == 0.0f || ys == 0.0f) return 0.0f;
  const float xm = xs / (float)n, ym = ys / (float)n;
  float xv = 0.0f, yv = 0.0f, cov = 0.0f;
  for (uint32_t i = 0; i < n; ++i) {
    const float dx = x[i] - xm, dy = y[i] - ym;
    xv = xv + dx * dx;
    yv = yv + dy * dy;
    cov = cov + dx * dy;
  }
  if (xv == 0.0f || yv == 0.0f) return 0.0f;
  const float c = cov / (sqrtf(xv) * sqrtf(yv));
  return fminf(fmaxf(c, -1.0f), 1.0f);
}

// src/peak_group_scoring/utils.rs:293-310
__device__ float gamma_ln_dev(float x, const LogfEntry *ltab) {
  if (x <= 0.0f) return 0.0f;
  if (fabsf(x - 1.0f) < 1e-6f) return 0.0f;
  if (fabsf(x - 2.0f) < 1e-6f) return 0.0f;
  const float ln_2pi = 1.837877f;
  return (x - 0.5f) * logf_exact(x, ltab) - x + 0.5f * ln_2pi;
}

// src/peak_group_scoring/utils.rs:193-279: the four hyperscores of a
// candidate (observed intensity, library intensity, observed intensity x
// inverse mass error weight, IDF) in one pass over the fragments, each with
// its own counts and sums in fragment order; then the 16 logarithms
// (ln_gamma of the counts, ln of the sums), one per lane, added in the
// reference's order.  Every lane returns the same four values.
__device__ void hyperscores_dev(const uint32_t *tn, const uint32_t *matched,
                                const float *oint, const float *fint,
                                const float *wie, const float *idf, uint32_t n,
                                const LogfEntry *ltab, float (&hs)[4]) {
  uint32_t nb[4] = {0u, 0u, 0u, 0u}, ny[4] = {0u, 0u, 0u, 0u};
  float sb[4] = {0.f, 0.f, 0.f, 0.f}, sy[4] = {0.f, 0.f, 0.f, 0.f};
  for (uint32_t i = 0; i < n; ++i) {
    if (!matched[i]) continue;
    const uint32_t t = tn[i] & 0xffu;
    const bool isb = t == (uint32_t)kFragB;
    if (!isb && t != (uint32_t)kFragY) continue;
    const float o = oint[i], fl = fint[i], d = idf[i];
    const float v[4] = {o, fl, o * wie[i], d};
    const bool on[4] = {o != 0.0f, fl != 0.0f, o != 0.0f, d != 0.0f};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (!on[k]) continue;
      if (isb) {
        ++nb[k];
        sb[k] = sb[k] + v[k];
      } else {
        ++ny[k];
        sy[k] = sy[k] + v[k];
      }
    }
  }
  const int lane = lane_id();
  const int k = (lane >> 2) & 3, m = lane & 3;  // term m of hyperscore k
  float term = 0.0f;
  if (lane < 16) {
    const uint32_t c = m == 0 ? (k == 0 ? nb[0] : k == 1 ? nb[1] : k == 2 ? nb[2] : nb[3])
                               : (k == 0 ? ny[0] : k == 1 ? ny[1] : k == 2 ? ny[2] : ny[3]);
    const float sv = m == 2 ? (k == 0 ? sb[0] : k == 1 ? sb[1] : k == 2 ? sb[2] : sb[3])
                            : (k == 0 ? sy[0] : k == 1 ? sy[1] : k == 2 ? sy[2] : sy[3]);
    if (m < 2) term = c > 0u ? gamma_ln_dev((float)c + 1.0f, ltab) : 0.0f;
    else term = sv > 0.0f ? logf_exact(sv, ltab) : 0.0f;
  }
  const int tb = __builtin_bit_cast(int, term);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float fb = __builtin_bit_cast(float, __builtin_amdgcn_readlane(tb, 4 * q));
    const float fy = __builtin_bit_cast(float, __builtin_amdgcn_readlane(tb, 4 * q + 1));
    const float lb = __builtin_bit_cast(float, __builtin_amdgcn_readlane(tb, 4 * q + 2));
    const float ly = __builtin_bit_cast(float, __builtin_amdgcn_readlane(tb, 4 * q + 3));
    const float h = fb + fy + lb + ly;
    hs[q] = (nb[q] == 0u && ny[q] == 0u) ? 0.0f : (isfinite(h) ? h : 0.0f);
  }
}

// Bin of an m/z (src/mz_index/mod.rs:55-90)
__device__ __forceinline__ uint32_t closest_bin(const KDia &dia, float x) {
  if (!(x > dia.grid_first)) return 0;
  if (x > dia.grid_last) return dia.n_grid - 1;
  return closest_from(dia, x,
                      dia.bucket1[(__float_as_uint(x) >> kBucketShift) - dia.base_key]);
}

// src/idf/mod.rs:13-33: fragments per bin over every fragment_mz_library
__global__ __launch_bounds__(256) void idf_count_kernel(KDia dia, const float *mzl,
                                                        uint64_t n, uint32_t *counts) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(&counts[closest_bin(dia, mzl[i])], 1u);
}

// speclib_flat/mod.rs:458-475 get_precursor_by_idx_filtered: binary search
// of the sorted precursor_idx, one thread per candidate
__global__ __launch_bounds__(256) void cand_row_kernel(const uint64_t *p_idx, uint32_t n_prec,
                                                       const uint64_t *c_pidx, uint64_t n,
                                                       uint32_t *row) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t key = c_pidx[i];
  uint32_t lo = 0, hi = n_prec;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (p_idx[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  row[i] = lo < n_prec && p_idx[lo] == key ? lo : 0xffffffffu;
}

struct ScoreLds {
  uint4 *rec;    // [64] bin items of a pass: first entry (u64), pool slot | count << 16, grid m/z
  uint2 *pool;   // [kScorePool] (cycle - c0, intensity) of the pass's in-window entries
  float *xic, *mz, *norm, *sortA, *sortF;  // [ks][L] rows / [L][ks] columns
  float *oint, *corr, *fint, *idf, *err, *wie, *buf;  // [ks]
  uint32_t *tn, *matched, *nz;                         // [ks]
  float *med, *medf;                                   // [kScoreLMax]
  uint32_t *ipre, *ibs, *inb;                          // [ks] bin items per fragment
  uint32_t *vo;                                        // [kScoreObsMax] valid windows
};

// The column sorts (2 ks L floats) reuse the bin-item records and the pool
// when they fit there: those are dead after the extraction.
constexpr uint32_t kScoreStageBytes = kWave * 16 + kScorePool * 8;

__host__ __device__ inline bool score_sorts_in_stage(uint32_t ks, uint32_t L) {
  return 2 * ks * L * 4 <= kScoreStageBytes;
}

__host__ __device__ inline uint32_t score_lds_bytes(uint32_t ks, uint32_t L) {
  return kScoreStageBytes + (score_sorts_in_stage(ks, L) ? 0u : 2 * ks * L * 4) +
         (3 * ks * L + 10 * ks + 2 * kScoreLMax + 3 * ks + kScoreObsMax) * 4;
}

// QUANT: PeakGroupQuantification.quantify (src/peak_group_quantification/
// mod.rs:86-168) -- the same extraction, filtered median profile,
// correlations, summed intensities and m/z errors, written per fragment.
//
// FPL: fragments per lane.  Lane l owns fragments l + 64 u (u < FPL), so a
// candidate holds up to 64 FPL selected fragments; every sum over fragments
// still runs in fragment order over the LDS columns.
template <bool QUANT, int FPL>
__global__ __launch_bounds__(256) void score_kernel(ScoreArgs a) {
  extern __shared__ __align__(16) char smem[];
  __shared__ LogfEntry ltab[16];
  if (threadIdx.x < 16) ltab[threadIdx.x] = kDevLogfTab[threadIdx.x];
  __syncthreads();
  const int lane = lane_id();
  const uint32_t wv = threadIdx.x / kWave;
  const uint32_t ci = uni(blockIdx.x * (blockDim.x / kWave) + wv);
  if (ci >= a.n) return;  // no block barrier follows
  const uint32_t ks = a.plan.ks;
  const uint64_t n_o = a.n;
  // feature k is gathered into lane k and the 41 are stored at once
  float fv = 0.0f;
  auto F_OUT = [&](int k, float v) {
    if (lane == k) fv = v;
  };
  // the library row, looked up by cand_row_kernel
  const uint32_t row = uni(a.c_row[ci]);
  if (row == 0xffffffffu) {  // filter_map -> None
    if (lane == 0) a.found[ci] = 0;
    return;
  }
  const uint64_t c0 = a.c_start[ci], c1 = a.c_stop[ci];
  const uint64_t L64 = c1 - c0;  // usize arithmetic wraps (release build)
  if constexpr (QUANT) {  // quantification/mod.rs:93-98 -> None
    const uint64_t cc = a.c_center[ci];
    if (L64 == 0 || cc < c0 || cc >= c1) {
      if (lane == 0) a.found[ci] = 0;
      return;
    }
  }
  if (L64 > (uint64_t)kScoreLMax) {
    if (lane == 0) {
      a.found[ci] = 0;
      atomicOr(a.status, kStScoreWindowTooLong);
    }
    return;
  }
  const uint32_t L = (uint32_t)L64;
  const uint32_t F = uni(a.plan.rec[row].z & 0xffffu);
  // LDS of this wave
  char *base = smem + (size_t)wv * a.lds_per_wave;
  ScoreLds s;
  {
    s.rec = reinterpret_cast<uint4 *>(base);
    s.pool = reinterpret_cast<uint2 *>(base + kWave * 16);
    float *p = reinterpret_cast<float *>(base + kScoreStageBytes);
    s.xic = p; p += ks * L;
    s.mz = p; p += ks * L;
    s.norm = p; p += ks * L;
    if (score_sorts_in_stage(ks, L)) {
      s.sortA = reinterpret_cast<float *>(base);
    } else {
      s.sortA = p;
      p += 2 * ks * L;
    }
    s.sortF = s.sortA + ks * L;
    s.oint = p; p += ks;
    s.corr = p; p += ks;
    s.fint = p; p += ks;
    s.idf = p; p += ks;
    s.err = p; p += ks;
    s.wie = p; p += ks;
    s.buf = p; p += ks;
    s.tn = reinterpret_cast<uint32_t *>(p); p += ks;
    s.matched = reinterpret_cast<uint32_t *>(p); p += ks;
    s.nz = reinterpret_cast<uint32_t *>(p); p += ks;
    s.med = p; p += kScoreLMax;
    s.medf = p; p += kScoreLMax;
    s.ipre = reinterpret_cast<uint32_t *>(p); p += ks;
    s.ibs = reinterpret_cast<uint32_t *>(p); p += ks;
    s.inb = reinterpret_cast<uint32_t *>(p); p += ks;
    s.vo = reinterpret_cast<uint32_t *>(p);
  }
  // fragment u of this lane (u < FPL) and whether the candidate has it
  auto fr = [&](int u) { return (uint32_t)lane + (uint32_t)(kWave * u); };
  auto has = [&](int u) { return fr(u) < F; };
  float fmz[FPL], fint[FPL];
  uint32_t ftype[FPL], fnum[FPL], df[FPL];
  // mod.rs:249-268 and quantification: RT at cycle_center, loaded early
  const uint64_t ccen = a.c_center[ci];
  const float rt_cen = ccen < a.dia.n_rt ? a.dia.rt[ccen] : 0.0f;
  const float prt = a.lib.p_rt[row];
#pragma unroll
  for (int u = 0; u < FPL; ++u) {
    fmz[u] = 0.0f;
    fint[u] = 0.0f;
    ftype[u] = 0;
    fnum[u] = 0;
    df[u] = 0;
    if (has(u)) {
      const uint32_t f = fr(u);
      const size_t so = (size_t)row * ks + f;
      const uint64_t j = a.lib.f_start[row] + a.plan.sel[so];
      fmz[u] = a.plan.sel_mz[so];
      fint[u] = a.plan.sel_w[so];
      ftype[u] = a.lib.f_type[j];
      fnum[u] = a.lib.f_num[j];
      const float fmzl = a.f_mzl[j];
      // src/idf/mod.rs:35-49, loaded here so that its latency overlaps the
      // extraction
      df[u] = a.idf_counts ? a.idf_counts[closest_bin(a.dia, fmzl)] : 0u;
      s.fint[f] = fint[u];
      s.tn[f] = ftype[u] | (fnum[u] << 8);
      for (uint32_t c = 0; c < L; ++c) {
        s.xic[f * L + c] = 0.0f;
        s.mz[f * L + c] = 0.0f;
      }
    }
  }
  // dense_xic_observation/mod.rs:98-137 with quadrupole_observation/mod.rs:
  // 152-211, valid windows in ascending order.  The (fragment, window, bin)
  // items are searched one per lane, 64 at a time (binary search for c0,
  // then the entries below c1, whose cycles and intensities are staged in
  // LDS); each fragment lane then adds its items' entries in the
  // reference's order: window, bin, entry.
  const float pmz = a.lib.p_mz[row];
  uint32_t nvo = 0;
  for (uint32_t o0 = 0; o0 < a.dia.n_obs; o0 += kWave) {
    const uint32_t o = o0 + (uint32_t)lane;
    const bool in = o < a.dia.n_obs && a.dia.iso_lo[o] <= pmz && a.dia.iso_hi[o] >= pmz;
    const uint64_t m = ballot64(in);
    const uint32_t pos = nvo + __builtin_amdgcn_mbcnt_hi(
                                   (uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    if (in && pos < kScoreObsMax) s.vo[pos] = o;  // windows in ascending order
    nvo += (uint32_t)__builtin_popcountll(m);
  }
  if (nvo > kScoreObsMax) {
    if (lane == 0) {
      a.found[ci] = 0;
      atomicOr(a.status, kStScoreTooManyWindows);
    }
    return;
  }
  // bin items of each fragment, numbered in fragment order
  uint32_t items_f[FPL], ipre_f[FPL];
  uint32_t NI = 0;
#pragma unroll
  for (int u = 0; u < FPL; ++u) {
    uint32_t nb_f = 0;
    if (has(u)) {
      const uint32_t f = fr(u);
      const float delta = fmz[u] * a.mass_tol * 1e-6f;
      uint32_t bs, be;
      mz_range(a.dia, fmz[u] - delta, fmz[u] + delta, bs, be);
      nb_f = be - bs;
      s.ibs[f] = bs;
      s.inb[f] = nb_f;
    }
    items_f[u] = nb_f * nvo;
    const uint32_t incl = wave_incl_scan(items_f[u]);
    ipre_f[u] = NI + incl - items_f[u];
    NI += rdlane(incl, kWave - 1);
    if (has(u)) s.ipre[fr(u)] = ipre_f[u];
  }
  __builtin_amdgcn_wave_barrier();
  const uint16_t key16 = (uint16_t)c0;
  for (uint32_t p0 = 0; p0 < NI; p0 += kWave) {
    const uint32_t k = p0 + (uint32_t)lane;
    uint64_t e = 0;
    uint32_t cnt = 0;
    float actual = 0.0f;
    if (k < NI) {
      // the item's fragment: the last one whose first item is <= k
      uint32_t g = 0, gh = F;
      while (gh - g > 1) {
        const uint32_t mid = (g + gh) >> 1;
        if (s.ipre[mid] <= k) g = mid;
        else gh = mid;
      }
      const uint32_t r = k - s.ipre[g], nb = s.inb[g];
      const uint32_t j = r / nb, b = s.ibs[g] + (r - j * nb);
      const uint32_t o = s.vo[j];
      const uint32_t *ssw = a.dia.ss + (size_t)o * (a.dia.n_grid + 1);
      const uint32_t e0 = ssw[b], n = ssw[b + 1] - e0;
      e = a.dia.entry_base[o] + e0;
      actual = a.dia.grid[b];
      const uint16_t *cy = a.dia.cycles + e;
      uint32_t t = rust_bsearch_u16(cy, n, key16);
      e += t;
      // a linear scan: most bins have no entry in the window, so it stops
      // at its first load (a bound search was measured slower)
      for (; t < n; ++t) {
        const uint64_t cyc = cy[t];
        if (cyc >= c1) break;
        if (cyc - c0 >= (uint64_t)L) break;  // only for a cycle_start beyond 65535
        ++cnt;
      }
    }
    // pool slots in item order; entries past the pool are read again below
    const uint32_t cincl = wave_incl_scan(cnt);
    const uint32_t off = cincl - cnt;
    if (k < NI) {
      s.rec[lane] = make_uint4((uint32_t)e, (uint32_t)(e >> 32), off | (cnt << 16),
                               __float_as_uint(actual));
      const uint32_t m = off < kScorePool ? min(cnt, kScorePool - off) : 0u;
      for (uint32_t i0 = 0; i0 < m; i0 += 4) {
        uint32_t cyc4[4];
        float v4[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          cyc4[u] = i0 + u < m ? a.dia.cycles[e + i0 + u] : 0u;
          v4[u] = i0 + u < m ? a.dia.intens[e + i0 + u] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (i0 + u < m)
            s.pool[off + i0 + u] =
                make_uint2((uint32_t)((uint64_t)cyc4[u] - c0), __float_as_uint(v4[u]));
      }
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int u = 0; u < FPL; ++u) {
    if (has(u) && items_f[u]) {
      const uint32_t f = fr(u);
      const uint32_t k0 = max(ipre_f[u], p0), k1 = min(ipre_f[u] + items_f[u], p0 + kWave);
      float *X = s.xic + f * L, *M = s.mz + f * L;
      for (uint32_t kk = k0; kk < k1; ++kk) {
        const uint4 rc = s.rec[kk - p0];
        const uint32_t c = rc.z >> 16;
        if (c == 0) continue;
        const uint32_t o0 = rc.z & 0xffffu;
        const uint64_t eg = ((uint64_t)rc.y << 32) | rc.x;
        const float act = __uint_as_float(rc.w);
        for (uint32_t i = 0; i < c; ++i) {
          uint32_t rel;
          float v;
          if (o0 + i < kScorePool) {
            const uint2 h = s.pool[o0 + i];
            rel = h.x;
            v = __uint_as_float(h.y);
          } else {
            rel = (uint32_t)((uint64_t)a.dia.cycles[eg + i] - c0);
            v = a.dia.intens[eg + i];
          }
          const float x = X[rel] + v;
          X[rel] = x;
          if (v > 0.0f) {
            const float prev = x - v;
            M[rel] = prev == 0.0f ? act : (M[rel] * prev + act * v) / x;
          }
        }
      }
    }
    }
    __builtin_amdgcn_wave_barrier();
  }
  // utils.rs:71-108 normalize_profiles(.., 1) and :7-29 filter_non_zero
  const uint32_t cidx = L / 2;
#pragma unroll
  for (int u = 0; u < FPL; ++u) {
    if (!has(u)) continue;
    const uint32_t f = fr(u);
    const uint32_t s0 = cidx >= 1 ? cidx - 1 : 0, s1 = min(cidx + 2, L);
    float sum = 0.0f;
    for (uint32_t j = s0; j < s1; ++j) sum = sum + s.xic[f * L + j];
    const float mean = s1 > s0 ? sum / (float)(s1 - s0) : 0.0f;
    uint32_t any = 0;
    for (uint32_t j = 0; j < L; ++j) {
      const float v = mean > 0.0f ? s.xic[f * L + j] / mean : 0.0f;
      s.norm[f * L + j] = v;
      any |= v != 0.0f ? 1u : 0u;
    }
    s.nz[f] = any;
  }
  __builtin_amdgcn_wave_barrier();
  // bit g: fragment g has a non-zero profile (lane g holds fragment g);
  // with more than 64 fragments the flags are read from s.nz
  const uint64_t nzm = ballot64(has(0) && s.nz[fr(0)] != 0u);
  uint32_t Ff = (uint32_t)__builtin_popcountll(nzm);
#pragma unroll
  for (int u = 1; u < FPL; ++u)
    Ff += (uint32_t)__builtin_popcountll(ballot64(has(u) && s.nz[fr(u)] != 0u));
  auto nz_of = [&](uint32_t g) {
    if constexpr (FPL == 1) return ((nzm >> g) & 1ull) != 0ull;
    else return s.nz[g] != 0u;
  };
  // utils.rs:35-67 median_axis_0 of all rows and of the non-zero rows:
  // every (column, fragment) pair, one per lane, places its value at its
  // stable rank in the column
  for (uint32_t q0 = 0; q0 < L * F; q0 += kWave) {
    const uint32_t q = q0 + (uint32_t)lane;
    if (q < L * F) {
      const uint32_t c = q / F, fq = q - c * F;
      const float v = s.norm[fq * L + c];
      uint32_t ra = 0, rf = 0;
      for (uint32_t g = 0; g < F; ++g) {
        const float u = s.norm[g * L + c];
        const bool before = u < v || (u == v && g < fq);
        ra += before ? 1u : 0u;
        rf += (before && nz_of(g)) ? 1u : 0u;
      }
      s.sortA[c * ks + ra] = v;
      if (nz_of(fq)) s.sortF[c * ks + rf] = v;
    }
  }
  __builtin_amdgcn_wave_barrier();
  if ((uint32_t)lane < L) {
    const uint32_t c = (uint32_t)lane;
    const float *sa = s.sortA + c * ks, *sf = s.sortF + c * ks;
    s.med[c] = F == 0 ? 0.0f
               : (F % 2 == 0 ? (sa[F / 2 - 1] + sa[F / 2]) / 2.0f : sa[F / 2]);
    s.medf[c] = Ff == 0 ? 0.0f
                : (Ff % 2 == 0 ? (sf[Ff / 2 - 1] + sf[Ff / 2]) / 2.0f : sf[Ff / 2]);
  }
  __builtin_amdgcn_wave_barrier();
  // mod.rs:142-149: median sums (Iterator::sum from -0.0) and FWHM
  float med_sum = -0.0f, medf_sum = -0.0f;
  for (uint32_t c = 0; c < L; ++c) {
    med_sum = med_sum + s.med[c];
    medf_sum = medf_sum + s.medf[c];
  }
  float fwhm = 0.0f;
  if (L > 0) {  // utils.rs:453-476, usize index arithmetic wraps
    const uint32_t h = L / 2;
    const float cen = s.medf[h];
    for (uint32_t i = 0; i < h; ++i) {
      const float m = (s.medf[h - i] + s.medf[h + i]) / 2.0f;
      if (m <= cen / 2.0f) {
        const uint64_t il = c0 + h - i, ir = c0 + h + i;
        if (il >= a.dia.n_rt || ir >= a.dia.n_rt) {
          if (lane == 0) atomicOr(a.status, kStScoreRtIndex);
        } else {
          fwhm = a.dia.rt[ir] - a.dia.rt[il];
        }
        break;
      }
    }
  }
  // mod.rs:152-160: correlations with the filtered median, summed rows
  float corr[FPL], oint[FPL];
#pragma unroll
  for (int u = 0; u < FPL; ++u) {
    corr[u] = 0.0f;
    oint[u] = 0.0f;
    if (!has(u)) continue;
    const uint32_t f = fr(u);
    corr[u] = corr_safe(s.medf, s.norm + f * L, L);
    oint[u] = ndarray_sum(s.xic + f * L, L);
    s.corr[f] = corr[u];
    s.oint[f] = oint[u];
    s.matched[f] = oint[u] > 0.0f ? 1u : 0u;
  }
  __builtin_amdgcn_wave_barrier();
  const float int_corr = corr_safe(s.oint, s.fint, F);
  // mod.rs:164-197
  float csum = -0.0f;
  for (uint32_t g = 0; g < F; ++g) csum = csum + s.corr[g];
  const float mean_corr = F ? csum / (float)F : 0.0f;
  // utils/mod.rs:120-137 median and :146-155 sample std of the correlations
#pragma unroll
  for (int u = 0; u < FPL; ++u) {
    if (!has(u)) continue;
    const uint32_t f = fr(u);
    uint32_t r = 0;
    for (uint32_t g = 0; g < F; ++g) {
      const float cg = s.corr[g];
      r += (cg < corr[u] || (cg == corr[u] && g < f)) ? 1u : 0u;
    }
    s.buf[r] = corr[u];
  }
  __builtin_amdgcn_wave_barrier();
  const float med_corr = F == 0 ? 0.0f
                         : (F % 2 == 0 ? (s.buf[F / 2 - 1] + s.buf[F / 2]) / 2.0f
                                       : s.buf[F / 2]);
  float std_corr = 0.0f;
  if (F >= 2) {
    const float m = csum / (float)F;
    float v = -0.0f;
    for (uint32_t g = 0; g < F; ++g) {
      const float d = s.corr[g] - m;
      v = v + d * d;
    }
    std_corr = sqrtf(v / (float)(F - 1));
  }
  // fragments over a correlation threshold among those sel(u) picks
  auto count = [&](float thr, auto sel) {
    uint32_t c = 0;
#pragma unroll
    for (int u = 0; u < FPL; ++u)
      c += (uint32_t)__builtin_popcountll(ballot64(has(u) && sel(u) && corr[u] > thr));
    return (float)c;
  };
  // utils/mod.rs:175-207 ranks by library intensity (descending, then
  // fragment order)
  uint32_t rint[FPL];
#pragma unroll
  for (int u = 0; u < FPL; ++u) {
    rint[u] = 0;
    if (!has(u)) continue;
    const uint32_t f = fr(u);
    for (uint32_t g = 0; g < F; ++g) {
      const float ig = s.fint[g];
      rint[u] += (ig > fint[u] || (ig == fint[u] && g < f)) ? 1u : 0u;
    }
  }
  // mod.rs:203-247
  bool matched[FPL];
#pragma unroll
  for (int u = 0; u < FPL; ++u) matched[u] = has(u) && oint[u] > 0.0f;
  // utils.rs:314-371 longest b / y series: matched numbers of a type at
  // their sorted rank, then one scan
  float longest[2];
  for (int k = 0; k < 2; ++k) {
    const uint32_t want = k == 0 ? (uint32_t)kFragB : (uint32_t)kFragY;
    uint32_t cnt = 0;
#pragma unroll
    for (int u = 0; u < FPL; ++u) {
      const bool in = matched[u] && ftype[u] == want;
      cnt += (uint32_t)__builtin_popcountll(ballot64(in));
      if (in) {
        const uint32_t f = fr(u);
        uint32_t r = 0;
        for (uint32_t g = 0; g < F; ++g) {
          const uint32_t tg = s.tn[g];
          if (!s.matched[g] || (tg & 0xffu) != want) continue;
          const uint32_t ng = tg >> 8;
          r += (ng < fnum[u] || (ng == fnum[u] && g < f)) ? 1u : 0u;
        }
        s.buf[r] = (float)fnum[u];
      }
    }
    __builtin_amdgcn_wave_barrier();
    uint32_t best = cnt ? 1u : 0u;
    if constexpr (FPL == 1) {
      // bit i: sorted number i continues number i - 1; the longest run of
      // such bits plus one is the longest series
      const uint32_t il = (uint32_t)lane;
      uint64_t run = ballot64(il >= 1u && il < cnt &&
                              (uint32_t)s.buf[il] == (uint32_t)s.buf[il - 1] + 1u);
      while (run) {
        run &= run >> 1;
        ++best;
      }
    } else {
      // more numbers than lanes: one scan over the sorted numbers
      uint32_t cur = 1;
      for (uint32_t i = 1; i < cnt; ++i) {
        if ((uint32_t)s.buf[i] == (uint32_t)s.buf[i - 1] + 1u) {
          ++cur;
          best = max(best, cur);
        } else {
          cur = 1;
        }
      }
    }
    longest[k] = (float)(best & 0xffu);
    __builtin_amdgcn_wave_barrier();
  }
  // utils/mod.rs:11-69 observed m/z and ppm error; :83-109 weighted mean
  // absolute error; utils.rs:374-399 inverse-mass-error weights
#pragma unroll
  for (int u = 0; u < FPL; ++u) {
    if (!has(u)) continue;
    const uint32_t f = fr(u);
    float ws = 0.0f, wt = 0.0f;
    for (uint32_t j = 0; j < L; ++j) {
      const float w = s.xic[f * L + j];
      if (w > 0.0f) {
        ws = ws + s.mz[f * L + j] * w;
        wt = wt + w;
      }
    }
    const float m = wt > 0.0f ? ws / wt : 0.0f;
    const float e = m > 0.0f ? ((m - fmz[u]) / fmz[u]) * 1e6f : 0.0f;
    s.err[f] = e;
    if constexpr (QUANT) {
      const size_t q = (size_t)ci * ks + f;
      a.q_int[q] = oint[u];
      a.q_mzo[q] = m;
      // quantification/mod.rs:121-128 correlates (profile, median)
      a.q_corr[q] = corr_safe(s.norm + f * L, s.medf, L);
      a.q_err[q] = e;
      a.q_raw[q] = (uint32_t)(a.lib.f_start[row] + a.plan.sel[(size_t)row * ks + f]);
    }
    s.wie[f] = 1.0f / (fabsf(e) + 0.1f);
    // src/idf/mod.rs:35-49
    s.idf[f] = a.idf_total == 0.0f ? 1.0f
               : logf_exact(a.idf_total / (float)(df[u] > 1u ? df[u] : 1u), ltab);
  }
  __builtin_amdgcn_wave_barrier();
  if constexpr (QUANT) {
    if (lane == 0) {
      a.q_nf[ci] = F;
      a.q_row[ci] = row;
      if (ccen >= a.dia.n_rt) atomicOr(a.status, kStScoreRtIndex);
      a.q_rt[ci] = rt_cen;
      a.found[ci] = 1;
    }
    return;
  }
  float wmae = 0.0f;
  {
    float ws = 0.0f, wt = 0.0f;
    for (uint32_t g = 0; g < F; ++g) {
      const float w = s.fint[g], e = s.err[g];
      if (w > 0.0f && e != 0.0f) {
        ws = ws + fabsf(e) * w;
        wt = wt + w;
      }
    }
    wmae = (F > 0 && wt > 0.0f) ? ws / wt : 0.0f;
  }
  float hs4[4];
  hyperscores_dev(s.tn, s.matched, s.oint, s.fint, s.wie, s.idf, F, ltab, hs4);
  const float hs_obs = hs4[0], hs_lib = hs4[1], hs_ime = hs4[2], idf_hs = hs4[3];
  // mod.rs:249-268
  if (ccen >= a.dia.n_rt && lane == 0) atomicOr(a.status, kStScoreRtIndex);
  const float rt_obs = rt_cen;
  const float delta_rt = rt_obs - prt;
  float ib = 0.0f, iy = 0.0f;
  for (uint32_t g = 0; g < F; ++g) {
    if (!s.matched[g] || !(s.oint[g] > 0.0f)) continue;
    const uint32_t t = s.tn[g] & 0xffu;
    if (t == (uint32_t)kFragB) ib = ib + s.oint[g];
    if (t == (uint32_t)kFragY) iy = iy + s.oint[g];
  }
  // mod.rs:272-293
  float dot_xic = -0.0f, dot_int = -0.0f;
  for (uint32_t g = 0; g < F; ++g) {
    dot_xic = dot_xic + s.idf[g] * s.corr[g];
    dot_int = dot_int + s.idf[g] * s.oint[g];
  }
  if (F == 0) dot_xic = dot_int = 0.0f;
  uint32_t ridf[FPL];
#pragma unroll
  for (int u = 0; u < FPL; ++u) {
    ridf[u] = 0;
    if (!has(u)) continue;
    const uint32_t f = fr(u);
    const float v = s.idf[f];
    for (uint32_t g = 0; g < F; ++g) {
      const float ig = s.idf[g];
      ridf[u] += (ig > v || (ig == v && g < f)) ? 1u : 0u;
    }
  }
  auto all = [&](int) { return true; };
  auto idf6 = [&](int u) { return ridf[u] < 6; };
  auto rk = [&](uint32_t lo, uint32_t hi) {
    return [&, lo, hi](int u) { return rint[u] >= lo && rint[u] < hi; };
  };
  const float cnt_o0_i6 = count(0.0f, idf6), cnt_o50_i6 = count(0.5f, idf6);
  const float c95 = count(0.95f, all), c90 = count(0.90f, all), c80 = count(0.80f, all);
  const float c50 = count(0.5f, all), c0v = count(0.0f, all);
  const float r0 = count(0.0f, rk(0, 6)), r1 = count(0.0f, rk(6, 12));
  const float r2 = count(0.0f, rk(12, 18)), r3 = count(0.0f, rk(18, 24));
  const float q0 = count(0.5f, rk(0, 6)), q1 = count(0.5f, rk(6, 12));
  const float q2 = count(0.5f, rk(12, 18)), q3 = count(0.5f, rk(18, 24));
  if (lane == 0) a.found[ci] = 1;
  F_OUT(ADSR_F_SCORE, a.c_score[ci]);
  F_OUT(ADSR_F_MEAN_CORRELATION, mean_corr);
  F_OUT(ADSR_F_MEDIAN_CORRELATION, med_corr);
  F_OUT(ADSR_F_CORRELATION_STD, std_corr);
  F_OUT(ADSR_F_INTENSITY_CORRELATION, int_corr);
  F_OUT(ADSR_F_NUM_FRAGMENTS, (float)F);
  F_OUT(ADSR_F_NUM_SCANS, (float)L64);
  F_OUT(ADSR_F_NUM_OVER_95, c95);
  F_OUT(ADSR_F_NUM_OVER_90, c90);
  F_OUT(ADSR_F_NUM_OVER_80, c80);
  F_OUT(ADSR_F_NUM_OVER_50, c50);
  F_OUT(ADSR_F_NUM_OVER_0, c0v);
  F_OUT(ADSR_F_NUM_OVER_0_RANK_0_5, r0);
  F_OUT(ADSR_F_NUM_OVER_0_RANK_6_11, r1);
  F_OUT(ADSR_F_NUM_OVER_0_RANK_12_17, r2);
  F_OUT(ADSR_F_NUM_OVER_0_RANK_18_23, r3);
  F_OUT(ADSR_F_NUM_OVER_50_RANK_0_5, q0);
  F_OUT(ADSR_F_NUM_OVER_50_RANK_6_11, q1);
  F_OUT(ADSR_F_NUM_OVER_50_RANK_12_17, q2);
  F_OUT(ADSR_F_NUM_OVER_50_RANK_18_23, q3);
  F_OUT(ADSR_F_HYPERSCORE_INTENSITY_OBSERVATION, hs_obs);
  F_OUT(ADSR_F_HYPERSCORE_INTENSITY_LIBRARY, hs_lib);
  F_OUT(ADSR_F_HYPERSCORE_INVERSE_MASS_ERROR, hs_ime);
  F_OUT(ADSR_F_RT_OBSERVED, rt_obs);
  F_OUT(ADSR_F_DELTA_RT, delta_rt);
  F_OUT(ADSR_F_LONGEST_B_SERIES, longest[0]);
  F_OUT(ADSR_F_LONGEST_Y_SERIES, longest[1]);
  F_OUT(ADSR_F_NAA, (float)a.p_naa[row]);
  F_OUT(ADSR_F_WEIGHTED_MASS_ERROR, wmae);
  // the three log10 features, one evaluation in their own lanes
  const float lg = log10f(lane == ADSR_F_LOG10_B_ION_INTENSITY   ? ib + 1e-8f
                          : lane == ADSR_F_LOG10_Y_ION_INTENSITY ? iy + 1e-8f
                                                                 : dot_int + 1e-8f);
  F_OUT(ADSR_F_LOG10_B_ION_INTENSITY, lg);
  F_OUT(ADSR_F_LOG10_Y_ION_INTENSITY, lg);
  F_OUT(ADSR_F_FWHM_RT, fwhm);
  F_OUT(ADSR_F_IDF_HYPERSCORE, idf_hs);
  F_OUT(ADSR_F_IDF_XIC_DOT_PRODUCT, dot_xic);
  F_OUT(ADSR_F_IDF_INTENSITY_DOT_PRODUCT, lg);
  F_OUT(ADSR_F_MEDIAN_PROFILE_SUM, med_sum);
  F_OUT(ADSR_F_MEDIAN_PROFILE_SUM_FILTERED, medf_sum);
  F_OUT(ADSR_F_NUM_PROFILES, (float)F);
  F_OUT(ADSR_F_NUM_PROFILES_FILTERED, (float)Ff);
  F_OUT(ADSR_F_NUM_OVER_0_TOP6_IDF, cnt_o0_i6);
  F_OUT(ADSR_F_NUM_OVER_50_TOP6_IDF, cnt_o50_i6);
  if (lane < ADSR_N_FEATURES) a.feat[(size_t)lane * n_o + ci] = fv;
}

static const char *const kFeatureNames[ADSR_N_FEATURES] = {
    "score", "mean_correlation", "median_correlation", "correlation_std",
    "intensity_correlation", "num_fragments", "num_scans", "num_over_95",
    "num_over_90", "num_over_80", "num_over_50", "num_over_0",
    "num_over_0_rank_0_5", "num_over_0_rank_6_11", "num_over_0_rank_12_17",
    "num_over_0_rank_18_23", "num_over_50_rank_0_5", "num_over_50_rank_6_11",
    "num_over_50_rank_12_17", "num_over_50_rank_18_23",
    "hyperscore_intensity_observation", "hyperscore_intensity_library",
    "hyperscore_inverse_mass_error", "rt_observed", "delta_rt",
    "longest_b_series", "longest_y_series", "naa", "weighted_mass_error",
    "log10_b_ion_intensity", "log10_y_ion_intensity", "fwhm_rt",
    "idf_hyperscore", "idf_xic_dot_product", "idf_intensity_dot_product",
    "median_profile_sum", "median_profile_sum_filtered", "num_profiles",
    "num_profiles_filtered", "num_over_0_top6_idf", "num_over_50_top6_idf"};

extern "C" const char *adsr_feature_name(size_t i) {
  return i < (size_t)ADSR_N_FEATURES ? kFeatureNames[i] : nullptr;
}

// src/peak_group_scoring/parameters.rs:24-33
extern "C" void adsr_scoring_params_default(adsr_scoring_params *p) {
  if (!p) return;
  p->mass_tolerance = 7.0f;
  p->top_k_fragments = 100;
}

// The library columns only scoring reads, and the IDF bin counts of every
// fragment_mz_library (src/speclib_flat/mod.rs:165-166, src/idf/mod.rs).
static adsr_status upload_lib_scoring(adsr_lib *l, const KDia &kd) {
  std::lock_guard<std::mutex> lk(l->mu);
  DeviceLib &v = l->dev;
  if (v.scoring) return ADSR_OK;
  size_t acc = 0;
  adsr_status st = dmalloc_copy(&v.p_naa, l->precursor_naa.data(), l->n_prec, acc);
  if (!st) st = dmalloc_copy(&v.f_mzl, l->fragment_mz_library.data(), l->n_frag, acc);
  if (st) return st;
  HIP_TRY(dev_malloc((void **)&v.idf_counts, (size_t)kd.n_grid * 4));
  HIP_TRY(hipMemset(v.idf_counts, 0, (size_t)kd.n_grid * 4));
  if (l->n_frag)
    idf_count_kernel<<<(uint32_t)((l->n_frag + 255) / 256), 256>>>(kd, v.f_mzl,
                                                                    l->n_frag,
                                                                    v.idf_counts);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  v.bytes += acc + (size_t)kd.n_grid * 4;
  v.scoring = true;
  return ADSR_OK;
}

// Both entry points: `hc` (host table) or `dc` (table in HBM).
// With qout set, PeakGroupQuantification.quantify with the tolerance and
// top-k of p (src/peak_group_quantification/mod.rs:39-168).
static adsr_status score_common(const adsr_dia *dc_, const adsr_lib *lc,
                                const adsr_candidates *hc,
                                const adsr_device_candidates *dc,
                                const adsr_scoring_params *p,
                                adsr_features **out,
                                adsr_quantified **qout = nullptr) {
  clear_error();
  const bool quant = qout != nullptr;
  if (!dc_ || !lc || !p || (!out && !quant) || (!hc && !dc))
    return set_error(ADSR_ERR_INVALID_ARGUMENT, "NULL argument");
  if (out) *out = nullptr;
  if (qout) *qout = nullptr;
  adsr_dia *d = const_cast<adsr_dia *>(dc_);
  adsr_lib *l = const_cast<adsr_lib *>(lc);
  adsr_status st = upload_dia(d);
  if (!st) st = upload_lib(l);
  if (st) return st;
  const KDia kd = make_kdia(d);
  st = upload_lib_scoring(l, kd);
  if (st) return st;
  const size_t n = hc ? hc->n : dc->n;
  const uint64_t nrt = d->rt_index.size();
  // the longest cycle range of the table bounds the LDS per candidate
  uint64_t lmax = 0;
  if (hc) {
    for (size_t i = 0; i < n; ++i) {
      const uint64_t li = hc->cycle_stop[i] - hc->cycle_start[i];
      // quantification skips these rows before any extraction (mod.rs:93-98)
      if (quant && (li == 0 || hc->cycle_center[i] < hc->cycle_start[i] ||
                    hc->cycle_center[i] >= hc->cycle_stop[i]))
        continue;
      lmax = std::max<uint64_t>(lmax, li);
      if (hc->cycle_center[i] >= nrt)
        return set_error(ADSR_ERR_INVALID_ARGUMENT,
                         "candidate %zu: cycle_center %llu is outside the RT "
                         "axis (%llu cycles)",
                         i, (unsigned long long)hc->cycle_center[i],
                         (unsigned long long)nrt);
    }
  } else {
    lmax = 2 * dc->pl + 1;
  }
  if (lmax > (uint64_t)kScoreLMax)
    return set_error(ADSR_ERR_UNSUPPORTED,
                     "candidate cycle ranges of %llu cycles exceed %d",
                     (unsigned long long)lmax, kScoreLMax);
  // fragment filter and m/z order with the scoring top-k: the selection's
  // plan pass (get_precursor_filtered(i, true, true, top_k))
  adsr_selection_params sp_p;
  adsr_selection_params_default(&sp_p);
  sp_p.mass_tolerance = p->mass_tolerance;
  sp_p.top_k_fragments = p->top_k_fragments;
  sp_p.rt_tolerance = 0.0f;
  sp_p.candidate_count = 1;
  Workspace w;
  SearchPlan sp;
  st = build_search(d, l, &sp_p, w, sp, false, kScoreMaxSel);
  if (st) {
    ws_free(w);
    return st;
  }
  auto fail = [&](adsr_status s2) {
    ws_free(w);
    return s2;
  };
  const uint32_t n_rows = (uint32_t)l->n_prec;
  (void)hipEventRecord(w.ev[0], w.stream);
  st = enqueue(w, sp, n_rows, LaunchShape(), nullptr, 0, 0, 0, false, true);
  if (st) return fail(st);
  (void)hipEventRecord(w.ev[1], w.stream);
  {
    uint32_t host[3] = {0, 0, 0};
    if (hipMemcpyAsync(host, w.small, 12, hipMemcpyDeviceToHost, w.stream) != hipSuccess ||
        hipStreamSynchronize(w.stream) != hipSuccess)
      return fail(set_error(ADSR_ERR_DEVICE, "fragment plan failed"));
    if (host[1] & kStTooManyFragments)
      return fail(set_error(ADSR_ERR_UNSUPPORTED,
                            "a precursor selects more than %d fragments", kScoreMaxSel));
    if (host[1] & kStSelectFallback)
      return fail(set_error(ADSR_ERR_UNSUPPORTED,
                            "a precursor's top-k fragment selection needs the "
                            "median-of-medians fallback of select_nth_unstable"));
  }
  // candidate columns on the device
  BuildBufs tmp;
  uint64_t *cpidx, *crank, *cs, *ce, *cc;
  float *csc, *feat;
  uint8_t *found;
  uint32_t *status, *crow;
  const size_t ks = sp.plan.ks;
  const size_t nq = quant ? n * ks : 0;
  float *q_int = nullptr, *q_mzo = nullptr, *q_corr = nullptr, *q_err = nullptr,
        *q_rt = nullptr;
  uint32_t *q_raw = nullptr, *q_nf = nullptr, *q_row = nullptr;
  if (tmp.alloc(&cpidx, n) || tmp.alloc(&crank, n) || tmp.alloc(&cs, n) ||
      tmp.alloc(&ce, n) || tmp.alloc(&cc, n) || tmp.alloc(&csc, n) ||
      tmp.alloc(&feat, quant ? 1 : (size_t)ADSR_N_FEATURES * n) ||
      tmp.alloc(&found, n) || tmp.alloc(&status, 1) || tmp.alloc(&crow, n))
    return fail(set_error(ADSR_ERR_OUT_OF_MEMORY, "scoring buffers"));
  if (quant && (tmp.alloc(&q_int, nq) || tmp.alloc(&q_mzo, nq) || tmp.alloc(&q_corr, nq) ||
                tmp.alloc(&q_err, nq) || tmp.alloc(&q_raw, nq) || tmp.alloc(&q_nf, n) ||
                tmp.alloc(&q_row, n) || tmp.alloc(&q_rt, n)))
    return fail(set_error(ADSR_ERR_OUT_OF_MEMORY, "quantification buffers"));
  if (hipMemsetAsync(status, 0, 4, w.stream) != hipSuccess)
    return fail(set_error(ADSR_ERR_DEVICE, "status reset failed"));
  if (n && hc) {
    auto up = [&](void *dst, const void *src, size_t bytes) {
      return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, w.stream) == hipSuccess;
    };
    if (!up(cpidx, hc->precursor_idx, n * 8) || !up(crank, hc->rank, n * 8) ||
        !up(cs, hc->cycle_start, n * 8) || !up(ce, hc->cycle_stop, n * 8) ||
        !up(cc, hc->cycle_center, n * 8) || !up(csc, hc->score, n * 4))
      return fail(set_error(ADSR_ERR_DEVICE, "candidate upload failed"));
  } else if (n) {
    export_kernel<<<(uint32_t)((n + 255) / 256), 256, 0, w.stream>>>(
        (const DevCand *)dc->rows, n, dc->pl, dc->nrt, cpidx, crank, csc, nullptr,
        nullptr, nullptr, cc, cs, ce);
    if (hipGetLastError() != hipSuccess)
      return fail(set_error(ADSR_ERR_DEVICE, "candidate export failed"));
  }
  ScoreArgs a;
  a.dia = sp.dia;
  a.lib = sp.lib;
  a.plan = sp.plan;
  a.p_naa = l->dev.p_naa;
  a.f_mzl = l->dev.f_mzl;
  a.idf_counts = l->dev.idf_counts;
  a.idf_total = (float)l->n_frag;
  a.mass_tol = p->mass_tolerance;
  a.n = (uint32_t)n;
  a.c_pidx = cpidx;
  a.c_start = cs;
  a.c_stop = ce;
  a.c_center = cc;
  a.c_row = crow;
  a.c_score = csc;
  a.feat = feat;
  a.found = found;
  a.status = status;
  a.q_int = q_int;
  a.q_mzo = q_mzo;
  a.q_corr = q_corr;
  a.q_err = q_err;
  a.q_raw = q_raw;
  a.q_nf = q_nf;
  a.q_row = q_row;
  a.q_rt = q_rt;
  a.lds_per_wave = (score_lds_bytes(sp.plan.ks, (uint32_t)lmax) + 15) / 16 * 16;
  // fragments per lane: the plan's stride is at most kScoreMaxSel = 4 x 64
  const int fpl = sp.plan.ks > 2u * kWave ? 4 : sp.plan.ks > (uint32_t)kWave ? 2 : 1;
  static_assert(kScoreMaxSel <= 4 * kWave, "four fragments per lane at most");
  int wpb = 4;
  while (wpb > 1 && (size_t)a.lds_per_wave * wpb > 64 * 1024) wpb >>= 1;
  if ((size_t)a.lds_per_wave > kMaxDynLds)
    return fail(set_error(ADSR_ERR_UNSUPPORTED, "scoring LDS of %u bytes", a.lds_per_wave));
  if ((size_t)a.lds_per_wave > 64 * 1024) {
    // one wave per block beyond 64 KB, with the dynamic LDS limit raised
    static std::once_flag once;
    std::call_once(once, [] {
      const void *fns[] = {(const void *)score_kernel<false, 1>, (const void *)score_kernel<true, 1>,
                           (const void *)score_kernel<false, 2>, (const void *)score_kernel<true, 2>,
                           (const void *)score_kernel<false, 4>, (const void *)score_kernel<true, 4>};
      for (const void *fn : fns)
        (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)kMaxDynLds);
    });
  }
  (void)hipEventRecord(w.ev[2], w.stream);
  if (n) {
    cand_row_kernel<<<(uint32_t)((n + 255) / 256), 256, 0, w.stream>>>(
        l->dev.p_idx, (uint32_t)l->n_prec, cpidx, n, crow);
    const dim3 grid((uint32_t)((n + wpb - 1) / wpb)), block(wpb * kWave);
    const size_t lds = (size_t)a.lds_per_wave * wpb;
    if (quant && fpl == 4)
      score_kernel<true, 4><<<grid, block, lds, w.stream>>>(a);
    else if (quant && fpl == 2)
      score_kernel<true, 2><<<grid, block, lds, w.stream>>>(a);
    else if (quant)
      score_kernel<true, 1><<<grid, block, lds, w.stream>>>(a);
    else if (fpl == 4)
      score_kernel<false, 4><<<grid, block, lds, w.stream>>>(a);
    else if (fpl == 2)
      score_kernel<false, 2><<<grid, block, lds, w.stream>>>(a);
    else
      score_kernel<false, 1><<<grid, block, lds, w.stream>>>(a);
    if (hipGetLastError() != hipSuccess)
      return fail(set_error(ADSR_ERR_DEVICE, "score kernel launch failed"));
  }
  (void)hipEventRecord(w.ev[3], w.stream);
  if (quant) {
    std::vector<float> qi(nq), qm(nq), qc(nq), qe(nq), qrt(n);
    std::vector<uint32_t> qraw(nq), qnf(n), qrow(n);
    std::vector<uint8_t> hfound(n);
    std::vector<uint64_t> hp(n), hr(n);
    uint32_t hstatus = 0;
    auto dl = [&](void *dst, const void *src, size_t bytes) {
      return bytes == 0 ||
             hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, w.stream) == hipSuccess;
    };
    bool ok = dl(&hstatus, status, 4) && dl(qi.data(), q_int, nq * 4) &&
              dl(qm.data(), q_mzo, nq * 4) && dl(qc.data(), q_corr, nq * 4) &&
              dl(qe.data(), q_err, nq * 4) && dl(qraw.data(), q_raw, nq * 4) &&
              dl(qnf.data(), q_nf, n * 4) && dl(qrow.data(), q_row, n * 4) &&
              dl(qrt.data(), q_rt, n * 4) && dl(hfound.data(), found, n) &&
              dl(hp.data(), cpidx, n * 8) && dl(hr.data(), crank, n * 8);
    if (!ok || hipStreamSynchronize(w.stream) != hipSuccess)
      return fail(set_error(ADSR_ERR_DEVICE, "quantification failed: %s",
                            hipGetErrorString(hipGetLastError())));
    if (hstatus & kStScoreRtIndex)
      return fail(set_error(ADSR_ERR_INVALID_ARGUMENT,
                            "a candidate's cycle_center is outside the RT axis"));
    if (hstatus & kStScoreWindowTooLong)
      return fail(set_error(ADSR_ERR_UNSUPPORTED, "a candidate cycle range exceeds %d",
                            kScoreLMax));
    if (hstatus & kStScoreTooManyWindows)
      return fail(set_error(ADSR_ERR_UNSUPPORTED,
                            "a precursor m/z lies in more than %u isolation windows",
                            kScoreObsMax));
    // precursor_quantified/mod.rs:26-94 (fragments with intensity > 0, none
    // -> no row) and speclib_flat_quantified/mod.rs:213-330 (flat fragments
    // in candidate order, precursor columns stably sorted by precursor_idx)
    auto *q = new adsr_quantified();
    std::vector<size_t> prow;  // library row of each quantified precursor
    uint64_t pos = 0;
    for (size_t i = 0; i < n; ++i) {
      if (!hfound[i]) continue;
      const size_t base = i * ks;
      size_t kept = 0;
      for (uint32_t f = 0; f < qnf[i]; ++f) kept += qi[base + f] > 0.0f ? 1 : 0;
      if (!kept) continue;
      const size_t r = qrow[i];
      prow.push_back(r);
      q->precursor_idx.push_back(l->precursor_idx[r]);
      q->rank.push_back(hr[i]);
      q->rt_observed.push_back(qrt[i]);
      q->frag_start.push_back(pos);
      pos += kept;
      q->frag_stop.push_back(pos);
      for (uint32_t f = 0; f < qnf[i]; ++f) {
        if (!(qi[base + f] > 0.0f)) continue;
        const size_t j = qraw[base + f];
        q->f_mz_library.push_back(l->fragment_mz_library[j]);
        q->f_mz.push_back(l->fragment_mz[j]);
        q->f_intensity.push_back(qi[base + f]);
        q->f_cardinality.push_back(l->fragment_cardinality[j]);
        q->f_charge.push_back(l->fragment_charge[j]);
        q->f_loss_type.push_back(l->fragment_loss_type[j]);
        q->f_number.push_back(l->fragment_number[j]);
        q->f_position.push_back(l->fragment_position[j]);
        q->f_type.push_back(l->fragment_type[j]);
        q->f_mz_observed.push_back(qm[base + f]);
        q->f_correlation.push_back(qc[base + f]);
        q->f_mass_error.push_back(qe[base + f]);
        q->f_precursor_idx.push_back(l->precursor_idx[r]);
        q->f_rank.push_back(hr[i]);
      }
    }
    const size_t m = prow.size();
    std::vector<size_t> ord(m);
    for (size_t i = 0; i < m; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](size_t x, size_t y) {
      return q->precursor_idx[x] < q->precursor_idx[y];
    });
    auto perm = [&](auto &v) {
      auto c = v;
      for (size_t i = 0; i < m; ++i) v[i] = c[ord[i]];
    };
    perm(q->precursor_idx);
    perm(q->rank);
    perm(q->rt_observed);
    perm(q->frag_start);
    perm(q->frag_stop);
    perm(prow);
    q->mz_library.resize(m);
    q->mz.resize(m);
    q->rt_library.resize(m);
    q->rt.resize(m);
    q->naa.resize(m);
    for (size_t i = 0; i < m; ++i) {
      q->mz_library[i] = l->precursor_mz_library[prow[i]];
      q->mz[i] = l->precursor_mz[prow[i]];
      q->rt_library[i] = l->precursor_rt_library[prow[i]];
      q->rt[i] = l->precursor_rt[prow[i]];
      q->naa[i] = l->precursor_naa[prow[i]];
    }
    float t_plan = 0, t_kern = 0;
    (void)hipEventElapsedTime(&t_plan, w.ev[0], w.ev[1]);
    (void)hipEventElapsedTime(&t_kern, w.ev[2], w.ev[3]);
    q->plan_ms = t_plan;
    q->kernel_ms = t_kern;
    ws_free(w);
    *qout = q;
    return ADSR_OK;
  }
  std::vector<float> hf((size_t)ADSR_N_FEATURES * n);
  std::vector<uint8_t> hfound(n);
  std::vector<uint64_t> hp(n), hr(n);
  uint32_t hstatus = 0;
  bool ok = hipMemcpyAsync(&hstatus, status, 4, hipMemcpyDeviceToHost, w.stream) == hipSuccess;
  if (n) {
    ok = ok && hipMemcpyAsync(hf.data(), feat, hf.size() * 4, hipMemcpyDeviceToHost,
                              w.stream) == hipSuccess;
    ok = ok && hipMemcpyAsync(hfound.data(), found, n, hipMemcpyDeviceToHost,
                              w.stream) == hipSuccess;
    ok = ok && hipMemcpyAsync(hp.data(), cpidx, n * 8, hipMemcpyDeviceToHost,
                              w.stream) == hipSuccess;
    ok = ok && hipMemcpyAsync(hr.data(), crank, n * 8, hipMemcpyDeviceToHost,
                              w.stream) == hipSuccess;
  }
  if (!ok || hipStreamSynchronize(w.stream) != hipSuccess)
    return fail(set_error(ADSR_ERR_DEVICE, "scoring failed: %s",
                          hipGetErrorString(hipGetLastError())));
  if (hstatus & kStScoreRtIndex)
    return fail(set_error(ADSR_ERR_INVALID_ARGUMENT,
                          "a candidate's cycle range indexes outside the RT axis"));
  if (hstatus & kStScoreWindowTooLong)
    return fail(set_error(ADSR_ERR_UNSUPPORTED, "a candidate cycle range exceeds %d",
                          kScoreLMax));
  if (hstatus & kStScoreTooManyWindows)
    return fail(set_error(ADSR_ERR_UNSUPPORTED,
                          "a precursor m/z lies in more than %u isolation windows",
                          kScoreObsMax));
  auto *f = new adsr_features();
  size_t m = 0;
  for (size_t i = 0; i < n; ++i) m += hfound[i] ? 1 : 0;
  f->precursor_idx.reserve(m);
  f->rank.reserve(m);
  f->columns.resize((size_t)ADSR_N_FEATURES * m);
  size_t k = 0;
  for (size_t i = 0; i < n; ++i) {
    if (!hfound[i]) continue;
    f->precursor_idx.push_back(hp[i]);
    f->rank.push_back(hr[i]);
    for (int c = 0; c < ADSR_N_FEATURES; ++c)
      f->columns[(size_t)c * m + k] = hf[(size_t)c * n + i];
    ++k;
  }
  float t_plan = 0, t_kern = 0;
  (void)hipEventElapsedTime(&t_plan, w.ev[0], w.ev[1]);
  (void)hipEventElapsedTime(&t_kern, w.ev[2], w.ev[3]);
  f->plan_ms = t_plan;
  f->kernel_ms = t_kern;
  ws_free(w);
  *out = f;
  return ADSR_OK;
}

extern "C" adsr_status adsr_score(const adsr_dia *d, const adsr_lib *l,
                                  const adsr_candidates *c,
                                  const adsr_scoring_params *p,
                                  adsr_features **out) {
  if (!c) {
    clear_error();
    return set_error(ADSR_ERR_INVALID_ARGUMENT, "NULL candidates");
  }
  return score_common(d, l, c, nullptr, p, out);
}

extern "C" adsr_status adsr_score_device(const adsr_dia *d, const adsr_lib *l,
                                         const adsr_device_candidates *c,
                                         const adsr_scoring_params *p,
                                         adsr_features **out) {
  if (!c) {
    clear_error();
    return set_error(ADSR_ERR_INVALID_ARGUMENT, "NULL candidates");
  }
  return score_common(d, l, nullptr, c, p, out);
}

extern "C" size_t adsr_features_len(const adsr_features *f) {
  return f ? f->precursor_idx.size() : 0;
}

extern "C" adsr_status adsr_features_export(const adsr_features *f,
                                            uint64_t *precursor_idx,
                                            uint64_t *rank, float *columns) {
  clear_error();
  if (!f) return set_error(ADSR_ERR_INVALID_ARGUMENT, "NULL handle");
  const size_t m = f->precursor_idx.size();
  if (precursor_idx && m) std::memcpy(precursor_idx, f->precursor_idx.data(), m * 8);
  if (rank && m) std::memcpy(rank, f->rank.data(), m * 8);
  if (columns && m) std::memcpy(columns, f->columns.data(), f->columns.size() * 4);
  return ADSR_OK;
}

extern "C" adsr_status adsr_features_timings(const adsr_features *f, double *plan_ms,
                                             double *kernel_ms) {
  clear_error();
  if (!f) return set_error(ADSR_ERR_INVALID_ARGUMENT, "NULL handle");
  if (plan_ms) *plan_ms = f->plan_ms;
  if (kernel_ms) *kernel_ms = f->kernel_ms;
  return ADSR_OK;
}

extern "C" void adsr_features_free(adsr_features *f) { delete f; }

// src/peak_group_quantification/parameters.rs:18-29
extern "C" void adsr_quant_params_default(adsr_quant_params *p) {
  if (!p) return;
  p->tolerance_ppm = 7.0f;
  p->top_k_fragments = 10000;
}

extern "C" adsr_status adsr_quantify(const adsr_dia *d, const adsr_lib *l,
                                     const adsr_candidates *c,
                                     const adsr_quant_params *p,
                                     adsr_quantified **out) {
  clear_error();
  if (!c || !p || !out) return set_error(ADSR_ERR_INVALID_ARGUMENT, "NULL argument");
  adsr_scoring_params sp{p->tolerance_ppm, p->top_k_fragments};
  return score_common(d, l, c, nullptr, &sp, nullptr, out);
}

extern "C" size_t adsr_quantified_num_precursors(const adsr_quantified *q) {
  return q ? q->precursor_idx.size() : 0;
}

extern "C" size_t adsr_quantified_num_fragments(const adsr_quantified *q) {
  return q ? q->f_mz.size() : 0;
}

template <class T>
static void copy_out(T *dst, const std::vector<T> &v) {
  if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(T));
}

extern "C" adsr_status adsr_quantified_export_precursors(
    const adsr_quantified *q, uint64_t *precursor_idx, float *mz_library,
    float *mz, float *rt_library, float *rt, uint8_t *naa, uint64_t *rank,
    float *rt_observed, uint64_t *flat_frag_start_idx,
    uint64_t *flat_frag_stop_idx) {
  clear_error();
  if (!q) return set_error(ADSR_ERR_INVALID_ARGUMENT, "NULL handle");
  copy_out(precursor_idx, q->precursor_idx);
  copy_out(mz_library, q->mz_library);
  copy_out(mz, q->mz);
  copy_out(rt_library, q->rt_library);
  copy_out(rt, q->rt);
  copy_out(naa, q->naa);
  copy_out(rank, q->rank);
  copy_out(rt_observed, q->rt_observed);
  copy_out(flat_frag_start_idx, q->frag_start);
  copy_out(flat_frag_stop_idx, q->frag_stop);
  return ADSR_OK;
}

extern "C" adsr_status adsr_quantified_export_fragments(
    const adsr_quantified *q, float *mz_library, float *mz, float *intensity,
    uint8_t *cardinality, uint8_t *charge, uint8_t *loss_type, uint8_t *number,
    uint8_t *position, uint8_t *type, float *mz_observed,
    float *correlation_observed, float *mass_error_observed,
    uint64_t *precursor_idx, uint64_t *rank) {
  clear_error();
  if (!q) return set_error(ADSR_ERR_INVALID_ARGUMENT, "NULL handle");
  copy_out(mz_library, q->f_mz_library);
  copy_out(mz, q->f_mz);
  copy_out(intensity, q->f_intensity);
  copy_out(cardinality, q->f_cardinality);
  copy_out(charge, q->f_charge);
  copy_out(loss_type, q->f_loss_type);
  copy_out(number, q->f_number);
  copy_out(position, q->f_position);
  copy_out(type, q->f_type);
  copy_out(mz_observed, q->f_mz_observed);
  copy_out(correlation_observed, q->f_correlation);
  copy_out(mass_error_observed, q->f_mass_error);
  copy_out(precursor_idx, q->f_precursor_idx);
  copy_out(rank, q->f_rank);
  return ADSR_OK;
}

extern "C" void adsr_quantified_free(adsr_quantified *q) { delete q; }

extern "C" adsr_status adsr_quantified_timings(const adsr_quantified *q, double *plan_ms,
                                               double *kernel_ms) {
  clear_error();
  if (!q) return set_error(ADSR_ERR_INVALID_ARGUMENT, "NULL handle");
  if (plan_ms) *plan_ms = q->plan_ms;
  if (kernel_ms) *kernel_ms = q->kernel_ms;
  return ADSR_OK;
}
