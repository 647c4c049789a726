// adsr_internal.hpp -- shared host/device declarations of the gfx950
// candidate-selection library (host builders in adsr_host.cpp, kernels and
// launches in adsr_device.hip).
#pragma once

#include <atomic>
#include <cstdlib>
#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/adsr.h"

namespace adsr {

// ---- m/z grid: src/mz_index/mod.rs:4-6 (RESOLUTION_PPM=1, 150..2000 Th) ----
// Bin lookups use a table over the top bits of the f32 pattern: for a
// positive float x, key = (bits(x) >> kBucketShift) - base, and
// bucket[key] = number of grid points below the smallest float with that key.
// One bucket spans 16 ulps while a 1-ppm bin spans at least 8.39 ulps, so a
// bucket holds at most two grid points and every lookup ends after at most
// two exact comparisons against the grid values.
constexpr int kBucketShift = 4;

struct Grid {
  std::vector<float> mz;        // the exact f32 recurrence
  std::vector<uint32_t> bucket;  // n_bucket + 1 entries
  uint32_t base_key = 0;         // bits(mz[0]) >> kBucketShift
  uint32_t n_bucket = 0;
};
const Grid &global_grid();
// Number of grid points < x (lower bound), exact.
uint32_t grid_lower_bound(const Grid &g, float x);
// src/mz_index/mod.rs:55-90 find_closest_index, exact.
uint32_t grid_find_closest(const Grid &g, float x);

// ---- errors -----------------------------------------------------------------
adsr_status set_error(adsr_status st, const char *fmt, ...);
void clear_error();
// True for the allocation that ADSR_FAIL_ALLOC names (a test hook).
bool alloc_should_fail();

// ---- device-resident copies -------------------------------------------------
struct DeviceDia {
  bool ready = false;
  int device = -1;
  float *grid = nullptr;
  uint32_t *bucket = nullptr;
  // per bucket: grid points below it and the first two grid values from it
  // (u32 count, f32, f32, unused)
  uint32_t *bucket2 = nullptr;
  uint32_t *bucket1 = nullptr;  // 4-byte form of the same buckets
  float *rt = nullptr;
  float *iso_lo = nullptr, *iso_hi = nullptr;
  uint32_t *slice_starts = nullptr;  // n_obs * (n_grid + 1)
  uint64_t *entry_base = nullptr;    // n_obs + 1
  uint16_t *cycles = nullptr;
  float *intens = nullptr;
  size_t bytes = 0;
};

struct DeviceLib {
  bool ready = false;
  int device = -1;
  uint64_t *p_idx = nullptr;
  float *p_mz = nullptr, *p_rt = nullptr;
  uint64_t *f_start = nullptr, *f_stop = nullptr;
  float *f_mz = nullptr, *f_int = nullptr;
  uint8_t *f_type = nullptr, *f_num = nullptr;
  // scoring only, uploaded by the first score call: precursor_naa,
  // fragment_mz_library and the IDF bin counts of all fragment_mz_library
  bool scoring = false;
  uint8_t *p_naa = nullptr;
  float *f_mzl = nullptr;
  uint32_t *idf_counts = nullptr;
  size_t bytes = 0;
};

void device_dia_free(DeviceDia &d);
void device_lib_free(DeviceLib &l);

}  // namespace adsr

struct adsr_dia;

namespace adsr {

// DIAData metadata shared by the host and the device CSR builders
// (src/dia_data/mod.rs:41-72, src/dia_data_builder/mod.rs:27-80): argument
// checks, RT axis, one observation per delta_scan_idx with its spectra in
// ascending spectrum order, entry bases, window bounds and cycle counts.
// With host_csr the CSR arrays are allocated (zeroed) for the host fill.
adsr_status dia_prepare(const int64_t *delta, const float *iso_lo,
                        const float *iso_hi, const int64_t *pstart,
                        const int64_t *pstop, const int64_t *cyc,
                        const float *srt, size_t n_spectra,
                        const float *peak_mz, const float *peak_int,
                        size_t n_peaks, const float *cycle,
                        const size_t cycle_shape[4], bool host_csr,
                        std::vector<std::vector<size_t>> &spectra,
                        adsr_dia **out);

// Copies observation o's CSR of a device-built handle to the host (any of
// the three pointers may be NULL).
adsr_status device_dia_window(const adsr_dia *d, size_t o, uint32_t *ss,
                              uint16_t *ci, float *in);

}  // namespace adsr

// ---- opaque handles ---------------------------------------------------------
struct adsr_dia {
  // src/dia_data/mod.rs:15-21
  std::vector<float> rt_values;  // all spectrum_rt (copied)
  std::vector<float> cycle;
  size_t cycle_shape[4] = {0, 0, 0, 0};
  // src/rt_index/mod.rs:11-14
  std::vector<float> rt_index;
  float delta_t = 1.0f;
  // src/quadrupole_observation/mod.rs:12-25, one per delta_scan_idx,
  // concatenated: observation o owns slice_starts[o*(B+1) .. (o+1)*(B+1))
  // and entries [entry_base[o], entry_base[o+1]).
  size_t n_obs = 0;
  size_t n_grid = 0;
  std::vector<float> iso_lo, iso_hi;
  std::vector<uint64_t> num_cycles;
  std::vector<uint32_t> slice_starts;
  std::vector<uint64_t> entry_base;
  std::vector<uint16_t> cycles;
  std::vector<float> intens;
  bool rt_sorted = true;
  bool rt_strict = true;  // strictly increasing
  // false: the CSR was built on the device and lives only in `dev`
  bool host_csr = true;
  double build_ms[3] = {0, 0, 0};  // device build: prepare, copy, kernels
  adsr::DeviceDia dev;
  std::mutex mu;
};

struct adsr_lib {
  // src/speclib_flat/mod.rs:6-70, precursor arrays sorted by precursor_idx
  size_t n_prec = 0, n_frag = 0;
  std::vector<uint64_t> precursor_idx;
  std::vector<float> precursor_mz_library, precursor_mz, precursor_rt_library,
      precursor_rt;
  std::vector<uint8_t> precursor_naa;
  std::vector<uint64_t> frag_start, frag_stop;
  std::vector<float> fragment_mz_library, fragment_mz, fragment_intensity;
  std::vector<uint8_t> fragment_cardinality, fragment_charge,
      fragment_loss_type, fragment_number, fragment_position, fragment_type;
  uint64_t max_raw_fragments = 0;
  // no fragment index is shared by two precursors (needed to scatter
  // per-fragment gradients without atomics)
  bool frag_ranges_disjoint = true;
  adsr::DeviceLib dev;
  std::mutex mu;
  // search workspaces kept for the next search on this library (each one
  // used by one search at a time; freed with the library)
  std::vector<void *> idle_workspaces;
};

namespace adsr {
// Frees the idle search workspaces of a library (adsr_device.hip).
void release_lib_workspaces(adsr_lib *l);
// Page-locked host blocks for candidate tables, recycled through a bounded
// pool: acquire returns a block of at least `bytes` and its capacity.
void *pinned_acquire(size_t bytes, size_t *cap);
void pinned_release(void *p, size_t cap);
}  // namespace adsr

// A candidate table left in HBM by adsr_select_device: `rows` holds n
// device records (precursor_idx, rank, score, cycle centre).
struct adsr_device_candidates {
  void *rows = nullptr;
  size_t n = 0;
  uint64_t pl = 0, nrt = 0;  // peak_length and RT-index length
};

// CandidateFeatureCollection (src/candidate/features.rs:196-430): one row
// per scored candidate, the f32 features column-major [ADSR_N_FEATURES][n].
struct adsr_features {
  std::vector<uint64_t> precursor_idx, rank;
  std::vector<float> columns;
  double plan_ms = 0, kernel_ms = 0;
};

// SpecLibFlatQuantified (src/speclib_flat_quantified/mod.rs:18-82):
// precursor columns sorted by precursor_idx, fragment columns in candidate
// order with per-precursor [start, stop) ranges.
struct adsr_quantified {
  std::vector<uint64_t> precursor_idx, rank, frag_start, frag_stop;
  std::vector<float> mz_library, mz, rt_library, rt, rt_observed;
  std::vector<uint8_t> naa;
  std::vector<float> f_mz_library, f_mz, f_intensity, f_mz_observed,
      f_correlation, f_mass_error;
  std::vector<uint8_t> f_cardinality, f_charge, f_loss_type, f_number,
      f_position, f_type;
  std::vector<uint64_t> f_precursor_idx, f_rank;
  double plan_ms = 0, kernel_ms = 0;
};

// CandidateCollection (src/candidate/entry.rs:5-16, 41-193): the nine
// to_arrays columns in one block, eight u64 columns then the f32 score,
// n entries each (68 bytes per row).  A selection's table is written into a
// page-locked block by one device-to-host copy; from_arrays uses malloc.
struct adsr_candidates {
  size_t n = 0;
  uint64_t *precursor_idx = nullptr, *rank = nullptr, *scan_center = nullptr,
           *scan_start = nullptr, *scan_stop = nullptr, *cycle_center = nullptr,
           *cycle_start = nullptr, *cycle_stop = nullptr;
  float *score = nullptr;
  void *block = nullptr;
  size_t block_cap = 0;
  bool pinned = false;

  static constexpr size_t kRowBytes = 8 * 8 + 4;
  // points the columns into `blk`: n rows, columns `stride` rows apart
  // (stride 0: n)
  void bind(void *blk, size_t rows, size_t stride = 0) {
    if (stride == 0) stride = rows;
    n = rows;
    block = blk;
    uint64_t *u = static_cast<uint64_t *>(blk);
    uint64_t **cols[8] = {&precursor_idx, &rank, &scan_center, &scan_start,
                          &scan_stop, &cycle_center, &cycle_start, &cycle_stop};
    for (int i = 0; i < 8; ++i) *cols[i] = u + (size_t)i * stride;
    score = reinterpret_cast<float *>(u + 8 * stride);
  }
  ~adsr_candidates() {
    if (!block) return;
    if (pinned) adsr::pinned_release(block, block_cap);
    else std::free(block);
  }
};
