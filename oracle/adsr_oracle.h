/*
 * adsr_oracle.h -- CPU restatement of alphadia-search-rs candidate selection.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path and the timed CPU baseline.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product library never links it.
 *
 * Every function restates one function of the reference (Rust crate
 * alphadia_search_rs 1.0.1-dev0) and cites the file:line it follows.  All
 * float arithmetic is f32, evaluated left to right with one rounding per
 * operation and no FMA contraction (Rust semantics; build with
 * -ffp-contract=off).  ln/exp are glibc logf/expf, the libm the Rust binary
 * calls on x86-64 Linux.
 *
 * How each part is pinned (tests/test_oracle.py unless stated):
 *   grid, nearest bin, bin range      reference KATs (length 2,995,974 at
 *                                     100..2000, clamp, nearest/range cases)
 *                                     + brute force over random queries
 *   RT index, cycle limits            reference KATs (src/rt_index/tests.rs)
 *   CSR build, fill_xic_slice,        reference KATs + brute force on tiny
 *   dense XIC                         runs (snap every peak, sum), including
 *                                     the duplicate-entry case below
 *   Gaussian kernel                   reference KATs (length, sum 1 +- 1e-6,
 *                                     centre max) + closed form + symmetry
 *   convolution                       direct 21-tap definition on random
 *                                     rows, constant-row closed form,
 *                                     L < kernel width -> zeros
 *   log-dot score                     reference KAT + closed form
 *   local maxima                      reference KATs + brute force
 *   fragment filter/top-k/m/z sort    reference KATs (multiset at ties)
 *   candidates                        reference KATs (to_arrays order)
 *   logf                              identical to glibc logf by definition
 *                                     (it calls it)
 * Parity unpinned (no fixture or closed form fixes the bits; the oracle
 * follows the reference's published algorithms):
 *   - kernel-weight bits: ndarray 0.16.1 sum() order, restated as an
 *     8-accumulator unrolled fold (orc_ndarray_sum)
 *   - which of several equal cycle entries Rust 1.88 binary_search returns
 *     (restated as the last one; orc_rust_bsearch_u16)
 *   - select_nth_unstable_by among tied intensities at the top-k cut:
 *     restated exactly for up to 16 fragments (insertion sort, the
 *     k == 0 and k == len-1 special cases); beyond 16 the introselect is
 *     replaced by a stable sort, which picks the same set whenever the
 *     intensities at the cut differ
 *   - rayon collect order (documented as order preserving)
 *
 * Scoring and quantification (adsr_oracle_score.c) are pinned by the
 * reference's unit tests of src/peak_group_scoring, src/utils, src/idf and
 * src/dense_xic_observation, by numpy/lgamma closed forms, and by an
 * independent float64 implementation of every feature
 * (tests/test_oracle_score.py).  Parity unpinned there: log10f bits (glibc
 * vs another libm), the sign of an all-zero f32 sum (Rust 1.88
 * Iterator::sum folds from -0.0, restated so), and which row binary_search
 * returns for a precursor_idx that two library rows share (the first).
 */
#ifndef ADSR_ORACLE_H
#define ADSR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- m/z grid: src/mz_index/mod.rs ---------------------------------- */
size_t orc_ppm_index(float resolution_ppm, float mz_start, float mz_end,
                     float *out, size_t cap);
size_t orc_find_closest_index(const float *grid, size_t n, float mz);
void orc_mz_range_indices(const float *grid, size_t n, float lower_mz,
                          float upper_mz, size_t *start, size_t *end);

/* ---- Rust core slice::binary_search_by (last-equal semantics) -------- */
/* returns 1 for Ok(idx), 0 for Err(idx) */
int orc_rust_bsearch_f32(const float *v, size_t n, float key, size_t *idx);
int orc_rust_bsearch_u16(const uint16_t *v, size_t n, uint16_t key,
                         size_t *idx);

/* ---- RT index: src/rt_index/mod.rs ---------------------------------- */
float orc_mean_delta_t(const float *rt, size_t n);
void orc_cycle_idx_limits(const float *rt, size_t n, float delta_t,
                          float precursor_rt, float rt_tolerance,
                          size_t num_cycles_padding, size_t *lower,
                          size_t *upper);

/* ---- DIA data: src/dia_data*, src/quadrupole_observation ------------ */
typedef struct orc_obs {
  float isolation_window[2];
  size_t num_cycles;
  uint32_t *slice_starts; /* n_grid + 1 */
  uint16_t *cycle_indices;
  float *intensities;
  size_t n_entries;
} orc_obs;

typedef struct orc_dia {
  size_t n_obs;
  orc_obs *obs;
  float *rt; /* RT index values (delta_scan_idx == 0 spectra) */
  size_t n_rt;
  float delta_t;
  const float *grid; /* global 1-ppm grid (shared) */
  size_t n_grid;
} orc_dia;

const float *orc_global_grid(size_t *n);
orc_dia *orc_dia_build(const int64_t *spectrum_delta_scan_idx,
                       const float *isolation_lower_mz,
                       const float *isolation_upper_mz,
                       const int64_t *spectrum_peak_start_idx,
                       const int64_t *spectrum_peak_stop_idx,
                       const int64_t *spectrum_cycle_idx,
                       const float *spectrum_rt, size_t n_spectra,
                       const float *peak_mz, const float *peak_intensity,
                       size_t n_peaks, int nthreads);
void orc_dia_free(orc_dia *d);
size_t orc_valid_observations(const orc_dia *d, float precursor_mz,
                              size_t *out);
void orc_fill_xic_slice_raw(const float *grid, size_t n_grid,
                            const uint32_t *slice_starts,
                            const uint16_t *cycle_indices,
                            const float *intensities, float *dense_xic,
                            size_t cycle_start_idx, size_t cycle_stop_idx,
                            float mass_tolerance, float mz);
void orc_fill_xic_slice(const orc_dia *d, size_t obs_idx, float *dense_xic,
                        size_t cycle_start_idx, size_t cycle_stop_idx,
                        float mass_tolerance, float mz);
void orc_dense_xic(const orc_dia *d, float precursor_mz,
                   size_t cycle_start_idx, size_t cycle_stop_idx,
                   float mass_tolerance, const float *fragment_mz, size_t nf,
                   float *out);

/* ---- kernel / convolution / score / maxima -------------------------- */
size_t orc_gaussian_kernel(float fwhm_rt, float sigma_scale_rt,
                           size_t kernel_width, float rt_resolution,
                           float *out, size_t cap);
size_t orc_gaussian_kernel_sigma(float sigma, size_t kernel_width, float *out,
                                 size_t cap);
float orc_ndarray_sum(const float *x, size_t n);
void orc_convolution(const float *kernel, size_t kernel_len, const float *xic,
                     size_t rows, size_t cols, float *out);
void orc_axis_log_dot_product(const float *a, size_t rows, size_t cols,
                              const float *w, float *out);
size_t orc_find_local_maxima(const float *a, size_t n, size_t offset,
                             size_t *indices, float *values);

/* ---- fragment filtering: src/speclib_flat/mod.rs:240-345 ------------ */
size_t orc_filter_sort_fragments(size_t n, const float *fragment_mz,
                                 const float *fragment_intensity,
                                 const uint8_t *fragment_number,
                                 const uint8_t *fragment_type, int non_zero,
                                 int filter_y1_ions, size_t top_k,
                                 size_t *out_order);

/* ---- spectral library: src/speclib_flat/mod.rs:98-188 --------------- */
typedef struct orc_lib {
  size_t n_prec, n_frag;
  uint64_t *precursor_idx;
  float *precursor_mz, *precursor_rt;
  uint64_t *frag_start, *frag_stop;
  float *fragment_mz, *fragment_intensity;
  uint8_t *fragment_number, *fragment_type;
  uint8_t *precursor_naa;      /* sorted order; scoring only (may be NULL) */
  float *fragment_mz_library;  /* fragment order; scoring only (may be NULL) */
  uint32_t *idf_counts;        /* built by the first orc_score (the
                                  reference builds it with the library) */
} orc_lib;

orc_lib *orc_lib_build(size_t n_prec, const uint64_t *precursor_idx,
                       const float *precursor_mz, const float *precursor_rt,
                       const uint64_t *flat_frag_start_idx,
                       const uint64_t *flat_frag_stop_idx, size_t n_frag,
                       const float *fragment_mz,
                       const float *fragment_intensity,
                       const uint8_t *fragment_number,
                       const uint8_t *fragment_type);
/* The same with the two columns scoring reads as well. */
orc_lib *orc_lib_build_full(size_t n_prec, const uint64_t *precursor_idx,
                            const float *precursor_mz,
                            const float *precursor_rt,
                            const uint8_t *precursor_naa,
                            const uint64_t *flat_frag_start_idx,
                            const uint64_t *flat_frag_stop_idx, size_t n_frag,
                            const float *fragment_mz,
                            const float *fragment_mz_library,
                            const float *fragment_intensity,
                            const uint8_t *fragment_number,
                            const uint8_t *fragment_type);
void orc_lib_free(orc_lib *l);

/* ---- selection: src/peak_group_selection/mod.rs:87-195 -------------- */
typedef struct orc_params {
  float fwhm_rt;
  size_t kernel_size;
  size_t peak_length;
  float mass_tolerance;
  float rt_tolerance;
  size_t candidate_count;
  size_t top_k_fragments;
  const float *rt_tolerance_per_precursor; /* NULL or n_prec (library order) */
  const float *sigma_per_window;           /* NULL or n_obs */
} orc_params;

typedef struct orc_cands {
  size_t n;
  uint64_t *precursor_idx, *rank, *cycle_center, *cycle_start, *cycle_stop;
  float *score;
} orc_cands;

orc_cands *orc_select(const orc_dia *d, const orc_lib *l, const orc_params *p,
                      size_t row_begin, size_t row_end, int nthreads);
void orc_cands_free(orc_cands *c);

/* One precursor, with intermediate arrays exposed for stage-wise checks.
 * Returns 0 when the RT range is empty (reference returns None).
 * xic/conv must hold top_k*L floats, score L floats.                     */
int orc_select_one(const orc_dia *d, const orc_lib *l, const orc_params *p,
                   size_t row, size_t *cycle_start, size_t *cycle_stop,
                   size_t *n_frag, float *frag_mz, float *frag_w, float *xic,
                   float *conv, float *score, size_t cap_cells);

float orc_logf(float x); /* glibc logf, as called by the reference */
/* Mismatches of y[i] against glibc logf(bits first + i), i < n (NaN == NaN);
 * smallest mismatching i in *first_bad (n if none).  Threaded checker. */
uint64_t orc_logf_check_range(uint32_t first, size_t n, const float *y,
                              int nthreads, size_t *first_bad);

/* ---- scoring: src/peak_group_scoring (adsr_oracle_score.c) ---------- */
/* CandidateFeature f32 columns (src/candidate/features.rs:51-96, to_dict_
 * arrays order after precursor_idx and rank) */
enum {
  ORC_F_SCORE, ORC_F_MEAN_CORRELATION, ORC_F_MEDIAN_CORRELATION,
  ORC_F_CORRELATION_STD, ORC_F_INTENSITY_CORRELATION, ORC_F_NUM_FRAGMENTS,
  ORC_F_NUM_SCANS, ORC_F_NUM_OVER_95, ORC_F_NUM_OVER_90, ORC_F_NUM_OVER_80,
  ORC_F_NUM_OVER_50, ORC_F_NUM_OVER_0, ORC_F_NUM_OVER_0_RANK_0_5,
  ORC_F_NUM_OVER_0_RANK_6_11, ORC_F_NUM_OVER_0_RANK_12_17,
  ORC_F_NUM_OVER_0_RANK_18_23, ORC_F_NUM_OVER_50_RANK_0_5,
  ORC_F_NUM_OVER_50_RANK_6_11, ORC_F_NUM_OVER_50_RANK_12_17,
  ORC_F_NUM_OVER_50_RANK_18_23, ORC_F_HYPERSCORE_INTENSITY_OBSERVATION,
  ORC_F_HYPERSCORE_INTENSITY_LIBRARY, ORC_F_HYPERSCORE_INVERSE_MASS_ERROR,
  ORC_F_RT_OBSERVED, ORC_F_DELTA_RT, ORC_F_LONGEST_B_SERIES,
  ORC_F_LONGEST_Y_SERIES, ORC_F_NAA, ORC_F_WEIGHTED_MASS_ERROR,
  ORC_F_LOG10_B_ION_INTENSITY, ORC_F_LOG10_Y_ION_INTENSITY, ORC_F_FWHM_RT,
  ORC_F_IDF_HYPERSCORE, ORC_F_IDF_XIC_DOT_PRODUCT,
  ORC_F_IDF_INTENSITY_DOT_PRODUCT, ORC_F_MEDIAN_PROFILE_SUM,
  ORC_F_MEDIAN_PROFILE_SUM_FILTERED, ORC_F_NUM_PROFILES,
  ORC_F_NUM_PROFILES_FILTERED, ORC_F_NUM_OVER_0_TOP6_IDF,
  ORC_F_NUM_OVER_50_TOP6_IDF, ORC_N_FEATURES
};

float orc_correlation_safe(const float *x, const float *y, size_t n);
float orc_gamma_ln(float x);
float orc_hyperscore_weighted(const uint8_t *types, const float *inten,
                              const uint8_t *matched, const float *w,
                              size_t n);
void orc_longest_ion_series(const uint8_t *types, const uint8_t *nums,
                            const uint8_t *matched, size_t n, uint8_t *b,
                            uint8_t *y);
float orc_intensity_ion_series(const uint8_t *types, const float *inten,
                               const uint8_t *matched, size_t n,
                               uint8_t target);
float orc_dot(const float *a, const float *b, size_t n);
float orc_fwhm_rt(const float *profile, size_t n, uint64_t cycle_start,
                  const float *rt);
void orc_normalize_profiles(const float *x, size_t rows, size_t cols,
                            size_t dil, float *out);
size_t orc_filter_non_zero(const float *x, size_t rows, size_t cols,
                           float *out);
void orc_median_axis_0(const float *x, size_t rows, size_t cols, float *out);
float orc_median(const float *v, size_t n);
float orc_std(const float *v, size_t n);
void orc_ranked_mask(const float *v, size_t n, size_t r1, size_t r2,
                     uint8_t *mask);
size_t orc_count_above(const float *v, size_t n, float thr,
                       const uint8_t *mask);
float orc_weighted_mae(const float *err, const float *w, size_t n);
void orc_fragment_mz_errors(const float *dmz, const float *dxic, size_t rows,
                            size_t cols, const float *mz_lib, float *mz_obs,
                            float *err);
void orc_fill_xic_and_mz_slice(const orc_dia *d, size_t obs_idx, float *xic,
                               float *dmz, uint64_t c0, uint64_t c1,
                               float tol, float mz);
void orc_dense_xic_mz(const orc_dia *d, float precursor_mz, uint64_t c0,
                      uint64_t c1, float tol, const float *fragment_mz,
                      size_t nf, float *xic, float *mz);
uint32_t *orc_idf_counts(const float *fragment_mz_library, size_t n);
float orc_idf(const uint32_t *counts, float total, float mz);
/* PeakGroupQuantification.quantify per candidate, before the intensity
 * filter and the flat layout (see adsr_oracle_score.c) */
void orc_quantify(const orc_dia *d, const orc_lib *l, float tolerance_ppm,
                  size_t top_k, size_t n, const uint64_t *precursor_idx,
                  const uint64_t *cycle_start, const uint64_t *cycle_stop,
                  const uint64_t *cycle_center, size_t ks, size_t *row,
                  size_t *nf, size_t *raw, float *intensity, float *mz_obs,
                  float *corr, float *mass_err, float *rt_observed, int nthreads);
/* PeakGroupScoring.score over n candidates: features [ORC_N_FEATURES][n]
 * (column-major), found[i] = 0 where precursor_idx is not in the library.
 * The library must come from orc_lib_build_full. */
void orc_score(const orc_dia *d, const orc_lib *l, float mass_tolerance,
               size_t top_k, size_t n, const uint64_t *precursor_idx,
               const float *score, const uint64_t *cycle_start,
               const uint64_t *cycle_stop, const uint64_t *cycle_center,
               float *features, uint8_t *found, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
