/*
 * adsr.h -- C ABI of the MI355X (gfx950) implementation of the
 * alphadia-search-rs candidate-selection path.
 *
 * Plain pointers and sizes only.  Every function returns an adsr_status;
 * on failure adsr_last_error() returns a thread-local message.  Handles own
 * host and device copies of their inputs; nothing borrowed outlives a call.
 *
 * Each entry point names the reference interface it replaces
 * (MannLabs/alphadia-search-rs 1.0.1-dev0, path:line).
 */
#ifndef ADSR_H
#define ADSR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum adsr_status {
  ADSR_OK = 0,
  ADSR_ERR_INVALID_ARGUMENT = 1, /* maps to ValueError */
  ADSR_ERR_DEVICE = 2,           /* HIP failure, no gfx950 device */
  ADSR_ERR_OUT_OF_MEMORY = 3,
  ADSR_ERR_UNSUPPORTED = 4,      /* input beyond a documented device limit */
  ADSR_ERR_INTERNAL = 5
} adsr_status;

typedef struct adsr_dia adsr_dia;               /* DIAData */
typedef struct adsr_lib adsr_lib;               /* SpecLibFlat */
typedef struct adsr_candidates adsr_candidates; /* CandidateCollection */
/* CandidateCollection kept in device memory (adsr_select_device) */
typedef struct adsr_device_candidates adsr_device_candidates;

/* SelectionParameters (src/peak_group_selection/parameters.rs:4-43) plus two
 * optional extensions; with both NULL the result equals the reference.    */
typedef struct adsr_selection_params {
  float fwhm_rt;            /* default 3.0 */
  uint64_t kernel_size;     /* default 20 */
  uint64_t peak_length;     /* default 5 */
  float mass_tolerance;     /* ppm, default 7.0 */
  float rt_tolerance;       /* seconds, default 200.0 */
  uint64_t candidate_count; /* default 3 */
  uint64_t top_k_fragments; /* default 12 */
  /* extension: per-precursor rt_tolerance in library (precursor_idx-sorted)
   * order, length = number of precursors; NULL = rt_tolerance for all.    */
  const float *rt_tolerance_per_precursor;
  /* extension: Gaussian sigma (cycles) per quadrupole observation, used for
   * precursors whose first valid observation is that window; NULL = the
   * kernel from fwhm_rt for every precursor.                               */
  const float *sigma_per_window;
  uint64_t n_sigma_per_window;
} adsr_selection_params;

/* Fills the reference defaults (parameters.rs:24-43). */
void adsr_selection_params_default(adsr_selection_params *p);

const char *adsr_last_error(void);

/* ---- device / backend registry (src/simd/mod.rs:80-134, src/lib.rs:53-72) */
/* 1 if a gfx950 device is visible and usable. */
int adsr_device_available(void);
/* Known backend names: "scalar", "neon" (the reference's, src/simd/mod.rs:
 * 53-58) and "hip", the only one implemented here.  set fails with
 * INVALID_ARGUMENT and the reference's messages (src/simd/mod.rs:97-119):
 * "Unknown backend ..." for other names, "Backend '<name>' is not available
 * on this system" for "scalar", "neon", and "hip" without a gfx950 device
 * (both are ValueError in Python, src/lib.rs:58-61).                       */
adsr_status adsr_backend_set(const char *name);
void adsr_backend_clear(void);
/* Writes the current / optimal backend name ("hip" or "none"). */
adsr_status adsr_backend_current(char *out, size_t cap);
adsr_status adsr_backend_optimal(char *out, size_t cap);

/* ---- DIAData (src/dia_data/mod.rs:41-135) ------------------------------ */
/* DIAData.from_arrays (src/dia_data/mod.rs:43-72): builds the 1-ppm binned
 * CSR per quadrupole window (src/dia_data_builder/mod.rs:11-139) on host
 * threads.  cycle is copied (shape cycle_shape[4]) and returned unchanged. */
adsr_status adsr_dia_create(const int64_t *spectrum_delta_scan_idx,
                            const float *isolation_lower_mz,
                            const float *isolation_upper_mz,
                            const int64_t *spectrum_peak_start_idx,
                            const int64_t *spectrum_peak_stop_idx,
                            const int64_t *spectrum_cycle_idx,
                            const float *spectrum_rt, size_t n_spectra,
                            const float *peak_mz, const float *peak_intensity,
                            size_t n_peaks, const float *cycle,
                            const size_t cycle_shape[4], adsr_dia **out);
/* DIAData.from_arrays with the CSR built on the device.  Same arguments,
 * checks and result as adsr_dia_create, bit for bit: DIADataBuilder's bin
 * snap and stable (bin, cycle) sort (src/dia_data_builder/mod.rs:52-139)
 * run as gfx950 kernels (a snap pass and a three-pass stable radix sort per
 * observation), and the CSR stays in HBM only; adsr_dia_observation copies
 * one observation back on request.  Needs a gfx950 device (DEVICE error
 * otherwise).  Peak arrays are read once and not kept.                    */
adsr_status adsr_dia_create_device(const int64_t *spectrum_delta_scan_idx,
                                   const float *isolation_lower_mz,
                                   const float *isolation_upper_mz,
                                   const int64_t *spectrum_peak_start_idx,
                                   const int64_t *spectrum_peak_stop_idx,
                                   const int64_t *spectrum_cycle_idx,
                                   const float *spectrum_rt, size_t n_spectra,
                                   const float *peak_mz,
                                   const float *peak_intensity, size_t n_peaks,
                                   const float *cycle,
                                   const size_t cycle_shape[4],
                                   adsr_dia **out);
/* Wall-clock milliseconds of a device build (0 for host-built handles):
 * host preparation (spectrum order, snap items), host-to-device copy of
 * the peaks, and the device kernels (snap, sort, slice_starts).           */
adsr_status adsr_dia_build_timings(const adsr_dia *d, double *prep_ms,
                                   double *copy_ms, double *build_ms);
/* Copies the CSR and indices into device memory (HBM); idempotent. */
adsr_status adsr_dia_upload(adsr_dia *d);
void adsr_dia_free(adsr_dia *d);
/* num_observations getter (src/dia_data/mod.rs:74-77) */
size_t adsr_dia_num_observations(const adsr_dia *d);
/* get_valid_observations (src/dia_data/mod.rs:79-87) */
adsr_status adsr_dia_valid_observations(const adsr_dia *d, float mz,
                                        uint64_t *out, size_t cap,
                                        size_t *n_out);
/* memory_footprint_bytes (src/dia_data/mod.rs:90-108) */
uint64_t adsr_dia_memory_footprint_bytes(const adsr_dia *d);
/* rt_values getter (src/dia_data/mod.rs:128-131): all spectrum_rt */
size_t adsr_dia_rt_values(const adsr_dia *d, float *out, size_t cap);
/* cycle getter (src/dia_data/mod.rs:133-135) */
void adsr_dia_cycle_shape(const adsr_dia *d, size_t shape[4]);
void adsr_dia_cycle_copy(const adsr_dia *d, float *out);
/* RT index (src/rt_index/mod.rs:53-75): values and mean delta_t */
size_t adsr_dia_rt_index(const adsr_dia *d, float *out, size_t cap,
                         float *delta_t);
/* Read access to one observation's CSR (src/quadrupole_observation/mod.rs:
 * 12-25).  Any output pointer may be NULL; sizes are always written.       */
adsr_status adsr_dia_observation(const adsr_dia *d, size_t obs,
                                 float isolation_window[2],
                                 uint64_t *num_cycles, uint32_t *slice_starts,
                                 size_t *n_slice_starts,
                                 uint16_t *cycle_indices, float *intensities,
                                 size_t *n_entries);

/* ---- m/z grid (src/mz_index/mod.rs:8-127) ----------------------------- */
/* Global 1-ppm grid over 150-2000 Th; copies up to cap values. */
size_t adsr_mz_grid(float *out, size_t cap);

/* ---- SpecLibFlat (src/speclib_flat/mod.rs:98-199) --------------------- */
/* SpecLibFlat.from_arrays: stable sort of the precursor arrays by
 * precursor_idx; fragment arrays keep their order.                        */
adsr_status adsr_lib_create(
    size_t n_precursors, const uint64_t *precursor_idx,
    const float *precursor_mz_library, const float *precursor_mz,
    const float *precursor_rt_library, const float *precursor_rt,
    const uint8_t *precursor_naa, const uint64_t *flat_frag_start_idx,
    const uint64_t *flat_frag_stop_idx, size_t n_fragments,
    const float *fragment_mz_library, const float *fragment_mz,
    const float *fragment_intensity, const uint8_t *fragment_cardinality,
    const uint8_t *fragment_charge, const uint8_t *fragment_loss_type,
    const uint8_t *fragment_number, const uint8_t *fragment_position,
    const uint8_t *fragment_type, adsr_lib **out);
adsr_status adsr_lib_upload(adsr_lib *l);
void adsr_lib_free(adsr_lib *l);
size_t adsr_lib_num_precursors(const adsr_lib *l);
size_t adsr_lib_num_fragments(const adsr_lib *l);

/* ---- PeakGroupSelection.search (src/peak_group_selection/mod.rs:87-128) */
/* Runs the fused gfx950 pipeline on the device-resident data and returns
 * the candidate table in library order.  Uploads lazily if needed.        */
adsr_status adsr_select(const adsr_dia *d, const adsr_lib *l,
                        const adsr_selection_params *p,
                        adsr_candidates **out);

/* Timing-oriented variant of adsr_select.  Runs `warmup` untimed searches,
 * then `steps` timed ones on the library's stream: *step_ms is the average
 * device time of one search (plan, bins, select, compaction), measured with
 * HIP events around the whole loop, and *select_kernel_ms the average
 * duration of the select kernel, from event pairs around each launch.  When
 * graph_step_ms is non-NULL the search is also captured once in a hipGraph
 * and `steps` replays are timed.  threads_per_block in {64,128,256} and
 * precursors_per_wave in {1,2,4,8,16} pick the launch shape; 0 = autotuned.
 * The chosen shape is written back when the pointers are non-NULL.  The
 * candidate table of the last search is returned.                        */
adsr_status adsr_select_timed(const adsr_dia *d, const adsr_lib *l,
                              const adsr_selection_params *p, int warmup,
                              int steps, int *threads_per_block,
                              int *precursors_per_wave, double *step_ms,
                              double *select_kernel_ms, double *graph_step_ms,
                              adsr_candidates **out);

/* Algorithmic HBM bytes of one search over the whole library: per
 * precursor 24 (m/z, RT, fragment range) + 10 per raw fragment + for every
 * valid window and selected fragment 4*(bins+1) (slice_starts) + 2 per CSR
 * entry in those bins (cycle index) + 4 per entry inside the RT window
 * (intensity), + 68 per candidate written.  Counted on the device.       */
adsr_status adsr_select_algorithmic_bytes(const adsr_dia *d,
                                          const adsr_lib *l,
                                          const adsr_selection_params *p,
                                          double *bytes, double *entries,
                                          double *hits);

/* Diagnostic: the fused kernel's per-cycle score vector for library rows
 * [row_begin, row_end), written densely (stride max_cycles, NaN-padded),
 * with each row's cycle_start and cycle_stop.                             */
adsr_status adsr_select_scores(const adsr_dia *d, const adsr_lib *l,
                               const adsr_selection_params *p,
                               size_t row_begin, size_t row_end,
                               size_t max_cycles, float *scores,
                               uint64_t *cycle_start, uint64_t *cycle_stop);

/* Diagnostic: the device plan for library rows [row_begin, row_end): the
 * RT cycle window (src/rt_index/mod.rs:89-138) and the selected fragments
 * (absolute indices into the fragment arrays, stride `stride`) in the order
 * of get_precursor_filtered(i, true, true, top_k_fragments)
 * (src/speclib_flat/mod.rs:388-456).                                      */
adsr_status adsr_select_plan(const adsr_dia *d, const adsr_lib *l,
                             const adsr_selection_params *p, size_t row_begin,
                             size_t row_end, uint64_t *cycle_start,
                             uint64_t *cycle_stop, uint32_t *n_selected,
                             uint64_t *selected, size_t stride);

/* ---- differentiable selection (no counterpart in the reference) ---------
 * Gradient of the rank-0 selection score with respect to a Gaussian sigma
 * per quadrupole observation and the library fragment intensities used as
 * score weights.  The forward runs PeakGroupSelection.search with
 * sigma_per_window and the weights in place of fragment_intensity (also for
 * the top-k fragment filter), and records each precursor's rank-0 apex; the
 * backward holds the apex and fragment selection fixed, recomputes the XIC
 * around the apex and differentiates the convolution and the weighted
 * ln(1+x) sum.  Device pointers, single GPU; calls on one state are not
 * concurrent.                                                             */
typedef struct adsr_diff adsr_diff;
adsr_status adsr_diff_create(const adsr_dia *d, const adsr_lib *l,
                             const adsr_selection_params *p,
                             adsr_diff **out);
/* sigma: host f32[num_observations]; weights: device f32[num_fragments];
 * s_star: device f32[num_precursors], rows in ascending precursor_idx
 * (stable), the order of the candidate table; 0 without a candidate;
 * has_candidate: optional device u8[num_precursors].                      */
adsr_status adsr_diff_forward(adsr_diff *s, const float *sigma,
                              const float *weights, float *s_star,
                              uint8_t *has_candidate);
/* adsr_diff_forward enqueued on `stream` (a hipStream_t; NULL: the null
 * stream) without waiting for the device, so that the caller can queue
 * the loss and its gradient behind it.  Argument errors are returned at
 * once; the status errors adsr_diff_forward reports from the device are
 * returned by the next adsr_diff_forward(_async), adsr_diff_backward or
 * adsr_diff_timings call on the state, which waits for this forward
 * (adsr_diff_free waits too).  weights must stay valid until the forward
 * has run on `stream`.  No counterpart in the reference.                  */
adsr_status adsr_diff_forward_async(adsr_diff *s, const float *sigma,
                                    const float *weights, float *s_star,
                                    uint8_t *has_candidate, void *stream);
/* grad_s: device f32[num_precursors] = dLoss/ds_star; outputs: device
 * f32[num_fragments] (fragments that were not selected get 0) and host
 * f64[num_observations].                                                  */
adsr_status adsr_diff_backward(adsr_diff *s, const float *grad_s,
                               float *grad_weights, double *grad_sigma);
/* Hinge loss over (target, decoy) pairs of s_star (device f32[n_rows];
 * targets, decoys: device i64[n_pairs]), for training the differentiable
 * selection; no counterpart in the reference.  terms (device
 * f32[n_pairs]) receives max(0, margin - (s[t] - s[d])) per pair (the loss
 * is their sum); grad_s (device f32[n_rows], zeroed by the caller) gets
 * dLoss/ds: -1 to the target and +1 to the decoy of every pair with
 * margin - (s[t] - s[d]) >= 0.  Pairs with an index outside [0, n_rows)
 * are skipped.  Enqueued on `stream` (a hipStream_t; NULL: the null
 * stream) without synchronizing.                                          */
adsr_status adsr_hinge_loss(const float *s_star, size_t n_rows,
                            const int64_t *targets, const int64_t *decoys,
                            size_t n_pairs, float margin, float *terms,
                            float *grad_s, void *stream);
/* Device time of the last forward / backward call (ms). */
adsr_status adsr_diff_timings(const adsr_diff *s, double *forward_ms,
                              double *backward_ms);
void adsr_diff_free(adsr_diff *s);

/* ---- CandidateCollection (src/candidate/entry.rs:41-193) -------------- */
size_t adsr_candidates_len(const adsr_candidates *c);
/* to_arrays (entry.rs:124-173): nine columns in reference order. */
adsr_status adsr_candidates_export(const adsr_candidates *c,
                                   uint64_t *precursor_idx, uint64_t *rank,
                                   float *score, uint64_t *scan_center,
                                   uint64_t *scan_start, uint64_t *scan_stop,
                                   uint64_t *cycle_center,
                                   uint64_t *cycle_start,
                                   uint64_t *cycle_stop);
/* from_arrays (entry.rs:70-120) */
adsr_status adsr_candidates_from_arrays(
    size_t n, const uint64_t *precursor_idx, const uint64_t *rank,
    const float *score, const uint64_t *scan_center,
    const uint64_t *scan_start, const uint64_t *scan_stop,
    const uint64_t *cycle_center, const uint64_t *cycle_start,
    const uint64_t *cycle_stop, adsr_candidates **out);
void adsr_candidates_free(adsr_candidates *c);

/* ---- candidate table in HBM (handoff to scoring without a host copy) --- */
/* PeakGroupSelection::search (src/peak_group_selection/mod.rs:87-128) with
 * the candidate table left in device memory: the same rows, in the same
 * order, as adsr_select.  Returns once the table is complete (its length is
 * known on the host).  The table does not reference the library handle.  */
adsr_status adsr_select_device(const adsr_dia *d, const adsr_lib *l,
                               const adsr_selection_params *p,
                               adsr_device_candidates **out);
size_t adsr_device_candidates_len(const adsr_device_candidates *c);
/* CandidateCollection::to_arrays (src/candidate/entry.rs:124-173) into
 * device memory: each pointer is NULL (column skipped) or a device buffer of
 * len elements (u64, score f32).  One kernel is enqueued on `stream` (a
 * hipStream_t; NULL is the null stream) and the call returns without
 * waiting for it.                                                          */
adsr_status adsr_device_candidates_export(
    const adsr_device_candidates *c, uint64_t *precursor_idx, uint64_t *rank,
    float *score, uint64_t *scan_center, uint64_t *scan_start,
    uint64_t *scan_stop, uint64_t *cycle_center, uint64_t *cycle_start,
    uint64_t *cycle_stop, void *stream);
/* The host CandidateCollection of the same table. */
adsr_status adsr_device_candidates_to_host(const adsr_device_candidates *c,
                                           adsr_candidates **out);
void adsr_device_candidates_free(adsr_device_candidates *c);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  PeakGroupSelection::search over library rows
 * [row_begin, row_end) (sorted order) with no allocation: the plan,
 * candidate slots and compaction live in a 64 MiB device array of the
 * library (one per GPU), so a call takes at most as many rows as fit it
 * (about 290k rows at top_k 12 / top-3; larger ranges fail with
 * INVALID_ARGUMENT and the row limit in the message).  The DIAData and the
 * library must already be on the current device (adsr_dia_upload,
 * adsr_lib_upload).  The candidates of those rows, in library order, are
 * written as the nine to_arrays columns into caller device buffers of
 * `capacity` rows on `stream` (a hipStream_t), and *n_out is their count;
 * more than `capacity` candidates fail with INVALID_ARGUMENT after the
 * first `capacity` are written (*n_out still the count).  The call waits
 * for the stream.  Calls that overlap on one GPU share the array and race.
 * Fragment lists longer than 48 entries are UNSUPPORTED here.             */
adsr_status adsr_select_rows_noalloc(
    const adsr_dia *d, const adsr_lib *l, const adsr_selection_params *p,
    size_t row_begin, size_t row_end, uint64_t *precursor_idx, uint64_t *rank,
    float *score, uint64_t *scan_center, uint64_t *scan_start,
    uint64_t *scan_stop, uint64_t *cycle_center, uint64_t *cycle_start,
    uint64_t *cycle_stop, size_t capacity, size_t *n_out, void *stream);

/* ---- PeakGroupScoring (src/peak_group_scoring/mod.rs:43-341) ---------- */
/* ScoringParameters (src/peak_group_scoring/parameters.rs:4-40) */
typedef struct adsr_scoring_params {
  float mass_tolerance;     /* ppm, default 7.0 */
  uint64_t top_k_fragments; /* default 100 */
} adsr_scoring_params;
void adsr_scoring_params_default(adsr_scoring_params *p);

/* CandidateFeatureCollection (src/candidate/features.rs:196-430) */
typedef struct adsr_features adsr_features;
/* f32 feature columns, in CandidateFeature / to_dict_arrays order after
 * precursor_idx and rank (features.rs:51-96, 314-410) */
enum {
  ADSR_F_SCORE, ADSR_F_MEAN_CORRELATION, ADSR_F_MEDIAN_CORRELATION,
  ADSR_F_CORRELATION_STD, ADSR_F_INTENSITY_CORRELATION, ADSR_F_NUM_FRAGMENTS,
  ADSR_F_NUM_SCANS, ADSR_F_NUM_OVER_95, ADSR_F_NUM_OVER_90,
  ADSR_F_NUM_OVER_80, ADSR_F_NUM_OVER_50, ADSR_F_NUM_OVER_0,
  ADSR_F_NUM_OVER_0_RANK_0_5, ADSR_F_NUM_OVER_0_RANK_6_11,
  ADSR_F_NUM_OVER_0_RANK_12_17, ADSR_F_NUM_OVER_0_RANK_18_23,
  ADSR_F_NUM_OVER_50_RANK_0_5, ADSR_F_NUM_OVER_50_RANK_6_11,
  ADSR_F_NUM_OVER_50_RANK_12_17, ADSR_F_NUM_OVER_50_RANK_18_23,
  ADSR_F_HYPERSCORE_INTENSITY_OBSERVATION, ADSR_F_HYPERSCORE_INTENSITY_LIBRARY,
  ADSR_F_HYPERSCORE_INVERSE_MASS_ERROR, ADSR_F_RT_OBSERVED, ADSR_F_DELTA_RT,
  ADSR_F_LONGEST_B_SERIES, ADSR_F_LONGEST_Y_SERIES, ADSR_F_NAA,
  ADSR_F_WEIGHTED_MASS_ERROR, ADSR_F_LOG10_B_ION_INTENSITY,
  ADSR_F_LOG10_Y_ION_INTENSITY, ADSR_F_FWHM_RT, ADSR_F_IDF_HYPERSCORE,
  ADSR_F_IDF_XIC_DOT_PRODUCT, ADSR_F_IDF_INTENSITY_DOT_PRODUCT,
  ADSR_F_MEDIAN_PROFILE_SUM, ADSR_F_MEDIAN_PROFILE_SUM_FILTERED,
  ADSR_F_NUM_PROFILES, ADSR_F_NUM_PROFILES_FILTERED,
  ADSR_F_NUM_OVER_0_TOP6_IDF, ADSR_F_NUM_OVER_50_TOP6_IDF, ADSR_N_FEATURES
};
/* Name of feature column i (the to_dict_arrays key), NULL past the end. */
const char *adsr_feature_name(size_t i);
/* PeakGroupScoring::score (mod.rs:43-102): every candidate in table order;
 * candidates whose precursor_idx is not in the library are skipped, as the
 * reference's filter_map does.  Limits: at most 256 selected fragments per
 * precursor and cycle_stop - cycle_start <= 32 (peak_length <= 15); a
 * candidate whose cycle indices fall outside the RT axis, where the
 * reference panics, returns INVALID_ARGUMENT.                             */
adsr_status adsr_score(const adsr_dia *d, const adsr_lib *l,
                       const adsr_candidates *c, const adsr_scoring_params *p,
                       adsr_features **out);
/* The same for a candidate table kept in HBM (adsr_select_device): no
 * host copy of the candidates. */
adsr_status adsr_score_device(const adsr_dia *d, const adsr_lib *l,
                              const adsr_device_candidates *c,
                              const adsr_scoring_params *p,
                              adsr_features **out);
size_t adsr_features_len(const adsr_features *f);
/* to_dict_arrays (features.rs:220-410): precursor_idx and rank (u64[len])
 * and the f32 columns, column-major [ADSR_N_FEATURES][len]; any pointer
 * may be NULL. */
adsr_status adsr_features_export(const adsr_features *f,
                                 uint64_t *precursor_idx, uint64_t *rank,
                                 float *columns);
/* Device time (ms) of the fragment plan and of the scoring kernel. */
adsr_status adsr_features_timings(const adsr_features *f, double *plan_ms,
                                  double *kernel_ms);
void adsr_features_free(adsr_features *f);

/* ---- PeakGroupQuantification (src/peak_group_quantification/mod.rs) -- */
/* QuantificationParameters (src/peak_group_quantification/parameters.rs) */
typedef struct adsr_quant_params {
  float tolerance_ppm;      /* default 7.0 */
  uint64_t top_k_fragments; /* default 10000 (all) */
} adsr_quant_params;
void adsr_quant_params_default(adsr_quant_params *p);
/* SpecLibFlatQuantified (src/speclib_flat_quantified/mod.rs:18-82) */
typedef struct adsr_quantified adsr_quantified;
/* PeakGroupQuantification::quantify (mod.rs:39-168): per candidate the
 * dense XIC and m/z, observed intensities (sum over cycles), m/z, mass
 * errors and correlations with the filtered median profile; fragments with
 * zero observed intensity dropped (precursor_quantified/mod.rs:26-94);
 * precursor columns sorted by precursor_idx.  Same limits as adsr_score. */
adsr_status adsr_quantify(const adsr_dia *d, const adsr_lib *l,
                          const adsr_candidates *c, const adsr_quant_params *p,
                          adsr_quantified **out);
size_t adsr_quantified_num_precursors(const adsr_quantified *q);
size_t adsr_quantified_num_fragments(const adsr_quantified *q);
/* to_dict_arrays precursor columns (speclib_flat_quantified/mod.rs:126-160),
 * num_precursors elements each; any pointer may be NULL. */
adsr_status adsr_quantified_export_precursors(
    const adsr_quantified *q, uint64_t *precursor_idx, float *mz_library,
    float *mz, float *rt_library, float *rt, uint8_t *naa, uint64_t *rank,
    float *rt_observed, uint64_t *flat_frag_start_idx,
    uint64_t *flat_frag_stop_idx);
/* to_dict_arrays fragment columns (mod.rs:161-205), num_fragments each. */
adsr_status adsr_quantified_export_fragments(
    const adsr_quantified *q, float *mz_library, float *mz, float *intensity,
    uint8_t *cardinality, uint8_t *charge, uint8_t *loss_type, uint8_t *number,
    uint8_t *position, uint8_t *type, float *mz_observed,
    float *correlation_observed, float *mass_error_observed,
    uint64_t *precursor_idx, uint64_t *rank);
/* Device time (ms) of the fragment plan and of the quantification kernel. */
adsr_status adsr_quantified_timings(const adsr_quantified *q, double *plan_ms,
                                    double *kernel_ms);
void adsr_quantified_free(adsr_quantified *q);

/* ---- per-stage device kernels (SimdBackend ops, src/simd/mod.rs:11-39) --
 * Host in/out buffers; the library stages them through HBM.               */
/* GaussianKernel::new (src/kernel/mod.rs:13-30); returns the length. */
size_t adsr_gaussian_kernel(float fwhm_rt, float sigma_scale_rt,
                            uint64_t kernel_width, float rt_resolution,
                            float *out, size_t cap);
/* gaussian_kernel_1d (src/kernel/mod.rs:32-56) for an explicit sigma in
 * cycles, as used by the sigma_per_window extension.                      */
size_t adsr_gaussian_kernel_sigma(float sigma, uint64_t kernel_width,
                                  float *out, size_t cap);
/* convolution (src/convolution/mod.rs:10-12 -> scalar.rs:5-56) */
adsr_status adsr_convolution(const float *kernel, size_t kernel_len,
                             const float *xic, size_t rows, size_t cols,
                             float *out);
/* axis_log_dot_product (src/score/mod.rs:15-17 -> scalar.rs:4-25) */
adsr_status adsr_axis_log_dot_product(const float *array, size_t rows,
                                      size_t cols, const float *weights,
                                      float *out);
/* find_local_maxima (src/peak_group_selection/mod.rs:23-64); values sorted
 * descending, ties in ascending index order.                              */
adsr_status adsr_find_local_maxima(const float *array, size_t n,
                                   uint64_t offset, uint64_t *indices,
                                   float *values, size_t *n_out);
/* DenseXICObservation::new (src/dense_xic_observation/mod.rs:30-64) on the
 * device; out is [nf, cycle_stop - cycle_start] row-major.                */
adsr_status adsr_dense_xic(const adsr_dia *d, float precursor_mz,
                           uint64_t cycle_start, uint64_t cycle_stop,
                           float mass_tolerance, const float *fragment_mz,
                           size_t nf, float *out);
/* Device ln(x) as used by the score; for exhaustive checks against libm. */
adsr_status adsr_device_logf(const float *x, size_t n, float *out);
/* The same ln evaluated by the host CPU (identical code and constants). */
float adsr_logf_host(float x);

#ifdef __cplusplus
}
#endif
#endif /* ADSR_H */
