/*
 * adsr_oracle.c -- CPU restatement of alphadia-search-rs candidate selection
 * (PeakGroupSelection.search).  TEST INFRASTRUCTURE: the checker for the HIP
 * product path and the timed CPU baseline; never linked into the product.
 *
 * Reference: MannLabs/alphadia-search-rs 1.0.1-dev0.  Each function names the
 * reference file:line it restates.  Semantics that depend on third-party code
 * absent from the reference tree are restated from the published algorithms:
 *   - Rust 1.88 core::slice::binary_search_by (branchless; Ok() returns the
 *     LAST of several equal elements),
 *   - Rust 1.88 core::slice::select::partition_at_index (min/max swap for
 *     index 0 / len-1, then introselect: insertion sort for len <= 16,
 *     recursive median-of-3 pivots, the branchless cyclic Lomuto partition
 *     of core::slice::sort::unstable::quicksort, median of medians after
 *     16 rounds),
 *   - ndarray 0.16.1 numeric_util::unrolled_fold (8 partial sums),
 *   - glibc logf/expf, called directly.
 * Build with -ffp-contract=off and without fast-math.
 */
#include "adsr_oracle.h"

/* ORC_MUTANT selects a planted mistake for the tests that show the
 * oracle's pinning checks catch it (tests/test_oracle_mutants.py); 0 is
 * the oracle itself. */
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ===================================================================== */
/* m/z grid -- src/mz_index/mod.rs                                        */
/* ===================================================================== */

/* src/mz_index/mod.rs:8-25  ppm_index(resolution_ppm, mz_start, mz_end) */
size_t orc_ppm_index(float resolution_ppm, float mz_start, float mz_end,
                     float *out, size_t cap) {
  float m = fmaxf(mz_start, 50.0f); /* :9 mz_start.max(50.0) */
  size_t n = 0;
  if (out && n < cap) out[n] = m;
  n++;
  while (m < mz_end) {                 /* :19 */
    float step = m * (resolution_ppm / 1e6f); /* :20 */
    m = m + step;
    if (out && n < cap) out[n] = m;
    n++;
  }
  return n;
}

static float *g_grid = NULL;
static size_t g_grid_n = 0;
static pthread_once_t g_grid_once = PTHREAD_ONCE_INIT;

static void grid_init(void) {
  /* src/mz_index/mod.rs:4-6,27,41-45: RESOLUTION_PPM=1, MZ_START=150,
   * MZ_END=2000, one lazily built global instance. */
  size_t n = orc_ppm_index(1.0f, 150.0f, 2000.0f, NULL, 0);
  g_grid = (float *)malloc(n * sizeof(float));
  orc_ppm_index(1.0f, 150.0f, 2000.0f, g_grid, n);
  g_grid_n = n;
}

const float *orc_global_grid(size_t *n) {
  pthread_once(&g_grid_once, grid_init);
  if (n) *n = g_grid_n;
  return g_grid;
}

/* src/mz_index/mod.rs:55-90  find_closest_index */
size_t orc_find_closest_index(const float *grid, size_t n, float mz) {
  size_t left = 0, right = n;
  while (left < right) {
    size_t mid = left + (right - left) / 2;
    if (grid[mid] == mz) return mid;
    if (grid[mid] < mz)
      left = mid + 1;
    else
      right = mid;
  }
  if (left == 0) return 0;
  if (left == n) return n - 1;
  float left_diff = fabsf(grid[left] - mz);
  float right_diff = fabsf(grid[left - 1] - mz);
  return (left_diff < right_diff) ? left : left - 1;
}

/* src/mz_index/mod.rs:95-127  mz_range_indices -> [start, end) */
void orc_mz_range_indices(const float *grid, size_t n, float lower_mz,
                          float upper_mz, size_t *start, size_t *end) {
  size_t s = 0, right = n;
  while (s < right) {
    size_t mid = s + (right - s) / 2;
    if (grid[mid] < lower_mz)
      s = mid + 1;
    else
      right = mid;
  }
  size_t e = s;
  if (s < n) {
    while (e < n && grid[e] <= upper_mz) e++;
  }
  *start = s;
  *end = e;
}

/* ===================================================================== */
/* Rust 1.88 core::slice::binary_search_by.  Used by                      */
/* src/rt_index/mod.rs:254-265 and src/quadrupole_observation/mod.rs:92.  */
/* base moves to mid whenever v[mid] <= key, so Ok() is the last equal.    */
/* ===================================================================== */
int orc_rust_bsearch_f32(const float *v, size_t n, float key, size_t *idx) {
  if (n == 0) {
    *idx = 0;
    return 0;
  }
  size_t size = n, base = 0;
  while (size > 1) {
    size_t half = size / 2;
    size_t mid = base + half;
    if (!(v[mid] > key)) base = mid; /* cmp != Greater */
    size -= half;
  }
  if (v[base] == key) {
    *idx = base;
    return 1;
  }
  *idx = base + (v[base] < key ? 1 : 0);
  return 0;
}

int orc_rust_bsearch_u16(const uint16_t *v, size_t n, uint16_t key,
                         size_t *idx) {
  if (n == 0) {
    *idx = 0;
    return 0;
  }
  size_t size = n, base = 0;
  while (size > 1) {
    size_t half = size / 2;
    size_t mid = base + half;
#if ORC_MUTANT == 5 /* planted: lower bound (first equal) instead of last */
    if (v[mid] < key) base = mid;
#else
    if (!(v[mid] > key)) base = mid;
#endif
    size -= half;
  }
  if (v[base] == key) {
    *idx = base;
    return 1;
  }
  *idx = base + (v[base] < key ? 1 : 0);
  return 0;
}

/* ===================================================================== */
/* RT index -- src/rt_index/mod.rs                                        */
/* ===================================================================== */

/* src/rt_index/mod.rs:20-26 calculate_mean_delta_t */
float orc_mean_delta_t(const float *rt, size_t n) {
  if (n > 1) return (rt[n - 1] - rt[0]) / (float)(n - 1);
  return 1.0f;
}

/* src/rt_index/mod.rs:89-138 get_cycle_idx_limits (MIN_CYCLES=10, :5) */
void orc_cycle_idx_limits(const float *rt, size_t n, float delta_t,
                          float precursor_rt, float rt_tolerance,
                          size_t num_cycles_padding, size_t *lower,
                          size_t *upper) {
  if (n == 0) {
    *lower = 0;
    *upper = 0;
    return;
  }
  float min_tol = (10.0f + (float)num_cycles_padding) * delta_t;
  float eff = fmaxf(rt_tolerance, min_tol);
  float lower_rt = precursor_rt - eff;
  float upper_rt = precursor_rt + eff;
  if (upper_rt < rt[0]) {
    *lower = 0;
    *upper = 0;
    return;
  }
  if (lower_rt > rt[n - 1]) {
    *lower = n;
    *upper = n;
    return;
  }
  size_t lo, hi;
  if (lower_rt <= rt[0])
    lo = 0;
  else
    orc_rust_bsearch_f32(rt, n, lower_rt, &lo);
  if (upper_rt >= rt[n - 1])
    hi = n;
  else
    orc_rust_bsearch_f32(rt, n, upper_rt, &hi);
  *lower = lo;
  *upper = hi;
}

/* ===================================================================== */
/* DIA data builder -- src/dia_data_builder/mod.rs                        */
/* ===================================================================== */

typedef struct build_ctx {
  const int64_t *delta, *pstart, *pstop, *cyc;
  const float *iso_lo, *iso_hi, *peak_mz, *peak_int;
  size_t n_spectra;
  orc_dia *d;
  size_t next;
  pthread_mutex_t mu;
} build_ctx;

static int cmp_i64(const void *a, const void *b) {
  int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
  return (x > y) - (x < y);
}

/* src/dia_data_builder/mod.rs:52-139 build_single_observation_rayon */
static void build_one_obs(build_ctx *c, int64_t target) {
  orc_obs *o = &c->d->obs[target];
  size_t B = c->d->n_grid;
  const float *grid = c->d->grid;
  size_t nm = 0;
  for (size_t s = 0; s < c->n_spectra; s++)
    if (c->delta[s] == target) nm++;
  o->slice_starts = (uint32_t *)calloc(B + 1, sizeof(uint32_t));
  if (nm == 0) { /* :71-80 empty observation, window [0,0] */
    o->isolation_window[0] = 0.0f;
    o->isolation_window[1] = 0.0f;
    o->num_cycles = 0;
    o->cycle_indices = NULL;
    o->intensities = NULL;
    o->n_entries = 0;
    return;
  }
  size_t *match = (size_t *)malloc(nm * sizeof(size_t));
  nm = 0;
  for (size_t s = 0; s < c->n_spectra; s++)
    if (c->delta[s] == target) match[nm++] = s;
  /* :83-85 window of the first matching spectrum */
  o->isolation_window[0] = c->iso_lo[match[0]];
  o->isolation_window[1] = c->iso_hi[match[0]];
  /* :87-91 number of unique cycles */
  int64_t *cy = (int64_t *)malloc(nm * sizeof(int64_t));
  for (size_t i = 0; i < nm; i++) cy[i] = c->cyc[match[i]];
  qsort(cy, nm, sizeof(int64_t), cmp_i64);
  size_t uniq = 0;
  for (size_t i = 0; i < nm; i++)
    if (i == 0 || cy[i] != cy[i - 1]) uniq++;
  o->num_cycles = uniq;
  free(cy);
  /* :94-109 collect (bin, cycle as u16, intensity) in spectrum/peak order */
  size_t total = 0;
  for (size_t i = 0; i < nm; i++)
    total += (size_t)(c->pstop[match[i]] - c->pstart[match[i]]);
  uint32_t *bin = (uint32_t *)malloc((total ? total : 1) * sizeof(uint32_t));
  uint16_t *cyc = (uint16_t *)malloc((total ? total : 1) * sizeof(uint16_t));
  float *inten = (float *)malloc((total ? total : 1) * sizeof(float));
  size_t k = 0;
  for (size_t i = 0; i < nm; i++) {
    size_t s = match[i];
    uint16_t cc = (uint16_t)c->cyc[s]; /* `as u16` truncation, :98 */
    for (int64_t p = c->pstart[s]; p < c->pstop[s]; p++) {
      bin[k] = (uint32_t)orc_find_closest_index(grid, B, c->peak_mz[p]);
      cyc[k] = cc;
      inten[k] = c->peak_int[p];
      k++;
    }
  }
  free(match);
  /* :113 stable sort by (bin, cycle): two stable counting passes (LSD). */
  size_t *ord = (size_t *)malloc((total ? total : 1) * sizeof(size_t));
  size_t *tmp = (size_t *)malloc((total ? total : 1) * sizeof(size_t));
  size_t *cnt = (size_t *)calloc(65537, sizeof(size_t));
  for (size_t i = 0; i < total; i++) cnt[cyc[i] + 1]++;
  for (size_t i = 0; i < 65536; i++) cnt[i + 1] += cnt[i];
  for (size_t i = 0; i < total; i++) tmp[cnt[cyc[i]]++] = i;
  free(cnt);
  size_t *bcnt = (size_t *)calloc(B + 1, sizeof(size_t));
  for (size_t i = 0; i < total; i++) bcnt[bin[i] + 1]++;
  for (size_t b = 0; b < B; b++) bcnt[b + 1] += bcnt[b];
  /* :123-136 slice_starts[b+1] = number of entries with bin <= b */
  for (size_t b = 0; b <= B; b++) o->slice_starts[b] = (uint32_t)bcnt[b];
  for (size_t i = 0; i < total; i++) {
    size_t e = tmp[i];
    ord[bcnt[bin[e]]++] = e;
  }
  free(bcnt);
  free(tmp);
  o->cycle_indices = (uint16_t *)malloc((total ? total : 1) * sizeof(uint16_t));
  o->intensities = (float *)malloc((total ? total : 1) * sizeof(float));
  for (size_t i = 0; i < total; i++) {
    o->cycle_indices[i] = cyc[ord[i]];
    o->intensities[i] = inten[ord[i]];
  }
  o->n_entries = total;
  free(ord);
  free(bin);
  free(cyc);
  free(inten);
}

static void *build_worker(void *arg) {
  build_ctx *c = (build_ctx *)arg;
  for (;;) {
    pthread_mutex_lock(&c->mu);
    size_t t = c->next++;
    pthread_mutex_unlock(&c->mu);
    if (t >= c->d->n_obs) break;
    build_one_obs(c, (int64_t)t);
  }
  return NULL;
}

/* src/dia_data_builder/mod.rs:11-49 from_alpha_raw and
 * src/rt_index/mod.rs:53-75 RTIndex::from_alpha_raw */
orc_dia *orc_dia_build(const int64_t *delta, const float *iso_lo,
                       const float *iso_hi, const int64_t *pstart,
                       const int64_t *pstop, const int64_t *cyc,
                       const float *srt, size_t n_spectra,
                       const float *peak_mz, const float *peak_int,
                       size_t n_peaks, int nthreads) {
  (void)n_peaks;
  orc_dia *d = (orc_dia *)calloc(1, sizeof(orc_dia));
  d->grid = orc_global_grid(&d->n_grid);
  size_t nrt = 0;
  for (size_t s = 0; s < n_spectra; s++)
    if (delta[s] == 0) nrt++;
  d->rt = (float *)malloc((nrt ? nrt : 1) * sizeof(float));
  d->n_rt = 0;
  for (size_t s = 0; s < n_spectra; s++)
    if (delta[s] == 0) d->rt[d->n_rt++] = srt[s];
  d->delta_t = orc_mean_delta_t(d->rt, d->n_rt);
  int64_t maxd = 0; /* :33-38 max().unwrap_or(0) */
  for (size_t s = 0; s < n_spectra; s++)
    if (s == 0 || delta[s] > maxd) maxd = delta[s];
  d->n_obs = (maxd < 0) ? 0 : (size_t)maxd + 1;
  d->obs = (orc_obs *)calloc(d->n_obs ? d->n_obs : 1, sizeof(orc_obs));
  build_ctx c;
  memset(&c, 0, sizeof(c));
  c.delta = delta;
  c.pstart = pstart;
  c.pstop = pstop;
  c.cyc = cyc;
  c.iso_lo = iso_lo;
  c.iso_hi = iso_hi;
  c.peak_mz = peak_mz;
  c.peak_int = peak_int;
  c.n_spectra = n_spectra;
  c.d = d;
  pthread_mutex_init(&c.mu, NULL);
  if (nthreads < 1) nthreads = 1;
  pthread_t th[256];
  if (nthreads > 256) nthreads = 256;
  for (int t = 0; t < nthreads; t++)
    pthread_create(&th[t], NULL, build_worker, &c);
  for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  pthread_mutex_destroy(&c.mu);
  return d;
}

void orc_dia_free(orc_dia *d) {
  if (!d) return;
  for (size_t i = 0; i < d->n_obs; i++) {
    free(d->obs[i].slice_starts);
    free(d->obs[i].cycle_indices);
    free(d->obs[i].intensities);
  }
  free(d->obs);
  free(d->rt);
  free(d);
}

/* src/dia_data/mod.rs:79-87 get_valid_observations (inclusive bounds) */
size_t orc_valid_observations(const orc_dia *d, float mz, size_t *out) {
  size_t n = 0;
  for (size_t i = 0; i < d->n_obs; i++)
    if (d->obs[i].isolation_window[0] <= mz &&
        d->obs[i].isolation_window[1] >= mz)
      out[n++] = i;
  return n;
}

/* src/quadrupole_observation/mod.rs:70-107 fill_xic_slice, on raw CSR arrays
 * (slice_starts has one entry per grid bin plus one). */
void orc_fill_xic_slice_raw(const float *grid, size_t n_grid,
                            const uint32_t *slice_starts,
                            const uint16_t *cycle_indices,
                            const float *intensities, float *dense_xic,
                            size_t c0, size_t c1, float mass_tolerance,
                            float mz) {
  float delta_mz = mz * mass_tolerance * 1e-6f; /* :79 (mz*tol)*1e-6 */
  float lower_mz = mz - delta_mz;
  float upper_mz = mz + delta_mz;
  size_t b0, b1;
  orc_mz_range_indices(grid, n_grid, lower_mz, upper_mz, &b0, &b1);
  for (size_t b = b0; b < b1; b++) {
    size_t start = slice_starts[b];
    size_t stop = slice_starts[b + 1];
    const uint16_t *ci = cycle_indices + start;
    const float *in = intensities + start;
    size_t len = stop - start;
    size_t pos;
    orc_rust_bsearch_u16(ci, len, (uint16_t)c0, &pos); /* :92-94 */
    for (size_t i = pos; i < len; i++) {
      size_t cyc = ci[i];
      if (cyc >= c1) break; /* :100-102 */
      dense_xic[cyc - c0] += in[i];
    }
  }
}

void orc_fill_xic_slice(const orc_dia *d, size_t obs_idx, float *dense_xic,
                        size_t c0, size_t c1, float mass_tolerance, float mz) {
  const orc_obs *o = &d->obs[obs_idx];
  orc_fill_xic_slice_raw(d->grid, d->n_grid, o->slice_starts, o->cycle_indices,
                         o->intensities, dense_xic, c0, c1, mass_tolerance,
                         mz);
}

/* src/dense_xic_observation/mod.rs:30-64 DenseXICObservation::new */
void orc_dense_xic(const orc_dia *d, float precursor_mz, size_t c0, size_t c1,
                   float mass_tolerance, const float *fragment_mz, size_t nf,
                   float *out) {
  size_t L = c1 - c0;
  memset(out, 0, nf * L * sizeof(float));
  size_t valid[1024];
  size_t nv = 0;
  for (size_t i = 0; i < d->n_obs && nv < 1024; i++)
    if (d->obs[i].isolation_window[0] <= precursor_mz &&
        d->obs[i].isolation_window[1] >= precursor_mz)
      valid[nv++] = i;
  for (size_t v = 0; v < nv; v++)
    for (size_t f = 0; f < nf; f++)
      orc_fill_xic_slice(d, valid[v], out + f * L, c0, c1, mass_tolerance,
                         fragment_mz[f]);
}

/* ===================================================================== */
/* Gaussian kernel -- src/kernel/mod.rs                                   */
/* ===================================================================== */

/* ndarray 0.16.1 ArrayBase::sum -> numeric_util::unrolled_fold: eight
 * partial sums over whole chunks of 8, combined as
 * acc += (p0+p4); acc += (p1+p5); acc += (p2+p6); acc += (p3+p7),
 * then the remaining (< 8) elements in order.  Called at src/kernel/mod.rs:52. */
float orc_ndarray_sum(const float *x, size_t n) {
  float p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  size_t i = 0;
  while (n - i >= 8) {
    for (int j = 0; j < 8; j++) p[j] = p[j] + x[i + j];
    i += 8;
  }
  float acc = 0.0f;
  acc = acc + (p[0] + p[4]);
  acc = acc + (p[1] + p[5]);
  acc = acc + (p[2] + p[6]);
  acc = acc + (p[3] + p[7]);
  for (; i < n; i++) acc = acc + x[i];
  return acc;
}

/* src/kernel/mod.rs:32-56 gaussian_kernel_1d(size, sigma) */
size_t orc_gaussian_kernel_sigma(float sigma, size_t size, float *out,
                                 size_t cap) {
  if (size > cap) return 0;
  int half = (int)(size / 2);
  const float PI_F = 3.14159265358979323846f; /* std::f32::consts::PI */
  float normalization = 1.0f / (sigma * sqrtf(2.0f * PI_F));
  for (int x = -half; x < half; x++) {
    float xf = (float)x;
#if ORC_MUTANT == 6 /* planted: the 1/2 of the exponent dropped */
    float exponent = (-1.0f * (xf * xf)) / (sigma * sigma);
#else
    float exponent = (-0.5f * (xf * xf)) / (sigma * sigma);
#endif
    out[x + half] = normalization * expf(exponent);
  }
  float sum = orc_ndarray_sum(out, size);
  for (size_t i = 0; i < size; i++) out[i] = out[i] / sum;
  return size;
}

/* src/kernel/mod.rs:13-30 GaussianKernel::new */
size_t orc_gaussian_kernel(float fwhm_rt, float sigma_scale_rt,
                           size_t kernel_width, float rt_resolution,
                           float *out, size_t cap) {
  size_t w = (size_t)ceilf((float)kernel_width / 2.0f) * 2;
  float sigma = fwhm_rt / 2.3548f;
  float rt_sigma = sigma * sigma_scale_rt / rt_resolution;
  return orc_gaussian_kernel_sigma(rt_sigma, w, out, cap);
}

/* ===================================================================== */
/* Convolution / score / maxima                                           */
/* ===================================================================== */

/* src/convolution/scalar.rs:5-56 convolution_scalar */
void orc_convolution(const float *k, size_t klen, const float *xic,
                     size_t rows, size_t cols, float *out) {
  size_t half = klen / 2;
  memset(out, 0, rows * cols * sizeof(float));
  if (rows == 0 || cols == 0 || cols < klen) return;
  size_t start = half;
  size_t end = cols >= half ? cols - half : 0; /* saturating_sub */
  for (size_t f = 0; f < rows; f++) {
    const float *x = xic + f * cols;
    float *y = out + f * cols;
    if (start < end) {
      for (size_t i = start; i < end; i++) {
        float sum = 0.0f;
#if ORC_MUTANT != 2 /* planted (2): the centre term dropped */
        sum = sum + x[i] * k[half];
#endif
        for (size_t kk = 0; kk < half; kk++) {
          float kv = k[kk];
          float l = x[i - (half - kk)];
          float r = x[i + (half - kk)];
#if ORC_MUTANT == 1 /* planted: tap index off by one */
          sum = sum + (l + r) * k[kk + 1];
#else
          sum = sum + (l + r) * kv;
#endif
        }
        y[i] = sum;
      }
    }
  }
}

/* glibc logf; f32::ln lowers to it on x86-64 Linux (src/score/scalar.rs:19) */
float orc_logf(float x) { return logf(x); }

/* Checker of a logf implementation over a contiguous range of inputs: y[i]
 * is the candidate's logf of the float with bits first + i.  Counts the i
 * where y[i] differs from glibc logf bit for bit (two NaNs count as equal:
 * the payload of a NaN result is not part of the reference's output), on
 * nthreads threads, and stores the smallest such i in *first_bad (n when
 * there is none). */
typedef struct {
  uint32_t first;
  size_t lo, hi;
  const float *y;
  uint64_t bad;
  size_t first_bad;
} LogfCheckPart;

static void *logf_check_worker(void *arg) {
  LogfCheckPart *p = (LogfCheckPart *)arg;
  p->bad = 0;
  p->first_bad = (size_t)-1;
  for (size_t i = p->lo; i < p->hi; i++) {
    uint32_t b = p->first + (uint32_t)i;
    float x;
    memcpy(&x, &b, 4);
    float want = logf(x), got = p->y[i];
    uint32_t wb, gb;
    memcpy(&wb, &want, 4);
    memcpy(&gb, &got, 4);
    if (wb != gb && !(want != want && got != got)) {
      if (p->bad == 0) p->first_bad = i;
      p->bad++;
    }
  }
  return NULL;
}

uint64_t orc_logf_check_range(uint32_t first, size_t n, const float *y,
                              int nthreads, size_t *first_bad) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  LogfCheckPart parts[256];
  pthread_t th[256];
  for (int t = 0; t < nthreads; t++) {
    parts[t].first = first;
    parts[t].lo = n * (size_t)t / (size_t)nthreads;
    parts[t].hi = n * (size_t)(t + 1) / (size_t)nthreads;
    parts[t].y = y;
    pthread_create(&th[t], NULL, logf_check_worker, &parts[t]);
  }
  uint64_t bad = 0;
  size_t fb = n;
  for (int t = 0; t < nthreads; t++) {
    pthread_join(th[t], NULL);
    bad += parts[t].bad;
    if (parts[t].bad && parts[t].first_bad < fb) fb = parts[t].first_bad;
  }
  if (first_bad) *first_bad = fb;
  return bad;
}

/* src/score/scalar.rs:4-25 axis_log_dot_product_scalar */
void orc_axis_log_dot_product(const float *a, size_t rows, size_t cols,
                              const float *w, float *out) {
  for (size_t j = 0; j < cols; j++) out[j] = 0.0f;
  for (size_t i = 0; i < rows; i++)
    for (size_t j = 0; j < cols; j++) {
#if ORC_MUTANT == 3 /* planted: ln(x) instead of ln(x + 1) */
      float val = logf(a[i * cols + j]);
#else
      float val = logf(a[i * cols + j] + 1.0f);
#endif
      out[j] = out[j] + val * w[i];
    }
}

/* src/peak_group_selection/mod.rs:23-64 find_local_maxima.  The stable
 * descending sort keeps ascending index order among equal values.         */
size_t orc_find_local_maxima(const float *a, size_t len, size_t offset,
                             size_t *indices, float *values) {
  size_t n = 0;
  if (len < 5) return 0;
  for (size_t i = 2; i < len - 2; i++) {
#if ORC_MUTANT == 4 /* planted: non-strict rise into the peak */
    if (a[i - 2] < a[i - 1] && a[i - 1] <= a[i] && a[i] > a[i + 1] &&
        a[i + 1] > a[i + 2]) {
#else
    if (a[i - 2] < a[i - 1] && a[i - 1] < a[i] && a[i] > a[i + 1] &&
        a[i + 1] > a[i + 2]) {
#endif
      indices[n] = i + offset;
      values[n] = a[i];
      n++;
    }
  }
  /* stable insertion sort, descending by value */
  for (size_t i = 1; i < n; i++) {
    size_t ii = indices[i];
    float vv = values[i];
    size_t j = i;
    while (j > 0 && vv > values[j - 1]) {
      indices[j] = indices[j - 1];
      values[j] = values[j - 1];
      j--;
    }
    indices[j] = ii;
    values[j] = vv;
  }
  return n;
}

/* ===================================================================== */
/* Fragment filtering -- src/speclib_flat/mod.rs:240-345                  */
/* ===================================================================== */

/* is_less for select_nth_unstable_by(k, |a,b| b.I.partial_cmp(&a.I)):
 * a < b  <=>  I[b] < I[a]  (descending intensity; NaN compares Equal).   */
static int frag_less(const float *I, size_t a, size_t b) { return I[b] < I[a]; }

/* Rust 1.88 core::slice::select (select_nth_unstable_by) on an array of
 * fragment indices ordered by frag_less.  `le` selects the comparison
 * !is_less(b, a) that partition_at_index_loop uses for the elements equal
 * to an ancestor pivot. */
typedef struct {
  const float *I;
  int le;
} sel_cmp;

static int sel_lt(const sel_cmp *c, size_t a, size_t b) {
  return c->le ? !frag_less(c->I, b, a) : frag_less(c->I, a, b);
}

static void swap_sz(size_t *v, size_t a, size_t b) {
  size_t t = v[a];
  v[a] = v[b];
  v[b] = t;
}

/* core::slice::sort::shared::smallsort::insertion_sort_shift_left(v, 1) */
static void sel_insertion_sort(size_t *v, size_t n, const float *I) {
  for (size_t i = 1; i < n; i++) {
    size_t tmp = v[i], j = i;
    while (j > 0 && frag_less(I, tmp, v[j - 1])) {
      v[j] = v[j - 1];
      j--;
    }
    v[j] = tmp;
  }
}

/* core::slice::sort::shared::pivot::median3: the position of the median */
static size_t sel_median3(const size_t *v, const float *I, size_t a, size_t b,
                          size_t c) {
  int x = frag_less(I, v[a], v[b]);
  int y = frag_less(I, v[a], v[c]);
  if (x == y) {
    int z = frag_less(I, v[b], v[c]);
    return (z ^ x) ? c : b;
  }
  return a;
}

/* pivot::median3_rec over three runs of n elements starting at a, b, c */
static size_t sel_median3_rec(const size_t *v, const float *I, size_t a,
                              size_t b, size_t c, size_t n) {
  if (n * 8 >= 64) {
    size_t n8 = n / 8;
    a = sel_median3_rec(v, I, a, a + n8 * 4, a + n8 * 7, n8);
    b = sel_median3_rec(v, I, b, b + n8 * 4, b + n8 * 7, n8);
    c = sel_median3_rec(v, I, c, c + n8 * 4, c + n8 * 7, n8);
  }
  return sel_median3(v, I, a, b, c);
}

/* pivot::choose_pivot (len >= 8) */
static size_t sel_choose_pivot(const size_t *v, size_t len, const float *I) {
  size_t d8 = len / 8;
  size_t a = 0, b = d8 * 4, c = d8 * 7;
  if (len < 64) return sel_median3(v, I, a, b, c);
  return sel_median3_rec(v, I, a, b, c, d8);
}

/* quicksort::partition with partition_lomuto_branchless_cyclic: the pivot
 * is swapped to the front, the rest is partitioned by a cyclic permutation
 * that lifts the first element out and processes it last, then the pivot
 * is swapped to its place.  Returns the number of elements less than it. */
static size_t sel_partition(size_t *v, size_t len, size_t pivot,
                            const sel_cmp *c) {
  if (len == 0) return 0;
  swap_sz(v, 0, pivot);
  const size_t pv = v[0];
  size_t *r = v + 1;
  const size_t n = len - 1;
  size_t num_lt = 0;
  if (n > 0) {
    const size_t gap_value = r[0];
    size_t gap = 0;
    for (size_t right = 1; right <= n; right++) {
      const int last = right == n;
      const size_t elem = last ? gap_value : r[right];
      const int lt = sel_lt(c, elem, pv);
      r[gap] = r[num_lt];
      r[num_lt] = elem;
      gap = right;
      num_lt += (size_t)lt;
    }
  }
  swap_sz(v, 0, num_lt);
  return num_lt;
}

static size_t sel_max_index(const size_t *v, size_t n, const float *I) {
  size_t m = 0;
  for (size_t t = 1; t < n; t++)
    if (frag_less(I, v[m], v[t])) m = t;
  return m;
}

static size_t sel_min_index(const size_t *v, size_t n, const float *I) {
  size_t m = 0;
  for (size_t t = 1; t < n; t++)
    if (frag_less(I, v[t], v[m])) m = t;
  return m;
}

/* select::median_idx, ninther, median_of_ninthers, median_of_medians */
static size_t sel_median_idx(const size_t *v, const float *I, size_t a, size_t b,
                             size_t c) {
  if (frag_less(I, v[c], v[a])) {
    size_t t = a;
    a = c;
    c = t;
  }
  if (frag_less(I, v[c], v[b])) return c;
  if (frag_less(I, v[b], v[a])) return a;
  return b;
}

static void sel_ninther(size_t *v, const float *I, size_t a, size_t b, size_t c,
                        size_t d, size_t e, size_t f, size_t g, size_t h,
                        size_t i) {
  b = sel_median_idx(v, I, a, b, c);
  h = sel_median_idx(v, I, g, h, i);
  if (frag_less(I, v[h], v[f])) {
    size_t t = f;
    f = h;
    h = t;
  }
  if (frag_less(I, v[f], v[d])) {
    size_t t = d;
    d = f;
    f = t;
    if (frag_less(I, v[h], v[f])) f = h;
    if (frag_less(I, v[f], v[b])) f = b;
  }
  swap_sz(v, f, e);
}

static void sel_median_of_medians(size_t *v, size_t len, size_t k, const float *I);

static size_t sel_median_of_ninthers(size_t *v, size_t len, const float *I) {
  size_t frac = len <= 1024 ? len / 12 : len <= 128 * 1024 ? len / 64 : len / 1024;
  size_t pivot = frac / 2;
  size_t lo = len / 2 - pivot;
  size_t hi = frac + lo;
  size_t gap = (len - 9 * frac) / 4;
  size_t a = lo - 4 * frac - gap;
  size_t b = hi + gap;
  for (size_t i = lo; i < hi; i++) {
    sel_ninther(v, I, a, i - frac, b, a + 1, i, b + 1, a + 2, i + frac, b + 2);
    a += 3;
    b += 3;
  }
  sel_median_of_medians(v + lo, frac, pivot, I);
  sel_cmp c = {I, 0};
  return sel_partition(v, len, lo + pivot, &c);
}

static void sel_median_of_medians(size_t *v, size_t len, size_t k, const float *I) {
  for (;;) {
    if (len <= 16) {
      if (len >= 2) sel_insertion_sort(v, len, I);
      return;
    }
    if (k == len - 1) {
      swap_sz(v, sel_max_index(v, len, I), k);
      return;
    }
    if (k == 0) {
      swap_sz(v, sel_min_index(v, len, I), 0);
      return;
    }
    size_t p = sel_median_of_ninthers(v, len, I);
    if (p == k) return;
    if (p > k) {
      len = p;
    } else {
      v += p + 1;
      len -= p + 1;
      k -= p + 1;
    }
  }
}

/* select::partition_at_index for 0 < index < len - 1:
 * partition_at_index_loop with its limit of 16 rounds. */
static void sel_partition_at_index_loop(size_t *v, size_t len, size_t index,
                                        const float *I) {
  int limit = 16;
  int have_ancestor = 0;
  size_t ancestor = 0;
  for (;;) {
    if (len <= 16) {
      if (len >= 2) sel_insertion_sort(v, len, I);
      return;
    }
    if (limit == 0) {
      sel_median_of_medians(v, len, index, I);
      return;
    }
    limit--;
    size_t pivot_pos = sel_choose_pivot(v, len, I);
    if (have_ancestor && !frag_less(I, ancestor, v[pivot_pos])) {
      sel_cmp c = {I, 1};
      size_t mid = sel_partition(v, len, pivot_pos, &c) + 1;
      if (mid > index) return;
      v += mid;
      len -= mid;
      index -= mid;
      have_ancestor = 0;
      continue;
    }
    sel_cmp c = {I, 0};
    size_t mid = sel_partition(v, len, pivot_pos, &c);
    if (mid == index) return;
    if (mid < index) {
      ancestor = v[mid];
      have_ancestor = 1;
      v += mid + 1;
      len -= mid + 1;
      index -= mid + 1;
    } else {
      len = mid;
    }
  }
}

/* Returns the number of selected fragments; out_order[j] is the original
 * (within-precursor) index of the j-th selected fragment in output order.
 * top_k == SIZE_MAX means no limit.                                      */
size_t orc_filter_sort_fragments(size_t n, const float *mz, const float *I,
                                 const uint8_t *number, const uint8_t *type,
                                 int non_zero, int filter_y1, size_t top_k,
                                 size_t *v) {
  size_t len = 0;
  for (size_t i = 0; i < n; i++) {
    if (non_zero && !(I[i] > 0.0f)) continue;           /* :273-275 */
    if (filter_y1 && type[i] == 121 && number[i] == 1) continue; /* :278-282 */
    v[len++] = i;
  }
  size_t k = top_k < len ? top_k : len; /* :285 */
  if (k < len) {                        /* :286-292 */
    if (k == len - 1) {
      /* max_index: first element with the smallest intensity */
      size_t mi = 0;
      for (size_t t = 1; t < len; t++)
        if (frag_less(I, v[mi], v[t])) mi = t;
      size_t s = v[mi];
      v[mi] = v[k];
      v[k] = s;
    } else if (k == 0) {
      /* min_index swap; the slice is truncated to length 0 below */
    } else {
      sel_partition_at_index_loop(v, len, k, I);
    }
    len = k;
  }
  /* :294-298 stable sort by fragment_mz ascending */
  for (size_t i = 1; i < len; i++) {
    size_t cur = v[i];
    size_t j = i;
    while (j > 0 && mz[cur] < mz[v[j - 1]]) {
      v[j] = v[j - 1];
      j--;
    }
    v[j] = cur;
  }
  return len;
}

/* ===================================================================== */
/* Spectral library -- src/speclib_flat/mod.rs:98-188                     */
/* ===================================================================== */

static const uint64_t *g_sort_keys;
static int cmp_sort_idx(const void *a, const void *b) {
  size_t x = *(const size_t *)a, y = *(const size_t *)b;
  uint64_t kx = g_sort_keys[x], ky = g_sort_keys[y];
  if (kx != ky) return kx < ky ? -1 : 1;
  return x < y ? -1 : (x > y); /* stable: original position breaks ties */
}
static pthread_mutex_t g_sort_mu = PTHREAD_MUTEX_INITIALIZER;

orc_lib *orc_lib_build(size_t np, const uint64_t *pidx, const float *pmz,
                       const float *prt, const uint64_t *fstart,
                       const uint64_t *fstop, size_t nfrag, const float *fmz,
                       const float *fint, const uint8_t *fnum,
                       const uint8_t *ftype) {
  return orc_lib_build_full(np, pidx, pmz, prt, NULL, fstart, fstop, nfrag,
                            fmz, NULL, fint, fnum, ftype);
}

orc_lib *orc_lib_build_full(size_t np, const uint64_t *pidx, const float *pmz,
                            const float *prt, const uint8_t *naa,
                            const uint64_t *fstart, const uint64_t *fstop,
                            size_t nfrag, const float *fmz, const float *fmzl,
                            const float *fint, const uint8_t *fnum,
                            const uint8_t *ftype) {
  orc_lib *l = (orc_lib *)calloc(1, sizeof(orc_lib));
  l->n_prec = np;
  l->n_frag = nfrag;
  size_t *ord = (size_t *)malloc((np ? np : 1) * sizeof(size_t));
  for (size_t i = 0; i < np; i++) ord[i] = i;
  pthread_mutex_lock(&g_sort_mu); /* :139-142 stable sort_by_key */
  g_sort_keys = pidx;
  qsort(ord, np, sizeof(size_t), cmp_sort_idx);
  pthread_mutex_unlock(&g_sort_mu);
  size_t sz = np ? np : 1;
  l->precursor_idx = (uint64_t *)malloc(sz * 8);
  l->precursor_mz = (float *)malloc(sz * 4);
  l->precursor_rt = (float *)malloc(sz * 4);
  l->frag_start = (uint64_t *)malloc(sz * 8);
  l->frag_stop = (uint64_t *)malloc(sz * 8);
  if (naa) l->precursor_naa = (uint8_t *)malloc(sz);
  for (size_t i = 0; i < np; i++) { /* :145-163 */
    l->precursor_idx[i] = pidx[ord[i]];
    l->precursor_mz[i] = pmz[ord[i]];
    l->precursor_rt[i] = prt[ord[i]];
    l->frag_start[i] = fstart[ord[i]];
    l->frag_stop[i] = fstop[ord[i]];
    if (naa) l->precursor_naa[i] = naa[ord[i]];
  }
  free(ord);
  size_t fz = nfrag ? nfrag : 1; /* fragment arrays keep their order */
  l->fragment_mz = (float *)malloc(fz * 4);
  l->fragment_intensity = (float *)malloc(fz * 4);
  l->fragment_number = (uint8_t *)malloc(fz);
  l->fragment_type = (uint8_t *)malloc(fz);
  if (fmzl) {
    l->fragment_mz_library = (float *)malloc(fz * 4);
    if (nfrag) memcpy(l->fragment_mz_library, fmzl, nfrag * 4);
  }
  if (nfrag) {
    memcpy(l->fragment_mz, fmz, nfrag * 4);
    memcpy(l->fragment_intensity, fint, nfrag * 4);
    memcpy(l->fragment_number, fnum, nfrag);
    memcpy(l->fragment_type, ftype, nfrag);
  }
  return l;
}

void orc_lib_free(orc_lib *l) {
  if (!l) return;
  free(l->precursor_naa);
  free(l->fragment_mz_library);
  free(l->idf_counts);
  free(l->precursor_idx);
  free(l->precursor_mz);
  free(l->precursor_rt);
  free(l->frag_start);
  free(l->frag_stop);
  free(l->fragment_mz);
  free(l->fragment_intensity);
  free(l->fragment_number);
  free(l->fragment_type);
  free(l);
}

/* ===================================================================== */
/* Selection -- src/peak_group_selection/mod.rs:87-195                    */
/* ===================================================================== */

typedef struct scratch {
  size_t cap;
  float *xic, *conv, *score;
  size_t *ord, *midx;
  float *mval, *fmz, *fw;
  size_t ord_cap;
} scratch;

static void scratch_need(scratch *s, size_t cells, size_t L, size_t nfrag) {
  if (cells > s->cap || L > s->cap) {
    size_t c = cells > L ? cells : L;
    c = c * 2 + 64;
    free(s->xic);
    free(s->conv);
    free(s->score);
    free(s->midx);
    free(s->mval);
    s->xic = (float *)malloc(c * 4);
    s->conv = (float *)malloc(c * 4);
    s->score = (float *)malloc(c * 4);
    s->midx = (size_t *)malloc(c * sizeof(size_t));
    s->mval = (float *)malloc(c * 4);
    s->cap = c;
  }
  if (nfrag > s->ord_cap) {
    size_t c = nfrag * 2 + 64;
    free(s->ord);
    free(s->fmz);
    free(s->fw);
    s->ord = (size_t *)malloc(c * sizeof(size_t));
    s->fmz = (float *)malloc(c * 4);
    s->fw = (float *)malloc(c * 4);
    s->ord_cap = c;
  }
}

static void scratch_free(scratch *s) {
  free(s->xic);
  free(s->conv);
  free(s->score);
  free(s->ord);
  free(s->midx);
  free(s->mval);
  free(s->fmz);
  free(s->fw);
}

typedef struct sel_ctx {
  const orc_dia *d;
  const orc_lib *l;
  const orc_params *p;
  float kernel[256];
  size_t klen;
  float *win_kernel; /* n_obs * klen, when sigma_per_window is given */
  size_t row_begin, row_end;
  size_t *counts;    /* per row */
  uint64_t *slot_center;
  float *slot_score;
  size_t next;
  pthread_mutex_t mu;
} sel_ctx;

/* src/peak_group_selection/mod.rs:131-195 search_precursor_generic, with the
 * fragment filter of :104-109.  Returns -1 for None, else #candidates. */
static long select_row(sel_ctx *c, scratch *s, size_t row, size_t *c0o,
                       size_t *c1o, size_t *nfo) {
  const orc_lib *l = c->l;
  const orc_dia *d = c->d;
  const orc_params *p = c->p;
  size_t fs = l->frag_start[row], fe = l->frag_stop[row];
  size_t nraw = fe - fs;
  scratch_need(s, 0, 0, nraw);
  size_t nf = orc_filter_sort_fragments(
      nraw, l->fragment_mz + fs, l->fragment_intensity + fs,
      l->fragment_number + fs, l->fragment_type + fs, 1, 1, p->top_k_fragments,
      s->ord);
  for (size_t j = 0; j < nf; j++) {
    s->fmz[j] = l->fragment_mz[fs + s->ord[j]];
    s->fw[j] = l->fragment_intensity[fs + s->ord[j]];
  }
  float rt_tol = p->rt_tolerance_per_precursor
                     ? p->rt_tolerance_per_precursor[row]
                     : p->rt_tolerance;
  size_t c0, c1;
  orc_cycle_idx_limits(d->rt, d->n_rt, d->delta_t, l->precursor_rt[row],
                       rt_tol, p->kernel_size, &c0, &c1);
  if (c0o) *c0o = c0;
  if (c1o) *c1o = c1;
  if (nfo) *nfo = nf;
  if (c1 <= c0) return -1; /* :146-148 */
  size_t L = c1 - c0;
  scratch_need(s, nf * L, L, nraw);
  float pmz = l->precursor_mz[row];
  orc_dense_xic(d, pmz, c0, c1, p->mass_tolerance, s->fmz, nf, s->xic);
  const float *k = c->kernel;
  if (c->win_kernel) {
    for (size_t o = 0; o < d->n_obs; o++)
      if (d->obs[o].isolation_window[0] <= pmz &&
          d->obs[o].isolation_window[1] >= pmz) {
        k = c->win_kernel + o * c->klen;
        break;
      }
  }
  orc_convolution(k, c->klen, s->xic, nf, L, s->conv);
  orc_axis_log_dot_product(s->conv, nf, L, s->fw, s->score);
  size_t nmax = orc_find_local_maxima(s->score, L, c0, s->midx, s->mval);
  size_t cc = p->candidate_count < nmax ? p->candidate_count : nmax;
  return (long)cc;
}

static void *sel_worker(void *arg) {
  sel_ctx *c = (sel_ctx *)arg;
  scratch s;
  memset(&s, 0, sizeof(s));
  size_t cc = c->p->candidate_count;
  for (;;) {
    pthread_mutex_lock(&c->mu);
    size_t r0 = c->next;
    c->next += 64; /* dynamic chunks of 64 precursors */
    pthread_mutex_unlock(&c->mu);
    if (r0 >= c->row_end) break;
    size_t r1 = r0 + 64 < c->row_end ? r0 + 64 : c->row_end;
    for (size_t row = r0; row < r1; row++) {
      long n = select_row(c, &s, row, NULL, NULL, NULL);
      size_t slot = row - c->row_begin;
      if (n < 0) n = 0;
      c->counts[slot] = (size_t)n;
      for (long r = 0; r < n; r++) {
        c->slot_center[slot * cc + r] = s.midx[r];
        c->slot_score[slot * cc + r] = s.mval[r];
      }
    }
  }
  scratch_free(&s);
  return NULL;
}

static void sel_ctx_init(sel_ctx *c, const orc_dia *d, const orc_lib *l,
                         const orc_params *p) {
  memset(c, 0, sizeof(*c));
  c->d = d;
  c->l = l;
  c->p = p;
  /* src/peak_group_selection/mod.rs:75-85 GaussianKernel::new(fwhm,1,ks,1) */
  c->klen = orc_gaussian_kernel(p->fwhm_rt, 1.0f, p->kernel_size, 1.0f,
                                c->kernel, 256);
  if (p->sigma_per_window && d->n_obs) {
    c->win_kernel = (float *)malloc(d->n_obs * c->klen * sizeof(float));
    for (size_t o = 0; o < d->n_obs; o++)
      orc_gaussian_kernel_sigma(p->sigma_per_window[o], c->klen,
                                c->win_kernel + o * c->klen, c->klen);
  }
}

orc_cands *orc_select(const orc_dia *d, const orc_lib *l, const orc_params *p,
                      size_t row_begin, size_t row_end, int nthreads) {
  sel_ctx c;
  sel_ctx_init(&c, d, l, p);
  if (row_end > l->n_prec) row_end = l->n_prec;
  if (row_begin > row_end) row_begin = row_end;
  size_t nr = row_end - row_begin;
  size_t cc = p->candidate_count;
  c.row_begin = row_begin;
  c.row_end = row_end;
  c.next = row_begin;
  c.counts = (size_t *)calloc(nr ? nr : 1, sizeof(size_t));
  size_t ns = (nr * cc) > 0 ? nr * cc : 1;
  c.slot_center = (uint64_t *)malloc(ns * 8);
  c.slot_score = (float *)malloc(ns * 4);
  pthread_mutex_init(&c.mu, NULL);
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  pthread_t th[256];
  for (int t = 0; t < nthreads; t++)
    pthread_create(&th[t], NULL, sel_worker, &c);
  for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  pthread_mutex_destroy(&c.mu);
  /* :118-119 flatten + collect preserves library order */
  size_t total = 0;
  for (size_t i = 0; i < nr; i++) total += c.counts[i];
  orc_cands *out = (orc_cands *)calloc(1, sizeof(orc_cands));
  size_t tz = total ? total : 1;
  out->n = total;
  out->precursor_idx = (uint64_t *)malloc(tz * 8);
  out->rank = (uint64_t *)malloc(tz * 8);
  out->cycle_center = (uint64_t *)malloc(tz * 8);
  out->cycle_start = (uint64_t *)malloc(tz * 8);
  out->cycle_stop = (uint64_t *)malloc(tz * 8);
  out->score = (float *)malloc(tz * 4);
  size_t k = 0;
  for (size_t i = 0; i < nr; i++)
    for (size_t r = 0; r < c.counts[i]; r++) {
      uint64_t center = c.slot_center[i * cc + r];
      /* :176-180 usize arithmetic (wrapping in release builds) */
      uint64_t start = center - (uint64_t)p->peak_length;
      uint64_t stop = center + (uint64_t)p->peak_length + 1;
      if (stop > d->n_rt) stop = d->n_rt;
      out->precursor_idx[k] = l->precursor_idx[row_begin + i];
      out->rank[k] = r;
      out->score[k] = c.slot_score[i * cc + r];
      out->cycle_center[k] = center;
      out->cycle_start[k] = start;
      out->cycle_stop[k] = stop;
      k++;
    }
  free(c.counts);
  free(c.slot_center);
  free(c.slot_score);
  free(c.win_kernel);
  return out;
}

void orc_cands_free(orc_cands *c) {
  if (!c) return;
  free(c->precursor_idx);
  free(c->rank);
  free(c->cycle_center);
  free(c->cycle_start);
  free(c->cycle_stop);
  free(c->score);
  free(c);
}

int orc_select_one(const orc_dia *d, const orc_lib *l, const orc_params *p,
                   size_t row, size_t *c0, size_t *c1, size_t *nf,
                   float *frag_mz, float *frag_w, float *xic, float *conv,
                   float *score, size_t cap_cells) {
  sel_ctx c;
  sel_ctx_init(&c, d, l, p);
  scratch s;
  memset(&s, 0, sizeof(s));
  long n = select_row(&c, &s, row, c0, c1, nf);
  int ok = 0;
  if (frag_mz)
    for (size_t j = 0; j < *nf; j++) {
      frag_mz[j] = s.fmz[j];
      frag_w[j] = s.fw[j];
    }
  if (n >= 0) {
    size_t L = *c1 - *c0;
    if (*nf * L <= cap_cells && L <= cap_cells) {
      if (xic) memcpy(xic, s.xic, *nf * L * 4);
      if (conv) memcpy(conv, s.conv, *nf * L * 4);
      if (score) memcpy(score, s.score, L * 4);
      ok = 1;
    } else {
      ok = -2;
    }
  }
  scratch_free(&s);
  free(c.win_kernel);
  return ok;
}
