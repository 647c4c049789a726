/*
 * adsr_oracle_score.c -- CPU restatement of PeakGroupScoring.score
 * (src/peak_group_scoring/mod.rs:43-341) and its helpers.
 *
 * TEST INFRASTRUCTURE ONLY (see adsr_oracle.h): the checker for the HIP
 * scoring kernel.  f32 arithmetic in the reference's order, one rounding
 * per operation (-ffp-contract=off); ln and log10 are glibc logf/log10f.
 *
 * Rust's `Iterator::sum` for f32 folds from -0.0 (core 1.88); ndarray's
 * `sum` on a contiguous lane is the 8-accumulator unrolled fold
 * (orc_ndarray_sum).  The two differ from a plain +0.0 fold only in the
 * sign of an all-zero result.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "adsr_oracle.h"

#define TYPE_B 98  /* src/constants.rs FragmentType::B ('b') */
#define TYPE_Y 121 /* FragmentType::Y ('y') */

static float iter_sum(const float *v, size_t n) {
  float s = -0.0f;
  for (size_t i = 0; i < n; i++) s = s + v[i];
  return s;
}

/* src/peak_group_scoring/utils.rs:126-172 calculate_correlation_safe */
float orc_correlation_safe(const float *x, const float *y, size_t n) {
  if (n == 0) return 0.0f;
  float xs = iter_sum(x, n), ys = iter_sum(y, n);
  if (xs == 0.0f || ys == 0.0f) return 0.0f;
  float xm = xs / (float)n, ym = ys / (float)n;
  float xv = 0.0f, yv = 0.0f, cov = 0.0f;
  for (size_t i = 0; i < n; i++) {
    float dx = x[i] - xm, dy = y[i] - ym;
    xv = xv + dx * dx;
    yv = yv + dy * dy;
    cov = cov + dx * dy;
  }
  if (xv == 0.0f || yv == 0.0f) return 0.0f;
#if ORC_MUTANT == 7 /* planted: one variance used twice */
  float c = cov / (sqrtf(xv) * sqrtf(xv));
#else
  float c = cov / (sqrtf(xv) * sqrtf(yv));
#endif
  if (c < -1.0f) c = -1.0f; /* clamp(-1, 1); NaN/inf panics there */
  if (c > 1.0f) c = 1.0f;
  return c;
}

/* src/peak_group_scoring/utils.rs:293-310 gamma_ln (Stirling) */
float orc_gamma_ln(float x) {
  if (x <= 0.0f) return 0.0f;
  if (fabsf(x - 1.0f) < 1e-6f) return 0.0f;
  if (fabsf(x - 2.0f) < 1e-6f) return 0.0f;
  const float ln_2pi = 1.837877f;
  return (x - 0.5f) * logf(x) - x + 0.5f * ln_2pi;
}

/* src/peak_group_scoring/utils.rs:193-279 calculate_hyperscore_weighted
 * (w == NULL: weight 1.0) */
float orc_hyperscore_weighted(const uint8_t *types, const float *inten,
                              const uint8_t *matched, const float *w,
                              size_t n) {
  uint32_t nb = 0, ny = 0;
  float sb = 0.0f, sy = 0.0f;
  for (size_t i = 0; i < n; i++) {
    if (!matched[i] || inten[i] == 0.0f) continue;
    float wi = w ? w[i] : 1.0f;
    float v = inten[i] * wi;
    if (types[i] == TYPE_B) {
      nb++;
      sb = sb + v;
    } else if (types[i] == TYPE_Y) {
      ny++;
      sy = sy + v;
    }
  }
  if (nb == 0 && ny == 0) return 0.0f;
  float fb = nb > 0 ? orc_gamma_ln((float)nb + 1.0f) : 0.0f;
  float fy = ny > 0 ? orc_gamma_ln((float)ny + 1.0f) : 0.0f;
  float lb = sb > 0.0f ? logf(sb) : 0.0f;
  float ly = sy > 0.0f ? logf(sy) : 0.0f;
#if ORC_MUTANT == 9 /* planted: factorial terms dropped */
  float h = lb + ly + 0.0f * (fb + fy);
#else
  float h = fb + fy + lb + ly;
#endif
  return isfinite(h) ? h : 0.0f;
}

/* utils.rs:339-365 find_longest_sequence over sorted fragment numbers */
static uint8_t longest_run(uint8_t *v, size_t n) {
  if (n == 0) return 0;
  for (size_t i = 1; i < n; i++) { /* sort_unstable: order of equals moot */
    uint8_t x = v[i];
    size_t j = i;
    while (j > 0 && v[j - 1] > x) {
      v[j] = v[j - 1];
      j--;
    }
    v[j] = x;
  }
  uint8_t best = 1, cur = 1;
  for (size_t i = 1; i < n; i++) {
    if ((unsigned)v[i] == (unsigned)v[i - 1] + 1u) {
      cur++; /* u8 arithmetic; at most 64 fragments here */
      if (cur > best) best = cur;
    } else {
      cur = 1;
    }
  }
  return best;
}

/* src/peak_group_scoring/utils.rs:314-371 calculate_longest_ion_series */
void orc_longest_ion_series(const uint8_t *types, const uint8_t *nums,
                            const uint8_t *matched, size_t n, uint8_t *b,
                            uint8_t *y) {
  uint8_t *bv = (uint8_t *)malloc(n ? n : 1), *yv = (uint8_t *)malloc(n ? n : 1);
  size_t nb = 0, ny = 0;
  for (size_t i = 0; i < n; i++) {
    if (!matched[i]) continue;
    if (types[i] == TYPE_B) bv[nb++] = nums[i];
    else if (types[i] == TYPE_Y) yv[ny++] = nums[i];
  }
  *b = longest_run(bv, nb);
  *y = longest_run(yv, ny);
  free(bv);
  free(yv);
}

/* src/peak_group_scoring/utils.rs:402-428 intensity_ion_series */
float orc_intensity_ion_series(const uint8_t *types, const float *inten,
                               const uint8_t *matched, size_t n,
                               uint8_t target) {
  float t = 0.0f;
  for (size_t i = 0; i < n; i++)
    if (matched[i] && inten[i] > 0.0f && types[i] == target) t = t + inten[i];
  return t;
}

/* src/peak_group_scoring/utils.rs:431-437 calculate_dot_product */
float orc_dot(const float *a, const float *b, size_t n) {
  if (n == 0) return 0.0f;
  float s = -0.0f;
  for (size_t i = 0; i < n; i++) s = s + a[i] * b[i];
  return s;
}

/* src/peak_group_scoring/utils.rs:453-476 calculate_fwhm_rt; index
 * arithmetic wraps like release-mode usize */
float orc_fwhm_rt(const float *p, size_t n, uint64_t cycle_start,
                  const float *rt) {
  if (n == 0) return 0.0f;
  size_t h = n / 2;
  float c = p[h];
  for (size_t i = 0; i < h; i++) {
    float m = (p[h - i] + p[h + i]) / 2.0f;
    if (m <= c / 2.0f) {
      float l = rt[cycle_start + h - i], r = rt[cycle_start + h + i];
      return r - l;
    }
  }
  return 0.0f;
}

/* src/peak_group_scoring/utils.rs:71-108 normalize_profiles */
void orc_normalize_profiles(const float *x, size_t rows, size_t cols,
                            size_t dil, float *out) {
  size_t c = cols / 2;
  for (size_t i = 0; i < rows; i++) {
    size_t s = c >= dil ? c - dil : 0;
    size_t e = c + dil + 1 < cols ? c + dil + 1 : cols;
    float sum = 0.0f;
    size_t cnt = 0;
    for (size_t j = s; j < e; j++) {
      sum = sum + x[i * cols + j];
      cnt++;
    }
    float mean = cnt > 0 ? sum / (float)cnt : 0.0f;
    for (size_t j = 0; j < cols; j++)
      out[i * cols + j] = mean > 0.0f ? x[i * cols + j] / mean : 0.0f;
  }
}

/* src/peak_group_scoring/utils.rs:7-29 filter_non_zero: rows kept, in order */
size_t orc_filter_non_zero(const float *x, size_t rows, size_t cols,
                           float *out) {
  size_t k = 0;
  for (size_t i = 0; i < rows; i++) {
    int any = 0;
    for (size_t j = 0; j < cols; j++) any |= x[i * cols + j] != 0.0f;
    if (any) memcpy(out + (k++) * cols, x + i * cols, cols * sizeof(float));
  }
  return k;
}

/* stable ascending sort by partial_cmp (NaN compares equal), as
 * slice::sort_by; insertion sort is stable */
static void sort_f32(float *v, size_t n) {
  for (size_t i = 1; i < n; i++) {
    float x = v[i];
    size_t j = i;
    while (j > 0 && v[j - 1] > x) {
      v[j] = v[j - 1];
      j--;
    }
    v[j] = x;
  }
}

/* src/peak_group_scoring/utils.rs:35-67 median_axis_0 */
void orc_median_axis_0(const float *x, size_t rows, size_t cols, float *out) {
  float *col = (float *)malloc((rows ? rows : 1) * sizeof(float));
  for (size_t j = 0; j < cols; j++) {
    if (rows == 0) {
      out[j] = 0.0f;
      continue;
    }
    for (size_t i = 0; i < rows; i++) col[i] = x[i * cols + j];
    sort_f32(col, rows);
    out[j] = rows % 2 == 0 ? (col[rows / 2 - 1] + col[rows / 2]) / 2.0f
                           : col[rows / 2];
  }
  free(col);
}

/* src/utils/mod.rs:120-137 calculate_median */
float orc_median(const float *v, size_t n) {
  if (n == 0) return 0.0f;
  float *s = (float *)malloc(n * sizeof(float));
  memcpy(s, v, n * sizeof(float));
  sort_f32(s, n);
  float m = n % 2 == 0 ? (s[n / 2 - 1] + s[n / 2]) / 2.0f : s[n / 2];
  free(s);
  return m;
}

/* src/utils/mod.rs:146-155 calculate_std (sample, n - 1); powi(2) is x*x */
float orc_std(const float *v, size_t n) {
  if (n < 2) return 0.0f;
  float mean = iter_sum(v, n) / (float)n;
  float s = -0.0f;
  for (size_t i = 0; i < n; i++) {
    float d = v[i] - mean;
    s = s + d * d;
  }
  return sqrtf(s / (float)(n - 1));
}

/* src/utils/mod.rs:175-207 create_ranked_mask: rank by value descending,
 * then index ascending; ranks [r1, r2) selected */
void orc_ranked_mask(const float *v, size_t n, size_t r1, size_t r2,
                     uint8_t *mask) {
  for (size_t i = 0; i < n; i++) {
    mask[i] = 0;
    if (r1 >= r2) continue;
    size_t r = 0;
    for (size_t j = 0; j < n; j++)
#if ORC_MUTANT == 8 /* planted: ascending rank */
      if (v[j] < v[i] || (!(v[j] < v[i]) && !(v[j] > v[i]) && j < i)) r++;
#else
      if (v[j] > v[i] || (!(v[j] < v[i]) && !(v[j] > v[i]) && j < i)) r++;
#endif
    mask[i] = r >= r1 && r < r2;
  }
}

/* src/utils/mod.rs:223-237 count_values_above (mask may be NULL) */
size_t orc_count_above(const float *v, size_t n, float thr,
                       const uint8_t *mask) {
  size_t c = 0;
  for (size_t i = 0; i < n; i++)
    if ((!mask || mask[i]) && v[i] > thr) c++;
  return c;
}

/* src/utils/mod.rs:83-109 calculate_weighted_mean_absolute_error */
float orc_weighted_mae(const float *err, const float *w, size_t n) {
  if (n == 0) return 0.0f;
  float ws = 0.0f, wt = 0.0f;
  for (size_t i = 0; i < n; i++) {
    if (w[i] > 0.0f && err[i] != 0.0f) {
      ws = ws + fabsf(err[i]) * w[i];
      wt = wt + w[i];
    }
  }
  return wt > 0.0f ? ws / wt : 0.0f;
}

/* src/utils/mod.rs:11-38 calculate_fragment_mz_and_errors with
 * :41-69 weighted_mean_nonzero_axis_1 */
void orc_fragment_mz_errors(const float *dmz, const float *dxic, size_t rows,
                            size_t cols, const float *mz_lib, float *mz_obs,
                            float *err) {
  for (size_t i = 0; i < rows; i++) {
    float ws = 0.0f, wt = 0.0f;
    for (size_t j = 0; j < cols; j++) {
      float w = dxic[i * cols + j];
      if (w > 0.0f) {
        ws = ws + dmz[i * cols + j] * w;
        wt = wt + w;
      }
    }
    float m = wt > 0.0f ? ws / wt : 0.0f;
    mz_obs[i] = m;
    err[i] = m > 0.0f ? ((m - mz_lib[i]) / mz_lib[i]) * 1e6f : 0.0f;
  }
}

/* src/quadrupole_observation/mod.rs:152-211 fill_xic_and_mz_slice; c0 and
 * the relative index wrap like release-mode usize */
void orc_fill_xic_and_mz_slice(const orc_dia *d, size_t obs_idx, float *xic,
                               float *dmz, uint64_t c0, uint64_t c1,
                               float tol, float mz) {
  const orc_obs *o = &d->obs[obs_idx];
  float delta = mz * tol * 1e-6f;
  size_t s, e;
  orc_mz_range_indices(d->grid, d->n_grid, mz - delta, mz + delta, &s, &e);
  for (size_t b = s; b < e; b++) {
    float actual = d->grid[b];
    size_t a = o->slice_starts[b], z = o->slice_starts[b + 1];
    const uint16_t *ci = o->cycle_indices + a;
    const float *in = o->intensities + a;
    size_t pos;
    orc_rust_bsearch_u16(ci, z - a, (uint16_t)c0, &pos);
    for (size_t i = pos; i < z - a; i++) {
      uint64_t cyc = ci[i];
      if (cyc >= c1) break;
      uint64_t rel = cyc - c0;
      float v = in[i];
      xic[rel] = xic[rel] + v;
      if (v > 0.0f) {
        float prev = xic[rel] - v;
        if (prev == 0.0f) dmz[rel] = actual;
#if ORC_MUTANT == 11 /* planted: unweighted running m/z */
        else dmz[rel] = (dmz[rel] + actual) / 2.0f;
#else
        else dmz[rel] = (dmz[rel] * prev + actual * v) / xic[rel];
#endif
      }
    }
  }
}

/* src/dense_xic_observation/mod.rs:91-138 DenseXICMZObservation::new:
 * xic and mz are [nf][c1 - c0], zeroed here */
void orc_dense_xic_mz(const orc_dia *d, float precursor_mz, uint64_t c0,
                      uint64_t c1, float tol, const float *fragment_mz,
                      size_t nf, float *xic, float *mz) {
  size_t L = (size_t)(c1 - c0);
  memset(xic, 0, nf * L * sizeof(float));
  memset(mz, 0, nf * L * sizeof(float));
  size_t *obs = malloc((d->n_obs ? d->n_obs : 1) * sizeof(size_t));
  size_t nobs = orc_valid_observations(d, precursor_mz, obs);
  for (size_t k = 0; k < nobs; k++)
    for (size_t f = 0; f < nf; f++)
      orc_fill_xic_and_mz_slice(d, obs[k], xic + f * L, mz + f * L, c0, c1, tol,
                                fragment_mz[f]);
  free(obs);
}

/* src/idf/mod.rs:13-33 InverseDocumentFrequency::new: per-bin counts of
 * all library fragment_mz_library values */
uint32_t *orc_idf_counts(const float *fragment_mz_library, size_t n) {
  size_t B;
  const float *g = orc_global_grid(&B);
  uint32_t *c = (uint32_t *)calloc(B, sizeof(uint32_t));
  for (size_t i = 0; i < n; i++)
    c[orc_find_closest_index(g, B, fragment_mz_library[i])]++;
  return c;
}

/* src/idf/mod.rs:35-49 get_idf */
float orc_idf(const uint32_t *counts, float total, float mz) {
  if (total == 0.0f) return 1.0f;
  size_t B;
  const float *g = orc_global_grid(&B);
  uint32_t df = counts[orc_find_closest_index(g, B, mz)];
#if ORC_MUTANT == 10 /* planted: document frequency not floored at 1 */
  return logf(total / (float)df);
#else
  return logf(total / (float)(df > 1 ? df : 1));
#endif
}

typedef struct {
  const orc_dia *d;
  const orc_lib *l;
  const uint32_t *idf_counts;
  float idf_total;
  float tol;
  size_t top_k;
  size_t n;
  const uint64_t *pidx, *start, *stop, *center;
  const float *score;
  float *feat; /* [ORC_N_FEATURES][n] */
  uint8_t *found;
  size_t next;
  pthread_mutex_t mu;
} score_ctx;

#define FEAT(k) c->feat[(size_t)(k) * c->n + ci]

/* src/peak_group_scoring/mod.rs:105-341 score_candidate_generic */
static void score_one(score_ctx *c, size_t ci) {
  const orc_lib *l = c->l;
  const orc_dia *d = c->d;
  /* speclib_flat/mod.rs:458-475 get_precursor_by_idx_filtered: binary
   * search of the sorted precursor_idx */
  uint64_t key = c->pidx[ci];
  size_t lo = 0, hi = l->n_prec;
  while (lo < hi) {
    size_t mid = lo + (hi - lo) / 2;
    if (l->precursor_idx[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  if (lo >= l->n_prec || l->precursor_idx[lo] != key) {
    c->found[ci] = 0;
    return;
  }
  c->found[ci] = 1;
  size_t row = lo;
  size_t fs = l->frag_start[row], nraw = l->frag_stop[row] - fs;
  size_t *ord = (size_t *)malloc((nraw ? nraw : 1) * sizeof(size_t));
  size_t F = orc_filter_sort_fragments(nraw, l->fragment_mz + fs,
                                       l->fragment_intensity + fs,
                                       l->fragment_number + fs,
                                       l->fragment_type + fs, 1, 1, c->top_k,
                                       ord);
  uint64_t c0 = c->start[ci], c1 = c->stop[ci];
  size_t L = (size_t)(c1 - c0);
  size_t cells = F * L, fz = F ? F : 1;
  if (cells == 0) cells = 1;
  float *fmz = malloc(fz * 4), *fint = malloc(fz * 4), *fmzl = malloc(fz * 4);
  uint8_t *ftype = malloc(fz), *fnum = malloc(fz), *matched = malloc(fz);
  for (size_t f = 0; f < F; f++) {
    size_t j = fs + ord[f];
    fmz[f] = l->fragment_mz[j];
    fint[f] = l->fragment_intensity[j];
    fmzl[f] = l->fragment_mz_library[j];
    ftype[f] = l->fragment_type[j];
    fnum[f] = l->fragment_number[j];
  }
  float *xic = calloc(cells, 4), *dmz = calloc(cells, 4);
  orc_dense_xic_mz(d, l->precursor_mz[row], c0, c1, c->tol, fmz, F, xic, dmz);
  /* mod.rs:130-143 */
  float *norm = malloc(cells * 4), *filt = malloc(cells * 4);
  float *med = malloc((L ? L : 1) * 4), *medf = malloc((L ? L : 1) * 4);
#if ORC_MUTANT == 12 /* planted: centre only, no dilation */
  orc_normalize_profiles(xic, F, L, 0, norm);
#else
  orc_normalize_profiles(xic, F, L, 1, norm);
#endif
  size_t Ff = orc_filter_non_zero(norm, F, L, filt);
  orc_median_axis_0(norm, F, L, med);
  orc_median_axis_0(filt, Ff, L, medf);
  float med_sum = iter_sum(med, L), medf_sum = iter_sum(medf, L);
  /* mod.rs:145-149 */
  float fwhm = orc_fwhm_rt(medf, L, c0, d->rt);
  /* mod.rs:152-160 */
  float *corr = malloc(fz * 4), *oint = malloc(fz * 4);
  for (size_t f = 0; f < F; f++) {
    corr[f] = orc_correlation_safe(medf, norm + f * L, L);
    oint[f] = orc_ndarray_sum(xic + f * L, L); /* sum_axis(Axis(1)) */
  }
  float int_corr = orc_correlation_safe(oint, fint, F);
  /* mod.rs:164-197 */
  float mean_corr = F ? iter_sum(corr, F) / (float)F : 0.0f;
  float med_corr = orc_median(corr, F);
  float std_corr = orc_std(corr, F);
  uint8_t *m1 = malloc(fz), *m2 = malloc(fz), *m3 = malloc(fz), *m4 = malloc(fz);
  orc_ranked_mask(fint, F, 0, 6, m1);
  orc_ranked_mask(fint, F, 6, 12, m2);
  orc_ranked_mask(fint, F, 12, 18, m3);
  orc_ranked_mask(fint, F, 18, 24, m4);
  /* mod.rs:203-247 */
  for (size_t f = 0; f < F; f++) matched[f] = oint[f] > 0.0f;
  float hs_obs = orc_hyperscore_weighted(ftype, oint, matched, NULL, F);
  float hs_lib = orc_hyperscore_weighted(ftype, fint, matched, NULL, F);
  uint8_t lb, ly;
  orc_longest_ion_series(ftype, fnum, matched, F, &lb, &ly);
  float *mzo = malloc(fz * 4), *err = malloc(fz * 4), *wie = malloc(fz * 4);
  orc_fragment_mz_errors(dmz, xic, F, L, fmz, mzo, err);
  float wmae = orc_weighted_mae(err, fint, F);
  for (size_t f = 0; f < F; f++) wie[f] = 1.0f / (fabsf(err[f]) + 0.1f);
  float hs_ime = orc_hyperscore_weighted(ftype, oint, matched, wie, F);
  /* mod.rs:249-268 */
  float rt_obs = d->rt[c->center[ci]];
  float delta_rt = rt_obs - l->precursor_rt[row];
  float ib = orc_intensity_ion_series(ftype, oint, matched, F, TYPE_B);
  float iy = orc_intensity_ion_series(ftype, oint, matched, F, TYPE_Y);
  float log_b = log10f(ib + 1e-8f), log_y = log10f(iy + 1e-8f);
  /* mod.rs:272-293 */
  float *idf = malloc(fz * 4);
  for (size_t f = 0; f < F; f++) idf[f] = orc_idf(c->idf_counts, c->idf_total, fmzl[f]);
  float idf_hs = orc_hyperscore_weighted(ftype, idf, matched, NULL, F);
  float idf_xic = orc_dot(idf, corr, F);
  float idf_int = orc_dot(idf, oint, F);
  float log_idf_int = log10f(idf_int + 1e-8f);
  uint8_t *m6 = malloc(fz);
  orc_ranked_mask(idf, F, 0, 6, m6);

  FEAT(ORC_F_SCORE) = c->score[ci];
  FEAT(ORC_F_MEAN_CORRELATION) = mean_corr;
  FEAT(ORC_F_MEDIAN_CORRELATION) = med_corr;
  FEAT(ORC_F_CORRELATION_STD) = std_corr;
  FEAT(ORC_F_INTENSITY_CORRELATION) = int_corr;
  FEAT(ORC_F_NUM_FRAGMENTS) = (float)F;
  FEAT(ORC_F_NUM_SCANS) = (float)(uint64_t)(c1 - c0);
  FEAT(ORC_F_NUM_OVER_95) = (float)orc_count_above(corr, F, 0.95f, NULL);
  FEAT(ORC_F_NUM_OVER_90) = (float)orc_count_above(corr, F, 0.90f, NULL);
  FEAT(ORC_F_NUM_OVER_80) = (float)orc_count_above(corr, F, 0.80f, NULL);
  FEAT(ORC_F_NUM_OVER_50) = (float)orc_count_above(corr, F, 0.50f, NULL);
  FEAT(ORC_F_NUM_OVER_0) = (float)orc_count_above(corr, F, 0.0f, NULL);
  FEAT(ORC_F_NUM_OVER_0_RANK_0_5) = (float)orc_count_above(corr, F, 0.0f, m1);
  FEAT(ORC_F_NUM_OVER_0_RANK_6_11) = (float)orc_count_above(corr, F, 0.0f, m2);
  FEAT(ORC_F_NUM_OVER_0_RANK_12_17) = (float)orc_count_above(corr, F, 0.0f, m3);
  FEAT(ORC_F_NUM_OVER_0_RANK_18_23) = (float)orc_count_above(corr, F, 0.0f, m4);
  FEAT(ORC_F_NUM_OVER_50_RANK_0_5) = (float)orc_count_above(corr, F, 0.5f, m1);
  FEAT(ORC_F_NUM_OVER_50_RANK_6_11) = (float)orc_count_above(corr, F, 0.5f, m2);
  FEAT(ORC_F_NUM_OVER_50_RANK_12_17) = (float)orc_count_above(corr, F, 0.5f, m3);
  FEAT(ORC_F_NUM_OVER_50_RANK_18_23) = (float)orc_count_above(corr, F, 0.5f, m4);
  FEAT(ORC_F_HYPERSCORE_INTENSITY_OBSERVATION) = hs_obs;
  FEAT(ORC_F_HYPERSCORE_INTENSITY_LIBRARY) = hs_lib;
  FEAT(ORC_F_HYPERSCORE_INVERSE_MASS_ERROR) = hs_ime;
  FEAT(ORC_F_RT_OBSERVED) = rt_obs;
  FEAT(ORC_F_DELTA_RT) = delta_rt;
  FEAT(ORC_F_LONGEST_B_SERIES) = (float)lb;
  FEAT(ORC_F_LONGEST_Y_SERIES) = (float)ly;
  FEAT(ORC_F_NAA) = (float)l->precursor_naa[row];
  FEAT(ORC_F_WEIGHTED_MASS_ERROR) = wmae;
  FEAT(ORC_F_LOG10_B_ION_INTENSITY) = log_b;
  FEAT(ORC_F_LOG10_Y_ION_INTENSITY) = log_y;
  FEAT(ORC_F_FWHM_RT) = fwhm;
  FEAT(ORC_F_IDF_HYPERSCORE) = idf_hs;
  FEAT(ORC_F_IDF_XIC_DOT_PRODUCT) = idf_xic;
  FEAT(ORC_F_IDF_INTENSITY_DOT_PRODUCT) = log_idf_int;
  FEAT(ORC_F_MEDIAN_PROFILE_SUM) = med_sum;
  FEAT(ORC_F_MEDIAN_PROFILE_SUM_FILTERED) = medf_sum;
  FEAT(ORC_F_NUM_PROFILES) = (float)F;
  FEAT(ORC_F_NUM_PROFILES_FILTERED) = (float)Ff;
  FEAT(ORC_F_NUM_OVER_0_TOP6_IDF) = (float)orc_count_above(corr, F, 0.0f, m6);
  FEAT(ORC_F_NUM_OVER_50_TOP6_IDF) = (float)orc_count_above(corr, F, 0.5f, m6);

  free(ord); free(fmz); free(fint); free(fmzl); free(ftype); free(fnum);
  free(matched); free(xic); free(dmz); free(norm); free(filt); free(med);
  free(medf); free(corr); free(oint); free(m1); free(m2); free(m3); free(m4);
  free(mzo); free(err); free(wie); free(idf); free(m6);
}

static void *score_worker(void *arg) {
  score_ctx *c = (score_ctx *)arg;
  for (;;) {
    pthread_mutex_lock(&c->mu);
    size_t a = c->next;
    c->next += 256;
    pthread_mutex_unlock(&c->mu);
    if (a >= c->n) break;
    size_t b = a + 256 < c->n ? a + 256 : c->n;
    for (size_t i = a; i < b; i++) score_one(c, i);
  }
  return NULL;
}

/* src/peak_group_scoring/mod.rs:55-102 score_generic: every candidate in
 * order; found[i] = 0 where the precursor is not in the library (skipped
 * by filter_map) */
void orc_score(const orc_dia *d, const orc_lib *l, float mass_tolerance,
               size_t top_k, size_t n, const uint64_t *precursor_idx,
               const float *score, const uint64_t *cycle_start,
               const uint64_t *cycle_stop, const uint64_t *cycle_center,
               float *features, uint8_t *found, int nthreads) {
  score_ctx c;
  memset(&c, 0, sizeof(c));
  c.d = d;
  c.l = l;
  /* src/speclib_flat/mod.rs:165-166 builds the IDF once per library */
  orc_lib *lm = (orc_lib *)l;
  if (!lm->idf_counts) lm->idf_counts = orc_idf_counts(l->fragment_mz_library, l->n_frag);
  c.idf_counts = lm->idf_counts;
  c.idf_total = (float)l->n_frag;
  c.tol = mass_tolerance;
  c.top_k = top_k;
  c.n = n;
  c.pidx = precursor_idx;
  c.score = score;
  c.start = cycle_start;
  c.stop = cycle_stop;
  c.center = cycle_center;
  c.feat = features;
  c.found = found;
  pthread_mutex_init(&c.mu, NULL);
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 64) nthreads = 64;
  pthread_t th[64];
  for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, score_worker, &c);
  for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  pthread_mutex_destroy(&c.mu);
}

/* ---- src/peak_group_quantification/mod.rs:39-168 ---------------------- */
/* Per candidate i: row[i] = library row (sorted order) or (size_t)-1 when
 * the candidate yields no PrecursorQuantified (precursor not found,
 * num_cycles == 0, centre outside [start, stop)); nf[i] selected fragments
 * and, in slots [i * ks, i * ks + nf[i]), the raw fragment index (library
 * fragment order) and the observed intensity (sum over cycles), m/z,
 * correlation with the filtered median profile and mass error (ppm).  The
 * intensity > 0 filter (precursor_quantified/mod.rs:26-94) and the flat
 * layout are left to the caller.
 * Pinned by tests/test_oracle_quant.py: an independent float64 numpy
 * implementation on 300 rank-0 candidates of workload 0 (intensity within
 * L * 2^-24, correlation within 1e-4, m/z and ppm error bit for bit, the
 * rt, the filter and the layout exactly), with planted mistakes 13-15. */
typedef struct {
  const orc_dia *d;
  const orc_lib *l;
  float tol;
  size_t top_k, n, ks;
  const uint64_t *pidx, *start, *stop, *center;
  size_t *row, *nf, *raw;
  float *intensity, *mz_obs, *corr, *mass_err, *rt_observed;
  size_t next;
  pthread_mutex_t mu;
} quant_ctx;

static void quantify_one(const quant_ctx *q, size_t ci) {
  const orc_dia *d = q->d;
  const orc_lib *l = q->l;
  const size_t ks = q->ks;
  size_t *row = q->row, *nf = q->nf, *raw = q->raw;
  float *intensity = q->intensity, *mz_obs = q->mz_obs, *corr = q->corr,
        *mass_err = q->mass_err, *rt_observed = q->rt_observed;
  const float tolerance_ppm = q->tol;
  const size_t top_k = q->top_k;
  {
    row[ci] = (size_t)-1;
    nf[ci] = 0;
    uint64_t key = q->pidx[ci];
    size_t lo = 0, hi = l->n_prec;
    while (lo < hi) {
      size_t mid = lo + (hi - lo) / 2;
      if (l->precursor_idx[mid] < key) lo = mid + 1;
      else hi = mid;
    }
    if (lo >= l->n_prec || l->precursor_idx[lo] != key) return;
    uint64_t c0 = q->start[ci], c1 = q->stop[ci], cc = q->center[ci];
    uint64_t L = c1 - c0; /* mod.rs:93-98 */
    if (L == 0 || cc < c0 || cc >= c1) return;
    size_t fs = l->frag_start[lo], nraw = l->frag_stop[lo] - fs;
    size_t *ord = malloc((nraw ? nraw : 1) * sizeof(size_t));
    size_t F = orc_filter_sort_fragments(nraw, l->fragment_mz + fs,
                                         l->fragment_intensity + fs,
                                         l->fragment_number + fs,
                                         l->fragment_type + fs, 1, 1, top_k, ord);
    if (F > ks) F = ks; /* caller sizes ks to the longest selection */
    size_t fz = F ? F : 1, cells = F * L;
    if (cells == 0) cells = 1;
    float *fmz = malloc(fz * 4), *xic = malloc(cells * 4), *dmz = malloc(cells * 4);
    float *norm = malloc(cells * 4), *filt = malloc(cells * 4), *med = malloc(L * 4);
    for (size_t f = 0; f < F; f++) fmz[f] = l->fragment_mz[fs + ord[f]];
    orc_dense_xic_mz(d, l->precursor_mz[lo], c0, c1, tolerance_ppm, fmz, F, xic, dmz);
    float *mo = malloc(fz * 4), *er = malloc(fz * 4);
    orc_fragment_mz_errors(dmz, xic, F, L, fmz, mo, er);
    orc_normalize_profiles(xic, F, L, 1, norm);
    size_t Ff = orc_filter_non_zero(norm, F, L, filt);
    orc_median_axis_0(filt, Ff, L, med);
#if ORC_MUTANT == 14 /* planted: median over every profile, not the non-zero ones */
    orc_median_axis_0(norm, F, L, med);
#endif
    for (size_t f = 0; f < F; f++) {
      size_t s = ci * ks + f;
      raw[s] = fs + ord[f];
      corr[s] = orc_correlation_safe(norm + f * L, med, L); /* :121-128 */
#if ORC_MUTANT == 13 /* planted: the last cycle left out of the sum */
      intensity[s] = orc_ndarray_sum(xic + f * L, L - 1);
#else
      intensity[s] = orc_ndarray_sum(xic + f * L, L);      /* :131 */
#endif
      mz_obs[s] = mo[f];
      mass_err[s] = er[f];
    }
#if ORC_MUTANT == 15 /* planted: RT of the first cycle instead of the centre */
    rt_observed[ci] = d->rt[c0];
#else
    rt_observed[ci] = d->rt[cc];
#endif
    row[ci] = lo;
    nf[ci] = F;
    free(ord); free(fmz); free(xic); free(dmz); free(norm); free(filt);
    free(med); free(mo); free(er);
  }
}

static void *quant_worker(void *arg) {
  quant_ctx *q = (quant_ctx *)arg;
  for (;;) {
    pthread_mutex_lock(&q->mu);
    size_t a = q->next;
    q->next += 256;
    pthread_mutex_unlock(&q->mu);
    if (a >= q->n) break;
    size_t b = a + 256 < q->n ? a + 256 : q->n;
    for (size_t i = a; i < b; i++) quantify_one(q, i);
  }
  return NULL;
}

/* Candidates are independent (src/peak_group_quantification/mod.rs:57-71
 * maps them with par_iter, order kept by collect); each writes only its own
 * slots, on nthreads threads. */
void orc_quantify(const orc_dia *d, const orc_lib *l, float tolerance_ppm,
                  size_t top_k, size_t n, const uint64_t *precursor_idx,
                  const uint64_t *cycle_start, const uint64_t *cycle_stop,
                  const uint64_t *cycle_center, size_t ks, size_t *row,
                  size_t *nf, size_t *raw, float *intensity, float *mz_obs,
                  float *corr, float *mass_err, float *rt_observed, int nthreads) {
  quant_ctx q;
  memset(&q, 0, sizeof(q));
  q.d = d;
  q.l = l;
  q.tol = tolerance_ppm;
  q.top_k = top_k;
  q.n = n;
  q.ks = ks;
  q.pidx = precursor_idx;
  q.start = cycle_start;
  q.stop = cycle_stop;
  q.center = cycle_center;
  q.row = row;
  q.nf = nf;
  q.raw = raw;
  q.intensity = intensity;
  q.mz_obs = mz_obs;
  q.corr = corr;
  q.mass_err = mass_err;
  q.rt_observed = rt_observed;
  pthread_mutex_init(&q.mu, NULL);
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 64) nthreads = 64;
  pthread_t th[64];
  for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, quant_worker, &q);
  for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  pthread_mutex_destroy(&q.mu);
}
