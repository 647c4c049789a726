// adsr_host.cpp -- host side of the gfx950 candidate-selection library:
// the exact 1-ppm m/z grid and its bucket table, the per-window CSR builder
// (DIAData.from_arrays), the spectral library (SpecLibFlat.from_arrays),
// candidate tables and the error plumbing.  No selection arithmetic runs
// here: PeakGroupSelection.search runs on the device (adsr_device.hip).
//
// Built with -ffp-contract=off: f32 expressions are evaluated one rounding
// at a time, as the Rust reference does.

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <thread>

#include "adsr_internal.hpp"

namespace adsr {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_last_error;

adsr_status set_error(adsr_status st, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return st;
}

void clear_error() { g_last_error.clear(); }

// ADSR_FAIL_ALLOC=n (n >= 1) makes the n-th allocation of the process that
// goes through this check (device buffers, page-locked and host candidate
// tables) report out of memory, so that tests can drive every error path
// of an allocation without exhausting memory; unset or 0: never.
bool alloc_should_fail() {
  static const long target = [] {
    const char *e = std::getenv("ADSR_FAIL_ALLOC");
    return e ? std::strtol(e, nullptr, 10) : 0L;
  }();
  if (target <= 0) return false;
  static std::atomic<long> count{0};
  return count.fetch_add(1) + 1 == target;
}

static inline uint32_t f2u(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  return u;
}
static inline float u2f(uint32_t u) {
  float x;
  std::memcpy(&x, &u, 4);
  return x;
}

// ---------------------------------------------------------------------------
// m/z grid -- src/mz_index/mod.rs:8-49
// ---------------------------------------------------------------------------
static Grid *build_grid() {
  Grid *g = new Grid();
  // ppm_index(1.0, 150.0, 2000.0): m0 = max(150, 50); m += m * (1/1e6)
  const float res = 1.0f, start = 150.0f, end = 2000.0f;
  float m = std::max(start, 50.0f);
  g->mz.reserve(2600000);
  g->mz.push_back(m);
  while (m < end) {
    float step = m * (res / 1e6f);
    m = m + step;
    g->mz.push_back(m);
  }
  const uint32_t n = (uint32_t)g->mz.size();
  g->base_key = f2u(g->mz[0]) >> kBucketShift;
  g->n_bucket = (f2u(g->mz[n - 1]) >> kBucketShift) - g->base_key + 1;
  g->bucket.resize((size_t)g->n_bucket + 1);
  uint32_t p = 0;
  for (uint32_t key = 0; key <= g->n_bucket; ++key) {
    float f = u2f((key + g->base_key) << kBucketShift);
    while (p < n && g->mz[p] < f) ++p;
    g->bucket[key] = p;
  }
  return g;
}

const Grid &global_grid() {
  static Grid *g = build_grid();  // thread-safe static init
  return *g;
}

uint32_t grid_lower_bound(const Grid &g, float x) {
  const float *m = g.mz.data();
  const uint32_t n = (uint32_t)g.mz.size();
  if (!(x > m[0])) return 0;  // also NaN: reference narrows to index 0
  if (x > m[n - 1]) return n;
  uint32_t b = g.bucket[(f2u(x) >> kBucketShift) - g.base_key];
  while (b < n && m[b] < x) ++b;
  return b;
}

// src/mz_index/mod.rs:55-90.  The reference binary search either hits the
// (unique) exact grid value or ends at the lower bound; ties between the two
// neighbours go to the lower index.
uint32_t grid_find_closest(const Grid &g, float x) {
  const float *m = g.mz.data();
  const uint32_t n = (uint32_t)g.mz.size();
  if (x != x) return 0;
  uint32_t lb = grid_lower_bound(g, x);
  if (lb < n && m[lb] == x) return lb;
  if (lb == 0) return 0;
  if (lb == n) return n - 1;
  float left_diff = std::fabs(m[lb] - x);
  float right_diff = std::fabs(m[lb - 1] - x);
  return left_diff < right_diff ? lb : lb - 1;
}

// ---------------------------------------------------------------------------
// small thread pool helper
// ---------------------------------------------------------------------------
static unsigned n_host_threads() {
  const char *e = std::getenv("ADSR_NUM_THREADS");
  if (e && *e) {
    long v = std::strtol(e, nullptr, 10);
    if (v > 0) return (unsigned)std::min<long>(v, 256);
  }
  unsigned h = std::thread::hardware_concurrency();
  return h ? std::min(h, 64u) : 4u;
}

template <class F>
static void parallel_for(size_t n, F &&fn) {
  unsigned nt = (unsigned)std::min<size_t>(n_host_threads(), n);
  if (nt <= 1) {
    for (size_t i = 0; i < n; ++i) fn(i);
    return;
  }
  std::atomic<size_t> next{0};
  std::vector<std::thread> th;
  th.reserve(nt);
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&] {
      for (;;) {
        size_t i = next.fetch_add(1);
        if (i >= n) break;
        fn(i);
      }
    });
  for (auto &t : th) t.join();
}

}  // namespace adsr

using namespace adsr;

// ===========================================================================
// C ABI: errors, params, grid
// ===========================================================================
extern "C" const char *adsr_last_error(void) {
  return adsr::g_last_error.c_str();
}

// src/peak_group_selection/parameters.rs:24-43
extern "C" void adsr_selection_params_default(adsr_selection_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->fwhm_rt = 3.0f;
  p->kernel_size = 20;
  p->peak_length = 5;
  p->mass_tolerance = 7.0f;
  p->rt_tolerance = 200.0f;
  p->candidate_count = 3;
  p->top_k_fragments = 12;
}

extern "C" size_t adsr_mz_grid(float *out, size_t cap) {
  const Grid &g = global_grid();
  if (out) std::memcpy(out, g.mz.data(), std::min(cap, g.mz.size()) * 4);
  return g.mz.size();
}

// src/kernel/mod.rs:13-57 GaussianKernel::new -> gaussian_kernel_1d.
// The normalising sum is ndarray's unrolled_fold (8 partial sums).
static size_t gaussian_kernel_sigma(float sigma, size_t size, float *out,
                                    size_t cap) {
  if (size > cap) return 0;
  const int half = (int)(size / 2);
  const float PI_F = 3.14159265358979323846f;
  const float normalization = 1.0f / (sigma * std::sqrt(2.0f * PI_F));
  for (int x = -half; x < half; ++x) {
    float xf = (float)x;
    float exponent = (-0.5f * (xf * xf)) / (sigma * sigma);
    out[x + half] = normalization * expf(exponent);
  }
  float p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  size_t i = 0;
  while (size - i >= 8) {
    for (int j = 0; j < 8; ++j) p[j] = p[j] + out[i + j];
    i += 8;
  }
  float acc = 0.0f;
  acc = acc + (p[0] + p[4]);
  acc = acc + (p[1] + p[5]);
  acc = acc + (p[2] + p[6]);
  acc = acc + (p[3] + p[7]);
  for (; i < size; ++i) acc = acc + out[i];
  for (size_t j = 0; j < size; ++j) out[j] = out[j] / acc;
  return size;
}

extern "C" size_t adsr_gaussian_kernel_sigma(float sigma, uint64_t width,
                                             float *out, size_t cap) {
  return gaussian_kernel_sigma(sigma, (size_t)width, out, cap);
}

extern "C" size_t adsr_gaussian_kernel(float fwhm_rt, float sigma_scale_rt,
                                       uint64_t kernel_width,
                                       float rt_resolution, float *out,
                                       size_t cap) {
  size_t w = (size_t)std::ceil((float)kernel_width / 2.0f) * 2;
  float sigma = fwhm_rt / 2.3548f;
  float rt_sigma = sigma * sigma_scale_rt / rt_resolution;
  if (!out) return w;
  return gaussian_kernel_sigma(rt_sigma, w, out, cap);
}

// ===========================================================================
// DIAData -- src/dia_data/mod.rs:41-135, src/dia_data_builder/mod.rs:11-139
// ===========================================================================
namespace adsr {

adsr_status dia_prepare(const int64_t *delta, const float *iso_lo,
                        const float *iso_hi, const int64_t *pstart,
                        const int64_t *pstop, const int64_t *cyc,
                        const float *srt, size_t n_spectra,
                        const float *peak_mz, const float *peak_int,
                        size_t n_peaks, const float *cycle,
                        const size_t cycle_shape[4], bool host_csr,
                        std::vector<std::vector<size_t>> &spectra,
                        adsr_dia **out) {
  if (!out) return set_error(ADSR_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  if (n_spectra && (!delta || !iso_lo || !iso_hi || !pstart || !pstop ||
                    !cyc || !srt))
    return set_error(ADSR_ERR_INVALID_ARGUMENT, "spectrum arrays are NULL");
  if (n_peaks && (!peak_mz || !peak_int))
    return set_error(ADSR_ERR_INVALID_ARGUMENT, "peak arrays are NULL");
  if (n_spectra == 0)  // src/rt_index/mod.rs:57-59 indexes len-1
    return set_error(ADSR_ERR_INVALID_ARGUMENT,
                     "spectrum arrays must not be empty");
  const Grid &g = global_grid();
  const size_t B = g.mz.size();
  auto *d = new adsr_dia();
  d->n_grid = B;
  d->rt_values.assign(srt, srt + n_spectra);
  size_t ncyc = 1;
  for (int i = 0; i < 4; ++i) {
    d->cycle_shape[i] = cycle_shape ? cycle_shape[i] : 0;
    ncyc *= d->cycle_shape[i];
  }
  if (ncyc && cycle) d->cycle.assign(cycle, cycle + ncyc);
  else d->cycle.assign(ncyc, 0.0f);

  // src/rt_index/mod.rs:53-75: RT axis from delta_scan_idx == 0 spectra
  for (size_t s = 0; s < n_spectra; ++s)
    if (delta[s] == 0) d->rt_index.push_back(srt[s]);
  const size_t nrt = d->rt_index.size();
  d->delta_t = nrt > 1 ? (d->rt_index[nrt - 1] - d->rt_index[0]) /
                             (float)(nrt - 1)
                       : 1.0f;
  for (size_t i = 1; i < nrt; ++i)
    if (d->rt_index[i] < d->rt_index[i - 1]) d->rt_sorted = false;
  for (size_t i = 1; i < nrt; ++i)
    if (!(d->rt_index[i] > d->rt_index[i - 1])) d->rt_strict = false;

  // src/dia_data_builder/mod.rs:33-46: observations 0..=max(delta)
  int64_t maxd = delta[0];
  for (size_t s = 1; s < n_spectra; ++s) maxd = std::max(maxd, delta[s]);
  const size_t n_obs = maxd < 0 ? 0 : (size_t)maxd + 1;
  if (n_obs > 65535) {
    delete d;
    return set_error(ADSR_ERR_UNSUPPORTED,
                     "at most 65535 quadrupole observations are supported");
  }
  d->n_obs = n_obs;
  d->host_csr = host_csr;
  spectra.assign(n_obs, {});
  std::vector<uint64_t> count(n_obs + 1, 0);
  for (size_t s = 0; s < n_spectra; ++s) {
    if (delta[s] < 0) continue;
    const uint64_t a = (uint64_t)pstart[s], b = (uint64_t)pstop[s];
    if (a < b && b > n_peaks) {
      delete d;
      return set_error(ADSR_ERR_INVALID_ARGUMENT,
                       "spectrum %zu: peak range [%lld, %lld) exceeds %zu "
                       "peaks",
                       s, (long long)pstart[s], (long long)pstop[s], n_peaks);
    }
    spectra[(size_t)delta[s]].push_back(s);
    if (a < b) count[(size_t)delta[s] + 1] += b - a;
  }
  for (size_t o = 0; o < n_obs; ++o) count[o + 1] += count[o];
  const uint64_t total = count[n_obs];
  try {
    d->iso_lo.assign(n_obs, 0.0f);
    d->iso_hi.assign(n_obs, 0.0f);
    d->num_cycles.assign(n_obs, 0);
    d->entry_base = count;
    if (host_csr) {
      d->slice_starts.assign(n_obs * (B + 1), 0u);
      d->cycles.resize(total);
      d->intens.resize(total);
    }
  } catch (const std::bad_alloc &) {
    delete d;
    return set_error(ADSR_ERR_OUT_OF_MEMORY, "host allocation failed");
  }
  for (size_t o = 0; o < n_obs; ++o)
    if (count[o + 1] - count[o] > 0xFFFFFFFFull) {
      delete d;
      return set_error(ADSR_ERR_UNSUPPORTED,
                       "observation %zu holds more than 2^32-1 peaks (u32 "
                       "slice_starts)",
                       o);
    }
  // src/dia_data_builder/mod.rs:71-80: window bounds from the first
  // spectrum, number of distinct cycles; an empty window keeps [0, 0]
  parallel_for(n_obs, [&](size_t o) {
    const std::vector<size_t> &sp = spectra[o];
    if (sp.empty()) return;
    d->iso_lo[o] = iso_lo[sp[0]];
    d->iso_hi[o] = iso_hi[sp[0]];
    std::vector<int64_t> c(sp.size());
    for (size_t i = 0; i < sp.size(); ++i) c[i] = cyc[sp[i]];
    std::sort(c.begin(), c.end());
    d->num_cycles[o] = (uint64_t)(std::unique(c.begin(), c.end()) - c.begin());
  });
  *out = d;
  return ADSR_OK;
}

}  // namespace adsr

extern "C" adsr_status adsr_dia_create(
    const int64_t *delta, const float *iso_lo, const float *iso_hi,
    const int64_t *pstart, const int64_t *pstop, const int64_t *cyc,
    const float *srt, size_t n_spectra, const float *peak_mz,
    const float *peak_int, size_t n_peaks, const float *cycle,
    const size_t cycle_shape[4], adsr_dia **out) {
  clear_error();
  std::vector<std::vector<size_t>> spectra;
  adsr_dia *d = nullptr;
  adsr_status st = dia_prepare(delta, iso_lo, iso_hi, pstart, pstop, cyc, srt,
                               n_spectra, peak_mz, peak_int, n_peaks, cycle,
                               cycle_shape, true, spectra, &d);
  if (st) return st;
  const Grid &g = global_grid();
  const size_t B = g.mz.size();
  const std::vector<uint64_t> &count = d->entry_base;

  // src/dia_data_builder/mod.rs:52-139, one observation per task
  parallel_for(d->n_obs, [&](size_t o) {
    const std::vector<size_t> &sp = spectra[o];
    uint32_t *ss = d->slice_starts.data() + o * (B + 1);
    if (sp.empty()) return;  // :71-80 empty: zero slices
    const size_t nw = count[o + 1] - count[o];
    if (nw == 0) return;
    // (bin, cycle as u16, intensity) in spectrum/peak order (:97-109)
    std::vector<uint32_t> bin(nw);
    std::vector<uint16_t> cy(nw);
    bool cyc_sorted = true;
    uint16_t last = 0;
    size_t k = 0;
    for (size_t i = 0; i < sp.size(); ++i) {
      const size_t s = sp[i];
      const uint16_t c = (uint16_t)cyc[s];
      if (i && c < last) cyc_sorted = false;
      last = c;
      const uint64_t a = (uint64_t)pstart[s], b = (uint64_t)pstop[s];
      for (uint64_t p = a; p < b; ++p) {
        bin[k] = grid_find_closest(g, peak_mz[p]);
        cy[k] = c;
        ++k;
      }
    }
    // peak intensity source index of entry k
    std::vector<uint64_t> src(nw);
    k = 0;
    for (size_t i = 0; i < sp.size(); ++i) {
      const size_t s = sp[i];
      const uint64_t a = (uint64_t)pstart[s], b = (uint64_t)pstop[s];
      for (uint64_t p = a; p < b; ++p) src[k++] = p;
    }
    // stable sort by (bin, cycle) (:113): if cycles are already
    // non-decreasing in push order one stable counting pass by bin
    // suffices, otherwise a stable pass by cycle goes first (LSD order).
    std::vector<uint32_t> perm;
    if (!cyc_sorted) {
      std::vector<uint64_t> cc(65537, 0);
      for (size_t i = 0; i < nw; ++i) cc[(size_t)cy[i] + 1]++;
      for (size_t c = 0; c < 65536; ++c) cc[c + 1] += cc[c];
      perm.resize(nw);
      for (size_t i = 0; i < nw; ++i) perm[cc[cy[i]]++] = (uint32_t)i;
    }
    for (size_t i = 0; i < nw; ++i) ss[(size_t)bin[i] + 1]++;
    for (size_t b = 0; b < B; ++b) ss[b + 1] += ss[b];
    std::vector<uint32_t> cursor(ss, ss + B);
    uint16_t *oc = d->cycles.data() + count[o];
    float *oi = d->intens.data() + count[o];
    for (size_t j = 0; j < nw; ++j) {
      const size_t i = perm.empty() ? j : perm[j];
      const uint32_t pos = cursor[bin[i]]++;
      oc[pos] = cy[i];
      oi[pos] = peak_int[src[i]];
    }
  });
  *out = d;
  return ADSR_OK;
}

extern "C" void adsr_dia_free(adsr_dia *d) {
  if (!d) return;
  device_dia_free(d->dev);
  delete d;
}

extern "C" size_t adsr_dia_num_observations(const adsr_dia *d) {
  return d ? d->n_obs : 0;
}

// src/dia_data/mod.rs:79-87 (inclusive bounds, ascending order)
extern "C" adsr_status adsr_dia_valid_observations(const adsr_dia *d, float mz,
                                                   uint64_t *out, size_t cap,
                                                   size_t *n_out) {
  clear_error();
  if (!d || !n_out) return set_error(ADSR_ERR_INVALID_ARGUMENT, "NULL handle");
  size_t n = 0;
  for (size_t o = 0; o < d->n_obs; ++o)
    if (d->iso_lo[o] <= mz && d->iso_hi[o] >= mz) {
      if (out && n < cap) out[n] = o;
      ++n;
    }
  *n_out = n;
  return ADSR_OK;
}

// src/dia_data/mod.rs:90-108 and src/quadrupole_observation/mod.rs:110-128
extern "C" uint64_t adsr_dia_memory_footprint_bytes(const adsr_dia *d) {
  if (!d) return 0;
  uint64_t total = d->rt_index.size() * 4;
  total += 24;  // size_of::<Vec<QuadrupoleObservation>>()
  for (size_t o = 0; o < d->n_obs; ++o) {
    total += 8 + 8 + 3 * 24;  // [f32;2], usize, three Vec headers
    total += (uint64_t)(d->n_grid + 1) * 4;
    total += (d->entry_base[o + 1] - d->entry_base[o]) * (2 + 4);
  }
  return total;
}

extern "C" size_t adsr_dia_rt_values(const adsr_dia *d, float *out,
                                     size_t cap) {
  if (!d) return 0;
  if (out)
    std::memcpy(out, d->rt_values.data(),
                std::min(cap, d->rt_values.size()) * 4);
  return d->rt_values.size();
}

extern "C" void adsr_dia_cycle_shape(const adsr_dia *d, size_t shape[4]) {
  for (int i = 0; i < 4; ++i) shape[i] = d ? d->cycle_shape[i] : 0;
}

extern "C" void adsr_dia_cycle_copy(const adsr_dia *d, float *out) {
  if (d && out && !d->cycle.empty())
    std::memcpy(out, d->cycle.data(), d->cycle.size() * 4);
}

extern "C" size_t adsr_dia_rt_index(const adsr_dia *d, float *out, size_t cap,
                                    float *delta_t) {
  if (!d) return 0;
  if (out)
    std::memcpy(out, d->rt_index.data(), std::min(cap, d->rt_index.size()) * 4);
  if (delta_t) *delta_t = d->delta_t;
  return d->rt_index.size();
}

extern "C" adsr_status adsr_dia_observation(
    const adsr_dia *d, size_t o, float win[2], uint64_t *num_cycles,
    uint32_t *ss, size_t *n_ss, uint16_t *ci, float *in, size_t *n_entries) {
  clear_error();
  if (!d || o >= d->n_obs)
    return set_error(ADSR_ERR_INVALID_ARGUMENT, "observation out of range");
  if (win) {
    win[0] = d->iso_lo[o];
    win[1] = d->iso_hi[o];
  }
  if (num_cycles) *num_cycles = d->num_cycles[o];
  if (n_ss) *n_ss = d->n_grid + 1;
  const size_t a = d->entry_base[o], b = d->entry_base[o + 1];
  if (n_entries) *n_entries = b - a;
  if (!d->host_csr) return device_dia_window(d, o, ss, ci, in);
  if (ss)
    std::memcpy(ss, d->slice_starts.data() + o * (d->n_grid + 1),
                (d->n_grid + 1) * 4);
  if (ci) std::memcpy(ci, d->cycles.data() + a, (b - a) * 2);
  if (in) std::memcpy(in, d->intens.data() + a, (b - a) * 4);
  return ADSR_OK;
}

// ===========================================================================
// SpecLibFlat -- src/speclib_flat/mod.rs:98-199
// ===========================================================================
extern "C" adsr_status adsr_lib_create(
    size_t np, const uint64_t *pidx, const float *pmz_lib, const float *pmz,
    const float *prt_lib, const float *prt, const uint8_t *pnaa,
    const uint64_t *fstart, const uint64_t *fstop, size_t nf,
    const float *fmz_lib, const float *fmz, const float *fint,
    const uint8_t *fcard, const uint8_t *fcharge, const uint8_t *floss,
    const uint8_t *fnum, const uint8_t *fpos, const uint8_t *ftype,
    adsr_lib **out) {
  clear_error();
  if (!out) return set_error(ADSR_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  if (np && (!pidx || !pmz_lib || !pmz || !prt_lib || !prt || !pnaa ||
             !fstart || !fstop))
    return set_error(ADSR_ERR_INVALID_ARGUMENT, "precursor arrays are NULL");
  if (nf && (!fmz_lib || !fmz || !fint || !fcard || !fcharge || !floss ||
             !fnum || !fpos || !ftype))
    return set_error(ADSR_ERR_INVALID_ARGUMENT, "fragment arrays are NULL");
  // get_precursor_filtered slices fragment_*[start..stop]
  // (src/speclib_flat/mod.rs:401-412): reject what would panic there.
  for (size_t i = 0; i < np; ++i)
    if (fstart[i] > fstop[i] || fstop[i] > nf)
      return set_error(ADSR_ERR_INVALID_ARGUMENT,
                       "precursor %zu: fragment range [%llu, %llu) is "
                       "invalid for %zu fragments",
                       i, (unsigned long long)fstart[i],
                       (unsigned long long)fstop[i], nf);
  auto *l = new adsr_lib();
  l->n_prec = np;
  l->n_frag = nf;
  // :138-163 stable sort of the precursor arrays by precursor_idx
  std::vector<size_t> ord(np);
  for (size_t i = 0; i < np; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(),
                   [&](size_t a, size_t b) { return pidx[a] < pidx[b]; });
  auto gather = [&](auto &dst, const auto *src) {
    dst.resize(np);
    for (size_t i = 0; i < np; ++i) dst[i] = src[ord[i]];
  };
  gather(l->precursor_idx, pidx);
  gather(l->precursor_mz_library, pmz_lib);
  gather(l->precursor_mz, pmz);
  gather(l->precursor_rt_library, prt_lib);
  gather(l->precursor_rt, prt);
  gather(l->precursor_naa, pnaa);
  gather(l->frag_start, fstart);
  gather(l->frag_stop, fstop);
  auto copy = [&](auto &dst, const auto *src) {
    if (nf) dst.assign(src, src + nf);
  };
  copy(l->fragment_mz_library, fmz_lib);
  copy(l->fragment_mz, fmz);
  copy(l->fragment_intensity, fint);
  copy(l->fragment_cardinality, fcard);
  copy(l->fragment_charge, fcharge);
  copy(l->fragment_loss_type, floss);
  copy(l->fragment_number, fnum);
  copy(l->fragment_position, fpos);
  copy(l->fragment_type, ftype);
  for (size_t i = 0; i < np; ++i)
    l->max_raw_fragments =
        std::max<uint64_t>(l->max_raw_fragments, l->frag_stop[i] - l->frag_start[i]);
  {
    std::vector<std::pair<uint64_t, uint64_t>> r;
    r.reserve(np);
    for (size_t i = 0; i < np; ++i)
      if (l->frag_stop[i] > l->frag_start[i])
        r.emplace_back(l->frag_start[i], l->frag_stop[i]);
    std::sort(r.begin(), r.end());
    for (size_t i = 1; i < r.size(); ++i)
      if (r[i].first < r[i - 1].second) l->frag_ranges_disjoint = false;
  }
  *out = l;
  return ADSR_OK;
}

extern "C" void adsr_lib_free(adsr_lib *l) {
  if (!l) return;
  release_lib_workspaces(l);
  device_lib_free(l->dev);
  delete l;
}

extern "C" size_t adsr_lib_num_precursors(const adsr_lib *l) {
  return l ? l->n_prec : 0;
}
extern "C" size_t adsr_lib_num_fragments(const adsr_lib *l) {
  return l ? l->n_frag : 0;
}

// ===========================================================================
// CandidateCollection -- src/candidate/entry.rs:41-193
// ===========================================================================
extern "C" size_t adsr_candidates_len(const adsr_candidates *c) {
  return c ? c->n : 0;
}

extern "C" adsr_status adsr_candidates_export(
    const adsr_candidates *c, uint64_t *precursor_idx, uint64_t *rank,
    float *score, uint64_t *scan_center, uint64_t *scan_start,
    uint64_t *scan_stop, uint64_t *cycle_center, uint64_t *cycle_start,
    uint64_t *cycle_stop) {
  clear_error();
  if (!c) return set_error(ADSR_ERR_INVALID_ARGUMENT, "NULL handle");
  const size_t n = c->n;
  auto put = [n](auto *dst, const auto *src) {
    if (dst && n) std::memcpy(dst, src, n * sizeof(src[0]));
  };
  put(precursor_idx, c->precursor_idx);
  put(rank, c->rank);
  put(score, c->score);
  put(scan_center, c->scan_center);
  put(scan_start, c->scan_start);
  put(scan_stop, c->scan_stop);
  put(cycle_center, c->cycle_center);
  put(cycle_start, c->cycle_start);
  put(cycle_stop, c->cycle_stop);
  return ADSR_OK;
}

extern "C" adsr_status adsr_candidates_from_arrays(
    size_t n, const uint64_t *precursor_idx, const uint64_t *rank,
    const float *score, const uint64_t *scan_center, const uint64_t *scan_start,
    const uint64_t *scan_stop, const uint64_t *cycle_center,
    const uint64_t *cycle_start, const uint64_t *cycle_stop,
    adsr_candidates **out) {
  clear_error();
  if (!out) return set_error(ADSR_ERR_INVALID_ARGUMENT, "out is NULL");
  if (n && (!precursor_idx || !rank || !score || !scan_center || !scan_start ||
            !scan_stop || !cycle_center || !cycle_start || !cycle_stop))
    return set_error(ADSR_ERR_INVALID_ARGUMENT, "column arrays are NULL");
  void *blk = alloc_should_fail()
                  ? nullptr
                  : std::malloc(std::max<size_t>(n * adsr_candidates::kRowBytes, 16));
  if (!blk) return set_error(ADSR_ERR_OUT_OF_MEMORY, "candidate table");
  auto *c = new adsr_candidates();
  c->bind(blk, n);
  const uint64_t *src[8] = {precursor_idx, rank, scan_center, scan_start,
                            scan_stop, cycle_center, cycle_start, cycle_stop};
  uint64_t *dst[8] = {c->precursor_idx, c->rank, c->scan_center, c->scan_start,
                      c->scan_stop, c->cycle_center, c->cycle_start, c->cycle_stop};
  if (n) {
    for (int i = 0; i < 8; ++i) std::memcpy(dst[i], src[i], n * 8);
    std::memcpy(c->score, score, n * 4);
  }
  *out = c;
  return ADSR_OK;
}

extern "C" void adsr_candidates_free(adsr_candidates *c) { delete c; }
