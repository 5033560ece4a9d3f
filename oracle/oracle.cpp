/*
 * oracle.cpp - CPU restatement of peasoup's per-DM-trial acceleration search.
 *
 * TEST INFRASTRUCTURE ONLY: the checker the HIP product is compared against.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Reference files restated (paths relative to the peasoup source tree):
 *   include/utils/utils.hpp:143-233          AccelerationPlan
 *   src/kernels.cu:34-100                    harmonic_sum_kernel
 *   src/kernels.cu:216-259                   power_series / bin_interbin kernels
 *   src/kernels.cu:315-386                   resample / resampleII
 *   src/kernels.cu:391-423                   device_find_peaks
 *   src/kernels.cu:427-521                   GPU_rms, GPU_mean, remove_baseline, normalise
 *   src/kernels.cu:902-1061                  median_scrunch5, linear_stretch, divide_c_by_f
 *   include/transforms/{resampler,ffter,spectrumformer,dereddener,harmonicfolder}.hpp
 *   include/transforms/peakfinder.hpp:27-94  PeakFinder
 *   include/transforms/distiller.hpp:26-196  Harmonic/Acceleration/DM distillers
 *   src/pipeline_multi.cu:155-295            Worker call order
 *
 * The FFT (cuFFT in the reference, ffter.hpp:54-110) is restated as an exact
 * double-precision DFT (mixed radix 4/2/3 Stockham), whose outputs are rounded
 * to float where the reference hands float spectra on.  For the timed CPU
 * baseline only (orc_search_params.fast_fft), a float four-step FFT
 * (fastfft.cpp) takes its place, as FFTW3f would on the reference's host.
 *
 * Build: oracle/Makefile (g++ -O3 -march=x86-64-v3 -ffp-contract=off).
 */
#include "oracle.h"

#include "fastfft.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

typedef std::complex<double> cd;

/* ------------------------------------------------------------------------- */
/* AccelerationPlan: include/utils/utils.hpp:143-233                         */
/* ------------------------------------------------------------------------- */
namespace {

struct AccPlan {
  float acc_lo, acc_hi, tol, pulse_width;
  uint64_t nsamps;
  float tsamp, cfreq, ch_bw;

  /* utils.hpp:163-173 */
  float dispersive_smear(float dm, float cdm) const {
    float ftop = (cfreq + ch_bw / 2) / 1e9;
    float fbottom = (cfreq - ch_bw / 2) / 1e9;
    float dm_diff = std::abs(cdm - dm);
    return 4.15e-3 * (1 / (fbottom * fbottom) - 1 / (ftop * ftop)) * dm_diff;
  }
  /* utils.hpp:175-183: float sqrt of a float sum */
  float pulse_broadening(float dm, float cdm) const {
    float tdm = dispersive_smear(dm, cdm);
    return std::sqrt((float)((tdm * tdm) + (pulse_width * pulse_width) + (tsamp * tsamp)));
  }
  /* utils.hpp:185-196: (tol*tol) - 1.0 is a double expression */
  float accel_step(float dm, float cdm) const {
    float broadening = pulse_broadening(dm, cdm);
    float factor = std::sqrt((double)(tol * tol) - 1.0);
    float tobs = (float)nsamps * tsamp;
    return 2.0 * broadening * 24.0 * 299792458.0 / (double)(tobs * tobs) * factor;
  }
  /* utils.hpp:198-232 */
  void generate(float dm, float cdm, std::vector<float>& acc_list) const {
    acc_list.clear();
    if (acc_hi == acc_lo) {
      acc_list.push_back(acc_hi);
      return;
    }
    if ((acc_hi >= 0) && (acc_lo <= 0)) acc_list.push_back(0.0f);
    float step = accel_step(dm, cdm);
    float acc = acc_lo;
    while (acc <= acc_hi) {
      acc_list.push_back(acc);
      acc += step;
    }
    acc_list.push_back(acc_hi);
    std::sort(acc_list.begin(), acc_list.end());
    acc_list.erase(std::unique(acc_list.begin(), acc_list.end()), acc_list.end());
  }
};

}  // namespace

extern "C" int orc_accel_list(float acc_lo, float acc_hi, float tol, float pulse_width,
                              uint64_t nsamps, float tsamp, float cfreq, float ch_bw,
                              float dm, float cdm, float* out, int cap) {
  AccPlan p{acc_lo, acc_hi, tol, pulse_width, nsamps, tsamp, cfreq, std::fabs(ch_bw)};
  std::vector<float> v;
  p.generate(dm, cdm, v);
  int n = (int)v.size();
  if (out) for (int i = 0; i < n && i < cap; ++i) out[i] = v[i];
  return n;
}

extern "C" float orc_accel_step(float acc_lo, float acc_hi, float tol, float pulse_width,
                                uint64_t nsamps, float tsamp, float cfreq, float ch_bw,
                                float dm, float cdm) {
  AccPlan p{acc_lo, acc_hi, tol, pulse_width, nsamps, tsamp, cfreq, std::fabs(ch_bw)};
  return p.accel_step(dm, cdm);
}

/* ------------------------------------------------------------------------- */
/* Resampling: src/kernels.cu:315-386                                        */
/* ------------------------------------------------------------------------- */
/* device_resampleII, kernels.cu:355-361: the float product a*tsamp is
 * widened to double before the divide by 2c. */
static inline double accel_fact(float acc, float tsamp) {
  float at = acc * tsamp;
  return (double)at / (2 * 299792458.0);
}

/* getAcceleratedIndexII, kernels.cu:321-324: id + id*af*(id-size), rounded to
 * nearest-even.  nvcc contracts the multiply-add into one fma (no
 * -fmad=false in the reference Makefile), restated with an explicit fma. */
extern "C" uint64_t orc_resample2_index(uint64_t i, uint64_t n, float acc, float tsamp) {
  double af = accel_fact(acc, tsamp);
  double di = (double)i;
  double v = std::fma(di * af, di - (double)n, di);
  return (uint64_t)std::nearbyint(v);
}

/* The same index without contraction, used only to count how many indices the
 * contraction choice can move. */
extern "C" uint64_t orc_resample2_index_nofma(uint64_t i, uint64_t n, float acc, float tsamp) {
  double af = accel_fact(acc, tsamp);
  double di = (double)i;
  volatile double prod = (di * af) * (di - (double)n);
  double v = di + prod;
  return (uint64_t)std::nearbyint(v);
}

/* resample_kernelII, kernels.cu:342-353 */
extern "C" void orc_resample2(const float* in, float* out, uint64_t n, float acc, float tsamp) {
  double af = accel_fact(acc, tsamp);
  double dn = (double)n;
  for (uint64_t i = 0; i < n; ++i) {
    double di = (double)i;
    uint64_t j = (uint64_t)std::nearbyint(std::fma(di * af, di - dn, di));
    out[i] = in[j];
  }
}

/* resample_kernel + getAcceleratedIndex, kernels.cu:315-318,327-339:
 * id + af*((id - n/2)^2 - (n/2)^2).  Contraction restated as
 * fma(af, fma(t, t, -(h*h)), id) with t = id - n/2, h = n/2. */
extern "C" void orc_resample1(const float* in, float* out, uint64_t n, float acc, float tsamp) {
  double af = accel_fact(acc, tsamp);
  double h = (double)n / 2.0;
  double hh = h * h;
  for (uint64_t i = 0; i < n; ++i) {
    double di = (double)i;
    double t = di - h;
    double q = std::fma(t, t, -hh);
    uint64_t j = (uint64_t)std::nearbyint(std::fma(af, q, di));
    out[i] = in[j];
  }
}

/* ------------------------------------------------------------------------- */
/* FFT: restates cuFFT R2C/C2R (ffter.hpp:54-110) as an exact double DFT.    */
/* ------------------------------------------------------------------------- */
namespace {

struct CPlan {
  uint64_t m = 0;
  std::vector<int> radices;
  std::vector<cd> tw;  // exp(-2 pi i t / m), t in [0, m)
};

struct RPlan {
  uint64_t n = 0;      // real length
  CPlan* cplan = nullptr;
  std::vector<cd> wn;  // exp(-2 pi i k / n), k in [0, n/2]
};

std::mutex g_plan_mu;
std::map<uint64_t, std::unique_ptr<CPlan>> g_cplans;
std::map<uint64_t, std::unique_ptr<RPlan>> g_rplans;

inline cd unit_root(uint64_t t, uint64_t m) {
  /* exp(-2 pi i t/m), evaluated in long double so that every entry is the
   * double nearest the exact root. */
  long double a = (long double)t / (long double)m;  // turns
  long double ang = -2.0L * 3.14159265358979323846264338327950288L * a;
  return cd((double)cosl(ang), (double)sinl(ang));
}

bool factorize(uint64_t m, std::vector<int>& r) {
  r.clear();
  while (m % 3 == 0) { r.push_back(3); m /= 3; }
  while (m % 4 == 0) { r.push_back(4); m /= 4; }
  while (m % 2 == 0) { r.push_back(2); m /= 2; }
  return m == 1;
}

CPlan* get_cplan(uint64_t m) {
  std::lock_guard<std::mutex> g(g_plan_mu);
  auto it = g_cplans.find(m);
  if (it != g_cplans.end()) return it->second.get();
  std::unique_ptr<CPlan> p(new CPlan);
  p->m = m;
  if (!factorize(m, p->radices)) return nullptr;
  p->tw.resize(m);
  for (uint64_t t = 0; t < m; ++t) p->tw[t] = unit_root(t, m);
  CPlan* raw = p.get();
  g_cplans[m] = std::move(p);
  return raw;
}

RPlan* get_rplan(uint64_t n) {
  if (n < 2 || (n & 1)) return nullptr;
  CPlan* cp = get_cplan(n / 2);
  if (!cp) return nullptr;
  std::lock_guard<std::mutex> g(g_plan_mu);
  auto it = g_rplans.find(n);
  if (it != g_rplans.end()) return it->second.get();
  std::unique_ptr<RPlan> p(new RPlan);
  p->n = n;
  p->cplan = cp;
  p->wn.resize(n / 2 + 1);
  for (uint64_t k = 0; k <= n / 2; ++k) p->wn[k] = unit_root(k, n);
  RPlan* raw = p.get();
  g_rplans[n] = std::move(p);
  return raw;
}

/* One Stockham autosort stage of radix R over the whole array. */
void stockham_stage(const cd* in, cd* out, uint64_t m, int R, uint64_t Ns, const cd* tw, int sign) {
  const uint64_t nbf = m / R;
  const uint64_t tstep = m / (Ns * R);
  for (uint64_t g = 0; g < nbf / Ns; ++g) {
    for (uint64_t k = 0; k < Ns; ++k) {
      const uint64_t j = g * Ns + k;
      cd v[4];
      for (int r = 0; r < R; ++r) v[r] = in[j + r * nbf];
      for (int r = 1; r < R; ++r) {
        cd w = tw[k * r * tstep];
        if (sign > 0) w = std::conj(w);
        v[r] *= w;
      }
      if (R == 2) {
        cd a = v[0] + v[1], b = v[0] - v[1];
        v[0] = a; v[1] = b;
      } else if (R == 4) {
        cd a0 = v[0] + v[2], a1 = v[0] - v[2], a2 = v[1] + v[3], d = v[1] - v[3];
        cd a3 = (sign < 0) ? cd(d.imag(), -d.real()) : cd(-d.imag(), d.real());
        v[0] = a0 + a2; v[2] = a0 - a2; v[1] = a1 + a3; v[3] = a1 - a3;
      } else {  // R == 3
        const double s3 = 0.86602540378443864676372317075294;
        cd t1 = v[1] + v[2], t2 = v[1] - v[2];
        cd base = v[0] - 0.5 * t1;
        cd rot = cd(-t2.imag(), t2.real()) * (sign * s3);  // i*sign*s3*t2
        v[0] = v[0] + t1;
        v[1] = base + rot;
        v[2] = base - rot;
      }
      const uint64_t idxD = g * Ns * R + k;
      for (int r = 0; r < R; ++r) out[idxD + r * Ns] = v[r];
    }
  }
}

/* In-place complex FFT using a caller-provided work buffer of m entries. */
void cfft(CPlan* p, cd* data, cd* work, int sign) {
  cd* in = data;
  cd* out = work;
  uint64_t Ns = 1;
  for (int R : p->radices) {
    stockham_stage(in, out, p->m, R, Ns, p->tw.data(), sign);
    Ns *= R;
    std::swap(in, out);
  }
  if (in != data) std::memcpy(data, in, sizeof(cd) * p->m);
}

/* R2C: half-length complex FFT + untangle.  out has n/2+1 entries. */
void rfft(RPlan* rp, const float* x, const double* xd, cd* out, std::vector<cd>& z,
          std::vector<cd>& work) {
  const uint64_t m = rp->n / 2;
  z.resize(m);
  work.resize(m);
  for (uint64_t i = 0; i < m; ++i)
    z[i] = xd ? cd(xd[2 * i], xd[2 * i + 1]) : cd((double)x[2 * i], (double)x[2 * i + 1]);
  cfft(rp->cplan, z.data(), work.data(), -1);
  for (uint64_t k = 0; k <= m; ++k) {
    cd zk = z[k % m];
    cd zm = std::conj(z[(m - k) % m]);
    cd e = 0.5 * (zk + zm);
    cd d = zk - zm;
    cd o = cd(0.5 * d.imag(), -0.5 * d.real());  // -i/2 * d
    out[k] = e + rp->wn[k] * o;
  }
}

/* C2R, unnormalised (output = n * x for X = R2C(x)). */
void irfft(RPlan* rp, const float* X, float* out, std::vector<cd>& z, std::vector<cd>& work) {
  const uint64_t m = rp->n / 2;
  z.resize(m);
  work.resize(m);
  for (uint64_t k = 0; k < m; ++k) {
    cd xk((double)X[2 * k], (double)X[2 * k + 1]);
    cd xm((double)X[2 * (m - k)], -(double)X[2 * (m - k) + 1]);
    cd s = xk + xm;
    cd d = (xk - xm) * std::conj(rp->wn[k]);
    z[k] = s + cd(-d.imag(), d.real());  // s + i*d
  }
  cfft(rp->cplan, z.data(), work.data(), +1);
  for (uint64_t i = 0; i < m; ++i) {
    out[2 * i] = (float)z[i].real();
    out[2 * i + 1] = (float)z[i].imag();
  }
}

}  // namespace

extern "C" int orc_fft_c2c_f64(double* data, uint64_t m, int sign) {
  CPlan* p = get_cplan(m);
  if (!p) return -1;
  std::vector<cd> buf(m), work(m);
  for (uint64_t i = 0; i < m; ++i) buf[i] = cd(data[2 * i], data[2 * i + 1]);
  cfft(p, buf.data(), work.data(), sign < 0 ? -1 : 1);
  for (uint64_t i = 0; i < m; ++i) {
    data[2 * i] = buf[i].real();
    data[2 * i + 1] = buf[i].imag();
  }
  return 0;
}

extern "C" int orc_fft_r2c(const float* in, float* out, uint64_t n) {
  RPlan* rp = get_rplan(n);
  if (!rp) return -1;
  std::vector<cd> z, work, X(n / 2 + 1);
  rfft(rp, in, nullptr, X.data(), z, work);
  for (uint64_t k = 0; k <= n / 2; ++k) {
    out[2 * k] = (float)X[k].real();
    out[2 * k + 1] = (float)X[k].imag();
  }
  return 0;
}

extern "C" int orc_fft_r2c_f64(const double* in, double* out, uint64_t n) {
  RPlan* rp = get_rplan(n);
  if (!rp) return -1;
  std::vector<cd> z, work, X(n / 2 + 1);
  rfft(rp, nullptr, in, X.data(), z, work);
  for (uint64_t k = 0; k <= n / 2; ++k) {
    out[2 * k] = X[k].real();
    out[2 * k + 1] = X[k].imag();
  }
  return 0;
}

extern "C" int orc_fft_c2r(const float* in, float* out, uint64_t n) {
  RPlan* rp = get_rplan(n);
  if (!rp) return -1;
  std::vector<cd> z, work;
  irfft(rp, in, out, z, work);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Spectrum formation: src/kernels.cu:216-259                                */
/* ------------------------------------------------------------------------- */
/* power_series_kernel (way 0): amplitude |X| (z*rsqrtf(z) in the reference,
 * restated as a correctly rounded sqrtf; z = re*re + im*im contracted to
 * fma(re, re, im*im) as nvcc does for device code). */
extern "C" void orc_form_amplitude(const float* X, float* A, uint64_t nb) {
  for (uint64_t i = 0; i < nb; ++i) {
    float re = X[2 * i], im = X[2 * i + 1];
    float z = std::fma(re, re, im * im);
    A[i] = (z == 0.0f) ? 0.0f : std::sqrt(z);
  }
}

/* bin_interbin_series_kernel (way 1): sqrt(max(|X_i|^2, 0.5*|X_i - X_{i-1}|^2)),
 * X_{-1} = 0.  0.5 is a double literal in the reference: the half is taken in
 * double and rounded back to float (exact). */
extern "C" void orc_form_interbin(const float* X, float* A, uint64_t nb) {
  for (uint64_t i = 0; i < nb; ++i) {
    float re_l = 0.0f, im_l = 0.0f;
    if (i > 0) {
      re_l = X[2 * i - 2];
      im_l = X[2 * i - 1];
    }
    float re = X[2 * i], im = X[2 * i + 1];
    float ampsq = std::fma(re, re, im * im);
    float dr = re - re_l, di = im - im_l;
    float ampsq_diff = (float)(0.5 * (double)std::fma(dr, dr, di * di));
    A[i] = std::sqrt(std::fmax(ampsq, ampsq_diff));
  }
}

/* ------------------------------------------------------------------------- */
/* Dereddening: src/kernels.cu:902-1061, dereddener.hpp:41-66                */
/* ------------------------------------------------------------------------- */
static inline float med3(float a, float b, float c) {
  float v[3] = {a, b, c};
  std::sort(v, v + 3);
  return v[1];
}
static inline float med4(float a, float b, float c, float d) {
  float v[4] = {a, b, c, d};
  std::sort(v, v + 4);
  return 0.5f * (v[1] + v[2]);
}
static inline float med5(const float* p) {
  float v[5] = {p[0], p[1], p[2], p[3], p[4]};
  std::sort(v, v + 5);
  return v[2];
}

/* median_scrunch5, kernels.cu:974-1008: counts 1..4 are special cases,
 * otherwise out[i] = median(in[5i..5i+4]) for i < count/5. */
extern "C" uint64_t orc_median_scrunch5(const float* in, uint64_t count, float* out) {
  if (count == 1) { out[0] = in[0]; return 1; }
  if (count == 2) { out[0] = 0.5f * (in[0] + in[1]); return 1; }
  if (count == 3) { out[0] = med3(in[0], in[1], in[2]); return 1; }
  if (count == 4) { out[0] = med4(in[0], in[1], in[2], in[3]); return 1; }
  uint64_t oc = count / 5;
  for (uint64_t i = 0; i < oc; ++i) out[i] = med5(in + 5 * i);
  return oc;
}

/* linear_stretch_functor, kernels.cu:1010-1038: step = float(in-1)/(out-1);
 * x = float(i)*step; j = (unsigned)x; in[j] + ((x-j > 1e-5f) ? (x-j)*(in[j+1]-in[j]) : 0).
 * The multiply-add is contracted by nvcc (device functor), restated as fma.
 * When x rounds past in_count-1 the reference reads in[in_count]
 * (out of bounds); the restatement clamps that neighbour to in[in_count-1]. */
extern "C" void orc_linear_stretch(const float* in, uint64_t in_count, float* out, uint64_t out_count) {
  float step = (float)(in_count - 1) / (float)(out_count - 1);
  for (uint64_t i = 0; i < out_count; ++i) {
    float x = (float)(unsigned)i * step;
    unsigned j = (unsigned)x;
    float fr = x - (float)j;
    float a = in[j];
    float b = (j + 1 < in_count) ? in[j + 1] : in[in_count - 1];
    out[i] = (fr > 1e-5f) ? std::fma(fr, b - a, a) : a + 0.f;
  }
}

/* Dereddener::calculate_median, dereddener.hpp:41-62.  Zone copies: the second
 * copy length is pos25 (:58), which the third copy overwrites from pos25 on,
 * so the effective zones are [0,pos5) [pos5,pos25) [pos25,nb). */
extern "C" void orc_calculate_median(const float* A, uint64_t nb, double bin_width,
                                     float b5, float b25, float* median) {
  int pos5 = (int)((double)b5 / bin_width);
  int pos25 = (int)((double)b25 / bin_width);
  std::vector<float> m5(nb / 5 + 1), m25(nb / 25 + 1), m125(nb / 125 + 1), inter(nb);
  orc_median_scrunch5(A, nb, m5.data());
  orc_median_scrunch5(m5.data(), nb / 5, m25.data());
  orc_median_scrunch5(m25.data(), nb / 5 / 5, m125.data());
  orc_linear_stretch(m5.data(), nb / 5, inter.data(), nb);
  for (int i = 0; i < pos5; ++i) median[i] = inter[i];
  orc_linear_stretch(m25.data(), nb / 5 / 5, inter.data(), nb);
  for (uint64_t i = pos5; i < (uint64_t)pos5 + (uint64_t)pos25 && i < nb; ++i) median[i] = inter[i];
  orc_linear_stretch(m125.data(), nb / 5 / 5 / 5, inter.data(), nb);
  for (uint64_t i = pos25; i < nb; ++i) median[i] = inter[i];
}

/* divide_c_by_f_kernel, kernels.cu:1040-1049: bins < 5 set to 0, otherwise
 * X/f (cuCdivf in the reference, restated as two IEEE divisions). */
extern "C" void orc_divide_c_by_f(float* X, const float* f, uint64_t nb) {
  for (uint64_t i = 0; i < nb; ++i) {
    if (i < 5) {
      X[2 * i] = 0.0f;
      X[2 * i + 1] = 0.0f;
    } else {
      X[2 * i] = X[2 * i] / f[i];
      X[2 * i + 1] = X[2 * i + 1] / f[i];
    }
  }
}

/* zap_birdies_kernel, kernels.cu:1063-1084: for each birdie, bins from
 * __float2int_rd((f-w)/bw) up to (excluding) __float2int_ru((f+w)/bw), with
 * the low end raised to 0, birdies starting at or past nb skipped and the
 * high end lowered to nb-1, are set to 1+0i. */
extern "C" void orc_zap_birdies(float* X, uint64_t nb, float bin_width, const float* freqs, const float* widths,
                                int n) {
  for (int b = 0; b < n; ++b) {
    float freq = freqs[b];
    float width = widths[b];
    long low_bin = (long)std::floor((freq - width) / bin_width);
    long high_bin = (long)std::ceil((freq + width) / bin_width);
    if (low_bin < 0) low_bin = 0;
    if (low_bin >= (long)nb) continue;
    if (high_bin >= (long)nb) high_bin = (long)nb - 1;
    for (long ii = low_bin; ii < high_bin; ii++) {
      X[2 * ii] = 1.0f;
      X[2 * ii + 1] = 0.0f;
    }
  }
}

/* ------------------------------------------------------------------------- */
/* Statistics: stats.hpp:6-40, kernels.cu:427-521                            */
/* ------------------------------------------------------------------------- */
/* GPU_mean / GPU_rms accumulate float terms; restated with a double
 * accumulator over the same float terms (x, and x*x rounded to float). */
static float mean_of(const float* p, uint64_t n, uint64_t first) {
  double s = 0.0;
  for (uint64_t i = first; i < n; ++i) s += (double)p[i];
  return (float)s / (float)(n - first);
}
static float rms_of(const float* p, uint64_t n, uint64_t first) {
  double s = 0.0;
  for (uint64_t i = first; i < n; ++i) {
    float sq = p[i] * p[i];
    s += (double)sq;
  }
  return std::sqrt((float)s / (float)(n - first));
}

extern "C" void orc_stats(const float* p, uint64_t n, uint64_t first, float* mean, float* rms, float* std_) {
  float r = rms_of(p, n, first);
  float m = mean_of(p, n, first);
  *rms = r;
  *mean = m;
  *std_ = std::sqrt(r * r - m * m);
}

/* normalisation_kernel, kernels.cu:496-507 */
extern "C" void orc_normalise(float* p, float mean, float sigma, uint64_t n) {
  for (uint64_t i = 0; i < n; ++i) {
    float v = p[i];
    v -= mean;
    v /= sigma;
    p[i] = v;
  }
}

/* GPU_remove_baseline, kernels.cu:463-479 */
extern "C" void orc_remove_baseline(float* x, uint64_t n) {
  float mean;
  int count = 0;
  do {
    mean = mean_of(x, n, 0);
    for (uint64_t i = 0; i < n; ++i) x[i] -= mean;
    if (count++ > 100) break;
  } while (std::fabs(mean) > 5e-7 * (double)n);
}

/* ------------------------------------------------------------------------- */
/* Harmonic summing: src/kernels.cu:34-100                                   */
/* ------------------------------------------------------------------------- */
/* Fold k (k = 0..4) adds P[(int)(i*j/2^(k+1) + 0.5)] for odd j, in the order
 * the reference lists them, to the running float sum, and stores the sum
 * times 2^-(k+1)/2 computed in double and rounded to float. */
static const int kFoldOrder1[2] = {3, 1};  // 0.75 then 0.25 (kernels.cu:52-53)
static const double kFoldScale[5] = {0.70710678118654752440, 0.5, 0.35355339059327376220, 0.25,
                                     0.17677669529663688110};

extern "C" void orc_harmonic_sum(const float* P, float* folds, uint64_t nb, unsigned nfolds) {
  if (nfolds > 5) nfolds = 5;
  for (uint64_t i = 0; i < nb; ++i) {
    float val = P[i];
    for (unsigned k = 0; k < nfolds; ++k) {
      const unsigned m = k + 1;
      const uint64_t half = 1ull << (m - 1);
      if (k == 1) {
        for (int t = 0; t < 2; ++t) val += P[(i * kFoldOrder1[t] + half) >> m];
      } else {
        for (uint64_t j = 1; j < (1ull << m); j += 2) val += P[(i * j + half) >> m];
      }
      folds[k * nb + i] = (float)((double)val * kFoldScale[k]);
    }
  }
}

/* ------------------------------------------------------------------------- */
/* Peak finding: kernels.cu:391-423, peakfinder.hpp:27-94                    */
/* ------------------------------------------------------------------------- */
/* device_find_peaks: stable compaction of (i, v) with v > thr, i in [start, n). */
extern "C" int orc_threshold(const float* p, int n, int start, float thr, int* idx, float* snr, int cap) {
  int c = 0;
  for (int i = std::max(start, 0); i < n; ++i) {
    if (p[i] > thr) {
      if (c < cap) {
        idx[c] = i;
        snr[c] = p[i];
      }
      ++c;
    }
  }
  return c;
}

/* identify_unique_peaks, peakfinder.hpp:27-56: a cluster continues while the
 * next index is within min_gap of the last index that set a new maximum. */
static int unique_peaks(const int* idxs, const float* snrs, int count, int min_gap, int* pidx, float* psnr) {
  int ii = 0, npeaks = 0;
  while (ii < count) {
    float cpeak = snrs[ii];
    int cpeakidx = idxs[ii];
    int lastidx = idxs[ii];
    ii++;
    while (ii < count && (idxs[ii] - lastidx) < min_gap) {
      if (snrs[ii] > cpeak) {
        cpeak = snrs[ii];
        cpeakidx = idxs[ii];
        lastidx = idxs[ii];
      }
      ii++;
    }
    pidx[npeaks] = cpeakidx;
    psnr[npeaks] = cpeak;
    npeaks++;
  }
  return npeaks;
}

/* PeakFinder::find_candidates, peakfinder.hpp:77-94 (max_cands = 100000). */
extern "C" int orc_find_candidates(const float* p, uint32_t nbins, double bin_width, uint32_t nh,
                                   float thr, float fmin, float fmax, int min_gap,
                                   int* peak_idx, float* peak_snr, float* peak_freq, int cap) {
  const int max_cands = 100000;
  int size = (int)nbins;
  float nyquist = bin_width * size;
  int orig_size = 2.0 * (size - 1.0);
  int max_bin = (int)((fmax / bin_width) * std::pow(2.0, (int)nh));
  int start_idx = (int)((double)((float)orig_size * (fmin / nyquist)) * std::pow(2.0, (int)nh));
  int n = std::min(size, max_bin);
  std::vector<int> idxs(max_cands);
  std::vector<float> snrs(max_cands);
  int count = (n > start_idx) ? orc_threshold(p, n, start_idx, thr, idxs.data(), snrs.data(), max_cands) : 0;
  if (count > max_cands) return -1;  // reference host buffers would overflow
  std::vector<int> pidx(count + 1);
  std::vector<float> psnr(count + 1);
  int npeaks = unique_peaks(idxs.data(), snrs.data(), count, min_gap, pidx.data(), psnr.data());
  float factor = 1.0 / size * nyquist / std::pow(2.0, (double)(float)nh);
  if (npeaks > cap) return -2;
  for (int i = 0; i < npeaks; ++i) {
    peak_idx[i] = pidx[i];
    peak_snr[i] = psnr[i];
    peak_freq[i] = pidx[i] * factor;
  }
  return npeaks;
}

/* ------------------------------------------------------------------------- */
/* Distillers: include/transforms/distiller.hpp:16-196                       */
/* ------------------------------------------------------------------------- */
namespace {

struct Cand {
  float dm;
  int dm_idx;
  float acc;
  int nh;
  float snr;
  float freq;
  std::vector<Cand> assoc;
  int count_assoc() const {
    int c = 0;
    for (const Cand& a : assoc) c += 1 + a.count_assoc();
    return c;
  }
};

struct SnrGreater {
  bool operator()(const Cand& x, const Cand& y) const { return x.snr > y.snr; }
};

/* BaseDistiller::distill, distiller.hpp:26-59 */
template <class Cond>
std::vector<Cand> distill(std::vector<Cand>& cands, Cond cond) {
  int size = (int)cands.size();
  std::vector<bool> unique(size, true);
  std::sort(cands.begin(), cands.end(), SnrGreater());
  int start = 0;
  while (true) {
    int idx = -1;
    for (int ii = start; ii < size; ii++) {
      if (unique[ii]) {
        start = ii + 1;
        idx = ii;
        break;
      }
    }
    if (idx == -1) break;
    cond(cands, idx, unique, size);
  }
  std::vector<Cand> out;
  for (int ii = 0; ii < size; ii++)
    if (unique[ii]) out.push_back(cands[ii]);
  return out;
}

/* HarmonicDistiller::condition, distiller.hpp:68-100 */
std::vector<Cand> harmonic_distill(std::vector<Cand>& c, float tolerance, float max_harm, bool keep_related,
                                   bool fractional) {
  return distill(c, [&](std::vector<Cand>& cands, int idx, std::vector<bool>& unique, int size) {
    double upper_tol = 1 + tolerance;
    double lower_tol = 1 - tolerance;
    double fundi_freq = cands[idx].freq;
    for (int ii = idx + 1; ii < size; ii++) {
      double freq = cands[ii].freq;
      int nh = cands[ii].nh;
      float max_denominator = fractional ? (float)std::pow(2.0, nh) : 1.0f;
      for (int jj = 1; jj <= max_harm; jj++) {
        for (int kk = 1; kk <= max_denominator; kk++) {
          double ratio = kk * freq / (jj * fundi_freq);
          if (ratio > lower_tol && ratio < upper_tol) {
            if (keep_related) cands[idx].assoc.push_back(cands[ii]);
            unique[ii] = false;
          }
        }
      }
    }
  });
}

/* AccelerationDistiller::condition, distiller.hpp:115-163; the accel
 * correction returns float (:121-123). */
std::vector<Cand> accel_distill(std::vector<Cand>& c, float tobs, float tolerance, bool keep_related) {
  double tobs_over_c = tobs / 299792458.0;
  return distill(c, [&](std::vector<Cand>& cands, int idx, std::vector<bool>& unique, int size) {
    double fundi_freq = cands[idx].freq;
    double fundi_acc = cands[idx].acc;
    double edge = fundi_freq * tolerance;
    for (int ii = idx + 1; ii < size; ii++) {
      double delta_acc = fundi_acc - cands[ii].acc;
      double acc_freq = (float)(fundi_freq + delta_acc * fundi_freq * tobs_over_c);
      if (acc_freq > fundi_freq) {
        if (cands[ii].freq > fundi_freq - edge && cands[ii].freq < acc_freq + edge) {
          if (keep_related) cands[idx].assoc.push_back(cands[ii]);
          unique[ii] = false;
        }
      } else {
        if (cands[ii].freq < fundi_freq + edge && cands[ii].freq > acc_freq - edge) {
          if (keep_related) cands[idx].assoc.push_back(cands[ii]);
          unique[ii] = false;
        }
      }
    }
  });
}

/* DMDistiller::condition, distiller.hpp:167-189 */
std::vector<Cand> dm_distill(std::vector<Cand>& c, float tolerance, bool keep_related) {
  return distill(c, [&](std::vector<Cand>& cands, int idx, std::vector<bool>& unique, int size) {
    double fundi_freq = cands[idx].freq;
    double upper_tol = 1 + tolerance;
    double lower_tol = 1 - tolerance;
    for (int ii = idx + 1; ii < size; ii++) {
      double ratio = cands[ii].freq / fundi_freq;
      if (ratio > lower_tol && ratio < upper_tol) {
        if (keep_related) cands[idx].assoc.push_back(cands[ii]);
        unique[ii] = false;
      }
    }
  });
}

std::vector<Cand> from_flat(const orc_cand* in, int n) {
  std::vector<Cand> v(n);
  for (int i = 0; i < n; ++i) {
    v[i].dm = in[i].dm;
    v[i].dm_idx = in[i].dm_idx;
    v[i].acc = in[i].acc;
    v[i].nh = in[i].nh;
    v[i].snr = in[i].snr;
    v[i].freq = in[i].freq;
  }
  return v;
}

int to_flat(const std::vector<Cand>& v, orc_cand* out, int cap) {
  int n = (int)v.size();
  for (int i = 0; i < n && i < cap; ++i) {
    out[i].dm = v[i].dm;
    out[i].dm_idx = v[i].dm_idx;
    out[i].acc = v[i].acc;
    out[i].nh = v[i].nh;
    out[i].snr = v[i].snr;
    out[i].freq = v[i].freq;
    out[i].nassoc = v[i].count_assoc();
    out[i]._pad = 0;
  }
  return n;
}

/* Pre-order forest (Candidate::collect_candidates order, candidates.hpp:78-84)
 * with each record's subtree size in nassoc.  Returns the record count. */
void forest_rec(const Cand& c, orc_cand* out, int cap, int& pos) {
  if (pos < cap) {
    out[pos].dm = c.dm;
    out[pos].dm_idx = c.dm_idx;
    out[pos].acc = c.acc;
    out[pos].nh = c.nh;
    out[pos].snr = c.snr;
    out[pos].freq = c.freq;
    out[pos].nassoc = c.count_assoc();
    out[pos]._pad = 0;
  }
  ++pos;
  for (const Cand& a : c.assoc) forest_rec(a, out, cap, pos);
}

int to_forest(const std::vector<Cand>& v, orc_cand* out, int cap) {
  int pos = 0;
  for (const Cand& c : v) forest_rec(c, out, cap, pos);
  return pos;
}

/* Inverse of to_forest; -1 if a subtree runs past n. */
int from_forest_rec(const orc_cand* in, int n, int i, Cand& c) {
  c.dm = in[i].dm;
  c.dm_idx = in[i].dm_idx;
  c.acc = in[i].acc;
  c.nh = in[i].nh;
  c.snr = in[i].snr;
  c.freq = in[i].freq;
  c.assoc.clear();
  if (in[i].nassoc < 0) return -1;
  const long end = (long)i + 1 + in[i].nassoc;
  if (end > n) return -1;
  int j = i + 1;
  while (j < end) {
    c.assoc.emplace_back();
    j = from_forest_rec(in, n, j, c.assoc.back());
    if (j < 0 || j > end) return -1;
  }
  return j;
}

/* CandidateScorer (include/transforms/scorer.hpp:8-87): float members,
 * double-literal expressions narrowed on assignment, as written there. */
struct Scores {
  float ddm_count_ratio, ddm_snr_ratio;
  bool is_adjacent, is_physical;
};

struct Scorer {
  float chan_partial, band_partial;
  Scorer(float cfreq, float foff, float bw) {
    float ftop = cfreq + bw / 2.0;     /* scorer.hpp:71 */
    float fbottom = cfreq - bw / 2.0;  /* :72 */
    chan_partial = 8300.0 * foff / std::pow((double)cfreq, 3.0);                                   /* :73 */
    band_partial = 4150.0 * (1.0 / std::pow((double)fbottom, 2.0) - 1.0 / std::pow((double)ftop, 2.0)); /* :74 */
  }
  Scores score(const Cand& c) const {
    Scores r;
    /* has_physical_period, :25-27 */
    float tdm = c.dm * chan_partial;
    r.is_physical = 1.0 / c.freq > tdm;
    /* has_adjacency, :29-45 */
    bool adjacent = false, unique = true;
    for (size_t ii = 0; ii < c.assoc.size(); ii++) {
      if (c.assoc[ii].dm_idx != c.dm_idx) unique = false;
      if (c.assoc[ii].dm_idx == c.dm_idx + 1 || c.assoc[ii].dm_idx == c.dm_idx - 1) {
        adjacent = true;
        break;
      }
    }
    r.is_adjacent = adjacent || unique;
    /* delta_dm_ratio, :47-65 */
    int inside_count = 1, total_count = 1;
    float inside_snr = c.snr, total_snr = c.snr;
    float ddm = 1.0 / (c.freq * band_partial);
    for (size_t ii = 0; ii < c.assoc.size(); ii++) {
      total_count++;
      total_snr += c.assoc[ii].snr;
      if (std::fabs(c.dm - c.assoc[ii].dm) <= ddm) {
        inside_count++;
        inside_snr += c.assoc[ii].snr;
      }
    }
    r.ddm_count_ratio = (float)inside_count / total_count;
    r.ddm_snr_ratio = inside_snr / total_snr;
    return r;
  }
};

}  // namespace

extern "C" int orc_score(const orc_cand* forest, int n, float cfreq, float foff, float bw, float* ratios,
                         int* flags) {
  int i = 0, t = 0;
  Scorer sc(cfreq, foff, bw);
  while (i < n) {
    Cand c;
    int j = from_forest_rec(forest, n, i, c);
    if (j < 0) return -1;
    Scores r = sc.score(c);
    ratios[2 * t] = r.ddm_count_ratio;
    ratios[2 * t + 1] = r.ddm_snr_ratio;
    flags[2 * t] = r.is_adjacent;
    flags[2 * t + 1] = r.is_physical;
    ++t;
    i = j;
  }
  return t;
}

extern "C" int orc_finalize(const orc_cand* in, int n, float freq_tol, float max_harm, float cfreq, float foff,
                            float bw, int limit, orc_cand* out, int cap, float* ratios, int* flags, int* nfinal) {
  std::vector<Cand> v;
  int i = 0;
  while (i < n) {
    v.emplace_back();
    i = from_forest_rec(in, n, i, v.back());
    if (i < 0) return -1;
  }
  /* pipeline_multi.cu:406-407,547-548 */
  v = dm_distill(v, freq_tol, true);
  v = harmonic_distill(v, freq_tol, max_harm, true, false);
  /* :550-552 */
  Scorer sc(cfreq, foff, bw);
  /* :557-558 */
  const size_t keep = std::min((size_t)std::max(limit, 0), v.size());
  v.resize(keep);
  for (size_t t = 0; t < v.size(); ++t) {
    Scores r = sc.score(v[t]);
    ratios[2 * t] = r.ddm_count_ratio;
    ratios[2 * t + 1] = r.ddm_snr_ratio;
    flags[2 * t] = r.is_adjacent;
    flags[2 * t + 1] = r.is_physical;
  }
  *nfinal = (int)v.size();
  return to_forest(v, out, cap);
}

extern "C" int orc_harmonic_distill(const orc_cand* in, int n, float tol, float max_harm, int keep_related,
                                    int fractional, orc_cand* out) {
  std::vector<Cand> v = from_flat(in, n);
  return to_flat(harmonic_distill(v, tol, max_harm, keep_related != 0, fractional != 0), out, n);
}

extern "C" int orc_accel_distill(const orc_cand* in, int n, float tobs, float tol, int keep_related,
                                 orc_cand* out) {
  std::vector<Cand> v = from_flat(in, n);
  return to_flat(accel_distill(v, tobs, tol, keep_related != 0), out, n);
}

extern "C" int orc_dm_distill(const orc_cand* in, int n, float tol, int keep_related, orc_cand* out) {
  std::vector<Cand> v = from_flat(in, n);
  return to_flat(dm_distill(v, tol, keep_related != 0), out, n);
}

/* ------------------------------------------------------------------------- */
/* Worker sequence: src/pipeline_multi.cu:155-295                            */
/* ------------------------------------------------------------------------- */
namespace {

struct TrialScratch {
  std::vector<float> tim_r, X, P, folds;
  std::vector<cd> z, work, Xd;
  std::vector<orc_fast::cf> fz, fzo, fwork;  // float FFT buffers (fast_fft)
};

/* pipeline_multi.cu:130-131: float tobs = size*tsamp; float bin_width = 1.0/tobs */
inline float tobs_of(const orc_search_params* p) { return (float)p->size * p->tsamp; }
inline float bin_width_of(const orc_search_params* p) { return (float)(1.0 / tobs_of(p)); }

void rfft_float(RPlan* rp, const float* in, float* X, TrialScratch& s) {
  const uint64_t nb = rp->n / 2 + 1;
  s.Xd.resize(nb);
  rfft(rp, in, nullptr, s.Xd.data(), s.z, s.work);
  for (uint64_t k = 0; k < nb; ++k) {
    X[2 * k] = (float)s.Xd[k].real();
    X[2 * k + 1] = (float)s.Xd[k].imag();
  }
}

/* One accel trial: pipeline_multi.cu:249-289 */
int run_trial(const float* tim, const orc_search_params* p, float acc, float mean, float std_,
              TrialScratch& s, std::vector<Cand>& out, int* raw_nh, int* raw_idx, float* raw_snr, int raw_cap,
              int* raw_count) {
  const uint64_t N = p->size;
  const uint64_t nb = N / 2 + 1;
  RPlan* rp = get_rplan(N);
  if (!rp) return -1;
  s.tim_r.resize(N);
  s.X.resize(2 * nb);
  s.P.resize(nb);
  const unsigned F = (unsigned)std::min(std::max(p->nharmonics, 0), 5);
  s.folds.resize((size_t)F * nb);
  orc_resample2(tim, s.tim_r.data(), N, acc, p->tsamp);
  if (p->fast_fft) {
    const orc_fast::RealPlan* fp = orc_fast::real_plan(N);
    if (!fp) return -1;
    orc_fast::rfft(*fp, s.tim_r.data(), s.X.data(), s.fz, s.fzo, s.fwork);
  } else {
    rfft_float(rp, s.tim_r.data(), s.X.data(), s);
  }
  orc_form_interbin(s.X.data(), s.P.data(), nb);
  /* stats::normalise(pspec, mean*size, std*size, size/2+1) */
  orc_normalise(s.P.data(), mean * (float)(unsigned)N, std_ * (float)(unsigned)N, nb);
  orc_harmonic_sum(s.P.data(), s.folds.data(), nb, F);
  const double bw = (double)bin_width_of(p);
  std::vector<Cand> trial;
  std::vector<int> pidx(100001);
  std::vector<float> psnr(100001), pfreq(100001);
  int nraw = 0;
  for (unsigned nh = 0; nh <= F; ++nh) {
    const float* spec = nh == 0 ? s.P.data() : s.folds.data() + (size_t)(nh - 1) * nb;
    int np = orc_find_candidates(spec, (uint32_t)nb, bw, nh, p->min_snr, p->min_freq, p->max_freq, p->min_gap,
                                 pidx.data(), psnr.data(), pfreq.data(), 100001);
    if (np < 0) return -2;
    for (int i = 0; i < np; ++i) {
      trial.push_back(Cand{p->dm, p->dm_idx, acc, (int)nh, psnr[i], pfreq[i], {}});
      if (raw_nh && nraw < raw_cap) {
        raw_nh[nraw] = (int)nh;
        raw_idx[nraw] = pidx[i];
        raw_snr[nraw] = psnr[i];
      }
      ++nraw;
    }
  }
  if (raw_count) *raw_count = nraw;
  out = harmonic_distill(trial, p->freq_tol, p->max_harm, false, true);
  return 0;
}

}  // namespace

extern "C" int orc_prologue(const float* series, uint64_t n_real, const orc_search_params* p, float* tim_out,
                            float* mean, float* rms, float* std_) {
  const uint64_t N = p->size;
  const uint64_t nb = N / 2 + 1;
  RPlan* rp = get_rplan(N);
  if (!rp) return -1;
  const uint64_t ncopy = std::min(n_real, N);
  std::vector<float> tim(N, 0.0f);
  std::memcpy(tim.data(), series, sizeof(float) * ncopy);
  orc_remove_baseline(tim.data(), ncopy);
  if (N > n_real)
    for (uint64_t i = n_real; i < N; ++i) tim[i] = 0.0f;
  TrialScratch s;
  std::vector<float> X(2 * nb), A(nb), med(nb), P(nb);
  const orc_fast::RealPlan* fp = p->fast_fft ? orc_fast::real_plan(N) : nullptr;
  if (p->fast_fft && !fp) return -1;
  if (fp)
    orc_fast::rfft(*fp, tim.data(), X.data(), s.fz, s.fzo, s.fwork);
  else
    rfft_float(rp, tim.data(), X.data(), s);
  orc_form_amplitude(X.data(), A.data(), nb);
  orc_calculate_median(A.data(), nb, (double)bin_width_of(p), p->boundary_5_freq, p->boundary_25_freq, med.data());
  orc_divide_c_by_f(X.data(), med.data(), nb);
  /* pipeline_multi.cu:222-226 */
  if (p->nbirdies > 0) orc_zap_birdies(X.data(), nb, bin_width_of(p), p->birdie_freq, p->birdie_width, p->nbirdies);
  orc_form_interbin(X.data(), P.data(), nb);
  orc_stats(P.data(), nb, 0, mean, rms, std_);
  if (fp) {
    orc_fast::irfft(*fp, X.data(), tim_out, s.fz, s.fzo, s.fwork);
    return 0;
  }
  std::vector<cd> z, work;
  irfft(rp, X.data(), tim_out, z, work);
  return 0;
}

extern "C" int orc_fft_r2c_fast(const float* in, float* out, uint64_t n) {
  const orc_fast::RealPlan* fp = orc_fast::real_plan(n);
  if (!fp) return -1;
  std::vector<orc_fast::cf> z, zo, work;
  orc_fast::rfft(*fp, in, out, z, zo, work);
  return 0;
}

extern "C" int orc_fft_c2r_fast(const float* in, float* out, uint64_t n) {
  const orc_fast::RealPlan* fp = orc_fast::real_plan(n);
  if (!fp) return -1;
  std::vector<orc_fast::cf> z, zo, work;
  orc_fast::irfft(*fp, in, out, z, zo, work);
  return 0;
}

extern "C" int orc_trial(const float* tim, const orc_search_params* p, float acc, float mean, float std_,
                         orc_cand* out, int cap, int* raw_nh, int* raw_idx, float* raw_snr, int raw_cap,
                         int* raw_count) {
  TrialScratch s;
  std::vector<Cand> v;
  int rc = run_trial(tim, p, acc, mean, std_, s, v, raw_nh, raw_idx, raw_snr, raw_cap, raw_count);
  if (rc) return rc;
  return to_flat(v, out, cap);
}

extern "C" int orc_search_accels(const float* tim, const orc_search_params* p, const float* accs, int nacc,
                                 float mean, float std_, orc_cand* out, int cap, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  if (!get_rplan(p->size)) return -1;
  std::vector<std::vector<Cand>> per(nacc);
  std::atomic<int> next(0), err(0);
  auto worker = [&]() {
    TrialScratch s;
    while (true) {
      int t = next.fetch_add(1);
      if (t >= nacc) break;
      if (run_trial(tim, p, accs[t], mean, std_, s, per[t], nullptr, nullptr, nullptr, 0, nullptr)) err = 1;
    }
  };
  std::vector<std::thread> th;
  for (int i = 0; i < nthreads; ++i) th.emplace_back(worker);
  for (auto& t : th) t.join();
  if (err) return -2;
  std::vector<Cand> all;
  for (auto& v : per) all.insert(all.end(), v.begin(), v.end());
  std::vector<Cand> d = accel_distill(all, tobs_of(p), p->freq_tol, true);
  return to_forest(d, out, cap);
}

extern "C" int orc_search_dm(const float* series, uint64_t n_real, const orc_search_params* p, const float* accs,
                             int nacc, orc_cand* out, int cap, int nthreads, double* timings) {
  auto t0 = std::chrono::steady_clock::now();
  std::vector<float> tim(p->size);
  float mean, rms, sd;
  int rc = orc_prologue(series, n_real, p, tim.data(), &mean, &rms, &sd);
  if (rc) return rc;
  auto t1 = std::chrono::steady_clock::now();
  int n = orc_search_accels(tim.data(), p, accs, nacc, mean, sd, out, cap, nthreads);
  auto t2 = std::chrono::steady_clock::now();
  if (timings) {
    timings[0] = std::chrono::duration<double>(t1 - t0).count();
    timings[1] = std::chrono::duration<double>(t2 - t1).count();
    timings[2] = 0.0;
  }
  return n;
}

extern "C" int orc_hardware_threads(void) { return (int)std::thread::hardware_concurrency(); }

/* ------------------------------------------------------------------------- */
/* Dedispersion (libdedisp, called through include/transforms/dedisperser.hpp) */
/* ------------------------------------------------------------------------- */
/* libdedisp is not part of the reference tree; these restate its published
 * host-side DM-list recurrence and delay table and its direct dedispersion
 * kernel with float output.  The plan holds tsamp, f0 and df as floats. */
extern "C" int orc_dm_list(float dm_start, float dm_end, float dt, float ti, float f0, float df, int nchans, float tol,
                           float* out, int cap) {
  /* the smearing of one DM step may grow the pulse width by at most tol;
   * times in microseconds, the centre frequency in GHz */
  double dt_us = (double)dt * 1e6;
  double f = ((double)f0 + ((double)(nchans / 2) - 0.5) * (double)df) * 1e-3;
  double tol2 = (double)tol * (double)tol;
  double a = 8.3 * (double)df / (f * f * f);
  double a2 = a * a;
  double b2 = a2 * (double)((double)nchans * (double)nchans / 16.0);
  double c = (dt_us * dt_us + (double)ti * (double)ti) * (tol2 - 1.0);
  std::vector<float> v;
  v.push_back(dm_start);
  while (v.back() < dm_end) {
    double prev = v.back();
    double prev2 = prev * prev;
    double k = c + tol2 * a2 * prev2;
    double dm = (b2 * prev + std::sqrt(-a2 * b2 * prev2 + (a2 + b2) * k)) / (a2 + b2);
    v.push_back((float)dm);
    if (v.size() > (1u << 24)) return -1;
  }
  for (size_t i = 0; i < v.size() && (int)i < cap; ++i) out[i] = v[i];
  return (int)v.size();
}

extern "C" void orc_delay_table(int nchans, float dt, float f0, float df, float* table) {
  for (int c = 0; c < nchans; ++c) {
    float a = 1.f / (f0 + c * df);
    float b = 1.f / f0;
    table[c] = 4.15e3 / dt * (a * a - b * b);
  }
}

/* Direct dedispersion of packed time-major samples (nbits per sample, lowest
 * bits first within a byte): out[d*nout + t] = sum over unmasked channels of
 * the sample at t + rn(dm_d * table[c]), divided by (2^nbits - 1) * nchans;
 * nout = nsamps - rn(dm_last * table[nchans-1]).  Returns nout (or -1). */
extern "C" long orc_dedisperse(const uint8_t* in, long nsamps, int nchans, int nbits, const float* dms, int ndm,
                               const int* killmask, float dt, float f0, float df, float* out) {
  std::vector<float> table(nchans);
  orc_delay_table(nchans, dt, f0, df, table.data());
  long max_delay = (long)std::nearbyint(dms[ndm - 1] * table[nchans - 1]);
  long nout = nsamps - max_delay;
  if (nout <= 0) return -1;
  long bytes_per_sample = (long)nchans * nbits / 8;
  unsigned mask = (1u << nbits) - 1u;
  float in_range = (float)mask;
  for (int d = 0; d < ndm; ++d) {
    std::vector<long> delay(nchans);
    for (int c = 0; c < nchans; ++c) delay[c] = (long)std::nearbyint(dms[d] * table[c]);
    for (long t = 0; t < nout; ++t) {
      unsigned sum = 0;
      for (int c = 0; c < nchans; ++c) {
        if (killmask && !killmask[c]) continue;
        long ts = t + delay[c];
        long bit = (long)c * nbits;
        unsigned byte = in[ts * bytes_per_sample + bit / 8];
        sum += (byte >> (bit % 8)) & mask;
      }
      out[(long)d * nout + t] = (float)sum * 1.0f / (in_range * nchans);
    }
  }
  return nout;
}
