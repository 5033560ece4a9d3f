/*
 * fastfft.cpp - single-precision complex FFT for the CPU baseline timing.
 *
 * TEST INFRASTRUCTURE ONLY (part of the CPU oracle library).  The reference
 * transforms in FP32 (cuFFT, include/transforms/ffter.hpp:54-110), and the
 * host-core baseline of the benchmark plays that role with an FFTW3f-class
 * transform: a four-step FFT (column FFTs, twiddle, row FFTs) whose
 * sub-transforms of at most a few thousand points run in cache, in float
 * arithmetic with correctly rounded float twiddles.  The exact double DFT in
 * oracle.cpp stays the checker; this one is checked against it.
 *
 * Lengths: m = 2^a * 3^b complex points (the search uses 2^k and 3 * 2^k).
 */
#include "fastfft.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>

namespace orc_fast {

namespace {

cf root(uint64_t t, uint64_t m) {
  const long double ang = -2.0L * 3.14159265358979323846264338327950288L * ((long double)t / (long double)m);
  return cf{(float)cosl(ang), (float)sinl(ang)};
}

inline cf mul(cf a, cf b) { return cf{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline cf add(cf a, cf b) { return cf{a.re + b.re, a.im + b.im}; }
inline cf sub(cf a, cf b) { return cf{a.re - b.re, a.im - b.im}; }

bool radices_of(uint64_t n, std::vector<int>& r) {
  r.clear();
  while (n % 4 == 0) { r.push_back(4); n /= 4; }
  while (n % 2 == 0) { r.push_back(2); n /= 2; }
  while (n % 3 == 0) { r.push_back(3); n /= 3; }
  return n == 1;
}

}  // namespace

bool Small::init(int len) {
  n = len;
  if (!radices_of((uint64_t)len, radices)) return false;
  tw.resize(len);
  for (int t = 0; t < len; ++t) tw[t] = root((uint64_t)t, (uint64_t)len);
  return true;
}

/* Forward transform of data (n points); work holds n points.  Stockham
 * autosort stages of radix 4, 2 and 3. */
void Small::forward(cf* data, cf* work) const {
  cf* in = data;
  cf* out = work;
  int Ns = 1;
  for (int R : radices) {
    const int nbf = n / R;
    const int tstep = n / (Ns * R);
    for (int j = 0; j < nbf; ++j) {
      const int k = j % Ns;
      cf v[4];
      for (int r = 0; r < R; ++r) v[r] = in[j + r * nbf];
      if (Ns > 1)
        for (int r = 1; r < R; ++r) v[r] = mul(v[r], tw[k * r * tstep]);
      const int o = (j / Ns) * Ns * R + k;
      if (R == 4) {
        const cf a0 = add(v[0], v[2]), a1 = sub(v[0], v[2]), a2 = add(v[1], v[3]), d = sub(v[1], v[3]);
        const cf a3{d.im, -d.re};  // -i d
        out[o] = add(a0, a2);
        out[o + Ns] = add(a1, a3);
        out[o + 2 * Ns] = sub(a0, a2);
        out[o + 3 * Ns] = sub(a1, a3);
      } else if (R == 2) {
        out[o] = add(v[0], v[1]);
        out[o + Ns] = sub(v[0], v[1]);
      } else {
        const float s3 = 0.866025403784438647f;
        const cf t1 = add(v[1], v[2]), t2 = sub(v[1], v[2]);
        const cf base{v[0].re - 0.5f * t1.re, v[0].im - 0.5f * t1.im};
        const cf rot{t2.im * s3, -t2.re * s3};  // -i s3 t2
        out[o] = add(v[0], t1);
        out[o + Ns] = add(base, rot);
        out[o + 2 * Ns] = sub(base, rot);
      }
    }
    Ns *= R;
    cf* t = in;
    in = out;
    out = t;
  }
  if (in != data) std::memcpy(data, in, sizeof(cf) * (size_t)n);
}

/* B transforms at once, split real/imaginary planes, point t of transform c
 * at [t * B + c]: every butterfly works on B independent transforms with the
 * same twiddle, so the inner loop over c is contiguous and vectorizes. */
void Small::forward_batch(float* re, float* im, float* wre, float* wim) const {
  constexpr int B = kBatch;
  float *ir = re, *ii = im, *orr = wre, *oi = wim;
  int Ns = 1;
  for (int R : radices) {
    const int nbf = n / R;
    const int tstep = n / (Ns * R);
    for (int j = 0; j < nbf; ++j) {
      const int k = j % Ns;
      const int o = (j / Ns) * Ns * R + k;
      float w1r = 1.f, w1i = 0.f, w2r = 1.f, w2i = 0.f, w3r = 1.f, w3i = 0.f;
      if (Ns > 1) {
        w1r = tw[k * tstep].re; w1i = tw[k * tstep].im;
        if (R > 2) { w2r = tw[2 * k * tstep].re; w2i = tw[2 * k * tstep].im; }
        if (R > 3) { w3r = tw[3 * k * tstep].re; w3i = tw[3 * k * tstep].im; }
      }
      const float* a0r = ir + (size_t)j * B;
      const float* a0i = ii + (size_t)j * B;
      if (R == 4) {
        const float *a1r = a0r + (size_t)nbf * B, *a1i = a0i + (size_t)nbf * B;
        const float *a2r = a1r + (size_t)nbf * B, *a2i = a1i + (size_t)nbf * B;
        const float *a3r = a2r + (size_t)nbf * B, *a3i = a2i + (size_t)nbf * B;
        float *o0r = orr + (size_t)o * B, *o0i = oi + (size_t)o * B;
        float *o1r = o0r + (size_t)Ns * B, *o1i = o0i + (size_t)Ns * B;
        float *o2r = o1r + (size_t)Ns * B, *o2i = o1i + (size_t)Ns * B;
        float *o3r = o2r + (size_t)Ns * B, *o3i = o2i + (size_t)Ns * B;
        for (int c = 0; c < B; ++c) {
          const float v0r = a0r[c], v0i = a0i[c];
          const float v1r = a1r[c] * w1r - a1i[c] * w1i, v1i = a1r[c] * w1i + a1i[c] * w1r;
          const float v2r = a2r[c] * w2r - a2i[c] * w2i, v2i = a2r[c] * w2i + a2i[c] * w2r;
          const float v3r = a3r[c] * w3r - a3i[c] * w3i, v3i = a3r[c] * w3i + a3i[c] * w3r;
          const float s0r = v0r + v2r, s0i = v0i + v2i, s1r = v0r - v2r, s1i = v0i - v2i;
          const float s2r = v1r + v3r, s2i = v1i + v3i, dr = v1r - v3r, di = v1i - v3i;
          o0r[c] = s0r + s2r; o0i[c] = s0i + s2i;
          o1r[c] = s1r + di;  o1i[c] = s1i - dr;   // s1 - i d
          o2r[c] = s0r - s2r; o2i[c] = s0i - s2i;
          o3r[c] = s1r - di;  o3i[c] = s1i + dr;   // s1 + i d
        }
      } else if (R == 2) {
        const float *a1r = a0r + (size_t)nbf * B, *a1i = a0i + (size_t)nbf * B;
        float *o0r = orr + (size_t)o * B, *o0i = oi + (size_t)o * B;
        float *o1r = o0r + (size_t)Ns * B, *o1i = o0i + (size_t)Ns * B;
        for (int c = 0; c < B; ++c) {
          const float v1r = a1r[c] * w1r - a1i[c] * w1i, v1i = a1r[c] * w1i + a1i[c] * w1r;
          o0r[c] = a0r[c] + v1r; o0i[c] = a0i[c] + v1i;
          o1r[c] = a0r[c] - v1r; o1i[c] = a0i[c] - v1i;
        }
      } else {
        const float s3 = 0.866025403784438647f;
        const float *a1r = a0r + (size_t)nbf * B, *a1i = a0i + (size_t)nbf * B;
        const float *a2r = a1r + (size_t)nbf * B, *a2i = a1i + (size_t)nbf * B;
        float *o0r = orr + (size_t)o * B, *o0i = oi + (size_t)o * B;
        float *o1r = o0r + (size_t)Ns * B, *o1i = o0i + (size_t)Ns * B;
        float *o2r = o1r + (size_t)Ns * B, *o2i = o1i + (size_t)Ns * B;
        for (int c = 0; c < B; ++c) {
          const float v1r = a1r[c] * w1r - a1i[c] * w1i, v1i = a1r[c] * w1i + a1i[c] * w1r;
          const float v2r = a2r[c] * w2r - a2i[c] * w2i, v2i = a2r[c] * w2i + a2i[c] * w2r;
          const float t1r = v1r + v2r, t1i = v1i + v2i, t2r = v1r - v2r, t2i = v1i - v2i;
          const float br = a0r[c] - 0.5f * t1r, bi = a0i[c] - 0.5f * t1i;
          const float rr = t2i * s3, ri = -t2r * s3;  // -i s3 t2
          o0r[c] = a0r[c] + t1r; o0i[c] = a0i[c] + t1i;
          o1r[c] = br + rr; o1i[c] = bi + ri;
          o2r[c] = br - rr; o2i[c] = bi - ri;
        }
      }
    }
    Ns *= R;
    std::swap(ir, orr);
    std::swap(ii, oi);
  }
  if (ir != re) {
    std::memcpy(re, ir, sizeof(float) * (size_t)n * B);
    std::memcpy(im, ii, sizeof(float) * (size_t)n * B);
  }
}

bool Plan::init(uint64_t len) {
  m = len;
  std::vector<int> r;
  if (!radices_of(len, r)) return false;
  if (len <= 4096) {
    n1 = 1;
    n2 = (int)len;
    return s2.init(n2);
  }
  // m = n1 * n2: the factor 3 (if any) in the column length n1, the row
  // length n2 the power of two nearest sqrt(m)
  uint64_t p2 = len, three = 1;
  while (p2 % 3 == 0) { p2 /= 3; three *= 3; }
  int lg = 0;
  while ((1ull << lg) < p2) ++lg;
  n2 = 1 << ((lg + 1) / 2);
  n1 = (int)(len / (uint64_t)n2);
  if (!s1.init(n1) || !s2.init(n2)) return false;
  // four-step twiddles W_m^(k1*b) as hi[x >> shift] * lo[x & mask]
  shift = 0;
  while ((1ull << (2 * (shift + 1))) <= len) ++shift;
  const uint64_t T = 1ull << shift;
  lo.resize(T);
  hi.resize(len / T + 2);
  for (uint64_t t = 0; t < T; ++t) lo[t] = root(t, len);
  for (uint64_t h = 0; h < hi.size(); ++h) hi[h] = root(h * T, len);
  return true;
}

/* Forward transform: z (m points, overwritten) -> out (m points, natural
 * order).  work holds work_points() points. */
void Plan::forward(cf* z, cf* out, cf* work) const {
  if (n1 == 1) {
    std::memcpy(out, z, sizeof(cf) * m);
    s2.forward(out, work);
    return;
  }
  constexpr int B = kBatch;  // columns / rows transformed together
  const size_t L = (size_t)std::max(n1, n2) * B;
  float* re = reinterpret_cast<float*>(work);
  float* im = re + L;
  float* wre = im + L;
  float* wim = wre + L;
  const uint64_t mask = (1ull << shift) - 1;
  // columns: Y[k1][b] = FFT_n1(z[. * n2 + b])[k1] * W_m^(k1 * b), in place
  for (int b0 = 0; b0 < n2; b0 += B) {
    for (int a = 0; a < n1; ++a) {
      const cf* row = z + (size_t)a * n2 + b0;
      for (int c = 0; c < B; ++c) {
        re[(size_t)a * B + c] = row[c].re;
        im[(size_t)a * B + c] = row[c].im;
      }
    }
    s1.forward_batch(re, im, wre, wim);
    for (int k1 = 0; k1 < n1; ++k1) {
      cf* row = z + (size_t)k1 * n2 + b0;
      for (int c = 0; c < B; ++c) {
        const uint64_t x = (uint64_t)k1 * (uint64_t)(b0 + c);
        const cf w = mul(hi[x >> shift], lo[x & mask]);
        const cf v{re[(size_t)k1 * B + c], im[(size_t)k1 * B + c]};
        row[c] = k1 ? mul(v, w) : v;
      }
    }
  }
  // rows: X[k1 + n1 * k2] = FFT_n2(Y[k1][.])[k2]
  for (int k10 = 0; k10 < n1; k10 += B) {
    for (int r = 0; r < B; ++r) {
      const cf* row = z + (size_t)(k10 + r) * n2;
      for (int b = 0; b < n2; ++b) {
        re[(size_t)b * B + r] = row[b].re;
        im[(size_t)b * B + r] = row[b].im;
      }
    }
    s2.forward_batch(re, im, wre, wim);
    for (int k2 = 0; k2 < n2; ++k2) {
      cf* o = out + (size_t)k2 * n1 + k10;
      for (int r = 0; r < B; ++r) o[r] = cf{re[(size_t)k2 * B + r], im[(size_t)k2 * B + r]};
    }
  }
}

size_t Plan::work_points() const { return (size_t)2 * kBatch * (size_t)std::max(n1, n2) + (size_t)n2; }

namespace {
std::mutex g_mu;
std::map<uint64_t, std::unique_ptr<RealPlan>> g_plans;
}  // namespace

const RealPlan* real_plan(uint64_t n) {
  if (n < 4 || (n & 1)) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_plans.find(n);
  if (it != g_plans.end()) return it->second.get();
  std::unique_ptr<RealPlan> p(new RealPlan);
  p->n = n;
  if (!p->c.init(n / 2)) return nullptr;
  p->wn.resize(n / 2 + 1);
  for (uint64_t k = 0; k <= n / 2; ++k) p->wn[k] = root(k, n);
  const RealPlan* raw = p.get();
  g_plans[n] = std::move(p);
  return raw;
}

/* R2C: x (n floats) -> X (n/2 + 1 complex), unnormalised.  z, zo, work
 * are resized as needed. */
void rfft(const RealPlan& p, const float* x, float* X, std::vector<cf>& z, std::vector<cf>& zo,
          std::vector<cf>& work) {
  const uint64_t m = p.n / 2;
  z.resize(m);
  zo.resize(m);
  work.resize(p.c.work_points());
  std::memcpy(z.data(), x, sizeof(float) * p.n);
  p.c.forward(z.data(), zo.data(), work.data());
  for (uint64_t k = 0; k <= m; ++k) {
    const cf zk = zo[k % m];
    const cf zmc{zo[(m - k) % m].re, -zo[(m - k) % m].im};  // conj(Z[m - k])
    const cf e{0.5f * (zk.re + zmc.re), 0.5f * (zk.im + zmc.im)};
    const cf d = sub(zk, zmc);
    const cf o{0.5f * d.im, -0.5f * d.re};  // -i/2 d
    const cf r = add(e, mul(p.wn[k], o));
    X[2 * k] = r.re;
    X[2 * k + 1] = r.im;
  }
}

/* C2R, unnormalised (n * x for X = R2C(x)): the inverse as conj(FFT(conj)). */
void irfft(const RealPlan& p, const float* X, float* out, std::vector<cf>& z, std::vector<cf>& zo,
           std::vector<cf>& work) {
  const uint64_t m = p.n / 2;
  z.resize(m);
  zo.resize(m);
  work.resize(p.c.work_points());
  for (uint64_t k = 0; k < m; ++k) {
    const cf xk{X[2 * k], X[2 * k + 1]};
    const cf xm{X[2 * (m - k)], -X[2 * (m - k) + 1]};
    const cf s = add(xk, xm);
    const cf w{p.wn[k].re, -p.wn[k].im};
    const cf d = mul(sub(xk, xm), w);
    const cf v{s.re - d.im, s.im + d.re};  // s + i d
    z[k] = cf{v.re, -v.im};                 // conjugated for the forward transform
  }
  p.c.forward(z.data(), zo.data(), work.data());
  for (uint64_t i = 0; i < m; ++i) {
    out[2 * i] = zo[i].re;
    out[2 * i + 1] = -zo[i].im;
  }
}

}  // namespace orc_fast
