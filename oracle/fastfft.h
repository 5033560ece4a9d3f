/* fastfft.h - single-precision FFT of the CPU baseline (test infrastructure,
 * see fastfft.cpp). */
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace orc_fast {

struct cf {
  float re, im;
};

constexpr int kBatch = 16;  // transforms of a four-step pass computed together

/* In-cache complex FFT of one length (Stockham, radix 4/2/3). */
struct Small {
  int n = 0;
  std::vector<int> radices;
  std::vector<cf> tw;
  bool init(int len);
  void forward(cf* data, cf* work) const;
  void forward_batch(float* re, float* im, float* wre, float* wim) const;
};

/* Four-step complex FFT of m = n1 * n2 points. */
struct Plan {
  uint64_t m = 0;
  int n1 = 1, n2 = 1;
  Small s1, s2;
  std::vector<cf> lo, hi;
  uint32_t shift = 0;
  bool init(uint64_t len);
  void forward(cf* z, cf* out, cf* work) const;
  size_t work_points() const;
};

struct RealPlan {
  uint64_t n = 0;
  Plan c;
  std::vector<cf> wn;  // W_n^k, k in [0, n/2]
};

const RealPlan* real_plan(uint64_t n);
void rfft(const RealPlan& p, const float* x, float* X, std::vector<cf>& z, std::vector<cf>& zo,
          std::vector<cf>& work);
void irfft(const RealPlan& p, const float* X, float* out, std::vector<cf>& z, std::vector<cf>& zo,
           std::vector<cf>& work);

}  // namespace orc_fast
