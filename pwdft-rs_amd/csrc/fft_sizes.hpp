// Line lengths with compiled register-resident FFT kernels ({2,3,5}-smooth).
#pragma once

// X(n) is expanded once per supported length.
#define PWK_FOR_EACH_FFT_SIZE(X) \
  X(8) X(9) X(10) X(12) X(15) X(16) X(18) X(20) X(24) X(25) X(27) X(30) X(32) X(36) X(40) X(45) X(48) X(50) \
  X(54) X(60) X(64) X(72) X(75) X(80) X(81)

namespace pwk {
constexpr int kMaxFftLen = 81;
inline bool fft_len_compiled(int n) {
  switch (n) {
#define PWK_CASE(n) case n:
    PWK_FOR_EACH_FFT_SIZE(PWK_CASE)
#undef PWK_CASE
    return true;
    default:
      return false;
  }
}
}  // namespace pwk
