// Line lengths with compiled register-resident FFT kernels ({2,3,5}-smooth).
#pragma once

// X(n) is expanded once per supported length.
#define PWK_FOR_EACH_FFT_SIZE(X) \
  X(8) X(9) X(10) X(12) X(15) X(16) X(18) X(20) X(24) X(25) X(27) X(30) X(32) X(36) X(40) X(45) X(48) X(50) \
  X(54) X(60) X(64) X(72) X(75) X(80) X(81)

// Longer {2,3,5}-smooth lines (up to 1024, the reference's largest grid axis) as L1 × L2 with both
// factors transformed in registers and the line staged in LDS (fft_long.hip): X(L, L1, L2).
#define PWK_FOR_EACH_LONG_FFT_SIZE(X)                                                                     \
  X(90, 9, 10) X(96, 8, 12) X(100, 10, 10) X(108, 9, 12) X(120, 10, 12) X(125, 5, 25) X(128, 8, 16)      \
  X(135, 9, 15) X(144, 12, 12) X(150, 10, 15) X(160, 10, 16) X(162, 9, 18) X(180, 12, 15) X(192, 12, 16) \
  X(200, 10, 20) X(216, 12, 18) X(225, 15, 15) X(240, 15, 16) X(243, 9, 27) X(250, 10, 25) X(256, 16, 16) \
  X(270, 10, 27) X(288, 9, 32) X(300, 10, 30) X(320, 10, 32) X(324, 12, 27) X(360, 12, 30) X(375, 15, 25) \
  X(384, 12, 32) X(400, 16, 25) X(405, 15, 27) X(432, 16, 27) X(450, 15, 30) X(480, 16, 30) X(486, 18, 27) \
  X(500, 20, 25) X(512, 16, 32) X(540, 20, 27) X(576, 24, 24) X(600, 20, 30) X(625, 25, 25) X(640, 20, 32) \
  X(648, 24, 27) X(675, 25, 27) X(720, 24, 30) X(729, 27, 27) X(750, 25, 30) X(768, 24, 32) X(800, 25, 32) \
  X(810, 27, 30) X(864, 27, 32) X(900, 30, 30) X(960, 30, 32) X(972, 27, 36) X(1000, 25, 40) X(1024, 32, 32)

namespace pwk {
constexpr int kMaxFftLen = 81;
constexpr int kMaxLongFftLen = 1024;
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
inline bool fft_len_long(int n) {
  switch (n) {
#define PWK_CASE(n, a, b) case n:
    PWK_FOR_EACH_LONG_FFT_SIZE(PWK_CASE)
#undef PWK_CASE
    return true;
    default:
      return false;
  }
}
// lengths any full-grid transform accepts: register lines up to 81, LDS-staged lines up to 1024
inline bool fft_len_supported(int n) { return fft_len_compiled(n) || fft_len_long(n); }
}  // namespace pwk
