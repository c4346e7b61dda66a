// Register-resident mixed-radix FFT of compile-time length N for one lane.
//
// In-place decimation in frequency with radices 4, 2, 3, 5 (largest power-of-
// four factor first). Every index is a compile-time constant after unrolling,
// so the N complex values stay in VGPRs; the digit-reversed output order is
// resolved at compile time by the caller through FftOrder<N>::pos(f).
//
// Convention (the reference's, src/fft.rs:19-114): SIGN = -1 computes
// X[f] = Σ_j x[j] e^{-2πi jf/N} (forward), SIGN = +1 the unnormalised inverse.
#pragma once

#include <hip/hip_runtime.h>

namespace pwk {

// ---- compile-time trigonometry (Taylor series after reduction to [-π/4, π/4]) ----
constexpr double kPi = 3.141592653589793238462643383279502884;

constexpr double taylor_sin(double x) {
  double term = x, sum = x, x2 = x * x;
  for (int n = 1; n < 14; ++n) {
    term *= -x2 / ((2.0 * n) * (2.0 * n + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double taylor_cos(double x) {
  double term = 1.0, sum = 1.0, x2 = x * x;
  for (int n = 1; n < 14; ++n) {
    term *= -x2 / ((2.0 * n - 1.0) * (2.0 * n));
    sum += term;
  }
  return sum;
}
// cos(2π·e/m) and sin(2π·e/m) for integers, reduced by octant so the series
// argument stays within [-π/4, π/4].
constexpr double cos_frac(long e, long m) {
  e %= m;
  if (e < 0) e += m;
  // angle = 2π e/m; octant q = round(8e/m)
  long q8 = (8 * e + m / 2) / m;  // nearest multiple of π/4
  double r = 2.0 * kPi * ((double)(8 * e - q8 * m) / (8.0 * (double)m));
  switch (q8 % 8) {
    case 0: return taylor_cos(r);
    case 1: return 0.70710678118654752440 * (taylor_cos(r) - taylor_sin(r));
    case 2: return -taylor_sin(r);
    case 3: return -0.70710678118654752440 * (taylor_cos(r) + taylor_sin(r));
    case 4: return -taylor_cos(r);
    case 5: return -0.70710678118654752440 * (taylor_cos(r) - taylor_sin(r));
    case 6: return taylor_sin(r);
    default: return 0.70710678118654752440 * (taylor_cos(r) + taylor_sin(r));
  }
}
constexpr double sin_frac(long e, long m) { return cos_frac(4 * e - m, 4 * m); }  // sin(a) = cos(a - π/2)

// ---- factorisation -----------------------------------------------------------
// Radix order of the decimation in frequency: ORD 0 takes the largest power-of-four factor
// first (4, then 2, 3, 5); ORD 1 the odd radices first (5, 3, then 4, 2). Both compute the same
// transform; which one runs faster for a length and kernel is measured, not assumed (the FFT
// autotuner in autotune.hip chooses per launch shape).
constexpr int first_radix(int n, int ord = 0) {
  if (ord == 1) return n % 5 == 0 ? 5 : n % 3 == 0 ? 3 : n % 4 == 0 ? 4 : n % 2 == 0 ? 2 : n;
  return n % 4 == 0 ? 4 : n % 2 == 0 ? 2 : n % 3 == 0 ? 3 : n % 5 == 0 ? 5 : n;
}
constexpr bool fft_supported(int n) {
  if (n < 1) return false;
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  while (n % 5 == 0) n /= 5;
  return n == 1;
}

template <int N, int ORD = 0>
struct FftOrder {
  // position of output frequency f in the DIF result
  static constexpr int pos_rec(int f, int m) {
    if (m == 1) return 0;
    int p = first_radix(m, ORD);
    int mm = m / p;
    return (f % p) * mm + pos_rec(f / p, mm);
  }
  struct Table {
    int p[N];
  };
  static constexpr Table make() {
    Table t{};
    for (int f = 0; f < N; ++f) t.p[f] = pos_rec(f, N);
    return t;
  }
  // compile-time table: after unrolling, tab.p[f] folds to a constant index
  static constexpr Table tab = make();
  static constexpr int pos(int f) { return tab.p[f]; }
};

// Twiddle table W_M^e = exp(SIGN·2πi·e/M), e in [0, M).
template <int M, int SIGN>
struct Twiddle {
  struct Table {
    double c[M];
    double s[M];
  };
  static constexpr Table make() {
    Table t{};
    for (int e = 0; e < M; ++e) {
      t.c[e] = cos_frac(e, M);
      t.s[e] = SIGN * sin_frac(e, M);
    }
    return t;
  }
  static constexpr Table tab = make();
};

template <typename T>
struct C2 {
  T x, y;
};

// ---- radix butterflies on values at stride STRIDE starting at BASE ----------
template <typename T, int SIGN>
__device__ __forceinline__ void bfly2(T* re, T* im, int i0, int i1) {
  T ar = re[i0], ai = im[i0], br = re[i1], bi = im[i1];
  re[i0] = ar + br;
  im[i0] = ai + bi;
  re[i1] = ar - br;
  im[i1] = ai - bi;
}

template <typename T, int SIGN>
__device__ __forceinline__ void bfly4(T* re, T* im, int i0, int i1, int i2, int i3) {
  T s02r = re[i0] + re[i2], s02i = im[i0] + im[i2];
  T d02r = re[i0] - re[i2], d02i = im[i0] - im[i2];
  T s13r = re[i1] + re[i3], s13i = im[i1] + im[i3];
  T d13r = re[i1] - re[i3], d13i = im[i1] - im[i3];
  // ω = SIGN·i ; ω·(a+ib) = -SIGN·b + i·SIGN·a
  T wr = (T)(-SIGN) * d13i, wi = (T)SIGN * d13r;
  re[i0] = s02r + s13r;
  im[i0] = s02i + s13i;
  re[i2] = s02r - s13r;
  im[i2] = s02i - s13i;
  re[i1] = d02r + wr;
  im[i1] = d02i + wi;
  re[i3] = d02r - wr;
  im[i3] = d02i - wi;
}

template <typename T, int SIGN>
__device__ __forceinline__ void bfly3(T* re, T* im, int i0, int i1, int i2) {
  const T h = (T)0.86602540378443864676;  // √3/2
  T tr = re[i1] + re[i2], ti = im[i1] + im[i2];
  T dr = re[i1] - re[i2], di = im[i1] - im[i2];
  T mr = re[i0] - (T)0.5 * tr, mi = im[i0] - (T)0.5 * ti;
  // i·SIGN·h·d = (-SIGN h di) + i (SIGN h dr)
  T xr = (T)(-SIGN) * h * di, xi = (T)SIGN * h * dr;
  re[i0] = re[i0] + tr;
  im[i0] = im[i0] + ti;
  re[i1] = mr + xr;
  im[i1] = mi + xi;
  re[i2] = mr - xr;
  im[i2] = mi - xi;
}

template <typename T, int SIGN>
__device__ __forceinline__ void bfly5(T* re, T* im, int i0, int i1, int i2, int i3, int i4) {
  const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
  const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
  T t1r = re[i1] + re[i4], t1i = im[i1] + im[i4];
  T t2r = re[i2] + re[i3], t2i = im[i2] + im[i3];
  T d1r = re[i1] - re[i4], d1i = im[i1] - im[i4];
  T d2r = re[i2] - re[i3], d2i = im[i2] - im[i3];
  T ar = re[i0] + c1 * t1r + c2 * t2r, ai = im[i0] + c1 * t1i + c2 * t2i;
  T br = re[i0] + c2 * t1r + c1 * t2r, bi = im[i0] + c2 * t1i + c1 * t2i;
  // e1 = SIGN·(s1 d1 + s2 d2), e2 = SIGN·(s2 d1 - s1 d2); U = a ± i e
  T e1r = (T)SIGN * (s1 * d1r + s2 * d2r), e1i = (T)SIGN * (s1 * d1i + s2 * d2i);
  T e2r = (T)SIGN * (s2 * d1r - s1 * d2r), e2i = (T)SIGN * (s2 * d1i - s1 * d2i);
  re[i0] = re[i0] + t1r + t2r;
  im[i0] = im[i0] + t1i + t2i;
  re[i1] = ar - e1i;
  im[i1] = ai + e1r;
  re[i4] = ar + e1i;
  im[i4] = ai - e1r;
  re[i2] = br - e2i;
  im[i2] = bi + e2r;
  re[i3] = br + e2i;
  im[i3] = bi - e2r;
}

// One DIF stage over all blocks of size M, then recurse on M/p.
template <typename T, int N, int M, int SIGN, int ORD>
__device__ __forceinline__ void dif_stages(T* re, T* im) {
  if constexpr (M > 1) {
    constexpr int p = first_radix(M, ORD);
    constexpr int m = M / p;
#pragma unroll
    for (int b = 0; b < N; b += M) {
#pragma unroll
      for (int k = 0; k < m; ++k) {
        const int i0 = b + k;
        if constexpr (p == 2) {
          bfly2<T, SIGN>(re, im, i0, i0 + m);
        } else if constexpr (p == 4) {
          bfly4<T, SIGN>(re, im, i0, i0 + m, i0 + 2 * m, i0 + 3 * m);
        } else if constexpr (p == 3) {
          bfly3<T, SIGN>(re, im, i0, i0 + m, i0 + 2 * m);
        } else if constexpr (p == 5) {
          bfly5<T, SIGN>(re, im, i0, i0 + m, i0 + 2 * m, i0 + 3 * m, i0 + 4 * m);
        }
        if (k != 0) {
#pragma unroll
          for (int q = 1; q < p; ++q) {
            const int e = (q * k) % M;
            const T wc = (T)Twiddle<M, SIGN>::tab.c[e];
            const T ws = (T)Twiddle<M, SIGN>::tab.s[e];
            const int i = i0 + q * m;
            T xr = re[i], xi = im[i];
            re[i] = xr * wc - xi * ws;
            im[i] = xr * ws + xi * wc;
          }
        }
      }
    }
    dif_stages<T, N, m, SIGN, ORD>(re, im);
  }
}

// In-place FFT; the result for frequency f sits at index FftOrder<N, ORD>::pos(f).
template <typename T, int N, int SIGN, int ORD = 0>
__device__ __forceinline__ void fft_dif(T (&re)[N], T (&im)[N]) {
  dif_stages<T, N, N, SIGN, ORD>(re, im);
}

}  // namespace pwk
