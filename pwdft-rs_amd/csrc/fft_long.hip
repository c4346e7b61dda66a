// Line transforms longer than the register-resident lengths (up to 1024 points, {2,3,5}-smooth),
// src/fft.rs:19-114 convention: X[k] = Σ_n x[n] e^{SIGN·2πi nk/L}.
//
// L = L1·L2 (the factor pair of fft_sizes.hpp), input index n = L2·n1 + n2, output index
// k = k1 + L1·k2 (the two-level decomposition of the DFT):
//   A: for each n2, the L1-point transform over n1 of x[L2·n1 + n2]  -> Y[n2][k1]
//   B: Y[n2][k1] ·= e^{SIGN·2πi n2·k1 / L}
//   C: for each k1, the L2-point transform over n2 of Y[n2][k1]      -> X[k1 + L1·k2]
// A workgroup stages LPB whole lines in LDS (loaded with consecutive lanes on consecutive lines,
// which are adjacent in memory for the strided passes), runs the L2·LPB transforms of step A with
// the twiddles of step B in place, then the L1·LPB transforms of step C, and writes the outputs.
// The sub-transforms are the register FFTs of fft_reg.hpp; the twiddle table is computed per
// workgroup in FP64 (sincospi) and rounded to the transform's precision.
#include "common.hpp"
#include "fft_reg.hpp"
#include "fft_sizes.hpp"

namespace pwk {

namespace {

template <typename T>
struct LongC;
template <>
struct LongC<double> {
  using t = double2;
};
template <>
struct LongC<float> {
  using t = float2;
};

constexpr int kLongNT = 256;

template <int L1, int L2>
struct LongShape {
  static constexpr int L = L1 * L2;
  static constexpr int TASKS = L1 > L2 ? L1 : L2;                 // transforms per line in a step
  static constexpr int LPB = kLongNT / TASKS > 0 ? kLongNT / TASKS : 1;  // lines per workgroup
};

template <typename T, int L1, int L2>
size_t long_lds_bytes() {
  using S = LongShape<L1, L2>;
  return ((size_t)S::LPB * S::L + S::L) * sizeof(typename LongC<T>::t);
}

// Line l starts at (l / cnt_in)·st_out + (l % cnt_in)·st_in; elements at stride es; batch item
// blockIdx.y at batch_stride.
template <typename T, int L1, int L2, int SIGN>
__global__ void __launch_bounds__(kLongNT) k_lines_long(typename LongC<T>::t* __restrict__ data, int nlines,
                                                        int cnt_in, long st_out, long st_in, long es,
                                                        long batch_stride) {
  using V2 = typename LongC<T>::t;
  using S = LongShape<L1, L2>;
  constexpr int L = S::L, LPB = S::LPB;
  extern __shared__ unsigned char smem_raw[];
  V2* s = reinterpret_cast<V2*>(smem_raw);  // LPB lines of L
  V2* tw = s + (size_t)LPB * L;              // e^{SIGN·2πi e/L}, e < L
  const int l0 = blockIdx.x * LPB;
  const int nl = min(LPB, nlines - l0);
  V2* base = data + (size_t)blockIdx.y * batch_stride;
  auto line_ptr = [&](int j) {
    const int l = l0 + j;
    return base + (size_t)(l / cnt_in) * st_out + (size_t)(l % cnt_in) * st_in;
  };
  for (int e = threadIdx.x; e < L; e += kLongNT) {
    double sn, cs;
    sincospi(2.0 * (double)e / (double)L, &sn, &cs);
    tw[e].x = (T)cs;
    tw[e].y = (T)(SIGN * sn);
  }
  for (int idx = threadIdx.x; idx < nl * L; idx += kLongNT) {
    const int j = idx % nl, i = idx / nl;
    s[j * L + i] = line_ptr(j)[(size_t)i * es];
  }
  __syncthreads();
  // A and B: task (j, n2) transforms the L1 entries L2·n1 + n2 of line j in place
  for (int t = threadIdx.x; t < nl * L2; t += kLongNT) {
    const int j = t / L2, n2 = t % L2;
    V2* q = s + j * L;
    T re[L1], im[L1];
#pragma unroll
    for (int n1 = 0; n1 < L1; ++n1) {
      re[n1] = q[L2 * n1 + n2].x;
      im[n1] = q[L2 * n1 + n2].y;
    }
    fft_dif<T, L1, SIGN>(re, im);
#pragma unroll
    for (int k1 = 0; k1 < L1; ++k1) {
      const T a = re[FftOrder<L1>::pos(k1)], b = im[FftOrder<L1>::pos(k1)];
      const V2 w = tw[(n2 * k1) % L];
      V2 o;
      o.x = a * w.x - b * w.y;
      o.y = a * w.y + b * w.x;
      q[L2 * k1 + n2] = o;
    }
  }
  __syncthreads();
  // C: task (j, k1) transforms the L2 entries L2·k1 + n2 of line j, output k1 + L1·k2
  for (int t = threadIdx.x; t < nl * L1; t += kLongNT) {
    const int j = t / L1, k1 = t % L1;
    const V2* q = s + j * L + L2 * k1;
    T re[L2], im[L2];
#pragma unroll
    for (int n2 = 0; n2 < L2; ++n2) {
      re[n2] = q[n2].x;
      im[n2] = q[n2].y;
    }
    fft_dif<T, L2, SIGN>(re, im);
    V2* p = line_ptr(j);
#pragma unroll
    for (int k2 = 0; k2 < L2; ++k2) {
      V2 o;
      o.x = re[FftOrder<L2>::pos(k2)];
      o.y = im[FftOrder<L2>::pos(k2)];
      p[(size_t)(k1 + L1 * k2) * es] = o;
    }
  }
}

template <typename T, int SIGN>
void long_pass_t(hipStream_t st, void* data, int batch, long bs, int n, int nlines, int cnt_in, long st_out,
                 long st_in, long es) {
  auto* d = static_cast<typename LongC<T>::t*>(data);
  switch (n) {
#define PWK_CASE(LL, A, B)                                                                                 \
  case LL: {                                                                                               \
    const size_t sh = long_lds_bytes<T, A, B>();                                                           \
    const dim3 grid(blocks_for((size_t)nlines, LongShape<A, B>::LPB), (unsigned)batch);                    \
    ensure_lds_limit((const void*)k_lines_long<T, A, B, SIGN>, sh);                                        \
    k_lines_long<T, A, B, SIGN><<<grid, kLongNT, sh, st>>>(d, nlines, cnt_in, st_out, st_in, es, bs);      \
    break;                                                                                                 \
  }
    PWK_FOR_EACH_LONG_FFT_SIZE(PWK_CASE)
#undef PWK_CASE
    default:
      throw Error(PWK_E_UNSUPPORTED, "FFT length " + std::to_string(n) + " has no compiled kernel");
  }
  PWK_HIP(hipGetLastError());
}

}  // namespace

void launch_long_lines(hipStream_t st, void* data, int batch, long bs, int n, int nlines, int cnt_in, long st_out,
                       long st_in, long es, int sign, int precision) {
  if (precision == PWK_FP64)
    sign < 0 ? long_pass_t<double, -1>(st, data, batch, bs, n, nlines, cnt_in, st_out, st_in, es)
             : long_pass_t<double, 1>(st, data, batch, bs, n, nlines, cnt_in, st_out, st_in, es);
  else
    sign < 0 ? long_pass_t<float, -1>(st, data, batch, bs, n, nlines, cnt_in, st_out, st_in, es)
             : long_pass_t<float, 1>(st, data, batch, bs, n, nlines, cnt_in, st_out, st_in, es);
}

}  // namespace pwk
