// Full-grid batched 3D FFT (src/fft.rs:19-114 convention): three sets of
// 1D transforms, one lane per line, the line held in registers.
#include "common.hpp"
#include "fft_reg.hpp"
#include "fft_sizes.hpp"

namespace pwk {

namespace {

template <typename T>
struct Vec2;
template <>
struct Vec2<double> {
  using t = double2;
};
template <>
struct Vec2<float> {
  using t = float2;
};

// Line l starts at (l / cnt_in)·st_out + (l % cnt_in)·st_in; elements at stride es.
template <typename T, int N, int SIGN>
__global__ void __launch_bounds__(64) k_lines(typename Vec2<T>::t* __restrict__ data, int nlines, int cnt_in,
                                              long st_out, long st_in, long es, long batch_stride) {
  int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nlines) return;
  typename Vec2<T>::t* p = data + (size_t)blockIdx.y * batch_stride + (size_t)(l / cnt_in) * st_out +
                           (size_t)(l % cnt_in) * st_in;
  T re[N], im[N];
#pragma unroll
  for (int j = 0; j < N; ++j) {
    typename Vec2<T>::t v = p[(size_t)j * es];
    re[j] = v.x;
    im[j] = v.y;
  }
  fft_dif<T, N, SIGN>(re, im);
#pragma unroll
  for (int f = 0; f < N; ++f) {
    typename Vec2<T>::t v;
    v.x = re[FftOrder<N>::pos(f)];
    v.y = im[FftOrder<N>::pos(f)];
    p[(size_t)f * es] = v;
  }
}

template <typename T, int SIGN>
void lines(Context& c, void* data, int batch, int n, int nlines, int cnt_in, long st_out, long st_in, long es) {
  dim3 grid(blocks_for((size_t)nlines, 64), (unsigned)batch);
  auto* d = static_cast<typename Vec2<T>::t*>(data);
  long bs = (long)c.N;
  switch (n) {
#define PWK_CASE(L)                                                                              \
  case L:                                                                                        \
    k_lines<T, L, SIGN><<<grid, 64, 0, c.stream>>>(d, nlines, cnt_in, st_out, st_in, es, bs); \
    break;
    PWK_FOR_EACH_FFT_SIZE(PWK_CASE)
#undef PWK_CASE
    default:
      throw Error(PWK_E_UNSUPPORTED, "FFT length " + std::to_string(n) + " has no compiled kernel");
  }
  PWK_HIP(hipGetLastError());
}

template <typename T, int SIGN>
void fft3d_t(Context& c, void* data, int batch) {
  int nx = c.dims[0], ny = c.dims[1], nz = c.dims[2];
  // x (stride ny·nz), y (stride nz), z (contiguous); the order does not change the result
  lines<T, SIGN>(c, data, batch, nx, ny * nz, ny * nz, 0, 1, (long)ny * nz);
  lines<T, SIGN>(c, data, batch, ny, nx * nz, nz, (long)ny * nz, 1, nz);
  lines<T, SIGN>(c, data, batch, nz, nx * ny, 1, nz, 0, 1);
}

}  // namespace

bool fft_size_supported(int n) { return fft_len_compiled(n); }

void launch_fft3d(Context& c, void* data, int batch, int sign, int precision) {
  c.check_grid();
  for (int d = 0; d < 3; ++d)
    PWK_REQUIRE(fft_len_compiled(c.dims[d]), PWK_E_UNSUPPORTED,
                "FFT length " + std::to_string(c.dims[d]) + " has no compiled kernel");
  if (precision == PWK_FP64) {
    if (sign < 0)
      fft3d_t<double, -1>(c, data, batch);
    else
      fft3d_t<double, 1>(c, data, batch);
  } else {
    if (sign < 0)
      fft3d_t<float, -1>(c, data, batch);
    else
      fft3d_t<float, 1>(c, data, batch);
  }
}

}  // namespace pwk
