// 3D FFT entry points (src/fft.rs:19-114 convention): small cubic grids as one workgroup per
// batch item with the grid in LDS; other grids through the planned line passes (autotune.hip).
#include "common.hpp"
#include "fft_reg.hpp"
#include "fft_sizes.hpp"
#include "lds_fft.hpp"

namespace pwk {

namespace {

// Small cubic grids: the whole transform of one batch item per workgroup, grid in LDS, one
// launch for the three passes. With real_in the input is real_in·scale + 0i (FP64 only).
template <typename T, int N, int SIGN>
__global__ void __launch_bounds__(SMALL_NT) k_fft3d_small(typename V2s<T>::t* __restrict__ data,
                                                          const double* __restrict__ real_in, double scale,
                                                          const typename V2s<T>::t* __restrict__ src) {
  using V2 = typename V2s<T>::t;
  extern __shared__ unsigned char smem_raw[];
  V2* G = reinterpret_cast<V2*>(smem_raw);
  constexpr int NN = N * N * N;
  V2* d = data + (size_t)blockIdx.x * NN;
  if (real_in) {
    grid_points<NN, double>([&](int i) { return real_in[i]; },
                            [&](int i, double x) {
                              G[i].x = (T)(x * scale);
                              G[i].y = 0;
                            });
  } else {
    const V2* in = src ? src + (size_t)blockIdx.x * NN : d;  // out of place with src
    grid_points<NN, V2>([&](int i) { return in[i]; }, [&](int i, V2 x) { G[i] = x; });
  }
  __syncthreads();
  lds_lines<T, N, 0, SIGN>(G, nullptr);
  __syncthreads();
  lds_lines<T, N, 1, SIGN>(G, nullptr);
  __syncthreads();
  lds_lines<T, N, 2, SIGN>(G, nullptr);
  __syncthreads();
  grid_points<NN, V2>([&](int i) { return G[i]; }, [&](int i, V2 x) { d[i] = x; });
}

#define PWK_FFT_SMALL_SIZES(X) X(8) X(9) X(10) X(12) X(15) X(16) X(18) X(20) X(24)

bool fft_small_len(int n) {
  switch (n) {
#define PWK_CASE(L) case L:
    PWK_FFT_SMALL_SIZES(PWK_CASE)
#undef PWK_CASE
    return true;
    default:
      return false;
  }
}

bool fft_small_ok(const Context& c, int precision) {
  const int n = c.dims[0];
  if (!c.small_hpsi || n != c.dims[1] || n != c.dims[2] || !fft_small_len(n)) return false;
  const size_t es = precision == PWK_FP64 ? sizeof(double2) : sizeof(float2);
  return (size_t)n * n * n * es <= 160 * 1024;
}

template <typename T, int SIGN>
void fft3d_small_t(Context& c, void* data, int batch, const double* real_in, double scale,
                   const void* src = nullptr) {
  using V2 = typename V2s<T>::t;
  const int n = c.dims[0];
  const size_t sh = (size_t)n * n * n * sizeof(V2);
  switch (n) {
#define PWK_CASE(L)                                                                                          \
  case L:                                                                                                    \
    ensure_lds_limit((const void*)k_fft3d_small<T, L, SIGN>, sh);                                           \
    k_fft3d_small<T, L, SIGN><<<(unsigned)batch, SMALL_NT, sh, c.stream>>>((V2*)data, real_in, scale,        \
                                                                          (const V2*)src);                 \
    break;
    PWK_FFT_SMALL_SIZES(PWK_CASE)
#undef PWK_CASE
    default:
      throw Error(PWK_E_UNSUPPORTED, "grid has no fused FFT kernel");
  }
  PWK_HIP(hipGetLastError());
}

__global__ void k_r2c_scale(size_t n, const double* __restrict__ r, double s, double2* __restrict__ c) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) c[i] = make_double2(r[i] * s, 0.0);
}

}  // namespace

bool fft_size_supported(int n) { return fft_len_supported(n); }

void launch_fft3d_r2c(Context& c, const double* r, double scale, void* out, int sign) {
  c.check_grid();
  if (fft_small_ok(c, PWK_FP64)) {
    if (sign < 0)
      fft3d_small_t<double, -1>(c, out, 1, r, scale);
    else
      fft3d_small_t<double, 1>(c, out, 1, r, scale);
    return;
  }
  k_r2c_scale<<<blocks_for(c.N, 256), 256, 0, c.stream>>>(c.N, r, scale, (double2*)out);
  PWK_HIP(hipGetLastError());
  launch_fft3d(c, out, 1, sign, PWK_FP64);
}

void launch_fft3d_copy(Context& c, const void* in, void* out, int sign) {
  c.check_grid();
  if (fft_small_ok(c, PWK_FP64)) {
    if (sign < 0)
      fft3d_small_t<double, -1>(c, out, 1, nullptr, 1.0, in);
    else
      fft3d_small_t<double, 1>(c, out, 1, nullptr, 1.0, in);
    return;
  }
  PWK_HIP(hipMemcpyAsync(out, in, c.N * sizeof(double2), hipMemcpyDeviceToDevice, c.stream));
  launch_fft3d(c, out, 1, sign, PWK_FP64);
}

void launch_fft3d(Context& c, void* data, int batch, int sign, int precision) {
  c.check_grid();
  for (int d = 0; d < 3; ++d)
    PWK_REQUIRE(fft_len_supported(c.dims[d]), PWK_E_UNSUPPORTED,
                "FFT length " + std::to_string(c.dims[d]) + " has no compiled kernel");
  const int fprec = precision == PWK_FP64 ? PWK_FP64 : PWK_FP32;
  if (fft_small_ok(c, fprec)) {
    if (fprec == PWK_FP64) {
      if (sign < 0)
        fft3d_small_t<double, -1>(c, data, batch, nullptr, 1.0);
      else
        fft3d_small_t<double, 1>(c, data, batch, nullptr, 1.0);
    } else {
      if (sign < 0)
        fft3d_small_t<float, -1>(c, data, batch, nullptr, 1.0);
      else
        fft3d_small_t<float, 1>(c, data, batch, nullptr, 1.0);
    }
    return;
  }
  fft3d_planned(c, data, batch, sign, fprec);
}

}  // namespace pwk
