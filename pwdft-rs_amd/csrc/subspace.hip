// Block kernels of the Chebyshev-filtered subspace eigensolver: the filter's
// three-term recurrence step, conjugate transposes for the Gram products
// Y^H·Y and Y^H·(HY), Hermitian symmetrisation and the Löwdin column scaling.
// All are HBM/L2-bound elementwise passes over npw × nb blocks.
#include "common.hpp"

namespace pwk {

namespace {

template <typename T>
struct C2;
template <>
struct C2<float> {
  using t = float2;
};
template <>
struct C2<double> {
  using t = double2;
};

// out = alpha·(HY - c·Y) - beta·X   (X may be null)
template <typename T>
__global__ void __launch_bounds__(256) k_cheb_step(size_t n, const typename C2<T>::t* __restrict__ HY,
                                                   const typename C2<T>::t* __restrict__ Y,
                                                   const typename C2<T>::t* __restrict__ X, T alpha, T cshift, T beta,
                                                   typename C2<T>::t* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const auto h = HY[i], y = Y[i];
  typename C2<T>::t o;
  o.x = alpha * (h.x - cshift * y.x);
  o.y = alpha * (h.y - cshift * y.y);
  if (X) {
    const auto x = X[i];
    o.x -= beta * x.x;
    o.y -= beta * x.y;
  }
  out[i] = o;
}

// out (cols × rows, column-major, ld ldo) = conj(A)^T of A (rows × cols, ld lda)
__global__ void __launch_bounds__(256) k_conj_transpose(int rows, int cols, const double2* __restrict__ A, int lda,
                                                        double2* __restrict__ out, int ldo) {
  __shared__ double2 tile[32][33];
  const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 × 8
  for (int j = ty; j < 32; j += 8) {
    const int r = r0 + tx, cc = c0 + j;
    if (r < rows && cc < cols) tile[j][tx] = A[(size_t)cc * lda + r];
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int cc = c0 + tx, r = r0 + j;  // out row = cc, out column = r
    if (r < rows && cc < cols) {
      const double2 v = tile[tx][j];
      out[(size_t)r * ldo + cc] = make_double2(v.x, -v.y);
    }
  }
}

// A = (A + A^H)/2 for an m × m column-major block
__global__ void __launch_bounds__(256) k_hermitize(int m, double2* __restrict__ A) {
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (size_t)m * m) return;
  const int i = (int)(idx % m), j = (int)(idx / m);
  if (i > j) return;
  const double2 a = A[(size_t)j * m + i], b = A[(size_t)i * m + j];
  const double2 h = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
  A[(size_t)j * m + i] = h;
  A[(size_t)i * m + j] = make_double2(h.x, -h.y);
}

// U[:, j] *= 1/sqrt(max(s_j, floor))  (Löwdin S^{-1/2} factor)
__global__ void __launch_bounds__(256) k_lowdin_scale(int m, const double* __restrict__ s, double floor_rel,
                                                      double2* __restrict__ U) {
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (size_t)m * m) return;
  const int j = (int)(idx / m);
  const double smax = s[m - 1];
  const double sj = fmax(s[j], floor_rel * smax);
  const double f = 1.0 / sqrt(sj);
  U[idx].x *= f;
  U[idx].y *= f;
}

// Deterministic start block: column j = unit vector on the j-th lowest kinetic
// energy plane wave plus a small seeded perturbation everywhere.
__global__ void __launch_bounds__(256) k_start_block(int npw, int nb, const int* __restrict__ lowest,
                                                     double2* __restrict__ V) {
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (size_t)npw * nb) return;
  const int g = (int)(idx % npw), j = (int)(idx / npw);
  uint32_t h = (uint32_t)g * 0x9E3779B1u ^ ((uint32_t)j * 0x85EBCA77u + 0x165667B1u);
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  h *= 0x297A2D39u;
  h ^= h >> 15;
  const double r1 = (double)(h & 0xFFFF) / 65536.0 - 0.5, r2 = (double)(h >> 16) / 65536.0 - 0.5;
  double2 v = make_double2(1e-2 * r1, 1e-2 * r2);
  if (g == lowest[j]) v.x += 1.0;
  V[idx] = v;
}

}  // namespace

void launch_cheb_step(Context& c, size_t n, const void* HY, const void* Y, const void* X, double alpha, double cshift,
                      double beta, void* out, int precision) {
  if (precision == PWK_FP64)
    k_cheb_step<double><<<blocks_for(n, 256), 256, 0, c.stream>>>(n, (const double2*)HY, (const double2*)Y,
                                                                  (const double2*)X, alpha, cshift, beta,
                                                                  (double2*)out);
  else
    k_cheb_step<float><<<blocks_for(n, 256), 256, 0, c.stream>>>(n, (const float2*)HY, (const float2*)Y,
                                                                 (const float2*)X, (float)alpha, (float)cshift,
                                                                 (float)beta, (float2*)out);
  PWK_HIP(hipGetLastError());
}

void launch_conj_transpose(Context& c, int rows, int cols, const void* A, int lda, void* out, int ldo) {
  dim3 g(blocks_for((size_t)rows, 32), blocks_for((size_t)cols, 32));
  k_conj_transpose<<<g, 256, 0, c.stream>>>(rows, cols, (const double2*)A, lda, (double2*)out, ldo);
  PWK_HIP(hipGetLastError());
}

void launch_hermitize(Context& c, int m, void* A) {
  k_hermitize<<<blocks_for((size_t)m * m, 256), 256, 0, c.stream>>>(m, (double2*)A);
  PWK_HIP(hipGetLastError());
}

void launch_lowdin_scale(Context& c, int m, const double* s, double floor_rel, void* U) {
  k_lowdin_scale<<<blocks_for((size_t)m * m, 256), 256, 0, c.stream>>>(m, s, floor_rel, (double2*)U);
  PWK_HIP(hipGetLastError());
}

void launch_start_block(Context& c, int npw, int nb, const int* lowest_dev, void* V) {
  k_start_block<<<blocks_for((size_t)npw * nb, 256), 256, 0, c.stream>>>(npw, nb, lowest_dev, (double2*)V);
  PWK_HIP(hipGetLastError());
}

}  // namespace pwk
