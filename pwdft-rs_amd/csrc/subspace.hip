// Block kernels of the Chebyshev-filtered subspace eigensolver, batched over
// k-points: the filter's three-term recurrence step, conjugate transposes for
// the Gram products Y^H·Y and Y^H·(HY), Hermitian symmetrisation, the Löwdin
// column scaling, per-block dot products and the seeded start block. All are
// HBM/L2-bound elementwise passes over nk blocks of npw × nb.
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

// out = α_k·(HY - c_k·Y) - β_k·X on block k = i / per (X may be null); coef holds
// (α, c, β) per block.
template <typename T>
__global__ void __launch_bounds__(256) k_cheb_step(size_t n, size_t per, const typename C2<T>::t* __restrict__ HY,
                                                   const typename C2<T>::t* __restrict__ Y,
                                                   const typename C2<T>::t* __restrict__ X,
                                                   const double* __restrict__ coef, typename C2<T>::t* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const size_t k = i / per;
  const T alpha = (T)coef[3 * k], cshift = (T)coef[3 * k + 1], beta = (T)coef[3 * k + 2];
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

// out_b (cols × rows, column-major, ld ldo) = conj(A_b)^T, A_b rows × cols (ld lda)
__global__ void __launch_bounds__(256) k_conj_transpose(int rows, int cols, const double2* __restrict__ A, int lda,
                                                        long long sA, double2* __restrict__ out, int ldo,
                                                        long long sO) {
  __shared__ double2 tile[32][33];
  A += (size_t)blockIdx.z * sA;
  out += (size_t)blockIdx.z * sO;
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

// A_b = (A_b + A_b^H)/2 for m × m column-major blocks
__global__ void __launch_bounds__(256) k_hermitize(int m, double2* __restrict__ A) {
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (size_t)m * m) return;
  A += (size_t)blockIdx.y * m * m;
  const int i = (int)(idx % m), j = (int)(idx / m);
  if (i > j) return;
  const double2 a = A[(size_t)j * m + i], b = A[(size_t)i * m + j];
  const double2 h = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
  A[(size_t)j * m + i] = h;
  A[(size_t)i * m + j] = make_double2(h.x, -h.y);
}

// U_b[:, j] *= 1/sqrt(max(s_bj, floor·s_b,max))  (Löwdin S^{-1/2} factor; s ascending)
__global__ void __launch_bounds__(256) k_lowdin_scale(int m, const double* __restrict__ s, double floor_rel,
                                                      double2* __restrict__ U) {
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (size_t)m * m) return;
  s += (size_t)blockIdx.y * m;
  U += (size_t)blockIdx.y * m * m;
  const int j = (int)(idx / m);
  const double sj = fmax(s[j], floor_rel * s[m - 1]);
  const double f = 1.0 / sqrt(sj);
  U[idx].x *= f;
  U[idx].y *= f;
}

// out[b] = Σ_i x_b[i]·y_b[i] over per reals of block b (one workgroup per block, fixed order)
__global__ void __launch_bounds__(256) k_block_dot(size_t per, const double* __restrict__ x,
                                                   const double* __restrict__ y, double* __restrict__ out) {
  __shared__ double sh[4];
  const size_t base = (size_t)blockIdx.x * per;
  double acc = 0.0;
  for (size_t i = threadIdx.x; i < per; i += 256) acc += x[base + i] * y[base + i];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// x_b *= f(s_b) per block: mode 0 multiplies by s, mode 1 by 1/sqrt(s)
__global__ void __launch_bounds__(256) k_block_scale(size_t n, size_t per, const double* __restrict__ s, int mode,
                                                     double* __restrict__ x) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double v = s[i / per];
  x[i] *= mode == 0 ? v : 1.0 / sqrt(v);
}

// out[k·nbs + b] = |HV_kb - λ_kb V_kb|, one workgroup per (band, k) for the first nres bands
__global__ void __launch_bounds__(256) k_resid_norms(int npw, int nbs, const double2* __restrict__ V,
                                                     const double2* __restrict__ HV, const double* __restrict__ lam,
                                                     double* __restrict__ out) {
  __shared__ double sh[4];
  const int b = blockIdx.x, k = blockIdx.y;
  const size_t off = ((size_t)k * nbs + b) * npw;
  const double l = lam[(size_t)k * nbs + b];
  double acc = 0.0;
  for (int i = threadIdx.x; i < npw; i += 256) {
    const double2 h = HV[off + i], v = V[off + i];
    const double rx = h.x - l * v.x, ry = h.y - l * v.y;
    acc += rx * rx + ry * ry;
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[(size_t)k * nbs + b] = sqrt((sh[0] + sh[1]) + (sh[2] + sh[3]));
}

// Deterministic start block per k: column j = unit vector on plane wave lowest[k][j]
// (skipped when negative) plus a small seeded perturbation everywhere.
__global__ void __launch_bounds__(256) k_start_block(int npw, int nb, const int* __restrict__ lowest,
                                                     double2* __restrict__ V) {
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (size_t)npw * nb) return;
  const int k = blockIdx.y;
  const int g = (int)(idx % npw), j = (int)(idx / npw);
  uint32_t h = (uint32_t)g * 0x9E3779B1u ^ ((uint32_t)j * 0x85EBCA77u + 0x165667B1u) ^ ((uint32_t)k * 0xC2B2AE3Du);
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  h *= 0x297A2D39u;
  h ^= h >> 15;
  const double r1 = (double)(h & 0xFFFF) / 65536.0 - 0.5, r2 = (double)(h >> 16) / 65536.0 - 0.5;
  double2 v = make_double2(1e-2 * r1, 1e-2 * r2);
  if (g == lowest[(size_t)k * nb + j]) v.x += 1.0;
  V[(size_t)k * npw * nb + idx] = v;
}

}  // namespace

void launch_cheb_step(Context& c, size_t n, size_t per, const void* HY, const void* Y, const void* X,
                      const double* coef_dev, void* out, int precision) {
  if (precision == PWK_FP64)
    k_cheb_step<double><<<blocks_for(n, 256), 256, 0, c.stream>>>(n, per, (const double2*)HY, (const double2*)Y,
                                                                  (const double2*)X, coef_dev, (double2*)out);
  else
    k_cheb_step<float><<<blocks_for(n, 256), 256, 0, c.stream>>>(n, per, (const float2*)HY, (const float2*)Y,
                                                                 (const float2*)X, coef_dev, (float2*)out);
  PWK_HIP(hipGetLastError());
}

void launch_conj_transpose(Context& c, int rows, int cols, const void* A, int lda, long long sA, void* out, int ldo,
                           long long sO, int batch) {
  dim3 g(blocks_for((size_t)rows, 32), blocks_for((size_t)cols, 32), (unsigned)batch);
  k_conj_transpose<<<g, 256, 0, c.stream>>>(rows, cols, (const double2*)A, lda, sA, (double2*)out, ldo, sO);
  PWK_HIP(hipGetLastError());
}

void launch_hermitize(Context& c, int m, void* A, int batch) {
  dim3 g(blocks_for((size_t)m * m, 256), (unsigned)batch);
  k_hermitize<<<g, 256, 0, c.stream>>>(m, (double2*)A);
  PWK_HIP(hipGetLastError());
}

void launch_lowdin_scale(Context& c, int m, const double* s, double floor_rel, void* U, int batch) {
  dim3 g(blocks_for((size_t)m * m, 256), (unsigned)batch);
  k_lowdin_scale<<<g, 256, 0, c.stream>>>(m, s, floor_rel, (double2*)U);
  PWK_HIP(hipGetLastError());
}

void launch_block_dot(Context& c, size_t per, int batch, const double* x, const double* y, double* out) {
  k_block_dot<<<(unsigned)batch, 256, 0, c.stream>>>(per, x, y, out);
  PWK_HIP(hipGetLastError());
}

void launch_block_scale(Context& c, size_t per, int batch, const double* s, int mode, double* x) {
  const size_t n = per * (size_t)batch;
  k_block_scale<<<blocks_for(n, 256), 256, 0, c.stream>>>(n, per, s, mode, x);
  PWK_HIP(hipGetLastError());
}

void launch_resid_norms(Context& c, int npw, int nbs, int nres, int nk, const void* V, const void* HV,
                        const double* lam, double* out) {
  dim3 g((unsigned)nres, (unsigned)nk);
  k_resid_norms<<<g, 256, 0, c.stream>>>(npw, nbs, (const double2*)V, (const double2*)HV, lam, out);
  PWK_HIP(hipGetLastError());
}

void launch_start_block(Context& c, int npw, int nb, int nk, const int* lowest_dev, void* V) {
  dim3 g(blocks_for((size_t)npw * nb, 256), (unsigned)nk);
  k_start_block<<<g, 256, 0, c.stream>>>(npw, nb, lowest_dev, (double2*)V);
  PWK_HIP(hipGetLastError());
}

}  // namespace pwk
