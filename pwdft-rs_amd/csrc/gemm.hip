// Complex GEMM, column-major: C (+)= A·B with A M×K, B K×N. LDS-tiled FMA
// kernel with a deterministic split-K (partials reduced in a fixed order) for
// the tall-skinny projector products of the nonlocal term.
#include <algorithm>

#include "common.hpp"

namespace pwk {

namespace {

template <typename T>
struct V2g;
template <>
struct V2g<double> {
  using t = double2;
};
template <>
struct V2g<float> {
  using t = float2;
};

constexpr int BM = 64, BN = 64, BK = 16;

template <typename T>
__global__ void __launch_bounds__(256) k_zgemm(int M, int N, int K, int kchunk, const typename V2g<T>::t* __restrict__ A,
                                               int lda, const typename V2g<T>::t* __restrict__ B, int ldb,
                                               typename V2g<T>::t* __restrict__ C, int ldc, int accumulate,
                                               typename V2g<T>::t* __restrict__ partial) {
  using V2 = typename V2g<T>::t;
  __shared__ V2 As[BK][BM + 1];
  __shared__ V2 Bs[BK][BN + 1];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int kbeg = blockIdx.z * kchunk;
  const int kend = min(K, kbeg + kchunk);
  T accr[4][4], acci[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      accr[i][j] = 0;
      acci[i][j] = 0;
    }
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    // A tile: 64 rows × 16 k; 256 threads × 4 elements
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      int idx = threadIdx.x + e * 256;
      int mm = idx & 63, kk = idx >> 6;
      int gm = m0 + mm, gk = k0 + kk;
      V2 v;
      v.x = 0;
      v.y = 0;
      if (gm < M && gk < kend) v = A[(size_t)gk * lda + gm];
      As[kk][mm] = v;
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      int idx = threadIdx.x + e * 256;
      int kk = idx & 15, nn = idx >> 4;
      int gk = k0 + kk, gn = n0 + nn;
      V2 v;
      v.x = 0;
      v.y = 0;
      if (gn < N && gk < kend) v = B[(size_t)gn * ldb + gk];
      Bs[kk][nn] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; ++kk) {
      V2 a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][tx + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kk][ty + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          accr[i][j] += a[i].x * b[j].x - a[i].y * b[j].y;
          acci[i][j] += a[i].x * b[j].y + a[i].y * b[j].x;
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      int gm = m0 + tx + 16 * i, gn = n0 + ty + 16 * j;
      if (gm >= M || gn >= N) continue;
      if (partial) {
        V2 o;
        o.x = accr[i][j];
        o.y = acci[i][j];
        partial[((size_t)blockIdx.z * N + gn) * M + gm] = o;
      } else {
        V2* c = &C[(size_t)gn * ldc + gm];
        V2 o;
        o.x = accr[i][j];
        o.y = acci[i][j];
        if (accumulate) {
          V2 old = *c;
          o.x += old.x;
          o.y += old.y;
        }
        *c = o;
      }
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_splitk_reduce(int M, int N, int nsplit,
                                                       const typename V2g<T>::t* __restrict__ partial,
                                                       typename V2g<T>::t* __restrict__ C, int ldc, int accumulate) {
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)M * N) return;
  int gm = (int)(idx % M), gn = (int)(idx / M);
  typename V2g<T>::t s;
  s.x = 0;
  s.y = 0;
  for (int z = 0; z < nsplit; ++z) {
    typename V2g<T>::t p = partial[(size_t)z * M * N + idx];
    s.x += p.x;
    s.y += p.y;
  }
  typename V2g<T>::t* c = &C[(size_t)gn * ldc + gm];
  if (accumulate) {
    typename V2g<T>::t old = *c;
    s.x += old.x;
    s.y += old.y;
  }
  *c = s;
}

template <typename T>
void zgemm_t(Context& c, int M, int N, int K, const void* A, int lda, const void* B, int ldb, void* C, int ldc,
             bool accumulate) {
  using V2 = typename V2g<T>::t;
  int tiles = ((M + BM - 1) / BM) * ((N + BN - 1) / BN);
  int target = 2 * (c.cu_count > 0 ? c.cu_count : 256);
  int nsplit = 1;
  if (tiles < target) {
    nsplit = (target + tiles - 1) / tiles;
    nsplit = std::min(nsplit, std::max(1, K / (4 * BK)));
  }
  int kchunk = (K + nsplit - 1) / nsplit;
  kchunk = ((kchunk + BK - 1) / BK) * BK;
  nsplit = (K + kchunk - 1) / kchunk;
  dim3 grid((unsigned)((M + BM - 1) / BM), (unsigned)((N + BN - 1) / BN), (unsigned)nsplit);
  if (nsplit == 1) {
    k_zgemm<T><<<grid, 256, 0, c.stream>>>(M, N, K, kchunk, (const V2*)A, lda, (const V2*)B, ldb, (V2*)C, ldc,
                                           accumulate ? 1 : 0, nullptr);
  } else {
    c.gemm_ws.ensure((size_t)nsplit * M * N * sizeof(V2));
    k_zgemm<T><<<grid, 256, 0, c.stream>>>(M, N, K, kchunk, (const V2*)A, lda, (const V2*)B, ldb, (V2*)C, ldc,
                                           0, c.gemm_ws.as<V2>());
    PWK_HIP(hipGetLastError());
    k_splitk_reduce<T><<<blocks_for((size_t)M * N, 256), 256, 0, c.stream>>>(M, N, nsplit, c.gemm_ws.as<V2>(),
                                                                             (V2*)C, ldc, accumulate ? 1 : 0);
  }
  PWK_HIP(hipGetLastError());
}

}  // namespace

void launch_zgemm(Context& c, int M, int N, int K, const void* A, int lda, const void* B, int ldb, void* C, int ldc,
                  bool accumulate, int precision) {
  if (M <= 0 || N <= 0) return;
  if (K <= 0) {
    PWK_REQUIRE(accumulate, PWK_E_INVALID, "zgemm with K = 0 and no accumulation");
    return;
  }
  if (precision == PWK_FP64)
    zgemm_t<double>(c, M, N, K, A, lda, B, ldb, C, ldc, accumulate);
  else
    zgemm_t<float>(c, M, N, K, A, lda, B, ldb, C, ldc, accumulate);
}

}  // namespace pwk
