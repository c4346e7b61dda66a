// Complex GEMM on the matrix cores, column-major: C (+)= A·B, A M×K, B K×N.
//
// The complex product is one real product of the embedded matrices
//   Ã (2M × 2K): [[Re a, -Im a], [Im a, Re a]] per entry,   B̃ (2K × N) = interleaved B,
//   C̃ (2M × N) = interleaved C,
// so B and C keep their interleaved (re, im) storage and only A is expanded,
// in registers on its way into LDS. Block tile: 64 complex rows × 16·NB16
// columns, 4 waves, each wave 32 real rows × 16·NB16 columns of 16×16 MFMA
// blocks. K is staged 16 complex (32 real) at a time.
//   MODE_F32 : v_mfma_f32_16x16x4_f32
//   MODE_F64 : v_mfma_f64_16x16x4_f64
//   MODE_BF16X3 : v_mfma_f32_16x16x32_bf16 on the split x = hi + lo,
//                 hi = bf16(x), lo = bf16(x - hi); C += Ahi·Bhi + Ahi·Blo + Alo·Bhi (FP32 accumulate)
// Split-K partial sums are reduced in a fixed order (deterministic).
#include <algorithm>
#include <type_traits>

#include "common.hpp"

namespace pwk {

namespace {

typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));

enum { MODE_F32 = 0, MODE_F64 = 1, MODE_BF16X3 = 2 };

constexpr int BMC = 64;  // complex rows per block
constexpr int BMR = 128; // real rows per block
constexpr int BKC = 16;  // complex K per stage
constexpr int BKR = 32;  // real K per stage

template <int MODE>
struct Traits {
  using T = float;
  using V2 = float2;
  using acc_t = f4v;
};
template <>
struct Traits<MODE_F64> {
  using T = double;
  using V2 = double2;
  using acc_t = d4v;
};

__device__ __forceinline__ __bf16 to_bf16(float x) { return (__bf16)x; }

// LDS row strides (elements). A rows keep one padding element (the expanded A tile is written
// column-wise, rows 33 apart spread the lanes over the banks); FP32 B rows are padded to 36
// floats so each lane's 8 k values are two aligned 16-byte reads (B is written along k, so
// its stores stay conflict-free); bf16 rows carry eight.
template <int MODE>
struct Lds {
  static constexpr int AK = (MODE == MODE_BF16X3) ? BKR + 8 : BKR + 1;                                // A [row][k]
  static constexpr int BK = (MODE == MODE_BF16X3) ? BKR + 8 : (MODE == MODE_F32 ? BKR + 4 : BKR + 1);  // B [col][k]
};

template <int MODE, int NB16>
__global__ void __launch_bounds__(256) k_zgemm_mfma(int M, int N, int K, int kchunk, int nsplit, long long sA,
                                                    long long sB, long long sC,
                                                    const typename Traits<MODE>::V2* __restrict__ A, int lda,
                                                    const typename Traits<MODE>::V2* __restrict__ B, int ldb,
                                                    typename Traits<MODE>::V2* __restrict__ C, int ldc, int accumulate,
                                                    typename Traits<MODE>::V2* __restrict__ partial) {
  using T = typename Traits<MODE>::T;
  using V2 = typename Traits<MODE>::V2;
  using acc_t = typename Traits<MODE>::acc_t;
  constexpr int BN = 16 * NB16;
  constexpr int AK = Lds<MODE>::AK, BKS = Lds<MODE>::BK;
  constexpr bool SPLIT = MODE == MODE_BF16X3;
  using ldsT = typename std::conditional<SPLIT, __bf16, T>::type;
  constexpr int NPART = SPLIT ? 2 : 1;
  __shared__ __attribute__((aligned(16))) ldsT As[NPART][BMR * AK];
  __shared__ __attribute__((aligned(16))) ldsT Bs[NPART][BN * BKS];

  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  // One-dimensional grid over (row tile, column tile, batch · nsplit + split). Workgroup i runs on
  // XCD i mod 8, and each XCD takes a contiguous range of tiles in an order where neighbours
  // share an operand in that XCD's L2: with split K, the row and column tiles of one K chunk
  // (they read the same B rows); without, the column tiles of one row tile (the same A rows).
  const int mt = (M + BMC - 1) / BMC, nt = (N + BN - 1) / BN, plane = mt * nt;
  const int total = (int)gridDim.x, xcd = (int)blockIdx.x & 7;
  const int lg = xcd * (total >> 3) + min(xcd, total & 7) + ((int)blockIdx.x >> 3);
  const int bz = lg / plane, rem = lg % plane;
  const int bx = nsplit > 1 ? rem % mt : rem / nt, by = nsplit > 1 ? rem / mt : rem % nt;
  const int m0 = bx * BMC, n0 = by * BN;
  // bz = batch · nsplit + split
  const int zb = bz / nsplit, zs = bz % nsplit;
  A += (size_t)zb * sA;
  B += (size_t)zb * sB;
  C += (size_t)zb * sC;
  if (partial) partial += (size_t)zb * nsplit * M * N;
  const int kbeg = zs * kchunk;
  const int kend = min(K, kbeg + kchunk);

  acc_t acc[2][NB16];
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int cb = 0; cb < NB16; ++cb)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[rb][cb][q] = 0;

  // global -> register prefetch of the next K stage overlaps the MFMA work of the current one
  constexpr int NA = (BMC * BKC) / 256, NBV = (BKC * BN + 255) / 256;
  V2 ra[NA], rbv[NBV];
  auto gload = [&](int k0) {
#pragma unroll
    for (int e = 0; e < NA; ++e) {
      const int idx = tid + 256 * e;
      const int mm = idx & (BMC - 1), kk = idx / BMC;
      const int gm = m0 + mm, gk = k0 + kk;
      V2 v;
      v.x = 0;
      v.y = 0;
      if (gm < M && gk < kend) v = A[(size_t)gk * lda + gm];
      ra[e] = v;
    }
#pragma unroll
    for (int e = 0; e < NBV; ++e) {
      const int idx = tid + 256 * e;
      V2 v;
      v.x = 0;
      v.y = 0;
      if (idx < BKC * BN) {
        const int kk = idx % BKC, nn = idx / BKC;
        const int gk = k0 + kk, gn = n0 + nn;
        if (gn < N && gk < kend) v = B[(size_t)gn * ldb + gk];
      }
      rbv[e] = v;
    }
  };
  // A tile: 64 complex rows × 16 complex k, expanded to 128 × 32 real; B tile: 16 complex k ×
  // BN columns, interleaved -> [col][2k+part]
  auto lstore = [&]() {
#pragma unroll
    for (int e = 0; e < NA; ++e) {
      const int idx = tid + 256 * e;
      const int mm = idx & (BMC - 1), kk = idx / BMC;
      const V2 v = ra[e];
      T vals[4] = {v.x, -v.y, v.y, v.x};  // (2m,2k) (2m,2k+1) (2m+1,2k) (2m+1,2k+1)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int r = 2 * mm + (t >> 1), c = 2 * kk + (t & 1);
        if constexpr (SPLIT) {
          __bf16 hi = to_bf16(vals[t]);
          As[0][r * AK + c] = hi;
          As[1][r * AK + c] = to_bf16(vals[t] - (float)hi);
        } else {
          As[0][r * AK + c] = vals[t];
        }
      }
    }
#pragma unroll
    for (int e = 0; e < NBV; ++e) {
      const int idx = tid + 256 * e;
      if (idx < BKC * BN) {
        const int kk = idx % BKC, nn = idx / BKC;
        const V2 v = rbv[e];
        if constexpr (SPLIT) {
          __bf16 hx = to_bf16(v.x), hy = to_bf16(v.y);
          Bs[0][nn * BKS + 2 * kk] = hx;
          Bs[0][nn * BKS + 2 * kk + 1] = hy;
          Bs[1][nn * BKS + 2 * kk] = to_bf16(v.x - (float)hx);
          Bs[1][nn * BKS + 2 * kk + 1] = to_bf16(v.y - (float)hy);
        } else {
          Bs[0][nn * BKS + 2 * kk] = v.x;
          Bs[0][nn * BKS + 2 * kk + 1] = v.y;
        }
      }
    }
  };
  if (kbeg < kend) gload(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += BKC) {
    lstore();
    __syncthreads();
    if (k0 + BKC < kend) gload(k0 + BKC);
    const int li = lane & 15, lk = lane >> 4;
    if constexpr (SPLIT) {
      // one 16x16x32 step covers the 32 real k of the stage; lane holds k = lk*8 .. lk*8+7
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) {
        const int row = w * 32 + rb * 16 + li;
        bf8v ahi = *reinterpret_cast<const bf8v*>(&As[0][row * AK + lk * 8]);
        bf8v alo = *reinterpret_cast<const bf8v*>(&As[1][row * AK + lk * 8]);
#pragma unroll
        for (int cb = 0; cb < NB16; ++cb) {
          const int col = cb * 16 + li;
          bf8v bhi = *reinterpret_cast<const bf8v*>(&Bs[0][col * BKS + lk * 8]);
          bf8v blo = *reinterpret_cast<const bf8v*>(&Bs[1][col * BKS + lk * 8]);
          acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, bhi, acc[rb][cb], 0, 0, 0);
          acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, blo, acc[rb][cb], 0, 0, 0);
          acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, bhi, acc[rb][cb], 0, 0, 0);
        }
      }
    } else if constexpr (MODE == MODE_F32) {
      // 8 MFMA k-steps; in step ks lane group lk supplies real k = lk*8 + ks (the sum over the
      // stage's 32 k values is split differently from the natural order, A and B alike), so
      // each lane reads its 8 k values of a B column with two 16-byte LDS reads
      f4v b[NB16][2];
#pragma unroll
      for (int cb = 0; cb < NB16; ++cb) {
        const float* p = &Bs[0][(cb * 16 + li) * BKS + lk * 8];
        b[cb][0] = *reinterpret_cast<const f4v*>(p);
        b[cb][1] = *reinterpret_cast<const f4v*>(p + 4);
      }
      float a[2][8];
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int ks = 0; ks < 8; ++ks) a[rb][ks] = As[0][(w * 32 + rb * 16 + li) * AK + lk * 8 + ks];
#pragma unroll
      for (int ks = 0; ks < 8; ++ks) {
#pragma unroll
        for (int cb = 0; cb < NB16; ++cb)
#pragma unroll
          for (int rb = 0; rb < 2; ++rb)
            acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[rb][ks], b[cb][ks >> 2][ks & 3], acc[rb][cb], 0, 0, 0);
      }
    } else {
#pragma unroll
      for (int ks = 0; ks < BKR / 4; ++ks) {
        const int kk = ks * 4 + lk;
        T a[2];
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) a[rb] = As[0][(w * 32 + rb * 16 + li) * AK + kk];
#pragma unroll
        for (int cb = 0; cb < NB16; ++cb) {
          T b = Bs[0][(cb * 16 + li) * BKS + kk];
#pragma unroll
          for (int rb = 0; rb < 2; ++rb) acc[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rb], b, acc[rb][cb], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  const int li = lane & 15, lq = lane >> 4;
  if constexpr (MODE == MODE_F64) {
    // f64 16x16x4 result layout (measured on gfx950): lane holds real rows lane/16 + 4q,
    // column lane%16, so a lane holds only real or only imaginary rows; the imaginary
    // partner of lane l is lane l+16.
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int cb = 0; cb < NB16; ++cb) {
        double partner[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) partner[q] = __shfl_down(acc[rb][cb][q], 16, 64);
        const int gn = n0 + cb * 16 + li;
        if ((lq & 1) != 0 || gn >= N) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int gm = m0 + (w * 32 + rb * 16 + lq + 4 * q) / 2;
          if (gm >= M) continue;
          V2 o;
          o.x = acc[rb][cb][q];
          o.y = partner[q];
          if (partial) {
            partial[((size_t)zs * N + gn) * M + gm] = o;
          } else {
            V2* c = &C[(size_t)gn * ldc + gm];
            if (accumulate) {
              V2 old = *c;
              o.x += old.x;
              o.y += old.y;
            }
            *c = o;
          }
        }
      }
    return;
  }
  // ---- store: lane holds real rows 4*(lane/16)+q of each 16×16 block, column lane%16
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int cb = 0; cb < NB16; ++cb) {
      const int gn = n0 + cb * 16 + li;
      if (gn >= N) continue;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int gm = m0 + (w * 32 + rb * 16) / 2 + 2 * lq + h;
        if (gm >= M) continue;
        V2 o;
        o.x = acc[rb][cb][2 * h];
        o.y = acc[rb][cb][2 * h + 1];
        if (partial) {
          partial[((size_t)zs * N + gn) * M + gm] = o;
        } else {
          V2* c = &C[(size_t)gn * ldc + gm];
          if (accumulate) {
            V2 old = *c;
            o.x += old.x;
            o.y += old.y;
          }
          *c = o;
        }
      }
    }
}

template <typename V2>
__global__ void __launch_bounds__(256) k_splitk_sum(int M, int N, int nsplit, const V2* __restrict__ partial,
                                                    V2* __restrict__ C, int ldc, long long sC, int accumulate) {
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)M * N) return;
  partial += (size_t)blockIdx.y * nsplit * M * N;
  C += (size_t)blockIdx.y * sC;
  int gm = (int)(idx % M), gn = (int)(idx / M);
  V2 s;
  s.x = 0;
  s.y = 0;
  for (int z = 0; z < nsplit; ++z) {
    V2 p = partial[(size_t)z * M * N + idx];
    s.x += p.x;
    s.y += p.y;
  }
  V2* c = &C[(size_t)gn * ldc + gm];
  if (accumulate) {
    V2 old = *c;
    s.x += old.x;
    s.y += old.y;
  }
  *c = s;
}

struct Batch {
  int count = 1;
  long long sA = 0, sB = 0, sC = 0;  // element strides between consecutive products
};

template <int MODE, int NB16>
void launch_nb(Context& c, dim3 grid, int M, int N, int K, int kchunk, int nsplit, const Batch& bt, const void* A,
               int lda, const void* B, int ldb, void* C, int ldc, int acc, void* part) {
  using V2 = typename Traits<MODE>::V2;
  k_zgemm_mfma<MODE, NB16><<<grid, 256, 0, c.stream>>>(M, N, K, kchunk, nsplit, bt.sA, bt.sB, bt.sC, (const V2*)A,
                                                        lda, (const V2*)B, ldb, (V2*)C, ldc, acc, (V2*)part);
}

template <int MODE>
void zgemm_mfma_t(Context& c, int M, int N, int K, const void* A, int lda, const void* B, int ldb, void* C, int ldc,
                  bool accumulate, const Batch& bt) {
  using V2 = typename Traits<MODE>::V2;
  const int cus = c.cu_count > 0 ? c.cu_count : 256;
  int nb16 = std::min(9, (N + 15) / 16);
  int mt = (M + BMC - 1) / BMC;
  // fewer tiles than CUs with wide column blocks: halve the block width so that every CU gets
  // work (the A tile is then read by two column blocks)
  if ((long)mt * ((N + 16 * nb16 - 1) / (16 * nb16)) * bt.count < cus && nb16 > 4 && (long)mt * bt.count >= cus / 4)
    nb16 = (nb16 + 1) / 2;
  int BN = 16 * nb16;
  int nt = (N + BN - 1) / BN;
  long tiles = (long)mt * nt * bt.count;
  int target = 2 * cus;
  int nsplit = 1;
  // split K only for K >= 256 (short reductions are cheaper than the extra partial-sum pass) and
  // when the tiles fill less than half the CUs (above that the partial-sum round trip costs
  // more than the idle CUs), then into chunks of at least 4 stages
  if (tiles < target / 4 && K >= 16 * BKC)
    nsplit = (int)std::min<long>((target + tiles - 1) / tiles, std::max(1, K / (4 * BKC)));
  int kchunk = (K + nsplit - 1) / nsplit;
  kchunk = ((kchunk + BKC - 1) / BKC) * BKC;
  nsplit = (K + kchunk - 1) / kchunk;
  PWK_REQUIRE((long)nsplit * bt.count <= 65535, PWK_E_UNSUPPORTED, "GEMM batch too large");
  PWK_REQUIRE((long)mt * nt * nsplit * bt.count <= 0x7fffffffL, PWK_E_UNSUPPORTED, "GEMM grid too large");
  dim3 grid((unsigned)((long)mt * nt * nsplit * bt.count));
  void* part = nullptr;
  if (nsplit > 1) {
    c.gemm_ws.ensure((size_t)bt.count * nsplit * M * N * sizeof(V2));
    part = c.gemm_ws.p;
  }
  int acc = (nsplit == 1 && accumulate) ? 1 : 0;
  switch (nb16) {
#define PWK_NB(X) \
  case X: launch_nb<MODE, X>(c, grid, M, N, K, kchunk, nsplit, bt, A, lda, B, ldb, C, ldc, acc, part); break;
    PWK_NB(1) PWK_NB(2) PWK_NB(3) PWK_NB(4) PWK_NB(5) PWK_NB(6) PWK_NB(7) PWK_NB(8)
#undef PWK_NB
    default: launch_nb<MODE, 9>(c, grid, M, N, K, kchunk, nsplit, bt, A, lda, B, ldb, C, ldc, acc, part); break;
  }
  PWK_HIP(hipGetLastError());
  if (nsplit > 1) {
    dim3 g(blocks_for((size_t)M * N, 256), (unsigned)bt.count);
    k_splitk_sum<V2><<<g, 256, 0, c.stream>>>(M, N, nsplit, (const V2*)part, (V2*)C, ldc, bt.sC,
                                              accumulate ? 1 : 0);
    PWK_HIP(hipGetLastError());
  }
}

}  // namespace

void launch_zgemm_mfma(Context& c, int M, int N, int K, const void* A, int lda, const void* B, int ldb, void* C,
                       int ldc, bool accumulate, int precision) {
  launch_zgemm_mfma_batched(c, M, N, K, A, lda, 0, B, ldb, 0, C, ldc, 0, 1, accumulate, precision);
}

void launch_zgemm_mfma_batched(Context& c, int M, int N, int K, const void* A, int lda, long long sA, const void* B,
                               int ldb, long long sB, void* C, int ldc, long long sC, int batch, bool accumulate,
                               int precision) {
  if (M <= 0 || N <= 0 || batch <= 0) return;
  if (K <= 0) {
    PWK_REQUIRE(accumulate, PWK_E_INVALID, "zgemm with K = 0 and no accumulation");
    return;
  }
  Batch bt;
  bt.count = batch;
  bt.sA = sA;
  bt.sB = sB;
  bt.sC = sC;
  if (precision == PWK_FP64)
    zgemm_mfma_t<MODE_F64>(c, M, N, K, A, lda, B, ldb, C, ldc, accumulate, bt);
  else if (precision == PWK_SPLIT_BF16)
    zgemm_mfma_t<MODE_BF16X3>(c, M, N, K, A, lda, B, ldb, C, ldc, accumulate, bt);
  else
    zgemm_mfma_t<MODE_F32>(c, M, N, K, A, lda, B, ldb, C, ldc, accumulate, bt);
}

}  // namespace pwk
