// Block kernels of the Chebyshev-filtered subspace eigensolver, batched over
// k-points: the filter's three-term recurrence step, conjugate transposes for
// the Gram products Y^H·Y and Y^H·(HY), Hermitian symmetrisation, the Löwdin
// column scaling, per-block dot products and the seeded start block. All are
// HBM/L2-bound elementwise passes over nk blocks of npw × nb.
#include <algorithm>
#include <cstring>

#include "common.hpp"
#include "jacobi.hpp"

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

// The Hermitian matrix of each block from its lower triangle (upper = conjugate of lower, real
// diagonal), which is all faer's self_adjoint_eigen(Side::Lower) reads of the projected
// Hamiltonian (src/eigensolver/dense.rs:66-75, 231-234).
__global__ void __launch_bounds__(256) k_lower_to_full(int m, double2* __restrict__ A) {
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (size_t)m * m) return;
  A += (size_t)blockIdx.y * m * m;
  const int i = (int)(idx % m), j = (int)(idx / m);  // row i, column j
  if (i < j) {
    const double2 l = A[(size_t)i * m + j];  // element (j, i) of the lower triangle
    A[(size_t)j * m + i] = make_double2(l.x, -l.y);
  } else if (i == j) {
    A[(size_t)j * m + i].y = 0.0;
  }
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
  double2 v = start_vector_entry(g, j, k);
  if (g == lowest[(size_t)k * nb + j]) v.x += 1.0;
  V[(size_t)k * npw * nb + idx] = v;
}

// Rayleigh-Ritz of one k-point's block in one workgroup, for m = nbs <= JMAX: overlap S = YᴴY and
// projected Hs = Yᴴ(HY), X with Xᴴ S X = I (X = L^{-H} from the Cholesky factor of S while its
// columns are independent; else X = U s^{-1/2}, S = U s Uᴴ by Jacobi, s floored at floor_rel·s_max),
// H' = Xᴴ Hs X = W λ Wᴴ (Jacobi), C = X·W in ascending λ, then V = Y·C and the residual norms
// |(HY)·C_j − λ_j V_j| of the first nres columns. One launch replaces the transposes, GEMMs,
// symmetrisations, two batched eigensolves and the residual pass of the general path.
constexpr int RR_NT = 256;
constexpr int RR_ROWS = 32;  // plane-wave rows staged in LDS per chunk of the Gram products

// sized for blocks of up to MB vectors (42 KB at MB = 16, 123 KB at 32), so a workgroup of the
// common 16-vector blocks fits on a CU beside the filter workgroups of concurrent solves
template <int MB>
struct RrLds {
  JacobiLdsT<MB> J;
  double2 Hs[MB][MB + 1];
  double2 X[MB][MB + 1];
  double2 C[MB][MB + 1];
  double2 Ys[RR_ROWS][MB + 1], HYs[RR_ROWS][MB + 1];
  double red[16][MB];
};

template <int MB>
__device__ __forceinline__ void rr_small_body(int k, int npw, int m, int nres, const double2* __restrict__ Yb,
                                              const double2* __restrict__ HYb, double2* __restrict__ Vb,
                                              double* __restrict__ lam_out, double* __restrict__ res_out,
                                              double floor_rel, const int* __restrict__ skip) {
  static_assert(MB == 16 || MB == 32, "blocks of up to 16 or 32 vectors");
  if (skip && *skip) return;  // a filter pass after the residual gate passed (uniform)
  extern __shared__ unsigned char smem_raw[];
  RrLds<MB>& L = *reinterpret_cast<RrLds<MB>*>(smem_raw);
  JacobiLdsT<MB>& J = L.J;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const double2* Y = Yb + (size_t)k * npw * m;
  const double2* HY = HYb + (size_t)k * npw * m;
  double2* V = Vb + (size_t)k * npw * m;
  // S(i, j) = Σ_G conj(Y_i) Y_j and Hs(i, j) = Σ_G conj(Y_i) (HY)_j over row chunks staged in LDS
  // (coalesced loads; the next chunk's loads are issued before this chunk's products). Wave w takes
  // rows [w·RR_ROWS/4, (w+1)·RR_ROWS/4) of every chunk and each lane a BLK × BLK block of entries,
  // so a row costs 3·BLK LDS reads per lane for 8·BLK² products; the waves' sums are added in wave
  // order at the end.
  constexpr int BLK = MB / 8, LPT = RR_ROWS * MB / RR_NT, WROWS = RR_ROWS / (RR_NT / 64);
  const int bi = (lane & 7) * BLK, bj = (lane >> 3) * BLK;
  double sre[BLK][BLK], sim[BLK][BLK], hre[BLK][BLK], him[BLK][BLK];
#pragma unroll
  for (int u = 0; u < BLK; ++u)
#pragma unroll
    for (int v = 0; v < BLK; ++v) sre[u][v] = sim[u][v] = hre[u][v] = him[u][v] = 0.0;
  double2 py[LPT], ph[LPT];
  auto load_chunk = [&](int g0) {
#pragma unroll
    for (int q = 0; q < LPT; ++q) {
      const int e = t + q * RR_NT;
      const int gg = e % RR_ROWS, i = e / RR_ROWS, g = g0 + gg;
      const bool in = e < RR_ROWS * m && g < npw;
      py[q] = in ? Y[(size_t)i * npw + g] : make_double2(0.0, 0.0);
      ph[q] = in ? HY[(size_t)i * npw + g] : make_double2(0.0, 0.0);
    }
  };
  load_chunk(0);
  for (int g0 = 0; g0 < npw; g0 += RR_ROWS) {
#pragma unroll
    for (int q = 0; q < LPT; ++q) {
      const int e = t + q * RR_NT;
      if (e < RR_ROWS * m) {
        L.Ys[e % RR_ROWS][e / RR_ROWS] = py[q];
        L.HYs[e % RR_ROWS][e / RR_ROWS] = ph[q];
      }
    }
    __syncthreads();
    if (g0 + RR_ROWS < npw) load_chunk(g0 + RR_ROWS);
    const int r0 = wave * WROWS, r1 = min(npw - g0, r0 + WROWS);
    for (int gg = r0; gg < r1; ++gg) {
      double2 a[BLK], b[BLK], h[BLK];
#pragma unroll
      for (int u = 0; u < BLK; ++u) {
        a[u] = L.Ys[gg][bi + u];  // columns past m hold stale values: their entries are dropped
        b[u] = L.Ys[gg][bj + u];
        h[u] = L.HYs[gg][bj + u];
      }
#pragma unroll
      for (int u = 0; u < BLK; ++u)
#pragma unroll
        for (int v = 0; v < BLK; ++v) {
          sre[u][v] += a[u].x * b[v].x + a[u].y * b[v].y;
          sim[u][v] += a[u].x * b[v].y - a[u].y * b[v].x;
          hre[u][v] += a[u].x * h[v].x + a[u].y * h[v].y;
          him[u][v] += a[u].x * h[v].y - a[u].y * h[v].x;
        }
    }
    __syncthreads();
  }
  for (int w = 0; w < RR_NT / 64; ++w) {
    if (wave == w) {
#pragma unroll
      for (int u = 0; u < BLK; ++u)
#pragma unroll
        for (int v = 0; v < BLK; ++v) {
          const int i = bi + u, j = bj + v;
          if (i < m && j < m) {
            if (w == 0) {
              J.A[i][j] = make_double2(sre[u][v], sim[u][v]);
              L.Hs[i][j] = make_double2(hre[u][v], him[u][v]);
            } else {
              J.A[i][j].x += sre[u][v];
              J.A[i][j].y += sim[u][v];
              L.Hs[i][j].x += hre[u][v];
              L.Hs[i][j].y += him[u][v];
            }
          }
        }
    }
    __syncthreads();
  }
  // Hermitian parts (subspace.hip k_hermitize), V = I
  for (int e = t; e < m * m; e += RR_NT) {
    const int i = e % m, j = e / m;
    if (i <= j) {
      const double2 a = J.A[i][j], b = J.A[j][i];
      const double2 h = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
      J.A[i][j] = h;
      J.A[j][i] = make_double2(h.x, -h.y);
      L.Ys[i][j] = h;  // S, kept for the Löwdin form
      L.Ys[j][i] = make_double2(h.x, -h.y);
      const double2 c = L.Hs[i][j], d = L.Hs[j][i];
      const double2 g = make_double2(0.5 * (c.x + d.x), 0.5 * (c.y - d.y));
      L.Hs[i][j] = g;
      L.Hs[j][i] = make_double2(g.x, -g.y);
    }
  }
  for (int e = t; e < m * m; e += RR_NT) J.V[e % m][e / m] = make_double2(e % m == e / m ? 1.0 : 0.0, 0.0);
  __syncthreads();
  // Cholesky S = L·Lᴴ, one barrier per column, while every pivot keeps more than 1e-8 of its
  // column's weight S_jj (the columns are far from dependent; the test ignores their scales).
  // X = L^{-H} satisfies Xᴴ·S·X = I and spans the same space as the Löwdin X, so the Ritz pairs
  // are the same. L goes to C (lower triangle); the trailing update reads column j, which step j
  // does not write.
  // L^{-1} by forward substitution on all columns of I at once (B = I in J.V): step j sets
  // L^{-1}_{jc} = B_jc / L_jj and B_ic -= L_ij·L^{-1}_{jc} below; X_cj = conj(L^{-1}_{jc}). Step j needs
  // only column j of L, complete once Cholesky step j has passed its barrier, and touches J.V and
  // L.X, which the factorisation does not: it runs in the barrier interval of Cholesky step j + 1.
  auto inverse_step = [&](int j) {
    const double ild = 1.0 / L.C[j][j].x;
    for (int e = t; e < m * m; e += RR_NT) {
      const int i = e % m, cc = e / m;
      const double2 bj = J.V[j][cc];
      const double2 x = make_double2(bj.x * ild, bj.y * ild);
      if (i == j) {
        L.X[cc][j] = make_double2(x.x, -x.y);
      } else if (i > j) {
        const double2 lij = L.C[i][j], b = J.V[i][cc];
        J.V[i][cc] = make_double2(b.x - (lij.x * x.x - lij.y * x.y), b.y - (lij.x * x.y + lij.y * x.x));
      }
    }
  };
  bool chol = true;
  for (int j = 0; j <= m; ++j) {
    if (j < m) {
      const double d = J.A[j][j].x;  // the same on every thread: the branch is block-uniform
      if (!(d > 1e-8 * L.Ys[j][j].x)) {
        chol = false;
        break;
      }
      const double sd = sqrt(d), isd = 1.0 / sd, id = isd * isd;
      for (int e = t; e < m * m; e += RR_NT) {
        const int i = e % m, l = e / m;
        if (l == j && i >= j) {
          const double2 a = J.A[i][j];
          L.C[i][j] = i == j ? make_double2(sd, 0.0) : make_double2(a.x * isd, a.y * isd);
        } else if (l > j && i >= l) {
          const double2 a = J.A[i][j], b = J.A[l][j];  // A_il -= A_ij·conj(A_lj)/d
          const double2 cur = J.A[i][l];
          J.A[i][l] = make_double2(cur.x - (a.x * b.x + a.y * b.y) * id, cur.y - (a.y * b.x - a.x * b.y) * id);
        }
      }
    }
    if (j > 0) inverse_step(j - 1);
    __syncthreads();
  }
  __syncthreads();  // every thread has read the failing pivot before the Löwdin form rewrites A
  if (!chol) {  // (with the Cholesky factor, X is complete: its last step ran in the loop's final interval)
    // Löwdin: S = U s Uᴴ (Jacobi), X = U diag(1/sqrt(max(s_j, floor_rel·s_max)))
    for (int e = t; e < m * m; e += RR_NT) {
      J.A[e % m][e / m] = L.Ys[e % m][e / m];
      J.V[e % m][e / m] = make_double2(e % m == e / m ? 1.0 : 0.0, 0.0);
    }
    __syncthreads();
    jacobi_lds(m, J);
    const double smax = J.A[J.order[m - 1]][J.order[m - 1]].x;
    for (int e = t; e < m * m; e += RR_NT) {
      const int r = e % m, j = e / m;
      const double f = 1.0 / sqrt(fmax(J.A[j][j].x, floor_rel * smax));
      L.X[r][j] = make_double2(J.V[r][j].x * f, J.V[r][j].y * f);
    }
    __syncthreads();
  }
  // T = Hs·X into C, then H' = Xᴴ·T into J.A (Hermitian part), V = I
  for (int e = t; e < m * m; e += RR_NT) {
    const int i = e % m, j = e / m;
    double re = 0.0, im = 0.0;
    for (int l = 0; l < m; ++l) {
      const double2 h = L.Hs[i][l], x = L.X[l][j];
      re += h.x * x.x - h.y * x.y;
      im += h.x * x.y + h.y * x.x;
    }
    L.C[i][j] = make_double2(re, im);
  }
  __syncthreads();
  for (int e = t; e < m * m; e += RR_NT) {
    const int i = e % m, j = e / m;
    double re = 0.0, im = 0.0;
    for (int l = 0; l < m; ++l) {
      const double2 x = L.X[l][i], c = L.C[l][j];  // conj(x)·c
      re += x.x * c.x + x.y * c.y;
      im += x.x * c.y - x.y * c.x;
    }
    J.A[i][j] = make_double2(re, im);
  }
  __syncthreads();
  for (int e = t; e < m * m; e += RR_NT) {
    const int i = e % m, j = e / m;
    if (i <= j) {
      const double2 a = J.A[i][j], b = J.A[j][i];
      const double2 h = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
      J.A[i][j] = h;
      J.A[j][i] = make_double2(h.x, -h.y);
    }
    J.V[i][j] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  jacobi_lds(m, J);
  // C[:, jj] = X·W[:, order[jj]] (ascending Ritz values)
  for (int e = t; e < m * m; e += RR_NT) {
    const int i = e % m, jj = e / m, w = J.order[jj];
    double re = 0.0, im = 0.0;
    for (int l = 0; l < m; ++l) {
      const double2 x = L.X[i][l], v = J.V[l][w];
      re += x.x * v.x - x.y * v.y;
      im += x.x * v.y + x.y * v.x;
    }
    L.C[i][jj] = make_double2(re, im);
  }
  if (t < m) lam_out[(size_t)k * m + t] = J.A[J.order[t]][J.order[t]].x;
  __syncthreads();
  double acc[MB];
#pragma unroll
  for (int jj = 0; jj < MB; ++jj) acc[jj] = 0.0;
  {
    // V = Y·C and the residual weights, one row G per thread, eight columns of C at a time (each
    // row of Y and HY is read m/8 times, coalesced across the wave)
    constexpr int CB = 8;
    for (int g = t; g < npw; g += RR_NT) {
#pragma unroll
      for (int jc = 0; jc < MB; jc += CB) {
        if (jc >= m) continue;  // fully unrolled: acc stays in registers
        double vr[CB], vi[CB], rr[CB], ri[CB];
#pragma unroll
        for (int q = 0; q < CB; ++q) vr[q] = vi[q] = rr[q] = ri[q] = 0.0;
        for (int i = 0; i < m; ++i) {
          const double2 y = Y[(size_t)i * npw + g], h = HY[(size_t)i * npw + g];
#pragma unroll
          for (int q = 0; q < CB; ++q) {
            const double2 c = L.C[i][min(jc + q, m - 1)];
            vr[q] += y.x * c.x - y.y * c.y;
            vi[q] += y.x * c.y + y.y * c.x;
            rr[q] += h.x * c.x - h.y * c.y;
            ri[q] += h.x * c.y + h.y * c.x;
          }
        }
#pragma unroll
        for (int q = 0; q < CB; ++q) {
          const int jj = jc + q;
          if (jj < m) V[(size_t)jj * npw + g] = make_double2(vr[q], vi[q]);
          if (jj < nres) {
            const double l = J.A[J.order[jj]][J.order[jj]].x;
            const double dx = rr[q] - l * vr[q], dy = ri[q] - l * vi[q];
            acc[jj] += dx * dx + dy * dy;
          }
        }
      }
    }
  }
#pragma unroll
  for (int jj = 0; jj < MB; ++jj) {
    if (jj >= nres) continue;
    double v = acc[jj];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((t & 63) == 0) L.red[wave][jj] = v;
  }
  __syncthreads();
  if (t < nres) {
    double v = 0.0;
    for (int w = 0; w < RR_NT / 64; ++w) v += L.red[w][t];
    res_out[(size_t)k * m + t] = sqrt(v);
  }
}

template <int MB>
__global__ void __launch_bounds__(RR_NT) k_rr_small(int npw, int m, int nres, const double2* __restrict__ Yb,
                                                    const double2* __restrict__ HYb, double2* __restrict__ Vb,
                                                    double* __restrict__ lam_out, double* __restrict__ res_out,
                                                    double floor_rel, const int* __restrict__ skip) {
  rr_small_body<MB>((int)blockIdx.x, npw, m, nres, Yb, HYb, Vb, lam_out, res_out, floor_rel, skip);
}

// the Rayleigh-Ritz steps of several solves' passes in one launch: blockIdx.y selects the request,
// blockIdx.x its k-point block (the code of k_rr_small per block)
template <int MB>
__global__ void __launch_bounds__(RR_NT) k_rr_multi(PassRrMulti mr, int npw, int m, int nres, double floor_rel) {
  const PassRrReq& q = mr.r[blockIdx.y];
  if ((int)blockIdx.x >= q.nk) return;  // uniform
  rr_small_body<MB>((int)blockIdx.x, npw, m, nres, q.Y, q.HY, q.V, q.lam, q.res, floor_rel, q.skip);
}

}  // namespace

bool rr_small_supported(int m) { return m >= 1 && m <= JMAX; }

// Warm-start Rayleigh-Ritz on dense Hamiltonians, batched over k: diagonalize_subspace
// (src/eigensolver/dense.rs:183-260) for every block. HV = H·V_prev, H_sub = V_prev^H·HV (its
// lower triangle), H_sub = U·diag(λ)·U^H, V = V_prev·U, residuals ||H v_i − λ_i v_i||₂
// (eigenvalue_residuals, :112-143); a block whose largest residual exceeds tol gets the full
// diagonalize_lowest instead (:250-254), from an untouched copy of its H. Without V_prev every
// block takes the full solve (:205-207). Synchronises the stream: the gate is read on the host.
void subspace_rr(Context& c, int n, int nb, int nk, const void* H, const void* Vprev, double tol, double* w, void* V,
                 int* used_full, double* max_resid) {
  PWK_REQUIRE(n >= 1 && nb >= 1 && nb <= n && nk >= 1, PWK_E_INVALID, "bad subspace shape");
  PWK_REQUIRE(heev_supported(n, nb) && heev_supported(nb, nb), PWK_E_UNSUPPORTED,
              "matrix too large for the batched eigensolver");
  const long long nn = (long long)n * n, nbk = (long long)n * nb, bb = (long long)nb * nb;
  std::vector<int> full((size_t)nk, 1);
  std::vector<double> worst((size_t)nk, 0.0);
  DevBuf hcopy;
  if (Vprev) {
    TimedScope ts(c, "subspace_rr");
    DevBuf HV, VH, Hs, U, res;
    HV.ensure((size_t)nk * nbk * 16);
    VH.ensure((size_t)nk * nbk * 16);
    Hs.ensure((size_t)nk * bb * 16);
    U.ensure((size_t)nk * bb * 16);
    res.ensure((size_t)nk * nb * 8);
    launch_zgemm_mfma_batched(c, n, nb, n, H, n, nn, Vprev, n, nbk, HV.p, n, nbk, nk, false, PWK_FP64);
    launch_conj_transpose(c, n, nb, Vprev, n, nbk, VH.p, nb, nbk, nk);
    launch_zgemm_mfma_batched(c, nb, nb, n, VH.p, nb, nbk, HV.p, n, nbk, Hs.p, nb, bb, nk, false, PWK_FP64);
    k_lower_to_full<<<dim3(blocks_for((size_t)bb, 256), (unsigned)nk), 256, 0, c.stream>>>(nb, Hs.as<double2>());
    PWK_HIP(hipGetLastError());
    c.eig_work.ensure(heev_work_doubles(nb, nb, nk) * 8);
    launch_heev_batched(c, nb, nb, nk, Hs.p, c.eig_work.p, w, U.p, nullptr);
    launch_zgemm_mfma_batched(c, n, nb, nb, Vprev, n, nbk, U.p, nb, bb, V, n, nbk, nk, false, PWK_FP64);
    launch_zgemm_mfma_batched(c, n, nb, n, H, n, nn, V, n, nbk, HV.p, n, nbk, nk, false, PWK_FP64);
    launch_resid_norms(c, n, nb, nb, nk, V, HV.p, w, res.as<double>());
    std::vector<double> r((size_t)nk * nb);
    PWK_HIP(hipMemcpyAsync(r.data(), res.p, r.size() * 8, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    for (int k = 0; k < nk; ++k) {
      double m = 0.0;
      for (int b = 0; b < nb; ++b) m = std::max(m, r[(size_t)k * nb + b]);
      worst[(size_t)k] = m;
      full[(size_t)k] = m > tol ? 1 : 0;
    }
  }
  std::vector<int> redo;
  for (int k = 0; k < nk; ++k)
    if (full[(size_t)k]) redo.push_back(k);
  if (!redo.empty()) {
    // the full solve of the gated blocks, batched: their H copied together (the solver
    // overwrites its input), eigenpairs scattered back
    TimedScope ts(c, "subspace_rr_full");
    const int nf = (int)redo.size();
    DevBuf hf, wf, vf;
    hf.ensure((size_t)nf * nn * 16);
    wf.ensure((size_t)nf * nb * 8);
    vf.ensure((size_t)nf * nbk * 16);
    for (int i = 0; i < nf; ++i)
      PWK_HIP(hipMemcpyAsync(hf.as<double2>() + (size_t)i * nn, (const double2*)H + (size_t)redo[(size_t)i] * nn,
                             (size_t)nn * 16, hipMemcpyDeviceToDevice, c.stream));
    c.eig_work.ensure(heev_work_doubles(n, nb, nf) * 8);
    launch_heev_batched(c, n, nb, nf, hf.p, c.eig_work.p, wf.as<double>(), vf.p, nullptr);
    for (int i = 0; i < nf; ++i) {
      const int k = redo[(size_t)i];
      PWK_HIP(hipMemcpyAsync(w + (size_t)k * nb, wf.as<double>() + (size_t)i * nb, (size_t)nb * 8,
                             hipMemcpyDeviceToDevice, c.stream));
      PWK_HIP(hipMemcpyAsync((double2*)V + (size_t)k * nbk, vf.as<double2>() + (size_t)i * nbk, (size_t)nbk * 16,
                             hipMemcpyDeviceToDevice, c.stream));
    }
    c.sync();
  }
  if (used_full) std::memcpy(used_full, full.data(), full.size() * sizeof(int));
  if (max_resid) std::memcpy(max_resid, worst.data(), worst.size() * 8);
}

void launch_rr_small(Context& c, int npw, int m, int nres, int nk, const void* Y, const void* HY, void* V,
                     double* lam, double* res, double floor_rel, const int* skip) {
  PWK_REQUIRE(rr_small_supported(m) && nres <= m && npw >= m && nk >= 1, PWK_E_INVALID, "bad Rayleigh-Ritz shape");
  auto go = [&](auto kern, size_t sh) {
    ensure_lds_limit((const void*)kern, sh);
    kern<<<(unsigned)nk, RR_NT, sh, c.stream>>>(npw, m, nres, (const double2*)Y, (const double2*)HY, (double2*)V,
                                               lam, res, floor_rel, skip);
  };
  if (m <= 16)
    go(k_rr_small<16>, sizeof(RrLds<16>));
  else
    go(k_rr_small<32>, sizeof(RrLds<32>));
  PWK_HIP(hipGetLastError());
}

void launch_rr_multi(hipStream_t stream, const PassShape& sh, const PassRrMulti& m, int n, int max_nk,
                     double floor_rel) {
  PWK_REQUIRE(n >= 1 && n <= MAX_PASS_BATCH && rr_small_supported(sh.nbs) && sh.nb <= sh.nbs && sh.npw >= sh.nbs,
              PWK_E_INVALID, "bad batched Rayleigh-Ritz shape");
  auto go = [&](auto kern, size_t lds) {
    ensure_lds_limit((const void*)kern, lds);
    kern<<<dim3((unsigned)max_nk, (unsigned)n), RR_NT, lds, stream>>>(m, sh.npw, sh.nbs, sh.nb, floor_rel);
  };
  if (sh.nbs <= 16)
    go(k_rr_multi<16>, sizeof(RrLds<16>));
  else
    go(k_rr_multi<32>, sizeof(RrLds<32>));
  PWK_HIP(hipGetLastError());
}

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
