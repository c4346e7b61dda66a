// Local part of H·ψ and the band density on the sphere-pruned FFT.
//
// The plane-wave sphere occupies S z-sticks (distinct (n1, n2)); the
// reference enumerates the basis n1 -> n2 -> n3 (src/basis.rs:47-59), so the
// coefficients of one stick are contiguous. One H·ψ is three kernels:
//   k_zinv   : per (band, stick) scatter ψ along z, inverse FFT along z
//              -> stick buffer T[band][i3][stick]            (S·nz values)
//   k_plane  : per (band, z-plane) in LDS: y-inverse on the rows carrying
//              sticks, x-inverse, × V(r), x-forward, y-forward -> T
//   k_zfwd   : per (band, stick) forward FFT along z, gather the sphere,
//              Hψ = T_kin ψ + (1/N)·(...)                  (src/scf/potentials.rs:163-190)
// which equals the dense fill_hamiltonian_with_v_eff product because the
// wrapped Miller difference there is a circular convolution.
// The density reuses k_zinv and a plane kernel that accumulates f·|ψ(r)|²
// (src/scf/density.rs:59-97).
#include <algorithm>
#include <map>

#include "common.hpp"
#include "fft_reg.hpp"
#include "fft_sizes.hpp"

namespace pwk {

namespace {

template <typename T>
struct V2t;
template <>
struct V2t<double> {
  using t = double2;
};
template <>
struct V2t<float> {
  using t = float2;
};

// index of n3 ≡ i3 (mod nz) inside a stick starting at n3 = lo, or >= cnt if absent
__device__ __forceinline__ int stick_slot(int i3, int lo, int nz) {
  int j = i3 - lo;
  if (j < 0) j += nz;
  if (j >= nz) j -= nz;
  return j;
}

template <typename T, int NZ>
__global__ void __launch_bounds__(64) k_zinv(const typename V2t<T>::t* __restrict__ psi,
                                             typename V2t<T>::t* __restrict__ tbuf, int npw, int S,
                                             const int* __restrict__ st_pw0, const int* __restrict__ st_cnt,
                                             const int* __restrict__ st_lo) {
  int s = blockIdx.x * 64 + threadIdx.x;
  int b = blockIdx.y;
  if (s >= S) return;
  int pw0 = st_pw0[s], cnt = st_cnt[s], lo = st_lo[s];
  const typename V2t<T>::t* src = psi + (size_t)b * npw + pw0;
  T re[NZ], im[NZ];
#pragma unroll
  for (int i3 = 0; i3 < NZ; ++i3) {
    int j = stick_slot(i3, lo, NZ);
    typename V2t<T>::t v;
    v.x = 0;
    v.y = 0;
    if (j < cnt) v = src[j];
    re[i3] = v.x;
    im[i3] = v.y;
  }
  fft_dif<T, NZ, 1>(re, im);
  typename V2t<T>::t* dst = tbuf + (size_t)b * NZ * S + s;
#pragma unroll
  for (int f = 0; f < NZ; ++f) {
    typename V2t<T>::t v;
    v.x = re[FftOrder<NZ>::pos(f)];
    v.y = im[FftOrder<NZ>::pos(f)];
    dst[(size_t)f * S] = v;
  }
}

template <typename T, int NZ>
__global__ void __launch_bounds__(64) k_zfwd(const typename V2t<T>::t* __restrict__ tbuf,
                                             const typename V2t<T>::t* __restrict__ psi,
                                             typename V2t<T>::t* __restrict__ hpsi, const T* __restrict__ kin,
                                             int npw, int S, const int* __restrict__ st_pw0,
                                             const int* __restrict__ st_cnt, const int* __restrict__ st_lo,
                                             T inv_n, int nbk) {
  int s = blockIdx.x * 64 + threadIdx.x;
  int b = blockIdx.y;
  if (s >= S) return;
  int pw0 = st_pw0[s], cnt = st_cnt[s], lo = st_lo[s];
  const typename V2t<T>::t* src = tbuf + (size_t)b * NZ * S + s;
  T re[NZ], im[NZ];
#pragma unroll
  for (int i3 = 0; i3 < NZ; ++i3) {
    typename V2t<T>::t v = src[(size_t)i3 * S];
    re[i3] = v.x;
    im[i3] = v.y;
  }
  fft_dif<T, NZ, -1>(re, im);
  const typename V2t<T>::t* p = psi + (size_t)b * npw + pw0;
  typename V2t<T>::t* h = hpsi + (size_t)b * npw + pw0;
  const T* t = kin + (size_t)(b / nbk) * npw + pw0;  // bands of k-point b / nbk
#pragma unroll
  for (int i3 = 0; i3 < NZ; ++i3) {
    int j = stick_slot(i3, lo, NZ);
    if (j < cnt) {
      typename V2t<T>::t ps = p[j];
      T tk = t[j];
      typename V2t<T>::t o;
      o.x = tk * ps.x + re[FftOrder<NZ>::pos(i3)] * inv_n;
      o.y = tk * ps.y + im[FftOrder<NZ>::pos(i3)] * inv_n;
      h[j] = o;
    }
  }
}

// LDS plane P[NX][NY+1] (one padding column against bank conflicts on row access)
template <typename T, int NX, int NY>
struct Plane {
  static constexpr int LD = NY + 1;
  static constexpr size_t bytes = (size_t)NX * LD * sizeof(typename V2t<T>::t);
};

// the y-lines of the plane's rows that carry sticks: row x1list[r] of P, or (x1list null) row r of
// a plane that holds only those rows
template <typename T, int NY, int SIGN>
__device__ __forceinline__ void plane_rows(typename V2t<T>::t* P, int nx1, const int* __restrict__ x1list) {
  constexpr int LD = NY + 1;
  for (int r = threadIdx.x; r < nx1; r += blockDim.x) {
    typename V2t<T>::t* row = P + (x1list ? x1list[r] : r) * LD;
    T re[NY], im[NY];
#pragma unroll
    for (int j = 0; j < NY; ++j) {
      re[j] = row[j].x;
      im[j] = row[j].y;
    }
    fft_dif<T, NY, SIGN>(re, im);
#pragma unroll
    for (int f = 0; f < NY; ++f) {
      row[f].x = re[FftOrder<NY>::pos(f)];
      row[f].y = im[FftOrder<NY>::pos(f)];
    }
  }
}

// H·ψ plane pass: one (z-plane, band) per workgroup. COMPACT: the rows of the plane that carry
// sticks are the two runs i1 < ra and i1 >= NX - rb (a sphere's columns), and only those
// nx1 = ra + rb rows are held in LDS, plus a zero row and a discard row. The other rows are zero
// through the y passes and their x-pass output is never gathered, so the x-lines read them from the
// zero row and write them to the discard row. Same arithmetic as the whole plane; about half the LDS
// for a sphere that spans half the axis, so more planes in flight per CU.
template <typename T, int NX, int NY, bool COMPACT>
__global__ void __launch_bounds__(64) k_plane_hpsi(typename V2t<T>::t* __restrict__ tbuf, const T* __restrict__ vr,
                                                   int S, int nz, const int* __restrict__ st_xy,
                                                   const int* __restrict__ st_x1, int nx1,
                                                   const int* __restrict__ x1list, int ra, int rb) {
  extern __shared__ unsigned char smem_raw[];
  using V2 = typename V2t<T>::t;
  constexpr int LD = NY + 1;
  V2* P = reinterpret_cast<V2*>(smem_raw);
  const int i3 = blockIdx.x, b = blockIdx.y;
  V2* tp = tbuf + ((size_t)b * nz + i3) * S;
  const int rows = COMPACT ? nx1 + 2 : NX;
  for (int i = threadIdx.x; i < rows * LD; i += blockDim.x) {
    P[i].x = 0;
    P[i].y = 0;
  }
  __syncthreads();
  for (int s = threadIdx.x; s < S; s += blockDim.x) {
    const int xy = st_xy[s];
    P[(COMPACT ? st_x1[s] : xy / NY) * LD + (xy % NY)] = tp[s];
  }
  __syncthreads();
  plane_rows<T, NY, 1>(P, nx1, COMPACT ? nullptr : x1list);
  __syncthreads();
  const T* v = vr + (size_t)i3 * NX * NY;
  // held row of plane row j (COMPACT); a row without sticks reads row nx1 (kept zero) and writes
  // row nx1 + 1 (discarded), so the loads and stores need no branch
  auto row = [&](int j, int absent) { return j < ra ? j : (j >= NX - rb ? j - (NX - rb) + ra : absent); };
  for (int i2 = threadIdx.x; i2 < NY; i2 += blockDim.x) {
    T re[NX], im[NX];
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      const int r = COMPACT ? row(j, nx1) : j;
      re[j] = P[r * LD + i2].x;
      im[j] = P[r * LD + i2].y;
    }
    fft_dif<T, NX, 1>(re, im);
    T r2[NX], m2[NX];
#pragma unroll
    for (int f = 0; f < NX; ++f) {
      T vv = v[f * NY + i2];
      r2[f] = re[FftOrder<NX>::pos(f)] * vv;
      m2[f] = im[FftOrder<NX>::pos(f)] * vv;
    }
    fft_dif<T, NX, -1>(r2, m2);
#pragma unroll
    for (int f = 0; f < NX; ++f) {
      const int r = COMPACT ? row(f, nx1 + 1) : f;
      P[r * LD + i2].x = r2[FftOrder<NX>::pos(f)];
      P[r * LD + i2].y = m2[FftOrder<NX>::pos(f)];
    }
  }
  __syncthreads();
  plane_rows<T, NY, -1>(P, nx1, COMPACT ? nullptr : x1list);
  __syncthreads();
  for (int s = threadIdx.x; s < S; s += blockDim.x) {
    const int xy = st_xy[s];
    tp[s] = P[(COMPACT ? st_x1[s] : xy / NY) * LD + (xy % NY)];
  }
}

// Density plane pass: one block per z-plane, loops over bands, accumulates
// Σ_b fw_b |ψ_b(r)|² in an FP64 LDS plane, then adds it to rho (row-major grid).
template <typename T, int NX, int NY>
__global__ void __launch_bounds__(64) k_plane_density(const typename V2t<T>::t* __restrict__ tbuf, int nb,
                                                      const double* __restrict__ fw, double* __restrict__ rho,
                                                      int S, int nz, const int* __restrict__ st_xy, int nx1,
                                                      const int* __restrict__ x1list) {
  extern __shared__ unsigned char smem_raw[];
  using V2 = typename V2t<T>::t;
  constexpr int LD = NY + 1;
  V2* P = reinterpret_cast<V2*>(smem_raw);
  double* acc = reinterpret_cast<double*>(smem_raw + Plane<T, NX, NY>::bytes);
  const int i3 = blockIdx.x;
  for (int i = threadIdx.x; i < NX * NY; i += blockDim.x) acc[i] = 0.0;
  for (int b = 0; b < nb; ++b) {
    double f = fw[b];
    if (f < 1e-15) continue;  // src/scf/density.rs:14,72
    const V2* tp = tbuf + ((size_t)b * nz + i3) * S;
    __syncthreads();
    for (int i = threadIdx.x; i < NX * LD; i += blockDim.x) {
      P[i].x = 0;
      P[i].y = 0;
    }
    __syncthreads();
    for (int s = threadIdx.x; s < S; s += blockDim.x) {
      int xy = st_xy[s];
      P[(xy / NY) * LD + (xy % NY)] = tp[s];
    }
    __syncthreads();
    plane_rows<T, NY, 1>(P, nx1, x1list);
    __syncthreads();
    for (int i2 = threadIdx.x; i2 < NY; i2 += blockDim.x) {
      T re[NX], im[NX];
#pragma unroll
      for (int j = 0; j < NX; ++j) {
        re[j] = P[j * LD + i2].x;
        im[j] = P[j * LD + i2].y;
      }
      fft_dif<T, NX, 1>(re, im);
#pragma unroll
      for (int q = 0; q < NX; ++q) {
        double x = (double)re[FftOrder<NX>::pos(q)], y = (double)im[FftOrder<NX>::pos(q)];
        acc[q * NY + i2] += f * (x * x + y * y);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < NX * NY; i += blockDim.x) {
    int i1 = i / NY, i2 = i % NY;
    rho[((size_t)i1 * NY + i2) * nz + i3] += acc[i];
  }
}

// V(r) = Re IFFT_unnorm(V_eff(G)) re-ordered to plane-major [i3][i1][i2] (stick path), and in
// grid order (fused small-grid path)
__global__ void __launch_bounds__(256) k_vr_reorder(const double2* __restrict__ vg_r, double* __restrict__ v64,
                                                    float* __restrict__ v32, double* __restrict__ n64,
                                                    float* __restrict__ n32, int nx, int ny, int nz) {
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t n = (size_t)nx * ny * nz;
  if (idx >= n) return;
  // idx in plane-major order
  int i2 = (int)(idx % ny);
  int i1 = (int)((idx / ny) % nx);
  int i3 = (int)(idx / ((size_t)nx * ny));
  const size_t g = ((size_t)i1 * ny + i2) * nz + i3;
  double v = vg_r[g].x;
  v64[idx] = v;
  v32[idx] = (float)v;
  n64[g] = v;
  n32[g] = (float)v;
}

// ---- full-grid path (grids with nx != ny, or an axis beyond the register lengths) --------
// the stick path's plane and z kernels hold whole lines in registers: square planes, every axis
// one of the register lengths
bool stick_path_ok(const Context& c) {
  return c.dims[0] == c.dims[1] && fft_len_compiled(c.dims[0]) && fft_len_compiled(c.dims[2]);
}

// ψ scattered on the whole grid of each band of a chunk, batched 3D FFTs (one compiled
// line length per axis), V(r) multiply and gather. Slower than the stick path but valid
// for every grid shape the line kernels compile.
template <typename T>
__global__ void __launch_bounds__(256) k_grid_zero(size_t n, typename V2t<T>::t* __restrict__ grid) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  typename V2t<T>::t z;
  z.x = 0;
  z.y = 0;
  grid[i] = z;
}

template <typename T>
__global__ void __launch_bounds__(256) k_grid_put(size_t n, size_t N, int npw, const typename V2t<T>::t* psi,
                                                  const int* __restrict__ g2fft, typename V2t<T>::t* __restrict__ grid) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;  // band · npw + g
  if (i >= n) return;
  const size_t b = i / npw, g = i % npw;
  grid[b * N + g2fft[g]] = psi[i];
}

template <typename T>
__global__ void __launch_bounds__(256) k_grid_mulv(size_t n, size_t N, const T* __restrict__ v,
                                                   typename V2t<T>::t* __restrict__ grid) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const T vv = v[i % N];
  grid[i].x *= vv;
  grid[i].y *= vv;
}

template <typename T>
__global__ void __launch_bounds__(256) k_grid_get(size_t n, size_t N, int npw, int nbk, int b0,
                                                  const typename V2t<T>::t* __restrict__ grid,
                                                  const int* __restrict__ g2fft,
                                                  const typename V2t<T>::t* __restrict__ psi, const T* __restrict__ kin,
                                                  typename V2t<T>::t* __restrict__ hpsi, T inv_n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;  // chunk band · npw + g
  if (i >= n) return;
  const size_t bl = i / npw, g = i % npw;
  const size_t b = (size_t)b0 + bl;
  const typename V2t<T>::t w = grid[bl * N + g2fft[g]], ps = psi[b * npw + g];
  const T tk = kin[(b / nbk) * npw + g];
  typename V2t<T>::t o;
  o.x = tk * ps.x + w.x * inv_n;
  o.y = tk * ps.y + w.y * inv_n;
  hpsi[b * npw + g] = o;
}

// ρ(r) += Σ_b fw_b |ψ_b(r)|² over the bands of a chunk in band order (src/scf/density.rs:59-97)
template <typename T>
__global__ void __launch_bounds__(256) k_grid_density(size_t N, int nbc, const typename V2t<T>::t* __restrict__ grid,
                                                      const double* __restrict__ fw, double* __restrict__ rho) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  double acc = rho[i];
  for (int b = 0; b < nbc; ++b) {
    const double f = fw[b];
    if (f < 1e-15) continue;  // src/scf/density.rs:14,72
    const typename V2t<T>::t v = grid[(size_t)b * N + i];
    const double x = (double)v.x, y = (double)v.y;
    acc += f * (x * x + y * y);
  }
  rho[i] = acc;
}

// bands per chunk of the full-grid path: at most 256 MiB of grids
template <typename T>
int grid_chunk(const Context& c, int nb) {
  const size_t per = c.N * sizeof(typename V2t<T>::t);
  return (int)std::max<size_t>(1, std::min<size_t>(std::min<size_t>((size_t)nb, 32768), ((size_t)256 << 20) / per));
}

template <typename T>
void hpsi_fullgrid(Context& c, const T* kin, int nb, int nbk, const void* psi, void* hpsi) {
  using V2 = typename V2t<T>::t;
  const int prec = std::is_same<T, double>::value ? PWK_FP64 : PWK_FP32;
  const T* vr = std::is_same<T, double>::value ? (const T*)c.vrn64.p : (const T*)c.vrn32.p;
  const size_t N = c.N;
  const int chunk = grid_chunk<T>(c, nb);
  c.work_a.ensure((size_t)chunk * N * sizeof(V2));
  V2* grid = c.work_a.as<V2>();
  TimedScope ts(c, "hpsi_fullgrid");
  for (int b0 = 0; b0 < nb; b0 += chunk) {
    const int nbc = std::min(chunk, nb - b0);
    const size_t ng = (size_t)nbc * N, np = (size_t)nbc * c.npw;
    k_grid_zero<T><<<blocks_for(ng, 256), 256, 0, c.stream>>>(ng, grid);
    k_grid_put<T><<<blocks_for(np, 256), 256, 0, c.stream>>>(np, N, c.npw, (const V2*)psi + (size_t)b0 * c.npw,
                                                             c.g2fft.as<int>(), grid);
    PWK_HIP(hipGetLastError());
    launch_fft3d(c, grid, nbc, +1, prec);
    k_grid_mulv<T><<<blocks_for(ng, 256), 256, 0, c.stream>>>(ng, N, vr, grid);
    PWK_HIP(hipGetLastError());
    launch_fft3d(c, grid, nbc, -1, prec);
    k_grid_get<T><<<blocks_for(np, 256), 256, 0, c.stream>>>(np, N, c.npw, nbk, b0, grid, c.g2fft.as<int>(),
                                                             (const V2*)psi, kin, (V2*)hpsi,
                                                             (T)(1.0 / (double)N));
    PWK_HIP(hipGetLastError());
  }
}

template <typename T>
void density_fullgrid(Context& c, int nb, const void* psi, const double* fw_host, void* rho) {
  using V2 = typename V2t<T>::t;
  const int prec = std::is_same<T, double>::value ? PWK_FP64 : PWK_FP32;
  const size_t N = c.N;
  const int chunk = grid_chunk<T>(c, nb);
  c.work_a.ensure((size_t)chunk * N * sizeof(V2));
  c.work_b.ensure((size_t)nb * sizeof(double));
  PWK_HIP(hipMemcpyAsync(c.work_b.p, fw_host, (size_t)nb * sizeof(double), hipMemcpyHostToDevice, c.stream));
  V2* grid = c.work_a.as<V2>();
  for (int b0 = 0; b0 < nb; b0 += chunk) {
    const int nbc = std::min(chunk, nb - b0);
    const size_t ng = (size_t)nbc * N, np = (size_t)nbc * c.npw;
    k_grid_zero<T><<<blocks_for(ng, 256), 256, 0, c.stream>>>(ng, grid);
    k_grid_put<T><<<blocks_for(np, 256), 256, 0, c.stream>>>(np, N, c.npw, (const V2*)psi + (size_t)b0 * c.npw,
                                                             c.g2fft.as<int>(), grid);
    PWK_HIP(hipGetLastError());
    launch_fft3d(c, grid, nbc, +1, prec);
    k_grid_density<T><<<blocks_for(N, 256), 256, 0, c.stream>>>(N, nbc, grid, c.work_b.as<double>() + b0,
                                                                (double*)rho);
    PWK_HIP(hipGetLastError());
  }
}

// ---- dispatch ------------------------------------------------------------------
// Local part of H·ψ for nb bands; band b uses the kinetic diagonal kin + (b / nbk)·npw,
// so consecutive groups of nbk bands may belong to different k-points.
template <typename T>
void hpsi_local(Context& c, const T* kin, int nb, int nbk, const void* psi, void* hpsi) {
  using V2 = typename V2t<T>::t;
  int nx = c.dims[0], ny = c.dims[1], nz = c.dims[2];
  PWK_REQUIRE(c.have_v, PWK_E_STATE, "V_eff not set (pwk_set_veff)");
  if (!stick_path_ok(c)) {
    hpsi_fullgrid<T>(c, kin, nb, nbk, psi, hpsi);
    return;
  }
  PWK_REQUIRE(nb <= 65535, PWK_E_UNSUPPORTED, "too many bands in one H·psi launch");
  size_t tb = (size_t)nb * nz * c.nsticks * sizeof(V2);
  c.work_a.ensure(tb);
  V2* T_ = c.work_a.as<V2>();
  const T* vr = std::is_same<T, double>::value ? (const T*)c.vr64.p : (const T*)c.vr32.p;
  dim3 gz(blocks_for((size_t)c.nsticks, 64), (unsigned)nb);
  {
    TimedScope ts(c, "hpsi_zinv");
    switch (nz) {
#define PWK_CASE(L)                                                                                     \
  case L:                                                                                               \
    k_zinv<T, L><<<gz, 64, 0, c.stream>>>((const V2*)psi, T_, c.npw, c.nsticks, c.st_pw0.as<int>(),     \
                                          c.st_cnt.as<int>(), c.st_n3lo.as<int>());                    \
    break;
      PWK_FOR_EACH_FFT_SIZE(PWK_CASE)
#undef PWK_CASE
      default:
        throw Error(PWK_E_UNSUPPORTED, "nz has no compiled kernel");
    }
    PWK_HIP(hipGetLastError());
  }
  dim3 gp((unsigned)nz, (unsigned)nb);
  {
  TimedScope ts_plane(c, "hpsi_plane");
  switch (nx) {
#define PWK_CASE(L)                                                                                      \
  case L: {                                                                                              \
    if (c.x1_runs) {                                                                                     \
      size_t sh = (size_t)(c.nx1 + 2) * (L + 1) * sizeof(V2);                                            \
      ensure_lds_limit((const void*)k_plane_hpsi<T, L, L, true>, sh);                                   \
      k_plane_hpsi<T, L, L, true><<<gp, 64, sh, c.stream>>>(T_, vr, c.nsticks, nz, c.st_xy.as<int>(),    \
                                                            c.st_x1.as<int>(), c.nx1, c.x1list.as<int>(), \
                                                            c.x1_ra, c.x1_rb);                           \
    } else {                                                                                             \
      size_t sh = Plane<T, L, L>::bytes;                                                                 \
      ensure_lds_limit((const void*)k_plane_hpsi<T, L, L, false>, sh);                                  \
      k_plane_hpsi<T, L, L, false><<<gp, 64, sh, c.stream>>>(T_, vr, c.nsticks, nz, c.st_xy.as<int>(),   \
                                                             c.st_x1.as<int>(), c.nx1,                  \
                                                             c.x1list.as<int>(), 0, 0);                 \
    }                                                                                                    \
    break;                                                                                               \
  }
    PWK_FOR_EACH_FFT_SIZE(PWK_CASE)
#undef PWK_CASE
    default:
      throw Error(PWK_E_UNSUPPORTED, "nx has no compiled kernel");
  }
  PWK_HIP(hipGetLastError());
  }  // the plane scope ends here, before the z-forward launch
  T inv_n = (T)(1.0 / (double)c.N);
  TimedScope ts_zfwd(c, "hpsi_zfwd");
  switch (nz) {
#define PWK_CASE(L)                                                                                     \
  case L:                                                                                               \
    k_zfwd<T, L><<<gz, 64, 0, c.stream>>>(T_, (const V2*)psi, (V2*)hpsi, kin, c.npw, c.nsticks,         \
                                          c.st_pw0.as<int>(), c.st_cnt.as<int>(), c.st_n3lo.as<int>(),  \
                                          inv_n, nbk);                                                  \
    break;
    PWK_FOR_EACH_FFT_SIZE(PWK_CASE)
#undef PWK_CASE
    default:
      throw Error(PWK_E_UNSUPPORTED, "nz has no compiled kernel");
  }
  PWK_HIP(hipGetLastError());
}

template <typename T>
void density_t(Context& c, int nb, const void* psi, const double* fw_host, void* rho) {
  using V2 = typename V2t<T>::t;
  int nx = c.dims[0], nz = c.dims[2];
  if (!stick_path_ok(c)) {
    density_fullgrid<T>(c, nb, psi, fw_host, rho);
    return;
  }
  size_t tb = (size_t)nb * nz * c.nsticks * sizeof(V2);
  c.work_a.ensure(tb);
  c.work_b.ensure((size_t)nb * sizeof(double));
  PWK_HIP(hipMemcpyAsync(c.work_b.p, fw_host, (size_t)nb * sizeof(double), hipMemcpyHostToDevice, c.stream));
  V2* T_ = c.work_a.as<V2>();
  dim3 gz(blocks_for((size_t)c.nsticks, 64), (unsigned)nb);
  switch (nz) {
#define PWK_CASE(L)                                                                                     \
  case L:                                                                                               \
    k_zinv<T, L><<<gz, 64, 0, c.stream>>>((const V2*)psi, T_, c.npw, c.nsticks, c.st_pw0.as<int>(),     \
                                          c.st_cnt.as<int>(), c.st_n3lo.as<int>());                    \
    break;
    PWK_FOR_EACH_FFT_SIZE(PWK_CASE)
#undef PWK_CASE
    default:
      throw Error(PWK_E_UNSUPPORTED, "nz has no compiled kernel");
  }
  PWK_HIP(hipGetLastError());
  switch (nx) {
#define PWK_CASE(L)                                                                                      \
  case L: {                                                                                              \
    size_t sh = Plane<T, L, L>::bytes + (size_t)L * L * sizeof(double);                                  \
    ensure_lds_limit((const void*)k_plane_density<T, L, L>, sh);                                        \
    k_plane_density<T, L, L><<<(unsigned)nz, 64, sh, c.stream>>>(T_, nb, c.work_b.as<double>(),          \
                                                                 (double*)rho, c.nsticks, nz,            \
                                                                 c.st_xy.as<int>(), c.nx1,              \
                                                                 c.x1list.as<int>());                   \
    break;                                                                                               \
  }
    PWK_FOR_EACH_FFT_SIZE(PWK_CASE)
#undef PWK_CASE
    default:
      throw Error(PWK_E_UNSUPPORTED, "nx has no compiled kernel");
  }
  PWK_HIP(hipGetLastError());
}

inline int wrap(int n, int d) { return ((n % d) + d) % d; }

}  // namespace

void setup_basis(Context& c, int npw, const int32_t* m) {
  c.check_grid();
  PWK_REQUIRE(npw > 0 && m != nullptr, PWK_E_INVALID, "empty basis");
  const int nx = c.dims[0], ny = c.dims[1], nz = c.dims[2];
  std::vector<int32_t> g2fft((size_t)npw), pw0, cnt, lo, xy, sx1;
  std::map<int, int> x1index;
  std::map<int, int> seen_xy;
  for (int i = 0; i < npw; ++i) {
    const int n1 = m[3 * i], n2 = m[3 * i + 1], n3 = m[3 * i + 2];
    g2fft[(size_t)i] = (int32_t)(((size_t)wrap(n1, nx) * ny + wrap(n2, ny)) * nz + wrap(n3, nz));
    bool cont = i > 0 && m[3 * i - 3] == n1 && m[3 * i - 2] == n2 && m[3 * i - 1] + 1 == n3;
    if (!cont) {
      int col = wrap(n1, nx) * ny + wrap(n2, ny);
      PWK_REQUIRE(seen_xy.find(col) == seen_xy.end(), PWK_E_UNSUPPORTED,
                  "basis is not ordered n1 -> n2 -> n3 or two sticks alias on the grid");
      seen_xy[col] = (int)pw0.size();
      pw0.push_back(i);
      cnt.push_back(0);
      lo.push_back(n3);
      xy.push_back(col);
      int i1 = wrap(n1, nx);
      if (!x1index.count(i1)) x1index[i1] = 0;
    }
    cnt.back() += 1;
    PWK_REQUIRE(cnt.back() <= nz, PWK_E_UNSUPPORTED, "a z-stick is longer than nz (basis aliases on the grid)");
  }
  std::vector<int32_t> x1list;
  for (auto& kv : x1index) {
    kv.second = (int)x1list.size();
    x1list.push_back(kv.first);
  }
  for (size_t s = 0; s < xy.size(); ++s) sx1.push_back(x1index[xy[s] / ny]);
  c.npw = npw;
  c.nsticks = (int)pw0.size();
  c.nx1 = (int)x1list.size();
  // the stick rows as two runs [0, ra) and [nx - rb, nx) (a sphere's columns): the plane pass
  // then holds only those rows
  {
    int ra = 0;
    while (ra < c.nx1 && x1list[(size_t)ra] == ra) ++ra;
    const int rb = c.nx1 - ra;
    bool runs = ra + rb == c.nx1 && c.nx1 < nx;
    for (int r = 0; r < rb && runs; ++r) runs = x1list[(size_t)(ra + r)] == nx - rb + r;
    c.x1_runs = runs;
    c.x1_ra = runs ? ra : 0;
    c.x1_rb = runs ? rb : 0;
  }
  auto up = [&](DevBuf& d, const void* h, size_t bytes) {
    d.ensure(bytes);
    PWK_HIP(hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, c.stream));
  };
  up(c.miller, m, (size_t)npw * 3 * sizeof(int32_t));
  up(c.g2fft, g2fft.data(), g2fft.size() * sizeof(int32_t));
  up(c.st_pw0, pw0.data(), pw0.size() * sizeof(int32_t));
  up(c.st_cnt, cnt.data(), cnt.size() * sizeof(int32_t));
  up(c.st_n3lo, lo.data(), lo.size() * sizeof(int32_t));
  up(c.st_xy, xy.data(), xy.size() * sizeof(int32_t));
  up(c.x1list, x1list.data(), x1list.size() * sizeof(int32_t));
  up(c.st_x1, sx1.data(), sx1.size() * sizeof(int32_t));
  c.sync();  // the host tables go out of scope here
}

void reserve_veff_slots(Context& c, int nslots) {
  c.check_grid();
  const size_t n = c.N * (size_t)std::max(1, nslots);
  c.vr64.ensure(n * sizeof(double));
  c.vr32.ensure(n * sizeof(float));
  c.vrn64.ensure(n * sizeof(double));
  c.vrn32.ensure(n * sizeof(float));
}

void set_veff_device(Context& c, const void* veff_g_dev, int slot, bool need_g) {
  c.check_grid();
  size_t n = c.N;
  PWK_REQUIRE(slot >= 0 && slot <= 1, PWK_E_INVALID, "V_eff slot must be 0 or 1");
  if (slot == 0) {
    reserve_veff_slots(c, 1);
    if (need_g) {
      c.veffg64.ensure(n * sizeof(double2));
      PWK_HIP(hipMemcpyAsync(c.veffg64.p, veff_g_dev, n * sizeof(double2), hipMemcpyDeviceToDevice, c.stream));
    }
  } else {
    PWK_REQUIRE(c.vrn64.bytes >= 2 * n * sizeof(double), PWK_E_STATE, "second V_eff slot not reserved");
  }
  c.work_b.ensure(n * sizeof(double2));
  launch_fft3d_copy(c, veff_g_dev, c.work_b.p, +1);
  const size_t o = (size_t)slot * n;
  k_vr_reorder<<<blocks_for(n, 256), 256, 0, c.stream>>>(c.work_b.as<double2>(), c.vr64.as<double>() + o,
                                                         c.vr32.as<float>() + o, c.vrn64.as<double>() + o,
                                                         c.vrn32.as<float>() + o, c.dims[0], c.dims[1], c.dims[2]);
  PWK_HIP(hipGetLastError());
  if (slot == 0) c.have_v = true;
}

void launch_hpsi(Context& c, int ik, int nb, const void* psi, void* hpsi, int precision) {
  c.check_basis();
  PWK_REQUIRE(nb > 0, PWK_E_INVALID, "nb must be positive");
  KData& kd = c.k(ik);
  PWK_REQUIRE(kd.have_kin, PWK_E_STATE, "kinetic diagonal of this k-point not set (pwk_set_kinetic)");
  if (precision != PWK_SPLIT_BF16 && c.have_v && hpsi_small_supported(c, precision, kd.nch) &&
      (kd.nch == 0 || (precision == PWK_FP64 ? kd.B64.p && kd.DBH64.p : kd.B32.p && kd.DBH32.p))) {
    const bool f64 = precision == PWK_FP64;
    launch_hpsi_small(c, nb, nb, psi, hpsi, precision, f64 ? kd.kin64.p : kd.kin32.p, kd.nch,
                      f64 ? kd.B64.p : kd.B32.p, f64 ? kd.DBH64.p : kd.DBH32.p);
    return;
  }
  if (precision == PWK_FP64) {
    hpsi_local<double>(c, kd.kin64.as<double>(), nb, nb, psi, hpsi);
  } else {
    hpsi_local<float>(c, kd.kin32.as<float>(), nb, nb, psi, hpsi);
  }
  if (kd.nch > 0) {
    // P = (D·B^H)·ψ  (nch × nb), then Hψ += B·P   (src/potential/nonlocal.rs:432-450)
    size_t es = precision == PWK_FP64 ? sizeof(double2) : sizeof(float2);
    c.proj.ensure((size_t)kd.nch * nb * es);
    const void* dbh = precision == PWK_FP64 ? kd.DBH64.p : kd.DBH32.p;
    const void* bm = precision == PWK_FP64 ? kd.B64.p : kd.B32.p;
    {
      TimedScope ts(c, "hpsi_nl_project");
      launch_zgemm_mfma(c, kd.nch, nb, c.npw, dbh, kd.nch, psi, c.npw, c.proj.p, kd.nch, false, precision);
    }
    TimedScope ts(c, "hpsi_nl_expand");
    launch_zgemm_mfma(c, c.npw, nb, kd.nch, bm, c.npw, c.proj.p, kd.nch, hpsi, c.npw, true, precision);
  }
}

// H·ψ of nb bands at each of nk k-points in one pass: psi and hpsi hold nk blocks of
// npw × nb (band-major), kin nk × npw in the precision of psi, B / D·B^H nk blocks of
// npw × nch / nch × npw. The local part runs as one launch sequence over nk·nb bands,
// the projector GEMMs as strided batches.
void launch_hpsi_multi(Context& c, int nk, int nb, const void* psi, void* hpsi, int precision, const void* kin,
                       int nch, const void* B, const void* DBH, void* proj) {
  c.check_basis();
  PWK_REQUIRE(nk > 0 && nb > 0, PWK_E_INVALID, "nk and nb must be positive");
  if (precision != PWK_SPLIT_BF16 && c.have_v && hpsi_small_supported(c, precision, nch)) {
    launch_hpsi_small(c, nk * nb, nb, psi, hpsi, precision, kin, nch, B, DBH);
    return;
  }
  const long long npw = c.npw;
  if (precision == PWK_FP64)
    hpsi_local<double>(c, (const double*)kin, nk * nb, nb, psi, hpsi);
  else
    hpsi_local<float>(c, (const float*)kin, nk * nb, nb, psi, hpsi);
  if (nch > 0) {
    {
      TimedScope ts(c, "hpsi_nl_project");
      launch_zgemm_mfma_batched(c, nch, nb, (int)npw, DBH, nch, (long long)nch * npw, psi, (int)npw, npw * nb, proj,
                                nch, (long long)nch * nb, nk, false, precision);
    }
    TimedScope ts(c, "hpsi_nl_expand");
    launch_zgemm_mfma_batched(c, (int)npw, nb, nch, B, (int)npw, npw * nch, proj, nch, (long long)nch * nb, hpsi,
                              (int)npw, npw * nb, nk, true, precision);
  }
}

namespace {

// Packs the kinetic diagonals and projector blocks of every k-point k-major into the context's
// pack buffers (device copies), once per change of the per-k data. False when the k-points do
// not share one channel count or miss data (the caller then goes k-point by k-point).
bool pack_kdata(Context& c) {
  if (c.pack_valid) return true;
  const int nk = (int)c.kd.size();
  if (nk == 0) return false;
  const int nch = c.kd[0].nch;
  for (const KData& kd : c.kd)
    if (!kd.have_kin || kd.nch != nch || (nch > 0 && !(kd.B64.p && kd.DBH64.p && kd.B32.p && kd.DBH32.p)))
      return false;
  const size_t npw = (size_t)c.npw, blk = npw * (size_t)nch;
  c.kin_pack64.ensure((size_t)nk * npw * 8);
  c.kin_pack32.ensure((size_t)nk * npw * 4);
  if (nch > 0) {
    c.B_pack64.ensure((size_t)nk * blk * 16);
    c.DBH_pack64.ensure((size_t)nk * blk * 16);
    c.B_pack32.ensure((size_t)nk * blk * 8);
    c.DBH_pack32.ensure((size_t)nk * blk * 8);
  }
  for (int k = 0; k < nk; ++k) {
    const KData& kd = c.kd[(size_t)k];
    PWK_HIP(hipMemcpyAsync(c.kin_pack64.as<double>() + (size_t)k * npw, kd.kin64.p, npw * 8, hipMemcpyDeviceToDevice,
                           c.stream));
    PWK_HIP(hipMemcpyAsync(c.kin_pack32.as<float>() + (size_t)k * npw, kd.kin32.p, npw * 4, hipMemcpyDeviceToDevice,
                           c.stream));
    if (nch == 0) continue;
    PWK_HIP(hipMemcpyAsync(c.B_pack64.as<double2>() + (size_t)k * blk, kd.B64.p, blk * 16, hipMemcpyDeviceToDevice,
                           c.stream));
    PWK_HIP(hipMemcpyAsync(c.DBH_pack64.as<double2>() + (size_t)k * blk, kd.DBH64.p, blk * 16,
                           hipMemcpyDeviceToDevice, c.stream));
    PWK_HIP(hipMemcpyAsync(c.B_pack32.as<float2>() + (size_t)k * blk, kd.B32.p, blk * 8, hipMemcpyDeviceToDevice,
                           c.stream));
    PWK_HIP(hipMemcpyAsync(c.DBH_pack32.as<float2>() + (size_t)k * blk, kd.DBH32.p, blk * 8, hipMemcpyDeviceToDevice,
                           c.stream));
  }
  c.pack_nch = nch;
  c.pack_valid = true;
  return true;
}

size_t batch_bands(Context& c, int nent, const int32_t* ik, const int32_t* nbands) {
  PWK_REQUIRE(nent >= 1 && ik && nbands, PWK_E_INVALID, "empty or null k-point batch");
  size_t total = 0;
  for (int e = 0; e < nent; ++e) {
    PWK_REQUIRE(nbands[e] >= 1, PWK_E_INVALID, "every batch entry needs at least one band");
    const KData& kd = c.k(ik[e]);
    PWK_REQUIRE(kd.have_kin, PWK_E_STATE, "kinetic diagonal of a batch k-point not set (pwk_set_kinetic)");
    total += (size_t)nbands[e];
  }
  PWK_REQUIRE(total <= (size_t)0x7fffffff, PWK_E_INVALID, "batch too large");
  return total;
}

}  // namespace

// H·ψ of a k × band batch with a band count per entry: entry e holds nbands[e] bands of k-point
// ik[e], the entries' bands packed back to back in psi and hpsi. On small cubic grids the whole
// batch is one fused launch that reads each band's k-point from a band table; elsewhere, or when
// the k-points' channel counts differ, one H·ψ per entry.
void launch_hpsi_batch(Context& c, int nent, const int32_t* ik, const int32_t* nbands, const void* psi, void* hpsi,
                       int precision) {
  c.check_basis();
  const size_t total = batch_bands(c, nent, ik, nbands);
  int nch = c.k(ik[0]).nch;
  if (precision != PWK_SPLIT_BF16 && c.have_v && hpsi_small_supported(c, precision, nch) && pack_kdata(c)) {
    nch = c.pack_nch;
    std::vector<int32_t> tab(total);
    size_t b = 0;
    for (int e = 0; e < nent; ++e)
      for (int j = 0; j < nbands[e]; ++j) tab[b++] = ik[e];
    if (tab != c.band_tab_host) {
      c.band_tab_host = tab;
      c.band_tab.ensure(total * sizeof(int32_t));
      PWK_HIP(hipMemcpyAsync(c.band_tab.p, c.band_tab_host.data(), total * sizeof(int32_t), hipMemcpyHostToDevice,
                             c.stream));
      PWK_HIP(hipStreamSynchronize(c.stream));  // the host table may change with the next batch
    }
    const bool f64 = precision == PWK_FP64;
    launch_hpsi_small(c, (int)total, 1, psi, hpsi, precision, f64 ? c.kin_pack64.p : c.kin_pack32.p, nch,
                      f64 ? c.B_pack64.p : c.B_pack32.p, f64 ? c.DBH_pack64.p : c.DBH_pack32.p, nullptr, nullptr, 0,
                      c.band_tab.as<int>());
    return;
  }
  const size_t es = precision == PWK_FP64 ? sizeof(double2) : sizeof(float2);
  size_t off = 0;
  for (int e = 0; e < nent; ++e) {
    launch_hpsi(c, ik[e], nbands[e], (const char*)psi + off * es, (char*)hpsi + off * es, precision);
    off += (size_t)nbands[e] * c.npw;
  }
}

// ρ += Σ_b fw_b |ψ_b(r)|² over every band of a k × band batch (fw per band, occupation × k
// weight, entries back to back): the density does not depend on k, so all bands go through one
// density launch.
void launch_density_batch(Context& c, int nent, const int32_t* ik, const int32_t* nbands, const void* psi,
                          const double* fw_host, void* rho, int precision) {
  c.check_basis();
  const size_t total = batch_bands(c, nent, ik, nbands);
  launch_density(c, (int)total, psi, fw_host, rho, precision);
}

void launch_density(Context& c, int nb, const void* psi, const double* fw_host, void* rho, int precision,
                    bool overwrite) {
  c.check_basis();
  PWK_REQUIRE(nb > 0, PWK_E_INVALID, "nb must be positive");
  const int dprec = precision == PWK_FP64 ? PWK_FP64 : PWK_FP32;
  if (density_small_supported(c, dprec)) {
    c.work_b.ensure((size_t)nb * sizeof(double));
    PWK_HIP(hipMemcpyAsync(c.work_b.p, fw_host, (size_t)nb * sizeof(double), hipMemcpyHostToDevice, c.stream));
    launch_density_small(c, nb, psi, c.work_b.as<double>(), rho, dprec, overwrite);
    return;
  }
  if (overwrite) PWK_HIP(hipMemsetAsync(rho, 0, c.N * sizeof(double), c.stream));
  if (precision == PWK_FP64)
    density_t<double>(c, nb, psi, fw_host, rho);
  else
    density_t<float>(c, nb, psi, fw_host, rho);
}

}  // namespace pwk
