// Pointwise grid kernels: Hartree, V_eff assembly, PZ-LDA and LSDA XC.
// Each is a single HBM pass (bandwidth-bound); FP64 and FP32 variants. On small cubic grids the
// SCF's chains of transforms and pointwise steps also run as single-workgroup kernels with the grid
// in LDS (the transforms of k_fft3d_small, the pointwise code below).
#include "common.hpp"
#include "lds_fft.hpp"
#include "xc.hpp"

namespace pwk {

namespace {

// Physical constants as in src/consts.rs:15-49 (RHO_FLOOR src/functionals/mod.rs:5)
constexpr double kPiD = 3.141592653589793;
constexpr double kHaToEv = 27.211386245988;
constexpr double kBohr = 0.529177210903;
constexpr double kBohr3 = kBohr * kBohr * kBohr;
constexpr double kRhoFloor = 1e-30;
constexpr double kG2Zero = 1e-12;

template <typename T>
struct Cx;
template <>
struct Cx<double> {
  using t = double2;
};
template <>
struct Cx<float> {
  using t = float2;
};

// v_eff = [v_local] + [4πe² ρ(G)/|G|²] + [v_xc]  (src/scf/energy.rs:293-301, 343-358)
template <typename T>
__global__ void __launch_bounds__(256) k_hartree_veff(size_t n, const typename Cx<T>::t* __restrict__ rho,
                                                      const T* __restrict__ g2, T fourpi_e2,
                                                      const typename Cx<T>::t* __restrict__ vl,
                                                      const typename Cx<T>::t* __restrict__ vxc,
                                                      typename Cx<T>::t* __restrict__ out) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  typename Cx<T>::t acc;
  acc.x = 0;
  acc.y = 0;
  if (rho) {
    T gg = g2[i];
    if (gg > (T)kG2Zero) {
      typename Cx<T>::t r = rho[i];
      acc.x = r.x * fourpi_e2 / gg;
      acc.y = r.y * fourpi_e2 / gg;
    }
  }
  if (vl) {
    typename Cx<T>::t a = vl[i];
    acc.x = a.x + acc.x;
    acc.y = a.y + acc.y;
  }
  if (vxc) {
    typename Cx<T>::t a = vxc[i];
    acc.x += a.x;
    acc.y += a.y;
  }
  out[i] = acc;
}

// (a + b) + d, the reference's left-to-right sum vl + vh + vxc
template <typename T>
__global__ void __launch_bounds__(256) k_sum3(size_t n, const typename Cx<T>::t* __restrict__ a,
                                              const typename Cx<T>::t* __restrict__ b,
                                              const typename Cx<T>::t* __restrict__ d,
                                              typename Cx<T>::t* __restrict__ out) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  typename Cx<T>::t x = a[i], y = b[i], z = d[i], r;
  r.x = (x.x + y.x) + z.x;
  r.y = (x.y + y.y) + z.y;
  out[i] = r;
}

// ---- PZ-LDA (src/potential/xc.rs:129-141, 223-288): xc.hpp ----------------------
using xc::lda_point;

template <typename T>
__global__ void __launch_bounds__(256) k_lda(size_t n, const T* __restrict__ rho, T* __restrict__ exc,
                                             T* __restrict__ vxc) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T e, v;
  lda_point(rho[i], e, v);
  exc[i] = e;
  vxc[i] = v;
}

// ρ_xc = max(ρ + cs·ρ_core, 0) (ρ without a core, src/scf/energy.rs:314-324) written out, and
// LDA of it, in one pass (FP64)
// g (optional): g[i] ← g[i]·gs alongside, for a ρ(G) the caller already has in another scale
__global__ void __launch_bounds__(256) k_lda_core(size_t n, const double* __restrict__ rho,
                                                  const double* __restrict__ core, double cs,
                                                  double* __restrict__ rho_xc, double* __restrict__ exc,
                                                  double* __restrict__ vxc, double2* __restrict__ g, double gs) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (g) {
    const double2 x = g[i];
    g[i] = make_double2(x.x * gs, x.y * gs);
  }
  const double r = core ? fmax(rho[i] + core[i] * cs, 0.0) : rho[i];
  rho_xc[i] = r;
  double e, v;
  lda_point(r, e, v);
  exc[i] = e;
  vxc[i] = v;
}

// ---- LSDA (src/potential/xc.rs:374-392, 465-596) ------------------------------
template <typename T>
__device__ __forceinline__ void pz_rs(T rs, bool pol, T& ec, T& vc) {
  if (rs >= (T)1) {
    T gamma = pol ? (T)-0.0843 : (T)-0.1423, b1 = pol ? (T)1.3981 : (T)1.0529, b2 = pol ? (T)0.2611 : (T)0.3334;
    T srs = sqrt(rs);
    T den = fma(b2, rs, fma(b1, srs, (T)1));
    ec = gamma / den;
    T dec = -gamma * (b1 / ((T)2 * srs) + b2) / (den * den);
    vc = fma(rs / (T)3, -dec, ec);
  } else {
    T a = pol ? (T)0.01555 : (T)0.0311, b = pol ? (T)-0.0269 : (T)-0.048, c = pol ? (T)0.0007 : (T)0.0020,
      d = pol ? (T)-0.0048 : (T)-0.0116;
    T lr = log(rs);
    ec = fma(d, rs, fma(c * rs, lr, fma(a, lr, b)));
    T dec = fma(c, lr + (T)1, a / rs) + d;
    vc = fma(rs / (T)3, -dec, ec);
  }
}

template <typename T>
__device__ __forceinline__ void lsda_point(T ru, T rd, T& exc, T& vu, T& vd) {
  const T floor_ = (T)kRhoFloor;
  T rho = ru + rd;
  if (rho < floor_) {
    exc = vu = vd = (T)0;
    return;
  }
  const T pi = (T)kPiD, b3 = (T)kBohr3, ha = (T)kHaToEv;
  T rub = ru * b3, rdb = rd * b3;
  T exu = rub > floor_ ? (T)-0.75 * cbrt((T)6 * rub / pi) : (T)0;
  T exd = rdb > floor_ ? (T)-0.75 * cbrt((T)6 * rdb / pi) : (T)0;
  T rb = rho * b3;
  T ex = fma(rub, exu, rdb * exd) / rb;
  T vxu = ((T)4 / (T)3) * exu, vxd = ((T)4 / (T)3) * exd;
  T rs = cbrt((T)3 / ((T)4 * pi * rb));
  T zeta = fmin(fmax((ru - rd) / rho, (T)-1), (T)1);
  T ecu, vcu, ecp, vcp;
  pz_rs<T>(rs, false, ecu, vcu);
  pz_rs<T>(rs, true, ecp, vcp);
  T fden = fma(cbrt((T)2), (T)2, (T)-2);
  T op = fmax((T)1 + zeta, (T)0), om = fmax((T)1 - zeta, (T)0);
  T cop = cbrt(op), com = cbrt(om);
  T fz = (cop * op + com * om - (T)2) / fden;
  T dfz = ((T)4 / (T)3) * (cop - com) / fden;
  T ec = ecu + fz * (ecp - ecu);
  T vrs = vcu + fz * (vcp - vcu);
  T dedz = dfz * (ecp - ecu);
  T vcu_ = fma((T)1 - zeta, dedz, vrs);
  T vcd_ = fma((T)1 + zeta, -dedz, vrs);
  exc = ex * ha + ec * ha;
  vu = vxu * ha + vcu_ * ha;
  vd = vxd * ha + vcd_ * ha;
}

template <typename T>
__global__ void __launch_bounds__(256) k_lsda(size_t n, const T* __restrict__ ru, const T* __restrict__ rd,
                                              T* __restrict__ exc, T* __restrict__ vu, T* __restrict__ vd) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T e, u, d;
  lsda_point<T>(ru[i], rd[i], e, u, d);
  exc[i] = e;
  vu[i] = u;
  vd[i] = d;
}

__global__ void __launch_bounds__(256) k_d2f(size_t n, const double* __restrict__ s, float* __restrict__ d) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = (float)s[i];
}
__global__ void __launch_bounds__(256) k_f2d(size_t n, const float* __restrict__ s, double* __restrict__ d) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = (double)s[i];
}


// The three LDS passes of k_fft3d_small (x, y, z lines) on the workgroup's grid.
template <int N, int SIGN>
__device__ __forceinline__ void small_fft3d(double2* G) {
  __syncthreads();
  lds_lines<double, N, 0, SIGN>(G, nullptr);
  __syncthreads();
  lds_lines<double, N, 1, SIGN>(G, nullptr);
  __syncthreads();
  lds_lines<double, N, 2, SIGN>(G, nullptr);
  __syncthreads();
}

// ρ(G) = FFT₋(ρ(r)/N) into rho_g, and ρ_xc = max(ρ + cs·ρ_core, 0) with its LDA ε_xc, v_xc; with
// VEFF also V_eff(G) = V_loc + 4πe² ρ(G)/|G|² + FFT₋(v_xc/N) into veff_g and V(r) = Re FFT₊(V_eff)
// into the plane-major and grid-order arrays (FP64 and FP32). One workgroup, the grid in LDS; each
// step is the arithmetic of the separate launches (r_to_g, k_lda_core, k_hartree_veff, the V(r)
// transform and k_vr_reorder), in the same order, so the results are theirs.
template <int N, bool VEFF>
__global__ void __launch_bounds__(SMALL_NT) k_rho_xc_veff_small(
    const double* __restrict__ rho_r, const double* __restrict__ core, double cs, double inv_n,
    double2* __restrict__ rho_g, double* __restrict__ rho_xc, double* __restrict__ exc, double* __restrict__ vxc,
    const double2* __restrict__ vloc_g, const double* __restrict__ g2, double fourpi_e2,
    double2* __restrict__ veff_g, double* __restrict__ v64, float* __restrict__ v32, double* __restrict__ n64,
    float* __restrict__ n32, int lda_ready) {
  extern __shared__ unsigned char smem_raw[];
  double2* G = reinterpret_cast<double2*>(smem_raw);
  constexpr int NN = N * N * N;
  // every pass issues the loads of a thread's points before it uses any of them (grid_points)
  grid_points<NN, double>([&](int i) { return rho_r[i]; },
                          [&](int i, double r) { G[i] = make_double2(r * inv_n, 0.0); });
  small_fft3d<N, -1>(G);
  struct RhoIn {
    double2 g;
    double r, c;
  };
  // lda_ready: rho_xc, exc and vxc already hold the LDA of this density (the mixing step that
  // produced it evaluated them); r then carries the given v_xc
  grid_points<NN, RhoIn>(
      [&](int i) {
        RhoIn x;
        x.g = G[i];
        x.r = lda_ready ? vxc[i] : rho_r[i];
        x.c = (core && !lda_ready) ? core[i] : 0.0;
        return x;
      },
      [&](int i, const RhoIn& x) {
        rho_g[i] = x.g;
        if (lda_ready) {
          if constexpr (VEFF) G[i] = make_double2(x.r * inv_n, 0.0);
          return;
        }
        const double r = core ? fmax(x.r + x.c * cs, 0.0) : x.r;
        rho_xc[i] = r;
        double e, v;
        lda_point(r, e, v);
        exc[i] = e;
        vxc[i] = v;
        if constexpr (VEFF) G[i] = make_double2(v * inv_n, 0.0);
      });
  if constexpr (VEFF) {
    small_fft3d<N, -1>(G);
    struct VIn {
      double2 rg, vl, x;
      double gg;
    };
    grid_points<NN, VIn>(
        [&](int i) {
          VIn y;
          y.gg = g2[i];
          y.rg = rho_g[i];  // written by this thread above
          y.vl = vloc_g[i];
          y.x = G[i];
          return y;
        },
        [&](int i, const VIn& y) {
          double2 acc = make_double2(0.0, 0.0);
          if (y.gg > kG2Zero) {
            acc.x = y.rg.x * fourpi_e2 / y.gg;
            acc.y = y.rg.y * fourpi_e2 / y.gg;
          }
          acc.x = y.vl.x + acc.x;
          acc.y = y.vl.y + acc.y;
          acc.x += y.x.x;
          acc.y += y.x.y;
          veff_g[i] = acc;
          G[i] = acc;
        });
    small_fft3d<N, 1>(G);
    grid_points<NN, double>([&](int g) { return G[g].x; },
                            [&](int g, double v) {
                              const int i1 = g / (N * N), i2 = (g / N) % N, i3 = g % N;
                              const size_t idx = ((size_t)i3 * N + i1) * N + i2;  // plane-major [i3][i1][i2]
                              v64[idx] = v;
                              v32[idx] = (float)v;
                              n64[g] = v;
                              n32[g] = (float)v;
                            });
  }
}

// out = Re FFT₊(in)·scale (FP64), the transform and k_c2r_scaled in one workgroup
template <int N>
__global__ void __launch_bounds__(SMALL_NT) k_fft_c2r_small(const double2* __restrict__ in, double scale,
                                                            double* __restrict__ out) {
  extern __shared__ unsigned char smem_raw[];
  double2* G = reinterpret_cast<double2*>(smem_raw);
  constexpr int NN = N * N * N;
  grid_points<NN, double2>([&](int i) { return in[i]; }, [&](int i, double2 x) { G[i] = x; });
  small_fft3d<N, 1>(G);
  grid_points<NN, double>([&](int i) { return G[i].x; }, [&](int i, double x) { out[i] = x * scale; });
}

#define PWK_GRID_SMALL_SIZES(X) X(8) X(9) X(10) X(12) X(15) X(16) X(18) X(20)

}  // namespace

void launch_hartree_veff(Context& c, const void* rho_g, double fourpi_e2, const void* v_local, const void* v_xc,
                         void* out, int precision, bool with_local, bool with_xc, bool with_hartree) {
  c.check_grid();
  size_t n = c.N;
  if (precision == PWK_FP64) {
    k_hartree_veff<double><<<blocks_for(n, 256), 256, 0, c.stream>>>(
        n, with_hartree ? (const double2*)rho_g : nullptr, c.g2_64.as<double>(), fourpi_e2,
        with_local ? (const double2*)v_local : nullptr, with_xc ? (const double2*)v_xc : nullptr, (double2*)out);
  } else {
    k_hartree_veff<float><<<blocks_for(n, 256), 256, 0, c.stream>>>(
        n, with_hartree ? (const float2*)rho_g : nullptr, c.g2_32.as<float>(), (float)fourpi_e2,
        with_local ? (const float2*)v_local : nullptr, with_xc ? (const float2*)v_xc : nullptr, (float2*)out);
  }
  PWK_HIP(hipGetLastError());
}

void launch_veff_sum(Context& c, const void* a, const void* b, const void* d, void* out, int precision) {
  c.check_grid();
  size_t n = c.N;
  if (precision == PWK_FP64)
    k_sum3<double><<<blocks_for(n, 256), 256, 0, c.stream>>>(n, (const double2*)a, (const double2*)b,
                                                             (const double2*)d, (double2*)out);
  else
    k_sum3<float><<<blocks_for(n, 256), 256, 0, c.stream>>>(n, (const float2*)a, (const float2*)b,
                                                            (const float2*)d, (float2*)out);
  PWK_HIP(hipGetLastError());
}

bool grid_small_supported(const Context& c) {
  const int n = c.dims[0];
  if (!c.small_hpsi || n != c.dims[1] || n != c.dims[2]) return false;
  switch (n) {
#define PWK_CASE(L) case L:
    PWK_GRID_SMALL_SIZES(PWK_CASE)
#undef PWK_CASE
    return (size_t)n * n * n * sizeof(double2) <= 160 * 1024;
    default:
      return false;
  }
}

void launch_rho_xc_veff_small(Context& c, const double* rho_r, const double* core, double cs, double2* rho_g,
                              double* rho_xc, double* exc, double* vxc, const double2* vloc_g, double fourpi_e2,
                              double2* veff_g, int slot, bool lda_ready) {
  PWK_REQUIRE(grid_small_supported(c), PWK_E_UNSUPPORTED, "grid has no fused single-workgroup kernels");
  const bool veff = veff_g != nullptr;
  const size_t o = (size_t)slot * c.N;
  if (veff) {
    if (slot == 0) reserve_veff_slots(c, 1);
    PWK_REQUIRE(slot >= 0 && c.vrn64.bytes >= (o + c.N) * sizeof(double) && c.vr32.bytes >= (o + c.N) * sizeof(float),
                PWK_E_STATE, "V_eff slot not reserved");
  }
  const size_t sh = c.N * sizeof(double2);
  const double inv_n = 1.0 / (double)c.N;
  switch (c.dims[0]) {
#define PWK_CASE(L)                                                                                            \
  case L:                                                                                                      \
    if (veff) {                                                                                                \
      ensure_lds_limit((const void*)k_rho_xc_veff_small<L, true>, sh);                                        \
      k_rho_xc_veff_small<L, true><<<1, SMALL_NT, sh, c.stream>>>(                                            \
          rho_r, core, cs, inv_n, rho_g, rho_xc, exc, vxc, vloc_g, c.g2_64.as<double>(), fourpi_e2, veff_g,     \
          c.vr64.as<double>() + o, c.vr32.as<float>() + o, c.vrn64.as<double>() + o, c.vrn32.as<float>() + o,   \
          lda_ready ? 1 : 0);                                                                                  \
    } else {                                                                                                   \
      ensure_lds_limit((const void*)k_rho_xc_veff_small<L, false>, sh);                                       \
      k_rho_xc_veff_small<L, false><<<1, SMALL_NT, sh, c.stream>>>(rho_r, core, cs, inv_n, rho_g, rho_xc, exc, \
                                                                   vxc, nullptr, nullptr, 0.0, nullptr,         \
                                                                   nullptr, nullptr, nullptr, nullptr,          \
                                                                   lda_ready ? 1 : 0);                          \
    }                                                                                                          \
    break;
    PWK_GRID_SMALL_SIZES(PWK_CASE)
#undef PWK_CASE
    default:
      throw Error(PWK_E_UNSUPPORTED, "grid has no fused single-workgroup kernels");
  }
  PWK_HIP(hipGetLastError());
  if (veff && slot == 0) c.have_v = true;
}

void launch_fft_c2r_small(Context& c, const double2* in, double scale, double* out) {
  PWK_REQUIRE(grid_small_supported(c), PWK_E_UNSUPPORTED, "grid has no fused single-workgroup kernels");
  const size_t sh = c.N * sizeof(double2);
  switch (c.dims[0]) {
#define PWK_CASE(L)                                                                    \
  case L:                                                                              \
    ensure_lds_limit((const void*)k_fft_c2r_small<L>, sh);                             \
    k_fft_c2r_small<L><<<1, SMALL_NT, sh, c.stream>>>(in, scale, out);                 \
    break;
    PWK_GRID_SMALL_SIZES(PWK_CASE)
#undef PWK_CASE
    default:
      throw Error(PWK_E_UNSUPPORTED, "grid has no fused single-workgroup kernels");
  }
  PWK_HIP(hipGetLastError());
}

void launch_lda_core(Context& c, const double* rho, const double* core, double cs, double* rho_xc, double* exc,
                     double* vxc, double2* g, double gs) {
  c.check_grid();
  k_lda_core<<<blocks_for(c.N, 256), 256, 0, c.stream>>>(c.N, rho, core, cs, rho_xc, exc, vxc, g, gs);
  PWK_HIP(hipGetLastError());
}

void launch_lda(Context& c, const void* rho, void* exc, void* vxc, int precision) {
  c.check_grid();
  size_t n = c.N;
  if (precision == PWK_FP64)
    k_lda<double><<<blocks_for(n, 256), 256, 0, c.stream>>>(n, (const double*)rho, (double*)exc, (double*)vxc);
  else
    k_lda<float><<<blocks_for(n, 256), 256, 0, c.stream>>>(n, (const float*)rho, (float*)exc, (float*)vxc);
  PWK_HIP(hipGetLastError());
}

void launch_lsda(Context& c, const void* ru, const void* rd, void* exc, void* vu, void* vd, int precision) {
  c.check_grid();
  size_t n = c.N;
  if (precision == PWK_FP64)
    k_lsda<double><<<blocks_for(n, 256), 256, 0, c.stream>>>(n, (const double*)ru, (const double*)rd,
                                                             (double*)exc, (double*)vu, (double*)vd);
  else
    k_lsda<float><<<blocks_for(n, 256), 256, 0, c.stream>>>(n, (const float*)ru, (const float*)rd, (float*)exc,
                                                            (float*)vu, (float*)vd);
  PWK_HIP(hipGetLastError());
}

void launch_convert(Context& c, const void* src, void* dst, size_t n, bool to_fp32) {
  if (n == 0) return;
  if (to_fp32)
    k_d2f<<<blocks_for(n, 256), 256, 0, c.stream>>>(n, (const double*)src, (float*)dst);
  else
    k_f2d<<<blocks_for(n, 256), 256, 0, c.stream>>>(n, (const float*)src, (double*)dst);
  PWK_HIP(hipGetLastError());
}

}  // namespace pwk
