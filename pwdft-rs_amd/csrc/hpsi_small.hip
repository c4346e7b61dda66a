// Fused H·ψ for small cubic grids: one workgroup per band holds the band's
// whole N³ grid in LDS and runs the complete product in one launch:
//   scatter ψ(G) -> inverse FFT z, y, x (the x pass multiplies V(r) between its
//   inverse and forward transforms) -> forward FFT y, z -> gather the sphere,
//   Hψ = T_kin ψ + (1/N)·FFT[V·IFFT ψ] + B·((D·B^H)·ψ)
// (src/scf/potentials.rs:163-190 and src/potential/nonlocal.rs:432-450, the
// same product as the three-kernel stick path plus the two projector GEMMs).
//
// Small decks (a few hundred plane waves, 15³-24³ grids) are launch- and
// latency-bound on the stick path: five launches and two HBM round trips of
// the stick buffer per H·ψ. Here ψ is read once, Hψ written once, and the
// projectors and V(r) stay in L2 across the bands of a launch.
//
// a*b + c is fused only where one source expression holds it (never across statements), so a
// line of the transforms or of the epilogues compiles to the same arithmetic in every kernel of
// this file that inlines it: the one-launch filter pass, the per-step H·ψ and the one-launch
// Lanczos then agree bit for bit with the step launches they replace.
#pragma clang fp contract(on)
#include "common.hpp"
#include "fft_reg.hpp"
#include "lds_fft.hpp"

namespace pwk {

namespace {

constexpr int kSmallWaves = SMALL_NT / 64;

// LDS entries (complex) of the nonlocal scratch: the projections and the waves' partial sums
inline size_t nl_lds_entries(int nch) { return (size_t)(1 + kSmallWaves) * nch; }

// One block of nl_partials: rows gb .. gb + PU − 1 of channel c (TAIL: rows from g1 on count zero),
// their products summed pairwise and the sum added to (ar, ai).
template <typename T, int PU, bool TAIL>
__device__ __forceinline__ void nl_block(T& ar, T& ai, const typename V2s<T>::t* v,
                                         const typename V2s<T>::t* __restrict__ dk, int nch, int c, int gb, int g1) {
  using V2 = typename V2s<T>::t;
  // every load is issued first (in a tail block, rows past the quarter's end read its last row,
  // which exists since gb < g1, and their products are replaced by zeros), so no load sits behind
  // a branch and the block waits for its loads once
  V2 d[PU], x[PU];
#pragma unroll
  for (int u = 0; u < PU; ++u) {
    const int gg = TAIL ? min(gb + u, g1 - 1) : gb + u;
    d[u] = (dk + gg * nch)[c];  // uniform row pointer (npw·nch < 2³¹), lane offset c
    x[u] = v[gg];
  }
  // scheduling groups: the PU global loads, then the PU LDS loads, ahead of the arithmetic
  // (left alone, the scheduler issues one load at a time to save registers)
  __builtin_amdgcn_sched_group_barrier(0x0020, PU, 0);
  __builtin_amdgcn_sched_group_barrier(0x0100, PU, 0);
  T pr[PU], pi[PU];
#pragma unroll
  for (int u = 0; u < PU; ++u) {
    pr[u] = d[u].x * x[u].x - d[u].y * x[u].y;
    pi[u] = d[u].x * x[u].y + d[u].y * x[u].x;
    if (TAIL && gb + u >= g1) pr[u] = pi[u] = 0;
  }
#pragma unroll
  for (int h = PU / 2; h > 0; h >>= 1)
#pragma unroll
    for (int u = 0; u < h; ++u) {
      pr[u] += pr[u + h];
      pi[u] += pi[u + h];
    }
  ar += pr[0];
  ai += pi[0];
}

// Partial projections P_c = Σ_G (D·B^H)[c, G] v(G) of a vector v held in LDS. D·B^H is column-major
// nch × npw (entry (c, G) at G·nch + c), so each of its columns is contiguous over c: lanes take
// channels, so that a wave reads whole columns, and the waves take contiguous quarters of the plane
// waves, with PU columns in flight per lane. The PU products of a block are summed pairwise and
// the block sums added in plane-wave order (rounding error grows with log2 PU + the block count,
// not with the quarter's length). Wave w's sums go to part[w·nch + c]; nl_combine adds them after a
// barrier.
template <typename T>
__device__ __forceinline__ void nl_partials(typename V2s<T>::t* part, const typename V2s<T>::t* v,
                                            const typename V2s<T>::t* __restrict__ dk, int npw, int nch) {
  constexpr int PU = sizeof(T) == 4 ? 16 : 8;  // rows in flight (register budget: 4·PU values)
  // the wave index is uniform over the wave: read it as such, so that the row range, the clamps
  // and the row offsets are scalar values and each load adds only the lane's channel offset
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int per = (npw + kSmallWaves - 1) / kSmallWaves, g0 = wave * per, g1 = min(npw, g0 + per);
  for (int c = lane; c < nch; c += 64) {
    T ar = 0, ai = 0;
    int gb = g0;
    for (; gb + PU <= g1; gb += PU) nl_block<T, PU, false>(ar, ai, v, dk, nch, c, gb, g1);
    if (gb < g1) nl_block<T, PU, true>(ar, ai, v, dk, nch, c, gb, g1);
    part[wave * nch + c].x = ar;
    part[wave * nch + c].y = ai;
  }
}

// proj[c] = the waves' partial sums added in wave order
template <typename T>
__device__ __forceinline__ void nl_combine(typename V2s<T>::t* proj, const typename V2s<T>::t* part, int nch) {
  using V2 = typename V2s<T>::t;
  for (int c = threadIdx.x; c < nch; c += SMALL_NT) {
    V2 a = part[c];
#pragma unroll
    for (int w = 1; w < kSmallWaves; ++w) {
      a.x += part[w * nch + c].x;
      a.y += part[w * nch + c].y;
    }
    proj[c] = a;
  }
}

// (ox, oy) += b·p (complex), as four FMAs in a fixed order: every kernel that applies the
// projectors uses this form, so that they agree bit for bit
template <typename T>
__device__ __forceinline__ void nl_term(T& ox, T& oy, const typename V2s<T>::t& b, const typename V2s<T>::t& p) {
  ox = fma(b.x, p.x, ox);
  ox = fma(-b.y, p.y, ox);
  oy = fma(b.x, p.y, oy);
  oy = fma(b.y, p.x, oy);
}

// One batch of nl_expand: channels c0 .. c0 + EC − 1 (TAIL: channels from nch on read the last
// channel and add the exact zero b·p with b = 0, which leaves (ox, oy) unchanged). All EC loads are
// issued before any is used.
template <typename T, int EC, bool TAIL>
__device__ __forceinline__ void nl_expand_batch(T& ox, T& oy, const typename V2s<T>::t* __restrict__ bk,
                                                const typename V2s<T>::t* proj, int g, int npw, int nch, int c0) {
  using V2 = typename V2s<T>::t;
  V2 bb[EC], pc[EC];
#pragma unroll
  for (int j = 0; j < EC; ++j) {
    const int cc = TAIL ? min(c0 + j, nch - 1) : c0 + j;
    bb[j] = (bk + cc * npw)[g];  // uniform column pointer (npw·nch < 2³¹), lane offset g
    pc[j] = proj[cc];
  }
  __builtin_amdgcn_sched_group_barrier(0x0020, EC, 0);
  __builtin_amdgcn_sched_group_barrier(0x0100, EC, 0);
#pragma unroll
  for (int j = 0; j < EC; ++j) {
    if (TAIL && c0 + j >= nch) bb[j].x = bb[j].y = 0;
    nl_term<T>(ox, oy, bb[j], pc[j]);
  }
}

// (ox, oy) += Σ_c B[G, c] P_c in channel order. B is column-major npw × nch (entry (G, c) at
// c·npw + G), so the lanes' loads of one channel coalesce; EC channels' loads are in flight at once.
template <typename T>
__device__ __forceinline__ void nl_expand(T& ox, T& oy, const typename V2s<T>::t* __restrict__ bk,
                                          const typename V2s<T>::t* proj, int g, int npw, int nch) {
  constexpr int EC = sizeof(T) == 4 ? 18 : 9;  // channels in flight (register budget: 4·EC values)
  int c0 = 0;
  for (; c0 + EC <= nch; c0 += EC) nl_expand_batch<T, EC, false>(ox, oy, bk, proj, g, npw, nch, c0);
  if (c0 < nch) nl_expand_batch<T, EC, true>(ox, oy, bk, proj, g, npw, nch, c0);
}

// Pruned transform passes (grids with one line per thread, N² <= SMALL_NT): a z-line (x, y) carries
// data only where the basis has plane waves in column (x, y), a y-line (x, z) only where plane x
// has such a column. LineAct holds the thread's two flags (its line l = threadIdx.x in either
// pass); small_line_activity derives them once per launch from the basis's grid indices, using the
// LDS grid as scratch (before the grid's first use; every write of a flag stores 1).
struct LineAct {
  bool z = true, y = true;
};

template <int N>
__device__ __forceinline__ LineAct small_line_activity(void* scratch, const int* __restrict__ g2fft, int npw) {
  LineAct a;
  if constexpr (N * N <= SMALL_NT) {
    int* fl = reinterpret_cast<int*>(scratch);
    for (int i = threadIdx.x; i < N * N + N; i += SMALL_NT) fl[i] = 0;
    lds_barrier();
    for (int g = threadIdx.x; g < npw; g += SMALL_NT) {
      const int col = g2fft[g] / N;  // x·N + y
      fl[col] = 1;
      fl[N * N + col / N] = 1;
    }
    lds_barrier();
    const int t = threadIdx.x;
    a.z = t < N * N && fl[t] != 0;
    a.y = t < N * N && fl[N * N + t / N] != 0;
    lds_barrier();
  }
  return a;
}

// The grid's state before the scatter of a vector: with pruned passes only the columns that carry
// plane waves are cleared (the first z-pass writes zeros on every other z-line), else all of it.
template <typename T, int N>
__device__ __forceinline__ void small_clear(typename V2s<T>::t* G, const LineAct& act) {
  if constexpr (N * N <= SMALL_NT) {
    if ((int)threadIdx.x < N * N && act.z) {
#pragma unroll
      for (int z = 0; z < N; ++z) {
        G[threadIdx.x * N + z].x = 0;
        G[threadIdx.x * N + z].y = 0;
      }
    }
  } else {
    for (int i = threadIdx.x; i < N * N * N; i += SMALL_NT) {
      G[i].x = 0;
      G[i].y = 0;
    }
  }
}

// The inverse z and y passes after the scatter (pruned where the grid allows: inactive lines are
// written as zeros) and the forward y and z passes before the gather (pruned: inactive lines are
// not needed).
template <typename T, int N, int SIGN>
__device__ __forceinline__ void small_zpass(typename V2s<T>::t* G, const LineAct& act) {
  if constexpr (N * N <= SMALL_NT)
    lds_lines_pruned<T, N, 2, SIGN, SIGN == 1>(G, act.z);
  else
    lds_lines<T, N, 2, SIGN>(G, nullptr);
}

template <typename T, int N, int SIGN>
__device__ __forceinline__ void small_ypass(typename V2s<T>::t* G, const LineAct& act) {
  if constexpr (N * N <= SMALL_NT)
    lds_lines_pruned<T, N, 1, SIGN, SIGN == 1>(G, act.y);
  else
    lds_lines<T, N, 1, SIGN>(G, nullptr);
}

// Band b (k-point b / nbk) of psi -> hpsi. vr is V(r) in natural grid order; g2fft the flat
// grid index of each plane wave; B (npw × nch) and DBH (nch × npw) column-major blocks per k.
// With coef (α, c, β per k-point) the output is the Chebyshev recurrence step
// α_k (Hψ − c_k ψ) − β_k X (X may be null) instead of Hψ.
template <typename T, int N>
__device__ __forceinline__ void hpsi_small_body(int bid, int nblocks, const typename V2s<T>::t* __restrict__ psi,
                                                         typename V2s<T>::t* __restrict__ hpsi,
                                                         const T* __restrict__ kin, const T* __restrict__ vr,
                                                         const int* __restrict__ g2fft, int npw, int nbk, int nch,
                                                         const typename V2s<T>::t* __restrict__ B,
                                                         const typename V2s<T>::t* __restrict__ DBH, T inv_n,
                                                         const double* __restrict__ coef,
                                                         const typename V2s<T>::t* __restrict__ X, int nkv,
                                                         const int* __restrict__ band_k,
                                                         const int* __restrict__ skip) {
  using V2 = typename V2s<T>::t;
  if (skip && *skip) return;  // a filter pass after the residual gate passed (uniform)
  extern __shared__ unsigned char smem_raw[];
  V2* G = reinterpret_cast<V2*>(smem_raw);
  V2* proj = G + N * N * N;   // nch projections
  V2* part = proj + nch;      // the waves' partial projections
  // XCD-aware band order: workgroups are dealt to the 8 XCDs round-robin, so workgroup i runs
  // on XCD i % 8; each XCD takes a contiguous range of bands, which keeps the bands of one
  // k-point (and its projector blocks) in one XCD's L2
  const int q = nblocks >> 3, r = nblocks & 7, xcd = bid & 7;
  const int b = xcd * q + min(xcd, r) + (bid >> 3), kb = b / nbk;
  // band block kb: k-point kb mod nkv, V(r) channel kb / nkv (coefficients per block); with a
  // band table (ragged k × band batches) the k-point of every band is read from it
  const int k = band_k ? band_k[b] : nkv > 0 ? kb % nkv : kb;
  if (nkv > 0) vr += (size_t)(kb / nkv) * N * N * N;
  const V2* p = psi + (size_t)b * npw;
  const T* t = kin + (size_t)k * npw;
  const V2* bk = B + (size_t)k * npw * nch;
  const V2* dk = DBH + (size_t)k * nch * npw;
  const V2* xb = X ? X + (size_t)b * npw : nullptr;
  // operands of the epilogue for this thread's first plane wave, loaded now so that their latency
  // hides behind the transforms (projector rows for up to PF channels)
  constexpr int PF = 16;
  const int g0 = threadIdx.x;
  const bool pf = g0 < npw && nch <= PF;
  V2 pf_b[PF], pf_v, pf_x;
  T pf_t = 0;
  pf_v.x = pf_v.y = pf_x.x = pf_x.y = 0;
  if (pf) {
    pf_v = p[g0];
    pf_t = t[g0];
    if (xb) pf_x = xb[g0];
#pragma unroll
    for (int ch = 0; ch < PF; ++ch)
      if (ch < nch) pf_b[ch] = bk[(size_t)ch * npw + g0];
  }
  // V(r) of the thread's x-line is loaded now too and used after the first transforms
  constexpr bool PRE_V = N * N <= SMALL_NT;
  T vline[PRE_V ? N : 1];
  if constexpr (PRE_V) {
    if (threadIdx.x < N * N) {
#pragma unroll
      for (int f = 0; f < N; ++f) vline[f] = vr[threadIdx.x + f * N * N];
    }
  }
  const LineAct act = small_line_activity<N>(G, g2fft, npw);
  small_clear<T, N>(G, act);
  lds_barrier();
  for (int g = threadIdx.x; g < npw; g += SMALL_NT) G[g2fft[g]] = p[g];
  nl_partials<T>(part, p, dk, npw, nch);
  lds_barrier();
  nl_combine<T>(proj, part, nch);
  small_zpass<T, N, 1>(G, act);
  lds_barrier();
  small_ypass<T, N, 1>(G, act);
  lds_barrier();
  if constexpr (PRE_V)
    lds_xlines_v<T, N, 1>(G, vline);
  else
    lds_lines<T, N, 0, 1>(G, vr);
  lds_barrier();
  small_ypass<T, N, -1>(G, act);
  lds_barrier();
  small_zpass<T, N, -1>(G, act);
  lds_barrier();
  V2* h = hpsi + (size_t)b * npw;
  T alpha = 0, cshift = 0, beta = 0;
  if (coef) {
    alpha = (T)coef[3 * kb];
    cshift = (T)coef[3 * kb + 1];
    beta = (T)coef[3 * kb + 2];
  }
  for (int g = threadIdx.x; g < npw; g += SMALL_NT) {
    const bool first = pf && g == g0;
    const V2 v = first ? pf_v : p[g], w = G[g2fft[g]];
    const T tk = first ? pf_t : t[g];
    T ox = fma(tk, v.x, w.x * inv_n), oy = fma(tk, v.y, w.y * inv_n);
    if (first) {
#pragma unroll
      for (int ch = 0; ch < PF; ++ch)
        if (ch < nch) nl_term<T>(ox, oy, pf_b[ch], proj[ch]);
    } else {
      nl_expand<T>(ox, oy, bk, proj, g, npw, nch);
    }
    V2 o;
    if (coef) {
      // the elementwise step of the filter (subspace.hip k_cheb_step), with its FMAs explicit
      o.x = alpha * fma(-cshift, v.x, ox);
      o.y = alpha * fma(-cshift, v.y, oy);
      if (xb) {
        const V2 xv = first ? pf_x : xb[g];
        o.x = fma(-beta, xv.x, o.x);
        o.y = fma(-beta, xv.y, o.y);
      }
    } else {
      o.x = ox;
      o.y = oy;
    }
    h[g] = o;
  }
}

// The local part of H·v for a vector held in LDS (cur, npw entries): scatter onto the grid G
// (cleared as small_clear does; the passes pruned by act where the grid allows),
// partial projections of v (nl_partials, summed into proj by nl_combine after the next barrier),
// inverse FFT z, y, x with V(r) applied between the x transforms (vline: V(r) of the thread's
// x-line, when N² <= SMALL_NT; else read from vr), forward FFT y, z. On return G holds
// FFT[V·IFFT v] (unnormalised) and proj the projections; the caller's epilogue adds the kinetic
// and nonlocal terms in the order k_hpsi_small uses. gi0 is the grid index of plane wave
// threadIdx.x (when threadIdx.x < npw), which the caller keeps in a register across steps.
template <typename T, int N>
__device__ __forceinline__ void small_h_core(typename V2s<T>::t* G, typename V2s<T>::t* proj,
                                             typename V2s<T>::t* part, const typename V2s<T>::t* cur,
                                             const T (&vline)[N * N <= SMALL_NT ? N : 1], const T* __restrict__ vr,
                                             const int* __restrict__ g2fft, int gi0, int npw, int nch,
                                             const typename V2s<T>::t* __restrict__ dk, const LineAct& act) {
  constexpr bool PRE_V = N * N <= SMALL_NT;
  small_clear<T, N>(G, act);
  lds_barrier();
  if ((int)threadIdx.x < npw) G[gi0] = cur[threadIdx.x];
  for (int g = threadIdx.x + SMALL_NT; g < npw; g += SMALL_NT) G[g2fft[g]] = cur[g];
  nl_partials<T>(part, cur, dk, npw, nch);
  lds_barrier();
  nl_combine<T>(proj, part, nch);
  small_zpass<T, N, 1>(G, act);
  lds_barrier();
  small_ypass<T, N, 1>(G, act);
  lds_barrier();
  if constexpr (PRE_V)
    lds_xlines_v<T, N, 1>(G, vline);
  else
    lds_lines<T, N, 0, 1>(G, vr);
  lds_barrier();
  small_ypass<T, N, -1>(G, act);
  lds_barrier();
  small_zpass<T, N, -1>(G, act);
  lds_barrier();
}

// Re⟨x, y⟩ over n complex entries held in LDS, summed as the 2n reals in the order of
// k_block_dot (subspace.hip): thread t takes reals t, t + SMALL_NT, ..., then the wave sums
// by shuffle-down and the four wave sums pairwise. Every thread returns the total.
__device__ __forceinline__ double lds_real_dot(const double2* x, const double2* y, int n, double* red) {
  const double* xr = reinterpret_cast<const double*>(x);
  const double* yr = reinterpret_cast<const double*>(y);
  double acc = 0.0;
  for (int i = threadIdx.x; i < 2 * n; i += SMALL_NT) acc += xr[i] * yr[i];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  const double s = (red[0] + red[1]) + (red[2] + red[3]);
  __syncthreads();
  return s;
}

// Lanczos recurrence for the upper spectral bound, every step in one launch: one workgroup per
// band block kb (k-point kb mod nkv with V(r) slot kb / nkv when nkv > 0, else k-point kb), the
// three vectors in LDS. Each step is the sequence the step launches of the general path run
// (H·v as in k_hpsi_small; α = Re⟨v, Hv⟩; w = (Hv − α v) − β v_prev; β' = |w|; v ← w/β', zero
// when β' < 1e-12), with the same reductions, so the coefficients are those of that path;
// α_j goes to ab[j·nke + kb], β_j to ab[(steps + j)·nke + kb].
template <int N>
__global__ void __launch_bounds__(SMALL_NT) k_lanczos_small(const int* __restrict__ g2fft, int npw, int nch,
                                                            const double* __restrict__ kin,
                                                            const double* __restrict__ vr,
                                                            const double2* __restrict__ B,
                                                            const double2* __restrict__ DBH, double inv_n, int nkv,
                                                            int steps, double* __restrict__ ab) {
  extern __shared__ unsigned char smem_raw[];
  __shared__ double red[4];
  double2* G = reinterpret_cast<double2*>(smem_raw);
  double2* proj = G + N * N * N;
  double2* part = proj + nch;
  double2* v0 = part + kSmallWaves * nch;
  double2* v1 = v0 + npw;
  double2* w = v1 + npw;
  const int kb = blockIdx.x, nke = gridDim.x;
  const int k = nkv > 0 ? kb % nkv : kb;
  if (nkv > 0) vr += (size_t)(kb / nkv) * N * N * N;
  const double* t = kin + (size_t)k * npw;
  const double2* bk = B + (size_t)k * npw * nch;
  const double2* dk = DBH + (size_t)k * nch * npw;
  constexpr bool PRE_V = N * N <= SMALL_NT;
  double vline[PRE_V ? N : 1];
  if constexpr (PRE_V) {
    if (threadIdx.x < N * N) {
#pragma unroll
      for (int f = 0; f < N; ++f) vline[f] = vr[threadIdx.x + f * N * N];
    }
  }
  // grid index and kinetic energy of plane wave threadIdx.x, kept for every step
  const int gi0 = (int)threadIdx.x < npw ? g2fft[threadIdx.x] : 0;
  const double tk0 = (int)threadIdx.x < npw ? t[threadIdx.x] : 0.0;
  const LineAct act = small_line_activity<N>(G, g2fft, npw);
  for (int g = threadIdx.x; g < npw; g += SMALL_NT) {
    v1[g] = start_vector_entry(g, 0, kb);
    v0[g] = make_double2(0.0, 0.0);
  }
  __syncthreads();
  {
    const double s = lds_real_dot(v1, v1, npw, red);
    for (int g = threadIdx.x; g < npw; g += SMALL_NT) {
      v1[g].x *= 1.0 / sqrt(s);
      v1[g].y *= 1.0 / sqrt(s);
    }
  }
  double beta = 0.0;
  for (int j = 0; j < steps; ++j) {
    lds_barrier();
    small_h_core<double, N>(G, proj, part, v1, vline, vr, g2fft, gi0, npw, nch, dk, act);
    auto epi = [&](int g, int gi, double tk) {
      const double2 v = v1[g], hw = G[gi];
      double ox = fma(tk, v.x, hw.x * inv_n), oy = fma(tk, v.y, hw.y * inv_n);
      nl_expand<double>(ox, oy, bk, proj, g, npw, nch);
      w[g] = make_double2(ox, oy);
    };
    if ((int)threadIdx.x < npw) epi(threadIdx.x, gi0, tk0);
    for (int g = threadIdx.x + SMALL_NT; g < npw; g += SMALL_NT) epi(g, g2fft[g], t[g]);
    __syncthreads();
    const double a = lds_real_dot(v1, w, npw, red);
    const double alpha = 1.0, cshift = a;
    for (int g = threadIdx.x; g < npw; g += SMALL_NT) {
      const double2 h = w[g], y = v1[g], x = v0[g];
      double2 o;
      o.x = alpha * (h.x - cshift * y.x);
      o.y = alpha * (h.y - cshift * y.y);
      o.x -= beta * x.x;
      o.y -= beta * x.y;
      w[g] = o;
    }
    __syncthreads();
    const double b = sqrt(fmax(lds_real_dot(w, w, npw, red), 0.0));
    if (threadIdx.x == 0) {
      ab[(size_t)j * nke + kb] = a;
      ab[(size_t)(steps + j) * nke + kb] = b;
    }
    beta = b;
    if (j + 1 == steps) break;
    const double inv = b < 1e-12 ? 0.0 : 1.0 / b;
    double2* s0 = v0;
    v0 = v1;
    for (int g = threadIdx.x; g < npw; g += SMALL_NT) {
      w[g].x *= inv;
      w[g].y *= inv;
    }
    v1 = w;
    w = s0;
  }
}

// The whole Chebyshev filter pass of one band in one workgroup: the `degree` steps of the scaled
// three-term recurrence
//   Y_1 = α_0 (H − c_0) Y_0,   Y_{i+1} = α_i (H − c_i) Y_i − β_i Y_{i−1}
// with Y_0 the band's FP64 vector of V rounded to T and the result written back as FP64 to Yout.
// The recurrence is independent per band, so the two live vectors stay in LDS between the steps
// (cur and prv below; the new vector overwrites prv element by element, since each thread owns
// its plane waves) and V(r) of the thread's x-line stays in registers for the whole pass. Each
// step is the H·ψ of k_hpsi_small with its Chebyshev epilogue, in the same order of operations,
// so the pass gives what `degree` launches of k_hpsi_small give. coef holds (α, c, β) per step and
// band block: coef[(i·nblk + kb)·3 + j], nblk = nblocks / nbk (nblocks: the bands of the pass).
template <typename T, int N>
__device__ __forceinline__ void cheb_filter_body(int bid, int nblocks, const double2* __restrict__ V,
                                                                double2* __restrict__ Yout, const T* __restrict__ kin,
                                                                const T* __restrict__ vr,
                                                                const int* __restrict__ g2fft, int npw, int nbk,
                                                                int nch, const typename V2s<T>::t* __restrict__ B,
                                                                const typename V2s<T>::t* __restrict__ DBH, T inv_n,
                                                                const double* __restrict__ coef, int degree, int nkv,
                                                                const int* __restrict__ skip) {
  using V2 = typename V2s<T>::t;
  if (skip && *skip) return;  // a filter pass after the residual gate passed (uniform)
  extern __shared__ unsigned char smem_raw[];
  V2* G = reinterpret_cast<V2*>(smem_raw);
  V2* proj = G + N * N * N;
  V2* part = proj + nch;
  V2* cur = part + kSmallWaves * nch;
  V2* prv = cur + npw;
  double* cfl = reinterpret_cast<double*>(prv + npw);  // this block's (α, c, β) of every step
  const int nblk = nblocks / nbk;
  // XCD-aware band order, as in k_hpsi_small
  const int q = nblocks >> 3, r = nblocks & 7, xcd = bid & 7;
  const int b = xcd * q + min(xcd, r) + (bid >> 3), kb = b / nbk;
  const int k = nkv > 0 ? kb % nkv : kb;
  if (nkv > 0) vr += (size_t)(kb / nkv) * N * N * N;
  const T* t = kin + (size_t)k * npw;
  const V2* bk = B + (size_t)k * npw * nch;
  const V2* dk = DBH + (size_t)k * nch * npw;
  constexpr bool PRE_V = N * N <= SMALL_NT;
  T vline[PRE_V ? N : 1];
  if constexpr (PRE_V) {
    if (threadIdx.x < N * N) {
#pragma unroll
      for (int f = 0; f < N; ++f) vline[f] = vr[threadIdx.x + f * N * N];
    }
  }
  // grid index and kinetic energy of plane wave threadIdx.x, kept for every step
  const int gi0 = (int)threadIdx.x < npw ? g2fft[threadIdx.x] : 0;
  const T tk0 = (int)threadIdx.x < npw ? t[threadIdx.x] : (T)0;
  const LineAct act = small_line_activity<N>(G, g2fft, npw);
  const double2* v0 = V + (size_t)b * npw;
  for (int g = threadIdx.x; g < npw; g += SMALL_NT) {
    const double2 x = v0[g];
    cur[g].x = (T)x.x;
    cur[g].y = (T)x.y;
  }
  // the coefficients are loaded now (they may sit in host memory) and stored to LDS after the
  // first transforms, so their latency hides behind them (3·degree <= SMALL_NT)
  const double cpre = (int)threadIdx.x < 3 * degree
                          ? coef[((size_t)(threadIdx.x / 3) * nblk + kb) * 3 + threadIdx.x % 3]
                          : 0.0;
  for (int step = 0; step < degree; ++step) {
    small_h_core<T, N>(G, proj, part, cur, vline, vr, g2fft, gi0, npw, nch, dk, act);
    if (step == 0) {
      if ((int)threadIdx.x < 3 * degree) cfl[threadIdx.x] = cpre;
      lds_barrier();
    }
    const T alpha = (T)cfl[3 * step], cshift = (T)cfl[3 * step + 1], beta = (T)cfl[3 * step + 2];
    auto epi = [&](int g, int gi, T tk) {
      const V2 v = cur[g], w = G[gi];
      T ox = fma(tk, v.x, w.x * inv_n), oy = fma(tk, v.y, w.y * inv_n);
      nl_expand<T>(ox, oy, bk, proj, g, npw, nch);
      V2 o;
      o.x = alpha * fma(-cshift, v.x, ox);
      o.y = alpha * fma(-cshift, v.y, oy);
      if (step > 0) {
        const V2 xv = prv[g];
        o.x = fma(-beta, xv.x, o.x);
        o.y = fma(-beta, xv.y, o.y);
      }
      prv[g] = o;
    };
    if ((int)threadIdx.x < npw) epi(threadIdx.x, gi0, tk0);
    for (int g = threadIdx.x + SMALL_NT; g < npw; g += SMALL_NT) epi(g, g2fft[g], t[g]);
    lds_barrier();
    V2* s = cur;
    cur = prv;
    prv = s;
  }
  double2* y = Yout + (size_t)b * npw;
  for (int g = threadIdx.x; g < npw; g += SMALL_NT) y[g] = make_double2((double)cur[g].x, (double)cur[g].y);
}

template <typename T, int N>
__global__ void __launch_bounds__(SMALL_NT) k_hpsi_small(const typename V2s<T>::t* __restrict__ psi,
                                                         typename V2s<T>::t* __restrict__ hpsi,
                                                         const T* __restrict__ kin, const T* __restrict__ vr,
                                                         const int* __restrict__ g2fft, int npw, int nbk, int nch,
                                                         const typename V2s<T>::t* __restrict__ B,
                                                         const typename V2s<T>::t* __restrict__ DBH, T inv_n,
                                                         const double* __restrict__ coef,
                                                         const typename V2s<T>::t* __restrict__ X, int nkv,
                                                         const int* __restrict__ band_k,
                                                         const int* __restrict__ skip) {
  hpsi_small_body<T, N>((int)blockIdx.x, (int)gridDim.x, psi, hpsi, kin, vr, g2fft, npw, nbk, nch, B, DBH, inv_n,
                        coef, X, nkv, band_k, skip);
}

template <typename T, int N>
__global__ void __launch_bounds__(SMALL_NT) k_cheb_filter_small(const double2* __restrict__ V,
                                                                double2* __restrict__ Yout, const T* __restrict__ kin,
                                                                const T* __restrict__ vr,
                                                                const int* __restrict__ g2fft, int npw, int nbk,
                                                                int nch, const typename V2s<T>::t* __restrict__ B,
                                                                const typename V2s<T>::t* __restrict__ DBH, T inv_n,
                                                                const double* __restrict__ coef, int degree, int nkv,
                                                                const int* __restrict__ skip) {
  cheb_filter_body<T, N>((int)blockIdx.x, (int)gridDim.x, V, Yout, kin, vr, g2fft, npw, nbk, nch, B, DBH, inv_n, coef,
                         degree, nkv, skip);
}

// The same products for several solves in one launch (the passes of concurrent requests that
// share grid, basis size, projector count, band counts and filter degree): blockIdx.y selects the
// request, blockIdx.x the band within it; each band runs exactly the code of the one-solve
// kernels above, so a request's result does not depend on the batch it ran in.
template <typename T, int N>
__global__ void __launch_bounds__(SMALL_NT) k_cheb_filter_multi(PassSmallMulti m, int npw, int nbk, int nch,
                                                                T inv_n, int degree) {
  using V2 = typename V2s<T>::t;
  const PassSmallReq& q = m.r[blockIdx.y];
  if ((int)blockIdx.x >= q.nbands) return;  // uniform: the request has fewer bands than the grid row
  cheb_filter_body<T, N>((int)blockIdx.x, q.nbands, (const double2*)q.in, (double2*)q.out, (const T*)q.kin,
                         (const T*)q.vr, q.g2fft, npw, nbk, nch, (const V2*)q.B, (const V2*)q.DBH, inv_n, q.coef,
                         degree, q.nkv, q.skip);
}

template <typename T, int N>
__global__ void __launch_bounds__(SMALL_NT) k_hpsi_multi(PassSmallMulti m, int npw, int nbk, int nch, T inv_n) {
  using V2 = typename V2s<T>::t;
  const PassSmallReq& q = m.r[blockIdx.y];
  if ((int)blockIdx.x >= q.nbands) return;  // uniform
  hpsi_small_body<T, N>((int)blockIdx.x, q.nbands, (const V2*)q.in, (V2*)q.out, (const T*)q.kin, (const T*)q.vr,
                        q.g2fft, npw, nbk, nch, (const V2*)q.B, (const V2*)q.DBH, inv_n, nullptr, nullptr, q.nkv,
                        nullptr, q.skip);
}

// Band density on small cubic grids: workgroup g takes bands [g·gsize, (g+1)·gsize), runs the
// inverse transform of each in LDS and keeps Σ_b fw_b |ψ_b(r)|² of its x-lines in FP64
// registers (bands with fw < 1e-15 skipped, src/scf/density.rs:14,72); the partial sums of
// the groups are added in group order by k_density_reduce, so the result does not depend on
// scheduling. A group whose bands are all skipped stores nothing: its partial is zero and the
// reduction passes over it (adding +0 to a sum that starts at +0 and adds no negative term
// changes nothing).
__device__ __forceinline__ bool density_group_empty(const double* __restrict__ fw, int g, int gsize, int nb) {
  const int b0 = g * gsize, b1 = min(nb, b0 + gsize);
  for (int b = b0; b < b1; ++b)
    if (!(fw[b] < 1e-15)) return false;
  return true;
}

template <typename T, int N, typename TIn>
__global__ void __launch_bounds__(SMALL_NT) k_density_small(const TIn* __restrict__ psi,
                                                            const int* __restrict__ g2fft, int npw, int nb, int gsize,
                                                            const double* __restrict__ fw,
                                                            double* __restrict__ partial) {
  using V2 = typename V2s<T>::t;
  extern __shared__ unsigned char smem_raw[];
  V2* G = reinterpret_cast<V2*>(smem_raw);
  constexpr int LPT = (N * N + SMALL_NT - 1) / SMALL_NT;  // x-lines per thread
  double acc[LPT][N];
#pragma unroll
  for (int li = 0; li < LPT; ++li)
#pragma unroll
    for (int x = 0; x < N; ++x) acc[li][x] = 0.0;
  const int b0 = blockIdx.x * gsize, b1 = min(nb, b0 + gsize);
  if (density_group_empty(fw, (int)blockIdx.x, gsize, nb)) return;  // uniform
  const LineAct act = small_line_activity<N>(G, g2fft, npw);
  for (int b = b0; b < b1; ++b) {
    const double f = fw[b];
    if (f < 1e-15) continue;  // uniform over the workgroup
    __syncthreads();
    small_clear<T, N>(G, act);
    __syncthreads();
    const TIn* p = psi + (size_t)b * npw;
    for (int g = threadIdx.x; g < npw; g += SMALL_NT) {  // FP64 input is rounded to T here
      const TIn v = p[g];
      V2 w;
      w.x = (T)v.x;
      w.y = (T)v.y;
      G[g2fft[g]] = w;
    }
    __syncthreads();
    small_zpass<T, N, 1>(G, act);
    __syncthreads();
    small_ypass<T, N, 1>(G, act);
    __syncthreads();
#pragma unroll
    for (int li = 0; li < LPT; ++li) {
      const int l = threadIdx.x + li * SMALL_NT;  // l = y·N + z
      if (l < N * N) {
        T re[N], im[N];
#pragma unroll
        for (int j = 0; j < N; ++j) {
          re[j] = G[l + j * N * N].x;
          im[j] = G[l + j * N * N].y;
        }
        fft_dif<T, N, 1>(re, im);
#pragma unroll
        for (int x = 0; x < N; ++x) {
          const double vr = (double)re[FftOrder<N>::pos(x)], vi = (double)im[FftOrder<N>::pos(x)];
          acc[li][x] += f * (vr * vr + vi * vi);
        }
      }
    }
  }
  double* out = partial + (size_t)blockIdx.x * N * N * N;
#pragma unroll
  for (int li = 0; li < LPT; ++li) {
    const int l = threadIdx.x + li * SMALL_NT;
    if (l < N * N) {
#pragma unroll
      for (int x = 0; x < N; ++x) out[(size_t)x * N * N + l] = acc[li][x];
    }
  }
}

constexpr int DRED_NT = 128;  // threads of a density reduction block; it takes up to 128 groups
// wave_sums (optional): the 64-lane shuffle-tree sum of ρ over each wave's points (0 past n), one
// value per wave in point order; a 256-point block reduction that adds its four waves' tree sums in
// wave order is rebuilt from them exactly.
__global__ void __launch_bounds__(DRED_NT) k_density_reduce(size_t n, int ngroups, int gsize, int nb,
                                                            const double* __restrict__ fw,
                                                            const double* __restrict__ partial,
                                                            double* __restrict__ rho, int overwrite,
                                                            double* __restrict__ wave_sums) {
  // the groups with a weighted band, in group order: thread t judges group t and the ballots of the
  // two waves give each live group its slot; the slots past the last live group name group 0, so
  // the loads below need no guard
  __shared__ int live[DRED_NT];
  __shared__ int wcount[DRED_NT / 64];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const bool on = t < ngroups && !density_group_empty(fw, t, gsize, nb);
  const unsigned long long bal = __ballot(on);
  if (lane == 0) wcount[w] = __popcll(bal);
  __syncthreads();
  const int m = wcount[0] + wcount[1];
  if (on) live[(w ? wcount[0] : 0) + __popcll(bal & ((1ull << lane) - 1ull))] = t;
  if (t >= m) live[t] = 0;
  __syncthreads();
  const size_t i = (size_t)blockIdx.x * DRED_NT + threadIdx.x;
  double s = 0.0;
  if (i < n) {
    s = overwrite ? 0.0 : rho[i];
    // every live partial of the point is loaded before they are added, in group order
    constexpr int U = 32;
    for (int g0 = 0; g0 < m; g0 += U) {
      double v[U];
#pragma unroll
      for (int j = 0; j < U; ++j) v[j] = partial[(size_t)live[min(g0 + j, DRED_NT - 1)] * n + i];
#pragma unroll
      for (int j = 0; j < U; ++j)
        if (g0 + j < m) s += v[j];
    }
    rho[i] = s;
  }
  if (wave_sums) {  // uniform
    double v = s;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) wave_sums[(size_t)blockIdx.x * (DRED_NT / 64) + w] = v;
  }
}

template <typename T>
size_t small_lds_bytes(int n, int nch) {
  return ((size_t)n * n * n + nl_lds_entries(nch)) * sizeof(typename V2s<T>::t);
}

constexpr size_t kLdsMax = 160 * 1024;

#define PWK_SMALL_SIZES(X) X(8) X(9) X(10) X(12) X(15) X(16) X(18) X(20) X(24)

bool small_len(int n) {
  switch (n) {
#define PWK_CASE(L) case L:
    PWK_SMALL_SIZES(PWK_CASE)
#undef PWK_CASE
    return true;
    default:
      return false;
  }
}

template <typename T>
void launch_small_t(Context& c, int nbands, int nbk, const void* psi, void* hpsi, const void* kin, int nch,
                    const void* B, const void* DBH, const double* coef, const void* X, int nkv, const int* band_k,
                    const int* skip) {
  using V2 = typename V2s<T>::t;
  const int n = c.dims[0];
  const size_t sh = small_lds_bytes<T>(n, nch);
  const T* vr = std::is_same<T, double>::value ? (const T*)c.vrn64.p : (const T*)c.vrn32.p;
  const T inv_n = (T)(1.0 / (double)c.N);
  TimedScope ts(c, "hpsi_small");
  switch (n) {
#define PWK_CASE(L)                                                                                            \
  case L:                                                                                                      \
    ensure_lds_limit((const void*)k_hpsi_small<T, L>, sh);                                                    \
    k_hpsi_small<T, L><<<(unsigned)nbands, SMALL_NT, sh, c.stream>>>(                                          \
        (const V2*)psi, (V2*)hpsi, (const T*)kin, vr, c.g2fft.as<int>(), c.npw, nbk, nch, (const V2*)B,        \
        (const V2*)DBH, inv_n, coef, (const V2*)X, nkv, band_k, skip);                                         \
    break;
    PWK_SMALL_SIZES(PWK_CASE)
#undef PWK_CASE
    default:
      throw Error(PWK_E_UNSUPPORTED, "grid has no fused small-grid kernel");
  }
  PWK_HIP(hipGetLastError());
}

constexpr int kMaxFilterDegree = 64;  // 3 coefficients per step, one thread each

template <typename T>
size_t filter_lds_bytes(int n, int nch, int npw) {
  return ((size_t)n * n * n + nl_lds_entries(nch) + 2 * (size_t)npw) * sizeof(typename V2s<T>::t) +
         3 * kMaxFilterDegree * sizeof(double);
}

template <typename T>
void launch_filter_small_t(Context& c, int nbands, int nbk, const void* V, void* Yout, const void* kin, int nch,
                           const void* B, const void* DBH, const double* coef, int degree, int nkv, const int* skip) {
  using V2 = typename V2s<T>::t;
  const int n = c.dims[0];
  const size_t sh = filter_lds_bytes<T>(n, nch, c.npw);
  const T* vr = std::is_same<T, double>::value ? (const T*)c.vrn64.p : (const T*)c.vrn32.p;
  const T inv_n = (T)(1.0 / (double)c.N);
  TimedScope ts(c, std::is_same<T, double>::value ? "cheb_filter_small_fp64" : "cheb_filter_small_fp32");
  switch (n) {
#define PWK_CASE(L)                                                                                            \
  case L:                                                                                                      \
    ensure_lds_limit((const void*)k_cheb_filter_small<T, L>, sh);                                             \
    k_cheb_filter_small<T, L><<<(unsigned)nbands, SMALL_NT, sh, c.stream>>>(                                   \
        (const double2*)V, (double2*)Yout, (const T*)kin, vr, c.g2fft.as<int>(), c.npw, nbk, nch, (const V2*)B, \
        (const V2*)DBH, inv_n, coef, degree, nkv, skip);                                                       \
    break;
    PWK_SMALL_SIZES(PWK_CASE)
#undef PWK_CASE
    default:
      throw Error(PWK_E_UNSUPPORTED, "grid has no fused small-grid kernel");
  }
  PWK_HIP(hipGetLastError());
}

template <typename T, typename TIn>
void launch_density_small_t(Context& c, int nb, const void* psi, const double* fw_dev, void* rho, bool overwrite,
                            double* wave_sums) {
  const int n = c.dims[0];
  const size_t sh = small_lds_bytes<T>(n, 0);
  const int gsize = std::max(1, (nb + 127) / 128);  // at most 128 groups, as the reduction takes
  const int ng = (nb + gsize - 1) / gsize;
  c.work_a.ensure((size_t)ng * c.N * sizeof(double));
  TimedScope ts(c, "density_small");
  switch (n) {
#define PWK_CASE(L)                                                                                            \
  case L:                                                                                                      \
    ensure_lds_limit((const void*)k_density_small<T, L, TIn>, sh);                                            \
    k_density_small<T, L, TIn><<<(unsigned)ng, SMALL_NT, sh, c.stream>>>((const TIn*)psi, c.g2fft.as<int>(),  \
                                                                         c.npw,                             \
                                                                    nb, gsize, fw_dev, c.work_a.as<double>());  \
    break;
    PWK_SMALL_SIZES(PWK_CASE)
#undef PWK_CASE
    default:
      throw Error(PWK_E_UNSUPPORTED, "grid has no fused small-grid kernel");
  }
  PWK_HIP(hipGetLastError());
  k_density_reduce<<<blocks_for(c.N, DRED_NT), DRED_NT, 0, c.stream>>>(c.N, ng, gsize, nb, fw_dev,
                                                                       c.work_a.as<double>(), (double*)rho,
                                                                       overwrite ? 1 : 0, wave_sums);
  PWK_HIP(hipGetLastError());
}

}  // namespace

void launch_density_small(Context& c, int nb, const void* psi, const double* fw_dev, void* rho, int precision,
                          bool overwrite, bool psi_fp64, double* wave_sums) {
  c.check_basis();
  PWK_REQUIRE(nb > 0, PWK_E_INVALID, "nb must be positive");
  PWK_REQUIRE(density_small_supported(c, precision), PWK_E_UNSUPPORTED, "grid too large for the fused kernel");
  if (precision == PWK_FP64)
    launch_density_small_t<double, double2>(c, nb, psi, fw_dev, rho, overwrite, wave_sums);
  else if (psi_fp64)
    launch_density_small_t<float, double2>(c, nb, psi, fw_dev, rho, overwrite, wave_sums);
  else
    launch_density_small_t<float, float2>(c, nb, psi, fw_dev, rho, overwrite, wave_sums);
}

size_t density_wave_sums(const Context& c) { return ((c.N + DRED_NT - 1) / DRED_NT) * (DRED_NT / 64); }

bool density_small_supported(const Context& c, int precision) {
  if (!c.small_hpsi || c.dims[0] != c.dims[1] || c.dims[1] != c.dims[2] || !small_len(c.dims[0])) return false;
  const size_t sh =
      precision == PWK_FP64 ? small_lds_bytes<double>(c.dims[0], 0) : small_lds_bytes<float>(c.dims[0], 0);
  return sh <= kLdsMax;
}

bool hpsi_small_supported(const Context& c, int precision, int nch) {
  if (!density_small_supported(c, precision)) return false;
  if (c.vrn64.p == nullptr || c.vrn32.p == nullptr) return false;
  const size_t sh = precision == PWK_FP64 ? small_lds_bytes<double>(c.dims[0], nch)
                                          : small_lds_bytes<float>(c.dims[0], nch);
  return sh <= kLdsMax;
}

bool cheb_filter_small_supported(const Context& c, int precision, int nch) {
  if (precision == PWK_SPLIT_BF16 || !hpsi_small_supported(c, precision, nch)) return false;
  const size_t sh = precision == PWK_FP64 ? filter_lds_bytes<double>(c.dims[0], nch, c.npw)
                                          : filter_lds_bytes<float>(c.dims[0], nch, c.npw);
  return sh <= kLdsMax;
}

void launch_cheb_filter_small(Context& c, int nbands, int nbk, const void* V, void* Yout, int precision,
                              const void* kin, int nch, const void* B, const void* DBH, const double* coef, int degree,
                              int nkv, const int* skip) {
  c.check_basis();
  PWK_REQUIRE(c.have_v, PWK_E_STATE, "V_eff not set (pwk_set_veff)");
  PWK_REQUIRE(nbands > 0 && nbk > 0 && nbands % nbk == 0, PWK_E_INVALID, "band counts must be positive");
  PWK_REQUIRE(degree >= 1 && degree <= kMaxFilterDegree, PWK_E_INVALID, "filter degree must be in 1..64");
  PWK_REQUIRE(cheb_filter_small_supported(c, precision, nch), PWK_E_UNSUPPORTED,
              "grid too large for the fused filter kernel");
  if (nkv > 0) {
    const size_t slots = ((size_t)(nbands / nbk) + nkv - 1) / nkv;
    PWK_REQUIRE(c.vrn64.bytes >= slots * c.N * sizeof(double), PWK_E_STATE, "V_eff slots not set");
  }
  if (precision == PWK_FP64)
    launch_filter_small_t<double>(c, nbands, nbk, V, Yout, kin, nch, B, DBH, coef, degree, nkv, skip);
  else
    launch_filter_small_t<float>(c, nbands, nbk, V, Yout, kin, nch, B, DBH, coef, degree, nkv, skip);
}

namespace {

template <typename T>
void launch_filter_multi_t(hipStream_t stream, const PassShape& sh, const PassSmallMulti& m, int n, int max_bands) {
  const size_t lds = filter_lds_bytes<T>(sh.n, sh.nch, sh.npw);
  const T inv_n = (T)sh.inv_n;
  const dim3 grid((unsigned)max_bands, (unsigned)n);
  switch (sh.n) {
#define PWK_CASE(L)                                                                                           \
  case L:                                                                                                     \
    ensure_lds_limit((const void*)k_cheb_filter_multi<T, L>, lds);                                           \
    k_cheb_filter_multi<T, L><<<grid, SMALL_NT, lds, stream>>>(m, sh.npw, sh.nbs, sh.nch, inv_n, sh.degree); \
    break;
    PWK_SMALL_SIZES(PWK_CASE)
#undef PWK_CASE
    default:
      throw Error(PWK_E_UNSUPPORTED, "grid has no fused small-grid kernel");
  }
  PWK_HIP(hipGetLastError());
}

}  // namespace

void launch_filter_multi(hipStream_t stream, const PassShape& sh, const PassSmallMulti& m, int n, int max_bands) {
  PWK_REQUIRE(n >= 1 && n <= MAX_PASS_BATCH && max_bands >= 1 && sh.degree >= 1 && sh.degree <= kMaxFilterDegree &&
                  small_len(sh.n),
              PWK_E_INVALID, "bad batched filter shape");
  const size_t lds = sh.fp == PWK_FP64 ? filter_lds_bytes<double>(sh.n, sh.nch, sh.npw)
                                       : filter_lds_bytes<float>(sh.n, sh.nch, sh.npw);
  PWK_REQUIRE(lds <= kLdsMax, PWK_E_UNSUPPORTED, "grid too large for the fused filter kernel");
  if (sh.fp == PWK_FP64)
    launch_filter_multi_t<double>(stream, sh, m, n, max_bands);
  else
    launch_filter_multi_t<float>(stream, sh, m, n, max_bands);
}

void launch_hpsi64_multi(hipStream_t stream, const PassShape& sh, const PassSmallMulti& m, int n, int max_bands) {
  PWK_REQUIRE(n >= 1 && n <= MAX_PASS_BATCH && max_bands >= 1 && small_len(sh.n), PWK_E_INVALID,
              "bad batched H·psi shape");
  const size_t lds = small_lds_bytes<double>(sh.n, sh.nch);
  PWK_REQUIRE(lds <= kLdsMax, PWK_E_UNSUPPORTED, "grid too large for the fused kernel");
  const dim3 grid((unsigned)max_bands, (unsigned)n);
  switch (sh.n) {
#define PWK_CASE(L)                                                                                       \
  case L:                                                                                                 \
    ensure_lds_limit((const void*)k_hpsi_multi<double, L>, lds);                                         \
    k_hpsi_multi<double, L><<<grid, SMALL_NT, lds, stream>>>(m, sh.npw, sh.nbs, sh.nch, sh.inv_n);       \
    break;
    PWK_SMALL_SIZES(PWK_CASE)
#undef PWK_CASE
    default:
      throw Error(PWK_E_UNSUPPORTED, "grid has no fused small-grid kernel");
  }
  PWK_HIP(hipGetLastError());
}

size_t lanczos_small_lds_bytes(int n, int nch, int npw) {
  return ((size_t)n * n * n + nl_lds_entries(nch) + 3 * (size_t)npw) * sizeof(double2);
}

bool lanczos_small_supported(const Context& c, int nch) {
  return hpsi_small_supported(c, PWK_FP64, nch) && lanczos_small_lds_bytes(c.dims[0], nch, c.npw) <= kLdsMax;
}

void launch_lanczos_small(Context& c, int nke, int nkv, int steps, const double* kin, int nch, const void* B,
                          const void* DBH, double* ab) {
  c.check_basis();
  PWK_REQUIRE(c.have_v, PWK_E_STATE, "V_eff not set (pwk_set_veff)");
  PWK_REQUIRE(nke > 0 && steps >= 1, PWK_E_INVALID, "bad Lanczos shape");
  PWK_REQUIRE(lanczos_small_supported(c, nch), PWK_E_UNSUPPORTED, "grid too large for the fused Lanczos kernel");
  if (nkv > 0) {
    const size_t slots = ((size_t)nke + nkv - 1) / nkv;
    PWK_REQUIRE(c.vrn64.bytes >= slots * c.N * sizeof(double), PWK_E_STATE, "V_eff slots not set");
  }
  const size_t sh = lanczos_small_lds_bytes(c.dims[0], nch, c.npw);
  TimedScope ts(c, "lanczos_small");
  switch (c.dims[0]) {
#define PWK_CASE(L)                                                                                              \
  case L:                                                                                                        \
    ensure_lds_limit((const void*)k_lanczos_small<L>, sh);                                                      \
    k_lanczos_small<L><<<(unsigned)nke, SMALL_NT, sh, c.stream>>>(c.g2fft.as<int>(), c.npw, nch, kin,           \
                                                                c.vrn64.as<double>(), (const double2*)B,        \
                                                                (const double2*)DBH, 1.0 / (double)c.N, nkv,   \
                                                                steps, ab);                                     \
    break;
    PWK_SMALL_SIZES(PWK_CASE)
#undef PWK_CASE
    default:
      throw Error(PWK_E_UNSUPPORTED, "grid has no fused small-grid kernel");
  }
  PWK_HIP(hipGetLastError());
}

void launch_hpsi_small(Context& c, int nbands, int nbk, const void* psi, void* hpsi, int precision, const void* kin,
                       int nch, const void* B, const void* DBH, const double* coef, const void* X, int nkv,
                       const int* band_k, const int* skip) {
  c.check_basis();
  PWK_REQUIRE(c.have_v, PWK_E_STATE, "V_eff not set (pwk_set_veff)");
  PWK_REQUIRE(nbands > 0 && nbk > 0 && nbands % nbk == 0, PWK_E_INVALID, "band counts must be positive");
  PWK_REQUIRE(hpsi_small_supported(c, precision, nch), PWK_E_UNSUPPORTED, "grid too large for the fused kernel");
  if (nkv > 0) {
    const size_t slots = ((size_t)(nbands / nbk) + nkv - 1) / nkv;
    PWK_REQUIRE(c.vrn64.bytes >= slots * c.N * sizeof(double), PWK_E_STATE, "V_eff slots not set");
  }
  PWK_REQUIRE(!band_k || (!coef && nkv == 0), PWK_E_INVALID, "a band table excludes filter coefficients");
  if (precision == PWK_FP64)
    launch_small_t<double>(c, nbands, nbk, psi, hpsi, kin, nch, B, DBH, coef, X, nkv, band_k, skip);
  else
    launch_small_t<float>(c, nbands, nbk, psi, hpsi, kin, nch, B, DBH, coef, X, nkv, band_k, skip);
}

}  // namespace pwk
