// Line passes of a 3D FFT held in LDS (grid G[x][y][z], z fastest), shared by the fused
// small-grid kernels: one lane per line, the line in registers (fft_reg.hpp).
#pragma once

#include "fft_reg.hpp"

namespace pwk {

template <typename T>
struct V2s;
template <>
struct V2s<double> {
  using t = double2;
};
template <>
struct V2s<float> {
  using t = float2;
};

constexpr int SMALL_NT = 256;

// Workgroup barrier that orders LDS only: it waits for this wave's LDS operations, not for its
// outstanding global loads, so operands prefetched into registers stay in flight across it.
// Sufficient wherever the workgroup exchanges data through LDS alone.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// A pass of one workgroup over the NN points of a grid, thread t taking points t + q·SMALL_NT:
// load(i) for up to sixteen of the thread's points is issued before use(i, value) runs on any of
// them, so the loads (global memory written by earlier launches, often on another XCD) are in
// flight together instead of one round trip per point.
template <int NN, typename V, typename Load, typename Use>
__device__ __forceinline__ void grid_points(Load load, Use use) {
  constexpr int PER = (NN + SMALL_NT - 1) / SMALL_NT, CH = PER < 16 ? PER : 16;
#pragma unroll
  for (int q0 = 0; q0 < PER; q0 += CH) {
    V v[CH];
#pragma unroll
    for (int q = 0; q < CH; ++q) {
      const int i = threadIdx.x + (q0 + q) * SMALL_NT;
      if (q0 + q < PER && i < NN) v[q] = load(i);
    }
#pragma unroll
    for (int q = 0; q < CH; ++q) {
      const int i = threadIdx.x + (q0 + q) * SMALL_NT;
      if (q0 + q < PER && i < NN) use(i, v[q]);
    }
  }
}

// One set of N² lines through the LDS grid G[x][y][z] (z fastest). AXIS 2: z-lines
// (stride 1), 1: y-lines (stride N), 0: x-lines (stride N²). With vr, the x pass runs
// inverse -> × V(r) -> forward on each line while it is in registers.
template <typename T, int N, int AXIS, int SIGN>
__device__ __forceinline__ void lds_lines(typename V2s<T>::t* G, const T* __restrict__ vr) {
  constexpr int es = AXIS == 2 ? 1 : AXIS == 1 ? N : N * N;
  for (int l = threadIdx.x; l < N * N; l += SMALL_NT) {
    int base;
    if constexpr (AXIS == 2)
      base = l * N;  // l = x·N + y
    else if constexpr (AXIS == 1)
      base = (l / N) * N * N + (l % N);  // l = x·N + z
    else
      base = l;  // l = y·N + z
    T re[N], im[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      re[j] = G[base + j * es].x;
      im[j] = G[base + j * es].y;
    }
    fft_dif<T, N, SIGN>(re, im);
    if constexpr (AXIS == 0) {
      if (vr) {
        T r2[N], m2[N];
#pragma unroll
        for (int f = 0; f < N; ++f) {
          const T v = vr[base + f * es];
          r2[f] = re[FftOrder<N>::pos(f)] * v;
          m2[f] = im[FftOrder<N>::pos(f)] * v;
        }
        fft_dif<T, N, -SIGN>(r2, m2);
#pragma unroll
        for (int f = 0; f < N; ++f) {
          G[base + f * es].x = r2[FftOrder<N>::pos(f)];
          G[base + f * es].y = m2[FftOrder<N>::pos(f)];
        }
        continue;
      }
    }
#pragma unroll
    for (int f = 0; f < N; ++f) {
      G[base + f * es].x = re[FftOrder<N>::pos(f)];
      G[base + f * es].y = im[FftOrder<N>::pos(f)];
    }
  }
}

// One pass of lds_lines on a grid with one line per thread (N² <= SMALL_NT), pruned to the lines
// that matter: an inactive line is left alone (ZERO = false: its result is never read) or filled
// with zeros (ZERO = true: its input is zero, so its transform is zero). Active lines run the
// transform of lds_lines exactly.
template <typename T, int N, int AXIS, int SIGN, bool ZERO>
__device__ __forceinline__ void lds_lines_pruned(typename V2s<T>::t* G, bool active) {
  static_assert(N * N <= SMALL_NT, "one line per thread");
  static_assert(AXIS == 1 || AXIS == 2, "y and z passes");
  constexpr int es = AXIS == 2 ? 1 : N;
  const int l = threadIdx.x;
  if (l >= N * N) return;
  const int base = AXIS == 2 ? l * N : (l / N) * N * N + (l % N);
  if (!active) {
    if constexpr (ZERO) {
#pragma unroll
      for (int f = 0; f < N; ++f) {
        G[base + f * es].x = 0;
        G[base + f * es].y = 0;
      }
    }
    return;
  }
  T re[N], im[N];
#pragma unroll
  for (int j = 0; j < N; ++j) {
    re[j] = G[base + j * es].x;
    im[j] = G[base + j * es].y;
  }
  fft_dif<T, N, SIGN>(re, im);
#pragma unroll
  for (int f = 0; f < N; ++f) {
    G[base + f * es].x = re[FftOrder<N>::pos(f)];
    G[base + f * es].y = im[FftOrder<N>::pos(f)];
  }
}

// The x pass of lds_lines with V(r) of the thread's line already in registers (vline[f] =
// V at x = f of line threadIdx.x); for grids with at most one x-line per thread.
template <typename T, int N, int SIGN>
__device__ __forceinline__ void lds_xlines_v(typename V2s<T>::t* G, const T (&vline)[N]) {
  static_assert(N * N <= SMALL_NT, "one x-line per thread");
  constexpr int es = N * N;
  const int l = threadIdx.x;
  if (l >= N * N) return;
  T re[N], im[N];
#pragma unroll
  for (int j = 0; j < N; ++j) {
    re[j] = G[l + j * es].x;
    im[j] = G[l + j * es].y;
  }
  fft_dif<T, N, SIGN>(re, im);
  T r2[N], m2[N];
#pragma unroll
  for (int f = 0; f < N; ++f) {
    r2[f] = re[FftOrder<N>::pos(f)] * vline[f];
    m2[f] = im[FftOrder<N>::pos(f)] * vline[f];
  }
  fft_dif<T, N, -SIGN>(r2, m2);
#pragma unroll
  for (int f = 0; f < N; ++f) {
    G[l + f * es].x = r2[FftOrder<N>::pos(f)];
    G[l + f * es].y = m2[FftOrder<N>::pos(f)];
  }
}

}  // namespace pwk
