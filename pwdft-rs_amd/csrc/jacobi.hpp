// Cyclic complex Jacobi eigensolver for small Hermitian matrices held in LDS,
// shared by the batched eigensolver (n <= 32) and the fused Rayleigh-Ritz
// kernel. Every thread of the block calls it; the rotations of one
// round-robin round are independent, so the block applies them together.
#pragma once

#include <hip/hip_runtime.h>

namespace pwk {

constexpr int JMAX = 32;

// LDS of the solver for matrices up to JM × JM
template <int JM>
struct JacobiLdsT {
  double2 A[JM][JM + 1];  // matrix (row, column); one padding column against bank conflicts
  double2 V[JM][JM + 1];  // accumulated rotations
  double rc[JM / 2], rs[JM / 2];
  double2 rph[JM / 2];
  int rp[JM / 2], rq[JM / 2], order[JM];
  double red[32];  // per-wave partial sums (off, total) for up to 16 waves
};
using JacobiLds = JacobiLdsT<JMAX>;

// Block-wide sum of (a, b) over all threads, in wave order (same result on every thread).
__device__ inline void jacobi_sum2(double& a, double& b, double* red) {
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o, 64);
    b += __shfl_xor(b, o, 64);
  }
  const int w = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
  if ((threadIdx.x & 63) == 0) {
    red[2 * w] = a;
    red[2 * w + 1] = b;
  }
  __syncthreads();
  a = 0.0;
  b = 0.0;
  for (int i = 0; i < nw; ++i) {
    a += red[2 * i];
    b += red[2 * i + 1];
  }
  __syncthreads();
}

// Diagonalises J.A (n <= JMAX) in place: on return the diagonal of J.A holds the eigenvalues,
// J.V = V_in · (product of the rotations), J.order[r] the index of the r-th smallest eigenvalue
// (ties by index). Sweeps stop when the off-diagonal weight falls below 1e-30 of the total.
template <int JM>
__device__ inline void jacobi_lds(int n, JacobiLdsT<JM>& J) {
  const int t = threadIdx.x, nt = blockDim.x;
  const int N = n + (n & 1);
  for (int sweep = 0; sweep < 30; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int e = t; e < n * n; e += nt) {
      const int r = e % n, col = e / n;
      const double2 a = J.A[r][col];
      const double m2 = a.x * a.x + a.y * a.y;
      tot += m2;
      if (r != col) off += m2;
    }
    jacobi_sum2(off, tot, J.red);
    if (off <= 1e-30 * tot || off < 1e-300) break;  // block-uniform
    for (int round = 0; round < N - 1; ++round) {
      // round-robin pairing: position 0 stays, positions 1..N-1 have rotated `round` times
      if (t < N / 2) {
        const int u = N - 1 - t;
        // (t - 1 - round) and (u - 1 - round) lie in (-(N - 1), N - 1): one wrap is enough
        int p = 0;
        if (t != 0) {
          p = t - 1 - round;
          p += p < 0 ? N - 1 : 0;
          p += 1;
        }
        int q = u - 1 - round;
        q += q < 0 ? N - 1 : 0;
        q += 1;
        if (p > q) {
          const int tt = p;
          p = q;
          q = tt;
        }
        double c = 1.0, sn = 0.0;
        double2 ph = make_double2(1.0, 0.0);
        if (q < n) {
          // with d = (A_qq - A_pp)/2, g = |A_pq|: tan θ = sign(d)·g/(|d| + sqrt(d² + g²)), so
          // c = den/sqrt(den² + g²), s = sign(d)·g/sqrt(den² + g²) with den = |d| + sqrt(d² + g²),
          // and the phase A_pq/g: one sqrt and two reciprocal square roots on the chain
          const double a = J.A[p][p].x, bq = J.A[q][q].x;
          const double2 apq = J.A[p][q];
          const double g2 = apq.x * apq.x + apq.y * apq.y;
          if (g2 > 1e-300) {  // |A_pq| above 1e-150: no denormal reaches rsqrt
            const double ig = rsqrt(g2);
            ph = make_double2(apq.x * ig, apq.y * ig);
            const double d = 0.5 * (bq - a);
            const double den = fabs(d) + sqrt(d * d + g2);
            const double inv = rsqrt(den * den + g2);
            c = den * inv;
            sn = (d >= 0.0 ? 1.0 : -1.0) * (g2 * ig) * inv;
          }
        }
        J.rp[t] = p;
        J.rq[t] = q < n ? q : -1;
        J.rc[t] = c;
        J.rs[t] = sn;
        J.rph[t] = ph;
      }
      __syncthreads();
      // A ← Wᴴ·A·W block by block: the entries in the rows of pair kr and the columns of pair kc
      // change only through rotations kr and kc, so each (kr, kc) block is updated by one thread
      // (the column rotation, then the row rotation, as a column pass followed by a row pass
      // would); V ← V·W column pairs alongside. W = [[c, s], [-s e, c e]], e = conj(ph).
      const int NP = N / 2;
      for (int it = t; it < NP * NP + NP * n; it += nt) {
        if (it < NP * NP) {
          const int kr = it / NP, kc = it % NP;
          const int pr = J.rp[kr], qr = J.rq[kr], pc = J.rp[kc], qc = J.rq[kc];
          double2 a00 = J.A[pr][pc], a01 = qc >= 0 ? J.A[pr][qc] : make_double2(0.0, 0.0);
          double2 a10 = qr >= 0 ? J.A[qr][pc] : make_double2(0.0, 0.0);
          double2 a11 = (qr >= 0 && qc >= 0) ? J.A[qr][qc] : make_double2(0.0, 0.0);
          if (qc >= 0) {  // columns (pc, qc)
            const double c = J.rc[kc], sn = J.rs[kc];
            const double2 e = make_double2(J.rph[kc].x, -J.rph[kc].y);
            double2 ey = make_double2(e.x * a01.x - e.y * a01.y, e.x * a01.y + e.y * a01.x);
            const double2 n00 = make_double2(c * a00.x - sn * ey.x, c * a00.y - sn * ey.y);
            const double2 n01 = make_double2(sn * a00.x + c * ey.x, sn * a00.y + c * ey.y);
            a00 = n00;
            a01 = n01;
            if (qr >= 0) {
              ey = make_double2(e.x * a11.x - e.y * a11.y, e.x * a11.y + e.y * a11.x);
              const double2 n10 = make_double2(c * a10.x - sn * ey.x, c * a10.y - sn * ey.y);
              const double2 n11 = make_double2(sn * a10.x + c * ey.x, sn * a10.y + c * ey.y);
              a10 = n10;
              a11 = n11;
            }
          }
          if (qr >= 0) {  // rows (pr, qr), Wᴴ = [[c, -s conj(e)], [s, c conj(e)]]
            const double c = J.rc[kr], sn = J.rs[kr];
            const double2 ce = J.rph[kr];
            double2 ey = make_double2(ce.x * a10.x - ce.y * a10.y, ce.x * a10.y + ce.y * a10.x);
            const double2 n00 = make_double2(c * a00.x - sn * ey.x, c * a00.y - sn * ey.y);
            const double2 n10 = make_double2(sn * a00.x + c * ey.x, sn * a00.y + c * ey.y);
            a00 = n00;
            a10 = n10;
            if (qc >= 0) {
              ey = make_double2(ce.x * a11.x - ce.y * a11.y, ce.x * a11.y + ce.y * a11.x);
              const double2 n01 = make_double2(c * a01.x - sn * ey.x, c * a01.y - sn * ey.y);
              const double2 n11 = make_double2(sn * a01.x + c * ey.x, sn * a01.y + c * ey.y);
              a01 = n01;
              a11 = n11;
            }
          }
          J.A[pr][pc] = a00;
          if (qc >= 0) J.A[pr][qc] = a01;
          if (qr >= 0) J.A[qr][pc] = a10;
          if (qr >= 0 && qc >= 0) J.A[qr][qc] = a11;
        } else {
          const int rem = it - NP * NP, k = rem / n, r = rem % n, q = J.rq[k];
          if (q < 0) continue;
          const int p = J.rp[k];
          const double c = J.rc[k], sn = J.rs[k];
          const double2 e = make_double2(J.rph[k].x, -J.rph[k].y);
          const double2 x = J.V[r][p], y = J.V[r][q];
          const double2 ey = make_double2(e.x * y.x - e.y * y.y, e.x * y.y + e.y * y.x);
          J.V[r][p] = make_double2(c * x.x - sn * ey.x, c * x.y - sn * ey.y);
          J.V[r][q] = make_double2(sn * x.x + c * ey.x, sn * x.y + c * ey.y);
        }
      }
      __syncthreads();
    }
  }
  // ascending order by rank (ties by index). The order is total even for a NaN diagonal (NaNs
  // after every number, among themselves by index), so the ranks are always a permutation and
  // every entry of J.order is written.
  if (t < n) {
    const double di = J.A[t][t].x;
    const bool ni = di != di;
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      const double dj = J.A[j][j].x;
      const bool nj = dj != dj;
      const bool before = (ni || nj) ? (nj == ni ? j < t : ni) : (dj < di || (dj == di && j < t));
      rank += before ? 1 : 0;
    }
    J.order[rank] = t;
  }
  __syncthreads();
}

}  // namespace pwk
