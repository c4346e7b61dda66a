// Self-consistent field solve on the device: run_scf_unpolarized
// (src/scf/driver.rs:353-735) with the dense eigensolver path, Fermi-Dirac
// smearing, G-space density symmetrisation and plain Anderson mixing.
//
// Setup (src/scf/context.rs:102-231) runs here too: the basis and the integer
// tables on the host (tables.hip), the radial transforms V_loc(G), ρ_core and
// the superposition-of-atoms start density as one device kernel over the FFT
// grid, Ewald as device FP64 reductions, the KB projectors per k-point on the
// device (nonlocal.hip).
//
// Per iteration every grid-sized array stays in HBM: V_eff assembly (Hartree,
// LDA, r→G FFTs), dense H per k, the batched eigensolver over all k at once,
// density accumulation, symmetrisation, energies (deterministic two-stage
// reductions) and the Anderson update. The host sees only nk·nb eigenvalues
// (Fermi level) and a handful of reduction results per iteration.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <set>
#include <unordered_map>

#include "common.hpp"
#include "xc.hpp"
#include "tables.hpp"

namespace pwk {

using namespace tab;

namespace {

constexpr int MAXSP = 8;
constexpr int RED_NT = 256;
constexpr int RED_BLOCKS = 128;
constexpr int MAX_TERMS = 48;   // dot products per reduction launch
constexpr int MAX_HIST = 32;    // Anderson history entries (mixing_ndim)
constexpr int MAX_DOTS = MAX_HIST - 1 + (MAX_HIST - 1) * MAX_HIST / 2;  // terms of one Anderson step

// ---------------------------------------------------------------------------------------
// radial transforms on the grid: V_loc(G) (src/scf/potentials.rs:18-51 with
// v_local_of_g, src/pseudopotential/upf/mod.rs:166-213), ρ_core(G)
// (src/scf/potentials.rs:72-130) and the atomic start density
// (src/scf/initial_density.rs:64-166). The Simpson coefficients and every
// G-independent factor are folded into per-species weight tables on the host.
// ---------------------------------------------------------------------------------------
struct RadialArgs {
  size_t n;
  int dims[3];
  double b[9];  // reciprocal vectors, rows
  double omega, sad_sigma;
  int nspecies, natoms, nshell;
  int off[MAXSP], mesh[MAXSP];
  double z[MAXSP], v0[MAXSP];
  int has_core[MAXSP], has_rho[MAXSP];
  const double* r;
  const double* wv;  // Simpson weight · rab · r² · (v_loc + Z e² erf(r)/r)
  const double* wc;  // Simpson weight · rab · r² · ρ_core
  const double* wr;  // Simpson weight · rab · 4πr²ρ_atom
  const double* tau;  // natoms × 3, Cartesian Å
  const int* sp;
  const double* shell_g2;  // |G|² of each shell
  const int* shell;        // shell of each grid point
  double* tab;             // nshell × nspecies × (V_loc, ρ_core, ρ_atom) form factors
  double2* vloc;
  double2* core;
  double2* sad;
};

__device__ inline double sinc_d(double x) { return x < 1e-10 ? 1.0 - x * x / 6.0 : sin(x) / x; }

// One wave per grid point or shell: the lanes split the radial mesh (and the atoms) and the
// partial sums meet in a butterfly, so the whole device works on the mesh integrals.
__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// The radial form factors of every species at every distinct |G| (one wave per shell): the
// transforms depend on |G| only, so they are evaluated once per shell of equal |G|² instead of
// once per grid point (src/scf/potentials.rs:18-51 evaluates v_local_of_g per point).
__global__ void __launch_bounds__(256) k_radial_shell(RadialArgs a) {
  const int sh = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (sh >= a.nshell) return;  // wave-uniform
  const double g2 = a.shell_g2[sh], gn = sqrt(g2);
  const double fourpi = 4.0 * 3.141592653589793;
  for (int s = 0; s < a.nspecies; ++s) {
    const double* r = a.r + a.off[s];
    const double* wv = a.wv + a.off[s];
    const double* wc = a.wc + a.off[s];
    const double* wr = a.wr + a.off[s];
    double iv = 0.0, ic = 0.0, ir = 0.0;
    for (int i = lane; i < a.mesh[s]; i += 64) {
      const double j0 = sinc_d(gn * r[i]);
      iv += wv[i] * j0;
      ic += wc[i] * j0;
      ir += wr[i] * j0;
    }
    iv = wave_sum(iv);
    ic = wave_sum(ic);
    ir = wave_sum(ir);
    if (lane == 0) {
      double vg;
      if (gn < 1e-12) {
        vg = a.v0[s];
      } else {
        vg = fourpi / a.omega * (iv / 3.0) - fourpi * a.z[s] * 14.399645351950548 * exp(-g2 / 4.0) / (a.omega * g2);
      }
      const double cg = a.has_core[s] ? fourpi * (ic / 3.0) / a.omega : 0.0;
      const double sg = a.has_rho[s] ? (ir / 3.0) / a.omega
                                     : a.z[s] * exp(-0.5 * g2 * a.sad_sigma * a.sad_sigma) / a.omega;
      double* t = a.tab + ((size_t)sh * a.nspecies + s) * 3;
      t[0] = vg;
      t[1] = cg;
      t[2] = sg;
    }
  }
}

// V_loc(G), ρ_core(G) and the atomic start density (src/scf/initial_density.rs:64-166) on the
// grid: Σ_atoms e^{-iG·τ} × the form factor of the point's shell, one wave per grid point with
// the lanes over the atoms.
__global__ void __launch_bounds__(256) k_radial(RadialArgs a) {
  const size_t idx = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (idx >= a.n) return;  // wave-uniform
  const int nx = a.dims[0], ny = a.dims[1], nz = a.dims[2];
  const int i1 = (int)(idx / ((size_t)ny * nz)), i2 = (int)((idx / nz) % ny), i3 = (int)(idx % nz);
  const double n1 = i1 > nx / 2 ? i1 - nx : i1, n2 = i2 > ny / 2 ? i2 - ny : i2, n3 = i3 > nz / 2 ? i3 - nz : i3;
  const double gx = (n1 * a.b[0] + n2 * a.b[3]) + n3 * a.b[6];
  const double gy = (n1 * a.b[1] + n2 * a.b[4]) + n3 * a.b[7];
  const double gz = (n1 * a.b[2] + n2 * a.b[5]) + n3 * a.b[8];
  const double* tab = a.tab + (size_t)a.shell[idx] * a.nspecies * 3;
  double2 vl = make_double2(0.0, 0.0), co = vl, sa = vl;
  for (int at = lane; at < a.natoms; at += 64) {
    const double* t = tab + (size_t)a.sp[at] * 3;
    const double vg = t[0], cg = t[1], sg = t[2];
    const double ph = -(gx * a.tau[3 * at] + gy * a.tau[3 * at + 1] + gz * a.tau[3 * at + 2]);
    double sn, cs;
    sincos(ph, &sn, &cs);
    vl.x += cs * vg;
    vl.y += sn * vg;
    co.x += cs * cg;
    co.y += sn * cg;
    sa.x += cs * sg;
    sa.y += sn * sg;
  }
  vl = make_double2(wave_sum(vl.x), wave_sum(vl.y));
  co = make_double2(wave_sum(co.x), wave_sum(co.y));
  sa = make_double2(wave_sum(sa.x), wave_sum(sa.y));
  if (lane == 0) {
    a.vloc[idx] = vl;
    a.core[idx] = co;
    a.sad[idx] = sa;
  }
}

// Lanczos recurrence coefficients per k for the spectrum bound of the filtered eigensolver, on
// the device so the steps need no host round trip. stage 0: α_j = red[k] and the step coefficients
// (1, α_j, β_{j-1}) of w ← H v_j − α_j v_j − β_{j−1} v_{j−1}; stage 1: β_j = sqrt(red[k]) and
// 1/β_j (0 below 1e-12, where the recurrence has broken down).
__global__ void k_lanczos_coef(int nke, int j, int stage, const double* __restrict__ red, double* __restrict__ beta,
                               double* __restrict__ cf, double* __restrict__ al, double* __restrict__ be,
                               double* __restrict__ inv) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nke) return;
  if (stage == 0) {
    const double a = red[k];
    al[(size_t)j * nke + k] = a;
    cf[3 * k] = 1.0;
    cf[3 * k + 1] = a;
    cf[3 * k + 2] = beta[k];
  } else {
    const double b = sqrt(fmax(red[k], 0.0));
    be[(size_t)j * nke + k] = b;
    beta[k] = b;
    inv[k] = b < 1e-12 ? 0.0 : 1.0 / b;
  }
}

// ---------------------------------------------------------------------------------------
// Ewald sums (src/ewald.rs:43-143): per-block partial sums, summed on the host in
// block order.
// ---------------------------------------------------------------------------------------
struct EwaldArgs {
  int na;
  int nr[3], nl[3];
  double eta, eta2;
  double b[9], l[9];
  const double* pos;  // na × 3
  const double* q;
};

__device__ double block_reduce(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int t = threadIdx.x;
  if ((t & 63) == 0) sh[t >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (t == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sh[w];
  return s;  // valid on thread 0
}

__global__ void __launch_bounds__(256) k_ewald_recip(EwaldArgs a, double* partial) {
  __shared__ double sh[4];
  const long w1 = 2 * a.nr[0] + 1, w2 = 2 * a.nr[1] + 1, w3 = 2 * a.nr[2] + 1;
  const long total = w1 * w2 * w3;
  double acc = 0.0;
  for (long idx = (long)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (long)gridDim.x * 256) {
    const int i1 = (int)(idx / (w2 * w3)) - a.nr[0], i2 = (int)((idx / w3) % w2) - a.nr[1],
              i3 = (int)(idx % w3) - a.nr[2];
    if (i1 == 0 && i2 == 0 && i3 == 0) continue;
    const double gx = (i1 * a.b[0] + i2 * a.b[3]) + i3 * a.b[6];
    const double gy = (i1 * a.b[1] + i2 * a.b[4]) + i3 * a.b[7];
    const double gz = (i1 * a.b[2] + i2 * a.b[5]) + i3 * a.b[8];
    const double g2 = gx * gx + gy * gy + gz * gz;
    if (g2 < 1e-12) continue;
    double sr = 0.0, si = 0.0;
    for (int i = 0; i < a.na; ++i) {
      double sn, cs;
      sincos(gx * a.pos[3 * i] + gy * a.pos[3 * i + 1] + gz * a.pos[3 * i + 2], &sn, &cs);
      sr += a.q[i] * cs;
      si += a.q[i] * sn;
    }
    acc += (sr * sr + si * si) * exp(-g2 / (4.0 * a.eta2)) / g2;
  }
  const double s = block_reduce(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void __launch_bounds__(256) k_ewald_real(EwaldArgs a, double* partial) {
  __shared__ double sh[4];
  const long w1 = 2 * a.nl[0] + 1, w2 = 2 * a.nl[1] + 1, w3 = 2 * a.nl[2] + 1;
  const long na2 = (long)a.na * a.na;
  const long total = w1 * w2 * w3 * na2;
  double acc = 0.0;
  for (long idx = (long)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (long)gridDim.x * 256) {
    const long cell = idx / na2, pair = idx % na2;
    const int i = (int)(pair / a.na), j = (int)(pair % a.na);
    const int t1 = (int)(cell / (w2 * w3)) - a.nl[0], t2 = (int)((cell / w3) % w2) - a.nl[1],
              t3 = (int)(cell % w3) - a.nl[2];
    const double tx = (t1 * a.l[0] + t2 * a.l[3]) + t3 * a.l[6];
    const double ty = (t1 * a.l[1] + t2 * a.l[4]) + t3 * a.l[7];
    const double tz = (t1 * a.l[2] + t2 * a.l[5]) + t3 * a.l[8];
    const double rx = a.pos[3 * i] - a.pos[3 * j] + tx, ry = a.pos[3 * i + 1] - a.pos[3 * j + 1] + ty,
                 rz = a.pos[3 * i + 2] - a.pos[3 * j + 2] + tz;
    const double rn = sqrt(rx * rx + ry * ry + rz * rz);
    if (rn < 1e-10) continue;
    acc += a.q[i] * a.q[j] * erfc(a.eta * rn) / rn;
  }
  const double s = block_reduce(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// ---------------------------------------------------------------------------------------
// grid helpers
// ---------------------------------------------------------------------------------------
__global__ void k_c2r_scaled(size_t n, const double2* __restrict__ c, double s, double* __restrict__ r,
                             int clamp_nonneg) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = c[i].x * s;
  if (clamp_nonneg && v < 0.0) v = 0.0;
  r[i] = v;
}

// src/scf/energy.rs:314-324: ρ_xc = max(ρ + cs·ρ_core, 0) (no clamp without a core); cs is
// 1, or 1/2 for one spin channel (src/scf/driver_spin.rs:260)
__global__ void k_add_core(size_t n, const double* __restrict__ rho, const double* __restrict__ core, double cs,
                           double* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = core ? fmax(rho[i] + core[i] * cs, 0.0) : rho[i];
}

// ρ_total = ρ↑ + ρ↓ and m = ρ↑ − ρ↓ (m may be null)
__global__ void k_spin_combine(size_t n, const double* __restrict__ up, const double* __restrict__ dn,
                               double* __restrict__ tot, double* __restrict__ mag) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double u = up[i], d = dn[i];
  tot[i] = u + d;
  if (mag) mag[i] = u - d;
}

// ρ↑ = (ρ_total + m)/2, ρ↓ = (ρ_total − m)/2  (src/scf/driver_spin.rs:820-831)
__global__ void k_spin_split(size_t n, const double* __restrict__ tot, const double* __restrict__ mag,
                             double* __restrict__ up, double* __restrict__ dn) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double t = tot[i], m = mag[i];
  up[i] = 0.5 * (t + m);
  dn[i] = 0.5 * (t - m);
}

// start split ρ↑ = ρ(1 + m)/2, ρ↓ = ρ(1 − m)/2 (src/scf/driver_spin.rs:74-75); rho may alias up
__global__ void k_spin_start(size_t n, const double* rho, double mag, double* up, double* __restrict__ dn) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double r = rho[i];
  up[i] = r * (1.0 + mag) / 2.0;
  dn[i] = r * (1.0 - mag) / 2.0;
}

// Kerker weight |G|²/(|G|² + q_TF²), zero at G = 0 (src/scf/mixing/anderson.rs:64-76)
__global__ void k_kerker(size_t n, const double* __restrict__ g2, double q2, double2* __restrict__ x) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double g = g2[i];
  const double w = g < 1e-12 ? 0.0 : g / (g + q2);
  x[i].x *= w;
  x[i].y *= w;
}

__global__ void k_scale(size_t n, double* __restrict__ x, double s) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] *= s;
}

// Anderson history entry: the input and its residual out − in
__global__ void k_hist_store(size_t n, const double* __restrict__ in, const double* __restrict__ out,
                             double* __restrict__ hin, double* __restrict__ hres) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = in[i];
  hin[i] = x;
  hres[i] = out[i] - x;
}

// ρ_sym(G) = (1/N_ops) Σ_S e^{-2πi n·τ_S} ρ(R_S^T n)  (src/symmetry/density/g_space.rs:147-252)
// The source point and the phase of every (point, operation) pair depend only on the grid and
// the operations, so they are tabulated once per solve (k_sym_table, [point][operation]); the
// gather of an iteration reads them (k_sym_gather). Sixteen lanes per grid point evaluate the
// operations (lane l takes l, l + 16, ...) into LDS; one lane then adds the terms in operation
// order, as one thread adding them in sequence would (nops <= 64).
constexpr int SYM_MAX_OPS = 64;

__global__ void __launch_bounds__(256) k_sym_table(size_t n, int nx, int ny, int nz, int nops,
                                                   const int* __restrict__ rot, const double* __restrict__ tau,
                                                   int* __restrict__ src_out, double2* __restrict__ ph_out) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= n * (size_t)nops) return;
  const size_t idx = e / (size_t)nops;
  const int o = (int)(e % (size_t)nops);
  const int i1 = (int)(idx / ((size_t)ny * nz)), i2 = (int)((idx / nz) % ny), i3 = (int)(idx % nz);
  const int nd[3] = {i1 > nx / 2 ? i1 - nx : i1, i2 > ny / 2 ? i2 - ny : i2, i3 > nz / 2 ? i3 - nz : i3};
  const int dd[3] = {nx, ny, nz};
  const int* R = rot + 9 * o;
  int src[3];
  for (int c = 0; c < 3; ++c) {
    const int ns = R[c] * nd[0] + R[3 + c] * nd[1] + R[6 + c] * nd[2];
    src[c] = ((ns % dd[c]) + dd[c]) % dd[c];
  }
  src_out[e] = (src[0] * ny + src[1]) * nz + src[2];
  const double arg = 6.283185307179586 * ((double)nd[0] * tau[3 * o] + (double)nd[1] * tau[3 * o + 1] +
                                          (double)nd[2] * tau[3 * o + 2]);
  double sn, cs;
  sincos(arg, &sn, &cs);
  ph_out[e] = make_double2(cs, sn);
}

__global__ void __launch_bounds__(256) k_sym_gather(size_t n, int nops, const int* __restrict__ src,
                                                    const double2* __restrict__ ph,
                                                    const double2* __restrict__ in, double2* __restrict__ out,
                                                    double inv) {
  __shared__ double2 term[16][SYM_MAX_OPS];
  const int pt = threadIdx.x >> 4, sub = threadIdx.x & 15;
  const size_t idx = (size_t)blockIdx.x * 16 + pt;
  constexpr int S = SYM_MAX_OPS / 16;
  if (idx < n) {
    int sx[S];
    double2 p[S], v[S];
#pragma unroll
    for (int q = 0; q < S; ++q) {
      const int o = sub + 16 * q;
      sx[q] = o < nops ? src[idx * nops + o] : 0;
      p[q] = o < nops ? ph[idx * nops + o] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int q = 0; q < S; ++q) v[q] = in[sx[q]];
#pragma unroll
    for (int q = 0; q < S; ++q) {
      const int o = sub + 16 * q;
      // (cos, -sin) · v
      if (o < nops)
        term[pt][o] = make_double2(p[q].x * v[q].x + p[q].y * v[q].y, p[q].x * v[q].y - p[q].y * v[q].x);
    }
  }
  __syncthreads();
  if (idx < n && sub == 0) {
    double ar = 0.0, ai = 0.0;
    for (int o = 0; o < nops; ++o) {
      const double2 t = term[pt][o];
      ar += t.x;
      ai += t.y;
    }
    out[idx] = make_double2(ar * inv, ai * inv);
  }
}

// ---------------------------------------------------------------------------------------
// deterministic dot products: out[t] = Σ_i (a_t - as_t)_i · (b_t - bs_t)_i, a missing b
// counts as 1, kind 1 is the Hartree sum Σ_{|G|²>1e-12} |ρ(G)|²/|G|² (a = ρ(G)
// interleaved, b = |G|²).
// ---------------------------------------------------------------------------------------
struct DotTerm {
  const double* a;
  const double* as;
  const double* b;
  const double* bs;
  int kind;
};
struct DotArgs {
  int nterms;
  DotTerm t[MAX_TERMS];
};

__global__ void __launch_bounds__(RED_NT) k_dots(size_t n, DotArgs args, double* __restrict__ partial) {
  __shared__ double sh[RED_NT / 64];
  const DotTerm T = args.t[blockIdx.y];
  double acc = 0.0;
  for (size_t i = (size_t)blockIdx.x * RED_NT + threadIdx.x; i < n; i += (size_t)gridDim.x * RED_NT) {
    if (T.kind == 1) {
      const double g2 = T.b[i];
      if (g2 > 1e-12) {
        const double re = T.a[2 * i], im = T.a[2 * i + 1];
        acc += (re * re + im * im) / g2;
      }
    } else {
      const double x = T.a[i] - (T.as ? T.as[i] : 0.0);
      const double y = T.b ? T.b[i] - (T.bs ? T.bs[i] : 0.0) : 1.0;
      acc += x * y;
    }
  }
  const double s = block_reduce(acc, sh);
  if (threadIdx.x == 0) partial[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = s;
}

// Setup uploads of one group in one launch: item blockIdx.y's 32-bit words go from the pinned host
// staging block (read by the device directly) to its device array.
constexpr int UNPACK_MAX = 16;
struct UnpackArgs {
  const unsigned* src[UNPACK_MAX];
  unsigned* dst[UNPACK_MAX];
  unsigned words[UNPACK_MAX];
};
__global__ void __launch_bounds__(256) k_unpack(UnpackArgs a) {
  const int j = blockIdx.y;
  const unsigned* __restrict__ s = a.src[j];
  unsigned* __restrict__ d = a.dst[j];
  const unsigned n = a.words[j];
  for (unsigned w = blockIdx.x * 256u + threadIdx.x; w < n; w += gridDim.x * 256u) d[w] = s[w];
}

// The output density's LDA pass (k_lda_core with the ρ(G) scale), optionally the Anderson history
// entry of the iteration (k_hist_store's), and the reduction partials of the iteration's energy
// terms in one launch: block b takes the RED_NT points of k_dots' block b (one
// point per thread, n <= RED_BLOCKS·RED_NT) and, after its LDA writes, forms each term's block partial
// with k_dots' expressions and block_reduce's additions, so the partials are those k_dots would
// write. A term may read what this launch writes: only the point's own thread writes and reads it.
constexpr int LDA_DOTS_MAX = 12;  // terms of the fused launch (the iteration's energy sums take 9)
// The Anderson step's sums over the new residual r_l = ρ − hist_in and the mm older residuals r_k
// (the terms the mixing step reduces, in its order: (r_k − r_l)·r_l for k < mm, then (r_i − r_l)·(r_j − r_l)
// for i ≤ j < mm); mm ≤ LDA_GRAM_MAX, each pair held at a fixed register slot
constexpr int LDA_GRAM_MAX = 7;
constexpr int LDA_GRAM_PAIRS = LDA_GRAM_MAX * (LDA_GRAM_MAX + 1) / 2;
struct GramArgs {
  int mm;
  const double* res[LDA_GRAM_MAX];
};
__device__ __forceinline__ int gram_pair_slot(int i, int j) { return i * LDA_GRAM_MAX - i * (i - 1) / 2 + (j - i); }
__global__ void __launch_bounds__(RED_NT) k_lda_dots(size_t n, const double* rho, const double* core, double cs,
                                                     double* rho_xc, double* exc, double* vxc, double2* g, double gs,
                                                     const double* hist_in, double* hin, double* hres, DotArgs args,
                                                     GramArgs gram, double* __restrict__ partial) {
  __shared__ double sh[RED_NT / 64][LDA_DOTS_MAX];
  __shared__ double shg[RED_NT / 64][LDA_GRAM_MAX + LDA_GRAM_PAIRS];
  const size_t i = (size_t)blockIdx.x * RED_NT + threadIdx.x;
  if (hin && i < n) {  // k_hist_store's entry (in = hist_in, out = rho)
    const double x = hist_in[i];
    hin[i] = x;
    hres[i] = rho[i] - x;
  }
  if (i < n) {
    const double2 x = g[i];
    g[i] = make_double2(x.x * gs, x.y * gs);
    const double r = core ? fmax(rho[i] + core[i] * cs, 0.0) : rho[i];
    rho_xc[i] = r;
    double e, v;
    xc::lda_point(r, e, v);
    exc[i] = e;
    vxc[i] = v;
  }
  // every term's value at this point first (their loads in flight together), then the 64-lane
  // trees of all terms, then each term's wave sums in wave order: block_reduce's arithmetic per term
  double acc[LDA_DOTS_MAX];
#pragma unroll
  for (int t = 0; t < LDA_DOTS_MAX; ++t) {
    acc[t] = 0.0;
    if (t < args.nterms && i < n) {
      const DotTerm& T = args.t[t];
      if (T.kind == 1) {
        const double g2 = T.b[i];
        if (g2 > 1e-12) {
          const double re = T.a[2 * i], im = T.a[2 * i + 1];
          acc[t] += (re * re + im * im) / g2;
        }
      } else {
        const double x = T.a[i] - (T.as ? T.as[i] : 0.0);
        const double y = T.b ? T.b[i] - (T.bs ? T.bs[i] : 0.0) : 1.0;
        acc[t] += x * y;
      }
    }
  }
  // the Anderson sums (k_dots' arithmetic: x = r_k − r_l, y = r_l − 0 or r_j − r_l, acc = 0 + x·y)
  constexpr int NG = LDA_GRAM_MAX + LDA_GRAM_PAIRS;
  double gacc[NG];
#pragma unroll
  for (int q = 0; q < NG; ++q) gacc[q] = 0.0;
  if (gram.mm > 0 && i < n) {
    const double rl = rho[i] - hist_in[i];  // the entry's residual, as stored above
    double xk[LDA_GRAM_MAX];
#pragma unroll
    for (int k = 0; k < LDA_GRAM_MAX; ++k) xk[k] = k < gram.mm ? gram.res[k][i] - rl : 0.0;
#pragma unroll
    for (int k = 0; k < LDA_GRAM_MAX; ++k)
      if (k < gram.mm) gacc[k] += xk[k] * (rl - 0.0);
#pragma unroll
    for (int a = 0; a < LDA_GRAM_MAX; ++a)
#pragma unroll
      for (int b = a; b < LDA_GRAM_MAX; ++b)
        if (b < gram.mm) gacc[LDA_GRAM_MAX + gram_pair_slot(a, b)] += xk[a] * xk[b];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int t = 0; t < LDA_DOTS_MAX; ++t) acc[t] += __shfl_down(acc[t], off, 64);
    if (gram.mm > 0)  // uniform
#pragma unroll
      for (int q = 0; q < NG; ++q) gacc[q] += __shfl_down(gacc[q], off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int t = 0; t < LDA_DOTS_MAX; ++t) sh[threadIdx.x >> 6][t] = acc[t];
#pragma unroll
    for (int q = 0; q < NG; ++q) shg[threadIdx.x >> 6][q] = gacc[q];
  }
  __syncthreads();
  const int ngram = gram.mm > 0 ? gram.mm + gram.mm * (gram.mm + 1) / 2 : 0;
  const int t = threadIdx.x;
  if (t < args.nterms + ngram) {
    int slot;  // the term's register slot
    bool energy = t < args.nterms;
    if (energy) {
      slot = t;
    } else if (t - args.nterms < gram.mm) {
      slot = t - args.nterms;
    } else {
      int u = t - args.nterms - gram.mm, a = 0;  // pair (a, b) in the mixing step's order
      while (u >= gram.mm - a) {
        u -= gram.mm - a;
        ++a;
      }
      slot = LDA_GRAM_MAX + gram_pair_slot(a, a + u);
    }
    double s = 0.0;
    for (int w = 0; w < RED_NT / 64; ++w) s += energy ? sh[w][slot] : shg[w][slot];
    partial[(size_t)t * gridDim.x + blockIdx.x] = s;
  }
}

// per-band expectation values: e[2b] = Σ_G |ψ|² t_G, e[2b+1] = Re Σ_G conj(ψ)·x; blockIdx.y = k
// steps every operand by its k stride (one launch for all k-points)
__global__ void __launch_bounds__(256) k_band_expect(int npw, const double2* __restrict__ psi,
                                                    const double* __restrict__ t, const double2* __restrict__ x,
                                                    double* __restrict__ e, size_t psi_k, size_t t_k, size_t x_k,
                                                    size_t e_k) {
  __shared__ double sh[4];
  const int b = blockIdx.x;
  psi += blockIdx.y * psi_k;
  t += blockIdx.y * t_k;
  x += blockIdx.y * x_k;
  e += blockIdx.y * e_k;
  const double2* p = psi + (size_t)b * npw;
  const double2* q = x + (size_t)b * npw;
  double kin = 0.0, nl = 0.0;
  for (int i = threadIdx.x; i < npw; i += 256) {
    const double2 v = p[i], w = q[i];
    kin += (v.x * v.x + v.y * v.y) * t[i];
    nl += v.x * w.x + v.y * w.y;
  }
  const double s1 = block_reduce(kin, sh);
  __syncthreads();
  const double s2 = block_reduce(nl, sh);
  if (threadIdx.x == 0) {
    e[2 * b] = s1;
    e[2 * b + 1] = s2;
  }
}

// Anderson update: out = c_last·(in_l + β r_l) + Σ_j α_j (in_j + β r_j)
struct MixArgs {
  int m;
  double beta;
  double coef[MAX_HIST];
  const double* in[MAX_HIST];
  const double* res[MAX_HIST];
};

// With rho_xc non-null, also the LDA of the result (ρ_xc = max(ρ + cs·ρ_core, 0), ε_xc, v_xc):
// the next iteration's V_eff needs exactly these, and evaluating them here spreads the
// transcendental work over the whole grid launch instead of the one workgroup of the V_eff kernel.
__global__ void k_anderson_combine(size_t n, MixArgs a, double* __restrict__ out, const double* __restrict__ core,
                                   double cs, double* __restrict__ rho_xc, double* __restrict__ exc,
                                   double* __restrict__ vxc) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int last = a.m - 1;
  double v = a.coef[last] * (a.in[last][i] + a.beta * a.res[last][i]);
  for (int j = 0; j < last; ++j) v += a.coef[j] * (a.in[j][i] + a.beta * a.res[j][i]);
  out[i] = v;
  if (rho_xc) {
    const double r = core ? fmax(v + core[i] * cs, 0.0) : v;
    rho_xc[i] = r;
    double e, x;
    xc::lda_point(r, e, x);
    exc[i] = e;
    vxc[i] = x;
  }
}

// ---------------------------------------------------------------------------------------
// Device-resident SCF iteration: the scalars of an iteration (reduction results, Fermi
// level, band energy, -TS, electrons per channel, mixing coefficients) live in one device
// array S, and the decisions the host path takes between launches (another filter pass or
// not, the filter coefficients, occupations, the DIIS solve, the history slot) are taken by
// kernels, so an iteration is a fixed sequence of launches with one host synchronisation at
// its end (and can be captured and replayed as a HIP graph).
// ---------------------------------------------------------------------------------------
constexpr int S_HAR = 0;     // Harris-side reductions (up to 4 terms)
constexpr int S_INT = 4;     // density integral per channel (2)
constexpr int S_OUT = 8;     // output-density reductions (up to 8 terms)
constexpr int S_EF = 16;     // Fermi level, then the level of each channel (3)
constexpr int S_EBAND = 19;  // Σ f w ε
constexpr int S_TS = 20;     // TS of the smearing
constexpr int S_NS = 21;     // electrons per channel (2)
constexpr int S_DELTA = 23;  // RMS Δρ of the latest iteration (filter target)
constexpr int S_PASSES = 24; // filter passes run in the iteration
constexpr int S_PASSED = 25; // 1 once the residual gate has passed in the iteration
constexpr int S_MIXC = 32;   // mixing coefficients, two mixers × MAX_HIST
constexpr int S_GRAM = S_MIXC + 2 * MAX_HIST;  // DIIS sums, two mixers × MAX_DOTS
constexpr int S_SIZE = S_GRAM + 2 * MAX_DOTS;

// Second stage of the fixed-order reduction: out[t] = Σ_b partial[t][b] in block order (the
// order the host path adds them in).
__global__ void k_dots_final(int nterms, int nblk, const double* __restrict__ partial, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nterms) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += partial[(size_t)t * nblk + b];
  out[t] = s;
}

// flag[0]: the residual gate has passed in this iteration (the remaining passes do nothing)
__global__ void k_iter_begin(int* flag, double* S) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    flag[0] = 0;
    S[S_PASSES] = 0.0;
    S[S_PASSED] = 0.0;
  }
}

__device__ double wg_sum(double v, double* sh) {  // deterministic sum over a 256-thread workgroup
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  return (sh[0] + sh[1]) + (sh[2] + sh[3]);
}
__device__ double wg_max(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  return fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

// After a Rayleigh-Ritz step (one workgroup): with gate set, the residual gate of the pass that
// ran (the largest residual norm of the nb bands used for occupations against the target 0.1·Δρ
// of the latest iteration, clamped to [1e-9, 1e-3]; once it passes, the remaining passes of the
// iteration do nothing) and the pass count; then, unless the gate has passed, the coefficients
// of the next pass from the current Ritz values (the host path's chebyshev_pass): per block k the
// interval [a, b] = [highest kept Ritz value, spectrum bound], scaled at the lowest Ritz value,
// and the three-term coefficients of every step.
__device__ __forceinline__ void pass_gate_body(int nke, int nbs, int nb, int degree, int gate,
                                               const double* __restrict__ lam, const double* __restrict__ res,
                                               const double* __restrict__ ub, double* __restrict__ S,
                                               int* __restrict__ flag, double* __restrict__ coef, double* sh) {
  const int done = flag[0];
  __syncthreads();  // every thread has read the flag before thread 0 may set it below
  if (done) return;  // uniform
  bool passed = false;
  if (gate) {
    double w = 0.0;
    for (int i = threadIdx.x; i < nke * nb; i += 256) w = fmax(w, res[(size_t)(i / nb) * nbs + i % nb]);
    w = wg_max(w, sh);
    passed = w < fmax(1e-9, fmin(1e-3, 0.1 * S[S_DELTA]));  // the same on every thread
    if (threadIdx.x == 0) {
      S[S_PASSES] += 1.0;
      if (passed) {
        flag[0] = 1;
        S[S_PASSED] = 1.0;
      }
    }
  }
  // the coefficients are computed also when the gate has passed: the next iteration's first
  // pass uses them
  for (int k = threadIdx.x; k < nke; k += 256) {
    const double* l = lam + (size_t)k * nbs;
    const double a0 = l[0];
    double a = l[nbs - 1], b = ub[k];
    if (!(b > a + 1e-6)) b = a + fmax(1.0, fabs(a));
    if (!(a > a0)) a = a0 + 1e-3 * (b - a0);
    const double e = 0.5 * (b - a), cc = 0.5 * (b + a);
    double sigma = e / (a0 - cc);
    const double tau = 2.0 / sigma;
    coef[(size_t)3 * k] = sigma / e;
    coef[(size_t)3 * k + 1] = cc;
    coef[(size_t)3 * k + 2] = 0.0;
    for (int i = 1; i < degree; ++i) {
      const double sn = 1.0 / (tau - sigma);
      double* cf = coef + ((size_t)i * nke + k) * 3;
      cf[0] = 2.0 * sn / e;
      cf[1] = cc;
      cf[2] = sigma * sn;
      sigma = sn;
    }
  }
}

__global__ void __launch_bounds__(256) k_pass_gate(int nke, int nbs, int nb, int degree, int gate,
                                                   const double* __restrict__ lam, const double* __restrict__ res,
                                                   const double* __restrict__ ub, double* __restrict__ S,
                                                   int* __restrict__ flag, double* __restrict__ coef) {
  __shared__ double sh[4];
  pass_gate_body(nke, nbs, nb, degree, gate, lam, res, ub, S, flag, coef, sh);
}

// src/scf/smearing.rs:112-130
__device__ inline double fd_occ_dev(double e, double ef, double sigma) {
  if (sigma < 1e-15) return e < ef ? 1.0 : (fabs(e - ef) < 1e-12 ? 0.5 : 0.0);
  const double x = (e - ef) / sigma;
  if (x > 40.0) return 0.0;
  if (x < -40.0) return 1.0;
  return 1.0 / (1.0 + exp(x));
}

struct FermiArgs {
  int nk, nb, nbs, ns, fixed;  // Ritz values of channel s, k-point k at lam[(s·nk + k)·nbs + b]
  double n_el, sigma, spin, tot_mag;
  const double* lam;
  const double* wk;
  double* S;
  double* occ;  // ns·nk·nb occupations f·spin
  double* fw;   // ns·nk·nbs density weights f·spin·w_k (0 on the buffer bands)
  double* ev;   // ns·nk·nb eigenvalues
};

// Fermi level over the entries [e0, e1) of the (channel, k, band) list: the electron count
// (src/scf/smearing.rs:49-90) bracketed from the lowest and highest eigenvalue padded by
// 10·max(σ, 0.1), then narrowed by multisection until the bracket is below 1e-14 wide (the
// bisection's stopping width), 16 probes per round instead of one: round r evaluates the count
// at lo + (j + 1)(hi − lo)/17, j = 0..15, one group of 16 lanes per probe, and keeps the
// sub-interval where the count crosses the target. Every reduction is a fixed shuffle tree, so
// the result does not depend on scheduling.
__device__ double fermi_bisect(const FermiArgs& a, int e0, int e1, double target, double* sh) {
  const int nb = a.nb;
  double lo = INFINITY, hi = -INFINITY;
  for (int i = e0 + (int)threadIdx.x; i < e1; i += 256) {
    const double e = a.lam[(size_t)(i / nb) * a.nbs + i % nb];
    lo = fmin(lo, e);
    hi = fmax(hi, e);
  }
  lo = -wg_max(-lo, sh);
  hi = wg_max(hi, sh);
  const double pad = 10.0 * fmax(a.sigma, 0.1);
  lo -= pad;
  hi += pad;
  __shared__ double cnt_sh[16];
  const int probe = threadIdx.x >> 4, sub = threadIdx.x & 15;
  for (int round = 0; round < 64 && !(fabs(hi - lo) < 1e-14); ++round) {
    const double step = (hi - lo) / 17.0;
    const double mid = lo + (double)(probe + 1) * step;
    double cnt = 0.0;
    for (int i = e0 + sub; i < e1; i += 16) {
      const int kk = i / nb;
      cnt += a.wk[kk % a.nk] * (fd_occ_dev(a.lam[(size_t)kk * a.nbs + i % nb], mid, a.sigma) * a.spin);
    }
    for (int o = 8; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 16);
    __syncthreads();
    if (sub == 0) cnt_sh[probe] = cnt;
    __syncthreads();
    // the first probe whose count reaches the target bounds the level from above
    int j = 16;
    for (int q = 15; q >= 0; --q)
      if (!(cnt_sh[q] < target)) j = q;
    const double nlo = j == 0 ? lo : lo + (double)j * step;
    const double nhi = j == 16 ? hi : lo + (double)(j + 1) * step;
    lo = nlo;
    hi = nhi;
  }
  __syncthreads();
  return 0.5 * (lo + hi);
}

// Fermi level(s), occupations, the density weights, the band energy, TS and the electrons per
// channel (src/scf/smearing.rs:49-90, 202-223; src/scf/driver_spin.rs:407-474 for the two
// levels of a fixed moment), one workgroup.
__global__ void __launch_bounds__(256) k_fermi(FermiArgs a) {
  __shared__ double sh[4];
  const int per = a.nk * a.nb, tot = a.ns * per;
  double ef, efs[2];
  if (a.ns == 2 && a.fixed) {
    efs[0] = fermi_bisect(a, 0, per, 0.5 * (a.n_el + a.tot_mag), sh);
    efs[1] = fermi_bisect(a, per, tot, (a.n_el - a.tot_mag) / 2.0, sh);
    ef = 0.5 * (efs[0] + efs[1]);
  } else {
    ef = fermi_bisect(a, 0, tot, a.n_el, sh);
    efs[0] = efs[1] = ef;
  }
  double eb = 0.0, ts = 0.0, n0 = 0.0, n1 = 0.0;
  for (int i = threadIdx.x; i < tot; i += 256) {
    const int kk = i / a.nb, b = i % a.nb, s = kk / a.nk;
    const double e = a.lam[(size_t)kk * a.nbs + b], w = a.wk[kk % a.nk];
    const double f = fd_occ_dev(e, efs[s], a.sigma) * a.spin;
    a.occ[i] = f;
    a.ev[i] = e;
    eb += f * w * e;
    (s == 0 ? n0 : n1) += f * w;
    if (a.sigma >= 1e-15) {
      const double x = (e - ef) / a.sigma;
      if (fabs(x) <= 30.0) {
        double g = 1.0 / (1.0 + exp(x));
        g = fmin(fmax(g, 1e-30), 1.0 - 1e-30);
        ts += w * (-(g * log(g) + (1.0 - g) * log(1.0 - g)));
      }
    }
  }
  // density weights: each thread evaluates its own entries (the same arithmetic as occ above),
  // so no thread reads what another one wrote
  for (int i = threadIdx.x; i < a.ns * a.nk * a.nbs; i += 256) {
    const int kk = i / a.nbs, b = i % a.nbs, s = kk / a.nk;
    a.fw[i] = b < a.nb ? (fd_occ_dev(a.lam[(size_t)kk * a.nbs + b], efs[s], a.sigma) * a.spin) * a.wk[kk % a.nk]
                       : 0.0;
  }
  eb = wg_sum(eb, sh);
  ts = wg_sum(ts, sh);
  n0 = wg_sum(n0, sh);
  n1 = wg_sum(n1, sh);
  if (threadIdx.x == 0) {
    a.S[S_EF] = ef;
    a.S[S_EF + 1] = efs[0];
    a.S[S_EF + 2] = efs[1];
    a.S[S_EBAND] = eb;
    a.S[S_TS] = a.sigma < 1e-15 ? 0.0 : ts * a.sigma * a.spin;
    a.S[S_NS] = n0;
    a.S[S_NS + 1] = n1;
  }
}

// ρ *= target / (∫ρ) when |∫ρ| > 1e-15 (src/scf/density.rs:99-108); the integral is
// S[i_int]·dvol, the target n_el or the channel's electron count S[i_target]
__global__ void k_scale_to(size_t n, double* __restrict__ rho, const double* __restrict__ S, int i_int, double dvol,
                           int i_target, double n_el) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double integral = S[i_int] * dvol;
  if (!(fabs(integral) > 1e-15)) return;
  const double target = i_target >= 0 ? S[i_target] : n_el;
  rho[i] *= target / integral;
}

// the latest RMS Δρ (max over channels) for the next iteration's filter target
__global__ void k_iter_end(int ns, double dvol, double omega, double* S) {
  if (threadIdx.x || blockIdx.x) return;
  double d = sqrt(S[S_OUT + 3] * dvol / omega);
  if (ns == 2) d = fmax(sqrt(S[S_OUT + 4] * dvol / omega), sqrt(S[S_OUT + 5] * dvol / omega));
  S[S_DELTA] = d;
}

// Anderson history on the device (src/scf/mixing/anderson.rs:112-140): slot order, oldest first
struct MixState {
  int m, next;
  int order[MAX_HIST];
};

__global__ void k_hist_advance(MixState* st, int ndim) {
  if (threadIdx.x || blockIdx.x) return;
  int slot;
  if (st->m < ndim) {
    slot = st->next++;
    st->order[st->m++] = slot;
  } else {
    slot = st->order[0];
    for (int j = 0; j + 1 < ndim; ++j) st->order[j] = st->order[j + 1];
    st->order[ndim - 1] = slot;
  }
}

// the new entry: input and residual (residual from `res`, or out − in when res is null)
__global__ void k_hist_put(size_t n, const double* __restrict__ in, const double* __restrict__ out,
                           const double* __restrict__ res, double* __restrict__ hist, const MixState* __restrict__ st) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double* hin = hist + (size_t)st->order[st->m - 1] * 2 * n;
  const double x = in[i];
  hin[i] = x;
  hin[n + i] = res ? res[i] : out[i] - x;
}

__global__ void k_resid(size_t n, const double* __restrict__ in, const double* __restrict__ out,
                        double* __restrict__ r) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = out[i] - in[i];
}

// DIIS sums of one Anderson step for the largest history: term t < M (M = ndim − 1) is
// Σ (r_t − r_l)·r_l, then the upper triangle Σ (r_i − r_l)(r_j − r_l), i ≤ j < M, row by row;
// terms beyond the current history are zero. Same per-element products and block partials as
// k_dots on the host path.
__global__ void __launch_bounds__(RED_NT) k_gram(size_t n, int M, const double* __restrict__ hist,
                                                 const MixState* __restrict__ st, double* __restrict__ partial) {
  __shared__ double sh[RED_NT / 64];
  const int t = blockIdx.y, mm = st->m - 1;
  int i, j;
  bool bterm = t < M;
  if (bterm) {
    i = t;
    j = -1;
  } else {
    int u = t - M;
    i = 0;
    while (u >= M - i) {
      u -= M - i;
      ++i;
    }
    j = i + u;
  }
  double acc = 0.0;
  if (i < mm && j < mm) {
    const double* rl = hist + (size_t)st->order[mm] * 2 * n + n;
    const double* ri = hist + (size_t)st->order[i] * 2 * n + n;
    const double* rj = bterm ? nullptr : hist + (size_t)st->order[j] * 2 * n + n;
    for (size_t e = (size_t)blockIdx.x * RED_NT + threadIdx.x; e < n; e += (size_t)gridDim.x * RED_NT) {
      const double x = ri[e] - rl[e];
      const double y = bterm ? rl[e] - 0.0 : rj[e] - rl[e];
      acc += x * y;
    }
  }
  const double s = block_reduce(acc, sh);
  if (threadIdx.x == 0) partial[(size_t)t * gridDim.x + blockIdx.x] = s;
}

// src/scf/mixing/linalg.rs:24-52 (the host path's small_solve), one thread
__device__ void small_solve_dev(double* A, double* b, int n, double* x) {
#pragma clang fp contract(off)
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (fabs(A[r * n + c]) > fabs(A[p * n + c])) p = r;
    if (p != c) {
      for (int j = 0; j < n; ++j) {
        const double t = A[c * n + j];
        A[c * n + j] = A[p * n + j];
        A[p * n + j] = t;
      }
      const double t = b[c];
      b[c] = b[p];
      b[p] = t;
    }
    if (fabs(A[c * n + c]) < 1e-15) {
      for (int i = 0; i < n; ++i) x[i] = 1.0 / (double)(n + 1);
      return;
    }
    for (int r = c + 1; r < n; ++r) {
      const double f = A[r * n + c] / A[c * n + c];
      for (int j = c; j < n; ++j) A[r * n + j] -= f * A[c * n + j];
      b[r] -= f * b[c];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = b[i];
    for (int j = i + 1; j < n; ++j) s -= A[i * n + j] * x[j];
    x[i] = s / A[i * n + i];
  }
}

// The Anderson coefficients of the current history from the DIIS sums (anderson.rs:170-203)
__global__ void k_diis(const MixState* __restrict__ st, int M, const double* __restrict__ gram,
                       double* __restrict__ coef) {
#pragma clang fp contract(off)
  __shared__ double A[(MAX_HIST - 1) * (MAX_HIST - 1)], b[MAX_HIST], x[MAX_HIST];
  if (threadIdx.x || blockIdx.x) return;
  const int m = st->m;
  if (m < 2) {
    coef[0] = 1.0;
    return;
  }
  const int mm = m - 1;
  for (int i = 0; i < mm; ++i) b[i] = -gram[i];
  for (int i = 0; i < mm; ++i)
    for (int j = i; j < mm; ++j) {
      const double v = gram[M + i * M - i * (i - 1) / 2 + (j - i)];
      A[i * mm + j] = v;
      A[j * mm + i] = v;
    }
  small_solve_dev(A, b, mm, x);
  double asum = 0.0;
  for (int j = 0; j < mm; ++j) asum += x[j];
  for (int j = 0; j < mm; ++j) coef[j] = x[j];
  coef[m - 1] = 1.0 - asum;
}

// out = c_last·(in_l + β r_l) + Σ_j α_j (in_j + β r_j) over the device history
__global__ void k_anderson_dev(size_t n, double beta, const double* __restrict__ hist, const MixState* __restrict__ st,
                               const double* __restrict__ coef, double* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int last = st->m - 1;
  const double* hl = hist + (size_t)st->order[last] * 2 * n;
  double v = coef[last] * (hl[i] + beta * hl[n + i]);
  for (int j = 0; j < last; ++j) {
    const double* hj = hist + (size_t)st->order[j] * 2 * n;
    v += coef[j] * (hj[i] + beta * hj[n + i]);
  }
  out[i] = v;
}

__global__ void k_publish(int n, const double* __restrict__ S, double* __restrict__ host) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) host[i] = S[i];
}

// ---------------------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------------------
std::vector<SymOp> identity_ops() {
  SymOp id{};
  id.rot[0] = id.rot[4] = id.rot[8] = 1;
  return {id};
}


// Simpson coefficients of src/numerics.rs:29-58 (odd mesh: 1,4,2,...,4,1; even mesh:
// the QE/DFTK end correction), scaled by 3 so that Σ c_i f_i rab_i / 3 is the rule.
std::vector<double> simpson_coeffs(int n) {
  std::vector<double> c((size_t)n, 0.0);
  if (n < 3) {
    for (auto& x : c) x = 3.0;
    return c;
  }
  for (int i = 1; i < n - 1; ++i) c[(size_t)i] = (i % 2 == 1) ? 4.0 : 2.0;
  c[0] = 1.0;
  if (n % 2 == 1) {
    c[(size_t)n - 1] = 1.0;
  } else {
    c[(size_t)n - 3] += -0.25;
    c[(size_t)n - 2] += 1.0;
    c[(size_t)n - 1] = 1.25;
  }
  return c;
}

// src/scf/smearing.rs:112-130
double fd_occ(double e, double ef, double sigma) {
  if (sigma < 1e-15) return e < ef ? 1.0 : (std::fabs(e - ef) < 1e-12 ? 0.5 : 0.0);
  const double x = (e - ef) / sigma;
  if (x > 40.0) return 0.0;
  if (x < -40.0) return 1.0;
  return 1.0 / (1.0 + std::exp(x));
}

// src/scf/smearing.rs:49-90: bisection on the electron count
double fermi_level(const std::vector<double>& ev, int nk, int nb, const std::vector<double>& wk, double n_el,
                   double sigma, double spin) {
  double lo = INFINITY, hi = -INFINITY;
  for (double e : ev) {
    lo = std::min(lo, e);
    hi = std::max(hi, e);
  }
  const double pad = 10.0 * std::max(sigma, 0.1);
  lo -= pad;
  hi += pad;
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    double count = 0.0;
    for (int k = 0; k < nk; ++k) {
      double sk = 0.0;
      for (int b = 0; b < nb; ++b) sk += wk[(size_t)k] * (fd_occ(ev[(size_t)k * nb + b], mid, sigma) * spin);
      count += sk;
    }
    if (count < n_el)
      lo = mid;
    else
      hi = mid;
    if (std::fabs(hi - lo) < 1e-14) break;
  }
  return 0.5 * (lo + hi);
}

// src/scf/smearing.rs:202-223: -TS of Fermi-Dirac smearing, returned as TS
double smearing_ts(const std::vector<double>& ev, int nk, int nb, const std::vector<double>& wk, double ef,
                   double sigma, double spin) {
  if (sigma < 1e-15) return 0.0;
  double s = 0.0;
  for (int k = 0; k < nk; ++k)
    for (int b = 0; b < nb; ++b) {
      const double x = (ev[(size_t)k * nb + b] - ef) / sigma;
      double w = 0.0;
      if (std::fabs(x) <= 30.0) {
        double f = 1.0 / (1.0 + std::exp(x));
        f = std::min(std::max(f, 1e-30), 1.0 - 1e-30);
        w = -(f * std::log(f) + (1.0 - f) * std::log(1.0 - f));
      }
      s += wk[(size_t)k] * w;
    }
  return s * sigma * spin;
}

// src/scf/mixing/linalg.rs:24-52: LU with partial pivoting; a singular pivot
// returns the uniform vector 1/(n+1).
std::vector<double> small_solve(std::vector<double> A, std::vector<double> b, int n) {
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(A[(size_t)r * n + c]) > std::fabs(A[(size_t)p * n + c])) p = r;
    if (p != c) {
      for (int j = 0; j < n; ++j) std::swap(A[(size_t)c * n + j], A[(size_t)p * n + j]);
      std::swap(b[(size_t)c], b[(size_t)p]);
    }
    if (std::fabs(A[(size_t)c * n + c]) < 1e-15) return std::vector<double>((size_t)n, 1.0 / (double)(n + 1));
    for (int r = c + 1; r < n; ++r) {
      const double f = A[(size_t)r * n + c] / A[(size_t)c * n + c];
      for (int j = c; j < n; ++j) A[(size_t)r * n + j] -= f * A[(size_t)c * n + j];
      b[(size_t)r] -= f * b[(size_t)c];
    }
  }
  std::vector<double> x((size_t)n);
  for (int i = n - 1; i >= 0; --i) {
    double s = b[(size_t)i];
    for (int j = i + 1; j < n; ++j) s -= A[(size_t)i * n + j] * x[(size_t)j];
    x[(size_t)i] = s / A[(size_t)i * n + i];
  }
  return x;
}

// ---------------------------------------------------------------------------------------
// the solve
// ---------------------------------------------------------------------------------------
struct Solve {
  Context& c;
  const pwk_crystal& cr;
  const pwk_scf_params& p;
  pwk_scf_result& res;

  Lat lat, rec;
  double omega = 0, n_el = 0;
  std::vector<int32_t> miller;
  std::vector<Vec3> gvec;
  int npw = 0, nb = 0, nk = 0;
  size_t N = 0;
  std::vector<KPt> kpts;
  std::vector<double> wk;
  std::vector<SymOp> ops;
  bool have_core = false;

  // spin channels: 1, or 2 for LSDA (src/scf/driver_spin.rs); per-channel arrays hold the
  // spin-up channel first
  int ns = 1;
  int sc = 0;  // channel the eigensolver works on
  // LSDA with the filtered solver on small grids: both channels go through each eigensolver
  // launch as 2·nk blocks (V(r) slot per channel); nke = blocks per eigensolver launch
  bool fold = false;
  int nke = 0;

  // device arrays: rho_in / rho_new / rho_xc / vxc / veff_g hold ns channels of N points;
  // t_in, m_in, t_out, m_out are ρ_total and m = ρ↑ − ρ↓ of the input and output density
  // (nspin = 2), xc_tot the clamped ρ_total + ρ_core
  DevBuf vloc_g, vloc_r, core_r, rho_in, rho_g_in, rho_new, rho_g_new, rho_xc, exc, vxc, vxc_g, veff_g, cwork;
  // LDA of the output density (one spin channel): kept apart from rho_xc / exc / vxc, which hold
  // the input density's until the iteration's energy sums have read them
  DevBuf rho_xc_o, exc_o, vxc_o;
  DevBuf t_in, m_in, t_out, m_out, xc_tot;
  DevBuf Hb, W, V, V32, sym_rot, sym_tau, sym_src, sym_ph, nl_p, nl_x, expect;
  PinnedBuf pin_small, pin_partial, pin_rr;
  // setup uploads: a group of host arrays is copied into the pinned staging block and delivered by
  // one k_unpack launch (instead of a copy launch per array); a group's region stays untouched until
  // the solve ends, so the launch may read it at any time
  PinnedBuf pin_up;
  size_t up_used = 0;
  struct UpItem {
    void* dst;
    const void* src;
    size_t bytes;
  };
  std::vector<UpItem> up_items;
  void up_add(void* dst, const void* src, size_t bytes) {
    PWK_REQUIRE(bytes % 4 == 0, PWK_E_INVALID, "setup upload of a partial word");
    if (bytes) up_items.push_back(UpItem{dst, src, bytes});
  }
  void up_flush() {
    size_t total = 0;
    for (const UpItem& it : up_items) total += (it.bytes + 255) & ~(size_t)255;
    if (up_used + total > pin_up.bytes) {
      c.sync();  // earlier groups' launches have read their regions
      up_used = 0;
      pin_up.ensure(std::max(total, (size_t)1 << 20));
    }
    for (size_t i0 = 0; i0 < up_items.size(); i0 += UNPACK_MAX) {
      UnpackArgs a{};
      const size_t n = std::min(up_items.size() - i0, (size_t)UNPACK_MAX);
      unsigned most = 0;
      for (size_t i = 0; i < n; ++i) {
        const UpItem& it = up_items[i0 + i];
        unsigned char* stage = (unsigned char*)pin_up.p + up_used;
        std::memcpy(stage, it.src, it.bytes);
        up_used += (it.bytes + 255) & ~(size_t)255;
        a.src[i] = (const unsigned*)stage;
        a.dst[i] = (unsigned*)it.dst;
        a.words[i] = (unsigned)(it.bytes / 4);
        most = std::max(most, a.words[i]);
      }
      const unsigned gx = std::min(64u, std::max(1u, (most + 255u) / 256u));
      k_unpack<<<dim3(gx, (unsigned)n), 256, 0, c.stream>>>(a);
      PWK_HIP(hipGetLastError());
    }
    up_items.clear();
  }
  double* host_small = nullptr;    // pinned: reduction results
  double* host_partial = nullptr;  // pinned: per-block partial sums

  // Anderson history of one mixed quantity (src/scf/mixing/anderson.rs:20-28)
  struct Mixer {
    DevBuf hist;             // ndim slots of (input, residual)
    std::vector<int> hslot;  // logical history order → slot
    int next_slot = 0, ndim = 1;
    bool kerker = false;
    // the slot the next entry goes to (a new one until ndim are in use, then the oldest)
    int peek_slot() const { return (int)hslot.size() < ndim ? next_slot : hslot.front(); }
  };
  Mixer mix_tot, mix_mag;
  double kerker_q2 = 0.0;

  Solve(Context& cc, const pwk_crystal& a, const pwk_scf_params& b, pwk_scf_result& r) : c(cc), cr(a), p(b), res(r) {}
  ~Solve() {
    for (auto& kv : pass_graphs) (void)hipGraphExecDestroy(kv.second);
    for (auto& kv : iter_graphs) (void)hipGraphExecDestroy(kv.second);
  }

  unsigned gb(size_t n) const { return blocks_for(n, 256); }

  // Σ terms on the device, results copied to host_small[0..nterms); launches of at most
  // MAX_TERMS terms each
  void dots(const std::vector<DotTerm>& terms, size_t n) {
    PWK_REQUIRE(terms.size() <= (size_t)MAX_DOTS, PWK_E_INVALID, "too many reduction terms");
    const int nblk = (int)std::min<size_t>(RED_BLOCKS, std::max<size_t>(1, (n + RED_NT - 1) / RED_NT));
    int ngram = 0;  // the fused launch's mixing sums, after the terms
    if (lda_pending) {
      // the output LDA pass and the partials of every term in one launch (one point per thread)
      lda_pending = false;
      PWK_REQUIRE(terms.size() <= (size_t)LDA_DOTS_MAX && n <= (size_t)RED_BLOCKS * RED_NT, PWK_E_INVALID,
                  "fused LDA reduction out of range");
      DotArgs a{};
      a.nterms = (int)terms.size();
      for (size_t i = 0; i < terms.size(); ++i) a.t[i] = terms[i];
      const LdaPass& q = lda_pass;
      const int mm = q.gram.mm;
      ngram = mm > 0 ? mm + mm * (mm + 1) / 2 : 0;
      k_lda_dots<<<(unsigned)nblk, RED_NT, 0, c.stream>>>(n, q.rho, q.core, 1.0, q.rho_xc, q.exc, q.vxc, q.g, q.gs,
                                                          q.hist_in, q.hin, q.hres, a, q.gram, host_partial);
      PWK_HIP(hipGetLastError());
    } else {
      // the per-block partials are written straight into pinned host memory (no copy launch) and
      // summed here in block order (the second stage of the fixed-order reduction)
      for (size_t t0 = 0; t0 < terms.size(); t0 += MAX_TERMS) {
        const size_t nt = std::min(terms.size() - t0, (size_t)MAX_TERMS);
        DotArgs a{};
        a.nterms = (int)nt;
        for (size_t i = 0; i < nt; ++i) a.t[i] = terms[t0 + i];
        k_dots<<<dim3((unsigned)nblk, (unsigned)nt), RED_NT, 0, c.stream>>>(n, a, host_partial + t0 * nblk);
        PWK_HIP(hipGetLastError());
      }
    }
    c.sync();
    for (size_t t = 0; t < terms.size() + (size_t)ngram; ++t) {
      double sum = 0.0;
      for (int b = 0; b < nblk; ++b) sum += host_partial[t * nblk + b];
      host_small[t] = sum;
    }
  }

  // ρ(r) → ρ(G) = FFT_fwd(ρ)/N  (src/scf/energy.rs:250-259)
  void r_to_g(const double* r, double2* g) { launch_fft3d_r2c(c, r, 1.0 / (double)N, g, -1); }

  // Chebyshev-filtered subspace iteration state
  bool chebyshev = false;
  std::vector<double> ev_sub;              // ns × nk × nbs Ritz values of the filtered subspace
  double* ev_host = nullptr;               // nk × nb eigenvalues of the current channel
  int nbs = 0;                             // bands per k held in V (nb, plus a buffer when filtering)
  int fprec = PWK_FP64;                // precision of the filter's H·ψ
  int degree = 10;
  std::vector<double> ub;              // per-(channel, k) upper bound of the spectrum
  DevBuf F0, F1, F2, HYf, Y64, HY64, YH, sS, sH, sU, sT, sX, sXH, sW, sC, s_ev, s_lam, s_work, lowest;
  DevBuf kin_all64, kin_all32;              // nk × npw kinetic diagonals
  DevBuf R64, resid;                        // H·V after Rayleigh-Ritz, residual norms
  std::vector<double> resid_host;
  double last_delta = INFINITY;             // density change of the previous SCF iteration
  int passes_total = 0;
  DevBuf Ball64, DBHall64, Ball32, DBHall32, projall, coef;
  int nch_all = 0;
  std::vector<double> coef_host;

  size_t vblk() const { return (size_t)nk * npw * nbs; }  // eigenvector elements per channel
  size_t eblk() const { return (size_t)nke * npw * nbs; }  // eigenvector elements per eigensolver call
  double2* Vc() { return V.as<double2>() + (size_t)sc * vblk(); }
  double* ev_sub_c() { return ev_sub.data() + (size_t)sc * nk * nbs; }
  double* ub_c() { return ub.data() + (size_t)sc * nk; }
  double* rho_in_s(int s) { return rho_in.as<double>() + (size_t)s * N; }
  double* rho_new_s(int s) { return rho_new.as<double>() + (size_t)s * N; }

  const void* B_of(int k) { return Ball64.as<double2>() + (size_t)k * npw * nch_all; }
  const void* DBH_of(int k) { return DBHall64.as<double2>() + (size_t)k * npw * nch_all; }
  void hpsi_all(const void* psi, void* hpsi, int nbands, int prec) {
    const bool f64 = prec == PWK_FP64;
    if (fold) {
      launch_hpsi_small(c, nke * nbands, nbands, psi, hpsi, prec, f64 ? kin_all64.p : kin_all32.p, nch_all,
                        f64 ? Ball64.p : Ball32.p, f64 ? DBHall64.p : DBHall32.p, nullptr, nullptr, nk);
      return;
    }
    launch_hpsi_multi(c, nk, nbands, psi, hpsi, prec, f64 ? kin_all64.p : kin_all32.p, nch_all,
                      f64 ? Ball64.p : Ball32.p, f64 ? DBHall64.p : DBHall32.p, projall.p);
  }
  // out = α_k (H·Y − c_k Y) − β_k X for all k: one fused launch on small grids, else H·ψ
  // into HYf and the elementwise step
  void hpsi_cheb(const void* Y, const void* X, const double* cf, void* out, int prec) {
    const bool f64 = prec == PWK_FP64;
    if (prec != PWK_SPLIT_BF16 && hpsi_small_supported(c, prec, nch_all)) {
      launch_hpsi_small(c, nke * nbs, nbs, Y, out, prec, f64 ? kin_all64.p : kin_all32.p, nch_all,
                        f64 ? Ball64.p : Ball32.p, f64 ? DBHall64.p : DBHall32.p, cf, X, fold ? nk : 0);
      return;
    }
    hpsi_all(Y, HYf.p, nbs, prec);
    launch_cheb_step(c, eblk(), (size_t)npw * nbs, HYf.p, Y, X, cf, out, prec);
  }

  void setup();
  double ewald();
  void rayleigh_ritz(const double2* Y, const double2* HY);
  void rr_launch(const double2* Y, const double2* HY);
  void rr_collect();
  void rr_general(const double2* Y, const double2* HY);
  std::map<int, hipGraphExec_t> pass_graphs;  // captured filter passes per (channel, precision)
  std::set<int> pass_warm;
  double* rr_pinned = nullptr;                // pinned: Ritz values and residual norms
  PinnedBuf coef_pin;                         // pinned: filter coefficients of the one-launch pass
  PinnedBuf fw_pin;                           // pinned: band weights of the small-grid density
  void lanczos_bounds();
  void lanczos_steps(int steps, size_t per, double2* v0, double2* v1, double2* w, DevBuf& none, DevBuf& red,
                     DevBuf& cf, DevBuf& beta, DevBuf& inv, double* al_d, double* be_d);
  void lanczos_finish(int steps, DevBuf& ab);
  double chebyshev_pass(int prec);
  void eig_chebyshev(bool first);
  void veff_from(const double* rho_r, const double2* rho_g, double* rho_xc_out, double* exc_out, double* vxc_out);
  void xc_spin(const double* up, const double* dn);
  // lda: also the LDA of the result into rho_xc / exc / vxc (one spin channel)
  // prestored: the entry (in, out − in) is already in the slot peek_slot() named
  void mix(Mixer& m, const double* in, const double* out, double* result, bool lda = false, bool prestored = false);
  void run(double* eig_out, double* rho_out);

  // stages shared by the host-synchronised and the device-resident iteration
  struct IterScalars {
    double eh_in, exc_in, e_h_out, exc_int, evxc, delta, e_loc, charge;
  };
  const double* rho_tot_new = nullptr;  // output total density of the latest iteration
  std::vector<DotTerm> veff_stage();
  std::vector<DotTerm> output_stage();
  void symmetrize(double* rn, double scale = 1.0);
  bool veff_r_ready = false;  // veff_stage wrote V(r) of channel 0 into slot 0 already
  // symmetrize left N·ρ_sym(G) of the single channel's output density in rho_g_new (the
  // transform output_stage would otherwise redo from ρ(r))
  bool g_from_sym = false;
  // output_stage leaves the output LDA pass (after a symmetrisation) to the iteration's reduction
  // launch (k_lda_dots) when this is set; lda_pending then records it
  bool fuse_out_lda = false;
  struct LdaPass {
    const double* rho;
    const double* core;
    double *rho_xc, *exc, *vxc;
    double2* g;
    double gs;
    const double* hist_in;  // with hin: the Anderson history entry (hist_in, rho − hist_in) too
    double *hin, *hres;
    GramArgs gram;  // gram.mm > 0: the mixing step's sums too, after the energy terms
  };
  bool lda_pending = false;
  LdaPass lda_pass{};
  bool hist_prestored = false;  // this iteration's history entry is already stored
  std::vector<double> gram_pre;  // the mixing step's sums, when the fused launch took them
  // the mixing step left the LDA of rho_in (single channel) in rho_xc / exc / vxc
  bool lda_in_ready = false;
  // V(r) of channel s into V_eff slot `slot` for H·ψ (from V_eff(G) of the channel)
  void upload_veff(int s, int slot, bool need_g) {
    if (veff_r_ready && s == 0 && slot == 0 && !need_g) return;
    set_veff_device(c, veff_g.as<double2>() + (size_t)s * N, slot, need_g);
  }
  IterScalars scalars(const double* har, const double* out) const;
  void finalize(const std::vector<double>& occ, const IterScalars& v, double e_total, double e_harris, double ef,
                double ts, double e_band, const double* n_s);
  void run_host(double e_ewald, std::vector<double>& ev);

  // device-resident iteration (filtered eigensolver on small cubic grids)
  DevBuf dS, dpart, dflag, dmix, lam_dev, ub_dev, wk_dev, fw_dev, occ_dev, ev_dev, rtmp;
  PinnedBuf pin_S;
  double* hS = nullptr;                        // pinned copy of the iteration scalars
  std::map<int, hipGraphExec_t> iter_graphs;   // captured iterations per filter precision
  std::set<int> iter_warm;
  bool resident_supported() const;
  void dots_dev(const std::vector<DotTerm>& terms, size_t n, int off);
  void mix_dev(int which, const double* in, const double* out, double* result);
  void pass_dev(int fp);
  void gate_dev(bool gate);
  int pass_budget = 6;  // filter passes launched per iteration after the first
  void resident_body(int fp, int max_passes);
  void resident_tail();
  void launch_iteration(int fp, int passes);
  void run_resident(double e_ewald, std::vector<double>& ev);
};

void Solve::setup() {
  up_used = 0;  // a previous run of this solve has synchronised: its staging regions are free
  up_items.clear();
  PWK_REQUIRE(cr.natoms >= 1 && cr.frac && cr.species && cr.nspecies >= 1 && cr.nspecies <= MAXSP &&
                  cr.species_data,
              PWK_E_INVALID, "bad crystal description (1..8 species)");
  for (int i = 0; i < 3; ++i) lat.v[i] = {cr.lattice[3 * i], cr.lattice[3 * i + 1], cr.lattice[3 * i + 2]};
  rec = reciprocal(lat);
  omega = volume(lat);
  PWK_REQUIRE(omega > 0.0, PWK_E_INVALID, "degenerate lattice");
  n_el = 0.0;
  for (int a = 0; a < cr.natoms; ++a) {
    PWK_REQUIRE(cr.species[a] >= 0 && cr.species[a] < cr.nspecies, PWK_E_INVALID, "atom species out of range");
    n_el += cr.species_data[cr.species[a]].z_valence;
  }
  // basis, grid (src/basis.rs:36-63, src/scf/grid.rs:44-86)
  PWK_REQUIRE(p.ecut_ev > 0.0, PWK_E_INVALID, "ecut must be positive");
  sphere_basis(lat, p.ecut_ev, miller, gvec);
  npw = (int)gvec.size();
  int dims[3];
  grid_dims(miller, p.ecutrho_ratio > 0 ? p.ecutrho_ratio : 4, p.fft_dims, dims);
  for (int d = 0; d < 3; ++d) PWK_REQUIRE(dims[d] > 0 && dims[d] <= 1024, PWK_E_INVALID, "FFT grid exceeds 1024");
  nb = std::min(p.n_bands, npw);
  PWK_REQUIRE(nb >= 1, PWK_E_INVALID, "n_bands must be positive");
  PWK_REQUIRE(p.nspin >= 0 && p.nspin <= 2, PWK_E_INVALID, "nspin must be 1 or 2");
  ns = p.nspin == 2 ? 2 : 1;
  PWK_REQUIRE(p.mixing_mode == 0 || p.mixing_mode == 1, PWK_E_INVALID, "mixing_mode must be 0 (plain) or 1 (Kerker)");
  PWK_REQUIRE(p.eigensolver != 0 || heev_supported(npw, nb), PWK_E_UNSUPPORTED,
              "basis too large for the dense eigensolver");
  // |G|² per grid point (src/scf/context.rs:146-156)
  c.dims[0] = dims[0];
  c.dims[1] = dims[1];
  c.dims[2] = dims[2];
  N = (size_t)dims[0] * dims[1] * dims[2];
  c.N = N;
  std::vector<double> g2(N);
  for (size_t idx = 0; idx < N; ++idx) {
    const int i1 = (int)(idx / ((size_t)dims[1] * dims[2])), i2 = (int)((idx / dims[2]) % dims[1]),
              i3 = (int)(idx % dims[2]);
    const int n1 = i1 > dims[0] / 2 ? i1 - dims[0] : i1, n2 = i2 > dims[1] / 2 ? i2 - dims[1] : i2,
              n3 = i3 > dims[2] / 2 ? i3 - dims[2] : i3;
    g2[idx] = vnorm2(miller_vec(rec, n1, n2, n3));
  }
  c.g2_64.ensure(N * 8);
  c.g2_32.ensure(N * 4);
  PWK_HIP(hipMemcpyAsync(c.g2_64.p, g2.data(), N * 8, hipMemcpyHostToDevice, c.stream));
  launch_convert(c, c.g2_64.p, c.g2_32.p, N, true);
  c.npw = 0;
  c.have_v = false;
  setup_basis(c, npw, miller.data());
  // k-points and symmetry (src/kpoints.rs:245-288, src/symmetry/*)
  std::vector<AtomSite> sites((size_t)cr.natoms);
  for (int a = 0; a < cr.natoms; ++a) {
    sites[(size_t)a].z = cr.species_data[cr.species[a]].atomic_number;
    for (int d = 0; d < 3; ++d) sites[(size_t)a].frac[d] = cr.frac[3 * a + d];
  }
  const double tol = p.symprec > 0.0 ? p.symprec : 1e-5;
  bool trev = false;
  if (p.use_symmetry) {
    ops = detect_symmetry(lat, sites, tol);
    trev = true;
  }
  if (ops.empty()) {
    SymOp id{};
    id.rot[0] = id.rot[4] = id.rot[8] = 1;
    ops.push_back(id);
  }
  int grid[3], shift[3];
  for (int d = 0; d < 3; ++d) {
    grid[d] = p.kgrid[d] > 0 ? p.kgrid[d] : 1;
    shift[d] = p.kshift[d];
  }
  kpts = kpoints(lat, grid, shift, p.use_symmetry == 2 ? identity_ops() : ops, p.use_symmetry == 2 ? false : trev);
  nk = (int)kpts.size();
  nke = nk;
  wk.resize((size_t)nk);
  for (int k = 0; k < nk; ++k) wk[(size_t)k] = kpts[(size_t)k].weight;
  c.kd.clear();
  c.kd.resize((size_t)nk);
  c.pack_valid = false;

  // radial tables → V_loc(G), ρ_core(r), start density
  RadialArgs ra{};
  ra.n = N;
  for (int d = 0; d < 3; ++d) ra.dims[d] = dims[d];
  for (int i = 0; i < 3; ++i) {
    ra.b[3 * i] = rec.v[i].x;
    ra.b[3 * i + 1] = rec.v[i].y;
    ra.b[3 * i + 2] = rec.v[i].z;
  }
  ra.omega = omega;
  ra.sad_sigma = p.sad_sigma > 0.0 ? p.sad_sigma : 1.0;
  ra.nspecies = cr.nspecies;
  ra.natoms = cr.natoms;
  std::vector<double> hr, hwv, hwc, hwr;
  int off = 0;
  const double fourpi = 4.0 * kPiC;
  for (int s = 0; s < cr.nspecies; ++s) {
    const pwk_species& sp = cr.species_data[s];
    PWK_REQUIRE(sp.mesh >= 1 && sp.r && sp.rab && sp.v_local, PWK_E_INVALID, "species without radial mesh");
    const int n = sp.mesh;
    std::vector<double> cf = simpson_coeffs(n);
    ra.off[s] = off;
    ra.mesh[s] = n;
    ra.z[s] = sp.z_valence;
    ra.has_core[s] = sp.rho_core != nullptr;
    bool rho_nonzero = false;
    if (sp.rho_atom)
      for (int i = 0; i < n; ++i) rho_nonzero |= std::fabs(sp.rho_atom[i]) > 1e-20;
    ra.has_rho[s] = rho_nonzero;
    have_core |= sp.rho_core != nullptr;
    // G = 0 term: bare Coulomb tail removed (upf/mod.rs:171-180)
    double s0 = 0.0;
    const double tsp = 2.0 / std::sqrt(kPiC);
    for (int i = 0; i < n; ++i) {
      const double r = sp.r[i];
      const double vs0 = sp.v_local[i] + sp.z_valence * kE2 / std::max(r, 1e-20);
      s0 += cf[(size_t)i] * sp.rab[i] * (r * r * vs0);
      const double eor = r < 1e-20 ? tsp : std::erf(r) / r;
      const double vs = sp.v_local[i] + sp.z_valence * kE2 * eor;
      hr.push_back(r);
      hwv.push_back(cf[(size_t)i] * sp.rab[i] * (r * r * vs));
      hwc.push_back(sp.rho_core ? cf[(size_t)i] * sp.rab[i] * (sp.rho_core[i] * r * r) : 0.0);
      hwr.push_back(rho_nonzero ? cf[(size_t)i] * sp.rab[i] * sp.rho_atom[i] : 0.0);
    }
    ra.v0[s] = fourpi / omega * (s0 / 3.0);
    off += n;
  }
  std::vector<double> tau((size_t)cr.natoms * 3);
  std::vector<int> spi((size_t)cr.natoms);
  for (int a = 0; a < cr.natoms; ++a) {
    const Vec3 t = frac_to_cart(lat, &cr.frac[3 * a]);
    tau[(size_t)3 * a] = t.x;
    tau[(size_t)3 * a + 1] = t.y;
    tau[(size_t)3 * a + 2] = t.z;
    spi[(size_t)a] = cr.species[a];
  }
  // shells of equal |G|² (the arithmetic of the grid kernels, in grid order of first appearance)
  std::vector<double> shell_g2;
  std::vector<int> shell_of((size_t)N);
  {
    std::unordered_map<uint64_t, int> seen;
    seen.reserve(256);
    for (size_t idx = 0; idx < N; ++idx) {
      const int i1 = (int)(idx / ((size_t)dims[1] * dims[2])), i2 = (int)((idx / dims[2]) % dims[1]),
                i3 = (int)(idx % dims[2]);
      const double n1 = i1 > dims[0] / 2 ? i1 - dims[0] : i1, n2 = i2 > dims[1] / 2 ? i2 - dims[1] : i2,
                   n3 = i3 > dims[2] / 2 ? i3 - dims[2] : i3;
      double g[3];
      for (int d = 0; d < 3; ++d) g[d] = (n1 * ra.b[d] + n2 * ra.b[3 + d]) + n3 * ra.b[6 + d];
      const double g2 = (g[0] * g[0] + g[1] * g[1]) + g[2] * g[2];
      uint64_t key;
      std::memcpy(&key, &g2, 8);
      auto it = seen.find(key);
      if (it == seen.end()) {
        it = seen.emplace(key, (int)shell_g2.size()).first;
        shell_g2.push_back(g2);
      }
      shell_of[idx] = it->second;
    }
  }
  ra.nshell = (int)shell_g2.size();
  DevBuf dr, dwv, dwc, dwr, dtau, dsp, core_g, sad_g, dshg2, dshell, dtab;
  dshg2.ensure(shell_g2.size() * 8);
  dshell.ensure(shell_of.size() * 4);
  dtab.ensure(shell_g2.size() * (size_t)cr.nspecies * 3 * 8);
  up_add(dshg2.p, shell_g2.data(), shell_g2.size() * 8);
  up_add(dshell.p, shell_of.data(), shell_of.size() * 4);
  ra.shell_g2 = dshg2.as<double>();
  ra.shell = dshell.as<int>();
  ra.tab = dtab.as<double>();
  dr.ensure(hr.size() * 8);
  dwv.ensure(hr.size() * 8);
  dwc.ensure(hr.size() * 8);
  dwr.ensure(hr.size() * 8);
  dtau.ensure(tau.size() * 8);
  dsp.ensure(spi.size() * 4);
  up_add(dr.p, hr.data(), hr.size() * 8);
  up_add(dwv.p, hwv.data(), hr.size() * 8);
  up_add(dwc.p, hwc.data(), hr.size() * 8);
  up_add(dwr.p, hwr.data(), hr.size() * 8);
  up_add(dtau.p, tau.data(), tau.size() * 8);
  up_add(dsp.p, spi.data(), spi.size() * 4);
  up_flush();
  vloc_g.ensure(N * 16);
  core_g.ensure(N * 16);
  sad_g.ensure(N * 16);
  ra.r = dr.as<double>();
  ra.wv = dwv.as<double>();
  ra.wc = dwc.as<double>();
  ra.wr = dwr.as<double>();
  ra.tau = dtau.as<double>();
  ra.sp = dsp.as<int>();
  ra.vloc = vloc_g.as<double2>();
  ra.core = core_g.as<double2>();
  ra.sad = sad_g.as<double2>();
  {
    TimedScope ts(c, "scf_radial");
    k_radial_shell<<<blocks_for((size_t)ra.nshell, 4), 256, 0, c.stream>>>(ra);
    k_radial<<<blocks_for(N, 4), 256, 0, c.stream>>>(ra);
    PWK_HIP(hipGetLastError());
  }
  // working arrays
  for (DevBuf* b : {&vloc_r, &core_r, &exc}) b->ensure(N * 8);
  for (DevBuf* b : {&rho_in, &rho_new, &rho_xc, &vxc}) b->ensure((size_t)ns * N * 8);
  for (DevBuf* b : {&rho_g_in, &rho_g_new, &vxc_g, &cwork}) b->ensure(N * 16);
  if (ns == 1)
    for (DevBuf* b : {&rho_xc_o, &exc_o, &vxc_o}) b->ensure(N * 8);
  veff_g.ensure((size_t)ns * N * 16);
  if (ns == 2)
    for (DevBuf* b : {&t_in, &m_in, &t_out, &m_out, &xc_tot}) b->ensure(N * 8);
  pin_small.ensure((size_t)MAX_DOTS * sizeof(double));
  pin_partial.ensure((size_t)MAX_DOTS * RED_BLOCKS * sizeof(double));
  host_small = pin_small.as<double>();
  host_partial = pin_partial.as<double>();
  // V_loc(r) for the local energy, ρ_core(r), the start density (IFFT, real part)
  PWK_HIP(hipMemcpyAsync(cwork.p, vloc_g.p, N * 16, hipMemcpyDeviceToDevice, c.stream));
  launch_fft3d(c, cwork.p, 1, +1, PWK_FP64);
  k_c2r_scaled<<<gb(N), 256, 0, c.stream>>>(N, cwork.as<double2>(), 1.0, vloc_r.as<double>(), 0);
  if (have_core) {
    launch_fft3d(c, core_g.p, 1, +1, PWK_FP64);
    k_c2r_scaled<<<gb(N), 256, 0, c.stream>>>(N, core_g.as<double2>(), 1.0, core_r.as<double>(), 0);
  }
  launch_fft3d(c, sad_g.p, 1, +1, PWK_FP64);
  k_c2r_scaled<<<gb(N), 256, 0, c.stream>>>(N, sad_g.as<double2>(), 1.0, rho_in.as<double>(), 1);
  PWK_HIP(hipGetLastError());
  // renormalise to N_el (initial_density.rs:150-166)
  dots({DotTerm{rho_in.as<double>(), nullptr, nullptr, nullptr, 0}}, N);
  const double post = host_small[0] * (omega / (double)N);
  const double scale = post > 1e-15 ? n_el / post : 1.0;
  k_scale<<<gb(N), 256, 0, c.stream>>>(N, rho_in.as<double>(), scale);
  PWK_HIP(hipGetLastError());
  if (ns == 2) {
    // split by the mean starting magnetisation over atoms (src/scf/driver_spin.rs:48-75)
    double msum = 0.0;
    for (int a = 0; a < cr.natoms; ++a) {
      const double m = p.starting_magnetization[cr.species[a]];
      PWK_REQUIRE(std::isfinite(m) && std::fabs(m) <= 1.0, PWK_E_INVALID, "starting magnetisation outside [-1, 1]");
      msum += m;
    }
    k_spin_start<<<gb(N), 256, 0, c.stream>>>(N, rho_in.as<double>(), msum / (double)cr.natoms, rho_in_s(0),
                                              rho_in_s(1));
    PWK_HIP(hipGetLastError());
  }
  // Kerker q_TF² (src/scf/mixing/kerker.rs:50-62)
  if (p.mixing_mode == 1) {
    if (p.kerker_q_tf > 0.0) {
      kerker_q2 = p.kerker_q_tf * p.kerker_q_tf;
    } else {
      const double bohr = 0.529177210903;
      const double rho_bohr = (n_el / omega) * (bohr * bohr * bohr);
      kerker_q2 = 4.0 * std::cbrt(3.0 * kPiC * kPiC * rho_bohr) / kPiC / (bohr * bohr);
    }
  }

  // per-k kinetic diagonal and KB projectors
  std::vector<int32_t> nproj((size_t)cr.nspecies), mesh((size_t)cr.nspecies), pl;
  std::vector<double> dij, rr, rab, chi;
  for (int s = 0; s < cr.nspecies; ++s) {
    const pwk_species& sp = cr.species_data[s];
    nproj[(size_t)s] = sp.nproj;
    mesh[(size_t)s] = sp.mesh;
    for (int i = 0; i < sp.nproj; ++i) pl.push_back(sp.proj_l[i]);
    for (int i = 0; i < sp.nproj * sp.nproj; ++i) dij.push_back(sp.dij[i]);
    rr.insert(rr.end(), sp.r, sp.r + sp.mesh);
    rab.insert(rab.end(), sp.rab, sp.rab + sp.mesh);
    for (int i = 0; i < sp.nproj * sp.mesh; ++i) chi.push_back(sp.beta[i]);
  }
  std::vector<double> gflat((size_t)npw * 3);
  for (int i = 0; i < npw; ++i) {
    gflat[(size_t)3 * i] = gvec[(size_t)i].x;
    gflat[(size_t)3 * i + 1] = gvec[(size_t)i].y;
    gflat[(size_t)3 * i + 2] = gvec[(size_t)i].z;
  }
  chebyshev = p.eigensolver == 1 || (p.eigensolver < 0 && !(npw <= 1500 && heev_supported(npw, nb)));
  kin_all64.ensure((size_t)nk * npw * 8);
  kin_all32.ensure((size_t)nk * npw * 4);
  // projector channels: (2l + 1) per projector per atom, as the build lays them out
  nch_all = 0;
  {
    std::vector<int> pofs((size_t)cr.nspecies + 1, 0);
    for (int s = 0; s < cr.nspecies; ++s) pofs[(size_t)s + 1] = pofs[(size_t)s] + nproj[(size_t)s];
    for (int a = 0; a < cr.natoms; ++a) {
      const int s = spi[(size_t)a];
      PWK_REQUIRE(s >= 0 && s < cr.nspecies, PWK_E_INVALID, "atom species out of range");
      for (int i = 0; i < nproj[(size_t)s]; ++i) nch_all += 2 * pl[(size_t)(pofs[(size_t)s] + i)] + 1;
    }
  }
  {
    const size_t e = (size_t)nk * npw * std::max(1, nch_all);
    Ball64.ensure(e * 16);
    DBHall64.ensure(e * 16);
    Ball32.ensure(e * 8);
    DBHall32.ensure(e * 8);
  }
  // kinetic diagonals of all k-points (the solve reads them from these blocks): one upload, one
  // conversion
  std::vector<double> tk_all((size_t)nk * npw);
  for (int k = 0; k < nk; ++k) {
    const Vec3 kc = kpts[(size_t)k].cart;
    for (int i = 0; i < npw; ++i)
      tk_all[(size_t)k * npw + i] = kHbar2Over2m * vnorm2(vadd(kc, gvec[(size_t)i]));
  }
  PWK_HIP(hipMemcpyAsync(kin_all64.p, tk_all.data(), tk_all.size() * 8, hipMemcpyHostToDevice, c.stream));
  launch_convert(c, kin_all64.p, kin_all32.p, tk_all.size(), true);
  {
    // projectors of all k-points in contiguous blocks for the batched products, built in place,
    // every k-point in the same launches
    std::vector<double> kvs((size_t)3 * nk);
    for (int k = 0; k < nk; ++k) {
      const Vec3 kc = kpts[(size_t)k].cart;
      kvs[(size_t)3 * k] = kc.x;
      kvs[(size_t)3 * k + 1] = kc.y;
      kvs[(size_t)3 * k + 2] = kc.z;
    }
    NonlocalDest dst;
    dst.nch = nch_all;
    dst.b64 = Ball64.p;
    dst.dbh64 = DBHall64.p;
    dst.b32 = Ball32.p;
    dst.dbh32 = DBHall32.p;
    build_nonlocal_batch(c, 0, nk, kvs.data(), gflat.data(), cr.natoms, tau.data(), spi.data(), cr.nspecies,
                         nproj.data(), pl.data(), dij.data(), mesh.data(), rr.data(), rab.data(), chi.data(), omega,
                         &dst);
    for (int k = 0; k < nk; ++k) c.k(k).nch = nch_all;
  }
  // symmetry tables for the density
  PWK_REQUIRE(ops.size() <= (size_t)SYM_MAX_OPS, PWK_E_UNSUPPORTED, "more symmetry operations than the gather kernel takes");
  std::vector<int> rot;
  std::vector<double> tr;
  for (const SymOp& o : ops) {
    rot.insert(rot.end(), o.rot, o.rot + 9);
    tr.insert(tr.end(), o.tau, o.tau + 3);
  }
  if (ops.size() > 1) {
    sym_rot.ensure(rot.size() * 4);
    sym_tau.ensure(tr.size() * 8);
    up_add(sym_rot.p, rot.data(), rot.size() * 4);
    up_add(sym_tau.p, tr.data(), tr.size() * 8);
    up_flush();
    const size_t ne = N * ops.size();
    sym_src.ensure(ne * 4);
    sym_ph.ensure(ne * 16);
    k_sym_table<<<blocks_for(ne, 256), 256, 0, c.stream>>>(N, c.dims[0], c.dims[1], c.dims[2], (int)ops.size(),
                                                          sym_rot.as<int>(), sym_tau.as<double>(), sym_src.as<int>(),
                                                          sym_ph.as<double2>());
    PWK_HIP(hipGetLastError());
  }
  // eigen-system buffers; the filtered solver carries buffer bands above the nb used for
  // occupations so that the filter's damped interval starts well above them
  nbs = chebyshev ? std::min(npw, nb + std::max(8, nb / 4)) : nb;
  W.ensure((size_t)ns * nk * nb * 8);
  V.ensure((size_t)ns * nk * npw * nbs * 16);
  if (p.precision != PWK_FP64) V32.ensure((size_t)ns * nk * npw * nbs * 8);
  if (!chebyshev) {
    Hb.ensure((size_t)nk * npw * npw * 16);
    c.eig_work.ensure(heev_work_doubles(npw, nb, nk) * 8);
  } else {
    PWK_REQUIRE(heev_supported(nbs, nbs), PWK_E_UNSUPPORTED, "too many bands for the subspace eigensolver");
    fprec = p.precision;
    // filter degree: on cubic grids up to 24³ a recurrence step costs little next to the fixed
    // cost of a pass (the Rayleigh-Ritz step and the launches), so passes of 14 steps; larger
    // grids, where every step is a full H·ψ, 10. Chosen by the grid alone, so every execution
    // mode of one deck runs the same recurrence.
    const bool small_cubic = c.dims[0] == c.dims[1] && c.dims[1] == c.dims[2] && c.dims[0] <= 24;
    degree = p.cheb_degree > 0 ? p.cheb_degree : small_cubic ? 14 : 10;
    // V(r) is resident from here on: the fused-kernel checks below (and the choice of the
    // device-resident iteration) see the buffers it goes to
    reserve_veff_slots(c, 1);
    if (ns == 2 && fprec != PWK_SPLIT_BF16) {
      reserve_veff_slots(c, 2);  // V(r) of both channels stays resident
      fold = hpsi_small_supported(c, PWK_FP64, nch_all) && hpsi_small_supported(c, fprec, nch_all);
    }
    nke = fold ? ns * nk : nk;
    const size_t blk = (size_t)nke * npw * nbs;
    // FP64-sized filter buffers: a filtered solve switches its H·ψ to FP64 once the residual
    // target falls below what FP32 resolves
    for (DevBuf* b : {&F0, &F1, &F2, &HYf}) b->ensure(blk * 16);
    for (DevBuf* b : {&Y64, &HY64, &YH, &R64}) b->ensure(blk * 16);
    resid.ensure((size_t)nke * nbs * 8);
    resid_host.assign((size_t)nke * nbs, 0.0);
    for (DevBuf* b : {&sS, &sH, &sU, &sT, &sX, &sXH, &sW, &sC}) b->ensure((size_t)nke * nbs * nbs * 16);
    s_ev.ensure((size_t)nke * nbs * 8);
    s_lam.ensure((size_t)2 * nke * nbs * 8);  // Ritz values, then residual norms (fused path)
    pin_rr.ensure((size_t)2 * nke * nbs * 8);
    rr_pinned = pin_rr.as<double>();
    s_work.ensure(heev_work_doubles(nbs, nbs, nke) * 8);
    projall.ensure((size_t)nk * std::max(1, nch_all) * nbs * 16);
    ev_sub.assign((size_t)ns * nk * nbs, 0.0);
    coef.ensure((size_t)std::max(degree, 1) * nke * 3 * 8);
    ub.assign((size_t)ns * nk, 0.0);
    // start block: the nb lowest-kinetic plane waves of each k (plus a seeded perturbation)
    lowest.ensure((size_t)nbs * nk * 4);
    std::vector<int> low((size_t)nbs * nk);
    for (int k = 0; k < nk; ++k) {
      std::vector<std::pair<double, int>> t((size_t)npw);
      for (int i = 0; i < npw; ++i) t[(size_t)i] = {vnorm2(vadd(kpts[(size_t)k].cart, gvec[(size_t)i])), i};
      std::partial_sort(t.begin(), t.begin() + nbs, t.end());
      for (int j = 0; j < nbs; ++j) low[(size_t)k * nbs + j] = t[(size_t)j].second;
    }
    PWK_HIP(hipMemcpyAsync(lowest.p, low.data(), low.size() * 4, hipMemcpyHostToDevice, c.stream));
    launch_start_block(c, npw, nbs, nk, lowest.as<int>(), V.p);
    if (ns == 2)  // the same start block for the spin-down channel
      PWK_HIP(hipMemcpyAsync(V.as<double2>() + vblk(), V.p, vblk() * 16, hipMemcpyDeviceToDevice, c.stream));
    c.sync();
  }
  c.sync();
}

// src/ewald.rs:43-143
double Solve::ewald() {
  EwaldArgs a{};
  a.na = cr.natoms;
  std::vector<double> pos((size_t)cr.natoms * 3), q((size_t)cr.natoms);
  for (int i = 0; i < cr.natoms; ++i) {
    const Vec3 t = frac_to_cart(lat, &cr.frac[3 * i]);
    pos[(size_t)3 * i] = t.x;
    pos[(size_t)3 * i + 1] = t.y;
    pos[(size_t)3 * i + 2] = t.z;
    q[(size_t)i] = cr.species_data[cr.species[i]].z_valence;
  }
  a.eta = std::cbrt((double)cr.natoms * kPiC / omega);
  a.eta2 = a.eta * a.eta;
  for (int i = 0; i < 3; ++i) {
    a.nr[i] = (int)std::ceil(10.0 * a.eta / vnorm(rec.v[i]));
    a.nl[i] = (int)std::ceil(10.0 / a.eta / vnorm(lat.v[i]));
    a.b[3 * i] = rec.v[i].x;
    a.b[3 * i + 1] = rec.v[i].y;
    a.b[3 * i + 2] = rec.v[i].z;
    a.l[3 * i] = lat.v[i].x;
    a.l[3 * i + 1] = lat.v[i].y;
    a.l[3 * i + 2] = lat.v[i].z;
  }
  DevBuf dpos, dq, part;
  dpos.ensure(pos.size() * 8);
  dq.ensure(q.size() * 8);
  const int nblk = 64;
  part.ensure(2 * nblk * 8);
  up_add(dpos.p, pos.data(), pos.size() * 8);
  up_add(dq.p, q.data(), q.size() * 8);
  up_flush();
  a.pos = dpos.as<double>();
  a.q = dq.as<double>();
  k_ewald_recip<<<nblk, 256, 0, c.stream>>>(a, part.as<double>());
  k_ewald_real<<<nblk, 256, 0, c.stream>>>(a, part.as<double>() + nblk);
  PWK_HIP(hipGetLastError());
  std::vector<double> hp(2 * nblk);
  PWK_HIP(hipMemcpyAsync(hp.data(), part.p, 2 * nblk * 8, hipMemcpyDeviceToHost, c.stream));
  c.sync();
  double er = 0.0, eq = 0.0;
  for (int i = 0; i < nblk; ++i) er += hp[(size_t)i];
  for (int i = 0; i < nblk; ++i) eq += hp[(size_t)nblk + i];
  er *= 2.0 * kPiC * kE2 / omega;
  eq *= 0.5 * kE2;
  double z2 = 0.0, zs = 0.0;
  for (double z : q) {
    z2 += z * z;
    zs += z;
  }
  const double self = -a.eta / std::sqrt(kPiC) * z2 * kE2;
  const double bg = -kPiC * zs * zs * kE2 / (2.0 * omega * a.eta2);
  return eq + er + self + bg;
}

// V_eff(G) = V_loc + V_H[ρ] + FFT(v_xc[ρ + ρ_core])/N  (src/scf/driver.rs:359-431)
void Solve::veff_from(const double* rho_r, const double2* rho_g, double* rho_xc_out, double* exc_out,
                      double* vxc_out) {
  launch_lda_core(c, rho_r, have_core ? core_r.as<double>() : nullptr, 1.0, rho_xc_out, exc_out, vxc_out);
  r_to_g(vxc_out, vxc_g.as<double2>());
  launch_hartree_veff(c, rho_g, 4.0 * kPiC * kE2, vloc_g.p, vxc_g.p, veff_g.p, PWK_FP64, true, true, true);
}

// LSDA on ρ_σ + ρ_core/2 per channel (src/scf/driver_spin.rs:273-318): ε_xc into exc, v_xc↑
// and v_xc↓ into the two halves of vxc, and the clamped ρ_total + ρ_core into xc_tot (the
// density the energy integrals weight ε_xc with)
void Solve::xc_spin(const double* up, const double* dn) {
  const double* core = have_core ? core_r.as<double>() : nullptr;
  k_add_core<<<gb(N), 256, 0, c.stream>>>(N, up, core, 0.5, rho_xc.as<double>());
  k_add_core<<<gb(N), 256, 0, c.stream>>>(N, dn, core, 0.5, rho_xc.as<double>() + N);
  PWK_HIP(hipGetLastError());
  launch_lsda(c, rho_xc.p, rho_xc.as<double>() + N, exc.p, vxc.p, vxc.as<double>() + N, PWK_FP64);
}

// One Anderson step (src/scf/mixing/anderson.rs:112-203): the input and its residual (Kerker-
// preconditioned when enabled, mixing/kerker.rs:5-29) enter the history, the oldest entry
// leaves past ndim, and result = c_last·(in_l + β r_l) + Σ_j α_j (in_j + β r_j). result may
// alias in.
void Solve::mix(Mixer& mx, const double* in, const double* out, double* result, bool lda, bool prestored) {
  TimedScope ts(c, "scf_mix");
  int slot;
  if ((int)mx.hslot.size() < mx.ndim) {
    slot = mx.next_slot++;
  } else {
    slot = mx.hslot.front();
    mx.hslot.erase(mx.hslot.begin());
  }
  mx.hslot.push_back(slot);
  double* hin = mx.hist.as<double>() + (size_t)slot * 2 * N;
  double* hres = hin + N;
  if (!prestored) {
    k_hist_store<<<gb(N), 256, 0, c.stream>>>(N, in, out, hin, hres);
    PWK_HIP(hipGetLastError());
  }
  if (mx.kerker) {
    launch_fft3d_r2c(c, hres, 1.0, cwork.p, -1);
    k_kerker<<<gb(N), 256, 0, c.stream>>>(N, c.g2_64.as<double>(), kerker_q2, cwork.as<double2>());
    launch_fft3d(c, cwork.p, 1, +1, PWK_FP64);
    k_c2r_scaled<<<gb(N), 256, 0, c.stream>>>(N, cwork.as<double2>(), 1.0 / (double)N, hres, 0);
    PWK_HIP(hipGetLastError());
  }
  const int m = (int)mx.hslot.size();
  MixArgs ma{};
  ma.m = m;
  ma.beta = p.mixing_beta;
  for (int j = 0; j < m; ++j) {
    ma.in[j] = mx.hist.as<double>() + (size_t)mx.hslot[(size_t)j] * 2 * N;
    ma.res[j] = ma.in[j] + N;
  }
  if (m < 2) {
    ma.coef[0] = 1.0;
  } else {
    const int mm = m - 1;
    const double* rl = ma.res[m - 1];
    std::vector<DotTerm> t;
    for (int i = 0; i < mm; ++i) t.push_back(DotTerm{ma.res[i], rl, rl, nullptr, 0});
    for (int i = 0; i < mm; ++i)
      for (int j = i; j < mm; ++j) t.push_back(DotTerm{ma.res[i], rl, ma.res[j], rl, 0});
    if (prestored && gram_pre.size() == t.size()) {
      // the iteration's fused launch took these sums (the same terms in the same order)
      std::memcpy(host_small, gram_pre.data(), t.size() * sizeof(double));
    } else {
      dots(t, N);
    }
    std::vector<double> A((size_t)mm * mm), bvec((size_t)mm);
    for (int i = 0; i < mm; ++i) bvec[(size_t)i] = -host_small[i];
    int q = mm;
    for (int i = 0; i < mm; ++i)
      for (int j = i; j < mm; ++j) {
        A[(size_t)i * mm + j] = host_small[q];
        A[(size_t)j * mm + i] = host_small[q];
        ++q;
      }
    std::vector<double> alpha = small_solve(A, bvec, mm);
    double asum = 0.0;
    for (double x : alpha) asum += x;
    for (int j = 0; j < mm; ++j) ma.coef[j] = alpha[(size_t)j];
    ma.coef[m - 1] = 1.0 - asum;
  }
  k_anderson_combine<<<gb(N), 256, 0, c.stream>>>(N, ma, result, have_core ? core_r.as<double>() : nullptr, 1.0,
                                                  lda ? rho_xc.as<double>() : nullptr, exc.as<double>(),
                                                  vxc.as<double>());
  PWK_HIP(hipGetLastError());
}

// Rayleigh-Ritz on span(Y_k) for every k at once, FP64, Löwdin orthonormalisation:
// S = Y^H Y = U s U^H, X = U s^{-1/2}, H' = X^H (Y^H HY) X = W λ W^H, V = Y·X·W; the Ritz values
// go to ev_sub, the residual norms |HV − λV| of the nb occupied-range bands to resid_host.
// Blocks of at most 32 bands per k take one fused launch per call.
void Solve::rayleigh_ritz(const double2* Y, const double2* HY) {
  rr_launch(Y, HY);
  rr_collect();
}

// the fused path leaves Ritz values and residual norms on the device for rr_collect
void Solve::rr_launch(const double2* Y, const double2* HY) {
  if (rr_small_supported(nbs)) {
    TimedScope ts(c, "scf_rayleigh_ritz");
    const size_t nkm = (size_t)nke * nbs;
    // Ritz values and residual norms go straight into pinned host memory
    launch_rr_small(c, npw, nbs, nb, nke, Y, HY, Vc(), rr_pinned, rr_pinned + nkm, 1e-14);
    return;
  }
  rr_general(Y, HY);
}

void Solve::rr_collect() {
  if (!rr_small_supported(nbs)) return;  // the general path has already copied its results
  const size_t nkm = (size_t)nke * nbs;
  c.sync();
  std::memcpy(ev_sub_c(), rr_pinned, nkm * 8);
  std::memcpy(resid_host.data(), rr_pinned + nkm, nkm * 8);
}

void Solve::rr_general(const double2* Y, const double2* HY) {
  TimedScope ts(c, "scf_rayleigh_ritz");
  const int m = nbs;
  const long long blk = (long long)npw * m, sm = (long long)m * m;
  launch_conj_transpose(c, npw, m, Y, npw, blk, YH.p, m, blk, nke);
  launch_zgemm_mfma_batched(c, m, m, npw, YH.p, m, blk, Y, npw, blk, sS.p, m, sm, nke, false, PWK_FP64);
  launch_zgemm_mfma_batched(c, m, m, npw, YH.p, m, blk, HY, npw, blk, sH.p, m, sm, nke, false, PWK_FP64);
  launch_hermitize(c, m, sS.p, nke);
  launch_hermitize(c, m, sH.p, nke);
  launch_heev_batched(c, m, m, nke, sS.p, s_work.p, s_ev.as<double>(), sU.p, nullptr);
  launch_lowdin_scale(c, m, s_ev.as<double>(), 1e-14, sU.p, nke);  // sU is now X
  launch_zgemm_mfma_batched(c, m, m, m, sH.p, m, sm, sU.p, m, sm, sT.p, m, sm, nke, false, PWK_FP64);
  launch_conj_transpose(c, m, m, sU.p, m, sm, sXH.p, m, sm, nke);
  launch_zgemm_mfma_batched(c, m, m, m, sXH.p, m, sm, sT.p, m, sm, sX.p, m, sm, nke, false, PWK_FP64);
  launch_hermitize(c, m, sX.p, nke);
  launch_heev_batched(c, m, m, nke, sX.p, s_work.p, s_lam.as<double>(), sW.p, nullptr);
  launch_zgemm_mfma_batched(c, m, m, m, sU.p, m, sm, sW.p, m, sm, sC.p, m, sm, nke, false, PWK_FP64);
  launch_zgemm_mfma_batched(c, npw, m, m, Y, npw, blk, sC.p, m, sm, Vc(), npw, blk, nke, false, PWK_FP64);
  // residuals of the nb bands used for occupations: H·V = (H·Y)·C
  launch_zgemm_mfma_batched(c, npw, m, m, HY, npw, blk, sC.p, m, sm, R64.p, npw, blk, nke, false, PWK_FP64);
  launch_resid_norms(c, npw, m, nb, nke, Vc(), R64.p, s_lam.as<double>(), resid.as<double>());
  PWK_HIP(hipMemcpyAsync(ev_sub_c(), s_lam.p, (size_t)nke * nbs * 8, hipMemcpyDeviceToHost, c.stream));
  PWK_HIP(hipMemcpyAsync(resid_host.data(), resid.p, resid_host.size() * 8, hipMemcpyDeviceToHost, c.stream));
  c.sync();
}

// Upper bound of the spectrum of every H_k: ten Lanczos steps from a seeded
// vector per k (all k in one batch); the Gershgorin bound of each Lanczos matrix
// plus its last off-diagonal element.
void Solve::lanczos_bounds() {
  TimedScope ts(c, "scf_lanczos");
  const int steps = std::min(10, npw);
  const size_t per = (size_t)npw * 2;  // reals per k
  DevBuf ab;
  ab.ensure((size_t)2 * steps * nke * 8);  // α_j (rows 0..steps), then β_j
  double* al_d = ab.as<double>();
  double* be_d = al_d + (size_t)steps * nke;
  if (c.fused_filter && lanczos_small_supported(c, nch_all)) {
    // small cubic grids: every step of every block in one launch, vectors in LDS
    launch_lanczos_small(c, nke, fold ? nk : 0, steps, kin_all64.as<double>(), nch_all, Ball64.p, DBHall64.p, al_d);
  } else {
    DevBuf a, b0, b1, none, red, cf, beta, inv;
    a.ensure((size_t)nke * npw * 16);
    b0.ensure((size_t)nke * npw * 16);
    b1.ensure((size_t)nke * npw * 16);
    none.ensure((size_t)nke * 4);
    red.ensure((size_t)nke * 8);
    cf.ensure((size_t)nke * 3 * 8);
    beta.ensure((size_t)nke * 8);
    inv.ensure((size_t)nke * 8);
    lanczos_steps(steps, per, a.as<double2>(), b0.as<double2>(), b1.as<double2>(), none, red, cf, beta, inv, al_d,
                  be_d);
  }
  lanczos_finish(steps, ab);
}

// The Lanczos recurrence as step launches (any grid): start vector, then per step H·v, α, the
// three-term update, β and the normalisation, batched over the band blocks.
void Solve::lanczos_steps(int steps, size_t per, double2* v0, double2* v1, double2* w, DevBuf& none, DevBuf& red,
                          DevBuf& cf, DevBuf& beta, DevBuf& inv, double* al_d, double* be_d) {
  const std::vector<int> minus((size_t)nke, -1);
  PWK_HIP(hipMemcpyAsync(none.p, minus.data(), minus.size() * 4, hipMemcpyHostToDevice, c.stream));
  launch_start_block(c, npw, 1, nke, none.as<int>(), v1);
  PWK_HIP(hipMemsetAsync(v0, 0, (size_t)nke * npw * 16, c.stream));
  PWK_HIP(hipMemsetAsync(beta.p, 0, (size_t)nke * 8, c.stream));
  launch_block_dot(c, per, nke, (const double*)v1, (const double*)v1, red.as<double>());
  launch_block_scale(c, per, nke, red.as<double>(), 1, (double*)v1);
  // the recurrence coefficients stay on the device; every step runs, and the host keeps the steps
  // up to the first β_j below 1e-12 on any k (the remaining ones act on zero vectors)
  const unsigned kb = blocks_for((size_t)nke, 64);
  for (int j = 0; j < steps; ++j) {
    hpsi_all(v1, w, 1, PWK_FP64);
    launch_block_dot(c, per, nke, (const double*)v1, (const double*)w, red.as<double>());
    k_lanczos_coef<<<kb, 64, 0, c.stream>>>(nke, j, 0, red.as<double>(), beta.as<double>(), cf.as<double>(), al_d,
                                            be_d, inv.as<double>());
    launch_cheb_step(c, (size_t)nke * npw, (size_t)npw, w, v1, v0, cf.as<double>(), w, PWK_FP64);
    launch_block_dot(c, per, nke, (const double*)w, (const double*)w, red.as<double>());
    k_lanczos_coef<<<kb, 64, 0, c.stream>>>(nke, j, 1, red.as<double>(), beta.as<double>(), cf.as<double>(), al_d,
                                            be_d, inv.as<double>());
    PWK_HIP(hipGetLastError());
    if (j + 1 == steps) break;
    std::swap(v0, v1);  // v0 <- v_j
    launch_block_scale(c, per, nke, inv.as<double>(), 0, (double*)w);
    std::swap(v1, w);  // v1 <- v_{j+1}
  }
}

// Upper bound per block from the recurrence coefficients (Gershgorin on the tridiagonal matrix
// of the steps kept, plus the last β)
void Solve::lanczos_finish(int steps, DevBuf& ab) {
  std::vector<double> abh((size_t)2 * steps * nke);
  PWK_HIP(hipMemcpyAsync(abh.data(), ab.p, abh.size() * 8, hipMemcpyDeviceToHost, c.stream));
  c.sync();
  int used = steps;
  for (int j = 0; j < steps && used == steps; ++j)
    for (int k = 0; k < nke; ++k)
      if (abh[(size_t)(steps + j) * nke + k] < 1e-12) used = j + 1;
  for (int k = 0; k < nke; ++k) {
    double bound = -INFINITY;
    for (int i = 0; i < used; ++i) {
      const double bl = i > 0 ? abh[(size_t)(steps + i - 1) * nke + k] : 0.0;
      const double br = i + 1 < used ? abh[(size_t)(steps + i) * nke + k] : 0.0;
      bound = std::max(bound, abh[(size_t)i * nke + k] + std::fabs(bl) + std::fabs(br));
    }
    ub_c()[k] = bound + std::fabs(abh[(size_t)(steps + used - 1) * nke + k]);
  }
}

// One filtered pass over all k: Y = p_k(H_k)·V_k with the scaled Chebyshev
// recurrence on [a_k, b_k] (a_k = highest Ritz value kept, b_k = spectrum bound,
// scaled at the lowest Ritz value), then Rayleigh-Ritz. The filter's H·ψ runs in
// the solve's precision, the Rayleigh-Ritz step in FP64.
double Solve::chebyshev_pass(int fprec) {
  TimedScope ts(c, "scf_chebyshev_filter");
  ++passes_total;
  const size_t blk = (size_t)nke * npw * nbs;
  coef_host.assign((size_t)degree * nke * 3, 0.0);
  for (int k = 0; k < nke; ++k) {
    const double* lam = ev_sub_c() + (size_t)k * nbs;
    const double a0 = lam[0];
    double a = lam[nbs - 1], b = ub_c()[k];
    if (!(b > a + 1e-6)) b = a + std::max(1.0, std::fabs(a));
    if (!(a > a0)) a = a0 + 1e-3 * (b - a0);
    const double e = 0.5 * (b - a), cc = 0.5 * (b + a);
    double sigma = e / (a0 - cc);
    const double tau = 2.0 / sigma;
    coef_host[(size_t)3 * k] = sigma / e;
    coef_host[(size_t)3 * k + 1] = cc;
    coef_host[(size_t)3 * k + 2] = 0.0;
    for (int i = 1; i < degree; ++i) {
      const double sn = 1.0 / (tau - sigma);
      double* cf = &coef_host[((size_t)i * nke + k) * 3];
      cf[0] = 2.0 * sn / e;
      cf[1] = cc;
      cf[2] = sigma * sn;
      sigma = sn;
    }
  }
  // the pass's launches: filter, FP64 H·ψ of the filtered block, Rayleigh-Ritz. The addresses
  // depend only on the channel and the filter precision, so on the fused small-grid path the
  // sequence is captured once per (channel, precision) as a HIP graph and replayed (one launch
  // call instead of fourteen); the coefficients are uploaded before each replay
  const bool fused_filter = c.fused_filter && cheb_filter_small_supported(c, fprec, nch_all);
  const bool graph = c.graphs && !c.prof.on && rr_small_supported(nbs) && hpsi_small_supported(c, PWK_FP64, nch_all) &&
                     hpsi_small_supported(c, fprec, nch_all) && fprec != PWK_SPLIT_BF16;
  // the one-launch filter reads its coefficients once, so it takes them from pinned host memory
  // (no upload); the previous pass has completed (its results were read), so the buffer is free
  const double* cf_dev = coef.as<double>();
  if (fused_filter && !graph) {
    coef_pin.ensure(coef_host.size() * 8);
    std::memcpy(coef_pin.p, coef_host.data(), coef_host.size() * 8);
    cf_dev = coef_pin.as<double>();
  } else {
    PWK_HIP(hipMemcpyAsync(coef.p, coef_host.data(), coef_host.size() * 8, hipMemcpyHostToDevice, c.stream));
  }
  auto body = [&]() {
    if (fused_filter) {
      // the whole recurrence of every band in one launch, result straight into Y64
      const bool f64 = fprec == PWK_FP64;
      launch_cheb_filter_small(c, nke * nbs, nbs, Vc(), Y64.p, fprec, f64 ? kin_all64.p : kin_all32.p, nch_all,
                               f64 ? Ball64.p : Ball32.p, f64 ? DBHall64.p : DBHall32.p, cf_dev, degree,
                               fold ? nk : 0);
      hpsi_all(Y64.p, HY64.p, nbs, PWK_FP64);
      rr_launch(Y64.as<double2>(), HY64.as<double2>());
      return;
    }
    void* buf[3] = {F0.p, F1.p, F2.p};
    if (fprec == PWK_FP64)
      PWK_HIP(hipMemcpyAsync(buf[0], Vc(), blk * 16, hipMemcpyDeviceToDevice, c.stream));
    else
      launch_convert(c, Vc(), buf[0], blk * 2, true);
    void *X = buf[0], *Y = buf[1], *Z = buf[2];
    hpsi_cheb(X, nullptr, coef.as<double>(), Y, fprec);
    for (int i = 1; i < degree; ++i) {
      hpsi_cheb(Y, X, coef.as<double>() + (size_t)i * nke * 3, Z, fprec);
      void* t = X;
      X = Y;
      Y = Z;
      Z = t;
    }
    if (fprec == PWK_FP64)
      PWK_HIP(hipMemcpyAsync(Y64.p, Y, blk * 16, hipMemcpyDeviceToDevice, c.stream));
    else
      launch_convert(c, Y, Y64.p, blk * 2, false);
    hpsi_all(Y64.p, HY64.p, nbs, PWK_FP64);
    rr_launch(Y64.as<double2>(), HY64.as<double2>());
  };
  if (c.batcher && fused_filter && !graph && rr_small_supported(nbs) && hpsi_small_supported(c, PWK_FP64, nch_all)) {
    // served solve: the pass may share its three launches with the passes of other solves
    // (the coefficients above are uploaded on this stream before the pass's start is marked)
    const bool f64 = fprec == PWK_FP64;
    const int nkv = fold ? nk : 0, nbands = (int)(nke * nbs);
    const size_t nkm = (size_t)nke * nbs;
    PassShape sh;
    sh.fp = f64 ? PWK_FP64 : PWK_FP32;
    sh.n = c.dims[0];
    sh.npw = npw;
    sh.nch = nch_all;
    sh.nbs = nbs;
    sh.nb = nb;
    sh.degree = degree;
    sh.inv_n = 1.0 / (double)c.N;
    PassReq r{};
    r.filt = {Vc(), Y64.p, f64 ? kin_all64.p : kin_all32.p, f64 ? c.vrn64.p : c.vrn32.p, c.g2fft.as<int>(),
              f64 ? Ball64.p : Ball32.p, f64 ? DBHall64.p : DBHall32.p, cf_dev, nullptr, nbands, nkv};
    r.hpsi = {Y64.p, HY64.p, kin_all64.p, c.vrn64.p, c.g2fft.as<int>(), Ball64.p, DBHall64.p, nullptr, nullptr,
              nbands, nkv};
    r.rr = {Y64.as<double2>(), HY64.as<double2>(), Vc(), rr_pinned, rr_pinned + nkm, nullptr, (int)nke};
    pass_batch_join(c, sh, r);
  } else if (graph) {
    const int key = sc * 4 + fprec;
    auto it = pass_graphs.find(key);
    if (it != pass_graphs.end()) {
      PWK_HIP(hipGraphLaunch(it->second, c.stream));
    } else if (!pass_warm.count(key)) {
      pass_warm.insert(key);  // first pass runs directly (one-time kernel attribute set-up)
      body();
    } else {
      hipGraph_t g = nullptr;
      hipGraphExec_t ex = nullptr;
      PWK_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
      try {
        body();
      } catch (...) {
        (void)hipStreamEndCapture(c.stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      PWK_HIP(hipStreamEndCapture(c.stream, &g));
      const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      PWK_HIP(ie);
      pass_graphs[key] = ex;
      PWK_HIP(hipGraphLaunch(ex, c.stream));
    }
  } else {
    body();
  }
  rr_collect();
  double worst = 0.0;
  for (int k = 0; k < nke; ++k)
    for (int b = 0; b < nb; ++b) worst = std::max(worst, resid_host[(size_t)k * nbs + b]);
  return worst;
}

void Solve::eig_chebyshev(bool first) {
  if (first) {
    lanczos_bounds();
    PWK_HIP(hipMemcpyAsync(Y64.p, Vc(), eblk() * 16, hipMemcpyDeviceToDevice, c.stream));
    hpsi_all(Y64.p, HY64.p, nbs, PWK_FP64);
    rayleigh_ritz(Y64.as<double2>(), HY64.as<double2>());
  }
  // passes until the residual of the occupied-range bands is below a target that tightens with
  // the SCF: 10% of the previous density change (eV per unit vector), between 1e-9 and 1e-3.
  // A reduced-precision filter hands over to FP64 below 3e-6, where FP32 noise dominates.
  const double target = std::max(1e-9, std::min(1e-3, 0.1 * last_delta));
  const int max_passes = first ? 12 : 6;
  for (int pass = 0; pass < max_passes; ++pass) {
    const int prec = (fprec != PWK_FP64 && target < 3e-6) ? PWK_FP64 : fprec;
    const double worst = chebyshev_pass(prec);
    if (p.verbose) std::fprintf(stderr, "[pwk scf]   filter pass %d (%s): max residual %.3e (target %.1e)\n", pass,
                                prec == PWK_FP64 ? "fp64" : "reduced", worst, target);
    if (worst < target) break;
  }
  for (int k = 0; k < nke; ++k)
    for (int b = 0; b < nb; ++b) ev_host[(size_t)k * nb + b] = ev_sub_c()[(size_t)k * nbs + b];
}

// V_eff of every channel from the input density (src/scf/driver.rs:359-431; nspin = 2: common
// Hartree of ρ↑ + ρ↓, LSDA per channel, src/scf/driver_spin.rs:263-339). Returns the Harris-side
// reduction terms: the Hartree sum of ρ_in(G), ∫ρ_xc,in·ε_xc,in and ∫ρ_σ,in·v_σ,in
// (driver_spin.rs:634-651).
std::vector<DotTerm> Solve::veff_stage() {
  const double* core = have_core ? core_r.as<double>() : nullptr;
  struct Once {  // the mixing step's LDA serves this stage only
    bool& f;
    ~Once() { f = false; }
  } once{lda_in_ready};
  if (ns == 1) {
    veff_r_ready = false;
    if (grid_small_supported(c)) {
      // small cubic grid: ρ(G), xc, V_eff(G) and V(r) (slot 0) in one launch
      launch_rho_xc_veff_small(c, rho_in.as<double>(), core, 1.0, rho_g_in.as<double2>(), rho_xc.as<double>(),
                               exc.as<double>(), vxc.as<double>(), vloc_g.as<double2>(), 4.0 * kPiC * kE2,
                               veff_g.as<double2>(), 0, lda_in_ready);
      veff_r_ready = true;
    } else {
      r_to_g(rho_in.as<double>(), rho_g_in.as<double2>());
      veff_from(rho_in.as<double>(), rho_g_in.as<double2>(), rho_xc.as<double>(), exc.as<double>(),
                vxc.as<double>());
    }
    return {DotTerm{rho_g_in.as<double>(), nullptr, c.g2_64.as<double>(), nullptr, 1},
            DotTerm{rho_xc.as<double>(), nullptr, exc.as<double>(), nullptr, 0},
            DotTerm{rho_in.as<double>(), nullptr, vxc.as<double>(), nullptr, 0}};
  }
  k_spin_combine<<<gb(N), 256, 0, c.stream>>>(N, rho_in_s(0), rho_in_s(1), t_in.as<double>(), m_in.as<double>());
  PWK_HIP(hipGetLastError());
  r_to_g(t_in.as<double>(), rho_g_in.as<double2>());
  xc_spin(rho_in_s(0), rho_in_s(1));
  k_add_core<<<gb(N), 256, 0, c.stream>>>(N, t_in.as<double>(), core, 1.0, xc_tot.as<double>());
  PWK_HIP(hipGetLastError());
  for (int s = 0; s < 2; ++s) {
    r_to_g(vxc.as<double>() + (size_t)s * N, vxc_g.as<double2>());
    launch_hartree_veff(c, rho_g_in.p, 4.0 * kPiC * kE2, vloc_g.p, vxc_g.p, veff_g.as<double2>() + (size_t)s * N,
                        PWK_FP64, true, true, true);
  }
  return {DotTerm{rho_g_in.as<double>(), nullptr, c.g2_64.as<double>(), nullptr, 1},
          DotTerm{xc_tot.as<double>(), nullptr, exc.as<double>(), nullptr, 0},
          DotTerm{rho_in_s(0), nullptr, vxc.as<double>(), nullptr, 0},
          DotTerm{rho_in_s(1), nullptr, vxc.as<double>() + N, nullptr, 0}};
}

// G-space symmetrisation of one channel's density (src/symmetry/density/g_space.rs)
// ρ ← symmetrised (scale·ρ); scale applies only with more than one operation (the caller scales
// otherwise)
void Solve::symmetrize(double* rn, double scale) {
  if (ops.size() <= 1) return;
  TimedScope ts(c, "scf_symmetrize");
  launch_fft3d_r2c(c, rn, scale, cwork.p, -1);
  k_sym_gather<<<blocks_for(N, 16), 256, 0, c.stream>>>(N, (int)ops.size(), sym_src.as<int>(), sym_ph.as<double2>(),
                                                       cwork.as<double2>(), rho_g_new.as<double2>(),
                                                       1.0 / (double)ops.size());
  g_from_sym = ns == 1;
  if (grid_small_supported(c)) {
    launch_fft_c2r_small(c, rho_g_new.as<double2>(), 1.0 / (double)N, rn);
    return;
  }
  // the transform runs on a copy: rho_g_new keeps the symmetrised ρ(G)
  PWK_HIP(hipMemcpyAsync(cwork.p, rho_g_new.p, N * 16, hipMemcpyDeviceToDevice, c.stream));
  launch_fft3d(c, cwork.p, 1, +1, PWK_FP64);
  k_c2r_scaled<<<gb(N), 256, 0, c.stream>>>(N, cwork.as<double2>(), 1.0 / (double)N, rn, 0);
  PWK_HIP(hipGetLastError());
}

// The output-density side (driver.rs:483-560; driver_spin.rs:525-604): ρ_total of the output
// density, its ρ(G) and xc, and the reduction terms: Hartree sum, ∫ρ_xc·ε_xc, ∫ρ_σ·v_σ, the
// Δρ norms per channel, ∫ρ·V_loc and ∫ρ.
std::vector<DotTerm> Solve::output_stage() {
  const double* core = have_core ? core_r.as<double>() : nullptr;
  if (ns == 1) {
    rho_tot_new = rho_new.as<double>();
    double *rx = rho_xc_o.as<double>(), *ex = exc_o.as<double>(), *vx = vxc_o.as<double>();
    if (g_from_sym && fuse_out_lda) {
      // as below, run by the iteration's reduction launch
      lda_pending = true;
      lda_pass = LdaPass{rho_tot_new, core, rx, ex, vx, rho_g_new.as<double2>(), 1.0 / (double)N, nullptr, nullptr,
                         nullptr, GramArgs{}};
    } else if (g_from_sym) {
      // ρ(G) of the output density is the symmetrised one already in rho_g_new (times N): the
      // forward transform of ρ(r) would only return it again
      launch_lda_core(c, rho_tot_new, core, 1.0, rx, ex, vx, rho_g_new.as<double2>(), 1.0 / (double)N);
    } else if (grid_small_supported(c)) {
      launch_rho_xc_veff_small(c, rho_tot_new, core, 1.0, rho_g_new.as<double2>(), rx, ex, vx, nullptr, 0.0, nullptr,
                               0);
    } else {
      r_to_g(rho_tot_new, rho_g_new.as<double2>());
      launch_lda_core(c, rho_tot_new, core, 1.0, rx, ex, vx);
    }
    g_from_sym = false;
    return {DotTerm{rho_g_new.as<double>(), nullptr, c.g2_64.as<double>(), nullptr, 1},
            DotTerm{rx, nullptr, ex, nullptr, 0},
            DotTerm{rho_new.as<double>(), nullptr, vx, nullptr, 0},
            DotTerm{rho_in.as<double>(), rho_new.as<double>(), rho_in.as<double>(), rho_new.as<double>(), 0},
            DotTerm{rho_new.as<double>(), nullptr, vloc_r.as<double>(), nullptr, 0},
            DotTerm{rho_new.as<double>(), nullptr, nullptr, nullptr, 0}};
  }
  rho_tot_new = t_out.as<double>();
  k_spin_combine<<<gb(N), 256, 0, c.stream>>>(N, rho_new_s(0), rho_new_s(1), t_out.as<double>(), m_out.as<double>());
  PWK_HIP(hipGetLastError());
  r_to_g(rho_tot_new, rho_g_new.as<double2>());
  xc_spin(rho_new_s(0), rho_new_s(1));
  k_add_core<<<gb(N), 256, 0, c.stream>>>(N, rho_tot_new, core, 1.0, xc_tot.as<double>());
  PWK_HIP(hipGetLastError());
  return {DotTerm{rho_g_new.as<double>(), nullptr, c.g2_64.as<double>(), nullptr, 1},
          DotTerm{xc_tot.as<double>(), nullptr, exc.as<double>(), nullptr, 0},
          DotTerm{rho_new_s(0), nullptr, vxc.as<double>(), nullptr, 0},
          DotTerm{rho_new_s(1), nullptr, vxc.as<double>() + N, nullptr, 0},
          DotTerm{rho_in_s(0), rho_new_s(0), rho_in_s(0), rho_new_s(0), 0},
          DotTerm{rho_in_s(1), rho_new_s(1), rho_in_s(1), rho_new_s(1), 0},
          DotTerm{rho_tot_new, nullptr, vloc_r.as<double>(), nullptr, 0},
          DotTerm{rho_tot_new, nullptr, nullptr, nullptr, 0}};
}

// Energies of one iteration from the Harris-side (har) and output-side (out) reduction results.
Solve::IterScalars Solve::scalars(const double* har, const double* out) const {
  const double dvol = omega / (double)N, fe = 4.0 * kPiC * kE2 * 0.5 * omega;
  IterScalars v;
  v.eh_in = har[0] * fe;
  v.e_h_out = out[0] * fe;
  v.exc_int = out[1] * dvol;
  if (ns == 1) {
    v.exc_in = har[1] * dvol - har[2] * dvol;
    v.evxc = out[2] * dvol;
    v.delta = std::sqrt(out[3] * dvol / omega);
    v.e_loc = out[4] * dvol;
    v.charge = out[5] * dvol;
  } else {
    v.exc_in = har[1] * dvol - (har[2] + har[3]) * dvol;
    v.evxc = (out[2] + out[3]) * dvol;
    v.delta = std::max(std::sqrt(out[4] * dvol / omega), std::sqrt(out[5] * dvol / omega));
    v.e_loc = out[6] * dvol;
    v.charge = out[7] * dvol;
  }
  return v;
}

// The converged result: the energy split, and the kinetic and nonlocal expectation values
// summed over channels (src/scf/energy.rs:382-510), all k-points of a channel at once: batched
// projector GEMMs, one expectation launch per k, one copy back.
void Solve::finalize(const std::vector<double>& occ, const IterScalars& v, double e_total, double e_harris, double ef,
                     double ts, double e_band, const double* n_s) {
  res.converged = 1;
  res.e_total = e_total;
  res.e_harris = e_harris;
  res.fermi = ef;
  res.ts = ts;
  res.e_band = e_band;
  res.e_hartree = v.e_h_out;
  res.e_xc = v.exc_int;
  res.e_vxc = v.evxc;
  res.e_smearing = -ts;
  res.e_local = v.e_loc;
  res.integrated_charge = v.charge;
  res.cheb_degree = chebyshev ? degree : 0;
  res.filter_passes = chebyshev ? passes_total : 0;
  if (ns == 2) {
    res.n_up = n_s[0];
    res.n_down = n_s[1];
    res.magnetization = std::fabs(n_s[0] - n_s[1]);
  }
  nl_x.ensure((size_t)nk * npw * nb * 16);
  expect.ensure((size_t)2 * nk * nb * 8);
  if (nch_all > 0) nl_p.ensure((size_t)nk * nch_all * nb * 16);
  std::vector<double> e2((size_t)2 * nk * nb);
  for (int s = 0; s < ns; ++s) {
    const double2* psi0 = V.as<double2>() + (size_t)s * vblk();
    if (nch_all > 0) {
      launch_zgemm_mfma_batched(c, nch_all, nb, npw, DBHall64.p, nch_all, (long long)nch_all * npw, psi0, npw,
                                (long long)npw * nbs, nl_p.p, nch_all, (long long)nch_all * nb, nk, false, PWK_FP64);
      launch_zgemm_mfma_batched(c, npw, nb, nch_all, Ball64.p, npw, (long long)npw * nch_all, nl_p.p, nch_all,
                                (long long)nch_all * nb, nl_x.p, npw, (long long)npw * nb, nk, false, PWK_FP64);
    } else {
      PWK_HIP(hipMemsetAsync(nl_x.p, 0, (size_t)nk * npw * nb * 16, c.stream));
    }
    k_band_expect<<<dim3((unsigned)nb, (unsigned)nk), 256, 0, c.stream>>>(
        npw, psi0, kin_all64.as<double>(), nl_x.as<double2>(), expect.as<double>(), (size_t)npw * nbs, (size_t)npw,
        (size_t)npw * nb, (size_t)2 * nb);
    PWK_HIP(hipGetLastError());
    PWK_HIP(hipMemcpyAsync(e2.data(), expect.p, e2.size() * 8, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    for (int k = 0; k < nk; ++k) {
      double ek = 0.0, enl = 0.0;
      for (int b = 0; b < nb; ++b) {
        const double f = occ[((size_t)s * nk + k) * nb + b];
        if (f == 0.0) continue;
        ek += f * wk[(size_t)k] * e2[(size_t)2 * k * nb + 2 * b];
        enl += f * wk[(size_t)k] * e2[(size_t)2 * k * nb + 2 * b + 1];
      }
      res.e_kinetic += ek;
      res.e_nonlocal += enl;
    }
  }
}

void Solve::run(double* eig_out, double* rho_out) {
  const auto t0 = std::chrono::steady_clock::now();
  setup();
  const double e_ewald = ewald();
  const int ndim = std::max(1, p.mixing_ndim);  // at most MAX_HIST (validate_scf_request)
  for (Mixer* m : {&mix_tot, &mix_mag}) {
    m->ndim = ndim;
    m->hslot.clear();
    m->next_slot = 0;
  }
  mix_tot.hist.ensure((size_t)2 * ndim * N * 8);
  mix_tot.kerker = p.mixing_mode == 1;
  if (ns == 2) mix_mag.hist.ensure((size_t)2 * ndim * N * 8);  // always plain (driver_spin.rs:76-98)
  res = pwk_scf_result{};
  res.nk = nk;
  res.n_bands = nb;
  res.npw = npw;
  res.nspin = ns;
  for (int d = 0; d < 3; ++d) res.dims[d] = c.dims[d];
  res.e_ewald = e_ewald;
  res.n_ops = (int)ops.size();
  std::vector<double> ev((size_t)ns * nk * nb);
  if (resident_supported())
    run_resident(e_ewald, ev);
  else
    run_host(e_ewald, ev);
  if (eig_out) std::memcpy(eig_out, ev.data(), ev.size() * 8);
  if (res.converged && rho_out) {
    PWK_HIP(hipMemcpyAsync(rho_out, rho_tot_new, N * 8, hipMemcpyDeviceToHost, c.stream));
    c.sync();
  }
  c.sync();
  res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// The iteration with the host between the launches (every eigensolver path and grid): the
// Fermi level, occupations and the DIIS solve on the host, reductions read back as they
// complete.
void Solve::run_host(double e_ewald, std::vector<double>& ev) {
  const double spin = ns == 2 ? 1.0 : 2.0;  // spin_factor = 2/nspin (src/scf/context.rs:188)
  const double dvol = omega / (double)N;
  const int nkb = ns * nk;  // (channel, k) entries, spin-up first
  std::vector<double> wk_all((size_t)nkb);
  for (int e = 0; e < nkb; ++e) wk_all[(size_t)e] = wk[(size_t)(e % nk)];
  lda_in_ready = false;
  g_from_sym = false;
  hist_prestored = false;
  gram_pre.clear();
  bool have_prev = false;
  double e_prev = 0.0;
  for (int iter = 0; iter < p.max_iter; ++iter) {
    // ---- V_eff per channel from the input density ---------------------------------------
    // one spin channel: the Harris-side sums are taken with the output-side ones at the end of the
    // iteration (one reduction launch and one host synchronisation fewer); each term is its own
    // reduction, so the values do not change
    const std::vector<DotTerm> hterms = veff_stage();
    const bool defer = ns == 1;
    double har[4] = {0.0, 0.0, 0.0, 0.0};
    if (!defer) {
      dots(hterms, N);
      std::memcpy(har, host_small, sizeof(har));
    }

    // ---- eigenpairs per channel (both channels in one pass of launches when folded) ----------
    if (fold) {
      for (int s = 0; s < ns; ++s) upload_veff(s, s, false);
      sc = 0;
      ev_host = ev.data();
      eig_chebyshev(iter == 0);
    }
    for (sc = 0; sc < (fold ? 0 : ns); ++sc) {
      upload_veff(sc, 0, !chebyshev);
      double* evc = ev.data() + (size_t)sc * nk * nb;
      if (chebyshev) {
        // Chebyshev-filtered subspace iteration on H·ψ, FP64 Rayleigh-Ritz
        ev_host = evc;
        eig_chebyshev(iter == 0);
      } else {
        // dense H per k, one batched eigensolve
        {
          TimedScope ts(c, "scf_dense_h");
          launch_dense_h_batched(c, nk, kin_all64.as<double>(), nch_all, Ball64.p, DBHall64.p, Hb.p);
        }
        launch_heev_batched(c, npw, nb, nk, Hb.p, c.eig_work.p, W.as<double>() + (size_t)sc * nk * nb, Vc(),
                            nullptr);
        PWK_HIP(hipMemcpyAsync(evc, W.as<double>() + (size_t)sc * nk * nb, (size_t)nk * nb * 8,
                               hipMemcpyDeviceToHost, c.stream));
        c.sync();
      }
    }
    sc = 0;

    // ---- Fermi level(s) and occupations (src/scf/smearing.rs, driver_spin.rs:407-474) ------
    double ef, ef_s[2];
    if (ns == 2 && p.fixed_magnetization) {
      const std::vector<double> eu(ev.begin(), ev.begin() + (long)nk * nb), ed(ev.begin() + (long)nk * nb, ev.end());
      ef_s[0] = fermi_level(eu, nk, nb, wk, 0.5 * (n_el + p.tot_magnetization), p.smearing_sigma, spin);
      ef_s[1] = fermi_level(ed, nk, nb, wk, (n_el - p.tot_magnetization) / 2.0, p.smearing_sigma, spin);
      ef = 0.5 * (ef_s[0] + ef_s[1]);
    } else {
      ef = fermi_level(ev, nkb, nb, wk_all, n_el, p.smearing_sigma, spin);
      ef_s[0] = ef_s[1] = ef;
    }
    std::vector<double> occ((size_t)nkb * nb);
    for (size_t i = 0; i < occ.size(); ++i) occ[i] = fd_occ(ev[i], ef_s[i / ((size_t)nk * nb)], p.smearing_sigma) * spin;
    double n_s[2] = {0.0, 0.0};  // electrons per channel
    for (int s = 0; s < ns; ++s)
      for (int k = 0; k < nk; ++k)
        for (int b = 0; b < nb; ++b) n_s[s] += occ[((size_t)s * nk + k) * nb + b] * wk[(size_t)k];

    // ---- density per channel (src/scf/density.rs:41-111) -----------------------------------
    const int dprec = p.precision == PWK_FP64 ? PWK_FP64 : PWK_FP32;
    const bool dsmall = density_small_supported(c, dprec);
    if (p.precision != PWK_FP64 && !dsmall) launch_convert(c, V.p, V32.p, (size_t)ns * vblk() * 2, true);
    if (dsmall) fw_pin.ensure((size_t)ns * nk * nbs * 8);
    for (int s = 0; s < ns; ++s) {
      double* rn = rho_new_s(s);
      bool waves = false;
      // the basis is the same G-sphere at every k, so all nk·nb bands go through one launch
      std::vector<double> fw((size_t)nk * nbs, 0.0);  // buffer bands carry no weight
      for (int k = 0; k < nk; ++k)
        for (int b = 0; b < nb; ++b)
          fw[(size_t)k * nbs + b] = occ[((size_t)s * nk + k) * nb + b] * wk[(size_t)k];
      {
        TimedScope ts(c, "scf_density");
        if (dsmall) {
          // small cubic grids: the weights from pinned host memory (the previous iteration's
          // density has completed), the FP64 bands read directly
          double* fwp = fw_pin.as<double>() + (size_t)s * nk * nbs;
          std::memcpy(fwp, fw.data(), fw.size() * 8);
          // the group reduction also leaves the wave sums the ∫ρ reduction below is made of
          waves = N <= (size_t)RED_BLOCKS * RED_NT && density_wave_sums(c) <= (size_t)MAX_DOTS * RED_BLOCKS;
          launch_density_small(c, nk * nbs, V.as<double2>() + (size_t)s * vblk(), fwp, rn, dprec, true, true,
                               waves ? host_partial : nullptr);
        } else if (p.precision == PWK_FP64) {
          launch_density(c, nk * nbs, V.as<double2>() + (size_t)s * vblk(), fw.data(), rn, PWK_FP64, true);
        } else {
          launch_density(c, nk * nbs, V32.as<float2>() + (size_t)s * vblk(), fw.data(), rn, PWK_FP32, true);
        }
      }
      double isum;
      if (waves) {
        // Σ ρ in the order of the dot-product reduction (blocks of RED_NT points, each the sum of its
        // waves' tree sums in wave order, blocks in order; one point per thread at this size)
        c.sync();
        const size_t nw = density_wave_sums(c);
        const int nblk = (int)std::max<size_t>(1, (N + RED_NT - 1) / RED_NT);
        isum = 0.0;
        for (int b = 0; b < nblk; ++b) {
          double part = 0.0;
          for (int w = 0; w < RED_NT / 64; ++w) {
            const size_t wi = (size_t)b * (RED_NT / 64) + w;
            part += wi < nw ? host_partial[wi] : 0.0;
          }
          isum += part;
        }
      } else {
        dots({DotTerm{rn, nullptr, nullptr, nullptr, 0}}, N);
        isum = host_small[0];
      }
      const double integral = isum * dvol;
      // normalised to the electron count: by the symmetrisation's first transform when there is
      // one (the same product), else in place
      double scale = 1.0;
      if (std::fabs(integral) > 1e-15) scale = (ns == 2 ? n_s[s] : n_el) / integral;
      if (ops.size() <= 1 && scale != 1.0) {
        k_scale<<<gb(N), 256, 0, c.stream>>>(N, rn, scale);
        PWK_HIP(hipGetLastError());
      }
      symmetrize(rn, scale);
    }

    // ---- output-density energies -------------------------------------------------------------
    const double* outs = host_small;
    if (defer) {
      std::vector<DotTerm> all = hterms;
      fuse_out_lda = N <= (size_t)RED_BLOCKS * RED_NT;
      const std::vector<DotTerm> oterms = output_stage();
      fuse_out_lda = false;
      if (lda_pending) {
        // the mixing step's history entry goes with it (written also when the iteration then
        // converges and does not mix: nothing reads the history after that)
        double* hin = mix_tot.hist.as<double>() + (size_t)mix_tot.peek_slot() * 2 * N;
        lda_pass.hist_in = rho_in.as<double>();
        lda_pass.hin = hin;
        lda_pass.hres = hin + N;
        hist_prestored = true;
        // and, without Kerker preconditioning, the mixing step's sums over the history as it will
        // stand with this entry (its logical order: the older entries, then this one)
        std::vector<int> order = mix_tot.hslot;
        if ((int)order.size() >= mix_tot.ndim) order.erase(order.begin());
        order.push_back(mix_tot.peek_slot());
        const int mm = (int)order.size() - 1;
        if (!mix_tot.kerker && mm >= 1 && mm <= LDA_GRAM_MAX) {
          lda_pass.gram.mm = mm;
          for (int k = 0; k < mm; ++k)
            lda_pass.gram.res[k] = mix_tot.hist.as<double>() + (size_t)order[(size_t)k] * 2 * N + N;
        }
      }
      all.insert(all.end(), oterms.begin(), oterms.end());
      const int mm = lda_pending ? lda_pass.gram.mm : 0;
      dots(all, N);
      std::memcpy(har, host_small, hterms.size() * sizeof(double));
      outs = host_small + hterms.size();
      gram_pre.assign(host_small + all.size(), host_small + all.size() + (mm > 0 ? mm + mm * (mm + 1) / 2 : 0));
    } else {
      dots(output_stage(), N);
    }
    const IterScalars v = scalars(har, outs);
    double e_band = 0.0;
    for (int e = 0; e < nkb; ++e) {
      double sk = 0.0;
      for (int b = 0; b < nb; ++b) sk += occ[(size_t)e * nb + b] * wk_all[(size_t)e] * ev[(size_t)e * nb + b];
      e_band += sk;
    }
    const double ts = smearing_ts(ev, nkb, nb, wk_all, ef, p.smearing_sigma, spin);
    const double e_xc_out = v.exc_int - v.evxc;
    const double e_total = e_band - v.e_h_out + e_xc_out + e_ewald + (-ts);
    const double e_harris = e_band - v.eh_in + v.exc_in + e_ewald + (-ts);
    const double de = have_prev ? std::fabs(e_total - e_prev) : INFINITY;
    const bool e_conv = have_prev && de < p.energy_threshold;
    have_prev = true;
    e_prev = e_total;
    res.n_iter = iter + 1;
    res.final_delta = v.delta;
    last_delta = v.delta;
    if (p.verbose)
      std::fprintf(stderr, "[pwk scf] iter %3d  E=%.10f  dE=%.3e  delta=%.3e  EF=%.6f%s\n", iter, e_total, de, v.delta,
                   ef, ns == 2 ? (std::string("  M=") + std::to_string(n_s[0] - n_s[1])).c_str() : "");
    if (v.delta < p.conv_threshold && e_conv) {
      finalize(occ, v, e_total, e_harris, ef, ts, e_band, n_s);
      break;
    }
    // ---- mixing: ρ (nspin = 1), or ρ_total and m with separate histories (CCMX,
    // src/scf/driver_spin.rs:800-832) ---------------------------------------------------------
    if (ns == 1) {
      lda_in_ready = grid_small_supported(c);
      mix(mix_tot, rho_in.as<double>(), rho_new.as<double>(), rho_in.as<double>(), lda_in_ready, hist_prestored);
      hist_prestored = false;
    } else {
      mix(mix_tot, t_in.as<double>(), t_out.as<double>(), t_in.as<double>());
      mix(mix_mag, m_in.as<double>(), m_out.as<double>(), m_in.as<double>());
      k_spin_split<<<gb(N), 256, 0, c.stream>>>(N, t_in.as<double>(), m_in.as<double>(), rho_in_s(0), rho_in_s(1));
      PWK_HIP(hipGetLastError());
    }
  }
}

// ---------------------------------------------------------------------------------------
// the device-resident iteration
// ---------------------------------------------------------------------------------------
bool Solve::resident_supported() const {
  if (!chebyshev || !c.resident || p.verbose || fprec == PWK_SPLIT_BF16 || !(ns == 1 || fold)) return false;
  const int dprec = p.precision == PWK_FP64 ? PWK_FP64 : PWK_FP32;
  return c.fused_filter && cheb_filter_small_supported(c, fprec, nch_all) &&
         cheb_filter_small_supported(c, PWK_FP64, nch_all) && hpsi_small_supported(c, PWK_FP64, nch_all) &&
         rr_small_supported(nbs) && density_small_supported(c, dprec);
}

// Σ terms on the device into S[off ...): block partials, then their sums in block order
void Solve::dots_dev(const std::vector<DotTerm>& terms, size_t n, int off) {
  const int nblk = (int)std::min<size_t>(RED_BLOCKS, std::max<size_t>(1, (n + RED_NT - 1) / RED_NT));
  for (size_t t0 = 0; t0 < terms.size(); t0 += MAX_TERMS) {
    const size_t nt = std::min(terms.size() - t0, (size_t)MAX_TERMS);
    DotArgs a{};
    a.nterms = (int)nt;
    for (size_t i = 0; i < nt; ++i) a.t[i] = terms[t0 + i];
    k_dots<<<dim3((unsigned)nblk, (unsigned)nt), RED_NT, 0, c.stream>>>(n, a, dpart.as<double>() + t0 * nblk);
  }
  k_dots_final<<<blocks_for(terms.size(), 64), 64, 0, c.stream>>>((int)terms.size(), nblk, dpart.as<double>(),
                                                                   dS.as<double>() + off);
  PWK_HIP(hipGetLastError());
}

// One Anderson step on the device (the host path's mix): history slot and order, the entry,
// the DIIS sums and solve, the combination.
void Solve::mix_dev(int which, const double* in, const double* out, double* result) {
  TimedScope ts(c, "scf_mix");
  Mixer& mx = which ? mix_mag : mix_tot;
  MixState* st = dmix.as<MixState>() + which;
  double* S = dS.as<double>();
  k_hist_advance<<<1, 64, 0, c.stream>>>(st, mx.ndim);
  const double* r = nullptr;
  if (mx.kerker) {
    k_resid<<<gb(N), 256, 0, c.stream>>>(N, in, out, rtmp.as<double>());
    launch_fft3d_r2c(c, rtmp.as<double>(), 1.0, cwork.p, -1);
    k_kerker<<<gb(N), 256, 0, c.stream>>>(N, c.g2_64.as<double>(), kerker_q2, cwork.as<double2>());
    launch_fft3d(c, cwork.p, 1, +1, PWK_FP64);
    k_c2r_scaled<<<gb(N), 256, 0, c.stream>>>(N, cwork.as<double2>(), 1.0 / (double)N, rtmp.as<double>(), 0);
    r = rtmp.as<double>();
  }
  k_hist_put<<<gb(N), 256, 0, c.stream>>>(N, in, out, r, mx.hist.as<double>(), st);
  const int M = mx.ndim - 1;
  double* gram = S + S_GRAM + which * MAX_DOTS;
  if (M > 0) {
    const int nterms = M + M * (M + 1) / 2;
    const int nblk = (int)std::min<size_t>(RED_BLOCKS, std::max<size_t>(1, (N + RED_NT - 1) / RED_NT));
    k_gram<<<dim3((unsigned)nblk, (unsigned)nterms), RED_NT, 0, c.stream>>>(N, M, mx.hist.as<double>(), st,
                                                                            dpart.as<double>());
    k_dots_final<<<blocks_for((size_t)nterms, 64), 64, 0, c.stream>>>(nterms, nblk, dpart.as<double>(), gram);
  }
  k_diis<<<1, 64, 0, c.stream>>>(st, M, gram, S + S_MIXC + which * MAX_HIST);
  k_anderson_dev<<<gb(N), 256, 0, c.stream>>>(N, p.mixing_beta, mx.hist.as<double>(), st,
                                              S + S_MIXC + which * MAX_HIST, result);
  PWK_HIP(hipGetLastError());
}

// One filter pass of every block with device-side control: coefficients from the current Ritz
// values, the one-launch filter, the FP64 H·ψ of the filtered block, Rayleigh-Ritz into the
// device Ritz values, then the residual gate (once it has passed, the following passes of the
// iteration launch but do nothing).
void Solve::pass_dev(int fp) {
  TimedScope ts(c, "scf_chebyshev_filter");
  const bool f64 = fp == PWK_FP64;
  const size_t nkm = (size_t)nke * nbs;
  int* flag = dflag.as<int>();
  const int nkv = fold ? nk : 0;
  launch_cheb_filter_small(c, nke * nbs, nbs, Vc(), Y64.p, fp, f64 ? kin_all64.p : kin_all32.p, nch_all,
                           f64 ? Ball64.p : Ball32.p, f64 ? DBHall64.p : DBHall32.p, coef.as<double>(), degree, nkv,
                           flag);
  launch_hpsi_small(c, nke * nbs, nbs, Y64.p, HY64.p, PWK_FP64, kin_all64.p, nch_all, Ball64.p, DBHall64.p, nullptr,
                    nullptr, nkv, nullptr, flag);
  launch_rr_small(c, npw, nbs, nb, nke, Y64.p, HY64.p, Vc(), lam_dev.as<double>(), lam_dev.as<double>() + nkm, 1e-14,
                  flag);
  gate_dev(true);
}

// the residual gate of the pass that ran and the coefficients of the next one
void Solve::gate_dev(bool gate) {
  TimedScope ts(c, "scf_pass_gate");
  const size_t nkm = (size_t)nke * nbs;
  k_pass_gate<<<1, 256, 0, c.stream>>>(nke, nbs, nb, degree, gate ? 1 : 0, lam_dev.as<double>(),
                                       lam_dev.as<double>() + nkm, ub_dev.as<double>(), dS.as<double>(),
                                       dflag.as<int>(), coef.as<double>());
  PWK_HIP(hipGetLastError());
}

// Everything of one iteration after the first: V_eff, the filter passes, Fermi level and
// occupations, density, symmetrisation, energies, mixing; no host synchronisation.
void Solve::resident_body(int fp, int max_passes) {
  double* S = dS.as<double>();
  k_iter_begin<<<1, 64, 0, c.stream>>>(dflag.as<int>(), S);
  dots_dev(veff_stage(), N, S_HAR);
  for (int s = 0; s < (fold ? ns : 1); ++s) upload_veff(s, s, false);
  for (int pass = 0; pass < max_passes; ++pass) pass_dev(fp);
  resident_tail();
}

// Fermi level to the end of the iteration (shared by the first iteration and the others)
void Solve::resident_tail() {
  double* S = dS.as<double>();
  const double dvol = omega / (double)N;
  FermiArgs fa{};
  fa.nk = nk;
  fa.nb = nb;
  fa.nbs = nbs;
  fa.ns = ns;
  fa.fixed = ns == 2 && p.fixed_magnetization;
  fa.n_el = n_el;
  fa.sigma = p.smearing_sigma;
  fa.spin = ns == 2 ? 1.0 : 2.0;
  fa.tot_mag = p.tot_magnetization;
  fa.lam = lam_dev.as<double>();
  fa.wk = wk_dev.as<double>();
  fa.S = S;
  fa.occ = occ_dev.as<double>();
  fa.fw = fw_dev.as<double>();
  fa.ev = ev_dev.as<double>();
  {
    TimedScope ts(c, "scf_fermi");
    k_fermi<<<1, 256, 0, c.stream>>>(fa);
    PWK_HIP(hipGetLastError());
  }
  const int dprec = p.precision == PWK_FP64 ? PWK_FP64 : PWK_FP32;
  for (int s = 0; s < ns; ++s) {
    double* rn = rho_new_s(s);
    {
      // the FP64 bands are read directly (rounded to FP32 on load for an FP32 density)
      TimedScope ts(c, "scf_density");
      launch_density_small(c, nk * nbs, V.as<double2>() + (size_t)s * vblk(), fw_dev.as<double>() + (size_t)s * nk * nbs,
                           rn, dprec, true, true);
    }
    dots_dev({DotTerm{rn, nullptr, nullptr, nullptr, 0}}, N, S_INT + s);
    k_scale_to<<<gb(N), 256, 0, c.stream>>>(N, rn, S, S_INT + s, dvol, ns == 2 ? S_NS + s : -1, n_el);
    PWK_HIP(hipGetLastError());
    symmetrize(rn);
  }
  dots_dev(output_stage(), N, S_OUT);
  k_iter_end<<<1, 64, 0, c.stream>>>(ns, dvol, omega, S);
  if (ns == 1) {
    mix_dev(0, rho_in.as<double>(), rho_new.as<double>(), rho_in.as<double>());
  } else {
    mix_dev(0, t_in.as<double>(), t_out.as<double>(), t_in.as<double>());
    mix_dev(1, m_in.as<double>(), m_out.as<double>(), m_in.as<double>());
    k_spin_split<<<gb(N), 256, 0, c.stream>>>(N, t_in.as<double>(), m_in.as<double>(), rho_in_s(0), rho_in_s(1));
  }
  k_publish<<<blocks_for((size_t)S_SIZE, 256), 256, 0, c.stream>>>(S_SIZE, S, hS);
  PWK_HIP(hipGetLastError());
}

void Solve::run_resident(double e_ewald, std::vector<double>& ev) {
  const size_t nkm = (size_t)nke * nbs;
  lda_in_ready = false;
  g_from_sym = false;
  dS.ensure((size_t)S_SIZE * 8);
  pin_S.ensure((size_t)S_SIZE * 8);
  hS = pin_S.as<double>();
  dpart.ensure((size_t)MAX_DOTS * RED_BLOCKS * 8);
  dflag.ensure(4 * sizeof(int));
  dmix.ensure(2 * sizeof(MixState));
  lam_dev.ensure(2 * nkm * 8);
  ub_dev.ensure((size_t)nke * 8);
  wk_dev.ensure((size_t)nk * 8);
  fw_dev.ensure((size_t)ns * nk * nbs * 8);
  occ_dev.ensure((size_t)ns * nk * nb * 8);
  ev_dev.ensure((size_t)ns * nk * nb * 8);
  if (mix_tot.kerker) rtmp.ensure(N * 8);
  // zeroed scalars and history state, the filter target of the first iteration (no previous
  // Δρ: the loosest target), the k-point weights
  std::vector<double> s0((size_t)S_SIZE, 0.0);
  s0[S_DELTA] = 1e300;
  const std::vector<MixState> m0(2, MixState{});
  PWK_HIP(hipMemcpyAsync(dS.p, s0.data(), s0.size() * 8, hipMemcpyHostToDevice, c.stream));
  PWK_HIP(hipMemcpyAsync(dmix.p, m0.data(), m0.size() * sizeof(MixState), hipMemcpyHostToDevice, c.stream));
  PWK_HIP(hipMemcpyAsync(wk_dev.p, wk.data(), wk.size() * 8, hipMemcpyHostToDevice, c.stream));

  // the first iteration: spectrum bounds (Lanczos, one host synchronisation), the Rayleigh-Ritz
  // of the start block, up to twelve filter passes
  double* S = dS.as<double>();
  k_iter_begin<<<1, 64, 0, c.stream>>>(dflag.as<int>(), S);
  dots_dev(veff_stage(), N, S_HAR);
  for (int s = 0; s < (fold ? ns : 1); ++s) upload_veff(s, s, false);
  sc = 0;
  lanczos_bounds();
  PWK_HIP(hipMemcpyAsync(ub_dev.p, ub.data(), (size_t)nke * 8, hipMemcpyHostToDevice, c.stream));
  PWK_HIP(hipMemcpyAsync(Y64.p, Vc(), eblk() * 16, hipMemcpyDeviceToDevice, c.stream));
  launch_hpsi_small(c, nke * nbs, nbs, Y64.p, HY64.p, PWK_FP64, kin_all64.p, nch_all, Ball64.p, DBHall64.p, nullptr,
                    nullptr, fold ? nk : 0);
  launch_rr_small(c, npw, nbs, nb, nke, Y64.p, HY64.p, Vc(), lam_dev.as<double>(), lam_dev.as<double>() + nkm, 1e-14);
  gate_dev(false);  // coefficients of the first pass
  // the first iteration launches directly, so its passes stop at the gate as on the host path
  for (int pass = 0; pass < 12; ++pass) {
    pass_dev(fprec);
    k_publish<<<1, 64, 0, c.stream>>>(S_PASSED + 1, S, hS);
    PWK_HIP(hipGetLastError());
    c.sync();
    if (hS[S_PASSED] != 0.0) break;
  }
  resident_tail();

  bool have_prev = false;
  double e_prev = 0.0;
  for (int iter = 0; iter < p.max_iter; ++iter) {
    if (iter > 0) {
      // a reduced-precision filter hands over to FP64 once the residual target (0.1 Δρ) falls
      // below 3e-6, as on the host path
      const double target = std::max(1e-9, std::min(1e-3, 0.1 * last_delta));
      const int fp = (fprec != PWK_FP64 && target < 3e-6) ? PWK_FP64 : fprec;
      launch_iteration(fp, pass_budget);
    }
    c.sync();
    // passes launched in the next iteration: one more than this one needed (at least two), all six
    // of the host path when this iteration's gate did not pass; a pass after the gate has passed
    // is an empty launch, so the budget follows what the iterations need
    const int used = (int)hS[S_PASSES];
    pass_budget = hS[S_PASSED] != 0.0 ? std::min(6, std::max(2, used + 1)) : 6;
    passes_total += used;
    const IterScalars v = scalars(hS + S_HAR, hS + S_OUT);
    const double e_band = hS[S_EBAND], ts = hS[S_TS], ef = hS[S_EF];
    const double e_xc_out = v.exc_int - v.evxc;
    const double e_total = e_band - v.e_h_out + e_xc_out + e_ewald + (-ts);
    const double e_harris = e_band - v.eh_in + v.exc_in + e_ewald + (-ts);
    const double de = have_prev ? std::fabs(e_total - e_prev) : INFINITY;
    const bool e_conv = have_prev && de < p.energy_threshold;
    have_prev = true;
    e_prev = e_total;
    res.n_iter = iter + 1;
    res.final_delta = v.delta;
    last_delta = v.delta;
    const bool done = v.delta < p.conv_threshold && e_conv;
    if (done || iter + 1 == p.max_iter) {
      std::vector<double> occ(ev.size());
      PWK_HIP(hipMemcpyAsync(ev.data(), ev_dev.p, ev.size() * 8, hipMemcpyDeviceToHost, c.stream));
      PWK_HIP(hipMemcpyAsync(occ.data(), occ_dev.p, occ.size() * 8, hipMemcpyDeviceToHost, c.stream));
      c.sync();
      if (done) finalize(occ, v, e_total, e_harris, ef, ts, e_band, hS + S_NS);
      break;
    }
  }
}

// An iteration after the first: launched directly, or replayed from a HIP graph captured
// once per filter precision and pass budget (after a direct run has set every kernel attribute
// and sized every buffer, so nothing is allocated while capturing).
void Solve::launch_iteration(int fp, int passes) {
  const bool graph = c.graphs && !c.prof.on;
  if (!graph) {
    resident_body(fp, passes);
    return;
  }
  const int key = fp * 64 + passes;  // one graph per filter precision and pass budget
  auto it = iter_graphs.find(key);
  if (it != iter_graphs.end()) {
    PWK_HIP(hipGraphLaunch(it->second, c.stream));
    return;
  }
  if (!iter_warm.count(key)) {
    iter_warm.insert(key);
    resident_body(fp, passes);
    return;
  }
  hipGraph_t g = nullptr;
  hipGraphExec_t ex = nullptr;
  PWK_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
  try {
    resident_body(fp, passes);
  } catch (...) {
    (void)hipStreamEndCapture(c.stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  PWK_HIP(hipStreamEndCapture(c.stream, &g));
  const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  PWK_HIP(ie);
  iter_graphs[key] = ex;
  PWK_HIP(hipGraphLaunch(ex, c.stream));
}

}  // namespace

// The filter passes of n solves: filter, FP64 H·ψ and Rayleigh-Ritz over every solve's k-point
// blocks, three launches in all (the launches Solve::chebyshev_pass makes for one solve on the
// fused small-grid path).
void launch_pass_batch(hipStream_t stream, const PassShape& sh, const PassReq* const* reqs, int n) {
  PWK_REQUIRE(n >= 1 && n <= MAX_PASS_BATCH, PWK_E_INVALID, "bad pass batch size");
  PassSmallMulti f{}, h{};
  PassRrMulti r{};
  int max_bands = 1, max_nk = 1;
  for (int i = 0; i < n; ++i) {
    f.r[i] = reqs[i]->filt;
    h.r[i] = reqs[i]->hpsi;
    r.r[i] = reqs[i]->rr;
    max_bands = std::max(max_bands, reqs[i]->filt.nbands);
    max_nk = std::max(max_nk, reqs[i]->rr.nk);
  }
  launch_filter_multi(stream, sh, f, n, max_bands);
  launch_hpsi64_multi(stream, sh, h, n, max_bands);
  launch_rr_multi(stream, sh, r, n, max_nk, 1e-14);
}

void scf_run(Context& c, const pwk_crystal& cr, const pwk_scf_params& p, pwk_scf_result& out, double* eig,
             double* rho) {
  validate_scf_request(cr, p);
  Solve s(c, cr, p, out);
  s.run(eig, rho);
}

// Checks a crystal and its parameters before any table is built from them (ScfParams::validate,
// src/scf/mod.rs:166-262, and the crystal checks of run_scf, :354-372): every pointer the solve
// dereferences, the species range, the radial meshes and the scalar parameters.
void validate_scf_request(const pwk_crystal& cr, const pwk_scf_params& p) {
  PWK_REQUIRE(cr.natoms >= 1 && cr.frac && cr.species && cr.species_data, PWK_E_INVALID,
              "bad crystal description (atoms, positions, species and species data are required)");
  PWK_REQUIRE(cr.nspecies >= 1 && cr.nspecies <= MAXSP, PWK_E_INVALID, "nspecies must be in 1..8");
  for (int a = 0; a < cr.natoms; ++a) {
    PWK_REQUIRE(cr.species[a] >= 0 && cr.species[a] < cr.nspecies, PWK_E_INVALID, "atom species out of range");
    for (int d = 0; d < 3; ++d) PWK_REQUIRE(std::isfinite(cr.frac[3 * a + d]), PWK_E_INVALID, "non-finite position");
  }
  for (int i = 0; i < 9; ++i) PWK_REQUIRE(std::isfinite(cr.lattice[i]), PWK_E_INVALID, "non-finite lattice");
  for (int s = 0; s < cr.nspecies; ++s) {
    const pwk_species& sp = cr.species_data[s];
    PWK_REQUIRE(sp.mesh >= 1 && sp.r && sp.rab && sp.v_local, PWK_E_INVALID, "species without radial mesh");
    PWK_REQUIRE(sp.nproj >= 0, PWK_E_INVALID, "negative projector count");
    PWK_REQUIRE(sp.nproj == 0 || (sp.proj_l && sp.beta && sp.dij), PWK_E_INVALID, "projectors without data");
    for (int i = 0; i < sp.nproj; ++i)
      PWK_REQUIRE(sp.proj_l[i] >= 0 && sp.proj_l[i] <= 3, PWK_E_UNSUPPORTED, "projector l must be in 0..3");
    PWK_REQUIRE(std::isfinite(sp.z_valence) && sp.z_valence >= 0.0, PWK_E_INVALID, "bad valence charge");
  }
  PWK_REQUIRE(std::isfinite(p.ecut_ev) && p.ecut_ev > 0.0, PWK_E_INVALID, "ecut must be positive");
  PWK_REQUIRE(p.n_bands >= 1, PWK_E_INVALID, "n_bands must be positive");
  PWK_REQUIRE(p.max_iter >= 1, PWK_E_INVALID, "max_iter must be positive");
  PWK_REQUIRE(p.conv_threshold > 0.0, PWK_E_INVALID, "conv_threshold must be positive");
  PWK_REQUIRE(p.mixing_beta > 0.0 && p.mixing_beta <= 1.0, PWK_E_INVALID, "mixing_beta must be in (0, 1]");
  PWK_REQUIRE(p.smearing_sigma >= 0.0, PWK_E_INVALID, "smearing_sigma must be non-negative");
  PWK_REQUIRE(p.precision == PWK_FP64 || p.precision == PWK_FP32 || p.precision == PWK_SPLIT_BF16, PWK_E_INVALID,
              "unknown precision");
  PWK_REQUIRE(p.nspin >= 0 && p.nspin <= 2, PWK_E_INVALID, "nspin must be 1 or 2");
  PWK_REQUIRE(p.mixing_mode == 0 || p.mixing_mode == 1, PWK_E_INVALID, "mixing_mode must be 0 (plain) or 1 (Kerker)");
  PWK_REQUIRE(p.mixing_ndim <= MAX_HIST, PWK_E_INVALID, "mixing_ndim above the 32 history entries the mixer holds");
  PWK_REQUIRE(p.eigensolver >= -1 && p.eigensolver <= 1, PWK_E_INVALID, "eigensolver must be -1, 0 or 1");
  PWK_REQUIRE(p.ecutrho_ratio >= 0 && p.cheb_degree >= 0 && p.cheb_degree <= 64, PWK_E_INVALID,
              "bad ecutrho_ratio or cheb_degree");
  for (int d = 0; d < 3; ++d)
    PWK_REQUIRE(p.kgrid[d] >= 0 && p.kgrid[d] <= 64 && p.fft_dims[d] >= 0 && p.fft_dims[d] <= 1024, PWK_E_INVALID,
                "bad k-grid or FFT dims");
}

void scf_tables(const pwk_crystal& cr, const pwk_scf_params& p, int* npw, int dims[3], int* nk, int* nops) {
  validate_scf_request(cr, p);
  Lat lat;
  for (int i = 0; i < 3; ++i) lat.v[i] = {cr.lattice[3 * i], cr.lattice[3 * i + 1], cr.lattice[3 * i + 2]};
  std::vector<int32_t> miller;
  std::vector<Vec3> g;
  sphere_basis(lat, p.ecut_ev, miller, g);
  *npw = (int)g.size();
  grid_dims(miller, p.ecutrho_ratio > 0 ? p.ecutrho_ratio : 4, p.fft_dims, dims);
  std::vector<AtomSite> sites((size_t)cr.natoms);
  for (int a = 0; a < cr.natoms; ++a) {
    sites[(size_t)a].z = cr.species_data[cr.species[a]].atomic_number;
    for (int d = 0; d < 3; ++d) sites[(size_t)a].frac[d] = cr.frac[3 * a + d];
  }
  std::vector<SymOp> ops;
  if (p.use_symmetry) ops = detect_symmetry(lat, sites, p.symprec > 0.0 ? p.symprec : 1e-5);
  const bool trev = p.use_symmetry != 0;
  if (ops.empty()) {
    SymOp id{};
    id.rot[0] = id.rot[4] = id.rot[8] = 1;
    ops.push_back(id);
  }
  int grid[3], shift[3];
  for (int d = 0; d < 3; ++d) {
    grid[d] = p.kgrid[d] > 0 ? p.kgrid[d] : 1;
    shift[d] = p.kshift[d];
  }
  *nk = (int)kpoints(lat, grid, shift, p.use_symmetry == 2 ? identity_ops() : ops, p.use_symmetry == 2 ? false : trev)
            .size();
  *nops = (int)ops.size();
}

}  // namespace pwk
