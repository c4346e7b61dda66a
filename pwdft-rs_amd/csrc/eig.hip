// Batched dense Hermitian eigensolver, FP64, one workgroup per matrix: the
// lowest nb eigenpairs of each n×n matrix, eigenvalues ascending, as
// diagonalize_lowest (src/eigensolver/dense.rs:87-98), which keeps the first
// n_bands pairs of a full self-adjoint decomposition.
//
//   1. Householder reduction to Hermitian tridiagonal form; the reflectors are
//      kept below the diagonal of the (overwritten) input, the trailing
//      rank-2 updates are spread over the workgroup.
//   2. A diagonal unitary D makes the sub-diagonal real: T = D·T_r·D^H.
//   3. The nb lowest eigenvalues of T_r by Sturm-count multisection, lane
//      groups of one wavefront per eigenvalue, one probe point per lane.
//   4. Their eigenvectors by inverse iteration (tridiagonal LU with partial
//      pivoting, one lane per vector, three solves from a fixed pseudo-random
//      start; factors in LDS when they fit), then modified Gram-Schmidt inside
//      clusters of close eigenvalues.
//   5. Back-transformation V = Q·D·Z through the reflectors in reverse order.
// Matrices of at most 32 rows (the Rayleigh-Ritz problems of the filtered
// eigensolver) take a different route: a one-wavefront cyclic Jacobi solver
// with the matrix and the eigenvectors in LDS (round-robin ordering, n/2
// disjoint complex rotations per round, at most 30 sweeps).
// Every loop has a fixed bound; nothing waits on data.
#include "common.hpp"
#include "jacobi.hpp"

namespace pwk {

namespace {

// Two workgroup shapes: 16 waves (memory-level parallelism for the latency-bound passes
// of larger matrices) and one wave for matrices of at most 64 rows, where barriers
// dominate and a single wave needs none of its reductions in LDS.
constexpr int EIG_NT = 1024;
constexpr int EIG_WAVES = EIG_NT / 64;  // LDS is sized for the larger shape

__host__ __device__ inline size_t eig_stride(int n, int nb) {
  // ecx (2n) | dph (2n) | z (n·nb) | LU factors (5·n·nb)
  size_t s = 4 * (size_t)n + 6 * (size_t)n * nb;
  return (s + 1) & ~(size_t)1;
}

template <int NT>
__device__ double block_sum(double v, double* red) {
  constexpr int EIG_WAVES = NT / 64;
  const int t = threadIdx.x;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if (NT == 64) return __shfl(v, 0, 64);
  __syncthreads();
  if ((t & 63) == 0) red[t >> 6] = v;
  __syncthreads();
  if (t == 0) {
    double s = 0.0;
    for (int w = 0; w < EIG_WAVES; ++w) s += red[w];
    red[EIG_WAVES] = s;
  }
  __syncthreads();
  return red[EIG_WAVES];
}

// Number of eigenvalues of the real symmetric tridiagonal (d, e) below x.
__device__ int sturm_count(int n, const double* d, const double* e2, double x, double pivmin) {
  int cnt = 0;
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  cnt += q < 0.0;
  for (int i = 1; i < n; ++i) {
    q = d[i] - x - e2[i - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += q < 0.0;
  }
  return cnt;
}

// Deterministic start vector entry in [-1, 1).
__device__ double start_entry(int i, int j) {
  uint32_t h = (uint32_t)i * 0x9E3779B1u ^ ((uint32_t)j + 0x7F4A7C15u) * 0x85EBCA77u;
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  h *= 0x297A2D39u;
  h ^= h >> 15;
  return (double)h * (2.0 / 4294967296.0) - 1.0;
}

template <int NT>
__global__ void __launch_bounds__(NT) k_heev(int n, int nb, int a_in_lds, int lu_in_lds, double2* __restrict__ Hb,
                                             double* __restrict__ work, double* __restrict__ w_out,
                                             double2* __restrict__ V_out, int* __restrict__ status) {
  constexpr int EIG_NT = NT;
  constexpr int EIG_WAVES = NT / 64;
  extern __shared__ unsigned char smem[];
  double2* v = reinterpret_cast<double2*>(smem);      // n
  double2* wv = v + n;                                // n
  double* d = reinterpret_cast<double*>(wv + n);      // n: diagonal of T_r
  double* e = d + n;                                  // n: sub-diagonal of T_r
  double* e2 = e + n;                                 // n: e²
  double* lam = e2 + n;                               // nb: eigenvalues
  double* red = lam + nb;                             // EIG_WAVES + 2
  // EIG_WAVES × 64 matvec partials, 16-byte aligned
  const size_t pp_off = ((size_t)(reinterpret_cast<unsigned char*>(red + EIG_WAVES + 2) - smem) + 15) & ~(size_t)15;
  double2* pp = reinterpret_cast<double2*>(smem + pp_off);
  double2* lds_a = pp + 16 * 64;                                          // n·n when a_in_lds
  double* lds_lu = reinterpret_cast<double*>(lds_a + (a_in_lds ? (size_t)n * n : 0));  // 6·n·nb when lu_in_lds
  __shared__ double s_bounds[3];

  const int b = blockIdx.x, t = threadIdx.x, lane = t & 63, wave = t >> 6;
  double2* A = Hb + (size_t)b * n * n;
  if (a_in_lds) {  // small matrices live in LDS for the whole solve (the input is not written back)
    for (int i = t; i < n * n; i += EIG_NT) lds_a[i] = A[i];
    __syncthreads();
    A = lds_a;
  }
  double* base = work + (size_t)b * eig_stride(n, nb);
  double2* ecx = reinterpret_cast<double2*>(base);    // n
  double2* dph = ecx + n;                              // n
  double* Z = lu_in_lds ? lds_lu : reinterpret_cast<double*>(dph + n);  // Z[i*nb + j]
  double* U0 = Z + (size_t)n * nb;                     // LU factors, interleaved by vector like Z
  double* U1 = U0 + (size_t)n * nb;
  double* U2 = U1 + (size_t)n * nb;
  double* ML = U2 + (size_t)n * nb;
  double* PV = ML + (size_t)n * nb;

  // ---- 1. Householder tridiagonalisation --------------------------------------------
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;
    double2* col = A + (size_t)k * n + k + 1;
    double part = 0.0;
    for (int i = t; i < m; i += EIG_NT) part += col[i].x * col[i].x + col[i].y * col[i].y;
    const double xn = sqrt(block_sum<NT>(part, red));
    if (xn == 0.0) {
      for (int i = t; i < m; i += EIG_NT) col[i] = make_double2(0.0, 0.0);  // identity reflector
      if (t == 0) ecx[k] = make_double2(0.0, 0.0);
      __syncthreads();
      continue;
    }
    const double2 x0 = col[0];
    const double ax0 = sqrt(x0.x * x0.x + x0.y * x0.y);
    const double2 ph = ax0 > 0.0 ? make_double2(x0.x / ax0, x0.y / ax0) : make_double2(1.0, 0.0);
    const double2 alpha = make_double2(-ph.x * xn, -ph.y * xn);
    // v = (x - alpha e1) / |x - alpha e1|;  |x - alpha e1|² = 2(|x|² + |x||x0|)
    const double vn = sqrt(2.0 * (xn * xn + xn * ax0));
    for (int i = t; i < m; i += EIG_NT) {
      double2 x = col[i];
      if (i == 0) {
        x.x -= alpha.x;
        x.y -= alpha.y;
      }
      v[i] = make_double2(x.x / vn, x.y / vn);
    }
    __syncthreads();
    // p = A_sub v: rows in chunks of 64 lanes, the columns split over the waves, then a
    // fixed-order sum of the per-wave partials
    for (int i0 = 0; i0 < m; i0 += 64) {
      const int i = i0 + lane;
      double pr = 0.0, pi = 0.0;
      if (i < m) {
        const double2* arow = A + (size_t)(k + 1) * n + k + 1 + i;
        for (int j = wave; j < m; j += EIG_WAVES) {
          const double2 a = arow[(size_t)j * n];
          const double2 vj = v[j];
          pr += a.x * vj.x - a.y * vj.y;
          pi += a.x * vj.y + a.y * vj.x;
        }
      }
      pp[wave * 64 + lane] = make_double2(pr, pi);
      __syncthreads();
      if (wave == 0 && i < m) {
        double sr = 0.0, si = 0.0;
        for (int w = 0; w < EIG_WAVES; ++w) {
          sr += pp[w * 64 + lane].x;
          si += pp[w * 64 + lane].y;
        }
        wv[i] = make_double2(sr, si);
      }
      __syncthreads();
    }
    part = 0.0;
    for (int i = t; i < m; i += EIG_NT) part += v[i].x * wv[i].x + v[i].y * wv[i].y;  // v^H p (real)
    const double K = block_sum<NT>(part, red);
    for (int i = t; i < m; i += EIG_NT) {
      wv[i].x -= K * v[i].x;
      wv[i].y -= K * v[i].y;
    }
    __syncthreads();
    // A_sub -= 2 (v w^H + w v^H)
    for (int idx = t; idx < m * m; idx += EIG_NT) {
      const int i = idx % m, j = idx / m;
      const double2 vi = v[i], wi = wv[i], vj = v[j], wj = wv[j];
      const double re = (vi.x * wj.x + vi.y * wj.y) + (wi.x * vj.x + wi.y * vj.y);
      const double im = (vi.y * wj.x - vi.x * wj.y) + (wi.y * vj.x - wi.x * vj.y);
      double2* a = &A[(size_t)(k + 1 + j) * n + k + 1 + i];
      a->x -= 2.0 * re;
      a->y -= 2.0 * im;
    }
    __syncthreads();
    for (int i = t; i < m; i += EIG_NT) col[i] = v[i];
    if (t == 0) ecx[k] = alpha;
    __syncthreads();
  }
  if (t == 0 && n >= 2) ecx[n - 2] = A[(size_t)(n - 2) * n + n - 1];
  for (int i = t; i < n; i += EIG_NT) d[i] = A[(size_t)i * n + i].x;
  __syncthreads();

  // ---- 2. real sub-diagonal, phases, Gershgorin bounds -------------------------------
  if (t == 0) {
    double2 p = make_double2(1.0, 0.0);
    dph[0] = p;
    double lo = 0.0, hi = 0.0, emax2 = 0.0;
    for (int k = 0; k < n; ++k) {
      double ab = 0.0;
      if (k + 1 < n) {
        const double2 ek = ecx[k];
        ab = sqrt(ek.x * ek.x + ek.y * ek.y);
        if (ab > 0.0) p = make_double2((p.x * ek.x - p.y * ek.y) / ab, (p.x * ek.y + p.y * ek.x) / ab);
        dph[k + 1] = p;
      }
      e[k] = ab;
      e2[k] = ab * ab;
      emax2 = fmax(emax2, ab * ab);
      const double r = ab + (k > 0 ? e[k - 1] : 0.0);
      lo = k == 0 ? d[k] - r : fmin(lo, d[k] - r);
      hi = k == 0 ? d[k] + r : fmax(hi, d[k] + r);
    }
    const double tnorm = fmax(fabs(lo), fabs(hi));
    const double pad = 2.220446049250313e-16 * tnorm * (double)n + 1e-300;
    s_bounds[0] = lo - pad;
    s_bounds[1] = hi + pad;
    s_bounds[2] = fmax(tnorm, 1e-300);
    (void)emax2;
  }
  __syncthreads();
  const double tnorm = s_bounds[2];
  const double pivmin = 1e-290 * fmax(1.0, tnorm * tnorm);

  // ---- 3. eigenvalues by multisection: groups of G lanes per eigenvalue, each lane one
  // probe; the groups of all waves cover as many eigenvalues at once as there are groups
  {
    int gpw = 1;  // groups per wave: the smallest power of two with NW·gpw >= nb (at most 64)
    while (gpw < 64 && gpw * EIG_WAVES < nb) gpw <<= 1;
    const int G = 64 / gpw, g = lane / G, sub = lane % G;
    const unsigned long long gmask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
    const int per_round = EIG_WAVES * gpw;
    for (int j0 = 0; j0 < nb; j0 += per_round) {
      const int j = j0 + wave * gpw + g;
      const bool active = j < nb;
      double a = s_bounds[0], bb = s_bounds[1];
      bool done = !active;
      for (int it = 0; it < 64; ++it) {
        if (__ballot(!done) == 0ull) break;  // wave-uniform exit
        const double x = a + (bb - a) * ((double)(sub + 1) / (double)(G + 1));
        const int cnt = done ? 0 : sturm_count(n, d, e2, x, pivmin);
        const unsigned long long le = (__ballot(cnt <= j) >> (g * G)) & gmask;
        // first probe of the group above the j-th eigenvalue
        const int k = (~le & gmask) == 0ull ? G : __builtin_ctzll(~le & gmask);
        const double na = __shfl(x, g * G + (k > 0 ? k - 1 : 0), 64);
        const double nbv = __shfl(x, g * G + (k < G ? k : G - 1), 64);
        if (!done) {
          const double a2 = k > 0 ? na : a, b2 = k < G ? nbv : bb;
          // converged to the precision the tridiagonal form carries: eps relative to the
          // eigenvalue or eps·|T| absolute, whichever is larger
          done = (a2 == a && b2 == bb) ||
                 b2 - a2 <= 4.440892098500626e-16 * fmax(fmax(fabs(a2), fabs(b2)), tnorm);
          a = a2;
          bb = b2;
        }
      }
      if (active && sub == 0) lam[j] = 0.5 * (a + bb);
    }
  }
  __syncthreads();

  // ---- 4. inverse iteration, one lane per eigenvector ------------------------------------
  const double tiny = 2.220446049250313e-16 * tnorm;
  for (int j = t; j < nb; j += EIG_NT) {
    const double lj = lam[j];
    // LU of T_r - lj I with row interchanges (upper factor has two super-diagonals)
    double cd = d[0] - lj, cs = n > 1 ? e[0] : 0.0;
    for (int i = 0; i + 1 < n; ++i) {
      const double sub = e[i];
      const double an = d[i + 1] - lj, bn = (i + 2 < n) ? e[i + 1] : 0.0;
      double u0, u1, u2, ml, pv;
      if (fabs(cd) >= fabs(sub)) {
        if (fabs(cd) < tiny) cd = cd < 0.0 ? -tiny : tiny;
        ml = sub / cd;
        u0 = cd;
        u1 = cs;
        u2 = 0.0;
        pv = 0.0;
        cd = an - ml * cs;
        cs = bn;
      } else {
        ml = cd / sub;
        u0 = sub;
        u1 = an;
        u2 = bn;
        pv = 1.0;
        cd = cs - ml * an;
        cs = -ml * bn;
      }
      const size_t o = (size_t)i * nb + j;
      U0[o] = u0;
      U1[o] = u1;
      U2[o] = u2;
      ML[o] = ml;
      PV[o] = pv;
    }
    if (fabs(cd) < tiny) cd = cd < 0.0 ? -tiny : tiny;
    {
      const size_t o = (size_t)(n - 1) * nb + j;
      U0[o] = cd;
      U1[o] = 0.0;
      U2[o] = 0.0;
      ML[o] = 0.0;
      PV[o] = 0.0;
    }
    // start vector
    double nrm = 0.0;
    for (int i = 0; i < n; ++i) {
      const double s = start_entry(i, j);
      Z[(size_t)i * nb + j] = s;
      nrm += s * s;
    }
    for (int sweep = 0; sweep < 3; ++sweep) {
      const double sc = 1.0 / sqrt(nrm);
      // forward: row interchanges and multipliers
      double carry = Z[j] * sc;
      for (int i = 0; i + 1 < n; ++i) {
        const size_t o = (size_t)i * nb + j;
        double yn = Z[o + nb] * sc;
        double yi = carry;
        if (PV[o] != 0.0) {
          const double tmp = yi;
          yi = yn;
          yn = tmp;
        }
        Z[o] = yi;
        carry = yn - ML[o] * yi;
      }
      Z[(size_t)(n - 1) * nb + j] = carry;
      // backward substitution
      double x2 = 0.0, x1 = 0.0;
      nrm = 0.0;
      for (int i = n - 1; i >= 0; --i) {
        const size_t o = (size_t)i * nb + j;
        const double xi = (Z[o] - U1[o] * x1 - U2[o] * x2) / U0[o];
        Z[o] = xi;
        nrm += xi * xi;
        x2 = x1;
        x1 = xi;
      }
      if (!(nrm > 0.0) || !isfinite(nrm)) nrm = 1.0;
    }
    const double sc = 1.0 / sqrt(nrm);
    for (int i = 0; i < n; ++i) Z[(size_t)i * nb + j] *= sc;
  }
  __syncthreads();
  // modified Gram-Schmidt inside clusters (gap below 1e-3·|T|)
  const double ortol = 1e-3 * tnorm;
  for (int j = 1; j < nb; ++j) {
    int first = j;
    while (first > 0 && lam[first] - lam[first - 1] <= ortol) --first;
    if (first == j) continue;
    for (int i = first; i < j; ++i) {
      double part = 0.0;
      for (int r = t; r < n; r += EIG_NT) part += Z[(size_t)r * nb + i] * Z[(size_t)r * nb + j];
      const double dp = block_sum<NT>(part, red);
      for (int r = t; r < n; r += EIG_NT) Z[(size_t)r * nb + j] -= dp * Z[(size_t)r * nb + i];
      __syncthreads();
    }
    double part = 0.0;
    for (int r = t; r < n; r += EIG_NT) part += Z[(size_t)r * nb + j] * Z[(size_t)r * nb + j];
    const double nn = block_sum<NT>(part, red);
    const double sc = nn > 0.0 ? 1.0 / sqrt(nn) : 0.0;
    for (int r = t; r < n; r += EIG_NT) Z[(size_t)r * nb + j] *= sc;
    __syncthreads();
  }

  // ---- 5. V = Q·D·Z --------------------------------------------------------------------
  double2* V = V_out + (size_t)b * n * nb;
  for (int idx = t; idx < n * nb; idx += EIG_NT) {
    const int r = idx % n, jj = idx / n;
    const double zr = Z[(size_t)r * nb + jj];
    const double2 p = dph[r];
    V[(size_t)jj * n + r] = make_double2(p.x * zr, p.y * zr);
  }
  for (int jj = t; jj < nb; jj += EIG_NT) w_out[(size_t)b * nb + jj] = lam[jj];
  __syncthreads();
  double2* sv = wv;  // nb ≤ n projections
  for (int k = n - 3; k >= 0; --k) {
    const int m = n - k - 1;
    const double2* vk = A + (size_t)k * n + k + 1;
    // one wave per vector, lanes over the reflector
    for (int jj = wave; jj < nb; jj += EIG_WAVES) {
      double sr = 0.0, si = 0.0;
      const double2* x = V + (size_t)jj * n + k + 1;
      for (int i = lane; i < m; i += 64) {
        const double2 a = vk[i], xi = x[i];
        sr += a.x * xi.x + a.y * xi.y;  // conj(v)·x
        si += a.x * xi.y - a.y * xi.x;
      }
      for (int off = 32; off > 0; off >>= 1) {
        sr += __shfl_down(sr, off, 64);
        si += __shfl_down(si, off, 64);
      }
      if (lane == 0) sv[jj] = make_double2(sr, si);
    }
    __syncthreads();
    for (int idx = t; idx < m * nb; idx += EIG_NT) {
      const int i = idx % m, jj = idx / m;
      const double2 a = vk[i], s = sv[jj];
      double2* x = &V[(size_t)jj * n + k + 1 + i];
      x->x -= 2.0 * (a.x * s.x - a.y * s.y);
      x->y -= 2.0 * (a.x * s.y + a.y * s.x);
    }
    __syncthreads();
  }
  if (t == 0 && status) status[b] = 0;
}

// One wavefront per matrix, n <= JMAX: cyclic Jacobi with the matrix and the vectors in LDS.
__global__ void __launch_bounds__(64) k_jacobi(int n, int nb, const double2* __restrict__ Hb, double* __restrict__ w_out,
                                               double2* __restrict__ V_out) {
  __shared__ JacobiLds J;
  const int b = blockIdx.x, lane = threadIdx.x;
  const double2* H = Hb + (size_t)b * n * n;
  for (int e = lane; e < n * n; e += 64) {
    const int r = e % n, col = e / n;
    J.A[r][col] = H[e];
    J.V[r][col] = make_double2(r == col ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  jacobi_lds(n, J);
  for (int jj = lane; jj < nb; jj += 64) w_out[(size_t)b * nb + jj] = J.A[J.order[jj]][J.order[jj]].x;
  for (int e = lane; e < n * nb; e += 64) {
    const int r = e % n, jj = e / n;
    V_out[(size_t)b * n * nb + e] = J.V[r][J.order[jj]];
  }
}

size_t heev_lds_base(int n, int nb) {
  size_t b = (size_t)2 * n * sizeof(double2) + ((size_t)3 * n + nb + EIG_WAVES + 2) * sizeof(double);
  b = (b + 15) & ~(size_t)15;
  return b + (size_t)16 * 64 * sizeof(double2);
}
constexpr size_t kLdsCap = 150 * 1024;
size_t a_bytes(int n) { return (size_t)n * n * sizeof(double2); }
size_t lu_bytes(int n, int nb) { return (size_t)6 * n * nb * sizeof(double); }
bool a_fits_lds(int n, int nb) { return heev_lds_base(n, nb) + a_bytes(n) <= kLdsCap; }
bool lu_fits_lds(int n, int nb) {
  return heev_lds_base(n, nb) + (a_fits_lds(n, nb) ? a_bytes(n) : 0) + lu_bytes(n, nb) <= kLdsCap;
}
size_t heev_lds_bytes(int n, int nb) {
  return heev_lds_base(n, nb) + (a_fits_lds(n, nb) ? a_bytes(n) : 0) + (lu_fits_lds(n, nb) ? lu_bytes(n, nb) : 0);
}

}  // namespace

size_t heev_work_doubles(int n, int nb, int nk) { return (size_t)nk * eig_stride(n, nb); }

// ---- the allocation-free form -----------------------------------------------------------
// Workspace of pwk_dev_dense_eigh_noalloc: one static device array per GPU (16 MiB, held from
// the first use of the library on that GPU until the process ends). Host code never passes the
// array's name to a kernel: its device address is looked up once per device, with that device
// current, when a context is created on it, and only that address is used afterwards. Every
// call on one GPU shares the array; the eigensolver writes each workspace entry before reading
// it, so nothing needs clearing between calls, and calls that overlap on one GPU race.
constexpr size_t kNoAllocEigDoubles = (size_t)1 << 21;
static __device__ __align__(256) double scratch[kNoAllocEigDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

void init_noalloc_scratch(int device) {
  PWK_REQUIRE(device >= 0 && device < kMaxDevices, PWK_E_NO_DEVICE, "device ordinal above 63");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  PWK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  PWK_REQUIRE(p != nullptr, PWK_E_DEVICE, "static eigensolver workspace has no device address");
  g_scratch_ptr[device] = static_cast<double*>(p);
}

size_t noalloc_heev_capacity_doubles() { return kNoAllocEigDoubles; }

void launch_heev_noalloc(Context& c, int n, int nb, int nk, void* H, double* w, void* V) {
  PWK_REQUIRE(n >= 1 && nb >= 1 && nb <= n && nk >= 1, PWK_E_INVALID, "bad eigensolver shape");
  PWK_REQUIRE(heev_work_doubles(n, nb, nk) <= kNoAllocEigDoubles, PWK_E_INVALID,
              "batch needs more than the 16 MiB static workspace (nk*(4n + 6n*nb) doubles above 2^21)");
  double* ws = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (c.device >= 0 && c.device < kMaxDevices) ws = g_scratch_ptr[c.device];
  }
  PWK_REQUIRE(ws != nullptr, PWK_E_STATE, "static eigensolver workspace not set up on this device");
  launch_heev_batched(c, n, nb, nk, H, ws, w, V, nullptr);
}

bool heev_supported(int n, int nb) { return n >= 1 && nb >= 1 && nb <= n && heev_lds_base(n, nb) <= kLdsCap; }

void launch_heev_batched(Context& c, int n, int nb, int nk, void* H, void* work, double* w, void* V, int* status) {
  PWK_REQUIRE(n >= 1 && nb >= 1 && nb <= n && nk >= 1, PWK_E_INVALID, "bad eigensolver shape");
  PWK_REQUIRE(heev_supported(n, nb), PWK_E_UNSUPPORTED, "matrix too large for the batched eigensolver");
  const size_t sh = heev_lds_bytes(n, nb);
  TimedScope ts(c, "dense_eigh");
  if (n <= JMAX) {
    k_jacobi<<<(unsigned)nk, 64, 0, c.stream>>>(n, nb, (const double2*)H, w, (double2*)V);
    PWK_HIP(hipGetLastError());
    if (status) PWK_HIP(hipMemsetAsync(status, 0, (size_t)nk * sizeof(int), c.stream));
    return;
  }
  const int al = a_fits_lds(n, nb) ? 1 : 0, lu = lu_fits_lds(n, nb) ? 1 : 0;
  if (n <= 64) {
    ensure_lds_limit((const void*)k_heev<64>, sh);
    k_heev<64><<<(unsigned)nk, 64, sh, c.stream>>>(n, nb, al, lu, (double2*)H, (double*)work, w, (double2*)V,
                                                   status);
  } else {
    ensure_lds_limit((const void*)k_heev<EIG_NT>, sh);
    k_heev<EIG_NT><<<(unsigned)nk, EIG_NT, sh, c.stream>>>(n, nb, al, lu, (double2*)H, (double*)work, w,
                                                           (double2*)V, status);
  }
  PWK_HIP(hipGetLastError());
}

}  // namespace pwk
