// Kleinman-Bylander projector matrices of a batch of k-points built on the device, each kernel
// over all of them at once (blockIdx.y = k-point), src/potential/nonlocal.rs:190-382:
//   F_i(q)   = 4π ∫ χ_i(r) j_l(q r) r dr   (Simpson, src/numerics.rs:29-58)
//   B[G, c]  = Ω^{-1/2} e^{-iG·τ_a} F_i(|k+G|) Y_lm(k+G)   (channel c = (atom a, projector i, m))
//   DBH[c,G] = Σ_{j: l_j = l_i} D_ij conj(B[G, (a, j, m)])
// One lane per plane wave; the radial mesh is read through L1/L2.
#include <algorithm>
#include <cstring>

#include "common.hpp"

namespace pwk {

namespace {

constexpr double kPi = 3.141592653589793;

// src/potential/nonlocal.rs:592-612
__device__ double sph_bessel(int l, double x) {
  if (fabs(x) < 1e-10) return l == 0 ? 1.0 : 0.0;
  double s = sin(x), c = cos(x);
  if (l == 0) return s / x;
  if (l == 1) return s / (x * x) - c / x;
  double jm1 = s / x, j = s / (x * x) - c / x;
  for (int n = 1; n < l; ++n) {
    double jp1 = fma((double)(2 * n + 1) / x, j, -jm1);
    jm1 = j;
    j = jp1;
  }
  return j;
}

// F[p][ig] = 4π ∫ χ_p(r) j_l(|k+G| r) r dr (Simpson with the even-mesh end correction,
// src/numerics.rs:29-58). One wavefront per (G, projector): lanes stride over the mesh,
// then a wave reduction, so every pair integrates in parallel.
__device__ inline double simpson_weight(int i, int n) {
  if (n < 3) return 3.0;  // plain sum, divided by 3 below
  if (i == 0) return 1.0;
  double w;
  if (i == n - 1) {
    w = (n % 2 == 1) ? 1.0 : 1.25;
  } else {
    w = (i % 2 == 1) ? 4.0 : 2.0;
    if (n % 2 == 0 && i == n - 3) w -= 0.25;
    if (n % 2 == 0 && i == n - 2) w += 1.0;
  }
  return w;
}

__global__ void __launch_bounds__(256) k_form_factors(int npw, int nprojtot, const double* __restrict__ qn,
                                                      const int* __restrict__ p_l, const int* __restrict__ p_mesh,
                                                      const long* __restrict__ p_rofs, const long* __restrict__ p_cofs,
                                                      const double* __restrict__ r, const double* __restrict__ rab,
                                                      const double* __restrict__ chi, double* __restrict__ F) {
  const int lane = threadIdx.x & 63;
  const long pair = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (pair >= (long)npw * nprojtot) return;  // wave-uniform
  qn += (size_t)blockIdx.y * npw;
  F += (size_t)blockIdx.y * nprojtot * npw;
  const int ig = (int)(pair % npw), p = (int)(pair / npw);
  const double q = qn[ig];
  const int l = p_l[p], n = p_mesh[p];
  const double* rr = r + p_rofs[p];
  const double* ra = rab + p_rofs[p];
  const double* ch = chi + p_cofs[p];
  double acc = 0.0;
  for (int i = lane; i < n; i += 64) acc += simpson_weight(i, n) * (ch[i] * sph_bessel(l, q * rr[i]) * rr[i]) * ra[i];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane == 0) F[(size_t)p * npw + ig] = 4.0 * kPi * (acc / 3.0);
}

// Real Y_lm for l <= 3 in the order (0,0),(1,0),(1,+1),(1,-1),(2,0),(2,+1),(2,-1),... (nonlocal.rs:469-552)
__device__ void ylm_dev(double qx, double qy, double qz, int lmax, double* out) {
  const double fpi = 4.0 * kPi;
  const double inv = sqrt(1.0 / fpi);
  int total = (lmax + 1) * (lmax + 1);
  for (int i = 0; i < total; ++i) out[i] = 0.0;
  out[0] = inv;
  if (lmax == 0) return;
  double g = sqrt(qx * qx + qy * qy + qz * qz);
  if (g < 1e-9) return;
  double cost = qz / g;
  double sint = sqrt(fmax(1.0 - cost * cost, 0.0));
  double phi = atan2(qy, qx);
  double Q[4][4] = {{0}};
  Q[0][0] = 1.0;
  Q[1][0] = cost;
  Q[1][1] = -sint / sqrt(2.0);
  double c1 = sqrt(3.0 / fpi);
  out[1] = c1 * Q[1][0];
  out[2] = c1 * sqrt(2.0) * Q[1][1] * cos(phi);
  out[3] = c1 * sqrt(2.0) * Q[1][1] * sin(phi);
  int lm = 4;
  for (int l = 2; l <= lmax; ++l) {
    double c = sqrt((double)(2 * l + 1) / fpi);
    double lf = (double)l;
    for (int m = 0; m + 2 <= l; ++m) {
      double mf = (double)m;
      double llmm = sqrt(lf * lf - mf * mf);
      double llm1 = sqrt((lf - 1.0) * (lf - 1.0) - mf * mf);
      Q[l][m] = (cost * (2.0 * lf - 1.0) * Q[l - 1][m] - llm1 * Q[l - 2][m]) / llmm;
    }
    Q[l][l - 1] = cost * sqrt(2.0 * lf - 1.0) * Q[l - 1][l - 1];
    Q[l][l] = -sqrt((2.0 * lf - 1.0) / (2.0 * lf)) * sint * Q[l - 1][l - 1];
    out[lm] = c * Q[l][0];
    for (int m = 1; m <= l; ++m) {
      double mf = (double)m;
      out[lm + 2 * m - 1] = c * sqrt(2.0) * Q[l][m] * cos(mf * phi);
      out[lm + 2 * m] = c * sqrt(2.0) * Q[l][m] * sin(mf * phi);
    }
    lm += 2 * l + 1;
  }
}

// B (npw × nch, column-major). Channel c: atom ch_atom[c], global projector
// ch_proj[c] (index into F), l = ch_l[c], m-offset ch_m[c].
__global__ void __launch_bounds__(256) k_build_b(int npw, int nch, int lmax, const double* __restrict__ kvec,
                                                 const double* __restrict__ gvec, const double* __restrict__ tau,
                                                 const int* __restrict__ ch_atom, const int* __restrict__ ch_proj,
                                                 const int* __restrict__ ch_l, const int* __restrict__ ch_m,
                                                 const double* __restrict__ F, int nprojtot, double inv_sqrt_omega,
                                                 double2* __restrict__ B) {
  int ig = blockIdx.x * blockDim.x + threadIdx.x;
  if (ig >= npw) return;
  kvec += 3 * blockIdx.y;
  F += (size_t)blockIdx.y * nprojtot * npw;
  B += (size_t)blockIdx.y * npw * nch;
  double gx = gvec[3 * ig], gy = gvec[3 * ig + 1], gz = gvec[3 * ig + 2];
  double y[16];
  ylm_dev(kvec[0] + gx, kvec[1] + gy, kvec[2] + gz, lmax, y);
  int last_atom = -1;
  double pc = 1.0, ps = 0.0;
  for (int c = 0; c < nch; ++c) {
    int a = ch_atom[c];
    if (a != last_atom) {
      double ph = -(gx * tau[3 * a] + gy * tau[3 * a + 1] + gz * tau[3 * a + 2]);
      pc = cos(ph);
      ps = sin(ph);
      last_atom = a;
    }
    int l = ch_l[c];
    double scale = inv_sqrt_omega * F[(size_t)ch_proj[c] * npw + ig];
    double yv = y[l * l + ch_m[c]];
    double2 o;
    o.x = pc * (scale * yv) - ps * 0.0;
    o.y = pc * 0.0 + ps * (scale * yv);
    B[(size_t)c * npw + ig] = o;
  }
}

// DBH[c_i, G] = Σ d·conj(B[G, c_j]) over (c_i, c_j, d) triplets in the reference's (i, j, m) order
__global__ void __launch_bounds__(256) k_build_dbh(int npw, int nch, int ntrip, const int* __restrict__ tri_ci,
                                                   const int* __restrict__ tri_cj, const double* __restrict__ tri_d,
                                                   const double2* __restrict__ B, double2* __restrict__ DBH) {
  int ig = blockIdx.x * blockDim.x + threadIdx.x;
  if (ig >= npw) return;
  B += (size_t)blockIdx.y * npw * nch;
  DBH += (size_t)blockIdx.y * npw * nch;
  double2* row = DBH + (size_t)ig * nch;
  for (int c = 0; c < nch; ++c) row[c] = make_double2(0.0, 0.0);
  for (int t = 0; t < ntrip; ++t) {
    double2 b = B[(size_t)tri_cj[t] * npw + ig];
    double d = tri_d[t];
    double2 acc = row[tri_ci[t]];
    // (d + 0i)·conj(b)
    acc.x += d * b.x - 0.0 * (-b.y);
    acc.y += d * (-b.y) + 0.0 * b.x;
    row[tri_ci[t]] = acc;
  }
}

__global__ void __launch_bounds__(256) k_qnorm(int npw, const double* __restrict__ kvec,
                                               const double* __restrict__ gvec, double* __restrict__ qn) {
  int ig = blockIdx.x * blockDim.x + threadIdx.x;
  if (ig >= npw) return;
  kvec += 3 * blockIdx.y;
  qn += (size_t)blockIdx.y * npw;
  double x = kvec[0] + gvec[3 * ig], y = kvec[1] + gvec[3 * ig + 1], z = kvec[2] + gvec[3 * ig + 2];
  qn[ig] = sqrt(x * x + y * y + z * z);
}

// Host tables of one build packed into a single upload (each array at a 256-byte aligned offset,
// at least one element), so a build costs one copy instead of one per table.
struct Pack {
  std::vector<unsigned char> blob;
  template <class T>
  size_t add(const T* h, size_t n) {
    const size_t off = (blob.size() + 255) & ~(size_t)255;
    blob.resize(off + std::max<size_t>(1, n) * sizeof(T), 0);
    if (n) std::memcpy(blob.data() + off, h, n * sizeof(T));
    return off;
  }
  template <class T>
  size_t add(const std::vector<T>& v) {
    return add(v.data(), v.size());
  }
};

// A typed view of one array inside the uploaded pack.
struct PackView {
  void* p;
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

}  // namespace

void build_nonlocal_device(Context& c, int ik, const double* kvec, const double* gvec, int natoms, const double* tau,
                           const int32_t* atom_species, int nspecies, const int32_t* nproj, const int32_t* proj_l,
                           const double* dij, const int32_t* mesh, const double* r, const double* rab,
                           const double* chi, double omega, const NonlocalDest* dest) {
  build_nonlocal_batch(c, ik, 1, kvec, gvec, natoms, tau, atom_species, nspecies, nproj, proj_l, dij, mesh, r, rab,
                       chi, omega, dest);
}

void build_nonlocal_batch(Context& c, int ik0, int nkb, const double* kvecs, const double* gvec, int natoms,
                          const double* tau, const int32_t* atom_species, int nspecies, const int32_t* nproj,
                          const int32_t* proj_l, const double* dij, const int32_t* mesh, const double* r,
                          const double* rab, const double* chi, double omega, const NonlocalDest* dest) {
  c.check_basis();
  PWK_REQUIRE(nkb >= 1 && (nkb == 1 || dest), PWK_E_INVALID, "a batch of k-points needs a destination");
  for (int i = 0; i < nkb; ++i) c.k(ik0 + i);  // range check
  const int npw = c.npw;
  PWK_REQUIRE(natoms >= 0 && nspecies > 0, PWK_E_INVALID, "bad atom/species counts");
  // per-species offsets into the concatenated arrays
  std::vector<int> p0(nspecies + 1, 0), d0(nspecies + 1, 0);
  std::vector<long> m0(nspecies + 1, 0), c0(nspecies + 1, 0);
  for (int s = 0; s < nspecies; ++s) {
    PWK_REQUIRE(nproj[s] >= 0 && mesh[s] > 0, PWK_E_INVALID, "bad projector table");
    p0[s + 1] = p0[s] + nproj[s];
    d0[s + 1] = d0[s] + nproj[s] * nproj[s];
    m0[s + 1] = m0[s] + mesh[s];
    c0[s + 1] = c0[s] + (long)nproj[s] * mesh[s];
  }
  int nprojtot = p0[nspecies];
  int lmax = 0;
  std::vector<int> pl((size_t)nprojtot), pmesh((size_t)nprojtot);
  std::vector<long> profs((size_t)nprojtot), pcofs((size_t)nprojtot);
  for (int s = 0; s < nspecies; ++s)
    for (int i = 0; i < nproj[s]; ++i) {
      int p = p0[s] + i;
      pl[(size_t)p] = proj_l[p];
      PWK_REQUIRE(proj_l[p] >= 0 && proj_l[p] <= 3, PWK_E_UNSUPPORTED, "projector l must be in 0..3");
      lmax = std::max(lmax, proj_l[p]);
      pmesh[(size_t)p] = mesh[s];
      profs[(size_t)p] = m0[s];
      pcofs[(size_t)p] = c0[s] + (long)i * mesh[s];
    }
  // channel layout (atom, projector, m) and D triplets, as nonlocal.rs:295-379
  std::vector<int> ch_atom, ch_proj, ch_l, ch_m, tri_ci, tri_cj;
  std::vector<double> tri_d;
  for (int a = 0; a < natoms; ++a) {
    int s = atom_species[a];
    PWK_REQUIRE(s >= 0 && s < nspecies, PWK_E_INVALID, "atom species out of range");
    std::vector<int> start((size_t)nproj[s]);
    for (int i = 0; i < nproj[s]; ++i) {
      start[(size_t)i] = (int)ch_atom.size();
      int l = proj_l[p0[s] + i];
      for (int m = 0; m < 2 * l + 1; ++m) {
        ch_atom.push_back(a);
        ch_proj.push_back(p0[s] + i);
        ch_l.push_back(l);
        ch_m.push_back(m);
      }
    }
    for (int i = 0; i < nproj[s]; ++i)
      for (int j = 0; j < nproj[s]; ++j) {
        int li = proj_l[p0[s] + i], lj = proj_l[p0[s] + j];
        if (li != lj) continue;
        double d = dij[d0[s] + i * nproj[s] + j];
        if (std::fabs(d) < 1e-20) continue;
        for (int m = 0; m < 2 * li + 1; ++m) {
          tri_ci.push_back(start[(size_t)i] + m);
          tri_cj.push_back(start[(size_t)j] + m);
          tri_d.push_back(d);
        }
      }
  }
  int nch = (int)ch_atom.size();
  c.pack_valid = false;
  for (int i = 0; i < nkb; ++i) c.k(ik0 + i).nch = nch;
  if (nch == 0) return;
  hipStream_t st = c.stream;
  Pack pk;
  const size_t o_k = pk.add(kvecs, (size_t)3 * nkb), o_g = pk.add(gvec, (size_t)npw * 3), o_tau = pk.add(tau, (size_t)natoms * 3);
  const size_t o_pl = pk.add(pl), o_pm = pk.add(pmesh), o_pro = pk.add(profs), o_pco = pk.add(pcofs);
  const size_t o_r = pk.add(r, (size_t)m0[nspecies]), o_rab = pk.add(rab, (size_t)m0[nspecies]);
  const size_t o_chi = pk.add(chi, (size_t)c0[nspecies]);
  const size_t o_ca = pk.add(ch_atom), o_cp = pk.add(ch_proj), o_cl = pk.add(ch_l), o_cm = pk.add(ch_m);
  const size_t o_ti = pk.add(tri_ci), o_tj = pk.add(tri_cj), o_td = pk.add(tri_d);
  DevBuf d_pack, d_qn, d_F;
  d_pack.ensure(pk.blob.size());
  PWK_HIP(hipMemcpyAsync(d_pack.p, pk.blob.data(), pk.blob.size(), hipMemcpyHostToDevice, st));
  auto at = [&](size_t off) { return PackView{static_cast<unsigned char*>(d_pack.p) + off}; };
  const PackView d_k = at(o_k), d_g = at(o_g), d_tau = at(o_tau), d_pl = at(o_pl), d_pm = at(o_pm), d_pro = at(o_pro),
             d_pco = at(o_pco), d_r = at(o_r), d_rab = at(o_rab), d_chi = at(o_chi), d_ca = at(o_ca),
             d_cp = at(o_cp), d_cl = at(o_cl), d_cm = at(o_cm), d_ti = at(o_ti), d_tj = at(o_tj), d_td = at(o_td);
  d_qn.ensure((size_t)nkb * npw * sizeof(double));
  d_F.ensure((size_t)nkb * std::max(1, nprojtot) * npw * sizeof(double));
  const dim3 gpw(blocks_for((size_t)npw, 256), (unsigned)nkb);
  k_qnorm<<<gpw, 256, 0, st>>>(npw, d_k.as<double>(), d_g.as<double>(), d_qn.as<double>());
  PWK_HIP(hipGetLastError());
  if (nprojtot > 0) {
    const dim3 g(blocks_for((size_t)npw * nprojtot, 4), (unsigned)nkb);
    k_form_factors<<<g, 256, 0, st>>>(npw, nprojtot, d_qn.as<double>(), d_pl.as<int>(), d_pm.as<int>(),
                                      d_pro.as<long>(), d_pco.as<long>(), d_r.as<double>(), d_rab.as<double>(),
                                      d_chi.as<double>(), d_F.as<double>());
    PWK_HIP(hipGetLastError());
  }
  void *b64, *dbh64, *b32, *dbh32;
  if (dest) {
    PWK_REQUIRE(dest->nch == nch, PWK_E_STATE, "projector destination sized for another channel count");
    b64 = dest->b64;
    dbh64 = dest->dbh64;
    b32 = dest->b32;
    dbh32 = dest->dbh32;
  } else {
    KData& kd = c.k(ik0);
    kd.B64.ensure((size_t)npw * nch * sizeof(double2));
    kd.DBH64.ensure((size_t)npw * nch * sizeof(double2));
    kd.B32.ensure((size_t)npw * nch * sizeof(float2));
    kd.DBH32.ensure((size_t)npw * nch * sizeof(float2));
    b64 = kd.B64.p;
    dbh64 = kd.DBH64.p;
    b32 = kd.B32.p;
    dbh32 = kd.DBH32.p;
  }
  // the destination holds the k-points' blocks back to back (npw × nch each)
  k_build_b<<<gpw, 256, 0, st>>>(npw, nch, lmax, d_k.as<double>(), d_g.as<double>(), d_tau.as<double>(),
                                 d_ca.as<int>(), d_cp.as<int>(), d_cl.as<int>(), d_cm.as<int>(), d_F.as<double>(),
                                 nprojtot, 1.0 / std::sqrt(omega), static_cast<double2*>(b64));
  PWK_HIP(hipGetLastError());
  k_build_dbh<<<gpw, 256, 0, st>>>(npw, nch, (int)tri_d.size(), d_ti.as<int>(), d_tj.as<int>(), d_td.as<double>(),
                                   static_cast<const double2*>(b64), static_cast<double2*>(dbh64));
  PWK_HIP(hipGetLastError());
  launch_convert(c, b64, b32, (size_t)nkb * npw * nch * 2, true);
  launch_convert(c, dbh64, dbh32, (size_t)nkb * npw * nch * 2, true);
  // temporaries are freed at scope exit; make sure the kernels have consumed them
  PWK_HIP(hipStreamSynchronize(st));
}

}  // namespace pwk
