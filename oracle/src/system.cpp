// Setup (ScfContext), Hamiltonian construction and application, density.
// Test infrastructure (oracle).
#include <cstring>
#include <stdexcept>
#include <unordered_map>

#include "oracle.hpp"

namespace orc {

namespace {
// Memoises a radial transform per (species, |G| bit pattern). The transform is
// a pure function of |G|, so a cached value is bit-identical to a recomputed one.
struct RadialCache {
  std::vector<std::unordered_map<uint64_t, double>> maps;
  explicit RadialCache(size_t n_species) : maps(n_species) {}
  template <class F>
  double get(size_t sp, double g, F&& f) {
    uint64_t key;
    std::memcpy(&key, &g, sizeof key);
    auto it = maps[sp].find(key);
    if (it != maps[sp].end()) return it->second;
    double v = f();
    maps[sp].emplace(key, v);
    return v;
  }
};
}  // namespace

System make_system(const Crystal& crystal, const std::vector<Upf>& pps, double ecut_ev, uint32_t ratio,
                   const size_t* explicit_dims) {
  System s;
  s.crystal = crystal;
  s.pps = pps;
  s.basis = make_basis(crystal.lat, ecut_ev);
  s.dims = fft_dims(s.basis, ratio, explicit_dims);
  s.n_grid = s.dims[0] * s.dims[1] * s.dims[2];
  s.g_to_fft = basis_to_fft(s.basis, s.dims);
  s.omega = crystal.lat.volume();
  s.n_electrons = 0;
  for (auto& a : crystal.atoms) s.n_electrons += pps[(size_t)a.species].z_valence;
  Lattice recip = crystal.lat.reciprocal();
  s.g2.resize(s.n_grid);
  for (size_t i = 0; i < s.n_grid; ++i) {
    V3 g = g_vector_at(i, s.dims, recip);
    s.g2[i] = g.x * g.x + g.y * g.y + g.z * g.z;  // scf/context.rs:146-156
  }
  s.v_local = compute_v_local(s);
  s.rho_core = compute_core_density(s);
  return s;
}

// src/scf/potentials.rs:18-51. The radial transform depends on |G| only;
// values are cached per distinct |G| bit pattern (same result per point).
std::vector<cplx> compute_v_local(const System& s) {
  Lattice recip = s.crystal.lat.reciprocal();
  std::vector<cplx> out(s.n_grid);
  std::vector<V3> tau;
  for (auto& a : s.crystal.atoms) tau.push_back(s.crystal.cart(a));
  RadialCache cache(s.pps.size());
  for (size_t idx = 0; idx < s.n_grid; ++idx) {
    V3 g = g_vector_at(idx, s.dims, recip);
    double gn = norm(g);
    cplx acc(0.0, 0.0);
    for (size_t ia = 0; ia < s.crystal.atoms.size(); ++ia) {
      size_t sp = (size_t)s.crystal.atoms[ia].species;
      const Upf& pp = s.pps[sp];
      double phase = -dot(g, tau[ia]);
      double vg = cache.get(sp, gn, [&] { return v_local_of_g(pp, gn, s.omega); });
      acc += cplx(std::cos(phase), std::sin(phase)) * vg;
    }
    out[idx] = acc;
  }
  return out;
}

// src/scf/potentials.rs:72-130
std::vector<double> compute_core_density(const System& s) {
  bool any = false;
  for (auto& a : s.crystal.atoms)
    if (s.pps[(size_t)a.species].has_nlcc()) any = true;
  if (!any) return {};
  Lattice recip = s.crystal.lat.reciprocal();
  const double four_pi = 4.0 * PI;
  std::vector<cplx> rho_g(s.n_grid);
  RadialCache cache(s.pps.size());
  for (size_t idx = 0; idx < s.n_grid; ++idx) {
    V3 g = g_vector_at(idx, s.dims, recip);
    double gn = norm(g);
    cplx acc(0.0, 0.0);
    for (auto& a : s.crystal.atoms) {
      const Upf& pp = s.pps[(size_t)a.species];
      if (!pp.has_nlcc()) continue;
      double integral = cache.get((size_t)a.species, gn, [&] {
        size_t n = pp.r.size();
        std::vector<double> integ(n);
        for (size_t i = 0; i < n; ++i) {
          double r = pp.r[i];
          double gr = gn * r;
          double j0 = gr < 1e-10 ? 1.0 - gr * gr / 6.0 : std::sin(gr) / gr;
          integ[i] = pp.core[i] * r * r * j0;
        }
        return simpson(integ.data(), pp.rab.data(), n);
      });
      double phase = -dot(g, s.crystal.cart(a));
      acc += cplx(std::cos(phase), std::sin(phase)) * (four_pi * integral / s.omega);
    }
    rho_g[idx] = acc;
  }
  fft3d(rho_g, s.dims, +1);
  std::vector<double> out(s.n_grid);
  for (size_t i = 0; i < s.n_grid; ++i) out[i] = rho_g[i].real();
  return out;
}

// src/potential/nonlocal.rs:190-382
NonlocalK build_nonlocal(const System& s, V3 k) {
  NonlocalK nl;
  size_t npw = s.basis.size();
  nl.npw = npw;
  double inv_sqrt_omega = 1.0 / std::sqrt(s.omega);
  int lmax = 0;
  for (auto& pp : s.pps)
    for (auto& b : pp.betas) lmax = std::max(lmax, b.l);
  size_t ys = (size_t)((lmax + 1) * (lmax + 1));
  std::vector<double> ylm(npw * ys);
  std::vector<double> qn(npw);
  for (size_t ig = 0; ig < npw; ++ig) {
    V3 q = k + s.basis.g[ig];
    qn[ig] = norm(q);
    real_sph_harmonics(q, lmax, &ylm[ig * ys]);
  }
  // form factors per species (computed once per species)
  std::vector<std::vector<std::vector<double>>> ff(s.pps.size());
  std::vector<bool> have(s.pps.size(), false);
  for (auto& a : s.crystal.atoms) {
    size_t sp = (size_t)a.species;
    if (have[sp]) continue;
    have[sp] = true;
    const Upf& pp = s.pps[sp];
    ff[sp].resize(pp.n_proj());
    for (size_t ip = 0; ip < pp.n_proj(); ++ip) {
      ff[sp][ip].resize(npw);
      RadialCache cache(1);
      for (size_t ig = 0; ig < npw; ++ig)
        ff[sp][ip][ig] = cache.get(0, qn[ig], [&] { return beta_form_factor(pp, (int)ip, qn[ig]); });
    }
  }
  // channel layout: (atom, projector, m)
  std::vector<std::vector<size_t>> starts(s.crystal.atoms.size());
  size_t nch = 0;
  for (size_t ia = 0; ia < s.crystal.atoms.size(); ++ia) {
    const Upf& pp = s.pps[(size_t)s.crystal.atoms[ia].species];
    for (size_t ip = 0; ip < pp.n_proj(); ++ip) {
      starts[ia].push_back(nch);
      int l = pp.betas[ip].l;
      for (int m = 0; m < 2 * l + 1; ++m) {
        nl.ch_atom.push_back((int)ia);
        nl.ch_proj.push_back((int)ip);
        nl.ch_m.push_back(m);
        nl.ch_l.push_back(l);
      }
      nch += (size_t)(2 * l + 1);
    }
  }
  nl.nch = nch;
  nl.B.assign(npw * nch, cplx(0.0, 0.0));
  for (size_t ia = 0; ia < s.crystal.atoms.size(); ++ia) {
    const Atom& at = s.crystal.atoms[ia];
    const Upf& pp = s.pps[(size_t)at.species];
    V3 tau = s.crystal.cart(at);
    for (size_t ip = 0; ip < pp.n_proj(); ++ip) {
      size_t l = (size_t)pp.betas[ip].l;
      size_t base_lm = l * l;
      size_t ch0 = starts[ia][ip];
      for (size_t ig = 0; ig < npw; ++ig) {
        double ph = -dot(s.basis.g[ig], tau);
        cplx phase(std::cos(ph), std::sin(ph));
        double scale = inv_sqrt_omega * ff[(size_t)at.species][ip][ig];
        for (size_t mo = 0; mo < 2 * l + 1; ++mo) {
          double y = ylm[ig * ys + base_lm + mo];
          nl.B[(ch0 + mo) * npw + ig] = phase * cplx(scale * y, 0.0);
        }
      }
    }
  }
  // D·B^H, block-diagonal in (atom, l, m)
  nl.DBH.assign(nch * npw, cplx(0.0, 0.0));
  for (size_t ia = 0; ia < s.crystal.atoms.size(); ++ia) {
    const Upf& pp = s.pps[(size_t)s.crystal.atoms[ia].species];
    size_t np = pp.n_proj();
    for (size_t i = 0; i < np; ++i)
      for (size_t j = 0; j < np; ++j) {
        if (pp.betas[i].l != pp.betas[j].l) continue;
        double d = pp.dij[i * np + j];
        if (std::fabs(d) < 1e-20) continue;
        size_t l = (size_t)pp.betas[i].l;
        for (size_t mo = 0; mo < 2 * l + 1; ++mo) {
          size_t ci = starts[ia][i] + mo, cj = starts[ia][j] + mo;
          cplx dc(d, 0.0);
          for (size_t ig = 0; ig < npw; ++ig) nl.DBH[ig * nch + ci] += dc * std::conj(nl.B[cj * npw + ig]);
        }
      }
  }
  return nl;
}

// src/ewald.rs:43-143
double ewald_energy(const System& s) {
  const Crystal& cr = s.crystal;
  double omega = s.omega;
  std::vector<double> q;
  std::vector<V3> pos;
  for (auto& a : cr.atoms) {
    q.push_back(s.pps[(size_t)a.species].z_valence);
    pos.push_back(cr.cart(a));
  }
  size_t na = pos.size();
  double eta = std::cbrt((double)na * PI / omega);
  double eta2 = eta * eta;
  Lattice recip = cr.lat.reciprocal();
  double g_max = 10.0 * eta;
  int n1 = (int)std::ceil(g_max / norm(recip.a)), n2 = (int)std::ceil(g_max / norm(recip.b)),
      n3 = (int)std::ceil(g_max / norm(recip.c));
  double e_recip = 0.0;
  for (int i1 = -n1; i1 <= n1; ++i1)
    for (int i2 = -n2; i2 <= n2; ++i2)
      for (int i3 = -n3; i3 <= n3; ++i3) {
        if (i1 == 0 && i2 == 0 && i3 == 0) continue;
        V3 g = (double)i1 * recip.a + (double)i2 * recip.b + (double)i3 * recip.c;
        double g2 = norm2(g);
        if (g2 < 1e-12) continue;
        cplx sf(0.0, 0.0);
        for (size_t i = 0; i < na; ++i) {
          double ph = dot(g, pos[i]);
          sf += q[i] * cplx(std::cos(ph), std::sin(ph));
        }
        e_recip += std::norm(sf) * std::exp(-g2 / (4.0 * eta2)) / g2;
      }
  e_recip *= 2.0 * PI * E2_COULOMB / omega;
  double r_max = 10.0 / eta;
  int l1 = (int)std::ceil(r_max / norm(cr.lat.a)), l2 = (int)std::ceil(r_max / norm(cr.lat.b)),
      l3 = (int)std::ceil(r_max / norm(cr.lat.c));
  double e_real = 0.0;
  for (int a1 = -l1; a1 <= l1; ++a1)
    for (int a2 = -l2; a2 <= l2; ++a2)
      for (int a3 = -l3; a3 <= l3; ++a3) {
        V3 t = (double)a1 * cr.lat.a + (double)a2 * cr.lat.b + (double)a3 * cr.lat.c;
        for (size_t i = 0; i < na; ++i)
          for (size_t j = 0; j < na; ++j) {
            V3 r = pos[i] - pos[j] + t;
            double rn = norm(r);
            if (rn < 1e-10) continue;
            e_real += q[i] * q[j] * std::erfc(eta * rn) / rn;
          }
      }
  e_real *= 0.5 * E2_COULOMB;
  double z2 = 0.0, zs = 0.0;
  for (double z : q) {
    z2 += z * z;
    zs += z;
  }
  double e_self = -eta / std::sqrt(PI) * z2 * E2_COULOMB;
  double e_bg = -PI * zs * zs * E2_COULOMB / (2.0 * omega * eta2);
  return e_real + e_recip + e_self + e_bg;
}

// src/scf/initial_density.rs:64-166 (non-magnetic SAD, clamp + renormalise)
std::vector<double> sad_density(const System& s, double sigma) {
  Lattice recip = s.crystal.lat.reciprocal();
  std::vector<cplx> rho_g(s.n_grid);
  RadialCache cache(s.pps.size());
  for (size_t idx = 0; idx < s.n_grid; ++idx) {
    V3 g = g_vector_at(idx, s.dims, recip);
    double gn = norm(g);
    cplx acc(0.0, 0.0);
    for (auto& a : s.crystal.atoms) {
      const Upf& pp = s.pps[(size_t)a.species];
      double rho_at;
      if (pp.has_rho_atom()) {
        rho_at = cache.get((size_t)a.species, gn, [&] {
          size_t n = pp.r.size();
          std::vector<double> integ(n);
          for (size_t i = 0; i < n; ++i) {
            double gr = gn * pp.r[i];
            double j0 = gr < 1e-10 ? 1.0 - gr * gr / 6.0 : std::sin(gr) / gr;
            integ[i] = pp.rho_atom[i] * j0;
          }
          return simpson(integ.data(), pp.rab.data(), n);
        });
      } else {
        rho_at = pp.z_valence * std::exp(-0.5 * gn * gn * sigma * sigma);
      }
      double ph = -dot(g, s.crystal.cart(a));
      acc += cplx(std::cos(ph), std::sin(ph)) * (rho_at / s.omega);
    }
    rho_g[idx] = acc;
  }
  fft3d(rho_g, s.dims, +1);
  double dvol = s.omega / (double)s.n_grid;
  std::vector<double> rho(s.n_grid);
  double post = 0.0;
  for (size_t i = 0; i < s.n_grid; ++i) {
    rho[i] = rho_g[i].real() < 0.0 ? 0.0 : rho_g[i].real();
    post += rho[i];
  }
  post *= dvol;
  double scale = post > 1e-15 ? s.n_electrons / post : 1.0;
  for (auto& v : rho) v *= scale;
  return rho;
}

// src/hamiltonian.rs:23-34 / scf/potentials.rs:178
std::vector<double> kinetic_diag(const System& s, V3 k) {
  std::vector<double> t(s.basis.size());
  for (size_t i = 0; i < t.size(); ++i) t[i] = HBAR2_OVER_2M * norm2(k + s.basis.g[i]);
  return t;
}

// src/scf/potentials.rs:163-190
void fill_dense_h(const System& s, V3 k, const cplx* v_eff_g, cplx* h) {
  size_t n = s.basis.size();
  for (size_t i = 0; i < n; ++i) {
    double ke = HBAR2_OVER_2M * norm2(k + s.basis.g[i]);
    auto& mi = s.basis.miller[i];
    for (size_t j = 0; j < n; ++j) {
      auto& mj = s.basis.miller[j];
      cplx v = v_eff_g[miller_to_idx(s.dims, mi[0] - mj[0], mi[1] - mj[1], mi[2] - mj[2])];
      h[j * n + i] = (i == j) ? cplx(ke, 0.0) + v : v;
    }
  }
}

// src/potential/nonlocal.rs:432-450: H += B · DBH
void add_nonlocal_dense(const NonlocalK& nl, cplx* h) {
  size_t n = nl.npw, c = nl.nch;
  for (size_t j = 0; j < n; ++j)
    for (size_t a = 0; a < c; ++a) {
      cplx d = nl.DBH[j * c + a];
      if (d == cplx(0.0, 0.0)) continue;
      const cplx* b = &nl.B[a * n];
      cplx* hc = &h[j * n];
      for (size_t i = 0; i < n; ++i) hc[i] += b[i] * d;
    }
}

// FFT-route H·ψ. Equals (dense H)·ψ exactly in exact arithmetic because the
// wrapped Miller difference in fill_dense_h is a circular convolution.
void hpsi_fft(const System& s, V3 k, const NonlocalK* nl, const cplx* v_eff_g, const cplx* psi, size_t nb,
              cplx* hpsi) {
  size_t npw = s.basis.size(), N = s.n_grid;
  std::vector<cplx> vr(v_eff_g, v_eff_g + N);
  fft3d(vr, s.dims, +1);  // V(r) = IFFT_unnorm(V_eff(G))
  std::vector<double> t = kinetic_diag(s, k);
  double invN = 1.0 / (double)N;
  // bands are independent: one band per thread
#pragma omp parallel for schedule(dynamic)
  for (size_t b = 0; b < nb; ++b) {
    std::vector<cplx> work(N);
    const cplx* p = psi + b * npw;
    cplx* hp = hpsi + b * npw;
    for (size_t ig = 0; ig < npw; ++ig) work[s.g_to_fft[ig]] = p[ig];
    fft3d(work, s.dims, +1);
    for (size_t i = 0; i < N; ++i) work[i] *= vr[i];
    fft3d(work, s.dims, -1);
    for (size_t ig = 0; ig < npw; ++ig) hp[ig] = t[ig] * p[ig] + work[s.g_to_fft[ig]] * invN;
    if (nl) {
      std::vector<cplx> proj(nl->nch, cplx(0.0, 0.0));
      for (size_t a = 0; a < nl->nch; ++a) {
        cplx acc(0.0, 0.0);
        for (size_t ig = 0; ig < npw; ++ig) acc += nl->DBH[ig * nl->nch + a] * p[ig];
        proj[a] = acc;
      }
      for (size_t a = 0; a < nl->nch; ++a) {
        const cplx* bcol = &nl->B[a * npw];
        for (size_t ig = 0; ig < npw; ++ig) hp[ig] += bcol[ig] * proj[a];
      }
    }
  }
}

// src/scf/density.rs:59-97 (per-k body of the fold)
void accumulate_density(const System& s, const cplx* psi, size_t nb, const double* fw,
                        std::vector<double>& rho_acc) {
  size_t npw = s.basis.size(), N = s.n_grid;
  std::vector<cplx> work(N);
  for (size_t b = 0; b < nb; ++b) {
    double f = fw[b];
    if (f < 1e-15) continue;
    std::fill(work.begin(), work.end(), cplx(0.0, 0.0));
    for (size_t ig = 0; ig < npw; ++ig) work[s.g_to_fft[ig]] = psi[b * npw + ig];
    fft3d(work, s.dims, +1);
    for (size_t i = 0; i < N; ++i) rho_acc[i] += f * std::norm(work[i]);
  }
}

}  // namespace orc
