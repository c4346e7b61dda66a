// nspin = 1 SCF driver, restating src/scf/driver.rs:266-742 with the dense
// eigensolver path (fill_hamiltonian_with_v_eff + add_to_hamiltonian +
// diagonalize_lowest), Fermi-Dirac occupations, G-space symmetrisation and
// plain Anderson mixing. Test infrastructure (oracle).
#include <cstdio>

#include "oracle.hpp"

namespace orc {

namespace {
struct XcOut {
  std::vector<double> exc, vxc;
};

XcOut eval_lda(const std::vector<double>& rho) {
  XcOut o;
  o.exc.resize(rho.size());
  o.vxc.resize(rho.size());
  for (size_t i = 0; i < rho.size(); ++i) lda_xc(rho[i], &o.exc[i], &o.vxc[i]);
  return o;
}

// src/scf/energy.rs:314-324
std::vector<double> add_core(const std::vector<double>& val, const std::vector<double>& core) {
  if (core.empty()) return val;
  std::vector<double> out(val.size());
  for (size_t i = 0; i < val.size(); ++i) out[i] = std::max(val[i] + core[i], 0.0);
  return out;
}

// dense H + lowest nb eigenpairs at every k for one V_eff (driver.rs:461-472, driver_spin.rs:354-389)
void diag_all(const System& s, const std::vector<KPoint>& kpts, const std::vector<NonlocalK>& vnl,
              const std::vector<cplx>& veff, size_t nb, std::vector<std::vector<double>>& evals,
              std::vector<std::vector<cplx>>& wfn) {
  const size_t npw = s.basis.size(), nk = kpts.size();
  evals.assign(nk, {});
  wfn.assign(nk, {});
#pragma omp parallel
  {
    std::vector<cplx> h(npw * npw), vecs(npw * nb);
    std::vector<double> wv(npw);
#pragma omp for schedule(dynamic)
    for (size_t k = 0; k < nk; ++k) {
      fill_dense_h(s, kpts[k].k, veff.data(), h.data());
      add_nonlocal_dense(vnl[k], h.data());
      eigh_lowest(npw, h.data(), nb, wv.data(), vecs.data());
      evals[k].assign(wv.begin(), wv.begin() + (long)nb);
      wfn[k].assign(vecs.begin(), vecs.begin() + (long)(npw * nb));
    }
  }
}

// Σ_k Σ_b f·w_k·|ψ(r)|², summed over k in a fixed order, rescaled to n_target
// (src/scf/density.rs:41-111)
std::vector<double> density_of(const System& s, const std::vector<std::vector<cplx>>& wfn,
                               const std::vector<std::vector<double>>& occ, const std::vector<double>& wk, size_t nb,
                               double n_target) {
  const size_t N = s.n_grid, nk = wfn.size();
  std::vector<std::vector<double>> per_k(nk);
#pragma omp parallel for schedule(dynamic)
  for (size_t k = 0; k < nk; ++k) {
    std::vector<double> fw(nb);
    for (size_t b = 0; b < nb; ++b) fw[b] = occ[k][b] * wk[k];
    per_k[k].assign(N, 0.0);
    accumulate_density(s, wfn[k].data(), nb, fw.data(), per_k[k]);
  }
  std::vector<double> rho(N, 0.0);
  for (size_t k = 0; k < nk; ++k)
    for (size_t i = 0; i < N; ++i) rho[i] += per_k[k][i];
  double integral = 0.0;
  for (double v : rho) integral += v;
  integral *= s.omega / (double)N;
  if (std::fabs(integral) > 1e-15) {
    const double sc = n_target / integral;
    for (auto& v : rho) v *= sc;
  }
  return rho;
}

// Σ_k Σ_b f·w_k ⟨ψ|T|ψ⟩ and ⟨ψ|V_NL|ψ⟩ (energy.rs:382-417, 467-510)
void kin_nl_expect(const System& s, const std::vector<KPoint>& kpts, const std::vector<NonlocalK>& vnl,
                   const std::vector<std::vector<cplx>>& wfn, const std::vector<std::vector<double>>& occ, size_t nb,
                   double& e_kin, double& e_nl) {
  const size_t npw = s.basis.size();
  for (size_t k = 0; k < kpts.size(); ++k) {
    std::vector<double> t = kinetic_diag(s, kpts[k].k);
    const double wk = kpts[k].weight;
    double ek = 0.0, enl = 0.0;
    for (size_t b = 0; b < nb; ++b) {
      const double f = occ[k][b];
      if (f == 0.0) continue;
      const cplx* psi = &wfn[k][b * npw];
      double tb = 0.0;
      for (size_t ig = 0; ig < npw; ++ig) tb += std::norm(psi[ig]) * t[ig];
      ek += f * wk * tb;
      cplx en(0.0, 0.0);
      for (size_t a = 0; a < vnl[k].nch; ++a) {
        cplx pr(0.0, 0.0), bp(0.0, 0.0);
        for (size_t ig = 0; ig < npw; ++ig) pr += vnl[k].DBH[ig * vnl[k].nch + a] * psi[ig];
        for (size_t ig = 0; ig < npw; ++ig) bp += std::conj(psi[ig]) * vnl[k].B[a * npw + ig];
        en += bp * pr;
      }
      enl += f * wk * en.real();
    }
    e_kin += ek;
    e_nl += enl;
  }
}

// per-grid-point Kerker weights |G|²/(|G|² + q_TF²), zero at G = 0 (mixing/anderson.rs:34-97)
std::vector<double> kerker_weights(const System& s, const ScfParams& prm) {
  if (!prm.kerker) return {};
  const double q2 = prm.q_tf > 0.0 ? prm.q_tf * prm.q_tf : kerker_auto_qtf2(s.n_electrons, s.omega);
  std::vector<double> w(s.n_grid);
  for (size_t i = 0; i < s.n_grid; ++i) w[i] = s.g2[i] < 1e-12 ? 0.0 : s.g2[i] / (s.g2[i] + q2);
  return w;
}
}  // namespace

ScfResult run_scf(const System& s, const std::vector<KPoint>& kpts, const ScfParams& prm, const Symmetry& sym) {
  size_t N = s.n_grid, npw = s.basis.size(), nk = kpts.size();
  size_t nb = std::min(prm.n_bands, npw);
  const double spin_factor = 2.0;
  std::vector<double> wk(nk);
  for (size_t k = 0; k < nk; ++k) wk[k] = kpts[k].weight;
  std::vector<NonlocalK> vnl;
  vnl.resize(nk);
#pragma omp parallel for schedule(dynamic)
  for (size_t k = 0; k < nk; ++k) vnl[k] = build_nonlocal(s, kpts[k].k);
  double e_ewald = ewald_energy(s);

  std::vector<double> rho_r = sad_density(s, prm.gaussian_sigma);
  std::vector<cplx> rho_g;
  density_r_to_g(rho_r, s.dims, rho_g);
  Anderson mixer;
  mixer.beta = prm.mixing_beta;
  mixer.max_history = prm.mixing_ndim;
  mixer.kerker_w = kerker_weights(s, prm);
  mixer.dims = s.dims;

  ScfResult res;
  bool have_prev = false;
  double e_prev = 0.0;
  for (size_t iter = 0; iter < prm.max_iter; ++iter) {
    std::vector<cplx> vh(N), vxc_g, veff(N);
    hartree(rho_g.data(), s.g2.data(), N, vh.data());
    std::vector<double> rho_xc = add_core(rho_r, s.rho_core);
    XcOut xin = eval_lda(rho_xc);
    density_r_to_g(xin.vxc, s.dims, vxc_g);
    assemble_veff(s.v_local.data(), vh.data(), vxc_g.data(), N, veff.data());

    std::vector<std::vector<double>> evals(nk);
    std::vector<std::vector<cplx>> wfn(nk);
    // one k-point per thread, as the reference's rayon par_iter_mut over k (driver.rs:461-472)
#pragma omp parallel
    {
      std::vector<cplx> h(npw * npw), vecs(npw * nb);
      std::vector<double> wv(npw);
#pragma omp for schedule(dynamic)
      for (size_t k = 0; k < nk; ++k) {
        fill_dense_h(s, kpts[k].k, veff.data(), h.data());
        add_nonlocal_dense(vnl[k], h.data());
        eigh_lowest(npw, h.data(), nb, wv.data(), vecs.data());
        evals[k].assign(wv.begin(), wv.begin() + (long)nb);
        wfn[k].assign(vecs.begin(), vecs.begin() + (long)(npw * nb));
      }
    }
    double ef = find_fermi_energy(evals, wk, s.n_electrons, prm.sigma, spin_factor);
    std::vector<std::vector<double>> occ(nk, std::vector<double>(nb));
    for (size_t k = 0; k < nk; ++k)
      for (size_t b = 0; b < nb; ++b) occ[k][b] = fermi_dirac_occ(evals[k][b], ef, prm.sigma) * spin_factor;

    // density (src/scf/density.rs:41-111)
    // per-k accumulation in parallel, then a fixed-order sum over k so the result is
    // independent of the thread count (the rayon fold/reduce of density.rs:59-97)
    std::vector<double> rho_new(N, 0.0);
    std::vector<std::vector<double>> per_k(nk);
#pragma omp parallel for schedule(dynamic)
    for (size_t k = 0; k < nk; ++k) {
      std::vector<double> fw(nb);
      for (size_t b = 0; b < nb; ++b) fw[b] = occ[k][b] * wk[k];
      per_k[k].assign(N, 0.0);
      accumulate_density(s, wfn[k].data(), nb, fw.data(), per_k[k]);
    }
    for (size_t k = 0; k < nk; ++k)
      for (size_t i = 0; i < N; ++i) rho_new[i] += per_k[k][i];
    double dvol = s.omega / (double)N;
    double integral = 0.0;
    for (double v : rho_new) integral += v;
    integral *= dvol;
    if (std::fabs(integral) > 1e-15) {
      double sc = s.n_electrons / integral;
      for (auto& v : rho_new) v *= sc;
    }
    symmetrize_density(rho_new, s.dims, sym);

    double delta = density_diff(rho_r, rho_new, s.omega);
    std::vector<cplx> rho_g_new;
    density_r_to_g(rho_new, s.dims, rho_g_new);
    std::vector<double> rho_new_xc = add_core(rho_new, s.rho_core);
    XcOut xout = eval_lda(rho_new_xc);
    double e_band = 0.0;
    for (size_t k = 0; k < nk; ++k) {
      double sk = 0.0;
      for (size_t b = 0; b < nb; ++b) sk += occ[k][b] * wk[k] * evals[k][b];
      e_band += sk;
    }
    double ts = entropy_ts(evals, wk, ef, prm.sigma, spin_factor);
    double e_h_out = hartree_energy(rho_g_new, s.g2, s.omega);
    double e_xc_out = xc_energy_corrected(rho_new_xc, rho_new, xout.exc, xout.vxc, s.omega);
    double e_total = e_band - e_h_out + e_xc_out + e_ewald + (-ts);
    double e_harris = e_band - hartree_energy(rho_g, s.g2, s.omega) +
                      xc_energy_corrected(rho_xc, rho_r, xin.exc, xin.vxc, s.omega) + e_ewald + (-ts);
    double de = have_prev ? std::fabs(e_total - e_prev) : INFINITY;
    bool energy_conv = have_prev && de < prm.energy_threshold;
    have_prev = true;
    e_prev = e_total;
    res.final_delta = delta;
    res.n_iter = iter + 1;
    if (prm.verbose)
      std::fprintf(stderr, "iter %3zu  E=%.10f  dE=%.3e  delta=%.3e  EF=%.6f\n", iter, e_total, de, delta, ef);
    if (delta < prm.conv_threshold && energy_conv) {
      res.converged = true;
      res.e_total = e_total;
      res.e_harris = e_harris;
      res.fermi = ef;
      res.ts = ts;
      res.eigenvalues = evals;
      res.rho_r = rho_new;
      EnergyComponents& c = res.comp;
      c.e_band = e_band;
      c.e_ewald = e_ewald;
      c.e_smearing = -ts;
      c.e_hartree = e_h_out;
      double exc_int = 0.0, evxc = 0.0;
      for (size_t i = 0; i < N; ++i) exc_int += rho_new_xc[i] * xout.exc[i] * dvol;
      for (size_t i = 0; i < N; ++i) evxc += rho_new[i] * xout.vxc[i] * dvol;
      c.e_xc = exc_int;
      c.e_vxc = evxc;
      // kinetic (energy.rs:382-417), local (energy.rs:439-446), nonlocal (energy.rs:467-510)
      for (size_t k = 0; k < nk; ++k) {
        std::vector<double> t = kinetic_diag(s, kpts[k].k);
        double ek = 0.0, enl = 0.0;
        for (size_t b = 0; b < nb; ++b) {
          double f = occ[k][b];
          if (f == 0.0) continue;
          const cplx* psi = &wfn[k][b * npw];
          double tb = 0.0;
          for (size_t ig = 0; ig < npw; ++ig) tb += std::norm(psi[ig]) * t[ig];
          ek += f * wk[k] * tb;
          std::vector<cplx> proj(vnl[k].nch);
          for (size_t a = 0; a < vnl[k].nch; ++a) {
            cplx acc(0.0, 0.0);
            for (size_t ig = 0; ig < npw; ++ig) acc += vnl[k].DBH[ig * vnl[k].nch + a] * psi[ig];
            proj[a] = acc;
          }
          cplx en(0.0, 0.0);
          for (size_t a = 0; a < vnl[k].nch; ++a) {
            cplx bp(0.0, 0.0);
            for (size_t ig = 0; ig < npw; ++ig) bp += std::conj(psi[ig]) * vnl[k].B[a * npw + ig];
            en += bp * proj[a];
          }
          enl += f * wk[k] * en.real();
        }
        c.e_kinetic += ek;
        c.e_nonlocal += enl;
      }
      std::vector<cplx> vl(s.v_local);
      fft3d(vl, s.dims, +1);
      double eloc = 0.0;
      for (size_t i = 0; i < N; ++i) eloc += rho_new[i] * vl[i].real() * dvol;
      c.e_local = eloc;
      return res;
    }
    rho_r = mixer.mix(rho_r, rho_new);
    density_r_to_g(rho_r, s.dims, rho_g);
  }
  res.converged = false;
  return res;
}


// LSDA SCF (src/scf/driver_spin.rs:37-839): spin_factor 1, one V_eff per channel from the
// common Hartree potential and lda_xc_spin on ρ_σ + ρ_core/2, occupations from one Fermi
// level over both channels (or one per channel at a fixed total moment), per-channel
// density, convergence on max(Δρ↑, Δρ↓), and Anderson mixing of ρ_total and m = ρ↑ − ρ↓
// with separate histories.
ScfResult run_scf_spin(const System& s, const std::vector<KPoint>& kpts, const ScfParams& prm, const Symmetry& sym) {
  const size_t N = s.n_grid, nk = kpts.size();
  const size_t nb = std::min(prm.n_bands, s.basis.size());
  const double spin_factor = 1.0;
  const double dvol = s.omega / (double)N;
  std::vector<double> wk(nk);
  for (size_t k = 0; k < nk; ++k) wk[k] = kpts[k].weight;
  std::vector<NonlocalK> vnl(nk);
#pragma omp parallel for schedule(dynamic)
  for (size_t k = 0; k < nk; ++k) vnl[k] = build_nonlocal(s, kpts[k].k);
  const double e_ewald = ewald_energy(s);

  // start: ρ_SAD split by the mean starting magnetisation (driver_spin.rs:67-75)
  const std::vector<double> rho0 = sad_density(s, prm.gaussian_sigma);
  std::vector<double> up(N), dn(N);
  for (size_t i = 0; i < N; ++i) {
    up[i] = rho0[i] * (1.0 + prm.start_mag) / 2.0;
    dn[i] = rho0[i] * (1.0 - prm.start_mag) / 2.0;
  }
  std::vector<double> core_half(s.rho_core.size());
  for (size_t i = 0; i < core_half.size(); ++i) core_half[i] = s.rho_core[i] / 2.0;
  Anderson mix_tot, mix_mag;
  mix_tot.beta = mix_mag.beta = prm.mixing_beta;
  mix_tot.max_history = mix_mag.max_history = prm.mixing_ndim;
  mix_tot.kerker_w = kerker_weights(s, prm);  // the magnetisation channel mixes plain (driver_spin.rs:76-98)
  mix_tot.dims = s.dims;

  auto lsda_grid = [&](const std::vector<double>& a, const std::vector<double>& b, std::vector<double>& e,
                       std::vector<double>& va, std::vector<double>& vb) {
    e.resize(N);
    va.resize(N);
    vb.resize(N);
    for (size_t i = 0; i < N; ++i) lsda_xc(a[i], b[i], &e[i], &va[i], &vb[i]);
  };

  ScfResult res;
  bool have_prev = false;
  double e_prev = 0.0;
  for (size_t iter = 0; iter < prm.max_iter; ++iter) {
    // V_eff per channel (driver_spin.rs:263-339)
    std::vector<double> tot_in(N);
    for (size_t i = 0; i < N; ++i) tot_in[i] = up[i] + dn[i];
    std::vector<cplx> tot_g_in;
    density_r_to_g(tot_in, s.dims, tot_g_in);
    std::vector<cplx> vh(N), vg_u, vg_d, veff_u(N), veff_d(N);
    hartree(tot_g_in.data(), s.g2.data(), N, vh.data());
    std::vector<double> exc_in, vu_in, vd_in;
    lsda_grid(add_core(up, core_half), add_core(dn, core_half), exc_in, vu_in, vd_in);
    density_r_to_g(vu_in, s.dims, vg_u);
    density_r_to_g(vd_in, s.dims, vg_d);
    assemble_veff(s.v_local.data(), vh.data(), vg_u.data(), N, veff_u.data());
    assemble_veff(s.v_local.data(), vh.data(), vg_d.data(), N, veff_d.data());

    // eigenpairs per channel (driver_spin.rs:354-392)
    std::vector<std::vector<double>> ev_u, ev_d;
    std::vector<std::vector<cplx>> wf_u, wf_d;
    diag_all(s, kpts, vnl, veff_u, nb, ev_u, wf_u);
    diag_all(s, kpts, vnl, veff_d, nb, ev_d, wf_d);
    std::vector<std::vector<double>> ev_all(ev_u);
    ev_all.insert(ev_all.end(), ev_d.begin(), ev_d.end());
    std::vector<double> w_all(wk);
    w_all.insert(w_all.end(), wk.begin(), wk.end());

    // Fermi level(s) and occupations (driver_spin.rs:407-474)
    double ef, ef_u, ef_d;
    if (prm.fixed_mag) {
      ef_u = find_fermi_energy(ev_u, wk, 0.5 * (s.n_electrons + prm.tot_mag), prm.sigma, spin_factor);
      ef_d = find_fermi_energy(ev_d, wk, (s.n_electrons - prm.tot_mag) / 2.0, prm.sigma, spin_factor);
      ef = 0.5 * (ef_u + ef_d);
    } else {
      ef = find_fermi_energy(ev_all, w_all, s.n_electrons, prm.sigma, spin_factor);
      ef_u = ef_d = ef;
    }
    std::vector<std::vector<double>> occ_u(nk, std::vector<double>(nb)), occ_d(nk, std::vector<double>(nb));
    double n_up = 0.0, n_dn = 0.0;
    for (size_t k = 0; k < nk; ++k)
      for (size_t b = 0; b < nb; ++b) {
        occ_u[k][b] = fermi_dirac_occ(ev_u[k][b], ef_u, prm.sigma) * spin_factor;
        occ_d[k][b] = fermi_dirac_occ(ev_d[k][b], ef_d, prm.sigma) * spin_factor;
      }
    for (size_t k = 0; k < nk; ++k)
      for (size_t b = 0; b < nb; ++b) n_up += occ_u[k][b] * wk[k];
    for (size_t k = 0; k < nk; ++k)
      for (size_t b = 0; b < nb; ++b) n_dn += occ_d[k][b] * wk[k];

    // per-channel densities, symmetrised (driver_spin.rs:477-523)
    std::vector<double> up_new = density_of(s, wf_u, occ_u, wk, nb, n_up);
    std::vector<double> dn_new = density_of(s, wf_d, occ_d, wk, nb, n_dn);
    symmetrize_density(up_new, s.dims, sym);
    symmetrize_density(dn_new, s.dims, sym);
    const double delta = std::max(density_diff(up, up_new, s.omega), density_diff(dn, dn_new, s.omega));
    std::vector<double> tot_new(N);
    for (size_t i = 0; i < N; ++i) tot_new[i] = up_new[i] + dn_new[i];
    std::vector<cplx> tot_g_new;
    density_r_to_g(tot_new, s.dims, tot_g_new);

    // energies: output-density xc (driver_spin.rs:546-604), Harris on the input side
    const std::vector<double> xc_tot_out = add_core(tot_new, s.rho_core);
    std::vector<double> exc_out, vu_out, vd_out;
    lsda_grid(add_core(up_new, core_half), add_core(dn_new, core_half), exc_out, vu_out, vd_out);
    double evxc_out = 0.0;
    for (size_t i = 0; i < N; ++i) evxc_out += (up_new[i] * vu_out[i] + dn_new[i] * vd_out[i]) * dvol;
    double exc_int_out = 0.0;
    for (size_t i = 0; i < N; ++i) exc_int_out += xc_tot_out[i] * exc_out[i] * dvol;
    double e_band = 0.0;
    for (size_t k = 0; k < nk; ++k)
      for (size_t b = 0; b < nb; ++b) e_band += occ_u[k][b] * wk[k] * ev_u[k][b];
    for (size_t k = 0; k < nk; ++k)
      for (size_t b = 0; b < nb; ++b) e_band += occ_d[k][b] * wk[k] * ev_d[k][b];
    const double ts = entropy_ts(ev_all, w_all, ef, prm.sigma, spin_factor);
    const double e_h_out = hartree_energy(tot_g_new, s.g2, s.omega);
    const double e_total = e_band - e_h_out + (exc_int_out - evxc_out) + e_ewald + (-ts);
    const std::vector<double> xc_tot_in = add_core(tot_in, s.rho_core);
    double exc_int_in = 0.0, evxc_in = 0.0;
    for (size_t i = 0; i < N; ++i) exc_int_in += xc_tot_in[i] * exc_in[i] * dvol;
    for (size_t i = 0; i < N; ++i) evxc_in += (up[i] * vu_in[i] + dn[i] * vd_in[i]) * dvol;
    const double e_harris =
        e_band - hartree_energy(tot_g_in, s.g2, s.omega) + (exc_int_in - evxc_in) + e_ewald + (-ts);
    const double de = have_prev ? std::fabs(e_total - e_prev) : INFINITY;
    const bool energy_conv = have_prev && de < prm.energy_threshold;
    have_prev = true;
    e_prev = e_total;
    res.final_delta = delta;
    res.n_iter = iter + 1;
    if (prm.verbose)
      std::fprintf(stderr, "iter %3zu  E=%.10f  dE=%.3e  delta=%.3e  EF=%.6f  M=%.6f\n", iter, e_total, de, delta, ef,
                   n_up - n_dn);
    if (delta < prm.conv_threshold && energy_conv) {
      res.converged = true;
      res.e_total = e_total;
      res.e_harris = e_harris;
      res.fermi = ef;
      res.ts = ts;
      res.eigenvalues = ev_all;
      res.rho_r = tot_new;
      res.n_up = n_up;
      res.n_dn = n_dn;
      res.magnetization = std::fabs(n_up - n_dn);
      EnergyComponents& c = res.comp;
      c.e_band = e_band;
      c.e_ewald = e_ewald;
      c.e_smearing = -ts;
      c.e_hartree = e_h_out;
      c.e_xc = exc_int_out;
      c.e_vxc = evxc_out;
      kin_nl_expect(s, kpts, vnl, wf_u, occ_u, nb, c.e_kinetic, c.e_nonlocal);
      kin_nl_expect(s, kpts, vnl, wf_d, occ_d, nb, c.e_kinetic, c.e_nonlocal);
      std::vector<cplx> vl(s.v_local);
      fft3d(vl, s.dims, +1);
      double eloc = 0.0;
      for (size_t i = 0; i < N; ++i) eloc += tot_new[i] * vl[i].real() * dvol;
      c.e_local = eloc;
      return res;
    }
    // coupled-channel mixing (driver_spin.rs:800-832)
    std::vector<double> m_in(N), m_out(N);
    for (size_t i = 0; i < N; ++i) {
      m_in[i] = up[i] - dn[i];
      m_out[i] = up_new[i] - dn_new[i];
    }
    const std::vector<double> t_mix = mix_tot.mix(tot_in, tot_new);
    const std::vector<double> m_mix = mix_mag.mix(m_in, m_out);
    for (size_t i = 0; i < N; ++i) {
      up[i] = 0.5 * (t_mix[i] + m_mix[i]);
      dn[i] = 0.5 * (t_mix[i] - m_mix[i]);
    }
  }
  res.converged = false;
  return res;
}

}  // namespace orc
