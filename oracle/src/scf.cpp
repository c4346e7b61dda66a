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

}  // namespace orc
