// XC (PZ-LDA, LSDA), Hartree, V_eff, energy reductions, smearing, Anderson
// mixing. Test infrastructure (oracle).
#include <stdexcept>

#include "oracle.hpp"

namespace orc {

// --- src/potential/xc.rs:223-238 -------------------------------------------
static void slater(double rho, double* ex, double* vx) {
  double rb = rho * BOHR3_TO_ANG3;
  double ex_ha = -0.75 * std::cbrt(3.0 * rb / PI);
  double vx_ha = (4.0 / 3.0) * ex_ha;
  *ex = ex_ha * HA_TO_EV;
  *vx = vx_ha * HA_TO_EV;
}

// --- src/potential/xc.rs:567-596 (pz_correlation_rs, Hartree units) --------
static void pz_rs(double rs, bool pol, double* ec, double* vc) {
  if (rs >= 1.0) {
    double gamma = pol ? -0.0843 : -0.1423, b1 = pol ? 1.3981 : 1.0529, b2 = pol ? 0.2611 : 0.3334;
    double srs = std::sqrt(rs);
    double den = std::fma(b2, rs, std::fma(b1, srs, 1.0));
    *ec = gamma / den;
    double dec = -gamma * (b1 / (2.0 * srs) + b2) / (den * den);
    *vc = std::fma(rs / 3.0, -dec, *ec);
  } else {
    double a = pol ? 0.01555 : 0.0311, b = pol ? -0.0269 : -0.048, c = pol ? 0.0007 : 0.0020,
           d = pol ? -0.0048 : -0.0116;
    double lr = std::log(rs);
    *ec = std::fma(d, rs, std::fma(c * rs, lr, std::fma(a, lr, b)));
    double dec = std::fma(c, lr + 1.0, a / rs) + d;
    *vc = std::fma(rs / 3.0, -dec, *ec);
  }
}

// --- src/potential/xc.rs:247-288 (unpolarised; separate expression form) ---
static void pz_unpol(double rho, double* ec, double* vc) {
  double rb = rho * BOHR3_TO_ANG3;
  double rs = std::cbrt(3.0 / (4.0 * PI * rb));
  double ec_ha, vc_ha;
  if (rs >= 1.0) {
    const double gamma = -0.1423, b1 = 1.0529, b2 = 0.3334;
    double srs = std::sqrt(rs);
    double den = 1.0 + b1 * srs + b2 * rs;
    ec_ha = gamma / den;
    double dec = -gamma * (b1 / (2.0 * srs) + b2) / (den * den);
    vc_ha = std::fma(rs / 3.0, -dec, ec_ha);
  } else {
    const double a = 0.0311, b = -0.048, c = 0.0020, d = -0.0116;
    double lr = std::log(rs);
    ec_ha = std::fma(d, rs, std::fma(c * rs, lr, std::fma(a, lr, b)));
    double dec = std::fma(c, lr + 1.0, a / rs) + d;
    vc_ha = std::fma(rs / 3.0, -dec, ec_ha);
  }
  *ec = ec_ha * HA_TO_EV;
  *vc = vc_ha * HA_TO_EV;
}

// src/potential/xc.rs:129-141
void lda_xc(double rho, double* exc, double* vxc) {
  if (rho < RHO_FLOOR) {
    *exc = 0.0;
    *vxc = 0.0;
    return;
  }
  double ex, vx, ec, vc;
  slater(rho, &ex, &vx);
  pz_unpol(rho, &ec, &vc);
  *exc = ex + ec;
  *vxc = vx + vc;
}

// src/potential/xc.rs:374-392, 465-561
void lsda_xc(double ru, double rd, double* exc, double* vup, double* vdn) {
  double rho = ru + rd;
  if (rho < RHO_FLOOR) {
    *exc = *vup = *vdn = 0.0;
    return;
  }
  // exchange (xc.rs:465-506)
  double rub = ru * BOHR3_TO_ANG3, rdb = rd * BOHR3_TO_ANG3;
  double exu = rub > RHO_FLOOR ? -0.75 * std::cbrt(6.0 * rub / PI) : 0.0;
  double exd = rdb > RHO_FLOOR ? -0.75 * std::cbrt(6.0 * rdb / PI) : 0.0;
  double rb = rho * BOHR3_TO_ANG3;
  double ex_ha = std::fma(rub, exu, rdb * exd) / rb;
  double vxu = (4.0 / 3.0) * exu, vxd = (4.0 / 3.0) * exd;
  // correlation (xc.rs:514-561)
  double rs = std::cbrt(3.0 / (4.0 * PI * rb));
  double zeta = std::min(std::max((ru - rd) / rho, -1.0), 1.0);
  double ecu, vcu, ecp, vcp;
  pz_rs(rs, false, &ecu, &vcu);
  pz_rs(rs, true, &ecp, &vcp);
  double fden = std::fma(std::cbrt(2.0), 2.0, -2.0);
  double op = std::max(1.0 + zeta, 0.0), om = std::max(1.0 - zeta, 0.0);
  double cop = std::cbrt(op), com = std::cbrt(om);
  double fz = (cop * op + com * om - 2.0) / fden;
  double dfz = (4.0 / 3.0) * (cop - com) / fden;
  double ec_ha = ecu + fz * (ecp - ecu);
  double vrs = vcu + fz * (vcp - vcu);
  double dedz = dfz * (ecp - ecu);
  double vcu_ha = std::fma(1.0 - zeta, dedz, vrs);
  double vcd_ha = std::fma(1.0 + zeta, -dedz, vrs);
  *exc = ex_ha * HA_TO_EV + ec_ha * HA_TO_EV;
  *vup = vxu * HA_TO_EV + vcu_ha * HA_TO_EV;
  *vdn = vxd * HA_TO_EV + vcd_ha * HA_TO_EV;
}

// src/scf/energy.rs:343-358
void hartree(const cplx* rho_g, const double* g2, size_t n, cplx* v_h) {
  const double fourpi_e2 = 4.0 * PI * E2_COULOMB;
  for (size_t i = 0; i < n; ++i) v_h[i] = g2[i] > G2_ZERO_THRESHOLD ? rho_g[i] * fourpi_e2 / g2[i] : cplx(0.0, 0.0);
}

// src/scf/energy.rs:293-301
void assemble_veff(const cplx* vl, const cplx* vh, const cplx* vxc, size_t n, cplx* out) {
  for (size_t i = 0; i < n; ++i) out[i] = vl[i] + vh[i] + vxc[i];
}

// src/scf/energy.rs:250-259
void density_r_to_g(const std::vector<double>& rho_r, const std::array<size_t, 3>& dims, std::vector<cplx>& rho_g) {
  size_t n = rho_r.size();
  rho_g.resize(n);
  for (size_t i = 0; i < n; ++i) rho_g[i] = cplx(rho_r[i], 0.0);
  fft3d(rho_g, dims, -1);
  double norm = 1.0 / (double)n;
  for (auto& v : rho_g) v *= norm;
}

// src/scf/energy.rs:81-96
double hartree_energy(const std::vector<cplx>& rho_g, const std::vector<double>& g2, double omega) {
  const double fourpi_e2 = 4.0 * PI * E2_COULOMB;
  double s = 0.0;
  for (size_t i = 0; i < rho_g.size(); ++i)
    if (g2[i] > G2_ZERO_THRESHOLD) s += std::norm(rho_g[i]) * fourpi_e2 / g2[i];
  return s * 0.5 * omega;
}

// src/scf/energy.rs:120-132 with lda_xc_energy (xc.rs:207-216)
double xc_energy_corrected(const std::vector<double>& rho_xc, const std::vector<double>& rho_val,
                           const std::vector<double>& exc, const std::vector<double>& vxc, double omega) {
  double dvol = omega / (double)rho_xc.size();
  double e_xc = 0.0, e_vxc = 0.0;
  for (size_t i = 0; i < rho_xc.size(); ++i) e_xc += rho_xc[i] * exc[i] * dvol;
  for (size_t i = 0; i < rho_val.size(); ++i) e_vxc += rho_val[i] * vxc[i] * dvol;
  return e_xc - e_vxc;
}

// src/scf/energy.rs:227-235
double density_diff(const std::vector<double>& a, const std::vector<double>& b, double omega) {
  double dvol = omega / (double)a.size();
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) {
    double d = a[i] - b[i];
    s += d * d * dvol;
  }
  return std::sqrt(s / omega);
}

// --- src/scf/smearing.rs:112-130 ---------------------------------------------
double fermi_dirac_occ(double e, double ef, double sigma) {
  if (sigma < 1e-15) return e < ef ? 1.0 : (std::fabs(e - ef) < 1e-12 ? 0.5 : 0.0);
  double x = (e - ef) / sigma;
  if (x > 40.0) return 0.0;
  if (x < -40.0) return 1.0;
  return 1.0 / (1.0 + std::exp(x));
}

// src/scf/smearing.rs:49-90
double find_fermi_energy(const std::vector<std::vector<double>>& evs, const std::vector<double>& w, double n_el,
                         double sigma, double spin_factor) {
  double e_min = INFINITY, e_max = -INFINITY;
  for (auto& v : evs)
    for (double e : v) {
      e_min = std::min(e_min, e);
      e_max = std::max(e_max, e);
    }
  e_min -= 10.0 * std::max(sigma, 0.1);
  e_max += 10.0 * std::max(sigma, 0.1);
  for (int it = 0; it < 200; ++it) {
    double mid = 0.5 * (e_min + e_max);
    double n = 0.0;
    for (size_t k = 0; k < evs.size(); ++k) {
      double sk = 0.0;
      for (double e : evs[k]) sk += w[k] * (fermi_dirac_occ(e, mid, sigma) * spin_factor);
      n += sk;
    }
    if (n < n_el)
      e_min = mid;
    else
      e_max = mid;
    if (std::fabs(e_max - e_min) < 1e-14) break;
  }
  return 0.5 * (e_min + e_max);
}

// src/scf/smearing.rs:202-223 (Fermi-Dirac entropy weight)
double entropy_ts(const std::vector<std::vector<double>>& evs, const std::vector<double>& w, double ef, double sigma,
                  double spin_factor) {
  if (sigma < 1e-15) return 0.0;
  double s = 0.0;
  for (size_t k = 0; k < evs.size(); ++k)
    for (double e : evs[k]) {
      double x = (e - ef) / sigma;
      double wt = 0.0;
      if (std::fabs(x) <= 30.0) {
        double f = 1.0 / (1.0 + std::exp(x));
        f = std::min(std::max(f, 1e-30), 1.0 - 1e-30);
        wt = -(f * std::log(f) + (1.0 - f) * std::log(1.0 - f));
      }
      s += w[k] * wt;
    }
  return s * sigma * spin_factor;
}

// --- small LU with partial pivoting (src/scf/mixing/linalg.rs:24-52) ---------
static std::vector<double> solve_small(std::vector<double> a, std::vector<double> b, size_t n) {
  if (n == 0) return {};
  std::vector<size_t> perm(n);
  for (size_t i = 0; i < n; ++i) perm[i] = i;
  for (size_t c = 0; c < n; ++c) {
    size_t piv = c;
    for (size_t r = c + 1; r < n; ++r)
      if (std::fabs(a[r * n + c]) > std::fabs(a[piv * n + c])) piv = r;
    if (piv != c) {
      for (size_t j = 0; j < n; ++j) std::swap(a[c * n + j], a[piv * n + j]);
      std::swap(b[c], b[piv]);
    }
    if (std::fabs(a[c * n + c]) < 1e-15) return std::vector<double>(n, 1.0 / (double)(n + 1));
    for (size_t r = c + 1; r < n; ++r) {
      double f = a[r * n + c] / a[c * n + c];
      for (size_t j = c; j < n; ++j) a[r * n + j] -= f * a[c * n + j];
      b[r] -= f * b[c];
    }
  }
  std::vector<double> x(n);
  for (size_t i = n; i-- > 0;) {
    double s = b[i];
    for (size_t j = i + 1; j < n; ++j) s -= a[i * n + j] * x[j];
    x[i] = s / a[i * n + i];
  }
  return x;
}

// src/scf/mixing/anderson.rs:170-266 (plain mode, no Kerker, fixed β)
double kerker_auto_qtf2(double n_electrons, double omega) {
  const double bohr = 0.529177210903;
  const double rho_bohr = n_electrons / omega * bohr * bohr * bohr;
  const double pi = 3.14159265358979323846;
  const double q2_bohr = 4.0 * std::cbrt(3.0 * pi * pi * rho_bohr) / pi;
  return q2_bohr / (bohr * bohr);
}

std::vector<double> Anderson::mix(const std::vector<double>& rho_in, const std::vector<double>& rho_out) {
  size_t n = rho_in.size();
  std::vector<double> res(n);
  for (size_t i = 0; i < n; ++i) res[i] = rho_out[i] - rho_in[i];
  if (!kerker_w.empty()) {
    // mixing/kerker.rs:5-29: forward FFT, weight per G, inverse FFT, 1/N
    std::vector<cplx> rg(n);
    for (size_t i = 0; i < n; ++i) rg[i] = cplx(res[i], 0.0);
    fft3d(rg, dims, -1);
    for (size_t i = 0; i < n; ++i) rg[i] *= kerker_w[i];
    fft3d(rg, dims, +1);
    const double inv = 1.0 / (double)n;
    for (size_t i = 0; i < n; ++i) res[i] = rg[i].real() * inv;
  }
  hist_in.push_back(rho_in);
  hist_res.push_back(res);
  if (hist_in.size() > max_history) {
    hist_in.erase(hist_in.begin());
    hist_res.erase(hist_res.begin());
  }
  size_t m = hist_in.size();
  if (m < 2) {
    std::vector<double> out(n);
    for (size_t i = 0; i < n; ++i) out[i] = hist_in[0][i] + beta * hist_res[0][i];
    return out;
  }
  size_t last = m - 1, mm = m - 1;
  const std::vector<double>& rl = hist_res[last];
  std::vector<std::vector<double>> dr(mm, std::vector<double>(n));
  for (size_t i = 0; i < mm; ++i)
    for (size_t g = 0; g < n; ++g) dr[i][g] = hist_res[i][g] - rl[g];
  std::vector<double> A(mm * mm), b(mm);
  for (size_t i = 0; i < mm; ++i) {
    double s = 0.0;
    for (size_t g = 0; g < n; ++g) s += dr[i][g] * rl[g];
    b[i] = -s;
    for (size_t j = 0; j < mm; ++j) {
      double t = 0.0;
      for (size_t g = 0; g < n; ++g) t += dr[i][g] * dr[j][g];
      A[i * mm + j] = t;
    }
  }
  std::vector<double> alpha = solve_small(A, b, mm);
  double asum = 0.0;
  for (double a : alpha) asum += a;
  double alast = 1.0 - asum;
  std::vector<double> out(n);
  for (size_t g = 0; g < n; ++g) out[g] = alast * (hist_in[last][g] + beta * rl[g]);
  for (size_t j = 0; j < mm; ++j)
    for (size_t g = 0; g < n; ++g) out[g] += alpha[j] * (hist_in[j][g] + beta * hist_res[j][g]);
  return out;
}

}  // namespace orc
