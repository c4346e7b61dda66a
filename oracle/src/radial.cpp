// UPF v2 parsing, radial quadrature, form factors, spherical harmonics.
// Test infrastructure (oracle).
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "oracle.hpp"

namespace orc {

// --- text helpers: src/pseudopotential/upf/xml.rs:10-91 ---------------------
static bool extract_attr(const std::string& s, const std::string& name, std::string& out) {
  std::string pat = name + "=\"";
  size_t p = s.find(pat);
  if (p == std::string::npos) return false;
  size_t start = p + pat.size();
  size_t end = s.find('"', start);
  if (end == std::string::npos) return false;
  out = s.substr(start, end - start);
  return true;
}

static std::vector<double> extract_block(const std::string& s, const std::string& tag, size_t expected,
                                         bool* ok = nullptr) {
  std::string open = "<" + tag, close = "</" + tag + ">";
  size_t tp = s.find(open);
  if (tp == std::string::npos) {
    if (ok) {
      *ok = false;
      return {};
    }
    throw std::runtime_error("missing tag <" + tag + ">");
  }
  size_t ds = s.find('>', tp) + 1;
  size_t de = s.find(close, ds);
  if (de == std::string::npos) throw std::runtime_error("missing closing tag " + tag);
  std::vector<double> v;
  v.reserve(expected);
  const char* c = s.c_str() + ds;
  const char* e = s.c_str() + de;
  while (c < e) {
    while (c < e && (*c == ' ' || *c == '\n' || *c == '\t' || *c == '\r')) ++c;
    if (c >= e) break;
    char* endp = nullptr;
    double x = std::strtod(c, &endp);
    if (endp == c) throw std::runtime_error("float parse error in " + tag);
    v.push_back(x);
    c = endp;
  }
  if (v.size() != expected) {
    if (ok) {
      *ok = false;
      return {};
    }
    throw std::runtime_error(tag + ": unexpected size");
  }
  if (ok) *ok = true;
  return v;
}

static std::string trim(const std::string& x) {
  size_t a = x.find_first_not_of(" \t\n\r"), b = x.find_last_not_of(" \t\n\r");
  return a == std::string::npos ? std::string() : x.substr(a, b - a + 1);
}

// src/pseudopotential/upf/parse.rs:20-140
Upf parse_upf(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  std::string s = ss.str();
  if (s.find("<UPF") == std::string::npos && s.find("<PP_HEADER") == std::string::npos)
    throw std::runtime_error("unrecognized pseudopotential format");
  Upf pp;
  std::string a;
  if (!extract_attr(s, "element", a)) throw std::runtime_error("missing element");
  pp.element = trim(a);
  if (!extract_attr(s, "z_valence", a)) throw std::runtime_error("missing z_valence");
  pp.z_valence = std::strtod(trim(a).c_str(), nullptr);
  if (!extract_attr(s, "l_max", a)) throw std::runtime_error("missing l_max");
  pp.l_max = std::atoi(trim(a).c_str());
  if (!extract_attr(s, "number_of_proj", a)) throw std::runtime_error("missing number_of_proj");
  size_t n_proj = (size_t)std::atol(trim(a).c_str());
  if (!extract_attr(s, "mesh_size", a)) throw std::runtime_error("missing mesh_size");
  size_t mesh = (size_t)std::atol(trim(a).c_str());

  pp.r = extract_block(s, "PP_R", mesh);
  for (auto& x : pp.r) x = x * BOHR_TO_ANG;
  pp.rab = extract_block(s, "PP_RAB", mesh);
  for (auto& x : pp.rab) x = x * BOHR_TO_ANG;
  pp.v_local = extract_block(s, "PP_LOCAL", mesh);
  for (auto& x : pp.v_local) x = x * RY_TO_EV;
  double sqrt_b = std::sqrt(BOHR_TO_ANG);
  for (size_t i = 1; i <= n_proj; ++i) {
    std::string tag = "PP_BETA." + std::to_string(i);
    size_t ts = s.find("<" + tag);
    if (ts == std::string::npos) throw std::runtime_error("missing " + tag);
    size_t te = s.find('>', ts);
    std::string tc = s.substr(ts, te - ts);
    if (!extract_attr(tc, "angular_momentum", a)) throw std::runtime_error("missing angular_momentum");
    Beta b;
    b.l = std::atoi(trim(a).c_str());
    if (b.l < 0) throw std::runtime_error("negative angular momentum");
    b.values = extract_block(s, tag, mesh);
    for (auto& x : b.values) x = x / sqrt_b;
    pp.betas.push_back(std::move(b));
  }
  pp.dij = extract_block(s, "PP_DIJ", n_proj * n_proj);
  for (auto& x : pp.dij) x = x * RY_TO_EV;
  bool ok = false;
  pp.rho_atom = extract_block(s, "PP_RHOATOM", mesh, &ok);
  if (ok)
    for (auto& x : pp.rho_atom) x = x / BOHR_TO_ANG;
  else
    pp.rho_atom.assign(mesh, 0.0);
  bool nlcc = s.find("core_correction=\"T\"") != std::string::npos ||
              s.find("core_correction=\"t\"") != std::string::npos || s.find("nlcc=.true.") != std::string::npos;
  if (nlcc) {
    pp.core = extract_block(s, "PP_NLCC", mesh, &ok);
    if (ok)
      for (auto& x : pp.core) x = x / BOHR3_TO_ANG3;
    else
      pp.core.assign(mesh, 0.0);
  }
  return pp;
}

bool Upf::has_rho_atom() const {
  for (double v : rho_atom)
    if (std::fabs(v) > 1e-20) return true;
  return false;
}

// src/numerics.rs:29-58 (QE/DFTK Simpson with the even-mesh correction)
double simpson(const double* f, const double* rab, size_t n) {
  if (n < 3) {
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += f[i] * rab[i];
    return s;
  }
  double sum = 0.0;
  for (size_t i = 1; i < n - 1; ++i) {
    double w = (i % 2 == 1) ? 4.0 : 2.0;
    sum += w * f[i] * rab[i];
  }
  if (n % 2 == 1) return std::fma(f[n - 1], rab[n - 1], std::fma(f[0], rab[0], sum)) / 3.0;
  double acc = std::fma(f[0], rab[0], sum);
  acc = std::fma(f[n - 3] * rab[n - 3], -0.25, acc);
  acc += f[n - 2] * rab[n - 2];
  acc = std::fma(f[n - 1] * rab[n - 1], 1.25, acc);
  return acc / 3.0;
}

// src/pseudopotential/upf/mod.rs:166-213
double v_local_of_g(const Upf& pp, double g, double omega) {
  const double four_pi = 4.0 * PI;
  size_t n = pp.r.size();
  std::vector<double> integ(n);
  if (g < G_ZERO_THRESHOLD) {
    for (size_t i = 0; i < n; ++i) {
      double r = pp.r[i];
      double vs = pp.v_local[i] + pp.z_valence * E2_COULOMB / std::max(r, 1e-20);
      integ[i] = r * r * vs;
    }
    return four_pi / omega * simpson(integ.data(), pp.rab.data(), n);
  }
  const double two_over_sqrt_pi = 2.0 / std::sqrt(PI);
  for (size_t i = 0; i < n; ++i) {
    double r = pp.r[i];
    double eor = r < 1e-20 ? two_over_sqrt_pi : std::erf(r) / r;
    double vs = pp.v_local[i] + pp.z_valence * E2_COULOMB * eor;
    double gr = g * r;
    double sinc = gr < 1e-10 ? 1.0 - gr * gr / 6.0 : std::sin(gr) / gr;
    integ[i] = r * r * vs * sinc;
  }
  double integral = simpson(integ.data(), pp.rab.data(), n);
  double g2 = g * g;
  return four_pi / omega * integral - four_pi * pp.z_valence * E2_COULOMB * std::exp(-g2 / 4.0) / (omega * g2);
}

// src/potential/nonlocal.rs:592-612
double spherical_bessel_j(int l, double x) {
  if (std::fabs(x) < 1e-10) return l == 0 ? 1.0 : 0.0;
  if (l == 0) return std::sin(x) / x;
  if (l == 1) return std::sin(x) / (x * x) - std::cos(x) / x;
  double jlm1 = std::sin(x) / x;
  double jl = std::sin(x) / (x * x) - std::cos(x) / x;
  for (int n = 1; n < l; ++n) {
    double jlp1 = std::fma((double)(2 * n + 1) / x, jl, -jlm1);
    jlm1 = jl;
    jl = jlp1;
  }
  return jl;
}

// src/potential/nonlocal.rs:566-583: F(q) = 4π ∫ χ(r) j_l(qr) r dr
double beta_form_factor(const Upf& pp, int iproj, double q) {
  const Beta& b = pp.betas[(size_t)iproj];
  size_t n = pp.r.size();
  std::vector<double> integ(n);
  for (size_t i = 0; i < n; ++i) {
    double r = pp.r[i];
    integ[i] = b.values[i] * spherical_bessel_j(b.l, q * r) * r;
  }
  return 4.0 * PI * simpson(integ.data(), pp.rab.data(), n);
}

// src/potential/nonlocal.rs:469-552 — real Y_lm in the order
// (0,0),(1,0),(1,+1),(1,-1),(2,0),(2,+1),(2,-1),(2,+2),(2,-2),...
void real_sph_harmonics(V3 q, int lmax, double* out) {
  const double fpi = 4.0 * PI;
  const double inv_sqrt_fpi = std::sqrt(1.0 / fpi);
  if (lmax == 0) {
    out[0] = inv_sqrt_fpi;
    return;
  }
  size_t total = (size_t)((lmax + 1) * (lmax + 1));
  for (size_t i = 0; i < total; ++i) out[i] = 0.0;
  out[0] = inv_sqrt_fpi;
  double gmod = norm(q);
  if (gmod < 1e-9) return;
  double cost = q.z / gmod;
  double sint = std::sqrt(std::max(1.0 - cost * cost, 0.0));
  double phi = std::atan2(q.y, q.x);
  size_t lm1 = (size_t)(lmax + 1);
  std::vector<double> Q(lm1 * lm1, 0.0);
  auto qi = [lm1](size_t l, size_t m) { return l * lm1 + m; };
  Q[qi(0, 0)] = 1.0;
  Q[qi(1, 0)] = cost;
  Q[qi(1, 1)] = -sint / std::sqrt(2.0);
  double c1 = std::sqrt(3.0 / fpi);
  out[1] = c1 * Q[qi(1, 0)];
  out[2] = c1 * std::sqrt(2.0) * Q[qi(1, 1)] * std::cos(phi);
  out[3] = c1 * std::sqrt(2.0) * Q[qi(1, 1)] * std::sin(phi);
  size_t lm = 4;
  for (size_t l = 2; l <= (size_t)lmax; ++l) {
    double c = std::sqrt((double)(2 * l + 1) / fpi);
    double lf = (double)l;
    for (size_t m = 0; m + 2 <= l; ++m) {
      double mf = (double)m;
      double llmm = std::sqrt(lf * lf - mf * mf);
      double llm1 = std::sqrt((lf - 1.0) * (lf - 1.0) - mf * mf);
      Q[qi(l, m)] = (cost * (2.0 * lf - 1.0) * Q[qi(l - 1, m)] - llm1 * Q[qi(l - 2, m)]) / llmm;
    }
    Q[qi(l, l - 1)] = cost * std::sqrt(2.0 * lf - 1.0) * Q[qi(l - 1, l - 1)];
    Q[qi(l, l)] = -std::sqrt((2.0 * lf - 1.0) / (2.0 * lf)) * sint * Q[qi(l - 1, l - 1)];
    out[lm] = c * Q[qi(l, 0)];
    for (size_t m = 1; m <= l; ++m) {
      double mf = (double)m;
      out[lm + 2 * m - 1] = c * std::sqrt(2.0) * Q[qi(l, m)] * std::cos(mf * phi);
      out[lm + 2 * m] = c * std::sqrt(2.0) * Q[qi(l, m)] * std::sin(mf * phi);
    }
    lm += 2 * l + 1;
  }
}

}  // namespace orc
