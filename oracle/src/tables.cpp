// Bit-exact tables: lattice, basis, FFT grid maps, Monkhorst-Pack k-points,
// symmetry detection and IBZ reduction. Test infrastructure (oracle).
#include <algorithm>
#include <cstdlib>
#include <stdexcept>

#include "oracle.hpp"

namespace orc {

// src/crystal.rs:108-110
double Lattice::volume() const { return std::fabs(dot(cross(a, b), c)); }

// src/crystal.rs:112-121 (signed triple product)
Lattice Lattice::reciprocal() const {
  double triple = dot(cross(a, b), c);
  double factor = 2.0 * PI / triple;
  return {cross(b, c) * factor, cross(c, a) * factor, cross(a, b) * factor};
}

// src/crystal.rs:77-80
V3 Crystal::cart(const Atom& at) const {
  return at.frac[0] * lat.a + at.frac[1] * lat.b + at.frac[2] * lat.c;
}

// src/basis.rs:36-63
Basis make_basis(const Lattice& lat, double ecut) {
  Lattice recip = lat.reciprocal();
  double g_max_sq = ecut / HBAR2_OVER_2M;
  double g_max = std::sqrt(g_max_sq);
  int n1_max = (int)std::ceil(g_max / norm(recip.a));
  int n2_max = (int)std::ceil(g_max / norm(recip.b));
  int n3_max = (int)std::ceil(g_max / norm(recip.c));
  Basis b;
  b.ecut = ecut;
  for (int n1 = -n1_max; n1 <= n1_max; ++n1)
    for (int n2 = -n2_max; n2 <= n2_max; ++n2)
      for (int n3 = -n3_max; n3 <= n3_max; ++n3) {
        V3 g = (double)n1 * recip.a + (double)n2 * recip.b + (double)n3 * recip.c;
        if (norm2(g) <= g_max_sq) {
          b.g.push_back(g);
          b.miller.push_back({n1, n2, n3});
        }
      }
  return b;
}

// src/fft.rs:269-279
bool is_fft_friendly(size_t n) {
  if (n == 0) return false;
  for (size_t p : {2, 3, 5})
    while (n % p == 0) n /= p;
  return n == 1;
}

// src/fft.rs:257-266
size_t fft_grid_size(uint32_t n_max) {
  size_t n = 2 * (size_t)n_max + 1;
  while (!is_fft_friendly(n)) ++n;
  return n;
}

// src/scf/grid.rs:44-86
std::array<size_t, 3> fft_dims(const Basis& basis, uint32_t ratio, const size_t* explicit_dims) {
  std::array<size_t, 3> dims{};
  if (explicit_dims) {
    dims = {explicit_dims[0], explicit_dims[1], explicit_dims[2]};
  } else {
    uint32_t n_max[3] = {0, 0, 0};
    for (auto& m : basis.miller)
      for (int d = 0; d < 3; ++d) n_max[d] = std::max(n_max[d], (uint32_t)std::abs(m[d]));
    uint32_t is = 0;
    while ((is + 1) * (is + 1) <= ratio) ++is;  // integer sqrt
    uint32_t scale = (is * is == ratio) ? is : is + 1;
    for (int d = 0; d < 3; ++d) dims[d] = fft_grid_size(scale * n_max[d]);
  }
  if (dims[0] > 1024 || dims[1] > 1024 || dims[2] > 1024) throw std::runtime_error("FFT grid exceeds MAX_FFT_DIM");
  return dims;
}

// src/scf/grid.rs:142-147
size_t miller_to_idx(const std::array<size_t, 3>& dims, int n1, int n2, int n3) {
  int d0 = (int)dims[0], d1 = (int)dims[1], d2 = (int)dims[2];
  size_t i1 = (size_t)((n1 % d0) + d0) % dims[0];
  size_t i2 = (size_t)((n2 % d1) + d1) % dims[1];
  size_t i3 = (size_t)((n3 % d2) + d2) % dims[2];
  return i1 * dims[1] * dims[2] + i2 * dims[2] + i3;
}

// src/scf/grid.rs:120-129
V3 g_vector_at(size_t idx, const std::array<size_t, 3>& dims, const Lattice& recip) {
  size_t nx = dims[0], ny = dims[1], nz = dims[2];
  size_t i1 = idx / (ny * nz), i2 = (idx / nz) % ny, i3 = idx % nz;
  int n1 = i1 > nx / 2 ? (int)i1 - (int)nx : (int)i1;
  int n2 = i2 > ny / 2 ? (int)i2 - (int)ny : (int)i2;
  int n3 = i3 > nz / 2 ? (int)i3 - (int)nz : (int)i3;
  return (double)n1 * recip.a + (double)n2 * recip.b + (double)n3 * recip.c;
}

// src/scf/grid.rs:100-106
std::vector<size_t> basis_to_fft(const Basis& basis, const std::array<size_t, 3>& dims) {
  std::vector<size_t> out;
  out.reserve(basis.size());
  for (auto& m : basis.miller) out.push_back(miller_to_idx(dims, m[0], m[1], m[2]));
  return out;
}

// ---------------------------------------------------------------------------
// k-points — src/kpoints.rs:245-288
// ---------------------------------------------------------------------------
void mp_fractional_coord(uint32_t i1, uint32_t i2, uint32_t i3, const uint32_t g[3], const uint32_t s[3],
                         double out[3]) {
  uint32_t idx[3] = {i1, i2, i3};
  for (int a = 0; a < 3; ++a) {
    double raw = (double)idx[a] / (double)g[a] + (double)s[a] / (2.0 * (double)g[a]);
    out[a] = raw - std::floor(raw + 0.5);
  }
}

std::vector<KPoint> monkhorst_pack(const uint32_t g[3], const uint32_t s[3], const Lattice& lat) {
  Lattice recip = lat.reciprocal();
  double weight = 1.0 / (double)(g[0] * g[1] * g[2]);
  std::vector<KPoint> out;
  for (uint32_t i1 = 0; i1 < g[0]; ++i1)
    for (uint32_t i2 = 0; i2 < g[1]; ++i2)
      for (uint32_t i3 = 0; i3 < g[2]; ++i3) {
        double f[3];
        mp_fractional_coord(i1, i2, i3, g, s, f);
        out.push_back({f[0] * recip.a + f[1] * recip.b + f[2] * recip.c, weight});
      }
  return out;
}

// ---------------------------------------------------------------------------
// Symmetry detection — src/symmetry/detect.rs:22-178, operations.rs
// ---------------------------------------------------------------------------
static void wrap_unit(double f[3]) {
  for (int i = 0; i < 3; ++i) f[i] = f[i] - std::floor(f[i]);
}

static double frac_distance(const double a[3], const double b[3]) {
  double d2 = 0.0;
  for (int i = 0; i < 3; ++i) {
    double dx = a[i] - b[i];
    dx -= std::round(dx);
    d2 += dx * dx;
  }
  return std::sqrt(d2);
}

static void apply_rot(const int r[3][3], const double f[3], double out[3]) {
  for (int i = 0; i < 3; ++i) out[i] = (double)r[i][0] * f[0] + (double)r[i][1] * f[1] + (double)r[i][2] * f[2];
}

// metric M = L^T L with L's columns the lattice vectors (nalgebra 3×3 product)
static void metric_of(const Lattice& lat, double m[3][3]) {
  const V3 col[3] = {lat.a, lat.b, lat.c};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) m[i][j] = col[i].x * col[j].x + col[i].y * col[j].y + col[i].z * col[j].z;
}

static double dot_metric(const int c[3], const int d[3], const double m[3][3]) {
  double sum = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) sum += (double)c[i] * m[i][j] * (double)d[j];
  return sum;
}

static std::vector<std::array<int, 3>> candidate_vectors(const double m[3][3], double target, double tol) {
  int range[3];
  for (int i = 0; i < 3; ++i) range[i] = (int)(std::sqrt(target / m[i][i]) + 1.5);
  std::vector<std::array<int, 3>> out;
  for (int n0 = -range[0]; n0 <= range[0]; ++n0)
    for (int n1 = -range[1]; n1 <= range[1]; ++n1)
      for (int n2 = -range[2]; n2 <= range[2]; ++n2) {
        int c[3] = {n0, n1, n2};
        if (std::fabs(dot_metric(c, c, m) - target) < tol) out.push_back({n0, n1, n2});
      }
  return out;
}

static bool all_atoms_map(const Crystal& cr, const int r[3][3], const double tau[3], double tol) {
  for (auto& at : cr.atoms) {
    double rp[3];
    apply_rot(r, at.frac, rp);
    double mapped[3] = {rp[0] + tau[0], rp[1] + tau[1], rp[2] + tau[2]};
    wrap_unit(mapped);
    bool found = false;
    for (auto& o : cr.atoms)
      if (o.z == at.z && frac_distance(mapped, o.frac) < tol) {
        found = true;
        break;
      }
    if (!found) return false;
  }
  return true;
}

Symmetry find_symmetry(const Crystal& cr, double tol) {
  double m[3][3];
  metric_of(cr.lat, m);
  std::vector<std::array<int, 3>> cand[3];
  for (int j = 0; j < 3; ++j) cand[j] = candidate_vectors(m, m[j][j], tol);
  Symmetry sym;
  sym.tolerance = tol;
  sym.time_reversal = true;  // symmetry/mod.rs:40
  for (auto& c0 : cand[0])
    for (auto& c1 : cand[1]) {
      if (std::fabs(dot_metric(c0.data(), c1.data(), m) - m[0][1]) > tol) continue;
      for (auto& c2 : cand[2]) {
        int det = c0[0] * (c1[1] * c2[2] - c1[2] * c2[1]) - c0[1] * (c1[0] * c2[2] - c1[2] * c2[0]) +
                  c0[2] * (c1[0] * c2[1] - c1[1] * c2[0]);
        if (det != 1 && det != -1) continue;
        if (std::fabs(dot_metric(c0.data(), c2.data(), m) - m[0][2]) > tol) continue;
        if (std::fabs(dot_metric(c1.data(), c2.data(), m) - m[1][2]) > tol) continue;
        int rot[3][3] = {{c0[0], c1[0], c2[0]}, {c0[1], c1[1], c2[1]}, {c0[2], c1[2], c2[2]}};
        // detect.rs:118-143: translation from the first atom onto each same-species atom
        if (cr.atoms.empty()) {
          SymOp op;
          std::copy(&rot[0][0], &rot[0][0] + 9, &op.rot[0][0]);
          op.tau[0] = op.tau[1] = op.tau[2] = 0.0;
          sym.ops.push_back(op);
          continue;
        }
        const Atom& ref = cr.atoms[0];
        double rpos[3];
        apply_rot(rot, ref.frac, rpos);
        for (auto& tgt : cr.atoms) {
          if (tgt.z != ref.z) continue;
          double tau[3] = {tgt.frac[0] - rpos[0], tgt.frac[1] - rpos[1], tgt.frac[2] - rpos[2]};
          wrap_unit(tau);
          if (all_atoms_map(cr, rot, tau, tol)) {
            SymOp op;
            std::copy(&rot[0][0], &rot[0][0] + 9, &op.rot[0][0]);
            double t[3] = {tau[0], tau[1], tau[2]};
            wrap_unit(t);  // SpaceGroupOp::new wraps again
            op.tau[0] = t[0];
            op.tau[1] = t[1];
            op.tau[2] = t[2];
            sym.ops.push_back(op);
            break;
          }
        }
      }
    }
  return sym;
}

Symmetry identity_symmetry() {
  Symmetry s;
  SymOp id{};
  id.rot[0][0] = id.rot[1][1] = id.rot[2][2] = 1;
  s.ops.push_back(id);
  s.time_reversal = false;
  return s;
}

static int det3(const int r[3][3]) {
  return r[0][0] * (r[1][1] * r[2][2] - r[1][2] * r[2][1]) - r[0][1] * (r[1][0] * r[2][2] - r[1][2] * r[2][0]) +
         r[0][2] * (r[1][0] * r[2][1] - r[1][1] * r[2][0]);
}

// (R^{-1})^T via the adjugate — symmetry/operations.rs SymmOp::inverse / inverse_transpose
static void inverse_transpose(const int r[3][3], int out[3][3]) {
  int d = det3(r);
  int adj[3][3] = {
      {(r[1][1] * r[2][2] - r[1][2] * r[2][1]) * d, (r[0][2] * r[2][1] - r[0][1] * r[2][2]) * d,
       (r[0][1] * r[1][2] - r[0][2] * r[1][1]) * d},
      {(r[1][2] * r[2][0] - r[1][0] * r[2][2]) * d, (r[0][0] * r[2][2] - r[0][2] * r[2][0]) * d,
       (r[0][2] * r[1][0] - r[0][0] * r[1][2]) * d},
      {(r[1][0] * r[2][1] - r[1][1] * r[2][0]) * d, (r[0][1] * r[2][0] - r[0][0] * r[2][1]) * d,
       (r[0][0] * r[1][1] - r[0][1] * r[1][0]) * d}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) out[i][j] = adj[j][i];
}

// src/symmetry/kpoints.rs:97-143
static long frac_to_grid_index(const double f3[3], const uint32_t g[3], const uint32_t s[3]) {
  uint32_t ind[3];
  for (int j = 0; j < 3; ++j) {
    double n = (double)g[j];
    double f = f3[j] - std::floor(f3[j]);
    double i_f = f * n - 0.5 * (double)s[j];
    double i_r = std::round(i_f);
    if (std::fabs(i_f - i_r) > 1e-6) return -1;
    int i = (int)i_r;
    ind[j] = (uint32_t)((i % (int)g[j]) + (int)g[j]) % g[j];
  }
  return (long)(ind[0] * g[1] * g[2] + ind[1] * g[2] + ind[2]);
}

// src/symmetry/kpoints.rs:31-95
std::vector<KPoint> reduce_kpoints(const std::vector<KPoint>& full, const uint32_t g[3], const uint32_t s[3],
                                   const Symmetry& sym, const Lattice& lat) {
  Lattice recip = lat.reciprocal();
  size_t n_total = full.size();
  std::vector<std::array<double, 3>> frac;
  for (uint32_t i1 = 0; i1 < g[0]; ++i1)
    for (uint32_t i2 = 0; i2 < g[1]; ++i2)
      for (uint32_t i3 = 0; i3 < g[2]; ++i3) {
        std::array<double, 3> f;
        mp_fractional_coord(i1, i2, i3, g, s, f.data());
        frac.push_back(f);
      }
  std::vector<bool> visited(n_total, false);
  std::vector<KPoint> out;
  for (size_t idx = 0; idx < n_total; ++idx) {
    if (visited[idx]) continue;
    const double* fr = frac[idx].data();
    int orbit = 0;
    for (auto& op : sym.ops) {
      int rit[3][3];
      inverse_transpose(op.rot, rit);
      double kr[3];
      apply_rot(rit, fr, kr);
      long ri = frac_to_grid_index(kr, g, s);
      if (ri >= 0 && !visited[(size_t)ri]) {
        visited[(size_t)ri] = true;
        ++orbit;
      }
      if (sym.time_reversal) {
        double kn[3] = {-kr[0], -kr[1], -kr[2]};
        long ni = frac_to_grid_index(kn, g, s);
        if (ni >= 0 && !visited[(size_t)ni]) {
          visited[(size_t)ni] = true;
          ++orbit;
        }
      }
    }
    V3 kc = fr[0] * recip.a + fr[1] * recip.b + fr[2] * recip.c;
    out.push_back({kc, (double)orbit / (double)n_total});
  }
  return out;
}

// src/symmetry/density/g_space.rs:147-252
void symmetrize_density(std::vector<double>& rho_r, const std::array<size_t, 3>& dims, const Symmetry& sym) {
  size_t n_ops = sym.ops.size();
  if (n_ops <= 1) return;
  size_t n = dims[0] * dims[1] * dims[2];
  std::vector<cplx> rho_g(n);
  for (size_t i = 0; i < n; ++i) rho_g[i] = cplx(rho_r[i], 0.0);
  fft3d(rho_g, dims, -1);
  std::vector<cplx> out(n);
  double inv = 1.0 / (double)n_ops;
  for (size_t idx = 0; idx < n; ++idx) {
    size_t i1 = idx / (dims[1] * dims[2]), i2 = (idx / dims[2]) % dims[1], i3 = idx % dims[2];
    int nd[3] = {i1 > dims[0] / 2 ? (int)i1 - (int)dims[0] : (int)i1, i2 > dims[1] / 2 ? (int)i2 - (int)dims[1] : (int)i2,
                 i3 > dims[2] / 2 ? (int)i3 - (int)dims[2] : (int)i3};
    cplx acc(0.0, 0.0);
    for (auto& op : sym.ops) {
      // source Miller = R^T · n_dst
      int ns[3];
      for (int a = 0; a < 3; ++a) ns[a] = op.rot[0][a] * nd[0] + op.rot[1][a] * nd[1] + op.rot[2][a] * nd[2];
      size_t src = miller_to_idx(dims, ns[0], ns[1], ns[2]);
      double arg = TAU * ((double)nd[0] * op.tau[0] + (double)nd[1] * op.tau[1] + (double)nd[2] * op.tau[2]);
      cplx phase(std::cos(arg), -std::sin(arg));
      acc += phase * rho_g[src];
    }
    out[idx] = acc * inv;
  }
  fft3d(out, dims, +1);
  double norm = 1.0 / (double)n;
  for (size_t i = 0; i < n; ++i) rho_r[i] = out[i].real() * norm;
}

}  // namespace orc
