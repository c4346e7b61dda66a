// Host-side integer tables of a calculation: lattice algebra, the G-sphere
// basis, the FFT grid rule, Monkhorst-Pack meshes, symmetry detection and the
// irreducible k-point set. These decide integers (basis membership, grid sizes,
// orbit assignment), so every floating-point expression keeps the reference's
// operation order and no multiply-add is contracted.
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "tables.hpp"

namespace pwk {
namespace tab {

static constexpr double kPi = 3.141592653589793;

// src/crystal.rs:108-110
double volume(const Lat& l) { return std::fabs(vdot(vcross(l.v[0], l.v[1]), l.v[2])); }

// src/crystal.rs:112-121: 2π (b×c, c×a, a×b) / (a·(b×c)), signed
Lat reciprocal(const Lat& l) {
  const double triple = vdot(vcross(l.v[0], l.v[1]), l.v[2]);
  const double f = 2.0 * kPi / triple;
  Lat r;
  r.v[0] = vscale(vcross(l.v[1], l.v[2]), f);
  r.v[1] = vscale(vcross(l.v[2], l.v[0]), f);
  r.v[2] = vscale(vcross(l.v[0], l.v[1]), f);
  return r;
}

// src/crystal.rs:77-80
Vec3 frac_to_cart(const Lat& l, const double f[3]) {
  return vadd(vadd(vscale2(f[0], l.v[0]), vscale2(f[1], l.v[1])), vscale2(f[2], l.v[2]));
}

Vec3 miller_vec(const Lat& recip, int n1, int n2, int n3) {
  return vadd(vadd(vscale2((double)n1, recip.v[0]), vscale2((double)n2, recip.v[1])), vscale2((double)n3, recip.v[2]));
}

// src/basis.rs:36-63: |G|² ≤ ecut/(ħ²/2m), loops n1 → n2 → n3 over ±ceil(g_max/|b_i|)
void sphere_basis(const Lat& lat, double ecut_ev, std::vector<int32_t>& miller, std::vector<Vec3>& g) {
  const Lat rc = reciprocal(lat);
  const double gmax2 = ecut_ev / kHbar2Over2m;
  const double gmax = std::sqrt(gmax2);
  int lim[3];
  for (int d = 0; d < 3; ++d) lim[d] = (int)std::ceil(gmax / vnorm(rc.v[d]));
  miller.clear();
  g.clear();
  for (int a = -lim[0]; a <= lim[0]; ++a)
    for (int b = -lim[1]; b <= lim[1]; ++b)
      for (int c = -lim[2]; c <= lim[2]; ++c) {
        const Vec3 v = miller_vec(rc, a, b, c);
        if (vnorm2(v) <= gmax2) {
          miller.push_back(a);
          miller.push_back(b);
          miller.push_back(c);
          g.push_back(v);
        }
      }
}

// src/fft.rs:257-279: smallest 2,3,5-smooth size ≥ 2·n_max+1
static bool smooth235(long n) {
  if (n <= 0) return false;
  for (long p : {2L, 3L, 5L})
    while (n % p == 0) n /= p;
  return n == 1;
}
int grid_size(int n_max) {
  long n = 2L * n_max + 1;
  while (!smooth235(n)) ++n;
  return (int)n;
}

// src/scf/grid.rs:44-86: scale = ceil(sqrt(ratio)), dims = grid_size(scale·n_max)
void grid_dims(const std::vector<int32_t>& miller, int ratio, const int* explicit_dims, int dims[3]) {
  if (explicit_dims && explicit_dims[0] > 0 && explicit_dims[1] > 0 && explicit_dims[2] > 0) {
    for (int d = 0; d < 3; ++d) dims[d] = explicit_dims[d];
    return;
  }
  int nmax[3] = {0, 0, 0};
  for (size_t i = 0; i < miller.size(); ++i) nmax[i % 3] = std::max(nmax[i % 3], std::abs(miller[i]));
  int root = 0;
  while ((root + 1) * (root + 1) <= ratio) ++root;
  const int scale = root * root == ratio ? root : root + 1;
  for (int d = 0; d < 3; ++d) dims[d] = grid_size(scale * nmax[d]);
}

// ---- k-points: src/kpoints.rs:245-288 ----------------------------------------------
static void mp_coord(const int idx[3], const int grid[3], const int shift[3], double f[3]) {
  for (int a = 0; a < 3; ++a) {
    const double raw = (double)idx[a] / (double)grid[a] + (double)shift[a] / (2.0 * (double)grid[a]);
    f[a] = raw - std::floor(raw + 0.5);  // wrapped to (-1/2, 1/2]
  }
}

// ---- symmetry detection: src/symmetry/detect.rs:22-178 -------------------------------
namespace {

struct Metric {
  double m[3][3];
  double of(const int a[3], const int b[3]) const {
    double s = 0.0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) s += (double)a[i] * m[i][j] * (double)b[j];
    return s;
  }
};

double wrapped_dist(const double a[3], const double b[3]) {
  double s = 0.0;
  for (int i = 0; i < 3; ++i) {
    double d = a[i] - b[i];
    d -= std::round(d);
    s += d * d;
  }
  return std::sqrt(s);
}

void to_unit_cell(double f[3]) {
  for (int i = 0; i < 3; ++i) f[i] = f[i] - std::floor(f[i]);
}

void rotate(const int r[9], const double f[3], double out[3]) {
  for (int i = 0; i < 3; ++i) out[i] = (double)r[3 * i] * f[0] + (double)r[3 * i + 1] * f[1] + (double)r[3 * i + 2] * f[2];
}

bool maps_onto_itself(const std::vector<AtomSite>& atoms, const int r[9], const double t[3], double tol) {
  for (const AtomSite& a : atoms) {
    double p[3];
    rotate(r, a.frac, p);
    for (int i = 0; i < 3; ++i) p[i] += t[i];
    to_unit_cell(p);
    bool hit = false;
    for (const AtomSite& o : atoms)
      if (o.z == a.z && wrapped_dist(p, o.frac) < tol) {
        hit = true;
        break;
      }
    if (!hit) return false;
  }
  return true;
}

int det3(const int r[9]) {
  return r[0] * (r[4] * r[8] - r[5] * r[7]) - r[1] * (r[3] * r[8] - r[5] * r[6]) + r[2] * (r[3] * r[7] - r[4] * r[6]);
}

}  // namespace

// Lattice-preserving integer matrices (columns = images of the lattice vectors
// with the same metric), each kept with the first translation, from atom 0 onto
// an atom of the same kind, that maps the crystal onto itself.
std::vector<SymOp> detect_symmetry(const Lat& lat, const std::vector<AtomSite>& atoms, double tol) {
  Metric M;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) M.m[i][j] = vdot(lat.v[i], lat.v[j]);
  std::vector<std::array<int, 3>> cand[3];
  for (int c = 0; c < 3; ++c) {
    const double target = M.m[c][c];
    int range[3];
    for (int i = 0; i < 3; ++i) range[i] = (int)(std::sqrt(target / M.m[i][i]) + 1.5);
    for (int a = -range[0]; a <= range[0]; ++a)
      for (int b = -range[1]; b <= range[1]; ++b)
        for (int d = -range[2]; d <= range[2]; ++d) {
          const int v[3] = {a, b, d};
          if (std::fabs(M.of(v, v) - target) < tol) cand[c].push_back({a, b, d});
        }
  }
  std::vector<SymOp> ops;
  for (const auto& c0 : cand[0])
    for (const auto& c1 : cand[1]) {
      if (std::fabs(M.of(c0.data(), c1.data()) - M.m[0][1]) > tol) continue;
      for (const auto& c2 : cand[2]) {
        const int r[9] = {c0[0], c1[0], c2[0], c0[1], c1[1], c2[1], c0[2], c1[2], c2[2]};
        const int det = det3(r);
        if (det != 1 && det != -1) continue;
        if (std::fabs(M.of(c0.data(), c2.data()) - M.m[0][2]) > tol) continue;
        if (std::fabs(M.of(c1.data(), c2.data()) - M.m[1][2]) > tol) continue;
        SymOp op;
        std::copy(r, r + 9, op.rot);
        if (atoms.empty()) {
          op.tau[0] = op.tau[1] = op.tau[2] = 0.0;
          ops.push_back(op);
          continue;
        }
        double r0[3];
        rotate(r, atoms[0].frac, r0);
        for (const AtomSite& target : atoms) {
          if (target.z != atoms[0].z) continue;
          double t[3] = {target.frac[0] - r0[0], target.frac[1] - r0[1], target.frac[2] - r0[2]};
          to_unit_cell(t);
          if (!maps_onto_itself(atoms, r, t, tol)) continue;
          to_unit_cell(t);
          std::copy(t, t + 3, op.tau);
          ops.push_back(op);
          break;
        }
      }
    }
  return ops;
}

// ---- IBZ reduction: src/symmetry/kpoints.rs:31-143 ------------------------------------
static long grid_index_of(const double f[3], const int grid[3], const int shift[3]) {
  int ind[3];
  for (int j = 0; j < 3; ++j) {
    const double n = (double)grid[j];
    const double w = f[j] - std::floor(f[j]);
    const double x = w * n - 0.5 * (double)shift[j];
    const double xr = std::round(x);
    if (std::fabs(x - xr) > 1e-6) return -1;
    ind[j] = (((int)xr % grid[j]) + grid[j]) % grid[j];
  }
  return ((long)ind[0] * grid[1] + ind[1]) * grid[2] + ind[2];
}

std::vector<KPt> kpoints(const Lat& lat, const int grid[3], const int shift[3], const std::vector<SymOp>& ops,
                         bool time_reversal) {
  const Lat rc = reciprocal(lat);
  const long total = (long)grid[0] * grid[1] * grid[2];
  std::vector<std::array<double, 3>> fr((size_t)total);
  for (int a = 0; a < grid[0]; ++a)
    for (int b = 0; b < grid[1]; ++b)
      for (int c = 0; c < grid[2]; ++c) {
        const int idx[3] = {a, b, c};
        mp_coord(idx, grid, shift, fr[((size_t)a * grid[1] + b) * grid[2] + c].data());
      }
  std::vector<char> seen((size_t)total, 0);
  std::vector<KPt> out;
  for (long i = 0; i < total; ++i) {
    if (seen[(size_t)i]) continue;
    const double* f = fr[(size_t)i].data();
    int orbit = 0;
    for (const SymOp& op : ops) {
      // k rotates with (R^{-1})^T = adj(R)^T · det(R) for det ±1
      const int* r = op.rot;
      const int d = det3(r);
      const int rit[9] = {(r[4] * r[8] - r[5] * r[7]) * d, (r[5] * r[6] - r[3] * r[8]) * d, (r[3] * r[7] - r[4] * r[6]) * d,
                          (r[2] * r[7] - r[1] * r[8]) * d, (r[0] * r[8] - r[2] * r[6]) * d, (r[1] * r[6] - r[0] * r[7]) * d,
                          (r[1] * r[5] - r[2] * r[4]) * d, (r[2] * r[3] - r[0] * r[5]) * d, (r[0] * r[4] - r[1] * r[3]) * d};
      double kr[3];
      rotate(rit, f, kr);
      const long j = grid_index_of(kr, grid, shift);
      if (j >= 0 && !seen[(size_t)j]) {
        seen[(size_t)j] = 1;
        ++orbit;
      }
      if (time_reversal) {
        const double km[3] = {-kr[0], -kr[1], -kr[2]};
        const long jm = grid_index_of(km, grid, shift);
        if (jm >= 0 && !seen[(size_t)jm]) {
          seen[(size_t)jm] = 1;
          ++orbit;
        }
      }
    }
    KPt k;
    k.frac[0] = f[0];
    k.frac[1] = f[1];
    k.frac[2] = f[2];
    k.cart = vadd(vadd(vscale2(f[0], rc.v[0]), vscale2(f[1], rc.v[1])), vscale2(f[2], rc.v[2]));
    k.weight = (double)orbit / (double)total;
    out.push_back(k);
  }
  return out;
}

}  // namespace tab
}  // namespace pwk
