// Host-side tables of a calculation (lattice, basis, grid, k-points, symmetry)
// and the physical constants of the reference (src/consts.rs:15-49).
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <vector>

namespace pwk {
namespace tab {

constexpr double kPiC = 3.141592653589793;
constexpr double kTwoPi = 6.283185307179586;
constexpr double kRyToEv = 13.605693122994;
constexpr double kBohrToAng = 0.529177210903;
constexpr double kE2 = 14.399645351950548;  // e²/(4πε0) in eV·Å
constexpr double kG2Zero = 1e-12;
constexpr double kHbarSI = 6.62607015e-34 / (2.0 * kPiC);
constexpr double kHbar2Over2m = kHbarSI * kHbarSI / (2.0 * 9.1093837015e-31) * 6.241509074e18 * 1e10 * 1e10;

// 3-vectors; dot and norm² sum x, y, z left to right.
struct Vec3 {
  double x = 0, y = 0, z = 0;
};
inline Vec3 vadd(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3 vsub(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec3 vscale(Vec3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline Vec3 vscale2(double s, Vec3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double vdot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double vnorm2(Vec3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline double vnorm(Vec3 a) { return std::sqrt(vnorm2(a)); }
inline Vec3 vcross(Vec3 a, Vec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

struct Lat {
  Vec3 v[3];  // a, b, c (Å) or b1, b2, b3 (1/Å)
};

struct AtomSite {
  int z = 0;  // atomic number: symmetry compares atoms by kind
  double frac[3] = {0, 0, 0};
};

struct SymOp {
  int rot[9];     // row-major integer matrix acting on fractional coordinates
  double tau[3];  // fractional translation in [0, 1)
};

struct KPt {
  double frac[3];
  Vec3 cart;  // 1/Å
  double weight;
};

double volume(const Lat& l);
Lat reciprocal(const Lat& l);
Vec3 frac_to_cart(const Lat& l, const double f[3]);
Vec3 miller_vec(const Lat& recip, int n1, int n2, int n3);
void sphere_basis(const Lat& lat, double ecut_ev, std::vector<int32_t>& miller, std::vector<Vec3>& g);
int grid_size(int n_max);
void grid_dims(const std::vector<int32_t>& miller, int ratio, const int* explicit_dims, int dims[3]);
std::vector<SymOp> detect_symmetry(const Lat& lat, const std::vector<AtomSite>& atoms, double tol);
std::vector<KPt> kpoints(const Lat& lat, const int grid[3], const int shift[3], const std::vector<SymOp>& ops,
                         bool time_reversal);

}  // namespace tab
}  // namespace pwk
