// FP64 CPU oracle for the Kohn-Sham hot path of pwdft-rs.
//
// TEST INFRASTRUCTURE ONLY. This library restates, in plain C++17, the
// reference's CPU algorithms (file:line citations at each function, paths
// relative to pwdft/pwdft-core/). It is used by tests/, by the smoke check
// and by bench.py's cpu_baseline leg as the checker; no product code links
// it or calls it.
//
// Arithmetic that decides an integer outcome (G-sphere membership, FFT grid
// size, symmetry and k-point tables) follows the reference's operation order
// exactly and is compiled with -ffp-contract=off; places where the reference
// uses f64::mul_add use std::fma.
#pragma once

#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <string>
#include <vector>

namespace orc {

using cplx = std::complex<double>;

// ---------------------------------------------------------------------------
// Constants — src/consts.rs:15-49 (RHO_FLOOR from src/functionals/mod.rs:5)
// ---------------------------------------------------------------------------
constexpr double PI = 3.141592653589793;
constexpr double TAU = 6.283185307179586;
constexpr double HA_TO_EV = 27.211386245988;
constexpr double RY_TO_EV = 13.605693122994;
constexpr double BOHR_TO_ANG = 0.529177210903;
constexpr double BOHR3_TO_ANG3 = BOHR_TO_ANG * BOHR_TO_ANG * BOHR_TO_ANG;
constexpr double E2_COULOMB = 14.399645351950548;
constexpr double G2_ZERO_THRESHOLD = 1e-12;
constexpr double G_ZERO_THRESHOLD = 1e-12;
constexpr double H_SI = 6.62607015e-34;
constexpr double HBAR_SI = H_SI / (2.0 * PI);
constexpr double M_E_SI = 9.1093837015e-31;
constexpr double EV_PER_J = 6.241509074e18;
constexpr double ANG_PER_M = 1e10;
constexpr double HBAR2_OVER_2M = HBAR_SI * HBAR_SI / (2.0 * M_E_SI) * EV_PER_J * ANG_PER_M * ANG_PER_M;
constexpr double RHO_FLOOR = 1e-30;

// ---------------------------------------------------------------------------
// 3-vectors with nalgebra's operation order (dot = (x*x + y*y) + z*z)
// ---------------------------------------------------------------------------
struct V3 {
  double x = 0, y = 0, z = 0;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double norm2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline double norm(V3 a) { return std::sqrt(norm2(a)); }
inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// src/crystal.rs:83-121
struct Lattice {
  V3 a, b, c;
  double volume() const;   // |a·(b×c)|  (crystal.rs:108-110)
  Lattice reciprocal() const;  // 2π(b×c)/triple ... signed triple (crystal.rs:112-121)
};

struct Atom {
  int species = 0;   // index into the pseudopotential table
  int z = 0;         // atomic number (symmetry compares species by z)
  double frac[3] = {0, 0, 0};
};

struct Crystal {
  Lattice lat;
  std::vector<Atom> atoms;
  V3 cart(const Atom& at) const;  // f1·a + f2·b + f3·c (crystal.rs:77-80)
};

// ---------------------------------------------------------------------------
// Basis and FFT grid maps — src/basis.rs:36-63, src/fft.rs:257-279,
// src/scf/grid.rs:44-147
// ---------------------------------------------------------------------------
struct Basis {
  std::vector<V3> g;                       // Cartesian G (1/Å)
  std::vector<std::array<int, 3>> miller;  // loop order n1 -> n2 -> n3
  double ecut = 0;                         // eV
  size_t size() const { return g.size(); }
};
Basis make_basis(const Lattice& lat, double ecut_ev);
size_t fft_grid_size(uint32_t n_max);
bool is_fft_friendly(size_t n);
std::array<size_t, 3> fft_dims(const Basis& basis, uint32_t ecutrho_ratio, const size_t* explicit_dims);
size_t miller_to_idx(const std::array<size_t, 3>& dims, int n1, int n2, int n3);
V3 g_vector_at(size_t idx, const std::array<size_t, 3>& dims, const Lattice& recip);
std::vector<size_t> basis_to_fft(const Basis& basis, const std::array<size_t, 3>& dims);

// ---------------------------------------------------------------------------
// 3D FFT (unnormalised both ways, forward e^{-iGr}) — src/fft.rs:19-114.
// Written from the DFT definition as a mixed-radix Cooley–Tukey transform.
// ---------------------------------------------------------------------------
void fft1d(cplx* data, size_t n, size_t stride, int sign);  // sign=-1 forward
void fft3d(std::vector<cplx>& data, const std::array<size_t, 3>& dims, int sign);

// ---------------------------------------------------------------------------
// Pseudopotential — src/pseudopotential/upf/{mod.rs,parse.rs,xml.rs}
// ---------------------------------------------------------------------------
struct Beta {
  int l = 0;
  std::vector<double> values;  // χ = r·β in Å^{-1/2}
};
struct Upf {
  std::string element;
  double z_valence = 0;
  int l_max = 0;
  std::vector<double> r, rab, v_local;  // Å, Å, eV
  std::vector<Beta> betas;
  std::vector<double> dij;       // eV, row-major n_proj²
  std::vector<double> rho_atom;  // 4πr²ρ in e/Å
  std::vector<double> core;      // bare ρ_core in e/Å³ (empty if no NLCC)
  bool has_nlcc() const { return !core.empty(); }
  bool has_rho_atom() const;
  size_t n_proj() const { return betas.size(); }
};
Upf parse_upf(const std::string& path);
double simpson(const double* f, const double* rab, size_t n);       // numerics.rs:29-58
double v_local_of_g(const Upf& pp, double g_norm, double omega);    // upf/mod.rs:166-213
double spherical_bessel_j(int l, double x);                         // nonlocal.rs:592-612
double beta_form_factor(const Upf& pp, int iproj, double q);        // nonlocal.rs:566-583
void real_sph_harmonics(V3 q, int lmax, double* out);               // nonlocal.rs:469-552

// ---------------------------------------------------------------------------
// System: everything ScfContext::new precomputes (src/scf/context.rs:102-231)
// ---------------------------------------------------------------------------
struct NonlocalK {
  size_t npw = 0, nch = 0;
  std::vector<cplx> B;    // npw × nch, column-major
  std::vector<cplx> DBH;  // nch × npw, column-major
  // channel metadata (atom, projector, m-offset, l)
  std::vector<int> ch_atom, ch_proj, ch_m, ch_l;
};

struct System {
  Crystal crystal;
  std::vector<Upf> pps;  // indexed by Atom::species
  Basis basis;
  std::array<size_t, 3> dims{};
  size_t n_grid = 0;
  std::vector<size_t> g_to_fft;
  std::vector<double> g2;         // |G|² per grid point
  std::vector<cplx> v_local;      // V_loc(G), eV
  std::vector<double> rho_core;   // NLCC ρ_core(r) (empty if none)
  double omega = 0;
  double n_electrons = 0;
};
System make_system(const Crystal& crystal, const std::vector<Upf>& pps, double ecut_ev,
                   uint32_t ecutrho_ratio, const size_t* explicit_dims);
std::vector<cplx> compute_v_local(const System& s);     // scf/potentials.rs:18-51
std::vector<double> compute_core_density(const System& s);  // scf/potentials.rs:72-130
NonlocalK build_nonlocal(const System& s, V3 k);        // potential/nonlocal.rs:190-382
double ewald_energy(const System& s);                   // ewald.rs:43-143
std::vector<double> sad_density(const System& s, double gaussian_sigma);  // scf/initial_density.rs:64-166

// ---------------------------------------------------------------------------
// Hamiltonian (dense and FFT route)
// ---------------------------------------------------------------------------
// H[i,j] = V_eff[wrap(m_i - m_j)] + δ_ij ħ²/2m |k+G_i|²  (scf/potentials.rs:163-190)
void fill_dense_h(const System& s, V3 k, const cplx* v_eff_g, cplx* h /*col-major npw²*/);
// H += B·(D·B^H)  (potential/nonlocal.rs:432-450)
void add_nonlocal_dense(const NonlocalK& nl, cplx* h);
// Hψ via FFT: T⊙ψ + (1/N) gather FFT_f[ IFFT(V_eff) ⊙ IFFT(scatter ψ) ] + B·(D·(B^H ψ))
void hpsi_fft(const System& s, V3 k, const NonlocalK* nl, const cplx* v_eff_g, const cplx* psi,
              size_t nb, cplx* hpsi);
std::vector<double> kinetic_diag(const System& s, V3 k);  // hamiltonian.rs:23-34

// ---------------------------------------------------------------------------
// Grid kernels — scf/energy.rs, potential/xc.rs
// ---------------------------------------------------------------------------
void hartree(const cplx* rho_g, const double* g2, size_t n, cplx* v_h);        // energy.rs:343-358
void assemble_veff(const cplx* vl, const cplx* vh, const cplx* vxc, size_t n, cplx* out);  // energy.rs:293-301
void lda_xc(double rho, double* exc, double* vxc);                             // xc.rs:129-141
void lsda_xc(double rho_up, double rho_dn, double* exc, double* vup, double* vdn);  // xc.rs:374-392
void density_r_to_g(const std::vector<double>& rho_r, const std::array<size_t, 3>& dims,
                    std::vector<cplx>& rho_g);                                  // energy.rs:250-259
double hartree_energy(const std::vector<cplx>& rho_g, const std::vector<double>& g2, double omega);
double xc_energy_corrected(const std::vector<double>& rho_xc, const std::vector<double>& rho_val,
                           const std::vector<double>& exc, const std::vector<double>& vxc, double omega);
double density_diff(const std::vector<double>& a, const std::vector<double>& b, double omega);

// ρ(r) += Σ f·|IFFT(scatter ψ)|² for one k (scf/density.rs:59-97); skip f < 1e-15
void accumulate_density(const System& s, const cplx* psi, size_t nb, const double* fw,
                        std::vector<double>& rho_acc);

// ---------------------------------------------------------------------------
// Dense Hermitian eigensolver (textbook: Householder tridiagonalisation +
// implicit QL). Eigenvalues ascending, eigenvectors column-major.
// Semantics of src/eigensolver/dense.rs:50-98.
// ---------------------------------------------------------------------------
void eigh(size_t n, const cplx* a_colmajor, double* w, cplx* v_colmajor);
// Same, all n eigenvalues but only the lowest n_keep eigenvectors (dense.rs:87-98).
void eigh_lowest(size_t n, const cplx* a_colmajor, size_t n_keep, double* w, cplx* v_colmajor);

// ---------------------------------------------------------------------------
// k-points and symmetry — src/kpoints.rs:245-288, src/symmetry/*
// ---------------------------------------------------------------------------
struct KPoint {
  V3 k;
  double weight = 0;
};
struct SymOp {
  int rot[3][3];
  double tau[3];
};
struct Symmetry {
  std::vector<SymOp> ops;
  bool time_reversal = false;
  double tolerance = 1e-5;
};
void mp_fractional_coord(uint32_t i1, uint32_t i2, uint32_t i3, const uint32_t grid[3],
                         const uint32_t shift[3], double out[3]);
std::vector<KPoint> monkhorst_pack(const uint32_t grid[3], const uint32_t shift[3], const Lattice& lat);
Symmetry find_symmetry(const Crystal& c, double tolerance);
Symmetry identity_symmetry();
std::vector<KPoint> reduce_kpoints(const std::vector<KPoint>& full, const uint32_t grid[3],
                                   const uint32_t shift[3], const Symmetry& sym, const Lattice& lat);
void symmetrize_density(std::vector<double>& rho_r, const std::array<size_t, 3>& dims, const Symmetry& sym);

// ---------------------------------------------------------------------------
// Smearing and mixing — src/scf/smearing.rs, src/scf/mixing/anderson.rs
// ---------------------------------------------------------------------------
double fermi_dirac_occ(double e, double ef, double sigma);  // [0,1]
double find_fermi_energy(const std::vector<std::vector<double>>& evs, const std::vector<double>& w,
                         double n_el, double sigma, double spin_factor);
double entropy_ts(const std::vector<std::vector<double>>& evs, const std::vector<double>& w, double ef,
                  double sigma, double spin_factor);
// q_TF² in Å⁻² from the mean valence density (mixing/kerker.rs:50-62)
double kerker_auto_qtf2(double n_electrons, double omega);

// Plain Anderson, or Kerker-preconditioned when kerker_w holds the per-grid-point weights
// |G|²/(|G|² + q_TF²) (0 at G = 0) of mixing/anderson.rs:34-97 and mixing/kerker.rs:31-48
struct Anderson {
  double beta = 0.3;
  size_t max_history = 8;
  std::vector<double> kerker_w;
  std::array<size_t, 3> dims{};
  std::vector<std::vector<double>> hist_in, hist_res;
  std::vector<double> mix(const std::vector<double>& rho_in, const std::vector<double>& rho_out);
};

// ---------------------------------------------------------------------------
// SCF driver, nspin = 1 — src/scf/driver.rs:266-742
// ---------------------------------------------------------------------------
struct ScfParams {
  size_t n_bands = 8;
  size_t max_iter = 100;
  double conv_threshold = 1e-6;
  double energy_threshold = 1e-5;
  double mixing_beta = 0.3;
  size_t mixing_ndim = 8;
  double sigma = 0.01;
  double gaussian_sigma = 1.0;
  int verbose = 0;
  // nspin = 2 (src/scf/driver_spin.rs): the start density splits as ρ(1 ± m)/2 with m the
  // mean of the per-atom starting magnetisations; a fixed total moment uses one Fermi
  // level per channel
  int nspin = 1;
  double start_mag = 0.0;
  bool fixed_mag = false;
  double tot_mag = 0.0;
  // mixing_mode Plain (0) or Kerker (1); q_tf ≤ 0 selects the automatic Thomas-Fermi estimate.
  // The magnetisation channel of nspin = 2 always mixes plain.
  int kerker = 0;
  double q_tf = 0.0;
};
struct EnergyComponents {
  double e_band = 0, e_kinetic = 0, e_local = 0, e_nonlocal = 0, e_hartree = 0, e_xc = 0, e_vxc = 0,
         e_ewald = 0, e_smearing = 0;
};
struct ScfResult {
  bool converged = false;
  size_t n_iter = 0;
  double e_total = 0, e_harris = 0, fermi = 0, ts = 0, final_delta = 0;
  double magnetization = 0, n_up = 0, n_dn = 0;  // nspin = 2
  EnergyComponents comp;
  std::vector<std::vector<double>> eigenvalues;
  std::vector<double> rho_r;
};
ScfResult run_scf(const System& s, const std::vector<KPoint>& kpts, const ScfParams& p, const Symmetry& sym);
// nspin = 2, LSDA with coupled-channel (ρ_total, m) Anderson mixing — src/scf/driver_spin.rs:37-839.
// Eigenvalues come back as [up k0.., down k0..].
ScfResult run_scf_spin(const System& s, const std::vector<KPoint>& kpts, const ScfParams& p, const Symmetry& sym);

}  // namespace orc
