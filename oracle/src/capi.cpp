// C entry points of the oracle library (liboracle.so) for the Python test
// harness (ctypes). Test infrastructure only.
#include <cstring>
#include <exception>
#include <string>

#include "oracle.hpp"

using namespace orc;

namespace {
thread_local std::string g_err;

Lattice lat_from(const double* l9) {
  return {{l9[0], l9[1], l9[2]}, {l9[3], l9[4], l9[5]}, {l9[6], l9[7], l9[8]}};
}

Crystal crystal_from(const double* lat9, int nat, const int* species, const int* z, const double* frac) {
  Crystal c;
  c.lat = lat_from(lat9);
  for (int i = 0; i < nat; ++i) {
    Atom a;
    a.species = species ? species[i] : 0;
    a.z = z[i];
    for (int d = 0; d < 3; ++d) a.frac[d] = frac[3 * i + d];
    c.atoms.push_back(a);
  }
  return c;
}

template <class F>
int guard(F&& f) {
  try {
    return f();
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

inline cplx* as_c(double* p) { return reinterpret_cast<cplx*>(p); }
inline const cplx* as_c(const double* p) { return reinterpret_cast<const cplx*>(p); }
}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

// ---- constants / scalar helpers ---------------------------------------------
double orc_const(const char* name) {
  std::string n(name);
  if (n == "HBAR2_OVER_2M") return HBAR2_OVER_2M;
  if (n == "RY_TO_EV") return RY_TO_EV;
  if (n == "HA_TO_EV") return HA_TO_EV;
  if (n == "BOHR_TO_ANG") return BOHR_TO_ANG;
  if (n == "E2_COULOMB") return E2_COULOMB;
  return 0.0;
}
int orc_fft_grid_size(int n_max) { return (int)fft_grid_size((uint32_t)n_max); }
double orc_bessel(int l, double x) { return spherical_bessel_j(l, x); }
void orc_ylm(const double* q, int lmax, double* out) { real_sph_harmonics({q[0], q[1], q[2]}, lmax, out); }

int orc_basis(const double* lat9, double ecut_ev, int cap, int* miller, double* g) {
  return guard([&] {
    Basis b = make_basis(lat_from(lat9), ecut_ev);
    int n = (int)b.size();
    if (miller && cap >= n)
      for (int i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) miller[3 * i + d] = b.miller[(size_t)i][(size_t)d];
        if (g) {
          g[3 * i] = b.g[(size_t)i].x;
          g[3 * i + 1] = b.g[(size_t)i].y;
          g[3 * i + 2] = b.g[(size_t)i].z;
        }
      }
    return n;
  });
}

void orc_lda_xc(long n, const double* rho, double* exc, double* vxc) {
  for (long i = 0; i < n; ++i) lda_xc(rho[i], &exc[i], &vxc[i]);
}
void orc_lsda_xc(long n, const double* ru, const double* rd, double* exc, double* vu, double* vd) {
  for (long i = 0; i < n; ++i) lsda_xc(ru[i], rd[i], &exc[i], &vu[i], &vd[i]);
}
void orc_hartree(long n, const double* rho_g, const double* g2, double* vh) {
  hartree(as_c(rho_g), g2, (size_t)n, as_c(vh));
}
void orc_veff(long n, const double* vl, const double* vh, const double* vxc, double* out) {
  assemble_veff(as_c(vl), as_c(vh), as_c(vxc), (size_t)n, as_c(out));
}
void orc_fft3d(const int* dims, double* data, int sign) {
  std::array<size_t, 3> d{(size_t)dims[0], (size_t)dims[1], (size_t)dims[2]};
  size_t n = d[0] * d[1] * d[2];
  std::vector<cplx> v(as_c(data), as_c(data) + n);
  fft3d(v, d, sign);
  std::memcpy(data, v.data(), n * sizeof(cplx));
}
int orc_eigh(int n, const double* a, double* w, double* v) {
  return guard([&] {
    eigh((size_t)n, as_c(a), w, as_c(v));
    return 0;
  });
}

// ---- pseudopotential ----------------------------------------------------------
void* orc_upf_load(const char* path) {
  try {
    return new Upf(parse_upf(path));
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}
void orc_upf_free(void* p) { delete static_cast<Upf*>(p); }
// out: z_valence, l_max, mesh, n_proj, has_nlcc
void orc_upf_info(void* p, double* out) {
  Upf* u = static_cast<Upf*>(p);
  out[0] = u->z_valence;
  out[1] = u->l_max;
  out[2] = (double)u->r.size();
  out[3] = (double)u->n_proj();
  out[4] = u->has_nlcc() ? 1.0 : 0.0;
}
int orc_upf_proj_l(void* p, int i) { return static_cast<Upf*>(p)->betas[(size_t)i].l; }
void orc_upf_dij(void* p, double* out) {
  Upf* u = static_cast<Upf*>(p);
  std::memcpy(out, u->dij.data(), u->dij.size() * sizeof(double));
}
// which: 0 r, 1 rab, 2 v_local, 3 rho_atom, 4 core, 10+i beta i
void orc_upf_array(void* p, int which, double* out) {
  Upf* u = static_cast<Upf*>(p);
  const std::vector<double>* v = nullptr;
  switch (which) {
    case 0: v = &u->r; break;
    case 1: v = &u->rab; break;
    case 2: v = &u->v_local; break;
    case 3: v = &u->rho_atom; break;
    case 4: v = &u->core; break;
    default: v = &u->betas[(size_t)(which - 10)].values;
  }
  std::memcpy(out, v->data(), v->size() * sizeof(double));
}
double orc_upf_vloc_g(void* p, double g, double omega) { return v_local_of_g(*static_cast<Upf*>(p), g, omega); }
double orc_upf_beta_ff(void* p, int iproj, double q) { return beta_form_factor(*static_cast<Upf*>(p), iproj, q); }
double orc_simpson(long n, const double* f, const double* rab) { return simpson(f, rab, (size_t)n); }

// ---- system -------------------------------------------------------------------
void* orc_sys_new(const double* lat9, int nat, const int* species, const int* z, const double* frac, int nsp,
                  const char* const* upf_paths, double ecut_ev, int ratio, const int* explicit_dims) {
  try {
    Crystal c = crystal_from(lat9, nat, species, z, frac);
    std::vector<Upf> pps;
    for (int i = 0; i < nsp; ++i) pps.push_back(parse_upf(upf_paths[i]));
    size_t ed[3];
    const size_t* edp = nullptr;
    if (explicit_dims) {
      for (int d = 0; d < 3; ++d) ed[d] = (size_t)explicit_dims[d];
      edp = ed;
    }
    return new System(make_system(c, pps, ecut_ev, (uint32_t)ratio, edp));
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}
void orc_sys_free(void* s) { delete static_cast<System*>(s); }
// out: npw, nx, ny, nz, omega, n_electrons
void orc_sys_info(void* sp, double* out) {
  System* s = static_cast<System*>(sp);
  out[0] = (double)s->basis.size();
  out[1] = (double)s->dims[0];
  out[2] = (double)s->dims[1];
  out[3] = (double)s->dims[2];
  out[4] = s->omega;
  out[5] = s->n_electrons;
}
void orc_sys_miller(void* sp, int* out) {
  System* s = static_cast<System*>(sp);
  for (size_t i = 0; i < s->basis.size(); ++i)
    for (size_t d = 0; d < 3; ++d) out[3 * i + d] = s->basis.miller[i][d];
}
void orc_sys_gvec(void* sp, double* out) {
  System* s = static_cast<System*>(sp);
  for (size_t i = 0; i < s->basis.size(); ++i) {
    out[3 * i] = s->basis.g[i].x;
    out[3 * i + 1] = s->basis.g[i].y;
    out[3 * i + 2] = s->basis.g[i].z;
  }
}
void orc_sys_g_to_fft(void* sp, long* out) {
  System* s = static_cast<System*>(sp);
  for (size_t i = 0; i < s->g_to_fft.size(); ++i) out[i] = (long)s->g_to_fft[i];
}
void orc_sys_g2(void* sp, double* out) {
  System* s = static_cast<System*>(sp);
  std::memcpy(out, s->g2.data(), s->g2.size() * sizeof(double));
}
void orc_sys_vloc(void* sp, double* out) {
  System* s = static_cast<System*>(sp);
  std::memcpy(out, s->v_local.data(), s->v_local.size() * sizeof(cplx));
}
int orc_sys_core(void* sp, double* out) {
  System* s = static_cast<System*>(sp);
  if (s->rho_core.empty()) return 0;
  if (out) std::memcpy(out, s->rho_core.data(), s->rho_core.size() * sizeof(double));
  return 1;
}
double orc_sys_ewald(void* sp) { return ewald_energy(*static_cast<System*>(sp)); }
void orc_sys_sad(void* sp, double sigma, double* out) {
  auto v = sad_density(*static_cast<System*>(sp), sigma);
  std::memcpy(out, v.data(), v.size() * sizeof(double));
}
void orc_sys_kinetic(void* sp, const double* k, double* out) {
  auto t = kinetic_diag(*static_cast<System*>(sp), {k[0], k[1], k[2]});
  std::memcpy(out, t.data(), t.size() * sizeof(double));
}
// Returns n_channels; fills B (npw×nch col-major, complex) and DBH (nch×npw) when non-null,
// plus channel metadata meta[4*nch] = (atom, proj, m_offset, l).
int orc_sys_nonlocal(void* sp, const double* k, double* B, double* DBH, int* meta) {
  System* s = static_cast<System*>(sp);
  NonlocalK nl = build_nonlocal(*s, {k[0], k[1], k[2]});
  if (B) std::memcpy(B, nl.B.data(), nl.B.size() * sizeof(cplx));
  if (DBH) std::memcpy(DBH, nl.DBH.data(), nl.DBH.size() * sizeof(cplx));
  if (meta)
    for (size_t c = 0; c < nl.nch; ++c) {
      meta[4 * c] = nl.ch_atom[c];
      meta[4 * c + 1] = nl.ch_proj[c];
      meta[4 * c + 2] = nl.ch_m[c];
      meta[4 * c + 3] = nl.ch_l[c];
    }
  return (int)nl.nch;
}
void orc_sys_dense_h(void* sp, const double* k, const double* veff, int with_nl, double* h) {
  System* s = static_cast<System*>(sp);
  V3 kv{k[0], k[1], k[2]};
  fill_dense_h(*s, kv, as_c(veff), as_c(h));
  if (with_nl) {
    NonlocalK nl = build_nonlocal(*s, kv);
    add_nonlocal_dense(nl, as_c(h));
  }
}
void orc_sys_hpsi(void* sp, const double* k, const double* veff, int with_nl, int nb, const double* psi,
                  double* hpsi) {
  System* s = static_cast<System*>(sp);
  V3 kv{k[0], k[1], k[2]};
  NonlocalK nl;
  if (with_nl) nl = build_nonlocal(*s, kv);
  hpsi_fft(*s, kv, with_nl ? &nl : nullptr, as_c(veff), as_c(psi), (size_t)nb, as_c(hpsi));
}
// H·ψ with caller-supplied B / DBH (col-major complex), as used for synthetic projector sets.
void orc_hpsi_with_nl(void* sp, const double* k, const double* veff, int nch, const double* B, const double* DBH,
                      int nb, const double* psi, double* hpsi) {
  System* s = static_cast<System*>(sp);
  V3 kv{k[0], k[1], k[2]};
  NonlocalK nl;
  nl.npw = s->basis.size();
  nl.nch = (size_t)nch;
  if (nch > 0) {
    nl.B.assign(as_c(B), as_c(B) + nl.npw * nl.nch);
    nl.DBH.assign(as_c(DBH), as_c(DBH) + nl.npw * nl.nch);
  }
  hpsi_fft(*s, kv, nch > 0 ? &nl : nullptr, as_c(veff), as_c(psi), (size_t)nb, as_c(hpsi));
}
void orc_sys_density(void* sp, int nb, const double* psi, const double* fw, double* rho_acc) {
  System* s = static_cast<System*>(sp);
  std::vector<double> acc(rho_acc, rho_acc + s->n_grid);
  accumulate_density(*s, as_c(psi), (size_t)nb, fw, acc);
  std::memcpy(rho_acc, acc.data(), acc.size() * sizeof(double));
}

// ---- k-points / symmetry ------------------------------------------------------
int orc_symmetry(const double* lat9, int nat, const int* z, const double* frac, double tol, int cap, int* rot,
                 double* tau) {
  Crystal c = crystal_from(lat9, nat, nullptr, z, frac);
  Symmetry s = find_symmetry(c, tol);
  int n = (int)s.ops.size();
  if (rot && cap >= n)
    for (int i = 0; i < n; ++i) {
      std::memcpy(&rot[9 * i], &s.ops[(size_t)i].rot[0][0], 9 * sizeof(int));
      std::memcpy(&tau[3 * i], s.ops[(size_t)i].tau, 3 * sizeof(double));
    }
  return n;
}
int orc_kpoints(const double* lat9, int nat, const int* z, const double* frac, const int* grid, const int* shift,
                int mode /*0 full grid, 1 identity-only reduction, 2 symmetry reduction*/, double tol, double* kout,
                double* wout) {
  Crystal c = crystal_from(lat9, nat, nullptr, z, frac);
  uint32_t g[3] = {(uint32_t)grid[0], (uint32_t)grid[1], (uint32_t)grid[2]};
  uint32_t s[3] = {(uint32_t)shift[0], (uint32_t)shift[1], (uint32_t)shift[2]};
  std::vector<KPoint> full = monkhorst_pack(g, s, c.lat);
  std::vector<KPoint> out = full;
  if (mode == 1) out = reduce_kpoints(full, g, s, identity_symmetry(), c.lat);
  if (mode == 2) out = reduce_kpoints(full, g, s, find_symmetry(c, tol), c.lat);
  for (size_t i = 0; i < out.size(); ++i) {
    if (kout) {
      kout[3 * i] = out[i].k.x;
      kout[3 * i + 1] = out[i].k.y;
      kout[3 * i + 2] = out[i].k.z;
    }
    if (wout) wout[i] = out[i].weight;
  }
  return (int)out.size();
}

// ---- SCF ----------------------------------------------------------------------
// params: n_bands, max_iter, conv, e_thr, beta, ndim, sigma, gaussian_sigma, use_sym, sym_tol
// out[0..]: e_total, e_harris, fermi, ts, n_iter, converged, final_delta, e_band, e_kin, e_loc,
//           e_nl, e_h, e_xc, e_vxc, e_ewald, e_smear
int orc_scf(void* sp, int nk, const double* k, const double* w, const double* params, double* out, double* evals,
            double* rho_out) {
  return guard([&] {
    System* s = static_cast<System*>(sp);
    std::vector<KPoint> kp;
    for (int i = 0; i < nk; ++i) kp.push_back({{k[3 * i], k[3 * i + 1], k[3 * i + 2]}, w[i]});
    ScfParams p;
    p.n_bands = (size_t)params[0];
    p.max_iter = (size_t)params[1];
    p.conv_threshold = params[2];
    p.energy_threshold = params[3];
    p.mixing_beta = params[4];
    p.mixing_ndim = (size_t)params[5];
    p.sigma = params[6];
    p.gaussian_sigma = params[7];
    p.verbose = (int)params[10];
    p.nspin = (int)params[11];
    p.start_mag = params[12];
    p.fixed_mag = params[13] != 0.0;
    p.tot_mag = params[14];
    p.kerker = (int)params[15];
    p.q_tf = params[16];
    Symmetry sym = params[8] != 0.0 ? find_symmetry(s->crystal, params[9]) : identity_symmetry();
    ScfResult r = p.nspin == 2 ? run_scf_spin(*s, kp, p, sym) : run_scf(*s, kp, p, sym);
    const EnergyComponents& c = r.comp;
    double vals[19] = {r.e_total, r.e_harris, r.fermi, r.ts, (double)r.n_iter, r.converged ? 1.0 : 0.0,
                       r.final_delta, c.e_band, c.e_kinetic, c.e_local, c.e_nonlocal, c.e_hartree, c.e_xc,
                       c.e_vxc, c.e_ewald, c.e_smearing, r.magnetization, r.n_up, r.n_dn};
    std::memcpy(out, vals, sizeof vals);
    if (evals && r.converged)
      for (size_t i = 0; i < r.eigenvalues.size(); ++i)
        for (size_t b = 0; b < r.eigenvalues[i].size(); ++b) evals[i * r.eigenvalues[i].size() + b] = r.eigenvalues[i][b];
    if (rho_out && r.converged) std::memcpy(rho_out, r.rho_r.data(), r.rho_r.size() * sizeof(double));
    return 0;
  });
}

}  // extern "C"

#include <omp.h>
extern "C" int orc_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}
