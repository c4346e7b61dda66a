/*
 * pwk — MI355X (gfx950) Kohn-Sham Hamiltonian kernels behind a plain C ABI.
 *
 * Drop-in boundary for the GPU-backend and Hamiltonian-apply calls of
 * pwdft-rs (paths relative to pwdft/pwdft-core/). Every entry point returns
 * an int status (PWK_OK = 0); the message of the last failure on the calling
 * thread is available from pwk_last_error(). No C++ exceptions cross the ABI.
 *
 * Conventions
 *  - Energies eV, lengths Å, densities e/Å³ (the reference's internal units).
 *  - Complex arrays are interleaved (re, im) pairs.
 *  - Grids are row-major (x slowest): flat index i1*ny*nz + i2*nz + i3
 *    (src/scf/grid.rs:120-147).
 *  - Wavefunctions are stored like the reference's faer Mat<c64> of shape
 *    npw × nbands in column-major order: band-major, each band contiguous.
 *  - Host-pointer entry points (pwk_*) copy in, run on the context's stream
 *    and synchronise before returning. Device-pointer entry points
 *    (pwk_dev_*) take HIP device pointers, enqueue on the context's stream and
 *    return without synchronising.
 *  - Precision: PWK_FP64 runs everything in FP64. PWK_FP32 runs FFTs,
 *    pointwise kernels and the nonlocal GEMMs in FP32 (device arrays are
 *    complex64 / float32). PWK_SPLIT_BF16 is PWK_FP32 with the nonlocal GEMMs
 *    as three-term split-bf16 MFMA products accumulated in FP32.
 *  - A context owns one device, one HIP stream and the device state of one
 *    calculation (grid, basis, k-point data, V_eff). Contexts are independent;
 *    calls on one context must not overlap.
 */
#ifndef PWK_H
#define PWK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pwk_ctx pwk_ctx;

enum pwk_status {
  PWK_OK = 0,
  PWK_E_INVALID = 1,     /* bad argument or shape */
  PWK_E_NO_DEVICE = 2,   /* no HIP device / bad ordinal */
  PWK_E_DEVICE = 3,      /* HIP runtime error */
  PWK_E_OOM = 4,         /* device allocation failed */
  PWK_E_UNSUPPORTED = 5, /* grid/basis shape the kernels do not handle */
  PWK_E_STATE = 6        /* call order violated (e.g. H·ψ before pwk_set_veff) */
};

enum pwk_precision { PWK_FP64 = 0, PWK_FP32 = 1, PWK_SPLIT_BF16 = 2 };

/* Message of the last failed call on this thread ("" if none). */
const char* pwk_last_error(void);
/* ABI version (major*100 + minor). */
int pwk_version(void);

/* ---------------------------------------------------------------------------
 * Context. Replaces GpuAccelerator::try_new (src/gpu/mod.rs:148-183): returns
 * PWK_E_NO_DEVICE where the reference returns None. There is no CPU fallback.
 * ------------------------------------------------------------------------- */
int pwk_ctx_create(int device, pwk_ctx** out);
int pwk_ctx_destroy(pwk_ctx* ctx);
int pwk_ctx_sync(pwk_ctx* ctx);
/* The context's hipStream_t, for callers that enqueue their own work. */
void* pwk_ctx_stream(pwk_ctx* ctx);
/* No counterpart in the reference. 1 (the default): H·psi on cubic grids of at
 * most 24 points per side runs as one fused launch per call (whole grid in LDS,
 * projectors included), and each Chebyshev filter pass of pwk_scf_run runs as one
 * launch for all bands and all steps of the recurrence; 2: the fused H·psi with
 * one launch per filter step; 0: the z-stick kernels and projector GEMMs used for
 * every other grid. All compute the same products; the switch exists for
 * comparisons. */
int pwk_ctx_set_fused_hpsi(pwk_ctx* ctx, int on);
/* No counterpart in the reference. On, pwk_scf_run captures every SCF iteration
 * after the first of a device-resident solve (below) as one HIP graph per filter
 * precision and replays it, and on the host-synchronised path each filter pass of
 * the Chebyshev-filtered solver (fused small-grid path) per spin channel and filter
 * precision; off (the default) launches the kernels one by one. Both run the same
 * kernels in the same order and give identical results. Capture must not overlap
 * allocations by other threads on the device (the serving workers never capture). */
int pwk_ctx_set_graphs(pwk_ctx* ctx, int on);
/* No counterpart in the reference. On (the default), a filtered solve on a small
 * cubic grid (one-launch filter, fused Rayleigh-Ritz and density) keeps every
 * scalar of an iteration on the device: the filter passes stop through a device
 * residual gate, the Fermi level, occupations, band energy, -TS, the density
 * normalisation and the Anderson/DIIS solve run as kernels, and the host
 * synchronises once per iteration to read the energies and test convergence. Off:
 * the host takes those steps between launches (the path every other solve uses).
 * Both follow the reference's algorithm; results agree to rounding. */
int pwk_ctx_set_resident(pwk_ctx* ctx, int on);

/* Full-grid FFT plans (grids that are not small and cubic). No counterpart in
 * the reference (rustfft plans its transforms on the CPU, src/fft.rs:19-114).
 * A plan is (ord: radix order of the line transforms, 0 = powers of four
 * first, 1 = odd radices first; tpb: lanes per workgroup of the line passes,
 * 64/128/256; zlds: 1 stages the contiguous pass through LDS). With autotune
 * on (the default), the first transform of a (grid, precision, batch) in the
 * process times every candidate plan on the context's stream and keeps the
 * fastest for the whole process; tpb and zlds change no arithmetic, ord
 * changes the rounding only. pwk_ctx_set_fft_plan forces one plan on this
 * context (ord < 0 clears it). pwk_fft_plan_query returns the plan the
 * context would use for its grid (tuning it now if needed) with the tuned and
 * the default plan's microseconds per transform and the candidate count
 * (0 when no tuning ran: forced plan, autotune off or stream capture). */
int pwk_ctx_set_fft_autotune(pwk_ctx* ctx, int on);
int pwk_ctx_set_fft_plan(pwk_ctx* ctx, int ord, int tpb, int zlds);
int pwk_fft_plan_query(pwk_ctx* ctx, int batch, int precision, int plan[3], double* us_tuned, double* us_default,
                       int* candidates);

/* ---------------------------------------------------------------------------
 * Kernel timing (no counterpart in the reference, which has no GPU timers):
 * when enabled, every kernel the context launches is bracketed with HIP
 * events on its stream; pwk_profile_read synchronises and returns the total
 * device time and launch count per kernel name.
 * ------------------------------------------------------------------------- */
int pwk_profile_enable(pwk_ctx* ctx, int on);
int pwk_profile_reset(pwk_ctx* ctx);
/* Fills up to cap entries: names (cap × 64 chars, NUL-terminated), total_ms,
 * launches. Returns the number of kernel names recorded (may exceed cap). */
int pwk_profile_read(pwk_ctx* ctx, int cap, char* names, double* total_ms, int64_t* launches);

/* ---------------------------------------------------------------------------
 * Grid kernels. Mirrors GpuAccelerator (src/gpu/mod.rs).
 * ------------------------------------------------------------------------- */
/* prepare_buffers (src/gpu/mod.rs:197-364): FFT dims and |G|² per grid point
 * (length N = nx*ny*nz, FP64, grid order) uploaded once. */
int pwk_prepare(pwk_ctx* ctx, const int dims[3], const double* g_squared);

/* hartree_potential (src/gpu/mod.rs:398-494; semantic src/scf/energy.rs:343-358):
 * v_h[i] = fourpi_e2 * rho_g[i] / |G_i|², 0 where |G|² <= 1e-12. Complex N. */
int pwk_hartree(pwk_ctx* ctx, const double* rho_g, double fourpi_e2, double* v_h, int precision);

/* v_eff_assembly (src/gpu/mod.rs:508-610; semantic src/scf/energy.rs:293-301):
 * v_eff = v_local + v_h + v_xc, complex N. */
int pwk_veff(pwk_ctx* ctx, const double* v_local, const double* v_h, const double* v_xc, double* v_eff,
             int precision);

/* Fused Hartree + V_eff (K1+K3): v_eff = v_local + 4πe²ρ(G)/|G|² + v_xc in one pass. */
int pwk_hartree_veff(pwk_ctx* ctx, const double* rho_g, double fourpi_e2, const double* v_local,
                     const double* v_xc, double* v_eff, int precision);

/* lda_xc (src/gpu/mod.rs:627-725; semantic src/potential/xc.rs:129-180):
 * PZ-LDA ε_xc and v_xc (eV) from ρ (e/Å³), real N. */
int pwk_lda_xc(pwk_ctx* ctx, const double* rho, double* exc, double* vxc, int precision);

/* lda_xc_spin_grid (src/potential/xc.rs:427-440): LSDA (ε_xc, v↑, v↓). */
int pwk_lsda_xc(pwk_ctx* ctx, const double* rho_up, const double* rho_dn, double* exc, double* v_up,
                double* v_dn, int precision);

/* Device-pointer forms (complex arrays are double2/float2 per precision). */
int pwk_dev_hartree_veff(pwk_ctx* ctx, const void* rho_g, double fourpi_e2, const void* v_local,
                         const void* v_xc, void* v_eff, int precision);
int pwk_dev_lda_xc(pwk_ctx* ctx, const void* rho, void* exc, void* vxc, int precision);
int pwk_dev_lsda_xc(pwk_ctx* ctx, const void* rho_up, const void* rho_dn, void* exc, void* v_up, void* v_dn,
                    int precision);

/* ---------------------------------------------------------------------------
 * 3D FFT with the reference's convention (src/fft.rs:19-114): unnormalised
 * both ways, forward e^{-iG·r}, inverse e^{+iG·r}. sign = -1 forward, +1
 * inverse. In place on N complex values (requires pwk_prepare).
 * ------------------------------------------------------------------------- */
int pwk_fft3d(pwk_ctx* ctx, double* data, int sign, int precision);
int pwk_dev_fft3d(pwk_ctx* ctx, void* data, int batch, int sign, int precision);

/* ---------------------------------------------------------------------------
 * Hamiltonian apply
 * ------------------------------------------------------------------------- */
/* Plane-wave basis (src/basis.rs:36-63): Miller indices in the reference's
 * enumeration order (n1 -> n2 -> n3 ascending), npw*3 int32. Builds the
 * G -> FFT map (src/scf/grid.rs:100-106, 142-147) and the z-stick tables on
 * the device. Requires pwk_prepare. */
int pwk_set_basis(pwk_ctx* ctx, int npw, const int32_t* miller);

/* Number of k-points whose kinetic diagonal / projectors are held. */
int pwk_set_nk(pwk_ctx* ctx, int nk);

/* Kinetic diagonal of k-point ik: t[i] = ħ²/2m |k+G_i|² (src/scf/potentials.rs:178). */
int pwk_set_kinetic(pwk_ctx* ctx, int ik, const double* t);

/* V_eff(G) (complex N, the reference's v_eff_fft) for the following H·ψ and
 * dense-H calls. The device keeps V_eff(G) and V(r) = IFFT_unnorm(V_eff(G)). */
int pwk_set_veff(pwk_ctx* ctx, const double* v_eff_g);
int pwk_dev_set_veff(pwk_ctx* ctx, const void* v_eff_g /* double2, device */);

/* Kleinman-Bylander projectors of k-point ik (src/potential/nonlocal.rs:110-126):
 * B (npw × nch, column-major complex) and D·B^H (nch × npw, column-major
 * complex), so that H_NL = B·(D·B^H) (nonlocal.rs:432-450). nch = 0 clears. */
int pwk_set_nonlocal(pwk_ctx* ctx, int ik, int nch, const double* B, const double* DBH);

/* Builds B and D·B^H for k-point ik on the device from radial projector
 * tables, as NonlocalPotential::new (src/potential/nonlocal.rs:190-382).
 *   kvec[3]: k (1/Å); gvec: Cartesian G per basis vector (npw*3, 1/Å);
 *   natoms, tau (natoms*3 Cartesian Å), atom_species (natoms);
 *   per species s: nproj[s], proj_l (concatenated), dij (concatenated nproj²
 *   row-major, eV), mesh[s], r, rab, chi (concatenated mesh-length arrays,
 *   χ = r·β in Å^{-1/2}, projector-major per species);
 *   omega (Å³). */
int pwk_build_nonlocal(pwk_ctx* ctx, int ik, const double* kvec, const double* gvec, int natoms,
                       const double* tau, const int32_t* atom_species, int nspecies, const int32_t* nproj,
                       const int32_t* proj_l, const double* dij, const int32_t* mesh, const double* r,
                       const double* rab, const double* chi, double omega);
/* Number of nonlocal channels held for k-point ik. */
int pwk_nonlocal_channels(pwk_ctx* ctx, int ik);

/* H·ψ for nb bands of k-point ik (semantic: dense H of
 * fill_hamiltonian_with_v_eff + add_to_hamiltonian applied to ψ):
 *   Hψ = T⊙ψ + (1/N)·gather FFT_f[V(r) ⊙ IFFT(scatter ψ)] + B·(D·B^H·ψ).
 * Host form: psi/hpsi complex128 npw×nb. */
int pwk_hpsi(pwk_ctx* ctx, int ik, int nb, const double* psi, double* hpsi, int precision);
/* Device form: psi/hpsi are double2 (PWK_FP64) or float2 (PWK_FP32,
 * PWK_SPLIT_BF16) device arrays of npw×nb. */
int pwk_dev_hpsi(pwk_ctx* ctx, int ik, int nb, const void* psi, void* hpsi, int precision);

/* No counterpart in the reference (its projector products are faer matmuls inside
 * NonlocalPotential, src/potential/nonlocal.rs:432-450): the complex GEMM on the
 * matrix cores that libpwk uses for the projectors and the subspace products,
 * C (+)= A·B column-major on device pointers, batch products with element strides
 * between them. complex128 for PWK_FP64, complex64 otherwise (PWK_SPLIT_BF16
 * computes with split-bf16 products). Enqueued on the context's stream. */
int pwk_dev_zgemm(pwk_ctx* ctx, int m, int n, int k, const void* a, int lda, long long stride_a, const void* b,
                  int ldb, long long stride_b, void* c, int ldc, long long stride_c, int batch, int accumulate,
                  int precision);

/* Dense Hamiltonian of k-point ik (src/scf/potentials.rs:163-190, then
 * nonlocal.rs:432-450 when with_nl): h is npw×npw column-major complex128. */
int pwk_dense_h(pwk_ctx* ctx, int ik, int with_nl, double* h);
int pwk_dev_dense_h(pwk_ctx* ctx, int ik, int with_nl, void* h /* double2 npw×npw */);

/* Lowest nb eigenpairs of nk Hermitian n×n matrices, FP64, eigenvalues
 * ascending: diagonalize_lowest (src/eigensolver/dense.rs:87-98) applied to
 * each matrix of a batch (the reference calls it once per k-point from
 * diagonalize_dispatch, src/scf/driver.rs:125-184). h holds the nk matrices
 * back to back, column-major complex128, and is overwritten; w receives nk×nb
 * eigenvalues, v nk column-major n×nb eigenvector blocks (unit norm, phase
 * arbitrary as in the reference). Returns PWK_E_UNSUPPORTED above the size the
 * one-workgroup-per-matrix solver holds in LDS (n about 2500). */
int pwk_dense_eigh(pwk_ctx* ctx, int n, int nb, int nk, double* h, double* w, double* v);
int pwk_dev_dense_eigh(pwk_ctx* ctx, int n, int nb, int nk, void* h, void* w, void* v);
/* No counterpart in the reference: pwk_dev_dense_eigh for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same arguments, kernels and results; the
 * eigensolver's workspace is a static 16 MiB array per GPU instead of memory
 * the context allocates, so the call allocates nothing. Batches needing more
 * than 2^21 workspace doubles (nk*(4n + 6n*nb)) return PWK_E_INVALID. All calls
 * on one GPU share that workspace: calls that overlap on one GPU (two streams,
 * two threads) race; serialise them. */
int pwk_dev_dense_eigh_noalloc(pwk_ctx* ctx, int n, int nb, int nk, void* h, void* w, void* v);

/* Density accumulation (src/scf/density.rs:59-97): for each band b with
 * fw[b] >= 1e-15, rho[r] += fw[b]·|IFFT_unnorm(scatter ψ_b)(r)|². rho is real
 * N (FP64 accumulator, host or device), psi as in pwk_hpsi. */
int pwk_density_accumulate(pwk_ctx* ctx, int nb, const double* psi, const double* fw, double* rho,
                           int precision);
int pwk_dev_density_accumulate(pwk_ctx* ctx, int nb, const void* psi, const double* fw_host, void* rho,
                               int precision);

/* ---- k × band batches ----------------------------------------------------------------
 * The reference applies its Hamiltonian and accumulates the density k-point by
 * k-point in a rayon loop (src/scf/driver.rs:461-472; compute_density takes one
 * eigenvector matrix per k-point, src/scf/density.rs:41-46). A batch names, per
 * entry, a k-point slot (pwk_set_nk) and a band count; the entries' bands are
 * packed back to back (entry 0's bands first), each band npw coefficients as in
 * pwk_hpsi. Band counts may differ between entries; npw is the context's basis,
 * which the reference shares between all k-points (src/basis.rs:54). */
typedef struct pwk_kbatch {
  int nentries;
  const int32_t* ik;      /* k-point slot of each entry */
  const int32_t* nbands;  /* bands of each entry (>= 1) */
} pwk_kbatch;

/* H·ψ of every band of a batch (semantic of pwk_hpsi per entry). On cubic grids of
 * at most 24 points per side all bands go through one fused launch that reads each
 * band's k-point from a table (the k-points must share one projector channel
 * count); elsewhere one H·ψ per entry. */
int pwk_hpsi_batch(pwk_ctx* ctx, const pwk_kbatch* batch, const double* psi, double* hpsi, int precision);
int pwk_dev_hpsi_batch(pwk_ctx* ctx, const pwk_kbatch* batch, const void* psi, void* hpsi, int precision);
/* compute_density's accumulation over a batch (src/scf/density.rs:59-97): rho +=
 * Σ_b fw_b |IFFT_unnorm(scatter ψ_b)|² with fw = occupation × k-point weight per band
 * (bands with fw < 1e-15 skipped), all bands in one density pass. The rescaling to
 * the electron count (density.rs:99-108) stays with the caller, as in
 * pwk_density_accumulate. */
int pwk_density_accumulate_batch(pwk_ctx* ctx, const pwk_kbatch* batch, const double* psi, const double* fw,
                                 double* rho, int precision);
int pwk_dev_density_accumulate_batch(pwk_ctx* ctx, const pwk_kbatch* batch, const void* psi, const double* fw_host,
                                     void* rho, int precision);

/* diagonalize_subspace (src/eigensolver/dense.rs:183-260), the warm-start
 * Rayleigh-Ritz of diagonalize_dispatch (src/scf/driver.rs:125-184), on nk dense
 * Hamiltonians at once, FP64: per k, HV = H·V_prev, H_sub = V_prev^H·HV (lower
 * triangle read), H_sub = U·diag(λ)·U^H, V = V_prev·U; when the largest residual
 * ||H v_i − λ_i v_i||₂ of a k-point exceeds tol (<= 0: WFRX_RESIDUAL_TOL = 1e-6,
 * dense.rs:37) that k-point gets the full diagonalize_lowest instead
 * (dense.rs:250-254), and with v_prev NULL every k-point does (dense.rs:205-207).
 *   h: nk column-major n×n complex128 matrices (not modified); v_prev, v: nk
 *   column-major n×nb blocks (must not overlap); w: nk×nb eigenvalues ascending;
 *   used_full (nk ints: 1 where the full solve ran) and max_residual (nk doubles:
 *   the gate's value, 0 without v_prev) may be NULL.
 * The device form synchronises the context's stream (the gate is read on the host,
 * as the reference decides it per call); used_full and max_residual are host
 * arrays in both forms. */
int pwk_subspace_rr(pwk_ctx* ctx, int n, int nb, int nk, const double* h, const double* v_prev, double tol, double* w,
                    double* v, int* used_full, double* max_residual);
int pwk_dev_subspace_rr(pwk_ctx* ctx, int n, int nb, int nk, const void* h, const void* v_prev, double tol, void* w,
                        void* v, int* used_full, double* max_residual);

/* ---- Integer tables (host; bit-exact with the reference) -------------------------- */

/* G-sphere basis |G|² ≤ ecut/(ħ²/2m), Miller indices in n1 → n2 → n3 loop
 * order (BasisSet::new, src/basis.rs:36-63). lattice: rows a, b, c in Å.
 * Writes min(npw, cap) triplets; *npw receives the full count. */
int pwk_basis(const double lattice[9], double ecut_ev, int cap, int32_t* miller, int* npw);

/* FFT grid of a basis: 2,3,5-smooth ≥ 2·ceil(√ratio)·n_max + 1 per axis
 * (FftGrid::new, src/scf/grid.rs:44-86; fft_grid_size, src/fft.rs:257-279). */
int pwk_fft_dims(const int32_t* miller, int npw, int ecutrho_ratio, int dims[3]);

/* Space-group operations (find_symmetry_operations, src/symmetry/detect.rs:22-178):
 * rot row-major 3×3 integer matrices on fractional coordinates, tau in [0,1).
 * z: atomic number per atom. Writes min(nops, cap) operations. */
int pwk_symmetry(const double lattice[9], int natoms, const double* frac, const int32_t* z, double symprec,
                 int cap, int32_t* rot, double* tau, int* nops);

/* Monkhorst-Pack mesh reduced to the irreducible wedge (monkhorst_pack,
 * src/kpoints.rs:245-288; reduce_kpoints, src/symmetry/kpoints.rs:31-143).
 * use_symmetry = 0 keeps the identity only (no time reversal); 1 detects the
 * operations and adds time reversal. kcart: Cartesian 1/Å, weights sum to 1. */
int pwk_kpoints(const double lattice[9], int natoms, const double* frac, const int32_t* z, const int kgrid[3],
                const int kshift[3], int use_symmetry, double symprec, int cap, double* kcart, double* weight,
                int* nk);

/* ---- Self-consistent field solve on the device ------------------------------------ */

/* One pseudopotential species on its radial mesh, in the units the reference's
 * UPF reader produces (src/pseudopotential/upf/parse.rs:20-140): r, rab in Å,
 * v_local in eV, beta as χ = r·β in Å^-1/2 (nproj × mesh, row-major), dij in
 * eV (nproj × nproj, row-major), rho_atom = 4πr²ρ in e/Å (NULL if absent),
 * rho_core in e/Å³ (NULL without a core correction). */
typedef struct pwk_species {
  int atomic_number;
  double z_valence;
  int mesh;
  const double* r;
  const double* rab;
  const double* v_local;
  int nproj;
  const int32_t* proj_l;
  const double* beta;
  const double* dij;
  const double* rho_atom;
  const double* rho_core;
} pwk_species;

#define PWK_MAX_SPECIES 8

/* UPF v2 reader: UpfPseudoPotential::load_from_path (src/pseudopotential/upf/mod.rs:98-109)
 * and parse_upf_body (src/pseudopotential/upf/parse.rs:20-140), host only. Converts to
 * the units above (Bohr -> Å, Ry -> eV, χ to Å^-1/2, 4πr²ρ_atom to e/Å, ρ_core to e/Å³).
 * A file without a <UPF or <PP_HEADER marker, a missing or malformed attribute or
 * block, or a block whose number count differs from the header's is PWK_E_INVALID
 * (message in pwk_last_error). PP_RHOATOM absent reads as zeros; rho_core is set only
 * with the core-correction flag (zeros when its block is absent). */
typedef struct pwk_upf pwk_upf;
int pwk_upf_load(const char* path, pwk_upf** out);
int pwk_upf_parse(const char* text, size_t len, pwk_upf** out);
/* Fills a pwk_species whose arrays point into the parsed file's arrays, valid until
 * pwk_upf_free. */
int pwk_upf_species(const pwk_upf* upf, pwk_species* out);
/* Element symbol (NUL-terminated, at most 3 characters) and l_max of the header. */
int pwk_upf_info(const pwk_upf* upf, char symbol[4], int* l_max);
int pwk_upf_free(pwk_upf* upf);

/* Crystal (src/crystal.rs:83-121): lattice rows a, b, c in Å, fractional
 * positions, a species index per atom into species_data (at most
 * PWK_MAX_SPECIES species). */
typedef struct pwk_crystal {
  double lattice[9];
  int natoms;
  const double* frac;
  const int32_t* species;
  int nspecies;
  const pwk_species* species_data;
} pwk_crystal;

/* Parameters of ScfParams (src/scf/mod.rs:72-402) used by the drivers. A
 * zero-initialised struct with ecut, bands and thresholds filled in is a plain
 * nspin = 1 solve. */
typedef struct pwk_scf_params {
  double ecut_ev;
  int ecutrho_ratio;      /* 4 in the reference */
  int fft_dims[3];        /* all > 0: explicit grid; else derived */
  int kgrid[3];
  int kshift[3];
  int use_symmetry;       /* 0: identity only; 1: detected operations + time reversal,
                             k-points reduced to the IBZ; 2: the full MP grid with the
                             density symmetrised by the detected operations */
  double symprec;         /* 1e-5 */
  int n_bands;
  int max_iter;
  double conv_threshold;  /* RMS Δρ in e/Å³ (density_diff, src/scf/energy.rs:227-235) */
  double energy_threshold;/* |ΔE_total| in eV */
  double mixing_beta;
  int mixing_ndim;
  double smearing_sigma;  /* Fermi-Dirac width, eV */
  double sad_sigma;       /* width of the Gaussian start density without PP_RHOATOM, Å */
  int precision;          /* PWK_FP64; PWK_FP32 / PWK_SPLIT_BF16: density FFTs and the
                             Chebyshev filter's H·ψ in that precision, FP64 elsewhere */
  int eigensolver;        /* 0: dense (diagonalize_lowest per k); 1: Chebyshev-filtered
                             subspace iteration on H·ψ with FP64 Rayleigh-Ritz; -1: dense
                             up to 1500 plane waves, filtered above */
  int cheb_degree;        /* filter polynomial degree (0: 14 on cubic grids up to 24³, else 10) */
  int verbose;
  /* spin and mixing (src/scf/mod.rs:90-107) */
  int nspin;              /* 1 (0 reads as 1) or 2: LSDA, src/scf/driver_spin.rs */
  double starting_magnetization[PWK_MAX_SPECIES]; /* per species, in [-1, 1]; the start
                             density splits as rho (1 +- m)/2 with m their mean over atoms */
  int fixed_magnetization;/* nonzero: hold the total moment at tot_magnetization with one
                             Fermi level per spin channel */
  double tot_magnetization;
  int mixing_mode;        /* 0: plain Anderson; 1: Anderson on Kerker-preconditioned residuals
                             (the magnetisation channel of nspin = 2 always mixes plain) */
  double kerker_q_tf;     /* Thomas-Fermi wave vector, 1/Angstrom; <= 0: estimated from the
                             mean valence density (mixing/kerker.rs:50-62) */
} pwk_scf_params;

/* ScfResult + EnergyComponents (src/scf/energy.rs:597-670), eV. */
typedef struct pwk_scf_result {
  int converged, n_iter, nk, n_bands, npw, n_ops;
  int dims[3];
  double e_total, e_harris, fermi, ts, final_delta, integrated_charge;
  double e_band, e_kinetic, e_local, e_nonlocal, e_hartree, e_xc, e_vxc, e_ewald, e_smearing;
  double seconds;
  int nspin;
  double magnetization;   /* |N_up - N_down| (nspin = 2) */
  double n_up, n_down;    /* electrons per spin channel (nspin = 2) */
  int cheb_degree;        /* filter degree used (filtered eigensolver; else 0) */
  int filter_passes;      /* filter passes run over the solve (filtered eigensolver; else 0) */
} pwk_scf_result;

/* run_scf (src/scf/mod.rs:326-400): nspin = 1 follows run_scf_unpolarized
 * (src/scf/driver.rs:353-735), nspin = 2 run_scf_spin (src/scf/driver_spin.rs:37-839).
 * Setup, V_eff, the eigensolve (dense H per k with a batched eigensolver, or
 * Chebyshev-filtered subspace iteration), Fermi-Dirac occupations, density,
 * symmetrisation, energies and Anderson mixing all run on the context's
 * device. eigenvalues (nspin × nk × n_bands, spin-up channel first) and rho
 * (the converged total density on the FFT grid) may be NULL; their sizes come
 * from pwk_scf_tables. Returns PWK_OK also when the solve did not converge
 * (result->converged = 0). */
int pwk_scf_run(pwk_ctx* ctx, const pwk_crystal* crystal, const pwk_scf_params* params, pwk_scf_result* result,
                double* eigenvalues, double* rho);

/* Sizes pwk_scf_run will use: plane waves, FFT grid, irreducible k-points and
 * symmetry operations. Host only. */
int pwk_scf_tables(const pwk_crystal* crystal, const pwk_scf_params* params, int* npw, int dims[3], int* nk,
                   int* nops);

/* ---- Serving: a request loop of SCF solves on one device --------------------------
 * No counterpart in the reference (it runs one solve per process). n_slots
 * worker threads each run one solve at a time in a context of their own;
 * queued requests are admitted in arrival order as slots and device memory
 * (hbm_fraction of the free HBM at creation) allow, so solves of different
 * sizes overlap on the device. The crystal's arrays must stay valid until the
 * request has been waited for. pwk_server_wait blocks until the request has
 * finished and returns its status (PWK_E_OOM when it can never fit); the
 * latency is from submission to completion, in seconds. */
typedef struct pwk_server pwk_server;
int pwk_server_create(int device, int n_slots, double hbm_fraction, pwk_server** out);
int pwk_server_submit(pwk_server* srv, const pwk_crystal* crystal, const pwk_scf_params* params, int64_t* ticket);
int pwk_server_wait(pwk_server* srv, int64_t ticket, pwk_scf_result* result, double* eigenvalues, double* rho,
                    double* latency_s);
int pwk_server_destroy(pwk_server* srv);
/* Cross-request pass batching. Served Chebyshev-filtered solves on small
 * cubic grids that share grid, basis size, projector channels, band counts,
 * filter degree and filter precision run their filter passes (filter, FP64
 * H·psi, Rayleigh-Ritz) together: the first solve to reach a pass waits up to
 * window_us microseconds for others (or until every running solve has
 * arrived, at most 16), then launches one kernel per step over all of them.
 * Served solves run the host-synchronised iteration (one synchronisation per
 * filter pass), in which each member waits for its batch on the host. The reference's k-point loop
 * (src/scf/driver.rs:461-472) is the dimension the batch extends across
 * requests. Results do not depend on the batching. window_us 0 runs every
 * pass alone; the default is 100. pwk_server_pass_stats returns the number
 * of batched launches and of solve passes they carried since creation. */
int pwk_server_set_pass_batching(pwk_server* srv, int window_us);
int pwk_server_pass_stats(pwk_server* srv, int64_t* launches, int64_t* passes);

#ifdef __cplusplus
}
#endif

#endif /* PWK_H */
