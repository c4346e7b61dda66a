// Internal declarations shared by the pwk translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pwk.h"

namespace pwk {

// Error carried to the ABI boundary and turned into a status code there.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define PWK_HIP(expr)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      throw ::pwk::Error(e_ == hipErrorOutOfMemory ? PWK_E_OOM : PWK_E_DEVICE,                    \
                         std::string(#expr) + ": " + hipGetErrorString(e_));                       \
  } while (0)

#define PWK_REQUIRE(cond, code, msg)                   \
  do {                                                 \
    if (!(cond)) throw ::pwk::Error((code), (msg));    \
  } while (0)

// Free list of device (or pinned host) blocks owned by one worker thread of the request
// loop. Buffers released on that thread go back to the list and are handed out again to
// its later allocations. All of them are used on the thread's one stream, so reuse is
// stream-ordered, and no hipFree (which waits for the whole device) runs between solves.
struct BlockCache {
  // limit: bytes kept on the free list; blocks released beyond it are freed at once
  BlockCache(bool host_pinned, size_t limit) : pinned(host_pinned), limit_(limit) {}
  BlockCache(const BlockCache&) = delete;
  BlockCache& operator=(const BlockCache&) = delete;
  ~BlockCache() { trim(0); }
  // a block of at least `bytes` (rounded up to 256); `bytes` returns the block's size
  void* take(size_t& bytes);
  void give(void* p, size_t bytes);
  void trim(size_t limit);  // frees the largest blocks until at most `limit` bytes stay cached
  size_t cached() const { return cached_; }
  size_t held() const { return held_; }  // bytes allocated through this cache, cached or in use

 private:
  bool pinned;
  size_t limit_;
  std::mutex mu;
  std::multimap<size_t, void*> free_;
  size_t cached_ = 0, held_ = 0;
};
// Caches that DevBuf / PinnedBuf allocations of the calling thread go through (none: plain
// hipMalloc / hipHostMalloc). Set by the request loop's workers.
extern thread_local BlockCache* tl_dev_cache;
extern thread_local BlockCache* tl_host_cache;

// Grow-only device buffer.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  BlockCache* owner = nullptr;  // cache the block returns to, if any
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes), owner(o.owner) {
    o.p = nullptr;
    o.bytes = 0;
    o.owner = nullptr;
  }
  ~DevBuf() { release(); }
  void ensure(size_t b) {
    if (b <= bytes) return;
    release();
    if (tl_dev_cache) {
      size_t got = b;
      p = tl_dev_cache->take(got);
      owner = tl_dev_cache;
      bytes = got;
    } else {
      PWK_HIP(hipMalloc(&p, b));
      bytes = b;
    }
  }
  void release() {
    if (p) {
      if (owner)
        owner->give(p, bytes);
      else
        (void)hipFree(p);
    }
    p = nullptr;
    bytes = 0;
    owner = nullptr;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// Grow-only pinned host buffer (hipHostMalloc), through the thread's cache when one is set.
struct PinnedBuf {
  void* p = nullptr;
  size_t bytes = 0;
  BlockCache* owner = nullptr;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf() { release(); }
  void ensure(size_t b) {
    if (b <= bytes) return;
    release();
    if (tl_host_cache) {
      size_t got = b;
      p = tl_host_cache->take(got);
      owner = tl_host_cache;
      bytes = got;
    } else {
      PWK_HIP(hipHostMalloc(&p, b, hipHostMallocDefault));
      bytes = b;
    }
  }
  void release() {
    if (p) {
      if (owner)
        owner->give(p, bytes);
      else
        (void)hipHostFree(p);
    }
    p = nullptr;
    bytes = 0;
    owner = nullptr;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// Per-kernel device timing with HIP events on the context stream (opt-in).
struct Profiler {
  bool on = false;
  struct Pending {
    const char* name;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, std::pair<double, long>> totals;  // name -> (ms, launches)
  hipEvent_t get() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    // timing-only events without the system-scope fence (cache writeback/invalidate) on record,
    // so short kernels are not inflated by the event itself; plain events where unsupported
    hipEvent_t e;
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) {
      (void)hipGetLastError();
      PWK_HIP(hipEventCreate(&e));
    }
    return e;
  }
  void collect() {  // callers synchronise the stream first
    for (auto& p : pending) {
      PWK_HIP(hipEventSynchronize(p.b));
      float ms = 0;
      PWK_HIP(hipEventElapsedTime(&ms, p.a, p.b));
      auto& t = totals[p.name];
      t.first += ms;
      t.second += 1;
      pool.push_back(p.a);
      pool.push_back(p.b);
    }
    pending.clear();
  }
  ~Profiler() {
    for (auto& p : pending) {
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};

// Per-k-point device data.
struct KData {
  bool have_kin = false;
  DevBuf kin64, kin32;  // ħ²/2m |k+G|², npw
  int nch = 0;
  DevBuf B64, DBH64;    // double2, npw×nch and nch×npw (column-major)
  DevBuf B32, DBH32;    // float2 copies
};

// The seeded perturbation of entry g of start vector j of block k (a hash of the three indices,
// components in [-0.005, 0.005)): the Chebyshev start blocks and the Lanczos start vector.
__device__ inline double2 start_vector_entry(int g, int j, int k) {
  uint32_t h = (uint32_t)g * 0x9E3779B1u ^ ((uint32_t)j * 0x85EBCA77u + 0x165667B1u) ^ ((uint32_t)k * 0xC2B2AE3Du);
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  h *= 0x297A2D39u;
  h ^= h >> 15;
  const double r1 = (double)(h & 0xFFFF) / 65536.0 - 0.5, r2 = (double)(h >> 16) / 65536.0 - 0.5;
  return make_double2(1e-2 * r1, 1e-2 * r2);
}

// ---- filter passes of several concurrent solves in one set of launches ----------------------
// A filter pass (filter, FP64 H·ψ, Rayleigh-Ritz; scf.hip) of every solve that shares the shape
// below can run in one launch per kernel, a grid row per solve. The operands travel by value in
// the kernel arguments, so a batch holds at most MAX_PASS_BATCH solves.
constexpr int MAX_PASS_BATCH = 16;

struct PassShape {
  int fp = 0;      // filter precision (PWK_FP64 / PWK_FP32)
  int n = 0;       // cubic grid edge
  int npw = 0, nch = 0, nbs = 0, nb = 0, degree = 0;
  double inv_n = 0.0;
  bool operator==(const PassShape& o) const {
    return fp == o.fp && n == o.n && npw == o.npw && nch == o.nch && nbs == o.nbs && nb == o.nb &&
           degree == o.degree && inv_n == o.inv_n;
  }
};
// one solve's operands of the small-grid filter or H·ψ launch (in -> out over nbands bands)
struct PassSmallReq {
  const void* in;
  void* out;
  const void* kin;
  const void* vr;
  const int* g2fft;
  const void* B;
  const void* DBH;
  const double* coef;
  const int* skip;
  int nbands, nkv;
};
struct PassSmallMulti {
  PassSmallReq r[MAX_PASS_BATCH];
};
struct PassRrReq {
  const double2* Y;
  const double2* HY;
  double2* V;
  double* lam;
  double* res;
  const int* skip;
  int nk;
};
struct PassRrMulti {
  PassRrReq r[MAX_PASS_BATCH];
};
struct PassReq {
  PassSmallReq filt, hpsi;
  PassRrReq rr;
};
// Groups the passes that concurrent solves reach at about the same time (serve.hip).
struct PassBatcher;

// How the full-grid 3D FFT runs (autotune.hip): radix order of the lines (0/1), lanes per
// workgroup of the line passes (64/128/256), and whether the contiguous pass is staged in LDS.
struct FftPlan {
  int ord = 0, tpb = 64, zlds = 0;
};

// Device state of one calculation.
struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  int cu_count = 0;
  // When set (by the request loop's workers, which own it), sync() sleeps on this event
  // (hipEventBlockingSync) instead of spinning, leaving the host cores to the other workers.
  hipEvent_t sleep_event = nullptr;

  void sync() {
    if (sleep_event) {
      PWK_HIP(hipEventRecord(sleep_event, stream));
      PWK_HIP(hipEventSynchronize(sleep_event));
    } else {
      PWK_HIP(hipStreamSynchronize(stream));
    }
  }

  // FFT grid (pwk_prepare)
  int dims[3] = {0, 0, 0};
  size_t N = 0;
  DevBuf g2_64, g2_32;

  // plane-wave basis (pwk_set_basis)
  int npw = 0;
  DevBuf miller;   // int32 npw*3
  DevBuf g2fft;    // int32 npw: flat grid index of each G
  // z-sticks: one per distinct (n1, n2); basis vectors of a stick are
  // contiguous in the reference's n1 -> n2 -> n3 enumeration.
  int nsticks = 0;
  DevBuf st_pw0;   // int32: first basis index of the stick
  DevBuf st_cnt;   // int32: number of basis vectors in the stick
  DevBuf st_n3lo;  // int32: n3 of the first basis vector
  DevBuf st_xy;    // int32: i1*ny + i2 (wrapped grid column)
  int nx1 = 0;     // distinct wrapped i1 values carrying sticks
  bool x1_runs = false;       // those values are the runs [0, x1_ra) and [nx - x1_rb, nx)
  int x1_ra = 0, x1_rb = 0;
  DevBuf x1list;   // int32 nx1
  DevBuf st_x1;    // int32: index into x1list of the stick's i1

  // effective potential (pwk_set_veff)
  bool have_v = false;
  DevBuf veffg64;      // double2 N, V_eff(G)
  DevBuf vr64, vr32;   // real N, V(r) in plane-major order [i3][i1][i2]
  DevBuf vrn64, vrn32; // real N, V(r) in grid order [i1][i2][i3]
  bool small_hpsi = true;  // fused one-launch H·ψ on small cubic grids (hpsi_small.hip)
  bool fused_filter = true;  // with small_hpsi: the whole Chebyshev filter pass in one launch
  bool resident = true;      // device-resident SCF iterations where the filtered solver allows them
  bool graphs = false;     // replay captured filter passes in the SCF (scf.hip), opt-in
  // set by the request loop's workers: filter passes go through this batcher, with the worker's
  // two events marking where the pass may start and when it has completed
  PassBatcher* batcher = nullptr;
  hipEvent_t batch_ready = nullptr, batch_done = nullptr;
  // full-grid FFT plans: tuned per (grid, precision, batch) on first use, or one forced plan
  bool fft_autotune = true;
  bool fft_plan_forced = false;
  FftPlan fft_plan;

  std::vector<KData> kd;
  // the kinetic diagonals and projector blocks of all k-points packed k-major (ragged k × band
  // H·ψ batches on small grids); rebuilt after any change to the per-k data
  bool pack_valid = false;
  int pack_nch = 0;
  DevBuf kin_pack64, kin_pack32, B_pack64, B_pack32, DBH_pack64, DBH_pack32;
  DevBuf band_tab;                  // int32 k-point per band of the last batch
  std::vector<int32_t> band_tab_host;

  // workspaces
  DevBuf work_a, work_b;  // stick buffers / FFT scratch
  DevBuf proj;            // nonlocal projections
  DevBuf gemm_ws;         // split-K partial sums
  DevBuf eig_work;        // dense eigensolver scratch
  DevBuf host_in, host_out;  // staging for host-pointer entry points

  Profiler prof;

  void check_grid() const { PWK_REQUIRE(N > 0, PWK_E_STATE, "pwk_prepare has not been called"); }
  void check_basis() const {
    check_grid();
    PWK_REQUIRE(npw > 0, PWK_E_STATE, "pwk_set_basis has not been called");
  }
  KData& k(int ik) {
    PWK_REQUIRE(ik >= 0 && ik < (int)kd.size(), PWK_E_INVALID, "k-point index out of range (pwk_set_nk)");
    return kd[(size_t)ik];
  }
};

// Thread-local last error (ABI)
void set_last_error(const std::string& m);

// ---- kernel launch helpers (defined in the .hip units) ------------------------
// grid.hip
void launch_hartree_veff(Context& c, const void* rho_g, double fourpi_e2, const void* v_local, const void* v_xc,
                         void* out, int precision, bool with_local, bool with_xc, bool with_hartree);
void launch_veff_sum(Context& c, const void* a, const void* b, const void* d, void* out, int precision);
void launch_lda(Context& c, const void* rho, void* exc, void* vxc, int precision);
// Small cubic grids (N ≤ 20, FP64 grid in LDS): one single-workgroup launch for ρ(G) = FFT₋(ρ/N),
// ρ_xc and its LDA ε_xc, v_xc, and (veff_g non-null) V_eff(G) = V_loc + Hartree + FFT₋(v_xc/N)
// with V(r) into V_eff slot `slot` (what pwk_set_veff's transform and reorder write); and
// out = Re FFT₊(in)·scale. Same arithmetic as the separate launches.
bool grid_small_supported(const Context& c);
// lda_ready: rho_xc, exc and vxc already hold the LDA of rho_r (+ core); it is not evaluated again
void launch_rho_xc_veff_small(Context& c, const double* rho_r, const double* core, double cs, double2* rho_g,
                              double* rho_xc, double* exc, double* vxc, const double2* vloc_g, double fourpi_e2,
                              double2* veff_g, int slot, bool lda_ready = false);
void launch_fft_c2r_small(Context& c, const double2* in, double scale, double* out);
// ρ_xc = max(ρ + cs·ρ_core, 0) (ρ when core is null) into rho_xc, then LDA of it (FP64, N points);
// with g, also g ← g·gs (N complex entries) in the same pass
void launch_lda_core(Context& c, const double* rho, const double* core, double cs, double* rho_xc, double* exc,
                     double* vxc, double2* g = nullptr, double gs = 1.0);
void launch_lsda(Context& c, const void* ru, const void* rd, void* exc, void* vu, void* vd, int precision);
void launch_convert(Context& c, const void* src, void* dst, size_t n, bool to_fp32);  // double<->float arrays
// autotune.hip: the full-grid batched 3D FFT with the context's plan (tuned on first use of a
// grid, precision and batch in the process when the context allows it); the plan it uses, with
// the tuned and the default plan's microseconds and the candidate count when it was tuned
void fft3d_planned(Context& c, void* data, int batch, int sign, int precision);
FftPlan fft_plan_for(Context& c, int batch, int precision, double* us_best = nullptr, double* us_default = nullptr,
                     int* ncand = nullptr);
// fft_long.hip: lines of a long length n (fft_sizes.hpp) through LDS, the line-pass contract of the
// planned transform (line l at (l / cnt_in)·st_out + (l % cnt_in)·st_in, element stride es)
void launch_long_lines(hipStream_t st, void* data, int batch, long bs, int n, int nlines, int cnt_in, long st_out,
                       long st_in, long es, int sign, int precision);
// fft3d.hip
bool fft_size_supported(int n);
void launch_fft3d(Context& c, void* data, int batch, int sign, int precision);
// out = FFT_sign(in), one batch item, FP64 (one launch on small cubic grids)
void launch_fft3d_copy(Context& c, const void* in, void* out, int sign);
// out (double2 N) = FFT_sign(r·scale + 0i): one launch on small cubic grids
void launch_fft3d_r2c(Context& c, const double* r, double scale, void* out, int sign);
// hpsi.hip
void setup_basis(Context& c, int npw, const int32_t* miller_host);
// V_eff(G) on the device -> V(r) for H·ψ; slot 1 holds a second channel (LSDA) for the fused
// small-grid H·ψ and needs reserve_veff_slots(c, 2) first
void reserve_veff_slots(Context& c, int nslots);
// (need_g false: skip the copy of V_eff(G) kept for the dense-H builder)
void set_veff_device(Context& c, const void* veff_g_dev, int slot = 0, bool need_g = true);
void launch_hpsi(Context& c, int ik, int nb, const void* psi, void* hpsi, int precision);
// rho += Σ_b fw_b |ψ_b(r)|² (overwrite: rho = Σ ...)
void launch_density(Context& c, int nb, const void* psi, const double* fw_host, void* rho, int precision,
                    bool overwrite = false);
void launch_hpsi_multi(Context& c, int nk, int nb, const void* psi, void* hpsi, int precision, const void* kin,
                       int nch, const void* B, const void* DBH, void* proj);
// k × band batches with a band count per entry (bands of the entries packed back to back)
void launch_hpsi_batch(Context& c, int nent, const int32_t* ik, const int32_t* nbands, const void* psi, void* hpsi,
                       int precision);
void launch_density_batch(Context& c, int nent, const int32_t* ik, const int32_t* nbands, const void* psi,
                          const double* fw_host, void* rho, int precision);
// hpsi_small.hip: the whole H·ψ of nbands bands (band b at k-point b / nbk) in one launch;
// with coef (α, c, β per k-point, device) the output is α(Hψ − cψ) − βX instead
bool hpsi_small_supported(const Context& c, int precision, int nch);
// nkv > 0: the band blocks of consecutive groups of nkv k-points use V(r) slot 0, 1, ...
// (spin channels of LSDA; kinetic and projector blocks are indexed by block mod nkv)
// band_k (device, nbands entries): k-point block of every band, for ragged k × band batches
void launch_hpsi_small(Context& c, int nbands, int nbk, const void* psi, void* hpsi, int precision, const void* kin,
                       int nch, const void* B, const void* DBH, const double* coef = nullptr,
                       const void* X = nullptr, int nkv = 0, const int* band_k = nullptr,
                       const int* skip = nullptr);
// the whole degree-`degree` Chebyshev filter of nbands bands in one launch on small cubic grids:
// Yout (FP64) = p(H)·V (FP64) with the recurrence in `precision`; coef (device) holds (α, c, β)
// per step and band block, coef[(i·nblocks + kb)·3 + j]
bool cheb_filter_small_supported(const Context& c, int precision, int nch);
void launch_cheb_filter_small(Context& c, int nbands, int nbk, const void* V, void* Yout, int precision,
                              const void* kin, int nch, const void* B, const void* DBH, const double* coef, int degree,
                              int nkv = 0, const int* skip = nullptr);
// the whole Lanczos recurrence for the spectral bound of nke band blocks in one launch on small
// cubic grids (FP64; block kb at k-point kb mod nkv and V(r) slot kb / nkv when nkv > 0):
// α_j → ab[j·nke + kb], β_j → ab[(steps + j)·nke + kb] (device)
bool lanczos_small_supported(const Context& c, int nch);
void launch_lanczos_small(Context& c, int nke, int nkv, int steps, const double* kin, int nch, const void* B,
                          const void* DBH, double* ab);
// ρ += Σ_b fw_b |ψ_b(r)|² on small cubic grids (fw device-readable: device or pinned host
// memory); psi_fp64: ψ is FP64 and, for an FP32 density, rounded to FP32 as it is loaded
bool density_small_supported(const Context& c, int precision);
// wave_sums (device-writable, density_wave_sums(c) entries): also the 64-lane shuffle-tree sum of the
// resulting ρ over each wave of 64 consecutive points, in point order (0 past the grid)
void launch_density_small(Context& c, int nb, const void* psi, const double* fw_dev, void* rho, int precision,
                          bool overwrite = false, bool psi_fp64 = false, double* wave_sums = nullptr);
size_t density_wave_sums(const Context& c);
// dense.hip
void launch_dense_h(Context& c, int ik, int with_nl, void* h);
void launch_dense_h_batched(Context& c, int nk, const double* kin, int nch, const void* B, const void* DBH, void* h);
// gemm_mfma.hip: complex C (+)= A·B, column-major, on the matrix cores (FP32, FP64 or 3-term
// split-bf16)
void launch_zgemm_mfma(Context& c, int M, int N, int K, const void* A, int lda, const void* B, int ldb, void* C,
                       int ldc, bool accumulate, int precision);
// `batch` independent products with element strides sA, sB, sC between them
void launch_zgemm_mfma_batched(Context& c, int M, int N, int K, const void* A, int lda, long long sA, const void* B,
                               int ldb, long long sB, void* C, int ldc, long long sC, int batch, bool accumulate,
                               int precision);
// eig.hip: lowest nb eigenpairs of nk Hermitian n×n matrices (H overwritten)
size_t heev_work_doubles(int n, int nb, int nk);
bool heev_supported(int n, int nb);
void launch_heev_batched(Context& c, int n, int nb, int nk, void* H, void* work, double* w, void* V, int* status);
// the same with the static per-device workspace instead of an allocated one (no allocation);
// init_noalloc_scratch looks up that workspace's device address, once per device, with the
// device current
void init_noalloc_scratch(int device);
size_t noalloc_heev_capacity_doubles();
void launch_heev_noalloc(Context& c, int n, int nb, int nk, void* H, double* w, void* V);
// nonlocal.hip
// Where a build writes its projectors: the k-point's own buffers (KData B64, DBH64, B32, DBH32)
// by default, or caller-owned arrays of npw × nch entries (the SCF's batched blocks).
struct NonlocalDest {
  int nch = 0;  // channel count the arrays are sized for (checked)
  void *b64 = nullptr, *dbh64 = nullptr, *b32 = nullptr, *dbh32 = nullptr;
};
void build_nonlocal_device(Context& c, int ik, const double* kvec, const double* gvec, int natoms, const double* tau,
                           const int32_t* atom_species, int nspecies, const int32_t* nproj, const int32_t* proj_l,
                           const double* dij, const int32_t* mesh, const double* r, const double* rab,
                           const double* chi, double omega, const NonlocalDest* dest = nullptr);
// the same for the nkb k-points ik0 .. ik0 + nkb - 1 (kvecs: nkb × 3) in one launch per kernel;
// dest (required for nkb > 1) holds their blocks back to back, npw × nch entries each
void build_nonlocal_batch(Context& c, int ik0, int nkb, const double* kvecs, const double* gvec, int natoms,
                          const double* tau, const int32_t* atom_species, int nspecies, const int32_t* nproj,
                          const int32_t* proj_l, const double* dij, const int32_t* mesh, const double* r,
                          const double* rab, const double* chi, double omega, const NonlocalDest* dest);

// subspace.hip: Chebyshev filter and Rayleigh-Ritz helpers, batched over blocks
void launch_cheb_step(Context& c, size_t n, size_t per, const void* HY, const void* Y, const void* X,
                      const double* coef_dev, void* out, int precision);
void launch_conj_transpose(Context& c, int rows, int cols, const void* A, int lda, long long sA, void* out, int ldo,
                           long long sO, int batch);
void launch_hermitize(Context& c, int m, void* A, int batch);
void launch_lowdin_scale(Context& c, int m, const double* s, double floor_rel, void* U, int batch);
void launch_block_dot(Context& c, size_t per, int batch, const double* x, const double* y, double* out);
void launch_block_scale(Context& c, size_t per, int batch, const double* s, int mode, double* x);
void launch_start_block(Context& c, int npw, int nb, int nk, const int* lowest_dev, void* V);
void launch_resid_norms(Context& c, int npw, int nbs, int nres, int nk, const void* V, const void* HV,
                        const double* lam, double* out);
// whole Rayleigh-Ritz step per k-point in one workgroup (m <= 32): V = Y·C, Ritz values lam
// (nk × m) and residual norms res (nk × m, first nres columns)
bool rr_small_supported(int m);
// skip (device, may be null): the launch does nothing while *skip is nonzero
void launch_rr_small(Context& c, int npw, int m, int nres, int nk, const void* Y, const void* HY, void* V,
                     double* lam, double* res, double floor_rel, const int* skip = nullptr);
// warm-start Rayleigh-Ritz of nk dense n×n Hamiltonians on n×nb start blocks (Vprev may be
// null: full solve), residual gate tol per block; w (nk×nb) and V (nk blocks of n×nb) on the
// device, used_full / max_resid (nk each, host, may be null)
void subspace_rr(Context& c, int n, int nb, int nk, const void* H, const void* Vprev, double tol, double* w, void* V,
                 int* used_full, double* max_resid);
// several solves' filter passes in one launch per kernel (grid row r = request r; max_* the
// largest band and k-block counts of the requests)
void launch_filter_multi(hipStream_t stream, const PassShape& sh, const PassSmallMulti& m, int n, int max_bands);
void launch_hpsi64_multi(hipStream_t stream, const PassShape& sh, const PassSmallMulti& m, int n, int max_bands);
void launch_rr_multi(hipStream_t stream, const PassShape& sh, const PassRrMulti& m, int n, int max_nk,
                     double floor_rel);
// scf.hip: the whole pass (the three launches above) of n solves on one stream
void launch_pass_batch(hipStream_t stream, const PassShape& sh, const PassReq* const* reqs, int n);
// serve.hip: runs this solve's pass through the batcher (c.batcher set); on return later work
// on c.stream is ordered after the pass, and after c.sync() its results (including those the
// pass writes to host memory) are readable
void pass_batch_join(Context& c, const PassShape& sh, const PassReq& r);
// scf.hip
void scf_run(Context& c, const pwk_crystal& cr, const pwk_scf_params& p, pwk_scf_result& out, double* eig,
             double* rho);
// checks every field the solve reads (pointers, species range, radial meshes, scalar parameters);
// scf_tables and scf_run call it first
void validate_scf_request(const pwk_crystal& cr, const pwk_scf_params& p);
void scf_tables(const pwk_crystal& cr, const pwk_scf_params& p, int* npw, int dims[3], int* nk, int* nops);

// Raises a kernel's dynamic-LDS limit once (the attribute call costs host time that
// would otherwise sit between back-to-back launches).
inline void ensure_lds_limit(const void* fn, size_t bytes) {
  static std::mutex mu;
  static std::map<const void*, size_t> set;
  std::lock_guard<std::mutex> lk(mu);
  size_t& cur = set[fn];
  if (bytes <= cur) return;
  PWK_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  cur = bytes;
}

inline unsigned blocks_for(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

// Brackets a launch with events when the context profiles.
struct TimedScope {
  Context& c;
  const char* name;
  hipEvent_t a = nullptr;
  TimedScope(Context& cc, const char* n) : c(cc), name(n) {
    if (c.prof.on) {
      a = c.prof.get();
      PWK_HIP(hipEventRecord(a, c.stream));
    }
  }
  ~TimedScope() noexcept(false) {
    if (a) {
      hipEvent_t b = c.prof.get();
      PWK_HIP(hipEventRecord(b, c.stream));
      c.prof.pending.push_back({name, a, b});
    }
  }
};

}  // namespace pwk
