// C-ABI entry points (include/pwk.h). Each converts exceptions into a status
// code plus a thread-local message; nothing here falls back to the CPU.
#include <cstring>
#include <memory>

#include <algorithm>

#include "common.hpp"
#include "fft_sizes.hpp"
#include "tables.hpp"

namespace pwk {
namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string& m) { g_last_error = m; }
}  // namespace pwk

struct pwk_ctx {
  pwk::Context c;
};

using pwk::Context;
using pwk::DevBuf;
using pwk::Error;

namespace {

template <class F>
int guarded(F&& f) {
  try {
    f();
    return PWK_OK;
  } catch (const Error& e) {
    pwk::set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    pwk::set_last_error("host allocation failed");
    return PWK_E_OOM;
  } catch (const std::exception& e) {
    pwk::set_last_error(e.what());
    return PWK_E_DEVICE;
  }
}

Context& ctx_of(pwk_ctx* p) {
  PWK_REQUIRE(p != nullptr, PWK_E_INVALID, "null context");
  PWK_HIP(hipSetDevice(p->c.device));
  return p->c;
}

void check_precision(int precision) {
  PWK_REQUIRE(precision == PWK_FP64 || precision == PWK_FP32 || precision == PWK_SPLIT_BF16, PWK_E_INVALID,
              "unknown precision");
}

size_t csize(int precision) { return precision == PWK_FP64 ? 16 : 8; }
size_t rsize(int precision) { return precision == PWK_FP64 ? 8 : 4; }

// Host double array (n doubles) -> device buffer in the requested precision.
void to_device(Context& c, DevBuf& d, const double* h, size_t ndoubles, int precision) {
  if (precision == PWK_FP64) {
    d.ensure(ndoubles * 8);
    PWK_HIP(hipMemcpyAsync(d.p, h, ndoubles * 8, hipMemcpyHostToDevice, c.stream));
  } else {
    c.host_in.ensure(ndoubles * 8);
    PWK_HIP(hipMemcpyAsync(c.host_in.p, h, ndoubles * 8, hipMemcpyHostToDevice, c.stream));
    d.ensure(ndoubles * 4);
    pwk::launch_convert(c, c.host_in.p, d.p, ndoubles, true);
  }
}

// Device buffer (n values in precision) -> host double array.
void to_host(Context& c, double* h, const DevBuf& d, size_t ndoubles, int precision) {
  if (precision == PWK_FP64) {
    PWK_HIP(hipMemcpyAsync(h, d.p, ndoubles * 8, hipMemcpyDeviceToHost, c.stream));
  } else {
    c.host_out.ensure(ndoubles * 8);
    pwk::launch_convert(c, d.p, c.host_out.p, ndoubles, false);
    PWK_HIP(hipMemcpyAsync(h, c.host_out.p, ndoubles * 8, hipMemcpyDeviceToHost, c.stream));
  }
  PWK_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace

extern "C" {

const char* pwk_last_error(void) { return pwk::g_last_error.c_str(); }

int pwk_version(void) { return 101; }

int pwk_ctx_create(int device, pwk_ctx** out) {
  return guarded([&] {
    PWK_REQUIRE(out != nullptr, PWK_E_INVALID, "null output pointer");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    PWK_REQUIRE(e == hipSuccess && n > 0, PWK_E_NO_DEVICE, "no HIP device available");
    PWK_REQUIRE(device >= 0 && device < n, PWK_E_NO_DEVICE, "device ordinal out of range");
    PWK_HIP(hipSetDevice(device));
    auto p = std::make_unique<pwk_ctx>();
    p->c.device = device;
    PWK_HIP(hipStreamCreateWithFlags(&p->c.stream, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    PWK_HIP(hipGetDeviceProperties(&prop, device));
    p->c.cu_count = prop.multiProcessorCount;
    pwk::init_noalloc_scratch(device);
    *out = p.release();
  });
}

int pwk_ctx_destroy(pwk_ctx* ctx) {
  return guarded([&] {
    if (!ctx) return;
    (void)hipSetDevice(ctx->c.device);
    (void)hipStreamSynchronize(ctx->c.stream);
    hipStream_t s = ctx->c.stream;
    delete ctx;
    (void)hipStreamDestroy(s);
  });
}

int pwk_ctx_sync(pwk_ctx* ctx) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    PWK_HIP(hipStreamSynchronize(c.stream));
  });
}

void* pwk_ctx_stream(pwk_ctx* ctx) { return ctx ? (void*)ctx->c.stream : nullptr; }

int pwk_ctx_set_fused_hpsi(pwk_ctx* ctx, int on) {
  return guarded([&] {
    PWK_REQUIRE(on >= 0 && on <= 2, PWK_E_INVALID, "mode must be 0, 1 or 2");
    Context& c = ctx_of(ctx);
    c.small_hpsi = on != 0;
    c.fused_filter = on == 1;
  });
}

int pwk_ctx_set_graphs(pwk_ctx* ctx, int on) {
  return guarded([&] { ctx_of(ctx).graphs = on != 0; });
}

int pwk_ctx_set_resident(pwk_ctx* ctx, int on) {
  return guarded([&] { ctx_of(ctx).resident = on != 0; });
}

int pwk_ctx_set_fft_autotune(pwk_ctx* ctx, int on) {
  return guarded([&] { ctx_of(ctx).fft_autotune = on != 0; });
}

int pwk_ctx_set_fft_plan(pwk_ctx* ctx, int ord, int tpb, int zlds) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    if (ord < 0) {
      c.fft_plan_forced = false;
      c.fft_plan = pwk::FftPlan{};
      return;
    }
    PWK_REQUIRE((ord == 0 || ord == 1) && (tpb == 64 || tpb == 128 || tpb == 256) && (zlds == 0 || zlds == 1),
                PWK_E_INVALID, "bad FFT plan (ord 0/1, tpb 64/128/256, zlds 0/1)");
    c.fft_plan.ord = ord;
    c.fft_plan.tpb = tpb;
    c.fft_plan.zlds = zlds;
    c.fft_plan_forced = true;
  });
}

int pwk_fft_plan_query(pwk_ctx* ctx, int batch, int precision, int plan[3], double* us_tuned, double* us_default,
                       int* candidates) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    check_precision(precision);
    c.check_grid();
    PWK_REQUIRE(batch > 0 && plan, PWK_E_INVALID, "batch must be positive and plan non-null");
    double tb = 0.0, td = 0.0;
    int nc = 0;
    const pwk::FftPlan pl = pwk::fft_plan_for(c, batch, precision, &tb, &td, &nc);
    plan[0] = pl.ord;
    plan[1] = pl.tpb;
    plan[2] = pl.zlds;
    if (us_tuned) *us_tuned = tb;
    if (us_default) *us_default = td;
    if (candidates) *candidates = nc;
  });
}

int pwk_profile_enable(pwk_ctx* ctx, int on) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    c.prof.on = on != 0;
  });
}

int pwk_profile_reset(pwk_ctx* ctx) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    PWK_HIP(hipStreamSynchronize(c.stream));
    c.prof.collect();
    c.prof.totals.clear();
  });
}

int pwk_profile_read(pwk_ctx* ctx, int cap, char* names, double* total_ms, int64_t* launches) {
  int n = 0;
  int st = guarded([&] {
    Context& c = ctx_of(ctx);
    PWK_HIP(hipStreamSynchronize(c.stream));
    c.prof.collect();
    for (auto& kv : c.prof.totals) {
      if (n < cap) {
        if (names) {
          std::strncpy(names + 64 * n, kv.first.c_str(), 63);
          names[64 * n + 63] = 0;
        }
        if (total_ms) total_ms[n] = kv.second.first;
        if (launches) launches[n] = kv.second.second;
      }
      ++n;
    }
  });
  return st == PWK_OK ? n : -st;
}

int pwk_prepare(pwk_ctx* ctx, const int dims[3], const double* g2) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    PWK_REQUIRE(dims && g2, PWK_E_INVALID, "null dims or |G|^2");
    for (int d = 0; d < 3; ++d) PWK_REQUIRE(dims[d] > 0 && dims[d] <= 1024, PWK_E_INVALID, "bad FFT dims");
    c.dims[0] = dims[0];
    c.dims[1] = dims[1];
    c.dims[2] = dims[2];
    c.N = (size_t)dims[0] * dims[1] * dims[2];
    c.g2_64.ensure(c.N * 8);
    c.g2_32.ensure(c.N * 4);
    PWK_HIP(hipMemcpyAsync(c.g2_64.p, g2, c.N * 8, hipMemcpyHostToDevice, c.stream));
    pwk::launch_convert(c, c.g2_64.p, c.g2_32.p, c.N, true);
    c.npw = 0;
    c.have_v = false;
    PWK_HIP(hipStreamSynchronize(c.stream));
  });
}

// ---- grid kernels ----------------------------------------------------------
int pwk_hartree(pwk_ctx* ctx, const double* rho_g, double fourpi_e2, double* v_h, int precision) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    check_precision(precision);
    c.check_grid();
    DevBuf in, out;
    to_device(c, in, rho_g, 2 * c.N, precision);
    out.ensure(2 * c.N * rsize(precision));
    pwk::launch_hartree_veff(c, in.p, fourpi_e2, nullptr, nullptr, out.p, precision, false, false, true);
    to_host(c, v_h, out, 2 * c.N, precision);
  });
}

int pwk_veff(pwk_ctx* ctx, const double* v_local, const double* v_h, const double* v_xc, double* v_eff,
             int precision) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    check_precision(precision);
    c.check_grid();
    DevBuf a, b, d, out;
    to_device(c, a, v_local, 2 * c.N, precision);
    to_device(c, b, v_h, 2 * c.N, precision);
    to_device(c, d, v_xc, 2 * c.N, precision);
    out.ensure(2 * c.N * rsize(precision));
    pwk::launch_veff_sum(c, a.p, b.p, d.p, out.p, precision);
    to_host(c, v_eff, out, 2 * c.N, precision);
  });
}

int pwk_hartree_veff(pwk_ctx* ctx, const double* rho_g, double fourpi_e2, const double* v_local, const double* v_xc,
                     double* v_eff, int precision) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    check_precision(precision);
    c.check_grid();
    DevBuf r, a, d, out;
    to_device(c, r, rho_g, 2 * c.N, precision);
    to_device(c, a, v_local, 2 * c.N, precision);
    to_device(c, d, v_xc, 2 * c.N, precision);
    out.ensure(2 * c.N * rsize(precision));
    pwk::launch_hartree_veff(c, r.p, fourpi_e2, a.p, d.p, out.p, precision, true, true, true);
    to_host(c, v_eff, out, 2 * c.N, precision);
  });
}

int pwk_lda_xc(pwk_ctx* ctx, const double* rho, double* exc, double* vxc, int precision) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    check_precision(precision);
    c.check_grid();
    DevBuf in, e, v;
    to_device(c, in, rho, c.N, precision);
    e.ensure(c.N * rsize(precision));
    v.ensure(c.N * rsize(precision));
    pwk::launch_lda(c, in.p, e.p, v.p, precision);
    to_host(c, exc, e, c.N, precision);
    to_host(c, vxc, v, c.N, precision);
  });
}

int pwk_lsda_xc(pwk_ctx* ctx, const double* rho_up, const double* rho_dn, double* exc, double* v_up, double* v_dn,
                int precision) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    check_precision(precision);
    c.check_grid();
    DevBuf u, d, e, vu, vd;
    to_device(c, u, rho_up, c.N, precision);
    to_device(c, d, rho_dn, c.N, precision);
    e.ensure(c.N * rsize(precision));
    vu.ensure(c.N * rsize(precision));
    vd.ensure(c.N * rsize(precision));
    pwk::launch_lsda(c, u.p, d.p, e.p, vu.p, vd.p, precision);
    to_host(c, exc, e, c.N, precision);
    to_host(c, v_up, vu, c.N, precision);
    to_host(c, v_dn, vd, c.N, precision);
  });
}

int pwk_dev_hartree_veff(pwk_ctx* ctx, const void* rho_g, double fourpi_e2, const void* v_local, const void* v_xc,
                         void* v_eff, int precision) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    check_precision(precision);
    pwk::launch_hartree_veff(c, rho_g, fourpi_e2, v_local, v_xc, v_eff, precision, v_local != nullptr,
                             v_xc != nullptr, rho_g != nullptr);
  });
}

int pwk_dev_lda_xc(pwk_ctx* ctx, const void* rho, void* exc, void* vxc, int precision) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    check_precision(precision);
    pwk::launch_lda(c, rho, exc, vxc, precision);
  });
}

int pwk_dev_lsda_xc(pwk_ctx* ctx, const void* rho_up, const void* rho_dn, void* exc, void* v_up, void* v_dn,
                    int precision) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    check_precision(precision);
    pwk::launch_lsda(c, rho_up, rho_dn, exc, v_up, v_dn, precision);
  });
}

// ---- FFT ------------------------------------------------------------------------
int pwk_fft3d(pwk_ctx* ctx, double* data, int sign, int precision) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    check_precision(precision);
    c.check_grid();
    PWK_REQUIRE(sign == 1 || sign == -1, PWK_E_INVALID, "sign must be +1 or -1");
    DevBuf d;
    to_device(c, d, data, 2 * c.N, precision);
    pwk::launch_fft3d(c, d.p, 1, sign, precision);
    to_host(c, data, d, 2 * c.N, precision);
  });
}

int pwk_dev_fft3d(pwk_ctx* ctx, void* data, int batch, int sign, int precision) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    check_precision(precision);
    PWK_REQUIRE(sign == 1 || sign == -1, PWK_E_INVALID, "sign must be +1 or -1");
    PWK_REQUIRE(batch > 0, PWK_E_INVALID, "batch must be positive");
    pwk::launch_fft3d(c, data, batch, sign, precision);
  });
}

// ---- Hamiltonian apply ---------------------------------------------------------
int pwk_set_basis(pwk_ctx* ctx, int npw, const int32_t* miller) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    pwk::setup_basis(c, npw, miller);
    c.kd.clear();
    c.pack_valid = false;
  });
}

int pwk_set_nk(pwk_ctx* ctx, int nk) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    PWK_REQUIRE(nk >= 0, PWK_E_INVALID, "nk must be non-negative");
    c.kd.resize((size_t)nk);
    c.pack_valid = false;
  });
}

int pwk_set_kinetic(pwk_ctx* ctx, int ik, const double* t) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    c.check_basis();
    pwk::KData& kd = c.k(ik);
    c.pack_valid = false;
    kd.kin64.ensure((size_t)c.npw * 8);
    kd.kin32.ensure((size_t)c.npw * 4);
    PWK_HIP(hipMemcpyAsync(kd.kin64.p, t, (size_t)c.npw * 8, hipMemcpyHostToDevice, c.stream));
    pwk::launch_convert(c, kd.kin64.p, kd.kin32.p, (size_t)c.npw, true);
    PWK_HIP(hipStreamSynchronize(c.stream));
    kd.have_kin = true;
  });
}

int pwk_set_veff(pwk_ctx* ctx, const double* v_eff_g) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    c.check_grid();
    DevBuf d;
    d.ensure(c.N * 16);
    PWK_HIP(hipMemcpyAsync(d.p, v_eff_g, c.N * 16, hipMemcpyHostToDevice, c.stream));
    pwk::set_veff_device(c, d.p);
    PWK_HIP(hipStreamSynchronize(c.stream));
  });
}

int pwk_dev_set_veff(pwk_ctx* ctx, const void* v_eff_g) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    pwk::set_veff_device(c, v_eff_g);
  });
}

int pwk_set_nonlocal(pwk_ctx* ctx, int ik, int nch, const double* B, const double* DBH) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    c.check_basis();
    pwk::KData& kd = c.k(ik);
    PWK_REQUIRE(nch >= 0, PWK_E_INVALID, "nch must be non-negative");
    c.pack_valid = false;
    kd.nch = nch;
    if (nch == 0) return;
    size_t n = (size_t)c.npw * nch;
    kd.B64.ensure(n * 16);
    kd.DBH64.ensure(n * 16);
    kd.B32.ensure(n * 8);
    kd.DBH32.ensure(n * 8);
    PWK_HIP(hipMemcpyAsync(kd.B64.p, B, n * 16, hipMemcpyHostToDevice, c.stream));
    PWK_HIP(hipMemcpyAsync(kd.DBH64.p, DBH, n * 16, hipMemcpyHostToDevice, c.stream));
    pwk::launch_convert(c, kd.B64.p, kd.B32.p, 2 * n, true);
    pwk::launch_convert(c, kd.DBH64.p, kd.DBH32.p, 2 * n, true);
    PWK_HIP(hipStreamSynchronize(c.stream));
  });
}

int pwk_build_nonlocal(pwk_ctx* ctx, int ik, const double* kvec, const double* gvec, int natoms, const double* tau,
                       const int32_t* atom_species, int nspecies, const int32_t* nproj, const int32_t* proj_l,
                       const double* dij, const int32_t* mesh, const double* r, const double* rab, const double* chi,
                       double omega) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    PWK_REQUIRE(kvec && gvec && nproj && mesh, PWK_E_INVALID, "null projector input");
    pwk::build_nonlocal_device(c, ik, kvec, gvec, natoms, tau, atom_species, nspecies, nproj, proj_l, dij, mesh, r,
                               rab, chi, omega);
  });
}

int pwk_nonlocal_channels(pwk_ctx* ctx, int ik) {
  int n = -1;
  int st = guarded([&] {
    Context& c = ctx_of(ctx);
    n = c.k(ik).nch;
  });
  return st == PWK_OK ? n : -st;
}

int pwk_hpsi(pwk_ctx* ctx, int ik, int nb, const double* psi, double* hpsi, int precision) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    check_precision(precision);
    c.check_basis();
    PWK_REQUIRE(nb > 0 && psi && hpsi, PWK_E_INVALID, "bad psi/hpsi");
    size_t nd = (size_t)2 * c.npw * nb;
    DevBuf in, out;
    to_device(c, in, psi, nd, precision);
    out.ensure(nd * rsize(precision));
    pwk::launch_hpsi(c, ik, nb, in.p, out.p, precision);
    to_host(c, hpsi, out, nd, precision);
  });
}

int pwk_dev_hpsi(pwk_ctx* ctx, int ik, int nb, const void* psi, void* hpsi, int precision) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    check_precision(precision);
    pwk::launch_hpsi(c, ik, nb, psi, hpsi, precision);
  });
}

int pwk_dev_zgemm(pwk_ctx* ctx, int m, int n, int k, const void* a, int lda, long long stride_a, const void* b,
                  int ldb, long long stride_b, void* cmat, int ldc, long long stride_c, int batch, int accumulate,
                  int precision) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    check_precision(precision);
    PWK_REQUIRE(m >= 0 && n >= 0 && k >= 0 && batch >= 1, PWK_E_INVALID, "bad GEMM shape");
    PWK_REQUIRE(lda >= std::max(1, m) && ldb >= std::max(1, k) && ldc >= std::max(1, m), PWK_E_INVALID,
                "leading dimension smaller than the matrix");
    PWK_REQUIRE(m == 0 || n == 0 || (a && b && cmat), PWK_E_INVALID, "null GEMM operand");
    pwk::launch_zgemm_mfma_batched(c, m, n, k, a, lda, stride_a, b, ldb, stride_b, cmat, ldc, stride_c, batch,
                                   accumulate != 0, precision);
  });
}

int pwk_dense_h(pwk_ctx* ctx, int ik, int with_nl, double* h) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    c.check_basis();
    DevBuf d;
    size_t nd = (size_t)2 * c.npw * c.npw;
    d.ensure(nd * 8);
    pwk::launch_dense_h(c, ik, with_nl, d.p);
    to_host(c, h, d, nd, PWK_FP64);
  });
}

int pwk_dev_dense_h(pwk_ctx* ctx, int ik, int with_nl, void* h) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    pwk::launch_dense_h(c, ik, with_nl, h);
  });
}

int pwk_dense_eigh(pwk_ctx* ctx, int n, int nb, int nk, double* h, double* w, double* v) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    PWK_REQUIRE(h && w && v, PWK_E_INVALID, "null eigensolver buffer");
    PWK_REQUIRE(n >= 1 && nb >= 1 && nb <= n && nk >= 1, PWK_E_INVALID, "bad eigensolver shape");
    PWK_REQUIRE(pwk::heev_supported(n, nb), PWK_E_UNSUPPORTED, "matrix too large for the batched eigensolver");
    DevBuf dh, dw, dv;
    size_t nh = (size_t)2 * n * n * nk, nv = (size_t)2 * n * nb * nk, nw = (size_t)nb * nk;
    dh.ensure(nh * 8);
    dw.ensure(nw * 8);
    dv.ensure(nv * 8);
    PWK_HIP(hipMemcpyAsync(dh.p, h, nh * 8, hipMemcpyHostToDevice, c.stream));
    c.eig_work.ensure(pwk::heev_work_doubles(n, nb, nk) * 8);
    pwk::launch_heev_batched(c, n, nb, nk, dh.p, c.eig_work.p, dw.as<double>(), dv.p, nullptr);
    PWK_HIP(hipMemcpyAsync(h, dh.p, nh * 8, hipMemcpyDeviceToHost, c.stream));
    PWK_HIP(hipMemcpyAsync(w, dw.p, nw * 8, hipMemcpyDeviceToHost, c.stream));
    PWK_HIP(hipMemcpyAsync(v, dv.p, nv * 8, hipMemcpyDeviceToHost, c.stream));
    PWK_HIP(hipStreamSynchronize(c.stream));
  });
}

int pwk_dev_dense_eigh(pwk_ctx* ctx, int n, int nb, int nk, void* h, void* w, void* v) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    PWK_REQUIRE(h && w && v, PWK_E_INVALID, "null eigensolver buffer");
    c.eig_work.ensure(pwk::heev_work_doubles(n, nb, nk) * 8);
    pwk::launch_heev_batched(c, n, nb, nk, h, c.eig_work.p, (double*)w, v, nullptr);
  });
}

int pwk_dev_dense_eigh_noalloc(pwk_ctx* ctx, int n, int nb, int nk, void* h, void* w, void* v) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    PWK_REQUIRE(h && w && v, PWK_E_INVALID, "null eigensolver buffer");
    PWK_REQUIRE(n >= 1 && nb >= 1 && nb <= n && nk >= 1, PWK_E_INVALID, "bad eigensolver shape");
    PWK_REQUIRE(pwk::heev_supported(n, nb), PWK_E_UNSUPPORTED, "matrix too large for the batched eigensolver");
    pwk::launch_heev_noalloc(c, n, nb, nk, h, (double*)w, v);
  });
}

int pwk_density_accumulate(pwk_ctx* ctx, int nb, const double* psi, const double* fw, double* rho, int precision) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    check_precision(precision);
    c.check_basis();
    PWK_REQUIRE(nb > 0 && psi && fw && rho, PWK_E_INVALID, "bad density arguments");
    size_t nd = (size_t)2 * c.npw * nb;
    DevBuf in, r;
    to_device(c, in, psi, nd, precision);
    r.ensure(c.N * 8);
    PWK_HIP(hipMemcpyAsync(r.p, rho, c.N * 8, hipMemcpyHostToDevice, c.stream));
    pwk::launch_density(c, nb, in.p, fw, r.p, precision);
    to_host(c, rho, r, c.N, PWK_FP64);
  });
}

int pwk_dev_density_accumulate(pwk_ctx* ctx, int nb, const void* psi, const double* fw_host, void* rho,
                               int precision) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    check_precision(precision);
    pwk::launch_density(c, nb, psi, fw_host, rho, precision);
  });
}

// ---- k × band batches and the warm-start Rayleigh-Ritz ----------------------------------
namespace {
size_t kbatch_bands(const pwk_kbatch* b) {
  PWK_REQUIRE(b && b->nentries >= 1 && b->ik && b->nbands, PWK_E_INVALID, "empty or null k-point batch");
  size_t total = 0;
  for (int e = 0; e < b->nentries; ++e) {
    PWK_REQUIRE(b->nbands[e] >= 1, PWK_E_INVALID, "every batch entry needs at least one band");
    total += (size_t)b->nbands[e];
  }
  return total;
}
}  // namespace

int pwk_hpsi_batch(pwk_ctx* ctx, const pwk_kbatch* batch, const double* psi, double* hpsi, int precision) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    check_precision(precision);
    c.check_basis();
    const size_t total = kbatch_bands(batch);
    PWK_REQUIRE(psi && hpsi, PWK_E_INVALID, "bad psi/hpsi");
    const size_t nd = (size_t)2 * c.npw * total;
    DevBuf in, out;
    to_device(c, in, psi, nd, precision);
    out.ensure(nd * rsize(precision));
    pwk::launch_hpsi_batch(c, batch->nentries, batch->ik, batch->nbands, in.p, out.p, precision);
    to_host(c, hpsi, out, nd, precision);
  });
}

int pwk_dev_hpsi_batch(pwk_ctx* ctx, const pwk_kbatch* batch, const void* psi, void* hpsi, int precision) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    check_precision(precision);
    kbatch_bands(batch);
    PWK_REQUIRE(psi && hpsi, PWK_E_INVALID, "bad psi/hpsi");
    pwk::launch_hpsi_batch(c, batch->nentries, batch->ik, batch->nbands, psi, hpsi, precision);
  });
}

int pwk_density_accumulate_batch(pwk_ctx* ctx, const pwk_kbatch* batch, const double* psi, const double* fw,
                                 double* rho, int precision) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    check_precision(precision);
    c.check_basis();
    const size_t total = kbatch_bands(batch);
    PWK_REQUIRE(psi && fw && rho, PWK_E_INVALID, "bad density arguments");
    const size_t nd = (size_t)2 * c.npw * total;
    DevBuf in, r;
    to_device(c, in, psi, nd, precision);
    r.ensure(c.N * 8);
    PWK_HIP(hipMemcpyAsync(r.p, rho, c.N * 8, hipMemcpyHostToDevice, c.stream));
    pwk::launch_density_batch(c, batch->nentries, batch->ik, batch->nbands, in.p, fw, r.p, precision);
    to_host(c, rho, r, c.N, PWK_FP64);
  });
}

int pwk_dev_density_accumulate_batch(pwk_ctx* ctx, const pwk_kbatch* batch, const void* psi, const double* fw_host,
                                     void* rho, int precision) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    check_precision(precision);
    kbatch_bands(batch);
    PWK_REQUIRE(psi && fw_host && rho, PWK_E_INVALID, "bad density arguments");
    pwk::launch_density_batch(c, batch->nentries, batch->ik, batch->nbands, psi, fw_host, rho, precision);
  });
}

int pwk_dev_subspace_rr(pwk_ctx* ctx, int n, int nb, int nk, const void* h, const void* v_prev, double tol, void* w,
                        void* v, int* used_full, double* max_residual) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    PWK_REQUIRE(h && w && v, PWK_E_INVALID, "null subspace buffer");
    PWK_REQUIRE(n >= 1 && nb >= 1 && nb <= n && nk >= 1, PWK_E_INVALID, "bad subspace shape");
    PWK_REQUIRE(v != v_prev, PWK_E_INVALID, "v and v_prev must not overlap");
    pwk::subspace_rr(c, n, nb, nk, h, v_prev, tol > 0.0 ? tol : 1e-6, (double*)w, v, used_full, max_residual);
  });
}

int pwk_subspace_rr(pwk_ctx* ctx, int n, int nb, int nk, const double* h, const double* v_prev, double tol, double* w,
                    double* v, int* used_full, double* max_residual) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    PWK_REQUIRE(h && w && v, PWK_E_INVALID, "null subspace buffer");
    PWK_REQUIRE(n >= 1 && nb >= 1 && nb <= n && nk >= 1, PWK_E_INVALID, "bad subspace shape");
    const size_t nh = (size_t)2 * n * n * nk, nv = (size_t)2 * n * nb * nk, nw = (size_t)nb * nk;
    DevBuf dh, dp, dw, dv;
    dh.ensure(nh * 8);
    dw.ensure(nw * 8);
    dv.ensure(nv * 8);
    PWK_HIP(hipMemcpyAsync(dh.p, h, nh * 8, hipMemcpyHostToDevice, c.stream));
    if (v_prev) {
      dp.ensure(nv * 8);
      PWK_HIP(hipMemcpyAsync(dp.p, v_prev, nv * 8, hipMemcpyHostToDevice, c.stream));
    }
    pwk::subspace_rr(c, n, nb, nk, dh.p, v_prev ? dp.p : nullptr, tol > 0.0 ? tol : 1e-6, dw.as<double>(), dv.p,
                     used_full, max_residual);
    PWK_HIP(hipMemcpyAsync(w, dw.p, nw * 8, hipMemcpyDeviceToHost, c.stream));
    PWK_HIP(hipMemcpyAsync(v, dv.p, nv * 8, hipMemcpyDeviceToHost, c.stream));
    PWK_HIP(hipStreamSynchronize(c.stream));
  });
}

// ---- integer tables --------------------------------------------------------------
namespace {
pwk::tab::Lat lat_of(const double* l) {
  pwk::tab::Lat a;
  for (int i = 0; i < 3; ++i) a.v[i] = {l[3 * i], l[3 * i + 1], l[3 * i + 2]};
  return a;
}
std::vector<pwk::tab::AtomSite> sites_of(int natoms, const double* frac, const int32_t* z) {
  PWK_REQUIRE(natoms >= 0 && (natoms == 0 || (frac && z)), PWK_E_INVALID, "bad atom arrays");
  std::vector<pwk::tab::AtomSite> s((size_t)natoms);
  for (int a = 0; a < natoms; ++a) {
    s[(size_t)a].z = z[a];
    for (int d = 0; d < 3; ++d) s[(size_t)a].frac[d] = frac[3 * a + d];
  }
  return s;
}
}  // namespace

int pwk_basis(const double lattice[9], double ecut_ev, int cap, int32_t* miller, int* npw) {
  return guarded([&] {
    PWK_REQUIRE(lattice && npw && ecut_ev > 0.0 && cap >= 0 && (cap == 0 || miller), PWK_E_INVALID,
                "bad basis arguments");
    std::vector<int32_t> m;
    std::vector<pwk::tab::Vec3> g;
    pwk::tab::sphere_basis(lat_of(lattice), ecut_ev, m, g);
    *npw = (int)g.size();
    const size_t n = std::min<size_t>((size_t)cap, g.size());
    std::memcpy(miller, m.data(), n * 3 * sizeof(int32_t));
  });
}

int pwk_fft_dims(const int32_t* miller, int npw, int ecutrho_ratio, int dims[3]) {
  return guarded([&] {
    PWK_REQUIRE(miller && npw > 0 && dims && ecutrho_ratio > 0, PWK_E_INVALID, "bad grid arguments");
    std::vector<int32_t> m(miller, miller + (size_t)npw * 3);
    pwk::tab::grid_dims(m, ecutrho_ratio, nullptr, dims);
  });
}

int pwk_symmetry(const double lattice[9], int natoms, const double* frac, const int32_t* z, double symprec,
                 int cap, int32_t* rot, double* tau, int* nops) {
  return guarded([&] {
    PWK_REQUIRE(lattice && nops && cap >= 0 && (cap == 0 || (rot && tau)), PWK_E_INVALID,
                "bad symmetry arguments");
    auto ops = pwk::tab::detect_symmetry(lat_of(lattice), sites_of(natoms, frac, z), symprec);
    *nops = (int)ops.size();
    for (size_t i = 0; i < std::min<size_t>((size_t)cap, ops.size()); ++i) {
      std::memcpy(rot + 9 * i, ops[i].rot, 9 * sizeof(int32_t));
      std::memcpy(tau + 3 * i, ops[i].tau, 3 * sizeof(double));
    }
  });
}

int pwk_kpoints(const double lattice[9], int natoms, const double* frac, const int32_t* z, const int kgrid[3],
                const int kshift[3], int use_symmetry, double symprec, int cap, double* kcart, double* weight,
                int* nk) {
  return guarded([&] {
    PWK_REQUIRE(lattice && kgrid && kshift && nk && cap >= 0 && (cap == 0 || (kcart && weight)), PWK_E_INVALID,
                "bad k-point arguments");
    for (int d = 0; d < 3; ++d) PWK_REQUIRE(kgrid[d] >= 1, PWK_E_INVALID, "k-grid must be positive");
    pwk::tab::Lat lat = lat_of(lattice);
    std::vector<pwk::tab::SymOp> ops;
    if (use_symmetry) ops = pwk::tab::detect_symmetry(lat, sites_of(natoms, frac, z), symprec);
    if (ops.empty()) {
      pwk::tab::SymOp id{};
      id.rot[0] = id.rot[4] = id.rot[8] = 1;
      ops.push_back(id);
    }
    auto ks = pwk::tab::kpoints(lat, kgrid, kshift, ops, use_symmetry != 0);
    *nk = (int)ks.size();
    for (size_t i = 0; i < std::min<size_t>((size_t)cap, ks.size()); ++i) {
      kcart[3 * i] = ks[i].cart.x;
      kcart[3 * i + 1] = ks[i].cart.y;
      kcart[3 * i + 2] = ks[i].cart.z;
      weight[i] = ks[i].weight;
    }
  });
}

// ---- SCF -------------------------------------------------------------------------
int pwk_scf_run(pwk_ctx* ctx, const pwk_crystal* crystal, const pwk_scf_params* params, pwk_scf_result* result,
                double* eigenvalues, double* rho) {
  return guarded([&] {
    Context& c = ctx_of(ctx);
    PWK_REQUIRE(crystal && params && result, PWK_E_INVALID, "null SCF argument");
    check_precision(params->precision);
    pwk::scf_run(c, *crystal, *params, *result, eigenvalues, rho);
  });
}

int pwk_scf_tables(const pwk_crystal* crystal, const pwk_scf_params* params, int* npw, int dims[3], int* nk,
                   int* nops) {
  return guarded([&] {
    PWK_REQUIRE(crystal && params && npw && dims && nk && nops, PWK_E_INVALID, "null argument");
    pwk::scf_tables(*crystal, *params, npw, dims, nk, nops);  // validates the request first
  });
}

}  // extern "C"
