// Autotuned full-grid 3D FFT (src/fft.rs:19-114 convention; the general path for grids that are
// not small and cubic).
//
// A plan fixes the three choices that shape the transform on the device:
//   ord   radix order of each line's decimation in frequency (0: powers of four first, 1: odd
//         radices first; fft_reg.hpp) — changes the rounding of the result, not its value;
//   tpb   lines (lanes) per workgroup of the strided passes: 64, 128 or 256;
//   zlds  the contiguous pass (lines along the fastest axis) staged through LDS: a workgroup
//         loads its tile of lines with coalesced loads, each lane transforms its line from LDS
//         and the tile is stored back coalesced (instead of one lane streaming its own line).
// tpb and zlds only move data and change no arithmetic. The autotuner times every candidate
// plan with HIP events on the caller's stream the first time a (device, grid, precision, batch)
// is transformed, keeps the fastest in a process-wide table (so every context of the process
// uses the same plan and results agree between them), and later transforms look the plan up.
// Tuning is skipped under stream capture (the default plan runs) and can be switched off per
// context or overridden by an explicit plan.
#include <chrono>
#include <map>
#include <type_traits>
#include <mutex>
#include <tuple>

#include "common.hpp"
#include "fft_reg.hpp"
#include "fft_sizes.hpp"

namespace pwk {

namespace {

template <typename T>
struct Cplx;
template <>
struct Cplx<double> {
  using t = double2;
};
template <>
struct Cplx<float> {
  using t = float2;
};

constexpr int kMaxTpb = 256;

// Line l starts at (l / cnt_in)·st_out + (l % cnt_in)·st_in; elements at stride es; batch item
// blockIdx.y at batch_stride. One lane per line, the line in registers.
template <typename T, int N, int SIGN, int ORD>
__global__ void __launch_bounds__(kMaxTpb) k_lines_t(typename Cplx<T>::t* __restrict__ data, int nlines, int cnt_in,
                                                     long st_out, long st_in, long es, long batch_stride) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nlines) return;
  typename Cplx<T>::t* p =
      data + (size_t)blockIdx.y * batch_stride + (size_t)(l / cnt_in) * st_out + (size_t)(l % cnt_in) * st_in;
  T re[N], im[N];
#pragma unroll
  for (int j = 0; j < N; ++j) {
    const typename Cplx<T>::t v = p[(size_t)j * es];
    re[j] = v.x;
    im[j] = v.y;
  }
  fft_dif<T, N, SIGN, ORD>(re, im);
#pragma unroll
  for (int f = 0; f < N; ++f) {
    typename Cplx<T>::t v;
    v.x = re[FftOrder<N, ORD>::pos(f)];
    v.y = im[FftOrder<N, ORD>::pos(f)];
    p[(size_t)f * es] = v;
  }
}

// Contiguous lines (line l = elements [l·N, l·N + N) of the batch item) staged through LDS: the
// workgroup's blockDim.x lines are loaded and stored with consecutive lanes on consecutive
// elements; each line sits in LDS with one element of padding after it.
template <typename T, int N, int SIGN, int ORD>
__global__ void __launch_bounds__(kMaxTpb) k_lines_lds(typename Cplx<T>::t* __restrict__ data, int nlines,
                                                       long batch_stride) {
  using V2 = typename Cplx<T>::t;
  extern __shared__ unsigned char smem_raw[];
  V2* s = reinterpret_cast<V2*>(smem_raw);
  const int tpb = blockDim.x, l0 = blockIdx.x * tpb;
  const int nl = min(tpb, nlines - l0);
  V2* base = data + (size_t)blockIdx.y * batch_stride + (size_t)l0 * N;
  for (int e = threadIdx.x; e < nl * N; e += tpb) s[(e / N) * (N + 1) + e % N] = base[e];
  __syncthreads();
  if ((int)threadIdx.x < nl) {
    V2* q = s + threadIdx.x * (N + 1);
    T re[N], im[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      re[j] = q[j].x;
      im[j] = q[j].y;
    }
    fft_dif<T, N, SIGN, ORD>(re, im);
#pragma unroll
    for (int f = 0; f < N; ++f) {
      V2 v;
      v.x = re[FftOrder<N, ORD>::pos(f)];
      v.y = im[FftOrder<N, ORD>::pos(f)];
      q[f] = v;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < nl * N; e += tpb) base[e] = s[(e / N) * (N + 1) + e % N];
}

// whether the two radix orders give different factor sequences for length n
constexpr bool orders_differ(int n) {
  int a = n, b = n;
  while (a > 1) {
    const int pa = first_radix(a, 0), pb = first_radix(b, 1);
    if (pa != pb) return true;
    a /= pa;
    b /= pb;
  }
  return false;
}

template <typename T, int SIGN, int ORD>
void pass_lines(hipStream_t st, void* data, int batch, long bs, int n, int tpb, int nlines, int cnt_in, long st_out,
                long st_in, long es) {
  if (!fft_len_compiled(n)) {  // lines beyond the register lengths: staged in LDS (fft_long.hip)
    launch_long_lines(st, data, batch, bs, n, nlines, cnt_in, st_out, st_in, es, SIGN,
                      std::is_same<T, double>::value ? PWK_FP64 : PWK_FP32);
    return;
  }
  const dim3 grid(blocks_for((size_t)nlines, tpb), (unsigned)batch);
  auto* d = static_cast<typename Cplx<T>::t*>(data);
  switch (n) {
#define PWK_CASE(L)                                                                                   \
  case L:                                                                                             \
    k_lines_t<T, L, SIGN, ORD><<<grid, tpb, 0, st>>>(d, nlines, cnt_in, st_out, st_in, es, bs);       \
    break;
    PWK_FOR_EACH_FFT_SIZE(PWK_CASE)
#undef PWK_CASE
    default:
      throw Error(PWK_E_UNSUPPORTED, "FFT length " + std::to_string(n) + " has no compiled kernel");
  }
  PWK_HIP(hipGetLastError());
}

template <typename T>
size_t zlds_bytes(int n, int tpb) {
  return (size_t)tpb * (n + 1) * sizeof(typename Cplx<T>::t);
}

template <typename T, int SIGN, int ORD>
void pass_lines_lds(hipStream_t st, void* data, int batch, long bs, int n, int tpb, int nlines) {
  const dim3 grid(blocks_for((size_t)nlines, tpb), (unsigned)batch);
  auto* d = static_cast<typename Cplx<T>::t*>(data);
  const size_t sh = zlds_bytes<T>(n, tpb);
  switch (n) {
#define PWK_CASE(L)                                                                      \
  case L:                                                                                \
    ensure_lds_limit((const void*)k_lines_lds<T, L, SIGN, ORD>, sh);                     \
    k_lines_lds<T, L, SIGN, ORD><<<grid, tpb, sh, st>>>(d, nlines, bs);                  \
    break;
    PWK_FOR_EACH_FFT_SIZE(PWK_CASE)
#undef PWK_CASE
    default:
      throw Error(PWK_E_UNSUPPORTED, "FFT length " + std::to_string(n) + " has no compiled kernel");
  }
  PWK_HIP(hipGetLastError());
}

template <typename T, int SIGN, int ORD>
void run_plan_t(Context& c, void* data, int batch, const FftPlan& pl) {
  const int nx = c.dims[0], ny = c.dims[1], nz = c.dims[2];
  const long bs = (long)c.N;
  // x (stride ny·nz), y (stride nz), z (contiguous)
  pass_lines<T, SIGN, ORD>(c.stream, data, batch, bs, nx, pl.tpb, ny * nz, ny * nz, 0, 1, (long)ny * nz);
  pass_lines<T, SIGN, ORD>(c.stream, data, batch, bs, ny, pl.tpb, nx * nz, nz, (long)ny * nz, 1, nz);
  if (pl.zlds && fft_len_compiled(nz))
    pass_lines_lds<T, SIGN, ORD>(c.stream, data, batch, bs, nz, pl.tpb, nx * ny);
  else
    pass_lines<T, SIGN, ORD>(c.stream, data, batch, bs, nz, pl.tpb, nx * ny, nx * ny, 0, nz, 1);
}

void run_plan(Context& c, void* data, int batch, int sign, int precision, const FftPlan& pl) {
  const bool f64 = precision == PWK_FP64;
  if (pl.ord == 0) {
    if (f64)
      sign < 0 ? run_plan_t<double, -1, 0>(c, data, batch, pl) : run_plan_t<double, 1, 0>(c, data, batch, pl);
    else
      sign < 0 ? run_plan_t<float, -1, 0>(c, data, batch, pl) : run_plan_t<float, 1, 0>(c, data, batch, pl);
  } else {
    if (f64)
      sign < 0 ? run_plan_t<double, -1, 1>(c, data, batch, pl) : run_plan_t<double, 1, 1>(c, data, batch, pl);
    else
      sign < 0 ? run_plan_t<float, -1, 1>(c, data, batch, pl) : run_plan_t<float, 1, 1>(c, data, batch, pl);
  }
}

bool any_orders_differ(const int* dims) {
  for (int d = 0; d < 3; ++d) {
    switch (dims[d]) {
#define PWK_CASE(L) \
  case L:           \
    if (orders_differ(L)) return true; \
    break;
      PWK_FOR_EACH_FFT_SIZE(PWK_CASE)
#undef PWK_CASE
      default:
        break;
    }
  }
  return false;
}

// the candidate plans for a grid: every radix order that differs, three lane counts, and the
// LDS-staged contiguous pass where its tile fits the LDS
std::vector<FftPlan> candidates(const int* dims, int precision) {
  std::vector<FftPlan> out;
  const int nord = any_orders_differ(dims) ? 2 : 1;
  const size_t es = precision == PWK_FP64 ? sizeof(double2) : sizeof(float2);
  for (int ord = 0; ord < nord; ++ord)
    for (int tpb : {64, 128, 256})
      for (int z = 0; z < 2; ++z) {
        if (z && (!fft_len_compiled(dims[2]) || (size_t)tpb * (dims[2] + 1) * es > 64 * 1024))
          continue;  // a register-length contiguous pass, at least two tiles per CU
        FftPlan p;
        p.ord = ord;
        p.tpb = tpb;
        p.zlds = z;
        out.push_back(p);
      }
  return out;
}

struct PlanKey {
  int device, nx, ny, nz, precision, batch;
  bool operator<(const PlanKey& o) const {
    return std::tie(device, nx, ny, nz, precision, batch) < std::tie(o.device, o.nx, o.ny, o.nz, o.precision, o.batch);
  }
};

struct PlanEntry {
  FftPlan plan;
  double us_best = 0.0, us_default = 0.0;
  int ncand = 0;
};

std::mutex g_plan_mu;
std::map<PlanKey, PlanEntry> g_plans;

template <typename T>
__global__ void k_fill_pattern(size_t n, T* x) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = (T)((double)((i * 2654435761u) % 1000u) * 1e-3 - 0.5);
}

void fill_pattern(Context& c, void* buf, size_t n, int precision) {
  if (precision == PWK_FP64)
    k_fill_pattern<double><<<blocks_for(n, 256), 256, 0, c.stream>>>(n, static_cast<double*>(buf));
  else
    k_fill_pattern<float><<<blocks_for(n, 256), 256, 0, c.stream>>>(n, static_cast<float*>(buf));
  PWK_HIP(hipGetLastError());
}

// average microseconds of one transform with plan pl on a freshly filled buffer (two warm runs,
// then five timed, forward and inverse alternating)
double time_plan(Context& c, void* buf, int batch, int precision, const FftPlan& pl, hipEvent_t e0, hipEvent_t e1) {
  fill_pattern(c, buf, (size_t)batch * c.N * 2, precision);
  for (int i = 0; i < 2; ++i) run_plan(c, buf, batch, i % 2 ? 1 : -1, precision, pl);
  PWK_HIP(hipEventRecord(e0, c.stream));
  constexpr int reps = 5;
  for (int i = 0; i < reps; ++i) run_plan(c, buf, batch, i % 2 ? 1 : -1, precision, pl);
  PWK_HIP(hipEventRecord(e1, c.stream));
  PWK_HIP(hipEventSynchronize(e1));
  float ms = 0.0f;
  PWK_HIP(hipEventElapsedTime(&ms, e0, e1));
  return 1e3 * (double)ms / reps;
}

PlanEntry tune(Context& c, int batch, int precision) {
  const std::vector<FftPlan> cands = candidates(c.dims, precision);
  const size_t es = precision == PWK_FP64 ? sizeof(double2) : sizeof(float2);
  DevBuf buf;
  buf.ensure((size_t)batch * c.N * es);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  PWK_HIP(hipEventCreate(&e0));
  PWK_HIP(hipEventCreate(&e1));
  PlanEntry best;
  best.ncand = (int)cands.size();
  try {
    double tbest = 1e300;
    for (size_t i = 0; i < cands.size(); ++i) {
      const double t = time_plan(c, buf.p, batch, precision, cands[i], e0, e1);
      if (i == 0) best.us_default = t;  // candidates[0] is the default plan
      if (t < tbest) {
        tbest = t;
        best.plan = cands[i];
      }
    }
    best.us_best = tbest;
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}

bool capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}

}  // namespace

FftPlan fft_plan_for(Context& c, int batch, int precision, double* us_best, double* us_default, int* ncand) {
  const int prec = precision == PWK_FP64 ? PWK_FP64 : PWK_FP32;
  if (c.fft_plan_forced) return c.fft_plan;
  if (!c.fft_autotune || capturing(c.stream)) return FftPlan{};
  const PlanKey key{c.device, c.dims[0], c.dims[1], c.dims[2], prec, batch};
  {
    std::lock_guard<std::mutex> lk(g_plan_mu);
    auto it = g_plans.find(key);
    if (it != g_plans.end()) {
      if (us_best) *us_best = it->second.us_best;
      if (us_default) *us_default = it->second.us_default;
      if (ncand) *ncand = it->second.ncand;
      return it->second.plan;
    }
  }
  // tuned outside the lock (concurrent first uses of one key may both tune; the first result
  // stored wins, so every later lookup returns the same plan)
  const PlanEntry e = tune(c, batch, prec);
  std::lock_guard<std::mutex> lk(g_plan_mu);
  const PlanEntry& kept = g_plans.emplace(key, e).first->second;
  if (us_best) *us_best = kept.us_best;
  if (us_default) *us_default = kept.us_default;
  if (ncand) *ncand = kept.ncand;
  return kept.plan;
}

void fft3d_planned(Context& c, void* data, int batch, int sign, int precision) {
  const FftPlan pl = fft_plan_for(c, batch, precision);
  if (pl.zlds) {
    const size_t es = precision == PWK_FP64 ? sizeof(double2) : sizeof(float2);
    PWK_REQUIRE((size_t)pl.tpb * (c.dims[2] + 1) * es <= 160 * 1024, PWK_E_UNSUPPORTED,
                "the FFT plan's LDS tile exceeds the LDS for this grid");
  }
  run_plan(c, data, batch, sign, precision, pl);
}


}  // namespace pwk
