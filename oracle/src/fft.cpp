// Textbook mixed-radix Cooley–Tukey DFT (decimation in time), FP64.
// Semantics of src/fft.rs:19-114: unnormalised in both directions,
// forward kernel e^{-2πi jk/n}, inverse e^{+2πi jk/n}; the 3D transform is
// three sets of 1D transforms along each axis of a row-major (x,y,z) array.
// Test infrastructure (oracle).
#include <map>
#include <mutex>

#include "oracle.hpp"

namespace orc {

namespace {

size_t pick_radix(size_t n) {
  if (n % 4 == 0) return 4;
  for (size_t p : {2, 3, 5, 7})
    if (n % p == 0) return p;
  for (size_t p = 11; p * p <= n; p += 2)
    if (n % p == 0) return p;
  return n;  // prime: direct DFT
}

// y[0..n) = DFT of x[0], x[xs], ..., x[(n-1)xs]; W[j] = exp(sign·2πi·j/N), n | N
void rec(const cplx* x, size_t xs, cplx* y, size_t n, const cplx* W, size_t N) {
  if (n == 1) {
    y[0] = x[0];
    return;
  }
  size_t p = pick_radix(n);
  size_t m = n / p;
  for (size_t r = 0; r < p; ++r) rec(x + r * xs, xs * p, y + r * m, m, W, N);
  size_t step = N / n;     // twiddle stride for this level
  size_t pstep = N / p;    // p-th roots of unity
  cplx t[64];
  std::vector<cplx> big;
  cplx* tp = t;
  if (p > 64) {
    big.resize(p);
    tp = big.data();
  }
  for (size_t k = 0; k < m; ++k) {
    for (size_t r = 0; r < p; ++r) tp[r] = y[r * m + k] * W[(r * k * step) % N];
    for (size_t q = 0; q < p; ++q) {
      cplx acc(0.0, 0.0);
      for (size_t r = 0; r < p; ++r) acc += tp[r] * W[((r * q) % p) * pstep];
      y[k + q * m] = acc;
    }
  }
}

struct Twiddles {
  std::vector<cplx> fwd, inv;
};

const Twiddles& twiddles(size_t n) {
  static std::mutex mu;
  static std::map<size_t, Twiddles> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(n);
  if (it != cache.end()) return it->second;
  Twiddles tw;
  tw.fwd.resize(n);
  tw.inv.resize(n);
  for (size_t j = 0; j < n; ++j) {
    double a = TAU * (double)j / (double)n;
    double c = std::cos(a), s = std::sin(a);
    tw.fwd[j] = cplx(c, -s);
    tw.inv[j] = cplx(c, s);
  }
  return cache.emplace(n, std::move(tw)).first->second;
}

}  // namespace

void fft1d(cplx* data, size_t n, size_t stride, int sign) {
  if (n <= 1) return;
  const Twiddles& tw = twiddles(n);
  std::vector<cplx> in(n), out(n);
  for (size_t i = 0; i < n; ++i) in[i] = data[i * stride];
  rec(in.data(), 1, out.data(), n, sign < 0 ? tw.fwd.data() : tw.inv.data(), n);
  for (size_t i = 0; i < n; ++i) data[i * stride] = out[i];
}

void fft3d(std::vector<cplx>& data, const std::array<size_t, 3>& dims, int sign) {
  size_t nx = dims[0], ny = dims[1], nz = dims[2];
  // axis 0 (x, stride ny·nz)
  for (size_t j = 0; j < ny; ++j)
    for (size_t k = 0; k < nz; ++k) fft1d(&data[j * nz + k], nx, ny * nz, sign);
  // axis 1 (y, stride nz)
  for (size_t i = 0; i < nx; ++i)
    for (size_t k = 0; k < nz; ++k) fft1d(&data[i * ny * nz + k], ny, nz, sign);
  // axis 2 (z, contiguous)
  for (size_t i = 0; i < nx; ++i)
    for (size_t j = 0; j < ny; ++j) fft1d(&data[(i * ny + j) * nz], nz, 1, sign);
}

}  // namespace orc
