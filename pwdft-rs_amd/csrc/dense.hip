// Dense Kohn-Sham Hamiltonian of one k-point, FP64, column-major npw×npw:
//   H[i,j] = V_eff[wrap(m_i - m_j)] + δ_ij ħ²/2m |k+G_i|²   (src/scf/potentials.rs:163-190)
// then H += B·(D·B^H)                                        (src/potential/nonlocal.rs:432-450)
// The gather is an HBM-write-bound pass (npw²·16 B); V_eff(G) stays in L2.
#include "common.hpp"

namespace pwk {

namespace {

__device__ __forceinline__ int wrapd(int n, int d) {
  int r = n % d;
  return r < 0 ? r + d : r;
}

// blockIdx.z selects the k-point: kinetic diagonal kin + z·npw, matrix H + z·npw²
__global__ void __launch_bounds__(256) k_dense_h(int npw, const int* __restrict__ miller,
                                                 const double2* __restrict__ veff, const double* __restrict__ kin,
                                                 int nx, int ny, int nz, double2* __restrict__ H) {
  int i = blockIdx.x * 64 + (threadIdx.x & 63);
  int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= npw || j >= npw) return;
  kin += (size_t)blockIdx.z * npw;
  H += (size_t)blockIdx.z * npw * npw;
  int d1 = miller[3 * i] - miller[3 * j];
  int d2 = miller[3 * i + 1] - miller[3 * j + 1];
  int d3 = miller[3 * i + 2] - miller[3 * j + 2];
  size_t idx = ((size_t)wrapd(d1, nx) * ny + wrapd(d2, ny)) * nz + wrapd(d3, nz);
  double2 v = veff[idx];
  if (i == j) {
    v.x = kin[i] + v.x;
    v.y = 0.0 + v.y;
  }
  H[(size_t)j * npw + i] = v;
}

}  // namespace

void launch_dense_h(Context& c, int ik, int with_nl, void* h) {
  c.check_basis();
  KData& kd = c.k(ik);
  PWK_REQUIRE(kd.have_kin, PWK_E_STATE, "kinetic diagonal of this k-point not set (pwk_set_kinetic)");
  PWK_REQUIRE(c.have_v, PWK_E_STATE, "V_eff not set (pwk_set_veff)");
  int n = c.npw;
  dim3 grid(blocks_for((size_t)n, 64), blocks_for((size_t)n, 4));
  k_dense_h<<<grid, 256, 0, c.stream>>>(n, c.miller.as<int>(), c.veffg64.as<double2>(), kd.kin64.as<double>(),
                                        c.dims[0], c.dims[1], c.dims[2], (double2*)h);
  PWK_HIP(hipGetLastError());
  if (with_nl && kd.nch > 0)
    launch_zgemm_mfma(c, n, n, kd.nch, kd.B64.p, n, kd.DBH64.p, kd.nch, h, n, true, PWK_FP64);
}

// Dense H of nk k-points at once: kin nk × npw, projectors in contiguous blocks
// (B: npw × nch, D·B^H: nch × npw per k), H nk column-major npw × npw blocks.
void launch_dense_h_batched(Context& c, int nk, const double* kin, int nch, const void* B, const void* DBH, void* h) {
  c.check_basis();
  PWK_REQUIRE(c.have_v, PWK_E_STATE, "V_eff not set (pwk_set_veff)");
  PWK_REQUIRE(nk >= 1 && nk <= 65535, PWK_E_INVALID, "bad k-point count");
  int n = c.npw;
  dim3 grid(blocks_for((size_t)n, 64), blocks_for((size_t)n, 4), (unsigned)nk);
  k_dense_h<<<grid, 256, 0, c.stream>>>(n, c.miller.as<int>(), c.veffg64.as<double2>(), kin, c.dims[0], c.dims[1],
                                        c.dims[2], (double2*)h);
  PWK_HIP(hipGetLastError());
  if (nch > 0)
    launch_zgemm_mfma_batched(c, n, n, nch, B, n, (long long)n * nch, DBH, nch, (long long)nch * n, h, n,
                              (long long)n * n, nk, true, PWK_FP64);
}

}  // namespace pwk
