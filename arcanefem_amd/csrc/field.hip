// field.hip -- log-normal coefficient fields for the heterogeneous Poisson
// workload: lambda_c = exp(g(x_c)) at the cell centroids x_c, with g a
// zero-mean Gaussian random field of squared-exponential covariance
//   C(r) = sigma^2 exp(-|r|^2 / (2 l^2)),
// sampled with K random Fourier features (Rahimi & Recht, "Random features
// for large-scale kernel machines", 2007):
//   g(x) = sigma sqrt(2/K) sum_k cos(w_k . x + phi_k),
//   w_k ~ N(0, I / l^2),  phi_k ~ U[0, 2 pi).
// E[g(x) g(y)] = C(x - y) exactly; the marginal is Gaussian as K grows.
// The reference has no such generator: its Fourier module fills one
// conductivity per cell group from the case options
// (modules/fourier/FemModule.cc:175-190), so this is an input generator for
// the workload, not a restated reference path.
#include "common.h"
#include "device_utils.h"

#include <cmath>
#include <vector>

namespace afem {

namespace {

uint64_t mix64(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// uniform in (0, 1]
double unit_open(uint64_t seed, uint64_t key)
{
  const uint64_t r = mix64(mix64(seed ^ 0x5DEECE66DULL) + key);
  return ((double)(r >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}

constexpr double kTwoPi = 6.283185307179586476925286766559;

void rff_parameters(int dim, int nfeatures, double corr_length, uint64_t seed, double* omega, double* phase)
{
  AFEM_REQUIRE(dim == 2 || dim == 3, AFEM_ERR_ARGUMENT, "dim must be 2 or 3");
  AFEM_REQUIRE(nfeatures >= 1 && nfeatures <= 4096, AFEM_ERR_ARGUMENT, "nfeatures must be in [1, 4096]");
  AFEM_REQUIRE(corr_length > 0.0, AFEM_ERR_ARGUMENT, "correlation length must be positive");
  for (int k = 0; k < nfeatures; ++k) {
    for (int d = 0; d < 3; ++d) {
      // Box-Muller on two independent uniforms per component
      const double u1 = unit_open(seed, (uint64_t)k * 8 + 2 * d);
      const double u2 = unit_open(seed, (uint64_t)k * 8 + 2 * d + 1);
      const double gauss = std::sqrt(-2.0 * std::log(u1)) * std::cos(kTwoPi * u2);
      omega[3 * k + d] = (d < dim) ? gauss / corr_length : 0.0;
    }
    phase[k] = kTwoPi * unit_open(seed, (uint64_t)k * 8 + 6);
  }
}

constexpr int kFieldBlock = 256;
constexpr int kFieldChunk = 512; // features staged in LDS per pass

// One thread per cell; features staged through LDS in chunks.  The phase is
// evaluated in FP32 (|w.x| stays below ~50 on the unit cube for l >= 0.05),
// the sum in FP32 per chunk and FP64 across chunks.
__global__ void __launch_bounds__(kFieldBlock)
k_lognormal(const double* __restrict__ coords, const int32_t* __restrict__ cells, int npc, int stride, int32_t nb_cell,
            const float4* __restrict__ feat, int nfeatures, double scale, double* __restrict__ lambda)
{
  __shared__ float4 s_feat[kFieldChunk];
  const int32_t c = blockIdx.x * kFieldBlock + threadIdx.x;
  float cx = 0.f, cy = 0.f, cz = 0.f;
  if (c < nb_cell) {
    double x = 0.0, y = 0.0, z = 0.0;
    for (int k = 0; k < npc; ++k) {
      const int64_t n = cells[stride * (int64_t)c + k];
      x += coords[4 * n];
      y += coords[4 * n + 1];
      z += coords[4 * n + 2];
    }
    cx = (float)(x / npc);
    cy = (float)(y / npc);
    cz = (float)(z / npc);
  }
  double g = 0.0;
  for (int k0 = 0; k0 < nfeatures; k0 += kFieldChunk) {
    const int cnt = min(kFieldChunk, nfeatures - k0);
    __syncthreads();
    for (int k = threadIdx.x; k < cnt; k += kFieldBlock)
      s_feat[k] = feat[k0 + k];
    __syncthreads();
    float part = 0.f;
    for (int k = 0; k < cnt; ++k) {
      const float4 f = s_feat[k];
      part += __cosf(f.x * cx + f.y * cy + f.z * cz + f.w);
    }
    g += part;
  }
  if (c < nb_cell)
    lambda[c] = exp(scale * g);
}

} // namespace

} // namespace afem

using namespace afem;

extern "C" {

AFEM_API int afem_rff_parameters(int dim, int nfeatures, double corr_length, uint64_t seed, double* omega,
                                 double* phase)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(omega && phase, AFEM_ERR_ARGUMENT, "null argument");
  rff_parameters(dim, nfeatures, corr_length, seed, omega, phase);
  AFEM_GUARD_END
}

AFEM_API int afem_lognormal_field(afem_mesh_t mesh, double sigma, double corr_length, int nfeatures, uint64_t seed,
                                  double* cell_lambda)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(mesh && cell_lambda, AFEM_ERR_ARGUMENT, "null argument");
  AFEM_REQUIRE(sigma >= 0.0, AFEM_ERR_ARGUMENT, "sigma must be >= 0");
  std::vector<double> omega((size_t)nfeatures * 3), phase((size_t)nfeatures);
  rff_parameters(mesh->dim, nfeatures, corr_length, seed, omega.data(), phase.data());
  std::vector<float> packed((size_t)nfeatures * 4);
  for (int k = 0; k < nfeatures; ++k) {
    packed[4 * k + 0] = (float)omega[3 * k + 0];
    packed[4 * k + 1] = (float)omega[3 * k + 1];
    packed[4 * k + 2] = (float)omega[3 * k + 2];
    packed[4 * k + 3] = (float)phase[k];
  }
  DBuf<float> feat(packed.size());
  feat.upload(packed.data(), packed.size());
  const double scale = sigma * std::sqrt(2.0 / nfeatures);
  hipLaunchKernelGGL(k_lognormal, dim3(grid_for(mesh->nb_cell, kFieldBlock)), dim3(kFieldBlock), 0, stream(),
                     mesh->coords.p, mesh->cells.p, mesh->npc, mesh->stride, mesh->nb_cell,
                     reinterpret_cast<const float4*>(feat.p), nfeatures, scale, cell_lambda);
  check_launch("lognormal_field");
  AFEM_HIP(hipStreamSynchronize(stream())); // feat is freed on return
  AFEM_GUARD_END
}

} // extern "C"
