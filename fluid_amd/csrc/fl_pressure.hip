// Pressure projection of libfluidhip: the reference's in-place lexicographic
// Gauss-Seidel/SOR sweeps (pressureJacobiIteration, fluid.go:188-234) reproduced in
// exactly the serial update order.
//
// Cell (i,j) in sweep k reads and writes U[i,j], U[i+1,j], V[i,j], V[i,j+1], p[i,j].
// Its dependences are (i-1,j,k), (i,j-1,k), (i+1,j,k-1), (i,j+1,k-1), (i,j,k-1), so the
// hyperplanes t = i + j + 2k order every face's updates as the serial loop does and
// two cells on one hyperplane never touch the same word: a schedule by t is
// bit-identical to the serial sweeps.
//
// This file holds the hyperplane kernel (one launch per t, in place) used for small
// grids and as the cross-check of the temporal-blocked kernel in fl_gs_wave.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>

#include "fl_internal.h"

namespace fl {

int gs_wave_solve(Ctx* c, int numIters, float cp);  // fl_gs_wave.hip
bool gs_wave_supported(Ctx* c, int numIters);

__device__ __forceinline__ float wave_max_f(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

// One GS cell update with the reference's operation order. Weights are the S values
// (exactly 0.0 or 1.0) of the four neighbours.
__global__ void k_gs_hyperplane(int NX, int NY, int t, int k0, const uint8_t* __restrict__ flags,
                                float* __restrict__ U, float* __restrict__ V, float* __restrict__ p,
                                const float* __restrict__ omega, float damping, float cp,
                                uint32_t* __restrict__ maxdiv) {
  const int k = k0 + blockIdx.y;
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = t - i - 2 * k;
  float ad = 0.0f;
  if (i <= NX - 2 && j >= 1 && j <= NY - 2) {
    const int n = NY, c = i * n + j;
    const uint32_t f = flags[c];
    if (f & CF_GS) {
      float sx0 = (f & CF_X0) ? 1.0f : 0.0f;
      float sx1 = (f & CF_X1) ? 1.0f : 0.0f;
      float sy0 = (f & CF_Y0) ? 1.0f : 0.0f;
      float sy1 = (f & CF_Y1) ? 1.0f : 0.0f;
      float s = sx0 + sx1 + sy0 + sy1;
      float div = U[c + n] - U[c] + V[c + 1] - V[c];
      float a = fabsf(div);
      if (a > ad) ad = a;
      float q = -div / s;
      q *= omega[k];
      q *= damping;
      p[c] += cp * q;
      U[c] -= sx0 * q;
      U[c + n] += sx1 * q;
      V[c] -= sy0 * q;
      V[c + 1] += sy1 * q;
    }
  }
  ad = wave_max_f(ad);
  if ((threadIdx.x & 63) == 0 && ad > 0.0f) atomicMax(&maxdiv[k], __float_as_uint(ad));
}

__global__ void k_zero_u32(uint32_t* __restrict__ d, int n) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) d[t] = 0u;
}

static int hyperplane_run(Ctx* c, int K, float cp, uint32_t* maxdiv) {
  const int NX = c->NX, NY = c->NY;
  hipLaunchKernelGGL(k_zero_u32, dim3((K + 255) / 256), dim3(256), 0, c->stream, maxdiv, K);
  const int tmin = 2, tmax = (NX - 2) + (NY - 2) + 2 * (K - 1);
  const int threads = 256;
  const int bx = (NX - 2 + threads - 1) / threads;
  for (int t = tmin; t <= tmax; t++) {
    // k such that 1 <= j = t - i - 2k <= NY-2 for some i in [1, NX-2]
    int klo = std::max(0, (t - (NX - 2) - (NY - 2) + 1) / 2);
    int khi = std::min(K - 1, (t - 2) / 2);
    if (klo > khi) continue;
    hipLaunchKernelGGL(k_gs_hyperplane, dim3(bx, khi - klo + 1), dim3(threads), 0, c->stream, NX, NY, t, klo,
                       c->flags, c->U, c->V, c->p, c->d_omega, c->prm.pressure_damping, cp, maxdiv);
  }
  FL_HIP(hipGetLastError());
  return FL_OK;
}

static int ensure_omega(Ctx* c, int numIters) {
  if (numIters > c->nmax && c->batch) {
    set_error("a batch member supports at most " + std::to_string(c->nmax) + " sweeps");
    return FL_ERR_UNSUPPORTED;
  }
  if (numIters > c->nmax) {
    if (c->d_omega) FL_HIP(hipFree(c->d_omega));
    if (c->d_maxdiv) FL_HIP(hipFree(c->d_maxdiv));
    c->d_omega = nullptr;
    c->d_maxdiv = nullptr;
    int cap = std::max(numIters, 64);
    FL_HIP(hipMalloc(&c->d_omega, sizeof(float) * cap));
    FL_HIP(hipMalloc(&c->d_maxdiv, sizeof(uint32_t) * 2 * cap));
    FL_HIP(hipMemsetAsync(c->d_maxdiv, 0, sizeof(uint32_t) * 2 * cap, c->stream));
    c->nmax = cap;
    c->omega_n = -1;
  }
  // Relaxation schedule of solveSingleGrid (fluid.go:162-170), in float32; uploaded only
  // when the sweep count or Relaxation changes, so a steady-state step never syncs.
  if (c->omega_n == numIters && c->omega_relax == c->prm.relaxation) return FL_OK;
  std::vector<float> om(numIters);
  const float init = c->prm.relaxation, minR = 1.2f;
  for (int k = 0; k < numIters; k++) {
    float prog = (float)k / (float)numIters;
    om[k] = init - (init - minR) * prog;
  }
  FL_HIP(hipMemcpyAsync(c->d_omega, om.data(), sizeof(float) * numIters, hipMemcpyHostToDevice, c->stream));
  // The host vector must outlive the async copy.
  FL_HIP(hipStreamSynchronize(c->stream));
  c->omega_n = numIters;
  c->omega_relax = c->prm.relaxation;
  return FL_OK;
}

// solveSingleGrid (fluid.go:157-186) with the early exit after the first sweep whose
// maxDiv < 1e-5: the sweeps run speculatively; if one before the last met the
// tolerance, the state is rolled back and exactly that many sweeps are re-run.
int pressure_solve(Ctx* c, int numIters, float dt) {
  c->last_n = numIters;
  if (numIters <= 0) return FL_OK;
  int rc = ensure_omega(c, numIters);
  if (rc) return rc;
  const float cp = c->density * c->h / dt;
  bool use_wave = c->prm.solver == 2 || (c->prm.solver == 0 && gs_wave_supported(c, numIters));
  if (use_wave) return gs_wave_solve(c, numIters, cp);

  c->host_sweeps = -1;
  const size_t bytes = sizeof(float) * c->N;
  if (c->prm.early_exit) {
    FL_HIP(hipMemcpyAsync(c->U2, c->U, bytes, hipMemcpyDeviceToDevice, c->stream));
    FL_HIP(hipMemcpyAsync(c->V2, c->V, bytes, hipMemcpyDeviceToDevice, c->stream));
    FL_HIP(hipMemcpyAsync(c->T3, c->p, bytes, hipMemcpyDeviceToDevice, c->stream));
  }
  rc = hyperplane_run(c, numIters, cp, c->d_maxdiv);
  if (rc) return rc;
  int kept = numIters;
  if (c->prm.early_exit) {
    std::vector<uint32_t> md(numIters);
    FL_HIP(hipMemcpyAsync(md.data(), c->d_maxdiv, 4 * numIters, hipMemcpyDeviceToHost, c->stream));
    FL_HIP(hipStreamSynchronize(c->stream));
    for (int k = 0; k < numIters; k++) {
      float v;
      std::memcpy(&v, &md[k], 4);
      if (v < 1e-5f) {
        kept = k + 1;
        break;
      }
    }
    if (kept < numIters) {
      FL_HIP(hipMemcpyAsync(c->U, c->U2, bytes, hipMemcpyDeviceToDevice, c->stream));
      FL_HIP(hipMemcpyAsync(c->V, c->V2, bytes, hipMemcpyDeviceToDevice, c->stream));
      FL_HIP(hipMemcpyAsync(c->p, c->T3, bytes, hipMemcpyDeviceToDevice, c->stream));
      rc = hyperplane_run(c, kept, cp, c->d_maxdiv + c->nmax);
      if (rc) return rc;
    }
  }
  c->host_sweeps = kept;
  return FL_OK;
}

}  // namespace fl
