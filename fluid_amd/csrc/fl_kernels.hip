// Step kernels of libfluidhip for gfx950: every pass of Simulate except the pressure
// solve, the edit/injection kernels and the diagnostic views.
//
// Mapping: 64x4 workgroups, threadIdx.x along j (the contiguous axis of the i*NY+j
// layout) so each wavefront reads and writes 64 consecutive floats of one row.
// The bilinear gathers hit rows i0, i0+1 at two j's each; at the back-trace lengths
// of the workloads (<= a few cells) these stay in L1/L2, so the passes run at the
// stream rate of their field reads and writes. Arithmetic follows the reference
// operation by operation (build with -ffp-contract=off).
#include <hip/hip_runtime.h>

#include "fl_internal.h"

namespace fl {

static constexpr int BX = 64, BY = 4;

static inline dim3 grid2d(int NX, int NY) { return dim3((NY + BX - 1) / BX, (NX + BY - 1) / BY); }

#define CELL_PROLOGUE(g)                             \
  const int j = blockIdx.x * BX + threadIdx.x;       \
  const int i = blockIdx.y * BY + threadIdx.y;       \
  if (i >= (g).NX || j >= (g).NY) return;            \
  const int n = (g).NY;                              \
  const int c = i * n + j;                           \
  (void)n;

// ---------------------------------------------------------------------------------------
__global__ void k_fill(float* __restrict__ d, float v, size_t n) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; t < n; t += stride) d[t] = v;
}
__global__ void k_copy(float* __restrict__ d, const float* __restrict__ s, size_t n) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; t < n; t += stride) d[t] = s[t];
}
__global__ void k_fill_u8(uint8_t* __restrict__ d, size_t n) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; t < n; t += stride) d[t] = 0;
}

int launch_fill(float* dst, float v, size_t n, hipStream_t s) {
  unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
  if (blocks == 0) return FL_OK;
  hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(256), 0, s, dst, v, n);
  FL_HIP(hipGetLastError());
  return FL_OK;
}
int launch_copy(float* dst, const float* src, size_t n, hipStream_t s) {
  unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
  if (blocks == 0) return FL_OK;
  hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, s, dst, src, n);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// ---- cell flags (from S, which holds exactly 0.0 or 1.0) ------------------------------
__global__ void k_flags(Grid g, const float* __restrict__ S, uint8_t* __restrict__ flags,
                        uint8_t* __restrict__ diag) {
  CELL_PROLOGUE(g);
  const int NX = g.NX, NY = g.NY;
  uint32_t f = 0;
  bool fluid = S[c] != 0.0f;
  if (fluid) f |= CF_FLUID;
  if (i > 0 && S[c - n] != 0.0f) f |= CF_X0;
  if (i < NX - 1 && S[c + n] != 0.0f) f |= CF_X1;
  if (j > 0 && S[c - 1] != 0.0f) f |= CF_Y0;
  if (j < NY - 1 && S[c + 1] != 0.0f) f |= CF_Y1;
  bool interior = i >= 1 && i <= NX - 2 && j >= 1 && j <= NY - 2;
  if (interior) f |= CF_INTERIOR;
  if (interior && fluid && (f & (CF_X0 | CF_X1 | CF_Y0 | CF_Y1))) f |= CF_GS;
  flags[c] = (uint8_t)f;
  diag[(size_t)(i + j) * NX + i] = (uint8_t)f;
}

int launch_flags(Ctx* c) {
  size_t nd = (size_t)(c->NX + c->NY - 1) * c->NX;
  hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)std::min<size_t>((nd + 255) / 256, 8192)), dim3(256), 0, c->stream,
                     c->flags_diag, nd);
  hipLaunchKernelGGL(k_flags, grid2d(c->NX, c->NY), dim3(BX, BY), 0, c->stream, c->grid, c->S, c->flags,
                     c->flags_diag);
  FL_HIP(hipGetLastError());
  c->flags_dirty = false;
  return FL_OK;
}

// ---- copyBorder (fluid.go:436-446) ------------------------------------------------------
__global__ void k_copy_border(Grid g, float* __restrict__ dst, const float* __restrict__ src) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = g.NY;
  if (t < g.NX) {
    dst[t * n + 0] = src[t * n + 0];
    dst[t * n + g.NY - 1] = src[t * n + g.NY - 1];
  }
  if (t < g.NY) {
    dst[t] = src[t];
    dst[(g.NX - 1) * n + t] = src[(g.NX - 1) * n + t];
  }
}
int launch_copy_border(Ctx* c, float* dst, const float* src) {
  int m = std::max(c->NX, c->NY);
  hipLaunchKernelGGL(k_copy_border, dim3((m + 255) / 256), dim3(256), 0, c->stream, c->grid, dst, src);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// ---- applyViscosity (fluid.go:112-142): result lands in both newU/newV and U2/V2 ------
__global__ void k_viscosity(Grid g, const uint8_t* __restrict__ flags, const float* __restrict__ U,
                            const float* __restrict__ V, float* __restrict__ nU, float* __restrict__ nV,
                            float* __restrict__ Uo, float* __restrict__ Vo, float visc) {
  CELL_PROLOGUE(g);
  float u = U[c], v = V[c];
  float ru = u, rv = v;
  uint32_t f = flags[c];
  if ((f & CF_INTERIOR) && (f & CF_FLUID)) {
    float uLap = U[c - n] + U[c + n] + U[c - 1] + U[c + 1] - 4.0f * u;
    ru = u + visc * uLap;
    float vLap = V[c - n] + V[c + n] + V[c - 1] + V[c + 1] - 4.0f * v;
    rv = v + visc * vLap;
  }
  nU[c] = ru;
  Uo[c] = ru;
  nV[c] = rv;
  Vo[c] = rv;
}
int launch_viscosity(Ctx* c, float dt) {
  float visc = c->prm.viscosity_diffusion * dt;
  hipLaunchKernelGGL(k_viscosity, grid2d(c->NX, c->NY), dim3(BX, BY), 0, c->stream, c->grid, c->flags, c->U, c->V,
                     c->nU, c->nV, c->U2, c->V2, visc);
  FL_HIP(hipGetLastError());
  std::swap(c->U, c->U2);
  std::swap(c->V, c->V2);
  return FL_OK;
}

// ---- vorticity confinement (fluid.go:449-493) + turbulence (fluid.go:496-526) ---------
__device__ __forceinline__ float curl_at(const float* __restrict__ U, const float* __restrict__ V, int c, int n,
                                         float h) {
  float dvdx = (V[c + n] - V[c - n]) * 0.5f / h;
  float dudy = (U[c + 1] - U[c - 1]) * 0.5f / h;
  return dvdx - dudy;
}

__global__ void k_curl(Grid g, const uint8_t* __restrict__ flags, const float* __restrict__ U,
                       const float* __restrict__ V, float* __restrict__ curl) {
  CELL_PROLOGUE(g);
  uint32_t f = flags[c];
  curl[c] = ((f & CF_INTERIOR) && (f & CF_FLUID)) ? curl_at(U, V, c, n, g.h) : 0.0f;
}
int launch_curl(Ctx* c, float* curl) {
  hipLaunchKernelGGL(k_curl, grid2d(c->NX, c->NY), dim3(BX, BY), 0, c->stream, c->grid, c->flags, c->U, c->V,
                     curl);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

__global__ void k_conf_turb(Grid g, const uint8_t* __restrict__ flags, float* __restrict__ U, float* __restrict__ V,
                            const float* __restrict__ curl, float conf, float dt, float turb, int do_conf,
                            int do_turb) {
  CELL_PROLOGUE(g);
  uint32_t f = flags[c];
  if (!((f & CF_INTERIOR) && (f & CF_FLUID))) return;
  float u = U[c], v = V[c];
  if (do_conf) {
    float gx = (fabsf(curl[c + n]) - fabsf(curl[c - n])) * 0.5f / g.h;
    float gy = (fabsf(curl[c + 1]) - fabsf(curl[c - 1])) * 0.5f / g.h;
    float mag = sqrtf(gx * gx + gy * gy) + 1e-5f;
    gx /= mag;
    gy /= mag;
    float vort = curl[c];
    float localVel = sqrtf(u * u + v * v);
    float strength = conf * (1.0f + localVel * 0.1f);
    u += strength * gy * vort * dt;
    v -= strength * gx * vort * dt;
  }
  if (do_turb) {
    float noiseU = turb_noise(i * 137 + j * 241) * turb;
    float noiseV = turb_noise(i * 157 + j * 263) * turb;
    float localVel = sqrtf(u * u + v * v);
    if (localVel > 0.1f) {
      float factor = go_minf(localVel * 0.5f, 1.0f);
      u += noiseU * factor;
      v += noiseV * factor;
    }
  }
  U[c] = u;
  V[c] = v;
}
int launch_conf_turb(Ctx* c, const float* curl, float dt, int do_conf, int do_turb) {
  float turb = c->prm.turbulence_strength * dt;
  hipLaunchKernelGGL(k_conf_turb, grid2d(c->NX, c->NY), dim3(BX, BY), 0, c->stream, c->grid, c->flags, c->U, c->V,
                     curl, c->prm.confinement, dt, turb, do_conf, do_turb);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// ---- handleBorders (fluid.go:236-289) -----------------------------------------------------
__global__ void k_borders(Grid g, const float* __restrict__ S, float* __restrict__ U, float* __restrict__ V) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = g.NY, NX = g.NX, NY = g.NY;
  if (t < NX) {
    int i = t;
    if (S[i * n + 0] == 0.0f || S[i * n + 1] == 0.0f) {
      U[i * n + 0] = 0.0f;
    } else if (i > 0 && i < NX - 1 && S[i * n + 2] > 0.0f) {
      U[i * n + 0] = 2.0f * U[i * n + 1] - U[i * n + 2];
    } else {
      U[i * n + 0] = U[i * n + 1];
    }
    if (S[i * n + NY - 1] == 0.0f || S[i * n + NY - 2] == 0.0f) {
      U[i * n + NY - 1] = 0.0f;
    } else if (i > 0 && i < NX - 1 && S[i * n + NY - 3] > 0.0f) {
      U[i * n + NY - 1] = 2.0f * U[i * n + NY - 2] - U[i * n + NY - 3];
    } else {
      U[i * n + NY - 1] = U[i * n + NY - 2];
    }
  }
  if (t < NY) {
    int j = t;
    if (S[j] == 0.0f || S[n + j] == 0.0f) {
      V[j] = 0.0f;
    } else if (j > 0 && j < NY - 1 && S[2 * n + j] > 0.0f) {
      V[j] = 2.0f * V[n + j] - V[2 * n + j];
    } else {
      V[j] = V[n + j];
    }
    if (S[(NX - 1) * n + j] == 0.0f || S[(NX - 2) * n + j] == 0.0f) {
      V[(NX - 1) * n + j] = 0.0f;
    } else if (j > 0 && j < NY - 1 && S[(NX - 3) * n + j] > 0.0f) {
      V[(NX - 1) * n + j] = 2.0f * V[(NX - 2) * n + j] - V[(NX - 3) * n + j];
    } else {
      V[(NX - 1) * n + j] = V[(NX - 2) * n + j];
    }
  }
}
int launch_borders(Ctx* c) {
  if (c->NX < 3 || c->NY < 3) return FL_ERR_UNSUPPORTED;
  int m = std::max(c->NX, c->NY);
  hipLaunchKernelGGL(k_borders, dim3((m + 255) / 256), dim3(256), 0, c->stream, c->grid, c->S, c->U, c->V);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// ---- semi-Lagrangian advection (fluid.go:291-434) ----------------------------------------
__device__ __forceinline__ float avg_u(const float* __restrict__ U, int c, int n) {
  return (U[c - 1] + U[c] + U[c + n - 1] + U[c + n]) * 0.25f;
}
__device__ __forceinline__ float avg_v(const float* __restrict__ V, int c, int n) {
  return (V[c - n] + V[c] + V[c - n + 1] + V[c + 1]) * 0.25f;
}

// advectVelocity: copyBorder(newU,U), the advected faces, stale newU elsewhere; then U=newU.
// Ut/Vt is the current velocity (traced and sampled); result goes to newU/newV and Uo/Vo.
__global__ void k_advect_velocity(Grid g, const uint8_t* __restrict__ flags, const float* __restrict__ Ut,
                                  const float* __restrict__ Vt, float* __restrict__ nU, float* __restrict__ nV,
                                  float* __restrict__ Uo, float* __restrict__ Vo, float dt, int* __restrict__ bad) {
  CELL_PROLOGUE(g);
  const int NX = g.NX, NY = g.NY;
  const bool ring = i == 0 || i == NX - 1 || j == 0 || j == NY - 1;
  const uint32_t f = flags[c];
  int badl = 0;
  float ou, ov;
  if (i >= 1 && j >= 1 && j <= NY - 2 && (f & CF_FLUID) && (f & CF_X0)) {
    float x = (float)i * g.h;
    float y = (float)j * g.h + g.ha;
    float u = Ut[c];
    float v = avg_v(Vt, c, n);
    x = x - dt * u;
    y = y - dt * v;
    ou = sample_field(x, y, Ut, 0.0f, g.hs, g, &badl);
  } else {
    ou = ring ? Ut[c] : nU[c];
  }
  if (j >= 1 && i >= 1 && i <= NX - 2 && (f & CF_FLUID) && (f & CF_Y0)) {
    float x = (float)i * g.h + g.ha;
    float y = (float)j * g.h;
    float u = avg_u(Ut, c, n);
    float v = Vt[c];
    x = x - dt * u;
    y = y - dt * v;
    ov = sample_field(x, y, Vt, g.hs, 0.0f, g, &badl);
  } else {
    ov = ring ? Vt[c] : nV[c];
  }
  nU[c] = ou;
  Uo[c] = ou;
  nV[c] = ov;
  Vo[c] = ov;
  if (badl) atomicOr(bad, 1);
}
int launch_advect_velocity(Ctx* c, const float* Ut, const float* Vt, float* Uo, float* Vo, float dt) {
  hipLaunchKernelGGL(k_advect_velocity, grid2d(c->NX, c->NY), dim3(BX, BY), 0, c->stream, c->grid, c->flags, Ut,
                     Vt, c->nU, c->nV, Uo, Vo, dt, c->d_ctrl + 2);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// clampToNeighbors (fluid.go:1094-1120) for an interior cell (the 3x3 block is in range).
__device__ __forceinline__ float clamp_neighbors(float val, const float* __restrict__ src, int c, int n) {
  float lo = src[c], hi = src[c];
#pragma unroll
  for (int di = -1; di <= 1; di++)
#pragma unroll
    for (int dj = -1; dj <= 1; dj++) {
      float v = src[c + di * n + dj];
      if (v < lo) lo = v;
      if (v > hi) hi = v;
    }
  if (val < lo) return lo;
  if (val > hi) return hi;
  return val;
}

// advectVelocityBFECC passes 2 and 3 (fluid.go:935-987): backward trace with the
// original velocity sampling the forward result, then correction and clamp.
__global__ void k_bfecc_velocity_corr(Grid g, const uint8_t* __restrict__ flags, const float* __restrict__ Uo,
                                      const float* __restrict__ Vo, const float* __restrict__ Uf,
                                      const float* __restrict__ Vf, float* __restrict__ Uc, float* __restrict__ Vc,
                                      float dt, int* __restrict__ bad) {
  CELL_PROLOGUE(g);
  const int NX = g.NX, NY = g.NY;
  const bool ring = i == 0 || i == NX - 1 || j == 0 || j == NY - 1;
  const uint32_t f = flags[c];
  int badl = 0;
  float bu, bv;
  if (i >= 1 && j >= 1 && j <= NY - 2 && (f & CF_FLUID) && (f & CF_X0)) {
    float x = (float)i * g.h;
    float y = (float)j * g.h + g.ha;
    float u = Uo[c];
    float v = avg_v(Vo, c, n);
    x = x + dt * u;
    y = y + dt * v;
    bu = sample_field(x, y, Uf, 0.0f, g.hs, g, &badl);
  } else {
    bu = ring ? Uf[c] : 0.0f;
  }
  if (j >= 1 && i >= 1 && i <= NX - 2 && (f & CF_FLUID) && (f & CF_Y0)) {
    float x = (float)i * g.h + g.ha;
    float y = (float)j * g.h;
    float u = avg_u(Uo, c, n);
    float v = Vo[c];
    x = x + dt * u;
    y = y + dt * v;
    bv = sample_field(x, y, Vf, g.hs, 0.0f, g, &badl);
  } else {
    bv = ring ? Vf[c] : 0.0f;
  }
  float ou = Uo[c], ov = Vo[c];
  float cu = ou, cv = ov;
  if (f & CF_INTERIOR) {
    cu = ou - (bu - ou) * 0.5f;
    cv = ov - (bv - ov) * 0.5f;
    cu = clamp_neighbors(cu, Uo, c, n);
    cv = clamp_neighbors(cv, Vo, c, n);
  }
  Uc[c] = cu;
  Vc[c] = cv;
  if (badl) atomicOr(bad, 1);
}
int launch_bfecc_velocity_corr(Ctx* c, const float* Uorig, const float* Vorig, const float* Ufwd, const float* Vfwd,
                               float* Ucorr, float* Vcorr, float dt) {
  hipLaunchKernelGGL(k_bfecc_velocity_corr, grid2d(c->NX, c->NY), dim3(BX, BY), 0, c->stream, c->grid, c->flags,
                     Uorig, Vorig, Ufwd, Vfwd, Ucorr, Vcorr, dt, c->d_ctrl + 2);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// advectSmoke (fluid.go:400-434) using the current (already advected) velocity.
__global__ void k_advect_smoke(Grid g, const uint8_t* __restrict__ flags, const float* __restrict__ U,
                               const float* __restrict__ V, const float* __restrict__ Ms, float* __restrict__ nM,
                               float* __restrict__ Mo, float dt, float smokeAdv, float visc,
                               int* __restrict__ bad) {
  CELL_PROLOGUE(g);
  const int NX = g.NX, NY = g.NY;
  const bool ring = i == 0 || i == NX - 1 || j == 0 || j == NY - 1;
  const uint32_t f = flags[c];
  float out;
  if ((f & CF_INTERIOR) && (f & CF_FLUID)) {
    int badl = 0;
    float u = (U[c] + U[c + n]) * 0.5f * smokeAdv;
    float v = (V[c] + V[c + 1]) * 0.5f * smokeAdv;
    float x = (float)i * g.h + g.hs - dt * u;
    float y = (float)j * g.h + g.hs - dt * v;
    float val = sample_field(x, y, Ms, g.hs, g.hs, g, &badl);
    if (visc > 0.0f) {
      float sd = visc * 0.3f * dt;
      float nb = Ms[c - n] + Ms[c + n] + Ms[c - 1] + Ms[c + 1] - 4.0f * Ms[c];
      val += sd * nb;
    }
    out = go_maxf(val, 0.0f);
    if (badl) atomicOr(bad, 1);
  } else {
    out = ring ? Ms[c] : nM[c];
  }
  nM[c] = out;
  Mo[c] = out;
}
int launch_advect_smoke(Ctx* c, const float* Msrc, float* Mo, float dt) {
  hipLaunchKernelGGL(k_advect_smoke, grid2d(c->NX, c->NY), dim3(BX, BY), 0, c->stream, c->grid, c->flags, c->U,
                     c->V, Msrc, c->nM, Mo, dt, c->prm.smoke_advection, c->prm.viscosity_diffusion, c->d_ctrl + 2);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// advectSmokeBFECC passes 2 and 3 (fluid.go:1013-1046).
__global__ void k_bfecc_smoke_corr(Grid g, const uint8_t* __restrict__ flags, const float* __restrict__ U,
                                   const float* __restrict__ V, const float* __restrict__ Mo,
                                   const float* __restrict__ Mf, float* __restrict__ Mc, float dt, float smokeAdv,
                                   int* __restrict__ bad) {
  CELL_PROLOGUE(g);
  const int NX = g.NX, NY = g.NY;
  const bool ring = i == 0 || i == NX - 1 || j == 0 || j == NY - 1;
  const uint32_t f = flags[c];
  float om = Mo[c];
  float cm = om;
  if (!ring) {
    float bm = 0.0f;
    if (f & CF_FLUID) {
      int badl = 0;
      float u = (U[c] + U[c + n]) * 0.5f * smokeAdv;
      float v = (V[c] + V[c + 1]) * 0.5f * smokeAdv;
      float x = (float)i * g.h + g.hs + dt * u;
      float y = (float)j * g.h + g.hs + dt * v;
      bm = sample_field(x, y, Mf, g.hs, g.hs, g, &badl);
      if (badl) atomicOr(bad, 1);
    }
    cm = om - (bm - om) * 0.5f;
    cm = clamp_neighbors(cm, Mo, c, n);
    if (cm < 0.0f) cm = 0.0f;
  }
  Mc[c] = cm;
}
int launch_bfecc_smoke_corr(Ctx* c, const float* Morig, const float* Mfwd, float* Mcorr, float dt) {
  hipLaunchKernelGGL(k_bfecc_smoke_corr, grid2d(c->NX, c->NY), dim3(BX, BY), 0, c->stream, c->grid, c->flags, c->U,
                     c->V, Morig, Mfwd, Mcorr, dt, c->prm.smoke_advection, c->d_ctrl + 2);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// ---- edits ------------------------------------------------------------------------------
struct FieldPtrs {
  float* f[8];
};

// One thread per target word; its ops are applied in call order.
__global__ void k_apply_edits(FieldPtrs fp, uint32_t ncells, const uint32_t* __restrict__ keys,
                              const float* __restrict__ vals, const int32_t* __restrict__ seg, int nseg) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  int b = seg[s], e = seg[s + 1];
  uint32_t key = keys[b] & 0x7fffffffu;
  uint32_t field = key / ncells, idx = key % ncells;
  float* dst = fp.f[field];
  float x = dst[idx];
  for (int k = b; k < e; k++) {
    bool add = (keys[k] >> 31) != 0;
    x = add ? x + vals[k] : vals[k];
  }
  dst[idx] = x;
}
int launch_apply_edits(Ctx* c, int nseg) {
  FieldPtrs fp;
  fp.f[FL_FIELD_U] = c->U;
  fp.f[FL_FIELD_V] = c->V;
  fp.f[FL_FIELD_NEWU] = c->nU;
  fp.f[FL_FIELD_NEWV] = c->nV;
  fp.f[FL_FIELD_P] = c->p;
  fp.f[FL_FIELD_S] = c->S;
  fp.f[FL_FIELD_M] = c->M;
  fp.f[FL_FIELD_NEWM] = c->nM;
  hipLaunchKernelGGL(k_apply_edits, dim3((nseg + 255) / 256), dim3(256), 0, c->stream, fp, (uint32_t)c->N,
                     c->d_edit_keys, c->d_edit_vals, c->d_edit_seg, nseg);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// Per-step sources: one thread per cell, ops in reference order, IsSolid read at apply time.
__global__ void k_apply_sources(const SrcOp* __restrict__ ops, const int32_t* __restrict__ seg, int nseg,
                                const float* __restrict__ S, float* __restrict__ U, float* __restrict__ V,
                                float* __restrict__ M) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  int b = seg[s], e = seg[s + 1];
  int cell = ops[b].cell;
  float u = U[cell], v = V[cell], m = M[cell];
  bool fluid = S[cell] != 0.0f;
  for (int k = b; k < e; k++) {
    SrcOp op = ops[k];
    if (op.check_solid && !fluid) continue;
    if (op.kind == 0) {
      u = op.a;
      v = op.b;
    } else {
      m = m + op.a;
    }
  }
  U[cell] = u;
  V[cell] = v;
  M[cell] = m;
}
int launch_apply_sources(Ctx* c) {
  if (c->n_src_seg == 0) return FL_OK;
  hipLaunchKernelGGL(k_apply_sources, dim3((c->n_src_seg + 255) / 256), dim3(256), 0, c->stream, c->d_src,
                     c->d_src_seg, c->n_src_seg, c->S, c->U, c->V, c->M);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// ---- reductions and views (fluid.go:529-557, 799-891) -----------------------------------
__device__ __forceinline__ float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

// kind 0: MaxDivergence over interior S!=0; kind 1: max(|U|+|V|) over interior S>0.
// All candidates are >= +0, so their float bits order like unsigned integers.
__global__ void k_max_reduce(Grid g, const uint8_t* __restrict__ flags, const float* __restrict__ U,
                             const float* __restrict__ V, int kind, uint32_t* __restrict__ out) {
  const int j = blockIdx.x * BX + threadIdx.x;
  const int i = blockIdx.y * BY + threadIdx.y;
  float m = 0.0f;
  if (i < g.NX && j < g.NY) {
    const int n = g.NY, c = i * n + j;
    uint32_t f = flags[c];
    if ((f & CF_INTERIOR) && (f & CF_FLUID)) {
      float a;
      if (kind == 0)
        a = fabsf(U[c + n] - U[c] + V[c + 1] - V[c]);
      else
        a = fabsf(U[c]) + fabsf(V[c]);
      if (a > m) m = a;
    }
  }
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0 && m > 0.0f) atomicMax(out, __float_as_uint(m));
}
int launch_max_reduce(Ctx* c, int kind, float* out) {
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, c->stream, (float*)c->d_red, 0.0f, (size_t)1);
  hipLaunchKernelGGL(k_max_reduce, grid2d(c->NX, c->NY), dim3(BX, BY), 0, c->stream, c->grid, c->flags, c->U, c->V,
                     kind, c->d_red);
  FL_HIP(hipGetLastError());
  uint32_t bits = 0;
  FL_HIP(hipMemcpyAsync(&bits, c->d_red, 4, hipMemcpyDeviceToHost, c->stream));
  FL_HIP(hipStreamSynchronize(c->stream));
  *out = __builtin_bit_cast(float, bits);
  return FL_OK;
}

__global__ void k_view(Grid g, const uint8_t* __restrict__ flags, const float* __restrict__ U,
                       const float* __restrict__ V, int kind, float* __restrict__ out) {
  CELL_PROLOGUE(g);
  uint32_t f = flags[c];
  float r = 0.0f;
  if ((f & CF_INTERIOR) && (f & CF_FLUID)) {
    if (kind == 0) {
      r = curl_at(U, V, c, n, g.h);
    } else {
      float u = (U[c] + U[c + n]) * 0.5f;
      float v = (V[c] + V[c + 1]) * 0.5f;
      r = sqrtf(u * u + v * v);
    }
  }
  out[c] = r;
}
int launch_view(Ctx* c, int kind, float* out) {
  hipLaunchKernelGGL(k_view, grid2d(c->NX, c->NY), dim3(BX, BY), 0, c->stream, c->grid, c->flags, c->U, c->V, kind,
                     out);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

__global__ void k_sample_batch(Grid g, const float* __restrict__ U, const float* __restrict__ V,
                               const float* __restrict__ xy, int n, float* __restrict__ uv, int* __restrict__ bad) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int badl = 0;
  float x = xy[2 * t], y = xy[2 * t + 1];
  uv[2 * t] = sample_field(x, y, U, 0.0f, g.hs, g, &badl);
  uv[2 * t + 1] = sample_field(x, y, V, g.hs, 0.0f, g, &badl);
  if (badl) atomicOr(bad, 1);
}
int launch_sample_batch(Ctx* c, const float* xy, int n, float* uv) {
  hipLaunchKernelGGL(k_sample_batch, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->grid, c->U, c->V, xy, n, uv,
                     c->d_ctrl + 2);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

}  // namespace fl
