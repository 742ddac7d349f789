// Step kernels of libfluidhip for gfx950: every pass of Simulate except the pressure
// solve, the edit/injection kernels and the diagnostic views.
//
// Mapping: 64x4 workgroups, threadIdx.x along j (the contiguous axis of the i*NY+j
// layout) so each wavefront reads and writes 64 consecutive floats of one row; the blocks
// are dealt to the XCDs in contiguous row bands (xcd_tile).
// The bilinear gathers hit rows i0, i0+1 at two j's each; at the back-trace lengths
// of the workloads (<= a few cells) these stay in L1/L2, so the passes run at the
// stream rate of their field reads and writes. Arithmetic follows the reference
// operation by operation (build with -ffp-contract=off).
//
// Every pass kernel is templated on its launch map: OneSim (one simulation passed by
// value, grid of 64x4 blocks) or ManySims (a batch: each workgroup takes a (simulation,
// block) item from a work-list), and addresses its fields by role (SimDesc::buf), so a
// batch of uneven grids runs each pass as one launch.
#include <hip/hip_runtime.h>

#include "fl_internal.h"

namespace fl {

static constexpr int BX = 64, BY = 4;

static inline dim3 grid2d(int NX, int NY) { return dim3((NY + BX - 1) / BX, (NX + BY - 1) / BY); }

// ---- launch maps --------------------------------------------------------------------------
__device__ __forceinline__ bool cell_of(const OneSim& m, const SimDesc*& d, int& i, int& j) {
  d = &m.d;
  const int gx = gridDim.x;
  const int t = xcd_tile(blockIdx.y * gx + blockIdx.x, gx * gridDim.y);
  const int by = t / gx, bx = t - by * gx;
  j = bx * BX + threadIdx.x;
  i = by * BY + threadIdx.y;
  return i < d->g.NX && j < d->g.NY;
}
__device__ __forceinline__ bool cell_of(const ManySims& m, const SimDesc*& d, int& i, int& j) {
  const CellItem it = m.cells[blockIdx.x];
  d = &m.sims[it.sim];
  j = it.bx * BX + threadIdx.x;
  i = it.by * BY + threadIdx.y;
  return i < d->g.NX && j < d->g.NY;
}
// Ring kernels: thread t of a simulation handles row i = t and column j = t.
__device__ __forceinline__ void ring_of(const OneSim& m, const SimDesc*& d, int& t) {
  d = &m.d;
  t = blockIdx.x * blockDim.x + threadIdx.x;
}
__device__ __forceinline__ void ring_of(const ManySims& m, const SimDesc*& d, int& t) {
  const RingItem it = m.rings[blockIdx.x];
  d = &m.sims[it.sim];
  t = it.blk * blockDim.x + threadIdx.x;
}

template <class L> static void launch_cells(Ctx* c, const L& fn) {
  OneSim m{c->desc()};
  fn(m, grid2d(c->NX, c->NY), dim3(BX, BY), c->stream);
}

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
int launch_zero_u8(uint8_t* dst, size_t n, hipStream_t s) {
  unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
  if (blocks == 0) return FL_OK;
  hipLaunchKernelGGL(k_fill_u8, dim3(blocks), dim3(256), 0, s, dst, n);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// Per-simulation fill and copy of a field role (batched steps copy results back instead of
// swapping buffers, so a batch's descriptors stay fixed).
template <class M> __global__ void k_fill_role(M m, int r, float v) {
  const SimDesc* d;
  int i, j;
  if (!cell_of(m, d, i, j)) return;
  d->buf[r][i * d->g.NY + j] = v;
}
template <class M> __global__ void k_copy_role(M m, int rdst, int rsrc) {
  const SimDesc* d;
  int i, j;
  if (!cell_of(m, d, i, j)) return;
  const int c = i * d->g.NY + j;
  d->buf[rdst][c] = d->buf[rsrc][c];
}

// ---- cell flags (from S, which holds exactly 0.0 or 1.0) ------------------------------
// The anti-diagonal copy is written only at grid cells; its other entries stay zero from
// allocation.
template <class M> __global__ void k_flags(M m) {
  const SimDesc* d;
  int i, j;
  if (!cell_of(m, d, i, j)) return;
  const int NX = d->g.NX, NY = d->g.NY, n = NY, c = i * n + j;
  const float* S = d->buf[B_S];
  uint32_t f = 0;
  const bool fluid = S[c] != 0.0f;
  if (fluid) f |= CF_FLUID;
  if (i > 0 && S[c - n] != 0.0f) f |= CF_X0;
  if (i < NX - 1 && S[c + n] != 0.0f) f |= CF_X1;
  if (j > 0 && S[c - 1] != 0.0f) f |= CF_Y0;
  if (j < NY - 1 && S[c + 1] != 0.0f) f |= CF_Y1;
  const bool interior = i >= 1 && i <= NX - 2 && j >= 1 && j <= NY - 2;
  if (interior) f |= CF_INTERIOR;
  if (interior && fluid && (f & (CF_X0 | CF_X1 | CF_Y0 | CF_Y1))) f |= CF_GS;
  d->flags[c] = (uint8_t)f;
  d->diag[(size_t)(i + j) * NX + i] = (uint8_t)f;
}

int launch_flags(Ctx* c) {
  launch_cells(c, [](const OneSim& m, dim3 g, dim3 b, hipStream_t s) { hipLaunchKernelGGL(k_flags<OneSim>, g, b, 0, s, m); });
  FL_HIP(hipGetLastError());
  c->flags_dirty = false;
  return FL_OK;
}

// ---- copyBorder (fluid.go:436-446) ------------------------------------------------------
template <class M> __global__ void k_copy_border(M m, int rdst, int rsrc) {
  const SimDesc* d;
  int t;
  ring_of(m, d, t);
  const int n = d->g.NY, NX = d->g.NX, NY = d->g.NY;
  float* dst = d->buf[rdst];
  const float* src = d->buf[rsrc];
  if (t < NX) {
    dst[t * n + 0] = src[t * n + 0];
    dst[t * n + NY - 1] = src[t * n + NY - 1];
  }
  if (t < NY) {
    dst[t] = src[t];
    dst[(NX - 1) * n + t] = src[(NX - 1) * n + t];
  }
}
static int role_of(Ctx* c, const float* p) {
  const SimDesc d = c->desc();
  for (int k = 0; k < NBUF; k++)
    if (d.buf[k] == p) return k;
  return -1;
}
int launch_copy_border(Ctx* c, float* dst, const float* src) {
  const int m = std::max(c->NX, c->NY);
  OneSim map{c->desc()};
  hipLaunchKernelGGL(k_copy_border<OneSim>, dim3((m + 255) / 256), dim3(256), 0, c->stream, map, role_of(c, dst),
                     role_of(c, src));
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// ---- applyViscosity (fluid.go:112-142): result lands in both newU/newV and U2/V2 ------
template <class M> __global__ void k_viscosity(M m, float visc) {
  const SimDesc* d;
  int i, j;
  if (!cell_of(m, d, i, j)) return;
  const int n = d->g.NY, c = i * n + j;
  const float* U = d->buf[B_U];
  const float* V = d->buf[B_V];
  float u = U[c], v = V[c];
  float ru = u, rv = v;
  const uint32_t f = d->flags[c];
  if ((f & CF_INTERIOR) && (f & CF_FLUID)) {
    float uLap = U[c - n] + U[c + n] + U[c - 1] + U[c + 1] - 4.0f * u;
    ru = u + visc * uLap;
    float vLap = V[c - n] + V[c + n] + V[c - 1] + V[c + 1] - 4.0f * v;
    rv = v + visc * vLap;
  }
  d->buf[B_NU][c] = ru;
  d->buf[B_U2][c] = ru;
  d->buf[B_NV][c] = rv;
  d->buf[B_V2][c] = rv;
}
int launch_viscosity(Ctx* c, float dt) {
  const float visc = c->prm.viscosity_diffusion * dt;
  launch_cells(c, [&](const OneSim& m, dim3 g, dim3 b, hipStream_t s) {
    hipLaunchKernelGGL(k_viscosity<OneSim>, g, b, 0, s, m, visc);
  });
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

// (rU, rV): buffer roles holding the current velocity (U, V for a context; the solve's outputs
// in a batched step).
template <class M> __global__ void k_curl(M m, int rcurl, int rU, int rV) {
  const SimDesc* d;
  int i, j;
  if (!cell_of(m, d, i, j)) return;
  const int n = d->g.NY, c = i * n + j;
  const uint32_t f = d->flags[c];
  d->buf[rcurl][c] =
      ((f & CF_INTERIOR) && (f & CF_FLUID)) ? curl_at(d->buf[rU], d->buf[rV], c, n, d->g.h) : 0.0f;
}
int launch_curl(Ctx* c, float* curl) {
  const int r = role_of(c, curl);
  launch_cells(c, [&](const OneSim& m, dim3 g, dim3 b, hipStream_t s) {
    hipLaunchKernelGGL(k_curl<OneSim>, g, b, 0, s, m, r, (int)B_U, (int)B_V);
  });
  FL_HIP(hipGetLastError());
  return FL_OK;
}

template <class M>
__global__ void k_conf_turb(M m, int rU, int rV, int rcurl, float conf, float dt, float turb, int do_conf,
                            int do_turb) {
  const SimDesc* d;
  int i, j;
  if (!cell_of(m, d, i, j)) return;
  const int n = d->g.NY, c = i * n + j;
  const uint32_t f = d->flags[c];
  if (!((f & CF_INTERIOR) && (f & CF_FLUID))) return;
  float* U = d->buf[rU];
  float* V = d->buf[rV];
  const float* curl = d->buf[rcurl];
  const float h = d->g.h;
  float u = U[c], v = V[c];
  if (do_conf) {
    float gx = (fabsf(curl[c + n]) - fabsf(curl[c - n])) * 0.5f / h;
    float gy = (fabsf(curl[c + 1]) - fabsf(curl[c - 1])) * 0.5f / h;
    float mag = sqrtf(gx * gx + gy * gy) + 1e-5f;
    gx /= mag;
    gy /= mag;
    float vort = curl[c];
    float localVel = sqrtf(u * u + v * v);
    float strength = conf * (1.0f + localVel * 0.1f);
    u += strength * gy * vort * dt;
    v -= strength * gx * vort * dt;
  }
  // the update is in place: a face the passes leave alone is not written back, and the
  // noise (two float64 sines) is evaluated only where the speed gate lets it in
  bool changed = do_conf != 0;
  if (do_turb) {
    float localVel = sqrtf(u * u + v * v);
    if (localVel > 0.1f) {
      float noiseU = turb_noise(i * 137 + j * 241) * turb;
      float noiseV = turb_noise(i * 157 + j * 263) * turb;
      float factor = go_minf(localVel * 0.5f, 1.0f);
      u += noiseU * factor;
      v += noiseV * factor;
      changed = true;
    }
  }
  if (changed) {
    U[c] = u;
    V[c] = v;
  }
}
int launch_conf_turb(Ctx* c, const float* curl, float dt, int do_conf, int do_turb) {
  const float turb = c->prm.turbulence_strength * dt, conf = c->prm.confinement;
  const int r = role_of(c, curl);
  launch_cells(c, [&](const OneSim& m, dim3 g, dim3 b, hipStream_t s) {
    hipLaunchKernelGGL(k_conf_turb<OneSim>, g, b, 0, s, m, (int)B_U, (int)B_V, r, conf, dt, turb, do_conf, do_turb);
  });
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// ---- handleBorders (fluid.go:236-289) -----------------------------------------------------
template <class M> __global__ void k_borders(M m, int rU, int rV) {
  const SimDesc* d;
  int t;
  ring_of(m, d, t);
  const int n = d->g.NY, NX = d->g.NX, NY = d->g.NY;
  const float* S = d->buf[B_S];
  float* U = d->buf[rU];
  float* V = d->buf[rV];
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
  const int m = std::max(c->NX, c->NY);
  OneSim map{c->desc()};
  hipLaunchKernelGGL(k_borders<OneSim>, dim3((m + 255) / 256), dim3(256), 0, c->stream, map, (int)B_U, (int)B_V);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// ---- semi-Lagrangian advection (fluid.go:291-434) ----------------------------------------
__device__ __forceinline__ float avg_u(const float* __restrict__ U, int c, int n) {
  return (ld_cell(U, c - 1) + ld_cell(U, c) + ld_cell(U, c + n - 1) + ld_cell(U, c + n)) * 0.25f;
}
__device__ __forceinline__ float avg_v(const float* __restrict__ V, int c, int n) {
  return (ld_cell(V, c - n) + ld_cell(V, c) + ld_cell(V, c - n + 1) + ld_cell(V, c + 1)) * 0.25f;
}

// advectVelocity: copyBorder(newU,U), the advected faces, stale newU elsewhere; then U=newU.
// (Ut,Vt) is the current velocity (traced and sampled); the result goes to newU/newV and
// to (Uo,Vo).
// Every load is issued unconditionally, from addresses in range whatever the cell (neighbour
// reads of cells outside a face's geometric range use a base cell inside the grid; samples of
// faces that are not advected clamp like any other), so a wave waits for two rounds of loads
// (the face neighbourhood, then both bilinear samples) instead of one round per branch.
template <class M> __global__ void k_advect_velocity(M m, int rUt, int rVt, int rUo, int rVo, float dt, int write_new) {
  const SimDesc* d;
  int i, j;
  if (!cell_of(m, d, i, j)) return;
  const Grid g = d->g;
  const int NX = g.NX, NY = g.NY, n = NY, c = i * n + j;
  const float* Ut = d->buf[rUt];
  const float* Vt = d->buf[rVt];
  float* nU = d->buf[B_NU];
  float* nV = d->buf[B_NV];
  const bool ring = i == 0 || i == NX - 1 || j == 0 || j == NY - 1;
  const uint32_t f = ld_flag(d->flags, c);
  // geometric ranges of the advected faces (u: 1 <= j <= NY-2, v: 1 <= i <= NX-2, both i, j >= 1)
  const bool gu = i >= 1 && j >= 1 && j <= NY - 2, gv = i >= 1 && j >= 1 && i <= NX - 2;
  const bool au = gu && (f & CF_FLUID) && (f & CF_X0), av = gv && (f & CF_FLUID) && (f & CF_Y0);
  const int cu = gu ? c : n + 1, cv = gv ? c : n + 1;  // (1,1): its neighbourhood is in range
  const float uc = ld_cell(Ut, c), vc = ld_cell(Vt, c);
  const float v_at_u = avg_v(Vt, cu, n), u_at_v = avg_u(Ut, cv, n);
  const float su = ring ? uc : ld_cell(nU, c), sv = ring ? vc : ld_cell(nV, c);  // kept where not advected
  int badu = 0, badv = 0;
  float ru, rv;
  if (g.precise) {
    ru = sample_cells(i, j, ((-dt) * uc) * g.h1, ((-dt) * v_at_u) * g.h1, Ut, false, true, g, &badu);
    rv = sample_cells(i, j, ((-dt) * u_at_v) * g.h1, ((-dt) * vc) * g.h1, Vt, true, false, g, &badv);
  } else {
    float xu = (float)i * g.h;
    float yu = (float)j * g.h + g.ha;
    xu = xu - dt * uc;
    yu = yu - dt * v_at_u;
    float xv = (float)i * g.h + g.ha;
    float yv = (float)j * g.h;
    xv = xv - dt * u_at_v;
    yv = yv - dt * vc;
    ru = sample_field(xu, yu, Ut, 0.0f, g.hs, g, &badu);
    rv = sample_field(xv, yv, Vt, g.hs, 0.0f, g, &badv);
  }
  const float ou = au ? ru : su, ov = av ? rv : sv;
  if (write_new) {
    st_cell(nU, c, ou);
    st_cell(nV, c, ov);
  }
  st_cell(d->buf[rUo], c, ou);
  st_cell(d->buf[rVo], c, ov);
  if ((au && badu) || (av && badv)) atomicOr(d->ctrl + 2, 1);
}
int launch_advect_velocity(Ctx* c, const float* Ut, const float* Vt, float* Uo, float* Vo, float dt, bool write_new) {
  const int a = role_of(c, Ut), b = role_of(c, Vt), e = role_of(c, Uo), f = role_of(c, Vo);
  launch_cells(c, [&](const OneSim& m, dim3 g, dim3 bl, hipStream_t s) {
    hipLaunchKernelGGL(k_advect_velocity<OneSim>, g, bl, 0, s, m, a, b, e, f, dt, (int)write_new);
  });
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// clampToNeighbors (fluid.go:1094-1120) for an interior cell (the 3x3 block is in range),
// in two parts: the 3x3 range of src (taken before the back-trace, whose loads it does not
// depend on) and the clamp of the corrected value to it. The comparisons are the reference's,
// so a NaN neighbour is skipped and a NaN centre leaves the value unclamped.
struct NbRange {
  float lo, hi;
};
__device__ __forceinline__ NbRange neighbor_range(const float* __restrict__ src, int c, int n) {
  NbRange r{ld_cell(src, c), ld_cell(src, c)};
#pragma unroll
  for (int di = -1; di <= 1; di++)
#pragma unroll
    for (int dj = -1; dj <= 1; dj++) {
      float v = ld_cell(src, c + di * n + dj);
      if (v < r.lo) r.lo = v;
      if (v > r.hi) r.hi = v;
    }
  return r;
}
__device__ __forceinline__ float clamp_to(float val, NbRange r) {
  if (val < r.lo) return r.lo;
  if (val > r.hi) return r.hi;
  return val;
}

// advectVelocityBFECC passes 2 and 3 (fluid.go:935-987): backward trace with the
// original velocity sampling the forward result, then correction and clamp. Loads are
// issued unconditionally from in-range addresses, as in k_advect_velocity.
template <class M>
__global__ void k_bfecc_velocity_corr(M m, int rUo, int rVo, int rUf, int rVf, int rUc, int rVc, float dt) {
  const SimDesc* d;
  int i, j;
  if (!cell_of(m, d, i, j)) return;
  const Grid g = d->g;
  const int NX = g.NX, NY = g.NY, n = NY, c = i * n + j;
  const float* Uo = d->buf[rUo];
  const float* Vo = d->buf[rVo];
  const float* Uf = d->buf[rUf];
  const float* Vf = d->buf[rVf];
  const bool ring = i == 0 || i == NX - 1 || j == 0 || j == NY - 1;
  const uint32_t f = ld_flag(d->flags, c);
  const bool gu = i >= 1 && j >= 1 && j <= NY - 2, gv = i >= 1 && j >= 1 && i <= NX - 2;
  const bool au = gu && (f & CF_FLUID) && (f & CF_X0), av = gv && (f & CF_FLUID) && (f & CF_Y0);
  const int cu = gu ? c : n + 1, cv = gv ? c : n + 1, ci = ring ? n + 1 : c;
  const float ou = ld_cell(Uo, c), ov = ld_cell(Vo, c);
  const float v_at_u = avg_v(Vo, cu, n), u_at_v = avg_u(Uo, cv, n);
  const float fu = ld_cell(Uf, c), fv = ld_cell(Vf, c);
  const NbRange nu = neighbor_range(Uo, ci, n), nv = neighbor_range(Vo, ci, n);
  int badu = 0, badv = 0;
  float ru, rv;
  if (g.precise) {
    ru = sample_cells(i, j, (dt * ou) * g.h1, (dt * v_at_u) * g.h1, Uf, false, true, g, &badu);
    rv = sample_cells(i, j, (dt * u_at_v) * g.h1, (dt * ov) * g.h1, Vf, true, false, g, &badv);
  } else {
    float xu = (float)i * g.h;
    float yu = (float)j * g.h + g.ha;
    xu = xu + dt * ou;
    yu = yu + dt * v_at_u;
    float xv = (float)i * g.h + g.ha;
    float yv = (float)j * g.h;
    xv = xv + dt * u_at_v;
    yv = yv + dt * ov;
    ru = sample_field(xu, yu, Uf, 0.0f, g.hs, g, &badu);
    rv = sample_field(xv, yv, Vf, g.hs, 0.0f, g, &badv);
  }
  const float bu = au ? ru : (ring ? fu : 0.0f), bv = av ? rv : (ring ? fv : 0.0f);
  // interior cells: corr = orig - (bwd - orig)/2 clamped to the 3x3 of orig; the ring keeps orig
  const float kcu = clamp_to(ou - (bu - ou) * 0.5f, nu);
  const float kcv = clamp_to(ov - (bv - ov) * 0.5f, nv);
  st_cell(d->buf[rUc], c, ring ? ou : kcu);
  st_cell(d->buf[rVc], c, ring ? ov : kcv);
  if ((au && badu) || (av && badv)) atomicOr(d->ctrl + 2, 1);
}
int launch_bfecc_velocity_corr(Ctx* c, const float* Uorig, const float* Vorig, const float* Ufwd, const float* Vfwd,
                               float* Ucorr, float* Vcorr, float dt) {
  const int a = role_of(c, Uorig), b = role_of(c, Vorig), e = role_of(c, Ufwd), f = role_of(c, Vfwd),
            x = role_of(c, Ucorr), y = role_of(c, Vcorr);
  launch_cells(c, [&](const OneSim& m, dim3 g, dim3 bl, hipStream_t s) {
    hipLaunchKernelGGL(k_bfecc_velocity_corr<OneSim>, g, bl, 0, s, m, a, b, e, f, x, y, dt);
  });
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// advectSmoke (fluid.go:400-434) using the current (already advected) velocity. Loads are
// issued unconditionally from in-range addresses (ring cells read around cell (1,1)).
template <class M>
__global__ void k_advect_smoke(M m, int rMs, int rMo, float dt, float smokeAdv, float visc, int write_new) {
  const SimDesc* d;
  int i, j;
  if (!cell_of(m, d, i, j)) return;
  const Grid g = d->g;
  const int NX = g.NX, NY = g.NY, n = NY, c = i * n + j;
  const float* U = d->buf[B_U];
  const float* V = d->buf[B_V];
  const float* Ms = d->buf[rMs];
  float* nM = d->buf[B_NM];
  const bool ring = i == 0 || i == NX - 1 || j == 0 || j == NY - 1;
  const uint32_t f = ld_flag(d->flags, c);
  const bool adv = (f & CF_INTERIOR) && (f & CF_FLUID);
  const int ci = ring ? n + 1 : c;
  const float mc = ld_cell(Ms, c);
  const float keep = ring ? mc : ld_cell(nM, c);
  float u = (ld_cell(U, ci) + ld_cell(U, ci + n)) * 0.5f * smokeAdv;
  float v = (ld_cell(V, ci) + ld_cell(V, ci + 1)) * 0.5f * smokeAdv;
  float nb = 0.0f;
  if (visc > 0.0f)
    nb = ld_cell(Ms, ci - n) + ld_cell(Ms, ci + n) + ld_cell(Ms, ci - 1) + ld_cell(Ms, ci + 1) - 4.0f * ld_cell(Ms, ci);
  int badl = 0;
  float x = (float)i * g.h + g.hs - dt * u;
  float y = (float)j * g.h + g.hs - dt * v;
  float val = g.precise ? sample_cells(i, j, ((-dt) * u) * g.h1, ((-dt) * v) * g.h1, Ms, true, true, g, &badl)
                        : sample_field(x, y, Ms, g.hs, g.hs, g, &badl);
  if (visc > 0.0f) {
    float sd = visc * 0.3f * dt;
    val += sd * nb;
  }
  const float out = adv ? go_maxf(val, 0.0f) : keep;
  if (write_new) st_cell(nM, c, out);
  st_cell(d->buf[rMo], c, out);
  if (adv && badl) atomicOr(d->ctrl + 2, 1);
}
int launch_advect_smoke(Ctx* c, const float* Msrc, float* Mo, float dt, bool write_new) {
  const int a = role_of(c, Msrc), b = role_of(c, Mo);
  const float sa = c->prm.smoke_advection, visc = c->prm.viscosity_diffusion;
  launch_cells(c, [&](const OneSim& m, dim3 g, dim3 bl, hipStream_t s) {
    hipLaunchKernelGGL(k_advect_smoke<OneSim>, g, bl, 0, s, m, a, b, dt, sa, visc, (int)write_new);
  });
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// advectSmokeBFECC passes 2 and 3 (fluid.go:1013-1046). Loads are issued unconditionally
// from in-range addresses (ring cells read around cell (1,1)).
template <class M> __global__ void k_bfecc_smoke_corr(M m, int rMo, int rMf, int rMc, float dt, float smokeAdv) {
  const SimDesc* d;
  int i, j;
  if (!cell_of(m, d, i, j)) return;
  const Grid g = d->g;
  const int NX = g.NX, NY = g.NY, n = NY, c = i * n + j;
  const float* U = d->buf[B_U];
  const float* V = d->buf[B_V];
  const float* Mo = d->buf[rMo];
  const float* Mf = d->buf[rMf];
  const bool ring = i == 0 || i == NX - 1 || j == 0 || j == NY - 1;
  const uint32_t f = ld_flag(d->flags, c);
  const bool adv = !ring && (f & CF_FLUID);
  const int ci = ring ? n + 1 : c;
  const float om = ld_cell(Mo, c);
  float u = (ld_cell(U, ci) + ld_cell(U, ci + n)) * 0.5f * smokeAdv;
  float v = (ld_cell(V, ci) + ld_cell(V, ci + 1)) * 0.5f * smokeAdv;
  const NbRange nr = neighbor_range(Mo, ci, n);
  int badl = 0;
  float x = (float)i * g.h + g.hs + dt * u;
  float y = (float)j * g.h + g.hs + dt * v;
  const float bs = g.precise ? sample_cells(i, j, (dt * u) * g.h1, (dt * v) * g.h1, Mf, true, true, g, &badl)
                             : sample_field(x, y, Mf, g.hs, g.hs, g, &badl);
  const float bm = adv ? bs : 0.0f;  // solid interior cells: the backward buffer's zero
  float cm = clamp_to(om - (bm - om) * 0.5f, nr);
  if (cm < 0.0f) cm = 0.0f;
  st_cell(d->buf[rMc], c, ring ? om : cm);
  if (adv && badl) atomicOr(d->ctrl + 2, 1);
}
int launch_bfecc_smoke_corr(Ctx* c, const float* Morig, const float* Mfwd, float* Mcorr, float dt) {
  const int a = role_of(c, Morig), b = role_of(c, Mfwd), e = role_of(c, Mcorr);
  const float sa = c->prm.smoke_advection;
  launch_cells(c, [&](const OneSim& m, dim3 g, dim3 bl, hipStream_t s) {
    hipLaunchKernelGGL(k_bfecc_smoke_corr<OneSim>, g, bl, 0, s, m, a, b, e, dt, sa);
  });
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// ---- batched passes: one launch over the work-list of all members ------------------------
static ManySims many(Batch* B) { return ManySims{B->d_sims, B->d_cells, B->d_rings}; }
#define FL_BATCH_CELLS(B) dim3((B)->ncell_items), dim3(BX, BY), 0, (B)->stream
#define FL_BATCH_RING(B) dim3((B)->nring_items), dim3(256), 0, (B)->stream

int batch_fill_role(Batch* B, int r, float v) {
  hipLaunchKernelGGL(k_fill_role<ManySims>, FL_BATCH_CELLS(B), many(B), r, v);
  FL_HIP(hipGetLastError());
  return FL_OK;
}
int batch_copy_role(Batch* B, int rdst, int rsrc) {
  hipLaunchKernelGGL(k_copy_role<ManySims>, FL_BATCH_CELLS(B), many(B), rdst, rsrc);
  FL_HIP(hipGetLastError());
  return FL_OK;
}
// Swap two roles in every member's descriptor: the ping-pong of a field whose pass writes
// its partner (semi-Lagrangian smoke writes M2), in place of copying it back. The host swaps
// the members' pointers to match after enqueueing the step (swap_member_roles).
__global__ void k_swap_roles(SimDesc* sims, int n, int a, int b) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  float* t = sims[e].buf[a];
  sims[e].buf[a] = sims[e].buf[b];
  sims[e].buf[b] = t;
}
int batch_swap_roles(Batch* B, int a, int b) {
  const int n = (int)B->members.size();
  hipLaunchKernelGGL(k_swap_roles, dim3((n + 63) / 64), dim3(64), 0, B->stream, B->d_sims, n, a, b);
  FL_HIP(hipGetLastError());
  return FL_OK;
}
int batch_flags(Batch* B) {
  hipLaunchKernelGGL(k_flags<ManySims>, FL_BATCH_CELLS(B), many(B));
  FL_HIP(hipGetLastError());
  return FL_OK;
}
int batch_viscosity(Batch* B, float dt) {
  hipLaunchKernelGGL(k_viscosity<ManySims>, FL_BATCH_CELLS(B), many(B), B->prm.viscosity_diffusion * dt);
  FL_HIP(hipGetLastError());
  return FL_OK;
}
int batch_conf_turb(Batch* B, float dt, int rU, int rV) {
  const fl_params& P = B->prm;
  const bool conf = P.confinement != 0.0f, turb = P.turbulence_strength > 0.0f;
  if (conf) hipLaunchKernelGGL(k_curl<ManySims>, FL_BATCH_CELLS(B), many(B), (int)B_T0, rU, rV);
  if (conf || turb)
    hipLaunchKernelGGL(k_conf_turb<ManySims>, FL_BATCH_CELLS(B), many(B), rU, rV, (int)B_T0, P.confinement, dt,
                       P.turbulence_strength * dt, (int)conf, (int)turb);
  FL_HIP(hipGetLastError());
  return FL_OK;
}
int batch_borders(Batch* B, int rU, int rV) {
  hipLaunchKernelGGL(k_borders<ManySims>, FL_BATCH_RING(B), many(B), rU, rV);
  FL_HIP(hipGetLastError());
  return FL_OK;
}
int batch_advect_velocity(Batch* B, int rUt, int rVt, int rUo, int rVo, float dt, bool write_new) {
  hipLaunchKernelGGL(k_advect_velocity<ManySims>, FL_BATCH_CELLS(B), many(B), rUt, rVt, rUo, rVo, dt,
                     (int)write_new);
  FL_HIP(hipGetLastError());
  return FL_OK;
}
int batch_bfecc_velocity_corr(Batch* B, int rUo, int rVo, int rUf, int rVf, float dt) {
  hipLaunchKernelGGL(k_bfecc_velocity_corr<ManySims>, FL_BATCH_CELLS(B), many(B), rUo, rVo, rUf, rVf, (int)B_T0,
                     (int)B_T1, dt);
  FL_HIP(hipGetLastError());
  return FL_OK;
}
int batch_advect_smoke(Batch* B, int rMs, int rMo, float dt, bool write_new) {
  hipLaunchKernelGGL(k_advect_smoke<ManySims>, FL_BATCH_CELLS(B), many(B), rMs, rMo, dt, B->prm.smoke_advection,
                     B->prm.viscosity_diffusion, (int)write_new);
  FL_HIP(hipGetLastError());
  return FL_OK;
}
int batch_bfecc_smoke_corr(Batch* B, float dt) {
  hipLaunchKernelGGL(k_bfecc_smoke_corr<ManySims>, FL_BATCH_CELLS(B), many(B), (int)B_M, (int)B_M2, (int)B_T2, dt,
                     B->prm.smoke_advection);
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
// Batched edits: segment s belongs to member segsim[s]; keys are member-local
// (field * cells + idx) and fields map to the member's current buffers.
__global__ void k_apply_edits_batch(const SimDesc* __restrict__ sims, const uint32_t* __restrict__ keys,
                                    const float* __restrict__ vals, const int32_t* __restrict__ seg,
                                    const int32_t* __restrict__ segsim, int nseg) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  const SimDesc& d = sims[segsim[s]];
  const uint32_t ncells = (uint32_t)d.g.NX * (uint32_t)d.g.NY;
  int b = seg[s], e = seg[s + 1];
  uint32_t key = keys[b] & 0x7fffffffu;
  uint32_t field = key / ncells, idx = key % ncells;
  float* dst = d.buf[field];
  float x = dst[idx];
  for (int k = b; k < e; k++) {
    bool add = (keys[k] >> 31) != 0;
    x = add ? x + vals[k] : vals[k];
  }
  dst[idx] = x;
}
int batch_apply_edits(Batch* B, const uint32_t* keys, const float* vals, const int32_t* seg, const int32_t* segsim,
                      int nseg) {
  if (nseg == 0) return FL_OK;
  hipLaunchKernelGGL(k_apply_edits_batch, dim3((nseg + 255) / 256), dim3(256), 0, B->stream, B->d_sims, keys, vals,
                     seg, segsim, nseg);
  FL_HIP(hipGetLastError());
  return FL_OK;
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
// `sim` (batched program) names the member each segment belongs to.
__device__ __forceinline__ void apply_source_seg(const SrcOp* __restrict__ ops, int b, int e,
                                                 const float* __restrict__ S, float* __restrict__ U,
                                                 float* __restrict__ V, float* __restrict__ M) {
  const int cell = ops[b].cell;
  float u = U[cell], v = V[cell], m = M[cell];
  const bool fluid = S[cell] != 0.0f;
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
__global__ void k_apply_sources(const SrcOp* __restrict__ ops, const int32_t* __restrict__ seg, int nseg,
                                const float* __restrict__ S, float* __restrict__ U, float* __restrict__ V,
                                float* __restrict__ M) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  apply_source_seg(ops, seg[s], seg[s + 1], S, U, V, M);
}
__global__ void k_apply_sources_batch(const SimDesc* __restrict__ sims, const SrcOp* __restrict__ ops,
                                      const int32_t* __restrict__ seg, const int32_t* __restrict__ segsim,
                                      int nseg) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  const SimDesc& d = sims[segsim[s]];
  apply_source_seg(ops, seg[s], seg[s + 1], d.buf[B_S], d.buf[B_U], d.buf[B_V], d.buf[B_M]);
}
int launch_apply_sources(Ctx* c) {
  if (c->n_src_seg == 0) return FL_OK;
  hipLaunchKernelGGL(k_apply_sources, dim3((c->n_src_seg + 255) / 256), dim3(256), 0, c->stream, c->d_src,
                     c->d_src_seg, c->n_src_seg, c->S, c->U, c->V, c->M);
  FL_HIP(hipGetLastError());
  return FL_OK;
}
int batch_apply_sources(Batch* B) {
  if (B->n_src_seg == 0) return FL_OK;
  hipLaunchKernelGGL(k_apply_sources_batch, dim3((B->n_src_seg + 255) / 256), dim3(256), 0, B->stream, B->d_sims,
                     B->d_src, B->d_src_seg, B->d_src_sim, B->n_src_seg);
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
    // CF_FLUID is S != 0; the reference tests S > 0 (fluid.go:536). S holds only +0.0 and 1.0
    // (fl_upload rejects every other bit pattern, SetSolid writes those two), so they agree.
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
  const int j = blockIdx.x * BX + threadIdx.x;
  const int i = blockIdx.y * BY + threadIdx.y;
  if (i >= g.NX || j >= g.NY) return;
  const int n = g.NY, c = i * n + j;
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
