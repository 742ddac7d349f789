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

#include <algorithm>
#include <map>
#include <mutex>
#include <utility>

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
  // the keep map: a wavefront is one row's 64-column block (BX = 64 lanes from j = 64*block)
  const bool kv = interior && !(fluid && (f & CF_X0) && (f & CF_Y0)), ks = interior && !fluid;
  const uint32_t kb = (wave_any(kv) ? 1u : 0u) | (wave_any(ks) ? 2u : 0u);
  if ((j & (BX - 1)) == 0) d->keep[i * ((NY + BX - 1) / BX) + j / BX] = kb;
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
// Gathering every neighbour and bilinear corner of every cell from global memory keeps the L1's
// address and data units ~90% busy while HBM runs at 3-4 TB/s. So the advection passes stage the
// windows of the fields they read in LDS with coalesced loads, a tile of TI x 64 cells plus a halo
// of HI rows and HJ columns (1 for fields read only by the face averages and the 3x3 clamp), and
// read stencils and corners from there. Corners outside the window (back-traces longer than the
// halo, near the lid) come from global memory, for those lanes only. Values are the same whichever
// memory they come from, so the result is bit-identical.
// TI is 16, 8 or 4 for one simulation (adv_rows), the work-list's 4 rows in a batch. A persistent
// variant that prefetched the next tile's windows into registers while computing the current one
// ran 20-30% slower: its registers cut occupancy, and any global load of the compute waits for the
// prefetch, since gfx9 counts outstanding loads in order.
// Halo measured at 4096^2 (cavity, BFECC; the six launches of a step): (HI, HJ) = (2, 1) 550 us,
// (2, 2) 552 us, (4, 2) 567 us. A wider halo serves a few more corners from LDS but stages more
// rows per tile than it saves.
constexpr int HI = 2, HJ = 1;

// A window of one field in LDS: grid rows [r0, r0 + R), columns [c0, c0 + C), row pitch C. Slots
// outside the grid hold copies of the nearest in-grid cells (the staging clamps its addresses), so
// slot (x + 1, y + 1) holds the cell (min(x + 1, NX - 1), min(y + 1, NY - 1)) that a bilinear
// sample pairs with (x, y).
template <int R, int C> struct Win {
  const float* s;
  int r0, c0;
  // (i, j) in the window: the local row is below 256, so its product is a 24-bit multiply
  __device__ __forceinline__ float at(int i, int j) const {
    return s[((unsigned)(i - r0) & 255u) * (unsigned)C + (unsigned)(j - c0)];
  }
};
// A window's elements in registers: wave w of WY loads rows w, w + WY, ... over the tile's 64 columns (the
// row offset is wave-uniform, so addresses cost one add per element), and the 2*HJ halo columns of
// all R rows come in one more load per thread.
template <int R, int C, int WY> struct Stage {
  static constexpr int H = (C - BX) / 2, RW = (R + WY - 1) / WY, PAD = R * C;
  static_assert(C - BX == 2 * H && 2 * H * R <= BX * WY, "halo columns in one pass");
  float v[RW], vh;
  __device__ __forceinline__ void load(const float* __restrict__ f, int r0, int c0, int NX, int NY) {
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y), lane = threadIdx.x;
    const int gj = min(max(c0 + H + lane, 0), NY - 1);
#pragma unroll
    for (int m = 0; m < RW; m++) {
      const int r = wy + WY * m;
      const int gi = min(max(r0 + min(r, R - 1), 0), NX - 1);
      v[m] = ld_cell(f, gi * NY + gj);
    }
    const int e = min(wy * BX + lane, 2 * H * R - 1), hr = e / (2 * H), side = e - hr * (2 * H);
    const int hi = min(max(r0 + hr, 0), NX - 1), hj = min(max(c0 + (side < H ? side : BX + side), 0), NY - 1);
    vh = ld_cell(f, hi * NY + hj);
  }
  __device__ __forceinline__ void store(float* s) const {
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y), lane = threadIdx.x;
#pragma unroll
    for (int m = 0; m < RW; m++) {
      const int r = wy + WY * m;
      if (r < R) s[r * C + H + lane] = v[m];
    }
    const int e = wy * BX + lane;
    if (e < 2 * H * R) {
      const int hr = e / (2 * H), side = e - hr * (2 * H);
      s[hr * C + (side < H ? side : BX + side)] = vh;
    }
  }
};
// Corner source of a sample: the LDS window where the 2x2 block at (x0, y0) lies in it, else global
// memory (only waves with such a lane issue the global loads). exact: x1, y1 are x0 + 1, y0 + 1 up
// to the grid's clamp (false only for the clamped negative indices of an invalid sample).
template <int R, int C> struct WinSrc {
  Win<R, C> w;
  const float* __restrict__ f;
  __device__ __forceinline__ Corners corners(int x0, int x1, int y0, int y1, int NY, bool exact) const {
    const unsigned a = (unsigned)(x0 - w.r0), cs = (unsigned)(y0 - w.c0);
    const bool in = exact && a < (unsigned)(R - 1) && cs < (unsigned)(C - 1);
    // every lane reads the window (lanes outside it at slot 0), so no branch guards the LDS reads
    const float* p = w.s + (in ? (a & 255u) * (unsigned)C + cs : 0u);
    Corners k{p[0], p[C], p[C + 1], p[1]};
    if (wave_any(!in) && !in) k = ld_corners(f, x0, x1, y0, y1, NY);
    return k;
  }
};

// Raise the invalid-sample flag (where the reference's slice index would panic). Once the flag
// is set, lanes only read it: on a field gone to NaN every wavefront of the pass has such a lane,
// and one atomic per wavefront on one address serialises in L2 (~0.25 ms per pass at 4096^2).
__device__ __forceinline__ void flag_invalid(int32_t* flag) {
  if (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) atomicOr(flag, 1);
}

// A tile: its simulation and its first row and column.
struct TileRef {
  const SimDesc* d;
  int i0, j0;
};
template <int TI> __device__ __forceinline__ TileRef tile_ref(const OneSim& m, int t, int ntiles) {
  const int gx = (m.d.g.NY + BX - 1) / BX;
  const int s = xcd_tile(t, ntiles);  // each XCD takes a contiguous band of tiles
  const int by = s / gx, bx = s - by * gx;
  return TileRef{&m.d, by * TI, bx * BX};
}
template <int TI> __device__ __forceinline__ TileRef tile_ref(const ManySims& m, int t, int) {
  static_assert(TI == BY, "work-list items are BY rows");
  const CellItem it = m.cells[t];
  return TileRef{&m.sims[it.sim], it.by * BY, it.bx * BX};
}

// The thread's cells are rows i0 + threadIdx.y + WY*k, k < NR, of column j0 + threadIdx.x. Their
// per-cell global inputs are prefetched with the windows (clamped addresses for cells outside the
// grid, which compute nothing).
template <int NR, int WY> __device__ __forceinline__ uint32_t row_flags(const TileRef& r) {
  static_assert(NR <= 4, "flags of up to 4 rows, one byte each");
  const int NX = r.d->g.NX, NY = r.d->g.NY, j = min(r.j0 + (int)threadIdx.x, NY - 1);
  uint32_t fr = 0;
#pragma unroll
  for (int k = 0; k < NR; k++)
    fr |= ld_flag(r.d->flags, min(r.i0 + (int)threadIdx.y + k * WY, NX - 1) * NY + j) << (8 * k);
  return fr;
}
// The cell-centred velocity of the smoke passes, U[ci], U[ci+n], V[ci], V[ci+1], loaded with the
// windows; ring cells (which do not advect) read around cell (1,1).
template <int NR, int WY> struct CellVel {
  float u0[NR], u1[NR], v0[NR], v1[NR];
  __device__ __forceinline__ void load(const TileRef& r) {
    const int NX = r.d->g.NX, NY = r.d->g.NY, j = r.j0 + (int)threadIdx.x;
    const float* U = r.d->buf[B_U];
    const float* V = r.d->buf[B_V];
#pragma unroll
    for (int k = 0; k < NR; k++) {
      const int i = r.i0 + (int)threadIdx.y + k * WY;
      const bool ring = i <= 0 || i >= NX - 1 || j <= 0 || j >= NY - 1;
      const int ci = ring ? NY + 1 : i * NY + j;
      const f2u vv = ld_pair(V, ci);
      u0[k] = ld_cell(U, ci);
      u1[k] = ld_cell(U, ci + NY);
      v0[k] = vv.x;
      v1[k] = vv.y;
    }
  }
};

// The stale values a pass keeps (newU/newV of faces it does not advect, newM of solid cells), for
// the rows whose keep-map entry has `bit`: the entry is a wave-uniform scalar load, so these loads
// are issued with the windows and no wave makes a second round trip for them.
template <int NR, int WY>
__device__ __forceinline__ void kept_rows(const TileRef& r, uint32_t bit, const float* __restrict__ a,
                                          const float* __restrict__ b, float (&va)[NR], float (&vb)[NR]) {
  const int NX = r.d->g.NX, NY = r.d->g.NY, wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int nb = (NY + BX - 1) / BX, c0 = min(r.j0 + (int)threadIdx.x, NY - 1);
#pragma unroll
  for (int k = 0; k < NR; k++) {
    const int i = r.i0 + wy + k * WY;
    va[k] = 0.0f;
    vb[k] = 0.0f;
    if (i < NX && (r.d->keep[i * nb + r.j0 / BX] & bit)) {
      va[k] = ld_cell(a, i * NY + c0);
      if (b) vb[k] = ld_cell(b, i * NY + c0);
    }
  }
}

// The tile loop shared by the advection passes: stage the tile's windows, then compute from LDS.
// P supplies the LDS windows (Lds), the staging registers (Pre), the per-cell inputs loaded with
// them (Rows) and the compute. One tile per workgroup: a workgroup's loads are all issued before its
// first wait, and the 8 workgroups resident per CU overlap each other's staging and compute.
template <class M, class P> __device__ __forceinline__ void run_tile(const M& m, const typename P::Args& a, int ntiles) {
  __shared__ typename P::Lds lds;
  const TileRef r = tile_ref<P::TI>(m, blockIdx.x, ntiles);
  typename P::Rows w;
  {
    typename P::Pre pre;
    P::prefetch(r, a, pre, w);
    P::store(pre, lds);
  }
  __syncthreads();
  P::compute(r, a, lds, w);
}
static inline int one_sim_tiles(int NX, int NY, int TI) { return ((NY + BX - 1) / BX) * ((NX + TI - 1) / TI); }
// Rows per tile of one simulation: 16 (four waves of four rows) where that leaves >= 32 tiles per
// CU, else 8 where that leaves >= 16, else 4: the tail of the last round of workgroups stays short
// on small grids. Measured at 4096^2 (cavity, BFECC, six launches): 16 rows x 4 waves 554 us;
// 32 x 8 585; 16 x 8 610; 8 x 4 589; 8 x 2 588.
static int adv_rows(const Ctx* c) {
  static int cus[64];
  const int dev = c->dev & 63;
  int n = __atomic_load_n(&cus[dev], __ATOMIC_RELAXED);
  if (n <= 0) {
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, c->dev) != hipSuccess || n <= 0) n = 256;
    __atomic_store_n(&cus[dev], n, __ATOMIC_RELAXED);
  }
  if (one_sim_tiles(c->NX, c->NY, 16) >= 32 * n) return 16;
  if (one_sim_tiles(c->NX, c->NY, 8) >= 16 * n) return 8;
  return 4;
}
#define FL_ADV_LAUNCH(K, c, a)                                                                                      \
  do {                                                                                                              \
    const int ti_ = adv_rows(c), nt_ = one_sim_tiles((c)->NX, (c)->NY, ti_);                                        \
    const OneSim m_{(c)->desc()};                                                                                   \
    if (ti_ == 16)                                                                                                  \
      hipLaunchKernelGGL((K<OneSim, 16, 4>), dim3(nt_), dim3(BX, 4), 0, (c)->stream, m_, a, nt_);                   \
    else if (ti_ == 8)                                                                                              \
      hipLaunchKernelGGL((K<OneSim, 8, 4>), dim3(nt_), dim3(BX, 4), 0, (c)->stream, m_, a, nt_);                    \
    else                                                                                                            \
      hipLaunchKernelGGL((K<OneSim, 4, 4>), dim3(nt_), dim3(BX, 4), 0, (c)->stream, m_, a, nt_);                    \
  } while (0)

// advectVelocity: copyBorder(newU,U), the advected faces, stale newU elsewhere; then U=newU.
// (Ut,Vt) is the current velocity (traced and sampled, staged with halo (HI, HJ)); the result goes to
// newU/newV and to (Uo,Vo). Faces outside a component's geometric range read the window around
// their own cell (values no output uses). The stale newU/newV are read only by waves holding a
// face that is kept (ring or not advected), as the keep map says: in open flow no wave reads them.
struct AdvVelArgs {
  int rUt, rVt, rUo, rVo;
  float dt;
  int write_new;
};
template <int TI_, int WY> struct AdvVelPass {
  static constexpr int TI = TI_, NR = TI / WY, R = TI + 2 * HI, C = BX + 2 * HJ;
  using Args = AdvVelArgs;
  struct Lds {
    float U[Stage<R, C, WY>::PAD], V[Stage<R, C, WY>::PAD];
  };
  struct Pre {
    Stage<R, C, WY> u, v;
  };
  struct Rows {
    uint32_t f;          // the NR rows' flags, a byte each
    float su[NR], sv[NR];  // stale newU/newV, loaded for rows whose keep bit 0 is set
  };
  static __device__ __forceinline__ void prefetch(const TileRef& r, const Args& a, Pre& p, Rows& w) {
    const int NX = r.d->g.NX, NY = r.d->g.NY;
    w.f = row_flags<NR, WY>(r);
    p.u.load(r.d->buf[a.rUt], r.i0 - HI, r.j0 - HJ, NX, NY);
    p.v.load(r.d->buf[a.rVt], r.i0 - HI, r.j0 - HJ, NX, NY);
    kept_rows<NR, WY>(r, 1u, r.d->buf[B_NU], r.d->buf[B_NV], w.su, w.sv);
  }
  static __device__ __forceinline__ void store(const Pre& p, Lds& l) {
    p.u.store(l.U);
    p.v.store(l.V);
  }
  static __device__ __forceinline__ void compute(const TileRef& r, const Args& a, const Lds& l, const Rows& w) {
    const SimDesc* d = r.d;
    const Grid g = d->g;
    const int NX = g.NX, NY = g.NY, n = NY;
    const float dt = a.dt;
    const Win<R, C> wU{l.U, r.i0 - HI, r.j0 - HJ}, wV{l.V, r.i0 - HI, r.j0 - HJ};
    const WinSrc<R, C> srcU{wU, d->buf[a.rUt]}, srcV{wV, d->buf[a.rVt]};
    float* nU = d->buf[B_NU];
    float* nV = d->buf[B_NV];
    float* Uo = d->buf[a.rUo];
    float* Vo = d->buf[a.rVo];
    const int j = r.j0 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < NR; k++) {
      const int i = r.i0 + threadIdx.y + k * WY;
      if (i >= NX || j >= NY) continue;
      const int c = i * n + j;
      const bool ring = i == 0 || i == NX - 1 || j == 0 || j == NY - 1;
      const uint32_t f = (w.f >> (8 * k)) & 0xffu;
      // geometric ranges of the advected faces (u: 1 <= j <= NY-2, v: 1 <= i <= NX-2, both i, j >= 1)
      const bool gu = i >= 1 && j >= 1 && j <= NY - 2, gv = i >= 1 && j >= 1 && i <= NX - 2;
      const bool au = gu && (f & CF_FLUID) && (f & CF_X0), av = gv && (f & CF_FLUID) && (f & CF_Y0);
      const float uc = wU.at(i, j), vc = wV.at(i, j);
      // avgV(i,j) and avgU(i,j) (fluid.go:336-350)
      const float v_at_u = (wV.at(i - 1, j) + wV.at(i, j) + wV.at(i - 1, j + 1) + wV.at(i, j + 1)) * 0.25f;
      const float u_at_v = (wU.at(i, j - 1) + wU.at(i, j) + wU.at(i + 1, j - 1) + wU.at(i + 1, j)) * 0.25f;
      // kept where not advected: the ring's copyBorder, else the stale newU/newV
      const float su = ring ? uc : w.su[k], sv = ring ? vc : w.sv[k];
      int badu = 0, badv = 0;
      float ru, rv;
      if (g.precise) {
        ru = sample_cells(i, j, ((-dt) * uc) * g.h1, ((-dt) * v_at_u) * g.h1, srcU, false, true, g, &badu);
        rv = sample_cells(i, j, ((-dt) * u_at_v) * g.h1, ((-dt) * vc) * g.h1, srcV, true, false, g, &badv);
      } else {
        float xu = (float)i * g.h;
        float yu = (float)j * g.h + g.ha;
        xu = xu - dt * uc;
        yu = yu - dt * v_at_u;
        float xv = (float)i * g.h + g.ha;
        float yv = (float)j * g.h;
        xv = xv - dt * u_at_v;
        yv = yv - dt * vc;
        ru = sample_field(xu, yu, srcU, 0.0f, g.hs, g, &badu);
        rv = sample_field(xv, yv, srcV, g.hs, 0.0f, g, &badv);
      }
      const float ou = au ? ru : su, ov = av ? rv : sv;
      if (a.write_new) {
        st_cell(nU, c, ou);
        st_cell(nV, c, ov);
      }
      st_cell(Uo, c, ou);
      st_cell(Vo, c, ov);
      if ((au && badu) || (av && badv)) flag_invalid(d->ctrl + 2);
    }
  }
};
template <class M, int TI, int WY> __global__ void __launch_bounds__(BX* WY) k_advect_velocity(M m, AdvVelArgs a, int ntiles) {
  run_tile<M, AdvVelPass<TI, WY>>(m, a, ntiles);
}
int launch_advect_velocity(Ctx* c, const float* Ut, const float* Vt, float* Uo, float* Vo, float dt, bool write_new) {
  const AdvVelArgs a{role_of(c, Ut), role_of(c, Vt), role_of(c, Uo), role_of(c, Vo), dt, (int)write_new};
  FL_ADV_LAUNCH(k_advect_velocity, c, a);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// clampToNeighbors (fluid.go:1094-1120) for an interior cell: the 3x3 range of src (taken
// before the back-trace, whose loads it does not depend on) and the clamp of the corrected
// value to it. The comparisons are the reference's, so a NaN neighbour is skipped and a NaN
// centre leaves the value unclamped.
struct NbRange {
  float lo, hi;
};
template <class W> __device__ __forceinline__ NbRange neighbor_range(const W& w, int i, int j) {
  NbRange r{w.at(i, j), w.at(i, j)};
#pragma unroll
  for (int di = -1; di <= 1; di++)
#pragma unroll
    for (int dj = -1; dj <= 1; dj++) {
      const float v = w.at(i + di, j + dj);
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

// advectVelocityBFECC passes 2 and 3 (fluid.go:935-987): backward trace with the original
// velocity (Uo,Vo; windows with halo 1 for the face averages and the 3x3 clamp ranges) sampling
// the forward result (Uf,Vf; windows with halo (HI, HJ)), then correction and clamp. Ring cells keep
// the original value; their traces still flag invalid samples where the reference's would fault.
struct CorrVelArgs {
  int rUo, rVo, rUf, rVf, rUc, rVc;
  float dt;
};
template <int TI_, int WY> struct CorrVelPass {
  static constexpr int TI = TI_, NR = TI / WY, R = TI + 2 * HI, C = BX + 2 * HJ, R1 = TI + 2, C1 = BX + 2;
  using Args = CorrVelArgs;
  struct Lds {
    float Uf[Stage<R, C, WY>::PAD], Vf[Stage<R, C, WY>::PAD], Uo[Stage<R1, C1, WY>::PAD], Vo[Stage<R1, C1, WY>::PAD];
  };
  struct Pre {
    Stage<R1, C1, WY> uo, vo;
    Stage<R, C, WY> uf, vf;
  };
  struct Rows {
    uint32_t f;  // the NR rows' flags, a byte each
  };
  static __device__ __forceinline__ void prefetch(const TileRef& r, const Args& a, Pre& p, Rows& w) {
    const int NX = r.d->g.NX, NY = r.d->g.NY;
    w.f = row_flags<NR, WY>(r);
    p.uo.load(r.d->buf[a.rUo], r.i0 - 1, r.j0 - 1, NX, NY);
    p.vo.load(r.d->buf[a.rVo], r.i0 - 1, r.j0 - 1, NX, NY);
    p.uf.load(r.d->buf[a.rUf], r.i0 - HI, r.j0 - HJ, NX, NY);
    p.vf.load(r.d->buf[a.rVf], r.i0 - HI, r.j0 - HJ, NX, NY);
  }
  static __device__ __forceinline__ void store(const Pre& p, Lds& l) {
    p.uo.store(l.Uo);
    p.vo.store(l.Vo);
    p.uf.store(l.Uf);
    p.vf.store(l.Vf);
  }
  static __device__ __forceinline__ void compute(const TileRef& r, const Args& a, const Lds& l, const Rows& w) {
    const SimDesc* d = r.d;
    const Grid g = d->g;
    const int NX = g.NX, NY = g.NY, n = NY;
    const float dt = a.dt;
    const Win<R1, C1> wU{l.Uo, r.i0 - 1, r.j0 - 1}, wV{l.Vo, r.i0 - 1, r.j0 - 1};
    const WinSrc<R, C> srcU{Win<R, C>{l.Uf, r.i0 - HI, r.j0 - HJ}, d->buf[a.rUf]},
        srcV{Win<R, C>{l.Vf, r.i0 - HI, r.j0 - HJ}, d->buf[a.rVf]};
    float* Uc = d->buf[a.rUc];
    float* Vc = d->buf[a.rVc];
    const int j = r.j0 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < NR; k++) {
      const int i = r.i0 + threadIdx.y + k * WY;
      if (i >= NX || j >= NY) continue;
      const int c = i * n + j;
      const bool ring = i == 0 || i == NX - 1 || j == 0 || j == NY - 1;
      const uint32_t f = (w.f >> (8 * k)) & 0xffu;
      const bool gu = i >= 1 && j >= 1 && j <= NY - 2, gv = i >= 1 && j >= 1 && i <= NX - 2;
      const bool au = gu && (f & CF_FLUID) && (f & CF_X0), av = gv && (f & CF_FLUID) && (f & CF_Y0);
      const float ou = wU.at(i, j), ov = wV.at(i, j);
      const float v_at_u = (wV.at(i - 1, j) + wV.at(i, j) + wV.at(i - 1, j + 1) + wV.at(i, j + 1)) * 0.25f;
      const float u_at_v = (wU.at(i, j - 1) + wU.at(i, j) + wU.at(i + 1, j - 1) + wU.at(i + 1, j)) * 0.25f;
      const NbRange nu = neighbor_range(wU, i, j), nv = neighbor_range(wV, i, j);
      int badu = 0, badv = 0;
      float ru, rv;
      if (g.precise) {
        ru = sample_cells(i, j, (dt * ou) * g.h1, (dt * v_at_u) * g.h1, srcU, false, true, g, &badu);
        rv = sample_cells(i, j, (dt * u_at_v) * g.h1, (dt * ov) * g.h1, srcV, true, false, g, &badv);
      } else {
        float xu = (float)i * g.h;
        float yu = (float)j * g.h + g.ha;
        xu = xu + dt * ou;
        yu = yu + dt * v_at_u;
        float xv = (float)i * g.h + g.ha;
        float yv = (float)j * g.h;
        xv = xv + dt * u_at_v;
        yv = yv + dt * ov;
        ru = sample_field(xu, yu, srcU, 0.0f, g.hs, g, &badu);
        rv = sample_field(xv, yv, srcV, g.hs, 0.0f, g, &badv);
      }
      // interior cells: corr = orig - (bwd - orig)/2 clamped to the 3x3 of orig, bwd being 0 at
      // faces the backward pass does not write; the ring keeps orig
      const float bu = au ? ru : 0.0f, bv = av ? rv : 0.0f;
      const float kcu = clamp_to(ou - (bu - ou) * 0.5f, nu);
      const float kcv = clamp_to(ov - (bv - ov) * 0.5f, nv);
      st_cell(Uc, c, ring ? ou : kcu);
      st_cell(Vc, c, ring ? ov : kcv);
      if ((au && badu) || (av && badv)) flag_invalid(d->ctrl + 2);
    }
  }
};
template <class M, int TI, int WY> __global__ void __launch_bounds__(BX* WY) k_bfecc_velocity_corr(M m, CorrVelArgs a, int ntiles) {
  run_tile<M, CorrVelPass<TI, WY>>(m, a, ntiles);
}
int launch_bfecc_velocity_corr(Ctx* c, const float* Uorig, const float* Vorig, const float* Ufwd, const float* Vfwd,
                               float* Ucorr, float* Vcorr, float dt) {
  const CorrVelArgs a{role_of(c, Uorig), role_of(c, Vorig), role_of(c, Ufwd), role_of(c, Vfwd),
                      role_of(c, Ucorr), role_of(c, Vcorr), dt};
  FL_ADV_LAUNCH(k_bfecc_velocity_corr, c, a);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// advectSmoke (fluid.go:400-434) using the current (already advected) velocity (CellVel); the
// source field is staged with halo (HI, HJ). The stale newM is loaded only for rows holding a solid
// interior cell (keep map).
struct AdvSmokeArgs {
  int rMs, rMo;
  float dt, smokeAdv, visc;
  int write_new;
};
template <int TI_, int WY> struct AdvSmokePass {
  static constexpr int TI = TI_, NR = TI / WY, R = TI + 2 * HI, C = BX + 2 * HJ;
  using Args = AdvSmokeArgs;
  struct Lds {
    float M[Stage<R, C, WY>::PAD];
  };
  struct Pre {
    Stage<R, C, WY> m;
  };
  struct Rows {
    uint32_t f;  // the NR rows' flags, a byte each
    CellVel<NR, WY> vel;
    float mn[NR];  // stale newM, loaded for rows whose keep bit 1 is set
  };
  static __device__ __forceinline__ void prefetch(const TileRef& r, const Args& a, Pre& p, Rows& w) {
    w.f = row_flags<NR, WY>(r);
    w.vel.load(r);
    p.m.load(r.d->buf[a.rMs], r.i0 - HI, r.j0 - HJ, r.d->g.NX, r.d->g.NY);
    float unused[NR];
    kept_rows<NR, WY>(r, 2u, r.d->buf[B_NM], nullptr, w.mn, unused);
  }
  static __device__ __forceinline__ void store(const Pre& p, Lds& l) { p.m.store(l.M); }
  static __device__ __forceinline__ void compute(const TileRef& r, const Args& a, const Lds& l, const Rows& w) {
    const SimDesc* d = r.d;
    const Grid g = d->g;
    const int NX = g.NX, NY = g.NY, n = NY;
    const float dt = a.dt;
    const Win<R, C> wM{l.M, r.i0 - HI, r.j0 - HJ};
    const WinSrc<R, C> srcM{wM, d->buf[a.rMs]};
    float* nM = d->buf[B_NM];
    float* Mo = d->buf[a.rMo];
    const int j = r.j0 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < NR; k++) {
      const int i = r.i0 + threadIdx.y + k * WY;
      if (i >= NX || j >= NY) continue;
      const int c = i * n + j;
      const bool ring = i == 0 || i == NX - 1 || j == 0 || j == NY - 1;
      const uint32_t f = (w.f >> (8 * k)) & 0xffu;
      const bool adv = (f & CF_INTERIOR) && (f & CF_FLUID);
      float u = (w.vel.u0[k] + w.vel.u1[k]) * 0.5f * a.smokeAdv;
      float v = (w.vel.v0[k] + w.vel.v1[k]) * 0.5f * a.smokeAdv;
      float nb = 0.0f;
      if (a.visc > 0.0f)  // used at interior cells only
        nb = wM.at(i - 1, j) + wM.at(i + 1, j) + wM.at(i, j - 1) + wM.at(i, j + 1) - 4.0f * wM.at(i, j);
      const float keep = ring ? wM.at(i, j) : w.mn[k];
      int badl = 0;
      float x = (float)i * g.h + g.hs - dt * u;
      float y = (float)j * g.h + g.hs - dt * v;
      float val = g.precise ? sample_cells(i, j, ((-dt) * u) * g.h1, ((-dt) * v) * g.h1, srcM, true, true, g, &badl)
                            : sample_field(x, y, srcM, g.hs, g.hs, g, &badl);
      if (a.visc > 0.0f) {
        float sd = a.visc * 0.3f * dt;
        val += sd * nb;
      }
      const float out = adv ? go_maxf(val, 0.0f) : keep;
      if (a.write_new) st_cell(nM, c, out);
      st_cell(Mo, c, out);
      if (adv && badl) flag_invalid(d->ctrl + 2);
    }
  }
};
template <class M, int TI, int WY> __global__ void __launch_bounds__(BX* WY) k_advect_smoke(M m, AdvSmokeArgs a, int ntiles) {
  run_tile<M, AdvSmokePass<TI, WY>>(m, a, ntiles);
}
int launch_advect_smoke(Ctx* c, const float* Msrc, float* Mo, float dt, bool write_new) {
  const AdvSmokeArgs a{role_of(c, Msrc), role_of(c, Mo), dt, c->prm.smoke_advection, c->prm.viscosity_diffusion,
                       (int)write_new};
  FL_ADV_LAUNCH(k_advect_smoke, c, a);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// advectSmokeBFECC passes 2 and 3 (fluid.go:1013-1046): the original smoke staged with halo 1
// (centre and 3x3 clamp range), the forward result with halo (HI, HJ) (backward sample); the velocity
// as in advectSmoke.
struct CorrSmokeArgs {
  int rMo, rMf, rMc;
  float dt, smokeAdv;
};
template <int TI_, int WY> struct CorrSmokePass {
  static constexpr int TI = TI_, NR = TI / WY, R = TI + 2 * HI, C = BX + 2 * HJ, R1 = TI + 2, C1 = BX + 2;
  using Args = CorrSmokeArgs;
  struct Lds {
    float Mf[Stage<R, C, WY>::PAD], Mo[Stage<R1, C1, WY>::PAD];
  };
  struct Pre {
    Stage<R1, C1, WY> mo;
    Stage<R, C, WY> mf;
  };
  struct Rows {
    uint32_t f;  // the NR rows' flags, a byte each
    CellVel<NR, WY> vel;
  };
  static __device__ __forceinline__ void prefetch(const TileRef& r, const Args& a, Pre& p, Rows& w) {
    const int NX = r.d->g.NX, NY = r.d->g.NY;
    w.f = row_flags<NR, WY>(r);
    w.vel.load(r);
    p.mo.load(r.d->buf[a.rMo], r.i0 - 1, r.j0 - 1, NX, NY);
    p.mf.load(r.d->buf[a.rMf], r.i0 - HI, r.j0 - HJ, NX, NY);
  }
  static __device__ __forceinline__ void store(const Pre& p, Lds& l) {
    p.mo.store(l.Mo);
    p.mf.store(l.Mf);
  }
  static __device__ __forceinline__ void compute(const TileRef& r, const Args& a, const Lds& l, const Rows& w) {
    const SimDesc* d = r.d;
    const Grid g = d->g;
    const int NX = g.NX, NY = g.NY, n = NY;
    const float dt = a.dt;
    const Win<R1, C1> wM{l.Mo, r.i0 - 1, r.j0 - 1};
    const WinSrc<R, C> srcM{Win<R, C>{l.Mf, r.i0 - HI, r.j0 - HJ}, d->buf[a.rMf]};
    float* Mc = d->buf[a.rMc];
    const int j = r.j0 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < NR; k++) {
      const int i = r.i0 + threadIdx.y + k * WY;
      if (i >= NX || j >= NY) continue;
      const int c = i * n + j;
      const bool ring = i == 0 || i == NX - 1 || j == 0 || j == NY - 1;
      const uint32_t f = (w.f >> (8 * k)) & 0xffu;
      const bool adv = !ring && (f & CF_FLUID);
      const float om = wM.at(i, j);
      float u = (w.vel.u0[k] + w.vel.u1[k]) * 0.5f * a.smokeAdv;
      float v = (w.vel.v0[k] + w.vel.v1[k]) * 0.5f * a.smokeAdv;
      const NbRange nr = neighbor_range(wM, i, j);  // used at interior cells only
      int badl = 0;
      float x = (float)i * g.h + g.hs + dt * u;
      float y = (float)j * g.h + g.hs + dt * v;
      const float bs = g.precise ? sample_cells(i, j, (dt * u) * g.h1, (dt * v) * g.h1, srcM, true, true, g, &badl)
                                 : sample_field(x, y, srcM, g.hs, g.hs, g, &badl);
      const float bm = adv ? bs : 0.0f;  // solid interior cells: the backward buffer's zero
      float cm = clamp_to(om - (bm - om) * 0.5f, nr);
      if (cm < 0.0f) cm = 0.0f;
      st_cell(Mc, c, ring ? om : cm);
      if (adv && badl) flag_invalid(d->ctrl + 2);
    }
  }
};
template <class M, int TI, int WY> __global__ void __launch_bounds__(BX* WY) k_bfecc_smoke_corr(M m, CorrSmokeArgs a, int ntiles) {
  run_tile<M, CorrSmokePass<TI, WY>>(m, a, ntiles);
}
int launch_bfecc_smoke_corr(Ctx* c, const float* Morig, const float* Mfwd, float* Mcorr, float dt) {
  const CorrSmokeArgs a{role_of(c, Morig), role_of(c, Mfwd), role_of(c, Mcorr), dt, c->prm.smoke_advection};
  FL_ADV_LAUNCH(k_bfecc_smoke_corr, c, a);
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
  const int nt = B->ncell_items;
  hipLaunchKernelGGL((k_advect_velocity<ManySims, BY, BY>), dim3(nt),
                     dim3(BX, BY), 0, B->stream, many(B), AdvVelArgs{rUt, rVt, rUo, rVo, dt, (int)write_new}, nt);
  FL_HIP(hipGetLastError());
  return FL_OK;
}
int batch_bfecc_velocity_corr(Batch* B, int rUo, int rVo, int rUf, int rVf, float dt) {
  const int nt = B->ncell_items;
  hipLaunchKernelGGL((k_bfecc_velocity_corr<ManySims, BY, BY>), dim3(nt),
                     dim3(BX, BY), 0, B->stream, many(B), CorrVelArgs{rUo, rVo, rUf, rVf, B_T0, B_T1, dt}, nt);
  FL_HIP(hipGetLastError());
  return FL_OK;
}
int batch_advect_smoke(Batch* B, int rMs, int rMo, float dt, bool write_new) {
  const int nt = B->ncell_items;
  hipLaunchKernelGGL((k_advect_smoke<ManySims, BY, BY>), dim3(nt), dim3(BX, BY), 0,
                     B->stream, many(B),
                     AdvSmokeArgs{rMs, rMo, dt, B->prm.smoke_advection, B->prm.viscosity_diffusion, (int)write_new}, nt);
  FL_HIP(hipGetLastError());
  return FL_OK;
}
int batch_bfecc_smoke_corr(Batch* B, float dt) {
  const int nt = B->ncell_items;
  hipLaunchKernelGGL((k_bfecc_smoke_corr<ManySims, BY, BY>), dim3(nt),
                     dim3(BX, BY), 0, B->stream, many(B), CorrSmokeArgs{B_M, B_M2, B_T2, dt, B->prm.smoke_advection},
                     nt);
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
