// Differentiable rollouts: the hand-written adjoint (vector-Jacobian product) of the step.
//
// No reference counterpart: the reference has no gradients. The rollout is the step as
// this library runs it: per step the scene's sources and jet (Game.Update, main.go:231-241),
// then U,V += F on fluid interior cells (an external force field with ApplyForce's semantics,
// fluid.go:761-772), then Simulate (fluid.go:79-109). The loss after the last step is
// sum over all cells of (M - target)^2; fl_vjp_backward returns dL/dF.
//
// Forward: the step's own kernels; every `every`-th step's state is checkpointed.
// Backward: per segment, the checkpoint is restored and the segment recomputed with a
// tape of the step's intermediate fields, then each step is reversed:
//   smoke BFECC / SL       scatter adjoint of the bilinear samples (atomics), the trace's
//                          velocity dependence, the diffusion term, the 3x3 clamp and
//                          max(.,0) branches
//   velocity BFECC / SL    same, with the back-trace and final pass tracing the field they
//                          sample
//   borders, turbulence, confinement (curl, normalised grad |curl|, speed factor)
//   Gauss-Seidel sweeps    transposed and run backwards. With face weights w (1 where both
//                          adjacent cells are fluid), mu = w * lambda evolves under the
//                          transposed sweeps exactly as the forward sweeps evolve (U,V) on
//                          the point-mirrored grid with the sweeps' relaxation reversed, so
//                          the temporal-blocked solver itself runs the adjoint; the p output
//                          of that run (cp = 1) is each cell's sum of updates, which gives
//                          the adjoint of the faces next to solid cells.
//   viscosity              the Laplacian step transposed (a gather); faces the advection
//                          does not advect keep the viscosity's newU/newV, so their adjoint
//                          joins it there.
//   sources and force      the force's gradient is the adjoint entering Simulate; cells set
//                          by SetVelocity lose theirs.
//   multigrid V-cycle      with its exit sweep fixed, linear in (U,V,p): smoothing sweeps,
//                          correction, prolongation, coarse sweeps, restriction and residual,
//                          each transposed and run in reverse (adj_multigrid below).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "fl_internal.h"

namespace fl {

int gs_wave_solve(Ctx* c, int numIters, float cp);  // fl_gs_wave.hip

static constexpr int BX = 64, BY = 4;
static inline dim3 grid_of(const Ctx* c) { return dim3((c->NY + BX - 1) / BX, (c->NX + BY - 1) / BY); }

// fields of one step's tape
enum TapeField { T_U2 = 0, T_V2, T_U3, T_V3, T_U5, T_V5, T_FU, T_FV, T_CU, T_CV, T_U6, T_V6, T_M1, T_FM, T_CM, NTAPE };
// adjoint work fields
enum LamField {
  L_U = 0, L_V, L_M, L_CU, L_CV, L_FU, L_FV, L_OU, L_OV, L_CM, L_FM, L_M1, L_CURL, L_LCD, L_GX, L_GY, L_LCURL,
  L_GFU, L_GFV, L_SU, L_SV, NLAM
};
// saved state: U, V, newU, newV, p, M, newM
static constexpr int NSTATE = 7;

struct Vjp {
  Ctx* c = nullptr;
  Ctx* mir = nullptr;            // point-mirrored grid for the transposed sweeps
  hipStream_t mir_stream = nullptr;
  int steps = 0, every = 0, nck = 0;
  size_t N = 0;
  float* Fu = nullptr;
  float* Fv = nullptr;
  float* target = nullptr;
  float* x0 = nullptr;           // [NSTATE][N] initial state
  float* fin = nullptr;          // [NSTATE][N] final state of the last forward
  float* ck = nullptr;           // [nck][NSTATE][N]
  float* tape = nullptr;         // [every][NTAPE][N]
  float* lam_mem = nullptr;      // [NLAM][N]
  float* lam[NLAM] = {};
  double* d_loss = nullptr;
  float dt = 0.0f;
  bool have_forward = false;
  int mir_K = -1;
  std::vector<int> sweeps;
};

static int verr(int code, const std::string& m) {
  set_error(m);
  return code;
}

// ---- small kernels -------------------------------------------------------------------------
__global__ void k_force_add(Grid g, const uint8_t* __restrict__ flags, float* __restrict__ U, float* __restrict__ V,
                            const float* __restrict__ Fu, const float* __restrict__ Fv) {
  const int j = blockIdx.x * BX + threadIdx.x, i = blockIdx.y * BY + threadIdx.y;
  if (i >= g.NX || j >= g.NY) return;
  const int c = i * g.NY + j;
  const uint32_t f = flags[c];
  if ((f & CF_INTERIOR) && (f & CF_FLUID)) {
    U[c] = U[c] + Fu[c];
    V[c] = V[c] + Fv[c];
  }
}

// loss = sum (M - T)^2 in double with a fixed summation tree (run to run identical),
// lamM = 2 (M - T). Stage 1: one partial per block; stage 2: one block sums them.
static constexpr int LOSS_BLOCKS = 1024, LOSS_THREADS = 256;
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double red[LOSS_THREADS];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int o = LOSS_THREADS / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  return red[0];
}
__global__ __launch_bounds__(LOSS_THREADS) void k_loss(const float* __restrict__ M, const float* __restrict__ T,
                                                       float* __restrict__ lamM, size_t n,
                                                       double* __restrict__ partial) {
  double acc = 0.0;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
    const float d = M[k] - T[k];
    lamM[k] = 2.0f * d;
    acc += (double)d * (double)d;
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}
__global__ __launch_bounds__(LOSS_THREADS) void k_loss_final(const double* __restrict__ partial, double* out) {
  double acc = 0.0;
  for (int k = threadIdx.x; k < LOSS_BLOCKS; k += LOSS_THREADS) acc += partial[k];
  const double s = block_sum(acc);
  if (threadIdx.x == 0) *out = s;
}

__device__ __forceinline__ void aadd(float* p, float v) {
  if (v != 0.0f) unsafeAtomicAdd(p, v);
}

// Clamped bilinear sample (fluid.go:357-398) with its corner weights and its partial
// derivatives in x and y (0 where the position was clamped to [h, N*h]).
struct SGeo {
  int k[4];
  float w[4];
  float val, ddx, ddy;
};
__device__ __forceinline__ SGeo sample_geo(float x, float y, const float* __restrict__ f, float dx, float dy,
                                           const Grid& g) {
  const float mx = (x >= g.h && x <= g.xmax) ? 1.0f : 0.0f;
  const float my = (y >= g.h && y <= g.ymax) ? 1.0f : 0.0f;
  x = go_maxf(go_minf(x, g.xmax), g.h);
  y = go_maxf(go_minf(y, g.ymax), g.h);
  int bad = 0;
  int x0 = max(min(floor_index((x - dx) * g.h1, &bad), g.NX - 1), 0);
  float tx = ((x - dx) - (float)x0 * g.h) * g.h1;
  int x1 = min(x0 + 1, g.NX - 1);
  int y0 = max(min(floor_index((y - dy) * g.h1, &bad), g.NY - 1), 0);
  float ty = ((y - dy) - (float)y0 * g.h) * g.h1;
  int y1 = min(y0 + 1, g.NY - 1);
  const float sx = 1.0f - tx, sy = 1.0f - ty;
  const int n = g.NY;
  SGeo s;
  s.k[0] = x0 * n + y0;
  s.k[1] = x1 * n + y0;
  s.k[2] = x1 * n + y1;
  s.k[3] = x0 * n + y1;
  s.w[0] = sx * sy;
  s.w[1] = tx * sy;
  s.w[2] = tx * ty;
  s.w[3] = sx * ty;
  const float a = f[s.k[0]], b = f[s.k[1]], c = f[s.k[2]], d = f[s.k[3]];
  s.val = sx * sy * a + tx * sy * b + tx * ty * c + sx * ty * d;
  s.ddx = (sy * (b - a) + ty * (c - d)) * g.h1 * mx;
  s.ddy = (sx * (d - a) + tx * (c - b)) * g.h1 * my;
  return s;
}
// The same at (i + dX, j + dY) in cell units (precise_advection): the derivative in cells
// times h1, 0 on an axis whose position was clamped.
__device__ __forceinline__ SGeo sample_geo_cells(int i, int j, float dX, float dY, const float* __restrict__ f,
                                                 bool hx, bool hy, const Grid& g) {
  int bad = 0, x0, y0;
  float tx, ty;
  const float mx = axis_cells(i, dX, g.NX, hx, x0, tx, &bad) ? 1.0f : 0.0f;
  const float my = axis_cells(j, dY, g.NY, hy, y0, ty, &bad) ? 1.0f : 0.0f;
  const int x1 = min(x0 + 1, g.NX - 1), y1 = min(y0 + 1, g.NY - 1);
  const float sx = 1.0f - tx, sy = 1.0f - ty;
  const int n = g.NY;
  SGeo s;
  s.k[0] = x0 * n + y0;
  s.k[1] = x1 * n + y0;
  s.k[2] = x1 * n + y1;
  s.k[3] = x0 * n + y1;
  s.w[0] = sx * sy;
  s.w[1] = tx * sy;
  s.w[2] = tx * ty;
  s.w[3] = sx * ty;
  const float a = f[s.k[0]], b = f[s.k[1]], c = f[s.k[2]], d = f[s.k[3]];
  s.val = sx * sy * a + tx * sy * b + tx * ty * c + sx * ty * d;
  s.ddx = (sy * (b - a) + ty * (c - d)) * g.h1 * mx;
  s.ddy = (sx * (d - a) + tx * (c - b)) * g.h1 * my;
  return s;
}
// A back-traced sample from (i,j): (x,y) is the reference's absolute position, offset plus
// (du,dv) = the trace step times the velocity; precise_advection samples at (du,dv)*h1 cells.
__device__ __forceinline__ SGeo trace_geo(float x, float y, int i, int j, float du, float dv,
                                          const float* __restrict__ f, float dx, float dy, const Grid& g) {
  if (g.precise) return sample_geo_cells(i, j, du * g.h1, dv * g.h1, f, dx != 0.0f, dy != 0.0f, g);
  return sample_geo(x, y, f, dx, dy, g);
}
__device__ __forceinline__ void scatter(const SGeo& s, float lam, float* __restrict__ dst) {
  for (int q = 0; q < 4; q++) aadd(dst + s.k[q], lam * s.w[q]);
}
__device__ __forceinline__ float avg_u4(const float* __restrict__ U, int c, int n) {
  return (U[c - 1] + U[c] + U[c + n - 1] + U[c + n]) * 0.25f;
}
__device__ __forceinline__ float avg_v4(const float* __restrict__ V, int c, int n) {
  return (V[c - n] + V[c] + V[c - n + 1] + V[c + 1]) * 0.25f;
}
__device__ __forceinline__ bool adv_u(uint32_t f, int i, int j, int NY) {
  return i >= 1 && j >= 1 && j <= NY - 2 && (f & CF_FLUID) && (f & CF_X0);
}
__device__ __forceinline__ bool adv_v(uint32_t f, int i, int j, int NX) {
  return i >= 1 && j >= 1 && i <= NX - 2 && (f & CF_FLUID) && (f & CF_Y0);
}

// Adjoint of a semi-Lagrangian velocity pass (advectVelocity, fluid.go:291-333, or BFECC's
// backward trace): trace with (Ut,Vt) and step s, sample (Su,Sv). lO: adjoint of the output;
// the sampled field's adjoint is added to lS, the tracing field's to lT; `ring`: non-advected
// ring faces are copies of the input (copyBorder). stU/stV (null without viscosity): the
// other non-advected faces keep newU/newV, which applyViscosity set this step; their adjoint
// is added there (each thread writes only its own face).
__global__ void k_adj_adv_vel(Grid g, const uint8_t* __restrict__ flags, const float* __restrict__ Ut,
                              const float* __restrict__ Vt, const float* __restrict__ Su,
                              const float* __restrict__ Sv, float s, const float* __restrict__ lOU,
                              const float* __restrict__ lOV, float* lSU, float* lSV, float* lTU, float* lTV,
                              int ring, float* __restrict__ stU, float* __restrict__ stV) {
  const int j = blockIdx.x * BX + threadIdx.x, i = blockIdx.y * BY + threadIdx.y;
  const int NX = g.NX, NY = g.NY, n = NY;
  if (i >= NX || j >= NY) return;
  const int c = i * n + j;
  const uint32_t f = flags[c];
  const bool onring = i == 0 || j == 0 || i == NX - 1 || j == NY - 1;
  if (adv_u(f, i, j, NY)) {
    const float lam = lOU[c];
    if (lam != 0.0f) {
      float x = (float)i * g.h;
      float y = (float)j * g.h + g.ha;
      const float du = s * Ut[c], dv = s * avg_v4(Vt, c, n);
      x = x + du;
      y = y + dv;
      const SGeo q = trace_geo(x, y, i, j, du, dv, Su, 0.0f, g.hs, g);
      scatter(q, lam, lSU);
      aadd(lTU + c, lam * q.ddx * s);
      const float gy = lam * q.ddy * s * 0.25f;
      aadd(lTV + c - n, gy);
      aadd(lTV + c, gy);
      aadd(lTV + c - n + 1, gy);
      aadd(lTV + c + 1, gy);
    }
  } else if (ring && onring) {
    aadd(lSU + c, lOU[c]);
  } else if (ring && stU) {
    stU[c] += lOU[c];
  }
  if (adv_v(f, i, j, NX)) {
    const float lam = lOV[c];
    if (lam != 0.0f) {
      float x = (float)i * g.h + g.ha;
      float y = (float)j * g.h;
      const float du = s * avg_u4(Ut, c, n), dv = s * Vt[c];
      x = x + du;
      y = y + dv;
      const SGeo q = trace_geo(x, y, i, j, du, dv, Sv, g.hs, 0.0f, g);
      scatter(q, lam, lSV);
      const float gx = lam * q.ddx * s * 0.25f;
      aadd(lTU + c - 1, gx);
      aadd(lTU + c, gx);
      aadd(lTU + c + n - 1, gx);
      aadd(lTU + c + n, gx);
      aadd(lTV + c, lam * q.ddy * s);
    }
  } else if (ring && onring) {
    aadd(lSV + c, lOV[c]);
  } else if (ring && stV) {
    stV[c] += lOV[c];
  }
}

// Adjoint of a smoke pass (advectSmoke, fluid.go:400-434; or BFECC's backward trace):
// trace with the face-averaged (U,V) * SA and step s, sample Sm, max(.,0) when `relu`.
// `ring`: the ring is a copy of the input. `diffuse`: advectSmoke's diffusion
// sd * (4-neighbour Laplacian of Sm), sd = ViscosityDiffusion * 0.3 * dt, is added before the
// max, in the forward's order so the clamp takes the same branch.
__global__ void k_adj_adv_smoke(Grid g, const uint8_t* __restrict__ flags, const float* __restrict__ U,
                                const float* __restrict__ V, const float* __restrict__ Sm, float s, float SA,
                                int relu, const float* __restrict__ lO, float* lS, float* lU, float* lV, int ring,
                                int diffuse, float sd) {
  const int j = blockIdx.x * BX + threadIdx.x, i = blockIdx.y * BY + threadIdx.y;
  const int NX = g.NX, NY = g.NY, n = NY;
  if (i >= NX || j >= NY) return;
  const int c = i * n + j;
  const uint32_t f = flags[c];
  if ((f & CF_INTERIOR) && (f & CF_FLUID)) {
    const float lam = lO[c];
    if (lam == 0.0f) return;
    const float u = (U[c] + U[c + n]) * 0.5f * SA;
    const float v = (V[c] + V[c + 1]) * 0.5f * SA;
    const float x = (float)i * g.h + g.hs + s * u;
    const float y = (float)j * g.h + g.hs + s * v;
    const SGeo q = trace_geo(x, y, i, j, s * u, s * v, Sm, g.hs, g.hs, g);
    float val = q.val;
    if (diffuse) {
      const float nb = Sm[c - n] + Sm[c + n] + Sm[c - 1] + Sm[c + 1] - 4.0f * Sm[c];
      val += sd * nb;
    }
    if (relu && !(val > 0.0f)) return;
    scatter(q, lam, lS);
    if (diffuse) {
      const float ls = lam * sd;
      aadd(lS + c - n, ls);
      aadd(lS + c + n, ls);
      aadd(lS + c - 1, ls);
      aadd(lS + c + 1, ls);
      aadd(lS + c, ls * -4.0f);
    }
    const float gu = lam * q.ddx * s * 0.5f * SA, gv = lam * q.ddy * s * 0.5f * SA;
    aadd(lU + c, gu);
    aadd(lU + c + n, gu);
    aadd(lV + c, gv);
    aadd(lV + c + 1, gv);
  } else if (ring && (i == 0 || j == 0 || i == NX - 1 || j == NY - 1)) {
    aadd(lS + c, lO[c]);
  }
}

// clampToNeighbors (fluid.go:1094-1120) with the index of the bound it returns, or -1.
__device__ __forceinline__ int clamp_src(float val, const float* __restrict__ src, int c, int n, float* out) {
  float lo = src[c], hi = src[c];
  int klo = c, khi = c;
#pragma unroll
  for (int di = -1; di <= 1; di++)
#pragma unroll
    for (int dj = -1; dj <= 1; dj++) {
      const int k = c + di * n + dj;
      const float v = src[k];
      if (v < lo) {
        lo = v;
        klo = k;
      }
      if (v > hi) {
        hi = v;
        khi = k;
      }
    }
  if (val < lo) {
    *out = lo;
    return klo;
  }
  if (val > hi) {
    *out = hi;
    return khi;
  }
  *out = val;
  return -1;
}

// BFECC velocity passes 2-3 transposed (fluid.go:935-987): for each interior face, the
// clamp branch of corr = orig - (bwd - orig)*0.5, then the backward trace's adjoint
// (sampling fwd, tracing with orig at +dt). Ring: corr = orig.
__global__ void k_adj_corr_vel(Grid g, const uint8_t* __restrict__ flags, const float* __restrict__ oU,
                               const float* __restrict__ oV, const float* __restrict__ fU,
                               const float* __restrict__ fV, float dt, const float* __restrict__ lCU,
                               const float* __restrict__ lCV, float* lOU, float* lOV, float* lFU, float* lFV) {
  const int j = blockIdx.x * BX + threadIdx.x, i = blockIdx.y * BY + threadIdx.y;
  const int NX = g.NX, NY = g.NY, n = NY;
  if (i >= NX || j >= NY) return;
  const int c = i * n + j;
  const uint32_t f = flags[c];
  if (!(f & CF_INTERIOR)) {
    aadd(lOU + c, lCU[c]);
    aadd(lOV + c, lCV[c]);
    return;
  }
  {
    const float lam = lCU[c];
    if (lam != 0.0f) {
      const bool a = adv_u(f, i, j, NY);
      float bwd = 0.0f;
      SGeo q;
      if (a) {
        float x = (float)i * g.h;
        float y = (float)j * g.h + g.ha;
        const float du = dt * oU[c], dv = dt * avg_v4(oV, c, n);
        x = x + du;
        y = y + dv;
        q = trace_geo(x, y, i, j, du, dv, fU, 0.0f, g.hs, g);
        bwd = q.val;
      }
      const float o = oU[c];
      float val;
      const int src = clamp_src(o - (bwd - o) * 0.5f, oU, c, n, &val);
      if (src >= 0) {
        aadd(lOU + src, lam);
      } else {
        aadd(lOU + c, 1.5f * lam);
        const float lb = -0.5f * lam;
        if (a) {
          scatter(q, lb, lFU);
          aadd(lOU + c, lb * q.ddx * dt);
          const float gy = lb * q.ddy * dt * 0.25f;
          aadd(lOV + c - n, gy);
          aadd(lOV + c, gy);
          aadd(lOV + c - n + 1, gy);
          aadd(lOV + c + 1, gy);
        }
      }
    }
  }
  {
    const float lam = lCV[c];
    if (lam != 0.0f) {
      const bool a = adv_v(f, i, j, NX);
      float bwd = 0.0f;
      SGeo q;
      if (a) {
        float x = (float)i * g.h + g.ha;
        float y = (float)j * g.h;
        const float du = dt * avg_u4(oU, c, n), dv = dt * oV[c];
        x = x + du;
        y = y + dv;
        q = trace_geo(x, y, i, j, du, dv, fV, g.hs, 0.0f, g);
        bwd = q.val;
      }
      const float o = oV[c];
      float val;
      const int src = clamp_src(o - (bwd - o) * 0.5f, oV, c, n, &val);
      if (src >= 0) {
        aadd(lOV + src, lam);
      } else {
        aadd(lOV + c, 1.5f * lam);
        const float lb = -0.5f * lam;
        if (a) {
          scatter(q, lb, lFV);
          const float gx = lb * q.ddx * dt * 0.25f;
          aadd(lOU + c - 1, gx);
          aadd(lOU + c, gx);
          aadd(lOU + c + n - 1, gx);
          aadd(lOU + c + n, gx);
          aadd(lOV + c, lb * q.ddy * dt);
        }
      }
    }
  }
}

// BFECC smoke passes 2-3 transposed (fluid.go:1017-1046), with the clamp's max(.,0).
__global__ void k_adj_corr_smoke(Grid g, const uint8_t* __restrict__ flags, const float* __restrict__ M1,
                                 const float* __restrict__ fM, const float* __restrict__ U,
                                 const float* __restrict__ V, float dt, float SA, const float* __restrict__ lC,
                                 float* lM1, float* lF, float* lU, float* lV) {
  const int j = blockIdx.x * BX + threadIdx.x, i = blockIdx.y * BY + threadIdx.y;
  const int NX = g.NX, NY = g.NY, n = NY;
  if (i >= NX || j >= NY) return;
  const int c = i * n + j;
  const uint32_t f = flags[c];
  if (!(f & CF_INTERIOR)) {
    aadd(lM1 + c, lC[c]);
    return;
  }
  const float lam = lC[c];
  if (lam == 0.0f) return;
  const bool a = (f & CF_FLUID) != 0;
  float bwd = 0.0f;
  SGeo q;
  if (a) {
    const float u = (U[c] + U[c + n]) * 0.5f * SA;
    const float v = (V[c] + V[c + 1]) * 0.5f * SA;
    const float x = (float)i * g.h + g.hs + dt * u;
    const float y = (float)j * g.h + g.hs + dt * v;
    q = trace_geo(x, y, i, j, dt * u, dt * v, fM, g.hs, g.hs, g);
    bwd = q.val;
  }
  const float o = M1[c];
  float val;
  const int src = clamp_src(o - (bwd - o) * 0.5f, M1, c, n, &val);
  if (val < 0.0f) return;
  if (src >= 0) {
    aadd(lM1 + src, lam);
    return;
  }
  aadd(lM1 + c, 1.5f * lam);
  if (a) {
    const float lb = -0.5f * lam;
    scatter(q, lb, lF);
    const float gu = lb * q.ddx * dt * 0.5f * SA, gv = lb * q.ddy * dt * 0.5f * SA;
    aadd(lU + c, gu);
    aadd(lU + c + n, gu);
    aadd(lV + c, gv);
    aadd(lV + c + 1, gv);
  }
}

// handleBorders (fluid.go:236-289) transposed, in place; thread t handles row i = t of U
// and column j = t of V.
__global__ void k_adj_borders(Grid g, const float* __restrict__ S, float* lU, float* lV) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int NX = g.NX, NY = g.NY, n = NY;
  if (t < NX) {
    const int i = t;
    const float g0 = lU[i * n + 0], g1 = lU[i * n + NY - 1];
    lU[i * n + 0] = 0.0f;
    lU[i * n + NY - 1] = 0.0f;
    if (!(S[i * n + 0] == 0.0f || S[i * n + 1] == 0.0f)) {
      if (i > 0 && i < NX - 1 && S[i * n + 2] > 0.0f) {
        lU[i * n + 1] += 2.0f * g0;
        lU[i * n + 2] -= g0;
      } else {
        lU[i * n + 1] += g0;
      }
    }
    if (!(S[i * n + NY - 1] == 0.0f || S[i * n + NY - 2] == 0.0f)) {
      if (i > 0 && i < NX - 1 && S[i * n + NY - 3] > 0.0f) {
        lU[i * n + NY - 2] += 2.0f * g1;
        lU[i * n + NY - 3] -= g1;
      } else {
        lU[i * n + NY - 2] += g1;
      }
    }
  }
  if (t < NY) {
    const int j = t;
    const float g0 = lV[j], g1 = lV[(NX - 1) * n + j];
    lV[j] = 0.0f;
    lV[(NX - 1) * n + j] = 0.0f;
    if (!(S[j] == 0.0f || S[n + j] == 0.0f)) {
      if (j > 0 && j < NY - 1 && S[2 * n + j] > 0.0f) {
        lV[n + j] += 2.0f * g0;
        lV[2 * n + j] -= g0;
      } else {
        lV[n + j] += g0;
      }
    }
    if (!(S[(NX - 1) * n + j] == 0.0f || S[(NX - 2) * n + j] == 0.0f)) {
      if (j > 0 && j < NY - 1 && S[(NX - 3) * n + j] > 0.0f) {
        lV[(NX - 2) * n + j] += 2.0f * g1;
        lV[(NX - 3) * n + j] -= g1;
      } else {
        lV[(NX - 2) * n + j] += g1;
      }
    }
  }
}

// addTurbulence (fluid.go:496-526) transposed at its input (U,V), in place.
__global__ void k_adj_turb(Grid g, const uint8_t* __restrict__ flags, const float* __restrict__ U,
                           const float* __restrict__ V, float ts, float* lU, float* lV) {
  const int j = blockIdx.x * BX + threadIdx.x, i = blockIdx.y * BY + threadIdx.y;
  if (i >= g.NX || j >= g.NY) return;
  const int n = g.NY, c = i * n + j;
  const uint32_t f = flags[c];
  if (!((f & CF_INTERIOR) && (f & CF_FLUID))) return;
  const float u = U[c], v = V[c];
  const float lv = sqrtf(u * u + v * v);
  if (!(lv > 0.1f) || !(lv * 0.5f < 1.0f)) return;
  const float nu = turb_noise(i * 137 + j * 241) * ts;
  const float nv = turb_noise(i * 157 + j * 263) * ts;
  const float d = (lU[c] * nu + lV[c] * nv) * 0.5f / lv;
  lU[c] += d * u;
  lV[c] += d * v;
}

// applyVorticityConfinement (fluid.go:449-493) transposed at its input (U,V).
__global__ void k_adj_curl(Grid g, const uint8_t* __restrict__ flags, const float* __restrict__ U,
                           const float* __restrict__ V, float* __restrict__ curl) {
  const int j = blockIdx.x * BX + threadIdx.x, i = blockIdx.y * BY + threadIdx.y;
  if (i >= g.NX || j >= g.NY) return;
  const int n = g.NY, c = i * n + j;
  const uint32_t f = flags[c];
  curl[c] = ((f & CF_INTERIOR) && (f & CF_FLUID))
                ? (V[c + n] - V[c - n]) * 0.5f / g.h - (U[c + 1] - U[c - 1]) * 0.5f / g.h
                : 0.0f;
}
// per cell: its own (U,V) terms in place, the direct curl adjoint, and dL/d(grad|curl|) * 0.5/h
__global__ void k_adj_conf_a(Grid g, const uint8_t* __restrict__ flags, const float* __restrict__ U,
                             const float* __restrict__ V, const float* __restrict__ curl, float C, float dt,
                             float* lU, float* lV, float* __restrict__ lcd, float* __restrict__ Gx,
                             float* __restrict__ Gy) {
  const int j = blockIdx.x * BX + threadIdx.x, i = blockIdx.y * BY + threadIdx.y;
  if (i >= g.NX || j >= g.NY) return;
  const int n = g.NY, c = i * n + j;
  const uint32_t f = flags[c];
  if (!((f & CF_INTERIOR) && (f & CF_FLUID))) {
    lcd[c] = 0.0f;
    Gx[c] = 0.0f;
    Gy[c] = 0.0f;
    return;
  }
  const float h = g.h, eps = 1e-5f;
  const float gx = (fabsf(curl[c + n]) - fabsf(curl[c - n])) * 0.5f / h;
  const float gy = (fabsf(curl[c + 1]) - fabsf(curl[c - 1])) * 0.5f / h;
  const float r = sqrtf(gx * gx + gy * gy), mag = r + eps;
  const float nx = gx / mag, ny = gy / mag;
  const float vort = curl[c];
  const float u = U[c], v = V[c];
  const float lv = sqrtf(u * u + v * v);
  const float str = C * (1.0f + lv * 0.1f);
  const float a = lU[c], b = lV[c];
  const float dstr = (a * ny - b * nx) * vort * dt;
  if (lv > 0.0f) {
    lU[c] = a + dstr * C * 0.1f * u / lv;
    lV[c] = b + dstr * C * 0.1f * v / lv;
  }
  lcd[c] = (a * ny - b * nx) * str * dt;
  const float dny = a * str * vort * dt, dnx = -b * str * vort * dt;
  float dgx = dnx / mag, dgy = dny / mag;
  if (r > 0.0f) {
    const float k = (dnx * gx + dny * gy) / (mag * mag * r);
    dgx -= k * gx;
    dgy -= k * gy;
  }
  Gx[c] = dgx * 0.5f / h;
  Gy[c] = dgy * 0.5f / h;
}
__device__ __forceinline__ float sgnf(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
__global__ void k_adj_conf_b(Grid g, const uint8_t* __restrict__ flags, const float* __restrict__ curl,
                             const float* __restrict__ lcd, const float* __restrict__ Gx,
                             const float* __restrict__ Gy, float* __restrict__ lcurl) {
  const int j = blockIdx.x * BX + threadIdx.x, i = blockIdx.y * BY + threadIdx.y;
  if (i >= g.NX || j >= g.NY) return;
  const int n = g.NY, c = i * n + j;
  const uint32_t f = flags[c];
  if (!((f & CF_INTERIOR) && (f & CF_FLUID))) {
    lcurl[c] = 0.0f;
    return;
  }
  lcurl[c] = lcd[c] + sgnf(curl[c]) * (Gx[c - n] - Gx[c + n] + Gy[c - 1] - Gy[c + 1]);
}
__global__ void k_adj_conf_c(Grid g, const float* __restrict__ lcurl, float* lU, float* lV) {
  const int j = blockIdx.x * BX + threadIdx.x, i = blockIdx.y * BY + threadIdx.y;
  const int NX = g.NX, NY = g.NY, n = NY;
  if (i >= NX || j >= NY) return;
  const int c = i * n + j;
  const float k = 0.5f / g.h;
  const float up = j + 1 < NY ? lcurl[c + 1] : 0.0f, dn = j >= 1 ? lcurl[c - 1] : 0.0f;
  const float rt = i + 1 < NX ? lcurl[c + n] : 0.0f, lf = i >= 1 ? lcurl[c - n] : 0.0f;
  lU[c] += up * k - dn * k;
  lV[c] += lf * k - rt * k;
}

// Transposed sweeps, part 1: mu = w * lambda written point-mirrored into the mirror grid's
// U, V (with p = 0). Mirrored faces: U'[i',j'] = mu_U[NX-i', NY-1-j'], V'[i',j'] =
// mu_V[NX-1-i', NY-j'] (index 0 of the shifted axis is a face no cell updates: 0).
__global__ void k_mirror_in(Grid g, const float* __restrict__ S, const float* __restrict__ lU,
                            const float* __restrict__ lV, float* __restrict__ Um, float* __restrict__ Vm,
                            float* __restrict__ Pm) {
  const int j = blockIdx.x * BX + threadIdx.x, i = blockIdx.y * BY + threadIdx.y;
  const int NX = g.NX, NY = g.NY, n = NY;
  if (i >= NX || j >= NY) return;
  const int cm = i * n + j;  // (i', j') in the mirror grid
  float u = 0.0f, v = 0.0f;
  if (i >= 1) {
    const int fi = NX - i, fj = NY - 1 - j, f = fi * n + fj;
    if (S[f] * S[f - n] != 0.0f) u = lU[f];
  }
  if (j >= 1) {
    const int fi = NX - 1 - i, fj = NY - j, f = fi * n + fj;
    if (S[f] * S[f - 1] != 0.0f) v = lV[f];
  }
  Um[cm] = u;
  Vm[cm] = v;
  Pm[cm] = 0.0f;
}
// Part 2: faces between two fluid cells take the mirrored result; the others collect the
// summed updates (mirror p) of their one adjacent updated cell: +p of the cell whose left
// or bottom face it is, -p of the cell whose right or top face it is.
__global__ void k_mirror_out(Grid g, const float* __restrict__ S, const float* __restrict__ Um,
                             const float* __restrict__ Vm, const float* __restrict__ Pm, float* lU, float* lV) {
  const int j = blockIdx.x * BX + threadIdx.x, i = blockIdx.y * BY + threadIdx.y;
  const int NX = g.NX, NY = g.NY, n = NY;
  if (i >= NX || j >= NY) return;
  const int c = i * n + j;
  auto pm = [&](int ci, int cj) { return Pm[(NX - 1 - ci) * n + (NY - 1 - cj)]; };
  if (i >= 1) {
    if (S[c] * S[c - n] != 0.0f)
      lU[c] = Um[(NX - i) * n + (NY - 1 - j)];
    else
      lU[c] += pm(i, j) - pm(i - 1, j);
  }
  if (j >= 1) {
    if (S[c] * S[c - 1] != 0.0f)
      lV[c] = Vm[(NX - 1 - i) * n + (NY - j)];
    else
      lV[c] += pm(i, j) - pm(i, j - 1);
  }
}

// Force and sources transposed: dL/dF += lambda on fluid interior cells, then the cells
// written by SetVelocity lose their adjoint.
// applyViscosity (fluid.go:112-142) transposed, as a gather: on fluid interior cells the
// forward sets new = u + visc * (sum of the 4 neighbours - 4u), elsewhere new = u. The
// adjoint of new is lO + st (st: faces that kept newU/newV through the advection).
__global__ void k_adj_visc(Grid g, const uint8_t* __restrict__ flags, float visc, const float* __restrict__ lOU,
                           const float* __restrict__ lOV, const float* __restrict__ stU,
                           const float* __restrict__ stV, float* __restrict__ oU, float* __restrict__ oV) {
  const int j = blockIdx.x * BX + threadIdx.x, i = blockIdx.y * BY + threadIdx.y;
  const int NX = g.NX, NY = g.NY, n = NY;
  if (i >= NX || j >= NY) return;
  const int c = i * n + j;
  auto upd = [&](int k) { return (flags[k] & CF_INTERIOR) && (flags[k] & CF_FLUID); };
  const float lu = lOU[c] + stU[c], lv = lOV[c] + stV[c];
  float ou = lu, ov = lv;
  if (upd(c)) {
    ou += lu * -4.0f * visc;
    ov += lv * -4.0f * visc;
  }
  const int nb[4] = {i > 0 ? c - n : -1, i < NX - 1 ? c + n : -1, j > 0 ? c - 1 : -1, j < NY - 1 ? c + 1 : -1};
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int k = nb[q];
    if (k >= 0 && upd(k)) {
      ou += (lOU[k] + stU[k]) * visc;
      ov += (lOV[k] + stV[k]) * visc;
    }
  }
  oU[c] = ou;
  oV[c] = ov;
}

__global__ void k_adj_force(Grid g, const uint8_t* __restrict__ flags, const float* __restrict__ lU,
                            const float* __restrict__ lV, float* __restrict__ gU, float* __restrict__ gV) {
  const int j = blockIdx.x * BX + threadIdx.x, i = blockIdx.y * BY + threadIdx.y;
  if (i >= g.NX || j >= g.NY) return;
  const int c = i * g.NY + j;
  const uint32_t f = flags[c];
  if ((f & CF_INTERIOR) && (f & CF_FLUID)) {
    gU[c] += lU[c];
    gV[c] += lV[c];
  }
}
__global__ void k_adj_sources(const SrcOp* __restrict__ ops, const int32_t* __restrict__ seg, int nseg,
                              const float* __restrict__ S, float* lU, float* lV) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  const int b = seg[s], e = seg[s + 1], cell = ops[b].cell;
  const bool fluid = S[cell] != 0.0f;
  bool set = false;
  for (int k = b; k < e; k++)
    if (ops[k].kind == 0 && !(ops[k].check_solid && !fluid)) set = true;
  if (set) {
    lU[cell] = 0.0f;
    lV[cell] = 0.0f;
  }
}

// ---- host side -------------------------------------------------------------------------
static float** state_ptrs(Ctx* c, float** out) {
  out[0] = c->U;
  out[1] = c->V;
  out[2] = c->nU;
  out[3] = c->nV;
  out[4] = c->p;
  out[5] = c->M;
  out[6] = c->nM;
  return out;
}
static int save_state(Vjp* J, float* dst) {
  float* p[NSTATE];
  state_ptrs(J->c, p);
  for (int k = 0; k < NSTATE; k++)
    FL_HIP(hipMemcpyAsync(dst + k * J->N, p[k], 4 * J->N, hipMemcpyDeviceToDevice, J->c->stream));
  return FL_OK;
}
static int load_state(Vjp* J, const float* src) {
  float* p[NSTATE];
  state_ptrs(J->c, p);
  for (int k = 0; k < NSTATE; k++)
    FL_HIP(hipMemcpyAsync(p[k], src + k * J->N, 4 * J->N, hipMemcpyDeviceToDevice, J->c->stream));
  return FL_OK;
}
static int dcopy(Vjp* J, float* dst, const float* src) {
  FL_HIP(hipMemcpyAsync(dst, src, 4 * J->N, hipMemcpyDeviceToDevice, J->c->stream));
  return FL_OK;
}

// One step of the rollout; with `tp` ([NTAPE][N]) the intermediate fields are recorded
// and the number of sweeps kept is returned in *kept (this synchronises).
static int vjp_step(Vjp* J, float dt, float* tp, int* kept) {
  Ctx* c = J->c;
  const fl_params& P = c->prm;
  const size_t N = J->N;
  int rc;
  if ((rc = ctx_prepare(c)) || (rc = ctx_build_sources(c)) || (rc = launch_apply_sources(c))) return rc;
  hipLaunchKernelGGL(k_force_add, grid_of(c), dim3(BX, BY), 0, c->stream, c->grid, c->flags, c->U, c->V, J->Fu,
                     J->Fv);
  FL_HIP(hipGetLastError());
  if ((rc = launch_fill(c->p, 0.0f, c->N, c->stream))) return rc;
  if (P.viscosity_diffusion > 0.0f && (rc = launch_viscosity(c, dt))) return rc;
  const bool mg = P.use_multigrid && P.multigrid_levels > 1;  // makeIncompressible (fluid.go:148-154)
  if ((rc = mg ? multigrid_solve(c, P.pressure_iters, dt) : pressure_solve(c, P.pressure_iters, dt))) return rc;
  if (tp && ((rc = dcopy(J, tp + T_U2 * N, c->U)) || (rc = dcopy(J, tp + T_V2 * N, c->V)))) return rc;
  const bool conf = P.confinement != 0.0f, turb = P.turbulence_strength > 0.0f;
  if (conf && ((rc = launch_curl(c, c->T0)) || (rc = launch_conf_turb(c, c->T0, dt, 1, 0)))) return rc;
  if (tp && ((rc = dcopy(J, tp + T_U3 * N, c->U)) || (rc = dcopy(J, tp + T_V3 * N, c->V)))) return rc;
  if (turb && (rc = launch_conf_turb(c, c->T0, dt, 0, 1))) return rc;
  if ((rc = launch_borders(c))) return rc;
  if (tp && ((rc = dcopy(J, tp + T_U5 * N, c->U)) || (rc = dcopy(J, tp + T_V5 * N, c->V)) ||
             (rc = dcopy(J, tp + T_M1 * N, c->M))))
    return rc;
  if (P.use_bfecc) {
    if ((rc = launch_advect_velocity(c, c->U, c->V, c->U2, c->V2, dt, false))) return rc;
    if (tp && ((rc = dcopy(J, tp + T_FU * N, c->U2)) || (rc = dcopy(J, tp + T_FV * N, c->V2)))) return rc;
    if ((rc = launch_bfecc_velocity_corr(c, c->U, c->V, c->U2, c->V2, c->T0, c->T1, dt))) return rc;
    if (tp && ((rc = dcopy(J, tp + T_CU * N, c->T0)) || (rc = dcopy(J, tp + T_CV * N, c->T1)))) return rc;
    if ((rc = launch_advect_velocity(c, c->T0, c->T1, c->U, c->V, dt))) return rc;
    if ((rc = launch_advect_smoke(c, c->M, c->M2, dt, false))) return rc;
    if (tp && (rc = dcopy(J, tp + T_FM * N, c->M2))) return rc;
    if ((rc = launch_bfecc_smoke_corr(c, c->M, c->M2, c->T2, dt))) return rc;
    if (tp && (rc = dcopy(J, tp + T_CM * N, c->T2))) return rc;
    if ((rc = launch_advect_smoke(c, c->T2, c->M, dt))) return rc;
  } else {
    if ((rc = launch_advect_velocity(c, c->U, c->V, c->U2, c->V2, dt))) return rc;
    std::swap(c->U, c->U2);
    std::swap(c->V, c->V2);
    if ((rc = launch_advect_smoke(c, c->M, c->M2, dt))) return rc;
    std::swap(c->M, c->M2);
  }
  if (tp) {
    if ((rc = dcopy(J, tp + T_U6 * N, c->U)) || (rc = dcopy(J, tp + T_V6 * N, c->V))) return rc;
    FL_HIP(hipStreamSynchronize(c->stream));
    int32_t ctrl[4];
    FL_HIP(hipMemcpy(ctrl, c->d_ctrl, 16, hipMemcpyDeviceToHost));
    *kept = c->host_sweeps >= 0 ? c->host_sweeps : (ctrl[1] > 0 ? ctrl[1] : P.pressure_iters);
  }
  return FL_OK;
}

// Relaxation of the mirror run: sweep k' of the transposed solve is forward sweep K-1-k'.
static int mirror_omega(Vjp* J, int K) {
  if (J->mir_K == K) return FL_OK;
  Ctx* m = J->mir;
  const fl_params& P = J->c->prm;
  const int Nf = P.pressure_iters;
  if (m->nmax < K) {
    if (m->d_omega) FL_HIP(hipFree(m->d_omega));
    if (m->d_maxdiv) FL_HIP(hipFree(m->d_maxdiv));
    const int cap = std::max(K, 64);
    FL_HIP(hipMalloc(&m->d_omega, sizeof(float) * cap));
    FL_HIP(hipMalloc(&m->d_maxdiv, sizeof(uint32_t) * 2 * cap));
    m->nmax = cap;
  }
  std::vector<float> om(K);
  const float init = P.relaxation, minR = 1.2f;
  for (int k = 0; k < K; k++) {
    const int kf = K - 1 - k;
    om[k] = init - (init - minR) * ((float)kf / (float)Nf);
  }
  FL_HIP(hipStreamSynchronize(J->c->stream));
  FL_HIP(hipMemcpy(m->d_omega, om.data(), sizeof(float) * K, hipMemcpyHostToDevice));
  m->omega_n = -1;  // the table is not the forward schedule
  J->mir_K = K;
  return FL_OK;
}

// ---- the multigrid V-cycle transposed (fluid.go:560-758, 1123-1149) -----------------------
// With the masks and the exit sweep fixed the cycle is linear in (U, V, p). p enters through
// the residual, so its adjoint lp is carried back through the smoothing sweeps, whose
// p += cp q gives q's adjoint a cp * lp term; p leaves the step unused, so lp starts at zero.
// The sweeps run as hyperplanes t = i + j + 2k in decreasing t: the cells of one hyperplane
// touch disjoint words (as in the forward schedule of fl_pressure.hip), so each launch is a
// set of independent transposed cell updates.

// Smoothing sweeps (pressureJacobiIteration, fluid.go:188-234) transposed: cell (i,j) of
// sweep k in hyperplane t.
__global__ void k_adj_gs_hp(int NX, int NY, int t, int k0, const uint8_t* __restrict__ flags, float om, float damping,
                            float cp, float* __restrict__ lU, float* __restrict__ lV, const float* __restrict__ lp) {
  const int k = k0 + blockIdx.y;
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = t - i - 2 * k;
  if (i > NX - 2 || j < 1 || j > NY - 2) return;
  const int n = NY, c = i * n + j;
  const uint32_t f = flags[c];
  if (!(f & CF_GS)) return;
  const float sx0 = (f & CF_X0) ? 1.0f : 0.0f, sx1 = (f & CF_X1) ? 1.0f : 0.0f;
  const float sy0 = (f & CF_Y0) ? 1.0f : 0.0f, sy1 = (f & CF_Y1) ? 1.0f : 0.0f;
  const float s = sx0 + sx1 + sy0 + sy1;
  const float alpha = -(om * damping) / s;
  const float g = alpha * (sx1 * lU[c + n] - sx0 * lU[c] + sy1 * lV[c + 1] - sy0 * lV[c] + cp * lp[c]);
  lU[c] -= g;
  lU[c + n] += g;
  lV[c] -= g;
  lV[c + 1] += g;
}

// Three sweeps at one relaxation, transposed.
static void adj_mg_sweeps(Ctx* c, float om, float cp, float* lU, float* lV, const float* lp) {
  const int NX = c->NX, NY = c->NY, K = 3, threads = 256;
  const int bx = (NX - 2 + threads - 1) / threads;
  for (int t = (NX - 2) + (NY - 2) + 2 * (K - 1); t >= 2; t--) {
    const int klo = std::max(0, (t - (NX - 2) - (NY - 2) + 1) / 2);
    const int khi = std::min(K - 1, (t - 2) / 2);
    if (klo > khi) continue;
    hipLaunchKernelGGL(k_adj_gs_hp, dim3(bx, khi - klo + 1), dim3(threads), 0, c->stream, NX, NY, t, klo, c->flags,
                       om, c->prm.pressure_damping, cp, lU, lV, lp);
  }
}

__device__ __forceinline__ bool mg_fluid_interior(const Grid& g, const float* __restrict__ S, int i, int j) {
  return i >= 1 && i <= g.NX - 2 && j >= 1 && j <= g.NY - 2 && S[i * g.NY + j] != 0.0f;
}

// applyPressureCorrection (fluid.go:1123-1149) transposed: the correction's adjoint at each
// fluid interior cell (p += cp corr; U, V faces -= / += S corr). The adjoints of U, V, p pass
// through unchanged.
__global__ void k_adj_mg_apply(Grid g, const float* __restrict__ S, float cp, const float* __restrict__ lU,
                               const float* __restrict__ lV, const float* __restrict__ lp, float* __restrict__ lcorr) {
  const int j = blockIdx.x * BX + threadIdx.x, i = blockIdx.y * BY + threadIdx.y;
  if (i >= g.NX || j >= g.NY) return;
  const int n = g.NY, c = i * n + j;
  float v = 0.0f;
  if (mg_fluid_interior(g, S, i, j))
    v = cp * lp[c] - S[c - n] * lU[c] + S[c + n] * lU[c + n] - S[c - 1] * lV[c] + S[c + 1] * lV[c + 1];
  lcorr[c] = v;
}

// prolongateCorrection (fluid.go:719-756) transposed: coarse value (I,J) gathers the fine
// values it was interpolated into, i in {2I-1, 2I, 2I+1} with weight 1/2, 1, 1/2 (likewise j),
// for fine cells with i/2 < cNX-1 and j/2 < cNY-1.
__global__ void k_adj_mg_prolong(int NX, int NY, int cNX, int cNY, const float* __restrict__ lcorr,
                                 float* __restrict__ lP) {
  const int J = blockIdx.x * BX + threadIdx.x, I = blockIdx.y * BY + threadIdx.y;
  if (I >= cNX || J >= cNY) return;
  float acc = 0.0f;
  for (int di = -1; di <= 1; di++) {
    const int i = 2 * I + di;
    if (i < 0 || i >= NX || i / 2 >= cNX - 1) continue;
    const float wi = di == 0 ? 1.0f : 0.5f;
    for (int dj = -1; dj <= 1; dj++) {
      const int j = 2 * J + dj;
      if (j < 0 || j >= NY || j / 2 >= cNY - 1) continue;
      acc += (wi * (dj == 0 ? 1.0f : 0.5f)) * lcorr[i * NY + j];
    }
  }
  lP[I * cNY + J] = acc;
}

// solveCoarseGrid (fluid.go:667-716) transposed. Its update of coarse cell c at sweep k,
// P_c <- P_c + relax/s_c (sum_d CS_d P_d - s_c P_c - rhs_c), read in reverse as a gather:
// the adjoint of P_c after sweep k is (1 - relax) times that after sweep k+1 (its own next
// update), plus relax/s_d times the adjoint after the update of each neighbour d that reads it:
// (i+1,j) and (i,j+1) at sweep k, (i-1,j) and (i,j-1) at sweep k+1. Every one of these lies
// on hyperplane t+1 or t+2, and lP holds each cell's adjoint in place (its neighbours' words
// are on hyperplanes of the other parity), so a decreasing-t schedule needs no atomics. The
// rhs adjoint collects -relax/s_c times each adjoint after an update.
__device__ __forceinline__ float mg_cs(const float* __restrict__ S, int NX, int NY, int a, int b) {
  return (2 * a < NX && 2 * b < NY) ? S[2 * a * NY + 2 * b] : 0.0f;  // fluid.go:672-684
}
__device__ __forceinline__ float mg_cs_sum(const float* __restrict__ S, int NX, int NY, int a, int b) {
  return mg_cs(S, NX, NY, a - 1, b) + mg_cs(S, NX, NY, a + 1, b) + mg_cs(S, NX, NY, a, b - 1) +
         mg_cs(S, NX, NY, a, b + 1);
}
__device__ __forceinline__ bool mg_cupd(const float* __restrict__ S, int NX, int NY, int cNX, int cNY, int a,
                                        int b) {
  return a >= 1 && a <= cNX - 2 && b >= 1 && b <= cNY - 2 && mg_cs(S, NX, NY, a, b) != 0.0f &&
         mg_cs_sum(S, NX, NY, a, b) != 0.0f;
}
__global__ void k_adj_mg_coarse(int NX, int NY, int cNX, int cNY, int t, int k0, int K, float relax,
                                const float* __restrict__ S, float* __restrict__ lP, float* __restrict__ lrhs) {
  const int k = k0 + blockIdx.y;
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = t - i - 2 * k;
  if (i > cNX - 2 || j < 1 || j > cNY - 2) return;
  if (!mg_cupd(S, NX, NY, cNX, cNY, i, j)) return;
  const int c = i * cNY + j;
  float L = lP[c];
  if (k < K - 1) L = L - relax * L;
  float acc = 0.0f;
  if (mg_cupd(S, NX, NY, cNX, cNY, i + 1, j)) acc += lP[c + cNY] / mg_cs_sum(S, NX, NY, i + 1, j);
  if (mg_cupd(S, NX, NY, cNX, cNY, i, j + 1)) acc += lP[c + 1] / mg_cs_sum(S, NX, NY, i, j + 1);
  if (k < K - 1) {
    if (mg_cupd(S, NX, NY, cNX, cNY, i - 1, j)) acc += lP[c - cNY] / mg_cs_sum(S, NX, NY, i - 1, j);
    if (mg_cupd(S, NX, NY, cNX, cNY, i, j - 1)) acc += lP[c - 1] / mg_cs_sum(S, NX, NY, i, j - 1);
  }
  L += relax * acc;
  lP[c] = L;
  lrhs[c] -= (relax / mg_cs_sum(S, NX, NY, i, j)) * L;
}

// restrictResidual (fluid.go:635-664) transposed: fine cell (i,j) gathers the coarse values
// whose 3x3 full-weighting stencil covers it, weights (1/2 or 1/4 per axis: 1/4, 1/8, 1/16).
// Only fluid interior cells carry a residual; the others get zero.
__global__ void k_adj_mg_restrict(Grid g, int cNX, int cNY, const float* __restrict__ S, const float* __restrict__ lrhs,
                                  float* __restrict__ lr) {
  const int j = blockIdx.x * BX + threadIdx.x, i = blockIdx.y * BY + threadIdx.y;
  if (i >= g.NX || j >= g.NY) return;
  float acc = 0.0f;
  if (mg_fluid_interior(g, S, i, j)) {
    for (int di = -1; di <= 1; di++) {
      const int v = i + di;
      if (v & 1) continue;
      const int I = v / 2;
      if (I < 1 || I > cNX - 2 || 2 * I >= g.NX - 1) continue;
      const float wi = di == 0 ? 0.5f : 0.25f;
      for (int dj = -1; dj <= 1; dj++) {
        const int u = j + dj;
        if (u & 1) continue;
        const int Jc = u / 2;
        if (Jc < 1 || Jc > cNY - 2 || 2 * Jc >= g.NY - 1) continue;
        acc += (wi * (dj == 0 ? 0.5f : 0.25f)) * lrhs[I * cNY + Jc];
      }
    }
  }
  lr[i * g.NY + j] = acc;
}

// computePressureResidual (fluid.go:601-632) transposed, r = -div - lap(p) on fluid interior
// cells, gathered per face and cell: U face (i,j) gets r(i,j) - r(i-1,j), V face (i,j) gets
// r(i,j) - r(i,j-1), and p at (i,j) gets s r(i,j) - S(i,j) (sum of its neighbours' r).
__global__ void k_adj_mg_residual(Grid g, const float* __restrict__ S, const float* __restrict__ lr,
                                  float* __restrict__ lU, float* __restrict__ lV, float* __restrict__ lp) {
  const int j = blockIdx.x * BX + threadIdx.x, i = blockIdx.y * BY + threadIdx.y;
  if (i >= g.NX || j >= g.NY) return;
  const int n = g.NY, c = i * n + j;
  const bool self = mg_fluid_interior(g, S, i, j);
  const float r = self ? lr[c] : 0.0f;
  float du = r, dv = r, nb = 0.0f;
  if (mg_fluid_interior(g, S, i - 1, j)) {
    du -= lr[c - n];
    nb += lr[c - n];
  }
  if (mg_fluid_interior(g, S, i, j - 1)) {
    dv -= lr[c - 1];
    nb += lr[c - 1];
  }
  if (mg_fluid_interior(g, S, i + 1, j)) nb += lr[c + n];
  if (mg_fluid_interior(g, S, i, j + 1)) nb += lr[c + 1];
  lU[c] += du;
  lV[c] += dv;
  float dp = -S[c] * nb;
  if (self) dp += (S[c - n] + S[c + n] + S[c - 1] + S[c + 1]) * r;
  lp[c] += dp;
}

// solveMultigridVCycle (fluid.go:560-599) transposed on L_OU/L_OV: `sweeps` fine sweeps ran,
// 6 per full cycle and 3 more when the exit test after a pre-smoothing stopped the solve.
// Work fields: L_CURL (lp), L_LCD (the correction's and the residual's adjoints), L_GX and
// L_GY (coarse P and rhs adjoints); all are free once confinement's adjoint is done.
static int adj_multigrid(Vjp* J, int sweeps, float dt) {
  Ctx* c = J->c;
  const int NX = c->NX, NY = c->NY, cNX = (NX + 1) / 2, cNY = (NY + 1) / 2, KC = 40;
  const size_t N = J->N, cN = (size_t)cNX * cNY;
  const Grid g = c->desc().g;
  const dim3 gr = grid_of(c), bl(BX, BY), cgr((cNY + BX - 1) / BX, (cNX + BY - 1) / BY);
  hipStream_t st = c->stream;
  float** L = J->lam;
  float *lU = L[L_OU], *lV = L[L_OV], *lp = L[L_CURL], *lf = L[L_LCD], *lP = L[L_GX], *lrhs = L[L_GY];
  const float cp = c->density * c->h / dt, relax = 1.6f;
  int rc;
  if ((rc = launch_fill(lp, 0.0f, N, st))) return rc;
  if (sweeps % 6 == 3) adj_mg_sweeps(c, 1.5f, cp, lU, lV, lp);
  for (int it = 0; it < sweeps / 6; it++) {
    adj_mg_sweeps(c, 1.2f, cp, lU, lV, lp);
    hipLaunchKernelGGL(k_adj_mg_apply, gr, bl, 0, st, g, c->S, cp, lU, lV, lp, lf);
    hipLaunchKernelGGL(k_adj_mg_prolong, cgr, bl, 0, st, NX, NY, cNX, cNY, lf, lP);
    if ((rc = launch_fill(lrhs, 0.0f, cN, st))) return rc;
    const int threads = 256, bx = (cNX - 2 + threads - 1) / threads;
    for (int t = (cNX - 2) + (cNY - 2) + 2 * (KC - 1); t >= 2; t--) {
      const int klo = std::max(0, (t - (cNX - 2) - (cNY - 2) + 1) / 2);
      const int khi = std::min(KC - 1, (t - 2) / 2);
      if (klo > khi) continue;
      hipLaunchKernelGGL(k_adj_mg_coarse, dim3(bx, khi - klo + 1), dim3(threads), 0, st, NX, NY, cNX, cNY, t, klo, KC,
                         relax, c->S, lP, lrhs);
    }
    hipLaunchKernelGGL(k_adj_mg_restrict, gr, bl, 0, st, g, cNX, cNY, c->S, lrhs, lf);
    hipLaunchKernelGGL(k_adj_mg_residual, gr, bl, 0, st, g, c->S, lf, lU, lV, lp);
    adj_mg_sweeps(c, 1.5f, cp, lU, lV, lp);
  }
  FL_HIP(hipGetLastError());
  return FL_OK;
}

static int adj_step(Vjp* J, const float* tp, int K, float dt) {
  Ctx* c = J->c;
  const fl_params& P = c->prm;
  const size_t N = J->N;
  const Grid g = c->desc().g;
  const dim3 gr = grid_of(c), bl(BX, BY);
  hipStream_t st = c->stream;
  float** L = J->lam;
  auto T = [&](int k) { return tp + (size_t)k * N; };
  auto zero = [&](int k) { return launch_fill(L[k], 0.0f, N, st); };
  const float SA = P.smoke_advection;
  const bool visc = P.viscosity_diffusion > 0.0f;
  const int dif = visc ? 1 : 0;
  const float sd = P.viscosity_diffusion * 0.3f * dt;  // advectSmoke's diffusion weight
  int rc;
  // smoke: adds the final velocity's share to L_U, L_V
  if ((rc = zero(L_M1))) return rc;
  if (P.use_bfecc) {
    if ((rc = zero(L_CM)) || (rc = zero(L_FM))) return rc;
    hipLaunchKernelGGL(k_adj_adv_smoke, gr, bl, 0, st, g, c->flags, T(T_U6), T(T_V6), T(T_CM), -dt, SA, 1, L[L_M],
                       L[L_CM], L[L_U], L[L_V], 1, dif, sd);
    hipLaunchKernelGGL(k_adj_corr_smoke, gr, bl, 0, st, g, c->flags, T(T_M1), T(T_FM), T(T_U6), T(T_V6), dt, SA,
                       L[L_CM], L[L_M1], L[L_FM], L[L_U], L[L_V]);
    hipLaunchKernelGGL(k_adj_adv_smoke, gr, bl, 0, st, g, c->flags, T(T_U6), T(T_V6), T(T_M1), -dt, SA, 1, L[L_FM],
                       L[L_M1], L[L_U], L[L_V], 1, dif, sd);
  } else {
    hipLaunchKernelGGL(k_adj_adv_smoke, gr, bl, 0, st, g, c->flags, T(T_U6), T(T_V6), T(T_M1), -dt, SA, 1, L[L_M],
                       L[L_M1], L[L_U], L[L_V], 1, dif, sd);
  }
  FL_HIP(hipGetLastError());
  std::swap(L[L_M], L[L_M1]);
  // velocity advection -> adjoint of the border-handled state (L_OU, L_OV)
  if ((rc = zero(L_OU)) || (rc = zero(L_OV))) return rc;
  // faces the passes do not advect keep the viscosity's output: their adjoint (L_SU, L_SV)
  float* stU = visc ? L[L_SU] : nullptr;
  float* stV = visc ? L[L_SV] : nullptr;
  if (visc && ((rc = zero(L_SU)) || (rc = zero(L_SV)))) return rc;
  if (P.use_bfecc) {
    if ((rc = zero(L_CU)) || (rc = zero(L_CV)) || (rc = zero(L_FU)) || (rc = zero(L_FV))) return rc;
    hipLaunchKernelGGL(k_adj_adv_vel, gr, bl, 0, st, g, c->flags, T(T_CU), T(T_CV), T(T_CU), T(T_CV), -dt, L[L_U],
                       L[L_V], L[L_CU], L[L_CV], L[L_CU], L[L_CV], 1, stU, stV);
    hipLaunchKernelGGL(k_adj_corr_vel, gr, bl, 0, st, g, c->flags, T(T_U5), T(T_V5), T(T_FU), T(T_FV), dt, L[L_CU],
                       L[L_CV], L[L_OU], L[L_OV], L[L_FU], L[L_FV]);
    hipLaunchKernelGGL(k_adj_adv_vel, gr, bl, 0, st, g, c->flags, T(T_U5), T(T_V5), T(T_U5), T(T_V5), -dt, L[L_FU],
                       L[L_FV], L[L_OU], L[L_OV], L[L_OU], L[L_OV], 1, stU, stV);
  } else {
    hipLaunchKernelGGL(k_adj_adv_vel, gr, bl, 0, st, g, c->flags, T(T_U5), T(T_V5), T(T_U5), T(T_V5), -dt, L[L_U],
                       L[L_V], L[L_OU], L[L_OV], L[L_OU], L[L_OV], 1, stU, stV);
  }
  FL_HIP(hipGetLastError());
  const int mr = std::max(c->NX, c->NY);
  hipLaunchKernelGGL(k_adj_borders, dim3((mr + 255) / 256), dim3(256), 0, st, g, c->S, L[L_OU], L[L_OV]);
  if (P.turbulence_strength > 0.0f)
    hipLaunchKernelGGL(k_adj_turb, gr, bl, 0, st, g, c->flags, T(T_U3), T(T_V3), P.turbulence_strength * dt, L[L_OU],
                       L[L_OV]);
  if (P.confinement != 0.0f) {
    hipLaunchKernelGGL(k_adj_curl, gr, bl, 0, st, g, c->flags, T(T_U2), T(T_V2), L[L_CURL]);
    hipLaunchKernelGGL(k_adj_conf_a, gr, bl, 0, st, g, c->flags, T(T_U2), T(T_V2), L[L_CURL], P.confinement, dt,
                       L[L_OU], L[L_OV], L[L_LCD], L[L_GX], L[L_GY]);
    hipLaunchKernelGGL(k_adj_conf_b, gr, bl, 0, st, g, c->flags, L[L_CURL], L[L_LCD], L[L_GX], L[L_GY], L[L_LCURL]);
    hipLaunchKernelGGL(k_adj_conf_c, gr, bl, 0, st, g, L[L_LCURL], L[L_OU], L[L_OV]);
  }
  FL_HIP(hipGetLastError());
  if (P.use_multigrid && P.multigrid_levels > 1) {
    if (K > 0 && (rc = adj_multigrid(J, K, dt))) return rc;
  } else if (K > 0) {
    Ctx* m = J->mir;
    if ((rc = mirror_omega(J, K))) return rc;
    hipLaunchKernelGGL(k_mirror_in, gr, bl, 0, st, g, c->S, L[L_OU], L[L_OV], m->U, m->V, m->p);
    FL_HIP(hipGetLastError());
    m->prm.pressure_damping = P.pressure_damping;
    m->prm.block_sweeps = P.block_sweeps;
    m->prm.handoff_lag = P.handoff_lag;
    m->prm.block_cols = P.block_cols;
    if ((rc = gs_wave_solve(m, K, 1.0f))) return rc;
    hipLaunchKernelGGL(k_mirror_out, gr, bl, 0, st, g, c->S, m->U, m->V, m->p, L[L_OU], L[L_OV]);
    FL_HIP(hipGetLastError());
  }
  if (visc) {  // L_CU/L_CV are free once the advection's adjoint is done
    hipLaunchKernelGGL(k_adj_visc, gr, bl, 0, st, g, c->flags, P.viscosity_diffusion * dt, L[L_OU], L[L_OV],
                       L[L_SU], L[L_SV], L[L_CU], L[L_CV]);
    FL_HIP(hipGetLastError());
    std::swap(L[L_OU], L[L_CU]);
    std::swap(L[L_OV], L[L_CV]);
  }
  hipLaunchKernelGGL(k_adj_force, gr, bl, 0, st, g, c->flags, L[L_OU], L[L_OV], L[L_GFU], L[L_GFV]);
  if (c->n_src_seg > 0)
    hipLaunchKernelGGL(k_adj_sources, dim3((c->n_src_seg + 255) / 256), dim3(256), 0, st, c->d_src, c->d_src_seg,
                       c->n_src_seg, c->S, L[L_OU], L[L_OV]);
  FL_HIP(hipGetLastError());
  std::swap(L[L_U], L[L_OU]);
  std::swap(L[L_V], L[L_OV]);
  return FL_OK;
}

static void vjp_free(Vjp* J) {
  if (!J) return;
  for (void* p : {(void*)J->Fu, (void*)J->Fv, (void*)J->target, (void*)J->x0, (void*)J->fin, (void*)J->ck,
                  (void*)J->tape, (void*)J->lam_mem, (void*)J->d_loss})
    if (p) (void)hipFree(p);
  if (J->mir) {
    J->mir->stream = J->mir_stream;  // its own stream again, destroyed with it
    free_ctx(J->mir);
  }
  delete J;
}

static int check_supported(Ctx* c) {
  if (c->NX < 6 || c->NY < 6) return verr(FL_ERR_UNSUPPORTED, "the adjoint needs at least 4x4 interior cells");
  return FL_OK;
}

}  // namespace fl

using namespace fl;

extern "C" {

int fl_vjp_create(fl_ctx* ctx, int steps, int checkpoint_every, fl_vjp** out) {
  if (!ctx || !out || steps < 1 || checkpoint_every < 1) return verr(FL_ERR_INVALID_ARG, "bad vjp arguments");
  *out = nullptr;
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (c->batch) return verr(FL_ERR_UNSUPPORTED, "rollouts of batch members are not supported");
  int rc = check_supported(c);
  if (rc) return rc;
  if ((rc = ctx_prepare(c))) return rc;
  Vjp* J = new Vjp();
  J->c = c;
  J->steps = steps;
  J->every = std::min(checkpoint_every, steps);
  J->nck = (steps + J->every - 1) / J->every;
  J->N = c->N;
  const size_t N = J->N;
  auto fail = [&](int code, const std::string& m) {
    vjp_free(J);
    return verr(code, m);
  };
  if (hipMalloc(&J->Fu, 4 * N) != hipSuccess || hipMalloc(&J->Fv, 4 * N) != hipSuccess ||
      hipMalloc(&J->target, 4 * N) != hipSuccess || hipMalloc(&J->x0, 4 * N * NSTATE) != hipSuccess ||
      hipMalloc(&J->fin, 4 * N * NSTATE) != hipSuccess ||
      hipMalloc(&J->ck, 4 * N * NSTATE * J->nck) != hipSuccess ||
      hipMalloc(&J->tape, 4 * N * NTAPE * J->every) != hipSuccess ||
      hipMalloc(&J->lam_mem, 4 * N * NLAM) != hipSuccess || hipMalloc(&J->d_loss, sizeof(double) * (1 + LOSS_BLOCKS)) != hipSuccess)
    return fail(FL_ERR_HIP, "hipMalloc vjp buffers");
  for (int k = 0; k < NLAM; k++) J->lam[k] = J->lam_mem + (size_t)k * N;
  // the point-mirrored grid for the transposed sweeps, on this context's stream
  fl_ctx* m = nullptr;
  if ((rc = fl_create(c->density, c->NX - 2, c->NY - 2, c->h, &m))) return fail(rc, "mirror grid");
  J->mir = reinterpret_cast<Ctx*>(m);
  std::vector<float> Sm(N);
  for (int i = 0; i < c->NX; i++)
    for (int j = 0; j < c->NY; j++) Sm[(size_t)(c->NX - 1 - i) * c->NY + (c->NY - 1 - j)] = c->hS[(size_t)i * c->NY + j];
  if ((rc = fl_upload(m, FL_FIELD_S, Sm.data(), N)) || (rc = fl_sync(m))) return fail(rc, "mirror grid S");
  J->mir_stream = J->mir->stream;
  J->mir->stream = c->stream;
  J->mir->prm.early_exit = 0;
  J->mir->prm.solver = 2;
  if ((rc = save_state(J, J->x0))) return fail(rc, "state");
  if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(FL_ERR_HIP, "vjp init");
  *out = reinterpret_cast<fl_vjp*>(J);
  return FL_OK;
}

int fl_vjp_destroy(fl_vjp* v) {
  vjp_free(reinterpret_cast<Vjp*>(v));
  return FL_OK;
}

int fl_vjp_forward(fl_vjp* v, float dt, const float* fu, const float* fv, const float* target, size_t n,
                   double* loss) {
  Vjp* J = reinterpret_cast<Vjp*>(v);
  if (!J || !fu || !fv || !target) return verr(FL_ERR_INVALID_ARG, "null argument");
  if (n != J->N) return verr(FL_ERR_INVALID_ARG, "arrays must hold NumX*NumY floats");
  Ctx* c = J->c;
  int rc = check_supported(c);
  if (rc) return rc;
  if ((rc = ctx_prepare(c))) return rc;
  FL_HIP(hipMemcpyAsync(J->Fu, fu, 4 * n, hipMemcpyHostToDevice, c->stream));
  FL_HIP(hipMemcpyAsync(J->Fv, fv, 4 * n, hipMemcpyHostToDevice, c->stream));
  FL_HIP(hipMemcpyAsync(J->target, target, 4 * n, hipMemcpyHostToDevice, c->stream));
  FL_HIP(hipStreamSynchronize(c->stream));  // the host arrays may be reused on return
  if ((rc = load_state(J, J->x0))) return rc;
  J->dt = dt;
  for (int s = 0; s < J->steps; s++) {
    if (s % J->every == 0 && (rc = save_state(J, J->ck + (size_t)(s / J->every) * NSTATE * J->N))) return rc;
    if ((rc = vjp_step(J, dt, nullptr, nullptr))) return rc;
  }
  if ((rc = save_state(J, J->fin))) return rc;
  hipLaunchKernelGGL(k_loss, dim3(LOSS_BLOCKS), dim3(LOSS_THREADS), 0, c->stream, c->M, J->target, J->lam[L_M],
                     J->N, J->d_loss + 1);
  hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(LOSS_THREADS), 0, c->stream, J->d_loss + 1, J->d_loss);
  FL_HIP(hipGetLastError());
  double l = 0.0;
  FL_HIP(hipMemcpyAsync(&l, J->d_loss, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  FL_HIP(hipStreamSynchronize(c->stream));
  if (loss) *loss = l;
  J->have_forward = true;
  return FL_OK;
}

int fl_vjp_backward(fl_vjp* v, float* grad_fu, float* grad_fv, size_t n) {
  Vjp* J = reinterpret_cast<Vjp*>(v);
  if (!J || !grad_fu || !grad_fv) return verr(FL_ERR_INVALID_ARG, "null argument");
  if (n != J->N) return verr(FL_ERR_INVALID_ARG, "arrays must hold NumX*NumY floats");
  if (!J->have_forward) return verr(FL_ERR_INVALID_ARG, "fl_vjp_forward must run first");
  Ctx* c = J->c;
  const size_t N = J->N;
  int rc;
  // L_M holds 2 (M - target) from the forward; the velocity adjoint starts at zero
  for (int k : {L_U, L_V, L_GFU, L_GFV})
    if ((rc = launch_fill(J->lam[k], 0.0f, N, c->stream))) return rc;
  J->sweeps.assign(J->every, 0);
  for (int seg = J->nck - 1; seg >= 0; seg--) {
    if ((rc = load_state(J, J->ck + (size_t)seg * NSTATE * N))) return rc;
    const int s0 = seg * J->every, s1 = std::min(J->steps, s0 + J->every);
    for (int s = s0; s < s1; s++)
      if ((rc = vjp_step(J, J->dt, J->tape + (size_t)(s - s0) * NTAPE * N, &J->sweeps[s - s0]))) return rc;
    for (int s = s1 - 1; s >= s0; s--)
      if ((rc = adj_step(J, J->tape + (size_t)(s - s0) * NTAPE * N, J->sweeps[s - s0], J->dt))) return rc;
  }
  if ((rc = copy_to_host(c, grad_fu, J->lam[L_GFU], 4 * N)) || (rc = copy_to_host(c, grad_fv, J->lam[L_GFV], 4 * N)))
    return rc;
  if ((rc = load_state(J, J->fin))) return rc;
  FL_HIP(hipStreamSynchronize(c->stream));
  // L_M was consumed: a second backward needs a new forward
  J->have_forward = false;
  return FL_OK;
}

}  // extern "C"
