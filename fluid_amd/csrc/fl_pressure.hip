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

// ---- multigrid V-cycle (fluid.go:559-758, 1123-1149) ---------------------------------------
namespace fl {

static constexpr int MBX = 64, MBY = 4;

// computePressureResidual (fluid.go:600-632): r = -div - lap(p) on fluid interior cells, 0 elsewhere.
__global__ void k_mg_residual(Grid g, const float* __restrict__ S, const float* __restrict__ U,
                              const float* __restrict__ V, const float* __restrict__ p, float* __restrict__ r) {
  const int j = blockIdx.x * MBX + threadIdx.x, i = blockIdx.y * MBY + threadIdx.y;
  const int NX = g.NX, NY = g.NY, n = NY;
  if (i >= NX || j >= NY) return;
  const int c = i * n + j;
  float out = 0.0f;
  if (i >= 1 && i <= NX - 2 && j >= 1 && j <= NY - 2 && S[c] != 0.0f) {
    const float div = U[c + n] - U[c] + V[c + 1] - V[c];
    const float sx0 = S[c - n], sx1 = S[c + n], sy0 = S[c - 1], sy1 = S[c + 1];
    const float lap = sx0 * (p[c - n] - p[c]) + sx1 * (p[c + n] - p[c]) + sy0 * (p[c - 1] - p[c]) +
                      sy1 * (p[c + 1] - p[c]);
    out = -div - lap;
  }
  r[c] = out;
}

// restrictResidual (fluid.go:635-664) and the coarse mask of solveCoarseGrid (fluid.go:672-684).
__global__ void k_mg_restrict(Grid g, int cNX, int cNY, const float* __restrict__ S, const float* __restrict__ r,
                              float* __restrict__ rhs, float* __restrict__ cS, float* __restrict__ cP) {
  const int j = blockIdx.x * MBX + threadIdx.x, i = blockIdx.y * MBY + threadIdx.y;
  if (i >= cNX || j >= cNY) return;
  const int NX = g.NX, NY = g.NY, cc = i * cNY + j;
  const int fi = 2 * i, fj = 2 * j;
  float out = 0.0f;
  if (i >= 1 && i <= cNX - 2 && j >= 1 && j <= cNY - 2 && fi < NX - 1 && fj < NY - 1) {
    const float center = r[fi * NY + fj] * 0.25f;
    const float nb = (r[(fi - 1) * NY + fj] + r[(fi + 1) * NY + fj] + r[fi * NY + fj - 1] + r[fi * NY + fj + 1]) *
                     0.125f;
    const float co = (r[(fi - 1) * NY + fj - 1] + r[(fi + 1) * NY + fj - 1] + r[(fi - 1) * NY + fj + 1] +
                      r[(fi + 1) * NY + fj + 1]) *
                     0.0625f;
    out = center + nb + co;
  }
  rhs[cc] = out;
  cS[cc] = (fi < NX && fj < NY) ? S[fi * NY + fj] : 0.0f;
  cP[cc] = 0.0f;
}

// solveCoarseGrid's 40 in-place lexicographic sweeps at relaxation 1.6 (fluid.go:686-716),
// bit-identical to the serial loop. The interior is cut into tiles (I, J) of 32 rows (i) by
// 64 columns (j), 26 KB of LDS, six workgroups to a CU. Tile (I, J) runs sweep k in the launch
// for diagonal T = I + J + 2k. Its four neighbour tiles then hold the values the serial order
// reads: (I-1, J) and (I, J-1) have done sweep k (diagonal T-1), (I+1, J) and (I, J+1) sweep
// k-1 but not k (T-1 and T+1). No two tiles of one diagonal are neighbours (their T would
// differ by an odd number), so a tile reads its one-cell halo once at the start and writes its
// interior back at the end. Inside, one wave walks the tile's anti-diagonals (the other three
// waves help load and store the tile): lane l updates column l, row d - l at step d, after
// (row - 1, l) and (row, l - 1) (step d - 1, this lane and lane l - 1) and before (row + 1, l)
// and (row, l + 1). Every step runs the reference's expression on the same operands, so the
// result is the serial loop's to the bit.
// ALL: the interior is one tile of at most TR x 64 cells, and the launch runs all 40 sweeps on
// it in LDS, the four waves walking successive sweeps three steps apart (below), instead of one
// launch per sweep.
static constexpr int MG_COARSE_ITERS = 40, MG_TR = 32, MG_TR_ALL = 64, MG_TC = 64, MG_PITCH = MG_TC + 2;

template <int TR, bool ALL>
__global__ __launch_bounds__(256) void k_mg_coarse_tiles(int cNX, int cNY, int nI, int nJ, int T, int klo,
                                                         const float* __restrict__ cS,
                                                         const float* __restrict__ rhs, float* __restrict__ cP) {
  constexpr int MG_TR = TR;
  const int k = klo + blockIdx.y, I = blockIdx.x, J = T - 2 * k - I;
  if (J < 0 || J >= nJ) return;
  __shared__ float lp[(MG_TR + 2) * MG_PITCH], ls[(MG_TR + 2) * MG_PITCH], lr[MG_TR * MG_TC];
  const int W = cNX - 2, H = cNY - 2;
  const int i0 = 1 + I * MG_TR, j0 = 1 + J * MG_TC;  // first interior cell of the tile
  const int rows = min(MG_TR, W - I * MG_TR), cols = min(MG_TC, H - J * MG_TC);
  const int tid = threadIdx.x, l = tid & 63;
  // Load: rows i0-1 .. i0+rows and columns j0-1 .. j0+cols, all inside the grid (i0+rows <=
  // W+1 = cNX-1), as one flat range over the four waves with every load in flight at once.
  // The corners belong to diagonal tiles and are never read.
  const int pw = cols + 2, n = (rows + 2) * pw;
  constexpr int LU = ((MG_TR + 2) * MG_PITCH + 255) / 256;
  float vp[LU], vs[LU], vr[LU];
#pragma unroll
  for (int u = 0; u < LU; u++) {
    const int e = tid + 256 * u, r = e / pw - 1, q = e - (r + 1) * pw;
    const bool in = e < n && !((r < 0 || r == rows) && (q == 0 || q == cols + 1));
    const size_t g = (size_t)(i0 + r) * cNY + j0 - 1 + q;
    vp[u] = in ? cP[g] : 0.0f;
    vs[u] = in ? cS[g] : 0.0f;
    vr[u] = (in && r >= 0 && r < rows && q >= 1 && q <= cols) ? rhs[g] : 0.0f;
  }
#pragma unroll
  for (int u = 0; u < LU; u++) {
    const int e = tid + 256 * u, r = e / pw - 1, q = e - (r + 1) * pw;
    if (e < n) {
      lp[(r + 1) * MG_PITCH + q] = vp[u];
      ls[(r + 1) * MG_PITCH + q] = vs[u];
      if (r >= 0 && r < rows && q >= 1 && q <= cols) lr[r * MG_TC + q - 1] = vr[u];
    }
  }
  __syncthreads();
  // The walk. Lane l is at row r = d - l at step d (also before and after its active rows, so
  // its registers are there for its neighbours) and holds its column's S[r-1..r+1], the old
  // P[r], P[r+1], rhs[r] and cur = P[r-1] after this sweep. Lane l-1 (row r+1) passes on the
  // new P[r][l-1] and S[r][l-1]; lane l+1 (row r-1) the old P[r][l+1] and S[r][l+1], by wave
  // shifts in full-wave control flow. Lane 0 and lane 63 take the halo columns instead; for a
  // narrower tile, lane `cols` walks the east halo column. LDS is only read here (the cells a
  // lane will still read are never the ones written), so the walk needs no barrier; results
  // go to LDS for the write-back.
  const float relax = 1.6f;
  auto row_of = [&](int x) { return min(max(x, -1), rows) + 1; };  // LDS row of row x, clamped
  const int col = l + 1;                                                // LDS column of this lane
  int r = 0;
  float sm1 = 0.0f, s0 = 0.0f, sp1 = 0.0f, po = 0.0f, po1 = 0.0f, cur = 0.0f, rh = 0.0f;
  auto init = [&]() {
    r = -l;
    sm1 = ls[row_of(r - 1) * MG_PITCH + col];
    s0 = ls[row_of(r) * MG_PITCH + col];
    sp1 = ls[row_of(r + 1) * MG_PITCH + col];
    po = lp[row_of(r) * MG_PITCH + col];
    po1 = lp[row_of(r + 1) * MG_PITCH + col];
    cur = lp[row_of(r - 1) * MG_PITCH + col];
    rh = lr[min(max(r, 0), rows - 1) * MG_TC + l];
  };
  auto step = [&]() {
    const int hrow = row_of(r) * MG_PITCH;
    const float pWh = lp[hrow], sWh = ls[hrow], pEh = lp[hrow + cols + 1], sEh = ls[hrow + cols + 1];
    const float pW = __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(pWh), __float_as_int(cur), 0x138, 0xF, 0xF, false));
    const float sy0 = __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(sWh), __float_as_int(sm1), 0x138, 0xF, 0xF, false));
    const float pE = __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(pEh), __float_as_int(po1), 0x130, 0xF, 0xF, false));
    const float sy1 = __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(sEh), __float_as_int(sp1), 0x130, 0xF, 0xF, false));
    // next row's values (prefetched before the update's dependent chain)
    const float sp2 = ls[row_of(r + 2) * MG_PITCH + col], po2 = lp[row_of(r + 2) * MG_PITCH + col];
    const float rh1 = lr[min(max(r + 1, 0), rows - 1) * MG_TC + l];
    float pn = po;
    const bool active = l < cols && r >= 0 && r < rows;
    if (active && s0 != 0.0f) {
      const float sx0 = sm1, sx1 = sp1;
      const float s = sx0 + sx1 + sy0 + sy1;
      if (s != 0.0f) {
        const float pc = po;
        const float lap = sx0 * (cur - pc) + sx1 * (po1 - pc) + sy0 * (pW - pc) + sy1 * (pE - pc);
        const float res = rh - lap;
        const float corr = -res / s * relax;
        pn = pc + corr;
      }
    }
    if (active) lp[(r + 1) * MG_PITCH + col] = pn;
    cur = pn;
    sm1 = s0;
    s0 = sp1;
    sp1 = sp2;
    po = po1;
    po1 = po2;
    rh = rh1;
    r++;
  };
  if constexpr (ALL) {
    // The 40 walks overlap on the four waves: wave w walks sweeps w, w+4, ..., and sweep k
    // starts at step (k / 4) P + 3 (k % 4) of a common clock with a barrier per step. A walk
    // at step d reads old values (the previous sweep's) on anti-diagonals up to d + 2, which
    // that sweep wrote at its step d + 2, at least one barrier earlier (lag 3 >= 2 + 1); it
    // writes anti-diagonal d after the previous sweep read it there (at that sweep's step d - 2).
    // P >= 12 keeps the lag from sweep 4m+3 to 4m+4 at 3 or more on small tiles.
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int D = rows + cols - 1, P = max(D, 12), M = MG_COARSE_ITERS / 4;
    const int total = (M - 1) * P + 9 + D;
    for (int tau = 0; tau < total; tau++) {
      const int x = tau - 3 * wv;
      if (x >= 0 && x < M * P) {
        const int d = x % P;
        if (d == 0) init();
        if (d < D) step();
      }
      __syncthreads();
    }
  } else if (tid < 64) {
    init();
    for (int d = 0; d < rows + cols - 1; d++) step();
  }
  __syncthreads();
  for (int e = tid; e < rows * cols; e += 256) {
    const int r = e / cols, q = e - r * cols;
    cP[(size_t)(i0 + r) * cNY + j0 + q] = lp[(r + 1) * MG_PITCH + q + 1];
  }
}

// All 40 coarse sweeps: one launch when the interior fits one tile of 64 x 64, else one
// launch per tile diagonal, in order on the stream.
static int mg_coarse_solve(Ctx* c, int cNX, int cNY) {
  const int W = cNX - 2, H = cNY - 2, K = MG_COARSE_ITERS;
  if (W < 1 || H < 1) return FL_OK;
  if (W <= MG_TR_ALL && H <= MG_TC) {
    hipLaunchKernelGGL((k_mg_coarse_tiles<MG_TR_ALL, true>), dim3(1, 1), dim3(256), 0, c->stream, cNX, cNY, 1, 1, 0,
                       0, c->mg_S, c->mg_R, c->mg_P);
    FL_HIP(hipGetLastError());
    return FL_OK;
  }
  const int nI = (W + MG_TR - 1) / MG_TR, nJ = (H + MG_TC - 1) / MG_TC;
  for (int T = 0; T <= (nI - 1) + (nJ - 1) + 2 * (K - 1); T++) {
    const int klo = std::max(0, (T - (nI - 1) - (nJ - 1) + 1) / 2), khi = std::min(K - 1, T / 2);
    if (klo > khi) continue;  // no tile on this diagonal (a one-tile grid has none on odd T)
    hipLaunchKernelGGL((k_mg_coarse_tiles<MG_TR, false>), dim3(nI, khi - klo + 1), dim3(256), 0, c->stream, cNX, cNY,
                       nI, nJ, T, klo, c->mg_S, c->mg_R, c->mg_P);
  }
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// prolongateCorrection (fluid.go:719-756) at one fine cell.
__device__ __forceinline__ float mg_corr(const float* __restrict__ cc, int i, int j, int cNX, int cNY) {
  const int ci = i / 2, cj = j / 2;
  if (ci >= cNX - 1 || cj >= cNY - 1) return 0.0f;
  const float fi = (float)(i % 2) * 0.5f, fj = (float)(j % 2) * 0.5f;
  return (1.0f - fi) * (1.0f - fj) * cc[ci * cNY + cj] + fi * (1.0f - fj) * cc[(ci + 1) * cNY + cj] +
         (1.0f - fi) * fj * cc[ci * cNY + cj + 1] + fi * fj * cc[(ci + 1) * cNY + cj + 1];
}
__device__ __forceinline__ bool mg_active(const float* __restrict__ S, int i, int j, int NX, int NY) {
  return i >= 1 && i <= NX - 2 && j >= 1 && j <= NY - 2 && S[i * NY + j] != 0.0f;
}

// applePressureCorrection (fluid.go:1123-1149), serial in the reference: face U[i,j] gets
// cell (i-1,j)'s "+= sx1*pc" and then cell (i,j)'s "-= sx0*pc"; V[i,j] gets (i,j-1)'s "+="
// then (i,j)'s "-=". Each face applies exactly the updates of its active cells in that order.
__global__ void k_mg_apply(Grid g, int cNX, int cNY, const float* __restrict__ S, const float* __restrict__ cc,
                           float cp, float* __restrict__ U, float* __restrict__ V, float* __restrict__ p) {
  const int j = blockIdx.x * MBX + threadIdx.x, i = blockIdx.y * MBY + threadIdx.y;
  const int NX = g.NX, NY = g.NY, n = NY;
  if (i >= NX || j >= NY) return;
  const int c = i * n + j;
  const bool self = mg_active(S, i, j, NX, NY);
  const float pcs = self ? mg_corr(cc, i, j, cNX, cNY) : 0.0f;
  if (self) p[c] += pcs * cp;
  if (i >= 1) {
    float u = U[c];
    if (mg_active(S, i - 1, j, NX, NY)) u += S[c] * mg_corr(cc, i - 1, j, cNX, cNY);
    if (self) u -= S[c - n] * pcs;
    U[c] = u;
  }
  if (j >= 1) {
    float v = V[c];
    if (mg_active(S, i, j - 1, NX, NY)) v += S[c] * mg_corr(cc, i, j - 1, cNX, cNY);
    if (self) v -= S[c - 1] * pcs;
    V[c] = v;
  }
}

int multigrid_solve(Ctx* c, int numIters, float dt) {
  c->last_n = 0;
  c->host_sweeps = 0;
  if (numIters <= 0) return FL_OK;
  int rc = ensure_omega(c, std::max(numIters, 3));
  if (rc) return rc;
  const int cNX = (c->NX + 1) / 2, cNY = (c->NY + 1) / 2;
  const size_t cN = (size_t)cNX * cNY;
  if (!c->mg_omega) {
    const float om[6] = {1.5f, 1.5f, 1.5f, 1.2f, 1.2f, 1.2f};
    FL_HIP(hipMalloc(&c->mg_omega, sizeof(om)));
    FL_HIP(hipMalloc(&c->mg_P, 4 * cN));
    FL_HIP(hipMalloc(&c->mg_S, 4 * cN));
    FL_HIP(hipMalloc(&c->mg_R, 4 * cN));
    FL_HIP(hipMemcpy(c->mg_omega, om, sizeof(om), hipMemcpyHostToDevice));
  }
  const float cp = c->density * c->h / dt;
  float* saved_omega = c->d_omega;
  const int saved_early = c->prm.early_exit;
  c->prm.early_exit = 0;  // the V-cycle checks the third pre-smoothing sweep itself
  auto restore = [&]() {
    c->d_omega = saved_omega;
    c->prm.early_exit = saved_early;
  };
  const dim3 fb(MBX, MBY), fg((c->NY + MBX - 1) / MBX, (c->NX + MBY - 1) / MBY);
  const dim3 cg((cNY + MBX - 1) / MBX, (cNX + MBY - 1) / MBY);
  int sweeps = 0;
  for (int it = 0; it < numIters; it++) {
    c->d_omega = c->mg_omega;
    if ((rc = gs_wave_solve(c, 3, cp))) return restore(), rc;
    sweeps += 3;
    uint32_t bits = 0;
    FL_HIP(hipMemcpyAsync(&bits, c->d_maxdiv + 2, 4, hipMemcpyDeviceToHost, c->stream));
    FL_HIP(hipStreamSynchronize(c->stream));
    float maxDiv;
    std::memcpy(&maxDiv, &bits, 4);
    if (maxDiv < 1e-5f) break;
    hipLaunchKernelGGL(k_mg_residual, fg, fb, 0, c->stream, c->grid, c->S, c->U, c->V, c->p, c->T0);
    hipLaunchKernelGGL(k_mg_restrict, cg, fb, 0, c->stream, c->grid, cNX, cNY, c->S, c->T0, c->mg_R, c->mg_S,
                       c->mg_P);
    if ((rc = mg_coarse_solve(c, cNX, cNY))) return restore(), rc;
    hipLaunchKernelGGL(k_mg_apply, fg, fb, 0, c->stream, c->grid, cNX, cNY, c->S, c->mg_P, cp, c->U, c->V, c->p);
    FL_HIP(hipGetLastError());
    c->d_omega = c->mg_omega + 3;
    if ((rc = gs_wave_solve(c, 3, cp))) return restore(), rc;
    sweeps += 3;
  }
  restore();
  c->last_n = sweeps;
  c->host_sweeps = sweeps;
  return FL_OK;
}

}  // namespace fl
