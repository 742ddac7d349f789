// Temporal-blocked Gauss-Seidel/SOR pressure solve for gfx950, bit-identical to the
// reference's serial in-place sweeps (pressureJacobiIteration, fluid.go:188-234).
//
// Cell (i,j) of sweep k needs (i-1,j,k), (i,j-1,k), (i+1,j,k-1), (i,j+1,k-1) and (i,j,k-1).
// With skewed rows i' = i + k these become forward dependences in (i', k), plus a
// j-lag of at most one column. A workgroup owns a tile of 64 skewed rows (the lanes a)
// by TK sweeps (the waves b) and marches along j: at step s, thread (a,b) updates cell
// (i = i0 + a - b, j = s - a - b) of sweep k0 + b. Every face then travels between
// neighbouring threads:
//   U right face, after the update, -> (a+1, b)   at s+1   (DPP wave_shr:1)
//   U left face                     -> (a, b+1)   at s+1   (LDS)
//   V bottom face and p             -> (a+1, b+1) at s+1   (LDS; p is used at s+2)
//   V top face                      -> (a, b) itself       (register)
// so the state of a tile lives in registers and LDS for all its TK sweeps; HBM sees the
// initial U,V once, the final U,V,p once, and per tile and column the boundary faces
// handed to the next tile along i' (G stream) and the next block of sweeps (K stream).
// Tiles run as a persistent kernel in ticket order (a topological order of the tile
// graph); a consumer waits on its producers' published progress, in blocks of 8 columns.
// Handoff data and progress counters are accessed only with agent-scope relaxed atomic
// loads and stores (device-coherent accesses, global_load/store sc1), so neither side
// needs an L2 writeback or invalidate; the producer waits for its data stores to
// complete (s_waitcnt vmcnt(0)) before it stores the new progress value, and consumers
// issue their data loads only after observing it.
//
// Cells outside the interior or solid pass their faces through unchanged, exactly like
// the skipped cells of the serial loop. Faces no interior cell touches keep their input
// values. The result (U, V and p) goes to separate output buffers, so the early exit (stop after the
// first sweep with max|div| < 1e-5, fluid.go:175-177) is a re-run of fewer sweeps from
// the untouched input, decided on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "fl_internal.h"

namespace fl {

static constexpr int LANES = 64;
static constexpr int BLK = 8;                  // columns per staging/publish block
static constexpr int SPIN_LIMIT = 1 << 22;     // bounded wait (about a second); then error and continue

struct GsArgs {
  int NX, NY, G, KB, rows, ntiles, which;
  int K;                         // sweeps of the main run; the re-run reads ctrl[1]
  const int32_t* tiles;          // ticket order, packed (g << 16) | kb
  const float* Uin;
  const float* Vin;
  const float* Pin;              // initial p (accumulated onto, like the serial loop)
  float* Uout;
  float* Vout;
  float* Pout;
  const uint8_t* fdiag;          // flags by anti-diagonal: [(i+j)*NX + i]
  const float* omega;
  float damping, cp;
  float* kout;                   // [G*KB][rows][3][65]: planes U, V, p per row
  float* gout;                   // [G*KB][rows][3][TK]
  int32_t* prog;                 // [G*KB]
  int32_t* ticket;
  uint32_t* maxdiv;              // [K]
  int32_t* ctrl;
};

__device__ __forceinline__ bool tile_empty(int g, int kb, int K, int TK, int NX) {
  int k0 = kb * TK;
  if (k0 >= K) return true;
  int kmax = min(k0 + TK, K) - 1;
  int rmin = LANES * g - kmax, rmax = LANES * g + LANES - 1 - k0;
  return rmax < 0 || rmin > NX - 1;
}

// Device-coherent (agent scope) accesses of the handoff buffers.
__device__ __forceinline__ float ld_dev(const float* p) {
  return __hip_atomic_load(const_cast<float*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(float* p, float v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// One handoff entry (U, V, p) stored as three planes of `stride` floats.
__device__ __forceinline__ float4 ld_dev3(const float* p, int stride) {
  return make_float4(ld_dev(p), ld_dev(p + stride), ld_dev(p + 2 * stride), 0.0f);
}
// Wait until every store this wave issued has been acknowledged by the memory system.
__device__ __forceinline__ void wait_stores() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

__device__ __forceinline__ int wait_progress(int32_t* p, int need, int32_t* err) {
  int v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (int it = 0; v < need; it++) {
    if (it > SPIN_LIMIT) {
      atomicOr(err, 1);
      return need;
    }
    __builtin_amdgcn_s_sleep(2);
    v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return v;
}

template <int TK>
__global__ __launch_bounds__(TK * LANES) void k_gs_wave(GsArgs A) {
  const int lane = threadIdx.x & (LANES - 1);
  const int wave = threadIdx.x / LANES;
  const int tid = threadIdx.x;
  constexpr int NT = TK * LANES;

  __shared__ float xU[2][TK][LANES];
  __shared__ float2 xVP[2][TK][LANES];
  __shared__ float4 Kst[2][BLK][LANES + 1];
  __shared__ float4 Gst[2][BLK][TK];
  __shared__ float4 Gos[2][BLK][TK];
  __shared__ int s_tile;
  __shared__ int s_seen[3];

  const int K = A.which ? A.ctrl[1] : A.K;
  if (K <= 0) return;
  const int NX = A.NX, NY = A.NY, rows = A.rows, KB = A.KB;
  const int nblk = rows / BLK;
  int32_t* err = A.ctrl + 3;

  for (;;) {
    if (tid == 0) s_tile = atomicAdd(A.ticket, 1);
    __syncthreads();
    const int t = s_tile;
    __syncthreads();
    if (t >= A.ntiles) break;
    const int g = A.tiles[t] >> 16, kb = A.tiles[t] & 0xffff;
    const int me = g * KB + kb;
    if (tile_empty(g, kb, K, TK, NX)) {
      if (tid == 0) __hip_atomic_store(&A.prog[me], rows, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      continue;
    }
    const int k0 = kb * TK;
    const int k = k0 + wave;
    const bool wave_on = k < K;
    const bool last = k == K - 1;
    const int i0 = LANES * g - k0;           // row of (lane 0, wave 0)
    const int row = i0 + lane - wave;        // this thread's row
    const bool real_row = row >= 0 && row <= NX - 1;
    const float om = wave_on ? A.omega[k] : 0.0f;
    const int d0 = LANES * g - k0 - 2 * wave;  // anti-diagonal of column 0 for this wave

    // Producers: G stream (g-1,kb), K stream (g,kb-1), corner (g-1,kb-1).
    const int pG = (g > 0 && !tile_empty(g - 1, kb, K, TK, NX)) ? (g - 1) * KB + kb : -1;
    const int pK = (kb > 0 && !tile_empty(g, kb - 1, K, TK, NX)) ? g * KB + kb - 1 : -1;
    const int pC = (g > 0 && kb > 0 && !tile_empty(g - 1, kb - 1, K, TK, NX)) ? (g - 1) * KB + kb - 1 : -1;
    if (tid < 3) s_seen[tid] = 0;
    __syncthreads();

    // Wait until every producer has published the rows block `blk` stages.
    auto wait_block = [&](int blk) {
      if (tid == 0) {
        const int s1 = blk * BLK + BLK - 1;  // last step of the block
        if (pG >= 0) {
          int need = min(s1 + 63 + 1, rows);
          if (s_seen[0] < need) s_seen[0] = wait_progress(&A.prog[pG], need, err);
        }
        if (pK >= 0) {
          int need = min(s1 + TK - 1 + 1, rows);
          if (s_seen[1] < need) s_seen[1] = wait_progress(&A.prog[pK], need, err);
        }
        if (pC >= 0) {
          int need = min(s1 + 63 + TK + 1, rows);
          if (s_seen[2] < need) s_seen[2] = wait_progress(&A.prog[pC], need, err);
        }
      }
    };

    // Staged boundary items of one block: K stream 8x65, G stream 8xTK.
    constexpr int NITEM = BLK * (LANES + 1) + BLK * TK;
    constexpr int PER = (NITEM + NT - 1) / NT;
    float4 stage[PER];
    auto load_block = [&](int blk) {
#pragma unroll
      for (int q = 0; q < PER; q++) {
        const int it = tid + q * NT;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (it < BLK * (LANES + 1)) {
          const int u = it / (LANES + 1), slot = it % (LANES + 1);
          const int s = blk * BLK + u;
          if (kb == 0) {
            // initial state: lane `slot` of wave 0 is row i0+slot at column s-slot
            const int r = i0 + slot, j = s - slot;
            if (slot < LANES && j >= 0 && j <= NY - 1) {
              if (r + 1 >= 0 && r + 1 <= NX - 1) v.x = A.Uin[(size_t)(r + 1) * NY + j];
              if (r >= 0 && r <= NX - 1 && j + 1 <= NY - 1) {
                v.y = A.Vin[(size_t)r * NY + j + 1];
                v.z = A.Pin[(size_t)r * NY + j + 1];
              }
            }
          } else {
            if (pK >= 0) {
              const int rr = min(s + TK - 1, rows - 1);
              v = ld_dev3(A.kout + ((size_t)pK * rows + rr) * 3 * (LANES + 1) + slot, LANES + 1);
            }
            if (slot == 0) {
              float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
              if (pC >= 0) {
                const int rr = min(s + 63 + TK, rows - 1);
                const float* cp = A.kout + ((size_t)pC * rows + rr) * 3 * (LANES + 1) + LANES;
                c.y = ld_dev(cp + (LANES + 1));
                c.z = ld_dev(cp + 2 * (LANES + 1));
              }
              v.y = c.y;
              v.z = c.z;
            }
          }
        } else if (it < NITEM) {
          const int e = it - BLK * (LANES + 1);
          const int u = e / TK, b = e % TK;
          const int s = blk * BLK + u;
          if (pG >= 0) {
            const int rr = min(s + 63, rows - 1);
            v = ld_dev3(A.gout + ((size_t)pG * rows + rr) * 3 * TK + b, TK);
          }
        }
        stage[q] = v;
      }
    };
    auto store_block = [&](int buf) {
#pragma unroll
      for (int q = 0; q < PER; q++) {
        const int it = tid + q * NT;
        if (it < BLK * (LANES + 1)) {
          Kst[buf][it / (LANES + 1)][it % (LANES + 1)] = stage[q];
        } else if (it < NITEM) {
          const int e = it - BLK * (LANES + 1);
          Gst[buf][e / TK][e % TK] = stage[q];
        }
      }
    };
    // Flags of this thread's cells for the 8 columns of a block (anti-diagonal layout).
    auto load_flags = [&](int blk, uint32_t* fl) {
#pragma unroll
      for (int u = 0; u < BLK; u++) {
        const int s = blk * BLK + u;
        const int d = d0 + s;
        uint32_t f = 0;
        if (wave_on && real_row && d >= 0 && d <= NX + NY - 2) f = A.fdiag[(size_t)d * NX + row];
        fl[u] = f;
      }
    };
    // Producer side: copy the G stream of block blk to global, then publish.
    auto publish = [&](int blk) {
      if (wave == TK - 1) {
        for (int e = lane; e < BLK * TK; e += LANES) {
          const int u = e / TK, b = e % TK;
          const float4 v = Gos[blk & 1][u][b];
          float* d = A.gout + ((size_t)me * rows + blk * BLK + u) * 3 * TK + b;
          st_dev(d, v.x);
          st_dev(d + TK, v.y);
          st_dev(d + 2 * TK, v.z);
        }
        wait_stores();
        if (lane == 0) __hip_atomic_store(&A.prog[me], (blk + 1) * BLK, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    };

    // Prologue: block 0 staged synchronously.
    wait_block(0);
    __syncthreads();
    load_block(0);
    store_block(0);
    uint32_t fcur[BLK], fnext[BLK];
    load_flags(0, fcur);
    __syncthreads();

    float UrOut = 0.f, Vreg = 0.f, pn = 0.f, md = 0.f;
    for (int blk = 0; blk < nblk; blk++) {
      const int cb = blk & 1;
      if (blk > 0) publish(blk - 1);
      const bool more = blk + 1 < nblk;
      if (more) {
        wait_block(blk + 1);
        __syncthreads();
        load_block(blk + 1);
        load_flags(blk + 1, fnext);
      }
#pragma unroll
      for (int u = 0; u < BLK; u++) {
        const int s = blk * BLK + u;
        const int pb = (s + 1) & 1;  // buffer written at step s-1
        const int wb = s & 1;
        const int j = s - lane - wave;
        // ---- inputs
        const float4 gin = Gst[cb][u][wave];
        float Ul = __int_as_float(
            __builtin_amdgcn_update_dpp(__float_as_int(gin.x), __float_as_int(UrOut), 0x138, 0xF, 0xF, false));
        float Ur, Vt, pnew;
        if (wave == 0) {
          const float4 kin = Kst[cb][u][lane];
          Ur = kin.x;
          Vt = kin.y;
          pnew = kin.z;
        } else {
          Ur = xU[pb][wave - 1][lane];
          if (lane == 0) {
            Vt = gin.y;
            pnew = gin.z;
          } else {
            const float2 vp = xVP[pb][wave - 1][lane - 1];
            Vt = vp.x;
            pnew = vp.y;
          }
        }
        float Vb = Vreg;
        float p = pn;
        pn = pnew;
        // ---- the reference cell update (fluid.go:196-229)
        const uint32_t f = fcur[u];
        if (f & CF_GS) {
          // spread neighbour bits 1..4 to bytes 0..3, then convert each byte (0 or 1)
          const uint32_t w = (((f >> 1) & 0xFu) * 0x00204081u) & 0x01010101u;
          const float sx0 = (float)(w & 0xFFu);
          const float sx1 = (float)((w >> 8) & 0xFFu);
          const float sy0 = (float)((w >> 16) & 0xFFu);
          const float sy1 = (float)(w >> 24);
          const float sum = sx0 + sx1 + sy0 + sy1;
          const float div = Ur - Ul + Vt - Vb;
          const float ad = fabsf(div);
          if (ad > md) md = ad;
          float q = -div / sum;
          q *= om;
          q *= A.damping;
          p += A.cp * q;
          Ul -= sx0 * q;
          Ur += sx1 * q;
          Vb -= sy0 * q;
          Vt += sy1 * q;
        }
        // ---- outputs
        UrOut = Ur;
        Vreg = Vt;
        xU[wb][wave][lane] = Ul;
        xVP[wb][wave][lane] = make_float2(Vb, p);
        if (lane == LANES - 1) {
          Gos[cb][u][wave].x = Ur;
          if (wave + 1 < TK) {
            Gos[cb][u][wave + 1].y = Vb;
            Gos[cb][u][wave + 1].z = p;
          }
        }
        if (wave == TK - 1) {
          float* kr = A.kout + ((size_t)me * rows + s) * 3 * (LANES + 1);
          st_dev(kr + lane, Ul);
          st_dev(kr + (LANES + 1) + lane + 1, Vb);
          st_dev(kr + 2 * (LANES + 1) + lane + 1, p);
        }
        if (last && real_row && j >= 0 && j <= NY - 1) {
          const size_t c = (size_t)row * NY + j;
          const bool tu = row >= 1 && j >= 1 && j <= NY - 2;
          const bool tv = row >= 1 && row <= NX - 2 && j >= 1;
          const bool tp = row >= 1 && row <= NX - 2 && j >= 1 && j <= NY - 2;
          if (tu) A.Uout[c] = Ul;
          if (tv) A.Vout[c] = Vb;
          if (tp) A.Pout[c] = p;
        }
        __syncthreads();
      }
      if (more) {
        store_block(cb ^ 1);
#pragma unroll
        for (int u = 0; u < BLK; u++) fcur[u] = fnext[u];
        __syncthreads();
      }
    }
    publish(nblk - 1);
    // per-sweep max |div|
    float m = md;
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if (lane == 0 && wave_on && m > 0.0f) atomicMax(&A.maxdiv[k], __float_as_uint(m));
    __syncthreads();
  }
}

// Faces and cells no interior cell touches keep their input values: U with i = 0 or
// j in {0, NY-1}; V with i in {0, NX-1} or j = 0; p on the ring. All lie on the ring.
__global__ void k_gs_ring(int NX, int NY, const float* __restrict__ Uin, const float* __restrict__ Vin,
                          const float* __restrict__ Pin, float* __restrict__ Uout, float* __restrict__ Vout,
                          float* __restrict__ Pout) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nring = 2 * NX + 2 * (NY - 2);
  if (t >= nring) return;
  int i, j;
  if (t < NX) {
    i = t;
    j = 0;
  } else if (t < 2 * NX) {
    i = t - NX;
    j = NY - 1;
  } else if (t < 2 * NX + NY - 2) {
    i = 0;
    j = t - 2 * NX + 1;
  } else {
    i = NX - 1;
    j = t - 2 * NX - (NY - 2) + 1;
  }
  const size_t c = (size_t)i * NY + j;
  if (!(i >= 1 && j >= 1 && j <= NY - 2)) Uout[c] = Uin[c];
  if (!(i >= 1 && i <= NX - 2 && j >= 1)) Vout[c] = Vin[c];
  Pout[c] = Pin[c];
}

// Zero the counters of both runs and the per-sweep maxima (graph-safe: no memset).
__global__ void k_gs_init(int32_t* prog, int nprog, int32_t* ticket, uint32_t* maxdiv, int nmd, int32_t* ctrl) {
  for (int t = threadIdx.x + blockIdx.x * blockDim.x; t < 2 * nprog; t += blockDim.x * gridDim.x) prog[t] = 0;
  for (int t = threadIdx.x + blockIdx.x * blockDim.x; t < nmd; t += blockDim.x * gridDim.x) maxdiv[t] = 0u;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ticket[0] = 0;
    ticket[1] = 0;
    ctrl[1] = 0;
  }
}

// Early exit of solveSingleGrid (fluid.go:175-177): the first sweep whose max |div| is
// below 1e-5 ends the solve; if that is not the last sweep, request a re-run of k*+1.
__global__ void k_gs_decide(const uint32_t* maxdiv, int N, int early_exit, int32_t* ctrl) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int kept = N;
  if (early_exit)
    for (int k = 0; k < N; k++)
      if (__uint_as_float(maxdiv[k]) < 1e-5f) {
        kept = k + 1;
        break;
      }
  ctrl[1] = kept < N ? kept : 0;
}

template <int TK>
static int launch_tk(Ctx* c, GsArgs& A, int grid) {
  hipLaunchKernelGGL(k_gs_wave<TK>, dim3(grid), dim3(TK * LANES), 0, c->stream, A);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

static bool tile_empty_h(int g, int kb, int K, int TK, int NX) {
  int k0 = kb * TK;
  if (k0 >= K) return true;
  int kmax = std::min(k0 + TK, K) - 1;
  int rmin = LANES * g - kmax, rmax = LANES * g + LANES - 1 - k0;
  return rmax < 0 || rmin > NX - 1;
}

static int pick_tk(Ctx* c, int N) {
  if (c->prm.block_sweeps == 8 || c->prm.block_sweeps == 16) return c->prm.block_sweeps;
  return N >= 64 ? 16 : 8;
}

static int layout(Ctx* c, int N) {
  GsWork& W = c->gs;
  const int tk = pick_tk(c, N);
  if (W.n_alloc == N && W.tk == tk) return FL_OK;
  pressure_free(c);
  W.tk = tk;
  W.n_alloc = N;
  W.G = (c->NX + N - 1 + LANES - 1) / LANES;
  W.KB = (N + tk - 1) / tk;
  const int steps = c->NY + LANES - 1 + tk - 1;
  W.rows = (steps + BLK - 1) / BLK * BLK;
  std::vector<std::pair<long, int32_t>> order;
  for (int g = 0; g < W.G; g++)
    for (int kb = 0; kb < W.KB; kb++)
      if (!tile_empty_h(g, kb, N, tk, c->NX)) order.push_back({(long)(LANES * g + kb * tk) * 65536 + kb, (g << 16) | kb});
  std::sort(order.begin(), order.end());
  std::vector<int32_t> tiles;
  for (auto& o : order) tiles.push_back(o.second);
  W.ntiles = (int)tiles.size();
  const size_t nt = (size_t)W.G * W.KB;
  W.kout_elems = nt * W.rows * 3 * (LANES + 1);
  W.gout_elems = nt * W.rows * 3 * tk;
  FL_HIP(hipMalloc(&W.d_tile, sizeof(int32_t) * std::max(1, W.ntiles)));
  FL_HIP(hipMalloc(&W.d_kout, sizeof(float) * W.kout_elems));
  FL_HIP(hipMalloc(&W.d_gout, sizeof(float) * W.gout_elems));
  FL_HIP(hipMalloc(&W.d_prog, sizeof(int32_t) * 2 * nt));
  FL_HIP(hipMalloc(&W.d_ticket, sizeof(int32_t) * 2));
  FL_HIP(hipMemcpy(W.d_tile, tiles.data(), sizeof(int32_t) * W.ntiles, hipMemcpyHostToDevice));
  int per_cu = 0, cus = 0;
  hipDeviceProp_t prop;
  FL_HIP(hipGetDeviceProperties(&prop, c->dev));
  cus = prop.multiProcessorCount;
  if (tk == 16)
    FL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_gs_wave<16>, 16 * LANES, 0));
  else
    FL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_gs_wave<8>, 8 * LANES, 0));
  W.grid = std::max(1, std::min(W.ntiles, std::max(1, per_cu) * cus));
  return FL_OK;
}

bool gs_wave_supported(Ctx* c, int numIters) { return numIters >= 1 && c->NX >= 3 && c->NY >= 3; }

int gs_wave_solve(Ctx* c, int N, float cp) {
  int rc = layout(c, N);
  if (rc) return rc;
  GsWork& W = c->gs;
  if (rc) return rc;
  if (c->flags_dirty && (rc = launch_flags(c))) return rc;
  const int nprog = W.G * W.KB;
  const int nring = 2 * c->NX + 2 * (c->NY - 2);
  hipLaunchKernelGGL(k_gs_ring, dim3((nring + 255) / 256), dim3(256), 0, c->stream, c->NX, c->NY, c->U, c->V, c->p,
                     c->U2, c->V2, c->T3);
  hipLaunchKernelGGL(k_gs_init, dim3(4), dim3(256), 0, c->stream, W.d_prog, nprog, W.d_ticket, c->d_maxdiv,
                     2 * c->nmax, c->d_ctrl);
  FL_HIP(hipGetLastError());
  GsArgs A;
  A.NX = c->NX;
  A.NY = c->NY;
  A.G = W.G;
  A.KB = W.KB;
  A.rows = W.rows;
  A.ntiles = W.ntiles;
  A.K = N;
  A.tiles = W.d_tile;
  A.Uin = c->U;
  A.Vin = c->V;
  A.Pin = c->p;
  A.Uout = c->U2;
  A.Vout = c->V2;
  A.Pout = c->T3;
  A.fdiag = c->flags_diag;
  A.omega = c->d_omega;
  A.damping = c->prm.pressure_damping;
  A.cp = cp;
  A.kout = W.d_kout;
  A.gout = W.d_gout;
  A.ctrl = c->d_ctrl;
  for (int which = 0; which < 2; which++) {
    A.which = which;
    A.prog = W.d_prog + which * nprog;
    A.ticket = W.d_ticket + which;
    A.maxdiv = c->d_maxdiv + which * c->nmax;
    rc = W.tk == 16 ? launch_tk<16>(c, A, W.grid) : launch_tk<8>(c, A, W.grid);
    if (rc) return rc;
    if (which == 0) {
      hipLaunchKernelGGL(k_gs_decide, dim3(1), dim3(64), 0, c->stream, c->d_maxdiv, N, c->prm.early_exit,
                         c->d_ctrl);
      FL_HIP(hipGetLastError());
    }
  }
  std::swap(c->U, c->U2);
  std::swap(c->V, c->V2);
  std::swap(c->p, c->T3);
  c->host_sweeps = -1;
  return FL_OK;
}

int pressure_free(Ctx* c) {
  GsWork& W = c->gs;
  if (W.d_tile) (void)hipFree(W.d_tile);
  if (W.d_kout) (void)hipFree(W.d_kout);
  if (W.d_gout) (void)hipFree(W.d_gout);
  if (W.d_prog) (void)hipFree(W.d_prog);
  if (W.d_ticket) (void)hipFree(W.d_ticket);
  W = GsWork{};
  return FL_OK;
}

}  // namespace fl
