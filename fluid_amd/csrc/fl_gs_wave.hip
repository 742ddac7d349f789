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
// complete before it stores the new progress value (publish_rows: a full wait at the next
// block, or a counted wait one block later), and consumers issue their data loads only after
// observing it.
//
// Cells outside the interior or solid pass their faces through unchanged, exactly like
// the skipped cells of the serial loop. Faces no interior cell touches keep their input
// values. The result (U, V and p) goes to separate output buffers, so the early exit (stop after the
// first sweep with max|div| < 1e-5, fluid.go:175-177) is a re-run of fewer sweeps from
// the untouched input, decided on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <type_traits>
#include <vector>

#include "fl_internal.h"

namespace fl {

static constexpr int LANES = 64;
static constexpr int SPIN_LIMIT = 1 << 22;     // bounded wait (about a second); then error and continue

// One launch of the solve (main run or early-exit re-run).
struct GsRun {
  const int2* tiles;  // ticket order: (simulation, (g << 16) | kb)
  int ntiles;
  int which;          // 0 main run (K sweeps), 1 re-run (ctrl[1] sweeps of each simulation)
  int K;
  int32_t* ticket;
  const int32_t* rerun;  // re-run launch only: nonzero when some simulation asked for one
  const float* omega;
  float damping;
  int rPin, rPout;    // roles of the initial p (read; < 0: p starts at zero) and of the solve's p
  int defer;          // when handoff rows are published (publish_rows below): 0 at the next
                      // block's start, 1 one block later, 2 halfway through the next block
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
// Wait until at most N of this wave's vector memory operations are in flight. The wave's
// stores retire in issue order, so when at least N stores were issued after a given store,
// that store is complete once this returns (loads issued in between only lengthen the wait).
template <int N> __device__ __forceinline__ void wait_stores_but() {
  static_assert(N >= 0 && N <= 63, "vmcnt is a 6-bit count");
  asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

__device__ __forceinline__ int ld_prog(const int32_t* p) {
  return __hip_atomic_load(const_cast<int32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wait until producer progress reaches what a block needs: seen[x] >= need[x] for the G and K
// producers (need 0: nothing to wait for). The counters still short are loaded together, so a
// block whose producers have both published costs one round trip. The wait is bounded: after
// SPIN_LIMIT polls, or as soon as another wait of this simulation has expired (err set), it
// sets err and gives up, so a solve whose producers stall ends within about a second with
// FL_ERR_INTERNAL instead of every later wait running out its own limit.
__device__ __forceinline__ void wait_producers(const int32_t* pg, int needG, const int32_t* pk, int needK, int& seenG,
                                               int& seenK, int32_t* err) {
  for (int it = 0;; it++) {
    const bool wG = seenG < needG, wK = seenK < needK;
    if (!wG && !wK) return;
    if (it > SPIN_LIMIT || ((it & 255) == 255 && ld_prog(err) != 0)) {
      atomicOr(err, 1);
      seenG = max(seenG, needG);
      seenK = max(seenK, needK);
      return;
    }
    if (it) __builtin_amdgcn_s_sleep(2);
    int vG = 0, vK = 0;
    if (wG) vG = ld_prog(pg);
    if (wK) vK = ld_prog(pk);
    if (wG) seenG = vG;
    if (wK) seenK = vK;
  }
}

__device__ __forceinline__ const SimDesc& sim_desc(const OneSim& m, int) { return m.d; }
__device__ __forceinline__ const SimDesc& sim_desc(const ManySims& m, int e) { return m.sims[e]; }

// TK sweeps per tile (one wave each), BLK columns per staging and publish block (BLK divides 64).
// SNAP: the main run also writes the first sweep's U, V, p to T0, T1, T2, so that an early exit
// after sweep 0 needs no re-run (k_gs_take_first); a separate instantiation, so the usual kernel
// carries none of its code or registers.
// 8-sweep tiles of up to 8 columns per block are held to 128 VGPRs (at least 4 waves per SIMD),
// so two workgroups share a CU: the batched kernel needs 137 otherwise and fits without spills.
template <int TK, int BLK, class M, bool SNAP = false>
__global__ __launch_bounds__(TK * LANES, (TK == 8 && BLK <= 8) ? 4 : 1) void k_gs_wave(M m, GsRun R) {
  static_assert(LANES % BLK == 0, "progress is published in whole blocks that divide the tile height");
  const int lane = threadIdx.x & (LANES - 1);
  // wave-uniform, and held in a scalar register so every test on it is a scalar branch
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / LANES);
  const int tid = threadIdx.x;
  constexpr int NT = TK * LANES;

  __shared__ float xU[2][TK][LANES];
  __shared__ float2 xVP[2][TK][LANES];
  __shared__ float4 Kst[2][BLK][LANES + 1];
  __shared__ float4 Gst[2][BLK][TK];
  __shared__ float4 Gos[2][BLK][TK];
  __shared__ int s_tile;
  // a re-run launch with nothing to re-run ends here, in every workgroup alike
  if (R.rerun && *R.rerun == 0) return;

  for (;;) {
    if (tid == 0) s_tile = atomicAdd(R.ticket, 1);
    __syncthreads();
    const int t = __builtin_amdgcn_readfirstlane(s_tile);
    __syncthreads();
    if (t >= R.ntiles) break;
    const int2 te = R.tiles[t];
    const int sim = te.x, g = te.y >> 16, kb = te.y & 0xffff;
    const SimDesc& D = sim_desc(m, sim);
    const int NX = D.g.NX, NY = D.g.NY, rows = D.rows, KB = D.KB;
    const int nblk = rows / BLK;
    const int me = g * KB + kb;
    int32_t* prog = D.prog + R.which * D.G * KB;
    int32_t* err = D.ctrl + 3;
    const int K = R.which ? D.ctrl[1] : R.K;
    if (tile_empty(g, kb, K, TK, NX)) {
      if (tid == 0) __hip_atomic_store(&prog[me], rows, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      continue;
    }
    const float* Uin = D.buf[B_U];
    const float* Vin = D.buf[B_V];
    const float* Pin = R.rPin >= 0 ? D.buf[R.rPin] : nullptr;
    float* Uout = D.buf[B_U2];
    float* Vout = D.buf[B_V2];
    float* Pout = D.buf[R.rPout];
    const uint8_t* fdiag = D.diag;
    float* kout = D.kout;
    float* gout = D.gout;
    const float cp = D.cp;
    const float damping = R.damping;
    const int k0 = kb * TK;
    const int k = k0 + wave;
    const bool wave_on = k < K;
    // wave-uniform in a scalar register: the other waves skip the output block with one branch
    const int last = __builtin_amdgcn_readfirstlane((int)(k == K - 1));
    const int first_out = SNAP ? __builtin_amdgcn_readfirstlane((int)(R.which == 0 && k == 0)) : 0;
    const int i0 = LANES * g - k0;           // row of (lane 0, wave 0)
    const int row = i0 + lane - wave;        // this thread's row
    const bool real_row = row >= 0 && row <= NX - 1;
    // this sweep's relaxation, read once into a scalar register: a vector-loaded value used in
    // the step loop would make every step wait for all of the wave's memory operations
    const float om = __int_as_float(
        __builtin_amdgcn_readfirstlane(__float_as_int(wave_on ? R.omega[min(k, K - 1)] : 0.0f)));
    const int d0 = LANES * g - k0 - 2 * wave;  // anti-diagonal of column 0 for this wave

    // Producers: G stream (g-1,kb), K stream (g,kb-1), corner (g-1,kb-1).
    const int pG = (g > 0 && !tile_empty(g - 1, kb, K, TK, NX)) ? (g - 1) * KB + kb : -1;
    const int pK = (kb > 0 && !tile_empty(g, kb - 1, K, TK, NX)) ? g * KB + kb - 1 : -1;
    const int pC = (g > 0 && kb > 0 && !tile_empty(g - 1, kb - 1, K, TK, NX)) ? (g - 1) * KB + kb - 1 : -1;
    int seenG = 0, seenK = 0;  // producer progress last observed (thread 0)

    // Wait until every producer has published the rows block `blk` stages: G stream rows up to
    // s1 + 63, K stream rows up to s1 + TK - 1, corner rows up to s1 + 63 + TK (s1 = the block's
    // last step). The corner producer (g-1, kb-1) is the G producer's K producer, and
    // progress is published in whole blocks: G progress >= s1 + 64 = BLK (blk+1) + 63 means
    // progress >= BLK (blk+1) + 64 = s1 + 65, which the G producer could publish only after it
    // waited for corner progress >= s1 + 64 + TK (both clamped to rows). So with a G producer
    // the corner needs no poll of its own; without one (g-1, kb) is empty, and the corner is
    // polled in the G slot.
    const int pGp = pG >= 0 ? pG : pC;
    const int lagG = pG >= 0 ? LANES : LANES + TK;
    auto wait_block = [&](int blk) {
      if (tid == 0) {
        const int s1 = blk * BLK + BLK - 1;  // last step of the block
        const int needG = pGp >= 0 ? min(s1 + lagG, rows) : 0;
        const int needK = pK >= 0 ? min(s1 + TK, rows) : 0;
        wait_producers(prog + max(pGp, 0), needG, prog + max(pK, 0), needK, seenG, seenK, err);
      }
    };

    // Staged boundary items of one block: K stream BLK x 65, G stream BLK x TK.
    // Items [0, KITEM) are the K stream; the G stream starts at the next wave boundary, so no
    // wave holds lanes of both (a wave running both branches would wait, in the second, for
    // the loads the first issued into the registers both use).
    constexpr int KITEM = BLK * (LANES + 1);
    constexpr int GITEM0 = (KITEM + LANES - 1) / LANES * LANES;
    constexpr int NITEM = GITEM0 + BLK * TK;
    constexpr int PER = (NITEM + NT - 1) / NT;
    float4 stage[PER];
    float2 corner[PER];  // K-stream lane 0: V and p of the corner producer, merged in store_block
    auto load_block = [&](int blk) {
#pragma unroll
      for (int q = 0; q < PER; q++) {
        const int it = tid + q * NT;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        float2 cz = make_float2(0.f, 0.f);
        if (it < KITEM) {
          const int u = it / (LANES + 1), slot = it % (LANES + 1);
          const int s = blk * BLK + u;
          if (kb == 0) {
            // initial state: lane `slot` of wave 0 is row i0+slot at column s-slot
            const int r = i0 + slot, j = s - slot;
            if (slot < LANES && j >= 0 && j <= NY - 1) {
              if (r + 1 >= 0 && r + 1 <= NX - 1) v.x = Uin[(size_t)(r + 1) * NY + j];
              if (r >= 0 && r <= NX - 1 && j + 1 <= NY - 1) {
                v.y = Vin[(size_t)r * NY + j + 1];
                v.z = Pin ? Pin[(size_t)r * NY + j + 1] : 0.0f;
              }
            }
          } else {
            if (pK >= 0) {
              const int rr = min(s + TK - 1, rows - 1);
              v = ld_dev3(kout + ((size_t)pK * rows + rr) * 3 * (LANES + 1) + slot, LANES + 1);
            }
            // lane 0 takes V and p from the corner producer instead (zero without one); the
            // values are merged in store_block, so no wait for these loads happens here
            if (slot == 0 && pC >= 0) {
              const int rr = min(s + 63 + TK, rows - 1);
              const float* cpp = kout + ((size_t)pC * rows + rr) * 3 * (LANES + 1) + LANES;
              cz.x = ld_dev(cpp + (LANES + 1));
              cz.y = ld_dev(cpp + 2 * (LANES + 1));
            }
          }
        } else if (it >= GITEM0 && it < NITEM) {
          const int e = it - GITEM0;
          const int u = e / TK, b = e % TK;
          const int s = blk * BLK + u;
          if (pG >= 0) {
            const int rr = min(s + 63, rows - 1);
            v = ld_dev3(gout + ((size_t)pG * rows + rr) * 3 * TK + b, TK);
          }
        }
        stage[q] = v;
        corner[q] = cz;
      }
    };
    auto store_block = [&](int buf) {
#pragma unroll
      for (int q = 0; q < PER; q++) {
        const int it = tid + q * NT;
        if (it < KITEM) {
          float4 v = stage[q];
          // lane 0 of each K-stream row takes V and p of the corner producer (zero without one)
          const bool lane0 = kb > 0 && it % (LANES + 1) == 0;
          v.y = lane0 ? corner[q].x : v.y;
          v.z = lane0 ? corner[q].y : v.z;
          Kst[buf][it / (LANES + 1)][it % (LANES + 1)] = v;
        } else if (it >= GITEM0 && it < NITEM) {
          const int e = it - GITEM0;
          Gst[buf][e / TK][e % TK] = stage[q];
        }
      }
    };
    // Flags of this thread's cells for the BLK columns of a block (anti-diagonal layout), one
    // byte load each into its own register.
    auto load_flags = [&](int blk, uint32_t* fl) {
#pragma unroll
      for (int u = 0; u < BLK; u++) {
        const int s = blk * BLK + u;
        const int dd = d0 + s;
        uint32_t f = 0;
        if (wave_on && real_row && dd >= 0 && dd <= NX + NY - 2) f = fdiag[(size_t)dd * NX + row];
        fl[u] = f;
      }
    };
    // The flags of the block being run: with 8-sweep tiles packed four to a register at the
    // block's end (when the loads have returned; packing them as they arrive would make the wave
    // wait for them, and for every load issued before them, at the block's start). That holds
    // the 8x8 kernel to 119 VGPRs, two workgroups per CU instead of one: 4096^2 cavity 12.42 ->
    // 11.89 ms/step. The unpacking costs a bit-field extract per step, which loses 5% where it
    // does not change the occupancy (16-sweep tiles), so other shapes keep one flag a register.
    constexpr bool PACK = TK == 8;
    constexpr int FW = PACK ? (BLK + 3) / 4 : BLK;
    auto take_flags = [&](const uint32_t* fl, uint32_t* cur) {
      if constexpr (PACK) {
#pragma unroll
        for (int w = 0; w < FW; w++) cur[w] = 0;
#pragma unroll
        for (int u = 0; u < BLK; u++) cur[u / 4] |= fl[u] << (8 * (u % 4));
      } else {
#pragma unroll
        for (int u = 0; u < BLK; u++) cur[u] = fl[u];
      }
    };
    // Producer side, in two halves run by the last wave (which also stores the K-stream
    // rows, three stores per step). publish_data(blk) copies the G stream of block blk to
    // global without waiting. publish_rows(n) makes handoff rows [0, n) visible: it waits
    // for the wave's stores of those rows, then stores the progress counter. On time
    // (R.defer 0), the rows of block blk are published at the start of block blk + 1 after a
    // full wait. With R.defer 1 they are published at the start of block blk + 2, right after
    // publish_data(blk + 1): by then the wave has issued at least NEWER stores after them
    // (the 3 * BLK K-stream stores of block blk + 1 and the G copy of block blk + 1), and
    // stores retire in issue order, so waiting until at most NEWER are in flight completes
    // them without waiting for the stores just issued. With R.defer 2 they are published
    // after step MID - 1 of block blk + 1, when the 3 * MID K-stream stores of that block's
    // first MID steps are the newer ones. Deferring removes a store round trip from every
    // block and lengthens the pipeline fill by the steps of delay per producer hop. After the
    // loop the last blocks are published with a full wait.
    constexpr int NEWER = 3 * BLK + 3 * ((BLK * TK + LANES - 1) / LANES);
    constexpr int MID = BLK / 2;
    auto publish_data = [&](int blk) {
      if (wave == TK - 1) {
        for (int e = lane; e < BLK * TK; e += LANES) {
          const int u = e / TK, b = e % TK;
          const float4 v = Gos[blk & 1][u][b];
          float* dst = gout + ((size_t)me * rows + blk * BLK + u) * 3 * TK + b;
          st_dev(dst, v.x);
          st_dev(dst + TK, v.y);
          st_dev(dst + 2 * TK, v.z);
        }
      }
    };
    auto publish_rows = [&](int n, bool all) {
      if (wave == TK - 1) {
        if (all)
          wait_stores();
        else
          wait_stores_but<NEWER>();
        if (lane == 0) __hip_atomic_store(&prog[me], n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    };
    auto publish_rows_mid = [&](int n) {
      if (wave == TK - 1) {
        wait_stores_but<3 * MID>();
        if (lane == 0) __hip_atomic_store(&prog[me], n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    };

    // Prologue: block 0 staged synchronously.
    wait_block(0);
    __syncthreads();
    load_block(0);
    store_block(0);
    uint32_t fcur[FW], fnext[BLK];
    if constexpr (PACK) {
      load_flags(0, fnext);
      take_flags(fnext, fcur);
    } else {
      load_flags(0, fcur);
    }
    __syncthreads();

    float UrOut = 0.f, Vreg = 0.f, pn = 0.f, md = 0.f;
    for (int blk = 0; blk < nblk; blk++) {
      const int cb = blk & 1;
      if (blk > 0) {
        publish_data(blk - 1);
        if (R.defer == 0)
          publish_rows(blk * BLK, true);
        else if (R.defer == 1 && blk > 1)
          publish_rows((blk - 1) * BLK, false);
      }
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
        const uint32_t f = PACK ? (fcur[u / 4] >> (8 * (u % 4))) & 0xFFu : fcur[u];
        const bool plain = __all((f & 0x3Fu) == 0x3Fu);
        if (std::is_same<M, OneSim>::value ? __builtin_expect(plain, 1) : plain) {
          // Every lane holds a fluid interior cell whose four neighbours are fluid: s = 4 and
          // all weights are 1, so (-div)/4 is the exact product with 0.25 and the face updates
          // are plain adds -- the same float results with a third of the instructions. For a
          // single simulation laid out as the fall-through path (1% per step on the 4096^2
          // cavity and the 256^2 jet); the batched 8x8 kernel would spill at 128 VGPRs with it.
          const float div = Ur - Ul + Vt - Vb;
          const float ad = fabsf(div);
          if (ad > md) md = ad;
          float q = -div * 0.25f;
          q *= om;
          q *= damping;
          p += cp * q;
          Ul -= q;
          Ur += q;
          Vb -= q;
          Vt += q;
        } else if (f & CF_GS) {
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
          q *= damping;
          p += cp * q;
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
          float* kr = kout + ((size_t)me * rows + s) * 3 * (LANES + 1);
          st_dev(kr + lane, Ul);
          st_dev(kr + (LANES + 1) + lane + 1, Vb);
          st_dev(kr + 2 * (LANES + 1) + lane + 1, p);
        }
        // halfway publication of the previous block's rows (R.defer 2): the K-stream stores of
        // steps 0..MID-1 of this block are the 3 * MID newest of this wave
        if (u == MID - 1 && R.defer == 2 && blk > 0) publish_rows_mid(blk * BLK);
        if (last) {  // only the wave of the final sweep writes the results
          const int j = s - lane - wave;
          if (real_row && j >= 0 && j <= NY - 1) {
            const size_t c = (size_t)row * NY + j;
            const bool tu = row >= 1 && j >= 1 && j <= NY - 2;
            const bool tv = row >= 1 && row <= NX - 2 && j >= 1;
            const bool tp = row >= 1 && row <= NX - 2 && j >= 1 && j <= NY - 2;
            if (tu) Uout[c] = Ul;
            if (tv) Vout[c] = Vb;
            if (tp) Pout[c] = p;
          }
        }
        if (first_out) {  // the first sweep's result, kept for an early exit after it
          const int j = s - lane - wave;
          if (real_row && j >= 0 && j <= NY - 1) {
            const size_t c = (size_t)row * NY + j;
            if (row >= 1 && j >= 1 && j <= NY - 2) D.buf[B_T0][c] = Ul;
            if (row >= 1 && row <= NX - 2 && j >= 1) D.buf[B_T1][c] = Vb;
            if (row >= 1 && row <= NX - 2 && j >= 1 && j <= NY - 2) D.buf[B_T2][c] = p;
          }
        }
        __syncthreads();
      }
      if (more) {
        store_block(cb ^ 1);
        take_flags(fnext, fcur);
        __syncthreads();
      }
    }
    publish_data(nblk - 1);
    publish_rows(rows, true);  // rows = nblk * BLK
    // per-sweep max |div|
    float mv = md;
    for (int o = 32; o > 0; o >>= 1) mv = fmaxf(mv, __shfl_xor(mv, o, 64));
    if (lane == 0 && wave_on && mv > 0.0f) atomicMax(&D.maxdiv[R.which * D.nmax + k], __float_as_uint(mv));
    __syncthreads();
  }
}

// Faces and cells no interior cell touches keep their input values: U with i = 0 or
// j in {0, NY-1}; V with i in {0, NX-1} or j = 0; p on the ring. All lie on the ring.
__device__ __forceinline__ void ring_copy(const SimDesc& D, int i, int j, int rPin, int rPout) {
  const int NX = D.g.NX, NY = D.g.NY;
  const size_t c = (size_t)i * NY + j;
  if (!(i >= 1 && j >= 1 && j <= NY - 2)) D.buf[B_U2][c] = D.buf[B_U][c];
  if (!(i >= 1 && i <= NX - 2 && j >= 1)) D.buf[B_V2][c] = D.buf[B_V][c];
  D.buf[rPout][c] = rPin >= 0 ? D.buf[rPin][c] : 0.0f;
}
__device__ __forceinline__ void ring_desc(const OneSim& m, const SimDesc*& D, int& t) {
  D = &m.d;
  t = blockIdx.x * blockDim.x + threadIdx.x;
}
__device__ __forceinline__ void ring_desc(const ManySims& m, const SimDesc*& D, int& t) {
  const RingItem it = m.rings[blockIdx.x];
  D = &m.sims[it.sim];
  t = it.blk * blockDim.x + threadIdx.x;
}
template <class M> __global__ void k_gs_ring(M m, int rPin, int rPout) {
  const SimDesc* D;
  int t;
  ring_desc(m, D, t);
  const int NX = D->g.NX, NY = D->g.NY;
  if (t < NX) {
    ring_copy(*D, t, 0, rPin, rPout);
    ring_copy(*D, t, NY - 1, rPin, rPout);
  }
  if (t < NY) {
    ring_copy(*D, 0, t, rPin, rPout);
    ring_copy(*D, NX - 1, t, rPin, rPout);
  }
}

// Zero the counters of both runs, the per-sweep maxima and the re-run requests of every
// simulation of the launch (graph-safe: no memset).
__global__ void k_gs_init(int32_t* prog, int nprog, uint32_t* maxdiv, int nmd, int32_t* ctrl, int nsim,
                          int32_t* ticket) {
  const int stride = blockDim.x * gridDim.x;
  for (int t = threadIdx.x + blockIdx.x * blockDim.x; t < nprog; t += stride) prog[t] = 0;
  for (int t = threadIdx.x + blockIdx.x * blockDim.x; t < nmd; t += stride) maxdiv[t] = 0u;
  for (int t = threadIdx.x + blockIdx.x * blockDim.x; t < nsim; t += stride) ctrl[4 * t] = ctrl[4 * t + 1] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ticket[0] = 0;  // main run
    ticket[1] = 0;  // re-run
    ticket[2] = 0;  // re-run requested by some simulation
  }
}

// k_gs_ring and k_gs_init of a single simulation in one launch (both only precede the solve).
__global__ void k_gs_prep(OneSim m, int rPin, int rPout, int32_t* prog, int nprog, uint32_t* maxdiv, int nmd,
                          int32_t* ctrl, int32_t* ticket) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const SimDesc& D = m.d;
  const int NX = D.g.NX, NY = D.g.NY;
  if (t < NX) {
    ring_copy(D, t, 0, rPin, rPout);
    ring_copy(D, t, NY - 1, rPin, rPout);
  }
  if (t < NY) {
    ring_copy(D, 0, t, rPin, rPout);
    ring_copy(D, NX - 1, t, rPin, rPout);
  }
  const int stride = blockDim.x * gridDim.x;
  for (int i = t; i < nprog; i += stride) prog[i] = 0;
  for (int i = t; i < nmd; i += stride) maxdiv[i] = 0u;
  if (t == 0) {
    ctrl[0] = ctrl[1] = 0;
    ticket[0] = ticket[1] = ticket[2] = 0;
  }
}

// Early exit of solveSingleGrid (fluid.go:175-177), one wave per simulation: the first
// sweep whose max |div| is below 1e-5 ends the solve; if that is not the last sweep,
// request a re-run of k*+1 sweeps (and flag the launch, rerun[0], for the re-run kernel),
// or, when the main run kept the first sweep's result (snap) and k* = 0, mark it to be
// taken instead (ctrl[0], k_gs_take_first). ctrl[1] holds the kept sweeps either way.
// hint (host memory, may be null): whether this solve ended after sweep 0, which decides
// whether the next solve keeps the first sweep's result.
// Lane l scans sweeps l, l+64, ...; the wave's minimum is the first such sweep.
template <class M>
__global__ void k_gs_decide(M m, int nsim, int N, int32_t* rerun, int snap, int32_t* hint) {
  const int e = blockIdx.x * (blockDim.x / LANES) + threadIdx.x / LANES;
  const int lane = threadIdx.x & (LANES - 1);
  if (e >= nsim) return;
  const SimDesc& D = sim_desc(m, e);
  int first = N;
  if (D.early_exit)
    for (int k = lane; k < N; k += LANES)
      if (__uint_as_float(D.maxdiv[k]) < 1e-5f) {
        first = k;
        break;
      }
  for (int o = LANES / 2; o > 0; o >>= 1) first = min(first, __shfl_xor(first, o, LANES));
  if (lane == 0) {
    const int kept = first < N ? first + 1 : N;
    const bool take = snap && kept == 1 && kept < N;
    D.ctrl[0] = take ? 1 : 0;
    D.ctrl[1] = kept < N ? kept : 0;
    if (kept < N && !take) atomicOr(rerun, 1);
    if (hint) *hint = kept == 1 && kept < N ? 1 : 0;
  }
}

// An early exit after sweep 0 with the first sweep's result kept: copy it into the solve's
// outputs (the faces and p the sweeps write; the ring is already there) and its max |div| into
// the re-run's record, which is where a re-run of one sweep would have put it.
template <class M> __global__ void k_gs_take_first(M m, int rPout) {
  const SimDesc& D = sim_desc(m, 0);
  if (D.ctrl[0] == 0) return;
  const int NX = D.g.NX, NY = D.g.NY;
  const size_t n = (size_t)NX * NY;
  for (size_t c = threadIdx.x + (size_t)blockIdx.x * blockDim.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
    const int row = (int)(c / NY), j = (int)(c % NY);
    if (row >= 1 && j >= 1 && j <= NY - 2) D.buf[B_U2][c] = D.buf[B_T0][c];
    if (row >= 1 && row <= NX - 2 && j >= 1) D.buf[B_V2][c] = D.buf[B_T1][c];
    if (row >= 1 && row <= NX - 2 && j >= 1 && j <= NY - 2) D.buf[rPout][c] = D.buf[B_T2][c];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) D.maxdiv[D.nmax] = D.maxdiv[0];
}

static bool tile_empty_h(int g, int kb, int K, int TK, int NX) {
  int k0 = kb * TK;
  if (k0 >= K) return true;
  int kmax = std::min(k0 + TK, K) - 1;
  int rmin = LANES * g - kmax, rmax = LANES * g + LANES - 1 - k0;
  return rmax < 0 || rmin > NX - 1;
}

// Tile geometry of one simulation for N sweeps with tk sweeps per tile and blk columns per
// block.
struct TileGeom {
  int G, KB, rows;
};
static TileGeom geom(int NX, int NY, int N, int tk, int blk) {
  TileGeom t;
  t.G = (NX + N - 1 + LANES - 1) / LANES;
  t.KB = (N + tk - 1) / tk;
  const int steps = NY + LANES - 1 + tk - 1;
  t.rows = (steps + blk - 1) / blk * blk;
  return t;
}
// Tiles are handed out in the order their producers let them start: a tile can start about
// 64 + 2 BLK columns after the tile one row-tile up and TK + 2 BLK after the tile one sweep block
// back (the consumer stages a block ahead and rows are published in whole blocks), so the key
// is that estimate of its start. Producers always get smaller keys (a topological order, which
// the persistent kernel needs: a workgroup never waits on a tile no workgroup has taken).
static void append_tiles(std::vector<std::pair<long long, int2>>& out, int sim, int NX, int N, int tk, int blk,
                         const TileGeom& t) {
  for (int g = 0; g < t.G; g++)
    for (int kb = 0; kb < t.KB; kb++)
      if (!tile_empty_h(g, kb, N, tk, NX)) {
        const long long start = (long long)(LANES + 2 * blk) * g + (long long)(tk + 2 * blk) * kb;
        long long key = (start << 32) | ((long long)sim << 12) | kb;
        out.push_back({key, make_int2(sim, (g << 16) | kb)});
      }
}

// Tile shapes the kernel is instantiated for: sweeps per tile (waves per workgroup) and columns
// per staging/publish block.
static bool shape_ok(int tk, int blk) {
  return (tk == 2 || tk == 4 || tk == 8 || tk == 16) && (blk == 4 || blk == 8 || blk == 16);
}
// Sweeps per tile: the knob if set; otherwise 4 for solves of at most 8 sweeps (Simulate's own
// count: one wave per SIMD, 300x251 frame 0.466 -> 0.447 ms), else 8: twice the tiles of 16 and
// a shorter pipeline fill, and with its flags packed two workgroups per CU. Measured against 16
// on the 4096^2 cavity (11.8 against 12.4 ms/step, each with its best publication) and on the
// 256-member ensemble (17.0 against 17.6).
int pick_tk(const fl_params& P, int N) {
  if (P.block_sweeps == 2 || P.block_sweeps == 4 || P.block_sweeps == 8 || P.block_sweeps == 16)
    return P.block_sweeps;
  return N <= 8 ? 4 : 8;
}
// Columns per block: the knob if set, else 8.
int pick_blk(const fl_params& P) { return (P.block_cols == 4 || P.block_cols == 16) ? P.block_cols : 8; }
// When handoff rows are published: on time (handoff_lag 1: GsRun.defer 0), one block late
// (2: defer 1) or halfway through the next block (3: defer 2); automatic (0) picks halfway
// when the tiles fill the persistent grid at least twice (the solve is throughput-bound and the
// per-block store wait costs more than the longer pipeline fill: 4096^2 cavity 12.10 ms/step on
// time, 12.15 one block late, 11.89 halfway), else on time.
static int defer_handoff(const fl_params& P, int ntiles, int grid) {
  if (P.handoff_lag >= 1 && P.handoff_lag <= 3) return P.handoff_lag - 1;
  return ntiles >= 2 * grid ? 2 : 0;
}

// The kernel of one tile shape and launch map, and its persistent grid for `ntiles` tiles: as
// many workgroups as fit on the device at once (and no more than there are tiles).
template <class M> struct GsKernels {
  using Fn = void (*)(M, GsRun);
  // snap: the instantiation that keeps the first sweep's result (single simulations only)
  static Fn get(int tk, int blk, bool snap = false) {
#define FL_GS_CASE(T, B)                               \
  if (tk == T && blk == B) {                           \
    if constexpr (std::is_same<M, OneSim>::value)      \
      if (snap) return k_gs_wave<T, B, M, true>;       \
    return snap ? nullptr : k_gs_wave<T, B, M, false>; \
  }
    FL_GS_CASE(2, 4) FL_GS_CASE(2, 8) FL_GS_CASE(2, 16)
    FL_GS_CASE(4, 4) FL_GS_CASE(4, 8) FL_GS_CASE(4, 16)
    FL_GS_CASE(8, 4) FL_GS_CASE(8, 8) FL_GS_CASE(8, 16)
    FL_GS_CASE(16, 4) FL_GS_CASE(16, 8) FL_GS_CASE(16, 16)
#undef FL_GS_CASE
    return nullptr;
  }
};
template <class M> static int occupancy_grid(int dev, int tk, int blk, int ntiles) {
  int per_cu = 0;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 1;
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, GsKernels<M>::get(tk, blk), tk * LANES, 0);
  return std::max(1, std::min(ntiles, std::max(1, per_cu) * prop.multiProcessorCount));
}
template <class M>
static int launch_gs(int tk, int blk, int grid, hipStream_t stream, const M& m, const GsRun& R, bool snap = false) {
  auto fn = GsKernels<M>::get(tk, blk, snap);
  if (!fn) {
    set_error("no Gauss-Seidel kernel for " + std::to_string(tk) + " sweeps x " + std::to_string(blk) + " columns");
    return FL_ERR_INTERNAL;
  }
  hipLaunchKernelGGL(fn, dim3(grid), dim3(tk * LANES), 0, stream, m, R);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

static int layout(Ctx* c, int N) {
  GsWork& W = c->gs;
  const int tk = pick_tk(c->prm, N);
  const int blk = pick_blk(c->prm);
  if (W.n_alloc == N && W.tk == tk && W.blk == blk) return FL_OK;
  if (!shape_ok(tk, blk)) return FL_ERR_INTERNAL;
  FL_HIP(hipStreamSynchronize(c->stream));  // earlier solves may still use the old layout
  pressure_free(c);
  const TileGeom t = geom(c->NX, c->NY, N, tk, blk);
  W.tk = tk;
  W.blk = blk;
  W.n_alloc = N;
  W.G = t.G;
  W.KB = t.KB;
  W.rows = t.rows;
  std::vector<std::pair<long long, int2>> order;
  append_tiles(order, 0, c->NX, N, tk, blk, t);
  std::sort(order.begin(), order.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
  std::vector<int2> tiles;
  for (auto& o : order) tiles.push_back(o.second);
  W.ntiles = (int)tiles.size();
  const size_t nt = (size_t)W.G * W.KB;
  W.kout_elems = nt * W.rows * 3 * (LANES + 1);
  W.gout_elems = nt * W.rows * 3 * tk;
  FL_HIP(hipMalloc(&W.d_tile, sizeof(int2) * std::max(1, W.ntiles)));
  FL_HIP(hipMalloc(&W.d_kout, sizeof(float) * W.kout_elems));
  FL_HIP(hipMalloc(&W.d_gout, sizeof(float) * W.gout_elems));
  FL_HIP(hipMalloc(&W.d_prog, sizeof(int32_t) * 2 * nt));
  FL_HIP(hipMalloc(&W.d_ticket, sizeof(int32_t) * 4));  // main, re-run, re-run flag
  FL_HIP(hipMemcpy(W.d_tile, tiles.data(), sizeof(int2) * W.ntiles, hipMemcpyHostToDevice));
  W.grid = occupancy_grid<OneSim>(c->dev, tk, blk, W.ntiles);
  return FL_OK;
}

bool gs_wave_supported(Ctx* c, int numIters) { return numIters >= 1 && c->NX >= 3 && c->NY >= 3; }

int gs_wave_solve(Ctx* c, int N, float cp) {
  int rc = layout(c, N);
  if (rc) return rc;
  GsWork& W = c->gs;
  if (c->flags_dirty && (rc = launch_flags(c))) return rc;
  SimDesc d = c->desc();
  d.cp = cp;
  OneSim m{d};
  const int nprog = W.G * W.KB;
  const int mr = std::max(c->NX, c->NY);
  // Simulate's fill(p, 0) left to this solve: the initial p is zero and is not read
  const int rPin = c->gs_p_zero ? -1 : (int)B_P;
  c->gs_p_zero = false;
  hipLaunchKernelGGL(k_gs_prep, dim3(std::max(4, (mr + 255) / 256)), dim3(256), 0, c->stream, m, rPin, (int)B_T3,
                     W.d_prog, 2 * nprog, c->d_maxdiv, 2 * c->nmax, c->d_ctrl, W.d_ticket);
  FL_HIP(hipGetLastError());
  GsRun R{};
  R.tiles = (const int2*)W.d_tile;
  R.ntiles = W.ntiles;
  R.K = N;
  R.omega = c->d_omega;
  R.damping = c->prm.pressure_damping;
  R.rPin = rPin;  // then p and T3 trade places
  R.rPout = B_T3;
  R.defer = defer_handoff(c->prm, W.ntiles, W.grid);
  W.defer = R.defer;
  // without the early exit no re-run can be requested: the decision and re-run are skipped
  const int runs = c->prm.early_exit ? 2 : 1;
  // Keep the first sweep's result (in T0-T2, scratch during the solve) when the last solve that
  // reported back ended after sweep 0: a quiescent field then takes it instead of a re-run
  // launch that costs as much as the main one on small grids.
  const bool snap = runs > 1 && N > 1 && c->snap_hint && __atomic_load_n(c->snap_hint, __ATOMIC_RELAXED) != 0;
  for (int which = 0; which < runs; which++) {
    R.which = which;
    R.ticket = W.d_ticket + which;
    R.rerun = which ? W.d_ticket + 2 : nullptr;
    if ((rc = launch_gs(W.tk, W.blk, W.grid, c->stream, m, R, snap && which == 0))) return rc;
    if (which == 0 && runs > 1) {
      hipLaunchKernelGGL(k_gs_decide<OneSim>, dim3(1), dim3(LANES), 0, c->stream, m, 1, N, W.d_ticket + 2,
                         (int)snap, N > 1 ? c->snap_hint : nullptr);
      FL_HIP(hipGetLastError());
      if (snap) {
        const int nb = (int)std::min<size_t>(1024, (c->N + 255) / 256);
        hipLaunchKernelGGL(k_gs_take_first<OneSim>, dim3(nb), dim3(256), 0, c->stream, m, (int)B_T3);
        FL_HIP(hipGetLastError());
      }
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

// ---- batched solve ------------------------------------------------------------------------
// All members' tiles in one ticket list, merged by hyperplane start; one persistent launch
// drains it (and one more for the members that early-exit). The initial p is read from role
// rPin (rPin < 0: all zero, Simulate's fill(p, 0)), the solve's p written to rPout, the
// velocity to U2, V2.
int batch_gs_layout(Batch* B, int N, std::vector<SimDesc>& descs) {
  const int tk = pick_tk(B->prm, N);
  const int blk = pick_blk(B->prm);
  if (!shape_ok(tk, blk)) return FL_ERR_INTERNAL;
  auto fill = [&]() {
    for (size_t e = 0; e < descs.size(); e++) {
      const SimDesc& f = B->gs_fields[e];
      SimDesc& d = descs[e];
      d.G = f.G;
      d.KB = f.KB;
      d.rows = f.rows;
      d.kout = f.kout;
      d.gout = f.gout;
      d.prog = f.prog;
    }
  };
  if (B->gs_n == N && B->tk == tk && B->blk == blk && B->gs_fields.size() == descs.size()) {
    fill();
    return FL_OK;
  }
  FL_HIP(hipStreamSynchronize(B->stream));  // earlier solves may still use the old layout
  if (B->d_tiles) (void)hipFree(B->d_tiles);
  if (B->kout) (void)hipFree(B->kout);
  if (B->gout) (void)hipFree(B->gout);
  if (B->prog) (void)hipFree(B->prog);
  if (B->ticket) (void)hipFree(B->ticket);
  B->d_tiles = nullptr;
  B->kout = B->gout = nullptr;
  B->prog = B->ticket = nullptr;
  std::vector<std::pair<long long, int2>> order;
  size_t kelems = 0, gelems = 0, nprog = 0;
  std::vector<TileGeom> tg(descs.size());
  for (size_t e = 0; e < descs.size(); e++) {
    tg[e] = geom(descs[e].g.NX, descs[e].g.NY, N, tk, blk);
    append_tiles(order, (int)e, descs[e].g.NX, N, tk, blk, tg[e]);
    const size_t nt = (size_t)tg[e].G * tg[e].KB;
    kelems += nt * tg[e].rows * 3 * (LANES + 1);
    gelems += nt * tg[e].rows * 3 * tk;
    nprog += 2 * nt;
  }
  std::sort(order.begin(), order.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
  std::vector<int2> tiles;
  for (auto& o : order) tiles.push_back(o.second);
  B->ntiles = (int)tiles.size();
  FL_HIP(hipMalloc(&B->d_tiles, sizeof(int2) * std::max(1, B->ntiles)));
  FL_HIP(hipMalloc(&B->kout, sizeof(float) * std::max<size_t>(kelems, 1)));
  FL_HIP(hipMalloc(&B->gout, sizeof(float) * std::max<size_t>(gelems, 1)));
  FL_HIP(hipMalloc(&B->prog, sizeof(int32_t) * std::max<size_t>(nprog, 1)));
  FL_HIP(hipMalloc(&B->ticket, sizeof(int32_t) * 4));  // main, re-run, re-run flag
  FL_HIP(hipMemcpy(B->d_tiles, tiles.data(), sizeof(int2) * B->ntiles, hipMemcpyHostToDevice));
  size_t ko = 0, go = 0, po = 0;
  B->gs_fields.assign(descs.size(), SimDesc{});
  for (size_t e = 0; e < descs.size(); e++) {
    SimDesc& d = B->gs_fields[e];
    d.G = tg[e].G;
    d.KB = tg[e].KB;
    d.rows = tg[e].rows;
    d.kout = B->kout + ko;
    d.gout = B->gout + go;
    d.prog = B->prog + po;
    const size_t nt = (size_t)d.G * d.KB;
    ko += nt * d.rows * 3 * (LANES + 1);
    go += nt * d.rows * 3 * tk;
    po += 2 * nt;
  }
  B->tk = tk;
  B->blk = blk;
  B->gs_n = N;
  B->gs_nprog = (int)nprog;
  B->grid = occupancy_grid<ManySims>(B->dev, tk, blk, B->ntiles);
  fill();
  return FL_OK;
}

int batch_gs_solve(Batch* B, int N, int rPin, int rPout) {
  ManySims m{B->d_sims, B->d_cells, B->d_rings};
  const int E = (int)B->members.size();
  hipLaunchKernelGGL(k_gs_ring<ManySims>, dim3(B->nring_items), dim3(256), 0, B->stream, m, rPin, rPout);
  // progress counters, per-sweep maxima and re-run requests of all members are contiguous
  hipLaunchKernelGGL(k_gs_init, dim3(256), dim3(256), 0, B->stream, B->prog, B->gs_nprog, B->maxdiv_pool,
                     2 * B->nmax * E, B->ctrl_pool, E, B->ticket);
  FL_HIP(hipGetLastError());
  GsRun R{};
  R.tiles = B->d_tiles;
  R.ntiles = B->ntiles;
  R.K = N;
  R.omega = B->d_omega;
  R.damping = B->prm.pressure_damping;
  R.rPin = rPin;
  R.rPout = rPout;
  R.defer = defer_handoff(B->prm, B->ntiles, B->grid);
  const int runs = B->prm.early_exit ? 2 : 1;  // no early exit: nothing to decide or re-run
  for (int which = 0; which < runs; which++) {
    R.which = which;
    R.ticket = B->ticket + which;
    R.rerun = which ? B->ticket + 2 : nullptr;
    int rc = launch_gs(B->tk, B->blk, B->grid, B->stream, m, R);
    if (rc) return rc;
    if (which == 0 && runs > 1) {
      // one wave per member, four per workgroup
      hipLaunchKernelGGL(k_gs_decide<ManySims>, dim3((E + 3) / 4), dim3(4 * LANES), 0, B->stream, m, E, N,
                         B->ticket + 2, 0, nullptr);
      FL_HIP(hipGetLastError());
    }
  }
  return FL_OK;
}

}  // namespace fl
