// Fluid oracle: CPU restatement of TheFellow/fluid `pkg/fluid` (Go) simulation step.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker for the HIP implementation in
// fluid_amd/. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// may load the library built from it. The product never calls into it.
//
// The restatement follows the reference algorithm operation by operation so that the
// float instantiation reproduces Go's float32 arithmetic bit for bit (Go rounds every
// float32 operation; build with -ffp-contract=off and without fast-math). The double
// instantiation is the FP64 model of the same algorithm: identical control flow and
// the same float32-valued constants, evaluated in double.
//
// Reference files restated here (paths relative to the reference root):
//   pkg/fluid/fluid.go      New, Simulate, every pass, multigrid, views, BFECC
//   pkg/fluid/walls.go      SetSolid, IsSolid, SetVelocity, AddSmoke, Reset
//   pkg/fluid/parallel.go   parallelRange (threading model of the CPU baseline)
//   pkg/fluid/pressure.go, smoke.go   ScalarField min/max conventions
//   main/main.go:200-249,474-486,747-858   presets, sources, jet, walls, strokes
//
// Additions with no reference counterpart, pinned by tests/test_precise_oracle.py against the
// reference sampling above (bit-equal where positions are exact in float32, equal in FP64 to
// rounding): PreciseAdvection (axisCells, sampleCells, disp), the back-trace in cell units.
//
// Third-party arithmetic: Go 1.24 `math.Sin` (pure Go, Cephes sin.c algorithm) is
// restated in go_sin() from the published Cephes method; it is pinned by the tests
// against the C library's sin over the whole argument domain the step uses.
// `math.Exp` (ApplyForceRadius) is evaluated with the C library's exp.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <thread>
#include <tuple>
#include <vector>

namespace orc {

// Planted mistakes. tests/test_oracle_pins.py shows that each pin of the restatement can
// fail: it loads a second build of this file (-DORC_PLANTABLE, _build/liboracle_fluid_plant.so)
// in which orc_set_plant(k) switches on mistake k, and checks that the pin meant for it
// reports a mismatch. In the default build PLANT(k) is the constant false and compiles away.
//   1 sampleField: corners (x1,y0) and (x0,y1) exchanged     2 sampleField: U/V half-cell offsets exchanged
//   3 BFECC velocity: backward trace with -dt               4 BFECC velocity: backward trace by fwd velocity
//   5 BFECC smoke: backward trace with -dt                  6 BFECC: 3x3 clamp skipped
//   7 BFECC: correction orig - (bwd - orig), no 1/2         8 confinement: force sign flipped
//   9 confinement: strength ignores the local speed        10 solveSingleGrid: omega at (iter+1)/N
//  11 applyViscosity: dt dropped from visc                 12 advectSmoke: diffusion factor 0.3 dropped
//  13 addTurbulence: speed |U|+|V|                         14 addTurbulence: factor min(|u|, 1)
//  15 residual: -div + lap                                 16 restriction: corner weight 1/8
//  17 coarse solve: relaxation 1                           18 prolongation: i and j fractions exchanged
//  19 correction: U faces with the opposite sign           20 GetAdaptiveTimeStep: upper clamp 1x basedt
//  21 GetAdaptiveTimeStep: speed sqrt(U^2+V^2)             22 drawLine: e2 > dy instead of e2 >= dy
//  23 cavity preset: source row at NY-3                    24 applySources: solid cells not skipped
//  25 Karman preset: disc centred at NX/2                  26 applyWallSettings: columns before rows
//  27 multigrid adjoint: residual's p term with the opposite sign
//  28 multigrid adjoint: coarse sweep's own (1 - relax) factor dropped
//  29 multigrid adjoint: correction's cp * lambda_p term dropped
#ifdef ORC_PLANTABLE
static int g_plant = 0;
#define PLANT(k) (g_plant == (k))
#else
#define PLANT(k) false
#endif

// ---------------------------------------------------------------------------
// Go builtin min/max for floats (Go 1.21 spec): NaN propagates,
// min(-0,+0) = -0, max(-0,+0) = +0.
template <class T> static inline T go_min(T a, T b) {
  if (a != a) return a;
  if (b != b) return b;
  if (a < b) return a;
  if (b < a) return b;
  return std::signbit(a) ? a : b;
}
template <class T> static inline T go_max(T a, T b) {
  if (a != a) return a;
  if (b != b) return b;
  if (a > b) return a;
  if (b > a) return b;
  return std::signbit(a) ? b : a;
}

// Go math.Sin restated from the Cephes sin.c algorithm (Moshier), as used by
// Go's pure-Go implementation for |x| < 2^29: octant reduction with the
// three-part Pi/4 split, then the sin or cos minimax polynomial.
static const double kSin[6] = {
    1.58962301576546568060e-10,  -2.50507477628578072866e-8,
    2.75573136213857245213e-6,   -1.98412698295895385996e-4,
    8.33333333332211858878e-3,   -1.66666666666666307295e-1,
};
static const double kCos[6] = {
    -1.13585365213876817300e-11, 2.08757008419747316778e-9,
    -2.75573141792967388112e-7,  2.48015872888517045348e-5,
    -1.38888888888730564116e-3,  4.16666666666665929218e-2,
};
static const double kPI4A = 7.85398125648498535156e-1;
static const double kPI4B = 3.77489470793079817668e-8;
static const double kPI4C = 2.69515142907905952645e-15;
static const double kFourOverPi = 0x1.45f306dc9c883p+0;  // 4/Pi rounded to double

double go_sin(double x) {
  if (x == 0 || x != x) return x;
  if (std::isinf(x)) return std::numeric_limits<double>::quiet_NaN();
  bool sign = false;
  if (x < 0) {
    x = -x;
    sign = true;
  }
  if (x >= 536870912.0) {
    // Payne-Hanek range; never reached by the step (seeds stay below 2^21).
    return std::numeric_limits<double>::quiet_NaN();
  }
  uint64_t j = (uint64_t)(x * kFourOverPi);
  double y = (double)j;
  if (j & 1) {
    j++;
    y++;
  }
  j &= 7;
  double z = ((x - y * kPI4A) - y * kPI4B) - y * kPI4C;
  if (j > 3) {
    sign = !sign;
    j -= 4;
  }
  double zz = z * z;
  if (j == 1 || j == 2) {
    y = 1.0 - 0.5 * zz +
        zz * zz * ((((((kCos[0] * zz) + kCos[1]) * zz + kCos[2]) * zz + kCos[3]) * zz + kCos[4]) * zz + kCos[5]);
  } else {
    y = z + z * zz * ((((((kSin[0] * zz) + kSin[1]) * zz + kSin[2]) * zz + kSin[3]) * zz + kSin[4]) * zz + kSin[5]);
  }
  if (sign) y = -y;
  return y;
}

// Turbulence noise for integer seed k: float32(math.Sin(float64(float32(k) * 0.01)))
// (fluid.go:509-514). Depends on position only.
static inline float turb_noise(int k) {
  float seed = (float)k * 0.01f;
  return (float)go_sin((double)seed);
}

// parallelRange (parallel.go:10-39): split [start,end) in ceil(total/workers) chunks.
static void prange(int threads, int start, int end, const std::function<void(int)>& fn) {
  int total = end - start;
  if (total <= 0) return;
  int workers = threads < 1 ? 1 : threads;
  if (workers > total) workers = total;
  if (workers == 1) {
    for (int i = start; i < end; i++) fn(i);
    return;
  }
  int chunk = (total + workers - 1) / workers;
  std::vector<std::thread> ts;
  for (int w = 0; w < workers; w++) {
    int s = start + w * chunk;
    int e = std::min(s + chunk, end);
    if (s >= end) break;
    ts.emplace_back([s, e, &fn]() {
      for (int i = s; i < e; i++) fn(i);
    });
  }
  for (auto& t : ts) t.join();
}

enum Field { F_U = 0, F_V, F_NEWU, F_NEWV, F_P, F_S, F_M, F_NEWM, F_COUNT };
enum SampleKind { SK_U = 0, SK_V, SK_M };

template <class T> struct Fluid {
  T density, h;
  int NX, NY, numCells;
  std::vector<T> U, V, newU, newV, p, S, M, newM;

  // Exported knobs (fluid.go:25-39) with New's defaults (fluid.go:59-66).
  T Confinement = 0;
  T ViscosityDiffusion = 0;
  T PressureDamping = 1;
  T TurbulenceStrength = (T)0.02f;
  T SmokeAdvection = 1;
  bool UseMultigrid = false;
  int MultigridLevels = 2;
  bool UseBFECC = false;
  // Package variable `Relaxation` (fluid.go:7-9).
  T Relaxation = (T)1.9f;
  // Additions: Simulate hard-codes 8 sweeps (fluid.go:81) and always early-exits.
  int PressureIters = 8;
  bool EarlyExit = true;
  // Addition (off in the reference): back-trace in cell units, see sampleCells.
  bool PreciseAdvection = false;

  int threads = 1;

  // Diagnostics of the last pressure solve.
  int lastSweeps = 0;
  std::vector<T> sweepMaxDiv;
  bool invalidIndex = false;  // a NaN reached a sample index (the reference panics there)

  // FP64 shadow decisions: run exactly forcedSweeps sweeps, take the turbulence gate
  // from an FP32 run of the same step.
  int forcedSweeps = -1;
  std::vector<uint8_t> turbGateIn;
  bool useTurbGateIn = false;
  std::vector<uint8_t> turbGateOut;

  // Optional taps of BFECC intermediates (the adjoint's tape): forward-pass and
  // corrected fields of velocity and smoke.
  struct Taps {
    std::vector<T>*fwdU = nullptr, *fwdV = nullptr, *corrU = nullptr, *corrV = nullptr;
    std::vector<T>*fwdM = nullptr, *corrM = nullptr;
  } taps;

  Fluid(T density_, int width, int height, T h_) : density(density_), h(h_) {
    NX = width + 2;
    NY = height + 2;
    numCells = NX * NY;
    for (auto* f : {&U, &V, &newU, &newV, &p, &S, &M, &newM}) f->assign(numCells, (T)0);
    turbGateOut.assign(numCells, 0);
  }

  std::vector<T>* field(int id) {
    switch (id) {
      case F_U: return &U;
      case F_V: return &V;
      case F_NEWU: return &newU;
      case F_NEWV: return &newV;
      case F_P: return &p;
      case F_S: return &S;
      case F_M: return &M;
      case F_NEWM: return &newM;
    }
    return nullptr;
  }

  // ---- Simulate (fluid.go:79-109) ---------------------------------------------------
  void Simulate(T dt) {
    int numIters = PressureIters;
    std::fill(p.begin(), p.end(), (T)0);
    if (ViscosityDiffusion > 0) applyViscosity(dt);
    makeIncompressible(numIters, dt);
    if (Confinement != 0) applyVorticityConfinement(dt);
    if (TurbulenceStrength > 0) addTurbulence(dt);
    handleBorders();
    if (UseBFECC) {
      advectVelocityBFECC(dt);
      advectSmokeBFECC(dt);
    } else {
      advectVelocity(dt);
      advectSmoke(dt);
    }
  }

  // fluid.go:112-142
  void applyViscosity(T dt) {
    if (ViscosityDiffusion <= 0) return;
    const int n = NY;
    T visc = PLANT(11) ? ViscosityDiffusion : ViscosityDiffusion * dt;
    newU = U;
    newV = V;
    prange(threads, 1, NX - 1, [&](int i) {
      for (int j = 1; j < NY - 1; j++) {
        int c = i * n + j;
        if (S[c] > 0) {
          T uLap = U[c - n] + U[c + n] + U[c - 1] + U[c + 1] - (T)4 * U[c];
          newU[c] = U[c] + visc * uLap;
          T vLap = V[c - n] + V[c + n] + V[c - 1] + V[c + 1] - (T)4 * V[c];
          newV[c] = V[c] + visc * vLap;
        }
      }
    });
    U = newU;
    V = newV;
  }

  // fluid.go:144-155
  void makeIncompressible(int numIters, T dt) {
    copyBorder(newU, U);
    copyBorder(newV, V);
    if (UseMultigrid && MultigridLevels > 1)
      solveMultigridVCycle(numIters, dt);
    else
      solveSingleGrid(numIters, dt);
  }

  // fluid.go:157-186 (lines 180-182 are dead: the relaxation is recomputed each sweep).
  void solveSingleGrid(int numIters, T dt) {
    T cp = density * h / dt;
    T tolerance = (T)1e-5f;
    T initialRelaxation = Relaxation;
    T minRelaxation = (T)1.2f;
    lastSweeps = 0;
    sweepMaxDiv.clear();
    for (int iter = 0; iter < numIters; iter++) {
      if (forcedSweeps >= 0 && iter >= forcedSweeps) break;
      T iterProgress = (T)(PLANT(10) ? iter + 1 : iter) / (T)numIters;
      T relax = initialRelaxation - (initialRelaxation - minRelaxation) * iterProgress;
      T maxDiv = pressureSweep(relax, cp);
      sweepMaxDiv.push_back(maxDiv);
      lastSweeps = iter + 1;
      if (forcedSweeps < 0 && EarlyExit && maxDiv < tolerance) break;
    }
  }

  // pressureJacobiIteration (fluid.go:188-234): in-place lexicographic Gauss-Seidel/SOR.
  T pressureSweep(T relaxation, T cp) {
    const int n = NY;
    T maxDiv = 0;
    for (int i = 1; i < NX - 1; i++) {
      for (int j = 1; j < NY - 1; j++) {
        int c = i * n + j;
        if (S[c] == 0) continue;
        T sx0 = S[c - n], sx1 = S[c + n], sy0 = S[c - 1], sy1 = S[c + 1];
        T s = sx0 + sx1 + sy0 + sy1;
        if (s == 0) continue;
        T div = U[c + n] - U[c] + V[c + 1] - V[c];
        T absDiv = std::fabs(div);
        if (absDiv > maxDiv) maxDiv = absDiv;
        T q = -div / s;
        q *= relaxation;
        q *= PressureDamping;
        p[c] += cp * q;
        U[c] -= sx0 * q;
        U[c + n] += sx1 * q;
        V[c] -= sy0 * q;
        V[c + 1] += sy1 * q;
      }
    }
    return maxDiv;
  }

  // fluid.go:236-289
  void handleBorders() {
    const int n = NY;
    prange(threads, 0, NX, [&](int i) {
      if (S[i * n + 0] == 0 || S[i * n + 1] == 0) {
        U[i * n + 0] = 0;
      } else if (i > 0 && i < NX - 1 && S[i * n + 2] > 0) {
        U[i * n + 0] = (T)2 * U[i * n + 1] - U[i * n + 2];
      } else {
        U[i * n + 0] = U[i * n + 1];
      }
      if (S[i * n + NY - 1] == 0 || S[i * n + NY - 2] == 0) {
        U[i * n + NY - 1] = 0;
      } else if (i > 0 && i < NX - 1 && S[i * n + NY - 3] > 0) {
        U[i * n + NY - 1] = (T)2 * U[i * n + NY - 2] - U[i * n + NY - 3];
      } else {
        U[i * n + NY - 1] = U[i * n + NY - 2];
      }
    });
    prange(threads, 0, NY, [&](int j) {
      if (S[0 * n + j] == 0 || S[1 * n + j] == 0) {
        V[0 * n + j] = 0;
      } else if (j > 0 && j < NY - 1 && S[2 * n + j] > 0) {
        V[0 * n + j] = (T)2 * V[1 * n + j] - V[2 * n + j];
      } else {
        V[0 * n + j] = V[1 * n + j];
      }
      if (S[(NX - 1) * n + j] == 0 || S[(NX - 2) * n + j] == 0) {
        V[(NX - 1) * n + j] = 0;
      } else if (j > 0 && j < NY - 1 && S[(NX - 3) * n + j] > 0) {
        V[(NX - 1) * n + j] = (T)2 * V[(NX - 2) * n + j] - V[(NX - 3) * n + j];
      } else {
        V[(NX - 1) * n + j] = V[(NX - 2) * n + j];
      }
    });
  }

  T avgU(const std::vector<T>& Uf, int i, int j) const {
    const int n = NY;
    return (Uf[i * n + j - 1] + Uf[i * n + j] + Uf[(i + 1) * n + j - 1] + Uf[(i + 1) * n + j]) * (T)0.25;
  }
  T avgV(const std::vector<T>& Vf, int i, int j) const {
    const int n = NY;
    return (Vf[(i - 1) * n + j] + Vf[i * n + j] + Vf[(i - 1) * n + j + 1] + Vf[i * n + j + 1]) * (T)0.25;
  }

  // int(math.Floor(float64(v))); a NaN or huge value would make the reference panic on
  // the slice index. Here it is flagged and mapped to 0.
  int floorIndex(T v) {
    double f = std::floor((double)v);
    if (!(f >= -2147483648.0 && f < 2147483648.0)) {
      invalidIndex = true;
      return 0;
    }
    return (int)f;
  }

  // sampleField / sampleFieldFrom (fluid.go:357-398, 1055-1091): clamped bilinear sample.
  T sample(T x, T y, const std::vector<T>& data, int kind) {
    const int n = NY;
    T h1 = (T)1 / h;
    T h2 = (T)0.5 * h;
    x = go_max(go_min(x, (T)NX * h), h);
    y = go_max(go_min(y, (T)NY * h), h);
    T dx = 0, dy = 0;
    if (kind == SK_U) dy = h2;
    if (kind == SK_V) dx = h2;
    if (kind == SK_M) {
      dx = h2;
      dy = h2;
    }
    if (PLANT(2) && kind != SK_M) std::swap(dx, dy);
    int x0 = std::min(floorIndex((x - dx) * h1), NX - 1);
    T tx = ((x - dx) - (T)x0 * h) * h1;
    int x1 = std::min(x0 + 1, NX - 1);
    int y0 = std::min(floorIndex((y - dy) * h1), NY - 1);
    T ty = ((y - dy) - (T)y0 * h) * h1;
    int y1 = std::min(y0 + 1, NY - 1);
    if (x0 < 0 || y0 < 0) {
      invalidIndex = true;
      x0 = std::max(x0, 0);
      y0 = std::max(y0, 0);
    }
    T sx = (T)1 - tx;
    T sy = (T)1 - ty;
    if (PLANT(1))
      return sx * sy * data[x0 * n + y0] + tx * sy * data[x0 * n + y1] + tx * ty * data[x1 * n + y1] +
             sx * ty * data[x1 * n + y0];
    return sx * sy * data[x0 * n + y0] + tx * sy * data[x1 * n + y0] + tx * ty * data[x1 * n + y1] +
           sx * ty * data[x0 * n + y1];
  }

  // PreciseAdvection (an addition; the reference always samples at absolute positions).
  // Every back-trace of the step lands at X = i + dX, Y = j + dY in cell units (the face
  // and sample offsets cancel), with dX = (s*u)*h1 for the trace step s = -dt or +dt. The
  // reference forms x = i*h + s*u and rounds it at ulp(N*h), i.e. to ~N*6e-8 cells; here
  // the integer part is split off exactly and only the fraction d - floor(d) is kept, so
  // the position error is that of dX alone. The clamp x in [h, N*h] becomes X in
  // [1 - o, N - o], o = 1/2 on an axis with a half-cell sample offset (y of U, x of V,
  // both of M); corners, weights and their evaluation order are sampleField's.
  // Returns whether X was inside [1 - o, N - o] (not clamped).
  bool axisCells(int base, T d, int N, bool half, int& i0, T& t) {
    if (d != d) {  // a NaN velocity: the reference flags the index and samples NaN
      invalidIndex = true;
      i0 = 0;
      t = d;
      return false;
    }
    bool inside = true;
    const int loI = half ? 0 : 1, hiI = half ? N - 1 : N;
    const T loF = half ? (T)0.5 : (T)0, hiF = half ? (T)0.5 : (T)0;
    const T fl = std::floor(d);
    if (fl < (T)-16777216) {
      i0 = loI;
      t = loF;
      inside = false;
    } else if (fl > (T)16777216) {
      i0 = hiI;
      t = hiF;
      inside = false;
    } else {
      i0 = base + (int)fl;
      t = d - fl;
      if (i0 < loI || (i0 == loI && t < loF)) {
        i0 = loI;
        t = loF;
        inside = false;
      } else if (i0 > hiI || (i0 == hiI && t > hiF)) {
        i0 = hiI;
        t = hiF;
        inside = false;
      }
    }
    if (i0 > N - 1) {  // X = N exactly: sampleField's min(x0, N-1) with tx = 1
      i0 = N - 1;
      t = (T)1;
    }
    return inside;
  }
  T sampleCells(int i, int j, T dX, T dY, const std::vector<T>& data, int kind) {
    const int n = NY;
    int x0, y0;
    T tx, ty;
    axisCells(i, dX, NX, kind != SK_U, x0, tx);
    axisCells(j, dY, NY, kind != SK_V, y0, ty);
    const int x1 = std::min(x0 + 1, NX - 1), y1 = std::min(y0 + 1, NY - 1);
    T sx = (T)1 - tx;
    T sy = (T)1 - ty;
    return sx * sy * data[x0 * n + y0] + tx * sy * data[x1 * n + y0] + tx * ty * data[x1 * n + y1] +
           sx * ty * data[x0 * n + y1];
  }
  // Displacement in cells of a trace step s (-dt forward, +dt for BFECC's backward pass):
  // s*u is the reference's dt*u with the trace's sign, exactly.
  T disp(T s, T u) const { return (s * u) * ((T)1 / h); }

  // fluid.go:291-333
  void advectVelocity(T dt) {
    copyBorder(newU, U);
    copyBorder(newV, V);
    const int n = NY;
    T h2 = h / (T)2;
    prange(threads, 1, NX, [&](int i) {
      for (int j = 1; j < NY; j++) {
        int c = i * n + j;
        if (S[c] != 0 && S[c - n] != 0 && j < NY - 1) {
          T x = (T)i * h;
          T y = (T)j * h + h2;
          T u = U[c];
          T v = avgV(V, i, j);
          x = x - dt * u;
          y = y - dt * v;
          newU[c] = PreciseAdvection ? sampleCells(i, j, disp(-dt, u), disp(-dt, v), U, SK_U) : sample(x, y, U, SK_U);
        }
        if (S[c] != 0 && S[c - 1] != 0 && i < NX - 1) {
          T x = (T)i * h + h2;
          T y = (T)j * h;
          T u = avgU(U, i, j);
          T v = V[c];
          x = x - dt * u;
          y = y - dt * v;
          newV[c] = PreciseAdvection ? sampleCells(i, j, disp(-dt, u), disp(-dt, v), V, SK_V) : sample(x, y, V, SK_V);
        }
      }
    });
    U = newU;
    V = newV;
  }

  // fluid.go:400-434
  void advectSmoke(T dt) {
    copyBorder(newM, M);
    const int n = NY;
    T h2 = (T)0.5 * h;
    prange(threads, 1, NX - 1, [&](int i) {
      for (int j = 1; j < NY - 1; j++) {
        int c = i * n + j;
        if (S[c] != 0) {
          T u = (U[c] + U[c + n]) * (T)0.5 * SmokeAdvection;
          T v = (V[c] + V[c + 1]) * (T)0.5 * SmokeAdvection;
          T x = (T)i * h + h2 - dt * u;
          T y = (T)j * h + h2 - dt * v;
          T val = PreciseAdvection ? sampleCells(i, j, disp(-dt, u), disp(-dt, v), M, SK_M) : sample(x, y, M, SK_M);
          if (ViscosityDiffusion > 0) {
            T smokeDiffusion = PLANT(12) ? ViscosityDiffusion * dt : ViscosityDiffusion * (T)0.3f * dt;
            T nb = M[c - n] + M[c + n] + M[c - 1] + M[c + 1] - (T)4 * M[c];
            val += smokeDiffusion * nb;
          }
          newM[c] = go_max(val, (T)0);
        }
      }
    });
    M = newM;
  }

  // fluid.go:436-446
  void copyBorder(std::vector<T>& dst, const std::vector<T>& src) {
    const int n = NY;
    for (int i = 0; i < NX; i++) {
      dst[i * n + 0] = src[i * n + 0];
      dst[i * n + NY - 1] = src[i * n + NY - 1];
    }
    for (int j = 0; j < NY; j++) {
      dst[0 * n + j] = src[0 * n + j];
      dst[(NX - 1) * n + j] = src[(NX - 1) * n + j];
    }
  }

  // fluid.go:449-493 (serial in the reference)
  void applyVorticityConfinement(T dt) {
    const int n = NY;
    std::vector<T> curl(numCells, (T)0);
    for (int i = 1; i < NX - 1; i++)
      for (int j = 1; j < NY - 1; j++) {
        int c = i * n + j;
        if (S[c] == 0) continue;
        T dvdx = (V[c + n] - V[c - n]) * (T)0.5 / h;
        T dudy = (U[c + 1] - U[c - 1]) * (T)0.5 / h;
        curl[c] = dvdx - dudy;
      }
    T eps = (T)1e-5f;
    for (int i = 1; i < NX - 1; i++)
      for (int j = 1; j < NY - 1; j++) {
        int c = i * n + j;
        if (S[c] == 0) continue;
        T gx = (std::fabs(curl[c + n]) - std::fabs(curl[c - n])) * (T)0.5 / h;
        T gy = (std::fabs(curl[c + 1]) - std::fabs(curl[c - 1])) * (T)0.5 / h;
        T mag = std::sqrt(gx * gx + gy * gy) + eps;
        gx /= mag;
        gy /= mag;
        T vort = curl[c];
        T localVel = std::sqrt(U[c] * U[c] + V[c] * V[c]);
        T adaptiveStrength = PLANT(9) ? Confinement : Confinement * ((T)1 + localVel * (T)0.1f);
        if (PLANT(8)) {
          U[c] -= adaptiveStrength * gy * vort * dt;
          V[c] += adaptiveStrength * gx * vort * dt;
          continue;
        }
        U[c] += adaptiveStrength * gy * vort * dt;
        V[c] -= adaptiveStrength * gx * vort * dt;
      }
  }

  // fluid.go:496-526
  void addTurbulence(T dt) {
    if (TurbulenceStrength <= 0) return;
    const int n = NY;
    T turbStrength = TurbulenceStrength * dt;
    prange(threads, 1, NX - 1, [&](int i) {
      for (int j = 1; j < NY - 1; j++) {
        int c = i * n + j;
        if (S[c] > 0) {
          T noiseU = (T)turb_noise(i * 137 + j * 241) * turbStrength;
          T noiseV = (T)turb_noise(i * 157 + j * 263) * turbStrength;
          T localVel = PLANT(13) ? std::fabs(U[c]) + std::fabs(V[c]) : std::sqrt(U[c] * U[c] + V[c] * V[c]);
          bool gate = localVel > (T)0.1f;
          turbGateOut[c] = gate ? 1 : 0;
          if (useTurbGateIn) gate = turbGateIn[c] != 0;
          if (gate) {
            T factor = go_min(PLANT(14) ? localVel : localVel * (T)0.5, (T)1);
            U[c] += noiseU * factor;
            V[c] += noiseV * factor;
          }
        }
      }
    });
  }

  // fluid.go:529-557
  T GetAdaptiveTimeStep(T basedt) {
    const int n = NY;
    T maxVel = 0;
    for (int i = 1; i < NX - 1; i++)
      for (int j = 1; j < NY - 1; j++) {
        int c = i * n + j;
        if (S[c] > 0) {
          T vel = PLANT(21) ? std::sqrt(U[c] * U[c] + V[c] * V[c]) : std::fabs(U[c]) + std::fabs(V[c]);
          if (vel > maxVel) maxVel = vel;
        }
      }
    if (maxVel == 0) return basedt;
    T cfl = (T)0.8f;
    T adt = cfl * h / maxVel;
    adt = go_max(go_min(adt, basedt * (T)(PLANT(20) ? 1 : 2)), basedt * (T)0.1f);
    return adt;
  }

  // ---- multigrid (fluid.go:560-758, 1123-1149) ------------------------------------
  void solveMultigridVCycle(int numIters, T dt) {
    T cp = density * h / dt;
    T tolerance = (T)1e-5f;
    lastSweeps = 0;
    sweepMaxDiv.clear();
    for (int iter = 0; iter < numIters; iter++) {
      T maxDiv = 0;
      for (int s = 0; s < 3; s++) {
        maxDiv = pressureSweep((T)1.5, cp);
        lastSweeps++;
      }
      sweepMaxDiv.push_back(maxDiv);
      if (maxDiv < tolerance) break;
      std::vector<T> residual = computePressureResidual();
      std::vector<T> coarseRHS = restrictResidual(residual);
      std::vector<T> coarse = solveCoarseGrid(coarseRHS);
      std::vector<T> corr = prolongateCorrection(coarse);
      applyPressureCorrection(corr, cp);
      for (int s = 0; s < 3; s++) {
        pressureSweep((T)1.2f, cp);
        lastSweeps++;
      }
    }
  }

  std::vector<T> computePressureResidual() {
    const int n = NY;
    std::vector<T> r(numCells, (T)0);
    for (int i = 1; i < NX - 1; i++)
      for (int j = 1; j < NY - 1; j++) {
        int c = i * n + j;
        if (S[c] == 0) continue;
        T div = U[c + n] - U[c] + V[c + 1] - V[c];
        T sx0 = S[c - n], sx1 = S[c + n], sy0 = S[c - 1], sy1 = S[c + 1];
        T lap = sx0 * (p[c - n] - p[c]) + sx1 * (p[c + n] - p[c]) + sy0 * (p[c - 1] - p[c]) +
                sy1 * (p[c + 1] - p[c]);
        r[c] = PLANT(15) ? -div + lap : -div - lap;
      }
    return r;
  }

  std::vector<T> restrictResidual(const std::vector<T>& fr) {
    int cNX = (NX + 1) / 2, cNY = (NY + 1) / 2;
    std::vector<T> out((size_t)cNX * cNY, (T)0);
    for (int i = 1; i < cNX - 1; i++)
      for (int j = 1; j < cNY - 1; j++) {
        int fi = i * 2, fj = j * 2;
        if (fi < NX - 1 && fj < NY - 1) {
          T center = fr[fi * NY + fj] * (T)0.25;
          T nb = (fr[(fi - 1) * NY + fj] + fr[(fi + 1) * NY + fj] + fr[fi * NY + fj - 1] + fr[fi * NY + fj + 1]) *
                 (T)0.125;
          T co = (fr[(fi - 1) * NY + fj - 1] + fr[(fi + 1) * NY + fj - 1] + fr[(fi - 1) * NY + fj + 1] +
                  fr[(fi + 1) * NY + fj + 1]) *
                 (T)(PLANT(16) ? 0.125 : 0.0625);
          out[i * cNY + j] = center + nb + co;
        }
      }
    return out;
  }

  std::vector<T> solveCoarseGrid(const std::vector<T>& rhs) {
    int cNX = (NX + 1) / 2, cNY = (NY + 1) / 2;
    std::vector<T> P((size_t)cNX * cNY, (T)0), CS((size_t)cNX * cNY, (T)0);
    for (int i = 0; i < cNX; i++)
      for (int j = 0; j < cNY; j++) {
        int fi = i * 2, fj = j * 2;
        CS[i * cNY + j] = (fi < NX && fj < NY) ? S[fi * NY + fj] : (T)0;
      }
    T relax = PLANT(17) ? (T)1 : (T)1.6f;
    for (int it = 0; it < 40; it++)
      for (int i = 1; i < cNX - 1; i++)
        for (int j = 1; j < cNY - 1; j++) {
          int c = i * cNY + j;
          if (CS[c] == 0) continue;
          T sx0 = CS[c - cNY], sx1 = CS[c + cNY], sy0 = CS[c - 1], sy1 = CS[c + 1];
          T s = sx0 + sx1 + sy0 + sy1;
          if (s == 0) continue;
          T lap = sx0 * (P[c - cNY] - P[c]) + sx1 * (P[c + cNY] - P[c]) + sy0 * (P[c - 1] - P[c]) +
                  sy1 * (P[c + 1] - P[c]);
          T res = rhs[c] - lap;
          T corr = -res / s * relax;
          P[c] += corr;
        }
    return P;
  }

  std::vector<T> prolongateCorrection(const std::vector<T>& cc) {
    std::vector<T> corr(numCells, (T)0);
    int cNY = (NY + 1) / 2, cNX = (NX + 1) / 2;
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++) {
        int ci = i / 2, cj = j / 2;
        if (ci >= cNX - 1 || cj >= cNY - 1) continue;
        T fi = (T)(i % 2) * (T)0.5, fj = (T)(j % 2) * (T)0.5;
        if (PLANT(18)) std::swap(fi, fj);
        corr[i * NY + j] = ((T)1 - fi) * ((T)1 - fj) * cc[ci * cNY + cj] + fi * ((T)1 - fj) * cc[(ci + 1) * cNY + cj] +
                           ((T)1 - fi) * fj * cc[ci * cNY + cj + 1] + fi * fj * cc[(ci + 1) * cNY + cj + 1];
      }
    return corr;
  }

  void applyPressureCorrection(const std::vector<T>& corr, T cp) {
    const int n = NY;
    for (int i = 1; i < NX - 1; i++)
      for (int j = 1; j < NY - 1; j++) {
        int c = i * n + j;
        if (S[c] == 0) continue;
        p[c] += corr[c] * cp;
        T sx0 = S[c - n], sx1 = S[c + n], sy0 = S[c - 1], sy1 = S[c + 1];
        T pc = corr[c];
        if (PLANT(19)) {
          U[c] += sx0 * pc;
          U[c + n] -= sx1 * pc;
        } else {
          U[c] -= sx0 * pc;
          U[c + n] += sx1 * pc;
        }
        V[c] -= sy0 * pc;
        V[c + 1] += sy1 * pc;
      }
  }

  // ---- BFECC (fluid.go:911-1051, 1094-1120) ----------------------------------------
  T clampToNeighbors(T val, const std::vector<T>& src, int i, int j) const {
    if (PLANT(6)) return val;
    const int n = NY;
    T lo = src[i * n + j], hi = src[i * n + j];
    for (int di = -1; di <= 1; di++)
      for (int dj = -1; dj <= 1; dj++) {
        int ni = i + di, nj = j + dj;
        if (ni >= 0 && ni < NX && nj >= 0 && nj < NY) {
          T v = src[ni * n + nj];
          if (v < lo) lo = v;
          if (v > hi) hi = v;
        }
      }
    if (val < lo) return lo;
    if (val > hi) return hi;
    return val;
  }

  void advectVelocityBFECC(T dt) {
    const int n = NY;
    T h2 = h / (T)2;
    std::vector<T> origU = U, origV = V;
    advectVelocity(dt);
    std::vector<T> fwdU = U, fwdV = V;
    if (taps.fwdU) *taps.fwdU = fwdU;
    if (taps.fwdV) *taps.fwdV = fwdV;
    U = origU;
    V = origV;
    std::vector<T> bwdU(numCells, (T)0), bwdV(numCells, (T)0);
    copyBorder(bwdU, fwdU);
    copyBorder(bwdV, fwdV);
    prange(threads, 1, NX, [&](int i) {
      for (int j = 1; j < NY; j++) {
        int c = i * n + j;
        if (S[c] != 0 && S[c - n] != 0 && j < NY - 1) {
          T x = (T)i * h;
          T y = (T)j * h + h2;
          T u = PLANT(4) ? fwdU[c] : U[c];
          T v = PLANT(4) ? avgV(fwdV, i, j) : avgV(V, i, j);
          const T st = PLANT(3) ? -dt : dt;
          x = x + st * u;
          y = y + st * v;
          bwdU[c] = PreciseAdvection ? sampleCells(i, j, disp(dt, u), disp(dt, v), fwdU, SK_U) : sample(x, y, fwdU, SK_U);
        }
        if (S[c] != 0 && S[c - 1] != 0 && i < NX - 1) {
          T x = (T)i * h + h2;
          T y = (T)j * h;
          T u = PLANT(4) ? avgU(fwdU, i, j) : avgU(U, i, j);
          T v = PLANT(4) ? fwdV[c] : V[c];
          const T st = PLANT(3) ? -dt : dt;
          x = x + st * u;
          y = y + st * v;
          bwdV[c] = PreciseAdvection ? sampleCells(i, j, disp(dt, u), disp(dt, v), fwdV, SK_V) : sample(x, y, fwdV, SK_V);
        }
      }
    });
    std::vector<T> corrU = origU, corrV = origV;
    for (int i = 1; i < NX - 1; i++)
      for (int j = 1; j < NY - 1; j++) {
        int c = i * n + j;
        const T half = PLANT(7) ? (T)1 : (T)0.5;
        corrU[c] = origU[c] - (bwdU[c] - origU[c]) * half;
        corrV[c] = origV[c] - (bwdV[c] - origV[c]) * half;
      }
    for (int i = 1; i < NX - 1; i++)
      for (int j = 1; j < NY - 1; j++) {
        int c = i * n + j;
        corrU[c] = clampToNeighbors(corrU[c], origU, i, j);
        corrV[c] = clampToNeighbors(corrV[c], origV, i, j);
      }
    if (taps.corrU) *taps.corrU = corrU;
    if (taps.corrV) *taps.corrV = corrV;
    U = corrU;
    V = corrV;
    advectVelocity(dt);
  }

  void advectSmokeBFECC(T dt) {
    const int n = NY;
    T h2 = (T)0.5 * h;
    std::vector<T> origM = M;
    advectSmoke(dt);
    std::vector<T> fwdM = M;
    if (taps.fwdM) *taps.fwdM = fwdM;
    std::vector<T> bwdM(numCells, (T)0);
    copyBorder(bwdM, fwdM);
    prange(threads, 1, NX - 1, [&](int i) {
      for (int j = 1; j < NY - 1; j++) {
        int c = i * n + j;
        if (S[c] != 0) {
          T u = (U[c] + U[c + n]) * (T)0.5 * SmokeAdvection;
          T v = (V[c] + V[c + 1]) * (T)0.5 * SmokeAdvection;
          const T st = PLANT(5) ? -dt : dt;
          T x = (T)i * h + h2 + st * u;
          T y = (T)j * h + h2 + st * v;
          bwdM[c] = PreciseAdvection ? sampleCells(i, j, disp(dt, u), disp(dt, v), fwdM, SK_M) : sample(x, y, fwdM, SK_M);
        }
      }
    });
    std::vector<T> corrM = origM;
    for (int i = 1; i < NX - 1; i++)
      for (int j = 1; j < NY - 1; j++) {
        int c = i * n + j;
        corrM[c] = origM[c] - (bwdM[c] - origM[c]) * (PLANT(7) ? (T)1 : (T)0.5);
      }
    for (int i = 1; i < NX - 1; i++)
      for (int j = 1; j < NY - 1; j++) {
        int c = i * n + j;
        corrM[c] = clampToNeighbors(corrM[c], origM, i, j);
        if (corrM[c] < 0) corrM[c] = 0;
      }
    if (taps.corrM) *taps.corrM = corrM;
    M = corrM;
    advectSmoke(dt);
  }

  // ---- mask / injection API (walls.go, fluid.go:761-796, 894-907) ----------------------
  bool inGrid(int i, int j) const { return i >= 0 && i < NX && j >= 0 && j < NY; }

  int SetSolid(int i, int j, bool value) {
    if (i < 0 || i >= NX) return 1;
    if (j < 0 || j >= NY) return 2;
    const int n = NY;
    int c = i * n + j;
    S[c] = value ? (T)0 : (T)1;
    if (value) {
      U[c] = 0;
      if (i + 1 < NX) U[c + n] = 0;
      V[c] = 0;
      if (j + 1 < NY) V[c + 1] = 0;
      newU[c] = 0;
      if (i + 1 < NX) newU[c + n] = 0;
      newV[c] = 0;
      if (j + 1 < NY) newV[c + 1] = 0;
    }
    return 0;
  }
  int IsSolid(int i, int j, bool* out) const {
    if (i < 0 || i >= NX) return 1;
    if (j < 0 || j >= NY) return 2;
    *out = S[i * NY + j] == 0;
    return 0;
  }
  int SetVelocity(int i, int j, T u, T v) {
    if (i < 0 || i >= NX) return 1;
    if (j < 0 || j >= NY) return 2;
    U[i * NY + j] = u;
    V[i * NY + j] = v;
    return 0;
  }
  int AddSmoke(int i, int j, T s) {
    if (i < 0 || i >= NX) return 1;
    if (j < 0 || j >= NY) return 2;
    M[i * NY + j] += s;
    return 0;
  }
  void Reset() {
    for (auto* f : {&U, &V, &newU, &newV, &p, &M, &newM}) std::fill(f->begin(), f->end(), (T)0);
  }
  void ApplyForce(int i, int j, T fx, T fy) {
    if (i < 1 || i >= NX - 1 || j < 1 || j >= NY - 1) return;
    int c = i * NY + j;
    if (S[c] == 0) return;
    U[c] += fx;
    V[c] += fy;
  }
  void ApplyForceRadius(int cx, int cy, T fx, T fy, int radius) {
    if (radius <= 0) {
      ApplyForce(cx, cy, fx, fy);
      return;
    }
    T r2 = (T)(radius * radius);
    for (int i = cx - radius; i <= cx + radius; i++)
      for (int j = cy - radius; j <= cy + radius; j++) {
        if (i < 1 || i >= NX - 1 || j < 1 || j >= NY - 1) continue;
        T dx = (T)(i - cx), dy = (T)(j - cy);
        T d2 = dx * dx + dy * dy;
        if (d2 > r2) continue;
        T w = (T)std::exp((double)((T)-3 * d2 / r2));
        ApplyForce(i, j, fx * w, fy * w);
      }
  }
  // Gaussian impulse of the ensemble workload (no reference counterpart): ApplyForce
  // (fluid.go:761-772) weighted by exp(-d^2/(2 sigma^2)) over cells with d <= 3 sigma,
  // visited in ApplyForceRadius's order (fluid.go:780-795).
  void ApplyForceGaussian(int cx, int cy, T fx, T fy, float sigma) {
    const float two_s2 = 2.0f * sigma * sigma, cut2 = 9.0f * sigma * sigma;
    const int R = (int)std::ceil(3.0f * sigma);
    for (int i = cx - R; i <= cx + R; i++)
      for (int j = cy - R; j <= cy + R; j++) {
        if (i < 1 || i >= NX - 1 || j < 1 || j >= NY - 1) continue;
        float dx = (float)(i - cx), dy = (float)(j - cy);
        float d2 = dx * dx + dy * dy;
        if (d2 > cut2) continue;
        T w = (T)std::exp(-(double)(d2 / two_s2));
        ApplyForce(i, j, fx * w, fy * w);
      }
  }
  void SetCircularObstacle(int cx, int cy, int radius) {
    for (int i = cx - radius; i <= cx + radius; i++)
      for (int j = cy - radius; j <= cy + radius; j++) {
        if (i < 0 || i >= NX || j < 0 || j >= NY) continue;
        T dx = (T)(i - cx), dy = (T)(j - cy);
        if (dx * dx + dy * dy <= (T)(radius * radius)) SetSolid(i, j, true);
      }
  }

  // ---- views (fluid.go:799-891, pressure.go, smoke.go) -------------------------------
  void SampleVelocity(T x, T y, T* u, T* v) {
    *u = sample(x, y, U, SK_U);
    *v = sample(x, y, V, SK_V);
  }
  void Vorticity(T* out, T* mn, T* mx) const {
    const int n = NY;
    T minV = (T)std::numeric_limits<float>::max(), maxV = -(T)std::numeric_limits<float>::max();
    for (int c = 0; c < numCells; c++) out[c] = 0;
    for (int i = 1; i < NX - 1; i++)
      for (int j = 1; j < NY - 1; j++) {
        int c = i * n + j;
        if (S[c] == 0) continue;
        T dvdx = (V[c + n] - V[c - n]) * (T)0.5 / h;
        T dudy = (U[c + 1] - U[c - 1]) * (T)0.5 / h;
        T curl = dvdx - dudy;
        out[c] = curl;
        if (curl < minV) minV = curl;
        if (curl > maxV) maxV = curl;
      }
    *mn = minV;
    *mx = maxV;
  }
  void VelocityMagnitude(T* out, T* mn, T* mx) const {
    const int n = NY;
    T minV = (T)std::numeric_limits<float>::max(), maxV = -(T)std::numeric_limits<float>::max();
    for (int c = 0; c < numCells; c++) out[c] = 0;
    for (int i = 1; i < NX - 1; i++)
      for (int j = 1; j < NY - 1; j++) {
        int c = i * n + j;
        if (S[c] == 0) continue;
        T u = (U[c] + U[c + n]) * (T)0.5;
        T v = (V[c] + V[c + 1]) * (T)0.5;
        T mag = std::sqrt(u * u + v * v);
        out[c] = mag;
        if (mag < minV) minV = mag;
        if (mag > maxV) maxV = mag;
      }
    *mn = minV;
    *mx = maxV;
  }
  T MaxDivergence() const {
    const int n = NY;
    T maxDiv = 0;
    for (int i = 1; i < NX - 1; i++)
      for (int j = 1; j < NY - 1; j++) {
        int c = i * n + j;
        if (S[c] == 0) continue;
        T div = U[c + n] - U[c] + V[c + 1] - V[c];
        T a = std::fabs(div);
        if (a > maxDiv) maxDiv = a;
      }
    return maxDiv;
  }
  void FieldMinMax(int id, T* mn, T* mx) {
    std::vector<T>& f = *field(id);
    T minV = (T)std::numeric_limits<float>::max(), maxV = -(T)std::numeric_limits<float>::max();
    for (int c = 0; c < numCells; c++) {
      if (f[c] < minV) minV = f[c];
      if (f[c] > maxV) maxV = f[c];
    }
    *mn = minV;
    *mx = maxV;
  }
};

// ---- scene layer lifted from main/main.go ------------------------------------------
struct Source {
  int I, J;
  double U, V;
};

template <class T> struct Scene {
  Fluid<T>* f;
  bool wallTop = true, wallBottom = true, wallLeft = true, wallRight = false;
  bool jet = false;
  int jetX = 1, jetLo = 0, jetHi = 0;
  T jetU = 4, jetSmoke = 1;
  std::vector<Source> sources, sinks;
  T sourceSmoke = (T)0.5;

  // main.go:840-849 (bottom is j=0, top is j=NY-1)
  void applyWallSettings() {
    if (PLANT(26))
      for (int j = 0; j < f->NY; j++) {
        f->SetSolid(0, j, wallLeft);
        f->SetSolid(f->NX - 1, j, wallRight);
      }
    for (int i = 0; i < f->NX; i++) {
      f->SetSolid(i, 0, wallBottom);
      f->SetSolid(i, f->NY - 1, wallTop);
    }
    if (!PLANT(26))
      for (int j = 0; j < f->NY; j++) {
        f->SetSolid(0, j, wallLeft);
        f->SetSolid(f->NX - 1, j, wallRight);
      }
  }
  // main.go:851-858 (also the cell initialisation of NewGame, main.go:222-227)
  void clearWalls() {
    for (int i = 0; i < f->NX; i++)
      for (int j = 0; j < f->NY; j++) f->SetSolid(i, j, false);
    applyWallSettings();
  }
  // main.go:747-793
  void resetToPreset(int preset, T lidU) {
    f->Reset();
    sources.clear();
    sinks.clear();
    for (int i = 0; i < f->NX; i++)
      for (int j = 0; j < f->NY; j++) f->SetSolid(i, j, false);
    if (preset == 0) {  // jet
      wallTop = wallBottom = wallLeft = true;
      wallRight = false;
    } else if (preset == 1) {  // cavity
      wallTop = wallBottom = wallLeft = wallRight = true;
      jet = false;
      for (int i = 2; i < f->NX - 2; i++) sources.push_back(Source{i, f->NY - (PLANT(23) ? 3 : 2), (double)lidU, 0.0});
    } else {  // karman
      wallTop = wallBottom = wallLeft = true;
      wallRight = false;
      f->SetCircularObstacle(f->NX / (PLANT(25) ? 2 : 4), f->NY / 2, 8);
    }
    applyWallSettings();
  }
  // main.go:474-486
  void applySources() {
    for (auto& s : sources) {
      bool solid = true;
      if (f->inGrid(s.I, s.J) && f->IsSolid(s.I, s.J, &solid) == 0 && (!solid || PLANT(24))) {
        f->SetVelocity(s.I, s.J, (T)(float)s.U, (T)(float)s.V);
        f->AddSmoke(s.I, s.J, sourceSmoke);
      }
    }
    for (auto& s : sinks) {
      bool solid = true;
      if (f->inGrid(s.I, s.J) && f->IsSolid(s.I, s.J, &solid) == 0 && !solid)
        f->SetVelocity(s.I, s.J, (T)(float)s.U, (T)(float)s.V);
    }
  }
  // main.go:231-249 (Update: sources, jet, Simulate)
  void update(T dt) {
    applySources();
    if (jet)
      for (int j = jetLo; j < jetHi; j++) {
        f->SetVelocity(jetX, j, jetU, 0);
        f->AddSmoke(jetX, j, jetSmoke);
      }
    f->Simulate(dt);
  }
  // main.go:809-838 (Bresenham stroke of SetSolid)
  void drawLine(int x0, int y0, int x1, int y1, bool value) {
    int dx = std::abs(x1 - x0), dy = -std::abs(y1 - y0);
    int sx = x0 < x1 ? 1 : -1, sy = y0 < y1 ? 1 : -1;
    int err = dx + dy;
    for (;;) {
      if (f->inGrid(x0, y0)) f->SetSolid(x0, y0, value);
      if (x0 == x1 && y0 == y1) break;
      int e2 = 2 * err;
      if (PLANT(22) ? e2 > dy : e2 >= dy) {
        err += dy;
        x0 += sx;
      }
      if (e2 <= dx) {
        err += dx;
        y0 += sy;
      }
    }
  }
  // ---- particles and rendering (main.go:512-546, 549-652; colors.go:10-84) -------------
  // Game.advectParticles: age, drop if too old, RK2 midpoint with SampleVelocity, drop if the
  // cell int(x/h), int(y/h) is outside the grid or solid. keep[k] = 1 for survivors.
  void advectParticles(T dt, int n, T* x, T* y, T* age, const T* maxAge, uint8_t* keep) {
    const T h = f->h;
    for (int k = 0; k < n; k++) {
      age[k] += dt;
      keep[k] = 0;
      if (age[k] > maxAge[k]) continue;
      T u1, v1, u2, v2;
      f->SampleVelocity(x[k], y[k], &u1, &v1);
      T midX = x[k] + (T)0.5 * dt * u1;
      T midY = y[k] + (T)0.5 * dt * v1;
      f->SampleVelocity(midX, midY, &u2, &v2);
      x[k] += dt * u2;
      y[k] += dt * v2;
      T qx = x[k] / h, qy = y[k] / h;
      if (!(qx > (T)-2147483648.0 && qx < (T)2147483648.0 && qy > (T)-2147483648.0 && qy < (T)2147483648.0))
        continue;
      int fi = (int)qx, fj = (int)qy;  // Go int() truncates toward zero
      if (fi < 0 || fi >= f->NX || fj < 0 || fj >= f->NY) continue;
      if (f->S[fi * f->NY + fj] == 0) continue;
      keep[k] = 1;
    }
  }
  static uint8_t u8(float v) {  // Go uint8(float32): truncation toward zero, modulo 256
    if (!(v > -2147483648.0f && v < 2147483648.0f)) return 0;
    return (uint8_t)(int64_t)v;
  }
  static void sciColor(float val, float mn, float mx, uint8_t* px) {
    val = go_min(go_max(val, mn), mx - 0.0001f);
    float d = mx - mn;
    if (d <= 0)
      val = 0.5f;
    else
      val = (val - mn) / d;
    const float m = 0.25f;
    float num = (float)std::floor((double)(val / m));
    float s = (val - num * m) / m;
    float r = 0, g = 0, b = 0;
    if (num == 0) {
      g = s;
      b = 1;
    } else if (num == 1) {
      g = 1;
      b = 1 - s;
    } else if (num == 2) {
      r = s;
      g = 1;
    } else if (num == 3) {
      r = 1;
      g = 1 - s;
    }
    px[0] = u8(255 * r);
    px[1] = u8(255 * g);
    px[2] = u8(255 * b);
    px[3] = 0xff;
  }
  static void divergingColor(float val, float mn, float mx, uint8_t* px) {
    float absMax = (float)std::fmax(std::fabs((double)mn), std::fabs((double)mx));
    if (std::isnan((double)mn) || std::isnan((double)mx)) absMax = std::numeric_limits<float>::quiet_NaN();
    if (absMax < 1e-8f) {
      px[0] = px[1] = px[2] = px[3] = 255;
      return;
    }
    float t = val / absMax;
    if (t > 1) t = 1;
    if (t < -1) t = -1;
    float r, g, b;
    if (t >= 0) {
      r = 1.0f;
      g = 1.0f - t;
      b = 1.0f - t;
    } else {
      float a = -t;
      r = 1.0f - a;
      g = 1.0f - a;
      b = 1.0f;
    }
    px[0] = u8(255 * r);
    px[1] = u8(255 * g);
    px[2] = u8(255 * b);
    px[3] = 0xff;
  }
  // Game.Draw's image (main.go:549-612): view 0 smoke, 1 pressure, 2 velocity magnitude (sci
  // colormap), 3 vorticity (diverging); image row j shows fluid row NumY-1-j; then solid cells
  // black and, with `overlay`, source pixels green and sink pixels red.
  void render(int view, bool overlay, uint8_t* pix) {
    const int NX = f->NX, NY = f->NY, stride = 4 * NX, len = 4 * NX * NY;
    std::vector<T> vals(f->numCells);
    T mn, mx;
    if (view == 0 || view == 1) {
      const std::vector<T>& src = view == 0 ? f->M : f->p;
      vals = src;
      f->FieldMinMax(view == 0 ? F_M : F_P, &mn, &mx);
    } else if (view == 2) {
      f->VelocityMagnitude(vals.data(), &mn, &mx);
    } else {
      f->Vorticity(vals.data(), &mn, &mx);
    }
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++) {
        float v = (float)vals[i * NY + (NY - j - 1)];
        uint8_t* px = pix + 4 * i + j * stride;
        if (view == 3)
          divergingColor(v, (float)mn, (float)mx, px);
        else
          sciColor(v, (float)mn, (float)mx, px);
      }
    for (int i = 0; i < NX; i++)
      for (int j = 0; j < NY; j++)
        if (f->S[i * NY + (NY - j - 1)] == 0) {
          uint8_t* px = pix + 4 * i + j * stride;
          px[0] = px[1] = px[2] = 0;
          px[3] = 0xff;
        }
    if (!overlay) return;
    auto mark = [&](const Source& s, uint8_t r, uint8_t g) {
      int idx = 4 * s.I + (NY - s.J - 1) * stride;
      if (idx >= 0 && idx + 3 < len) {
        pix[idx] = r;
        pix[idx + 1] = g;
        pix[idx + 2] = 0;
        pix[idx + 3] = 255;
      }
    };
    for (auto& s : sources) mark(s, 0, 255);
    for (auto& s : sinks) mark(s, 255, 0);
  }

  // main.go:453-467 (smoke brush: AddSmoke(strength*0.3) over a disc)
  void smokeBrush(int ci, int cj, int r, T strength) {
    for (int i = ci - r; i <= ci + r; i++)
      for (int j = cj - r; j <= cj + r; j++) {
        if (!f->inGrid(i, j)) continue;
        T dx = (T)(i - ci), dy = (T)(j - cj);
        if (dx * dx + dy * dy > (T)(r * r)) continue;
        f->AddSmoke(i, j, strength * (T)0.3f);
      }
  }
};

#include "fluid_adjoint.inc"

}  // namespace orc

// ===========================================================================================
// C API for the ctypes wrapper (oracle/pyoracle.py). Handles are typed by precision.
// ===========================================================================================
using orc::Fluid;
using orc::Scene;

struct OrcHandle {
  int prec;
  Fluid<float>* f32 = nullptr;
  Fluid<double>* f64 = nullptr;
  Scene<float> s32;
  Scene<double> s64;
};

#define DISPATCH(H, ...)         \
  do {                           \
    if ((H)->prec == 32) {       \
      auto* F = (H)->f32;        \
      auto& SC = (H)->s32;       \
      (void)F;                   \
      (void)SC;                  \
      __VA_ARGS__;               \
    } else {                     \
      auto* F = (H)->f64;        \
      auto& SC = (H)->s64;       \
      (void)F;                   \
      (void)SC;                  \
      __VA_ARGS__;               \
    }                            \
  } while (0)

extern "C" {

double orc_go_sin(double x) { return orc::go_sin(x); }
#ifdef ORC_PLANTABLE
void orc_set_plant(int k) { orc::g_plant = k; }
#endif
// Turbulence noise for every integer seed k in [0, n): the table checked against libm.
void orc_noise_table(int n, float* out) {
  for (int k = 0; k < n; k++) out[k] = orc::turb_noise(k);
}
void orc_go_sin_many(int n, const double* x, double* out) {
  for (int k = 0; k < n; k++) out[k] = orc::go_sin(x[k]);
}
float orc_turb_noise(int k) { return orc::turb_noise(k); }

void* orc_new(int prec, float density, int width, int height, float h) {
  if (width < 1 || height < 1 || (prec != 32 && prec != 64)) return nullptr;
  OrcHandle* H = new OrcHandle();
  H->prec = prec;
  if (prec == 32) {
    H->f32 = new Fluid<float>(density, width, height, h);
    H->s32.f = H->f32;
  } else {
    H->f64 = new Fluid<double>((double)density, width, height, (double)h);
    H->s64.f = H->f64;
  }
  return H;
}
void orc_free(void* h) {
  OrcHandle* H = (OrcHandle*)h;
  if (!H) return;
  delete H->f32;
  delete H->f64;
  delete H;
}
int orc_nx(void* h) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, return F->NX);
}
int orc_ny(void* h) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, return F->NY);
}
// Field I/O: `buf` holds float for prec 32 and double for prec 64.
int orc_set_field(void* h, int id, const void* buf) {
  OrcHandle* H = (OrcHandle*)h;
  if (id < 0 || id >= orc::F_COUNT) return 1;
  DISPATCH(H, {
    auto* fld = F->field(id);
    std::memcpy(fld->data(), buf, sizeof((*fld)[0]) * fld->size());
  });
  return 0;
}
int orc_get_field(void* h, int id, void* buf) {
  OrcHandle* H = (OrcHandle*)h;
  if (id < 0 || id >= orc::F_COUNT) return 1;
  DISPATCH(H, {
    auto* fld = F->field(id);
    std::memcpy(buf, fld->data(), sizeof((*fld)[0]) * fld->size());
  });
  return 0;
}
// params: confinement, viscosity, damping, turbulence, smokeAdvection, relaxation
void orc_set_params(void* h, const double* fp, const int* ip) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, {
    using T = std::remove_reference_t<decltype(F->h)>;
    F->Confinement = (T)(float)fp[0];
    F->ViscosityDiffusion = (T)(float)fp[1];
    F->PressureDamping = (T)(float)fp[2];
    F->TurbulenceStrength = (T)(float)fp[3];
    F->SmokeAdvection = (T)(float)fp[4];
    F->Relaxation = (T)(float)fp[5];
    F->UseMultigrid = ip[0] != 0;
    F->MultigridLevels = ip[1];
    F->UseBFECC = ip[2] != 0;
    F->PressureIters = ip[3];
    F->EarlyExit = ip[4] != 0;
    F->PreciseAdvection = ip[5] != 0;
  });
}
void orc_set_threads(void* h, int n) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, F->threads = n);
}
void orc_simulate(void* h, float dt) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, F->Simulate(dt));
}
// White-box pass entry points, mirroring the unexported passes the reference tests call.
enum {
  P_COPYBORDER_UV = 0,
  P_VISCOSITY,
  P_MAKE_INCOMPRESSIBLE,
  P_CONFINEMENT,
  P_TURBULENCE,
  P_BORDERS,
  P_ADVECT_VELOCITY,
  P_ADVECT_SMOKE,
  P_ADVECT_VELOCITY_BFECC,
  P_ADVECT_SMOKE_BFECC,
  P_CLEAR_P,
};
int orc_pass(void* h, int pass, float dt, int iters) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, {
    switch (pass) {
      case P_COPYBORDER_UV:
        F->copyBorder(F->newU, F->U);
        F->copyBorder(F->newV, F->V);
        break;
      case P_VISCOSITY: F->applyViscosity(dt); break;
      case P_MAKE_INCOMPRESSIBLE: F->makeIncompressible(iters, dt); break;
      case P_CONFINEMENT: F->applyVorticityConfinement(dt); break;
      case P_TURBULENCE: F->addTurbulence(dt); break;
      case P_BORDERS: F->handleBorders(); break;
      case P_ADVECT_VELOCITY: F->advectVelocity(dt); break;
      case P_ADVECT_SMOKE: F->advectSmoke(dt); break;
      case P_ADVECT_VELOCITY_BFECC: F->advectVelocityBFECC(dt); break;
      case P_ADVECT_SMOKE_BFECC: F->advectSmokeBFECC(dt); break;
      case P_CLEAR_P: std::fill(F->p.begin(), F->p.end(), 0); break;
      default: return 1;
    }
  });
  return 0;
}
int orc_last_sweeps(void* h) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, return F->lastSweeps);
}
int orc_sweep_maxdiv(void* h, double* out, int cap) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, {
    int n = std::min<int>(cap, (int)F->sweepMaxDiv.size());
    for (int k = 0; k < n; k++) out[k] = (double)F->sweepMaxDiv[k];
    return (int)F->sweepMaxDiv.size();
  });
}
int orc_invalid_index(void* h) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, return F->invalidIndex ? 1 : 0);
}
// Shadow decisions for the FP64 model.
void orc_set_forced_sweeps(void* h, int n) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, F->forcedSweeps = n);
}
void orc_get_turb_gate(void* h, uint8_t* out) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, std::memcpy(out, F->turbGateOut.data(), F->turbGateOut.size()));
}
void orc_set_turb_gate(void* h, const uint8_t* in) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, {
    if (in) {
      F->turbGateIn.assign(in, in + F->numCells);
      F->useTurbGateIn = true;
    } else {
      F->useTurbGateIn = false;
    }
  });
}

// Mask / injection API. Returns 0, or 1/2 for an invalid x/y index (the reference panics).
int orc_set_solid(void* h, int i, int j, int v) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, return F->SetSolid(i, j, v != 0));
}
int orc_is_solid(void* h, int i, int j, int* out) {
  OrcHandle* H = (OrcHandle*)h;
  bool b = false;
  int rc;
  DISPATCH(H, rc = F->IsSolid(i, j, &b));
  *out = b ? 1 : 0;
  return rc;
}
int orc_set_velocity(void* h, int i, int j, double u, double v) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, {
    using T = std::remove_reference_t<decltype(F->h)>;
    return F->SetVelocity(i, j, (T)(float)u, (T)(float)v);
  });
}
int orc_add_smoke(void* h, int i, int j, double s) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, {
    using T = std::remove_reference_t<decltype(F->h)>;
    return F->AddSmoke(i, j, (T)(float)s);
  });
}
void orc_reset(void* h) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, F->Reset());
}
void orc_apply_force(void* h, int i, int j, double fx, double fy) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, {
    using T = std::remove_reference_t<decltype(F->h)>;
    F->ApplyForce(i, j, (T)(float)fx, (T)(float)fy);
  });
}
void orc_apply_force_radius(void* h, int cx, int cy, double fx, double fy, int r) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, {
    using T = std::remove_reference_t<decltype(F->h)>;
    F->ApplyForceRadius(cx, cy, (T)(float)fx, (T)(float)fy, r);
  });
}
void orc_apply_force_gaussian(void* h, int cx, int cy, double fx, double fy, double sigma) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, {
    using T = std::remove_reference_t<decltype(F->h)>;
    F->ApplyForceGaussian(cx, cy, (T)(float)fx, (T)(float)fy, (float)sigma);
  });
}
// Differentiable rollout (oracle/fluid_adjoint.inc): buffers are in the handle's
// precision, NX*NY each; gFu/gFv may be null for the loss alone.
double orc_rollout(void* h, int steps, float dt, const void* Fu, const void* Fv, const void* target, void* gFu,
                   void* gFv) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, {
    using T = std::remove_reference_t<decltype(F->h)>;
    const size_t n = (size_t)F->numCells;
    std::vector<T> fu((const T*)Fu, (const T*)Fu + n), fv((const T*)Fv, (const T*)Fv + n);
    std::vector<T> tg((const T*)target, (const T*)target + n), gu, gv;
    orc::Rollout<T> R(&SC);
    const bool grad = gFu && gFv;
    double loss = R.run(steps, (T)dt, fu, fv, tg, grad ? &gu : nullptr, grad ? &gv : nullptr);
    if (grad) {
      std::memcpy(gFu, gu.data(), sizeof(T) * n);
      std::memcpy(gFv, gv.data(), sizeof(T) * n);
    }
    return loss;
  });
  return 0.0;
}
// Particles in float32 (the reference's Particle fields are float32).
void orc_advect_particles(void* h, float dt, int n, float* x, float* y, float* age, const float* maxAge,
                          uint8_t* keep) {
  OrcHandle* H = (OrcHandle*)h;
  if (H->prec != 32) return;
  H->s32.advectParticles(dt, n, x, y, age, maxAge, keep);
}
void orc_render(void* h, int view, int overlay, uint8_t* pix) {
  OrcHandle* H = (OrcHandle*)h;
  if (H->prec != 32) return;
  H->s32.render(view, overlay != 0, pix);
}
void orc_set_circular_obstacle(void* h, int cx, int cy, int r) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, F->SetCircularObstacle(cx, cy, r));
}
double orc_adaptive_dt(void* h, float basedt) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, return (double)F->GetAdaptiveTimeStep(basedt));
}
double orc_max_divergence(void* h) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, return (double)F->MaxDivergence());
}
void orc_sample_velocity(void* h, float x, float y, double* u, double* v) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, {
    using T = std::remove_reference_t<decltype(F->h)>;
    T uu, vv;
    F->SampleVelocity(x, y, &uu, &vv);
    *u = uu;
    *v = vv;
  });
}
// sampleField (fluid.go:357-398) of one field at (x, y) with the offsets of `kind`
// (0 = U, 1 = V, 2 = M): the bilinear sampler on its own, for the closed-form pins.
double orc_sample_field(void* h, int kind, int id, double x, double y) {
  OrcHandle* H = (OrcHandle*)h;
  if (id < 0 || id >= orc::F_COUNT || kind < 0 || kind > 2) return 0.0;
  DISPATCH(H, {
    using T = std::remove_reference_t<decltype(F->h)>;
    return (double)F->sample((T)(float)x, (T)(float)y, *F->field(id), kind);
  });
}
// Multigrid helpers one at a time (fluid.go:601-758, 1123-1149). Buffers have the handle's
// precision; coarse buffers hold ((NX+1)/2) x ((NY+1)/2) values.
void orc_mg_dims(void* h, int* cnx, int* cny) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, {
    *cnx = (F->NX + 1) / 2;
    *cny = (F->NY + 1) / 2;
  });
}
void orc_mg_residual(void* h, void* out) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, {
    auto r = F->computePressureResidual();
    std::memcpy(out, r.data(), sizeof(r[0]) * r.size());
  });
}
void orc_mg_restrict(void* h, const void* fine, void* coarse) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, {
    using T = std::remove_reference_t<decltype(F->h)>;
    std::vector<T> in((const T*)fine, (const T*)fine + F->numCells);
    auto r = F->restrictResidual(in);
    std::memcpy(coarse, r.data(), sizeof(T) * r.size());
  });
}
void orc_mg_coarse_solve(void* h, const void* rhs, void* out) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, {
    using T = std::remove_reference_t<decltype(F->h)>;
    const size_t nc = (size_t)((F->NX + 1) / 2) * ((F->NY + 1) / 2);
    std::vector<T> in((const T*)rhs, (const T*)rhs + nc);
    auto r = F->solveCoarseGrid(in);
    std::memcpy(out, r.data(), sizeof(T) * r.size());
  });
}
void orc_mg_prolong(void* h, const void* coarse, void* fine) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, {
    using T = std::remove_reference_t<decltype(F->h)>;
    const size_t nc = (size_t)((F->NX + 1) / 2) * ((F->NY + 1) / 2);
    std::vector<T> in((const T*)coarse, (const T*)coarse + nc);
    auto r = F->prolongateCorrection(in);
    std::memcpy(fine, r.data(), sizeof(T) * r.size());
  });
}
void orc_mg_correct(void* h, const void* corr, double cp) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, {
    using T = std::remove_reference_t<decltype(F->h)>;
    std::vector<T> in((const T*)corr, (const T*)corr + F->numCells);
    F->applyPressureCorrection(in, (T)(float)cp);
  });
}
// kind 0 = vorticity, 1 = velocity magnitude; out has the precision of the handle.
void orc_view(void* h, int kind, void* out, double* mn, double* mx) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, {
    using T = std::remove_reference_t<decltype(F->h)>;
    T a, b;
    if (kind == 0)
      F->Vorticity((T*)out, &a, &b);
    else
      F->VelocityMagnitude((T*)out, &a, &b);
    *mn = a;
    *mx = b;
  });
}
void orc_field_minmax(void* h, int id, double* mn, double* mx) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, {
    using T = std::remove_reference_t<decltype(F->h)>;
    T a, b;
    F->FieldMinMax(id, &a, &b);
    *mn = a;
    *mx = b;
  });
}

// Scene layer (main.go).
void orc_scene_walls(void* h, int top, int bottom, int left, int right) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, {
    SC.wallTop = top;
    SC.wallBottom = bottom;
    SC.wallLeft = left;
    SC.wallRight = right;
  });
}
void orc_scene_apply_walls(void* h) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, SC.applyWallSettings());
}
void orc_scene_clear_walls(void* h) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, SC.clearWalls());
}
void orc_scene_preset(void* h, int preset, float lidU) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, SC.resetToPreset(preset, lidU));
}
void orc_scene_jet(void* h, int on, int x, int lo, int hi, float u, float smoke) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, {
    SC.jet = on != 0;
    SC.jetX = x;
    SC.jetLo = lo;
    SC.jetHi = hi;
    SC.jetU = u;
    SC.jetSmoke = smoke;
  });
}
int orc_scene_jet_on(void* h) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, return SC.jet ? 1 : 0);
}
void orc_scene_add_source(void* h, int sink, int i, int j, float u, float v) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, {
    orc::Source s{i, j, (double)u, (double)v};
    if (sink)
      SC.sinks.push_back(s);
    else
      SC.sources.push_back(s);
  });
}
void orc_scene_clear_sources(void* h) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, {
    SC.sources.clear();
    SC.sinks.clear();
  });
}
int orc_scene_num_sources(void* h, int sink) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, return (int)(sink ? SC.sinks.size() : SC.sources.size()));
}
void orc_scene_apply_sources(void* h) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, SC.applySources());
}
void orc_scene_update(void* h, float dt) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, SC.update(dt));
}
void orc_scene_draw_line(void* h, int x0, int y0, int x1, int y1, int v) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, SC.drawLine(x0, y0, x1, y1, v != 0));
}
void orc_scene_smoke_brush(void* h, int ci, int cj, int r, float strength) {
  OrcHandle* H = (OrcHandle*)h;
  DISPATCH(H, SC.smokeBrush(ci, cj, r, strength));
}

}  // extern "C"
