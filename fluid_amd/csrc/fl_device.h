// Device-side helpers shared by the step kernels: Go float semantics, the Go/Cephes
// sine, and the clamped bilinear sample of the reference. Everything here is compiled
// with -ffp-contract=off so each float operation rounds exactly as Go's float32 does.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fl {

// Grid description handed to every step kernel by value.
struct Grid {
  int NX, NY;     // NumX, NumY (width+2, height+2)
  float h;        // cell size
  float h1;       // float32(1.0/h)          fluid.go:360
  float hs;       // float32(0.5*h)          fluid.go:361 (sample offset)
  float ha;       // h/2                     fluid.go:298 (advect face offset)
  float xmax;     // float32(NumX)*h         fluid.go:363
  float ymax;     // float32(NumY)*h         fluid.go:364
  int precise;    // back-trace in cell units (sample_cells) instead of absolute positions
};

// Cell flag bits (uint8 per cell), derived from S (S must be exactly 0.0 or 1.0).
enum : uint32_t {
  CF_FLUID = 1u << 0,   // S[i,j] != 0
  CF_X0 = 1u << 1,      // S[i-1,j] != 0
  CF_X1 = 1u << 2,      // S[i+1,j] != 0
  CF_Y0 = 1u << 3,      // S[i,j-1] != 0
  CF_Y1 = 1u << 4,      // S[i,j+1] != 0
  CF_GS = 1u << 5,      // interior, fluid, and at least one fluid neighbour: updated by a GS sweep
  CF_INTERIOR = 1u << 6 // 1 <= i <= NX-2 and 1 <= j <= NY-2
};

// Go builtin min/max on float32: NaN propagates, min(-0,+0) = -0, max(-0,+0) = +0. That is
// IEEE 754-2019 minimum/maximum, one v_minimum3_f32 / v_maximum3_f32 on gfx950. (Which NaN
// comes back is unspecified in both; NaN payloads already differ between x86 and the GPU.)
__device__ __forceinline__ float go_minf(float a, float b) { return __builtin_elementwise_minimum(a, b); }
__device__ __forceinline__ float go_maxf(float a, float b) { return __builtin_elementwise_maximum(a, b); }

// Go math.Sin for 0 <= |x| < 2^29 (Cephes sin.c octant reduction and polynomials),
// evaluated in double without contraction.
__device__ __forceinline__ double go_sin(double x) {
  if (x == 0.0 || x != x) return x;
  bool sign = false;
  if (x < 0) {
    x = -x;
    sign = true;
  }
  const double PI4A = 7.85398125648498535156e-1;
  const double PI4B = 3.77489470793079817668e-8;
  const double PI4C = 2.69515142907905952645e-15;
  uint64_t j = (uint64_t)(x * 0x1.45f306dc9c883p+0);
  double y = (double)j;
  if (j & 1) {
    j++;
    y += 1.0;
  }
  j &= 7;
  double z = ((x - y * PI4A) - y * PI4B) - y * PI4C;
  if (j > 3) {
    sign = !sign;
    j -= 4;
  }
  double zz = z * z;
  if (j == 1 || j == 2) {
    double c = -1.13585365213876817300e-11;
    c = c * zz + 2.08757008419747316778e-9;
    c = c * zz + -2.75573141792967388112e-7;
    c = c * zz + 2.48015872888517045348e-5;
    c = c * zz + -1.38888888888730564116e-3;
    c = c * zz + 4.16666666666665929218e-2;
    y = 1.0 - 0.5 * zz + zz * zz * c;
  } else {
    double s = 1.58962301576546568060e-10;
    s = s * zz + -2.50507477628578072866e-8;
    s = s * zz + 2.75573136213857245213e-6;
    s = s * zz + -1.98412698295895385996e-4;
    s = s * zz + 8.33333333332211858878e-3;
    s = s * zz + -1.66666666666666307295e-1;
    y = z + z * zz * s;
  }
  return sign ? -y : y;
}

// float32(math.Sin(float64(float32(k) * 0.01)))   (fluid.go:509-514)
__device__ __forceinline__ float turb_noise(int k) {
  float seed = (float)k * 0.01f;
  return (float)go_sin((double)seed);
}

// A field element at a 32-bit cell index from a kernel-uniform base. The byte offset stays a
// 32-bit unsigned value, so the compiler addresses with the base in scalar registers instead of
// per-lane 64-bit arithmetic. Grids are limited to 2^28 cells (fl_create, fl_batch_create), so
// byte offsets stay below 2^30.
__device__ __forceinline__ float ld_cell(const float* __restrict__ base, int idx) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + (uint32_t)idx * 4u);
}
__device__ __forceinline__ void st_cell(float* __restrict__ base, int idx, float v) {
  *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + (uint32_t)idx * 4u) = v;
}
__device__ __forceinline__ uint32_t ld_flag(const uint8_t* __restrict__ base, int idx) {
  return *(base + (uint32_t)idx);
}
// Wave-uniform: some active lane of the wavefront has p.
__device__ __forceinline__ bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0; }

// Two and three consecutive elements of a row in one load (global_load_dwordx2 / dwordx3),
// 4-byte aligned. A wavefront's gather or stencil load costs the L1 about the same whatever its
// width, so loading a cell's neighbours along j together cuts the vector-memory work per cell.
typedef float f2u __attribute__((ext_vector_type(2), aligned(4)));
typedef float f3u __attribute__((ext_vector_type(3), aligned(4)));
__device__ __forceinline__ f2u ld_pair(const float* __restrict__ base, int idx) {
  return *reinterpret_cast<const f2u*>(reinterpret_cast<const char*>(base) + (uint32_t)idx * 4u);
}
__device__ __forceinline__ f3u ld_trip(const float* __restrict__ base, int idx) {
  return *reinterpret_cast<const f3u*>(reinterpret_cast<const char*>(base) + (uint32_t)idx * 4u);
}

// The corners (x0,y0), (x1,y0), (x1,y1), (x0,y1) of a bilinear sample, y1 = y0 or y0 + 1,
// both <= NY-1: one pair per row, starting at ys = min(y0, NY-2) so that it holds both columns.
struct Corners {
  float a, b, c, d;
};
__device__ __forceinline__ Corners ld_corners(const float* __restrict__ f, int x0, int x1, int y0, int y1,
                                              int NY) {
  const int ys = min(y0, NY - 2);
  const f2u p0 = ld_pair(f, x0 * NY + ys), p1 = ld_pair(f, x1 * NY + ys);
  const bool a0 = y0 == ys, a1 = y1 == ys;
  return Corners{a0 ? p0.x : p0.y, a0 ? p1.x : p1.y, a1 ? p1.x : p1.y, a1 ? p0.x : p0.y};
}

// int(math.Floor(float64(v))) with NaN/overflow mapped to 0 and flagged (the
// reference would panic on the resulting slice index).
__device__ __forceinline__ int floor_index(float v, int* bad) {
  float f = floorf(v);
  if (!(f >= -2147483648.0f && f < 2147483648.0f)) {
    *bad = 1;
    return 0;
  }
  return (int)f;
}

// sampleField / sampleFieldFrom (fluid.go:357-398, 1055-1091). (dx,dy) is (0,hs) for
// U, (hs,0) for V and (hs,hs) for M. Evaluation order matches the reference exactly.
// Corner source of a sample: the field in global memory.
struct GlobalSrc {
  const float* __restrict__ f;
  __device__ __forceinline__ Corners corners(int x0, int x1, int y0, int y1, int NY, bool) const {
    return ld_corners(f, x0, x1, y0, y1, NY);
  }
};

template <class S>
__device__ __forceinline__ float sample_field(float x, float y, const S& src, float dx, float dy, const Grid& g,
                                              int* bad) {
  x = go_maxf(go_minf(x, g.xmax), g.h);
  y = go_maxf(go_minf(y, g.ymax), g.h);
  int x0 = min(floor_index((x - dx) * g.h1, bad), g.NX - 1);
  float tx = ((x - dx) - (float)x0 * g.h) * g.h1;
  int x1 = min(x0 + 1, g.NX - 1);
  int y0 = min(floor_index((y - dy) * g.h1, bad), g.NY - 1);
  float ty = ((y - dy) - (float)y0 * g.h) * g.h1;
  int y1 = min(y0 + 1, g.NY - 1);
  const bool exact = x0 >= 0 && y0 >= 0;  // else x1 or y1 may equal x0 or y0 below the clamp
  if (x0 < 0 || y0 < 0) {
    *bad = 1;
    x0 = max(x0, 0);
    y0 = max(y0, 0);
    x1 = max(x1, 0);
    y1 = max(y1, 0);
  }
  float sx = 1.0f - tx;
  float sy = 1.0f - ty;
  const Corners k = src.corners(x0, x1, y0, y1, g.NY, exact);
  return sx * sy * k.a + tx * sy * k.b + tx * ty * k.c + sx * ty * k.d;
}

// Back-trace in cell units (the precise_advection parameter; the reference always samples
// at absolute positions). Every back-trace of the step lands at X = i + dX, Y = j + dY in
// cell units (the face and sample offsets cancel), dX = (s*u)*h1 for the trace step
// s = -dt or +dt. The reference forms x = i*h + s*u and rounds it at ulp(N*h), ~N*6e-8
// cells; here floor(d) is split off and the fraction d - floor(d) is exact, so the
// position carries only dX's own rounding. The clamp x in [h, N*h] becomes X in
// [1 - o, N - o], o = 1/2 on an axis with a half-cell sample offset (y of U, x of V, both
// of M). X = N exactly maps to sampleField's x0 = N-1, tx = 1. Returns whether X was
// inside the clamp range.
__device__ __forceinline__ bool axis_cells(int base, float d, int N, bool half, int& i0, float& t, int* bad) {
  if (d != d) {  // a NaN velocity: the reference flags the index and samples NaN
    *bad = 1;
    i0 = 0;
    t = d;
    return false;
  }
  bool inside = true;
  const int loI = half ? 0 : 1, hiI = half ? N - 1 : N;
  const float loF = half ? 0.5f : 0.0f, hiF = loF;
  const float fl = floorf(d);
  if (fl < -16777216.0f) {
    i0 = loI;
    t = loF;
    inside = false;
  } else if (fl > 16777216.0f) {
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
  if (i0 > N - 1) {
    i0 = N - 1;
    t = 1.0f;
  }
  return inside;
}
// The bilinear sample of sampleField (fluid.go:389-397) at (i + dX, j + dY); hx/hy: the
// axis carries a half-cell sample offset.
template <class S>
__device__ __forceinline__ float sample_cells(int i, int j, float dX, float dY, const S& src, bool hx, bool hy,
                                              const Grid& g, int* bad) {
  int x0, y0;
  float tx, ty;
  axis_cells(i, dX, g.NX, hx, x0, tx, bad);
  axis_cells(j, dY, g.NY, hy, y0, ty, bad);
  const int x1 = min(x0 + 1, g.NX - 1), y1 = min(y0 + 1, g.NY - 1);
  float sx = 1.0f - tx;
  float sy = 1.0f - ty;
  const Corners k = src.corners(x0, x1, y0, y1, g.NY, true);  // x1 = min(x0 + 1, NX - 1), y1 likewise
  return sx * sy * k.a + tx * sy * k.b + tx * ty * k.c + sx * ty * k.d;
}

__device__ __forceinline__ float sample_field(float x, float y, const float* __restrict__ f, float dx, float dy,
                                              const Grid& g, int* bad) {
  return sample_field(x, y, GlobalSrc{f}, dx, dy, g, bad);
}
__device__ __forceinline__ float sample_cells(int i, int j, float dX, float dY, const float* __restrict__ f, bool hx,
                                              bool hy, const Grid& g, int* bad) {
  return sample_cells(i, j, dX, dY, GlobalSrc{f}, hx, hy, g, bad);
}

}  // namespace fl
