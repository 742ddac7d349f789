// Rendering, field ranges and particles on the device (the application layer around the
// step: main/main.go:512-652, main/colors.go:10-84, pkg/fluid/smoke.go, pressure.go).
//
// - Field min/max as the reference scans them (first value in index order wins among equals,
//   NaN never selected), reduced on the device in two fixed passes, so Pressure()/Smoke()
//   ranges and the colormaps need no field download.
// - Game.Draw's image: the scientific (blue-cyan-green-yellow-red) or diverging colormap of a
//   field, flipped vertically, with solid cells black; the host adds the few source/sink
//   marker pixels.
// - Game.advectParticles: age, RK2 midpoint trace with SampleVelocity, survival test.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "fl_internal.h"

namespace fl {

static constexpr int RB = 256;          // threads per reduction block
static constexpr int RBLOCKS = 2048;    // partials of the first pass

struct MinMax {
  float mn, mx;
  long long imn, imx;  // index of the chosen element; -1 = the scan's initial value
};

__device__ __forceinline__ void mm_take_min(MinMax& a, float v, long long i) {
  if (v < a.mn || (v == a.mn && i < a.imn)) {
    a.mn = v;
    a.imn = i;
  }
}
__device__ __forceinline__ void mm_take_max(MinMax& a, float v, long long i) {
  if (v > a.mx || (v == a.mx && i < a.imx)) {
    a.mx = v;
    a.imx = i;
  }
}
__device__ __forceinline__ MinMax mm_init() { return MinMax{FLT_MAX, -FLT_MAX, -1, -1}; }

// The sequential scan keeps the first element that compares below (above) the running
// value, starting from +-MaxFloat32 (smoke.go:6-15): i.e. among the elements equal to the
// extreme the lowest index, and the initial value when nothing beats it. `interior`: only
// fluid interior cells take part (Vorticity, VelocityMagnitude, fluid.go:806-872).
__device__ __forceinline__ void mm_block(MinMax m, MinMax* out) {
  __shared__ MinMax red[RB];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int o = RB / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      MinMax b = red[threadIdx.x + o];
      MinMax& a = red[threadIdx.x];
      if (b.imn >= 0) mm_take_min(a, b.mn, b.imn);
      if (b.imx >= 0) mm_take_max(a, b.mx, b.imx);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}
// A thread scans its elements in increasing index order (four at a time, grid-strided, when
// `vec` says f is 16-byte and flags 4-byte aligned; then the rest one at a time, grid-strided,
// all above the vector part's indices), so a strict comparison keeps the lowest index among
// equal values; indices are compared only when threads merge. NaN and values tying the
// initial bound are never taken.
__device__ __forceinline__ void mm_scan(MinMax& m, float v, bool take, long long k) {
  if (take && v < m.mn) {
    m.mn = v;
    m.imn = k;
  }
  if (take && v > m.mx) {
    m.mx = v;
    m.imx = k;
  }
}
__device__ __forceinline__ bool mm_cell(uint32_t fl) { return (fl & CF_INTERIOR) && (fl & CF_FLUID); }
__global__ __launch_bounds__(RB) void k_minmax_partial(const float* __restrict__ f, const uint8_t* __restrict__ flags,
                                                       int interior, int vec, size_t n, MinMax* __restrict__ part) {
  MinMax m = mm_init();
  const size_t n4 = vec ? n / 4 : 0, stride = (size_t)gridDim.x * RB, t = blockIdx.x * (size_t)RB + threadIdx.x;
  const float4* f4 = reinterpret_cast<const float4*>(f);
  const uchar4* fl4 = reinterpret_cast<const uchar4*>(flags);
  for (size_t q = t; q < n4; q += stride) {
    const float4 v = f4[q];
    const uchar4 fl = interior ? fl4[q] : make_uchar4(0, 0, 0, 0);
    const long long k = 4 * (long long)q;
    mm_scan(m, v.x, !interior || mm_cell(fl.x), k);
    mm_scan(m, v.y, !interior || mm_cell(fl.y), k + 1);
    mm_scan(m, v.z, !interior || mm_cell(fl.z), k + 2);
    mm_scan(m, v.w, !interior || mm_cell(fl.w), k + 3);
  }
  for (size_t k = 4 * n4 + t; k < n; k += stride) mm_scan(m, f[k], !interior || mm_cell(flags[k]), (long long)k);
  mm_block(m, part + blockIdx.x);
}
__global__ __launch_bounds__(RB) void k_minmax_final(const MinMax* __restrict__ part, int nparts,
                                                     float* __restrict__ out) {
  MinMax m = mm_init();
  for (int k = threadIdx.x; k < nparts; k += RB) {
    const MinMax b = part[k];
    if (b.imn >= 0) mm_take_min(m, b.mn, b.imn);
    if (b.imx >= 0) mm_take_max(m, b.mx, b.imx);
  }
  __shared__ MinMax res;
  mm_block(m, &res);
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = res.imn >= 0 ? res.mn : FLT_MAX;
    out[1] = res.imx >= 0 ? res.mx : -FLT_MAX;
  }
}

// Go's uint8(float32): truncation toward zero, modulo 256 (values in range here).
__device__ __forceinline__ uint32_t to_u8(float v) {
  if (!(v > -2147483648.0f && v < 2147483648.0f)) return 0u;
  return (uint32_t)(uint8_t)(long long)v;
}
__device__ __forceinline__ uint32_t rgba(float r, float g, float b) {
  return to_u8(255.0f * r) | (to_u8(255.0f * g) << 8) | (to_u8(255.0f * b) << 16) | (0xffu << 24);
}
// getSciValue (colors.go:46-84)
__device__ __forceinline__ uint32_t sci_color(float val, float mn, float mx) {
  val = go_minf(go_maxf(val, mn), mx - 0.0001f);
  const float d = mx - mn;
  if (d <= 0.0f)
    val = 0.5f;
  else
    val = (val - mn) / d;
  const float m = 0.25f;
  const float num = (float)floor((double)(val / m));
  const float s = (val - num * m) / m;
  float r = 0.0f, g = 0.0f, b = 0.0f;
  if (num == 0.0f) {
    g = s;
    b = 1.0f;
  } else if (num == 1.0f) {
    g = 1.0f;
    b = 1.0f - s;
  } else if (num == 2.0f) {
    r = s;
    g = 1.0f;
  } else if (num == 3.0f) {
    r = 1.0f;
    g = 1.0f - s;
  }
  return rgba(r, g, b);
}
// getDivergingColor (colors.go:10-44)
__device__ __forceinline__ uint32_t diverging_color(float val, float mn, float mx) {
  float absMax;
  if (mn != mn || mx != mx)
    absMax = __int_as_float(0x7fc00000);
  else
    absMax = (float)fmax(fabs((double)mn), fabs((double)mx));
  if (absMax < 1e-8f) return 0xffffffffu;
  float t = val / absMax;
  if (t > 1.0f) t = 1.0f;
  if (t < -1.0f) t = -1.0f;
  if (t >= 0.0f) return rgba(1.0f, 1.0f - t, 1.0f - t);
  const float a = -t;
  return rgba(1.0f - a, 1.0f - a, 1.0f);
}

// One pixel per thread: image (x = i, row j) shows fluid cell (i, NumY-1-j) (main.go:614-646);
// solid cells are painted black afterwards in the reference, i.e. win (main.go:560-572).
// Image row r shows field row j = NY-1-r, pixel i field column i: a transpose of the j-contiguous
// field. A 64x64 block of cells is read along j, coloured, and written along i through LDS.
static constexpr int RT = 64;
__global__ __launch_bounds__(256) void k_render(int NX, int NY, const float* __restrict__ vals,
                                                const float* __restrict__ S, const float* __restrict__ range,
                                                int diverging, uint32_t* __restrict__ pix) {
  __shared__ uint32_t tile[RT][RT + 1];
  const int j0 = blockIdx.x * RT, i0 = blockIdx.y * RT;
  const int lane = threadIdx.x & (RT - 1), row = threadIdx.x / RT;
  const float mn = range[0], mx = range[1];
  for (int r = row; r < RT; r += 256 / RT) {
    const int i = i0 + r, j = j0 + lane;
    if (i < NX && j < NY) {
      const size_t c = (size_t)i * NY + j;
      uint32_t px;
      if (S[c] == 0.0f)
        px = 0xff000000u;
      else
        px = diverging ? diverging_color(vals[c], mn, mx) : sci_color(vals[c], mn, mx);
      tile[lane][r] = px;
    }
  }
  __syncthreads();
  for (int r = row; r < RT; r += 256 / RT) {
    const int j = j0 + r, i = i0 + lane;
    if (i < NX && j < NY) pix[(size_t)(NY - 1 - j) * NX + i] = tile[r][lane];
  }
}

// Game.advectParticles (main.go:512-546) for one particle.
__global__ void k_particles(Grid g, const float* __restrict__ U, const float* __restrict__ V,
                            const float* __restrict__ S, float dt, int n, float* __restrict__ x,
                            float* __restrict__ y, float* __restrict__ age, const float* __restrict__ maxAge,
                            uint8_t* __restrict__ keep) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float a = age[k] + dt;
  age[k] = a;
  keep[k] = 0;
  if (a > maxAge[k]) return;
  int bad = 0;
  float px = x[k], py = y[k];
  const float u1 = sample_field(px, py, U, 0.0f, g.hs, g, &bad);
  const float v1 = sample_field(px, py, V, g.hs, 0.0f, g, &bad);
  const float mx = px + 0.5f * dt * u1;
  const float my = py + 0.5f * dt * v1;
  const float u2 = sample_field(mx, my, U, 0.0f, g.hs, g, &bad);
  const float v2 = sample_field(mx, my, V, g.hs, 0.0f, g, &bad);
  px = px + dt * u2;
  py = py + dt * v2;
  x[k] = px;
  y[k] = py;
  const float qx = px / g.h, qy = py / g.h;
  if (!(qx > -2147483648.0f && qx < 2147483648.0f && qy > -2147483648.0f && qy < 2147483648.0f)) return;
  const int fi = (int)qx, fj = (int)qy;  // truncation toward zero, as Go's int()
  if (fi < 0 || fi >= g.NX || fj < 0 || fj >= g.NY) return;
  keep[k] = S[fi * g.NY + fj] != 0.0f ? 1 : 0;
}

// ---- host side -------------------------------------------------------------------------
static int rerr(int code, const std::string& m) {
  set_error(m);
  return code;
}

// Scratch of these calls: [NX*NY floats of view values, padded to 256 B][RBLOCKS MinMax]
// [2 floats range, 16 B][pixels or particle arrays], grown on demand.
static int grow_scratch(Ctx* c, size_t bytes, char** out) {
  if (bytes > c->rs_cap) {
    if (c->rs_mem) FL_HIP(hipFree(c->rs_mem));
    c->rs_mem = nullptr;
    FL_HIP(hipMalloc(&c->rs_mem, bytes));
    c->rs_cap = bytes;
  }
  *out = c->rs_mem;
  return FL_OK;
}

// Static scratch of fl_advect_particles_noalloc, for callers that cannot allocate: the
// particles' x, y, age and max_age, FL_PARTICLES_NOALLOC_MAX floats each, then their keep
// bytes. Each GPU holds its own copy for the life of the process; every call on a GPU uses
// it, so overlapping calls on one GPU (two contexts, two threads) race. Nothing in it must
// start at zero: every call copies all inputs in and k_particles writes every keep byte.
static __device__ __align__(256) char scratch[17 * (size_t)FL_PARTICLES_NOALLOC_MAX];
static constexpr int MAX_DEVICES = 64;
static char* g_scratch_ptr[MAX_DEVICES] = {};
static std::mutex g_scratch_mu;

// Device address of `scratch` on device `dev`, looked up once per device with that device
// current (the caller has made it current). The array's name is a host-side stand-in and is
// never passed to a kernel or a copy.
static int particle_scratch(int dev, char** out) {
  if (dev < 0 || dev >= MAX_DEVICES) return rerr(FL_ERR_INTERNAL, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return rerr(FL_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<char*>(p);
  }
  *out = g_scratch_ptr[dev];
  return FL_OK;
}

// Game.advectParticles for n particles whose device arrays are carved from `mem` (17 n
// bytes: x, y, age, max_age as float [n] each, keep as uint8 [n]). Copies below the staging
// threshold are issued asynchronously with one synchronisation at the end; larger ones go
// through the pinned staging of copy_to_device / copy_to_host.
static int advect_particles_in(Ctx* c, float dt, int n, float* x, float* y, float* age, const float* max_age,
                               uint8_t* keep, char* mem) {
  const size_t fb = 4 * (size_t)n;
  float* dx = reinterpret_cast<float*>(mem);
  float* dy = dx + n;
  float* da = dy + n;
  float* dm = da + n;
  uint8_t* dk = reinterpret_cast<uint8_t*>(dm + n);
  const bool direct = copy_is_direct(c, fb);
  int rc;
  if (direct) {
    FL_HIP(hipMemcpyAsync(dx, x, fb, hipMemcpyHostToDevice, c->stream));
    FL_HIP(hipMemcpyAsync(dy, y, fb, hipMemcpyHostToDevice, c->stream));
    FL_HIP(hipMemcpyAsync(da, age, fb, hipMemcpyHostToDevice, c->stream));
    FL_HIP(hipMemcpyAsync(dm, max_age, fb, hipMemcpyHostToDevice, c->stream));
  } else if ((rc = copy_to_device(c, dx, x, fb)) || (rc = copy_to_device(c, dy, y, fb)) ||
             (rc = copy_to_device(c, da, age, fb)) || (rc = copy_to_device(c, dm, max_age, fb))) {
    return rc;
  }
  hipLaunchKernelGGL(k_particles, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->grid, c->U, c->V, c->S, dt, n,
                     dx, dy, da, dm, dk);
  FL_HIP(hipGetLastError());
  if (direct) {
    FL_HIP(hipMemcpyAsync(x, dx, fb, hipMemcpyDeviceToHost, c->stream));
    FL_HIP(hipMemcpyAsync(y, dy, fb, hipMemcpyDeviceToHost, c->stream));
    FL_HIP(hipMemcpyAsync(age, da, fb, hipMemcpyDeviceToHost, c->stream));
    FL_HIP(hipMemcpyAsync(keep, dk, (size_t)n, hipMemcpyDeviceToHost, c->stream));
    FL_HIP(hipStreamSynchronize(c->stream));
    return FL_OK;
  }
  if ((rc = copy_to_host(c, x, dx, fb)) || (rc = copy_to_host(c, y, dy, fb)) || (rc = copy_to_host(c, age, da, fb)))
    return rc;
  return copy_to_host(c, keep, dk, (size_t)n);
}

static int device_minmax(Ctx* c, const float* f, int interior, MinMax* part, float* range) {
  const int vec = (uintptr_t)f % 16 == 0 && (!interior || (uintptr_t)c->flags % 4 == 0);
  hipLaunchKernelGGL(k_minmax_partial, dim3(RBLOCKS), dim3(RB), 0, c->stream, f, c->flags, interior, vec, c->N,
                     part);
  hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(RB), 0, c->stream, part, RBLOCKS, range);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

// The scan's MinValue/MaxValue of a device field, over every cell or (`interior`) the fluid
// interior cells, reduced on the device; the flags must be current.
int field_range(Ctx* c, const float* f, int interior, float* mn, float* mx) {
  char* mem;
  const size_t off = sizeof(MinMax) * RBLOCKS;
  int rc = grow_scratch(c, off + 2 * sizeof(float), &mem);
  if (rc) return rc;
  if ((rc = device_minmax(c, f, interior, reinterpret_cast<MinMax*>(mem), reinterpret_cast<float*>(mem + off))))
    return rc;
  float r[2];
  FL_HIP(hipMemcpyAsync(r, mem + off, sizeof(r), hipMemcpyDeviceToHost, c->stream));
  FL_HIP(hipStreamSynchronize(c->stream));
  if (mn) *mn = r[0];
  if (mx) *mx = r[1];
  return FL_OK;
}

}  // namespace fl

using namespace fl;

extern "C" {

int fl_field_minmax(fl_ctx* ctx, int field, float* mn, float* mx) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  if (!ctx) return rerr(FL_ERR_INVALID_ARG, "null context");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  float* f = nullptr;
  switch (field) {
    case FL_FIELD_U: f = c->U; break;
    case FL_FIELD_V: f = c->V; break;
    case FL_FIELD_NEWU: f = c->nU; break;
    case FL_FIELD_NEWV: f = c->nV; break;
    case FL_FIELD_P: f = c->p; break;
    case FL_FIELD_S: f = c->S; break;
    case FL_FIELD_M: f = c->M; break;
    case FL_FIELD_NEWM: f = c->nM; break;
    default: return rerr(FL_ERR_INVALID_ARG, "bad field");
  }
  int rc = ctx_prepare(c);
  if (rc) return rc;
  return field_range(c, f, 0, mn, mx);
}

int fl_render(fl_ctx* ctx, int view, int overlay, uint8_t* rgba, size_t nbytes) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  if (!ctx || !rgba) return rerr(FL_ERR_INVALID_ARG, "null argument");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (view < 0 || view > 3) return rerr(FL_ERR_INVALID_ARG, "view must be 0 smoke, 1 pressure, 2 speed, 3 vorticity");
  if (nbytes != 4 * c->N) return rerr(FL_ERR_INVALID_ARG, "the image holds NumX*NumY RGBA pixels");
  int rc = ctx_prepare(c);
  if (rc) return rc;
  char* mem;
  const size_t o_part = (4 * c->N + 255) / 256 * 256, o_range = o_part + sizeof(MinMax) * RBLOCKS,
               o_pix = o_range + 16;
  if ((rc = grow_scratch(c, o_pix + 4 * c->N, &mem))) return rc;
  float* vals = reinterpret_cast<float*>(mem);
  float* range = reinterpret_cast<float*>(mem + o_range);
  uint32_t* pix = reinterpret_cast<uint32_t*>(mem + o_pix);
  const float* f;
  if (view == 0 || view == 1) {
    f = view == 0 ? c->M : c->p;
  } else {
    if ((rc = launch_view(c, view == 2 ? 1 : 0, vals))) return rc;
    f = vals;
  }
  if ((rc = device_minmax(c, f, view >= 2, reinterpret_cast<MinMax*>(mem + o_part), range))) return rc;
  hipLaunchKernelGGL(k_render, dim3((c->NY + RT - 1) / RT, (c->NX + RT - 1) / RT), dim3(256), 0, c->stream, c->NX, c->NY, f, c->S,
                     range, view == 3 ? 1 : 0, pix);
  FL_HIP(hipGetLastError());
  if ((rc = copy_to_host(c, rgba, pix, nbytes))) return rc;
  if (overlay) {  // source and sink markers (main.go:574-595), in the reference's order
    const long long stride = 4LL * c->NX, len = (long long)nbytes;
    auto mark = [&](const Source& s, uint8_t r, uint8_t g) {
      const long long idx = 4LL * s.I + (long long)(c->NY - s.J - 1) * stride;
      if (idx >= 0 && idx + 3 < len) {
        rgba[idx] = r;
        rgba[idx + 1] = g;
        rgba[idx + 2] = 0;
        rgba[idx + 3] = 255;
      }
    };
    for (const Source& s : c->sources) mark(s, 0, 255);
    for (const Source& s : c->sinks) mark(s, 255, 0);
  }
  return FL_OK;
}

int fl_advect_particles(fl_ctx* ctx, float dt, int n, float* x, float* y, float* age, const float* max_age,
                        uint8_t* keep) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  if (!ctx) return rerr(FL_ERR_INVALID_ARG, "null context");
  if (n < 0 || (n > 0 && (!x || !y || !age || !max_age || !keep))) return rerr(FL_ERR_INVALID_ARG, "bad arguments");
  if (n == 0) return FL_OK;
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (hipSetDevice(c->dev) != hipSuccess) return rerr(FL_ERR_HIP, "hipSetDevice");
  int rc = ctx_prepare(c);
  if (rc) return rc;
  char* mem;
  if ((rc = grow_scratch(c, 17 * (size_t)n, &mem))) return rc;
  return advect_particles_in(c, dt, n, x, y, age, max_age, keep, mem);
}

int fl_advect_particles_noalloc(fl_ctx* ctx, float dt, int n, float* x, float* y, float* age,
                                const float* max_age, uint8_t* keep) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  if (!ctx) return rerr(FL_ERR_INVALID_ARG, "null context");
  if (n < 0 || (n > 0 && (!x || !y || !age || !max_age || !keep))) return rerr(FL_ERR_INVALID_ARG, "bad arguments");
  if (n > FL_PARTICLES_NOALLOC_MAX)
    return rerr(FL_ERR_INVALID_ARG, "fl_advect_particles_noalloc: n = " + std::to_string(n) +
                                        " exceeds the static scratch (" + std::to_string(FL_PARTICLES_NOALLOC_MAX) +
                                        " particles); use fl_advect_particles");
  if (n == 0) return FL_OK;
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (hipSetDevice(c->dev) != hipSuccess) return rerr(FL_ERR_HIP, "hipSetDevice");
  // Queued edits are flushed through the context's edit buffers, which grow by allocation.
  if (c->edits.size() > c->edit_cap)
    return rerr(FL_ERR_INVALID_ARG, "fl_advect_particles_noalloc: " + std::to_string(c->edits.size()) +
                                        " queued edits exceed the edit buffer (" + std::to_string(c->edit_cap) +
                                        "); flush them with a step or fl_sync first");
  char* mem;
  int rc = particle_scratch(c->dev, &mem);
  if (rc) return rc;
  if ((rc = ctx_prepare(c))) return rc;
  return advect_particles_in(c, dt, n, x, y, age, max_age, keep, mem);
}

}  // extern "C"
