// libfluidhip host runtime: the C ABI of include/fluid_hip.h.
//
// A context owns the device copies of the reference's eight buffers (fluid.go:17-23)
// and a host mirror of S. Mutating API calls (SetSolid, SetVelocity, AddSmoke,
// ApplyForce...) are recorded as ordered word-level edits and applied on the device
// by one kernel before the next device operation, so their effect and order match the
// reference's direct slice writes exactly. Simulate enqueues the step kernels on the
// context's stream in the reference's pass order (fluid.go:79-109).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <thread>
#include <limits>
#include <numeric>
#include <string>
#include <vector>

#include "fl_internal.h"

namespace fl {

static thread_local std::string g_err;
static thread_local Ctx* g_err_ctx = nullptr;  // context of the call in progress on this thread
void set_error(const std::string& msg) {
  g_err = msg;
  if (g_err_ctx) g_err_ctx->last_err = msg;
}
CtxErrScope::CtxErrScope(Ctx* c) : prev(g_err_ctx) { g_err_ctx = c; }
CtxErrScope::~CtxErrScope() { g_err_ctx = prev; }

static int err(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

static float go_minf_h(float a, float b) {
  if (a != a) return a;
  if (b != b) return b;
  if (a < b) return a;
  if (b < a) return b;
  return std::signbit(a) ? a : b;
}
static float go_maxf_h(float a, float b) {
  if (a != a) return a;
  if (b != b) return b;
  if (a > b) return a;
  if (b > a) return b;
  return std::signbit(a) ? b : a;
}

static int check_xy(Ctx* c, int i, int j) {
  if (i < 0 || i >= c->NX) return err(FL_ERR_X_RANGE, "invalid x-index: " + std::to_string(i));
  if (j < 0 || j >= c->NY) return err(FL_ERR_Y_RANGE, "invalid y-index: " + std::to_string(j));
  return FL_OK;
}

static void push_op(Ctx* c, int field, size_t idx, bool add, float v) {
  WordOp op;
  op.key = (uint32_t)(field * c->N + idx) | (add ? 0x80000000u : 0u);
  op.kind = add ? 1 : 0;
  op.val = v;
  c->edits.push_back(op);
}

// Apply queued edits on the device: ops are grouped per target word, order kept.
// Copies of 2 MB and more (16 MB for a batch member, so that hundreds of members do not each
// pin 16 MB) go through two pinned 8 MB buffers: the DMA of chunk k+1 runs while the host
// copies chunk k out, instead of the runtime's staging of pageable memory (about 5 GB/s
// here). Smaller copies go directly.
static constexpr size_t STAGE_CHUNK = size_t(8) << 20;
static size_t stage_min(const Ctx* c) { return c->batch ? size_t(16) << 20 : size_t(2) << 20; }
bool copy_is_direct(const Ctx* c, size_t bytes) { return bytes < stage_min(c); }
// A chunk's host copy on four threads: one thread copies (and first-touches fresh pages of a
// caller's buffer) at about 12 GB/s, below what the DMA delivers.
// If a thread cannot be started (std::system_error) its part is copied on the calling thread:
// no exception leaves the C ABI.
static void host_copy(char* d, const char* s, size_t len) {
  constexpr int T = 4;
  const size_t part = (len / T + 63) / 64 * 64;
  std::thread th[T - 1];
  for (int t = 1; t < T; t++) {
    const size_t a = std::min(len, t * part), b = std::min(len, (t + 1) * part);
    try {
      th[t - 1] = std::thread([=] { std::memcpy(d + a, s + a, b - a); });
    } catch (...) {
      std::memcpy(d + a, s + a, b - a);
    }
  }
  std::memcpy(d, s, std::min(len, part));
  for (auto& x : th)
    if (x.joinable()) x.join();
}
// On a failed call inside a staged copy, wait for the DMAs already queued on the stream before
// returning, so that none still reads or writes the pinned buffers the next copy refills.
#define FL_HIP_STAGED(expr)                                                             \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) {                                                             \
      ::fl::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));               \
      (void)hipStreamSynchronize(c->stream);                                            \
      return FL_ERR_HIP;                                                                \
    }                                                                                   \
  } while (0)
int copy_to_host(Ctx* c, void* dst, const void* src, size_t bytes) {
  if (bytes < stage_min(c)) {
    FL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    FL_HIP(hipStreamSynchronize(c->stream));
    return FL_OK;
  }
  if (!c->stage) FL_HIP_STAGED(hipHostMalloc((void**)&c->stage, 2 * STAGE_CHUNK, hipHostMallocDefault));
  for (auto& e : c->stage_ev)
    if (!e) FL_HIP_STAGED(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  const char* s = static_cast<const char*>(src);
  char* d = static_cast<char*>(dst);
  const size_t nch = (bytes + STAGE_CHUNK - 1) / STAGE_CHUNK;
  auto len = [&](size_t k) { return std::min(STAGE_CHUNK, bytes - k * STAGE_CHUNK); };
  FL_HIP_STAGED(hipMemcpyAsync(c->stage, s, len(0), hipMemcpyDeviceToHost, c->stream));
  FL_HIP_STAGED(hipEventRecord(c->stage_ev[0], c->stream));
  for (size_t k = 0; k < nch; k++) {
    if (k + 1 < nch) {  // its buffer was emptied by the host copy of chunk k-1
      char* b = c->stage + ((k + 1) & 1) * STAGE_CHUNK;
      FL_HIP_STAGED(hipMemcpyAsync(b, s + (k + 1) * STAGE_CHUNK, len(k + 1), hipMemcpyDeviceToHost, c->stream));
      FL_HIP_STAGED(hipEventRecord(c->stage_ev[(k + 1) & 1], c->stream));
    }
    FL_HIP_STAGED(hipEventSynchronize(c->stage_ev[k & 1]));
    host_copy(d + k * STAGE_CHUNK, c->stage + (k & 1) * STAGE_CHUNK, len(k));
  }
  return FL_OK;
}

// The same for host-to-device copies: the host fills one pinned buffer while the DMA of the
// other runs; the call returns when the data is on the device.
int copy_to_device(Ctx* c, void* dst, const void* src, size_t bytes) {
  if (bytes < stage_min(c)) {
    FL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    FL_HIP(hipStreamSynchronize(c->stream));
    return FL_OK;
  }
  if (!c->stage) FL_HIP_STAGED(hipHostMalloc((void**)&c->stage, 2 * STAGE_CHUNK, hipHostMallocDefault));
  for (auto& e : c->stage_ev)
    if (!e) FL_HIP_STAGED(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  const char* s = static_cast<const char*>(src);
  char* d = static_cast<char*>(dst);
  const size_t nch = (bytes + STAGE_CHUNK - 1) / STAGE_CHUNK;
  for (size_t k = 0; k < nch; k++) {
    const size_t len = std::min(STAGE_CHUNK, bytes - k * STAGE_CHUNK);
    char* b = c->stage + (k & 1) * STAGE_CHUNK;
    if (k >= 2) FL_HIP_STAGED(hipEventSynchronize(c->stage_ev[k & 1]));  // the DMA of chunk k-2 left b
    host_copy(b, s + k * STAGE_CHUNK, len);
    FL_HIP_STAGED(hipMemcpyAsync(d + k * STAGE_CHUNK, b, len, hipMemcpyHostToDevice, c->stream));
    FL_HIP_STAGED(hipEventRecord(c->stage_ev[k & 1], c->stream));
  }
  FL_HIP_STAGED(hipStreamSynchronize(c->stream));
  return FL_OK;
}

int flush_edits(Ctx* c) {
  if (c->edits.empty()) return FL_OK;
  const size_t n = c->edits.size();
  std::vector<uint32_t> order(n);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return (c->edits[a].key & 0x7fffffffu) < (c->edits[b].key & 0x7fffffffu);
  });
  std::vector<uint32_t> keys(n);
  std::vector<float> vals(n);
  std::vector<int32_t> seg;
  seg.reserve(n + 1);
  uint32_t prev = 0xffffffffu;
  for (size_t k = 0; k < n; k++) {
    const WordOp& op = c->edits[order[k]];
    keys[k] = op.key;
    vals[k] = op.val;
    if ((op.key & 0x7fffffffu) != prev) {
      seg.push_back((int32_t)k);
      prev = op.key & 0x7fffffffu;
    }
  }
  const int nseg = (int)seg.size();
  seg.push_back((int32_t)n);
  if (n > c->edit_cap) {
    if (c->d_edit_keys) FL_HIP(hipFree(c->d_edit_keys));
    if (c->d_edit_vals) FL_HIP(hipFree(c->d_edit_vals));
    if (c->d_edit_seg) FL_HIP(hipFree(c->d_edit_seg));
    size_t cap = std::max<size_t>(n, 1024);
    FL_HIP(hipMalloc(&c->d_edit_keys, 4 * cap));
    FL_HIP(hipMalloc(&c->d_edit_vals, 4 * cap));
    FL_HIP(hipMalloc(&c->d_edit_seg, 4 * (cap + 1)));
    c->edit_cap = cap;
  }
  FL_HIP(hipMemcpyAsync(c->d_edit_keys, keys.data(), 4 * n, hipMemcpyHostToDevice, c->stream));
  FL_HIP(hipMemcpyAsync(c->d_edit_vals, vals.data(), 4 * n, hipMemcpyHostToDevice, c->stream));
  FL_HIP(hipMemcpyAsync(c->d_edit_seg, seg.data(), 4 * (nseg + 1), hipMemcpyHostToDevice, c->stream));
  int rc = launch_apply_edits(c, nseg);
  if (rc) return rc;
  FL_HIP(hipStreamSynchronize(c->stream));  // host staging vectors die here
  c->edits.clear();
  return FL_OK;
}

int ctx_prepare(Ctx* c) {
  int rc = flush_edits(c);
  if (rc) return rc;
  if (c->flags_dirty) return launch_flags(c);
  return FL_OK;
}

// Per-step source ops of one context in the order the reference applies them: sources,
// sinks, then the jet (main.go:231-241,474-486). Cells are local to the context.
void source_ops(Ctx* c, std::vector<SrcOp>& ops) {
  auto inb = [&](int i, int j) { return i >= 0 && i < c->NX && j >= 0 && j < c->NY; };
  for (auto& s : c->sources) {
    if (!inb(s.I, s.J)) continue;
    int cell = s.I * c->NY + s.J;
    ops.push_back(SrcOp{cell, 0, 1, s.U, s.V});
    ops.push_back(SrcOp{cell, 1, 1, 0.5f, 0.0f});
  }
  for (auto& s : c->sinks) {
    if (!inb(s.I, s.J)) continue;
    ops.push_back(SrcOp{s.I * c->NY + s.J, 0, 1, s.U, s.V});
  }
  if (c->jet) {
    for (int j = c->jetLo; j < c->jetHi; j++) {
      int cell = c->jetX * c->NY + j;
      ops.push_back(SrcOp{cell, 0, 0, c->jetU, 0.0f});
      ops.push_back(SrcOp{cell, 1, 0, c->jetSmoke, 0.0f});
    }
  }
}

int ctx_build_sources(Ctx* c) {
  if (!c->src_dirty) return FL_OK;
  std::vector<SrcOp> ops;
  source_ops(c, ops);
  std::stable_sort(ops.begin(), ops.end(), [](const SrcOp& a, const SrcOp& b) { return a.cell < b.cell; });
  std::vector<int32_t> seg;
  for (size_t k = 0; k < ops.size(); k++)
    if (k == 0 || ops[k].cell != ops[k - 1].cell) seg.push_back((int32_t)k);
  const int nseg = (int)seg.size();
  seg.push_back((int32_t)ops.size());
  if (ops.size() > c->src_cap) {
    if (c->d_src) FL_HIP(hipFree(c->d_src));
    if (c->d_src_seg) FL_HIP(hipFree(c->d_src_seg));
    size_t cap = std::max<size_t>(ops.size(), 256);
    FL_HIP(hipMalloc(&c->d_src, sizeof(SrcOp) * cap));
    FL_HIP(hipMalloc(&c->d_src_seg, 4 * (cap + 1)));
    c->src_cap = cap;
  }
  if (!ops.empty()) {
    FL_HIP(hipMemcpyAsync(c->d_src, ops.data(), sizeof(SrcOp) * ops.size(), hipMemcpyHostToDevice, c->stream));
    FL_HIP(hipMemcpyAsync(c->d_src_seg, seg.data(), 4 * (nseg + 1), hipMemcpyHostToDevice, c->stream));
    FL_HIP(hipStreamSynchronize(c->stream));
  }
  c->n_src_ops = (int)ops.size();
  c->n_src_seg = nseg;
  c->src_dirty = false;
  return FL_OK;
}

// ---- the step ---------------------------------------------------------------------------
static int make_incompressible(Ctx* c, int iters, float dt, bool copy_border) {
  int rc;
  if (copy_border) {
    // copyBorder(newU,U), copyBorder(newV,V) (fluid.go:145-146). Inside Simulate the
    // velocity advection rewrites newU/newV's ring before anything reads it, so the
    // step skips these two copies.
    if ((rc = launch_copy_border(c, c->nU, c->U))) return rc;
    if ((rc = launch_copy_border(c, c->nV, c->V))) return rc;
  }
  if (c->prm.use_multigrid && c->prm.multigrid_levels > 1) return multigrid_solve(c, iters, dt);
  return pressure_solve(c, iters, dt);
}

static int advect_velocity_sl(Ctx* c, float dt) {
  int rc = launch_advect_velocity(c, c->U, c->V, c->U2, c->V2, dt);
  if (rc) return rc;
  std::swap(c->U, c->U2);
  std::swap(c->V, c->V2);
  return FL_OK;
}
static int advect_smoke_sl(Ctx* c, float dt) {
  int rc = launch_advect_smoke(c, c->M, c->M2, dt);
  if (rc) return rc;
  std::swap(c->M, c->M2);
  return FL_OK;
}
// advectVelocityBFECC (fluid.go:911-994): forward into U2/V2, backward + correction + clamp
// into T0/T1, final forward pass from the corrected field into U/V and newU/newV. The forward
// pass leaves newU/newV alone: the final pass writes every cell of them and reads them only
// at faces neither pass advects, whose newU/newV the forward pass would keep as they are.
static int advect_velocity_bfecc(Ctx* c, float dt) {
  int rc;
  if ((rc = launch_advect_velocity(c, c->U, c->V, c->U2, c->V2, dt, false))) return rc;
  if ((rc = launch_bfecc_velocity_corr(c, c->U, c->V, c->U2, c->V2, c->T0, c->T1, dt))) return rc;
  return launch_advect_velocity(c, c->T0, c->T1, c->U, c->V, dt);
}
// advectSmokeBFECC (fluid.go:997-1051)
static int advect_smoke_bfecc(Ctx* c, float dt) {
  int rc;
  if ((rc = launch_advect_smoke(c, c->M, c->M2, dt, false))) return rc;
  if ((rc = launch_bfecc_smoke_corr(c, c->M, c->M2, c->T2, dt))) return rc;
  return launch_advect_smoke(c, c->T2, c->M, dt);
}

static int mark(Ctx* c, int k) {
  if (c->timing) FL_HIP(hipEventRecord(c->ev[k], c->stream));
  return FL_OK;
}

static int simulate_passes(Ctx* c, float dt) {
  int rc;
  const fl_params& P = c->prm;
  // fill(p,0): a temporal-blocked solve starts from zero p without reading it and writes every
  // p cell (viscosity, in between, does not touch p), so the fill is only launched for the other
  // solves, or when there is no solve
  const bool mg = P.use_multigrid && P.multigrid_levels > 1;
  const int N = P.pressure_iters;
  const bool p_by_solve = N > 0 && !mg && (P.solver == 2 || (P.solver == 0 && gs_wave_supported(c, N)));
  if (!p_by_solve && (rc = launch_fill(c->p, 0.0f, c->N, c->stream))) return rc;         // fill(p,0)
  if (P.viscosity_diffusion > 0 && (rc = launch_viscosity(c, dt))) return rc;          // applyViscosity
  if ((rc = mark(c, 1))) return rc;
  c->gs_p_zero = p_by_solve;
  rc = make_incompressible(c, P.pressure_iters, dt, false);                              // makeIncompressible
  c->gs_p_zero = false;
  if (rc) return rc;
  if ((rc = mark(c, 2))) return rc;
  const bool conf = P.confinement != 0.0f, turb = P.turbulence_strength > 0.0f;
  if (conf && (rc = launch_curl(c, c->T0))) return rc;
  if ((conf || turb) && (rc = launch_conf_turb(c, c->T0, dt, conf, turb))) return rc;  // confinement, turbulence
  if ((rc = launch_borders(c))) return rc;                                             // handleBorders
  if ((rc = mark(c, 3))) return rc;
  if (P.use_bfecc) {
    if ((rc = advect_velocity_bfecc(c, dt))) return rc;
    return advect_smoke_bfecc(c, dt);
  }
  if ((rc = advect_velocity_sl(c, dt))) return rc;
  return advect_smoke_sl(c, dt);
}

// Simulate (fluid.go:79-109). Event 0 is recorded by the caller before any per-step
// source injection so that it is counted as part of the step.
static int simulate(Ctx* c, float dt, bool start_marked = false) {
  int rc = ctx_prepare(c);
  if (rc) return rc;
  if (c->NX < 3 || c->NY < 3) return err(FL_ERR_UNSUPPORTED, "grid must have at least 1x1 interior cells");
  if (!start_marked && (rc = mark(c, 0))) return rc;
  if ((rc = simulate_passes(c, dt))) return rc;
  if (c->timing) {
    if ((rc = mark(c, 4))) return rc;
    FL_HIP(hipEventSynchronize(c->ev[4]));
    float t[4];
    for (int k = 0; k < 4; k++) FL_HIP(hipEventElapsedTime(&t[k], c->ev[k], c->ev[k + 1]));
    c->t_other += t[0] + t[2];
    c->t_pressure += t[1];
    c->t_advect += t[3];
    c->t_total += t[0] + t[1] + t[2] + t[3];
    c->t_steps++;
  }
  return FL_OK;
}

void free_ctx(Ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->dev);
  if (!c->batch) {  // a batch member's fields, flags, control block and stream belong to the batch
    float** bufs[] = {&c->U, &c->V, &c->nU, &c->nV, &c->p, &c->S, &c->M, &c->nM,
                      &c->U2, &c->V2, &c->M2, &c->T0, &c->T1, &c->T2, &c->T3};
    for (auto** b : bufs)
      if (*b) (void)hipFree(*b);
    if (c->flags) (void)hipFree(c->flags);
    if (c->flags_diag) (void)hipFree(c->flags_diag);
    if (c->keep) (void)hipFree(c->keep);
    if (c->snap_hint) (void)hipHostFree(c->snap_hint);
    if (c->d_ctrl) (void)hipFree(c->d_ctrl);
    if (c->d_maxdiv) (void)hipFree(c->d_maxdiv);
  }
  for (float** b : {&c->d_omega, &c->d_edit_vals, &c->d_xy, &c->mg_omega, &c->mg_P, &c->mg_S, &c->mg_R})
    if (*b) (void)hipFree(*b);
  if (c->d_red) (void)hipFree(c->d_red);
  if (c->rs_mem) (void)hipFree(c->rs_mem);
  if (c->stage) (void)hipHostFree(c->stage);
  for (auto& e : c->stage_ev)
    if (e) (void)hipEventDestroy(e);
  if (c->d_edit_keys) (void)hipFree(c->d_edit_keys);
  if (c->d_edit_seg) (void)hipFree(c->d_edit_seg);
  if (c->d_src) (void)hipFree(c->d_src);
  if (c->d_src_seg) (void)hipFree(c->d_src_seg);
  pressure_free(c);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream && !c->batch) (void)hipStreamDestroy(c->stream);
  delete c;
}

fl_params default_params() {
  fl_params p{};
  p.confinement = 0.0f;
  p.viscosity_diffusion = 0.0f;
  p.pressure_damping = 1.0f;
  p.turbulence_strength = 0.02f;
  p.smoke_advection = 1.0f;
  p.relaxation = 1.9f;
  p.use_multigrid = 0;
  p.multigrid_levels = 2;
  p.use_bfecc = 0;
  p.pressure_iters = 8;
  p.early_exit = 1;
  p.solver = 0;
  p.block_sweeps = 0;
  p.precise_advection = 0;
  p.handoff_lag = 0;
  p.block_cols = 0;
  return p;
}

static int set_solid(Ctx* c, int i, int j, bool solid) {
  int rc = check_xy(c, i, j);
  if (rc) return rc;
  const size_t cell = (size_t)i * c->NY + j;
  const float sv = solid ? 0.0f : 1.0f;
  c->hS[cell] = sv;
  push_op(c, FL_FIELD_S, cell, false, sv);
  c->flags_dirty = true;
  if (solid) {
    const int n = c->NY;
    for (int f : {FL_FIELD_U, FL_FIELD_NEWU}) {
      push_op(c, f, cell, false, 0.0f);
      if (i + 1 < c->NX) push_op(c, f, cell + n, false, 0.0f);
    }
    for (int f : {FL_FIELD_V, FL_FIELD_NEWV}) {
      push_op(c, f, cell, false, 0.0f);
      if (j + 1 < c->NY) push_op(c, f, cell + 1, false, 0.0f);
    }
  }
  return FL_OK;
}

static void apply_force(Ctx* c, int i, int j, float fx, float fy) {
  if (i < 1 || i >= c->NX - 1 || j < 1 || j >= c->NY - 1) return;
  const size_t cell = (size_t)i * c->NY + j;
  if (c->hS[cell] == 0.0f) return;
  push_op(c, FL_FIELD_U, cell, true, fx);
  push_op(c, FL_FIELD_V, cell, true, fy);
}

static void apply_wall_settings(Ctx* c) {
  for (int i = 0; i < c->NX; i++) {
    set_solid(c, i, 0, c->wallBottom);
    set_solid(c, i, c->NY - 1, c->wallTop);
  }
  for (int j = 0; j < c->NY; j++) {
    set_solid(c, 0, j, c->wallLeft);
    set_solid(c, c->NX - 1, j, c->wallRight);
  }
}

static int reset(Ctx* c) {
  int rc = flush_edits(c);
  if (rc) return rc;
  for (float* b : {c->U, c->V, c->nU, c->nV, c->p, c->M, c->nM})
    if ((rc = launch_fill(b, 0.0f, c->N, c->stream))) return rc;
  return FL_OK;
}

static void set_circular_obstacle(Ctx* c, int cx, int cy, int r) {
  for (int i = cx - r; i <= cx + r; i++)
    for (int j = cy - r; j <= cy + r; j++) {
      if (i < 0 || i >= c->NX || j < 0 || j >= c->NY) continue;
      float dx = (float)(i - cx), dy = (float)(j - cy);
      if (dx * dx + dy * dy <= (float)(r * r)) set_solid(c, i, j, true);
    }
}

// fl_fill_cells: one thread per column j of the range, rows strided over blockIdx.y.
__global__ void k_fill_cells(float* __restrict__ f, int NY, int i0, int i1, int j0, int j1, float value) {
  const int j = j0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= j1) return;
  for (int i = i0 + blockIdx.y; i < i1; i += gridDim.y) f[(size_t)i * NY + j] = value;
}

}  // namespace fl

using namespace fl;

#define CTX_CHECK(c)                                                                                   \
  do {                                                                                                 \
    if (!(c)) return err(FL_ERR_INVALID_ARG, "null context");                                          \
    hipError_t e__ = hipSetDevice(reinterpret_cast<Ctx*>(c)->dev);                                      \
    if (e__ != hipSuccess) return err(FL_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e__)); \
  } while (0)
#define C_OF(x) reinterpret_cast<Ctx*>(x)

extern "C" {

const char* fl_last_error(void) { return g_err.c_str(); }

int fl_ctx_last_error(fl_ctx* ctx, char* buf, size_t n) {
  if (!ctx || (!buf && n)) return err(FL_ERR_INVALID_ARG, "null argument");
  if (n == 0) return FL_OK;
  const std::string& e = C_OF(ctx)->last_err;
  const size_t k = std::min(n - 1, e.size());
  std::memcpy(buf, e.data(), k);
  buf[k] = 0;
  return FL_OK;
}

int fl_device_count(int* n) {
  int k = 0;
  hipError_t e = hipGetDeviceCount(&k);
  if (e != hipSuccess) k = 0;
  if (n) *n = k;
  return FL_OK;
}

int fl_create(float density, int width, int height, float h, fl_ctx** out) {
  if (!out) return err(FL_ERR_INVALID_ARG, "null output pointer");
  *out = nullptr;
  if (width < 1 || height < 1) return err(FL_ERR_INVALID_ARG, "width and height must be >= 1");
  const int64_t cells = (int64_t)(width + 2) * (height + 2);
  if (cells > (int64_t)1 << 28) return err(FL_ERR_INVALID_ARG, "grid too large (more than 2^28 cells)");
  if (!(h > 0.0f)) return err(FL_ERR_INVALID_ARG, "h must be > 0");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return err(FL_ERR_NO_DEVICE, "no HIP device available: libfluidhip requires an MI355X (gfx950) GPU");
  Ctx* c = new Ctx();
  if (hipGetDevice(&c->dev) != hipSuccess) c->dev = 0;
  c->NX = width + 2;
  c->NY = height + 2;
  c->N = (size_t)c->NX * c->NY;
  c->density = density;
  c->h = h;
  c->prm = default_params();
  c->grid.NX = c->NX;
  c->grid.NY = c->NY;
  c->grid.h = h;
  c->grid.h1 = 1.0f / h;
  c->grid.hs = 0.5f * h;
  c->grid.ha = h / 2.0f;
  c->grid.xmax = (float)c->NX * h;
  c->grid.ymax = (float)c->NY * h;
  c->hS.assign(c->N, 0.0f);
  auto fail = [&](hipError_t e, const char* what) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    free_ctx(c);
    return (int)FL_ERR_HIP;
  };
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return fail(e, "hipStreamCreate");
  float** bufs[] = {&c->U, &c->V, &c->nU, &c->nV, &c->p, &c->S, &c->M, &c->nM,
                    &c->U2, &c->V2, &c->M2, &c->T0, &c->T1, &c->T2, &c->T3};
  for (auto** b : bufs) {
    if ((e = hipMalloc(b, sizeof(float) * c->N)) != hipSuccess) return fail(e, "hipMalloc field");
    if ((e = hipMemsetAsync(*b, 0, sizeof(float) * c->N, c->stream)) != hipSuccess) return fail(e, "hipMemset");
  }
  if ((e = hipMalloc(&c->flags, c->N)) != hipSuccess) return fail(e, "hipMalloc flags");
  if ((e = hipMalloc(&c->keep, sizeof(uint32_t) * c->NX * ((c->NY + 63) / 64))) != hipSuccess)
    return fail(e, "hipMalloc keep");
  if ((e = hipHostMalloc((void**)&c->snap_hint, 64, hipHostMallocCoherent)) != hipSuccess)
    return fail(e, "hipHostMalloc hint");
  *c->snap_hint = 0;
  if ((e = hipMalloc(&c->flags_diag, (size_t)(c->NX + c->NY - 1) * c->NX)) != hipSuccess)
    return fail(e, "hipMalloc flags_diag");
  // entries off the grid stay zero for the context's lifetime
  if ((e = hipMemsetAsync(c->flags_diag, 0, (size_t)(c->NX + c->NY - 1) * c->NX, c->stream)) != hipSuccess)
    return fail(e, "hipMemset flags_diag");
  if ((e = hipMalloc(&c->d_ctrl, 64)) != hipSuccess) return fail(e, "hipMalloc ctrl");
  if ((e = hipMemsetAsync(c->d_ctrl, 0, 64, c->stream)) != hipSuccess) return fail(e, "hipMemset ctrl");
  if ((e = hipMalloc(&c->d_red, 64)) != hipSuccess) return fail(e, "hipMalloc red");
  if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return fail(e, "hipStreamSynchronize");
  c->flags_dirty = true;
  *out = reinterpret_cast<fl_ctx*>(c);
  return FL_OK;
}

int fl_destroy(fl_ctx* ctx) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  if (!ctx) return FL_OK;
  if (C_OF(ctx)->batch) return err(FL_ERR_INVALID_ARG, "batch members are destroyed with their batch");
  free_ctx(C_OF(ctx));
  return FL_OK;
}

int fl_dims(fl_ctx* ctx, int* nx, int* ny) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  if (!ctx) return err(FL_ERR_INVALID_ARG, "null context");
  if (nx) *nx = C_OF(ctx)->NX;
  if (ny) *ny = C_OF(ctx)->NY;
  return FL_OK;
}
int fl_get_h(fl_ctx* ctx, float* h) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  if (!ctx || !h) return err(FL_ERR_INVALID_ARG, "null argument");
  *h = C_OF(ctx)->h;
  return FL_OK;
}
int fl_get_params(fl_ctx* ctx, fl_params* out) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  if (!ctx || !out) return err(FL_ERR_INVALID_ARG, "null argument");
  *out = C_OF(ctx)->prm;
  return FL_OK;
}
int fl_set_params(fl_ctx* ctx, const fl_params* in) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  if (!ctx || !in) return err(FL_ERR_INVALID_ARG, "null argument");
  if (in->pressure_iters < 0 || in->pressure_iters > 100000)
    return err(FL_ERR_INVALID_ARG, "pressure_iters out of range");
  if (in->solver < 0 || in->solver > 2) return err(FL_ERR_INVALID_ARG, "solver must be 0, 1 or 2");
  C_OF(ctx)->prm = *in;
  return FL_OK;
}

static float* field_ptr(Ctx* c, int f) {
  switch (f) {
    case FL_FIELD_U: return c->U;
    case FL_FIELD_V: return c->V;
    case FL_FIELD_NEWU: return c->nU;
    case FL_FIELD_NEWV: return c->nV;
    case FL_FIELD_P: return c->p;
    case FL_FIELD_S: return c->S;
    case FL_FIELD_M: return c->M;
    case FL_FIELD_NEWM: return c->nM;
  }
  return nullptr;
}

int fl_upload(fl_ctx* ctx, int field, const float* src, size_t n) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  CTX_CHECK(ctx);
  Ctx* c = C_OF(ctx);
  float* dst = field_ptr(c, field);
  if (!dst || !src) return err(FL_ERR_INVALID_ARG, "bad field or null source");
  if (n != c->N) return err(FL_ERR_INVALID_ARG, "size must be NumX*NumY");
  if (field == FL_FIELD_S) {
    for (size_t k = 0; k < n; k++) {
      uint32_t b;
      std::memcpy(&b, &src[k], 4);
      if (b != 0u && b != 0x3f800000u)
        return err(FL_ERR_INVALID_ARG, "S values must be exactly 0.0 (solid) or 1.0 (fluid)");
    }
  }
  int rc = flush_edits(c);
  if (rc) return rc;
  if ((rc = copy_to_device(c, dst, src, 4 * n))) return rc;
  if (field == FL_FIELD_S) {
    std::memcpy(c->hS.data(), src, 4 * n);
    c->flags_dirty = true;
  }
  return FL_OK;
}

int fl_fill_cells(fl_ctx* ctx, int field, int i0, int i1, int j0, int j1, float value) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  CTX_CHECK(ctx);
  Ctx* c = C_OF(ctx);
  float* dst = field_ptr(c, field);
  if (!dst || field == FL_FIELD_S) return err(FL_ERR_INVALID_ARG, "bad field (S changes through fl_set_solid)");
  if (i0 < 0 || j0 < 0 || i1 > c->NX || j1 > c->NY) return err(FL_ERR_INVALID_ARG, "cell range outside the grid");
  if (i0 >= i1 || j0 >= j1) return FL_OK;
  int rc = flush_edits(c);
  if (rc) return rc;
  const int w = j1 - j0, bx = w < 256 ? ((w + 63) / 64) * 64 : 256;
  hipLaunchKernelGGL(k_fill_cells, dim3((w + bx - 1) / bx, std::min(i1 - i0, 4096)), dim3(bx), 0, c->stream, dst,
                     c->NY, i0, i1, j0, j1, value);
  FL_HIP(hipGetLastError());
  return FL_OK;
}

int fl_download(fl_ctx* ctx, int field, float* dst, size_t n) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  CTX_CHECK(ctx);
  Ctx* c = C_OF(ctx);
  float* src = field_ptr(c, field);
  if (!src || !dst) return err(FL_ERR_INVALID_ARG, "bad field or null destination");
  if (n != c->N) return err(FL_ERR_INVALID_ARG, "size must be NumX*NumY");
  int rc = flush_edits(c);
  if (rc) return rc;
  return copy_to_host(c, dst, src, 4 * n);
}

int fl_simulate(fl_ctx* ctx, float dt) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  CTX_CHECK(ctx);
  return simulate(C_OF(ctx), dt);
}

int fl_sync(fl_ctx* ctx) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  CTX_CHECK(ctx);
  FL_HIP(hipStreamSynchronize(C_OF(ctx)->stream));
  return FL_OK;
}

int fl_get_stats(fl_ctx* ctx, fl_stats* out) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  CTX_CHECK(ctx);
  if (!out) return err(FL_ERR_INVALID_ARG, "null output");
  Ctx* c = C_OF(ctx);
  FL_HIP(hipStreamSynchronize(c->stream));
  int32_t ctrl[4] = {0, 0, 0, 0};
  FL_HIP(hipMemcpy(ctrl, c->d_ctrl, 16, hipMemcpyDeviceToHost));
  *out = fl_stats{};
  out->invalid_sample = ctrl[2] != 0;
  int n = c->last_n, kept = n;
  if (c->host_sweeps >= 0) {
    kept = c->host_sweeps;
  } else if (ctrl[1] > 0) {
    kept = ctrl[1];
  }
  out->sweeps_run = kept;
  out->reran = kept < n;
  if (kept > 0 && c->d_maxdiv) {
    uint32_t bits = 0;
    FL_HIP(hipMemcpy(&bits, c->d_maxdiv + (out->reran ? c->nmax : 0) + (kept - 1), 4, hipMemcpyDeviceToHost));
    std::memcpy(&out->max_div_last, &bits, 4);
  }
  if (ctrl[3] != 0) return err(FL_ERR_INTERNAL, "pressure kernel reported a dependency wait timeout");
  return FL_OK;
}

int fl_set_timing(fl_ctx* ctx, int on) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  CTX_CHECK(ctx);
  Ctx* c = C_OF(ctx);
  if (on && !c->ev[0])
    for (auto& e : c->ev) FL_HIP(hipEventCreate(&e));
  c->timing = on != 0;
  c->t_total = c->t_pressure = c->t_advect = c->t_other = 0;
  c->t_steps = 0;
  return FL_OK;
}

int fl_get_timing(fl_ctx* ctx, fl_timing* out) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  if (!ctx || !out) return err(FL_ERR_INVALID_ARG, "null argument");
  Ctx* c = C_OF(ctx);
  *out = fl_timing{};
  out->total_ms = c->t_total;
  out->pressure_ms = c->t_pressure;
  out->advect_ms = c->t_advect;
  out->other_ms = c->t_other;
  out->steps = c->t_steps;
  return FL_OK;
}

int fl_sweep_maxdiv(fl_ctx* ctx, float* out, int cap, int* count) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  CTX_CHECK(ctx);
  Ctx* c = C_OF(ctx);
  FL_HIP(hipStreamSynchronize(c->stream));
  int n = c->last_n;
  if (count) *count = n;
  if (out && n > 0 && c->d_maxdiv) {
    int m = std::min(cap, n);
    FL_HIP(hipMemcpy(out, c->d_maxdiv, 4 * m, hipMemcpyDeviceToHost));
  }
  return FL_OK;
}

int fl_get_gs_layout(fl_ctx* ctx, fl_gs_layout* out) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  CTX_CHECK(ctx);
  if (!out) return err(FL_ERR_INVALID_ARG, "null argument");
  const GsWork& W = C_OF(ctx)->gs;
  if (W.defer < 0) return err(FL_ERR_INVALID_ARG, "no temporal-blocked solve has run on this context");
  *out = fl_gs_layout{};
  out->sweeps_per_tile = W.tk;
  out->cols_per_block = W.blk;
  out->handoff_lag = W.defer + 1;
  out->row_tiles = W.G;
  out->sweep_blocks = W.KB;
  out->steps_per_tile = W.rows;
  out->tiles = W.ntiles;
  out->workgroups = W.grid;
  return FL_OK;
}

int fl_set_solid(fl_ctx* ctx, int i, int j, int solid) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  if (!ctx) return err(FL_ERR_INVALID_ARG, "null context");
  return set_solid(C_OF(ctx), i, j, solid != 0);
}
int fl_is_solid(fl_ctx* ctx, int i, int j, int* out) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  if (!ctx || !out) return err(FL_ERR_INVALID_ARG, "null argument");
  Ctx* c = C_OF(ctx);
  int rc = check_xy(c, i, j);
  if (rc) return rc;
  *out = c->hS[(size_t)i * c->NY + j] == 0.0f;
  return FL_OK;
}
int fl_set_velocity(fl_ctx* ctx, int i, int j, float u, float v) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  if (!ctx) return err(FL_ERR_INVALID_ARG, "null context");
  Ctx* c = C_OF(ctx);
  int rc = check_xy(c, i, j);
  if (rc) return rc;
  const size_t cell = (size_t)i * c->NY + j;
  push_op(c, FL_FIELD_U, cell, false, u);
  push_op(c, FL_FIELD_V, cell, false, v);
  return FL_OK;
}
int fl_add_smoke(fl_ctx* ctx, int i, int j, float s) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  if (!ctx) return err(FL_ERR_INVALID_ARG, "null context");
  Ctx* c = C_OF(ctx);
  int rc = check_xy(c, i, j);
  if (rc) return rc;
  push_op(c, FL_FIELD_M, (size_t)i * c->NY + j, true, s);
  return FL_OK;
}
int fl_reset(fl_ctx* ctx) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  CTX_CHECK(ctx);
  return reset(C_OF(ctx));
}
int fl_apply_force(fl_ctx* ctx, int i, int j, float fx, float fy) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  if (!ctx) return err(FL_ERR_INVALID_ARG, "null context");
  apply_force(C_OF(ctx), i, j, fx, fy);
  return FL_OK;
}
int fl_apply_force_radius(fl_ctx* ctx, int cx, int cy, float fx, float fy, int radius) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  if (!ctx) return err(FL_ERR_INVALID_ARG, "null context");
  Ctx* c = C_OF(ctx);
  if (radius <= 0) {
    apply_force(c, cx, cy, fx, fy);
    return FL_OK;
  }
  const float r2 = (float)(radius * radius);
  for (int i = cx - radius; i <= cx + radius; i++)
    for (int j = cy - radius; j <= cy + radius; j++) {
      if (i < 1 || i >= c->NX - 1 || j < 1 || j >= c->NY - 1) continue;
      float dx = (float)(i - cx), dy = (float)(j - cy);
      float d2 = dx * dx + dy * dy;
      if (d2 > r2) continue;
      float w = (float)std::exp((double)(-3.0f * d2 / r2));
      apply_force(c, i, j, fx * w, fy * w);
    }
  return FL_OK;
}
int fl_apply_force_gaussian(fl_ctx* ctx, int cx, int cy, float fx, float fy, float sigma) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  if (!ctx) return err(FL_ERR_INVALID_ARG, "null context");
  if (!(sigma > 0.0f) || sigma > 1e6f) return err(FL_ERR_INVALID_ARG, "sigma must be in (0, 1e6]");
  Ctx* c = C_OF(ctx);
  const float two_s2 = 2.0f * sigma * sigma, cut2 = 9.0f * sigma * sigma;
  const int R = (int)std::ceil(3.0f * sigma);
  const int ilo = std::max(cx - R, 1), ihi = std::min(cx + R, c->NX - 2);
  const int jlo = std::max(cy - R, 1), jhi = std::min(cy + R, c->NY - 2);
  for (int i = ilo; i <= ihi; i++)
    for (int j = jlo; j <= jhi; j++) {
      float dx = (float)(i - cx), dy = (float)(j - cy);
      float d2 = dx * dx + dy * dy;
      if (d2 > cut2) continue;
      float w = (float)std::exp(-(double)(d2 / two_s2));
      apply_force(c, i, j, fx * w, fy * w);
    }
  return FL_OK;
}
int fl_set_circular_obstacle(fl_ctx* ctx, int cx, int cy, int radius) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  if (!ctx) return err(FL_ERR_INVALID_ARG, "null context");
  set_circular_obstacle(C_OF(ctx), cx, cy, radius);
  return FL_OK;
}

int fl_get_adaptive_time_step(fl_ctx* ctx, float basedt, float* out) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  CTX_CHECK(ctx);
  if (!out) return err(FL_ERR_INVALID_ARG, "null output");
  Ctx* c = C_OF(ctx);
  int rc = ctx_prepare(c);
  if (rc) return rc;
  float maxVel = 0.0f;
  if ((rc = launch_max_reduce(c, 1, &maxVel))) return rc;
  if (maxVel == 0.0f) {
    *out = basedt;
    return FL_OK;
  }
  float adt = 0.8f * c->h / maxVel;
  adt = go_maxf_h(go_minf_h(adt, basedt * 2.0f), basedt * 0.1f);
  *out = adt;
  return FL_OK;
}
int fl_max_divergence(fl_ctx* ctx, float* out) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  CTX_CHECK(ctx);
  if (!out) return err(FL_ERR_INVALID_ARG, "null output");
  Ctx* c = C_OF(ctx);
  int rc = ctx_prepare(c);
  if (rc) return rc;
  return launch_max_reduce(c, 0, out);
}

int fl_sample_velocity_batch(fl_ctx* ctx, const float* xs, const float* ys, int n, float* us, float* vs) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  CTX_CHECK(ctx);
  if (n < 0 || (n > 0 && (!xs || !ys || !us || !vs))) return err(FL_ERR_INVALID_ARG, "bad arguments");
  if (n == 0) return FL_OK;
  Ctx* c = C_OF(ctx);
  int rc = ctx_prepare(c);
  if (rc) return rc;
  if (n > c->xy_cap) {
    if (c->d_xy) FL_HIP(hipFree(c->d_xy));
    FL_HIP(hipMalloc(&c->d_xy, sizeof(float) * 4 * n));
    c->xy_cap = n;
  }
  std::vector<float> xy(2 * (size_t)n), uv(2 * (size_t)n);
  for (int k = 0; k < n; k++) {
    xy[2 * k] = xs[k];
    xy[2 * k + 1] = ys[k];
  }
  if ((rc = copy_to_device(c, c->d_xy, xy.data(), 8 * (size_t)n))) return rc;
  if ((rc = launch_sample_batch(c, c->d_xy, n, c->d_xy + 2 * (size_t)n))) return rc;
  if ((rc = copy_to_host(c, uv.data(), c->d_xy + 2 * (size_t)n, 8 * (size_t)n))) return rc;
  for (int k = 0; k < n; k++) {
    us[k] = uv[2 * k];
    vs[k] = uv[2 * k + 1];
  }
  return FL_OK;
}
int fl_sample_velocity(fl_ctx* ctx, float x, float y, float* u, float* v) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  return fl_sample_velocity_batch(ctx, &x, &y, 1, u, v);
}

// Vorticity()/VelocityMagnitude(): field on the device, MinValue/MaxValue scanned in
// the reference's loop order over interior fluid cells (fluid.go:810-828, 844-863).
static int view(fl_ctx* ctx, int kind, float* out, size_t n, float* mn, float* mx) {
  CTX_CHECK(ctx);
  Ctx* c = C_OF(ctx);
  if (!out || n != c->N) return err(FL_ERR_INVALID_ARG, "output must hold NumX*NumY floats");
  int rc = ctx_prepare(c);
  if (rc) return rc;
  if ((rc = launch_view(c, kind, c->T3))) return rc;
  // MinValue/MaxValue over the fluid interior cells (fluid.go:806-872), on the device
  if ((rc = field_range(c, c->T3, 1, mn, mx))) return rc;
  return copy_to_host(c, out, c->T3, 4 * n);
}
int fl_vorticity(fl_ctx* ctx, float* out, size_t n, float* mn, float* mx) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx)); return view(ctx, 0, out, n, mn, mx); }
int fl_velocity_magnitude(fl_ctx* ctx, float* out, size_t n, float* mn, float* mx) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  return view(ctx, 1, out, n, mn, mx);
}
int fl_pass(fl_ctx* ctx, int pass, float dt, int iters) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  CTX_CHECK(ctx);
  Ctx* c = C_OF(ctx);
  int rc = ctx_prepare(c);
  if (rc) return rc;
  switch (pass) {
    case FL_PASS_COPY_BORDER_UV:
      if ((rc = launch_copy_border(c, c->nU, c->U))) return rc;
      return launch_copy_border(c, c->nV, c->V);
    case FL_PASS_VISCOSITY:
      if (c->prm.viscosity_diffusion <= 0) return FL_OK;
      return launch_viscosity(c, dt);
    case FL_PASS_MAKE_INCOMPRESSIBLE: return make_incompressible(c, iters, dt, true);
    case FL_PASS_CONFINEMENT:
      if ((rc = launch_curl(c, c->T0))) return rc;
      return launch_conf_turb(c, c->T0, dt, 1, 0);
    case FL_PASS_TURBULENCE:
      if (c->prm.turbulence_strength <= 0) return FL_OK;
      return launch_conf_turb(c, c->T0, dt, 0, 1);
    case FL_PASS_BORDERS: return launch_borders(c);
    case FL_PASS_ADVECT_VELOCITY: return advect_velocity_sl(c, dt);
    case FL_PASS_ADVECT_SMOKE: return advect_smoke_sl(c, dt);
    case FL_PASS_ADVECT_VELOCITY_BFECC: return advect_velocity_bfecc(c, dt);
    case FL_PASS_ADVECT_SMOKE_BFECC: return advect_smoke_bfecc(c, dt);
    case FL_PASS_CLEAR_P: return launch_fill(c->p, 0.0f, c->N, c->stream);
  }
  return err(FL_ERR_INVALID_ARG, "unknown pass");
}

// ---- scene layer (package main) ------------------------------------------------------------
int fl_scene_set_walls(fl_ctx* ctx, int top, int bottom, int left, int right) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  if (!ctx) return err(FL_ERR_INVALID_ARG, "null context");
  Ctx* c = C_OF(ctx);
  c->wallTop = top;
  c->wallBottom = bottom;
  c->wallLeft = left;
  c->wallRight = right;
  return FL_OK;
}
int fl_scene_apply_walls(fl_ctx* ctx) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  if (!ctx) return err(FL_ERR_INVALID_ARG, "null context");
  apply_wall_settings(C_OF(ctx));
  return FL_OK;
}
int fl_scene_clear_walls(fl_ctx* ctx) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  if (!ctx) return err(FL_ERR_INVALID_ARG, "null context");
  Ctx* c = C_OF(ctx);
  for (int i = 0; i < c->NX; i++)
    for (int j = 0; j < c->NY; j++) set_solid(c, i, j, false);
  apply_wall_settings(c);
  return FL_OK;
}
int fl_scene_preset(fl_ctx* ctx, int preset, float lid_u) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  CTX_CHECK(ctx);
  Ctx* c = C_OF(ctx);
  if (preset < 0 || preset > 2) return err(FL_ERR_INVALID_ARG, "preset must be 0 (jet), 1 (cavity) or 2 (karman)");
  int rc = reset(c);
  if (rc) return rc;
  c->sources.clear();
  c->sinks.clear();
  c->src_dirty = true;
  c->src_gen++;
  for (int i = 0; i < c->NX; i++)
    for (int j = 0; j < c->NY; j++) set_solid(c, i, j, false);
  if (preset == 0) {
    c->wallTop = c->wallBottom = c->wallLeft = true;
    c->wallRight = false;
  } else if (preset == 1) {
    c->wallTop = c->wallBottom = c->wallLeft = c->wallRight = true;
    c->jet = false;
    for (int i = 2; i < c->NX - 2; i++) c->sources.push_back(Source{i, c->NY - 2, lid_u, 0.0f});
  } else {
    c->wallTop = c->wallBottom = c->wallLeft = true;
    c->wallRight = false;
    set_circular_obstacle(c, c->NX / 4, c->NY / 2, 8);
  }
  apply_wall_settings(c);
  return FL_OK;
}
int fl_scene_set_jet(fl_ctx* ctx, int on, int x, int j_lo, int j_hi, float u, float smoke) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  if (!ctx) return err(FL_ERR_INVALID_ARG, "null context");
  Ctx* c = C_OF(ctx);
  if (on) {
    int rc = check_xy(c, x, 0);
    if (rc) return rc;
    if (j_lo < j_hi && ((rc = check_xy(c, x, j_lo)) || (rc = check_xy(c, x, j_hi - 1)))) return rc;
  }
  c->jet = on != 0;
  c->jetX = x;
  c->jetLo = j_lo;
  c->jetHi = j_hi;
  c->jetU = u;
  c->jetSmoke = smoke;
  c->src_dirty = true;
  c->src_gen++;
  return FL_OK;
}
int fl_scene_add_source(fl_ctx* ctx, int sink, int i, int j, float u, float v) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  if (!ctx) return err(FL_ERR_INVALID_ARG, "null context");
  Ctx* c = C_OF(ctx);
  (sink ? c->sinks : c->sources).push_back(Source{i, j, u, v});
  c->src_dirty = true;
  c->src_gen++;
  return FL_OK;
}
int fl_scene_clear_sources(fl_ctx* ctx) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  if (!ctx) return err(FL_ERR_INVALID_ARG, "null context");
  Ctx* c = C_OF(ctx);
  c->sources.clear();
  c->sinks.clear();
  c->src_dirty = true;
  c->src_gen++;
  return FL_OK;
}
int fl_scene_apply_sources(fl_ctx* ctx) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  if (!ctx) return err(FL_ERR_INVALID_ARG, "null context");
  Ctx* c = C_OF(ctx);
  auto inb = [&](int i, int j) { return i >= 0 && i < c->NX && j >= 0 && j < c->NY; };
  for (auto& s : c->sources)
    if (inb(s.I, s.J) && c->hS[(size_t)s.I * c->NY + s.J] != 0.0f) {
      fl_set_velocity(ctx, s.I, s.J, s.U, s.V);
      fl_add_smoke(ctx, s.I, s.J, 0.5f);
    }
  for (auto& s : c->sinks)
    if (inb(s.I, s.J) && c->hS[(size_t)s.I * c->NY + s.J] != 0.0f) fl_set_velocity(ctx, s.I, s.J, s.U, s.V);
  return FL_OK;
}
int fl_scene_update(fl_ctx* ctx, float dt) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  CTX_CHECK(ctx);
  Ctx* c = C_OF(ctx);
  int rc = flush_edits(c);
  if (rc) return rc;
  if ((rc = ctx_build_sources(c))) return rc;
  if ((rc = mark(c, 0))) return rc;
  if ((rc = launch_apply_sources(c))) return rc;
  return simulate(c, dt, true);
}
int fl_scene_draw_line(fl_ctx* ctx, int x0, int y0, int x1, int y1, int solid) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  if (!ctx) return err(FL_ERR_INVALID_ARG, "null context");
  Ctx* c = C_OF(ctx);
  int dx = std::abs(x1 - x0), dy = -std::abs(y1 - y0);
  int sx = x0 < x1 ? 1 : -1, sy = y0 < y1 ? 1 : -1;
  int e = dx + dy;
  for (;;) {
    if (x0 >= 0 && x0 < c->NX && y0 >= 0 && y0 < c->NY) set_solid(c, x0, y0, solid != 0);
    if (x0 == x1 && y0 == y1) break;
    int e2 = 2 * e;
    if (e2 >= dy) {
      e += dy;
      x0 += sx;
    }
    if (e2 <= dx) {
      e += dx;
      y0 += sy;
    }
  }
  return FL_OK;
}
int fl_scene_smoke_brush(fl_ctx* ctx, int ci, int cj, int r, float strength) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  if (!ctx) return err(FL_ERR_INVALID_ARG, "null context");
  Ctx* c = C_OF(ctx);
  for (int i = ci - r; i <= ci + r; i++)
    for (int j = cj - r; j <= cj + r; j++) {
      if (i < 0 || i >= c->NX || j < 0 || j >= c->NY) continue;
      float dx = (float)(i - ci), dy = (float)(j - cj);
      if (dx * dx + dy * dy > (float)(r * r)) continue;
      fl_add_smoke(ctx, i, j, strength * 0.3f);
    }
  return FL_OK;
}

}  // extern "C"
