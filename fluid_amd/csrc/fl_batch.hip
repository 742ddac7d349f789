// Batched ensembles: many independent simulations of uneven grid sizes stepped together.
//
// A batch carves every field of its members out of shared pools. Each member is a full
// context (fl_ctx) over its slice, so the whole single-simulation API (edits, presets,
// sources, views and its own Simulate) works per member. fl_batch_update advances all
// members with one launch per pass: cell kernels take (member, 64x4 block) items from a
// work-list, and the pressure solve is one persistent launch over the merged tile list of
// every member (fl_gs_wave.hip). A batched step copies results back instead of swapping
// buffers, so the member descriptors stay fixed and the step can be captured once as a
// HIP graph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "fl_internal.h"

namespace fl {

static constexpr int MEMBER_NMAX = 1024;  // sweeps a member supports (size of its max|div| record)

static int berr(int code, const std::string& m) {
  set_error(m);
  return code;
}

static void batch_free(Batch* B) {
  if (!B) return;
  (void)hipSetDevice(B->dev);
  if (B->graph) (void)hipGraphExecDestroy(B->graph);
  for (Ctx* c : B->members) free_ctx(c);
  for (float* p : B->pool)
    if (p) (void)hipFree(p);
  for (void* p : {(void*)B->flags_pool, (void*)B->diag_pool, (void*)B->keep_pool, (void*)B->ctrl_pool, (void*)B->maxdiv_pool,
                  (void*)B->d_sims, (void*)B->d_cells, (void*)B->d_rings, (void*)B->d_tiles, (void*)B->kout,
                  (void*)B->gout, (void*)B->prog, (void*)B->ticket, (void*)B->d_omega, (void*)B->d_src,
                  (void*)B->d_src_seg, (void*)B->d_src_sim, (void*)B->d_ekeys, (void*)B->d_evals,
                  (void*)B->d_eseg, (void*)B->d_esim})
    if (p) (void)hipFree(p);
  if (B->stream) (void)hipStreamDestroy(B->stream);
  delete B;
}

// ---- per-step preparation (host side, outside any graph) -----------------------------------
static int upload_omega(Batch* B, int N) {
  if (B->omega_n == N && B->omega_relax == B->prm.relaxation) return FL_OK;
  FL_HIP(hipStreamSynchronize(B->stream));  // earlier steps may still read the table
  std::vector<float> om(std::max(N, 1));
  const float init = B->prm.relaxation, minR = 1.2f;
  for (int k = 0; k < N; k++) om[k] = init - (init - minR) * ((float)k / (float)N);
  FL_HIP(hipMemcpy(B->d_omega, om.data(), sizeof(float) * N, hipMemcpyHostToDevice));
  B->omega_n = N;
  B->omega_relax = B->prm.relaxation;
  return FL_OK;
}

static int build_batch_sources(Batch* B) {
  bool dirty = false;
  for (size_t e = 0; e < B->members.size(); e++)
    if (B->members[e]->src_gen != B->src_gen_seen[e]) dirty = true;
  if (!dirty) return FL_OK;
  FL_HIP(hipStreamSynchronize(B->stream));  // earlier steps may still read the program
  std::vector<SrcOp> all;
  std::vector<int32_t> seg, segsim;
  for (size_t e = 0; e < B->members.size(); e++) {
    std::vector<SrcOp> ops;
    source_ops(B->members[e], ops);
    std::stable_sort(ops.begin(), ops.end(), [](const SrcOp& a, const SrcOp& b) { return a.cell < b.cell; });
    for (size_t k = 0; k < ops.size(); k++) {
      if (k == 0 || ops[k].cell != ops[k - 1].cell) {
        seg.push_back((int32_t)all.size());
        segsim.push_back((int32_t)e);
      }
      all.push_back(ops[k]);
    }
    B->src_gen_seen[e] = B->members[e]->src_gen;
  }
  const int nseg = (int)seg.size();
  seg.push_back((int32_t)all.size());
  if (all.size() + 1 > B->src_cap || (size_t)nseg + 1 > B->src_cap) {
    for (void* p : {(void*)B->d_src, (void*)B->d_src_seg, (void*)B->d_src_sim})
      if (p) FL_HIP(hipFree(p));
    const size_t cap = std::max<size_t>(all.size() + 1, 1024);
    FL_HIP(hipMalloc(&B->d_src, sizeof(SrcOp) * cap));
    FL_HIP(hipMalloc(&B->d_src_seg, sizeof(int32_t) * (cap + 1)));
    FL_HIP(hipMalloc(&B->d_src_sim, sizeof(int32_t) * (cap + 1)));
    B->src_cap = cap;
  }
  if (!all.empty()) {
    FL_HIP(hipMemcpy(B->d_src, all.data(), sizeof(SrcOp) * all.size(), hipMemcpyHostToDevice));
    FL_HIP(hipMemcpy(B->d_src_seg, seg.data(), sizeof(int32_t) * seg.size(), hipMemcpyHostToDevice));
    FL_HIP(hipMemcpy(B->d_src_sim, segsim.data(), sizeof(int32_t) * segsim.size(), hipMemcpyHostToDevice));
  }
  B->n_src_seg = nseg;
  B->graph_gen++;
  return FL_OK;
}

// The members' queued edits (SetSolid, SetVelocity, AddSmoke, ApplyForce...) in one
// launch: per member the ops are grouped by target word in call order, as flush_edits
// does for one context.
static int flush_batch_edits(Batch* B) {
  size_t total = 0;
  for (Ctx* c : B->members) total += c->edits.size();
  if (total == 0) return FL_OK;
  std::vector<uint32_t> keys;
  std::vector<float> vals;
  std::vector<int32_t> seg, segsim;
  keys.reserve(total);
  vals.reserve(total);
  for (size_t e = 0; e < B->members.size(); e++) {
    Ctx* c = B->members[e];
    if (c->edits.empty()) continue;
    std::vector<uint32_t> order(c->edits.size());
    for (size_t k = 0; k < order.size(); k++) order[k] = (uint32_t)k;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
      return (c->edits[a].key & 0x7fffffffu) < (c->edits[b].key & 0x7fffffffu);
    });
    uint32_t prev = 0xffffffffu;
    for (uint32_t k : order) {
      const WordOp& op = c->edits[k];
      if ((op.key & 0x7fffffffu) != prev) {
        seg.push_back((int32_t)keys.size());
        segsim.push_back((int32_t)e);
        prev = op.key & 0x7fffffffu;
      }
      keys.push_back(op.key);
      vals.push_back(op.val);
    }
    c->edits.clear();
  }
  const int nseg = (int)seg.size();
  seg.push_back((int32_t)keys.size());
  FL_HIP(hipStreamSynchronize(B->stream));  // the staging buffers may still be read
  if (total + 1 > B->edit_cap) {
    for (void* p : {(void*)B->d_ekeys, (void*)B->d_evals, (void*)B->d_eseg, (void*)B->d_esim})
      if (p) FL_HIP(hipFree(p));
    const size_t cap = std::max<size_t>(total + 1, 4096);
    FL_HIP(hipMalloc(&B->d_ekeys, sizeof(uint32_t) * cap));
    FL_HIP(hipMalloc(&B->d_evals, sizeof(float) * cap));
    FL_HIP(hipMalloc(&B->d_eseg, sizeof(int32_t) * (cap + 1)));
    FL_HIP(hipMalloc(&B->d_esim, sizeof(int32_t) * (cap + 1)));
    B->edit_cap = cap;
  }
  FL_HIP(hipMemcpy(B->d_ekeys, keys.data(), sizeof(uint32_t) * keys.size(), hipMemcpyHostToDevice));
  FL_HIP(hipMemcpy(B->d_evals, vals.data(), sizeof(float) * vals.size(), hipMemcpyHostToDevice));
  FL_HIP(hipMemcpy(B->d_eseg, seg.data(), sizeof(int32_t) * seg.size(), hipMemcpyHostToDevice));
  FL_HIP(hipMemcpy(B->d_esim, segsim.data(), sizeof(int32_t) * segsim.size(), hipMemcpyHostToDevice));
  return batch_apply_edits(B, B->d_ekeys, B->d_evals, B->d_eseg, B->d_esim, nseg);
}

static int prepare_batch(Batch* B, float dt) {
  int rc;
  const int N = B->prm.pressure_iters;
  if (N > MEMBER_NMAX) return berr(FL_ERR_UNSUPPORTED, "batched solves support at most 1024 sweeps");
  if ((rc = upload_omega(B, N))) return rc;
  // descriptors of all members with this step's cp and the GS layout for N sweeps
  std::vector<SimDesc> d(B->members.size());
  for (size_t e = 0; e < B->members.size(); e++) {
    Ctx* c = B->members[e];
    d[e] = c->desc();
    d[e].cp = c->density * c->h / dt;
    d[e].early_exit = B->prm.early_exit;
    d[e].g.precise = B->prm.precise_advection != 0;
    d[e].nmax = MEMBER_NMAX;
  }
  // the batch's own tile layout replaces whatever GS workspace a member holds from
  // stepping on its own
  const int old_n = B->gs_n, old_tk = B->tk, old_blk = B->blk, old_grid = B->grid;
  if (N > 0 && (rc = batch_gs_layout(B, N, d))) return rc;
  if (N > 0 && (B->gs_n != old_n || B->tk != old_tk || B->blk != old_blk || B->grid != old_grid)) B->graph_gen++;
  if (B->descs.size() != d.size() || std::memcmp(B->descs.data(), d.data(), sizeof(SimDesc) * d.size()) != 0) {
    FL_HIP(hipStreamSynchronize(B->stream));  // earlier steps may still read the descriptors
    FL_HIP(hipMemcpy(B->d_sims, d.data(), sizeof(SimDesc) * d.size(), hipMemcpyHostToDevice));
    B->descs = d;
  }
  if ((rc = flush_batch_edits(B))) return rc;
  bool any_flags = false;
  for (Ctx* c : B->members) any_flags |= c->flags_dirty;
  if (any_flags) {
    if ((rc = batch_flags(B))) return rc;
    for (Ctx* c : B->members) c->flags_dirty = false;
  }
  return build_batch_sources(B);
}

// Simulate (fluid.go:79-109) of every member, preceded by each member's per-step sources
// and jet (main.go:231-241); the results end in the members' own buffers.
static int enqueue_step(Batch* B, float dt) {
  int rc;
  const fl_params& P = B->prm;
  if ((rc = batch_apply_sources(B))) return rc;
  // fill(p, 0): the solve starts from p = 0 without reading it and writes p itself
  if (P.pressure_iters <= 0 && (rc = batch_fill_role(B, B_P, 0.0f))) return rc;
  if (P.viscosity_diffusion > 0) {
    if ((rc = batch_viscosity(B, dt))) return rc;
    if ((rc = batch_copy_role(B, B_U, B_U2)) || (rc = batch_copy_role(B, B_V, B_V2))) return rc;
  }
  // The velocity after the solve stays in (U2, V2): confinement, turbulence and borders work
  // there in place and the (first) velocity advection writes the step's result into (U, V).
  // Semi-Lagrangian smoke writes M2, which then trades roles with M (no copy).
  int cu = B_U, cv = B_V;
  if (P.pressure_iters > 0) {
    if ((rc = batch_gs_solve(B, P.pressure_iters, -1, B_P))) return rc;
    cu = B_U2;
    cv = B_V2;
  }
  if ((rc = batch_conf_turb(B, dt, cu, cv))) return rc;
  if ((rc = batch_borders(B, cu, cv))) return rc;
  // forward (or only) pass: from the current velocity into the other pair of roles
  const int fu = cu == B_U ? B_U2 : B_U, fv = cv == B_V ? B_V2 : B_V;
  if (P.use_bfecc) {
    if ((rc = batch_advect_velocity(B, cu, cv, fu, fv, dt, false))) return rc;
    if ((rc = batch_bfecc_velocity_corr(B, cu, cv, fu, fv, dt))) return rc;
    if ((rc = batch_advect_velocity(B, B_T0, B_T1, B_U, B_V, dt))) return rc;
    if ((rc = batch_advect_smoke(B, B_M, B_M2, dt, false))) return rc;
    if ((rc = batch_bfecc_smoke_corr(B, dt))) return rc;
    return batch_advect_smoke(B, B_T2, B_M, dt);
  }
  if ((rc = batch_advect_velocity(B, cu, cv, fu, fv, dt))) return rc;
  if (fu != B_U && ((rc = batch_copy_role(B, B_U, fu)) || (rc = batch_copy_role(B, B_V, fv)))) return rc;
  if ((rc = batch_advect_smoke(B, B_M, B_M2, dt))) return rc;
  return batch_swap_roles(B, B_M, B_M2);  // the host side follows in swap_member_roles
}

// The host half of the step's role swaps: what k_swap_roles did to the device descriptors,
// applied to the members' pointers and to the copy prepare_batch compares against, so the
// next step's descriptors match the device's without an upload.
static void swap_member_roles(Batch* B) {
  if (B->prm.use_bfecc) return;  // the BFECC smoke's last pass writes M itself
  for (size_t e = 0; e < B->members.size(); e++) {
    Ctx* c = B->members[e];
    std::swap(c->M, c->M2);
    if (e < B->descs.size()) std::swap(B->descs[e].buf[B_M], B->descs[e].buf[B_M2]);
  }
}

// UseMultigrid (fluid.go:148-151): the V-cycle decides on the host between cycles whether to go
// on, which neither the merged persistent launch nor graph capture allows, so each member takes
// its own update (its sources, then Simulate, on the batch's stream) with the batch's knobs. The
// members' buffers trade places as in a member's own step; the next batched step reads their
// new descriptors.
static int batch_update_members(Batch* B, float dt) {
  for (Ctx* c : B->members) {
    const fl_params saved = c->prm;
    c->prm = B->prm;
    const int rc = fl_scene_update(reinterpret_cast<fl_ctx*>(c), dt);
    c->prm = saved;
    if (rc) return rc;
  }
  B->steps++;
  return FL_OK;
}

static int batch_update(Batch* B, float dt) {
  if (B->prm.use_multigrid && B->prm.multigrid_levels > 1) return batch_update_members(B, dt);
  int rc = prepare_batch(B, dt);
  if (rc) return rc;
  if (B->use_graph) {
    const bool same = B->graph && B->graph_dt == dt && B->graph_key_gen == B->graph_gen &&
                      std::memcmp(&B->graph_prm, &B->prm, sizeof(fl_params)) == 0;
    if (!same) {
      if (B->graph) {
        FL_HIP(hipGraphExecDestroy(B->graph));
        B->graph = nullptr;
      }
      hipGraph_t g;
      FL_HIP(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
      rc = enqueue_step(B, dt);
      hipError_t e = hipStreamEndCapture(B->stream, &g);
      if (rc) return rc;
      if (e != hipSuccess) return berr(FL_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
      e = hipGraphInstantiate(&B->graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) return berr(FL_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      B->graph_dt = dt;
      B->graph_key_gen = B->graph_gen;
      B->graph_prm = B->prm;
    }
    FL_HIP(hipGraphLaunch(B->graph, B->stream));
  } else {
    if ((rc = enqueue_step(B, dt))) return rc;
  }
  swap_member_roles(B);
  for (Ctx* c : B->members) {
    c->last_n = B->prm.pressure_iters;
    c->host_sweeps = -1;
  }
  B->steps++;
  return FL_OK;
}

// ---- autotuning of the Gauss-Seidel tile depth ---------------------------------------------
// Tile shapes the autotuners time: sweeps per tile (2, 4, 8, 16), columns per block (4, 8,
// 16) and handoff publication (1 on time, 2 a block late, 3 halfway), ordered so that the
// layout changes only with the first two.
// Shallow tiles multiply the handoff traffic and the tile count, so 2-deep tiles are tried
// only where the 8-deep tiling has at most two tiles per CU (the critical path, not the CUs'
// throughput, bounds such solves).
struct TuneShape {
  int tk, blk, lag;
};
static std::vector<TuneShape> tune_shapes(long long tiles8, int cus) {
  std::vector<TuneShape> v;
  for (int tk : {2, 4, 8, 16}) {
    if (tk == 2 && tiles8 > 2LL * cus) continue;
    for (int blk : {4, 8, 16})
      for (int lag : {1, 2, 3}) v.push_back(TuneShape{tk, blk, lag});
  }
  return v;
}
static long long tiles_at(int NX, int N, int tk) { return (long long)((NX + N - 1 + 63) / 64) * ((N + tk - 1) / tk); }
static int device_cus(int dev) {
  int n = 0;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
  return n;
}

// The solve reads U, V, p and writes only scratch outputs (U2, V2, T3) and its statistics,
// so candidates are timed on the live state without changing it.
static float time_ms(hipStream_t s, const std::function<int()>& fn, int reps, int* rc) {
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  *rc = fn();  // warm-up (layout, first touch)
  (void)hipEventRecord(a, s);
  for (int r = 0; r < reps && !*rc; r++) *rc = fn();
  (void)hipEventRecord(b, s);
  (void)hipEventSynchronize(b);
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms / reps;
}

// Time every candidate shape (apply(shape) sets the knobs, then `solve` is timed over 2 runs
// after a warm-up), then re-time the three fastest twice more, interleaved, and keep the shape
// with the lowest total: shapes whose solves differ by less than the run-to-run spread of one
// timing (a few percent on large batches) are otherwise ranked by chance.
static int tune(const std::vector<TuneShape>& shapes, hipStream_t s, const std::function<int(const TuneShape&)>& apply,
                const std::function<int()>& solve, TuneShape* best, float* best_ms) {
  int rc = 0;
  std::vector<std::pair<float, int>> t;
  for (size_t i = 0; i < shapes.size(); i++) {
    if ((rc = apply(shapes[i]))) return rc;
    const float ms = time_ms(s, solve, 2, &rc);
    if (rc) return rc;
    t.push_back({ms, (int)i});
  }
  if (t.empty()) return FL_OK;
  std::sort(t.begin(), t.end());
  const size_t top = std::min<size_t>(3, t.size());
  std::vector<float> total(top);
  for (size_t k = 0; k < top; k++) total[k] = t[k].first;
  for (int pass = 0; pass < 2; pass++)
    for (size_t k = 0; k < top; k++) {
      if ((rc = apply(shapes[t[k].second]))) return rc;
      total[k] += time_ms(s, solve, 2, &rc);
      if (rc) return rc;
    }
  size_t kb = 0;
  for (size_t k = 1; k < top; k++)
    if (total[k] < total[kb]) kb = k;
  *best = shapes[t[kb].second];
  *best_ms = total[kb] / 3;
  return FL_OK;
}

}  // namespace fl

using namespace fl;

#define B_OF(x) reinterpret_cast<Batch*>(x)

extern "C" {

int fl_batch_create(int n, const int* widths, const int* heights, float density, float h, fl_batch** out) {
  if (!out || n < 1 || !widths || !heights) return berr(FL_ERR_INVALID_ARG, "bad batch arguments");
  *out = nullptr;
  if (!(h > 0.0f)) return berr(FL_ERR_INVALID_ARG, "h must be > 0");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return berr(FL_ERR_NO_DEVICE, "no HIP device available: libfluidhip requires an MI355X (gfx950) GPU");
  Batch* B = new Batch();
  if (hipGetDevice(&B->dev) != hipSuccess) B->dev = 0;
  B->prm = default_params();
  B->nmax = MEMBER_NMAX;
  auto fail = [&](int code, const std::string& m) {
    batch_free(B);
    return berr(code, m);
  };
  size_t ncells = 0, ndiag = 0, nkeep = 0;
  for (int e = 0; e < n; e++) {
    if (widths[e] < 1 || heights[e] < 1) return fail(FL_ERR_INVALID_ARG, "member sizes must be >= 1");
    if ((int64_t)(widths[e] + 2) * (heights[e] + 2) > (int64_t)1 << 28)
      return fail(FL_ERR_INVALID_ARG, "member grid too large (more than 2^28 cells)");
    ncells += (size_t)(widths[e] + 2) * (heights[e] + 2);
    ndiag += (size_t)(widths[e] + heights[e] + 3) * (widths[e] + 2);
    nkeep += (size_t)(widths[e] + 2) * ((heights[e] + 2 + 63) / 64);
  }
  if (ncells * NBUF > ((size_t)1 << 31) * 8) return fail(FL_ERR_INVALID_ARG, "batch too large");
  B->ncells = ncells;
  hipError_t er = hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking);
  if (er != hipSuccess) return fail(FL_ERR_HIP, "hipStreamCreate");
  for (int k = 0; k < NBUF; k++) {
    if (hipMalloc(&B->pool[k], sizeof(float) * ncells) != hipSuccess) return fail(FL_ERR_HIP, "hipMalloc pool");
    (void)hipMemsetAsync(B->pool[k], 0, sizeof(float) * ncells, B->stream);
  }
  if (hipMalloc(&B->flags_pool, ncells) != hipSuccess || hipMalloc(&B->diag_pool, ndiag) != hipSuccess ||
      hipMalloc(&B->keep_pool, sizeof(uint32_t) * nkeep) != hipSuccess ||
      hipMalloc(&B->ctrl_pool, sizeof(int32_t) * 4 * n) != hipSuccess ||
      hipMalloc(&B->maxdiv_pool, sizeof(uint32_t) * 2 * MEMBER_NMAX * n) != hipSuccess ||
      hipMalloc(&B->d_sims, sizeof(SimDesc) * n) != hipSuccess ||
      hipMalloc(&B->d_omega, sizeof(float) * MEMBER_NMAX) != hipSuccess)
    return fail(FL_ERR_HIP, "hipMalloc batch");
  (void)hipMemsetAsync(B->diag_pool, 0, ndiag, B->stream);
  (void)hipMemsetAsync(B->ctrl_pool, 0, sizeof(int32_t) * 4 * n, B->stream);
  (void)hipMemsetAsync(B->maxdiv_pool, 0, sizeof(uint32_t) * 2 * MEMBER_NMAX * n, B->stream);
  std::vector<CellItem> cells;
  std::vector<RingItem> rings;
  size_t off = 0, doff = 0, koff = 0;
  for (int e = 0; e < n; e++) {
    Ctx* c = new Ctx();
    c->batch = B;
    c->member = e;
    c->dev = B->dev;
    c->stream = B->stream;
    c->NX = widths[e] + 2;
    c->NY = heights[e] + 2;
    c->N = (size_t)c->NX * c->NY;
    c->density = density;
    c->h = h;
    c->prm = default_params();
    c->grid = Grid{c->NX, c->NY, h, 1.0f / h, 0.5f * h, h / 2.0f, (float)c->NX * h, (float)c->NY * h};
    c->hS.assign(c->N, 0.0f);
    float** fields[NBUF] = {&c->U, &c->V, &c->nU, &c->nV, &c->p, &c->S, &c->M, &c->nM,
                            &c->U2, &c->V2, &c->M2, &c->T0, &c->T1, &c->T2, &c->T3};
    for (int k = 0; k < NBUF; k++) *fields[k] = B->pool[k] + off;
    c->flags = B->flags_pool + off;
    c->flags_diag = B->diag_pool + doff;
    c->keep = B->keep_pool + koff;
    c->d_ctrl = B->ctrl_pool + 4 * e;
    c->d_maxdiv = B->maxdiv_pool + (size_t)2 * MEMBER_NMAX * e;
    c->nmax = MEMBER_NMAX;
    B->members.push_back(c);
    if (hipMalloc(&c->d_omega, sizeof(float) * MEMBER_NMAX) != hipSuccess ||
        hipMalloc(&c->d_red, 64) != hipSuccess)
      return fail(FL_ERR_HIP, "hipMalloc member");
    for (int by = 0; by < (c->NX + 3) / 4; by++)
      for (int bx = 0; bx < (c->NY + 63) / 64; bx++) cells.push_back(CellItem{e, bx, by, 0});
    const int m = std::max(c->NX, c->NY);
    for (int blk = 0; blk < (m + 255) / 256; blk++) rings.push_back(RingItem{e, blk});
    off += c->N;
    doff += (size_t)(c->NX + c->NY - 1) * c->NX;
    koff += (size_t)c->NX * ((c->NY + 63) / 64);
  }
  {  // item k runs on XCD k % 8: give each XCD a contiguous run of the list
    std::vector<CellItem> dealt(cells.size());
    for (size_t k = 0; k < cells.size(); k++) dealt[k] = cells[xcd_tile((int)k, (int)cells.size())];
    cells.swap(dealt);
  }
  B->src_gen_seen.assign(n, -1);
  B->ncell_items = (int)cells.size();
  B->nring_items = (int)rings.size();
  if (hipMalloc(&B->d_cells, sizeof(CellItem) * cells.size()) != hipSuccess ||
      hipMalloc(&B->d_rings, sizeof(RingItem) * rings.size()) != hipSuccess)
    return fail(FL_ERR_HIP, "hipMalloc work-lists");
  (void)hipMemcpy(B->d_cells, cells.data(), sizeof(CellItem) * cells.size(), hipMemcpyHostToDevice);
  (void)hipMemcpy(B->d_rings, rings.data(), sizeof(RingItem) * rings.size(), hipMemcpyHostToDevice);
  if (hipStreamSynchronize(B->stream) != hipSuccess) return fail(FL_ERR_HIP, "batch init");
  *out = reinterpret_cast<fl_batch*>(B);
  return FL_OK;
}

int fl_batch_destroy(fl_batch* b) {
  batch_free(B_OF(b));
  return FL_OK;
}

int fl_batch_size(fl_batch* b, int* n) {
  if (!b || !n) return berr(FL_ERR_INVALID_ARG, "null argument");
  *n = (int)B_OF(b)->members.size();
  return FL_OK;
}

int fl_batch_member(fl_batch* b, int index, fl_ctx** out) {
  if (!b || !out) return berr(FL_ERR_INVALID_ARG, "null argument");
  Batch* B = B_OF(b);
  if (index < 0 || index >= (int)B->members.size()) return berr(FL_ERR_INVALID_ARG, "member index out of range");
  *out = reinterpret_cast<fl_ctx*>(B->members[index]);
  return FL_OK;
}

int fl_batch_get_params(fl_batch* b, fl_params* p) {
  if (!b || !p) return berr(FL_ERR_INVALID_ARG, "null argument");
  *p = B_OF(b)->prm;
  return FL_OK;
}

int fl_batch_set_params(fl_batch* b, const fl_params* p) {
  if (!b || !p) return berr(FL_ERR_INVALID_ARG, "null argument");
  if (p->pressure_iters < 0 || p->pressure_iters > MEMBER_NMAX)
    return berr(FL_ERR_INVALID_ARG, "pressure_iters must be in [0, 1024] for a batch");
  B_OF(b)->prm = *p;
  return FL_OK;
}

int fl_batch_update(fl_batch* b, float dt) {
  if (!b) return berr(FL_ERR_INVALID_ARG, "null batch");
  Batch* B = B_OF(b);
  if (hipSetDevice(B->dev) != hipSuccess) return berr(FL_ERR_HIP, "hipSetDevice");
  for (Ctx* c : B->members)
    if (c->NX < 3 || c->NY < 3) return berr(FL_ERR_UNSUPPORTED, "members need at least 1x1 interior cells");
  return batch_update(B, dt);
}

int fl_batch_sync(fl_batch* b) {
  if (!b) return berr(FL_ERR_INVALID_ARG, "null batch");
  FL_HIP(hipStreamSynchronize(B_OF(b)->stream));
  return FL_OK;
}

int fl_batch_set_graph(fl_batch* b, int on) {
  if (!b) return berr(FL_ERR_INVALID_ARG, "null batch");
  Batch* B = B_OF(b);
  B->use_graph = on != 0;
  if (!B->use_graph && B->graph) {
    FL_HIP(hipGraphExecDestroy(B->graph));
    B->graph = nullptr;
  }
  return FL_OK;
}

int fl_batch_autotune(fl_batch* b, float dt, int* block_sweeps, float* ms) {
  if (!b) return berr(FL_ERR_INVALID_ARG, "null batch");
  Batch* B = B_OF(b);
  const int N = B->prm.pressure_iters;
  if (N <= 0 || (B->prm.use_multigrid && B->prm.multigrid_levels > 1)) return FL_OK;
  if (hipSetDevice(B->dev) != hipSuccess) return berr(FL_ERR_HIP, "hipSetDevice");
  long long tiles8 = 0;
  for (Ctx* c : B->members) tiles8 += tiles_at(c->NX, N, 8);
  TuneShape best{};
  float best_ms = 1e30f;
  const fl_params saved = B->prm;
  // trial solves write only scratch roles (U2, V2, T2), so the members' state is unchanged
  int rc = tune(
      tune_shapes(tiles8, device_cus(B->dev)), B->stream,
      [&](const TuneShape& sh) {
        B->prm.block_sweeps = sh.tk;
        B->prm.block_cols = sh.blk;
        B->prm.handoff_lag = sh.lag;
        return prepare_batch(B, dt);
      },
      [&]() { return batch_gs_solve(B, N, -1, B_T2); }, &best, &best_ms);
  B->prm = saved;
  if (rc) return rc;
  B->prm.block_sweeps = best.tk;
  B->prm.block_cols = best.blk;
  B->prm.handoff_lag = best.lag;
  if ((rc = prepare_batch(B, dt))) return rc;
  if (block_sweeps) *block_sweeps = best.tk;
  if (ms) *ms = best_ms;
  return FL_OK;
}

int fl_autotune(fl_ctx* ctx, float dt, int* block_sweeps, float* ms) {
  CtxErrScope err_scope_(reinterpret_cast<fl::Ctx*>(ctx));
  if (!ctx) return berr(FL_ERR_INVALID_ARG, "null context");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  const int N = c->prm.pressure_iters;
  if (N <= 0 || c->prm.solver == 1 || (c->prm.use_multigrid && c->prm.multigrid_levels > 1)) return FL_OK;
  if (hipSetDevice(c->dev) != hipSuccess) return berr(FL_ERR_HIP, "hipSetDevice");
  int rc = flush_edits(c);
  if (rc) return rc;
  if (c->flags_dirty && (rc = launch_flags(c))) return rc;
  TuneShape best{};
  float best_ms = 1e30f;
  const fl_params saved = c->prm;
  c->prm.solver = 2;
  rc = tune(
      tune_shapes(tiles_at(c->NX, N, 8), device_cus(c->dev)), c->stream,
      [&](const TuneShape& sh) {
        c->prm.block_sweeps = sh.tk;
        c->prm.block_cols = sh.blk;
        c->prm.handoff_lag = sh.lag;
        return FL_OK;
      },
      [&]() {
        // the solve swaps U/U2, V/V2, p/T3 towards its outputs; swapping back restores the state
        int r = pressure_solve(c, N, dt);
        std::swap(c->U, c->U2);
        std::swap(c->V, c->V2);
        std::swap(c->p, c->T3);
        return r;
      },
      &best, &best_ms);
  c->prm = saved;
  if (rc) return rc;
  c->prm.block_sweeps = best.tk;
  c->prm.block_cols = best.blk;
  c->prm.handoff_lag = best.lag;
  if (block_sweeps) *block_sweeps = best.tk;
  if (ms) *ms = best_ms;
  return FL_OK;
}

}  // extern "C"
