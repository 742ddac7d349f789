// Internal declarations of libfluidhip: the context that owns the device state of one
// simulation, the host-side edit queue, and the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fluid_hip.h"
#include "fl_device.h"

namespace fl {

void set_error(const std::string& msg);

#define FL_HIP(expr)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) {                                                             \
      ::fl::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));               \
      return FL_ERR_HIP;                                                                \
    }                                                                                   \
  } while (0)

// One word-level edit: set or add `val` at field[idx]. Host API calls (SetSolid,
// SetVelocity, AddSmoke, ApplyForce) expand to these, in call order.
struct WordOp {
  uint32_t key;   // field * numCells + idx
  uint32_t kind;  // 0 = set, 1 = add
  float val;
};

// Per-step source program (jet, sources, sinks): ops grouped per cell, in the
// order the reference applies them (sources, sinks, jet; main.go:231-241,474-486).
struct SrcOp {
  int32_t cell;
  int32_t kind;  // 0 = SetVelocity(u,v), 1 = AddSmoke(a)
  int32_t check_solid;
  float a, b;
};

struct Source {
  int I, J;
  float U, V;
};

// Buffer roles of one simulation (indices into SimDesc::buf).
enum Buf { B_U = 0, B_V, B_NU, B_NV, B_P, B_S, B_M, B_NM, B_U2, B_V2, B_M2, B_T0, B_T1, B_T2, B_T3, NBUF };

// Everything a kernel needs to know about one simulation: its grid, its buffers and,
// for the Gauss-Seidel kernel, its tile layout and handoff buffers.
struct SimDesc {
  Grid g;
  float* buf[NBUF];
  uint8_t* flags;
  uint8_t* diag;
  uint32_t* keep;     // [NX][ceil(NY/64)]: per row and 64-column block, bit 0 some interior velocity
                      // face is not advected (advection keeps newU/newV), bit 1 some interior cell
                      // is solid (keeps newM); written by k_flags with the flags
  int32_t* ctrl;      // [4]: [1] re-run sweeps, [2] invalid sample flag, [3] GS wait timeout
  uint32_t* maxdiv;   // [2][nmax]
  float* kout;        // GS K-stream handoff planes
  float* gout;        // GS G-stream handoff planes
  int32_t* prog;      // [2][G*KB] GS progress counters
  int G, KB, rows, nmax;
  float cp;           // density*h/dt of this step
  int early_exit;
};

// A 64x4 block of cells of one simulation (batched cell kernels) and a 256-cell chunk
// of one simulation's ring (batched ring kernels).
struct CellItem {
  int sim, bx, by, pad;
};
struct RingItem {
  int sim, blk;
};

// Workgroups are dispatched to the 8 XCDs in turn (workgroup b runs on XCD b % 8), and each
// XCD has its own L2. xcd_tile(b, B) hands XCD x a contiguous run of the B tiles (row bands of a
// grid, consecutive items of a work-list), so the neighbour rows a tile's stencils and gathers
// read were mostly brought into the same L2 by the tiles just before it. A bijection on [0, B):
// XCD x takes q + (x < B % 8) tiles, starting at x * q + min(x, B % 8), q = B / 8.
__host__ __device__ inline int xcd_tile(int b, int B) {
  const int q = B >> 3, rem = B & 7, x = b & 7, r = b >> 3;
  return x * q + (x < rem ? x : rem) + r;
}

// Launch maps: one simulation passed by value, or many through a work-list.
struct OneSim {
  SimDesc d;
};
struct ManySims {
  const SimDesc* sims;
  const CellItem* cells;
  const RingItem* rings;
};

// Pressure workspace of the temporal-blocked Gauss-Seidel kernel.
struct GsWork {
  int tk = 0;              // sweeps per tile (waves per workgroup)
  int blk = 0;             // columns per staging/publish block
  int n_alloc = -1;        // sweep count the layout was built for
  int G = 0, KB = 0;       // tiles along skewed rows and along sweeps
  int rows = 0;            // handoff rows per tile (= steps per tile)
  int ntiles = 0;          // non-empty tiles in ticket order
  int grid = 0;            // persistent workgroups launched
  int defer = -1;          // handoff publication of the last solve (GsRun.defer; -1: none yet)
  int32_t* d_tile = nullptr;    // [ntiles] packed (g << 16 | kb), ticket order
  float* d_kout = nullptr;      // [G*KB][rows][3][65]  K-stream handoff planes U, V, p
  float* d_gout = nullptr;      // [G*KB][rows][3][tk]  G-stream handoff planes
  int32_t* d_prog = nullptr;    // [2][G*KB] progress counters (main run, rerun)
  int32_t* d_ticket = nullptr;  // [2]
  size_t kout_elems = 0, gout_elems = 0;
};

struct Batch;

struct Ctx {
  Batch* batch = nullptr;  // set for members of a batch: buffers live in the batch's pools
  int member = -1;
  int dev = 0;
  hipStream_t stream = nullptr;
  int NX = 0, NY = 0;
  size_t N = 0;  // NX*NY
  float density = 0, h = 0;
  fl_params prm{};
  Grid grid{};
  // Simulate's fill(p, 0) is left to its temporal-blocked solve, which then starts from p = 0
  // without reading p (set by the step around its solve only)
  bool gs_p_zero = false;

  // Device state: the 8 reference buffers plus ping-pong partners and scratch.
  float *U = nullptr, *V = nullptr, *nU = nullptr, *nV = nullptr, *p = nullptr, *S = nullptr, *M = nullptr,
        *nM = nullptr;
  float *U2 = nullptr, *V2 = nullptr, *M2 = nullptr;
  float *T0 = nullptr, *T1 = nullptr, *T2 = nullptr, *T3 = nullptr;  // BFECC / curl scratch
  uint8_t* flags = nullptr;      // cell flags, i-major
  uint32_t* keep = nullptr;      // SimDesc::keep
  int32_t* snap_hint = nullptr;  // host-pinned: the last solve that reported back ended after sweep 0
  uint8_t* flags_diag = nullptr; // cell flags, anti-diagonal-major [d=i+j][i]
  bool flags_dirty = true;

  // Solver control block on the device.
  int32_t* d_ctrl = nullptr;       // [0] sweeps to run, [1] rerun sweeps, [2] invalid sample flag, [3] gs error
  uint32_t* d_maxdiv = nullptr;    // [2][nmax] float bits of per-sweep max |div| (main, rerun)
  float* d_omega = nullptr;        // [nmax] relaxation of each sweep
  int nmax = 0;
  int omega_n = -1;                // sweep count and Relaxation the table holds
  float omega_relax = 0.0f;
  uint32_t* d_red = nullptr;       // scratch for reductions
  float* d_xy = nullptr;           // scratch for batched sampling
  int xy_cap = 0;

  GsWork gs;

  // Multigrid V-cycle workspace (coarse grid of ((NX+1)/2) x ((NY+1)/2) cells).
  float* mg_omega = nullptr;  // [6]: the pre-smoothing (1.5) and post-smoothing (1.2) relaxation
  float* mg_P = nullptr;      // coarse correction
  float* mg_S = nullptr;      // coarse solid mask
  float* mg_R = nullptr;      // coarse right-hand side

  // Scratch of rendering, range reductions and particles (fl_render.hip).
  char* rs_mem = nullptr;
  size_t rs_cap = 0;

  // Pinned staging of large device-to-host copies (copy_to_host): two chunk buffers.
  char* stage = nullptr;
  hipEvent_t stage_ev[2] = {nullptr, nullptr};

  // Host mirror of S (always valid) and pending edits.
  std::vector<float> hS;
  std::vector<WordOp> edits;
  uint32_t* d_edit_keys = nullptr;
  float* d_edit_vals = nullptr;
  int32_t* d_edit_seg = nullptr;
  size_t edit_cap = 0;

  // Scene (package main) state.
  bool wallTop = true, wallBottom = true, wallLeft = true, wallRight = false;
  bool jet = false;
  int jetX = 1, jetLo = 0, jetHi = 0;
  float jetU = 4.0f, jetSmoke = 1.0f;
  std::vector<Source> sources, sinks;
  bool src_dirty = true;
  int src_gen = 0;  // bumped with every change of the source program
  SrcOp* d_src = nullptr;
  int32_t* d_src_seg = nullptr;
  int n_src_ops = 0, n_src_seg = 0;
  size_t src_cap = 0;

  // Optional per-pass timing (fl_set_timing).
  bool timing = false;
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  double t_total = 0, t_pressure = 0, t_advect = 0, t_other = 0;
  int t_steps = 0;

  // Text of the last failing call on this context (fl_ctx_last_error).
  std::string last_err;

  // Stats of the last solve.
  int last_n = 0;
  bool last_rollback_checked = true;
  int host_sweeps = -1;  // sweeps known on the host (hyperplane path), -1 if on device only

  // Descriptor of this simulation with its current buffers.
  SimDesc desc() const {
    SimDesc d{};
    d.g = grid;
    d.g.precise = prm.precise_advection != 0;
    float* b[NBUF] = {U, V, nU, nV, p, S, M, nM, U2, V2, M2, T0, T1, T2, T3};
    for (int k = 0; k < NBUF; k++) d.buf[k] = b[k];
    d.flags = flags;
    d.diag = flags_diag;
    d.keep = keep;
    d.ctrl = d_ctrl;
    d.maxdiv = d_maxdiv;
    d.kout = gs.d_kout;
    d.gout = gs.d_gout;
    d.prog = gs.d_prog;
    d.G = gs.G;
    d.KB = gs.KB;
    d.rows = gs.rows;
    d.nmax = nmax;
    d.early_exit = prm.early_exit;
    return d;
  }
};

// A batch: E member simulations whose buffers are carved from shared pools, stepped
// together with one launch per pass (fl_batch.hip).
struct Batch {
  int dev = 0;
  hipStream_t stream = nullptr;
  std::vector<Ctx*> members;
  float* pool[NBUF] = {};
  uint8_t* flags_pool = nullptr;
  uint8_t* diag_pool = nullptr;
  uint32_t* keep_pool = nullptr;
  int32_t* ctrl_pool = nullptr;
  uint32_t* maxdiv_pool = nullptr;
  size_t ncells = 0;
  fl_params prm{};
  int nmax = 0;
  // work-lists
  SimDesc* d_sims = nullptr;
  CellItem* d_cells = nullptr;
  RingItem* d_rings = nullptr;
  int ncell_items = 0, nring_items = 0;
  std::vector<SimDesc> descs;  // what d_sims holds
  // batched Gauss-Seidel layout
  int tk = 0, blk = 0, gs_n = -1, gs_nprog = 0;
  int ntiles = 0, grid = 0;
  std::vector<SimDesc> gs_fields;  // per member: G, KB, rows and its handoff/progress slices
  int2* d_tiles = nullptr;
  float* kout = nullptr;
  float* gout = nullptr;
  int32_t* prog = nullptr;
  int32_t* ticket = nullptr;
  float* d_omega = nullptr;
  int omega_n = -1;
  float omega_relax = 0.0f;
  // per-step sources of all members
  SrcOp* d_src = nullptr;
  int32_t* d_src_seg = nullptr;
  int32_t* d_src_sim = nullptr;
  int n_src_seg = 0;
  size_t src_cap = 0;
  std::vector<int> src_gen_seen;
  // staging of the members' queued edits, applied by one launch per step
  uint32_t* d_ekeys = nullptr;
  float* d_evals = nullptr;
  int32_t* d_eseg = nullptr;
  int32_t* d_esim = nullptr;
  size_t edit_cap = 0;
  // captured step: replayed while dt, the parameters and every launch shape are unchanged
  bool use_graph = true;
  hipGraphExec_t graph = nullptr;
  float graph_dt = -1.0f;
  fl_params graph_prm{};
  int graph_gen = 0, graph_key_gen = -1;  // bumped when a launch shape changes
  int steps = 0;
};

// Errors of a call that takes a context are also recorded in that context, so that a caller
// whose next call may run on another OS thread (a cgo binding) still reads them: each such
// entry point opens a CtxErrScope for the duration of the call.
struct CtxErrScope {
  Ctx* prev;
  explicit CtxErrScope(Ctx* c);
  ~CtxErrScope();
};

// ---- host runtime shared by single contexts and batches (fl_capi.hip) ------------------
int flush_edits(Ctx* c);
// Device-to-host copy on the context's stream that returns when `dst` holds the data.
int copy_to_host(Ctx* c, void* dst, const void* src, size_t bytes);
int copy_to_device(Ctx* c, void* dst, const void* src, size_t bytes);
// True when copy_to_host / copy_to_device move `bytes` with one plain async copy (no pinned
// staging buffer, which those calls allocate on first use).
bool copy_is_direct(const Ctx* c, size_t bytes);
// MinValue/MaxValue of a device field (all cells, or the fluid interior), reduced on the device.
int field_range(Ctx* c, const float* f, int interior, float* mn, float* mx);
int ctx_prepare(Ctx* c);        // flush edits, refresh flags
int ctx_build_sources(Ctx* c);  // upload the per-step source program if it changed
void source_ops(Ctx* c, std::vector<SrcOp>& ops);
void free_ctx(Ctx* c);
fl_params default_params();

// ---- launchers (fl_kernels.hip) ------------------------------------------------------
int launch_fill(float* dst, float v, size_t n, hipStream_t s);
int launch_zero_u8(uint8_t* dst, size_t n, hipStream_t s);
int launch_copy(float* dst, const float* src, size_t n, hipStream_t s);
int launch_flags(Ctx* c);
int launch_copy_border(Ctx* c, float* dst, const float* src);
int launch_viscosity(Ctx* c, float dt);
int launch_curl(Ctx* c, float* curl);
int launch_conf_turb(Ctx* c, const float* curl, float dt, int do_conf, int do_turb);
int launch_borders(Ctx* c);
// SL velocity advection: trace with (Ut,Vt), sample (Us,Vs); write nU/nV and (Uo,Vo).
// write_new = false: leave newU/newV (newM) alone. Only for BFECC's forward pass, whose final
// pass rewrites them at every cell and reads them only where the forward pass did not change them.
int launch_advect_velocity(Ctx* c, const float* Ut, const float* Vt, float* Uo, float* Vo, float dt,
                           bool write_new = true);
// BFECC backward pass fused with the correction and the clamp.
int launch_bfecc_velocity_corr(Ctx* c, const float* Uorig, const float* Vorig, const float* Ufwd,
                               const float* Vfwd, float* Ucorr, float* Vcorr, float dt);
int launch_advect_smoke(Ctx* c, const float* Msrc, float* Mo, float dt, bool write_new = true);
int launch_bfecc_smoke_corr(Ctx* c, const float* Morig, const float* Mfwd, float* Mcorr, float dt);
int launch_apply_edits(Ctx* c, int nseg);
int launch_apply_sources(Ctx* c);
int launch_max_reduce(Ctx* c, int kind, float* out);  // 0 MaxDivergence, 1 max(|U|+|V|)
int launch_view(Ctx* c, int kind, float* out);        // 0 vorticity, 1 velocity magnitude
int launch_sample_batch(Ctx* c, const float* xy, int n, float* uv);

// ---- batched passes (fl_kernels.hip): one launch over all members of a batch -----------
int batch_fill_role(Batch* B, int r, float v);
int batch_copy_role(Batch* B, int rdst, int rsrc);
int batch_swap_roles(Batch* B, int a, int b);  // in every member's device descriptor
int batch_flags(Batch* B);
int batch_viscosity(Batch* B, float dt);
int batch_conf_turb(Batch* B, float dt, int rU, int rV);
int batch_borders(Batch* B, int rU, int rV);
int batch_advect_velocity(Batch* B, int rUt, int rVt, int rUo, int rVo, float dt, bool write_new = true);
int batch_bfecc_velocity_corr(Batch* B, int rUo, int rVo, int rUf, int rVf, float dt);
int batch_advect_smoke(Batch* B, int rMs, int rMo, float dt, bool write_new = true);
int batch_bfecc_smoke_corr(Batch* B, float dt);
int batch_apply_sources(Batch* B);
int batch_apply_edits(Batch* B, const uint32_t* keys, const float* vals, const int32_t* seg, const int32_t* segsim,
                      int nseg);
// batched pressure solve (fl_gs_wave.hip): sets the GS fields of `descs`
int pick_tk(const fl_params& P, int N);
bool gs_wave_supported(Ctx* c, int numIters);
int pick_blk(const fl_params& P);
int batch_gs_layout(Batch* B, int N, std::vector<SimDesc>& descs);
int batch_gs_solve(Batch* B, int N, int rPin, int rPout);  // p from rPin (< 0: zero), to rPout

// ---- pressure (fl_pressure.hip) -------------------------------------------------------
// Runs the single-grid projection: sweeps with the reference order on (U,V,p).
// On return the projected velocity is in (c->U, c->V) and the pressure in c->p.
int pressure_solve(Ctx* c, int numIters, float dt);
// solveMultigridVCycle (fluid.go:559-596): on return the result is in (c->U, c->V, c->p).
int multigrid_solve(Ctx* c, int numIters, float dt);
int pressure_free(Ctx* c);

}  // namespace fl
