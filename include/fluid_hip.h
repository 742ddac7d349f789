/*
 * fluid_hip.h — C ABI of libfluidhip.so, the MI355X (gfx950) implementation of the
 * TheFellow/fluid `pkg/fluid` simulation step.
 *
 * The reference has no FFI: its interface is the Go package API
 * (github.com/TheFellow/fluid/pkg/fluid). Each entry point below replaces one Go
 * function or method; the reference file:line it stands for is given beside it.
 * A cgo binding restores the exact Go API on top of these calls. Plain pointers and sizes only; no Go pointer is retained.
 *
 * Layout: every field is float32 [NumX*NumY] with index i*NumY + j (j contiguous),
 * NumX = width+2, NumY = height+2, as in the reference (fluid.go:42-58).
 *
 * Error behaviour: every call returns an fl_status. Where the reference panics
 * (walls.go:6-11 and its siblings) the call returns FL_ERR_X_RANGE / FL_ERR_Y_RANGE
 * and fl_last_error() carries the reference's panic text. ApplyForce ignores
 * out-of-range and solid cells like the reference (fluid.go:762-768).
 *
 * Threading: a context is not safe for concurrent use (the reference Fluid is not
 * either); it owns one HIP stream. Calls on one context are ordered.
 */
#ifndef FLUID_HIP_H
#define FLUID_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fl_ctx fl_ctx;

typedef enum fl_status {
  FL_OK = 0,
  FL_ERR_INVALID_ARG = 1,
  FL_ERR_X_RANGE = 2,     /* reference: panic("invalid x-index: %d") */
  FL_ERR_Y_RANGE = 3,     /* reference: panic("invalid y-index: %d") */
  FL_ERR_NO_DEVICE = 4,
  FL_ERR_HIP = 5,
  FL_ERR_UNSUPPORTED = 6,
  FL_ERR_INTERNAL = 7
} fl_status;

typedef enum fl_field {
  FL_FIELD_U = 0,    /* Fluid.U    (fluid.go:17) */
  FL_FIELD_V = 1,    /* Fluid.V    (fluid.go:17) */
  FL_FIELD_NEWU = 2, /* Fluid.newU (fluid.go:18), stale-value semantics kept */
  FL_FIELD_NEWV = 3, /* Fluid.newV (fluid.go:18) */
  FL_FIELD_P = 4,    /* Fluid.p    (fluid.go:19) */
  FL_FIELD_S = 5,    /* Fluid.S    (fluid.go:20); values must be exactly 0.0 or 1.0 */
  FL_FIELD_M = 6,    /* Fluid.M    (fluid.go:22) */
  FL_FIELD_NEWM = 7  /* Fluid.newM (fluid.go:23) */
} fl_field;

/* Exported knobs of Fluid (fluid.go:25-39) plus the package variable Relaxation
 * (fluid.go:7-9). pressure_iters and early_exit are additions: the reference
 * hard-codes 8 sweeps (fluid.go:81) with the maxDiv<1e-5 exit (fluid.go:175-177). */
typedef struct fl_params {
  float confinement;         /* Confinement, default 0 */
  float viscosity_diffusion; /* ViscosityDiffusion, default 0 */
  float pressure_damping;    /* PressureDamping, default 1 */
  float turbulence_strength; /* TurbulenceStrength, default 0.02 */
  float smoke_advection;     /* SmokeAdvection, default 1 */
  float relaxation;          /* Relaxation (package var), default 1.9 */
  int32_t use_multigrid;     /* UseMultigrid, default 0 */
  int32_t multigrid_levels;  /* MultigridLevels, default 2 */
  int32_t use_bfecc;         /* UseBFECC, default 0 */
  int32_t pressure_iters;    /* sweeps per Simulate, default 8 */
  int32_t early_exit;        /* 1: stop after the first sweep with maxDiv < 1e-5 */
  int32_t solver;            /* 0 auto, 1 hyperplane kernel, 2 temporal-blocked kernel */
  int32_t block_sweeps;      /* sweeps per temporal block: 0 auto (8; 4 for at most 8 sweeps),
                                2, 4, 8 or 16 */
  int32_t precise_advection; /* 1: back-trace in cell units (no reference counterpart), default 0 */
  int32_t handoff_lag;       /* temporal-blocked solve: 0 auto, 1 publish each block's handoff rows
                                at the next block's start, 2 one block later (fewer waits, longer
                                pipeline fill), 3 halfway through the next block; auto picks 3
                                from two rounds of tiles up, else 1 */
  int32_t block_cols;        /* temporal-blocked solve: columns per staging and publish block,
                                0 auto (8), 4, 8 or 16 */
} fl_params;

/* Statistics of the last Simulate (no reference counterpart). */
typedef struct fl_stats {
  int32_t sweeps_run;     /* pressure sweeps whose result was kept */
  int32_t reran;          /* 1 if the early exit rolled back and re-ran fewer sweeps */
  int32_t invalid_sample; /* 1 if a NaN/Inf reached a sample index (the reference panics) */
  float max_div_last;     /* maxDiv of the last kept sweep */
} fl_stats;

const char* fl_last_error(void);
/* The text of the last failing call on `ctx` (the reference's panic text for FL_ERR_X_RANGE /
 * FL_ERR_Y_RANGE, walls.go:6-11), copied into buf (n bytes, NUL-terminated, truncated). No
 * reference counterpart: fl_last_error is per OS thread, and a cgo caller's next call may run
 * on another thread; this one can be read from any thread. */
int fl_ctx_last_error(fl_ctx* ctx, char* buf, size_t n);
int fl_device_count(int* n);

/* New(density, width, height, h) — fluid.go:42-68. S starts all zero (all solid). */
int fl_create(float density, int width, int height, float h, fl_ctx** out);
int fl_destroy(fl_ctx* ctx);
/* NumX, NumY (fluid.go:15) and H() (fluid.go:71). */
int fl_dims(fl_ctx* ctx, int* num_x, int* num_y);
int fl_get_h(fl_ctx* ctx, float* h);
int fl_get_params(fl_ctx* ctx, fl_params* out);
int fl_set_params(fl_ctx* ctx, const fl_params* in);

/* Whole-field access to the exported slices U, V, S, M and the internal buffers
 * (the reference exposes U,V,S,M as mutable slices, fluid.go:17-22). n = NumX*NumY. */
int fl_upload(fl_ctx* ctx, int field, const float* src, size_t n);
int fl_download(fl_ctx* ctx, int field, float* dst, size_t n);
/* A loop of direct writes f.F[i*NumY + j] = value over i in [i0, i1), j in [j0, j1) to any
 * field but S (whose writes go through fl_set_solid), as the reference's callers do on the
 * exported slices each frame (fluid_bench_test.go:36-39, fluid_test.go:1097-1101). Asynchronous
 * on the context's stream, after the queued edits; an empty range does nothing. */
int fl_fill_cells(fl_ctx* ctx, int field, int i0, int i1, int j0, int j1, float value);

/* Simulate(dt) — fluid.go:79-109. Asynchronous on the context's stream. */
int fl_simulate(fl_ctx* ctx, float dt);
int fl_sync(fl_ctx* ctx);
int fl_get_stats(fl_ctx* ctx, fl_stats* out);
/* Per-pass device time (no reference counterpart): when enabled, every Simulate records
 * HIP events on the context's stream around its pass groups and waits for them. */
typedef struct fl_timing {
  double total_ms;     /* whole steps */
  double pressure_ms;  /* pressure projection (all sweeps, rollback included) */
  double advect_ms;    /* velocity + smoke advection (SL or BFECC) */
  double other_ms;     /* p clear, viscosity, confinement, turbulence, borders, sources */
  int32_t steps;
  int32_t pad;
} fl_timing;
int fl_set_timing(fl_ctx* ctx, int on); /* also resets the accumulators */
int fl_get_timing(fl_ctx* ctx, fl_timing* out);

/* maxDiv of every sweep of the last solve (speculative run, before a rollback). */
int fl_sweep_maxdiv(fl_ctx* ctx, float* out, int cap, int* count);

/* The tiling of the last temporal-blocked Gauss-Seidel solve on `ctx` (no reference
 * counterpart: the reference's solve is a serial loop, fluid.go:188-234). For tuning and for
 * the critical-path model of the solve's time. FL_ERR_INVALID_ARG before the first such solve. */
typedef struct fl_gs_layout {
  int32_t sweeps_per_tile;  /* TK */
  int32_t cols_per_block;   /* BLK: columns per staging and publication block */
  int32_t handoff_lag;      /* publication used: 1 on time, 2 one block late, 3 halfway */
  int32_t row_tiles;        /* G: tiles along the skewed rows (64 rows each) */
  int32_t sweep_blocks;     /* KB: tiles along the sweeps */
  int32_t steps_per_tile;   /* column steps each tile runs */
  int32_t tiles;            /* non-empty tiles */
  int32_t workgroups;       /* persistent workgroups launched */
} fl_gs_layout;
int fl_get_gs_layout(fl_ctx* ctx, fl_gs_layout* out);

/* walls.go:5-93 */
int fl_set_solid(fl_ctx* ctx, int i, int j, int solid);   /* SetSolid  walls.go:5  */
int fl_is_solid(fl_ctx* ctx, int i, int j, int* out);      /* IsSolid   walls.go:51 */
int fl_set_velocity(fl_ctx* ctx, int i, int j, float u, float v); /* walls.go:62 */
int fl_add_smoke(fl_ctx* ctx, int i, int j, float s);      /* AddSmoke  walls.go:74 */
int fl_reset(fl_ctx* ctx);                                 /* Reset     walls.go:85 */

/* fluid.go:761-907 */
int fl_apply_force(fl_ctx* ctx, int i, int j, float fx, float fy);            /* :761 */
int fl_apply_force_radius(fl_ctx* ctx, int cx, int cy, float fx, float fy, int radius); /* :774 */
/* Gaussian impulse: ApplyForce(i, j, fx*w, fy*w) (fluid.go:761) for every interior cell
 * within 3*sigma of (cx, cy), w = exp(-d^2 / (2 sigma^2)), cells in i-major order like
 * ApplyForceRadius (fluid.go:774, whose weight exp(-3 d^2/r^2) is this with sigma =
 * r/sqrt(6) but not truncated at 3 sigma). An addition of this library. */
int fl_apply_force_gaussian(fl_ctx* ctx, int cx, int cy, float fx, float fy, float sigma);
int fl_set_circular_obstacle(fl_ctx* ctx, int cx, int cy, int radius);        /* :894 */
int fl_get_adaptive_time_step(fl_ctx* ctx, float basedt, float* out);         /* :529 */
int fl_max_divergence(fl_ctx* ctx, float* out);                               /* :876 */
int fl_sample_velocity(fl_ctx* ctx, float x, float y, float* u, float* v);    /* :799 */
/* Batched SampleVelocity for tracer particles (main.go:527-530 calls it per particle). */
int fl_sample_velocity_batch(fl_ctx* ctx, const float* xs, const float* ys, int n, float* us, float* vs);
/* Vorticity() (:806) and VelocityMagnitude() (:841): the field plus MinValue/MaxValue. */
int fl_vorticity(fl_ctx* ctx, float* out, size_t n, float* min_value, float* max_value);
int fl_velocity_magnitude(fl_ctx* ctx, float* out, size_t n, float* min_value, float* max_value);
/* Pressure() (pressure.go:5) / Smoke() (smoke.go:5): MinValue/MaxValue over all cells. */
int fl_field_minmax(fl_ctx* ctx, int field, float* min_value, float* max_value);

/* Game.Draw's image (main.go:549-652, colors.go:10-84): NumY rows x NumX RGBA8 pixels, row j
 * showing fluid row NumY-1-j; view 0 Smoke(), 1 Pressure(), 2 VelocityMagnitude() with the
 * scientific colormap (getSciValue), 3 Vorticity() with the diverging one (getDivergingColor);
 * solid cells black; with `overlay` source pixels green and sink pixels red. Computed on the
 * device, field ranges included: nbytes = 4*NumX*NumY. */
int fl_render(fl_ctx* ctx, int view, int overlay, uint8_t* rgba, size_t nbytes);
/* Game.advectParticles (main.go:512-546) for n particles in one call: age += dt; a particle
 * older than max_age, or whose cell int(x/h), int(y/h) after the RK2 midpoint step (with
 * SampleVelocity, fluid.go:799) is outside the grid or solid, gets keep = 0. x, y, age are
 * updated in place; the caller drops the particles with keep = 0 (keeping the order). */
int fl_advect_particles(fl_ctx* ctx, float dt, int n, float* x, float* y, float* age, const float* max_age,
                        uint8_t* keep);
/* fl_advect_particles for callers that cannot allocate device memory (code under HIP graph
 * capture, per-frame loops, harnesses that forbid hipMalloc). No counterpart in the reference.
 * The particle arrays are staged in a static device array of FL_PARTICLES_NOALLOC_MAX
 * particles (2.2 MB, held on each GPU used until the process ends) instead of the context's
 * scratch, which fl_advect_particles grows with hipMalloc. Same results as
 * fl_advect_particles. n > FL_PARTICLES_NOALLOC_MAX returns FL_ERR_INVALID_ARG, and so do
 * queued edits too many for the context's edit buffer (flush them with a step first). All
 * calls on one GPU share the array: calls that overlap on one GPU race. */
#define FL_PARTICLES_NOALLOC_MAX 131072
int fl_advect_particles_noalloc(fl_ctx* ctx, float dt, int n, float* x, float* y, float* age,
                                const float* max_age, uint8_t* keep);

/* Unexported passes, for white-box tests in the manner of fluid_test.go. */
typedef enum fl_pass_id {
  FL_PASS_COPY_BORDER_UV = 0,   /* copyBorder(newU,U); copyBorder(newV,V)  fluid.go:436 */
  FL_PASS_VISCOSITY = 1,        /* applyViscosity          fluid.go:112 */
  FL_PASS_MAKE_INCOMPRESSIBLE = 2, /* makeIncompressible(iters,dt) fluid.go:144 */
  FL_PASS_CONFINEMENT = 3,      /* applyVorticityConfinement fluid.go:449 */
  FL_PASS_TURBULENCE = 4,       /* addTurbulence           fluid.go:496 */
  FL_PASS_BORDERS = 5,          /* handleBorders           fluid.go:236 */
  FL_PASS_ADVECT_VELOCITY = 6,  /* advectVelocity          fluid.go:291 */
  FL_PASS_ADVECT_SMOKE = 7,     /* advectSmoke             fluid.go:400 */
  FL_PASS_ADVECT_VELOCITY_BFECC = 8, /* advectVelocityBFECC  fluid.go:911 */
  FL_PASS_ADVECT_SMOKE_BFECC = 9,    /* advectSmokeBFECC     fluid.go:997 */
  FL_PASS_CLEAR_P = 10          /* fill(p, 0)              fluid.go:83 */
} fl_pass_id;
int fl_pass(fl_ctx* ctx, int pass, float dt, int iters);

/* Scene layer lifted from package main (main/main.go), which the reference keeps
 * outside pkg/fluid. Presets: 0 jet, 1 cavity, 2 karman (main.go:747-793). */
int fl_scene_set_walls(fl_ctx* ctx, int top, int bottom, int left, int right);
int fl_scene_apply_walls(fl_ctx* ctx);   /* applyWallSettings main.go:840 */
int fl_scene_clear_walls(fl_ctx* ctx);   /* clearWalls        main.go:851 */
int fl_scene_preset(fl_ctx* ctx, int preset, float lid_u); /* resetToPreset main.go:747 (lid_u=3 there) */
/* Jet (main.go:236-241): each Update does SetVelocity(x,j,u,0); AddSmoke(x,j,smoke)
 * for j in [j_lo, j_hi). */
int fl_scene_set_jet(fl_ctx* ctx, int on, int x, int j_lo, int j_hi, float u, float smoke);
/* Sources (sink=0) and sinks (sink=1), main.go:147-150,474-486. */
int fl_scene_add_source(fl_ctx* ctx, int sink, int i, int j, float u, float v);
int fl_scene_clear_sources(fl_ctx* ctx);
int fl_scene_apply_sources(fl_ctx* ctx); /* applySources main.go:474 */
/* Game.Update without input and particles (main.go:231-249): sources, jet, Simulate. */
int fl_scene_update(fl_ctx* ctx, float dt);
int fl_scene_draw_line(fl_ctx* ctx, int x0, int y0, int x1, int y1, int solid); /* drawLine main.go:809 */
int fl_scene_smoke_brush(fl_ctx* ctx, int ci, int cj, int radius, float strength); /* main.go:453 */

/* Gauss-Seidel tile-shape autotuner for one context: times the candidate shapes (sweeps per
 * tile, columns per block, handoff publication) of the temporal-blocked solve on the current
 * state (which it leaves unchanged), re-times the three fastest, and keeps the best in
 * params.block_sweeps, block_cols and handoff_lag. The reference has no counterpart (its
 * solve is a serial loop, fluid.go:188-234). */
int fl_autotune(fl_ctx* ctx, float dt, int* block_sweeps, float* ms);

/* Ensembles. The reference steps one Fluid per call (Simulate, fluid.go:79); an
 * ensemble of independent simulations of uneven sizes is a loop over Fluids there
 * (main.go:245 per frame). A batch advances all of them with one launch per pass.
 * Each member is an ordinary fl_ctx (fl_batch_member) for setup, edits, scenes, views
 * and downloads; it must not be fl_destroy'd (the batch owns it). fl_batch_update
 * applies every member's per-step sources and jet (as fl_scene_update does) and then
 * Simulate(dt) with the batch's parameters, identically to stepping each member on its
 * own. The step is captured as a HIP graph and replayed while dt, the parameters and
 * the members' source programs are unchanged (fl_batch_set_graph(b, 0) turns this off). */
typedef struct fl_batch fl_batch;
int fl_batch_create(int n, const int* widths, const int* heights, float density, float h, fl_batch** out);
int fl_batch_destroy(fl_batch* b);
int fl_batch_size(fl_batch* b, int* n);
int fl_batch_member(fl_batch* b, int index, fl_ctx** out);
int fl_batch_get_params(fl_batch* b, fl_params* out);
int fl_batch_set_params(fl_batch* b, const fl_params* in); /* pressure_iters <= 1024 */
int fl_batch_update(fl_batch* b, float dt);
int fl_batch_sync(fl_batch* b);
int fl_batch_set_graph(fl_batch* b, int on);
int fl_batch_autotune(fl_batch* b, float dt, int* block_sweeps, float* ms);

/* Differentiable rollouts (no reference counterpart: the reference has no gradients).
 * A rollout of `steps` steps from the state the context holds at fl_vjp_create; each step
 * is Game.Update without input (the scene's sources and jet, main.go:231-241), then
 * U,V += F on fluid interior cells (ApplyForce semantics, fluid.go:761-772, for a whole
 * force field F = (fu, fv)), then Simulate (fluid.go:79-109). loss = sum over all cells of
 * (M - target)^2 after the last step. fl_vjp_forward runs it (from the captured state,
 * checkpointing every `checkpoint_every` steps) and leaves the final state in the context;
 * fl_vjp_backward returns dL/dF by the hand-written adjoint of every pass, recomputing
 * each segment from its checkpoint. Arrays hold NumX*NumY floats (index i*NumY + j).
 * With UseMultigrid the V-cycle (fluid.go:560-758) is differentiated too. */
typedef struct fl_vjp fl_vjp;
int fl_vjp_create(fl_ctx* ctx, int steps, int checkpoint_every, fl_vjp** out);
int fl_vjp_destroy(fl_vjp* v);
int fl_vjp_forward(fl_vjp* v, float dt, const float* fu, const float* fv, const float* target, size_t n,
                   double* loss);
int fl_vjp_backward(fl_vjp* v, float* grad_fu, float* grad_fv, size_t n);

#ifdef __cplusplus
}
#endif

#endif /* FLUID_HIP_H */
