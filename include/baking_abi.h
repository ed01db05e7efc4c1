/*
 * baking_abi.h -- C ABI of libbaking_hip.so, the MI355X (gfx950) MLS-MPM
 * dough-baking substep.
 *
 * Every entry point takes plain device pointers, sizes and a HIP stream
 * (hipStream_t passed as void*; NULL = the null stream), enqueues its work
 * and returns without synchronising.  Return codes: 0 on success, a negative
 * value on a launch or argument error (-1 bad argument, -2 HIP error; call
 * bk_last_error() for text).  Physical failure conditions are reported the
 * reference's way, through device int flag arrays with the reference's
 * indices; the caller reads them where the reference raises.
 *
 * Layouts follow the reference's Warp arrays: vec3 = float[3], mat33 =
 * row-major float[9] per element, wp.float64 = double, int = int32.  The
 * per-kernel entries take the reference's dense node arrays (node id
 * (i*res + j)*res + k).  The fused substep keeps its grid "tiled": 4x4x4
 * node blocks stored contiguously, node id ((bi*nb + bj)*nb + bk)*64 +
 * (i&3)*16 + (j&3)*4 + (k&3), nb = ceil(res/4).
 *
 * Reference = bbbpimasheep/baking (Python + Warp 1.18); each entry names the
 * Warp kernel / function it replaces (file:line).
 */
#ifndef BAKING_ABI_H
#define BAKING_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BK_ABI_VERSION 3

/* Flag block of the fused substep (one int32[8] array):
 *   [0] J nonpositive / nonfinite        (reference MpmSimulator.error_flags[0], physics/mpm.py:245)
 *   [1] particle stencil left the grid   (MpmSimulator.error_flags[1], physics/transfer.py:50-52)
 *   [2] Ding porosity domain             (DingMaterial.error_flags[0], ding.py:156-158,210-211)
 *   [3] coupled porosity unresolved      (DingMaterial.error_flags[1], ding.py:193-194)
 *   [4] gas acoustic CFL                 (DingMaterial.error_flags[2], gas.py:34-35)
 *   [5] heat explicit bound              (ThermalWaterTransport.error_flags[0], thermal_water.py:184-185)
 *   [6] water explicit bound             (ThermalWaterTransport.error_flags[1], thermal_water.py:186-188)
 *   [7] reserved                                                                  */
enum { BK_FLAG_J = 0, BK_FLAG_STENCIL = 1, BK_FLAG_POROSITY = 2, BK_FLAG_COUPLED = 3, BK_FLAG_GAS_CFL = 4,
       BK_FLAG_HEAT = 5, BK_FLAG_WATER = 6, BK_FLAG_COUNT = 8 };

/* Scalar parameters of one substep (config.py MaterialConfig/SimulationConfig/DingConfig/TransportConfig). */
typedef struct bk_params {
  float origin[3];
  float dx;
  int32_t res;
  int32_t boundary_cells;
  float gravity[3];
  float dt;
  float cfl_limit;
  int32_t model;              /* 0 = neo_hookean, 1 = ding2019 */
  float mu, lam, density;     /* shear modulus, Lame lambda, material density */
  /* Ding (config.py:109-145) */
  float yield_stress, n0;
  int32_t temperature_enabled, gas_enabled;
  float knots[4], yield_factors[4], viscosity_factors[4];
  float p_ambient, gas_peak, gas_width;
  /* transport (config.py:148-196).  settling: 0 baking; 1 resting (thermal_water.py:296-355 with settling=True:
   * no heat or tray transfer, no surface loss, no conduction or water diffusion, no gas production); 2 rest
   * while the dough's settling controller (bk_settling) is in phase 1, then bake */
  int32_t transport_enabled, water_enabled, settling;
  float T_oven, T_tray, tray_h, loss_coef, rho_w, latent, kappa, diffusivity, s0, water_stress, contact_distance;
  /* sizes */
  int32_t n, n_vertices, n_faces, n_slots;
  /* coupled-porosity solve: 0 = the reference's 64-step bisection (porous_return.py:43-61),
   * 1 = a safeguarded Newton solve of the same equation inside the bisection's bracket */
  int32_t return_solver;
  /* the return-map network (set with bk_workspace_set_surrogate; no reference counterpart) in the coupled
   * porosity solve: 0 not used (the solve starts from the old porosity); 1 guided: every solve starts from the
   * network's prediction and converges to the exact root; 2 stand-in: the prediction, corrected by one Newton
   * step of the porosity equation, is the result when that step is at most 2e-5 (within about a float ulp of
   * the root), else the guided and then the exact solve run */
  int32_t surrogate;
  /* 1: sticky floor -- after the reference's frictionless unilateral walls (physics/grid.py:32-37), every node
   * of the floor's boundary_cells layers (j < boundary_cells) gets velocity 0 (no reference counterpart; 0 keeps
   * the reference's walls) */
  int32_t sticky_floor;
} bk_params;

/* One dough of a batch (no reference counterpart: the reference runs one
 * dough per MpmSimulator).  Doughs are packed in order: dough d owns particles
 * [particle_offset, particle_offset + params.n) and node blocks
 * [block_offset, block_offset + ceil(params.res/4)^3) of the batch grid, with
 * particle_offset and block_offset the running sums over the doughs before it. */
typedef struct bk_dough {
  bk_params params;
  int32_t particle_offset;
  int32_t block_offset;
} bk_dough;

/* Launch tuning of a workspace (no reference counterpart):
 *   tile_chunk      most particles per P2G tile, at most 2^20 (a particle block holding more is split)
 *   particle_block  workgroup size of the per-particle passes (64, 128, 192 or 256)
 *   grid_cap        workgroups of the tile and node-block passes (grid-stride)
 *   deferred_wgs    workgroups of the deferred return-map pass where it runs beside
 *                   the next substep (pipelined graphs; it strides over its queue);
 *                   0: pipelined graphs are the plain substep                      */
typedef struct bk_tuning {
  int32_t tile_chunk;
  int32_t particle_block;
  int32_t grid_cap;
  int32_t deferred_wgs;
} bk_tuning;

/* Per-particle state of one step (physics/state.py ParticleState, ding.py DingState, thermal_water.py TransportState). */
typedef struct bk_state {
  float *x, *v, *F, *C;                                    /* N x 3, N x 3, N x 9, N x 9 */
  float *Fe, *Fp, *tau, *residual, *porosity, *T_mat;      /* DingState (f32) */
  int32_t* gelled;
  float *yield_stress, *viscosity, *porosity_residual, *density;
  double *moles, *pressure;                                /* gas (f64) */
  double *T, *W, *s, *losses, *exchanges;                  /* TransportState (f64) */
  float* masses;                                           /* TransportState.particle_masses */
} bk_state;

/* Static scene data (set once per simulator). */
typedef struct bk_scene {
  const float *V0, *m0;                      /* reference volumes, initial masses */
  const double *capacities, *dry;            /* thermal_water.py:272-274 */
  const float* visc_param;                   /* DingMaterial.viscosity_parameter [1] */
  const float* prod_param;                   /* DingMaterial.production_parameter [1] */
  const float* heat_param;                   /* ThermalWaterTransport.heat_transfer_parameter [1] */
  /* mesh reconstruction CSR (geometry/mesh_tracking.py:41-75) */
  const float* ref_vertices;                 /* V x 3 */
  const int32_t *bind_offsets, *bind_particles;
  const float *bind_weights, *bind_ref_offsets;
  /* boundary faces (thermal_water.py:51-74) grouped by owner slot */
  const int32_t* faces;                      /* F x 3 */
  const int32_t *slot_face_offsets, *slot_faces, *slot_owner, *particle_slot;
  /* batch only: dough of every particle (NULL for a batch of one).  In a
   * batch, the per-dough parameters above (visc/prod/heat_param) hold one value
   * per dough and every mesh, face, slot and particle index is global. */
  const int32_t* dough_of;
} bk_scene;

/* Tiled grid of the fused substep (GridState + transport workspace); a
 * batch grid holds every dough's blocks, 64 nodes per block. */
typedef struct bk_grid {
  float *m, *p, *v;                          /* node mass, momentum (x3), velocity (x3) */
  double *V, *C, *E, *W, *H, *L, *X;         /* thermal P2G: volume, capacity, energy, water, heat, loss, exchange */
  double *dT, *dW;                           /* diffusion increments: temperature change, and water change divided by
                                                the node's water (0 where it has none) */
} bk_grid;

typedef struct bk_workspace bk_workspace;

/* ---------------------------------------------------------------- library */
int bk_abi_version(void);
const char* bk_last_error(void);
int bk_device_info(char* name, int name_len, int* cu_count, int* xcd_count);

/* Workspace: the only library-owned device memory (sort scratch, tile and
 * active-block lists, per-step stress and surface buffers, flags). */
bk_workspace* bk_workspace_create(int32_t max_particles, int32_t res, int32_t n_vertices, int32_t n_slots);
/* Workspace for batches of up to max_doughs doughs whose grids total at most
 * total_blocks node blocks (no reference counterpart). */
bk_workspace* bk_workspace_create_batch(int32_t max_particles, int32_t total_blocks, int32_t max_doughs, int32_t n_vertices,
                                        int32_t n_slots);
void bk_workspace_destroy(bk_workspace* ws);
/* Upload the dough table used by bk_batch_substep (synchronises the device
 * first).  A table with another grid layout than the last one starts from a
 * zeroed batch grid. */
int bk_workspace_set_doughs(bk_workspace* ws, const bk_dough* table, int32_t n_doughs);
int bk_workspace_set_tuning(bk_workspace* ws, const bk_tuning* tuning);
int bk_workspace_get_tuning(bk_workspace* ws, bk_tuning* tuning);
int32_t* bk_workspace_flags(bk_workspace* ws);          /* device int32[BK_FLAG_COUNT * max_doughs], dough-major */
float* bk_workspace_vertices(bk_workspace* ws);         /* device float[V*3]: the last substep's positions of the
                                                           vertices its boundary faces use (the others are not updated) */
/* Copies {tiles, active node blocks, deferred return maps} of the last substep; synchronises. */
int bk_workspace_counts(bk_workspace* ws, int32_t* out3, void* stream);
/* Device-to-device copies (any may be null) of the last substep's cell keys in particle order (dough's first block
 * << 6 | its 4x4x4-block-major cell key, the key of physics/transfer.py:16-33 on the tiled grid), the keys in sorted
 * order and the sort permutation (sorted position -> particle): a stable sort of the keys. n <= the workspace's
 * particle capacity; enqueued on `stream`. */
int bk_workspace_sort_order(bk_workspace* ws, uint32_t* keys, uint32_t* sorted_keys, int32_t* order, int32_t n, void* stream);
/* Blocking copy of the flag block to host int32[8] after the work already on `stream`; and the reverse. */
int bk_workspace_read_flags(bk_workspace* ws, int32_t* host8, void* stream);
int bk_workspace_write_flags(bk_workspace* ws, const int32_t* host8, void* stream);
/* The same for the first n_doughs flag blocks (int32[8 * n_doughs], dough-major). */
int bk_workspace_read_flags_all(bk_workspace* ws, int32_t* host, int32_t n_doughs, void* stream);
int bk_workspace_write_flags_all(bk_workspace* ws, const int32_t* host, int32_t n_doughs, void* stream);
/* Stopped doughs (no reference counterpart; the reference stops the whole run at an error).  With freezing on
 * (off by default; batches turn it on), a dough with any flag raised stops at the start of the next substep:
 * its particles get no key (no tile, grid or material work) and its state is the one before the substep that
 * raised the flag -- on the first stopped substep nothing is written, so the output buffer keeps the last
 * accepted state, and later substeps copy the state through.  bk_workspace_freeze_modes reads (write = 0) or
 * writes host int32[2 * n_doughs] pairs per dough: the mode of the last substep (0 ran, 1 first stopped
 * substep, 2 copying) and its age, the substeps run since it stopped.  A dough with a flag and mode 0 raised
 * it in the last substep run: its accepted state is in the other buffer (the caller restores it and writes
 * mode 1).  Writing flags also sets each written dough's mode (0 clear, 1 flagged) and age 0.  Both calls
 * synchronise `stream`. */
int bk_workspace_set_freezing(bk_workspace* ws, int32_t on);
int bk_workspace_freeze_modes(bk_workspace* ws, int32_t* host, int32_t n_doughs, int32_t write, void* stream);

/* ---------------------------------------------------------------- settling controller
 * The rest before baking (physics/mpm.py:57-109 settling_steps/IsSettling/UpdateSettling with
 * physics/quasistatic.py:11-21 SettlingMetrics), one record per dough kept in the workspace.  While a dough's
 * params.settling is 2, every substep of bk_substep / bk_batch_substep (and their graphs) rests if the record's
 * phase is 1, and ends with the controller's kernels, which reduce the three settling metrics over the dough's particles
 * (auto mode) and advances the record exactly as UpdateSettling advances the simulator: the host reads it only
 * when it wants the outcome, never per substep.  Auto mode reads state.Fp (the Ding model).  On a timeout the
 * record keeps phase 1 and sets reason 3; the caller raises the reference's RuntimeError. */
typedef struct bk_settling {
  /* configuration (TransportConfig, config.py:150-155) */
  int32_t auto_mode;          /* settling_auto */
  int32_t settling_steps;     /* ceil(settling_duration_s / time_step - 1e-10): fixed duration, or the timeout */
  double time_step;
  double min_duration_s, hold_duration_s, speed_tolerance, rate_tolerance;
  /* state (the simulator attributes of physics/mpm.py:150-157) */
  int32_t phase;              /* 1 while resting (IsSettling), 0 once baking has started */
  int32_t reason;             /* settling_reason: 0 pending, 1 fixed_duration, 2 quasi_static, 3 timeout */
  int64_t step_count;         /* substeps run, counted after each one as step_count is */
  int64_t reference_step;     /* settling_reference_step */
  int64_t quiet_steps;        /* settling_quiet_steps */
  int64_t baking_start_step;  /* _baking_start_step, -1 while resting */
  double plastic_rate;        /* settling_plastic_rate */
  float diagnostics[3];       /* settling_diagnostics: max speed, max strain rate, max plastic rate */
  int32_t ref_slot;           /* half of the reference buffer holding settling_reference_plastic */
  int32_t copy_pending;       /* the last substep closed a plastic-rate window: the next one takes its input Fp as
                                 the reference (and stores it in the ref_slot half) */
  int32_t reserved;
} bk_settling;

/* Register the caller-owned reference buffer, device float[2][max_particles][9] in batch particle order: the
 * two halves of settling_reference_plastic (record.ref_slot names the live one; a closed window writes the
 * other half and flips ref_slot, so restoring a saved record restores its reference).  Synchronises the device.
 * Auto mode needs it. */
int bk_workspace_settling_reference(bk_workspace* ws, float* reference);
/* Write the whole records of doughs [first, first + count) after the work already on `stream`; _get copies them
 * back (blocking). */
int bk_workspace_settling_set(bk_workspace* ws, int32_t first, int32_t count, const bk_settling* records, void* stream);
int bk_workspace_settling_get(bk_workspace* ws, int32_t first, int32_t count, bk_settling* records, void* stream);

/* ---------------------------------------------------------------- return-map surrogate (no reference counterpart)
 * A 7 -> 4 x 64 SiLU -> 4 network standing in for the start of the coupled (gas)
 * Ding return map (physics/materials/ding.py:168-221 with porous_return.py:43-61).
 * Rows: 7 input floats + 1 unused per particle (principal trial log-strains in
 * descending order, temperature of the yield/viscosity factors, saturation,
 * porosity before the substep, |Fp - I|_F).  Outputs: 4 floats (returned
 * log-strains in the same order, porosity increment).  Parameters: FP32 vector
 * of bk_mlp_theta_count() floats, W1[64][8] b1 W2[64][64] b2 W3 b3 W4 b4 W5[4][64]
 * b5[4].  The inference blob (bk_mlp_blob_bytes() bytes: normalisation, biases,
 * BF16 weights in matrix-core operand order) is built by the host; its layout
 * is given in baking_amd/csrc/bk_surrogate.h. */
int32_t bk_mlp_theta_count(void);
int32_t bk_mlp_blob_bytes(void);
int32_t bk_mlp_grad_chunk(void);
/* BF16 matrix-core inference of n rows (rows: n x 8 floats, out: n x 4 floats). */
int bk_mlp_infer(const void* blob, const float* rows, float* out, int32_t n, void* stream);
/* FP32 forward of normalised rows (n x 8) to normalised outputs (n x 4); z: scratch of 4 * n * 64 floats. */
int bk_mlp_forward(const float* theta, const float* rows, float* out, int32_t n, float* z, void* stream);
/* Gradient of the mean squared error between the FP32 forward of normalised rows and normalised targets
 * (n x 4) with respect to theta; loss += the error (double on device).  Scratch: z and d of 4 * n * 64
 * floats, y and d_out of n * 4, partial of ceil(n / bk_mlp_grad_chunk()) * bk_mlp_theta_count(). */
int bk_mlp_gradient(const float* theta, const float* rows, const float* targets, int32_t n, float* z, float* y, float* d,
                    float* d_out, float* partial, float* grad, double* loss, void* stream);
/* One Adam update (bias-corrected, step >= 1) of count parameters. */
int bk_adam_step(float* theta, const float* grad, float* m, float* v, int32_t count, float lr, float beta1, float beta2, float eps,
                 int32_t step, void* stream);
/* Use the network (blob in device memory, kept alive by the caller; NULL: none) in the substeps of
 * doughs with params.surrogate = 1.  Synchronises the device. */
int bk_workspace_set_surrogate(bk_workspace* ws, const void* blob);
/* Emit one row of 12 floats (7 inputs, the 4 exact outputs, 1 unused) per coupled return map of the
 * following substeps into buffer, up to capacity rows (NULL: stop); resets the emitted count. */
int bk_workspace_set_samples(bk_workspace* ws, float* buffer, int32_t capacity);
/* Counters summed over the substeps since the last reset: {porosity-equation evaluations, guided solves
 * accepted, guided solves handed to the exact solve, queued return maps, rows emitted}; synchronises. */
int bk_workspace_stats(bk_workspace* ws, int64_t* out5, int32_t reset, void* stream);

/* ---------------------------------------------------------------- sort (build-only; no reference counterpart)
 * Cell key of every particle: flatten_node(stencil_base((x-o)/dx)) of
 * physics/transfer.py:16-33 (tiled=0), or the same cell numbered block-major
 * over 4x4x4 cell blocks (tiled=1).  Off-grid particles get key 0 and set
 * flags[1]. */
int bk_cell_keys(const float* x, int32_t n, const float* origin3_host, float dx, int32_t res, int32_t tiled, uint32_t* keys,
                 int32_t* flags, void* stream);
/* Stable LSD radix sort of (key, value) pairs on the low key_bits bits; values
 * = NULL means 0..n-1.  Outputs go to keys_out/values_out. */
int bk_sort_pairs(bk_workspace* ws, const uint32_t* keys_in, const int32_t* values_in, uint32_t* keys_out, int32_t* values_out,
                  int32_t n, int32_t key_bits, void* stream);
/* No counterpart in the reference.  bk_sort_pairs without a workspace, for callers that cannot allocate (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc): the same sort and the same results, with its
 * scratch in a static device array (16.5 MiB per GPU, held until the process ends) instead of a workspace.  It
 * calls no allocating routine.  At most bk_sort_pairs_noalloc_capacity() = 2^20 pairs; more return -1 with the
 * reason in bk_noalloc_last_error().  Every call on one GPU shares that GPU's scratch: calls that overlap on one
 * GPU (concurrent streams or threads) race and must be serialised by the caller. */
int bk_sort_pairs_noalloc(const uint32_t* keys_in, const int32_t* values_in, uint32_t* keys_out, int32_t* values_out, int32_t n,
                          int32_t key_bits, void* stream);
int32_t bk_sort_pairs_noalloc_capacity(void);
const char* bk_noalloc_last_error(void);

/* ---------------------------------------------------------------- fused substep
 * One MpmSimulator.Step (physics/mpm.py:163-225) in forward mode: transport
 * pre-pass, gas and water stress, P2G, grid update, G2P, Ding return map, gas
 * refresh and saturations, from state `in` to state `out`.  The tiled grid
 * keeps this step's values afterwards (inactive blocks are zero). */
int bk_substep(const bk_params* params, const bk_scene* scene, const bk_state* in, bk_state* out, bk_grid* grid,
               bk_workspace* ws, void* stream);
/* The same substep captured once into a HIP graph (on a library-owned
 * capture stream) and replayed with one launch: the state/grid pointers and
 * parameters are frozen at capture, so keep one graph per ping-pong direction
 * and recapture when a buffer or parameter changes. */
typedef struct bk_graph bk_graph;
bk_graph* bk_substep_graph_create(const bk_params* params, const bk_scene* scene, const bk_state* in, bk_state* out,
                                  bk_grid* grid, bk_workspace* ws);
/* A graph keeps the dough table it was captured with and re-uploads it at
 * launch if another table was set on its workspace since. */
int bk_graph_launch(bk_graph* graph, void* stream);
void bk_graph_destroy(bk_graph* graph);
/* One substep of every dough of the uploaded table, each as MpmSimulator.Step
 * (physics/mpm.py:163-225) of that dough alone, in one launch per stage; flags
 * are raised in the dough's own block.  bk_substep is this with a table of one. */
int bk_batch_substep(bk_workspace* ws, const bk_scene* scene, const bk_state* in, bk_state* out, bk_grid* grid, void* stream);
bk_graph* bk_batch_graph_create(bk_workspace* ws, const bk_scene* scene, const bk_state* in, bk_state* out, bk_grid* grid);
/* Pipelined substeps (no reference counterpart; a schedule of the same
 * substeps): a graph whose transport pre-pass and P2G set-up (sort, mesh, thermal
 * P2G, diffusion) run beside the PREVIOUS substep's deferred return maps, which
 * write only Ding, gas and saturation fields of the particles they finish.  The
 * substep's own deferred return maps are left pending, so a run of graph
 * launches on alternating ping-pong directions must end with bk_*_finish on the
 * last substep's in / out (a no-op when nothing is pending) before its state,
 * flags or statistics are read.  Flags of a pending deferred pass are merged at
 * the next substep's join; a dough it flags stops there exactly as at the start
 * of that substep.  Without pore gas, or with the settling controller, the
 * graph is the plain substep. */
bk_graph* bk_substep_graph_create_pipelined(const bk_params* params, const bk_scene* scene, const bk_state* in, bk_state* out,
                                            bk_grid* grid, bk_workspace* ws);
bk_graph* bk_batch_graph_create_pipelined(bk_workspace* ws, const bk_scene* scene, const bk_state* in, bk_state* out,
                                          bk_grid* grid);
/* `steps` substeps (1-64) captured into one graph, alternating a -> b -> a ... (plain or, with pipelined != 0,
 * pipelined): one launch replays them all, saving the per-launch gap between substeps.  An even count ends on
 * `a`.  After a pipelined one, bk_*_finish takes the last substep's in / out as above. */
bk_graph* bk_substep_graph_create_steps(const bk_params* params, const bk_scene* scene, const bk_state* a, bk_state* b,
                                        bk_grid* grid, bk_workspace* ws, int32_t steps, int32_t pipelined);
bk_graph* bk_batch_graph_create_steps(bk_workspace* ws, const bk_scene* scene, const bk_state* a, bk_state* b, bk_grid* grid,
                                      int32_t steps, int32_t pipelined);
int bk_substep_finish(const bk_params* params, const bk_scene* scene, const bk_state* in, bk_state* out, bk_workspace* ws,
                      void* stream);
int bk_batch_finish(bk_workspace* ws, const bk_scene* scene, const bk_state* in, bk_state* out, void* stream);
/* Per-stage timing of bk_substep with HIP events on its stream: begin() arms
 * recording for the next max_steps substeps; end() waits for them and returns
 * the summed milliseconds per stage (bk_profile_stage_count() entries, named
 * by bk_profile_stage_name). */
int bk_profile_begin(bk_workspace* ws, int32_t max_steps);
int bk_profile_end(bk_workspace* ws, float* ms_per_stage, int32_t* steps_recorded);
int32_t bk_profile_stage_count(void);
const char* bk_profile_stage_name(int32_t stage);
/* Acceptance aggregates as f64 sums per dough (experiments/recording.py:46-112 Measure and the
 * CheckState statistics of physics/mpm.py:292-308), out[8 * d + k] on the device:
 * k = 0 sum V0 det F (particle volume), 1 sum capacity * T, 2 sum capacity (mean T = [1] / [2]),
 * 3 sum W (water), 4 sum water losses, 5 sum heat exchanges, 6 sum V0 * gas moles, 7 sum particle mass.
 * NULL state fields contribute 0 (mass falls back to scene.m0); scene.dough_of NULL = one dough. */
int bk_measure(const bk_scene* scene, const bk_state* state, int32_t n, int32_t n_doughs, double* out, void* stream);
/* CheckState in one pass on the device (physics/mpm.py:243-308 with DingMaterial.GasWaveSpeed, ding.py:285-292):
 * out[BK_CHECK_COUNT] (device doubles) after the work on `stream`:
 *   NONFINITE_STATE       1 if any x, v, F or C component is not finite, else 0
 *   MIN_J, MAX_J          extremes of det F (float, as numpy's det of the float32 state)
 *   OUTSIDE_STENCIL       1 if any (x - origin) / dx < 0.5 or >= res - 1.5 (double, origin/dx from grid_origin_dx)
 *   MAX_SPEED             max |v|
 *   GAS_WAVE              max sqrt((lam + 2 mu + p (1 + (1 - n0 + n) / n)) J / rho) (Ding with gas; else -inf)
 *   WATER_WAVE_SQ         max water_stress s J / rho (Ding with water transport; else -inf)
 *   NONFINITE_DING        1 if any Fe or Fp component is not finite (Ding)
 *   MAX_RESIDUAL          max return residual (Ding; else -inf)
 *   NONFINITE_VERTICES    1 if any of the n_vertices vertex positions is not finite
 *   INVERTED_TETS         tets whose det([v1-v0, v2-v0, v3-v0]) / 6 <= 0 (double)
 *   VOLUME, MASS          sum V0 det F, sum of the masses (state.masses with transport, else scene.m0), f64
 *   GRID_MASS             sum of node_masses[0, node_count), f64
 *   KINETIC_ENERGY        0.5 sum m |v|^2, f64
 * The sums are combined in a fixed order (the same on every run).  grid_origin_dx = host double[4]: the
 * configuration's grid origin and spacing.  scratch = bk_check_scratch_size(...) device doubles. */
enum { BK_CHECK_NONFINITE_STATE = 0, BK_CHECK_MIN_J, BK_CHECK_MAX_J, BK_CHECK_OUTSIDE_STENCIL, BK_CHECK_MAX_SPEED,
       BK_CHECK_GAS_WAVE, BK_CHECK_WATER_WAVE_SQ, BK_CHECK_NONFINITE_DING, BK_CHECK_MAX_RESIDUAL,
       BK_CHECK_NONFINITE_VERTICES, BK_CHECK_INVERTED_TETS, BK_CHECK_VOLUME, BK_CHECK_MASS, BK_CHECK_GRID_MASS,
       BK_CHECK_KINETIC_ENERGY, BK_CHECK_COUNT };
int64_t bk_check_scratch_size(int32_t n, int64_t node_count, int32_t n_vertices, int32_t n_tets);
int bk_check_state(const bk_params* P, const bk_scene* scene, const bk_state* state, const double* grid_origin_dx,
                   const float* node_masses, int64_t node_count, const float* vertices, int32_t n_vertices, const int32_t* tets,
                   int32_t n_tets, double* scratch, double* out, void* stream);
/* Convert the tiled grid's mass/velocity/momentum to the reference's dense layout. */
int bk_grid_untile(const float* tiled, float* dense, int32_t res, int32_t components, void* stream);
int bk_grid_untile_f64(const double* tiled, double* dense, int32_t res, void* stream);

/* ---------------------------------------------------------------- per-kernel entries (reference layouts)
 * Same arguments, order and semantics as the Warp kernel named; n = launch dim. */
/* physics/transfer.py:36-70 */
int bk_ParticleToGrid(const float* x, const float* v, const float* C, const float* tau, const float* masses, const float* V0,
                      float* node_masses, float* node_momenta, const float* origin3_host, float dx, int32_t res, float dt,
                      int32_t* flags, int32_t n, void* stream);
/* physics/grid.py:20-38 */
int bk_UpdateGrid(const float* node_masses, const float* node_momenta, float* node_velocities, int32_t res, int32_t boundary_cells,
                  const float* gravity3_host, float dt, int64_t node_count, void* stream);
/* physics/transfer.py:73-109 */
int bk_GridToParticle(const float* x_in, const float* F_in, const float* node_velocities, float* x_out, float* v_out, float* C_out,
                      float* F_out, const float* origin3_host, float dx, int32_t res, float dt, int32_t* flags, int32_t n,
                      void* stream);
/* physics/quasistatic.py:11-21; metrics = float[3] running maxima (the caller zeroes them) */
int bk_SettlingMetrics(const float* velocities, const float* affine, const float* old_plastic, const float* new_plastic,
                       float* metrics, float time_step, int32_t n, void* stream);
/* physics/materials/elastic.py:13-26 */
int bk_ComputeElasticStress(const float* F, float* tau, float mu, float lam, int32_t* flags, int32_t n, void* stream);
/* physics/materials/ding.py:140-221; ding flags = int32[3] */
int bk_UpdateDing(const float* old_Fe, const float* old_porosity, const int32_t* old_gelled, const float* F_total,
                  const float* C_affine, float* Fe, float* Fp, float* tau, float* residuals, float* porosity, float* temperatures,
                  int32_t* gelled, float* yield_stresses, float* viscosities, const float* viscosity_parameter, int32_t* flags,
                  float* porosity_residuals, int32_t coupled, float mu, float lam, float yield_stress, float n0,
                  int32_t temperature_enabled, const float* knots4_host, const float* yield4_host, const float* visc4_host,
                  float dt, const float* local_temperatures, int32_t n, void* stream);
/* physics/materials/gas.py:16-35 */
int bk_CombineGasStress(const float* solid, const float* F, const float* porosity, const double* moles, const float* T,
                        float* total, int32_t* flags, float p_ambient, float n0, const float* density, float elastic_modulus,
                        float dt, float dx, float cfl, int32_t n, void* stream);
/* physics/materials/gas.py:38-45 */
int bk_RefreshGas(const double* old_moles, double* new_moles, const float* F, const float* porosity, double* pressures,
                  const float* T, int32_t n, void* stream);
/* physics/thermal_water.py:77-92 */
int bk_SurfaceAreas(const float* vertices, const int32_t* faces, const int32_t* owners, double* exposed, double* contact,
                    float floor_height, float contact_distance, int32_t n_faces, void* stream);
/* physics/thermal_water.py:95-137 */
int bk_ScalarToGrid(const float* x, const float* F, const float* V0, const double* capacities, const double* T, const double* W,
                    const double* exposed, const double* contact, const float* heat_parameter, float T_oven, float T_tray,
                    float tray_h, float loss_coef, float rho_w, float latent, int32_t water_enabled, double* nV, double* nC,
                    double* nE, double* nW, double* nH, double* nL, double* nX, double* particle_heat, double* particle_loss,
                    const float* origin3_host, float dx, int32_t res, int32_t n, void* stream);
/* physics/thermal_water.py:140-147 */
int bk_GridFields(const double* V, const double* C, const double* E, const double* W, double* T, double* S, float rho_w,
                  int64_t node_count, void* stream);
/* physics/thermal_water.py:150-190; errors = int32[2] */
int bk_DiffuseGrid(const double* V, const double* C, const double* T, const double* S, const double* W, const double* heat,
                   const double* loss, const double* exchange, double* dT, double* dW, int32_t* errors, float conductivity,
                   float diffusivity, float rho_w, int32_t water_enabled, float dx, int32_t res, float dt, int64_t node_count,
                   void* stream);
/* physics/thermal_water.py:193-229 */
int bk_ScalarToParticles(const float* x, const double* T_old, const double* W_old, const double* L_old, const double* X_old,
                         const double* dry, const float* V0, const float* F, const double* node_water, const double* dT,
                         const double* dW, const double* particle_heat, const double* particle_loss, double* T_new, float* T_mat,
                         double* W_new, double* L_new, double* X_new, float* m_new, float* densities, double* s_new, float rho_w,
                         const float* origin3_host, float dx, int32_t res, float dt, int32_t n, void* stream);
/* physics/thermal_water.py:232-245 */
int bk_LocalGas(const double* old_moles, const double* T_old, const double* T_new, double* new_moles, double* pressures,
                const float* F, const float* porosity, const float* parameter, float density, float peak, float width, float dt,
                int32_t produce, int32_t n, void* stream);
/* physics/thermal_water.py:248-253 */
int bk_WaterStress(const float* F, const double* saturations, const float* old_stress, float* new_stress, float s0,
                   float stiffness, int32_t n, void* stream);
/* physics/thermal_water.py:256-260 */
int bk_Saturations(const double* W, const float* V0, const float* F, double* saturations, float rho_w, int32_t n, void* stream);
/* geometry/mesh_tracking.py:13-38 */
int bk_ReconstructMeshVertices(const float* x, const float* F, const float* ref_vertices, const int32_t* offsets,
                               const int32_t* particles, const float* weights, const float* ref_offsets, float* out,
                               int32_t n_vertices, void* stream);
/* Scalar functions evaluated on device for arrays of inputs:
 * physics/materials/gas.py:11-13, ding.py:68-74, porous_return.py:43-61 and 34-40.
 * bk_solve_porosity: args6 = (q, p, dev_factor, mean_scale, exponent, initial_porosity) per element;
 * solver 0 = the reference's bisection, 1 = the bracketed Newton solve of bk_params.return_solver. */
int bk_gas_pressure(const double* moles, const float* T, const float* J, const float* porosity, double* out, int32_t n, void* stream);
int bk_porosity_coefficient(const float* porosity, float* out, int32_t n, void* stream);
int bk_solve_porosity(const float* args6, const double* plastic_volume, float* out_porosity, double* out_prediction,
                      int32_t solver, int32_t n, void* stream);

/* ---------------------------------------------------------------- adjoint (reverse mode)
 * Replaces the reference's wp.Tape over MpmSimulator.Step in history mode
 * (physics/mpm.py:18-22, 163-225; tests/test_warp_autodiff.py, tests/test_ding.py:81-170,
 * tests/test_gas.py:131-160, tests/test_transport.py:96-121,
 * tests/test_ding_thermal.py:166-191): adjoints of one substep with the
 * neo-Hookean or the Ding model (pore gas and the coupled porosity return
 * included) and the heat and water transport (mesh areas, Robin terms,
 * thermal P2G, FV diffusion, scalar G2P, gas source).  Adjoint arrays are
 * FP64 and particle-major like the state. */
typedef struct bk_adjoint {
  double *x, *v, *C, *F;                   /* N x 3, N x 3, N x 9, N x 9 */
  double *Fe, *tau, *porosity;              /* Ding model only: N x 9, N x 9, N */
  double* moles;                            /* Ding with pore gas only: N */
  double *T, *W, *losses, *exchanges;       /* transport only: N each */
  /* transport only, static per scene: the mesh bindings (scene.bind_*) indexed by particle --
   * rev_offsets[N+1], rev_bindings / rev_vertices[B] (binding index and its vertex, grouped by
   * bind_particles) -- and anchor_coef[V] = 1 - sum of the vertex's binding weights */
  const int32_t *rev_offsets, *rev_bindings, *rev_vertices;
  const double* anchor_coef;
} bk_adjoint;
/* In: adj = the adjoint of the substep's output state `out`; the grid and the
 * workspace `ws` as the forward substep from `in` to `out` left them, with
 * nothing run on `ws` since (single dough, tiled nodes; the adjoint reuses
 * that substep's sorted P2G tiles, active blocks, vertices and areas).
 * Out: adj = the adjoint of `in`; param_grad[0] += the gradient of the Ding
 * viscosity parameter (scene.visc_param), [1] += the gas production
 * parameter's (scene.prod_param), [2] += the heat-transfer parameter's
 * (scene.heat_param), [3] += the oven temperature's (params.T_oven, the
 * oven_temperature of physics/thermal_water.py:100, passed at :325); param_grad
 * holds 4 doubles.  scratch: bk_adjoint_scratch_size(params) doubles, all
 * zero before the first call; each call leaves it ready for the next. */
int64_t bk_adjoint_scratch_size(const bk_params* params);
int bk_adjoint_substep(const bk_params* params, const bk_scene* scene, const bk_state* in, const bk_state* out, const bk_grid* grid,
                       bk_workspace* ws, const bk_adjoint* adj, double* scratch, double* param_grad, void* stream);
/* Adjoint of UpdateDing without pore gas (physics/materials/ding.py:140-221; arguments as bk_UpdateDing):
 * given adjoints of its outputs Fe, tau and porosity (any may be NULL = 0), ADDS the adjoints of
 * old_Fe, old_porosity, F_total, C_affine (NULL = not wanted) and of the viscosity parameter (one double). */
int bk_UpdateDing_adjoint(const float* old_Fe, const float* old_porosity, const int32_t* old_gelled, const float* F_total,
                          const float* C_affine, const float* temperatures, const float* viscosity_parameter, const double* adj_Fe,
                          const double* adj_tau, const double* adj_porosity, double* adj_old_Fe, double* adj_old_porosity,
                          double* adj_F_total, double* adj_C_affine, double* adj_viscosity_parameter, float mu, float lam,
                          float yield_stress, float n0, int32_t temperature_enabled, const float* knots4_host,
                          const float* yield4_host, const float* visc4_host, float dt, int32_t n, void* stream);
/* Implicit-function adjoint of the overstress solve (ding.py:90-100): args4 = (q2, p2, dev_factor,
 * mean_factor) per element, gbar = adjoint of the overstress; out4 = adjoints of the four arguments. */
int bk_overstress_adjoint(const double* args4, const double* gbar, double* out4, int32_t n, void* stream);
/* Implicit-function adjoint of the coupled porosity solve (porous_return.py:64-100): arguments as
 * bk_solve_porosity, gbar = adjoint of the solved porosity; out7 = adjoints of (q, p, dev_factor,
 * mean_scale, exponent, initial_porosity, plastic_volume) at the bracketed solve's root. */
int bk_solve_porosity_adjoint(const float* args6, const double* plastic_volume, const double* gbar, double* out7, int32_t n,
                              void* stream);

#ifdef __cplusplus
}
#endif
#endif /* BAKING_ABI_H */
