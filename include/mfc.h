/* mfc.h — C ABI of the MI355X-native 2D SPH orthogonal-cutting solver.
 *
 * The reference (iwf-inspire/mfree_iwf-ul_cut_gpu) has no FFI: main() calls C++ free
 * functions and classes directly. The entry points below are the C-ABI form of that
 * call sequence, one per reference interface:
 *
 *   setup_ref_cut(ny, &grid, rake, clear, chamfer, speed, feed)  benchmarks.h:48,
 *       benchmarks.cpp:255-537                         -> mfc_config_ref_cut + mfc_create
 *   setup_rings(nbox, &grid) / setup_impact(nbox, &grid)  benchmarks.h:42,45
 *                                                      -> mfc_config_rings / mfc_config_impact
 *                                                         + mfc_create
 *   new leap_frog(N, num_cell)                         leap_frog.h:49, leap_frog.cu:157-166
 *                                                      -> part of mfc_create
 *   leap_frog::step(particles, grid)                   leap_frog.h:48, leap_frog.cu:94-155
 *       + global_step++ (mfree_gpu.cpp:308-309)        -> mfc_step
 *   float2_t report_force(particle_gpu*)               output.h:26, output.cu:31-39
 *                                                      -> mfc_tool_force
 *   cudaMemcpy of particle_gpu members (debug.cpp:33-41, vtk_writer.cpp:59-71)
 *                                                      -> mfc_get_field
 *   grid_base::get_cells(particles, start, end)        grid.h:280, grid_gpu_green.cu:207-210
 *                                                      -> mfc_get_cells
 *   tool::get_segments / get_fillet (tool.h:75-76)     -> mfc_get_tool
 *   tool_wear::eval_usui(particles, dt, min, max, avg) tool_wear.h:53, tool_wear.cu:20-60
 *       + get_accum_wear (tool_wear.cu:62-64)          -> mfc_tool_wear
 *   vtk_writer_write(particles, step)                  vtk_writer.h, vtk_writer.cpp:20-176
 *                                                      -> mfc_write_vtk
 *   vtk_writer_write(tool, step)                       vtk_writer.cpp:183-263
 *                                                      -> mfc_write_tool_vtk
 *   (no counterpart: the reference's only state dump is the text dump_state, debug.cpp:20-64)
 *                                                      -> mfc_save / mfc_load
 *
 * Conventions: every function returns 0 on success and a negative MFC_E_* code on
 * failure (mfc_last_error() gives the message of the calling thread's last failure);
 * nothing exits or asserts. The handle owns all device memory; the caller owns host
 * buffers. All constants live in the handle, so several handles may coexist in one
 * process. One handle is driven by one host thread at a time.
 *
 * Units are the reference's "bomb units" (benchmarks.h:47): cm, microseconds, grams,
 * Mbar; temperatures in K.
 */
#ifndef MFC_H_
#define MFC_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MFC_OK            0
#define MFC_E_INVALID    -1   /* bad argument or configuration */
#define MFC_E_HIP        -2   /* HIP runtime failure */
#define MFC_E_NOMEM      -3   /* allocation failed */
#define MFC_E_RANGE      -4   /* size or index out of range */
#define MFC_E_STATE      -5   /* call not valid in the handle's current state */

/* thermal solver, compile-time switch in the reference (types.h:94-95) */
#define MFC_THERMAL_BROOKSHAW 0
#define MFC_THERMAL_PSE       1

/* flow stress of the radial return's secant (plasticity.cuh:72, m_s_yield) */
#define MFC_JC_PLAIN 0  /* Johnson-Cook (plasticity.cuh:30-46), the reference's choice */
#define MFC_JC_TANH  1  /* tanh-modified Johnson-Cook after Calamaz 2008 (plasticity.cuh:48-70) */

/* cell grid (grid.h): both sort the particles stably by cell id and give the same cell ranges */
#define MFC_GRID_GREEN   0  /* grid_gpu_green (grid_gpu_green.cu): ranges from the counting sort */
#define MFC_GRID_ROTHLIN 1  /* grid_gpu_rothlin (grid_gpu_rothlin.cu:28-250): ranges rebuilt from the
                               sorted ids by a unique / scatter / scan pass, and, as its sort does
                               (:183-247), the friction force ft is not carried along with its
                               particle: contact reads the history at the sorted index */

/* arithmetic of the step */
#define MFC_PREC_FP64   0   /* plain FP64, bit-exact with the oracle */
#define MFC_PREC_FP32C  1   /* FP32 neighbour sums with Kahan compensation, FP64 islands */

/* lattice rule */
#define MFC_LATTICE_REF      0  /* setup_ref_cut: lo_y = hi_y - 3 feed, dx = (hi_y-lo_y)/(ny-1),
                                   nx = (hi_x-lo_x)/dx (benchmarks.cpp:266-273) */
#define MFC_LATTICE_EXPLICIT 1  /* nx x ny particles at spacing dx, top row at hi_y */
#define MFC_LATTICE_RINGS    2  /* setup_rings (benchmarks.cpp:155-253): two rings, radii 0.03 and
                                   0.04, centres at -+0.041, on an ny x ny lattice of spacing
                                   0.08/(ny-1), moving at +-200 towards each other; fixed grid
                                   (-0.2,-0.08)-(0.2,0.08); no tool; dt from cfl and speed */
#define MFC_LATTICE_IMPACT   3  /* setup_impact (benchmarks.cpp:23-153): a disk of radius 0.2 at
                                   x = -0.25 flying at +1e3 into a wall [0, ny/2 dx) x [-1, 1]
                                   whose top and bottom rows are fixed, dx = 0.4/(ny-1); locked
                                   grid (-0.5,-1.2)-(1,1.2) (benchmarks.cpp:136) unless
                                   grid_adaptive; no tool */

typedef struct mfc_config {
  /* workpiece lattice */
  int32_t lattice;        /* MFC_LATTICE_* */
  int32_t nx, ny;         /* particle rows/columns (nx ignored for MFC_LATTICE_REF) */
  double dx;              /* spacing (MFC_LATTICE_EXPLICIT) */
  double lo_x, hi_x, hi_y;/* workpiece extent; hi_x only used by MFC_LATTICE_REF */
  double hdx;             /* h = hdx dx */
  double jitter;          /* initial positions += U(-jitter, jitter) dx per coordinate */
  uint64_t seed;          /* SplitMix64 seed of the jitter */
  int32_t fixed_bottom_rows; /* rows fixed at the bottom (reference: 1, benchmarks.cpp:374) */
  int32_t fix_left;       /* fix left edge below half height (benchmarks.cpp:376) */
  int32_t fix_right;      /* fix right edge (benchmarks.cpp:378) */

  /* material (phys_constants, constants_structs.h:27-35) */
  double E, nu, rho0;
  /* Johnson-Cook (joco_constants, constants_structs.h:61-75) */
  double jc_A, jc_B, jc_C, jc_n, jc_m, jc_Tmelt, jc_Tref, jc_eps_dot_ref;
  /* workpiece thermal constants (trml_constants, constants_structs.h:39-46) */
  double cp, tq, k_wp, eta, T_init;
  /* artificial viscosity / XSPH / artificial stress (corr_constants, constants_structs.h:50-57) */
  double av_alpha, av_beta, av_eta, xsph_eps, stress_eps;

  /* tool (tool.cpp:342-375, benchmarks.cpp:389-421) */
  double tool_tl_x, tool_tl_y, tool_length, tool_height;
  double rake_deg, clear_deg, edge_radius;  /* edge_radius == 0: sharp tool */
  double feed;            /* depth of cut */
  double speed;           /* cm/us */
  double mu;              /* Coulomb friction coefficient */
  double contact_alpha;   /* penalty scaling (benchmarks.cpp:419) */
  int32_t sample_tool;    /* tool particles for thermal contact (benchmarks.cpp:263,424-459) */
  double rho_tool, cp_tool, k_tool;

  /* solver */
  int32_t thermal_solver; /* MFC_THERMAL_* */
  int32_t precision;      /* MFC_PREC_* */
  double cfl;             /* dt = cfl h / (c0 + sqrt(speed)) (benchmarks.cpp:530 uses 0.5) */
  int32_t pse_flag_quirk; /* 1: PSE tool flag reads 0 for index >= N/2 (interactions_gpu.cu:405) */
  int32_t promote_fp64;   /* 1: evaluate the reference's single-precision truncations of its FP64
                             build (sqrtf, fmaxf, atan2f/cosf/sinf) in FP64; 0: as the reference */
  double secant_tol;      /* absolute tolerance of the radial-return secant; 0 selects the
                             reference's 1e-6 (actions_gpu.cu:261) */
  int32_t jc_model;       /* MFC_JC_*: the flow stress the secant solves with. As in the
                             reference, the yield test and the plastic heat use plain JC */
  int32_t elastic;        /* 1: no Johnson-Cook return (the reference without
                             actions_setup_johnson_cook_constants, setup_rings) */
  double jc_tanh_a, jc_tanh_b, jc_tanh_c, jc_tanh_d;  /* joco_constants tanh_a..d
                                                         (constants_structs.h:71-74) */
  int32_t grid_adaptive;  /* 1: the cell grid follows the particles every step: bounding box of
                             all positions +- 1e-6 and cell size max h, recomputed before the
                             binning (grid_base::update_geometry, grid.cu:97-136; setup_impact's
                             active grid, benchmarks.cpp:135); 0: the scenario's locked grid */
  int32_t max_cell_factor;/* adaptive grid: a cell table of max_cell_factor * N cells (0: 10, the
                             reference's grid_gpu_green(10*n, n)); a step whose grid would need
                             as many cells keeps the previous geometry and is counted in
                             mfc_info.grid_overflow (the reference asserts) */
  int32_t grid_kind;      /* MFC_GRID_*: the reference's two cell grids */
  int32_t reserved[1];
} mfc_config;

typedef struct mfc_info {
  int32_t n_particles;    /* workpiece + tool particles */
  int32_t n_tool;         /* sampled tool particles (stored after the workpiece initially) */
  int32_t grid_nx, grid_ny, num_cell;
  int32_t precision;
  double dt, t_final;     /* global_dt, global_t_final (benchmarks.cpp:530-531) */
  double mass;            /* particle mass dx^2 rho0 */
  double bbmin_x, bbmin_y, cell_size;
  int64_t step;           /* number of completed steps (global_step) */
  int64_t out_of_grid;    /* particles binned so far whose position lay outside the grid on
                             an axis (counted once per binning). Their cell id is the
                             reference's ix * ny + iy, clamped into the cell table; such a
                             particle loses neighbours, so a valid run keeps this at 0 */
  int64_t grid_overflow;  /* adaptive grid: steps whose geometry did not fit the cell table
                             (0 in a valid run) */
  int32_t grid_adaptive;  /* the configuration's grid_adaptive; grid_nx, grid_ny, num_cell,
                             bbmin_* then describe the geometry of the last step */
  int32_t cell_table;     /* cells of the binning table (num_cell for a locked grid) */
} mfc_info;

/* particle fields, returned in the handle's current (cell-sorted) order */
typedef enum mfc_field {
  MFC_FIELD_POS = 0,      /* 2 doubles */
  MFC_FIELD_VEL = 1,      /* 2 */
  MFC_FIELD_S = 2,        /* 4: Sxx, Sxy, Syy, Szz */
  MFC_FIELD_RHO = 3,      /* 1 */
  MFC_FIELD_P = 4,        /* 1 */
  MFC_FIELD_T = 5,        /* 1 */
  MFC_FIELD_EPS_PL = 6,   /* 1 */
  MFC_FIELD_EPS_PL_DOT = 7, /* 1 */
  MFC_FIELD_FC = 8,       /* 2 contact force */
  MFC_FIELD_FT = 9,       /* 2 friction force */
  MFC_FIELD_N = 10,       /* 2 contact normal (positional, as particle_gpu::n) */
  MFC_FIELD_H = 11,       /* 1 */
  MFC_FIELD_FLAGS = 12,   /* 1 int32: bit0 fixed, bit1 blanked, bit2 tool particle */
  MFC_FIELD_ID = 13,      /* 1 int32: particle index at creation */
  MFC_FIELD_HASH = 14,    /* 1 int32: cell id the current order was sorted by */
  MFC_FIELD_POS_T = 15,   /* 2 */
  MFC_FIELD_VEL_T = 16,   /* 2 */
  MFC_FIELD_RHO_T = 17,   /* 1 */
  MFC_FIELD_S_T = 18,     /* 4 */
  MFC_FIELD_T_T = 19,     /* 1 */
  MFC_FIELD_R = 20,       /* 3 artificial stress Rxx, Rxy, Ryy */
  MFC_FIELD_COUNT = 21
} mfc_field;

#define MFC_FLAG_FIXED   1
#define MFC_FLAG_BLANKED 2
#define MFC_FLAG_TOOL    4

/* tool state: up to 5 segments + 1 fillet (tool_gpu.cuh:32-52) */
typedef struct mfc_tool_state {
  double left[5][2], right[5][2], n[5][2];
  double line_a[5], line_b[5];
  int32_t vertical[5];
  int32_t num_seg, has_fillet;
  double fillet_p[2], fillet_r, fillet_t1, fillet_t2;
  double vel[2];
} mfc_tool_state;

typedef struct mfc_sim mfc_sim;

/* number of components of a field (1..4), or 0 for an unknown field */
int mfc_field_components(int field);

/* the defaults of setup_ref_cut(ny, grid, rake, clear, chamfer, speed, feed)
 * (benchmarks.cpp:255-537, benchmarks.h:48): Ti6Al4V after Johnson 1985, thermal
 * constants, tool sampling, Brookshaw conduction, FP64, CFL 0.5. speed == 0 selects
 * the reference's 500 m/min (benchmarks.cpp:384). */
int mfc_config_ref_cut(mfc_config *cfg, int ny, double rake, double clear, double chamfer,
                       double speed, double feed);

/* the constants of setup_rings(nbox, grid) (benchmarks.cpp:155-253, benchmarks.h:42): rubber-like
 * rings (E 1e7, nu 0.4, rho0 1), hdx 1.7, no plasticity, CFL 0.5 with the rings' 170 as the
 * velocity of the time-step rule, FP64. The reference's main() names nbox = 80. */
int mfc_config_rings(mfc_config *cfg, int nbox);

/* the constants of setup_impact(nbox, grid) (benchmarks.cpp:23-153, benchmarks.h:45): steel 4340
 * in SI units with its Johnson-Cook constants, no heat, hdx 1.7, CFL 0.3 with the flyer's 1e3,
 * FP64, and the adaptive grid of its active line (grid_gpu_green(10*n, n), benchmarks.cpp:135).
 * The reference's main() names nbox = 20. */
int mfc_config_impact(mfc_config *cfg, int nbox);

/* setup_ref_cut + particle_gpu + grid_gpu_green + leap_frog construction */
int mfc_create(const mfc_config *cfg, mfc_sim **out);
void mfc_destroy(mfc_sim *sim);

/* leap_frog::step, nsteps times (hipGraph replay when nsteps > 1) */
int mfc_step(mfc_sim *sim, int nsteps);
/* wait for queued steps to finish */
int mfc_sync(mfc_sim *sim);

/* run nsteps steps without graphs, timing each phase with HIP events on the handle's
 * stream; ms[phase] receives the mean milliseconds per step of: 0 binning (cell offsets and
 * counting sort), 1 gather/predict/EOS/artificial stress, 2 the fused interaction kernel
 * (neighbour sums + per-particle update + contact + the next step's cell ids and counts), or
 * the neighbour-sum kernel when split_interact is set, 3 the update kernel (split_interact
 * only, else 0; it then computes the next step's cell ids), 4 tool-force reduction. Phases 2
 * and 3 are the kernels' own durations (events recorded by their dispatch); the others are
 * measured between events on the stream and include the gaps between launches. */
#define MFC_PROFILE_PHASES 5
int mfc_profile(mfc_sim *sim, int nsteps, double *ms);

/* report_force: F = sum_i fc_i + sum_i ft_i after the last step, summed in a fixed
 * order: groups of 64 particles in storage order, each reduced by halving strides; lane l
 * of 256 adds the group results l, l + 256, ... in ascending order (4 groups per 256
 * particles started); the 256 lane sums reduced by halving strides; sum of fc and sum of ft
 * added last. The order does not depend on the work-group size of the step kernels. */
int mfc_tool_force(mfc_sim *sim, double F[2]);
/* forces of every step since creation or the last reset, as F[step][2];
 * *n receives the number of steps copied (at most cap); with F == NULL, *n receives the
 * number of recorded steps (at most 2^20 are kept) */
int mfc_force_history(mfc_sim *sim, double *F, int64_t cap, int64_t *n);
int mfc_force_history_reset(mfc_sim *sim);

/* tool_wear::eval_usui on the device, over the current state: the Usui wear rate
 * K exp(alpha / T) p |v_r| of every particle in contact (|fc.x|, |fc.y| or |ft.x| non-zero: the
 * reference tests fc.y twice and never ft.y), with v_r = v - v*n (v the particle velocity minus
 * tool_vel, the product component-wise as in the reference's glm code) and p = |fc| / (2 pi h).
 * out[0] = min, out[1] = max, out[2] = mean over the particles in contact (all 0 when none);
 * the handle then accumulates dt * mean (get_accum_wear), returned in out[3]. The reference's
 * setup passes K = 0.078, alpha = 2500 and tool_vel = (0, tool speed) (benchmarks.cpp:508-512).
 * The mean's sum is a fixed-order tree over tiles of 256 particles, not the reference's
 * sequential host loop. */
int mfc_tool_wear(mfc_sim *sim, double K, double alpha, const double tool_vel[2], double dt, double out[4]);
/* mfc_tool_wear for callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid allocation); no counterpart in the reference. The same
 * kernels and results as mfc_tool_wear, with the per-tile partial sums in a fixed device array of
 * the library (one per GPU, held until the process ends) instead of memory allocated on the first
 * call: it calls no allocating routine. Handles of up to MFC_NOALLOC_WEAR_MAX_TILES tiles of 256
 * particles (2,097,152 particles); larger ones are refused with MFC_E_RANGE. All calls on a GPU
 * share the array: calls that overlap on one GPU race. */
#define MFC_NOALLOC_WEAR_MAX_TILES 8192
int mfc_tool_wear_noalloc(mfc_sim *sim, double K, double alpha, const double tool_vel[2], double dt, double out[4]);

/* vtk_writer_write(particles, step): the current state's particles that are not blanked, as
 * the reference's legacy ASCII VTK file at path (vertex cells; density, Temperature, Fixed,
 * Tool, EquivAccumPlasticStrain, Velocity, SvM). Unlike the reference, the blanked test reads
 * the state's flags (the reference never copies them) and cell k is written point k (the
 * reference writes the particle's slot, which points past the points once any are skipped). */
int mfc_write_vtk(mfc_sim *sim, const char *path);
/* vtk_writer_write(tool, step): the tool's body and fillet as triangles (legacy ASCII VTK) */
int mfc_write_tool_vtk(mfc_sim *sim, const char *path);

/* ---- reverse mode ----
 * Gradient of the mean tool force over a window of steps with respect to
 * theta = (JC A, JC B, JC n, JC C, JC m, friction mu), in the solver's units. The reference
 * has no derivatives; this is the calibration interface of the solver. The handle must use
 * precision MFC_PREC_FP64 with promote_fp64 = 1 (the smooth form of the step) and should use
 * a tight secant_tol (1e-14): the radial return is differentiated at its converged root.
 * Branches (elastic/plastic, contact, stick/slip, kernel support) are held at the ones the
 * forward run took.
 *
 * mfc_vjp_forward: nsteps steps from the current state (as mfc_step), keeping the state at
 * the start of every checkpoint_every-th step on the device; Fmean receives the mean of F over
 * the steps [w0, w1) of the run (0 <= w0 < w1 <= nsteps).
 * mfc_vjp_backward: grad[j] = cot . dFmean/dtheta_j for the last mfc_vjp_forward; may be
 * called again with other cotangents; leaves the handle's state as mfc_vjp_forward left it.
 * Memory: about 290 B per particle per checkpoint plus checkpoint_every snapshots more.
 * mfc_vjp: both calls for a cotangent known in advance. */
typedef struct mfc_vjp_req {
  int nsteps;
  int checkpoint_every;
  int w0, w1;
  double cot[2];
  double Fmean[2]; /* out */
} mfc_vjp_req;
int mfc_vjp_forward(mfc_sim *sim, int nsteps, int checkpoint_every, int w0, int w1, double Fmean[2]);
int mfc_vjp_backward(mfc_sim *sim, const double cot[2], double grad[6]);
int mfc_vjp(mfc_sim *sim, mfc_vjp_req *req, double grad[6]);

/* launch configuration of the step kernels (results do not depend on it: every variant
 * evaluates the same expressions in the same order). split_interact = 0 (default): one fused
 * interaction kernel; 1: a neighbour-sum kernel and a per-particle update kernel with the sums
 * stored in between (fewer registers per kernel, 224 B more traffic per particle).
 * compact_pairs = 1: each lane first marks its in-support neighbours in a bit mask, then runs
 * the pair terms of the marked ones in order; 0 (default): the pair terms run inside the window
 * walk. stage_limit: the most window candidates a tile of 256 particles copies into LDS
 * (0 = MFC_STAGE_CAP); a tile whose candidates span more sorted indices reads them from the
 * records in memory. xcd_chunk: the interaction kernels deal their tiles to the 8 XCDs in chunks
 * of this many consecutive tiles (1..MFC_XCD_CHUNK_MAX; 0 = 8); longer chunks fetch fewer
 * neighbour columns into two L2s, shorter ones balance the XCDs better. The Python package's
 * autotuner (pymfc.autotune) times the combinations on a live handle and keeps the fastest.
 * sort_fast and sort_list bound the counting sort's paths (0 = the defaults; no autotuned
 * setting: they exist so that every path of the sort can be run on any cut): a chunk of cells
 * with at most sort_fast particles (1..MFC_SORT_FAST_MAX) sorts them in registers and LDS, a
 * denser one walks its particle list with a bucket in memory, and a chunk with more than
 * sort_list particles (>= 1) passes over every particle's cell id instead of its list. */
#define MFC_STAGE_CAP 2048
#define MFC_XCD_CHUNK_MAX 64
#define MFC_SORT_FAST_MAX 1536
typedef struct mfc_tuning {
  int split_interact;
  int compact_pairs;
  int stage_limit;
  int xcd_chunk;
  int sort_fast;
  int sort_list;
  int reserved[2];
} mfc_tuning;
int mfc_set_tuning(mfc_sim *sim, const mfc_tuning *tuning);
int mfc_get_tuning(const mfc_sim *sim, mfc_tuning *tuning);

/* ---- batches of independent instances ----
 * n cutting instances (any sizes, one precision) stepped together: every phase of a step is
 * one launch over all instances' tiles, and two steps are one hipGraph launch. Each instance
 * gives exactly the results of its own single run (same kernels, same order). The reference
 * runs one simulation per process (module-level __constant__ state); this is the solver's
 * interface for parameter sweeps and calibration ensembles.
 * mfc_batch_instance returns instance i's handle for the read-out calls (mfc_get_field,
 * mfc_tool_force, mfc_force_history, mfc_get_info, ...); it is owned by the batch, must not be
 * destroyed or stepped on its own, and shares the batch's stream.
 * mfc_batch_vjp_forward/backward: the reverse mode of every instance (Fmean: 2 n doubles,
 * cot: 2 n, grad: 6 n); the taped forward runs batched, the backward instance by instance. */
typedef struct mfc_batch mfc_batch;
int mfc_batch_create(const mfc_config *cfgs, int n, mfc_batch **batch);
void mfc_batch_destroy(mfc_batch *batch);
int mfc_batch_size(const mfc_batch *batch);
int mfc_batch_instance(mfc_batch *batch, int i, mfc_sim **sim);
int mfc_batch_set_tuning(mfc_batch *batch, const mfc_tuning *tuning);
int mfc_batch_step(mfc_batch *batch, int nsteps);
int mfc_batch_sync(mfc_batch *batch);
int mfc_batch_vjp_forward(mfc_batch *batch, int nsteps, int checkpoint_every, int w0, int w1, double *Fmean);
int mfc_batch_vjp_backward(mfc_batch *batch, const double *cot, double *grad);

/* ---- checkpoints ----
 * The reference has no restart format (its only dump is the text dump_state, debug.cpp:20-64).
 * mfc_save writes the handle's state after its queued steps to a binary file: the creating
 * configuration, launch tuning, completed steps, the current particle fields in storage order,
 * the positional arrays (contact normal, pressure, artificial stress), the cell ids of the last
 * sort, the tool state, the last tool force and the force history, with an FNV-1a checksum of
 * the payload. mfc_load creates a new handle from such a file; it continues bit for bit as the
 * saved run would have (the next step recomputes the cell tables and neighbour records).
 * A file of another format version, a truncated file or a checksum mismatch is refused with
 * MFC_E_INVALID. */
int mfc_save(mfc_sim *sim, const char *path);
int mfc_load(const char *path, mfc_sim **sim);
/* the configuration a handle was created from (for a loaded handle: the saved run's) */
int mfc_get_config(const mfc_sim *sim, mfc_config *cfg);

int mfc_get_info(const mfc_sim *sim, mfc_info *info);
/* copy a field to host memory: bytes must be N * components * sizeof(double)
 * (sizeof(int32_t) for FLAGS, ID, HASH) */
int mfc_get_field(mfc_sim *sim, int field, void *host, size_t bytes);
/* overwrite a field from host memory (same layout); used to inject states in tests */
int mfc_set_field(mfc_sim *sim, int field, const void *host, size_t bytes);
/* cell ranges of the last binning: offsets[c]..offsets[c+1] is cell c in storage order,
 * an array of num_cell + 1 ints (num_cell of mfc_get_info: for an adaptive grid the last
 * step's geometry) */
int mfc_get_cells(mfc_sim *sim, int32_t *offsets, size_t count);
int mfc_get_tool(mfc_sim *sim, mfc_tool_state *tool);

/* host-side setup only (no device work): the initial particles and tool of a
 * configuration, in creation order. Arrays may be NULL; non-NULL arrays hold
 * info->n_particles entries (pos: 2 per particle). */
int mfc_scenario(const mfc_config *cfg, mfc_info *info, double *pos, double *rho, double *h, double *T,
                 int32_t *flags, mfc_tool_state *tool);

/* run the shared transcendental library on the device (diagnostics):
 * fn 0 exp, 1 log, 2 pow(x,y), 3 atan2(x,y), 4 sin, 5 cos, 6 atan2f, 7 cosf, 8 sinf,
 * 9 sqrt, 10 sqrtf, 11 x/y, 12 tanh */
int mfc_device_math(int fn, const double *x, const double *y, double *out, int64_t n);

const char *mfc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* MFC_H_ */
