/*
 * psst_gpu.h -- C-ABI of the MI355X (gfx950) explicit finite-volume Euler
 * solver.  This is the drop-in boundary under the reference's host C++
 * (PSST, nkmashaev/SupersonicFlowSolver): the host keeps the Fluent reader
 * and the settings files, flattens the grid once, and hands the whole
 * iteration loop of psstGridFV::explicit_solve() to the device.
 *
 * Plain C types only; no HIP or torch types cross this boundary.  All
 * pointers passed in are read during the call and never retained.  Arrays
 * are structure-of-arrays: `[k][n]` means k rows of n contiguous values.
 *
 * Cell numbering ("canonical"): interior cells 0..N-1 in the order of the
 * reference's update loop (cell zones of type 1 in ascending zone id, then
 * zone order; psstFiniteVolumeGrid.cpp:312-349), followed by ghost cells
 * N..N+G-1 (boundary zones in ascending zone id, then zone order).
 * Face numbering: the reference's face-loop order (face zones in ascending
 * zone id, zone type 31 skipped, then zone order; e.g.
 * psstConvectiveFlow.cpp:118-123).
 *
 * Error convention: every entry point returns a psst_status; no C++
 * exception crosses the ABI.  The reference throws C++ exceptions instead
 * (e.g. psstConvectiveFlow.cpp:285, psstFiniteVolumeGrid.cpp:42); the host
 * adapter (psstGridFV in the package) rethrows std::runtime_error with
 * psst_last_error().
 */
#ifndef PSST_GPU_H
#define PSST_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum psst_status {
    PSST_OK = 0,
    PSST_BAD_ARG = 1,         /* invalid argument or inconsistent mesh/settings */
    PSST_HIP_ERROR = 2,       /* a HIP runtime call failed */
    PSST_NAN_DETECTED = 3,    /* a non-finite or non-physical state was produced */
    PSST_NO_CONVERGENCE = 4,  /* tolerance not reached within the iteration budget */
    PSST_UNSUPPORTED = 5,     /* a setting this build does not implement */
    PSST_IO_ERROR = 6         /* host file could not be read or written */
} psst_status;

/* Strategy enums; the values are the reference's `.inp` keys
 * (psstFiniteVolumeGrid.cpp:50-192, Description.md:177-195). */
enum { PSST_FLUX_AUSM = 1, PSST_FLUX_HLL = 2, PSST_FLUX_HLLC = 3 };
enum { PSST_GRAD_GREEN_GAUSS = 1, PSST_GRAD_LEAST_SQUARES = 2 };
enum { PSST_LIMITER_VAN_ALBADA = 1, PSST_LIMITER_MINMOD = 2, PSST_LIMITER_VAN_LEER = 3,
       PSST_LIMITER_VENKATAKRISHNAN = 4 };
enum { PSST_DT_UNIFORM = 0, PSST_DT_LOCAL1 = 1, PSST_DT_LOCAL2 = 2 };
/* Boundary kinds; the values are the reference's .bvp/.btq ids
 * (psstFiniteVolumeGrid.cpp:432-439, :509-515). */
enum { PSST_BC_SS_INLET = 23, PSST_BC_SS_OUTLET = 26, PSST_BC_SLIP_WALL = 15,
       PSST_BC_INLET = 27, PSST_BC_OUTLET = 25 };
enum { PSST_BCT_WALL_T = 11, PSST_BCT_WALL_Q = 18, PSST_BCT_INLET = 27, PSST_BCT_OUTLET = 25 };

/* Flattened grid.  Replaces the shared_ptr object graph that the strategy
 * objects walk through psstGridFV's accessors (psstFiniteVolumeGrid.h:54-86;
 * psstFace/psstVolume, psstFiniteVolumeGridTools.h:148-360).  Geometry is
 * computed on the host in FP64 exactly as psstFace::initialize and
 * psstVolume::initialize do (psstFiniteVolumeGridTools.cpp:397-505, :709-827). */
typedef struct psst_mesh_desc {
    int32_t dim;                  /* 2 or 3 */
    int32_t n_cells;              /* N interior cells */
    int32_t n_ghosts;             /* G ghost cells (one per boundary face) */
    int32_t n_faces;              /* F faces */
    const double *cell_volume;    /* [N+G]; ghosts 0 */
    const double *cell_centre;    /* [dim][N+G]; a ghost's centre is its face's centre */
    const double *face_normal;    /* [dim][F] unit normal, pointing toward neighbour 0 */
    const double *face_area;      /* [F] */
    const int32_t *face_cells;    /* [2][F] neighbour 0 (Fluent c0) and neighbour 1 (c1) */
    const double *face_beta;      /* [F] Green-Gauss weight of neighbour 0, as the reference
                                     indexes it (psstGradCalculation.cpp:30-70) */
    const int32_t *visit_offsets; /* [N+1] CSR: faces of each interior cell ... */
    const int32_t *visit_faces;   /* ... in face-loop order (sum order of fluxes and GG) */
    const int32_t *list_offsets;  /* [N+1] CSR: faces of each interior cell ... */
    const int32_t *list_faces;    /* ... in the cell's own face-list order (psstVolume::add_face,
                                     psstAnsysGridReader.cpp:825-883): order of the time step
                                     and least-squares sums */
    const int32_t *ghost_face;    /* [G] the boundary face of each ghost */
    const int32_t *ghost_bc;      /* [G] index into psst_settings.bc */
    const double *lsq_r;          /* [6][N] least-squares factors a,b,d,c,e,g per cell
                                     (psstGradCalculation.cpp:191-260); NULL unless grad == 2 */
    const double *lsq_beta;       /* [N] 3D least-squares beta; NULL unless grad == 2 */
    double min_volume;            /* smallest interior cell volume */
    const double *face_shift;     /* [3][F] translation that carries neighbour 1's centre into
                                     neighbour 0's frame across a periodic face (zero elsewhere);
                                     NULL: no periodic faces.  No reference counterpart (the
                                     reader knows Fluent's periodic zone type 12 by name only,
                                     psstAnsysGridReader.h:140-143). */
} psst_mesh_desc;

/* One boundary zone: dynamic + thermal condition (psstBoundary, psstBoundaryCondition.h). */
typedef struct psst_bc {
    int32_t dynamic_kind;         /* PSST_BC_* */
    int32_t thermal_kind;         /* PSST_BCT_* */
    double velocity[3];           /* fixed velocity (supersonic inlet) */
    double pressure;              /* fixed pressure (supersonic inlet, subsonic outlet) */
    double temperature;           /* fixed temperature (thermal inlet, wall_T) */
} psst_bc;

/* Solver settings: the `.inp`/`.prop` content that psstGridFV::solver_parameters
 * turns into strategy objects (psstFiniteVolumeGrid.cpp:36-194), plus the
 * device choices of this build. */
typedef struct psst_settings {
    int32_t flux;                 /* PSST_FLUX_* (.inp line 9) */
    int32_t order;                /* 1 or 2 (.inp line 10) */
    int32_t grad;                 /* PSST_GRAD_* (.inp line 11) */
    int32_t limiter;              /* PSST_LIMITER_* (.inp line 12) */
    int32_t time_mode;            /* PSST_DT_* (.inp line 3) */
    int32_t rk_levels;            /* 1..8 */
    double rk_coeff[8];           /* alpha_k (psstRungeKutta.cpp:10-57) */
    double cfl;
    double gamma, rm, cp, cv;     /* psstFiniteVolumeGrid.cpp:649-654 */
    double uniform_dt;            /* time step of PSST_DT_UNIFORM (psstLocalTimeCalculation.cpp:26-30) */
    int32_t precision;            /* 64: FP64 state (parity); 32: FP32 state/flux, FP64 geometry and norms */
    int32_t dt_ordering;          /* 0: reference order (PSST_DT_LOCAL1 sees neighbours already updated
                                     in the same stage, psstFiniteVolumeGrid.cpp:317-347); 1: Jacobi */
    int32_t fused;                /* 0: face-flux + cell-gather kernels; 1: tiled kernels (the fused stage
                                     kernel; with time_mode 1 and dt_ordering 0 the tiled gradient and
                                     residual passes before the ordered sweep) */
    int32_t n_bc;                 /* entries of bc */
    const psst_bc *bc;
    double venkat_k;              /* Venkatakrishnan constant K (limiter 4 only) */
    int32_t tile_cells;           /* fused kernels: cells per tile (0 = the first of 224, 192, 160, 128, 96
                                     cells in FP64, 768 ... 128 in FP32, whose stage kernel LDS lets a CU
                                     hold 16 waves) */
    int32_t tile_threads;         /* fused kernels: threads per tile workgroup, 256 or 512 (0 = default) */
} psst_settings;

typedef struct psst_ctx psst_ctx;

/* Creates a device context: uploads geometry, orders and settings, builds
 * the per-stage schedule.  Replaces the strategy construction and their
 * initialize() calls in psstGridFV::solver_parameters
 * (psstFiniteVolumeGrid.cpp:36-194). */
psst_status psst_ctx_create(const psst_mesh_desc *mesh, const psst_settings *settings, psst_ctx **out);

/* Loads a primitive state [(dim+3)][N+G] (u_0..u_{dim-1}, p, T, rho) and
 * forms the conservative variables from it, as explicit_solve does before
 * its loop (psstFiniteVolumeGrid.cpp:279-300). */
psst_status psst_set_state(psst_ctx *ctx, const double *prim);

/* Runs n_iter iterations of psstGridFV::explicit_solve's loop
 * (psstFiniteVolumeGrid.cpp:305-397): gradients, then per RK stage the
 * convective flux, the time step and update, the ghost boundary states.
 * res_hist (may be NULL) receives [n_iter][dim+2] residual max-norms in the
 * reference's column order (momentum..., mass, energy; :331-332, :377-382).
 * If tol > 0, stops after the first iteration whose max_j res_j/res0_j <= tol,
 * where res0 is the row set by psst_set_residual_reference (or the first row
 * this context produced) and columns with res0_j <= 1e-12 max_k res0_k (a
 * component zero up to rounding, such as the z momentum of a z-symmetric
 * flow) are left out; *iters_done reports how many ran.  A non-finite
 * residual row ends the run at that iteration (PSST_NAN_DETECTED).
 * Synchronous. */
psst_status psst_run(psst_ctx *ctx, int32_t n_iter, double tol, double *res_hist, int32_t *iters_done);

/* Same as psst_run, allocating nothing: no hipMalloc, hipFree or other
 * allocating call.  No counterpart in the reference; it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid it).  The context must already exist
 * (psst_ctx_create allocates); the residual slots, control words and
 * residual rows of the call live in a fixed workspace of about 44 KB per
 * GPU (one static device array, held until the process ends), whose address
 * psst_ctx_create takes once per device.  Limits: n_iter <= 1024 per call
 * (PSST_BAD_ARG beyond), and kernel timing (psst_kernel_timing) must be off.
 * Every call on one GPU shares that GPU's workspace: calls that overlap on
 * one GPU (several host threads, or contexts on several streams) race.
 * Results are bitwise those of psst_run with the same arguments. */
psst_status psst_run_noalloc(psst_ctx *ctx, int32_t n_iter, double tol, double *res_hist, int32_t *iters_done);

/* Copies the primitive state [(dim+3)][N+G] to the host. */
psst_status psst_get_state(psst_ctx *ctx, double *prim);

/* Copies the conservative state [(dim+2)][N] of the interior cells. */
psst_status psst_get_conservative(psst_ctx *ctx, double *w);

/* Sets the normalising residual row used by psst_run's tolerance test. */
psst_status psst_set_residual_reference(psst_ctx *ctx, const double *res0);

/* Diagnostics for per-kernel parity tests.  psst_debug_gradients computes
 * the gradients of the current state ([(dim+2)][dim][N]: u_0..u_{dim-1}, p, T)
 * as psstGreenGauss/psstLeastSquare::calculate do (psstGradCalculation.cpp:80-177,
 * :268-368); psst_debug_residual computes one convective pass
 * ([(dim+2)][N]) with the gradients last computed
 * (psstConvective*::calculate, psstConvectiveFlow.cpp:106-871). */
psst_status psst_debug_gradients(psst_ctx *ctx, double *grads);
psst_status psst_debug_residual(psst_ctx *ctx, double *residual);

/* Diagnostic, host only (no device needed): builds the schedule of the
 * reference-ordered time step (time step mode 1, dt_ordering 0) that
 * psst_ctx_create builds for a large mesh -- dependency levels of the
 * in-place update loop (psstFiniteVolumeGrid.cpp:317-347 with
 * psstLocalTimeCalculation.cpp:90-131), cells in chunks of about
 * `chunk_cells` (<= 0: the default), the waits between chunks and the
 * levels each chunk publishes -- and checks it by simulation: every earlier
 * neighbour of a cell is finished before the cell, and every chunk
 * finishes.  partition: 1 = ranges of the device (Morton) numbering, 2 =
 * boxes of per-axis dependency potentials, 0 = the one with the shorter
 * modelled critical path (as psst_ctx_create picks).  stats[10] receives:
 * levels, chunks, entries (levels of a chunk), waits, publishes, widest
 * entry (cells), makespan in workgroup passes with free waits and with
 * waits costing four passes, chunk cells, partition used.  PSST_BAD_ARG
 * with psst_last_error(NULL) on a violation. */
psst_status psst_debug_sweep_schedule(const psst_mesh_desc *mesh, int32_t chunk_cells, int32_t partition,
                                      int64_t *stats);

/* Kernel timing with HIP events on the context's stream.  While enabled,
 * every launch is bracketed by events; psst_kernel_stats returns, for the
 * kernel named `name`, the launch count and the mean duration in ms. */
psst_status psst_kernel_timing(psst_ctx *ctx, int32_t enable);
psst_status psst_kernel_stats(psst_ctx *ctx, const char *name, int64_t *count, double *mean_ms);
/* Names the kernels a psst_run issues (";"-separated) into buf. */
psst_status psst_kernel_names(psst_ctx *ctx, char *buf, int32_t buf_len);

/* Diagnostic: times a plain device-to-device copy kernel ("k_copy") moving
 * `bytes` (read) + `bytes` (write), reps times; used to calibrate hardware
 * traffic counters against a kernel of known traffic.  No reference
 * counterpart. */
psst_status psst_copy_calibration(int64_t bytes, int32_t reps, double *mean_ms);

void psst_ctx_destroy(psst_ctx *ctx);

/* Last error of a context (or of the last failed create when ctx is NULL). */
const char *psst_last_error(const psst_ctx *ctx);

/* Switches the state between FP64 and FP32 in place (geometry and residual
 * norms stay FP64).  An FP32 state cannot drive the residual much below
 * about 1e-4 of its first value (the rounding of the face-flux sums), so a
 * solve to a tight tolerance can run FP32 first and finish in FP64.  No
 * reference counterpart (the reference is FP64 only). */
psst_status psst_set_precision(psst_ctx *ctx, int32_t precision);

/* Autotuner of the fused kernels' tile shape: runs `trial_iters` iterations
 * of the given mesh, settings and state for each candidate (cells per tile x
 * threads per workgroup) and returns the fastest in best[0..1].  cand
 * ([max_cand][2]) and ms_per_iter ([max_cand]) receive every candidate and
 * its time (-1: did not run).  No reference counterpart. */
psst_status psst_autotune(const psst_mesh_desc *mesh, const psst_settings *settings, const double *prim,
                          int32_t trial_iters, int32_t max_cand, int32_t *cand, double *ms_per_iter,
                          int32_t *n_cand, int32_t *best);

/* ---------------------------------------------------------------------------
 * Batches: many independent instances (meshes, boundary values) advanced
 * together, one launch per kernel for all of them.  The reference solves one
 * case per process (main.cpp:9-63); a batch is the same explicit_solve loop
 * (psstFiniteVolumeGrid.cpp:305-397) over every instance, each instance
 * bitwise identical to its own psst_ctx run in FP64.  The instances share
 * the scheme (every psst_settings field but the boundary table); they are
 * stored as one union mesh whose tiles never straddle two instances.  After
 * every iteration the device computes each instance's residual norms and
 * freezes the instances that reached the tolerance; one workgroup runs each
 * remaining tile, an eighth of them per XCD, and each iteration is replayed
 * from a captured hipGraph.
 * ------------------------------------------------------------------------- */
typedef struct psst_batch psst_batch;

psst_status psst_batch_create(int32_t n, const psst_mesh_desc *meshes, const psst_settings *settings,
                              psst_batch **out);
/* Per-instance state, in the instance's own canonical numbering
 * ([(dim+3)][N_i+G_i], as psst_set_state / psst_get_state). */
psst_status psst_batch_set_state(psst_batch *b, int32_t inst, const double *prim);
psst_status psst_batch_get_state(psst_batch *b, int32_t inst, double *prim);
psst_status psst_batch_set_residual_reference(psst_batch *b, int32_t inst, const double *res0);
/* Runs every instance (those with a finite state) for up to max_iter
 * iterations; with tol > 0 each stops after its first iteration whose
 * max_j res_j / res0_j <= tol.  Per instance (each may be NULL): iters =
 * iterations run, status = PSST_OK / PSST_NO_CONVERGENCE / PSST_NAN_DETECTED,
 * drop = last max_j res_j / res0_j; res_hist = [max_iter][n][dim+2] residual
 * rows (zero once an instance stopped).  Returns the worst status. */
psst_status psst_batch_run(psst_batch *b, int32_t max_iter, double tol, int32_t *iters, int32_t *status,
                           double *drop, double *res_hist);
psst_status psst_batch_set_precision(psst_batch *b, int32_t precision);
/* Excludes (0) or includes (1, default) an instance in later psst_batch_run
 * calls; an excluded instance keeps its state and reports 0 iterations. */
psst_status psst_batch_set_enabled(psst_batch *b, int32_t inst, int32_t enabled);
/* 1: replay iterations from a captured hipGraph (default); 0: launch each kernel. */
psst_status psst_batch_set_graph(psst_batch *b, int32_t enable);
/* info[0..7] = instances, union N, G, F, tiles, cells per tile, threads per tile, iterations run */
psst_status psst_batch_info(const psst_batch *b, int64_t *info);
/* The batch's device context (for psst_kernel_timing / psst_kernel_stats). */
psst_ctx *psst_batch_context(psst_batch *b);
void psst_batch_destroy(psst_batch *b);
const char *psst_batch_last_error(const psst_batch *b);

/* ---------------------------------------------------------------------------
 * Host-side helpers of the same library (the build's own host C++): Fluent
 * reader, settings files, wedge generator.  They replace main.cpp's setup
 * (main.cpp:9-53) for FFI callers.
 * ------------------------------------------------------------------------- */
typedef struct psst_case psst_case;

/* Reads a Fluent ASCII mesh and the five settings files `<prefix>.inp/.prop/
 * .par/.bvp/.btq` (and `.sav` when .inp asks for a restart), like
 * psstAnsysDirector::construct_grid + create_finite_volume_grid +
 * psstGridFV::solver_parameters. */
psst_status psst_case_open(const char *mesh_path, const char *case_prefix, psst_case **out);
/* Same, for a wedge mesh generated in memory (see psst_wedge_params). */
typedef struct psst_wedge_params {
    int32_t nx, ny, nz;           /* cells along x, y, z (nz = 0: 2D) */
    double length, height, depth;
    double ramp_start, ramp_length, ramp_deg;
    double jitter;                /* node jitter, fraction of local spacing */
    uint64_t jitter_seed;
    int32_t numbering;            /* 0 column-major, 1 RCM, 2 random permutation */
    uint64_t perm_seed;
    int32_t tets;                 /* 3D: split hexes into 6 tetrahedra */
    int32_t periodic_z;           /* 3D: z = 0 and z = depth faces form one periodic pair */
} psst_wedge_params;
psst_status psst_wedge_write_msh(const psst_wedge_params *p, const char *path);
psst_status psst_case_open_wedge(const psst_wedge_params *p, const char *case_prefix, psst_case **out);

psst_status psst_case_mesh(const psst_case *c, psst_mesh_desc *mesh);
psst_status psst_case_settings(const psst_case *c, psst_settings *settings);
/* Initial primitive state [(dim+3)][N+G]. */
psst_status psst_case_initial_state(const psst_case *c, double *prim);
/* info[0..7] = dim, N, G, F, rk_levels, first iteration, last iteration, save step */
psst_status psst_case_info(const psst_case *c, int64_t *info);
/* Raw mesh arrays (vertices, faces, cells, zones) of the case, for writers
 * and external checkers: counts[0..3] = vertices, faces, cells, zones. */
psst_status psst_case_raw_counts(const psst_case *c, int64_t *counts);
psst_status psst_case_raw(const psst_case *c, double *xyz, int8_t *face_nv, int64_t *face_v, int64_t *face_c,
                          int64_t *face_zone, int64_t *cell_zone, int64_t *zone_id, int64_t *zone_range,
                          int64_t *zone_bc, char *zone_names, int32_t names_len);
/* One zone of a mesh flattened by the caller: its Fluent id and type
 * (1 fluid cells, 2 interior faces, a boundary type, 31/32 parents), its
 * element count (cells for types 1 and 32, faces otherwise) and the name the
 * .bvp/.btq files use (psstFaceStorage / psstVolumeStorage,
 * psstFiniteVolumeGridTools.h:336-360). */
typedef struct psst_zone {
    int64_t id;
    int32_t type;
    int32_t count;
    const char *name;
} psst_zone;

/* Reads the five settings files `<prefix>.inp/.prop/.par/.bvp/.btq` (and
 * `.sav` when .inp asks for a restart) over a mesh the caller flattened
 * itself, e.g. from the reference's own psstGridFV with
 * include/psst_ref_adapter.hpp: the settings and initial state that
 * psstGridFV::solver_parameters derives (psstFiniteVolumeGrid.cpp:36-194,
 * :403-717).  zones lists every zone in ascending id; mesh->ghost_bc indexes
 * it.  The case has no Fluent file behind it, so the writers and
 * psst_case_solve are PSST_UNSUPPORTED for it (the reference's own
 * save_data writes the files). */
psst_status psst_case_from_mesh(const psst_mesh_desc *mesh, int32_t n_zones, const psst_zone *zones,
                                const char *case_prefix, psst_case **out);

/* Writes the reference's `.sav` checkpoint for a primitive state. */
psst_status psst_case_write_sav(const psst_case *c, const double *prim, int64_t iter_value, const char *path);
/* Writes the reference's `.dat` field file for a primitive state [(dim+3)][N+G]:
 * p, T, rho, u, v (, w) of every interior and ghost cell in Fluent data
 * sections (psstAnsysOutputManager::private_write_field,
 * psstOutputManager.cpp:71-128). */
psst_status psst_case_write_dat(const psst_case *c, const double *prim, const char *path);
/* Writes the reference's `.cas` file of the case's mesh: vertices, faces and
 * cells per zone, the trees of an adapted mesh, zone names
 * (psstAnsysOutputManager::private_write_cas_file, psstOutputManager.cpp:379-398). */
psst_status psst_case_write_cas(const psst_case *c, const char *path);
/* Replaces the settings of a case (e.g. to change the scheme or precision).
 * A boundary table in `settings` replaces the case's entry for entry; one with
 * a different n_bc is PSST_BAD_ARG and leaves the case unchanged. */
psst_status psst_case_set_settings(psst_case *c, const psst_settings *settings);
void psst_case_close(psst_case *c);
const char *psst_case_last_error(void);

/* Runs a whole case the way the reference's main() does: explicit_solve for
 * the .inp iteration count on the device, writing `<prefix>_residual.dat`
 * and, at every save, `<prefix>.dat`, `<prefix>.cas` and `<prefix>.sav` in
 * the reference's formats (psstFiniteVolumeGrid.cpp:240-400). */
psst_status psst_case_solve(psst_case *c);

#ifdef __cplusplus
}
#endif
#endif /* PSST_GPU_H */
