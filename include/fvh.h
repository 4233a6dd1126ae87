/*
 * fvh.h -- C ABI of the MI355X (gfx950) explicit finite-volume library.
 *
 * One C-ABI shared library (libfvh.so) replaces fvhyper's explicit
 * residual-and-update path on one GPU.  Plain pointers and sizes only; every
 * entry point returns FVH_OK (0) or a negative FVH_E* code and leaves a
 * message for fvh_last_error().  No C++ exception crosses this boundary.
 *
 * States cross the boundary in the reference layout: AoS doubles,
 * q[4*cell + k], over the reference's cell numbering (real cells in file
 * order, then one boundary ghost cell per boundary element,
 * fvhyper/src/mesh.cpp:623-669).  The library renumbers cells internally
 * (reverse Cuthill-McKee) and inverts the permutation on the way out.
 *
 * Threading: a solver owns one HIP stream; a handle is not thread-safe, two
 * handles may be used from two threads.
 */
#ifndef FVH_H
#define FVH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FVH_ABI_VERSION 3

enum {
    FVH_OK = 0,
    FVH_E_ARG = -1,       /* invalid argument */
    FVH_E_IO = -2,        /* file cannot be read or written */
    FVH_E_FORMAT = -3,    /* malformed or unsupported mesh or restart file */
    FVH_E_MESH = -4,      /* inconsistent mesh (boundary edge not in mesh, ...) */
    FVH_E_BC = -5,        /* physical name without a boundary condition */
    FVH_E_HIP = -6,       /* HIP runtime error */
    FVH_E_UNSUPPORTED = -7,
    FVH_E_STATE = -8      /* non-finite state detected */
};

/* Roe flux with the entropy cure; the two reference spellings differ in the
 * association of the central flux, which changes the last bits. */
enum { FVH_FLUX_ROE_SPLIT = 0,   /* examples/shockTube/main.cpp:68-135, naca/main.cpp:71-137 */
       FVH_FLUX_ROE_JOINED = 1,  /* examples/forwardStep/main.cpp:72-135 */
       /* the split Roe flux plus the laminar viscous terms of
        * examples/flatPlate/main.cpp:126-225 (= nacaTransient/main.cpp:130-229),
        * constants mu, pr, cp, r_gas of the config; the face gradient is that of
        * gradient_for_diffusion (explicit.cpp:70-109) with FVH_FLAG_DIFFUSIVE,
        * zero without */
       FVH_FLUX_ROE_VISCOUS = 2 };
enum { FVH_PHI_ZERO = 0, FVH_PHI_MICHALAK = 1, FVH_PHI_VENKATAKRISHNAN = 2, FVH_PHI_MINMOD = 3 };
enum { FVH_DT_FIXED = 0,         /* shockTube/main.cpp:140-147 */
       FVH_DT_SUM_CFL = 1,       /* forwardStep/main.cpp:140-185 */
       FVH_DT_EDGE_MIN = 2,      /* naca/main.cpp:142-183 */
       FVH_DT_EDGE_MIN_VISCOUS = 3,  /* flatPlate/main.cpp:230-278 */
       FVH_DT_SUM_VISCOUS = 4    /* nacaTransient/main.cpp:234-286 */ };
enum { FVH_BC_COPY = 0,          /* shockTube/main.cpp:158-163 */
       FVH_BC_MIRROR = 1,        /* forwardStep/main.cpp:262-269 */
       FVH_BC_INLET_OUTLET = 2,  /* forwardStep/main.cpp:194-261 */
       FVH_BC_FARFIELD = 3,      /* naca/main.cpp:192-259 */
       FVH_BC_NOSLIP = 4,        /* flatPlate/main.cpp:361-366: momentum reversed */
       FVH_BC_FARFIELD_V = 5     /* flatPlate/main.cpp:286-353: the farfield with
                                    b[2] = b[0]*(...) in both subsonic branches */ };
/* Precision: FP64 throughout (reference-exact); FP32 states, geometry and
 * arithmetic; FP32 states with FP64 arithmetic and FP64 geometry (each stage
 * rounds the stored state to FP32, nothing else is rounded to FP32). */
enum { FVH_F64 = 0, FVH_F32 = 1, FVH_F32_F64 = 2 };
/* Device cell order (results do not depend on it, only locality does). */
enum { FVH_ORDER_FILE = 0,       /* reference numbering */
       FVH_ORDER_RCM = 1,        /* reverse Cuthill-McKee */
       FVH_ORDER_TILES = 2       /* recursive coordinate bisection into 256-cell tiles */ };
/* Stage kernels.  Second order runs a gradient/limiter pass and a flux pass
 * per stage, with the flux pass in one of two forms: each tile of 256 cells
 * computes every edge flux it touches once into LDS (TILE), or each cell
 * recomputes the flux of each of its edges (CELL).  FUSED does the whole
 * stage in one pass per tile, recomputing the gradients of the tile's
 * neighbour ring in LDS instead of exchanging face states through HBM.
 * AUTO picks FUSED for linear reconstruction when the halo fits in LDS,
 * TILE otherwise.  All forms give identical results. */
enum { FVH_FLUX_MODE_AUTO = 0, FVH_FLUX_MODE_CELL = 1, FVH_FLUX_MODE_TILE = 2, FVH_FLUX_MODE_FUSED = 3 };

/* solver:: flags of fvhyper/explicit.h:35-40 */
#define FVH_FLAG_GRADIENTS   1u   /* do_calc_gradients */
#define FVH_FLAG_LIMITERS    2u   /* do_calc_limiters */
#define FVH_FLAG_LINEAR      4u   /* linear_interpolate */
#define FVH_FLAG_GLOBAL_DT   8u   /* global_dt */
#define FVH_FLAG_DIFFUSIVE  64u   /* diffusive_gradients: face gradients for the
                                     viscous flux (explicit.cpp:70-109, 324-338) */
#define FVH_FLAG_SMOOTH     32u   /* smooth_residuals: two Jacobi sweeps of implicit
                                     residual smoothing, eps 0.6 (explicit.cpp:32-66,
                                     743); not with the FUSED stage form */
/* Extension without a reference counterpart: where a limited linear
 * reconstruction yields a face state with rho <= 0 or p <= 0, that face uses
 * the cell value instead (first order).  Off for reference parity. */
#define FVH_FLAG_POSITIVITY 16u
/* Extension without a reference counterpart, for derivative checks of the
 * adjoint: every kink of the second-order path becomes an infinitely
 * differentiable function of width w,
 *   |x| -> x^2 / sqrt(x^2 + w^2),  min/max(a, b) -> (a + b -/+ |a - b|_w) / 2
 * (exact at x = 0 and a = b), in the entropy cure's |min(V-c, VL-c)|, |V|
 * and |max(V+c, VR+c)|, the summed-CFL wave speeds c + |u.n| and their max,
 * and the limiter's neighbour extrema and running minimum; the limiter's
 * sigma blend becomes 1 - y^5 (126 - 420y + 540y^2 - 315y^3 + 70y^4)
 * (four times differentiable where the reference's cubic is once); the
 * limiter's choice of qmax or qmin by the sign of the gradient projection dqg becomes
 * a blend by dqg / sqrt(dqg^2 + w^2) with phi(d/dqg) written without the
 * quotient and regularised by w^2 (Venkatakrishnan's original form); the
 * global time step's minimum (CFL kinds) becomes the p-norm
 * (sum_i dt_i^-128)^(-1/128).  Widths: kink_width times kink_scale.  Takes
 * the Venkatakrishnan limiter with linear reconstruction, the summed-CFL or
 * fixed time step and FVH_F64; the positivity fallback stays a switch. */
#define FVH_FLAG_SMOOTH_KINKS 128u

#define FVH_MAX_ALPHA 8
#define FVH_MAX_BC 16
#define FVH_NAME_LEN 64

typedef struct fvh_bc_spec {
    char name[FVH_NAME_LEN];   /* physical name in the mesh (boundaries::bounds key) */
    int32_t kind;              /* FVH_BC_* */
    double rho, u, v, p;       /* freestream of the characteristic kinds */
} fvh_bc_spec;

typedef struct fvh_config {
    int32_t precision;         /* FVH_F64 (reference-exact), FVH_F32 or FVH_F32_F64 */
    int32_t flux;              /* FVH_FLUX_* */
    double gamma;
    int32_t phi;               /* FVH_PHI_* (limiter_func) */
    double phi_yt;             /* Michalak threshold, 2 in the reference */
    double K;                  /* limiter_k_value, explicit.cpp:27 (7.5) */
    int32_t dt_kind;           /* FVH_DT_* (calc_dt) */
    double dt_fixed;
    double cfl;
    uint32_t flags;            /* FVH_FLAG_* */
    int32_t n_alpha;           /* RK stage coefficients, explicit.cpp:681-687 */
    double alpha[FVH_MAX_ALPHA];
    int32_t n_bc;
    fvh_bc_spec bc[FVH_MAX_BC];
    int32_t renumber;          /* device cell order: FVH_ORDER_* */
    int32_t flux_mode;         /* FVH_FLUX_MODE_* */
    int32_t tile_cells;        /* cells per tile (workgroup), 32..256; 0 = automatic (the default): 256, or
                                  fewer when the fused stage pass runs 256-lane workgroups and a tile with
                                  its first neighbour ring would outnumber the lanes */
    /* laminar viscous constants of FVH_FLUX_ROE_VISCOUS and the viscous time
     * steps (consts:: of flatPlate/main.cpp:28-35): viscosity, Prandtl number,
     * heat capacity, gas constant */
    double mu, pr, cp, r_gas;
    /* workgroup size of the fused stage pass: 0 = by the tile halo's LDS
     * (256 lanes when four workgroups fit a CU, else 512), or 256, 512 */
    int32_t stage_threads;
    /* launch form of the fused stage pass: 0 = one workgroup per tile (the
     * default), 1 = persistent workgroups pulling tiles from per-XCD work
     * queues (same results) */
    int32_t stage_launch;
    /* FVH_FLAG_SMOOTH_KINKS: relative width (0: 1e-2) and the absolute scales
     * it multiplies: rho, rho u, rho v, rho E (limiter extrema and blend) and
     * a wave speed (entropy cure, time step); a scale of 0 takes the
     * freestream of the first inlet/outlet or farfield boundary (rho,
     * rho (|V| + c) twice, E, |V| + c), or 1 without one.  The limiter's running
     * minimum takes the relative width itself. */
    double kink_width;
    double kink_scale[5];
} fvh_config;

/* solverOptions of fvhyper/explicit.h:152-160 (minus the output fields) */
typedef struct fvh_run_options {
    double max_time;
    uint32_t max_step;
    uint32_t print_interval;
    double tolerance;
    int32_t stop_var;          /* -1: stop when max_k R[k] < tolerance, as explicit.cpp:718-723;
                                  k in 0..3: when R[k] < tolerance (e.g. 0, the density residual) */
    int32_t reserved;          /* 0 */
} fvh_run_options;

typedef struct fvh_status {
    uint32_t step;             /* steps taken */
    uint32_t stopped;          /* 1 once the stop test of explicit.cpp:718-726 held */
    double time;               /* sum of dt[cell 0] over steps (explicit.cpp:800) */
    double time_prev;          /* time before the last step (what the reference prints for it) */
    double R[4];               /* residual, normalised by R0 after step 0 */
    double R0[4];              /* step-0 residual (explicit.cpp:752-754) */
} fvh_status;

typedef struct fvh_mesh fvh_mesh;
typedef struct fvh_solver fvh_solver;

/* Message of the last failing call on this thread. */
const char* fvh_last_error(void);
int32_t fvh_abi_version(void);

/* ---- mesh (replaces fvhyper::mesh::read_file, fvhyper/mesh.h:193,
 *      fvhyper/src/mesh.cpp:803-855; single device, so no halo maps) ---- */

/* Read a gmsh MSH 4.1 ASCII file (the subset of mesh.cpp:264-617). */
int32_t fvh_mesh_read_msh(const char* path, fvh_mesh** out);

/* A 2-D SU2 ASCII mesh (triangles and quadrilaterals, line markers), as the
 * reference's examples/naca/NACA_M.su2; no counterpart in the reference, which
 * converts such files to MSH 4.1 with gmsh (examples/naca/naca.geo:1-5).
 * Cells, nodes and marker elements keep file order; marker tags are the
 * physical names bound to boundary kinds. */
int32_t fvh_mesh_read_su2(const char* path, fvh_mesh** out);

/* Same mesh from file-order arrays: x,y per node; cells 4 node indices each
 * (0-based, the 4th -1 for a triangle); bnd 2 node indices per boundary
 * element with bnd_tag indexing names[]. */
int32_t fvh_mesh_from_arrays(int32_t n_nodes, const double* x, const double* y,
                             int32_t n_cells, const int32_t* cells,
                             int32_t n_bnd, const int32_t* bnd, const int32_t* bnd_tag,
                             int32_t n_names, const char* const* names, fvh_mesh** out);
void fvh_mesh_free(fvh_mesh* m);

/* counts[0..5] = nodes, real cells, cells incl. ghosts, edges, boundary
 * elements, physical names */
int32_t fvh_mesh_counts(const fvh_mesh* m, int64_t* counts);

/* Copy one mesh member in reference order into dst (capacity n elements).
 * double fields: nodesX nodesY edgesLengths edgesNormalsX edgesNormalsY
 *   edgesCentersX edgesCentersY cellsAreas cellsCentersX cellsCentersY
 * int32 fields: edgesNodes(2/edge) edgesCells(2/edge) cellsNodes(4/cell)
 *   cellsIsTriangle boundaryEdges boundaryTags rcmPermutation tilePermutation
 *   (the permutations give each reference cell's device position)
 * Returns the number of elements (dst may be NULL to query it). */
int64_t fvh_mesh_export(const fvh_mesh* m, const char* field, void* dst, int64_t n);

/* Physical name idx into buf (NUL-terminated). */
int32_t fvh_mesh_name(const fvh_mesh* m, int32_t idx, char* buf, int32_t len);

/* ---- solver (replaces fvhyper::run and its phases, fvhyper/explicit.h:88-181) */

void fvh_config_default(fvh_config* c);

/* Bind the BC table by physical name (error FVH_E_BC as mesh.cpp:662-664),
 * upload geometry and allocate device state on the current HIP device. */
int32_t fvh_solver_create(const fvh_mesh* m, const fvh_config* c, fvh_solver** out);
void fvh_solver_free(fvh_solver* s);

/* State in the reference layout (4 doubles per cell, all cells). */
int32_t fvh_solver_set_state(fvh_solver* s, const double* q);
int32_t fvh_solver_get_state(fvh_solver* s, double* q);

/* Binary restart (no reference counterpart: the reference writes VTU only,
 * post.cpp:21-200).  save writes the state of every cell in the reference
 * layout plus each instance's run bookkeeping (step, time, residuals, the
 * step-0 residuals, stop flag); load restores both into a solver built on the
 * same meshes, so that fvh_solver_run_continue after load reproduces the
 * uninterrupted run bit for bit.  Format: "FVHSTAT1", int32 {1, precision,
 * instances, 0}, int64 value count, 104-byte record per instance, doubles. */
int32_t fvh_solver_save(fvh_solver* s, const char* path);
int32_t fvh_solver_load(fvh_solver* s, const char* path);

/* Fill ghost cells from boundary conditions with zero gradients
 * (update_bounds(q, 0, 0, 0), explicit.cpp:712) and reset step/time. */
int32_t fvh_solver_apply_bounds(fvh_solver* s);

/* One stage of every phase from the current state, results in reference
 * layout (explicit.cpp:729-747 with a single alpha); any output may be NULL.
 * dt: 1 value per cell; the others 4 per cell.  The state is not changed. */
int32_t fvh_solver_phases(fvh_solver* s, double alpha, double* dt, double* gx, double* gy,
                          double* lim, double* qt, double* qk);

/* ---- the phases of explicit.cpp one at a time (explicit.h:88-149), for
 * callers that compose their own loop.  Every array is in the reference
 * layout (4 values per cell over all cells: real cells, then ghosts); each
 * call copies its inputs to the device, runs the phase there and copies its
 * outputs back, so it costs PCIe transfers and is meant for composing, not
 * for speed.  FP64, single-mesh solvers; the solver's own state is not
 * touched.  The solver's config supplies the plugin kinds: flux, limiter
 * function, boundary kinds and the solver:: flags. */

/* calc_gradients (explicit.cpp:112-171): Green-Gauss gradients of the real
 * cells, 0 on the ghost cells. */
int32_t fvh_phase_gradients(fvh_solver* s, const double* q, double* gx, double* gy);
/* calc_limiters (explicit.cpp:174-258) from given gradients: the limiter of
 * every real cell, 1 on the ghost cells; qmin, qmax (may be NULL) the
 * neighbour minima and maxima of every cell. */
int32_t fvh_phase_limiters(fvh_solver* s, const double* q, const double* gx, const double* gy, double* lim,
                           double* qmin, double* qmax);
/* calc_time_derivatives (explicit.cpp:261-364) from given gradients and
 * limiters: qt of the real cells, 0 on the ghost cells. */
int32_t fvh_phase_time_derivatives(fvh_solver* s, const double* q, const double* gx, const double* gy,
                                   const double* lim, double* qt);
/* update_cells (explicit.cpp:368-378): q[i] = ql[i] + qt[i] dt[i] v over n
 * values (n at most 4 per cell).  s may be NULL (the reference's
 * update_cells takes no mesh): the call then uses temporary device buffers
 * on the current device. */
int32_t fvh_phase_update_cells(fvh_solver* s, double* q, const double* ql, const double* qt, const double* dt,
                               double v, int64_t n);
/* update_bounds (explicit.cpp:380-417): the ghost cells of q from the
 * boundary functions of the interior face states (given gradients and
 * limiters with linear_interpolate); the real cells are unchanged. */
int32_t fvh_phase_update_bounds(fvh_solver* s, double* q, const double* gx, const double* gy, const double* lim);
/* min_dt (explicit.cpp:555-564): every one of the n values becomes their
 * minimum (NaN if dt[0] is NaN, other NaNs ignored, as std::min from dt[0]). */
int32_t fvh_phase_min_dt(fvh_solver* s, double* dt, int64_t n);
/* The same minimum on a device array, in place, allocating nothing.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops).  dt: a device
 * pointer to n doubles on the current device (at most 2^31 values);
 * stream: a hipStream_t (NULL: the null stream); returns once the kernels
 * are queued.  The per-block minima go to a file-scope device array of 1024
 * doubles (8 KiB on each GPU the library runs on, held until the process
 * ends) shared by every call on that GPU: calls that overlap in time on one
 * GPU race. */
int32_t fvh_min_dt_device(double* dt, int64_t n, void* stream);

/* calc_residuals (explicit.cpp:481-552), one rank: R[k] = sqrt(sum over the
 * real cells of qt^2 A), summed in a fixed tree order (not ascending cell
 * order: equal to ~1e-15 relative). */
int32_t fvh_phase_residuals(fvh_solver* s, const double* qt, double* R);

/* The run() loop of explicit.cpp:657-817 from the current state: applies the
 * initial boundary update, then steps until the stop test holds.
 * Failure detection (no reference counterpart: the reference runs on with
 * NaN): every stage update checks the real cells' states, and when one has
 * rho <= 0, p <= 0 or a non-finite value, or a residual is not finite, the
 * run stops after the group of steps under way (at most 256) and returns
 * FVH_E_STATE naming the instance and step; the status and the state are
 * those at that point.  Applies to run, run_continue and run_log. */
int32_t fvh_solver_run(fvh_solver* s, const fvh_run_options* o, fvh_status* st);

/* run() in pieces, for callers that print or save between steps:
 * run_begin applies the initial boundary update and resets step/time;
 * run_continue executes up to n more steps of the same loop (stop test,
 * residuals on print steps) and returns the status after them. */
int32_t fvh_solver_run_begin(fvh_solver* s);
int32_t fvh_solver_run_continue(fvh_solver* s, const fvh_run_options* o, uint32_t n, fvh_status* st);

/* run() with its printed rows as a callback (explicit.cpp:752-785): applies
 * the initial boundary update, steps until the stop test holds, and calls
 * cb(step, status, user) after step 0 and after every step whose index is a
 * multiple of print_interval, with that step's residuals (R0 after step 0).
 * cb may be NULL. */
typedef void (*fvh_row_cb)(uint32_t step, const fvh_status* st, void* user);
int32_t fvh_solver_run_log(fvh_solver* s, const fvh_run_options* o, fvh_row_cb cb, void* user, fvh_status* st);

/* Measurement helper without a reference counterpart: the HBM bandwidth a
 * plain device-to-device copy of mbytes MB achieves on the current device,
 * read plus write bytes per second over reps launches (GB/s), the best of
 * a few launch shapes with plain and non-temporal accesses. */
int32_t fvh_stream_copy(double mbytes, int32_t reps, double* gbs);

/* Enqueue n full steps (dt, all RK stages, swap) without residuals or stop
 * test; returns without waiting.  Used for fixed-length timing. */
int32_t fvh_solver_advance(fvh_solver* s, int32_t n);
int32_t fvh_solver_synchronize(fvh_solver* s);
int32_t fvh_solver_status(fvh_solver* s, fvh_status* st);

/* Kernel timing with HIP events on the solver's stream.  While enabled, each
 * launch is bracketed by events; fvh_solver_kernel_times returns, per kernel
 * (names in fvh_solver_kernel_name), the launch count and total ms. */
int32_t fvh_solver_profile(fvh_solver* s, int32_t enable);
int32_t fvh_solver_kernel_times(fvh_solver* s, int32_t max, int64_t* counts, double* total_ms);
const char* fvh_solver_kernel_name(int32_t k);

/* Algorithmic bytes per RK stage for this mesh and config (the HBM roofline
 * numerator): first order C(3Vs+s+g) + E(3g+8) + B(2Vs+2g+4); second order
 * adds C(7Vs+2g) + E(2g). */
double fvh_solver_stage_bytes(const fvh_solver* s);

/* Cells per tile the solver uses (resolves tile_cells = 0).  No reference
 * counterpart: a tuning query. */
int32_t fvh_solver_tile_cells(const fvh_solver* s);

/* Layout of the fused stage pass (no reference counterpart: a tuning
 * query).  info[0..7] = fused pass in use (0/1), real tiles, LDS bytes per
 * workgroup, lanes per workgroup, largest local cell list (tile and two
 * neighbour rings), most edges of a tile, cells per tile, mean tile plus
 * real ring-1 cells. */
int32_t fvh_solver_stage_info(const fvh_solver* s, int64_t* info);

/* Capture each RK step (every kernel of it) in a HIP graph on first use and
 * replay it afterwards (one graph per state buffer, residual flag and run
 * options).  No reference counterpart: it removes per-launch host overhead
 * on small meshes.  Results are unchanged. */
int32_t fvh_solver_use_graphs(fvh_solver* s, int32_t enable);

/* Autotuner (no reference counterpart): times each candidate stage form
 * (TILE, FUSED, CELL) at each tile size (128, 192, 256 cells) on these meshes and
 * configs from state q (as for fvh_solver_set_state of the batch) for
 * `steps` steps, and returns the fastest in *best_mode / *best_tile.
 * report, if non-NULL, receives up to max_report rows (mode, tile,
 * ms per step; -1 for a candidate that does not apply); returns the number
 * of rows.  The configs' own flux_mode and tile_cells are ignored. */
int32_t fvh_autotune(int32_t n, const fvh_mesh* const* meshes, const fvh_config* configs, const double* q,
                     int32_t steps, int32_t* best_mode, int32_t* best_tile, double* report, int32_t max_report);

/* ---- discrete adjoint (no reference counterpart: fvhyper has no objective
 * and no derivatives).  The objective is a boundary pressure force after
 * n_steps steps of run() from state q (reference layout, all cells):
 *   J = sum over boundary elements b of p(ghost_b) (wx[b] n_x,b + wy[b] n_y,b) len_b
 * with p the ghost cell's pressure, n the boundary edge's outward normal and
 * wx, wy one weight per boundary element (reference order; NULL = 0).
 * fvh_adjoint returns J and dJ/dq (4 values per cell; ghost entries are 0,
 * since run() overwrites the ghost cells before the first step), computed by
 * reverse-mode kernels over a replay of the forward steps.  With grad_nodes
 * non-NULL it also returns dJ/d(node coordinates), (d/dx, d/dy) per node in
 * file order: the adjoint of every geometric quantity the run uses (normals,
 * lengths, areas, centres, Green-Gauss weights, ghost centres), chained to
 * the nodes; boundary nodes included (moving them moves the boundary).  Forward states
 * are kept at binomially placed step checkpoints: `snaps` slots (0 = as many
 * as mem_cap_bytes allows, at most one per step; then the budget left over
 * keeps the last steps' whole stage records, which the reverse reads without
 * replaying those steps).  FP64, single-mesh
 * solvers, global or fixed time step. */
typedef struct fvh_adjoint_stats {
    int32_t steps;             /* steps of the objective's run */
    int32_t snaps;             /* checkpoint slots used */
    int64_t forward_steps;     /* forward steps executed, replays included */
    double checkpoint_bytes;   /* device memory of the checkpoints */
    double work_bytes;         /* device memory of one step's replay and adjoint buffers */
    double forward_ms;         /* the plain forward run to the objective */
    double total_ms;           /* the whole gradient evaluation */
    int32_t recorded_steps;    /* last steps whose stage records the forward sweep kept (no replay) */
    int32_t reserved;
} fvh_adjoint_stats;

int32_t fvh_objective(fvh_solver* s, const double* q, int32_t n_steps, const double* wx, const double* wy,
                      double* J);
int32_t fvh_adjoint(fvh_solver* s, const double* q, int32_t n_steps, const double* wx, const double* wy,
                    int32_t snaps, double mem_cap_bytes, double* J, double* grad, double* grad_nodes,
                    fvh_adjoint_stats* st);

/* ---- batches: many independent meshes advanced together (no reference
 * counterpart; a parameter sweep in the reference is one process per mesh).
 * The handle is a fvh_solver and every fvh_solver_* call applies to all
 * instances at once.  Configs may differ only in their boundary tables
 * (e.g. a freestream per instance); each instance keeps its own step count,
 * time, residuals and stop test, and stops on its own.  States cross as the
 * concatenation of the instances' reference-layout states, in instance
 * order.  Each instance's results are bit-identical to a single-mesh solver
 * on that mesh.  fvh_solver_status reports instance 0 with `stopped` set
 * once all instances have stopped and `step` the largest step count. */
int32_t fvh_batch_create(int32_t n, const fvh_mesh* const* meshes, const fvh_config* configs,
                         fvh_solver** out);
int32_t fvh_batch_size(const fvh_solver* s);
/* info[0..2] = cells incl. ghosts, real cells, first cell in the concatenated state */
int32_t fvh_batch_instance(const fvh_solver* s, int32_t k, int64_t* info);
int32_t fvh_batch_status(fvh_solver* s, int32_t k, fvh_status* st);
/* Conserved totals sum_c A_c q_ck over each instance's real cells, summed in
 * FP64 in a fixed order (4 per instance), from the current state. */
int32_t fvh_batch_sums(fvh_solver* s, double* out);

#ifdef __cplusplus
}
#endif
#endif
