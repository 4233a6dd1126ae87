/*
 * fvh_oracle -- CPU restatement of fvhyper's explicit residual-and-update
 * path, in plain C99.  TEST INFRASTRUCTURE ONLY: the product never links,
 * loads or calls it.  Tests, the smoke check and the bench's CPU-baseline leg
 * use it as the checker the HIP library is compared against.
 *
 * Every function follows the reference operation by operation (same
 * association, same summation order, std::min/std::max semantics, no
 * contraction) so that, on the same mesh and state, it reproduces the
 * compiled reference bit for bit.  Citations are to files under the
 * reference tree.
 *
 * Layout: states are the reference's AoS arrays, q[4*cell + k], over the
 * n_cells = real + boundary-ghost cells of the reference numbering.
 */
#ifndef FVH_ORACLE_H
#define FVH_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_FLUX_ROE_SPLIT = 0, ORC_FLUX_ROE_JOINED = 1, ORC_FLUX_ROE_VISCOUS = 2 };
enum { ORC_PHI_ZERO = 0, ORC_PHI_MICHALAK = 1, ORC_PHI_VENKAT = 2, ORC_PHI_MINMOD = 3 };
enum { ORC_DT_FIXED = 0, ORC_DT_SUM_CFL = 1, ORC_DT_EDGE_MIN = 2, ORC_DT_EDGE_MIN_VISCOUS = 3,
       ORC_DT_SUM_VISCOUS = 4 };
enum { ORC_BC_COPY = 0, ORC_BC_MIRROR = 1, ORC_BC_INLET_OUTLET = 2, ORC_BC_FARFIELD = 3, ORC_BC_NOSLIP = 4,
       ORC_BC_FARFIELD_V = 5 };

#define ORC_MAX_ALPHA 8
#define ORC_MAX_BC 16

typedef struct {
    int flux;            /* ORC_FLUX_* */
    int phi;             /* ORC_PHI_* */
    double phi_yt;       /* Michalak threshold y_t */
    int dt_kind;         /* ORC_DT_* */
    double dt_fixed;
    double cfl;
    int grad, lim, lin, gdt;   /* solver:: flags (explicit.h:35-40) */
    int pos;             /* extension: first-order face state where the
                            reconstruction has rho <= 0 or p <= 0 */
    double K;            /* limiter_k_value (explicit.cpp:27) */
    double gamma;
    int n_alpha;
    double alpha[ORC_MAX_ALPHA];
    int n_bc;
    int bc_kind[ORC_MAX_BC];
    double bc_param[ORC_MAX_BC][4];   /* freestream rho, u, v, p */
    int smooth;          /* solver::smooth_residuals (explicit.h:40) */
    double mu, pr, cp, r;   /* consts:: of the viscous examples (flatPlate/main.cpp:28-35) */
    int diff;            /* solver::diffusive_gradients (explicit.h:38) */
    /* extension: smoothed kinks (the library's FVH_FLAG_SMOOTH_KINKS), no
     * reference counterpart: |x|_w = x^2/sqrt(x^2+w^2) and min/max_w in the
     * entropy cure, the summed-CFL wave speeds and the limiter; the limiter's
     * extremum chosen by a blend on the sign of dqg; the global step as the
     * p-norm (sum dt_i^-128)^(-1/128).  Widths per conserved variable, for
     * wave speeds and for the limiter's running minimum. */
    int kinks;
    double kw[4], kwv, kwl;
} orc_config;

typedef struct {
    int32_t n_nodes;
    double *x, *y;
    int32_t n_real, n_cells, n_edges, n_bnd;
    int32_t *edge_nodes;   /* 2 per edge */
    int32_t *edge_cells;   /* 2 per edge */
    double *edge_len, *edge_nx, *edge_ny, *edge_cx, *edge_cy;
    int32_t *cell_nodes;   /* 4 per cell */
    uint8_t *cell_tri;
    double *cell_area, *cell_cx, *cell_cy;
    int32_t *bnd_edge;     /* reference boundaryEdges */
    int32_t *bnd_bc;       /* index into orc_config bc table */
} orc_mesh;

/* Build the reference numbering and geometry from file-order arrays.
 * cells: 4 node indices per cell, the 4th -1 for triangles.
 * bnd: 2 node indices per boundary line element, bnd_bc its BC table slot.
 * Returns NULL and sets *err (1 = boundary edge not in mesh) on failure. */
orc_mesh* orc_mesh_build(int32_t n_nodes, const double* x, const double* y,
                         int32_t n_cells, const int32_t* cells,
                         int32_t n_bnd, const int32_t* bnd, const int32_t* bnd_bc,
                         int* err);
void orc_mesh_free(orc_mesh* m);

/* Per-phase functions (reference semantics, AoS q of 4*n_cells). */
void orc_calc_dt(const orc_mesh* m, const orc_config* c, const double* q, double* dt);
void orc_min_dt(const orc_mesh* m, double* dt);
/* the limiter's sigma blend at y = spread / K3a - 1 in (0, 1] */
double orc_sigma(const orc_config* cf, double y);
/* extension (smoothed kinks): the p-norm form of min_dt */
void orc_min_dt_soft(const orc_mesh* m, double* dt);
void orc_calc_gradients(const orc_mesh* m, const double* q, double* gx, double* gy);
void orc_calc_limiters(const orc_mesh* m, const orc_config* c, const double* q,
                       const double* gx, const double* gy,
                       double* lim, double* qmin, double* qmax);
void orc_calc_time_derivatives(const orc_mesh* m, const orc_config* c, const double* q,
                               const double* gx, const double* gy, const double* lim,
                               double* qt);
void orc_update_cells(int64_t n, double* q, const double* ql, const double* qt,
                      const double* dt, double a);
void orc_update_bounds(const orc_mesh* m, const orc_config* c, double* q,
                       const double* gx, const double* gy, const double* lim);
void orc_calc_residuals(const orc_mesh* m, const double* qt, double* R);
/* smooth_residuals (explicit.cpp:32-66) in place on qt. */
void orc_smooth_residuals(const orc_mesh* m, double* qt);

/* Scalar plugin kinds, for known-answer tests.  orc_flux passes zero
 * gradients to the viscous kind; orc_flux_grad takes the face gradient. */
void orc_flux(const orc_config* c, double* f, const double* qi, const double* qj,
              const double* n);
void orc_flux_grad(const orc_config* c, double* f, const double* qi, const double* qj,
                   const double* gx, const double* gy, const double* n);
/* gradient_for_diffusion (explicit.cpp:70-109) */
void orc_face_gradient(double* gx, double* gy, const double* gxi, const double* gyi,
                       const double* gxj, const double* gyj, const double* qi, const double* qj,
                       const double* ci, const double* cj);
double orc_phi(const orc_config* c, double r);
void orc_bc(const orc_config* c, int slot, double* b, const double* qint, const double* n);

/* run() of explicit.cpp:657-817 with the config's alpha list.  q holds the
 * initial solution over all cells on entry and the final state on exit.
 * stop_var < 0 is the reference's stop test (max_k R[k] < tol,
 * explicit.cpp:718-723); stop_var = k tests R[k] alone.
 * out: step count, final time, R[4] (normalised), R0[4]. */
int orc_run(const orc_mesh* m, const orc_config* c, double* q,
            uint32_t max_step, double max_time, double tol, int32_t stop_var, uint32_t print_interval,
            uint32_t* steps_out, double* time_out, double* R_out, double* R0_out,
            double* qt_out);

#ifdef __cplusplus
}
#endif
#endif
