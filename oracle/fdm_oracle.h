/*
 * fdm_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU FP64 restatement of the reference solver's algorithm (2-D Euler, WCNS3 flux
 * derivative, WLS-WENO level-jump interpolation, SSP-RK3, quadtree AMR), written in plain
 * C from the reference's Fortran sources (the src directory). It is the checker for the product
 * library: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * The product (libfdmamr) never links, loads or calls anything here.
 *
 * Every routine cites the reference file:line it restates. Floating-point operations are
 * written in the reference's evaluation order (Fortran left-to-right for equal precedence,
 * matmul as an ascending-k sum) and the library is built with -ffp-contract=off, so results
 * are reproducible bit-for-bit.
 */
#ifndef FDM_ORACLE_H
#define FDM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Boundary kinds. 6 = periodic is the reference's only kind (module.F90:7,
 * solution.F90:81-104). 1 (transmissive) and 2 (fixed inflow state) are new: the reference
 * has no counterpart; they are needed by the shock configurations. */
enum { ORC_BC_TRANSMISSIVE = 1, ORC_BC_INFLOW = 2, ORC_BC_PERIODIC = 6 };

/* Initial conditions. 0 = isentropic vortex (solution.F90:31-69); 1 = seeded 4-quadrant
 * Riemann problem; 2 = shock + bubbles; 3 = density wave (the exact solution that
 * solution.F90:241-264 compares against). 1-3 are new. */
enum { ORC_IC_VORTEX = 0, ORC_IC_RIEMANN4 = 1, ORC_IC_SHOCK_BUBBLES = 2, ORC_IC_DENSITY_WAVE = 3 };

typedef struct OrcParams {
    /* physical domain before the 3-cell ghost extension (MAIN.F90:35-36,55-56) */
    double xmin, xmax, ymin, ymax;
    double h;              /* base spacing (MAIN.F90:37) */
    double gamma;          /* MAIN.F90:38 (pass (double)1.4f for the reference literal) */
    double Ma;             /* MAIN.F90:61 */
    double tstep, cfl, t_end;
    int32_t iflag_localdt; /* solution.F90:195-197 */
    int32_t iflag_amr, lvl_amr, pstep_adapt, nstep_adapt;
    double refine_goal_rot, refine_goal_div, coarse_goal_rot, coarse_goal_div;
    int32_t amr_sensor;    /* 0: rot OR div (solution.F90:525); 1: density gradient only */
    int32_t bc_xm, bc_xp, bc_ym, bc_yp; /* kind_xmbc, kind_xpbc, kind_ymbc, kind_ypbc */
    double inflow[4];      /* rho,u,v,p for ORC_BC_INFLOW */
    int32_t ic_kind;
    uint64_t seed;
    double ic_param[8];    /* IC-specific (see fdm_oracle.c: orc_ic_*) */
    int32_t pre_adapt;     /* refine the initial condition this many times before step 1 */
    /* WLS-WENO frame: 0 = absolute coordinates (XuanGua.F90:224-229, the reference);
     * 1 = coordinates relative to the WLS cell centre, scaled by its size (new; the absolute
     * frame is numerically singular on fine meshes far from the origin) */
    int32_t wls_local;
} OrcParams;

typedef struct Orc Orc;

Orc*    orc_new(const OrcParams* p);
void    orc_free(Orc* o);
/* creat_mesh + idneighbor + initial (+ optional IC pre-adaptation) + IDgrid */
int32_t orc_init(Orc* o);
/* one pass of the solute loop body (solution.F90:183-235); 0 ok, -1 NaN residual */
int32_t orc_step(Orc* o);
/* Time_advance_RK only, with caller-supplied dt and lambda (no AMR, no lambda update) */
int32_t orc_advance_rk3(Orc* o, double dt, double lam_x, double lam_y);

int32_t orc_n_leaf(const Orc* o);
int32_t orc_n_bc(const Orc* o);
int32_t orc_n_node(const Orc* o);
int32_t orc_base_dims(const Orc* o, int32_t* m, int32_t* n);
/* interior leaves in cell_NS order: centre, level, canonical node id, primitive (rho,u,v,T,p) */
int32_t orc_get_leaves(const Orc* o, double* x, double* y, int32_t* lvl, int32_t* node, double* prim5);
/* ghost leaves in cell_BC order */
int32_t orc_get_bc_leaves(const Orc* o, int32_t* node, double* prim5);
/* overwrite the primitive state of interior leaves (cell_NS order), then refill the ghosts */
int32_t orc_set_leaves(Orc* o, const double* prim5);
/* one forward-Euler substep from the current state: Ut of every interior leaf (cell_NS
 * order); the state is not changed */
int32_t orc_substep(Orc* o, double dt, double lam_x, double lam_y, double* prim5);
int32_t orc_get_scalars(const Orc* o, double* out8); /* tt, tstep, lam_x, lam_y, l_max, it, 0, 0 */
int32_t orc_get_residual(const Orc* o, double* res_max4, double* res_rms4);

/* Flattened tree in canonical numbering: base cells 0..m*n-1 in reference order, then a
 * pre-order walk (son1..son4) of every refined base cell. son[k] = first-son id (sons are
 * consecutive son1..son4) or -1. prim7 = U(1:7) of every node (internal nodes keep the value
 * they had when refined, as the reference does). */
int32_t orc_tree_export(const Orc* o, int32_t* father, int32_t* son, int32_t* lvl, int32_t* sp,
                        int32_t* num2, int32_t* cross, double* x, double* y, double* prim7);

/* Stencil classification for interior leaf i (cell_NS order), position k in
 * {0:+1x,1:-1x,2:+1y,3:-1y,4:+2x,5:-2x,6:+2y,7:-2y}.
 * kind: 0 copy of leaf `node`; 1 parent WLS-WENO on `node`; 2 parent interpA fallback;
 *       3 child WLS-WENO on coarse leaf `node` at (px,py); 4 child fallback copy of `node`.
 * nb receives up to 9 canonical ids of the WLS stencil (list1 then list2); m1,m2 sizes. */
int32_t orc_stencil_info(const Orc* o, int32_t i, int32_t k, int32_t* kind, int32_t* node,
                         double* px, double* py, int32_t* m1, int32_t* m2, int32_t* nb);
/* value the reference's Tip/Tim/.../Tjmm produce for that position (var(1:7); var(4) is
 * 0 where the reference leaves it undefined) */
int32_t orc_stencil_value(const Orc* o, int32_t i, int32_t k, double* var7);

/* per-leaf vorticity / density-gradient sensors (solution.F90:424-466,558-580), cell_NS
 * order, computed on the current state; returns Ne and writes Srot,Sdiv */
int32_t orc_sensors(Orc* o, double* rot, double* div, double* s_rot_div);
/* amr_regrid with the interior leaves' sensors given in cell_NS order (no ghost refill) */
int32_t orc_regrid_sensors(Orc* o, const double* rot, const double* div);

/* ---- kernel-level restatements (pure functions) ---- */
void    orc_wcns_z3(const double* u1, const double* u2, const double* u3, const double* R,
                    double* hp);                                   /* nnd_center_RK.F90:310-353 */
void    orc_eigen(int32_t dir, const double* Ucons, double gamma, double cv, double* R,
                  double* Rinv);                                   /* nnd_center_RK.F90:114-169,236-291 */
/* dE (dir 0) or dF (dir 1) from a 5-point primitive column prim5x5[pt*5+var] */
void    orc_flux_derivative(int32_t dir, const double* prim5x5, double d, double lam, double gamma,
                            double cv, double* dEF);               /* nnd_center_RK.F90:66-112,187-234 */
int32_t orc_brinv(int32_t n, const double* a, double* r);          /* XuanGua.F90:593-660; returns flag */
int32_t orc_get_direction(double xx, double yy, double x, double y); /* XuanGua.F90:33-43 */
/* WLSI_US_WENO on explicit stencil points (XuanGua.F90:179-356). pts: m rows of
 * (x, y, rho, u, v, p); first m1 rows are list 1. cell box from (cx, cy, lvl). out4 = rho,u,v,p */
void    orc_wls_weno(int32_t m, int32_t m1, const double* pts, double xx, double yy, double cx,
                     double cy, int32_t lvl, double h, double x_factor, double* out4);

/* smoothness indicator of compute_smoothness (XuanGua.F90:309-355): order 2 takes 6
 * coefficients, order 1 takes 3; box of the cell (cx, cy, lvl) */
double  orc_smoothness(int32_t order, const double* coeff, double cx, double cy, int32_t lvl,
                       double h, double x_factor);
/* the 8 cached neighbours (idnei, neighbor.F90:412-449) of canonical node `node`:
 * up, down, right, left, upleft, upright, downleft, downright; -1 = null */
int32_t orc_neighbors(const Orc* o, int32_t node, int32_t* out8);

uint64_t orc_splitmix64(uint64_t* state);

/* ---- test harness ---- */
/* OpenMP threads of the leaf loop (n <= 0 keeps the current count); returns the count */
int32_t orc_set_threads(int32_t n);
/* Refine_cell (mesh.F90:199-260) on canonical node `node`; then orc_rebuild */
int32_t orc_refine(Orc* o, int32_t node);
/* idneighbor + IDgrid after hand edits (reinit_ic: evaluate the initial condition again on
 * every leaf), then the ghost fill */
int32_t orc_rebuild(Orc* o, int32_t reinit_ic);
/* U(1:7) of every node in canonical numbering; ghosts are not refilled */
int32_t orc_set_node_values(Orc* o, const double* prim7);
/* replace the tree by a flattened one in canonical numbering (base cells 0..m*n-1 of this
 * oracle's m, n); idneighbor + IDgrid follow, ghosts keep the given values */
int32_t orc_tree_import(Orc* o, int32_t n_node, const int32_t* father, const int32_t* son,
                        const int32_t* lvl, const int32_t* sp, const int32_t* num2, const int32_t* cross,
                        const double* x, const double* y, const double* prim7);
/* orc_neighbors of every node: out8[8 * canonical id + k] */
int32_t orc_neighbors_all(const Orc* o, int32_t* out8);
/* orc_stencil_info / orc_stencil_value of every interior leaf and position (N*8 entries;
 * nb N*8*9, var N*8*7, either may be NULL) */
int32_t orc_stencil_all(const Orc* o, int32_t* kind, int32_t* node, double* px, double* py, int32_t* m1,
                        int32_t* m2, int32_t* nb, double* var);
/* Time_advance_RK on interior leaves [i0, i1) only, no ghost fill (CPU-baseline timing) */
int32_t orc_rk3_range(Orc* o, int32_t i0, int32_t i1, double dt, double lam_x, double lam_y);

#ifdef __cplusplus
}
#endif
#endif
