/*
 * fdmamr.h -- C-ABI of the MI355X (gfx950) hot path of the cell-based quadtree AMR
 * finite-difference Euler solver (reference: Jian-cd/FDM-cell-based-AMR-Cartesian, Fortran).
 *
 * The reference has no FFI: its hot path is the argument-less module procedure
 * Time_advance_RK (solution.F90:267-298) working on module globals and a pointer quadtree
 * (module.F90:24-42). This header is the boundary that replaces it. A Fortran driver binds
 * it with ISO_C_BINDING (fdm-cell-based-amr-cartesian_amd/fortran/fdmamr_c.f90); Python
 * binds it with ctypes (fdm-cell-based-amr-cartesian_amd/fdmamr/).
 *
 * Conventions
 *  - Every entry point returns 0 on success and a negative code on failure; the message is
 *    available from fdmamr_last_error. Codes: -1 invalid argument, -2 HIP error,
 *    -3 NaN residual (the reference stops, solution.F90:209-214), -4 positivity failure,
 *    -5 no device / library built without a usable GPU.
 *  - Host arrays are borrowed for the duration of the call only. Device memory and the
 *    context are owned by the library. One context = one HIP stream; not thread-safe per
 *    context; call from the driver's serial part (as the reference calls Time_advance_RK).
 *  - Leaf order is the reference's: interior leaves in cell_NS order (base cells row-major,
 *    sons son1,son2,son4,son3; solution.F90:593-670), ghost leaves in cell_BC order
 *    (solution.F90:672-693).
 *  - State arrays are primitive (rho,u,v,T,p) as in U(1:5) (module.F90:32), AoS: prim[5*i+k].
 */
#ifndef FDMAMR_H
#define FDMAMR_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fdmamr_ctx fdmamr_ctx;

/* boundary kinds: 6 periodic (the reference's BCcyc, module.F90:7); 1 transmissive and
 * 2 fixed inflow state are new (no reference counterpart) */
#define FDMAMR_BC_TRANSMISSIVE 1
#define FDMAMR_BC_INFLOW 2
#define FDMAMR_BC_PERIODIC 6

#define FDMAMR_DEVICE_HOST_ONLY (-1)

/* flux precision modes */
#define FDMAMR_FLUX_FP64 0   /* plain method, reference arithmetic */
#define FDMAMR_FLUX_FP32 1   /* WCNS reconstruction + LF flux in FP32, RK update in FP64 */

/* initial conditions (host-side, driver support): 0 isentropic vortex (solution.F90:31-69),
 * 1 seeded 4-quadrant Riemann, 2 shock + seeded bubbles, 3 density wave (1..3 are new) */
#define FDMAMR_IC_VORTEX 0
#define FDMAMR_IC_RIEMANN4 1
#define FDMAMR_IC_SHOCK_BUBBLES 2
#define FDMAMR_IC_DENSITY_WAVE 3

typedef struct fdmamr_params {
    /* physical domain before the 3-cell ghost extension (MAIN.F90:35-36,55-56) */
    double xmin, xmax, ymin, ymax;
    double h;                 /* base spacing (MAIN.F90:37) */
    double gamma;             /* MAIN.F90:38 */
    double Ma;                /* MAIN.F90:61, cv = 1/(gamma (gamma-1) Ma^2) */
    int32_t bc_xm, bc_xp, bc_ym, bc_yp;  /* kind_xmbc, kind_xpbc, kind_ymbc, kind_ypbc (MAIN.F90:48) */
    double inflow[4];         /* rho,u,v,p for FDMAMR_BC_INFLOW */
    int32_t flux_precision;   /* FDMAMR_FLUX_FP64 / FDMAMR_FLUX_FP32 */
    int32_t selector;         /* 0 off; 1 learned troubled-cell selector (FP32 mode only) */
    int32_t device;           /* HIP device ordinal, or FDMAMR_DEVICE_HOST_ONLY for a context
                                 that only builds meshes and tables (no device calls) */
    /* WLS-WENO fitting frame: 0 = absolute coordinates, as the reference
     * (XuanGua.F90:224-229); 1 = coordinates about the WLS cell centre scaled by its size
     * (new: the absolute frame's normal equations are numerically singular on fine meshes
     * far from the origin; the fitted values are the same in exact arithmetic, the
     * smoothness indicators are taken in the scaled frame) */
    int32_t wls_local;
} fdmamr_params;

/* ---- lifecycle ---- */
/* replaces set_parameter's derived quantities (MAIN.F90:55-61) and allocates a stream */
int32_t fdmamr_create(const fdmamr_params* p, fdmamr_ctx** out);
void    fdmamr_destroy(fdmamr_ctx* ctx);
int32_t fdmamr_last_error(const fdmamr_ctx* ctx, char* buf, int32_t len);

/* ---- mesh (host quadtree owned by the context; regrid-time, not per step) ---- */
/* creat_mesh (mesh.F90:9-47) + idneighbor (neighbor.F90:399-409) */
int32_t fdmamr_creat_mesh(fdmamr_ctx* ctx);
/* initial (solution.F90:9-29): host IC on every leaf, then initial_boundary. ic_param as in
 * the driver's config (vortex: none; Riemann: split x,y; shock: x_s, Mach, n_bubble,
 * r_min, r_max, ratio_min, ratio_max) */
int32_t fdmamr_initial(fdmamr_ctx* ctx, int32_t ic_kind, uint64_t seed, const double* ic_param8);
/* Replace the whole quadtree with a caller-flattened one (for drivers that own their tree):
 * n_node nodes; base cells are nodes 0..m*n-1 in the reference's cell(i) order; son[k] is the
 * id of son1 of node k (son1..son4 consecutive) or -1; lvl/sp/num/cross as type(grid)
 * (module.F90:24-35); prim7 = U(1:7) of every node. Rebuilds neighbour caches. The whole input
 * is validated first (father/son/sp/lvl/num consistency, ranges); on -1 the context keeps its
 * old tree. Either way the tables must be rebuilt (fdmamr_build_tables) before the next step. */
int32_t fdmamr_tree_import(fdmamr_ctx* ctx, int32_t m, int32_t n, int32_t n_node,
                           const int32_t* father, const int32_t* son, const int32_t* lvl,
                           const int32_t* sp, const int32_t* num2, const int32_t* cross,
                           const double* x, const double* y, const double* prim7);
/* Export in canonical numbering (base cells, then son blocks in a pre-order walk). Any
 * pointer may be NULL. Returns the node count through n_node. */
int32_t fdmamr_tree_export(fdmamr_ctx* ctx, int32_t* n_node, int32_t* father, int32_t* son,
                           int32_t* lvl, int32_t* sp, int32_t* num2, int32_t* cross, double* x,
                           double* y, double* prim7);
int32_t fdmamr_tree_size(fdmamr_ctx* ctx, int32_t* n_node);

/* IDgrid (solution.F90:593-693) + flattening of the neighbour tables into device-resident
 * leaf arrays: stencil ids for +-1,+-2 in x and y, WLS-WENO slot tables (FP64 coefficient
 * setup, XuanGua.F90:179-255), restriction and boundary tables. Uploads the host state. */
int32_t fdmamr_build_tables(fdmamr_ctx* ctx);
int32_t fdmamr_counts(fdmamr_ctx* ctx, int32_t* n_leaf, int32_t* n_ghost, int32_t* n_slot_wls,
                      int32_t* n_slot_restrict, int32_t* l_max);

/* ---- state transfer (at regrid / output only) ---- */
int32_t fdmamr_set_state(fdmamr_ctx* ctx, const double* prim5_leaf);   /* interior leaves */
int32_t fdmamr_get_state(fdmamr_ctx* ctx, double* prim5_leaf);         /* interior leaves */
int32_t fdmamr_get_ghost_state(fdmamr_ctx* ctx, double* prim5_ghost);  /* cell_BC leaves */
/* leaf geometry: centre x,y and level, interior leaves; canonical node ids optional */
int32_t fdmamr_leaf_geometry(fdmamr_ctx* ctx, double* x, double* y, int32_t* lvl, int32_t* node);

/* ---- the hot path ---- */
/* lambda/l_max loop of solute (solution.F90:185-193): max over interior leaves of |u|+c and
 * |v|+c. The caller keeps the running maximum, as the reference does. */
int32_t fdmamr_max_wavespeed(fdmamr_ctx* ctx, double* lam_x, double* lam_y, int32_t* l_max);
/* No counterpart in the reference: fdmamr_max_wavespeed for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same result, bit for bit; its partials go to a static device scratch array of 2 MiB that each
 * GPU holds until the process ends, so it calls no hipMalloc/hipFree. Meshes of more than
 * FDMAMR_NOALLOC_MAX_LEAVES interior leaves are rejected (-1). Calls that overlap on one GPU
 * share the scratch and race: serialise them (one context's calls are serial already). */
#define FDMAMR_NOALLOC_MAX_LEAVES 33554432
int32_t fdmamr_max_wavespeed_noalloc(fdmamr_ctx* ctx, double* lam_x, double* lam_y, int32_t* l_max);
/* Time_advance_RK (solution.F90:267-298): three SSP-RK3 stages, each = level-jump slot
 * values (WLS-WENO / restriction), fused WCNS3 x+y flux derivative + forward-Euler substep +
 * primitive RK blend (nnd_center_RK.F90:15-112, solution.F90:386-404), then the ghost fill
 * (solution.F90:288-293). res_max/res_rms_sum receive Res_max(1:4) and Res_rms(1:4) of the
 * last stage before the sqrt (nnd_center_RK.F90:31-36). Returns -4 when a stage's FP64 result
 * has rho <= 0 or p <= 0 at any leaf (new check, folded into the stage kernels), else -3 on a
 * NaN residual. */
int32_t fdmamr_advance_rk3(fdmamr_ctx* ctx, double dt, double lam_x, double lam_y,
                           double* res_max4, double* res_rms_sum4);
/* n passes of the solute loop body without AMR (solution.F90:183-215) kept on the device:
 * running-max lambda, dt = cfl/(lx/dx_min + ly/dy_min) when localdt != 0 (else fixed dt),
 * clamp to t_end, RK3. No host synchronisation inside; scalars[8] (may be NULL) receive
 * tt, dt, lam_x, lam_y, Res_rms(1:4) after the last step (sqrt applied). A step whose stages
 * leave rho <= 0 or p <= 0 at a leaf stops the clock (tt no longer advances) and the call
 * returns -4; a NaN residual does the same with -3 (the reference's stop, solution.F90:209-214). */
int32_t fdmamr_run_steps(fdmamr_ctx* ctx, int32_t n_steps, int32_t localdt, double cfl,
                         double dt_fixed, double t_end, double* tt_inout, double* lam_inout2,
                         double* scalars8);
/* FP64 positivity check of rho and p over interior leaves; returns count of bad leaves
 * through n_bad (new; the reference only checks for NaN residuals) */
int32_t fdmamr_check_positivity(fdmamr_ctx* ctx, int32_t* n_bad);

/* ---- regrid (host refine/coarsen with device sensors) ---- */
/* rot_div (solution.F90:424-466,558-580): per-interior-leaf vorticity and density-gradient
 * sensors from the +-1 stencil values on the device; rot/div may be NULL */
int32_t fdmamr_sensors(fdmamr_ctx* ctx, double* rot, double* div);
/* AMR body (solution.F90:413-419): sensors, solute_ref (refine/coarsen by goals,
 * solution.F90:468-535), solute_gridmodify (mesh.F90:305-395), idneighbor, IDgrid, tables,
 * ghost fill. sensor_mode 0: rot OR div (reference); 1: density gradient only. */
int32_t fdmamr_regrid(fdmamr_ctx* ctx, int32_t lvl_amr, double refine_rot, double refine_div,
                      double coarse_rot, double coarse_div, int32_t sensor_mode);
/* The same AMR body with caller-supplied sensors instead of rot_div's: rot/div per interior
 * leaf in cell_NS order, already scaled by dy**1.5 as fdmamr_sensors returns them (the
 * reference's c%rot, c%div read by solute_ref, solution.F90:468-535). A driver that computes
 * its own sensors, or that replays another run's regrid decisions on the same mesh, uses it.
 * On a host-only context (FDMAMR_DEVICE_HOST_ONLY) it regrids the host tree and rebuilds the
 * host tables: the leaves keep the values refine and coarse give them (no ghost refill). */
int32_t fdmamr_regrid_sensors(fdmamr_ctx* ctx, int32_t lvl_amr, double refine_rot, double refine_div,
                              double coarse_rot, double coarse_div, int32_t sensor_mode, const double* rot,
                              const double* div);

/* ---- diagnostics for parity tests ---- */
/* stencil table of interior leaf i: kind[8] (0 copy, 1 parent WLS, 2 parent restriction,
 * 3 child WLS, 4 child fallback copy), canonical node[8]; same classification as the
 * reference's Tip..Tjmm (XuanGua.F90:401-590) */
int32_t fdmamr_stencil_info(fdmamr_ctx* ctx, int32_t i, int32_t* kind8, int32_t* node8,
                            double* px8, double* py8, int32_t* m1_8, int32_t* m2_8,
                            int32_t* nb8x9);
/* device-evaluated stencil values (rho,u,v,p) of interior leaf i, 8 positions */
int32_t fdmamr_stencil_values(fdmamr_ctx* ctx, int32_t i, double* val8x4);
/* FP32 contexts: the values (rho,u,v,p) the fast path's stage reads for interior leaf i, 8
 * positions, after the launch that evaluates the WLS slots and refills the ghosts of the
 * current state (no reference counterpart; diagnostics) */
int32_t fdmamr_stencil_values_fast(fdmamr_ctx* ctx, int32_t i, float* val8x4);
/* regrid-time table facts (no reference counterpart; diagnostics), 11 entries: info8[0] WLS
 * slots that read no ghost leaf, [1] WLS slots, [2] ghost leaves read through the leaf they copy
 * (or the first inflow ghost), [3] items the fast path refills per stage (the other ghosts and
 * the restriction slots), [4] ghost-reading slots without a fixed-size entry list, [5] selector
 * tiles gathered without LDS staging, [6] selector tiles reading more than 64 distinct values,
 * [7] the most distinct values of a staged selector tile, [8] / [9] interior leaves whose x-1/2
 * / y-1/2 face the FP64 stage takes from the neighbour (fdmamr_set_face_share), [10] interior
 * leaves whose 16-byte stencil record cannot hold their ids (read from the 32-byte one) */
int32_t fdmamr_table_info(fdmamr_ctx* ctx, int32_t* info8);
/* the fast path's leaf lists as the last flagged step built them (no reference counterpart;
 * diagnostics): lists[0, *n_clear) the cleared interior leaves, then the troubled ones, each
 * in ascending order (N ints) */
int32_t fdmamr_leaf_lists(fdmamr_ctx* ctx, int32_t* lists, int32_t* n_clear);
/* FP32 contexts: enable = 0 keeps the WLS-slot launch and the ghost refill as separate launches
 * even on meshes where the refill can ride along in the slot launch (no reference counterpart;
 * the default, 1, merges them where possible) */
int32_t fdmamr_set_slot_merge(fdmamr_ctx* ctx, int32_t enable);

/* FP64 contexts: enable = 0 computes both faces of each direction in the update kernel of every
 * leaf, as the reference does; the default, 1, computes a face once where both adjacent leaves
 * read it from the same four state indices (same results bit for bit; no reference
 * counterpart, diagnostics and timing) */
int32_t fdmamr_set_face_share(fdmamr_ctx* ctx, int32_t enable);

/* ---- troubled-cell selector (no reference counterpart: the reference applies the nonlinear
 * WCNS weights everywhere, nnd_center_RK.F90:310-353). FP32 mode only. A leaf is "troubled"
 * when its nonlinear WCNS3 weights deviate from the linear ones; troubled leaves take the full
 * characteristic reconstruction, cleared leaves the linear-weight one. ---- */
/* set / read the per-leaf flags (bit 0 = troubled; bit 2 = forced troubled by a judge
 * rejection until the tables are rebuilt; bit 3 = made by the last regrid, troubled through
 * the selector's next decision), interior leaves in cell_NS order, and switch flag use on
 * (use != 0) or off (every leaf takes the full path). After a table build every leaf is
 * troubled, and the leaves the last regrid made (refined sons and coarsened parents, whose
 * values are copies or means) carry bit 3, as do the leaves whose stencil reads one of them. */
int32_t fdmamr_selector_set_flags(fdmamr_ctx* ctx, const uint8_t* flags, int32_t use);
int32_t fdmamr_selector_get_flags(fdmamr_ctx* ctx, uint8_t* flags);
/* label of every interior leaf on the current state: the largest deviation |w1 - w1_lin|
 * over the leaf's 32 WCNS3 reconstructions (FP32, smoothness floored at 1e-6 of the local
 * mean square so flat data reports 0) */
int32_t fdmamr_labels(fdmamr_ctx* ctx, float* dev);

/* The network: 75 -> 64 -> 64 -> 64 -> 1, GELU (tanh form), sigmoid; BF16 MFMA with FP32 accumulation.
 * Parameters (FP32 master copy, FDMAMR_SELECTOR_NPARAM floats): W1 [80][64] (input rows 75..79
 * are padding and stay zero), b1[64], W2 [64][64], b2, W3 [64][64], b3, w4[64], b4. */
#define FDMAMR_SELECTOR_NPARAM 13569
#define FDMAMR_SELECTOR_NFEAT 80
int32_t fdmamr_selector_init(fdmamr_ctx* ctx, uint64_t seed);                 /* random init */
int32_t fdmamr_selector_set_weights(fdmamr_ctx* ctx, const float* w, int32_t n);
int32_t fdmamr_selector_get_weights(fdmamr_ctx* ctx, float* w, int32_t n);
/* threshold: probability >= threshold is troubled (default 0.5); judge_fraction: share of
 * cleared leaves re-checked per step (default 0.02); label_bound: |w - w_lin| above which a
 * leaf is troubled (default 0.05) */
int32_t fdmamr_selector_config(fdmamr_ctx* ctx, float threshold, float judge_fraction,
                               float label_bound);
/* run the network on the current state and set the flags (and switch flag use on). Leaves
 * whose 5x5 stencil is near-uniform (every channel within 1e-6 of the centre value, relative)
 * are cleared without the network and report probability 0. */
int32_t fdmamr_selector_infer(fdmamr_ctx* ctx, float* prob, int32_t* n_troubled);
/* BF16 feature rows (FDMAMR_SELECTOR_NFEAT per interior leaf) of the current state */
int32_t fdmamr_selector_features(fdmamr_ctx* ctx, uint16_t* x_bf16);
/* append features + labels of every interior leaf of the current state to the device-resident
 * training set; clear it with fdmamr_selector_clear_dataset */
int32_t fdmamr_selector_collect(fdmamr_ctx* ctx, int32_t* n_rows, int32_t* n_positive);
int32_t fdmamr_selector_clear_dataset(fdmamr_ctx* ctx);
/* Adam on the training set: iters minibatches of `batch` rows drawn with `seed`; weighted
 * binary cross-entropy (pos_weight on troubled rows); loss of the last minibatch */
int32_t fdmamr_selector_train(fdmamr_ctx* ctx, int32_t iters, int32_t batch, float lr,
                              float pos_weight, uint64_t seed, float* loss);
/* gradient of the mean weighted loss over n given rows (for gradient checks) */
int32_t fdmamr_selector_grad(fdmamr_ctx* ctx, const uint16_t* x_bf16, const uint8_t* y, int32_t n,
                             float pos_weight, float* grad, float* loss);
/* counters since the last reset: leaves flagged troubled, judged, rejected by the judge */
int32_t fdmamr_selector_stats(fdmamr_ctx* ctx, int32_t* troubled, int32_t* judged,
                              int32_t* rejected, int32_t reset);
/* the judge's sample counter (the step number its hash sample is drawn with; each learned step
 * advances it): *current receives the value, then set_to >= 0 replaces it. A checkpoint keeps
 * it so that a resumed run draws the same samples as an uninterrupted one. */
int32_t fdmamr_selector_judge_counter(fdmamr_ctx* ctx, int64_t set_to, uint32_t* current);

/* ---- measurement (no reference counterpart; the reference times only block refinement,
 * mesh.F90:55,76-77) ---- */
/* bracket the flux + update launches of every stage with HIP events on the context's stream */
int32_t fdmamr_profile(fdmamr_ctx* ctx, int32_t enable);
/* total milliseconds and number of stage-kernel launches recorded since the last read */
int32_t fdmamr_profile_read(fdmamr_ctx* ctx, double* stage_ms_total, int32_t* stage_launches);
/* the same for the selector's inference launches (k_select), with the number of 16-leaf tiles
 * the network ran on since the selector counters were last reset (fdmamr_selector_stats with
 * reset); tiles whose leaves are all near-uniform are cleared without the network */
int32_t fdmamr_profile_read_select(fdmamr_ctx* ctx, double* ms_total, int32_t* launches, int64_t* network_tiles);

#ifdef __cplusplus
}
#endif
#endif
