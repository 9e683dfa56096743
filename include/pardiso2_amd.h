/*
 * pardiso2_amd.h -- C ABI of the MI355X (gfx950) variable-viscosity Stokes
 * multigrid library.
 *
 * Two surfaces:
 *
 *  1. The reference's own entry points (same names, same C signatures, same
 *     argument meaning), acting on the context selected with stk_bind().
 *     Each declaration cites the reference interface it replaces.  Where the
 *     reference reads global arrays (sol0, fre0, fre1, nd, bondm, ...) the
 *     library uses the bound context's copies of them, laid out exactly like
 *     the reference's (node-major, y fastest, 4 unknowns P,vx,vy,vz per node,
 *     levels concatenated at mgp[l]).  Accessors stk_vec_set/get move those
 *     arrays between host and device.
 *
 *  2. A context API (stk_*) for the matrix-free GPU solve: build the level
 *     hierarchy, load boundary equations and node properties, prepare the
 *     coarse operators, and solve to a relative residual.
 *
 * All pointers are host pointers unless a name says _device.  Functions
 * return 0 on success and a negative STK_E* code on failure; nothing calls
 * exit() (the reference's fatal paths, e.g. parallel/gaus3mg_MG.c:173-204,
 * become status codes).
 */
#ifndef PARDISO2_AMD_H
#define PARDISO2_AMD_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define STK_OK 0
#define STK_EINVAL -1      /* invalid argument or unsupported configuration */
#define STK_ENOMEM -2      /* device allocation failed */
#define STK_EHIP -3        /* HIP runtime error */
#define STK_ESINGULAR -4   /* coarsest-level matrix singular */
#define STK_ENOCONV -5     /* solve did not reach the tolerance */
#define STK_ENAN -6        /* NaN/Inf in the iteration */
#define STK_ESTATE -7      /* call order violated (e.g. solve before prepare) */
#define STK_ENOCTX -8      /* reference-named entry point without stk_bind() */
#define STK_EBCFORM -9     /* an unknown of the outer layer has no boundary equation: the
                              reference's walls-on-the-outer-lines layout (its free-slip rows
                              would read nodes outside the grid); give the outer layer
                              boundary equations (walls on lines 1 and n-2, ghost copies) */

typedef struct stk_ctx stk_ctx;

/* Solver knobs with the reference's meaning (parallel/mode.t3c:84-96,
 * parallel/load3mg.c:116-131).  gx/gy/gz are GXKOEF/GYKOEF/GZKOEF.
 * divvmin/stoksmin/divvmax/stoksmax are DIVVMIN/STOKSMIN/DIVVMAX/STOKSMAX
 * (parallel/mode.t3c:79-82, parallel/load3mg.c:112-115): the thresholds of
 * the reference's stop test, reported in stk_norms.converged; divvmax > 0
 * replaces the velocity scale maxvxyz of the norms.  Zero-initialised
 * trailing fields (a caller that sets only the first eleven) report
 * converged = 0 unless every norm is zero. */
typedef struct {
  double p0koef, p1koef, p2koef, v0koef, v1koef;
  double nubeg, nuend;
  int viscmod;
  double gx, gy, gz;
  double divvmin, stoksmin, divvmax, stoksmax;
} stk_knobs;

typedef struct {
  int precision;     /* 0: FP32 V-cycle inside FP64 outer iteration; 1: FP64 throughout */
  int restart;       /* Krylov directions kept before a restart (1..256) */
  int pre[12];       /* pre-smoothing sweeps per level */
  int post[12];      /* post-smoothing sweeps per level */
  int tile[3];       /* smoother tile (y, x, z) in nodes; 0 0 0 = per level: autotuned, else 32 4 4 */
  double omega;      /* pressure block scale of the preconditioner (schur = 0) */
  int cycle;         /* multigrid cycle of the velocity block: 0 or 1, a V-cycle (other values: STK_EINVAL) */
  int coarse_viscmod; /* averaging of coarse-level viscosities: -1 = knobs.viscmod, 0/1/2 */
  double tol;        /* target relative residual ||W r|| / ||W b|| */
  int maxit;         /* outer iteration cap */
  int verbose;
  int schur;         /* pressure Schur approximation: 0 omega*eta_cell scaling,
                        1 weighted BFBt with one aggregation-multigrid V-cycle per K solve */
  int ksweeps;       /* red/black Gauss-Seidel sweeps before and after each K-cycle level */
  double kscale;     /* coarse-correction scale of the aggregation K-cycle */
  int krylov;        /* outer Krylov method: 0 GCR, 1 flexible GMRES */
  int graph;         /* 1: replay the preconditioner from a captured HIP graph (captured on the
                        first solve of a shape; captures, and the library's device allocations and
                        releases, are serialised across the process's threads), 0: eager */
  int history;       /* 0: the context keeps its Krylov vectors between solves; 1: it borrows them
                        from a process-wide pool for the solve (many contexts, few solving at once:
                        memory follows the solves in flight; a solve waits while the device has no
                        room for another set) */
  int basis;         /* precision 0 only: 1 stores the flexible-GMRES basis and preconditioned vectors
                        in FP32 (the orthogonalisation and the update of x run in FP64 on the stored
                        values; the true residual is FP64), 0 stores them in FP64 */
} stk_solver_opts;

/* The reference's own convergence norms of a finest-level solution
 * (parallel/move3mg.c:3061-3086, 3090-3238), FP64:
 *   maxvxyz   RMS of the vx, vy, vz ranges (1e-8 if zero; DIVVMAX if > 0)
 *   stoksum   RMS momentum error (divergence of the recomputed stresses
 *             minus pressure gradient minus gravity) over momentum rows
 *   stoksum1  RMS of that error times h^2/eta, over maxvxyz
 *   stoksum2  RMS of the product of the two, over maxvxyz
 *   contsum   RMS of div(v) times the mean cell size over maxvxyz, cells
 *             with a continuity equation; contsumb: cells with a pressure
 *             boundary equation
 *   bondsum   RMS error of the velocity boundary equations
 * with the counts of rows and cells, the extrema the reference prints
 * (velocities, pressure, cell viscosity), and converged = the reference's
 * stop test (parallel/move3mg.c:3282): stoksum2 <= STOKSMIN*STOKSMAX,
 * stoksum*stoksum1 <= STOKSMIN*STOKSMAX and contsum <= DIVVMIN. */
typedef struct {
  double stoksum, stoksum1, stoksum2, contsum, contsumb, bondsum;
  double stoknum, contnum, contnumb, bondnum;
  double maxvxyz;
  double minvx, maxvx, minvy, maxvy, minvz, maxvz, minpr, maxpr, minnu, maxnu;
  int converged;
} stk_norms;

typedef struct {
  int iterations;
  int status;
  double relres;     /* final relative residual, recomputed from b - A x in FP64 */
  double relres0;
  double seconds;    /* wall time of the solve on the device stream */
  double blocks[3];  /* ||r_momentum||, ||W r_continuity||, ||r_boundary|| */
  double hist[256];  /* relative residual after each outer iteration */
  stk_norms norms;   /* the reference's norms of the final solution */
} stk_stats;

/* ---- context API ------------------------------------------------------ */

/* Level hierarchy from the fine grid lines (restates gridcheck,
 * parallel/load3mg.c:2011-2195): mg[l] = (mg[l-1]-5)/2+5, coarse lines keep
 * the three outer lines per side.  levels = coarsest level index (multinum).
 * Grid lines must increase strictly (STK_EINVAL otherwise).  A periodic axis
 * (the reference's xperiodic/yperiodic/zperiodic) identifies its last grid
 * line with line 0: the stencils wrap as the reference's periodic branches
 * do (parallel/move3mg.c:27-34, 400-429, 834-847, 1481-1483), the normal
 * velocity on the last line needs the copy equation of a periodic boundary
 * line (shift magnitude > 1, parallel/in3mg.c:339-341), and marker averaging
 * folds the last line into line 0 (parallel/mark3mg.c:893-955, 1032-1111). */
int stk_create(stk_ctx **ctx, int nx, int ny, int nz, int levels, const double *gx,
               const double *gy, const double *gz, int xperiodic, int yperiodic, int zperiodic);
void stk_destroy(stk_ctx *ctx);
/* mgx, mgy, mgz, mgxy, mgn, mgp of level l */
int stk_level_info(const stk_ctx *ctx, int level, int out6[6]);
int stk_set_knobs(stk_ctx *ctx, const stk_knobs *k);
void stk_default_opts(stk_solver_opts *o, int levels);

/* One boundary-equation line of the reference grammar, projected on every
 * level (restates parallel/in3mg.c:234-390).  var: 0 P, 1 vx, 2 vy, 3 vz,
 * 4 T (the temperature rows of the heat solve, in3mg.c:369-385). */
int stk_bc_line(stk_ctx *ctx, int var, int m10, int m11, int m20, int m21, int m30, int m31,
                float cnst, float koef, int dm1, int dm2, int dm3);
/* Boundary tables in the reference's own layout (parallel/head3mg.c:110-112):
 * bondm[4*ntot] (0 = none), bondv1[2*(nbond+1)], bondn1[nbond+1] (1-based). */
int stk_set_bc_tables(stk_ctx *ctx, const int *bondm, int nbond, const float *bondv1,
                      const int *bondn1);
int stk_get_bc_tables(const stk_ctx *ctx, int *bondm, int *nbond, float *bondv1, int *bondn1);

/* Finest-level node properties (parallel/head3mg.c:94-96): nu, ro at nodes,
 * nxx at cell centres (upper-corner node index), nxy/nxz/nyz on edges. */
int stk_set_fields(stk_ctx *ctx, const float *nu, const float *ro, const float *nxx,
                   const float *nxy, const float *nxz, const float *nyz);
/* Marker -> node averaging on the device (restates ronurecalc,
 * parallel/mark3mg.c:630-1132, for per-marker viscosity and density). */
int stk_markers_to_nodes(stk_ctx *ctx, long nmark, const float *mx, const float *my,
                         const float *mz, const float *mnu, const float *mro, double nukoef);
/* The same averaging, with the same results, without allocating: every device
 * buffer of the call (staged and sorted marker fields, cell keys, cell starts,
 * sort storage) lives in one static device array per GPU, about 73 MB, held
 * until the process ends.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  Limits: nmark <= 2^20 markers and
 * (nx-1)(ny-1)(nz-1) <= 2^18 cells, else STK_EINVAL.  Calls on one GPU share
 * its array: calls that overlap in time on the same GPU race. */
int stk_markers_to_nodes_noalloc(stk_ctx *ctx, long nmark, const float *mx, const float *my,
                                 const float *mz, const float *mnu, const float *mro, double nukoef);
/* node arrays over all levels: nu, ro (ntot), nd (ntot), level-0 nxx..nyz */
int stk_get_fields(stk_ctx *ctx, float *nu, float *ro, float *nd, float *nxx, float *nxy,
                   float *nxz, float *nyz);

/* Restrict nu/ro to all levels, build stencil viscosities, right-hand side
 * and the coarsest-level inverse.  Must follow any field/BC/knob change. */
int stk_prepare(stk_ctx *ctx);

/* Finest-level solution in the reference's sol0 layout (4*mgn[0] doubles). */
int stk_set_solution(stk_ctx *ctx, const double *sol0_level0);
int stk_get_solution(stk_ctx *ctx, double *sol0_level0);
/* Solve A x = b on the finest level from the current solution. */
int stk_solve(stk_ctx *ctx, const stk_solver_opts *opts, stk_stats *stats);
/* Relative residual of the current finest solution (FP64, device). */
int stk_relres(stk_ctx *ctx, double *relres, double blocks[3]);
/* The reference's convergence norms of the current solution (the post-solve
 * pass of vpiterate, parallel/move3mg.c:2835-3238; every stk_solve reports
 * them for its result in stk_stats.norms), periodic copies included. */
int stk_norms_eval(stk_ctx *ctx, stk_norms *out);
/* Stress and strain-rate recompute of the current solution
 * (parallel/move3mg.c:2908-3046): sig and eps get 6 planes of mgn[0]
 * doubles each, in the order xx yy zz xy xz yz, node layout of level 0
 * (sxx..szz at the cell whose upper corner is the node, shear stresses at
 * the node's edge); zero where the reference's recompute writes nothing.
 * On a periodic axis the shear stresses of the last line are line 0's and
 * the normal ones of line 0 the last cell's (move3mg.c:2933-3044,
 * 3321-3379).  Either pointer may be NULL. */
int stk_stress_strain(stk_ctx *ctx, double *sig, double *eps);
/* One marker-advection step through the current solution: movemark
 * (parallel/mark3mg.c:174-629) with the velocity and strain-rate
 * interpolations vxyzcalc1 and epscalc1 (parallel/mark3mg.c:2963, 3244),
 * walled boxes.  Velocities and strain rates are read as the reference's
 * FP32 node arrays hold them.  mx/my/mz (marker coordinates) are updated in
 * place; markmod 1: one Euler step, else RK4 (the knob markmod); outgrid as
 * the reference's knob (1: markers outside the box stay, 2: they keep only
 * their motion across the faces they lie beyond).  Optional (NULL):
 * marke, the accumulated strain, grows by the RK4 average of the strain-rate
 * invariant times timestep where positive; markk is set to 0 for a normal
 * marker that leaves the box; markt, the marker types (< 50 normal, >= 100
 * immobile; NULL: all normal); eii receives the strain-rate invariant at the
 * start position (the reference's val1) of every marker the step moves (the
 * others keep their entries).  The water-marker branch and the
 * immobile-marker reset, both disabled in the reference, are not built.
 * Periodic axes (mark3mg.c:337-485, 2992-3760): the interpolations pair the
 * centres or cells across the wrap as the reference does, and the RK stage
 * points and the moved markers wrap by one period; STK_EINVAL after the step
 * if a point is still outside then (moved more than a period: the reference
 * exits). */
int stk_move_markers(stk_ctx *ctx, long nmark, float *mx, float *my, float *mz, float *marke, float *markk,
                     const int *markt, float *eii, double timestep, int markmod, int outgrid);
/* Hydrostatic initial guess: v = 0, P integrated up each column through the
 * z-momentum rows (the reference seeds P this way, parallel/in3mg.c:927-938). */
int stk_hydrostatic_guess(stk_ctx *ctx);
/* Average device time (ms) of one launch of a kernel of level `level`, HIP
 * events on the library stream.  which: 0 FP32 smoother sweep as the V-cycle
 * runs it there (the level's tuned tile, boundary rows set in the tiles where
 * the level allows), 1 FP64 smoother sweep (32x4x4 tile), 2 FP64 operator
 * apply, 3 FP32 residual, 4 FP64 residual, 5 plain copy of 4*mgn[l] doubles
 * (32 B read + 32 B written per node). */
int stk_time_kernel(stk_ctx *ctx, int which, int level, int reps, double *ms_avg);
/* Shift P so that P(px,py,pz) = pinit (parallel/move3mg.c:2835-2856). */
int stk_pressure_gauge(stk_ctx *ctx, int px, int py, int pz, double pinit);

/* ---- temperature step (parallel/heat3mg.c) ---------------------------- */
/* The implicit heat-equation step of titerate on levels 0..levels of the
 * context's hierarchy: 7-point rows of heaterr (heat3mg.c:195-249, rho cp
 * clamped below at 2e6, face conductivities the two-node means, x wraps on
 * an x-periodic box), boundary rows of the var 4 lines, Gauss-Seidel in two
 * colours and boundary passes, residual restriction and correction
 * prolongation (trestrict, tprolong), property restriction (rocprestrict).
 * Every node without a T boundary row needs its six neighbours on the grid
 * (STK_EINVAL otherwise), i.e. T rows on every face except x line 0 of an
 * x-periodic box. */
typedef struct {
  int levels;          /* multinumt: coarse levels under the finest (0: Gauss-Seidel alone) */
  int cycles;          /* multicyct: V-cycles per iteration */
  int pre[12];         /* multittt[0][l]: sweeps on level l going down (level 0: every iteration) */
  int post[12];        /* multittt[1][l]: sweeps on level l >= 1 going up */
  double heatmin;      /* HEATMIN: iterate until the RMS heat-equation error (K) is at most this */
  double t0koef;       /* relaxation of a sweep's update */
  double t1koef;       /* weight of the prolonged correction */
  double maxtkstep;    /* largest temperature change (K) a sub-step may take from the source term */
  int maxit;           /* iterations per sub-step before STK_ENOCONV (the reference has no cap) */
} stk_heat_opts;

typedef struct {
  int substeps;        /* sub-steps taken (0: timestep or the sub-step limit was 0) */
  int iterations;      /* iterations over all sub-steps */
  int converged;       /* 1 if every sub-step reached heatmin */
  double heatsum, heatnum;  /* RMS heat-equation error of the last iteration, its node count */
  double bondsum, bondnum;  /* RMS boundary-row error, its row count */
  double tkstep;       /* sub-step limit maxtkstep / largest rate (1e30 without a source) */
  double tkmin, tkmax; /* temperature range after the step */
  double seconds;      /* wall time of the step */
} stk_heat_stats;

/* The reference's shipped values: levels 0, cycles 1, pre {10}, post {0},
 * heatmin 1e-4, t0koef = t1koef = 1, maxtkstep 25 K, maxit 2500 (cyc2max). */
void stk_heat_default_opts(stk_heat_opts *o);
/* Level-0 properties (FP32, mgn[0] each): density ro, heat capacity cp,
 * conductivity kt, temperature tk, source term ht (tk1: W/m^3, the right part
 * of heaterr, heat3mg.c:54-62, 183). */
int stk_heat_set_fields(stk_ctx *ctx, const float *ro, const float *cp, const float *kt, const float *tk,
                        const float *ht);
/* T boundary equations from the reference's own tables (parallel/head3mg.c:
 * 110-112, in3mg.c:369-385): tbondm = bondm + nodenum8, one entry per node of
 * every level (0 = none), bondv1[2*(nbond+1)] and bondn1[nbond+1] (partner
 * node + 1, 0 = none) shared with the Stokes rows.  Replaces the T rows of
 * earlier stk_bc_line(var 4) calls. */
int stk_heat_set_bc_tables(stk_ctx *ctx, const int *tbondm, int nbond, const float *bondv1,
                           const int *bondn1);
/* Marker temperatures after the step (titerate, parallel/heat3mg.c:618-950),
 * from the step's new and starting node temperatures: fluid markers (type 50..99)
 * inside take the new temperature; normal markers (type < 50) strictly inside
 * with markk <= 0 take the old one (tdeep where y > zdeep and colder); with
 * heatdif, the subgrid numerical diffusion toward the nodes (the exponent
 * -heatdif k/(rho cp) 2 (1/dx^2+1/dy^2+1/dz^2) timestep clamped at -150, the
 * boundary cells capped at the nodes' largest temperature) scattered to the
 * nodes and taken back from every marker (antidiffusion); then the nodal
 * change.  markt may be NULL (all normal); markk (FP32) is updated in place.
 * The node sums run over the markers of each node's 8 cells in cell order,
 * markers of a cell in input order: the reference's single-thread order for
 * markers stored cell by cell.  exp is the device's (ULPs from the host's). */
int stk_heat_markers(stk_ctx *ctx, long nmark, const float *mx, const float *my, const float *mz,
                     const int *markt, float *markk, double heatdif, double timestep, double zdeep, double tdeep);
/* ro, cp, kt of every level (ntot each, restricted by the last step or
 * stk_heat_prepare); tk, tk0 and the source term tk1 of level 0.  Any
 * pointer may be NULL. */
int stk_heat_get_fields(stk_ctx *ctx, float *ro, float *cp, float *kt, float *tk, float *tk0, float *tk1);
/* The right part of heaterr (ynerr = 0, parallel/heat3mg.c:54-187) from the
 * context's current Stokes solution: tk1 = ht (stk_heat_set_fields) plus the
 * adiabatic term (adiabyn 2: et tk v.grad P, the cells' pressure gradient at
 * the node; adiabyn 1: et tk ro v.g with the knobs' gravity) and the
 * frictional heating (frictyn 1: sigma:epsilon from the recomputed stresses
 * and strain rates), both where ro > 2e3, with the velocity vxyzcalc1
 * interpolates at the node.  et: the thermal expansion (mgn[0] floats; may be
 * NULL when adiabyn is 0).  Call it after the solve and before
 * stk_heat_titerate, as titerate evaluates it at the start of the step. */
int stk_heat_sources(stk_ctx *ctx, const float *et, int adiabyn, int frictyn);
/* The whole temperature step (titerate, parallel/heat3mg.c:292-616): the
 * sub-step limit from the largest heating rate (331-382), rocprestrict on
 * levels 1..levels, then sub-steps, each from tk0 = tk: groups of `cycles`
 * V-cycles, tk = sol0 and the errors after each group, until heatsum <=
 * heatmin.  STK_ENOCONV if a sub-step reached maxit first (the step still
 * runs to its end).  st may be NULL. */
int stk_heat_titerate(stk_ctx *ctx, double timestep, const stk_heat_opts *o, stk_heat_stats *st);
/* The pieces, for inspection.  prepare: rocprestrict on levels 1..levels,
 * then the start of one sub-step of `timestep` (tk0 = tk, sol0 = tk, the
 * level-0 right-hand side fre1).  gausmatt: `sweeps` sweeps on level l, then
 * its residual fre0 (heat3mg.c:1057-1120).  trestrict: level l-1's fre0 into
 * level l's fre1, sol0 = 0 there (1122-1232).  tprolong: level l's sol0 into
 * level l-1 (954-1050).  err: tk = (float) sol0, then the RMS errors
 * (heat3mg.c:507-553). */
int stk_heat_prepare(stk_ctx *ctx, int levels, double timestep, double t0koef, double t1koef);
int stk_heat_gausmatt(stk_ctx *ctx, int level, int sweeps);
int stk_heat_trestrict(stk_ctx *ctx, int level);
int stk_heat_tprolong(stk_ctx *ctx, int level, int ynreset);
int stk_heat_err(stk_ctx *ctx, double *heatsum, double *bondsum);
/* sol0 / fre1 / fre0 (STK_SOL0, STK_FRE1, STK_FRE0) of the heat solve on
 * level l (mgn[l] doubles): dir 0 device to host, 1 host to device. */
int stk_heat_vec(stk_ctx *ctx, int which, int level, double *host, int dir);
/* Average device time (ms) of one launch on level `level` after
 * stk_heat_prepare or a step, HIP events on the library stream: which 0 one
 * colour of a sweep (k_heat_gs over the whole level), 1 a whole sweep (both
 * colours and the boundary passes), 2 the residual, 3 a plain copy of
 * mgn[level] doubles (8 B read + 8 B written per node). */
int stk_heat_time(stk_ctx *ctx, int which, int level, int reps, double *ms_avg);

/* ---- reference arrays on the context (host <-> device) ---------------- */
#define STK_SOL0 0
#define STK_FRE1 1
#define STK_FRE0 2
int stk_vec_set(stk_ctx *ctx, int which, int level, const double *host);
int stk_vec_get(stk_ctx *ctx, int which, int level, double *host);
/* y = A x on a level (FP64, matrix-free), x/y host arrays of 4*mgn[l] */
int stk_apply(stk_ctx *ctx, int level, const double *x, double *y);
/* Smoother sweeps in tile-coloured order on the context's sol0/fre1 of one
 * level, FP64, all rows (pressure Uzawa step included), then fre0.
 * ty = tx = tz = 0 selects one tile covering the whole level (the LDS-resident
 * kernel the V-cycle uses on levels of at most 2048 nodes in FP64); tz < 0
 * selects the tile (ty, tx, -tz) with its viscosities staged in LDS. */
int stk_smooth(stk_ctx *ctx, int level, int sweeps, int ty, int tx, int tz);
/* The same in FP32 arithmetic (precision 0: the mixed-precision V-cycle's
 * smoother, on FP32 roundings of sol0/fre1; sol0 is widened back and fre0 is
 * the FP64 residual) or FP64 (precision 1), on all rows or (velocity_only)
 * on the velocity rows as the V-cycle of the velocity block does.
 * stk_smooth = stk_smooth_ex(..., 1, 0).  Same reference routine: gausmat4,
 * parallel/move3mg.c:3855-3960. */
int stk_smooth_ex(stk_ctx *ctx, int level, int sweeps, int ty, int tx, int tz, int precision, int velocity_only);
int stk_coarse_solve(stk_ctx *ctx);

/* ---- smoother tile autotuning ------------------------------------------ */
/* Time the tile-coloured smoother of every level that is not LDS-resident on
 * a list of tile shapes, each with and without the stencil viscosities staged
 * in LDS (HIP events, reps launches each), and keep the fastest per level for
 * solves whose opts->tile is 0 0 0.  precision 0 FP32, 1 FP64. */
int stk_autotune(stk_ctx *ctx, int precision, int reps);
/* tile (y, x, z, viscosities staged) a solve uses on a level */
int stk_level_tile(const stk_ctx *ctx, int level, int precision, int out4[4]);
/* set that tile by hand (instead of or after stk_autotune); STK_EINVAL if
 * its LDS would exceed 64 KB.  No counterpart in the reference: a tuning hook. */
int stk_set_level_tile(stk_ctx *ctx, int level, int precision, const int tile4[4]);
/* 1 when the tile sweeps of shape ty x tx x tz on a level set the level's
 * boundary rows themselves (each tile evaluates the chains it owns in
 * composed form), 0 when a separate chain pass follows every sweep (no
 * boundary rows, or a chain whose row and partner fall in different tiles).
 * Builds the tile lists if they do not exist yet.  No counterpart in the
 * reference: an inspection hook for the tests, like stk_level_tile. */
int stk_tile_chains(stk_ctx *ctx, int level, int ty, int tx, int tz);

/* ---- many independent instances --------------------------------------- */
/* Solve n contexts concurrently: `threads` host workers (<= 0: 16) take the
 * contexts in order, each context on its own stream, so the device overlaps
 * the instances.  mode 0: solve from the current solution; mode 1: prepare,
 * hydrostatic guess, then solve.  opts holds 1 (shared) or n entries, stats n
 * entries (or NULL).  Returns STK_OK, else the first hard error, else
 * STK_ENOCONV if some instance did not converge.  Results equal those of
 * handling each context alone. */
int stk_solve_batch(int n, stk_ctx *const *ctxs, const stk_solver_opts *opts, int nopts, stk_stats *stats,
                    int threads, int mode);
/* free and total device memory (admission control of a request loop) */
int stk_mem_info(size_t *free_bytes, size_t *total_bytes);

/* ---- Schur approximation, for inspection ------------------------------ */
/* y = K_l x with K_l the pressure-cell operator of level l (level 0:
 * B D^-1 G, coarser: aggregated Galerkin), cell vectors of mgn[l] doubles */
int stk_kop_apply(stk_ctx *ctx, int level, const double *x, double *y);
/* x = one K-cycle applied to b (level-0 cell vectors), FP32 (precision 0) or FP64 */
int stk_kcycle_apply(stk_ctx *ctx, int precision, int ksweeps, double kscale, const double *b, double *x);
/* pressure slots of z = the preconditioner's Schur block applied to the
 * pressure slots of r (opts->schur 1 or 2), before the boundary-row
 * substitution; velocity slots of z are 0 */
int stk_schur_apply(stk_ctx *ctx, const stk_solver_opts *opts, const double *r, double *z);

/* ---- reference entry points (act on the bound context) ---------------- */
int stk_bind(stk_ctx *ctx);

/* The caller's copies of the reference's global arrays (declared at
 * parallel/head3mg.c:59-62, 94-112, 196), in the reference's layout and
 * types: solver vectors and node properties over all levels (node index
 * mgp[l] + m3*mgxy[l] + m1*mgy[l] + m2, 4 unknowns P,vx,vy,vz per node),
 * the row store of gausmat2, the boundary tables (bondv1 holds 2 floats per
 * equation, equation 0 unused) and the count of boundary equations.  The
 * level-0 outputs of vpiterate's post-solve pass may be NULL. */
typedef struct {
  double *sol0, *fre0, *fre1, *val1;
  int *lin0, *num0, *pos0;
  float *nu, *ro, *nd;
  float *nxx, *nxy, *nxz, *nyz;
  int *bondm;
  float *bondv1;
  int *bondn1;
  long *bondnum;
  double *pr;
  float *vx, *vy, *vz;
  float *sxx, *syy, *szz, *sxy, *sxz, *syz;
  float *exx, *eyy, *ezz, *exy, *exz, *eyz;
} stk_ref_arrays;
/* Bind the caller's arrays (NULL unbinds; the struct is copied, the arrays
 * are not).  With arrays bound the reference-named entry points below read
 * and write them in place, as the reference's own routines do its globals:
 *   gausmat2   stores a row into val1/lin0/num0/pos0 and fre1 (host only);
 *   gausmat4   sweeps level mgi of that row store on the device (multicolour
 *              order: vx, vy, vz rows red then black, continuity rows, then
 *              boundary rows by dependency) and writes sol0 and fre0 back;
 *   vprestrict, vpprolong, ronurestrict, pardisomat(0)
 *              move the levels they read and write between the arrays and
 *              the device around the device kernel;
 *   vpiterate  loads nu/ro/nxx..nyz of level 0 and the boundary tables,
 *              prepares, solves from sol0 and writes sol0, pr, vx, vy, vz
 *              and the stress and strain-rate planes of level 0 back.
 * gausmat4 .. vpiterate also need a prepared context selected by stk_bind,
 * built on the same grid, knobs and boundary equations. */
int stk_bind_arrays(const stk_ref_arrays *arrays);

/* parallel/gaus3mg_MG.c:2 -- clear an rm x (rm+1) dense matrix */
void matclear(int rm, double *adr);
/* parallel/gaus3mg_MG.c:21 -- dense Gauss elimination, no pivoting */
void gausmat(int rm, double *adr, double *adr1);
/* parallel/gaus3mg_MG.c:311 -- store one assembled row (un1/ui1 format) in
 * the bound arrays (STK_ENOCTX without them) */
int gausmat2(int am, long mcmax, long pos0cur1, long *un1, double *ui1);
/* parallel/gaus3mg_MG.c:64 -- am=1: set one coarsest-level row from its
 * list (duplicate columns add up; a later call for the row replaces it);
 * am=0: solve the coarsest system (device dense inverse) and store the
 * solution at sol0[4*(mgp[L]+node)+c] of the bound context. */
int pardisomat(int am, long mcmax, long mcmax2, long pos0cur1, int tid, long *un1, double *ui1);
/* parallel/move3mg.c:3855 (prototype parallel/head3mg.c:350-377) -- am
 * smoother sweeps on level mgi, then fre0 = fre1 - A sol0 on [mcmin,mcmax]:
 * over the bound arrays' row store if arrays are bound, else over the bound
 * context's matrix-free rows and its own sol0/fre1/fre0 copies */
int gausmat4(int am, long mcmax, long mcmin, int mgi);
/* parallel/move3mg.c:3966 -- residual restriction to level mgi */
void vprestrict(int mgi);
/* The mixed-precision V-cycle's restriction of the velocity rows (FP32,
 * pressure rows 0) on the context's copies: fre0 of level l-1 (rounded to
 * FP32) -> fre1 of level l (widened).  Same reference routine as vprestrict
 * (parallel/move3mg.c:3966-4354), summed axis by axis. */
int stk_vrestrict32(stk_ctx *ctx, int level);
/* parallel/move3mg.c:3535 -- correction prolongation from level mgi */
void vpprolong(int mgi, int ynreset);
/* parallel/move3mg.c:3387 -- nu/ro restriction to level mgi */
void ronurestrict(int mgi);
/* parallel/move3mg.c:1791 -- full Stokes solve of the bound context */
void vpiterate(int m0, int f0);

#ifdef __cplusplus
}
#endif
#endif
