/*
 * stokes_oracle.h -- CPU restatement of the reference's 3D variable-viscosity
 * Stokes multigrid solve (marker-in-cell staggered grid), used ONLY as test
 * infrastructure: the checker that the HIP library is compared against, and
 * the timed CPU baseline.  Nothing in the product library links this file.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference tree).  The restatement keeps the reference's storage types
 * (FP32 node properties and boundary tables, FP64 solver vectors and matrix
 * rows) so that single-threaded results can be compared bit-for-bit where the
 * reference's arithmetic order is sequential.
 *
 * Layout (parallel/head3mg.c:59-113, parallel/load3mg.c:2011-2320):
 *   level l has mgx[l] x mgy[l] x mgz[l] nodes, node index
 *     n = mgp[l] + m3*mgxy[l] + m1*mgy[l] + m2     (y fastest)
 *   unknown r = 4*n + c, c = 0:P 1:vx 2:vy 3:vz, levels concatenated.
 *   Assembled rows live in a 57-slot-per-node ELL store
 *     pos0[4n+c] = 57n + {0,6,23,40}[c], num0 = entry count.
 */
#ifndef STOKES_ORACLE_H
#define STOKES_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXLVL 12
#define ORC_MAXPOS 64

typedef struct orc_s {
  int L; /* index of the coarsest level (reference: multinum) */
  int mgx[ORC_MAXLVL], mgy[ORC_MAXLVL], mgz[ORC_MAXLVL], mgxy[ORC_MAXLVL];
  long mgn[ORC_MAXLVL], mgp[ORC_MAXLVL], mgs[ORC_MAXLVL];
  double *mggx[ORC_MAXLVL], *mggy[ORC_MAXLVL], *mggz[ORC_MAXLVL];
  int xper, yper, zper;
  double xsize, ysize, zsize;
  double gxk, gyk, gzk; /* GXKOEF, GYKOEF, GZKOEF */
  double p0koef, p1koef, p2koef, v0koef, v1koef, nubeg, nuend;
  int viscmod;
  long ntot; /* nodes over all levels */
  /* node properties (FP32, like the reference) */
  float *nu, *ro, *nd;          /* ntot */
  float *nxx, *nxy, *nxz, *nyz; /* mgn[0] */
  /* boundary equations  x_r = bondv1[b][0] + bondv1[b][1] * x_{bondn1[b]-1} */
  int *bondm; /* 4*ntot, 0 = no boundary equation */
  int nbond, capbond;
  float *bondv1; /* 2*(capbond+1) */
  int *bondn1;   /* capbond+1 */
  /* solver vectors (FP64) */
  double *sol0, *sol1, *fre0, *fre1; /* 4*ntot */
  /* assembled rows */
  double *val1; /* 57*ntot */
  int *lin0;    /* 57*ntot */
  int *num0, *pos0; /* 4*ntot */
  /* coarsest dense LU */
  int nc;
  double *lu;
  int *piv;
  /* velocity-block mode: smoothers skip pressure rows, the coarsest LU
     treats pressure unknowns as fixed (identity rows), prolongation skips P */
  int vonly;
  /* scratch */
  double *tmp;
} orc_t;

/* hierarchy (parallel/load3mg.c:2011-2195) */
orc_t *orc_create(int nx, int ny, int nz, int L, const double *gx, const double *gy,
                  const double *gz, int xper, int yper, int zper);
void orc_destroy(orc_t *o);
void orc_set_vonly(orc_t *o, int flag);
void orc_set_knobs(orc_t *o, double p0, double p1, double p2, double v0, double v1,
                   double nubeg, double nuend, int viscmod, double gxk, double gyk, double gzk);
int orc_level_info(const orc_t *o, int l, int *out6); /* mgx,mgy,mgz,mgxy,mgn,mgp */
double *orc_grid(orc_t *o, int l, int axis);
void *orc_array(orc_t *o, const char *name);

/* boundary-equation grammar projected on every level (parallel/in3mg.c:234-390) */
int orc_bc_line(orc_t *o, int var, int m10, int m11, int m20, int m21, int m30, int m31,
                float cnst, float koef, int dm1, int dm2, int dm3);

/* marker -> node averaging (parallel/mark3mg.c:630-1132) */
int orc_ronurecalc(orc_t *o, long nmark, const float *mx, const float *my, const float *mz,
                   const float *mnu, const float *mro, double nukoef);
/* viscosity/density restriction (parallel/move3mg.c:3387-3528) */
void orc_ronurestrict(orc_t *o, int l);

/* row assembly (parallel/move3mg.c:796-1787 + 2329-2663, parallel/gaus3mg_MG.c:311-365) */
int orc_row(orc_t *o, int l, int m1, int m2, int m3, int c, long *un1, double *ui1);
void orc_assemble(orc_t *o);
int orc_get_row(const orc_t *o, long r, int *idx, double *val, double *rhs);

/* smoother + residual (parallel/move3mg.c:3855-3960) */
void orc_gausmat4(orc_t *o, int am, int l);
void orc_residual(orc_t *o, int l);
/* coloured tile-hybrid sweep of the same row updates (the GPU ordering) */
void orc_eq_box(const orc_t *o, int l, int box[6]);
void orc_sweep_tiles(orc_t *o, int am, int l, int ty, int tx, int tz);
void orc_bc_pass(orc_t *o, int l, double *x);

/* transfers (parallel/move3mg.c:3535-3850, 3966-4354) */
void orc_vprestrict(orc_t *o, int l);
void orc_vpprolong(orc_t *o, int l, int ynreset);

/* coarsest direct solve (replaces parallel/gaus3mg_MG.c:94-249) */
int orc_coarse_factor(orc_t *o);
int orc_coarse_solve(orc_t *o);
int orc_dense_lu(int n, double *a, int *piv);
void orc_dense_lu_solve(int n, const double *lu, const int *piv, double *b);

/* gausmat4 in the multicolour order of the device's sweep over a caller's
 * row store (velocity rows by component and colour, continuity rows,
 * boundary rows by dependency depth); -1 if boundary rows form a cycle */
int orc_gausmat4_colored(orc_t *o, int am, int l);

/* V-cycle (canonical shape of mg/move3mgold.c:2486-2600, fixes for K1/K2) */
void orc_vcycle(orc_t *o, const int *pre, const int *post, int mode, int ty, int tx, int tz);
double orc_relres(orc_t *o, double *blocks3);
/* the same for a finest-level state x, rows built on the fly (no row store) */
double orc_relres_rows(orc_t *o, const double *x, double *blocks3);
int orc_solve(orc_t *o, const int *pre, const int *post, int mode, double tol, int maxcyc,
              double *relres_hist);

/* FP64 GCR with block preconditioner (velocity V-cycle + eta-scaled P) */
void orc_weights(orc_t *o, double *w2);
void orc_hydrostatic(orc_t *o, double *x);
int orc_gcr(orc_t *o, const int *pre, const int *post, int mode, double omega, double tol, int maxit,
            int restart, int hydro, double *x_out, double *hist, double *seconds_per_it);

/* Post-solve pass of vpiterate on level 0 for a solution x (4 per node):
 * stress and strain-rate recompute (parallel/move3mg.c:2908-3046) into
 * sig/eps (6 planes of mgn[0] each: xx yy zz xy xz yz; either may be NULL;
 * zero outside the reference's ranges) and the sums of the reference's
 * convergence norms (parallel/move3mg.c:3090-3238) into out[ORC_POSTN]:
 *   0 sum ival^2 (momentum error)  1 sum eps^2 (error * h^2/eta)  2 sum eps^2 ival^2
 *   3 momentum rows  4 sum c^2 (div v * h) over cells with a continuity row
 *   5 those cells  6 sum c^2 over cells with a pressure boundary equation
 *   7 those cells  8 sum of squared velocity boundary-equation errors  9 those rows
 *   10..17 min/max of vx, vy, vz, P (move3mg.c:2858-2900)  18/19 min/max
 *   of the cell viscosity of the normal stresses (move3mg.c:2926-2938). */
#define ORC_POSTN 20
void orc_post(const orc_t *o, const double *x, double *sig, double *eps, double *out);

/* Marker advection (movemark, parallel/mark3mg.c:174-629) through the
 * interpolations vxyzcalc1 (:2963) and epscalc1 (:3244), walled and periodic
 * boxes (periodic axes wrap the RK stage points and the moved markers;
 * returns the number of wraps that left a point outside, where the reference
 * exits, 0 normally).
 * x: level-0 solution (4 per node); eps: the six strain-rate planes of
 * orc_post.  Velocities and strain rates are read as the reference's FP32
 * arrays hold them.  marke (accumulated strain), markk (set to 0 for a
 * normal marker leaving the box), markt (types: < 50 normal, >= 100
 * immobile), eii (strain-rate invariant at the start position) may be NULL.
 * markmod 1: one Euler step, else RK4; outgrid as the reference's knob. */
int orc_move_markers(const orc_t *o, const double *x, const double *eps, long nmark, float *mx, float *my,
                     float *mz, float *marke, float *markk, const int *markt, float *eii, double timestep,
                     int markmod, int outgrid);
/* the two interpolations alone: v[3] at (px,py,pz); e[6] the strain rates there */
void orc_interp(const orc_t *o, const double *x, const double *eps, double px, double py, double pz, double *v,
                double *e);

/* banded direct solver restatement (parallel/gausjomp.c:70-254) */
int orc_band_solve(orc_t *o, double *x, long *band_out);
/* the same with the reference's thread structure (rows round-robin, each
   waiting for the rows it eliminates with); bit-identical to orc_band_solve */
int orc_band_solve_mt(orc_t *o, double *x, long *band_out, int threads);

int orc_threads(void);

/* ---- heat-equation multigrid (parallel/heat3mg.c), heat_oracle.c ---------- */
typedef struct orh_s orh_t;
/* heat solve on levels 0..Lt of o's hierarchy (Lt: the reference's multinumt) */
orh_t *orh_create(orc_t *o, int Lt);
void orh_destroy(orh_t *h);
void orh_set(orh_t *h, double timestep, double t0koef, double t1koef);
void *orh_array(orh_t *h, const char *name);
int orh_nbond(const orh_t *h);
/* T boundary line of the reference grammar (parallel/in3mg.c:302-385, var T) */
int orh_bc_line(orh_t *h, int m10, int m11, int m20, int m21, int m30, int m31, float cnst, float koef, int dm1,
                int dm2, int dm3);
void orh_rocprestrict(orh_t *h, int l);                                  /* heat3mg.c:1235-1330 */
int orh_row(orh_t *h, int l, int m1, int m2, int m3, long *un1, double *ui1); /* heaterr / tbonderr */
void orh_assemble(orh_t *h);
void orh_residual(orh_t *h, int l);
void orh_gausmatt(orh_t *h, int am, int l);                              /* heat3mg.c:1057-1120 */
int orh_gausmatt_colored(orh_t *h, int am, int l);
void orh_trestrict(orh_t *h, int l);                                     /* heat3mg.c:1122-1232 */
void orh_tprolong(orh_t *h, int l, int ynreset);                         /* heat3mg.c:954-1050 */
void orh_err(orh_t *h, double *heatsum, double *bondsum);                /* heat3mg.c:528-565 */
int orh_step(orh_t *h, const int *pre, const int *post, int multicyct, double heatmin, int maxit, int colored,
             double *heatsum, double *bondsum);                          /* titerate, heat3mg.c:384-590 */
/* the right part of heaterr (ynerr = 0, heat3mg.c:54-187) into tk1 from a
   Stokes solution x (4 per node), its stress and strain-rate planes (6 n0
   each) and FP32 ht, et; gxk..gzk the gravity (GXKOEF..GZKOEF) */
void orh_sources(orh_t *h, const double *x, const double *sig, const double *eps, const float *ht, const float *et,
                 int adiabyn, int frictyn, double gxk, double gyk, double gzk);
/* marker temperatures after the step (heat3mg.c:618-950), markers in index
   order; tk2 (array "tk2") is the temperature at the start of the step */
void orh_markers(orh_t *h, long nmark, const float *mx, const float *my, const float *mz, const int *markt,
                 float *markk, double heatdif, double timestep, double zdeep, double tdeep);
/* the whole thermal step with the sub-step limit (heat3mg.c:313-616); out[8] */
int orh_titerate(orh_t *h, double timestep, double maxtkstep, const int *pre, const int *post, int multicyct,
                 double heatmin, int maxit, int colored, double *out);

#ifdef __cplusplus
}
#endif
#endif
