/*
 * heat_oracle.c -- CPU restatement of the reference's heat-equation
 * multigrid (parallel/heat3mg.c), test infrastructure only: the checker the
 * HIP library's heat solve is compared against.  Nothing in the product
 * library links this file.
 *
 * The heat solve runs on the Stokes hierarchy of an orc_t (grid lines,
 * levels 0..Lt, Lt <= L) with its own node arrays: FP32 ro, cp, kt on all
 * levels, FP32 temperature tk, start-of-step temperature tk0 and source term
 * tk1 on level 0 (parallel/head3mg.c:95-104), FP64 solver vectors and a
 * 7-slot row store per node (pos0 = 7 * node, parallel/heat3mg.c:427).
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "stokes_oracle.h"

#define NIDH(o, l, m1, m2, m3) ((o)->mgp[l] + (long)(m3) * (o)->mgxy[l] + (long)(m1) * (o)->mgy[l] + (m2))

struct orh_s {
  orc_t *o;
  int Lt;
  long ntot;
  int *tbondm; /* ntot, 0 = no T boundary equation (the reference's bondm[nodenum8 + node]) */
  int nbond, cap;
  float *bv1; /* 2*(cap+1): constant, coefficient */
  int *bn1;   /* cap+1: partner node + 1 (0: none) */
  float *ro, *cp, *kt;    /* ntot */
  float *tk, *tk0, *tk1;  /* mgn[0] */
  float *tk2;             /* mgn[0]: the temperature at the start of the whole step (heat3mg.c:377-381) */
  double *sol0, *sol1, *fre0, *fre1; /* ntot */
  double *val1; /* 7*ntot */
  int *lin0;    /* 7*ntot */
  int *num0;    /* ntot */
  double timestep, t0koef, t1koef;
};

orh_t *orh_create(orc_t *o, int Lt) {
  if (!o || Lt < 0 || Lt > o->L) return NULL;
  orh_t *h = (orh_t *)calloc(1, sizeof(orh_t));
  h->o = o;
  h->Lt = Lt;
  h->ntot = o->mgp[Lt] + o->mgn[Lt];
  const long nt = h->ntot, n0 = o->mgn[0];
  h->tbondm = (int *)calloc(nt, sizeof(int));
  h->cap = 1024;
  h->bv1 = (float *)calloc(2 * (h->cap + 1), sizeof(float));
  h->bn1 = (int *)calloc(h->cap + 1, sizeof(int));
  h->ro = (float *)calloc(nt, sizeof(float));
  h->cp = (float *)calloc(nt, sizeof(float));
  h->kt = (float *)calloc(nt, sizeof(float));
  h->tk = (float *)calloc(n0, sizeof(float));
  h->tk0 = (float *)calloc(n0, sizeof(float));
  h->tk1 = (float *)calloc(n0, sizeof(float));
  h->tk2 = (float *)calloc(n0, sizeof(float));
  h->sol0 = (double *)calloc(nt, sizeof(double));
  h->sol1 = (double *)calloc(nt, sizeof(double));
  h->fre0 = (double *)calloc(nt, sizeof(double));
  h->fre1 = (double *)calloc(nt, sizeof(double));
  h->val1 = (double *)calloc(7 * nt, sizeof(double));
  h->lin0 = (int *)calloc(7 * nt, sizeof(int));
  h->num0 = (int *)calloc(nt, sizeof(int));
  h->timestep = 1.0;
  h->t0koef = 1.0;
  h->t1koef = 1.0;
  return h;
}

void orh_destroy(orh_t *h) {
  if (!h) return;
  free(h->tbondm); free(h->bv1); free(h->bn1);
  free(h->ro); free(h->cp); free(h->kt); free(h->tk); free(h->tk0); free(h->tk1); free(h->tk2);
  free(h->sol0); free(h->sol1); free(h->fre0); free(h->fre1);
  free(h->val1); free(h->lin0); free(h->num0);
  free(h);
}

void orh_set(orh_t *h, double timestep, double t0koef, double t1koef) {
  h->timestep = timestep;
  h->t0koef = t0koef;
  h->t1koef = t1koef;
}

void *orh_array(orh_t *h, const char *n) {
  if (!strcmp(n, "tbondm")) return h->tbondm;
  if (!strcmp(n, "bv1")) return h->bv1;
  if (!strcmp(n, "bn1")) return h->bn1;
  if (!strcmp(n, "ro")) return h->ro;
  if (!strcmp(n, "cp")) return h->cp;
  if (!strcmp(n, "kt")) return h->kt;
  if (!strcmp(n, "tk")) return h->tk;
  if (!strcmp(n, "tk0")) return h->tk0;
  if (!strcmp(n, "tk1")) return h->tk1;
  if (!strcmp(n, "tk2")) return h->tk2;
  if (!strcmp(n, "sol0")) return h->sol0;
  if (!strcmp(n, "fre0")) return h->fre0;
  if (!strcmp(n, "fre1")) return h->fre1;
  if (!strcmp(n, "val1")) return h->val1;
  if (!strcmp(n, "lin0")) return h->lin0;
  if (!strcmp(n, "num0")) return h->num0;
  return NULL;
}

int orh_nbond(const orh_t *h) { return h->nbond; }

/* ------------------------------------------------------------------------- */
/* T boundary equations on levels 0..Lt: parallel/in3mg.c:302-324 (index     */
/* ranges kept 3 outer lines per side), 369-385 (case 4, T):                  */
/* T_n = const + koef * T_{n + shift}, |shift| > 1 periodic (in3mg.c:339-341). */
/* ------------------------------------------------------------------------- */
/* cell search of m1serch/m2serch/m3serch (mark3mg.c:1138-1205): bisection
   for the largest a <= n-2 with g[a] <= x */
static long serch_h(const double *g, int n, double x) {
  long a = 0, b = n - 1, m;
  do {
    m = (a + b) / 2;
    if (g[m] > x) b = m; else a = m;
  } while ((b - a) > 1);
  if (a > n - 2) a = n - 2;
  return a;
}

static int lvl_index_h(int m, int nfine, int nlev, long mgs) {
  int r = (m - 2) / (int)mgs + 2;
  if (m < 3) r = m;
  if (m > nfine - 4) r = nlev - (nfine - m);
  return r;
}

int orh_bc_line(orh_t *h, int m10, int m11, int m20, int m21, int m30, int m31, float cnst, float koef, int dm1,
                int dm2, int dm3) {
  const orc_t *o = h->o;
  for (int l = 0; l <= h->Lt; l++) {
    int m1min = lvl_index_h(m10, o->mgx[0], o->mgx[l], o->mgs[l]);
    int m1max = lvl_index_h(m11, o->mgx[0], o->mgx[l], o->mgs[l]);
    int m2min = lvl_index_h(m20, o->mgy[0], o->mgy[l], o->mgs[l]);
    int m2max = lvl_index_h(m21, o->mgy[0], o->mgy[l], o->mgs[l]);
    int m3min = lvl_index_h(m30, o->mgz[0], o->mgz[l], o->mgs[l]);
    int m3max = lvl_index_h(m31, o->mgz[0], o->mgz[l], o->mgs[l]);
    for (int m3 = m3min; m3 <= m3max; m3++)
      for (int m1 = m1min; m1 <= m1max; m1++)
        for (int m2 = m2min; m2 <= m2max; m2++) {
          long m4 = NIDH(o, l, m1, m2, m3);
          int sx = dm1, sy = dm2, sz = dm3;
          if (abs(dm1) > 1) sx = -(o->mgx[l] - 1);
          if (abs(dm2) > 1) sy = -(o->mgy[l] - 1);
          if (abs(dm3) > 1) sz = -(o->mgz[l] - 1);
          long m6 = m4 + (long)sz * o->mgxy[l] + (long)sx * o->mgy[l] + sy;
          if (h->nbond + 1 > h->cap) {
            int nc = h->cap * 2;
            h->bv1 = (float *)realloc(h->bv1, 2 * (size_t)(nc + 1) * sizeof(float));
            h->bn1 = (int *)realloc(h->bn1, (size_t)(nc + 1) * sizeof(int));
            h->cap = nc;
          }
          h->nbond++;
          h->tbondm[m4] = h->nbond;
          h->bv1[2 * h->nbond + 0] = cnst;
          if (dm1 || dm2 || dm3) {
            h->bv1[2 * h->nbond + 1] = koef;
            h->bn1[h->nbond] = (int)(m6 + 1);
          } else {
            h->bv1[2 * h->nbond + 1] = 0;
            h->bn1[h->nbond] = 0;
          }
        }
  }
  return h->nbond;
}

/* ------------------------------------------------------------------------- */
/* Ro, Cp, kt restriction to level l: parallel/heat3mg.c:1235-1330.  Every   */
/* fine node adds to the 8 coarse nodes of its cell (index rules 1254-1272,  */
/* weights 1-|d|/dx per axis, 1282-1302), then each coarse value is divided  */
/* by its weight sum.  FP32 arrays, sums in double then stored as float.     */
/* ------------------------------------------------------------------------- */
static void cell_lo(int m, int nf, int nc, int *lo) {
  int r = (m - 2) / 2 + 2;
  if (m < 3) r = m;
  if (m > nf - 4) r = nc - (nf - m);
  if (r > nc - 2) r = nc - 2;
  *lo = r;
}

void orh_rocprestrict(orh_t *h, int l) {
  const orc_t *o = h->o;
  const int f = l - 1;
  double *S = h->sol1;
  for (long m1 = o->mgp[l]; m1 < o->mgp[l] + o->mgn[l]; m1++) {
    h->cp[m1] = 0; h->kt[m1] = 0; h->ro[m1] = 0; S[m1] = 0;
  }
  for (int m3 = 0; m3 < o->mgz[f]; m3++)
    for (int m1 = 0; m1 < o->mgx[f]; m1++)
      for (int m2 = 0; m2 < o->mgy[f]; m2++) {
        long m4 = NIDH(o, f, m1, m2, m3);
        int m10min, m20min, m30min;
        cell_lo(m1, o->mgx[f], o->mgx[l], &m10min);
        cell_lo(m2, o->mgy[f], o->mgy[l], &m20min);
        cell_lo(m3, o->mgz[f], o->mgz[l], &m30min);
        double dx = o->mggx[l][m10min + 1] - o->mggx[l][m10min];
        double dy = o->mggy[l][m20min + 1] - o->mggy[l][m20min];
        double dz = o->mggz[l][m30min + 1] - o->mggz[l][m30min];
        for (int m30 = m30min; m30 <= m30min + 1; m30++) {
          double ddz = (o->mggz[l][m30] - o->mggz[f][m3]) / dz;
          ddz = 1.0 - fabs(ddz);
          for (int m10 = m10min; m10 <= m10min + 1; m10++) {
            double ddx = (o->mggx[l][m10] - o->mggx[f][m1]) / dx;
            ddx = 1.0 - fabs(ddx);
            for (int m20 = m20min; m20 <= m20min + 1; m20++) {
              double ddy = (o->mggy[l][m20] - o->mggy[f][m2]) / dy;
              ddy = 1.0 - fabs(ddy);
              long m40 = NIDH(o, l, m10, m20, m30);
              double swt = ddx * ddy * ddz;
              h->cp[m40] += h->cp[m4] * swt;
              h->kt[m40] += h->kt[m4] * swt;
              h->ro[m40] += h->ro[m4] * swt;
              S[m40] += swt;
            }
          }
        }
      }
  for (long m1 = o->mgp[l]; m1 < o->mgp[l] + o->mgn[l]; m1++) {
    h->cp[m1] /= S[m1];
    h->kt[m1] /= S[m1];
    h->ro[m1] /= S[m1];
    S[m1] = 0;
  }
}

/* ------------------------------------------------------------------------- */
/* Row of node (m1, m2, m3) of level l: the implicit heat equation           */
/* heaterr(ynerr = 1), parallel/heat3mg.c:195-249 -- slots: 1 the node        */
/* (rho cp / dt plus the six face conductances, added x-, x+, y-, y+, z-, z+), */
/* then -k_face xkf / dx for x-, x+, y-, y+, z-, z+ (k_face the two-node      */
/* mean); x wraps on a periodic x axis (heat3mg.c:21, 36) -- or the boundary  */
/* equation tbonderr (heat3mg.c:869-893).  ui1[0]: tk1 + tk0 rho cp / dt on   */
/* level 0, 0 on coarser levels.                                              */
/* ------------------------------------------------------------------------- */
static double rocp_of(const orh_t *h, long v) {
  double r = (float)(h->ro[v] * h->cp[v]);
  if (r < 2e+6) r = 2e+6;
  return r;
}

int orh_row(orh_t *h, int l, int m1, int m2, int m3, long *un1, double *ui1) {
  const orc_t *o = h->o;
  long v3 = NIDH(o, l, m1, m2, m3);
  int b = h->tbondm[v3];
  if (b) {
    un1[0] = 1;
    ui1[0] = h->bv1[2 * b + 0];
    un1[1] = v3;
    ui1[1] = 1.0;
    if (h->bn1[b]) {
      un1[0] += 1;
      un1[un1[0]] = h->bn1[b] - 1;
      ui1[un1[0]] = -h->bv1[2 * b + 1];
    }
    return 1;
  }
  const double *gx = o->mggx[l], *gy = o->mggy[l], *gz = o->mggz[l];
  const int X = o->mgx[l];
  double xkf;
  if (o->xper && m1 == 0) xkf = 2.0 / (gx[1] + gx[X - 1] - gx[X - 2]);
  else xkf = 2.0 / (gx[m1 + 1] - gx[m1 - 1]);
  double ykf = 2.0 / (gy[m2 + 1] - gy[m2 - 1]);
  double zkf = 2.0 / (gz[m3 + 1] - gz[m3 - 1]);
  long v[7];
  v[3] = v3;
  v[0] = v3 - o->mgxy[l];
  if (o->xper && m1 == 0) v[1] = NIDH(o, l, X - 2, m2, m3);
  else v[1] = v3 - o->mgy[l];
  v[2] = v3 - 1;
  v[4] = v3 + 1;
  v[5] = v3 + o->mgy[l];
  v[6] = v3 + o->mgxy[l];
  const float *kt = h->kt;
  double rocpkf = rocp_of(h, v3), ival;
  if (l == 0) ui1[0] = h->tk1[v3 - o->mgp[0]] + h->tk0[v3 - o->mgp[0]] / h->timestep * rocpkf;
  else ui1[0] = 0;
  ui1[1] = rocpkf / h->timestep;
  un1[1] = v3;
  if (o->xper && m1 == 0) ival = (float)(kt[v[1]] + kt[v[3]]) / 2.0 * xkf / (gx[X - 1] - gx[X - 2]);
  else ival = (float)(kt[v[1]] + kt[v[3]]) / 2.0 * xkf / (gx[m1] - gx[m1 - 1]);
  ui1[2] = -ival; un1[2] = v[1];
  ui1[1] += ival;
  ival = (float)(kt[v[3]] + kt[v[5]]) / 2.0 * xkf / (gx[m1 + 1] - gx[m1]);
  ui1[1] += ival;
  ui1[3] = -ival; un1[3] = v[5];
  ival = (float)(kt[v[2]] + kt[v[3]]) / 2.0 * ykf / (gy[m2] - gy[m2 - 1]);
  ui1[4] = -ival; un1[4] = v[2];
  ui1[1] += ival;
  ival = (float)(kt[v[3]] + kt[v[4]]) / 2.0 * ykf / (gy[m2 + 1] - gy[m2]);
  ui1[1] += ival;
  ui1[5] = -ival; un1[5] = v[4];
  ival = (float)(kt[v[0]] + kt[v[3]]) / 2.0 * zkf / (gz[m3] - gz[m3 - 1]);
  ui1[6] = -ival; un1[6] = v[0];
  ui1[1] += ival;
  ival = (float)(kt[v[3]] + kt[v[6]]) / 2.0 * zkf / (gz[m3 + 1] - gz[m3]);
  ui1[1] += ival;
  ui1[7] = -ival; un1[7] = v[6];
  un1[0] = 7;
  return 0;
}

/* rows of levels 0..Lt into the 7-slot store (parallel/heat3mg.c:416-451) */
void orh_assemble(orh_t *h) {
  const orc_t *o = h->o;
  for (int l = 0; l <= h->Lt; l++) {
#pragma omp parallel for schedule(static)
    for (int m3 = 0; m3 < o->mgz[l]; m3++) {
      long un1[16];
      double ui1[16];
      for (int m1 = 0; m1 < o->mgx[l]; m1++)
        for (int m2 = 0; m2 < o->mgy[l]; m2++) {
          long n = NIDH(o, l, m1, m2, m3);
          orh_row(h, l, m1, m2, m3, un1, ui1);
          h->fre1[n] = ui1[0];
          h->num0[n] = (int)un1[0];
          for (int k = 0; k < un1[0]; k++) {
            h->lin0[7 * n + k] = (int)un1[k + 1];
            h->val1[7 * n + k] = ui1[k + 1];
          }
        }
    }
  }
}

static inline void row_update(orh_t *h, long m1) {
  double ival = -h->fre1[m1];
  double ival1 = h->val1[7 * m1];
  for (int m2 = 0; m2 < h->num0[m1]; m2++) ival += h->val1[7 * m1 + m2] * h->sol0[h->lin0[7 * m1 + m2]];
  h->sol0[m1] -= ival / ival1 * h->t0koef;
}

/* residual fre0 = fre1 - A sol0 on level l (parallel/heat3mg.c:1101-1117) */
void orh_residual(orh_t *h, int l) {
  const orc_t *o = h->o;
  const long a = o->mgp[l], e = o->mgp[l] + o->mgn[l];
#pragma omp parallel for schedule(static)
  for (long m1 = a; m1 < e; m1++)
    if (h->num0[m1]) {
      double ival = h->fre1[m1];
      for (int m2 = 0; m2 < h->num0[m1]; m2++) ival -= h->val1[7 * m1 + m2] * h->sol0[h->lin0[7 * m1 + m2]];
      h->fre0[m1] = ival;
    }
}

/* gausmatt (parallel/heat3mg.c:1057-1120): am Gauss-Seidel sweeps over the   */
/* level's rows in index order (the reference's loop on one thread), then    */
/* the residual                                                               */
void orh_gausmatt(orh_t *h, int am, int l) {
  const orc_t *o = h->o;
  const long a = o->mgp[l], e = o->mgp[l] + o->mgn[l];
  for (int s = 0; s < am; s++)
    for (long m1 = a; m1 < e; m1++)
      if (h->num0[m1]) row_update(h, m1);
  orh_residual(h, l);
}

static int tbc_depth(const orh_t *h, long r, int lim) {
  int d = 0;
  while (d <= lim) {
    int b = h->tbondm[r];
    if (!b || !h->bn1[b]) return d;
    long j = h->bn1[b] - 1;
    if (!h->tbondm[j]) return d;
    r = j;
    d++;
  }
  return -1;
}

/* The same row updates in the device's multicolour order: heat-equation    */
/* rows of colour (m1+m2+m3)&1 = 0, then of colour 1, then the boundary rows */
/* by dependency depth (a row's partner first); -1 if boundary rows cycle.   */
int orh_gausmatt_colored(orh_t *h, int am, int l) {
  const orc_t *o = h->o;
  const long a = o->mgp[l], e = o->mgp[l] + o->mgn[l];
  int maxd = 0;
  for (long r = a; r < e; r++)
    if (h->tbondm[r]) {
      int d = tbc_depth(h, r, 64);
      if (d < 0) return -1;
      if (d > maxd) maxd = d;
    }
  for (int s = 0; s < am; s++) {
    for (int col = 0; col < 2; col++)
#pragma omp parallel for schedule(static)
      for (int m3 = 0; m3 < o->mgz[l]; m3++)
        for (int m1 = 0; m1 < o->mgx[l]; m1++)
          for (int m2 = 0; m2 < o->mgy[l]; m2++) {
            if (((m1 + m2 + m3) & 1) != col) continue;
            long n = NIDH(o, l, m1, m2, m3);
            if (h->num0[n] && !h->tbondm[n]) row_update(h, n);
          }
    for (int d = 0; d <= maxd; d++)
      for (long r = a; r < e; r++)
        if (h->tbondm[r] && tbc_depth(h, r, 64) == d) row_update(h, r);
  }
  orh_residual(h, l);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Residual restriction to level l (parallel/heat3mg.c:1122-1232): every     */
/* fine residual adds to the coarse nodes of its cell without a boundary     */
/* equation, weights 1 -+ d/dx per axis (1170-1192); then division by the    */
/* weight sums, coarse solution 0.                                            */
/* ------------------------------------------------------------------------- */
void orh_trestrict(orh_t *h, int l) {
  const orc_t *o = h->o;
  const int f = l - 1;
  for (long m1 = o->mgp[l]; m1 < o->mgp[l] + o->mgn[l]; m1++) { h->fre1[m1] = 0; h->sol1[m1] = 0; }
  for (int m3 = 0; m3 < o->mgz[f]; m3++)
    for (int m1 = 0; m1 < o->mgx[f]; m1++)
      for (int m2 = 0; m2 < o->mgy[f]; m2++) {
        long m4 = NIDH(o, f, m1, m2, m3);
        int m10min, m20min, m30min;
        cell_lo(m1, o->mgx[f], o->mgx[l], &m10min);
        cell_lo(m2, o->mgy[f], o->mgy[l], &m20min);
        cell_lo(m3, o->mgz[f], o->mgz[l], &m30min);
        double dx = o->mggx[l][m10min + 1] - o->mggx[l][m10min];
        double dy = o->mggy[l][m20min + 1] - o->mggy[l][m20min];
        double dz = o->mggz[l][m30min + 1] - o->mggz[l][m30min];
        for (int m30 = m30min; m30 <= m30min + 1; m30++) {
          double ddz = (o->mggz[l][m30] - o->mggz[f][m3]) / dz;
          if (m30 == m30min) ddz = -ddz;
          ddz = 1.0 - ddz;
          for (int m10 = m10min; m10 <= m10min + 1; m10++) {
            double ddx = (o->mggx[l][m10] - o->mggx[f][m1]) / dx;
            if (m10 == m10min) ddx = -ddx;
            ddx = 1.0 - ddx;
            for (int m20 = m20min; m20 <= m20min + 1; m20++) {
              double ddy = (o->mggy[l][m20] - o->mggy[f][m2]) / dy;
              if (m20 == m20min) ddy = -ddy;
              ddy = 1.0 - ddy;
              long m40 = NIDH(o, l, m10, m20, m30);
              if (!h->tbondm[m40]) {
                double swt = ddx * ddy * ddz;
                h->fre1[m40] += h->fre0[m4] * swt;
                h->sol1[m40] += swt;
              }
            }
          }
        }
      }
  for (long m1 = o->mgp[l]; m1 < o->mgp[l] + o->mgn[l]; m1++) {
    if (h->sol1[m1]) {
      h->fre1[m1] /= h->sol1[m1];
      h->sol1[m1] = 0;
    }
    h->sol0[m1] = 0;
  }
}

/* Correction prolongation from level l (parallel/heat3mg.c:954-1050): every */
/* fine node without a boundary equation adds t1koef * sum w sol0_c over the */
/* 8 coarse nodes of its cell, weights as the restriction's, unnormalised.   */
void orh_tprolong(orh_t *h, int l, int ynreset) {
  const orc_t *o = h->o;
  const int f = l - 1;
#pragma omp parallel for schedule(static)
  for (int m3 = 0; m3 < o->mgz[f]; m3++)
    for (int m1 = 0; m1 < o->mgx[f]; m1++)
      for (int m2 = 0; m2 < o->mgy[f]; m2++) {
        long m4 = NIDH(o, f, m1, m2, m3);
        if (h->tbondm[m4]) continue;
        int m10min, m20min, m30min;
        cell_lo(m1, o->mgx[f], o->mgx[l], &m10min);
        cell_lo(m2, o->mgy[f], o->mgy[l], &m20min);
        cell_lo(m3, o->mgz[f], o->mgz[l], &m30min);
        double dx = o->mggx[l][m10min + 1] - o->mggx[l][m10min];
        double dy = o->mggy[l][m20min + 1] - o->mggy[l][m20min];
        double dz = o->mggz[l][m30min + 1] - o->mggz[l][m30min];
        if (ynreset) h->sol0[m4] = 0;
        for (int m30 = m30min; m30 <= m30min + 1; m30++) {
          double ddz = (o->mggz[l][m30] - o->mggz[f][m3]) / dz;
          if (m30 == m30min) ddz = -ddz;
          ddz = 1.0 - ddz;
          for (int m10 = m10min; m10 <= m10min + 1; m10++) {
            double ddx = (o->mggx[l][m10] - o->mggx[f][m1]) / dx;
            if (m10 == m10min) ddx = -ddx;
            ddx = 1.0 - ddx;
            for (int m20 = m20min; m20 <= m20min + 1; m20++) {
              double ddy = (o->mggy[l][m20] - o->mggy[f][m2]) / dy;
              if (m20 == m20min) ddy = -ddy;
              ddy = 1.0 - ddy;
              long m40 = NIDH(o, l, m10, m20, m30);
              double swt = ddx * ddy * ddz;
              h->sol0[m4] += h->sol0[m40] * swt * h->t1koef;
            }
          }
        }
      }
}

/* ------------------------------------------------------------------------- */
/* The right part of heaterr (ynerr = 0, parallel/heat3mg.c:54-187) on the   */
/* level-0 heat-equation nodes: tk1 = ht, plus the adiabatic term (adiabyn 2: */
/* et tk0 v.grad(P) with the pressure gradient of the 8 cells around the node */
/* at the node's relative position, :100-115; adiabyn 1: et tk0 ro v.g,      */
/* :131-134) and the frictional heating (frictyn: the normal stress-strain   */
/* products of the 8 cells and twice the shear ones of the node's 6 edges,  */
/* :140-175), both where ro > 2e3, with the velocity vxyzcalc1 interpolates  */
/* at the node (orc_interp on the FP32 velocities of x).  The reference's   */
/* types: FP32 ht, et, ro, tk0, stresses and strain rates (sig, eps: 6      */
/* planes each, xx yy zz xy xz yz, rounded to FP32 here), FP64 pressure     */
/* (x[4 n]).  Boundary-row nodes keep tk1 = ht.                             */
/* ------------------------------------------------------------------------- */
void orh_sources(orh_t *h, const double *x, const double *sig, const double *eps, const float *ht, const float *et,
                 int adiabyn, int frictyn, double gxk, double gyk, double gzk) {
  const orc_t *o = h->o;
  const int X = o->mgx[0], Y = o->mgy[0], Z = o->mgz[0];
  const long XY = o->mgxy[0], n0 = o->mgn[0];
  const double *gx = o->mggx[0], *gy = o->mggy[0], *gz = o->mggz[0];
#define F32(a, k, i) ((float)(a)[(size_t)(k) * n0 + (i)])
#pragma omp parallel for schedule(static)
  for (int m3 = 0; m3 < Z; m3++)
    for (int m1 = 0; m1 < X; m1++)
      for (int m2 = 0; m2 < Y; m2++) {
        const long v3 = NIDH(o, 0, m1, m2, m3);
        if (h->tbondm[v3]) { h->tk1[v3] = ht[v3]; continue; }
        double ival = 0;
        ival += ht[v3];
        double xkf;
        if (o->xper && m1 == 0) xkf = 2.0 / (gx[1] + gx[X - 1] - gx[X - 2]);
        else xkf = 2.0 / (gx[m1 + 1] - gx[m1 - 1]);
        const double ykf = 2.0 / (gy[m2 + 1] - gy[m2 - 1]);
        const double zkf = 2.0 / (gz[m3 + 1] - gz[m3 - 1]);
        long v[7];
        v[3] = v3;
        v[0] = v3 - XY;
        v[1] = (o->xper && m1 == 0) ? NIDH(o, 0, X - 2, m2, m3) : v3 - Y;
        v[2] = v3 - 1;
        v[4] = v3 + 1;
        v[5] = v3 + Y;
        v[6] = v3 + XY;
        long p[8];
        double rx = 0, ry = 0, rz = 0;
        if (adiabyn == 2 || frictyn) {
          p[0] = v3; p[1] = p[0] + 1; p[2] = p[0] + Y; p[3] = p[2] + 1;
          p[4] = p[0] + XY; p[5] = p[4] + 1; p[6] = p[4] + Y; p[7] = p[6] + 1;
          if (o->xper && m1 == 0) rx = (gx[X - 1] - gx[X - 2]) / 2.0 * xkf;
          else rx = (gx[m1] - gx[m1 - 1]) / 2.0 * xkf;
          ry = (gy[m2] - gy[m2 - 1]) / 2.0 * ykf;
          rz = (gz[m3] - gz[m3 - 1]) / 2.0 * zkf;
        }
        double vv[3] = {0, 0, 0};
        if (adiabyn) orc_interp(o, x, NULL, gx[m1], gy[m2], gz[m3], vv, NULL);
        const float r = h->ro[v3], tk0 = h->tk[v3];
        if (adiabyn == 2 && r > 2e+3) {
#define PR(k) x[4 * p[k]]
          const double dpdx = xkf * ((1.0 - ry) * (1.0 - rz) * (PR(2) - PR(0)) + ry * (1.0 - rz) * (PR(3) - PR(1)) +
                                     (1.0 - ry) * rz * (PR(6) - PR(4)) + ry * rz * (PR(7) - PR(5)));
          const double dpdy = ykf * ((1.0 - rx) * (1.0 - rz) * (PR(1) - PR(0)) + rx * (1.0 - rz) * (PR(3) - PR(2)) +
                                     (1.0 - rx) * rz * (PR(5) - PR(4)) + rx * rz * (PR(7) - PR(6)));
          const double dpdz = zkf * ((1.0 - rx) * (1.0 - ry) * (PR(4) - PR(0)) + rx * (1.0 - ry) * (PR(6) - PR(2)) +
                                     (1.0 - rx) * ry * (PR(5) - PR(1)) + rx * ry * (PR(7) - PR(3)));
#undef PR
          ival += et[v3] * tk0 * (vv[0] * dpdx + vv[1] * dpdy + vv[2] * dpdz);
        }
        if (adiabyn == 1 && r > 2e+3) ival += et[v3] * tk0 * r * (vv[0] * gxk + vv[1] * gyk + vv[2] * gzk);
        if (frictyn && r > 2e+3) {
#define NRM(k) (F32(sig, 0, p[k]) * F32(eps, 0, p[k]) + F32(sig, 1, p[k]) * F32(eps, 1, p[k]) + \
                F32(sig, 2, p[k]) * F32(eps, 2, p[k]))
          double epsval;
          epsval = (1.0 - rx) * (1.0 - ry) * (1.0 - rz) * NRM(0);
          epsval += (1.0 - rx) * (ry) * (1.0 - rz) * NRM(1);
          epsval += (rx) * (1.0 - ry) * (1.0 - rz) * NRM(2);
          epsval += (rx) * (ry) * (1.0 - rz) * NRM(3);
          epsval += (1.0 - rx) * (1.0 - ry) * (rz) * NRM(4);
          epsval += (1.0 - rx) * (ry) * (rz) * NRM(5);
          epsval += (rx) * (1.0 - ry) * (rz) * NRM(6);
          epsval += (rx) * (ry) * (rz) * NRM(7);
#undef NRM
          epsval += 2.0 * ((1.0 - rz) * (F32(sig, 3, v[0]) * F32(eps, 3, v[0])) + rz * (F32(sig, 3, v3) * F32(eps, 3, v3)));
          epsval += 2.0 * ((1.0 - ry) * (F32(sig, 4, v[2]) * F32(eps, 4, v[2])) + ry * (F32(sig, 4, v3) * F32(eps, 4, v3)));
          epsval += 2.0 * ((1.0 - rx) * (F32(sig, 5, v[1]) * F32(eps, 5, v[1])) + rx * (F32(sig, 5, v3) * F32(eps, 5, v3)));
          ival += epsval;
        }
        h->tk1[v3] = (float)ival;
      }
#undef F32
}

/* ------------------------------------------------------------------------- */
/* Marker temperatures after the step (titerate, parallel/heat3mg.c:618-950),  */
/* markers in index order (the reference's loops on one thread):              */
/*  - tkcalc1 (mark3mg.c:1212-1278): the marker's cell by serch, trilinear   */
/*    node weights 1 - |d| (z outer, x, y inner), tk and tk2 interpolated;   */
/*  - fluid markers (50 <= type < 100) inside take the new temperature;      */
/*  - normal markers (type < 50) strictly inside: markk <= 0 takes the old   */
/*    one (tdeep where y > zdeep and it is colder); with heatdif, the        */
/*    boundary-cell cap at the nodes' largest tk, the numerical-diffusion    */
/*    change dtk1 = (Told - markk) (1 - exp(-heatdif kt/(rho cp) 2 (1/dx^2 + */
/*    1/dy^2 + 1/dz^2) dt)) (the exponent clamped at -150), added to the     */
/*    marker and scattered to the 8 nodes with the trilinear weights (Sol1:  */
/*    dtk1 w, Sol0: w); then the nodal change Tnew - Told;                    */
/*  - with heatdif: Sol1 / Sol0 at the nodes, and each normal marker loses   */
/*    sum w Sol1 over its nodes with Sol1 != 0 (antidiffusion).              */
/* FP32 markk, node fields as the reference's arrays hold them.  exp is the  */
/* host libm's.                                                               */
/* ------------------------------------------------------------------------- */
static void tkcalc(const orh_t *h, double x, double y, double z, long *wn, double *wi, double *tn, double *to) {
  const orc_t *o = h->o;
  const double *gx = o->mggx[0], *gy = o->mggy[0], *gz = o->mggz[0];
  const long m10 = serch_h(gx, o->mgx[0], x), m20 = serch_h(gy, o->mgy[0], y), m30 = serch_h(gz, o->mgz[0], z);
  wn[0] = m10; wn[1] = m20; wn[2] = m30;
  const double dx = gx[m10 + 1] - gx[m10], dy = gy[m20 + 1] - gy[m20], dz = gz[m30 + 1] - gz[m30];
  wi[0] = dx; wi[1] = dy; wi[2] = dz;
  double a = 0, b = 0;
  int n1 = 0;
  for (long m3 = m30; m3 <= m30 + 1; m3++) {
    double ddz = (gz[m3] - z) / dz;
    ddz = 1.0 - fabs(ddz);
    for (long m1 = m10; m1 <= m10 + 1; m1++) {
      double ddx = (gx[m1] - x) / dx;
      ddx = 1.0 - fabs(ddx);
      for (long m2 = m20; m2 <= m20 + 1; m2++) {
        double ddy = (gy[m2] - y) / dy;
        ddy = 1.0 - fabs(ddy);
        const double swt = ddx * ddy * ddz;
        const long m4 = m3 * o->mgxy[0] + m1 * o->mgy[0] + m2;
        wi[3 + n1] = swt;
        wn[3 + n1] = m4;
        a += swt * h->tk[m4];
        b += swt * h->tk2[m4];
        n1++;
      }
    }
  }
  *tn = a;
  *to = b;
}

void orh_markers(orh_t *h, long nmark, const float *mx, const float *my, const float *mz, const int *markt,
                 float *markk, double heatdif, double timestep, double zdeep, double tdeep) {
  const orc_t *o = h->o;
  const int X = o->mgx[0], Y = o->mgy[0], Z = o->mgz[0];
  const double xs = o->mggx[0][X - 1], ys = o->mggy[0][Y - 1], zs = o->mggz[0][Z - 1];
  const long n0 = o->mgn[0];
  double *S0 = (double *)calloc(n0, sizeof(double)), *S1 = (double *)calloc(n0, sizeof(double));
  long wn[11];
  double wi[11], tn, to;
  for (long i = 0; i < nmark; i++)
    if (markt[i] >= 50 && markt[i] < 100 && mx[i] > 0 && my[i] > 0 && mz[i] > 0 && (double)mx[i] < xs &&
        (double)my[i] < ys && (double)mz[i] < zs) {
      tkcalc(h, mx[i], my[i], mz[i], wn, wi, &tn, &to);
      markk[i] = (float)tn;
    }
  for (long i = 0; i < nmark; i++) {
    if (!(markt[i] < 50 && mx[i] > 0 && my[i] > 0 && mz[i] > 0 && (double)mx[i] < xs && (double)my[i] < ys &&
          (double)mz[i] < zs))
      continue;
    tkcalc(h, mx[i], my[i], mz[i], wn, wi, &tn, &to);
    if (markk[i] <= 0) {
      markk[i] = (float)to;
      if (my[i] > zdeep && markk[i] < tdeep) markk[i] = (float)tdeep;
    }
    if (heatdif) {
      double mkt = 0, mcp = 0, mro = 0, mtk = h->tk[wn[3]];
      for (int n1 = 0; n1 < 8; n1++) {
        const long m4 = wn[3 + n1];
        const double ival = wi[3 + n1];
        mkt += h->kt[m4] * ival;
        mcp += h->cp[m4] * ival;
        mro += h->ro[m4] * ival;
        if (h->tk[m4] > mtk) mtk = h->tk[m4];
      }
      if ((wn[0] <= 0 || wn[0] >= X - 2 || wn[1] <= 0 || wn[1] >= Y - 2 || wn[2] <= 0 || wn[2] >= Z - 2) &&
          (markk[i] + tn - to) > mtk)
        markk[i] = (float)(mtk - tn + to);
      const double dtk = to - (double)markk[i];
      double rocpkf = mro * mcp;
      if (rocpkf < 2e+6) rocpkf = 2e+6;
      double dtk1 = -heatdif * mkt / rocpkf * 2.0 * (1.0 / wi[0] / wi[0] + 1.0 / wi[1] / wi[1] + 1.0 / wi[2] / wi[2]) *
                    timestep;
      if (dtk1 < -150.0) dtk1 = -150.0;
      dtk1 = dtk * (1.0 - exp(dtk1));
      markk[i] += (float)dtk1;
      for (int n1 = 0; n1 < 8; n1++) {
        const long m4 = wn[3 + n1];
        const double ival = wi[3 + n1];
        S1[m4] += dtk1 * ival;
        S0[m4] += ival;
      }
    }
    markk[i] += (float)(tn - to);
  }
  if (heatdif) {
    for (long n = 0; n < n0; n++)
      if (S0[n]) S1[n] /= S0[n];
    for (long i = 0; i < nmark; i++) {
      if (!(markt[i] < 50 && mx[i] > 0 && my[i] > 0 && mz[i] > 0 && (double)mx[i] < xs && (double)my[i] < ys &&
            (double)mz[i] < zs))
        continue;
      tkcalc(h, mx[i], my[i], mz[i], wn, wi, &tn, &to);
      for (int n1 = 0; n1 < 8; n1++) {
        const long m4 = wn[3 + n1];
        if (S1[m4]) markk[i] -= (float)(S1[m4] * wi[3 + n1]);
      }
    }
  }
  free(S0);
  free(S1);
}

/* ------------------------------------------------------------------------- */
/* Errors of the level-0 temperature tk (parallel/heat3mg.c:528-565):        */
/* heatsum = RMS over heat-equation nodes of heaterr(ynerr = 2) * timestep   */
/* (heat3mg.c:251-287) on the FP32 tk, tk0, tk1 as the reference's, and     */
/* bondsum = RMS of the boundary equations' errors on the FP64 solution      */
/* (tbonderr ynerr = 1, 854-867).                                             */
/* ------------------------------------------------------------------------- */
void orh_err(orh_t *h, double *heatsum, double *bondsum) {
  const orc_t *o = h->o;
  const int X = o->mgx[0];
  const double *gx = o->mggx[0], *gy = o->mggy[0], *gz = o->mggz[0];
  double hs = 0, hn = 0, bs = 0, bn = 0;
  const float *tk = h->tk, *tk0 = h->tk0, *tk1 = h->tk1, *kt = h->kt;
  for (int m3 = 0; m3 < o->mgz[0]; m3++)
    for (int m1 = 0; m1 < X; m1++)
      for (int m2 = 0; m2 < o->mgy[0]; m2++) {
        long n = NIDH(o, 0, m1, m2, m3);
        int b = h->tbondm[n];
        if (b) {
          double lb = h->sol0[n] - h->bv1[2 * b + 0];
          if (h->bn1[b]) lb -= h->bv1[2 * b + 1] * h->sol0[h->bn1[b] - 1];
          bs += lb * lb;
          bn += 1.0;
          continue;
        }
        double xkf;
        if (o->xper && m1 == 0) xkf = 2.0 / (gx[1] + gx[X - 1] - gx[X - 2]);
        else xkf = 2.0 / (gx[m1 + 1] - gx[m1 - 1]);
        double ykf = 2.0 / (gy[m2 + 1] - gy[m2 - 1]);
        double zkf = 2.0 / (gz[m3 + 1] - gz[m3 - 1]);
        long v[7];
        v[3] = n;
        v[0] = n - o->mgxy[0];
        v[1] = (o->xper && m1 == 0) ? NIDH(o, 0, X - 2, m2, m3) : n - o->mgy[0];
        v[2] = n - 1; v[4] = n + 1; v[5] = n + o->mgy[0]; v[6] = n + o->mgxy[0];
        double rocpkf = rocp_of(h, n), ival, ival1;
        ival1 = (tk[n] - tk0[n]) / h->timestep;
        if (o->xper && m1 == 0) ival = (float)(kt[v[1]] + kt[v[3]]) / 2.0 * xkf / (gx[X - 1] - gx[X - 2]) / rocpkf;
        else ival = (float)(kt[v[1]] + kt[v[3]]) / 2.0 * xkf / (gx[m1] - gx[m1 - 1]) / rocpkf;
        ival1 += (tk[v[3]] - tk[v[1]]) * ival;
        ival = (float)(kt[v[3]] + kt[v[5]]) / 2.0 * xkf / (gx[m1 + 1] - gx[m1]) / rocpkf;
        ival1 -= (tk[v[5]] - tk[v[3]]) * ival;
        ival = (float)(kt[v[2]] + kt[v[3]]) / 2.0 * ykf / (gy[m2] - gy[m2 - 1]) / rocpkf;
        ival1 += (tk[v[3]] - tk[v[2]]) * ival;
        ival = (float)(kt[v[3]] + kt[v[4]]) / 2.0 * ykf / (gy[m2 + 1] - gy[m2]) / rocpkf;
        ival1 -= (tk[v[4]] - tk[v[3]]) * ival;
        ival = (float)(kt[v[0]] + kt[v[3]]) / 2.0 * zkf / (gz[m3] - gz[m3 - 1]) / rocpkf;
        ival1 += (tk[v[3]] - tk[v[0]]) * ival;
        ival = (float)(kt[v[3]] + kt[v[6]]) / 2.0 * zkf / (gz[m3 + 1] - gz[m3]) / rocpkf;
        ival1 -= (tk[v[6]] - tk[v[3]]) * ival;
        ival1 -= tk1[n] / rocpkf;
        ival1 *= h->timestep;
        hs += ival1 * ival1;
        hn += 1.0;
      }
  *heatsum = hn > 0 ? sqrt(hs / hn) : 0.0;
  *bondsum = bn > 0 ? sqrt(bs / bn) : 0.0;
}

/* ------------------------------------------------------------------------- */
/* One implicit temperature step (titerate, parallel/heat3mg.c:384-590):     */
/* ro/cp/kt restricted to levels 1..Lt, rows assembled, sol0 = tk, then      */
/* groups of multicyct V-cycles (up: smooth pre[l] + restrict, down: smooth  */
/* post[l] + prolong, level 0 not smoothed after the last prolongation,      */
/* heat3mg.c:463-492; Lt = 0: pre[0] sweeps) until heatsum <= heatmin or     */
/* maxit groups; tk = sol0 after each group.  colored: the device's sweep    */
/* order.  Returns the number of groups (negative: not converged).            */
/* ------------------------------------------------------------------------- */
static int orh_substep(orh_t *h, const int *pre, const int *post, int multicyct, double heatmin, int maxit,
                       int colored, double *heatsum, double *bondsum) {
  const orc_t *o = h->o;
  for (long m1 = 0; m1 < o->mgn[0]; m1++) h->tk0[m1] = h->tk[m1];
  for (long m1 = 0; m1 < o->mgn[0]; m1++) h->sol0[m1] = h->tk[m1];
  orh_assemble(h);
  double hs = 0, bs = 0;
  int it = 0;
  do {
    for (int c = 0; c < multicyct; c++) {
      if (h->Lt > 0) {
        for (int l = 0; l < h->Lt; l++) {
          if (colored) orh_gausmatt_colored(h, pre[l], l);
          else orh_gausmatt(h, pre[l], l);
          orh_trestrict(h, l + 1);
        }
        for (int l = h->Lt; l > 0; l--) {
          if (colored) orh_gausmatt_colored(h, post[l], l);
          else orh_gausmatt(h, post[l], l);
          orh_tprolong(h, l, 0);
        }
      } else {
        if (colored) orh_gausmatt_colored(h, pre[0], 0);
        else orh_gausmatt(h, pre[0], 0);
      }
    }
    for (long m1 = 0; m1 < o->mgn[0]; m1++) h->tk[m1] = (float)h->sol0[m1];
    orh_err(h, &hs, &bs);
    it++;
  } while (hs > heatmin && it < maxit);
  if (heatsum) *heatsum = hs;
  if (bondsum) *bondsum = bs;
  return hs > heatmin ? -it : it;
}

int orh_step(orh_t *h, const int *pre, const int *post, int multicyct, double heatmin, int maxit, int colored,
             double *heatsum, double *bondsum) {
  for (int l = 1; l <= h->Lt; l++) orh_rocprestrict(h, l);
  return orh_substep(h, pre, post, multicyct, heatmin, maxit, colored, heatsum, bondsum);
}

/* ------------------------------------------------------------------------- */
/* The whole thermal step of titerate (parallel/heat3mg.c:313-616): the      */
/* largest heating rate |tk1 / rho cp| over the heat-equation nodes           */
/* (heaterr ynerr = 0 with the source term tk1 alone, 54-62 and 183-187)     */
/* limits the sub-step to maxtkstep / rate (373-382); nothing is done for a  */
/* zero step (387); ro/cp/kt restricted once (397); sub-steps of             */
/* min(timestep, limit), the last one cut to what is left (403-405), each    */
/* the iteration of orh_step from tk0 = tk (408-411) until heatsum <=        */
/* heatmin.  out[8]: sub-steps, iterations (all sub-steps), heatsum and      */
/* bondsum of the last iteration, the sub-step limit, tk min and max after,  */
/* and 1 if every sub-step converged.  Returns the sub-step count.           */
/* ------------------------------------------------------------------------- */
int orh_titerate(orh_t *h, double timestep, double maxtkstep, const int *pre, const int *post, int multicyct,
                 double heatmin, int maxit, int colored, double *out) {
  const orc_t *o = h->o;
  double rate = 0, tkstep = 1e+30;
  for (long n = 0; n < o->mgn[0]; n++)
    if (!h->tbondm[n]) {
      double v = fabs(h->tk1[n] / rocp_of(h, n));
      if (v > rate) rate = v;
    }
  if (rate) tkstep = maxtkstep / rate;
  int sub = 0, its = 0, conv = 1;
  double hs = 0, bs = 0;
  if (timestep != 0 && tkstep != 0) {
    for (long n = 0; n < o->mgn[0]; n++) h->tk2[n] = h->tk[n];
    for (int l = 1; l <= h->Lt; l++) orh_rocprestrict(h, l);
    double tsum = 0, tsum0 = timestep;
    do {
      double dt = tsum0 < tkstep ? tsum0 : tkstep;
      if (dt > tsum0 - tsum) dt = tsum0 - tsum;
      h->timestep = dt;
      int it = orh_substep(h, pre, post, multicyct, heatmin, maxit, colored, &hs, &bs);
      if (it < 0) conv = 0;
      its += it < 0 ? -it : it;
      sub++;
      tsum += dt;
    } while (tsum < tsum0);
    h->timestep = tsum0;
  }
  double mn = 1e30, mx = -1e30;
  for (long n = 0; n < o->mgn[0]; n++) {
    if (h->tk[n] < mn) mn = h->tk[n];
    if (h->tk[n] > mx) mx = h->tk[n];
  }
  if (out) {
    out[0] = sub; out[1] = its; out[2] = hs; out[3] = bs; out[4] = tkstep; out[5] = mn; out[6] = mx; out[7] = conv;
  }
  return sub;
}
