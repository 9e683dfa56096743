/*
 * stokes_oracle.c -- CPU restatement of the reference Stokes multigrid path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle that the HIP library is
 * checked against and the CPU baseline that bench.py times; the product
 * library never links it.  Each function names the reference file:line it
 * restates.  Arithmetic order and storage types follow the reference's
 * single-thread (OMP_NUM_THREADS=1) execution so that results can be compared
 * exactly where the reference itself is deterministic.
 */
#include "stokes_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#define OMP_FOR _Pragma("omp parallel for schedule(static)")

#define ABSV(a) ((a) >= 0 ? (a) : -(a))
#define NID(o, l, m1, m2, m3) ((o)->mgp[l] + (long)(m3) * (o)->mgxy[l] + (long)(m1) * (o)->mgy[l] + (m2))

/* ------------------------------------------------------------------------- */
/* Hierarchy: parallel/load3mg.c:2011-2195 (gridcheck)                        */
/* ------------------------------------------------------------------------- */
orc_t *orc_create(int nx, int ny, int nz, int L, const double *gx, const double *gy,
                  const double *gz, int xper, int yper, int zper) {
  if (L < 0 || L >= ORC_MAXLVL || nx < 5 || ny < 5 || nz < 5) return NULL;
  orc_t *o = (orc_t *)calloc(1, sizeof(orc_t));
  o->L = L;
  o->xper = xper;
  o->yper = yper;
  o->zper = zper;
  o->mgn[0] = (long)nx * ny * nz;
  o->mgp[0] = 0;
  o->mgs[0] = 1;
  o->mgx[0] = nx;
  o->mgy[0] = ny;
  o->mgz[0] = nz;
  o->mgxy[0] = nx * ny;
  /* level sizes, load3mg.c:2064-2075 */
  for (int l = 1; l <= L; l++) {
    o->mgp[l] = o->mgp[l - 1] + o->mgn[l - 1];
    o->mgs[l] = o->mgs[l - 1] * 2;
    o->mgx[l] = ((o->mgx[l - 1] - 5) / 2) + 5;
    o->mgy[l] = ((o->mgy[l - 1] - 5) / 2) + 5;
    o->mgz[l] = ((o->mgz[l - 1] - 5) / 2) + 5;
    o->mgxy[l] = o->mgx[l] * o->mgy[l];
    o->mgn[l] = (long)o->mgx[l] * o->mgy[l] * o->mgz[l];
  }
  /* coarse coordinates keep the 3 outer lines per side, load3mg.c:2078-2190 */
  const double *g[3] = {gx, gy, gz};
  int nf[3] = {nx, ny, nz};
  for (int l = 0; l <= L; l++) {
    int nl[3] = {o->mgx[l], o->mgy[l], o->mgz[l]};
    double **dst[3] = {&o->mggx[l], &o->mggy[l], &o->mggz[l]};
    for (int a = 0; a < 3; a++) {
      double *c = (double *)calloc(nl[a], sizeof(double));
      c[0] = g[a][0];
      c[1] = g[a][1];
      c[2] = g[a][2];
      c[nl[a] - 1] = g[a][nf[a] - 1];
      c[nl[a] - 2] = g[a][nf[a] - 2];
      c[nl[a] - 3] = g[a][nf[a] - 3];
      for (int n2 = 3; n2 < nl[a] - 3; n2++) c[n2] = g[a][(n2 - 2) * o->mgs[l] + 2];
      *dst[a] = c;
    }
  }
  o->xsize = gx[nx - 1] - gx[0];
  o->ysize = gy[ny - 1] - gy[0];
  o->zsize = gz[nz - 1] - gz[0];
  long nt = o->mgp[L] + o->mgn[L];
  o->ntot = nt;
  o->nu = (float *)calloc(nt, sizeof(float));
  o->ro = (float *)calloc(nt, sizeof(float));
  o->nd = (float *)calloc(nt, sizeof(float));
  o->nxx = (float *)calloc(o->mgn[0], sizeof(float));
  o->nxy = (float *)calloc(o->mgn[0], sizeof(float));
  o->nxz = (float *)calloc(o->mgn[0], sizeof(float));
  o->nyz = (float *)calloc(o->mgn[0], sizeof(float));
  o->bondm = (int *)calloc(4 * nt, sizeof(int));
  o->capbond = 1024;
  o->bondv1 = (float *)calloc(2 * (o->capbond + 1), sizeof(float));
  o->bondn1 = (int *)calloc(o->capbond + 1, sizeof(int));
  o->sol0 = (double *)calloc(4 * nt, sizeof(double));
  o->sol1 = (double *)calloc(4 * nt, sizeof(double));
  o->fre0 = (double *)calloc(4 * nt, sizeof(double));
  o->fre1 = (double *)calloc(4 * nt, sizeof(double));
  o->tmp = (double *)calloc(4 * nt, sizeof(double));
  o->val1 = (double *)calloc(57 * nt, sizeof(double));
  o->lin0 = (int *)calloc(57 * nt, sizeof(int));
  o->num0 = (int *)calloc(4 * nt, sizeof(int));
  o->pos0 = (int *)calloc(4 * nt, sizeof(int));
  /* defaults: shipped effective knobs (parallel/mode.t3c:87-95, per-file overrides) */
  orc_set_knobs(o, 0.10, 0.0, 0.10, 1.0, 1.0, 1.0, 1e4, 0, 0.0, 0.0, 0.0);
  return o;
}

void orc_destroy(orc_t *o) {
  if (!o) return;
  for (int l = 0; l <= o->L; l++) {
    free(o->mggx[l]);
    free(o->mggy[l]);
    free(o->mggz[l]);
  }
  free(o->nu); free(o->ro); free(o->nd);
  free(o->nxx); free(o->nxy); free(o->nxz); free(o->nyz);
  free(o->bondm); free(o->bondv1); free(o->bondn1);
  free(o->sol0); free(o->sol1); free(o->fre0); free(o->fre1); free(o->tmp);
  free(o->val1); free(o->lin0); free(o->num0); free(o->pos0);
  free(o->lu); free(o->piv);
  free(o);
}

void orc_set_vonly(orc_t *o, int flag) { o->vonly = flag; }

void orc_set_knobs(orc_t *o, double p0, double p1, double p2, double v0, double v1,
                   double nubeg, double nuend, int viscmod, double gxk, double gyk, double gzk) {
  o->p0koef = p0; o->p1koef = p1; o->p2koef = p2; o->v0koef = v0; o->v1koef = v1;
  o->nubeg = nubeg; o->nuend = nuend; o->viscmod = viscmod;
  o->gxk = gxk; o->gyk = gyk; o->gzk = gzk;
}

int orc_level_info(const orc_t *o, int l, int *out6) {
  if (l < 0 || l > o->L) return -1;
  out6[0] = o->mgx[l]; out6[1] = o->mgy[l]; out6[2] = o->mgz[l];
  out6[3] = o->mgxy[l]; out6[4] = (int)o->mgn[l]; out6[5] = (int)o->mgp[l];
  return 0;
}

double *orc_grid(orc_t *o, int l, int axis) {
  if (l < 0 || l > o->L) return NULL;
  return axis == 0 ? o->mggx[l] : axis == 1 ? o->mggy[l] : o->mggz[l];
}

void *orc_array(orc_t *o, const char *n) {
  if (!strcmp(n, "nu")) return o->nu;
  if (!strcmp(n, "ro")) return o->ro;
  if (!strcmp(n, "nd")) return o->nd;
  if (!strcmp(n, "nxx")) return o->nxx;
  if (!strcmp(n, "nxy")) return o->nxy;
  if (!strcmp(n, "nxz")) return o->nxz;
  if (!strcmp(n, "nyz")) return o->nyz;
  if (!strcmp(n, "bondm")) return o->bondm;
  if (!strcmp(n, "bondv1")) return o->bondv1;
  if (!strcmp(n, "bondn1")) return o->bondn1;
  if (!strcmp(n, "sol0")) return o->sol0;
  if (!strcmp(n, "sol1")) return o->sol1;
  if (!strcmp(n, "fre0")) return o->fre0;
  if (!strcmp(n, "fre1")) return o->fre1;
  if (!strcmp(n, "val1")) return o->val1;
  if (!strcmp(n, "lin0")) return o->lin0;
  if (!strcmp(n, "num0")) return o->num0;
  if (!strcmp(n, "pos0")) return o->pos0;
  if (!strcmp(n, "nbond")) return &o->nbond;
  return NULL;
}

/* ------------------------------------------------------------------------- */
/* Boundary equations on every level: parallel/in3mg.c:234-390                */
/* var: 0 P, 1 vx, 2 vy, 3 vz.  Fine-grid index ranges are mapped to each     */
/* level keeping the 3 outer lines (in3mg.c:307-324); |shift|>1 = periodic    */
/* (in3mg.c:339-341).  Later equations overwrite bondm of earlier ones.      */
/* ------------------------------------------------------------------------- */
static int lvl_index(int m, int nfine, int nlev, long mgs) {
  int r = (m - 2) / (int)mgs + 2;
  if (m < 3) r = m;
  if (m > nfine - 4) r = nlev - (nfine - m);
  return r;
}

int orc_bc_line(orc_t *o, int var, int m10, int m11, int m20, int m21, int m30, int m31,
                float cnst, float koef, int dm1, int dm2, int dm3) {
  if (var < 0 || var > 3) return -1;
  for (int l = 0; l <= o->L; l++) {
    int m1min = lvl_index(m10, o->mgx[0], o->mgx[l], o->mgs[l]);
    int m1max = lvl_index(m11, o->mgx[0], o->mgx[l], o->mgs[l]);
    int m2min = lvl_index(m20, o->mgy[0], o->mgy[l], o->mgs[l]);
    int m2max = lvl_index(m21, o->mgy[0], o->mgy[l], o->mgs[l]);
    int m3min = lvl_index(m30, o->mgz[0], o->mgz[l], o->mgs[l]);
    int m3max = lvl_index(m31, o->mgz[0], o->mgz[l], o->mgs[l]);
    for (int m3 = m3min; m3 <= m3max; m3++)
      for (int m1 = m1min; m1 <= m1max; m1++)
        for (int m2 = m2min; m2 <= m2max; m2++) {
          long m4 = NID(o, l, m1, m2, m3);
          int sx = dm1, sy = dm2, sz = dm3;
          if (ABSV(dm1) > 1) sx = -(o->mgx[l] - 1);
          if (ABSV(dm2) > 1) sy = -(o->mgy[l] - 1);
          if (ABSV(dm3) > 1) sz = -(o->mgz[l] - 1);
          long m6 = m4 + (long)sz * o->mgxy[l] + (long)sx * o->mgy[l] + sy;
          if (o->nbond + 1 > o->capbond) {
            int nc = o->capbond * 2;
            o->bondv1 = (float *)realloc(o->bondv1, 2 * (size_t)(nc + 1) * sizeof(float));
            o->bondn1 = (int *)realloc(o->bondn1, (size_t)(nc + 1) * sizeof(int));
            o->capbond = nc;
          }
          long m5 = m4 * 4 + var;
          o->nbond++;
          o->bondm[m5] = o->nbond;
          o->bondv1[2 * o->nbond + 0] = cnst;
          if (dm1 || dm2 || dm3) {
            o->bondv1[2 * o->nbond + 1] = koef;
            o->bondn1[o->nbond] = (int)(m6 * 4 + var + 1);
          } else {
            o->bondv1[2 * o->nbond + 1] = 0;
            o->bondn1[o->nbond] = 0;
          }
        }
  }
  return o->nbond;
}

/* ------------------------------------------------------------------------- */
/* Marker -> node averaging: parallel/mark3mg.c:630-1132 (ronurecalc)         */
/* Marker material values are given per marker (the outputs of viscalc and   */
/* dencalc1, mark3mg.c:1418-1498).  Cell search follows m1serch/m2serch/      */
/* m3serch (mark3mg.c:1138-1205) via prcalc1 (mark3mg.c:1285-1304).           */
/* Periodic axes: the sums of the last grid line are added to line 0 after  */
/* the marker loop (mark3mg.c:893-955) and line 0's values are copied back   */
/* to the last line after normalising (mark3mg.c:1032-1111); each property  */
/* folds along the axes whose grid lines it sits on (nu, ro: all; nxy: x, y; */
/* nxz: x, z; nyz: y, z; nxx: none).                                          */
/* ------------------------------------------------------------------------- */
static long serch(const double *g, int n, double x) {
  long a = 0, b = n - 1, m;
  do {
    m = (a + b) / 2;
    if (g[m] > x) b = m; else a = m;
  } while ((b - a) > 1);
  if (a > n - 2) a = n - 2;
  return a;
}

int orc_ronurecalc(orc_t *o, long nmark, const float *mx, const float *my, const float *mz,
                   const float *mnu_in, const float *mro_in, double nukoef) {
  const int nx = o->mgx[0], ny = o->mgy[0], nz = o->mgz[0];
  const long nn = o->mgn[0];
  const double *gx = o->mggx[0], *gy = o->mggy[0], *gz = o->mggz[0];
  const double xs = gx[nx - 1], ys = gy[ny - 1], zs = gz[nz - 1];
  /* accumulators: FP32 property sums (Nu,Ro,Nxx.. are float**), FP64 weights */
  float *anu = (float *)calloc(nn, sizeof(float)), *aro = (float *)calloc(nn, sizeof(float));
  float *axx = (float *)calloc(nn, sizeof(float)), *axy = (float *)calloc(nn, sizeof(float));
  float *axz = (float *)calloc(nn, sizeof(float)), *ayz = (float *)calloc(nn, sizeof(float));
  double *w = (double *)calloc(6 * nn, sizeof(double));
  double *wro = w, *wnu = w + nn, *wxx = w + 2 * nn, *wxy = w + 3 * nn, *wxz = w + 4 * nn,
         *wyz = w + 5 * nn;
  const int vm = o->viscmod;
  for (long mm1 = 0; mm1 < nmark; mm1++) {
    if (!(mx[mm1] > 0 && my[mm1] > 0 && mz[mm1] > 0 && mx[mm1] < xs && my[mm1] < ys &&
          mz[mm1] < zs))
      continue;
    double x = mx[mm1], y = my[mm1], z = mz[mm1];
    long m10 = serch(gx, nx, x), m20 = serch(gy, ny, y), m30 = serch(gz, nz, z);
    double mnu = mnu_in[mm1];
    double mro = mro_in[mm1];
    if (mnu < o->nubeg) mnu = o->nubeg;
    if (mnu > o->nuend) mnu = o->nuend;
    double dx = gx[m10 + 1] - gx[m10], dy = gy[m20 + 1] - gy[m20], dz = gz[m30 + 1] - gz[m30];
    for (long m3 = m30; m3 <= m30 + 1; m3++) {
      double ddz = (gz[m3] - z) / dz;
      ddz = 1.0 - ABSV(ddz);
      for (long m1 = m10; m1 <= m10 + 1; m1++) {
        double ddx = (gx[m1] - x) / dx;
        ddx = 1.0 - ABSV(ddx);
        for (long m2 = m20; m2 <= m20 + 1; m2++) {
          double ddy = (gy[m2] - y) / dy;
          ddy = 1.0 - ABSV(ddy);
          double swt = ddx * ddy * ddz, swt1, ival2 = 0;
          long m4 = m3 * (long)nx * ny + m1 * ny + m2;
          if (ddx > nukoef && ddy > nukoef && ddz > nukoef) {
            if (vm == 0) ival2 = mnu * swt;
            if (vm == 1) ival2 = log(mnu) * swt;
            if (vm == 2) ival2 = 1.0 / mnu * swt;
            anu[m4] += ival2;
            wnu[m4] += swt;
          }
          if (m1 > m10 && m2 > m20 && m3 > m30) {
            swt1 = 8.0 * (0.5 - ABSV(ddx - 0.5)) * (0.5 - ABSV(ddy - 0.5)) * (0.5 - ABSV(ddz - 0.5));
            if (vm == 0) ival2 = mnu * swt1;
            if (vm == 1) ival2 = log(mnu) * swt1;
            if (vm == 2) ival2 = 1.0 / mnu * swt1;
            axx[m4] += ival2;
            wxx[m4] += swt1;
          }
          if (m3 == m30 && ddx > 0.5 && ddy > 0.5) {
            swt1 = 8.0 * (ddx - 0.5) * (ddy - 0.5) * (0.5 - ABSV(ddz - 0.5));
            if (vm == 0) ival2 = mnu * swt1;
            if (vm == 1) ival2 = log(mnu) * swt1;
            if (vm == 2) ival2 = 1.0 / mnu * swt1;
            axy[m4] += ival2;
            wxy[m4] += swt1;
          }
          if (m2 == m20 && ddx > 0.5 && ddz > 0.5) {
            swt1 = 8.0 * (ddx - 0.5) * (0.5 - ABSV(ddy - 0.5)) * (ddz - 0.5);
            if (vm == 0) ival2 = mnu * swt1;
            if (vm == 1) ival2 = log(mnu) * swt1;
            if (vm == 2) ival2 = 1.0 / mnu * swt1;
            axz[m4] += ival2;
            wxz[m4] += swt1;
          }
          if (m1 == m10 && ddy > 0.5 && ddz > 0.5) {
            swt1 = 8.0 * (0.5 - ABSV(ddx - 0.5)) * (ddy - 0.5) * (ddz - 0.5);
            if (vm == 0) ival2 = mnu * swt1;
            if (vm == 1) ival2 = log(mnu) * swt1;
            if (vm == 2) ival2 = 1.0 / mnu * swt1;
            ayz[m4] += ival2;
            wyz[m4] += swt1;
          }
          if (ddx > 0.5 && ddy > 0.5 && ddz > 0.5) {
            swt1 = 8.0 * (ddx - 0.5) * (ddy - 0.5) * (ddz - 0.5);
            aro[m4] += mro * swt1;
            wro[m4] += swt1;
          }
        }
      }
    }
  }
  /* periodic folds, in the reference's order x, y, z (mark3mg.c:893-955) */
  const long nxy = (long)nx * ny;
  if (o->xper)
    for (long m3 = 0; m3 < nz; m3++)
      for (long m2 = 0; m2 < ny; m2++) {
        long a = m3 * nxy + m2, b = a + (long)(nx - 1) * ny;
        wro[a] += wro[b]; wnu[a] += wnu[b]; wxy[a] += wxy[b]; wxz[a] += wxz[b];
        aro[a] += aro[b]; anu[a] += anu[b]; axy[a] += axy[b]; axz[a] += axz[b];
      }
  if (o->yper)
    for (long m3 = 0; m3 < nz; m3++)
      for (long m1 = 0; m1 < nx; m1++) {
        long a = m3 * nxy + m1 * ny, b = a + ny - 1;
        wro[a] += wro[b]; wnu[a] += wnu[b]; wxy[a] += wxy[b]; wyz[a] += wyz[b];
        aro[a] += aro[b]; anu[a] += anu[b]; axy[a] += axy[b]; ayz[a] += ayz[b];
      }
  if (o->zper)
    for (long m1 = 0; m1 < nx; m1++)
      for (long m2 = 0; m2 < ny; m2++) {
        long a = m1 * ny + m2, b = a + (long)(nz - 1) * nxy;
        wro[a] += wro[b]; wnu[a] += wnu[b]; wxz[a] += wxz[b]; wyz[a] += wyz[b];
        aro[a] += aro[b]; anu[a] += anu[b]; axz[a] += axz[b]; ayz[a] += ayz[b];
      }
  /* normalise, back-transform, clamp: mark3mg.c:962-1031 */
  for (long m3 = 0; m3 < nn; m3++) {
    if (wnu[m3] > 0) {
      o->nu[m3] = anu[m3] / wnu[m3];
      if (vm == 1) o->nu[m3] = exp(o->nu[m3]);
      if (vm == 2) o->nu[m3] = 1.0 / o->nu[m3];
    }
    if (wxx[m3] > 0) {
      o->nxx[m3] = axx[m3] / wxx[m3];
      if (vm == 1) o->nxx[m3] = exp(o->nxx[m3]);
      if (vm == 2) o->nxx[m3] = 1.0 / o->nxx[m3];
    }
    if (wxy[m3] > 0) {
      o->nxy[m3] = axy[m3] / wxy[m3];
      if (vm == 1) o->nxy[m3] = exp(o->nxy[m3]);
      if (vm == 2) o->nxy[m3] = 1.0 / o->nxy[m3];
    }
    if (wxz[m3] > 0) {
      o->nxz[m3] = axz[m3] / wxz[m3];
      if (vm == 1) o->nxz[m3] = exp(o->nxz[m3]);
      if (vm == 2) o->nxz[m3] = 1.0 / o->nxz[m3];
    }
    if (wyz[m3] > 0) {
      o->nyz[m3] = ayz[m3] / wyz[m3];
      if (vm == 1) o->nyz[m3] = exp(o->nyz[m3]);
      if (vm == 2) o->nyz[m3] = 1.0 / o->nyz[m3];
    }
    if (wro[m3] > 0) o->ro[m3] = aro[m3] / wro[m3];
    if (o->nu[m3] < o->nubeg) o->nu[m3] = o->nubeg;
    if (o->nu[m3] > o->nuend) o->nu[m3] = o->nuend;
    if (o->nxx[m3] < o->nubeg) o->nxx[m3] = o->nubeg;
    if (o->nxx[m3] > o->nuend) o->nxx[m3] = o->nuend;
    if (o->nxy[m3] < o->nubeg) o->nxy[m3] = o->nubeg;
    if (o->nxy[m3] > o->nuend) o->nxy[m3] = o->nuend;
    if (o->nxz[m3] < o->nubeg) o->nxz[m3] = o->nubeg;
    if (o->nxz[m3] > o->nuend) o->nxz[m3] = o->nuend;
    if (o->nyz[m3] < o->nubeg) o->nyz[m3] = o->nubeg;
    if (o->nyz[m3] > o->nuend) o->nyz[m3] = o->nuend;
  }
  /* the last line of a periodic axis takes line 0's values (mark3mg.c:1032-1111) */
  if (o->xper)
    for (long m3 = 0; m3 < nz; m3++)
      for (long m2 = 0; m2 < ny; m2++) {
        long a = m3 * nxy + m2, b = a + (long)(nx - 1) * ny;
        o->ro[b] = o->ro[a]; o->nu[b] = o->nu[a]; o->nxy[b] = o->nxy[a]; o->nxz[b] = o->nxz[a];
      }
  if (o->yper)
    for (long m3 = 0; m3 < nz; m3++)
      for (long m1 = 0; m1 < nx; m1++) {
        long a = m3 * nxy + m1 * ny, b = a + ny - 1;
        o->ro[b] = o->ro[a]; o->nu[b] = o->nu[a]; o->nxy[b] = o->nxy[a]; o->nyz[b] = o->nyz[a];
      }
  if (o->zper)
    for (long m1 = 0; m1 < nx; m1++)
      for (long m2 = 0; m2 < ny; m2++) {
        long a = m1 * ny + m2, b = a + (long)(nz - 1) * nxy;
        o->ro[b] = o->ro[a]; o->nu[b] = o->nu[a]; o->nxz[b] = o->nxz[a]; o->nyz[b] = o->nyz[a];
      }
  if (o->xper && o->zper)
    for (long m2 = 0; m2 < ny; m2++) {
      const long b3[3] = {m2 + (long)(nx - 1) * ny, m2 + (long)(nz - 1) * nxy, m2 + (long)(nz - 1) * nxy + (long)(nx - 1) * ny};
      for (int k = 0; k < 3; k++) { o->ro[b3[k]] = o->ro[m2]; o->nu[b3[k]] = o->nu[m2]; o->nxz[b3[k]] = o->nxz[m2]; }
    }
  free(anu); free(aro); free(axx); free(axy); free(axz); free(ayz); free(w);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Viscosity/density restriction: parallel/move3mg.c:3387-3528                */
/* ------------------------------------------------------------------------- */
void orc_ronurestrict(orc_t *o, int l) {
  const int f = l - 1;
  const long nc = o->mgn[l], pc = o->mgp[l];
  float *Nu = (float *)calloc(nc, sizeof(float)), *Ro = (float *)calloc(nc, sizeof(float));
  double *S1 = (double *)calloc(nc, sizeof(double));
  const double *cx = o->mggx[l], *cy = o->mggy[l], *cz = o->mggz[l];
  const double *fx = o->mggx[f], *fy = o->mggy[f], *fz = o->mggz[f];
  for (int m3 = 0; m3 < o->mgz[f]; m3++)
    for (int m1 = 0; m1 < o->mgx[f]; m1++)
      for (int m2 = 0; m2 < o->mgy[f]; m2++) {
        long m4 = NID(o, f, m1, m2, m3);
        int m10min = (m1 - 2) / 2 + 2;
        if (m1 < 3) m10min = m1;
        if (m1 > o->mgx[f] - 4) m10min = o->mgx[l] - (o->mgx[f] - m1);
        if (m10min > o->mgx[l] - 2) m10min = o->mgx[l] - 2;
        int m20min = (m2 - 2) / 2 + 2;
        if (m2 < 3) m20min = m2;
        if (m2 > o->mgy[f] - 4) m20min = o->mgy[l] - (o->mgy[f] - m2);
        if (m20min > o->mgy[l] - 2) m20min = o->mgy[l] - 2;
        int m30min = (m3 - 2) / 2 + 2;
        if (m3 < 3) m30min = m3;
        if (m3 > o->mgz[f] - 4) m30min = o->mgz[l] - (o->mgz[f] - m3);
        if (m30min > o->mgz[l] - 2) m30min = o->mgz[l] - 2;
        double dx = cx[m10min + 1] - cx[m10min];
        double dy = cy[m20min + 1] - cy[m20min];
        double dz = cz[m30min + 1] - cz[m30min];
        for (int m30 = m30min; m30 <= m30min + 1; m30++) {
          double ddz = (cz[m30] - fz[m3]) / dz;
          ddz = 1.0 - ABSV(ddz);
          for (int m10 = m10min; m10 <= m10min + 1; m10++) {
            double ddx = (cx[m10] - fx[m1]) / dx;
            ddx = 1.0 - ABSV(ddx);
            for (int m20 = m20min; m20 <= m20min + 1; m20++) {
              double ddy = (cy[m20] - fy[m2]) / dy;
              ddy = 1.0 - ABSV(ddy);
              long m40 = NID(o, l, m10, m20, m30) - pc;
              double swt = ddx * ddy * ddz, ival2 = 0;
              if (o->viscmod == 0) ival2 = o->nu[m4] * swt;
              if (o->viscmod == 1) ival2 = log(o->nu[m4]) * swt;
              if (o->viscmod == 2) ival2 = 1.0 / o->nu[m4] * swt;
              Nu[m40] += ival2;
              ival2 = o->ro[m4] * swt;
              Ro[m40] += ival2;
              S1[m40] += swt;
            }
          }
        }
      }
  for (long m = 0; m < nc; m++) {
    o->nu[pc + m] = 0;
    o->ro[pc + m] = 0;
    o->nu[pc + m] += Nu[m];
    o->ro[pc + m] += Ro[m];
    o->nu[pc + m] /= S1[m];
    if (o->viscmod == 1) o->nu[pc + m] = exp(o->nu[pc + m]);
    if (o->viscmod == 2) o->nu[pc + m] = 1.0 / o->nu[pc + m];
    o->ro[pc + m] /= S1[m];
  }
  free(Nu); free(Ro); free(S1);
}

/* ------------------------------------------------------------------------- */
/* Stress-point stencils, coefficient mode: parallel/move3mg.c:3-788          */
/* Each appends (index, coefficient) pairs at un1[un1[0]+1..] and advances    */
/* un1[0]; ynval weights the contribution.                                   */
/* ------------------------------------------------------------------------- */
static double cellvisc(const orc_t *o, const long *v) {
  const float *nu = o->nu;
  double e = 0;
  if (o->viscmod == 0) e = (nu[v[0]] + nu[v[1]] + nu[v[2]] + nu[v[3]] + nu[v[4]] + nu[v[5]] + nu[v[6]] + nu[v[7]]) / 8.0;
  if (o->viscmod == 1)
    e = exp((log(nu[v[0]]) + log(nu[v[1]]) + log(nu[v[2]]) + log(nu[v[3]]) + log(nu[v[4]]) + log(nu[v[5]]) + log(nu[v[6]]) + log(nu[v[7]])) / 8.0);
  if (o->viscmod == 2)
    e = 1.0 / ((1.0 / nu[v[0]] + 1.0 / nu[v[1]] + 1.0 / nu[v[2]] + 1.0 / nu[v[3]] + 1.0 / nu[v[4]] + 1.0 / nu[v[5]] + 1.0 / nu[v[6]] + 1.0 / nu[v[7]]) / 8.0);
  return e;
}

static double edgevisc(const orc_t *o, long a, long b) {
  const float *nu = o->nu;
  double e = 0;
  if (o->viscmod == 0) e = (nu[a] + nu[b]) / 2.0;
  if (o->viscmod == 1) e = exp((log(nu[a]) + log(nu[b])) / 2.0);
  if (o->viscmod == 2) e = 1.0 / ((1.0 / nu[a] + 1.0 / nu[b]) / 2.0);
  return e;
}

static double clampnu(const orc_t *o, double e) {
  if (e > o->nuend) e = o->nuend;
  if (e < o->nubeg) e = o->nubeg;
  return e;
}

/* cell node octet around cell (m1,m2,m3), with periodic wraps (move3mg.c:27-34) */
static void cell_octet(const orc_t *o, int l, long m1, long m2, long m3, long *v) {
  v[0] = o->mgp[l] + (m3 - 1) * o->mgxy[l] + (m1 - 1) * o->mgy[l] + (m2 - 1);
  v[1] = v[0] + 1;
  v[2] = v[0] + o->mgy[l];
  v[3] = v[2] + 1;
  v[4] = v[0] + o->mgxy[l];
  v[5] = v[4] + 1;
  v[6] = v[4] + o->mgy[l];
  v[7] = v[6] + 1;
  if (o->xper && m1 == o->mgx[l] - 1) v[2] = o->mgp[l] + (m3 - 1) * o->mgxy[l] + m2 - 1;
  if (o->yper && m2 == o->mgy[l] - 1) v[1] = o->mgp[l] + (m3 - 1) * o->mgxy[l] + (m1 - 1) * o->mgy[l];
  if (o->zper && m3 == o->mgz[l] - 1) v[4] = o->mgp[l] + (m1 - 1) * o->mgy[l] + m2 - 1;
}

/* Sxx/Syy/Szz, move3mg.c:3-380 (axis 0/1/2 selects which strain leads) */
static void snn_coef(const orc_t *o, int axis, long m1, long m2, long m3, double w, int l,
                     long *un1, double *ui1) {
  double xkf = 1.0 / (o->mggx[l][m1] - o->mggx[l][m1 - 1]);
  double ykf = 1.0 / (o->mggy[l][m2] - o->mggy[l][m2 - 1]);
  double zkf = 1.0 / (o->mggz[l][m3] - o->mggz[l][m3 - 1]);
  long v[8];
  cell_octet(o, l, m1, m2, m3, v);
  double ne = (l == 0) ? (double)o->nxx[v[7]] : cellvisc(o, v);
  ne = clampnu(o, ne);
  long k = un1[0];
  /* the reference writes the signs literally per stencil; keep its factor order */
  if (axis == 0) {
    un1[k + 1] = v[0] * 4 + 1; ui1[k + 1] = -w * ne * 4.0 / 3.0 * xkf;
    un1[k + 2] = v[2] * 4 + 1; ui1[k + 2] = +w * ne * 4.0 / 3.0 * xkf;
    un1[k + 3] = v[0] * 4 + 2; ui1[k + 3] = +w * ne * 2.0 / 3.0 * ykf;
    un1[k + 4] = v[1] * 4 + 2; ui1[k + 4] = -w * ne * 2.0 / 3.0 * ykf;
    un1[k + 5] = v[0] * 4 + 3; ui1[k + 5] = +w * ne * 2.0 / 3.0 * zkf;
    un1[k + 6] = v[4] * 4 + 3; ui1[k + 6] = -w * ne * 2.0 / 3.0 * zkf;
  } else if (axis == 1) {
    un1[k + 1] = v[0] * 4 + 1; ui1[k + 1] = +w * ne * 2.0 / 3.0 * xkf;
    un1[k + 2] = v[2] * 4 + 1; ui1[k + 2] = -w * ne * 2.0 / 3.0 * xkf;
    un1[k + 3] = v[0] * 4 + 2; ui1[k + 3] = -w * ne * 4.0 / 3.0 * ykf;
    un1[k + 4] = v[1] * 4 + 2; ui1[k + 4] = +w * ne * 4.0 / 3.0 * ykf;
    un1[k + 5] = v[0] * 4 + 3; ui1[k + 5] = +w * ne * 2.0 / 3.0 * zkf;
    un1[k + 6] = v[4] * 4 + 3; ui1[k + 6] = -w * ne * 2.0 / 3.0 * zkf;
  } else {
    un1[k + 1] = v[0] * 4 + 1; ui1[k + 1] = +w * ne * 2.0 / 3.0 * xkf;
    un1[k + 2] = v[2] * 4 + 1; ui1[k + 2] = -w * ne * 2.0 / 3.0 * xkf;
    un1[k + 3] = v[0] * 4 + 2; ui1[k + 3] = +w * ne * 2.0 / 3.0 * ykf;
    un1[k + 4] = v[1] * 4 + 2; ui1[k + 4] = -w * ne * 2.0 / 3.0 * ykf;
    un1[k + 5] = v[0] * 4 + 3; ui1[k + 5] = -w * ne * 4.0 / 3.0 * zkf;
    un1[k + 6] = v[4] * 4 + 3; ui1[k + 6] = +w * ne * 4.0 / 3.0 * zkf;
  }
  un1[0] += 6;
}

/* Sxy, move3mg.c:388-512 */
static void sxy_coef(const orc_t *o, long m1, long m2, long m3, double w, int l, long *un1,
                     double *ui1) {
  const int X = o->mgx[l], Y = o->mgy[l];
  const double *gx = o->mggx[l], *gy = o->mggy[l];
  double xkf, ykf;
  if (o->xper && m1 == 0) xkf = 1.0 / (gx[1] - gx[0] + gx[X - 1] - gx[X - 2]);
  else xkf = 1.0 / (gx[m1 + 1] - gx[m1 - 1]);
  if (o->yper && m2 == 0) ykf = 1.0 / (gy[1] - gy[0] + gy[Y - 1] - gy[Y - 2]);
  else ykf = 1.0 / (gy[m2 + 1] - gy[m2 - 1]);
  long p = o->mgp[l], xy = o->mgxy[l];
  long v[8];
  v[0] = p + m3 * xy + (m1 - 1) * Y + (m2 - 1);
  v[1] = v[0] + 1; v[2] = v[0] + Y; v[3] = v[2] + 1;
  v[4] = v[0] + xy; v[5] = v[4] + 1; v[6] = v[4] + Y; v[7] = v[6] + 1;
  if (o->xper && m1 == 0) {
    v[1] = p + m3 * xy + (X - 2) * Y + m2;
    v[2] = p + m3 * xy + (m2 - 1); v[3] = v[2] + 1; v[7] = v[3] + xy;
  }
  if (o->yper && m2 == 0) {
    v[2] = p + m3 * xy + m1 * Y + Y - 2;
    v[1] = p + m3 * xy + (m1 - 1) * Y; v[3] = v[1] + Y; v[7] = v[3] + xy;
  }
  if (o->xper && m1 == 0 && o->yper && m2 == 0) {
    v[1] = p + m3 * xy + (X - 2) * Y;
    v[2] = p + m3 * xy + Y - 2; v[3] = p + m3 * xy; v[7] = v[3] + xy;
  }
  double ne = (l == 0) ? (double)o->nxy[v[3]] : edgevisc(o, v[3], v[7]);
  ne = clampnu(o, ne);
  long k = un1[0];
  un1[k + 1] = v[2] * 4 + 1; ui1[k + 1] = -w * ne * 2.0 * ykf;
  un1[k + 2] = v[3] * 4 + 1; ui1[k + 2] = +w * ne * 2.0 * ykf;
  un1[k + 3] = v[1] * 4 + 2; ui1[k + 3] = -w * ne * 2.0 * xkf;
  un1[k + 4] = v[3] * 4 + 2; ui1[k + 4] = +w * ne * 2.0 * xkf;
  un1[0] += 4;
}

/* Sxz, move3mg.c:520-651 */
static void sxz_coef(const orc_t *o, long m1, long m2, long m3, double w, int l, long *un1,
                     double *ui1) {
  const int X = o->mgx[l], Y = o->mgy[l], Z = o->mgz[l];
  const double *gx = o->mggx[l], *gz = o->mggz[l];
  double xkf, zkf;
  if (o->xper && m1 == 0) xkf = 1.0 / (gx[1] - gx[0] + gx[X - 1] - gx[X - 2]);
  else xkf = 1.0 / (gx[m1 + 1] - gx[m1 - 1]);
  if (o->zper && m3 == 0) zkf = 1.0 / (gz[1] - gz[0] + gz[Z - 1] - gz[Z - 2]);
  else zkf = 1.0 / (gz[m3 + 1] - gz[m3 - 1]);
  long p = o->mgp[l], xy = o->mgxy[l];
  long v[8];
  v[0] = p + (m3 - 1) * xy + (m1 - 1) * Y + m2;
  v[1] = v[0] + 1; v[2] = v[0] + Y; v[3] = v[2] + 1;
  v[4] = v[0] + xy; v[5] = v[4] + 1; v[6] = v[4] + Y; v[7] = v[6] + 1;
  if (o->xper && m1 == 0) {
    v[4] = p + m3 * xy + (X - 2) * Y + m2;
    v[2] = p + (m3 - 1) * xy + m2; v[6] = v[2] + xy; v[7] = v[6] + 1;
  }
  if (o->zper && m3 == 0) {
    v[2] = p + (Z - 2) * xy + m1 * Y + m2;
    v[4] = p + (m1 - 1) * Y + m2; v[6] = v[4] + Y; v[7] = v[6] + 1;
  }
  if (o->xper && m1 == 0 && o->zper && m3 == 0) {
    v[2] = p + (Z - 2) * xy + m2;
    v[4] = p + (X - 2) * Y + m2;
    v[6] = p + m2; v[7] = v[6] + 1;
  }
  double ne = (l == 0) ? (double)o->nxz[v[6]] : edgevisc(o, v[6], v[7]);
  ne = clampnu(o, ne);
  long k = un1[0];
  un1[k + 1] = v[2] * 4 + 1; ui1[k + 1] = -w * ne * 2.0 * zkf;
  un1[k + 2] = v[6] * 4 + 1; ui1[k + 2] = +w * ne * 2.0 * zkf;
  un1[k + 3] = v[4] * 4 + 3; ui1[k + 3] = -w * ne * 2.0 * xkf;
  un1[k + 4] = v[6] * 4 + 3; ui1[k + 4] = +w * ne * 2.0 * xkf;
  un1[0] += 4;
}

/* Syz, move3mg.c:659-788 */
static void syz_coef(const orc_t *o, long m1, long m2, long m3, double w, int l, long *un1,
                     double *ui1) {
  const int Y = o->mgy[l], Z = o->mgz[l];
  const double *gy = o->mggy[l], *gz = o->mggz[l];
  double ykf, zkf;
  if (o->yper && m2 == 0) ykf = 1.0 / (gy[1] - gy[0] + gy[Y - 1] - gy[Y - 2]);
  else ykf = 1.0 / (gy[m2 + 1] - gy[m2 - 1]);
  if (o->zper && m3 == 0) zkf = 1.0 / (gz[1] - gz[0] + gz[Z - 1] - gz[Z - 2]);
  else zkf = 1.0 / (gz[m3 + 1] - gz[m3 - 1]);
  long p = o->mgp[l], xy = o->mgxy[l];
  long v[8];
  v[0] = p + (m3 - 1) * xy + m1 * Y + (m2 - 1);
  v[1] = v[0] + 1; v[2] = v[0] + Y; v[3] = v[2] + 1;
  v[4] = v[0] + xy; v[5] = v[4] + 1; v[6] = v[4] + Y; v[7] = v[6] + 1;
  if (o->yper && m2 == 0) {
    v[1] = p + (m3 - 1) * xy + m1 * Y; v[5] = v[1] + xy; v[7] = v[5] + Y;
    v[4] = v[5] + Y - 2;
  }
  if (o->zper && m3 == 0) {
    v[4] = p + m1 * Y + (m2 - 1); v[5] = v[4] + 1; v[7] = v[5] + Y;
    v[1] = v[5] + (Z - 2) * xy;
  }
  if (o->yper && m2 == 0 && o->zper && m3 == 0) {
    v[5] = p + m1 * Y; v[7] = v[5] + Y;
    v[1] = v[5] + (Z - 2) * xy;
    v[4] = v[5] + Y - 2;
  }
  double ne = (l == 0) ? (double)o->nyz[v[5]] : edgevisc(o, v[5], v[7]);
  ne = clampnu(o, ne);
  long k = un1[0];
  un1[k + 1] = v[1] * 4 + 2; ui1[k + 1] = -w * ne * 2.0 * zkf;
  un1[k + 2] = v[5] * 4 + 2; ui1[k + 2] = +w * ne * 2.0 * zkf;
  un1[k + 3] = v[4] * 4 + 3; ui1[k + 3] = -w * ne * 2.0 * ykf;
  un1[k + 4] = v[5] * 4 + 3; ui1[k + 4] = +w * ne * 2.0 * ykf;
  un1[0] += 4;
}

/* grad-div contribution, move3mg.c:1580-1674.  Its level-0 nxx branch is    */
/* dead (mgi==-10), so the 8-node nu average is used on every level.         */
static void continadd_coef(const orc_t *o, long m1, long m2, long m3, double w, int l,
                           long *un1, double *ui1) {
  double xkf = 1.0 / (o->mggx[l][m1] - o->mggx[l][m1 - 1]);
  double ykf = 1.0 / (o->mggy[l][m2] - o->mggy[l][m2 - 1]);
  double zkf = 1.0 / (o->mggz[l][m3] - o->mggz[l][m3 - 1]);
  long v[8];
  cell_octet(o, l, m1, m2, m3, v);
  double ne = clampnu(o, cellvisc(o, v));
  long k = un1[0];
  ui1[0] += 0.0 * w; /* rightc = 0 */
  un1[k + 1] = v[0] * 4 + 1; ui1[k + 1] = -xkf * w * ne;
  un1[k + 2] = v[2] * 4 + 1; ui1[k + 2] = +xkf * w * ne;
  un1[k + 3] = v[0] * 4 + 2; ui1[k + 3] = -ykf * w * ne;
  un1[k + 4] = v[1] * 4 + 2; ui1[k + 4] = +ykf * w * ne;
  un1[k + 5] = v[0] * 4 + 3; ui1[k + 5] = -zkf * w * ne;
  un1[k + 6] = v[4] * 4 + 3; ui1[k + 6] = +zkf * w * ne;
  un1[0] += 6;
}

/* ------------------------------------------------------------------------- */
/* Momentum rows: x move3mg.c:796-1003, y 1011-1223, z 1231-1447              */
/* Slot layout (1..17) and merge order follow the reference exactly: the     */
/* assembled ELL rows and the coarsest CSR depend on it.                     */
/* ------------------------------------------------------------------------- */
#define SETS(s, k) (un1[s] = un1[k], ui1[s] = ui1[k])
#define ADDS(s, k) (ui1[s] += ui1[k])

static void xstokes_row(const orc_t *o, long m1, long m2, long m3, int l, long *un1, double *ui1) {
  const int X = o->mgx[l], Y = o->mgy[l], Z = o->mgz[l];
  const double *gx = o->mggx[l], *gy = o->mggy[l], *gz = o->mggz[l];
  double xkf = 2.0 / (gx[m1 + 1] - gx[m1 - 1]);
  if (o->xper && m1 == 0) xkf = 2.0 / (gx[1] - gx[0] + gx[X - 1] - gx[X - 2]);
  double ykf = 1.0 / (gy[m2 + 1] - gy[m2]);
  double zkf = 1.0 / (gz[m3 + 1] - gz[m3]);
  long v[8];
  v[0] = o->mgp[l] + m3 * o->mgxy[l] + m1 * Y + m2;
  v[1] = v[0] + 1; v[2] = v[0] + Y; v[3] = v[2] + 1;
  v[4] = v[0] + o->mgxy[l]; v[5] = v[4] + 1; v[6] = v[4] + Y; v[7] = v[6] + 1;
  if (o->xper && m1 == 0) v[5] = o->mgp[l] + (m3 + 1) * o->mgxy[l] + (X - 1) * Y + m2 + 1;
  const float *ro = o->ro;
  ui1[0] = -o->gxk * (ro[v[0]] + ro[v[1]] + ro[v[4]] + ro[v[5]]) / 4.0;
  un1[0] = 17;
  un1[16] = v[5] * 4 + 0; ui1[16] = +xkf;
  un1[17] = v[7] * 4 + 0; ui1[17] = -xkf;
  if (o->xper && m1 == 0) snn_coef(o, 0, X - 1, m2 + 1, m3 + 1, -xkf, l, un1, ui1);
  else snn_coef(o, 0, m1, m2 + 1, m3 + 1, -xkf, l, un1, ui1);
  SETS(2, 18); SETS(1, 19); SETS(8, 20); SETS(9, 21); SETS(12, 22); SETS(13, 23);
  un1[0] = 17;
  snn_coef(o, 0, m1 + 1, m2 + 1, m3 + 1, xkf, l, un1, ui1);
  ADDS(1, 18); SETS(3, 19); SETS(10, 20); SETS(11, 21); SETS(14, 22); SETS(15, 23);
  un1[0] = 17;
  sxy_coef(o, m1, m2, m3, -ykf, l, un1, ui1);
  ADDS(1, 19); SETS(4, 18); ADDS(8, 20); ADDS(10, 21);
  un1[0] = 17;
  if (o->yper && m2 == Y - 2) sxy_coef(o, m1, 0, m3, ykf, l, un1, ui1);
  else sxy_coef(o, m1, m2 + 1, m3, ykf, l, un1, ui1);
  ADDS(1, 18); SETS(5, 19); ADDS(9, 20); ADDS(11, 21);
  un1[0] = 17;
  sxz_coef(o, m1, m2, m3, -zkf, l, un1, ui1);
  ADDS(1, 19); SETS(6, 18); ADDS(12, 20); ADDS(14, 21);
  un1[0] = 17;
  if (o->zper && m3 == Z - 2) sxz_coef(o, m1, m2, 0, zkf, l, un1, ui1);
  else sxz_coef(o, m1, m2, m3 + 1, zkf, l, un1, ui1);
  ADDS(1, 18); SETS(7, 19); ADDS(13, 20); ADDS(15, 21);
  if (o->p2koef != 0) {
    un1[0] = 17;
    if (o->xper && m1 == 0) continadd_coef(o, X - 1, m2 + 1, m3 + 1, -xkf * o->p2koef, l, un1, ui1);
    else continadd_coef(o, m1, m2 + 1, m3 + 1, -xkf * o->p2koef, l, un1, ui1);
    ADDS(2, 18); ADDS(1, 19); ADDS(8, 20); ADDS(9, 21); ADDS(12, 22); ADDS(13, 23);
    un1[0] = 17;
    continadd_coef(o, m1 + 1, m2 + 1, m3 + 1, xkf * o->p2koef, l, un1, ui1);
    ADDS(1, 18); ADDS(3, 19); ADDS(10, 20); ADDS(11, 21); ADDS(14, 22); ADDS(15, 23);
  }
  un1[0] = 17;
}

static void ystokes_row(const orc_t *o, long m1, long m2, long m3, int l, long *un1, double *ui1) {
  const int X = o->mgx[l], Y = o->mgy[l], Z = o->mgz[l];
  const double *gx = o->mggx[l], *gy = o->mggy[l], *gz = o->mggz[l];
  double xkf = 1.0 / (gx[m1 + 1] - gx[m1]);
  double ykf = 2.0 / (gy[m2 + 1] - gy[m2 - 1]);
  if (o->yper && m2 == 0) ykf = 2.0 / (gy[1] - gy[0] + gy[Y - 1] - gy[Y - 2]);
  double zkf = 1.0 / (gz[m3 + 1] - gz[m3]);
  long v[8];
  v[0] = o->mgp[l] + m3 * o->mgxy[l] + m1 * Y + m2;
  v[1] = v[0] + 1; v[2] = v[0] + Y; v[3] = v[2] + 1;
  v[4] = v[0] + o->mgxy[l]; v[5] = v[4] + 1; v[6] = v[4] + Y; v[7] = v[6] + 1;
  if (o->yper && m2 == 0) v[6] = o->mgp[l] + (m3 + 1) * o->mgxy[l] + (m1 + 1) * Y + Y - 1;
  const float *ro = o->ro;
  ui1[0] = -o->gyk * (ro[v[0]] + ro[v[2]] + ro[v[4]] + ro[v[6]]) / 4.0;
  un1[0] = 17;
  un1[16] = v[6] * 4 + 0; ui1[16] = +ykf;
  un1[17] = v[7] * 4 + 0; ui1[17] = -ykf;
  if (o->yper && m2 == 0) snn_coef(o, 1, m1 + 1, Y - 1, m3 + 1, -ykf, l, un1, ui1);
  else snn_coef(o, 1, m1 + 1, m2, m3 + 1, -ykf, l, un1, ui1);
  SETS(8, 18); SETS(10, 19); SETS(4, 20); SETS(1, 21); SETS(12, 22); SETS(14, 23);
  un1[0] = 17;
  snn_coef(o, 1, m1 + 1, m2 + 1, m3 + 1, ykf, l, un1, ui1);
  SETS(9, 18); SETS(11, 19); ADDS(1, 20); SETS(5, 21); SETS(13, 22); SETS(15, 23);
  un1[0] = 17;
  sxy_coef(o, m1, m2, m3, -xkf, l, un1, ui1);
  ADDS(8, 18); ADDS(9, 19); SETS(2, 20); ADDS(1, 21);
  un1[0] = 17;
  if (o->xper && m1 == X - 2) sxy_coef(o, 0, m2, m3, xkf, l, un1, ui1);
  else sxy_coef(o, m1 + 1, m2, m3, xkf, l, un1, ui1);
  ADDS(10, 18); ADDS(11, 19); ADDS(1, 20); SETS(3, 21);
  un1[0] = 17;
  syz_coef(o, m1, m2, m3, -zkf, l, un1, ui1);
  SETS(6, 18); ADDS(1, 19); ADDS(12, 20); ADDS(13, 21);
  un1[0] = 17;
  if (o->zper && m3 == Z - 2) syz_coef(o, m1, m2, 0, zkf, l, un1, ui1);
  else syz_coef(o, m1, m2, m3 + 1, zkf, l, un1, ui1);
  ADDS(1, 18); SETS(7, 19); ADDS(14, 20); ADDS(15, 21);
  if (o->p2koef != 0) {
    un1[0] = 17;
    if (o->yper && m2 == 0) continadd_coef(o, m1 + 1, Y - 1, m3 + 1, -ykf * o->p2koef, l, un1, ui1);
    else continadd_coef(o, m1 + 1, m2, m3 + 1, -ykf * o->p2koef, l, un1, ui1);
    ADDS(8, 18); ADDS(10, 19); ADDS(4, 20); ADDS(1, 21); ADDS(12, 22); ADDS(14, 23);
    un1[0] = 17;
    continadd_coef(o, m1 + 1, m2 + 1, m3 + 1, ykf * o->p2koef, l, un1, ui1);
    ADDS(9, 18); ADDS(11, 19); ADDS(1, 20); ADDS(5, 21); ADDS(13, 22); ADDS(15, 23);
  }
  un1[0] = 17;
}

static void zstokes_row(const orc_t *o, long m1, long m2, long m3, int l, long *un1, double *ui1) {
  const int X = o->mgx[l], Y = o->mgy[l], Z = o->mgz[l];
  const double *gx = o->mggx[l], *gy = o->mggy[l], *gz = o->mggz[l];
  double xkf = 1.0 / (gx[m1 + 1] - gx[m1]);
  double ykf = 1.0 / (gy[m2 + 1] - gy[m2]);
  double zkf = 2.0 / (gz[m3 + 1] - gz[m3 - 1]);
  if (o->zper && m3 == 0) zkf = 2.0 / (gz[1] - gz[0] + gz[Z - 1] - gz[Z - 2]);
  long v[8];
  v[0] = o->mgp[l] + m3 * o->mgxy[l] + m1 * Y + m2;
  v[1] = v[0] + 1; v[2] = v[0] + Y; v[3] = v[2] + 1;
  v[4] = v[0] + o->mgxy[l]; v[5] = v[4] + 1; v[6] = v[4] + Y; v[7] = v[6] + 1;
  if (o->zper && m3 == 0) v[3] = o->mgp[l] + (Z - 1) * o->mgxy[l] + (m1 + 1) * Y + m2 + 1;
  const float *ro = o->ro;
  ui1[0] = -o->gzk * (ro[v[0]] + ro[v[1]] + ro[v[2]] + ro[v[3]]) / 4.0;
  un1[0] = 17;
  un1[16] = v[3] * 4 + 0; ui1[16] = +zkf;
  un1[17] = v[7] * 4 + 0; ui1[17] = -zkf;
  if (o->zper && m3 == 0) snn_coef(o, 2, m1 + 1, m2 + 1, Z - 1, -zkf, l, un1, ui1);
  else snn_coef(o, 2, m1 + 1, m2 + 1, m3, -zkf, l, un1, ui1);
  SETS(12, 18); SETS(14, 19); SETS(8, 20); SETS(9, 21); SETS(6, 22); SETS(1, 23);
  un1[0] = 17;
  snn_coef(o, 2, m1 + 1, m2 + 1, m3 + 1, zkf, l, un1, ui1);
  SETS(13, 18); SETS(15, 19); SETS(10, 20); SETS(11, 21); ADDS(1, 22); SETS(7, 23);
  un1[0] = 17;
  sxz_coef(o, m1, m2, m3, -xkf, l, un1, ui1);
  ADDS(12, 18); ADDS(13, 19); SETS(2, 20); ADDS(1, 21);
  un1[0] = 17;
  if (o->xper && m1 == X - 2) sxz_coef(o, 0, m2, m3, xkf, l, un1, ui1);
  else sxz_coef(o, m1 + 1, m2, m3, xkf, l, un1, ui1);
  ADDS(14, 18); ADDS(15, 19); ADDS(1, 20); SETS(3, 21);
  un1[0] = 17;
  syz_coef(o, m1, m2, m3, -ykf, l, un1, ui1);
  ADDS(8, 18); ADDS(10, 19); SETS(4, 20); ADDS(1, 21);
  un1[0] = 17;
  if (o->yper && m2 == Y - 2) syz_coef(o, m1, 0, m3, ykf, l, un1, ui1);
  else syz_coef(o, m1, m2 + 1, m3, ykf, l, un1, ui1);
  ADDS(9, 18); ADDS(11, 19); ADDS(1, 20); SETS(5, 21);
  if (o->p2koef != 0) {
    un1[0] = 17;
    if (o->zper && m3 == 0) continadd_coef(o, m1 + 1, m2 + 1, Z - 1, -zkf * o->p2koef, l, un1, ui1);
    else continadd_coef(o, m1 + 1, m2 + 1, m3, -zkf * o->p2koef, l, un1, ui1);
    ADDS(12, 18); ADDS(14, 19); ADDS(8, 20); ADDS(9, 21); ADDS(6, 22); ADDS(1, 23);
    un1[0] = 17;
    continadd_coef(o, m1 + 1, m2 + 1, m3 + 1, zkf * o->p2koef, l, un1, ui1);
    ADDS(13, 18); ADDS(15, 19); ADDS(10, 20); ADDS(11, 21); ADDS(1, 22); ADDS(7, 23);
  }
  un1[0] = 17;
}

/* continuity row, move3mg.c:1455-1573; side effect nd[cell] = cell viscosity */
static void contin_row(orc_t *o, long m1, long m2, long m3, int l, long *un1, double *ui1) {
  double xkf = 1.0 / (o->mggx[l][m1] - o->mggx[l][m1 - 1]);
  double ykf = 1.0 / (o->mggy[l][m2] - o->mggy[l][m2 - 1]);
  double zkf = 1.0 / (o->mggz[l][m3] - o->mggz[l][m3 - 1]);
  long v[8];
  cell_octet(o, l, m1, m2, m3, v);
  double ne = (l == 0) ? (double)o->nxx[v[7]] : cellvisc(o, v);
  ne = clampnu(o, ne);
  ui1[0] = 0;
  un1[0] = 0;
  o->nd[v[7]] = ne;
  un1[1] = v[0] * 4 + 1; ui1[1] = -xkf;
  un1[2] = v[2] * 4 + 1; ui1[2] = +xkf;
  un1[3] = v[0] * 4 + 2; ui1[3] = -ykf;
  un1[4] = v[1] * 4 + 2; ui1[4] = +ykf;
  un1[5] = v[0] * 4 + 3; ui1[5] = -zkf;
  un1[6] = v[4] * 4 + 3; ui1[6] = +zkf;
  un1[0] = 6;
}

/* boundary row x_r - k x_j = c, move3mg.c:1682-1729 */
static void bc_row(const orc_t *o, long r, long *un1, double *ui1) {
  int b = o->bondm[r];
  un1[0] = 1;
  ui1[0] = o->bondv1[2 * b + 0];
  un1[1] = r;
  ui1[1] = 1.0;
  if (o->bondn1[b]) {
    un1[0] += 1;
    un1[un1[0]] = o->bondn1[b] - 1;
    ui1[un1[0]] = -o->bondv1[2 * b + 1];
  }
}

/* one row of level l at node (m1,m2,m3), component c.  Returns 0 if the      */
/* unknown is a ghost (no row), 1 for an equation row, 2 for a boundary row.  */
/* Row selection follows move3mg.c:2372-2663.                                */
int orc_row(orc_t *o, int l, int m1, int m2, int m3, int c, long *un1, double *ui1) {
  long n = NID(o, l, m1, m2, m3);
  long r = n * 4 + c;
  const int X = o->mgx[l], Y = o->mgy[l], Z = o->mgz[l];
  int exists;
  if (c == 0) exists = (m1 && m2 && m3);
  else if (c == 1) exists = (m2 < Y - 1 && m3 < Z - 1);
  else if (c == 2) exists = (m1 < X - 1 && m3 < Z - 1);
  else exists = (m1 < X - 1 && m2 < Y - 1);
  if (!exists) {
    un1[0] = 0;
    ui1[0] = 0;
    return 0;
  }
  if (o->bondm[r]) {
    bc_row(o, r, un1, ui1);
    return 2;
  }
  if (c == 0) contin_row(o, m1, m2, m3, l, un1, ui1);
  else if (c == 1) xstokes_row(o, m1, m2, m3, l, un1, ui1);
  else if (c == 2) ystokes_row(o, m1, m2, m3, l, un1, ui1);
  else zstokes_row(o, m1, m2, m3, l, un1, ui1);
  return 1;
}

/* gausmat2 row store, parallel/gaus3mg_MG.c:311-365 */
static void gausmat2_store(orc_t *o, long mcmax, long pos0cur1, const long *un1, const double *ui1) {
  o->fre1[mcmax] = ui1[0];
  o->pos0[mcmax] = (int)pos0cur1;
  o->num0[mcmax] = 0;
  for (long m2 = 1; m2 <= un1[0]; m2++) {
    o->lin0[pos0cur1] = (int)un1[m2];
    o->val1[pos0cur1] = ui1[m2];
    pos0cur1++;
    o->num0[mcmax]++;
  }
}

/* assembly of all levels (move3mg.c:2329-2663); ghost unknowns get no row */
void orc_assemble(orc_t *o) {
  static const int slot[4] = {0, 6, 23, 40};
  long un1[ORC_MAXPOS];
  double ui1[ORC_MAXPOS];
  memset(o->num0, 0, sizeof(int) * 4 * o->ntot);
  for (int l = 0; l <= o->L; l++)
    for (int m3 = 0; m3 < o->mgz[l]; m3++)
      for (int m1 = 0; m1 < o->mgx[l]; m1++)
        for (int m2 = 0; m2 < o->mgy[l]; m2++) {
          long n = NID(o, l, m1, m2, m3);
          for (int c = 0; c < 4; c++) {
            memset(un1, 0, sizeof(un1));
            memset(ui1, 0, sizeof(ui1));
            int k = orc_row(o, l, m1, m2, m3, c, un1, ui1);
            if (k == 0) {
              o->pos0[n * 4 + c] = (int)(n * 57 + slot[c]);
              o->num0[n * 4 + c] = 0;
              o->fre1[n * 4 + c] = 0;
              continue;
            }
            gausmat2_store(o, n * 4 + c, n * 57 + slot[c], un1, ui1);
          }
        }
}

int orc_get_row(const orc_t *o, long r, int *idx, double *val, double *rhs) {
  int n = o->num0[r];
  for (int k = 0; k < n; k++) {
    idx[k] = o->lin0[o->pos0[r] + k];
    val[k] = o->val1[o->pos0[r] + k];
  }
  *rhs = o->fre1[r];
  return n;
}

/* ------------------------------------------------------------------------- */
/* Smoother + residual: parallel/move3mg.c:3855-3960 (gausmat4), one thread   */
/* ------------------------------------------------------------------------- */
static inline void row_update(orc_t *o, long m1, double *x) {
  double ival = -o->fre1[m1];
  double ival1 = o->val1[o->pos0[m1]];
  for (int m2 = 0; m2 < o->num0[m1]; m2++)
    ival += o->val1[o->pos0[m1] + m2] * x[o->lin0[o->pos0[m1] + m2]];
  if (!(m1 & 3)) {
    if (!o->bondm[m1]) x[m1] -= ival * o->nd[m1 / 4] * o->p0koef;
    else x[m1] -= ival;
  } else {
    if (!o->bondm[m1]) x[m1] -= ival / ival1 * o->v0koef;
    else x[m1] -= ival;
  }
}

void orc_residual(orc_t *o, int l) {
  long mcmin = o->mgp[l] * 4, mcmax = (o->mgp[l] + o->mgn[l]) * 4 - 1;
  /* rows are independent: the thread count does not change any value */
#pragma omp parallel for schedule(static)
  for (long m1 = mcmin; m1 <= mcmax; m1++)
    if (o->num0[m1]) {
      double ival = o->fre1[m1];
      for (int m2 = 0; m2 < o->num0[m1]; m2++)
        ival -= o->val1[o->pos0[m1] + m2] * o->sol0[o->lin0[o->pos0[m1] + m2]];
      o->fre0[m1] = ival;
    }
}

void orc_gausmat4(orc_t *o, int am, int l) {
  for (int s = 0; s < am; s++)
    for (long m30 = 0; m30 < o->mgn[l]; m30++)
      for (int m40 = 0; m40 < 4; m40++) {
        long m1 = (o->mgp[l] + m30) * 4 + m40;
        if (o->vonly && m40 == 0) continue;
        if (o->num0[m1]) row_update(o, m1, o->sol0);
      }
  orc_residual(o, l);
}

/* ------------------------------------------------------------------------- */
/* gausmat4 over the assembled row store in the multicolour order the device  */
/* runs on a caller's arrays: per sweep the momentum rows of vx, vy, vz, each */
/* red then black by (m1+m2+m3)&1, then all continuity rows, then the         */
/* boundary rows in dependency order (a row after the boundary row it copies). */
/* Rows of one phase read no unknown of the same phase (momentum rows couple */
/* face neighbours, continuity rows only velocities), so any order inside a  */
/* phase gives the same result; then fre0 = fre1 - A sol0 (move3mg.c:3935-3957). */
/* ------------------------------------------------------------------------- */
static int bc_row_depth(const orc_t *o, long r, int lim) {
  int d = 0;
  while (lim-- > 0) {
    long j = -1;
    for (int k = 0; k < o->num0[r]; k++)
      if (o->lin0[o->pos0[r] + k] != r) j = o->lin0[o->pos0[r] + k];
    if (j < 0 || !o->bondm[j]) return d;
    r = j;
    d++;
  }
  return -1;
}

int orc_gausmat4_colored(orc_t *o, int am, int l) {
  const long r0 = 4 * o->mgp[l], r1 = 4 * (o->mgp[l] + o->mgn[l]);
  int maxd = 0;
  for (long r = r0; r < r1; r++)
    if (o->num0[r] && o->bondm[r]) {
      int d = bc_row_depth(o, r, 64);
      if (d < 0) return -1;
      if (d > maxd) maxd = d;
    }
  for (int s = 0; s < am; s++) {
    for (int c = 1; c < 4; c++)
      for (int col = 0; col < 2; col++)
        for (int m3 = 0; m3 < o->mgz[l]; m3++)
          for (int m1 = 0; m1 < o->mgx[l]; m1++)
            for (int m2 = 0; m2 < o->mgy[l]; m2++) {
              if (((m1 + m2 + m3) & 1) != col) continue;
              long r = 4 * NID(o, l, m1, m2, m3) + c;
              if (o->num0[r] && !o->bondm[r]) row_update(o, r, o->sol0);
            }
    if (!o->vonly)
      for (long r = r0; r < r1; r += 4)
        if (o->num0[r] && !o->bondm[r]) row_update(o, r, o->sol0);
    for (int d = 0; d <= maxd; d++)
      for (long r = r0; r < r1; r++)
        if (o->num0[r] && o->bondm[r] && !(o->vonly && !(r & 3)) && bc_row_depth(o, r, 64) == d)
          row_update(o, r, o->sol0);
  }
  orc_residual(o, l);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Coloured tile-hybrid sweep: the same gausmat4 row updates in the order the */
/* GPU smoother uses.  Nodes are grouped in ty x tx x tz tiles (y fastest)   */
/* laid from the lower corner of the equation-row box (orc_eq_box).          */
/* Inside a tile the phases are vx(red, black), vy(red, black), vz(red,       */
/* black), P, colour = (m1+m2+m3)&1; unknowns of other tiles are read from    */
/* the state at the start of the sweep (block-Jacobi between tiles).          */
/* Boundary rows are left out of the tile phases and updated afterwards in    */
/* dependency order (orc_bc_pass).                                           */
/* ------------------------------------------------------------------------- */
typedef struct {
  long mgp;
  int Y, XY;
  int y0, x0, z0, ty, tx, tz;
} tilectx_t;

static inline int in_tile(const tilectx_t *t, long u) {
  long n = u / 4 - t->mgp;
  if (n < 0) return 0;
  int m3 = (int)(n / t->XY);
  int rem = (int)(n - (long)m3 * t->XY);
  int m1 = rem / t->Y, m2 = rem - m1 * t->Y;
  return m2 >= t->y0 && m2 < t->y0 + t->ty && m1 >= t->x0 && m1 < t->x0 + t->tx && m3 >= t->z0 &&
         m3 < t->z0 + t->tz;
}

static void tile_row_update(orc_t *o, long r, double *cur, const double *old, const tilectx_t *t) {
  double ival = -o->fre1[r];
  double ival1 = o->val1[o->pos0[r]];
  for (int m2 = 0; m2 < o->num0[r]; m2++) {
    long u = o->lin0[o->pos0[r] + m2];
    double xu = in_tile(t, u) ? cur[u] : old[u];
    ival += o->val1[o->pos0[r] + m2] * xu;
  }
  if (!(r & 3)) cur[r] -= ival * o->nd[r / 4] * o->p0koef;
  else cur[r] -= ival / ival1 * o->v0koef;
}

static int bc_depth(const orc_t *o, long r, int lim) {
  int d = 0;
  while (d <= lim) {
    int b = o->bondm[r];
    if (!b || !o->bondn1[b]) return d;
    long j = o->bondn1[b] - 1;
    if (!o->bondm[j] || !o->num0[j]) return d;
    r = j;
    d++;
  }
  return -1;
}

void orc_bc_pass(orc_t *o, int l, double *x) {
  long a = o->mgp[l] * 4, e = (o->mgp[l] + o->mgn[l]) * 4;
  int maxd = 0;
  for (long r = a; r < e; r++)
    if (o->num0[r] && o->bondm[r]) {
      int d = bc_depth(o, r, 64);
      if (d > maxd) maxd = d;
    }
  for (int d = 0; d <= maxd; d++)
    for (long r = a; r < e; r++)
      if (o->num0[r] && o->bondm[r] && !(o->vonly && !(r & 3)) && bc_depth(o, r, 64) == d) {
        double ival = -o->fre1[r];
        for (int m2 = 0; m2 < o->num0[r]; m2++)
          ival += o->val1[o->pos0[r] + m2] * x[o->lin0[o->pos0[r] + m2]];
        x[r] -= ival;
      }
}

/* Bounding box {y0, x0, z0, y1, x1, z1} (upper bounds exclusive) of the     */
/* nodes that hold an equation row (a row that is not a boundary equation);  */
/* the tiles start at its lower corner.  The whole level if there is none.   */
void orc_eq_box(const orc_t *o, int l, int box[6]) {
  const int X = o->mgx[l], Y = o->mgy[l], Z = o->mgz[l];
  int b[6] = {Y, X, Z, 0, 0, 0};
  for (int m3 = 0; m3 < Z; m3++)
    for (int m1 = 0; m1 < X; m1++)
      for (int m2 = 0; m2 < Y; m2++) {
        long r0 = NID(o, l, m1, m2, m3) * 4;
        int eq = 0;
        for (int c = 0; c < 4; c++) eq |= o->num0[r0 + c] && !o->bondm[r0 + c];
        if (!eq) continue;
        const int m[3] = {m2, m1, m3};
        for (int k = 0; k < 3; k++) {
          if (m[k] < b[k]) b[k] = m[k];
          if (m[k] + 1 > b[k + 3]) b[k + 3] = m[k] + 1;
        }
      }
  if (b[3] == 0) { b[0] = b[1] = b[2] = 0; b[3] = Y; b[4] = X; b[5] = Z; }
  for (int k = 0; k < 6; k++) box[k] = b[k];
}

void orc_sweep_tiles(orc_t *o, int am, int l, int ty, int tx, int tz) {
  const int X = o->mgx[l], Y = o->mgy[l], Z = o->mgz[l];
  long a = o->mgp[l] * 4, e = (o->mgp[l] + o->mgn[l]) * 4;
  double *old = o->tmp;
  double *cur = o->sol0;
  static const int comp[4] = {1, 2, 3, 0};
  int box[6];
  orc_eq_box(o, l, box);
  for (int s = 0; s < am; s++) {
    memcpy(old + a, cur + a, sizeof(double) * (e - a));
    /* tiles only read their own nodes in cur and the rest in old, so they are
       independent within a sweep and may run in any order or in parallel */
    const int ntz = (box[5] - box[2] + tz - 1) / tz, ntx = (box[4] - box[1] + tx - 1) / tx,
              nty = (box[3] - box[0] + ty - 1) / ty;
#pragma omp parallel for collapse(3) schedule(static)
    for (int kz = 0; kz < ntz; kz++)
      for (int kx = 0; kx < ntx; kx++)
        for (int ky = 0; ky < nty; ky++)
          for (int ph = 0; ph < (o->vonly ? 6 : 7); ph++) {
            tilectx_t t;
            t.mgp = o->mgp[l]; t.Y = Y; t.XY = o->mgxy[l]; t.ty = ty; t.tx = tx; t.tz = tz;
            t.z0 = box[2] + kz * tz; t.x0 = box[1] + kx * tx; t.y0 = box[0] + ky * ty;
            int c = comp[ph / 2];
            int colour = ph & 1;
            for (int m3 = t.z0; m3 < t.z0 + tz && m3 < Z; m3++)
              for (int m1 = t.x0; m1 < t.x0 + tx && m1 < X; m1++)
                for (int m2 = t.y0; m2 < t.y0 + ty && m2 < Y; m2++) {
                  if (c != 0 && ((m1 + m2 + m3) & 1) != colour) continue;
                  long r = NID(o, l, m1, m2, m3) * 4 + c;
                  if (!o->num0[r] || o->bondm[r]) continue;
                  tile_row_update(o, r, cur, old, &t);
                }
          }
    orc_bc_pass(o, l, cur);
  }
  orc_residual(o, l);
}

/* ------------------------------------------------------------------------- */
/* Residual restriction: parallel/move3mg.c:3966-4354 (one thread)            */
/* Per-axis rules: 'N' = staggered on grid lines (vx in x), 'C' = cell centre */
/* with lower index (vx in y/z), 'P' = pressure cell (upper-corner index).    */
/* ------------------------------------------------------------------------- */
typedef struct {
  int lo;
  double d;
} axmap_t;

static void ax_rule(char kind, int m, int nf, int nc, const double *gf, const double *gc, int *lo_out,
                    double *d_out) {
  int lo = (m - 2) / 2 + 2;
  if (m < 3) lo = m;
  if (m > nf - 4) lo = nc - (nf - m);
  if (kind == 'N') {
    if (lo > nc - 2) lo = nc - 2;
    *d_out = gc[lo + 1] - gc[lo];
  } else if (kind == 'C') {
    if (lo > nc - 3) lo = nc - 3;
    if ((gf[m] + gf[m + 1]) / 2.0 < (gc[lo] + gc[lo + 1]) / 2.0 && lo > 0) lo -= 1;
    *d_out = (gc[lo + 2] - gc[lo]) / 2.0;
  } else {
    if (lo > nc - 2) lo = nc - 2;
    if ((gf[m] + gf[m - 1]) / 2.0 < (gc[lo] + gc[lo - 1]) / 2.0 && lo > 1) lo -= 1;
    *d_out = (gc[lo + 1] - gc[lo - 1]) / 2.0;
  }
  *lo_out = lo;
}

static double ax_w(char kind, int cc, int lo, int m, const double *gf, const double *gc, double d) {
  double dd;
  if (kind == 'N') dd = (gc[cc] - gf[m]) / d;
  else if (kind == 'C') dd = (gc[cc] + gc[cc + 1] - gf[m] - gf[m + 1]) / 2.0 / d;
  else dd = (gc[cc] + gc[cc - 1] - gf[m] - gf[m - 1]) / 2.0 / d;
  if (cc == lo) dd = -dd;
  return 1.0 - dd;
}

static const char KIND[4][3] = {{'P', 'P', 'P'}, {'N', 'C', 'C'}, {'C', 'N', 'C'}, {'C', 'C', 'N'}};

static int comp_exists(const orc_t *o, int l, int m1, int m2, int m3, int c) {
  if (c == 0) return m1 && m2 && m3;
  if (c == 1) return m2 < o->mgy[l] - 1 && m3 < o->mgz[l] - 1;
  if (c == 2) return m1 < o->mgx[l] - 1 && m3 < o->mgz[l] - 1;
  return m1 < o->mgx[l] - 1 && m2 < o->mgy[l] - 1;
}

void orc_vprestrict(orc_t *o, int l) {
  const int f = l - 1;
  long a = o->mgp[l] * 4, e = (o->mgp[l] + o->mgn[l]) * 4;
  double *S0 = (double *)calloc(e - a, sizeof(double)), *S1 = (double *)calloc(e - a, sizeof(double));
  for (long m = a; m < e; m++) { o->fre1[m] = 0; o->sol1[m] = 0; }
  const double *gf[3] = {o->mggx[f], o->mggy[f], o->mggz[f]};
  const double *gc[3] = {o->mggx[l], o->mggy[l], o->mggz[l]};
  const int nfa[3] = {o->mgx[f], o->mgy[f], o->mgz[f]}, nca[3] = {o->mgx[l], o->mgy[l], o->mgz[l]};
  for (int m3 = 0; m3 < o->mgz[f]; m3++)
    for (int m1 = 0; m1 < o->mgx[f]; m1++)
      for (int m2 = 0; m2 < o->mgy[f]; m2++) {
        long m4 = NID(o, f, m1, m2, m3);
        static const int order[4] = {1, 2, 3, 0};
        for (int q = 0; q < 4; q++) {
          int c = order[q];
          if (!comp_exists(o, f, m1, m2, m3, c) || o->bondm[m4 * 4 + c]) continue;
          int mm[3] = {m1, m2, m3}, lo[3];
          double d[3];
          for (int ax = 0; ax < 3; ax++)
            ax_rule(KIND[c][ax], mm[ax], nfa[ax], nca[ax], gf[ax], gc[ax], &lo[ax], &d[ax]);
          for (int m30 = lo[2]; m30 <= lo[2] + 1; m30++) {
            double ddz = ax_w(KIND[c][2], m30, lo[2], m3, gf[2], gc[2], d[2]);
            for (int m10 = lo[0]; m10 <= lo[0] + 1; m10++) {
              double ddx = ax_w(KIND[c][0], m10, lo[0], m1, gf[0], gc[0], d[0]);
              for (int m20 = lo[1]; m20 <= lo[1] + 1; m20++) {
                double ddy = ax_w(KIND[c][1], m20, lo[1], m2, gf[1], gc[1], d[1]);
                long m40 = NID(o, l, m10, m20, m30);
                if (o->bondm[m40 * 4 + c]) continue;
                double swt = ddx * ddy * ddz;
                double ival2;
                if (c == 0) ival2 = o->fre0[m4 * 4] * o->nd[m4] / o->nd[m40] * swt;
                else ival2 = o->fre0[m4 * 4 + c] * swt;
                S0[m40 * 4 + c - a] += ival2;
                S1[m40 * 4 + c - a] += swt;
              }
            }
          }
        }
      }
  for (long m = a; m < e; m++) {
    o->fre1[m] += S0[m - a];
    o->sol1[m] += S1[m - a];
  }
  for (long m = a; m < e; m++) {
    if (o->sol1[m]) {
      o->fre1[m] /= o->sol1[m];
      o->sol1[m] = 0;
    }
    o->sol0[m] = 0;
  }
  free(S0);
  free(S1);
}

/* correction prolongation: parallel/move3mg.c:3535-3850 */
void orc_vpprolong(orc_t *o, int l, int ynreset) {
  const int f = l - 1;
  const double *gf[3] = {o->mggx[f], o->mggy[f], o->mggz[f]};
  const double *gc[3] = {o->mggx[l], o->mggy[l], o->mggz[l]};
  const int nfa[3] = {o->mgx[f], o->mgy[f], o->mgz[f]}, nca[3] = {o->mgx[l], o->mgy[l], o->mgz[l]};
  static const int order[4] = {1, 2, 3, 0};
  for (int m3 = 0; m3 < o->mgz[f]; m3++)
    for (int m1 = 0; m1 < o->mgx[f]; m1++)
      for (int m2 = 0; m2 < o->mgy[f]; m2++) {
        long m4 = NID(o, f, m1, m2, m3);
        for (int q = 0; q < 4; q++) {
          int c = order[q];
          if (!comp_exists(o, f, m1, m2, m3, c)) continue;
          if (o->vonly && c == 0) continue;
          int mm[3] = {m1, m2, m3}, lo[3];
          double d[3];
          for (int ax = 0; ax < 3; ax++)
            ax_rule(KIND[c][ax], mm[ax], nfa[ax], nca[ax], gf[ax], gc[ax], &lo[ax], &d[ax]);
          double k = (c == 0) ? o->p1koef : o->v1koef;
          if (ynreset) o->sol0[m4 * 4 + c] = 0;
          for (int m30 = lo[2]; m30 <= lo[2] + 1; m30++) {
            double ddz = ax_w(KIND[c][2], m30, lo[2], m3, gf[2], gc[2], d[2]);
            for (int m10 = lo[0]; m10 <= lo[0] + 1; m10++) {
              double ddx = ax_w(KIND[c][0], m10, lo[0], m1, gf[0], gc[0], d[0]);
              for (int m20 = lo[1]; m20 <= lo[1] + 1; m20++) {
                double ddy = ax_w(KIND[c][1], m20, lo[1], m2, gf[1], gc[1], d[1]);
                long m40 = NID(o, l, m10, m20, m30);
                double swt = ddx * ddy * ddz;
                o->sol0[m4 * 4 + c] += o->sol0[m40 * 4 + c] * swt * k;
              }
            }
          }
        }
      }
}

/* ------------------------------------------------------------------------- */
/* Coarsest-level direct solve.  The reference hands the coarsest CSR system */
/* to MKL PARDISO (parallel/gaus3mg_MG.c:94-249); this is a dense LU with     */
/* partial pivoting (row interchanges, first maximal |a| wins), on the same  */
/* assembled rows; ghost unknowns get identity rows (move3mg.c:2417-2425).    */
/* The solution is stored at the hierarchy offset 4*(mgp[L]+local)+c.          */
/* ------------------------------------------------------------------------- */
int orc_dense_lu(int n, double *a, int *piv) {
  for (int k = 0; k < n; k++) {
    int p = k;
    double best = fabs(a[(long)k * n + k]);
    for (int i = k + 1; i < n; i++) {
      double v = fabs(a[(long)i * n + k]);
      if (v > best) { best = v; p = i; }
    }
    piv[k] = p;
    if (best == 0.0) return k + 1;
    if (p != k)
      for (int j = 0; j < n; j++) {
        double t = a[(long)k * n + j];
        a[(long)k * n + j] = a[(long)p * n + j];
        a[(long)p * n + j] = t;
      }
    double inv = a[(long)k * n + k];
    for (int i = k + 1; i < n; i++) a[(long)i * n + k] /= inv;
    for (int i = k + 1; i < n; i++) {
      double lik = a[(long)i * n + k];
      if (lik == 0.0) continue;
      double *ri = a + (long)i * n, *rk = a + (long)k * n;
      for (int j = k + 1; j < n; j++) ri[j] -= lik * rk[j];
    }
  }
  return 0;
}

void orc_dense_lu_solve(int n, const double *lu, const int *piv, double *b) {
  for (int k = 0; k < n; k++)
    if (piv[k] != k) { double t = b[k]; b[k] = b[piv[k]]; b[piv[k]] = t; }
  for (int i = 0; i < n; i++) {
    double s = b[i];
    for (int j = 0; j < i; j++) s -= lu[(long)i * n + j] * b[j];
    b[i] = s;
  }
  for (int i = n - 1; i >= 0; i--) {
    double s = b[i];
    for (int j = i + 1; j < n; j++) s -= lu[(long)i * n + j] * b[j];
    b[i] = s / lu[(long)i * n + i];
  }
}

int orc_coarse_factor(orc_t *o) {
  int L = o->L;
  int n = (int)(4 * o->mgn[L]);
  long base = 4 * o->mgp[L];
  free(o->lu);
  free(o->piv);
  o->nc = n;
  o->lu = (double *)calloc((size_t)n * n, sizeof(double));
  o->piv = (int *)calloc(n, sizeof(int));
  for (int i = 0; i < n; i++) {
    long r = base + i;
    if (!o->num0[r] || (o->vonly && !(r & 3))) {
      o->lu[(long)i * n + i] = 1.0;
      continue;
    }
    for (int k = 0; k < o->num0[r]; k++) {
      long c = o->lin0[o->pos0[r] + k] - base;
      if (c < 0 || c >= n) return -2;
      o->lu[(long)i * n + c] += o->val1[o->pos0[r] + k];
    }
  }
  return orc_dense_lu(n, o->lu, o->piv);
}

int orc_coarse_solve(orc_t *o) {
  int L = o->L;
  long base = 4 * o->mgp[L];
  int n = o->nc;
  double *b = (double *)malloc(sizeof(double) * n);
  for (int i = 0; i < n; i++)
    b[i] = (o->num0[base + i] && !(o->vonly && !((base + i) & 3))) ? o->fre1[base + i] : 0.0;
  orc_dense_lu_solve(n, o->lu, o->piv, b);
  for (int i = 0; i < n; i++) o->sol0[base + i] = b[i];
  free(b);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* V-cycle (move3mg.c:2710-2792 with the coarse solve every cycle on the     */
/* restricted residual, as sketched at pardiso/move3mg.c:2718,2756-2763).     */
/* mode 0: lexicographic gausmat4; mode 1: coloured tile sweep.               */
/* ------------------------------------------------------------------------- */
static void smooth(orc_t *o, int am, int l, int mode, int ty, int tx, int tz) {
  if (am <= 0) { orc_residual(o, l); return; }
  if (mode == 0) orc_gausmat4(o, am, l);
  else if (mode == 2) orc_gausmat4_colored(o, am, l);
  else orc_sweep_tiles(o, am, l, ty, tx, tz);
}

void orc_vcycle(orc_t *o, const int *pre, const int *post, int mode, int ty, int tx, int tz) {
  int L = o->L;
  if (L == 0) {
    smooth(o, pre[0], 0, mode, ty, tx, tz);
    return;
  }
  for (int l = 0; l < L; l++) {
    smooth(o, pre[l], l, mode, ty, tx, tz);
    orc_vprestrict(o, l + 1);
  }
  orc_coarse_solve(o);
  orc_vpprolong(o, L, 0);
  for (int l = L - 1; l > 0; l--) {
    if (post[l] > 0) smooth(o, post[l], l, mode, ty, tx, tz);
    orc_vpprolong(o, l, 0);
  }
  if (post[0] > 0) smooth(o, post[0], 0, mode, ty, tx, tz);
}

/* Relative residual on the finest level: ||W r|| / ||W b||, W = 1 on        */
/* momentum and boundary rows, W = eta_cell / h_cell on continuity rows       */
/* (h_cell = mean cell edge).  blocks3 = {||r_mom||, ||W r_cont||, ||r_bc||}. */
double orc_relres(orc_t *o, double *blocks3) {
  double sr = 0, sb = 0, bm = 0, bc = 0, bb = 0;
  const double *gx = o->mggx[0], *gy = o->mggy[0], *gz = o->mggz[0];
  for (int m3 = 0; m3 < o->mgz[0]; m3++)
    for (int m1 = 0; m1 < o->mgx[0]; m1++)
      for (int m2 = 0; m2 < o->mgy[0]; m2++) {
        long n = NID(o, 0, m1, m2, m3);
        for (int c = 0; c < 4; c++) {
          long r = n * 4 + c;
          if (!o->num0[r]) continue;
          double res = o->fre1[r];
          for (int k = 0; k < o->num0[r]; k++) res -= o->val1[o->pos0[r] + k] * o->sol0[o->lin0[o->pos0[r] + k]];
          double w = 1.0;
          if (c == 0 && !o->bondm[r]) {
            double h = ((gx[m1] - gx[m1 - 1]) + (gy[m2] - gy[m2 - 1]) + (gz[m3] - gz[m3 - 1])) / 3.0;
            w = o->nd[n] / h;
          }
          sr += (w * res) * (w * res);
          sb += (w * o->fre1[r]) * (w * o->fre1[r]);
          if (o->bondm[r]) bb += res * res;
          else if (c == 0) bc += (w * res) * (w * res);
          else bm += res * res;
        }
      }
  if (blocks3) { blocks3[0] = sqrt(bm); blocks3[1] = sqrt(bc); blocks3[2] = sqrt(bb); }
  return sb > 0 ? sqrt(sr) / sqrt(sb) : sqrt(sr);
}

/* The same W-weighted relative residual for a finest-level state x (4 mgn[0]
   doubles), with every row built on the fly by orc_row instead of read from
   the row store, so that levels too large to assemble (57 slots per node) can
   be scored.  Rows are independent; per-thread partial sums over z planes are
   added in plane order, so the result does not depend on the thread count.
   Returns the relative residual; blocks3 as orc_relres. */
double orc_relres_rows(orc_t *o, const double *x, double *blocks3) {
  const int X = o->mgx[0], Y = o->mgy[0], Z = o->mgz[0];
  const double *gx = o->mggx[0], *gy = o->mggy[0], *gz = o->mggz[0];
  double *part = (double *)calloc((size_t)5 * Z, sizeof(double));
#pragma omp parallel for schedule(dynamic, 1)
  for (int m3 = 0; m3 < Z; m3++) {
    long un1[ORC_MAXPOS];
    double ui1[ORC_MAXPOS];
    double sr = 0, sb = 0, bm = 0, bc = 0, bb = 0;
    for (int m1 = 0; m1 < X; m1++)
      for (int m2 = 0; m2 < Y; m2++) {
        long n = NID(o, 0, m1, m2, m3);
        for (int c = 0; c < 4; c++) {
          int k = orc_row(o, 0, m1, m2, m3, c, un1, ui1);
          if (!k) continue;
          long r = n * 4 + c;
          double res = ui1[0];
          for (long q = 1; q <= un1[0]; q++) res -= ui1[q] * x[un1[q]];
          double w = 1.0;
          if (c == 0 && !o->bondm[r]) {
            double h = ((gx[m1] - gx[m1 - 1]) + (gy[m2] - gy[m2 - 1]) + (gz[m3] - gz[m3 - 1])) / 3.0;
            w = o->nd[n] / h;
          }
          sr += (w * res) * (w * res);
          sb += (w * ui1[0]) * (w * ui1[0]);
          if (o->bondm[r]) bb += res * res;
          else if (c == 0) bc += (w * res) * (w * res);
          else bm += res * res;
        }
      }
    double *p = part + 5 * (size_t)m3;
    p[0] = sr; p[1] = sb; p[2] = bm; p[3] = bc; p[4] = bb;
  }
  double t[5] = {0, 0, 0, 0, 0};
  for (int m3 = 0; m3 < Z; m3++)
    for (int q = 0; q < 5; q++) t[q] += part[5 * (size_t)m3 + q];
  free(part);
  if (blocks3) { blocks3[0] = sqrt(t[2]); blocks3[1] = sqrt(t[3]); blocks3[2] = sqrt(t[4]); }
  return t[1] > 0 ? sqrt(t[0]) / sqrt(t[1]) : sqrt(t[0]);
}

int orc_solve(orc_t *o, const int *pre, const int *post, int mode, double tol, int maxcyc,
              double *hist) {
  double r = orc_relres(o, NULL);
  if (hist) hist[0] = r;
  int it = 0;
  while (r > tol && it < maxcyc) {
    orc_vcycle(o, pre, post, mode, 64, 4, 4);
    it++;
    r = orc_relres(o, NULL);
    if (hist) hist[it] = r;
    if (!(r == r)) return -it;
  }
  return it;
}

/* ------------------------------------------------------------------------- */
/* Banded direct solve of the finest assembled system: restatement of the    */
/* frontal banded Gaussian elimination gausmat3 (parallel/gausjomp.c:70-254), */
/* single thread.  Rows are processed in order: the row is gathered into a   */
/* dense buffer, entries left of the diagonal are eliminated with the already */
/* normalised rows (gausjomp.c:150-183), the row is normalised and its right  */
/* band (width rightnum = max(col-row)) stored (gausjomp.c:196-204); then     */
/* back substitution (gausjomp.c:219-250).  No pivoting, as in the reference. */
/* Returns 0; -1 if the band store exceeds the reference's MAXMAT=1e9         */
/* (parallel/head3mg.c:13); -2 on allocation failure; -3 singular line.       */
/* ------------------------------------------------------------------------- */
int orc_band_solve(orc_t *o, double *x, long *band_out) {
  long n = 4 * o->mgn[0];
  long rightn = 0;
  for (long r = 0; r < n; r++)
    for (int k = 0; k < o->num0[r]; k++) {
      long d = (long)o->lin0[o->pos0[r] + k] - r;
      if (d > rightn) rightn = d;
    }
  if (band_out) *band_out = rightn;
  if ((double)n * (double)rightn > 1e9) return -1;
  double *U = (double *)calloc((size_t)n * (rightn ? rightn : 1), sizeof(double));
  double *buf = (double *)calloc((size_t)(n + rightn + 1), sizeof(double));
  double *f = (double *)malloc(sizeof(double) * n);
  if (!U || !buf || !f) { free(U); free(buf); free(f); return -2; }
  for (long r = 0; r < n; r++) f[r] = o->fre1[r];
  for (long m1 = 0; m1 < n; m1++) {
    if (!o->num0[m1]) continue;
    long m6 = m1;
    for (int k = 0; k < o->num0[m1]; k++) {
      long m4 = o->lin0[o->pos0[m1] + k];
      buf[m4] += o->val1[o->pos0[m1] + k];
      if (m4 < m6) m6 = m4;
    }
    for (long m2 = m6; m2 < m1; m2++)
      if (buf[m2]) {
        double ival = buf[m2];
        buf[m2] = 0;
        if (!o->num0[m2]) { free(U); free(buf); free(f); return -3; }
        const double *u = U + m2 * rightn;
        for (long m3 = m2 + 1; m3 <= m2 + rightn; m3++) buf[m3] -= ival * u[m3 - m2 - 1];
        f[m1] -= f[m2] * ival;
      }
    double ival = buf[m1];
    buf[m1] = 0;
    if (!ival) { free(U); free(buf); free(f); return -3; }
    double *u = U + m1 * rightn;
    for (long m2 = m1 + 1; m2 <= m1 + rightn; m2++) {
      u[m2 - m1 - 1] = buf[m2] / ival;
      buf[m2] = 0;
    }
    f[m1] /= ival;
  }
  for (long m1 = n - 1; m1 >= 0; m1--) {
    if (o->num0[m1]) {
      const double *u = U + m1 * rightn;
      double ival = f[m1];
      for (long m2 = m1 + 1; m2 < m1 + 1 + rightn && m2 < n; m2++) ival -= u[m2 - m1 - 1] * x[m2];
      x[m1] = ival;
    } else {
      x[m1] = 0;
    }
  }
  free(U); free(buf); free(f);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* The same banded elimination with the reference's thread structure          */
/* (parallel/gausjomp.c:117-215): rows go round-robin to the threads, each    */
/* thread accumulates its row in a private dense buffer and, before           */
/* eliminating with an earlier row, waits until that row is finished (the     */
/* reference's pos0[m2] == -1 flag, here one ready byte per row with          */
/* acquire/release ordering).  Every row sees the same finished rows in the   */
/* same order as the serial restatement, so the result is bit-identical to    */
/* orc_band_solve whatever the thread count; back substitution on one        */
/* thread (gausjomp.c:219-250).  threads <= 0: the OpenMP default.            */
/* ------------------------------------------------------------------------- */
int orc_band_solve_mt(orc_t *o, double *x, long *band_out, int threads) {
  long n = 4 * o->mgn[0];
  long rightn = 0;
  for (long r = 0; r < n; r++)
    for (int k = 0; k < o->num0[r]; k++) {
      long d = (long)o->lin0[o->pos0[r] + k] - r;
      if (d > rightn) rightn = d;
    }
  if (band_out) *band_out = rightn;
  if (rightn < 1) rightn = 1;
  double *U = (double *)calloc((size_t)n * rightn, sizeof(double));
  double *f = (double *)malloc(sizeof(double) * n);
  unsigned char *ready = (unsigned char *)calloc((size_t)n, 1);
  int nt = threads > 0 ? threads : omp_get_max_threads();
  double *bufs = (double *)calloc((size_t)nt * (size_t)(n + rightn + 1), sizeof(double));
  if (!U || !f || !ready || !bufs) { free(U); free(f); free(ready); free(bufs); return -2; }
  for (long r = 0; r < n; r++) f[r] = o->fre1[r];
  int err = 0;
#pragma omp parallel num_threads(nt)
  {
    const int t = omp_get_thread_num(), tn = omp_get_num_threads();
    double *buf = bufs + (size_t)t * (size_t)(n + rightn + 1);
    for (long m1 = t; m1 < n; m1 += tn) {
      if (o->num0[m1] && !__atomic_load_n(&err, __ATOMIC_RELAXED)) {
        long m6 = m1;
        for (int k = 0; k < o->num0[m1]; k++) {
          long m4 = o->lin0[o->pos0[m1] + k];
          buf[m4] += o->val1[o->pos0[m1] + k];
          if (m4 < m6) m6 = m4;
        }
        for (long m2 = m6; m2 < m1; m2++)
          if (buf[m2]) {
            double ival = buf[m2];
            buf[m2] = 0;
            while (!__atomic_load_n(&ready[m2], __ATOMIC_ACQUIRE)) {
              if (__atomic_load_n(&err, __ATOMIC_RELAXED)) break;
            }
            if (!o->num0[m2]) { __atomic_store_n(&err, -3, __ATOMIC_RELAXED); break; }
            const double *u = U + m2 * rightn;
            for (long m3 = m2 + 1; m3 <= m2 + rightn; m3++) buf[m3] -= ival * u[m3 - m2 - 1];
            f[m1] -= f[m2] * ival;
          }
        double ival = buf[m1];
        buf[m1] = 0;
        if (!ival) __atomic_store_n(&err, -3, __ATOMIC_RELAXED);
        double *u = U + m1 * rightn;
        for (long m2 = m1 + 1; m2 <= m1 + rightn; m2++) {
          if (ival) u[m2 - m1 - 1] = buf[m2] / ival;
          buf[m2] = 0;
        }
        if (ival) f[m1] /= ival;
      }
      __atomic_store_n(&ready[m1], 1, __ATOMIC_RELEASE);
    }
  }
  if (!err)
    for (long m1 = n - 1; m1 >= 0; m1--) {
      if (o->num0[m1]) {
        const double *u = U + m1 * rightn;
        double ival = f[m1];
        for (long m2 = m1 + 1; m2 < m1 + 1 + rightn && m2 < n; m2++) ival -= u[m2 - m1 - 1] * x[m2];
        x[m1] = ival;
      } else {
        x[m1] = 0;
      }
    }
  free(U); free(f); free(ready); free(bufs);
  return err;
}

/* ------------------------------------------------------------------------- */
/* FP64 outer iteration on the assembled finest operator: GCR with the right  */
/* preconditioner z_p = omega*eta_cell*r_p (continuity rows), pressure        */
/* boundary rows by substitution, z_v = one velocity-block V-cycle of the     */
/* reference smoother/transfers for r_v - G z_p.  The same algorithm the      */
/* device solver runs; used as the CPU baseline and as a solution check.      */
/* ------------------------------------------------------------------------- */
static void ell_apply(const orc_t *o, long n4, const double *x, double *y) {
#pragma omp parallel for schedule(static)
  for (long r = 0; r < n4; r++) {
    double s = 0;
    for (int k = 0; k < o->num0[r]; k++) s += o->val1[o->pos0[r] + k] * x[o->lin0[o->pos0[r] + k]];
    y[r] = o->num0[r] ? s : 0.0;
  }
}

static double wdot(long n4, const double *w2, const double *a, const double *b) {
  double s = 0;
#pragma omp parallel for schedule(static) reduction(+ : s)
  for (long i = 0; i < n4; i++) s += w2[i] * a[i] * b[i];
  return s;
}

void orc_weights(orc_t *o, double *w2) {
  const double *gx = o->mggx[0], *gy = o->mggy[0], *gz = o->mggz[0];
  for (int m3 = 0; m3 < o->mgz[0]; m3++)
    for (int m1 = 0; m1 < o->mgx[0]; m1++)
      for (int m2 = 0; m2 < o->mgy[0]; m2++) {
        long n = NID(o, 0, m1, m2, m3);
        for (int c = 0; c < 4; c++) {
          long r = 4 * n + c;
          double w = o->num0[r] ? 1.0 : 0.0;
          if (c == 0 && o->num0[r] && !o->bondm[r]) {
            double h = ((gx[m1] - gx[m1 - 1]) + (gy[m2] - gy[m2 - 1]) + (gz[m3] - gz[m3 - 1])) / 3.0;
            w = o->nd[n] / h;
          }
          w2[r] = w * w;
        }
      }
}

/* v = 0 and P integrated up each column through the z rows (in3mg.c:927-938) */
void orc_hydrostatic(orc_t *o, double *x) {
  long n4 = 4 * o->mgn[0];
  memset(x, 0, sizeof(double) * n4);
  for (int m1 = 0; m1 < o->mgx[0]; m1++)
    for (int m2 = 0; m2 < o->mgy[0]; m2++)
      for (int m3 = 0; m3 + 1 < o->mgz[0]; m3++) {
        long r = 4 * NID(o, 0, m1, m2, m3) + 3;
        if (!o->num0[r] || o->bondm[r] || m1 + 1 >= o->mgx[0] || m2 + 1 >= o->mgy[0]) continue;
        long plo = 4 * NID(o, 0, m1 + 1, m2 + 1, m3), phi = 4 * NID(o, 0, m1 + 1, m2 + 1, m3 + 1);
        double zkf = 2.0 / (o->mggz[0][m3 + 1] - o->mggz[0][m3 - 1]);
        x[phi] = x[plo] - o->fre1[r] / zkf;
      }
  int vonly = o->vonly;
  o->vonly = 0;
  orc_bc_pass(o, 0, x);
  o->vonly = vonly;
}

int orc_gcr(orc_t *o, const int *pre, const int *post, int mode, double omega, double tol, int maxit,
            int restart, int hydro, double *x_out, double *hist, double *seconds_per_it) {
  long n4 = 4 * o->mgn[0];
  double *b = (double *)malloc(sizeof(double) * n4), *x = (double *)calloc(n4, sizeof(double));
  double *r = (double *)malloc(sizeof(double) * n4), *w2 = (double *)malloc(sizeof(double) * n4);
  double *Z = (double *)malloc(sizeof(double) * n4 * restart), *Q = (double *)malloc(sizeof(double) * n4 * restart);
  double *fsave = (double *)malloc(sizeof(double) * 4 * o->ntot);
  memcpy(b, o->fre1, sizeof(double) * n4);
  memcpy(fsave, o->fre1, sizeof(double) * 4 * o->ntot);
  orc_weights(o, w2);
  o->vonly = 1;
  if (orc_coarse_factor(o)) return -1000;
  if (hydro) orc_hydrostatic(o, x);
  ell_apply(o, n4, x, r);
  OMP_FOR
  for (long i = 0; i < n4; i++) r[i] = b[i] - r[i];
  double nb = sqrt(wdot(n4, w2, b, b));
  if (nb == 0) nb = 1;
  double rel = sqrt(wdot(n4, w2, r, r)) / nb;
  if (hist) hist[0] = rel;
  int it = 0, k = 0;
  struct timespec ts0, ts1;
  clock_gettime(CLOCK_MONOTONIC, &ts0);
  while (rel > tol && it < maxit) {
    double *z = Z + (long)k * n4, *q = Q + (long)k * n4;
    /* pressure part */
    OMP_FOR
    for (long i = 0; i < n4; i++) z[i] = 0;
    OMP_FOR
    for (long i = 0; i < n4; i += 4)
      if (o->num0[i] && !o->bondm[i]) z[i] = omega * o->nd[i / 4] * r[i];
    {
      /* pressure boundary rows: z_r = r_r + k z_j in dependency order */
      double *ftmp = o->fre1;
      memcpy(o->tmp, o->fre1, sizeof(double) * n4);
      memcpy(o->fre1, r, sizeof(double) * n4);
      int vo = o->vonly;
      o->vonly = 0;
      /* run the pass on P rows only: velocity BC rows of z are overwritten below */
      orc_bc_pass(o, 0, z);
      o->vonly = vo;
      memcpy(ftmp, o->tmp, sizeof(double) * n4);
      OMP_FOR
      for (long i = 0; i < n4; i++)
        if (i & 3) z[i] = 0;
    }
    /* velocity right-hand side r_v - G z_p */
    ell_apply(o, n4, z, q);
    OMP_FOR
    for (long i = 0; i < n4; i++) o->fre1[i] = (i & 3) ? ((o->num0[i] && !o->bondm[i]) ? r[i] - q[i] : r[i]) : 0.0;
    memset(o->sol0, 0, sizeof(double) * 4 * o->ntot);
    orc_vcycle(o, pre, post, mode, 32, 4, 4);
    OMP_FOR
    for (long i = 0; i < n4; i++)
      if (i & 3) z[i] = o->sol0[i];
    memcpy(o->fre1, fsave, sizeof(double) * 4 * o->ntot);
    ell_apply(o, n4, z, q);
    for (int j = 0; j < k; j++) {
      double *zj = Z + (long)j * n4, *qj = Q + (long)j * n4;
      double be = wdot(n4, w2, q, qj);
      OMP_FOR
      for (long i = 0; i < n4; i++) { q[i] -= be * qj[i]; z[i] -= be * zj[i]; }
    }
    double nq = sqrt(wdot(n4, w2, q, q));
    OMP_FOR
    for (long i = 0; i < n4; i++) { q[i] /= nq; z[i] /= nq; }
    double al = wdot(n4, w2, r, q);
    OMP_FOR
    for (long i = 0; i < n4; i++) { x[i] += al * z[i]; r[i] -= al * q[i]; }
    it++;
    k++;
    if (k == restart) {
      k = 0;
      ell_apply(o, n4, x, r);
      OMP_FOR
      for (long i = 0; i < n4; i++) r[i] = b[i] - r[i];
    }
    rel = sqrt(wdot(n4, w2, r, r)) / nb;
    if (hist) hist[it] = rel;
    if (!(rel == rel)) break;
  }
  clock_gettime(CLOCK_MONOTONIC, &ts1);
  if (seconds_per_it)
    *seconds_per_it = it ? ((double)(ts1.tv_sec - ts0.tv_sec) + 1e-9 * (double)(ts1.tv_nsec - ts0.tv_nsec)) / it : 0.0;
  o->vonly = 0;
  memcpy(o->fre1, fsave, sizeof(double) * 4 * o->ntot);
  memcpy(o->sol0, x, sizeof(double) * n4);
  if (x_out) memcpy(x_out, x, sizeof(double) * n4);
  free(b); free(x); free(r); free(w2); free(Z); free(Q); free(fsave);
  return it;
}

/* ------------------------------------------------------------------------- */
/* Post-solve pass (parallel/move3mg.c:2835-3238) on level 0.                 */
/* Stress values: sxxcalc/syycalc/szzcalc/sxycalc/sxzcalc/syzcalc in value    */
/* mode (ynval = 0; move3mg.c:3-131, 132-258, 259-387, 388-519, 520-658,      */
/* 659-795), strain rate in e, effective viscosity in nuo.                    */
/* ------------------------------------------------------------------------- */
static double post_sig(const orc_t *o, int kind, long m1, long m2, long m3, const double *x, double *e,
                       double *nuo) {
  const int X = o->mgx[0], Y = o->mgy[0], Z = o->mgz[0];
  const long XY = o->mgxy[0], p = o->mgp[0];
  const double *gx = o->mggx[0], *gy = o->mggy[0], *gz = o->mggz[0];
  long v[8];
  double nueff, val;
#define VX(i) x[4 * (i) + 1]
#define VY(i) x[4 * (i) + 2]
#define VZ(i) x[4 * (i) + 3]
  if (kind < 3) {
    /* cell normal stress, v7 = the node itself (move3mg.c:23-30, 85-101) */
    double xkf = 1.0 / (gx[m1] - gx[m1 - 1]), ykf = 1.0 / (gy[m2] - gy[m2 - 1]), zkf = 1.0 / (gz[m3] - gz[m3 - 1]);
    v[0] = p + (m3 - 1) * XY + (m1 - 1) * Y + (m2 - 1); v[1] = v[0] + 1; v[2] = v[0] + Y; v[3] = v[2] + 1;
    v[4] = v[0] + XY; v[5] = v[4] + 1; v[6] = v[4] + Y; v[7] = v[6] + 1;
    if (o->xper && m1 == X - 1) v[2] = p + (m3 - 1) * XY + m2 - 1;
    if (o->yper && m2 == Y - 1) v[1] = p + (m3 - 1) * XY + (m1 - 1) * Y;
    if (o->zper && m3 == Z - 1) v[4] = p + (m1 - 1) * Y + m2 - 1;
    nueff = o->nxx[v[7]];
    if (nueff > o->nuend) nueff = o->nuend;
    if (nueff < o->nubeg) nueff = o->nubeg;
    double dvxdx = (VX(v[2]) - VX(v[0])) * xkf;
    double dvydy = (VY(v[1]) - VY(v[0])) * ykf;
    double dvzdz = (VZ(v[4]) - VZ(v[0])) * zkf;
    double d = kind == 0 ? dvxdx : kind == 1 ? dvydy : dvzdz;
    val = d - (dvxdx + dvydy + dvzdz) / 3.0;
  } else if (kind == 3) {
    /* sxy at the node (move3mg.c:400-479) */
    double xkf = (o->xper && m1 == 0) ? 1.0 / (gx[1] - gx[0] + gx[X - 1] - gx[X - 2]) : 1.0 / (gx[m1 + 1] - gx[m1 - 1]);
    double ykf = (o->yper && m2 == 0) ? 1.0 / (gy[1] - gy[0] + gy[Y - 1] - gy[Y - 2]) : 1.0 / (gy[m2 + 1] - gy[m2 - 1]);
    v[0] = p + m3 * XY + (m1 - 1) * Y + (m2 - 1); v[1] = v[0] + 1; v[2] = v[0] + Y; v[3] = v[2] + 1;
    v[7] = v[3] + XY;
    if (o->xper && m1 == 0) { v[1] = p + m3 * XY + (X - 2) * Y + m2; v[2] = p + m3 * XY + (m2 - 1); v[3] = v[2] + 1; }
    if (o->yper && m2 == 0) { v[2] = p + m3 * XY + m1 * Y + Y - 2; v[1] = p + m3 * XY + (m1 - 1) * Y; v[3] = v[1] + Y; }
    if (o->xper && m1 == 0 && o->yper && m2 == 0) { v[1] = p + m3 * XY + (X - 2) * Y; v[2] = p + m3 * XY + Y - 2; v[3] = p + m3 * XY; }
    nueff = o->nxy[v[3]];
    if (nueff > o->nuend) nueff = o->nuend;
    if (nueff < o->nubeg) nueff = o->nubeg;
    double dvxdy = (VX(v[3]) - VX(v[2])) * ykf;
    double dvydx = (VY(v[3]) - VY(v[1])) * xkf;
    val = dvxdy + dvydx;
  } else if (kind == 4) {
    /* sxz at the node (move3mg.c:532-615) */
    double xkf = (o->xper && m1 == 0) ? 1.0 / (gx[1] - gx[0] + gx[X - 1] - gx[X - 2]) : 1.0 / (gx[m1 + 1] - gx[m1 - 1]);
    double zkf = (o->zper && m3 == 0) ? 1.0 / (gz[1] - gz[0] + gz[Z - 1] - gz[Z - 2]) : 1.0 / (gz[m3 + 1] - gz[m3 - 1]);
    v[0] = p + (m3 - 1) * XY + (m1 - 1) * Y + m2; v[2] = v[0] + Y; v[4] = v[0] + XY; v[6] = v[4] + Y;
    if (o->xper && m1 == 0) { v[4] = p + m3 * XY + (X - 2) * Y + m2; v[2] = p + (m3 - 1) * XY + m2; v[6] = v[2] + XY; }
    if (o->zper && m3 == 0) { v[2] = p + (Z - 2) * XY + m1 * Y + m2; v[4] = p + (m1 - 1) * Y + m2; v[6] = v[4] + Y; }
    if (o->xper && m1 == 0 && o->zper && m3 == 0) { v[2] = p + (Z - 2) * XY + m2; v[4] = p + (X - 2) * Y + m2; v[6] = p + m2; }
    nueff = o->nxz[v[6]];
    if (nueff > o->nuend) nueff = o->nuend;
    if (nueff < o->nubeg) nueff = o->nubeg;
    double dvxdz = (VX(v[6]) - VX(v[2])) * zkf;
    double dvzdx = (VZ(v[6]) - VZ(v[4])) * xkf;
    val = dvxdz + dvzdx;
  } else {
    /* syz at the node (move3mg.c:671-754) */
    double ykf = (o->yper && m2 == 0) ? 1.0 / (gy[1] - gy[0] + gy[Y - 1] - gy[Y - 2]) : 1.0 / (gy[m2 + 1] - gy[m2 - 1]);
    double zkf = (o->zper && m3 == 0) ? 1.0 / (gz[1] - gz[0] + gz[Z - 1] - gz[Z - 2]) : 1.0 / (gz[m3 + 1] - gz[m3 - 1]);
    v[0] = p + (m3 - 1) * XY + m1 * Y + (m2 - 1); v[1] = v[0] + 1; v[4] = v[0] + XY; v[5] = v[4] + 1;
    if (o->yper && m2 == 0) { v[1] = p + (m3 - 1) * XY + m1 * Y; v[5] = v[1] + XY; v[4] = v[5] + Y - 2; }
    if (o->zper && m3 == 0) { v[4] = p + m1 * Y + (m2 - 1); v[5] = v[4] + 1; v[1] = v[5] + (Z - 2) * XY; }
    if (o->yper && m2 == 0 && o->zper && m3 == 0) { v[5] = p + m1 * Y; v[1] = v[5] + (Z - 2) * XY; v[4] = v[5] + Y - 2; }
    nueff = o->nyz[v[5]];
    if (nueff > o->nuend) nueff = o->nuend;
    if (nueff < o->nubeg) nueff = o->nubeg;
    double dvydz = (VY(v[5]) - VY(v[1])) * zkf;
    double dvzdy = (VZ(v[5]) - VZ(v[4])) * ykf;
    val = dvydz + dvzdy;
  }
  if (e) *e = val;
  if (nuo) *nuo = nueff;
  return 2.0 * nueff * val;
}

/* which stress planes hold a value at node (m1,m2,m3) and where it is taken
   from (the periodic copies of the outer line; move3mg.c:2919-3044) */
static int post_src(const orc_t *o, int kind, long m1, long m2, long m3, long *s1, long *s2, long *s3) {
  const int X = o->mgx[0], Y = o->mgy[0], Z = o->mgz[0];
  *s1 = m1; *s2 = m2; *s3 = m3;
  if (kind < 3) return m1 > 0 && m2 > 0 && m3 > 0;
  int pa, pb;     /* periodicity of the two in-plane axes */
  long *ia, *ib;  /* their indices */
  long na, nb, nc, kc;
  if (kind == 3) { pa = o->xper; pb = o->yper; ia = s1; ib = s2; na = X; nb = Y; nc = Z; kc = m3; }
  else if (kind == 4) { pa = o->xper; pb = o->zper; ia = s1; ib = s3; na = X; nb = Z; nc = Y; kc = m2; }
  else { pa = o->yper; pb = o->zper; ia = s2; ib = s3; na = Y; nb = Z; nc = X; kc = m1; }
  if (kc >= nc - 1) return 0;
  if (pa && pb) {
    if (*ia == na - 1) *ia = 0;
    if (*ib == nb - 1) *ib = 0;
    return 1;
  }
  if (pa) {
    if (*ib <= 0 || *ib >= nb - 1) return 0;
    if (*ia == na - 1) *ia = 0;
    return 1;
  }
  if (pb) {
    if (*ia <= 0 || *ia >= na - 1) return 0;
    if (*ib == nb - 1) *ib = 0;
    return 1;
  }
  return *ia > 0 && *ia < na - 1 && *ib > 0 && *ib < nb - 1;
}

/* the stored stress plane of kind at node n (zero where the recompute writes nothing) */
static double post_stored(const orc_t *o, int kind, long n, const double *x) {
  long q = n - o->mgp[0];
  long m3 = q / o->mgxy[0], m1 = (q % o->mgxy[0]) / o->mgy[0], m2 = q % o->mgy[0];
  long s1, s2, s3;
  if (!post_src(o, kind, m1, m2, m3, &s1, &s2, &s3)) return 0.0;
  return post_sig(o, kind, s1, s2, s3, x, NULL, NULL);
}

static double nu48(const orc_t *o, const long *lo, const long *mid, const long *hi) {
  const float *nu = o->nu;
  double nueff;
  if (o->viscmod == 1)
    nueff = exp((log(nu[lo[0]]) + log(nu[lo[1]]) + log(nu[lo[2]]) + log(nu[lo[3]]) +
                 10.0 * (log(nu[mid[0]]) + log(nu[mid[1]]) + log(nu[mid[2]]) + log(nu[mid[3]])) + log(nu[hi[0]]) +
                 log(nu[hi[1]]) + log(nu[hi[2]]) + log(nu[hi[3]])) / 48.0);
  else if (o->viscmod == 2)
    nueff = 1.0 / ((1.0 / nu[lo[0]] + 1.0 / nu[lo[1]] + 1.0 / nu[lo[2]] + 1.0 / nu[lo[3]] +
                    10.0 * (1.0 / nu[mid[0]] + 1.0 / nu[mid[1]] + 1.0 / nu[mid[2]] + 1.0 / nu[mid[3]]) +
                    1.0 / nu[hi[0]] + 1.0 / nu[hi[1]] + 1.0 / nu[hi[2]] + 1.0 / nu[hi[3]]) / 48.0);
  else
    nueff = ((double)nu[lo[0]] + nu[lo[1]] + nu[lo[2]] + nu[lo[3]] +
             10.0 * ((double)nu[mid[0]] + nu[mid[1]] + nu[mid[2]] + nu[mid[3]]) + nu[hi[0]] + nu[hi[1]] + nu[hi[2]] +
             nu[hi[3]]) / 48.0;
  if (nueff > o->nuend) nueff = o->nuend;
  if (nueff < o->nubeg) nueff = o->nubeg;
  return nueff;
}

/* x/y/z-Stokes error (xstokserr/ystokserr/zstokserr with ynerr = 1,
   move3mg.c:796-905, 1011-1125, 1231-1345): the residual of the momentum
   equation from the stored stresses, and in *eps that residual times
   h^2/eta (skf1) */
static double post_stokes(const orc_t *o, int c, long m1, long m2, long m3, const double *x, double *eps) {
  const int X = o->mgx[0], Y = o->mgy[0], Z = o->mgz[0];
  const long XY = o->mgxy[0], p = o->mgp[0];
  const double *gx = o->mggx[0], *gy = o->mggy[0], *gz = o->mggz[0];
  const float *ro = o->ro;
  long v[8];
  double xkf, ykf, zkf, left, right;
  v[0] = p + m3 * XY + m1 * Y + m2; v[1] = v[0] + 1; v[2] = v[0] + Y; v[3] = v[2] + 1;
  v[4] = v[0] + XY; v[5] = v[4] + 1; v[6] = v[4] + Y; v[7] = v[6] + 1;
#define PR(i) x[4 * (i)]
  if (c == 1) {
    xkf = 2.0 / (gx[m1 + 1] - gx[m1 - 1]);
    if (o->xper && m1 == 0) xkf = 2.0 / (gx[1] - gx[0] + gx[X - 1] - gx[X - 2]);
    ykf = 1.0 / (gy[m2 + 1] - gy[m2]);
    zkf = 1.0 / (gz[m3 + 1] - gz[m3]);
    if (o->xper && m1 == 0) v[5] = p + (m3 + 1) * XY + (X - 1) * Y + m2 + 1;
    right = -o->gxk * (ro[v[0]] + ro[v[1]] + ro[v[4]] + ro[v[5]]) / 4.0;
    long lo[4] = {v[0] - Y, v[1] - Y, v[4] - Y, v[5] - Y}, mid[4] = {v[0], v[1], v[4], v[5]},
         hi[4] = {v[2], v[3], v[6], v[7]};
    for (int k = 0; k < 4; k++) if (lo[k] < 0) lo[k] = 0;  /* the reference reads before its array here */
    double nueff = nu48(o, lo, mid, hi);
    double skf1 = 1.0 / nueff * (1.0 / xkf / xkf + 1.0 / ykf / ykf + 1.0 / zkf / zkf) / 3.0;
    left = (post_stored(o, 0, v[7], x) - post_stored(o, 0, v[5], x)) * xkf;
    left += (post_stored(o, 3, v[1], x) - post_stored(o, 3, v[0], x)) * ykf;
    left += (post_stored(o, 4, v[4], x) - post_stored(o, 4, v[0], x)) * zkf;
    left -= (PR(v[7]) - PR(v[5])) * xkf;
    left = left - right;
    *eps = left * skf1;
    return left;
  }
  if (c == 2) {
    xkf = 1.0 / (gx[m1 + 1] - gx[m1]);
    ykf = 2.0 / (gy[m2 + 1] - gy[m2 - 1]);
    if (o->yper && m2 == 0) ykf = 2.0 / (gy[1] - gy[0] + gy[Y - 1] - gy[Y - 2]);
    zkf = 1.0 / (gz[m3 + 1] - gz[m3]);
    if (o->yper && m2 == 0) v[6] = p + (m3 + 1) * XY + (m1 + 1) * Y + Y - 1;
    right = -o->gyk * (ro[v[0]] + ro[v[2]] + ro[v[4]] + ro[v[6]]) / 4.0;
    long lo[4] = {v[0] - 1, v[2] - 1, v[4] - 1, v[6] - 1}, mid[4] = {v[0], v[2], v[4], v[6]},
         hi[4] = {v[1], v[3], v[5], v[7]};
    for (int k = 0; k < 4; k++) if (lo[k] < 0) lo[k] = 0;
    double nueff = nu48(o, lo, mid, hi);
    double skf1 = 1.0 / nueff * (1.0 / xkf / xkf + 1.0 / ykf / ykf + 1.0 / zkf / zkf) / 3.0;
    left = (post_stored(o, 1, v[7], x) - post_stored(o, 1, v[6], x)) * ykf;
    left += (post_stored(o, 3, v[2], x) - post_stored(o, 3, v[0], x)) * xkf;
    left += (post_stored(o, 5, v[4], x) - post_stored(o, 5, v[0], x)) * zkf;
    left -= (PR(v[7]) - PR(v[6])) * ykf;
    left = left - right;
    *eps = left * skf1;
    return left;
  }
  xkf = 1.0 / (gx[m1 + 1] - gx[m1]);
  ykf = 1.0 / (gy[m2 + 1] - gy[m2]);
  zkf = 2.0 / (gz[m3 + 1] - gz[m3 - 1]);
  if (o->zper && m3 == 0) zkf = 2.0 / (gz[1] - gz[0] + gz[Z - 1] - gz[Z - 2]);
  if (o->zper && m3 == 0) v[3] = p + (Z - 1) * XY + (m1 + 1) * Y + m2 + 1;
  right = -o->gzk * (ro[v[0]] + ro[v[1]] + ro[v[2]] + ro[v[3]]) / 4.0;
  long lo[4] = {v[0] - XY, v[1] - XY, v[2] - XY, v[3] - XY}, mid[4] = {v[0], v[1], v[2], v[3]},
       hi[4] = {v[4], v[5], v[6], v[7]};
  for (int k = 0; k < 4; k++) if (lo[k] < 0) lo[k] = 0;
  double nueff = nu48(o, lo, mid, hi);
  double skf1 = 1.0 / nueff * (1.0 / xkf / xkf + 1.0 / ykf / ykf + 1.0 / zkf / zkf) / 3.0;
  left = (post_stored(o, 2, v[7], x) - post_stored(o, 2, v[3], x)) * zkf;
  left += (post_stored(o, 4, v[2], x) - post_stored(o, 4, v[0], x)) * xkf;
  left += (post_stored(o, 5, v[1], x) - post_stored(o, 5, v[0], x)) * ykf;
  left -= (PR(v[7]) - PR(v[3])) * zkf;
  left = left - right;
  *eps = left * skf1;
  return left;
#undef PR
}

void orc_post(const orc_t *o, const double *x, double *sig, double *eps, double *out) {
  const int X = o->mgx[0], Y = o->mgy[0], Z = o->mgz[0];
  const double *gx = o->mggx[0], *gy = o->mggy[0], *gz = o->mggz[0];
  const long n0 = o->mgn[0];
  double *part = (double *)calloc((size_t)ORC_POSTN * Z, sizeof(double));
#pragma omp parallel for schedule(dynamic, 1)
  for (int m3 = 0; m3 < Z; m3++) {
    double *pp = part + (size_t)ORC_POSTN * m3;
    for (int q = 10; q < ORC_POSTN; q += 2) { pp[q] = 1e30; pp[q + 1] = -1e30; }
    for (int m1 = 0; m1 < X; m1++)
      for (int m2 = 0; m2 < Y; m2++) {
        long n = NID(o, 0, m1, m2, m3), r = 4 * n;
        /* stress and strain-rate planes; on line 0 of a periodic axis the  */
        /* normal ones are copied from the last line after the solve        */
        /* (move3mg.c:3321-3379, x then y then z)                            */
        for (int k = 0; k < 6; k++) {
          long s1, s2, s3;
          double e = 0.0, s = 0.0, nuk = 0.0;
          if (post_src(o, k, m1, m2, m3, &s1, &s2, &s3)) {
            s = post_sig(o, k, s1, s2, s3, x, &e, &nuk);
            if (k < 3) { if (nuk < pp[18]) pp[18] = nuk; if (nuk > pp[19]) pp[19] = nuk; }
          }
          if (k < 3) {
            long c1 = (o->xper && m1 == 0) ? X - 1 : m1, c2 = (o->yper && m2 == 0) ? Y - 1 : m2;
            long c3 = (o->zper && m3 == 0) ? Z - 1 : m3;
            if ((c1 != m1 || c2 != m2 || c3 != m3)) {
              e = 0.0; s = 0.0;
              if (post_src(o, k, c1, c2, c3, &s1, &s2, &s3)) s = post_sig(o, k, s1, s2, s3, x, &e, NULL);
            }
          }
          if (sig) sig[(size_t)k * n0 + (n - o->mgp[0])] = s;
          if (eps) eps[(size_t)k * n0 + (n - o->mgp[0])] = e;
        }
        /* velocity and pressure extrema (move3mg.c:2858-2900) */
        if (m1 && m2 && m3) { if (x[r] < pp[16]) pp[16] = x[r]; if (x[r] > pp[17]) pp[17] = x[r]; }
        if (m2 < Y - 1 && m3 < Z - 1) { if (x[r + 1] < pp[10]) pp[10] = x[r + 1]; if (x[r + 1] > pp[11]) pp[11] = x[r + 1]; }
        if (m1 < X - 1 && m3 < Z - 1) { if (x[r + 2] < pp[12]) pp[12] = x[r + 2]; if (x[r + 2] > pp[13]) pp[13] = x[r + 2]; }
        if (m1 < X - 1 && m2 < Y - 1) { if (x[r + 3] < pp[14]) pp[14] = x[r + 3]; if (x[r + 3] > pp[15]) pp[15] = x[r + 3]; }
        /* continuity (continerr ynerr = 1, move3mg.c:1500-1507; sums 3130-3147) */
        if (m1 && m2 && m3) {
          double xkf = 1.0 / (gx[m1] - gx[m1 - 1]), ykf = 1.0 / (gy[m2] - gy[m2 - 1]), zkf = 1.0 / (gz[m3] - gz[m3 - 1]);
          long v[8];
          v[0] = NID(o, 0, m1 - 1, m2 - 1, m3 - 1); v[1] = v[0] + 1; v[2] = v[0] + Y; v[4] = v[0] + o->mgxy[0];
          if (o->xper && m1 == X - 1) v[2] = NID(o, 0, 0, m2 - 1, m3 - 1);
          if (o->yper && m2 == Y - 1) v[1] = NID(o, 0, m1 - 1, 0, m3 - 1);
          if (o->zper && m3 == Z - 1) v[4] = NID(o, 0, m1 - 1, m2 - 1, 0);
          double leftc = (x[4 * v[2] + 1] - x[4 * v[0] + 1]) * xkf + (x[4 * v[1] + 2] - x[4 * v[0] + 2]) * ykf +
                         (x[4 * v[4] + 3] - x[4 * v[0] + 3]) * zkf;
          double skf = (1.0 / xkf + 1.0 / ykf + 1.0 / zkf) / 3.0;
          double cv = (leftc - 0.0) * skf;
          if (!o->bondm[r]) { pp[4] += cv * cv; pp[5] += 1.0; }
          else { pp[6] += cv * cv; pp[7] += 1.0; }
        }
        /* momentum and velocity boundary rows (move3mg.c:3149-3227) */
        const int inr[4] = {0, m2 < Y - 1 && m3 < Z - 1, m1 < X - 1 && m3 < Z - 1, m1 < X - 1 && m2 < Y - 1};
        for (int c = 1; c < 4; c++) {
          if (!inr[c]) continue;
          long rr = r + c;
          if (!o->bondm[rr]) {
            double e1;
            double iv = post_stokes(o, c, m1, m2, m3, x, &e1);
            pp[0] += iv * iv; pp[1] += e1 * e1; pp[2] += e1 * e1 * iv * iv; pp[3] += 1.0;
          } else {
            int b = o->bondm[rr];
            double lb = x[rr] - o->bondv1[2 * b + 0];
            if (o->bondn1[b]) lb -= o->bondv1[2 * b + 1] * x[o->bondn1[b] - 1];
            pp[8] += lb * lb; pp[9] += 1.0;
          }
        }
      }
  }
  for (int q = 0; q < 10; q++) out[q] = 0.0;
  for (int q = 10; q < ORC_POSTN; q += 2) { out[q] = 1e30; out[q + 1] = -1e30; }
  for (int m3 = 0; m3 < Z; m3++) {
    const double *pp = part + (size_t)ORC_POSTN * m3;
    for (int q = 0; q < 10; q++) out[q] += pp[q];
    for (int q = 10; q < ORC_POSTN; q += 2) {
      if (pp[q] < out[q]) out[q] = pp[q];
      if (pp[q + 1] > out[q + 1]) out[q + 1] = pp[q + 1];
    }
  }
  free(part);
#undef VX
#undef VY
#undef VZ
}

/* ------------------------------------------------------------------------- */
/* Marker advection: movemark (parallel/mark3mg.c:174-629) with vxyzcalc1     */
/* (:2963-3236) and epscalc1 (:3244-3760), walled and periodic boxes.  Along  */
/* each axis a quantity sits on grid lines or on cell centres, and the        */
/* reference brackets the point with one of five rules (names below);        */
/* weights are 1 - |distance| / spacing, accumulated z outer, x, y inner, as */
/* the reference's loops run.                                                */
/* ------------------------------------------------------------------------- */
enum { AX_VN, AX_VC, AX_CELL, AX_SN, AX_SC };
/* AX_VN   velocity component along its own axis: lines m, m+1 (m clamped to
           0..n-2), the point itself                                (:3004-3007)
   AX_VC   velocity across an axis: centres of cells below and above (:3008-3023);
           periodic: below centre 0 or above centre n-3 the pair is (n-2, 0)
           across the wrap, the point shifted by the box size (:2992-3001)
   AX_CELL normal strain rates: cell centres m, m+1 (1..n-2)        (:3272-3285);
           periodic: below centre 1 or above centre n-2 the pair is (n-1, 1)
           across the wrap (:3282-3291)
   AX_SN   shear strain rates along a line axis: lines 1..n-3, point clamped
           to the clamped line                                      (:3390-3398);
           periodic: lines 0..n-2 (:3385-3389)
   AX_SC   shear strain rates across: centres, the lower bracket clamped from
           below at 1 to 0 with the point on centre 0               (:3411-3427);
           periodic: below centre 0 the pair is (n-2, 0) across the wrap, the
           rest as walled (the wrap test there is > n-2, never true) (:3410-3416)
   Across the wrap the second node's distance is taken from the point minus
   the box size (:3031, 3044, 3336-3351).  dvar selects the spacing the vz
   interpolation takes across a periodic x wrap, (g[1] + g[n-1] - g[n-2]) / 2
   (:3192), where every other wrap takes (g[1] - g[0] + g[n-1] - g[n-2]) / 2. */
typedef struct { long i0, i1; double w0, w1; } axw_t;

static axw_t axis_weights_p(const double *g, int n, long m, double p, int kind, int per, double size, int dvar,
                            int ylow) {
  axw_t r;
  double P = p, d, t0, t1;
  long i0 = m, i1;
  int wrap = 0;
  if (kind == AX_VN) {
    if (i0 < 0) i0 = 0;
    if (i0 > n - 2) i0 = n - 2;
    i1 = i0 + 1;
    d = g[i0 + 1] - g[i0];
    t0 = (g[i0] - p) / d; t1 = (g[i1] - p) / d;
  } else if (kind == AX_SN) {
    if (per) {
      if (i0 < 0) { i0 = 0; P = ylow ? g[0] : 0.0; }
      if (i0 > n - 2) { i0 = n - 2; P = g[i0 + 1]; }
    } else {
      if (i0 < 1) { i0 = 1; P = g[i0]; }
      if (i0 > n - 3) { i0 = n - 3; P = g[i0 + 1]; }
    }
    i1 = i0 + 1;
    d = g[i0 + 1] - g[i0];
    t0 = (g[i0] - P) / d; t1 = (g[i1] - P) / d;
  } else if (kind == AX_CELL) {
    if (P > (g[i0] + g[i0 + 1]) / 2.0) i0 += 1;
    if (per && (i0 < 1 || i0 > n - 2)) {
      if (i0 < 1) P = size + p;
      i0 = n - 1; i1 = 1; wrap = 1;
      d = (g[1] - g[0] + g[n - 1] - g[n - 2]) / 2.0;
    } else {
      if (i0 < 1) { i0 = 1; P = (g[i0] + g[i0 - 1]) / 2.0; }
      if (i0 > n - 2) { i0 = n - 2; P = (g[i0] + g[i0 + 1]) / 2.0; }
      i1 = i0 + 1;
      d = (g[i0 + 1] - g[i0 - 1]) / 2.0;
    }
    t0 = ((g[i0] + g[i0 - 1]) / 2.0 - P) / d;
    t1 = ((g[i1] + g[i1 - 1]) / 2.0 - (wrap ? P - size : P)) / d;
  } else {  /* AX_VC, AX_SC */
    if (P < (g[i0] + g[i0 + 1]) / 2.0) i0 -= 1;
    if (per && (i0 < 0 || i0 > (kind == AX_VC ? n - 3 : n - 2))) {
      if (i0 < 0) P = size + p;
      i0 = n - 2; i1 = 0; wrap = 1;
      d = dvar ? (g[1] + g[n - 1] - g[n - 2]) / 2.0 : (g[1] - g[0] + g[n - 1] - g[n - 2]) / 2.0;
    } else {
      if (kind == AX_VC ? i0 < 0 : i0 < 1) { i0 = 0; P = (g[i0] + g[i0 + 1]) / 2.0; }
      if (i0 > n - 3) { i0 = n - 3; P = (g[i0 + 1] + g[i0 + 2]) / 2.0; }
      i1 = i0 + 1;
      d = (g[i0 + 2] - g[i0]) / 2.0;
    }
    t0 = ((g[i0] + g[i0 + 1]) / 2.0 - P) / d;
    t1 = ((g[i1] + g[i1 + 1]) / 2.0 - (wrap ? P - size : P)) / d;
  }
  r.i0 = i0;
  r.i1 = i1;
  r.w0 = 1.0 - (-t0);
  r.w1 = 1.0 - t1;
  return r;
}

/* sum over the 8 bracketing nodes of val(node) * wx * wy * wz, z outer, x, y inner */
static double interp8(const orc_t *o, axw_t ax, axw_t ay, axw_t az, const double *vals, int stride, int comp,
                      int as_float) {
  const long Y = o->mgy[0], XY = o->mgxy[0];
  double acc = 0.0;
  for (int kz = 0; kz < 2; kz++) {
    const double wz = kz ? az.w1 : az.w0;
    for (int kx = 0; kx < 2; kx++) {
      const double wx = kx ? ax.w1 : ax.w0;
      for (int ky = 0; ky < 2; ky++) {
        const double wy = ky ? ay.w1 : ay.w0;
        const long n = (kz ? az.i1 : az.i0) * XY + (kx ? ax.i1 : ax.i0) * Y + (ky ? ay.i1 : ay.i0);
        double v = vals[(size_t)n * stride + comp];
        if (as_float) v = (double)(float)v;
        acc += v * (wx * wy * wz);
      }
    }
  }
  return acc;
}

void orc_interp(const orc_t *o, const double *x, const double *eps, double px, double py, double pz, double *v,
                double *e) {
  const int X = o->mgx[0], Y = o->mgy[0], Z = o->mgz[0];
  const double *gx = o->mggx[0], *gy = o->mggy[0], *gz = o->mggz[0];
  const long n0 = o->mgn[0];
  const double xs = gx[X - 1] - gx[0], ys = gy[Y - 1] - gy[0], zs = gz[Z - 1] - gz[0];
  const int px_ = o->xper, py_ = o->yper, pz_ = o->zper;
  /* out of grid: onto the box (:2975-2977, :3255-3257) */
  px = px < 0 ? 0 : (px > xs ? xs : px);
  py = py < 0 ? 0 : (py > ys ? ys : py);
  pz = pz < 0 ? 0 : (pz > zs ? zs : pz);
  const long m1 = serch(gx, X, px), m2 = serch(gy, Y, py), m3 = serch(gz, Z, pz);
#define AW(g, n, m, p, k, per, sz, dv, yl) axis_weights_p(g, n, m, p, k, per, sz, dv, yl)
  if (v) {
    v[0] = interp8(o, AW(gx, X, m1, px, AX_VN, px_, xs, 0, 0), AW(gy, Y, m2, py, AX_VC, py_, ys, 0, 0),
                   AW(gz, Z, m3, pz, AX_VC, pz_, zs, 0, 0), x, 4, 1, 1);
    v[1] = interp8(o, AW(gx, X, m1, px, AX_VC, px_, xs, 0, 0), AW(gy, Y, m2, py, AX_VN, py_, ys, 0, 0),
                   AW(gz, Z, m3, pz, AX_VC, pz_, zs, 0, 0), x, 4, 2, 1);
    v[2] = interp8(o, AW(gx, X, m1, px, AX_VC, px_, xs, 1, 0), AW(gy, Y, m2, py, AX_VC, py_, ys, 0, 0),
                   AW(gz, Z, m3, pz, AX_VN, pz_, zs, 0, 0), x, 4, 3, 1);
  }
  if (e) {
    const axw_t cx = AW(gx, X, m1, px, AX_CELL, px_, xs, 0, 0), cy = AW(gy, Y, m2, py, AX_CELL, py_, ys, 0, 0),
                cz = AW(gz, Z, m3, pz, AX_CELL, pz_, zs, 0, 0);
    for (int k = 0; k < 3; k++) e[k] = interp8(o, cx, cy, cz, eps + (size_t)k * n0, 1, 0, 1);
    e[3] = interp8(o, AW(gx, X, m1, px, AX_SN, px_, xs, 0, 0), AW(gy, Y, m2, py, AX_SN, py_, ys, 0, 1),
                   AW(gz, Z, m3, pz, AX_SC, pz_, zs, 0, 0), eps + 3 * (size_t)n0, 1, 0, 1);
    e[4] = interp8(o, AW(gx, X, m1, px, AX_SN, px_, xs, 0, 0), AW(gy, Y, m2, py, AX_SC, py_, ys, 0, 0),
                   AW(gz, Z, m3, pz, AX_SN, pz_, zs, 0, 0), eps + 4 * (size_t)n0, 1, 0, 1);
    e[5] = interp8(o, AW(gx, X, m1, px, AX_SC, px_, xs, 0, 0), AW(gy, Y, m2, py, AX_SN, py_, ys, 0, 1),
                   AW(gz, Z, m3, pz, AX_SN, pz_, zs, 0, 0), eps + 5 * (size_t)n0, 1, 0, 1);
  }
#undef AW
}

/* second invariant of the strain rate (movemark, mark3mg.c:229): sqrt of
   (exx^2+eyy^2+ezz^2)/2 + exy^2+exz^2+eyz^2; pow(., 0.5) there, sqrt here */
static double eii_of(const double *e) {
  return sqrt(0.5 * (e[0] * e[0] + e[1] * e[1] + e[2] * e[2]) + e[3] * e[3] + e[4] * e[4] + e[5] * e[5]);
}

/* periodic wrap of an RK stage position (mark3mg.c:337-355): one period
   back into [0, size]; the z test compares with the x size as the reference
   does (:353); 1 if the point is still outside (the reference exits) */
static int wrap_stage(double *P, double size, double cmp) {
  if (*P < 0) *P = size + *P;
  if (*P > cmp) *P = *P - size;
  return *P < 0 || *P > size;
}

/* the same for a marker coordinate after the step (:468-485) */
static int wrap_mark(float *m, double size) {
  if (*m < 0) *m = (float)(size + *m);
  if (*m > size) *m = (float)(*m - size);
  return *m < 0 || *m > size;
}

int orc_move_markers(const orc_t *o, const double *x, const double *eps, long nmark, float *mx, float *my,
                     float *mz, float *marke, float *markk, const int *markt, float *eii, double timestep,
                     int markmod, int outgrid) {
  const int X = o->mgx[0], Y = o->mgy[0], Z = o->mgz[0];
  const double xs = o->mggx[0][X - 1] - o->mggx[0][0], ys = o->mggy[0][Y - 1] - o->mggy[0][0],
               zs = o->mggz[0][Z - 1] - o->mggz[0][0];
  const int xp = o->xper, yp = o->yper, zp = o->zper;
  long bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
  for (long i = 0; i < nmark; i++) {
    const int t = markt ? markt[i] : 0;
    const int inside = mx[i] >= 0 && my[i] >= 0 && mz[i] >= 0 && (double)mx[i] <= xs && (double)my[i] <= ys &&
                       (double)mz[i] <= zs;
    if (!(inside || (outgrid != 1 && (t < 50 || t >= 100)))) continue;
    const int wantE = marke && marke[i] > 0;
    double v1[3], e[6], ee1, ee0, vv[3];
    orc_interp(o, x, eps, mx[i], my[i], mz[i], v1, e);
    ee1 = eii_of(e);
    if (eii) eii[i] = (float)ee1;
    if (markmod == 1) {
      vv[0] = v1[0]; vv[1] = v1[1]; vv[2] = v1[2];
      ee0 = ee1;
    } else {
      double vs[4][3], es[4] = {ee1, 0, 0, 0};
      for (int k = 0; k < 3; k++) vs[0][k] = v1[k];
      for (int s = 1; s < 4; s++) {
        /* stage points (:333-335, 363-365, 393-395): x + v dt / 2, x + v dt / 2, x + v dt */
        double P[3];
        for (int k = 0; k < 3; k++) {
          const double mk = k == 0 ? mx[i] : (k == 1 ? my[i] : mz[i]);
          P[k] = s < 3 ? mk + vs[s - 1][k] * timestep / 2.0 : mk + vs[s - 1][k] * timestep;
        }
        if (xp) bad += wrap_stage(&P[0], xs, xs);
        if (yp) bad += wrap_stage(&P[1], ys, ys);
        if (zp) bad += wrap_stage(&P[2], zs, xs);
        orc_interp(o, x, eps, P[0], P[1], P[2], vs[s], wantE ? e : NULL);
        if (wantE) es[s] = eii_of(e);
      }
      for (int k = 0; k < 3; k++) vv[k] = (vs[0][k] + 2.0 * vs[1][k] + 2.0 * vs[2][k] + vs[3][k]) / 6.0;
      ee0 = (es[0] + 2.0 * es[1] + 2.0 * es[2] + es[3]) / 6.0;
    }
    if (outgrid == 2) {  /* orthogonal motion only outside the box (mark3mg.c:433-457) */
      if (xp) {
        if (my[i] < 0 || (double)my[i] > ys) vv[0] = vv[2] = 0;
        if (mz[i] < 0 || (double)mz[i] > zs) vv[0] = vv[1] = 0;
      } else if (yp) {
        if (mx[i] < 0 || (double)mx[i] > xs) vv[1] = vv[2] = 0;
        if (mz[i] < 0 || (double)mz[i] > zs) vv[0] = vv[1] = 0;
      } else if (zp) {
        if (mx[i] < 0 || (double)mx[i] > xs) vv[1] = vv[2] = 0;
        if (my[i] < 0 || (double)my[i] > zs) vv[0] = vv[2] = 0;
      } else {
        if (mx[i] < 0 || (double)mx[i] > xs) vv[1] = vv[2] = 0;
        if (my[i] < 0 || (double)my[i] > ys) vv[0] = vv[2] = 0;
        if (mz[i] < 0 || (double)mz[i] > zs) vv[0] = vv[1] = 0;
      }
    }
    mx[i] += timestep * vv[0];
    my[i] += timestep * vv[1];
    mz[i] += timestep * vv[2];
    if (xp) bad += wrap_mark(&mx[i], xs);
    if (yp) bad += wrap_mark(&my[i], ys);
    if (zp) bad += wrap_mark(&mz[i], zs);
    if (t < 100) {
      if (wantE) marke[i] += (float)(ee0 * timestep);
      if (markk && (mx[i] < 0 || my[i] < 0 || mz[i] < 0 || (double)mx[i] > xs || (double)my[i] > ys ||
                    (double)mz[i] > zs))
        markk[i] = 0;
    }
  }
  return (int)(bad > 0x7fffffff ? 0x7fffffff : bad);
}

/* threads the OpenMP loops of this file use (1 without OpenMP) */
int orc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
