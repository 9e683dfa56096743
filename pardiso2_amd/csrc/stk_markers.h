// stk_markers.h -- marker -> node averaging on the device (restates
// ronurecalc, parallel/mark3mg.c:630-1132, for per-marker viscosity/density).
//
// Markers are binned by cell with a stable radix sort, then every node
// gathers the markers of the 8 cells around it, cells in (z, x, y) order and
// markers of a cell in input order.  For markers stored cell by cell (the
// layout in3mg.c:173-212 produces) this is the reference's single-thread
// accumulation order, so the FP32 sums round identically.
#pragma once
#include <hip/hip_runtime.h>

namespace stk {

// cell search of m1serch/m2serch/m3serch (mark3mg.c:1138-1205)
__device__ __forceinline__ int serch(const double *g, int n, double x) {
  int a = 0, b = n - 1, m;
  do {
    m = (a + b) / 2;
    if (g[m] > x) b = m; else a = m;
  } while ((b - a) > 1);
  if (a > n - 2) a = n - 2;
  return a;
}

__global__ void k_marker_cells(long nm, const float *__restrict__ mx, const float *__restrict__ my,
                               const float *__restrict__ mz, const double *__restrict__ gx,
                               const double *__restrict__ gy, const double *__restrict__ gz, int X, int Y,
                               int Z, int *key, int *val) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= nm) return;
  float x = mx[i], y = my[i], z = mz[i];
  int k = 0x7fffffff;  // outside: sorted to the end and ignored
  if (x > 0 && y > 0 && z > 0 && x < gx[X - 1] && y < gy[Y - 1] && z < gz[Z - 1]) {
    int c1 = serch(gx, X, x), c2 = serch(gy, Y, y), c3 = serch(gz, Z, z);
    k = (c3 * (X - 1) + c1) * (Y - 1) + c2;
  }
  key[i] = k;
  val[i] = (int)i;
}

__global__ void k_marker_permute(long nm, const int *__restrict__ idx, const float *__restrict__ a0,
                                 const float *__restrict__ a1, const float *__restrict__ a2,
                                 const float *__restrict__ a3, const float *__restrict__ a4, float *o0, float *o1,
                                 float *o2, float *o3, float *o4) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= nm) return;
  int j = idx[i];
  o0[i] = a0[j]; o1[i] = a1[j]; o2[i] = a2[j]; o3[i] = a3[j]; o4[i] = a4[j];
}

// first marker of every cell in the sorted key array (cells with no marker
// get the start of the next occupied cell)
__global__ void k_cell_start(long nm, const int *__restrict__ key, int ncell, int *start) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i > nm) return;
  int kc = (i < nm) ? key[i] : 0x7fffffff;
  int kp = (i > 0) ? key[i - 1] : -1;
  if (kc == kp) return;
  int lo = kp + 1, hi = (kc < ncell) ? kc : ncell;
  for (int c = lo; c <= hi; c++) start[c] = (int)i;
}

__device__ __forceinline__ double vlaw(double mnu, double w, int vm) {
  if (vm == 1) return log(mnu) * w;
  if (vm == 2) return 1.0 / mnu * w;
  return mnu * w;
}

__global__ void k_markers_gather(int X, int Y, int Z, const double *__restrict__ gx,
                                 const double *__restrict__ gy, const double *__restrict__ gz,
                                 const int *__restrict__ start, const float *__restrict__ mx,
                                 const float *__restrict__ my, const float *__restrict__ mz,
                                 const float *__restrict__ mnu, const float *__restrict__ mro, double nubeg,
                                 double nuend, double nukoef, int vm, float *nu, float *ro, float *nxx,
                                 float *nxy, float *nxz, float *nyz) {
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int XY = X * Y;
  if (n >= (long)XY * Z) return;
  const int m3 = (int)(n / XY), rem = (int)(n - (long)m3 * XY), m1 = rem / Y, m2 = rem - m1 * Y;
  float anu = 0.f, aro = 0.f, axx = 0.f, axy = 0.f, axz = 0.f, ayz = 0.f;
  double wnu = 0, wro = 0, wxx = 0, wxy = 0, wxz = 0, wyz = 0;
  for (int c3 = m3 - 1; c3 <= m3; c3++) {
    if (c3 < 0 || c3 > Z - 2) continue;
    for (int c1 = m1 - 1; c1 <= m1; c1++) {
      if (c1 < 0 || c1 > X - 2) continue;
      for (int c2 = m2 - 1; c2 <= m2; c2++) {
        if (c2 < 0 || c2 > Y - 2) continue;
        int cell = (c3 * (X - 1) + c1) * (Y - 1) + c2;
        const double dx = gx[c1 + 1] - gx[c1], dy = gy[c2 + 1] - gy[c2], dz = gz[c3 + 1] - gz[c3];
        for (int i = start[cell]; i < start[cell + 1]; i++) {
          double mnuv = mnu[i];
          if (mnuv < nubeg) mnuv = nubeg;
          if (mnuv > nuend) mnuv = nuend;
          double ddz = (gz[m3] - (double)mz[i]) / dz;
          ddz = 1.0 - fabs(ddz);
          double ddx = (gx[m1] - (double)mx[i]) / dx;
          ddx = 1.0 - fabs(ddx);
          double ddy = (gy[m2] - (double)my[i]) / dy;
          ddy = 1.0 - fabs(ddy);
          double swt = ddx * ddy * ddz, swt1;
          if (ddx > nukoef && ddy > nukoef && ddz > nukoef) {
            anu = (float)((double)anu + vlaw(mnuv, swt, vm));
            wnu += swt;
          }
          if (m1 > c1 && m2 > c2 && m3 > c3) {
            swt1 = 8.0 * (0.5 - fabs(ddx - 0.5)) * (0.5 - fabs(ddy - 0.5)) * (0.5 - fabs(ddz - 0.5));
            axx = (float)((double)axx + vlaw(mnuv, swt1, vm));
            wxx += swt1;
          }
          if (m3 == c3 && ddx > 0.5 && ddy > 0.5) {
            swt1 = 8.0 * (ddx - 0.5) * (ddy - 0.5) * (0.5 - fabs(ddz - 0.5));
            axy = (float)((double)axy + vlaw(mnuv, swt1, vm));
            wxy += swt1;
          }
          if (m2 == c2 && ddx > 0.5 && ddz > 0.5) {
            swt1 = 8.0 * (ddx - 0.5) * (0.5 - fabs(ddy - 0.5)) * (ddz - 0.5);
            axz = (float)((double)axz + vlaw(mnuv, swt1, vm));
            wxz += swt1;
          }
          if (m1 == c1 && ddy > 0.5 && ddz > 0.5) {
            swt1 = 8.0 * (0.5 - fabs(ddx - 0.5)) * (ddy - 0.5) * (ddz - 0.5);
            ayz = (float)((double)ayz + vlaw(mnuv, swt1, vm));
            wyz += swt1;
          }
          if (ddx > 0.5 && ddy > 0.5 && ddz > 0.5) {
            swt1 = 8.0 * (ddx - 0.5) * (ddy - 0.5) * (ddz - 0.5);
            aro = (float)((double)aro + (double)mro[i] * swt1);
            wro += swt1;
          }
        }
      }
    }
  }
  // normalise, back-transform, clamp (mark3mg.c:962-1031); nodes without
  // markers keep their previous value before the clamp
  float v;
  v = nu[n];
  if (wnu > 0) { v = (float)(anu / wnu); if (vm == 1) v = (float)exp((double)v); if (vm == 2) v = (float)(1.0 / v); }
  if (v < nubeg) v = (float)nubeg;
  if (v > nuend) v = (float)nuend;
  nu[n] = v;
  v = nxx[n];
  if (wxx > 0) { v = (float)(axx / wxx); if (vm == 1) v = (float)exp((double)v); if (vm == 2) v = (float)(1.0 / v); }
  if (v < nubeg) v = (float)nubeg;
  if (v > nuend) v = (float)nuend;
  nxx[n] = v;
  v = nxy[n];
  if (wxy > 0) { v = (float)(axy / wxy); if (vm == 1) v = (float)exp((double)v); if (vm == 2) v = (float)(1.0 / v); }
  if (v < nubeg) v = (float)nubeg;
  if (v > nuend) v = (float)nuend;
  nxy[n] = v;
  v = nxz[n];
  if (wxz > 0) { v = (float)(axz / wxz); if (vm == 1) v = (float)exp((double)v); if (vm == 2) v = (float)(1.0 / v); }
  if (v < nubeg) v = (float)nubeg;
  if (v > nuend) v = (float)nuend;
  nxz[n] = v;
  v = nyz[n];
  if (wyz > 0) { v = (float)(ayz / wyz); if (vm == 1) v = (float)exp((double)v); if (vm == 2) v = (float)(1.0 / v); }
  if (v < nubeg) v = (float)nubeg;
  if (v > nuend) v = (float)nuend;
  nyz[n] = v;
  if (wro > 0) ro[n] = (float)(aro / wro);
}

}  // namespace stk
