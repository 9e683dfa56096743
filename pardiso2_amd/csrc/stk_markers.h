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

// Sums of one node over the markers of its 8 cells: FP32 property sums and
// FP64 weights, index 0 nu, 1 ro, 2 nxx, 3 nxy, 4 nxz, 5 nyz.
struct MSum {
  float a[6];
  double w[6];
};

struct MGather {
  int X, Y, Z;
  const double *gx, *gy, *gz;
  const int *start;
  const float *mx, *my, *mz, *mnu, *mro;
  double nubeg, nuend, nukoef;
  int vm;
};

__device__ __forceinline__ void msum_node(const MGather &q, int m1, int m2, int m3, MSum &s) {
  const int X = q.X, Y = q.Y, Z = q.Z;
  const double *gx = q.gx, *gy = q.gy, *gz = q.gz;
  float anu = 0.f, aro = 0.f, axx = 0.f, axy = 0.f, axz = 0.f, ayz = 0.f;
  double wnu = 0, wro = 0, wxx = 0, wxy = 0, wxz = 0, wyz = 0;
  const int vm = q.vm;
  for (int c3 = m3 - 1; c3 <= m3; c3++) {
    if (c3 < 0 || c3 > Z - 2) continue;
    for (int c1 = m1 - 1; c1 <= m1; c1++) {
      if (c1 < 0 || c1 > X - 2) continue;
      for (int c2 = m2 - 1; c2 <= m2; c2++) {
        if (c2 < 0 || c2 > Y - 2) continue;
        int cell = (c3 * (X - 1) + c1) * (Y - 1) + c2;
        const double dx = gx[c1 + 1] - gx[c1], dy = gy[c2 + 1] - gy[c2], dz = gz[c3 + 1] - gz[c3];
        for (int i = q.start[cell]; i < q.start[cell + 1]; i++) {
          double mnuv = q.mnu[i];
          if (mnuv < q.nubeg) mnuv = q.nubeg;
          if (mnuv > q.nuend) mnuv = q.nuend;
          double ddz = (gz[m3] - (double)q.mz[i]) / dz;
          ddz = 1.0 - fabs(ddz);
          double ddx = (gx[m1] - (double)q.mx[i]) / dx;
          ddx = 1.0 - fabs(ddx);
          double ddy = (gy[m2] - (double)q.my[i]) / dy;
          ddy = 1.0 - fabs(ddy);
          double swt = ddx * ddy * ddz, swt1;
          if (ddx > q.nukoef && ddy > q.nukoef && ddz > q.nukoef) {
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
            aro = (float)((double)aro + (double)q.mro[i] * swt1);
            wro += swt1;
          }
        }
      }
    }
  }
  s.a[0] = anu; s.a[1] = aro; s.a[2] = axx; s.a[3] = axy; s.a[4] = axz; s.a[5] = ayz;
  s.w[0] = wnu; s.w[1] = wro; s.w[2] = wxx; s.w[3] = wxy; s.w[4] = wxz; s.w[5] = wyz;
}

// Properties folded along each periodic axis (those on its grid lines):
// x nu ro nxy nxz, y nu ro nxy nyz, z nu ro nxz nyz.
constexpr unsigned kFoldX = 0x1b, kFoldY = 0x2b, kFoldZ = 0x33;

__device__ __noinline__ void msum_node_far(const MGather &q, int m1, int m2, int m3, MSum &s) {
  msum_node(q, m1, m2, m3, s);
}

// s += t on the properties of mask (FP32 sums, FP64 weights: the reference's
// sol0[m] += sol0[m4] and nu[nodenum+m] += nu[nodenum+m4], mark3mg.c:893-955)
__device__ __forceinline__ void msum_add(MSum &s, const MSum &t, unsigned mask) {
#pragma unroll
  for (int k = 0; k < 6; k++)
    if ((mask >> k) & 1) {
      s.a[k] = s.a[k] + t.a[k];
      s.w[k] += t.w[k];
    }
}

// sums of node (m1, m2, m3) after the folds of the periodic axes up to AX
// (0: x only, 1: x then y, 2: x, y, z), each fold adding the last line's
// (already folded) sums to line 0's
template <int AX>
__device__ __forceinline__ void msum_fold(const MGather &q, int per, int m1, int m2, int m3, MSum &s, bool far) {
  if constexpr (AX < 0) {
    if (far) msum_node_far(q, m1, m2, m3, s);
    else msum_node(q, m1, m2, m3, s);
  } else {
    msum_fold<AX - 1>(q, per, m1, m2, m3, s, far);
    const int n = AX == 0 ? q.X : (AX == 1 ? q.Y : q.Z);
    const int m = AX == 0 ? m1 : (AX == 1 ? m2 : m3);
    if (!((per >> AX) & 1) || m != 0) return;
    MSum t;
    msum_fold<AX - 1>(q, per, AX == 0 ? n - 1 : m1, AX == 1 ? n - 1 : m2, AX == 2 ? n - 1 : m3, t, true);
    msum_add(s, t, AX == 0 ? kFoldX : (AX == 1 ? kFoldY : kFoldZ));
  }
}

__device__ __forceinline__ float mnorm(float prev, float a, double w, int vm, double nubeg, double nuend) {
  float v = prev;
  if (w > 0) { v = (float)(a / w); if (vm == 1) v = (float)exp((double)v); if (vm == 2) v = (float)(1.0 / v); }
  if (v < nubeg) v = (float)nubeg;
  if (v > nuend) v = (float)nuend;
  return v;
}

__global__ void k_markers_gather(MGather q, int per, float *nu, float *ro, float *nxx, float *nxy, float *nxz,
                                 float *nyz) {
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int X = q.X, Y = q.Y, Z = q.Z, XY = X * Y;
  if (n >= (long)XY * Z) return;
  const int m3 = (int)(n / XY), rem = (int)(n - (long)m3 * XY), m1 = rem / Y, m2 = rem - m1 * Y;
  MSum s;
  if (per) msum_fold<2>(q, per, m1, m2, m3, s, false);
  else msum_node(q, m1, m2, m3, s);
  // normalise, back-transform, clamp (mark3mg.c:962-1031); nodes without
  // markers keep their previous value before the clamp
  const int vm = q.vm;
  nu[n] = mnorm(nu[n], s.a[0], s.w[0], vm, q.nubeg, q.nuend);
  nxx[n] = mnorm(nxx[n], s.a[2], s.w[2], vm, q.nubeg, q.nuend);
  nxy[n] = mnorm(nxy[n], s.a[3], s.w[3], vm, q.nubeg, q.nuend);
  nxz[n] = mnorm(nxz[n], s.a[4], s.w[4], vm, q.nubeg, q.nuend);
  nyz[n] = mnorm(nyz[n], s.a[5], s.w[5], vm, q.nubeg, q.nuend);
  if (s.w[1] > 0) ro[n] = (float)(s.a[1] / s.w[1]);
}

// the last line of each periodic axis takes line 0's values of the
// properties folded along it (mark3mg.c:1032-1111; applied x, y, z in turn,
// each copy reads the state the previous one left, which composes to one
// source node per property); sources are never copied onto themselves
__global__ void k_markers_wrap(int X, int Y, int Z, int per, float *nu, float *ro, float *nxy, float *nxz,
                               float *nyz) {
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int XY = X * Y;
  if (n >= (long)XY * Z) return;
  const int m3 = (int)(n / XY), rem = (int)(n - (long)m3 * XY), m1 = rem / Y, m2 = rem - m1 * Y;
  const bool ex = (per & 1) && m1 == X - 1, ey = (per & 2) && m2 == Y - 1, ez = (per & 4) && m3 == Z - 1;
  if (!ex && !ey && !ez) return;
  auto src = [&](unsigned folds) {
    const int s1 = (ex && (folds & 1)) ? 0 : m1, s2 = (ey && (folds & 2)) ? 0 : m2, s3 = (ez && (folds & 4)) ? 0 : m3;
    return (long)s3 * XY + (long)s1 * Y + s2;
  };
  // axes each property is copied along: nu, ro x y z; nxy x y; nxz x z; nyz y z
  const long snr = src(7), sxy = src(3), sxz = src(5), syz = src(6);
  if (snr != n) { nu[n] = nu[snr]; ro[n] = ro[snr]; }
  if (sxy != n) nxy[n] = nxy[sxy];
  if (sxz != n) nxz[n] = nxz[sxz];
  if (syz != n) nyz[n] = nyz[syz];
}

}  // namespace stk
