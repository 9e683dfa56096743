// stk_post.h -- post-solve pass on the finest level (gfx950): the stress and
// strain-rate recompute of vpiterate (parallel/move3mg.c:2908-3046) and the
// reference's own convergence norms stoksum/stoksum1/stoksum2/contsum/
// contsumb/bondsum (parallel/move3mg.c:3090-3238).
//
// Both run in FP64 on the solution in its reference layout (4 unknowns per
// node).  The stress values follow the value mode (ynval = 0) of sxxcalc ..
// syzcalc operation for operation; the momentum errors take the stresses of
// their neighbour cells and edges as the reference's stored arrays hold them
// (zero where the recompute writes nothing), so the norms need no stored
// stress planes.  Periodic axes follow the reference's branches: the
// stencils wrap (move3mg.c:27-29, 400-429, 834-847, 1481-1483) and a shear
// stress on the last line of a periodic axis is the one on line 0
// (move3mg.c:2933-3044).
#pragma once
#include <hip/hip_runtime.h>
#include "stk_kernels.h"

namespace stk {

// sums and extrema of the norm pass: see kPost* below
constexpr int kPostN = 20;
constexpr int kPostBlocks = 1024;  // fixed grid: the sums' order does not depend on the level size

struct PostGeo {
  int X, Y, Z, XY;
  const double *iAx, *iBx, *iAy, *iBy, *iAz, *iBz;  // 1/(g[i]-g[i-1]), 1/(g[i+1]-g[i-1])
  const double *ec, *exy, *exz, *eyz;               // clamped nxx, nxy, nxz, nyz of level 0
  const float *nu;                                   // node viscosity of level 0
  double nubeg, nuend;
  int vm;   // viscmod
  int per;  // periodic axes (bit 0 x, bit 1 y, bit 2 z)
  __device__ __forceinline__ long nid(int m1, int m2, int m3) const { return (long)m3 * XY + (long)m1 * Y + m2; }
};

// Normal stress of the cell whose upper corner is node (a1, a2, a3):
// sxxcalc/syycalc/szzcalc value mode (move3mg.c:85-101, 210-226, 337-353).
// AX 0/1/2 = xx/yy/zz.  Zero outside the recompute's range (a cell needs
// a1, a2, a3 >= 1, move3mg.c:2919).
template <int AX>
__device__ __forceinline__ double post_sn(const PostGeo &g, const double *__restrict__ x, int a1, int a2, int a3,
                                          double *eo = nullptr) {
  if (a1 < 1 || a2 < 1 || a3 < 1) {
    if (eo) *eo = 0.0;
    return 0.0;
  }
  const long n = g.nid(a1, a2, a3);
  const long v0 = n - g.XY - g.Y - 1;
  const double xkf = g.iAx[a1], ykf = g.iAy[a2], zkf = g.iAz[a3];
  // upper faces; on the last cell of a periodic axis the one on line 0
  const long vxu = ((g.per & 1) && a1 == g.X - 1) ? g.nid(0, a2 - 1, a3 - 1) : v0 + g.Y;
  const long vyu = ((g.per & 2) && a2 == g.Y - 1) ? g.nid(a1 - 1, 0, a3 - 1) : v0 + 1;
  const long vzu = ((g.per & 4) && a3 == g.Z - 1) ? g.nid(a1 - 1, a2 - 1, 0) : v0 + g.XY;
  const double dvxdx = (x[4 * vxu + 1] - x[4 * v0 + 1]) * xkf;
  const double dvydy = (x[4 * vyu + 2] - x[4 * v0 + 2]) * ykf;
  const double dvzdz = (x[4 * vzu + 3] - x[4 * v0 + 3]) * zkf;
  const double d = AX == 0 ? dvxdx : (AX == 1 ? dvydy : dvzdz);
  const double e = d - (dvxdx + dvydy + dvzdz) / 3.0;
  if (eo) *eo = e;
  return 2.0 * g.ec[n] * e;
}

// Where the recompute stores a shear stress of the plane with in-plane axes
// (ia, ib) and normal axis kc (move3mg.c:2933-3044): below the last line of
// kc; in-plane indices 1 .. n-2 on walled axes, and on a periodic axis any
// line, the last one holding the value of line 0 (so the position is moved
// there).  False: the plane holds zero at this node.
__device__ __forceinline__ bool post_shear_at(int &ia, int &ib, int kc, int na, int nb, int nc, bool pa, bool pb) {
  if (kc >= nc - 1) return false;
  if (!pa && (ia < 1 || ia > na - 2)) return false;
  if (!pb && (ib < 1 || ib > nb - 2)) return false;
  if (pa && ia == na - 1) ia = 0;
  if (pb && ib == nb - 1) ib = 0;
  return true;
}

// Shear stresses at node (a1, a2, a3): sxycalc (move3mg.c:400-479),
// sxzcalc (532-615), syzcalc (671-754), value mode, as the recompute stores
// them (post_shear_at); the lower neighbours wrap on periodic axes.
__device__ __forceinline__ double post_sxy(const PostGeo &g, const double *__restrict__ x, int a1, int a2, int a3,
                                           double *eo = nullptr) {
  if (!post_shear_at(a1, a2, a3, g.X, g.Y, g.Z, g.per & 1, g.per & 2)) {
    if (eo) *eo = 0.0;
    return 0.0;
  }
  const long n = g.nid(a1, a2, a3);
  const int l1 = (g.per & 1) ? vwrap(a1 - 1, g.X) : a1 - 1, l2 = (g.per & 2) ? vwrap(a2 - 1, g.Y) : a2 - 1;
  const double xkf = g.iBx[a1], ykf = g.iBy[a2];
  const double dvxdy = (x[4 * n + 1] - x[4 * g.nid(a1, l2, a3) + 1]) * ykf;
  const double dvydx = (x[4 * n + 2] - x[4 * g.nid(l1, a2, a3) + 2]) * xkf;
  const double e = dvxdy + dvydx;
  if (eo) *eo = e;
  return 2.0 * g.exy[n] * e;
}

__device__ __forceinline__ double post_sxz(const PostGeo &g, const double *__restrict__ x, int a1, int a2, int a3,
                                           double *eo = nullptr) {
  if (!post_shear_at(a1, a3, a2, g.X, g.Z, g.Y, g.per & 1, g.per & 4)) {
    if (eo) *eo = 0.0;
    return 0.0;
  }
  const long n = g.nid(a1, a2, a3);
  const int l1 = (g.per & 1) ? vwrap(a1 - 1, g.X) : a1 - 1, l3 = (g.per & 4) ? vwrap(a3 - 1, g.Z) : a3 - 1;
  const double xkf = g.iBx[a1], zkf = g.iBz[a3];
  const double dvxdz = (x[4 * n + 1] - x[4 * g.nid(a1, a2, l3) + 1]) * zkf;
  const double dvzdx = (x[4 * n + 3] - x[4 * g.nid(l1, a2, a3) + 3]) * xkf;
  const double e = dvxdz + dvzdx;
  if (eo) *eo = e;
  return 2.0 * g.exz[n] * e;
}

__device__ __forceinline__ double post_syz(const PostGeo &g, const double *__restrict__ x, int a1, int a2, int a3,
                                           double *eo = nullptr) {
  if (!post_shear_at(a2, a3, a1, g.Y, g.Z, g.X, g.per & 2, g.per & 4)) {
    if (eo) *eo = 0.0;
    return 0.0;
  }
  const long n = g.nid(a1, a2, a3);
  const int l2 = (g.per & 2) ? vwrap(a2 - 1, g.Y) : a2 - 1, l3 = (g.per & 4) ? vwrap(a3 - 1, g.Z) : a3 - 1;
  const double ykf = g.iBy[a2], zkf = g.iBz[a3];
  const double dvydz = (x[4 * n + 2] - x[4 * g.nid(a1, a2, l3) + 2]) * zkf;
  const double dvzdy = (x[4 * n + 3] - x[4 * g.nid(a1, l2, a3) + 3]) * ykf;
  const double e = dvydz + dvzdy;
  if (eo) *eo = e;
  return 2.0 * g.eyz[n] * e;
}

// 12-node viscosity of the momentum error scale (move3mg.c:867-869,
// 1080-1082, 1300-1302): lines lo, mid (weight 10), hi, clamped.
__device__ __forceinline__ double post_nu48(const PostGeo &g, const long lo[4], const long mid[4], const long hi[4]) {
  const float *nu = g.nu;
  double ne;
  if (g.vm == 1) {
    ne = exp((log((double)nu[lo[0]]) + log((double)nu[lo[1]]) + log((double)nu[lo[2]]) + log((double)nu[lo[3]]) +
              10.0 * (log((double)nu[mid[0]]) + log((double)nu[mid[1]]) + log((double)nu[mid[2]]) +
                      log((double)nu[mid[3]])) +
              log((double)nu[hi[0]]) + log((double)nu[hi[1]]) + log((double)nu[hi[2]]) + log((double)nu[hi[3]])) /
             48.0);
  } else if (g.vm == 2) {
    ne = 1.0 / ((1.0 / nu[lo[0]] + 1.0 / nu[lo[1]] + 1.0 / nu[lo[2]] + 1.0 / nu[lo[3]] +
                 10.0 * (1.0 / nu[mid[0]] + 1.0 / nu[mid[1]] + 1.0 / nu[mid[2]] + 1.0 / nu[mid[3]]) +
                 1.0 / nu[hi[0]] + 1.0 / nu[hi[1]] + 1.0 / nu[hi[2]] + 1.0 / nu[hi[3]]) /
                48.0);
  } else {
    ne = ((double)nu[lo[0]] + nu[lo[1]] + nu[lo[2]] + nu[lo[3]] +
          10.0 * ((double)nu[mid[0]] + nu[mid[1]] + nu[mid[2]] + nu[mid[3]]) + nu[hi[0]] + nu[hi[1]] + nu[hi[2]] +
          nu[hi[3]]) /
         48.0;
  }
  if (ne > g.nuend) ne = g.nuend;
  if (ne < g.nubeg) ne = g.nubeg;
  return ne;
}

// Momentum error of the c-row at node (m1, m2, m3) (xstokserr/ystokserr/
// zstokserr with ynerr = 1, move3mg.c:863-903, 1077-1122, 1297-1342):
// divergence of the stored stresses minus the pressure gradient minus the
// gravity term b; *eps = error * h^2 / eta (skf1).
template <int C>
__device__ __forceinline__ double post_err(const PostGeo &g, const double *__restrict__ x,
                                           const double *__restrict__ b, int m1, int m2, int m3, double *eps) {
  const long n = g.nid(m1, m2, m3), Y = g.Y, XY = g.XY;
  long v0 = n, v1 = n + 1, v2 = n + Y, v3 = n + Y + 1, v4 = n + XY, v5 = n + XY + 1, v6 = n + XY + Y,
       v7 = n + XY + Y + 1;
  // a row on line 0 of its periodic axis: the lower cell is the last one
  // (x: v5, y: v6, z: v3 moved, move3mg.c:842, 1056, 1276), which also
  // enters the viscosity lines as the reference has it
  const bool wrap = ((g.per >> (C - 1)) & 1) && (C == 1 ? m1 : (C == 2 ? m2 : m3)) == 0;
  double xkf, ykf, zkf, left, ne;
  // the reference reads the viscosity lines by plain offsets, clamped at the array start
  auto lo0 = [](long v) { return v < 0 ? 0L : v; };
  if (C == 1) {
    if (wrap) v5 = g.nid(g.X - 1, m2 + 1, m3 + 1);
    xkf = 2.0 * g.iBx[m1];
    ykf = g.iAy[m2 + 1];
    zkf = g.iAz[m3 + 1];
    left = (post_sn<0>(g, x, m1 + 1, m2 + 1, m3 + 1) - post_sn<0>(g, x, wrap ? g.X - 1 : m1, m2 + 1, m3 + 1)) * xkf;
    left += (post_sxy(g, x, m1, m2 + 1, m3) - post_sxy(g, x, m1, m2, m3)) * ykf;
    left += (post_sxz(g, x, m1, m2, m3 + 1) - post_sxz(g, x, m1, m2, m3)) * zkf;
    left -= (x[4 * v7] - x[4 * v5]) * xkf;
    const long lo[4] = {lo0(v0 - Y), lo0(v1 - Y), lo0(v4 - Y), lo0(v5 - Y)}, mid[4] = {v0, v1, v4, v5},
               hi[4] = {v2, v3, v6, v7};
    ne = post_nu48(g, lo, mid, hi);
  } else if (C == 2) {
    if (wrap) v6 = g.nid(m1 + 1, g.Y - 1, m3 + 1);
    xkf = g.iAx[m1 + 1];
    ykf = 2.0 * g.iBy[m2];
    zkf = g.iAz[m3 + 1];
    left = (post_sn<1>(g, x, m1 + 1, m2 + 1, m3 + 1) - post_sn<1>(g, x, m1 + 1, wrap ? g.Y - 1 : m2, m3 + 1)) * ykf;
    left += (post_sxy(g, x, m1 + 1, m2, m3) - post_sxy(g, x, m1, m2, m3)) * xkf;
    left += (post_syz(g, x, m1, m2, m3 + 1) - post_syz(g, x, m1, m2, m3)) * zkf;
    left -= (x[4 * v7] - x[4 * v6]) * ykf;
    const long lo[4] = {lo0(v0 - 1), lo0(v2 - 1), lo0(v4 - 1), lo0(v6 - 1)}, mid[4] = {v0, v2, v4, v6},
               hi[4] = {v1, v3, v5, v7};
    ne = post_nu48(g, lo, mid, hi);
  } else {
    if (wrap) v3 = g.nid(m1 + 1, m2 + 1, g.Z - 1);
    xkf = g.iAx[m1 + 1];
    ykf = g.iAy[m2 + 1];
    zkf = 2.0 * g.iBz[m3];
    left = (post_sn<2>(g, x, m1 + 1, m2 + 1, m3 + 1) - post_sn<2>(g, x, m1 + 1, m2 + 1, wrap ? g.Z - 1 : m3)) * zkf;
    left += (post_sxz(g, x, m1 + 1, m2, m3) - post_sxz(g, x, m1, m2, m3)) * xkf;
    left += (post_syz(g, x, m1, m2 + 1, m3) - post_syz(g, x, m1, m2, m3)) * ykf;
    left -= (x[4 * v7] - x[4 * v3]) * zkf;
    const long lo[4] = {lo0(v0 - XY), lo0(v1 - XY), lo0(v2 - XY), lo0(v3 - XY)}, mid[4] = {v0, v1, v2, v3},
               hi[4] = {v4, v5, v6, v7};
    ne = post_nu48(g, lo, mid, hi);
  }
  left = left - b[4 * n + C];
  const double skf1 = 1.0 / ne * (1.0 / xkf / xkf + 1.0 / ykf / ykf + 1.0 / zkf / zkf) / 3.0;
  *eps = left * skf1;
  return left;
}

// Stress and strain-rate planes (xx yy zz xy xz yz, n each) of the
// recompute; zero outside its ranges.  On line 0 of a periodic axis the
// normal stresses and strain rates are those of the last cell, as the
// reference copies them after the solve (move3mg.c:3321-3379).
__global__ void __launch_bounds__(256) k_post_stress(PostGeo g, const double *__restrict__ x, double *sig,
                                                     double *eps) {
  const long N = (long)g.XY * g.Z;
  const long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= N) return;
  int m1, m2, m3;
  decode(n, g.X, g.Y, g.XY, m1, m2, m3);
  double e;
  const int c1 = (g.per & 1) ? pwrap(m1, g.X) : m1, c2 = (g.per & 2) ? pwrap(m2, g.Y) : m2;
  const int c3 = (g.per & 4) ? pwrap(m3, g.Z) : m3;
  sig[0 * N + n] = post_sn<0>(g, x, c1, c2, c3, &e); eps[0 * N + n] = e;
  sig[1 * N + n] = post_sn<1>(g, x, c1, c2, c3, &e); eps[1 * N + n] = e;
  sig[2 * N + n] = post_sn<2>(g, x, c1, c2, c3, &e); eps[2 * N + n] = e;
  sig[3 * N + n] = post_sxy(g, x, m1, m2, m3, &e); eps[3 * N + n] = e;
  sig[4 * N + n] = post_sxz(g, x, m1, m2, m3, &e); eps[4 * N + n] = e;
  sig[5 * N + n] = post_syz(g, x, m1, m2, m3, &e); eps[5 * N + n] = e;
}

// Norm pass: per-block sums and extrema (layout of kPostN, the oracle's
// orc_post): 0 sum err^2, 1 sum eps^2, 2 sum eps^2 err^2, 3 momentum rows,
// 4/5 continuity sum/cells, 6/7 the same over cells with a pressure boundary
// equation, 8/9 velocity boundary-equation sum/rows, 10..17 min/max vx vy
// vz P, 18/19 min/max cell viscosity.  Grid-stride over nodes with a fixed
// grid, then over the boundary rows; the block's partials go to
// partial[blockIdx.x * kPostN + q].
__global__ void __launch_bounds__(256) k_post_norms(PostGeo g, const unsigned char *__restrict__ flags,
                                                    const double *__restrict__ x, const double *__restrict__ b,
                                                    int nbc, const int *__restrict__ bc_r,
                                                    const int *__restrict__ bc_j, const double *__restrict__ bc_negk,
                                                    const double *__restrict__ bc_c, double *partial) {
  double a[kPostN];
#pragma unroll
  for (int q = 0; q < 10; q++) a[q] = 0.0;
#pragma unroll
  for (int q = 10; q < kPostN; q += 2) { a[q] = 1e30; a[q + 1] = -1e30; }
  const long N = (long)g.XY * g.Z;
  const long G = (long)gridDim.x * blockDim.x;
  const int X = g.X, Y = g.Y, Z = g.Z;
  for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < N; n += G) {
    int m1, m2, m3;
    decode(n, X, Y, g.XY, m1, m2, m3);
    const unsigned char f = flags[n];
    const double *xn = x + 4 * n;
    if (m1 && m2 && m3) {
      const double p = xn[0];
      a[16] = fmin(a[16], p); a[17] = fmax(a[17], p);
      const double ne = g.ec[n];
      a[18] = fmin(a[18], ne); a[19] = fmax(a[19], ne);
      // continerr, ynerr = 1 (move3mg.c:1500-1507)
      const double xkf = g.iAx[m1], ykf = g.iAy[m2], zkf = g.iAz[m3];
      const long v0 = n - g.XY - Y - 1;
      const long vxu = ((g.per & 1) && m1 == X - 1) ? g.nid(0, m2 - 1, m3 - 1) : v0 + Y;
      const long vyu = ((g.per & 2) && m2 == Y - 1) ? g.nid(m1 - 1, 0, m3 - 1) : v0 + 1;
      const long vzu = ((g.per & 4) && m3 == Z - 1) ? g.nid(m1 - 1, m2 - 1, 0) : v0 + g.XY;
      const double leftc = (x[4 * vxu + 1] - x[4 * v0 + 1]) * xkf + (x[4 * vyu + 2] - x[4 * v0 + 2]) * ykf +
                           (x[4 * vzu + 3] - x[4 * v0 + 3]) * zkf;
      const double skf = (1.0 / xkf + 1.0 / ykf + 1.0 / zkf) / 3.0;
      const double cv = (leftc - 0.0) * skf;
      if (kind_of(f, 0) != KIND_BC) { a[4] += cv * cv; a[5] += 1.0; }
      else { a[6] += cv * cv; a[7] += 1.0; }
    }
    double e1, iv;
    if (m2 < Y - 1 && m3 < Z - 1) {
      a[10] = fmin(a[10], xn[1]); a[11] = fmax(a[11], xn[1]);
      if (kind_of(f, 1) == KIND_EQ) {
        iv = post_err<1>(g, x, b, m1, m2, m3, &e1);
        a[0] += iv * iv; a[1] += e1 * e1; a[2] += e1 * e1 * iv * iv; a[3] += 1.0;
      }
    }
    if (m1 < X - 1 && m3 < Z - 1) {
      a[12] = fmin(a[12], xn[2]); a[13] = fmax(a[13], xn[2]);
      if (kind_of(f, 2) == KIND_EQ) {
        iv = post_err<2>(g, x, b, m1, m2, m3, &e1);
        a[0] += iv * iv; a[1] += e1 * e1; a[2] += e1 * e1 * iv * iv; a[3] += 1.0;
      }
    }
    if (m1 < X - 1 && m2 < Y - 1) {
      a[14] = fmin(a[14], xn[3]); a[15] = fmax(a[15], xn[3]);
      if (kind_of(f, 3) == KIND_EQ) {
        iv = post_err<3>(g, x, b, m1, m2, m3, &e1);
        a[0] += iv * iv; a[1] += e1 * e1; a[2] += e1 * e1 * iv * iv; a[3] += 1.0;
      }
    }
  }
  // bonderr, ynerr = 1 (move3mg.c:1693-1706), velocity rows in the loop's ranges
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nbc; i += G) {
    const int r = bc_r[i], c = r & 3;
    if (c == 0) continue;
    int m1, m2, m3;
    decode(r >> 2, X, Y, g.XY, m1, m2, m3);
    const bool inr = c == 1 ? (m2 < Y - 1 && m3 < Z - 1) : (c == 2 ? (m1 < X - 1 && m3 < Z - 1) : (m1 < X - 1 && m2 < Y - 1));
    if (!inr) continue;
    double lb = x[r] - bc_c[i];
    if (bc_j[i] >= 0) lb += bc_negk[i] * x[bc_j[i]];
    a[8] += lb * lb;
    a[9] += 1.0;
  }
  // block reduction in a fixed order (wave shuffles, then the four waves in turn)
  __shared__ double red[4][kPostN];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < kPostN; q++) {
    double v = a[q];
    for (int s = 32; s > 0; s >>= 1) {
      const double o = __shfl_down(v, s, 64);
      v = q < 10 ? v + o : ((q & 1) ? fmax(v, o) : fmin(v, o));
    }
    if (lane == 0) red[w][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < kPostN) {
    const int q = threadIdx.x;
    double v = red[0][q];
    for (int k = 1; k < 4; k++) v = q < 10 ? v + red[k][q] : ((q & 1) ? fmax(v, red[k][q]) : fmin(v, red[k][q]));
    partial[(long)blockIdx.x * kPostN + q] = v;
  }
}

// partials of nb blocks -> out[kPostN]: one 256-thread block per quantity,
// thread t folds blocks t, t+256, ... in order, then a fixed tree
__global__ void __launch_bounds__(256) k_post_reduce(const double *__restrict__ partial, int nb, double *out) {
  __shared__ double red[256];
  const int q = blockIdx.x, t = threadIdx.x;
  auto op = [q](double a, double b) { return q < 10 ? a + b : ((q & 1) ? fmax(a, b) : fmin(a, b)); };
  double v = q < 10 ? 0.0 : ((q & 1) ? -1e30 : 1e30);
  for (int k = t; k < nb; k += 256) v = op(v, partial[(long)k * kPostN + q]);
  red[t] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) red[t] = op(red[t], red[t + s]);
    __syncthreads();
  }
  if (t == 0) out[q] = red[0];
}

}  // namespace stk
