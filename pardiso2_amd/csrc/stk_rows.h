// stk_rows.h -- matrix-free rows of the staggered-grid Stokes operator.
//
// Each row is produced in coefficient form: 17 slots for a momentum row,
// 6 for a continuity row, in the slot order and with the floating-point
// merge order of the reference assembly (x row parallel/move3mg.c:796-1003,
// y row 1011-1223, z row 1231-1447, continuity 1455-1573, stress stencils
// 3-788, grad-div 1580-1674).  The slot order is also the reference's ELL
// row layout (gausmat2, parallel/gaus3mg_MG.c:311-365), so a residual
// accumulated slot by slot repeats gausmat4's sum (move3mg.c:3937-3958).
//
// Geometry enters through per-axis inverse-spacing tables
//   iA[i] = 1/(g[i]-g[i-1]),  iB[i] = 1/(g[i+1]-g[i-1]);
// the reference's 2/(g[i+1]-g[i-1]) equals 2*iB[i] exactly.
// Viscosities enter through per-level stencil arrays (cell, three edge
// families, grad-div cell) computed once per viscosity field.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>

namespace stk {

// (dm1, dm2, dm3, component) of each slot 1..17 (index 0 unused); rows of
// component 1 (x), 2 (y), 3 (z); component 0 uses slots 1..6.
__constant__ signed char kSlot[4][18][4] = {
    {{0, 0, 0, 0},
     {-1, -1, -1, 1}, {0, -1, -1, 1}, {-1, -1, -1, 2}, {-1, 0, -1, 2}, {-1, -1, -1, 3}, {-1, -1, 0, 3},
     {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0},
     {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}},
    {{0, 0, 0, 0},
     {0, 0, 0, 1}, {-1, 0, 0, 1}, {1, 0, 0, 1}, {0, -1, 0, 1}, {0, 1, 0, 1}, {0, 0, -1, 1}, {0, 0, 1, 1},
     {-1, 0, 0, 2}, {-1, 1, 0, 2}, {0, 0, 0, 2}, {0, 1, 0, 2},
     {-1, 0, 0, 3}, {-1, 0, 1, 3}, {0, 0, 0, 3}, {0, 0, 1, 3},
     {0, 1, 1, 0}, {1, 1, 1, 0}},
    {{0, 0, 0, 0},
     {0, 0, 0, 2}, {-1, 0, 0, 2}, {1, 0, 0, 2}, {0, -1, 0, 2}, {0, 1, 0, 2}, {0, 0, -1, 2}, {0, 0, 1, 2},
     {0, -1, 0, 1}, {0, 0, 0, 1}, {1, -1, 0, 1}, {1, 0, 0, 1},
     {0, -1, 0, 3}, {0, 0, 0, 3}, {0, -1, 1, 3}, {0, 0, 1, 3},
     {1, 0, 1, 0}, {1, 1, 1, 0}},
    {{0, 0, 0, 0},
     {0, 0, 0, 3}, {-1, 0, 0, 3}, {1, 0, 0, 3}, {0, -1, 0, 3}, {0, 1, 0, 3}, {0, 0, -1, 3}, {0, 0, 1, 3},
     {0, 0, -1, 2}, {0, 1, -1, 2}, {0, 0, 0, 2}, {0, 1, 0, 2},
     {0, 0, -1, 1}, {0, 0, 0, 1}, {1, 0, -1, 1}, {1, 0, 0, 1},
     {1, 1, 0, 0}, {1, 1, 1, 0}},
};

// Multiply-accumulate: FP32 rows fuse (one rounding), FP64 rows round the
// product and the sum separately, as the reference's C does.
__device__ __forceinline__ float mac(float acc, float a, float b) { return __builtin_fmaf(a, b, acc); }
__device__ __forceinline__ double mac(double acc, double a, double b) { return acc + a * b; }
__device__ __forceinline__ float msub(float acc, float a, float b) { return __builtin_fmaf(-a, b, acc); }
__device__ __forceinline__ double msub(double acc, double a, double b) { return acc - a * b; }

// Per-level geometry and stencil viscosities in arithmetic type T.
template <typename T>
struct Geo {
  int X, Y, Z, XY;
  const T *iAx, *iBx, *iAy, *iBy, *iAz, *iBz;
  const T *ec, *exy, *exz, *eyz, *ed;  // stencil viscosities (clamped)
  T p2;                                // grad-div weight (p2koef)
  __device__ __forceinline__ long nid(int m1, int m2, int m3) const {
    return (long)m3 * XY + (long)m1 * Y + m2;
  }
};

// Normal-stress contributions of a cell (sxxcalc/syycalc/szzcalc, coefficient
// mode): a[0..5] = vx v0, vx v2, vy v0, vy v1, vz v0, vz v4 with weight w.
template <typename T, int AX>
__device__ __forceinline__ void snn(T w, T ne, T xk, T yk, T zk, T a[6]) {
  if constexpr (std::is_same<T, float>::value) {
    // FP32 rows (the mixed-precision smoother): 4/3 and 2/3 as constants, no divisions
    const float wn = w * ne, f4 = wn * (4.0f / 3.0f), f2 = wn * (2.0f / 3.0f);
    const float p = (AX == 0) ? f4 : f2, q = (AX == 1) ? f4 : f2, r = (AX == 2) ? f4 : f2;
    const float sp = (AX == 0) ? -1.0f : 1.0f, sq = (AX == 1) ? -1.0f : 1.0f, sr = (AX == 2) ? -1.0f : 1.0f;
    a[0] = sp * p * xk; a[1] = -a[0];
    a[2] = sq * q * yk; a[3] = -a[2];
    a[4] = sr * r * zk; a[5] = -a[4];
    return;
  }
  if (AX == 0) {
    a[0] = -w * ne * T(4) / T(3) * xk; a[1] = +w * ne * T(4) / T(3) * xk;
    a[2] = +w * ne * T(2) / T(3) * yk; a[3] = -w * ne * T(2) / T(3) * yk;
    a[4] = +w * ne * T(2) / T(3) * zk; a[5] = -w * ne * T(2) / T(3) * zk;
  } else if (AX == 1) {
    a[0] = +w * ne * T(2) / T(3) * xk; a[1] = -w * ne * T(2) / T(3) * xk;
    a[2] = -w * ne * T(4) / T(3) * yk; a[3] = +w * ne * T(4) / T(3) * yk;
    a[4] = +w * ne * T(2) / T(3) * zk; a[5] = -w * ne * T(2) / T(3) * zk;
  } else {
    a[0] = +w * ne * T(2) / T(3) * xk; a[1] = -w * ne * T(2) / T(3) * xk;
    a[2] = +w * ne * T(2) / T(3) * yk; a[3] = -w * ne * T(2) / T(3) * yk;
    a[4] = -w * ne * T(4) / T(3) * zk; a[5] = +w * ne * T(4) / T(3) * zk;
  }
}

// Shear-stress contributions (sxycalc/sxzcalc/syzcalc): a[0..3] =
// first comp lower, first comp upper, second comp lower, second comp upper.
template <typename T>
__device__ __forceinline__ void sshear(T w, T ne, T k1, T k2, T a[4]) {
  a[0] = -w * ne * T(2) * k1; a[1] = +w * ne * T(2) * k1;
  a[2] = -w * ne * T(2) * k2; a[3] = +w * ne * T(2) * k2;
}

// grad-div contribution (continadd): same 6 positions as snn.
template <typename T>
__device__ __forceinline__ void cadd(T w, T ne, T xk, T yk, T zk, T a[6]) {
  a[0] = -xk * w * ne; a[1] = +xk * w * ne;
  a[2] = -yk * w * ne; a[3] = +yk * w * ne;
  a[4] = -zk * w * ne; a[5] = +zk * w * ne;
}

// Row builders: (m1,m2,m3) index the spacing tables; v is the node's index
// in the viscosity arrays (strides g.Y, g.XY), from which every viscosity the
// row reads lies at offset 0 or +1 per axis.
// x-momentum row at vx(m1,m2,m3); s[1..17]
template <typename T>
__device__ __forceinline__ void row_x(const Geo<T> &g, int m1, int m2, int m3, long v, T s[18]) {
  const long Y = g.Y, XY = g.XY;
  const T xkf = T(2) * g.iBx[m1];
  const T ykf = g.iAy[m2 + 1];
  const T zkf = g.iAz[m3 + 1];
  T a[6], e[4];
  s[16] = +xkf; s[17] = -xkf;
  snn<T, 0>(-xkf, g.ec[v + (XY + 1)], g.iAx[m1], g.iAy[m2 + 1], g.iAz[m3 + 1], a);
  s[2] = a[0]; s[1] = a[1]; s[8] = a[2]; s[9] = a[3]; s[12] = a[4]; s[13] = a[5];
  snn<T, 0>(xkf, g.ec[v + (XY + Y + 1)], g.iAx[m1 + 1], g.iAy[m2 + 1], g.iAz[m3 + 1], a);
  s[1] += a[0]; s[3] = a[1]; s[10] = a[2]; s[11] = a[3]; s[14] = a[4]; s[15] = a[5];
  sshear<T>(-ykf, g.exy[v], g.iBy[m2], g.iBx[m1], e);
  s[1] += e[1]; s[4] = e[0]; s[8] += e[2]; s[10] += e[3];
  sshear<T>(ykf, g.exy[v + 1], g.iBy[m2 + 1], g.iBx[m1], e);
  s[1] += e[0]; s[5] = e[1]; s[9] += e[2]; s[11] += e[3];
  sshear<T>(-zkf, g.exz[v], g.iBz[m3], g.iBx[m1], e);
  s[1] += e[1]; s[6] = e[0]; s[12] += e[2]; s[14] += e[3];
  sshear<T>(zkf, g.exz[v + XY], g.iBz[m3 + 1], g.iBx[m1], e);
  s[1] += e[0]; s[7] = e[1]; s[13] += e[2]; s[15] += e[3];
  if (g.p2 != T(0)) {
    cadd<T>(-xkf * g.p2, g.ed[v + (XY + 1)], g.iAx[m1], g.iAy[m2 + 1], g.iAz[m3 + 1], a);
    s[2] += a[0]; s[1] += a[1]; s[8] += a[2]; s[9] += a[3]; s[12] += a[4]; s[13] += a[5];
    cadd<T>(xkf * g.p2, g.ed[v + (XY + Y + 1)], g.iAx[m1 + 1], g.iAy[m2 + 1], g.iAz[m3 + 1], a);
    s[1] += a[0]; s[3] += a[1]; s[10] += a[2]; s[11] += a[3]; s[14] += a[4]; s[15] += a[5];
  }
}

// y-momentum row at vy(m1,m2,m3)
template <typename T>
__device__ __forceinline__ void row_y(const Geo<T> &g, int m1, int m2, int m3, long v, T s[18]) {
  const long Y = g.Y, XY = g.XY;
  const T xkf = g.iAx[m1 + 1];
  const T ykf = T(2) * g.iBy[m2];
  const T zkf = g.iAz[m3 + 1];
  T a[6], e[4];
  s[16] = +ykf; s[17] = -ykf;
  snn<T, 1>(-ykf, g.ec[v + (XY + Y)], g.iAx[m1 + 1], g.iAy[m2], g.iAz[m3 + 1], a);
  s[8] = a[0]; s[10] = a[1]; s[4] = a[2]; s[1] = a[3]; s[12] = a[4]; s[14] = a[5];
  snn<T, 1>(ykf, g.ec[v + (XY + Y + 1)], g.iAx[m1 + 1], g.iAy[m2 + 1], g.iAz[m3 + 1], a);
  s[9] = a[0]; s[11] = a[1]; s[1] += a[2]; s[5] = a[3]; s[13] = a[4]; s[15] = a[5];
  // sxy(m1,m2,m3,-xkf): vx lower/upper along y, vy lower/upper along x
  sshear<T>(-xkf, g.exy[v], g.iBy[m2], g.iBx[m1], e);
  s[8] += e[0]; s[9] += e[1]; s[2] = e[2]; s[1] += e[3];
  sshear<T>(xkf, g.exy[v + Y], g.iBy[m2], g.iBx[m1 + 1], e);
  s[10] += e[0]; s[11] += e[1]; s[1] += e[2]; s[3] = e[3];
  // syz(m1,m2,m3,-zkf): vy lower/upper along z, vz lower/upper along y
  sshear<T>(-zkf, g.eyz[v], g.iBz[m3], g.iBy[m2], e);
  s[6] = e[0]; s[1] += e[1]; s[12] += e[2]; s[13] += e[3];
  sshear<T>(zkf, g.eyz[v + XY], g.iBz[m3 + 1], g.iBy[m2], e);
  s[1] += e[0]; s[7] = e[1]; s[14] += e[2]; s[15] += e[3];
  if (g.p2 != T(0)) {
    cadd<T>(-ykf * g.p2, g.ed[v + (XY + Y)], g.iAx[m1 + 1], g.iAy[m2], g.iAz[m3 + 1], a);
    s[8] += a[0]; s[10] += a[1]; s[4] += a[2]; s[1] += a[3]; s[12] += a[4]; s[14] += a[5];
    cadd<T>(ykf * g.p2, g.ed[v + (XY + Y + 1)], g.iAx[m1 + 1], g.iAy[m2 + 1], g.iAz[m3 + 1], a);
    s[9] += a[0]; s[11] += a[1]; s[1] += a[2]; s[5] += a[3]; s[13] += a[4]; s[15] += a[5];
  }
}

// z-momentum row at vz(m1,m2,m3)
template <typename T>
__device__ __forceinline__ void row_z(const Geo<T> &g, int m1, int m2, int m3, long v, T s[18]) {
  const long Y = g.Y, XY = g.XY;
  const T xkf = g.iAx[m1 + 1];
  const T ykf = g.iAy[m2 + 1];
  const T zkf = T(2) * g.iBz[m3];
  T a[6], e[4];
  s[16] = +zkf; s[17] = -zkf;
  snn<T, 2>(-zkf, g.ec[v + (Y + 1)], g.iAx[m1 + 1], g.iAy[m2 + 1], g.iAz[m3], a);
  s[12] = a[0]; s[14] = a[1]; s[8] = a[2]; s[9] = a[3]; s[6] = a[4]; s[1] = a[5];
  snn<T, 2>(zkf, g.ec[v + (XY + Y + 1)], g.iAx[m1 + 1], g.iAy[m2 + 1], g.iAz[m3 + 1], a);
  s[13] = a[0]; s[15] = a[1]; s[10] = a[2]; s[11] = a[3]; s[1] += a[4]; s[7] = a[5];
  // sxz(m1,m2,m3,-xkf): vx lower/upper along z, vz lower/upper along x
  sshear<T>(-xkf, g.exz[v], g.iBz[m3], g.iBx[m1], e);
  s[12] += e[0]; s[13] += e[1]; s[2] = e[2]; s[1] += e[3];
  sshear<T>(xkf, g.exz[v + Y], g.iBz[m3], g.iBx[m1 + 1], e);
  s[14] += e[0]; s[15] += e[1]; s[1] += e[2]; s[3] = e[3];
  // syz(m1,m2,m3,-ykf): vy lower/upper along z, vz lower/upper along y
  sshear<T>(-ykf, g.eyz[v], g.iBz[m3], g.iBy[m2], e);
  s[8] += e[0]; s[10] += e[1]; s[4] = e[2]; s[1] += e[3];
  sshear<T>(ykf, g.eyz[v + 1], g.iBz[m3], g.iBy[m2 + 1], e);
  s[9] += e[0]; s[11] += e[1]; s[1] += e[2]; s[5] = e[3];
  if (g.p2 != T(0)) {
    cadd<T>(-zkf * g.p2, g.ed[v + (Y + 1)], g.iAx[m1 + 1], g.iAy[m2 + 1], g.iAz[m3], a);
    s[12] += a[0]; s[14] += a[1]; s[8] += a[2]; s[9] += a[3]; s[6] += a[4]; s[1] += a[5];
    cadd<T>(zkf * g.p2, g.ed[v + (XY + Y + 1)], g.iAx[m1 + 1], g.iAy[m2 + 1], g.iAz[m3 + 1], a);
    s[13] += a[0]; s[15] += a[1]; s[10] += a[2]; s[11] += a[3]; s[1] += a[4]; s[7] += a[5];
  }
}

// continuity row of cell (m1,m2,m3) (upper-corner index)
template <typename T>
__device__ __forceinline__ void row_p(const Geo<T> &g, int m1, int m2, int m3, T s[18]) {
  const T xk = g.iAx[m1], yk = g.iAy[m2], zk = g.iAz[m3];
  s[1] = -xk; s[2] = +xk; s[3] = -yk; s[4] = +yk; s[5] = -zk; s[6] = +zk;
}

template <typename T>
__device__ __forceinline__ int row_coef(const Geo<T> &g, int c, int m1, int m2, int m3, T s[18]) {
  const long v = g.nid(m1, m2, m3);
  if (c == 1) { row_x<T>(g, m1, m2, m3, v, s); return 17; }
  if (c == 2) { row_y<T>(g, m1, m2, m3, v, s); return 17; }
  if (c == 3) { row_z<T>(g, m1, m2, m3, v, s); return 17; }
  row_p<T>(g, m1, m2, m3, s);
  return 6;
}

// compile-time slot tables: Slot<C>::d[q] = {dm1, dm2, dm3, comp}
template <int C> struct Slot;
template <> struct Slot<0> {
  static constexpr int ns = 6;
  static constexpr signed char d[18][4] = {
      {0, 0, 0, 0}, {-1, -1, -1, 1}, {0, -1, -1, 1}, {-1, -1, -1, 2}, {-1, 0, -1, 2}, {-1, -1, -1, 3},
      {-1, -1, 0, 3}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0},
      {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
};
template <> struct Slot<1> {
  static constexpr int ns = 17;
  static constexpr signed char d[18][4] = {
      {0, 0, 0, 0}, {0, 0, 0, 1}, {-1, 0, 0, 1}, {1, 0, 0, 1}, {0, -1, 0, 1}, {0, 1, 0, 1}, {0, 0, -1, 1},
      {0, 0, 1, 1}, {-1, 0, 0, 2}, {-1, 1, 0, 2}, {0, 0, 0, 2}, {0, 1, 0, 2}, {-1, 0, 0, 3}, {-1, 0, 1, 3},
      {0, 0, 0, 3}, {0, 0, 1, 3}, {0, 1, 1, 0}, {1, 1, 1, 0}};
};
template <> struct Slot<2> {
  static constexpr int ns = 17;
  static constexpr signed char d[18][4] = {
      {0, 0, 0, 0}, {0, 0, 0, 2}, {-1, 0, 0, 2}, {1, 0, 0, 2}, {0, -1, 0, 2}, {0, 1, 0, 2}, {0, 0, -1, 2},
      {0, 0, 1, 2}, {0, -1, 0, 1}, {0, 0, 0, 1}, {1, -1, 0, 1}, {1, 0, 0, 1}, {0, -1, 0, 3}, {0, 0, 0, 3},
      {0, -1, 1, 3}, {0, 0, 1, 3}, {1, 0, 1, 0}, {1, 1, 1, 0}};
};
template <> struct Slot<3> {
  static constexpr int ns = 17;
  static constexpr signed char d[18][4] = {
      {0, 0, 0, 0}, {0, 0, 0, 3}, {-1, 0, 0, 3}, {1, 0, 0, 3}, {0, -1, 0, 3}, {0, 1, 0, 3}, {0, 0, -1, 3},
      {0, 0, 1, 3}, {0, 0, -1, 2}, {0, 1, -1, 2}, {0, 0, 0, 2}, {0, 1, 0, 2}, {0, 0, -1, 1}, {0, 0, 0, 1},
      {1, 0, -1, 1}, {1, 0, 0, 1}, {1, 1, 0, 0}, {1, 1, 1, 0}};
};

template <typename T, int C>
__device__ __forceinline__ void row_coef_c(const Geo<T> &g, int m1, int m2, int m3, long v, T s[18]) {
  if (C == 1) row_x<T>(g, m1, m2, m3, v, s);
  else if (C == 2) row_y<T>(g, m1, m2, m3, v, s);
  else if (C == 3) row_z<T>(g, m1, m2, m3, v, s);
  else row_p<T>(g, m1, m2, m3, s);
}
template <typename T, int C>
__device__ __forceinline__ void row_coef_c(const Geo<T> &g, int m1, int m2, int m3, T s[18]) {
  row_coef_c<T, C>(g, m1, m2, m3, g.nid(m1, m2, m3), s);
}

// sum over the slots of a row against an AoS vector x (4 per node):
// RES: b - sum s*x in slot order, otherwise sum s*x
template <typename T, int C, bool RES>
__device__ __forceinline__ T row_dot(const Geo<T> &g, const T s[18], const T *__restrict__ x, long n, T b) {
  T acc = RES ? b : T(0);
#pragma unroll
  for (int q = 1; q <= Slot<C>::ns; q++) {
    const long uq = 4 * (n + (long)Slot<C>::d[q][2] * g.XY + (long)Slot<C>::d[q][0] * g.Y + Slot<C>::d[q][1]) +
                    Slot<C>::d[q][3];
    acc = RES ? msub(acc, s[q], x[uq]) : mac(acc, s[q], x[uq]);
  }
  return acc;
}

// unknown kind flags: two bits per component in one byte per node
enum { KIND_NONE = 0, KIND_EQ = 1, KIND_BC = 2 };
__host__ __device__ __forceinline__ int kind_of(unsigned char f, int c) { return (f >> (2 * c)) & 3; }

}  // namespace stk
