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
__host__ __device__ __forceinline__ float mac(float acc, float a, float b) { return __builtin_fmaf(a, b, acc); }
__host__ __device__ __forceinline__ double mac(double acc, double a, double b) { return acc + a * b; }
__host__ __device__ __forceinline__ float msub(float acc, float a, float b) { return __builtin_fmaf(-a, b, acc); }
__host__ __device__ __forceinline__ double msub(double acc, double a, double b) { return acc - a * b; }

// Per-level geometry and stencil viscosities in arithmetic type T.
template <typename T>
struct Geo {
  int X, Y, Z, XY;
  const T *iAx, *iBx, *iAy, *iBy, *iAz, *iBz;
  const T *ec, *exy, *exz, *eyz, *ed;  // stencil viscosities (clamped)
  T p2;                                // grad-div weight (p2koef)
  int per;                             // periodic axes: bit 0 x, bit 1 y, bit 2 z
  __device__ __forceinline__ long nid(int m1, int m2, int m3) const {
    return (long)m3 * XY + (long)m1 * Y + m2;
  }
};

// Periodic axes (the reference's xperiodic/yperiodic/zperiodic branches,
// parallel/move3mg.c:27-34, 400-429, 834-847, 1481-1483) amount to one index
// map per axis of n lines: grid line n-1 is line 0 again (its normal
// velocity carries a copy equation), so velocities live on lines 0..n-2 and
// pressure cells on 1..n-1, and a stencil index i outside those ranges names
// the unknown at
//   velocity:  vwrap(i) = i mod (n-1)        (-1 -> n-2, n-1 -> 0, n -> 1)
//   pressure:  pwrap(i) = (i-1) mod (n-1) + 1 (0 -> n-1, n -> 1)
// for i in [-1, n].  Stress points follow the same maps: a cell (indexed by
// its upper corner, like P) by pwrap, an edge on a grid line by vwrap.
__host__ __device__ __forceinline__ int vwrap(int i, int n) {
  return i < 0 ? i + (n - 1) : (i >= n - 1 ? i - (n - 1) : i);
}
__host__ __device__ __forceinline__ int pwrap(int i, int n) {
  return i < 1 ? i + (n - 1) : (i > n - 1 ? i - (n - 1) : i);
}
// node that holds unknown c read at stencil position (m1, m2, m3)
template <typename T>
__device__ __forceinline__ long unode(const Geo<T> &g, int m1, int m2, int m3, int c) {
  if (g.per & 1) m1 = c ? vwrap(m1, g.X) : pwrap(m1, g.X);
  if (g.per & 2) m2 = c ? vwrap(m2, g.Y) : pwrap(m2, g.Y);
  if (g.per & 4) m3 = c ? vwrap(m3, g.Z) : pwrap(m3, g.Z);
  return g.nid(m1, m2, m3);
}

// Normal-stress contributions of a cell (sxxcalc/syycalc/szzcalc, coefficient
// mode): a[0..5] = vx v0, vx v2, vy v0, vy v1, vz v0, vz v4 with weight w.
template <typename T, int AX>
__host__ __device__ __forceinline__ void snn(T w, T ne, T xk, T yk, T zk, T a[6]) {
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
__host__ __device__ __forceinline__ void sshear(T w, T ne, T k1, T k2, T a[4]) {
  a[0] = -w * ne * T(2) * k1; a[1] = +w * ne * T(2) * k1;
  a[2] = -w * ne * T(2) * k2; a[3] = +w * ne * T(2) * k2;
}

// grad-div contribution (continadd): same 6 positions as snn.
template <typename T>
__host__ __device__ __forceinline__ void cadd(T w, T ne, T xk, T yk, T zk, T a[6]) {
  a[0] = -xk * w * ne; a[1] = +xk * w * ne;
  a[2] = -yk * w * ne; a[3] = +yk * w * ne;
  a[4] = -zk * w * ne; a[5] = +zk * w * ne;
}

// Spacing values a node's velocity rows read: iA and iB at the node's own
// index and the next one along each axis (ax0 = iAx[m1], ax1 = iAx[m1+1],
// bx0 = iBx[m1], bx1 = iBx[m1+1], and likewise in y and z).
template <typename T>
struct Sp {
  T ax0, ax1, bx0, bx1, ay0, ay1, by0, by1, az0, az1, bz0, bz1;
};
template <typename T>
__device__ __forceinline__ Sp<T> sp_load(const Geo<T> &g, int m1, int m2, int m3) {
  return Sp<T>{g.iAx[m1], g.iAx[m1 + 1], g.iBx[m1], g.iBx[m1 + 1], g.iAy[m2], g.iAy[m2 + 1],
               g.iBy[m2], g.iBy[m2 + 1], g.iAz[m3], g.iAz[m3 + 1], g.iBz[m3], g.iBz[m3 + 1]};
}

// The viscosities one momentum row reads: the two normal-stress cells (c),
// the two shear edges along the first (s) and second (t) transverse axis,
// and the two grad-div cells (d).  Every one lies at offset 0 or +1 per axis
// from the row's node v (strides g.Y, g.XY).
template <typename T>
struct Vis {
  T c0, c1, s0, s1, t0, t1, d0, d1;
};
template <typename T, int C>
__device__ __forceinline__ Vis<T> vis_load(const Geo<T> &g, long v) {
  const long Y = g.Y, XY = g.XY;
  Vis<T> e;
  // without a grad-div term the d values are not used: they are read from
  // the cell plane at the c offsets (the same addresses, no extra traffic)
  // so that no load sits behind a branch
  const T *ed = g.p2 != T(0) ? g.ed : g.ec;
  if (C == 1) {
    e.c0 = g.ec[v + (XY + 1)]; e.c1 = g.ec[v + (XY + Y + 1)];
    e.s0 = g.exy[v]; e.s1 = g.exy[v + 1]; e.t0 = g.exz[v]; e.t1 = g.exz[v + XY];
    e.d0 = ed[v + (XY + 1)]; e.d1 = ed[v + (XY + Y + 1)];
  } else if (C == 2) {
    e.c0 = g.ec[v + (XY + Y)]; e.c1 = g.ec[v + (XY + Y + 1)];
    e.s0 = g.exy[v]; e.s1 = g.exy[v + Y]; e.t0 = g.eyz[v]; e.t1 = g.eyz[v + XY];
    e.d0 = ed[v + (XY + Y)]; e.d1 = ed[v + (XY + Y + 1)];
  } else {
    e.c0 = g.ec[v + (Y + 1)]; e.c1 = g.ec[v + (XY + Y + 1)];
    e.s0 = g.exz[v]; e.s1 = g.exz[v + Y]; e.t0 = g.eyz[v]; e.t1 = g.eyz[v + 1];
    e.d0 = ed[v + (Y + 1)]; e.d1 = ed[v + (XY + Y + 1)];
  }
  return e;
}

// Row builders from the row's spacing values and viscosities; s[1..17].
// x-momentum row at vx(m1,m2,m3)
template <typename T>
__host__ __device__ __forceinline__ void row_x_v(const Sp<T> &p, const Vis<T> &w, T p2, T s[18]) {
  const T xkf = T(2) * p.bx0;
  const T ykf = p.ay1;
  const T zkf = p.az1;
  T a[6], e[4];
  s[16] = +xkf; s[17] = -xkf;
  snn<T, 0>(-xkf, w.c0, p.ax0, p.ay1, p.az1, a);
  s[2] = a[0]; s[1] = a[1]; s[8] = a[2]; s[9] = a[3]; s[12] = a[4]; s[13] = a[5];
  snn<T, 0>(xkf, w.c1, p.ax1, p.ay1, p.az1, a);
  s[1] += a[0]; s[3] = a[1]; s[10] = a[2]; s[11] = a[3]; s[14] = a[4]; s[15] = a[5];
  sshear<T>(-ykf, w.s0, p.by0, p.bx0, e);
  s[1] += e[1]; s[4] = e[0]; s[8] += e[2]; s[10] += e[3];
  sshear<T>(ykf, w.s1, p.by1, p.bx0, e);
  s[1] += e[0]; s[5] = e[1]; s[9] += e[2]; s[11] += e[3];
  sshear<T>(-zkf, w.t0, p.bz0, p.bx0, e);
  s[1] += e[1]; s[6] = e[0]; s[12] += e[2]; s[14] += e[3];
  sshear<T>(zkf, w.t1, p.bz1, p.bx0, e);
  s[1] += e[0]; s[7] = e[1]; s[13] += e[2]; s[15] += e[3];
  if (p2 != T(0)) {
    cadd<T>(-xkf * p2, w.d0, p.ax0, p.ay1, p.az1, a);
    s[2] += a[0]; s[1] += a[1]; s[8] += a[2]; s[9] += a[3]; s[12] += a[4]; s[13] += a[5];
    cadd<T>(xkf * p2, w.d1, p.ax1, p.ay1, p.az1, a);
    s[1] += a[0]; s[3] += a[1]; s[10] += a[2]; s[11] += a[3]; s[14] += a[4]; s[15] += a[5];
  }
}

// y-momentum row at vy(m1,m2,m3)
template <typename T>
__host__ __device__ __forceinline__ void row_y_v(const Sp<T> &p, const Vis<T> &w, T p2, T s[18]) {
  const T xkf = p.ax1;
  const T ykf = T(2) * p.by0;
  const T zkf = p.az1;
  T a[6], e[4];
  s[16] = +ykf; s[17] = -ykf;
  snn<T, 1>(-ykf, w.c0, p.ax1, p.ay0, p.az1, a);
  s[8] = a[0]; s[10] = a[1]; s[4] = a[2]; s[1] = a[3]; s[12] = a[4]; s[14] = a[5];
  snn<T, 1>(ykf, w.c1, p.ax1, p.ay1, p.az1, a);
  s[9] = a[0]; s[11] = a[1]; s[1] += a[2]; s[5] = a[3]; s[13] = a[4]; s[15] = a[5];
  // sxy(m1,m2,m3,-xkf): vx lower/upper along y, vy lower/upper along x
  sshear<T>(-xkf, w.s0, p.by0, p.bx0, e);
  s[8] += e[0]; s[9] += e[1]; s[2] = e[2]; s[1] += e[3];
  sshear<T>(xkf, w.s1, p.by0, p.bx1, e);
  s[10] += e[0]; s[11] += e[1]; s[1] += e[2]; s[3] = e[3];
  // syz(m1,m2,m3,-zkf): vy lower/upper along z, vz lower/upper along y
  sshear<T>(-zkf, w.t0, p.bz0, p.by0, e);
  s[6] = e[0]; s[1] += e[1]; s[12] += e[2]; s[13] += e[3];
  sshear<T>(zkf, w.t1, p.bz1, p.by0, e);
  s[1] += e[0]; s[7] = e[1]; s[14] += e[2]; s[15] += e[3];
  if (p2 != T(0)) {
    cadd<T>(-ykf * p2, w.d0, p.ax1, p.ay0, p.az1, a);
    s[8] += a[0]; s[10] += a[1]; s[4] += a[2]; s[1] += a[3]; s[12] += a[4]; s[14] += a[5];
    cadd<T>(ykf * p2, w.d1, p.ax1, p.ay1, p.az1, a);
    s[9] += a[0]; s[11] += a[1]; s[1] += a[2]; s[5] += a[3]; s[13] += a[4]; s[15] += a[5];
  }
}

// z-momentum row at vz(m1,m2,m3)
template <typename T>
__host__ __device__ __forceinline__ void row_z_v(const Sp<T> &p, const Vis<T> &w, T p2, T s[18]) {
  const T xkf = p.ax1;
  const T ykf = p.ay1;
  const T zkf = T(2) * p.bz0;
  T a[6], e[4];
  s[16] = +zkf; s[17] = -zkf;
  snn<T, 2>(-zkf, w.c0, p.ax1, p.ay1, p.az0, a);
  s[12] = a[0]; s[14] = a[1]; s[8] = a[2]; s[9] = a[3]; s[6] = a[4]; s[1] = a[5];
  snn<T, 2>(zkf, w.c1, p.ax1, p.ay1, p.az1, a);
  s[13] = a[0]; s[15] = a[1]; s[10] = a[2]; s[11] = a[3]; s[1] += a[4]; s[7] = a[5];
  // sxz(m1,m2,m3,-xkf): vx lower/upper along z, vz lower/upper along x
  sshear<T>(-xkf, w.s0, p.bz0, p.bx0, e);
  s[12] += e[0]; s[13] += e[1]; s[2] = e[2]; s[1] += e[3];
  sshear<T>(xkf, w.s1, p.bz0, p.bx1, e);
  s[14] += e[0]; s[15] += e[1]; s[1] += e[2]; s[3] = e[3];
  // syz(m1,m2,m3,-ykf): vy lower/upper along z, vz lower/upper along y
  sshear<T>(-ykf, w.t0, p.bz0, p.by0, e);
  s[8] += e[0]; s[10] += e[1]; s[4] = e[2]; s[1] += e[3];
  sshear<T>(ykf, w.t1, p.bz0, p.by1, e);
  s[9] += e[0]; s[11] += e[1]; s[1] += e[2]; s[5] = e[3];
  if (p2 != T(0)) {
    cadd<T>(-zkf * p2, w.d0, p.ax1, p.ay1, p.az0, a);
    s[12] += a[0]; s[14] += a[1]; s[8] += a[2]; s[9] += a[3]; s[6] += a[4]; s[1] += a[5];
    cadd<T>(zkf * p2, w.d1, p.ax1, p.ay1, p.az1, a);
    s[13] += a[0]; s[15] += a[1]; s[10] += a[2]; s[11] += a[3]; s[1] += a[4]; s[7] += a[5];
  }
}

// ---------------------------------------------------------------------------
// Flux form of the same velocity rows (the FP32 smoother's).  Every stencil
// contribution above writes its coefficients in +/- pairs: snn's a[1] = -a[0],
// a[3] = -a[2], a[5] = -a[4]; sshear's e[1] = -e[0], e[3] = -e[2]; cadd's the
// same; the pressure pair s[16] = -s[17].  So the row's sum over slots is
//   S = sum_q s[q] x[q] = sum over pairs (a, b) of c (x[a] - x[b]),
// and its diagonal s[1] is the sum of the pair coefficients that touch slot 1
// (+c as first member, -c as second).  Each pair below names the slots the
// coefficient form assigns (row_x_v / row_y_v / row_z_v), so a sweep can use
// S and D = s[1] without building the 17 coefficients: about half the
// arithmetic.  The sums round differently from the coefficient form.
// xs[q] = x at slot q (q = 1..17).
template <typename T>
__host__ __device__ __forceinline__ T fmat(T a, T b, T c) {
  if constexpr (std::is_same<T, float>::value) return __builtin_fmaf(a, b, c);
  return a * b + c;
}
template <typename T, int A, int B>
__host__ __device__ __forceinline__ void fpair(T c, const T *xs, T &S, T &D) {
  S = fmat<T>(c, xs[A] - xs[B], S);
  if (A == 1) D += c;
  if (B == 1) D -= c;
}
template <typename T, int AX, int A0, int A1, int A2, int A3, int A4, int A5>
__host__ __device__ __forceinline__ void fsnn(T w, T ne, T xk, T yk, T zk, const T *xs, T &S, T &D) {
  const T wn = w * ne, f4 = wn * T(4.0 / 3.0), f2 = wn * T(2.0 / 3.0);
  const T p = AX == 0 ? f4 : f2, q = AX == 1 ? f4 : f2, r = AX == 2 ? f4 : f2;
  fpair<T, A0, A1>((AX == 0 ? -p : p) * xk, xs, S, D);
  fpair<T, A2, A3>((AX == 1 ? -q : q) * yk, xs, S, D);
  fpair<T, A4, A5>((AX == 2 ? -r : r) * zk, xs, S, D);
}
template <typename T, int A0, int A1, int A2, int A3>
__host__ __device__ __forceinline__ void fshear(T w, T ne, T k1, T k2, const T *xs, T &S, T &D) {
  const T t = -(w * ne) * T(2);
  fpair<T, A0, A1>(t * k1, xs, S, D);
  fpair<T, A2, A3>(t * k2, xs, S, D);
}
template <typename T, int A0, int A1, int A2, int A3, int A4, int A5>
__host__ __device__ __forceinline__ void fcadd(T w, T ne, T xk, T yk, T zk, const T *xs, T &S, T &D) {
  const T t = -(w * ne);
  fpair<T, A0, A1>(t * xk, xs, S, D);
  fpair<T, A2, A3>(t * yk, xs, S, D);
  fpair<T, A4, A5>(t * zk, xs, S, D);
}

// S and D of the x-momentum row (pairs of row_x_v)
template <typename T>
__host__ __device__ __forceinline__ void flux_x(const Sp<T> &p, const Vis<T> &w, T p2, const T *xs, T &S, T &D) {
  const T xkf = T(2) * p.bx0, ykf = p.ay1, zkf = p.az1;
  S = T(0); D = T(0);
  fsnn<T, 0, 2, 1, 8, 9, 12, 13>(-xkf, w.c0, p.ax0, p.ay1, p.az1, xs, S, D);
  fsnn<T, 0, 1, 3, 10, 11, 14, 15>(xkf, w.c1, p.ax1, p.ay1, p.az1, xs, S, D);
  fshear<T, 4, 1, 8, 10>(-ykf, w.s0, p.by0, p.bx0, xs, S, D);
  fshear<T, 1, 5, 9, 11>(ykf, w.s1, p.by1, p.bx0, xs, S, D);
  fshear<T, 6, 1, 12, 14>(-zkf, w.t0, p.bz0, p.bx0, xs, S, D);
  fshear<T, 1, 7, 13, 15>(zkf, w.t1, p.bz1, p.bx0, xs, S, D);
  if (p2 != T(0)) {
    fcadd<T, 2, 1, 8, 9, 12, 13>(-xkf * p2, w.d0, p.ax0, p.ay1, p.az1, xs, S, D);
    fcadd<T, 1, 3, 10, 11, 14, 15>(xkf * p2, w.d1, p.ax1, p.ay1, p.az1, xs, S, D);
  }
  fpair<T, 16, 17>(xkf, xs, S, D);
}
// S and D of the y-momentum row (pairs of row_y_v)
template <typename T>
__host__ __device__ __forceinline__ void flux_y(const Sp<T> &p, const Vis<T> &w, T p2, const T *xs, T &S, T &D) {
  const T xkf = p.ax1, ykf = T(2) * p.by0, zkf = p.az1;
  S = T(0); D = T(0);
  fsnn<T, 1, 8, 10, 4, 1, 12, 14>(-ykf, w.c0, p.ax1, p.ay0, p.az1, xs, S, D);
  fsnn<T, 1, 9, 11, 1, 5, 13, 15>(ykf, w.c1, p.ax1, p.ay1, p.az1, xs, S, D);
  fshear<T, 8, 9, 2, 1>(-xkf, w.s0, p.by0, p.bx0, xs, S, D);
  fshear<T, 10, 11, 1, 3>(xkf, w.s1, p.by0, p.bx1, xs, S, D);
  fshear<T, 6, 1, 12, 13>(-zkf, w.t0, p.bz0, p.by0, xs, S, D);
  fshear<T, 1, 7, 14, 15>(zkf, w.t1, p.bz1, p.by0, xs, S, D);
  if (p2 != T(0)) {
    fcadd<T, 8, 10, 4, 1, 12, 14>(-ykf * p2, w.d0, p.ax1, p.ay0, p.az1, xs, S, D);
    fcadd<T, 9, 11, 1, 5, 13, 15>(ykf * p2, w.d1, p.ax1, p.ay1, p.az1, xs, S, D);
  }
  fpair<T, 16, 17>(ykf, xs, S, D);
}
// S and D of the z-momentum row (pairs of row_z_v)
template <typename T>
__host__ __device__ __forceinline__ void flux_z(const Sp<T> &p, const Vis<T> &w, T p2, const T *xs, T &S, T &D) {
  const T xkf = p.ax1, ykf = p.ay1, zkf = T(2) * p.bz0;
  S = T(0); D = T(0);
  fsnn<T, 2, 12, 14, 8, 9, 6, 1>(-zkf, w.c0, p.ax1, p.ay1, p.az0, xs, S, D);
  fsnn<T, 2, 13, 15, 10, 11, 1, 7>(zkf, w.c1, p.ax1, p.ay1, p.az1, xs, S, D);
  fshear<T, 12, 13, 2, 1>(-xkf, w.s0, p.bz0, p.bx0, xs, S, D);
  fshear<T, 14, 15, 1, 3>(xkf, w.s1, p.bz0, p.bx1, xs, S, D);
  fshear<T, 8, 10, 4, 1>(-ykf, w.t0, p.bz0, p.by0, xs, S, D);
  fshear<T, 9, 11, 1, 5>(ykf, w.t1, p.bz0, p.by1, xs, S, D);
  if (p2 != T(0)) {
    fcadd<T, 12, 14, 8, 9, 6, 1>(-zkf * p2, w.d0, p.ax1, p.ay1, p.az0, xs, S, D);
    fcadd<T, 13, 15, 10, 11, 1, 7>(zkf * p2, w.d1, p.ax1, p.ay1, p.az1, xs, S, D);
  }
  fpair<T, 16, 17>(zkf, xs, S, D);
}

// On a level whose spacing is uniform along each axis, a row's pair
// coefficients are a viscosity times a constant of the level: the
// contributions with w = -k and w = +k differ only in sign, so each row kind
// needs 11 constants (FluxU::g[C-1]): the normal-stress pairs' w (-4/3 or
// 2/3) k for w = -k (3), the two shear families' -2 w k1, -2 w k2 for
// w = -k (2 + 2), the grad-div pairs' -w k for w = -k p2 (3), and the
// pressure pair's k.  Computed once per level on the host.
struct FluxU {
  float g[3][11];
};
template <int AX, int A0, int A1, int A2, int A3, int A4, int A5>
__host__ __device__ __forceinline__ void usnn(float ne, const float *G, const float *xs, float &S, float &D) {
  fpair<float, A0, A1>(ne * G[0], xs, S, D);
  fpair<float, A2, A3>(ne * G[1], xs, S, D);
  fpair<float, A4, A5>(ne * G[2], xs, S, D);
}
template <int A0, int A1, int A2, int A3>
__host__ __device__ __forceinline__ void ushear(float ne, const float *G, const float *xs, float &S, float &D) {
  fpair<float, A0, A1>(ne * G[0], xs, S, D);
  fpair<float, A2, A3>(ne * G[1], xs, S, D);
}
// the x, y and z rows with the pairs of flux_x / flux_y / flux_z; the second
// contribution of each family enters with its viscosity negated (w = +k)
template <int C>
__host__ __device__ __forceinline__ void flux_u(const Vis<float> &w, const float *G, bool p2, const float *xs,
                                                float &S, float &D) {
  S = 0.f; D = 0.f;
  if (C == 1) {
    usnn<0, 2, 1, 8, 9, 12, 13>(w.c0, G, xs, S, D);
    usnn<0, 1, 3, 10, 11, 14, 15>(-w.c1, G, xs, S, D);
    ushear<4, 1, 8, 10>(w.s0, G + 3, xs, S, D);
    ushear<1, 5, 9, 11>(-w.s1, G + 3, xs, S, D);
    ushear<6, 1, 12, 14>(w.t0, G + 5, xs, S, D);
    ushear<1, 7, 13, 15>(-w.t1, G + 5, xs, S, D);
    if (p2) {
      usnn<0, 2, 1, 8, 9, 12, 13>(w.d0, G + 7, xs, S, D);
      usnn<0, 1, 3, 10, 11, 14, 15>(-w.d1, G + 7, xs, S, D);
    }
  } else if (C == 2) {
    usnn<1, 8, 10, 4, 1, 12, 14>(w.c0, G, xs, S, D);
    usnn<1, 9, 11, 1, 5, 13, 15>(-w.c1, G, xs, S, D);
    ushear<8, 9, 2, 1>(w.s0, G + 3, xs, S, D);
    ushear<10, 11, 1, 3>(-w.s1, G + 3, xs, S, D);
    ushear<6, 1, 12, 13>(w.t0, G + 5, xs, S, D);
    ushear<1, 7, 14, 15>(-w.t1, G + 5, xs, S, D);
    if (p2) {
      usnn<1, 8, 10, 4, 1, 12, 14>(w.d0, G + 7, xs, S, D);
      usnn<1, 9, 11, 1, 5, 13, 15>(-w.d1, G + 7, xs, S, D);
    }
  } else {
    usnn<2, 12, 14, 8, 9, 6, 1>(w.c0, G, xs, S, D);
    usnn<2, 13, 15, 10, 11, 1, 7>(-w.c1, G, xs, S, D);
    ushear<12, 13, 2, 1>(w.s0, G + 3, xs, S, D);
    ushear<14, 15, 1, 3>(-w.s1, G + 3, xs, S, D);
    ushear<8, 10, 4, 1>(w.t0, G + 5, xs, S, D);
    ushear<9, 11, 1, 5>(-w.t1, G + 5, xs, S, D);
    if (p2) {
      usnn<2, 12, 14, 8, 9, 6, 1>(w.d0, G + 7, xs, S, D);
      usnn<2, 13, 15, 10, 11, 1, 7>(-w.d1, G + 7, xs, S, D);
    }
  }
  fpair<float, 16, 17>(G[10], xs, S, D);
}
// FluxU of a level from its uniform spacings (iA, iB per axis), in double
// and rounded once; the constants of the w = -k contributions
inline FluxU flux_u_consts(double ax, double bx, double ay, double by, double az, double bz, double p2) {
  FluxU u;
  auto fill = [&](float *G, int AX, double k, double xk, double yk, double zk, double s1, double s2, double t1,
                  double t2, double ks, double kt) {
    const double w = -k;
    G[0] = (float)(w * (AX == 0 ? -4.0 / 3.0 : 2.0 / 3.0) * xk);
    G[1] = (float)(w * (AX == 1 ? -4.0 / 3.0 : 2.0 / 3.0) * yk);
    G[2] = (float)(w * (AX == 2 ? -4.0 / 3.0 : 2.0 / 3.0) * zk);
    G[3] = (float)(-2.0 * (-ks) * s1); G[4] = (float)(-2.0 * (-ks) * s2);
    G[5] = (float)(-2.0 * (-kt) * t1); G[6] = (float)(-2.0 * (-kt) * t2);
    const double wd = -k * p2;
    G[7] = (float)(-wd * xk); G[8] = (float)(-wd * yk); G[9] = (float)(-wd * zk);
    G[10] = (float)k;
  };
  // x row: xkf = 2 bx, ykf = ay, zkf = az; shears (by, bx) and (bz, bx)
  fill(u.g[0], 0, 2.0 * bx, ax, ay, az, by, bx, bz, bx, ay, az);
  // y row: xkf = ax, ykf = 2 by, zkf = az; shears (by, bx) with w = -xkf and (bz, by) with w = -zkf
  fill(u.g[1], 1, 2.0 * by, ax, ay, az, by, bx, bz, by, ax, az);
  // z row: xkf = ax, ykf = ay, zkf = 2 bz; shears (bz, bx) with w = -xkf and (bz, by) with w = -ykf
  fill(u.g[2], 2, 2.0 * bz, ax, ay, az, bz, bx, bz, by, ax, ay);
  return u;
}

// The same rows with their values read from the level's tables: (m1,m2,m3)
// index the spacing tables; v is the node's index in the viscosity arrays.
template <typename T>
__device__ __forceinline__ void row_x(const Geo<T> &g, int m1, int m2, int m3, long v, T s[18]) {
  row_x_v<T>(sp_load<T>(g, m1, m2, m3), vis_load<T, 1>(g, v), g.p2, s);
}
template <typename T>
__device__ __forceinline__ void row_y(const Geo<T> &g, int m1, int m2, int m3, long v, T s[18]) {
  row_y_v<T>(sp_load<T>(g, m1, m2, m3), vis_load<T, 2>(g, v), g.p2, s);
}
template <typename T>
__device__ __forceinline__ void row_z(const Geo<T> &g, int m1, int m2, int m3, long v, T s[18]) {
  row_z_v<T>(sp_load<T>(g, m1, m2, m3), vis_load<T, 3>(g, v), g.p2, s);
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
