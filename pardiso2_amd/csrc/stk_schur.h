// Pressure Schur-complement approximation of the block preconditioner:
// weighted BFBt with the velocity diagonal D as weight,
//     S^-1 ~ K^-1 (B D^-1 A D^-1 G) K^-1,   K = B D^-1 G,
// B the continuity rows (parallel/move3mg.c:1455-1573), G the pressure
// columns of the momentum rows (slots 16/17, move3mg.c:923-925), A the
// velocity block, D = |diagonal| of the momentum equations (boundary rows
// excluded, so walls act as Neumann faces of K).  K is a 7-point operator on
// the pressure cells of a level; its inverse is approximated by one
// aggregation multigrid V-cycle (piecewise-constant transfers along the
// level hierarchy's own cell nesting, Galerkin coarse operators, red/black
// Gauss-Seidel, dense inverse on the coarsest level).
#pragma once
#include <hip/hip_runtime.h>
#include "stk_rows.h"

namespace stk {

template <typename T> struct KPlanes { T *p[7]; };  // diag, -x, +x, -y, +y, -z, +z
template <typename T> struct KPlanesC {
  const T *p[7];
  int per;  // periodic axes of the level (bit 0 x, bit 1 y, bit 2 z)
};

// Neighbour offset of a cell across its lower (LO) or upper face along an
// axis of nn lines (cells 1 .. nn-1) with stride st; false: no neighbour.
// On a periodic axis cells 1 and nn-1 are neighbours across the wrap (the
// face velocity on line 0), and a Gauss-Seidel half-sweep (GS) takes that
// coupling only when the two cells have different colours (nn odd), so no
// cell reads a value its own half-sweep writes.
template <bool LO, bool GS>
__device__ __forceinline__ bool knb(int m, int nn, bool per, long st, long &off) {
  if (per && m == (LO ? 1 : nn - 1)) {
    if (GS && !(nn & 1)) return false;
    off = LO ? (long)(nn - 2) * st : -(long)(nn - 2) * st;
    return true;
  }
  if (LO ? m > 0 : m < nn - 1) {
    off = LO ? -st : st;
    return true;
  }
  return false;
}

// 1 / |diagonal| of the level's momentum equations; 0 on other unknowns
__global__ void k_bf_dinv(Geo<double> g, const unsigned char *__restrict__ flags, double *d64, float *d32) {
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= (long)g.XY * g.Z) return;
  int m1, m2, m3;
  decode(n, g.X, g.Y, g.XY, m1, m2, m3);
  const unsigned char f = flags[n];
  double s[18], v[4] = {0.0, 0.0, 0.0, 0.0};
  if (kind_of(f, 1) == KIND_EQ) { row_coef_c<double, 1>(g, m1, m2, m3, s); v[1] = 1.0 / fabs(s[1]); }
  if (kind_of(f, 2) == KIND_EQ) { row_coef_c<double, 2>(g, m1, m2, m3, s); v[2] = 1.0 / fabs(s[1]); }
  if (kind_of(f, 3) == KIND_EQ) { row_coef_c<double, 3>(g, m1, m2, m3, s); v[3] = 1.0 / fabs(s[1]); }
  for (int c = 0; c < 4; c++) {
    d64[4 * n + c] = v[c];
    d32[4 * n + c] = (float)v[c];
  }
}

// K = B D^-1 G on the finest level's continuity cells.  For axis a the +face
// of cell (m1,m2,m3) is the velocity at node (m - 1 + e_a), coefficient
// +iA_a[m_a] in B; the -face is at node (m - 1), coefficient -iA_a[m_a].  A
// face velocity u at axis index k has G = +2 iB_a[k] on its lower cell and
// -2 iB_a[k] on its upper cell.
__global__ void k_bf_kop(Geo<double> g, const unsigned char *__restrict__ flags, const double *__restrict__ dinv,
                         KPlanes<double> k64, KPlanes<float> k32) {
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= (long)g.XY * g.Z) return;
  int m1, m2, m3;
  decode(n, g.X, g.Y, g.XY, m1, m2, m3);
  double kc[7] = {0, 0, 0, 0, 0, 0, 0};
  if (kind_of(flags[n], 0) == KIND_EQ) {
    const long base = g.nid(m1 - 1, m2 - 1, m3 - 1);  // node of the three -faces
    const long step[3] = {(long)g.Y, 1L, (long)g.XY};
    const int mm[3] = {m1, m2, m3};
    const double *iA[3] = {g.iAx, g.iAy, g.iAz};
    const double *iB[3] = {g.iBx, g.iBy, g.iBz};
    const int nn[3] = {g.X, g.Y, g.Z};
    for (int a = 0; a < 3; a++) {
      const int m = mm[a];
      // on a periodic axis the +face of the last cell is the velocity on line 0
      // (the coupling across the wrap is left out of K: its diagonal keeps both faces)
      const long up = ((g.per >> a) & 1) && m == nn[a] - 1 ? base - (long)(nn[a] - 2) * step[a] : base + step[a];
      const double ap = iA[a][m] * dinv[4 * up + 1 + a] * 2.0 * iB[a][m];
      const double am = iA[a][m] * dinv[4 * base + 1 + a] * 2.0 * iB[a][m - 1];
      kc[0] += ap + am;
      kc[1 + 2 * a] = -am;
      kc[2 + 2 * a] = -ap;
    }
  }
  for (int q = 0; q < 7; q++) {
    k64.p[q][n] = kc[q];
    k32.p[q][n] = (float)kc[q];
  }
}

// y = D^-1 G p on the velocity unknowns of level 0 (P slots and
// non-equation unknowns 0); p is a cell vector (one value per node)
template <typename T>
__global__ void k_bf_grad(int X, int Y, int Z, const T *__restrict__ iBx, const T *__restrict__ iBy,
                          const T *__restrict__ iBz, const T *__restrict__ dinv, const T *__restrict__ p, T *y,
                          int per) {
  long n = xcd_block() * blockDim.x + threadIdx.x;
  const int XY = X * Y;
  if (n >= (long)XY * Z) return;
  int m1, m2, m3;
  decode(n, X, Y, XY, m1, m2, m3);
  y[4 * n] = T(0);
  const long up = n + XY + Y + 1;  // cell (m1+1, m2+1, m3+1)
  const long stp[3] = {(long)Y, 1L, (long)XY};
  const int mm[3] = {m1, m2, m3};
  const int nn[3] = {X, Y, Z};
  const T *iB[3] = {iBx, iBy, iBz};
  for (int a = 0; a < 3; a++) {
    const T d = dinv[4 * n + 1 + a];
    // the lower cell of a velocity on line 0 of a periodic axis is cell n-1
    const long lo = ((per >> a) & 1) && mm[a] == 0 ? up + (long)(nn[a] - 2) * stp[a] : up - stp[a];
    T v = T(0);
    if (d != T(0)) v = d * (T(2) * iB[a][mm[a]]) * (p[lo] - p[up]);
    y[4 * n + 1 + a] = v;
  }
}

// t = B D^-1 w on the continuity cells of level 0 (0 elsewhere)
template <typename T>
__global__ void k_bf_div(int X, int Y, int Z, const unsigned char *__restrict__ flags, const T *__restrict__ iAx,
                         const T *__restrict__ iAy, const T *__restrict__ iAz, const T *__restrict__ dinv,
                         const T *__restrict__ w, T *t, int per) {
  long n = xcd_block() * blockDim.x + threadIdx.x;
  const int XY = X * Y;
  if (n >= (long)XY * Z) return;
  if (kind_of(flags[n], 0) != KIND_EQ) { t[n] = T(0); return; }
  int m1, m2, m3;
  decode(n, X, Y, XY, m1, m2, m3);
  const long base = n - XY - Y - 1;
  const long stp[3] = {(long)Y, 1L, (long)XY};
  const int mm[3] = {m1, m2, m3};
  const int nn[3] = {X, Y, Z};
  const T *iA[3] = {iAx, iAy, iAz};
  T acc = T(0);
  for (int a = 0; a < 3; a++) {
    // the +face of the last cell of a periodic axis is the velocity on line 0
    const long un = ((per >> a) & 1) && mm[a] == nn[a] - 1 ? base - (long)(nn[a] - 2) * stp[a] : base + stp[a];
    const long up = 4 * un + 1 + a, lo = 4 * base + 1 + a;
    acc += iA[a][mm[a]] * (dinv[up] * w[up] - dinv[lo] * w[lo]);
  }
  t[n] = acc;
}

// cell vector from / to the pressure slots of an AoS vector (continuity cells only)
template <typename T>
__global__ void k_bf_gather(long nn, const unsigned char *__restrict__ flags, const double *__restrict__ r, T *b) {
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= nn) return;
  b[n] = (kind_of(flags[n], 0) == KIND_EQ) ? (T)r[4 * n] : T(0);
}

template <typename T>
__global__ void k_bf_scatter(long nn, const unsigned char *__restrict__ flags, const T *__restrict__ x, double sign,
                             double *z) {
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= nn) return;
  z[4 * n + 0] = (kind_of(flags[n], 0) == KIND_EQ) ? sign * (double)x[n] : 0.0;
  z[4 * n + 1] = 0.0;
  z[4 * n + 2] = 0.0;
  z[4 * n + 3] = 0.0;
}

// Galerkin coarse K by aggregation: coarse cell (cx,cy,cz) owns fine cells
// lo_a[c] .. lo_a[c]+cnt_a[c]-1 per axis; couplings inside the aggregate
// fold into its diagonal, couplings across a face go to that face's plane.
struct KAgg {
  const int *lo[3];
  const int *cnt[3];
};

template <typename T>
__global__ void k_kagg(int Xf, int Yf, int Xc, int Yc, int Zc, KAgg t, KPlanesC<T> kf, KPlanes<T> kc) {
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int XYc = Xc * Yc, XYf = Xf * Yf;
  if (n >= (long)XYc * Zc) return;
  int cx, cy, cz;
  decode(n, Xc, Yc, XYc, cx, cy, cz);
  T acc[7] = {T(0), T(0), T(0), T(0), T(0), T(0), T(0)};
  const int c3[3] = {cx, cy, cz};
  int lo[3], cnt[3];
  for (int a = 0; a < 3; a++) {
    lo[a] = t.lo[a][c3[a]];
    cnt[a] = t.cnt[a][c3[a]];
  }
  for (int iz = 0; iz < cnt[2]; iz++)
    for (int ix = 0; ix < cnt[0]; ix++)
      for (int iy = 0; iy < cnt[1]; iy++) {
        const long f = (long)(lo[2] + iz) * XYf + (long)(lo[0] + ix) * Yf + (lo[1] + iy);
        acc[0] += kf.p[0][f];
        const int ii[3] = {ix, iy, iz};
        for (int a = 0; a < 3; a++) {
          const T km = kf.p[1 + 2 * a][f], kp = kf.p[2 + 2 * a][f];
          if (ii[a] > 0) acc[0] += km; else acc[1 + 2 * a] += km;
          if (ii[a] < cnt[a] - 1) acc[0] += kp; else acc[2 + 2 * a] += kp;
        }
      }
  for (int q = 0; q < 7; q++) kc.p[q][n] = acc[q];
}

// one red/black Gauss-Seidel half-sweep of K x = b on cells of one colour
// new value of cell n (false: the cell keeps its value)
template <typename T>
__device__ __forceinline__ bool kgs_at(int X, int Y, int Z, const KPlanesC<T> &k, const T *__restrict__ b,
                                       const T *x, long n, int m1, int m2, int m3, T &out) {
  const int XY = X * Y;
  const T d = k.p[0][n];
  if (d == T(0)) return false;
  T acc = b[n];
  long o;
  if (knb<true, true>(m1, X, k.per & 1, Y, o)) acc -= k.p[1][n] * x[n + o];
  if (knb<false, true>(m1, X, k.per & 1, Y, o)) acc -= k.p[2][n] * x[n + o];
  if (knb<true, true>(m2, Y, k.per & 2, 1, o)) acc -= k.p[3][n] * x[n + o];
  if (knb<false, true>(m2, Y, k.per & 2, 1, o)) acc -= k.p[4][n] * x[n + o];
  if (knb<true, true>(m3, Z, k.per & 4, XY, o)) acc -= k.p[5][n] * x[n + o];
  if (knb<false, true>(m3, Z, k.per & 4, XY, o)) acc -= k.p[6][n] * x[n + o];
  out = acc / d;
  return true;
}

template <typename T>
__device__ __forceinline__ bool kgs_val(int X, int Y, int Z, const KPlanesC<T> &k, const T *__restrict__ b,
                                        const T *x, int colour, long n, T &out) {
  int m1, m2, m3;
  decode(n, X, Y, X * Y, m1, m2, m3);
  if (((m1 + m2 + m3) & 1) != colour) return false;  // before any load: half the cells stop here
  return kgs_at<T>(X, Y, Z, k, b, x, n, m1, m2, m3, out);
}

template <typename T>
__device__ __forceinline__ void kgs_cell(int X, int Y, int Z, const KPlanesC<T> &k, const T *__restrict__ b, T *x,
                                         int colour, long n) {
  T v;
  if (kgs_val<T>(X, Y, Z, k, b, x, colour, n, v)) x[n] = v;
}

template <typename T>
__global__ void k_kgs(int X, int Y, int Z, KPlanesC<T> k, const T *__restrict__ b, T *x, int colour) {
  long n = xcd_block() * blockDim.x + threadIdx.x;
  if (n >= (long)X * Y * Z) return;
  kgs_cell<T>(X, Y, Z, k, b, x, colour, n);
}

// The same half-sweep with one thread per cell of the colour (no lanes idle
// on the other colour): thread i is cell m2 = 2 kk + parity of the row
// q = (m3, m1), kk = i mod hy, q = i / hy with hy = ceil(Y/2).  The
// multiply-high divisions are exact while hy^2 X Z < 2^32 (the caller checks).
// The seven planes come colour-compact (kc[(p * 2 + colour) * ncol + i], from
// k_kplanes_colour), so a half-sweep reads only its own colour's
// coefficients, each once and contiguous; the arithmetic is kgs_at's.
__device__ __forceinline__ void kcell_of(unsigned i, int colour, int X, int Y, FDiv dhy, FDiv dx, int &m1, int &m2,
                                         int &m3) {
  const int hy = (Y + 1) / 2;
  const int q = (int)fdiv(i, dhy), kk = (int)i - q * hy;
  m3 = (int)fdiv(q, dx);
  m1 = q - m3 * X;
  m2 = 2 * kk + ((colour - m1 - m3) & 1);
}

template <typename T>
__global__ void k_kgs_c(int X, int Y, int Z, FDiv dhy, FDiv dx, const T *__restrict__ kc, const T *__restrict__ b,
                        T *x, int colour, int per) {
  const unsigned i = (unsigned)(xcd_block() * blockDim.x + threadIdx.x);
  const unsigned ncol = (unsigned)((Y + 1) / 2) * X * Z;
  if (i >= ncol) return;
  int m1, m2, m3;
  kcell_of(i, colour, X, Y, dhy, dx, m1, m2, m3);
  if (m2 >= Y) return;
  const long n = ((long)m3 * X + m1) * Y + m2;
  const int XY = X * Y;
  const T *__restrict__ p = kc + (size_t)colour * ncol + i;  // plane q at p[2 q ncol]
  const T d = p[0];
  if (d == T(0)) return;
  T acc = b[n];
  long o;
  if (knb<true, true>(m1, X, per & 1, Y, o)) acc -= p[2 * ncol] * x[n + o];
  if (knb<false, true>(m1, X, per & 1, Y, o)) acc -= p[4 * (size_t)ncol] * x[n + o];
  if (knb<true, true>(m2, Y, per & 2, 1, o)) acc -= p[6 * (size_t)ncol] * x[n + o];
  if (knb<false, true>(m2, Y, per & 2, 1, o)) acc -= p[8 * (size_t)ncol] * x[n + o];
  if (knb<true, true>(m3, Z, per & 4, XY, o)) acc -= p[10 * (size_t)ncol] * x[n + o];
  if (knb<false, true>(m3, Z, per & 4, XY, o)) acc -= p[12 * (size_t)ncol] * x[n + o];
  x[n] = acc / d;
}

// The colour-compact copy of a level's seven K planes (0 past the row ends).
template <typename T>
__global__ void k_kplanes_colour(int X, int Y, int Z, FDiv dhy, FDiv dx, KPlanesC<T> k, T *kc) {
  const unsigned t = (unsigned)(blockIdx.x * blockDim.x + threadIdx.x);
  const unsigned ncol = (unsigned)((Y + 1) / 2) * X * Z;
  if (t >= 2 * ncol) return;
  const int colour = t >= ncol;
  const unsigned i = t - (colour ? ncol : 0u);
  int m1, m2, m3;
  kcell_of(i, colour, X, Y, dhy, dx, m1, m2, m3);
  const long n = ((long)m3 * X + m1) * Y + m2;
#pragma unroll
  for (int q = 0; q < 7; q++) kc[(size_t)(q * 2 + colour) * ncol + i] = m2 < Y ? k.p[q][n] : T(0);
}

// r = b - K x
template <typename T>
__device__ __forceinline__ T kres_val(int X, int Y, int Z, const KPlanesC<T> &k, const T *__restrict__ b,
                                      const T *__restrict__ x, long n) {
  const int XY = X * Y;
  const T d = k.p[0][n];
  if (d == T(0)) return T(0);
  int m1, m2, m3;
  decode(n, X, Y, XY, m1, m2, m3);
  T acc = b[n] - d * x[n];
  long o;
  if (knb<true, false>(m1, X, k.per & 1, Y, o)) acc -= k.p[1][n] * x[n + o];
  if (knb<false, false>(m1, X, k.per & 1, Y, o)) acc -= k.p[2][n] * x[n + o];
  if (knb<true, false>(m2, Y, k.per & 2, 1, o)) acc -= k.p[3][n] * x[n + o];
  if (knb<false, false>(m2, Y, k.per & 2, 1, o)) acc -= k.p[4][n] * x[n + o];
  if (knb<true, false>(m3, Z, k.per & 4, XY, o)) acc -= k.p[5][n] * x[n + o];
  if (knb<false, false>(m3, Z, k.per & 4, XY, o)) acc -= k.p[6][n] * x[n + o];
  return acc;
}

template <typename T>
__device__ __forceinline__ void kres_cell(int X, int Y, int Z, const KPlanesC<T> &k, const T *__restrict__ b,
                                          const T *__restrict__ x, T *r, long n) {
  r[n] = kres_val<T>(X, Y, Z, k, b, x, n);
}

template <typename T>
__global__ void k_kres(int X, int Y, int Z, KPlanesC<T> k, const T *__restrict__ b, const T *__restrict__ x, T *r) {
  long n = xcd_block() * blockDim.x + threadIdx.x;
  if (n >= (long)X * Y * Z) return;
  kres_cell<T>(X, Y, Z, k, b, x, r, n);
}

// b_c = sum of the aggregate's fine residuals, x_c = 0
template <typename T>
__device__ __forceinline__ void krestrict_cell(int Xf, int Yf, int Xc, int Yc, const KAgg &t,
                                               const T *__restrict__ rf, T *bc, T *xc, long n) {
  const int XYc = Xc * Yc, XYf = Xf * Yf;
  int cx, cy, cz;
  decode(n, Xc, Yc, XYc, cx, cy, cz);
  const int lx = t.lo[0][cx], ly = t.lo[1][cy], lz = t.lo[2][cz];
  const int nx = t.cnt[0][cx], ny = t.cnt[1][cy], nz = t.cnt[2][cz];
  T acc = T(0);
  for (int iz = 0; iz < nz; iz++)
    for (int ix = 0; ix < nx; ix++)
      for (int iy = 0; iy < ny; iy++) acc += rf[(long)(lz + iz) * XYf + (long)(lx + ix) * Yf + (ly + iy)];
  bc[n] = acc;
  xc[n] = T(0);
}

template <typename T>
__global__ void k_krestrict(int Xf, int Yf, int Xc, int Yc, int Zc, KAgg t, const T *__restrict__ rf, T *bc, T *xc) {
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= (long)Xc * Yc * Zc) return;
  krestrict_cell<T>(Xf, Yf, Xc, Yc, t, rf, bc, xc, n);
}

// x_f += scale * x_c(parent) on fine cells with an equation
template <typename T>
__device__ __forceinline__ bool kprolong_val(int Xf, int Yf, int Xc, int Yc, const int *__restrict__ px,
                                             const int *__restrict__ py, const int *__restrict__ pz,
                                             const T *__restrict__ kdiag, const T *__restrict__ xc, const T *xf,
                                             T scale, long n, T &out) {
  if (kdiag[n] == T(0)) return false;
  const int XYf = Xf * Yf;
  int m1, m2, m3;
  decode(n, Xf, Yf, XYf, m1, m2, m3);
  out = xf[n] + scale * xc[(long)pz[m3] * Xc * Yc + (long)px[m1] * Yc + py[m2]];
  return true;
}

template <typename T>
__device__ __forceinline__ void kprolong_cell(int Xf, int Yf, int Xc, int Yc, const int *__restrict__ px,
                                              const int *__restrict__ py, const int *__restrict__ pz,
                                              const T *__restrict__ kdiag, const T *__restrict__ xc, T *xf, T scale,
                                              long n) {
  T v;
  if (kprolong_val<T>(Xf, Yf, Xc, Yc, px, py, pz, kdiag, xc, xf, scale, n, v)) xf[n] = v;
}

template <typename T>
__global__ void k_kprolong(int Xf, int Yf, int Zf, int Xc, int Yc, const int *__restrict__ px,
                           const int *__restrict__ py, const int *__restrict__ pz, const T *__restrict__ kdiag,
                           const T *__restrict__ xc, T *xf, T scale) {
  long n = xcd_block() * blockDim.x + threadIdx.x;
  if (n >= (long)Xf * Yf * Zf) return;
  kprolong_cell<T>(Xf, Yf, Xc, Yc, px, py, pz, kdiag, xc, xf, scale, n);
}

// coarsest level: x[cell[i]] = sum_j inv[i][j] b[cell[j]], from the
// transposed inverse invT[j][i] (lanes read neighbouring rows) and b of the
// active cells staged in LDS (nc values)
template <typename T>
__global__ void k_kcoarse(int nc, const int *__restrict__ cell, const T *__restrict__ invT, const T *__restrict__ b,
                          T *x) {
  extern __shared__ unsigned char smem_raw[];
  T *sb = reinterpret_cast<T *>(smem_raw);
  for (int j = threadIdx.x; j < nc; j += blockDim.x) sb[j] = b[cell[j]];
  __syncthreads();
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  T acc = T(0);
  for (int j = 0; j < nc; j++) acc += invT[(long)j * nc + i] * sb[j];
  x[cell[i]] = acc;
}

// ---------------------------------------------------------------------------
// The small levels of the K-cycle in one workgroup: the same cell updates,
// transfers and coarsest solve as the kernels above, level after level, with
// workgroup barriers where the separate kernels had launch boundaries.
// lv[0 .. nlev-1] are the smoothed levels, lv[nlev] the coarsest; agg of
// lv[i+1] groups the cells of lv[i], px/py/pz of lv[i] name the parents.
// A thread owns cells tid, tid + 1024, ... (at most kTailPT); it computes
// their new values kTailB at a time before storing them, so their loads
// overlap (a phase reads only cells it does not write).  The coarsest product reads the
// transposed inverse (one thread per row, coalesced) in k_kcoarse's order.
// ---------------------------------------------------------------------------
constexpr int kKTail = 4;        // smoothed levels in the tail
constexpr int kTailPT = 2;       // cells per thread (levels of at most 2048 cells: one workgroup is
                                 // faster than launches only while a phase is about one memory round trip)
constexpr int kTailB = 2;        // cells computed together
constexpr int kTailCoarse = 2048;
template <typename T> struct KTailLevel {
  int X, Y, Z;
  KPlanesC<T> k;
  T *b, *x, *r;
  KAgg agg;
  const int *px, *py, *pz;
};
template <typename T> struct KTail {
  int nlev, nu, nkc;
  int stage;  // 1: the first level's planes and vectors are staged in dynamic LDS
  T scale;
  const int *kcell;
  const T *kinvT;  // [j * nkc + i] = inverse(i, j)
  KTailLevel<T> lv[kKTail + 1];
};

template <typename T>
__device__ __forceinline__ void tail_kgs(const KTailLevel<T> &V, int col) {
  const long n = (long)V.X * V.Y * V.Z;
  // thread index the compiler may not assume constant across phases: keeps
  // it from holding every cell's plane addresses in registers for the kernel
  long t0 = threadIdx.x;
  asm volatile("" : "+v"(t0));
#pragma unroll 1
  for (int q0 = 0; q0 < kTailPT; q0 += kTailB) {
    T v[kTailB];
    unsigned ok = 0;
#pragma unroll
    for (int q = 0; q < kTailB; q++) {
      const long u = t0 + (long)(q0 + q) * blockDim.x;
      if (u < n && kgs_val<T>(V.X, V.Y, V.Z, V.k, V.b, V.x, col, u, v[q])) ok |= 1u << q;
    }
#pragma unroll
    for (int q = 0; q < kTailB; q++)
      if (ok & (1u << q)) V.x[t0 + (long)(q0 + q) * blockDim.x] = v[q];
  }
}

template <typename T>
__device__ __forceinline__ void tail_kres(const KTailLevel<T> &V) {
  const long n = (long)V.X * V.Y * V.Z;
  // thread index the compiler may not assume constant across phases: keeps
  // it from holding every cell's plane addresses in registers for the kernel
  long t0 = threadIdx.x;
  asm volatile("" : "+v"(t0));
#pragma unroll 1
  for (int q0 = 0; q0 < kTailPT; q0 += kTailB) {
    T v[kTailB];
#pragma unroll
    for (int q = 0; q < kTailB; q++) {
      const long u = t0 + (long)(q0 + q) * blockDim.x;
      if (u < n) v[q] = kres_val<T>(V.X, V.Y, V.Z, V.k, V.b, V.x, u);
    }
#pragma unroll
    for (int q = 0; q < kTailB; q++) {
      const long u = t0 + (long)(q0 + q) * blockDim.x;
      if (u < n) V.r[u] = v[q];
    }
  }
}

template <typename T>
__device__ __forceinline__ void tail_kprolong(const KTailLevel<T> &V, const KTailLevel<T> &C, T scale) {
  const long n = (long)V.X * V.Y * V.Z;
  // thread index the compiler may not assume constant across phases: keeps
  // it from holding every cell's plane addresses in registers for the kernel
  long t0 = threadIdx.x;
  asm volatile("" : "+v"(t0));
#pragma unroll 1
  for (int q0 = 0; q0 < kTailPT; q0 += kTailB) {
    T v[kTailB];
    unsigned ok = 0;
#pragma unroll
    for (int q = 0; q < kTailB; q++) {
      const long u = t0 + (long)(q0 + q) * blockDim.x;
      if (u < n && kprolong_val<T>(V.X, V.Y, C.X, C.Y, V.px, V.py, V.pz, V.k.p[0], C.x, V.x, scale, u, v[q]))
        ok |= 1u << q;
    }
#pragma unroll
    for (int q = 0; q < kTailB; q++)
      if (ok & (1u << q)) V.x[t0 + (long)(q0 + q) * blockDim.x] = v[q];
  }
}

// NL = t.nlev as a constant, so that every lv[] index is a constant and the
// argument block stays in scalar registers
template <typename T, int NL>
__global__ void __launch_bounds__(1024) k_ktail(KTail<T> t) {
  __shared__ T sb[kTailCoarse];
  extern __shared__ unsigned char smem_raw[];
  const int tid = threadIdx.x, nt = blockDim.x;
  // stage: the first level's seven planes, b, x and r in LDS for all its
  // phases (x goes back to HBM at the end); the same code runs on either copy
  KTailLevel<T> lv0 = t.lv[0];
  if (NL >= 1 && t.stage) {
    const long n0 = (long)lv0.X * lv0.Y * lv0.Z;
    T *sl = reinterpret_cast<T *>(smem_raw);
    for (long u = tid; u < n0; u += nt) {
#pragma unroll
      for (int q = 0; q < 7; q++) sl[q * n0 + u] = t.lv[0].k.p[q][u];
      sl[7 * n0 + u] = t.lv[0].b[u];
      sl[8 * n0 + u] = t.lv[0].x[u];
    }
    for (int q = 0; q < 7; q++) lv0.k.p[q] = sl + q * n0;
    lv0.b = sl + 7 * n0;
    lv0.x = sl + 8 * n0;
    lv0.r = sl + 9 * n0;
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < NL; i++) {
    const KTailLevel<T> &V = i == 0 ? lv0 : t.lv[i];
    const long n = (long)V.X * V.Y * V.Z;
#pragma unroll 1
    for (int s = 0; s < t.nu; s++)
      for (int col = 0; col < 2; col++) {
        tail_kgs<T>(V, col);
        __syncthreads();
      }
    tail_kres<T>(V);
    __syncthreads();
    const KTailLevel<T> &C = t.lv[i + 1];
    const long nc = (long)C.X * C.Y * C.Z;
    for (long u = tid; u < nc; u += nt) krestrict_cell<T>(V.X, V.Y, C.X, C.Y, C.agg, V.r, C.b, C.x, u);
    __syncthreads();
  }
  {
    const KTailLevel<T> &C = t.lv[NL];
    for (int j = tid; j < t.nkc; j += nt) sb[j] = C.b[t.kcell[j]];
    __syncthreads();
    for (int row = tid; row < t.nkc; row += nt) {
      T acc = T(0);
      for (int j = 0; j < t.nkc; j++) acc += t.kinvT[(long)j * t.nkc + row] * sb[j];
      C.x[t.kcell[row]] = acc;
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = NL - 1; i >= 0; i--) {
    const KTailLevel<T> &V = i == 0 ? lv0 : t.lv[i], &C = t.lv[i + 1];
    tail_kprolong<T>(V, C, t.scale);
    __syncthreads();
#pragma unroll 1
    for (int s = 0; s < t.nu; s++)
      for (int col = 1; col >= 0; col--) {
        tail_kgs<T>(V, col);
        __syncthreads();
      }
  }
  if (NL >= 1 && t.stage) {
    const long n0 = (long)lv0.X * lv0.Y * lv0.Z;
    for (long u = tid; u < n0; u += nt) t.lv[0].x[u] = lv0.x[u];
  }
}

}  // namespace stk
