// stk_compat.h -- device sweeps over a caller's assembled row store (the
// reference's val1/lin0/num0/pos0/fre1, parallel/head3mg.c:59-62), for the
// reference-named gausmat4 acting on bound arrays.
//
// The rows of one level are gathered on the host into a padded row-major
// ELL block (kEllMax entries per row, level-local columns) in the order the
// reference stored them, so each row update sums its entries in the
// reference's order (move3mg.c:3893-3899).
#pragma once
#include <hip/hip_runtime.h>
#include "stk_kernels.h"

namespace stk {

constexpr int kEllMax = 24;  // momentum rows carry 17 entries (move3mg.c:903 etc.)

// gausmat4 row update (move3mg.c:3892-3923) of momentum or boundary row r in place
__device__ __forceinline__ void ell_update(long r, int k, const int *__restrict__ col, const double *__restrict__ val,
                                           const double *__restrict__ f, bool bc, double v0, double *x) {
  const int *cr = col + r * kEllMax;
  const double *vr = val + r * kEllMax;
  double ival = -f[r];
  const double ival1 = vr[0];
  for (int q = 0; q < k; q++) ival += vr[q] * x[cr[q]];
  if (bc) x[r] -= ival;
  else x[r] -= ival / ival1 * v0;
}

// One phase of the multicolour sweep on a level: momentum rows of
// component c (1..3) whose node has colour (m1+m2+m3)&1 == col.  Rows of
// one phase read no unknown another row of the phase writes.
__global__ void k_ell_phase(long nn, int Y, int XY, int c, int col, const int *__restrict__ cnt,
                            const int *__restrict__ ecol, const double *__restrict__ val, const double *__restrict__ f,
                            const unsigned char *__restrict__ isbc, double v0, double *x) {
  const long node = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (node >= nn) return;
  const int m3 = (int)(node / XY), rem = (int)(node - (long)m3 * XY), m1 = rem / Y, m2 = rem - m1 * Y;
  if (((m1 + m2 + m3) & 1) != col) return;
  const long r = 4 * node + c;
  const int k = cnt[r];
  if (!k || isbc[r]) return;
  ell_update(r, k, ecol, val, f, false, v0, x);
}

// continuity rows: relaxed by (ival * nd) * p0koef as the reference groups it
__global__ void k_ell_pphase(long nn, const int *__restrict__ cnt, const int *__restrict__ ecol,
                             const double *__restrict__ val, const double *__restrict__ f,
                             const unsigned char *__restrict__ isbc, const float *__restrict__ nd, double p0,
                             double *x) {
  const long node = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (node >= nn) return;
  const long r = 4 * node;
  const int k = cnt[r];
  if (!k || isbc[r]) return;
  const int *cr = ecol + r * kEllMax;
  const double *vr = val + r * kEllMax;
  double ival = -f[r];
  for (int q = 0; q < k; q++) ival += vr[q] * x[cr[q]];
  x[r] -= ival * nd[node] * p0;
}

// boundary rows of one dependency depth (rows[i], level-local)
__global__ void k_ell_bc(int nr, const int *__restrict__ rows, const int *__restrict__ cnt,
                         const int *__restrict__ ecol, const double *__restrict__ val, const double *__restrict__ f,
                         double *x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nr) return;
  const long r = rows[i];
  ell_update(r, cnt[r], ecol, val, f, true, 0.0, x);
}

// fre0 = fre1 - A sol0 on level-local rows [r0, r1) with entries
// (move3mg.c:3935-3957)
__global__ void k_ell_res(long r0, long r1, const int *__restrict__ cnt, const int *__restrict__ ecol,
                          const double *__restrict__ val, const double *__restrict__ f, const double *__restrict__ x,
                          double *f0) {
  const long r = r0 + blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (r >= r1) return;
  const int k = cnt[r];
  if (!k) return;
  const int *cr = ecol + r * kEllMax;
  const double *vr = val + r * kEllMax;
  double ival = f[r];
  for (int q = 0; q < k; q++) ival -= vr[q] * x[cr[q]];
  f0[r] = ival;
}

}  // namespace stk
