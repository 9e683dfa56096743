// stk_heat.h -- the implicit temperature step on the device (restates the
// heat-equation multigrid of parallel/heat3mg.c: heaterr 6-290, titerate
// 292-616, tprolong 954-1050, gausmatt 1057-1120, trestrict 1122-1232,
// rocprestrict 1235-1330).
//
// Rows are not stored: a node's 7-point row is rebuilt from the FP32
// conductivity kt and the clamped FP32 rho*cp wherever it is needed, in the
// reference's operation order (face mean as an FP32 sum, then / 2.0 * kf /
// spacing in FP64; diagonal rho cp / dt plus the faces in the order x-, x+,
// y-, y+, z-, z+), so every coefficient equals the reference's assembled
// one bit for bit.  A sweep reads 8 bytes of temperature, 8 of right-hand
// side and 8 of properties per node instead of a 56-byte row.
//
// Gauss-Seidel runs in two colours of (m1 + m2 + m3) parity -- the 7-point
// rows of one colour read only the other colour -- then the boundary rows by
// dependency depth.  On an x-periodic level with an even line count the
// rows on line 0 read line X-2 of their own colour; they run first, so they
// read it before its update, as the reference's loop order does.
//
// Blocks of the node kernels are renumbered by xcd_block (each XCD owns a
// contiguous slab of lines, so the x and z neighbour lines a block reads
// were read by blocks of the same L2).
//
// Layout: the solve's arrays hold every y line split by parity, the nodes
// of even m2 first (position m2/2) and those of odd m2 after them (position
// (Y+1)/2 + m2/2); lines, planes and levels are where the reference puts
// them.  One colour of a line is then one contiguous run, its y neighbours
// the other run of the same line, its x and z neighbours the same run of
// the next lines: a colour pass reads and writes only what it uses, in
// whole cache lines.  The boundary converts to and from the reference's
// order.

namespace stk {

// unsigned n / d for n < 2^31 as a multiply and shift: m = ceil(2^(32+s) / d)
// with 2^s >= d (m < 2^33, so n m stays below 2^64; the error term
// n (m d - 2^(32+s)) / (d 2^(32+s)) stays below 1/d)
struct HDiv {
  unsigned long long m;
  int s;
};
__host__ __device__ __forceinline__ unsigned hdiv(unsigned n, const HDiv &f) {
  return (unsigned)(((unsigned long long)n * f.m) >> (32 + f.s));
}
inline HDiv make_hdiv(unsigned d) {
  HDiv f;
  f.s = 0;
  while ((1ull << f.s) < d) f.s++;
  const unsigned long long p = 1ull << (32 + f.s);
  f.m = (p + d - 1) / d;
  return f;
}

struct HeatGeo {
  int X, Y, Z, XY;
  int hy;                   // (Y + 1) / 2: position of the first odd-m2 node of a line
  HDiv fY, fX;           // / Y, / X
  int xper;                 // x wraps: line 0's x- neighbour is line X-2 (heat3mg.c:21, 36)
  const double *kf[3];      // 2 / (g[m+1] - g[m-1]) per line (x line 0 of a periodic axis: the wrapped span)
  const double *dm[3];      // g[m] - g[m-1] (x line 0 periodic: g[X-1] - g[X-2])
  const double *dp[3];      // g[m+1] - g[m]
};



// index of node (m1, m2, m3) in the split-line layout
__host__ __device__ __forceinline__ long hidx(int X, int Y, int hy, int m1, int m2, int m3) {
  return ((long)m3 * X + m1) * Y + ((m2 & 1) ? hy : 0) + (m2 >> 1);
}
__device__ __forceinline__ void hdecode(const HeatGeo &g, long n, int &m1, int &m2, int &m3) {
  const unsigned L = hdiv((unsigned)n, g.fY);
  const int pos = (int)(n - (long)L * g.Y);
  m2 = pos < g.hy ? 2 * pos : 2 * (pos - g.hy) + 1;
  m3 = (int)hdiv(L, g.fX);
  m1 = (int)L - m3 * g.X;
}

// the row of heat-equation node n = (m1, m2, m3): neighbours x-, x+, y-, y+,
// z-, z+, their conductances (the negated off-diagonals) and the diagonal
// (heat3mg.c:195-249)
struct HeatRow {
  long nb[6];
  double c[6];
  double d;
};

__device__ __forceinline__ void heat_row(const HeatGeo &g, const float *__restrict__ kt, float rcp, double dt,
                                         int m1, int m2, int m3, long n, HeatRow &h) {
  h.nb[0] = (g.xper && m1 == 0) ? n + (long)(g.X - 2) * g.Y : n - g.Y;
  h.nb[1] = n + g.Y;
  // y neighbours: the other parity run of the same line
  const long k = m2 >> 1;
  const long base = n - ((m2 & 1) ? g.hy : 0) - k;
  h.nb[2] = (m2 & 1) ? base + k : base + g.hy + k - 1;
  h.nb[3] = (m2 & 1) ? base + k + 1 : base + g.hy + k;
  h.nb[4] = n - g.XY;
  h.nb[5] = n + g.XY;
  const float k0 = kt[n];
  const int mm[3] = {m1, m2, m3};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const float sm = kt[h.nb[2 * a]] + k0;      // FP32 sums, as the reference's float arrays add
    const float sp = k0 + kt[h.nb[2 * a + 1]];
    const int m = mm[a];
    const double kf = g.kf[a][m];
    h.c[2 * a] = (double)sm / 2.0 * kf / g.dm[a][m];
    h.c[2 * a + 1] = (double)sp / 2.0 * kf / g.dp[a][m];
  }
  double d = (double)rcp / dt;
#pragma unroll
  for (int k = 0; k < 6; k++) d += h.c[k];
  h.d = d;
}

// rho cp of a node, FP32 product clamped below at 2e6 (heat3mg.c:46-47)
__global__ void k_heat_rcp(long n, const float *__restrict__ ro, const float *__restrict__ cp, float *rcp) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float r = ro[i] * cp[i];
  rcp[i] = r < 2e+6f ? 2e+6f : r;
}

// rocprestrict (heat3mg.c:1235-1330): every coarse node gathers the fine
// nodes of the 8 cells around it, (z, x, y) in the reference's scan order,
// weights (1 - |d|) per axis multiplied x * y * z; FP32 accumulators of FP64
// increments and an FP64 weight sum, as the reference's float arrays; then
// the coarse rho cp
__global__ void k_heat_props(int Xf, int Yf, int Xc, int Yc, int Zc, CompTab t, const float *__restrict__ rof,
                             const float *__restrict__ cpf, const float *__restrict__ ktf, float *roc, float *cpc,
                             float *ktc, float *rcpc) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int XYc = Xc * Yc, XYf = Xf * Yf, hyf = (Yf + 1) / 2;
  if (q >= (long)XYc * Zc) return;
  int cx, cy, cz;
  decode(q, Xc, Yc, XYc, cx, cy, cz);
  const long n = hidx(Xc, Yc, (Yc + 1) / 2, cx, cy, cz);
  float acp = 0.f, akt = 0.f, aro = 0.f;
  double ws = 0.0;
  for (int a = t.ax[2].cbeg[cz]; a < t.ax[2].cbeg[cz + 1]; a++) {
    const int k = t.ax[2].cfi[a];
    const double wz = t.ax[2].cw64[a];
    for (int bq = t.ax[0].cbeg[cx]; bq < t.ax[0].cbeg[cx + 1]; bq++) {
      const int i = t.ax[0].cfi[bq];
      const double wx = t.ax[0].cw64[bq];
      for (int e = t.ax[1].cbeg[cy]; e < t.ax[1].cbeg[cy + 1]; e++) {
        const int j = t.ax[1].cfi[e];
        const double wy = t.ax[1].cw64[e];
        const long nf = (long)k * XYf + (long)i * Yf + ((j & 1) ? hyf : 0) + (j >> 1);
        const double swt = wx * wy * wz;
        acp = (float)((double)acp + cpf[nf] * swt);
        akt = (float)((double)akt + ktf[nf] * swt);
        aro = (float)((double)aro + rof[nf] * swt);
        ws += swt;
      }
    }
  }
  const float c = (float)((double)acp / ws), k = (float)((double)akt / ws), r = (float)((double)aro / ws);
  cpc[n] = c;
  ktc[n] = k;
  roc[n] = r;
  const float rc = r * c;
  rcpc[n] = rc < 2e+6f ? 2e+6f : rc;
}

// block reduction of up to 4 values (sum, sum, min, max) over 256 threads
__device__ __forceinline__ void heat_block_red(double v0, double v1, double v2, double v3, double *out) {
  __shared__ double s[4][256];
  const int t = threadIdx.x;
  s[0][t] = v0; s[1][t] = v1; s[2][t] = v2; s[3][t] = v3;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      s[0][t] += s[0][t + w];
      s[1][t] += s[1][t + w];
      s[2][t] = fmin(s[2][t], s[2][t + w]);
      s[3][t] = fmax(s[3][t], s[3][t + w]);
    }
    __syncthreads();
  }
  if (t == 0) {
    out[4 * blockIdx.x + 0] = s[0][0];
    out[4 * blockIdx.x + 1] = s[1][0];
    out[4 * blockIdx.x + 2] = s[2][0];
    out[4 * blockIdx.x + 3] = s[3][0];
  }
}

// largest heating rate |tk1 / rho cp| over the heat-equation nodes (heaterr
// ynerr = 0 with the source term alone, heat3mg.c:54-62, 183-187; the
// titerate pass 331-365); per block in out[4 b + 3]
__global__ void __launch_bounds__(256) k_heat_rate(long n, const unsigned char *__restrict__ isbc,
                                                   const float *__restrict__ tk1, const float *__restrict__ rcp,
                                                   double *out) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  double v = 0.0;
  if (i < n && !isbc[i]) v = fabs(tk1[i] / (double)rcp[i]);
  heat_block_red(0.0, 0.0, 0.0, v, out);
}

// start of a sub-step (heat3mg.c:408-416, 429-451): tk0 = tk, sol0 = tk, and
// the level-0 right-hand side of the heat rows, tk1 + tk0 / dt * rho cp
// (heat3mg.c:204); boundary rows take their constant from k_heat_bcc
__global__ void k_heat_rhs(long n, const unsigned char *__restrict__ isbc, const float *__restrict__ tk,
                           float *tk0, const float *__restrict__ tk1, const float *__restrict__ rcp, double dt,
                           double *x, double *b) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float t = tk[i];
  tk0[i] = t;
  x[i] = (double)t;
  b[i] = isbc[i] ? 0.0 : (double)tk1[i] + (double)t / dt * (double)rcp[i];
}

__global__ void k_heat_bcc(int nbc, const int *__restrict__ r, const double *__restrict__ c, double *b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nbc) b[r[i]] = c[i];
}

// one colour of a Gauss-Seidel sweep over the heat rows on lines
// m1lo <= m1 < m1lo + nx (gausmatt's update, heat3mg.c:1083-1099: the row's
// products summed from -fre1 in slot order, sol0 -= sum / diagonal *
// t0koef).  Thread t: line t / hy of the range, k = t mod hy the k-th node of
// the line's run of this colour.
__global__ void __launch_bounds__(256) k_heat_gs(HeatGeo g, int col, int m1lo, int nx, HDiv fhy, HDiv fnx,
                                                 const unsigned char *__restrict__ isbc,
                                                 const float *__restrict__ kt, const float *__restrict__ rcp,
                                                 double dt, double t0, const double *__restrict__ b,
                                                 double *x) {
  const unsigned t = (unsigned)xcd_block() * blockDim.x + threadIdx.x;
  if (t >= (unsigned)g.hy * nx * g.Z) return;
  const unsigned L = hdiv(t, fhy);
  const int k = (int)(t - L * g.hy);
  const int m3 = (int)hdiv(L, fnx);
  const int m1 = m1lo + (int)L - m3 * nx;
  const int p = (m1 + m3 + col) & 1;
  if (p && k >= (g.Y >> 1)) return;
  const int m2 = 2 * k + p;
  const long n = ((long)m3 * g.X + m1) * g.Y + (p ? g.hy : 0) + k;
  if (isbc[n]) return;
  HeatRow h;
  heat_row(g, kt, rcp[n], dt, m1, m2, m3, n, h);
  const double xn = x[n];
  double iv = -b[n];
  iv += h.d * xn;
#pragma unroll
  for (int s = 0; s < 6; s++) iv += (-h.c[s]) * x[h.nb[s]];
  x[n] = xn - iv / h.d * t0;
}

// boundary rows lo <= i < hi of one dependency depth: x_r = c + koef x_j
// through the same update (tbonderr rows, heat3mg.c:869-893)
__global__ void k_heat_bcpass(int lo, int hi, const int *__restrict__ r, const int *__restrict__ j,
                              const double *__restrict__ negk, const double *__restrict__ b, double *x, double t0) {
  const int i = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= hi) return;
  const int rr = r[i], jj = j[i];
  const double xr = x[rr];
  double iv = -b[rr];
  iv += 1.0 * xr;
  if (jj >= 0) iv += negk[i] * x[jj];
  x[rr] = xr - iv / 1.0 * t0;
}

// All depths of a level's boundary rows in one workgroup (levels with few
// of them): depth after depth, a barrier between (it orders the global
// writes of one depth before the reads of the next within the workgroup).
constexpr int kHeatBcDepths = 8;
constexpr int kHeatBcSmall = 2048;
struct HeatBcDepths {
  int nd;
  int off[kHeatBcDepths + 1];
};
__global__ void __launch_bounds__(1024) k_heat_bc_small(HeatBcDepths D, const int *__restrict__ r,
                                                        const int *__restrict__ j, const double *__restrict__ negk,
                                                        const double *__restrict__ b, double *x, double t0) {
  for (int d = 0; d < D.nd; d++) {
    for (int i = D.off[d] + threadIdx.x; i < D.off[d + 1]; i += blockDim.x) {
      const int rr = r[i], jj = j[i];
      const double xr = x[rr];
      double iv = -b[rr];
      iv += 1.0 * xr;
      if (jj >= 0) iv += negk[i] * x[jj];
      x[rr] = xr - iv / 1.0 * t0;
    }
    __syncthreads();
  }
}

// residual fre0 = fre1 - A sol0 of the heat rows (heat3mg.c:1101-1117)
__global__ void __launch_bounds__(256) k_heat_res(HeatGeo g, const unsigned char *__restrict__ isbc,
                                                  const float *__restrict__ kt, const float *__restrict__ rcp,
                                                  double dt, const double *__restrict__ b,
                                                  const double *__restrict__ x, double *r) {
  const long n = xcd_block() * blockDim.x + threadIdx.x;
  if (n >= (long)g.XY * g.Z) return;
  if (isbc[n]) return;
  int m1, m2, m3;
  hdecode(g, n, m1, m2, m3);
  HeatRow h;
  heat_row(g, kt, rcp[n], dt, m1, m2, m3, n, h);
  double iv = b[n];
  iv -= h.d * x[n];
#pragma unroll
  for (int s = 0; s < 6; s++) iv -= (-h.c[s]) * x[h.nb[s]];
  r[n] = iv;
}

__global__ void k_heat_bcres(int nbc, const int *__restrict__ r, const int *__restrict__ j,
                             const double *__restrict__ negk, const double *__restrict__ b,
                             const double *__restrict__ x, double *res) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nbc) return;
  const int rr = r[i], jj = j[i];
  double iv = b[rr];
  iv -= 1.0 * x[rr];
  if (jj >= 0) iv -= negk[i] * x[jj];
  res[rr] = iv;
}

// trestrict (heat3mg.c:1122-1232): a coarse node without a boundary row
// gathers the fine residuals of the 8 cells around it in scan order, weights
// (1 -+ d) per axis multiplied x * y * z, and divides by the weight sum; the
// coarse boundary rows get 0, the coarse solution 0
__global__ void __launch_bounds__(256) k_heat_restrict(int Xf, int Yf, int Xc, int Yc, int Zc, CompTab t,
                                                       const unsigned char *__restrict__ isbc_c,
                                                       const double *__restrict__ rf, double *bc, double *xc) {
  const long q = xcd_block() * blockDim.x + threadIdx.x;
  const int XYc = Xc * Yc, XYf = Xf * Yf, hyf = (Yf + 1) / 2;
  if (q >= (long)XYc * Zc) return;
  int cx, cy, cz;
  decode(q, Xc, Yc, XYc, cx, cy, cz);
  const long n = hidx(Xc, Yc, (Yc + 1) / 2, cx, cy, cz);
  xc[n] = 0.0;
  if (isbc_c[n]) { bc[n] = 0.0; return; }
  double acc = 0.0, ws = 0.0;
  for (int a = t.ax[2].cbeg[cz]; a < t.ax[2].cbeg[cz + 1]; a++) {
    const long oz = (long)t.ax[2].cfi[a] * XYf;
    const double wz = t.ax[2].cw64[a];
    for (int bq = t.ax[0].cbeg[cx]; bq < t.ax[0].cbeg[cx + 1]; bq++) {
      const long ox = oz + (long)t.ax[0].cfi[bq] * Yf;
      const double wx = t.ax[0].cw64[bq];
      for (int e = t.ax[1].cbeg[cy]; e < t.ax[1].cbeg[cy + 1]; e++) {
        const double swt = wx * t.ax[1].cw64[e] * wz;
        const int j = t.ax[1].cfi[e];
        acc += rf[ox + ((j & 1) ? hyf : 0) + (j >> 1)] * swt;
        ws += swt;
      }
    }
  }
  bc[n] = ws != 0.0 ? acc / ws : acc;
}

// tprolong (heat3mg.c:954-1050): a fine node without a boundary row adds
// sol0_c * w * t1koef of the 8 coarse nodes of its cell, (z, x, y) order
__global__ void __launch_bounds__(256) k_heat_prolong(int Xf, int Yf, int Zf, int Xc, int Yc, CompTab t,
                                                      const unsigned char *__restrict__ isbc_f,
                                                      const double *__restrict__ xc, double *xf, double t1,
                                                      int ynreset) {
  const long q = xcd_block() * blockDim.x + threadIdx.x;
  const int XYf = Xf * Yf, XYc = Xc * Yc, hyc = (Yc + 1) / 2;
  if (q >= (long)XYf * Zf) return;
  int m1, m2, m3;
  decode(q, Xf, Yf, XYf, m1, m2, m3);
  const long n = hidx(Xf, Yf, (Yf + 1) / 2, m1, m2, m3);
  if (isbc_f[n]) return;
  const int lx = t.ax[0].lo[m1], ly = t.ax[1].lo[m2], lz = t.ax[2].lo[m3];
  double v = ynreset ? 0.0 : xf[n];
  for (int dz = 0; dz < 2; dz++) {
    const double wz = t.ax[2].w64[2 * m3 + dz];
    for (int dx = 0; dx < 2; dx++) {
      const double wx = t.ax[0].w64[2 * m1 + dx];
      for (int dy = 0; dy < 2; dy++) {
        const double swt = wx * t.ax[1].w64[2 * m2 + dy] * wz;
        const int j = ly + dy;
        v += xc[(long)(lz + dz) * XYc + (long)(lx + dx) * Yc + ((j & 1) ? hyc : 0) + (j >> 1)] * swt * t1;
      }
    }
  }
  xf[n] = v;
}

// tk = (float) sol0 (heat3mg.c:507-512)
__global__ void k_heat_tk(long n, const double *__restrict__ x, float *tk) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) tk[i] = (float)x[i];
}

// heat-equation error of the FP32 temperature (heaterr ynerr = 2,
// heat3mg.c:251-287, times the sub-step, 540-548) squared and summed per
// block, with the node count and the tk min / max: out[4 b + 0..3]
__global__ void __launch_bounds__(256) k_heat_err(HeatGeo g, const unsigned char *__restrict__ isbc,
                                                  const float *__restrict__ kt, const float *__restrict__ rcp,
                                                  const float *__restrict__ tk, const float *__restrict__ tk0,
                                                  const float *__restrict__ tk1, double dt, double *out) {
  const long n = xcd_block() * blockDim.x + threadIdx.x;
  double sq = 0.0, cnt = 0.0, mn = 1e30, mx = -1e30;
  if (n < (long)g.XY * g.Z) {
    const float t3 = tk[n];
    mn = mx = (double)t3;
    if (!isbc[n]) {
      int m1, m2, m3;
      hdecode(g, n, m1, m2, m3);
      const double rc = (double)rcp[n];
      const long k = m2 >> 1, base = n - ((m2 & 1) ? g.hy : 0) - k;
      const long nb[6] = {(g.xper && m1 == 0) ? n + (long)(g.X - 2) * g.Y : n - g.Y, n + g.Y,
                          (m2 & 1) ? base + k : base + g.hy + k - 1, (m2 & 1) ? base + k + 1 : base + g.hy + k,
                          n - g.XY, n + g.XY};
      const int mm[3] = {m1, m2, m3};
      const float k0 = kt[n];
      double e = (double)(t3 - tk0[n]) / dt;
#pragma unroll
      for (int a = 0; a < 3; a++) {
        const int m = mm[a];
        const double kf = g.kf[a][m];
        const float sm = kt[nb[2 * a]] + k0, sp = k0 + kt[nb[2 * a + 1]];
        const double cm = (double)sm / 2.0 * kf / g.dm[a][m] / rc;
        e += (double)(t3 - tk[nb[2 * a]]) * cm;
        const double cp = (double)sp / 2.0 * kf / g.dp[a][m] / rc;
        e -= (double)(tk[nb[2 * a + 1]] - t3) * cp;
      }
      e -= (double)tk1[n] / rc;
      e *= dt;
      sq = e * e;
      cnt = 1.0;
    }
  }
  heat_block_red(sq, cnt, mn, mx, out);
}

// boundary-row errors on the FP64 solution (tbonderr ynerr = 1,
// heat3mg.c:854-867), squared and summed per block
__global__ void __launch_bounds__(256) k_heat_bonderr(int nbc, const int *__restrict__ r, const int *__restrict__ j,
                                                      const double *__restrict__ c, const double *__restrict__ kf,
                                                      const double *__restrict__ x, double *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double sq = 0.0, cnt = 0.0;
  if (i < nbc) {
    double lb = x[r[i]] - c[i];
    if (j[i] >= 0) lb -= kf[i] * x[j[i]];
    sq = lb * lb;
    cnt = 1.0;
  }
  heat_block_red(sq, cnt, 0.0, 0.0, out);
}

// The right part of heaterr (ynerr = 0, heat3mg.c:54-187) on the level-0
// heat rows: tk1 = ht, plus the adiabatic term (adiabyn 2: et tk0 v.grad P,
// the pressure gradient of the 8 cells around the node at the node's
// relative position, :100-115; adiabyn 1: et tk0 ro v.g, :131-134) and the
// frictional heating (frictyn: the cells' normal stress-strain products and
// twice the edges' shear ones, :140-175), both where ro > 2e3.  v at the node
// as vxyzcalc1 interpolates it (adv_interp on the FP32 velocity planes), the
// pressure from the FP64 solution, stresses and strain rates rounded to FP32
// as the reference's arrays hold them; FP32 products and sums where the
// reference multiplies float arrays.  Natural node order for the Stokes
// fields, the split-line layout for the temperature arrays.
__global__ void __launch_bounds__(256) k_heat_sources(HeatGeo hg, AdvGeo ag, const unsigned char *__restrict__ isbc,
                                                      const double *__restrict__ x, const float *__restrict__ v3,
                                                      const double *__restrict__ sig,
                                                      const double *__restrict__ eps, const float *__restrict__ ro,
                                                      const float *__restrict__ tk, const float *__restrict__ ht,
                                                      const float *__restrict__ et, float *tk1, int adiabyn,
                                                      int frictyn, double gxk, double gyk, double gzk) {
  const long n = xcd_block() * blockDim.x + threadIdx.x;
  const long n0 = (long)hg.XY * hg.Z;
  if (n >= n0) return;
  int m1, m2, m3;
  decode(n, hg.X, hg.Y, hg.XY, m1, m2, m3);
  const long hn = hidx(hg.X, hg.Y, hg.hy, m1, m2, m3);
  if (isbc[hn]) return;
  double ival = 0.0;
  ival += ht[hn];
  const double xkf = hg.kf[0][m1], ykf = hg.kf[1][m2], zkf = hg.kf[2][m3];
  const long Y = hg.Y, XY = hg.XY;
  const long v0 = n - XY, v1 = (hg.xper && m1 == 0) ? n + (long)(hg.X - 2) * Y : n - Y, v2 = n - 1;
  long p[8];
  double rx = 0.0, ry = 0.0, rz = 0.0;
  if (adiabyn == 2 || frictyn) {
    p[0] = n; p[1] = n + 1; p[2] = n + Y; p[3] = p[2] + 1;
    p[4] = n + XY; p[5] = p[4] + 1; p[6] = p[4] + Y; p[7] = p[6] + 1;
    rx = hg.dm[0][m1] / 2.0 * xkf;
    ry = hg.dm[1][m2] / 2.0 * ykf;
    rz = hg.dm[2][m3] / 2.0 * zkf;
  }
  double vv[3] = {0.0, 0.0, 0.0};
  if (adiabyn) adv_interp(ag, v3, nullptr, ag.gx[m1], ag.gy[m2], ag.gz[m3], vv, nullptr);
  const float r = ro[hn], t0 = tk[hn];
  if (adiabyn == 2 && r > 2e+3) {
    double P[8];
#pragma unroll
    for (int k = 0; k < 8; k++) P[k] = x[4 * p[k]];
    const double dpdx = xkf * ((1.0 - ry) * (1.0 - rz) * (P[2] - P[0]) + ry * (1.0 - rz) * (P[3] - P[1]) +
                               (1.0 - ry) * rz * (P[6] - P[4]) + ry * rz * (P[7] - P[5]));
    const double dpdy = ykf * ((1.0 - rx) * (1.0 - rz) * (P[1] - P[0]) + rx * (1.0 - rz) * (P[3] - P[2]) +
                               (1.0 - rx) * rz * (P[5] - P[4]) + rx * rz * (P[7] - P[6]));
    const double dpdz = zkf * ((1.0 - rx) * (1.0 - ry) * (P[4] - P[0]) + rx * (1.0 - ry) * (P[6] - P[2]) +
                               (1.0 - rx) * ry * (P[5] - P[1]) + rx * ry * (P[7] - P[3]));
    const float ek = et[hn] * t0;
    ival += (double)ek * (vv[0] * dpdx + vv[1] * dpdy + vv[2] * dpdz);
  }
  if (adiabyn == 1 && r > 2e+3) {
    const float ekr = et[hn] * t0 * r;
    ival += (double)ekr * (vv[0] * gxk + vv[1] * gyk + vv[2] * gzk);
  }
  if (frictyn && r > 2e+3) {
    float q[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const float a = (float)sig[p[k]] * (float)eps[p[k]];
      const float b = (float)sig[n0 + p[k]] * (float)eps[n0 + p[k]];
      const float c = (float)sig[2 * n0 + p[k]] * (float)eps[2 * n0 + p[k]];
      q[k] = a + b + c;
    }
    double e;
    e = (1.0 - rx) * (1.0 - ry) * (1.0 - rz) * (double)q[0];
    e += (1.0 - rx) * (ry) * (1.0 - rz) * (double)q[1];
    e += (rx) * (1.0 - ry) * (1.0 - rz) * (double)q[2];
    e += (rx) * (ry) * (1.0 - rz) * (double)q[3];
    e += (1.0 - rx) * (1.0 - ry) * (rz) * (double)q[4];
    e += (1.0 - rx) * (ry) * (rz) * (double)q[5];
    e += (rx) * (1.0 - ry) * (rz) * (double)q[6];
    e += (rx) * (ry) * (rz) * (double)q[7];
    const float sxy0 = (float)sig[3 * n0 + v0] * (float)eps[3 * n0 + v0], sxy3 = (float)sig[3 * n0 + n] * (float)eps[3 * n0 + n];
    const float sxz2 = (float)sig[4 * n0 + v2] * (float)eps[4 * n0 + v2], sxz3 = (float)sig[4 * n0 + n] * (float)eps[4 * n0 + n];
    const float syz1 = (float)sig[5 * n0 + v1] * (float)eps[5 * n0 + v1], syz3 = (float)sig[5 * n0 + n] * (float)eps[5 * n0 + n];
    e += 2.0 * ((1.0 - rz) * (double)sxy0 + rz * (double)sxy3);
    e += 2.0 * ((1.0 - ry) * (double)sxz2 + ry * (double)sxz3);
    e += 2.0 * ((1.0 - rx) * (double)syz1 + rx * (double)syz3);
    ival += e;
  }
  tk1[hn] = (float)ival;
}

// ---- marker temperatures after the step (heat3mg.c:618-950) --------------
// tkcalc1 (mark3mg.c:1212-1278): the marker's cell by the cell search,
// trilinear node weights 1 - |d| per axis multiplied x * y * z, z outer,
// x, y inner; tk (new) and tk2 (the step's start) interpolated
struct HeatMark {
  int c1, c2, c3;
  double dx, dy, dz;
  double tn, to;
};
__device__ __forceinline__ double tk_w(const double *g, int m, double p, double d) {
  double t = (g[m] - p) / d;
  return 1.0 - fabs(t);
}
__device__ __forceinline__ HeatMark heat_tkcalc(const HeatGeo &hg, const AdvGeo &ag, const float *__restrict__ tk,
                                               const float *__restrict__ tk2, double x, double y, double z) {
  HeatMark r;
  r.c1 = adv_serch(ag.gx, ag.X, ag.bx, x);
  r.c2 = adv_serch(ag.gy, ag.Y, ag.by, y);
  r.c3 = adv_serch(ag.gz, ag.Z, ag.bz, z);
  r.dx = ag.gx[r.c1 + 1] - ag.gx[r.c1];
  r.dy = ag.gy[r.c2 + 1] - ag.gy[r.c2];
  r.dz = ag.gz[r.c3 + 1] - ag.gz[r.c3];
  double a = 0.0, b = 0.0;
  for (int m3 = r.c3; m3 <= r.c3 + 1; m3++) {
    const double ddz = tk_w(ag.gz, m3, z, r.dz);
    for (int m1 = r.c1; m1 <= r.c1 + 1; m1++) {
      const double ddx = tk_w(ag.gx, m1, x, r.dx);
      for (int m2 = r.c2; m2 <= r.c2 + 1; m2++) {
        const double ddy = tk_w(ag.gy, m2, y, r.dy);
        const double swt = ddx * ddy * ddz;
        const long h = hidx(hg.X, hg.Y, hg.hy, m1, m2, m3);
        a += swt * tk[h];
        b += swt * tk2[h];
      }
    }
  }
  r.tn = a;
  r.to = b;
  return r;
}

__device__ __forceinline__ bool mark_inside(const AdvGeo &ag, float x, float y, float z) {
  return x > 0 && y > 0 && z > 0 && (double)x < ag.gx[ag.X - 1] && (double)y < ag.gy[ag.Y - 1] &&
         (double)z < ag.gz[ag.Z - 1];
}

// per marker: fluid markers (50 <= type < 100) inside take the new
// temperature; normal ones (type < 50) strictly inside: markk <= 0 takes the
// old one (tdeep where y > zdeep and it is colder), the numerical diffusion
// (heatdif: the boundary-cell cap at the 8 nodes' largest tk, dtk1 =
// (Told - markk)(1 - exp(-heatdif kt/(rho cp) 2 (1/dx^2+1/dy^2+1/dz^2) dt)),
// the exponent clamped at -150, added to the marker; dtk1 kept for the
// nodes' gather), then the nodal change Tnew - Told
__global__ void __launch_bounds__(256) k_heat_mark_a(HeatGeo hg, AdvGeo ag, long nm, const float *__restrict__ mx,
                                                     const float *__restrict__ my, const float *__restrict__ mz,
                                                     const int *__restrict__ markt, float *markk,
                                                     const float *__restrict__ tk, const float *__restrict__ tk2,
                                                     const float *__restrict__ kt, const float *__restrict__ cp,
                                                     const float *__restrict__ ro, double heatdif, double dt,
                                                     double zdeep, double tdeep, unsigned char *flag, double *dtk1o) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= nm) return;
  flag[i] = 0;
  const int t = markt ? markt[i] : 0;
  const float fx = mx[i], fy = my[i], fz = mz[i];
  if (!mark_inside(ag, fx, fy, fz) || t >= 100) return;
  const HeatMark r = heat_tkcalc(hg, ag, tk, tk2, fx, fy, fz);
  if (t >= 50) {
    markk[i] = (float)r.tn;
    return;
  }
  float mk = markk[i];
  if (mk <= 0) {
    mk = (float)r.to;
    if (fy > zdeep && mk < tdeep) mk = (float)tdeep;
  }
  if (heatdif != 0.0) {
    double mkt = 0.0, mcp = 0.0, mro = 0.0;
    double mtk = tk[hidx(hg.X, hg.Y, hg.hy, r.c1, r.c2, r.c3)];
    for (int m3 = r.c3; m3 <= r.c3 + 1; m3++) {
      const double ddz = tk_w(ag.gz, m3, fz, r.dz);
      for (int m1 = r.c1; m1 <= r.c1 + 1; m1++) {
        const double ddx = tk_w(ag.gx, m1, fx, r.dx);
        for (int m2 = r.c2; m2 <= r.c2 + 1; m2++) {
          const double w = ddx * tk_w(ag.gy, m2, fy, r.dy) * ddz;
          const long h = hidx(hg.X, hg.Y, hg.hy, m1, m2, m3);
          mkt += kt[h] * w;
          mcp += cp[h] * w;
          mro += ro[h] * w;
          mtk = fmax(mtk, (double)tk[h]);
        }
      }
    }
    if ((r.c1 <= 0 || r.c1 >= ag.X - 2 || r.c2 <= 0 || r.c2 >= ag.Y - 2 || r.c3 <= 0 || r.c3 >= ag.Z - 2) &&
        ((double)mk + r.tn - r.to) > mtk)
      mk = (float)(mtk - r.tn + r.to);
    const double dtk = r.to - (double)mk;
    double rocpkf = mro * mcp;
    if (rocpkf < 2e+6) rocpkf = 2e+6;
    double d1 = -heatdif * mkt / rocpkf * 2.0 * (1.0 / r.dx / r.dx + 1.0 / r.dy / r.dy + 1.0 / r.dz / r.dz) * dt;
    if (d1 < -150.0) d1 = -150.0;
    d1 = dtk * (1.0 - exp(d1));
    mk = mk + (float)d1;
    dtk1o[i] = d1;
    flag[i] = 1;
  }
  mk = mk + (float)(r.tn - r.to);
  markk[i] = mk;
}

// per node: Sol0 = sum w, Sol1 = sum dtk1 w over the diffusing markers of
// the 8 cells around it (cells in (z, x, y) order, markers of a cell in the
// sorted order -- the input order for markers stored cell by cell, the
// reference's single-thread accumulation order), then Sol1 / Sol0 (natural
// node order)
__global__ void __launch_bounds__(256) k_heat_mark_gather(AdvGeo ag, const int *__restrict__ start,
                                                          const int *__restrict__ sidx,
                                                          const float *__restrict__ mx, const float *__restrict__ my,
                                                          const float *__restrict__ mz,
                                                          const unsigned char *__restrict__ flag,
                                                          const double *__restrict__ dtk1, double *sol1) {
  const long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= ag.n0) return;
  int m1, m2, m3;
  decode(n, ag.X, ag.Y, ag.XY, m1, m2, m3);
  double s0 = 0.0, s1 = 0.0;
  for (int c3 = m3 - 1; c3 <= m3; c3++) {
    if (c3 < 0 || c3 > ag.Z - 2) continue;
    for (int c1 = m1 - 1; c1 <= m1; c1++) {
      if (c1 < 0 || c1 > ag.X - 2) continue;
      for (int c2 = m2 - 1; c2 <= m2; c2++) {
        if (c2 < 0 || c2 > ag.Y - 2) continue;
        const int cell = (c3 * (ag.X - 1) + c1) * (ag.Y - 1) + c2;
        for (int k = start[cell]; k < start[cell + 1]; k++) {
          const int i = sidx[k];
          if (!flag[i]) continue;
          const double ddz = tk_w(ag.gz, m3, mz[i], ag.gz[c3 + 1] - ag.gz[c3]);
          const double ddx = tk_w(ag.gx, m1, mx[i], ag.gx[c1 + 1] - ag.gx[c1]);
          const double ddy = tk_w(ag.gy, m2, my[i], ag.gy[c2 + 1] - ag.gy[c2]);
          const double w = ddx * ddy * ddz;
          s1 += dtk1[i] * w;
          s0 += w;
        }
      }
    }
  }
  sol1[n] = s0 != 0.0 ? s1 / s0 : s1;
}

// antidiffusion: a normal marker inside loses sum w Sol1 over its nodes with
// Sol1 != 0 (heat3mg.c:795-830)
__global__ void __launch_bounds__(256) k_heat_mark_anti(AdvGeo ag, long nm, const float *__restrict__ mx,
                                                        const float *__restrict__ my, const float *__restrict__ mz,
                                                        const int *__restrict__ markt, float *markk,
                                                        const double *__restrict__ sol1) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= nm) return;
  const int t = markt ? markt[i] : 0;
  const float fx = mx[i], fy = my[i], fz = mz[i];
  if (t >= 50 || !mark_inside(ag, fx, fy, fz)) return;
  const int c1 = adv_serch(ag.gx, ag.X, ag.bx, fx), c2 = adv_serch(ag.gy, ag.Y, ag.by, fy),
            c3 = adv_serch(ag.gz, ag.Z, ag.bz, fz);
  const double dx = ag.gx[c1 + 1] - ag.gx[c1], dy = ag.gy[c2 + 1] - ag.gy[c2], dz = ag.gz[c3 + 1] - ag.gz[c3];
  float mk = markk[i];
  for (int m3 = c3; m3 <= c3 + 1; m3++) {
    const double ddz = tk_w(ag.gz, m3, fz, dz);
    for (int m1 = c1; m1 <= c1 + 1; m1++) {
      const double ddx = tk_w(ag.gx, m1, fx, dx);
      for (int m2 = c2; m2 <= c2 + 1; m2++) {
        const double w = ddx * tk_w(ag.gy, m2, fy, dy) * ddz;
        const double s = sol1[(long)m3 * ag.XY + (long)m1 * ag.Y + m2];
        if (s != 0.0) mk = mk - (float)(s * w);
      }
    }
  }
  markk[i] = mk;
}

}  // namespace stk
