// stk_advect.h -- marker advection on the device (gfx950): one step of
// movemark (parallel/mark3mg.c:174-629) per marker, with the velocity and
// strain-rate interpolations vxyzcalc1 (mark3mg.c:2963-3236) and epscalc1
// (mark3mg.c:3244-3760), walled and periodic boxes.
//
// One thread per marker.  The staggered quantities are read from FP32 node
// planes, as the reference's arrays hold them (k_adv_planes rounds the FP64
// solution and strain rates once per step), and every interpolation weight
// and sum follows the reference's operation order, so a marker lands where
// the oracle's restatement puts it bit for bit.
#pragma once
#include <hip/hip_runtime.h>

namespace stk {

// bins of the cell search: an axis [g[0], g[n-1]] cut into nb equal bins,
// bin b holding the largest line index a <= n-2 with g[a] <= the bin's start
struct AdvBins {
  const int *t;
  int nb;
  double g0, inv;  // first line, nb / (g[n-1] - g[0])
};

struct AdvGeo {
  int X, Y, Z, XY;
  const double *gx, *gy, *gz;  // grid lines of level 0
  double xs, ys, zs;           // box sizes (last line minus first)
  long n0;
  AdvBins bx, by, bz;
  int px, py, pz;              // periodic axes
  int *wrap_out;               // set when a wrapped point stays outside (the reference exits there)
};

// bins per axis: 8 per cell, so a search takes a step or two on grids whose
// spacing varies by less than 8x
constexpr int kAdvBinsPerCell = 8;

// the rules by which an axis brackets a point (mark3mg.c:3004-3023,
// 3272-3285, 3390-3398, 3411-3427): a velocity component along its own
// axis (lines, index clamped), across an axis (centres, point clamped to the
// outer centres), normal strain rates (cell centres 1..n-1), shear strain
// rates along a line axis (lines 1..n-2, point clamped to the line) and
// across (centres; the lower bracket clamped from 1 down to 0).  On a
// periodic axis (mark3mg.c:2992-3001, 3282-3291, 3385-3389, 3410-3416):
// across-axis velocities pair centres n-2 and 0 over the wrap below centre 0
// and above centre n-3; normal rates pair cells n-1 and 1 below cell 1 and
// above cell n-2; shear rates along the axis take lines 0..n-2; shear rates
// across it wrap only below centre 0 (the reference's test above is > n-2,
// never true), the rest as walled.  Over the wrap the second node's
// distance is taken from the point minus the box size.
enum { kAxVN = 0, kAxVC = 1, kAxCell = 2, kAxSN = 3, kAxSC = 4 };

struct AxW {
  int i0, i1;
  double w0, w1;
};

// cell search of m1serch/m2serch/m3serch (mark3mg.c:1138-1205).  Their
// bisection over strictly increasing lines returns the largest a <= n-2 with
// g[a] <= p, 0 below g[0] and n-2 for a NaN; here the bin of p gives a start
// and short walks up and down end on that same index whatever the start.
__device__ __forceinline__ int adv_serch(const double *__restrict__ g, int n, const AdvBins &B, double p) {
  if (p != p) return n - 2;
  const double f = (p - B.g0) * B.inv;
  const int b = f <= 0.0 ? 0 : (f >= (double)(B.nb - 1) ? B.nb - 1 : (int)f);
  int a = B.t[b];
  while (a < n - 2 && g[a + 1] <= p) a++;
  while (a > 0 && g[a] > p) a--;
  return a;
}

// per: periodic axis; size: its box size; dvar: the spacing over an x wrap
// of the vz interpolation, (g[1] + g[n-1] - g[n-2]) / 2 (mark3mg.c:3192),
// where every other wrap takes (g[1] - g[0] + g[n-1] - g[n-2]) / 2; ylow:
// the low clamp of a periodic line axis at g[0] (the y axis, :3399) rather
// than at 0 (x and z, :3387, :3503)
template <int K>
__device__ __forceinline__ AxW adv_axis(const double *__restrict__ g, int n, int m, double p, int per = 0,
                                        double size = 0.0, int dvar = 0, int ylow = 0) {
  double P = p, d, t0, t1;
  int i0 = m, i1;
  bool wrap = false;
  if (K == kAxVN) {
    if (i0 < 0) i0 = 0;
    if (i0 > n - 2) i0 = n - 2;
    i1 = i0 + 1;
    d = g[i0 + 1] - g[i0];
    t0 = (g[i0] - p) / d;
    t1 = (g[i1] - p) / d;
  } else if (K == kAxSN) {
    if (per) {
      if (i0 < 0) { i0 = 0; P = ylow ? g[0] : 0.0; }
      if (i0 > n - 2) { i0 = n - 2; P = g[i0 + 1]; }
    } else {
      if (i0 < 1) { i0 = 1; P = g[i0]; }
      if (i0 > n - 3) { i0 = n - 3; P = g[i0 + 1]; }
    }
    i1 = i0 + 1;
    d = g[i0 + 1] - g[i0];
    t0 = (g[i0] - P) / d;
    t1 = (g[i1] - P) / d;
  } else if (K == kAxCell) {
    if (P > (g[i0] + g[i0 + 1]) / 2.0) i0 += 1;
    if (per && (i0 < 1 || i0 > n - 2)) {
      if (i0 < 1) P = size + p;
      i0 = n - 1;
      i1 = 1;
      wrap = true;
      d = (g[1] - g[0] + g[n - 1] - g[n - 2]) / 2.0;
    } else {
      if (i0 < 1) { i0 = 1; P = (g[i0] + g[i0 - 1]) / 2.0; }
      if (i0 > n - 2) { i0 = n - 2; P = (g[i0] + g[i0 + 1]) / 2.0; }
      i1 = i0 + 1;
      d = (g[i0 + 1] - g[i0 - 1]) / 2.0;
    }
    t0 = ((g[i0] + g[i0 - 1]) / 2.0 - P) / d;
    t1 = ((g[i1] + g[i1 - 1]) / 2.0 - (wrap ? P - size : P)) / d;
  } else {
    if (P < (g[i0] + g[i0 + 1]) / 2.0) i0 -= 1;
    if (per && (i0 < 0 || i0 > (K == kAxVC ? n - 3 : n - 2))) {
      if (i0 < 0) P = size + p;
      i0 = n - 2;
      i1 = 0;
      wrap = true;
      d = dvar ? (g[1] + g[n - 1] - g[n - 2]) / 2.0 : (g[1] - g[0] + g[n - 1] - g[n - 2]) / 2.0;
    } else {
      if (K == kAxVC ? i0 < 0 : i0 < 1) { i0 = 0; P = (g[i0] + g[i0 + 1]) / 2.0; }
      if (i0 > n - 3) { i0 = n - 3; P = (g[i0 + 1] + g[i0 + 2]) / 2.0; }
      i1 = i0 + 1;
      d = (g[i0 + 2] - g[i0]) / 2.0;
    }
    t0 = ((g[i0] + g[i0 + 1]) / 2.0 - P) / d;
    t1 = ((g[i1] + g[i1 + 1]) / 2.0 - (wrap ? P - size : P)) / d;
  }
  AxW r;
  r.i0 = i0;
  r.i1 = i1;
  r.w0 = 1.0 - (-t0);
  r.w1 = 1.0 - t1;
  return r;
}

// vx, vy, vz planes (3 n0) from the solution and, with eps, the six
// strain-rate planes (6 n0), rounded to FP32
__global__ void k_adv_planes(long n0, const double *__restrict__ x, const double *__restrict__ eps, float *v3,
                             float *e6) {
  const long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= n0) return;
  const double4 q = reinterpret_cast<const double4 *>(x)[n];
  v3[n] = (float)q.y;
  v3[n0 + n] = (float)q.z;
  v3[2 * n0 + n] = (float)q.w;
  if (eps)
    for (int k = 0; k < 6; k++) e6[k * n0 + n] = (float)eps[k * n0 + n];
}

// sum over the 8 bracketing nodes, z outer, x, y inner, of the node value
// times (wx * wy) * wz
__device__ __forceinline__ double adv_sum8(const AdvGeo &g, AxW ax, AxW ay, AxW az, const float *__restrict__ v) {
  double acc = 0.0;
#pragma unroll
  for (int kz = 0; kz < 2; kz++) {
    const double wz = kz ? az.w1 : az.w0;
#pragma unroll
    for (int kx = 0; kx < 2; kx++) {
      const double wx = kx ? ax.w1 : ax.w0;
#pragma unroll
      for (int ky = 0; ky < 2; ky++) {
        const double wy = ky ? ay.w1 : ay.w0;
        const long n = (long)(kz ? az.i1 : az.i0) * g.XY + (long)(kx ? ax.i1 : ax.i0) * g.Y + (ky ? ay.i1 : ay.i0);
        acc += (double)v[n] * (wx * wy * wz);
      }
    }
  }
  return acc;
}

// velocity (v, if non-null) and strain rates (e, if non-null) at a point;
// PER = false compiles the walled box only (the periodic branches fold away)
template <bool PER = true>
__device__ __forceinline__ void adv_interp(const AdvGeo &g, const float *__restrict__ v3,
                                           const float *__restrict__ e6, double px, double py, double pz, double *v,
                                           double *e) {
  px = px < 0 ? 0 : (px > g.xs ? g.xs : px);
  py = py < 0 ? 0 : (py > g.ys ? g.ys : py);
  pz = pz < 0 ? 0 : (pz > g.zs ? g.zs : pz);
  const int m1 = adv_serch(g.gx, g.X, g.bx, px), m2 = adv_serch(g.gy, g.Y, g.by, py),
            m3 = adv_serch(g.gz, g.Z, g.bz, pz);
  // each bracketing serves two of the three sums that use it (the vz sum
  // takes its own x bracketing on a periodic x axis: another spacing there)
  const int px_ = PER ? g.px : 0, py_ = PER ? g.py : 0, pz_ = PER ? g.pz : 0;
  if (v) {
    const AxW nx = adv_axis<kAxVN>(g.gx, g.X, m1, px), ny = adv_axis<kAxVN>(g.gy, g.Y, m2, py),
              nz = adv_axis<kAxVN>(g.gz, g.Z, m3, pz);
    const AxW cx = adv_axis<kAxVC>(g.gx, g.X, m1, px, px_, g.xs),
              cy = adv_axis<kAxVC>(g.gy, g.Y, m2, py, py_, g.ys),
              cz = adv_axis<kAxVC>(g.gz, g.Z, m3, pz, pz_, g.zs);
    v[0] = adv_sum8(g, nx, cy, cz, v3);
    v[1] = adv_sum8(g, cx, ny, cz, v3 + g.n0);
    const AxW cxz = px_ ? adv_axis<kAxVC>(g.gx, g.X, m1, px, 1, g.xs, 1) : cx;
    v[2] = adv_sum8(g, cxz, cy, nz, v3 + 2 * g.n0);
  }
  if (e) {
    {
      const AxW cx = adv_axis<kAxCell>(g.gx, g.X, m1, px, px_, g.xs),
                cy = adv_axis<kAxCell>(g.gy, g.Y, m2, py, py_, g.ys),
                cz = adv_axis<kAxCell>(g.gz, g.Z, m3, pz, pz_, g.zs);
      for (int k = 0; k < 3; k++) e[k] = adv_sum8(g, cx, cy, cz, e6 + (long)k * g.n0);
    }
    const AxW nx = adv_axis<kAxSN>(g.gx, g.X, m1, px, px_), ny = adv_axis<kAxSN>(g.gy, g.Y, m2, py, py_, 0.0, 0, 1),
              nz = adv_axis<kAxSN>(g.gz, g.Z, m3, pz, pz_);
    const AxW cx = adv_axis<kAxSC>(g.gx, g.X, m1, px, px_, g.xs),
              cy = adv_axis<kAxSC>(g.gy, g.Y, m2, py, py_, g.ys),
              cz = adv_axis<kAxSC>(g.gz, g.Z, m3, pz, pz_, g.zs);
    e[3] = adv_sum8(g, nx, ny, cz, e6 + 3 * g.n0);
    e[4] = adv_sum8(g, nx, cy, nz, e6 + 4 * g.n0);
    e[5] = adv_sum8(g, cx, ny, nz, e6 + 5 * g.n0);
  }
}

// second invariant of the strain rate (mark3mg.c:229)
__device__ __forceinline__ double adv_eii(const double *e) {
  return sqrt(0.5 * (e[0] * e[0] + e[1] * e[1] + e[2] * e[2]) + e[3] * e[3] + e[4] * e[4] + e[5] * e[5]);
}

// one movemark step per marker (mark3mg.c:208-520): the velocity at the
// marker (Euler, markmod 1) or the RK4 combination of four stages, the
// strain-rate invariant at the start (eii) and, for markers with
// accumulated strain, its RK4 average added times dt; outgrid 2 keeps only
// the motion across the faces an outside marker lies beyond; a normal
// marker (type < 100) leaving the box gets markk = 0.  e6 may be null when
// neither eii nor marke is asked for (the strain rates are then not read).
template <bool PER>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_move_markers(AdvGeo g, const float *__restrict__ v3,
                                                      const float *__restrict__ e6, long nmark, float *mx,
                                                      float *my, float *mz, float *marke, float *markk,
                                                      const int *__restrict__ markt, float *eii, double dt,
                                                      int markmod, int outgrid) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= nmark) return;
  const int t = markt ? markt[i] : 0;
  const float fx = mx[i], fy = my[i], fz = mz[i];
  const bool inside = fx >= 0 && fy >= 0 && fz >= 0 && (double)fx <= g.xs && (double)fy <= g.ys && (double)fz <= g.zs;
  if (!(inside || (outgrid != 1 && (t < 50 || t >= 100)))) return;
  const bool wantE = e6 && marke && marke[i] > 0;
  double v1[3], e[6], vv[3];
  adv_interp<PER>(g, v3, e6, fx, fy, fz, v1, e6 ? e : nullptr);
  const double ee1 = e6 ? adv_eii(e) : 0.0;
  if (eii) eii[i] = (float)ee1;
  double ee0;
  bool out = false;
  if (markmod == 1) {
    vv[0] = v1[0]; vv[1] = v1[1]; vv[2] = v1[2];
    ee0 = ee1;
  } else {
    // stage points (mark3mg.c:333-335, 363-365, 393-395), wrapped on periodic axes (:337-355;
    // the z test compares with the x size, as the reference's does)
    auto stage = [&](const double *vin, bool half, double *vout) -> double {
      double P[3] = {half ? fx + vin[0] * dt / 2.0 : fx + vin[0] * dt, half ? fy + vin[1] * dt / 2.0 : fy + vin[1] * dt,
                     half ? fz + vin[2] * dt / 2.0 : fz + vin[2] * dt};
      if (PER) {
        const double sz[3] = {g.xs, g.ys, g.zs}, cmp[3] = {g.xs, g.ys, g.xs};
        const int pr[3] = {g.px, g.py, g.pz};
#pragma unroll
        for (int a = 0; a < 3; a++)
          if (pr[a]) {
            if (P[a] < 0) P[a] = sz[a] + P[a];
            if (P[a] > cmp[a]) P[a] = P[a] - sz[a];
            out |= P[a] < 0 || P[a] > sz[a];
          }
      }
      adv_interp<PER>(g, v3, e6, P[0], P[1], P[2], vout, wantE ? e : nullptr);
      return wantE ? adv_eii(e) : 0.0;
    };
    double v2[3], w3[3], v4[3];
    const double ee2 = stage(v1, true, v2);
    const double ee3 = stage(v2, true, w3);
    const double ee4 = stage(w3, false, v4);
    for (int k = 0; k < 3; k++) vv[k] = (v1[k] + 2.0 * v2[k] + 2.0 * w3[k] + v4[k]) / 6.0;
    ee0 = (ee1 + 2.0 * ee2 + 2.0 * ee3 + ee4) / 6.0;
  }
  if (outgrid == 2) {  // orthogonal motion only (mark3mg.c:433-457): one periodic branch, the first set
    if (PER && g.px) {
      if (fy < 0 || (double)fy > g.ys) vv[0] = vv[2] = 0;
      if (fz < 0 || (double)fz > g.zs) vv[0] = vv[1] = 0;
    } else if (PER && g.py) {
      if (fx < 0 || (double)fx > g.xs) vv[1] = vv[2] = 0;
      if (fz < 0 || (double)fz > g.zs) vv[0] = vv[1] = 0;
    } else if (PER && g.pz) {
      if (fx < 0 || (double)fx > g.xs) vv[1] = vv[2] = 0;
      if (fy < 0 || (double)fy > g.zs) vv[0] = vv[2] = 0;
    } else {
      if (fx < 0 || (double)fx > g.xs) vv[1] = vv[2] = 0;
      if (fy < 0 || (double)fy > g.ys) vv[0] = vv[2] = 0;
      if (fz < 0 || (double)fz > g.zs) vv[0] = vv[1] = 0;
    }
  }
  float nx = fx + dt * vv[0], ny = fy + dt * vv[1], nz = fz + dt * vv[2];
  // periodic wrap of the moved marker (mark3mg.c:468-485)
  if (PER && g.px) {
    if (nx < 0) nx = (float)(g.xs + nx);
    if (nx > g.xs) nx = (float)(nx - g.xs);
    out |= nx < 0 || nx > g.xs;
  }
  if (PER && g.py) {
    if (ny < 0) ny = (float)(g.ys + ny);
    if (ny > g.ys) ny = (float)(ny - g.ys);
    out |= ny < 0 || ny > g.ys;
  }
  if (PER && g.pz) {
    if (nz < 0) nz = (float)(g.zs + nz);
    if (nz > g.zs) nz = (float)(nz - g.zs);
    out |= nz < 0 || nz > g.zs;
  }
  if (PER && out && g.wrap_out) g.wrap_out[0] = 1;
  mx[i] = nx;
  my[i] = ny;
  mz[i] = nz;
  if (t < 100) {
    if (wantE) marke[i] = marke[i] + (float)(ee0 * dt);
    if (markk && (nx < 0 || ny < 0 || nz < 0 || (double)nx > g.xs || (double)ny > g.ys || (double)nz > g.zs))
      markk[i] = 0.0f;
  }
}

}  // namespace stk
