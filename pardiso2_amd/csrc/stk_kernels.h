// stk_kernels.h -- HIP kernels of the Stokes multigrid library (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include "stk_rows.h"

namespace stk {

template <typename T> struct V4;
template <> struct V4<float> { using t = float4; };
template <> struct V4<double> { using t = double4; };

// Workgroups go to the 8 XCDs round-robin (block b on XCD b % 8), and each
// XCD has its own L2.  xcd_block() renumbers blocks so that every XCD owns a
// contiguous run of the index space (a slab of tiles or nodes): neighbouring
// blocks then share halos and neighbour viscosities in the same L2.  The map
// is a bijection of [0, gridDim.x).
constexpr int kXcds = 8;
__device__ __forceinline__ long xcd_block() {
  const long nb = gridDim.x, b = blockIdx.x;
  const long x = b % kXcds, j = b / kXcds;
  const long q = nb / kXcds, r = nb % kXcds;
  return x * q + (x < r ? x : r) + j;
}

// Tile of a workgroup for the tile sweeps: the blocks (renumbered by
// xcd_block, so each XCD owns a contiguous run) walk the tiles in pencils
// of kPencil tiles along x: within a pencil, y fastest, then x, then z.  An
// XCD then works on a pencil's z-layers in turn, so the halo planes a tile
// shares with the tiles below it and beside it were read by tiles that ran
// shortly before on the same XCD (its L2), not a whole z-layer of the level
// earlier.  The map is a bijection of the tile indices; results do not
// depend on it (tiles are independent within a sweep).
constexpr int kPencil = 8;
__device__ __forceinline__ void tile_of_block(int p, int nty, int ntx, int ntz, int &by, int &bx, int &bz) {
  const int full = nty * ntz * kPencil;
  const int pp = p / full, r = p - pp * full;
  const int w = min(kPencil, ntx - pp * kPencil);
  by = r % nty;
  const int q = r / nty;
  const int kxi = q % w;
  bz = q / w;
  bx = pp * kPencil + kxi;
}

__device__ __forceinline__ void decode(long n, int X, int Y, int XY, int &m1, int &m2, int &m3) {
  m3 = (int)(n / XY);
  int rem = (int)(n - (long)m3 * XY);
  m1 = rem / Y;
  m2 = rem - m1 * Y;
}

// ---------------------------------------------------------------------------
// Stencil viscosities of one level (nueff of the stress/continuity stencils,
// parallel/move3mg.c:36-50, 432-447, 567-580, 708-721, 1609-1627), clamped
// to [nubeg, nuend].  Level 0 reads nxx/nxy/nxz/nyz, coarse levels average
// node nu by viscmod; the grad-div cell value always averages nu.
// Sums of FP32 node values are formed in FP32 as in the reference.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double clampnu(double e, double nb, double ne) {
  if (e > ne) e = ne;
  if (e < nb) e = nb;
  return e;
}

// 8-node mean of the cell whose lower corner is node v0.  On the last cell
// of a periodic axis the reference swaps one node of the octet for its
// line-0 copy (x: v[2], y: v[1], z: v[4]; move3mg.c:27-29), and so does
// this (wx, wy, wz: the replacement nodes, or -1).
__device__ __forceinline__ double mean8(const float *nu, long v0, int Y, int XY, int vm, long wx = -1,
                                        long wy = -1, long wz = -1) {
  long v[8];
  v[0] = v0; v[1] = v0 + 1; v[2] = v0 + Y; v[3] = v[2] + 1;
  v[4] = v0 + XY; v[5] = v[4] + 1; v[6] = v[4] + Y; v[7] = v[6] + 1;
  if (wx >= 0) v[2] = wx;
  if (wy >= 0) v[1] = wy;
  if (wz >= 0) v[4] = wz;
  if (vm == 1) {
    double s = log((double)nu[v[0]]);
    for (int k = 1; k < 8; k++) s = s + log((double)nu[v[k]]);
    return exp(s / 8.0);
  }
  if (vm == 2) {
    double s = 1.0 / nu[v[0]];
    for (int k = 1; k < 8; k++) s = s + 1.0 / nu[v[k]];
    return 1.0 / (s / 8.0);
  }
  float s = nu[v[0]];
  for (int k = 1; k < 8; k++) s = s + nu[v[k]];
  return s / 8.0;
}

__device__ __forceinline__ double mean2(const float *nu, long a, long b, int vm) {
  if (vm == 1) return exp((log((double)nu[a]) + log((double)nu[b])) / 2.0);
  if (vm == 2) return 1.0 / ((1.0 / nu[a] + 1.0 / nu[b]) / 2.0);
  return (nu[a] + nu[b]) / 2.0;
}

__global__ void k_visc_prep(int X, int Y, int Z, int level0, const float *__restrict__ nu,
                            const float *__restrict__ nxx, const float *__restrict__ nxy,
                            const float *__restrict__ nxz, const float *__restrict__ nyz,
                            double nubeg, double nuend, int vm, double *e64c, double *e64xy,
                            double *e64xz, double *e64yz, double *e64d, float *e32c, float *e32xy,
                            float *e32xz, float *e32yz, float *e32d, int per) {
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int XY = X * Y;
  if (n >= (long)XY * Z) return;
  int m1, m2, m3;
  decode(n, X, Y, XY, m1, m2, m3);
  double ec = nubeg, ed = nubeg, exy = nubeg, exz = nubeg, eyz = nubeg;
  if (m1 >= 1 && m2 >= 1 && m3 >= 1) {
    long v0 = n - XY - Y - 1;
    const long wx = ((per & 1) && m1 == X - 1) ? (long)(m3 - 1) * XY + (m2 - 1) : -1;
    const long wy = ((per & 2) && m2 == Y - 1) ? (long)(m3 - 1) * XY + (long)(m1 - 1) * Y : -1;
    const long wz = ((per & 4) && m3 == Z - 1) ? (long)(m1 - 1) * Y + (m2 - 1) : -1;
    ed = clampnu(mean8(nu, v0, Y, XY, vm, wx, wy, wz), nubeg, nuend);
    ec = level0 ? clampnu((double)nxx[n], nubeg, nuend) : ed;
  }
  if (level0) {
    exy = clampnu((double)nxy[n], nubeg, nuend);
    exz = clampnu((double)nxz[n], nubeg, nuend);
    eyz = clampnu((double)nyz[n], nubeg, nuend);
  } else {
    if (m3 < Z - 1) exy = clampnu(mean2(nu, n, n + XY, vm), nubeg, nuend);
    if (m2 < Y - 1) exz = clampnu(mean2(nu, n, n + 1, vm), nubeg, nuend);
    if (m1 < X - 1) eyz = clampnu(mean2(nu, n, n + Y, vm), nubeg, nuend);
  }
  e64c[n] = ec; e64xy[n] = exy; e64xz[n] = exz; e64yz[n] = eyz; e64d[n] = ed;
  e32c[n] = (float)ec; e32xy[n] = (float)exy; e32xz[n] = (float)exz; e32yz[n] = (float)eyz;
  e32d[n] = (float)ed;
}

// Periodic copies of the stencil viscosities (after k_visc_prep): a cell
// plane at index 0 of a periodic axis holds the value of its cell n-1, an
// edge plane at index n-1 the value of its edge on line 0 (pwrap / vwrap),
// so rows next to the wrap read the wrapped stress points at the offsets
// they use elsewhere.  Sources are never copies themselves.
__global__ void k_visc_wrap(int X, int Y, int Z, int per, double *e64c, double *e64xy, double *e64xz,
                            double *e64yz, double *e64d, float *e32c, float *e32xy, float *e32xz, float *e32yz,
                            float *e32d) {
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int XY = X * Y;
  if (n >= (long)XY * Z) return;
  int m1, m2, m3;
  decode(n, X, Y, XY, m1, m2, m3);
  const int c1 = (per & 1) ? pwrap(m1, X) : m1, c2 = (per & 2) ? pwrap(m2, Y) : m2, c3 = (per & 4) ? pwrap(m3, Z) : m3;
  const long sc = (long)c3 * XY + (long)c1 * Y + c2;
  if (sc != n) {
    e64c[n] = e64c[sc]; e64d[n] = e64d[sc];
    e32c[n] = e32c[sc]; e32d[n] = e32d[sc];
  }
  const int v1 = (per & 1) ? vwrap(m1, X) : m1, v2 = (per & 2) ? vwrap(m2, Y) : m2, v3 = (per & 4) ? vwrap(m3, Z) : m3;
  const long se = (long)v3 * XY + (long)v1 * Y + v2;
  if (se != n) {
    e64xy[n] = e64xy[se]; e64xz[n] = e64xz[se]; e64yz[n] = e64yz[se];
    e32xy[n] = e32xy[se]; e32xz[n] = e32xz[se]; e32yz[n] = e32yz[se];
  }
}

// ---------------------------------------------------------------------------
// Finest-level right-hand side: gravity terms of the momentum rows
// (move3mg.c:848, 1061, 1281), zero for continuity, the constant of
// boundary equations (move3mg.c:1710-1712).
// ---------------------------------------------------------------------------
__global__ void k_rhs(int X, int Y, int Z, const unsigned char *__restrict__ flags,
                      const float *__restrict__ ro, double gx, double gy, double gz, double *b, int per) {
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int XY = X * Y;
  if (n >= (long)XY * Z) return;
  int m1, m2, m3;
  decode(n, X, Y, XY, m1, m2, m3);
  unsigned char f = flags[n];
  double *bn = b + 4 * n;
  // rows on line 0 of a periodic axis take the fourth density node from the
  // last line (the node the reference also moves into the pressure slot:
  // x row v[5], y row v[6], z row v[3]; move3mg.c:842, 1056, 1276)
  const long wx = ((per & 1) && m1 == 0) ? n + XY + 1 + (long)(X - 1) * Y : n + XY + 1;
  const long wy = ((per & 2) && m2 == 0) ? n + XY + Y + (Y - 1) : n + XY + Y;
  const long wz = ((per & 4) && m3 == 0) ? n + Y + 1 + (long)(Z - 1) * XY : n + Y + 1;
  if (kind_of(f, 0) != KIND_BC) bn[0] = 0.0;
  if (kind_of(f, 1) == KIND_EQ) bn[1] = -gx * (ro[n] + ro[n + 1] + ro[n + XY] + ro[wx]) / 4.0;
  else if (kind_of(f, 1) == KIND_NONE) bn[1] = 0.0;
  if (kind_of(f, 2) == KIND_EQ) bn[2] = -gy * (ro[n] + ro[n + Y] + ro[n + XY] + ro[wy]) / 4.0;
  else if (kind_of(f, 2) == KIND_NONE) bn[2] = 0.0;
  if (kind_of(f, 3) == KIND_EQ) bn[3] = -gz * (ro[n] + ro[n + 1] + ro[n + Y] + ro[wz]) / 4.0;
  else if (kind_of(f, 3) == KIND_NONE) bn[3] = 0.0;
}

__global__ void k_bc_rhs(int nbc, const int *__restrict__ r, const double *__restrict__ c, double *b) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nbc) b[r[i]] = c[i];
}

// boundary rows  x_r - k x_j = b_r  (bonderr, move3mg.c:1682-1729)
template <typename T, bool RES>
__global__ void k_bc_rowop(int nbc, const int *__restrict__ br, const int *__restrict__ bj,
                           const double *__restrict__ negk, const T *__restrict__ x,
                           const T *__restrict__ b, T *__restrict__ y, int withp) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nbc) return;
  int r = br[i];
  if (!withp && (r & 3) == 0) { y[r] = T(0); return; }
  T acc;
  if (RES) {
    acc = b[r];
    acc -= T(1) * x[r];
    if (bj[i] >= 0) acc -= (T)negk[i] * x[bj[i]];
  } else {
    acc = T(0);
    acc += T(1) * x[r];
    if (bj[i] >= 0) acc += (T)negk[i] * x[bj[i]];
  }
  y[r] = acc;
}

// ---------------------------------------------------------------------------
// Smoother: one sweep of the reference's row updates (gausmat4,
// move3mg.c:3893-3923) in tile-coloured order.  A block owns a ty*tx*tz node
// tile, stages it with a one-node halo in LDS from xold, and updates in
// phases vx(red,black), vy(red,black), vz(red,black)[, P], red/black by
// (m1+m2+m3)&1; halo values stay those of xold (block-Jacobi across tiles),
// so the sweep is deterministic.  Boundary rows are skipped here and done
// by k_bc_update afterwards.  The tile is written to xnew.
// Division by a small runtime divisor d (2 <= d < 2^16) for numerators
// i < 2^16: i / d = umulhi(i, ceil(2^32 / d)).  With m = (2^32 + e) / d,
// 0 <= e < d, the product exceeds i/d by i e / (d 2^32) < 1/d, less than the
// distance from i/d to the next integer, so the floor is exact.  d == 1 is
// passed through.
struct FDiv {
  unsigned d, m;
};
__host__ __device__ inline FDiv make_fdiv(unsigned d) {
  return FDiv{d, d > 1 ? (unsigned)((0x100000000ull + d - 1) / d) : 0u};
}
__device__ __forceinline__ unsigned fdiv(unsigned i, FDiv f) { return f.d == 1 ? i : __umulhi(i, f.m); }
// i / D for a divisor known at compile time (D > 0), else by f
template <int D>
__device__ __forceinline__ unsigned tdiv(unsigned i, FDiv f) { return D ? i / (unsigned)(D ? D : 1) : fdiv(i, f); }

// Velocity update residual / diagonal: FP64 divides as the reference does;
// the FP32 smoother multiplies by the hardware reciprocal (1 ulp).
__device__ __forceinline__ double diag_step(double r, double d) { return r / d; }
__device__ __forceinline__ float diag_step(float r, float d) { return r * __builtin_amdgcn_rcpf(d); }

// Smoother tile of ty*tx*tz nodes (y fastest) with its one-node halo
// (hy = ty+2, hx = tx+2).  The ntx*nty*ntz tiles cover the box of nodes that
// hold equation rows, from its lower corner (oy, ox, oz): nodes outside it
// carry only boundary rows (set by the boundary chains after each sweep) or
// no unknowns.
struct TileGeo {
  int ty, tx, tz, ntx, nty, ntz, oy, ox, oz;
  FDiv dty, dtx, dhy, dhx, dhty;  // ty, tx, ty+2, tx+2, ceil(ty/2)
};
// box = {y0, x0, z0, y1, x1, z1}, upper bounds exclusive
inline TileGeo make_tile(int ty, int tx, int tz, const int box[6]) {
  TileGeo t;
  t.ty = ty; t.tx = tx; t.tz = tz;
  t.oy = box[0]; t.ox = box[1]; t.oz = box[2];
  t.nty = (box[3] - box[0] + ty - 1) / ty; t.ntx = (box[4] - box[1] + tx - 1) / tx;
  t.ntz = (box[5] - box[2] + tz - 1) / tz;
  t.dty = make_fdiv(ty); t.dtx = make_fdiv(tx); t.dhy = make_fdiv(ty + 2); t.dhx = make_fdiv(tx + 2);
  t.dhty = make_fdiv((ty + 1) / 2);
  return t;
}

// ---------------------------------------------------------------------------
// one phase of the tile smoother for component C (colour for velocities)
template <typename T, int C, bool SV, int TY, int TX, int TZ>
__device__ __forceinline__ void smooth_phase(const Geo<T> &g, const Geo<T> &gv, const unsigned char *__restrict__ sf,
                                             const T *__restrict__ b, T *sx, int nh, const TileGeo &tg,
                                             int x0, int y0, int z0, int colour, T v0koef, T p0koef) {
  int ty = TY ? TY : tg.ty, tx = TX ? TX : tg.tx, tz = TZ ? TZ : tg.tz;
  // values derived from a run-time tile shape are rebuilt per phase rather
  // than hoisted out of the caller's phase loop and held in scalar registers
  if constexpr (!TY) asm volatile("" : "+s"(ty), "+s"(tx), "+s"(tz));
  const int hy = ty + 2, hx = tx + 2;
  const int hty = (ty + 1) / 2;
  const int ncand = (C == 0) ? ty * tx * tz : hty * tx * tz;
  for (int i = threadIdx.x; i < ncand; i += blockDim.x) {
    int ly, lx, lz;
    // index products in 24-bit multiplies (full rate): tile and level
    // extents stay below 2^24 and node numbers below 2^32
    if (C == 0) {
      const int q = (int)tdiv<TY>(i, tg.dty);
      ly = i - (int)__umul24(q, ty); lz = (int)tdiv<TX>(q, tg.dtx); lx = q - (int)__umul24(lz, tx);
    } else {
      const int q = (int)tdiv<(TY + 1) / 2>(i, tg.dhty);
      const int kk = i - (int)__umul24(q, hty);
      lz = (int)tdiv<TX>(q, tg.dtx); lx = q - (int)__umul24(lz, tx);
      int par = (colour - (x0 + lx) - (z0 + lz) - y0) & 1;
      ly = 2 * kk + par;
      if (ly >= ty) continue;
    }
    int m1 = x0 + lx, m2 = y0 + ly, m3 = z0 + lz;
    if (m1 >= g.X || m2 >= g.Y || m3 >= g.Z) continue;
    const unsigned n = __umul24(m3, g.XY) + __umul24(m1, g.Y) + m2;
    if (kind_of(sf[__umul24(__umul24(lz, tx) + lx, ty) + ly], C) != KIND_EQ) continue;
    T s[18];
    // the row's viscosities: in the tile's LDS block (SV) or in HBM
    const unsigned v = SV ? __umul24(__umul24(lz, tx + 1) + lx, ty + 1) + ly : n;
    row_coef_c<T, C>(gv, m1, m2, m3, (long)v, s);
    T ival = -b[4 * (size_t)n + C];
    const int base = (int)__umul24(__umul24(lz + 1, hx) + lx + 1, hy) + ly + 1;
#pragma unroll
    for (int q = 1; q <= Slot<C>::ns; q++) {
      const int li = base + (Slot<C>::d[q][2] * hx + Slot<C>::d[q][0]) * hy + Slot<C>::d[q][1];
      ival = mac(ival, s[q], sx[Slot<C>::d[q][3] * nh + li]);
    }
    if (C == 0) sx[base] -= ival * g.ec[n] * p0koef;
    else sx[C * nh + base] -= diag_step(ival, s[1]) * v0koef;
  }
}

// Boundary rows along their dependency chains in one launch.  Row i's chain
// is r_0 (the row itself), r_1, ..., r_d, each x_{r_t} = b_{r_t} + k_t x_{r_{t+1}},
// ending at x_{j_i} with j_i not a boundary row (or no partner, j_i = -1).
// Every thread evaluates its whole chain from the final partner, so all rows
// of a chain see the same values without ordering the launch by depth; the
// partners read are never written here.
struct BcChain {
  int nbc, dm;
  const int *r, *len, *j;
  const double *k;
};

// A smoothing call's boundary chains in composed form, evaluated by the tile
// sweeps themselves: with b fixed for the call, chain i is
// x_{r_0} = c_i + K_i x_{j_i}, where c_i and K_i fold the chain's b and k
// terms from the partner end (K_i = 0 without a partner).  A chain whose row
// lies in a tile's interior belongs to that tile, and its partner too (the
// tile sets the row in LDS before its write-back); any other chain belongs
// to the tile whose interior holds its partner (or, without one, the tile
// nearest its row) and is stored to HBM directly.  Everything but off is in
// list order (the chains of tile t at positions off[t] .. off[t+1]-1); in
// device memory, passed by one pointer.
template <typename T>
struct TileChains {
  const int *off;   // [ntiles + 1]
  const int *idx;   // chain number at each position
  const int *head;  // its r_0
  const int *rl;    // LDS index of r_0 in the owning tile (-1: outside every tile)
  const int *pl;    // LDS index of the partner (-1: none)
  T *cc, *ck;       // its composed terms
};

// c and K at every list position from the current b (mode 1: velocity rows
// only; pressure chains are marked by K = NaN and left alone)
template <typename T>
__global__ void k_bc_compose(BcChain ch, const int *__restrict__ idx, const T *__restrict__ b, T *cc, T *ck,
                             int mode) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ch.nbc) return;
  const int i = idx[q];
  const long o = (long)i * ch.dm;
  T c = T(0), K = ch.j[i] >= 0 ? T(1) : T(0);
  for (int t = ch.len[i] - 1; t >= 0; t--) {
    const T k = (T)ch.k[o + t];
    c = b[ch.r[o + t]] + k * c;
    K = k * K;
  }
  if (mode == 1 && (ch.r[o] & 3) == 0) K = __builtin_nan("");
  cc[q] = c;
  ck[q] = K;
}

// Stage a tile and its one-node halo of (P, vx, vy, vz) in LDS, component-
// major (zeros outside the level), and the unknown kinds of the interior.
// On a periodic axis a position outside the level's lines, or on a line
// whose unknown is a copy, holds the unknown it names (vwrap for the
// velocities, pwrap for P, stk_rows.h), so the rows next to the wrap read
// what the reference's periodic stencils read; such positions in the halo
// hold the values of xold like every other halo position.
template <typename T>
struct StagePos {
  bool lev;         // inside the level (interior kinds are staged only there)
  bool ld;          // a value is staged (else zeros)
  unsigned nv, np;  // node of the velocities and of P
};
template <typename T>
__device__ __forceinline__ StagePos<T> stage_pos(const Geo<T> &g, int m1, int m2, int m3, bool any) {
  StagePos<T> s;
  s.lev = any && m1 >= 0 && m1 < g.X && m2 >= 0 && m2 < g.Y && m3 >= 0 && m3 < g.Z;
  if (!g.per) {
    s.ld = s.lev;
    s.nv = s.np = s.lev ? __umul24(m3, g.XY) + __umul24(m1, g.Y) + m2 : 0u;
    return s;
  }
  const int per = g.per;
  const bool ox = (per & 1) ? (m1 >= -1 && m1 <= g.X) : (m1 >= 0 && m1 < g.X);
  const bool oy = (per & 2) ? (m2 >= -1 && m2 <= g.Y) : (m2 >= 0 && m2 < g.Y);
  const bool oz = (per & 4) ? (m3 >= -1 && m3 <= g.Z) : (m3 >= 0 && m3 < g.Z);
  s.ld = any && ox && oy && oz;
  const int a1 = (per & 1) ? vwrap(m1, g.X) : m1, b1 = (per & 1) ? pwrap(m1, g.X) : m1;
  const int a2 = (per & 2) ? vwrap(m2, g.Y) : m2, b2 = (per & 2) ? pwrap(m2, g.Y) : m2;
  const int a3 = (per & 4) ? vwrap(m3, g.Z) : m3, b3 = (per & 4) ? pwrap(m3, g.Z) : m3;
  s.nv = s.ld ? __umul24(a3, g.XY) + __umul24(a1, g.Y) + a2 : 0u;
  s.np = s.ld ? __umul24(b3, g.XY) + __umul24(b1, g.Y) + b2 : 0u;
  return s;
}

template <typename T, int TY, int TX, int TZ = 0, int BS = 256>
__device__ __forceinline__ void tile_stage(const Geo<T> &g, const unsigned char *__restrict__ flags,
                                           const T *__restrict__ xold, T *sx, unsigned char *sf, const TileGeo &tg,
                                           int x0, int y0, int z0, int ty, int tx, int tz, int nh) {
  using V = typename V4<T>::t;
  const int hy = ty + 2, hx = tx + 2;
  constexpr int kNh = (TY && TX && TZ) ? (TY + 2) * (TX + 2) * (TZ + 2) : 0;
  if constexpr (kNh > 0) {
    // shape compiled in (walled levels only; periodic levels run the
    // run-time-shape instances): every load of the thread's share is issued
    // before the first LDS store, from a clamped (always valid) address, and
    // the values outside the level are replaced by zeros afterwards (no
    // branch around a load, so no wait between them)
    constexpr int kIt = (kNh + BS - 1) / BS;
    V v[kIt];
    unsigned char f[kIt];
    bool inl[kIt], intr[kIt];
#pragma unroll
    for (int k = 0; k < kIt; k++) {
      const int i = k * BS + (int)threadIdx.x;
      const int q = i / (TY + 2);
      const int ly = i - q * (TY + 2), lz = q / (TX + 2), lx = q - lz * (TX + 2);
      const int m2 = y0 + ly - 1, m1 = x0 + lx - 1, m3 = z0 + lz - 1;
      inl[k] = i < kNh && m1 >= 0 && m1 < g.X && m2 >= 0 && m2 < g.Y && m3 >= 0 && m3 < g.Z;
      intr[k] = i < kNh && lx >= 1 && lx <= TX && ly >= 1 && ly <= TY && lz >= 1 && lz <= TZ;
      const unsigned n = inl[k] ? __umul24(m3, g.XY) + __umul24(m1, g.Y) + m2 : 0u;
      v[k] = reinterpret_cast<const V *>(xold)[n];
      f[k] = flags[n];
    }
#pragma unroll
    for (int k = 0; k < kIt; k++) {
      const int i = k * BS + (int)threadIdx.x;
      if (i >= kNh) break;
      const V z = {T(0), T(0), T(0), T(0)};
      const V w = inl[k] ? v[k] : z;
      sx[i] = w.x; sx[kNh + i] = w.y; sx[2 * kNh + i] = w.z; sx[3 * kNh + i] = w.w;
      if (intr[k]) {
        const int q = i / (TY + 2);
        const int ly = i - q * (TY + 2), lz = q / (TX + 2), lx = q - lz * (TX + 2);
        sf[((lz - 1) * TX + lx - 1) * TY + ly - 1] = inl[k] ? f[k] : (unsigned char)0;
      }
    }
    return;
  }
  for (int i = threadIdx.x; i < nh; i += blockDim.x) {
    const int q = (int)tdiv<TY ? TY + 2 : 0>(i, tg.dhy);
    const int ly = i - q * hy, lz = (int)tdiv<TX ? TX + 2 : 0>(q, tg.dhx), lx = q - lz * hx;
    int m2 = y0 + ly - 1, m1 = x0 + lx - 1, m3 = z0 + lz - 1;
    V v = {T(0), T(0), T(0), T(0)};
    unsigned char f = 0;
    const bool interior = lx >= 1 && lx <= tx && ly >= 1 && ly <= ty && lz >= 1 && lz <= tz;
    const StagePos<T> sp = stage_pos<T>(g, m1, m2, m3, true);
    if (sp.ld) {
      v = reinterpret_cast<const V *>(xold)[sp.nv];
      if (g.per) v.x = xold[4 * (size_t)sp.np];
    }
    if (interior && sp.lev) f = flags[g.nid(m1, m2, m3)];  // kinds of the halo are not needed
    sx[i] = v.x; sx[nh + i] = v.y; sx[2 * nh + i] = v.z; sx[3 * nh + i] = v.w;
    if (interior) sf[__umul24(__umul24(lz - 1, tx) + lx - 1, ty) + ly - 1] = f;
  }
}

// Write-back value of a node of a tile: on a periodic level the unknowns a
// node does not hold (kind none) are written as zeros, not as the copies
// the staging put there.
template <typename V>
__device__ __forceinline__ V wb_mask(V o, unsigned char f) {
  if (kind_of(f, 0) == KIND_NONE) o.x = 0;
  if (kind_of(f, 1) == KIND_NONE) o.y = 0;
  if (kind_of(f, 2) == KIND_NONE) o.z = 0;
  if (kind_of(f, 3) == KIND_NONE) o.w = 0;
  return o;
}

// SV: the tile's stencil viscosities (nodes [x0, x0+tx] x [y0, y0+ty] x
// [z0, z0+tz], every offset a row uses is 0 or +1) are staged in LDS after the
// unknowns, and the rows read them through a Geo whose arrays are that block
// (strides ty+1, (ty+1)(tx+1)), so each viscosity comes from HBM once per tile.
// The interior's unknown kinds are staged too (last in LDS, one byte per node),
// so a phase's row loads do not wait on a flag load from HBM.
template <typename T, bool SV, int TY = 0, int TX = 0, int TZ = 0, bool CH = false>
__global__ void __launch_bounds__(256) k_smooth(Geo<T> g, const unsigned char *__restrict__ flags,
                                                const T *__restrict__ xold, T *__restrict__ xnew,
                                                const T *__restrict__ b, TileGeo tg, int withp, T v0koef,
                                                T p0koef, const TileChains<T> *__restrict__ tcs) {
  extern __shared__ unsigned char smem_raw[];
  T *sx = reinterpret_cast<T *>(smem_raw);
  const int ty = TY ? TY : tg.ty, tx = TX ? TX : tg.tx, tz = TZ ? TZ : tg.tz, ntx = tg.ntx, nty = tg.nty;
  const int hy = ty + 2, hx = tx + 2, hz = tz + 2;
  const int bt = (int)xcd_block();
  const int by = bt % nty, bx = (bt / nty) % ntx, bz = bt / (nty * ntx);
  const int y0 = tg.oy + by * ty, x0 = tg.ox + bx * tx, z0 = tg.oz + bz * tz;
  const int nh = hy * hx * hz;
  int q0 = 0, q1 = 0;
  if (CH) { q0 = tcs->off[bt]; q1 = tcs->off[bt + 1]; }  // this tile's chains, read early
  unsigned char *sf = smem_raw + (4 * nh + (SV ? 5 * (ty + 1) * (tx + 1) * (tz + 1) : 0)) * sizeof(T);
  Geo<T> gv = g;
  if (SV) {
    const int vy = ty + 1, vxy = (ty + 1) * (tx + 1), nv = vxy * (tz + 1);
    T *sv = sx + 4 * nh;
    const int np = (g.p2 != T(0)) ? 5 : 4;
    const T *src[5] = {g.ec, g.exy, g.exz, g.eyz, g.ed};
    for (int i = threadIdx.x; i < nv; i += blockDim.x) {
      const int ly = i % vy, lx = (i / vy) % (tx + 1), lz = i / vxy;
      const int m2 = y0 + ly, m1 = x0 + lx, m3 = z0 + lz;
      const bool in = m1 < g.X && m2 < g.Y && m3 < g.Z;
      const long n = g.nid(in ? m1 : 0, in ? m2 : 0, in ? m3 : 0);
      for (int q = 0; q < np; q++) sv[q * nv + i] = in ? src[q][n] : T(0);
    }
    gv.XY = vxy;
    gv.Y = vy;
    gv.ec = sv;
    gv.exy = sv + nv;
    gv.exz = sv + 2 * nv;
    gv.eyz = sv + 3 * nv;
    gv.ed = sv + 4 * nv;
  }
  tile_stage<T, TY, TX, TZ>(g, flags, xold, sx, sf, tg, x0, y0, z0, ty, tx, tz, nh);
  __syncthreads();
  const int nphase = withp ? 7 : 6;
  for (int ph = 0; ph < nphase; ph++) {
    const int colour = ph & 1;
    if (ph < 2) smooth_phase<T, 1, SV, TY, TX, TZ>(g, gv, sf, b, sx, nh, tg, x0, y0, z0, colour, v0koef, p0koef);
    else if (ph < 4) smooth_phase<T, 2, SV, TY, TX, TZ>(g, gv, sf, b, sx, nh, tg, x0, y0, z0, colour, v0koef, p0koef);
    else if (ph < 6) smooth_phase<T, 3, SV, TY, TX, TZ>(g, gv, sf, b, sx, nh, tg, x0, y0, z0, colour, v0koef, p0koef);
    else smooth_phase<T, 0, SV, TY, TX, TZ>(g, gv, sf, b, sx, nh, tg, x0, y0, z0, colour, v0koef, p0koef);
    __syncthreads();
  }
  if (CH) {
    // this tile's chains on the new values of their partners; partners are
    // never chain rows, so the LDS writes here read no value they change
    const int *__restrict__ rl = tcs->rl;
    const int *__restrict__ pl = tcs->pl;
    const T *__restrict__ cc = tcs->cc;
    const T *__restrict__ ck = tcs->ck;
    for (int q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
      const int r = rl[q], j = pl[q];
      const T K = ck[q];
      if (K != K) continue;  // a pressure chain in a velocity-only sweep
      const T v = j >= 0 ? cc[q] + K * sx[j] : cc[q];
      if (r >= 0) sx[r] = v;
      else xnew[tcs->head[q]] = v;
    }
    __syncthreads();
  }
  // write the tile interior
  const int nt = ty * tx * tz;
  for (int i = threadIdx.x; i < nt; i += blockDim.x) {
    const int q = (int)tdiv<TY>(i, tg.dty);
    const int ly = i - q * ty, lz = (int)tdiv<TX>(q, tg.dtx), lx = q - lz * tx;
    int m1 = x0 + lx, m2 = y0 + ly, m3 = z0 + lz;
    if (m1 >= g.X || m2 >= g.Y || m3 >= g.Z) continue;
    int li = ((lz + 1) * hx + (lx + 1)) * hy + (ly + 1);
    using V = typename V4<T>::t;
    V o = V{sx[li], sx[nh + li], sx[2 * nh + li], sx[3 * nh + li]};
    if (g.per) o = wb_mask(o, sf[i]);
    reinterpret_cast<V *>(xnew)[g.nid(m1, m2, m3)] = o;
  }
}


// ---------------------------------------------------------------------------
// The FP32 velocity sweep of the V-cycle (pressure rows left alone) with the
// rows' data prefetched.  The colour phases give every thread the same pair
// of nodes in y: node A (ly = 2kk + p) in the red phases and node B
// (ly = 2kk + 1 - p) in the black ones, for all three components.  So each
// thread loads, before the tile is staged, everything its six rows read from
// global memory: the right-hand sides, the 13 stencil viscosities of each
// node (+ grad-div cells) and the spacing values.  The phases then only read
// LDS, and the loads' latency overlaps the staging.  The rows are evaluated in
// flux form (stk_rows.h: the same coefficients, summed as pair differences,
// about half the arithmetic of building the 17 coefficients).  One node pair
// per thread: ceil(ty/2) * tx * tz <= BS, the block size.
template <typename T>
struct PfNode {
  bool ok;             // inside the level and below the upper faces (every velocity row's stencil fits)
  int m2;              // y index (m1, m3 are the pair's)
  unsigned n;          // node number
  T b1, b2, b3;        // right-hand sides of vx, vy, vz
  T ay0, ay1, by0, by1;
  Vis<T> vx, vy, vz;
};

template <typename T, bool UNI>
__device__ __forceinline__ void pf_load(const Geo<T> &g, const T *__restrict__ b, int m1, int m2, int m3, bool in,
                                        PfNode<T> &d) {
  d.m2 = m2;
  d.ok = in && m1 <= g.X - 2 && m2 <= g.Y - 2 && m3 <= g.Z - 2;
  // loads without a branch: a node that has no velocity row reads node 0's
  // neighbourhood (inside every level) and its values are never used
  d.n = d.ok ? __umul24(m3, g.XY) + __umul24(m1, g.Y) + m2 : 0u;
  const int my = d.ok ? m2 : 1;
  const float4 bb = reinterpret_cast<const float4 *>(b)[d.n];
  d.b1 = bb.y; d.b2 = bb.z; d.b3 = bb.w;
  if (!UNI) { d.ay0 = g.iAy[my]; d.ay1 = g.iAy[my + 1]; d.by0 = g.iBy[my]; d.by1 = g.iBy[my + 1]; }
  d.vx = vis_load<T, 1>(g, d.n);
  d.vy = vis_load<T, 2>(g, d.n);
  d.vz = vis_load<T, 3>(g, d.n);
}

// S (sum over the row's slots) and D (its diagonal) in flux form
template <int C>
__device__ __forceinline__ void pf_flux(const PfNode<float> &d, const Sp<float> &pxz, float p2, const float *xs,
                                        float &S, float &D) {
  const Sp<float> p{pxz.ax0, pxz.ax1, pxz.bx0, pxz.bx1, d.ay0, d.ay1, d.by0, d.by1,
                    pxz.az0, pxz.az1, pxz.bz0, pxz.bz1};
  if (C == 1) flux_x<float>(p, d.vx, p2, xs, S, D);
  else if (C == 2) flux_y<float>(p, d.vy, p2, xs, S, D);
  else flux_z<float>(p, d.vz, p2, xs, S, D);
}

// UNI: the level's spacing is uniform along each axis; the rows use the
// level constants fu (FluxU) instead of the spacing tables.
template <int TY, int TX, int TZ, bool CH, bool UNI, int BS = 256>
__global__ void __launch_bounds__(BS) k_vsmooth32(Geo<float> g, const unsigned char *__restrict__ flags,
                                                   const float *__restrict__ xold, float *__restrict__ xnew,
                                                   const float *__restrict__ b, TileGeo tg, float v0koef,
                                                   const TileChains<float> *__restrict__ tcs, FluxU fu) {
  extern __shared__ unsigned char smem_raw[];
  float *sx = reinterpret_cast<float *>(smem_raw);
  const int ty = TY ? TY : tg.ty, tx = TX ? TX : tg.tx, tz = TZ ? TZ : tg.tz, ntx = tg.ntx, nty = tg.nty;
  const int hy = ty + 2, hx = tx + 2, hz = tz + 2;
  int by, bx, bz;
  tile_of_block((int)xcd_block(), nty, ntx, tg.ntz, by, bx, bz);
  const int bt = (bz * ntx + bx) * nty + by;  // the tile's number (chain lists)
  const int y0 = tg.oy + by * ty, x0 = tg.ox + bx * tx, z0 = tg.oz + bz * tz;
  const int nh = hy * hx * hz;
  int q0 = 0, q1 = 0;
  if (CH) { q0 = tcs->off[bt]; q1 = tcs->off[bt + 1]; }
  unsigned char *sf = smem_raw + 4 * nh * sizeof(float);
  // this thread's node pair and everything its rows read from global memory
  const int hty = (ty + 1) / 2;
  const int i = threadIdx.x;
  const bool act = i < hty * tx * tz;
  const int q = (int)tdiv<(TY + 1) / 2>(i, tg.dhty);
  const int kk = i - (int)__umul24(q, hty);
  const int lz = (int)tdiv<TX>(q, tg.dtx), lx = q - (int)__umul24(lz, tx);
  const int m1 = x0 + lx, m3 = z0 + lz;
  const int par = (-m1 - m3 - y0) & 1;  // ly of the red node is 2kk + par
  const int lyA = 2 * kk + par, lyB = 2 * kk + 1 - par;
  const bool in13 = act && m1 < g.X && m3 < g.Z;
  Sp<float> pxz{};
  if (!UNI) {
    const bool okxz = in13 && m1 <= g.X - 2 && m3 <= g.Z - 2;
    const int k1 = okxz ? m1 : 1, k3 = okxz ? m3 : 1;
    pxz.ax0 = g.iAx[k1]; pxz.ax1 = g.iAx[k1 + 1]; pxz.bx0 = g.iBx[k1]; pxz.bx1 = g.iBx[k1 + 1];
    pxz.az0 = g.iAz[k3]; pxz.az1 = g.iAz[k3 + 1]; pxz.bz0 = g.iBz[k3]; pxz.bz1 = g.iBz[k3 + 1];
  }
  PfNode<float> A, B;
  pf_load<float, UNI>(g, b, m1, y0 + lyA, m3, in13 && lyA < ty && y0 + lyA < g.Y, A);
  pf_load<float, UNI>(g, b, m1, y0 + lyB, m3, in13 && lyB < ty && y0 + lyB < g.Y, B);
  tile_stage<float, TY, TX, TZ, BS>(g, flags, xold, sx, sf, tg, x0, y0, z0, ty, tx, tz, nh);
  __syncthreads();
  const int sbase = (int)__umul24(__umul24(lz, tx) + lx, ty);
  const int lbase = (int)__umul24(__umul24(lz + 1, hx) + lx + 1, hy) + 1;
  auto phase = [&](auto cc, const PfNode<float> &d, int ly) {
    constexpr int C = decltype(cc)::value;
    if (!d.ok || kind_of(sf[sbase + ly], C) != KIND_EQ) return;
    const int base = lbase + ly;
    float xs[18];
    xs[0] = 0.f;
#pragma unroll
    for (int k = 1; k <= Slot<C>::ns; k++) {
      const int li = base + (Slot<C>::d[k][2] * hx + Slot<C>::d[k][0]) * hy + Slot<C>::d[k][1];
      xs[k] = sx[Slot<C>::d[k][3] * nh + li];
    }
    float S, D;
    if (UNI) {
      const Vis<float> &w = C == 1 ? d.vx : (C == 2 ? d.vy : d.vz);
      flux_u<C>(w, fu.g[C - 1], g.p2 != 0.f, xs, S, D);
    } else {
      pf_flux<C>(d, pxz, g.p2, xs, S, D);
    }
    const float ival = S - (C == 1 ? d.b1 : (C == 2 ? d.b2 : d.b3));
    sx[C * nh + base] = xs[1] - diag_step(ival, D) * v0koef;
  };
  phase(std::integral_constant<int, 1>{}, A, lyA);
  __syncthreads();
  phase(std::integral_constant<int, 1>{}, B, lyB);
  __syncthreads();
  phase(std::integral_constant<int, 2>{}, A, lyA);
  __syncthreads();
  phase(std::integral_constant<int, 2>{}, B, lyB);
  __syncthreads();
  phase(std::integral_constant<int, 3>{}, A, lyA);
  __syncthreads();
  phase(std::integral_constant<int, 3>{}, B, lyB);
  __syncthreads();
  if (CH) {
    const int *__restrict__ rl = tcs->rl;
    const int *__restrict__ pl = tcs->pl;
    const float *__restrict__ cc = tcs->cc;
    const float *__restrict__ ck = tcs->ck;
    for (int qq = q0 + threadIdx.x; qq < q1; qq += blockDim.x) {
      const int r = rl[qq], j = pl[qq];
      const float K = ck[qq];
      if (K != K) continue;  // a pressure chain: velocity-only sweep
      const float v = j >= 0 ? cc[qq] + K * sx[j] : cc[qq];
      if (r >= 0) sx[r] = v;
      else xnew[tcs->head[qq]] = v;
    }
    __syncthreads();
  }
  const int nt = ty * tx * tz;
  for (int t = threadIdx.x; t < nt; t += blockDim.x) {
    const int qt = (int)tdiv<TY>(t, tg.dty);
    const int ly = t - qt * ty, lzt = (int)tdiv<TX>(qt, tg.dtx), lxt = qt - lzt * tx;
    const int n1 = x0 + lxt, n2 = y0 + ly, n3 = z0 + lzt;
    if (n1 >= g.X || n2 >= g.Y || n3 >= g.Z) continue;
    const int li = ((lzt + 1) * hx + (lxt + 1)) * hy + (ly + 1);
    float4 o = float4{sx[li], sx[nh + li], sx[2 * nh + li], sx[3 * nh + li]};
    if (g.per) o = wb_mask(o, sf[t]);
    reinterpret_cast<float4 *>(xnew)[g.nid(n1, n2, n3)] = o;
  }
}

// ---------------------------------------------------------------------------
// Row operator on equation rows: RES -> y = b - A x (gausmat4's residual
// order, move3mg.c:3937-3958), otherwise y = A x.  withp=0 leaves pressure
// rows at zero (velocity block).  Tiles of the whole level: x and the
// unknown kinds staged in LDS as in the smoother, then every equation row of
// the tile's nodes from LDS in slot order.  The four unknowns of a node are
// written with one vector store; boundary rows get a placeholder 0 there and
// their values from k_bc_rowop, which always follows.
// ---------------------------------------------------------------------------
template <typename T, bool RES, int TY = 0, int TX = 0, int TZ = 0>
__global__ void __launch_bounds__(256) k_rowop_tile(Geo<T> g, const unsigned char *__restrict__ flags,
                                                    const T *__restrict__ x, const T *__restrict__ b,
                                                    T *__restrict__ y, TileGeo tg, int withp) {
  using V = typename V4<T>::t;
  extern __shared__ unsigned char smem_raw[];
  T *sx = reinterpret_cast<T *>(smem_raw);
  const int ty = TY ? TY : tg.ty, tx = TX ? TX : tg.tx, tz = TZ ? TZ : tg.tz;
  const int hy = ty + 2, hx = tx + 2, nh = hy * hx * (tz + 2);
  int by, bx, bz;
  tile_of_block((int)xcd_block(), tg.nty, tg.ntx, tg.ntz, by, bx, bz);
  const int y0 = tg.oy + by * ty, x0 = tg.ox + bx * tx, z0 = tg.oz + bz * tz;
  unsigned char *sf = smem_raw + 4 * nh * sizeof(T);
  tile_stage<T, TY, TX, TZ>(g, flags, x, sx, sf, tg, x0, y0, z0, ty, tx, tz, nh);
  __syncthreads();
  const int nt = ty * tx * tz;
  for (int i = threadIdx.x; i < nt; i += blockDim.x) {
    const int q = (int)tdiv<TY>(i, tg.dty);
    const int ly = i - q * ty, lz = (int)tdiv<TX>(q, tg.dtx), lx = q - lz * tx;
    const int m1 = x0 + lx, m2 = y0 + ly, m3 = z0 + lz;
    if (m1 >= g.X || m2 >= g.Y || m3 >= g.Z) continue;
    const unsigned n = __umul24(m3, g.XY) + __umul24(m1, g.Y) + m2;
    const unsigned char f = sf[i];
    const int base = (int)__umul24(__umul24(lz + 1, hx) + lx + 1, hy) + ly + 1;
    V bn = {T(0), T(0), T(0), T(0)};
    if (RES) bn = reinterpret_cast<const V *>(b)[n];
    T o[4] = {T(0), T(0), T(0), T(0)};
    const T bb[4] = {bn.x, bn.y, bn.z, bn.w};
    auto row = [&](auto cc) {
      constexpr int C = decltype(cc)::value;
      if (kind_of(f, C) != KIND_EQ || (!withp && C == 0)) return;
      T s[18];
      row_coef_c<T, C>(g, m1, m2, m3, (long)n, s);
      T acc = RES ? bb[C] : T(0);
#pragma unroll
      for (int q2 = 1; q2 <= Slot<C>::ns; q2++) {
        const int li = base + (Slot<C>::d[q2][2] * hx + Slot<C>::d[q2][0]) * hy + Slot<C>::d[q2][1];
        const T xv = sx[Slot<C>::d[q2][3] * nh + li];
        acc = RES ? msub(acc, s[q2], xv) : mac(acc, s[q2], xv);
      }
      o[C] = acc;
    };
    row(std::integral_constant<int, 0>{});
    row(std::integral_constant<int, 1>{});
    row(std::integral_constant<int, 2>{});
    row(std::integral_constant<int, 3>{});
    reinterpret_cast<V *>(y)[n] = V{o[0], o[1], o[2], o[3]};
  }
}

template <typename T>
__device__ __forceinline__ void bc_chain_one(const BcChain &ch, int i, const T *__restrict__ b, T *x, int mode) {
  const long o = (long)i * ch.dm;
  const int r0 = ch.r[o];
  if ((mode == 1 && (r0 & 3) == 0) || (mode == 2 && (r0 & 3) != 0)) return;
  const int jf = ch.j[i];
  T v = (jf >= 0) ? x[jf] : T(0);
  for (int t = ch.len[i] - 1; t >= 0; t--) v = b[ch.r[o + t]] + (T)ch.k[o + t] * v;
  x[r0] = v;
}

template <typename T>
__global__ void k_bc_chain(BcChain ch, const T *__restrict__ b, T *x, int mode) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ch.nbc) bc_chain_one<T>(ch, i, b, x, mode);
}

// Whole-level smoother for small levels: one workgroup holds the level's
// (P,vx,vy,vz) in LDS in the global layout and runs all sweeps there, each
// sweep the phases vx(red,black), vy(red,black), vz(red,black)[, P] over the
// whole level, then the boundary chains.  It is the tile-coloured sweep with
// one tile covering the level.
// divisors of the whole-level index decode: Y, X, ceil(Y/2)
struct LevelDiv {
  FDiv dy, dx, dhty;
};

// Velocity rows of a small level in FP32, computed once per viscosity field:
// coef[((C-1)*18 + q)*N + n] = slot q of the C row at node n (q = 1..17),
// slot 0 = the reciprocal of the diagonal the FP32 update multiplies by.
__global__ void k_level_coef(Geo<float> g, const unsigned char *__restrict__ flags, float *coef) {
  const int N = g.XY * g.Z;
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  int m1, m2, m3;
  decode(n, g.X, g.Y, g.XY, m1, m2, m3);
  float s[18];
  for (int C = 1; C <= 3; C++) {
    for (int q = 0; q < 18; q++) s[q] = 0.0f;
    if (kind_of(flags[n], C) == KIND_EQ) {
      if (C == 1) row_coef_c<float, 1>(g, m1, m2, m3, s);
      else if (C == 2) row_coef_c<float, 2>(g, m1, m2, m3, s);
      else row_coef_c<float, 3>(g, m1, m2, m3, s);
      s[0] = __builtin_amdgcn_rcpf(s[1]);
    }
    for (int q = 0; q < 18; q++) coef[(long)((C - 1) * 18 + q) * N + n] = s[q];
  }
}

template <typename T, int C>
__device__ __forceinline__ void level_phase(const Geo<T> &g, const unsigned char *__restrict__ flags,
                                            const T *__restrict__ b, T *sx, int colour, T v0koef, T p0koef,
                                            const LevelDiv &ld, const float *__restrict__ coef) {
  const int N = g.XY * g.Z;
  const int hty = (g.Y + 1) / 2;
  const int ncand = (C == 0) ? g.XY * g.Z : hty * g.X * g.Z;
  for (int i = threadIdx.x; i < ncand; i += blockDim.x) {
    int m1, m2, m3;
    if (C == 0) {
      const int q = (int)fdiv(i, ld.dy);
      m2 = i - q * g.Y; m3 = (int)fdiv(q, ld.dx); m1 = q - m3 * g.X;
    } else {
      const int q = (int)fdiv(i, ld.dhty);
      const int kk = i - q * hty;
      m3 = (int)fdiv(q, ld.dx); m1 = q - m3 * g.X;
      m2 = 2 * kk + ((colour - m1 - m3) & 1);
      if (m2 >= g.Y) continue;
    }
    const long n = g.nid(m1, m2, m3);
    if (kind_of(flags[n], C) != KIND_EQ) continue;
    T s[18];
    if constexpr (std::is_same<T, float>::value && C != 0) {
      if (coef) {
        // precomputed rows: slot 0 holds 1/diagonal
#pragma unroll
        for (int q = 0; q <= Slot<C>::ns; q++) s[q] = coef[(long)((C - 1) * 18 + q) * N + n];
        T ival = -b[4 * n + C];
#pragma unroll
        for (int q = 1; q <= Slot<C>::ns; q++) {
          const long nb = n + (long)Slot<C>::d[q][2] * g.XY + (long)Slot<C>::d[q][0] * g.Y + Slot<C>::d[q][1];
          ival = mac(ival, s[q], sx[Slot<C>::d[q][3] * N + nb]);
        }
        sx[C * N + n] -= ival * s[0] * v0koef;
        continue;
      }
    }
    row_coef_c<T, C>(g, m1, m2, m3, s);
    T ival = -b[4 * n + C];
#pragma unroll
    for (int q = 1; q <= Slot<C>::ns; q++) {
      const long nb = n + (long)Slot<C>::d[q][2] * g.XY + (long)Slot<C>::d[q][0] * g.Y + Slot<C>::d[q][1];
      ival = mac(ival, s[q], sx[Slot<C>::d[q][3] * N + nb]);
    }
    if (C == 0) sx[n] -= ival * g.ec[n] * p0koef;
    else sx[C * N + n] -= diag_step(ival, s[1]) * v0koef;
  }
}

template <typename T>
__global__ void __launch_bounds__(1024) k_smooth_level(Geo<T> g, const unsigned char *__restrict__ flags, T *x,
                                                       const T *__restrict__ b, int sweeps, int withp, T v0koef,
                                                       T p0koef, BcChain ch, LevelDiv ld,
                                                       const float *__restrict__ coef) {
  extern __shared__ unsigned char smem_raw[];
  T *sx = reinterpret_cast<T *>(smem_raw);  // component-major: sx[c * N + node]
  const int N = g.XY * g.Z, n4 = 4 * N;
  for (int i = threadIdx.x; i < n4; i += blockDim.x) sx[(i & 3) * N + (i >> 2)] = x[i];
  __syncthreads();
  const int mode = withp ? 0 : 1;
  for (int s = 0; s < sweeps; s++) {
    for (int col = 0; col < 2; col++) { level_phase<T, 1>(g, flags, b, sx, col, v0koef, p0koef, ld, coef); __syncthreads(); }
    for (int col = 0; col < 2; col++) { level_phase<T, 2>(g, flags, b, sx, col, v0koef, p0koef, ld, coef); __syncthreads(); }
    for (int col = 0; col < 2; col++) { level_phase<T, 3>(g, flags, b, sx, col, v0koef, p0koef, ld, coef); __syncthreads(); }
    if (withp) { level_phase<T, 0>(g, flags, b, sx, 0, v0koef, p0koef, ld, coef); __syncthreads(); }
    for (int i = threadIdx.x; i < ch.nbc; i += blockDim.x) {
      const long o = (long)i * ch.dm;
      const int r0 = ch.r[o];
      if ((mode == 1 && (r0 & 3) == 0) || (mode == 2 && (r0 & 3) != 0)) continue;
      const int jf = ch.j[i];
      T v = (jf >= 0) ? sx[(jf & 3) * N + (jf >> 2)] : T(0);
      for (int t = ch.len[i] - 1; t >= 0; t--) v = b[ch.r[o + t]] + (T)ch.k[o + t] * v;
      sx[(r0 & 3) * N + (r0 >> 2)] = v;
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < n4; i += blockDim.x) x[i] = sx[(i & 3) * N + (i >> 2)];
}

// Whole-level FP32 velocity smoother (the V-cycle's small levels) when each
// thread owns at most one candidate per colour: rows come from k_level_coef,
// and the rows and right-hand side a phase needs are loaded during the phase
// before it, so a phase costs its LDS reads and one barrier.  The boundary
// chains are composed once to x_r = c + K x_j and kept in LDS after the
// unknowns (4 words per chain).
template <int C>
__device__ __forceinline__ void lrow_load(const float *__restrict__ coef, const float *__restrict__ b, int N, int n,
                                          int f, float s[18], float &bv) {
  if (n < 0 || kind_of((unsigned char)f, C) != KIND_EQ) return;
#pragma unroll
  for (int q = 0; q <= Slot<C>::ns; q++) s[q] = coef[(long)((C - 1) * 18 + q) * N + n];
  bv = b[4 * n + C];
}

// n: this thread's candidate node for the phase, or -1; f: its flags
template <int C>
__device__ __forceinline__ void lrow_update(const Geo<float> &g, float *sx, int N, int n, int f, const float s[18],
                                            float bv, float v0koef) {
  if (n < 0 || kind_of((unsigned char)f, C) != KIND_EQ) return;
  float ival = -bv;
#pragma unroll
  for (int q = 1; q <= Slot<C>::ns; q++) {
    const int nb = n + Slot<C>::d[q][2] * g.XY + Slot<C>::d[q][0] * g.Y + Slot<C>::d[q][1];
    ival = mac(ival, s[q], sx[Slot<C>::d[q][3] * N + nb]);
  }
  sx[C * N + n] -= ival * s[0] * v0koef;
}

__global__ void __launch_bounds__(1024) k_smooth_level_pc(Geo<float> g, const unsigned char *__restrict__ flags,
                                                          float *x, const float *__restrict__ b, int sweeps,
                                                          float v0koef, BcChain ch, LevelDiv ld,
                                                          const float *__restrict__ coef) {
  extern __shared__ unsigned char smem_raw[];
  float *sx = reinterpret_cast<float *>(smem_raw);  // component-major: sx[c * N + node]
  const int N = g.XY * g.Z, n4 = 4 * N;
  int *cr = reinterpret_cast<int *>(sx + n4);       // chain target in sx (-1: pressure row)
  int *cj = cr + ch.nbc;                            // partner in sx
  float *cc = reinterpret_cast<float *>(cj + ch.nbc), *ck = cc + ch.nbc;
  for (int i = threadIdx.x; i < n4; i += blockDim.x) sx[(i & 3) * N + (i >> 2)] = x[i];
  for (int i = threadIdx.x; i < ch.nbc; i += blockDim.x) {
    const long o = (long)i * ch.dm;
    const int r0 = ch.r[o], jf = ch.j[i];
    float c = 0.0f, K = jf >= 0 ? 1.0f : 0.0f;
    for (int u = ch.len[i] - 1; u >= 0; u--) {
      const float k = (float)ch.k[o + u];
      c = b[ch.r[o + u]] + k * c;
      K = k * K;
    }
    cr[i] = (r0 & 3) == 0 ? -1 : (r0 & 3) * N + (r0 >> 2);
    cj[i] = jf >= 0 ? (jf & 3) * N + (jf >> 2) : 0;
    cc[i] = c; ck[i] = K;
  }
  // this thread's candidate node of each colour
  const int hty = (g.Y + 1) / 2, ncand = hty * g.X * g.Z;
  int nc0 = -1, nc1 = -1;
  if ((int)threadIdx.x < ncand) {
    const int i = threadIdx.x;
    const int q = (int)fdiv(i, ld.dhty), kk = i - q * hty;
    const int m3 = (int)fdiv(q, ld.dx), m1 = q - m3 * g.X;
    const int p0 = (0 - m1 - m3) & 1;
    const int m2a = 2 * kk + p0, m2b = 2 * kk + (1 - p0);
    if (m2a < g.Y) nc0 = (int)g.nid(m1, m2a, m3);
    if (m2b < g.Y) nc1 = (int)g.nid(m1, m2b, m3);
  }
  const int f0 = nc0 >= 0 ? flags[nc0] : 0, f1 = nc1 >= 0 ? flags[nc1] : 0;
  float sa[18], sb[18], ba = 0.0f, bb = 0.0f;
#pragma unroll
  for (int q = 0; q < 18; q++) { sa[q] = 0.0f; sb[q] = 0.0f; }
  lrow_load<1>(coef, b, N, nc0, f0, sa, ba);
  __syncthreads();
  for (int s = 0; s < sweeps; s++) {
    // keep the 6 x 17 neighbour addresses from being hoisted out of the
    // sweep loop into registers (an empty asm that may change the nodes)
    asm volatile("" : "+v"(nc0), "+v"(nc1));
    lrow_load<1>(coef, b, N, nc1, f1, sb, bb); lrow_update<1>(g, sx, N, nc0, f0, sa, ba, v0koef); __syncthreads();
    lrow_load<2>(coef, b, N, nc0, f0, sa, ba); lrow_update<1>(g, sx, N, nc1, f1, sb, bb, v0koef); __syncthreads();
    lrow_load<2>(coef, b, N, nc1, f1, sb, bb); lrow_update<2>(g, sx, N, nc0, f0, sa, ba, v0koef); __syncthreads();
    lrow_load<3>(coef, b, N, nc0, f0, sa, ba); lrow_update<2>(g, sx, N, nc1, f1, sb, bb, v0koef); __syncthreads();
    lrow_load<3>(coef, b, N, nc1, f1, sb, bb); lrow_update<3>(g, sx, N, nc0, f0, sa, ba, v0koef); __syncthreads();
    lrow_load<1>(coef, b, N, nc0, f0, sa, ba); lrow_update<3>(g, sx, N, nc1, f1, sb, bb, v0koef); __syncthreads();
    for (int i = threadIdx.x; i < ch.nbc; i += blockDim.x)
      if (cr[i] >= 0) sx[cr[i]] = cc[i] + ck[i] * sx[cj[i]];
    __syncthreads();
  }
  for (int i = threadIdx.x; i < n4; i += blockDim.x) x[i] = sx[(i & 3) * N + (i >> 2)];
}

// boundary-row update x_r -= (x_r - k x_j - b_r), one dependency depth
// mode 0: all boundary rows, 1: velocity rows only, 2: pressure rows only
template <typename T>
__global__ void k_bc_update(int beg, int end, const int *__restrict__ br, const int *__restrict__ bj,
                            const double *__restrict__ negk, const T *__restrict__ b, T *x, int mode) {
  int i = beg + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= end) return;
  int r = br[i];
  if ((mode == 1 && (r & 3) == 0) || (mode == 2 && (r & 3) != 0)) return;
  T ival = -b[r];
  ival += T(1) * x[r];
  if (bj[i] >= 0) ival += (T)negk[i] * x[bj[i]];
  x[r] -= ival;
}

// ---------------------------------------------------------------------------
// Transfers in gather form.  Per axis and staggering kind a fine index i has
// a coarse base lo(i) and weights for lo and lo+1 (the clamped index rules of
// vprestrict/vpprolong, move3mg.c:3563-3576, 4015-4028, ...).  For the
// restriction each coarse index keeps the list of (fine index, weight) pairs
// in increasing fine order, so the sums run in the reference's
// single-thread scatter order (fine nodes z, x, y).
// ---------------------------------------------------------------------------
struct Ptrs16 {
  const double *p[16];
};

constexpr int kPad = 6;  // padded gather-list width (2:1 hierarchies need at most 5)
struct AxisTab {
  const int *lo;        // [nf]
  const double *w64;    // [2*nf] weights for lo, lo+1
  const float *w32;
  const int *cbeg;      // [nc+1]
  const int *cfi;       // fine index of each entry
  const double *cw64;   // weight of each entry
  const float *cw32;
  const int *pfi;       // [kPad*nc] the same lists padded to kPad entries (fine index -1: none)
  const double *pw64;
  const float *pw32;
};

struct CompTab {
  AxisTab ax[3];  // x, y, z
};

template <typename T> __device__ __forceinline__ T tw(const double *a, const float *b, long i);
template <> __device__ __forceinline__ double tw<double>(const double *a, const float *, long i) { return a[i]; }
template <> __device__ __forceinline__ float tw<float>(const double *, const float *b, long i) { return b[i]; }

// vprestrict: coarse b from fine residual r (parallel/move3mg.c:3966-4354);
// one thread per coarse unknown, summed over its fine contributors in the
// reference's order.  The four components of a node run on neighbouring
// lanes through one code path (the component picks the axis tables).
template <typename T>
__global__ void k_restrict(int Xf, int Yf, int Zf, int Xc, int Yc, int Zc, CompTab t0, CompTab t1,
                           CompTab t2, CompTab t3, const unsigned char *__restrict__ ff,
                           const unsigned char *__restrict__ fc, const T *__restrict__ rf,
                           const float *__restrict__ ndf, const float *__restrict__ ndc, T *bc, T *xc,
                           int withp) {
  const long u = xcd_block() * blockDim.x + threadIdx.x;
  const long n = u >> 2;
  const int XYc = Xc * Yc, XYf = Xf * Yf;
  if (n >= (long)XYc * Zc) return;
  const int c = (int)(u & 3);
  xc[u] = T(0);
  if (kind_of(fc[n], c) != KIND_EQ || (!withp && c == 0)) { bc[u] = T(0); return; }
  int cx, cy, cz;
  decode(n, Xc, Yc, XYc, cx, cy, cz);
  const CompTab &t = (c == 0) ? t0 : (c == 1) ? t1 : (c == 2) ? t2 : t3;
  T acc = T(0), ws = T(0);
  for (int a = t.ax[2].cbeg[cz]; a < t.ax[2].cbeg[cz + 1]; a++) {
    const int k = t.ax[2].cfi[a];
    const T wz = tw<T>(t.ax[2].cw64, t.ax[2].cw32, a);
    for (int bq = t.ax[0].cbeg[cx]; bq < t.ax[0].cbeg[cx + 1]; bq++) {
      const int i = t.ax[0].cfi[bq];
      const T wx = tw<T>(t.ax[0].cw64, t.ax[0].cw32, bq);
      for (int e = t.ax[1].cbeg[cy]; e < t.ax[1].cbeg[cy + 1]; e++) {
        const int j = t.ax[1].cfi[e];
        const long nf = (long)k * XYf + (long)i * Yf + j;
        if (kind_of(ff[nf], c) != KIND_EQ) continue;
        const T wy = tw<T>(t.ax[1].cw64, t.ax[1].cw32, e);
        const T swt = wx * wy * wz;
        T iv;
        if (c == 0) iv = rf[4 * nf] * (T)ndf[nf] / (T)ndc[n] * swt;
        else iv = rf[4 * nf + c] * swt;
        acc += iv;
        ws += swt;
      }
    }
  }
  bc[u] = (ws != T(0)) ? acc / ws : acc;
}

// The same sums from the padded lists when no coarse index has more than K
// fine entries per axis (K = 4 on the fine levels of a 2:1 hierarchy, 5 near
// the coarsest): the KxKxK candidate loop has fixed trip counts and its loads
// are independent of each other.  Padding and non-equation candidates add an
// exact zero, so the sums and their order are those of k_restrict.
template <typename T, int K>
__global__ void __launch_bounds__(256) k_restrictK(int Xf, int Yf, int Zf, int Xc, int Yc, int Zc, CompTab t0,
                                                   CompTab t1, CompTab t2, CompTab t3,
                                                   const unsigned char *__restrict__ ff,
                                                   const unsigned char *__restrict__ fc, const T *__restrict__ rf,
                                                   const float *__restrict__ ndf, const float *__restrict__ ndc,
                                                   T *bc, T *xc, int withp) {
  const long u = xcd_block() * blockDim.x + threadIdx.x;
  const long n = u >> 2;
  const int XYc = Xc * Yc, XYf = Xf * Yf;
  if (n >= (long)XYc * Zc) return;
  const int c = (int)(u & 3);
  xc[u] = T(0);
  if (kind_of(fc[n], c) != KIND_EQ || (!withp && c == 0)) { bc[u] = T(0); return; }
  int cx, cy, cz;
  decode(n, Xc, Yc, XYc, cx, cy, cz);
  const CompTab &t = (c == 0) ? t0 : (c == 1) ? t1 : (c == 2) ? t2 : t3;
  long oz[K], ox[K];
  int jy[K];
  T wz[K], wx[K], wy[K];
#pragma unroll
  for (int q = 0; q < K; q++) {
    const int k = t.ax[2].pfi[kPad * cz + q], i = t.ax[0].pfi[kPad * cx + q];
    jy[q] = t.ax[1].pfi[kPad * cy + q];
    oz[q] = k < 0 ? -1 : (long)k * XYf;
    ox[q] = i < 0 ? -1 : (long)i * Yf;
    wz[q] = tw<T>(t.ax[2].pw64, t.ax[2].pw32, kPad * cz + q);
    wx[q] = tw<T>(t.ax[0].pw64, t.ax[0].pw32, kPad * cx + q);
    wy[q] = tw<T>(t.ax[1].pw64, t.ax[1].pw32, kPad * cy + q);
  }
  const T ndn = (c == 0) ? (T)ndc[n] : T(1);
  T acc = T(0), ws = T(0);
#pragma unroll
  for (int a = 0; a < K; a++)
#pragma unroll
    for (int b = 0; b < K; b++)
#pragma unroll
      for (int e = 0; e < K; e++) {
        const bool ok = oz[a] >= 0 && ox[b] >= 0 && jy[e] >= 0;
        const long nf = ok ? oz[a] + ox[b] + jy[e] : 0;
        const bool eq = ok && kind_of(ff[nf], c) == KIND_EQ;
        const T swt = wx[b] * wy[e] * wz[a];
        T iv;
        if (c == 0) iv = rf[4 * nf] * (T)ndf[nf] / ndn * swt;
        else iv = rf[4 * nf + c] * swt;
        acc += eq ? iv : T(0);
        ws += eq ? swt : T(0);
      }
  bc[u] = (ws != T(0)) ? acc / ws : acc;
}

// Separable FP32 restriction of the V-cycle's velocity rows.  A coarse
// unknown's candidates are a tensor product of per-axis lists, so its sum
// over the equation-row fine residuals, sum_z wz sum_x wx sum_y wy r, runs one
// axis at a time: y (k_rsep_y, which also drops the fine rows that are not
// equations), x (k_rsep_x), then z with the division by the weight sum
// (k_rsep_z).  The weight sums come from the same passes with an indicator
// in place of r (mode 1), once per prepare.  Sums are rounded in this order,
// not the reference's, which the FP32 path does not promise.
struct SepTab {
  const int *fi;    // [kPad * nc] padded fine indices (-1: none)
  const float *w;
};
struct SepComp {
  SepTab ax[3];     // x, y, z
};
struct SepComps {
  SepComp c[3];     // vx, vy, vz
};

// out[(k * Xf + i) * Yc + cy] = (0, sum_y for vx, vy, vz)
__global__ void k_rsep_y(int Xf, int Yf, int Zf, int Yc, SepComps t, const unsigned char *__restrict__ ff,
                         const float *__restrict__ rf, float4 *__restrict__ out, int mode) {
  const long n = xcd_block() * blockDim.x + threadIdx.x;
  if (n >= (long)Zf * Xf * Yc) return;
  const int cy = (int)(n % Yc);
  const long q = n / Yc;
  const long row = q * Yf;  // (k * Xf + i) * Yf
  float o[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const SepTab &ty = t.c[c].ax[1];
#pragma unroll
    for (int e = 0; e < kPad; e++) {
      const int j = ty.fi[kPad * cy + e];
      if (j < 0) break;
      const long nf = row + j;
      const bool eq = kind_of(ff[nf], c + 1) == KIND_EQ;
      const float v = mode ? 1.0f : rf[4 * nf + c + 1];
      o[c] += eq ? ty.w[kPad * cy + e] * v : 0.0f;
    }
  }
  out[n] = make_float4(0.0f, o[0], o[1], o[2]);
}

// out[(k * Xc + cx) * Yc + cy] = sum over x of in[(k * Xf + i) * Yc + cy]
__global__ void k_rsep_x(int Xf, int Xc, int Yc, int Zf, SepComps t, const float4 *__restrict__ in,
                         float4 *__restrict__ out) {
  const long n = xcd_block() * blockDim.x + threadIdx.x;
  if (n >= (long)Zf * Xc * Yc) return;
  const int cy = (int)(n % Yc);
  const long q = n / Yc;
  const int cx = (int)(q % Xc);
  const long k = q / Xc;
  float o[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const SepTab &tx = t.c[c].ax[0];
#pragma unroll
    for (int e = 0; e < kPad; e++) {
      const int i = tx.fi[kPad * cx + e];
      if (i < 0) break;
      const float4 v = in[(k * Xf + i) * Yc + cy];
      o[c] += tx.w[kPad * cx + e] * (c == 0 ? v.y : c == 1 ? v.z : v.w);
    }
  }
  out[n] = make_float4(0.0f, o[0], o[1], o[2]);
}

// mode 1: ws[n] = the sums (weight sums); mode 0: bc = sum / ws on the
// coarse equation rows (0 elsewhere and on pressure rows), xc = 0
__global__ void k_rsep_z(int Xc, int Yc, int Zc, SepComps t, const float4 *__restrict__ in,
                         const unsigned char *__restrict__ fc, float4 *ws, float *bc, float *xc, int mode) {
  const long n = xcd_block() * blockDim.x + threadIdx.x;
  const long XYc = (long)Xc * Yc;
  if (n >= XYc * Zc) return;
  const long cz = n / XYc, rem = n - cz * XYc;  // rem = cx * Yc + cy
  float o[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const SepTab &tz = t.c[c].ax[2];
#pragma unroll
    for (int e = 0; e < kPad; e++) {
      const int k = tz.fi[kPad * cz + e];
      if (k < 0) break;
      const float4 v = in[k * XYc + rem];
      o[c] += tz.w[kPad * cz + e] * (c == 0 ? v.y : c == 1 ? v.z : v.w);
    }
  }
  if (mode) {
    ws[n] = make_float4(0.0f, o[0], o[1], o[2]);
    return;
  }
  const float4 s = ws[n];
  const unsigned char f = fc[n];
  const float sw[3] = {s.y, s.z, s.w};
  float r[3];
#pragma unroll
  for (int c = 0; c < 3; c++)
    r[c] = kind_of(f, c + 1) == KIND_EQ ? (sw[c] != 0.0f ? o[c] / sw[c] : o[c]) : 0.0f;
  reinterpret_cast<float4 *>(bc)[n] = make_float4(0.0f, r[0], r[1], r[2]);
  reinterpret_cast<float4 *>(xc)[n] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// vpprolong: fine x += sum coarse x * w * koef (parallel/move3mg.c:3535-3850),
// one thread per fine node and its four unknowns
template <typename T, int C>
__device__ __forceinline__ T prolong_one(const CompTab &t, int m1, int m2, int m3, int XYc, int Yc,
                                         const T *__restrict__ xc, T v, T k) {
  const int lx = t.ax[0].lo[m1], ly = t.ax[1].lo[m2], lz = t.ax[2].lo[m3];
  for (int dz = 0; dz < 2; dz++) {
    const T wz = tw<T>(t.ax[2].w64, t.ax[2].w32, 2 * m3 + dz);
    for (int dx = 0; dx < 2; dx++) {
      const T wx = tw<T>(t.ax[0].w64, t.ax[0].w32, 2 * m1 + dx);
      for (int dy = 0; dy < 2; dy++) {
        const T wy = tw<T>(t.ax[1].w64, t.ax[1].w32, 2 * m2 + dy);
        const int nc = (lz + dz) * XYc + (lx + dx) * Yc + (ly + dy);  // coarse levels stay below 2^29 nodes
        const T swt = wx * wy * wz;
        v += xc[4 * nc + C] * swt * k;
      }
    }
  }
  return v;
}

// (the node's kinds, its four fine values and the result in one load or store each)
template <typename T>
__global__ void k_prolong(int Xf, int Yf, int Zf, int Xc, int Yc, int Zc, CompTab t0, CompTab t1,
                          CompTab t2, CompTab t3, const unsigned char *__restrict__ ff,
                          const T *__restrict__ xc, T *xf, T pk, T vk, int withp, int ynreset) {
  using V = typename V4<T>::t;
  const long n = xcd_block() * blockDim.x + threadIdx.x;
  const int XYc = Xc * Yc, XYf = Xf * Yf;
  if (n >= (long)XYf * Zf) return;
  int m1, m2, m3;
  decode(n, Xf, Yf, XYf, m1, m2, m3);
  const unsigned char f = ff[n];
  V x = reinterpret_cast<const V *>(xf)[n];
  if (withp && kind_of(f, 0) != KIND_NONE) x.x = prolong_one<T, 0>(t0, m1, m2, m3, XYc, Yc, xc, ynreset ? T(0) : x.x, pk);
  if (kind_of(f, 1) != KIND_NONE) x.y = prolong_one<T, 1>(t1, m1, m2, m3, XYc, Yc, xc, ynreset ? T(0) : x.y, vk);
  if (kind_of(f, 2) != KIND_NONE) x.z = prolong_one<T, 2>(t2, m1, m2, m3, XYc, Yc, xc, ynreset ? T(0) : x.z, vk);
  if (kind_of(f, 3) != KIND_NONE) x.w = prolong_one<T, 3>(t3, m1, m2, m3, XYc, Yc, xc, ynreset ? T(0) : x.w, vk);
  reinterpret_cast<V *>(xf)[n] = x;
}

// ronurestrict: coarse nu/ro from fine nodes (parallel/move3mg.c:3387-3528),
// FP32 accumulators and FP64 weights as in the reference
__global__ void k_ronurestrict(int Xf, int Yf, int Xc, int Yc, int Zc, CompTab t,
                               const float *__restrict__ nuf, const float *__restrict__ rof,
                               float *nuc, float *roc, int vm) {
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int XYc = Xc * Yc, XYf = Xf * Yf;
  if (n >= (long)XYc * Zc) return;
  int cx, cy, cz;
  decode(n, Xc, Yc, XYc, cx, cy, cz);
  float anu = 0.f, aro = 0.f;
  double ws = 0.0;
  for (int a = t.ax[2].cbeg[cz]; a < t.ax[2].cbeg[cz + 1]; a++) {
    int k = t.ax[2].cfi[a];
    double wz = t.ax[2].cw64[a];
    for (int bq = t.ax[0].cbeg[cx]; bq < t.ax[0].cbeg[cx + 1]; bq++) {
      int i = t.ax[0].cfi[bq];
      double wx = t.ax[0].cw64[bq];
      for (int e = t.ax[1].cbeg[cy]; e < t.ax[1].cbeg[cy + 1]; e++) {
        int j = t.ax[1].cfi[e];
        double wy = t.ax[1].cw64[e];
        long nf = (long)k * XYf + (long)i * Yf + j;
        double swt = wx * wy * wz, iv = 0.0;
        if (vm == 0) iv = nuf[nf] * swt;
        if (vm == 1) iv = log((double)nuf[nf]) * swt;
        if (vm == 2) iv = 1.0 / nuf[nf] * swt;
        anu = (float)((double)anu + iv);
        iv = rof[nf] * swt;
        aro = (float)((double)aro + iv);
        ws += swt;
      }
    }
  }
  float nv = (float)((double)anu / ws);
  if (vm == 1) nv = (float)exp((double)nv);
  if (vm == 2) nv = (float)(1.0 / nv);
  nuc[n] = nv;
  roc[n] = (float)((double)aro / ws);
}

// ---------------------------------------------------------------------------
// Coarsest level: dense operator rows (velocity block or all rows), then a
// Gauss-Jordan inverse with partial pivoting (first maximal |a| wins).
// ---------------------------------------------------------------------------
__global__ void k_dense_rows(Geo<double> g, const unsigned char *__restrict__ flags, int nbc,
                             const int *__restrict__ br, const int *__restrict__ bj,
                             const double *__restrict__ negk, double *A, int n4, int withp) {
  long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (u >= n4) return;
  long n = u >> 2;
  int c = (int)(u & 3);
  double *row = A + u * (long)n4;
  int k = kind_of(flags[n], c);
  if (k == KIND_NONE || (!withp && c == 0)) { row[u] += 1.0; return; }
  if (k == KIND_BC) {
    for (int i = 0; i < nbc; i++)
      if (br[i] == u) {
        row[u] += 1.0;
        if (bj[i] >= 0) row[bj[i]] += negk[i];
        break;
      }
    return;
  }
  int m1, m2, m3;
  decode(n, g.X, g.Y, g.XY, m1, m2, m3);
  double s[18];
  int ns = row_coef<double>(g, c, m1, m2, m3, s);
  for (int q = 1; q <= ns; q++) {
    const int cq = kSlot[c][q][3];
    long uq = 4 * unode(g, m1 + kSlot[c][q][0], m2 + kSlot[c][q][1], m3 + kSlot[c][q][2], cq) + cq;
    row[uq] += s[q];
  }
}

// Gauss-Jordan inverse with partial pivoting and no row swaps, two light
// launches per column k.  k_gj_select: the pivot p is the unused row with
// the largest |A[i][k]| (lowest index on ties), from the per-row candidates
// cand[i] the previous elimination left (step 0 reads column 0); it records
// p_k and d_k = A[p][k].  k_gj_elim: one thread per element of rows i != p
// applies A[i][j>k] -= f A[p][j] and B[i][:] -= f B[p][:], f = A[i][k] / d_k,
// and the thread of column k+1 leaves cand[i] = |A[i][k+1]| (-1 for used
// rows).  Row p is only read in its step.  Afterwards row k of the inverse is B[p_k][:] / d_k (k_gj_final).
__device__ __forceinline__ void gj_better(double v, int i, double &bv, int &bi) {
  if (v > bv || (v == bv && i < bi)) { bv = v; bi = i; }
}

__global__ void __launch_bounds__(1024) k_gj_select(const double *__restrict__ A, int n, int k, int *used, int *prow,
                                                    double *pval, const double *__restrict__ cand, int *sing) {
  __shared__ double rv[1024];
  __shared__ int ri[1024];
  double best = -1.0;
  int bidx = n;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double v = (k == 0) ? fabs(A[(long)i * n]) : cand[i];
    gj_better(v, i, best, bidx);
  }
  rv[threadIdx.x] = best;
  ri[threadIdx.x] = bidx;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) gj_better(rv[threadIdx.x + s], ri[threadIdx.x + s], rv[threadIdx.x], ri[threadIdx.x]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int p = ri[0];
    if (!(rv[0] > 0.0) || p >= n) {
      if (!*sing) *sing = k + 1;
      prow[k] = -1;
      return;
    }
    prow[k] = p;
    pval[k] = A[(long)p * n + k];
    used[p] = 1;
  }
}

__global__ void k_gj_elim(double *A, double *B, int n, int k, const int *__restrict__ used,
                          const int *__restrict__ prow, const double *__restrict__ pval, double *cand) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long nn = (long)n * n;
  if (t >= 2 * nn) return;
  const int p = prow[k];
  if (p < 0) return;  // singular: no further updates
  const bool inA = t < nn;
  const long e = inA ? t : t - nn;
  const int i = (int)(e / n), j = (int)(e - (long)i * n);
  if (inA && j <= k) return;  // columns up to k of A are no longer needed
  if (i == p) {
    if (inA && j == k + 1) cand[i] = -1.0;
    return;
  }
  const double aik = A[(long)i * n + k];  // column k is not written in this step
  if (aik == 0.0) {  // nothing to eliminate in this row (most rows: the matrix is sparse)
    if (inA && j == k + 1) cand[i] = used[i] ? -1.0 : fabs(A[e]);
    return;
  }
  const double f = aik / pval[k];
  if (inA) {
    const double v = A[e] - f * A[(long)p * n + j];
    A[e] = v;
    if (j == k + 1) cand[i] = used[i] ? -1.0 : fabs(v);
  } else {
    B[e] -= f * B[(long)p * n + j];
  }
}

// Blocked form of the same elimination (same pivot rule): a panel of kb
// columns is eliminated in one workgroup (k_gjb_panel, the panel in LDS),
// recording the factors F[i][kk] = A[i][k]/d_k (0 on the pivot row); the
// panel's kb steps then reach the rest of [A | B] as one rank-kb update,
// T -= F R'', where R'' are the pivot rows as each step saw them:
// R''[kk] = T[p_kk] - sum_{q<kk} F[p_kk][q] R''[q] (k_gjb_rows).  Rounding
// differs from the column-by-column order; the pivots, d_k and the final
// scaling are those of k_gj_final.
constexpr int kGjbMax = 16;  // panel width limit
__global__ void __launch_bounds__(1024) k_gjb_panel(const double *__restrict__ A, int n, int k0, int kb, int *used,
                                                    int *prow, double *pval, double *F, int *sing) {
  extern __shared__ unsigned char smem_raw[];
  double *P = reinterpret_cast<double *>(smem_raw);  // P[kk * n + i]: column-major, no bank conflicts
  unsigned char *su = smem_raw + (size_t)n * kb * 8;  // used rows, kept in LDS for the panel
  __shared__ double rv[16];
  __shared__ int ri[16];
  __shared__ int piv;
  if (*sing) return;  // an earlier panel found no pivot
  for (int e = threadIdx.x; e < n * kb; e += blockDim.x) {
    const int i = e / kb, kk = e - i * kb;
    P[kk * n + i] = A[(long)i * n + k0 + kk];
  }
  for (int i = threadIdx.x; i < n; i += blockDim.x) su[i] = (unsigned char)used[i];
  __syncthreads();
  for (int kk = 0; kk < kb; kk++) {
    double best = -1.0;
    int bidx = n;
    for (int i = threadIdx.x; i < n; i += blockDim.x)
      if (!su[i]) gj_better(fabs(P[kk * n + i]), i, best, bidx);
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(best, off);
      const int oi = __shfl_xor(bidx, off);
      gj_better(ov, oi, best, bidx);
    }
    if ((threadIdx.x & 63) == 0) { rv[threadIdx.x >> 6] = best; ri[threadIdx.x >> 6] = bidx; }
    __syncthreads();
    if (threadIdx.x == 0) {
      double bv = -1.0;
      int bi = n;
      for (int w = 0; w < (int)(blockDim.x >> 6); w++) gj_better(rv[w], ri[w], bv, bi);
      if (!(bv > 0.0) || bi >= n) {
        if (!*sing) *sing = k0 + kk + 1;
        prow[k0 + kk] = -1;
        piv = -1;
      } else {
        piv = bi;
        prow[k0 + kk] = bi;
        pval[k0 + kk] = P[kk * n + bi];
        su[bi] = 1;
        used[bi] = 1;
      }
    }
    __syncthreads();
    const int p = piv;
    if (p < 0) return;  // singular: every thread leaves here together
    const double d = P[kk * n + p];
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const double f = (i == p) ? 0.0 : P[kk * n + i] / d;
      F[(long)i * kGjbMax + kk] = f;
      if (f != 0.0)
        for (int q = kk + 1; q < kb; q++) P[q * n + i] -= f * P[q * n + p];
    }
    __syncthreads();
  }
}

// column c of the trailing part T = [A[:, k0+kb ..) | B]
__device__ __forceinline__ double *gjb_col(double *A, double *B, int n, int k0, int kb, int c, long i) {
  const int na = n - k0 - kb;
  return c < na ? A + i * n + k0 + kb + c : B + i * n + (c - na);
}

// R''[kk][c] for every trailing column c (one thread per column)
__global__ void k_gjb_rows(double *A, double *B, int n, int k0, int kb, const int *__restrict__ prow,
                           const double *__restrict__ F, double *R, const int *__restrict__ sing) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = 2 * n - k0 - kb;
  if (c >= m || *sing) return;  // after a failed pivot some prow are -1
  double r[kGjbMax];
#pragma unroll
  for (int kk = 0; kk < kGjbMax; kk++) {
    if (kk >= kb) break;
    const int p = prow[k0 + kk];
    double v = *gjb_col(A, B, n, k0, kb, c, p);
    for (int q = 0; q < kk; q++) v -= F[(long)p * kGjbMax + q] * r[q];
    r[kk] = v;
    R[(long)kk * m + c] = v;
  }
}

// T[i][c] -= sum_kk F[i][kk] R''[kk][c]: a workgroup takes 256 columns and
// kGjbRows rows, each thread one column with its R'' values in registers
constexpr int kGjbRows = 8;
__global__ void __launch_bounds__(256) k_gjb_update(double *A, double *B, int n, int k0, int kb,
                                                    const double *__restrict__ F, const double *__restrict__ R,
                                                    const int *__restrict__ sing) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = 2 * n - k0 - kb;
  if (c >= m || *sing) return;
  double r[kGjbMax];
#pragma unroll
  for (int kk = 0; kk < kGjbMax; kk++) r[kk] = kk < kb ? R[(long)kk * m + c] : 0.0;
  const int i0 = blockIdx.y * kGjbRows;
  for (int i = i0; i < i0 + kGjbRows && i < n; i++) {
    double acc = 0.0;
    bool any = false;
#pragma unroll
    for (int kk = 0; kk < kGjbMax; kk++) {
      const double f = kk < kb ? F[(long)i * kGjbMax + kk] : 0.0;
      any = any || f != 0.0;
      acc += f * r[kk];
    }
    if (any) *gjb_col(A, B, n, k0, kb, c, i) -= acc;
  }
}

// inv[k][:] = B[p_k][:] / A[p_k][k]
__global__ void k_gj_final(const double *__restrict__ B, int n, const int *__restrict__ prow,
                           const double *__restrict__ pval, double *inv) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)n * n) return;
  const int k = (int)(t / n), j = (int)(t % n);
  inv[t] = B[(long)prow[k] * n + j] / pval[k];
}

// B = identity
__global__ void k_eye(double *B, int n) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)n * n) return;
  B[t] = (t / n == t % n) ? 1.0 : 0.0;
}

// Compact dense rows of the coarsest velocity block: row i is unknown act[i]
// (velocity unknowns with an equation or a boundary row; pressure and absent
// unknowns are left out), columns mapped through idx (-1: left out)
__global__ void k_dense_rows_c(Geo<double> g, const unsigned char *__restrict__ flags, int nbc,
                               const int *__restrict__ br, const int *__restrict__ bj,
                               const double *__restrict__ negk, const int *__restrict__ act,
                               const int *__restrict__ idx, double *A, int N) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int u = act[i];
  const long n = u >> 2;
  const int c = u & 3;
  double *row = A + (long)i * N;
  if (kind_of(flags[n], c) == KIND_BC) {
    for (int q = 0; q < nbc; q++)
      if (br[q] == u) {
        row[i] += 1.0;
        if (bj[q] >= 0 && idx[bj[q]] >= 0) row[idx[bj[q]]] += negk[q];
        break;
      }
    return;
  }
  int m1, m2, m3;
  decode(n, g.X, g.Y, g.XY, m1, m2, m3);
  double s[18];
  const int ns = row_coef<double>(g, c, m1, m2, m3, s);
  for (int q = 1; q <= ns; q++) {
    const int cq = kSlot[c][q][3];
    const long uq = 4 * unode(g, m1 + kSlot[c][q][0], m2 + kSlot[c][q][1], m3 + kSlot[c][q][2], cq) + cq;
    const int j = idx[uq];
    if (j >= 0) row[j] += s[q];
  }
}

// x[act[i]] = sum_j M[i][j] b[act[j]], one wavefront per row
template <typename TM, typename T>
__global__ void k_gemv_c(const TM *__restrict__ M, const int *__restrict__ act, const T *__restrict__ b, T *x, int N) {
  const int row = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (row >= N) return;
  const TM *mr = M + (long)row * N;
  double acc = 0.0;
  for (int j = lane; j < N; j += 64) acc += (double)mr[j] * (double)b[act[j]];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane == 0) x[act[row]] = (T)acc;
}

// y = M x for a dense row-major n x n matrix, one wavefront per row
template <typename TM, typename T>
__global__ void k_gemv(const TM *__restrict__ M, const T *__restrict__ x, T *y, int n) {
  int row = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  int lane = threadIdx.x & 63;
  if (row >= n) return;
  const TM *mr = M + (long)row * n;
  double acc = 0.0;
  for (int j = lane; j < n; j += 64) acc += (double)mr[j] * (double)x[j];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane == 0) y[row] = (T)acc;
}

// ---------------------------------------------------------------------------
// Preconditioner pieces and vector kernels of the FP64 outer iteration.
// ---------------------------------------------------------------------------
// z_p = omega * eta_cell * r_p on continuity rows, zero elsewhere
__global__ void k_pscale(long n4, const unsigned char *__restrict__ flags, const double *__restrict__ ec,
                         const double *__restrict__ r, double *z, double omega) {
  long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (u >= n4) return;
  long n = u >> 2;
  int c = (int)(u & 3);
  z[u] = (c == 0 && kind_of(flags[n], 0) == KIND_EQ) ? omega * ec[n] * r[u] : 0.0;
}

// velocity right-hand side r_v - G z_p of the preconditioner (momentum rows
// have P only in slots 16/17), boundary rows take r; stored in type T
// (one thread per node: its four entries in one vector store)
template <typename T>
__global__ void k_vrhs(Geo<double> g, const unsigned char *__restrict__ flags, const double *__restrict__ r,
                       const double *__restrict__ zp, T *bv) {
  using V = typename V4<T>::t;
  const long n = xcd_block() * blockDim.x + threadIdx.x;
  if (n >= (long)g.XY * g.Z) return;
  const unsigned char f = flags[n];
  const double4 rn = reinterpret_cast<const double4 *>(r)[n];
  const double rr[4] = {rn.x, rn.y, rn.z, rn.w};
  int m1, m2, m3;
  decode(n, g.X, g.Y, g.XY, m1, m2, m3);
  T o[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
  for (int c = 1; c < 4; c++) {
    const int k = kind_of(f, c);
    if (k == KIND_BC) o[c] = (T)rr[c];
    if (k != KIND_EQ) continue;
    const double kf = (c == 1) ? 2.0 * g.iBx[m1] : (c == 2) ? 2.0 * g.iBy[m2] : 2.0 * g.iBz[m3];
    double acc = rr[c];
#pragma unroll
    for (int q = 16; q <= 17; q++) {
      const long uq = 4 * unode(g, m1 + kSlot[c][q][0], m2 + kSlot[c][q][1], m3 + kSlot[c][q][2], 0);
      acc -= (q == 16 ? kf : -kf) * zp[uq];
    }
    o[c] = (T)acc;
  }
  reinterpret_cast<V *>(bv)[n] = V{o[0], o[1], o[2], o[3]};
}

// z_v = x (velocity unknowns), pressure entries of z kept
template <typename T>
__global__ void k_combine(long n4, const T *__restrict__ xv, double *z) {
  long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (u >= n4 || (u & 3) == 0) return;
  z[u] = (double)xv[u];
}

// partial sums of w2*a*b over blocks, for k vectors b_i = B[i*n4 ..]
__global__ void __launch_bounds__(256) k_multidot(long n4, const double *__restrict__ w2,
                                                  const double *__restrict__ a, Ptrs16 Bs,
                                                  int k, double *partial) {
  __shared__ double red[256];
  double acc[16];
  for (int i = 0; i < 16; i++) acc[i] = 0.0;
  for (long u = blockIdx.x * (long)blockDim.x + threadIdx.x; u < n4; u += (long)gridDim.x * blockDim.x) {
    double wa = w2[u] * a[u];
    for (int i = 0; i < k; i++) acc[i] += wa * Bs.p[i][u];
  }
  for (int i = 0; i < k; i++) {
    red[threadIdx.x] = acc[i];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[(long)i * gridDim.x + blockIdx.x] = red[0];
    __syncthreads();
  }
}

__global__ void k_reduce_rows(const double *__restrict__ partial, int nb, int k, double *out) {
  __shared__ double red[256];
  int i = blockIdx.x;
  if (i >= k) return;
  double acc = 0.0;
  for (int j = threadIdx.x; j < nb; j += blockDim.x) acc += partial[(long)i * nb + j];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[i] = red[0];
}

// ---------------------------------------------------------------------------
// Krylov basis kernels: one pass over the vectors for all k directions
// ---------------------------------------------------------------------------
constexpr int kMaxDirs = 257;  // restart + 1
// pointers to the Krylov vectors (basis or preconditioned), stored as VT
template <typename VT>
struct PtrsKT {
  const VT *p[kMaxDirs];
};
using PtrsK = PtrsKT<double>;
template <typename VT> struct Pair2;
template <> struct Pair2<double> { using t = double2; };
template <> struct Pair2<float> { using t = float2; };

// partial[i*gridDim + block] = sum over the block's slices of w2 * a * V_i
__global__ void __launch_bounds__(256) k_mdot(long n4, const double *__restrict__ w2, const double *__restrict__ a,
                                              PtrsK V, int k, double *partial) {
  __shared__ double acc[kMaxDirs][4];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int i = threadIdx.x; i < kMaxDirs * 4; i += blockDim.x) acc[i >> 2][i & 3] = 0.0;
  __syncthreads();
  constexpr int E = 8;
  for (long base = (long)blockIdx.x * 256 * E; base < n4; base += (long)gridDim.x * 256 * E) {
    double wa[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
      const long u = base + threadIdx.x + 256L * e;
      wa[e] = (u < n4) ? w2[u] * a[u] : 0.0;
    }
    for (int i = 0; i < k; i++) {
      const double *__restrict__ v = V.p[i];
      double sum = 0.0;
#pragma unroll
      for (int e = 0; e < E; e++) {
        const long u = base + threadIdx.x + 256L * e;
        if (u < n4) sum += wa[e] * v[u];
      }
      for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
      if (lane == 0) acc[i][wave] += sum;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < k; i += blockDim.x)
    partial[(long)i * gridDim.x + blockIdx.x] = (acc[i][0] + acc[i][1]) + (acc[i][2] + acc[i][3]);
}

// out = a - sum_i coef_i V_i (out may alias a); partial[block] = sum of w2 out^2
template <typename VT>
__global__ void __launch_bounds__(256) k_maxpy_norm(long n4, int k, const double *__restrict__ coef, PtrsKT<VT> V,
                                                    const double *a, double *out, const double *__restrict__ w2,
                                                    double *partial) {
  __shared__ double sc[kMaxDirs];
  __shared__ double red[4];
  for (int i = threadIdx.x; i < k; i += blockDim.x) sc[i] = coef[i];
  __syncthreads();
  double ss = 0.0;
  for (long u = (long)blockIdx.x * blockDim.x + threadIdx.x; u < n4; u += (long)gridDim.x * blockDim.x) {
    double v = a[u];
    for (int i = 0; i < k; i++) v -= sc[i] * (double)V.p[i][u];
    out[u] = v;
    ss += w2[u] * v * v;
  }
  for (int off = 32; off > 0; off >>= 1) ss += __shfl_xor(ss, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ss;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// Delayed classical Gram-Schmidt with reorthogonalisation (one pass for the
// dots, one for the updates per Arnoldi step).  Vectors are read as double2
// (n4 is a multiple of 4).  Pass 1: partial sums of a_i = <V_i,u>,
// b_i = <V_i,w> (i < k), <u,w>, <u,u> in the W2 inner product.
template <typename VT>
__global__ void __launch_bounds__(256) k_dots2(long n4, const double *__restrict__ w2, const double *__restrict__ u,
                                               const double *__restrict__ w, PtrsKT<VT> V, int k, double *partial) {
  using P2 = typename Pair2<VT>::t;
  __shared__ double acc[2 * kMaxDirs + 2][4];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nd = 2 * k + 2;
  for (int i = threadIdx.x; i < (2 * kMaxDirs + 2) * 4; i += blockDim.x) acc[i >> 2][i & 3] = 0.0;
  __syncthreads();
  constexpr int E = 4;  // double2 per thread per slice; two basis vectors per step
  const long n2 = n4 >> 1;
  const double2 *__restrict__ W2 = reinterpret_cast<const double2 *>(w2);
  const double2 *__restrict__ U = reinterpret_cast<const double2 *>(u);
  const double2 *__restrict__ Wv = reinterpret_cast<const double2 *>(w);
  for (long base = (long)blockIdx.x * 256 * E; base < n2; base += (long)gridDim.x * 256 * E) {
    double2 wu[E], ww[E];
    double suw = 0.0, suu = 0.0;
#pragma unroll
    for (int e = 0; e < E; e++) {
      const long p = base + threadIdx.x + 256L * e;
      double2 q = make_double2(0.0, 0.0), uu = q, wv = q;
      if (p < n2) { q = W2[p]; uu = U[p]; wv = Wv[p]; }
      wu[e] = make_double2(q.x * uu.x, q.y * uu.y);
      ww[e] = make_double2(q.x * wv.x, q.y * wv.y);
      suw += wu[e].x * wv.x + wu[e].y * wv.y;
      suu += wu[e].x * uu.x + wu[e].y * uu.y;
    }
    for (int off = 32; off > 0; off >>= 1) {
      suw += __shfl_xor(suw, off);
      suu += __shfl_xor(suu, off);
    }
    if (lane == 0) { acc[2 * k][wave] += suw; acc[2 * k + 1][wave] += suu; }
    for (int i = 0; i < k; i += 2) {
      const bool two = i + 1 < k;
      const P2 *__restrict__ v0 = reinterpret_cast<const P2 *>(V.p[i]);
      const P2 *__restrict__ v1 = reinterpret_cast<const P2 *>(V.p[two ? i + 1 : i]);
      double2 a0[E], a1[E];
#pragma unroll
      for (int e = 0; e < E; e++) {
        const long p = base + threadIdx.x + 256L * e;
        a0[e] = make_double2(0.0, 0.0);
        a1[e] = a0[e];
        if (p < n2) {
          const P2 q0 = v0[p], q1 = v1[p];
          a0[e] = make_double2((double)q0.x, (double)q0.y);
          a1[e] = make_double2((double)q1.x, (double)q1.y);
        }
      }
      double sa0 = 0.0, sb0 = 0.0, sa1 = 0.0, sb1 = 0.0;
#pragma unroll
      for (int e = 0; e < E; e++) {
        sa0 += wu[e].x * a0[e].x + wu[e].y * a0[e].y;
        sb0 += ww[e].x * a0[e].x + ww[e].y * a0[e].y;
        sa1 += wu[e].x * a1[e].x + wu[e].y * a1[e].y;
        sb1 += ww[e].x * a1[e].x + ww[e].y * a1[e].y;
      }
      for (int off = 32; off > 0; off >>= 1) {
        sa0 += __shfl_xor(sa0, off);
        sb0 += __shfl_xor(sb0, off);
        sa1 += __shfl_xor(sa1, off);
        sb1 += __shfl_xor(sb1, off);
      }
      if (lane == 0) {
        acc[i][wave] += sa0; acc[k + i][wave] += sb0;
        if (two) { acc[i + 1][wave] += sa1; acc[k + i + 1][wave] += sb1; }
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nd; i += blockDim.x)
    partial[(long)i * gridDim.x + blockIdx.x] = (acc[i][0] + acc[i][1]) + (acc[i][2] + acc[i][3]);
}

// Pass 2: u <- (u - sum_i a_i V_i) * su (the reorthogonalised, normalised
// basis vector v_k), w <- w - sum_i h_i V_i - h_k v_k, partial sums of w2 w^2.
// coef = [a_0..a_{k-1}, h_0..h_{k-1}, h_k, su]
// With a VT = float basis, the new basis vector v_k is stored in vout
// (rounded) and w subtracts h_k times the stored value, so the Arnoldi
// relation holds for the basis as stored; u keeps its FP64 value.
template <typename VT>
__global__ void __launch_bounds__(256) k_update2(long n4, int k, const double *__restrict__ coef, PtrsKT<VT> V,
                                                 double *u, double *w, const double *__restrict__ w2, double *partial,
                                                 VT *vout) {
  using P2 = typename Pair2<VT>::t;
  __shared__ double sc[2 * kMaxDirs + 2];
  __shared__ double red[4];
  for (int i = threadIdx.x; i < 2 * k + 2; i += blockDim.x) sc[i] = coef[i];
  __syncthreads();
  const double hk = sc[2 * k], su = sc[2 * k + 1];
  const long n2 = n4 >> 1;
  double2 *U = reinterpret_cast<double2 *>(u);
  double2 *Wv = reinterpret_cast<double2 *>(w);
  const double2 *__restrict__ W2 = reinterpret_cast<const double2 *>(w2);
  double ss = 0.0;
  // E double2 per thread per slice, basis vectors two at a time (their loads
  // in flight together); each element subtracts the terms in basis order
  constexpr int E = 4;
  for (long base = (long)blockIdx.x * 256 * E; base < n2; base += (long)gridDim.x * 256 * E) {
    double2 uu[E], wv[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
      const long p = base + threadIdx.x + 256L * e;
      uu[e] = make_double2(0.0, 0.0);
      wv[e] = uu[e];
      if (p < n2) { uu[e] = U[p]; wv[e] = Wv[p]; }
    }
    for (int i = 0; i < k; i += 2) {
      const bool two = i + 1 < k;
      const P2 *__restrict__ v0 = reinterpret_cast<const P2 *>(V.p[i]);
      const P2 *__restrict__ v1 = reinterpret_cast<const P2 *>(V.p[two ? i + 1 : i]);
      double2 a0[E], a1[E];
#pragma unroll
      for (int e = 0; e < E; e++) {
        const long p = base + threadIdx.x + 256L * e;
        a0[e] = make_double2(0.0, 0.0);
        a1[e] = a0[e];
        if (p < n2) {
          const P2 q0 = v0[p], q1 = v1[p];
          a0[e] = make_double2((double)q0.x, (double)q0.y);
          a1[e] = make_double2((double)q1.x, (double)q1.y);
        }
      }
      const double c0 = sc[i], d0 = sc[k + i];
#pragma unroll
      for (int e = 0; e < E; e++) {
        uu[e].x -= c0 * a0[e].x; uu[e].y -= c0 * a0[e].y;
        wv[e].x -= d0 * a0[e].x; wv[e].y -= d0 * a0[e].y;
      }
      if (two) {
        const double c1 = sc[i + 1], d1 = sc[k + i + 1];
#pragma unroll
        for (int e = 0; e < E; e++) {
          uu[e].x -= c1 * a1[e].x; uu[e].y -= c1 * a1[e].y;
          wv[e].x -= d1 * a1[e].x; wv[e].y -= d1 * a1[e].y;
        }
      }
    }
#pragma unroll
    for (int e = 0; e < E; e++) {
      const long p = base + threadIdx.x + 256L * e;
      if (p >= n2) continue;
      uu[e].x *= su;
      uu[e].y *= su;
      if constexpr (std::is_same<VT, float>::value) {
        const float2 vs = make_float2((float)uu[e].x, (float)uu[e].y);
        reinterpret_cast<float2 *>(vout)[p] = vs;
        wv[e].x -= hk * (double)vs.x;
        wv[e].y -= hk * (double)vs.y;
      } else {
        wv[e].x -= hk * uu[e].x;
        wv[e].y -= hk * uu[e].y;
      }
      U[p] = uu[e];
      Wv[p] = wv[e];
      const double2 q = W2[p];
      ss += q.x * wv[e].x * wv[e].x + q.y * wv[e].y * wv[e].y;
    }
  }
  for (int off = 32; off > 0; off >>= 1) ss += __shfl_xor(ss, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ss;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// ---------------------------------------------------------------------------
// The same two passes for an FP32 basis, four basis vectors per step and
// 16-byte loads (float4 of the basis, double4 of u, w, W2): each thread
// keeps E element quads of W2 u and W2 w in registers across the k basis
// vectors.  The eight dots of a step (4 vectors x {u, w}) are reduced over
// the wave by halving: at each of the first three exchanges a lane keeps
// half of its sums and adds the partner's copy of them (xor 32, 16, 8), so
// 10 shuffles give the wave total of quantity lane >> 3 in every lane.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum8(const double v[8], int lane) {
  const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8;
  double k4[4], k2[2];
#pragma unroll
  for (int j = 0; j < 4; j++) k4[j] = (b5 ? v[4 + j] : v[j]) + __shfl_xor(b5 ? v[j] : v[4 + j], 32);
#pragma unroll
  for (int j = 0; j < 2; j++) k2[j] = (b4 ? k4[2 + j] : k4[j]) + __shfl_xor(b4 ? k4[j] : k4[2 + j], 16);
  double k1 = (b3 ? k2[1] : k2[0]) + __shfl_xor(b3 ? k2[0] : k2[1], 8);
  k1 += __shfl_xor(k1, 4);
  k1 += __shfl_xor(k1, 2);
  k1 += __shfl_xor(k1, 1);
  return k1;
}

// a basis vector's 16 bytes, loaded non-temporally: each pass streams the
// FP32 basis once (j vectors of 4 n0 values, far beyond the caches), and
// keeping it out of the caches leaves them to u, w and W2 (3 % of a 257^3
// solve)
typedef float f4v_nt __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld_basis(const float4 *p) {
  const f4v_nt v = __builtin_nontemporal_load(reinterpret_cast<const f4v_nt *>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}

constexpr int kGsE = 4;  // element quads per thread per slice (4: 1.4 % faster 257^3 solves than 2)

// partial[i * gridDim + block], i < 2k+2: <V_i, W2 u> (i < k), <V_i, W2 w>
// (k + i), <u, W2 w>, <u, W2 u>
__global__ void __launch_bounds__(256) k_dots2f(long n4, const double *__restrict__ w2, const double *__restrict__ u,
                                                const double *__restrict__ w, PtrsKT<float> V, int k,
                                                double *partial) {
  __shared__ double acc[2 * kMaxDirs + 2][4];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nd = 2 * k + 2;
  for (int i = threadIdx.x; i < (2 * kMaxDirs + 2) * 4; i += blockDim.x) acc[i >> 2][i & 3] = 0.0;
  __syncthreads();
  constexpr int E = kGsE;
  const long nq = n4 >> 2;
  const double4 *__restrict__ W2 = reinterpret_cast<const double4 *>(w2);
  const double4 *__restrict__ U = reinterpret_cast<const double4 *>(u);
  const double4 *__restrict__ Wv = reinterpret_cast<const double4 *>(w);
  for (long base = (long)blockIdx.x * 256 * E; base < nq; base += (long)gridDim.x * 256 * E) {
    double wu[E][4], ww[E][4];
    double suw = 0.0, suu = 0.0;
#pragma unroll
    for (int e = 0; e < E; e++) {
      const long p = base + threadIdx.x + 256L * e;
      double4 q = make_double4(0.0, 0.0, 0.0, 0.0), uu = q, wv = q;
      if (p < nq) { q = W2[p]; uu = U[p]; wv = Wv[p]; }
      wu[e][0] = q.x * uu.x; wu[e][1] = q.y * uu.y; wu[e][2] = q.z * uu.z; wu[e][3] = q.w * uu.w;
      ww[e][0] = q.x * wv.x; ww[e][1] = q.y * wv.y; ww[e][2] = q.z * wv.z; ww[e][3] = q.w * wv.w;
      suw += wu[e][0] * wv.x + wu[e][1] * wv.y + wu[e][2] * wv.z + wu[e][3] * wv.w;
      suu += wu[e][0] * uu.x + wu[e][1] * uu.y + wu[e][2] * uu.z + wu[e][3] * uu.w;
    }
    for (int off = 32; off > 0; off >>= 1) {
      suw += __shfl_xor(suw, off);
      suu += __shfl_xor(suu, off);
    }
    if (lane == 0) { acc[2 * k][wave] += suw; acc[2 * k + 1][wave] += suu; }
    for (int i = 0; i < k; i += 4) {
      float4 a[4][E];
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const float4 *__restrict__ vt = reinterpret_cast<const float4 *>(V.p[min(i + t, k - 1)]);
#pragma unroll
        for (int e = 0; e < E; e++) {
          const long p = base + threadIdx.x + 256L * e;
          a[t][e] = p < nq ? ld_basis(vt + p) : make_float4(0.f, 0.f, 0.f, 0.f);
        }
      }
      double sv[8];
#pragma unroll
      for (int t = 0; t < 4; t++) {
        double su = 0.0, sw = 0.0;
#pragma unroll
        for (int e = 0; e < E; e++) {
          const double ax = a[t][e].x, ay = a[t][e].y, az = a[t][e].z, aw = a[t][e].w;
          su += wu[e][0] * ax + wu[e][1] * ay + wu[e][2] * az + wu[e][3] * aw;
          sw += ww[e][0] * ax + ww[e][1] * ay + ww[e][2] * az + ww[e][3] * aw;
        }
        sv[2 * t] = su;
        sv[2 * t + 1] = sw;
      }
      const double tot = wave_sum8(sv, lane);
      if ((lane & 7) == 0) {
        const int q = lane >> 3, t = q >> 1;
        if (i + t < k) acc[(q & 1) ? k + i + t : i + t][wave] += tot;
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nd; i += blockDim.x)
    partial[(long)i * gridDim.x + blockIdx.x] = (acc[i][0] + acc[i][1]) + (acc[i][2] + acc[i][3]);
}

// u <- (u - sum_i a_i V_i) su stored rounded in vout (and in u with keep_u),
// w <- w - sum_i b_i V_i - h_k vout, partial sums of W2 w^2 (coef = [a, b,
// h_k, su] as k_update2); each element subtracts the terms in basis order
__global__ void __launch_bounds__(256) k_update2f(long n4, int k, const double *__restrict__ coef, PtrsKT<float> V,
                                                  double *u, double *w, const double *__restrict__ w2,
                                                  double *partial, float *vout, int keep_u) {
  __shared__ double sc[2 * kMaxDirs + 2];
  __shared__ double red[4];
  for (int i = threadIdx.x; i < 2 * k + 2; i += blockDim.x) sc[i] = coef[i];
  __syncthreads();
  const double hk = sc[2 * k], su = sc[2 * k + 1];
  constexpr int E = kGsE;
  const long nq = n4 >> 2;
  double4 *U = reinterpret_cast<double4 *>(u);
  double4 *Wv = reinterpret_cast<double4 *>(w);
  const double4 *__restrict__ W2 = reinterpret_cast<const double4 *>(w2);
  double ss = 0.0;
  for (long base = (long)blockIdx.x * 256 * E; base < nq; base += (long)gridDim.x * 256 * E) {
    double uu[E][4], wv[E][4];
#pragma unroll
    for (int e = 0; e < E; e++) {
      const long p = base + threadIdx.x + 256L * e;
      double4 a = make_double4(0.0, 0.0, 0.0, 0.0), b = a;
      if (p < nq) { a = U[p]; b = Wv[p]; }
      uu[e][0] = a.x; uu[e][1] = a.y; uu[e][2] = a.z; uu[e][3] = a.w;
      wv[e][0] = b.x; wv[e][1] = b.y; wv[e][2] = b.z; wv[e][3] = b.w;
    }
    for (int i = 0; i < k; i += 4) {
      const int nt = min(4, k - i);
      float4 a[4][E];
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const float4 *__restrict__ vt = reinterpret_cast<const float4 *>(V.p[min(i + t, k - 1)]);
#pragma unroll
        for (int e = 0; e < E; e++) {
          const long p = base + threadIdx.x + 256L * e;
          a[t][e] = p < nq ? ld_basis(vt + p) : make_float4(0.f, 0.f, 0.f, 0.f);
        }
      }
#pragma unroll
      for (int t = 0; t < 4; t++) {
        if (t >= nt) break;
        const double c = sc[i + t], d = sc[k + i + t];
#pragma unroll
        for (int e = 0; e < E; e++) {
          const double av[4] = {a[t][e].x, a[t][e].y, a[t][e].z, a[t][e].w};
#pragma unroll
          for (int q = 0; q < 4; q++) {
            uu[e][q] -= c * av[q];
            wv[e][q] -= d * av[q];
          }
        }
      }
    }
#pragma unroll
    for (int e = 0; e < E; e++) {
      const long p = base + threadIdx.x + 256L * e;
      if (p >= nq) continue;
      float vs[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        uu[e][q] *= su;
        vs[q] = (float)uu[e][q];
        wv[e][q] -= hk * (double)vs[q];
      }
      reinterpret_cast<float4 *>(vout)[p] = make_float4(vs[0], vs[1], vs[2], vs[3]);
      if (keep_u) U[p] = make_double4(uu[e][0], uu[e][1], uu[e][2], uu[e][3]);
      Wv[p] = make_double4(wv[e][0], wv[e][1], wv[e][2], wv[e][3]);
      const double4 q = W2[p];
      ss += q.x * wv[e][0] * wv[e][0] + q.y * wv[e][1] * wv[e][1] + q.z * wv[e][2] * wv[e][2] +
            q.w * wv[e][3] * wv[e][3];
    }
  }
  for (int off = 32; off > 0; off >>= 1) ss += __shfl_xor(ss, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ss;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// x = v (a fill kernel instead of a memset node, so the preconditioner can be
// captured in a HIP graph)
template <typename T>
__global__ void k_fill(long n, T v, T *x) {
  long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (u < n) x[u] = v;
}

// y = s * x
template <typename A, typename B>
__global__ void k_cast(long n, const A *__restrict__ x, B *__restrict__ y) {
  long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (u < n) y[u] = (B)x[u];
}

// z rounded to FP32: zs = (float)z, and z = (double)zs, so that products
// formed from z and from its stored FP32 copy agree
// The preconditioner's last pass in mixed precision with FP32 storage: the
// velocity slots from the FP32 V-cycle, the pressure slots (already in z)
// rounded to FP32, each value stored in z and in the FP32 copy zs -- what
// k_combine followed by k_round32 leave, in one pass.
__global__ void k_combine_round(long n4, const float *__restrict__ xv, double *z, float *zs) {
  long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (u >= n4) return;
  const float f = (u & 3) == 0 ? (float)z[u] : xv[u];
  zs[u] = f;
  z[u] = (double)f;
}

__global__ void k_round32(long n, double *z, float *zs) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float f = (float)z[i];
  zs[i] = f;
  z[i] = (double)f;
}

__global__ void k_scale_to(long n4, double s, const double *__restrict__ x, double *y) {
  long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (u < n4) y[u] = s * x[u];
}

// a -= sum_i c_i A_i ; b -= sum_i c_i B_i
__global__ void k_multiaxpy(long n4, int k, const double *__restrict__ coef, Ptrs16 As, Ptrs16 Bs,
                            double *a, double *b) {
  long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (u >= n4) return;
  double va = a[u], vb = b[u];
  for (int i = 0; i < k; i++) {
    double c = coef[i];
    va -= c * As.p[i][u];
    vb -= c * Bs.p[i][u];
  }
  a[u] = va;
  b[u] = vb;
}

__global__ void k_scale2(long n4, const double *__restrict__ sp, double *a, double *b) {
  long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (u >= n4) return;
  double s = *sp;
  a[u] *= s;
  b[u] *= s;
}

// GCR scalars from <q,q>, <q,r>: s = 1/sqrt(qq), alpha = qr * s
__global__ void k_gcr_scalars(const double *qqr, double *s_alpha) {
  double s = 1.0 / sqrt(qqr[0]);
  s_alpha[0] = s;
  s_alpha[1] = qqr[1] * s;
}

// x += alpha z ; r -= alpha q
__global__ void k_update(long n4, const double *__restrict__ alpha, const double *__restrict__ z,
                         const double *__restrict__ q, double *x, double *r) {
  long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (u >= n4) return;
  double a = *alpha;
  x[u] += a * z[u];
  r[u] -= a * q[u];
}

__global__ void k_sub(long n4, const double *__restrict__ a, const double *__restrict__ b, double *c) {
  long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (u < n4) c[u] = a[u] - b[u];
}

// residual blocks: sum w2 r^2 over momentum / continuity / boundary rows and w2 b^2
__global__ void __launch_bounds__(256) k_blocks(long n4, const unsigned char *__restrict__ flags,
                                                const double *__restrict__ w2, const double *__restrict__ r,
                                                const double *__restrict__ b, double *partial) {
  __shared__ double red[256];
  double acc[4] = {0, 0, 0, 0};
  for (long u = blockIdx.x * (long)blockDim.x + threadIdx.x; u < n4; u += (long)gridDim.x * blockDim.x) {
    int c = (int)(u & 3);
    int k = kind_of(flags[u >> 2], c);
    double rr = w2[u] * r[u] * r[u];
    if (k == KIND_BC) acc[2] += rr;
    else if (k == KIND_EQ) { if (c == 0) acc[1] += rr; else acc[0] += rr; }
    acc[3] += w2[u] * b[u] * b[u];
  }
  for (int i = 0; i < 4; i++) {
    red[threadIdx.x] = acc[i];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[(long)i * gridDim.x + blockIdx.x] = red[0];
    __syncthreads();
  }
}

// W^2 of the relative-residual norm: 1 on momentum and boundary rows,
// (eta_cell / h_cell)^2 on continuity rows, 0 on ghost unknowns
__global__ void k_weights(int X, int Y, int Z, const unsigned char *__restrict__ flags,
                          const double *__restrict__ ec, const double *__restrict__ gx,
                          const double *__restrict__ gy, const double *__restrict__ gz, double *w2) {
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int XY = X * Y;
  if (n >= (long)XY * Z) return;
  int m1, m2, m3;
  decode(n, X, Y, XY, m1, m2, m3);
  unsigned char f = flags[n];
  for (int c = 0; c < 4; c++) {
    int k = kind_of(f, c);
    double w = (k == KIND_NONE) ? 0.0 : 1.0;
    if (c == 0 && k == KIND_EQ) {
      double h = ((gx[m1] - gx[m1 - 1]) + (gy[m2] - gy[m2 - 1]) + (gz[m3] - gz[m3 - 1])) / 3.0;
      w = ec[n] / h;
    }
    w2[4 * n + c] = w * w;
  }
}

template <typename TS, typename TD>
__global__ void k_convert(long n, const TS *__restrict__ a, TD *b) {
  long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (u < n) b[u] = (TD)a[u];
}

// plain device copy: the known-size kernel that HBM byte counters are scaled by
__global__ void __launch_bounds__(256) k_stream_copy(long n, const double *__restrict__ a, double *__restrict__ b) {
  long u = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (u < n) b[u] = a[u];
}

__global__ void k_pshift(long nn, const unsigned char *__restrict__ flags, double *x, const double *ref,
                         double pinit) {
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= nn) return;
  if (kind_of(flags[n], 0) != KIND_NONE) x[4 * n] += pinit - *ref;
}

// Hydrostatic initial guess: velocities zero, P integrated column by column
// through the z-momentum rows (with v = 0 a z row reads
// zkf*(P(k) - P(k+1)) = b), the way the reference seeds its first pressure
// field (parallel/in3mg.c:927-938).  One thread per (m1, m2) column.
__global__ void k_hydro(Geo<double> g, const unsigned char *__restrict__ flags, const double *__restrict__ b,
                        double *x) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.X * g.Y) return;
  int m1 = t / g.Y, m2 = t - m1 * g.Y;
  for (int m3 = 0; m3 < g.Z; m3++) {
    long n = g.nid(m1, m2, m3);
    for (int c = 1; c < 4; c++) x[4 * n + c] = 0.0;
  }
  if (m1 + 1 >= g.X || m2 + 1 >= g.Y) return;
  for (int m3 = 0; m3 + 1 < g.Z; m3++) {
    long n = g.nid(m1, m2, m3);
    if (kind_of(flags[n], 3) != KIND_EQ) continue;
    double zkf = 2.0 * g.iBz[m3];
    long plo = 4 * g.nid(m1 + 1, m2 + 1, m3), phi = 4 * g.nid(m1 + 1, m2 + 1, m3 + 1);
    x[phi] = x[plo] - b[4 * n + 3] / zkf;
  }
}

}  // namespace stk
