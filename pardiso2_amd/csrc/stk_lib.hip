// stk_lib.hip -- host side of the MI355X Stokes multigrid library: level
// hierarchy, boundary tables, transfer tables, multigrid cycle, FP64 outer
// iteration and the C ABI declared in include/pardiso2_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <array>
#include <atomic>
#include <memory>
#include <mutex>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/pardiso2_amd.h"
#include "stk_kernels.h"
#include "stk_advect.h"
#include "stk_compat.h"
#include "stk_heat.h"
#include "stk_markers.h"
#include "stk_post.h"
#include "stk_schur.h"

#include <hipcub/hipcub.hpp>

using namespace stk;

#define MAXLV 12

namespace {

#define HIPCHK(x)                                  \
  do {                                             \
    hipError_t e_ = (x);                           \
    if (e_ != hipSuccess) {                        \
      fprintf(stderr, "pardiso2_amd: HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
      return STK_EHIP;                             \
    }                                              \
  } while (0)

// Graph capture and the runtime calls that can wait on other streams
// (device allocation and release, pinned host memory, stream, event and
// graph lifetime) exclude each other across host threads: a release that
// synchronises the device while another thread's stream is capturing breaks
// that capture.  Recursive, so a holder may allocate.
std::recursive_mutex g_rt_mu;
using RtLock = std::lock_guard<std::recursive_mutex>;

template <typename T>
int dalloc(T **p, size_t n) {
  RtLock lk(g_rt_mu);
  *p = nullptr;
  if (n == 0) n = 1;
  if (hipMalloc((void **)p, n * sizeof(T)) != hipSuccess) return STK_ENOMEM;
  return STK_OK;
}

template <typename T>
void dfree(T *&p) {
  RtLock lk(g_rt_mu);
  if (p) (void)hipFree(p);
  p = nullptr;
}

inline unsigned blocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// one axis of a transfer: fine->coarse index rule and gather lists
struct AxisHost {
  std::vector<int> lo;
  std::vector<double> w;  // 2 per fine index
  std::vector<int> cbeg, cfi;
  std::vector<double> cw;
  int kmax = 0;           // longest gather list
  std::vector<int> pfi;   // lists padded to kPad entries (if kmax <= kPad)
  std::vector<double> pw;
};

struct AxisDev {
  int *lo = nullptr, *cbeg = nullptr, *cfi = nullptr, *pfi = nullptr;
  double *w64 = nullptr, *cw64 = nullptr, *pw64 = nullptr;
  float *w32 = nullptr, *cw32 = nullptr, *pw32 = nullptr;
  int kmax = 0;
  AxisTab tab() const { return AxisTab{lo, w64, w32, cbeg, cfi, cw64, cw32, pfi, pw64, pw32}; }
  void release() {
    dfree(lo); dfree(cbeg); dfree(cfi); dfree(w64); dfree(cw64); dfree(w32); dfree(cw32);
    dfree(pfi); dfree(pw64); dfree(pw32);
    kmax = 0;
  }
};

struct Level {
  int X = 0, Y = 0, Z = 0, XY = 0;
  int per = 0;  // periodic axes (bit 0 x, bit 1 y, bit 2 z), the same on every level
  long n = 0, off = 0;
  std::vector<double> g[3];
  // device geometry
  double *iA64[3] = {}, *iB64[3] = {}, *g64[3] = {};
  float *iA32[3] = {}, *iB32[3] = {};
  double *e64[5] = {};
  float *e32[5] = {};
  float *lc32 = nullptr;  // FP32 velocity rows of an LDS-resident level (54 planes), from its viscosities
  // separable FP32 restriction into this level: the y and x partial sums and the weight sums
  float4 *rs_t1 = nullptr, *rs_t2 = nullptr, *rs_ws = nullptr;
  unsigned char *flags = nullptr;
  std::vector<unsigned char> hflags;
  // bounding box of the nodes with equation rows {y0, x0, z0, y1, x1, z1}: the smoother's tiles cover it
  int box[6] = {0, 0, 0, 0, 0, 0};
  // boundary rows, sorted by dependency depth
  int nbc = 0;
  std::vector<int> bcoff;
  int *bc_r = nullptr, *bc_j = nullptr;
  double *bc_negk = nullptr, *bc_c = nullptr;
  std::vector<int> hbc_r, hbc_j;
  std::vector<double> hbc_negk, hbc_c;
  // the same rows as dependency chains (one launch per pass)
  int ch_dm = 1;
  int *ch_r = nullptr, *ch_len = nullptr, *ch_j = nullptr;
  double *ch_k = nullptr;
  BcChain chain() const { return BcChain{nbc, ch_dm, ch_r, ch_len, ch_j, ch_k}; }
  std::vector<int> hch_r0, hch_j;  // head row and final partner of each chain (host)
  // chains per smoother tile of shape bt_key, for sweeps that set them in place
  int bt_key[3] = {0, 0, 0};
  bool bt_ok = false;
  int *bt_off = nullptr, *bt_idx = nullptr, *bt_head = nullptr, *bt_rl = nullptr, *bt_pl = nullptr;
  float *bt_cc32 = nullptr, *bt_ck32 = nullptr;
  double *bt_cc64 = nullptr, *bt_ck64 = nullptr;
  TileChains<float> *bt_tc32 = nullptr;
  TileChains<double> *bt_tc64 = nullptr;
  void free_tile_chains() {
    dfree(bt_off); dfree(bt_idx); dfree(bt_head); dfree(bt_rl); dfree(bt_pl); dfree(bt_cc32); dfree(bt_ck32); dfree(bt_cc64); dfree(bt_ck64);
    dfree(bt_tc32); dfree(bt_tc64);
    bt_key[0] = bt_key[1] = bt_key[2] = 0;
    bt_ok = false;
  }
  // vectors
  float *x32[2] = {}, *b32 = nullptr, *r32 = nullptr;
  double *x64[2] = {}, *b64 = nullptr, *r64 = nullptr;
  int cur32 = 0, cur64 = 0;
  // smoother tile (y, x, z, stage viscosities in LDS) picked by stk_autotune (0: default)
  int tune32[4] = {0, 0, 0, 0}, tune64[4] = {0, 0, 0, 0};
  // reference arrays (compat): sol0, fre1, fre0
  double *ref[3] = {};
  // transfer from level-1 (kinds N, C, P, R per axis)
  AxisDev ax[4][3];
  // pressure-cell operator K of the Schur approximation and its K-cycle vectors
  double *k64[7] = {};
  float *k32[7] = {};
  double *kc64 = nullptr;  // the seven planes colour-compact (Gauss-Seidel half-sweeps), or null
  float *kc32 = nullptr;
  float *kx32 = nullptr, *kb32 = nullptr, *kr32 = nullptr;
  double *kx64 = nullptr, *kb64 = nullptr, *kr64 = nullptr;
  int *klo[3] = {}, *kcnt[3] = {};  // cells of level-1 owned by each cell of this level, per axis
  int *kpar[3] = {};                // cell of level+1 owning each cell of this level, per axis
  template <typename T> Geo<T> geo(double p2) const;
  // spacing uniform along each axis (FP32 tables): the FP32 velocity sweep
  // then uses the level constants fu (set by stk_prepare for the knobs' p2koef)
  bool uni = false;
  float ua[3] = {0, 0, 0}, ub[3] = {0, 0, 0};  // the uniform iA, iB per axis (FP32)
  FluxU fu{};
  // a caller's row store of this level (gausmat4 on bound arrays): ELL rows, state, boundary rows by depth
  int *ell_cnt = nullptr, *ell_col = nullptr, *ell_bcrows = nullptr;
  double *ell_val = nullptr, *ell_x = nullptr, *ell_f = nullptr, *ell_f0 = nullptr;
  unsigned char *ell_isbc = nullptr;
  float *ell_nd = nullptr;
};

template <>
Geo<double> Level::geo<double>(double p2) const {
  Geo<double> g;
  g.X = X; g.Y = Y; g.Z = Z; g.XY = XY;
  g.iAx = iA64[0]; g.iBx = iB64[0]; g.iAy = iA64[1]; g.iBy = iB64[1]; g.iAz = iA64[2]; g.iBz = iB64[2];
  g.ec = e64[0]; g.exy = e64[1]; g.exz = e64[2]; g.eyz = e64[3]; g.ed = e64[4];
  g.p2 = p2;
  g.per = per;
  return g;
}
template <>
Geo<float> Level::geo<float>(double p2) const {
  Geo<float> g;
  g.X = X; g.Y = Y; g.Z = Z; g.XY = XY;
  g.iAx = iA32[0]; g.iBx = iB32[0]; g.iAy = iA32[1]; g.iBy = iB32[1]; g.iAz = iA32[2]; g.iBz = iB32[2];
  g.ec = e32[0]; g.exy = e32[1]; g.exz = e32[2]; g.eyz = e32[3]; g.ed = e32[4];
  g.p2 = (float)p2;
  g.per = per;
  return g;
}

// Temperature solve on levels 0..Lt of the hierarchy: per-level spacing
// tables of the heat rows, boundary-row flags and the boundary rows by
// dependency depth
struct HeatLevel {
  double *kf[3] = {}, *dm[3] = {}, *dp[3] = {};
  unsigned char *isbc = nullptr;
  int nbc = 0;
  std::vector<int> bcoff;  // depth d: rows bcoff[d] .. bcoff[d+1]-1
  int *bc_r = nullptr, *bc_j = nullptr;  // row and partner (-1: none), level-local nodes
  double *bc_negk = nullptr, *bc_k = nullptr, *bc_c = nullptr;
  bool split0 = false;  // x-periodic with an even line count: line 0 of a colour first
  void release() {
    for (int a = 0; a < 3; a++) { dfree(kf[a]); dfree(dm[a]); dfree(dp[a]); }
    dfree(isbc); dfree(bc_r); dfree(bc_j); dfree(bc_negk); dfree(bc_k); dfree(bc_c);
    nbc = 0;
    bcoff.clear();
    split0 = false;
  }
};

struct HeatState {
  HeatLevel lv[MAXLV];
  int built = -1;                       // levels 0..built have their tables
  float *ro = nullptr, *cp = nullptr, *kt = nullptr, *rcp = nullptr;  // every level
  float *tk = nullptr, *tk0 = nullptr, *tk1 = nullptr;                // level 0
  float *ht = nullptr;                  // the caller's source term (tk1 = ht + heating terms)
  float *tk2 = nullptr;                 // the temperature at the start of the last step
  double *x = nullptr, *b = nullptr, *r = nullptr;                    // sol0, fre1, fre0 of every level
  double *part = nullptr;               // block partials of the reductions
  std::vector<double> hpart;
  bool fields = false;
  int Lt = -1;                          // levels of the current sub-step (-1: none started)
  double dt = 1.0, t0 = 1.0, t1 = 1.0;
};

}  // namespace

struct stk_ctx {
  int L = 0;
  size_t lds_max = 65536;  // dynamic LDS one workgroup may use on this device
  int per[3] = {0, 0, 0};
  Level lv[MAXLV];
  long ntot = 0;
  stk_knobs kn{};
  // boundary tables in the reference layout
  std::vector<int> bondm;  // 4*ntot
  std::vector<float> bondv1;  // 2*(nbond+1)
  std::vector<int> bondn1;    // nbond+1
  // fields
  std::vector<float> hnu, hro, hnxx, hnxy, hnxz, hnyz;  // level 0 (nu/ro over all levels after prepare)
  float *nu = nullptr, *ro = nullptr;                  // all levels
  float *nxx = nullptr, *nxy = nullptr, *nxz = nullptr, *nyz = nullptr;
  float *nd = nullptr;                                 // all levels (cell viscosity)
  bool fields_set = false, prepared = false, rows_dirty = true;
  // coarsest inverse
  int nc = 0;
  double *ainv64 = nullptr, *ainvP64 = nullptr;  // velocity block (compact, nca x nca) / all rows
  float *ainv32 = nullptr;
  int *cact = nullptr;  // coarsest velocity unknowns of the compact block
  int nca = 0;
  // outer iteration vectors (finest level)
  double *X = nullptr, *B = nullptr, *R = nullptr, *W2 = nullptr, *T1 = nullptr;
  double *wu = nullptr, *ww = nullptr, *wz = nullptr;  // FP64 working u, w, z of a solve with an FP32 basis
  std::vector<double *> Zs, Qs;
  double *red = nullptr, *scal = nullptr;
  int nred = 1024;
  double *kred = nullptr, *hcol = nullptr;  // flexible GMRES: per-block dots, Hessenberg column / update coefficients
  double *hpin = nullptr;  // pinned host staging of the per-step scalars (2 kMaxDirs + 8 doubles)
  // preconditioner captured as a HIP graph: z = M^-1 gin written to gout
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool graph_ok = false;
  stk_solver_opts gkey{};
  double *gin = nullptr, *gout = nullptr;
  int cvm = -1;  // coarse-level averaging actually used by the prepared operators
  hipStream_t st = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // Schur approximation: level-0 velocity D^-1, work vectors, coarsest K inverse
  double *dinv64 = nullptr, *bfy64 = nullptr, *bfw64 = nullptr;
  float *dinv32 = nullptr, *bfy32 = nullptr, *bfw32 = nullptr;
  double *kinvT64 = nullptr;  // coarsest K inverse, transposed: [j * nkc + i] = inverse(i, j)
  float *kinvT32 = nullptr;
  int *kcell = nullptr;
  int nkc = 0;
  bool ktabs = false;
  // compat row store (gausmat2 / pardisomat)
  std::vector<double> hval1, hfre1;
  std::vector<int> hlin0, hnum0, hpos0;
  std::vector<double> ca;           // coarsest CSR values (dense accumulation)
  std::vector<double> cb;
  double *cinv = nullptr;           // inverse of ca, kept until a row changes
  bool cinv_ok = false;
  // post-solve pass: block partials of the norms, stress and strain-rate planes (6 n0 each)
  double *post_part = nullptr, *sig = nullptr, *epsr = nullptr;
  int *adv_bins = nullptr;  // cell-search bins of the three axes (marker advection)
  // temperature boundary equations in the reference's layout (bondm[nodenum8 + node],
  // parallel/in3mg.c:369-385): tbondm[ntot] (0 = none), tbv1[2*(n+1)], tbn1[n+1] (node + 1)
  std::vector<int> tbondm;
  std::vector<float> tbv1;
  std::vector<int> tbn1;
  bool heat_dirty = true;
  HeatState *heat = nullptr;
};

namespace {

stk_ctx *g_bound = nullptr;
stk_ref_arrays g_arr{};  // the caller's reference arrays (stk_bind_arrays)
bool g_arr_set = false;

// the level whose FP32 stencil viscosities a Geo names
const Level *level_of(const stk_ctx *c, const Geo<float> &g) {
  for (int l = 0; l <= c->L; l++)
    if (c->lv[l].e32[0] == g.ec) return &c->lv[l];
  return nullptr;
}
template <typename T> const Level *level_of(const stk_ctx *, const Geo<T> &) { return nullptr; }

// ---------------------------------------------------------------------------
// index rules (parallel/move3mg.c:3563-3576 etc.) and weights
// kind 0 'N' grid-line staggering, 1 'C' cell centre, 2 'P' pressure cell,
// 3 'R' node restriction of ronurestrict (move3mg.c:3433-3444)
// ---------------------------------------------------------------------------
bool axis_rule(int kind, int m, int nf, int nc, const std::vector<double> &gf,
               const std::vector<double> &gc, int &lo, double &w0, double &w1) {
  if (kind == 1 && m > nf - 2) return false;
  if (kind == 2 && m < 1) return false;
  lo = (m - 2) / 2 + 2;
  if (m < 3) lo = m;
  if (m > nf - 4) lo = nc - (nf - m);
  double d, dd0, dd1;
  if (kind == 0 || kind == 3) {
    if (lo > nc - 2) lo = nc - 2;
    d = gc[lo + 1] - gc[lo];
    dd0 = (gc[lo] - gf[m]) / d;
    dd1 = (gc[lo + 1] - gf[m]) / d;
    if (kind == 3) {
      w0 = 1.0 - std::fabs(dd0);
      w1 = 1.0 - std::fabs(dd1);
      return true;
    }
  } else if (kind == 1) {
    if (lo > nc - 3) lo = nc - 3;
    if ((gf[m] + gf[m + 1]) / 2.0 < (gc[lo] + gc[lo + 1]) / 2.0 && lo > 0) lo -= 1;
    d = (gc[lo + 2] - gc[lo]) / 2.0;
    dd0 = (gc[lo] + gc[lo + 1] - gf[m] - gf[m + 1]) / 2.0 / d;
    dd1 = (gc[lo + 1] + gc[lo + 2] - gf[m] - gf[m + 1]) / 2.0 / d;
  } else {
    if (lo > nc - 2) lo = nc - 2;
    if ((gf[m] + gf[m - 1]) / 2.0 < (gc[lo] + gc[lo - 1]) / 2.0 && lo > 1) lo -= 1;
    d = (gc[lo + 1] - gc[lo - 1]) / 2.0;
    dd0 = (gc[lo] + gc[lo - 1] - gf[m] - gf[m - 1]) / 2.0 / d;
    dd1 = (gc[lo + 1] + gc[lo] - gf[m] - gf[m - 1]) / 2.0 / d;
  }
  w0 = 1.0 - (-dd0);
  w1 = 1.0 - dd1;
  return true;
}

AxisHost build_axis(int kind, int nf, int nc, const std::vector<double> &gf, const std::vector<double> &gc) {
  AxisHost a;
  a.lo.assign(nf, 0);
  a.w.assign(2 * nf, 0.0);
  std::vector<char> ok(nf, 0);
  for (int m = 0; m < nf; m++) {
    int lo;
    double w0, w1;
    if (!axis_rule(kind, m, nf, nc, gf, gc, lo, w0, w1)) continue;
    ok[m] = 1;
    a.lo[m] = lo;
    a.w[2 * m] = w0;
    a.w[2 * m + 1] = w1;
  }
  a.cbeg.assign(nc + 1, 0);
  for (int C = 0; C < nc; C++) {
    a.cbeg[C] = (int)a.cfi.size();
    for (int m = 0; m < nf; m++) {
      if (!ok[m]) continue;
      if (a.lo[m] == C) { a.cfi.push_back(m); a.cw.push_back(a.w[2 * m]); }
      else if (a.lo[m] + 1 == C) { a.cfi.push_back(m); a.cw.push_back(a.w[2 * m + 1]); }
    }
  }
  a.cbeg[nc] = (int)a.cfi.size();
  for (int C = 0; C < nc; C++) a.kmax = std::max(a.kmax, a.cbeg[C + 1] - a.cbeg[C]);
  if (a.kmax <= kPad) {
    a.pfi.assign(kPad * (size_t)nc, -1);
    a.pw.assign(kPad * (size_t)nc, 0.0);
    for (int C = 0; C < nc; C++)
      for (int e = a.cbeg[C]; e < a.cbeg[C + 1]; e++) {
        a.pfi[kPad * C + (e - a.cbeg[C])] = a.cfi[e];
        a.pw[kPad * C + (e - a.cbeg[C])] = a.cw[e];
      }
  }
  return a;
}

template <typename T>
int upload(T **d, const std::vector<T> &h) {
  int rc = dalloc(d, h.size());
  if (rc) return rc;
  if (!h.empty() && hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
    return STK_EHIP;
  return STK_OK;
}

// stream-ordered copies that complete before returning (the library stream
// does not synchronise with the null stream)
#define H2D(c, dst, src, bytes)                                                         \
  do {                                                                                  \
    HIPCHK(hipMemcpyAsync((dst), (src), (bytes), hipMemcpyHostToDevice, (c)->st));      \
    HIPCHK(hipStreamSynchronize((c)->st));                                              \
  } while (0)
#define D2H(c, dst, src, bytes)                                                         \
  do {                                                                                  \
    HIPCHK(hipMemcpyAsync((dst), (src), (bytes), hipMemcpyDeviceToHost, (c)->st));      \
    HIPCHK(hipStreamSynchronize((c)->st));                                              \
  } while (0)

int upload_axis(AxisDev &d, const AxisHost &h) {
  d.release();
  std::vector<float> w32(h.w.begin(), h.w.end()), cw32(h.cw.begin(), h.cw.end());
  int rc = 0;
  rc |= upload(&d.lo, h.lo);
  rc |= upload(&d.w64, h.w);
  rc |= upload(&d.w32, w32);
  rc |= upload(&d.cbeg, h.cbeg);
  rc |= upload(&d.cfi, h.cfi);
  rc |= upload(&d.cw64, h.cw);
  rc |= upload(&d.cw32, cw32);
  d.kmax = h.kmax;
  if (h.kmax <= kPad) {
    std::vector<float> pw32(h.pw.begin(), h.pw.end());
    rc |= upload(&d.pfi, h.pfi);
    rc |= upload(&d.pw64, h.pw);
    rc |= upload(&d.pw32, pw32);
  }
  return rc ? STK_ENOMEM : STK_OK;
}

int lvl_index(int m, int nfine, int nlev, long mgs) {
  int r = (m - 2) / (int)mgs + 2;
  if (m < 3) r = m;
  if (m > nfine - 4) r = nlev - (nfine - m);
  return r;
}

bool comp_exists(const Level &L, int m1, int m2, int m3, int c) {
  if (c == 0) return m1 && m2 && m3;
  if (c == 1) return m2 < L.Y - 1 && m3 < L.Z - 1;
  if (c == 2) return m1 < L.X - 1 && m3 < L.Z - 1;
  return m1 < L.X - 1 && m2 < L.Y - 1;
}

// equation rows must keep their stencil inside the grid (the reference has
// no guard either: boundary equations cover those unknowns)
// (on a periodic axis the stencil wraps: velocity rows on lines 0..n-2)
bool eq_inside(const Level &L, int m1, int m2, int m3, int c) {
  if (c == 0) return true;
  int lo1 = 1, lo2 = 1, lo3 = 1;
  int h1 = L.X - 3, h2 = L.Y - 3, h3 = L.Z - 3;
  if (c == 1) h1 = L.X - 2;
  if (c == 2) h2 = L.Y - 2;
  if (c == 3) h3 = L.Z - 2;
  if (L.per & 1) { lo1 = 0; h1 = L.X - 2; }
  if (L.per & 2) { lo2 = 0; h2 = L.Y - 2; }
  if (L.per & 4) { lo3 = 0; h3 = L.Z - 2; }
  return m1 >= lo1 && m1 <= h1 && m2 >= lo2 && m2 <= h2 && m3 >= lo3 && m3 <= h3;
}

template <typename T> T *lvx(Level &L);
template <> float *lvx<float>(Level &L) { return L.x32[L.cur32]; }
template <> double *lvx<double>(Level &L) { return L.x64[L.cur64]; }
template <typename T> T *lvxo(Level &L);
template <> float *lvxo<float>(Level &L) { return L.x32[1 - L.cur32]; }
template <> double *lvxo<double>(Level &L) { return L.x64[1 - L.cur64]; }
template <typename T> void lvflip(Level &L);
template <> void lvflip<float>(Level &L) { L.cur32 ^= 1; }
template <> void lvflip<double>(Level &L) { L.cur64 ^= 1; }
template <typename T> T *lvb(Level &L);
template <> float *lvb<float>(Level &L) { return L.b32; }
template <> double *lvb<double>(Level &L) { return L.b64; }
template <typename T> T *lvr(Level &L);
template <> float *lvr<float>(Level &L) { return L.r32; }
template <> double *lvr<double>(Level &L) { return L.r64; }

int ensure_vectors(stk_ctx *c, bool want32, bool want64) {
  for (int l = 0; l <= c->L; l++) {
    Level &L = c->lv[l];
    size_t n4 = 4 * (size_t)L.n;
    if (want32 && !L.b32) {
      if (dalloc(&L.x32[0], n4) || dalloc(&L.x32[1], n4) || dalloc(&L.b32, n4) || dalloc(&L.r32, n4))
        return STK_ENOMEM;
      hipMemsetAsync(L.x32[0], 0, n4 * 4, c->st); hipMemsetAsync(L.x32[1], 0, n4 * 4, c->st);
      hipMemsetAsync(L.b32, 0, n4 * 4, c->st); hipMemsetAsync(L.r32, 0, n4 * 4, c->st);
    }
    if (want64 && !L.b64) {
      if (dalloc(&L.x64[0], n4) || dalloc(&L.x64[1], n4) || dalloc(&L.b64, n4) || dalloc(&L.r64, n4))
        return STK_ENOMEM;
      hipMemsetAsync(L.x64[0], 0, n4 * 8, c->st); hipMemsetAsync(L.x64[1], 0, n4 * 8, c->st);
      hipMemsetAsync(L.b64, 0, n4 * 8, c->st); hipMemsetAsync(L.r64, 0, n4 * 8, c->st);
    }
  }
  return STK_OK;
}

// ---------------------------------------------------------------------------
// multigrid building blocks
// ---------------------------------------------------------------------------
const int kTileDefault[3] = {32, 4, 4};

// boundary rows after a sweep: one launch over the dependency chains
template <typename T>
void bc_passes(stk_ctx *c, Level &L, const T *b, T *x, int mode) {
  if (L.nbc) hipLaunchKernelGGL(k_bc_chain<T>, dim3(blocks(L.nbc, 256)), dim3(256), 0, c->st, L.chain(), b, x, mode);
}

// whole-level LDS smoother for levels whose four unknowns fit in 64 KB
const size_t kLevelLds = 65536;

// LDS of one smoother tile: unknowns with a one-node halo, plus (sv) five
// viscosity planes of (ty+1)(tx+1)(tz+1) nodes
template <typename T>
size_t smooth_lds(const int *tile, bool sv) {
  const size_t nh = (size_t)(tile[0] + 2) * (tile[1] + 2) * (tile[2] + 2);
  const size_t nv = (size_t)(tile[0] + 1) * (tile[1] + 1) * (tile[2] + 1);
  const size_t nt = (size_t)tile[0] * tile[1] * tile[2];  // interior kinds, one byte per node
  return (4 * nh + (sv ? 5 * nv : 0)) * sizeof(T) + ((nt + 15) & ~(size_t)15);
}

const int kTileDefault4[4] = {32, 4, 4, 0};
template <typename T> const int *tuned_tile(const Level &L);
template <> const int *tuned_tile<float>(const Level &L) { return L.tune32[0] ? L.tune32 : kTileDefault4; }
template <> const int *tuned_tile<double>(const Level &L) { return L.tune64[0] ? L.tune64 : kTileDefault4; }

// One launch of the tile smoother.  The 32x4x4 tile (the default, and the
// usual choice on the finest levels) and the other tiles the autotuner picks
// on large levels run instances with their shape compiled in: the rows' LDS
// offsets become immediates and index divisions shifts.
// tcs: the tile chains of this shape (the sweep sets the boundary rows), or
// nullptr (a separate chain pass follows).
template <typename T>
void launch_smooth(stk_ctx *c, const int *tile, bool sv, const TileGeo &tg, const Geo<T> &g, const unsigned char *flags,
                   const T *xo, T *xn, const T *b, int withp, const TileChains<T> *tcs = nullptr) {
  const dim3 grid(tg.nty * tg.ntx * tg.ntz), blk(256);
  const size_t lds = smooth_lds<T>(tile, sv);
  const T v0 = (T)c->kn.v0koef, p0 = (T)c->kn.p0koef;
  // the compiled-shape instances stage walled levels only
  const bool t344 = tile[0] == 32 && tile[1] == 4 && tile[2] == 4 && !g.per;
  if constexpr (std::is_same<T, float>::value) {
    // the V-cycle's velocity sweep: rows' data prefetched per node pair
    // (one pair per thread: 256 threads, or 512 for the deeper tiles)
    const long pairs = (long)((tile[0] + 1) / 2) * tile[1] * tile[2];
    if (!withp && !sv && pairs <= 512) {
      const Level *L = level_of(c, g);
      const bool uni = L && L->uni;
      const FluxU fu = uni ? L->fu : FluxU{};
      const bool t348 = tile[0] == 32 && tile[1] == 4 && tile[2] == 8 && !g.per;
      const bool t384 = tile[0] == 32 && tile[1] == 8 && tile[2] == 4 && !g.per;
      const bool t642 = tile[0] == 64 && tile[1] == 4 && tile[2] == 2 && !g.per;
      const bool t624 = tile[0] == 64 && tile[1] == 2 && tile[2] == 4 && !g.per;
      const dim3 blk2(pairs <= 256 ? 256 : 512);
#define STK_VSMOOTH(TY, TX, TZ, CHV, BS)                                                                      \
  do {                                                                                                        \
    if (uni)                                                                                                  \
      hipLaunchKernelGGL((k_vsmooth32<TY, TX, TZ, CHV, true, BS>), grid, blk2, lds, c->st, g, flags, xo, xn, b, \
                         tg, v0, tcs, fu);                                                                    \
    else                                                                                                      \
      hipLaunchKernelGGL((k_vsmooth32<TY, TX, TZ, CHV, false, BS>), grid, blk2, lds, c->st, g, flags, xo, xn,  \
                         b, tg, v0, tcs, fu);                                                                 \
  } while (0)
      if (t344) { if (tcs) STK_VSMOOTH(32, 4, 4, true, 256); else STK_VSMOOTH(32, 4, 4, false, 256); }
      else if (t348) { if (tcs) STK_VSMOOTH(32, 4, 8, true, 512); else STK_VSMOOTH(32, 4, 8, false, 512); }
      else if (t384) { if (tcs) STK_VSMOOTH(32, 8, 4, true, 512); else STK_VSMOOTH(32, 8, 4, false, 512); }
      else if (t642) { if (tcs) STK_VSMOOTH(64, 4, 2, true, 256); else STK_VSMOOTH(64, 4, 2, false, 256); }
      else if (t624) { if (tcs) STK_VSMOOTH(64, 2, 4, true, 256); else STK_VSMOOTH(64, 2, 4, false, 256); }
      else if (pairs <= 256) { if (tcs) STK_VSMOOTH(0, 0, 0, true, 256); else STK_VSMOOTH(0, 0, 0, false, 256); }
      else { if (tcs) STK_VSMOOTH(0, 0, 0, true, 512); else STK_VSMOOTH(0, 0, 0, false, 512); }
#undef STK_VSMOOTH
      return;
    }
  }
#define STK_SMOOTH(...) \
  hipLaunchKernelGGL((k_smooth<__VA_ARGS__>), grid, blk, lds, c->st, g, flags, xo, xn, b, tg, withp, v0, p0, tcs)
  if (sv && tcs) STK_SMOOTH(T, true, 0, 0, 0, true);
  else if (sv) STK_SMOOTH(T, true);
  else if (t344 && tcs) STK_SMOOTH(T, false, 32, 4, 4, true);
  else if (t344) STK_SMOOTH(T, false, 32, 4, 4);
  else if (tcs) STK_SMOOTH(T, false, 0, 0, 0, true);
  else STK_SMOOTH(T, false);
#undef STK_SMOOTH
}

template <typename T> TileChains<T> *tile_chains(Level &L);
template <> TileChains<float> *tile_chains<float>(Level &L) { return L.bt_tc32; }
template <> TileChains<double> *tile_chains<double>(Level &L) { return L.bt_tc64; }
template <typename T> T *chain_c(Level &L);
template <> float *chain_c<float>(Level &L) { return L.bt_cc32; }
template <> double *chain_c<double>(Level &L) { return L.bt_cc64; }
template <typename T> T *chain_k(Level &L);
template <> float *chain_k<float>(Level &L) { return L.bt_ck32; }
template <> double *chain_k<double>(Level &L) { return L.bt_ck64; }

// The per-tile chain lists of `tile` on level L, built on first use.  Not
// while the stream is being captured (then the sweeps keep the separate
// pass; the graph is captured after the lists of its tiles exist).  Owners
// as TileChains describes; false: the separate pass (no chains, or a chain
// whose row and partner lie in different tiles, or whose partner lies
// outside every tile).
bool tile_bc(stk_ctx *c, Level &L, const int *tile) {
  if (!L.nbc) return false;
  if (L.bt_key[0] == tile[0] && L.bt_key[1] == tile[1] && L.bt_key[2] == tile[2]) return L.bt_ok;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(c->st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return false;
  if (L.bt_ok) {  // a captured sweep may name the old lists
    c->graph_ok = false;
    (void)hipStreamSynchronize(c->st);
  }
  L.free_tile_chains();
  for (int a = 0; a < 3; a++) L.bt_key[a] = tile[a];
  const int ty = tile[0], tx = tile[1], tz = tile[2], hy = ty + 2, hx = tx + 2, nh = hy * hx * (tz + 2);
  const TileGeo tg = make_tile(ty, tx, tz, L.box);
  const int nt = tg.nty * tg.ntx * tg.ntz;
  auto clampi = [](int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); };
  // tile of unknown u (clamped into the tile grid), whether u is in its
  // interior, and u's LDS index there
  auto place = [&](int u, int &t, int &lds) {
    const long n = u >> 2;
    const int m3 = (int)(n / L.XY), rem = (int)(n - (long)m3 * L.XY), m1 = rem / L.Y, m2 = rem - m1 * L.Y;
    const int ky = clampi((m2 - tg.oy) / ty, tg.nty - 1);
    const int kx = clampi((m1 - tg.ox) / tx, tg.ntx - 1);
    const int kz = clampi((m3 - tg.oz) / tz, tg.ntz - 1);
    const int ly = m2 - (tg.oy + ky * ty), lx = m1 - (tg.ox + kx * tx), lz = m3 - (tg.oz + kz * tz);
    t = (kz * tg.ntx + kx) * tg.nty + ky;
    lds = (u & 3) * nh + ((lz + 1) * hx + (lx + 1)) * hy + (ly + 1);
    return ly >= 0 && ly < ty && lx >= 0 && lx < tx && lz >= 0 && lz < tz;
  };
  std::vector<int> owner(L.nbc), rl(L.nbc), pl(L.nbc), off(nt + 1, 0), idx(L.nbc);
  for (int i = 0; i < L.nbc; i++) {
    int tr, lr, tp = -1, lp = -1;
    const bool rin = place(L.hch_r0[i], tr, lr);
    if (L.hch_j[i] >= 0 && !place(L.hch_j[i], tp, lp)) return false;
    if (rin && L.hch_j[i] >= 0 && tp != tr) return false;
    owner[i] = rin ? tr : (L.hch_j[i] >= 0 ? tp : tr);
    rl[i] = rin ? lr : -1;
    pl[i] = lp;
    off[owner[i] + 1]++;
  }
  for (int t = 0; t < nt; t++) off[t + 1] += off[t];
  std::vector<int> fill(off.begin(), off.end() - 1);
  for (int i = 0; i < L.nbc; i++) idx[fill[owner[i]]++] = i;
  std::vector<int> head(L.nbc), rlq(L.nbc), plq(L.nbc);  // in list order
  for (int q = 0; q < L.nbc; q++) { head[q] = L.hch_r0[idx[q]]; rlq[q] = rl[idx[q]]; plq[q] = pl[idx[q]]; }
  if (upload(&L.bt_off, off) || upload(&L.bt_idx, idx) || upload(&L.bt_head, head) || upload(&L.bt_rl, rlq) ||
      upload(&L.bt_pl, plq) || dalloc(&L.bt_cc32, L.nbc) || dalloc(&L.bt_ck32, L.nbc) || dalloc(&L.bt_cc64, L.nbc) ||
      dalloc(&L.bt_ck64, L.nbc))
    return false;
  const std::vector<TileChains<float>> t32{
      TileChains<float>{L.bt_off, L.bt_idx, L.bt_head, L.bt_rl, L.bt_pl, L.bt_cc32, L.bt_ck32}};
  const std::vector<TileChains<double>> t64{
      TileChains<double>{L.bt_off, L.bt_idx, L.bt_head, L.bt_rl, L.bt_pl, L.bt_cc64, L.bt_ck64}};
  if (upload(&L.bt_tc32, t32) || upload(&L.bt_tc64, t64)) return false;
  L.bt_ok = true;
  return true;
}

// tile == nullptr: the level's autotuned tile (or the default one)
template <typename T>
void smooth(stk_ctx *c, int l, int sweeps, const int *tile, int withp, bool whole_level = false) {
  Level &L = c->lv[l];
  const bool tuned = !tile;
  if (!tile) tile = tuned_tile<T>(L);
  // (the whole-level kernels read neighbours by plain offsets: not on periodic levels)
  if (whole_level && sweeps > 0 && !L.per && 4 * (size_t)L.n * sizeof(T) <= kLevelLds) {
    Geo<T> g = L.geo<T>(c->kn.p2koef);
    const LevelDiv ld{make_fdiv(L.Y), make_fdiv(L.X), make_fdiv((L.Y + 1) / 2)};
    const float *coef = std::is_same<T, float>::value ? L.lc32 : nullptr;
    const long ncand = (long)((L.Y + 1) / 2) * L.X * L.Z;
    if constexpr (std::is_same<T, float>::value) {
      const size_t lds_pc = 4 * (size_t)L.n * sizeof(float) + 16 * (size_t)L.nbc;
      if (coef && !withp && ncand <= 1024 && lds_pc <= c->lds_max) {
        hipLaunchKernelGGL(k_smooth_level_pc, dim3(1), dim3(1024), lds_pc, c->st, g, L.flags, lvx<float>(L),
                           lvb<float>(L), sweeps, (float)c->kn.v0koef, L.chain(), ld, coef);
        return;
      }
    }
    hipLaunchKernelGGL(k_smooth_level<T>, dim3(1), dim3(1024), 4 * (size_t)L.n * sizeof(T), c->st, g, L.flags,
                       lvx<T>(L), lvb<T>(L), sweeps, withp, (T)c->kn.v0koef, (T)c->kn.p0koef, L.chain(), ld, coef);
    return;
  }
  int ty = tile[0], tx = tile[1], tz = tile[2];
  const TileGeo tg = make_tile(ty, tx, tz, L.box);
  const bool sv = tuned && tile[3] && smooth_lds<T>(tile, true) <= kLevelLds;
  Geo<T> g = L.geo<T>(c->kn.p2koef);
  const bool inplace = sweeps > 0 && tile_bc(c, L, tile);
  if (inplace)  // the chains of this call's b, for all its sweeps
    hipLaunchKernelGGL(k_bc_compose<T>, dim3(blocks(L.nbc, 256)), dim3(256), 0, c->st, L.chain(), L.bt_idx,
                       lvb<T>(L), chain_c<T>(L), chain_k<T>(L), withp ? 0 : 1);
  for (int s = 0; s < sweeps; s++) {
    launch_smooth<T>(c, tile, sv, tg, g, L.flags, lvx<T>(L), lvxo<T>(L), lvb<T>(L), withp,
                     inplace ? tile_chains<T>(L) : nullptr);
    lvflip<T>(L);
    if (!inplace) bc_passes<T>(c, L, lvb<T>(L), lvx<T>(L), withp ? 0 : 1);
  }
}

// equation rows over 32x4x4 tiles of the whole level (x staged in LDS), then
// the boundary rows; y must not alias x
template <typename T, bool RES>
void rowop_rows(stk_ctx *c, const Level &L, const T *x, const T *b, T *y, int withp) {
  const int full[6] = {0, 0, 0, L.Y, L.X, L.Z};
  const TileGeo tg = make_tile(32, 4, 4, full);
  if (L.per)  // the run-time-shape instance stages periodic levels
    hipLaunchKernelGGL((k_rowop_tile<T, RES>), dim3(tg.nty * tg.ntx * tg.ntz), dim3(256),
                       smooth_lds<T>(kTileDefault4, false), c->st, L.geo<T>(c->kn.p2koef), L.flags, x, b, y, tg,
                       withp);
  else
    hipLaunchKernelGGL((k_rowop_tile<T, RES, 32, 4, 4>), dim3(tg.nty * tg.ntx * tg.ntz), dim3(256),
                       smooth_lds<T>(kTileDefault4, false), c->st, L.geo<T>(c->kn.p2koef), L.flags, x, b, y, tg,
                       withp);
}

template <typename T, bool RES>
void rowop(stk_ctx *c, int l, const T *x, const T *b, T *y, int withp) {
  Level &L = c->lv[l];
  rowop_rows<T, RES>(c, L, x, b, y, withp);
  if (L.nbc)
    hipLaunchKernelGGL((k_bc_rowop<T, RES>), dim3(blocks(L.nbc, 256)), dim3(256), 0, c->st, L.nbc, L.bc_r,
                       L.bc_j, L.bc_negk, x, b, y, withp);
}

CompTab comptab(const Level &L, int comp) {
  static const int K[4][3] = {{2, 2, 2}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}};
  CompTab t;
  for (int a = 0; a < 3; a++) t.ax[a] = L.ax[K[comp][a]][a].tab();
  return t;
}

// longest gather list of the axis tables the restriction uses (kinds N, C,
// P), or kPad + 1 if a table is not padded
int gather_width(const Level &C) {
  int k = 0;
  for (int kind = 0; kind < 3; kind++)
    for (int a = 0; a < 3; a++) k = std::max(k, C.ax[kind][a].pfi ? C.ax[kind][a].kmax : kPad + 1);
  return k;
}

SepComps sepcomps(const Level &C) {
  static const int K[4][3] = {{2, 2, 2}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}};
  SepComps t;
  for (int c = 0; c < 3; c++)
    for (int a = 0; a < 3; a++) {
      const AxisDev &d = C.ax[K[c + 1][a]][a];
      t.c[c].ax[a] = SepTab{d.pfi, d.pw32};
    }
  return t;
}

// the three passes of the separable restriction F = level l-1 -> C = level l
// (mode 1: weight sums into C.rs_ws)
void rsep(stk_ctx *c, int l, const float *rf, float *bcoarse, float *xcoarse, int mode) {
  Level &F = c->lv[l - 1], &C = c->lv[l];
  const SepComps t = sepcomps(C);
  const long n1 = (long)F.Z * F.X * C.Y, n2 = (long)F.Z * C.X * C.Y;
  hipLaunchKernelGGL(k_rsep_y, dim3(blocks(n1, 256)), dim3(256), 0, c->st, F.X, F.Y, F.Z, C.Y, t, F.flags, rf,
                     C.rs_t1, mode);
  hipLaunchKernelGGL(k_rsep_x, dim3(blocks(n2, 256)), dim3(256), 0, c->st, F.X, C.X, C.Y, F.Z, t, C.rs_t1, C.rs_t2);
  hipLaunchKernelGGL(k_rsep_z, dim3(blocks(C.n, 256)), dim3(256), 0, c->st, C.X, C.Y, C.Z, t, C.rs_t2, C.flags,
                     C.rs_ws, bcoarse, xcoarse, mode);
}

template <typename T>
void restrict_to(stk_ctx *c, int l, const T *rf, T *bcoarse, T *xcoarse, int withp) {
  Level &F = c->lv[l - 1], &C = c->lv[l];
  if constexpr (std::is_same<T, float>::value) {
    if (!withp && C.rs_ws) {  // the V-cycle's velocity rows
      rsep(c, l, rf, bcoarse, xcoarse, 0);
      return;
    }
  }
  (void)F;
  const int kw = gather_width(C);
  if (kw <= 4)
    hipLaunchKernelGGL((k_restrictK<T, 4>), dim3(blocks(4 * C.n, 256)), dim3(256), 0, c->st, F.X, F.Y, F.Z, C.X,
                       C.Y, C.Z, comptab(C, 0), comptab(C, 1), comptab(C, 2), comptab(C, 3), F.flags, C.flags, rf,
                       c->nd + F.off, c->nd + C.off, bcoarse, xcoarse, withp);
  else if (kw <= kPad)
    hipLaunchKernelGGL((k_restrictK<T, kPad>), dim3(blocks(4 * C.n, 256)), dim3(256), 0, c->st, F.X, F.Y, F.Z, C.X,
                       C.Y, C.Z, comptab(C, 0), comptab(C, 1), comptab(C, 2), comptab(C, 3), F.flags, C.flags, rf,
                       c->nd + F.off, c->nd + C.off, bcoarse, xcoarse, withp);
  else
    hipLaunchKernelGGL(k_restrict<T>, dim3(blocks(4 * C.n, 256)), dim3(256), 0, c->st, F.X, F.Y, F.Z, C.X, C.Y, C.Z,
                       comptab(C, 0), comptab(C, 1), comptab(C, 2), comptab(C, 3), F.flags, C.flags, rf,
                       c->nd + F.off, c->nd + C.off, bcoarse, xcoarse, withp);
}

template <typename T>
void prolong_from(stk_ctx *c, int l, const T *xcoarse, T *xf, int withp, int ynreset) {
  Level &F = c->lv[l - 1], &C = c->lv[l];
  hipLaunchKernelGGL(k_prolong<T>, dim3(blocks(F.n, 128)), dim3(128), 0, c->st, F.X, F.Y, F.Z, C.X, C.Y, C.Z,
                     comptab(C, 0), comptab(C, 1), comptab(C, 2), comptab(C, 3), F.flags, xcoarse, xf,
                     (T)c->kn.p1koef, (T)c->kn.v1koef, withp, ynreset);
}

template <typename T> const T *ainv(stk_ctx *c);
template <> const float *ainv<float>(stk_ctx *c) { return c->ainv32; }
template <> const double *ainv<double>(stk_ctx *c) { return c->ainv64; }

// velocity-block V-cycle on level l from the current x_l; shape of the
// reference's V-cycle (move3mg.c:2710-2792) with the coarsest level solved
// exactly every visit
template <typename T>
void mg_cycle(stk_ctx *c, const stk_solver_opts *o, const int *tile, int l) {
  const int L = c->L;
  if (l == L) {
    Level &C = c->lv[L];
    int n4 = (int)(4 * C.n);
    (void)n4;  // unknowns outside the compact block keep the 0 the restriction wrote
    hipLaunchKernelGGL((k_gemv_c<T, T>), dim3(blocks(c->nca, 4)), dim3(256), 0, c->st, ainv<T>(c), c->cact, lvb<T>(C),
                       lvx<T>(C), c->nca);
    return;
  }
  Level &V = c->lv[l];
  smooth<T>(c, l, o->pre[l], tile, 0, true);
  rowop<T, true>(c, l, lvx<T>(V), lvb<T>(V), lvr<T>(V), 0);
  restrict_to<T>(c, l + 1, lvr<T>(V), lvb<T>(c->lv[l + 1]), lvx<T>(c->lv[l + 1]), 0);
  mg_cycle<T>(c, o, tile, l + 1);
  prolong_from<T>(c, l + 1, lvx<T>(c->lv[l + 1]), lvx<T>(V), 0, 0);
  smooth<T>(c, l, o->post[l], tile, 0, true);
}

template <typename T>
void vcycle_v(stk_ctx *c, const stk_solver_opts *o, const int *tile) {
  Level &V = c->lv[0];
  hipLaunchKernelGGL(k_fill<T>, dim3(blocks(4 * V.n, 256)), dim3(256), 0, c->st, 4 * V.n, T(0), lvx<T>(V));
  mg_cycle<T>(c, o, tile, 0);
}

// ---------------------------------------------------------------------------
// Schur approximation (weighted BFBt) and its aggregation K-cycle
// ---------------------------------------------------------------------------
// cells of one colour of a K level, and whether the colour-compact sweep's
// index decode is exact there (hy^2 X Z < 2^32)
inline unsigned long kcolour_n(const Level &V) { return (unsigned long)((V.Y + 1) / 2) * V.X * V.Z; }
inline bool kcolour_ok(const Level &V) { return (unsigned long)((V.Y + 1) / 2) * kcolour_n(V) < (1ul << 32); }
template <typename T> T *kcol(Level &L);
template <> float *kcol<float>(Level &L) { return L.kc32; }
template <> double *kcol<double>(Level &L) { return L.kc64; }

template <typename T> KPlanesC<T> kplanes(const Level &L);
template <> KPlanesC<float> kplanes<float>(const Level &L) {
  KPlanesC<float> k;
  for (int q = 0; q < 7; q++) k.p[q] = L.k32[q];
  k.per = L.per;
  return k;
}
template <> KPlanesC<double> kplanes<double>(const Level &L) {
  KPlanesC<double> k;
  for (int q = 0; q < 7; q++) k.p[q] = L.k64[q];
  k.per = L.per;
  return k;
}
template <typename T> T *kx(Level &L);
template <> float *kx<float>(Level &L) { return L.kx32; }
template <> double *kx<double>(Level &L) { return L.kx64; }
template <typename T> T *kb(Level &L);
template <> float *kb<float>(Level &L) { return L.kb32; }
template <> double *kb<double>(Level &L) { return L.kb64; }
template <typename T> T *kr(Level &L);
template <> float *kr<float>(Level &L) { return L.kr32; }
template <> double *kr<double>(Level &L) { return L.kr64; }
template <typename T> T *bfy(stk_ctx *c);
template <> float *bfy<float>(stk_ctx *c) { return c->bfy32; }
template <> double *bfy<double>(stk_ctx *c) { return c->bfy64; }
template <typename T> T *bfw(stk_ctx *c);
template <> float *bfw<float>(stk_ctx *c) { return c->bfw32; }
template <> double *bfw<double>(stk_ctx *c) { return c->bfw64; }
template <typename T> const T *dinvt(stk_ctx *c);
template <> const float *dinvt<float>(stk_ctx *c) { return c->dinv32; }
template <> const double *dinvt<double>(stk_ctx *c) { return c->dinv64; }
template <typename T> const T *kinvT(stk_ctx *c);
template <> const float *kinvT<float>(stk_ctx *c) { return c->kinvT32; }
template <> const double *kinvT<double>(stk_ctx *c) { return c->kinvT64; }
template <typename T> const T *iAt(const Level &L, int a);
template <> const float *iAt<float>(const Level &L, int a) { return L.iA32[a]; }
template <> const double *iAt<double>(const Level &L, int a) { return L.iA64[a]; }
template <typename T> const T *iBt(const Level &L, int a);
template <> const float *iBt<float>(const Level &L, int a) { return L.iB32[a]; }
template <> const double *iBt<double>(const Level &L, int a) { return L.iB64[a]; }

// levels of at most this many cells finish the K-cycle in one workgroup
const long kKTailCells = (long)kTailPT * 1024;

KAgg kagg(const Level &C) {
  KAgg t;
  for (int a = 0; a < 3; a++) { t.lo[a] = C.klo[a]; t.cnt[a] = C.kcnt[a]; }
  return t;
}

// one aggregation V-cycle for K x = b on level l (x_l = 0 on entry)
template <typename T>
void kcycle(stk_ctx *c, int l, int nu, T scale) {
  Level &V = c->lv[l];
  // the remaining levels are small: one workgroup runs them without launches
  int small = 0;
  while (l + small < c->L && c->lv[l + small].n <= kKTailCells) small++;
  if (l + small == c->L && small <= kKTail && c->nkc <= kTailCoarse) {
    KTail<T> t;
    t.nlev = small; t.nu = nu; t.nkc = c->nkc; t.scale = scale; t.kcell = c->kcell; t.kinvT = kinvT<T>(c);
    // the first tail level in LDS where it fits beside the coarse vector
    const size_t lds = small >= 1 ? 10 * (size_t)c->lv[l].n * sizeof(T) : 0;
    t.stage = small >= 1 && lds + kTailCoarse * sizeof(T) + 1024 <= c->lds_max;
    const size_t dyn = t.stage ? lds : 0;
    for (int i = 0; i <= small; i++) {
      Level &W = c->lv[l + i];
      KTailLevel<T> &d = t.lv[i];
      d.X = W.X; d.Y = W.Y; d.Z = W.Z;
      d.k = kplanes<T>(W);
      d.b = kb<T>(W); d.x = kx<T>(W); d.r = kr<T>(W);
      d.agg = i > 0 ? kagg(W) : KAgg{};
      d.px = W.kpar[0]; d.py = W.kpar[1]; d.pz = W.kpar[2];
    }
    switch (small) {
      case 0: hipLaunchKernelGGL((k_ktail<T, 0>), dim3(1), dim3(1024), dyn, c->st, t); break;
      case 1: hipLaunchKernelGGL((k_ktail<T, 1>), dim3(1), dim3(1024), dyn, c->st, t); break;
      case 2: hipLaunchKernelGGL((k_ktail<T, 2>), dim3(1), dim3(1024), dyn, c->st, t); break;
      case 3: hipLaunchKernelGGL((k_ktail<T, 3>), dim3(1), dim3(1024), dyn, c->st, t); break;
      default: hipLaunchKernelGGL((k_ktail<T, 4>), dim3(1), dim3(1024), dyn, c->st, t); break;
    }
    return;
  }
  if (l == c->L) {
    hipLaunchKernelGGL(k_kcoarse<T>, dim3(blocks(c->nkc, 64)), dim3(64), c->nkc * sizeof(T), c->st, c->nkc,
                       c->kcell, kinvT<T>(c), kb<T>(V), kx<T>(V));
    return;
  }
  const KPlanesC<T> k = kplanes<T>(V);
  const unsigned nb = blocks(V.n, 256);
  // one thread per cell of the colour on colour-compact planes where they exist
  const T *kc = kcol<T>(V);
  const FDiv dhy = make_fdiv((V.Y + 1) / 2), dx = make_fdiv(V.X);
  auto sweeps = [&](int c1) {
    for (int s = 0; s < nu; s++)
      for (int q = 0; q < 2; q++) {
        const int col = q == 0 ? c1 : 1 - c1;
        if (kc)
          hipLaunchKernelGGL(k_kgs_c<T>, dim3(blocks(kcolour_n(V), 256)), dim3(256), 0, c->st, V.X, V.Y, V.Z, dhy,
                             dx, kc, kb<T>(V), kx<T>(V), col, V.per);
        else
          hipLaunchKernelGGL(k_kgs<T>, dim3(nb), dim3(256), 0, c->st, V.X, V.Y, V.Z, k, kb<T>(V), kx<T>(V), col);
      }
  };
  sweeps(0);
  hipLaunchKernelGGL(k_kres<T>, dim3(nb), dim3(256), 0, c->st, V.X, V.Y, V.Z, k, kb<T>(V), kx<T>(V), kr<T>(V));
  Level &C = c->lv[l + 1];
  hipLaunchKernelGGL(k_krestrict<T>, dim3(blocks(C.n, 256)), dim3(256), 0, c->st, V.X, V.Y, C.X, C.Y, C.Z, kagg(C),
                     kr<T>(V), kb<T>(C), kx<T>(C));
  kcycle<T>(c, l + 1, nu, scale);
  hipLaunchKernelGGL(k_kprolong<T>, dim3(nb), dim3(256), 0, c->st, V.X, V.Y, V.Z, C.X, C.Y, V.kpar[0], V.kpar[1],
                     V.kpar[2], k.p[0], kx<T>(C), kx<T>(V), scale);
  sweeps(1);
}

// z_p = sign K^-1 (B D^-1 A D^-1 G) K^-1 r_p on the continuity cells,
// velocity slots of z cleared.  The BFBt product approximates (B A^-1 G)^-1;
// the preconditioner's pressure block is S = -B A^-1 G, hence sign = -1.
template <typename T>
void schur_bfbt(stk_ctx *c, const stk_solver_opts *o, const double *r, double *z, double sign) {
  Level &F = c->lv[0];
  const unsigned nb = blocks(F.n, 256);
  const T sc = (T)o->kscale;
  const int nu = o->ksweeps;
  hipLaunchKernelGGL(k_bf_gather<T>, dim3(nb), dim3(256), 0, c->st, F.n, F.flags, r, kb<T>(F));
  hipLaunchKernelGGL(k_fill<T>, dim3(nb), dim3(256), 0, c->st, F.n, T(0), kx<T>(F));
  kcycle<T>(c, 0, nu, sc);
  hipLaunchKernelGGL(k_bf_grad<T>, dim3(nb), dim3(256), 0, c->st, F.X, F.Y, F.Z, iBt<T>(F, 0), iBt<T>(F, 1),
                     iBt<T>(F, 2), dinvt<T>(c), kx<T>(F), bfy<T>(c), F.per);
  rowop<T, false>(c, 0, bfy<T>(c), nullptr, bfw<T>(c), 0);
  hipLaunchKernelGGL(k_bf_div<T>, dim3(nb), dim3(256), 0, c->st, F.X, F.Y, F.Z, F.flags, iAt<T>(F, 0),
                     iAt<T>(F, 1), iAt<T>(F, 2), dinvt<T>(c), bfw<T>(c), kb<T>(F), F.per);
  hipLaunchKernelGGL(k_fill<T>, dim3(nb), dim3(256), 0, c->st, F.n, T(0), kx<T>(F));
  kcycle<T>(c, 0, nu, sc);
  hipLaunchKernelGGL(k_bf_scatter<T>, dim3(nb), dim3(256), 0, c->st, F.n, F.flags, kx<T>(F), sign, z);
}

// z = M^{-1} r : pressure by the Schur approximation, velocity by one V-cycle
template <typename T>
void precond(stk_ctx *c, const stk_solver_opts *o, const double *r, double *z, const int *tile, float *z32 = nullptr) {
  Level &F = c->lv[0];
  long n4 = 4 * F.n;
  if (o->schur >= 1)
    schur_bfbt<T>(c, o, r, z, o->schur == 1 ? -1.0 : 1.0);
  else
    hipLaunchKernelGGL(k_pscale, dim3(blocks(n4, 256)), dim3(256), 0, c->st, n4, F.flags, F.e64[0], r, z, o->omega);
  bc_passes<double>(c, F, r, z, 2);
  Geo<double> g = F.geo<double>(c->kn.p2koef);
  hipLaunchKernelGGL(k_vrhs<T>, dim3(blocks(F.n, 256)), dim3(256), 0, c->st, g, F.flags, r, z, lvb<T>(F));
  vcycle_v<T>(c, o, tile);
  if constexpr (std::is_same<T, float>::value) {
    if (z32) {  // z rounded to FP32 and its FP32 copy, in the same pass
      hipLaunchKernelGGL(k_combine_round, dim3(blocks(n4, 256)), dim3(256), 0, c->st, n4, lvx<float>(F), z, z32);
      return;
    }
  }
  hipLaunchKernelGGL(k_combine<T>, dim3(blocks(n4, 256)), dim3(256), 0, c->st, n4, lvx<T>(F), z);
}

// weighted dots: out[i] = <a, bs[i]>_W for k vectors (device result)
void multidot(stk_ctx *c, const double *a, const std::vector<const double *> &bs, double *out) {
  long n4 = 4 * c->lv[0].n;
  int k = (int)bs.size();
  for (int i0 = 0; i0 < k; i0 += 16) {
    int kk = std::min(16, k - i0);
    Ptrs16 P;
    for (int i = 0; i < 16; i++) P.p[i] = (i < kk) ? bs[i0 + i] : nullptr;
    hipLaunchKernelGGL(k_multidot, dim3(c->nred), dim3(256), 0, c->st, n4, c->W2, a, P, kk, c->red);
    hipLaunchKernelGGL(k_reduce_rows, dim3(kk), dim3(256), 0, c->st, c->red, c->nred, kk, out + i0);
  }
}

void multiaxpy(stk_ctx *c, int k, const double *coef, double *a, double *b) {
  long n4 = 4 * c->lv[0].n;
  for (int i0 = 0; i0 < k; i0 += 16) {
    int kk = std::min(16, k - i0);
    Ptrs16 A, B;
    for (int i = 0; i < 16; i++) {
      A.p[i] = (i < kk) ? c->Qs[i0 + i] : nullptr;
      B.p[i] = (i < kk) ? c->Zs[i0 + i] : nullptr;
    }
    hipLaunchKernelGGL(k_multiaxpy, dim3(blocks(n4, 256)), dim3(256), 0, c->st, n4, kk, coef + i0, A, B, a, b);
  }
}

int blocks_norms(stk_ctx *c, const double *r, const double *b, double out4[4]) {
  long n4 = 4 * c->lv[0].n;
  hipLaunchKernelGGL(k_blocks, dim3(c->nred), dim3(256), 0, c->st, n4, c->lv[0].flags, c->W2, r, b, c->red);
  hipLaunchKernelGGL(k_reduce_rows, dim3(4), dim3(256), 0, c->st, c->red, c->nred, 4, c->scal + 8);
  HIPCHK(hipMemcpyAsync(c->hpin, c->scal + 8, 4 * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  memcpy(out4, c->hpin, 4 * sizeof(double));
  return STK_OK;
}

int true_residual(stk_ctx *c) {
  rowop<double, true>(c, 0, c->X, c->B, c->R, 1);
  return STK_OK;
}

// ---------------------------------------------------------------------------
// Krylov vectors borrowed for one solve (history = 1).  Sets are keyed by
// vector length and kept for reuse; a set needs m + 1 basis and m
// preconditioned vectors.  When the device has no room for a new set, the
// request waits for another solve to return one, and frees idle sets of
// other lengths first.
// ---------------------------------------------------------------------------
struct HistSet {
  int dev = 0;
  long n4 = 0;
  int m = 0;
  std::vector<double *> Z, Q;
};
struct HistPool {
  std::mutex mu;
  std::condition_variable cv;
  std::vector<HistSet *> idle;
  int lent = 0;
};
static HistPool g_hist;

static void hist_free(HistSet *h) {
  for (auto p : h->Z) dfree(p);
  for (auto p : h->Q) dfree(p);
  delete h;
}

static HistSet *hist_alloc(int dev, long n4, int m) {
  HistSet *h = new HistSet();
  h->dev = dev; h->n4 = n4; h->m = m;
  for (int i = 0; i <= m; i++) {
    double *q = nullptr, *z = nullptr;
    if (dalloc(&q, n4) || (i < m && dalloc(&z, n4))) {
      dfree(q);
      hist_free(h);
      return nullptr;
    }
    h->Q.push_back(q);
    if (i < m) h->Z.push_back(z);
  }
  return h;
}

static HistSet *hist_acquire(long n4, int m) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::unique_lock<std::mutex> lk(g_hist.mu);
  for (;;) {
    for (size_t i = 0; i < g_hist.idle.size(); i++) {
      HistSet *h = g_hist.idle[i];
      if (h->dev == dev && h->n4 == n4 && h->m >= m) {
        g_hist.idle.erase(g_hist.idle.begin() + i);
        g_hist.lent++;
        return h;
      }
    }
    HistSet *h = hist_alloc(dev, n4, m);
    if (!h) {  // no room: free idle sets of this device, then wait for a lent one
      bool freed = false;
      for (size_t i = 0; i < g_hist.idle.size();)
        if (g_hist.idle[i]->dev == dev) { hist_free(g_hist.idle[i]); g_hist.idle.erase(g_hist.idle.begin() + i); freed = true; }
        else i++;
      if (freed) h = hist_alloc(dev, n4, m);
    }
    if (h) { g_hist.lent++; return h; }
    if (g_hist.lent == 0) return nullptr;  // nothing will come back: out of memory
    g_hist.cv.wait(lk);
  }
}

static void hist_release(HistSet *h) {
  {
    std::lock_guard<std::mutex> lk(g_hist.mu);
    g_hist.idle.push_back(h);
    g_hist.lent--;
  }
  g_hist.cv.notify_all();
}

// the context's solve-time view of a borrowed set, returned on scope exit
struct HistLease {
  stk_ctx *c = nullptr;
  HistSet *h = nullptr;
  std::vector<double *> keepZ, keepQ;
  ~HistLease();
};

int ensure_outer(stk_ctx *c, int m) {
  long n4 = 4 * c->lv[0].n;
  if (!c->X) {
    if (dalloc(&c->X, n4) || dalloc(&c->B, n4) || dalloc(&c->R, n4) || dalloc(&c->W2, n4) || dalloc(&c->T1, n4) ||
        dalloc(&c->wu, n4) || dalloc(&c->ww, n4) || dalloc(&c->wz, n4) ||
        dalloc(&c->red, (size_t)c->nred * 16) || dalloc(&c->scal, 256) ||
        dalloc(&c->kred, (size_t)4 * c->nred * (2 * kMaxDirs + 2)) || dalloc(&c->hcol, 2 * kMaxDirs + 2))
      return STK_ENOMEM;
    RtLock lk(g_rt_mu);
    if (hipHostMalloc((void **)&c->hpin, (2 * kMaxDirs + 8) * sizeof(double), hipHostMallocDefault) != hipSuccess)
      return STK_ENOMEM;
    hipMemsetAsync(c->X, 0, n4 * 8, c->st);
  }
  while ((int)c->Zs.size() < m) {
    double *z, *q;
    if (dalloc(&z, n4) || dalloc(&q, n4)) return STK_ENOMEM;
    c->Zs.push_back(z);
    c->Qs.push_back(q);
  }
  return STK_OK;
}

HistLease::~HistLease() {
  if (!h) return;
  (void)hipStreamSynchronize(c->st);  // the stream is done with the vectors
  c->Zs = std::move(keepZ);
  c->Qs = std::move(keepQ);
  hist_release(h);
}

// ---------------------------------------------------------------------------
// coarsest-level inverse (Gauss-Jordan with partial pivoting, FP64)
// ---------------------------------------------------------------------------
// inv = A^-1 for a device n x n row-major A (overwritten); STK_ESINGULAR if a
// pivot column has no nonzero left.  Two launches per column.
int gj_invert(stk_ctx *c, double *A, int n, double *inv) {
  double *B = nullptr, *pval = nullptr, *cand = nullptr;
  int *used = nullptr, *prow = nullptr, *sing = nullptr;
  if (dalloc(&B, (size_t)n * n) || dalloc(&pval, n) || dalloc(&cand, n) || dalloc(&used, n) || dalloc(&prow, n) ||
      dalloc(&sing, 1)) {
    dfree(B); dfree(pval); dfree(cand); dfree(used); dfree(prow); dfree(sing);
    return STK_ENOMEM;
  }
  hipLaunchKernelGGL(k_eye, dim3(blocks((long)n * n, 256)), dim3(256), 0, c->st, B, n);
  HIPCHK(hipMemsetAsync(used, 0, (size_t)n * 4, c->st));
  HIPCHK(hipMemsetAsync(sing, 0, 4, c->st));
  // panels of kb columns in one workgroup's LDS (about 3 launches per kb
  // columns); column by column where a panel of 4 would not fit
  const int kbl = (int)std::min<long>(kGjbMax, ((long)c->lds_max - 4096 - n) / (8L * n));
  double *F = nullptr, *R = nullptr;
  if (kbl >= 4 && (dalloc(&F, (size_t)n * kGjbMax) || dalloc(&R, (size_t)kGjbMax * 2 * n))) {
    dfree(B); dfree(pval); dfree(cand); dfree(used); dfree(prow); dfree(sing); dfree(F); dfree(R);
    return STK_ENOMEM;
  }
  if (kbl >= 4) {
    (void)hipFuncSetAttribute((const void *)k_gjb_panel, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)((size_t)n * kbl * 8 + n));
    for (int k0 = 0; k0 < n; k0 += kbl) {
      const int kb = std::min(kbl, n - k0), m = 2 * n - k0 - kb;
      hipLaunchKernelGGL(k_gjb_panel, dim3(1), dim3(1024), (size_t)n * kb * 8 + n, c->st, A, n, k0, kb, used, prow,
                         pval, F, sing);
      if (m <= 0) continue;
      hipLaunchKernelGGL(k_gjb_rows, dim3(blocks(m, 256)), dim3(256), 0, c->st, A, B, n, k0, kb, prow, F, R, sing);
      hipLaunchKernelGGL(k_gjb_update, dim3(blocks(m, 256), (n + kGjbRows - 1) / kGjbRows), dim3(256), 0, c->st, A,
                         B, n, k0, kb, F, R, sing);
    }
  } else {
    for (int k = 0; k < n; k++) {
      hipLaunchKernelGGL(k_gj_select, dim3(1), dim3(1024), 0, c->st, A, n, k, used, prow, pval, cand, sing);
      hipLaunchKernelGGL(k_gj_elim, dim3(blocks(2L * n * n, 256)), dim3(256), 0, c->st, A, B, n, k, used, prow,
                         pval, cand);
    }
  }
  dfree(F);
  dfree(R);
  int hs = 0;
  D2H(c, &hs, sing, 4);
  if (!hs)
    hipLaunchKernelGGL(k_gj_final, dim3(blocks((long)n * n, 256)), dim3(256), 0, c->st, B, n, prow, pval, inv);
  HIPCHK(hipStreamSynchronize(c->st));
  dfree(B); dfree(pval); dfree(cand); dfree(used); dfree(prow); dfree(sing);
  return hs ? STK_ESINGULAR : (hipGetLastError() == hipSuccess ? STK_OK : STK_EHIP);
}

// velocity block of the coarsest level: Gauss-Jordan on the compact system
// of velocity unknowns with an equation or boundary row (N x N, N < 4 mgn)
int coarse_inverse_v(stk_ctx *c) {
  Level &C = c->lv[c->L];
  const int n4 = (int)(4 * C.n);
  std::vector<int> act, idx(n4, -1);
  for (int u = 0; u < n4; u++) {
    const int cc = u & 3;
    if (cc != 0 && kind_of(C.hflags[u >> 2], cc) != KIND_NONE) {
      idx[u] = (int)act.size();
      act.push_back(u);
    }
  }
  const int N = (int)act.size();
  if (N == 0) return STK_EINVAL;
  int *dact = nullptr, *didx = nullptr;
  if (upload(&dact, act) || upload(&didx, idx)) return STK_ENOMEM;
  double *A = nullptr, *inv = nullptr;
  if (dalloc(&A, (size_t)N * N) || dalloc(&inv, (size_t)N * N)) { dfree(A); dfree(dact); dfree(didx); return STK_ENOMEM; }
  HIPCHK(hipMemsetAsync(A, 0, (size_t)N * N * 8, c->st));
  Geo<double> g = C.geo<double>(c->kn.p2koef);
  hipLaunchKernelGGL(k_dense_rows_c, dim3(blocks(N, 128)), dim3(128), 0, c->st, g, C.flags, C.nbc, C.bc_r, C.bc_j,
                     C.bc_negk, dact, didx, A, N);
  const int rc = gj_invert(c, A, N, inv);
  dfree(A); dfree(didx);
  if (rc) { dfree(inv); dfree(dact); return rc; }
  if (N != c->nca || !c->ainv64) {  // same size: refill in place (captured graphs keep their pointers)
    dfree(c->ainv64); dfree(c->ainv32); dfree(c->cact);
    if (dalloc(&c->ainv64, (size_t)N * N) || dalloc(&c->ainv32, (size_t)N * N) || dalloc(&c->cact, N))
      return STK_ENOMEM;
    c->graph_ok = false;
  }
  HIPCHK(hipMemcpyAsync(c->ainv64, inv, (size_t)N * N * 8, hipMemcpyDeviceToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->cact, dact, (size_t)N * 4, hipMemcpyDeviceToDevice, c->st));
  c->nca = N;
  hipLaunchKernelGGL((k_convert<double, float>), dim3(blocks((long)N * N, 256)), dim3(256), 0, c->st, (long)N * N,
                     c->ainv64, c->ainv32);
  HIPCHK(hipStreamSynchronize(c->st));
  dfree(inv); dfree(dact);
  return STK_OK;
}

int coarse_inverse(stk_ctx *c, int withp, double **out) {
  Level &C = c->lv[c->L];
  int n = (int)(4 * C.n);
  double *A = nullptr, *inv = nullptr;
  if (dalloc(&A, (size_t)n * n) || dalloc(&inv, (size_t)n * n)) { dfree(A); return STK_ENOMEM; }
  HIPCHK(hipMemsetAsync(A, 0, (size_t)n * n * 8, c->st));
  Geo<double> g = C.geo<double>(c->kn.p2koef);
  hipLaunchKernelGGL(k_dense_rows, dim3(blocks(n, 128)), dim3(128), 0, c->st, g, C.flags, C.nbc, C.bc_r, C.bc_j,
                     C.bc_negk, A, n, withp);
  const int rc = gj_invert(c, A, n, inv);
  dfree(A);
  if (rc) { dfree(inv); return rc; }
  dfree(*out);
  *out = inv;
  return STK_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
// smoother tile autotuning (timed on the library stream)
template <typename T>
static double time_smooth(stk_ctx *c, int l, const int *tile, int reps) {
  Level &L = c->lv[l];
  const int ty = tile[0], tx = tile[1], tz = tile[2];
  const TileGeo tg = make_tile(ty, tx, tz, L.box);
  const bool sv = tile[3] && smooth_lds<T>(tile, true) <= kLevelLds;
  Geo<T> g = L.geo<T>(c->kn.p2koef);
  T *x0 = lvx<T>(L), *x1 = lvxo<T>(L);
  // as the V-cycle runs a sweep of this shape: boundary rows set inside the
  // tiles (chains composed once per smoothing call) where the shape allows,
  // else a chain pass after the sweep
  const bool inplace = tile_bc(c, L, tile);
  if (inplace)
    hipLaunchKernelGGL(k_bc_compose<T>, dim3(blocks(L.nbc, 256)), dim3(256), 0, c->st, L.chain(), L.bt_idx,
                       lvb<T>(L), chain_c<T>(L), chain_k<T>(L), 1);
  auto launch = [&]() {
    launch_smooth<T>(c, tile, sv, tg, g, L.flags, x0, x1, lvb<T>(L), 0, inplace ? tile_chains<T>(L) : nullptr);
    if (!inplace) bc_passes<T>(c, L, lvb<T>(L), x1, 1);
  };
  launch();
  hipEventRecord(c->ev0, c->st);
  for (int i = 0; i < reps; i++) launch();
  hipEventRecord(c->ev1, c->st);
  hipEventSynchronize(c->ev1);
  float ms = 0.f;
  hipEventElapsedTime(&ms, c->ev0, c->ev1);
  return ms / reps;
}

template <typename T>
static int autotune_levels(stk_ctx *c, int reps) {
  static const int shape[][3] = {{32, 4, 4}, {32, 4, 8}, {32, 8, 4}, {16, 8, 8}, {64, 4, 2}, {64, 2, 4},
                                 {16, 4, 4}, {32, 2, 4}, {16, 4, 8}, {8, 8, 8}, {64, 2, 2}, {32, 2, 8}};
  static const int xz[][2] = {{4, 4}, {2, 4}, {4, 2}, {4, 8}, {8, 4}, {2, 8}, {2, 2}};
  const int ns = (int)(sizeof(shape) / sizeof(shape[0]));
  for (int l = 0; l <= c->L; l++) {
    Level &L = c->lv[l];
    if (!L.per && 4 * (size_t)L.n * sizeof(T) <= kLevelLds) continue;  // smoothed by the whole-level kernel
    // fixed shapes, plus tile heights that split the y extent of the equation-row box evenly
    std::vector<std::array<int, 4>> cand;
    auto add = [&](int ty, int tx, int tz) {
      for (int sv = 0; sv < 2; sv++) {
        const std::array<int, 4> t = {ty, tx, tz, sv};
        if (ty >= 2 && smooth_lds<T>(t.data(), sv) <= kLevelLds) cand.push_back(t);
      }
    };
    for (int k = 0; k < ns; k++) add(shape[k][0], shape[k][1], shape[k][2]);
    for (int parts = 2; parts <= 8; parts++)
      for (const auto &q : xz) add((L.box[3] - L.box[0] + parts - 1) / parts, q[0], q[1]);
    // candidate 0 is the default 32x4x4 tile; another shape must beat it by 3 %
    // so that timing noise does not move the choice
    double best = 1e30, base = 0.0;
    int bi = 0;
    for (int k = 0; k < (int)cand.size(); k++) {
      const double ms = time_smooth<T>(c, l, cand[k].data(), reps);
      if (k == 0) base = ms;
      if (ms < best && (k == 0 || ms < 0.97 * base)) { best = ms; bi = k; }
    }
    int *t = std::is_same<T, float>::value ? L.tune32 : L.tune64;
    for (int a = 0; a < 4; a++) t[a] = cand[bi][a];
  }
  return hipGetLastError() == hipSuccess ? STK_OK : STK_EHIP;
}

// one K-cycle on host cell vectors (inspection entry point)
template <typename T>
static void kcycle_host(stk_ctx *c, int nu, double scale, const double *b, double *x) {
  Level &F = c->lv[0];
  std::vector<T> hb(b, b + F.n), hx(F.n);
  hipMemcpyAsync(kb<T>(F), hb.data(), F.n * sizeof(T), hipMemcpyHostToDevice, c->st);
  hipMemsetAsync(kx<T>(F), 0, F.n * sizeof(T), c->st);
  kcycle<T>(c, 0, nu, (T)scale);
  hipStreamSynchronize(c->st);
  hipMemcpyAsync(hx.data(), kx<T>(F), F.n * sizeof(T), hipMemcpyDeviceToHost, c->st);
  hipStreamSynchronize(c->st);
  for (long i = 0; i < F.n; i++) x[i] = (double)hx[i];
}

// ---------------------------------------------------------------------------
// marker -> node averaging: device buffers of one call, and the kernels
// ---------------------------------------------------------------------------
struct MarkerBufs {
  float *in[5] = {}, *so[5] = {};
  int *key = nullptr, *val = nullptr, *skey = nullptr, *sval = nullptr, *start = nullptr;
  void *tmp = nullptr;
  size_t tmpb = 0;
};

// host marker arrays staged into b.in, binned by cell (stable radix sort),
// permuted, cell starts, then the per-node gather into the context's fields
static int markers_run(stk_ctx *c, long nmark, const float *mx, const float *my, const float *mz, const float *mnu,
                       const float *mro, double nukoef, MarkerBufs &b) {
  Level &F = c->lv[0];
  const int ncell = (F.X - 1) * (F.Y - 1) * (F.Z - 1);
  const float *src[5] = {mx, my, mz, mnu, mro};
  for (int k = 0; k < 5; k++) HIPCHK(hipMemcpyAsync(b.in[k], src[k], nmark * 4, hipMemcpyHostToDevice, c->st));
  hipLaunchKernelGGL(k_marker_cells, dim3(blocks(nmark, 256)), dim3(256), 0, c->st, nmark, b.in[0], b.in[1], b.in[2],
                     F.g64[0], F.g64[1], F.g64[2], F.X, F.Y, F.Z, b.key, b.val);
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(b.tmp, b.tmpb, b.key, b.skey, b.val, b.sval, (int)nmark, 0, 32, c->st));
  hipLaunchKernelGGL(k_marker_permute, dim3(blocks(nmark, 256)), dim3(256), 0, c->st, nmark, b.sval, b.in[0],
                     b.in[1], b.in[2], b.in[3], b.in[4], b.so[0], b.so[1], b.so[2], b.so[3], b.so[4]);
  // every start[0..ncell] is written (no entry needs clearing first)
  hipLaunchKernelGGL(k_cell_start, dim3(blocks(nmark + 1, 256)), dim3(256), 0, c->st, nmark, b.skey, ncell + 1,
                     b.start);
  const MGather q{F.X, F.Y, F.Z, F.g64[0], F.g64[1], F.g64[2], b.start, b.so[0], b.so[1], b.so[2], b.so[3], b.so[4],
                  c->kn.nubeg, c->kn.nuend, nukoef, c->kn.viscmod};
  hipLaunchKernelGGL(k_markers_gather, dim3(blocks(F.n, 256)), dim3(256), 0, c->st, q, F.per, c->nu, c->ro, c->nxx,
                     c->nxy, c->nxz, c->nyz);
  if (F.per)
    hipLaunchKernelGGL(k_markers_wrap, dim3(blocks(F.n, 256)), dim3(256), 0, c->st, F.X, F.Y, F.Z, F.per, c->nu,
                       c->ro, c->nxy, c->nxz, c->nyz);
  HIPCHK(hipStreamSynchronize(c->st));
  HIPCHK(hipGetLastError());
  c->fields_set = true;
  c->prepared = false;
  return STK_OK;
}

// Scratch of stk_markers_to_nodes_noalloc: 14 arrays of kScratchMarkers
// 4-byte words (5 staged and 5 sorted marker fields, keys, indices and their
// sorted copies), the cell starts, and the radix sort's temporary storage,
// each 256-byte aligned.  One copy per GPU, held until the process ends.
constexpr long kScratchMarkers = 1L << 20;
constexpr long kScratchCells = 1L << 18;
constexpr size_t kScratchArr = ((size_t)kScratchMarkers * 4 + 255) & ~(size_t)255;
constexpr size_t kScratchStart = ((size_t)(kScratchCells + 2) * 4 + 255) & ~(size_t)255;
constexpr size_t kScratchSort = (size_t)16 << 20;
constexpr size_t kScratchBytes = 14 * kScratchArr + kScratchStart + kScratchSort;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of this GPU's scratch array, looked up once per device
// (host code never uses the array's name as an address).
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[64] = {};

static int scratch_base(unsigned char **out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return STK_EINVAL;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      fprintf(stderr, "pardiso2_amd: scratch address on device %d: %s\n", dev, hipGetErrorString(e));
      return STK_EHIP;
    }
    g_scratch_ptr[dev] = (unsigned char *)p;
  }
  *out = g_scratch_ptr[dev];
  return STK_OK;
}

// ---------------------------------------------------------------------------
// Temperature solve (parallel/heat3mg.c) on levels 0..Lt of the hierarchy
// ---------------------------------------------------------------------------
namespace {

void heat_free(HeatState *h) {
  if (!h) return;
  for (int l = 0; l < MAXLV; l++) h->lv[l].release();
  dfree(h->ro); dfree(h->cp); dfree(h->kt); dfree(h->rcp);
  dfree(h->tk); dfree(h->tk0); dfree(h->tk1); dfree(h->ht); dfree(h->tk2);
  dfree(h->x); dfree(h->b); dfree(h->r); dfree(h->part);
  delete h;
}

int heat_alloc(stk_ctx *c) {
  if (c->heat) return STK_OK;
  HeatState *h = new HeatState();
  const long nt = c->ntot, n0 = c->lv[0].n;
  // every level's rows read all of its level-0 sized partial buffer at most
  const long np = 4 * (2 * (long)blocks(n0, 256) + 2);
  if (dalloc(&h->ro, nt) || dalloc(&h->cp, nt) || dalloc(&h->kt, nt) || dalloc(&h->rcp, nt) ||
      dalloc(&h->tk, n0) || dalloc(&h->tk0, n0) || dalloc(&h->tk1, n0) || dalloc(&h->ht, n0) ||
      dalloc(&h->tk2, n0) || dalloc(&h->x, nt) ||
      dalloc(&h->b, nt) || dalloc(&h->r, nt) || dalloc(&h->part, np)) {
    heat_free(h);
    return STK_ENOMEM;
  }
  h->hpart.assign(np, 0.0);
  for (double *p : {h->x, h->b, h->r}) HIPCHK(hipMemsetAsync(p, 0, nt * sizeof(double), c->st));
  for (float *p : {h->ro, h->cp, h->kt, h->rcp}) HIPCHK(hipMemsetAsync(p, 0, nt * sizeof(float), c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  c->heat = h;
  return STK_OK;
}

// level arrays between the reference's node order and the split-line
// layout of the temperature solve (even m2 of a line first, then odd m2)
template <typename T>
void to_split(const T *src, T *dst, int X, int Y, int Z) {
  const int hy = (Y + 1) / 2;
  for (long L = 0; L < (long)X * Z; L++)
    for (int m2 = 0; m2 < Y; m2++) dst[L * Y + ((m2 & 1) ? hy : 0) + (m2 >> 1)] = src[L * Y + m2];
}
template <typename T>
void from_split(const T *src, T *dst, int X, int Y, int Z) {
  const int hy = (Y + 1) / 2;
  for (long L = 0; L < (long)X * Z; L++)
    for (int m2 = 0; m2 < Y; m2++) dst[L * Y + m2] = src[L * Y + ((m2 & 1) ? hy : 0) + (m2 >> 1)];
}
inline long split_pos(const Level &L, long n) {
  const long line = n / L.Y;
  const int m2 = (int)(n - line * L.Y);
  return line * L.Y + ((m2 & 1) ? (L.Y + 1) / 2 : 0) + (m2 >> 1);
}

HeatGeo heat_geo(const stk_ctx *c, int l) {
  const Level &L = c->lv[l];
  const HeatLevel &H = c->heat->lv[l];
  HeatGeo g;
  g.X = L.X; g.Y = L.Y; g.Z = L.Z; g.XY = L.XY;
  g.hy = (L.Y + 1) / 2;
  g.fY = make_hdiv((unsigned)L.Y);
  g.fX = make_hdiv((unsigned)L.X);
  g.xper = c->per[0];
  for (int a = 0; a < 3; a++) { g.kf[a] = H.kf[a]; g.dm[a] = H.dm[a]; g.dp[a] = H.dp[a]; }
  return g;
}

// Tables of levels 0..Lt: spacings of the heat rows (heat3mg.c:21-24 and the
// denominators of 216-243), boundary flags, and the boundary rows sorted by
// dependency depth (a row after its partner).  Every node without a
// boundary row must have its six neighbours on the grid (x line 0 of an
// x-periodic box wraps to line X-2): the reference reads outside the level
// otherwise.
int heat_build(stk_ctx *c, int Lt) {
  HeatState *h = c->heat;
  if (!c->heat_dirty && h->built >= Lt) return STK_OK;
  for (int l = 0; l < MAXLV; l++) h->lv[l].release();
  h->built = -1;
  for (int l = 0; l <= Lt; l++) {
    const Level &L = c->lv[l];
    HeatLevel &H = h->lv[l];
    if (L.n >= (1l << 31)) return STK_EINVAL;  // node indices of the sweeps are 31-bit
    const int dims[3] = {L.X, L.Y, L.Z};
    for (int a = 0; a < 3; a++) {
      const int n = dims[a];
      const std::vector<double> &g = L.g[a];
      std::vector<double> kf(n, 1.0), dm(n, 1.0), dp(n, 1.0);
      for (int m = 1; m + 1 < n; m++) {
        kf[m] = 2.0 / (g[m + 1] - g[m - 1]);
        dm[m] = g[m] - g[m - 1];
        dp[m] = g[m + 1] - g[m];
      }
      if (a == 0 && c->per[0]) {
        kf[0] = 2.0 / (g[1] + g[n - 1] - g[n - 2]);
        dm[0] = g[n - 1] - g[n - 2];
        dp[0] = g[1] - g[0];
      }
      if (upload(&H.kf[a], kf) || upload(&H.dm[a], dm) || upload(&H.dp[a], dp)) return STK_ENOMEM;
    }
    std::vector<unsigned char> isbc(L.n, 0);  // reference order here, split-line on the device
    std::vector<int> rows;
    for (int m3 = 0; m3 < L.Z; m3++)
      for (int m1 = 0; m1 < L.X; m1++)
        for (int m2 = 0; m2 < L.Y; m2++) {
          const long n = (long)m3 * L.XY + (long)m1 * L.Y + m2;
          if (c->tbondm[L.off + n]) {
            isbc[n] = 1;
            rows.push_back((int)n);
            continue;
          }
          const bool x_ok = (m1 >= 1 || c->per[0]) && m1 <= L.X - 2;
          if (!x_ok || m2 < 1 || m2 > L.Y - 2 || m3 < 1 || m3 > L.Z - 2) return STK_EINVAL;
        }
    // partner and depth of every boundary row
    const int nb = (int)rows.size();
    std::vector<int> part(nb, -1), depth(nb, 0);
    for (int i = 0; i < nb; i++) {
      const int b = c->tbondm[L.off + rows[i]];
      if (c->tbn1[b]) {
        const long j = (long)c->tbn1[b] - 1 - L.off;
        if (j < 0 || j >= L.n) return STK_EINVAL;
        part[i] = (int)j;
      }
    }
    for (int i = 0; i < nb; i++) {
      long r = rows[i];
      int d = 0;
      for (;;) {
        const int b = c->tbondm[L.off + r];
        if (!c->tbn1[b]) break;
        const long j = (long)c->tbn1[b] - 1 - L.off;
        if (!isbc[j]) break;
        if (++d > 64) return STK_EINVAL;  // boundary rows that refer to each other in a cycle
        r = j;
      }
      depth[i] = d;
    }
    std::vector<int> order(nb);
    for (int i = 0; i < nb; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return depth[a] < depth[b]; });
    std::vector<int> br(nb), bj(nb);
    std::vector<double> bnk(nb), bk(nb), bcv(nb);
    const int maxd = nb ? depth[order[nb - 1]] : -1;
    H.bcoff.assign(maxd + 2, nb);
    for (int d = 0, q = 0; d <= maxd; d++) {
      while (q < nb && depth[order[q]] < d) q++;
      H.bcoff[d] = q;
    }
    for (int q = 0; q < nb; q++) {
      const int i = order[q];
      const int b = c->tbondm[L.off + rows[i]];
      br[q] = (int)split_pos(L, rows[i]);
      bj[q] = part[i] < 0 ? -1 : (int)split_pos(L, part[i]);
      bk[q] = (double)c->tbv1[2 * b + 1];
      bnk[q] = -(double)c->tbv1[2 * b + 1];
      bcv[q] = (double)c->tbv1[2 * b];
    }
    H.nbc = nb;
    std::vector<unsigned char> isbc_s(L.n);
    to_split(isbc.data(), isbc_s.data(), L.X, L.Y, L.Z);
    if (upload(&H.isbc, isbc_s) || upload(&H.bc_r, br) || upload(&H.bc_j, bj) || upload(&H.bc_negk, bnk) ||
        upload(&H.bc_k, bk) || upload(&H.bc_c, bcv))
      return STK_ENOMEM;
    H.split0 = c->per[0] && (L.X % 2 == 0);
  }
  h->built = Lt;
  c->heat_dirty = false;
  return STK_OK;
}

// am Gauss-Seidel sweeps of level l: colour 0, colour 1, boundary rows by depth
void heat_gs(stk_ctx *c, int l, int sweeps) {
  HeatState *h = c->heat;
  const Level &L = c->lv[l];
  const HeatLevel &H = h->lv[l];
  const HeatGeo g = heat_geo(c, l);
  const float *kt = h->kt + L.off, *rcp = h->rcp + L.off;
  const double *b = h->b + L.off;
  double *x = h->x + L.off;
  const int hy = (L.Y + 1) / 2, cut = H.split0 ? 1 : 0;
  const HDiv fhy = make_hdiv((unsigned)hy), f1 = make_hdiv(1u), fn = make_hdiv((unsigned)(L.X - cut));
  for (int s = 0; s < sweeps; s++) {
    for (int col = 0; col < 2; col++) {
      if (cut)
        hipLaunchKernelGGL(k_heat_gs, dim3(blocks((long)hy * L.Z, 256)), dim3(256), 0, c->st, g, col, 0, 1, fhy, f1,
                           H.isbc, kt, rcp, h->dt, h->t0, b, x);
      hipLaunchKernelGGL(k_heat_gs, dim3(blocks((long)hy * (L.X - cut) * L.Z, 256)), dim3(256), 0, c->st, g, col,
                         cut, L.X - cut, fhy, fn, H.isbc, kt, rcp, h->dt, h->t0, b, x);
    }
    const int nd = (int)H.bcoff.size() - 1;
    if (H.nbc && H.nbc <= kHeatBcSmall && nd <= kHeatBcDepths) {
      HeatBcDepths D;
      D.nd = nd;
      for (int d = 0; d <= nd; d++) D.off[d] = H.bcoff[d];
      hipLaunchKernelGGL(k_heat_bc_small, dim3(1), dim3(1024), 0, c->st, D, H.bc_r, H.bc_j, H.bc_negk, b, x, h->t0);
    } else {
      for (int d = 0; d < nd; d++) {
        const int lo = H.bcoff[d], hi = H.bcoff[d + 1];
        if (hi > lo)
          hipLaunchKernelGGL(k_heat_bcpass, dim3(blocks(hi - lo, 256)), dim3(256), 0, c->st, lo, hi, H.bc_r,
                             H.bc_j, H.bc_negk, b, x, h->t0);
      }
    }
  }
}

void heat_res(stk_ctx *c, int l) {
  HeatState *h = c->heat;
  const Level &L = c->lv[l];
  const HeatLevel &H = h->lv[l];
  hipLaunchKernelGGL(k_heat_res, dim3(blocks(L.n, 256)), dim3(256), 0, c->st, heat_geo(c, l), H.isbc,
                     h->kt + L.off, h->rcp + L.off, h->dt, h->b + L.off, h->x + L.off, h->r + L.off);
  if (H.nbc)
    hipLaunchKernelGGL(k_heat_bcres, dim3(blocks(H.nbc, 256)), dim3(256), 0, c->st, H.nbc, H.bc_r, H.bc_j,
                       H.bc_negk, h->b + L.off, h->x + L.off, h->r + L.off);
}

CompTab comp_tab(const Level &C, int kind) {
  CompTab t;
  for (int a = 0; a < 3; a++) t.ax[a] = C.ax[kind][a].tab();
  return t;
}

void heat_restrict(stk_ctx *c, int l) {
  HeatState *h = c->heat;
  const Level &F = c->lv[l - 1], &C = c->lv[l];
  hipLaunchKernelGGL(k_heat_restrict, dim3(blocks(C.n, 256)), dim3(256), 0, c->st, F.X, F.Y, C.X, C.Y, C.Z,
                     comp_tab(C, 0), h->lv[l].isbc, h->r + F.off, h->b + C.off, h->x + C.off);
}

void heat_prolong(stk_ctx *c, int l, int ynreset) {
  HeatState *h = c->heat;
  const Level &F = c->lv[l - 1], &C = c->lv[l];
  hipLaunchKernelGGL(k_heat_prolong, dim3(blocks(F.n, 256)), dim3(256), 0, c->st, F.X, F.Y, F.Z, C.X, C.Y,
                     comp_tab(C, 0), h->lv[l - 1].isbc, h->x + C.off, h->x + F.off, h->t1, ynreset);
}

void heat_props(stk_ctx *c, int Lt) {
  HeatState *h = c->heat;
  for (int l = 1; l <= Lt; l++) {
    const Level &F = c->lv[l - 1], &C = c->lv[l];
    hipLaunchKernelGGL(k_heat_props, dim3(blocks(C.n, 256)), dim3(256), 0, c->st, F.X, F.Y, C.X, C.Y, C.Z,
                       comp_tab(C, 3), h->ro + F.off, h->cp + F.off, h->kt + F.off, h->ro + C.off, h->cp + C.off,
                       h->kt + C.off, h->rcp + C.off);
  }
}

// tk0 = tk, sol0 = tk, level-0 right-hand side for a sub-step of dt
void heat_start(stk_ctx *c, double dt) {
  HeatState *h = c->heat;
  const Level &F = c->lv[0];
  const HeatLevel &H = h->lv[0];
  h->dt = dt;
  hipLaunchKernelGGL(k_heat_rhs, dim3(blocks(F.n, 256)), dim3(256), 0, c->st, F.n, H.isbc, h->tk, h->tk0, h->tk1,
                     h->rcp, dt, h->x, h->b);
  if (H.nbc)
    hipLaunchKernelGGL(k_heat_bcc, dim3(blocks(H.nbc, 256)), dim3(256), 0, c->st, H.nbc, H.bc_r, H.bc_c, h->b);
}

// sums of nb block partials (4 per block) in block order
int heat_partials(stk_ctx *c, int nb, double out4[4]) {
  HeatState *h = c->heat;
  D2H(c, h->hpart.data(), h->part, (size_t)4 * nb * sizeof(double));
  out4[0] = out4[1] = 0.0;
  out4[2] = 1e30;
  out4[3] = -1e30;
  for (int i = 0; i < nb; i++) {
    out4[0] += h->hpart[4 * i];
    out4[1] += h->hpart[4 * i + 1];
    out4[2] = std::min(out4[2], h->hpart[4 * i + 2]);
    out4[3] = std::max(out4[3], h->hpart[4 * i + 3]);
  }
  return STK_OK;
}

// tk = (float) sol0, then the reference's errors (heat3mg.c:507-553)
int heat_errs(stk_ctx *c, double *hs, double *bs, double *hn, double *bn, double *tmin, double *tmax) {
  HeatState *h = c->heat;
  const Level &F = c->lv[0];
  const HeatLevel &H = h->lv[0];
  hipLaunchKernelGGL(k_heat_tk, dim3(blocks(F.n, 256)), dim3(256), 0, c->st, F.n, h->x, h->tk);
  const int nb = (int)blocks(F.n, 256);
  hipLaunchKernelGGL(k_heat_err, dim3(nb), dim3(256), 0, c->st, heat_geo(c, 0), H.isbc, h->kt, h->rcp, h->tk,
                     h->tk0, h->tk1, h->dt, h->part);
  double e[4];
  int rc = heat_partials(c, nb, e);
  if (rc) return rc;
  *hs = e[1] > 0 ? std::sqrt(e[0] / e[1]) : 0.0;
  if (hn) *hn = e[1];
  if (tmin) *tmin = e[2];
  if (tmax) *tmax = e[3];
  double q[4] = {0, 0, 0, 0};
  if (H.nbc) {
    const int nq = (int)blocks(H.nbc, 256);
    hipLaunchKernelGGL(k_heat_bonderr, dim3(nq), dim3(256), 0, c->st, H.nbc, H.bc_r, H.bc_j, H.bc_c, H.bc_k, h->x,
                       h->part);
    rc = heat_partials(c, nq, q);
    if (rc) return rc;
  }
  *bs = q[1] > 0 ? std::sqrt(q[0] / q[1]) : 0.0;
  if (bn) *bn = q[1];
  return STK_OK;
}

// largest heating rate |tk1 / rho cp| over the level-0 heat rows
int heat_rate(stk_ctx *c, double *rate) {
  HeatState *h = c->heat;
  const Level &F = c->lv[0];
  const int nb = (int)blocks(F.n, 256);
  hipLaunchKernelGGL(k_heat_rate, dim3(nb), dim3(256), 0, c->st, F.n, h->lv[0].isbc, h->tk1, h->rcp, h->part);
  double e[4];
  int rc = heat_partials(c, nb, e);
  *rate = e[3] > 0 ? e[3] : 0.0;
  return rc;
}

// one V-cycle (heat3mg.c:463-492): up, smooth pre[l] and restrict the
// residual; down, smooth post[l] and prolong; level 0 is not smoothed after
// the last prolongation; one level: pre[0] sweeps
void heat_cycle(stk_ctx *c, const stk_heat_opts *o) {
  const int Lt = o->levels;
  if (Lt > 0) {
    for (int l = 0; l < Lt; l++) {
      heat_gs(c, l, o->pre[l]);
      heat_res(c, l);
      heat_restrict(c, l + 1);
    }
    for (int l = Lt; l > 0; l--) {
      heat_gs(c, l, o->post[l]);
      heat_prolong(c, l, 0);
    }
  } else {
    heat_gs(c, 0, o->pre[0]);
  }
}

int heat_ready(stk_ctx *c, int Lt) {
  if (!c || !c->heat || !c->heat->fields) return STK_ESTATE;
  if (Lt < 0 || Lt > c->L) return STK_EINVAL;
  return heat_build(c, Lt);
}

}  // namespace

extern "C" {

int stk_create(stk_ctx **out, int nx, int ny, int nz, int levels, const double *gx, const double *gy,
               const double *gz, int xper, int yper, int zper) {
  if (!out || levels < 0 || levels >= MAXLV || nx < 5 || ny < 5 || nz < 5 || !gx || !gy || !gz)
    return STK_EINVAL;
  {
    // grid lines strictly increasing (every spacing is a divisor)
    const double *gl[3] = {gx, gy, gz};
    const int nl[3] = {nx, ny, nz};
    for (int a = 0; a < 3; a++)
      for (int i = 0; i + 1 < nl[a]; i++)
        if (!(gl[a][i + 1] > gl[a][i])) return STK_EINVAL;
  }
  stk_ctx *c = new stk_ctx();
  c->L = levels;
  c->per[0] = xper ? 1 : 0; c->per[1] = yper ? 1 : 0; c->per[2] = zper ? 1 : 0;
  const int pmask = c->per[0] | (c->per[1] << 1) | (c->per[2] << 2);
  {
    int dev = 0, lds = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lds > 0)
      c->lds_max = (size_t)lds;
  }
  const double *g0[3] = {gx, gy, gz};
  int nf[3] = {nx, ny, nz};
  long off = 0;
  long mgs = 1;
  int dims[3] = {nx, ny, nz};
  for (int l = 0; l <= levels; l++) {
    Level &L = c->lv[l];
    if (l > 0) {
      mgs *= 2;
      for (int a = 0; a < 3; a++) dims[a] = (dims[a] - 5) / 2 + 5;
    }
    L.X = dims[0]; L.Y = dims[1]; L.Z = dims[2]; L.XY = L.X * L.Y;
    L.per = pmask;
    L.n = (long)L.XY * L.Z;
    L.off = off;
    off += L.n;
    // coarse lines keep 3 outer lines per side (load3mg.c:2078-2190)
    for (int a = 0; a < 3; a++) {
      int nl = dims[a];
      std::vector<double> &gg = L.g[a];
      gg.assign(nl, 0.0);
      gg[0] = g0[a][0]; gg[1] = g0[a][1]; gg[2] = g0[a][2];
      gg[nl - 1] = g0[a][nf[a] - 1]; gg[nl - 2] = g0[a][nf[a] - 2]; gg[nl - 3] = g0[a][nf[a] - 3];
      for (int n2 = 3; n2 < nl - 3; n2++) gg[n2] = g0[a][(n2 - 2) * mgs + 2];
    }
  }
  c->ntot = off;
  c->bondm.assign(4 * off, 0);
  c->bondv1.assign(2, 0.f);
  c->bondn1.assign(1, 0);
  c->tbondm.assign(off, 0);
  c->tbv1.assign(2, 0.f);
  c->tbn1.assign(1, 0);
  c->kn = stk_knobs{0.10, 0.0, 0.10, 1.0, 1.0, 1.0, 1e4, 0, 0.0, 0.0, 0.0, 9e-3, 5e1, 0.0, 3e-3};
  {
    RtLock lk(g_rt_mu);
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) { delete c; return STK_EHIP; }
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
      (void)hipStreamDestroy(c->st);
      delete c;
      return STK_EHIP;
    }
  }
  // per-level geometry tables
  for (int l = 0; l <= levels; l++) {
    Level &L = c->lv[l];
    for (int a = 0; a < 3; a++) {
      int nl = (int)L.g[a].size();
      std::vector<double> iA(nl, 0.0), iB(nl, 0.0);
      for (int i = 1; i < nl; i++) iA[i] = 1.0 / (L.g[a][i] - L.g[a][i - 1]);
      for (int i = 1; i + 1 < nl; i++) iB[i] = 1.0 / (L.g[a][i + 1] - L.g[a][i - 1]);
      if (c->per[a]) {
        // across the wrap: cell 0 is cell nl-1, and the spacing at line 0
        // (line nl-1) spans lines nl-2 and 1 (move3mg.c:17-19 with the
        // periodic cell, 400-403, 834), summed in the reference's order
        const std::vector<double> &gg = L.g[a];
        iA[0] = iA[nl - 1];
        iB[0] = 1.0 / (gg[1] - gg[0] + gg[nl - 1] - gg[nl - 2]);
        iB[nl - 1] = iB[0];
      }
      std::vector<float> fA(iA.begin(), iA.end()), fB(iB.begin(), iB.end());
      // uniform along this axis: one iA over 1..nl-1 and one iB over 1..nl-2
      // (and over the wrap entries of a periodic axis)
      bool u = nl >= 3;
      for (int i = 2; u && i < nl; i++) u = fA[i] == fA[1];
      for (int i = 2; u && i + 1 < nl; i++) u = fB[i] == fB[1];
      if (u && c->per[a]) u = fA[0] == fA[1] && fB[0] == fB[1] && fB[nl - 1] == fB[1];
      L.ua[a] = u ? fA[1] : 0.f;
      L.ub[a] = u ? fB[1] : 0.f;
      if (upload(&L.iA64[a], iA) || upload(&L.iB64[a], iB) || upload(&L.iA32[a], fA) || upload(&L.iB32[a], fB) ||
          upload(&L.g64[a], L.g[a])) {
        stk_destroy(c);
        return STK_ENOMEM;
      }
    }
    for (int k = 0; k < 5; k++)
      if (dalloc(&L.e64[k], L.n) || dalloc(&L.e32[k], L.n)) { stk_destroy(c); return STK_ENOMEM; }
    if (dalloc(&L.flags, L.n)) { stk_destroy(c); return STK_ENOMEM; }
  }
  if (dalloc(&c->nu, off) || dalloc(&c->ro, off) || dalloc(&c->nd, off) || dalloc(&c->nxx, c->lv[0].n) ||
      dalloc(&c->nxy, c->lv[0].n) || dalloc(&c->nxz, c->lv[0].n) || dalloc(&c->nyz, c->lv[0].n)) {
    stk_destroy(c);
    return STK_ENOMEM;
  }
  hipMemsetAsync(c->nu, 0, off * 4, c->st); hipMemsetAsync(c->ro, 0, off * 4, c->st);
  hipMemsetAsync(c->nd, 0, off * 4, c->st);
  // stress-point viscosities no marker reaches read as 0 before the clamp
  for (float *p : {c->nxx, c->nxy, c->nxz, c->nyz}) hipMemsetAsync(p, 0, c->lv[0].n * 4, c->st);
  hipStreamSynchronize(c->st);
  // transfer tables
  for (int l = 1; l <= levels; l++) {
    Level &F = c->lv[l - 1], &C = c->lv[l];
    int nfa[3] = {F.X, F.Y, F.Z}, nca[3] = {C.X, C.Y, C.Z};
    for (int kind = 0; kind < 4; kind++)
      for (int a = 0; a < 3; a++) {
        AxisHost h = build_axis(kind, nfa[a], nca[a], F.g[a], C.g[a]);
        if (upload_axis(C.ax[kind][a], h)) { stk_destroy(c); return STK_ENOMEM; }
      }
  }
  *out = c;
  return STK_OK;
}

void stk_destroy(stk_ctx *c) {
  if (!c) return;
  if (g_bound == c) g_bound = nullptr;
  if (c->st) (void)hipStreamSynchronize(c->st);  // all of the context's work runs on its stream
  for (int l = 0; l <= c->L; l++) {
    Level &L = c->lv[l];
    for (int a = 0; a < 3; a++) { dfree(L.iA64[a]); dfree(L.iB64[a]); dfree(L.iA32[a]); dfree(L.iB32[a]); dfree(L.g64[a]); }
    for (int k = 0; k < 5; k++) { dfree(L.e64[k]); dfree(L.e32[k]); }
    dfree(L.lc32); dfree(L.rs_t1); dfree(L.rs_t2); dfree(L.rs_ws);
    L.free_tile_chains();
    dfree(L.flags);
    dfree(L.bc_r); dfree(L.bc_j); dfree(L.bc_negk); dfree(L.bc_c);
    dfree(L.ch_r); dfree(L.ch_len); dfree(L.ch_j); dfree(L.ch_k);
    for (int k = 0; k < 2; k++) { dfree(L.x32[k]); dfree(L.x64[k]); }
    dfree(L.b32); dfree(L.r32); dfree(L.b64); dfree(L.r64);
    for (int k = 0; k < 3; k++) dfree(L.ref[k]);
    for (int k = 0; k < 4; k++)
      for (int a = 0; a < 3; a++) L.ax[k][a].release();
    for (int q = 0; q < 7; q++) { dfree(L.k64[q]); dfree(L.k32[q]); }
    dfree(L.kc64); dfree(L.kc32);
    dfree(L.kx32); dfree(L.kb32); dfree(L.kr32); dfree(L.kx64); dfree(L.kb64); dfree(L.kr64);
    dfree(L.ell_cnt); dfree(L.ell_col); dfree(L.ell_bcrows); dfree(L.ell_val); dfree(L.ell_x); dfree(L.ell_f);
    dfree(L.ell_f0); dfree(L.ell_isbc); dfree(L.ell_nd);
    for (int a = 0; a < 3; a++) { dfree(L.klo[a]); dfree(L.kcnt[a]); dfree(L.kpar[a]); }
  }
  dfree(c->dinv64); dfree(c->bfy64); dfree(c->bfw64); dfree(c->dinv32); dfree(c->bfy32); dfree(c->bfw32);
  dfree(c->kcell); dfree(c->kinvT64); dfree(c->kinvT32);
  dfree(c->nu); dfree(c->ro); dfree(c->nd); dfree(c->nxx); dfree(c->nxy); dfree(c->nxz); dfree(c->nyz);
  dfree(c->ainv64); dfree(c->ainvP64); dfree(c->ainv32); dfree(c->cact); dfree(c->cinv);
  dfree(c->X); dfree(c->B); dfree(c->R); dfree(c->W2); dfree(c->T1); dfree(c->red); dfree(c->scal);
  dfree(c->wu); dfree(c->ww); dfree(c->wz);
  dfree(c->post_part); dfree(c->sig); dfree(c->epsr); dfree(c->adv_bins);
  heat_free(c->heat);
  c->heat = nullptr;
  dfree(c->kred); dfree(c->hcol); dfree(c->gin); dfree(c->gout);
  RtLock lk(g_rt_mu);
  if (c->hpin) hipHostFree(c->hpin);
  if (c->gexec) hipGraphExecDestroy(c->gexec);
  if (c->graph) hipGraphDestroy(c->graph);
  for (auto p : c->Zs) dfree(p);
  for (auto p : c->Qs) dfree(p);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  if (c->st) hipStreamDestroy(c->st);
  delete c;
}

int stk_level_info(const stk_ctx *c, int l, int out6[6]) {
  if (!c || l < 0 || l > c->L) return STK_EINVAL;
  const Level &L = c->lv[l];
  out6[0] = L.X; out6[1] = L.Y; out6[2] = L.Z; out6[3] = L.XY; out6[4] = (int)L.n; out6[5] = (int)L.off;
  return STK_OK;
}

int stk_set_knobs(stk_ctx *c, const stk_knobs *k) {
  if (!c || !k || k->viscmod < 0 || k->viscmod > 2 || !(k->nubeg > 0) || !(k->nuend >= k->nubeg)) return STK_EINVAL;
  c->graph_ok = c->graph_ok && !memcmp(&c->kn, k, sizeof *k);  // relaxation weights are captured by value
  c->kn = *k;
  c->prepared = false;
  return STK_OK;
}

void stk_default_opts(stk_solver_opts *o, int levels) {
  memset(o, 0, sizeof(*o));
  o->precision = 0;
  o->restart = 128;
  // per level from the finest; measured over random-field, inclusion and sinker boxes (33^3..257^3)
  int pre[12] = {3, 6, 8, 8, 12, 12, 12, 12, 12, 12, 12, 12};
  for (int l = 0; l < 12; l++) { o->pre[l] = pre[l]; o->post[l] = pre[l]; }
  (void)levels;
  o->tile[0] = o->tile[1] = o->tile[2] = 0;  // per-level autotuned tile, else 32x4x4
  o->omega = 4.0 / 3.0;  // isoviscous Schur constant of the 2eta(e - div/3) stress form
  o->cycle = 1;
  o->coarse_viscmod = -1;
  o->tol = 1e-8;
  o->maxit = 800;
  o->schur = 1;
  o->krylov = 1;  // GCR can stagnate when the preconditioned operator is indefinite (sharp strong inclusions)
  // 3 red/black sweeps per K-cycle level and coarse corrections scaled by 1.8 (piecewise-constant
  // aggregation undershoots by about 2): measured against 2 and 2.0 over random-field, inclusion and
  // sinker boxes from 33^3 to 257^3 (257^3 spheres: 153 -> 104 outer iterations)
  o->ksweeps = 3;
  o->graph = 0;  // replay gains ~5% on small levels
  o->kscale = 1.8;
  o->basis = 1;  // FP32 Krylov vectors in mixed precision (half the Gram-Schmidt traffic)
}

// boundary grammar line projected on every level (parallel/in3mg.c:299-375)
int stk_bc_line(stk_ctx *c, int var, int m10, int m11, int m20, int m21, int m30, int m31, float cnst,
                float koef, int dm1, int dm2, int dm3) {
  if (!c || var < 0 || var > 4) return STK_EINVAL;
  const Level &F = c->lv[0];
  long mgs = 1;
  for (int l = 0; l <= c->L; l++, mgs *= 2) {
    const Level &L = c->lv[l];
    int a1 = lvl_index(m10, F.X, L.X, mgs), b1 = lvl_index(m11, F.X, L.X, mgs);
    int a2 = lvl_index(m20, F.Y, L.Y, mgs), b2 = lvl_index(m21, F.Y, L.Y, mgs);
    int a3 = lvl_index(m30, F.Z, L.Z, mgs), b3 = lvl_index(m31, F.Z, L.Z, mgs);
    for (int m3 = a3; m3 <= b3; m3++)
      for (int m1 = a1; m1 <= b1; m1++)
        for (int m2 = a2; m2 <= b2; m2++) {
          if (m1 < 0 || m1 >= L.X || m2 < 0 || m2 >= L.Y || m3 < 0 || m3 >= L.Z) return STK_EINVAL;
          long m4 = L.off + (long)m3 * L.XY + (long)m1 * L.Y + m2;
          int sx = dm1, sy = dm2, sz = dm3;
          if (std::abs(dm1) > 1) sx = -(L.X - 1);
          if (std::abs(dm2) > 1) sy = -(L.Y - 1);
          if (std::abs(dm3) > 1) sz = -(L.Z - 1);
          long m6 = m4 + (long)sz * L.XY + (long)sx * L.Y + sy;
          if (var == 4) {
            // temperature rows (in3mg.c:369-385): partner node + 1
            int nt = (int)c->tbn1.size();
            c->tbondm[m4] = nt;
            c->tbv1.push_back(cnst);
            c->tbv1.push_back((dm1 || dm2 || dm3) ? koef : 0.f);
            c->tbn1.push_back((dm1 || dm2 || dm3) ? (int)(m6 + 1) : 0);
            continue;
          }
          int nb = (int)c->bondn1.size();
          c->bondm[m4 * 4 + var] = nb;
          c->bondv1.push_back(cnst);
          if (dm1 || dm2 || dm3) {
            c->bondv1.push_back(koef);
            c->bondn1.push_back((int)(m6 * 4 + var + 1));
          } else {
            c->bondv1.push_back(0.f);
            c->bondn1.push_back(0);
          }
        }
  }
  if (var == 4) {
    c->heat_dirty = true;
    return STK_OK;
  }
  c->prepared = false;
  c->rows_dirty = true;
  return STK_OK;
}

int stk_set_bc_tables(stk_ctx *c, const int *bondm, int nbond, const float *bondv1, const int *bondn1) {
  if (!c || !bondm || nbond < 0 || (nbond && (!bondv1 || !bondn1))) return STK_EINVAL;
  c->bondm.assign(bondm, bondm + 4 * c->ntot);
  c->bondv1.assign(bondv1, bondv1 + 2 * (nbond + 1));
  c->bondn1.assign(bondn1, bondn1 + nbond + 1);
  for (long r = 0; r < 4 * c->ntot; r++)
    if (c->bondm[r] < 0 || c->bondm[r] > nbond) return STK_EINVAL;
  c->prepared = false;
  c->rows_dirty = true;
  return STK_OK;
}

int stk_get_bc_tables(const stk_ctx *c, int *bondm, int *nbond, float *bondv1, int *bondn1) {
  if (!c || !nbond) return STK_EINVAL;
  *nbond = (int)c->bondn1.size() - 1;
  if (bondm) memcpy(bondm, c->bondm.data(), c->bondm.size() * sizeof(int));
  if (bondv1) memcpy(bondv1, c->bondv1.data(), c->bondv1.size() * sizeof(float));
  if (bondn1) memcpy(bondn1, c->bondn1.data(), c->bondn1.size() * sizeof(int));
  return STK_OK;
}

int stk_set_fields(stk_ctx *c, const float *nu, const float *ro, const float *nxx, const float *nxy,
                   const float *nxz, const float *nyz) {
  if (!c || !nu || !ro || !nxx || !nxy || !nxz || !nyz) return STK_EINVAL;
  long n0 = c->lv[0].n;
  for (long i = 0; i < n0; i++)
    if (!(nu[i] > 0) || !std::isfinite(nu[i]) || !std::isfinite(ro[i]) || !(nxx[i] > 0) || !(nxy[i] > 0) ||
        !(nxz[i] > 0) || !(nyz[i] > 0))
      return STK_EINVAL;
  H2D(c, c->nu, nu, n0 * 4);
  H2D(c, c->ro, ro, n0 * 4);
  H2D(c, c->nxx, nxx, n0 * 4);
  H2D(c, c->nxy, nxy, n0 * 4);
  H2D(c, c->nxz, nxz, n0 * 4);
  H2D(c, c->nyz, nyz, n0 * 4);
  c->fields_set = true;
  c->prepared = false;
  return STK_OK;
}

int stk_get_fields(stk_ctx *c, float *nu, float *ro, float *nd, float *nxx, float *nxy, float *nxz, float *nyz) {
  if (!c) return STK_EINVAL;
  long n0 = c->lv[0].n;
  HIPCHK(hipStreamSynchronize(c->st));
  if (nu) D2H(c, nu, c->nu, c->ntot * 4);
  if (ro) D2H(c, ro, c->ro, c->ntot * 4);
  if (nd) D2H(c, nd, c->nd, c->ntot * 4);
  if (nxx) D2H(c, nxx, c->nxx, n0 * 4);
  if (nxy) D2H(c, nxy, c->nxy, n0 * 4);
  if (nxz) D2H(c, nxz, c->nxz, n0 * 4);
  if (nyz) D2H(c, nyz, c->nyz, n0 * 4);
  return STK_OK;
}

int stk_markers_to_nodes(stk_ctx *c, long nmark, const float *mx, const float *my, const float *mz,
                         const float *mnu, const float *mro, double nukoef) {
  if (!c || nmark <= 0 || nmark > 0x7ffffff0L || !mx || !my || !mz || !mnu || !mro) return STK_EINVAL;
  Level &F = c->lv[0];
  const int ncell = (F.X - 1) * (F.Y - 1) * (F.Z - 1);
  MarkerBufs b;
  size_t tmpb = 0;
  int rc = STK_OK;
  for (int k = 0; k < 5 && !rc; k++) rc = dalloc(&b.in[k], nmark) || dalloc(&b.so[k], nmark) ? STK_ENOMEM : 0;
  if (!rc) rc = (dalloc(&b.key, nmark) || dalloc(&b.val, nmark) || dalloc(&b.skey, nmark) || dalloc(&b.sval, nmark) ||
                 dalloc(&b.start, ncell + 2)) ? STK_ENOMEM : 0;
  if (!rc && hipcub::DeviceRadixSort::SortPairs(nullptr, tmpb, b.key, b.skey, b.val, b.sval, (int)nmark, 0, 32,
                                                c->st) != hipSuccess)
    rc = STK_EHIP;
  if (!rc && dalloc((unsigned char **)&b.tmp, tmpb)) rc = STK_ENOMEM;
  b.tmpb = tmpb;
  if (!rc) rc = markers_run(c, nmark, mx, my, mz, mnu, mro, nukoef, b);
  for (int k = 0; k < 5; k++) { dfree(b.in[k]); dfree(b.so[k]); }
  dfree(b.key); dfree(b.val); dfree(b.skey); dfree(b.sval); dfree(b.start);
  if (b.tmp) {
    RtLock lk(g_rt_mu);
    (void)hipFree(b.tmp);
  }
  return rc;
}

// The same averaging with every device buffer carved from this GPU's static
// scratch array (no hipMalloc, hipFree or allocating library call).
int stk_markers_to_nodes_noalloc(stk_ctx *c, long nmark, const float *mx, const float *my, const float *mz,
                                 const float *mnu, const float *mro, double nukoef) {
  if (!c || nmark <= 0 || !mx || !my || !mz || !mnu || !mro) return STK_EINVAL;
  Level &F = c->lv[0];
  const long ncell = (long)(F.X - 1) * (F.Y - 1) * (F.Z - 1);
  if (nmark > kScratchMarkers || ncell > kScratchCells) {
    fprintf(stderr,
            "pardiso2_amd: stk_markers_to_nodes_noalloc: %ld markers and %ld cells exceed the scratch limit "
            "(%ld markers, %ld cells)\n",
            nmark, ncell, kScratchMarkers, kScratchCells);
    return STK_EINVAL;
  }
  unsigned char *base = nullptr;
  int rc = scratch_base(&base);
  if (rc) return rc;
  MarkerBufs b;
  size_t off = 0;
  for (int k = 0; k < 5; k++) { b.in[k] = (float *)(base + off); off += kScratchArr; }
  for (int k = 0; k < 5; k++) { b.so[k] = (float *)(base + off); off += kScratchArr; }
  int **ints[4] = {&b.key, &b.val, &b.skey, &b.sval};
  for (int k = 0; k < 4; k++) { *ints[k] = (int *)(base + off); off += kScratchArr; }
  b.start = (int *)(base + off);
  off += kScratchStart;
  b.tmp = base + off;
  size_t tmpb = 0;
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmpb, b.key, b.skey, b.val, b.sval, (int)nmark, 0, 32, c->st));
  if (tmpb > kScratchSort) {
    fprintf(stderr, "pardiso2_amd: stk_markers_to_nodes_noalloc: the sort needs %zu bytes, the scratch holds %zu\n",
            tmpb, kScratchSort);
    return STK_ENOMEM;
  }
  b.tmpb = tmpb;
  return markers_run(c, nmark, mx, my, mz, mnu, mro, nukoef, b);
}

// nesting of level-l cells in level-(l-1) cells along one axis: coarse line
// i lies on the fine line with the same coordinate, so coarse cell m (lines
// m-1, m) owns fine cells map[m-1]+1 .. map[m]
static int axis_nesting(const std::vector<double> &gf, const std::vector<double> &gc, std::vector<int> &lo,
                        std::vector<int> &cnt, std::vector<int> &par) {
  const int nf = (int)gf.size(), nc = (int)gc.size();
  std::vector<int> map(nc);
  int k = 0;
  for (int i = 0; i < nc; i++) {
    while (k < nf && gf[k] != gc[i]) k++;
    if (k == nf) return STK_EINVAL;
    map[i] = k;
  }
  if (map[0] != 0 || map[nc - 1] != nf - 1) return STK_EINVAL;
  lo.assign(nc, 0);
  cnt.assign(nc, 0);
  par.assign(nf, 0);
  for (int m = 1; m < nc; m++) {
    lo[m] = map[m - 1] + 1;
    cnt[m] = map[m] - map[m - 1];
    for (int f = lo[m]; f < lo[m] + cnt[m]; f++) par[f] = m;
  }
  return STK_OK;
}

// Schur approximation of the prepared operator: D^-1, K on level 0, its
// aggregated Galerkin hierarchy and the dense inverse of the coarsest K
// (regularised by s/N * 1 1^T, s the mean diagonal, so the constant pressure
// mode of the Neumann operator gets a bounded response)
static int build_schur(stk_ctx *c) {
  const int L = c->L;
  Level &F = c->lv[0];
  const long n4 = 4 * F.n;
  if (!c->dinv64) {
    if (dalloc(&c->dinv64, n4) || dalloc(&c->dinv32, n4) || dalloc(&c->bfy64, n4) || dalloc(&c->bfw64, n4) ||
        dalloc(&c->bfy32, n4) || dalloc(&c->bfw32, n4))
      return STK_ENOMEM;
  }
  for (int l = 0; l <= L; l++) {
    Level &V = c->lv[l];
    if (V.kx32) continue;
    for (int q = 0; q < 7; q++)
      if (dalloc(&V.k64[q], V.n) || dalloc(&V.k32[q], V.n)) return STK_ENOMEM;
    if (dalloc(&V.kx32, V.n) || dalloc(&V.kb32, V.n) || dalloc(&V.kr32, V.n) || dalloc(&V.kx64, V.n) ||
        dalloc(&V.kb64, V.n) || dalloc(&V.kr64, V.n))
      return STK_ENOMEM;
    HIPCHK(hipMemsetAsync(V.kx32, 0, V.n * 4, c->st));
    HIPCHK(hipMemsetAsync(V.kx64, 0, V.n * 8, c->st));
    if (l < L && kcolour_ok(V) && (dalloc(&V.kc64, 14 * kcolour_n(V)) || dalloc(&V.kc32, 14 * kcolour_n(V))))
      return STK_ENOMEM;
  }
  if (!c->ktabs) {
    for (int l = 1; l <= L; l++) {
      Level &P = c->lv[l - 1], &C = c->lv[l];
      for (int a = 0; a < 3; a++) {
        std::vector<int> lo, cnt, par;
        int rc = axis_nesting(P.g[a], C.g[a], lo, cnt, par);
        if (rc) return rc;
        if (upload(&C.klo[a], lo) || upload(&C.kcnt[a], cnt) || upload(&P.kpar[a], par)) return STK_ENOMEM;
      }
    }
    c->ktabs = true;
  }
  Geo<double> g = F.geo<double>(c->kn.p2koef);
  const unsigned nb = blocks(F.n, 256);
  hipLaunchKernelGGL(k_bf_dinv, dim3(nb), dim3(256), 0, c->st, g, F.flags, c->dinv64, c->dinv32);
  KPlanes<double> k64;
  KPlanes<float> k32;
  for (int q = 0; q < 7; q++) { k64.p[q] = F.k64[q]; k32.p[q] = F.k32[q]; }
  hipLaunchKernelGGL(k_bf_kop, dim3(nb), dim3(256), 0, c->st, g, F.flags, c->dinv64, k64, k32);
  for (int l = 1; l <= L; l++) {
    Level &P = c->lv[l - 1], &C = c->lv[l];
    KPlanes<double> kc;
    for (int q = 0; q < 7; q++) kc.p[q] = C.k64[q];
    hipLaunchKernelGGL(k_kagg<double>, dim3(blocks(C.n, 256)), dim3(256), 0, c->st, P.X, P.Y, C.X, C.Y, C.Z,
                       kagg(C), kplanes<double>(P), kc);
    for (int q = 0; q < 7; q++)
      hipLaunchKernelGGL((k_convert<double, float>), dim3(blocks(C.n, 256)), dim3(256), 0, c->st, C.n, C.k64[q],
                         C.k32[q]);
  }
  for (int l = 0; l < L; l++) {  // colour-compact copies for the Gauss-Seidel half-sweeps
    Level &V = c->lv[l];
    if (!V.kc64) continue;
    const FDiv dhy = make_fdiv((V.Y + 1) / 2), dx = make_fdiv(V.X);
    const unsigned nt = blocks(2 * kcolour_n(V), 256);
    hipLaunchKernelGGL(k_kplanes_colour<double>, dim3(nt), dim3(256), 0, c->st, V.X, V.Y, V.Z, dhy, dx,
                       kplanes<double>(V), V.kc64);
    hipLaunchKernelGGL(k_kplanes_colour<float>, dim3(nt), dim3(256), 0, c->st, V.X, V.Y, V.Z, dhy, dx,
                       kplanes<float>(V), V.kc32);
  }
  // coarsest K: dense rows assembled on the host (at most a few hundred cells), inverted on the device
  Level &C = c->lv[L];
  std::vector<double> h[7];
  HIPCHK(hipStreamSynchronize(c->st));
  for (int q = 0; q < 7; q++) {
    h[q].resize(C.n);
    D2H(c, h[q].data(), C.k64[q], C.n * 8);
  }
  std::vector<int> idx(C.n, -1), cell;
  for (long n = 0; n < C.n; n++)
    if (h[0][n] != 0.0) { idx[n] = (int)cell.size(); cell.push_back((int)n); }
  const int N = (int)cell.size();
  if (N == 0 || N > 4096) return STK_EINVAL;
  std::vector<double> A((size_t)N * N, 0.0), I((size_t)N * N, 0.0);
  double dmean = 0.0;
  for (int i = 0; i < N; i++) {
    const long n = cell[i];
    int m1, m2, m3;
    m3 = (int)(n / C.XY); m1 = (int)((n % C.XY) / C.Y); m2 = (int)(n % C.Y);
    A[(size_t)i * N + i] += h[0][n];
    dmean += h[0][n];
    long nbh[6] = {m1 > 0 ? n - C.Y : -1, m1 < C.X - 1 ? n + C.Y : -1, m2 > 0 ? n - 1 : -1,
                   m2 < C.Y - 1 ? n + 1 : -1, m3 > 0 ? n - C.XY : -1, m3 < C.Z - 1 ? n + C.XY : -1};
    // periodic axes: cells 1 and n-1 are neighbours across the wrap
    const int mm[3] = {m1, m2, m3}, nn[3] = {C.X, C.Y, C.Z};
    const long st[3] = {C.Y, 1, C.XY};
    for (int a = 0; a < 3; a++) {
      if (!((C.per >> a) & 1)) continue;
      if (mm[a] == 1) nbh[2 * a] = n + (long)(nn[a] - 2) * st[a];
      if (mm[a] == nn[a] - 1) nbh[2 * a + 1] = n - (long)(nn[a] - 2) * st[a];
    }
    for (int q = 0; q < 6; q++)
      if (nbh[q] >= 0 && idx[nbh[q]] >= 0) A[(size_t)i * N + idx[nbh[q]]] += h[q + 1][n];
    I[(size_t)i * N + i] = 1.0;
  }
  dmean /= N;
  for (auto &v : A) v += dmean / N;
  // inverse on the device (the same Gauss-Jordan as the velocity block's)
  {
    double *dA = nullptr, *dI = nullptr;
    if (dalloc(&dA, (size_t)N * N) || dalloc(&dI, (size_t)N * N)) { dfree(dA); dfree(dI); return STK_ENOMEM; }
    const size_t bytes = (size_t)N * N * 8;
    int rc = (hipMemcpyAsync(dA, A.data(), bytes, hipMemcpyHostToDevice, c->st) == hipSuccess &&
              hipStreamSynchronize(c->st) == hipSuccess) ? STK_OK : STK_EHIP;
    if (!rc) rc = gj_invert(c, dA, N, dI);
    if (!rc && (hipMemcpyAsync(I.data(), dI, bytes, hipMemcpyDeviceToHost, c->st) != hipSuccess ||
                hipStreamSynchronize(c->st) != hipSuccess))
      rc = STK_EHIP;
    dfree(dA);
    dfree(dI);
    if (rc) return rc;
  }
  std::vector<float> I32(I.begin(), I.end());
  if (N != c->nkc || !c->kinvT64) {  // same size: refill in place (captured graphs keep their pointers)
    dfree(c->kcell); dfree(c->kinvT64); dfree(c->kinvT32);
    if (dalloc(&c->kcell, N) || dalloc(&c->kinvT64, (size_t)N * N) || dalloc(&c->kinvT32, (size_t)N * N))
      return STK_ENOMEM;
    c->graph_ok = false;
  }
  {
    std::vector<double> IT((size_t)N * N);
    std::vector<float> IT32((size_t)N * N);
    for (long i = 0; i < N; i++)
      for (long j = 0; j < N; j++) {
        IT[j * N + i] = I[i * N + j];
        IT32[j * N + i] = I32[i * N + j];
      }
    H2D(c, c->kinvT64, IT.data(), (size_t)N * N * 8);
    H2D(c, c->kinvT32, IT32.data(), (size_t)N * N * 4);
  }
  H2D(c, c->kcell, cell.data(), (size_t)N * 4);
  c->nkc = N;
  return STK_OK;
}

// flags + boundary lists of every level from the reference tables
static int build_rows(stk_ctx *c) {
  c->graph_ok = false;  // boundary lists are reallocated
  for (int l = 0; l <= c->L; l++) {
    Level &L = c->lv[l];
    L.hflags.assign(L.n, 0);
    std::vector<long> bcs;
    for (int m3 = 0; m3 < L.Z; m3++)
      for (int m1 = 0; m1 < L.X; m1++)
        for (int m2 = 0; m2 < L.Y; m2++) {
          long n = (long)m3 * L.XY + (long)m1 * L.Y + m2;
          unsigned char f = 0;
          for (int cc = 0; cc < 4; cc++) {
            if (!comp_exists(L, m1, m2, m3, cc)) continue;
            long r = 4 * (L.off + n) + cc;
            if (c->bondm[r]) {
              f |= KIND_BC << (2 * cc);
              bcs.push_back(r);
            } else {
              if (!eq_inside(L, m1, m2, m3, cc)) {
                fprintf(stderr, "pardiso2_amd: level %d unknown (%d,%d,%d,%d) needs a boundary equation\n", l, m1,
                        m2, m3, cc);
                return STK_EBCFORM;
              }
              f |= KIND_EQ << (2 * cc);
            }
          }
          L.hflags[n] = f;
        }
    // equation-row box (the whole level if there are none)
    int bx[6] = {L.Y, L.X, L.Z, 0, 0, 0};
    for (int m3 = 0; m3 < L.Z; m3++)
      for (int m1 = 0; m1 < L.X; m1++)
        for (int m2 = 0; m2 < L.Y; m2++) {
          const unsigned char f = L.hflags[(long)m3 * L.XY + (long)m1 * L.Y + m2];
          bool eq = false;
          for (int cc = 0; cc < 4; cc++) eq = eq || kind_of(f, cc) == KIND_EQ;
          if (!eq) continue;
          const int m[3] = {m2, m1, m3};
          for (int a = 0; a < 3; a++) {
            bx[a] = std::min(bx[a], m[a]);
            bx[a + 3] = std::max(bx[a + 3], m[a] + 1);
          }
        }
    if (bx[3] == 0) { bx[0] = bx[1] = bx[2] = 0; bx[3] = L.Y; bx[4] = L.X; bx[5] = L.Z; }
    for (int a = 0; a < 6; a++) L.box[a] = bx[a];
    // dependency depth of each boundary row
    long base = 4 * L.off;
    std::vector<int> depth(bcs.size(), 0);
    int maxd = 0;
    for (size_t i = 0; i < bcs.size(); i++) {
      long r = bcs[i];
      int d = 0;
      while (true) {
        int b = c->bondm[r];
        if (!b || !c->bondn1[b]) break;
        long j = c->bondn1[b] - 1;
        if (j < base || j >= base + 4 * L.n) return STK_EINVAL;  // partner outside the level
        long nj = (j - base) / 4;
        int cj = (int)((j - base) & 3);
        if (kind_of(L.hflags[nj], cj) != KIND_BC) break;
        r = j;
        if (++d > 64) return STK_EINVAL;  // cyclic boundary equations
      }
      depth[i] = d;
      maxd = std::max(maxd, d);
    }
    L.hbc_r.clear(); L.hbc_j.clear(); L.hbc_negk.clear(); L.hbc_c.clear();
    L.bcoff.assign(1, 0);
    for (int d = 0; d <= maxd; d++) {
      for (size_t i = 0; i < bcs.size(); i++) {
        if (depth[i] != d) continue;
        long r = bcs[i];
        int b = c->bondm[r];
        L.hbc_r.push_back((int)(r - base));
        L.hbc_j.push_back(c->bondn1[b] ? (int)(c->bondn1[b] - 1 - base) : -1);
        L.hbc_negk.push_back((double)(-c->bondv1[2 * b + 1]));
        L.hbc_c.push_back((double)c->bondv1[2 * b + 0]);
      }
      L.bcoff.push_back((int)L.hbc_r.size());
    }
    L.nbc = (int)L.hbc_r.size();
    {
      std::vector<int> pos(4 * L.n, -1);
      for (int i = 0; i < L.nbc; i++) pos[L.hbc_r[i]] = i;
      L.ch_dm = maxd + 1;
      std::vector<int> cr((size_t)std::max(L.nbc, 1) * L.ch_dm, 0), cl(std::max(L.nbc, 1), 0), cj(std::max(L.nbc, 1), -1);
      std::vector<double> ck((size_t)std::max(L.nbc, 1) * L.ch_dm, 0.0);
      for (int i = 0; i < L.nbc; i++) {
        int p = i, t = 0, j;
        while (true) {
          cr[(size_t)i * L.ch_dm + t] = L.hbc_r[p];
          ck[(size_t)i * L.ch_dm + t] = -L.hbc_negk[p];
          t++;
          j = L.hbc_j[p];
          if (j < 0 || pos[j] < 0 || t >= L.ch_dm) break;
          p = pos[j];
        }
        cl[i] = t;
        cj[i] = j;
      }
      L.hch_r0.assign(L.nbc, 0);
      L.hch_j.assign(L.nbc, -1);
      for (int i = 0; i < L.nbc; i++) { L.hch_r0[i] = cr[(size_t)i * L.ch_dm]; L.hch_j[i] = cj[i]; }
      L.free_tile_chains();  // rebuilt on the next tile sweep
      dfree(L.ch_r); dfree(L.ch_len); dfree(L.ch_j); dfree(L.ch_k);
      if (upload(&L.ch_r, cr) || upload(&L.ch_len, cl) || upload(&L.ch_j, cj) || upload(&L.ch_k, ck))
        return STK_ENOMEM;
    }
    dfree(L.bc_r); dfree(L.bc_j); dfree(L.bc_negk); dfree(L.bc_c);
    if (upload(&L.bc_r, L.hbc_r) || upload(&L.bc_j, L.hbc_j) || upload(&L.bc_negk, L.hbc_negk) ||
        upload(&L.bc_c, L.hbc_c))
      return STK_ENOMEM;
    H2D(c, L.flags, L.hflags.data(), L.n);
  }
  return STK_OK;
}

int stk_prepare_cvm(stk_ctx *c, int cvm);
int stk_prepare(stk_ctx *c) { return stk_prepare_cvm(c, -1); }

int stk_prepare_cvm(stk_ctx *c, int cvm) {
  if (!c || cvm < -1 || cvm > 2) return STK_EINVAL;
  const int vmc = (cvm < 0) ? c->kn.viscmod : cvm;
  if (!c->fields_set) return STK_ESTATE;
  int rc = STK_OK;
  if (c->rows_dirty) {
    rc = build_rows(c);
    if (rc) return rc;
    c->rows_dirty = false;
  }
  // nu/ro restriction to every level (parallel/move3mg.c:2247)
  for (int l = 1; l <= c->L; l++) {
    Level &F = c->lv[l - 1], &C = c->lv[l];
    CompTab t;
    for (int a = 0; a < 3; a++) t.ax[a] = C.ax[3][a].tab();
    hipLaunchKernelGGL(k_ronurestrict, dim3(blocks(C.n, 256)), dim3(256), 0, c->st, F.X, F.Y, C.X, C.Y, C.Z, t,
                       c->nu + F.off, c->ro + F.off, c->nu + C.off, c->ro + C.off, vmc);
  }
  for (int l = 0; l <= c->L; l++) {
    Level &L = c->lv[l];
    hipLaunchKernelGGL(k_visc_prep, dim3(blocks(L.n, 256)), dim3(256), 0, c->st, L.X, L.Y, L.Z, l == 0 ? 1 : 0,
                       c->nu + L.off, c->nxx, c->nxy, c->nxz, c->nyz, c->kn.nubeg, c->kn.nuend,
                       l == 0 ? c->kn.viscmod : vmc,
                       L.e64[0], L.e64[1], L.e64[2], L.e64[3], L.e64[4], L.e32[0], L.e32[1], L.e32[2], L.e32[3],
                       L.e32[4], L.per);
    if (L.per)
      hipLaunchKernelGGL(k_visc_wrap, dim3(blocks(L.n, 256)), dim3(256), 0, c->st, L.X, L.Y, L.Z, L.per, L.e64[0],
                         L.e64[1], L.e64[2], L.e64[3], L.e64[4], L.e32[0], L.e32[1], L.e32[2], L.e32[3], L.e32[4]);
    L.uni = L.ua[0] > 0 && L.ua[1] > 0 && L.ua[2] > 0;
    if (L.uni) L.fu = flux_u_consts(L.ua[0], L.ub[0], L.ua[1], L.ub[1], L.ua[2], L.ub[2], c->kn.p2koef);
    // nd = cell viscosity of the continuity stencil (continerr, move3mg.c:1549)
    hipLaunchKernelGGL((k_convert<float, float>), dim3(blocks(L.n, 256)), dim3(256), 0, c->st, L.n, L.e32[0],
                       c->nd + L.off);
    // rows of the FP32 whole-level smoother
    if (!L.per && 4 * (size_t)L.n * sizeof(float) <= kLevelLds) {
      if (!L.lc32 && dalloc(&L.lc32, (size_t)54 * L.n)) return STK_ENOMEM;
      hipLaunchKernelGGL(k_level_coef, dim3(blocks(L.n, 256)), dim3(256), 0, c->st, L.geo<float>(c->kn.p2koef),
                         L.flags, L.lc32);
    }
    // separable FP32 restriction into this level: buffers and weight sums (flags only)
    if (l >= 1 && gather_width(L) <= kPad) {
      const Level &F = c->lv[l - 1];
      if (!L.rs_ws && (dalloc(&L.rs_t1, (size_t)F.Z * F.X * L.Y) || dalloc(&L.rs_t2, (size_t)F.Z * L.X * L.Y) ||
                       dalloc(&L.rs_ws, (size_t)L.n)))
        return STK_ENOMEM;
      rsep(c, l, nullptr, nullptr, nullptr, 1);
    }
  }
  rc = ensure_outer(c, 0);
  if (rc) return rc;
  Level &F = c->lv[0];
  hipLaunchKernelGGL(k_rhs, dim3(blocks(F.n, 256)), dim3(256), 0, c->st, F.X, F.Y, F.Z, F.flags, c->ro,
                     c->kn.gx, c->kn.gy, c->kn.gz, c->B, F.per);
  if (F.nbc) hipLaunchKernelGGL(k_bc_rhs, dim3(blocks(F.nbc, 256)), dim3(256), 0, c->st, F.nbc, F.bc_r, F.bc_c, c->B);
  hipLaunchKernelGGL(k_weights, dim3(blocks(F.n, 256)), dim3(256), 0, c->st, F.X, F.Y, F.Z, F.flags, F.e64[0],
                     F.g64[0], F.g64[1], F.g64[2], c->W2);
  rc = coarse_inverse_v(c);
  if (rc) return rc;
  dfree(c->ainvP64);
  rc = build_schur(c);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(c->st));
  c->prepared = true;
  c->cvm = vmc;
  return STK_OK;
}

int stk_set_solution(stk_ctx *c, const double *x) {
  if (!c || !x) return STK_EINVAL;
  int rc = ensure_outer(c, 0);
  if (rc) return rc;
  H2D(c, c->X, x, 4 * c->lv[0].n * 8);
  return STK_OK;
}

int stk_get_solution(stk_ctx *c, double *x) {
  if (!c || !x || !c->X) return STK_EINVAL;
  HIPCHK(hipStreamSynchronize(c->st));
  D2H(c, x, c->X, 4 * c->lv[0].n * 8);
  return STK_OK;
}

int stk_relres(stk_ctx *c, double *rel, double blk[3]) {
  if (!c || !c->prepared) return STK_ESTATE;
  true_residual(c);
  double s[4];
  int rc = blocks_norms(c, c->R, c->B, s);
  if (rc) return rc;
  double nb = std::sqrt(s[3]);
  if (rel) *rel = std::sqrt(s[0] + s[1] + s[2]) / (nb > 0 ? nb : 1.0);
  if (blk) { blk[0] = std::sqrt(s[0]); blk[1] = std::sqrt(s[1]); blk[2] = std::sqrt(s[2]); }
  return STK_OK;
}

static PostGeo post_geo(const stk_ctx *c);
static int post_norms(stk_ctx *c, const double *x, stk_norms *out);

int stk_norms_eval(stk_ctx *c, stk_norms *out) {
  if (!c || !out) return STK_EINVAL;
  if (!c->prepared || !c->X || !c->B) return STK_ESTATE;
  return post_norms(c, c->X, out);
}

int stk_stress_strain(stk_ctx *c, double *sig, double *eps) {
  if (!c) return STK_EINVAL;
  if (!c->prepared || !c->X) return STK_ESTATE;
  Level &F = c->lv[0];
  const size_t n6 = (size_t)6 * F.n;
  if (!c->sig && dalloc(&c->sig, n6)) return STK_ENOMEM;
  if (!c->epsr && dalloc(&c->epsr, n6)) return STK_ENOMEM;
  hipLaunchKernelGGL(k_post_stress, dim3(blocks(F.n, 256)), dim3(256), 0, c->st, post_geo(c), c->X, c->sig, c->epsr);
  if (sig) HIPCHK(hipMemcpyAsync(sig, c->sig, n6 * sizeof(double), hipMemcpyDeviceToHost, c->st));
  if (eps) HIPCHK(hipMemcpyAsync(eps, c->epsr, n6 * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return STK_OK;
}

// interpolation geometry of level 0 (vxyzcalc1 / epscalc1); the cell-search
// bins are built once per context from the level-0 lines
static int adv_geo(stk_ctx *c, AdvGeo &g) {
  const Level &F = c->lv[0];
  const int nl[3] = {F.X, F.Y, F.Z};
  const int nbin[3] = {kAdvBinsPerCell * (F.X - 1), kAdvBinsPerCell * (F.Y - 1), kAdvBinsPerCell * (F.Z - 1)};
  if (!c->adv_bins) {
    std::vector<int> t;
    for (int a = 0; a < 3; a++) {
      const double *gl = F.g[a].data();
      const double h = (gl[nl[a] - 1] - gl[0]) / nbin[a];
      int i = 0;
      for (int b = 0; b < nbin[a]; b++) {
        const double s0 = gl[0] + b * h;
        while (i < nl[a] - 2 && gl[i + 1] <= s0) i++;
        t.push_back(i);
      }
    }
    int rc = dalloc(&c->adv_bins, t.size());
    if (!rc) rc = hipMemcpy(c->adv_bins, t.data(), t.size() * sizeof(int), hipMemcpyHostToDevice) ? STK_EHIP : 0;
    if (rc) {
      dfree(c->adv_bins);
      c->adv_bins = nullptr;
      return rc;
    }
  }
  g.X = F.X; g.Y = F.Y; g.Z = F.Z; g.XY = F.XY;
  g.gx = F.g64[0]; g.gy = F.g64[1]; g.gz = F.g64[2];
  g.xs = F.g[0][F.X - 1] - F.g[0][0]; g.ys = F.g[1][F.Y - 1] - F.g[1][0]; g.zs = F.g[2][F.Z - 1] - F.g[2][0];
  g.n0 = F.n;
  g.px = c->per[0]; g.py = c->per[1]; g.pz = c->per[2];
  g.wrap_out = nullptr;
  AdvBins *bn[3] = {&g.bx, &g.by, &g.bz};
  for (int a = 0, o = 0; a < 3; o += nbin[a], a++) {
    bn[a]->t = c->adv_bins + o;
    bn[a]->nb = nbin[a];
    bn[a]->g0 = F.g[a][0];
    bn[a]->inv = nbin[a] / (F.g[a][nl[a] - 1] - F.g[a][0]);
  }
  return STK_OK;
}

int stk_move_markers(stk_ctx *c, long nmark, float *mx, float *my, float *mz, float *marke, float *markk,
                     const int *markt, float *eii, double timestep, int markmod, int outgrid) {
  if (!c || nmark < 0 || (nmark && (!mx || !my || !mz)) || outgrid < 0 || outgrid > 2) return STK_EINVAL;
  if (!c->prepared || !c->X) return STK_ESTATE;
  if (nmark == 0) return STK_OK;
  Level &F = c->lv[0];
  const size_t n6 = (size_t)6 * F.n;
  // FP32 planes of vx, vy, vz and, when an invariant or strain is asked for, of the strain rates
  const bool need_e = eii || marke;
  if (need_e) {
    if (!c->sig && dalloc(&c->sig, n6)) return STK_ENOMEM;
    if (!c->epsr && dalloc(&c->epsr, n6)) return STK_ENOMEM;
    hipLaunchKernelGGL(k_post_stress, dim3(blocks(F.n, 256)), dim3(256), 0, c->st, post_geo(c), c->X, c->sig,
                       c->epsr);
  }
  float *planes = nullptr;
  if (dalloc(&planes, (size_t)(need_e ? 9 : 3) * F.n)) return STK_ENOMEM;
  hipLaunchKernelGGL(k_adv_planes, dim3(blocks(F.n, 256)), dim3(256), 0, c->st, F.n, c->X,
                     need_e ? c->epsr : nullptr, planes, planes + 3 * F.n);
  float *d[5] = {};
  int *dt = nullptr;
  float *de = nullptr;
  float *h[5] = {mx, my, mz, marke, markk};
  int rc = STK_OK;
  for (int k = 0; k < 5 && !rc; k++)
    if (h[k]) {
      rc = dalloc(&d[k], nmark);
      if (!rc) rc = hipMemcpyAsync(d[k], h[k], nmark * sizeof(float), hipMemcpyHostToDevice, c->st) ? STK_EHIP : 0;
    }
  if (!rc && markt) {
    rc = dalloc(&dt, nmark);
    if (!rc) rc = hipMemcpyAsync(dt, markt, nmark * sizeof(int), hipMemcpyHostToDevice, c->st) ? STK_EHIP : 0;
  }
  // eii of the markers the step skips keeps the caller's values
  if (!rc && eii) {
    rc = dalloc(&de, nmark);
    if (!rc) rc = hipMemcpyAsync(de, eii, nmark * sizeof(float), hipMemcpyHostToDevice, c->st) ? STK_EHIP : 0;
  }
  int *wrap_out = nullptr;
  const bool periodic = c->per[0] || c->per[1] || c->per[2];
  if (!rc && periodic) {
    rc = dalloc(&wrap_out, 1);
    if (!rc) rc = hipMemsetAsync(wrap_out, 0, sizeof(int), c->st) ? STK_EHIP : 0;
  }
  AdvGeo g;
  if (!rc) rc = adv_geo(c, g);
  if (!rc) {
    g.wrap_out = wrap_out;
    if (g.px || g.py || g.pz)
      hipLaunchKernelGGL(k_move_markers<true>, dim3(blocks(nmark, 256)), dim3(256), 0, c->st, g, planes,
                         need_e ? planes + 3 * F.n : nullptr, nmark, d[0], d[1], d[2], d[3], d[4], dt, de, timestep,
                         markmod, outgrid);
    else
      hipLaunchKernelGGL(k_move_markers<false>, dim3(blocks(nmark, 256)), dim3(256), 0, c->st, g, planes,
                         need_e ? planes + 3 * F.n : nullptr, nmark, d[0], d[1], d[2], d[3], d[4], dt, de, timestep,
                         markmod, outgrid);
    for (int k = 0; k < 5 && !rc; k++)
      if (h[k]) rc = hipMemcpyAsync(h[k], d[k], nmark * sizeof(float), hipMemcpyDeviceToHost, c->st) ? STK_EHIP : 0;
    if (!rc && eii) rc = hipMemcpyAsync(eii, de, nmark * sizeof(float), hipMemcpyDeviceToHost, c->st) ? STK_EHIP : 0;
    int out = 0;
    if (!rc && wrap_out) rc = hipMemcpyAsync(&out, wrap_out, sizeof(int), hipMemcpyDeviceToHost, c->st) ? STK_EHIP : 0;
    if (!rc) rc = hipStreamSynchronize(c->st) ? STK_EHIP : 0;
    // a point still outside after one period's wrap: the reference stops there (mark3mg.c:340, 471)
    if (!rc && out) rc = STK_EINVAL;
  }
  (void)hipStreamSynchronize(c->st);
  for (int k = 0; k < 5; k++) dfree(d[k]);
  dfree(dt);
  dfree(de);
  dfree(planes);
  dfree(wrap_out);
  return rc;
}

int stk_pressure_gauge(stk_ctx *c, int px, int py, int pz, double pinit) {
  if (!c || !c->X) return STK_EINVAL;
  Level &F = c->lv[0];
  if (px < 1 || py < 1 || pz < 1 || px >= F.X || py >= F.Y || pz >= F.Z) return STK_EINVAL;
  long n = (long)pz * F.XY + (long)px * F.Y + py;
  hipLaunchKernelGGL((k_convert<double, double>), dim3(1), dim3(1), 0, c->st, 1L, c->X + 4 * n, c->scal + 16);
  hipLaunchKernelGGL(k_pshift, dim3(blocks(F.n, 256)), dim3(256), 0, c->st, F.n, F.flags, c->X, c->scal + 16, pinit);
  HIPCHK(hipStreamSynchronize(c->st));
  return STK_OK;
}

int stk_hydrostatic_guess(stk_ctx *c) {
  if (!c || !c->prepared) return STK_ESTATE;
  Level &F = c->lv[0];
  long n4 = 4 * F.n;
  HIPCHK(hipMemsetAsync(c->X, 0, n4 * 8, c->st));
  Geo<double> g = F.geo<double>(c->kn.p2koef);
  hipLaunchKernelGGL(k_hydro, dim3(blocks((long)F.X * F.Y, 64)), dim3(64), 0, c->st, g, F.flags, c->B, c->X);
  bc_passes<double>(c, F, c->B, c->X, 0);
  HIPCHK(hipStreamSynchronize(c->st));
  return STK_OK;
}

// average device time of one level-0 kernel launch, measured with HIP events
// on the library stream: which = 0 FP32 smoother sweep (velocity block),
// 1 FP64 smoother sweep, 2 FP64 operator apply (A x, all rows),
// 3 FP32 residual, 4 FP64 residual (b - A x)
int stk_time_kernel(stk_ctx *c, int which, int level, int reps, double *ms_avg) {
  if (!c || !c->prepared || level < 0 || level > c->L || reps < 1 || !ms_avg || which < 0 || which > 5)
    return STK_EINVAL;
  int rc = ensure_vectors(c, true, true);
  if (rc) return rc;
  Level &L = c->lv[level];
  const int *tt = kTileDefault4;  // FP64 sweeps: the default tile
  int tile[3] = {tt[0], tt[1], tt[2]};
  int ty = tile[0], tx = tile[1], tz = tile[2];
  const TileGeo tg = make_tile(ty, tx, tz, L.box);
  // FP32 sweep: the instance the V-cycle runs on this level (the level's
  // tuned tile, boundary rows set inside the tiles where the level allows),
  // its chains composed once before the timed launches
  const int *t32 = tuned_tile<float>(L);
  const TileGeo tg32 = make_tile(t32[0], t32[1], t32[2], L.box);
  const bool sv32 = t32[3] && smooth_lds<float>(t32, true) <= kLevelLds;
  const bool in32 = which == 0 && tile_bc(c, L, t32);
  if (in32)
    hipLaunchKernelGGL(k_bc_compose<float>, dim3(blocks(L.nbc, 256)), dim3(256), 0, c->st, L.chain(), L.bt_idx,
                       L.b32, chain_c<float>(L), chain_k<float>(L), 1);
  auto launch = [&]() {
    if (which == 0) {
      Geo<float> g = L.geo<float>(c->kn.p2koef);
      launch_smooth<float>(c, t32, sv32, tg32, g, L.flags, L.x32[0], L.x32[1], L.b32, 0,
                           in32 ? tile_chains<float>(L) : nullptr);
    } else if (which == 1) {
      Geo<double> g = L.geo<double>(c->kn.p2koef);
      const bool sv = tt[3] && smooth_lds<double>(tile, true) <= kLevelLds;
      launch_smooth<double>(c, tile, sv, tg, g, L.flags, L.x64[0], L.x64[1], L.b64, 0);
    } else if (which == 2) {
      rowop_rows<double, false>(c, L, L.x64[0], nullptr, L.r64, 1);
    } else if (which == 3) {
      rowop_rows<float, true>(c, L, L.x32[0], L.b32, L.r32, 0);
    } else if (which == 4) {
      rowop_rows<double, true>(c, L, L.x64[0], L.b64, L.r64, 1);
    } else {
      hipLaunchKernelGGL(k_stream_copy, dim3(blocks(4 * L.n, 256)), dim3(256), 0, c->st, 4 * L.n, L.x64[0],
                         L.x64[1]);
    }
  };
  launch();
  HIPCHK(hipEventRecord(c->ev0, c->st));
  for (int i = 0; i < reps; i++) launch();
  HIPCHK(hipEventRecord(c->ev1, c->st));
  HIPCHK(hipEventSynchronize(c->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  *ms_avg = ms / reps;
  return STK_OK;
}

// ---- preconditioner as a HIP graph ---------------------------------------
// The preconditioner is a fixed sequence of launches on the library stream
// with no host synchronisation, memsets or allocations, so it is captured
// once (input gin, output gout) and replayed every Krylov step.  The capture
// stays valid while the options that shape the sequence and the buffers it
// names are unchanged (a prepare that keeps the sizes refills in place).
static bool same_shape(const stk_solver_opts &a, const stk_solver_opts &b) {
  return a.precision == b.precision && !memcmp(a.pre, b.pre, sizeof a.pre) && !memcmp(a.post, b.post, sizeof a.post) &&
         !memcmp(a.tile, b.tile, sizeof a.tile) && a.omega == b.omega && a.cycle == b.cycle && a.schur == b.schur &&
         a.ksweeps == b.ksweeps && a.kscale == b.kscale;
}

static int ensure_graph(stk_ctx *c, const stk_solver_opts *o, const int *tile) {
  if (c->graph_ok && same_shape(c->gkey, *o)) return STK_OK;
  const long n4 = 4 * c->lv[0].n;
  if (!c->gin && (dalloc(&c->gin, n4) || dalloc(&c->gout, n4))) return STK_ENOMEM;
  c->graph_ok = false;
  {
    RtLock lk(g_rt_mu);
    if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
    if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
  }
  // the tile sweeps' chain lists exist before capture, so the graph uses them
  for (int l = 0; l < c->L; l++) {
    Level &L = c->lv[l];
    (void)tile_bc(c, L, tile ? tile : (o->precision == 1 ? tuned_tile<double>(L) : tuned_tile<float>(L)));
  }
  HIPCHK(hipStreamSynchronize(c->st));
  // one capture at a time in the process, and no allocation or release on
  // another thread while it runs; the preconditioner only launches kernels
  RtLock lk(g_rt_mu);
  hipError_t e = hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    fprintf(stderr, "pardiso2_amd: hipStreamBeginCapture: %s\n", hipGetErrorString(e));
    return STK_EHIP;
  }
  if (o->precision == 1) precond<double>(c, o, c->gin, c->gout, tile);
  else precond<float>(c, o, c->gin, c->gout, tile);
  hipGraph_t gr = nullptr;
  e = hipStreamEndCapture(c->st, &gr);  // always ends the capture, so the stream leaves capture mode
  if (e == hipSuccess) e = hipGetLastError();
  if (e == hipSuccess) e = hipGraphInstantiate(&c->gexec, gr, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    fprintf(stderr, "pardiso2_amd: preconditioner capture: %s\n", hipGetErrorString(e));
    if (gr) (void)hipGraphDestroy(gr);
    c->gexec = nullptr;
    return STK_EHIP;
  }
  c->graph = gr;
  c->gkey = *o;
  c->graph_ok = true;
  return STK_OK;
}

// z = M^-1 r, replayed from the captured graph or launched eagerly
// z32: in mixed precision with FP32 storage, the FP32 copy of z to write
// (z itself then holds the rounded values)
static int apply_precond(stk_ctx *c, const stk_solver_opts *o, const int *tile, const double *r, double *z,
                         float *z32 = nullptr) {
  const long n4 = 4 * c->lv[0].n;
  if (o->graph) {
    HIPCHK(hipMemcpyAsync(c->gin, r, n4 * 8, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipGraphLaunch(c->gexec, c->st));
    HIPCHK(hipMemcpyAsync(z, c->gout, n4 * 8, hipMemcpyDeviceToDevice, c->st));
  } else if (o->precision == 1) {
    precond<double>(c, o, r, z, tile);
  } else {
    precond<float>(c, o, r, z, tile, z32);
  }
  if (z32 && o->graph)  // the captured preconditioner ends with k_combine: round here
    hipLaunchKernelGGL(k_round32, dim3(blocks(n4, 256)), dim3(256), 0, c->st, n4, z, z32);
  return STK_OK;
}

// Flexible GMRES(m), right-preconditioned, in the W-weighted inner product,
// with delayed classical Gram-Schmidt reorthogonalisation: step j
// preconditions the once-orthogonalised candidate u (Z_j = M^-1 u), forms
// w = A Z_j, then one pass computes the dots of u and w with the basis and a
// second pass reorthogonalises u into v_j and orthogonalises w against
// v_0..v_j.  The Hessenberg column of step j-1 is corrected by u's
// reorthogonalisation before its Givens rotation, so the residual estimate
// lags one step (the last step's preconditioning is discarded on exit).
// Returns the number of Arnoldi steps used; *rel is the true relative
// residual after the last cycle.
static int fgmres(stk_ctx *c, const stk_solver_opts *o, const int *tile, double nb, double *rel, stk_stats *S,
                  int &status) {
  const long n4 = 4 * c->lv[0].n;
  const int m = o->restart;
  // FP32 storage of the basis and of the preconditioned vectors (mixed
  // precision): the Krylov vectors' buffers hold FP32 values, u and w stay
  // FP64 working vectors
  const bool b32 = o->precision == 0 && o->basis == 1;
  auto v32 = [&](int i) { return reinterpret_cast<float *>(c->Qs[i]); };
  auto z32 = [&](int i) { return reinterpret_cast<float *>(c->Zs[i]); };
  // dots and updates: one block per 1024-element slice up to 4 nred blocks
  // (partials of 2j+2 sums per block for the dots)
  const long slice = b32 ? 256L * kGsE : 1024L;  // quads (FP32 basis) or double2 pairs per block slice
  const long units = b32 ? n4 / 4 : n4 / 2;
  const unsigned nbk = (unsigned)std::min<long>(4L * c->nred, std::max<long>(1, (units + slice - 1) / slice));
  const unsigned nbu = nbk;
  std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m), sn(m), g(m + 1), hh(2 * m + 2), y(m), coef(2 * m + 2);
  auto h = [&](int i, int j) -> double & { return H[(size_t)i * m + j]; };
  auto givens = [&](int col) {
    for (int i = 0; i < col; i++) {
      const double a = h(i, col), b = h(i + 1, col);
      h(i, col) = cs[i] * a + sn[i] * b;
      h(i + 1, col) = -sn[i] * a + cs[i] * b;
    }
    const double a = h(col, col), b = h(col + 1, col);
    const double d = std::hypot(a, b);
    cs[col] = d > 0 ? a / d : 1.0;
    sn[col] = d > 0 ? b / d : 0.0;
    h(col, col) = d;
    h(col + 1, col) = 0.0;
    g[col + 1] = -sn[col] * g[col];
    g[col] = cs[col] * g[col];
  };
  int it = 0;
  double r = *rel;
  while (r > o->tol && it < o->maxit) {
    const double beta = r * nb;
    hipLaunchKernelGGL(k_scale_to, dim3(blocks(n4, 256)), dim3(256), 0, c->st, n4, 1.0 / beta, c->R,
                       b32 ? c->wu : c->Qs[0]);
    std::fill(g.begin(), g.end(), 0.0);
    std::fill(H.begin(), H.end(), 0.0);
    g[0] = beta;
    int j = 0, ncols = 0;
    bool conv = false;
    while (j < m && it < o->maxit) {
      double *u = b32 ? c->wu : c->Qs[j], *Z = b32 ? c->wz : c->Zs[j], *W = b32 ? c->ww : c->Qs[j + 1];
      // FP32 storage: A is applied to the stored (rounded) z, so the update x += Z y uses the same vectors
      int prc = apply_precond(c, o, tile, u, Z, b32 ? z32(j) : nullptr);
      if (prc) { status = prc; break; }
      rowop<double, false>(c, 0, Z, nullptr, W, 1);
      PtrsK P{};
      PtrsKT<float> P32{};
      for (int i = 0; i < j; i++) { P.p[i] = c->Qs[i]; P32.p[i] = v32(i); }
      if (b32)
        hipLaunchKernelGGL(k_dots2f, dim3(nbk), dim3(256), 0, c->st, n4, c->W2, u, W, P32, j, c->kred);
      else
        hipLaunchKernelGGL(k_dots2<double>, dim3(nbk), dim3(256), 0, c->st, n4, c->W2, u, W, P, j, c->kred);
      hipLaunchKernelGGL(k_reduce_rows, dim3(2 * j + 2), dim3(256), 0, c->st, c->kred, (int)nbk, 2 * j + 2, c->hcol);
      HIPCHK(hipMemcpyAsync(c->hpin, c->hcol, (2 * j + 2) * sizeof(double), hipMemcpyDeviceToHost, c->st));
      HIPCHK(hipStreamSynchronize(c->st));
      memcpy(hh.data(), c->hpin, (2 * j + 2) * sizeof(double));
      const double *a = hh.data(), *b = hh.data() + j;
      const double cuw = hh[2 * j], cuu = hh[2 * j + 1];
      double asq = 0.0, ab = 0.0;
      for (int i = 0; i < j; i++) { asq += a[i] * a[i]; ab += a[i] * b[i]; }
      const double nu = std::sqrt(cuu - asq > 0.5 * cuu ? cuu - asq : cuu);
      if (!std::isfinite(nu) || nu <= 0) { status = STK_ENAN; break; }
      if (j >= 1) {
        // u = V a + nu v_j: fold the reorthogonalisation into column j-1
        const double hj = h(j, j - 1);
        for (int i = 0; i < j; i++) h(i, j - 1) += hj * a[i];
        h(j, j - 1) = hj * nu;
        givens(j - 1);
        const double est = std::fabs(g[j]) / nb;
        if (it < 256) S->hist[it] = est;
        if (o->verbose) fprintf(stderr, "pardiso2_amd: it %d relres~ %.3e\n", it, est);
        ncols = j;
        if (est <= o->tol) { conv = true; break; }
      }
      const double hjj = (cuw - ab) / nu;
      for (int i = 0; i < j; i++) { coef[i] = a[i]; coef[j + i] = b[i]; }
      coef[2 * j] = hjj;
      coef[2 * j + 1] = 1.0 / nu;
      memcpy(c->hpin, coef.data(), (2 * j + 2) * sizeof(double));  // copied before the stream reuses hpin
      HIPCHK(hipMemcpyAsync(c->hcol, c->hpin, (2 * j + 2) * sizeof(double), hipMemcpyHostToDevice, c->st));
      if (b32)
        // the FP64 v_j is not read again (the next candidate overwrites u), only its stored FP32 copy
        hipLaunchKernelGGL(k_update2f, dim3(nbu), dim3(256), 0, c->st, n4, j, c->hcol, P32, u, W, c->W2, c->kred,
                           v32(j), 0);
      else
        hipLaunchKernelGGL(k_update2<double>, dim3(nbu), dim3(256), 0, c->st, n4, j, c->hcol, P, u, W, c->W2,
                           c->kred, (double *)nullptr);
      hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(256), 0, c->st, c->kred, (int)nbu, 1, c->hcol + 2 * m + 1);
      HIPCHK(hipMemcpyAsync(c->hpin + 2 * kMaxDirs + 4, c->hcol + 2 * m + 1, sizeof(double), hipMemcpyDeviceToHost,
                            c->st));
      HIPCHK(hipStreamSynchronize(c->st));
      const double hn2 = c->hpin[2 * kMaxDirs + 4];
      const double hn = std::sqrt(std::max(hn2, 0.0));
      if (!std::isfinite(hn)) { status = STK_ENAN; break; }
      for (int i = 0; i < j; i++) h(i, j) = b[i];
      h(j, j) = hjj;
      h(j + 1, j) = hn;
      it++;
      j++;
      if (hn == 0.0) break;
      // the next candidate: w / |w| (in place, or into the working u)
      hipLaunchKernelGGL(k_scale_to, dim3(blocks(n4, 256)), dim3(256), 0, c->st, n4, 1.0 / hn, W, b32 ? u : W);
    }
    if (status) break;
    if (!conv && j > ncols) {
      givens(j - 1);  // the last column keeps its single orthogonalisation
      ncols = j;
      if (it < 256) S->hist[it] = std::fabs(g[j]) / nb;
    }
    // x += Z y with H y = g over the ncols finished columns
    const int k = ncols;
    for (int i = k - 1; i >= 0; i--) {
      double acc = g[i];
      for (int q = i + 1; q < k; q++) acc -= h(i, q) * y[q];
      y[i] = acc / h(i, i);
    }
    for (int i = 0; i < k; i++) y[i] = -y[i];
    HIPCHK(hipStreamSynchronize(c->st));  // the staging buffer may still feed the last coefficient copy
    memcpy(c->hpin, y.data(), k * sizeof(double));
    HIPCHK(hipMemcpyAsync(c->hcol, c->hpin, k * sizeof(double), hipMemcpyHostToDevice, c->st));
    if (b32) {
      PtrsKT<float> P{};
      for (int i = 0; i < k; i++) P.p[i] = z32(i);
      hipLaunchKernelGGL(k_maxpy_norm<float>, dim3(nbu), dim3(256), 0, c->st, n4, k, c->hcol, P, c->X, c->X, c->W2,
                         c->kred);
    } else {
      PtrsK P{};
      for (int i = 0; i < k; i++) P.p[i] = c->Zs[i];
      hipLaunchKernelGGL(k_maxpy_norm<double>, dim3(nbu), dim3(256), 0, c->st, n4, k, c->hcol, P, c->X, c->X, c->W2,
                         c->kred);
    }
    true_residual(c);
    double s4[4];
    int rc = blocks_norms(c, c->R, c->B, s4);
    if (rc) { status = rc; break; }
    r = std::sqrt(s4[0] + s4[1] + s4[2]) / nb;
    if (!std::isfinite(r)) { status = STK_ENAN; break; }
  }
  *rel = r;
  return it;
}

// allocations a solve needs, and the preconditioner capture when replay is on
// ---- post-solve pass: stresses and the reference's norms -------------------
static PostGeo post_geo(const stk_ctx *c) {
  const Level &F = c->lv[0];
  PostGeo g;
  g.X = F.X; g.Y = F.Y; g.Z = F.Z; g.XY = F.XY;
  g.iAx = F.iA64[0]; g.iBx = F.iB64[0]; g.iAy = F.iA64[1]; g.iBy = F.iB64[1]; g.iAz = F.iA64[2]; g.iBz = F.iB64[2];
  g.ec = F.e64[0]; g.exy = F.e64[1]; g.exz = F.e64[2]; g.eyz = F.e64[3];
  g.nu = c->nu;
  g.nubeg = c->kn.nubeg; g.nuend = c->kn.nuend; g.vm = c->kn.viscmod;
  g.per = F.per;
  return g;
}

// norms of the finest-level solution x (move3mg.c:3061-3086, 3090-3238, stop test 3282)
static int post_norms(stk_ctx *c, const double *x, stk_norms *out) {
  Level &F = c->lv[0];
  if (!c->post_part && dalloc(&c->post_part, (size_t)(kPostBlocks + 1) * kPostN)) return STK_ENOMEM;
  const int nb = (int)std::min<long>(kPostBlocks, (F.n + 255) / 256);
  hipLaunchKernelGGL(k_post_norms, dim3(nb), dim3(256), 0, c->st, post_geo(c), F.flags, x, c->B, F.nbc, F.bc_r,
                     F.bc_j, F.bc_negk, F.bc_c, c->post_part);
  double *dout = c->post_part + (size_t)kPostBlocks * kPostN;
  hipLaunchKernelGGL(k_post_reduce, dim3(kPostN), dim3(256), 0, c->st, c->post_part, nb, dout);
  double s[kPostN];
  HIPCHK(hipMemcpyAsync(s, dout, sizeof(s), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  const double dvx = s[11] - s[10], dvy = s[13] - s[12], dvz = s[15] - s[14];
  double mv = std::sqrt((dvx * dvx + dvy * dvy + dvz * dvz) / 3.0);
  if (mv <= 0) mv = 1e-8;
  if (c->kn.divvmax > 0) mv = c->kn.divvmax;
  auto rms = [](double sum, double n) { return n > 0 ? std::sqrt(sum / n) : 0.0; };
  stk_norms &N = *out;
  N.stoksum = rms(s[0], s[3]);
  N.stoksum1 = rms(s[1], s[3]) / mv;
  N.stoksum2 = rms(s[2], s[3]) / mv;
  N.contsum = rms(s[4], s[5]) / mv;
  N.contsumb = rms(s[6], s[7]) / mv;
  N.bondsum = rms(s[8], s[9]);
  N.stoknum = s[3]; N.contnum = s[5]; N.contnumb = s[7]; N.bondnum = s[9];
  N.maxvxyz = mv;
  N.minvx = s[10]; N.maxvx = s[11]; N.minvy = s[12]; N.maxvy = s[13]; N.minvz = s[14]; N.maxvz = s[15];
  N.minpr = s[16]; N.maxpr = s[17]; N.minnu = s[18]; N.maxnu = s[19];
  const double lim = c->kn.stoksmin * c->kn.stoksmax;
  N.converged = !(N.stoksum2 > lim || N.stoksum * N.stoksum1 > lim || N.contsum > c->kn.divvmin);
  return STK_OK;
}

static int solve_setup(stk_ctx *c, const stk_solver_opts *o, HistLease &lease) {
  const bool fp64 = o->precision == 1;
  int rc = ensure_vectors(c, !fp64, fp64);
  if (rc) return rc;
  if (o->history == 1) {
    rc = ensure_outer(c, 0);
    if (rc) return rc;
    HistSet *h = hist_acquire(4 * c->lv[0].n, o->restart);
    if (!h) return STK_ENOMEM;
    lease.c = c;
    lease.h = h;
    lease.keepZ = std::move(c->Zs);
    lease.keepQ = std::move(c->Qs);
    c->Zs.assign(h->Z.begin(), h->Z.begin() + o->restart);
    c->Qs.assign(h->Q.begin(), h->Q.begin() + o->restart + 1);
  } else {
    rc = ensure_outer(c, o->restart + 1);
    if (rc) return rc;
  }
  if (o->graph) {
    const bool explicit_tile = o->tile[0] || o->tile[1] || o->tile[2];
    rc = ensure_graph(c, o, explicit_tile ? o->tile : nullptr);
  }
  return rc;
}

int stk_solve(stk_ctx *c, const stk_solver_opts *o, stk_stats *st) {
  if (!c || !o) return STK_EINVAL;
  if (!c->prepared) return STK_ESTATE;
  if (o->restart < 1 || o->restart > kMaxDirs - 1 || o->maxit < 1 || !(o->tol > 0) || o->cycle < 0 || o->cycle > 1 ||
      o->coarse_viscmod < -1 || o->coarse_viscmod > 2 || o->schur < 0 || o->schur > 2 || o->ksweeps < 0 ||
      o->ksweeps > 16 || !(o->kscale > 0) || o->krylov < 0 || o->krylov > 1 || o->graph < 0 || o->graph > 1 ||
      o->history < 0 || o->history > 1 || o->basis < 0 || o->basis > 1)
    return STK_EINVAL;
  {
    const int want = (o->coarse_viscmod < 0) ? c->kn.viscmod : o->coarse_viscmod;
    if (want != c->cvm) {
      int rc0 = stk_prepare_cvm(c, want);
      if (rc0) return rc0;
    }
  }
  const bool explicit_tile = o->tile[0] || o->tile[1] || o->tile[2];
  const int *tile = explicit_tile ? o->tile : nullptr;
  if (explicit_tile && (o->tile[0] < 2 || o->tile[1] < 1 || o->tile[2] < 1 ||
                        (size_t)(o->tile[0] + 2) * (o->tile[1] + 2) * (o->tile[2] + 2) * 32 > 65536))
    return STK_EINVAL;
  HistLease lease;  // borrowed Krylov vectors (history = 1) go back when the solve returns
  int rc = solve_setup(c, o, lease);
  if (rc) return rc;
  stk_stats local;
  stk_stats *S = st ? st : &local;
  memset(S, 0, sizeof(*S));
  long n4 = 4 * c->lv[0].n;
  auto t0 = std::chrono::steady_clock::now();
  HIPCHK(hipEventRecord(c->ev0, c->st));
  true_residual(c);
  double s4[4];
  rc = blocks_norms(c, c->R, c->B, s4);
  if (rc) return rc;
  const double nb = std::sqrt(s4[3]) > 0 ? std::sqrt(s4[3]) : 1.0;
  double rel = std::sqrt(s4[0] + s4[1] + s4[2]) / nb;
  S->relres0 = rel;
  S->hist[0] = rel;
  int it = 0, k = 0;
  double hr[2];
  if (o->krylov == 1) {
    int status = 0;
    it = fgmres(c, o, tile, nb, &rel, S, status);
    S->status = status;
  }
  while (o->krylov == 0 && rel > o->tol && it < o->maxit) {
    double *Z = c->Zs[k], *Q = c->Qs[k];
    rc = apply_precond(c, o, tile, c->R, Z);
    if (rc) return rc;
    rowop<double, false>(c, 0, Z, nullptr, Q, 1);
    if (k > 0) {
      std::vector<const double *> bs(c->Qs.begin(), c->Qs.begin() + k);
      multidot(c, Q, bs, c->scal + 32);
      multiaxpy(c, k, c->scal + 32, Q, Z);
    }
    multidot(c, Q, {Q, c->R}, c->scal);
    hipLaunchKernelGGL(k_gcr_scalars, dim3(1), dim3(1), 0, c->st, c->scal, c->scal + 2);
    hipLaunchKernelGGL(k_scale2, dim3(blocks(n4, 256)), dim3(256), 0, c->st, n4, c->scal + 2, Q, Z);
    hipLaunchKernelGGL(k_update, dim3(blocks(n4, 256)), dim3(256), 0, c->st, n4, c->scal + 3, Z, Q, c->X, c->R);
    multidot(c, c->R, {c->R}, c->scal + 4);
    HIPCHK(hipMemcpyAsync(c->hpin, c->scal + 4, 8, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    hr[0] = c->hpin[0];
    double nrel = std::sqrt(std::max(hr[0], 0.0)) / nb;
    it++;
    k++;
    if (!std::isfinite(nrel)) { S->status = STK_ENAN; break; }
    if (k == o->restart) {
      k = 0;
      true_residual(c);
      multidot(c, c->R, {c->R}, c->scal + 4);
      HIPCHK(hipMemcpyAsync(c->hpin, c->scal + 4, 8, hipMemcpyDeviceToHost, c->st));
      HIPCHK(hipStreamSynchronize(c->st));
      hr[0] = c->hpin[0];
      nrel = std::sqrt(hr[0]) / nb;
    }
    rel = nrel;
    if (it < 256) S->hist[it] = rel;
    if (o->verbose) fprintf(stderr, "pardiso2_amd: it %d relres %.3e\n", it, rel);
  }
  // final residual recomputed from b - A x
  true_residual(c);
  rc = blocks_norms(c, c->R, c->B, s4);
  if (rc) return rc;
  HIPCHK(hipEventRecord(c->ev1, c->st));
  HIPCHK(hipEventSynchronize(c->ev1));
  float ms = 0;
  hipEventElapsedTime(&ms, c->ev0, c->ev1);
  (void)t0;
  S->seconds = ms * 1e-3;
  S->iterations = it;
  S->relres = std::sqrt(s4[0] + s4[1] + s4[2]) / nb;
  S->blocks[0] = std::sqrt(s4[0]); S->blocks[1] = std::sqrt(s4[1]); S->blocks[2] = std::sqrt(s4[2]);
  if (!S->status) S->status = (S->relres <= o->tol * 1.0000001) ? STK_OK : STK_ENOCONV;
  rc = post_norms(c, c->X, &S->norms);
  if (rc) return rc;
  return S->status;
}

// ---- reference arrays and kernels on the context -------------------------
static int ensure_ref(stk_ctx *c, int l) {
  Level &L = c->lv[l];
  for (int k = 0; k < 3; k++)
    if (!L.ref[k]) {
      if (dalloc(&L.ref[k], 4 * L.n)) return STK_ENOMEM;
      hipMemsetAsync(L.ref[k], 0, 4 * L.n * 8, c->st);
    }
  return STK_OK;
}

int stk_vec_set(stk_ctx *c, int which, int l, const double *h) {
  if (!c || which < 0 || which > 2 || l < 0 || l > c->L || !h) return STK_EINVAL;
  int rc = ensure_ref(c, l);
  if (rc) return rc;
  H2D(c, c->lv[l].ref[which], h, 4 * c->lv[l].n * 8);
  return STK_OK;
}

int stk_vec_get(stk_ctx *c, int which, int l, double *h) {
  if (!c || which < 0 || which > 2 || l < 0 || l > c->L || !h) return STK_EINVAL;
  int rc = ensure_ref(c, l);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(c->st));
  D2H(c, h, c->lv[l].ref[which], 4 * c->lv[l].n * 8);
  return STK_OK;
}

int stk_apply(stk_ctx *c, int l, const double *x, double *y) {
  if (!c || !c->prepared || l < 0 || l > c->L || !x || !y) return STK_EINVAL;
  Level &L = c->lv[l];
  double *dx, *dy;
  if (dalloc(&dx, 4 * L.n) || dalloc(&dy, 4 * L.n)) return STK_ENOMEM;
  H2D(c, dx, x, 4 * L.n * 8);
  rowop<double, false>(c, l, dx, nullptr, dy, 1);
  HIPCHK(hipStreamSynchronize(c->st));
  D2H(c, y, dy, 4 * L.n * 8);
  dfree(dx); dfree(dy);
  return STK_OK;
}

// FP64 coupled sweeps on the reference arrays of level l, then fre0
int stk_smooth(stk_ctx *c, int l, int sweeps, int ty, int tx, int tz) {
  return stk_smooth_ex(c, l, sweeps, ty, tx, tz, 1, 0);
}

int stk_smooth_ex(stk_ctx *c, int l, int sweeps, int ty, int tx, int tz, int precision, int velocity_only) {
  if (!c || !c->prepared || l < 0 || l > c->L || sweeps < 0 || precision < 0 || precision > 1) return STK_EINVAL;
  const int withp = velocity_only ? 0 : 1;
  const bool whole = (ty == 0 && tx == 0 && tz == 0);
  const bool staged = tz < 0;
  if (staged) tz = -tz;
  const size_t tsize = precision ? sizeof(double) : sizeof(float);
  if (whole && (4 * (size_t)c->lv[l].n * tsize > kLevelLds || c->lv[l].per)) return STK_EINVAL;
  if (!whole && (ty < 1 || tx < 1 || tz < 1 || (size_t)(ty + 2) * (tx + 2) * (tz + 2) * 32 > 65536)) return STK_EINVAL;
  int rc = ensure_ref(c, l) || ensure_vectors(c, !precision, true);
  if (rc) return STK_ENOMEM;
  Level &L = c->lv[l];
  const long n4 = 4 * L.n;
  const size_t bytes = n4 * 8;
  int tile[4] = {ty, tx, tz, 1};
  if (precision) {
    HIPCHK(hipMemcpyAsync(L.x64[L.cur64], L.ref[0], bytes, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemcpyAsync(L.b64, L.ref[1], bytes, hipMemcpyDeviceToDevice, c->st));
    if (staged) {
      if (smooth_lds<double>(tile, true) > kLevelLds) return STK_EINVAL;
      const int keep[4] = {L.tune64[0], L.tune64[1], L.tune64[2], L.tune64[3]};
      for (int a = 0; a < 4; a++) L.tune64[a] = tile[a];
      smooth<double>(c, l, sweeps, nullptr, withp, false);
      for (int a = 0; a < 4; a++) L.tune64[a] = keep[a];
    } else {
      smooth<double>(c, l, sweeps, tile, withp, whole);
    }
  } else {
    // FP32 arithmetic (the mixed-precision V-cycle's smoother) on rounded copies
    hipLaunchKernelGGL((k_cast<double, float>), dim3(blocks(n4, 256)), dim3(256), 0, c->st, n4, L.ref[0],
                       L.x32[L.cur32]);
    hipLaunchKernelGGL((k_cast<double, float>), dim3(blocks(n4, 256)), dim3(256), 0, c->st, n4, L.ref[1], L.b32);
    if (staged) {
      if (smooth_lds<float>(tile, true) > kLevelLds) return STK_EINVAL;
      const int keep[4] = {L.tune32[0], L.tune32[1], L.tune32[2], L.tune32[3]};
      for (int a = 0; a < 4; a++) L.tune32[a] = tile[a];
      smooth<float>(c, l, sweeps, nullptr, withp, false);
      for (int a = 0; a < 4; a++) L.tune32[a] = keep[a];
    } else {
      smooth<float>(c, l, sweeps, tile, withp, whole);
    }
    hipLaunchKernelGGL((k_cast<double, double>), dim3(blocks(n4, 256)), dim3(256), 0, c->st, n4, L.ref[1], L.b64);
    hipLaunchKernelGGL((k_cast<float, double>), dim3(blocks(n4, 256)), dim3(256), 0, c->st, n4, L.x32[L.cur32],
                       L.x64[L.cur64]);
  }
  rowop<double, true>(c, l, L.x64[L.cur64], L.b64, L.ref[2], 1);
  HIPCHK(hipMemcpyAsync(L.ref[0], L.x64[L.cur64], bytes, hipMemcpyDeviceToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return hipGetLastError() == hipSuccess ? STK_OK : STK_EHIP;
}

int stk_coarse_solve(stk_ctx *c) {
  if (!c || !c->prepared) return STK_ESTATE;
  int rc = ensure_ref(c, c->L);
  if (rc) return rc;
  if (!c->ainvP64) {
    rc = coarse_inverse(c, 1, &c->ainvP64);
    if (rc) return rc;
  }
  Level &C = c->lv[c->L];
  int n4 = (int)(4 * C.n);
  hipLaunchKernelGGL((k_gemv<double, double>), dim3(blocks(n4, 4)), dim3(256), 0, c->st, c->ainvP64, C.ref[1],
                     C.ref[0], n4);
  HIPCHK(hipStreamSynchronize(c->st));
  return STK_OK;
}

// ---- smoother tile autotuning -----------------------------------------------
int stk_autotune(stk_ctx *c, int precision, int reps) {
  if (!c || !c->prepared || reps < 1 || precision < 0 || precision > 1) return STK_EINVAL;
  c->graph_ok = false;  // captured launches use the tiles of the time of capture
  int rc = ensure_vectors(c, precision == 0, precision == 1);
  if (rc) return rc;
  return precision ? autotune_levels<double>(c, reps) : autotune_levels<float>(c, reps);
}

int stk_level_tile(const stk_ctx *c, int level, int precision, int out4[4]) {
  if (!c || level < 0 || level > c->L || !out4) return STK_EINVAL;
  const Level &L = c->lv[level];
  const int *t = precision ? tuned_tile<double>(L) : tuned_tile<float>(L);
  for (int a = 0; a < 4; a++) out4[a] = t[a];
  return STK_OK;
}

int stk_set_level_tile(stk_ctx *c, int level, int precision, const int tile4[4]) {
  if (!c || level < 0 || level > c->L || !tile4 || precision < 0 || precision > 1) return STK_EINVAL;
  const int ty = tile4[0], tx = tile4[1], tz = tile4[2], sv = tile4[3] ? 1 : 0;
  if (ty < 2 || tx < 1 || tz < 1) return STK_EINVAL;
  const int t[4] = {ty, tx, tz, sv};
  const size_t lds = precision ? smooth_lds<double>(t, sv) : smooth_lds<float>(t, sv);
  if (lds > kLevelLds) return STK_EINVAL;
  c->graph_ok = false;  // captured launches use the tiles of the time of capture
  int *d = precision ? c->lv[level].tune64 : c->lv[level].tune32;
  for (int a = 0; a < 4; a++) d[a] = t[a];
  return STK_OK;
}

int stk_tile_chains(stk_ctx *c, int level, int ty, int tx, int tz) {
  if (!c || level < 0 || level >= c->L || ty < 1 || tx < 1 || tz < 1) return STK_EINVAL;
  const int t[3] = {ty, tx, tz};
  return tile_bc(c, c->lv[level], t) ? 1 : 0;
}

// ---- many independent instances ---------------------------------------------
// Each context owns its stream and buffers, so instances solve concurrently:
// `threads` host workers take contexts in order and run stk_solve on each;
// the device overlaps the work of the different streams.
int stk_solve_batch(int n, stk_ctx *const *ctxs, const stk_solver_opts *opts, int nopts, stk_stats *stats,
                    int threads, int mode) {
  if (mode < 0 || mode > 1) return STK_EINVAL;
  if (n < 0 || (n > 0 && (!ctxs || !opts || (nopts != 1 && nopts != n)))) return STK_EINVAL;
  for (int i = 0; i < n; i++)
    if (!ctxs[i]) return STK_EINVAL;
  if (n == 0) return STK_OK;
  if (threads <= 0) threads = 16;
  threads = std::min(threads, n);
  std::vector<int> rcs(n, STK_OK);
  // graph replay: capture every context's preconditioner here, one at a time
  // on this thread (prepared and with its buffers allocated), so the workers
  // only replay
  for (int i = 0; i < n; i++) {
    const stk_solver_opts &oi = opts[nopts == 1 ? 0 : i];
    if (!oi.graph) continue;
    stk_ctx *c = ctxs[i];
    int rc = c->prepared ? STK_OK : stk_prepare(c);
    if (!rc) {
      HistLease lease;  // a borrowed set goes straight back: the capture does not use it
      rc = solve_setup(c, &oi, lease);
    }
    if (rc) return rc;
  }
  std::atomic<int> next{0};
  auto work = [&]() {
    int dev = 0;
    (void)hipGetDevice(&dev);
    for (int i; (i = next.fetch_add(1)) < n;) {
      int rc = STK_OK;
      if (mode == 1) {
        rc = stk_prepare(ctxs[i]);
        if (!rc) rc = stk_hydrostatic_guess(ctxs[i]);
      }
      const stk_solver_opts &oi = opts[nopts == 1 ? 0 : i];
      rcs[i] = rc ? rc : stk_solve(ctxs[i], &oi, stats ? &stats[i] : nullptr);
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < threads; t++) pool.emplace_back(work);
  work();
  for (auto &t : pool) t.join();
  int rc = STK_OK;
  for (int i = 0; i < n; i++)
    if (rcs[i] != STK_OK && (rc == STK_OK || rc == STK_ENOCONV)) rc = rcs[i];
  return rc;
}

int stk_mem_info(size_t *free_bytes, size_t *total_bytes) {
  if (!free_bytes || !total_bytes) return STK_EINVAL;
  HIPCHK(hipMemGetInfo(free_bytes, total_bytes));
  return STK_OK;
}

// ---- inspection of the Schur approximation (tests) ----------------------
int stk_kop_apply(stk_ctx *c, int l, const double *x, double *y) {
  if (!c || !c->prepared || l < 0 || l > c->L || !x || !y) return STK_EINVAL;
  Level &V = c->lv[l];
  double *dx, *dy, *db;
  if (dalloc(&dx, V.n) || dalloc(&dy, V.n) || dalloc(&db, V.n)) return STK_ENOMEM;
  H2D(c, dx, x, V.n * 8);
  HIPCHK(hipMemsetAsync(db, 0, V.n * 8, c->st));
  hipLaunchKernelGGL(k_kres<double>, dim3(blocks(V.n, 256)), dim3(256), 0, c->st, V.X, V.Y, V.Z, kplanes<double>(V),
                     db, dx, dy);
  HIPCHK(hipStreamSynchronize(c->st));
  D2H(c, y, dy, V.n * 8);
  for (long i = 0; i < V.n; i++) y[i] = -y[i];
  dfree(dx); dfree(dy); dfree(db);
  return STK_OK;
}

int stk_kcycle_apply(stk_ctx *c, int precision, int ksweeps, double kscale, const double *b, double *x) {
  if (!c || !c->prepared || !b || !x || ksweeps < 0 || ksweeps > 16 || !(kscale > 0)) return STK_EINVAL;
  if (precision) kcycle_host<double>(c, ksweeps, kscale, b, x);
  else kcycle_host<float>(c, ksweeps, kscale, b, x);
  HIPCHK(hipGetLastError());
  return STK_OK;
}

int stk_schur_apply(stk_ctx *c, const stk_solver_opts *o, const double *r, double *z) {
  if (!c || !c->prepared || !o || !r || !z || o->schur < 1 || o->schur > 2) return STK_EINVAL;
  const long n4 = 4 * c->lv[0].n;
  double *dr, *dz;
  if (dalloc(&dr, n4) || dalloc(&dz, n4)) return STK_ENOMEM;
  H2D(c, dr, r, n4 * 8);
  if (o->precision) schur_bfbt<double>(c, o, dr, dz, o->schur == 1 ? -1.0 : 1.0);
  else schur_bfbt<float>(c, o, dr, dz, o->schur == 1 ? -1.0 : 1.0);
  HIPCHK(hipStreamSynchronize(c->st));
  D2H(c, z, dz, n4 * 8);
  dfree(dr); dfree(dz);
  return STK_OK;
}

int stk_bind(stk_ctx *c) {
  g_bound = c;
  return STK_OK;
}

int stk_bind_arrays(const stk_ref_arrays *a) {
  if (!a) {
    g_arr = stk_ref_arrays{};
    g_arr_set = false;
    return STK_OK;
  }
  if (!a->sol0 || !a->fre0 || !a->fre1 || !a->val1 || !a->lin0 || !a->num0 || !a->pos0) return STK_EINVAL;
  g_arr = *a;
  g_arr_set = true;
  return STK_OK;
}

// ---- reference-named entry points ----------------------------------------

// parallel/gaus3mg_MG.c:2-17
void matclear(int rm, double *adr) {
  for (int m1 = 0; m1 < rm; m1++)
    for (int m2 = 0; m2 <= rm; m2++) *adr++ = 0;
}

// parallel/gaus3mg_MG.c:21-61: row-combination elimination without pivoting
// on the rm x (rm+1) augmented matrix, back substitution skipping zero pivots
void gausmat(int rm, double *adr, double *adr1) {
  const int w = rm + 1;
  for (int m1 = 0; m1 < rm - 1; m1++) {
    double v1 = adr[m1 * w + m1];
    if (v1 == 0) continue;
    for (int m2 = m1 + 1; m2 < rm; m2++) {
      double v2 = adr[m2 * w + m1];
      if (v2 == 0) continue;
      for (int m3 = m1 + 1; m3 <= rm; m3++) adr[m2 * w + m3] = adr[m2 * w + m3] / v2 - adr[m1 * w + m3] / v1;
    }
  }
  for (int m1 = rm - 1; m1 >= 0; m1--) {
    adr1[m1] = 0;
    double v1 = adr[m1 * w + m1];
    if (v1 != 0) {
      double v2 = adr[m1 * w + rm] / v1;
      adr1[m1] = v2;
      for (int m2 = 0; m2 <= m1 - 1; m2++) {
        double a = adr[m2 * w + m1];
        if (a != 0) adr[m2 * w + rm] -= a * v2;
      }
    }
  }
}

// parallel/gaus3mg_MG.c:311-365: store the row in the caller's row store
// (fre1[mcmax] = ui1[0]; entries from position pos0cur1 on)
int gausmat2(int am, long mcmax, long pos0cur1, long *un1, double *ui1) {
  (void)am;
  if (!g_arr_set) return STK_ENOCTX;
  if (mcmax < 0 || pos0cur1 < 0 || !un1 || !ui1) return STK_EINVAL;
  g_arr.fre1[mcmax] = ui1[0];
  g_arr.pos0[mcmax] = (int)pos0cur1;
  g_arr.num0[mcmax] = 0;
  for (long m2 = 1; m2 <= un1[0]; m2++) {
    g_arr.lin0[pos0cur1] = (int)un1[m2];
    g_arr.val1[pos0cur1] = ui1[m2];
    pos0cur1++;
    g_arr.num0[mcmax]++;
  }
  return 0;
}

// parallel/gaus3mg_MG.c:64-307.  am=1 sets one coarsest-level row
// (level-relative indices, duplicate columns summed, zeros dropped);
// am=0 solves the accumulated system with the device dense inverse and
// writes sol0 of the coarsest level at 4*(mgp[L]+node)+c (the reference
// writes mgp[L]+4*node+c, see parallel/gaus3mg_MG.c:234-238).
int pardisomat(int am, long mcmax, long mcmax2, long pos0cur1, int tid, long *un1, double *ui1) {
  (void)mcmax2; (void)pos0cur1; (void)tid;
  stk_ctx *c = g_bound;
  if (!c) return STK_ENOCTX;
  Level &C = c->lv[c->L];
  int n = (int)(4 * C.n);
  if (am == 0) {
    if (c->ca.size() != (size_t)n * n) return STK_ESTATE;
    // the inverse is built once and reused until pardisomat(1, ...) changes a row
    if (!c->cinv_ok) {
      double *A = nullptr;
      if (!c->cinv && dalloc(&c->cinv, (size_t)n * n)) return STK_ENOMEM;
      if (dalloc(&A, (size_t)n * n)) return STK_ENOMEM;
      H2D(c, A, c->ca.data(), (size_t)n * n * 8);
      const int hs = gj_invert(c, A, n, c->cinv);
      dfree(A);
      if (hs) return hs;
      c->cinv_ok = true;
    }
    double *db = nullptr, *dx = nullptr;
    if (dalloc(&db, n) || dalloc(&dx, n)) { dfree(db); dfree(dx); return STK_ENOMEM; }
    H2D(c, db, c->cb.data(), (size_t)n * 8);
    hipLaunchKernelGGL((k_gemv<double, double>), dim3(blocks(n, 4)), dim3(256), 0, c->st, c->cinv, db, dx, n);
    int rc = ensure_ref(c, c->L);
    if (!rc) {
      HIPCHK(hipMemcpyAsync(C.ref[0], dx, (size_t)n * 8, hipMemcpyDeviceToDevice, c->st));
      if (g_arr_set) HIPCHK(hipMemcpyAsync(g_arr.sol0 + 4 * C.off, dx, (size_t)n * 8, hipMemcpyDeviceToHost, c->st));
      HIPCHK(hipStreamSynchronize(c->st));
    }
    dfree(db); dfree(dx);
    return rc;
  }
  if (mcmax < 0 || mcmax >= n) return STK_EINVAL;
  if (c->ca.size() != (size_t)n * n) {
    c->ca.assign((size_t)n * n, 0.0);
    c->cb.assign(n, 0.0);
  }
  // the call sets row mcmax (duplicates within the list add up), as the
  // reference rebuilds the row from its list (parallel/gaus3mg_MG.c:256-306);
  // the kept inverse stays valid while the matrix part of every row is unchanged
  for (long m2 = 1; m2 <= un1[0]; m2++)
    if (un1[m2] < 0 || un1[m2] >= n) return STK_EINVAL;
  c->cb[mcmax] = ui1[0];
  std::vector<double> nrow(n, 0.0);
  for (long m2 = 1; m2 <= un1[0]; m2++)
    if (ui1[m2]) nrow[un1[m2]] += ui1[m2];
  double *row = c->ca.data() + (size_t)mcmax * n;
  if (!std::equal(nrow.begin(), nrow.end(), row)) {
    std::copy(nrow.begin(), nrow.end(), row);
    c->cinv_ok = false;
  }
  return 0;
}

// gausmat4 on the caller's row store of level l (move3mg.c:3855-3960):
// rows gathered into ELL form, am multicolour sweeps on the device, then
// fre0 = fre1 - A sol0 on the rows [mcmin, mcmax] of the level
static int ell_gausmat4(stk_ctx *c, int am, long mcmax, long mcmin, int l) {
  Level &L = c->lv[l];
  const long n4 = 4 * L.n, base = 4 * L.off;
  std::vector<int> cnt(n4, 0), col((size_t)n4 * kEllMax, 0);
  std::vector<double> val((size_t)n4 * kEllMax, 0.0);
  std::vector<unsigned char> isbc(n4, 0);
  for (long r = 0; r < n4; r++) {
    const long R = base + r;
    const int k = g_arr.num0[R];
    if (k < 0 || k > kEllMax) return STK_EINVAL;
    cnt[r] = k;
    isbc[r] = g_arr.bondm && g_arr.bondm[R] != 0;
    for (int q = 0; q < k; q++) {
      const long j = (long)g_arr.lin0[g_arr.pos0[R] + q] - base;
      if (j < 0 || j >= n4) return STK_EINVAL;  // a row of this level reads only this level
      col[(size_t)r * kEllMax + q] = (int)j;
      val[(size_t)r * kEllMax + q] = g_arr.val1[g_arr.pos0[R] + q];
    }
  }
  // boundary rows by dependency depth: a row after the boundary row it copies
  auto partner = [&](long r) -> long {
    long j = -1;
    for (int q = 0; q < cnt[r]; q++)
      if (col[(size_t)r * kEllMax + q] != r) j = col[(size_t)r * kEllMax + q];
    return j;
  };
  std::vector<int> depth(n4, -1);
  int maxd = -1;
  for (long r = 0; r < n4; r++) {
    if (!cnt[r] || !isbc[r]) continue;
    int d = 0;
    long q = r;
    for (;;) {
      const long j = partner(q);
      if (j < 0 || !isbc[j] || !cnt[j]) break;
      if (++d > 64) return STK_EINVAL;  // boundary rows copying each other in a cycle
      q = j;
    }
    depth[r] = d;
    maxd = std::max(maxd, d);
  }
  std::vector<int> bcrows, bcoff(1, 0);
  for (int d = 0; d <= maxd; d++) {
    for (long r = 0; r < n4; r++)
      if (depth[r] == d) bcrows.push_back((int)r);
    bcoff.push_back((int)bcrows.size());
  }
  if (!L.ell_cnt) {
    if (dalloc(&L.ell_cnt, n4) || dalloc(&L.ell_col, (size_t)n4 * kEllMax) || dalloc(&L.ell_bcrows, n4) ||
        dalloc(&L.ell_val, (size_t)n4 * kEllMax) || dalloc(&L.ell_x, n4) || dalloc(&L.ell_f, n4) ||
        dalloc(&L.ell_f0, n4) || dalloc(&L.ell_isbc, n4) || dalloc(&L.ell_nd, L.n))
      return STK_ENOMEM;
  }
  H2D(c, L.ell_cnt, cnt.data(), n4 * sizeof(int));
  H2D(c, L.ell_col, col.data(), col.size() * sizeof(int));
  H2D(c, L.ell_val, val.data(), val.size() * sizeof(double));
  H2D(c, L.ell_isbc, isbc.data(), n4);
  if (!bcrows.empty()) H2D(c, L.ell_bcrows, bcrows.data(), bcrows.size() * sizeof(int));
  H2D(c, L.ell_x, g_arr.sol0 + base, n4 * 8);
  H2D(c, L.ell_f, g_arr.fre1 + base, n4 * 8);
  H2D(c, L.ell_f0, g_arr.fre0 + base, n4 * 8);
  if (g_arr.nd) H2D(c, L.ell_nd, g_arr.nd + L.off, L.n * sizeof(float));
  else HIPCHK(hipMemsetAsync(L.ell_nd, 0, L.n * sizeof(float), c->st));
  const unsigned nb = blocks(L.n, 256);
  for (int sw = 0; sw < am; sw++) {
    for (int cc = 1; cc < 4; cc++)
      for (int colr = 0; colr < 2; colr++)
        hipLaunchKernelGGL(k_ell_phase, dim3(nb), dim3(256), 0, c->st, L.n, L.Y, L.XY, cc, colr, L.ell_cnt,
                           L.ell_col, L.ell_val, L.ell_f, L.ell_isbc, c->kn.v0koef, L.ell_x);
    hipLaunchKernelGGL(k_ell_pphase, dim3(nb), dim3(256), 0, c->st, L.n, L.ell_cnt, L.ell_col, L.ell_val, L.ell_f,
                       L.ell_isbc, L.ell_nd, c->kn.p0koef, L.ell_x);
    for (int d = 0; d <= maxd; d++) {
      const int nr = bcoff[d + 1] - bcoff[d];
      if (nr)
        hipLaunchKernelGGL(k_ell_bc, dim3(blocks(nr, 256)), dim3(256), 0, c->st, nr, L.ell_bcrows + bcoff[d],
                           L.ell_cnt, L.ell_col, L.ell_val, L.ell_f, L.ell_x);
    }
  }
  const long r0 = std::max(mcmin - base, 0L), r1 = std::min(mcmax - base + 1, n4);
  if (r1 > r0)
    hipLaunchKernelGGL(k_ell_res, dim3(blocks(r1 - r0, 256)), dim3(256), 0, c->st, r0, r1, L.ell_cnt, L.ell_col,
                       L.ell_val, L.ell_f, L.ell_x, L.ell_f0);
  HIPCHK(hipGetLastError());
  D2H(c, g_arr.sol0 + base, L.ell_x, n4 * 8);
  D2H(c, g_arr.fre0 + base, L.ell_f0, n4 * 8);
  return STK_OK;
}

// level l of a bound all-levels array <-> the context's copy of it
static int arr_up(stk_ctx *c, int which, int l, const double *h) {
  int rc = ensure_ref(c, l);
  if (rc) return rc;
  H2D(c, c->lv[l].ref[which], h + 4 * c->lv[l].off, 4 * c->lv[l].n * 8);
  return STK_OK;
}
static int arr_down(stk_ctx *c, int which, int l, double *h) {
  D2H(c, h + 4 * c->lv[l].off, c->lv[l].ref[which], 4 * c->lv[l].n * 8);
  return STK_OK;
}

// parallel/move3mg.c:3855-3960: sweeps on level mgi, then fre0 = fre1 - A sol0
// (rows [mcmin, mcmax]): over the bound arrays' row store, else the bound
// context's matrix-free rows on its own copies
int gausmat4(int am, long mcmax, long mcmin, int mgi) {
  stk_ctx *c = g_bound;
  if (!c) return STK_ENOCTX;
  if (mgi < 0 || mgi > c->L || am < 0) return STK_EINVAL;
  if (g_arr_set) return ell_gausmat4(c, am, mcmax, mcmin, mgi);
  return stk_smooth(c, mgi, am, kTileDefault[0], kTileDefault[1], kTileDefault[2]);
}

static int vprestrict_arrays(stk_ctx *c, int mgi) {
  int rc = arr_up(c, 2, mgi - 1, g_arr.fre0) || arr_up(c, 1, mgi, g_arr.fre1) || arr_up(c, 0, mgi, g_arr.sol0);
  if (rc) return STK_ENOMEM;
  restrict_to<double>(c, mgi, c->lv[mgi - 1].ref[2], c->lv[mgi].ref[1], c->lv[mgi].ref[0], 1);
  HIPCHK(hipGetLastError());
  return arr_down(c, 1, mgi, g_arr.fre1) || arr_down(c, 0, mgi, g_arr.sol0);
}

// parallel/move3mg.c:3966-4354
void vprestrict(int mgi) {
  stk_ctx *c = g_bound;
  if (!c || !c->prepared || mgi < 1 || mgi > c->L) return;
  if (g_arr_set) {
    (void)vprestrict_arrays(c, mgi);
    return;
  }
  if (ensure_ref(c, mgi) || ensure_ref(c, mgi - 1)) return;
  restrict_to<double>(c, mgi, c->lv[mgi - 1].ref[2], c->lv[mgi].ref[1], c->lv[mgi].ref[0], 1);
  hipStreamSynchronize(c->st);
}

int stk_vrestrict32(stk_ctx *c, int l) {
  if (!c || !c->prepared || l < 1 || l > c->L) return STK_EINVAL;
  int rc = ensure_ref(c, l) || ensure_ref(c, l - 1) || ensure_vectors(c, true, false);
  if (rc) return STK_ENOMEM;
  Level &F = c->lv[l - 1], &C = c->lv[l];
  const long nf4 = 4 * F.n, nc4 = 4 * C.n;
  hipLaunchKernelGGL((k_cast<double, float>), dim3(blocks(nf4, 256)), dim3(256), 0, c->st, nf4, F.ref[2], F.r32);
  restrict_to<float>(c, l, F.r32, C.b32, lvx<float>(C), 0);
  hipLaunchKernelGGL((k_cast<float, double>), dim3(blocks(nc4, 256)), dim3(256), 0, c->st, nc4, C.b32, C.ref[1]);
  HIPCHK(hipStreamSynchronize(c->st));
  return hipGetLastError() == hipSuccess ? STK_OK : STK_EHIP;
}

static int vpprolong_arrays(stk_ctx *c, int mgi, int ynreset) {
  int rc = arr_up(c, 0, mgi, g_arr.sol0) || arr_up(c, 0, mgi - 1, g_arr.sol0);
  if (rc) return STK_ENOMEM;
  prolong_from<double>(c, mgi, c->lv[mgi].ref[0], c->lv[mgi - 1].ref[0], 1, ynreset);
  HIPCHK(hipGetLastError());
  return arr_down(c, 0, mgi - 1, g_arr.sol0);
}

// parallel/move3mg.c:3535-3850
void vpprolong(int mgi, int ynreset) {
  stk_ctx *c = g_bound;
  if (!c || !c->prepared || mgi < 1 || mgi > c->L) return;
  if (g_arr_set) {
    (void)vpprolong_arrays(c, mgi, ynreset);
    return;
  }
  if (ensure_ref(c, mgi) || ensure_ref(c, mgi - 1)) return;
  prolong_from<double>(c, mgi, c->lv[mgi].ref[0], c->lv[mgi - 1].ref[0], 1, ynreset);
  hipStreamSynchronize(c->st);
}

static int ronurestrict_run(stk_ctx *c, int mgi) {
  Level &F = c->lv[mgi - 1], &C = c->lv[mgi];
  const bool arr = g_arr_set && g_arr.nu && g_arr.ro;
  if (arr) {
    H2D(c, c->nu + F.off, g_arr.nu + F.off, F.n * sizeof(float));
    H2D(c, c->ro + F.off, g_arr.ro + F.off, F.n * sizeof(float));
  }
  CompTab t;
  for (int a = 0; a < 3; a++) t.ax[a] = C.ax[3][a].tab();
  hipLaunchKernelGGL(k_ronurestrict, dim3(blocks(C.n, 256)), dim3(256), 0, c->st, F.X, F.Y, C.X, C.Y, C.Z, t,
                     c->nu + F.off, c->ro + F.off, c->nu + C.off, c->ro + C.off, c->kn.viscmod);
  HIPCHK(hipGetLastError());
  if (arr) {
    D2H(c, g_arr.nu + C.off, c->nu + C.off, C.n * sizeof(float));
    D2H(c, g_arr.ro + C.off, c->ro + C.off, C.n * sizeof(float));
  }
  HIPCHK(hipStreamSynchronize(c->st));
  return STK_OK;
}

// parallel/move3mg.c:3387-3528
void ronurestrict(int mgi) {
  stk_ctx *c = g_bound;
  if (!c || mgi < 1 || mgi > c->L) return;
  (void)ronurestrict_run(c, mgi);
}

// the reference's post-solve reload into the bound arrays (move3mg.c:2858-2900
// without the pressure shift, then the stress recompute 2908-3046)
static int vpiterate_outputs(stk_ctx *c) {
  Level &F = c->lv[0];
  const double *x = g_arr.sol0;
  for (int m3 = 0; m3 < F.Z; m3++)
    for (int m1 = 0; m1 < F.X; m1++)
      for (int m2 = 0; m2 < F.Y; m2++) {
        const long n = (long)m3 * F.XY + (long)m1 * F.Y + m2;
        if (g_arr.pr && m1 && m2 && m3) g_arr.pr[n] = x[4 * n];
        if (g_arr.vx && m2 < F.Y - 1 && m3 < F.Z - 1) g_arr.vx[n] = (float)x[4 * n + 1];
        if (g_arr.vy && m1 < F.X - 1 && m3 < F.Z - 1) g_arr.vy[n] = (float)x[4 * n + 2];
        if (g_arr.vz && m1 < F.X - 1 && m2 < F.Y - 1) g_arr.vz[n] = (float)x[4 * n + 3];
      }
  // pressure of line 0 of a periodic axis from the last line (move3mg.c:3321-3379, x, y, z in turn)
  if (g_arr.pr) {
    double *pr = g_arr.pr;
    if (c->per[0])
      for (int m3 = 0; m3 < F.Z; m3++)
        for (int m2 = 0; m2 < F.Y; m2++) pr[(long)m3 * F.XY + m2] = pr[(long)m3 * F.XY + (long)(F.X - 1) * F.Y + m2];
    if (c->per[1])
      for (int m3 = 0; m3 < F.Z; m3++)
        for (int m1 = 0; m1 < F.X; m1++) pr[(long)m3 * F.XY + (long)m1 * F.Y] = pr[(long)m3 * F.XY + (long)m1 * F.Y + F.Y - 1];
    if (c->per[2])
      for (int m1 = 0; m1 < F.X; m1++)
        for (int m2 = 0; m2 < F.Y; m2++) pr[(long)m1 * F.Y + m2] = pr[(long)(F.Z - 1) * F.XY + (long)m1 * F.Y + m2];
  }
  float *sp[6] = {g_arr.sxx, g_arr.syy, g_arr.szz, g_arr.sxy, g_arr.sxz, g_arr.syz};
  float *ep[6] = {g_arr.exx, g_arr.eyy, g_arr.ezz, g_arr.exy, g_arr.exz, g_arr.eyz};
  bool any = false;
  for (int k = 0; k < 6; k++) any = any || sp[k] || ep[k];
  if (!any) return STK_OK;
  std::vector<double> sig(6 * (size_t)F.n), eps(6 * (size_t)F.n);
  int rc = stk_stress_strain(c, sig.data(), eps.data());
  if (rc) return rc;
  for (int k = 0; k < 6; k++)
    for (long n = 0; n < F.n; n++) {
      if (sp[k]) sp[k][n] = (float)sig[(size_t)k * F.n + n];
      if (ep[k]) ep[k][n] = (float)eps[(size_t)k * F.n + n];
    }
  return STK_OK;
}

// parallel/move3mg.c:1791-3382: prepare and solve the bound context; with
// bound arrays from their level-0 fields, boundary tables and sol0
void vpiterate(int m0, int f0) {
  (void)m0; (void)f0;
  stk_ctx *c = g_bound;
  if (!c) return;
  if (g_arr_set) {
    const long n0 = c->lv[0].n;
    if (g_arr.nu && g_arr.ro && g_arr.nxx && g_arr.nxy && g_arr.nxz && g_arr.nyz &&
        stk_set_fields(c, g_arr.nu, g_arr.ro, g_arr.nxx, g_arr.nxy, g_arr.nxz, g_arr.nyz))
      return;
    if (g_arr.bondm && g_arr.bondv1 && g_arr.bondn1 && g_arr.bondnum &&
        stk_set_bc_tables(c, g_arr.bondm, (int)*g_arr.bondnum, g_arr.bondv1, g_arr.bondn1))
      return;
    (void)n0;
    if (stk_prepare(c) || stk_set_solution(c, g_arr.sol0)) return;
  } else if (!c->prepared && stk_prepare(c)) {
    return;
  }
  stk_solver_opts o;
  stk_default_opts(&o, c->L);
  stk_solve(c, &o, nullptr);
  if (g_arr_set && !stk_get_solution(c, g_arr.sol0)) (void)vpiterate_outputs(c);
}

// ---- temperature step (parallel/heat3mg.c) --------------------------------

void stk_heat_default_opts(stk_heat_opts *o) {
  if (!o) return;
  memset(o, 0, sizeof *o);
  o->levels = 0;
  o->cycles = 1;
  o->pre[0] = 10;
  o->heatmin = 1e-4;
  o->t0koef = 1.0;
  o->t1koef = 1.0;
  o->maxtkstep = 25.0;
  o->maxit = 2500;
}

int stk_heat_set_fields(stk_ctx *c, const float *ro, const float *cp, const float *kt, const float *tk,
                        const float *ht) {
  if (!c || !ro || !cp || !kt || !tk || !ht) return STK_EINVAL;
  const long n0 = c->lv[0].n;
  for (long i = 0; i < n0; i++)
    if (!std::isfinite(ro[i]) || !std::isfinite(cp[i]) || !std::isfinite(kt[i]) || !std::isfinite(tk[i]) ||
        !std::isfinite(ht[i]))
      return STK_EINVAL;
  int rc = heat_alloc(c);
  if (rc) return rc;
  HeatState *h = c->heat;
  const Level &F = c->lv[0];
  std::vector<float> sp(n0);
  const std::pair<const float *, float *> up[8] = {{ro, h->ro}, {cp, h->cp}, {kt, h->kt}, {tk, h->tk},
                                                   {tk, h->tk0}, {ht, h->tk1}, {ht, h->ht}, {tk, h->tk2}};
  for (const auto &u : up) {
    to_split(u.first, sp.data(), F.X, F.Y, F.Z);
    H2D(c, u.second, sp.data(), n0 * 4);
  }
  hipLaunchKernelGGL(k_heat_rcp, dim3(blocks(n0, 256)), dim3(256), 0, c->st, n0, h->ro, h->cp, h->rcp);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->st));
  h->fields = true;
  h->Lt = -1;
  return STK_OK;
}

int stk_heat_set_bc_tables(stk_ctx *c, const int *tbondm, int nbond, const float *bondv1, const int *bondn1) {
  if (!c || !tbondm || nbond < 0 || (nbond && (!bondv1 || !bondn1))) return STK_EINVAL;
  std::vector<int> tb(c->ntot, 0);
  std::vector<float> v1(2, 0.f);
  std::vector<int> n1(1, 0);
  for (long n = 0; n < c->ntot; n++) {
    const int b = tbondm[n];
    if (!b) continue;
    if (b < 0 || b > nbond || bondn1[b] < 0 || bondn1[b] > c->ntot) return STK_EINVAL;
    tb[n] = (int)n1.size();
    v1.push_back(bondv1[2 * b]);
    v1.push_back(bondn1[b] ? bondv1[2 * b + 1] : 0.f);
    n1.push_back(bondn1[b]);
  }
  c->tbondm.swap(tb);
  c->tbv1.swap(v1);
  c->tbn1.swap(n1);
  c->heat_dirty = true;
  return STK_OK;
}

int stk_heat_sources(stk_ctx *c, const float *et, int adiabyn, int frictyn) {
  if (!c || !c->heat || !c->heat->fields) return STK_ESTATE;
  if (adiabyn < 0 || adiabyn > 2 || frictyn < 0 || frictyn > 1 || (adiabyn && !et)) return STK_EINVAL;
  if ((adiabyn || frictyn) && (!c->prepared || !c->X)) return STK_ESTATE;
  int rc = heat_build(c, std::max(c->heat->built, 0));
  if (rc) return rc;
  HeatState *h = c->heat;
  const Level &F = c->lv[0];
  float *et_d = nullptr, *v3 = nullptr;
  AdvGeo ag;
  memset(&ag, 0, sizeof ag);
  if (dalloc(&et_d, F.n)) return STK_ENOMEM;
  if (et) {
    std::vector<float> sp(F.n);
    to_split(et, sp.data(), F.X, F.Y, F.Z);
    rc = hipMemcpyAsync(et_d, sp.data(), F.n * sizeof(float), hipMemcpyHostToDevice, c->st) ? STK_EHIP : 0;
    if (!rc) rc = hipStreamSynchronize(c->st) ? STK_EHIP : 0;
  } else {
    rc = hipMemsetAsync(et_d, 0, F.n * sizeof(float), c->st) ? STK_EHIP : 0;
  }
  if (!rc && adiabyn) {
    rc = dalloc(&v3, (size_t)3 * F.n);
    if (!rc) rc = adv_geo(c, ag);
    if (!rc)
      hipLaunchKernelGGL(k_adv_planes, dim3(blocks(F.n, 256)), dim3(256), 0, c->st, F.n, c->X, nullptr, v3,
                         nullptr);
  }
  if (!rc && frictyn) {
    const size_t n6 = (size_t)6 * F.n;
    if (!c->sig && dalloc(&c->sig, n6)) rc = STK_ENOMEM;
    if (!rc && !c->epsr && dalloc(&c->epsr, n6)) rc = STK_ENOMEM;
    if (!rc)
      hipLaunchKernelGGL(k_post_stress, dim3(blocks(F.n, 256)), dim3(256), 0, c->st, post_geo(c), c->X, c->sig,
                         c->epsr);
  }
  if (!rc) {
    hipLaunchKernelGGL(k_heat_sources, dim3(blocks(F.n, 256)), dim3(256), 0, c->st, heat_geo(c, 0), ag,
                       h->lv[0].isbc, c->X, v3, c->sig, c->epsr, h->ro, h->tk, h->ht, et_d, h->tk1, adiabyn,
                       frictyn, c->kn.gx, c->kn.gy, c->kn.gz);
    rc = hipGetLastError() ? STK_EHIP : 0;
    if (!rc) rc = hipStreamSynchronize(c->st) ? STK_EHIP : 0;
  }
  (void)hipStreamSynchronize(c->st);
  dfree(et_d);
  dfree(v3);
  return rc;
}

int stk_heat_markers(stk_ctx *c, long nmark, const float *mx, const float *my, const float *mz, const int *markt,
                     float *markk, double heatdif, double timestep, double zdeep, double tdeep) {
  if (!c || !c->heat || !c->heat->fields) return STK_ESTATE;
  if (nmark < 0 || (nmark && (!mx || !my || !mz || !markk)) || !std::isfinite(heatdif) || !std::isfinite(timestep))
    return STK_EINVAL;
  if (nmark == 0) return STK_OK;
  if (nmark >= (1l << 31)) return STK_EINVAL;
  int rc = heat_build(c, std::max(c->heat->built, 0));
  if (rc) return rc;
  HeatState *h = c->heat;
  const Level &F = c->lv[0];
  const int ncell = (F.X - 1) * (F.Y - 1) * (F.Z - 1);
  AdvGeo ag;
  rc = adv_geo(c, ag);
  if (rc) return rc;
  float *d[4] = {};
  int *dt = nullptr, *key = nullptr, *val = nullptr, *skey = nullptr, *sval = nullptr, *start = nullptr;
  unsigned char *flag = nullptr;
  double *dtk1 = nullptr, *sol1 = nullptr;
  void *tmp = nullptr;
  size_t tmpb = 0;
  const float *hsrc[4] = {mx, my, mz, markk};
  for (int k = 0; k < 4 && !rc; k++) {
    rc = dalloc(&d[k], nmark);
    if (!rc) rc = hipMemcpyAsync(d[k], hsrc[k], nmark * sizeof(float), hipMemcpyHostToDevice, c->st) ? STK_EHIP : 0;
  }
  if (!rc && markt) {
    rc = dalloc(&dt, nmark);
    if (!rc) rc = hipMemcpyAsync(dt, markt, nmark * sizeof(int), hipMemcpyHostToDevice, c->st) ? STK_EHIP : 0;
  }
  if (!rc && (dalloc(&flag, nmark) || dalloc(&dtk1, nmark))) rc = STK_ENOMEM;
  if (!rc)
    hipLaunchKernelGGL(k_heat_mark_a, dim3(blocks(nmark, 256)), dim3(256), 0, c->st, heat_geo(c, 0), ag, nmark, d[0],
                       d[1], d[2], dt, d[3], h->tk, h->tk2, h->kt, h->cp, h->ro, heatdif, timestep, zdeep, tdeep, flag,
                       dtk1);
  if (!rc && heatdif != 0.0) {
    // markers by cell (stable radix sort of the cell keys), then the nodes' gather
    if (dalloc(&key, nmark) || dalloc(&val, nmark) || dalloc(&skey, nmark) || dalloc(&sval, nmark) ||
        dalloc(&start, (size_t)ncell + 2) || dalloc(&sol1, F.n))
      rc = STK_ENOMEM;
    if (!rc) {
      hipLaunchKernelGGL(k_marker_cells, dim3(blocks(nmark, 256)), dim3(256), 0, c->st, nmark, d[0], d[1], d[2],
                         F.g64[0], F.g64[1], F.g64[2], F.X, F.Y, F.Z, key, val);
      rc = hipcub::DeviceRadixSort::SortPairs(nullptr, tmpb, key, skey, val, sval, (int)nmark, 0, 32, c->st)
               ? STK_EHIP : 0;
      if (!rc) rc = dalloc((unsigned char **)&tmp, tmpb);
      if (!rc)
        rc = hipcub::DeviceRadixSort::SortPairs(tmp, tmpb, key, skey, val, sval, (int)nmark, 0, 32, c->st)
                 ? STK_EHIP : 0;
    }
    if (!rc) {
      hipLaunchKernelGGL(k_cell_start, dim3(blocks(nmark + 1, 256)), dim3(256), 0, c->st, nmark, skey, ncell + 1,
                         start);
      hipLaunchKernelGGL(k_heat_mark_gather, dim3(blocks(F.n, 256)), dim3(256), 0, c->st, ag, start, sval, d[0],
                         d[1], d[2], flag, dtk1, sol1);
      hipLaunchKernelGGL(k_heat_mark_anti, dim3(blocks(nmark, 256)), dim3(256), 0, c->st, ag, nmark, d[0], d[1],
                         d[2], dt, d[3], sol1);
    }
  }
  if (!rc) rc = hipGetLastError() ? STK_EHIP : 0;
  if (!rc) rc = hipMemcpyAsync(markk, d[3], nmark * sizeof(float), hipMemcpyDeviceToHost, c->st) ? STK_EHIP : 0;
  if (!rc) rc = hipStreamSynchronize(c->st) ? STK_EHIP : 0;
  (void)hipStreamSynchronize(c->st);
  for (int k = 0; k < 4; k++) dfree(d[k]);
  dfree(dt); dfree(key); dfree(val); dfree(skey); dfree(sval); dfree(start); dfree(flag); dfree(dtk1); dfree(sol1);
  unsigned char *tb = (unsigned char *)tmp;
  dfree(tb);
  return rc;
}

int stk_heat_get_fields(stk_ctx *c, float *ro, float *cp, float *kt, float *tk, float *tk0, float *tk1) {
  if (!c || !c->heat) return STK_ESTATE;
  HeatState *h = c->heat;
  HIPCHK(hipStreamSynchronize(c->st));
  std::vector<float> sp;
  const std::pair<float *, const float *> dn[3] = {{ro, h->ro}, {cp, h->cp}, {kt, h->kt}};
  for (const auto &d : dn) {
    if (!d.first) continue;
    for (int l = 0; l <= c->L; l++) {
      const Level &L = c->lv[l];
      sp.resize(L.n);
      D2H(c, sp.data(), d.second + L.off, L.n * 4);
      from_split(sp.data(), d.first + L.off, L.X, L.Y, L.Z);
    }
  }
  const Level &F = c->lv[0];
  const std::pair<float *, const float *> d0[3] = {{tk, h->tk}, {tk0, h->tk0}, {tk1, h->tk1}};
  for (const auto &d : d0) {
    if (!d.first) continue;
    sp.resize(F.n);
    D2H(c, sp.data(), d.second, F.n * 4);
    from_split(sp.data(), d.first, F.X, F.Y, F.Z);
  }
  return STK_OK;
}

int stk_heat_prepare(stk_ctx *c, int levels, double timestep, double t0koef, double t1koef) {
  int rc = heat_ready(c, levels);
  if (rc) return rc;
  if (!(timestep > 0) || !std::isfinite(t0koef) || !std::isfinite(t1koef)) return STK_EINVAL;
  HeatState *h = c->heat;
  h->t0 = t0koef;
  h->t1 = t1koef;
  heat_props(c, levels);
  heat_start(c, timestep);
  h->Lt = levels;
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->st));
  return STK_OK;
}

int stk_heat_gausmatt(stk_ctx *c, int level, int sweeps) {
  if (!c || !c->heat || c->heat->Lt < 0) return STK_ESTATE;
  if (level < 0 || level > c->heat->Lt || sweeps < 0) return STK_EINVAL;
  heat_gs(c, level, sweeps);
  heat_res(c, level);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->st));
  return STK_OK;
}

int stk_heat_trestrict(stk_ctx *c, int level) {
  if (!c || !c->heat || c->heat->Lt < 0) return STK_ESTATE;
  if (level < 1 || level > c->heat->Lt) return STK_EINVAL;
  heat_restrict(c, level);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->st));
  return STK_OK;
}

int stk_heat_tprolong(stk_ctx *c, int level, int ynreset) {
  if (!c || !c->heat || c->heat->Lt < 0) return STK_ESTATE;
  if (level < 1 || level > c->heat->Lt) return STK_EINVAL;
  heat_prolong(c, level, ynreset ? 1 : 0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->st));
  return STK_OK;
}

int stk_heat_err(stk_ctx *c, double *heatsum, double *bondsum) {
  if (!c || !c->heat || c->heat->Lt < 0) return STK_ESTATE;
  double hs = 0, bs = 0;
  int rc = heat_errs(c, &hs, &bs, nullptr, nullptr, nullptr, nullptr);
  if (rc) return rc;
  HIPCHK(hipGetLastError());
  if (heatsum) *heatsum = hs;
  if (bondsum) *bondsum = bs;
  return STK_OK;
}

int stk_heat_vec(stk_ctx *c, int which, int level, double *host, int dir) {
  if (!c || !c->heat) return STK_ESTATE;
  if (!host || level < 0 || level > c->L || which < 0 || which > 2) return STK_EINVAL;
  HeatState *h = c->heat;
  const Level &L = c->lv[level];
  double *d = (which == STK_SOL0 ? h->x : which == STK_FRE1 ? h->b : h->r) + L.off;
  std::vector<double> sp(L.n);
  if (dir) {
    to_split(host, sp.data(), L.X, L.Y, L.Z);
    H2D(c, d, sp.data(), L.n * sizeof(double));
  } else {
    D2H(c, sp.data(), d, L.n * sizeof(double));
    from_split(sp.data(), host, L.X, L.Y, L.Z);
  }
  return STK_OK;
}

int stk_heat_time(stk_ctx *c, int which, int level, int reps, double *ms_avg) {
  if (!c || !c->heat || c->heat->Lt < 0) return STK_ESTATE;
  if (level < 0 || level > c->heat->Lt || reps < 1 || !ms_avg || which < 0 || which > 3) return STK_EINVAL;
  HeatState *h = c->heat;
  const Level &L = c->lv[level];
  const HeatLevel &H = h->lv[level];
  const HeatGeo g = heat_geo(c, level);
  const int hy = (L.Y + 1) / 2;
  auto launch = [&]() {
    if (which == 0) {
      hipLaunchKernelGGL(k_heat_gs, dim3(blocks((long)hy * L.X * L.Z, 256)), dim3(256), 0, c->st, g, 0, 0, L.X,
                         make_hdiv((unsigned)hy), make_hdiv((unsigned)L.X), H.isbc, h->kt + L.off, h->rcp + L.off,
                         h->dt, h->t0, h->b + L.off, h->x + L.off);
    } else if (which == 1) {
      heat_gs(c, level, 1);
    } else if (which == 2) {
      heat_res(c, level);
    } else {
      hipLaunchKernelGGL(k_stream_copy, dim3(blocks(L.n, 256)), dim3(256), 0, c->st, L.n, h->b + L.off,
                         h->r + L.off);
    }
  };
  launch();
  HIPCHK(hipEventRecord(c->ev0, c->st));
  for (int i = 0; i < reps; i++) launch();
  HIPCHK(hipEventRecord(c->ev1, c->st));
  HIPCHK(hipEventSynchronize(c->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  *ms_avg = ms / reps;
  return STK_OK;
}

int stk_heat_titerate(stk_ctx *c, double timestep, const stk_heat_opts *o, stk_heat_stats *st) {
  if (!o) return STK_EINVAL;
  int rc = heat_ready(c, o->levels);
  if (rc) return rc;
  if (!std::isfinite(timestep) || timestep < 0 || o->cycles < 1 || o->maxit < 1 || !(o->heatmin >= 0) ||
      !std::isfinite(o->t0koef) || !std::isfinite(o->t1koef) || !(o->maxtkstep >= 0))
    return STK_EINVAL;
  for (int l = 0; l <= o->levels; l++)
    if (o->pre[l] < 0 || o->post[l] < 0) return STK_EINVAL;
  const auto t_start = std::chrono::steady_clock::now();
  HeatState *h = c->heat;
  h->t0 = o->t0koef;
  h->t1 = o->t1koef;
  stk_heat_stats S;
  memset(&S, 0, sizeof S);
  S.converged = 1;
  double rate = 0;
  rc = heat_rate(c, &rate);
  if (rc) return rc;
  S.tkstep = 1e+30;
  if (rate) S.tkstep = o->maxtkstep / rate;
  if (timestep != 0 && S.tkstep != 0) {
    // the step's starting temperature, for the markers' update (heat3mg.c:377-381)
    HIPCHK(hipMemcpyAsync(h->tk2, h->tk, c->lv[0].n * sizeof(float), hipMemcpyDeviceToDevice, c->st));
    heat_props(c, o->levels);
    h->Lt = o->levels;
    double tsum = 0;
    const double tsum0 = timestep;
    do {
      double dt = std::min(tsum0, S.tkstep);
      if (dt > tsum0 - tsum) dt = tsum0 - tsum;
      heat_start(c, dt);
      int it = 0;
      double hs = 0;
      do {
        for (int k = 0; k < o->cycles; k++) heat_cycle(c, o);
        HIPCHK(hipGetLastError());
        rc = heat_errs(c, &hs, &S.bondsum, &S.heatnum, &S.bondnum, &S.tkmin, &S.tkmax);
        if (rc) return rc;
        if (!std::isfinite(hs)) return STK_ENAN;
        it++;
      } while (hs > o->heatmin && it < o->maxit);
      if (hs > o->heatmin) S.converged = 0;
      S.heatsum = hs;
      S.iterations += it;
      S.substeps++;
      tsum += dt;
    } while (tsum < tsum0);
  }
  HIPCHK(hipStreamSynchronize(c->st));
  S.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  if (st) *st = S;
  return S.converged ? STK_OK : STK_ENOCONV;
}

}  // extern "C"
