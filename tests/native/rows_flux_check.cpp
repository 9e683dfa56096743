// Host check of the flux form of the velocity rows (stk_rows.h): for random
// spacings, viscosities, grad-div weights and slot values, the flux form's
// S and D equal the coefficient form's sum over slots and its diagonal
// s[1] (FP64, to round-off).  With a one-slot mistake planted in a copy of
// the x row (a pair with its members swapped), the check must fail.
// Exit 0 on success; prints the worst relative differences.
#include <cmath>
#include <cstdio>
#include <random>

#include "../../pardiso2_amd/csrc/stk_rows.h"

using namespace stk;

template <typename T>
void planted_x(const Sp<T> &p, const Vis<T> &w, T p2, const T *xs, T &S, T &D) {
  const T xkf = T(2) * p.bx0, ykf = p.ay1, zkf = p.az1;
  S = T(0); D = T(0);
  fsnn<T, 0, 2, 1, 8, 9, 12, 13>(-xkf, w.c0, p.ax0, p.ay1, p.az1, xs, S, D);
  fsnn<T, 0, 1, 3, 10, 11, 14, 15>(xkf, w.c1, p.ax1, p.ay1, p.az1, xs, S, D);
  fshear<T, 4, 1, 10, 8>(-ykf, w.s0, p.by0, p.bx0, xs, S, D);  // members of the vy pair swapped
  fshear<T, 1, 5, 9, 11>(ykf, w.s1, p.by1, p.bx0, xs, S, D);
  fshear<T, 6, 1, 12, 14>(-zkf, w.t0, p.bz0, p.bx0, xs, S, D);
  fshear<T, 1, 7, 13, 15>(zkf, w.t1, p.bz1, p.bx0, xs, S, D);
  fpair<T, 16, 17>(xkf, xs, S, D);
}

int main() {
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> u(0.5, 2.0), v(-1.0, 1.0);
  double worst[3] = {0, 0, 0}, worst_planted = 1e30;
  for (int trial = 0; trial < 2000; trial++) {
    Sp<double> p{u(rng), u(rng), u(rng), u(rng), u(rng), u(rng), u(rng), u(rng), u(rng), u(rng), u(rng), u(rng)};
    Vis<double> w{u(rng), u(rng), u(rng), u(rng), u(rng), u(rng), u(rng), u(rng)};
    const double p2 = (trial & 1) ? 0.1 : 0.0;
    double xs[18];
    for (int q = 0; q < 18; q++) xs[q] = v(rng);
    for (int c = 1; c <= 3; c++) {
      double s[18] = {0}, S = 0, D = 0;
      if (c == 1) { row_x_v<double>(p, w, p2, s); flux_x<double>(p, w, p2, xs, S, D); }
      if (c == 2) { row_y_v<double>(p, w, p2, s); flux_y<double>(p, w, p2, xs, S, D); }
      if (c == 3) { row_z_v<double>(p, w, p2, s); flux_z<double>(p, w, p2, xs, S, D); }
      double ref = 0, mag = 0;
      for (int q = 1; q <= 17; q++) { ref += s[q] * xs[q]; mag += std::fabs(s[q] * xs[q]); }
      worst[c - 1] = std::fmax(worst[c - 1], std::fabs(S - ref) / mag + std::fabs(D - s[1]) / std::fabs(s[1]));
      if (c == 1 && !(trial & 1)) {
        double Sp_ = 0, Dp = 0;
        planted_x<double>(p, w, p2, xs, Sp_, Dp);
        worst_planted = std::fmin(worst_planted, std::fabs(Sp_ - ref) / mag + std::fabs(Dp - s[1]) / std::fabs(s[1]));
      }
    }
  }
  // uniform spacing: the level-constant form (FP32) against the FP64 coefficient form
  double worst_u[3] = {0, 0, 0};
  std::uniform_real_distribution<double> h(0.01, 0.1);
  for (int trial = 0; trial < 2000; trial++) {
    const double hx = h(rng), hy = h(rng), hz = h(rng);
    const double ax = 1.0 / hx, bx = 0.5 / hx, ay = 1.0 / hy, by = 0.5 / hy, az = 1.0 / hz, bz = 0.5 / hz;
    Sp<double> p{ax, ax, bx, bx, ay, ay, by, by, az, az, bz, bz};
    Vis<double> w{u(rng), u(rng), u(rng), u(rng), u(rng), u(rng), u(rng), u(rng)};
    const double p2 = (trial & 1) ? 0.1 : 0.0;
    const FluxU fu = flux_u_consts(ax, bx, ay, by, az, bz, p2);
    const Vis<float> wf{(float)w.c0, (float)w.c1, (float)w.s0, (float)w.s1, (float)w.t0, (float)w.t1, (float)w.d0,
                        (float)w.d1};
    double xs[18];
    float xf[18];
    for (int q = 0; q < 18; q++) { xs[q] = v(rng); xf[q] = (float)xs[q]; xs[q] = xf[q]; }
    for (int c = 1; c <= 3; c++) {
      double s[18] = {0};
      float S = 0, D = 0;
      if (c == 1) { row_x_v<double>(p, w, p2, s); flux_u<1>(wf, fu.g[0], p2 != 0, xf, S, D); }
      if (c == 2) { row_y_v<double>(p, w, p2, s); flux_u<2>(wf, fu.g[1], p2 != 0, xf, S, D); }
      if (c == 3) { row_z_v<double>(p, w, p2, s); flux_u<3>(wf, fu.g[2], p2 != 0, xf, S, D); }
      double ref = 0, mag = 0;
      for (int q = 1; q <= 17; q++) { ref += s[q] * xs[q]; mag += std::fabs(s[q] * xs[q]); }
      // FP32: a few ulps of the largest terms (viscosities rounded to FP32 too)
      worst_u[c - 1] = std::fmax(worst_u[c - 1], std::fabs(S - ref) / mag + std::fabs(D - s[1]) / std::fabs(s[1]));
    }
  }
  std::printf("worst relative difference x %.3e y %.3e z %.3e; planted mistake at least %.3e; "
              "uniform FP32 x %.3e y %.3e z %.3e\n", worst[0], worst[1], worst[2], worst_planted, worst_u[0],
              worst_u[1], worst_u[2]);
  const bool ok = worst[0] < 1e-14 && worst[1] < 1e-14 && worst[2] < 1e-14 && worst_u[0] < 2e-6 &&
                  worst_u[1] < 2e-6 && worst_u[2] < 2e-6;
  // the planted swap changes S in all but a vanishing fraction of the trials
  return ok ? (worst_planted > 0 ? 0 : 2) : 1;
}
