// Atom data (numerical radial functions) shared by the atom-file reader and
// the problem builder of libdfx.  Not installed.
#pragma once
#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "common.hpp"

struct dfx_atom
{
  struct Function
  {
    int                 q[3]; // n, l, m
    double              cutoff, smoothness;
    std::vector<double> x;          // knots (empty: the file's radial grid)
    std::vector<double> y, b, c, d; // spline: y + b h + c h^2 + d h^3 on [x_i, x_i+1)
    double              cLeft = 0;  // left of x_0: y_0 + b_0 h + cLeft h^2
    // mixed (UPF local potential): Y00 (spline(r) for r <= x_last, tail / r
    // beyond), no cutoff function
    bool                mixed = false;
    double              tail  = 0;
    const std::vector<double> &knots(const std::vector<double> &grid) const
    {
      return x.empty() ? grid : x;
    }
  };
  std::string                                     file;
  std::map<std::string, std::vector<std::string>> text; // direct children of <atom>
  std::vector<double>                             r;
  std::map<std::string, std::vector<Function>>   fields;
};

namespace dfx
{
  // a spline on the device: knots x and coefficients y, b, c, d (n each)
  struct SplineDev
  {
    const double *x, *y, *b, *c, *d;
    u32           n;
    double        cl = 0.0; // left extrapolation y_0 + b_0 h + cl h^2
  };

  // value and slope at r (left of x_0 linear with slope b_0, right of
  // x_{n-1} quadratic: utils::Spline::operator() / deriv,
  // src/utils/Spline.cpp:490-518)
  __device__ inline void splineEval(const SplineDev &S, double r, double &v, double &dv)
  {
    if (!(r >= S.x[0])) // also NaN
      {
        const double h = r - S.x[0];
        v              = (S.cl * h + S.b[0]) * h + S.y[0];
        dv             = 2.0 * S.cl * h + S.b[0];
        return;
      }
    u32 lo = 0, hi = S.n; // x[lo] <= r < x[hi] (x[n] = +inf)
    while (hi - lo > 1)
      {
        const u32 mid = (lo + hi) >> 1;
        if (S.x[mid] <= r)
          lo = mid;
        else
          hi = mid;
      }
    const double h = r - S.x[lo];
    if (lo == S.n - 1)
      {
        v  = (S.c[lo] * h + S.b[lo]) * h + S.y[lo];
        dv = 2.0 * S.c[lo] * h + S.b[lo];
        return;
      }
    v  = ((S.d[lo] * h + S.c[lo]) * h + S.b[lo]) * h + S.y[lo];
    dv = (3.0 * S.d[lo] * h + 2.0 * S.c[lo]) * h + S.b[lo];
  }

  // smooth cutoff f_c(r) and its slope (SmoothCutoffFunctions.cpp:25-70)
  __device__ inline double smoothCutoff(double r, double rc, double d, double &dv)
  {
    auto f1  = [](double x) { return x <= 0.0 ? 0.0 : exp(-1.0 / x); };
    auto f1d = [&](double x) { return x <= 0.0 ? 0.0 : f1(x) / (x * x); };
    const double y = 1.0 - d * (r - rc) / rc;
    const double a = f1(y), b = f1(1.0 - y);
    const double s = a + b;
    dv             = ((f1d(y) * b + a * f1d(1.0 - y)) / (s * s)) * (-d / rc);
    return a / s;
  }

  // f = R(r) Y_lm(dx / r) f_c(r) and grad f at offset dx from the centre,
  // l <= 2 (SphericalDataNumerical.cpp:40-140): zero past rc (1 + 1/sm);
  // at r = 0 the gradient is 0
  __device__ inline void radialYlm(const SplineDev &S, int l, int m, double rc, double sm,
                                   const double *dx, double &f, double *g)
  {
    const double r = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
    f              = 0.0;
    g[0] = g[1] = g[2] = 0.0;
    if (!(r <= rc + rc / sm))
      return;
    double R, dR, dfc;
    splineEval(S, r, R, dR);
    const double fc = smoothCutoff(r, rc, sm, dfc);
    const double u = R * fc, du = dR * fc + R * dfc; // radial part and its slope
    if (l == 0)
      {
        const double y00 = 0.28209479177387814; // 1 / (2 sqrt(pi))
        f                = y00 * u;
        if (r > 0.0)
          for (int k = 0; k < 3; ++k)
            g[k] = y00 * du * dx[k] / r;
      }
    else if (r == 0.0)
      {
        // at the centre the reference takes theta = phi = 0: Y_l0 = sqrt((2l + 1)
        // / (4 pi)), Y_lm = 0 for m != 0 (the gradient is not defined there: 0)
        if (m == 0)
          f = (l == 1 ? 0.48860251190291992 : 0.6307831305050401) * u;
      }
    else if (l == 2)
      {
        // Y_2m = k_m S_m(x) / r^2 with S = xy, yz, 2z^2 - x^2 - y^2, xz, x^2 - y^2
        // for m = -2 .. 2 (B_2|m| D_m P_2^|m| Q_m without the Condon-Shortley phase)
        const double x = dx[0], y = dx[1], z = dx[2];
        double       k, S, dS[3];
        switch (m)
          {
            case -2: k = 1.0925484305920792, S = x * y, dS[0] = y, dS[1] = x, dS[2] = 0.0; break;
            case -1: k = 1.0925484305920792, S = y * z, dS[0] = 0.0, dS[1] = z, dS[2] = y; break;
            case 0:
              k = 0.31539156525252005, S = 2.0 * z * z - x * x - y * y;
              dS[0] = -2.0 * x, dS[1] = -2.0 * y, dS[2] = 4.0 * z;
              break;
            case 1: k = 1.0925484305920792, S = x * z, dS[0] = z, dS[1] = 0.0, dS[2] = x; break;
            default:
              k = 0.5462742152960396, S = x * x - y * y;
              dS[0] = 2.0 * x, dS[1] = -2.0 * y, dS[2] = 0.0;
              break;
          }
        const double r2 = r * r;
        const double q = u / r2, dq = du / r2 - 2.0 * u / (r2 * r); // f = k S q(r)
        f              = k * S * q;
        for (int c = 0; c < 3; ++c)
          g[c] = k * (q * dS[c] + S * dq * dx[c] / r);
      }
    else
      {
        // Y_1m = sqrt(3 / (4 pi)) (y, z, x)_m / r for m = -1, 0, 1
        const double c1   = 0.48860251190291992;
        const int    comp = m == -1 ? 1 : (m == 0 ? 2 : 0);
        const double q = u / r, dq = du / r - u / (r * r); // f = c1 x_comp q(r)
        f              = c1 * dx[comp] * q;
        for (int k = 0; k < 3; ++k)
          g[k] = c1 * (dx[comp] * dq * dx[k] / r + (k == comp ? q : 0.0));
      }
  }

  // the enrichment functions atom a of a problem spec carries: their count
  // and the radius of the atom's support (largest rc (1 + 1/sm)); the
  // analytic Slater set (five functions, the spec's cutoff) unless the spec
  // names atom data
  inline const std::vector<dfx_atom::Function> &atomFunctions(const dfx_problem_spec &s, u32 a)
  {
    const char *field = s.enrichField ? s.enrichField : "orbital";
    const auto  it    = s.atomData[a]->fields.find(field);
    if (it == s.atomData[a]->fields.end())
      throw Error(DFX_E_ARG, std::string("atom data of atom ") + std::to_string(a) +
                               " holds no field " + field);
    return it->second;
  }
  inline void atomEnrichment(const dfx_problem_spec &s, u32 a, u32 &nFun, double &support)
  {
    if (!s.atomData)
      {
        nFun    = 5;
        support = s.rcut + s.rcut / s.smooth;
        return;
      }
    const auto &F = atomFunctions(s, a);
    nFun          = (u32)F.size();
    support       = 0.0;
    for (const auto &f : F)
      support = std::max(support, f.cutoff + f.cutoff / f.smoothness);
  }
} // namespace dfx
