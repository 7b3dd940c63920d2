// Scalar math of the BFE path, usable on host and device (gfx950).
//
// Coordinate maps follow exptool/basis/compatibility.py:16-99; cell location
// and interpolation weights follow exptool/basis/spheresl.py:301-335 (SL) and
// exptool/basis/eof.py:354-457 (EOF); Legendre recurrences follow
// spheresl.py:664-770.  All geometry is float64: bin indices must agree with
// the float64 reference, and this part of the work is small next to the table
// gathers it feeds.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define BFE_HD __host__ __device__ __forceinline__

namespace bfe {

constexpr double kPi = 3.14159265358979323846;
constexpr double kFourPi = 4.0 * kPi;

// Logarithms for the coordinate maps.  The library log and asinh carry
// double-double steps for the last half ulp (98 and 193 VALU instructions on
// gfx950); the maps need a few ulp.  log(f 2^k) = k ln2 + 2 atanh(s) with
// s = (f - 1)/(f + 1) and f in [sqrt(1/2), sqrt(2)), so |s| <= 0.1716; the
// series 2 (s + s^3/3 + ... + s^21/21) is truncated below 1e-18 relative.
// ln2 is split as float32(ln 2) + the rest, so that k ln2_hi is exact.
constexpr double kLn2 = 0.6931471805599453;
constexpr double kLn2Hi = 0.693147182464599609375;
constexpr double kLn2Lo = -1.904654299957768e-09;
constexpr double kSqrtHalf = 0.7071067811865476;

// k ln2 + log((1 + s)/(1 - s)) for s = num/den
BFE_HD double log_series(double num, double den, int k) {
  const double s = num / den;
  const double z = s * s;
  double R = 2.0 / 21.0;
  R = fma(R, z, 2.0 / 19.0);
  R = fma(R, z, 2.0 / 17.0);
  R = fma(R, z, 2.0 / 15.0);
  R = fma(R, z, 2.0 / 13.0);
  R = fma(R, z, 2.0 / 11.0);
  R = fma(R, z, 2.0 / 9.0);
  R = fma(R, z, 2.0 / 7.0);
  R = fma(R, z, 2.0 / 5.0);
  R = fma(R, z, 2.0 / 3.0);
  const double kd = (double)k;
  return fma(kd, kLn2Hi, fma(s * z, R, 2.0 * s) + kd * kLn2Lo);
}

// log(u), u > 0 finite; 0 -> -inf, inf -> inf, negative or NaN -> NaN.
BFE_HD double log_pos(double u) {
  int k;
  double f = frexp(u, &k);   // u = f 2^k, f in [0.5, 1)
  if (f < kSqrtHalf) {
    f += f;
    --k;
  }
  const double v = log_series(f - 1.0, f + 1.0, k);   // f - 1 is exact
  if (u > 0.0 && u < INFINITY) return v;
  return u == 0.0 ? -INFINITY : (u > 0.0 ? u : NAN);
}

// log(1 + w) for w >= 0: below 1 + w = sqrt(2) the series takes w itself
// (s = w/(2 + w)), so the rounding of 1 + w does not enter.
BFE_HD double log1p_nonneg(double w) {
  int k;
  double f = frexp(1.0 + w, &k);
  if (f < kSqrtHalf) {
    f += f;
    --k;
  }
  const bool near1 = w < 0.41421356237309503;   // sqrt(2) - 1
  return log_series(near1 ? w : f - 1.0, near1 ? 2.0 + w : f + 1.0, near1 ? 0 : k);
}

// asinh(q) for q >= 0: log1p(q + q^2/(1 + sqrt(q^2 + 1))), and log(2q) past
// 2^28 where sqrt(q^2 + 1) = q in float64.  Within 3 ulp of numpy.arcsinh
// over [1e-310, 1e300]; inf -> inf, NaN -> NaN.
BFE_HD double asinh_nonneg(double q) {
  const double q2 = q * q;
  const double w = q + q2 / (1.0 + sqrt(q2 + 1.0));
  const bool big = q > 268435456.0;
  const double v = log1p_nonneg(big ? q - 1.0 : w) + (big ? kLn2 : 0.0);
  return q < INFINITY ? v : q + q;
}

// compatibility.py:16-47 (scalar): negative r maps to 0.
BFE_HD double r_to_xi(double r, int cmap, double scale) {
  if (r < 0.0) return 0.0;
  if (cmap == 1) return (r - scale) / (r + scale);   // (r/a - 1)/(r/a + 1), one division
  if (cmap == 2) return log_pos(r);
  return r;
}

// compatibility.py:65-80 (returns d xi / d r).
BFE_HD double d_xi_to_r(double xi, int cmap, double scale) {
  if (cmap == 1) return 0.5 * (1.0 - xi) * (1.0 - xi) / scale;
  if (cmap == 2) return exp(-xi);
  return 1.0;
}

// compatibility.py:83-91 (eps 1e-8, the Python map that runs).
// ihscale = 1 / hscale.
BFE_HD double z_to_y(double z, double ihscale) {
  return (z / (fabs(z) + 1.e-8)) * asinh_nonneg(fabs(z * ihscale));
}

// Planar angle of (x, y) as cos/sin, with atan2 semantics: the zero vector
// goes through atan2 itself so that signed zeros behave as in numpy.arctan2.
BFE_HD void planar_angle(double x, double y, double& c, double& s) {
  double rho = sqrt(x * x + y * y);
  if (rho > 0.0) {
    const double irho = 1.0 / rho;
    c = x * irho;
    s = y * irho;
  } else {
    double ph = atan2(y, x);
    c = cos(ph);
    s = sin(ph);
  }
}

// cos(m phi), sin(m phi) for m = 0..mmax by the Chebyshev recurrence
// (spheresl.py:773-795).
template <int MAXM>
BFE_HD void chebyshev(double c1, double s1, int mmax, double* cm, double* sm) {
  cm[0] = 1.0;
  sm[0] = 0.0;
  if (mmax >= 1) {
    cm[1] = c1;
    sm[1] = s1;
  }
  for (int m = 2; m <= mmax && m <= MAXM; ++m) {
    cm[m] = 2.0 * c1 * cm[m - 1] - cm[m - 2];
    sm[m] = 2.0 * c1 * sm[m - 1] - sm[m - 2];
  }
}

// Index of P_l^m in a packed triangular array (l >= m).
BFE_HD int lm_index(int l, int m) { return l * (l + 1) / 2 + m; }

// legendre_R (spheresl.py:664-700): P_l^m(x) with Condon-Shortley phase into
// p[lm_index(l, m)].  Non-finite values (x NaN, i.e. r = 0 in the batch
// evaluator) are replaced by 0 as at spheresl.py:698.
BFE_HD void legendre(int lmax, double x, double* p) {
  const bool bad = !(x == x);
  p[0] = 1.0;
  if (bad) {
    for (int i = 1; i < (lmax + 1) * (lmax + 2) / 2; ++i) p[i] = 0.0;
    return;
  }
  double pll = 1.0;
  if (lmax > 0) {
    double somx2 = sqrt((1.0 - x) * (1.0 + x));
    double fact = 1.0;
    for (int m = 1; m <= lmax; ++m) {
      pll *= -fact * somx2;
      p[lm_index(m, m)] = pll;
      fact += 2.0;
    }
  }
  for (int m = 0; m < lmax; ++m) {
    double pl2 = p[lm_index(m, m)];
    double pl1 = x * (2.0 * m + 1.0) * pl2;
    p[lm_index(m + 1, m)] = pl1;
    for (int l = m + 2; l <= lmax; ++l) {
      double v = (x * (2 * l - 1) * pl1 - (l + m - 1) * pl2) / (l - m);
      p[lm_index(l, m)] = v;
      pl2 = pl1;
      pl1 = v;
    }
  }
}

// dlegendre_R (spheresl.py:706-770): derivative with x clamped to
// +-(1 - 1e-8) for the derivative only; p must already hold legendre(x).
BFE_HD void dlegendre(int lmax, double x, const double* p, double* dp) {
  dp[0] = 0.0;
  if (!(x == x)) {
    for (int i = 1; i < (lmax + 1) * (lmax + 2) / 2; ++i) dp[i] = 0.0;
    return;
  }
  const double mineps = 1.e-8;
  if (1.0 - fabs(x) < mineps) x = (x > 0) ? 1.0 - mineps : -(1.0 - mineps);
  double somx2 = 1.0 / (x * x - 1.0);
  for (int l = 1; l <= lmax; ++l) {
    for (int m = 0; m < l; ++m)
      dp[lm_index(l, m)] = somx2 * (x * l * p[lm_index(l, m)] - (l + m) * p[lm_index(l - 1, m)]);
    dp[lm_index(l, l)] = somx2 * x * l * p[lm_index(l, l)];
  }
  for (int i = 1; i < (lmax + 1) * (lmax + 2) / 2; ++i)
    if (!isfinite(dp[i])) dp[i] = 0.0;
}

// Cell location of the SL radial tables (spheresl.py:308-328 / :123-140):
// clamped xi, cell i in [0, numr-2], weights x1, x2.
struct SLCell {
  double x;   // mapped, clamped coordinate
  int i;
  double x1, x2;
};

// idxi = 1 / (xi[1] - xi[0]).
BFE_HD SLCell sl_locate(double r, int cmap, double scale, const double* xi, int numr,
                        double idxi) {
  SLCell c;
  double x = r_to_xi(r, cmap, scale);
  if (cmap == 1) {
    if (x < -1.0) x = -1.0;
    if (x >= 1.0) x = 1.0 - 1.0e-08;
  } else if (cmap == 2) {
    // intended semantics of spheresl.py:128-130 (NameError as written)
    if (!(x >= xi[0])) x = xi[0];
    if (x > xi[numr - 1]) x = xi[numr - 1];
  }
  double f = floor((x - xi[0]) * idxi);
  int i;
  if (!(f == f)) i = 0;
  else if (f < 0.0) i = 0;
  else if (f > (double)(numr - 2)) i = numr - 2;
  else i = (int)f;
  c.x = x;
  c.i = i;
  c.x1 = (xi[i + 1] - x) * idxi;
  c.x2 = (x - xi[i]) * idxi;
  return c;
}

// EOF bin location (eof.py:354-427) and bilinear weights (eof.py:443-451).
// Negative X clamps to 0; ix clamps to numx-1 while X is left unclamped above,
// so the weights extrapolate linearly past the outer edge (eof.py:414-415).
struct EOFCell {
  int ix, iy;
  double c00, c10, c01, c11;
};

// ihscale, idx, idy: reciprocals of hscale and of the grid steps dx, dy.
BFE_HD EOFCell eof_locate(double R, double z, int cmap, double ascale, double ihscale,
                          double xmin, double idx, double ymin, double idy, int numx, int numy) {
  EOFCell c;
  double X = (r_to_xi(R, cmap, ascale) - xmin) * idx;
  double Y = (z_to_y(z, ihscale) - ymin) * idy;
  if (!(X >= 0.0)) X = (X == X) ? 0.0 : X;
  if (!(Y >= 0.0)) Y = (Y == Y) ? 0.0 : Y;
  double fx = X < (double)(numx - 1) ? trunc(X) : (double)(numx - 1);
  double fy = Y < (double)(numy - 1) ? trunc(Y) : (double)(numy - 1);
  c.ix = (fx == fx) ? (int)fx : 0;
  c.iy = (fy == fy) ? (int)fy : 0;
  double dx0 = c.ix + 1.0 - X, dy0 = c.iy + 1.0 - Y;
  double dx1 = X - c.ix, dy1 = Y - c.iy;
  c.c00 = dx0 * dy0;
  c.c10 = dx1 * dy0;
  c.c01 = dx0 * dy1;
  c.c11 = dx1 * dy1;
  return c;
}

}  // namespace bfe
