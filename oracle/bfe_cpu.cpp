// FP64 C++/OpenMP restatement of the BFE hot path on the host.
//
// TEST INFRASTRUCTURE ONLY: bench.py times it on all host cores as the
// multi-core CPU baseline, and tests/ check it against the NumPy oracle
// (oracle/bfe_oracle.py), which the reference's own vectors pin.  Nothing in
// exptool_amd links or loads it.
//
// Semantics are those of the NumPy oracle without compat flags:
//   cpu_sl_accumulate  compute_coefficients_solitary  spheresl.py:567-656
//   cpu_eof_accumulate accumulate                      eof.py:492-640
//   cpu_field_eval     Fields.return_forces_cart       potential.py:445-497
//                      (eof.force_eval eof.py:756-872, spheresl.force_eval
//                      spheresl.py:1107-1234 with cos/sin(m phi), the
//                      quadratic derivative stencil of
//                      get_halo_dens_pot_force spheresl.py:106-160 taken at
//                      the particle in the first cell)
// Particles are split over threads in contiguous blocks; each thread sums
// into its own coefficient array and the arrays are added in thread order.
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <string.h>

#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;

struct SL {
  int lmax, nmax, numr, cmap;
  double scale;
  const double *xi, *ev, *ef, *p0;   // ev [l][n], ef [l][n][numr]
};

struct EOF {
  int mmax, norder, numx, numy, cmap;
  double ascale, hscale, xmin, dx, ymin, dy;
  const double *potC, *rfC, *zfC, *potS, *rfS, *zfS;   // [m][n][numx+1][numy+1]
};

double r_to_xi(double r, int cmap, double s) {
  if (r < 0.0) return 0.0;
  if (cmap == 1) return (r / s - 1.0) / (r / s + 1.0);
  if (cmap == 2) return log(r);
  return r;
}

double dxi_dr(double xi, int cmap, double s) {
  if (cmap == 1) return 0.5 * (1.0 - xi) * (1.0 - xi) / s;
  if (cmap == 2) return exp(-xi);
  return 1.0;
}

double z_to_y(double z, double h) { return z / (fabs(z) + 1.0e-8) * asinh(fabs(z / h)); }

// N_lm = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!), x sqrt 2 for m > 0 (spheresl.py:823-863)
std::vector<double> norms(int lmax) {
  std::vector<double> f((lmax + 1) * (lmax + 1), 0.0);
  for (int l = 0; l <= lmax; ++l)
    for (int m = 0; m <= l; ++m) {
      double ratio = 1.0;
      for (int k = l - m + 1; k <= l + m; ++k) ratio /= (double)k;
      double v = sqrt((0.5 * l + 0.25) / kPi * ratio);
      if (m) v *= sqrt(2.0);
      f[l * (lmax + 1) + m] = v;
    }
  return f;
}

// P_l^m(x) with the Condon-Shortley phase, non-finite -> 0 (spheresl.py:664-700);
// with dp, also dP/dx with x clamped to +-(1 - 1e-8) (spheresl.py:706-770).
void legendre(int L, double x, double* p, double* dp) {
  const int W = L + 1;
  for (int k = 0; k < W * W; ++k) p[k] = 0.0;
  p[0] = 1.0;
  double pll = 1.0;
  if (L > 0) {
    const double somx2 = sqrt((1.0 - x) * (1.0 + x));
    double fact = 1.0;
    for (int m = 1; m <= L; ++m) {
      pll *= -fact * somx2;
      p[m * W + m] = pll;
      fact += 2.0;
    }
  }
  for (int m = 0; m < L; ++m) {
    double pl2 = p[m * W + m];
    double pl1 = x * (2.0 * m + 1) * pl2;
    p[(m + 1) * W + m] = pl1;
    for (int l = m + 2; l <= L; ++l) {
      const double v = (x * (2 * l - 1) * pl1 - (l + m - 1) * pl2) / (l - m);
      p[l * W + m] = v;
      pl2 = pl1;
      pl1 = v;
    }
  }
  for (int k = 0; k < W * W; ++k)
    if (!isfinite(p[k])) p[k] = 0.0;
  if (!dp) return;
  for (int k = 0; k < W * W; ++k) dp[k] = 0.0;
  double xc = x;
  if (1.0 - fabs(xc) < 1.0e-8) xc = xc > 0 ? 1.0 - 1.0e-8 : -(1.0 - 1.0e-8);
  const double s = 1.0 / (xc * xc - 1.0);
  for (int l = 1; l <= L; ++l) {
    for (int m = 0; m < l; ++m) dp[l * W + m] = s * (xc * l * p[l * W + m] - (l + m) * p[(l - 1) * W + m]);
    dp[l * W + l] = s * xc * l * p[l * W + l];
  }
  for (int k = 0; k < W * W; ++k)
    if (!isfinite(dp[k])) dp[k] = 0.0;
}

// SL radial cell (spheresl.py:308-328): clamped xi, cell index, x1, x2.
void sl_cell(const SL& t, double r, double& xv, int& i, double& x1, double& x2) {
  double x = r_to_xi(r, t.cmap, t.scale);
  if (t.cmap == 1) {
    if (x < -1.0) x = -1.0;
    if (x >= 1.0) x = 1.0 - 1.0e-8;
  } else if (t.cmap == 2) {
    if (x < t.xi[0]) x = t.xi[0];
    if (x > t.xi[t.numr - 1]) x = t.xi[t.numr - 1];
  }
  double xmin = t.xi[0];
  for (int k = 1; k < t.numr; ++k) xmin = t.xi[k] < xmin ? t.xi[k] : xmin;
  const double dxi = t.xi[1] - t.xi[0];
  double fl = floor((x - xmin) / dxi);
  if (!isfinite(fl)) fl = 0.0;
  if (fl < 0) fl = 0;
  if (fl > t.numr - 2) fl = t.numr - 2;
  i = (int)fl;
  x1 = (t.xi[i + 1] - x) / dxi;
  x2 = (x - t.xi[i]) / dxi;
  xv = x;
}

// EOF bins and bilinear weights (eof.py:354-457, upper edge extrapolated).
void eof_cell(const EOF& t, double R, double z, int& ix, int& iy, double c[4]) {
  double X = (r_to_xi(R, t.cmap, t.ascale) - t.xmin) / t.dx;
  double Y = (z_to_y(z, t.hscale) - t.ymin) / t.dy;
  if (X < 0) X = 0;
  if (Y < 0) Y = 0;
  ix = (int)fmin(trunc(fmin(X, (double)t.numx)), (double)(t.numx - 1));
  iy = (int)fmin(trunc(fmin(Y, (double)t.numy)), (double)(t.numy - 1));
  const double dx0 = ix + 1.0 - X, dy0 = iy + 1.0 - Y, dx1 = X - ix, dy1 = Y - iy;
  c[0] = dx0 * dy0;
  c[1] = dx1 * dy0;
  c[2] = dx0 * dy1;
  c[3] = dx1 * dy1;
}

inline double bilin(const double* T, size_t plane, int ny1, int ix, int iy, const double c[4]) {
  const double* a = T + (size_t)ix * ny1 + iy;
  (void)plane;
  return c[0] * a[0] + c[1] * a[ny1] + c[2] * a[1] + c[3] * a[ny1 + 1];
}

void sum_threads(std::vector<std::vector<double>>& part, double* out, size_t K) {
  for (size_t k = 0; k < K; ++k) out[k] = 0.0;
  for (auto& p : part)
    for (size_t k = 0; k < K; ++k) out[k] += p[k];
}

}  // namespace

extern "C" {

int cpu_threads() { return omp_get_max_threads(); }

int cpu_sl_accumulate(int lmax, int nmax, int numr, int cmap, double scale, const double* xi,
                      const double* ev, const double* ef, const double* p0, const double* px,
                      const double* py, const double* pz, const double* pm, int64_t n, int no_odd,
                      double* out) {
  const SL t{lmax, nmax, numr, cmap, scale, xi, ev, ef, p0};
  const int W = lmax + 1;
  const size_t K = (size_t)W * W * nmax;
  const std::vector<double> N = norms(lmax);
  std::vector<double> isev((size_t)W * nmax);
  for (size_t k = 0; k < isev.size(); ++k) isev[k] = 1.0 / sqrt(ev[k]);
  const int nt = omp_get_max_threads();
  std::vector<std::vector<double>> part(nt, std::vector<double>(K, 0.0));
#pragma omp parallel
  {
    const int tid = omp_get_thread_num();
    double* acc = part[tid].data();
    std::vector<double> P(W * W), pot((size_t)W * nmax);
#pragma omp for schedule(static)
    for (int64_t q = 0; q < n; ++q) {
      const double x = px[q], y = py[q], z = pz[q], mass = pm[q];
      double r = sqrt(x * x + y * y + z * z);
      if (!(r >= 1.0e-10)) r = 1.0e-10;   // nanmax([r, 1e-10])
      const double costh = z / r, phi = atan2(y, x);
      legendre(lmax, costh, P.data(), nullptr);
      double xv, x1, x2;
      int i;
      sl_cell(t, r, xv, i, x1, x2);
      const double pp = x1 * p0[i] + x2 * p0[i + 1];
      for (int l = 0; l <= lmax; ++l)
        for (int k = 0; k < nmax; ++k) {
          const double* e = ef + ((size_t)l * nmax + k) * numr;
          pot[(size_t)l * nmax + k] = (x1 * e[i] + x2 * e[i + 1]) * isev[(size_t)l * nmax + k] * pp;
        }
      for (int l = 0; l <= lmax; ++l) {
        if (no_odd && (l & 1)) continue;
        const double* pl = &pot[(size_t)l * nmax];
        for (int m = 0; m <= l; ++m) {
          const double w = N[l * W + m] * P[l * W + m] * (-4.0 * kPi) * mass;
          if (m == 0) {
            double* o = acc + (size_t)(l * l) * nmax;
            for (int k = 0; k < nmax; ++k) o[k] += w * pl[k];
          } else {
            const double wc = w * cos(phi * m), ws = w * sin(phi * m);
            double* oc = acc + (size_t)(l * l + 2 * m - 1) * nmax;
            double* os = acc + (size_t)(l * l + 2 * m) * nmax;
            for (int k = 0; k < nmax; ++k) {
              oc[k] += wc * pl[k];
              os[k] += ws * pl[k];
            }
          }
        }
      }
    }
  }
  sum_threads(part, out, K);
  return 0;
}

int cpu_eof_accumulate(int mmax, int norder, int numx, int numy, int cmap, double ascale,
                       double hscale, double xmin, double dx, double ymin, double dy,
                       const double* potC, const double* potS, const double* px,
                       const double* py, const double* pz, const double* pm, int64_t n,
                       double* cos_out, double* sin_out) {
  const EOF t{mmax, norder, numx, numy, cmap, ascale, hscale, xmin, dx, ymin, dy,
              potC, nullptr, nullptr, potS, nullptr, nullptr};
  const size_t K = (size_t)(mmax + 1) * norder;
  const int ny1 = numy + 1;
  const size_t plane = (size_t)(numx + 1) * ny1;
  const int nt = omp_get_max_threads();
  std::vector<std::vector<double>> part(nt, std::vector<double>(2 * K, 0.0));
#pragma omp parallel
  {
    double* acc = part[omp_get_thread_num()].data();
#pragma omp for schedule(static)
    for (int64_t q = 0; q < n; ++q) {
      const double x = px[q], y = py[q], z = pz[q];
      const double R = sqrt(x * x + y * y + 1.0e-10), phi = atan2(y, x);
      int ix, iy;
      double c[4];
      eof_cell(t, R, z, ix, iy, c);
      for (int m = 0; m <= mmax; ++m) {
        const double wc = -4.0 * kPi * cos(phi * m) * pm[q];
        const double ws = -4.0 * kPi * sin(phi * m) * pm[q];
        for (int k = 0; k < norder; ++k) {
          const size_t row = ((size_t)m * norder + k) * plane;
          acc[(size_t)m * norder + k] += wc * bilin(potC + row, plane, ny1, ix, iy, c);
          acc[K + (size_t)m * norder + k] += ws * bilin(potS + row, plane, ny1, ix, iy, c);
        }
      }
    }
  }
  std::vector<double> tot(2 * K);
  sum_threads(part, tot.data(), 2 * K);
  memcpy(cos_out, tot.data(), K * sizeof(double));
  memcpy(sin_out, tot.data() + K, K * sizeof(double));
  return 0;
}

int cpu_field_eval(int lmax, int nmax, int numr, int cmap, double scale, const double* xi,
                   const double* ev, const double* ef, const double* p0, int mmax, int norder,
                   int numx, int numy, int ecmap, double ascale, double hscale, double xmin,
                   double dx, double ymin, double dy, const double* potC, const double* rfC,
                   const double* zfC, const double* potS, const double* rfS, const double* zfS,
                   const double* eof_cos, const double* eof_sin, const double* sl_coef,
                   const double* px, const double* py, const double* pz, int64_t n, double* fx,
                   double* fy, double* fz, double* pot) {
  const SL t{lmax, nmax, numr, cmap, scale, xi, ev, ef, p0};
  const EOF e{mmax, norder, numx, numy, ecmap, ascale, hscale, xmin, dx, ymin, dy,
              potC, rfC, zfC, potS, rfS, zfS};
  const int W = lmax + 1;
  const int ny1 = numy + 1;
  const size_t plane = (size_t)(numx + 1) * ny1;
  const std::vector<double> N = norms(lmax);
  std::vector<double> isev((size_t)W * nmax);
  for (size_t k = 0; k < isev.size(); ++k) isev[k] = 1.0 / sqrt(ev[k]);
  double ximin = xi[0];
  for (int k = 1; k < numr; ++k) ximin = xi[k] < ximin ? xi[k] : ximin;
  const double dxi = xi[1] - xi[0];
#pragma omp parallel
  {
    std::vector<double> P(W * W), dP(W * W), pot_l((size_t)W * nmax), dpot_l((size_t)W * nmax);
#pragma omp for schedule(static)
    for (int64_t q = 0; q < n; ++q) {
      const double x = px[q], y = py[q], z = pz[q];
      const double r2 = sqrt(x * x + y * y) + 1.0e-15;
      const double r3 = sqrt(r2 * r2 + z * z) + 1.0e-15;
      const double costh = z / r3, phi = atan2(y, x);
      // ---- disc: eof.force_eval at (r2, z, phi)
      int ix, iy;
      double c[4];
      eof_cell(e, r2, z, ix, iy, c);
      double dfr = 0, dfz = 0, dfp = 0, dp = 0;
      for (int m = 0; m <= mmax; ++m) {
        const double cm = cos(phi * m), sm = sin(phi * m);
        for (int k = 0; k < norder; ++k) {
          const size_t row = ((size_t)m * norder + k) * plane;
          const double ac = eof_cos[(size_t)m * norder + k];
          const double vp = bilin(potC + row, plane, ny1, ix, iy, c);
          const double vr = bilin(rfC + row, plane, ny1, ix, iy, c);
          const double vz = bilin(zfC + row, plane, ny1, ix, iy, c);
          if (m == 0) {
            dp += ac * vp;
            dfr += ac * vr;
            dfz += ac * vz;
            continue;
          }
          const double as = eof_sin[(size_t)m * norder + k];
          const double sp = bilin(potS + row, plane, ny1, ix, iy, c);
          const double sr = bilin(rfS + row, plane, ny1, ix, iy, c);
          const double sz = bilin(zfS + row, plane, ny1, ix, iy, c);
          dp += ac * cm * vp + as * sm * sp;
          dfr += ac * cm * vr + as * sm * sr;
          dfz += ac * cm * vz + as * sm * sz;
          dfp += m * (ac * sm * vp - as * cm * sp);
        }
      }
      // ---- halo: spheresl.force_eval at (r3, costh, phi)
      double xv, x1, x2;
      int i;
      sl_cell(t, r3, xv, i, x1, x2);
      const double pp = x1 * p0[i] + x2 * p0[i + 1];
      const double fac = dxi_dr(xv, cmap, scale) / dxi;
      const int cc = i == 0 ? 1 : i;                 // stencil centre
      const double ps = i == 0 ? x2 - 1.0 : x2;      // the particle's offset from it
      for (int l = 0; l <= lmax; ++l)
        for (int k = 0; k < nmax; ++k) {
          const double* f = ef + ((size_t)l * nmax + k) * numr;
          const double is = isev[(size_t)l * nmax + k];
          pot_l[(size_t)l * nmax + k] = (x1 * f[i] + x2 * f[i + 1]) * is * pp;
          dpot_l[(size_t)l * nmax + k] =
              fac * ((ps - 0.5) * f[cc - 1] * p0[cc - 1] - 2.0 * ps * f[cc] * p0[cc] +
                     (ps + 0.5) * f[cc + 1] * p0[cc + 1]) * is;
        }
      legendre(lmax, costh, P.data(), dP.data());
      double hp0 = 0, hpr = 0, hp1 = 0, hpt = 0, hpp = 0;
      for (int l = 0; l <= lmax; ++l)
        for (int m = 0; m <= l; ++m) {
          const double f = N[l * W + m];
          const double* a = sl_coef + (size_t)(l * l + (m ? 2 * m - 1 : 0)) * nmax;
          const double* b = sl_coef + (size_t)(l * l + 2 * m) * nmax;
          double pc = 0, pS = 0, dc = 0, dS = 0;
          for (int k = 0; k < nmax; ++k) {
            pc += a[k] * pot_l[(size_t)l * nmax + k];
            dc += a[k] * dpot_l[(size_t)l * nmax + k];
            if (m) {
              pS += b[k] * pot_l[(size_t)l * nmax + k];
              dS += b[k] * dpot_l[(size_t)l * nmax + k];
            }
          }
          if (l == 0) {
            hp0 = f * pc;
            hpr = f * dc;
            continue;
          }
          if (m == 0) {
            hp1 += f * P[l * W] * pc;
            hpr += f * P[l * W] * dc;
            hpt += f * dP[l * W] * pc;
          } else {
            const double cm = cos(phi * m), sm = sin(phi * m);
            hp1 += f * P[l * W + m] * (pc * cm + pS * sm);
            hpr += f * P[l * W + m] * (dc * cm + dS * sm);
            hpt += f * dP[l * W + m] * (pc * cm + pS * sm);
            hpp += f * P[l * W + m] * m * (-pc * sm + pS * cm);
          }
        }
      // ---- return_forces_cart assembly (potential.py:455-494)
      if (r3 < ximin) {
        hpp = 0.0;
        dfp = 0.0;
      }
      const double fxd = dfr * (x / r2) - dfp * (y / (r2 * r2));
      const double fxh = -(hpr * (x / r3) - hpt * (x * z / (r3 * r3 * r3))) + hpp * (y / (r2 * r2));
      const double fyd = dfr * (y / r2) + dfp * (x / (r2 * r2));
      const double fyh = -(hpr * (y / r3) - hpt * (y * z / (r3 * r3 * r3))) - hpp * (x / (r2 * r2));
      const double fzh = -(hpr * (z / r3) + hpt * ((r2 * r2) / (r3 * r3 * r3)));
      fx[q] = fxd + fxh;
      fy[q] = fyd + fyh;
      fz[q] = dfz + fzh;
      pot[q] = dp + hp1 + hp0;
    }
  }
  return 0;
}

}  // extern "C"
