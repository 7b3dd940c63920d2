// Point evaluation of folded SL and EOF tables (device side).
//
// "Folded" tables hold sum_n a_n T_n per table node for a fixed coefficient
// set (an exact linear rewrite of the reference sums over n, whose only effect
// is rounding), so evaluation gathers (2 mmax + 1) values per EOF node and
// (lmax + 1)^2 per SL node instead of norder / nmax times that.
#pragma once

#include "bfe_internal.hpp"

namespace bfe {

struct SLOut {
  double den0, den1, pot0, pot1, potr, pott, potp;
};

struct EOFOut {
  double p0, p, fr, fp, fz, d0, d;
};

// SL evaluation at (r, cos theta, phi) from a folded table F (groups E, D1,
// D2[, Dn] of `stride` floats per node).  Follows get_halo_dens_pot_force
// (spheresl.py:106-160) and the l,m sums of all_eval_particles / force_eval
// (spheresl.py:1276-1341, :1199-1228); N_lm and the coefficients are folded in.
//   flags: BFE_COMPAT_SL_EDGE  first-cell stencil one cell outward (:150-153)
//          BFE_COMPAT_SL_COS_L cos/sin(l phi) for 0 < m (:1222-1225)
//          BFE_COMPAT_SL_DENS  den1 uses P_1^0 for m = 0 (:1323)
template <bool DENS>
__device__ __forceinline__ void sl_eval_folded(const SLGeom& g, const float* __restrict__ F,
                                               int stride, int groups, int lmax_use, double r,
                                               double costh, double c1, double s1,
                                               uint32_t flags, SLOut& o) {
  const SLCell cell = sl_locate(r, g.cmap, g.scale, g.xi, g.numr);
  const int i = cell.i;
  const double x1 = cell.x1, x2 = cell.x2;
  const double P0 = x1 * g.p0[i] + x2 * g.p0[i + 1];
  const double D0 = DENS ? x1 * g.d0[i] + x2 * g.d0[i + 1] : 0.0;
  const double dxi = g.xi[1] - g.xi[0];
  const double fac = d_xi_to_r(cell.x, g.cmap, g.scale) / dxi;
  const int ctr = (i == 0) ? 1 : i;
  const double ps = (i == 0) ? ((flags & BFE_COMPAT_SL_EDGE) ? x2 : x2 - 1.0) : x2;
  const int gs = groups * stride;
  const float* __restrict__ Fi = F + (size_t)i * gs;
  const float* __restrict__ Fi1 = Fi + gs;
  const float* __restrict__ Fc = F + (size_t)ctr * gs;

  double P[kMaxLM], dP[kMaxLM];
  legendre(lmax_use, costh, P);
  dlegendre(lmax_use, costh, P, dP);
  double cm[kMaxL + 1], sm[kMaxL + 1];
  chebyshev<kMaxL>(c1, s1, lmax_use, cm, sm);

  auto pot = [&](int row) { return (x1 * (double)Fi[row] + x2 * (double)Fi1[row]) * P0; };
  auto dpot = [&](int row) { return fac * ((double)Fc[stride + row] + ps * (double)Fc[2 * stride + row]); };
  auto den = [&](int row) {
    return (x1 * (double)Fi[3 * stride + row] + x2 * (double)Fi1[3 * stride + row]) * D0;
  };

  o.pot0 = pot(0);
  o.potr = dpot(0);
  o.den0 = DENS ? den(0) : 0.0;
  o.pot1 = o.pott = o.potp = o.den1 = 0.0;
  const bool cosl = (flags & BFE_COMPAT_SL_COS_L) != 0;
  const bool dq = (flags & BFE_COMPAT_SL_DENS) != 0;
  for (int l = 1; l <= lmax_use; ++l) {
    const int r0 = l * l;
    {
      const double pv = pot(r0), dv = dpot(r0);
      const double Pl = P[lm_index(l, 0)];
      o.pot1 += Pl * pv;
      o.potr += Pl * dv;
      o.pott += dP[lm_index(l, 0)] * pv;
      if (DENS) o.den1 += (dq ? P[lm_index(1, 0)] : Pl) * den(r0);
    }
    for (int m = 1; m <= l; ++m) {
      const int rc = r0 + 2 * m - 1, rs = rc + 1;
      const int ang = cosl ? l : m;
      const double cmv = cm[ang], smv = sm[ang];
      const double pc = pot(rc), pss = pot(rs);
      const double dc = dpot(rc), ds = dpot(rs);
      const double Plm = P[lm_index(l, m)], dPlm = dP[lm_index(l, m)];
      o.pot1 += Plm * (pc * cmv + pss * smv);
      o.potr += Plm * (dc * cmv + ds * smv);
      o.pott += dPlm * (pc * cmv + pss * smv);
      o.potp += Plm * m * (-pc * smv + pss * cmv);
      if (DENS) o.den1 += Plm * (den(rc) * cmv + den(rs) * smv);
    }
  }
}

// EOF evaluation at planar radius R, height z, angle phi from a folded table
// F (per node: pot, rforce, zforce[, dens] groups of q = 2 mmax + 1 entries,
// cos m=0..M then sin m=1..M).  Bilinear weights as eof.py:443-451; sums as
// force_eval (eof.py:842-870) and accumulated_eval_particles (:1092-1138):
// fp = sum m (C sin m phi potC - S cos m phi potS) = -dPhi/dphi.
template <bool DENS>
__device__ __forceinline__ void eof_eval_folded(const EOFGeom& g, const float* __restrict__ F,
                                                int stride, double R, double z, double c1,
                                                double s1, EOFOut& o) {
  const EOFCell cell = eof_locate(R, z, g.cmap, g.ascale, g.hscale, g.xmin, g.dx, g.ymin, g.dy,
                                  g.numx, g.numy);
  const int ny1 = g.numy + 1;
  const size_t n00 = (size_t)cell.ix * ny1 + cell.iy;
  const float* __restrict__ A = F + n00 * stride;           // (ix, iy)
  const float* __restrict__ B = A + stride;                  // (ix, iy+1)
  const float* __restrict__ C = A + (size_t)ny1 * stride;    // (ix+1, iy)
  const float* __restrict__ D = C + stride;                  // (ix+1, iy+1)
  const int M = g.mmax;
  const int q = 2 * M + 1;
  auto v = [&](int k) {
    return cell.c00 * (double)A[k] + cell.c10 * (double)C[k] + cell.c01 * (double)B[k] +
           cell.c11 * (double)D[k];
  };
  double cm[kMaxM + 1], sm[kMaxM + 1];
  chebyshev<kMaxM>(c1, s1, M, cm, sm);
  o.p0 = v(0);
  o.fr = v(q);
  o.fz = v(2 * q);
  o.d0 = DENS ? v(3 * q) : 0.0;
  o.p = o.fp = o.d = 0.0;
  for (int m = 1; m <= M; ++m) {
    const double pc = v(m), pss = v(M + m);
    o.p += cm[m] * pc + sm[m] * pss;
    o.fr += cm[m] * v(q + m) + sm[m] * v(q + M + m);
    o.fz += cm[m] * v(2 * q + m) + sm[m] * v(2 * q + M + m);
    o.fp += m * (sm[m] * pc - cm[m] * pss);
    if (DENS) o.d += cm[m] * v(3 * q + m) + sm[m] * v(3 * q + M + m);
  }
}

// Frozen disc + halo field at a Cartesian point, Fields.return_forces_cart
// (potential.py:445-497).  comp receives the 8 reference outputs.
struct FieldTables {
  SLGeom slg;
  EOFGeom eg;
  const float* slF;
  const float* eofF;
  int sl_stride, sl_groups, lmax_use, eof_stride;
  uint32_t flags;
  double ximin;      // min(xi_halo) for the azimuthal-force guard (:483-485)
  double crot, srot; // cos/sin of rotpos
};

__device__ __forceinline__ void field_eval_point(const FieldTables& T, double x, double y,
                                                 double z, double comp[8]) {
  const double r2 = sqrt(x * x + y * y) + 1.e-15;
  const double r3 = sqrt(r2 * r2 + z * z) + 1.e-15;
  const double costh = z / r3;
  double c0, s0;
  planar_angle(x, y, c0, s0);
  const double c1 = c0 * T.crot - s0 * T.srot;   // cos(phi + rotpos)
  const double s1 = s0 * T.crot + c0 * T.srot;   // sin(phi + rotpos)
  EOFOut d;
  eof_eval_folded<false>(T.eg, T.eofF, T.eof_stride, r2, z, c1, s1, d);
  SLOut h;
  sl_eval_folded<false>(T.slg, T.slF, T.sl_stride, T.sl_groups, T.lmax_use, r3, costh, c1, s1,
                        T.flags, h);
  double dfp = d.fp, hfp = h.potp;
  if (r3 < T.ximin) {
    hfp = 0.0;
    dfp = 0.0;
  }
  const double dfr = d.fr, dfz = d.fz, dp = d.p + d.p0;
  const double hfr = h.potr, hft = h.pott, hp = h.pot1 + h.pot0;
  const double r2sq = r2 * r2, r3cu = r3 * r3 * r3;
  comp[0] = dfr * (x / r2) - dfp * (y / r2sq);
  comp[1] = -1.0 * (hfr * (x / r3) - hft * (x * z / r3cu)) + hfp * (y / r2sq);
  comp[2] = dfr * (y / r2) + dfp * (x / r2sq);
  comp[3] = -1.0 * (hfr * (y / r3) - hft * (y * z / r3cu)) - hfp * (x / r2sq);
  comp[4] = dfz;
  comp[5] = -1.0 * (hfr * (z / r3) + hft * (r2sq / r3cu));
  comp[6] = dp;
  comp[7] = hp;
}

}  // namespace bfe
