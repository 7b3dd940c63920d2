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
  double fr0, fz0;   // the m = 0 parts of fr and fz (eof.force_eval perturb=True)
};

// Folded table layouts (float32, streamed as float4):
//  SL : per radial node i, one 8-float block per (l, m) pair in m-major order
//       (m = 0..L, l = m..L):
//         [E_c(i), E_s(i), dE_c(i), dE_s(i), D1_c, D1_s, D2_c, D2_s]
//       with dE(i) = E(i+1) - E(i), so that E at x2 is E(i) + x2 dE(i)
//       with D1, D2 at the stencil centre c = max(i, 1); after the numr
//       nodes of these, the density blocks per node and pair:
//         [Dn_c(i), Dn_s(i), Dn_c(i+1), Dn_s(i+1)].
//       E = sum_n a N_lm ef/sqrt(ev), Dn = sum_n a N_lm ef sqrt(ev),
//       D1 = (G_{c+1} - G_{c-1})/2, D2 = G_{c-1} - 2 G_c + G_{c+1},
//       G = sum_n a N_lm ef p0/sqrt(ev)   (S entries are 0 for m = 0).
//       Forces read two float4 per pair, both from node i.
//  EOF: per grid node, one 8-float block per m = 0..M:
//       [potC, rforceC, zforceC, densC, potS, rforceS, zforceS, densS];
//       then, in float64, the m = 0 cos block of each edge node
//       (eof_edge_index): [pot, rforce, zforce, dens].
// SL table values are interpolated in float32 (they are stored in float32;
// the weights are convex) and combined in float64.  EOF values are
// interpolated in float64: far from the disc plane the mapped grid cells are
// wide and the force is small against the corner values, where float32
// interpolation gave force errors of 3e-5.
__host__ __device__ inline int sl_npairs(int L) { return (L + 1) * (L + 2) / 2; }

// SL evaluation at (r, cos theta, phi): get_halo_dens_pot_force
// (spheresl.py:106-160) and the l,m sums of all_eval_particles / force_eval
// (spheresl.py:1276-1341, :1199-1228) with N_lm and the coefficients folded in.
// The associated Legendre functions are generated column by column (m outer,
// l inner) with the recurrences of legendre_R / dlegendre_R
// (spheresl.py:664-770), in the order the table is stored.
//   flags: BFE_COMPAT_SL_EDGE  first-cell stencil one cell outward (:150-153)
//          BFE_COMPAT_SL_COS_L cos/sin(l phi) for 0 < m (:1222-1225)
//          BFE_COMPAT_SL_DENS  den1 uses P_1^0 for m = 0 (:1323)
// LDS view of staged float4 data: reads through it are LDS instructions, not
// flat ones (which would pass through the vector-memory address path).
typedef float F4v __attribute__((ext_vector_type(4)));
typedef const F4v __attribute__((address_space(3)))* LdsF4;
__device__ __forceinline__ float4 ld4(LdsF4 p, int k) {
  const F4v v = p[k];
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float4 ld4(const float4* __restrict__ p, int k) { return p[k]; }

// Angular sums of sl_eval_folded.  The azimuthal factors cos/sin(m phi) do
// not depend on l, so the l sums are taken per m first (sum_l P_l^m E,
// sum_l P_l^m D, sum_l dP_l^m E for cos and sin) and combined with the
// factors once per m.  The l loops are branch-free: m = 0 runs alone (l = 0
// peeled) and each m > 0 peels l = m, so no accumulator passes through a
// per-lane or per-iteration branch.  A NaN cos(theta) is replaced by 0 for
// the recurrences and its l > 0 sums are zeroed at the end, which is what
// skipping them term by term gives.
template <bool DENS, typename PA>
__device__ __forceinline__ void sl_angular(const SLGeom& g, PA A, const float4* __restrict__ Nd,
                                           int L, float x1f, float x2f, float ps, double x,
                                           double c1, double s1, bool dq, double* acc) {
  const bool good = (x == x);
  const double xs = good ? x : 0.0;
  double xc = xs;   // dlegendre_R clamps x for the derivative only
  if (1.0 - fabs(xc) < 1.e-8) xc = (xc > 0) ? 1.0 - 1.e-8 : -(1.0 - 1.e-8);
  const double somx2d = 1.0 / (xc * xc - 1.0);
  const double somx2 = sqrt((1.0 - xs) * (1.0 + xs));
  // m = 0, l = 0
  const float4 a0 = ld4(A, 0), b0 = ld4(A, 1);
  const double pot0 = (double)fmaf(x2f, a0.z, a0.x);
  const double potr0 = (double)fmaf(ps, b0.z, b0.x);
  double den0 = 0.0;
  if (DENS) {
    const float4 nd = Nd[0];
    den0 = (double)fmaf(x2f, nd.z, x1f * nd.x);
  }
  double pot1 = 0, potr = 0, pott = 0, potp = 0, den1 = 0;
  // m = 0, l >= 1
  {
    double plm1 = 0.0, pl = 1.0;
#pragma unroll 2
    for (int l = 1; l <= L; ++l) {
      const int lm = lm_index(l, 0);
      const double pn = xs * g.lmc[kLmcA + lm] * pl - g.lmc[kLmcB + lm] * plm1;
      plm1 = pl;
      pl = pn;
      const float4 a = ld4(A, 2 * l), b = ld4(A, 2 * l + 1);
      const double ec = (double)fmaf(x2f, a.z, a.x);
      const double dc = (double)fmaf(ps, b.z, b.x);
      // dP/dx without its factor 1/(x^2 - 1), applied once to pott below
      const double dpl = xc * g.lmc[kLmcL + lm] * pl - g.lmc[kLmcS + lm] * plm1;
      pot1 += pl * ec;
      potr += pl * dc;
      pott += dpl * ec;
      if (DENS) {
        const float4 nd = Nd[l];
        den1 += (dq ? xs : pl) * (double)fmaf(x2f, nd.z, x1f * nd.x);
      }
    }
  }
  double pmm = 1.0, fact = 1.0;
  double cmm = 1.0, smm = 0.0, cmm1 = c1, smm1 = -s1;
  int p = L + 1;
#pragma unroll 1
  for (int m = 1; m <= L; ++m) {
    pmm *= -fact * somx2;
    fact += 2.0;
    const double cn = 2.0 * c1 * cmm - cmm1;   // m = 1 gives c1, s1 exactly
    const double sn = 2.0 * c1 * smm - smm1;
    cmm1 = cmm; smm1 = smm; cmm = cn; smm = sn;
    double Pc = 0, Ps = 0, Rc = 0, Rs = 0, Tc = 0, Ts = 0, Nc = 0, Ns = 0;
    auto term = [&](int q, int lm, double pl, double plm1) {
      const float4 a = ld4(A, 2 * q), b = ld4(A, 2 * q + 1);
      const double ec = (double)fmaf(x2f, a.z, a.x);
      const double dc = (double)fmaf(ps, b.z, b.x);
      const double es = (double)fmaf(x2f, a.w, a.y);
      const double ds = (double)fmaf(ps, b.w, b.y);
      const double dpl = xc * g.lmc[kLmcL + lm] * pl - g.lmc[kLmcS + lm] * plm1;
      Pc += pl * ec;
      Ps += pl * es;
      Rc += pl * dc;
      Rs += pl * ds;
      Tc += dpl * ec;
      Ts += dpl * es;
      if (DENS) {
        const float4 nd = Nd[q];
        Nc += pl * (double)fmaf(x2f, nd.z, x1f * nd.x);
        Ns += pl * (double)fmaf(x2f, nd.w, x1f * nd.y);
      }
    };
    double plm1 = 0.0, pl = pmm;
    term(p, lm_index(m, m), pl, plm1);
    ++p;
#pragma unroll 2
    for (int l = m + 1; l <= L; ++l, ++p) {
      const int lm = lm_index(l, m);
      const double pn = xs * g.lmc[kLmcA + lm] * pl - g.lmc[kLmcB + lm] * plm1;
      plm1 = pl;
      pl = pn;
      term(p, lm, pl, plm1);
    }
    pot1 += cmm * Pc + smm * Ps;
    potr += cmm * Rc + smm * Rs;
    pott += cmm * Tc + smm * Ts;
    potp += m * (cmm * Ps - smm * Pc);
    if (DENS) den1 += cmm * Nc + smm * Ns;
  }
  pott *= somx2d;
  if (!good) {   // P_l^m = 0 for l > 0 when cos(theta) is NaN (spheresl.py:698)
    pot1 = potr = pott = potp = den1 = 0.0;
  }
  acc[0] = pot0;
  acc[1] = potr0 + potr;
  acc[2] = den0;
  acc[3] = pot1;
  acc[4] = pott;
  acc[5] = potp;
  acc[6] = den1;
}

// The sums with BFE_COMPAT_SL_COS_L: the azimuthal factor of term (l, m > 0)
// is cos/sin(l phi), so it enters per term.  Same recurrences as sl_angular.
template <bool DENS, typename PA>
__device__ __forceinline__ void sl_angular_cosl(const SLGeom& g, PA A,
                                                const float4* __restrict__ Nd, int L, float x1f,
                                                float x2f, float ps, double x, double c1,
                                                double s1, bool dq, double* acc) {
  const bool good = (x == x);
  double xc = x;
  if (good && 1.0 - fabs(xc) < 1.e-8) xc = (xc > 0) ? 1.0 - 1.e-8 : -(1.0 - 1.e-8);
  const double somx2d = 1.0 / (xc * xc - 1.0);
  const double somx2 = good ? sqrt((1.0 - x) * (1.0 + x)) : 0.0;
  double pot0 = 0, potr0 = 0, den0 = 0, pot1 = 0, potr = 0, pott = 0, potp = 0, den1 = 0;
  double pmm = 1.0, fact = 1.0;
  double cmm = 1.0, smm = 0.0, cmm1 = c1, smm1 = -s1;
  int p = 0;
#pragma unroll 1
  for (int m = 0; m <= L; ++m) {
    if (m > 0) {
      pmm *= -fact * somx2;
      fact += 2.0;
      const double cn = 2.0 * c1 * cmm - cmm1, sn = 2.0 * c1 * smm - smm1;
      cmm1 = cmm; smm1 = smm; cmm = cn; smm = sn;
    }
    double plm1 = 0.0, pl = pmm;
    double cl = cmm, sl = smm, clm1 = cmm1, slm1 = smm1;
#pragma unroll 1
    for (int l = m; l <= L; ++l, ++p) {
      const int lm = lm_index(l, m);
      if (l > m) {
        const double pn = x * g.lmc[kLmcA + lm] * pl - g.lmc[kLmcB + lm] * plm1;
        plm1 = pl;
        pl = pn;
        const double cn = 2.0 * c1 * cl - clm1, sn = 2.0 * c1 * sl - slm1;
        clm1 = cl; slm1 = sl; cl = cn; sl = sn;
      }
      const float4 a = ld4(A, 2 * p), b = ld4(A, 2 * p + 1);
      const double ec = (double)fmaf(x2f, a.z, a.x);
      const double dc = (double)fmaf(ps, b.z, b.x);
      if (l == 0) {
        pot0 = ec;
        potr0 = dc;
        if (DENS) {
          const float4 nd = Nd[p];
          den0 = (double)fmaf(x2f, nd.z, x1f * nd.x);
        }
        continue;
      }
      if (!good) continue;  // P_l^m = 0 for l > 0 when cos(theta) is NaN (spheresl.py:698)
      const double dpl = somx2d * (xc * g.lmc[kLmcL + lm] * pl - g.lmc[kLmcS + lm] * plm1);
      if (m == 0) {
        pot1 += pl * ec;
        potr += pl * dc;
        pott += dpl * ec;
        if (DENS) {
          const float4 nd = Nd[p];
          den1 += (dq ? x : pl) * (double)fmaf(x2f, nd.z, x1f * nd.x);
        }
        continue;
      }
      const double es = (double)fmaf(x2f, a.w, a.y);
      const double ds = (double)fmaf(ps, b.w, b.y);
      const double pa = ec * cl + es * sl;
      pot1 += pl * pa;
      potr += pl * (dc * cl + ds * sl);
      pott += dpl * pa;
      potp += pl * m * (-ec * sl + es * cl);
      if (DENS) {
        const float4 nd = Nd[p];
        const double nc = (double)fmaf(x2f, nd.z, x1f * nd.x);
        const double ns = (double)fmaf(x2f, nd.w, x1f * nd.y);
        den1 += pl * (nc * cl + ns * sl);
      }
    }
  }
  acc[0] = pot0;
  acc[1] = potr0 + potr;
  acc[2] = den0;
  acc[3] = pot1;
  acc[4] = pott;
  acc[5] = potp;
  acc[6] = den1;
}

// pcell: the radial cell of r when the caller has located it already.
// STAGED: node i's force blocks are read from the LDS copy Abase at
// (i - ibase) * 2 np float4 (a block stages its nodes there); otherwise from
// the folded table itself.
template <bool DENS, bool STAGED = false>
__device__ __forceinline__ void sl_eval_folded(const SLGeom& g, const float* __restrict__ F,
                                               int L, double r, double costh, double c1,
                                               double s1, uint32_t flags, SLOut& o,
                                               const SLCell* pcell = nullptr,
                                               const float4* Abase = nullptr, int ibase = 0) {
  const SLCell cell = pcell ? *pcell : sl_locate(r, g.cmap, g.scale, g.xi, g.numr, g.idxi);
  const int i = cell.i;
  const double x1 = cell.x1, x2 = cell.x2;
  const double P0 = x1 * g.p0[i] + x2 * g.p0[i + 1];
  const double D0 = DENS ? x1 * g.d0[i] + x2 * g.d0[i + 1] : 0.0;
  const double fac = (g.cmap == 1) ? (1.0 - cell.x) * (1.0 - cell.x) * g.dfac
                                   : d_xi_to_r(cell.x, g.cmap, g.scale) * g.idxi;
  const float ps = (float)((i == 0) ? ((flags & BFE_COMPAT_SL_EDGE) ? x2 : x2 - 1.0) : x2);
  const int np = sl_npairs(L);
  const float4* __restrict__ Nd =
      reinterpret_cast<const float4*>(F + (size_t)g.numr * np * 8) + (size_t)i * np;
  const bool dq = (flags & BFE_COMPAT_SL_DENS) != 0;
  double acc[7];
  if (STAGED) {
    const LdsF4 A = (LdsF4)Abase + (size_t)(i - ibase) * 2 * np;
    if (flags & BFE_COMPAT_SL_COS_L)
      sl_angular_cosl<DENS>(g, A, Nd, L, (float)x1, (float)x2, ps, costh, c1, s1, dq, acc);
    else
      sl_angular<DENS>(g, A, Nd, L, (float)x1, (float)x2, ps, costh, c1, s1, dq, acc);
  } else {
    const float4* __restrict__ A = reinterpret_cast<const float4*>(F) + (size_t)i * 2 * np;
    if (flags & BFE_COMPAT_SL_COS_L)
      sl_angular_cosl<DENS>(g, A, Nd, L, (float)x1, (float)x2, ps, costh, c1, s1, dq, acc);
    else
      sl_angular<DENS>(g, A, Nd, L, (float)x1, (float)x2, ps, costh, c1, s1, dq, acc);
  }
  o.pot0 = acc[0] * P0;
  o.potr = acc[1] * fac;
  o.den0 = acc[2] * D0;
  o.pot1 = acc[3] * P0;
  o.pott = acc[4] * P0;
  o.potp = acc[5] * P0;
  o.den1 = acc[6] * D0;
}

// EOF evaluation at planar radius R, height z, angle phi from the folded
// table (8-float block per m).  Bilinear weights as eof.py:443-451; sums as
// force_eval (eof.py:842-870) and accumulated_eval_particles (:1092-1138):
// fp = sum m (C sin m phi potC - S cos m phi potS) = -dPhi/dphi.  MU: unroll
// count of the m loop (2 measured 7 % faster in the RK4 kernel, slower in
// the leapfrog one).  Cells of the last grid column or row take their m = 0
// corner values from the float64 edge blocks that follow the float32 table:
// past the outer edges the weights extrapolate linearly (eof.py:414-415) and
// grow without bound, and with them the rounding of the table values; the
// m = 0 terms carry the largest values.  The m > 0 terms stay float32.
template <bool DENS, int MU = 1>
__device__ __forceinline__ void eof_eval_folded(const EOFGeom& g, const float* __restrict__ F,
                                                double R, double z, double c1, double s1,
                                                EOFOut& o) {
  const EOFCell cell = eof_locate(R, z, g.cmap, g.ascale, g.ihscale, g.xmin, g.idx, g.ymin, g.idy,
                                  g.numx, g.numy);
  const int ny1 = g.numy + 1;
  const int M = g.mmax;
  const int ns4 = 2 * (M + 1);
  const size_t n00 = (size_t)cell.ix * ny1 + cell.iy;
  const float4* __restrict__ A = reinterpret_cast<const float4*>(F) + n00 * ns4;  // (ix, iy)
  const float4* __restrict__ B = A + ns4;                                          // (ix, iy+1)
  const float4* __restrict__ C = A + (size_t)ny1 * ns4;                            // (ix+1, iy)
  const float4* __restrict__ D = C + ns4;                                          // (ix+1, iy+1)
  const double w00 = cell.c00, w10 = cell.c10, w01 = cell.c01, w11 = cell.c11;
  // m = 0: [pot, rforce, zforce, dens] of the four corners
  double4 a0, b0, c0, e0;
  if (cell.ix == g.numx - 1 || cell.iy == g.numy - 1) {
    const double4* __restrict__ E =
        reinterpret_cast<const double4*>(F + eof_fold32_floats(g.numx, g.numy, M));
    const int x0 = cell.ix, y0 = cell.iy;
    a0 = E[eof_edge_index(x0, y0, g.numx, g.numy)];
    b0 = E[eof_edge_index(x0, y0 + 1, g.numx, g.numy)];
    c0 = E[eof_edge_index(x0 + 1, y0, g.numx, g.numy)];
    e0 = E[eof_edge_index(x0 + 1, y0 + 1, g.numx, g.numy)];
  } else {
    const float4 a = A[0], b = B[0], c = C[0], e = D[0];
    a0 = make_double4(a.x, a.y, a.z, a.w);
    b0 = make_double4(b.x, b.y, b.z, b.w);
    c0 = make_double4(c.x, c.y, c.z, c.w);
    e0 = make_double4(e.x, e.y, e.z, e.w);
  }
  const double p0 = w00 * a0.x + w10 * c0.x + w01 * b0.x + w11 * e0.x;
  const double fr0 = w00 * a0.y + w10 * c0.y + w01 * b0.y + w11 * e0.y;
  const double fz0 = w00 * a0.z + w10 * c0.z + w01 * b0.z + w11 * e0.z;
  const double d0 = DENS ? w00 * a0.w + w10 * c0.w + w01 * b0.w + w11 * e0.w : 0.0;
  double fr = fr0, fz = fz0, p = 0, fp = 0, d = 0;
  double cm = 1.0, sm = 0.0, cm1 = c1, sm1 = -s1;
  // m > 0: the corners of the float32 table interpolated in float32 (one
  // conversion per value instead of four), the sums over m in float64
  const float v00 = (float)w00, v10 = (float)w10, v01 = (float)w01, v11 = (float)w11;
  auto bil = [&](float a, float c, float b, float e) {
    return (double)fmaf(v11, e, fmaf(v01, b, fmaf(v10, c, v00 * a)));
  };
#pragma unroll MU
  for (int m = 1; m <= M; ++m) {
    const double cn = 2.0 * c1 * cm - cm1;   // m = 1 gives c1, s1 exactly
    const double sn = 2.0 * c1 * sm - sm1;
    cm1 = cm; sm1 = sm; cm = cn; sm = sn;
    const float4 a = A[2 * m], b = B[2 * m], c = C[2 * m], e = D[2 * m];
    const double pc = bil(a.x, c.x, b.x, e.x);
    const double rc = bil(a.y, c.y, b.y, e.y);
    const double zc = bil(a.z, c.z, b.z, e.z);
    const float4 a2 = A[2 * m + 1], b2 = B[2 * m + 1], c2 = C[2 * m + 1], e2 = D[2 * m + 1];
    const double ps = bil(a2.x, c2.x, b2.x, e2.x);
    const double rs = bil(a2.y, c2.y, b2.y, e2.y);
    const double zs = bil(a2.z, c2.z, b2.z, e2.z);
    p += cm * pc + sm * ps;
    fr += cm * rc + sm * rs;
    fz += cm * zc + sm * zs;
    fp += m * (sm * pc - cm * ps);
    if (DENS) {
      const double dc = bil(a.w, c.w, b.w, e.w);
      const double ds = bil(a2.w, c2.w, b2.w, e2.w);
      d += cm * dc + sm * ds;
    }
  }
  o.p0 = p0;
  o.p = p;
  o.fr = fr;
  o.fz = fz;
  o.fp = fp;
  o.d0 = d0;
  o.d = d;
  o.fr0 = fr0;
  o.fz0 = fz0;
}

// The SL radial cell field_eval_point locates for the point (x, y, z).
__device__ __forceinline__ SLCell field_sl_cell(const SLGeom& g, double x, double y, double z) {
  const double r2 = sqrt(x * x + y * y) + 1.e-15;
  const double r3 = sqrt(r2 * r2 + z * z) + 1.e-15;
  return sl_locate(r3, g.cmap, g.scale, g.xi, g.numr, g.idxi);
}

// Frozen disc + halo field at a Cartesian point, Fields.return_forces_cart
// (potential.py:445-497).  comp receives the 8 reference outputs.
struct FieldTables {
  SLGeom slg;
  EOFGeom eg;
  const float* slF;
  const float* eofF;
  int lmax;          // folded SL layout size (table lmax)
  uint32_t flags;
  double ximin;      // min(xi_halo) for the azimuthal-force guard (:483-485)
  double crot, srot; // cos/sin of rotpos
};

// pcell/slA/slbase: the point's SL cell (field_sl_cell) and SL node blocks
// staged by the caller in LDS (STAGED, see sl_eval_folded).  MU: unroll count
// of the EOF m loop (eof_eval_folded).
template <bool STAGED = false, int MU = 1>
__device__ __forceinline__ void field_eval_point(const FieldTables& T, double x, double y,
                                                 double z, double comp[8],
                                                 const SLCell* pcell = nullptr,
                                                 const float4* slA = nullptr, int slbase = 0) {
  const double r2 = sqrt(x * x + y * y) + 1.e-15;
  const double r3 = sqrt(r2 * r2 + z * z) + 1.e-15;
  // reciprocals once; the force combination below multiplies by them
  const double ir2 = 1.0 / r2, ir3 = 1.0 / r3;
  const double costh = z * ir3;
  double c0, s0;
  planar_angle(x, y, c0, s0);
  const double c1 = c0 * T.crot - s0 * T.srot;   // cos(phi + rotpos)
  const double s1 = s0 * T.crot + c0 * T.srot;   // sin(phi + rotpos)
  EOFOut d;
  eof_eval_folded<false, MU>(T.eg, T.eofF, r2, z, c1, s1, d);
  SLOut h;
  sl_eval_folded<false, STAGED>(T.slg, T.slF, T.lmax, r3, costh, c1, s1, T.flags, h, pcell, slA,
                                slbase);
  double dfp = d.fp, hfp = h.potp;
  if (r3 < T.ximin) {
    hfp = 0.0;
    dfp = 0.0;
  }
  const double dfr = d.fr, dfz = d.fz, dp = d.p + d.p0;
  const double hfr = h.potr, hft = h.pott, hp = h.pot1 + h.pot0;
  const double r2sq = r2 * r2, ir2sq = ir2 * ir2, ir3cu = ir3 * ir3 * ir3;
  comp[0] = dfr * (x * ir2) - dfp * (y * ir2sq);
  comp[1] = -1.0 * (hfr * (x * ir3) - hft * (x * z * ir3cu)) + hfp * (y * ir2sq);
  comp[2] = dfr * (y * ir2) + dfp * (x * ir2sq);
  comp[3] = -1.0 * (hfr * (y * ir3) - hft * (y * z * ir3cu)) - hfp * (x * ir2sq);
  comp[4] = dfz;
  comp[5] = -1.0 * (hfr * (z * ir3) + hft * (r2sq * ir3cu));
  comp[6] = dp;
  comp[7] = hp;
}

// LDS a block may fill with SL node force blocks (2 np float4 per node): 40
// nodes of lmax 6 (35 KB, four blocks per CU); measured on the workload,
// 16, 32, 40 and 48 nodes gave 243, 222, 219 and 230 us.
constexpr int kStageFloat4 = 40 * 56;

// Block-wide minimum and maximum of v (all 256 threads call it).
__device__ __forceinline__ void block_minmax(int v_min, int v_max, int* smin, int* smax,
                                             int& out_min, int& out_max) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    v_min = min(v_min, __shfl_xor(v_min, o, 64));
    v_max = max(v_max, __shfl_xor(v_max, o, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    smin[threadIdx.x >> 6] = v_min;
    smax[threadIdx.x >> 6] = v_max;
  }
  __syncthreads();
  out_min = min(min(smin[0], smin[1]), min(smin[2], smin[3]));
  out_max = max(max(smax[0], smax[1]), max(smax[2], smax[3]));
}

// Copies the force blocks of the SL nodes used by the block's points to
// stage when they fit (all 256 threads call it; live: the thread has a
// point, node: its SL node).  Returns whether they were staged; imin is then
// the first staged node.  Ends with a barrier.
__device__ __forceinline__ bool stage_sl_nodes(const FieldTables& T, bool live, int node,
                                               float4* stage, int* smin, int* smax, int& imin) {
  const int np2 = sl_npairs(T.lmax) * 2;   // float4 per SL node
  int imax;
  block_minmax(live ? node : 0x7fffffff, live ? node : -1, smin, smax, imin, imax);
  bool staged = false;
  if (imax >= imin && (imax - imin + 1) * np2 <= kStageFloat4) {
    const float4* src = reinterpret_cast<const float4*>(T.slF) + (size_t)imin * np2;
    const int cnt = (imax - imin + 1) * np2;
    for (int k = threadIdx.x; k < cnt; k += 256) stage[k] = src[k];
    staged = true;
  }
  __syncthreads();
  return staged;
}

}  // namespace bfe
