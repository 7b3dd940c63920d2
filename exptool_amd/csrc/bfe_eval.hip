// Evaluation: coefficient folding and the batch / field evaluators.
#include <math.h>

#include <algorithm>
#include <vector>

#include "bfe_field.hpp"

namespace bfe {

struct SLFoldArgs {
  double lmul[kMaxL + 1];  // per-l multiplier (halofac, 0 for skipped l)
  int lmax_use, nmax_use, rows, stride, groups;
};

// One thread per (node i, row).  E, G (at i-1, i, i+1) and Dn are summed over
// n in float64 from the float64 master tables, then D1/D2 are formed and the
// results rounded once to float32.
__global__ void sl_fold_kernel(const double* __restrict__ ef, const double* __restrict__ ev,
                               const double* __restrict__ p0, const double* __restrict__ coef,
                               int nmax, int numr, SLFoldArgs a, float* __restrict__ out) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)numr * a.rows) return;
  const int i = (int)(idx / a.rows);
  const int row = (int)(idx - (int64_t)i * a.rows);
  int l = 0;
  while ((l + 1) * (l + 1) <= row) ++l;
  int m = (row - l * l + 1) / 2;
  // N_lm (spheresl.py:858-861) folded into the table
  double ratio = 1.0;
  for (int k = 1; k <= m; ++k) ratio /= (double)(l - k + 1) * (double)(l + k);
  double N = sqrt((0.5 * l + 0.25) / kPi * ratio);
  if (m > 0) N *= 1.4142135623730951;
  const double mul = (l <= a.lmax_use) ? a.lmul[l] * N : 0.0;
  double E = 0, Dn = 0, Gm = 0, G0 = 0, Gp = 0;
  const bool inner = (i >= 1 && i <= numr - 2);
  for (int n = 0; n < a.nmax_use; ++n) {
    const double c = coef[(size_t)row * nmax + n] * mul;
    const size_t base = ((size_t)l * nmax + n) * numr;
    const double sev = sqrt(ev[l * nmax + n]);
    const double e = ef[base + i];
    E += c * e / sev;
    Dn += c * e * sev;
    G0 += c * e * p0[i] / sev;
    if (inner) {
      Gm += c * ef[base + i - 1] * p0[i - 1] / sev;
      Gp += c * ef[base + i + 1] * p0[i + 1] / sev;
    }
  }
  float* o = out + (size_t)i * a.groups * a.stride;
  o[row] = (float)E;
  o[a.stride + row] = inner ? (float)(0.5 * (Gp - Gm)) : 0.f;
  o[2 * a.stride + row] = inner ? (float)(Gm - 2.0 * G0 + Gp) : 0.f;
  if (a.groups > 3) o[3 * a.stride + row] = (float)Dn;
}

struct EOFFoldArgs {
  double mmul[kMaxM + 1];  // per-m multiplier (range / no_odd mask)
  int nuse, groups, stride;
};

// One thread per (node, j) with j = cos m (0..M) or sin m (M+1..2M); sums the
// n < nuse terms of every group in float64.
__global__ void eof_fold_kernel(const double* __restrict__ T0, const double* __restrict__ T1,
                                const double* __restrict__ T2, const double* __restrict__ T3,
                                const double* __restrict__ ccos, const double* __restrict__ csin,
                                int mmax, int norder, int64_t nodes, EOFFoldArgs a,
                                float* __restrict__ out) {
  const int q = 2 * mmax + 1;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nodes * q) return;
  const int64_t node = idx / q;
  const int j = (int)(idx - node * q);
  const int m = (j <= mmax) ? j : j - mmax;
  const double* cf = (j <= mmax) ? ccos + (size_t)m * norder : csin + (size_t)m * norder;
  const double mul = a.mmul[m];
  const int K = q * norder;
  const double* tabs[4] = {T0, T1, T2, T3};
  float* o = out + node * a.stride;
  for (int g = 0; g < a.groups; ++g) {
    const double* t = tabs[g] + node * K + (size_t)j * norder;
    double s = 0.0;
    for (int n = 0; n < a.nuse; ++n) s += cf[n] * t[n];
    o[g * q + j] = (float)(s * mul);
  }
}

static int round4(int v) { return (v + 3) & ~3; }

static int sl_fold(Context* c, int slot, const double* coef, double halofac, int lmax_use,
                   int nmax_use, int L1, int L2, bool no_odd, bool dens, hipStream_t s) {
  SLFolded& f = c->slf[slot];
  const int rows = (c->lmax + 1) * (c->lmax + 1);
  const int stride = round4(rows);
  const int groups = dens ? 4 : 3;
  const size_t need = (size_t)c->numr * groups * stride;
  if (!f.d || f.rows != rows || f.groups < groups) {
    if (f.d) {
      BFE_HIP(c, hipStreamSynchronize(s));
      (void)hipFree(f.d);
      f.d = nullptr;
    }
    BFE_HIP(c, hipMalloc(&f.d, need * sizeof(float)));
    BFE_HIP(c, hipMemsetAsync(f.d, 0, need * sizeof(float), s));
  }
  f.rows = rows;
  f.stride = stride;
  f.groups = groups;
  f.lmax_use = lmax_use;
  SLFoldArgs a;
  for (int l = 0; l <= kMaxL; ++l) {
    bool keep = l <= lmax_use;
    if (l >= 1 && (l > L2 || l < L1)) keep = false;
    if (l >= 1 && no_odd && (l & 1)) keep = false;
    a.lmul[l] = keep ? halofac : 0.0;
  }
  a.lmax_use = lmax_use;
  a.nmax_use = nmax_use;
  a.rows = rows;
  a.stride = stride;
  a.groups = groups;
  const int64_t total = (int64_t)c->numr * rows;
  hipLaunchKernelGGL(sl_fold_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                     c->sl_ef64, c->sl_ev64, c->sl_p0, coef, c->nmax, c->numr, a, f.d);
  BFE_HIP(c, hipGetLastError());
  f.valid = true;
  return 0;
}

static int eof_fold(Context* c, int slot, const double* ccos, const double* csin, int mmax_use,
                    int nuse, int m1, int m2, bool no_odd, bool compat_mask, bool dens,
                    hipStream_t s) {
  EOFFolded& f = c->eoff[slot];
  const int q = 2 * c->mmax + 1;
  const int groups = dens ? 4 : 3;
  const int stride = round4(groups * q);
  const int64_t nodes = (int64_t)(c->numx + 1) * (c->numy + 1);
  if (!f.d || f.q != q || f.stride != stride) {
    if (f.d) {
      BFE_HIP(c, hipStreamSynchronize(s));
      (void)hipFree(f.d);
      f.d = nullptr;
    }
    BFE_HIP(c, hipMalloc(&f.d, (size_t)nodes * stride * sizeof(float)));
    BFE_HIP(c, hipMemsetAsync(f.d, 0, (size_t)nodes * stride * sizeof(float), s));
  }
  f.q = q;
  f.groups = groups;
  f.stride = stride;
  EOFFoldArgs a;
  for (int m = 0; m <= kMaxM; ++m) {
    double mul = 1.0;
    if (m > mmax_use || m > m2 || m < m1) mul = 0.0;
    if (no_odd && m >= 1 && (m & 1)) {
      // eof.force_eval mask = |cos(pi m / 2)| (eof.py:833): ~6e-17 for odd m
      mul *= compat_mask ? fabs(cos((kPi / 2.0) * m)) : 0.0;
    }
    a.mmul[m] = mul;
  }
  a.nuse = nuse;
  a.groups = groups;
  a.stride = stride;
  const int64_t total = nodes * q;
  hipLaunchKernelGGL(eof_fold_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                     c->eof64[0], c->eof64[1], c->eof64[2], dens ? c->eof64[3] : c->eof64[0],
                     ccos, csin, c->mmax, c->norder, nodes, a, f.d);
  BFE_HIP(c, hipGetLastError());
  f.valid = true;
  return 0;
}

// ---------------------------------------------------------------------------
// batch evaluators
// ---------------------------------------------------------------------------
template <bool DENS>
__global__ __launch_bounds__(256) void sl_eval_kernel(
    const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
    int64_t n, SLGeom g, const float* __restrict__ F, int stride, int groups, int lmax_use,
    uint32_t flags, double densfac, float* den0, float* den1, float* pot0, float* pot1,
    float* potr, float* pott, float* potp) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double x = px[p], y = py[p], z = pz[p];
  const double r = sqrt(x * x + y * y + z * z);  // no floor (spheresl.py:1257)
  const double costh = z / r;                     // NaN at r = 0, as the reference
  double c1, s1;
  planar_angle(x, y, c1, s1);
  SLOut o;
  sl_eval_folded<DENS>(g, F, stride, groups, lmax_use, r, costh, c1, s1, flags, o);
  if (den0) den0[p] = (float)(o.den0 * densfac);
  if (den1) den1[p] = (float)(o.den1 * densfac);
  if (pot0) pot0[p] = (float)o.pot0;
  if (pot1) pot1[p] = (float)o.pot1;
  if (potr) potr[p] = (float)o.potr;
  if (pott) pott[p] = (float)o.pott;
  if (potp) potp[p] = (float)o.potp;
}

template <bool DENS>
__global__ __launch_bounds__(256) void eof_eval_kernel(
    const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
    int64_t n, EOFGeom g, const float* __restrict__ F, int stride, float* p0, float* p, float* fr,
    float* fp, float* fz, float* d0, float* d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = px[i], y = py[i], z = pz[i];
  const double R = sqrt(x * x + y * y + 1.e-10);  // eof.py:1070
  double c1, s1;
  planar_angle(x, y, c1, s1);
  EOFOut o;
  eof_eval_folded<DENS>(g, F, stride, R, z, c1, s1, o);
  if (p0) p0[i] = (float)o.p0;
  if (p) p[i] = (float)o.p;
  if (fr) fr[i] = (float)o.fr;
  if (fp) fp[i] = (float)o.fp;
  if (fz) fz[i] = (float)o.fz;
  if (DENS) {
    if (d0) d0[i] = (float)o.d0;
    if (d) d[i] = (float)o.d;
  }
}

__global__ __launch_bounds__(256) void field_eval_kernel(
    const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
    int64_t n, FieldTables T, float* fx, float* fy, float* fz, float* pot, float* comp) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double cv[8];
  field_eval_point(T, px[i], py[i], pz[i], cv);
  if (fx) fx[i] = (float)(cv[0] + cv[1]);
  if (fy) fy[i] = (float)(cv[2] + cv[3]);
  if (fz) fz[i] = (float)(cv[4] + cv[5]);
  if (pot) pot[i] = (float)(cv[6] + cv[7]);
  if (comp) {
#pragma unroll
    for (int k = 0; k < 8; ++k) comp[(size_t)k * n + i] = (float)cv[k];
  }
}

int field_tables(Context* c, double rotpos, FieldTables& T) {
  if (!c->slf[1].valid || !c->eoff[1].valid)
    return set_error(c, BFE_ENOTABLE, "field not set (call bfe_field_set)");
  T.slg = sl_geom(c);
  T.eg = eof_geom(c);
  T.slF = c->slf[1].d;
  T.eofF = c->eoff[1].d;
  T.sl_stride = c->slf[1].stride;
  T.sl_groups = c->slf[1].groups;
  T.lmax_use = c->slf[1].lmax_use;
  T.eof_stride = c->eoff[1].stride;
  T.flags = (uint32_t)c->field_flags;
  T.ximin = *std::min_element(c->h_xi.begin(), c->h_xi.end());
  T.crot = cos(rotpos);
  T.srot = sin(rotpos);
  return 0;
}

}  // namespace bfe

using namespace bfe;

static unsigned nblocks256(int64_t n) { return (unsigned)((n + 255) / 256); }

extern "C" int bfe_sl_eval(bfe_ctx* h, const float* x, const float* y, const float* z, int64_t n,
                           const double* coef, int L1, int L2, uint32_t flags, float* den0,
                           float* den1, float* pot0, float* pot1, float* potr, float* pott,
                           float* potp, void* stream) {
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (!c->sl_ok) return set_error(c, BFE_ENOTABLE, "sl_eval: SL tables not set");
  if (n < 0 || !coef) return set_error(c, BFE_EINVAL, "sl_eval: bad n or NULL coef");
  if (n > 0 && (!x || !y || !z)) return set_error(c, BFE_EINVAL, "sl_eval: NULL input");
  BFE_HIP(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream;
  const bool dens = den0 || den1;
  int rc = sl_fold(c, 0, coef, 1.0, c->lmax, c->nmax, L1, L2, (flags & BFE_NO_ODD) != 0, dens, s);
  if (rc) return rc;
  if (n == 0) return 0;
  const SLFolded& f = c->slf[0];
  const double densfac = (flags & BFE_COMPAT_SL_DENS) ? 0.25 * kPi : 0.25 / kPi;
  const uint32_t ef = flags & (BFE_COMPAT_SL_DENS | BFE_COMPAT_SL_EDGE);
  if (dens)
    hipLaunchKernelGGL(sl_eval_kernel<true>, dim3(nblocks256(n)), dim3(256), 0, s, x, y, z, n,
                       sl_geom(c), f.d, f.stride, f.groups, c->lmax, ef, densfac, den0, den1,
                       pot0, pot1, potr, pott, potp);
  else
    hipLaunchKernelGGL(sl_eval_kernel<false>, dim3(nblocks256(n)), dim3(256), 0, s, x, y, z, n,
                       sl_geom(c), f.d, f.stride, f.groups, c->lmax, ef, densfac, den0, den1,
                       pot0, pot1, potr, pott, potp);
  BFE_HIP(c, hipGetLastError());
  return 0;
}

extern "C" int bfe_eof_eval(bfe_ctx* h, const float* x, const float* y, const float* z, int64_t n,
                            const double* cos_coef, const double* sin_coef, int m1, int m2,
                            uint32_t flags, float* p0, float* p, float* fr, float* fp, float* fz,
                            float* d0, float* d, void* stream) {
  (void)flags;
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (!c->eof_ok) return set_error(c, BFE_ENOTABLE, "eof_eval: EOF tables not set");
  if (n < 0 || !cos_coef || !sin_coef) return set_error(c, BFE_EINVAL, "eof_eval: bad n or NULL coef");
  if (n > 0 && (!x || !y || !z)) return set_error(c, BFE_EINVAL, "eof_eval: NULL input");
  const bool dens = d0 || d;
  if (dens && !c->eof_dens) return set_error(c, BFE_ENOTABLE, "eof_eval: no density tables");
  BFE_HIP(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream;
  int rc = eof_fold(c, 0, cos_coef, sin_coef, c->mmax, c->norder, m1, m2, false, false, dens, s);
  if (rc) return rc;
  if (n == 0) return 0;
  const EOFFolded& f = c->eoff[0];
  if (dens)
    hipLaunchKernelGGL(eof_eval_kernel<true>, dim3(nblocks256(n)), dim3(256), 0, s, x, y, z, n,
                       eof_geom(c), f.d, f.stride, p0, p, fr, fp, fz, d0, d);
  else
    hipLaunchKernelGGL(eof_eval_kernel<false>, dim3(nblocks256(n)), dim3(256), 0, s, x, y, z, n,
                       eof_geom(c), f.d, f.stride, p0, p, fr, fp, fz, d0, d);
  BFE_HIP(c, hipGetLastError());
  return 0;
}

extern "C" int bfe_field_set(bfe_ctx* h, const double* eof_cos, const double* eof_sin,
                             const double* sl_coef, double halofac, int disk_m, int disk_n,
                             int halo_l, int halo_n, uint32_t flags, void* stream) {
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (!c->sl_ok || !c->eof_ok) return set_error(c, BFE_ENOTABLE, "field_set: tables not set");
  if (!eof_cos || !eof_sin || !sl_coef) return set_error(c, BFE_EINVAL, "field_set: NULL coef");
  const int mm = disk_m < 0 ? c->mmax : std::min(disk_m, c->mmax);
  const int nn = disk_n < 0 ? c->norder : std::min(disk_n, c->norder);
  const int ll = halo_l < 0 ? c->lmax : std::min(halo_l, c->lmax);
  const int hn = halo_n < 0 ? c->nmax : std::min(halo_n, c->nmax);
  BFE_HIP(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream;
  const bool no_odd = (flags & BFE_NO_ODD) != 0;
  int rc = eof_fold(c, 1, eof_cos, eof_sin, mm, nn, 0, 1 << 30, no_odd,
                    (flags & BFE_COMPAT_EOF_MASK) != 0, false, s);
  if (rc) return rc;
  rc = sl_fold(c, 1, sl_coef, halofac, ll, hn, -(1 << 30), 1 << 30, no_odd, false, s);
  if (rc) return rc;
  c->field_halofac = halofac;
  c->field_flags = (int)(flags & (BFE_COMPAT_SL_COS_L | BFE_COMPAT_SL_EDGE));
  return 0;
}

extern "C" int bfe_field_eval_cart(bfe_ctx* h, const float* x, const float* y, const float* z,
                                   int64_t n, double rotpos, float* fx, float* fy, float* fz,
                                   float* pot, float* comp, void* stream) {
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (n < 0) return set_error(c, BFE_EINVAL, "field_eval: n < 0");
  if (n > 0 && (!x || !y || !z)) return set_error(c, BFE_EINVAL, "field_eval: NULL input");
  FieldTables T;
  int rc = field_tables(c, rotpos, T);
  if (rc) return rc;
  if (n == 0) return 0;
  BFE_HIP(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(field_eval_kernel, dim3(nblocks256(n)), dim3(256), 0, s, x, y, z, n, T, fx, fy,
                     fz, pot, comp);
  BFE_HIP(c, hipGetLastError());
  return 0;
}
