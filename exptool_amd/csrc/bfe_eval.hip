// Evaluation: coefficient folding and the batch / field evaluators.
#include <math.h>

#include <algorithm>
#include <vector>

#include "bfe_field.hpp"

namespace bfe {

struct SLFoldArgs {
  double lmul[kMaxL + 1];  // per-l multiplier (halofac, 0 for skipped l)
  int lmax, nmax_use;
};

// Four lanes per (radial node i, (l, m) pair p in m-major order), lane q
// summing n = q, q + 4, ...: E and Dn at i and i+1 and G at c-1, c, c+1
// (c = max(i, 1)) are summed over n in float64 from the float64 scaled
// tables S = [ef/sqrt(ev), ef sqrt(ev), ef p0/sqrt(ev)] ([l][i][n] each),
// the four partial sums are combined across the lanes, D1/D2 formed, and the
// blocks rounded once to float32 (layout in bfe_field.hpp; the last node,
// never a cell start, is zero).  grid.y runs over coefficient sets: set q
// reads coef + q * coef_stride and writes out + q * out_stride.
constexpr int kFoldLanes = 4;
__global__ __launch_bounds__(256) void sl_fold_kernel(const double* __restrict__ S,
                                                      const double* __restrict__ coef, int nmax,
                                                      int numr, SLFoldArgs a,
                                                      float* __restrict__ out, int64_t coef_stride,
                                                      int64_t out_stride) {
  coef += (size_t)blockIdx.y * coef_stride;
  out += (size_t)blockIdx.y * out_stride;
  const int np = sl_npairs(a.lmax);
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int q = (int)(gid & (kFoldLanes - 1));
  const int64_t idx = gid / kFoldLanes;
  // groups of four lanes are whole within a wave, so a group leaves together
  if (idx >= (int64_t)numr * np) return;
  const int i = (int)(idx / np);
  const int p = (int)(idx - (int64_t)i * np);
  float* o = out + ((size_t)i * np + p) * 8;
  float* od = out + (size_t)numr * np * 8 + ((size_t)i * np + p) * 4;
  if (i == numr - 1) {
    if (q == 0)
      for (int k = 0; k < 8; ++k) o[k] = 0.f;
    if (q == 1)
      for (int k = 0; k < 4; ++k) od[k] = 0.f;
    return;
  }
  int m = 0, rem = p;
  while (rem > a.lmax - m) { rem -= a.lmax - m + 1; ++m; }
  const int l = m + rem;
  // N_lm (spheresl.py:858-861) folded into the table
  double ratio = 1.0;
  for (int k = 1; k <= m; ++k) ratio /= (double)(l - k + 1) * (double)(l + k);
  double N = sqrt((0.5 * l + 0.25) / kPi * ratio);
  if (m > 0) N *= 1.4142135623730951;
  const double mul = a.lmul[l] * N;
  const double* cc = coef + (size_t)(l * l + (m > 0 ? 2 * m - 1 : 0)) * nmax;
  const double* cs = coef + (size_t)(l * l + 2 * m) * nmax;
  const bool two = m > 0;
  const int ctr = i == 0 ? 1 : i;
  const size_t plane = (size_t)(a.lmax + 1) * numr * nmax;
  const double* E = S + ((size_t)l * numr + i) * nmax;   // rows i, i+1
  const double* D = E + plane;
  const double* G = S + 2 * plane + ((size_t)l * numr + ctr) * nmax;  // rows c-1, c, c+1
  // v[cs]: E_i, E_i+1, Dn_i, Dn_i+1, G_c-1, G_c, G_c+1
  double v[2][7] = {{0, 0, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0, 0}};
#pragma unroll 2
  for (int n = q; n < a.nmax_use; n += kFoldLanes) {
    const double c0 = cc[n] * mul;
    const double c1 = two ? cs[n] * mul : 0.0;
    const double t[7] = {E[n], E[n + nmax], D[n], D[n + nmax], G[n - nmax], G[n], G[n + nmax]};
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      v[0][k] += c0 * t[k];
      v[1][k] += c1 * t[k];
    }
  }
  // fixed-order combination: (q0 + q1) + (q2 + q3) on every lane of the group
#pragma unroll
  for (int k = 0; k < 7; ++k) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      v[h][k] += __shfl_xor(v[h][k], 1, kFoldLanes);
      v[h][k] += __shfl_xor(v[h][k], 2, kFoldLanes);
    }
  }
  if (q == 0) {
    // node i and the step to node i+1, so that the evaluation interpolates
    // with one FMA: E_i + x2 (E_i+1 - E_i)
    o[0] = (float)v[0][0];
    o[1] = (float)v[1][0];
    o[2] = (float)(v[0][1] - v[0][0]);
    o[3] = (float)(v[1][1] - v[1][0]);
  } else if (q == 1) {
    o[4] = (float)(0.5 * (v[0][6] - v[0][4]));
    o[5] = (float)(0.5 * (v[1][6] - v[1][4]));
    o[6] = (float)(v[0][4] - 2.0 * v[0][5] + v[0][6]);
    o[7] = (float)(v[1][4] - 2.0 * v[1][5] + v[1][6]);
  } else if (q == 2) {
    od[0] = (float)v[0][2];
    od[1] = (float)v[1][2];
    od[2] = (float)v[0][3];
    od[3] = (float)v[1][3];
  }
}

struct EOFFoldArgs {
  double mmul[kMaxM + 1];  // per-m multiplier (range / no_odd mask)
  int nuse, dens;
  int numx, numy;          // grid, for the float64 edge blocks
};

// The float64 edge blocks of the folded field (after the float32 blocks of
// each set): the m = 0 cos sums [pot, rforce, zforce, dens] of edge node e
// (eof_edge_index) from its float64 fold inputs E[f][n][edge], as the float32
// blocks sum them for m = 0.
__device__ __forceinline__ void eof_fold_edge(const double* __restrict__ E, int64_t nedge,
                                              int64_t e, const double* __restrict__ cf, int mmax,
                                              int norder, const EOFFoldArgs& a,
                                              float* __restrict__ out_set) {
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (int n = 0; n < a.nuse; ++n) {
    const double c = cf[n];
#pragma unroll
    for (int f = 0; f < 4; ++f)
      if (f < 3 || a.dens) s[f] += c * E[((size_t)f * norder + n) * nedge + e];
  }
  const double mul = a.mmul[0];
  double4* o = reinterpret_cast<double4*>(out_set + eof_fold32_floats(a.numx, a.numy, mmax)) + e;
  *o = make_double4(s[0] * mul, s[1] * mul, s[2] * mul, s[3] * mul);
}

// One thread per (row group q, node): q = m for cos m = 0..mmax, mmax + m
// for sin m = 1..mmax; nodes fastest so the function-major float32 tables
// are read coalesced (sums in float64), and q uniform per block so the coefficients are
// scalar loads.  Writes its half [pot, rf, zf, dens] of the 8-float block
// [potC, rfC, zfC, dC, potS, rfS, zfS, dS]; cos m = 0 also zeroes the sin
// half.  grid.x = ceil(nedge / 256) blocks for the float64 edge blocks
// (eof_fold_edge) of the m = 0 cos rows, first so that they do not trail,
// then (2 mmax + 1) * chunks blocks of these; grid.y = coefficient sets.
__global__ __launch_bounds__(256) void eof_fold_kernel(
    const float* __restrict__ T0, const float* __restrict__ T1, const float* __restrict__ T2,
    const float* __restrict__ T3, const double* __restrict__ ccos, const double* __restrict__ csin,
    int mmax, int norder, int64_t nodes, int chunks, EOFFoldArgs a, float* __restrict__ out,
    int64_t coef_stride, int64_t out_stride, const double* __restrict__ E, int64_t nedge) {
  const unsigned eblocks = (unsigned)((nedge + 255) / 256);
  if (blockIdx.x < eblocks) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e < nedge)
      eof_fold_edge(E, nedge, e, ccos + (size_t)blockIdx.y * coef_stride, mmax, norder, a,
                    out + (size_t)blockIdx.y * out_stride);
    return;
  }
  const unsigned b = blockIdx.x - eblocks;
  const int q = (int)(b / chunks);
  const int64_t node = (int64_t)(b - (unsigned)q * chunks) * 256 + threadIdx.x;
  if (node >= nodes) return;
  const bool sn = q > mmax;
  const int m = sn ? q - mmax : q;
  const double* cf = (sn ? csin : ccos) + (size_t)blockIdx.y * coef_stride + (size_t)m * norder;
  const double mul = a.mmul[m];
  const size_t row = (size_t)q * norder * nodes + node;
  const float* t0 = T0 + row;
  const float* t1 = T1 + row;
  const float* t2 = T2 + row;
  const float* t3 = T3 + row;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  if (a.dens) {
#pragma unroll 6
    for (int n = 0; n < a.nuse; ++n) {
      const double c = cf[n];
      const size_t o = (size_t)n * nodes;
      s0 += c * t0[o];
      s1 += c * t1[o];
      s2 += c * t2[o];
      s3 += c * t3[o];
    }
  } else {
#pragma unroll 6
    for (int n = 0; n < a.nuse; ++n) {
      const double c = cf[n];
      const size_t o = (size_t)n * nodes;
      s0 += c * t0[o];
      s1 += c * t1[o];
      s2 += c * t2[o];
    }
  }
  float* o = out + (size_t)blockIdx.y * out_stride + ((size_t)node * (mmax + 1) + m) * 8 + (sn ? 4 : 0);
  o[0] = (float)(s0 * mul);
  o[1] = (float)(s1 * mul);
  o[2] = (float)(s2 * mul);
  o[3] = (float)(s3 * mul);
  if (q == 0) o[4] = o[5] = o[6] = o[7] = 0.f;
}

static int ensure_folded(Context* c, float** d, size_t* have, size_t floats, hipStream_t s) {
  if (*d && *have >= floats) return 0;
  if (*d) {
    BFE_HIP(c, hipStreamSynchronize(s));
    (void)hipFree(*d);
    *d = nullptr;
    *have = 0;
  }
  BFE_HIP(c, hipMalloc(d, floats * sizeof(float)));
  *have = floats;
  return 0;
}

static size_t sl_fold_floats(const Context* c) { return (size_t)c->numr * sl_npairs(c->lmax) * 12; }
// float32 blocks of every node, then the float64 m = 0 cos blocks (four
// values, eight floats) of the edge nodes
static size_t eof_fold_floats(const Context* c) {
  return (size_t)eof_fold32_floats(c->numx, c->numy, c->mmax) +
         (size_t)eof_edge_count(c->numx, c->numy) * 8;
}

// Folds nset coefficient sets (stride (lmax+1)^2 nmax) into out (stride sl_fold_floats).
static int sl_fold_into(Context* c, const double* coef, int nset, double halofac, int lmax_use,
                        int nmax_use, int L1, int L2, bool no_odd, float* out, hipStream_t s) {
  const int np = sl_npairs(c->lmax);
  SLFoldArgs a;
  for (int l = 0; l <= kMaxL; ++l) {
    bool keep = l <= lmax_use;
    if (l >= 1 && (l > L2 || l < L1)) keep = false;
    if (l >= 1 && no_odd && (l & 1)) keep = false;
    a.lmul[l] = keep ? halofac : 0.0;
  }
  a.lmax = c->lmax;
  a.nmax_use = nmax_use;
  const int64_t total = (int64_t)c->numr * np * kFoldLanes;
  hipLaunchKernelGGL(sl_fold_kernel, dim3((unsigned)((total + 255) / 256), nset), dim3(256), 0, s,
                     c->sl_fold64, coef, c->nmax, c->numr, a, out,
                     (int64_t)(c->lmax + 1) * (c->lmax + 1) * c->nmax, (int64_t)sl_fold_floats(c));
  BFE_HIP(c, hipGetLastError());
  return 0;
}

static int sl_fold(Context* c, int slot, const double* coef, double halofac, int lmax_use,
                   int nmax_use, int L1, int L2, bool no_odd, hipStream_t s) {
  SLFolded& f = c->slf[slot];
  int rc = ensure_folded(c, &f.d, &f.floats, sl_fold_floats(c), s);
  if (rc) return rc;
  rc = sl_fold_into(c, coef, 1, halofac, lmax_use, nmax_use, L1, L2, no_odd, f.d, s);
  if (rc) return rc;
  f.valid = true;
  return 0;
}

static int eof_fold_into(Context* c, const double* ccos, const double* csin, int nset,
                         int mmax_use, int nuse, int m1, int m2, bool no_odd, bool compat_mask,
                         bool dens, float* out, hipStream_t s) {
  const int64_t nodes = (int64_t)(c->numx + 1) * (c->numy + 1);
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
  a.dens = dens ? 1 : 0;
  a.numx = c->numx;
  a.numy = c->numy;
  const int chunks = (int)((nodes + 255) / 256);
  const int64_t nedge = eof_edge_count(c->numx, c->numy);
  const unsigned grid = (unsigned)((2 * c->mmax + 1) * chunks + (nedge + 255) / 256);
  hipLaunchKernelGGL(eof_fold_kernel, dim3(grid, nset), dim3(256), 0, s, c->eof_fold32[0],
                     c->eof_fold32[1], c->eof_fold32[2],
                     dens ? c->eof_fold32[3] : c->eof_fold32[0], ccos, csin, c->mmax, c->norder,
                     nodes, chunks, a, out, (int64_t)(c->mmax + 1) * c->norder,
                     (int64_t)eof_fold_floats(c), c->eof_edge64, nedge);
  BFE_HIP(c, hipGetLastError());
  return 0;
}

static int eof_fold(Context* c, int slot, const double* ccos, const double* csin, int mmax_use,
                    int nuse, int m1, int m2, bool no_odd, bool compat_mask, bool dens,
                    hipStream_t s) {
  EOFFolded& f = c->eoff[slot];
  int rc = ensure_folded(c, &f.d, &f.floats, eof_fold_floats(c), s);
  if (rc) return rc;
  rc = eof_fold_into(c, ccos, csin, 1, mmax_use, nuse, m1, m2, no_odd, compat_mask, dens, f.d, s);
  if (rc) return rc;
  f.valid = true;
  return 0;
}

// ---------------------------------------------------------------------------
// batch evaluators
// ---------------------------------------------------------------------------
template <bool DENS>
__global__ __launch_bounds__(256) void sl_eval_kernel(
    const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
    int64_t n, SLGeom g, const float* __restrict__ F, int L, uint32_t flags, double densfac, float* den0, float* den1, float* pot0, float* pot1,
    float* potr, float* pott, float* potp) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double x = px[p], y = py[p], z = pz[p];
  const double r = sqrt(x * x + y * y + z * z);  // no floor (spheresl.py:1257)
  const double costh = z / r;                     // NaN at r = 0, as the reference
  double c1, s1;
  planar_angle(x, y, c1, s1);
  SLOut o;
  sl_eval_folded<DENS>(g, F, L, r, costh, c1, s1, flags, o);
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
    int64_t n, EOFGeom g, const float* __restrict__ F, float* p0, float* p, float* fr,
    float* fp, float* fz, float* d0, float* d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = px[i], y = py[i], z = pz[i];
  const double R = sqrt(x * x + y * y + 1.e-10);  // eof.py:1070
  double c1, s1;
  planar_angle(x, y, c1, s1);
  EOFOut o;
  eof_eval_folded<DENS>(g, F, R, z, c1, s1, o);
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

// Point forms: the caller gives the coordinates the reference's point
// evaluators take, in float64, and receives float64 rows of n values.
// spheresl.force_eval / all_eval (spheresl.py:1107-1234, :987-1102) at
// (r, cos theta, phi): out rows den0, den1 (l >= 1), pot0, pot1, potr, pott,
// potp; densities are written only with DENS (zero rows otherwise).
template <bool DENS>
__global__ __launch_bounds__(256) void sl_eval_sph_kernel(
    const double* __restrict__ pr, const double* __restrict__ pct, const double* __restrict__ pph,
    int64_t n, SLGeom g, const float* __restrict__ F, int L, uint32_t flags, double densfac,
    double* __restrict__ out) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double s1, c1;
  sincos(pph[p], &s1, &c1);
  SLOut o;
  sl_eval_folded<DENS>(g, F, L, pr[p], pct[p], c1, s1, flags, o);
  out[p] = DENS ? o.den0 * densfac : 0.0;
  out[n + p] = DENS ? o.den1 * densfac : 0.0;
  out[2 * n + p] = o.pot0;
  out[3 * n + p] = o.pot1;
  out[4 * n + p] = o.potr;
  out[5 * n + p] = o.pott;
  out[6 * n + p] = o.potp;
}

// eof.force_eval / accumulated_eval (eof.py:756-872, :874-929) at planar
// radius R (as given), z, phi: out rows p0, p (m >= 1), fr, fp, fz, d0,
// d (m >= 1), fr0, fz0 (the m = 0 parts of fr, fz).
template <bool DENS>
__global__ __launch_bounds__(256) void eof_eval_cyl_kernel(
    const double* __restrict__ pR, const double* __restrict__ pz, const double* __restrict__ pph,
    int64_t n, EOFGeom g, const float* __restrict__ F, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s1, c1;
  sincos(pph[i], &s1, &c1);
  EOFOut o;
  eof_eval_folded<DENS>(g, F, pR[i], pz[i], c1, s1, o);
  out[i] = o.p0;
  out[n + i] = o.p;
  out[2 * n + i] = o.fr;
  out[3 * n + i] = o.fp;
  out[4 * n + i] = o.fz;
  out[5 * n + i] = DENS ? o.d0 : 0.0;
  out[6 * n + i] = DENS ? o.d : 0.0;
  out[7 * n + i] = o.fr0;
  out[8 * n + i] = o.fz0;
}

// Fields.return_forces_cart at n points.  pts (may be NULL): the points in
// table-cell order as float4 (x, y, z, input index bits), outputs scattered
// back to input order; block b then takes logical block xcd_block(nb) so that
// each XCD evaluates a contiguous range of cells.  Without pts, point j of
// px/py/pz.  In cell order a block's points usually share a few SL radial
// nodes: when their force blocks fit in kStageFloat4, the block copies them
// to LDS and the SL sums read them there.
__global__ __launch_bounds__(256) void field_eval_kernel(
    const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
    int64_t n, const float4* __restrict__ pts, FieldTables T, float* fx, float* fy, float* fz,
    float* pot, float* comp, float4* __restrict__ quad) {
  __shared__ float4 stage[kStageFloat4];
  __shared__ int smin[4], smax[4];
  const int64_t nb = (n + 255) / 256;
  const int64_t j = (pts ? xcd_block(nb) : (int64_t)blockIdx.x) * 256 + threadIdx.x;
  if (pts && (j - threadIdx.x) >= n) return;   // whole block past the end (XCD mapping)
  const bool live = j < n;
  float qx = 0.f, qy = 0.f, qz = 0.f;
  int64_t i = 0;
  if (live) {
    if (pts) {
      const float4 q = pts[j];
      qx = q.x; qy = q.y; qz = q.z;
      i = (int64_t)__float_as_int(q.w);
    } else {
      qx = px[j]; qy = py[j]; qz = pz[j];
      i = j;
    }
  }
  bool staged = false;
  int imin = 0;
  SLCell cell{0.0, 0, 0.0, 0.0};
  if (pts) {
    if (live) cell = field_sl_cell(T.slg, qx, qy, qz);
    staged = stage_sl_nodes(T, live, cell.i, stage, smin, smax, imin);
  }
  if (!live) return;
  double cv[8];
  if (staged)
    field_eval_point<true>(T, qx, qy, qz, cv, &cell, stage, imin);
  else
    field_eval_point(T, qx, qy, qz, cv);
  if (quad) {
    // one 16-byte store per point; quad_split_kernel makes the four arrays
    quad[i] = make_float4((float)(cv[0] + cv[1]), (float)(cv[2] + cv[3]), (float)(cv[4] + cv[5]),
                          (float)(cv[6] + cv[7]));
    return;
  }
  if (fx) fx[i] = (float)(cv[0] + cv[1]);
  if (fy) fy[i] = (float)(cv[2] + cv[3]);
  if (fz) fz[i] = (float)(cv[4] + cv[5]);
  if (pot) pot[i] = (float)(cv[6] + cv[7]);
  if (comp) {
#pragma unroll
    for (int k = 0; k < 8; ++k) comp[(size_t)k * n + i] = (float)cv[k];
  }
}

// Batched frozen fields (one coefficient set per snapshot): point j of the
// (snapshot, cell) order belongs to snapshot sorted[j] / ncell and uses that
// snapshot's folded tables.
__global__ __launch_bounds__(256) void field_eval_batch_kernel(
    int64_t n, const float4* __restrict__ pts, const uint32_t* __restrict__ sorted, uint32_t ncell,
    FieldTables T, const float* __restrict__ slB, int64_t sl_stride,
    const float* __restrict__ eofB, int64_t eof_stride, float* fx, float* fy, float* fz,
    float* pot, float4* __restrict__ quad) {
  __shared__ float4 stage[kStageFloat4];
  __shared__ int smin[4], smax[4];
  const int64_t nb = (n + 255) / 256;
  const int64_t j = xcd_block(nb) * 256 + threadIdx.x;
  if (j - threadIdx.x >= n) return;   // whole block past the end (XCD mapping)
  const bool live = j < n;
  float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
  int snap = 0;
  if (live) {
    q = pts[j];
    snap = (int)(sorted[j] / ncell);
  }
  // SL nodes are staged when all the block's points belong to one snapshot
  int s0, s1;
  block_minmax(live ? snap : 0x7fffffff, live ? snap : -1, smin, smax, s0, s1);
  __syncthreads();   // smin/smax are reused below
  T.slF = slB + (size_t)(live ? snap : s0) * sl_stride;
  T.eofF = eofB + (size_t)(live ? snap : s0) * eof_stride;
  SLCell cell{0.0, 0, 0.0, 0.0};
  bool staged = false;
  int imin = 0;
  if (s0 == s1) {
    if (live) cell = field_sl_cell(T.slg, q.x, q.y, q.z);
    staged = stage_sl_nodes(T, live, cell.i, stage, smin, smax, imin);
  }
  if (!live) return;
  const int64_t i = (int64_t)__float_as_int(q.w);
  double cv[8];
  if (staged)
    field_eval_point<true>(T, q.x, q.y, q.z, cv, &cell, stage, imin);
  else
    field_eval_point(T, q.x, q.y, q.z, cv);
  quad[i] = make_float4((float)(cv[0] + cv[1]), (float)(cv[2] + cv[3]), (float)(cv[4] + cv[5]),
                        (float)(cv[6] + cv[7]));
}

static int ensure_batch(Context* c, float** d, size_t* have, size_t floats) {
  if (*d && *have >= floats) return 0;
  if (*d) {
    BFE_HIP(c, hipDeviceSynchronize());
    (void)hipFree(*d);
    *d = nullptr;
    *have = 0;
  }
  BFE_HIP(c, hipMalloc(d, floats * sizeof(float)));
  *have = floats;
  return 0;
}

void batch_free(Context* c) {
  if (c->batch_sl) (void)hipFree(c->batch_sl);
  if (c->batch_eof) (void)hipFree(c->batch_eof);
  if (c->batch_off) (void)hipFree(c->batch_off);
  c->batch_sl = nullptr;
  c->batch_eof = nullptr;
  c->batch_off = nullptr;
  c->batch_sl_floats = c->batch_eof_floats = 0;
  c->batch_off_cap = 0;
}

int field_tables(Context* c, double rotpos, FieldTables& T) {
  if (!c->slf[1].valid || !c->eoff[1].valid)
    return set_error(c, BFE_ENOTABLE, "field not set (call bfe_field_set)");
  T.slg = sl_geom(c);
  T.eg = eof_geom(c);
  T.slF = c->slf[1].d;
  T.eofF = c->eoff[1].d;
  T.lmax = c->lmax;
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
  int rc = sl_fold(c, 0, coef, 1.0, c->lmax, c->nmax, L1, L2, (flags & BFE_NO_ODD) != 0, s);
  if (rc) return rc;
  if (n == 0) return 0;
  const SLFolded& f = c->slf[0];
  const double densfac = (flags & BFE_COMPAT_SL_DENS) ? 0.25 * kPi : 0.25 / kPi;
  const uint32_t ef = flags & (BFE_COMPAT_SL_DENS | BFE_COMPAT_SL_EDGE | BFE_COMPAT_SL_COS_L);
  if (dens)
    hipLaunchKernelGGL(sl_eval_kernel<true>, dim3(nblocks256(n)), dim3(256), 0, s, x, y, z, n,
                       sl_geom(c), f.d, c->lmax, ef, densfac, den0, den1, pot0, pot1, potr, pott,
                       potp);
  else
    hipLaunchKernelGGL(sl_eval_kernel<false>, dim3(nblocks256(n)), dim3(256), 0, s, x, y, z, n,
                       sl_geom(c), f.d, c->lmax, ef, densfac, den0, den1, pot0, pot1, potr, pott,
                       potp);
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
                       eof_geom(c), f.d, p0, p, fr, fp, fz, d0, d);
  else
    hipLaunchKernelGGL(eof_eval_kernel<false>, dim3(nblocks256(n)), dim3(256), 0, s, x, y, z, n,
                       eof_geom(c), f.d, p0, p, fr, fp, fz, d0, d);
  BFE_HIP(c, hipGetLastError());
  return 0;
}

extern "C" int bfe_sl_eval_sph(bfe_ctx* h, const double* r, const double* costh,
                               const double* phi, int64_t n, const double* coef, int lmax_use,
                               int nmax_use, uint32_t flags, double* out, void* stream) {
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (!c->sl_ok) return set_error(c, BFE_ENOTABLE, "sl_eval_sph: SL tables not set");
  if (n < 0 || !coef) return set_error(c, BFE_EINVAL, "sl_eval_sph: bad n or NULL coef");
  if (n > 0 && (!r || !costh || !phi || !out))
    return set_error(c, BFE_EINVAL, "sl_eval_sph: NULL input or output");
  BFE_HIP(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream;
  const int ll = lmax_use < 0 ? c->lmax : std::min(lmax_use, c->lmax);
  const int nn = nmax_use < 0 ? c->nmax : std::min(nmax_use, c->nmax);
  int rc = sl_fold(c, 0, coef, 1.0, ll, nn, -(1 << 30), 1 << 30, (flags & BFE_NO_ODD) != 0, s);
  if (rc) return rc;
  if (n == 0) return 0;
  const SLFolded& f = c->slf[0];
  const uint32_t ef = flags & (BFE_COMPAT_SL_EDGE | BFE_COMPAT_SL_COS_L);
  if (flags & BFE_EVAL_DENS)
    hipLaunchKernelGGL(sl_eval_sph_kernel<true>, dim3(nblocks256(n)), dim3(256), 0, s, r, costh,
                       phi, n, sl_geom(c), f.d, c->lmax, ef, 0.25 / kPi, out);
  else
    hipLaunchKernelGGL(sl_eval_sph_kernel<false>, dim3(nblocks256(n)), dim3(256), 0, s, r, costh,
                       phi, n, sl_geom(c), f.d, c->lmax, ef, 0.25 / kPi, out);
  BFE_HIP(c, hipGetLastError());
  return 0;
}

extern "C" int bfe_eof_eval_cyl(bfe_ctx* h, const double* R, const double* z, const double* phi,
                                int64_t n, const double* cos_coef, const double* sin_coef,
                                int mmax_use, int nmax_use, uint32_t flags, double* out,
                                void* stream) {
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (!c->eof_ok) return set_error(c, BFE_ENOTABLE, "eof_eval_cyl: EOF tables not set");
  if (n < 0 || !cos_coef || !sin_coef)
    return set_error(c, BFE_EINVAL, "eof_eval_cyl: bad n or NULL coef");
  if (n > 0 && (!R || !z || !phi || !out))
    return set_error(c, BFE_EINVAL, "eof_eval_cyl: NULL input or output");
  const bool dens = (flags & BFE_EVAL_DENS) != 0;
  if (dens && !c->eof_dens) return set_error(c, BFE_ENOTABLE, "eof_eval_cyl: no density tables");
  BFE_HIP(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream;
  const int mm = mmax_use < 0 ? c->mmax : std::min(mmax_use, c->mmax);
  const int nn = nmax_use < 0 ? c->norder : std::min(nmax_use, c->norder);
  int rc = eof_fold(c, 0, cos_coef, sin_coef, mm, nn, 0, 1 << 30, (flags & BFE_NO_ODD) != 0,
                    (flags & BFE_COMPAT_EOF_MASK) != 0, dens, s);
  if (rc) return rc;
  if (n == 0) return 0;
  const EOFFolded& f = c->eoff[0];
  if (dens)
    hipLaunchKernelGGL(eof_eval_cyl_kernel<true>, dim3(nblocks256(n)), dim3(256), 0, s, R, z, phi,
                       n, eof_geom(c), f.d, out);
  else
    hipLaunchKernelGGL(eof_eval_cyl_kernel<false>, dim3(nblocks256(n)), dim3(256), 0, s, R, z,
                       phi, n, eof_geom(c), f.d, out);
  BFE_HIP(c, hipGetLastError());
  return 0;
}

namespace bfe {

// Fold one coefficient set into the frozen field (bfe_field_set).
static int field_set_run(Context* c, const double* eof_cos, const double* eof_sin,
                         const double* sl_coef, double halofac, int disk_m, int disk_n,
                         int halo_l, int halo_n, uint32_t flags, hipStream_t s) {
  const int mm = disk_m < 0 ? c->mmax : std::min(disk_m, c->mmax);
  const int nn = disk_n < 0 ? c->norder : std::min(disk_n, c->norder);
  const int ll = halo_l < 0 ? c->lmax : std::min(halo_l, c->lmax);
  const int hn = halo_n < 0 ? c->nmax : std::min(halo_n, c->nmax);
  const bool no_odd = (flags & BFE_NO_ODD) != 0;
  // the SL fold (a few thousand waves) runs beside the HBM-bound EOF fold
  hipStream_t s2 = nullptr;
  int rc = fork_begin(c, s, &s2);
  if (rc) return rc;
  const int rs = sl_fold(c, 1, sl_coef, halofac, ll, hn, -(1 << 30), 1 << 30, no_odd, s2);
  const int re = rs ? 0 : eof_fold(c, 1, eof_cos, eof_sin, mm, nn, 0, 1 << 30, no_odd,
                                   (flags & BFE_COMPAT_EOF_MASK) != 0, false, s);
  rc = fork_end(c, s);
  if (rs || re || rc) return rs ? rs : (re ? re : rc);
  c->field_halofac = halofac;
  c->field_flags = (int)(flags & (BFE_COMPAT_SL_COS_L | BFE_COMPAT_SL_EDGE));
  return 0;
}

// Evaluate n points: sorted (pts, the index in .w) or in input order.
// (fx, fy, fz, pot) rows of the float4 quad array, coalesced both ways.
__global__ __launch_bounds__(256) void quad_split_kernel(const float4* __restrict__ quad, int64_t n,
                                                         float* fx, float* fy, float* fz,
                                                         float* pot) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const float4 v = quad[i];
    if (fx) fx[i] = v.x;
    if (fy) fy[i] = v.y;
    if (fz) fz[i] = v.z;
    if (pot) pot[i] = v.w;
  }
}

// Evaluate n points: sorted (pts, the index in .w) or in input order.  In
// sorted order the outputs land at scattered indices; without comp, each
// point writes one float4 to scratch (one 16-byte store instead of four
// 4-byte ones, a quarter of the partial-line HBM writes) and a coalesced
// pass splits it into fx, fy, fz, pot.
static int field_eval_run(Context* c, const FieldTables& T, const float* x, const float* y,
                          const float* z, int64_t n, const float4* pts, float* fx, float* fy,
                          float* fz, float* pot, float* comp, hipStream_t s) {
  if (n == 0) return 0;
  float4* quad = nullptr;
  if (pts && !comp && (fx || fy || fz || pot)) {
    int rc = ensure_scratch(c, (size_t)n * sizeof(float4));
    if (rc) return rc;
    quad = (float4*)c->scratch;
  }
  const unsigned grid = pts ? xcd_grid(nblocks256(n)) : nblocks256(n);
  timer_begin(c, "field_eval", s);
  hipLaunchKernelGGL(field_eval_kernel, dim3(grid), dim3(256), 0, s, x, y, z, n, pts, T, fx, fy,
                     fz, pot, comp, quad);
  timer_end(c, "field_eval", s);
  BFE_HIP(c, hipGetLastError());
  if (quad) {
    hipLaunchKernelGGL(quad_split_kernel, dim3(nblocks256(n)), dim3(256), 0, s, quad, n, fx, fy,
                       fz, pot);
    BFE_HIP(c, hipGetLastError());
  }
  return 0;
}

}  // namespace bfe

extern "C" int bfe_field_set(bfe_ctx* h, const double* eof_cos, const double* eof_sin,
                             const double* sl_coef, double halofac, int disk_m, int disk_n,
                             int halo_l, int halo_n, uint32_t flags, void* stream) {
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (!c->sl_ok || !c->eof_ok) return set_error(c, BFE_ENOTABLE, "field_set: tables not set");
  if (!eof_cos || !eof_sin || !sl_coef) return set_error(c, BFE_EINVAL, "field_set: NULL coef");
  BFE_HIP(c, hipSetDevice(c->device));
  return field_set_run(c, eof_cos, eof_sin, sl_coef, halofac, disk_m, disk_n, halo_l, halo_n,
                       flags, (hipStream_t)stream);
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
  const float4* pts = nullptr;
  if (n >= c->order_min) {
    rc = order_points(c, T, x, y, z, n, s, nullptr, nullptr, 1, nullptr, &pts);
    if (rc) return rc;
  }
  return field_eval_run(c, T, x, y, z, n, pts, fx, fy, fz, pot, comp, s);
}

// Fields.total_coefficients followed by Fields.return_forces_cart at the
// same particles (potential.py:98-252, :445-497).  Two concurrent branches,
// each folding its coefficients as soon as they are reduced:
//   caller stream  disc EOF accumulate, the EOF fold, then the evaluation
//                  order of all nd + nh points taken as one set (it needs
//                  only positions and table geometry; the disc branch is
//                  the shorter one, so the sort fills its idle tail);
//   internal       halo SL accumulate (second buffer set), then the SL fold;
// then, after the join, the evaluation.  Results equal
// bfe_accumulate_disc_halo + bfe_field_set + bfe_field_eval_cart over the
// concatenated points.
extern "C" int bfe_forces_disc_halo(bfe_ctx* h, const float* dx, const float* dy,
                                    const float* dz, const float* dm, int64_t nd, const float* hx,
                                    const float* hy, const float* hz, const float* hm, int64_t nh,
                                    uint32_t acc_flags, double* cos_out, double* sin_out,
                                    double* coef_out, double halofac, int disk_m, int disk_n,
                                    int halo_l, int halo_n, uint32_t field_flags, double rotpos,
                                    float* fx, float* fy, float* fz, float* pot, float* comp,
                                    void* stream) {
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (!c->eof_ok || !c->sl_ok) return set_error(c, BFE_ENOTABLE, "forces_disc_halo: tables not set");
  if (!cos_out || !sin_out || !coef_out)
    return set_error(c, BFE_EINVAL, "forces_disc_halo: NULL coefficient output");
  if (nd < 0 || nh < 0) return set_error(c, BFE_EINVAL, "forces_disc_halo: n < 0");
  if ((nd > 0 && (!dx || !dy || !dz || !dm)) || (nh > 0 && (!hx || !hy || !hz || !hm)))
    return set_error(c, BFE_EINVAL, "forces_disc_halo: NULL input");
  if (nd + nh > (int64_t)0x7fffffff) return set_error(c, BFE_ELIMIT, "forces_disc_halo: nd + nh > 2^31 - 1");
  BFE_HIP(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream, s1 = nullptr;
  const int64_t n = nd + nh;
  const int mm = disk_m < 0 ? c->mmax : std::min(disk_m, c->mmax);
  const int nn = disk_n < 0 ? c->norder : std::min(disk_n, c->norder);
  const int ll = halo_l < 0 ? c->lmax : std::min(halo_l, c->lmax);
  const int hn = halo_n < 0 ? c->nmax : std::min(halo_n, c->nmax);
  const bool no_odd = (field_flags & BFE_NO_ODD) != 0;
  int rc = fork_begin(c, s, &s1);
  if (rc) return rc;
  // halo: accumulate and fold on the first internal stream
  swap_work(c);
  int r1 = sl_accumulate_one(c, hx, hy, hz, hm, nh, acc_flags, coef_out, s1);
  swap_work(c);
  if (!r1) r1 = sl_fold(c, 1, coef_out, halofac, ll, hn, -(1 << 30), 1 << 30, no_odd, s1);
  // disc: accumulate and fold on the caller's stream
  int r0 = eof_accumulate_one(c, dx, dy, dz, dm, nd, cos_out, sin_out, s);
  if (!r0) r0 = eof_fold(c, 1, cos_out, sin_out, mm, nn, 0, 1 << 30, no_odd,
                         (field_flags & BFE_COMPAT_EOF_MASK) != 0, false, s);
  // the two inputs are evaluated as one set, so they are ordered at every size
  const float4* pts = nullptr;
  if (!r0 && n > 0) {
    FieldTables G;
    G.slg = sl_geom(c);
    G.eg = eof_geom(c);
    r0 = order_points2(c, G, dx, dy, dz, nd, hx, hy, hz, nh, s, &pts);
  }
  rc = fork_end(c, s);
  if (r0 || r1 || rc) return r0 ? r0 : (r1 ? r1 : rc);
  c->field_halofac = halofac;
  c->field_flags = (int)(field_flags & (BFE_COMPAT_SL_COS_L | BFE_COMPAT_SL_EDGE));
  FieldTables T;
  rc = field_tables(c, rotpos, T);
  if (rc || n == 0) return rc;
  return field_eval_run(c, T, nullptr, nullptr, nullptr, n, pts, fx, fy, fz, pot, comp, s);
}

extern "C" int bfe_field_set_batch(bfe_ctx* h, const double* eof_cos, const double* eof_sin,
                                   const double* sl_coef, int nsnap, double halofac, int disk_m,
                                   int disk_n, int halo_l, int halo_n, uint32_t flags,
                                   void* stream) {
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (!c->sl_ok || !c->eof_ok) return set_error(c, BFE_ENOTABLE, "field_set_batch: tables not set");
  if (!eof_cos || !eof_sin || !sl_coef) return set_error(c, BFE_EINVAL, "field_set_batch: NULL coef");
  if (nsnap < 1 || nsnap > 65535) return set_error(c, BFE_EINVAL, "field_set_batch: nsnap outside [1, 65535]");
  const int mm = disk_m < 0 ? c->mmax : std::min(disk_m, c->mmax);
  const int nn = disk_n < 0 ? c->norder : std::min(disk_n, c->norder);
  const int ll = halo_l < 0 ? c->lmax : std::min(halo_l, c->lmax);
  const int hn = halo_n < 0 ? c->nmax : std::min(halo_n, c->nmax);
  BFE_HIP(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream;
  c->batch_nsnap = 0;
  const size_t slf = sl_fold_floats(c), eoff = eof_fold_floats(c);
  int rc = ensure_batch(c, &c->batch_sl, &c->batch_sl_floats, slf * nsnap);
  if (rc) return rc;
  rc = ensure_batch(c, &c->batch_eof, &c->batch_eof_floats, eoff * nsnap);
  if (rc) return rc;
  const bool no_odd = (flags & BFE_NO_ODD) != 0;
  rc = eof_fold_into(c, eof_cos, eof_sin, nsnap, mm, nn, 0, 1 << 30, no_odd,
                     (flags & BFE_COMPAT_EOF_MASK) != 0, false, c->batch_eof, s);
  if (rc) return rc;
  rc = sl_fold_into(c, sl_coef, nsnap, halofac, ll, hn, -(1 << 30), 1 << 30, no_odd, c->batch_sl, s);
  if (rc) return rc;
  c->batch_nsnap = nsnap;
  c->batch_flags = (int)(flags & (BFE_COMPAT_SL_COS_L | BFE_COMPAT_SL_EDGE));
  return 0;
}

extern "C" int bfe_field_eval_batch(bfe_ctx* h, const float* x, const float* y, const float* z,
                                    const int64_t* offsets, int nsnap, double rotpos, float* fx,
                                    float* fy, float* fz, float* pot, void* stream) {
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (!c->sl_ok || !c->eof_ok) return set_error(c, BFE_ENOTABLE, "field_eval_batch: tables not set");
  if (c->batch_nsnap < 1) return set_error(c, BFE_ENOTABLE, "field_eval_batch: no batch field (bfe_field_set_batch)");
  if (c->batch_sl_floats < sl_fold_floats(c) * (size_t)c->batch_nsnap ||
      c->batch_eof_floats < eof_fold_floats(c) * (size_t)c->batch_nsnap)
    return set_error(c, BFE_ENOTABLE, "field_eval_batch: batch field folded for other tables (call bfe_field_set_batch again)");
  if (!offsets || nsnap != c->batch_nsnap)
    return set_error(c, BFE_EINVAL, "field_eval_batch: need offsets for the %d snapshots of the batch field",
                     c->batch_nsnap);
  if (offsets[0] != 0) return set_error(c, BFE_EINVAL, "field_eval_batch: offsets[0] must be 0");
  for (int q = 0; q < nsnap; ++q)
    if (offsets[q + 1] < offsets[q]) return set_error(c, BFE_EINVAL, "field_eval_batch: offsets must not decrease");
  const int64_t n = offsets[nsnap];
  if (n > 0 && (!x || !y || !z)) return set_error(c, BFE_EINVAL, "field_eval_batch: NULL input");
  if (n == 0) return 0;
  BFE_HIP(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream;
  if (c->batch_off_cap < (size_t)nsnap + 1) {
    if (c->batch_off) {
      BFE_HIP(c, hipDeviceSynchronize());
      (void)hipFree(c->batch_off);
      c->batch_off = nullptr;
    }
    BFE_HIP(c, hipMalloc(&c->batch_off, ((size_t)nsnap + 1) * sizeof(int64_t)));
    c->batch_off_cap = (size_t)nsnap + 1;
  }
  // the offsets buffer may still be read by an earlier call on this stream
  BFE_HIP(c, hipStreamSynchronize(s));
  BFE_HIP(c, hipMemcpy(c->batch_off, offsets, ((size_t)nsnap + 1) * sizeof(int64_t),
                       hipMemcpyHostToDevice));
  const size_t slf = sl_fold_floats(c), eoff = eof_fold_floats(c);
  FieldTables T;
  T.slg = sl_geom(c);
  T.eg = eof_geom(c);
  T.slF = c->batch_sl;
  T.eofF = c->batch_eof;
  T.lmax = c->lmax;
  T.flags = (uint32_t)c->batch_flags;
  T.ximin = *std::min_element(c->h_xi.begin(), c->h_xi.end());
  T.crot = cos(rotpos);
  T.srot = sin(rotpos);
  const uint32_t* sorted = nullptr;
  const float4* pts = nullptr;
  int rc = order_points(c, T, x, y, z, n, s, nullptr, c->batch_off, nsnap, &sorted, &pts, offsets);
  if (rc) return rc;
  // results go through a float4 per point, split coalesced (as field_eval_run)
  rc = ensure_scratch(c, (size_t)n * sizeof(float4));
  if (rc) return rc;
  float4* quad = (float4*)c->scratch;
  hipLaunchKernelGGL(field_eval_batch_kernel, dim3(xcd_grid(nblocks256(n))), dim3(256), 0, s, n,
                     pts, sorted, order_cell_count(T, nsnap), T, c->batch_sl, (int64_t)slf, c->batch_eof,
                     (int64_t)eoff, fx, fy, fz, pot, quad);
  BFE_HIP(c, hipGetLastError());
  hipLaunchKernelGGL(quad_split_kernel, dim3(nblocks256(n)), dim3(256), 0, s, quad, n, fx, fy, fz,
                     pot);
  BFE_HIP(c, hipGetLastError());
  return 0;
}
