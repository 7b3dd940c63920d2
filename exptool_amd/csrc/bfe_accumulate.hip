// Coefficient accumulation kernels.
//
// SL  (spheresl.py:567-656): expcoef[row][n] = sum_p Y_row(p) phi_{l(row) n}(r_p),
//      Y_row = -4 pi m_p N_lm P_l^m(cos th) {1, cos m phi, sin m phi}.
// EOF (eof.py:492-640):     accum[m][n] = sum_p -4 pi m_p {cos,sin}(m phi) V_mn(R_p, z_p),
//      V = bilinear interpolation of potC/potS (eof.py:430-457).
//
// Both kernels stage per-particle geometry (float64) in LDS for a tile of
// particles, then lanes own coefficient entries and stream the tile.  Each
// lane sums one tile in float32, adds the tile sum into a float64 register,
// and every block writes float64 partials; a second kernel reduces the
// partials over blocks in a fixed order with Kahan compensation, so results
// are bitwise reproducible for a given n.
#include "bfe_internal.hpp"

namespace bfe {

constexpr int kTile = 256;   // EOF: particles per tile (one per lane for geometry)
constexpr int kSLTile = 64;  // SL: particles per tile

// ---------------------------------------------------------------------------
// EOF accumulate
// ---------------------------------------------------------------------------
template <int OPT>
__global__ __launch_bounds__(256) void eof_accumulate_kernel(
    const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
    const float* __restrict__ pm, int64_t n, EOFGeom g, const float* __restrict__ T, int K,
    double* __restrict__ partials) {
  __shared__ int s_base[kTile];
  __shared__ float s_c[4][kTile];
  __shared__ float s_w[2 * kMaxM + 1][kTile + 1];  // +1: spread rows over banks

  const int tid = threadIdx.x;
  const int ny1 = g.numy + 1;
  const int J = 2 * g.mmax + 1;
  float acc[OPT];
  double accd[OPT];
  int rowg[OPT];
#pragma unroll
  for (int o = 0; o < OPT; ++o) {
    acc[o] = 0.f;
    accd[o] = 0.0;
    const int k = tid + o * 256;
    rowg[o] = (k < K) ? k / g.norder : 0;
  }
  const int64_t ntiles = (n + kTile - 1) / kTile;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int64_t p = t * kTile + tid;
    if (p < n) {
      const double x = px[p], y = py[p], z = pz[p], mass = pm[p];
      const double R = sqrt(x * x + y * y + 1.e-10);
      double c1, s1;
      planar_angle(x, y, c1, s1);
      EOFCell cell = eof_locate(R, z, g.cmap, g.ascale, g.hscale, g.xmin, g.dx, g.ymin, g.dy,
                                g.numx, g.numy);
      s_base[tid] = cell.ix * ny1 + cell.iy;
      s_c[0][tid] = (float)cell.c00;
      s_c[1][tid] = (float)cell.c10;
      s_c[2][tid] = (float)cell.c01;
      s_c[3][tid] = (float)cell.c11;
      const double w = -kFourPi * mass;
      double cm = 1.0, sm = 0.0, cm1 = c1, sm1 = -s1;
      s_w[0][tid] = (float)w;
      for (int m = 1; m <= g.mmax; ++m) {
        const double cn = (m == 1) ? c1 : 2.0 * c1 * cm - cm1;
        const double sn = (m == 1) ? s1 : 2.0 * c1 * sm - sm1;
        cm1 = cm; sm1 = sm; cm = cn; sm = sn;
        s_w[m][tid] = (float)(w * cm);
        s_w[g.mmax + m][tid] = (float)(w * sm);
      }
    } else {
      s_base[tid] = 0;
      s_c[0][tid] = s_c[1][tid] = s_c[2][tid] = s_c[3][tid] = 0.f;
      for (int j = 0; j < J; ++j) s_w[j][tid] = 0.f;
    }
    __syncthreads();
    const int64_t left = n - t * kTile;
    const int np = left < kTile ? (int)left : kTile;
    for (int q = 0; q < np; ++q) {
      const size_t b = (size_t)s_base[q];
      const float* T00 = T + b * K;
      const float* T01 = T00 + K;
      const float* T10 = T00 + (size_t)ny1 * K;
      const float* T11 = T10 + K;
      const float c00 = s_c[0][q], c10 = s_c[1][q], c01 = s_c[2][q], c11 = s_c[3][q];
#pragma unroll
      for (int o = 0; o < OPT; ++o) {
        const int k = tid + o * 256;
        if (k < K) {
          const float v = c00 * T00[k] + c10 * T10[k] + c01 * T01[k] + c11 * T11[k];
          acc[o] = fmaf(s_w[rowg[o]][q], v, acc[o]);
        }
      }
    }
#pragma unroll
    for (int o = 0; o < OPT; ++o) {
      accd[o] += (double)acc[o];
      acc[o] = 0.f;
    }
    __syncthreads();
  }
#pragma unroll
  for (int o = 0; o < OPT; ++o) {
    const int k = tid + o * 256;
    if (k < K) partials[(size_t)blockIdx.x * K + k] = accd[o];
  }
}

// Fixed-order reduction of partials[nb][K] -> out[K]: one block per k; lane t
// sums b = t, t + 256, ... with Kahan compensation, then a fixed pairwise tree
// in LDS combines the 256 lane sums.  The order depends only on nb, so results
// are bitwise reproducible.
__global__ __launch_bounds__(256) void reduce_partials_kernel(const double* __restrict__ partials,
                                                              int nb, int K,
                                                              double* __restrict__ out) {
  __shared__ double sh[256];
  const int k = blockIdx.x;
  const int t = threadIdx.x;
  double s = 0.0, comp = 0.0;
  for (int b = t; b < nb; b += 256) {
    const double yv = partials[(size_t)b * K + k] - comp;
    const double tv = s + yv;
    comp = (tv - s) - yv;
    s = tv;
  }
  sh[t] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) sh[t] += sh[t + w];
    __syncthreads();
  }
  if (t == 0) out[k] = sh[0];
}

// Scatter the K-ordered EOF sums into (mmax+1, norder) cos and sin arrays.
__global__ void eof_split_kernel(const double* __restrict__ sums, int mmax, int norder,
                                 double* __restrict__ cos_out, double* __restrict__ sin_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int nc = (mmax + 1) * norder;
  if (i < nc) {
    cos_out[i] = sums[i];
    sin_out[i] = (i < norder) ? 0.0 : sums[nc + i - norder];
  }
}

// ---------------------------------------------------------------------------
// SL accumulate
// ---------------------------------------------------------------------------
template <int OPT>
__global__ __launch_bounds__(256) void sl_accumulate_kernel(
    const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
    const float* __restrict__ pm, int64_t n, SLGeom g, const float* __restrict__ E,
    const double* __restrict__ lmc, int no_odd, double* __restrict__ partials) {
  extern __shared__ float smem[];
  const int R = (g.lmax + 1) * (g.lmax + 1);
  const int LN = (g.lmax + 1) * g.nmax;
  const int K = R * g.nmax;
  constexpr int YS = kSLTile + 1;  // padded row stride of s_Y (bank spread)
  float* s_Y = smem;                           // [R][YS]
  float* s_phi = s_Y + R * YS;                 // [kSLTile][LN]
  float* s_w1 = s_phi + kSLTile * LN;          // [kSLTile]
  float* s_w2 = s_w1 + kSLTile;                // [kSLTile]
  int* s_i = (int*)(s_w2 + kSLTile);           // [kSLTile]

  const int tid = threadIdx.x;
  const int qp = tid & (kSLTile - 1);  // particle of the tile
  const int h = tid >> 6;              // this wave builds the Y columns m = h, h + 4, ...
  float acc[OPT];
  double accd[OPT];
  int yrow[OPT], prow[OPT];
#pragma unroll
  for (int o = 0; o < OPT; ++o) {
    acc[o] = 0.f;
    accd[o] = 0.0;
    const int k = tid + o * 256;
    const int row = (k < K) ? k / g.nmax : 0;
    const int nn = (k < K) ? k % g.nmax : 0;
    int l = 0;
    while ((l + 1) * (l + 1) <= row) ++l;
    yrow[o] = row * YS;
    prow[o] = l * g.nmax + nn;
  }
  const int64_t ntiles = (n + kSLTile - 1) / kSLTile;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int64_t p = t * kSLTile + qp;
    {
      // padding particles: unit position, zero mass, so every Y row is 0
      const bool live = p < n;
      const double x = live ? (double)px[p] : 1.0, y = live ? (double)py[p] : 0.0;
      const double z = live ? (double)pz[p] : 0.0, mass = live ? (double)pm[p] : 0.0;
      double r = sqrt(x * x + y * y + z * z);
      r = (r == r) ? (r > 1.0e-10 ? r : 1.0e-10) : 1.0e-10;   // np.nanmax([r, 1e-10])
      const double costh = z / r;
      double c1, s1;
      planar_angle(x, y, c1, s1);
      if (h == 0) {
        SLCell cell = sl_locate(r, g.cmap, g.scale, g.xi, g.numr);
        const double P0 = cell.x1 * g.p0[cell.i] + cell.x2 * g.p0[cell.i + 1];
        s_i[qp] = cell.i;
        s_w1[qp] = live ? (float)(cell.x1 * P0) : 0.f;
        s_w2[qp] = live ? (float)(cell.x2 * P0) : 0.f;
      }
      // Y rows -4 pi m N_lm P_l^m(cos th) {1, cos m phi, sin m phi}: P_m^m in
      // closed form, then the legendre_R recurrence in l (spheresl.py:678-696)
      // with tabulated coefficients; e^{i m phi} by complex powers.
      const bool bad = !(costh == costh);
      const double somx2 = sqrt((1.0 - costh) * (1.0 + costh));
      double ch = 1.0, sh = 0.0, pw = 1.0;
      for (int j = 0; j < h; ++j) {
        const double cn = ch * c1 - sh * s1;
        sh = sh * c1 + ch * s1;
        ch = cn;
        pw *= somx2;
      }
      const double c2 = c1 * c1 - s1 * s1, s2 = 2.0 * c1 * s1;
      const double c4 = c2 * c2 - s2 * s2, s4 = 2.0 * c2 * s2;
      const double sq2 = somx2 * somx2, p4 = sq2 * sq2;
      for (int m = h; m <= g.lmax; m += 4) {
        double plm1 = 0.0, pl = lmc[kLmcD + m] * pw;
        for (int l = m; l <= g.lmax; ++l) {
          const int lm = lm_index(l, m);
          if (l > m) {
            const double pn = costh * lmc[kLmcA + lm] * pl - lmc[kLmcB + lm] * plm1;
            plm1 = pl;
            pl = pn;
          }
          const bool skip = (no_odd && (l & 1)) || (bad && l > 0);
          const double base = mass * lmc[kLmcY + lm] * pl;
          if (m == 0) {
            s_Y[(l * l) * YS + qp] = skip ? 0.f : (float)base;
          } else {
            s_Y[(l * l + 2 * m - 1) * YS + qp] = skip ? 0.f : (float)(base * ch);
            s_Y[(l * l + 2 * m) * YS + qp] = skip ? 0.f : (float)(base * sh);
          }
        }
        const double cn = ch * c4 - sh * s4;
        sh = sh * c4 + ch * s4;
        ch = cn;
        pw *= p4;
      }
    }
    __syncthreads();
    // radial functions phi_ln(r_p) for the tile, coalesced over ln
    for (int idx = tid; idx < kSLTile * LN; idx += 256) {
      const int q = idx / LN, ln = idx - q * LN;
      const size_t i = (size_t)s_i[q];
      s_phi[idx] = s_w1[q] * E[i * LN + ln] + s_w2[q] * E[(i + 1) * LN + ln];
    }
    __syncthreads();
    const int64_t left = n - t * kSLTile;
    const int np = left < kSLTile ? (int)left : kSLTile;
    for (int q = 0; q < np; ++q) {
#pragma unroll
      for (int o = 0; o < OPT; ++o) acc[o] = fmaf(s_Y[yrow[o] + q], s_phi[q * LN + prow[o]], acc[o]);
    }
#pragma unroll
    for (int o = 0; o < OPT; ++o) {
      accd[o] += (double)acc[o];
      acc[o] = 0.f;
    }
    __syncthreads();
  }
#pragma unroll
  for (int o = 0; o < OPT; ++o) {
    const int k = tid + o * 256;
    if (k < K) partials[(size_t)blockIdx.x * K + k] = accd[o];
  }
}

static int accumulate_blocks(int64_t ntiles) {
  const int64_t cap = 2048;
  return (int)(ntiles < cap ? (ntiles > 0 ? ntiles : 1) : cap);
}

}  // namespace bfe

using namespace bfe;

extern "C" int bfe_eof_accumulate(bfe_ctx* h, const float* x, const float* y, const float* z,
                                  const float* m, int64_t n, uint32_t flags, double* cos_out,
                                  double* sin_out, void* stream) {
  (void)flags;  // eof.accumulate computes a no_odd mask but never applies it (eof.py:545-551)
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (!c->eof_ok) return set_error(c, BFE_ENOTABLE, "eof_accumulate: EOF tables not set");
  if (n < 0) return set_error(c, BFE_EINVAL, "eof_accumulate: n < 0");
  if (!cos_out || !sin_out) return set_error(c, BFE_EINVAL, "eof_accumulate: NULL output");
  if (n > 0 && (!x || !y || !z || !m)) return set_error(c, BFE_EINVAL, "eof_accumulate: NULL input");
  BFE_HIP(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream;
  const int K = c->eof_K;
  const int64_t ntiles = (n + kTile - 1) / kTile;
  const int nb = accumulate_blocks(ntiles);
  const size_t pbytes = (size_t)nb * K * sizeof(double);
  int rc = ensure_scratch(c, pbytes + (size_t)K * sizeof(double));
  if (rc) return rc;
  double* partials = (double*)c->scratch;
  double* sums = partials + (size_t)nb * K;
  EOFGeom g = eof_geom(c);
  const int opt = (K + 255) / 256;
  if (n == 0) {
    BFE_HIP(c, hipMemsetAsync(partials, 0, pbytes, s));
  } else if (opt == 1) {
    hipLaunchKernelGGL(eof_accumulate_kernel<1>, dim3(nb), dim3(256), 0, s, x, y, z, m, n, g, c->eof_pot32, K, partials);
  } else if (opt == 2) {
    hipLaunchKernelGGL(eof_accumulate_kernel<2>, dim3(nb), dim3(256), 0, s, x, y, z, m, n, g, c->eof_pot32, K, partials);
  } else {
    hipLaunchKernelGGL(eof_accumulate_kernel<4>, dim3(nb), dim3(256), 0, s, x, y, z, m, n, g, c->eof_pot32, K, partials);
  }
  BFE_HIP(c, hipGetLastError());
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(K), dim3(256), 0, s, partials, nb, K, sums);
  const int nc = (c->mmax + 1) * c->norder;
  hipLaunchKernelGGL(eof_split_kernel, dim3((nc + 255) / 256), dim3(256), 0, s, sums, c->mmax, c->norder, cos_out, sin_out);
  BFE_HIP(c, hipGetLastError());
  return 0;
}

extern "C" int bfe_sl_accumulate(bfe_ctx* h, const float* x, const float* y, const float* z,
                                 const float* m, int64_t n, uint32_t flags, double* coef_out,
                                 void* stream) {
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (!c->sl_ok) return set_error(c, BFE_ENOTABLE, "sl_accumulate: SL tables not set");
  if (n < 0) return set_error(c, BFE_EINVAL, "sl_accumulate: n < 0");
  if (!coef_out) return set_error(c, BFE_EINVAL, "sl_accumulate: NULL output");
  if (n > 0 && (!x || !y || !z || !m)) return set_error(c, BFE_EINVAL, "sl_accumulate: NULL input");
  const int R = (c->lmax + 1) * (c->lmax + 1);
  const int LN = (c->lmax + 1) * c->nmax;
  const int K = R * c->nmax;
  if (K > 1024) return set_error(c, BFE_ELIMIT, "sl_accumulate: (lmax+1)^2 nmax = %d > 1024", K);
  const size_t lds = ((size_t)R * (kSLTile + 1) + (size_t)kSLTile * LN + 3 * kSLTile) * 4;
  if (lds > 64 * 1024) return set_error(c, BFE_ELIMIT, "sl_accumulate: LDS %zu B > 64 KiB", lds);
  BFE_HIP(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream;
  const int64_t ntiles = (n + kSLTile - 1) / kSLTile;
  const int nb = accumulate_blocks(ntiles);
  const size_t pbytes = (size_t)nb * K * sizeof(double);
  int rc = ensure_scratch(c, pbytes);
  if (rc) return rc;
  double* partials = (double*)c->scratch;
  SLGeom g = sl_geom(c);
  const int no_odd = (flags & BFE_NO_ODD) ? 1 : 0;
  const int opt = (K + 255) / 256;
  if (n == 0) {
    BFE_HIP(c, hipMemsetAsync(partials, 0, pbytes, s));
  } else if (opt == 1) {
    hipLaunchKernelGGL(sl_accumulate_kernel<1>, dim3(nb), dim3(256), lds, s, x, y, z, m, n, g, c->sl_e32, c->sl_lmc, no_odd, partials);
  } else if (opt == 2) {
    hipLaunchKernelGGL(sl_accumulate_kernel<2>, dim3(nb), dim3(256), lds, s, x, y, z, m, n, g, c->sl_e32, c->sl_lmc, no_odd, partials);
  } else {
    hipLaunchKernelGGL(sl_accumulate_kernel<4>, dim3(nb), dim3(256), lds, s, x, y, z, m, n, g, c->sl_e32, c->sl_lmc, no_odd, partials);
  }
  BFE_HIP(c, hipGetLastError());
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(K), dim3(256), 0, s, partials, nb, K, coef_out);
  BFE_HIP(c, hipGetLastError());
  return 0;
}
