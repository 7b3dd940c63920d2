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
#include <algorithm>
#include <vector>

#include "bfe_internal.hpp"

namespace bfe {

constexpr int kTile = 256;   // EOF: particles per tile (one per lane for geometry)
constexpr int kSLTile = 64;  // SL: particles per tile
constexpr int kFlush = 16;   // float32 partial sums are moved to float64 every kFlush particles

// ---------------------------------------------------------------------------
// EOF accumulate
// ---------------------------------------------------------------------------
// Work is split into segments of consecutive particles (segment sg covers
// [segb[sg], segb[sg + 1]), or [sg * seglen, min((sg + 1) * seglen, n)) when
// segb is NULL); each segment writes one float64 partial row, so a batch of
// snapshots whose particles are concatenated can share one launch.
struct Segments {
  const int64_t* segb;  // nseg + 1 boundaries, or NULL
  int64_t seglen;
  int64_t n;
  int64_t nseg;
};

__device__ __forceinline__ void segment_range(const Segments& S, int64_t sg, int64_t& p0,
                                              int64_t& p1) {
  if (S.segb) {
    p0 = S.segb[sg];
    p1 = S.segb[sg + 1];
  } else {
    p0 = sg * S.seglen;
    p1 = p0 + S.seglen < S.n ? p0 + S.seglen : S.n;
  }
}

template <int OPT>
__global__ __launch_bounds__(256) void eof_accumulate_kernel(
    const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
    const float* __restrict__ pm, Segments S, EOFGeom g, const float* __restrict__ T, int K,
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
    const int k = tid + o * 256;
    rowg[o] = (k < K) ? k / g.norder : 0;
  }
  for (int64_t sg = blockIdx.x; sg < S.nseg; sg += gridDim.x) {
  int64_t s0, s1;
  segment_range(S, sg, s0, s1);
#pragma unroll
  for (int o = 0; o < OPT; ++o) accd[o] = 0.0;
  for (int64_t t0 = s0; t0 < s1; t0 += kTile) {
    const int64_t p = t0 + tid;
    if (p < s1) {
      const double x = px[p], y = py[p], z = pz[p], mass = pm[p];
      const double R = sqrt(x * x + y * y + 1.e-10);
      double c1, s1;
      planar_angle(x, y, c1, s1);
      EOFCell cell = eof_locate(R, z, g.cmap, g.ascale, g.ihscale, g.xmin, g.idx, g.ymin, g.idy,
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
    const int64_t left = s1 - t0;
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
      // float32 sums over at most kFlush particles: for noise-dominated
      // coefficients (|a| ~ sqrt(N) |term|) the rounding of long float32
      // sums, not of the terms, sets the error
      if ((q & (kFlush - 1)) == kFlush - 1) {
#pragma unroll
        for (int o = 0; o < OPT; ++o) {
          accd[o] += (double)acc[o];
          acc[o] = 0.f;
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
    if (k < K) partials[(size_t)sg * K + k] = accd[o];
  }
  }
}

// Fixed-order float64 reduction of the per-segment partials, column k
// (a column's rows are K doubles apart): row lane j Kahan-sums rows j,
// j + kRedLanes, ... of its column, then a fixed tree adds the lanes.  The
// order depends only on the row count, so results are bitwise reproducible.
// With rowb, grid.y runs over snapshots: snapshot q sums rows
// [rowb[q], rowb[q+1]) into out + q * K.  SPLIT (the EOF sums): column k <
// nc = (mmax+1) norder goes to cos_out, the others to sin_out from row m = 1
// on, and the m = 0 sin row is zeroed, each set q at q * nc.
constexpr int kRedCols = 8;    // a row read is 64 contiguous bytes
constexpr int kRedLanes = 32;
template <bool SPLIT>
__global__ __launch_bounds__(256) void reduce_partials_kernel(const double* __restrict__ partials,
                                                              int nb, int K,
                                                              double* __restrict__ out,
                                                              const int64_t* __restrict__ rowb,
                                                              double* __restrict__ sin_out,
                                                              int nc, int norder) {
  __shared__ double sh[kRedLanes][kRedCols];
  const int c = threadIdx.x % kRedCols;
  const int j = threadIdx.x / kRedCols;
  const int k = blockIdx.x * kRedCols + c;
  int64_t r0 = 0, r1 = nb;
  if (rowb) {
    r0 = rowb[blockIdx.y];
    r1 = rowb[blockIdx.y + 1];
  }
  double s = 0.0, comp = 0.0;
  if (k < K) {
#pragma unroll 4
    for (int64_t b = r0 + j; b < r1; b += kRedLanes) {
      const double yv = partials[(size_t)b * K + k] - comp;
      const double tv = s + yv;
      comp = (tv - s) - yv;
      s = tv;
    }
  }
  sh[j][c] = s;
  __syncthreads();
  for (int w = kRedLanes / 2; w > 0; w >>= 1) {
    if (j < w) sh[j][c] += sh[j + w][c];
    __syncthreads();
  }
  if (j == 0 && k < K) {
    const double v = sh[0][c];
    if (!SPLIT) {
      out[(size_t)blockIdx.y * K + k] = v;
    } else {
      const size_t q = (size_t)blockIdx.y * nc;
      if (k < nc) {
        out[q + k] = v;
        if (k < norder) sin_out[q + k] = 0.0;
      } else {
        sin_out[q + (k - nc) + norder] = v;
      }
    }
  }
}

__global__ void fill_f64_kernel(double* __restrict__ a, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

// ---------------------------------------------------------------------------
// SL accumulate
// ---------------------------------------------------------------------------
template <int OPT>
__global__ __launch_bounds__(256) void sl_accumulate_kernel(
    const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
    const float* __restrict__ pm, Segments S, SLGeom g, const float* __restrict__ E,
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
    const int k = tid + o * 256;
    const int row = (k < K) ? k / g.nmax : 0;
    const int nn = (k < K) ? k % g.nmax : 0;
    int l = 0;
    while ((l + 1) * (l + 1) <= row) ++l;
    yrow[o] = row * YS;
    prow[o] = l * g.nmax + nn;
  }
  for (int64_t sg = blockIdx.x; sg < S.nseg; sg += gridDim.x) {
  int64_t s0, s1;
  segment_range(S, sg, s0, s1);
#pragma unroll
  for (int o = 0; o < OPT; ++o) accd[o] = 0.0;
  for (int64_t t0 = s0; t0 < s1; t0 += kSLTile) {
    const int64_t p = t0 + qp;
    {
      // padding particles: unit position, zero mass, so every Y row is 0
      const bool live = p < s1;
      const double x = live ? (double)px[p] : 1.0, y = live ? (double)py[p] : 0.0;
      const double z = live ? (double)pz[p] : 0.0, mass = live ? (double)pm[p] : 0.0;
      double r = sqrt(x * x + y * y + z * z);
      r = (r == r) ? (r > 1.0e-10 ? r : 1.0e-10) : 1.0e-10;   // np.nanmax([r, 1e-10])
      const double costh = z / r;
      double c1, s1;
      planar_angle(x, y, c1, s1);
      if (h == 0) {
        SLCell cell = sl_locate(r, g.cmap, g.scale, g.xi, g.numr, g.idxi);
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
    const int64_t left = s1 - t0;
    const int np = left < kSLTile ? (int)left : kSLTile;
    for (int q = 0; q < np; ++q) {
#pragma unroll
      for (int o = 0; o < OPT; ++o) acc[o] = fmaf(s_Y[yrow[o] + q], s_phi[q * LN + prow[o]], acc[o]);
      if ((q & (kFlush - 1)) == kFlush - 1) {
#pragma unroll
        for (int o = 0; o < OPT; ++o) {
          accd[o] += (double)acc[o];
          acc[o] = 0.f;
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
    if (k < K) partials[(size_t)sg * K + k] = accd[o];
  }
  }
}

// ---------------------------------------------------------------------------
// Cell-ordered accumulate (particles sorted by cell, packed as float4)
// ---------------------------------------------------------------------------
// Each wave works on its own 64-particle tiles of the block's segment: lane q
// computes particle q's geometry and writes its per-row weights, its cell and
// its cell weights to a wave-private LDS slab.  Lanes then own (row, chunk of
// CH radial orders) slots and walk the particles in order; as the particles
// are sorted by cell, consecutive particles share the table rows they need,
// so the sums are regrouped exactly per run of equal cells:
//   EOF: sum_p w_j(p) sum_c c_c(p) T_c[k] = sum_c T_c[k] (sum_p w_j(p) c_c(p))
//   SL : sum_p Y_row(p) (x1 P0 E[i] + x2 P0 E[i+1])[ln]
//        = E[i][ln] sum_p Y_row x1 P0 + E[i+1][ln] sum_p Y_row x2 P0
// The per-particle sums run in float32 over at most kFlush particles and are
// then added to float64 run sums; the run sums are applied to the table
// entries (float64) when the cell changes and at the end of the segment, so
// table rows are read once per run.  A run may continue across the wave's
// tiles.  Per-particle values are read from LDS at wave-uniform addresses
// (broadcast), four particles at a time; a group of four in the current cell
// takes a branch-free path.  At the end of a segment the four waves' float64
// sums are added in wave order.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int kDepStride = 68;     // LDS row stride (floats): 16-byte rows on distinct banks
constexpr int kDepExtra = 5 * 64;  // EOF: per-wave LDS after the rows (corner weights, cells)
// SL: rows of the per-wave LDS region (the padding of 48 rows holds three
// per-particle arrays)
__host__ __device__ constexpr int sl_dep_rows(int R) { return R > 48 ? R : 48; }

// CH consecutive floats of a padded deposit-table row (16-byte aligned for
// CH % 4 == 0, 8-byte for CH == 2).
template <int CH>
__device__ __forceinline__ void dep_row(const float* __restrict__ p, float (&v)[CH]) {
  if constexpr (CH % 4 == 0) {
#pragma unroll
    for (int k = 0; k < CH / 4; ++k) {
      const float4 q = reinterpret_cast<const float4*>(p)[k];
      v[4 * k] = q.x;
      v[4 * k + 1] = q.y;
      v[4 * k + 2] = q.z;
      v[4 * k + 3] = q.w;
    }
  } else {
    static_assert(CH == 2, "chunk of 2 or a multiple of 4");
    const float2 q = reinterpret_cast<const float2*>(p)[0];
    v[0] = q.x;
    v[1] = q.y;
  }
}

// Applies the float64 corner run sums GD of the cell at node base to the
// lane's coefficients and clears them.  T: padded table [node][J][np].
template <int PPT, int CH>
__device__ __forceinline__ void eof_apply(const float* __restrict__ T, int base, int rowlen,
                                          int ny1, int np, const int (&j_)[PPT],
                                          const int (&n0_)[PPT], double (&GD)[PPT][4],
                                          double (&acc)[PPT][CH]) {
#pragma unroll
  for (int o = 0; o < PPT; ++o) {
    if (j_[o] >= 0) {
      const float* T00 = T + (size_t)base * rowlen + j_[o] * np + n0_[o];
      float t00[CH], t10[CH], t01[CH], t11[CH];
      dep_row<CH>(T00, t00);
      dep_row<CH>(T00 + (size_t)ny1 * rowlen, t10);
      dep_row<CH>(T00 + rowlen, t01);
      dep_row<CH>(T00 + (size_t)(ny1 + 1) * rowlen, t11);
#pragma unroll
      for (int c = 0; c < CH; ++c)
        acc[o][c] += GD[o][0] * t00[c] + GD[o][1] * t10[c] + GD[o][2] * t01[c] + GD[o][3] * t11[c];
    }
    GD[o][0] = GD[o][1] = GD[o][2] = GD[o][3] = 0.0;
  }
}

template <int PPT>
__device__ __forceinline__ void eof_to_run(float (&G)[PPT][4], double (&GD)[PPT][4]) {
#pragma unroll
  for (int o = 0; o < PPT; ++o)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      GD[o][k] += (double)G[o][k];
      G[o][k] = 0.f;
    }
}

// Per-particle part of the cell-ordered EOF accumulates: lane's particle p
// of the sorted input gives its row weights -4 pi m {1, cos k phi, sin k phi}
// (rows 0, k, mmax + k of sW), its four bilinear corner weights (sE, 64 per
// corner: c00, c10, c01, c11) and its cell's first node (sB, also returned).
__device__ __forceinline__ int eof_dep_setup(const float4* __restrict__ P, int64_t p, bool live,
                                             const EOFGeom& g, int ny1, int lane, float* sW,
                                             float* sE, int* sB) {
  const float4 q4 = live ? P[p] : make_float4(1.f, 0.f, 0.f, 0.f);
  const double x = q4.x, y = q4.y, z = q4.z, mass = live ? (double)q4.w : 0.0;
  const double R = sqrt(x * x + y * y + 1.e-10);
  double c1, s1v;
  planar_angle(x, y, c1, s1v);
  const EOFCell cell = eof_locate(R, z, g.cmap, g.ascale, g.ihscale, g.xmin, g.idx, g.ymin,
                                  g.idy, g.numx, g.numy);
  const double w = -kFourPi * mass;
  double cm = 1.0, sm = 0.0, cm1 = c1, sm1 = -s1v;
  sW[lane] = (float)w;
  sE[lane] = (float)cell.c00;
  sE[64 + lane] = (float)cell.c10;
  sE[128 + lane] = (float)cell.c01;
  sE[192 + lane] = (float)cell.c11;
  const int node = cell.ix * ny1 + cell.iy;
  sB[lane] = node;
  for (int m = 1; m <= g.mmax; ++m) {
    const double cn = 2.0 * c1 * cm - cm1;   // m = 1 gives c1, s1 exactly
    const double sn = 2.0 * c1 * sm - sm1;
    cm1 = cm; sm1 = sm; cm = cn; sm = sn;
    sW[m * kDepStride + lane] = (float)(w * cm);
    sW[(g.mmax + m) * kDepStride + lane] = (float)(w * sm);
  }
  return node;
}

// Four particles' row weights (per owned row) times their four corner
// weights, summed into the float32 corner sums in particle order.
template <int PPT>
__device__ __forceinline__ void eof_group(const float4 (&wv)[PPT], float4 e00, float4 e10,
                                          float4 e01, float4 e11, float (&G)[PPT][4]) {
#pragma unroll
  for (int o = 0; o < PPT; ++o) {
    G[o][0] = fmaf(wv[o].x, e00.x, G[o][0]);
    G[o][1] = fmaf(wv[o].x, e10.x, G[o][1]);
    G[o][2] = fmaf(wv[o].x, e01.x, G[o][2]);
    G[o][3] = fmaf(wv[o].x, e11.x, G[o][3]);
    G[o][0] = fmaf(wv[o].y, e00.y, G[o][0]);
    G[o][1] = fmaf(wv[o].y, e10.y, G[o][1]);
    G[o][2] = fmaf(wv[o].y, e01.y, G[o][2]);
    G[o][3] = fmaf(wv[o].y, e11.y, G[o][3]);
    G[o][0] = fmaf(wv[o].z, e00.z, G[o][0]);
    G[o][1] = fmaf(wv[o].z, e10.z, G[o][1]);
    G[o][2] = fmaf(wv[o].z, e01.z, G[o][2]);
    G[o][3] = fmaf(wv[o].z, e11.z, G[o][3]);
    G[o][0] = fmaf(wv[o].w, e00.w, G[o][0]);
    G[o][1] = fmaf(wv[o].w, e10.w, G[o][1]);
    G[o][2] = fmaf(wv[o].w, e01.w, G[o][2]);
    G[o][3] = fmaf(wv[o].w, e11.w, G[o][3]);
  }
}

// Body of the cell-ordered EOF accumulate; dsm is the block's dynamic LDS.
template <int PPT, int CH>
__device__ __forceinline__ void eof_deposit_body(
    unsigned char* dsm, const float4* __restrict__ P, Segments S, EOFGeom g,
    const float* __restrict__ T, int rp, int K, double* __restrict__ partials) {
  // rp: padded row length of T, norder rounded up to a multiple of CH
  const int rowlen = (2 * g.mmax + 1) * rp;
  const int J = 2 * g.mmax + 1;
  const int N = g.norder;
  const int NC = (N + CH - 1) / CH;
  const int ny1 = g.numy + 1;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // per wave: row weights [J][stride], then corner weights [4][64] and cells [64]
  float* sW = reinterpret_cast<float*>(dsm) + (size_t)wave * (J * kDepStride + kDepExtra);
  float* sE = sW + J * kDepStride;
  int* sB = reinterpret_cast<int*>(sE + 4 * 64);
  double* sRed = reinterpret_cast<double*>(dsm);  // [4][K] after the tiles
  int j_[PPT], n0_[PPT];
#pragma unroll
  for (int o = 0; o < PPT; ++o) {
    const int pr = lane + 64 * o;
    j_[o] = pr / NC < J ? pr / NC : -1;
    n0_[o] = (pr % NC) * CH;
  }
  double acc[PPT][CH];
  for (int64_t sg = blockIdx.x; sg < S.nseg; sg += gridDim.x) {
    int64_t s0, s1;
    segment_range(S, sg, s0, s1);
    float G[PPT][4];
    double GD[PPT][4];
#pragma unroll
    for (int o = 0; o < PPT; ++o) {
#pragma unroll
      for (int c = 0; c < CH; ++c) acc[o][c] = 0.0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        G[o][k] = 0.f;
        GD[o][k] = 0.0;
      }
    }
    int cur = -1, cnt = 0;
    for (int64_t t0 = s0 + 64 * wave; t0 < s1; t0 += 256) {
      const int64_t p = t0 + lane;
      const bool live = p < s1;
      const int node = eof_dep_setup(P, p, live, g, ny1, lane, sW, sE, sB);
      wave_sync();
      const int64_t left = s1 - t0;
      const int np = left < 64 ? (int)left : 64;
      if (cur < 0) cur = __builtin_amdgcn_readfirstlane(node);
      // the tile's leading particles in the current cell (a prefix, the
      // input being sorted by cell), eight at a time with their LDS reads
      // issued together and no per-group cell check
      const uint64_t other = __ballot(live && node != cur);
      const int pre = other ? (int)__builtin_ctzll(other) : np;
      int q0 = 0;
#pragma unroll 1
      for (; q0 + 8 <= pre; q0 += 8) {
        const float4 a00 = *reinterpret_cast<const float4*>(sE + q0);
        const float4 a10 = *reinterpret_cast<const float4*>(sE + 64 + q0);
        const float4 a01 = *reinterpret_cast<const float4*>(sE + 128 + q0);
        const float4 a11 = *reinterpret_cast<const float4*>(sE + 192 + q0);
        const float4 b00 = *reinterpret_cast<const float4*>(sE + q0 + 4);
        const float4 b10 = *reinterpret_cast<const float4*>(sE + 68 + q0);
        const float4 b01 = *reinterpret_cast<const float4*>(sE + 132 + q0);
        const float4 b11 = *reinterpret_cast<const float4*>(sE + 196 + q0);
        float4 w0[PPT], w1[PPT];
#pragma unroll
        for (int o = 0; o < PPT; ++o) {
          const float* wr = sW + (j_[o] < 0 ? 0 : j_[o]) * kDepStride + q0;
          w0[o] = *reinterpret_cast<const float4*>(wr);
          w1[o] = *reinterpret_cast<const float4*>(wr + 4);
        }
        if (cnt + 4 > kFlush) {
          eof_to_run<PPT>(G, GD);
          cnt = 0;
        }
        eof_group<PPT>(w0, a00, a10, a01, a11, G);
        cnt += 4;
        if (cnt + 4 > kFlush) {
          eof_to_run<PPT>(G, GD);
          cnt = 0;
        }
        eof_group<PPT>(w1, b00, b10, b01, b11, G);
        cnt += 4;
      }
#pragma unroll 1
      for (; q0 < np; q0 += 4) {
        const int4 bb = *reinterpret_cast<const int4*>(sB + q0);
        const bool run = q0 + 4 <= np && bb.x == cur && bb.y == cur && bb.z == cur && bb.w == cur;
        if (__builtin_amdgcn_readfirstlane((int)run)) {
          if (cnt + 4 > kFlush) {
            eof_to_run<PPT>(G, GD);
            cnt = 0;
          }
          const float4 e00 = *reinterpret_cast<const float4*>(sE + q0);
          const float4 e10 = *reinterpret_cast<const float4*>(sE + 64 + q0);
          const float4 e01 = *reinterpret_cast<const float4*>(sE + 128 + q0);
          const float4 e11 = *reinterpret_cast<const float4*>(sE + 192 + q0);
          float4 wv[PPT];
#pragma unroll
          for (int o = 0; o < PPT; ++o)
            wv[o] = *reinterpret_cast<const float4*>(sW + (j_[o] < 0 ? 0 : j_[o]) * kDepStride + q0);
          eof_group<PPT>(wv, e00, e10, e01, e11, G);
          cnt += 4;
          continue;
        }
        // one particle at a time: cell changes
#pragma unroll 1
        for (int q = q0; q < q0 + 4 && q < np; ++q) {
          const int b = sB[q];
          if (b != cur) {
            eof_to_run<PPT>(G, GD);
            eof_apply<PPT, CH>(T, cur, rowlen, ny1, rp, j_, n0_, GD, acc);
            cur = b;
            cnt = 0;
          } else if (cnt == kFlush) {
            eof_to_run<PPT>(G, GD);
            cnt = 0;
          }
          const float f00 = sE[q], f10 = sE[64 + q], f01 = sE[128 + q], f11 = sE[192 + q];
#pragma unroll
          for (int o = 0; o < PPT; ++o) {
            const float w = sW[(j_[o] < 0 ? 0 : j_[o]) * kDepStride + q];
            G[o][0] = fmaf(w, f00, G[o][0]);
            G[o][1] = fmaf(w, f10, G[o][1]);
            G[o][2] = fmaf(w, f01, G[o][2]);
            G[o][3] = fmaf(w, f11, G[o][3]);
          }
          ++cnt;
        }
      }
      wave_sync();
    }
    if (cur >= 0) {
      eof_to_run<PPT>(G, GD);
      eof_apply<PPT, CH>(T, cur, rowlen, ny1, rp, j_, n0_, GD, acc);
    }
    // fixed-order sum of the four waves' partial sums
    __syncthreads();
#pragma unroll
    for (int o = 0; o < PPT; ++o)
      if (j_[o] >= 0)
#pragma unroll
        for (int c = 0; c < CH; ++c)
          if (n0_[o] + c < N) sRed[(size_t)wave * K + j_[o] * N + n0_[o] + c] = acc[o][c];
    __syncthreads();
    for (int k = threadIdx.x; k < K; k += 256)
      partials[(size_t)sg * K + k] = ((sRed[k] + sRed[K + k]) + sRed[2 * K + k]) + sRed[3 * K + k];
    __syncthreads();
  }
}

// As for the SL deposit: eof_deposit_kernel beside the SL accumulate (the
// fused calls), eof_deposit_alone_kernel when it has the GPU to itself.
// Alone, four waves per SIMD (128 registers, about 100 B spilled) measured
// 80 against 83 us for 1M disc particles; beside the SL deposit they lost.
// Beside it, the one-slot shapes up to 8 orders are held at three waves
// (168 registers, a few bytes spilled): unconstrained they took 169 and
// fell to two, which made the fused call slower; the wider shapes keep
// their natural register counts.
template <int PPT, int CH>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(PPT == 1 && CH <= 8 ? 3 : 1))) void eof_deposit_kernel(
    const float4* __restrict__ P, Segments S, EOFGeom g, const float* __restrict__ T, int rp,
    int K, double* __restrict__ partials) {
  extern __shared__ __align__(16) unsigned char dsm[];
  eof_deposit_body<PPT, CH>(dsm, P, S, g, T, rp, K, partials);
}

template <int PPT, int CH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void
eof_deposit_alone_kernel(const float4* __restrict__ P, Segments S, EOFGeom g,
                         const float* __restrict__ T, int rp, int K,
                         double* __restrict__ partials) {
  extern __shared__ __align__(16) unsigned char dsm[];
  eof_deposit_body<PPT, CH>(dsm, P, S, g, T, rp, K, partials);
}

// eof_dep32[node][j][n < norder] = eof_pot32[node][j norder + n], zero for
// n in [norder, rp): one thread per (node, j, n).
__global__ __launch_bounds__(256) void eof_repad_kernel(const float* __restrict__ src,
                                                        float* __restrict__ dst, int64_t rows,
                                                        int norder, int rp) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows * rp) return;
  const int64_t row = t / rp;
  const int n = (int)(t - row * rp);
  dst[t] = n < norder ? src[row * norder + n] : 0.f;
}

// Applies the float64 radial run sums D1, D2 of cell i to the lane's
// coefficients and clears them.  E: padded table [node][lmax + 1][np].
template <int PPT, int CH>
__device__ __forceinline__ void sl_apply(const float* __restrict__ E, int i, int rowlen, int np,
                                         const int (&row_)[PPT], const int (&l_)[PPT],
                                         const int (&n0_)[PPT], double (&D1)[PPT],
                                         double (&D2)[PPT], double (&acc)[PPT][CH]) {
#pragma unroll
  for (int o = 0; o < PPT; ++o) {
    if (row_[o] >= 0) {
      const float* e1 = E + (size_t)i * rowlen + l_[o] * np + n0_[o];
      float a[CH], b[CH];
      dep_row<CH>(e1, a);
      dep_row<CH>(e1 + rowlen, b);
#pragma unroll
      for (int c = 0; c < CH; ++c) acc[o][c] += D1[o] * a[c] + D2[o] * b[c];
    }
    D1[o] = D2[o] = 0.0;
  }
}

// Four particles' Y values (per owned row) times their two radial weights,
// summed into the float32 run sums in particle order.
template <int PPT>
__device__ __forceinline__ void sl_group(const float4 (&yv)[PPT], float4 wa, float4 wb,
                                         float (&A1)[PPT], float (&A2)[PPT]) {
#pragma unroll
  for (int o = 0; o < PPT; ++o) {
    A1[o] = fmaf(yv[o].x, wa.x, A1[o]);
    A2[o] = fmaf(yv[o].x, wb.x, A2[o]);
    A1[o] = fmaf(yv[o].y, wa.y, A1[o]);
    A2[o] = fmaf(yv[o].y, wb.y, A2[o]);
    A1[o] = fmaf(yv[o].z, wa.z, A1[o]);
    A2[o] = fmaf(yv[o].z, wb.z, A2[o]);
    A1[o] = fmaf(yv[o].w, wa.w, A1[o]);
    A2[o] = fmaf(yv[o].w, wb.w, A2[o]);
  }
}

template <int PPT>
__device__ __forceinline__ void sl_to_run(float (&A1)[PPT], float (&A2)[PPT], double (&D1)[PPT],
                                          double (&D2)[PPT]) {
#pragma unroll
  for (int o = 0; o < PPT; ++o) {
    D1[o] += (double)A1[o];
    D2[o] += (double)A2[o];
    A1[o] = A2[o] = 0.f;
  }
}

// Body of the cell-ordered SL accumulate; dsm is the block's dynamic LDS.
// LM > 0: g.lmax == LM, known at compile time; LM = 0: any lmax.
template <int PPT, int CH, int LM>
__device__ __forceinline__ void sl_deposit_body(
    unsigned char* dsm, const float4* __restrict__ P, Segments S, SLGeom g,
    const float* __restrict__ E, const double* __restrict__ lmc, int no_odd,
    double* __restrict__ partials) {
  const int R = (g.lmax + 1) * (g.lmax + 1);
  const int K = R * g.nmax;
  const int N = g.nmax;
  const int NC = (N + CH - 1) / CH;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // per wave: Y rows [max(R, 48)][stride]; the per-particle radial weights
  // and cells live in the row padding (columns 64..67): value q of array k
  // at row 16 k + q / 4, column 64 + q % 4, so four particles are one float4
  float* sY = reinterpret_cast<float*>(dsm) + (size_t)wave * sl_dep_rows(R) * kDepStride;
  float* sP = sY + 64;
  double* sRed = reinterpret_cast<double*>(dsm);  // [4][K] after the tiles
  const int rp = dep_np(N);  // padded row length of E
  const int rowlen = (g.lmax + 1) * rp;
  int row_[PPT], l_[PPT], n0_[PPT];
#pragma unroll
  for (int o = 0; o < PPT; ++o) {
    const int pr = lane + 64 * o;
    const int row = pr / NC;
    row_[o] = row < R ? row : -1;
    int l = 0;
    while ((l + 1) * (l + 1) <= row) ++l;
    n0_[o] = (pr % NC) * CH;
    l_[o] = l;
  }
  double acc[PPT][CH];
  for (int64_t sg = blockIdx.x; sg < S.nseg; sg += gridDim.x) {
    int64_t s0, s1;
    segment_range(S, sg, s0, s1);
    float A1[PPT], A2[PPT];
    double D1[PPT], D2[PPT];
#pragma unroll
    for (int o = 0; o < PPT; ++o) {
#pragma unroll
      for (int c = 0; c < CH; ++c) acc[o][c] = 0.0;
      A1[o] = A2[o] = 0.f;
      D1[o] = D2[o] = 0.0;
    }
    int cur = -1, cnt = 0;
    for (int64_t t0 = s0 + 64 * wave; t0 < s1; t0 += 256) {
      const int64_t p = t0 + lane;
      const bool live = p < s1;
      int ci = 0;
      float w1 = 0.f, w2 = 0.f;
      {
        const float4 q4 = live ? P[p] : make_float4(1.f, 0.f, 0.f, 0.f);
        const double x = q4.x, y = q4.y, z = q4.z, mass = live ? (double)q4.w : 0.0;
        double r = sqrt(x * x + y * y + z * z);
        r = (r == r) ? (r > 1.0e-10 ? r : 1.0e-10) : 1.0e-10;   // np.nanmax([r, 1e-10])
        const double costh = z / r;
        double c1, s1v;
        planar_angle(x, y, c1, s1v);
        const SLCell cell = sl_locate(r, g.cmap, g.scale, g.xi, g.numr, g.idxi);
        const double P0 = cell.x1 * g.p0[cell.i] + cell.x2 * g.p0[cell.i + 1];
        ci = cell.i;
        w1 = live ? (float)(cell.x1 * P0) : 0.f;
        w2 = live ? (float)(cell.x2 * P0) : 0.f;
        const int pad = (lane >> 2) * kDepStride + (lane & 3);
        sP[pad] = w1;
        sP[16 * kDepStride + pad] = w2;
        reinterpret_cast<int*>(sP)[32 * kDepStride + pad] = ci;
        // Y rows: Q_l^m = -4 pi N_lm P_l^m by its recurrence in l from the
        // closed form at l = m (legendre_R, spheresl.py:678-696, with the
        // normalisation folded into the constants), in float64; then one
        // float32 rounding of Q and of m_p {1, cos m phi, sin m phi} (the
        // Chebyshev recurrence, float64) and a float32 product.  Each
        // rounding is to nearest, so the sums over particles see no bias.
        const bool bad = !(costh == costh);
        const double somx2 = sqrt((1.0 - costh) * (1.0 + costh));
        double pw = 1.0, cmm = 1.0, smm = 0.0, cmm1 = c1, smm1 = -s1v;
        auto yrows = [&](int m, int lmax) {
          if (m > 0) {
            pw *= somx2;
            const double cn = (m == 1) ? c1 : 2.0 * c1 * cmm - cmm1;
            const double sn = (m == 1) ? s1v : 2.0 * c1 * smm - smm1;
            cmm1 = cmm; smm1 = smm; cmm = cn; smm = sn;
          }
          const float wc = (float)(mass * cmm), ws = (float)(mass * smm);
          double plm1 = 0.0, pl = lmc[kLmcN0 + m] * pw;
          for (int l = m; l <= lmax; ++l) {
            const int lm = lm_index(l, m);
            if (l > m) {
              const double pn = costh * lmc[kLmcNA + lm] * pl - lmc[kLmcNB + lm] * plm1;
              plm1 = pl;
              pl = pn;
            }
            const bool skip = (no_odd && (l & 1)) || (bad && l > 0);
            const float q = (float)pl;
            if (m == 0) {
              sY[(l * l) * kDepStride + lane] = skip ? 0.f : q * wc;
            } else {
              sY[(l * l + 2 * m - 1) * kDepStride + lane] = skip ? 0.f : q * wc;
              sY[(l * l + 2 * m) * kDepStride + lane] = skip ? 0.f : q * ws;
            }
          }
        };
        // With lmax known at compile time the loops unroll and the
        // recurrence constants sit at fixed offsets, so they are loaded once
        // per block, before the tile loop, instead of as two dependent
        // scalar loads per (l, m) and particle tile, each a wait that also
        // drains the wave's LDS stores: 74 -> 62 us alone for 1M particles
        // at lmax 6, nmax 20 (same sums, bitwise).
        if constexpr (LM > 0) {
#pragma unroll
          for (int m = 0; m <= LM; ++m) yrows(m, LM);
        } else {
          for (int m = 0; m <= g.lmax; ++m) yrows(m, g.lmax);
        }
      }
      wave_sync();
      const int64_t left = s1 - t0;
      const int np = left < 64 ? (int)left : 64;
      if (cur < 0) cur = __builtin_amdgcn_readfirstlane(ci);
      // The particles of the tile still in the current cell are a prefix
      // (the input is sorted by cell): they take eight at a time with all
      // their LDS reads issued together and no per-group cell check.
      const uint64_t other = __ballot(live && ci != cur);
      const int pre = other ? (int)__builtin_ctzll(other) : np;
      int q0 = 0;
#pragma unroll 1
      for (; q0 + 8 <= pre; q0 += 8) {
        const float* pq = sP + (q0 >> 2) * kDepStride;
        const float4 wa0 = *reinterpret_cast<const float4*>(pq);
        const float4 wb0 = *reinterpret_cast<const float4*>(pq + 16 * kDepStride);
        const float4 wa1 = *reinterpret_cast<const float4*>(pq + kDepStride);
        const float4 wb1 = *reinterpret_cast<const float4*>(pq + 17 * kDepStride);
        float4 y0[PPT], y1[PPT];
#pragma unroll
        for (int o = 0; o < PPT; ++o) {
          const float* yr = sY + (row_[o] < 0 ? 0 : row_[o]) * kDepStride + q0;
          y0[o] = *reinterpret_cast<const float4*>(yr);
          y1[o] = *reinterpret_cast<const float4*>(yr + 4);
        }
        if (cnt + 4 > kFlush) {
          sl_to_run<PPT>(A1, A2, D1, D2);
          cnt = 0;
        }
        sl_group<PPT>(y0, wa0, wb0, A1, A2);
        cnt += 4;
        if (cnt + 4 > kFlush) {
          sl_to_run<PPT>(A1, A2, D1, D2);
          cnt = 0;
        }
        sl_group<PPT>(y1, wa1, wb1, A1, A2);
        cnt += 4;
      }
#pragma unroll 1
      for (; q0 < np; q0 += 4) {
        const float* pq = sP + (q0 >> 2) * kDepStride;
        const int4 cc = *reinterpret_cast<const int4*>(pq + 32 * kDepStride);
        const bool run = q0 + 4 <= np && cc.x == cur && cc.y == cur && cc.z == cur && cc.w == cur;
        if (__builtin_amdgcn_readfirstlane((int)run)) {
          if (cnt + 4 > kFlush) {
            sl_to_run<PPT>(A1, A2, D1, D2);
            cnt = 0;
          }
          const float4 wa = *reinterpret_cast<const float4*>(pq);
          const float4 wb = *reinterpret_cast<const float4*>(pq + 16 * kDepStride);
          float4 yv[PPT];
#pragma unroll
          for (int o = 0; o < PPT; ++o)
            yv[o] = *reinterpret_cast<const float4*>(sY + (row_[o] < 0 ? 0 : row_[o]) * kDepStride + q0);
          sl_group<PPT>(yv, wa, wb, A1, A2);
          cnt += 4;
          continue;
        }
        // one particle at a time: cell changes
#pragma unroll 1
        for (int q = q0; q < q0 + 4 && q < np; ++q) {
          const float* pj = sP + (q >> 2) * kDepStride + (q & 3);
          const int i = reinterpret_cast<const int*>(pj)[32 * kDepStride];
          if (i != cur) {
            sl_to_run<PPT>(A1, A2, D1, D2);
            sl_apply<PPT, CH>(E, cur, rowlen, rp, row_, l_, n0_, D1, D2, acc);
            cur = i;
            cnt = 0;
          } else if (cnt == kFlush) {
            sl_to_run<PPT>(A1, A2, D1, D2);
            cnt = 0;
          }
          const float a = pj[0], b = pj[16 * kDepStride];
#pragma unroll
          for (int o = 0; o < PPT; ++o) {
            const float yq = sY[(row_[o] < 0 ? 0 : row_[o]) * kDepStride + q];
            A1[o] = fmaf(yq, a, A1[o]);
            A2[o] = fmaf(yq, b, A2[o]);
          }
          ++cnt;
        }
      }
      wave_sync();
    }
    if (cur >= 0) {
      sl_to_run<PPT>(A1, A2, D1, D2);
      sl_apply<PPT, CH>(E, cur, rowlen, rp, row_, l_, n0_, D1, D2, acc);
    }
    __syncthreads();
#pragma unroll
    for (int o = 0; o < PPT; ++o)
      if (row_[o] >= 0)
#pragma unroll
        for (int c = 0; c < CH; ++c)
          if (n0_[o] + c < N) sRed[(size_t)wave * K + row_[o] * N + n0_[o] + c] = acc[o][c];
    __syncthreads();
    for (int k = threadIdx.x; k < K; k += 256)
      partials[(size_t)sg * K + k] = ((sRed[k] + sRed[K + k]) + sRed[2 * K + k]) + sRed[3 * K + k];
    __syncthreads();
  }
}

// Two instantiations of the body: sl_deposit_kernel when the SL accumulate
// runs beside the EOF one (the fused calls), sl_deposit_alone_kernel when it
// has the GPU to itself.  Alone, three waves per SIMD (at most 168 registers,
// a few spilled) measured 85 against 104 us for 1M particles at lmax 6,
// nmax 20; beside the EOF accumulate the two-wave form gave the shorter call.
template <int PPT, int CH, int LM>
__global__ __launch_bounds__(256) void sl_deposit_kernel(
    const float4* __restrict__ P, Segments S, SLGeom g, const float* __restrict__ E,
    const double* __restrict__ lmc, int no_odd, double* __restrict__ partials) {
  extern __shared__ __align__(16) unsigned char dsm[];
  sl_deposit_body<PPT, CH, LM>(dsm, P, S, g, E, lmc, no_odd, partials);
}

template <int PPT, int CH, int LM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void
sl_deposit_alone_kernel(const float4* __restrict__ P, Segments S, SLGeom g,
                        const float* __restrict__ E, const double* __restrict__ lmc, int no_odd,
                        double* __restrict__ partials) {
  extern __shared__ __align__(16) unsigned char dsm[];
  sl_deposit_body<PPT, CH, LM>(dsm, P, S, g, E, lmc, no_odd, partials);
}

// ---------------------------------------------------------------------------
// SL accumulate as a matrix product on the bf16 MFMA units
// ---------------------------------------------------------------------------
// For each l the tile contribution is C_l[row][n] += sum_p Y_row(p) Phi_ln(p),
// a (2l+1) x 64 by 64 x nmax product.  Operands are split into NP bf16 pieces
// x = x_0 + x_1 (+ x_2) (each the bf16 rounding of the remainder) and the
// products x_a y_b with a + b < NP are accumulated on
// v_mfma_f32_16x16x32_bf16 (smallest first): NP = 3 keeps about 24 bits of
// each product (float32-class), NP = 2 about 16 bits (the bf16 mode).  Sums
// are float32 inside the MFMA over 32 particles, then float64.
//   A (16 x 32): A[c][8g + j] = Y_{l^2 + c}(particle 32 s + 8 g + j), c < 2l+1
//   B (32 x 16): B[8g + j][c] = Phi_{l, 16 nt + c}(particle 32 s + 8 g + j)
//   D (16 x 16): D[4g + i][c]
// Y rows are built per particle (four threads by m column) and stored as
// bf16 pieces in LDS; each lane computes its own B entries from the radial
// table (every Phi value is used by exactly one output tile).
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));

constexpr int kMT = 64;        // particles per tile (two MFMA K-steps)
constexpr int kMS = 72;        // LDS row stride (bf16): 144 B rows spread over banks
constexpr int kMSlots = 4;     // output tiles per wave

__device__ __forceinline__ uint16_t bf16_bits(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_float(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

template <int NP>
__device__ __forceinline__ void bf16_split(float x, uint16_t (&p)[NP]) {
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    p[k] = bf16_bits(x);
    x -= bf16_float(p[k]);
  }
}

template <int NP>
__global__ __launch_bounds__(256) void sl_mfma_kernel(
    const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
    const float* __restrict__ pm, Segments S, SLGeom g, const float* __restrict__ E,
    const double* __restrict__ lmc, int no_odd, double* __restrict__ partials) {
  extern __shared__ __align__(16) unsigned char msm[];
  const int L = g.lmax, R = (L + 1) * (L + 1), N = g.nmax, LN = (L + 1) * N;
  const int K = R * N;
  const int NT = (N + 15) / 16;
  const int ntiles = (L + 1) * NT;
  uint16_t* sY = reinterpret_cast<uint16_t*>(msm);                  // [NP][R][kMS]
  int* s_i = reinterpret_cast<int*>(msm + (size_t)NP * R * kMS * 2);   // [kMT]
  float* s_w1 = reinterpret_cast<float*>(s_i + kMT);
  float* s_w2 = s_w1 + kMT;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6, gq = lane >> 4, cc = lane & 15;
  const int qp = tid & (kMT - 1), h = tid >> 6;
  double accd[kMSlots][4];
  for (int64_t sg = blockIdx.x; sg < S.nseg; sg += gridDim.x) {
    int64_t s0, s1;
    segment_range(S, sg, s0, s1);
#pragma unroll
    for (int t = 0; t < kMSlots; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) accd[t][i] = 0.0;
    for (int64_t t0 = s0; t0 < s1; t0 += kMT) {
      // ---- geometry: Y rows (bf16 pieces), radial cell and weights ----------
      {
        const int64_t p = t0 + qp;
        const bool live = p < s1;
        const double x = live ? (double)px[p] : 1.0, y = live ? (double)py[p] : 0.0;
        const double z = live ? (double)pz[p] : 0.0, mass = live ? (double)pm[p] : 0.0;
        double r = sqrt(x * x + y * y + z * z);
        r = (r == r) ? (r > 1.0e-10 ? r : 1.0e-10) : 1.0e-10;   // np.nanmax([r, 1e-10])
        const double costh = z / r;
        double c1, s1v;
        planar_angle(x, y, c1, s1v);
        if (h == 0) {
          SLCell cell = sl_locate(r, g.cmap, g.scale, g.xi, g.numr, g.idxi);
          const double P0 = cell.x1 * g.p0[cell.i] + cell.x2 * g.p0[cell.i + 1];
          s_i[qp] = cell.i;
          s_w1[qp] = live ? (float)(cell.x1 * P0) : 0.f;
          s_w2[qp] = live ? (float)(cell.x2 * P0) : 0.f;
        }
        const bool bad = !(costh == costh);
        const double somx2 = sqrt((1.0 - costh) * (1.0 + costh));
        double ch = 1.0, sh = 0.0, pw = 1.0;
        for (int j = 0; j < h; ++j) {
          const double cn = ch * c1 - sh * s1v;
          sh = sh * c1 + ch * s1v;
          ch = cn;
          pw *= somx2;
        }
        const double c2 = c1 * c1 - s1v * s1v, s2 = 2.0 * c1 * s1v;
        const double c4 = c2 * c2 - s2 * s2, s4 = 2.0 * c2 * s2;
        const double sq2 = somx2 * somx2, p4 = sq2 * sq2;
        for (int m = h; m <= L; m += 4) {
          double plm1 = 0.0, pl = lmc[kLmcD + m] * pw;
          for (int l = m; l <= L; ++l) {
            const int lm = lm_index(l, m);
            if (l > m) {
              const double pn = costh * lmc[kLmcA + lm] * pl - lmc[kLmcB + lm] * plm1;
              plm1 = pl;
              pl = pn;
            }
            const bool skip = (no_odd && (l & 1)) || (bad && l > 0);
            const double base = mass * lmc[kLmcY + lm] * pl;
            const int nrow = m == 0 ? 1 : 2;
            for (int k = 0; k < nrow; ++k) {
              const int row = m == 0 ? l * l : l * l + 2 * m - 1 + k;
              const float v = skip ? 0.f : (float)(m == 0 ? base : base * (k ? sh : ch));
              uint16_t pc[NP];
              bf16_split<NP>(v, pc);
#pragma unroll
              for (int q = 0; q < NP; ++q) sY[((size_t)q * R + row) * kMS + qp] = pc[q];
            }
          }
          const double cn = ch * c4 - sh * s4;
          sh = sh * c4 + ch * s4;
          ch = cn;
          pw *= p4;
        }
      }
      __syncthreads();
      // ---- products: wave w takes output tiles w, w + 4, ... ------------------
#pragma unroll
      for (int slot = 0; slot < kMSlots; ++slot) {
        const int t = wave + 4 * slot;
        if (t >= ntiles) break;
        const int l = t / NT, nt = t - l * NT;
        const bool arow = cc < 2 * l + 1;
        const int n = 16 * nt + cc;
        const bool bcol = n < N;
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
          const int q0 = 32 * ks + 8 * gq;
          bf16x8_t a[NP], b[NP];
#pragma unroll
          for (int q = 0; q < NP; ++q) {
            if (arow) {
              a[q] = *reinterpret_cast<const bf16x8_t*>(sY + ((size_t)q * R + l * l + cc) * kMS + q0);
            } else {
#pragma unroll
              for (int j = 0; j < 8; ++j) a[q][j] = (__bf16)0.f;
            }
          }
          uint16_t bp[NP][8];
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            float phi = 0.f;
            if (bcol) {
              const int qq = q0 + j;
              const float* e = E + (size_t)s_i[qq] * LN + l * N + n;
              phi = s_w1[qq] * e[0] + s_w2[qq] * e[LN];
            }
            uint16_t pc[NP];
            bf16_split<NP>(phi, pc);
#pragma unroll
            for (int q = 0; q < NP; ++q) bp[q][j] = pc[q];
          }
#pragma unroll
          for (int q = 0; q < NP; ++q)
#pragma unroll
            for (int j = 0; j < 8; ++j) b[q][j] = __builtin_bit_cast(__bf16, bp[q][j]);
          floatx4_t acc = {0.f, 0.f, 0.f, 0.f};
          if (NP == 3) {
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], b[0], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[2], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[1], acc, 0, 0, 0);
          }
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[0], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[1], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[0], acc, 0, 0, 0);
#pragma unroll
          for (int i = 0; i < 4; ++i) accd[slot][i] += (double)acc[i];
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int slot = 0; slot < kMSlots; ++slot) {
      const int t = wave + 4 * slot;
      if (t >= ntiles) break;
      const int l = t / NT, nt = t - l * NT;
      const int n = 16 * nt + cc;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int rl = 4 * gq + i;
        if (rl < 2 * l + 1 && n < N) partials[(size_t)sg * K + (size_t)(l * l + rl) * N + n] = accd[slot][i];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// EOF accumulate as matrix products on the bf16 MFMA units (BFE_ACC_BF16)
// ---------------------------------------------------------------------------
// With the corner sums of each cell
//   S[s][cell][j][c] = sum_{p in cell} w_j(p) c_c(p)
// (w_j = -4 pi m {1, cos k phi, sin k phi}, c_c the bilinear corner weights)
// and their node sums G[s][j][node] (each node collects the corner of the up
// to four cells that touch it), the coefficients of snapshot s are
//   a[s][j][n] = sum_node G[s][j][node] T[node][j][n],
// for each row j a (snapshots x nodes) x (nodes x norder) matrix product.
//   eof_cell_sums_kernel   S from particles in cell order (any order is
//                          correct, cell order makes the runs long)
//   eof_node_pieces_kernel G, split into bf16 pieces hi + lo, [j][s][node]
//   eof_mfma_kernel        per row and 16 snapshots, hi*hi + hi*lo + lo*hi
//                          on v_mfma_f32_16x16x32_bf16, float32 sums over a
//                          block's 512 nodes
//   eof_mfma_reduce_kernel float64 sum of the block partials in block order
// T is the float32 potential table split into bf16 pieces once per table set
// (eof_mfma_prepare).  The pieces carry about 16 bits, so the mode is good
// to about 1e-5 of the larger terms (the float32 mode's 1e-7 class is not
// its aim).
constexpr int kMfmaKBlock = 512;    // nodes per block of the product (4 waves x 4 K-steps x 32)
constexpr int kCellSlotsMax = 128;  // (row, corner) slots: 4 (2 mmax + 1), two per lane

// Lanes own slots sl = lane + 64 o (row sl / 4, corner sl % 4).  Each wave
// walks a contiguous quarter of its block's segment, 64 particles per tile:
// lane q computes particle q's cell, corner weights and row weights into
// LDS, then every lane sums its slot over the tile's particles in order,
// adding the run sum to S (float64 hardware atomics) where the cell changes.
// The sums run in float32 over at most kFlush particles, then in float64.
__global__ __launch_bounds__(256) void eof_cell_sums_kernel(
    const float4* __restrict__ P, Segments S, const int64_t* __restrict__ off, int nsnap,
    EOFGeom g, float* __restrict__ cs) {
  extern __shared__ __align__(16) unsigned char dsm[];
  const int J = 2 * g.mmax + 1, NS = 4 * J;
  const int64_t ncell = (int64_t)g.numx * g.numy;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  float* sW = reinterpret_cast<float*>(dsm) + (size_t)wave * (J * kDepStride + kDepExtra);
  float* sE = sW + J * kDepStride;
  int* sC = reinterpret_cast<int*>(sE + 4 * 64);
  int jrow[2], corner[2];
#pragma unroll
  for (int o = 0; o < 2; ++o) {
    const int sl = lane + 64 * o;
    jrow[o] = sl < NS ? sl >> 2 : 0;
    corner[o] = sl & 3;
  }
  const bool own0 = lane < NS, own1 = lane + 64 < NS;
  for (int64_t sg = blockIdx.x; sg < S.nseg; sg += gridDim.x) {
    int64_t s0, s1;
    segment_range(S, sg, s0, s1);
    if (s1 <= s0) continue;
    int snap = 0;
    if (off) {
      int lo = 0, hi = nsnap - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= s0) lo = mid; else hi = mid - 1;
      }
      snap = lo;
    }
    float* out = cs + (size_t)snap * ncell * NS;
    const int64_t quarter = ((s1 - s0 + 3) / 4 + 63) / 64 * 64;
    const int64_t w0 = s0 + wave * quarter;
    const int64_t w1 = w0 + quarter < s1 ? w0 + quarter : s1;
    int cur = -1, cnt = 0;
    float acc[2] = {0.f, 0.f};
    double accd[2] = {0.0, 0.0};
    auto flush = [&]() {
      if (cur >= 0) {
        float* o = out + (size_t)cur * NS;
        if (own0) unsafeAtomicAdd(o + lane, (float)(accd[0] + (double)acc[0]));
        if (own1) unsafeAtomicAdd(o + lane + 64, (float)(accd[1] + (double)acc[1]));
      }
      acc[0] = acc[1] = 0.f;
      accd[0] = accd[1] = 0.0;
      cnt = 0;
    };
    for (int64_t t0 = w0; t0 < w1; t0 += 64) {
      const int64_t p = t0 + lane;
      if (p < w1) {
        const float4 q4 = P[p];
        const double x = q4.x, y = q4.y, z = q4.z;
        const double R = sqrt(x * x + y * y + 1.e-10);
        double c1, s1v;
        planar_angle(x, y, c1, s1v);
        const EOFCell cell = eof_locate(R, z, g.cmap, g.ascale, g.ihscale, g.xmin, g.idx, g.ymin,
                                        g.idy, g.numx, g.numy);
        const double w = -kFourPi * (double)q4.w;
        sC[lane] = cell.ix * g.numy + cell.iy;
        sE[lane] = (float)cell.c00;
        sE[64 + lane] = (float)cell.c10;
        sE[128 + lane] = (float)cell.c01;
        sE[192 + lane] = (float)cell.c11;
        sW[lane] = (float)w;
        double cm = 1.0, sm = 0.0, cm1 = c1, sm1 = -s1v;
        for (int m = 1; m <= g.mmax; ++m) {
          const double cn = (m == 1) ? c1 : 2.0 * c1 * cm - cm1;
          const double sn = (m == 1) ? s1v : 2.0 * c1 * sm - sm1;
          cm1 = cm; sm1 = sm; cm = cn; sm = sn;
          sW[m * kDepStride + lane] = (float)(w * cm);
          sW[(g.mmax + m) * kDepStride + lane] = (float)(w * sm);
        }
      }
      wave_sync();
      const int np = w1 - t0 < 64 ? (int)(w1 - t0) : 64;
      if (cur < 0) cur = sC[0];
#pragma unroll 1
      for (int q0 = 0; q0 < np; q0 += 4) {
        const int4 cc = *reinterpret_cast<const int4*>(sC + q0);
        const bool run = q0 + 4 <= np && cc.x == cur && cc.y == cur && cc.z == cur && cc.w == cur;
        if (__builtin_amdgcn_readfirstlane((int)run)) {
          if (cnt + 4 > kFlush) {
            accd[0] += (double)acc[0];
            accd[1] += (double)acc[1];
            acc[0] = acc[1] = 0.f;
            cnt = 0;
          }
#pragma unroll
          for (int o = 0; o < 2; ++o) {
            const float4 wv = *reinterpret_cast<const float4*>(sW + jrow[o] * kDepStride + q0);
            const float4 ev = *reinterpret_cast<const float4*>(sE + corner[o] * 64 + q0);
            acc[o] = fmaf(wv.x, ev.x, acc[o]);
            acc[o] = fmaf(wv.y, ev.y, acc[o]);
            acc[o] = fmaf(wv.z, ev.z, acc[o]);
            acc[o] = fmaf(wv.w, ev.w, acc[o]);
          }
          cnt += 4;
          continue;
        }
#pragma unroll 1
        for (int q = q0; q < q0 + 4 && q < np; ++q) {
          const int b = sC[q];
          if (b != cur) {
            flush();
            cur = b;
          } else if (cnt == kFlush) {
            accd[0] += (double)acc[0];
            accd[1] += (double)acc[1];
            acc[0] = acc[1] = 0.f;
            cnt = 0;
          }
#pragma unroll
          for (int o = 0; o < 2; ++o)
            acc[o] = fmaf(sW[jrow[o] * kDepStride + q], sE[corner[o] * 64 + q], acc[o]);
          ++cnt;
        }
      }
      wave_sync();
    }
    flush();
    cur = -1;
  }
}

// Node sums of the cell corner sums as bf16 pieces hi = bf16(G),
// lo = bf16(G - hi) in [j][s][node] order.  Block (X, s): the nodes of grid
// column X of snapshot s.  The cells of columns X and X-1 (numy rows of NS
// floats each, contiguous) are staged in LDS with an odd row stride; node
// (X, Y) takes corner c00 of cell (X, Y), c10 of (X-1, Y), c01 of (X, Y-1)
// and c11 of (X-1, Y-1).  The nodes past the grid are written as zero by the
// last column's blocks; rows of snapshots past nsnap (grid.y < nsnap) are
// neither written nor read (the product masks them).
__global__ __launch_bounds__(256) void eof_node_pieces_kernel(const float* __restrict__ cs,
                                                              int nsnap, int nsp, int64_t nodes_pad,
                                                              EOFGeom g, __bf16* __restrict__ Ahi,
                                                              __bf16* __restrict__ Alo) {
  extern __shared__ float sc[];
  const int X = blockIdx.x, s = blockIdx.y;
  const int J = 2 * g.mmax + 1, NS = 4 * J, RS = NS + 1;
  const int ny = g.numy, ny1 = ny + 1;
  const int64_t nodes = (int64_t)(g.numx + 1) * ny1;
  const bool live = s < nsnap;
  float* cur = sc;               // [ny][RS] cells (X, Y)
  float* prv = sc + ny * RS;     // [ny][RS] cells (X-1, Y)
  if (live) {
    const float* base = cs + (size_t)s * g.numx * ny * NS;
    for (int t = threadIdx.x; t < ny * NS; t += blockDim.x) {
      const int y = t / NS, k = t - y * NS;
      cur[y * RS + k] = X < g.numx ? base[((size_t)X * ny) * NS + t] : 0.f;
      prv[y * RS + k] = X >= 1 ? base[((size_t)(X - 1) * ny) * NS + t] : 0.f;
    }
  }
  __syncthreads();
  for (int Y = threadIdx.x; Y < ny1; Y += blockDim.x) {
    const int64_t node = (int64_t)X * ny1 + Y;
    for (int j = 0; j < J; ++j) {
      float G = 0.f;
      if (live) {
        if (Y < ny) G = cur[Y * RS + 4 * j] + prv[Y * RS + 4 * j + 1];
        if (Y >= 1) G += cur[(Y - 1) * RS + 4 * j + 2] + prv[(Y - 1) * RS + 4 * j + 3];
      }
      const __bf16 hi = (__bf16)G;
      const __bf16 lo = (__bf16)(G - (float)hi);
      const size_t o = ((size_t)j * nsp + s) * nodes_pad + node;
      Ahi[o] = hi;
      Alo[o] = lo;
    }
  }
  if (X == g.numx) {
    for (int64_t node = nodes + threadIdx.x; node < nodes_pad; node += blockDim.x)
      for (int j = 0; j < J; ++j) {
        const size_t o = ((size_t)j * nsp + s) * nodes_pad + node;
        Ahi[o] = (__bf16)0.f;
        Alo[o] = (__bf16)0.f;
      }
  }
}

// Per row j (grid.z), 16 snapshots (grid.y) and 512 nodes (grid.x): the
// product of the node pieces with the table pieces, three MFMAs per K-step
// and radial tile (hi*hi, hi*lo, lo*hi), float32 accumulation; the four
// waves' sums are added in wave order into part[kb][s][j][np] (np = norder
// rounded up to 16).  Operand layout of v_mfma_f32_16x16x32_bf16 (lane l,
// g = l / 16, c = l % 16): A[c][8g + i] = G[s0 + c][node0 + 8g + i],
// B[8g + i][c] = T[node0 + 8g + i][n0 + c], D[4g + i][c] = a[s0 + 4g + i][n0 + c].
// Snapshot rows s >= nsnap and radial columns n >= norder are not stored:
// their lanes take zero operands.
__global__ __launch_bounds__(256) void eof_mfma_kernel(const __bf16* __restrict__ Ahi,
                                                       const __bf16* __restrict__ Alo,
                                                       const __bf16* __restrict__ Thi,
                                                       const __bf16* __restrict__ Tlo, int nsnap,
                                                       int nsp, int J, int norder, int np,
                                                       int64_t nodes_pad, float* __restrict__ part) {
  __shared__ float red[4][4][64][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, c = lane & 15;
  const int kb = blockIdx.x, s0 = 16 * blockIdx.y, j = blockIdx.z;
  const int ntn = np / 16;
  const bool arow_ok = s0 + c < nsnap;
  const size_t arow = ((size_t)j * nsp + s0 + c) * nodes_pad;
  const bf16x8_t zero8 = {};
  floatx4_t acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = floatx4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
    const int64_t node0 = (int64_t)kb * kMfmaKBlock + wave * 128 + ks * 32 + 8 * g;
    const bf16x8_t ah = arow_ok ? *reinterpret_cast<const bf16x8_t*>(Ahi + arow + node0) : zero8;
    const bf16x8_t al = arow_ok ? *reinterpret_cast<const bf16x8_t*>(Alo + arow + node0) : zero8;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (t < ntn) {
        const bool bok = 16 * t + c < norder;
        const size_t brow = ((size_t)j * norder + 16 * t + c) * nodes_pad + node0;
        const bf16x8_t bh = bok ? *reinterpret_cast<const bf16x8_t*>(Thi + brow) : zero8;
        const bf16x8_t bl = bok ? *reinterpret_cast<const bf16x8_t*>(Tlo + brow) : zero8;
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc[t], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) red[wave][t][lane][i] = acc[t][i];
  __syncthreads();
  if (wave == 0) {
    for (int t = 0; t < ntn; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float v = ((red[0][t][lane][i] + red[1][t][lane][i]) + red[2][t][lane][i]) + red[3][t][lane][i];
        const int s = s0 + 4 * g + i;
        part[(((size_t)kb * nsp + s) * J + j) * np + 16 * t + c] = v;
      }
  }
}

// Thread per (s, j, n): the float64 sum of the block partials in block
// order, written as accum_cos[s][j][n] (j <= mmax) or accum_sin[s][j-mmax][n]
// (sin row m = 0 is zero).
__global__ __launch_bounds__(256) void eof_mfma_reduce_kernel(const float* __restrict__ part,
                                                              int nkb, int nsnap, int nsp, int J,
                                                              int np, int mmax, int norder,
                                                              double* __restrict__ cos_out,
                                                              double* __restrict__ sin_out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nsnap * J * norder) return;
  const int n = (int)(t % norder);
  const int j = (int)((t / norder) % J);
  const int s = (int)(t / ((int64_t)norder * J));
  double v = 0.0;
  for (int kb = 0; kb < nkb; ++kb) v += (double)part[(((size_t)kb * nsp + s) * J + j) * np + n];
  const size_t rows = (size_t)(mmax + 1) * norder;
  if (j <= mmax) {
    cos_out[(size_t)s * rows + (size_t)j * norder + n] = v;
    if (j == 0) sin_out[(size_t)s * rows + n] = 0.0;
  } else {
    sin_out[(size_t)s * rows + (size_t)(j - mmax) * norder + n] = v;
  }
}

// bf16 pieces of the potential table, [2][J][np][nodes_pad] from the
// function-major float32 table fold32[k = j norder + n][node]; padding zero.
__global__ __launch_bounds__(256) void eof_mfma_table_kernel(const float* __restrict__ fold32,
                                                             int J, int norder, int np,
                                                             int64_t nodes, int64_t nodes_pad,
                                                             __bf16* __restrict__ T) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)J * np * nodes_pad;
  if (t >= total) return;
  const int64_t node = t % nodes_pad;
  const int n = (int)((t / nodes_pad) % np);
  const int j = (int)(t / ((int64_t)np * nodes_pad));
  const float v = (n < norder && node < nodes) ? fold32[((size_t)j * norder + n) * nodes + node] : 0.f;
  const __bf16 hi = (__bf16)v;
  T[t] = hi;
  T[total + t] = (__bf16)(v - (float)hi);
}

// Segment plan of one accumulate launch.  Single snapshot: at most
// `cap` segments of whole tiles, one per block.  Batch: every
// snapshot is cut into segments of seg_tiles tiles (never spanning two
// snapshots); the blocks loop over all segments and rowb[s] is the first
// segment of snapshot s, so each snapshot's partial rows are reduced on
// their own, in a fixed order.
struct AccPlan {
  Segments S;
  int grid;
  std::vector<int64_t> segb, rowb;  // batch only (host copies)
};

static AccPlan plan_single(int64_t cap, int64_t n, int tile) {
  AccPlan P;
  const int64_t ntiles = (n + tile - 1) / tile;
  int64_t nseg = ntiles < cap ? ntiles : cap;
  if (nseg < 1) nseg = 1;
  const int64_t per = (ntiles + nseg - 1) / nseg;
  P.S.segb = nullptr;
  P.S.seglen = (per > 0 ? per : 1) * tile;
  P.S.n = n;
  P.S.nseg = (n + P.S.seglen - 1) / P.S.seglen;
  if (P.S.nseg < 1) P.S.nseg = 1;
  P.grid = (int)P.S.nseg;
  return P;
}

static AccPlan plan_batch(int64_t cap, int64_t seg_tiles, const int64_t* off, int nsnap, int tile) {
  AccPlan P;
  int64_t total_tiles = 0;
  for (int q = 0; q < nsnap; ++q) total_tiles += (off[q + 1] - off[q] + tile - 1) / tile;
  if (seg_tiles <= 0) {
    seg_tiles = (total_tiles + cap - 1) / cap;
    if (seg_tiles < 1) seg_tiles = 1;
  }
  P.rowb.push_back(0);
  P.segb.push_back(off[0]);
  for (int q = 0; q < nsnap; ++q) {
    const int64_t b = off[q], e = off[q + 1];
    for (int64_t p = b; p < e; p += seg_tiles * tile) {
      const int64_t pe = p + seg_tiles * tile < e ? p + seg_tiles * tile : e;
      P.segb.push_back(pe);
    }
    if (e == b) P.segb.push_back(e);   // empty snapshot: one empty segment
    P.rowb.push_back((int64_t)P.segb.size() - 1);
  }
  P.S.segb = nullptr;  // device copy set by upload_plan
  P.S.seglen = 0;
  P.S.n = off[nsnap];
  P.S.nseg = (int64_t)P.segb.size() - 1;
  P.grid = (int)(P.S.nseg < cap ? P.S.nseg : cap);
  return P;
}

// Copies the batch plan and the snapshot offsets to device memory behind the
// partials in scratch (waiting first for earlier work that may still read it).
static int upload_plan(Context* c, AccPlan& P, size_t pbytes, const int64_t* off, int nsnap,
                       hipStream_t s, const int64_t** rowb_dev, const int64_t** off_dev) {
  const size_t segb_off = (pbytes + 255) & ~(size_t)255;
  const size_t rowb_off = segb_off + ((P.segb.size() * 8 + 255) & ~(size_t)255);
  const size_t offs_off = rowb_off + ((P.rowb.size() * 8 + 255) & ~(size_t)255);
  int rc = ensure_scratch(c, offs_off + ((size_t)nsnap + 1) * 8 + 256);
  if (rc) return rc;
  BFE_HIP(c, hipStreamSynchronize(s));
  int64_t* segb = (int64_t*)((char*)c->scratch + segb_off);
  int64_t* rowb = (int64_t*)((char*)c->scratch + rowb_off);
  int64_t* offs = (int64_t*)((char*)c->scratch + offs_off);
  BFE_HIP(c, hipMemcpy(segb, P.segb.data(), P.segb.size() * 8, hipMemcpyHostToDevice));
  BFE_HIP(c, hipMemcpy(rowb, P.rowb.data(), P.rowb.size() * 8, hipMemcpyHostToDevice));
  BFE_HIP(c, hipMemcpy(offs, off, ((size_t)nsnap + 1) * 8, hipMemcpyHostToDevice));
  P.S.segb = segb;
  *rowb_dev = rowb;
  *off_dev = offs;
  return 0;
}

static int check_offsets(Context* c, const int64_t* off, int nsnap, const char* who) {
  if (!off || nsnap < 1) return set_error(c, BFE_EINVAL, "%s: need offsets and nsnap >= 1", who);
  if (off[0] != 0) return set_error(c, BFE_EINVAL, "%s: offsets[0] must be 0", who);
  for (int q = 0; q < nsnap; ++q)
    if (off[q + 1] < off[q]) return set_error(c, BFE_EINVAL, "%s: offsets must not decrease", who);
  return 0;
}

// (slots per lane, radial orders per slot) of the cell-ordered kernels.  The
// shape must give rows * ceil(N / CH) slots <= 64 * PPT.  With a fitting
// tuned index it is used; otherwise the first fitting shape of the list
// (ordered by preference).  {0, 0}: none fits.
struct DepShape {
  int ppt, ch;
};
static bool dep_fits(int rows, int N, DepShape d) { return rows * ((N + d.ch - 1) / d.ch) <= 64 * d.ppt; }
static DepShape dep_shape(int rows, int N, const DepShape* cands, int ncand, int64_t tuned) {
  if (tuned >= 0 && tuned < ncand && dep_fits(rows, N, cands[tuned])) return cands[tuned];
  for (int k = 0; k < ncand; ++k)
    if (dep_fits(rows, N, cands[k])) return cands[k];
  return DepShape{0, 0};
}
static const DepShape kEofShapes[kEofDepShapes] = {{1, 4}, {1, 8}, {1, 16}, {2, 2}, {2, 4}, {2, 8}, {2, 16}, {4, 8}};
static const DepShape kSlShapes[kSlDepShapes] = {{1, 8}, {1, 16}, {1, 24}, {2, 8}, {2, 16}, {2, 24}, {4, 16}};

int eof_dep_prepare(Context* c) {
  const DepShape dsh = dep_shape(2 * c->mmax + 1, c->norder, kEofShapes, kEofDepShapes,
                                 c->eof_dep_shape);
  const int ch = dsh.ppt > 0 ? dsh.ch : 4;
  const int rp = (c->norder + ch - 1) / ch * ch;
  if (c->eof_dep32 && c->eof_dep_rp == rp) return 0;
  const int64_t rows = (int64_t)(c->numx + 1) * (c->numy + 1) * (2 * c->mmax + 1);
  float* d = nullptr;
  BFE_HIP(c, hipDeviceSynchronize());   // the old layout may still be read
  BFE_HIP(c, hipMalloc(&d, (size_t)rows * rp * sizeof(float)));
  hipLaunchKernelGGL(eof_repad_kernel, dim3((unsigned)((rows * rp + 255) / 256)), dim3(256), 0,
                     nullptr, c->eof_pot32, d, rows, c->norder, rp);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    (void)hipFree(d);
    return check_hip(c, e, "eof_repad_kernel");
  }
  if (c->eof_dep32) (void)hipFree(c->eof_dep32);
  c->eof_dep32 = d;
  c->eof_dep_rp = rp;
  return 0;
}

// beside: another accumulate runs concurrently (selects the deposit kernel).
// bf16 pieces of the potential table for the MFMA mode (once per table set).
static int eof_mfma_prepare(Context* c, hipStream_t s) {
  if (c->eof_mfmaT) return 0;
  const int J = 2 * c->mmax + 1;
  const int64_t nodes = (int64_t)(c->numx + 1) * (c->numy + 1);
  const int64_t nodes_pad = (nodes + kMfmaKBlock - 1) / kMfmaKBlock * kMfmaKBlock;
  const int64_t total = (int64_t)J * c->norder * nodes_pad;
  BFE_HIP(c, hipMalloc(&c->eof_mfmaT, 2 * (size_t)total * sizeof(__bf16)));
  hipLaunchKernelGGL(eof_mfma_table_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                     c->eof_fold32[0], J, c->norder, c->norder, nodes, nodes_pad, c->eof_mfmaT);
  BFE_HIP(c, hipGetLastError());
  c->eof_mfma_np = (c->norder + 15) / 16 * 16;
  c->eof_mfma_nodes = nodes_pad;
  return 0;
}

static int ensure_mfma_work(Context* c, size_t bytes) {
  if (bytes <= c->mfma_work_bytes) return 0;
  if (c->mfma_work) {
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return check_hip(c, e, "hipDeviceSynchronize (mfma work grow)");
    (void)hipFree(c->mfma_work);
    c->mfma_work = nullptr;
    c->mfma_work_bytes = 0;
  }
  const size_t want = bytes + bytes / 8;
  BFE_HIP(c, hipMalloc(&c->mfma_work, want));
  c->mfma_work_bytes = want;
  return 0;
}

// The EOF accumulate in the bf16 MFMA mode (BFE_ACC_BF16) for one snapshot
// (off NULL) or nsnap concatenated ones; the same results as
// eof_accumulate_run to about 1e-5 of each row's larger terms.
static int eof_accumulate_mfma_run(Context* c, const float* x, const float* y, const float* z,
                                   const float* m, int64_t n, const int64_t* off, int nsnap,
                                   double* cos_out, double* sin_out, hipStream_t s) {
  const int J = 2 * c->mmax + 1, NS = 4 * J;
  if (NS > kCellSlotsMax)
    return set_error(c, BFE_ELIMIT, "eof_accumulate: BFE_ACC_BF16 needs mmax <= 15");
  if (c->norder > 64)
    return set_error(c, BFE_ELIMIT, "eof_accumulate: BFE_ACC_BF16 needs norder <= 64");
  const size_t plds = (size_t)2 * c->numy * (NS + 1) * sizeof(float);
  if (plds > 64 * 1024)
    return set_error(c, BFE_ELIMIT, "eof_accumulate: BFE_ACC_BF16 needs 2 numy (8 mmax + 5) <= 16384");
  int rc = eof_mfma_prepare(c, s);
  if (rc) return rc;
  const int np = c->eof_mfma_np;
  const int64_t nodes_pad = c->eof_mfma_nodes;
  const int nsp = (nsnap + 15) / 16 * 16;
  const int64_t ncell = (int64_t)c->numx * c->numy;
  const int nkb = (int)(nodes_pad / kMfmaKBlock);
  // work: cell sums [nsnap][cell][NS] f32 | node pieces [2][J][nsp][nodes_pad]
  // bf16 | block partials [nkb][nsp][J][np] f32
  const size_t cs_bytes = (size_t)nsnap * ncell * NS * sizeof(float);
  const size_t a_elems = (size_t)J * nsp * nodes_pad;
  const size_t a_off = (cs_bytes + 255) & ~(size_t)255;
  const size_t p_off = a_off + ((2 * a_elems * sizeof(__bf16) + 255) & ~(size_t)255);
  const size_t total = p_off + (size_t)nkb * nsp * J * np * sizeof(float);
  AccPlan P = off ? plan_batch(c->eof_acc_blocks, c->eof_seg_tiles, off, nsnap, kTile)
                  : plan_single(c->eof_acc_blocks, n, kTile);
  const int64_t* rowb = nullptr;
  const int64_t* off_dev = nullptr;
  if (off && (rc = upload_plan(c, P, 0, off, nsnap, s, &rowb, &off_dev))) return rc;
  if ((rc = ensure_mfma_work(c, total))) return rc;
  float* cs = (float*)c->mfma_work;
  __bf16* Ahi = (__bf16*)((char*)c->mfma_work + a_off);
  __bf16* Alo = Ahi + a_elems;
  float* part = (float*)((char*)c->mfma_work + p_off);
  const EOFGeom g = eof_geom(c);
  // zero the cell sums (NS = 4 J floats per cell: an even count, cleared as doubles)
  const int64_t ncs2 = (int64_t)nsnap * ncell * NS / 2;
  hipLaunchKernelGGL(fill_f64_kernel, dim3((unsigned)((ncs2 + 255) / 256)), dim3(256), 0, s,
                     (double*)cs, ncs2);
  BFE_HIP(c, hipGetLastError());
  if (n > 0) {
    const float4* sorted = nullptr;
    if ((rc = order_particles(c, 0, x, y, z, m, n, off_dev, nsnap, s, &sorted, off))) return rc;
    const size_t lds = (size_t)4 * (J * kDepStride + kDepExtra) * sizeof(float);
    timer_begin(c, "eof_cell_sums", s);
    hipLaunchKernelGGL(eof_cell_sums_kernel, dim3(P.grid), dim3(256), lds, s, sorted, P.S,
                       off_dev, nsnap, g, cs);
    timer_end(c, "eof_cell_sums", s);
    BFE_HIP(c, hipGetLastError());
  }
  hipLaunchKernelGGL(eof_node_pieces_kernel, dim3((unsigned)(c->numx + 1), (unsigned)nsnap), dim3(256),
                     plds, s, cs, nsnap, nsp, nodes_pad, g, Ahi, Alo);
  BFE_HIP(c, hipGetLastError());
  const __bf16* Thi = c->eof_mfmaT;
  const __bf16* Tlo = Thi + (size_t)J * c->norder * nodes_pad;
  timer_begin(c, "eof_mfma", s);
  hipLaunchKernelGGL(eof_mfma_kernel, dim3((unsigned)nkb, (unsigned)(nsp / 16), (unsigned)J), dim3(256),
                     0, s, Ahi, Alo, Thi, Tlo, nsnap, nsp, J, c->norder, np, nodes_pad, part);
  timer_end(c, "eof_mfma", s);
  BFE_HIP(c, hipGetLastError());
  const int64_t nout = (int64_t)nsnap * J * c->norder;
  hipLaunchKernelGGL(eof_mfma_reduce_kernel, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, s,
                     part, nkb, nsnap, nsp, J, np, c->mmax, c->norder, cos_out, sin_out);
  BFE_HIP(c, hipGetLastError());
  return 0;
}

static int eof_accumulate_run(Context* c, const float* x, const float* y, const float* z,
                              const float* m, int64_t n, const int64_t* off, int nsnap,
                              double* cos_out, double* sin_out, hipStream_t s,
                              bool beside = false) {
  const int K = c->eof_K;
  AccPlan P = off ? plan_batch(c->eof_acc_blocks, c->eof_seg_tiles, off, nsnap, kTile)
                  : plan_single(c->eof_acc_blocks, n, kTile);
  const size_t pbytes = (size_t)P.S.nseg * K * sizeof(double);
  const size_t need = pbytes;
  const int64_t* rowb = nullptr;
  const int64_t* off_dev = nullptr;
  int rc;
  if (off) {
    rc = upload_plan(c, P, need, off, nsnap, s, &rowb, &off_dev);   // partials | plan
  } else {
    rc = ensure_scratch(c, need);
  }
  if (rc) return rc;
  double* partials = (double*)c->scratch;
  EOFGeom g = eof_geom(c);
  const int opt = (K + 255) / 256;
  const float4* sorted = nullptr;
  // cell-ordered form: (2 mmax + 1) rows x chunks of radial orders per lane,
  // LDS for the per-wave weights and the float64 wave reduction
  const size_t dlds = std::max((size_t)4 * ((2 * c->mmax + 1) * kDepStride + kDepExtra) * sizeof(float),
                               (size_t)4 * K * sizeof(double));
  const DepShape dsh = dep_shape(2 * c->mmax + 1, c->norder, kEofShapes, kEofDepShapes, c->eof_dep_shape);
  if (n >= c->order_min && n > 0 && dsh.ppt > 0 && dlds <= 64 * 1024) {
    if ((rc = eof_dep_prepare(c))) return rc;
    rc = order_particles(c, 0, x, y, z, m, n, off_dev, nsnap, s, &sorted, off);
    if (rc) return rc;
  }
  timer_begin(c, "eof_accumulate", s);
  if (sorted) {
#define BFE_EOF_DEP(PP, CC)                                                                   \
  else if (dsh.ppt == PP && dsh.ch == CC) {                                                   \
    if (beside)                                                                               \
      hipLaunchKernelGGL((eof_deposit_kernel<PP, CC>), dim3(P.grid), dim3(256), dlds, s,       \
                         sorted, P.S, g, c->eof_dep32, c->eof_dep_rp, K, partials);           \
    else                                                                                      \
      hipLaunchKernelGGL((eof_deposit_alone_kernel<PP, CC>), dim3(P.grid), dim3(256), dlds, s, \
                         sorted, P.S, g, c->eof_dep32, c->eof_dep_rp, K, partials);           \
  }
    if (false) {}
    BFE_EOF_DEP(1, 4) BFE_EOF_DEP(1, 8) BFE_EOF_DEP(1, 16) BFE_EOF_DEP(2, 2) BFE_EOF_DEP(2, 4)
    BFE_EOF_DEP(2, 8) BFE_EOF_DEP(2, 16) BFE_EOF_DEP(4, 8)
#undef BFE_EOF_DEP
  } else if (n == 0) {
    hipLaunchKernelGGL(fill_f64_kernel, dim3((unsigned)((P.S.nseg * K + 255) / 256)), dim3(256), 0, s,
                       partials, P.S.nseg * (int64_t)K);
  } else if (opt == 1) {
    hipLaunchKernelGGL(eof_accumulate_kernel<1>, dim3(P.grid), dim3(256), 0, s, x, y, z, m, P.S, g, c->eof_pot32, K, partials);
  } else if (opt == 2) {
    hipLaunchKernelGGL(eof_accumulate_kernel<2>, dim3(P.grid), dim3(256), 0, s, x, y, z, m, P.S, g, c->eof_pot32, K, partials);
  } else {
    hipLaunchKernelGGL(eof_accumulate_kernel<4>, dim3(P.grid), dim3(256), 0, s, x, y, z, m, P.S, g, c->eof_pot32, K, partials);
  }
  timer_end(c, "eof_accumulate", s);
  BFE_HIP(c, hipGetLastError());
  const int nc = (c->mmax + 1) * c->norder;
  hipLaunchKernelGGL(reduce_partials_kernel<true>, dim3((K + kRedCols - 1) / kRedCols, nsnap), dim3(256), 0,
                     s, partials, (int)P.S.nseg, K, cos_out, rowb, sin_out, nc, c->norder);
  BFE_HIP(c, hipGetLastError());
  return 0;
}

// beside: another accumulate runs concurrently (selects the deposit kernel).
static int sl_accumulate_run(Context* c, const float* x, const float* y, const float* z,
                             const float* m, int64_t n, const int64_t* off, int nsnap, uint32_t flags,
                             double* coef_out, hipStream_t s, bool beside = false) {
  const int R = (c->lmax + 1) * (c->lmax + 1);
  const int LN = (c->lmax + 1) * c->nmax;
  const int K = R * c->nmax;
  if (K > 1024) return set_error(c, BFE_ELIMIT, "sl_accumulate: (lmax+1)^2 nmax = %d > 1024", K);
  const size_t lds = ((size_t)R * (kSLTile + 1) + (size_t)kSLTile * LN + 3 * kSLTile) * 4;
  if (lds > 64 * 1024) return set_error(c, BFE_ELIMIT, "sl_accumulate: LDS %zu B > 64 KiB", lds);
  AccPlan P = off ? plan_batch(c->sl_acc_blocks, c->sl_seg_tiles, off, nsnap, kSLTile)
                  : plan_single(beside ? c->sl_acc_blocks_beside : c->sl_acc_blocks, n, kSLTile);
  const size_t pbytes = (size_t)P.S.nseg * K * sizeof(double);
  const int64_t* rowb = nullptr;
  const int64_t* off_dev = nullptr;
  int rc = off ? upload_plan(c, P, pbytes, off, nsnap, s, &rowb, &off_dev) : ensure_scratch(c, pbytes);
  if (rc) return rc;
  double* partials = (double*)c->scratch;
  SLGeom g = sl_geom(c);
  const int no_odd = (flags & BFE_NO_ODD) ? 1 : 0;
  const int opt = (K + 255) / 256;
  // matrix-product form on the MFMA units (lmax <= 7, at most 16 output tiles)
  const int mtiles = (c->lmax + 1) * ((c->nmax + 15) / 16);
  const bool bf16_mode = (flags & BFE_ACC_BF16) != 0;
  const bool mfma_ok = c->lmax <= 7 && mtiles <= 4 * kMSlots;
  if (bf16_mode && !mfma_ok)
    return set_error(c, BFE_ELIMIT, "sl_accumulate: BFE_ACC_BF16 needs lmax <= 7 and (lmax+1) ceil(nmax/16) <= 16");
  // measured at lmax 6, nmax 20, 1M particles on MI355X: MFMA 0.59 ms, the
  // cell-ordered vector-ALU path 0.19 ms + 0.09 ms sort, so "auto" takes the
  // vector-ALU paths (building the bf16 operand pieces costs more than the
  // products save at these sizes)
  if (n > 0 && mfma_ok && (bf16_mode || c->sl_acc_path == 1)) {
    const int np = bf16_mode ? 2 : 3;
    const size_t mlds = (size_t)np * R * kMS * 2 + 3 * kMT * 4;
    timer_begin(c, "sl_accumulate", s);
    if (np == 3)
      hipLaunchKernelGGL(sl_mfma_kernel<3>, dim3(P.grid), dim3(256), mlds, s, x, y, z, m, P.S, g, c->sl_e32, c->sl_lmc, no_odd, partials);
    else
      hipLaunchKernelGGL(sl_mfma_kernel<2>, dim3(P.grid), dim3(256), mlds, s, x, y, z, m, P.S, g, c->sl_e32, c->sl_lmc, no_odd, partials);
    timer_end(c, "sl_accumulate", s);
    BFE_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(reduce_partials_kernel<false>, dim3((K + kRedCols - 1) / kRedCols, nsnap), dim3(256), 0,
                       s, partials, (int)P.S.nseg, K, coef_out, rowb, nullptr, 0, 0);
    BFE_HIP(c, hipGetLastError());
    return 0;
  }
  // cell-ordered form: (lmax+1)^2 rows x chunks of radial orders per lane,
  // LDS for the per-wave Y rows and the float64 wave reduction
  const DepShape dsh = dep_shape(R, c->nmax, kSlShapes, kSlDepShapes, c->sl_dep_shape);
  const size_t dlds = std::max((size_t)4 * sl_dep_rows(R) * kDepStride * sizeof(float),
                               (size_t)4 * K * sizeof(double));
  const float4* sorted = nullptr;
  if (n >= c->order_min && n > 0 && dsh.ppt > 0 && dlds <= 64 * 1024) {
    rc = order_particles(c, 1, x, y, z, m, n, off_dev, nsnap, s, &sorted, off);
    if (rc) return rc;
  }
  timer_begin(c, "sl_accumulate", s);
  if (sorted) {
#define BFE_SL_DEP_LM(PP, CC, LL)                                                              \
  if (beside)                                                                                  \
    hipLaunchKernelGGL((sl_deposit_kernel<PP, CC, LL>), dim3(P.grid), dim3(256), dlds, s, sorted, \
                       P.S, g, c->sl_dep32, c->sl_lmc, no_odd, partials);                      \
  else                                                                                         \
    hipLaunchKernelGGL((sl_deposit_alone_kernel<PP, CC, LL>), dim3(P.grid), dim3(256), dlds, s,  \
                       sorted, P.S, g, c->sl_dep32, c->sl_lmc, no_odd, partials);
#define BFE_SL_DEP(PP, CC)                                                 \
  else if (dsh.ppt == PP && dsh.ch == CC) {                                \
    if (c->lmax == 6) { BFE_SL_DEP_LM(PP, CC, 6) }                         \
    else if (c->lmax == 4) { BFE_SL_DEP_LM(PP, CC, 4) }                    \
    else { BFE_SL_DEP_LM(PP, CC, 0) }                                      \
  }
    if (false) {}
    BFE_SL_DEP(1, 8) BFE_SL_DEP(1, 16) BFE_SL_DEP(1, 24) BFE_SL_DEP(2, 8) BFE_SL_DEP(2, 16)
    BFE_SL_DEP(2, 24) BFE_SL_DEP(4, 16)
#undef BFE_SL_DEP
#undef BFE_SL_DEP_LM
  } else if (n == 0) {
    hipLaunchKernelGGL(fill_f64_kernel, dim3((unsigned)((P.S.nseg * K + 255) / 256)), dim3(256), 0, s,
                       partials, P.S.nseg * (int64_t)K);
  } else if (opt == 1) {
    hipLaunchKernelGGL(sl_accumulate_kernel<1>, dim3(P.grid), dim3(256), lds, s, x, y, z, m, P.S, g, c->sl_e32, c->sl_lmc, no_odd, partials);
  } else if (opt == 2) {
    hipLaunchKernelGGL(sl_accumulate_kernel<2>, dim3(P.grid), dim3(256), lds, s, x, y, z, m, P.S, g, c->sl_e32, c->sl_lmc, no_odd, partials);
  } else {
    hipLaunchKernelGGL(sl_accumulate_kernel<4>, dim3(P.grid), dim3(256), lds, s, x, y, z, m, P.S, g, c->sl_e32, c->sl_lmc, no_odd, partials);
  }
  timer_end(c, "sl_accumulate", s);
  BFE_HIP(c, hipGetLastError());
  hipLaunchKernelGGL(reduce_partials_kernel<false>, dim3((K + kRedCols - 1) / kRedCols, nsnap), dim3(256), 0,
                     s, partials, (int)P.S.nseg, K, coef_out, rowb, nullptr, 0, 0);
  BFE_HIP(c, hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// Allocation-free EOF accumulate (bfe_eof_accumulate_noalloc)
// ---------------------------------------------------------------------------
// Per-segment partial sums of the input-order kernel live in this file-scope
// array (4 MiB per device, held until the process ends), so the call needs
// no device allocation.  Host code must not pass the array's name to a
// kernel: the name is a host-side stand-in, not a device address.
// noalloc_scratch_init takes the device address with hipGetSymbolAddress
// once per context (bfe_create, with the context's device current) and the
// entry point passes that address.  Every context on one device shares the
// one array: calls that overlap on one GPU race.
namespace noalloc {
constexpr int64_t kScratchDoubles = (int64_t)1 << 19;
static __device__ __align__(256) double scratch[kScratchDoubles];
}  // namespace noalloc

int noalloc_scratch_init(Context* c) {
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc::scratch));
  if (e != hipSuccess) {
    c->noalloc_scratch = nullptr;
    return check_hip(c, e, "hipGetSymbolAddress (allocation-free accumulate scratch)");
  }
  c->noalloc_scratch = (double*)p;
  return 0;
}

static int eof_accumulate_noalloc_run(Context* c, const float* x, const float* y, const float* z,
                                      const float* m, int64_t n, double* cos_out,
                                      double* sin_out, hipStream_t s) {
  const int K = c->eof_K;
  if (K > 1024)
    return set_error(c, BFE_ELIMIT, "eof_accumulate_noalloc: (2 mmax + 1) norder = %d > 1024", K);
  // at most kScratchDoubles / K segments of K partial sums each
  const int64_t cap = std::min<int64_t>(std::max<int64_t>(c->eof_acc_blocks, 1),
                                        noalloc::kScratchDoubles / K);
  AccPlan P = plan_single(cap, n, kTile);
  if (P.S.nseg * (int64_t)K > noalloc::kScratchDoubles)
    return set_error(c, BFE_ELIMIT, "eof_accumulate_noalloc: %lld partial sums > %lld",
                     (long long)(P.S.nseg * (int64_t)K), (long long)noalloc::kScratchDoubles);
  double* partials = c->noalloc_scratch;
  const EOFGeom g = eof_geom(c);
  const int opt = (K + 255) / 256;
  if (n == 0) {
    hipLaunchKernelGGL(fill_f64_kernel, dim3((unsigned)((P.S.nseg * K + 255) / 256)), dim3(256), 0, s,
                       partials, P.S.nseg * (int64_t)K);
  } else if (opt == 1) {
    hipLaunchKernelGGL(eof_accumulate_kernel<1>, dim3(P.grid), dim3(256), 0, s, x, y, z, m, P.S, g, c->eof_pot32, K, partials);
  } else if (opt == 2) {
    hipLaunchKernelGGL(eof_accumulate_kernel<2>, dim3(P.grid), dim3(256), 0, s, x, y, z, m, P.S, g, c->eof_pot32, K, partials);
  } else {
    hipLaunchKernelGGL(eof_accumulate_kernel<4>, dim3(P.grid), dim3(256), 0, s, x, y, z, m, P.S, g, c->eof_pot32, K, partials);
  }
  BFE_HIP(c, hipGetLastError());
  const int nc = (c->mmax + 1) * c->norder;
  hipLaunchKernelGGL(reduce_partials_kernel<true>, dim3((K + kRedCols - 1) / kRedCols, 1), dim3(256), 0,
                     s, partials, (int)P.S.nseg, K, cos_out, nullptr, sin_out, nc, c->norder);
  BFE_HIP(c, hipGetLastError());
  return 0;
}

int eof_accumulate_one(Context* c, const float* x, const float* y, const float* z,
                       const float* m, int64_t n, double* cos_out, double* sin_out, hipStream_t s) {
  return eof_accumulate_run(c, x, y, z, m, n, nullptr, 1, cos_out, sin_out, s, true);
}

int sl_accumulate_one(Context* c, const float* x, const float* y, const float* z, const float* m,
                      int64_t n, uint32_t flags, double* coef_out, hipStream_t s) {
  return sl_accumulate_run(c, x, y, z, m, n, nullptr, 1, flags, coef_out, s, true);
}

}  // namespace bfe

using namespace bfe;

static int acc_common_checks(Context* c, const float* x, const float* y, const float* z,
                             const float* m, int64_t n, const char* who) {
  if (n < 0) return set_error(c, BFE_EINVAL, "%s: n < 0", who);
  if (n > 0 && (!x || !y || !z || !m)) return set_error(c, BFE_EINVAL, "%s: NULL input", who);
  return 0;
}

extern "C" int bfe_eof_accumulate(bfe_ctx* h, const float* x, const float* y, const float* z,
                                  const float* m, int64_t n, uint32_t flags, double* cos_out,
                                  double* sin_out, void* stream) {
  // BFE_NO_ODD: eof.accumulate computes a no_odd mask but never applies it
  // (eof.py:545-551), so it is ignored; BFE_ACC_BF16 selects the MFMA mode
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (!c->eof_ok) return set_error(c, BFE_ENOTABLE, "eof_accumulate: EOF tables not set");
  if (!cos_out || !sin_out) return set_error(c, BFE_EINVAL, "eof_accumulate: NULL output");
  int rc = acc_common_checks(c, x, y, z, m, n, "eof_accumulate");
  if (rc) return rc;
  BFE_HIP(c, hipSetDevice(c->device));
  if (flags & BFE_ACC_BF16)
    return eof_accumulate_mfma_run(c, x, y, z, m, n, nullptr, 1, cos_out, sin_out, (hipStream_t)stream);
  return eof_accumulate_run(c, x, y, z, m, n, nullptr, 1, cos_out, sin_out, (hipStream_t)stream);
}

// The input-order path of bfe_eof_accumulate with its partial sums in the
// library's static scratch array: no device allocation, memset or sort.
extern "C" int bfe_eof_accumulate_noalloc(bfe_ctx* h, const float* x, const float* y,
                                          const float* z, const float* m, int64_t n,
                                          uint32_t flags, double* cos_out, double* sin_out,
                                          void* stream) {
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  // BFE_NO_ODD has no effect, as in bfe_eof_accumulate (eof.py:545-551)
  if (flags & BFE_ACC_BF16)
    return set_error(c, BFE_EINVAL, "eof_accumulate_noalloc: BFE_ACC_BF16 needs work memory (use bfe_eof_accumulate)");
  if (!c->eof_ok) return set_error(c, BFE_ENOTABLE, "eof_accumulate_noalloc: EOF tables not set");
  if (!cos_out || !sin_out) return set_error(c, BFE_EINVAL, "eof_accumulate_noalloc: NULL output");
  int rc = acc_common_checks(c, x, y, z, m, n, "eof_accumulate_noalloc");
  if (rc) return rc;
  if (!c->noalloc_scratch)
    return set_error(c, BFE_EHIP, "eof_accumulate_noalloc: no device address for the scratch array");
  BFE_HIP(c, hipSetDevice(c->device));
  return eof_accumulate_noalloc_run(c, x, y, z, m, n, cos_out, sin_out, (hipStream_t)stream);
}

extern "C" int bfe_eof_accumulate_batch(bfe_ctx* h, const float* x, const float* y, const float* z,
                                        const float* m, const int64_t* offsets, int nsnap,
                                        uint32_t flags, double* cos_out, double* sin_out,
                                        void* stream) {
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (!c->eof_ok) return set_error(c, BFE_ENOTABLE, "eof_accumulate_batch: EOF tables not set");
  if (!cos_out || !sin_out) return set_error(c, BFE_EINVAL, "eof_accumulate_batch: NULL output");
  int rc = check_offsets(c, offsets, nsnap, "eof_accumulate_batch");
  if (rc) return rc;
  rc = acc_common_checks(c, x, y, z, m, offsets[nsnap], "eof_accumulate_batch");
  if (rc) return rc;
  if (nsnap > 65535) return set_error(c, BFE_ELIMIT, "eof_accumulate_batch: nsnap > 65535");
  BFE_HIP(c, hipSetDevice(c->device));
  if (flags & BFE_ACC_BF16)
    return eof_accumulate_mfma_run(c, x, y, z, m, offsets[nsnap], offsets, nsnap, cos_out,
                                   sin_out, (hipStream_t)stream);
  return eof_accumulate_run(c, x, y, z, m, offsets[nsnap], offsets, nsnap, cos_out, sin_out,
                            (hipStream_t)stream);
}

extern "C" int bfe_sl_accumulate(bfe_ctx* h, const float* x, const float* y, const float* z,
                                 const float* m, int64_t n, uint32_t flags, double* coef_out,
                                 void* stream) {
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (!c->sl_ok) return set_error(c, BFE_ENOTABLE, "sl_accumulate: SL tables not set");
  if (!coef_out) return set_error(c, BFE_EINVAL, "sl_accumulate: NULL output");
  int rc = acc_common_checks(c, x, y, z, m, n, "sl_accumulate");
  if (rc) return rc;
  BFE_HIP(c, hipSetDevice(c->device));
  return sl_accumulate_run(c, x, y, z, m, n, nullptr, 1, flags, coef_out, (hipStream_t)stream);
}

// Fields.total_coefficients (potential.py:98-252) for one snapshot: the disc
// EOF and halo SL coefficients in one call, the two accumulates running
// concurrently (the SL half on the context's internal stream with its own
// work buffers).  Results are those of bfe_eof_accumulate and
// bfe_sl_accumulate.
extern "C" int bfe_accumulate_disc_halo(bfe_ctx* h, const float* dx, const float* dy,
                                        const float* dz, const float* dm, int64_t nd,
                                        const float* hx, const float* hy, const float* hz,
                                        const float* hm, int64_t nh, uint32_t sl_flags,
                                        double* cos_out, double* sin_out, double* coef_out,
                                        void* stream) {
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (!c->eof_ok || !c->sl_ok) return set_error(c, BFE_ENOTABLE, "accumulate_disc_halo: tables not set");
  if (!cos_out || !sin_out || !coef_out)
    return set_error(c, BFE_EINVAL, "accumulate_disc_halo: NULL output");
  int rc = acc_common_checks(c, dx, dy, dz, dm, nd, "accumulate_disc_halo (disc)");
  if (rc) return rc;
  rc = acc_common_checks(c, hx, hy, hz, hm, nh, "accumulate_disc_halo (halo)");
  if (rc) return rc;
  BFE_HIP(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream, s2 = nullptr;
  rc = fork_begin(c, s, &s2);
  if (rc) return rc;
  swap_work(c);
  const int rs = sl_accumulate_run(c, hx, hy, hz, hm, nh, nullptr, 1, sl_flags, coef_out, s2, true);
  swap_work(c);
  const int re = rs ? 0 : eof_accumulate_run(c, dx, dy, dz, dm, nd, nullptr, 1, cos_out, sin_out, s, true);
  rc = fork_end(c, s);
  return rs ? rs : (re ? re : rc);
}

extern "C" int bfe_sl_accumulate_batch(bfe_ctx* h, const float* x, const float* y, const float* z,
                                       const float* m, const int64_t* offsets, int nsnap,
                                       uint32_t flags, double* coef_out, void* stream) {
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (!c->sl_ok) return set_error(c, BFE_ENOTABLE, "sl_accumulate_batch: SL tables not set");
  if (!coef_out) return set_error(c, BFE_EINVAL, "sl_accumulate_batch: NULL output");
  int rc = check_offsets(c, offsets, nsnap, "sl_accumulate_batch");
  if (rc) return rc;
  rc = acc_common_checks(c, x, y, z, m, offsets[nsnap], "sl_accumulate_batch");
  if (rc) return rc;
  if (nsnap > 65535) return set_error(c, BFE_ELIMIT, "sl_accumulate_batch: nsnap > 65535");
  BFE_HIP(c, hipSetDevice(c->device));
  return sl_accumulate_run(c, x, y, z, m, offsets[nsnap], offsets, nsnap, flags, coef_out,
                           (hipStream_t)stream);
}
