// Cell ordering of evaluation points and particles.
//
// The folded field tables (SL 1.8 MB, EOF 7.4 MB at the workload size) do
// not fit in one XCD's 4 MB L2, and points in input order send every lane of
// a wave to a different table cell.  Sorting the points by table cell lets a
// wave's lanes share cells and, with the XCD-aware block mapping, each XCD
// work on one contiguous range of cells.  Accumulate sorts its particles the
// same way so that runs of equal cells can be summed before the table reads.
//
// Key (kind 1, evaluation and the SL accumulate): the SL radial cell of the
// point times a number of cos(theta) bins, plus its bin (key_sl_theta);
// (kind 0, the EOF accumulate): the EOF (ix, iy) cell; plus snapshot * cells
// for concatenated snapshots.
//
// The sort is a stable least-significant-digit counting sort over 8-bit
// digits, ceil(bits / 8) passes of three kernels:
//   count   - per 2048-item tile, the digit histogram by LDS atomics (the
//             first pass also computes the keys from the coordinates);
//   rowscan - per digit, the exclusive prefix of the tile counts;
//   scatter - each tile ranks its items per digit, stages them in LDS in
//             digit order and writes them out in runs.
// In the scatter, items of a wave with equal digits are found with eight
// ballots and ranked by popcount; a per-wave histogram in LDS carries the
// ranks across the wave's item slots.  The order inside a tile is (wave,
// slot, lane), which is input order, so the sort is stable and the result
// reproducible.  Concatenated snapshots sort segmented: tiles (at most 2048
// items) stay inside one snapshot and the digit prefixes restart at each.
// There are no global atomics (a counting sort with them was slow here
// because thousands of points share the busiest cells), no spin-waits and no
// memsets: every counter is written by a kernel.  The packed float4 (x, y,
// z, mass or input index) travels with the key, so no gather pass is needed.
#include <algorithm>
#include <vector>

#include "bfe_field.hpp"

namespace bfe {

constexpr int kSortThreads = 256;
constexpr int kSortIPT = 8;                              // items per thread (2 and 4 measured slower)
constexpr int kSortTile = kSortThreads * kSortIPT;       // items per tile (block)
constexpr int kRadix = 256;                              // 8-bit digits

// Snapshot of point i for concatenated snapshots: the last s with off[s] <= i.
__device__ __forceinline__ int snapshot_of(const int64_t* __restrict__ off, int nsnap, int64_t i) {
  int lo = 0, hi = nsnap - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= i) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Table geometry for the sort keys, in float32.  A key only decides the
// order: the accumulate kernels recompute each particle's cell in float64
// and start a new run wherever it changes, and evaluation is per point, so a
// key that differs from the float64 cell next to a cell edge costs a shorter
// run, never a different result.
struct KeyGeom {
  int ecmap, numx, numy;
  float ea, eih, exmin, eidx, eymin, eidy;   // EOF: ascale, 1/hscale, xmin, 1/dx, ymin, 1/dy
  int scmap, numr;
  float ss, sxi0, sidxi;                     // SL: scale, xi[0], 1/(xi[1] - xi[0])
  uint32_t tbins;                            // cos(theta) bins per SL cell (kind 1)
};

static KeyGeom key_geom(const EOFGeom& e, const SLGeom& s, const double* h_xi) {
  KeyGeom g;
  g.ecmap = e.cmap;
  g.numx = e.numx;
  g.numy = e.numy;
  g.ea = (float)e.ascale;
  g.eih = (float)(1.0 / e.hscale);
  g.exmin = (float)e.xmin;
  g.eidx = (float)(1.0 / e.dx);
  g.eymin = (float)e.ymin;
  g.eidy = (float)(1.0 / e.dy);
  g.scmap = s.cmap;
  g.numr = s.numr;
  g.ss = (float)s.scale;
  g.sxi0 = h_xi ? (float)h_xi[0] : 0.f;
  g.sidxi = h_xi ? (float)(1.0 / (h_xi[1] - h_xi[0])) : 0.f;
  g.tbins = 1;
  return g;
}

// cos(theta) bins per SL cell: kThetaBins unless the keys of nsnap
// snapshots would then pass 32 bits
constexpr uint32_t kThetaBins = 32;
static uint32_t theta_bins(int numr, int nsnap) {
  const uint64_t full = (uint64_t)(numr - 1) * kThetaBins * (uint64_t)nsnap;
  return full <= 0xffffffffull ? kThetaBins : 1u;
}

// r -> xi of the maps (compatibility.py r_to_xi): (r/a - 1)/(r/a + 1), ln r, r
__device__ __forceinline__ float key_xi(float r, int cmap, float a) {
  if (cmap == 1) return __fdividef(r - a, r + a);
  if (cmap == 2) return __logf(r);
  return r;
}

// z -> y = sign(z) asinh(|z| / h)
__device__ __forceinline__ float key_y(float z, float ih) {
  const float q = fabsf(z) * ih;
  const float v = __logf(q + sqrtf(fmaf(q, q, 1.f)));
  return z < 0.f ? -v : v;
}

// floor(v) clamped to [0, hi]; NaN -> 0
__device__ __forceinline__ uint32_t key_clamp(float v, int hi) {
  return v >= 0.f ? (uint32_t)fminf(v, (float)hi) : 0u;
}

__device__ __forceinline__ uint32_t key_sl_cell(const KeyGeom& g, float r) {
  return key_clamp((key_xi(r, g.scmap, g.ss) - g.sxi0) * g.sidxi, g.numr - 2);
}

// SL radial cell, then one of g.tbins equal bins of cos(theta): points of
// one key share their SL nodes and lie in a small (R, z) patch of the EOF
// grid, which serves the accumulate (runs of one SL cell) and evaluation.
__device__ __forceinline__ uint32_t key_sl_theta(const KeyGeom& g, float x, float y, float z) {
  const float r = sqrtf(fmaf(x, x, fmaf(y, y, z * z)));
  const float ct = r > 0.f ? __fdividef(z, r) : 0.f;
  return key_sl_cell(g, fmaxf(r, 1.e-10f)) * g.tbins +
         key_clamp((ct + 1.f) * (0.5f * (float)g.tbins), (int)g.tbins - 1);
}

// Where the keys and packed values of the first pass come from.
struct KeySrc {
  const float* x;
  const float* y;
  const float* z;
  const float* m;        // mass for the packed .w; NULL: the input index as int bits
  const int64_t* off;    // device snapshot offsets [nsnap + 1], NULL: one snapshot
  int nsnap;
  uint32_t ncell;        // cells per snapshot
  int kind;              // 0 EOF accumulate cell, 1 SL accumulate / evaluation cell
  KeyGeom g;
  // items n1.. come from a second set of arrays (x2, y2, z2; no mass), so
  // two point sets are ordered as one: the index in .w runs over both
  const float* x2;
  const float* y2;
  const float* z2;
  int64_t n1;
};

__device__ __forceinline__ void source_xyz(const KeySrc& S, int64_t i, float& x, float& y, float& z) {
  if (i < S.n1) {
    x = S.x[i]; y = S.y[i]; z = S.z[i];
  } else {
    const int64_t k = i - S.n1;
    x = S.x2[k]; y = S.y2[k]; z = S.z2[k];
  }
}

// Packed value of item i on the first pass: (x, y, z, mass or index bits).
__device__ __forceinline__ float4 source_value(const KeySrc& S, int64_t i) {
  float x, y, z;
  source_xyz(S, i, x, y, z);
  return make_float4(x, y, z, S.m ? S.m[i] : __int_as_float((int)i));
}

// kind 0: the EOF (ix, iy) cell of (R = sqrt(x^2 + y^2 + 1e-10), z), clamped
// to the grid as eof.return_bins; kind 1: key_sl_theta of r = max(|x|, 1e-10)
// (the SL accumulate and evaluation order).
// snap: the point's snapshot when the caller knows it (a tile of the
// segmented sort lies in one snapshot), else -1 (searched in S.off).
__device__ __forceinline__ uint32_t source_key(const KeySrc& S, int64_t i, int snap_known) {
  float x, y, z;
  source_xyz(S, i, x, y, z);
  const KeyGeom& g = S.g;
  uint32_t cell;
  if (S.kind == 1) {
    cell = key_sl_theta(g, x, y, z);
  } else {
    const float R = sqrtf(fmaf(x, x, fmaf(y, y, 1.e-10f)));
    const float X = (key_xi(R, g.ecmap, g.ea) - g.exmin) * g.eidx;
    const float Y = (key_y(z, g.eih) - g.eymin) * g.eidy;
    cell = key_clamp(X, g.numx - 1) * (uint32_t)g.numy + key_clamp(Y, g.numy - 1);
  }
  const uint32_t snap = snap_known >= 0 ? (uint32_t)snap_known
                       : (S.off ? (uint32_t)snapshot_of(S.off, S.nsnap, i) : 0u);
  return snap * S.ncell + cell;
}

// Lanes of the wave (among the live ones) whose digit equals this lane's.
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool live) {
  uint64_t mask = __ballot(live);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const bool bit = (d >> b) & 1u;
    const uint64_t v = __ballot(live && bit);
    mask &= bit ? v : ~v;
  }
  return mask;
}

// Inclusive scan over the 256 threads of the block (all must call it).
__device__ __forceinline__ uint32_t block_scan_incl(uint32_t v, uint32_t* wsum) {
  const int t = threadIdx.x, l = t & 63, w = t >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(v, o, 64);
    if (l >= o) v += u;
  }
  if (l == 63) wsum[w] = v;
  __syncthreads();
  uint32_t add = 0;
  for (int q = 0; q < w; ++q) add += wsum[q];
  __syncthreads();
  return v + add;
}

// Items [t0, end) of tile b: b kSortTile and on, or the bounds tb[2b], tb[2b + 1]
// (at most kSortTile items) of a segmented sort.  Slot i of thread (w, l)
// holds item t0 + (w kSortIPT + i) 64 + l.
struct TileRange {
  int64_t t0, end;
};
__device__ __forceinline__ TileRange sort_tile(const int64_t* __restrict__ tb, int64_t n) {
  if (tb) return TileRange{tb[2 * blockIdx.x], tb[2 * blockIdx.x + 1]};
  const int64_t t0 = (int64_t)blockIdx.x * kSortTile;
  return TileRange{t0, t0 + kSortTile < n ? t0 + kSortTile : n};
}
__device__ __forceinline__ int64_t sort_item(const TileRange& r, int i) {
  const int t = threadIdx.x;
  return r.t0 + (int64_t)((t >> 6) * kSortIPT + i) * 64 + (t & 63);
}

// Digit histogram of each tile: counts[d * nb + tile].  First pass (kin ==
// NULL): keys computed from the source and written to kfirst.
__global__ __launch_bounds__(kSortThreads) void sort_count_kernel(
    KeySrc S, const uint32_t* __restrict__ kin, uint32_t* __restrict__ kfirst, int64_t n,
    const int64_t* __restrict__ tb, const int* __restrict__ tseg, int shift,
    uint32_t* __restrict__ counts, int nb) {
  __shared__ uint32_t hist[4][kRadix];
  const int t = threadIdx.x, w = t >> 6, l = t & 63;
  const TileRange tr = sort_tile(tb, n);
  // segmented sort: digits of the key less its segment's first key, so
  // that the passes cover the cell bits only
  const int tsnap = tseg ? tseg[blockIdx.x] : -1;
  const uint32_t kb = tseg ? (uint32_t)tsnap * S.ncell : 0u;
  for (int j = t; j < 4 * kRadix; j += kSortThreads) (&hist[0][0])[j] = 0u;
  uint32_t key[kSortIPT];
#pragma unroll
  for (int i = 0; i < kSortIPT; ++i) {
    const int64_t idx = sort_item(tr, i);
    key[i] = 0u;
    if (idx < tr.end) {
      if (kin) {
        key[i] = kin[idx];
      } else {
        key[i] = source_key(S, idx, tsnap);
        kfirst[idx] = key[i];
      }
    }
  }
  __syncthreads();
  // counts only, so LDS atomics (lanes on one digit serialise in the LDS,
  // no worse than the eight ballots of a rank)
#pragma unroll
  for (int i = 0; i < kSortIPT; ++i) {
    const bool live = sort_item(tr, i) < tr.end;
    const uint32_t d = ((key[i] - kb) >> shift) & (kRadix - 1);
    if (live) atomicAdd(&hist[w][d], 1u);
  }
  __syncthreads();
  counts[(int64_t)t * nb + blockIdx.x] = hist[0][t] + hist[1][t] + hist[2][t] + hist[3][t];
}

// Per digit (block), the exclusive prefix of its tile counts, in place, and
// the digit total.
__global__ __launch_bounds__(kSortThreads) void sort_rowscan_kernel(uint32_t* __restrict__ counts,
                                                                    int nb,
                                                                    uint32_t* __restrict__ totals) {
  __shared__ uint32_t wsum[4];
  const int t = threadIdx.x;
  uint32_t* row = counts + (int64_t)blockIdx.x * nb;
  const int per = (nb + kSortThreads - 1) / kSortThreads;
  const int j0 = min(t * per, nb), j1 = min(j0 + per, nb);
  uint32_t sum = 0u;
  for (int j = j0; j < j1; ++j) sum += row[j];
  const uint32_t incl = block_scan_incl(sum, wsum);
  uint32_t run = incl - sum;
  for (int j = j0; j < j1; ++j) {
    const uint32_t v = row[j];
    row[j] = run;
    run += v;
  }
  if (t == kSortThreads - 1) totals[blockIdx.x] = incl;
}

// Segmented form (tile j in segment tseg[j], each segment's tiles
// consecutive): per digit, the exclusive prefix of the tile counts within
// each segment, in place, and each segment's digit total at
// totals[segment * kRadix + digit].  The block walks its digit's row in
// rounds of 256 x kSegIPT consecutive tiles, thread t taking kSegIPT
// consecutive ones (all their loads issued together): a segmented scan of
// each thread's tiles, then over the threads of (sum since the last segment
// start, start seen), plus the running sum carried from the previous round
// where no segment starts before the tile.  (One run of ~150 tiles per
// thread took 0.25 ms per pass for a 64-snapshot batch, each load of a wave
// touching 64 cache lines; one tile per thread and round, 153 rounds of
// dependent load latency, took as long.)
constexpr int kSegIPT = 8;
__global__ __launch_bounds__(kSortThreads) void sort_rowscan_seg_kernel(
    uint32_t* __restrict__ counts, int nb, const int* __restrict__ tseg,
    uint32_t* __restrict__ totals) {
  __shared__ uint32_t wv[4], wincl[4];
  __shared__ int wf[4];
  const int t = threadIdx.x, l = t & 63, w = t >> 6;
  uint32_t* row = counts + (int64_t)blockIdx.x * nb;
  uint32_t carry = 0u;   // sum since the last segment start, before this round
  for (int j0 = 0; j0 < nb; j0 += kSortThreads * kSegIPT) {
    const int base = j0 + kSegIPT * t;
    uint32_t c[kSegIPT];
    int sg[kSegIPT];
#pragma unroll
    for (int k = 0; k < kSegIPT; ++k) {
      const int j = base + k;
      c[k] = j < nb ? row[j] : 0u;
      sg[k] = j < nb ? tseg[j] : -1;
    }
    const int sprev = (base > 0 && base <= nb) ? tseg[base - 1] : -2;
    const int snext = base + kSegIPT < nb ? tseg[base + kSegIPT] : -3;
    // the thread's own tiles: sums since the last start within them
    uint32_t loc[kSegIPT];
    uint32_t v = 0u;
    int fl = 0;
    int first = kSegIPT;   // first of the thread's tiles that starts a segment
#pragma unroll
    for (int k = 0; k < kSegIPT; ++k) {
      const bool in = base + k < nb;
      const bool st = in && (base + k == 0 || sg[k] != (k ? sg[k - 1] : sprev));
      if (st) {
        v = 0u;
        fl = 1;
        if (first == kSegIPT) first = k;
      }
      v += c[k];
      loc[k] = v;
    }
    // inclusive segmented scan of (v, fl) over the threads
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t pv = __shfl_up(v, o, 64);
      const int pf = __shfl_up(fl, o, 64);
      if (l >= o) {
        if (!fl) v += pv;
        fl |= pf;
      }
    }
    if (l == 63) {
      wv[w] = v;
      wf[w] = fl;
    }
    __syncthreads();
    bool seen = fl != 0;
    if (!seen) {
      for (int q = w - 1; q >= 0; --q) {
        v += wv[q];
        if (wf[q]) {
          seen = true;
          break;
        }
      }
    }
    if (!seen) v += carry;
    if (l == 63) wincl[w] = v;
    __syncthreads();
    // the running sum before the thread's first tile
    const uint32_t up = __shfl_up(v, 1, 64);
    const uint32_t pre = l > 0 ? up : (w > 0 ? wincl[w - 1] : carry);
#pragma unroll
    for (int k = 0; k < kSegIPT; ++k) {
      const int j = base + k;
      if (j < nb) {
        const uint32_t val = k < first ? pre + loc[k] : loc[k];
        row[j] = val - c[k];
        if (j + 1 == nb || (k + 1 < kSegIPT ? sg[k + 1] : snext) != sg[k])
          totals[(int64_t)sg[k] * kRadix + blockIdx.x] = val;
      }
    }
    carry = wincl[3];
    __syncthreads();   // wv, wf, wincl are rewritten by the next round
  }
}

// Stable scatter of one tile by digit.  vin == NULL: the values are packed
// from the source (first pass).  kout and perm may be NULL on the last pass.
// Segmented sort (tb != NULL): tile b belongs to segment tseg[b], whose digit
// totals are totals[tseg[b] * kRadix + d] and whose items start at
// segoff[tseg[b]].
__global__ __launch_bounds__(kSortThreads) void sort_scatter_kernel(
    KeySrc S, const uint32_t* __restrict__ kin, const float4* __restrict__ vin, int64_t n,
    const int64_t* __restrict__ tb, const int* __restrict__ tseg,
    const int64_t* __restrict__ segoff, int shift, const uint32_t* __restrict__ rowpre,
    const uint32_t* __restrict__ totals, int nb, uint32_t* __restrict__ kout,
    float4* __restrict__ vout, int* __restrict__ perm) {
  __shared__ uint32_t hist[4][kRadix];   // per-wave running counts, then wave prefixes
  __shared__ uint32_t dstart[kRadix];    // tile-local start of each digit
  __shared__ uint32_t gpos[kRadix];      // output position of the tile's first item of each digit
  __shared__ uint32_t wsum[4];
  __shared__ uint32_t skey[kSortTile];
  __shared__ float4 sval[kSortTile];
  const int t = threadIdx.x, w = t >> 6, l = t & 63;
  const TileRange tr = sort_tile(tb, n);
  const int sg = tseg ? tseg[blockIdx.x] : 0;
  const uint32_t kb = tseg ? (uint32_t)sg * S.ncell : 0u;
  for (int j = t; j < 4 * kRadix; j += kSortThreads) (&hist[0][0])[j] = 0u;
  uint32_t key[kSortIPT];
  float4 val[kSortIPT];
#pragma unroll
  for (int i = 0; i < kSortIPT; ++i) {
    const int64_t idx = sort_item(tr, i);
    key[i] = 0u;
    val[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (idx < tr.end) {
      key[i] = kin[idx];
      val[i] = vin ? vin[idx] : source_value(S, idx);
    }
  }
  __syncthreads();
  // rank within the wave: items of earlier slots, then lower lanes of this slot
  const uint64_t lt = (1ull << l) - 1ull;
  uint32_t rank[kSortIPT];
#pragma unroll
  for (int i = 0; i < kSortIPT; ++i) {
    const bool live = sort_item(tr, i) < tr.end;
    const uint32_t d = ((key[i] - kb) >> shift) & (kRadix - 1);
    const uint64_t mk = match_digit(d, live);
    uint32_t before = 0u;
    if (live) before = hist[w][d];
    if (live && (mk & lt) == 0ull) hist[w][d] = before + (uint32_t)__popcll(mk);
    rank[i] = before + (uint32_t)__popcll(mk & lt);
  }
  __syncthreads();
  const uint32_t c0 = hist[0][t], c1 = hist[1][t], c2 = hist[2][t], c3 = hist[3][t];
  const uint32_t tot = c0 + c1 + c2 + c3;
  const uint32_t gt = totals[(int64_t)sg * kRadix + t];
  const uint32_t pre = rowpre[(int64_t)t * nb + blockIdx.x];
  __syncthreads();
  hist[0][t] = 0u;
  hist[1][t] = c0;
  hist[2][t] = c0 + c1;
  hist[3][t] = c0 + c1 + c2;
  const uint32_t ds = block_scan_incl(tot, wsum) - tot;
  const uint32_t gb = block_scan_incl(gt, wsum) - gt + (segoff ? (uint32_t)segoff[sg] : 0u);
  dstart[t] = ds;
  gpos[t] = gb + pre;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kSortIPT; ++i) {
    if (sort_item(tr, i) < tr.end) {
      const uint32_t d = ((key[i] - kb) >> shift) & (kRadix - 1);
      const uint32_t lp = dstart[d] + hist[w][d] + rank[i];
      skey[lp] = key[i];
      sval[lp] = val[i];
    }
  }
  __syncthreads();
  const int cnt = (int)(tr.end - tr.t0);
  for (int j = t; j < cnt; j += kSortThreads) {
    const uint32_t k = skey[j];
    const uint32_t d = ((k - kb) >> shift) & (kRadix - 1);
    const uint32_t pos = gpos[d] + (uint32_t)j - dstart[d];
    const float4 v = sval[j];
    if (kout) kout[pos] = k;
    vout[pos] = v;
    if (perm) perm[pos] = __float_as_int(v.w);
  }
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Ordering buffer: first-pass keys | keys A | keys B | values A | values B |
// perm | per-point f64 | tile counts | digit totals (per segment) | tile
// bounds and segments (segmented sorts).
struct OrderLayout {
  size_t k0, ka, kb, va, vb, perm, energy, counts, totals, tb, ts, total;
};

static int sort_tiles(int64_t n) { return (int)((n + kSortTile - 1) / kSortTile); }

static OrderLayout order_layout(int64_t n, int nb, int nseg) {
  OrderLayout L;
  const size_t kbytes = align256((size_t)n * 4), vbytes = align256((size_t)n * 16);
  L.k0 = 0;
  L.ka = L.k0 + kbytes;
  L.kb = L.ka + kbytes;
  L.va = L.kb + kbytes;
  L.vb = L.va + vbytes;
  L.perm = L.vb + vbytes;
  L.energy = L.perm + kbytes;
  L.counts = L.energy + align256((size_t)n * 8);
  L.totals = L.counts + align256((size_t)nb * kRadix * 4);
  L.tb = L.totals + align256((size_t)nseg * kRadix * 4);
  L.ts = L.tb + (nseg > 1 ? align256((size_t)nb * 16) : 0);
  L.total = L.ts + (nseg > 1 ? align256((size_t)nb * 4) : 0);
  return L;
}

static int ensure_order(Context* c, size_t bytes) {
  if (bytes <= c->order_bytes) return 0;
  if (c->order) {
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return check_hip(c, e, "hipDeviceSynchronize (order grow)");
    (void)hipFree(c->order);
    c->order = nullptr;
    c->order_bytes = 0;
  }
  const size_t want = bytes + bytes / 4;
  hipError_t e = hipMalloc(&c->order, want);
  if (e != hipSuccess) return check_hip(c, e, "hipMalloc order buffer");
  c->order_bytes = want;
  return 0;
}

struct SortOut {
  const float4* values;
  const uint32_t* keys;
  const int* perm;
};

// Sort the n items of S by key (nkeys distinct keys at most); writes the
// sorted values, and when asked the sorted keys and the permutation (values
// packed with the input index).  With off_host (S.off's host copy, more than
// one snapshot) the sort is segmented: the snapshots' items stay in their
// ranges, tiles never straddle two, and the passes cover only the S.ncell
// cells of a snapshot (two 8-bit passes for up to 65536 cells, where the
// (snapshot, cell) key of 64 snapshots needs three).
static int sort_source(Context* c, const KeySrc& S, int64_t n, uint64_t nkeys, bool want_keys,
                       bool want_perm, hipStream_t s, SortOut* out,
                       const int64_t* off_host = nullptr) {
  out->values = nullptr;
  out->keys = nullptr;
  out->perm = nullptr;
  if (n > (int64_t)0x7fffffff) return set_error(c, BFE_ELIMIT, "order: n = %lld > 2^31 - 1", (long long)n);
  if (nkeys > 0xffffffffull) return set_error(c, BFE_ELIMIT, "order: too many table cells");
  const bool segmented = off_host && S.off && S.nsnap > 1;
  std::vector<int64_t> tbv;
  std::vector<int> tsv;
  if (segmented) {
    for (int q = 0; q < S.nsnap; ++q)
      for (int64_t a = off_host[q]; a < off_host[q + 1]; a += kSortTile) {
        tbv.push_back(a);
        tbv.push_back(std::min(a + (int64_t)kSortTile, off_host[q + 1]));
        tsv.push_back(q);
      }
    if ((int64_t)tsv.size() > (int64_t)0x7fffffff)
      return set_error(c, BFE_ELIMIT, "order: too many tiles");
  }
  const uint64_t span = segmented ? (uint64_t)S.ncell : nkeys;
  int bits = 1;
  while (bits < 32 && (1ull << bits) < span) ++bits;
  const int passes = (bits + 7) / 8;
  const int nb = segmented ? (int)tsv.size() : sort_tiles(n);
  const OrderLayout L = order_layout(n, nb, segmented ? S.nsnap : 1);
  int rc = ensure_order(c, L.total);
  if (rc) return rc;
  const int64_t* tb = nullptr;
  const int* ts = nullptr;
  if (segmented && nb > 0) {
    // the tile table is read by this call's kernels only; earlier work on s
    // may still read the buffer
    BFE_HIP(c, hipStreamSynchronize(s));
    int64_t* tbd = (int64_t*)((char*)c->order + L.tb);
    int* tsd = (int*)((char*)c->order + L.ts);
    BFE_HIP(c, hipMemcpy(tbd, tbv.data(), tbv.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    BFE_HIP(c, hipMemcpy(tsd, tsv.data(), tsv.size() * sizeof(int), hipMemcpyHostToDevice));
    tb = tbd;
    ts = tsd;
  }
  char* base = (char*)c->order;
  uint32_t* k0 = (uint32_t*)(base + L.k0);
  uint32_t* kbuf[2] = {(uint32_t*)(base + L.ka), (uint32_t*)(base + L.kb)};
  float4* vbuf[2] = {(float4*)(base + L.va), (float4*)(base + L.vb)};
  int* perm = (int*)(base + L.perm);
  uint32_t* counts = (uint32_t*)(base + L.counts);
  uint32_t* totals = (uint32_t*)(base + L.totals);
  for (int p = 0; p < passes && n > 0 && nb > 0; ++p) {
    const int shift = 8 * p;
    const bool last = p == passes - 1;
    const uint32_t* kin = p == 0 ? (const uint32_t*)k0 : kbuf[(p - 1) & 1];
    const float4* vin = p == 0 ? nullptr : vbuf[(p - 1) & 1];
    uint32_t* kout = (!last || want_keys) ? kbuf[p & 1] : nullptr;
    hipLaunchKernelGGL(sort_count_kernel, dim3((unsigned)nb), dim3(kSortThreads), 0, s, S,
                       p == 0 ? nullptr : kin, k0, n, tb, ts, shift, counts, nb);
    BFE_HIP(c, hipGetLastError());
    if (segmented)
      hipLaunchKernelGGL(sort_rowscan_seg_kernel, dim3(kRadix), dim3(kSortThreads), 0, s, counts,
                         nb, ts, totals);
    else
      hipLaunchKernelGGL(sort_rowscan_kernel, dim3(kRadix), dim3(kSortThreads), 0, s, counts, nb,
                         totals);
    BFE_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(sort_scatter_kernel, dim3((unsigned)nb), dim3(kSortThreads), 0, s, S, kin,
                       vin, n, tb, ts, segmented ? S.off : nullptr, shift, counts, totals, nb,
                       kout, vbuf[p & 1], (last && want_perm) ? perm : nullptr);
    BFE_HIP(c, hipGetLastError());
  }
  out->values = vbuf[(passes - 1) & 1];
  if (want_keys) out->keys = kbuf[(passes - 1) & 1];
  if (want_perm) out->perm = perm;
  return 0;
}

uint32_t order_cell_count(const FieldTables& T, int nsnap) {
  return (uint32_t)(T.slg.numr - 1) * theta_bins(T.slg.numr, nsnap);
}

int order_points(Context* c, const FieldTables& T, const float* x, const float* y, const float* z,
                 int64_t n, hipStream_t s, const int** perm_out, const int64_t* off_dev,
                 int nsnap, const uint32_t** sorted_out, const float4** points_out,
                 const int64_t* off_host) {
  if (perm_out) *perm_out = nullptr;
  if (sorted_out) *sorted_out = nullptr;
  if (points_out) *points_out = nullptr;
  const int ns = off_dev ? nsnap : 1;
  const uint32_t tb = theta_bins(T.slg.numr, ns);
  const uint64_t ncell = (uint64_t)(T.slg.numr - 1) * tb;
  if (ncell > 0xffffffffull) return set_error(c, BFE_ELIMIT, "order: too many table cells");
  KeySrc S{x, y, z, nullptr, off_dev, ns, (uint32_t)ncell, 1,
           key_geom(T.eg, T.slg, c->h_xi.size() >= 2 ? c->h_xi.data() : nullptr),
           x, y, z, n};
  S.g.tbins = tb;
  SortOut o;
  int rc = sort_source(c, S, n, ncell * (uint64_t)ns, sorted_out != nullptr, perm_out != nullptr,
                       s, &o, off_dev ? off_host : nullptr);
  if (rc) return rc;
  if (perm_out) *perm_out = o.perm;
  if (sorted_out) *sorted_out = o.keys;
  if (points_out) *points_out = o.values;
  return 0;
}

int order_points2(Context* c, const FieldTables& T, const float* x, const float* y,
                  const float* z, int64_t n1, const float* x2, const float* y2, const float* z2,
                  int64_t n2, hipStream_t s, const float4** points_out) {
  *points_out = nullptr;
  const uint32_t tb = theta_bins(T.slg.numr, 1);
  const uint64_t ncell = (uint64_t)(T.slg.numr - 1) * tb;
  if (ncell > 0xffffffffull) return set_error(c, BFE_ELIMIT, "order: too many table cells");
  KeySrc S{x, y, z, nullptr, nullptr, 1, (uint32_t)ncell, 1,
           key_geom(T.eg, T.slg, c->h_xi.size() >= 2 ? c->h_xi.data() : nullptr), x2, y2, z2, n1};
  S.g.tbins = tb;
  SortOut o;
  int rc = sort_source(c, S, n1 + n2, ncell, false, false, s, &o);
  if (rc) return rc;
  *points_out = o.values;
  return 0;
}

int order_particles(Context* c, int kind, const float* x, const float* y, const float* z,
                    const float* m, int64_t n, const int64_t* off_dev, int nsnap, hipStream_t s,
                    const float4** sorted_out, const int64_t* off_host) {
  *sorted_out = nullptr;
  const EOFGeom eg = eof_geom(c);
  const SLGeom sg = sl_geom(c);
  const int ns = off_dev ? nsnap : 1;
  const uint32_t tb = kind == 0 ? 1u : theta_bins(sg.numr, ns);
  const uint64_t ncell = kind == 0 ? (uint64_t)eg.numx * eg.numy : (uint64_t)(sg.numr - 1) * tb;
  if (ncell > 0xffffffffull) return set_error(c, BFE_ELIMIT, "order: too many table cells");
  KeySrc S{x, y, z, m, off_dev, ns, (uint32_t)ncell, kind,
           key_geom(eg, sg, c->h_xi.size() >= 2 ? c->h_xi.data() : nullptr), x, y, z, n};
  S.g.tbins = tb;
  SortOut o;
  int rc = sort_source(c, S, n, ncell * (uint64_t)(off_dev ? nsnap : 1), false, false, s, &o,
                       off_dev ? off_host : nullptr);
  if (rc) return rc;
  *sorted_out = o.values;
  return 0;
}

double* order_energy_buffer(Context* c, int64_t n) {
  return (double*)((char*)c->order + order_layout(n, sort_tiles(n), 1).energy);
}

void order_free(Context* c) {
  if (c->order) (void)hipFree(c->order);
  c->order = nullptr;
  c->order_bytes = 0;
}

}  // namespace bfe
