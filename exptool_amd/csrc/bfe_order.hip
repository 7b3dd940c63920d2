// Cell ordering of evaluation points.
//
// The folded field tables (SL 1.8 MB, EOF 7.4 MB at the workload size) do
// not fit in one XCD's 4 MB L2, and points in input order send every lane of
// a wave to a different table cell.  Sorting the points by table cell lets a
// wave's lanes share cells and, with the XCD-aware block mapping, each XCD
// work on one contiguous range of cells.
//
// Key: the EOF (ix, iy) cell for points inside the disc grid, else
// numx * numy + the SL radial cell.  The keys are sorted with rocPRIM's radix
// sort (values: the point indices); a counting sort with global atomics was
// slower here because thousands of points share the busiest cells.  No
// result depends on the order (evaluation and orbit integration are per
// point, and energy sums are taken in input order).  The sort issues
// hipMemsetAsync, so it is never placed inside a captured graph.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "bfe_field.hpp"

namespace bfe {

// Onesweep radix sort at every size: rocPRIM's default switches to a merge
// sort up to 2^20 items, which took several times longer for our keys.
using SortConfig = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                              rocprim::default_config, 0>;

__device__ __forceinline__ uint32_t cell_key(const FieldTables& T, float xf, float yf, float zf,
                                             uint32_t neof) {
  const double x = xf, y = yf, z = zf;
  const double R = sqrt(x * x + y * y);
  const EOFGeom& g = T.eg;
  const double X = (r_to_xi(R, g.cmap, g.ascale) - g.xmin) / g.dx;
  const double Y = (z_to_y(z, g.hscale) - g.ymin) / g.dy;
  if (X >= 0.0 && X < (double)g.numx && Y >= 0.0 && Y < (double)g.numy)
    return (uint32_t)X * (uint32_t)g.numy + (uint32_t)Y;
  const SLGeom& s = T.slg;
  const double xi = r_to_xi(sqrt(R * R + z * z), s.cmap, s.scale);
  const double f = floor((xi - s.xi[0]) / (s.xi[1] - s.xi[0]));
  int i = 0;
  if (f == f) i = f < 0.0 ? 0 : (f > (double)(s.numr - 2) ? s.numr - 2 : (int)f);
  return neof + (uint32_t)i;
}

// Snapshot of point i for concatenated snapshots: the last s with off[s] <= i.
__device__ __forceinline__ int snapshot_of(const int64_t* __restrict__ off, int nsnap, int64_t i) {
  int lo = 0, hi = nsnap - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= i) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// key = snapshot * nkeys + cell (snapshot 0 without offsets); with packed,
// also (x, y, z, index bits) for a gather in key order
__global__ __launch_bounds__(256) void order_key_kernel(
    const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
    int64_t n, FieldTables T, uint32_t neof, uint32_t nkeys, const int64_t* __restrict__ off,
    int nsnap, uint32_t* __restrict__ keys, float4* __restrict__ packed) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float xf = x[i], yf = y[i], zf = z[i];
  const uint32_t snap = off ? (uint32_t)snapshot_of(off, nsnap, i) : 0u;
  keys[i] = snap * nkeys + cell_key(T, xf, yf, zf, neof);
  if (packed) packed[i] = make_float4(xf, yf, zf, __int_as_float((int)i));
}

__global__ __launch_bounds__(256) void gather_points_kernel(const float4* __restrict__ in,
                                                            const int* __restrict__ perm, int64_t n,
                                                            float4* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) out[j] = in[perm[j]];
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

struct OrderLayout {
  size_t keys, sorted, perm, energy, pin, pout, temp, total;
};

static OrderLayout order_layout(int64_t n, size_t temp_bytes) {
  OrderLayout L;
  L.keys = 0;
  L.sorted = L.keys + align256((size_t)n * 4);
  L.perm = L.sorted + align256((size_t)n * 4);
  L.energy = L.perm + align256((size_t)n * 4);
  L.pin = L.energy + align256((size_t)n * 8);
  L.pout = L.pin + align256((size_t)n * 16);
  L.temp = L.pout + align256((size_t)n * 16);
  L.total = L.temp + align256(temp_bytes);
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

uint32_t order_cell_count(const FieldTables& T) {
  return (uint32_t)T.eg.numx * (uint32_t)T.eg.numy + (uint32_t)(T.slg.numr - 1);
}

int order_points(Context* c, const FieldTables& T, const float* x, const float* y, const float* z,
                 int64_t n, hipStream_t s, const int** perm_out, const int64_t* off_dev,
                 int nsnap, const uint32_t** sorted_out, const float4** points_out) {
  *perm_out = nullptr;
  if (sorted_out) *sorted_out = nullptr;
  if (points_out) *points_out = nullptr;
  if (n > (int64_t)0x7fffffff) return set_error(c, BFE_ELIMIT, "order: n = %lld > 2^31 - 1", (long long)n);
  const uint64_t ncell = (uint64_t)T.eg.numx * T.eg.numy + (uint64_t)(T.slg.numr - 1);
  const uint64_t nkeys = ncell * (uint64_t)(off_dev ? nsnap : 1);
  if (nkeys > 0xffffffffull) return set_error(c, BFE_ELIMIT, "order: too many table cells");
  unsigned bits = 1;
  while (bits < 32 && (1ull << bits) < nkeys) ++bits;
  const int ni = (int)n;
  rocprim::counting_iterator<int> iota(0);
  size_t temp_bytes = 0;
  BFE_HIP(c, rocprim::radix_sort_pairs<SortConfig>(nullptr, temp_bytes, (const uint32_t*)nullptr,
                                       (uint32_t*)nullptr, iota, (int*)nullptr, ni, 0u, bits, s));
  const OrderLayout L = order_layout(n, temp_bytes);
  int rc = ensure_order(c, L.total);
  if (rc) return rc;
  char* base = (char*)c->order;
  uint32_t* keys = (uint32_t*)(base + L.keys);
  uint32_t* sorted = (uint32_t*)(base + L.sorted);
  int* perm = (int*)(base + L.perm);
  const uint32_t neof = (uint32_t)T.eg.numx * (uint32_t)T.eg.numy;
  float4* pin = points_out ? (float4*)(base + L.pin) : nullptr;
  hipLaunchKernelGGL(order_key_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, y, z,
                     n, T, neof, (uint32_t)ncell, off_dev, nsnap, keys, pin);
  BFE_HIP(c, hipGetLastError());
  BFE_HIP(c, rocprim::radix_sort_pairs<SortConfig>((void*)(base + L.temp), temp_bytes, keys, sorted, iota, perm,
                                       ni, 0u, bits, s));
  *perm_out = perm;
  if (sorted_out) *sorted_out = sorted;
  if (points_out) {
    float4* pout = (float4*)(base + L.pout);
    hipLaunchKernelGGL(gather_points_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pin,
                       perm, n, pout);
    BFE_HIP(c, hipGetLastError());
    *points_out = pout;
  }
  return 0;
}

// ---- accumulate: a sorted, packed copy of the particles ---------------------

// key = snapshot * ncell + accumulate cell: kind 0 the EOF (ix, iy) cell of
// (R = sqrt(x^2 + y^2 + 1e-10), z) as eof.return_bins; kind 1 the SL radial
// cell of r = max(|x|, 1e-10).
__global__ __launch_bounds__(256) void pack_key_kernel(
    const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
    const float* __restrict__ m, int64_t n, int kind, EOFGeom eg, SLGeom sg, uint32_t ncell,
    const int64_t* __restrict__ off, int nsnap, uint32_t* __restrict__ keys,
    float4* __restrict__ packed) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float xf = x[i], yf = y[i], zf = z[i];
  packed[i] = make_float4(xf, yf, zf, m[i]);
  const double xd = xf, yd = yf, zd = zf;
  uint32_t cell;
  if (kind == 0) {
    const double R = sqrt(xd * xd + yd * yd + 1.e-10);
    const EOFCell ce = eof_locate(R, zd, eg.cmap, eg.ascale, eg.hscale, eg.xmin, eg.dx, eg.ymin,
                                  eg.dy, eg.numx, eg.numy);
    cell = (uint32_t)ce.ix * (uint32_t)eg.numy + (uint32_t)ce.iy;
  } else {
    double r = sqrt(xd * xd + yd * yd + zd * zd);
    r = (r == r) ? (r > 1.0e-10 ? r : 1.0e-10) : 1.0e-10;
    cell = (uint32_t)sl_locate(r, sg.cmap, sg.scale, sg.xi, sg.numr).i;
  }
  const uint32_t snap = off ? (uint32_t)snapshot_of(off, nsnap, i) : 0u;
  keys[i] = snap * ncell + cell;
}

__global__ __launch_bounds__(256) void gather_packed_kernel(const float4* __restrict__ in,
                                                            const int* __restrict__ perm, int64_t n,
                                                            float4* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) out[j] = in[perm[j]];
}

// The keys are sorted with the point indices as values (rocPRIM's fast path
// for 4-byte values; 16-byte values took several times longer), then the
// packed particles are gathered in sorted order.
int order_particles(Context* c, int kind, const float* x, const float* y, const float* z,
                    const float* m, int64_t n, const int64_t* off_dev, int nsnap, hipStream_t s,
                    const float4** sorted_out) {
  *sorted_out = nullptr;
  if (n > (int64_t)0x7fffffff) return set_error(c, BFE_ELIMIT, "order: n = %lld > 2^31 - 1", (long long)n);
  const EOFGeom eg = eof_geom(c);
  const SLGeom sg = sl_geom(c);
  const uint64_t ncell = kind == 0 ? (uint64_t)eg.numx * eg.numy : (uint64_t)(sg.numr - 1);
  const uint64_t nkeys = ncell * (uint64_t)(off_dev ? nsnap : 1);
  if (nkeys > 0xffffffffull) return set_error(c, BFE_ELIMIT, "order: too many table cells");
  unsigned bits = 1;
  while (bits < 32 && (1ull << bits) < nkeys) ++bits;
  const int ni = (int)n;
  rocprim::counting_iterator<int> iota(0);
  size_t temp_bytes = 0;
  BFE_HIP(c, rocprim::radix_sort_pairs<SortConfig>(nullptr, temp_bytes, (const uint32_t*)nullptr,
                                       (uint32_t*)nullptr, iota, (int*)nullptr, ni, 0u, bits, s));
  const size_t kb = align256((size_t)n * 4), pb = align256((size_t)n * 16);
  const size_t total = 3 * kb + 2 * pb + align256(temp_bytes);
  int rc = ensure_order(c, total);
  if (rc) return rc;
  char* base = (char*)c->order;
  uint32_t* keys = (uint32_t*)base;
  uint32_t* sorted = (uint32_t*)(base + kb);
  int* perm = (int*)(base + 2 * kb);
  float4* pin = (float4*)(base + 3 * kb);
  float4* pout = (float4*)(base + 3 * kb + pb);
  void* temp = base + 3 * kb + 2 * pb;
  const unsigned grid = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(pack_key_kernel, dim3(grid), dim3(256), 0, s, x, y, z, m, n, kind, eg, sg,
                     (uint32_t)ncell, off_dev, nsnap, keys, pin);
  BFE_HIP(c, hipGetLastError());
  BFE_HIP(c, rocprim::radix_sort_pairs<SortConfig>(temp, temp_bytes, keys, sorted, iota, perm, ni, 0u, bits, s));
  hipLaunchKernelGGL(gather_packed_kernel, dim3(grid), dim3(256), 0, s, pin, perm, n, pout);
  BFE_HIP(c, hipGetLastError());
  *sorted_out = pout;
  return 0;
}

double* order_energy_buffer(Context* c, int64_t n) {
  return (double*)((char*)c->order + order_layout(n, 0).energy);
}

void order_free(Context* c) {
  if (c->order) (void)hipFree(c->order);
  c->order = nullptr;
  c->order_bytes = 0;
}

}  // namespace bfe
