// Internal state of the BFE library: the context, device table layouts and
// the error helpers shared by the translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <map>
#include <string>
#include <utility>
#include <vector>

#include "bfe_math.hpp"
#include "../../include/exptool_bfe.h"

namespace bfe {

// Device limits of the kernels (checked by the host before every launch).
constexpr int kMaxL = 16;          // SL lmax
constexpr int kMaxM = 16;          // EOF mmax
constexpr int kMaxLM = (kMaxL + 1) * (kMaxL + 2) / 2;
// Layout of Context::sl_lmc (index lm_index(l, m)):
//   [kLmcY + lm]  -4 pi N_lm (x sqrt 2 for m > 0), N_lm of factorial_return
//   [kLmcA + lm]  (2l - 1)/(l - m)      P_l^m = x A P_{l-1}^m - B P_{l-2}^m
//   [kLmcB + lm]  (l + m - 1)/(l - m)
//   [kLmcD + m]   (-1)^m (2m - 1)!!     P_m^m = D_m (1 - x^2)^{m/2}
//   [kLmcL + lm]  l, [kLmcS + lm] l + m  (as doubles: scalar operands of the
//                 derivative recurrence instead of per-lane conversions)
//   [kLmcNA + lm], [kLmcNB + lm]  the same recurrence for Q_l^m = -4 pi N_lm
//                 P_l^m:  Q_l^m = x A N_lm/N_{l-1,m} Q_{l-1}^m
//                                 - B N_lm/N_{l-2,m} Q_{l-2}^m
//   [kLmcN0 + m]  -4 pi N_mm (-1)^m (2m - 1)!!   Q_m^m = N0_m (1 - x^2)^{m/2}
constexpr int kLmcY = 0, kLmcA = kMaxLM, kLmcB = 2 * kMaxLM, kLmcD = 3 * kMaxLM;
constexpr int kLmcL = 3 * kMaxLM + kMaxL + 1, kLmcS = kLmcL + kMaxLM;
constexpr int kLmcNA = kLmcS + kMaxLM, kLmcNB = kLmcNA + kMaxLM, kLmcN0 = kLmcNB + kMaxLM;
constexpr int kLmcSize = kLmcN0 + kMaxL + 1;

// Row length of the padded deposit tables: radial orders rounded up to a
// multiple of 48, so that for every chunk length of the lane shapes (2, 4, 8,
// 16, 24: divisors of 48) all chunks lie inside the row and start on a
// 16-byte boundary (chunks of 2: 8 bytes).
__host__ __device__ constexpr int dep_np(int n) { return (n + 47) / 48 * 48; }

// Lane shapes (slots per lane, radial orders per slot) instantiated for the
// cell-ordered accumulate kernels (bfe_accumulate.hip), selectable by index
// with the tuning knobs eof_dep_shape / sl_dep_shape.
constexpr int kEofDepShapes = 8;
constexpr int kSlDepShapes = 7;

// Folded tables (layouts documented in bfe_field.hpp): coefficients summed
// into the basis tables once per coefficient set.
struct SLFolded {
  float* d = nullptr;
  size_t floats = 0;
  bool valid = false;
};

struct EOFFolded {
  float* d = nullptr;
  size_t floats = 0;
  bool valid = false;
};

// Analytic base of the MLP field (bfe_mlp_set_base): at most kMaxMN
// Miyamoto-Nagai discs; values nmn, (M, a, b) x nmn, Mh, ah, phi0.
constexpr int kMaxMN = 4;
constexpr int kMLPBaseMax = 4 + 3 * kMaxMN;

// MLP surrogate state (float32 masters, bf16 MFMA operand copies, Adam moments
// and the per-batch training buffers).
struct MLPState {
  float* params = nullptr;
  float* grads = nullptr;
  float* m = nullptr;
  float* v = nullptr;
  __bf16* Wp = nullptr;
  __bf16* Wt = nullptr;
  float* feat = nullptr;
  float* Z = nullptr;
  float* gZ = nullptr;
  float* gY = nullptr;
  float* dwpart = nullptr;
  double* loss_part = nullptr;
  int64_t bcap = 0;
  int64_t step = 0;
  double cfg[11] = {1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0};
  double base[kMLPBaseMax] = {0};   // no base: nmn 0, Mh 0, phi0 0
  int conservative = 0;             // forces as -grad of the potential output
  double lw[4] = {1, 1, 1, 1};      // training loss weight per output
};

struct Context {
  int device = 0;
  std::string err;

  // ---- SL tables ---------------------------------------------------------
  bool sl_ok = false;
  int lmax = 0, nmax = 0, numr = 0, cmap = 0;
  double scale = 1.0;
  double* sl_xi = nullptr;     // numr
  double* sl_p0 = nullptr;     // numr
  double* sl_d0 = nullptr;     // numr
  double* sl_fold64 = nullptr; // [3][lmax+1][numr][nmax] float64: ef/sqrt(ev), ef sqrt(ev), ef p0/sqrt(ev)
  float* sl_e32 = nullptr;     // [numr][(lmax+1)*nmax]  ef/sqrt(ev), node-major
  float* sl_dep32 = nullptr;   // [numr][lmax+1][dep_np(nmax)] the same, rows zero-padded
  double* sl_lmc = nullptr;    // per-(l, m) constants of the accumulate recurrences
  std::vector<double> h_xi;    // host copy (min xi for the field guard)

  // ---- EOF tables --------------------------------------------------------
  bool eof_ok = false;
  bool eof_dens = false;
  int mmax = 0, norder = 0, numx = 0, numy = 0, ecmap = 0;
  double ascale = 1.0, hscale = 1.0, xmin = 0, dx = 1, ymin = 0, dy = 1;
  int eof_K = 0;               // (2*mmax+1)*norder rows of the node-major tables
  // fold inputs pot, rf, zf, dens: [K][node], float32 (the folded field is
  // float32 too; the fold sums in float64)
  float* eof_fold32[4] = {nullptr, nullptr, nullptr, nullptr};
  // the m = 0 cos fold inputs of the edge nodes (eof_edge_index) in float64:
  // [4][norder][eof_edge_count], for the float64 edge blocks of the folded field
  double* eof_edge64 = nullptr;
  float* eof_pot32 = nullptr;  // [node][K] potential for accumulate
  float* eof_dep32 = nullptr;  // [node][2 mmax + 1][eof_dep_rp] the same, rows zero-padded
  int eof_dep_rp = 0;          // row length of eof_dep32: norder rounded up to the
                               // lane shape's chunk length (eof_dep_prepare)
  // bf16 MFMA mode of the EOF accumulate (BFE_ACC_BF16): the potential table
  // as bf16 pieces hi, lo, each [2 mmax + 1][norder][eof_mfma_nodes] (nodes
  // padded to the product's K block), built on first use
  // (eof_mfma_prepare); eof_mfma_np: norder rounded up to 16
  __bf16* eof_mfmaT = nullptr;
  int eof_mfma_np = 0;
  int64_t eof_mfma_nodes = 0;

  // ---- folded tables: slot 0 = batch evaluators, slot 1 = frozen field ----
  SLFolded slf[2];
  EOFFolded eoff[2];
  double field_halofac = 1.0;
  int field_flags = 0;

  // ---- MLP surrogate -----------------------------------------------------
  MLPState mlp;

  // ---- scratch (grown on demand, freed by bfe_destroy) -------------------
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  // device address of the static array of the allocation-free accumulate
  // (bfe_accumulate.hip), taken once at bfe_create; NULL if unavailable
  double* noalloc_scratch = nullptr;
  // work memory of the EOF bf16 MFMA mode (cell sums, node pieces, partials)
  void* mfma_work = nullptr;
  size_t mfma_work_bytes = 0;

  // ---- tuning (bfe_set_tuning) --------------------------------------------
  int64_t eof_acc_blocks = 2048;   // EOF accumulate: blocks (single) / segment budget (batch)
  int64_t sl_acc_blocks = 2048;    // SL accumulate: the same
  // SL accumulate beside the EOF one (the fused disc + halo calls): fewer
  // blocks leave CUs to the EOF branch, the longer one (fused call
  // 427 -> 422 us at 1M + 1M against 2 048 blocks)
  int64_t sl_acc_blocks_beside = 1024;
  int64_t eof_seg_tiles = 0;       // batch accumulate: tiles per segment, 0 = from the budget
  int64_t sl_seg_tiles = 0;
  int64_t order_min = 65536;       // evaluate in cell order from this many points
  int64_t sl_acc_path = 0;         // SL accumulate: 0 auto (vector ALU), 1 MFMA, 2 vector ALU
  int64_t eof_dep_shape = -1;      // cell-ordered EOF accumulate lane shape (index), -1 auto
  int64_t sl_dep_shape = -1;       // cell-ordered SL accumulate lane shape (index), -1 auto

  // ---- optional HIP-event timing of the main compute kernels ---------------
  bool kernel_timing = false;
  std::map<std::string, std::vector<std::pair<hipEvent_t, hipEvent_t>>> timer_events;

  // ---- a second set of work buffers and an internal stream, for the two
  // independent halves of a call that runs them concurrently (disc and
  // halo accumulate, EOF and SL folds); see fork_begin / fork_end
  void* alt_scratch = nullptr;
  size_t alt_scratch_bytes = 0;
  void* alt_order = nullptr;
  size_t alt_order_bytes = 0;
  hipStream_t aux = nullptr;   // created on first use, default priority
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;

  // ---- spatial ordering of evaluation points (bfe_order.hip) -------------
  void* order = nullptr;       // sort keys and packed points (two buffers each) | perm | per-point f64 | tile counts
  size_t order_bytes = 0;

  // ---- batched field evaluation: per-snapshot folded tables ---------------
  float* batch_sl = nullptr;   // [nsnap][SL folded floats]
  float* batch_eof = nullptr;  // [nsnap][EOF folded floats]
  size_t batch_sl_floats = 0, batch_eof_floats = 0;
  int64_t* batch_off = nullptr;  // device copy of the snapshot offsets
  size_t batch_off_cap = 0;
  int batch_nsnap = 0;           // coefficient sets folded by bfe_field_set_batch
  int batch_flags = 0;
};

// Table geometry passed to kernels by value.
struct SLGeom {
  const double* xi;
  const double* p0;
  const double* d0;
  int numr, cmap, lmax, nmax;
  double scale;
  const double* lmc;   // Context::sl_lmc
  double idxi;         // 1 / (xi[1] - xi[0])
  double dfac;         // cmap 1: 0.5 / (scale (xi[1] - xi[0])), d xi/d r over the step
};

struct EOFGeom {
  int cmap, numx, numy, mmax, norder;
  double ascale, hscale, xmin, dx, ymin, dy;
  double ihscale, idx, idy;   // reciprocals of hscale, dx, dy
};

// Folded EOF table of one coefficient set: the float32 blocks of every node,
// then, as float64, the m = 0 cos block of each edge node; the edge nodes are
// the corners of every cell that can extrapolate: the last two grid columns
// X >= numx - 1 (all rows) and, in the other columns, the last two rows
// Y >= numy - 1.
__host__ __device__ inline int64_t eof_fold32_floats(int numx, int numy, int mmax) {
  return (int64_t)(numx + 1) * (numy + 1) * (mmax + 1) * 8;
}
__host__ __device__ inline int64_t eof_edge_count(int numx, int numy) {
  return 2 * (int64_t)(numy + 1) + 2 * (int64_t)(numx - 1);
}
// Edge index of node (X, Y), or -1 for an interior node.
__host__ __device__ inline int eof_edge_index(int X, int Y, int numx, int numy) {
  if (X >= numx - 1) return (X - (numx - 1)) * (numy + 1) + Y;
  if (Y >= numy - 1) return 2 * (numy + 1) + 2 * X + (Y - (numy - 1));
  return -1;
}

inline SLGeom sl_geom(const Context* c) {
  const double dxi = c->h_xi.size() >= 2 ? c->h_xi[1] - c->h_xi[0] : 1.0;
  return SLGeom{c->sl_xi, c->sl_p0, c->sl_d0, c->numr, c->cmap, c->lmax, c->nmax, c->scale,
                c->sl_lmc, 1.0 / dxi, 0.5 / (c->scale * dxi)};
}
inline EOFGeom eof_geom(const Context* c) {
  return EOFGeom{c->ecmap, c->numx, c->numy, c->mmax, c->norder, c->ascale, c->hscale,
                 c->xmin, c->dx, c->ymin, c->dy, 1.0 / c->hscale, 1.0 / c->dx, 1.0 / c->dy};
}

struct FieldTables;
// Permutation perm[j] = i listing the points in order of their SL radial
// cell and cos(theta) bin, so that the lanes of a block share SL nodes and a
// small patch of the EOF grid (perm_out may be NULL).
// Rebuilt by each call; the returned pointers stay valid until the next call
// on this context.
// With off_dev (device int64[nsnap + 1], concatenated snapshots) the key is
// (snapshot, cell), and sorted_out (may be NULL) receives the sorted keys:
// snapshot = key / order_cell_count(T).  points_out (may be NULL) receives
// the points in key order as float4 (x, y, z, input index as int bits).
// off_host (the host copy of off_dev) makes the sort segmented: the
// snapshots keep their ranges and the passes cover only the cells of one.
int order_points(Context* c, const FieldTables& T, const float* x, const float* y,
                 const float* z, int64_t n, hipStream_t s, const int** perm_out,
                 const int64_t* off_dev = nullptr, int nsnap = 1,
                 const uint32_t** sorted_out = nullptr, const float4** points_out = nullptr,
                 const int64_t* off_host = nullptr);
uint32_t order_cell_count(const FieldTables& T, int nsnap);
// order_points over two point sets taken as one (n1 points of x, y, z then
// n2 of x2, y2, z2): points_out in key order with the index over both.
int order_points2(Context* c, const FieldTables& T, const float* x, const float* y,
                  const float* z, int64_t n1, const float* x2, const float* y2, const float* z2,
                  int64_t n2, hipStream_t s, const float4** points_out);
// Sorted copy of the particles packed as float4 (x, y, z, m) in accumulate
// cell order: kind 0 EOF (ix, iy) cells, kind 1 SL radial cells and
// cos(theta) bins (the evaluation order); with off_dev the key is
// (snapshot, cell), so snapshot s keeps its range [off[s], off[s+1]).  The
// sort is stable, so the order is reproducible.
// The sort is segmented when off_host (host copy of off_dev) is given.
int order_particles(Context* c, int kind, const float* x, const float* y, const float* z,
                    const float* m, int64_t n, const int64_t* off_dev, int nsnap, hipStream_t s,
                    const float4** sorted_out, const int64_t* off_host = nullptr);
// Single-snapshot accumulates (bfe_eof_accumulate / bfe_sl_accumulate
// without the argument checks), enqueued on s with the current buffer set.
// Both are for callers that run them side by side (they take the deposit
// kernels sized for sharing the GPU).
int eof_accumulate_one(Context* c, const float* x, const float* y, const float* z,
                       const float* m, int64_t n, double* cos_out, double* sin_out, hipStream_t s);
int sl_accumulate_one(Context* c, const float* x, const float* y, const float* z, const float* m,
                      int64_t n, uint32_t flags, double* coef_out, hipStream_t s);
// float64 per-point scratch in the ordering buffer (n entries), valid after
// order_points for the same n.
double* order_energy_buffer(Context* c, int64_t n);
// Release the ordering buffer (bfe_destroy).
void order_free(Context* c);
// Release the batched-field buffers (bfe_destroy).
void batch_free(Context* c);

// Workgroups are dispatched to the 8 XCDs round-robin; this maps them so that
// each XCD runs one contiguous eighth of the logical blocks (and so of the
// cell-ordered points), keeping each XCD's L2 on its own part of the tables.
// Launch with xcd_grid(nb) blocks; logical blocks >= nb do nothing.
inline unsigned xcd_grid(int64_t nb) { return (unsigned)(8 * ((nb + 7) / 8)); }
__device__ __forceinline__ int64_t xcd_block(int64_t nb) {
  const int64_t per = (nb + 7) / 8;
  return (int64_t)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
}

// With kernel_timing on, record HIP events around a compute kernel launch on
// stream s (read back by bfe_kernel_times).
void timer_begin(Context* c, const char* name, hipStream_t s);
void timer_end(Context* c, const char* name, hipStream_t s);
void timers_free(Context* c);

// Concurrent halves of one call: fork_begin makes the internal stream wait
// for the caller's stream s and returns it; work enqueued on it between
// swap_work calls uses the second buffer set.  fork_end makes s wait for the
// internal stream again (always call it once fork_begin succeeded).
int fork_begin(Context* c, hipStream_t s, hipStream_t* aux);
int fork_end(Context* c, hipStream_t s);
void swap_work(Context* c);

int set_error(Context* c, int code, const char* fmt, ...);
int check_hip(Context* c, hipError_t e, const char* what);
int ensure_scratch(Context* c, size_t bytes);
// Takes the device address of the allocation-free accumulate's static
// scratch array for the current device (bfe_create).
int noalloc_scratch_init(Context* c);
// Lays out eof_dep32 for the cell-ordered EOF accumulate's current lane
// shape (rows of norder rounded up to the shape's chunk length), rebuilding
// it from eof_pot32 when the shape's row length differs (synchronous).
int eof_dep_prepare(Context* c);
void mlp_free(Context* c);

#define BFE_HIP(ctx, call)                                   \
  do {                                                       \
    hipError_t _e = (call);                                  \
    if (_e != hipSuccess) return check_hip((ctx), _e, #call); \
  } while (0)

}  // namespace bfe

struct bfe_ctx {
  bfe::Context c;
};
