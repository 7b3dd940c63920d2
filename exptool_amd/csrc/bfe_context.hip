// Context management and table upload.
#include <stdarg.h>
#include <string.h>

#include <cmath>

#include <utility>

#include "bfe_internal.hpp"

namespace bfe {

static thread_local std::string g_create_err;

int set_error(Context* c, int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) c->err = buf; else g_create_err = buf;
  return code;
}

int check_hip(Context* c, hipError_t e, const char* what) {
  return set_error(c, BFE_EHIP, "HIP error %d (%s) in %s", (int)e, hipGetErrorString(e), what);
}

int ensure_scratch(Context* c, size_t bytes) {
  if (bytes <= c->scratch_bytes) return 0;
  if (c->scratch) {
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return check_hip(c, e, "hipDeviceSynchronize (scratch grow)");
    (void)hipFree(c->scratch);
    c->scratch = nullptr;
    c->scratch_bytes = 0;
  }
  size_t want = bytes + bytes / 4;
  hipError_t e = hipMalloc(&c->scratch, want);
  if (e != hipSuccess) return check_hip(c, e, "hipMalloc scratch");
  c->scratch_bytes = want;
  return 0;
}

template <typename T>
static int dev_upload(Context* c, T** dst, const T* src, size_t count) {
  if (*dst) {
    (void)hipFree(*dst);
    *dst = nullptr;
  }
  BFE_HIP(c, hipMalloc((void**)dst, count * sizeof(T)));
  BFE_HIP(c, hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

static void free_ptr(void* p) {
  if (p) (void)hipFree(p);
}

static void free_sl(Context* c) {
  free_ptr(c->sl_xi); free_ptr(c->sl_p0); free_ptr(c->sl_d0);
  free_ptr(c->sl_fold64); free_ptr(c->sl_e32); free_ptr(c->sl_dep32); free_ptr(c->sl_lmc);
  c->sl_dep32 = nullptr;
  c->sl_xi = c->sl_p0 = c->sl_d0 = c->sl_fold64 = nullptr;
  c->sl_e32 = nullptr;
  c->sl_lmc = nullptr;
  for (auto& f : c->slf) { free_ptr(f.d); f = SLFolded(); }
  c->sl_ok = false;
  c->batch_nsnap = 0;   // batch fields were folded against these tables
}

static void free_eof(Context* c) {
  for (auto& p : c->eof_fold32) { free_ptr(p); p = nullptr; }
  free_ptr(c->eof_pot32);
  free_ptr(c->eof_dep32);
  free_ptr(c->eof_mfmaT);
  free_ptr(c->eof_edge64);
  c->eof_edge64 = nullptr;
  c->eof_pot32 = nullptr;
  c->eof_dep32 = nullptr;
  c->eof_mfmaT = nullptr;
  c->eof_dep_rp = 0;
  c->eof_mfma_np = 0;
  c->eof_mfma_nodes = 0;
  for (auto& f : c->eoff) { free_ptr(f.d); f = EOFFolded(); }
  c->eof_ok = false;
  c->batch_nsnap = 0;   // batch fields were folded against these tables
}

void timer_begin(Context* c, const char* name, hipStream_t s) {
  if (!c->kernel_timing) return;
  hipEvent_t a = nullptr, b = nullptr;
  if (hipEventCreate(&a) != hipSuccess) return;
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return;
  }
  (void)hipEventRecord(a, s);
  c->timer_events[name].push_back({a, b});
}

void timer_end(Context* c, const char* name, hipStream_t s) {
  if (!c->kernel_timing) return;
  auto it = c->timer_events.find(name);
  if (it == c->timer_events.end() || it->second.empty()) return;
  (void)hipEventRecord(it->second.back().second, s);
}

void timers_free(Context* c) {
  for (auto& kv : c->timer_events)
    for (auto& ab : kv.second) {
      (void)hipEventDestroy(ab.first);
      (void)hipEventDestroy(ab.second);
    }
  c->timer_events.clear();
}

}  // namespace bfe

using namespace bfe;

extern "C" int bfe_create(int device, bfe_ctx** out) {
  if (!out) return set_error(nullptr, BFE_EINVAL, "bfe_create: out is NULL");
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0)
    return set_error(nullptr, BFE_EHIP, "bfe_create: no HIP device (%s)", hipGetErrorString(e));
  if (device < 0 || device >= ndev)
    return set_error(nullptr, BFE_EINVAL, "bfe_create: device %d out of range [0,%d)", device, ndev);
  e = hipSetDevice(device);
  if (e != hipSuccess)
    return set_error(nullptr, BFE_EHIP, "bfe_create: hipSetDevice: %s", hipGetErrorString(e));
  bfe_ctx* h = new bfe_ctx();
  h->c.device = device;
  // with the context's device current: the allocation-free accumulate's
  // scratch address on this device (a failure is reported by that entry
  // point, which then refuses to run)
  (void)noalloc_scratch_init(&h->c);
  h->c.err.clear();
  *out = h;
  return 0;
}

extern "C" int bfe_destroy(bfe_ctx* h) {
  if (!h) return 0;
  Context* c = &h->c;
  (void)hipSetDevice(c->device);
  (void)hipDeviceSynchronize();
  free_sl(c);
  free_eof(c);
  mlp_free(c);
  order_free(c);
  batch_free(c);
  timers_free(c);
  if (c->scratch) (void)hipFree(c->scratch);
  if (c->alt_scratch) (void)hipFree(c->alt_scratch);
  if (c->mfma_work) (void)hipFree(c->mfma_work);
  if (c->alt_order) (void)hipFree(c->alt_order);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->aux) (void)hipStreamDestroy(c->aux);
  delete h;
  return 0;
}

namespace bfe {

int fork_begin(Context* c, hipStream_t s, hipStream_t* aux) {
  if (!c->aux) {
    BFE_HIP(c, hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking));
    BFE_HIP(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    BFE_HIP(c, hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
  }
  BFE_HIP(c, hipEventRecord(c->ev_fork, s));
  BFE_HIP(c, hipStreamWaitEvent(c->aux, c->ev_fork, 0));
  *aux = c->aux;
  return 0;
}

int fork_end(Context* c, hipStream_t s) {
  BFE_HIP(c, hipEventRecord(c->ev_join, c->aux));
  BFE_HIP(c, hipStreamWaitEvent(s, c->ev_join, 0));
  return 0;
}

void swap_work(Context* c) {
  std::swap(c->scratch, c->alt_scratch);
  std::swap(c->scratch_bytes, c->alt_scratch_bytes);
  std::swap(c->order, c->alt_order);
  std::swap(c->order_bytes, c->alt_order_bytes);
}

}  // namespace bfe

extern "C" const char* bfe_last_error(const bfe_ctx* h) {
  return h ? h->c.err.c_str() : g_create_err.c_str();
}

// Builds the node-major float32 table ef/sqrt(ev) used by accumulate.
extern "C" int bfe_sl_tables_set(bfe_ctx* h, int lmax, int nmax, int numr, int cmap, double scale,
                                 const double* xi, const double* evtable, const double* eftable,
                                 const double* p0, const double* d0) {
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (lmax < 0 || lmax > kMaxL) return set_error(c, BFE_ELIMIT, "sl: lmax=%d outside [0,%d]", lmax, kMaxL);
  if (nmax < 1 || nmax > 64) return set_error(c, BFE_ELIMIT, "sl: nmax=%d outside [1,64]", nmax);
  if (numr < 3) return set_error(c, BFE_EINVAL, "sl: numr=%d < 3", numr);
  if (cmap < 0 || cmap > 2) return set_error(c, BFE_EINVAL, "sl: cmap=%d", cmap);
  if (!xi || !evtable || !eftable || !p0 || !d0) return set_error(c, BFE_EINVAL, "sl: NULL table");
  if (!(xi[1] > xi[0])) return set_error(c, BFE_EINVAL, "sl: xi must increase");
  for (int i = 0; i < (lmax + 1) * nmax; ++i)
    if (!(evtable[i] > 0.0)) return set_error(c, BFE_EINVAL, "sl: evtable[%d] <= 0", i);
  BFE_HIP(c, hipSetDevice(c->device));
  free_sl(c);
  const size_t LN = (size_t)(lmax + 1) * nmax;
  std::vector<float> e32(LN * numr);
  for (size_t ln = 0; ln < LN; ++ln) {
    double inv = 1.0 / sqrt(evtable[ln]);
    for (int i = 0; i < numr; ++i) e32[(size_t)i * LN + ln] = (float)(eftable[ln * numr + i] * inv);
  }
  // the three scaled tables the coefficient fold sums over n, with n fastest
  const size_t plane = LN * numr;
  std::vector<double> fold64(3 * plane);
  for (int l = 0; l <= lmax; ++l)
    for (int n = 0; n < nmax; ++n) {
      const size_t ln = (size_t)l * nmax + n;
      const double sev = sqrt(evtable[ln]);
      for (int i = 0; i < numr; ++i) {
        const double e = eftable[ln * numr + i];
        const size_t o = ((size_t)l * numr + i) * nmax + n;
        fold64[o] = e / sev;
        fold64[plane + o] = e * sev;
        fold64[2 * plane + o] = e * p0[i] / sev;
      }
    }
  // recurrence and normalisation constants (spheresl.py:664-700, :823-863);
  // N_lm^2 = (2l+1)/(4 pi) (l-m)!/(l+m)! from an exact product of integers
  std::vector<double> lmc(kLmcSize, 0.0);
  for (int l = 0; l <= kMaxL; ++l)
    for (int m = 0; m <= l; ++m) {
      long double ratio = 1.0L;
      for (int k = l - m + 1; k <= l + m; ++k) ratio /= (long double)k;
      long double n2 = (2.0L * l + 1.0L) / (4.0L * 3.14159265358979323846264338327950288L) * ratio;
      long double nlm = sqrtl(n2) * (m > 0 ? sqrtl(2.0L) : 1.0L);
      const int lm = lm_index(l, m);
      lmc[kLmcY + lm] = (double)(-4.0L * 3.14159265358979323846264338327950288L * nlm);
      lmc[kLmcL + lm] = (double)l;
      lmc[kLmcS + lm] = (double)(l + m);
      if (l > m) {
        lmc[kLmcA + lm] = (double)(2 * l - 1) / (double)(l - m);
        lmc[kLmcB + lm] = (double)(l + m - 1) / (double)(l - m);
      }
    }
  {
    double d = 1.0;
    for (int m = 0; m <= kMaxL; ++m) {
      if (m > 0) d *= -(2.0 * m - 1.0);
      lmc[kLmcD + m] = d;
    }
  }
  // the recurrence of Q_l^m = -4 pi N_lm P_l^m: ratios of the normalisations
  // (from the same exact integer products) folded into its constants
  for (int m = 0; m <= kMaxL; ++m) {
    lmc[kLmcN0 + m] = lmc[kLmcY + lm_index(m, m)] * lmc[kLmcD + m];
    for (int l = m + 1; l <= kMaxL; ++l) {
      const int lm = lm_index(l, m);
      const double y = lmc[kLmcY + lm];
      lmc[kLmcNA + lm] = lmc[kLmcA + lm] * y / lmc[kLmcY + lm_index(l - 1, m)];
      lmc[kLmcNB + lm] = l - 2 >= m ? lmc[kLmcB + lm] * y / lmc[kLmcY + lm_index(l - 2, m)] : 0.0;
    }
  }
  int rc;
  if ((rc = dev_upload(c, &c->sl_lmc, lmc.data(), lmc.size()))) return rc;
  if ((rc = dev_upload(c, &c->sl_xi, xi, numr))) return rc;
  if ((rc = dev_upload(c, &c->sl_p0, p0, numr))) return rc;
  if ((rc = dev_upload(c, &c->sl_d0, d0, numr))) return rc;
  if ((rc = dev_upload(c, &c->sl_fold64, fold64.data(), fold64.size()))) return rc;
  if ((rc = dev_upload(c, &c->sl_e32, e32.data(), e32.size()))) return rc;
  {
    const int np = dep_np(nmax);
    std::vector<float> d32((size_t)numr * (lmax + 1) * np, 0.f);
    for (int i = 0; i < numr; ++i)
      for (int l = 0; l <= lmax; ++l)
        for (int n = 0; n < nmax; ++n)
          d32[((size_t)i * (lmax + 1) + l) * np + n] = e32[(size_t)i * LN + (size_t)l * nmax + n];
    if ((rc = dev_upload(c, &c->sl_dep32, d32.data(), d32.size()))) return rc;
  }
  c->h_xi.assign(xi, xi + numr);
  c->lmax = lmax; c->nmax = nmax; c->numr = numr; c->cmap = cmap; c->scale = scale;
  c->sl_ok = true;
  return 0;
}

// Reorders (mmax+1, norder, numx+1, numy+1) C and S tables into node-major
// [node][K] rows with K = (2 mmax + 1) norder: cos m=0..mmax, then sin m=1..mmax.
extern "C" int bfe_eof_tables_set(bfe_ctx* h, int mmax, int norder, int numx, int numy, int cmap,
                                  double ascale, double hscale, double xmin, double dx, double ymin,
                                  double dy, const double* potC, const double* rforceC,
                                  const double* zforceC, const double* densC, const double* potS,
                                  const double* rforceS, const double* zforceS, const double* densS) {
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (mmax < 0 || mmax > kMaxM) return set_error(c, BFE_ELIMIT, "eof: mmax=%d outside [0,%d]", mmax, kMaxM);
  if (norder < 1 || norder > 64) return set_error(c, BFE_ELIMIT, "eof: norder=%d outside [1,64]", norder);
  if (numx < 1 || numy < 1) return set_error(c, BFE_EINVAL, "eof: numx/numy < 1");
  if (cmap < 0 || cmap > 2) return set_error(c, BFE_EINVAL, "eof: cmap=%d", cmap);
  if (!potC || !rforceC || !zforceC || !potS || !rforceS || !zforceS)
    return set_error(c, BFE_EINVAL, "eof: NULL table");
  if (!(dx > 0) || !(dy > 0)) return set_error(c, BFE_EINVAL, "eof: dx, dy must be > 0");
  const bool dens = densC && densS;
  const int K = (2 * mmax + 1) * norder;
  if (K > 1024) return set_error(c, BFE_ELIMIT, "eof: (2 mmax+1) norder = %d > 1024", K);
  BFE_HIP(c, hipSetDevice(c->device));
  free_eof(c);
  const size_t nodes = (size_t)(numx + 1) * (numy + 1);
  const double* srcC[4] = {potC, rforceC, zforceC, densC};
  const double* srcS[4] = {potS, rforceS, zforceS, densS};
  // fold inputs are function-major [K][node] (coalesced folding); the
  // float32 accumulate table is node-major [node][K] (coalesced gathers).
  std::vector<double> buf(nodes * K);
  std::vector<float> buf32(nodes * K);
  std::vector<float> pot32;
  // edge nodes (eof_edge_index) keep their m = 0 cos fold inputs in float64
  const int64_t nedge = eof_edge_count(numx, numy);
  std::vector<int64_t> edge_node(nedge);
  for (int X = 0; X <= numx; ++X)
    for (int Y = 0; Y <= numy; ++Y) {
      const int ei = eof_edge_index(X, Y, numx, numy);
      if (ei >= 0) edge_node[ei] = (int64_t)X * (numy + 1) + Y;
    }
  std::vector<double> edge64((size_t)4 * norder * nedge, 0.0);
  for (int q = 0; q < (dens ? 4 : 3); ++q) {
    for (int m = 0; m <= mmax; ++m)
      for (int n = 0; n < norder; ++n) {
        const double* s = srcC[q] + ((size_t)m * norder + n) * nodes;
        const int k = m * norder + n;
        memcpy(&buf[(size_t)k * nodes], s, nodes * sizeof(double));
      }
    for (int m = 1; m <= mmax; ++m)
      for (int n = 0; n < norder; ++n) {
        const double* s = srcS[q] + ((size_t)m * norder + n) * nodes;
        const int k = (mmax + m) * norder + n;
        memcpy(&buf[(size_t)k * nodes], s, nodes * sizeof(double));
      }
    for (size_t j = 0; j < buf.size(); ++j) buf32[j] = (float)buf[j];
    int rc = dev_upload(c, &c->eof_fold32[q], buf32.data(), buf32.size());
    if (rc) return rc;
    for (int n = 0; n < norder; ++n)   // rows k = n: cos m = 0
      for (int64_t e = 0; e < nedge; ++e)
        edge64[((size_t)q * norder + n) * nedge + e] = buf[(size_t)n * nodes + edge_node[e]];
    if (q == 0) {
      pot32.resize(buf.size());
      for (int k = 0; k < K; ++k)
        for (size_t node = 0; node < nodes; ++node)
          pot32[node * K + k] = (float)buf[(size_t)k * nodes + node];
      if ((rc = dev_upload(c, &c->eof_pot32, pot32.data(), pot32.size()))) return rc;
    }
  }
  {
    int rc = dev_upload(c, &c->eof_edge64, edge64.data(), edge64.size());
    if (rc) return rc;
  }
  c->mmax = mmax; c->norder = norder; c->numx = numx; c->numy = numy; c->ecmap = cmap;
  c->ascale = ascale; c->hscale = hscale; c->xmin = xmin; c->dx = dx; c->ymin = ymin; c->dy = dy;
  c->eof_K = K;
  c->eof_dens = dens;
  // padded rows for the cell-ordered accumulate's lane shape
  if (int rc = eof_dep_prepare(c)) return rc;
  c->eof_ok = true;
  return 0;
}

// Launch-shape knobs, set by hand or by the autotuner (exptool_amd/autotune.py).
extern "C" int bfe_set_tuning(bfe_ctx* h, const char* name, int64_t value) {
  if (!h || !name) return BFE_EINVAL;
  Context* c = &h->c;
  if (!strcmp(name, "eof_acc_blocks") || !strcmp(name, "sl_acc_blocks")) {
    if (value < 1 || value > (1 << 20)) return set_error(c, BFE_EINVAL, "%s outside [1, 2^20]", name);
    (name[0] == 'e' ? c->eof_acc_blocks : c->sl_acc_blocks) = value;
  } else if (!strcmp(name, "sl_acc_blocks_beside")) {
    if (value < 1 || value > (1 << 20)) return set_error(c, BFE_EINVAL, "%s outside [1, 2^20]", name);
    c->sl_acc_blocks_beside = value;
  } else if (!strcmp(name, "eof_seg_tiles") || !strcmp(name, "sl_seg_tiles")) {
    if (value < 0) return set_error(c, BFE_EINVAL, "%s < 0", name);
    (name[0] == 'e' ? c->eof_seg_tiles : c->sl_seg_tiles) = value;
  } else if (!strcmp(name, "order_min")) {
    if (value < 0) return set_error(c, BFE_EINVAL, "order_min < 0");
    c->order_min = value;
  } else if (!strcmp(name, "kernel_timing")) {
    c->kernel_timing = value != 0;
  } else if (!strcmp(name, "sl_acc_path")) {
    if (value < 0 || value > 2) return set_error(c, BFE_EINVAL, "sl_acc_path outside [0, 2]");
    c->sl_acc_path = value;
  } else if (!strcmp(name, "eof_dep_shape") || !strcmp(name, "sl_dep_shape")) {
    const bool e = name[0] == 'e';
    const int64_t count = e ? kEofDepShapes : kSlDepShapes;
    if (value < -1 || value >= count)
      return set_error(c, BFE_EINVAL, "%s outside [-1, %lld]", name, (long long)(count - 1));
    (e ? c->eof_dep_shape : c->sl_dep_shape) = value;
    // the EOF table rows are laid out for the shape's chunk length
    if (e && c->eof_ok) {
      (void)hipSetDevice(c->device);
      if (int rc = eof_dep_prepare(c)) return rc;
    }
  } else {
    return set_error(c, BFE_EINVAL, "unknown tuning parameter '%s'", name);
  }
  return 0;
}

extern "C" int bfe_get_tuning(const bfe_ctx* h, const char* name, int64_t* value) {
  if (!h || !name || !value) return BFE_EINVAL;
  const Context* c = &h->c;
  if (!strcmp(name, "eof_acc_blocks")) *value = c->eof_acc_blocks;
  else if (!strcmp(name, "sl_acc_blocks")) *value = c->sl_acc_blocks;
  else if (!strcmp(name, "sl_acc_blocks_beside")) *value = c->sl_acc_blocks_beside;
  else if (!strcmp(name, "eof_seg_tiles")) *value = c->eof_seg_tiles;
  else if (!strcmp(name, "sl_seg_tiles")) *value = c->sl_seg_tiles;
  else if (!strcmp(name, "order_min")) *value = c->order_min;
  else if (!strcmp(name, "kernel_timing")) *value = c->kernel_timing ? 1 : 0;
  else if (!strcmp(name, "sl_acc_path")) *value = c->sl_acc_path;
  else if (!strcmp(name, "eof_dep_shape")) *value = c->eof_dep_shape;
  else if (!strcmp(name, "sl_dep_shape")) *value = c->sl_dep_shape;
  else return BFE_EINVAL;
  return 0;
}

// Total HIP-event time (ms) and count of the timed launches of kernel `name`
// since the last call; waits for them to finish and forgets them.
extern "C" int bfe_kernel_times(bfe_ctx* h, const char* name, double* total_ms, int64_t* count) {
  if (!h || !name || !total_ms || !count) return BFE_EINVAL;
  Context* c = &h->c;
  *total_ms = 0.0;
  *count = 0;
  auto it = c->timer_events.find(name);
  if (it == c->timer_events.end()) return 0;
  for (auto& ab : it->second) {
    BFE_HIP(c, hipEventSynchronize(ab.second));
    float ms = 0.f;
    BFE_HIP(c, hipEventElapsedTime(&ms, ab.first, ab.second));
    *total_ms += ms;
    *count += 1;
    (void)hipEventDestroy(ab.first);
    (void)hipEventDestroy(ab.second);
  }
  it->second.clear();
  return 0;
}
