// Orbit integration in the frozen disc+halo field (integrate.py:53-190).
//
// One lane per orbit.  A launch advances every orbit by a chunk of steps with
// the state (position, velocity, acceleration) in registers; the field is
// read from the folded tables, which stay resident in L2 / Infinity Cache.
// Chunks end on energy checkpoints; the float64 energy 0.5 v^2 + Phi of each
// orbit is compared with its initial value and per-block totals are kept for
// a deterministic energy log.  Optionally the chunk launches are captured in
// a hipGraph and replayed as one unit.
#include <math.h>

#include <algorithm>
#include <vector>

#include "bfe_field.hpp"

namespace bfe {

int field_tables(Context* c, double rotpos, FieldTables& T);

struct OrbitAux {
  float ax, ay, az, pad;
  double e0, drift;
};

struct LeapArgs {
  double dt, t0, omega;  // omega = 2 pi rotfreq
  const double* dtv;     // per-orbit step sizes [n] (NULL: every orbit uses dt)
  int64_t n;
  int64_t k0, k1;        // steps (k0, k1] are taken; step k is at time t0 + k dt
  int energy_every;      // 0 = no checks
  int check0;            // index of the checkpoint at step k0
  int record_every;      // 0 = no trajectory
  int nblocks;
};

__device__ __forceinline__ double step_of(const LeapArgs& a, int64_t i) {
  return a.dtv ? a.dtv[i] : a.dt;
}

// rotation of the field at step k of an orbit with step size dt
__device__ __forceinline__ void rot_of(const LeapArgs& a, double dt, int64_t k, FieldTables& T) {
  if (a.omega != 0.0) {
    const double ang = a.omega * (a.t0 + (double)k * dt);
    T.crot = cos(ang);
    T.srot = sin(ang);
  } else {
    T.crot = 1.0;
    T.srot = 0.0;
  }
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (tid < w) sh[tid] += sh[tid + w];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// Orbits are processed in cell order: lane j of logical block xcd_block(nb)
// advances orbit perm[j].  Energies at a checkpoint go to ebuf[orbit] and are
// summed per block of 256 consecutive orbits by energy_block_kernel, so the
// energy log does not depend on the order.

// Step 0: initial acceleration, energy and trajectory record.
__global__ __launch_bounds__(256) void leapfrog_init_kernel(
    const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
    const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz,
    const int* __restrict__ perm, FieldTables T, LeapArgs a, OrbitAux* __restrict__ aux,
    double* __restrict__ ebuf, float* __restrict__ traj) {
  const int64_t j = xcd_block(a.nblocks) * 256 + threadIdx.x;
  if (j < a.n) {
    const int64_t i = perm[j];
    double e;
    rot_of(a, step_of(a, i), a.k0, T);
    double cv[8];
    field_eval_point(T, x[i], y[i], z[i], cv);
    const float fx = (float)(cv[0] + cv[1]), fy = (float)(cv[2] + cv[3]), fz = (float)(cv[4] + cv[5]);
    const float pot = (float)(cv[6] + cv[7]);
    const double v2 = (double)vx[i] * vx[i] + (double)vy[i] * vy[i] + (double)vz[i] * vz[i];
    e = 0.5 * v2 + (double)pot;
    if (ebuf) ebuf[i] = e;
    OrbitAux o;
    o.ax = fx; o.ay = fy; o.az = fz; o.pad = 0.f;
    o.e0 = e;
    o.drift = 0.0;
    aux[i] = o;
    if (traj) {
      const size_t n = (size_t)a.n;
      float* t = traj;  // record 0
      t[0 * n + i] = x[i]; t[1 * n + i] = y[i]; t[2 * n + i] = z[i];
      t[3 * n + i] = vx[i]; t[4 * n + i] = vy[i]; t[5 * n + i] = vz[i];
      t[6 * n + i] = pot;
    }
  }
}

// PER_ORBIT: each orbit reads its step size from a.dtv (an instantiation of
// its own, so the common one keeps dt in scalar registers).  Three waves per
// SIMD (at most 168 registers, 340 B per lane spilled outside the field
// sums): 1.11e10 against 9.1e9 orbit-steps/s at two waves, 1.07e10 at four
// (1M orbits, 10k steps, same results, when 364 B were spilled).
template <bool PER_ORBIT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void leapfrog_chunk_kernel(
    float* __restrict__ x, float* __restrict__ y, float* __restrict__ z, float* __restrict__ vx,
    float* __restrict__ vy, float* __restrict__ vz, const int* __restrict__ perm, FieldTables T,
    LeapArgs a, OrbitAux* __restrict__ aux, double* __restrict__ ebuf, float* __restrict__ traj) {
  __shared__ float4 stage[kStageFloat4];
  __shared__ int smin[4], smax[4];
  const int64_t j = xcd_block(a.nblocks) * 256 + threadIdx.x;
  if (j - threadIdx.x >= a.n) return;   // whole block past the end (XCD mapping)
  const bool live = j < a.n;
  const int64_t i = live ? (int64_t)perm[j] : 0;
  float px = 0, py = 0, pz = 0, ux = 0, uy = 0, uz = 0, ax = 0, ay = 0, az = 0;
  double dtd = a.dt;
  if (live) {
    px = x[i]; py = y[i]; pz = z[i];
    ux = vx[i]; uy = vy[i]; uz = vz[i];
    // the energy fields of aux are read after the steps (fewer registers
    // live across the field evaluations)
    const float4 acc = *reinterpret_cast<const float4*>(&aux[i].ax);
    ax = acc.x; ay = acc.y; az = acc.z;
    if (PER_ORBIT) dtd = a.dtv[i];
  }
  const float dt = (float)dtd;
  const float hdt2 = 0.5f * dt * dt;
  const float hdt = 0.5f * dt;
  float pot = 0.f;
  // every thread of the block runs the same steps: per step the block's SL
  // nodes are staged in LDS when they fit (orbits start a chunk in cell order)
  for (int64_t k = a.k0 + 1; k <= a.k1; ++k) {
    SLCell cell{0.0, 0, 0.0, 0.0};
    if (live) {
      // integrate.py:129-143
      px = px + ux * dt + ax * hdt2;
      py = py + uy * dt + ay * hdt2;
      pz = pz + uz * dt + az * hdt2;
      cell = field_sl_cell(T.slg, px, py, pz);
    }
    int imin = 0;
    const bool staged = stage_sl_nodes(T, live, cell.i, stage, smin, smax, imin);
    if (live) {
      rot_of(a, dtd, k, T);
      double cv[8];
      if (staged)
        field_eval_point<true>(T, px, py, pz, cv, &cell, stage, imin);
      else
        field_eval_point(T, px, py, pz, cv);
      const float nx = (float)(cv[0] + cv[1]), ny = (float)(cv[2] + cv[3]), nz = (float)(cv[4] + cv[5]);
      pot = (float)(cv[6] + cv[7]);
      ux = ux + (ax + nx) * hdt;
      uy = uy + (ay + ny) * hdt;
      uz = uz + (az + nz) * hdt;
      ax = nx; ay = ny; az = nz;
      if (traj && a.record_every > 0 && (k % a.record_every) == 0) {
        const size_t n = (size_t)a.n;
        float* t = traj + (size_t)(k / a.record_every) * 7 * n;
        t[0 * n + i] = px; t[1 * n + i] = py; t[2 * n + i] = pz;
        t[3 * n + i] = ux; t[4 * n + i] = uy; t[5 * n + i] = uz;
        t[6 * n + i] = pot;
      }
    }
    __syncthreads();   // the stage is rewritten by the next step
  }
  if (live) {
    x[i] = px; y[i] = py; z[i] = pz;
    vx[i] = ux; vy[i] = uy; vz[i] = uz;
    OrbitAux o = aux[i];
    o.ax = ax; o.ay = ay; o.az = az;
    if (a.energy_every > 0 && a.k1 > a.k0) {
      const double e = 0.5 * ((double)ux * ux + (double)uy * uy + (double)uz * uz) + (double)pot;
      const double rel = fabs(e - o.e0) / fabs(o.e0);
      if (rel > o.drift || !(rel == rel)) o.drift = rel;
      if (ebuf) ebuf[i] = e;
    }
    aux[i] = o;
  }
}

// Block sums of the checkpoint energies over 256 consecutive orbits (input
// order), fixed tree: partE[check * nblocks + b].
__global__ __launch_bounds__(256) void energy_block_kernel(const double* __restrict__ ebuf,
                                                           int64_t n, double* __restrict__ partE) {
  __shared__ double sh[256];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const double s = block_sum(i < n ? ebuf[i] : 0.0, sh);
  if (threadIdx.x == 0) partE[blockIdx.x] = s;
}

__global__ void leapfrog_finish_kernel(const OrbitAux* __restrict__ aux, int64_t n,
                                       double* __restrict__ drift_out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) drift_out[i] = aux[i].drift;
}

__global__ void energy_reduce_kernel(const double* __restrict__ partE, int nchecks, int nblocks,
                                     double* __restrict__ log_out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nchecks) return;
  double s = 0.0, comp = 0.0;
  for (int b = 0; b < nblocks; ++b) {
    const double yv = partE[(size_t)c * nblocks + b] - comp;
    const double tv = s + yv;
    comp = (tv - s) - yv;
    s = tv;
  }
  log_out[c] = s;
}

// ---------------------------------------------------------------------------
// Fourth-order Runge-Kutta (rk4.py:26-135): four field evaluations per step,
// optionally with the pseudo-forces of a frame rotating at omegab about z.

struct RK4Args {
  double dt;
  double omegab;         // 0 = inertial frame (rk4); else rk4_rotation
  int frame_sign;        // +1 physical pseudo-forces, -1 the reference's signs
  int64_t n;
  int64_t k0, k1;        // steps (k0, k1] are taken
  int record_every;      // 0 = no trajectory
  int nblocks;
};

// Field force at (px, py, pz) for every live lane of the block, the block's
// SL nodes staged in LDS when they fit.  Called by every thread of the block.
__device__ __forceinline__ void rk4_force(const FieldTables& T, bool live, float px, float py,
                                          float pz, float4* stage, int* smin, int* smax,
                                          float& fx, float& fy, float& fz) {
  SLCell cell{0.0, 0, 0.0, 0.0};
  if (live) cell = field_sl_cell(T.slg, px, py, pz);
  int imin = 0;
  const bool staged = stage_sl_nodes(T, live, cell.i, stage, smin, smax, imin);
  if (live) {
    double cv[8];
    if (staged)
      field_eval_point<true, 2>(T, px, py, pz, cv, &cell, stage, imin);
    else
      field_eval_point<false, 2>(T, px, py, pz, cv);
    fx = (float)(cv[0] + cv[1]);
    fy = (float)(cv[2] + cv[3]);
    fz = (float)(cv[4] + cv[5]);
  }
  __syncthreads();   // the stage is rewritten by the next evaluation
}

// Pseudo-forces of the rotating frame in the plane (rk4.py:66-90): physical
// (Binney & Tremaine 3.117) Omega^2 R - 2 Omega x v; the reference returns
// the negative of both terms (frame_sign = -1).
__device__ __forceinline__ void rk4_frame(const RK4Args& a, float px, float py, float ux, float uy,
                                          float& fx, float& fy) {
  if (a.omegab == 0.0) return;
  const float w = (float)a.omegab, s = (float)a.frame_sign;
  fx += s * (w * w * px + 2.0f * w * uy);
  fy += s * (w * w * py - 2.0f * w * ux);
}

// Three waves per SIMD (168 registers, about 300 B per lane spilled):
// 1.94e9 against 1.73e9 steps/s at the two waves 255 registers allowed.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void rk4_chunk_kernel(
    float* __restrict__ x, float* __restrict__ y, float* __restrict__ z, float* __restrict__ vx,
    float* __restrict__ vy, float* __restrict__ vz, const int* __restrict__ perm, FieldTables T,
    RK4Args a, float* __restrict__ traj) {
  __shared__ float4 stage[kStageFloat4];
  __shared__ int smin[4], smax[4];
  const int64_t j = xcd_block(a.nblocks) * 256 + threadIdx.x;
  if (j - threadIdx.x >= a.n) return;   // whole block past the end (XCD mapping)
  const bool live = j < a.n;
  const int64_t i = live ? (int64_t)perm[j] : 0;
  float px = 0, py = 0, pz = 0, ux = 0, uy = 0, uz = 0;
  if (live) {
    px = x[i]; py = y[i]; pz = z[i];
    ux = vx[i]; uy = vy[i]; uz = vz[i];
  }
  const float h = (float)a.dt, hh = 0.5f * h, h6 = h / 6.0f;
  for (int64_t k = a.k0 + 1; k <= a.k1; ++k) {
    // rk4.py:48-60: k1 at x, k2 and k3 at the half steps, k4 at the full step
    float f1x = 0, f1y = 0, f1z = 0;
    rk4_force(T, live, px, py, pz, stage, smin, smax, f1x, f1y, f1z);
    rk4_frame(a, px, py, ux, uy, f1x, f1y);
    const float r2x = ux + f1x * hh, r2y = uy + f1y * hh, r2z = uz + f1z * hh;
    const float q2x = px + ux * hh, q2y = py + uy * hh, q2z = pz + uz * hh;
    float f2x = 0, f2y = 0, f2z = 0;
    rk4_force(T, live, q2x, q2y, q2z, stage, smin, smax, f2x, f2y, f2z);
    rk4_frame(a, q2x, q2y, r2x, r2y, f2x, f2y);
    const float r3x = ux + f2x * hh, r3y = uy + f2y * hh, r3z = uz + f2z * hh;
    const float q3x = px + r2x * hh, q3y = py + r2y * hh, q3z = pz + r2z * hh;
    float f3x = 0, f3y = 0, f3z = 0;
    rk4_force(T, live, q3x, q3y, q3z, stage, smin, smax, f3x, f3y, f3z);
    rk4_frame(a, q3x, q3y, r3x, r3y, f3x, f3y);
    const float r4x = ux + f3x * h, r4y = uy + f3y * h, r4z = uz + f3z * h;
    const float q4x = px + r3x * h, q4y = py + r3y * h, q4z = pz + r3z * h;
    float f4x = 0, f4y = 0, f4z = 0;
    rk4_force(T, live, q4x, q4y, q4z, stage, smin, smax, f4x, f4y, f4z);
    rk4_frame(a, q4x, q4y, r4x, r4y, f4x, f4y);
    if (live) {
      const float nux = ux + h6 * (f1x + 2.0f * f2x + 2.0f * f3x + f4x);
      const float nuy = uy + h6 * (f1y + 2.0f * f2y + 2.0f * f3y + f4y);
      const float nuz = uz + h6 * (f1z + 2.0f * f2z + 2.0f * f3z + f4z);
      px = px + h6 * (ux + 2.0f * r2x + 2.0f * r3x + r4x);
      py = py + h6 * (uy + 2.0f * r2y + 2.0f * r3y + r4y);
      pz = pz + h6 * (uz + 2.0f * r2z + 2.0f * r3z + r4z);
      ux = nux; uy = nuy; uz = nuz;
      if (traj && a.record_every > 0 && (k % a.record_every) == 0) {
        const size_t n = (size_t)a.n;
        float* t = traj + (size_t)(k / a.record_every) * 6 * n;
        t[0 * n + i] = px; t[1 * n + i] = py; t[2 * n + i] = pz;
        t[3 * n + i] = ux; t[4 * n + i] = uy; t[5 * n + i] = uz;
      }
    }
  }
  if (live) {
    x[i] = px; y[i] = py; z[i] = pz;
    vx[i] = ux; vy[i] = uy; vz[i] = uz;
  }
}

// Record 0 of an RK4 trajectory: the initial state.
__global__ void rk4_record0_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                   const float* __restrict__ z, const float* __restrict__ vx,
                                   const float* __restrict__ vy, const float* __restrict__ vz,
                                   int64_t n, float* __restrict__ traj) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    traj[0 * n + i] = x[i]; traj[1 * n + i] = y[i]; traj[2 * n + i] = z[i];
    traj[3 * n + i] = vx[i]; traj[4 * n + i] = vy[i]; traj[5 * n + i] = vz[i];
  }
}

}  // namespace bfe

using namespace bfe;

static int leapfrog_run(bfe_ctx* h, float* x, float* y, float* z, float* vx, float* vy,
                        float* vz, int64_t n, double dt, const double* dtv, int64_t nsteps,
                        double t0, double rotfreq, int energy_every, double* drift_out,
                        double* energy_log, int record_every, float* traj, int use_graph,
                        void* stream) {
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (n < 0 || nsteps < 0) return set_error(c, BFE_EINVAL, "leapfrog: n or nsteps < 0");
  if (n > 0 && (!x || !y || !z || !vx || !vy || !vz)) return set_error(c, BFE_EINVAL, "leapfrog: NULL state");
  if (energy_every < 0 || record_every < 0) return set_error(c, BFE_EINVAL, "leapfrog: negative interval");
  if (traj && record_every == 0) return set_error(c, BFE_EINVAL, "leapfrog: traj needs record_every > 0");
  if (energy_log && energy_every == 0) return set_error(c, BFE_EINVAL, "leapfrog: energy_log needs energy_every > 0");
  if (use_graph && stream == nullptr)
    return set_error(c, BFE_EINVAL, "leapfrog: use_graph needs a non-default stream");
  FieldTables T;
  int rc = field_tables(c, 0.0, T);
  if (rc) return rc;
  if (n == 0) return 0;
  BFE_HIP(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream;
  const int nb = (int)((n + 255) / 256);
  const int chunk = energy_every > 0 ? energy_every : (int)std::min<int64_t>(nsteps > 0 ? nsteps : 1, 100);
  const int nchecks = energy_every > 0 ? (int)(nsteps / energy_every) + 1 : 1;
  const size_t aux_bytes = (size_t)n * sizeof(OrbitAux);
  const size_t part_bytes = (size_t)nchecks * nb * sizeof(double);
  rc = ensure_scratch(c, aux_bytes + part_bytes + 256);
  if (rc) return rc;
  OrbitAux* aux = (OrbitAux*)c->scratch;
  double* partE = (double*)((char*)c->scratch + ((aux_bytes + 255) & ~(size_t)255));
  // cell order of the initial positions, rebuilt from the current positions
  // before every chunk
  const int* perm = nullptr;
  rc = order_points(c, T, x, y, z, n, s, &perm);
  if (rc) return rc;
  double* ebuf = energy_log ? order_energy_buffer(c, n) : nullptr;
  const unsigned grid = xcd_grid(nb);
  LeapArgs a;
  a.dt = dt;
  a.dtv = dtv;
  a.t0 = t0;
  a.omega = 2.0 * kPi * rotfreq;
  a.n = n;
  a.k0 = 0;
  a.k1 = 0;
  a.energy_every = energy_every;
  a.check0 = 0;
  a.record_every = record_every;
  a.nblocks = nb;
  hipLaunchKernelGGL(leapfrog_init_kernel, dim3(grid), dim3(256), 0, s, x, y, z, vx, vy, vz, perm,
                     T, a, aux, ebuf, traj);
  if (ebuf)
    hipLaunchKernelGGL(energy_block_kernel, dim3(nb), dim3(256), 0, s, ebuf, n, partE);
  BFE_HIP(c, hipGetLastError());

  std::vector<LeapArgs> chunks;
  for (int64_t k0 = 0, ci = 0; k0 < nsteps; k0 += chunk, ++ci) {
    LeapArgs b = a;
    b.k0 = k0;
    b.k1 = std::min<int64_t>(nsteps, k0 + chunk);
    b.check0 = (int)ci;
    // a partial final chunk is not a checkpoint (nsteps % energy_every != 0)
    if (energy_every > 0 && (b.k1 - b.k0) != energy_every) b.energy_every = 0;
    chunks.push_back(b);
  }
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (use_graph) BFE_HIP(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  for (const LeapArgs& b : chunks) {
    // re-sort by the current cells between chunks (the sort allocates
    // nothing here: the buffer was sized by the first sort, before any capture)
    if (b.k0 > 0) {
      rc = order_points(c, T, x, y, z, n, s, &perm);
      if (rc) {
        // leave the caller's stream out of capture mode
        if (use_graph && hipStreamEndCapture(s, &graph) == hipSuccess && graph)
          (void)hipGraphDestroy(graph);
        return rc;
      }
    }
    const bool check = ebuf && b.energy_every > 0;
    if (dtv)
      hipLaunchKernelGGL(leapfrog_chunk_kernel<true>, dim3(grid), dim3(256), 0, s, x, y, z, vx, vy,
                         vz, perm, T, b, aux, check ? ebuf : nullptr, traj);
    else
      hipLaunchKernelGGL(leapfrog_chunk_kernel<false>, dim3(grid), dim3(256), 0, s, x, y, z, vx, vy,
                         vz, perm, T, b, aux, check ? ebuf : nullptr, traj);
    if (check)
      hipLaunchKernelGGL(energy_block_kernel, dim3(nb), dim3(256), 0, s, ebuf, n,
                         partE + (size_t)(b.check0 + 1) * nb);
  }
  if (use_graph) {
    hipError_t e = hipStreamEndCapture(s, &graph);
    if (e != hipSuccess) return check_hip(c, e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(graph);
      return check_hip(c, e, "hipGraphInstantiate");
    }
    e = hipGraphLaunch(exec, s);
    if (e != hipSuccess) return check_hip(c, e, "hipGraphLaunch");
    // the graph is released after the stream has drained it
    e = hipStreamSynchronize(s);
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return check_hip(c, e, "hipStreamSynchronize (graph)");
  }
  BFE_HIP(c, hipGetLastError());
  if (drift_out)
    hipLaunchKernelGGL(leapfrog_finish_kernel, dim3(nb), dim3(256), 0, s, aux, n, drift_out);
  if (energy_log)
    hipLaunchKernelGGL(energy_reduce_kernel, dim3((nchecks + 63) / 64), dim3(64), 0, s, partE,
                       nchecks, nb, energy_log);
  BFE_HIP(c, hipGetLastError());
  return 0;
}

extern "C" int bfe_leapfrog(bfe_ctx* h, float* x, float* y, float* z, float* vx, float* vy,
                            float* vz, int64_t n, double dt, int64_t nsteps, double t0,
                            double rotfreq, int energy_every, double* drift_out,
                            double* energy_log, int record_every, float* traj, int use_graph,
                            void* stream) {
  return leapfrog_run(h, x, y, z, vx, vy, vz, n, dt, nullptr, nsteps, t0, rotfreq, energy_every,
                      drift_out, energy_log, record_every, traj, use_graph, stream);
}

extern "C" int bfe_leapfrog_dt(bfe_ctx* h, float* x, float* y, float* z, float* vx, float* vy,
                               float* vz, int64_t n, const double* dt, int64_t nsteps, double t0,
                               double rotfreq, int energy_every, double* drift_out,
                               double* energy_log, int record_every, float* traj, int use_graph,
                               void* stream) {
  if (h && n > 0 && !dt) return set_error(&h->c, BFE_EINVAL, "leapfrog_dt: NULL dt");
  return leapfrog_run(h, x, y, z, vx, vy, vz, n, 0.0, dt, nsteps, t0, rotfreq, energy_every,
                      drift_out, energy_log, record_every, traj, use_graph, stream);
}

extern "C" int bfe_rk4(bfe_ctx* h, float* x, float* y, float* z, float* vx, float* vy, float* vz,
                       int64_t n, double dt, int64_t nsteps, double omegab, uint32_t flags,
                       int record_every, float* traj, void* stream) {
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (n < 0 || nsteps < 0) return set_error(c, BFE_EINVAL, "rk4: n or nsteps < 0");
  if (n > 0 && (!x || !y || !z || !vx || !vy || !vz)) return set_error(c, BFE_EINVAL, "rk4: NULL state");
  if (record_every < 0) return set_error(c, BFE_EINVAL, "rk4: negative record_every");
  if (traj && record_every == 0) return set_error(c, BFE_EINVAL, "rk4: traj needs record_every > 0");
  if (flags & ~BFE_COMPAT_RK4_FRAME) return set_error(c, BFE_EINVAL, "rk4: unknown flags");
  FieldTables T;
  int rc = field_tables(c, 0.0, T);
  if (rc) return rc;
  if (n == 0) return 0;
  BFE_HIP(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream;
  const int nb = (int)((n + 255) / 256);
  const unsigned grid = xcd_grid(nb);
  if (traj)
    hipLaunchKernelGGL(rk4_record0_kernel, dim3(nb), dim3(256), 0, s, x, y, z, vx, vy, vz, n, traj);
  RK4Args a;
  a.dt = dt;
  a.omegab = omegab;
  a.frame_sign = (flags & BFE_COMPAT_RK4_FRAME) ? -1 : 1;
  a.n = n;
  a.record_every = record_every;
  a.nblocks = nb;
  // chunks of 25 steps (100 field evaluations), re-sorted by cell in between
  const int64_t chunk = 25;
  for (int64_t k0 = 0; k0 < nsteps; k0 += chunk) {
    const int* perm = nullptr;
    rc = order_points(c, T, x, y, z, n, s, &perm);
    if (rc) return rc;
    a.k0 = k0;
    a.k1 = std::min<int64_t>(nsteps, k0 + chunk);
    hipLaunchKernelGGL(rk4_chunk_kernel, dim3(grid), dim3(256), 0, s, x, y, z, vx, vy, vz, perm, T,
                       a, traj);
  }
  BFE_HIP(c, hipGetLastError());
  return 0;
}
