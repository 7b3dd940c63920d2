/* exptool_bfe.h -- C ABI of the MI355X basis-function-expansion library.
 *
 * The reference (exptool) has no FFI on this path: its interface is plain
 * Python functions over NumPy arrays.  Each entry point below replaces the
 * reference function cited beside it; the Python package exptool_amd keeps the
 * reference signatures and calls these through ctypes.
 *
 * Conventions
 *  - Particle arrays (x, y, z, m, velocities) and outputs are DEVICE pointers
 *    owned by the caller (float32 SoA, contiguous).  Coefficient arrays are
 *    DEVICE float64 pointers.  Table arrays passed to *_tables_set are HOST
 *    float64 pointers; the context copies them to device memory it owns.
 *  - Every call returns 0 on success and a negative BFE_E* code on error; the
 *    message is available from bfe_last_error().  No exception crosses the ABI.
 *  - Work is enqueued on `stream` (a hipStream_t; NULL = default stream).
 *    Calls are asynchronous unless stated otherwise.
 *  - One context per host thread; a context is bound to one device.
 *  - A context's work buffers (sort buffers, partial sums, the staging of
 *    evaluated forces) are shared by all its calls: issue the calls of one
 *    context on one stream, or order them yourself.  Calls on different
 *    streams that overlap on one context race.  Use one context per
 *    concurrent stream instead.
 *  - Results are deterministic: reductions use fixed orders, no float atomics.
 */
#ifndef EXPTOOL_BFE_H
#define EXPTOOL_BFE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bfe_ctx bfe_ctx;

/* error codes */
#define BFE_OK 0
#define BFE_EINVAL (-1)   /* bad argument / shape */
#define BFE_ENOTABLE (-2) /* tables or field not set */
#define BFE_EHIP (-3)     /* HIP runtime error */
#define BFE_ENOMEM (-4)
#define BFE_ELIMIT (-5)   /* size beyond a kernel limit */

/* flags (bit mask) */
#define BFE_COMPAT_SL_COS_L 0x1u   /* spheresl.force_eval cos/sin(l phi), spheresl.py:1222-1225 */
#define BFE_COMPAT_SL_EDGE 0x2u    /* first-cell derivative stencil, spheresl.py:150-153 */
#define BFE_COMPAT_SL_DENS 0x4u    /* all_eval_particles den1 P_1^m + 0.25*pi, spheresl.py:1323,1351 */
#define BFE_COMPAT_EOF_MASK 0x8u   /* eof.force_eval odd-m mask |cos(pi m/2)|, eof.py:833 */
#define BFE_NO_ODD 0x10u           /* skip odd l (SL) / odd m (EOF eval) */
#define BFE_ACC_BF16 0x20u         /* accumulate in the bf16 mode: two-piece bf16 MFMA
                                      products (about 1e-5 relative), see bfe_sl_accumulate
                                      and bfe_eof_accumulate */
#define BFE_COMPAT_ALL 0xFu
#define BFE_COMPAT_RK4_FRAME 0x40u /* rk4_rotation pseudo-force signs, rk4.py:66-90 */
#define BFE_EVAL_DENS 0x80u        /* point evaluators: also the density */

/* ---- context ------------------------------------------------------------ */
/* replaces: the Pool set-up of spheresl.compute_coefficients (spheresl.py:456)
 * and eof.compute_coefficients (eof.py:1215); there is no per-call process. */
int bfe_create(int device, bfe_ctx** out);
int bfe_destroy(bfe_ctx* ctx);
/* last error message of ctx (or of the last failed bfe_create when ctx=NULL) */
const char* bfe_last_error(const bfe_ctx* ctx);

/* ---- tables ------------------------------------------------------------- */
/* replaces: halo_methods.read_cached_table + init_table as consumed by
 * spheresl.compute_coefficients_solitary (spheresl.py:587-588).
 * evtable (lmax+1, nmax), eftable (lmax+1, nmax, numr), xi/p0/d0 (numr). */
int bfe_sl_tables_set(bfe_ctx* ctx, int lmax, int nmax, int numr, int cmap, double scale,
                      const double* xi, const double* evtable, const double* eftable,
                      const double* p0, const double* d0);

/* replaces: eof.parse_eof + eof_params + set_table_params as consumed by
 * eof.accumulate / force_eval (eof.py:1225-1227).  Each table is
 * (mmax+1, norder, numx+1, numy+1) float64; S tables have an unused m=0 slab.
 * densC/densS may be NULL. */
int bfe_eof_tables_set(bfe_ctx* ctx, int mmax, int norder, int numx, int numy, int cmap,
                       double ascale, double hscale, double xmin, double dx, double ymin,
                       double dy, const double* potC, const double* rforceC,
                       const double* zforceC, const double* densC, const double* potS,
                       const double* rforceS, const double* zforceS, const double* densS);

/* ---- accumulate --------------------------------------------------------- */
/* replaces: spheresl.compute_coefficients_solitary (spheresl.py:567-656) and
 * its Pool driver multi_compute_coefficients + np.sum (spheresl.py:411-471).
 * coef_out: device ((lmax+1)^2, nmax) float64.  flags: BFE_NO_ODD, BFE_ACC_BF16.
 * The sums over particles run on the vector ALUs in float32 with float64
 * partial sums (cell-ordered from "order_min" particles on).  For lmax <= 7
 * they can run as bf16 MFMA products with float32 accumulation instead:
 * three-piece split operands with tuning knob sl_acc_path = 1 (float32-class,
 * <= 1e-6 relative), two pieces with BFE_ACC_BF16 (the bf16 mode, <= 1e-4). */
int bfe_sl_accumulate(bfe_ctx* ctx, const float* x, const float* y, const float* z,
                      const float* m, int64_t n, uint32_t flags, double* coef_out,
                      void* stream);

/* replaces: eof.accumulate (eof.py:492-640) and make_coefficients_multi
 * (eof.py:1415-1455).  cos_out/sin_out: device (mmax+1, norder) float64.
 * flags: BFE_ACC_BF16 selects the bf16 mode (mmax <= 15, norder <= 64):
 * the corner sums of each table cell, their node sums, then per row the
 * product with the potential table on the bf16 MFMA units, both operands as
 * bf16 hi + lo pieces (<= 1e-4, about 1e-5, per m block).  Otherwise float32
 * sums with float64 partials (<= 1e-6).  BFE_NO_ODD has no effect (the
 * reference computes the mask but never applies it, eof.py:545-551). */
int bfe_eof_accumulate(bfe_ctx* ctx, const float* x, const float* y, const float* z,
                       const float* m, int64_t n, uint32_t flags, double* cos_out,
                       double* sin_out, void* stream);

/* No counterpart in the reference: bfe_eof_accumulate for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and results as
 * bfe_eof_accumulate in its input-order form (no cell sort): the partial sums
 * live in a static 4 MiB device array of the library, so the call allocates
 * nothing and issues no memset.  Limit: (2 mmax + 1) norder <= 1024; any n.
 * The float32 mode only: BFE_ACC_BF16 returns BFE_EINVAL.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (any context, any stream) race. */
int bfe_eof_accumulate_noalloc(bfe_ctx* ctx, const float* x, const float* y, const float* z,
                               const float* m, int64_t n, uint32_t flags, double* cos_out,
                               double* sin_out, void* stream);

/* ---- batch evaluation --------------------------------------------------- */
/* replaces: spheresl.all_eval_particles (spheresl.py:1240-1362) /
 * eval_particles (spheresl.py:502-528).  coef: device ((lmax+1)^2, nmax).
 * Terms with l outside [L1, L2] are skipped (l >= 1).  Any output may be NULL.
 * flags: BFE_NO_ODD, BFE_COMPAT_SL_DENS, BFE_COMPAT_SL_EDGE, and
 * BFE_COMPAT_SL_COS_L for the force_eval azimuthal factor (spheresl.py:1107-1234). */
int bfe_sl_eval(bfe_ctx* ctx, const float* x, const float* y, const float* z, int64_t n,
                const double* coef, int L1, int L2, uint32_t flags, float* den0, float* den1,
                float* pot0, float* pot1, float* potr, float* pott, float* potp, void* stream);

/* replaces: eof.accumulated_eval_particles (eof.py:989-1144) / compute_forces
 * (eof.py:1263-1317), with the intended semantics (the reference cannot run).
 * Orders m outside [m1, m2] are skipped.  d0/d need density tables. */
int bfe_eof_eval(bfe_ctx* ctx, const float* x, const float* y, const float* z, int64_t n,
                 const double* cos_coef, const double* sin_coef, int m1, int m2,
                 uint32_t flags, float* p0, float* p, float* fr, float* fp, float* fz,
                 float* d0, float* d, void* stream);

/* ---- point evaluators (float64 coordinates in, float64 rows out) ------- */
/* replaces: spheresl.force_eval (spheresl.py:1107-1234) and spheresl.all_eval
 * (spheresl.py:987-1102) at n points given as the reference takes them:
 * r (the radius as passed, no floor), cos theta and phi, DEVICE float64 [n]
 * each.  coef: device ((lmax+1)^2, nmax); l > lmax_use and n >= nmax_use are
 * dropped (< 0: all).  out: device float64 [7][n] = den0, den1 (the l >= 1
 * density), pot0, pot1, potr (dPhi/dr), pott (dPhi/dcos theta), potp
 * (dPhi/dphi); densities times 0.25/pi (all_eval, :1092), written only with
 * BFE_EVAL_DENS (zero rows otherwise).  flags: BFE_NO_ODD, BFE_EVAL_DENS,
 * BFE_COMPAT_SL_EDGE (the stencil get_halo_dens_pot_force uses, :150-153),
 * BFE_COMPAT_SL_COS_L (force_eval's cos/sin(l phi), :1222-1225). */
int bfe_sl_eval_sph(bfe_ctx* ctx, const double* r, const double* costh, const double* phi,
                    int64_t n, const double* coef, int lmax_use, int nmax_use, uint32_t flags,
                    double* out, void* stream);
/* replaces: eof.force_eval (eof.py:756-872) and eof.accumulated_eval
 * (eof.py:874-929) at n points given as the reference takes them: planar
 * radius R (as passed), z and phi, DEVICE float64 [n] each.  Orders
 * m > mmax_use and n >= nmax_use are dropped (< 0: all).  out: device float64
 * [9][n] = p0, p (m >= 1), fr, fp (= -dPhi/dphi), fz, d0, d (m >= 1), fr0,
 * fz0 (the m = 0 parts of fr and fz).  d0/d need BFE_EVAL_DENS and density
 * tables (zero rows otherwise).  flags: BFE_NO_ODD (odd m dropped),
 * BFE_COMPAT_EOF_MASK (with BFE_NO_ODD: force_eval's |cos(pi m/2)| weights,
 * eof.py:833), BFE_EVAL_DENS. */
int bfe_eof_eval_cyl(bfe_ctx* ctx, const double* R, const double* z, const double* phi,
                     int64_t n, const double* cos_coef, const double* sin_coef, int mmax_use,
                     int nmax_use, uint32_t flags, double* out, void* stream);

/* ---- frozen disc+halo field -------------------------------------------- */
/* replaces: Fields.prep_tables + set_field_parameters (potential.py:257-294,
 * :365-378).  Folds the coefficients into the tables (exact linear rewrite).
 * disk_m/disk_n/halo_l/halo_n < 0 mean "all".  flags: BFE_NO_ODD and the
 * BFE_COMPAT_* bits of force_eval. */
int bfe_field_set(bfe_ctx* ctx, const double* eof_cos, const double* eof_sin,
                  const double* sl_coef, double halofac, int disk_m, int disk_n, int halo_l,
                  int halo_n, uint32_t flags, void* stream);

/* replaces: Fields.return_forces_cart (potential.py:445-497) over many points.
 * fx/fy/fz/pot receive the disc+halo sums; comp (may be NULL) receives the 8
 * reference outputs as [8][n]: fx_d, fx_h, fy_d, fy_h, fz_d, fz_h, pot_d, pot_h. */
int bfe_field_eval_cart(bfe_ctx* ctx, const float* x, const float* y, const float* z,
                        int64_t n, double rotpos, float* fx, float* fy, float* fz,
                        float* pot, float* comp, void* stream);

/* ---- orbit integration -------------------------------------------------- */
/* replaces: utils/integrate.leapfrog_integrate (integrate.py:53-190) for many
 * orbits at once.  State (x..vz) is advanced in place by nsteps steps of
 * x += v dt + a dt^2/2; a' = F(x', 2 pi rotfreq t); v += (a + a') dt/2,
 * starting at time t0.  Every `energy_every` steps (0 = never) and at the
 * start, the float64 energy 0.5 v^2 + Phi of every orbit is compared with its
 * initial value; drift_out (device float64 [n], may be NULL) receives
 * max |E - E0| / |E0| per orbit and energy_log (device float64
 * [nsteps/energy_every + 1], may be NULL) the total energy at each check.
 * traj (device float32 [nrec][7][n], may be NULL) records t-ordered
 * (x, y, z, vx, vy, vz, Phi) every record_every steps (step 0 included).
 * use_graph: capture the chunk launches in a hipGraph and replay it. */
int bfe_leapfrog(bfe_ctx* ctx, float* x, float* y, float* z, float* vx, float* vy, float* vz,
                 int64_t n, double dt, int64_t nsteps, double t0, double rotfreq,
                 int energy_every, double* drift_out, double* energy_log, int record_every,
                 float* traj, int use_graph, void* stream);
/* replaces: the per-orbit loop of utils/integrate.integrate_grid (integrate.py:
 * 848-889, and its launch-from-y and 3D forms :760-846), where every orbit of
 * the grid runs leapfrog_integrate with its own dtime (:871).  Same as
 * bfe_leapfrog with dt (device float64 [n]) giving each orbit its step size:
 * orbit i is at time t0 + k dt[i] after k steps and sees the field rotated by
 * 2 pi rotfreq (t0 + k dt[i]). */
int bfe_leapfrog_dt(bfe_ctx* ctx, float* x, float* y, float* z, float* vx, float* vy, float* vz,
                    int64_t n, const double* dt, int64_t nsteps, double t0, double rotfreq,
                    int energy_every, double* drift_out, double* energy_log, int record_every,
                    float* traj, int use_graph, void* stream);

/* replaces: utils/rk4.rk4 and rk4_rotation (rk4.py:26-63, 94-135) for many
 * orbits in the frozen disc+halo field (not rotated: the field is fixed in
 * the frame the orbits are integrated in).  State (x..vz) is advanced in
 * place by nsteps classical RK4 steps of size dt: four field evaluations per
 * step, at x, at the two half steps and at the full step.  omegab != 0 adds
 * the in-plane pseudo-forces of a frame rotating at omegab about z:
 * Omega^2 R - 2 Omega x v (Binney & Tremaine 3.117); with
 * BFE_COMPAT_RK4_FRAME in flags the reference's add_rotation_acceleration
 * signs, -Omega^2 R + 2 Omega x v.  traj (device float32 [nrec][6][n], may
 * be NULL) records (x, y, z, vx, vy, vz) every record_every steps, step 0
 * included. */
int bfe_rk4(bfe_ctx* ctx, float* x, float* y, float* z, float* vx, float* vy, float* vz,
            int64_t n, double dt, int64_t nsteps, double omegab, uint32_t flags,
            int record_every, float* traj, void* stream);

/* ---- batched snapshots ---------------------------------------------------- */
/* Many snapshots in one launch: the particles of nsnap snapshots are
 * concatenated, snapshot s holding [offsets[s], offsets[s+1]) (offsets: HOST
 * int64[nsnap + 1], offsets[0] = 0).  Replaces a loop of the single-snapshot
 * calls over snapshots (e.g. eof.compute_coefficients / spheresl
 * .compute_coefficients per dump, eof.py:1158, spheresl.py:439); each
 * snapshot's coefficients are reduced on their own in a fixed order, equal to
 * what the single-snapshot call returns for the same particles up to the
 * float32 tile sums.  Outputs: [nsnap][...] in the single-call layouts.
 * flags as the single-snapshot calls; in the EOF bf16 mode (BFE_ACC_BF16)
 * the product runs over all snapshots at once (16 per MFMA tile). */
int bfe_eof_accumulate_batch(bfe_ctx* ctx, const float* x, const float* y, const float* z,
                             const float* m, const int64_t* offsets, int nsnap, uint32_t flags,
                             double* cos_out, double* sin_out, void* stream);
int bfe_sl_accumulate_batch(bfe_ctx* ctx, const float* x, const float* y, const float* z,
                            const float* m, const int64_t* offsets, int nsnap, uint32_t flags,
                            double* coef_out, void* stream);
/* Fields.total_coefficients (potential.py:98-252) for one snapshot: disc
 * (x, y, z, m)[nd] -> EOF cos/sin [mmax+1][norder] as bfe_eof_accumulate,
 * halo [nh] -> SL coef [(lmax+1)^2][nmax] as bfe_sl_accumulate (sl_flags:
 * BFE_NO_ODD, BFE_ACC_BF16).  The two run concurrently: the halo half on the
 * context's internal stream, joined back into `stream` before return. */
int bfe_accumulate_disc_halo(bfe_ctx* ctx, const float* dx, const float* dy, const float* dz,
                             const float* dm, int64_t nd, const float* hx, const float* hy,
                             const float* hz, const float* hm, int64_t nh, uint32_t sl_flags,
                             double* cos_out, double* sin_out, double* coef_out, void* stream);

/* replaces: Fields.total_coefficients followed by Fields.return_forces_cart
 * at the same particles (potential.py:98-252, :445-497) -- the force on
 * every particle of the snapshot from the expansion of that snapshot.  The
 * coefficients are those of bfe_accumulate_disc_halo (acc_flags: BFE_NO_ODD,
 * BFE_ACC_BF16), the field that of bfe_field_set (halofac, truncations,
 * field_flags: BFE_NO_ODD, BFE_COMPAT_*), and fx, fy, fz, pot, comp (device
 * float32, [nd + nh] and [8][nd + nh], disc particles first, each may be
 * NULL) those of bfe_field_eval_cart at rotpos over the concatenated points
 * (no concatenated copy is needed).  Each half folds its coefficients as
 * soon as its accumulate ends, so the fold of one overlaps the accumulate
 * of the other.  The frozen field stays set afterwards, as after
 * bfe_field_set. */
int bfe_forces_disc_halo(bfe_ctx* ctx, const float* dx, const float* dy, const float* dz,
                         const float* dm, int64_t nd, const float* hx, const float* hy,
                         const float* hz, const float* hm, int64_t nh, uint32_t acc_flags,
                         double* cos_out, double* sin_out, double* coef_out, double halofac,
                         int disk_m, int disk_n, int halo_l, int halo_n, uint32_t field_flags,
                         double rotpos, float* fx, float* fy, float* fz, float* pot, float* comp,
                         void* stream);
/* Fields.prep_tables + set_field_parameters (potential.py:257-294, :365-378)
 * for nsnap frozen fields at once: eof_cos/eof_sin [nsnap][mmax+1][norder],
 * sl_coef [nsnap][(lmax+1)^2][nmax]; truncation and flags as bfe_field_set.
 * Independent of the field set by bfe_field_set. */
int bfe_field_set_batch(bfe_ctx* ctx, const double* eof_cos, const double* eof_sin,
                        const double* sl_coef, int nsnap, double halofac, int disk_m, int disk_n,
                        int halo_l, int halo_n, uint32_t flags, void* stream);
/* Fields.return_forces_cart (potential.py:445-497) of the batch fields: the
 * points of snapshot s, [offsets[s], offsets[s+1]) (HOST offsets, nsnap as
 * given to bfe_field_set_batch), are evaluated in field s.  Outputs as
 * bfe_field_eval_cart (sums; no components).  Blocks the host until earlier
 * work on `stream` is done (the offsets are copied to the device). */
int bfe_field_eval_batch(bfe_ctx* ctx, const float* x, const float* y, const float* z,
                         const int64_t* offsets, int nsnap, double rotpos, float* fx, float* fy,
                         float* fz, float* pot, void* stream);

/* ---- launch-shape tuning (no counterpart in the reference) ---------------- */
/* Knobs: "sl_acc_path" (0 auto, 1 MFMA, 2 vector ALU), "eof_acc_blocks",
 * "sl_acc_blocks" (accumulate blocks / batch segment budget per kernel,
 * default 2048), "eof_seg_tiles", "sl_seg_tiles" (tiles per batch segment,
 * 0 = derived from the budget), "eof_dep_shape" (0..7), "sl_dep_shape" (0..6)
 * (lane shape of the cell-ordered accumulate kernels, -1 = automatic; a
 * shape that does not fit the tables falls back to the automatic one),
 * "order_min" (points from which evaluation runs in table-cell order, 65536).
 * Results do not depend on "order_min"; accumulate sums depend on the segment
 * plan and the lane shape only through the float32 partial sums. */
int bfe_set_tuning(bfe_ctx* ctx, const char* name, int64_t value);
int bfe_get_tuning(const bfe_ctx* ctx, const char* name, int64_t* value);
/* With tuning knob "kernel_timing" = 1, the main compute kernel of
 * bfe_field_eval_cart ("field_eval"), bfe_eof_accumulate ("eof_accumulate")
 * and bfe_sl_accumulate ("sl_accumulate") is bracketed by HIP events on the
 * call's stream (cell sort and reductions excluded).  Returns the total time
 * (ms) and number of such launches since the last call for `name`; waits for
 * them. */
int bfe_kernel_times(bfe_ctx* ctx, const char* name, double* total_ms, int64_t* count);

/* ---- MLP neural field (no counterpart in the reference) ----------------- */
/* A 3 -> 128 -> 128 -> 128 -> 128 -> 4 SiLU network mapping (x, y, z) to
 * (Phi, F_R, F_phi, F_z) of a frozen field, trained only on evaluations of the
 * plain BFE.  Inputs are features (x/sxy, y/sxy, asinh(z/hz)/wz); outputs are
 * physical = os[k] * y_k + ob[k].  cfg = {sxy, hz, wz, os[4], ob[4]}. */
#define BFE_MLP_NPARAMS 50564
/* Parameter vector (float32): W0[128][3], b0[128], W1[128][128], b1[128],
 * W2, b2, W3, b3, W4[4][128], b4[4].  Resets the Adam moments. */
int bfe_mlp_set_params(bfe_ctx* ctx, const float* params_host, int64_t count, const double* cfg);
int bfe_mlp_get_params(bfe_ctx* ctx, float* params_host, int64_t count);
/* Analytic base the network's outputs are added to (the network learns the
 * field minus the base; default: no base).  base_host: nmn, then (M, a, b)
 * for each of nmn <= 4 Miyamoto-Nagai discs Phi = -M / sqrt(R^2 + (a +
 * sqrt(z^2 + b^2))^2), then Mh, ah of a Hernquist sphere Phi = -Mh / (r +
 * ah) (Mh = 0: none) and a constant phi0; count = 4 + 3 nmn.  Evaluated in
 * float64 by bfe_mlp_forward and bfe_mlp_leapfrog: physical output k =
 * os[k] y_k + ob[k] + base_k, with base (Phi, F_R, 0, F_z). */
int bfe_mlp_set_base(bfe_ctx* ctx, const double* base_host, int count);
/* conservative = 1: the forces of bfe_mlp_forward and bfe_mlp_leapfrog are
 * minus the gradient of the potential output (plus the base), taken by
 * forward-mode tangents through the network (three more bf16 MFMA passes per
 * hidden layer); 0: the four outputs are read directly.  loss_weights (HOST
 * float64 [4], may be NULL = all 1) weight each output's squared error in
 * bfe_mlp_gradient / bfe_mlp_train_step. */
int bfe_mlp_set_mode(bfe_ctx* ctx, int conservative, const double* loss_weights);
/* out: device float32 [4][n] = Phi, F_R, F_phi, F_z (F_phi = -dPhi/dphi / R). */
int bfe_mlp_forward(bfe_ctx* ctx, const float* x, const float* y, const float* z, int64_t n,
                    float* out, void* stream);
/* Gradient of L = sum_p w_p sum_k (y_k - t_k)^2 / n (t: device [4][n] normalised
 * targets, w: device [n] or NULL) into grad_out (device float32 [BFE_MLP_NPARAMS]);
 * loss_out (device float64 [1], may be NULL) receives n L. */
int bfe_mlp_gradient(bfe_ctx* ctx, const float* x, const float* y, const float* z,
                     const float* target, const float* weight, int64_t n, float* grad_out,
                     double* loss_out, void* stream);
/* One Adam step on the batch (gradient as bfe_mlp_gradient). */
int bfe_mlp_train_step(bfe_ctx* ctx, const float* x, const float* y, const float* z,
                       const float* target, const float* weight, int64_t n, float lr,
                       float beta1, float beta2, float eps, double* loss_out, void* stream);

/* Orbit integration with the surrogate field (same update rule and state
 * conventions as bfe_leapfrog; needs bfe_field_set and bfe_mlp_set_params).
 * Points with R >= rmax or |z| >= zmax use the plain BFE field.  Every step,
 * one wave in check_period (rotating) also evaluates the BFE and records the
 * force error |F_mlp - F_bfe| / |F_bfe|.  Energies use the BFE potential.
 * Outputs (device, each may be NULL): drift_out float64 [n] max relative
 * energy drift, maxerr_out float32 [n] largest checked force error,
 * counts_out int32 [n][2] (BFE fallback evaluations, checked evaluations),
 * energy_log float64 [nsteps/energy_every + 1] total energy. */
int bfe_mlp_leapfrog(bfe_ctx* ctx, float* x, float* y, float* z, float* vx, float* vy, float* vz,
                     int64_t n, double dt, int64_t nsteps, double t0, double rotfreq,
                     int energy_every, float rmax, float zmax, int check_period,
                     double* drift_out, float* maxerr_out, int* counts_out, double* energy_log,
                     void* stream);

#ifdef __cplusplus
}
#endif

#endif /* EXPTOOL_BFE_H */
