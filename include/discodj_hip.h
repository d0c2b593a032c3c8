/*
 * discodj_hip — C ABI of the MI355X (gfx950) particle-mesh N-body step.
 *
 * Drop-in boundary for DiscoDJNative's PM path (reference: native/DiscoDJNative/src/nbody/).
 * The reference selects a solver object and calls it through one generic function
 * (run_nbody.jl:140-141):
 *       solver = lean ? PMSolverLean(cfg, psi_ini) : PMSolver(cfg, psi_ini)
 *       acc(x) = pm_acceleration!(solver, x)
 * and steps  Psi += dd1 Mom;  Mom = alpha Mom + beta acc(Psi);  Psi += dd2 Mom  (run_nbody.jl:155-158)
 * with FP64 coefficients from stepper_coefficients (run_nbody.jl:48-101) kept on the host.
 * A third solver type on the Julia side (PMSolverHIP, see the integration notes) ccalls these entry
 * points; the particle state stays resident on the device between steps.
 *
 * Conventions
 *   - Every function returns dj_status: 0 on success, a negative DJ_ERR_* code otherwise; no C++
 *     exception crosses the ABI.  dj_last_error() describes the last failure of a context.
 *   - Particle arrays are the reference's column-major (n,n,n,3) arrays: element (i,j,k,c) at
 *     i + n*j + n^2*k + n^3*c, in the context's precision (float for precision 32, double for 64).
 *   - Meshes are column-major (N,N,N): cell (ix,iy,iz) at ix + N*iy + N^2*iz.
 *   - Host pointers are borrowed for the duration of the call; the call copies and blocks.
 *   - The library allocates and owns all device memory.  A context owns one HIP stream of default
 *     priority and is single-threaded; separate contexts may be used from separate threads.
 */
#ifndef DISCODJ_HIP_H
#define DISCODJ_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t dj_status;

#define DJ_OK 0
#define DJ_ERR_INVALID_CONFIG (-1)   /* option outside what the reference accepts (pm.jl error paths) */
#define DJ_ERR_INVALID_ARGUMENT (-2) /* null pointer, bad size */
#define DJ_ERR_UNSUPPORTED (-3)      /* valid reference option not built into this library */
#define DJ_ERR_OUT_OF_MEMORY (-4)    /* device allocation failed */
#define DJ_ERR_DEVICE (-5)           /* HIP runtime error */
#define DJ_ERR_NONFINITE (-6)        /* NaN/Inf found in the particle state (dj_check_finite) */

#define DJ_DEPOSIT_FLOAT 0   /* density mesh in the working precision.  FP32: each tile sums the
                              * float weight products w, as the integers nearest to w 2^31, in 64
                              * bits and rounds once to float (independent of the order of the
                              * adds); FP64: double atomics */
#define DJ_DEPOSIT_FIXED64 1 /* int64 fixed-point weights: bit-reproducible density */

/* PMConfig (pm.jl:35-46) plus the device-side choices.  Field meanings and accepted values are the
 * reference's: worder 2 (CIC) / 3 (TSC) / 4 (PCS); antialias -1, 0, 1, 2, 3; grad_order 0 (ik) or
 * 2/4/6 (finite differences); lap_order 0 (-1/k^2) or 2/4/6; resampling 0 = fourier, 1 = linear.
 * Sheet resampling is built for n_resample 1..4 (DJ_ERR_UNSUPPORTED above); the Fourier form needs
 * n_part in {16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024}. */
typedef struct dj_pm_config {
    int32_t res_pm;
    int32_t n_part;
    double boxsize;
    int32_t worder;
    int32_t deconvolve;
    int32_t antialias;
    int32_t grad_order;
    int32_t lap_order;
    int32_t n_resample;
    int32_t resampling;
    int32_t precision;       /* 32 or 64 */
    int32_t deposit;         /* DJ_DEPOSIT_FLOAT or DJ_DEPOSIT_FIXED64 */
    int32_t fixed_frac_bits; /* fractional bits of the fixed-point deposit (0 -> 32) */
    int32_t lean;            /* 1: PMSolverLean semantics and option checks (pm.jl:569-571) */
    int32_t device;          /* HIP device ordinal */
} dj_pm_config;

typedef struct dj_ctx dj_ctx;

/* Called by dj_run after step k (k = 0 before the first step) when want[k] is set; the callback may
 * call dj_get_state.  a_k is not known to the library: the host keeps all_a. */
typedef void (*dj_step_cb)(int32_t k, void *user);

/* Per-kernel timing of a context (dj_set_profiling): HIP events on the context's stream. */
typedef struct dj_kernel_stat {
    char name[32];
    int32_t calls;
    double ms_total;        /* summed event time */
    double bytes_per_call;  /* algorithmic HBM bytes of one launch */
} dj_kernel_stat;

/* PMSolver(cfg, like) / PMSolverLean(cfg, like) (pm.jl:385-414, 569-585): validates the options,
 * builds the k-space multipliers in FP64 on the host, allocates the device buffers. */
dj_status dj_create(const dj_pm_config *cfg, dj_ctx **out);
dj_status dj_destroy(dj_ctx *ctx);
/* message of the last failing call on ctx (ctx NULL: of the last failing dj_create) */
const char *dj_last_error(const dj_ctx *ctx);

/* particle state (Psi, Pi): host arrays of 3 n^3 elements; pi NULL sets zero momentum */
dj_status dj_set_state(dj_ctx *ctx, const void *psi, const void *pi);
dj_status dj_get_state(dj_ctx *ctx, void *psi, void *pi);
/* the same from / to device memory (same layout, same precision) */
dj_status dj_set_state_device(dj_ctx *ctx, const void *d_psi, const void *d_pi);
dj_status dj_get_state_device(dj_ctx *ctx, void *d_psi, void *d_pi);

/* pm_acceleration!(S, psi) (pm.jl:507-521, 623-641): -grad phi at X = q + psi, nabla^2 phi = delta,
 * averaged over the interlacing shifts.  psi NULL uses the context's Psi; acc_out is a host array. */
dj_status dj_acceleration(dj_ctx *ctx, const void *psi, void *acc_out);

/* one DKD step of run_nbody (run_nbody.jl:156-158) on the resident state, FP64 coefficients */
dj_status dj_step(dj_ctx *ctx, double dd1, double alpha, double beta, double dd2);

/* the step loop of run_nbody (run_nbody.jl:151-163): n_steps DKD steps with per-step coefficients
 * (arrays of n_steps).  want (n_steps + 1 entries, or NULL) selects the steps after which cb runs
 * (callback_when); between such steps the next drift is fused into the kick. */
dj_status dj_run(dj_ctx *ctx, int32_t n_steps, const double *dd1, const double *alpha, const double *beta,
                 const double *dd2, const uint8_t *want, dj_step_cb cb, void *user);

/* pm_scatter! (pm.jl:97-106) of X = q + psi + shift (shift NULL = 0): the unit-mass density
 * mesh, unnormalised.  mesh_out is a host array of N^3 elements of the context's precision, or of
 * int64 fixed-point values (fixed_frac_bits fractional bits) in DJ_DEPOSIT_FIXED64 mode. */
dj_status dj_deposit(dj_ctx *ctx, const void *psi, const double *shift, void *mesh_out);
/* dj_deposit of the context's Psi (no shift) into device memory d_mesh_out (N^3 values of the context's
 * precision, or int64 in DJ_DEPOSIT_FIXED64 mode), for callers that keep the mesh on the device (a
 * ROCArray behind the rrule of the density, a PyTorch tensor); returns after the copy */
dj_status dj_deposit_device(dj_ctx *ctx, void *d_mesh_out);

/* Initial conditions into the context's state, computed on the device in FP64 and rounded to the
 * context's precision: the IC pipeline of the reference (ngenic_ic, ngenic_gsl.jl:187-191; ic_operator /
 * white_noise_to_fphi, grf.jl:118-144; compute_core, nlpt_core.jl:311-362 and nlpt_fast.jl:208-249;
 * nbody_ics_lpt, run_nbody.jl:180-208).  White noise: `white` (host, n^3 real values, column-major) or,
 * when NULL, the N-GenIC field of `seed` (GSL ranlxd1, generated on the device).  phi(k) = rfft(w) sqrt(P(k) (n/L)^3) (-1/k^2)
 * with P interpolated linearly in the table (k_tab, pk_tab; ntab entries, k ascending).  order 1 or 2
 * (de-aliased on the 3n/2 grid).  Psi = sum_m D^m psi_m, Pi = sum_m m D^(m-1) psi_m (EdS growth). */
dj_status dj_ics_lpt(dj_ctx *ctx, uint32_t seed, const double *white, const double *k_tab, const double *pk_tab,
                     int32_t ntab, double D, int32_t order);
/* dj_ics_lpt with the white noise in device memory (d_white: n^3 FP64 values, or NULL for the N-GenIC field
 * of seed).  The IC pipeline keeps its device buffers in the context between calls (about 20 n^3 doubles
 * for order 2: 5.3 GB at n = 256, 42 GB at n = 512), until dj_destroy.  For order 2 it also keeps the nine
 * second-order fields A_ik on the 3n/2 grid and a copy of the white noise (31.4 n^3 doubles more: 4.2 GB
 * at n = 256, 34 GB at n = 512) for dj_vjp_ics_lpt; without room for them it runs as before. */
dj_status dj_ics_lpt_device(dj_ctx *ctx, uint32_t seed, const double *d_white, const double *k_tab,
                            const double *pk_tab, int32_t ntab, double D, int32_t order);

/* Power spectrum of the CIC (W = 2; TSC 3, PCS 4) density of X = q + psi on the context's mesh
 * (psi NULL: the context's Psi): delta = mesh N^3/n^3 - 1, P(k) = V <|delta_k|^2 / |W(k)|^2> / N^6
 * averaged with the half-spectrum mode weights over linear bins [kf (b + 1/2), kf (b + 3/2)),
 * b < nbins, kf = 2 pi / L (evaluate_power_spectrum, power_spectrum.jl:25-79, with the window
 * deconvolution of the study estimator).  W = 0 deposits CIC without deconvolution.  Outputs k_out
 * (bin centres kf (b + 1)), pk_out and the mode counts (nmodes_out may be NULL); nbins <= 2048. */
dj_status dj_power_spectrum(dj_ctx *ctx, const void *psi, int32_t W, int32_t nbins, double *k_out, double *pk_out,
                            int64_t *nmodes_out);

/* Binned power spectra on the context's mesh (res_pm^3, boxsize L).  A mode k != 0 of the half spectrum
 * (the reference halves z: kz >= 0) is in bin b when edges[b] <= |k| < edges[b + 1], |k| formed as the
 * reference does in FP64; at most DJ_PK_MAX_BINS bins.  Host arrays, blocking calls. */
#define DJ_PK_MAX_BINS 256

/* evaluate_power_spectrum(field, boxsize; bins, logarithmic) (analysis/power_spectrum.jl:25-79) of a real
 * (N,N,N) host field of the context's precision, N = res_pm: P(k) = V sum w |f_k|^2 / sum w / N^6 with
 * w = 1/2 on the kz = 0 and kz = N/2 planes and 1 elsewhere, nbins bins from kf = 2 pi / L to
 * k_Nyq sqrt(3) (k_Nyq = pi N / L), equal in k or, logarithmic != 0, in log10 k; k_out the mean |k| of a
 * bin's modes (the geometric centre of an empty bin), nmodes_out (may be NULL) their number */
dj_status dj_field_power_spectrum(dj_ctx *ctx, const void *field, int32_t nbins, int32_t logarithmic, double *k_out,
                                  double *pk_out, int64_t *nmodes_out);
/* evaluate_cross_power_spectrum(field1, field2, boxsize; bins, logarithmic) (power_spectrum.jl:86-123):
 * P12(k) = V sum Re(conj(f1_k) f2_k) / n_k / N^6, the unweighted mean over a bin's n_k modes; bins and
 * k_out as dj_field_power_spectrum */
dj_status dj_field_cross_power_spectrum(dj_ctx *ctx, const void *field1, const void *field2, int32_t nbins,
                                        int32_t logarithmic, double *k_out, double *pk_out);
/* pk_full of the study estimator (studies/eft_shell_crossing/src/measure.jl:403-447) of X = q + psi (psi
 * NULL: the context's Psi) over the given edges (nbins + 1, ascending, finite): pk_delta_out the CIC
 * density's P(k) on the mesh, delta = mesh N^3 / n^3 - 1, divided by the CIC window |W(k)|^2; pk_psi_out
 * P_Psi(k) = V sum_c sum w |Psi_c,k|^2 / (sum w / 3) / n^6 of the displacement on the particle grid
 * (n_part in the transform sizes of dj_pm_config's Fourier resampling); mode weights 1 on the kz = 0 and
 * kz = N/2 planes, 2 elsewhere; k_out the geometric centres sqrt(edges[b] edges[b + 1]).  k_out and
 * either spectrum may be NULL. */
dj_status dj_power_spectra_binned(dj_ctx *ctx, const void *psi, int32_t nbins, const double *edges, double *k_out,
                                  double *pk_delta_out, double *pk_psi_out);

/* Allocation-free dj_field_power_spectrum, for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc); no counterpart in the
 * reference.  All pointers are device pointers: the field (res_pm^3 values of the context's precision)
 * and the outputs (d_k_out and d_nmodes_out may be NULL).  The call only enqueues work on the
 * context's stream (results are ready after dj_synchronize) and allocates nothing: its bin sums live in
 * a static device array of DJ_PK_MAX_BINS x 4 doubles, one per GPU, held until the process ends, which
 * every call on that GPU shares, so calls that overlap on one GPU (contexts used from several threads)
 * race.  nbins above DJ_PK_MAX_BINS returns DJ_ERR_INVALID_ARGUMENT.  It uses the context's own
 * spectrum buffers, allocated by dj_create. */
dj_status dj_field_power_spectrum_noalloc(dj_ctx *ctx, const void *d_field, int32_t nbins, int32_t logarithmic,
                                          double *d_k_out, double *d_pk_out, int64_t *d_nmodes_out);

/* Host helpers of the IC pipeline: the N-GenIC Fourier white noise of `seed` (ngenic_field_gsl,
 * ngenic_gsl.jl:126-165), (res, res, res/2+1) interleaved complex, halved along z, element (i,j,k)
 * at i + res*(j + res*k); and the GSL ranlxd1 stream (ngenic_gsl.jl:34-98). */
dj_status dj_ngenic_field(uint32_t seed, int32_t res, double *out);
/* the same field generated on GPU `device` (ngenic_field_gsl, ngenic_gsl.jl:126-165): the seed table on the
 * host, each (i, j) column's ranlxd1 generator in one device thread; the uniforms are the host's bits,
 * sqrt(-log u), cos and sin the device's FP64 functions (within 2 ulp of the host's).  dj_ics_lpt without
 * white noise generates its field this way. */
dj_status dj_ngenic_field_on_device(uint32_t seed, int32_t res, int32_t device, double *out);
dj_status dj_ranlxd1_uniforms(uint32_t seed, int32_t n, double *out);

/* Reverse mode (no counterpart in the reference, whose PM path has no adjoint: NR/README.md:169; its
 * conventions follow the CIC rrule of field/sheet_deposit.jl:118-132 and the FFT rrules of
 * lpt/nlpt_core.jl:682-713).  Built for finite-difference and spectral (ik) gradients, with sheet resampling
 * (Fourier or linear; the resampled points' cotangents go back to Psi through the transposed interpolation)
 * when (n_resample n_part)^3 < 2^31, and without the learned conv net; other options return
 * DJ_ERR_UNSUPPORTED.  All arrays are host arrays
 * in the context's layout and precision, except where an entry point's name ends in _device. */

/* psi_bar_out = (d acc / d psi)^T acc_bar of dj_acceleration at psi (NULL: the context's Psi) */
dj_status dj_vjp_acceleration(dj_ctx *ctx, const void *psi, const void *acc_bar, void *psi_bar_out);
/* psi_bar_out = (d mesh / d psi)^T mesh_bar of dj_deposit (no shift) at psi (NULL: the context's Psi).
 * psi_bar_out NULL keeps the result on the device as the context's resident cotangents (Psibar, Pibar = 0),
 * which dj_vjp_run and dj_vjp_ics_lpt take when their cotangent pointers are NULL (DJ_ERR_INVALID_ARGUMENT
 * when none are resident: dj_vjp_acceleration, and calls given host cotangents, overwrite them). */
dj_status dj_vjp_deposit(dj_ctx *ctx, const void *psi, const void *mesh_bar, void *psi_bar_out);
/* dj_vjp_deposit at the context's Psi with mesh_bar in device memory (N^3 values of the context's precision);
 * the result stays resident */
dj_status dj_vjp_deposit_device(dj_ctx *ctx, const void *d_mesh_bar);
/* The differentiation workload's loss on the device: L = sum (delta - target)^2 over the mesh, delta =
 * mesh / rho - 1 the unshifted deposit of the context's Psi, rho = n^3 / N^3 (float deposit mode only).
 * d_target: N^3 FP64 values in device memory, or NULL for zero.  loss_out (host) gets L, summed in FP64 in
 * a fixed order; the cotangent of Psi, (d mesh / d psi)^T (2 / rho)(delta - target), is left resident for
 * dj_vjp_run / dj_vjp_ics_lpt(_device).  Replaces dj_deposit, a loss on the host and dj_vjp_deposit. */
dj_status dj_density_loss_device(dj_ctx *ctx, const double *d_target, double *loss_out);
/* dj_run's n_steps DKD steps from the current state, keeping (Psi, Pi) before each step on the device
 * (2 x 3 n^3 values per step) for dj_vjp_run, and, when the device has room for them, each step's tile
 * lists (about 2.5 n^3 entries of 4 values per step), which dj_vjp_run then uses instead of sorting the
 * step's positions again; the final state is the context's state as after dj_run */
dj_status dj_run_checkpointed(dj_ctx *ctx, int32_t n_steps, const double *dd1, const double *alpha,
                              const double *beta, const double *dd2);
/* reverse of the last dj_run_checkpointed: (psi_bar, pi_bar), the cotangents of the final (Psi, Pi), are
 * replaced by those of the initial state; each step's force is recomputed from its checkpoint.  Both NULL:
 * the context's resident cotangents, left resident */
dj_status dj_vjp_run(dj_ctx *ctx, void *psi_bar, void *pi_bar);
/* reverse of dj_ics_lpt (same seed / white, tables, D, order): white_bar_out (n^3 FP64) = the cotangent of
 * the real white-noise field given psi_bar, pi_bar of the ICs (both NULL: the resident cotangents).  When
 * the context's last dj_ics_lpt(_device) had the same inputs (n, L, order, table values, and the same seed
 * or white-noise values, compared bit for bit on the device), the A_ik fields it kept are used instead of
 * recomputing them; the result is the same bits either way. */
dj_status dj_vjp_ics_lpt(dj_ctx *ctx, uint32_t seed, const double *white, const double *k_tab, const double *pk_tab,
                         int32_t ntab, double D, int32_t order, const void *psi_bar, const void *pi_bar,
                         double *white_bar_out);
/* dj_vjp_ics_lpt from the resident cotangents with the white noise (or NULL: seed) and its cotangent
 * d_white_bar_out (n^3 FP64) in device memory */
dj_status dj_vjp_ics_lpt_device(dj_ctx *ctx, uint32_t seed, const double *d_white, const double *k_tab,
                                const double *pk_tab, int32_t ntab, double D, int32_t order,
                                double *d_white_bar_out);

/* Learned force correction (no counterpart in the reference; for callers that run the corrected PM of
 * the learned-correction workload).  A residual 3-D convolution net on the mesh, BF16 weights and
 * activations with FP32 accumulation on MFMA, adds s_out NN(log(1 + delta), a) to the potential of every
 * force pass.  Built for FP32 contexts with the float deposit, finite-difference gradients and res_pm a
 * multiple of 64.  Parameters: DJ_NN_PARAMS FP32 values, 4 layers of W[27][16][16] (tap (dz,dy,dx)
 * z-major, out, in; layer 1 uses inputs 0, 1 and layer 4 output 0, the rest is zero) then b[16]. */
#define DJ_NN_PARAMS (4 * (27 * 16 * 16 + 16))
dj_status dj_nn_set_params(dj_ctx *ctx, const float *params);  /* also resets the optimiser state */
dj_status dj_nn_get_params(dj_ctx *ctx, float *params);
/* on: add the correction, scaled by out_scale, in every force evaluation */
dj_status dj_nn_enable(dj_ctx *ctx, int32_t on, double out_scale);
/* the net's time input per step of dj_run (n values: the scale factor at each step's force) */
dj_status dj_nn_set_schedule(dj_ctx *ctx, int32_t n, const double *a_force);
/* the correction (N^3 floats) of a density mesh (N^3 floats, the unnormalised deposit) at scale factor a */
dj_status dj_nn_forward(dj_ctx *ctx, const float *dens, double a, float *dphi_out);
/* parameter gradient (DJ_NN_PARAMS floats) of <gphi, dphi(dens, a)> */
dj_status dj_nn_backward(dj_ctx *ctx, const float *dens, double a, const float *gphi, float *grad_out);
/* one Adam step (learning rate lr, step >= 1) on L = weight mean((acc(psi) - acc_target)^2) of the corrected
 * force at the snapshot psi (antialias 0); loss_out receives L before the update */
dj_status dj_nn_train_step(dj_ctx *ctx, const void *psi, const void *acc_target, double a, double weight, double lr,
                           int32_t step, double *loss_out);

/* Fitted transfer function of the Poisson kernel (no counterpart in the reference; part of the learned
 * correction): every force pass multiplies the kernel by g(|k|), with g given at nbins >= 2 points
 * |k| = i kmax / (nbins - 1), i < nbins (k in h/Mpc), linear in between and g[nbins - 1] beyond kmax.
 * gain NULL (or nbins 0) removes it.  Works with every gradient and deconvolution option. */
dj_status dj_set_kgain(dj_ctx *ctx, const double *gain, int32_t nbins, double kmax);

/* NaN/Inf scan of the particle state (failure detection); DJ_ERR_NONFINITE if any */
dj_status dj_check_finite(dj_ctx *ctx);

dj_status dj_synchronize(dj_ctx *ctx);
dj_status dj_set_profiling(dj_ctx *ctx, int32_t on);
dj_status dj_reset_stats(dj_ctx *ctx);
dj_status dj_kernel_stats(dj_ctx *ctx, dj_kernel_stat *out, int32_t cap, int32_t *n_out);

/* Stand-alone 3-D transform with the library's kernels (tests and tools): rfft (direction -1)
 * of a real (N,N,N) host array into the x-halved spectrum out[kx + (N/2+1)*(ky + N*kz)], or the
 * unnormalised inverse (direction +1) back to real.  precision 32 or 64. */
dj_status dj_fft3d(int32_t N, int32_t precision, int32_t direction, const void *in, void *out, int32_t device);

/* library version string */
const char *dj_version(void);

#ifdef __cplusplus
}
#endif
#endif
