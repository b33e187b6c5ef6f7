/*
 * wlc_mde.h -- C ABI of libwlc_mde.so, the MI355X (gfx950) implementation of the
 * wormlike-chain modified diffusion equation (MDE) propagator solve of
 * tjz2026/wormlike-chain-triblock.
 *
 * The reference exposes the solve as Fortran module procedures that read and write
 * module globals (module wormlike_MDE, wormlike_MDE_2dcomp_mod.f90:227-244, 961-973):
 *   MDE_q(R2C_FFT)      forward propagator  qA, qB          (:231-956)
 *   MDE_qstar(R2C_FFT)  backward propagator qAstar, qBstar  (:961-1557)
 * with its inputs set beforehand by initialize/init_inverse_sparse_GAB and its
 * outputs consumed by `density` (SCFT_Density_FE.f90:3-271).  Every entry point
 * below replaces one of those pieces; the Fortran binding that keeps the reference
 * names lives in wormlike-chain-triblock_amd/fortran/wormlike_mde_gpu.f90.
 *
 * Conventions: all arrays are FP64, fields are [nz][ny][nx] with x fastest (the
 * reference's kD map, grid_mod.f90:356-379); block fields are stacked [nblk][...];
 * propagator slices are [M][nz][ny][nx] with harmonic i <-> (l, m), i = l*l + l + m
 * (SPH_mod.f90:214-222).  Host pointers are read/written synchronously; the handle
 * owns all device memory.  Every function returns WLC_OK (0) or a negative status;
 * nothing aborts.  A handle is not thread-safe; separate handles are independent.
 */
#ifndef WLC_MDE_H
#define WLC_MDE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WLC_OK 0
#define WLC_EINVAL (-1)       /* bad parameters or arguments */
#define WLC_ENOMEM (-2)       /* device allocation failed */
#define WLC_EHIP (-3)         /* a HIP runtime call failed */
#define WLC_ESTATE (-4)       /* call out of order (e.g. density before the sweeps) */
#define WLC_ENONFINITE (-5)   /* Q, totden or phi not finite */

#define WLC_MAX_BLOCKS 3
#define WLC_MAX_L 8

typedef struct wlc_params {
  int nx, ny, nz;                 /* grid; even edges with factors 2 and 3 only */
  double box[3];                  /* box edges (x, y, z) in units of chain length */
  int nblk;                       /* 1..3 blocks (A, B, C) */
  int Ns;                         /* contour steps of the chain (reference Nmax); ds = 1/Ns */
  int Nblk[WLC_MAX_BLOCKS];       /* contour steps per block (NA, NB, NC); sum = Ns */
  int Lmax[WLC_MAX_BLOCKS];       /* harmonic cut-off per block, 0..WLC_MAX_L */
  double lambda[WLC_MAX_BLOCKS];  /* angular coefficient Loa/(2 kappa) per block */
  int keep_qstar;                 /* 1: keep every backward slice (for wlc_mde_get_q) */
  int ckpt_every;                 /* forward store: 0 = every slice when it fits in device
                                     memory, else segments of 16 steps: the lowest segments
                                     that fit are kept whole, the others checkpointed and
                                     recomputed during the backward sweep; K >= 3 = every
                                     segment of K steps checkpointed (no counterpart in the
                                     reference, which stores every slice,
                                     init_array_2dcomp_mod.f90:47-51) */
  int spectrum_fp32;              /* 1: store the half spectrum (the FFT / solve intermediate)
                                     as FP32, every arithmetic step stays FP64; halves the
                                     spectrum traffic at ~1e-7 rounding per store.  No
                                     counterpart in the reference beyond its own FP32 staging
                                     of the G-applied spectrum (wormlike_MDE_2dcomp_mod.f90:429) */
  int reserved[5];                /* must be zero */
} wlc_params;

typedef struct wlc_handle wlc_handle;

/* Create a solver for one chain architecture and grid; builds the orientation
 * tables and FFT plans and allocates the forward propagator store.
 * Replaces: initialize's derived parameters (initialize_mod.f90:140-199), the
 * allocations of init_arrays_2decomp (init_array_2dcomp_mod.f90:47-51) and
 * decomp2d_mpi_3d_dft_r2c_init (2dcomp_fft_mod.f90:39-73). */
int wlc_mde_create(const wlc_params* p, wlc_handle** out);

/* Change the box.  Replaces init_inverse_sparse_GAB(R2C_FFT)
 * (Gmatrix_2dcomp_mod.f90:222-650): there are no stored G^-1 tables, the k-aligned
 * solve reads the box through its wave-number factors. */
int wlc_mde_set_box(wlc_handle* h, const double box[3]);

/* Set the block fields w [nblk][nz][ny][nx] from host / device memory.
 * Replaces the module globals wA, wB (global_parameters.f90; ABC: w_blk). */
int wlc_mde_set_w(wlc_handle* h, const double* w_host);
int wlc_mde_set_w_device(wlc_handle* h, const double* w_dev);

/* Forward propagator q(s), s = 0..Ns, q(0) = delta_{i0}.
 * Replaces MDE_q(R2C_FFT) (wormlike_MDE_2dcomp_mod.f90:231-956). */
int wlc_mde_forward(wlc_handle* h);

/* Backward propagator q*(s), q*(Ns) = delta_{i0}; also accumulates the contour
 * quadrature of sum_i q_i q*_i per block (needs a forward sweep on the same w).
 * Replaces MDE_qstar(R2C_FFT) (wormlike_MDE_2dcomp_mod.f90:961-1557) and the
 * contour loop of density (SCFT_Density_FE.f90:97-118). */
int wlc_mde_backward(wlc_handle* h);

/* Q, totden and phi [nblk][nz][ny][nx] (host or device pointer; phi may be NULL).
 * Replaces density's Q / totden / phi part (SCFT_Density_FE.f90:48-159,
 * SCFT_Density_FE_ABC.f90:98-139). */
int wlc_mde_density(wlc_handle* h, double* phi_host, double* Q, double* totden);
int wlc_mde_density_device(wlc_handle* h, double* phi_dev, double* Q, double* totden);

/* Copy slice s of the forward (dir = +1) or backward (dir = -1) propagator to host
 * memory as [M(s)][nz][ny][nx].  The backward store needs keep_qstar = 1, except for
 * s = 0 which is always kept.  Replaces reading qA/qB/qAstar/qBstar
 * (init_array_2dcomp_mod.f90:48-51) for tests and diagnostics. */
int wlc_mde_get_q(wlc_handle* h, int dir, int s, double* q_host);

/* harmonics held by slice s (M of the block that contains contour point s) */
int wlc_mde_slice_M(const wlc_handle* h, int s);

/* forward slices 0 .. n-1 kept whole by the forward sweep: Ns + 1 with the full store;
 * in checkpoint mode the slices of the lowest segments that fitted beside the checkpoints
 * (0 when every segment is recomputed).  No counterpart in the reference. */
int wlc_mde_forward_stored(const wlc_handle* h);

/* block index (0-based) that owns contour point s */
int wlc_mde_block_of(const wlc_handle* h, int s);

/* wait for all work queued on the handle's stream */
int wlc_mde_sync(wlc_handle* h);

/* the HIP stream (hipStream_t) all work of the handle is queued on */
void* wlc_mde_stream(wlc_handle* h);

/* Switch the storage precision of the half spectrum (see wlc_params.spectrum_fp32) between
 * sweeps, e.g. FP32 while an SCFT iteration is far from convergence and FP64 for the final
 * iterations and the free energy.  Waits for queued work; the next wlc_mde_density needs
 * new sweeps.  No counterpart in the reference. */
int wlc_mde_set_spectrum_fp32(wlc_handle* h, int on);

/* Per-kernel timing with HIP events.  enable: 0 off; 1 every launch of the step kernels
 * is bracketed by events on the stream it runs on, read back when the host waits anyway
 * (wlc_mde_density*, wlc_mde_get_timing, wlc_mde_set_timing), so the sweeps keep running
 * concurrently and a duration includes the time a kernel shares the GPU; 2 as 1, but
 * wlc_mde_forward waits for its sweep, so each step kernel is timed with the GPU to
 * itself (the density terms still overlap the backward sweep).  wlc_mde_get_timing returns,
 * for kernel k in {0: rhs+xy FFT, 1: z FFT+solve, 2: inverse xy FFT, 3: density
 * accumulation, 4: inverse xy FFT of one step fused with rhs+xy FFT of the next}, the
 * summed milliseconds and the launch count since enabling.
 * No counterpart in the reference (which times only the whole run, SCFT_main.f90:51-55). */
int wlc_mde_set_timing(wlc_handle* h, int enable);
int wlc_mde_get_timing(wlc_handle* h, double* ms, long long* count, int nkernels);

/* Event pair around queued work on the handle's stream; elapsed milliseconds
 * between the two records.  No counterpart in the reference. */
int wlc_mde_event_record(wlc_handle* h, int slot);
int wlc_mde_event_elapsed(wlc_handle* h, int slot_a, int slot_b, double* ms);

/* One contour step of the handle's chain on caller device buffers: the step of direction
 * dir (+1 forward, -1 backward) whose target is contour point s, i.e. the BDF right-hand
 * side from hist_dev[j] = slice s - dir (j + 1) (j < BDF order = min(s, 3) forward,
 * min(Ns - s, 3) backward; entries beyond the order are not read), the R2C FFT, the solve
 * of G(k) q = P and the C2R FFT into out_dev.  Slices are dense [M][nz][ny][nx] with the M
 * of the block owning their contour point, w_dev is [nblk][nz][ny][nx]; the target takes
 * its block's field, Lmax and lambda and the history is zero-padded / truncated to its
 * harmonics, as in the sweeps.  Queued on the handle's stream (no host wait); the handle's
 * half spectrum is the scratch.  Replaces the body of one iteration of the contour loops of
 * MDE_q / MDE_qstar (wormlike_MDE_2dcomp_mod.f90:323-455, :584-650; :1035-1110). */
int wlc_mde_step(wlc_handle* h, int dir, int s, const double* const hist_dev[3],
                 const double* w_dev, double* out_dev);

/* wlc_mde_step without a handle and without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The chain is described by p
 * (validated as in wlc_mde_create), the work is queued on `stream` (a hipStream_t, NULL =
 * the null stream).  The scratch (FFT twiddles, one zero field, the half spectrum of the
 * target's harmonics) lives in a fixed 128 MiB device array that each GPU the library runs
 * on holds until the process ends; a step whose zero field plus half spectrum exceeds it
 * (e.g. 64^3 with Lmax above 6) is rejected with WLC_EINVAL.  Every call on one GPU shares
 * that array, so calls that overlap on one GPU (different streams or threads) race. */
int wlc_mde_step_noalloc(const wlc_params* p, int dir, int s, const double* const hist_dev[3],
                         const double* w_dev, double* out_dev, void* stream);

/* ---- Batched SCFT: many independent instances on one GPU -------------------------------
 * The reference runs one SCFT per process (cal_scft, SCFT_control.f90:235-262: MDE_q,
 * MDE_qstar, density, free_energy, iterate; field update AB_iterate, SCFT_iterate.f90:27-333).
 * A batch holds n instances, each with its own wlc_params (grid, box, blocks, Lmax, lambda;
 * keep_qstar = ckpt_every = 0: the full forward store must fit) and interaction matrix
 * chi[i][3][3] (chi_ab N, row a, column b), and advances all of them together: each contour
 * step is one launch per kernel variant covering every instance, and the field update
 * (simple, then Anderson mixing; FP64) runs on the device.  One SCFT iteration of all
 * instances is a fixed launch sequence, replayed from a HIP graph when use_graph is set.
 * The field update is that of the Fortran host module scft_iterate (forces, residuals,
 * convergence test; Anderson coefficients from a Gram matrix of the dW history). */
typedef struct wlc_batch wlc_batch;

typedef struct wlc_scft_opts {
  int max_iter;         /* iterations per instance (simple + Anderson) */
  int n_simple;         /* simple-mixing iterations before Anderson (Num_simple_mixing_AB) */
  int n_hist;           /* Anderson history depth, 1..16 (Anderson_nim_AB) */
  int form;             /* 0: w = sum c Wout + lam sum c dW (the reference's update);
                           1: w = sum c w + lam sum c dW */
  double lam_simple;    /* simple-mixing step (lanbtWA / lanbtWB) */
  double lam_anderson;  /* Anderson step (lambda_WAB_anderson) */
  double tol;           /* converged when every species' residual is below tol */
  int adaptive;         /* 1: the Anderson step grows 5 % per iteration that lowers the
                           largest residual, up to lam_max; when that residual exceeds ten
                           times the best since the last restart the history restarts from the
                           current entry and the step halves (no counterpart in the reference,
                           whose step is fixed) */
  int no_compact;       /* 0 (default): once at most half of the instances in the launch
                           sequence still advance, rebuild it over those (fewer launches per
                           iteration, same results); 1: keep every instance in it */
  int reserved[2];      /* must be zero */
  double lam_max;       /* adaptive mode: largest Anderson step (0: 1.0) */
} wlc_scft_opts;

typedef struct wlc_scft_status {
  int iterations;       /* iterations done */
  int converged;        /* 1: residual below tol */
  int done;             /* converged or max_iter reached: the instance no longer advances */
  int pad;
  double F;             /* free energy per chain of the last iteration (kT) */
  double lnQ;
  double resid[3];      /* residual per species of the last iteration */
  int restarts;         /* Anderson history restarts (adaptive mode) */
  int pad2;
  double time_ms;       /* wall time from the first iteration to the end of the iteration in
                           which the instance finished (-1 while it runs) */
  double elapsed_ms;    /* wall time of all iterations so far */
} wlc_scft_status;

/* Create a batch of n instances (params[n], chi[n][9]); opts NULL = max_iter 500,
 * n_simple 20, n_hist 10, form 0, lam_simple 0.05, lam_anderson 0.1, tol 1e-6.
 * Replaces starting n reference processes (SCFT_main.f90:1-55). */
int wlc_batch_create(const wlc_params* params, const double* chi, int n, const wlc_scft_opts* opts,
                     wlc_batch** out);
/* number of instances */
int wlc_batch_size(const wlc_batch* b);
/* launches of one sweep pair with observables (with_update = 0) or of one SCFT iteration */
int wlc_batch_launches(const wlc_batch* b, int with_update);
/* fields of instance i, [nblk][nz][ny][nx] host arrays (the module globals wA, wB / w_blk) */
int wlc_batch_set_w(wlc_batch* b, int i, const double* w_host);
int wlc_batch_get_w(wlc_batch* b, int i, double* w_host);
/* forward and backward sweeps, Q, totden and phi of every instance still advancing
 * (MDE_q, MDE_qstar and density of each, wormlike_MDE_2dcomp_mod.f90:231-1557,
 * SCFT_Density_FE.f90:3-159) */
int wlc_batch_sweep(wlc_batch* b);
/* Q, totden and phi [nblk][nz][ny][nx] of instance i after wlc_batch_sweep or an iteration */
int wlc_batch_density(wlc_batch* b, int i, double* phi_host, double* Q, double* totden);
/* forward slice s of instance i, [M(s)][nz][ny][nx] (tests) */
int wlc_batch_get_q(wlc_batch* b, int i, int s, double* q_host);
/* up to niter SCFT iterations of every instance still advancing (sweeps, density, free
 * energy, field update, convergence test); stops early when none is left; *n_active =
 * instances still advancing */
int wlc_batch_iterate(wlc_batch* b, int niter, int use_graph, int* n_active);
int wlc_batch_status(wlc_batch* b, int i, wlc_scft_status* st);
void wlc_batch_destroy(wlc_batch* b);

/* ---- Warm-start network ------------------------------------------------------------------
 * A 3D convolutional U-Net (levels n, n/2, n/4, n/8; base 32 channels; 1.9 M parameters)
 * that proposes initial SCFT fields.  The reference has no counterpart: every SCFT starts
 * from the hand-built fields of init_w (init_array_2dcomp_mod.f90:843-1065), which this
 * replaces as the source of w at iteration 0; plain SCFT still decides the converged state.
 * Inputs x are [batch][n][n][n][11] (3 seed fields, then f_A, f_B, f_C, chi_AB, chi_BC,
 * chi_AC, lambda, box broadcast), outputs and targets [batch][n][n][n][3]; FP32 host arrays.
 * Training and inference run hand-written kernels: bf16 matrix-core convolutions with FP32
 * accumulation, FP32 master weights and Adam. */
typedef struct wlc_unet wlc_unet;
int wlc_unet_create(int n, int batch, unsigned long long seed, wlc_unet** out);
/* new grid edge (multiple of 8) and batch for the next calls; parameters and Adam state kept
 * (the network is fully convolutional) */
int wlc_unet_resize(wlc_unet* u, int n, int batch);
/* trainable parameters (padded = 1: the flat layout's length, including zero padding) */
long long wlc_unet_num_params(const wlc_unet* u, int padded);
/* layer i: info[8] = cin, cout (padded), cin, cout (real), level, relu, weight offset, bias
 * offset in the flat parameters (W [cout][27][cin], b [cout]) */
int wlc_unet_layer(const wlc_unet* u, int i, int* info);
int wlc_unet_forward(wlc_unet* u, const float* x, float* y);
/* forward + backward: parameter gradients of the mean-squared error against t; *loss */
int wlc_unet_gradients(wlc_unet* u, const float* x, const float* t, float* loss);
/* one Adam step (beta 0.9 / 0.999) with the last gradients */
int wlc_unet_adam(wlc_unet* u, float lr);
int wlc_unet_train_step(wlc_unet* u, const float* x, const float* t, float lr, float* loss);
/* flat parameters (which = 0) or gradients (which = 1) to host; parameters from host */
int wlc_unet_get(wlc_unet* u, int which, float* out);
int wlc_unet_set_params(wlc_unet* u, const float* in);
void wlc_unet_destroy(wlc_unet* u);
/* Test hook: one convolution (bf16 operands, FP32 accumulation) and its gradients on host
 * arrays: x [V][cin], w [cout][27][cin], b [cout], dy [V][cout] -> y, dx, dw, db. */
int wlc_test_conv3d(int B, int D, int H, int W, int cin, int cout, const float* x, const float* w,
                    const float* b, const float* dy, float* y, float* dx, float* dw, float* db);

void wlc_mde_destroy(wlc_handle* h);

/* message of the last failure on this thread ("" if none) */
const char* wlc_mde_last_error(void);

/* Test hook: the k-aligned solve of G(k) x = v for nk k-vectors on the device
 * (kvec [nk][3] in wave-number units, v [nk][M] complex interleaved, in place).
 * Replaces one G^-1 apply (4 x mkl_dcsrsymv, wormlike_MDE_2dcomp_mod.f90:395-409). */
int wlc_test_solve(int L, int nk, const double* kvec_host, double cdelta, double sigma,
                   double ds, double dslam, double* v_host);

#ifdef __cplusplus
}
#endif
#endif
