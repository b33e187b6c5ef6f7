/*
 * wlc_oracle.h -- CPU oracle (TEST INFRASTRUCTURE ONLY) for the wormlike-chain
 * MDE propagator solve of tjz2026/wormlike-chain-triblock.
 *
 * This header and wlc_oracle.c restate, in plain C with OpenMP, the reference's
 * per-contour-step dataflow (BDF1/2/3 right-hand side, per-harmonic 3D R2C FFT,
 * per-k dense G(k)^-1 applied through its lower triangle, C2R inverse FFT / N),
 * the real-spherical-harmonic algebra it is built on, and the contour/volume
 * quadrature of the density routine.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker: the product
 * (libwlc_mde.so) never links or calls it.
 */
#ifndef WLC_ORACLE_H
#define WLC_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
  int nx, ny, nz;     /* grid, x fastest in memory */
  double box[3];      /* box edges (x, y, z), units of chain length */
  int nblk;           /* number of blocks, 1..3 (A, B, C) */
  int Ns;             /* contour steps of the whole chain (reference Nmax) */
  int Nblk[3];        /* contour steps per block (reference NA, NB, NC) */
  int L[3];           /* Lmax per block; M = (L+1)^2 harmonics */
  double lam[3];      /* angular coefficient Loa/(2 kappa) per block */
  int literal;        /* 0: intended FP64, 1: emulate the reference's FP32 literals/staging */
  int nthreads;       /* OpenMP threads (<=0: runtime default) */
} orc_params;

typedef struct orc_ctx orc_ctx;

orc_ctx* orc_create(const orc_params* p);
void orc_destroy(orc_ctx* c);

/* layout helpers: slice s of a propagator store has orc_slice_M(c,s) harmonics,
 * each an [nz][ny][nx] real field; slices are concatenated in s order. */
int orc_slice_M(const orc_ctx* c, int s);
long long orc_slice_offset(const orc_ctx* c, int s);   /* in doubles */
long long orc_store_size(const orc_ctx* c);            /* in doubles */
int orc_block_of(const orc_ctx* c, int s);

/* full sweeps; w is [nblk][nz][ny][nx] */
int orc_forward(orc_ctx* c, const double* w, double* q);
int orc_backward(orc_ctx* c, const double* w, double* qs);
/* density: phi [nblk][N]; Q, totden scalars */
int orc_density(orc_ctx* c, const double* q, const double* qs, double* phi,
                double* Q, double* totden);
/* ABC free energy (per chain, kT): -ln Q + (1/V) sum_r W_r [sum_{i<j} chi_ij phi_i phi_j
 *   - sum_i w_i phi_i - xi (1 - sum_i phi_i)], xi = mean of the block fields. chi is [3][3]. */
double orc_free_energy(const orc_ctx* c, const double* w, const double* phi,
                       double Q, const double* chi);

/* one contour step on given history (for tests); dir=+1 forward (target s from
 * s-1,s-2,s-3), dir=-1 backward (target s from s+1,s+2,s+3).  hist[j] points at
 * slice s-dir*(j+1) (NULL when unused); out receives slice s. */
int orc_step(orc_ctx* c, const double* w, int dir, int s, const double* const hist[3],
             double* out);

/* building blocks */
double orc_three_j(int l1, int l2, int l3, int m1, int m2, int m3);
double orc_triple_product(int l1, int l2, int l3, int m1, int m2, int m3, int literal);
void orc_rmatrices(int L, int literal, double* Rx, double* Ry, double* Rz); /* M*M row-major */
/* dense complex inverse of G = diag(c + ds lam l(l+1)) + i sigma ds (k . R); out M*M complex interleaved */
int orc_ginv(int L, double cdelta, double sigma, double ds, double lam, double kx, double ky,
             double kz, int literal, double* out);
/* unnormalised 3D transforms; spectrum is [nz][ny][nx/2+1] complex interleaved */
void orc_rfft3(int nx, int ny, int nz, const double* in, double* out);
void orc_irfft3(int nx, int ny, int nz, const double* in, double* out);
/* complex 1D FFT, in place, sign -1 forward / +1 inverse, unnormalised */
void orc_fft1(int n, double* x, int sign);
/* contour Simpson weights for a block of n intervals (n+1 points), spacing ds */
void orc_simpson_weights(int n, double ds, double* wts);
/* periodic 3D volume integral sum_r W_r f(r) */
double orc_volume_integral(int nx, int ny, int nz, const double* box, const double* f);

/* CPU-baseline helper: time nsteps BDF3 steps of block 0 in the forward direction
 * (factor build excluded); returns seconds, or <0 on error. */
double orc_time_bdf3_steps(orc_ctx* c, const double* w, int nsteps);
/* OpenMP threads for the following calls (n > 0) */
void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
