// Kernel launch interface of libwlc_mde.so (internal).
#pragma once
#include <hip/hip_runtime.h>

#include "wlc_internal.h"

namespace wlc {

struct AxisDev {
  int n, npass;
  int radix[8];
  const double2* tw;  // exp(-2 pi i t / n), t < n
};

// Half spectrum layout: [i][ky][kx / 4][z][kx % 4] (complex FP64).  Plane kernels then
// move whole 64-byte runs of 4 kx, and the four waves of a kernel-B workgroup (one per
// kx of a tile) read interleaved pieces of the same cache lines.
constexpr int kKxTile = 4;
__host__ __device__ inline int kx_tiles(int nxh) { return (nxh + kKxTile - 1) / kKxTile; }
__host__ __device__ inline long long spec_index(int i, int ky, int kx, int z, int ny, int nt, int nz) {
  return ((((long long)i * ny + ky) * nt + (kx >> 2)) * nz + z) * kKxTile + (kx & 3);
}

// The half spectrum is stored as double2 (FP64) or, in the FP32-spectrum mode, as float2;
// every kernel computes in FP64 and converts where it loads or stores the spectrum
// (round to nearest on the store).
__device__ __forceinline__ double2 spec_ld(const double2* p) { return *p; }
__device__ __forceinline__ double2 spec_ld(const float2* p) {
  const float2 v = *p;
  return make_double2((double)v.x, (double)v.y);
}
__device__ __forceinline__ void spec_st(double2* p, double2 v) { *p = v; }
__device__ __forceinline__ void spec_st(float2* p, double2 v) {
  *p = make_float2((float)v.x, (float)v.y);
}

struct Geom {
  int nx, ny, nz, nxh;
  int nt;          // kx tiles
  int spec32;      // 1: the half spectrum is stored as float2
  long long zs;    // real-space z-plane stride (doubles): nx*ny + pad
  long long hs;    // real-space field / harmonic stride (doubles): nz*zs + pad
  AxisDev ax, ay, az;
};

// Kernel A: BDF right-hand side of one propagator step fused with the x (R2C, two
// rows packed per complex transform) and y passes of the forward 3D FFT.
struct RhsArgs {
  const double* h[3];  // history slices [M_h][nz][ny][nx] (nullptr beyond the BDF order)
  int hM[3];           // harmonics held by each history slice
  double a[3], b[3];   // P = sum_j a_j q_j - ds w sum_j b_j q_j
  double ds;
  const double* w;     // field of the target block [nz][ny][nx]
  const double* zero;  // a field of zeros, read in place of a history plane that is zero
  int M;               // harmonics of the target slice
  double2* spec;       // half spectrum, layout of spec_index
};

// Kernel B: z pass of the FFT, per-k solve of G(k) x = P(k), inverse z pass.
struct SolveArgs {
  double2* spec;
  int M, L, TX;        // harmonics, Lmax, kx columns per workgroup (TX * nz threads)
  double cdelta, sigma, ds, dslam;
  double kfac[3];      // 2 pi / box per axis
  int solve;           // 0: FFT round trip only (test hook)
  int generic;         // 1: force the generic (any nz) kernel
  int occ;             // nz = 64 kernel: target waves per SIMD (1 or 2)
  int split;           // harmonics of a k-point split over two waves / threads by degree
                       // (nz = 64: Lmax 5-7; nz = 128: Lmax 5-8; other nz: Lmax 4-7): 0 off;
                       // nz = 64: 1 two columns per workgroup, loaded directly, 2 a whole 4-kx
                       // tile per workgroup, 3 half a tile per workgroup, both through LDS
                       // (2 and 3: FP64 spectrum; FP32 runs as 1); elsewhere any nonzero
                       // value selects the split
};

// Kernel C: inverse y and x (C2R) passes, scaling by 1/N, store of q(s).
struct InvArgs {
  const double2* spec;
  int M;
  double scale;
  double* out;         // [M][nz][ny][nx]
};

// One instance of a batched launch: the arguments a single launch of the same kernel would
// take.  A batched launch runs the kernel's grid once per job (job = blockIdx.z); every job
// of one launch selects the same kernel (same grid sizes, Lmax, spectrum type and BDF
// order).  A job whose *skip is nonzero returns at once (an instance that has converged).
struct PlaneJob {
  Geom g;
  InvArgs c;         // kernel C part (inverse xy, store of q(s))
  RhsArgs a;         // kernel A part (right-hand side, forward xy)
  const int* skip;   // nullptr: always run
};
struct ZJob {
  Geom g;
  SolveArgs a;
  const int* skip;
};

// LDS of one workgroup on gfx950
constexpr size_t kMaxLdsBytes = 160 * 1024;
// true when the plane kernels can hold an nx x ny half plane (64 x 64 and 128 x 128 have
// their own kernels; other sizes keep the whole ny (nx/2 + 1) half plane in LDS)
bool plane_supported(int nx, int ny);

hipError_t launch_rhs_fft(const Geom& g, const RhsArgs& a, int order, hipStream_t st);
hipError_t launch_zsolve(const Geom& g, const SolveArgs& a, hipStream_t st);
hipError_t launch_ifft_store(const Geom& g, const InvArgs& a, hipStream_t st);
// C of one step fused with A of the next: inverts and stores c.M harmonics of the
// solved step, whose q stays on chip as history slice 0 of `a` (a.h[0] is not read),
// and writes the next step's half spectrum for a.M harmonics
hipError_t launch_inv_rhs(const Geom& g, const InvArgs& c, const RhsArgs& a, int order,
                          hipStream_t st);

// batched twins: `jobs` is a device array of n jobs whose kernel selection equals that of
// the representative rep (the host groups jobs so); same stream semantics
hipError_t launch_rhs_fft_batch(const PlaneJob& rep, int order, const PlaneJob* jobs, int n,
                                hipStream_t st);
hipError_t launch_ifft_store_batch(const PlaneJob& rep, const PlaneJob* jobs, int n, hipStream_t st);
hipError_t launch_inv_rhs_batch(const PlaneJob& rep, int order, const PlaneJob* jobs, int n,
                                hipStream_t st);
hipError_t launch_zsolve_batch(const ZJob& rep, const ZJob* jobs, int n, hipStream_t st);
// key of the kernel (and grid) a plane / z-solve launch with these arguments selects: jobs
// with equal keys may share a batched launch
struct LaunchKey {
  const void* k;
  unsigned gx, gy, threads;
  size_t lds;
  bool operator==(const LaunchKey& o) const {
    return k == o.k && gx == o.gx && gy == o.gy && threads == o.threads && lds == o.lds;
  }
};
LaunchKey plane_key(const Geom& g, int kind, int order, int planes);
LaunchKey zsolve_key(const Geom& g, const SolveArgs& a);

// small elementwise/reduction kernels
hipError_t launch_set_delta(const Geom& g, double* slice, int M, hipStream_t st);
hipError_t launch_fill(double* dst, double v, long long n, hipStream_t st);
// y += a x over n doubles
hipError_t launch_axpy(double* y, double a, const double* x, long long n, hipStream_t st);
// rho_t += wgt_t * sum_{i < lim_t} q_i qs_i for up to two targets (t = 0, 1)
hipError_t launch_density_acc(const Geom& g, const double* q, const double* qs, double* rho0,
                              int lim0, double w0, double* rho1, int lim1, double w1,
                              hipStream_t st);
// partial[b] = sum over a chunk of r of W_r * f(r), f = sum_{t<nf} src[t*N + r]
hipError_t launch_weighted_sum(const Geom& g, const double* src, int nf, double dV,
                               double* partial, int nparts, double* result, hipStream_t st);
// phi_b = V * rho_b / totden (totden read from device)
hipError_t launch_phi(const Geom& g, const double* rho, int nblk, double V, const double* totden,
                      double* phi, hipStream_t st);

hipError_t upload_orientation_tables(const OrientTables& t);

// test hook: rotated-frame solve of nk vectors (M complex each, interleaved)
hipError_t launch_test_solve(int L, int nk, const double* kvec, double cdelta, double sigma,
                             double ds, double dslam, double2* v, hipStream_t st);

}  // namespace wlc
