// Plane kernels of the propagator step (gfx950, FP64) and the small elementwise /
// reduction kernels of the density.
//
// A  k_rhs_fft_xy : per (z-plane, harmonic): BDF right-hand side of the reference
//                   (wormlike_MDE_2dcomp_mod.f90:325-343, 484-502, 584-608; junction
//                   padding :697-806) from the 1-3 history planes and w, then the x
//                   (R2C, two real rows per complex transform) and y passes of the
//                   forward FFT in LDS; writes the half spectrum plane.
// C  k_ifft_xy    : per (z-plane, harmonic): inverse y and x (C2R) passes, 1/N
//                   (wormlike_MDE_2dcomp_mod.f90:433-455), store of q(s).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "wlc_device.h"

namespace wlc {

namespace {
constexpr int kThreads = 256;
constexpr int kChunk = 8;  // plane elements each thread loads per batch
constexpr int kPairs = 4;  // x pairs each thread loads per batch (kernel A)
}  // namespace

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2).  When
// nz is a multiple of 8, XCD k takes the z-planes [k nz/8, (k+1) nz/8) of every
// harmonic, z fastest: the w planes of its z range are reused from L2 by every
// harmonic, and the two z-planes that share a half-spectrum cache line run back to
// back on one XCD.  Otherwise each XCD walks a contiguous range of (harmonic, z).
__device__ __forceinline__ void plane_of_block(int nz, int& z, int& grp) {
  const int nb = gridDim.x;
  const int b = blockIdx.x;
  if ((nz & 7) == 0) {
    const int zr = nz >> 3, t = b >> 3;
    grp = t / zr;
    z = (b & 7) * zr + (t - grp * zr);
    return;
  }
  const int q = nb >> 3;
  const int logical = (b < (q << 3)) ? ((b & 7) * q + (b >> 3)) : b;
  grp = logical / nz;
  z = logical - grp * nz;
}

// x then y forward transform of the packed plane in LDS (sizes compile-time when
// NX, NY are nonzero)
template <int NX, int NY, int MAXE>
__device__ __forceinline__ void plane_fft_x(double2* lds, const Geom& g) {
  if constexpr (NX > 0) FftCt<-1, NX, NY / 2, 1, NX, kThreads, true>::run(lds, g.ax.tw);
  else fft_axis<-1, MAXE>(lds, g.ax, g.ny / 2, 1, g.nx);
}
template <int NX, int NY, int MAXE, int DIR>
__device__ __forceinline__ void plane_fft_y(double2* lds, const Geom& g) {
  if constexpr (NX > 0) FftCt<DIR, NY, NX / 2 + 1, NX / 2 + 1, 1, kThreads, true>::run(lds, g.ay.tw);
  else fft_axis<DIR, MAXE>(lds, g.ay, g.nxh, g.nxh, 1);
}
template <int NX, int NY, int MAXE>
__device__ __forceinline__ void plane_ifft_x(double2* lds, const Geom& g) {
  if constexpr (NX > 0) FftCt<+1, NX, NY / 2, 1, NX, kThreads, true>::run(lds, g.ax.tw);
  else fft_axis<+1, MAXE>(lds, g.ax, g.ny / 2, 1, g.nx);
}

// LDS index of logical complex slot u (padded for the compile-time sizes)
template <int NX>
__device__ __forceinline__ int lp(int u) { return NX > 0 ? lds_pad(u) : u; }

// ---------------------------------------------------------------------------
// plane stages shared by kernels A, C and the fused C+A kernel
// ---------------------------------------------------------------------------
// The real plane lives in LDS as packed rows: row pair (2j, 2j+1) is the complex row
// j, value (x, y) at double index 2 lp(u) + (y & 1) with u = (y >> 1) nx + x.
template <int NX>
__device__ __forceinline__ int pk_at(int y, int x, int nx) {
  return 2 * lp<NX>((y >> 1) * nx + x) + (y & 1);
}

// packed real plane in LDS -> x pass (two real rows per complex transform), unpack,
// y pass -> half-spectrum plane (i, z)
template <int NX, int NY, int MAXE, class S>
__device__ __forceinline__ void forward_xy(double2* lds, const Geom& g, S* __restrict__ spec,
                                          int i, int z) {
  const int nx = NX ? NX : g.nx, ny = NY ? NY : g.ny;
  const int nxh = nx / 2 + 1;
  __syncthreads();
  plane_fft_x<NX, NY, MAXE>(lds, g);
  // unpack Z = A + iB into A_k = (Z_k + conj Z_{n-k})/2, B_k = (Z_k - conj Z_{n-k})/(2i)
  {
    constexpr int UPT = NX > 0 ? ((NY / 2) * (NX / 2 + 1) + kThreads - 1) / kThreads
                               : (MAXE + 1) / 2 + 1;
    double2 ra[UPT], rb[UPT];
    int dst[UPT];
    const int tu = (ny / 2) * nxh;
#pragma unroll
    for (int c = 0; c < UPT; ++c) {
      const int idx = threadIdx.x + c * kThreads;
      dst[c] = -1;
      if (idx < tu) {
        const int j = idx / nxh, k = idx - j * nxh;
        const double2 zk = lds[lp<NX>(j * nx + k)];
        const double2 zn = lds[lp<NX>(j * nx + (k == 0 ? 0 : nx - k))];
        ra[c] = make_double2(0.5 * (zk.x + zn.x), 0.5 * (zk.y - zn.y));
        rb[c] = make_double2(0.5 * (zk.y + zn.y), -0.5 * (zk.x - zn.x));
        dst[c] = (2 * j) * nxh + k;
      }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < UPT; ++c)
      if (dst[c] >= 0) {
        lds[lp<NX>(dst[c])] = ra[c];
        lds[lp<NX>(dst[c] + nxh)] = rb[c];
      }
    __syncthreads();
  }
  plane_fft_y<NX, NY, MAXE, -1>(lds, g);
  // half spectrum, 64-byte runs of 4 kx per (ky, tile, z)
  for (int e = threadIdx.x; e < ny * nxh; e += kThreads) {
    const int ky = e / nxh, kx = e - ky * nxh;
    spec_st(spec + spec_index(i, ky, kx, z, ny, g.nt, g.nz), lds[lp<NX>(e)]);
  }
}

// half-spectrum plane (i, z) -> y pass, Hermitian pack, x pass -> packed real plane
// in LDS (unscaled)
template <int NX, int NY, int MAXE, class S>
__device__ __forceinline__ void inverse_xy(double2* lds, const Geom& g,
                                          const S* __restrict__ spec, int i, int z) {
  const int nx = NX ? NX : g.nx, ny = NY ? NY : g.ny;
  const int nxh = nx / 2 + 1;
  const int tin = ny * nxh;
  for (int base = 0; base < tin; base += kChunk * kThreads) {
    double2 r[kChunk];
#pragma unroll
    for (int c = 0; c < kChunk; ++c) {
      const int e = base + c * kThreads + threadIdx.x;
      const int ky = e / nxh, kx = e - ky * nxh;
      r[c] = e < tin ? spec_ld(spec + spec_index(i, ky, kx, z, ny, g.nt, g.nz))
                     : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int c = 0; c < kChunk; ++c) {
      const int e = base + c * kThreads + threadIdx.x;
      if (e < tin) lds[lp<NX>(e)] = r[c];
    }
  }
  __syncthreads();
  plane_fft_y<NX, NY, MAXE, +1>(lds, g);
  // pack rows 2j, 2j+1 (Hermitian-extended along x; imaginary parts of kx = 0 and
  // kx = nx/2 dropped) into Z = A + iB
  {
    constexpr int UPT = NX > 0 ? ((NY / 2) * NX + kThreads - 1) / kThreads : MAXE + 1;
    double2 rz[UPT];
    int dst[UPT];
    const int tot = (ny / 2) * nx;
    const int half = nx / 2;
#pragma unroll
    for (int c = 0; c < UPT; ++c) {
      const int idx = threadIdx.x + c * kThreads;
      dst[c] = -1;
      if (idx < tot) {
        const int j = idx / nx, k = idx - j * nx;
        double2 A, B;
        if (k <= half) {
          A = lds[lp<NX>((2 * j) * nxh + k)];
          B = lds[lp<NX>((2 * j + 1) * nxh + k)];
          if (k == 0 || k == half) {
            A.y = 0.0;
            B.y = 0.0;
          }
        } else {
          A = lds[lp<NX>((2 * j) * nxh + nx - k)];
          B = lds[lp<NX>((2 * j + 1) * nxh + nx - k)];
          A.y = -A.y;
          B.y = -B.y;
        }
        rz[c] = make_double2(A.x - B.y, A.y + B.x);
        dst[c] = j * nx + k;
      }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < UPT; ++c)
      if (dst[c] >= 0) lds[lp<NX>(dst[c])] = rz[c];
    __syncthreads();
  }
  plane_ifft_x<NX, NY, MAXE>(lds, g);
}

// One point of the BDF right-hand side with a fixed rounding sequence (explicit fma, so
// that every kernel forming it, fused or not, produces the same bits).
__device__ __forceinline__ double bdf_rhs(double q1, double q2, double q3, double w, double a0,
                                          double a1, double a2, double b0, double b1, double b2,
                                          double ds) {
  const double t = __builtin_fma(a2, q3, __builtin_fma(a1, q2, a0 * q1));
  const double u = __builtin_fma(b2, q3, __builtin_fma(b1, q2, b0 * q1));
  return __builtin_fma(-(ds * w), u, t);
}

// BDF right-hand side of harmonic i on z-plane `plane` into the packed LDS plane:
//   P = a0 q1 + a1 q2 + a2 q3 - ds w (b0 q1 + b1 q2 + b2 q3)
// q1 is read from history slice 0, or (FROM_LDS) taken from the packed LDS plane times
// `scale` (the freshly inverted q(s), also stored to `out` when `store`).  have1 = 0
// makes q1 zero; rhs = 0 skips the right-hand side (store only).
template <int NX, int NY, int ORDER, bool FROM_LDS>
__device__ __forceinline__ void rhs_to_lds(double* pk, const Geom& g, const RhsArgs& a, int i,
                                           long long plane, bool have1, double scale,
                                           double* __restrict__ out, bool store, bool rhs) {
  const int nx = NX ? NX : g.nx, ny = NY ? NY : g.ny;
  const long long hs = g.hs;
  const int tot2 = (nx * ny) >> 1;  // nx is even: x pairs (2p, 2p+1) share a row
  const double ds = a.ds;
  const double a0 = a.a[0], a1 = a.a[1], a2 = a.a[2];
  const double b0 = a.b[0], b1 = a.b[1], b2 = a.b[2];
  const bool u1 = have1 && (FROM_LDS || i < a.hM[0]);
  const bool u2 = rhs && ORDER >= 2 && i < a.hM[1];
  const bool u3 = rhs && ORDER >= 3 && i < a.hM[2];
  // a zero history plane is read from the zero field: the loads depend only on `in`
  const double2* w2 = reinterpret_cast<const double2*>(a.w + plane);
  const double2* q1 =
      reinterpret_cast<const double2*>(u1 && !FROM_LDS ? a.h[0] + i * hs + plane : a.zero + plane);
  const double2* q2 = reinterpret_cast<const double2*>(u2 ? a.h[1] + i * hs + plane : a.zero + plane);
  const double2* q3 = reinterpret_cast<const double2*>(u3 ? a.h[2] + i * hs + plane : a.zero + plane);
  double2* o2 = reinterpret_cast<double2*>(out);
  const double2 zero = make_double2(0.0, 0.0);
  for (int base = 0; base < tot2; base += kPairs * kThreads) {
    double2 v1[kPairs], v2[kPairs], v3[kPairs], ww[kPairs];
#pragma unroll
    for (int c = 0; c < kPairs; ++c) {
      const int e = base + c * kThreads + threadIdx.x;
      const bool in = e < tot2;
      v1[c] = v2[c] = v3[c] = ww[c] = zero;
      if (in) {
        if (!FROM_LDS) v1[c] = q1[e];
        if (rhs) {
          if (ORDER >= 2) v2[c] = q2[e];
          if (ORDER >= 3) v3[c] = q3[e];
          ww[c] = w2[e];
        }
      }
    }
#pragma unroll
    for (int c = 0; c < kPairs; ++c) {
      const int e = base + c * kThreads + threadIdx.x;
      if (e < tot2) {
        const int y = (2 * e) / nx, x = 2 * e - y * nx;
        const int p0 = pk_at<NX>(y, x, nx), p1 = pk_at<NX>(y, x + 1, nx);
        if (FROM_LDS && u1) {
          v1[c] = make_double2(pk[p0] * scale, pk[p1] * scale);
          if (store) o2[e] = v1[c];
        }
        if (rhs) {
          pk[p0] = bdf_rhs(v1[c].x, v2[c].x, v3[c].x, ww[c].x, a0, a1, a2, b0, b1, b2, ds);
          pk[p1] = bdf_rhs(v1[c].y, v2[c].y, v3[c].y, ww[c].y, a0, a1, a2, b0, b1, b2, ds);
        }
      }
    }
  }
}

// A: per (z-plane, harmonic): BDF right-hand side from the history slices, forward
// x and y passes
template <int NX, int NY, int MAXE, int ORDER, class S>
__device__ __forceinline__ void k_rhs_fft_xy_body(const Geom& g, const RhsArgs& a) {
  extern __shared__ double2 lds[];
  int z, i;
  plane_of_block(g.nz, z, i);
  const long long plane = (long long)z * g.zs;
  rhs_to_lds<NX, NY, ORDER, false>(reinterpret_cast<double*>(lds), g, a, i, plane, true, 1.0,
                               nullptr, false, true);
  forward_xy<NX, NY, MAXE>(lds, g, reinterpret_cast<S*>(a.spec), i, z);
}

template <int NX, int NY, int MAXE, int ORDER, class S>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(NX > 0 ? 2 : 1))) k_rhs_fft_xy(Geom g, RhsArgs a) {
  k_rhs_fft_xy_body<NX, NY, MAXE, ORDER, S>(g, a);
}

// the same for one instance of a batch per blockIdx.z
template <int NX, int NY, int MAXE, int ORDER, class S>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(NX > 0 ? 2 : 1))) k_rhs_fft_xy_b(const PlaneJob* __restrict__ jobs) {
  const PlaneJob& j = jobs[blockIdx.z];
  if (j.skip && *j.skip) return;
  k_rhs_fft_xy_body<NX, NY, MAXE, ORDER, S>(j.g, j.a);
}

// C: per (z-plane, harmonic): inverse y and x (C2R) passes, 1/N, store of q(s)
template <int NX, int NY, int MAXE, class S>
__device__ __forceinline__ void k_ifft_xy_body(const Geom& g, const InvArgs& a) {
  extern __shared__ double2 lds[];
  int z, i;
  plane_of_block(g.nz, z, i);
  const int nx = NX ? NX : g.nx, ny = NY ? NY : g.ny;
  inverse_xy<NX, NY, MAXE>(lds, g, reinterpret_cast<const S*>(a.spec), i, z);
  double* out = a.out + i * g.hs + (long long)z * g.zs;
  const double* pk = reinterpret_cast<const double*>(lds);
  const double sc = a.scale;
  for (int e = threadIdx.x; e < nx * ny; e += kThreads) {
    const int y = e / nx, x = e - y * nx;
    out[e] = pk[pk_at<NX>(y, x, nx)] * sc;
  }
}

template <int NX, int NY, int MAXE, class S>
__global__ void __launch_bounds__(kThreads) k_ifft_xy(Geom g, InvArgs a) {
  k_ifft_xy_body<NX, NY, MAXE, S>(g, a);
}

// the same for one instance of a batch per blockIdx.z
template <int NX, int NY, int MAXE, class S>
__global__ void __launch_bounds__(kThreads) k_ifft_xy_b(const PlaneJob* __restrict__ jobs) {
  const PlaneJob& j = jobs[blockIdx.z];
  if (j.skip && *j.skip) return;
  k_ifft_xy_body<NX, NY, MAXE, S>(j.g, j.c);
}

// C of step s fused with A of the next step: per (z-plane, harmonic i <
// max(M(s), M(next))).  For i < M(s) the solved plane is inverted and q_i(s) stored;
// it stays in LDS as the newest history plane of the next right-hand side (zero for
// i >= M(s): a junction that grows the basis).  For i < M(next) the right-hand side
// is formed in place and transformed forward into the same spectrum plane.
template <int NX, int NY, int MAXE, int ORDER, class S>
__device__ __forceinline__ void k_inv_rhs_body(const Geom& g, const InvArgs& c, const RhsArgs& a) {
  extern __shared__ double2 lds[];
  int z, i;
  plane_of_block(g.nz, z, i);
  const long long plane = (long long)z * g.zs;
  const bool inv = i < c.M;
  const bool rhs = i < a.M;
  if (inv) inverse_xy<NX, NY, MAXE>(lds, g, reinterpret_cast<const S*>(c.spec), i, z);
  rhs_to_lds<NX, NY, ORDER, true>(reinterpret_cast<double*>(lds), g, a, i, plane, inv, c.scale,
                              c.out + i * g.hs + plane, inv, rhs);
  if (rhs) forward_xy<NX, NY, MAXE>(lds, g, reinterpret_cast<S*>(a.spec), i, z);
}

template <int NX, int NY, int MAXE, int ORDER, class S>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(NX > 0 ? 2 : 1))) k_inv_rhs(Geom g, InvArgs c, RhsArgs a) {
  k_inv_rhs_body<NX, NY, MAXE, ORDER, S>(g, c, a);
}

// the same for one instance of a batch per blockIdx.z
template <int NX, int NY, int MAXE, int ORDER, class S>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(NX > 0 ? 2 : 1))) k_inv_rhs_b(const PlaneJob* __restrict__ jobs) {
  const PlaneJob& j = jobs[blockIdx.z];
  if (j.skip && *j.skip) return;
  k_inv_rhs_body<NX, NY, MAXE, ORDER, S>(j.g, j.c, j.a);
}

namespace {
int pick_maxe(long long elems, int nthr) {
  const long long need = (elems + nthr - 1) / nthr;
  if (need <= 8) return 8;
  if (need <= 16) return 16;
  if (need <= 32) return 32;
  if (need <= 64) return 64;
  return -1;
}
}  // namespace

// B: the batched twin (jobs per blockIdx.z) instead of the single launch
template <int NX, int NY, int MAXE, class S, bool B>
static void* rhs_kernel(int order) {
  switch (order) {
    case 1: return B ? (void*)k_rhs_fft_xy_b<NX, NY, MAXE, 1, S> : (void*)k_rhs_fft_xy<NX, NY, MAXE, 1, S>;
    case 2: return B ? (void*)k_rhs_fft_xy_b<NX, NY, MAXE, 2, S> : (void*)k_rhs_fft_xy<NX, NY, MAXE, 2, S>;
    case 3: return B ? (void*)k_rhs_fft_xy_b<NX, NY, MAXE, 3, S> : (void*)k_rhs_fft_xy<NX, NY, MAXE, 3, S>;
  }
  return nullptr;
}

template <int NX, int NY, int MAXE, class S, bool B>
static void* inv_rhs_kernel(int order) {
  switch (order) {
    case 1: return B ? (void*)k_inv_rhs_b<NX, NY, MAXE, 1, S> : (void*)k_inv_rhs<NX, NY, MAXE, 1, S>;
    case 2: return B ? (void*)k_inv_rhs_b<NX, NY, MAXE, 2, S> : (void*)k_inv_rhs<NX, NY, MAXE, 2, S>;
    case 3: return B ? (void*)k_inv_rhs_b<NX, NY, MAXE, 3, S> : (void*)k_inv_rhs<NX, NY, MAXE, 3, S>;
  }
  return nullptr;
}

// 24 x 24, 32 x 32 and 48 x 48 planes (the batching workload's other grids) take
// compile-time Stockham passes (64 x 64 and 128 x 128 planes have their own kernels),
// everything else the runtime kernels
static bool specialised(const Geom& g) {
  return g.nx == g.ny && (g.nx == 24 || g.nx == 32 || g.nx == 48);
}

static size_t plane_lds_bytes(const Geom& g) {
  const size_t n = (size_t)g.ny * g.nxh;
  return sizeof(double2) * (specialised(g) ? n + n / 8 + 1 : n);
}

bool plane_supported(int nx, int ny) {
  if ((nx == 64 && ny == 64) || (nx == 128 && ny == 128)) return true;
  Geom g{};
  g.nx = nx;
  g.ny = ny;
  g.nxh = nx / 2 + 1;
  return plane_lds_bytes(g) <= kMaxLdsBytes && pick_maxe((long long)ny * g.nxh, kThreads) > 0;
}

}  // namespace wlc
#include "wlc_plane64.h"
#include "wlc_plane128.h"
namespace wlc {

template <int ORDER, class S, bool B>
static void* p64_kernel(int kind) {
  if (B) return kind == 0 ? (void*)p64::k_rhs64_b<ORDER, S> : (void*)p64::k_inv_rhs64_b<ORDER, S>;
  return kind == 0 ? (void*)p64::k_rhs64<ORDER, S> : (void*)p64::k_inv_rhs64<ORDER, S>;
}

template <int ORDER, class S, bool B>
static void* p128_kernel(int kind) {
  if (B) return kind == 0 ? (void*)p128::k_rhs128_b<ORDER, S> : (void*)p128::k_inv_rhs128_b<ORDER, S>;
  return kind == 0 ? (void*)p128::k_rhs128<ORDER, S> : (void*)p128::k_inv_rhs128<ORDER, S>;
}

// kind 0: A (order), 1: C, 2: fused C + A (order of the next step); *threads = workgroup
// size; S = storage type of the half spectrum; B: the batched twin
template <class S, bool B>
static void* pick_plane_kernel_t(const Geom& g, int kind, int order, size_t* lds, int* threads) {
  *lds = plane_lds_bytes(g);
  *threads = kThreads;
  if (g.nx == 64 && g.ny == 64) {  // static 32 KB plane
    *lds = 0;
    if (kind == 1) return B ? (void*)p64::k_inv64_b<S> : (void*)p64::k_inv64<S>;
    switch (order) {
      case 1: return p64_kernel<1, S, B>(kind);
      case 2: return p64_kernel<2, S, B>(kind);
      case 3: return p64_kernel<3, S, B>(kind);
    }
    return nullptr;
  }
  if (g.nx == 128 && g.ny == 128) {  // 128 KB plane in dynamic LDS, 512 threads
    *lds = sizeof(double2) * p128::kPlane;
    *threads = p128::kThreads;
    if (kind == 1) return B ? (void*)p128::k_inv128_b<S> : (void*)p128::k_inv128<S>;
    switch (order) {
      case 1: return p128_kernel<1, S, B>(kind);
      case 2: return p128_kernel<2, S, B>(kind);
      case 3: return p128_kernel<3, S, B>(kind);
    }
    return nullptr;
  }
#define WLC_PICK(NX, NY, MAXE)                                                            \
  (kind == 0 ? rhs_kernel<NX, NY, MAXE, S, B>(order)                                      \
   : kind == 1 ? (B ? (void*)k_ifft_xy_b<NX, NY, MAXE, S> : (void*)k_ifft_xy<NX, NY, MAXE, S>) \
               : inv_rhs_kernel<NX, NY, MAXE, S, B>(order))
  if (specialised(g)) {
    if (g.nx == 24) return WLC_PICK(24, 24, 0);
    if (g.nx == 48) return WLC_PICK(48, 48, 0);
    return WLC_PICK(32, 32, 0);
  }
  switch (pick_maxe((long long)g.ny * g.nxh, kThreads)) {
    case 8: return WLC_PICK(0, 0, 8);
    case 16: return WLC_PICK(0, 0, 16);
    case 32: return WLC_PICK(0, 0, 32);
    case 64: return WLC_PICK(0, 0, 64);
  }
#undef WLC_PICK
  return nullptr;
}

static void* pick_plane_kernel(const Geom& g, int kind, int order, size_t* lds, int* threads,
                               bool batch = false) {
  if (batch)
    return g.spec32 ? pick_plane_kernel_t<float2, true>(g, kind, order, lds, threads)
                    : pick_plane_kernel_t<double2, true>(g, kind, order, lds, threads);
  return g.spec32 ? pick_plane_kernel_t<float2, false>(g, kind, order, lds, threads)
                  : pick_plane_kernel_t<double2, false>(g, kind, order, lds, threads);
}

static hipError_t launch_plane(const Geom& g, int kind, int order, int planes, void** args,
                               hipStream_t st) {
  size_t lds = 0;
  int threads = kThreads;
  void* k = pick_plane_kernel(g, kind, order, &lds, &threads);
  if (!k) return hipErrorInvalidValue;
  if (lds) {
    hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  return hipLaunchKernel(k, dim3(g.nz * planes), dim3(threads), args, lds, st);
}

hipError_t launch_rhs_fft(const Geom& g, const RhsArgs& a, int order, hipStream_t st) {
  Geom gg = g;
  RhsArgs aa = a;
  void* args[] = {&gg, &aa};
  return launch_plane(g, 0, order, a.M, args, st);
}

hipError_t launch_ifft_store(const Geom& g, const InvArgs& a, hipStream_t st) {
  Geom gg = g;
  InvArgs aa = a;
  void* args[] = {&gg, &aa};
  return launch_plane(g, 1, 0, a.M, args, st);
}

hipError_t launch_inv_rhs(const Geom& g, const InvArgs& c, const RhsArgs& a, int order,
                          hipStream_t st) {
  Geom gg = g;
  InvArgs cc = c;
  RhsArgs aa = a;
  void* args[] = {&gg, &cc, &aa};
  return launch_plane(g, 2, order, c.M > a.M ? c.M : a.M, args, st);
}

LaunchKey plane_key(const Geom& g, int kind, int order, int planes) {
  size_t lds = 0;
  int threads = kThreads;
  const void* k = pick_plane_kernel(g, kind, order, &lds, &threads);
  return LaunchKey{k, (unsigned)(g.nz * planes), 1u, (unsigned)threads, lds};
}

static hipError_t launch_plane_batch(const Geom& g, int kind, int order, int planes,
                                     const PlaneJob* jobs, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  size_t lds = 0;
  int threads = kThreads;
  void* k = pick_plane_kernel(g, kind, order, &lds, &threads, true);
  if (!k) return hipErrorInvalidValue;
  if (lds) {
    hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  const PlaneJob* jp = jobs;
  void* args[] = {&jp};
  return hipLaunchKernel(k, dim3(g.nz * planes, 1, n), dim3(threads), args, lds, st);
}

hipError_t launch_rhs_fft_batch(const PlaneJob& rep, int order, const PlaneJob* jobs, int n,
                                hipStream_t st) {
  return launch_plane_batch(rep.g, 0, order, rep.a.M, jobs, n, st);
}

hipError_t launch_ifft_store_batch(const PlaneJob& rep, const PlaneJob* jobs, int n, hipStream_t st) {
  return launch_plane_batch(rep.g, 1, 0, rep.c.M, jobs, n, st);
}

hipError_t launch_inv_rhs_batch(const PlaneJob& rep, int order, const PlaneJob* jobs, int n,
                                hipStream_t st) {
  return launch_plane_batch(rep.g, 2, order, rep.c.M > rep.a.M ? rep.c.M : rep.a.M, jobs, n, st);
}

// ---------------------------------------------------------------------------
// elementwise / reduction kernels
// ---------------------------------------------------------------------------
// real-space kernels visit the valid points of the padded layout: grid (blocks over
// the nx*ny plane, nz * planes-per-field)
__global__ void k_set_delta(Geom g, double* s, int M) {
  const int nxy = g.nx * g.ny;
  const int z = blockIdx.y % g.nz, i = blockIdx.y / g.nz;
  double* p = s + i * g.hs + (long long)z * g.zs;
  const double v = i == 0 ? 1.0 : 0.0;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nxy; e += gridDim.x * blockDim.x) p[e] = v;
}

hipError_t launch_set_delta(const Geom& g, double* slice, int M, hipStream_t st) {
  const int nxy = g.nx * g.ny;
  dim3 grid((nxy + 255) / 256, g.nz * M);
  hipLaunchKernelGGL(k_set_delta, grid, dim3(256), 0, st, g, slice, M);
  return hipGetLastError();
}

__global__ void k_fill(double* d, double v, long long n) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x)
    d[e] = v;
}

__global__ void k_axpy(double* y, double a, const double* x, long long n) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x)
    y[e] += a * x[e];
}

hipError_t launch_axpy(double* y, double a, const double* x, long long n, hipStream_t st) {
  int blocks = (int)std::min<long long>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_axpy, dim3(blocks), dim3(256), 0, st, y, a, x, n);
  return hipGetLastError();
}

hipError_t launch_fill(double* dst, double v, long long n, hipStream_t st) {
  int blocks = (int)std::min<long long>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(256), 0, st, dst, v, n);
  return hipGetLastError();
}

// rho_t += w_t * sum_{i < lim_t} q_i qs_i at every grid point, for up to two blocks.
// One thread per x pair (16-byte loads); the harmonic loop is unrolled so that many
// independent loads are in flight, the sum itself runs in harmonic order.
__global__ void __launch_bounds__(256) k_density_acc(Geom g, const double* __restrict__ q,
                                                     const double* __restrict__ qs, double* rho0,
                                                     int lim0, double w0, double* rho1, int lim1,
                                                     double w1) {
  const int nxy2 = (g.nx * g.ny) >> 1;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nxy2) return;
  const long long r = (long long)blockIdx.y * g.zs + 2 * e;
  const long long hs = g.hs;
  double2 s = make_double2(0.0, 0.0), s0 = s, s1 = s;
  const int top = lim0 > lim1 ? lim0 : lim1;
#pragma unroll 8
  for (int i = 0; i < top; ++i) {
    const double2 a = *reinterpret_cast<const double2*>(q + i * hs + r);
    const double2 b = *reinterpret_cast<const double2*>(qs + i * hs + r);
    s.x = __builtin_fma(a.x, b.x, s.x);
    s.y = __builtin_fma(a.y, b.y, s.y);
    if (i + 1 == lim0) s0 = s;
    if (i + 1 == lim1) s1 = s;
  }
  if (rho0) {
    double2* p = reinterpret_cast<double2*>(rho0 + r);
    const double2 o = *p;
    *p = make_double2(__builtin_fma(w0, s0.x, o.x), __builtin_fma(w0, s0.y, o.y));
  }
  if (rho1) {
    double2* p = reinterpret_cast<double2*>(rho1 + r);
    const double2 o = *p;
    *p = make_double2(__builtin_fma(w1, s1.x, o.x), __builtin_fma(w1, s1.y, o.y));
  }
}

hipError_t launch_density_acc(const Geom& g, const double* q, const double* qs, double* rho0,
                              int lim0, double w0, double* rho1, int lim1, double w1,
                              hipStream_t st) {
  if (!rho1) lim1 = 0;
  const int nxy2 = (g.nx * g.ny) / 2;  // nx is even
  dim3 grid((nxy2 + 255) / 256, g.nz);
  hipLaunchKernelGGL(k_density_acc, grid, dim3(256), 0, st, g, q, qs, rho0, lim0, w0, rho1,
                     lim1, w1);
  return hipGetLastError();
}

// periodic-Simpson weights c(j) = 2/3 (even j), 4/3 (odd j) per axis
__device__ __forceinline__ double simp_w(int j) { return (j & 1) ? (4.0 / 3.0) : (2.0 / 3.0); }

__global__ void k_weighted_partial(Geom g, const double* __restrict__ src, int nf, double dV,
                                   double* partial) {
  __shared__ double red[256];
  const long long N = (long long)g.nx * g.ny * g.nz;
  double acc = 0.0;
  for (long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x; r < N;
       r += (long long)gridDim.x * blockDim.x) {
    const int x = (int)(r % g.nx);
    const long long t = r / g.nx;
    const int y = (int)(t % g.ny), z = (int)(t / g.ny);
    const long long pr = (long long)z * g.zs + (long long)y * g.nx + x;
    double f = 0.0;
    for (int k = 0; k < nf; ++k) f += src[k * g.hs + pr];
    acc += simp_w(x) * simp_w(y) * simp_w(z) * f;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0] * dV;
}

__global__ void k_final_sum(const double* partial, int n, double* result) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partial[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *result = red[0];
}

hipError_t launch_weighted_sum(const Geom& g, const double* src, int nf, double dV,
                               double* partial, int nparts, double* result, hipStream_t st) {
  hipLaunchKernelGGL(k_weighted_partial, dim3(nparts), dim3(256), 0, st, g, src, nf, dV, partial);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(256), 0, st, partial, nparts, result);
  return hipGetLastError();
}

__global__ void k_phi(Geom g, const double* __restrict__ rho, double V,
                      const double* __restrict__ totden, double* phi) {
  const int nxy = g.nx * g.ny;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nxy) return;
  const long long r = (long long)(blockIdx.y % g.nz) * g.zs + (blockIdx.y / g.nz) * g.hs + e;
  phi[r] = rho[r] * (V / *totden);
}

hipError_t launch_phi(const Geom& g, const double* rho, int nblk, double V, const double* totden,
                      double* phi, hipStream_t st) {
  const int nxy = g.nx * g.ny;
  dim3 grid((nxy + 255) / 256, g.nz * nblk);
  hipLaunchKernelGGL(k_phi, grid, dim3(256), 0, st, g, rho, V, totden, phi);
  return hipGetLastError();
}

}  // namespace wlc
