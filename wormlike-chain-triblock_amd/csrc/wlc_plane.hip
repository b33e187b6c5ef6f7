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
}  // namespace

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2).
// Remap the linear block id so that every XCD walks a contiguous range of
// (harmonic, z-plane) pairs, z fastest: the z-fastest half spectrum then has the
// 16-byte pieces of each cache line written/read by workgroups of one XCD at
// about the same time.
__device__ __forceinline__ void plane_of_block(int nz, int& z, int& i) {
  const int nb = gridDim.x;
  const int b = blockIdx.x;
  const int q = nb >> 3;
  const int logical = (b < (q << 3)) ? ((b & 7) * q + (b >> 3)) : b;
  i = logical / nz;
  z = logical - i * nz;
}

template <int MAXE, int ORDER>
__global__ void __launch_bounds__(kThreads) k_rhs_fft_xy(Geom g, RhsArgs a) {
  extern __shared__ double2 lds[];
  int z, i;
  plane_of_block(g.nz, z, i);
  const int nx = g.nx, ny = g.ny, nxh = g.nxh;
  const long long N = (long long)nx * ny * g.nz;
  const long long plane = (long long)z * nx * ny;
  const bool u1 = i < a.hM[0];
  const bool u2 = ORDER >= 2 && i < a.hM[1];
  const bool u3 = ORDER >= 3 && i < a.hM[2];
  const double* q1 = u1 ? a.h[0] + i * N + plane : nullptr;
  const double* q2 = u2 ? a.h[1] + i * N + plane : nullptr;
  const double* q3 = u3 ? a.h[2] + i * N + plane : nullptr;
  const double* w = a.w + plane;
  double* pk = reinterpret_cast<double*>(lds);  // packed rows: (row 2j, row 2j+1) -> complex
  const double ds = a.ds;
  const double a0 = a.a[0], a1 = a.a[1], a2 = a.a[2];
  const double b0 = a.b[0], b1 = a.b[1], b2 = a.b[2];
  const int tot = nx * ny;
  for (int base = 0; base < tot; base += kChunk * kThreads) {
    double v1[kChunk], v2[kChunk], v3[kChunk], ww[kChunk];
#pragma unroll
    for (int c = 0; c < kChunk; ++c) {
      const int e = base + c * kThreads + threadIdx.x;
      const bool in = e < tot;
      v1[c] = (u1 && in) ? q1[e] : 0.0;
      v2[c] = (u2 && in) ? q2[e] : 0.0;
      v3[c] = (u3 && in) ? q3[e] : 0.0;
      ww[c] = in ? w[e] : 0.0;
    }
#pragma unroll
    for (int c = 0; c < kChunk; ++c) {
      const int e = base + c * kThreads + threadIdx.x;
      if (e < tot) {
        const double t = a0 * v1[c] + a1 * v2[c] + a2 * v3[c];
        const double u = b0 * v1[c] + b1 * v2[c] + b2 * v3[c];
        const int y = e / nx, x = e - y * nx;
        pk[2 * ((y >> 1) * nx + x) + (y & 1)] = t - (ds * ww[c]) * u;
      }
    }
  }
  __syncthreads();
  fft_axis<-1, MAXE>(lds, g.ax, ny / 2, 1, nx);
  // unpack Z = A + iB into A_k = (Z_k + conj Z_{n-k})/2, B_k = (Z_k - conj Z_{n-k})/(2i)
  {
    constexpr int UPT = (MAXE + 1) / 2 + 1;
    double2 ra[UPT], rb[UPT];
    int dst[UPT];
    const int tu = (ny / 2) * nxh;
#pragma unroll
    for (int c = 0; c < UPT; ++c) {
      const int idx = threadIdx.x + c * kThreads;
      dst[c] = -1;
      if (idx < tu) {
        const int j = idx / nxh, k = idx - j * nxh;
        const double2 zk = lds[j * nx + k];
        const double2 zn = lds[j * nx + (k == 0 ? 0 : nx - k)];
        ra[c] = make_double2(0.5 * (zk.x + zn.x), 0.5 * (zk.y - zn.y));
        rb[c] = make_double2(0.5 * (zk.y + zn.y), -0.5 * (zk.x - zn.x));
        dst[c] = (2 * j) * nxh + k;
      }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < UPT; ++c)
      if (dst[c] >= 0) {
        lds[dst[c]] = ra[c];
        lds[dst[c] + nxh] = rb[c];
      }
    __syncthreads();
  }
  fft_axis<-1, MAXE>(lds, g.ay, nxh, nxh, 1);
  // half spectrum [i][ky][kx][z]
  const int nz = g.nz;
  double2* out = a.spec + (long long)i * ny * nxh * nz + z;
  for (int e = threadIdx.x; e < ny * nxh; e += kThreads) out[(long long)e * nz] = lds[e];
}

template <int MAXE>
__global__ void __launch_bounds__(kThreads) k_ifft_xy(Geom g, InvArgs a) {
  extern __shared__ double2 lds[];
  int z, i;
  plane_of_block(g.nz, z, i);
  const int nx = g.nx, ny = g.ny, nxh = g.nxh;
  const int nz = g.nz;
  const double2* in = a.spec + (long long)i * ny * nxh * nz + z;
  const int tin = ny * nxh;
  for (int base = 0; base < tin; base += kChunk * kThreads) {
    double2 r[kChunk];
#pragma unroll
    for (int c = 0; c < kChunk; ++c) {
      const int e = base + c * kThreads + threadIdx.x;
      r[c] = e < tin ? in[(long long)e * nz] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int c = 0; c < kChunk; ++c) {
      const int e = base + c * kThreads + threadIdx.x;
      if (e < tin) lds[e] = r[c];
    }
  }
  __syncthreads();
  fft_axis<+1, MAXE>(lds, g.ay, nxh, nxh, 1);
  // pack rows 2j, 2j+1 (Hermitian-extended along x; imaginary parts of kx = 0 and
  // kx = nx/2 dropped) into Z = A + iB
  {
    constexpr int UPT = MAXE + 1;
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
          A = lds[(2 * j) * nxh + k];
          B = lds[(2 * j + 1) * nxh + k];
          if (k == 0 || k == half) {
            A.y = 0.0;
            B.y = 0.0;
          }
        } else {
          A = lds[(2 * j) * nxh + nx - k];
          B = lds[(2 * j + 1) * nxh + nx - k];
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
      if (dst[c] >= 0) lds[dst[c]] = rz[c];
    __syncthreads();
  }
  fft_axis<+1, MAXE>(lds, g.ax, ny / 2, 1, nx);
  const long long N = (long long)nx * ny * g.nz;
  double* out = a.out + i * N + (long long)z * nx * ny;
  const double* pk = reinterpret_cast<const double*>(lds);
  const double sc = a.scale;
  for (int e = threadIdx.x; e < nx * ny; e += kThreads) {
    const int y = e / nx, x = e - y * nx;
    out[e] = pk[2 * ((y >> 1) * nx + x) + (y & 1)] * sc;
  }
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

template <int MAXE>
static void* rhs_kernel(int order) {
  switch (order) {
    case 1: return (void*)k_rhs_fft_xy<MAXE, 1>;
    case 2: return (void*)k_rhs_fft_xy<MAXE, 2>;
    case 3: return (void*)k_rhs_fft_xy<MAXE, 3>;
  }
  return nullptr;
}

hipError_t launch_rhs_fft(const Geom& g, const RhsArgs& a, int order, hipStream_t st) {
  const size_t lds = sizeof(double2) * (size_t)g.ny * g.nxh;
  const int me = pick_maxe((long long)g.ny * g.nxh, kThreads);
  void* k = nullptr;
  switch (me) {
    case 8: k = rhs_kernel<8>(order); break;
    case 16: k = rhs_kernel<16>(order); break;
    case 32: k = rhs_kernel<32>(order); break;
    case 64: k = rhs_kernel<64>(order); break;
    default: return hipErrorInvalidValue;
  }
  if (!k) return hipErrorInvalidValue;
  hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  Geom gg = g;
  RhsArgs aa = a;
  void* args[] = {&gg, &aa};
  return hipLaunchKernel(k, dim3(g.nz * a.M), dim3(kThreads), args, lds, st);
}

hipError_t launch_ifft_store(const Geom& g, const InvArgs& a, hipStream_t st) {
  const size_t lds = sizeof(double2) * (size_t)g.ny * g.nxh;
  const int me = pick_maxe((long long)g.ny * g.nxh, kThreads);
  void* k = nullptr;
  switch (me) {
    case 8: k = (void*)k_ifft_xy<8>; break;
    case 16: k = (void*)k_ifft_xy<16>; break;
    case 32: k = (void*)k_ifft_xy<32>; break;
    case 64: k = (void*)k_ifft_xy<64>; break;
    default: return hipErrorInvalidValue;
  }
  hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  Geom gg = g;
  InvArgs aa = a;
  void* args[] = {&gg, &aa};
  return hipLaunchKernel(k, dim3(g.nz * a.M), dim3(kThreads), args, lds, st);
}

// ---------------------------------------------------------------------------
// elementwise / reduction kernels
// ---------------------------------------------------------------------------
__global__ void k_set_delta(double* s, int M, long long N) {
  const long long tot = (long long)M * N;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot;
       e += (long long)gridDim.x * blockDim.x)
    s[e] = (e < N) ? 1.0 : 0.0;
}

hipError_t launch_set_delta(double* slice, int M, long long N, hipStream_t st) {
  const long long tot = (long long)M * N;
  int blocks = (int)std::min<long long>((tot + 255) / 256, 65536);
  hipLaunchKernelGGL(k_set_delta, dim3(blocks), dim3(256), 0, st, slice, M, N);
  return hipGetLastError();
}

__global__ void k_fill(double* d, double v, long long n) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x)
    d[e] = v;
}

hipError_t launch_fill(double* dst, double v, long long n, hipStream_t st) {
  int blocks = (int)std::min<long long>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(256), 0, st, dst, v, n);
  return hipGetLastError();
}

__global__ void k_density_acc(const double* __restrict__ q, const double* __restrict__ qs,
                              long long N, double* rho0, int lim0, double w0, double* rho1,
                              int lim1, double w1) {
  for (long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x; r < N;
       r += (long long)gridDim.x * blockDim.x) {
    double s = 0.0, s0 = 0.0, s1 = 0.0;
    const int top = lim0 > lim1 ? lim0 : lim1;
    for (int i = 0; i < top; ++i) {
      s += q[i * N + r] * qs[i * N + r];
      if (i + 1 == lim0) s0 = s;
      if (i + 1 == lim1) s1 = s;
    }
    if (rho0) rho0[r] += w0 * s0;
    if (rho1) rho1[r] += w1 * s1;
  }
}

hipError_t launch_density_acc(const double* q, const double* qs, long long N, double* rho0,
                              int lim0, double w0, double* rho1, int lim1, double w1,
                              hipStream_t st) {
  if (!rho1) lim1 = 0;
  int blocks = (int)std::min<long long>((N + 255) / 256, 16384);
  hipLaunchKernelGGL(k_density_acc, dim3(blocks), dim3(256), 0, st, q, qs, N, rho0, lim0, w0,
                     rho1, lim1, w1);
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
    double f = 0.0;
    for (int k = 0; k < nf; ++k) f += src[k * N + r];
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

__global__ void k_phi(const double* __restrict__ rho, long long tot, double V,
                      const double* __restrict__ totden, double* phi) {
  const double f = V / *totden;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot;
       e += (long long)gridDim.x * blockDim.x)
    phi[e] = rho[e] * f;
}

hipError_t launch_phi(const double* rho, int nblk, long long N, double V, const double* totden,
                      double* phi, hipStream_t st) {
  const long long tot = (long long)nblk * N;
  int blocks = (int)std::min<long long>((tot + 255) / 256, 65536);
  hipLaunchKernelGGL(k_phi, dim3(blocks), dim3(256), 0, st, rho, tot, V, totden, phi);
  return hipGetLastError();
}

}  // namespace wlc
