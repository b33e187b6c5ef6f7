// Warm-start network for the SCFT fields: a 3D convolutional U-Net trained and run with
// hand-written kernels.  The reference has no counterpart (it starts every SCFT from the
// hand-built fields of init_w, init_array_2dcomp_mod.f90:843-1065); this network proposes
// initial fields from a seed field and the instance's parameters, and plain SCFT from them
// still decides the converged state.
//
// Layout: activations channels-last [B][D][H][W][C] with C a multiple of 16, stored as
// bfloat16 where a convolution reads them; pre-activations, gradients and master weights
// FP32.  Convolutions are 3x3x3, zero padding 1, stride 1, weights [Cout][27][Cin] (tap
// t = 9 kz + 3 ky + kx).  Every convolution (forward, gradient with respect to the input,
// gradient with respect to the weights) is an implicit GEMM on the bf16 matrix cores
// (v_mfma_f32_16x16x16_bf16, FP32 accumulation):
//   forward   M = voxels, N = Cout, K = 27 Cin
//   input     the forward kernel on the output gradient with the flipped, transposed kernel
//             W'[ci][26 - t][co] = W[co][t][ci]
//   weights   M = Cout, N = Cin per tap, K = voxels, split over voxel chunks whose partial
//             sums are added in a fixed order (deterministic)
// Architecture (levels n, n/2, n/4, n/8; base 32 channels):
//   e0: conv 16->32, conv 32->32 (skip s0) | pool | e1: 32->64, 64->64 (s1) | pool |
//   e2: 64->128, 128->128 (s2) | pool | bottleneck 128->128 | up, concat s2 -> conv 256->64 |
//   up, concat s1 -> conv 128->32 | up, concat s0 -> conv 64->32 | conv 32->16 (output)
// ReLU after every convolution but the last; 2x2x2 average pooling; nearest 2x upsampling.
// Input channels: 3 seed fields and 8 broadcast scalars (11, padded to 16); output: the
// first 3 of 16 channels (the others have zero weights and no loss).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/wlc_mde.h"
#include "wlc_host.h"

namespace wlc {
namespace nn {

typedef unsigned short bf16;
typedef short bf16x4 __attribute__((__vector_size__(4 * sizeof(short))));
typedef float f32x4 __attribute__((__vector_size__(4 * sizeof(float))));
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// round to nearest even (finite inputs)
__host__ __device__ inline bf16 f2bf(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}
__host__ __device__ inline float bf2f(bf16 h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

struct Vol {
  int B, D, H, W;
  __host__ __device__ long long voxels() const { return (long long)B * D * H * W; }
};

// neighbour of voxel v at tap t (zero padding: -1 outside)
__device__ __forceinline__ long long neighbour(const Vol& g, long long v, int t) {
  const int x = (int)(v % g.W);
  long long r = v / g.W;
  const int y = (int)(r % g.H);
  r /= g.H;
  const int z = (int)(r % g.D);
  const long long b = r / g.D;
  const int zz = z + t / 9 - 1, yy = y + (t / 3) % 3 - 1, xx = x + t % 3 - 1;
  if (zz < 0 || zz >= g.D || yy < 0 || yy >= g.H || xx < 0 || xx >= g.W) return -1;
  return ((b * g.D + zz) * g.H + yy) * g.W + xx;
}

// Y[v][co] = bias[co] + sum_{t, ci} W[co][t][ci] X[nbr(v, t)][ci] (zero outside the grid).
// Workgroup = 64 consecutive voxels x 16 NT output channels, wave w its voxels 16 w .. 16 w + 15.
// Matrix-core step (16 voxels x 16 channels x KI input channels): lane l holds the input
// channels KI/4 (l / 16) .. of voxel l % 16 (A) and of output channel l % 16 (B), contiguous
// in both tensors, so each operand is one 8- or 16-byte load; it receives D[4 (l / 16) + r]
// [l % 16].  KI = 32 (v_mfma_f32_16x16x32_bf16) when Cin is a multiple of 32, else 16.  The
// voxel's coordinates are decoded once; a tap only adds an offset and checks the bounds.
template <int NT, int KI>
__global__ void __launch_bounds__(256) k_conv_fwd(const bf16* __restrict__ X, const bf16* __restrict__ Wt,
                                                  const float* __restrict__ bias, float* __restrict__ Y,
                                                  Vol g, int Cin, int Cout) {
  constexpr int CPL = KI / 4;  // input channels per lane and step
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int row = lane & 15, kq = lane >> 4;
  const int V = (int)g.voxels();
  const int vb = blockIdx.x * 64 + 16 * wv;
  const int v = vb + row;
  const int co0 = blockIdx.y * 16 * NT;
  const bool vin = v < V;
  int x = 0, y = 0, z = 0;
  if (vin) {
    x = v % g.W;
    const int r = v / g.W;
    y = r % g.H;
    z = (r / g.H) % g.D;
  }
  f32x4 acc[NT];
#pragma unroll
  for (int j = 0; j < NT; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const size_t wstride = (size_t)16 * 27 * Cin;  // between the NT output-channel tiles
  for (int t = 0; t < 27; ++t) {
    const int dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
    const bool ok = vin && (unsigned)(z + dz) < (unsigned)g.D && (unsigned)(y + dy) < (unsigned)g.H &&
                    (unsigned)(x + dx) < (unsigned)g.W;
    const int nb = ok ? v + (dz * g.H + dy) * g.W + dx : 0;
    const bf16* xr = X + (size_t)nb * Cin + CPL * kq;
    const bf16* wr = Wt + ((size_t)(co0 + row) * 27 + t) * Cin + CPL * kq;
    for (int c0 = 0; c0 < Cin; c0 += KI) {
      if constexpr (KI == 32) {
        const s16x8 a = ok ? *reinterpret_cast<const s16x8*>(xr + c0) : s16x8{0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < NT; ++j) {
          const s16x8 b = *reinterpret_cast<const s16x8*>(wr + j * wstride + c0);
          acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                           __builtin_bit_cast(bf16x8, b), acc[j], 0, 0, 0);
        }
      } else {
        const bf16x4 a = ok ? *reinterpret_cast<const bf16x4*>(xr + c0) : bf16x4{0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < NT; ++j) {
          const bf16x4 b = *reinterpret_cast<const bf16x4*>(wr + j * wstride + c0);
          acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[j], 0, 0, 0);
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    const int co = co0 + 16 * j + row;
    const float bv = bias ? bias[co] : 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int vv = vb + 4 * kq + r;
      if (vv < V) Y[(size_t)vv * Cout + co] = acc[j][r] + bv;
    }
  }
}

// The same convolution (Cin a multiple of 32) with the weights staged through LDS and the
// next step's operands in flight: per K step (tap t, 32 input channels) the workgroup's
// 16 NT x 32 weight block (NT KB) is loaded once by 64 NT threads into one of two LDS buffers
// instead of once per wave, every wave reads its B fragments from there (16 lanes of equal
// kq read 256 contiguous bytes), and the global loads of step k + 1 (weights and the gathered
// input run) are issued before the matrix-core steps of step k.  One barrier per step: a
// thread reaching barrier k has finished its reads of step k - 1, so the buffer written next
// is free.
template <int NT>
__global__ void __launch_bounds__(256) k_conv_fwd_l(const bf16* __restrict__ X, const bf16* __restrict__ Wt,
                                                    const float* __restrict__ bias, float* __restrict__ Y,
                                                    Vol g, int Cin, int Cout) {
  __shared__ s16x8 wbuf[2][4 * 16 * NT];  // [kq][output channel] 16-byte pieces
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int row = lane & 15, kq = lane >> 4;
  const int V = (int)g.voxels();
  const int vb = blockIdx.x * 64 + 16 * wv;
  const int v = vb + row;
  const int co0 = blockIdx.y * 16 * NT;
  const bool vin = v < V;
  int x = 0, y = 0, z = 0;
  if (vin) {
    x = v % g.W;
    const int r = v / g.W;
    y = r % g.H;
    z = (r / g.H) % g.D;
  }
  const int nch = Cin >> 5;
  const int K = 27 * nch;
  // staging role: thread st < 64 NT moves the 8 channels 8 (st & 3) .. of output channel st >> 2
  const int st = threadIdx.x;
  const bool stager = st < 64 * NT;
  const int sco = st >> 2, skq = st & 3;
  const bf16* wrow = Wt + (size_t)(co0 + (stager ? sco : 0)) * 27 * Cin + 8 * skq;
  const s16x8 zero = {0, 0, 0, 0, 0, 0, 0, 0};
  auto load_w = [&](int ks) -> s16x8 {
    const int t = ks / nch, c0 = (ks - t * nch) << 5;
    return stager ? *reinterpret_cast<const s16x8*>(wrow + (size_t)t * Cin + c0) : zero;
  };
  auto load_a = [&](int ks) -> s16x8 {
    const int t = ks / nch, c0 = (ks - t * nch) << 5;
    const int dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
    const bool ok = vin && (unsigned)(z + dz) < (unsigned)g.D && (unsigned)(y + dy) < (unsigned)g.H &&
                    (unsigned)(x + dx) < (unsigned)g.W;
    const int nb = ok ? v + (dz * g.H + dy) * g.W + dx : 0;
    return ok ? *reinterpret_cast<const s16x8*>(X + (size_t)nb * Cin + c0 + 8 * kq) : zero;
  };
  f32x4 acc[NT];
#pragma unroll
  for (int j = 0; j < NT; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
  s16x8 wn = load_w(0), an = load_a(0);
  for (int ks = 0; ks < K; ++ks) {
    const s16x8 wc = wn, ac = an;
    if (stager) wbuf[ks & 1][skq * (16 * NT) + sco] = wc;
    if (ks + 1 < K) {
      wn = load_w(ks + 1);
      an = load_a(ks + 1);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const s16x8 b = wbuf[ks & 1][kq * (16 * NT) + 16 * j + row];
      acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, ac),
                                                       __builtin_bit_cast(bf16x8, b), acc[j], 0, 0, 0);
    }
  }
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    const int co = co0 + 16 * j + row;
    const float bv = bias ? bias[co] : 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int vv = vb + 4 * kq + r;
      if (vv < V) Y[(size_t)vv * Cout + co] = acc[j][r] + bv;
    }
  }
}

// out[c][v] = in[v][c] (bf16), 32 x 32 tiles through LDS
__global__ void __launch_bounds__(256) k_transpose(const bf16* __restrict__ in, bf16* __restrict__ out,
                                                   int V, int C) {
  __shared__ bf16 tile[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int v0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int v = v0 + ty + 8 * r, c = c0 + tx;
    if (v < V && c < C) tile[ty + 8 * r][tx] = in[(size_t)v * C + c];
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int c = c0 + ty + 8 * r, v = v0 + tx;
    if (v < V && c < C) out[(size_t)c * V + v] = tile[tx][ty + 8 * r];
  }
}

// part[chunk][co][t][ci] = sum_{v in chunk} dY[v][co] X[nbr(v, t)][ci] from the transposed
// operands dYT [Cout][V] and XT [Cin][V]: the voxels are the reduction index, so lane l's
// eight consecutive voxels 8 (l / 16) .. of a 32-voxel step are one 16-byte load of a
// channel row.  Needs W % 8 == 0 (eight consecutive voxels lie in one x row: their tap
// neighbours are the same run shifted by dx, read as the aligned run and its left or right
// neighbour run) and V, chunk multiples of 32.  One wave per (tap, 16 MC output channels,
// 16 NC input channels, voxel chunk).
template <int MC, int NC>
__global__ void __launch_bounds__(64) k_conv_dw_t(const bf16* __restrict__ dYT, const bf16* __restrict__ XT,
                                                  float* __restrict__ part, Vol g, int Cin, int Cout, int chunk) {
  const int lane = threadIdx.x, row = lane & 15, kq = lane >> 4;
  const int nci = Cin / (16 * NC);
  int idx = blockIdx.x;
  const int t = idx % 27;
  idx /= 27;
  const int ci0 = (idx % nci) * 16 * NC, co0 = (idx / nci) * 16 * MC;
  const int dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
  const int V = (int)g.voxels();
  const int vb0 = blockIdx.y * chunk;
  const int ve = vb0 + chunk < V ? vb0 + chunk : V;
  const int rowshift = (dz * g.H + dy) * g.W;
  const s16x8 zero = {0, 0, 0, 0, 0, 0, 0, 0};
  f32x4 acc[MC][NC];
#pragma unroll
  for (int i = 0; i < MC; ++i)
#pragma unroll
    for (int j = 0; j < NC; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int vb = vb0; vb < ve; vb += 32) {
    const int v = vb + 8 * kq;  // this lane's eight voxels v .. v + 7: one x-row run
    const int x = v % g.W;
    const int r1 = v / g.W;
    const int y = r1 % g.H, z = (r1 / g.H) % g.D;
    const bool rowok = (unsigned)(z + dz) < (unsigned)g.D && (unsigned)(y + dy) < (unsigned)g.H;
    const int base = v + rowshift;  // neighbour run at dx = 0 (a multiple of 8)
    s16x8 a[MC], b[NC];
#pragma unroll
    for (int i = 0; i < MC; ++i) a[i] = *reinterpret_cast<const s16x8*>(dYT + (size_t)(co0 + 16 * i + row) * V + v);
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      const bf16* p = XT + (size_t)(ci0 + 16 * j + row) * V;
      const s16x8 c = rowok ? *reinterpret_cast<const s16x8*>(p + base) : zero;
      if (dx == 0) {
        b[j] = c;
      } else if (dx < 0) {  // voxels v - 1 .. v + 6: the last of the run to the left, then c[0..6]
        const s16x8 l = (rowok && x > 0) ? *reinterpret_cast<const s16x8*>(p + base - 8) : zero;
        b[j] = s16x8{l[7], c[0], c[1], c[2], c[3], c[4], c[5], c[6]};
      } else {  // voxels v + 1 .. v + 8: c[1..7], then the first of the run to the right
        const s16x8 r = (rowok && x + 8 < g.W) ? *reinterpret_cast<const s16x8*>(p + base + 8) : zero;
        b[j] = s16x8{c[1], c[2], c[3], c[4], c[5], c[6], c[7], r[0]};
      }
    }
#pragma unroll
    for (int i = 0; i < MC; ++i)
#pragma unroll
      for (int j = 0; j < NC; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a[i]),
                                                            __builtin_bit_cast(bf16x8, b[j]), acc[i][j], 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < MC; ++i)
#pragma unroll
    for (int j = 0; j < NC; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        part[(((size_t)blockIdx.y * Cout + co0 + 16 * i + 4 * kq + r) * 27 + t) * Cin + ci0 + 16 * j + row] =
            acc[i][j][r];
}

// part[chunk][co][t][ci] = sum_{v in chunk} dY[v][co] X[nbr(v, t)][ci] on the channels-last
// operands, element by element (grids whose x edge is not a multiple of 8, i.e. the coarsest
// levels of small grids): one wave per (tap, 16 co, 16 ci) and voxel chunk; lane l holds
// A[co = l % 16][v = 4 (l / 16) + j] and B[v = 4 (l / 16) + j][ci = l % 16]
__global__ void __launch_bounds__(64) k_conv_dw(const bf16* __restrict__ dY, const bf16* __restrict__ X,
                                                float* __restrict__ part, Vol g, int Cin, int Cout,
                                                long long chunk) {
  const int l = threadIdx.x, row = l & 15, kq = l >> 4;
  const int nci = Cin / 16;
  int idx = blockIdx.x;
  const int t = idx % 27;
  idx /= 27;
  const int ci0 = (idx % nci) * 16, co0 = (idx / nci) * 16;
  const long long V = g.voxels();
  const long long vb0 = (long long)blockIdx.y * chunk;
  const long long ve = vb0 + chunk < V ? vb0 + chunk : V;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (long long vb = vb0; vb < ve; vb += 16) {
    bf16x4 a, b;
    for (int j = 0; j < 4; ++j) {
      const long long vv = vb + 4 * kq + j;
      const bool ok = vv < ve;
      a[j] = ok ? (short)dY[vv * Cout + co0 + row] : (short)0;
      const long long nb = ok ? neighbour(g, vv, t) : -1;
      b[j] = nb >= 0 ? (short)X[nb * Cin + ci0 + row] : (short)0;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  for (int r = 0; r < 4; ++r)
    part[(((long long)blockIdx.y * Cout + co0 + 4 * kq + r) * 27 + t) * Cin + ci0 + row] = acc[r];
}

// out[i] = sum_c part[c][i], chunks in order
__global__ void k_sum_chunks(const float* __restrict__ part, float* __restrict__ out, long long n, int nch) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int c = 0; c < nch; ++c) s += part[(long long)c * n + i];
  out[i] = s;
}

// bpart[p][c] = sum of dY[v][c] over the voxels of part p (FP32): the 256 threads are
// 256 / C voxel lanes x C channels (C <= 256), so each pass reads 256 / C whole voxel rows;
// every lane keeps four partial sums (independent loads in flight), and the sums are added
// in a fixed order (deterministic)
__global__ void __launch_bounds__(256) k_bias_partial(const float* __restrict__ dY, float* __restrict__ bpart,
                                                      int V, int C, int vchunk) {
  __shared__ float red[256];
  const int nr = 256 / C, c = threadIdx.x % C, r = threadIdx.x / C;
  const int v0 = blockIdx.x * vchunk, v1 = v0 + vchunk < V ? v0 + vchunk : V;
  float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
  if (r < nr) {
    int v = v0 + r;
    for (; v + 3 * nr < v1; v += 4 * nr) {
      s0 += dY[(size_t)v * C + c];
      s1 += dY[(size_t)(v + nr) * C + c];
      s2 += dY[(size_t)(v + 2 * nr) * C + c];
      s3 += dY[(size_t)(v + 3 * nr) * C + c];
    }
    for (; v < v1; v += nr) s0 += dY[(size_t)v * C + c];
  }
  red[threadIdx.x] = (s0 + s1) + (s2 + s3);
  __syncthreads();
  if (threadIdx.x < C) {
    float t = 0.f;
    for (int k = 0; k < nr; ++k) t += red[k * C + c];
    bpart[(size_t)blockIdx.x * C + c] = t;
  }
}

// db[c] = sum over the np parts of bpart[p][c]: one workgroup per channel, each thread adds
// the parts t, t + 256, ... in order, then a tree in shared memory (fixed order)
__global__ void __launch_bounds__(256) k_bias_final(const float* __restrict__ bpart, float* __restrict__ db,
                                                    int np, int C) {
  __shared__ float red[256];
  const int c = blockIdx.x;
  float s = 0.f;
  for (int p = threadIdx.x; p < np; p += 256) s += bpart[(size_t)p * C + c];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[c] = red[0];
}

// W'[ci][26 - t][co] = W[co][t][ci] (bf16)
__global__ void k_flip_weights(const bf16* __restrict__ W, bf16* __restrict__ Wf, int Cin, int Cout) {
  const long long n = (long long)Cout * 27 * Cin;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int ci = (int)(i % Cin);
  const int t = (int)((i / Cin) % 27);
  const int co = (int)(i / ((long long)Cin * 27));
  Wf[((long long)ci * 27 + (26 - t)) * Cout + co] = W[i];
}

// a = bf16(relu(y)) (relu = 0: a = bf16(y)) into channels [off, off + C) of a tensor of Ct
// channels
__global__ void k_act(const float* __restrict__ y, bf16* __restrict__ a, long long V, int C, int Ct,
                      int off, int relu) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= V * C) return;
  const long long v = i / C;
  const int c = (int)(i - v * C);
  float x = y[i];
  if (relu && x < 0.f) x = 0.f;
  a[v * Ct + off + c] = f2bf(x);
}

// dy = (y > 0) * da[channels off .. off + C of Ct]  (dy FP32; dyb its bf16 copy)
__global__ void k_relu_back(const float* __restrict__ y, const float* __restrict__ da, float* __restrict__ dy,
                            bf16* __restrict__ dyb, long long V, int C, int Ct, int off, int relu) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= V * C) return;
  const long long v = i / C;
  const int c = (int)(i - v * C);
  const float g = da[v * Ct + off + c];
  const float d = (!relu || y[i] > 0.f) ? g : 0.f;
  dy[i] = d;
  dyb[i] = f2bf(d);
}

// 2x2x2 average pooling of bf16 activations (out bf16)
__global__ void k_pool(const bf16* __restrict__ x, bf16* __restrict__ o, Vol go, int C) {
  const long long n = go.voxels() * C;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = (int)(i % C);
  long long v = i / C;
  const int X = (int)(v % go.W);
  v /= go.W;
  const int Y = (int)(v % go.H);
  v /= go.H;
  const int Z = (int)(v % go.D);
  const long long b = v / go.D;
  const int Wi = 2 * go.W, Hi = 2 * go.H, Di = 2 * go.D;
  float s = 0.f;
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx)
        s += bf2f(x[(((b * Di + 2 * Z + dz) * Hi + 2 * Y + dy) * Wi + 2 * X + dx) * C + c]);
  o[i] = f2bf(s * 0.125f);
}

// adjoint of the pooling: dx (fine, FP32) += do / 8 into channels [off, off + C) of Ct
__global__ void k_pool_back(const float* __restrict__ dout, float* __restrict__ dx, Vol gi, int C, int Ct,
                            int off) {
  const long long n = gi.voxels() * C;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = (int)(i % C);
  long long v = i / C;
  const int X = (int)(v % gi.W);
  long long r = v / gi.W;
  const int Y = (int)(r % gi.H);
  r /= gi.H;
  const int Z = (int)(r % gi.D);
  const long long b = r / gi.D;
  const long long vc = ((b * (gi.D / 2) + Z / 2) * (gi.H / 2) + Y / 2) * (gi.W / 2) + X / 2;
  dx[v * Ct + off + c] += dout[vc * C + c] * 0.125f;
}

// nearest 2x upsampling of bf16 activations into channels [off, off + C) of Ct (fine grid)
__global__ void k_up(const bf16* __restrict__ x, bf16* __restrict__ o, Vol gf, int C, int Ct, int off) {
  const long long n = gf.voxels() * C;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = (int)(i % C);
  long long v = i / C;
  const int X = (int)(v % gf.W);
  long long r = v / gf.W;
  const int Y = (int)(r % gf.H);
  r /= gf.H;
  const int Z = (int)(r % gf.D);
  const long long b = r / gf.D;
  const long long vc = ((b * (gf.D / 2) + Z / 2) * (gf.H / 2) + Y / 2) * (gf.W / 2) + X / 2;
  o[v * Ct + off + c] = x[vc * C + c];
}

// adjoint of the upsampling: dx (coarse, FP32) = sum of the 8 fine gradients of channels
// [off, off + C) of Ct
__global__ void k_up_back(const float* __restrict__ dfine, float* __restrict__ dx, Vol gc, int C, int Ct,
                          int off) {
  const long long n = gc.voxels() * C;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = (int)(i % C);
  long long v = i / C;
  const int X = (int)(v % gc.W);
  v /= gc.W;
  const int Y = (int)(v % gc.H);
  v /= gc.H;
  const int Z = (int)(v % gc.D);
  const long long b = v / gc.D;
  const int Wf = 2 * gc.W, Hf = 2 * gc.H, Df = 2 * gc.D;
  float s = 0.f;
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx2 = 0; dx2 < 2; ++dx2)
        s += dfine[((((b * Df + 2 * Z + dz) * Hf + 2 * Y + dy) * Wf + 2 * X + dx2) * Ct) + off + c];
  dx[i] += s;
}

// loss partials: sum over the first nout channels of (y - t)^2 and the gradient
// 2 (y - t) / count (zero on the padding channels)
__global__ void __launch_bounds__(256) k_mse(const float* __restrict__ y, const float* __restrict__ t,
                                             float* __restrict__ dy, float* __restrict__ part, long long V,
                                             int C, int nout, float inv_count) {
  __shared__ float red[256];
  float s = 0.f;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < V * C; i += (long long)gridDim.x * 256) {
    const long long v = i / C;
    const int c = (int)(i - v * C);
    float g = 0.f;
    if (c < nout) {
      const float d = y[i] - t[v * nout + c];
      s += d * d;
      g = 2.f * d * inv_count;
    }
    dy[i] = g;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Adam on FP32 master weights with a bf16 copy for the matrix cores; `mask` zeroes the
// padding entries (kept exactly zero)
__global__ void k_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, bf16* __restrict__ pb, const unsigned char* __restrict__ mask,
                       long long n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (mask && !mask[i]) {
    p[i] = 0.f;
    if (pb) pb[i] = 0;
    return;
  }
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  p[i] = pi;
  if (pb) pb[i] = f2bf(pi);
}

__global__ void k_to_bf16(const float* __restrict__ x, bf16* __restrict__ o, long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) o[i] = f2bf(x[i]);
}

// dst[v][c] += src[v][off + c] (channels [off, off + C) of a tensor of Cs channels)
__global__ void k_slice_add(const float* __restrict__ src, float* __restrict__ dst, long long V, int C, int Cs,
                            int off) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= V * C) return;
  const long long v = i / C;
  const int c = (int)(i - v * C);
  dst[i] += src[v * Cs + off + c];
}

inline unsigned blocks(long long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

struct Layer {
  int cin, cout;          // padded channel counts (multiples of 16)
  int cin_real, cout_real;
  int level;              // grid n >> level
  int relu;
  long long woff, boff;   // offsets of W [cout][27][cin] and b [cout] in the flat parameters
};

// the network: 11 convolutions (see the file header)
std::vector<Layer> architecture() {
  const int L[11][6] = {
      // cin, cout, cin_real, cout_real, level, relu
      {16, 32, 11, 32, 0, 1}, {32, 32, 32, 32, 0, 1},    // e0
      {32, 64, 32, 64, 1, 1}, {64, 64, 64, 64, 1, 1},    // e1
      {64, 128, 64, 128, 2, 1}, {128, 128, 128, 128, 2, 1},  // e2
      {128, 128, 128, 128, 3, 1},                          // bottleneck
      {256, 64, 256, 64, 2, 1},                            // d2
      {128, 32, 128, 32, 1, 1},                            // d1
      {64, 32, 64, 32, 0, 1},                              // d0
      {32, 16, 32, 3, 0, 0},                               // output
  };
  std::vector<Layer> out;
  long long off = 0;
  for (auto& r : L) {
    Layer l{r[0], r[1], r[2], r[3], r[4], r[5], 0, 0};
    l.woff = off;
    off += (long long)l.cout * 27 * l.cin;
    l.boff = off;
    off += l.cout;
    out.push_back(l);
  }
  return out;
}

}  // namespace nn
}  // namespace wlc

using namespace wlc::nn;

struct wlc_unet {
  int device = 0;
  hipStream_t st = nullptr;
  int n = 0, B = 0;
  std::vector<Layer> layers;
  long long np = 0;       // flat parameter count (padded)
  float *p = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;
  bf16* pb = nullptr;     // bf16 copy of the parameters (weights used by the kernels)
  unsigned char* mask = nullptr;
  bf16* wflip = nullptr;  // flipped weights of one layer
  float* part = nullptr;  // weight-gradient partial sums
  float* bpart = nullptr; // bias-gradient partial sums
  bf16 *xt = nullptr, *dyt = nullptr;  // a layer's input and output gradient, transposed
  long long part_n = 0;
  float* lpart = nullptr;
  int adam_t = 0;
  // activations (bf16) per tensor and level
  bf16 *x0 = nullptr, *a0a = nullptr, *s0 = nullptr, *p1 = nullptr, *a1a = nullptr, *s1 = nullptr,
       *p2 = nullptr, *a2a = nullptr, *s2 = nullptr, *p3 = nullptr, *bt = nullptr, *c2 = nullptr,
       *d2 = nullptr, *c1 = nullptr, *d1 = nullptr, *c0 = nullptr, *d0 = nullptr;
  // their gradients (FP32)
  float *ga0a = nullptr, *gs0 = nullptr, *gp1 = nullptr, *ga1a = nullptr, *gs1 = nullptr, *gp2 = nullptr,
        *ga2a = nullptr, *gs2 = nullptr, *gp3 = nullptr, *gbt = nullptr, *gc2 = nullptr, *gd2 = nullptr,
        *gc1 = nullptr, *gd1 = nullptr, *gc0 = nullptr, *gd0 = nullptr;
  float* y[11] = {};      // pre-activations
  float* dy[11] = {};     // their gradients
  bf16* dyb[11] = {};     // bf16 copies (matrix-core operand)
  float* target = nullptr;
  std::vector<void*> allocs;  // parameters and optimiser state
  std::vector<void*> acts;    // activations, gradients and scratch of the current grid and batch
  Vol vol(int level) const { return Vol{B, n >> level, n >> level, n >> level}; }
};

namespace {

int ufail(int code, const std::string& m) { return wlc::host_fail(code, m); }

#define UCHK(call)                                                                          \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) return ufail(WLC_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
int ualloc(std::vector<void*>& list, T** ptr, long long count) {
  void* q = nullptr;
  if (hipMalloc(&q, sizeof(T) * (size_t)std::max<long long>(count, 1)) != hipSuccess)
    return ufail(WLC_ENOMEM, "device allocation failed (warm-start network)");
  if (hipMemset(q, 0, sizeof(T) * (size_t)std::max<long long>(count, 1)) != hipSuccess)
    return ufail(WLC_EHIP, "device initialisation failed (warm-start network)");
  list.push_back(q);
  *ptr = static_cast<T*>(q);
  return WLC_OK;
}

void free_acts(wlc_unet* u) {
  for (void* q : u->acts) (void)hipFree(q);
  u->acts.clear();
}

void unet_free(wlc_unet* u) {
  if (!u) return;
  (void)hipSetDevice(u->device);
  if (u->st) (void)hipStreamSynchronize(u->st);
  free_acts(u);
  for (void* q : u->allocs) (void)hipFree(q);
  if (u->st) (void)hipStreamDestroy(u->st);
  delete u;
}

// voxel chunks of the weight-gradient reduction (at most 64, about 4096 voxels each)
int wgrad_chunks(long long V) { return (int)std::min<long long>(64, std::max<long long>(1, V / 4096)); }
// partial sums of the bias gradient (at most; about 256 voxels each)
constexpr int kBiasParts = 1024;

// activation, gradient and scratch buffers for grid edge n and batch B
int alloc_acts(wlc_unet* u, int n, int B) {
  free_acts(u);
  u->n = n;
  u->B = B;
  int rc;
#define UA(ptr, cnt) if ((rc = ualloc(u->acts, &(ptr), (cnt)))) { free_acts(u); return rc; }
  long long maxpart = 0, maxx = 0, maxdy = 0;
  for (const Layer& l : u->layers) {
    const long long nw = (long long)l.cout * 27 * l.cin;
    const long long V = u->vol(l.level).voxels();
    maxpart = std::max(maxpart, wgrad_chunks(V) * nw);
    maxx = std::max(maxx, V * l.cin);
    maxdy = std::max(maxdy, V * l.cout);
  }
  UA(u->part, maxpart); UA(u->lpart, 256); UA(u->bpart, (long long)kBiasParts * 256);
  UA(u->xt, maxx); UA(u->dyt, maxdy);
  const long long V0 = u->vol(0).voxels(), V1 = u->vol(1).voxels(), V2 = u->vol(2).voxels(), V3 = u->vol(3).voxels();
  UA(u->x0, V0 * 16); UA(u->a0a, V0 * 32); UA(u->s0, V0 * 32); UA(u->p1, V1 * 32); UA(u->a1a, V1 * 64);
  UA(u->s1, V1 * 64); UA(u->p2, V2 * 64); UA(u->a2a, V2 * 128); UA(u->s2, V2 * 128); UA(u->p3, V3 * 128);
  UA(u->bt, V3 * 128); UA(u->c2, V2 * 256); UA(u->d2, V2 * 64); UA(u->c1, V1 * 128); UA(u->d1, V1 * 32);
  UA(u->c0, V0 * 64); UA(u->d0, V0 * 32);
  UA(u->ga0a, V0 * 32); UA(u->gs0, V0 * 32); UA(u->gp1, V1 * 32); UA(u->ga1a, V1 * 64); UA(u->gs1, V1 * 64);
  UA(u->gp2, V2 * 64); UA(u->ga2a, V2 * 128); UA(u->gs2, V2 * 128); UA(u->gp3, V3 * 128); UA(u->gbt, V3 * 128);
  UA(u->gc2, V2 * 256); UA(u->gd2, V2 * 64); UA(u->gc1, V1 * 128); UA(u->gd1, V1 * 32); UA(u->gc0, V0 * 64);
  UA(u->gd0, V0 * 32);
  for (int li = 0; li < 11; ++li) {
    const Layer& l = u->layers[li];
    const long long V = u->vol(l.level).voxels();
    UA(u->y[li], V * l.cout); UA(u->dy[li], V * l.cout); UA(u->dyb[li], V * l.cout);
  }
  UA(u->target, V0 * 3);
#undef UA
  return WLC_OK;
}

// the LDS-staged forward convolution (default; WLC_CONV_LDS=0 selects the direct one, for
// measurements and tests)
bool conv_lds_on() {
  const char* e = std::getenv("WLC_CONV_LDS");
  return !(e && e[0] == '0');
}

// launches shared by the network and the convolution test entry point (stream st)

// y = conv(X) (+ bias); X has cin channels, Y cout (multiples of 16)
int launch_conv_fwd(hipStream_t st, const bf16* X, const bf16* W, const float* bias, float* Y, Vol g, int cin,
                    int cout) {
  const long long V = g.voxels();
  const int nt = cout % 64 == 0 ? 4 : (cout % 32 == 0 ? 2 : 1);
  const dim3 grid(blocks(V, 64), cout / (16 * nt)), block(256);
#define CONV_FWD(NT, KI) hipLaunchKernelGGL((k_conv_fwd<NT, KI>), grid, block, 0, st, X, W, bias, Y, g, cin, cout)
#define CONV_FWD_L(NT) hipLaunchKernelGGL((k_conv_fwd_l<NT>), grid, block, 0, st, X, W, bias, Y, g, cin, cout)
  if (cin % 32 == 0) {
    if (conv_lds_on()) {
      if (nt == 4) CONV_FWD_L(4); else if (nt == 2) CONV_FWD_L(2); else CONV_FWD_L(1);
    } else {
      if (nt == 4) CONV_FWD(4, 32); else if (nt == 2) CONV_FWD(2, 32); else CONV_FWD(1, 32);
    }
  } else {
    if (nt == 4) CONV_FWD(4, 16); else if (nt == 2) CONV_FWD(2, 16); else CONV_FWD(1, 16);
  }
#undef CONV_FWD
#undef CONV_FWD_L
  UCHK(hipGetLastError());
  return WLC_OK;
}

// dW [cout][27][cin] and db [cout] of one convolution (overwritten) from its input X
// (channels-last bf16), the output gradient dYb (bf16) and dYf (FP32).  Scratch: xt >= V cin
// and dyt >= V cout bf16, part >= wgrad_chunks(V) cout 27 cin floats, bpart >= kBiasParts cout.
int launch_conv_wgrad(hipStream_t st, const bf16* X, const bf16* dYb, const float* dYf, Vol g, int cin, int cout,
                      bf16* xt, bf16* dyt, float* part, float* bpart, float* dW, float* db) {
  const long long V = g.voxels();
  const long long nw = (long long)cout * 27 * cin;
  const int nch = wgrad_chunks(V);
  int nrun;
  if (g.W % 8 == 0 && V % 32 == 0) {
    const long long chunk = ((V + nch - 1) / nch + 31) / 32 * 32;
    nrun = (int)((V + chunk - 1) / chunk);
    hipLaunchKernelGGL(k_transpose, dim3(blocks(V, 32), blocks(cin, 32)), dim3(256), 0, st, X, xt, (int)V, cin);
    hipLaunchKernelGGL(k_transpose, dim3(blocks(V, 32), blocks(cout, 32)), dim3(256), 0, st, dYb, dyt, (int)V, cout);
    const int mc = cout % 32 == 0 ? 2 : 1, nc = cin % 32 == 0 ? 2 : 1;
    const dim3 grid(27 * (cout / (16 * mc)) * (cin / (16 * nc)), nrun), block(64);
#define CONV_DW(MC, NC) \
  hipLaunchKernelGGL((k_conv_dw_t<MC, NC>), grid, block, 0, st, dyt, xt, part, g, cin, cout, (int)chunk)
    if (mc == 2) {
      if (nc == 2) CONV_DW(2, 2); else CONV_DW(2, 1);
    } else {
      if (nc == 2) CONV_DW(1, 2); else CONV_DW(1, 1);
    }
#undef CONV_DW
  } else {
    const long long chunk = ((V + nch - 1) / nch + 15) / 16 * 16;
    nrun = (int)((V + chunk - 1) / chunk);
    hipLaunchKernelGGL(k_conv_dw, dim3(27 * (cout / 16) * (cin / 16), nrun), dim3(64), 0, st, dYb, X, part, g, cin,
                       cout, chunk);
  }
  UCHK(hipGetLastError());
  hipLaunchKernelGGL(k_sum_chunks, dim3(blocks(nw)), dim3(256), 0, st, part, dW, nw, nrun);
  const int nbp = (int)std::min<long long>(kBiasParts, std::max<long long>(1, V / 256));
  const int vchunk = (int)((V + nbp - 1) / nbp);
  hipLaunchKernelGGL(k_bias_partial, dim3(nbp), dim3(256), 0, st, dYf, bpart, (int)V, cout, vchunk);
  hipLaunchKernelGGL(k_bias_final, dim3(cout), dim3(256), 0, st, bpart, db, nbp, cout);
  UCHK(hipGetLastError());
  return WLC_OK;
}

int conv_forward(wlc_unet* u, const bf16* X, const bf16* W, const float* bias, float* Y, Vol g, int cin,
                 int cout) {
  return launch_conv_fwd(u->st, X, W, bias, Y, g, cin, cout);
}

// gradients of layer li: dW, db into u->g; dX (FP32, overwritten) when dX != nullptr
int conv_backward(wlc_unet* u, const Layer& l, const bf16* X, float* dX) {
  const int li = (int)(&l - u->layers.data());
  const Vol g = u->vol(l.level);
  const long long nw = (long long)l.cout * 27 * l.cin;
  int rc = launch_conv_wgrad(u->st, X, u->dyb[li], u->dy[li], g, l.cin, l.cout, u->xt, u->dyt, u->part, u->bpart,
                             u->g + l.woff, u->g + l.boff);
  if (rc) return rc;
  if (dX) {
    hipLaunchKernelGGL(k_flip_weights, dim3(blocks(nw)), dim3(256), 0, u->st, u->pb + l.woff, u->wflip, l.cin, l.cout);
    rc = conv_forward(u, u->dyb[li], u->wflip, nullptr, dX, g, l.cout, l.cin);
    if (rc) return rc;
  }
  UCHK(hipGetLastError());
  return WLC_OK;
}

int act(wlc_unet* u, int li, bf16* a, int Ct, int off) {
  const Layer& l = u->layers[li];
  const long long V = u->vol(l.level).voxels();
  hipLaunchKernelGGL(k_act, dim3(blocks(V * l.cout)), dim3(256), 0, u->st, u->y[li], a, V, l.cout, Ct, off, l.relu);
  UCHK(hipGetLastError());
  return WLC_OK;
}

int relu_back(wlc_unet* u, int li, const float* da, int Ct, int off) {
  const Layer& l = u->layers[li];
  const long long V = u->vol(l.level).voxels();
  hipLaunchKernelGGL(k_relu_back, dim3(blocks(V * l.cout)), dim3(256), 0, u->st, u->y[li], da, u->dy[li],
                     u->dyb[li], V, l.cout, Ct, off, l.relu);
  UCHK(hipGetLastError());
  return WLC_OK;
}

int forward(wlc_unet* u) {
  auto& Ls = u->layers;
  int rc;
  auto C = [&](int li, const bf16* X) {
    const Layer& l = Ls[li];
    return conv_forward(u, X, u->pb + l.woff, u->p + l.boff, u->y[li], u->vol(l.level), l.cin, l.cout);
  };
  const Vol g1 = u->vol(1), g2 = u->vol(2), g3 = u->vol(3), g0 = u->vol(0);
  if ((rc = C(0, u->x0)) || (rc = act(u, 0, u->a0a, 32, 0))) return rc;
  if ((rc = C(1, u->a0a)) || (rc = act(u, 1, u->s0, 32, 0)) || (rc = act(u, 1, u->c0, 64, 32))) return rc;
  hipLaunchKernelGGL(k_pool, dim3(blocks(g1.voxels() * 32)), dim3(256), 0, u->st, u->s0, u->p1, g1, 32);
  if ((rc = C(2, u->p1)) || (rc = act(u, 2, u->a1a, 64, 0))) return rc;
  if ((rc = C(3, u->a1a)) || (rc = act(u, 3, u->s1, 64, 0)) || (rc = act(u, 3, u->c1, 128, 64))) return rc;
  hipLaunchKernelGGL(k_pool, dim3(blocks(g2.voxels() * 64)), dim3(256), 0, u->st, u->s1, u->p2, g2, 64);
  if ((rc = C(4, u->p2)) || (rc = act(u, 4, u->a2a, 128, 0))) return rc;
  if ((rc = C(5, u->a2a)) || (rc = act(u, 5, u->s2, 128, 0)) || (rc = act(u, 5, u->c2, 256, 128))) return rc;
  hipLaunchKernelGGL(k_pool, dim3(blocks(g3.voxels() * 128)), dim3(256), 0, u->st, u->s2, u->p3, g3, 128);
  if ((rc = C(6, u->p3)) || (rc = act(u, 6, u->bt, 128, 0))) return rc;
  hipLaunchKernelGGL(k_up, dim3(blocks(g2.voxels() * 128)), dim3(256), 0, u->st, u->bt, u->c2, g2, 128, 256, 0);
  if ((rc = C(7, u->c2)) || (rc = act(u, 7, u->d2, 64, 0))) return rc;
  hipLaunchKernelGGL(k_up, dim3(blocks(g1.voxels() * 64)), dim3(256), 0, u->st, u->d2, u->c1, g1, 64, 128, 0);
  if ((rc = C(8, u->c1)) || (rc = act(u, 8, u->d1, 32, 0))) return rc;
  hipLaunchKernelGGL(k_up, dim3(blocks(g0.voxels() * 32)), dim3(256), 0, u->st, u->d1, u->c0, g0, 32, 64, 0);
  if ((rc = C(9, u->c0)) || (rc = act(u, 9, u->d0, 32, 0))) return rc;
  if ((rc = C(10, u->d0))) return rc;
  UCHK(hipGetLastError());
  return WLC_OK;
}

// gradients of all parameters for the loss mse(y10[:, :3], target); *loss on the host
int backward(wlc_unet* u, float* loss) {
  auto& Ls = u->layers;
  const Vol g0 = u->vol(0), g1 = u->vol(1), g2 = u->vol(2), g3 = u->vol(3);
  const long long V0 = g0.voxels();
  int rc;
  // zero the accumulated activation gradients
  auto Z = [&](float* q, long long n) { return hipMemsetAsync(q, 0, sizeof(float) * (size_t)n, u->st); };
  UCHK(Z(u->gs0, V0 * 32));
  UCHK(Z(u->gs1, g1.voxels() * 64));
  UCHK(Z(u->gs2, g2.voxels() * 128));
  UCHK(Z(u->gbt, g3.voxels() * 128));
  UCHK(Z(u->gd2, g2.voxels() * 64));
  UCHK(Z(u->gd1, g1.voxels() * 32));
  const int nparts = 256;
  hipLaunchKernelGGL(k_mse, dim3(nparts), dim3(256), 0, u->st, u->y[10], u->target, u->dy[10], u->lpart, V0, 16, 3,
                     1.0f / (float)(V0 * 3));
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks(V0 * 16)), dim3(256), 0, u->st, u->dy[10], u->dyb[10], V0 * 16);
  if ((rc = conv_backward(u, Ls[10], u->d0, u->gd0))) return rc;
  if ((rc = relu_back(u, 9, u->gd0, 32, 0)) || (rc = conv_backward(u, Ls[9], u->c0, u->gc0))) return rc;
  hipLaunchKernelGGL(k_up_back, dim3(blocks(g1.voxels() * 32)), dim3(256), 0, u->st, u->gc0, u->gd1, g1, 32, 64, 0);
  hipLaunchKernelGGL(k_slice_add, dim3(blocks(V0 * 32)), dim3(256), 0, u->st, u->gc0, u->gs0, V0, 32, 64, 32);
  if ((rc = relu_back(u, 8, u->gd1, 32, 0)) || (rc = conv_backward(u, Ls[8], u->c1, u->gc1))) return rc;
  hipLaunchKernelGGL(k_up_back, dim3(blocks(g2.voxels() * 64)), dim3(256), 0, u->st, u->gc1, u->gd2, g2, 64, 128, 0);
  hipLaunchKernelGGL(k_slice_add, dim3(blocks(g1.voxels() * 64)), dim3(256), 0, u->st, u->gc1, u->gs1, g1.voxels(), 64, 128, 64);
  if ((rc = relu_back(u, 7, u->gd2, 64, 0)) || (rc = conv_backward(u, Ls[7], u->c2, u->gc2))) return rc;
  hipLaunchKernelGGL(k_up_back, dim3(blocks(g3.voxels() * 128)), dim3(256), 0, u->st, u->gc2, u->gbt, g3, 128, 256, 0);
  hipLaunchKernelGGL(k_slice_add, dim3(blocks(g2.voxels() * 128)), dim3(256), 0, u->st, u->gc2, u->gs2, g2.voxels(), 128, 256, 128);
  if ((rc = relu_back(u, 6, u->gbt, 128, 0)) || (rc = conv_backward(u, Ls[6], u->p3, u->gp3))) return rc;
  hipLaunchKernelGGL(k_pool_back, dim3(blocks(g2.voxels() * 128)), dim3(256), 0, u->st, u->gp3, u->gs2, g2, 128, 128, 0);
  if ((rc = relu_back(u, 5, u->gs2, 128, 0)) || (rc = conv_backward(u, Ls[5], u->a2a, u->ga2a))) return rc;
  if ((rc = relu_back(u, 4, u->ga2a, 128, 0)) || (rc = conv_backward(u, Ls[4], u->p2, u->gp2))) return rc;
  hipLaunchKernelGGL(k_pool_back, dim3(blocks(g1.voxels() * 64)), dim3(256), 0, u->st, u->gp2, u->gs1, g1, 64, 64, 0);
  if ((rc = relu_back(u, 3, u->gs1, 64, 0)) || (rc = conv_backward(u, Ls[3], u->a1a, u->ga1a))) return rc;
  if ((rc = relu_back(u, 2, u->ga1a, 64, 0)) || (rc = conv_backward(u, Ls[2], u->p1, u->gp1))) return rc;
  hipLaunchKernelGGL(k_pool_back, dim3(blocks(V0 * 32)), dim3(256), 0, u->st, u->gp1, u->gs0, g0, 32, 32, 0);
  if ((rc = relu_back(u, 1, u->gs0, 32, 0)) || (rc = conv_backward(u, Ls[1], u->a0a, u->ga0a))) return rc;
  if ((rc = relu_back(u, 0, u->ga0a, 32, 0)) || (rc = conv_backward(u, Ls[0], u->x0, nullptr))) return rc;
  UCHK(hipGetLastError());
  float lp[256];
  UCHK(hipMemcpyAsync(lp, u->lpart, sizeof(lp), hipMemcpyDeviceToHost, u->st));
  UCHK(hipStreamSynchronize(u->st));
  double s = 0.0;
  for (float x : lp) s += x;
  if (loss) *loss = (float)(s / (double)(V0 * 3));
  return WLC_OK;
}

// input [B][n^3][11] (host FP32) -> bf16 [B n^3][16]; targets [B][n^3][3]
int upload(wlc_unet* u, const float* x, const float* t) {
  const long long V = u->vol(0).voxels();
  if (x) {
    std::vector<bf16> h((size_t)V * 16, 0);
    for (long long v = 0; v < V; ++v)
      for (int c = 0; c < 11; ++c) h[(size_t)v * 16 + c] = f2bf(x[v * 11 + c]);
    UCHK(hipMemcpyAsync(u->x0, h.data(), sizeof(bf16) * h.size(), hipMemcpyHostToDevice, u->st));
    UCHK(hipStreamSynchronize(u->st));
  }
  if (t) {
    UCHK(hipMemcpyAsync(u->target, t, sizeof(float) * (size_t)V * 3, hipMemcpyHostToDevice, u->st));
    UCHK(hipStreamSynchronize(u->st));
  }
  return WLC_OK;
}

}  // namespace

extern "C" {

int wlc_unet_create(int n, int batch, unsigned long long seed, wlc_unet** out) {
  if (!out) return ufail(WLC_EINVAL, "out is NULL");
  *out = nullptr;
  if (n < 8 || n % 8 || batch < 1) return ufail(WLC_EINVAL, "grid edge must be a multiple of 8 and batch >= 1");
  if ((long long)batch * n * n * n * 256 >= (1LL << 31))
    return ufail(WLC_EINVAL, "batch x grid too large (the kernels index an activation tensor with 32 bits)");
  wlc_unet* u = new wlc_unet();
  u->n = n;
  u->B = batch;
  if (hipGetDevice(&u->device) != hipSuccess || hipStreamCreateWithFlags(&u->st, hipStreamNonBlocking) != hipSuccess) {
    delete u;
    return ufail(WLC_EHIP, "no HIP device / stream");
  }
  u->layers = architecture();
  const Layer& last = u->layers.back();
  u->np = last.boff + last.cout;
  int rc;
#define UA(ptr, cnt) if ((rc = ualloc(u->allocs, &(ptr), (cnt)))) { unet_free(u); return rc; }
  UA(u->p, u->np); UA(u->m, u->np); UA(u->v, u->np); UA(u->g, u->np); UA(u->pb, u->np); UA(u->mask, u->np);
  long long maxw = 0;
  for (const Layer& l : u->layers) maxw = std::max(maxw, (long long)l.cout * 27 * l.cin);
  UA(u->wflip, maxw);
#undef UA
  if ((rc = alloc_acts(u, n, batch))) {
    unet_free(u);
    return rc;
  }
  // He-normal weights over the real fan-in, zero biases, zero padding entries; the output
  // layer starts small
  std::vector<float> hp((size_t)u->np, 0.f);
  std::vector<unsigned char> hm((size_t)u->np, 0);
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> nd(0.f, 1.f);
  for (size_t li = 0; li < u->layers.size(); ++li) {
    const Layer& l = u->layers[li];
    const float sd = std::sqrt(2.f / (27.f * l.cin_real)) * (li + 1 == u->layers.size() ? 0.1f : 1.f);
    for (int co = 0; co < l.cout; ++co)
      for (int t = 0; t < 27; ++t)
        for (int ci = 0; ci < l.cin; ++ci) {
          const long long i = l.woff + ((long long)co * 27 + t) * l.cin + ci;
          const bool real = co < l.cout_real && ci < l.cin_real;
          const float w = nd(rng);
          hp[i] = real ? sd * w : 0.f;
          hm[i] = real ? 1 : 0;
        }
    for (int co = 0; co < l.cout; ++co) hm[l.boff + co] = co < l.cout_real ? 1 : 0;
  }
  std::vector<bf16> hb((size_t)u->np);
  for (size_t i = 0; i < hb.size(); ++i) hb[i] = f2bf(hp[i]);
  if (hipMemcpy(u->p, hp.data(), sizeof(float) * hp.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(u->pb, hb.data(), sizeof(bf16) * hb.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(u->mask, hm.data(), hm.size(), hipMemcpyHostToDevice) != hipSuccess) {
    unet_free(u);
    return ufail(WLC_EHIP, "parameter upload failed");
  }
  *out = u;
  return WLC_OK;
}

void wlc_unet_destroy(wlc_unet* u) { unet_free(u); }

int wlc_unet_resize(wlc_unet* u, int n, int batch) {
  if (!u) return ufail(WLC_EINVAL, "NULL network");
  if (n < 8 || n % 8 || batch < 1) return ufail(WLC_EINVAL, "grid edge must be a multiple of 8 and batch >= 1");
  if ((long long)batch * n * n * n * 256 >= (1LL << 31))
    return ufail(WLC_EINVAL, "batch x grid too large (the kernels index an activation tensor with 32 bits)");
  UCHK(hipSetDevice(u->device));
  UCHK(hipStreamSynchronize(u->st));
  return alloc_acts(u, n, batch);
}

long long wlc_unet_num_params(const wlc_unet* u, int padded) {
  if (!u) return WLC_EINVAL;
  if (padded) return u->np;
  long long n = 0;
  for (const Layer& l : u->layers) n += (long long)l.cout_real * 27 * l.cin_real + l.cout_real;
  return n;
}

int wlc_unet_layer(const wlc_unet* u, int i, int* info) {
  if (!u || !info || i < 0 || i >= (int)u->layers.size()) return ufail(WLC_EINVAL, "bad layer");
  const Layer& l = u->layers[i];
  info[0] = l.cin; info[1] = l.cout; info[2] = l.cin_real; info[3] = l.cout_real;
  info[4] = l.level; info[5] = l.relu; info[6] = (int)l.woff; info[7] = (int)l.boff;
  return WLC_OK;
}

int wlc_unet_forward(wlc_unet* u, const float* x, float* y) {
  if (!u || !x) return ufail(WLC_EINVAL, "NULL argument");
  UCHK(hipSetDevice(u->device));
  int rc = upload(u, x, nullptr);
  if (rc || (rc = forward(u))) return rc;
  if (y) {
    const long long V = u->vol(0).voxels();
    std::vector<float> h((size_t)V * 16);
    UCHK(hipMemcpyAsync(h.data(), u->y[10], sizeof(float) * h.size(), hipMemcpyDeviceToHost, u->st));
    UCHK(hipStreamSynchronize(u->st));
    for (long long v = 0; v < V; ++v)
      for (int c = 0; c < 3; ++c) y[v * 3 + c] = h[(size_t)v * 16 + c];
  }
  return WLC_OK;
}

int wlc_unet_gradients(wlc_unet* u, const float* x, const float* t, float* loss) {
  if (!u || !x || !t) return ufail(WLC_EINVAL, "NULL argument");
  UCHK(hipSetDevice(u->device));
  int rc = upload(u, x, t);
  if (rc || (rc = forward(u)) || (rc = backward(u, loss))) return rc;
  return WLC_OK;
}

int wlc_unet_adam(wlc_unet* u, float lr) {
  if (!u) return ufail(WLC_EINVAL, "NULL argument");
  UCHK(hipSetDevice(u->device));
  u->adam_t += 1;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.f - std::pow(b1, (float)u->adam_t), c2 = 1.f - std::pow(b2, (float)u->adam_t);
  hipLaunchKernelGGL(k_adam, dim3(blocks(u->np)), dim3(256), 0, u->st, u->p, u->m, u->v, u->g, u->pb, u->mask,
                     u->np, lr, b1, b2, eps, c1, c2);
  UCHK(hipGetLastError());
  UCHK(hipStreamSynchronize(u->st));
  return WLC_OK;
}

int wlc_unet_train_step(wlc_unet* u, const float* x, const float* t, float lr, float* loss) {
  int rc = wlc_unet_gradients(u, x, t, loss);
  if (rc) return rc;
  return wlc_unet_adam(u, lr);
}

// flat parameters / gradients (padded layout: per layer W [cout][27][cin] then b [cout])
int wlc_unet_get(wlc_unet* u, int which, float* out) {
  if (!u || !out || which < 0 || which > 1) return ufail(WLC_EINVAL, "bad arguments");
  UCHK(hipSetDevice(u->device));
  UCHK(hipStreamSynchronize(u->st));
  UCHK(hipMemcpy(out, which == 0 ? u->p : u->g, sizeof(float) * (size_t)u->np, hipMemcpyDeviceToHost));
  return WLC_OK;
}

int wlc_unet_set_params(wlc_unet* u, const float* in) {
  if (!u || !in) return ufail(WLC_EINVAL, "NULL argument");
  UCHK(hipSetDevice(u->device));
  UCHK(hipMemcpy(u->p, in, sizeof(float) * (size_t)u->np, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks(u->np)), dim3(256), 0, u->st, u->p, u->pb, u->np);
  UCHK(hipGetLastError());
  UCHK(hipStreamSynchronize(u->st));
  return WLC_OK;
}

// Test hook: one convolution and its gradients on host arrays (x [V][cin], w [cout][27][cin],
// b [cout], dy [V][cout], all rounded to bf16 where the matrix cores read them, as in the
// network): y = conv(x), dx, dw, db.  cin, cout multiples of 16.
int wlc_test_conv3d(int B, int D, int H, int W, int cin, int cout, const float* x, const float* w,
                    const float* b, const float* dy, float* y, float* dx, float* dw, float* db) {
  if (cin % 16 || cout % 16 || cin > 256 || cout > 256 || B < 1 || D < 1 || H < 1 || W < 1 || !x || !w || !b ||
      !dy || !y || !dx || !dw || !db)
    return ufail(WLC_EINVAL, "bad arguments");
  const Vol g{B, D, H, W};
  const long long V = g.voxels(), nw = (long long)cout * 27 * cin;
  if (V * std::max(cin, cout) >= (1LL << 31)) return ufail(WLC_EINVAL, "tensor too large");
  const int nch = wgrad_chunks(V);
  std::vector<bf16> hx((size_t)(V * cin)), hw((size_t)nw), hdy((size_t)(V * cout));
  for (size_t i = 0; i < hx.size(); ++i) hx[i] = f2bf(x[i]);
  for (size_t i = 0; i < hw.size(); ++i) hw[i] = f2bf(w[i]);
  for (size_t i = 0; i < hdy.size(); ++i) hdy[i] = f2bf(dy[i]);
  bf16 *dX = nullptr, *dW = nullptr, *dDY = nullptr, *dWf = nullptr, *dXT = nullptr, *dDYT = nullptr;
  float *dB = nullptr, *dYo = nullptr, *dDX = nullptr, *dDW = nullptr, *dDB = nullptr, *dPart = nullptr,
        *dDYf = nullptr, *dBpart = nullptr;
  hipError_t e = hipSuccess;
  auto M = [&](void** q, size_t bytes) { if (e == hipSuccess) e = hipMalloc(q, bytes); };
  M((void**)&dX, sizeof(bf16) * hx.size()); M((void**)&dW, sizeof(bf16) * nw); M((void**)&dDY, sizeof(bf16) * hdy.size());
  M((void**)&dWf, sizeof(bf16) * nw); M((void**)&dB, sizeof(float) * cout); M((void**)&dYo, sizeof(float) * V * cout);
  M((void**)&dDX, sizeof(float) * V * cin); M((void**)&dDW, sizeof(float) * nw); M((void**)&dDB, sizeof(float) * cout);
  M((void**)&dPart, sizeof(float) * nw * nch); M((void**)&dDYf, sizeof(float) * V * cout);
  M((void**)&dXT, sizeof(bf16) * hx.size()); M((void**)&dDYT, sizeof(bf16) * hdy.size());
  M((void**)&dBpart, sizeof(float) * kBiasParts * cout);
  if (e == hipSuccess) e = hipMemcpy(dX, hx.data(), sizeof(bf16) * hx.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dW, hw.data(), sizeof(bf16) * nw, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dDY, hdy.data(), sizeof(bf16) * hdy.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dB, b, sizeof(float) * cout, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dDYf, dy, sizeof(float) * V * cout, hipMemcpyHostToDevice);
  int rc = WLC_OK;
  if (e == hipSuccess) {
    rc = launch_conv_fwd(nullptr, dX, dW, dB, dYo, g, cin, cout);
    if (!rc) rc = launch_conv_wgrad(nullptr, dX, dDY, dDYf, g, cin, cout, dXT, dDYT, dPart, dBpart, dDW, dDB);
    if (!rc) {
      hipLaunchKernelGGL(k_flip_weights, dim3(blocks(nw)), dim3(256), 0, nullptr, dW, dWf, cin, cout);
      rc = launch_conv_fwd(nullptr, dDY, dWf, nullptr, dDX, g, cout, cin);
    }
    if (!rc) e = hipGetLastError();
  }
  if (!rc && e == hipSuccess) e = hipDeviceSynchronize();
  if (!rc && e == hipSuccess) e = hipMemcpy(y, dYo, sizeof(float) * V * cout, hipMemcpyDeviceToHost);
  if (!rc && e == hipSuccess) e = hipMemcpy(dx, dDX, sizeof(float) * V * cin, hipMemcpyDeviceToHost);
  if (!rc && e == hipSuccess) e = hipMemcpy(dw, dDW, sizeof(float) * nw, hipMemcpyDeviceToHost);
  if (!rc && e == hipSuccess) e = hipMemcpy(db, dDB, sizeof(float) * cout, hipMemcpyDeviceToHost);
  for (void* q : {(void*)dX, (void*)dW, (void*)dDY, (void*)dWf, (void*)dB, (void*)dYo, (void*)dDX, (void*)dDW,
                  (void*)dDB, (void*)dPart, (void*)dDYf, (void*)dXT, (void*)dDYT, (void*)dBpart})
    if (q) (void)hipFree(q);
  if (rc) return rc;
  if (e != hipSuccess) return ufail(WLC_EHIP, std::string("test conv3d: ") + hipGetErrorString(e));
  return WLC_OK;
}

}  // extern "C"
