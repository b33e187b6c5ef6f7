// Warm-start network for the SCFT fields: a 3D convolutional U-Net trained and run with
// hand-written kernels.  The reference has no counterpart (it starts every SCFT from the
// hand-built fields of init_w, init_array_2dcomp_mod.f90:843-1065); this network proposes
// initial fields from a seed field and the instance's parameters, and plain SCFT stays the
// judge of the result.
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

// Y[v][co] = bias[co] + sum_{t, ci} W[co][t][ci] X[nbr(v, t)][ci]: one wave per 16 voxels x
// 16 output channels; lane l holds A[l % 16][4 (l / 16) + j] (voxel row, channel group) and
// B[4 (l / 16) + j][l % 16] (channel group, output column) and receives
// D[4 (l / 16) + r][l % 16]
__global__ void __launch_bounds__(64) k_conv_fwd(const bf16* __restrict__ X, const bf16* __restrict__ Wt,
                                                 const float* __restrict__ bias, float* __restrict__ Y,
                                                 Vol g, int Cin, int Cout) {
  const int l = threadIdx.x, row = l & 15, kq = l >> 4;
  const long long V = g.voxels();
  const long long v0 = (long long)blockIdx.x * 16;
  const int co0 = blockIdx.y * 16;
  const long long v = v0 + row;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  const bf16x4 zero = {0, 0, 0, 0};
  for (int t = 0; t < 27; ++t) {
    const long long nb = v < V ? neighbour(g, v, t) : -1;
    const bf16* xr = X + (nb < 0 ? 0 : nb) * Cin + 4 * kq;
    const bf16* wr = Wt + ((long long)(co0 + row) * 27 + t) * Cin + 4 * kq;
    for (int c0 = 0; c0 < Cin; c0 += 16) {
      const bf16x4 a = nb >= 0 ? *reinterpret_cast<const bf16x4*>(xr + c0) : zero;
      const bf16x4 b = *reinterpret_cast<const bf16x4*>(wr + c0);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
  }
  const float bv = bias ? bias[co0 + row] : 0.f;
  for (int r = 0; r < 4; ++r) {
    const long long vv = v0 + 4 * kq + r;
    if (vv < V) Y[vv * Cout + co0 + row] = acc[r] + bv;
  }
}

// part[chunk][co][t][ci] = sum_{v in chunk} dY[v][co] X[nbr(v, t)][ci]: one wave per (tap,
// 16 co, 16 ci) and voxel chunk; lane l holds A[co = l % 16][v = 4 (l / 16) + j] and
// B[v = 4 (l / 16) + j][ci = l % 16]
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

// db[co] = sum_v dY[v][co] (FP32), one workgroup per channel, fixed order
__global__ void __launch_bounds__(256) k_bias_grad(const float* __restrict__ dY, float* __restrict__ db,
                                                   long long V, int C) {
  __shared__ float red[256];
  const int co = blockIdx.x;
  float s = 0.f;
  for (long long v = threadIdx.x; v < V; v += 256) s += dY[v * C + co];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[co] = red[0];
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

// activation, gradient and scratch buffers for grid edge n and batch B
int alloc_acts(wlc_unet* u, int n, int B) {
  free_acts(u);
  u->n = n;
  u->B = B;
  int rc;
#define UA(ptr, cnt) if ((rc = ualloc(u->acts, &(ptr), (cnt)))) { free_acts(u); return rc; }
  long long maxpart = 0;
  for (const Layer& l : u->layers) {
    const long long nw = (long long)l.cout * 27 * l.cin;
    const long long V = u->vol(l.level).voxels();
    const long long nch = std::min<long long>(64, std::max<long long>(1, V / 4096));
    maxpart = std::max(maxpart, nch * nw);
  }
  UA(u->part, maxpart); UA(u->lpart, 256);
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

// y = conv(X) (+ bias) of layer li on level lv; X has l.cin channels
int conv_forward(wlc_unet* u, const bf16* X, const bf16* W, const float* bias, float* Y, Vol g, int cin,
                 int cout) {
  const long long V = g.voxels();
  hipLaunchKernelGGL(k_conv_fwd, dim3(blocks(V, 16), cout / 16), dim3(64), 0, u->st, X, W, bias, Y, g, cin, cout);
  UCHK(hipGetLastError());
  return WLC_OK;
}

// gradients of layer li: dW, db into u->g; dX (FP32, overwritten) when dX != nullptr
int conv_backward(wlc_unet* u, const Layer& l, const bf16* X, float* dX) {
  const int li = (int)(&l - u->layers.data());
  const Vol g = u->vol(l.level);
  const long long V = g.voxels();
  const int nch = (int)std::min<long long>(64, std::max<long long>(1, V / 4096));
  const long long chunk = ((V + nch - 1) / nch + 15) / 16 * 16;
  const long long nw = (long long)l.cout * 27 * l.cin;
  hipLaunchKernelGGL(k_conv_dw, dim3(27 * (l.cout / 16) * (l.cin / 16), nch), dim3(64), 0, u->st, u->dyb[li], X,
                     u->part, g, l.cin, l.cout, chunk);
  UCHK(hipGetLastError());
  hipLaunchKernelGGL(k_sum_chunks, dim3(blocks(nw)), dim3(256), 0, u->st, u->part, u->g + l.woff, nw, nch);
  hipLaunchKernelGGL(k_bias_grad, dim3(l.cout), dim3(256), 0, u->st, u->dy[li], u->g + l.boff, V, l.cout);
  if (dX) {
    hipLaunchKernelGGL(k_flip_weights, dim3(blocks(nw)), dim3(256), 0, u->st, u->pb + l.woff, u->wflip, l.cin, l.cout);
    int rc = conv_forward(u, u->dyb[li], u->wflip, nullptr, dX, g, l.cout, l.cin);
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
  if (cin % 16 || cout % 16 || B < 1 || D < 1 || H < 1 || W < 1 || !x || !w || !b || !dy || !y || !dx || !dw || !db)
    return ufail(WLC_EINVAL, "bad arguments");
  const Vol g{B, D, H, W};
  const long long V = g.voxels(), nw = (long long)cout * 27 * cin;
  const int nch = (int)std::min<long long>(64, std::max<long long>(1, V / 4096));
  const long long chunk = ((V + nch - 1) / nch + 15) / 16 * 16;
  std::vector<bf16> hx((size_t)(V * cin)), hw((size_t)nw), hdy((size_t)(V * cout));
  for (size_t i = 0; i < hx.size(); ++i) hx[i] = f2bf(x[i]);
  for (size_t i = 0; i < hw.size(); ++i) hw[i] = f2bf(w[i]);
  for (size_t i = 0; i < hdy.size(); ++i) hdy[i] = f2bf(dy[i]);
  bf16 *dX = nullptr, *dW = nullptr, *dDY = nullptr, *dWf = nullptr;
  float *dB = nullptr, *dYo = nullptr, *dDX = nullptr, *dDW = nullptr, *dDB = nullptr, *dPart = nullptr, *dDYf = nullptr;
  hipError_t e = hipSuccess;
  auto M = [&](void** q, size_t bytes) { if (e == hipSuccess) e = hipMalloc(q, bytes); };
  M((void**)&dX, sizeof(bf16) * hx.size()); M((void**)&dW, sizeof(bf16) * nw); M((void**)&dDY, sizeof(bf16) * hdy.size());
  M((void**)&dWf, sizeof(bf16) * nw); M((void**)&dB, sizeof(float) * cout); M((void**)&dYo, sizeof(float) * V * cout);
  M((void**)&dDX, sizeof(float) * V * cin); M((void**)&dDW, sizeof(float) * nw); M((void**)&dDB, sizeof(float) * cout);
  M((void**)&dPart, sizeof(float) * nw * nch); M((void**)&dDYf, sizeof(float) * V * cout);
  if (e == hipSuccess) e = hipMemcpy(dX, hx.data(), sizeof(bf16) * hx.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dW, hw.data(), sizeof(bf16) * nw, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dDY, hdy.data(), sizeof(bf16) * hdy.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dB, b, sizeof(float) * cout, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dDYf, dy, sizeof(float) * V * cout, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_conv_fwd, dim3(blocks(V, 16), cout / 16), dim3(64), 0, nullptr, dX, dW, dB, dYo, g, cin, cout);
    hipLaunchKernelGGL(k_conv_dw, dim3(27 * (cout / 16) * (cin / 16), nch), dim3(64), 0, nullptr, dDY, dX, dPart, g,
                       cin, cout, chunk);
    hipLaunchKernelGGL(k_sum_chunks, dim3(blocks(nw)), dim3(256), 0, nullptr, dPart, dDW, nw, nch);
    hipLaunchKernelGGL(k_bias_grad, dim3(cout), dim3(256), 0, nullptr, dDYf, dDB, V, cout);
    hipLaunchKernelGGL(k_flip_weights, dim3(blocks(nw)), dim3(256), 0, nullptr, dW, dWf, cin, cout);
    hipLaunchKernelGGL(k_conv_fwd, dim3(blocks(V, 16), cin / 16), dim3(64), 0, nullptr, dDY, dWf, nullptr, dDX, g,
                       cout, cin);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(y, dYo, sizeof(float) * V * cout, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(dx, dDX, sizeof(float) * V * cin, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(dw, dDW, sizeof(float) * nw, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(db, dDB, sizeof(float) * cout, hipMemcpyDeviceToHost);
  for (void* q : {(void*)dX, (void*)dW, (void*)dDY, (void*)dWf, (void*)dB, (void*)dYo, (void*)dDX, (void*)dDW,
                  (void*)dDB, (void*)dPart, (void*)dDYf})
    if (q) (void)hipFree(q);
  if (e != hipSuccess) return ufail(WLC_EHIP, std::string("test conv3d: ") + hipGetErrorString(e));
  return WLC_OK;
}

}  // extern "C"
